// The two-box grid through the C++ header with the reference's signature
// (NPME_RecSumGrid_FullTrans_Func, src/RecSumGrid.h:99-112): a user kernel
// f = r sampled on the host.  Prints the grid values, one per line, for the
// caller to compare with the oracle; a built-in kernel object takes the
// device path and is not run here.
#include <cmath>
#include <cstdio>
#include <vector>

#include "npme_amd.hpp"

using namespace NPME_Library;

class RFunc : public NPME_KfuncReal {
 public:
  void Print(std::ostream &) const override {}
  void Calc(const size_t N, double *x_f0, const double *y, const double *z) const override
  {
    for (size_t i = 0; i < N; i++) x_f0[i] = std::sqrt(x_f0[i] * x_f0[i] + y[i] * y[i] + z[i] * z[i]);
  }
  void Calc(const size_t N, double *f0, double *x, double *y, double *z) const override
  {
    for (size_t i = 0; i < N; i++) {
      const double r = std::sqrt(x[i] * x[i] + y[i] * y[i] + z[i] * z[i]);
      f0[i] = r;
      x[i] /= r;
      y[i] /= r;
      z[i] /= r;
    }
  }
};

int main()
{
  const long N1 = 12, N2 = 8, N3 = 10;
  std::vector<_Complex double> f(N1 * N2 * N3);
  const double R0[3] = {20.0, -2.0, 1.0};
  RFunc func;
  if (!NPME_RecSumGrid_FullTrans_Func(f.data(), N1, 4.0, 3.0, N2, 4.0, 3.0, N3, 4.0, 3.0, 4.0, 3.5,
                                      4.5, func, R0, 1, false, std::cout))
    return 1;
  for (const _Complex double &v : f) std::printf("%.17g %.17g\n", __real__ v, __imag__ v);
  // a separation inside the grid box is refused (the reference exits)
  const double close[3] = {1.0, 0.0, 0.0};
  std::ostringstream err;
  if (NPME_RecSumGrid_FullTrans_Func(f.data(), N1, 4.0, 3.0, N2, 4.0, 3.0, N3, 4.0, 3.0, 4.0, 3.5, 4.5,
                                     func, close, 1, false, err))
    return 2;
  return err.str().find("separation") == std::string::npos ? 3 : 0;
}
