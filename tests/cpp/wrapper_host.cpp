// Host-only use of the C++ drop-in header: keywords, NPME_AddMissingKeywords
// (RSEM grid choice) and kernel selection, written as a caller of the
// reference's NPME_Library interface would write it.  Prints the chosen
// parameters; makes no GPU call.
#include <cstdio>
#include <cstdlib>
#include <fstream>

#include "npme_amd.hpp"

int main(int argc, char **argv)
{
  using namespace NPME_Library;
  if (argc != 2) return 64;
  // coordinates and charges: "x y z q" per line
  std::ifstream in(argv[1]);
  std::vector<double> coord, q;
  double x, y, z, c;
  while (in >> x >> y >> z >> c) {
    coord.push_back(x);
    coord.push_back(y);
    coord.push_back(z);
    q.push_back(c);
  }
  const size_t n = q.size();
  NPME_KeywordInput kw;
  kw.AddKeyword("calcType", "pme");
  kw.AddKeyword("funcType", "Laplace");
  kw.AddKeyword("Rdir", 8.0);
  kw.AddKeyword("BnOrder", 6);
  kw.AddKeyword("tol", 1e-6);
  if (!NPME_AddMissingKeywords("input", kw, n, coord, q, true, std::cout)) return 1;
  NPME_KeywordInput copy = kw;  // deep copy
  long N[3], nb[3];
  int nDeriv = 0;
  const char *kN[3] = {"N1", "N2", "N3"}, *kn[3] = {"n1", "n2", "n3"};
  for (int d = 0; d < 3; d++)
    if (!copy.GetValue(kN[d], N[d]) || !copy.GetValue(kn[d], nb[d])) return 2;
  if (!copy.GetValue("nDeriv", nDeriv)) return 3;
  NPME_KernelList kernels;
  NPME_KfuncReal *f = nullptr, *fLR = nullptr, *fSR = nullptr;
  if (!NPME_Interface_SelectKernelPtr(f, fLR, fSR, kernels, kw, true, std::cout)) return 4;
  if (f != &kernels.funcLaplace || fLR != &kernels.funcLaplace_LR_DM) return 5;
  NPME_KfuncComplex *g = nullptr, *gLR = nullptr, *gSR = nullptr;
  std::ostringstream quiet;
  if (NPME_Interface_SelectKernelPtr(g, gLR, gSR, kernels, kw, true, quiet)) return 6;
  std::printf("%ld %ld %ld %ld %ld %ld %d\n", N[0], N[1], N[2], nb[0], nb[1], nb[2], nDeriv);
  return 0;
}
