// npme_count_pairs_noalloc: the in-cutoff pair count of a set-up without any
// device allocation.  The per-workgroup partial counts and the total live in
// a file-scope device array (`scratch`), one copy per GPU, whose device
// address is looked up once per device with hipGetSymbolAddress.
//
// Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU (from several threads or handles) race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "npme_internal.h"

namespace npme {
namespace {

constexpr int kCountBlock = 256;
// partial counts of up to kCountMaxGroups workgroups of kCountBlock targets,
// then the total: 1 MiB per GPU, at most 33,554,432 charges
constexpr long kCountMaxGroups = 131072;
constexpr long kCountMaxCharges = kCountMaxGroups * kCountBlock;

static __device__ __align__(256) unsigned long long scratch[kCountMaxGroups + 1];

constexpr int kMaxDevices = 64;
std::mutex g_scratchMtx;
unsigned long long *g_scratchPtr[kMaxDevices] = {nullptr};

// device address of `scratch` on device dev, looked up on first use with
// dev current (the caller makes it current)
unsigned long long *scratch_ptr(int dev)
{
  if (dev < 0 || dev >= kMaxDevices) throw Error{NPME_ERR_DEVICE, "device index out of range"};
  std::lock_guard<std::mutex> lk(g_scratchMtx);
  if (!g_scratchPtr[dev]) {
    void *p = nullptr;
    NPME_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error{NPME_ERR_DEVICE, "hipGetSymbolAddress returned no address"};
    g_scratchPtr[dev] = static_cast<unsigned long long *>(p);
  }
  return g_scratchPtr[dev];
}

struct CountGeo {
  double R2;
  int nNeigh;
  int nb0, nb1, nb2;
};

// pairs (s, j), j != s, r < Rdir of every target s of the workgroup, over
// the same stencil cells as the direct sum; one partial per workgroup
__global__ __launch_bounds__(kCountBlock) void k_pairs_partial(CountGeo g, size_t n,
                                                              const int4 *__restrict__ cell,
                                                              const int2 *__restrict__ cellSE,
                                                              const double4 *__restrict__ xs,
                                                              unsigned long long *__restrict__ part)
{
  __shared__ unsigned long long wsum[kCountBlock / 64];
  const size_t s = (size_t)blockIdx.x * kCountBlock + threadIdx.x;
  unsigned long long cnt = 0;
  if (s < n) {
    const int4 c = cell[s];
    const double xi = xs[s].x, yi = xs[s].y, zi = xs[s].z;
    for (int dx = -g.nNeigh; dx <= g.nNeigh; dx++)
      for (int dy = -g.nNeigh; dy <= g.nNeigh; dy++)
        for (int dz = -g.nNeigh; dz <= g.nNeigh; dz++) {
          const int gx = c.x + dx, gy = c.y + dy, gz = c.z + dz;
          if (gx < 0 || gy < 0 || gz < 0 || gx >= g.nb0 || gy >= g.nb1 || gz >= g.nb2) continue;
          const int2 se = cellSE[((long)gx * g.nb1 + gy) * g.nb2 + gz];
          for (int j = se.x; j < se.x + se.y; j++) {
            if ((size_t)j == s) continue;
            const double4 o = xs[j];
            const double rx = xi - o.x, ry = yi - o.y, rz = zi - o.z;
            if (rx * rx + ry * ry + rz * rz < g.R2) cnt++;
          }
        }
  }
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_down(cnt, off, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < kCountBlock / 64; w++) t += wsum[w];
    part[blockIdx.x] = t;
  }
}

// total of nPart partials into *total (integer sum, order-independent)
__global__ __launch_bounds__(kCountBlock) void k_pairs_total(long nPart,
                                                            const unsigned long long *__restrict__ part,
                                                            unsigned long long *__restrict__ total)
{
  __shared__ unsigned long long wsum[kCountBlock / 64];
  unsigned long long t = 0;
  for (long i = threadIdx.x; i < nPart; i += kCountBlock) t += part[i];
  for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long a = 0;
    for (int w = 0; w < kCountBlock / 64; w++) a += wsum[w];
    *total = a;
  }
}

// makes dev current for the scope and restores the caller's device
struct DeviceScope {
  int prev = -1;
  explicit DeviceScope(int dev)
  {
    NPME_HIP_CHECK(hipGetDevice(&prev));
    if (prev != dev) NPME_HIP_CHECK(hipSetDevice(dev));
  }
  ~DeviceScope()
  {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

}  // namespace

double handle_count_pairs_noalloc(Handle &h)
{
  if (!h.ready) throw Error{NPME_ERR_ARG, "no valid set-up: the last SetUp or ResetCoord failed"};
  if ((long)h.n > kCountMaxCharges)
    throw Error{NPME_ERR_ARG, "npme_count_pairs_noalloc: nCharge " + std::to_string(h.n) +
                                  " is above its limit of " + std::to_string(kCountMaxCharges) +
                                  " (static scratch of " + std::to_string(kCountMaxGroups) +
                                  " workgroup partials); use npme_count_pairs"};
  DeviceScope ds(h.device);
  unsigned long long *part = scratch_ptr(h.device);
  unsigned long long *total = part + kCountMaxGroups;
  const DirectParams dp = direct_params(h);
  CountGeo g;
  g.R2 = dp.Rdir * dp.Rdir;
  g.nNeigh = dp.nNeigh;
  g.nb0 = dp.nb[0];
  g.nb1 = dp.nb[1];
  g.nb2 = dp.nb[2];
  const long nPart = (long)((h.n + kCountBlock - 1) / kCountBlock);
  // every partial in [0, nPart) is written before it is read, and the total
  // is written by the second kernel: nothing of an earlier call is read
  hipLaunchKernelGGL(k_pairs_partial, dim3((unsigned)nPart), dim3(kCountBlock), 0, h.stream, g, h.n,
                     h.d_cell, (const int2 *)h.d_cellStart, h.d_xs, part);
  NPME_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_pairs_total, dim3(1), dim3(kCountBlock), 0, h.stream, nPart,
                     (const unsigned long long *)part, total);
  NPME_HIP_CHECK(hipGetLastError());
  unsigned long long tot = 0;
  NPME_HIP_CHECK(hipMemcpyAsync(&tot, total, sizeof(tot), hipMemcpyDeviceToHost, h.stream));
  NPME_HIP_CHECK(hipStreamSynchronize(h.stream));
  return (double)tot;
}

}  // namespace npme
