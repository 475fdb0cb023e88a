// Caching device allocator.  Handles are created and destroyed per request
// when serving; hipMalloc/hipFree of the grid buffers would then dominate a
// small request's latency, so freed blocks are kept in per-device free lists
// keyed by size class and handed out again.  Size classes: powers of two up
// to 1 MiB, then multiples of 2 MiB rounded up by at most 1/8.
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "npme_internal.h"

namespace npme {

namespace {

struct Pool {
  std::mutex mtx;
  // (device, class bytes) -> free blocks
  std::map<std::pair<int, size_t>, std::vector<void *>> freeList;
  std::unordered_map<void *, std::pair<int, size_t>> live;
  size_t cached = 0, inUse = 0;
};

Pool &pool()
{
  static Pool *p = new Pool;  // never destroyed: blocks live until process exit
  return *p;
}

size_t size_class(size_t bytes)
{
  if (bytes <= 256) return 256;
  if (bytes <= (1u << 20)) {
    size_t c = 256;
    while (c < bytes) c <<= 1;
    return c;
  }
  const size_t unit = 2u << 20;
  size_t units = (bytes + unit - 1) / unit;
  // round the number of units up to 4 significant bits (<= 12.5% waste)
  size_t shift = 0;
  while ((units >> shift) > 15) shift++;
  units = ((units + ((size_t)1 << shift) - 1) >> shift) << shift;
  return units * unit;
}

thread_local long g_failAfter = -1;

}  // namespace

void dev_fail_after(long k) { g_failAfter = k < 0 ? -1 : k; }

void *dev_alloc(size_t bytes)
{
  if (g_failAfter >= 0 && g_failAfter-- == 0) return nullptr;
  int dev = 0;
  NPME_HIP_CHECK(hipGetDevice(&dev));
  const size_t cls = size_class(std::max<size_t>(bytes, 1));
  Pool &p = pool();
  {
    std::lock_guard<std::mutex> g(p.mtx);
    auto it = p.freeList.find({dev, cls});
    if (it != p.freeList.end() && !it->second.empty()) {
      void *ptr = it->second.back();
      it->second.pop_back();
      p.cached -= cls;
      p.inUse += cls;
      p.live[ptr] = {dev, cls};
      return ptr;
    }
  }
  void *ptr = nullptr;
  hipError_t e = hipMalloc(&ptr, cls);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    // release the cache of this device and try once more
    dev_trim();
    e = hipMalloc(&ptr, cls);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  std::lock_guard<std::mutex> g(pool().mtx);
  p.inUse += cls;
  p.live[ptr] = {dev, cls};
  return ptr;
}

void dev_free(void *ptr)
{
  if (!ptr) return;
  Pool &p = pool();
  std::lock_guard<std::mutex> g(p.mtx);
  auto it = p.live.find(ptr);
  if (it == p.live.end()) return;
  const auto key = it->second;
  p.live.erase(it);
  p.inUse -= key.second;
  p.cached += key.second;
  p.freeList[key].push_back(ptr);
}

size_t dev_block_bytes(size_t bytes) { return size_class(std::max<size_t>(bytes, 1)); }

void dev_trim()
{
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return;
  Pool &p = pool();
  std::vector<void *> rel;
  {
    std::lock_guard<std::mutex> g(p.mtx);
    for (auto &kv : p.freeList)
      if (kv.first.first == dev) {
        for (void *ptr : kv.second) {
          rel.push_back(ptr);
          p.cached -= kv.first.second;
        }
        kv.second.clear();
      }
  }
  if (!rel.empty()) (void)hipDeviceSynchronize();
  for (void *ptr : rel) (void)hipFree(ptr);
}

void dev_stats(size_t *inUse, size_t *cached)
{
  Pool &p = pool();
  std::lock_guard<std::mutex> g(p.mtx);
  if (inUse) *inUse = p.inUse;
  if (cached) *cached = p.cached;
}

}  // namespace npme
