// Device memory of the library's own buffers: a caching allocator.
//
// A lattice's context holds ~30 device buffers (state, lists, shares,
// per-site work arrays) and its optimizer 7 + 2m vectors; with one
// hipMalloc each, admitting 64 lattices spent ~0.1 s in hipMalloc and
// releasing them ~0.1 s per 1,000 hipFree.  Here blocks are carved from
// 64 MiB slabs (or taken whole for large requests) and kept on free lists
// keyed by device and rounded size when released, so a batch or a serving
// loop that admits and retires lattices reuses its memory without calling
// the runtime.  Sizes round up to a quarter of their power of two (at most
// 25% over) and to 256 bytes, so every block is 256-byte aligned.
//
// Release synchronises the block's device first, as hipFree does: a block
// may still be read by work in flight on any stream, and it must not be
// handed out again before that work ends.  Memory stays with the process
// and is reused; when the device runs out, the cached blocks that are whole
// allocations (requests above 16 MiB) go back to it and the request is tried
// once more, so a loop over many lattice sizes cannot exhaust HBM through
// the cache's large blocks.  Pinned host staging buffers
// (hmalloc / hfree) are cached the same way, whole (hipHostMalloc pins
// pages, ~0.1-0.5 ms a call; release waits for the device, whose copies
// may still read them).
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
#include <unordered_map>
#include <utility>
#include <vector>

#include "fe_internal.hpp"

namespace a2c {
namespace {

constexpr size_t GRANULE = 256;
constexpr size_t SLAB = size_t(64) << 20;  // carved into blocks of at most SLAB / 4
constexpr size_t CARVE_MAX = SLAB / 4;     // larger requests get an allocation of their own

size_t round_size(size_t bytes) {
  if (bytes <= GRANULE) return GRANULE;
  size_t p = GRANULE;
  while (p < bytes) p <<= 1;  // p / 2 < bytes <= p
  const size_t q = p / 4 < GRANULE ? GRANULE : p / 4;
  return (bytes + q - 1) / q * q;
}

struct Slab {
  char* base = nullptr;
  size_t used = 0;
};

struct Pool {
  std::mutex mu;
  std::map<std::pair<int, size_t>, std::vector<void*>> free_blocks;  // (device, rounded size)
  std::unordered_map<const void*, std::pair<int, size_t>> owner;      // block -> (device, rounded size)
  std::map<int, Slab> slab;                                           // current slab per device
  std::map<size_t, std::vector<void*>> free_host;                     // pinned host, by rounded size
  std::unordered_map<const void*, size_t> host_owner;
};

Pool& pool() {
  static Pool* p = new Pool;  // never destroyed: blocks may be released during static teardown
  return *p;
}

// hipMalloc, and on failure once more after returning the device's cached
// whole allocations (caller holds the pool lock; those blocks are idle)
hipError_t device_alloc(Pool& P, int dev, void** out, size_t bytes) {
  if (hipMalloc(out, bytes) == hipSuccess) return hipSuccess;
  (void)hipGetLastError();
  for (auto it = P.free_blocks.begin(); it != P.free_blocks.end(); ++it)
    if (it->first.first == dev && it->first.second > CARVE_MAX) {
      for (void* b : it->second) (void)hipFree(b);
      it->second.clear();
    }
  return hipMalloc(out, bytes);
}

}  // namespace

hipError_t dmalloc(void** ptr, size_t bytes) {
  if (!ptr) return hipErrorInvalidValue;
  *ptr = nullptr;
  int dev = 0;
  if (hipError_t e = hipGetDevice(&dev)) return e;
  const size_t sz = round_size(bytes);
  Pool& P = pool();
  std::lock_guard<std::mutex> lk(P.mu);
  auto it = P.free_blocks.find({dev, sz});
  if (it != P.free_blocks.end() && !it->second.empty()) {
    *ptr = it->second.back();
    it->second.pop_back();
    P.owner[*ptr] = {dev, sz};
    return hipSuccess;
  }
  void* blk = nullptr;
  if (sz > CARVE_MAX) {
    if (hipError_t e = device_alloc(P, dev, &blk, sz)) return e;
  } else {
    Slab& s = P.slab[dev];
    if (!s.base || s.used + sz > SLAB) {  // the rest of a full slab stays unused
      void* b = nullptr;
      if (hipError_t e = device_alloc(P, dev, &b, SLAB)) return e;
      s.base = static_cast<char*>(b);
      s.used = 0;
    }
    blk = s.base + s.used;
    s.used += sz;
  }
  P.owner[blk] = {dev, sz};
  *ptr = blk;
  return hipSuccess;
}

hipError_t dfree(const void* ptr) {
  if (!ptr) return hipSuccess;
  Pool& P = pool();
  std::pair<int, size_t> o;
  {
    std::lock_guard<std::mutex> lk(P.mu);
    auto it = P.owner.find(ptr);
    if (it == P.owner.end()) return hipErrorInvalidValue;
    o = it->second;
  }
  int cur = 0;
  if (hipError_t e = hipGetDevice(&cur)) return e;
  if (cur != o.first) (void)hipSetDevice(o.first);
  const hipError_t e = hipDeviceSynchronize();  // work in flight may still use the block
  if (cur != o.first) (void)hipSetDevice(cur);
  std::lock_guard<std::mutex> lk(P.mu);
  P.owner.erase(ptr);
  P.free_blocks[o].push_back(const_cast<void*>(ptr));
  return e;
}

hipError_t hmalloc(void** ptr, size_t bytes) {
  if (!ptr) return hipErrorInvalidValue;
  *ptr = nullptr;
  const size_t sz = round_size(bytes);
  Pool& P = pool();
  std::lock_guard<std::mutex> lk(P.mu);
  auto it = P.free_host.find(sz);
  if (it != P.free_host.end() && !it->second.empty()) {
    *ptr = it->second.back();
    it->second.pop_back();
  } else if (hipError_t e = hipHostMalloc(ptr, sz)) {
    return e;
  }
  P.host_owner[*ptr] = sz;
  return hipSuccess;
}

hipError_t hfree(const void* ptr) {
  if (!ptr) return hipSuccess;
  Pool& P = pool();
  size_t sz = 0;
  {
    std::lock_guard<std::mutex> lk(P.mu);
    auto it = P.host_owner.find(ptr);
    if (it == P.host_owner.end()) return hipErrorInvalidValue;
    sz = it->second;
  }
  const hipError_t e = hipDeviceSynchronize();  // copies in flight may still read it
  std::lock_guard<std::mutex> lk(P.mu);
  P.host_owner.erase(ptr);
  P.free_host[sz].push_back(const_cast<void*>(ptr));
  return e;
}

hipError_t device_tables(int device, DeviceTables* out) {
  static std::mutex mu;
  static std::map<int, DeviceTables> made;
  std::lock_guard<std::mutex> lk(mu);
  auto it = made.find(device);
  if (it == made.end()) {
    int cur = 0;
    if (hipError_t e = hipGetDevice(&cur)) return e;
    if (cur != device)
      if (hipError_t e = hipSetDevice(device)) return e;
    std::vector<float4> t32;
    std::vector<double> t64;
    host_tables(t32);
    host_tables64(t64);
    DeviceTables d;
    hipError_t e = hipMalloc(&d.t32, sizeof(float4) * t32.size());
    if (e == hipSuccess) e = hipMalloc(&d.t64, sizeof(double) * t64.size());
    if (e == hipSuccess) e = hipMemcpy(d.t32, t32.data(), sizeof(float4) * t32.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d.t64, t64.data(), sizeof(double) * t64.size(), hipMemcpyHostToDevice);
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess) {
      if (d.t32) (void)hipFree(d.t32);
      if (d.t64) (void)hipFree(d.t64);
      return e;
    }
    it = made.emplace(device, d).first;
  }
  *out = it->second;
  return hipSuccess;
}

}  // namespace a2c
