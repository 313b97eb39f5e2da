// Device-wide scan / radix sort / unique for gfx950 (wave64).
#include "prims.h"

#include <time.h>

#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <utility>
#include <algorithm>

namespace shci {

namespace {

bool allocTiming() {
  static const bool on = std::getenv("SHCI_ALLOC_TIMING") != nullptr;
  return on;
}

double monoSeconds() {
  timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

void reportSlow(const char* what, size_t bytes, double t0) {
  const double dt = monoSeconds() - t0;
  if (dt > 5e-3) std::fprintf(stderr, "  %s %.2f GiB took %.1f ms\n", what, bytes / 1073741824.0, 1e3 * dt);
}

}  // namespace

namespace {

// Idle physical chunks per device (process lifetime; the driver reclaims
// them at exit).
struct ChunkPool {
  std::mutex mu;
  std::vector<std::pair<int, hipMemGenericAllocationHandle_t>> idle;
};

ChunkPool& chunkPool() {
  static ChunkPool* pool = new ChunkPool;
  return *pool;
}

hipMemAllocationProp chunkProp(int device) {
  hipMemAllocationProp prop = {};
  prop.type = hipMemAllocationTypePinned;
  prop.location.type = hipMemLocationTypeDevice;
  prop.location.id = device;
  return prop;
}

size_t roundUp(size_t v, size_t m) { return (v + m - 1) / m * m; }

hipMemGenericAllocationHandle_t takeChunk(int device) {
  ChunkPool& pool = chunkPool();
  {
    std::lock_guard<std::mutex> g(pool.mu);
    for (size_t k = pool.idle.size(); k-- > 0;)
      if (pool.idle[k].first == device) {
        hipMemGenericAllocationHandle_t h = pool.idle[k].second;
        pool.idle.erase(pool.idle.begin() + (ptrdiff_t)k);
        return h;
      }
  }
  hipMemAllocationProp prop = chunkProp(device);
  hipMemGenericAllocationHandle_t h{};
  const double t0 = allocTiming() ? monoSeconds() : 0.0;
  hipError_t e = hipMemCreate(&h, kBigChunk, &prop, 0);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    trimChunkPool();
    HIP_CHECK(hipMemCreate(&h, kBigChunk, &prop, 0));
  }
  if (allocTiming()) reportSlow("hipMemCreate", kBigChunk, t0);
  return h;
}

void giveChunk(int device, hipMemGenericAllocationHandle_t h) {
  ChunkPool& pool = chunkPool();
  std::lock_guard<std::mutex> g(pool.mu);
  pool.idle.emplace_back(device, h);
}

// maps chunks [first, last) of `handles` at base + k * kBigChunk
void mapChunks(void* base, const std::vector<hipMemGenericAllocationHandle_t>& handles, size_t first, size_t last,
               int device) {
  hipMemAccessDesc acc = {};
  acc.location.type = hipMemLocationTypeDevice;
  acc.location.id = device;
  acc.flags = hipMemAccessFlagsProtReadWrite;
  for (size_t k = first; k < last; k++) {
    void* at = static_cast<char*>(base) + k * kBigChunk;
    HIP_CHECK(hipMemMap(at, kBigChunk, 0, handles[k], 0));
    HIP_CHECK(hipMemSetAccess(at, kBigChunk, &acc, 1));
  }
}

void unmapChunks(void* base, size_t count) {
  for (size_t k = 0; k < count; k++) HIP_CHECK(hipMemUnmap(static_cast<char*>(base) + k * kBigChunk, kBigChunk));
}

hipError_t mallocRetry(void** q, size_t bytes) {
  hipError_t e = hipMalloc(q, bytes);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    trimChunkPool();
    e = hipMalloc(q, bytes);
  }
  return e;
}

}  // namespace

void trimChunkPool() {
  ChunkPool& pool = chunkPool();
  std::lock_guard<std::mutex> g(pool.mu);
  for (auto& c : pool.idle) (void)hipMemRelease(c.second);
  pool.idle.clear();
}

size_t deviceFreeBytes(size_t* total) {
  size_t freeB = 0, totalB = 0;
  HIP_CHECK(hipMemGetInfo(&freeB, &totalB));
  int device = 0;
  HIP_CHECK(hipGetDevice(&device));
  ChunkPool& pool = chunkPool();
  std::lock_guard<std::mutex> g(pool.mu);
  for (auto& c : pool.idle)
    if (c.first == device) freeB += kBigChunk;
  if (total) *total = totalB;
  return freeB;
}

void DevMem::release() {
  // (runs in destructors: errors are not thrown from here)
  if (p && !chunks.empty()) {
    // kernels still queued may read the range: wait before unmapping it
    (void)hipDeviceSynchronize();
    bool unmapped = true;
    for (size_t k = 0; k < chunks.size(); k++)
      unmapped = hipMemUnmap(static_cast<char*>(p) + k * kBigChunk, kBigChunk) == hipSuccess && unmapped;
    (void)hipMemAddressFree(p, vaBytes);
    // a chunk that could not be unmapped is released rather than reused
    for (auto h : chunks) {
      if (unmapped)
        giveChunk(device, h);
      else
        (void)hipMemRelease(h);
    }
    chunks.clear();
  } else if (p) {
    const double t0 = allocTiming() ? monoSeconds() : 0.0;
    (void)hipFree(p);
    if (allocTiming()) reportSlow("hipFree", bytes, t0);
  }
  p = nullptr;
  bytes = 0;
  vaBytes = 0;
  device = -1;
}

void DevMem::swap(DevMem& o) {
  std::swap(p, o.p);
  std::swap(bytes, o.bytes);
  std::swap(vaBytes, o.vaBytes);
  std::swap(device, o.device);
  chunks.swap(o.chunks);
}

void DevMem::reserve(size_t nb, hipStream_t s, bool keep) {
  if (nb <= bytes) return;
  if (nb < kBigBuffer && chunks.empty()) {
    size_t cap = bytes ? bytes : 256;
    while (cap < nb) cap += cap / 2 + 256;
    void* q = nullptr;
    const double t0 = allocTiming() ? monoSeconds() : 0.0;
    HIP_CHECK(mallocRetry(&q, cap));
    if (allocTiming()) reportSlow("hipMalloc", cap, t0);
    if (keep && p && bytes) HIP_CHECK(hipMemcpyAsync(q, p, bytes, hipMemcpyDeviceToDevice, s));
    if (p) {
      if (keep) HIP_CHECK(hipStreamSynchronize(s));
      release();
    }
    p = q;
    bytes = cap;
    return;
  }
  const size_t need = roundUp(nb, kBigChunk);
  const size_t have = chunks.size();
  if (have == 0) {  // a hipMalloc'ed (or empty) buffer becomes chunk-mapped
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    const size_t va = roundUp(std::max(2 * need, size_t(4) << 30), kBigChunk);
    void* base = nullptr;
    HIP_CHECK(hipMemAddressReserve(&base, va, kBigChunk, nullptr, 0));
    std::vector<hipMemGenericAllocationHandle_t> hs;
    for (size_t k = 0; k < need / kBigChunk; k++) hs.push_back(takeChunk(dev));
    mapChunks(base, hs, 0, hs.size(), dev);
    if (keep && p && bytes) {
      HIP_CHECK(hipMemcpyAsync(base, p, bytes, hipMemcpyDeviceToDevice, s));
      HIP_CHECK(hipStreamSynchronize(s));
    }
    release();
    p = base;
    bytes = need;
    vaBytes = va;
    device = dev;
    chunks.swap(hs);
    return;
  }
  if (need > vaBytes) {  // a larger address range: remap the same chunks
    const size_t va = roundUp(2 * need, kBigChunk);
    void* base = nullptr;
    HIP_CHECK(hipMemAddressReserve(&base, va, kBigChunk, nullptr, 0));
    HIP_CHECK(hipDeviceSynchronize());  // queued work may still use the old range
    mapChunks(base, chunks, 0, have, device);
    unmapChunks(p, have);
    HIP_CHECK(hipMemAddressFree(p, vaBytes));
    p = base;
    vaBytes = va;
  }
  for (size_t k = have; k < need / kBigChunk; k++) chunks.push_back(takeChunk(device));
  mapChunks(p, chunks, have, chunks.size(), device);
  bytes = need;
}

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ int64_t waveInclusiveScan(int64_t v) {
  const int lane = __lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int64_t t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// block-wide exclusive scan of one value per thread; returns exclusive
// prefix, writes the block total to *total
__device__ int64_t blockExclusiveScan(int64_t v, int64_t* total) {
  __shared__ int64_t wsum[kScanThreads / 64];
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  int64_t inc = waveInclusiveScan(v);
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  int64_t off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / 64; w++) {
    int64_t x = wsum[w];
    if (w < wave) off += x;
    tot += x;
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

__global__ void __launch_bounds__(kScanThreads) kScanReduce(const int64_t* d, int64_t n, int64_t* partial) {
  int64_t base = (int64_t)blockIdx.x * kScanTile;
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    int64_t i = base + (int64_t)k * kScanThreads + threadIdx.x;
    if (i < n) s += d[i];
  }
  int64_t tot;
  (void)blockExclusiveScan(s, &tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kScanThreads) kScanDown(int64_t* d, int64_t n, const int64_t* partial) {
  int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int64_t v[kScanItems];
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    int64_t i = base + k;
    v[k] = i < n ? d[i] : 0;
    s += v[k];
  }
  int64_t tot;
  int64_t ex = blockExclusiveScan(s, &tot) + (partial ? partial[blockIdx.x] : 0);
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    int64_t i = base + k;
    if (i < n) d[i] = ex;
    ex += v[k];
  }
}

}  // namespace

void exclusiveScanI64(int64_t* d, int64_t n, Workspace& ws, hipStream_t s) {
  if (n <= 0) return;
  // level sizes: n, ceil(n/T), ceil(ceil(n/T)/T), ... until one tile
  int64_t sizes[16];
  int levels = 0;
  int64_t m = n;
  size_t need = 0;
  while (m > kScanTile && levels < 15) {
    m = (m + kScanTile - 1) / kScanTile;
    sizes[levels++] = m;
    need += (size_t)m;
  }
  ws.scan.reserve(need * sizeof(int64_t) + 64);
  int64_t* lev[16];
  int64_t* base = ws.scan.as<int64_t>();
  for (int l = 0; l < levels; l++) {
    lev[l] = base;
    base += sizes[l];
  }
  // up: reduce each level into the next
  const int64_t* src = d;
  int64_t srcN = n;
  for (int l = 0; l < levels; l++) {
    kScanReduce<<<(unsigned)sizes[l], kScanThreads, 0, s>>>(src, srcN, lev[l]);
    src = lev[l];
    srcN = sizes[l];
  }
  // top level fits one tile
  kScanDown<<<1, kScanThreads, 0, s>>>(levels ? lev[levels - 1] : d, levels ? sizes[levels - 1] : n, nullptr);
  // down: scan each level with the offsets of the level above
  for (int l = levels - 1; l >= 0; l--) {
    int64_t* tgt = l ? lev[l - 1] : d;
    int64_t tgtN = l ? sizes[l - 1] : n;
    kScanDown<<<(unsigned)sizes[l], kScanThreads, 0, s>>>(tgt, tgtN, lev[l]);
  }
}

// ---------------------------------------------------------------------------
// radix sort

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortRounds = 16;  // keys per thread per tile
constexpr int kSortTile = kSortThreads * kSortRounds;

template <int KW>
__device__ __forceinline__ uint32_t digitOf(const uint64_t* keys, int64_t i, int bit, bool signedOrder) {
  int w = bit >> 6;
  uint64_t x = keys[i * KW + w];
  if (signedOrder) x ^= 0x8000000000000000ull;
  return (uint32_t)(x >> (bit & 63)) & 0xFFu;
}

template <int KW>
__global__ void __launch_bounds__(kSortThreads) kRadixCount(const uint64_t* keys, int64_t n, int bit,
                                                            bool signedOrder, int64_t* hist, int64_t numTiles) {
  __shared__ int cnt[256];
  for (int d = threadIdx.x; d < 256; d += kSortThreads) cnt[d] = 0;
  __syncthreads();
  int64_t base = (int64_t)blockIdx.x * kSortTile;
  for (int r = 0; r < kSortRounds; r++) {
    int64_t i = base + (int64_t)r * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&cnt[digitOf<KW>(keys, i, bit, signedOrder)], 1);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < 256; d += kSortThreads) hist[(int64_t)d * numTiles + blockIdx.x] = cnt[d];
}

template <int KW, bool HASVAL>
__global__ void __launch_bounds__(kSortThreads) kRadixScatter(const uint64_t* kin, uint64_t* kout, const uint32_t* vin,
                                                              uint32_t* vout, int64_t n, int bit, bool signedOrder,
                                                              const int64_t* offs, int64_t numTiles) {
  __shared__ int64_t base[256];
  __shared__ int wcnt[kSortThreads / 64][256];
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  for (int d = threadIdx.x; d < 256; d += kSortThreads) base[d] = offs[(int64_t)d * numTiles + blockIdx.x];
  const int64_t tileBase = (int64_t)blockIdx.x * kSortTile;
  const uint64_t ltMask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  for (int r = 0; r < kSortRounds; r++) {
    for (int k = threadIdx.x; k < (kSortThreads / 64) * 256; k += kSortThreads) (&wcnt[0][0])[k] = 0;
    __syncthreads();
    int64_t i = tileBase + (int64_t)r * kSortThreads + threadIdx.x;
    bool valid = i < n;
    uint32_t dg = valid ? digitOf<KW>(kin, i, bit, signedOrder) : 0u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      uint64_t m = __ballot((dg >> b) & 1u);
      peers &= ((dg >> b) & 1u) ? m : ~m;
    }
    int rank = __popcll(peers & ltMask);
    if (valid && rank == 0) wcnt[wave][dg] = __popcll(peers);
    __syncthreads();
    int64_t pos = 0;
    if (valid) {
      pos = base[dg] + rank;
      for (int w = 0; w < wave; w++) pos += wcnt[w][dg];
    }
    __syncthreads();
    for (int d = threadIdx.x; d < 256; d += kSortThreads) {
      int s = 0;
#pragma unroll
      for (int w = 0; w < kSortThreads / 64; w++) s += wcnt[w][d];
      base[d] += s;
    }
    if (valid) {
#pragma unroll
      for (int w = 0; w < KW; w++) kout[pos * KW + w] = kin[i * KW + w];
      if (HASVAL) vout[pos] = vin[i];
    }
    __syncthreads();
  }
}

template <int KW>
void radixSortImpl(uint64_t* keys, uint32_t* vals, int64_t n, int nbits, bool signedOrder, Workspace& ws,
                   hipStream_t s) {
  if (n <= 1 || nbits <= 0) return;
  int64_t numTiles = (n + kSortTile - 1) / kSortTile;
  ws.a.reserve((size_t)n * KW * sizeof(uint64_t));
  if (vals) ws.b.reserve((size_t)n * sizeof(uint32_t));
  ws.c.reserve((size_t)(256 * numTiles + 1) * sizeof(int64_t));
  uint64_t* k0 = keys;
  uint64_t* k1 = ws.a.as<uint64_t>();
  uint32_t* v0 = vals;
  uint32_t* v1 = vals ? ws.b.as<uint32_t>() : nullptr;
  int64_t* hist = ws.c.as<int64_t>();
  int passes = 0;
  for (int bit = 0; bit < nbits; bit += 8, passes++) {
    kRadixCount<KW><<<(unsigned)numTiles, kSortThreads, 0, s>>>(k0, n, bit, signedOrder, hist, numTiles);
    exclusiveScanI64(hist, 256 * numTiles, ws, s);
    if (vals)
      kRadixScatter<KW, true><<<(unsigned)numTiles, kSortThreads, 0, s>>>(k0, k1, v0, v1, n, bit, signedOrder, hist,
                                                                         numTiles);
    else
      kRadixScatter<KW, false><<<(unsigned)numTiles, kSortThreads, 0, s>>>(k0, k1, nullptr, nullptr, n, bit,
                                                                          signedOrder, hist, numTiles);
    std::swap(k0, k1);
    std::swap(v0, v1);
  }
  if (passes & 1) {
    HIP_CHECK(hipMemcpyAsync(keys, k0, (size_t)n * KW * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
    if (vals) HIP_CHECK(hipMemcpyAsync(vals, v0, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  }
}

template <int KW>
__global__ void kUniqueFlags(const uint64_t* keys, int64_t n, int64_t* flags) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  if (i == n) {
    flags[n] = 0;
    return;
  }
  bool first = (i == 0);
  if (!first) {
    first = false;
#pragma unroll
    for (int w = 0; w < KW; w++)
      if (keys[i * KW + w] != keys[(i - 1) * KW + w]) first = true;
  }
  flags[i] = first ? 1 : 0;
}

template <int KW>
__global__ void kUniqueScatter(const uint64_t* keys, int64_t n, const int64_t* pos, uint64_t* out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (pos[i + 1] != pos[i]) {
#pragma unroll
    for (int w = 0; w < KW; w++) out[pos[i] * KW + w] = keys[i * KW + w];
  }
}

template <int KW>
int64_t uniqueImpl(uint64_t* keys, int64_t n, Workspace& ws, hipStream_t s) {
  if (n <= 1) return n;
  ws.d.reserve((size_t)(n + 1) * sizeof(int64_t));
  int64_t* flags = ws.d.as<int64_t>();
  unsigned g = (unsigned)((n + 1 + 255) / 256);
  kUniqueFlags<KW><<<g, 256, 0, s>>>(keys, n, flags);
  exclusiveScanI64(flags, n + 1, ws, s);
  ws.a.reserve((size_t)n * KW * sizeof(uint64_t));
  kUniqueScatter<KW><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(keys, n, flags, ws.a.as<uint64_t>());
  int64_t m = 0;
  HIP_CHECK(hipMemcpyAsync(&m, flags + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  HIP_CHECK(hipMemcpyAsync(keys, ws.a.p, (size_t)m * KW * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
  return m;
}

}  // namespace

void radixSortKeys(uint64_t* keys, uint32_t* vals, int64_t n, int KW, int nbits, bool signedOrder, Workspace& ws,
                   hipStream_t s) {
  if (KW == 1)
    radixSortImpl<1>(keys, vals, n, nbits, signedOrder, ws, s);
  else if (KW == 2)
    radixSortImpl<2>(keys, vals, n, nbits, signedOrder, ws, s);
  else
    throw HipError("radixSortKeys: unsupported key width");
}

int64_t uniqueKeys(uint64_t* keys, int64_t n, int KW, Workspace& ws, hipStream_t s) {
  if (KW == 1) return uniqueImpl<1>(keys, n, ws, s);
  if (KW == 2) return uniqueImpl<2>(keys, n, ws, s);
  throw HipError("uniqueKeys: unsupported key width");
}

}  // namespace shci
