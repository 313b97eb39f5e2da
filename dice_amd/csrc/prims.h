// Device-wide primitives used by the SHCI kernels: scan, radix sort,
// unique.  All calls are stream-ordered; scratch comes from a caller-owned
// Workspace so nothing is allocated per call once buffers have grown.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>
#include <vector>

namespace shci {

struct HipError : std::runtime_error {
  explicit HipError(const std::string& m) : std::runtime_error(m) {}
};

#define HIP_CHECK(expr)                                                                   \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      throw ::shci::HipError(std::string(#expr) + ": " + hipGetErrorString(e_) + " (" +   \
                             __FILE__ + ":" + std::to_string(__LINE__) + ")");            \
  } while (0)

// Growable device buffer (bytes).  Buffers below kBigBuffer come from
// hipMalloc and grow by reallocating (contents kept only with `keep`).
// Larger buffers are a reserved virtual address range onto which physical
// chunks of kBigChunk bytes are mapped as the buffer grows: growth maps
// more chunks at the end and keeps the contents in place, and the chunks of
// a released buffer go to a per-process pool that later buffers map again,
// so the multi-gigabyte Hamiltonian buffers are neither copied nor freed
// between iterations (freeing and reallocating them costs 30 ms per GB or
// more on this part).  Idle chunks are returned to the driver when an
// allocation fails, then the allocation is retried.  With SHCI_ALLOC_TIMING
// set, every driver allocation or free that takes longer than 5 ms is
// reported on stderr.
constexpr size_t kBigChunk = size_t(256) << 20;
constexpr size_t kBigBuffer = kBigChunk;

struct DevMem {
  void* p = nullptr;
  size_t bytes = 0;
  // chunk-mapped buffers: reserved address range and mapped chunks
  size_t vaBytes = 0;
  int device = -1;
  std::vector<hipMemGenericAllocationHandle_t> chunks;
  DevMem() = default;
  DevMem(const DevMem&) = delete;
  DevMem& operator=(const DevMem&) = delete;
  ~DevMem() { release(); }
  // waits for the device before unmapping a chunk-mapped buffer
  void release();
  // after reserve(nb), at least nb bytes are addressable from p; the address
  // may change (callers that captured it must refresh)
  void reserve(size_t nb, hipStream_t s = 0, bool keep = false);
  void swap(DevMem& o);
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// free device memory as hipMemGetInfo reports it plus the idle pooled chunks
// of the current device (which any buffer can map without a driver call)
size_t deviceFreeBytes(size_t* total = nullptr);
// returns the idle pooled chunks of every device to the driver
void trimChunkPool();

// scratch for the primitives
struct Workspace {
  DevMem a, b, c, d;
  DevMem scan;  // partial sums of the multi-level scan
};

// exclusive scan of n int64 values in place; returns nothing, total can be
// read at index n if the buffer holds n+1 entries (the caller sets d[n]=0
// before the call and the scan includes it).
void exclusiveScanI64(int64_t* d, int64_t n, Workspace& ws, hipStream_t s);

// Stable LSD radix sort of `n` keys of KW 64-bit words each, over key bits
// [0, nbits) of the composite key (word KW-1 most significant).  With
// `signedOrder`, every word is compared as a signed int64 (the determinant
// order of the reference).  Optional 32-bit payload.  Results land back in
// keys / vals.
void radixSortKeys(uint64_t* keys, uint32_t* vals, int64_t n, int KW, int nbits, bool signedOrder,
                   Workspace& ws, hipStream_t s);

// In-place unique of a sorted array of KW-word keys; returns the new count
// (synchronises the stream to read it).
int64_t uniqueKeys(uint64_t* keys, int64_t n, int KW, Workspace& ws, hipStream_t s);

}  // namespace shci
