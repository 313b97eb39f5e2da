// Device-wide primitives used by the SHCI kernels: scan, radix sort,
// unique.  All calls are stream-ordered; scratch comes from a caller-owned
// Workspace so nothing is allocated per call once buffers have grown.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>

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

// Growable device buffer (bytes).  Grows by reallocating; contents are not
// preserved unless `keep` is set.
struct DevMem {
  void* p = nullptr;
  size_t bytes = 0;
  DevMem() = default;
  DevMem(const DevMem&) = delete;
  DevMem& operator=(const DevMem&) = delete;
  ~DevMem() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void reserve(size_t nb, hipStream_t s = 0, bool keep = false) {
    if (nb <= bytes) return;
    size_t cap = bytes ? bytes : 256;
    while (cap < nb) cap += cap / 2 + 256;
    void* q = nullptr;
    HIP_CHECK(hipMalloc(&q, cap));
    if (keep && p && bytes) HIP_CHECK(hipMemcpyAsync(q, p, bytes, hipMemcpyDeviceToDevice, s));
    if (p) {
      if (keep) HIP_CHECK(hipStreamSynchronize(s));
      HIP_CHECK(hipFree(p));
    }
    p = q;
    bytes = cap;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

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
