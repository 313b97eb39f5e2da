// SHCI variational-step kernels for CDNA4 (gfx950, wave64).
//
//   energy / Hij batches      SHCI/Determinants.cpp:252-300,406-505,564-631
//   heat-bath screening       SHCI/SHCIgetdeterminants.cpp:586-739
//   connected-pair CSR        SHCI/SHCImakeHamiltonian.cpp:486-715 (row content)
//   sigma = H x               SHCI/Hmult.h:52-150
//   Davidson vector algebra   SHCI/Davidson.cpp:85-348
#include <algorithm>
#include <type_traits>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <string>

#include "kernels.h"

namespace shci {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

inline unsigned gridFor(int64_t n, int per = kBlock, int64_t cap = 1 << 20) {
  int64_t g = (n + per - 1) / per;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

// ---------------------------------------------------------------------------
// determinant primitives

template <int W>
__global__ void kEnergy(DevSystem S, const uint64_t* dets, int64_t n, double* out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t d[W];
#pragma unroll
    for (int w = 0; w < W; w++) d[w] = dets[i * W + w];
    out[i] = detEnergy<W>(S, d);
  }
}

template <int W>
__global__ void kHijPairs(DevSystem S, const uint64_t* bra, const uint64_t* ket, int64_t n, double* out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t b[W], k[W];
#pragma unroll
    for (int w = 0; w < W; w++) {
      b[w] = bra[i * W + w];
      k[w] = ket[i * W + w];
    }
    out[i] = hijElement<W>(S, b, k);
  }
}

// ---------------------------------------------------------------------------
// membership table

__global__ void kFillNeg(int32_t* t, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    t[i] = -1;
}

template <int W>
__device__ __forceinline__ bool detEqual(const uint64_t* a, const uint64_t* b) {
  bool e = true;
#pragma unroll
  for (int w = 0; w < W; w++) e &= (a[w] == b[w]);
  return e;
}

template <int W>
__global__ void kHashInsert(const uint64_t* dets, int64_t begin, int64_t end, int32_t* table, uint64_t mask) {
  for (int64_t i = begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < end;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t d[W];
#pragma unroll
    for (int w = 0; w < W; w++) d[w] = dets[i * W + w];
    uint64_t h = detHash<W>(d) & mask;
    while (true) {
      int32_t prev = atomicCAS(&table[h], -1, (int32_t)i);
      if (prev == -1) break;
      if (detEqual<W>(dets + (int64_t)prev * W, d)) break;
      h = (h + 1) & mask;
    }
  }
}

template <int W>
__device__ __forceinline__ bool hashContains(const int32_t* table, uint64_t mask, const uint64_t* dets,
                                             const uint64_t* key) {
  uint64_t h = detHash<W>(key) & mask;
  while (true) {
    int32_t v = table[h];
    if (v < 0) return false;
    if (detEqual<W>(dets + (int64_t)v * W, key)) return true;
    h = (h + 1) & mask;
  }
}

// ---------------------------------------------------------------------------
// screening

__global__ void kCmax(const double* ci, int nroots, int64_t ld, int64_t n, double* cmax) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int r = 0; r < nroots; r++) {
      double c = fabs(ci[r * ld + j]);
      s = __dadd_rn(s, __dmul_rn(c, c));
    }
    cmax[j] = sqrt(s);
  }
}

// emit one candidate per lane with `want`, skipping determinants already in
// the variational space; wave-aggregated append
template <int W>
__device__ __forceinline__ void emitCandidate(bool want, const uint64_t* candIn, const int32_t* table, uint64_t mask,
                                              const uint64_t* dets, uint64_t* out, unsigned long long* counter,
                                              int64_t cap, bool trev = false) {
  // with time reversal the space holds standard forms (SHCI/SHCIbasics.cpp:984-996)
  uint64_t cand[W];
#pragma unroll
  for (int w = 0; w < W; w++) cand[w] = candIn[w];
  if (trev) makeStandardForm<W>(cand);
  bool keep = want && !hashContains<W>(table, mask, dets, cand);
  uint64_t m = __ballot(keep);
  if (m == 0) return;
  const int lane = __lane_id();
  unsigned long long base = 0;
  int leader = __ffsll((unsigned long long)m) - 1;
  if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(m));
  base = __shfl(base, leader, 64);
  if (keep) {
    unsigned long long pos = base + __popcll(m & lanemaskLt());
    if ((int64_t)pos < cap) {
#pragma unroll
      for (int w = 0; w < W; w++) out[pos * W + w] = cand[w];
    }
  }
}

__device__ __forceinline__ void emitRecord(bool want, const CandRecord& r, CandRecord* out,
                                           unsigned long long* counter, int64_t cap) {
  uint64_t m = __ballot(want);
  if (m == 0) return;
  const int lane = __lane_id();
  unsigned long long base = 0;
  int leader = __ffsll((unsigned long long)m) - 1;
  if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(m));
  base = __shfl(base, leader, 64);
  if (want) {
    unsigned long long pos = base + __popcll(m & lanemaskLt());
    if ((int64_t)pos < cap) out[pos] = r;
  }
}

// MODE 0: singles and doubles emitted as determinants (the reference's
// screening).  MODE 1: singles only.  MODE 2: no singles; every double
// excitation whose heat-bath value passes eps / loosen (a superset of the
// accepted ones) is emitted as a CandRecord for the learned ranking.
template <int W, int MODE>
__global__ void __launch_bounds__(kBlock) kScreen(DevSystem S, const uint64_t* dets, int64_t srcBegin,
                                                  int64_t ndets, const double* cmax, double eps1, int ncore,
                                                  int nact, const int32_t* table, uint64_t mask, uint64_t* out,
                                                  unsigned long long* counter, int64_t cap, double loosen,
                                                  CandRecord* rec) {
  __shared__ uint8_t sClosed[kWaves][128];
  __shared__ uint8_t sOpen[kWaves][128];
  __shared__ int sExcl[kWaves][65];
  __shared__ int64_t sStart[kWaves][64];
  __shared__ int sPair[kWaves][64];
  const int wave = threadIdx.x >> 6, lane = __lane_id();
  const int nso = S.nso;
  for (int64_t di = srcBegin + (int64_t)blockIdx.x * kWaves + wave; di < ndets; di += (int64_t)gridDim.x * kWaves) {
    uint64_t d[W];
#pragma unroll
    for (int w = 0; w < W; w++) d[w] = dets[di * W + w];
    const double eps = eps1 / fabs(cmax[di]);
    // ascending closed / open spin-orbital lists
    int nclosed = 0, nopen = 0;
    for (int base = 0; base < nso; base += 64) {
      int k = base + lane;
      bool valid = k < nso;
      bool o = valid && occ<W>(d, k);
      uint64_t mc = __ballot(o), mo = __ballot(valid && !o);
      if (o) sClosed[wave][nclosed + __popcll(mc & lanemaskLt())] = (uint8_t)k;
      if (valid && !o) sOpen[wave][nopen + __popcll(mo & lanemaskLt())] = (uint8_t)k;
      nclosed += __popcll(mc);
      nopen += __popcll(mo);
    }
    __syncwarp();
    int cur[SHCI_MAX_RESTRICT];
    for (int r = 0; r < S.nrestrict; r++) {
      int c = 0;
      for (int k = 0; k < nclosed; k++) c += S.rmult[r * nso + sClosed[wave][k]];
      cur[r] = c;
    }
    // ---- single excitations
    const int nsingle = MODE == 2 ? 0 : nclosed * nopen;
    for (int t0 = 0; t0 < nsingle; t0 += 64) {
      int t = t0 + lane;
      bool want = false;
      uint64_t cand[W];
#pragma unroll
      for (int w = 0; w < W; w++) cand[w] = d[w];
      if (t < nsingle) {
        int i = t / nopen, a = t - (t / nopen) * nopen;
        int ci = sClosed[wave][i], oa = sOpen[wave][a];
        bool ok = !((ci >> 1) < ncore || (oa >> 1) >= ncore + nact);
        for (int r = 0; r < S.nrestrict && ok; r++) {
          int e = cur[r] - S.rmult[r * nso + ci] + S.rmult[r * nso + oa];
          if (e < S.rmin[r] || e > S.rmax[r]) ok = false;
        }
        if (ok) {
          double integral = S.singles[oa * nso + ci];
          if (fabs(integral) > eps)
            if (((ci ^ oa) & 1) == 0) integral = hijSingle<W>(S, d, oa, ci);
          if (fabs(integral) > eps) {
            want = true;
            cand[oa >> 6] |= (1ull << (oa & 63));
            cand[ci >> 6] &= ~(1ull << (ci & 63));
          }
        }
      }
      emitCandidate<W>(want, cand, table, mask, dets, out, counter, cap, S.trev);
    }
    // ---- double excitations
    if (MODE == 1) continue;
    const double epsD = MODE == 2 ? eps / loosen : eps;
    if (S.maxEntryAbs < epsD) continue;
    const int npairs = nclosed * (nclosed - 1) / 2;
    for (int p0 = 0; p0 < npairs; p0 += 64) {
      int p = p0 + lane;
      int len = 0;
      int64_t st = 0;
      if (p < npairs) {
        // p = i*(i-1)/2 + j with i > j
        int i = (int)((1.0 + sqrt(1.0 + 8.0 * (double)p)) * 0.5);
        while (i * (i - 1) / 2 > p) i--;
        while ((i + 1) * i / 2 <= p) i++;
        int j = p - i * (i - 1) / 2;
        int ci = sClosed[wave][i], cj = sClosed[wave][j];
        if (!((ci >> 1) < ncore || (cj >> 1) < ncore)) {
          int I = ci >> 1, J = cj >> 1;
          int X = I > J ? I : J, Y = I > J ? J : I;
          int pairIndex = X * (X + 1) / 2 + Y;
          int sameSpin = ((ci ^ cj) & 1) == 0 ? 1 : 0;
          int64_t s0 = S.hbStart[sameSpin][pairIndex], s1 = S.hbStart[sameSpin][pairIndex + 1];
          const float* v = S.hbVal[sameSpin];
          // number of leading entries with |v| >= eps (list is |v|-descending)
          int64_t lo = s0, hi = s1;
          while (lo < hi) {
            int64_t mid = (lo + hi) >> 1;
            if ((double)fabsf(v[mid]) < epsD)
              hi = mid;
            else
              lo = mid + 1;
          }
          len = (int)(lo - s0);
          st = s0;
        }
        sPair[wave][lane] = i | (j << 8) | ((((sClosed[wave][i] ^ sClosed[wave][j]) & 1) == 0 ? 1 : 0) << 16);
      }
      // exclusive prefix of len over lanes
      int inc = len;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
      }
      sExcl[wave][lane] = inc - len;
      sStart[wave][lane] = st;
      int total = __shfl(inc, 63, 64);
      __syncwarp();
      for (int k0 = 0; k0 < total; k0 += 64) {
        int k = k0 + lane;
        bool want = false;
        uint64_t cand[W];
        CandRecord recTmp{};
#pragma unroll
        for (int w = 0; w < W; w++) cand[w] = d[w];
        if (k < total) {
          // owner lane q: last q with sExcl[q] <= k
          int lo = 0, hi = 63;
          while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (sExcl[wave][mid] <= k)
              lo = mid;
            else
              hi = mid - 1;
          }
          int q = lo;
          int pk = sPair[wave][q];
          int i = pk & 0xFF, j = (pk >> 8) & 0xFF, same = (pk >> 16) & 1;
          int ci = sClosed[wave][i], cj = sClosed[wave][j];
          int64_t idx = sStart[wave][q] + (k - sExcl[wave][q]);
          uint32_t ab = S.hbPair[same][idx];
          int a = 2 * (int)(ab & 0xFFFF) + (ci & 1), b = 2 * (int)(ab >> 16) + (cj & 1);
          bool ok = !((a >> 1) >= ncore + nact || (b >> 1) >= ncore + nact);
          for (int r = 0; r < S.nrestrict && ok; r++) {
            const int8_t* m = S.rmult + r * nso;
            int e = cur[r] - m[ci] - m[cj] + m[a] + m[b];
            if (e < S.rmin[r] || e > S.rmax[r]) ok = false;
          }
          if (ok && !(occ<W>(d, a) || occ<W>(d, b))) {
            want = true;
            if (MODE == 2) {
              recTmp.src = (uint32_t)di;
              recTmp.ci = (uint8_t)ci;
              recTmp.cj = (uint8_t)cj;
              recTmp.a = (uint8_t)a;
              recTmp.b = (uint8_t)b;
              recTmp.v = S.hbVal[same][idx];
            }
            cand[a >> 6] |= (1ull << (a & 63));
            cand[b >> 6] |= (1ull << (b & 63));
            cand[ci >> 6] &= ~(1ull << (ci & 63));
            cand[cj >> 6] &= ~(1ull << (cj & 63));
          }
        }
        if (MODE == 2)
          emitRecord(want, recTmp, rec, counter, cap);
        else
          emitCandidate<W>(want, cand, table, mask, dets, out, counter, cap, S.trev);
      }
      __syncwarp();
    }
  }
}

// ---------------------------------------------------------------------------
// deterministic EN-PT2 (SHCI/SHCIbasics.cpp:453-728 with the candidate rules of
// SHCI/SHCIgetdeterminants.cpp:40-171)

// partition of a candidate among nparts passes (hash bits above the ones the
// membership table uses)
template <int W>
__device__ __forceinline__ uint32_t ptPart(const uint64_t* d, uint32_t nparts) {
  return nparts == 1 ? 0u : (uint32_t)((detHash<W>(d) >> 40) % nparts);
}

// Count pass: per-lane counters.  Fill pass: wave-uniform cursor with ballot
// positions, so a source's candidates land in the reference's generation
// order (singles by (closed i, open a), then doubles by (i > j) pair and
// heat-bath order).
template <int W, bool FILL>
struct PtSink {
  int64_t pos;
  int cnt;
  uint64_t* outD;
  double* outN;
  double* outN2;     // semistochastic: second-moment numerator
  uint8_t* outP;     // semistochastic: term passes epsilon2Large
  __device__ __forceinline__ void put(bool keep, const uint64_t* cand, double num, double num2 = 0.0,
                                      bool present = false) {
    if (!FILL) {
      cnt += keep ? 1 : 0;
      return;
    }
    const uint64_t m = __ballot(keep);
    if (keep) {
      const int64_t p = pos + __popcll(m & lanemaskLt());
#pragma unroll
      for (int w = 0; w < W; w++) outD[p * W + w] = cand[w];
      outN[p] = num;
      if (outN2) {
        outN2[p] = num2;
        outP[p] = present ? 1 : 0;
      }
    }
    pos += __popcll(m);
  }
};

// Semistochastic mode (SHCI/SHCIgetdeterminants.cpp:889-1046): source k is
// determinant src[k] with sample weight w[k]; num = H_ai w, num2 = H_ai^2 w
// (w Nmc / (Nmc - 1) - c) with a double's H_ai^2 formed in float as there,
// present = |H_ai| > epsilon2Large / |c|.
struct PtStoch {
  const uint32_t* src;
  const double* w;
  double eps2Large;
  double nmc;
  double* outN2;
  uint8_t* outP;
};

// Candidates D_a of sources [0, ndets) with numerator H_ai c_i: singles with
// |Hij_1Excite| > eps, doubles while the float heat-bath value passes eps
// (numerator = value x parity of the excitation x c_i), eps = |eps2 / c_i|,
// no active-space limits, the PT restrictions in S; kept when D_a is not a
// variational determinant and falls in partition `part`.
template <int W, bool FILL, bool STOCH>
__global__ void __launch_bounds__(kBlock) kPtGen(DevSystem S, const uint64_t* dets, int64_t ndets, const double* ci,
                                                 double eps2, const int32_t* table, uint64_t mask, uint32_t nparts,
                                                 uint32_t part, int64_t* counts, const int64_t* fillStart,
                                                 uint64_t* outD, double* outN, PtStoch st) {
  __shared__ uint8_t sClosed[kWaves][128];
  __shared__ uint8_t sOpen[kWaves][128];
  __shared__ int sExcl[kWaves][65];
  __shared__ int64_t sStart[kWaves][64];
  __shared__ int sPair[kWaves][64];
  const int wave = threadIdx.x >> 6, lane = __lane_id();
  const int nso = S.nso;
  for (int64_t di = (int64_t)blockIdx.x * kWaves + wave; di < ndets; di += (int64_t)gridDim.x * kWaves) {
    const int64_t src = STOCH ? (int64_t)st.src[di] : di;
    uint64_t d[W];
#pragma unroll
    for (int w = 0; w < W; w++) d[w] = dets[src * W + w];
    const double c = ci[src];
    const double wt = STOCH ? st.w[di] : c;
    const double eps = STOCH ? __ddiv_rn(eps2, fabs(c)) : fabs(eps2 / c);
    const double epsL = STOCH ? __ddiv_rn(st.eps2Large, fabs(c)) : 0.0;
    const double wfac = STOCH ? __dsub_rn(__ddiv_rn(__dmul_rn(wt, st.nmc), st.nmc - 1.0), c) : 0.0;
    PtSink<W, FILL> sink;
    sink.pos = FILL ? fillStart[di] : 0;
    sink.cnt = 0;
    sink.outD = outD;
    sink.outN = outN;
    sink.outN2 = STOCH ? st.outN2 : nullptr;
    sink.outP = STOCH ? st.outP : nullptr;
    int nclosed = 0, nopen = 0;
    for (int base = 0; base < nso; base += 64) {
      int k = base + lane;
      bool valid = k < nso;
      bool o = valid && occ<W>(d, k);
      uint64_t mc = __ballot(o), mo = __ballot(valid && !o);
      if (o) sClosed[wave][nclosed + __popcll(mc & lanemaskLt())] = (uint8_t)k;
      if (valid && !o) sOpen[wave][nopen + __popcll(mo & lanemaskLt())] = (uint8_t)k;
      nclosed += __popcll(mc);
      nopen += __popcll(mo);
    }
    __syncwarp();
    int cur[SHCI_MAX_RESTRICT];
    for (int r = 0; r < S.nrestrict; r++) {
      int cc = 0;
      for (int k = 0; k < nclosed; k++) cc += S.rmult[r * nso + sClosed[wave][k]];
      cur[r] = cc;
    }
    // ---- single excitations
    const int nsingle = nclosed * nopen;
    for (int t0 = 0; t0 < nsingle; t0 += 64) {
      int t = t0 + lane;
      bool keep = false, pres = false;
      double num = 0.0, num2 = 0.0;
      uint64_t cand[W];
#pragma unroll
      for (int w = 0; w < W; w++) cand[w] = d[w];
      if (t < nsingle) {
        int i = t / nopen, a = t - (t / nopen) * nopen;
        int cI = sClosed[wave][i], oa = sOpen[wave][a];
        bool ok = true;
        for (int r = 0; r < S.nrestrict && ok; r++) {
          int e = cur[r] - S.rmult[r * nso + cI] + S.rmult[r * nso + oa];
          if (e < S.rmin[r] || e > S.rmax[r]) ok = false;
        }
        // (a spin-changing single has only the one-electron term, which is 0)
        const double integral = (ok && ((cI ^ oa) & 1) == 0) ? hijSingle<W>(S, d, oa, cI) : 0.0;
        if (ok && fabs(integral) > eps) {
          cand[oa >> 6] |= (1ull << (oa & 63));
          cand[cI >> 6] &= ~(1ull << (cI & 63));
          num = __dmul_rn(integral, wt);
          if (STOCH) {
            num2 = __dmul_rn(__dmul_rn(__dmul_rn(integral, integral), wt), wfac);
            pres = fabs(integral) > epsL;
          }
          keep = ptPart<W>(cand, nparts) == part && !hashContains<W>(table, mask, dets, cand);
        }
      }
      sink.put(keep, cand, num, num2, pres);
    }
    // ---- double excitations
    if (!(S.maxEntryAbs < eps)) {
      const int npairs = nclosed * (nclosed - 1) / 2;
      for (int p0 = 0; p0 < npairs; p0 += 64) {
        int p = p0 + lane;
        int len = 0;
        int64_t st = 0;
        if (p < npairs) {
          int i = (int)((1.0 + sqrt(1.0 + 8.0 * (double)p)) * 0.5);
          while (i * (i - 1) / 2 > p) i--;
          while ((i + 1) * i / 2 <= p) i++;
          int j = p - i * (i - 1) / 2;
          int cI = sClosed[wave][i], cJ = sClosed[wave][j];
          int I = cI >> 1, J = cJ >> 1;
          int X = I > J ? I : J, Y = I > J ? J : I;
          int pairIndex = X * (X + 1) / 2 + Y;
          int sameSpin = ((cI ^ cJ) & 1) == 0 ? 1 : 0;
          int64_t s0 = S.hbStart[sameSpin][pairIndex], s1 = S.hbStart[sameSpin][pairIndex + 1];
          const float* v = S.hbVal[sameSpin];
          int64_t lo = s0, hi = s1;
          while (lo < hi) {
            int64_t mid = (lo + hi) >> 1;
            if ((double)fabsf(v[mid]) < eps)
              hi = mid;
            else
              lo = mid + 1;
          }
          len = (int)(lo - s0);
          st = s0;
          sPair[wave][lane] = i | (j << 8) | (sameSpin << 16);
        }
        int inc = len;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          int tt = __shfl_up(inc, o, 64);
          if (lane >= o) inc += tt;
        }
        sExcl[wave][lane] = inc - len;
        sStart[wave][lane] = st;
        int total = __shfl(inc, 63, 64);
        __syncwarp();
        for (int k0 = 0; k0 < total; k0 += 64) {
          int k = k0 + lane;
          bool keep = false, pres = false;
          double num = 0.0, num2 = 0.0;
          uint64_t cand[W];
#pragma unroll
          for (int w = 0; w < W; w++) cand[w] = d[w];
          if (k < total) {
            int lo = 0, hi = 63;
            while (lo < hi) {
              int mid = (lo + hi + 1) >> 1;
              if (sExcl[wave][mid] <= k)
                lo = mid;
              else
                hi = mid - 1;
            }
            int q = lo;
            int pk = sPair[wave][q];
            int i = pk & 0xFF, j = (pk >> 8) & 0xFF, same = (pk >> 16) & 1;
            int cI = sClosed[wave][i], cJ = sClosed[wave][j];
            int64_t idx = sStart[wave][q] + (k - sExcl[wave][q]);
            uint32_t ab = S.hbPair[same][idx];
            int a = 2 * (int)(ab & 0xFFFF) + (cI & 1), b = 2 * (int)(ab >> 16) + (cJ & 1);
            bool ok = true;
            for (int r = 0; r < S.nrestrict && ok; r++) {
              const int8_t* m = S.rmult + r * nso;
              int e = cur[r] - m[cI] - m[cJ] + m[a] + m[b];
              if (e < S.rmin[r] || e > S.rmax[r]) ok = false;
            }
            if (ok && !(occ<W>(d, a) || occ<W>(d, b))) {
              cand[a >> 6] |= (1ull << (a & 63));
              cand[b >> 6] |= (1ull << (b & 63));
              cand[cI >> 6] &= ~(1ull << (cI & 63));
              cand[cJ >> 6] &= ~(1ull << (cJ & 63));
              // Determinant::parity(a, b, cI, cJ) on the new determinant
              // (SHCI/Determinants.cpp:332-357): a -> cI, then b -> cJ
              int nb = countBetween<W>(cand, a < cI ? a : cI, a < cI ? cI : a);
              uint64_t mid2[W];
#pragma unroll
              for (int w = 0; w < W; w++) mid2[w] = cand[w];
              mid2[a >> 6] &= ~(1ull << (a & 63));
              mid2[cI >> 6] |= (1ull << (cI & 63));
              nb += countBetween<W>(mid2, b < cJ ? b : cJ, b < cJ ? cJ : b);
              const double sgn = (nb & 1) ? -1.0 : 1.0;
              const float hv = S.hbVal[same][idx];
              num = __dmul_rn(__dmul_rn((double)hv, sgn), wt);
              if (STOCH) {
                num2 = __dmul_rn(__dmul_rn((double)__fmul_rn(hv, hv), wt), wfac);
                pres = (double)fabsf(hv) > epsL;
              }
              keep = ptPart<W>(cand, nparts) == part && !hashContains<W>(table, mask, dets, cand);
            }
          }
          sink.put(keep, cand, num, num2, pres);
        }
        __syncwarp();
      }
    }
    if (!FILL) {
      int cc = sink.cnt;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) cc += __shfl_xor(cc, o, 64);
      if (lane == 0) counts[di] = cc;
    }
  }
}

// ---------------------------------------------------------------------------
// RDMs of <bra| ... |ket> (the quantities of SHCI/SHCIrdm.cpp:863-1100):
//   g1[P*n + R]                 = sum_s   <a+_{P s} a_{R s}>
//   g2[((P*n + Q)*n + R)*n + S] = sum_s,t <a+_{P s} a+_{Q t} a_{S t} a_{R s}>
// and optionally the spin-orbital ones
//   s1[p*nso + r]               = <a+_p a_r>
//   s2[pair(p,q)*npair + pair(r,s)] = <a+_p a+_q a_s a_r>, p > q, r > s,
//   pair(p,q) = p(p+1)/2 + q, npair = nso(nso+1)/2
// Off-diagonal pairs come from the J < I entries of the full symmetric CSR
// (an element with bra[I] ket[J] and its transpose with bra[J] ket[I]); the
// diagonal is a rank-n product of the occupation matrix.
//
// g2 is assembled from three accumulators, so that each connection costs
// few FP64 atomics:
//   A[P,Q,R,S]  one element of every exchange pair G[PQRS] = G[QPSR] (and,
//               when bra = ket, of every hermitian pair G[PQRS] = G[RSPQ]);
//               G = A + A[QPSR] (+ the hermitian images)
//   Z[X,Y,K]    single excitations X <- Y times the number of common
//               occupied spin-orbitals of spatial orbital K; it adds to
//               G[X,K,Y,K] and G[K,X,K,Y]
//   Zs[X,Y,K]   the same over the common spin-orbitals with the spin of the
//               excitation; it subtracts from G[X,K,K,Y] and G[K,X,Y,K]

// parity of the occupied spin-orbitals of d below k
template <int W>
__device__ __forceinline__ int belowParity(const uint64_t* d, int k) {
  int c = 0;
#pragma unroll
  for (int w = 0; w < W; w++) {
    const int lo = 64 * w;
    if (k >= lo + 64)
      c += __popcll(d[w]);
    else if (k > lo)
      c += __popcll(d[w] & ((1ull << (k - lo)) - 1ull));
  }
  return c & 1;
}

struct RdmOut {
  double *g1, *g2, *s1, *s2;  // s1 / s2 may be null
  double *A, *Z, *Zs;         // accumulators of g2
  int n, nso;
  int64_t npair;
  bool herm;  // bra == ket: only one element of each hermitian pair is accumulated
  __device__ __forceinline__ void acc(int P, int Q, int R, int S, double v) const {
    atomicAdd(A + (((int64_t)P * n + Q) * n + R) * n + S, v);
  }
  __device__ __forceinline__ void add2(int P, int Q, int R, int S, double v) const {
    atomicAdd(g2 + (((int64_t)P * n + Q) * n + R) * n + S, v);
  }
  __device__ __forceinline__ void spin2(int p, int q, int r, int s, double v) const {
    // p > q, r > s
    atomicAdd(s2 + ((int64_t)p * (p + 1) / 2 + q) * npair + ((int64_t)r * (r + 1) / 2 + s), v);
  }
};

// diagonal part: M[p*nso + q] = sum_I bra[I] ket[I] occ_p(I) occ_q(I), a
// rank-n product of the occupation matrix; thread t of a workgroup owns row
// p = t / TPR and QPT consecutive columns of it in registers over the
// workgroup's slice of determinants, then adds them to M
template <int W>
__global__ void __launch_bounds__(kBlock) kRdmOcc(const uint64_t* dets, const double* bra, const double* ket, int64_t n,
                                                  double* M) {
  constexpr int QPT = W == 2 ? 64 : 16, TPR = 64 * W / QPT, NSO = 64 * W;
  const int p = threadIdx.x / TPR, h = threadIdx.x % TPR;
  const int qw = (h * QPT) >> 6, qs = (h * QPT) & 63;
  double acc[QPT];
#pragma unroll
  for (int j = 0; j < QPT; j++) acc[j] = 0.0;
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  const int64_t b = (int64_t)blockIdx.x * per, e = b + per < n ? b + per : n;
  for (int64_t I = b; I < e; I++) {
    uint64_t d[W];
#pragma unroll
    for (int w = 0; w < W; w++) d[w] = dets[I * W + w];
    if (!((d[p >> 6] >> (p & 63)) & 1ull)) continue;
    const double c2 = bra[I] * ket[I];
    const uint64_t row = d[qw] >> qs;
#pragma unroll
    for (int j = 0; j < QPT; j++)
      if ((row >> j) & 1ull) acc[j] += c2;
  }
#pragma unroll
  for (int j = 0; j < QPT; j++)
    if (acc[j] != 0.0) atomicAdd(M + (int64_t)p * NSO + h * QPT + j, acc[j]);
}

// scatters M (NSO x NSO) into the RDMs: <a+_p a_p> = n_p,
// <a+_p a+_q a_q a_p> = n_p n_q and <a+_p a+_q a_p a_q> = -n_p n_q (p != q)
__global__ void kRdmOccScatter(RdmOut o, const double* M, int NSO) {
  const int nso = o.nso;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nso * nso; t += gridDim.x * blockDim.x) {
    const int p = t / nso, q = t % nso;
    const double v = M[(int64_t)p * NSO + q];
    if (v == 0.0) continue;
    if (p == q) {
      atomicAdd(o.g1 + (int64_t)(p >> 1) * o.n + (p >> 1), v);
      if (o.s1) atomicAdd(o.s1 + (int64_t)p * nso + p, v);
      continue;
    }
    o.add2(p >> 1, q >> 1, p >> 1, q >> 1, v);
    if (((p ^ q) & 1) == 0) o.add2(p >> 1, q >> 1, q >> 1, p >> 1, -v);
    if (o.s2 && p > q) o.spin2(p, q, p, q, v);
  }
}

template <int W>
__global__ void __launch_bounds__(kBlock) kRdm(RdmOut o, const uint64_t* dets, const int64_t* rp, const int32_t* cols,
                                               const double* bra, const double* ket, int64_t ndets) {
  const int wave = threadIdx.x >> 6, lane = __lane_id();
  const int nso = o.nso, norb = o.n;
  for (int64_t I = (int64_t)blockIdx.x * kWaves + wave; I < ndets; I += (int64_t)gridDim.x * kWaves) {
    uint64_t dI[W];
#pragma unroll
    for (int w = 0; w < W; w++) dI[w] = dets[I * W + w];
    const double bI = bra[I], kI = ket[I];
    const int64_t b = rp[I], e = rp[I + 1];
    for (int64_t p = b + lane; p < e; p += 64) {
      const int64_t J = cols[p];
      if (J >= I) continue;
      uint64_t dJ[W];
#pragma unroll
      for (int w = 0; w < W; w++) dJ[w] = dets[J * W + w];
      int cre[2] = {0, 0}, des[2] = {0, 0}, nc = 0, nd = 0;
#pragma unroll
      for (int w = 0; w < W; w++) {
        uint64_t x = dI[w] & ~dJ[w], y = dJ[w] & ~dI[w];
        while (x) {
          const int q = 64 * w + __ffsll((unsigned long long)x) - 1;
          x &= x - 1;
          if (nc < 2) cre[nc] = q;
          nc++;
        }
        while (y) {
          const int q = 64 * w + __ffsll((unsigned long long)y) - 1;
          y &= y - 1;
          if (nd < 2) des[nd] = q;
          nd++;
        }
      }
      if (nc == 1 && nd == 1) {
        // I = a+_a a_i J: sign = parity of the occupied of J strictly between
        const int a = cre[0], i = des[0];
        const int lo = a < i ? a : i, hi = a < i ? i : a;
        const double sg = (countBetween<W>(dJ, lo, hi) & 1) ? -1.0 : 1.0;
        const double vE = sg * bI * ket[J], vT = sg * bra[J] * kI;
        const int A = a >> 1, Ii = i >> 1;
        atomicAdd(o.g1 + (int64_t)A * norb + Ii, vE);
        atomicAdd(o.g1 + (int64_t)Ii * norb + A, vT);
        if (o.s1) {
          atomicAdd(o.s1 + (int64_t)a * nso + i, vE);
          atomicAdd(o.s1 + (int64_t)i * nso + a, vT);
        }
        // common occupied spin-orbitals, per spatial orbital: count and
        // count with the excitation's spin
        const int64_t zE = ((int64_t)A * norb + Ii) * norb, zT = ((int64_t)Ii * norb + A) * norb;
        const int sa = a & 1;
#pragma unroll
        for (int w = 0; w < W; w++) {
          const uint64_t common = dJ[w] & dI[w];
          uint64_t sp = (common | (common >> 1)) & 0x5555555555555555ull;
          while (sp) {
            const int bit = __ffsll((unsigned long long)sp) - 1;
            sp &= sp - 1;
            const int K = (64 * w + bit) >> 1;
            const uint64_t pairBits = (common >> bit) & 3ull;
            const double cnt = (double)__popcll(pairBits);
            atomicAdd(o.Z + zE + K, cnt * vE);
            if (!o.herm) atomicAdd(o.Z + zT + K, cnt * vT);
            if ((pairBits >> sa) & 1ull) {
              atomicAdd(o.Zs + zE + K, vE);
              if (!o.herm) atomicAdd(o.Zs + zT + K, vT);
            }
          }
        }
        if (o.s2) {
#pragma unroll
          for (int w = 0; w < W; w++) {
            uint64_t common = dJ[w] & dI[w];
            while (common) {
              const int k = 64 * w + __ffsll((unsigned long long)common) - 1;
              common &= common - 1;
              // <a+_max(k,a) a+_min(k,a) a_min(k,i) a_max(k,i)>: one swap
              // when k lies between a and i
              const double f = (k > lo && k < hi) ? -1.0 : 1.0;
              const int ah = k > a ? k : a, al = k > a ? a : k, ih = k > i ? k : i, il = k > i ? i : k;
              o.spin2(ah, al, ih, il, f * vE);
              o.spin2(ih, il, ah, al, f * vT);
            }
          }
        }
      } else if (nc == 2 && nd == 2) {
        // I = a+_a a+_b a_j a_i J (a < b, i < j): sign from applying the
        // operators one by one
        const int a = cre[0], bb = cre[1], i = des[0], j = des[1];
        uint64_t d[W];
#pragma unroll
        for (int w = 0; w < W; w++) d[w] = dJ[w];
        int par = belowParity<W>(d, i);
        d[i >> 6] ^= 1ull << (i & 63);
        par += belowParity<W>(d, j);
        d[j >> 6] ^= 1ull << (j & 63);
        par += belowParity<W>(d, bb);
        d[bb >> 6] ^= 1ull << (bb & 63);
        par += belowParity<W>(d, a);
        const double sg = (par & 1) ? -1.0 : 1.0;
        const double vE = sg * bI * ket[J], vT = sg * bra[J] * kI;
        const int A = a >> 1, B = bb >> 1, Ii = i >> 1, Jj = j >> 1;
        // G[A,B,I,J] += vE (exchange image G[B,A,J,I]), G[I,J,A,B] += vT;
        // G[A,B,J,I] -= vE (image G[B,A,I,J]), G[J,I,A,B] -= vT
        if (((a ^ i) & 1) == 0 && ((bb ^ j) & 1) == 0) {
          o.acc(A, B, Ii, Jj, vE);
          if (!o.herm) o.acc(Ii, Jj, A, B, vT);
        }
        if (((a ^ j) & 1) == 0 && ((bb ^ i) & 1) == 0) {
          o.acc(A, B, Jj, Ii, -vE);
          if (!o.herm) o.acc(Jj, Ii, A, B, -vT);
        }
        if (o.s2) {  // <a+_b a+_a a_i a_j> = <a+_a a+_b a_j a_i>
          o.spin2(bb, a, j, i, vE);
          o.spin2(j, i, bb, a, vT);
        }
      }
    }
  }
}

// g2 from the accumulators (one thread per element; overwrites g2)
__global__ void kRdmFinish(RdmOut o) {
  const int n = o.n;
  const int64_t n2 = (int64_t)n * n, n4 = n2 * n2;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n4; t += (int64_t)gridDim.x * blockDim.x) {
    const int S = (int)(t % n), R = (int)((t / n) % n), Q = (int)((t / n2) % n), P = (int)(t / (n2 * n));
    auto at = [&](int p, int q, int r, int s) { return o.A[(((int64_t)p * n + q) * n + r) * n + s]; };
    auto z = [&](const double* Z, int x, int y, int k) {
      const double v = Z[((int64_t)x * n + y) * n + k];
      return o.herm ? v + Z[((int64_t)y * n + x) * n + k] : v;
    };
    double g = at(P, Q, R, S) + at(Q, P, S, R);
    if (o.herm) g += at(R, S, P, Q) + at(S, R, Q, P);
    if (Q == S) g += z(o.Z, P, R, Q);
    if (P == R) g += z(o.Z, Q, S, P);
    if (Q == R) g -= z(o.Zs, P, S, Q);
    if (P == S) g -= z(o.Zs, Q, R, P);
    o.g2[t] = g;
  }
}

// ---------------------------------------------------------------------------
// time reversal: the adapted basis of the standard determinants S and the
// plain basis E = S followed by the flipped copies of the open-shell ones
// (in determinant order).  Basis function of I: |I> when I is closed-shell,
// else (|I> + p_I |flip I>) / sqrt 2, p_I = the flip parity (the Hamiltonian
// of SHCI/Determinants.cpp:32-121 and SHCI/SHCImakeHamiltonian.cpp:411-483,
// which leaves out the Treversal sign).

template <int W>
__global__ void kTrevMark(const uint64_t* dets, int64_t i0, int64_t i1, int64_t* flag) {
  for (int64_t i = i0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < i1; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t d[W];
#pragma unroll
    for (int w = 0; w < W; w++) d[w] = dets[i * W + w];
    flag[i - i0] = hasUnpaired<W>(d) ? 1 : 0;
  }
}

// After an exclusive scan of the flags of [i0, i1): the block of E for these
// determinants (the cnt = i1 - i0 determinants, then the flipped copies of
// the open-shell ones), placed at E index base; each determinant's index in
// E, its partner's (-1 when closed-shell) and its flip parity
template <int W>
__global__ void kTrevBlock(const uint64_t* dets, int64_t i0, int64_t i1, const int64_t* pos, int64_t base,
                           uint64_t* block, int64_t* eIdx, int64_t* partner, double* parity) {
  const int64_t cnt = i1 - i0;
  for (int64_t i = i0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < i1; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = i - i0;
    uint64_t d[W];
#pragma unroll
    for (int w = 0; w < W; w++) {
      d[w] = dets[i * W + w];
      block[k * W + w] = d[w];
    }
    const bool u = pos[k + 1] != pos[k];
    eIdx[i] = base + k;
    parity[i] = flipParity<W>(d);
    partner[i] = u ? base + cnt + pos[k] : -1;
    if (u) {
#pragma unroll
      for (int w = 0; w < W; w++) block[(cnt + pos[k]) * W + w] = flipWord(d[w]);
    }
  }
}

// xE = U x (k columns, ld / ldE)
__global__ void kTrevExpand(const double* x, int64_t n, int k, int64_t ld, const int64_t* eIdx,
                            const int64_t* partner, const double* parity, double* xE, int64_t ldE) {
  const double r2 = 1.0 / sqrt(2.0);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n * k; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t % n, c = t / n;
    const double v = x[c * ld + i];
    const int64_t q = partner[i], e = eIdx[i];
    if (q < 0) {
      xE[c * ldE + e] = v;
    } else {
      xE[c * ldE + e] = v * r2;
      xE[c * ldE + q] = parity[i] * v * r2;
    }
  }
}

// y = U^T yE
__global__ void kTrevContract(const double* yE, int64_t ldE, int64_t n, int k, const int64_t* eIdx,
                              const int64_t* partner, const double* parity, double* y, int64_t ld) {
  const double r2 = 1.0 / sqrt(2.0);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n * k; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t % n, c = t / n;
    const int64_t q = partner[i], e = eIdx[i];
    y[c * ld + i] = q < 0 ? yE[c * ldE + e] : (yE[c * ldE + e] + parity[i] * yE[c * ldE + q]) * r2;
  }
}

// unpackTrevState (SHCI/SHCIbasics.cpp:730-776): c_I / sqrt 2 and
// Treversal p_I c_I / sqrt 2 for the partner, per root (ld n -> ldE)
__global__ void kTrevUnpack(const double* ci, int64_t n, int nroots, int64_t ld, const int64_t* partner,
                            const double* parity, int trev, double* out, int64_t ldE) {
  const double s2 = sqrt(2.0);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n * nroots;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t % n, c = t / n;
    const double v = ci[c * ld + i];
    const int64_t q = partner[i];
    if (q < 0) {
      out[c * ldE + i] = v;
    } else {
      out[c * ldE + i] = v / s2;
      out[c * ldE + q] = trev * parity[i] * v / s2;
    }
  }
}

// ---------------------------------------------------------------------------
// learned screening: features, targets and the exact test of ranked records

// X[r * din + f]: f < nso occupation bit f of the source determinant; then
// one-hot blocks (nso each) of ci, cj, a, b; then log10 |cmax[src]|; zero
// padding up to din
template <int W>
__global__ void kCandFeatures(const CandRecord* rec, int64_t n, const uint64_t* dets, int nso, const double* cmax,
                              float* X, int din) {
  const int64_t total = n * din;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = q / din;
    const int f = (int)(q - r * din);
    const CandRecord c = rec[r];
    float v = 0.f;
    if (f < nso) {
      v = (float)((dets[(int64_t)c.src * W + (f >> 6)] >> (f & 63)) & 1ull);
    } else if (f < 5 * nso) {
      const int blk = (f - nso) / nso, o = (f - nso) - blk * nso;
      const int which = blk == 0 ? c.ci : blk == 1 ? c.cj : blk == 2 ? c.a : c.b;
      v = which == o ? 1.f : 0.f;
    } else if (f == 5 * nso) {
      v = (float)log10(fabs(cmax[c.src]) + 1e-300);
    }
    X[q] = v;
  }
}

// t[r] = log10(|cmax[src]| * |v|): the screening value of the record
__global__ void kCandTargets(const CandRecord* rec, int64_t n, const double* cmax, float* t) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
    t[r] = (float)log10(fabs(cmax[rec[r].src]) * (double)fabsf(rec[r].v) + 1e-300);
}

// records whose score >= tau are evaluated with the reference's double test
// (|v| >= eps1 / |cmax|) and, when accepted, emitted as determinants not yet
// held; counts[0] += evaluated, counts[1] += accepted
template <int W>
__global__ void __launch_bounds__(kBlock) kCandExact(const CandRecord* rec, int64_t n, const float* score, float tau,
                                                     const uint64_t* dets, const double* cmax, double eps1,
                                                     const int32_t* table, uint64_t mask, uint64_t* out,
                                                     unsigned long long* counter, int64_t cap,
                                                     unsigned long long* counts) {
  for (int64_t r0 = (int64_t)blockIdx.x * blockDim.x; r0 < n; r0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = r0 + threadIdx.x;
    bool sel = false, pass = false;
    uint64_t cand[W];
#pragma unroll
    for (int w = 0; w < W; w++) cand[w] = 0;
    if (r < n && score[r] >= tau) {
      sel = true;
      const CandRecord c = rec[r];
      const double eps = eps1 / fabs(cmax[c.src]);
      pass = (double)fabsf(c.v) >= eps;
      if (pass) {
#pragma unroll
        for (int w = 0; w < W; w++) cand[w] = dets[(int64_t)c.src * W + w];
        cand[c.a >> 6] |= (1ull << (c.a & 63));
        cand[c.b >> 6] |= (1ull << (c.b & 63));
        cand[c.ci >> 6] &= ~(1ull << (c.ci & 63));
        cand[c.cj >> 6] &= ~(1ull << (c.cj & 63));
      }
    }
    const uint64_t ms = __ballot(sel), mp = __ballot(pass);
    if (__lane_id() == 0 && ms) {
      atomicAdd(&counts[0], (unsigned long long)__popcll(ms));
      atomicAdd(&counts[1], (unsigned long long)__popcll(mp));
    }
    emitCandidate<W>(pass, cand, table, mask, dets, out, counter, cap);  // (no time reversal in this form)
  }
}

__global__ void kGatherRecords(const CandRecord* rec, const int64_t* idx, int64_t m, CandRecord* out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = rec[idx[i]];
}

// sampled records: flags[i] = (pass ? 1 : 0) | (score >= tau ? 2 : 0) for
// record idx[i]; vals[i] = its score
__global__ void kCandSample(const CandRecord* rec, const float* score, const int64_t* idx, int m, const double* cmax,
                            double eps1, float tau, int* flags, float* vals) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
    const CandRecord c = rec[idx[i]];
    const float sc = score[idx[i]];
    const double eps = eps1 / fabs(cmax[c.src]);
    flags[i] = ((double)fabsf(c.v) >= eps ? 1 : 0) | (sc >= tau ? 2 : 0);
    vals[i] = sc;
  }
}

// ---------------------------------------------------------------------------
// string tables

__device__ __forceinline__ uint64_t compressEven(uint64_t x) {
  x &= 0x5555555555555555ull;
  x = (x | (x >> 1)) & 0x3333333333333333ull;
  x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
  x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
  x = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
  return x;
}

template <int W>
__global__ void kExtractStrings(const uint64_t* dets, int64_t n, uint64_t* alpha, uint64_t* beta) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t a = 0, b = 0;
#pragma unroll
    for (int w = 0; w < W; w++) {
      uint64_t x = dets[i * W + w];
      a |= compressEven(x) << (32 * w);
      b |= compressEven(x >> 1) << (32 * w);
    }
    alpha[i] = a;
    beta[i] = b;
  }
}

__device__ __forceinline__ int64_t lowerBound(const uint64_t* arr, int64_t n, uint64_t x) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (arr[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__global__ void kIndexOf(const uint64_t* str, int64_t n, const uint64_t* sorted, int64_t ns, uint32_t* idx) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    idx[i] = (uint32_t)lowerBound(sorted, ns, str[i]);
}

__global__ void kGroupKeys(const uint32_t* major, int64_t n, uint64_t* key, uint32_t* val) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    key[i] = (uint64_t)major[i] * (uint64_t)n + (uint64_t)i;
    val[i] = (uint32_t)i;
  }
}

// key = major * nminor + minor (members of a group ordered by minor)
__global__ void kGroupKeys2(const uint32_t* major, const uint32_t* minor, int64_t n, uint64_t nminor, uint64_t* key,
                            uint32_t* val) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    key[i] = (uint64_t)major[i] * nminor + (uint64_t)minor[i];
    val[i] = (uint32_t)i;
  }
}

__global__ void kGroupStarts(const uint64_t* key, int64_t n, uint64_t nminor, int64_t ngroups, int64_t* start,
                             const uint32_t* val, const uint64_t* partnerStr, uint64_t* memberPartner) {
  for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m <= n; m += (int64_t)gridDim.x * blockDim.x) {
    if (m == n) {
      start[ngroups] = n;
      continue;
    }
    uint64_t g = key[m] / nminor;
    if (m == 0 || key[m - 1] / nminor != g) start[g] = m;
    memberPartner[m] = partnerStr[val[m]];
  }
}

// alpha strings one excitation away from each unique alpha string
template <bool FILL>
__global__ void kStringSingles(const uint64_t* sorted, int64_t ns, int norb, const int64_t* start, int64_t* count,
                               uint32_t* out) {
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < ns; s += (int64_t)gridDim.x * blockDim.x) {
    uint64_t str = sorted[s];
    int64_t c = 0, pos = FILL ? start[s] : 0;
    uint64_t occd = str;
    while (occd) {
      int j = __ffsll((unsigned long long)occd) - 1;
      occd &= occd - 1;
      uint64_t emp = ~str & ((norb >= 64) ? ~0ull : ((1ull << norb) - 1ull));
      while (emp) {
        int k = __ffsll((unsigned long long)emp) - 1;
        emp &= emp - 1;
        uint64_t t = (str & ~(1ull << j)) | (1ull << k);
        int64_t f = lowerBound(sorted, ns, t);
        if (f < ns && sorted[f] == t) {
          if (FILL) out[pos++] = (uint32_t)f;
          c++;
        }
      }
    }
    if (!FILL) count[s] = c;
  }
}

// ---------------------------------------------------------------------------
// connected-pair rows: one wavefront per row

struct RowTables {
  const uint64_t* alpha;
  const uint64_t* beta;
  const uint32_t* aIdx;
  const uint32_t* bIdx;
  const int64_t* aStart;
  const int64_t* bStart;
  const uint32_t* aMajorDet;
  const uint32_t* bMajorDet;
  const uint64_t* aMajorB;
  const uint64_t* bMajorA;
  const int64_t* aSinglesStart;
  const uint32_t* aSingles;
  const int64_t* bSinglesStart;
  const uint32_t* bSingles;
  const uint64_t* sortedB;
  const uint64_t* aSortB;
  const uint32_t* aSortDet;
  int64_t n;
};

// Collects the connections of row I.  Count pass: per-lane counters; fill
// pass: wave-uniform cursor, positions by ballot (deterministic order).
// SPLIT stores each FP64 value as FP32 hi = fp32(v) and lo = fp32(v - hi).
template <int W, bool FILL, bool SPLIT, bool LOWER>
struct RowSink {
  const DevSystem* S;
  const uint64_t* dets;
  const uint64_t* dI;
  int64_t I;
  int32_t* cols;
  double* vals;
  float* hi;
  float* lo;
  int64_t pos;  // wave-uniform write cursor (fill)
  int cnt;      // per-lane count (count)
  __device__ __forceinline__ void hit(bool h, uint32_t J) {
    if (LOWER) h = h && (int64_t)J < I;  // half storage: row max(I, J) only
    if (!FILL) {
      cnt += h ? 1 : 0;
      return;
    }
    uint64_t m = __ballot(h);
    if (h) {
      int64_t p = pos + __popcll(m & lanemaskLt());
      uint64_t dJ[W];
#pragma unroll
      for (int w = 0; w < W; w++) dJ[w] = dets[(int64_t)J * W + w];
      double v = ((int64_t)J < I) ? hijElement<W>(*S, dJ, dI) : hijElement<W>(*S, dI, dJ);
      cols[p] = (int32_t)J;
      if (SPLIT) {
        const float h = (float)v;
        hi[p] = h;
        lo[p] = (float)(v - (double)h);
      } else {
        vals[p] = v;
      }
    }
    pos += __popcll(m);
  }
};

// first position in [lo, hi) whose determinant index is >= x (members of a
// group are stored in increasing determinant index)
__device__ __forceinline__ int64_t firstDetAtLeast(const uint32_t* a, int64_t lo, int64_t hi, uint32_t x) {
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (a[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Visits every connection of row I (all J != I one or two excitations
// apart, restricted to J >= jmin): sink.hit(bool, J) is called by all 64
// lanes together, once per candidate slot.  (a) the same-beta group: alpha
// singles and alpha-alpha doubles; (b) the same-alpha group: beta singles
// and beta-beta doubles; (c) alpha single x beta single: members of the
// groups of alpha strings one excitation away whose beta string is one
// excitation away.  Group members are in increasing determinant index, so
// jmin skips a prefix of each group (binary search).
template <class Sink>
__device__ __forceinline__ void enumerateRow(const RowTables& T, int64_t I, uint32_t jmin, int wave, int lane,
                                             int (*sExcl)[65], int64_t (*sBeg)[64], int (*sLook)[64], Sink& sink) {
  const uint64_t astr = T.alpha[I], bstr = T.beta[I];
  const uint32_t ai = T.aIdx[I], bi = T.bIdx[I];
  {
    int64_t b0 = T.bStart[bi], b1 = T.bStart[bi + 1];
    if (jmin) b0 = firstDetAtLeast(T.bMajorDet, b0, b1, jmin);
    for (int64_t m0 = b0; m0 < b1; m0 += 64) {
      int64_t m = m0 + lane;
      bool h = false;
      uint32_t J = 0;
      if (m < b1) {
        int dd = __popcll(T.bMajorA[m] ^ astr);
        h = (dd == 2 || dd == 4);
        J = T.bMajorDet[m];
      }
      sink.hit(h, J);
    }
  }
  {
    int64_t a0 = T.aStart[ai], a1 = T.aStart[ai + 1];
    if (jmin) a0 = firstDetAtLeast(T.aMajorDet, a0, a1, jmin);
    for (int64_t m0 = a0; m0 < a1; m0 += 64) {
      int64_t m = m0 + lane;
      bool h = false;
      uint32_t J = 0;
      if (m < a1) {
        int dd = __popcll(T.aMajorB[m] ^ bstr);
        h = (dd == 2 || dd == 4);
        J = T.aMajorDet[m];
      }
      sink.hit(h, J);
    }
  }
  {
    // alpha single x beta single: per alpha single group either scan its
    // members for a beta one excitation from the row's (the members in
    // determinant order, from the first at or after jmin), or, when the
    // group is large, look up each of the row's beta singles in the group's
    // beta-ordered members (binary search)
    int64_t s0 = T.aSinglesStart[ai], s1 = T.aSinglesStart[ai + 1];
    const int64_t bs0 = T.bSinglesStart[bi];
    const int nbs = (int)(T.bSinglesStart[bi + 1] - bs0);
    for (int64_t c0 = s0; c0 < s1; c0 += 64) {
      int64_t c = c0 + lane;
      int len = 0, look = 0, gsz = 0;
      int64_t beg = 0;
      if (c < s1) {
        uint32_t sa = T.aSingles[c];
        const int64_t g0 = T.aStart[sa], g1 = T.aStart[sa + 1];
        beg = g0;
        if (jmin) beg = firstDetAtLeast(T.aMajorDet, beg, g1, jmin);
        len = (int)(g1 - beg);
        gsz = (int)(g1 - g0);
        const int steps = 66 - __clzll((unsigned long long)gsz);  // binary-search depth + 2
        if ((int64_t)nbs * steps < len) {
          look = 1;
          len = nbs;
          beg = g0;
        }
      }
      int inc = len;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
      }
      sExcl[wave][lane] = inc - len;
      sBeg[wave][lane] = beg;
      sLook[wave][lane] = look ? gsz : -1;
      int total = __shfl(inc, 63, 64);
      __syncwarp();
      for (int k0 = 0; k0 < total; k0 += 64) {
        int k = k0 + lane;
        bool h = false;
        uint32_t J = 0;
        if (k < total) {
          int lo = 0, hi = 63;
          while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (sExcl[wave][mid] <= k)
              lo = mid;
            else
              hi = mid - 1;
          }
          const int o = k - sExcl[wave][lo];
          const int g = sLook[wave][lo];
          if (g < 0) {
            int64_t m = sBeg[wave][lo] + o;
            h = __popcll(T.aMajorB[m] ^ bstr) == 2;
            J = T.aMajorDet[m];
          } else {
            const uint64_t want = T.sortedB[T.bSingles[bs0 + o]];
            int64_t a = sBeg[wave][lo], b = a + g;  // [a, b)
            while (a < b) {
              const int64_t mid = (a + b) >> 1;
              if (T.aSortB[mid] < want)
                a = mid + 1;
              else
                b = mid;
            }
            if (a < sBeg[wave][lo] + g && T.aSortB[a] == want) {
              J = T.aSortDet[a];
              h = J >= jmin;
            }
          }
        }
        sink.hit(h, J);
      }
      __syncwarp();
    }
  }
}

// Rows I < nOld enumerate only connections to J >= nOld (the columns the
// previous Hamiltonian lacks); rows I >= nOld enumerate all connections.
// LOWER (half storage) keeps only J < I, so rows I < nOld gain nothing and
// are skipped.
template <int W, bool FILL, bool SPLIT, bool LOWER>
__global__ void __launch_bounds__(kBlock) kCsrRows(DevSystem S, const uint64_t* dets, RowTables T, int64_t nOld,
                                                   int64_t* rowCount, const int64_t* fillStart, int32_t* cols,
                                                   double* vals, float* hi, float* lo) {
  __shared__ int sExcl[kWaves][65];
  __shared__ int64_t sBeg[kWaves][64];
  __shared__ int sLook[kWaves][64];
  const int wave = threadIdx.x >> 6, lane = __lane_id();
  for (int64_t I = (int64_t)blockIdx.x * kWaves + wave; I < T.n; I += (int64_t)gridDim.x * kWaves) {
    uint64_t dI[W];
#pragma unroll
    for (int w = 0; w < W; w++) dI[w] = dets[I * W + w];
    if (LOWER && I < nOld) {
      if (!FILL && lane == 0) rowCount[I] = 0;
      continue;
    }
    const uint32_t jmin = I >= nOld ? 0u : (uint32_t)nOld;
    RowSink<W, FILL, SPLIT, LOWER> sink;
    sink.S = &S;
    sink.dets = dets;
    sink.dI = dI;
    sink.I = I;
    sink.cols = cols;
    sink.vals = vals;
    sink.hi = hi;
    sink.lo = lo;
    sink.pos = FILL ? fillStart[I] : 0;
    sink.cnt = 0;
    enumerateRow(T, I, jmin, wave, lane, sExcl, sBeg, sLook, sink);
    if (!FILL) {
      int c = sink.cnt;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
      if (lane == 0) rowCount[I] = c;
    }
  }
}

// Direct sigma (no stored H): y[:, r] = diag .* x[:, r] + sum over the
// connections J of row I of Hij(bra = D_I, ket = D_J) x[J, r], the
// orientation of the reference's direct product; each element is evaluated
// once for all k <= 8 vectors.
template <int W>
struct DirectSink {
  const DevSystem* S;
  const uint64_t* dets;
  const uint64_t* dI;
  const double* x;
  int64_t ld;
  int k;
  double acc[8];
  __device__ __forceinline__ void hit(bool h, uint32_t J) {
    if (!h) return;
    uint64_t dJ[W];
#pragma unroll
    for (int w = 0; w < W; w++) dJ[w] = dets[(int64_t)J * W + w];
    const double v = hijElement<W>(*S, dI, dJ);
#pragma unroll
    for (int r = 0; r < 8; r++)
      if (r < k) acc[r] = fma(v, x[r * ld + J], acc[r]);
  }
};

template <int W>
__global__ void __launch_bounds__(kBlock) kSigmaDirect(DevSystem S, const uint64_t* dets, RowTables T,
                                                       const double* diag, const double* x, double* y, int k,
                                                       int64_t ld) {
  __shared__ int sExcl[kWaves][65];
  __shared__ int64_t sBeg[kWaves][64];
  __shared__ int sLook[kWaves][64];
  const int wave = threadIdx.x >> 6, lane = __lane_id();
  for (int64_t I = (int64_t)blockIdx.x * kWaves + wave; I < T.n; I += (int64_t)gridDim.x * kWaves) {
    uint64_t dI[W];
#pragma unroll
    for (int w = 0; w < W; w++) dI[w] = dets[I * W + w];
    DirectSink<W> sink;
    sink.S = &S;
    sink.dets = dets;
    sink.dI = dI;
    sink.x = x;
    sink.ld = ld;
    sink.k = k;
#pragma unroll
    for (int r = 0; r < 8; r++) sink.acc[r] = 0.0;
    enumerateRow(T, I, 0u, wave, lane, sExcl, sBeg, sLook, sink);
#pragma unroll
    for (int r = 0; r < 8; r++) {
      if (r >= k) break;
      double v = sink.acc[r];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) y[r * ld + I] = fma(diag[I], x[r * ld + I], v);
    }
  }
}

// new row length = kept old entries + new connections
__global__ void kNewRowLengths(const int64_t* oldRowPtr, int64_t nOld, int64_t n, int64_t* cnt) {
  for (int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; I < n; I += (int64_t)gridDim.x * blockDim.x)
    if (I < nOld) cnt[I] += oldRowPtr[I + 1] - oldRowPtr[I];
}

__global__ void kFillStart(const int64_t* oldRowPtr, int64_t nOld, int64_t n, const int64_t* rowPtr, int64_t* fs) {
  for (int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; I < n; I += (int64_t)gridDim.x * blockDim.x)
    fs[I] = rowPtr[I] + (I < nOld ? oldRowPtr[I + 1] - oldRowPtr[I] : 0);
}

// copy the old rows to the front of the new rows (one wavefront per row)
template <typename T1>
__global__ void __launch_bounds__(kBlock) kCopyOldRows(const int64_t* oldRowPtr, const int32_t* oldCols,
                                                       const T1* oldA, const T1* oldB, int64_t nOld,
                                                       const int64_t* rowPtr, int32_t* cols, T1* a, T1* b) {
  const int wave = threadIdx.x >> 6, lane = __lane_id();
  for (int64_t I = (int64_t)blockIdx.x * kWaves + wave; I < nOld; I += (int64_t)gridDim.x * kWaves) {
    const int64_t s = oldRowPtr[I], e = oldRowPtr[I + 1], d = rowPtr[I];
    for (int64_t p = s + lane; p < e; p += 64) {
      cols[d + (p - s)] = __builtin_nontemporal_load(oldCols + p);
      a[d + (p - s)] = __builtin_nontemporal_load(oldA + p);
      if (oldB) b[d + (p - s)] = __builtin_nontemporal_load(oldB + p);
    }
  }
}

__global__ void kCast(const double* v, float* out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (float)v[i];
}

// per-row insertion/shell sort of columns (rows are short relative to n)
__global__ void kSortRows(const int64_t* rowPtr, int64_t n, int32_t* cols, double* vals) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    int64_t b = rowPtr[r], e = rowPtr[r + 1];
    int64_t len = e - b;
    for (int64_t gap = len / 2; gap > 0; gap /= 2) {
      for (int64_t i = b + gap; i < e; i++) {
        int32_t c = cols[i];
        double v = vals[i];
        int64_t j = i;
        while (j - gap >= b && cols[j - gap] > c) {
          cols[j] = cols[j - gap];
          vals[j] = vals[j - gap];
          j -= gap;
        }
        cols[j] = c;
        vals[j] = v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// sigma = H x over the full symmetric CSR (gather only, no atomics): one
// wavefront per row, eight loads in flight per lane, FP64 accumulation, the
// diagonal term fused into the row's single store.  Values are VT,
// optionally plus an FP32 low part (hi + lo reproduces the FP64 value to
// ~2^-48 relative).  (Splitting the columns into per-XCD blocks, LDS-tiled
// gathers and 4-wide vector loads were all measured slower on a
// 1.2M-determinant matrix: the gathers of x are L2-request bound.)
template <typename VT, bool LO>
__global__ void __launch_bounds__(kBlock) kSigma(const int64_t* __restrict__ rowPtr, const int32_t* __restrict__ cols,
                                                 const VT* __restrict__ vals, const float* __restrict__ lo,
                                                 const double* __restrict__ diag, int64_t n,
                                                 const double* __restrict__ x, double* __restrict__ y, int k,
                                                 int64_t ld) {
  const int wave = threadIdx.x >> 6, lane = __lane_id();
  constexpr int U = 8;  // loads in flight per lane
  for (int64_t I = (int64_t)blockIdx.x * kWaves + wave; I < n; I += (int64_t)gridDim.x * kWaves) {
    const int64_t b = rowPtr[I], e = rowPtr[I + 1];
    const double d = diag[I];
    for (int r = 0; r < k; r++) {
      const double* xr = x + r * ld;
      double acc[U];
#pragma unroll
      for (int u = 0; u < U; u++) acc[u] = 0.0;
      int64_t p = b + lane;
      for (; p + 64 * (U - 1) < e; p += 64 * U) {
        int32_t c[U];
        double v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          c[u] = __builtin_nontemporal_load(cols + p + 64 * u);
          v[u] = (double)__builtin_nontemporal_load(vals + p + 64 * u);
          if (LO) v[u] += (double)__builtin_nontemporal_load(lo + p + 64 * u);
        }
#pragma unroll
        for (int u = 0; u < U; u++) acc[u] = fma(v[u], xr[c[u]], acc[u]);
      }
      for (; p < e; p += 64) acc[0] = fma((double)vals[p] + (LO ? (double)lo[p] : 0.0), xr[cols[p]], acc[0]);
      double sum = 0.0;
#pragma unroll
      for (int u = 0; u < U; u++) sum += acc[u];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
      if (lane == 0) y[r * ld + I] = fma(d, xr[I], sum);
    }
  }
}

// ---------------------------------------------------------------------------
// sigma over the alpha-major relabelled copy of H.  Row m of the copy is
// determinant perm[m] (members of one alpha string are consecutive, in
// increasing determinant index) and its columns are alpha-major positions,
// so the x values a row gathers (same-alpha group members, members of the
// groups of alpha strings one excitation away) sit in a few contiguous
// ranges, and consecutive rows gather from the same ranges.  A work item is
// a chunk of consecutive rows, so that reuse stays within one CU.

__global__ void kInversePerm(const uint32_t* perm, int64_t n, uint32_t* pos) {
  for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < n; m += (int64_t)gridDim.x * blockDim.x)
    pos[perm[m]] = (uint32_t)m;
}

__global__ void kPermRowLengths(const int64_t* rowPtr, const uint32_t* perm, int64_t n, int64_t* len) {
  for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < n; m += (int64_t)gridDim.x * blockDim.x) {
    const int64_t I = perm[m];
    len[m] = rowPtr[I + 1] - rowPtr[I];
  }
}

// row m of the copy = row perm[m] of H with columns mapped through pos; with
// bIdx / bPos given, a column whose determinant has the row's beta string
// maps to n + bPos[J] instead (its place in a beta-major copy of x that
// follows the alpha-major one: a row's same-beta columns all lie in one beta
// group there, so their gathers share lines)
template <typename T1>
__global__ void __launch_bounds__(kBlock) kPermRows(const int64_t* rowPtr, const int32_t* cols, const T1* vals,
                                                    const uint32_t* perm, const uint32_t* pos,
                                                    const uint32_t* bIdx, const uint32_t* bPos, int64_t n,
                                                    const int64_t* rpP, int32_t* colsP, T1* valsP) {
  const int wave = threadIdx.x >> 6, lane = __lane_id();
  for (int64_t m = (int64_t)blockIdx.x * kWaves + wave; m < n; m += (int64_t)gridDim.x * kWaves) {
    const int64_t I = perm[m];
    const int64_t b = rowPtr[I], e = rowPtr[I + 1], d = rpP[m];
    const uint32_t bi = bIdx ? bIdx[I] : 0u;
    for (int64_t p = b + lane; p < e; p += 64) {
      const int32_t J = __builtin_nontemporal_load(cols + p);
      colsP[d + (p - b)] = bIdx && bIdx[J] == bi ? (int32_t)(n + bPos[J]) : (int32_t)pos[J];
      valsP[d + (p - b)] = __builtin_nontemporal_load(vals + p);
    }
  }
}

// x in alpha-major order for the sigma kernels, followed (bPerm given) by
// x in beta-major order: row m < nx of the copy holds x[perm[m]] and row
// n + m' holds x[bPerm[m']], nx = 2n (n without bPerm); K == 1: xP[r * nx +
// m] for r < k; K > 1 (SpMM): interleaved, xP[m * K + r] (zero for k <= r <
// K).  Also zeroes the work counter of a persistent sigma launch that
// follows on the same stream.
// the persistent sigma's work counters: 8 words kSchedStride apart (the
// counter buffer holds 64 words)
constexpr int kSchedStride = 8;

template <typename XT, int K>
__global__ void kGatherPerm(const double* x, int k, int64_t ld, const uint32_t* perm, const uint32_t* bPerm, int64_t n,
                            XT* xP, unsigned* counter) {
  if (counter && blockIdx.x == 0 && threadIdx.x < 8) counter[kSchedStride * threadIdx.x] = 0u;
  const int64_t nx = bPerm ? 2 * n : n;
  for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < nx; m += (int64_t)gridDim.x * blockDim.x) {
    const int64_t I = m < n ? perm[m] : bPerm[m - n];
    if (K == 1) {
      for (int r = 0; r < k; r++) xP[r * nx + m] = (XT)x[r * ld + I];
    } else {
#pragma unroll
      for (int r = 0; r < K; r++) xP[m * K + r] = r < k ? (XT)x[r * ld + I] : (XT)0;
    }
  }
}

// Row-chunk scheduler of the sigma kernels: CHUNK consecutive rows per work
// item; static (work item = workgroup) or persistent (workgroups take work
// items from an atomic counter until none is left, which evens out rows of
// very different lengths).
template <int CHUNK, bool PERSIST>
struct ChunkSched {
  int64_t nChunks;
  unsigned* counter;
  int64_t cur;
  __device__ __forceinline__ bool next() {
    if (!PERSIST) {
      if (cur < 0) {
        cur = blockIdx.x;
        return cur < nChunks;
      }
      return false;
    }
    // XCD-aware: workgroups are dispatched round-robin over the 8 XCDs, so
    // blockIdx mod 8 names the XCD.  Each XCD takes the chunks of its own
    // eighth of the rows first (consecutive alpha groups, whose x gathers
    // then share its L2), then helps the next XCDs; one counter per eighth
    // (kSchedCounters words apart, cleared by the gather kernel).  Results
    // do not depend on which workgroup takes a chunk.
    __shared__ int64_t sCur;
    const int x = blockIdx.x & 7;
    __syncthreads();
    if (threadIdx.x == 0) {
      sCur = -1;
      for (; off < 8; off++) {
        const int t = (x + off) & 7;
        const int64_t lo = nChunks * t / 8, hi = nChunks * (t + 1) / 8;
        const int64_t c = lo + (int64_t)atomicAdd(counter + kSchedStride * t, 1u);
        if (c < hi) {
          sCur = c;
          break;
        }
      }
    }
    __syncthreads();
    cur = sCur;
    return cur >= 0;
  }
  int off = 0;  // (thread 0) eighths already exhausted
};

// y[r*ld + perm[m]] = diag .* x + sum over row m of the copy (gathers from
// xP, one vector at a time, vectors nx apart)
template <typename VT, typename XT, int CHUNK, bool PERSIST>
__global__ void __launch_bounds__(kBlock) kSigmaPerm(const int64_t* __restrict__ rp, const int32_t* __restrict__ cols,
                                                     const VT* __restrict__ vals, const double* __restrict__ diag,
                                                     const uint32_t* __restrict__ perm, int64_t n,
                                                     const XT* __restrict__ xP, int64_t nx, const double* __restrict__ x,
                                                     double* __restrict__ y, int k, int64_t ld, unsigned* counter) {
  const int wave = threadIdx.x >> 6, lane = __lane_id();
  constexpr int U = 8;
  ChunkSched<CHUNK, PERSIST> sched{(n + CHUNK - 1) / CHUNK, counter, -1};
  while (sched.next()) {
    const int64_t r0 = sched.cur * CHUNK;
    const int64_t r1 = r0 + CHUNK < n ? r0 + CHUNK : n;
    for (int64_t m = r0 + wave; m < r1; m += kWaves) {
      const int64_t b = rp[m], e = rp[m + 1];
      const int64_t I = perm[m];
      const double d = diag[I];
      for (int r = 0; r < k; r++) {
        const XT* xr = xP + r * nx;
        double acc[U];
#pragma unroll
        for (int u = 0; u < U; u++) acc[u] = 0.0;
        int64_t p = b + lane;
        for (; p + 64 * (U - 1) < e; p += 64 * U) {
          int32_t c[U];
          double v[U];
#pragma unroll
          for (int u = 0; u < U; u++) {
            c[u] = __builtin_nontemporal_load(cols + p + 64 * u);
            v[u] = (double)__builtin_nontemporal_load(vals + p + 64 * u);
          }
#pragma unroll
          for (int u = 0; u < U; u++) acc[u] = fma(v[u], (double)xr[c[u]], acc[u]);
        }
        for (; p < e; p += 64) acc[0] = fma((double)vals[p], (double)xr[cols[p]], acc[0]);
        double sum = 0.0;
#pragma unroll
        for (int u = 0; u < U; u++) sum += acc[u];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
        if (lane == 0) y[r * ld + I] = fma(d, x[r * ld + I], sum);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Half storage: each unordered pair {I, J} stored once, in row max(I, J)
// (the reference's lower-triangular rows, SHCI/SHCImakeHamiltonian.cpp:549-711,
// without their diagonal), for spaces whose two-triangle Hamiltonian does not
// fit.  Sigma applies every stored element twice, as Hmult2 does
// (SHCI/Hmult.h:110-121): y[I] += h x[J] from a gather and y[J] += h x[I]
// by an FP64 atomic add at the L2, so its sums are not in a fixed order.

// rows [nOld, n] of a grown half-storage CSR: rowPtr[I] = rowPtr[nOld] +
// cnt[I], cnt the exclusive scan of new-row lengths (zero below nOld)
__global__ void kLowerRowPtr(const int64_t* cnt, int64_t nOld, int64_t n, int64_t* rowPtr) {
  const int64_t base = nOld ? rowPtr[nOld] : 0;
  for (int64_t I = nOld + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; I <= n; I += (int64_t)gridDim.x * blockDim.x)
    if (I > nOld || nOld == 0) rowPtr[I] = base + cnt[I];
}

__global__ void kZeroVectors(double* y, int k, int64_t ld, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n * k; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / n;
    y[r * ld + (i - r * n)] = 0.0;
  }
}

template <typename VT, bool LO>
__global__ void __launch_bounds__(kBlock) kSigmaLower(const int64_t* __restrict__ rowPtr,
                                                      const int32_t* __restrict__ cols, const VT* __restrict__ vals,
                                                      const float* __restrict__ lo, const double* __restrict__ diag,
                                                      int64_t n, const double* __restrict__ x, double* __restrict__ y,
                                                      int k, int64_t ld) {
  const int wave = threadIdx.x >> 6, lane = __lane_id();
  constexpr int U = 4;
  for (int64_t I = (int64_t)blockIdx.x * kWaves + wave; I < n; I += (int64_t)gridDim.x * kWaves) {
    const int64_t b = rowPtr[I], e = rowPtr[I + 1];
    for (int r = 0; r < k; r++) {
      const double* xr = x + r * ld;
      double* yr = y + r * ld;
      const double xi = xr[I];
      double acc = 0.0;
      int64_t p = b + lane;
      for (; p + 64 * (U - 1) < e; p += 64 * U) {
        int32_t c[U];
        double v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          c[u] = __builtin_nontemporal_load(cols + p + 64 * u);
          v[u] = (double)__builtin_nontemporal_load(vals + p + 64 * u);
          if (LO) v[u] += (double)__builtin_nontemporal_load(lo + p + 64 * u);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
          acc = fma(v[u], xr[c[u]], acc);
          unsafeAtomicAdd(yr + c[u], v[u] * xi);
        }
      }
      for (; p < e; p += 64) {
        const int32_t c = cols[p];
        const double v = (double)vals[p] + (LO ? (double)lo[p] : 0.0);
        acc = fma(v, xr[c], acc);
        unsafeAtomicAdd(yr + c, v * xi);
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
      if (lane == 0) unsafeAtomicAdd(yr + I, fma(diag[I], xi, acc));
    }
  }
}

// ---------------------------------------------------------------------------
// Tiled sigma.  The alpha-major rows are cut into blocks of kTileRows
// consecutive rows, and each block's entries are regrouped by column tile
// (kTileCols consecutive values of the permuted x).  One wave instruction
// takes up to 64 entries of one (block, tile) segment, so its gathers stay
// within one tile, and the tile's x lines, brought into L1 by the first
// instruction, serve the rest of the block's entries in that tile.  On the
// 1.18M-determinant benchmark space a row's entries touch 0.44 distinct
// 128-byte x lines per entry, a 256-row block's 0.038.  Within a segment
// the entries are ordered by the wave that placed them (row mod 4), then by
// row, then in the row's own order, so each row is one contiguous run; an
// instruction's products are summed per row by a segmented scan across
// lanes (each entry carries its distance to its run's first entry) and
// added into the executing wave's partial row sums in LDS, and the four
// partials are added in a fixed order at the end of the block.  Which wave
// takes which instructions is fixed by the layout, so the result does not
// depend on scheduling.  Segments of many instructions can gather from a
// copy of their x tile in LDS instead of from L1 (kSigmaTiled, STAGE).
constexpr int kTileRows = 256;
constexpr int kTileLog = 10;
constexpr int kTileCols = 1 << kTileLog;
constexpr int kTileMaxWords = 4096;  // bitmap of up to 131,072 tiles: x of up to 134M values
constexpr int kTileInstPad = 16;  // zero descriptors after the last one (the largest step of kSigmaTiled)
constexpr int kTileLdsSegs = 1024;  // blocks with at most this many segments count and fill in LDS
// an entry's 32-bit word: column in tile (bits 0-9), distance to the first
// entry of its row run within the instruction (bits 10-15), row in block
// (bits 16-23), last entry of its run within the instruction (bit 31)
constexpr int kTileDistShift = 10;
// an instruction descriptor: its first entry (offset within the block),
// and a second word with the entry count (bits 0-6), the tile (bits 7-30),
// and bit 31 set when its segment has at least stageMin
// instructions (the staged kernel copies such a segment's x tile to LDS)
constexpr uint32_t kTileDescMask = 0xffffffu;

// the permuted-x column of H's column J in a row with beta string index bi
// (pos null: the input is the relabelled copy, whose columns are mapped)
__device__ __forceinline__ uint32_t tileMapCol(int32_t J, uint32_t bi, const uint32_t* __restrict__ pos,
                                               const uint32_t* __restrict__ bIdx,
                                               const uint32_t* __restrict__ bPos, int64_t n) {
  if (!pos) return (uint32_t)J;
  return (bIdx && bIdx[J] == bi) ? (uint32_t)(n + bPos[J]) : pos[J];
}

// inclusive scan over the block's threads (kBlock); returns the thread's
// inclusive value and the block total in *total
template <typename T>
__device__ __forceinline__ T blockInclusiveScan(T v, T* total) {
  __shared__ T waveSum[kWaves];
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const T o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  if (lane == 63) waveSum[wave] = v;
  __syncthreads();
  T before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kWaves; w++) {
    if (w < wave) before += waveSum[w];
    all += waveSum[w];
  }
  __syncthreads();
  *total = all;
  return v + before;
}

// marks the tiles the block's entries touch in bits[nwords] (LDS, cleared
// here); with pre given, pre[k] = set bits before word k; returns the
// number of touched tiles
__device__ int tileBitmap(const int64_t* __restrict__ rowPtr, const int32_t* __restrict__ cols,
                          const uint32_t* __restrict__ perm, const uint32_t* __restrict__ pos,
                          const uint32_t* __restrict__ bIdx, const uint32_t* __restrict__ bPos, int64_t n,
                          int64_t b, int nwords, uint32_t* bits, uint32_t* pre) {
  const int wave = threadIdx.x >> 6, lane = __lane_id();
  for (int k = threadIdx.x; k < nwords; k += kBlock) bits[k] = 0u;
  __syncthreads();
  const int64_t m0 = b * kTileRows;
  const int rows = (int)(n - m0 < kTileRows ? n - m0 : kTileRows);
  for (int r = wave; r < rows; r += kWaves) {
    const int64_t I = perm ? (int64_t)perm[m0 + r] : m0 + r;
    const int64_t e0 = rowPtr[I], e1 = rowPtr[I + 1];
    const uint32_t bi = bIdx ? bIdx[I] : 0u;
    for (int64_t p = e0 + lane; p < e1; p += 64) {
      const uint32_t t = tileMapCol(cols[p], bi, pos, bIdx, bPos, n) >> kTileLog;
      atomicOr(&bits[t >> 5], 1u << (t & 31));
    }
  }
  __syncthreads();
  const int per = (nwords + kBlock - 1) / kBlock;
  const int k0 = threadIdx.x * per, k1 = k0 + per < nwords ? k0 + per : nwords;
  int local = 0;
  for (int k = k0; k < k1; k++) local += __popc(bits[k]);
  int total = 0;
  int run = blockInclusiveScan(local, &total) - local;
  if (pre)
    for (int k = k0; k < k1; k++) {
      pre[k] = (uint32_t)run;
      run += __popc(bits[k]);
    }
  __syncthreads();
  return total;
}

__device__ __forceinline__ uint32_t tileRank(const uint32_t* bits, const uint32_t* pre, uint32_t t) {
  return pre[t >> 5] + __popc(bits[t >> 5] & ((1u << (t & 31)) - 1u));
}

// pass 1: touched tiles per block (segCount[b])
__global__ void __launch_bounds__(kBlock) kTileMark(const int64_t* rowPtr, const int32_t* cols, const uint32_t* perm,
                                                    const uint32_t* pos, const uint32_t* bIdx, const uint32_t* bPos,
                                                    int64_t n, int nwords, int64_t* segCount) {
  __shared__ uint32_t bits[kTileMaxWords];
  const int total = tileBitmap(rowPtr, cols, perm, pos, bIdx, bPos, n, blockIdx.x, nwords, bits, nullptr);
  if (threadIdx.x == 0) segCount[blockIdx.x] = total;
}

// pass 2: per (segment, wave) entry counts; segment tiles; per segment its
// entry and instruction offsets within the block; per (segment, wave) the
// fill cursor: the block-relative position of the wave's first entry in
// the segment; per block its entries and instructions
__global__ void __launch_bounds__(kBlock) kTileCount(const int64_t* rowPtr, const int32_t* cols, const uint32_t* perm,
                                                     const uint32_t* pos, const uint32_t* bIdx, const uint32_t* bPos,
                                                     int64_t n, int nwords, const int64_t* segBase, uint32_t* segTile,
                                                     uint32_t* segCnt, uint32_t* segStart, uint32_t* segInst,
                                                     uint32_t* segTotal, uint32_t* fill, int64_t* blockE,
                                                     int64_t* blockI) {
  __shared__ uint32_t bits[kTileMaxWords];
  __shared__ uint32_t pre[kTileMaxWords];
  __shared__ uint32_t cntL[kWaves][kTileLdsSegs];
  const int64_t b = blockIdx.x;
  tileBitmap(rowPtr, cols, perm, pos, bIdx, bPos, n, b, nwords, bits, pre);
  const int wave = threadIdx.x >> 6, lane = __lane_id();
  const int64_t base = segBase[b];
  const bool inLds = segBase[b + 1] - base <= kTileLdsSegs;
  if (inLds)
    for (int k = threadIdx.x; k < kWaves * kTileLdsSegs; k += kBlock) (&cntL[0][0])[k] = 0u;
  __syncthreads();
  const int64_t m0 = b * kTileRows;
  const int rows = (int)(n - m0 < kTileRows ? n - m0 : kTileRows);
  for (int r = wave; r < rows; r += kWaves) {
    const int64_t I = perm ? (int64_t)perm[m0 + r] : m0 + r;
    const int64_t e0 = rowPtr[I], e1 = rowPtr[I + 1];
    const uint32_t bi = bIdx ? bIdx[I] : 0u;
    for (int64_t p = e0 + lane; p < e1; p += 64) {
      const uint32_t t = tileMapCol(cols[p], bi, pos, bIdx, bPos, n) >> kTileLog;
      const uint32_t sg = tileRank(bits, pre, t);
      if (inLds)
        atomicAdd(&cntL[wave][sg], 1u);
      else
        atomicAdd(&segCnt[(base + sg) * 4 + wave], 1u);
    }
  }
  if (inLds) {
    __syncthreads();
    const int64_t D = segBase[b + 1] - base;
    for (int64_t k = threadIdx.x; k < D * 4; k += kBlock) segCnt[base * 4 + k] = cntL[k & 3][k >> 2];
  }
  for (int k = threadIdx.x; k < nwords; k += kBlock) {
    uint32_t w = bits[k];
    uint32_t j = pre[k];
    while (w) {
      segTile[base + j] = ((uint32_t)k << 5) + (uint32_t)(__ffs(w) - 1);
      j++;
      w &= w - 1u;
    }
  }
  __threadfence();
  __syncthreads();
  const int64_t D = segBase[b + 1] - base;
  uint64_t carry = 0;  // entries (low 32 bits) and instructions (high 32 bits) so far
  for (int64_t s0 = 0; s0 < D; s0 += kBlock) {
    const int64_t s = s0 + threadIdx.x;
    uint64_t mine = 0;
    uint32_t c[4] = {0u, 0u, 0u, 0u};
    if (s < D) {
#pragma unroll
      for (int w = 0; w < 4; w++)
        c[w] = __hip_atomic_load(&segCnt[(base + s) * 4 + w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t tot = c[0] + c[1] + c[2] + c[3];
      segTotal[base + s] = tot;
      mine = (uint64_t)tot | ((uint64_t)((tot + 63u) / 64u) << 32);
    }
    uint64_t chunk = 0;
    const uint64_t incl = blockInclusiveScan(mine, &chunk);
    if (s < D) {
      const uint64_t ex = carry + incl - mine;
      const uint32_t st = (uint32_t)ex;
      segStart[base + s] = st;
      segInst[base + s] = (uint32_t)(ex >> 32);
      // the segment's entries are ordered by the wave that places them
      fill[(base + s) * 4 + 0] = st;
      fill[(base + s) * 4 + 1] = st + c[0];
      fill[(base + s) * 4 + 2] = st + c[0] + c[1];
      fill[(base + s) * 4 + 3] = st + c[0] + c[1] + c[2];
    }
    carry += chunk;
  }
  if (threadIdx.x == 0) {
    blockE[b] = (int64_t)(uint32_t)carry;
    blockI[b] = (int64_t)(carry >> 32);
  }
}

// pass 3: entries to their places, and the instruction descriptors
template <typename VT>
__global__ void __launch_bounds__(kBlock) kTileFill(const int64_t* rowPtr, const int32_t* cols, const VT* vals,
                                                    const uint32_t* perm, const uint32_t* pos, const uint32_t* bIdx,
                                                    const uint32_t* bPos, int64_t n, int nwords,
                                                    const int64_t* segBase, const uint32_t* segTile,
                                                    const uint32_t* segStart, const uint32_t* segInst,
                                                    const uint32_t* segTotal, uint32_t* fill,
                                                    const int64_t* blockE, const int64_t* blockI, uint32_t* tIdx,
                                                    VT* tVal, uint2* inst, int stageMin, const float* lo,
                                                    float* tLo) {
  __shared__ uint32_t bits[kTileMaxWords];
  __shared__ uint32_t pre[kTileMaxWords];
  __shared__ uint32_t fillL[kWaves][kTileLdsSegs];
  const int64_t b = blockIdx.x;
  tileBitmap(rowPtr, cols, perm, pos, bIdx, bPos, n, b, nwords, bits, pre);
  const int wave = threadIdx.x >> 6, lane = __lane_id();
  const int64_t base = segBase[b], eb = blockE[b];
  const bool inLds = segBase[b + 1] - base <= kTileLdsSegs;
  if (inLds)
    for (int64_t k = threadIdx.x; k < (segBase[b + 1] - base) * 4; k += kBlock) fillL[k & 3][k >> 2] = fill[base * 4 + k];
  __syncthreads();
  const int64_t m0 = b * kTileRows;
  const int rows = (int)(n - m0 < kTileRows ? n - m0 : kTileRows);
  for (int r = wave; r < rows; r += kWaves) {
    const int64_t I = perm ? (int64_t)perm[m0 + r] : m0 + r;
    const int64_t e0 = rowPtr[I], e1 = rowPtr[I + 1];
    const uint32_t bi = bIdx ? bIdx[I] : 0u;
    for (int64_t p0 = e0; p0 < e1; p0 += 64) {
      const int64_t p = p0 + lane;
      const bool valid = p < e1;
      uint32_t c = 0, s = 0;
      VT v = (VT)0;
      if (valid) {
        c = tileMapCol(cols[p], bi, pos, bIdx, bPos, n);
        s = tileRank(bits, pre, c >> kTileLog);
        v = vals[p];
      }
      // the entry's place in the block: the wave's fill cursor of the
      // segment, taken by an atomic add with return (lanes of one
      // instruction that share a cursor receive consecutive values in lane
      // order)
      uint32_t got = 0;
      if (valid) {
        if (inLds)
          got = atomicAdd(&fillL[wave][s], 1u);
        else
          got = atomicAdd(&fill[(base + s) * 4 + wave], 1u);
      }
      if (valid) {
        const int64_t q = eb + got;
        tIdx[q] = ((uint32_t)r << 16) | (c & (uint32_t)(kTileCols - 1));
        tVal[q] = v;
        if (lo) tLo[q] = lo[p];
      }
    }
  }
  const int64_t D = segBase[b + 1] - base, ib = blockI[b];
  for (int64_t s = threadIdx.x; s < D; s += kBlock) {
    const uint32_t tot = segTotal[base + s], st = segStart[base + s], tile = segTile[base + s];
    const int64_t i0 = ib + segInst[base + s];
    const uint32_t big = (tot + 63u) / 64u >= (uint32_t)stageMin ? 0x80000000u : 0u;
    for (uint32_t j = 0; 64u * j < tot; j++) {
      const uint32_t cnt = tot - 64u * j < 64u ? tot - 64u * j : 64u;
      inst[i0 + j] = make_uint2(st + 64u * j, big | (tile << 7) | cnt);
    }
  }
}

// pass 4: per entry, its distance to the first entry of its row run within
// its instruction and whether it is the run's last entry there (one wave
// per instruction)
__global__ void __launch_bounds__(kBlock) kTileRunBits(const uint2* __restrict__ inst,
                                                       const int64_t* __restrict__ blockI,
                                                       const int64_t* __restrict__ blockE, uint32_t* tIdx) {
  const int64_t b = blockIdx.x;
  const int wave = threadIdx.x >> 6, lane = __lane_id();
  const int64_t i1 = blockI[b + 1], e0 = blockE[b];
  for (int64_t i = blockI[b] + wave; i < i1; i += kWaves) {
    const uint2 d = inst[i];
    const int cnt = (int)(d.y & 127u);
    const bool valid = lane < cnt;
    const int64_t p = e0 + d.x + lane;
    const uint32_t ix = valid ? tIdx[p] : 0u;
    const int key = valid ? (int)((ix >> 16) & (kTileRows - 1)) : -1 - lane;
    const int prev = __shfl_up(key, 1, 64), next = __shfl_down(key, 1, 64);
    const bool start = lane == 0 || prev != key;
    const bool last = lane == 63 || next != key;
    const unsigned long long starts = __ballot(start);
    // the highest run start at or below this lane
    const unsigned long long upTo = starts & (~0ull >> (63 - lane));
    const int first = 63 - __clzll(upTo);
    const uint32_t dist = (uint32_t)(lane - first);
    if (valid)
      tIdx[p] = (ix & ~((63u << kTileDistShift) | 0x80000000u)) | (dist << kTileDistShift) | (last ? 0x80000000u : 0u);
  }
}

// One step of a segmented inclusive scan across the wave by DPP: lane i
// adds the value of the lane the DPP control CTRL names (within the rows of
// 16 lanes ROWMASK enables) when that lane lies in its run, i.e. when the
// lane's distance to the first lane of its run is at least `need` (need is
// larger than any distance where CTRL brings no value into the lane, so the
// value a disabled or out-of-row lane receives is never used).  Runs are
// contiguous, so the steps row_shr 1, 2, 4, 8 (distance counted within the
// row of 16), then row_bcast 15 and row_bcast 31 give every lane the sum of
// its run up to itself.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ void segScanStep(double& s, int dist, int need) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(s), CTRL, ROWMASK, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(s), CTRL, ROWMASK, 0xf, false);
  if (dist >= need) s += __hiloint2double(hi, lo);
}

// dist: the lane's distance to the first lane of its run (runs contiguous);
// c15, c31: per-lane thresholds of the two broadcast steps (tileScanConsts)
__device__ __forceinline__ double segInclusiveScan(double s, int dist, int c15, int c31) {
  const int inRow = dist < (int)(__lane_id() & 15u) ? dist : (int)(__lane_id() & 15u);
  segScanStep<0x111, 0xf>(s, inRow, 1);  // row_shr:1
  segScanStep<0x112, 0xf>(s, inRow, 2);  // row_shr:2
  segScanStep<0x114, 0xf>(s, inRow, 4);  // row_shr:4
  segScanStep<0x118, 0xf>(s, inRow, 8);  // row_shr:8
  segScanStep<0x142, 0xa>(s, dist, c15);  // row_bcast:15 (lane 15 into row 1, lane 47 into row 3)
  segScanStep<0x143, 0xc>(s, dist, c31);  // row_bcast:31 (lane 31 into rows 2 and 3)
  return s;
}

// the broadcast steps' thresholds: a lane of row 1 or 3 takes the previous
// row's last lane when that lane is in its run (distance > its index in the
// row); a lane of row 2 or 3 takes lane 31 when the distance reaches back to
// it; other lanes never take the broadcast value
__device__ __forceinline__ void tileScanConsts(int& c15, int& c31) {
  const int lane = (int)__lane_id();
  c15 = ((lane >> 4) & 1) ? (lane & 15) + 1 : 64;
  c31 = lane >= 32 ? lane - 31 : 64;
}

// y[perm[m]] = diag .* x + (H x) over the tiled copy, one vector; xP holds x
// in the permuted order of the copy's columns
// One x tile (kTileCols values from xP at tile << kTileLog, nx values in all)
// into LDS by one wave, in 16-byte pieces; the last tile's values past nx
// read as 0 (no entry points at them)
template <typename XT>
__device__ __forceinline__ void stageTile(const XT* __restrict__ xP, int64_t nx, uint32_t tile, XT* xs) {
  constexpr int per = 16 / (int)sizeof(XT);  // values per 16-byte piece
  const int lane = __lane_id();
  const int64_t base = (int64_t)tile << kTileLog;
  if (base + kTileCols <= nx) {
    typedef XT V __attribute__((ext_vector_type(per)));
#pragma unroll
    for (int k = 0; k < kTileCols / (64 * per); k++) {
      const int j = (k * 64 + lane) * per;
      *(V*)(xs + j) = *(const V*)(xP + base + j);
    }
  } else {
    for (int j = lane; j < kTileCols; j += 64) xs[j] = base + j < nx ? xP[base + j] : (XT)0;
  }
}

// y[perm[m]] = diag .* x + (H x) over the tiled copy, one vector; xP holds x
// in the permuted order of the copy's columns (nx values).  STAGE: each wave
// takes a contiguous quarter of the block's instructions, and before an
// instruction of a large segment (descriptor bit 31) gathers, the wave
// copies that segment's x tile into its own LDS (once per tile) and gathers
// from there; otherwise the waves take turns of U instructions and gather
// from global memory.
template <typename VT, typename XT, bool PERSIST, int U, bool STAGE>
__global__ void __launch_bounds__(kBlock) kSigmaTiled(const uint2* __restrict__ inst, const int64_t* __restrict__ blockI,
                                                      const int64_t* __restrict__ blockE,
                                                      const uint32_t* __restrict__ tIdx, const VT* __restrict__ tVal,
                                                      const double* __restrict__ diag, const uint32_t* __restrict__ perm,
                                                      int64_t n, int64_t nblocks, const XT* __restrict__ xP, int64_t nx,
                                                      const double* __restrict__ x, double* __restrict__ y,
                                                      unsigned* counter) {
  __shared__ double yw[kWaves][kTileRows];
  __shared__ __align__(16) XT xs[STAGE ? kWaves : 1][STAGE ? kTileCols : 1];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = __lane_id();
  static_assert(U <= kTileInstPad, "step larger than the descriptor padding");
  int c15, c31;
  tileScanConsts(c15, c31);
  uint32_t staged = 0xffffffffu;  // the tile in this wave's LDS (x does not change within a launch)
  ChunkSched<1, PERSIST> sched{nblocks, counter, -1};
  while (sched.next()) {
    const int64_t b = sched.cur;
    for (int j = lane; j < kTileRows; j += 64) yw[wave][j] = 0.0;
    const int64_t i0 = blockI[b], i1 = blockI[b + 1], e0 = blockE[b];
    const int64_t len = i1 - i0;
    const int64_t w0 = STAGE ? i0 + len * wave / kWaves : i0 + (int64_t)wave * U;
    const int64_t wEnd = STAGE ? i0 + len * (wave + 1) / kWaves : i1;
    // the U instructions' loads are issued without branches, all matrix loads
    // before all gathers (vector loads return in order, so a gather issued
    // between them would make every later wait include it); absent
    // instructions and lanes past an instruction's count read the block's
    // first entry and their tile's first x value, and are dropped
    // each step takes U consecutive instructions (one scalar load of their
    // descriptors; the array has U zero descriptors of padding)
    for (int64_t i = w0; i < wEnd; i += STAGE ? (int64_t)U : (int64_t)kWaves * U) {
      uint32_t dx[U], dy[U];
      const uint2* dp = inst + i;
#pragma unroll
      for (int u = 0; u < U; u++) {
        const uint2 d = dp[u];
        dx[u] = d.x;
        dy[u] = i + u < wEnd ? d.y : (d.y & ~127u);  // count 0 past the range
      }
      uint32_t ix[U];
      double v[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const bool valid = (uint32_t)lane < (dy[u] & 127u);
        const int64_t p = e0 + (valid ? (int64_t)dx[u] + lane : 0);
        ix[u] = __builtin_nontemporal_load(tIdx + p);
        v[u] = (double)__builtin_nontemporal_load(tVal + p);
      }
      double xv[U];
      if (STAGE) {
        // the tile of the step's first instruction from a large segment
        uint32_t want = 0xffffffffu;
#pragma unroll
        for (int u = 0; u < U; u++)
          if (want == 0xffffffffu && (dy[u] & 127u) && (dy[u] >> 31)) want = (dy[u] >> 7) & kTileDescMask;
        if (want != 0xffffffffu && want != staged) {
          stageTile<XT>(xP, nx, want, &xs[wave][0]);
          staged = want;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
          const bool valid = (uint32_t)lane < (dy[u] & 127u);
          const uint32_t col = valid ? (ix[u] & (uint32_t)(kTileCols - 1)) : 0u;
          const uint32_t tile = (dy[u] >> 7) & kTileDescMask;
          if (tile == staged)
            xv[u] = (double)xs[wave][col];
          else
            xv[u] = (double)xP[((size_t)tile << kTileLog) + col];
        }
      } else {
#pragma unroll
        for (int u = 0; u < U; u++) {
          const bool valid = (uint32_t)lane < (dy[u] & 127u);
          const size_t tb = (size_t)((dy[u] >> 7) & kTileDescMask) << kTileLog;
          xv[u] = (double)xP[tb + (valid ? (ix[u] & (uint32_t)(kTileCols - 1)) : 0u)];
        }
      }
      // products summed per row run by the segmented scan; the run's last
      // lane adds the sum into the wave's partial row (lanes past the count
      // sit above every valid lane, so no valid lane takes their values, and
      // are dropped)
#pragma unroll
      for (int u = 0; u < U; u++) {
        const bool valid = (uint32_t)lane < (dy[u] & 127u);
        const double sum = segInclusiveScan(v[u] * xv[u], (int)((ix[u] >> kTileDistShift) & 63u), c15, c31);
        if (valid && (int)ix[u] < 0) yw[wave][(ix[u] >> 16) & (kTileRows - 1)] += sum;
      }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kTileRows; t += kBlock) {
      const int64_t m = b * kTileRows + t;
      if (m < n) {
        const int64_t I = perm[m];
        y[I] = fma(diag[I], x[I], ((yw[0][t] + yw[1][t]) + yw[2][t]) + yw[3][t]);
      }
    }
    if (!PERSIST) break;
  }
}

// Half storage over the tiled copy: the copy holds each stored pair (I, J)
// once, in row I's block and J's column tile, and sigma needs each entry
// twice: v x[J] into row I and v x[I] into row J.  Both go into one
// accumulator per row, yT (the permuted order of x: row I's sum is split
// between its alpha-major and, with split columns, beta-major positions),
// kept in 64-bit fixed point so that the order of the sums does not matter
// and the result is the same on every run: a value t becomes the integer
// nearest t 2^-e, with 2^(e+61) above B max|x| (B: the largest off-diagonal
// row sum of |H_IJ|), so no row's sum of integers reaches 2^62 in
// magnitude, and each integer is off by at most 2^(e-1) (below
// 2^-61 B max|x|).  Row parts: each instruction's products are summed per
// row run in FP64 by the segmented scan, and each run's sum is added as an
// integer into the block's row sums in LDS, which go to yT at the end.
// Transposed parts: each wave sums its current tile's integers in LDS
// (segments with the large-segment bit; the others add theirs into yT
// directly) and adds the nonzero sums into yT with 64-bit integer atomics
// when its tile changes.  Each wave takes a contiguous quarter of its
// block's instructions.  yT: nx sums (cleared first), then the bits of
// max|x|; bound: the bits of B.
__device__ __forceinline__ int symFixExp(const unsigned long long* yT, int64_t nx, const unsigned long long* bound) {
  const double S = __longlong_as_double((long long)yT[nx]) * __longlong_as_double((long long)*bound);
  return S > 0.0 && S <= DBL_MAX ? ilogb(S) + 1 - 61 : 0;  // (x holding inf: no scale is right)
}

template <typename VT, typename XT, int U, bool LO>
__global__ void __launch_bounds__(kBlock) kSigmaTiledSym(const uint2* __restrict__ inst,
                                                         const int64_t* __restrict__ blockI,
                                                         const int64_t* __restrict__ blockE,
                                                         const uint32_t* __restrict__ tIdx,
                                                         const VT* __restrict__ tVal,
                                                         const float* __restrict__ tLo, int64_t n, int64_t nx,
                                                         const XT* __restrict__ xP, unsigned long long* yT,
                                                         const unsigned long long* __restrict__ bound) {
  __shared__ unsigned long long yr[kTileRows];          // the block's row sums
  __shared__ unsigned long long ya[kWaves][kTileCols];  // the wave's current tile of transposed sums
  __shared__ double xr[kTileRows];                      // x of the block's rows, times 2^-e
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = __lane_id();
  static_assert(U <= kTileInstPad, "step larger than the descriptor padding");
  int c15, c31;
  tileScanConsts(c15, c31);
  const int e = symFixExp(yT, nx, bound);
  const int64_t b = blockIdx.x;
  const int64_t m0 = b * kTileRows;
  for (int j = lane; j < kTileCols; j += 64) ya[wave][j] = 0ull;
  for (int t = threadIdx.x; t < kTileRows; t += kBlock) {
    yr[t] = 0ull;
    xr[t] = m0 + t < n ? ldexp((double)xP[m0 + t], -e) : 0.0;
  }
  __syncthreads();
  uint32_t cur = 0xffffffffu;
  auto flush = [&](uint32_t tile) {
    for (int j = lane; j < kTileCols; j += 64) {
      const unsigned long long a = ya[wave][j];
      if (a != 0ull) {
        atomicAdd(&yT[((size_t)tile << kTileLog) + j], a);
        ya[wave][j] = 0ull;
      }
    }
  };
  const int64_t i0 = blockI[b], i1 = blockI[b + 1], e0 = blockE[b];
  const int64_t len = i1 - i0;
  const int64_t w0 = i0 + len * wave / kWaves, wEnd = i0 + len * (wave + 1) / kWaves;
  for (int64_t i = w0; i < wEnd; i += U) {
    uint32_t dx[U], dy[U];
    const uint2* dp = inst + i;
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint2 d = dp[u];
      dx[u] = d.x;
      dy[u] = i + u < wEnd ? d.y : (d.y & ~127u);  // count 0 past the range
    }
    uint32_t ix[U];
    double v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const bool valid = (uint32_t)lane < (dy[u] & 127u);
      const int64_t p = e0 + (valid ? (int64_t)dx[u] + lane : 0);
      ix[u] = __builtin_nontemporal_load(tIdx + p);
      v[u] = (double)__builtin_nontemporal_load(tVal + p);
      if (LO) v[u] += (double)__builtin_nontemporal_load(tLo + p);
    }
    double xv[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const bool valid = (uint32_t)lane < (dy[u] & 127u);
      const size_t tb = (size_t)((dy[u] >> 7) & kTileDescMask) << kTileLog;
      xv[u] = (double)xP[tb + (valid ? (ix[u] & (uint32_t)(kTileCols - 1)) : 0u)];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if ((dy[u] & 127u) == 0) continue;
      const bool valid = (uint32_t)lane < (dy[u] & 127u);
      const uint32_t row = (ix[u] >> 16) & (kTileRows - 1);
      const double sum = segInclusiveScan(v[u] * xv[u], (int)((ix[u] >> kTileDistShift) & 63u), c15, c31);
      if (valid && (int)ix[u] < 0) atomicAdd(&yr[row], (unsigned long long)__double2ll_rn(ldexp(sum, -e)));
      const uint32_t tile = (dy[u] >> 7) & kTileDescMask;
      const uint32_t col = ix[u] & (uint32_t)(kTileCols - 1);
      const unsigned long long q = (unsigned long long)__double2ll_rn(v[u] * xr[row]);
      if (!(dy[u] >> 31)) {
        if (valid && q != 0ull) atomicAdd(&yT[((size_t)tile << kTileLog) + col], q);
        continue;
      }
      if (tile != cur) {
        if (cur != 0xffffffffu) flush(cur);
        cur = tile;
      }
      if (valid) atomicAdd(&ya[wave][col], q);
    }
  }
  if (cur != 0xffffffffu) flush(cur);
  __syncthreads();
  for (int t = threadIdx.x; t < kTileRows; t += kBlock)
    if (m0 + t < n && yr[t] != 0ull) atomicAdd(&yT[m0 + t], yr[t]);
}

// y[I] = diag[I] x[I] + 2^e (yT[pos[I]] (+ yT[n + bPos[I]] with split
// columns))
__global__ void kSymFinish(const double* __restrict__ diag, const double* __restrict__ x, double* __restrict__ y,
                           const unsigned long long* __restrict__ yT, const unsigned long long* __restrict__ bound,
                           const uint32_t* __restrict__ pos, const uint32_t* __restrict__ bPos, int64_t n,
                           int64_t nx) {
  const int e = symFixExp(yT, nx, bound);
  for (int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; I < n; I += (int64_t)gridDim.x * blockDim.x) {
    long long t = (long long)yT[pos[I]];
    if (bPos) t += (long long)yT[n + bPos[I]];
    y[I] = fma(diag[I], x[I], ldexp((double)t, e));
  }
}

__global__ void kFillZeroU64(unsigned long long* v, int64_t m) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = 0ull;
}

// xP = x in the permuted order (as kGatherPerm for one vector), and the
// bits of max|x| into *mx (cleared first; the bits of non-negative doubles
// order as their values)
template <typename XT>
__global__ void kGatherPermMax(const double* __restrict__ x, const uint32_t* __restrict__ perm,
                               const uint32_t* __restrict__ bPerm, int64_t n, XT* __restrict__ xP,
                               unsigned long long* mx) {
  const int64_t nx = bPerm ? 2 * n : n;
  double m = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nx; i += (int64_t)gridDim.x * blockDim.x) {
    const double v = x[i < n ? perm[i] : bPerm[i - n]];
    xP[i] = (XT)v;
    m = fmax(m, fabs(v));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  // one atomic per workgroup (the launch is capped at a few per CU):
  // atomics on one address are serialised
  __shared__ double wm[kWaves];
  if (__lane_id() == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kWaves; w++) m = fmax(m, wm[w]);
    if (m > 0.0) atomicMax(mx, (unsigned long long)__double_as_longlong(m));
  }
}

// off-diagonal row sums of |H| over a stored half grown by rows
// [nOld, n): row I's stored entries add into acc[I] and each into its
// column's acc[J]; every |v| is rounded up to a multiple of 2^-32 and
// summed as an integer (exactly, so in any order).  One wave per row.
template <typename VT>
__global__ void kRowAbsSum(const int64_t* __restrict__ rowPtr, const int32_t* __restrict__ cols,
                           const VT* __restrict__ vals, int64_t nOld, int64_t n, unsigned long long* acc) {
  const int lane = __lane_id();
  for (int64_t I = nOld + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64; I < n;
       I += ((int64_t)gridDim.x * blockDim.x) / 64) {
    unsigned long long r = 0ull;
    for (int64_t p = rowPtr[I] + lane; p < rowPtr[I + 1]; p += 64) {
      const unsigned long long q = (unsigned long long)ceil(ldexp(fabs((double)vals[p]), 32));
      atomicAdd(&acc[cols[p]], q);
      r += q;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) r += __shfl_xor(r, o, 64);
    if (lane == 0 && r) atomicAdd(&acc[I], r);
  }
}

__global__ void kMaxU64(const unsigned long long* __restrict__ a, int64_t n, unsigned long long* out) {
  unsigned long long m = 0ull;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    m = a[i] > m ? a[i] : m;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long t = __shfl_xor(m, o, 64);
    m = t > m ? t : m;
  }
  if (__lane_id() == 0 && m) atomicMax(out, m);
}

__global__ void kU64ScaledToDouble(unsigned long long* v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *v = (unsigned long long)__double_as_longlong(ldexp((double)*v, -32));
}

// SpMM: K interleaved vectors per gather (one K*sizeof(XT)-byte load per
// stored element serves all K products); rows read once.
template <typename XT, int K>
struct XVec {
  typedef XT T __attribute__((ext_vector_type(K)));
};

template <typename VT, typename XT, int K, int CHUNK, bool PERSIST>
__global__ void __launch_bounds__(kBlock) kSpmmPerm(const int64_t* __restrict__ rp, const int32_t* __restrict__ cols,
                                                    const VT* __restrict__ vals, const double* __restrict__ diag,
                                                    const uint32_t* __restrict__ perm, int64_t n,
                                                    const XT* __restrict__ xP, const double* __restrict__ x,
                                                    double* __restrict__ y, int k, int64_t ld, unsigned* counter) {
  typedef typename XVec<XT, K>::T XV;
  const int wave = threadIdx.x >> 6, lane = __lane_id();
  constexpr int U = K >= 8 ? 1 : 8 / K;
  ChunkSched<CHUNK, PERSIST> sched{(n + CHUNK - 1) / CHUNK, counter, -1};
  while (sched.next()) {
    const int64_t r0 = sched.cur * CHUNK;
    const int64_t r1 = r0 + CHUNK < n ? r0 + CHUNK : n;
    for (int64_t m = r0 + wave; m < r1; m += kWaves) {
      const int64_t b = rp[m], e = rp[m + 1];
      const int64_t I = perm[m];
      double acc[K];
#pragma unroll
      for (int r = 0; r < K; r++) acc[r] = 0.0;
      int64_t p = b + lane;
      for (; p + 64 * (U - 1) < e; p += 64 * U) {
        int32_t c[U];
        double v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          c[u] = __builtin_nontemporal_load(cols + p + 64 * u);
          v[u] = (double)__builtin_nontemporal_load(vals + p + 64 * u);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
          const XV xv = *reinterpret_cast<const XV*>(xP + (int64_t)c[u] * K);
#pragma unroll
          for (int r = 0; r < K; r++) acc[r] = fma(v[u], (double)xv[r], acc[r]);
        }
      }
      for (; p < e; p += 64) {
        const XV xv = *reinterpret_cast<const XV*>(xP + (int64_t)cols[p] * K);
        const double v = (double)vals[p];
#pragma unroll
        for (int r = 0; r < K; r++) acc[r] = fma(v, (double)xv[r], acc[r]);
      }
#pragma unroll
      for (int r = 0; r < K; r++) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc[r] += __shfl_xor(acc[r], o, 64);
      }
      if (lane < k && lane < K) {
        double sum = acc[0];
#pragma unroll
        for (int r = 1; r < K; r++)
          if (lane == r) sum = acc[r];
        y[lane * ld + I] = fma(diag[I], x[lane * ld + I], sum);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Davidson vector kernels

constexpr int kRedBlocks = 512;
// the PT reductions evaluate E_a of every unique external determinant: enough
// workgroups to keep every SIMD busy with latency-bound work
constexpr int kPtRedBlocks = 8192;

// per-block partial Gram entries, then a fixed-order final sum
template <int MA, int MB>
__global__ void __launch_bounds__(kBlock) kGramPartial(const double* A, const double* B, int ma, int mb, int64_t n,
                                                       int64_t ld, double* partial) {
  __shared__ double red[kWaves][MA * MB];
  double acc[MA * MB];
#pragma unroll
  for (int q = 0; q < MA * MB; q++) acc[q] = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    double a[MA], bb[MB];
#pragma unroll
    for (int i = 0; i < MA; i++) a[i] = i < ma ? A[i * ld + k] : 0.0;
#pragma unroll
    for (int j = 0; j < MB; j++) bb[j] = j < mb ? B[j * ld + k] : 0.0;
#pragma unroll
    for (int i = 0; i < MA; i++)
#pragma unroll
      for (int j = 0; j < MB; j++) acc[i * MB + j] = fma(a[i], bb[j], acc[i * MB + j]);
  }
  const int wave = threadIdx.x >> 6, lane = __lane_id();
#pragma unroll
  for (int q = 0; q < MA * MB; q++) {
    double v = acc[q];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) red[wave][q] = v;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < MA * MB; q += blockDim.x) {
    double v = 0.0;
    for (int w = 0; w < kWaves; w++) v += red[w][q];
    partial[(int64_t)blockIdx.x * MA * MB + q] = v;
  }
}

// one block per output: fixed-order strided partial sums, then a block tree
__global__ void __launch_bounds__(kBlock) kGramFinal(const double* partial, int nblocks, int MA, int MB, int ma,
                                                     int mb, double* G) {
  __shared__ double red[kWaves];
  const int q = blockIdx.x;
  const int i = q / mb, j = q % mb;
  double v = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) v += partial[(int64_t)b * MA * MB + i * MB + j];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (__lane_id() == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kWaves; w++) t += red[w];
    G[i * mb + j] = t;
  }
}

__global__ void kRotate(double* X, int m, int64_t n, int64_t ld, const double* V) {
  __shared__ double sV[16 * 16];
  for (int q = threadIdx.x; q < m * m; q += blockDim.x) sV[q] = V[q];
  __syncthreads();
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    double xin[16], xo[16];
    for (int l = 0; l < m; l++) xin[l] = X[l * ld + k];
    for (int j = 0; j < m; j++) {
      double s = 0.0;
      for (int l = 0; l < m; l++) s = fma(xin[l], sV[l * m + j], s);
      xo[j] = s;
    }
    for (int j = 0; j < m; j++) X[j * ld + k] = xo[j];
  }
}

// Gram tiles: blockIdx.y = tile (ta, tb) of 8 x 8 entries
__global__ void __launch_bounds__(kBlock) kGramTile(const double* A, int ma, const double* B, int mb, int64_t n,
                                                    int64_t ld, double* partial) {
  constexpr int T = 8;
  __shared__ double red[kWaves][T * T];
  const int tilesB = (mb + T - 1) / T;
  const int ta = blockIdx.y / tilesB, tb = blockIdx.y % tilesB;
  const int na = min(T, ma - ta * T), nb = min(T, mb - tb * T);
  const double* A0 = A + (int64_t)ta * T * ld;
  const double* B0 = B + (int64_t)tb * T * ld;
  double acc[T * T];
#pragma unroll
  for (int q = 0; q < T * T; q++) acc[q] = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    double a[T], bb[T];
#pragma unroll
    for (int i = 0; i < T; i++) a[i] = i < na ? A0[i * ld + k] : 0.0;
#pragma unroll
    for (int j = 0; j < T; j++) bb[j] = j < nb ? B0[j * ld + k] : 0.0;
#pragma unroll
    for (int i = 0; i < T; i++)
#pragma unroll
      for (int j = 0; j < T; j++) acc[i * T + j] = fma(a[i], bb[j], acc[i * T + j]);
  }
  const int wave = threadIdx.x >> 6, lane = __lane_id();
#pragma unroll
  for (int q = 0; q < T * T; q++) {
    double v = acc[q];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) red[wave][q] = v;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < T * T; q += blockDim.x) {
    double v = 0.0;
    for (int w = 0; w < kWaves; w++) v += red[w][q];
    partial[((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * T * T + q] = v;
  }
}

// one block per output entry (i, j): fixed-order sum over the tile's partials
__global__ void __launch_bounds__(kBlock) kGramTileFinal(const double* partial, int nblocks, int ma, int mb,
                                                         double* G) {
  constexpr int T = 8;
  __shared__ double red[kWaves];
  const int i = blockIdx.x / mb, j = blockIdx.x % mb;
  const int tilesB = (mb + T - 1) / T;
  const int tile = (i / T) * tilesB + j / T;
  const int q = (i % T) * T + (j % T);
  double v = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) v += partial[((int64_t)tile * nblocks + b) * T * T + q];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (__lane_id() == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kWaves; w++) t += red[w];
    G[i * mb + j] = t;
  }
}

// out = alpha * A W + beta * out, one row per thread, 8 output columns per pass
__global__ void __launch_bounds__(kBlock) kGemmTS(const double* A, int ma, const double* W, int mb, int64_t n,
                                                  int64_t lda, double* out, int64_t ldo, double alpha, double beta) {
  __shared__ double sW[64 * 64];
  for (int q = threadIdx.x; q < ma * mb; q += blockDim.x) sW[q] = W[q];
  __syncthreads();
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    for (int j0 = 0; j0 < mb; j0 += 8) {
      double acc[8];
#pragma unroll
      for (int jj = 0; jj < 8; jj++) acc[jj] = 0.0;
      for (int l = 0; l < ma; l++) {
        const double a = A[l * lda + k];
#pragma unroll
        for (int jj = 0; jj < 8; jj++)
          if (j0 + jj < mb) acc[jj] = fma(a, sW[l * mb + j0 + jj], acc[jj]);
      }
#pragma unroll
      for (int jj = 0; jj < 8; jj++)
        if (j0 + jj < mb) {
          double* o = out + (int64_t)(j0 + jj) * ldo + k;
          *o = beta == 0.0 ? alpha * acc[jj] : fma(beta, *o, alpha * acc[jj]);
        }
    }
  }
}

__device__ __forceinline__ double blockSum(double v) {
  __shared__ double red[kWaves];
  const int wave = threadIdx.x >> 6, lane = __lane_id();
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kWaves; w++) t += red[w];
  __syncthreads();
  return t;
}

__global__ void __launch_bounds__(kBlock) kResidual(const double* sig, const double* b, const double* e, int64_t n,
                                                    double* r, double* partial) {
  const double ev = e[0];
  double acc = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    double v = sig[k] - ev * b[k];
    r[k] = v;
    acc = fma(v, v, acc);
  }
  double t = blockSum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// one block per output value (blockIdx.x = q < nvals)
__global__ void __launch_bounds__(kBlock) kFinalSum(const double* partial, int nb, int stride, int nvals,
                                                    double* out) {
  __shared__ double red[kWaves];
  const int q = blockIdx.x;
  double v = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) v += partial[(int64_t)b * stride + q];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (__lane_id() == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kWaves; w++) t += red[w];
    out[q] = t;
  }
}

// Semistochastic reduction (SHCI/SHCIbasics.cpp:264-385): per run of equal
// determinants (sorted, generation order kept), a1/a2 = sums of num / num2,
// b1/b2 = the same over the terms present at epsilon2Large; per-block partial
// sums of (|a1|^2 Nmc/(Nmc-1) - a2) / (E_a - E0) and of the b version.
// The energy tables of E_a in LDS: hdiag (n), direct (n^2), exchange (n^2)
// when they fit (LDS_TAB), else read from global memory.
constexpr size_t kPtTabMaxBytes = 64 * 1024;
inline size_t ptTabBytes(int n) { return ((size_t)n + 2 * (size_t)n * n) * sizeof(double); }

template <int W, bool LDS_TAB>
__device__ __forceinline__ double ptEnergy(const DevSystem& S, const double* tab, const uint64_t* d) {
  // (the same additions in the same order as Energy(): E_a is bit-identical,
  // which matters for external determinants whose E0 - E_a is small)
  if (LDS_TAB) return detEnergyTab<W>(tab, tab + S.norb, tab + S.norb + S.norb * S.norb, S.norb, d);
  return detEnergy<W>(S, d);
}

template <bool LDS_TAB>
__device__ __forceinline__ double* ptStageTables(const DevSystem& S) {
  extern __shared__ double tab[];
  if (LDS_TAB) {
    const int n = S.norb, nn = n * n;
    for (int k = threadIdx.x; k < n + 2 * nn; k += blockDim.x)
      tab[k] = k < n ? S.h1[k * n + k] : (k < n + nn ? S.direct[k - n] : S.exch[k - n - nn]);
    __syncthreads();
  }
  return tab;
}

template <int W, bool LDS_TAB>
__global__ void __launch_bounds__(kBlock) kPtStochReduce(DevSystem S, const uint64_t* keys, const uint32_t* idx,
                                                         const double* num, const double* num2, const uint8_t* pres,
                                                         int64_t m, double e0, double nmc, double* partial) {
  const double* tab = ptStageTables<LDS_TAB>(S);
  double sA = 0.0, sB = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    if (i > 0 && detEqual<W>(keys + i * W, keys + (i - 1) * W)) continue;
    double a1 = 0.0, a2 = 0.0, b1 = 0.0, b2 = 0.0;
    for (int64_t j = i; j < m && detEqual<W>(keys + j * W, keys + i * W); j++) {
      const uint32_t q = idx[j];
      a1 = __dadd_rn(a1, num[q]);
      a2 = __dadd_rn(a2, num2[q]);
      if (pres[q]) {
        b1 = __dadd_rn(b1, num[q]);
        b2 = __dadd_rn(b2, num2[q]);
      }
    }
    uint64_t d[W];
#pragma unroll
    for (int w = 0; w < W; w++) d[w] = keys[i * W + w];
    const double den = __dsub_rn(ptEnergy<W, LDS_TAB>(S, tab, d), e0);
    sA = __dadd_rn(sA, __ddiv_rn(__dsub_rn(__ddiv_rn(__dmul_rn(__dmul_rn(fabs(a1), fabs(a1)), nmc), nmc - 1.0), a2),
                                 den));
    sB = __dadd_rn(sB, __ddiv_rn(__dsub_rn(__ddiv_rn(__dmul_rn(__dmul_rn(fabs(b1), fabs(b1)), nmc), nmc - 1.0), b2),
                                 den));
  }
  const double t0 = blockSum(sA), t1 = blockSum(sB);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = t0;
    partial[2 * blockIdx.x + 1] = t1;
  }
}

__global__ void kIota(uint32_t* v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = (uint32_t)i;
}

// Over candidates sorted by determinant (stable, so equal determinants keep
// their generation order): the first entry of each run sums the run's
// numerators in that order (the reference's merge) and contributes
// num^2 / (E0 - E_a) and (num / (E0 - E_a))^2; per-block partial sums of
// (E_PT, psi1 norm, number of external determinants) in fixed order.

// This is evaluated over the run starts only (compacted in order first), so
// that every lane of the E_a evaluation works on a unique determinant
// instead of one lane in ~6 (runs average six candidates on the 1.18M-
// determinant space).
constexpr int kStartChunk = kBlock * 16;

template <int W>
__global__ void __launch_bounds__(kBlock) kPtCountStarts(const uint64_t* keys, int64_t m, int64_t* counts) {
  const int64_t b0 = (int64_t)blockIdx.x * kStartChunk;
  const int64_t b1 = b0 + kStartChunk < m ? b0 + kStartChunk : m;
  int c = 0;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += blockDim.x)
    if (i == 0 || !detEqual<W>(keys + i * W, keys + (i - 1) * W)) c++;
  const double t = blockSum((double)c);
  if (threadIdx.x == 0) counts[blockIdx.x] = (int64_t)t;
}

template <int W>
__global__ void __launch_bounds__(kBlock) kPtWriteStarts(const uint64_t* keys, int64_t m, const int64_t* offs,
                                                         uint32_t* starts) {
  __shared__ int sWave[kWaves];
  __shared__ int64_t sBase;
  const int wave = threadIdx.x >> 6, lane = __lane_id();
  const int64_t b0 = (int64_t)blockIdx.x * kStartChunk;
  const int64_t b1 = b0 + kStartChunk < m ? b0 + kStartChunk : m;
  if (threadIdx.x == 0) sBase = offs[blockIdx.x];
  __syncthreads();
  for (int64_t t0 = b0; t0 < b1; t0 += blockDim.x) {
    const int64_t i = t0 + threadIdx.x;
    const bool f = i < b1 && (i == 0 || !detEqual<W>(keys + i * W, keys + (i - 1) * W));
    const uint64_t bal = __ballot(f);
    if (lane == 0) sWave[wave] = __popcll(bal);
    __syncthreads();
    int64_t base = sBase;
    for (int w = 0; w < wave; w++) base += sWave[w];
    if (f) starts[base + __popcll(bal & lanemaskLt())] = (uint32_t)i;
    __syncthreads();
    if (threadIdx.x == 0)
      for (int w = 0; w < kWaves; w++) sBase += sWave[w];
    __syncthreads();
  }
}

template <int W, bool LDS_TAB>
__global__ void __launch_bounds__(kBlock) kPtReduceStarts(DevSystem S, const uint64_t* keys, const uint32_t* idx,
                                                          const double* num, const uint32_t* starts, int64_t nu,
                                                          int64_t m, double e0, double* partial) {
  const double* tab = ptStageTables<LDS_TAB>(S);
  double sPt = 0.0, sPsi = 0.0, sCnt = 0.0;
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < nu; u += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = starts[u];
    double sum = num[idx[i]];
    for (int64_t j = i + 1; j < m && detEqual<W>(keys + j * W, keys + i * W); j++) sum = __dadd_rn(sum, num[idx[j]]);
    uint64_t d[W];
#pragma unroll
    for (int w = 0; w < W; w++) d[w] = keys[i * W + w];
    const double den = __dsub_rn(e0, ptEnergy<W, LDS_TAB>(S, tab, d));
    sPt = __dadd_rn(sPt, __ddiv_rn(__dmul_rn(fabs(sum), fabs(sum)), den));
    const double r = __ddiv_rn(sum, den);
    sPsi = __dadd_rn(sPsi, __dmul_rn(fabs(r), fabs(r)));
    sCnt += 1.0;
  }
  const double t0 = blockSum(sPt), t1 = blockSum(sPsi), t2 = blockSum(sCnt);
  if (threadIdx.x == 0) {
    partial[3 * blockIdx.x] = t0;
    partial[3 * blockIdx.x + 1] = t1;
    partial[3 * blockIdx.x + 2] = t2;
  }
}


__global__ void kPrecondition(double* r, const double* diag, const double* e, int64_t n) {
  const double ev = e[0];
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    double den = ev - diag[k];
    if (fabs(den) > 1e-12)
      r[k] = r[k] / den;
    else
      r[k] = r[k] / (den - 1.e-12);
  }
}

__global__ void __launch_bounds__(kBlock) kDot2(const double* b, const double* r, int64_t n, double* partial) {
  double s0 = 0.0, s1 = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    double bv = b[k];
    s0 = fma(bv, r[k], s0);
    s1 = fma(bv, bv, s1);
  }
  double t0 = blockSum(s0);
  double t1 = blockSum(s1);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = t0;
    partial[2 * blockIdx.x + 1] = t1;
  }
}

__global__ void kOrthoStep(double* r, const double* b, const double* d, int64_t n) {
  const double ov = d[0], nn = d[1];
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    r[k] = r[k] - ov * b[k] / nn;
}

__global__ void __launch_bounds__(kBlock) kNorm2(const double* x, int64_t n, double* partial) {
  double s = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    s = fma(x[k], x[k], s);
  double t = blockSum(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ void kScaleInto(double* dst, const double* src, const double* nrm2, int64_t n) {
  const double nr = sqrt(nrm2[0]);
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    dst[k] = src[k] / nr;
}

__global__ void kAxpy(double* x, const double* y, double alpha, int64_t n) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    x[k] = x[k] + alpha * y[k];
}

__global__ void kSymUpper(const double* G, int m, double* A) {
  for (int q = threadIdx.x; q < m * m; q += blockDim.x) {
    const int i = q / m, j = q % m;
    A[q] = i <= j ? G[i * m + j] : G[j * m + i];
  }
}

// Cyclic Jacobi for m <= kJacobiMax on one wavefront with the matrices in
// LDS: rotations are applied in the sequential (p, q) order, lane k updating
// element k of the two rows / columns, so every element sees the same
// operations in the same order as a one-thread sweep.
constexpr int kJacobiMax = 32;

__global__ void __launch_bounds__(64) kJacobi(const double* A, int m, double* w, double* V) {
  __shared__ double a[kJacobiMax * kJacobiMax], v[kJacobiMax * kJacobiMax];
  __shared__ double sOff;
  __shared__ int order[kJacobiMax];
  const int lane = threadIdx.x;
  for (int q = lane; q < m * m; q += 64) {
    a[q] = A[q];
    v[q] = 0.0;
  }
  __syncthreads();
  if (lane < m) v[lane * m + lane] = 1.0;
  __syncthreads();
  for (int sweep = 0; sweep < 100; sweep++) {
    if (lane == 0) {
      double off = 0.0;
      for (int p = 0; p < m; p++)
        for (int q = p + 1; q < m; q++) off += fabs(a[p * m + q]);
      sOff = off;
    }
    __syncthreads();
    if (sOff == 0.0) break;
    for (int p = 0; p < m; p++)
      for (int q = p + 1; q < m; q++) {
        const double apq = a[p * m + q], app = a[p * m + p], aqq = a[q * m + q];
        const double g = 100.0 * fabs(apq);
        if (sweep > 3 && fabs(app) + g == fabs(app) && fabs(aqq) + g == fabs(aqq)) {
          __syncthreads();
          if (lane == 0) {
            a[p * m + q] = 0.0;
            a[q * m + p] = 0.0;
          }
          __syncthreads();
          continue;
        }
        if (apq == 0.0) continue;
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
        __syncthreads();
        if (lane < m) {
          const double akp = a[lane * m + p], akq = a[lane * m + q];
          a[lane * m + p] = c * akp - sn * akq;
          a[lane * m + q] = sn * akp + c * akq;
        }
        __syncthreads();
        if (lane < m) {
          const double apk = a[p * m + lane], aqk = a[q * m + lane];
          a[p * m + lane] = c * apk - sn * aqk;
          a[q * m + lane] = sn * apk + c * aqk;
          const double vkp = v[lane * m + p], vkq = v[lane * m + q];
          v[lane * m + p] = c * vkp - sn * vkq;
          v[lane * m + q] = sn * vkp + c * vkq;
        }
        __syncthreads();
        if (lane == 0) {
          a[p * m + q] = 0.0;
          a[q * m + p] = 0.0;
        }
        __syncthreads();
      }
    __syncthreads();
  }
  if (lane == 0) {
    for (int i = 0; i < m; i++) order[i] = i;
    for (int i = 1; i < m; i++) {  // stable insertion sort by eigenvalue
      const int o = order[i];
      int j = i;
      while (j > 0 && a[order[j - 1] * m + order[j - 1]] > a[o * m + o]) {
        order[j] = order[j - 1];
        j--;
      }
      order[j] = o;
    }
  }
  __syncthreads();
  if (lane < m) w[lane] = a[order[lane] * m + order[lane]];
  for (int q = lane; q < m * m; q += 64) {
    const int k = q / m, j = q % m;
    V[k * m + j] = v[k * m + order[j]];
  }
}

// The same eigenproblem with the rotations of a sweep in round-robin
// (tournament) order: the m'/2 disjoint pairs of a step (m' = m rounded up
// to even) are rotated at once, m'-1 steps per sweep, so a sweep needs
// m'-1 rounds of barriers instead of m(m-1)/2.  Another rotation order than
// kJacobi's, so the eigenpairs agree with it to rounding, not bit for bit;
// used by the block Davidson.
__global__ void __launch_bounds__(64) kJacobiRR(const double* A, int m, double* w, double* V) {
  __shared__ double a[kJacobiMax * kJacobiMax], v[kJacobiMax * kJacobiMax];
  __shared__ double rc[kJacobiMax / 2], rs[kJacobiMax / 2];
  __shared__ int rp[kJacobiMax / 2], rq[kJacobiMax / 2];
  __shared__ double sOff;
  __shared__ int order[kJacobiMax];
  const int lane = threadIdx.x;
  const int mm = (m + 1) & ~1, half = mm / 2;
  for (int q = lane; q < m * m; q += 64) {
    a[q] = A[q];
    v[q] = 0.0;
  }
  __syncthreads();
  if (lane < m) v[lane * m + lane] = 1.0;
  __syncthreads();
  for (int sweep = 0; sweep < 100; sweep++) {
    double part = 0.0;
    for (int t = lane; t < m * m; t += 64) {
      const int p = t / m, q = t % m;
      if (q > p) part += fabs(a[t]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
    if (lane == 0) sOff = part;
    __syncthreads();
    if (sOff == 0.0) break;
    for (int step = 0; step < mm - 1; step++) {
      if (lane < half) {
        // round-robin: player 0 fixed, the others rotate by step
        const int i = lane, j = mm - 1 - lane;
        int p = i == 0 ? 0 : (i - 1 + step) % (mm - 1) + 1;
        int q = (j - 1 + step) % (mm - 1) + 1;
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        double c = 1.0, sn = 0.0;
        int pp = -1;
        if (q < m) {
          const double apq = a[p * m + q], app = a[p * m + p], aqq = a[q * m + q];
          const double g = 100.0 * fabs(apq);
          if (sweep > 3 && fabs(app) + g == fabs(app) && fabs(aqq) + g == fabs(aqq)) {
            pp = -2;  // negligible: zero it without a rotation
          } else if (apq != 0.0) {
            const double theta = (aqq - app) / (2.0 * apq);
            const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            c = 1.0 / sqrt(t * t + 1.0);
            sn = t * c;
            pp = p;
          }
        }
        rp[lane] = pp == -2 ? -2 - p : pp;
        rq[lane] = q;
        rc[lane] = c;
        rs[lane] = sn;
      }
      __syncthreads();
      // columns p, q of every row k (a), then rows p, q and columns of v
      for (int t = lane; t < half * m; t += 64) {
        const int r = t / m, k = t % m;
        const int p = rp[r], q = rq[r];
        if (p < 0) continue;
        const double c = rc[r], sn = rs[r];
        const double akp = a[k * m + p], akq = a[k * m + q];
        a[k * m + p] = c * akp - sn * akq;
        a[k * m + q] = sn * akp + c * akq;
      }
      __syncthreads();
      for (int t = lane; t < half * m; t += 64) {
        const int r = t / m, k = t % m;
        const int p = rp[r], q = rq[r];
        if (p < 0) continue;
        const double c = rc[r], sn = rs[r];
        const double apk = a[p * m + k], aqk = a[q * m + k];
        a[p * m + k] = c * apk - sn * aqk;
        a[q * m + k] = sn * apk + c * aqk;
        const double vkp = v[k * m + p], vkq = v[k * m + q];
        v[k * m + p] = c * vkp - sn * vkq;
        v[k * m + q] = sn * vkp + c * vkq;
      }
      __syncthreads();
      if (lane < half) {
        int p = rp[lane];
        const int q = rq[lane];
        if (p <= -2) p = -2 - p;  // negligible pair
        if (p >= 0 && q < m) {
          a[p * m + q] = 0.0;
          a[q * m + p] = 0.0;
        }
      }
      __syncthreads();
    }
  }
  if (lane == 0) {
    for (int i = 0; i < m; i++) order[i] = i;
    for (int i = 1; i < m; i++) {  // stable insertion sort by eigenvalue
      const int o = order[i];
      int j = i;
      while (j > 0 && a[order[j - 1] * m + order[j - 1]] > a[o * m + o]) {
        order[j] = order[j - 1];
        j--;
      }
      order[j] = o;
    }
  }
  __syncthreads();
  if (lane < m) w[lane] = a[order[lane] * m + order[lane]];
  for (int q = lane; q < m * m; q += 64) {
    const int k = q / m, j = q % m;
    V[k * m + j] = v[k * m + order[j]];
  }
}

}  // namespace

// ===========================================================================
// launchers

#define DISPATCH_W(W_, ...)                  \
  do {                                       \
    if ((W_) == 1) {                         \
      constexpr int WW = 1;                  \
      __VA_ARGS__;                           \
    } else if ((W_) == 2) {                  \
      constexpr int WW = 2;                  \
      __VA_ARGS__;                           \
    } else {                                 \
      throw HipError("unsupported det width"); \
    }                                        \
  } while (0)

void launchEnergy(const DevSystem& S, const uint64_t* dets, int64_t n, double* out, hipStream_t s) {
  if (n <= 0) return;
  DISPATCH_W(S.W, kEnergy<WW><<<gridFor(n), kBlock, 0, s>>>(S, dets, n, out));
  HIP_CHECK(hipGetLastError());
}

void launchHijPairs(const DevSystem& S, const uint64_t* bra, const uint64_t* ket, int64_t n, double* out,
                    hipStream_t s) {
  if (n <= 0) return;
  DISPATCH_W(S.W, kHijPairs<WW><<<gridFor(n), kBlock, 0, s>>>(S, bra, ket, n, out));
  HIP_CHECK(hipGetLastError());
}

void hashClear(int32_t* table, uint64_t slots, hipStream_t s) {
  kFillNeg<<<gridFor((int64_t)slots), kBlock, 0, s>>>(table, slots);
  HIP_CHECK(hipGetLastError());
}

void hashInsert(const DevSystem& S, const uint64_t* dets, int64_t begin, int64_t end, int32_t* table,
                uint64_t slots, hipStream_t s) {
  if (end <= begin) return;
  DISPATCH_W(S.W, kHashInsert<WW><<<gridFor(end - begin), kBlock, 0, s>>>(dets, begin, end, table, slots - 1));
  HIP_CHECK(hipGetLastError());
}

void launchCmax(const double* ci, int nroots, int64_t ld, int64_t n, double* cmax, hipStream_t s) {
  if (n <= 0) return;
  kCmax<<<gridFor(n), kBlock, 0, s>>>(ci, nroots, ld, n, cmax);
  HIP_CHECK(hipGetLastError());
}

void launchScreen(const DevSystem& S, const uint64_t* dets, int64_t begin, int64_t end, const double* cmax,
                  double eps1, int ncore, int nact, const int32_t* table, uint64_t slots, uint64_t* out,
                  unsigned long long* counter, int64_t cap, hipStream_t s) {
  if (end <= begin) return;
  unsigned g = gridFor(end - begin, kWaves, 1 << 16);
  DISPATCH_W(S.W, kScreen<WW, 0><<<g, kBlock, 0, s>>>(S, dets, begin, end, cmax, eps1, ncore, nact, table,
                                                         slots - 1, out, counter, cap, 1.0, nullptr));
  HIP_CHECK(hipGetLastError());
}

void launchPtGen(const DevSystem& S, const uint64_t* dets, int64_t n, const double* ci, double eps2,
                 const int32_t* table, uint64_t slots, uint32_t nparts, uint32_t part, int64_t* counts,
                 const int64_t* fillStart, uint64_t* outD, double* outN, hipStream_t s) {
  if (n <= 0) return;
  if (nparts == 0 || part >= nparts) throw HipError("launchPtGen: bad partition");
  unsigned g = gridFor(n, kWaves, 1 << 16);
  const PtStoch st{nullptr, nullptr, 0.0, 0.0, nullptr, nullptr};
  if (fillStart)
    DISPATCH_W(S.W, kPtGen<WW, true, false><<<g, kBlock, 0, s>>>(S, dets, n, ci, eps2, table, slots - 1, nparts, part,
                                                                   counts, fillStart, outD, outN, st));
  else
    DISPATCH_W(S.W, kPtGen<WW, false, false><<<g, kBlock, 0, s>>>(S, dets, n, ci, eps2, table, slots - 1, nparts,
                                                                    part, counts, nullptr, nullptr, nullptr, st));
  HIP_CHECK(hipGetLastError());
}

void launchPtGenStoch(const DevSystem& S, const uint64_t* dets, const uint32_t* src, const double* w, int64_t nsrc,
                      const double* ci, double eps2, double eps2Large, double nmc, const int32_t* table, uint64_t slots,
                      int64_t* counts, const int64_t* fillStart, uint64_t* outD, double* outN, double* outN2,
                      uint8_t* outP, hipStream_t s) {
  if (nsrc <= 0) return;
  unsigned g = gridFor(nsrc, kWaves, 1 << 16);
  const PtStoch st{src, w, eps2Large, nmc, outN2, outP};
  if (fillStart)
    DISPATCH_W(S.W, kPtGen<WW, true, true><<<g, kBlock, 0, s>>>(S, dets, nsrc, ci, eps2, table, slots - 1, 1u, 0u,
                                                                  counts, fillStart, outD, outN, st));
  else
    DISPATCH_W(S.W, kPtGen<WW, false, true><<<g, kBlock, 0, s>>>(S, dets, nsrc, ci, eps2, table, slots - 1, 1u, 0u,
                                                                   counts, nullptr, nullptr, nullptr, st));
  HIP_CHECK(hipGetLastError());
}

void launchPtStochReduce(const DevSystem& S, const uint64_t* keys, const uint32_t* idx, const double* num,
                         const double* num2, const uint8_t* pres, int64_t m, double e0, double nmc, DevMem& partial,
                         double* out, hipStream_t s) {
  const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(kPtRedBlocks, (m + kBlock - 1) / kBlock));
  partial.reserve((size_t)nb * 2 * sizeof(double));
  const size_t tb = ptTabBytes(S.norb);
  if (tb <= kPtTabMaxBytes)
    DISPATCH_W(S.W, kPtStochReduce<WW, true><<<nb, kBlock, tb, s>>>(S, keys, idx, num, num2, pres, m, e0, nmc,
                                                                    partial.as<double>()));
  else
    DISPATCH_W(S.W, kPtStochReduce<WW, false><<<nb, kBlock, 0, s>>>(S, keys, idx, num, num2, pres, m, e0, nmc,
                                                                     partial.as<double>()));
  kFinalSum<<<2, kBlock, 0, s>>>(partial.as<double>(), nb, 2, 2, out);
  HIP_CHECK(hipGetLastError());
}

size_t rdmScratchDoubles(int norb, int W) {
  const size_t n = norb, NSO = 64 * W;
  return NSO * NSO + n * n * n * n + 2 * n * n * n;
}

void launchRdm(int norb, int W, const uint64_t* dets, const int64_t* rp, const int32_t* cols, const double* bra,
               const double* ket, int64_t n, double* g1, double* g2, double* s1, double* s2, double* scratch,
               hipStream_t s) {
  if (norb > 32 * W) throw HipError("launchRdm: more orbitals than the determinant words hold");
  const int NSO = 64 * W;
  const int64_t nn = norb, n3 = nn * nn * nn;
  double* M = scratch;
  RdmOut o{g1, g2, s1, s2, M + (int64_t)NSO * NSO, M + (int64_t)NSO * NSO + n3 * nn,
           M + (int64_t)NSO * NSO + n3 * nn + n3, norb, 2 * norb, (int64_t)(2 * norb) * (2 * norb + 1) / 2,
           bra == ket};
  HIP_CHECK(hipMemsetAsync(scratch, 0, rdmScratchDoubles(norb, W) * 8, s));
  if (n > 0) {
    const unsigned go = (unsigned)std::min<int64_t>(std::max<int64_t>(n / 64, 1), 2048);
    DISPATCH_W(W, kRdmOcc<WW><<<go, kBlock, 0, s>>>(dets, bra, ket, n, M));
    const unsigned g = gridFor(n, kWaves, 1 << 16);
    DISPATCH_W(W, kRdm<WW><<<g, kBlock, 0, s>>>(o, dets, rp, cols, bra, ket, n));
  }
  kRdmFinish<<<gridFor(n3 * nn), kBlock, 0, s>>>(o);
  kRdmOccScatter<<<gridFor((int64_t)o.nso * o.nso), kBlock, 0, s>>>(o, M, NSO);
  HIP_CHECK(hipGetLastError());
}

int64_t launchTrevBasis(const uint64_t* dets, int64_t i0, int64_t i1, int W, DevMem& flags, uint64_t* block,
                        int64_t base, int64_t* eIdx, int64_t* partner, double* parity, Workspace& ws, hipStream_t s) {
  const int64_t cnt = i1 - i0;
  if (cnt <= 0) return 0;
  flags.reserve((size_t)(cnt + 1) * 8, s);
  int64_t* f = flags.as<int64_t>();
  HIP_CHECK(hipMemsetAsync(f + cnt, 0, 8, s));
  DISPATCH_W(W, kTrevMark<WW><<<gridFor(cnt), kBlock, 0, s>>>(dets, i0, i1, f));
  exclusiveScanI64(f, cnt + 1, ws, s);
  int64_t nu = 0;
  HIP_CHECK(hipMemcpyAsync(&nu, f + cnt, 8, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  if (block)
    DISPATCH_W(W, kTrevBlock<WW><<<gridFor(cnt), kBlock, 0, s>>>(dets, i0, i1, f, base, block, eIdx, partner, parity));
  HIP_CHECK(hipGetLastError());
  return nu;
}

void launchTrevExpand(const double* x, int64_t n, int k, int64_t ld, const int64_t* eIdx, const int64_t* partner,
                      const double* parity, double* xE, int64_t ldE, hipStream_t s) {
  if (n <= 0) return;
  kTrevExpand<<<gridFor(n * k), kBlock, 0, s>>>(x, n, k, ld, eIdx, partner, parity, xE, ldE);
  HIP_CHECK(hipGetLastError());
}

void launchTrevContract(const double* yE, int64_t ldE, int64_t n, int k, const int64_t* eIdx, const int64_t* partner,
                        const double* parity, double* y, int64_t ld, hipStream_t s) {
  if (n <= 0) return;
  kTrevContract<<<gridFor(n * k), kBlock, 0, s>>>(yE, ldE, n, k, eIdx, partner, parity, y, ld);
  HIP_CHECK(hipGetLastError());
}

void launchTrevUnpack(const double* ci, int64_t n, int nroots, int64_t ld, const int64_t* partner,
                      const double* parity, int trev, double* out, int64_t ldE, hipStream_t s) {
  if (n <= 0) return;
  kTrevUnpack<<<gridFor(n * nroots), kBlock, 0, s>>>(ci, n, nroots, ld, partner, parity, trev, out, ldE);
  HIP_CHECK(hipGetLastError());
}

static int computeUnits();

void launchIota(uint32_t* v, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  kIota<<<gridFor(n, kBlock, 8192), kBlock, 0, s>>>(v, n);
  HIP_CHECK(hipGetLastError());
}

void launchPtReduce(const DevSystem& S, const uint64_t* keys, const uint32_t* idx, const double* num, int64_t m,
                    double e0, DevMem& partial, DevMem& counts, DevMem& starts, Workspace& ws, double* out,
                    hipStream_t s) {
  if (m <= 0) {
    HIP_CHECK(hipMemsetAsync(out, 0, 3 * sizeof(double), s));
    return;
  }
  // run starts, compacted in order
  const int64_t nbS = (m + kStartChunk - 1) / kStartChunk;
  counts.reserve((size_t)(nbS + 1) * 8, s);
  int64_t* cnt = counts.as<int64_t>();
  HIP_CHECK(hipMemsetAsync(cnt + nbS, 0, 8, s));
  DISPATCH_W(S.W, kPtCountStarts<WW><<<(unsigned)nbS, kBlock, 0, s>>>(keys, m, cnt));
  exclusiveScanI64(cnt, nbS + 1, ws, s);
  int64_t nu = 0;
  HIP_CHECK(hipMemcpyAsync(&nu, cnt + nbS, 8, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  starts.reserve((size_t)std::max<int64_t>(nu, 1) * 4, s);
  DISPATCH_W(S.W, kPtWriteStarts<WW><<<(unsigned)nbS, kBlock, 0, s>>>(keys, m, cnt, starts.as<uint32_t>()));
  const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(kPtRedBlocks, (nu + kBlock - 1) / kBlock));
  partial.reserve((size_t)nb * 3 * sizeof(double));
  const size_t tb = ptTabBytes(S.norb);
  if (tb <= kPtTabMaxBytes)
    DISPATCH_W(S.W, kPtReduceStarts<WW, true><<<nb, kBlock, tb, s>>>(S, keys, idx, num, starts.as<uint32_t>(), nu, m,
                                                                     e0, partial.as<double>()));
  else
    DISPATCH_W(S.W, kPtReduceStarts<WW, false><<<nb, kBlock, 0, s>>>(S, keys, idx, num, starts.as<uint32_t>(), nu,
                                                                      m, e0, partial.as<double>()));
  kFinalSum<<<3, kBlock, 0, s>>>(partial.as<double>(), nb, 3, 3, out);
  HIP_CHECK(hipGetLastError());
}

void launchScreenSingles(const DevSystem& S, const uint64_t* dets, int64_t begin, int64_t end, const double* cmax,
                         double eps1, int ncore, int nact, const int32_t* table, uint64_t slots, uint64_t* out,
                         unsigned long long* counter, int64_t cap, hipStream_t s) {
  if (end <= begin) return;
  unsigned g = gridFor(end - begin, kWaves, 1 << 16);
  DISPATCH_W(S.W, kScreen<WW, 1><<<g, kBlock, 0, s>>>(S, dets, begin, end, cmax, eps1, ncore, nact, table,
                                                         slots - 1, out, counter, cap, 1.0, nullptr));
  HIP_CHECK(hipGetLastError());
}

void launchScreenRecords(const DevSystem& S, const uint64_t* dets, int64_t begin, int64_t end, const double* cmax,
                         double eps1, double loosen, int ncore, int nact, CandRecord* rec,
                         unsigned long long* counter, int64_t cap, hipStream_t s) {
  if (end <= begin) return;
  unsigned g = gridFor(end - begin, kWaves, 1 << 16);
  DISPATCH_W(S.W, kScreen<WW, 2><<<g, kBlock, 0, s>>>(S, dets, begin, end, cmax, eps1, ncore, nact, nullptr, 0,
                                                         nullptr, counter, cap, loosen, rec));
  HIP_CHECK(hipGetLastError());
}

void launchCandFeatures(const DevSystem& S, const CandRecord* rec, int64_t n, const uint64_t* dets,
                        const double* cmax, float* X, int din, hipStream_t s) {
  if (n <= 0) return;
  DISPATCH_W(S.W, kCandFeatures<WW><<<gridFor(n * din), kBlock, 0, s>>>(rec, n, dets, S.nso, cmax, X, din));
  HIP_CHECK(hipGetLastError());
}

void launchCandTargets(const CandRecord* rec, int64_t n, const double* cmax, float* t, hipStream_t s) {
  if (n <= 0) return;
  kCandTargets<<<gridFor(n), kBlock, 0, s>>>(rec, n, cmax, t);
  HIP_CHECK(hipGetLastError());
}

void launchCandExact(const DevSystem& S, const CandRecord* rec, int64_t n, const float* score, float tau,
                     const uint64_t* dets, const double* cmax, double eps1, const int32_t* table, uint64_t slots,
                     uint64_t* out, unsigned long long* counter, int64_t cap, unsigned long long* counts,
                     hipStream_t s) {
  if (n <= 0) return;
  DISPATCH_W(S.W, kCandExact<WW><<<gridFor(n), kBlock, 0, s>>>(rec, n, score, tau, dets, cmax, eps1, table,
                                                                  slots - 1, out, counter, cap, counts));
  HIP_CHECK(hipGetLastError());
}

void launchCandSample(const CandRecord* rec, const float* score, const int64_t* idx, int m, const double* cmax,
                      double eps1, float tau, int* flags, float* vals, hipStream_t s) {
  if (m <= 0) return;
  kCandSample<<<gridFor(m), kBlock, 0, s>>>(rec, score, idx, m, cmax, eps1, tau, flags, vals);
  HIP_CHECK(hipGetLastError());
}

void launchGatherRecords(const CandRecord* rec, const int64_t* idx, int64_t m, CandRecord* out, hipStream_t s) {
  if (m <= 0) return;
  kGatherRecords<<<gridFor(m), kBlock, 0, s>>>(rec, idx, m, out);
  HIP_CHECK(hipGetLastError());
}

void buildStringTables(const DevSystem& S, const uint64_t* dets, int64_t n, StringTables& T, Workspace& ws,
                       hipStream_t s) {
  T.n = n;
  const int norb = S.norb;
  T.alpha.reserve(n * 8);
  T.beta.reserve(n * 8);
  T.aIdx.reserve(n * 4);
  T.bIdx.reserve(n * 4);
  T.sortedA.reserve(n * 8);
  T.sortedB.reserve(n * 8);
  T.key.reserve(n * 8);
  T.val.reserve(n * 4);
  T.aMajorDet.reserve(n * 4);
  T.bMajorDet.reserve(n * 4);
  T.aMajorB.reserve(n * 8);
  T.bMajorA.reserve(n * 8);
  DISPATCH_W(S.W, kExtractStrings<WW><<<gridFor(n), kBlock, 0, s>>>(dets, n, T.alpha.as<uint64_t>(),
                                                                        T.beta.as<uint64_t>()));
  HIP_CHECK(hipMemcpyAsync(T.sortedA.p, T.alpha.p, n * 8, hipMemcpyDeviceToDevice, s));
  HIP_CHECK(hipMemcpyAsync(T.sortedB.p, T.beta.p, n * 8, hipMemcpyDeviceToDevice, s));
  radixSortKeys(T.sortedA.as<uint64_t>(), nullptr, n, 1, norb, false, ws, s);
  T.nA = uniqueKeys(T.sortedA.as<uint64_t>(), n, 1, ws, s);
  radixSortKeys(T.sortedB.as<uint64_t>(), nullptr, n, 1, norb, false, ws, s);
  T.nB = uniqueKeys(T.sortedB.as<uint64_t>(), n, 1, ws, s);
  kIndexOf<<<gridFor(n), kBlock, 0, s>>>(T.alpha.as<uint64_t>(), n, T.sortedA.as<uint64_t>(), T.nA,
                                          T.aIdx.as<uint32_t>());
  kIndexOf<<<gridFor(n), kBlock, 0, s>>>(T.beta.as<uint64_t>(), n, T.sortedB.as<uint64_t>(), T.nB,
                                          T.bIdx.as<uint32_t>());
  auto bitsFor = [](uint64_t v) {
    int b = 0;
    while (b < 64 && (v >> b) != 0) b++;
    return b;
  };
  T.aStart.reserve((T.nA + 1) * 8);
  T.bStart.reserve((T.nB + 1) * 8);
  // alpha-major groups, members in increasing determinant index:
  // key = aIdx * n + det
  kGroupKeys<<<gridFor(n), kBlock, 0, s>>>(T.aIdx.as<uint32_t>(), n, T.key.as<uint64_t>(), T.val.as<uint32_t>());
  radixSortKeys(T.key.as<uint64_t>(), T.val.as<uint32_t>(), n, 1, bitsFor((uint64_t)T.nA * n), false, ws, s);
  HIP_CHECK(hipMemcpyAsync(T.aMajorDet.p, T.val.p, n * 4, hipMemcpyDeviceToDevice, s));
  kGroupStarts<<<gridFor(n + 1), kBlock, 0, s>>>(T.key.as<uint64_t>(), n, (uint64_t)n, T.nA,
                                                  T.aStart.as<int64_t>(), T.val.as<uint32_t>(),
                                                  T.beta.as<uint64_t>(), T.aMajorB.as<uint64_t>());
  // beta-major groups: key = bIdx * n + det
  kGroupKeys<<<gridFor(n), kBlock, 0, s>>>(T.bIdx.as<uint32_t>(), n, T.key.as<uint64_t>(), T.val.as<uint32_t>());
  radixSortKeys(T.key.as<uint64_t>(), T.val.as<uint32_t>(), n, 1, bitsFor((uint64_t)T.nB * n), false, ws, s);
  HIP_CHECK(hipMemcpyAsync(T.bMajorDet.p, T.val.p, n * 4, hipMemcpyDeviceToDevice, s));
  kGroupStarts<<<gridFor(n + 1), kBlock, 0, s>>>(T.key.as<uint64_t>(), n, (uint64_t)n, T.nB,
                                                  T.bStart.as<int64_t>(), T.val.as<uint32_t>(),
                                                  T.alpha.as<uint64_t>(), T.bMajorA.as<uint64_t>());
  // alpha-major groups with members ordered by beta string (for lookups of
  // a row's beta singles in large groups): key = aIdx * nB + bIdx (the
  // group starts it writes equal the ones above)
  T.aSortB.reserve(n * 8);
  T.aSortDet.reserve(n * 4);
  kGroupKeys2<<<gridFor(n), kBlock, 0, s>>>(T.aIdx.as<uint32_t>(), T.bIdx.as<uint32_t>(), n, (uint64_t)T.nB,
                                             T.key.as<uint64_t>(), T.val.as<uint32_t>());
  radixSortKeys(T.key.as<uint64_t>(), T.val.as<uint32_t>(), n, 1, bitsFor((uint64_t)T.nA * T.nB), false, ws, s);
  HIP_CHECK(hipMemcpyAsync(T.aSortDet.p, T.val.p, n * 4, hipMemcpyDeviceToDevice, s));
  kGroupStarts<<<gridFor(n + 1), kBlock, 0, s>>>(T.key.as<uint64_t>(), n, (uint64_t)T.nB, T.nA,
                                                  T.aStart.as<int64_t>(), T.val.as<uint32_t>(),
                                                  T.beta.as<uint64_t>(), T.aSortB.as<uint64_t>());
  // beta single-excitation adjacency
  {
    T.bSinglesStart.reserve((T.nB + 1) * 8);
    int64_t* bc = T.bSinglesStart.as<int64_t>();
    HIP_CHECK(hipMemsetAsync(bc + T.nB, 0, 8, s));
    kStringSingles<false><<<gridFor(T.nB), kBlock, 0, s>>>(T.sortedB.as<uint64_t>(), T.nB, norb, nullptr, bc,
                                                            nullptr);
    exclusiveScanI64(bc, T.nB + 1, ws, s);
    int64_t tb = 0;
    HIP_CHECK(hipMemcpyAsync(&tb, bc + T.nB, 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    T.bSingles.reserve(std::max<int64_t>(tb, 1) * 4);
    kStringSingles<true><<<gridFor(T.nB), kBlock, 0, s>>>(T.sortedB.as<uint64_t>(), T.nB, norb, bc, nullptr,
                                                           T.bSingles.as<uint32_t>());
  }
  // alpha single-excitation adjacency
  T.aSinglesStart.reserve((T.nA + 1) * 8);
  int64_t* cnt = T.aSinglesStart.as<int64_t>();
  HIP_CHECK(hipMemsetAsync(cnt + T.nA, 0, 8, s));
  kStringSingles<false><<<gridFor(T.nA), kBlock, 0, s>>>(T.sortedA.as<uint64_t>(), T.nA, norb, nullptr, cnt,
                                                          nullptr);
  exclusiveScanI64(cnt, T.nA + 1, ws, s);
  int64_t tot = 0;
  HIP_CHECK(hipMemcpyAsync(&tot, cnt + T.nA, 8, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  T.aSingles.reserve(std::max<int64_t>(tot, 1) * 4);
  kStringSingles<true><<<gridFor(T.nA), kBlock, 0, s>>>(T.sortedA.as<uint64_t>(), T.nA, norb, cnt, nullptr,
                                                         T.aSingles.as<uint32_t>());
  HIP_CHECK(hipGetLastError());
}

static RowTables rowTables(const StringTables& T) {
  RowTables r;
  r.alpha = T.alpha.as<uint64_t>();
  r.beta = T.beta.as<uint64_t>();
  r.aIdx = T.aIdx.as<uint32_t>();
  r.bIdx = T.bIdx.as<uint32_t>();
  r.aStart = T.aStart.as<int64_t>();
  r.bStart = T.bStart.as<int64_t>();
  r.aMajorDet = T.aMajorDet.as<uint32_t>();
  r.bMajorDet = T.bMajorDet.as<uint32_t>();
  r.aMajorB = T.aMajorB.as<uint64_t>();
  r.bMajorA = T.bMajorA.as<uint64_t>();
  r.aSinglesStart = T.aSinglesStart.as<int64_t>();
  r.aSingles = T.aSingles.as<uint32_t>();
  r.bSinglesStart = T.bSinglesStart.as<int64_t>();
  r.bSingles = T.bSingles.as<uint32_t>();
  r.sortedB = T.sortedB.as<uint64_t>();
  r.aSortB = T.aSortB.as<uint64_t>();
  r.aSortDet = T.aSortDet.as<uint32_t>();
  r.n = T.n;
  return r;
}

void launchCsrCount(const DevSystem& S, const uint64_t* dets, const StringTables& T, int64_t nOld, int64_t* rowCount,
                    hipStream_t s, bool lower) {
  if (T.n <= 0) return;
  unsigned g = gridFor(T.n, kWaves, 1 << 16);
  RowTables r = rowTables(T);
  if (lower)
    DISPATCH_W(S.W, kCsrRows<WW, false, false, true><<<g, kBlock, 0, s>>>(S, dets, r, nOld, rowCount, nullptr,
                                                                          nullptr, nullptr, nullptr, nullptr));
  else
    DISPATCH_W(S.W, kCsrRows<WW, false, false, false><<<g, kBlock, 0, s>>>(S, dets, r, nOld, rowCount, nullptr,
                                                                           nullptr, nullptr, nullptr, nullptr));
  HIP_CHECK(hipGetLastError());
}

void launchCsrFill(const DevSystem& S, const uint64_t* dets, const StringTables& T, int64_t nOld,
                   const int64_t* fillStart, int32_t* cols, double* vals, hipStream_t s, bool lower) {
  if (T.n <= 0) return;
  unsigned g = gridFor(T.n, kWaves, 1 << 16);
  RowTables r = rowTables(T);
  if (lower)
    DISPATCH_W(S.W, kCsrRows<WW, true, false, true><<<g, kBlock, 0, s>>>(S, dets, r, nOld, nullptr, fillStart, cols,
                                                                         vals, nullptr, nullptr));
  else
    DISPATCH_W(S.W, kCsrRows<WW, true, false, false><<<g, kBlock, 0, s>>>(S, dets, r, nOld, nullptr, fillStart, cols,
                                                                          vals, nullptr, nullptr));
  HIP_CHECK(hipGetLastError());
}

void launchCsrFillSplit(const DevSystem& S, const uint64_t* dets, const StringTables& T, int64_t nOld,
                        const int64_t* fillStart, int32_t* cols, float* hi, float* lo, hipStream_t s, bool lower) {
  if (T.n <= 0) return;
  unsigned g = gridFor(T.n, kWaves, 1 << 16);
  RowTables r = rowTables(T);
  if (lower)
    DISPATCH_W(S.W, kCsrRows<WW, true, true, true><<<g, kBlock, 0, s>>>(S, dets, r, nOld, nullptr, fillStart, cols,
                                                                        nullptr, hi, lo));
  else
    DISPATCH_W(S.W, kCsrRows<WW, true, true, false><<<g, kBlock, 0, s>>>(S, dets, r, nOld, nullptr, fillStart, cols,
                                                                         nullptr, hi, lo));
  HIP_CHECK(hipGetLastError());
}

void launchLowerRowPtr(const int64_t* cnt, int64_t nOld, int64_t n, int64_t* rowPtr, hipStream_t s) {
  if (n + 1 - nOld <= 0) return;
  kLowerRowPtr<<<gridFor(n + 1 - nOld), kBlock, 0, s>>>(cnt, nOld, n, rowPtr);
  HIP_CHECK(hipGetLastError());
}

void launchSigmaLowerF64(const int64_t* rowPtr, const int32_t* cols, const double* vals, const double* diag, int64_t n,
                         const double* x, double* y, int k, int64_t ld, hipStream_t s) {
  if (n <= 0 || k <= 0) return;
  kZeroVectors<<<gridFor(n * k), kBlock, 0, s>>>(y, k, ld, n);
  kSigmaLower<double, false><<<gridFor(n, kWaves), kBlock, 0, s>>>(rowPtr, cols, vals, nullptr, diag, n, x, y, k, ld);
  HIP_CHECK(hipGetLastError());
}

void launchSigmaLowerF32(const int64_t* rowPtr, const int32_t* cols, const float* vals, const float* lo,
                         const double* diag, int64_t n, const double* x, double* y, int k, int64_t ld, hipStream_t s) {
  if (n <= 0 || k <= 0) return;
  kZeroVectors<<<gridFor(n * k), kBlock, 0, s>>>(y, k, ld, n);
  if (lo)
    kSigmaLower<float, true><<<gridFor(n, kWaves), kBlock, 0, s>>>(rowPtr, cols, vals, lo, diag, n, x, y, k, ld);
  else
    kSigmaLower<float, false><<<gridFor(n, kWaves), kBlock, 0, s>>>(rowPtr, cols, vals, nullptr, diag, n, x, y, k,
                                                                    ld);
  HIP_CHECK(hipGetLastError());
}

void launchSigmaDirect(const DevSystem& S, const uint64_t* dets, const StringTables& T, const double* diag,
                       const double* x, double* y, int k, int64_t ld, hipStream_t s) {
  if (T.n <= 0 || k <= 0) return;
  unsigned g = gridFor(T.n, kWaves, 1 << 16);
  RowTables r = rowTables(T);
  for (int c0 = 0; c0 < k; c0 += 8) {
    const int kk = std::min(8, k - c0);
    DISPATCH_W(S.W, kSigmaDirect<WW><<<g, kBlock, 0, s>>>(S, dets, r, diag, x + (int64_t)c0 * ld,
                                                             y + (int64_t)c0 * ld, kk, ld));
  }
  HIP_CHECK(hipGetLastError());
}

void launchNewRowLengths(const int64_t* oldRowPtr, int64_t nOld, int64_t n, int64_t* cnt, hipStream_t s) {
  if (nOld <= 0) return;
  kNewRowLengths<<<gridFor(n), kBlock, 0, s>>>(oldRowPtr, nOld, n, cnt);
  HIP_CHECK(hipGetLastError());
}

void launchFillStart(const int64_t* oldRowPtr, int64_t nOld, int64_t n, const int64_t* rowPtr, int64_t* fs,
                     hipStream_t s) {
  kFillStart<<<gridFor(n), kBlock, 0, s>>>(oldRowPtr, nOld, n, rowPtr, fs);
  HIP_CHECK(hipGetLastError());
}

void launchCopyOldRowsF64(const int64_t* oldRowPtr, const int32_t* oldCols, const double* oldVals, int64_t nOld,
                          const int64_t* rowPtr, int32_t* cols, double* vals, hipStream_t s) {
  if (nOld <= 0) return;
  kCopyOldRows<double><<<gridFor(nOld, kWaves, 1 << 16), kBlock, 0, s>>>(oldRowPtr, oldCols, oldVals, nullptr, nOld,
                                                                          rowPtr, cols, vals, nullptr);
  HIP_CHECK(hipGetLastError());
}

void launchCopyOldRowsF32(const int64_t* oldRowPtr, const int32_t* oldCols, const float* oldHi, const float* oldLo,
                          int64_t nOld, const int64_t* rowPtr, int32_t* cols, float* hi, float* lo, hipStream_t s) {
  if (nOld <= 0) return;
  kCopyOldRows<float><<<gridFor(nOld, kWaves, 1 << 16), kBlock, 0, s>>>(oldRowPtr, oldCols, oldHi, oldLo, nOld, rowPtr,
                                                                         cols, hi, lo);
  HIP_CHECK(hipGetLastError());
}

void launchCastValues(const double* v, float* out, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  kCast<<<gridFor(n), kBlock, 0, s>>>(v, out, n);
  HIP_CHECK(hipGetLastError());
}

void launchSortRows(const int64_t* rowPtr, int64_t n, int32_t* cols, double* vals, hipStream_t s) {
  if (n <= 0) return;
  kSortRows<<<gridFor(n, 64), 64, 0, s>>>(rowPtr, n, cols, vals);
  HIP_CHECK(hipGetLastError());
}

template <typename VT, bool LO>
static void launchSigmaT(const int64_t* rowPtr, const int32_t* cols, const VT* vals, const float* lo,
                         const double* diag, int64_t n, const double* x, double* y, int k, int64_t ld, hipStream_t s) {
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>((n + kWaves - 1) / kWaves, 1 << 15);
  kSigma<VT, LO><<<(unsigned)blocks, kBlock, 0, s>>>(rowPtr, cols, vals, lo, diag, n, x, y, k, ld);
  HIP_CHECK(hipGetLastError());
}

void launchSigmaF64(const int64_t* rowPtr, const int32_t* cols, const double* vals, const double* diag, int64_t n,
                    const double* x, double* y, int k, int64_t ld, hipStream_t s) {
  launchSigmaT<double, false>(rowPtr, cols, vals, nullptr, diag, n, x, y, k, ld, s);
}

void launchSigmaF32(const int64_t* rowPtr, const int32_t* cols, const float* vals, const float* lo,
                    const double* diag, int64_t n, const double* x, double* y, int k, int64_t ld, hipStream_t s) {
  if (lo)
    launchSigmaT<float, true>(rowPtr, cols, vals, lo, diag, n, x, y, k, ld, s);
  else
    launchSigmaT<float, false>(rowPtr, cols, vals, nullptr, diag, n, x, y, k, ld, s);
}

__global__ void kSplitHiLo(const double* v, float* hi, float* lo, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double d = v[i];
    const float h = (float)d;
    hi[i] = h;
    lo[i] = (float)(d - (double)h);
  }
}

void launchSplitHiLo(const double* v, float* hi, float* lo, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  kSplitHiLo<<<gridFor(n), kBlock, 0, s>>>(v, hi, lo, n);
  HIP_CHECK(hipGetLastError());
}

void launchInversePerm(const uint32_t* perm, int64_t n, uint32_t* pos, hipStream_t s) {
  if (n <= 0) return;
  kInversePerm<<<gridFor(n), kBlock, 0, s>>>(perm, n, pos);
  HIP_CHECK(hipGetLastError());
}

void launchPermRowLengths(const int64_t* rowPtr, const uint32_t* perm, int64_t n, int64_t* len, hipStream_t s) {
  if (n <= 0) return;
  kPermRowLengths<<<gridFor(n), kBlock, 0, s>>>(rowPtr, perm, n, len);
  HIP_CHECK(hipGetLastError());
}

void launchPermRowsF64(const int64_t* rowPtr, const int32_t* cols, const double* vals, const uint32_t* perm,
                       const uint32_t* pos, const uint32_t* bIdx, const uint32_t* bPos, int64_t n,
                       const int64_t* rpP, int32_t* colsP, double* valsP, hipStream_t s) {
  if (n <= 0) return;
  if (bIdx && 2 * n > 0x7fffffffLL) throw HipError("too many rows for the split relabelled columns");
  kPermRows<double><<<gridFor(n, kWaves, 1 << 16), kBlock, 0, s>>>(rowPtr, cols, vals, perm, pos, bIdx, bPos, n, rpP,
                                                                     colsP, valsP);
  HIP_CHECK(hipGetLastError());
}

void launchPermRowsF32(const int64_t* rowPtr, const int32_t* cols, const float* vals, const uint32_t* perm,
                       const uint32_t* pos, const uint32_t* bIdx, const uint32_t* bPos, int64_t n,
                       const int64_t* rpP, int32_t* colsP, float* valsP, hipStream_t s) {
  if (n <= 0) return;
  if (bIdx && 2 * n > 0x7fffffffLL) throw HipError("too many rows for the split relabelled columns");
  kPermRows<float><<<gridFor(n, kWaves, 1 << 16), kBlock, 0, s>>>(rowPtr, cols, vals, perm, pos, bIdx, bPos, n, rpP,
                                                                    colsP, valsP);
  HIP_CHECK(hipGetLastError());
}

static int computeUnits() {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (cus <= 0) cus = 256;
  }
  return cus;
}

void* XScratch::get(size_t bytes) const {
  if (mem) {
    mem->reserve(bytes);
    return mem->p;
  }
  if (bytes > fixedBytes)
    throw HipError("sigma needs " + std::to_string(bytes) + " bytes of x scratch, more than the fixed " +
                   std::to_string(fixedBytes));
  return fixed;
}

size_t sigmaPermScratchBytes(int64_t n, bool split, int k, size_t xBytes) {
  const int K = k <= 1 ? 1 : k == 2 ? 2 : k <= 4 ? 4 : 8;
  return (size_t)(split ? 2 * n : n) * K * xBytes;
}

// Fixed scratch of the allocation-free sigma: 256 bytes of work counters,
// then up to 64 MiB of permuted x (one FP64 vector of up to 4,194,304
// determinants with split columns, or n * vectors per pass up to that).
// Each GPU holds its own copy for the life of the process.
constexpr size_t kFixedScratchBytes = ((size_t)64 << 20) + 256;
static __device__ __align__(256) unsigned char scratch[kFixedScratchBytes];

void* fixedSigmaScratch(size_t* bytes) {
  static std::mutex mu;
  static void* addr[64] = {};
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) throw HipError("device index out of range for the fixed sigma scratch");
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw HipError("hipGetSymbolAddress returned no address for the fixed sigma scratch");
    addr[dev] = p;
  }
  if (bytes) *bytes = kFixedScratchBytes;
  return addr[dev];
}

template <typename VT, typename XT, int CHUNK, bool PERSIST>
static void sigmaPermChunk(const int64_t* rp, const int32_t* cols, const VT* vals, const double* diag,
                           const uint32_t* perm, const uint32_t* bPerm, int64_t n, const double* x, double* y, int k,
                           int64_t ld, const XScratch& xP, unsigned* counter, hipStream_t s) {
  const int64_t nx = bPerm ? 2 * n : n;
  const int64_t chunks = (n + CHUNK - 1) / CHUNK;
  if (chunks > 0x7fffffffLL) throw HipError("too many rows for the relabelled sigma");
  unsigned grid = (unsigned)chunks;
  if (PERSIST) grid = (unsigned)std::min<int64_t>(chunks, (int64_t)computeUnits() * 8);
  for (int c0 = 0; c0 < k;) {
    const int kk = k - c0;
    const double* xc = x + (int64_t)c0 * ld;
    double* yc = y + (int64_t)c0 * ld;
    auto spmm = [&](auto KT) {
      constexpr int K = decltype(KT)::value;
      const int take = kk < K ? kk : K;
      XT* xs = (XT*)xP.get((size_t)nx * K * sizeof(XT));
      kGatherPerm<XT, K><<<gridFor(nx), kBlock, 0, s>>>(xc, take, ld, perm, bPerm, n, xs, counter);
      kSpmmPerm<VT, XT, K, CHUNK, PERSIST><<<grid, kBlock, 0, s>>>(rp, cols, vals, diag, perm, n, xs, xc, yc, take,
                                                                     ld, counter);
      return take;
    };
    if (kk == 1) {
      XT* xs = (XT*)xP.get((size_t)nx * sizeof(XT));
      kGatherPerm<XT, 1><<<gridFor(nx), kBlock, 0, s>>>(xc, 1, ld, perm, bPerm, n, xs, counter);
      kSigmaPerm<VT, XT, CHUNK, PERSIST><<<grid, kBlock, 0, s>>>(rp, cols, vals, diag, perm, n, xs, nx, xc, yc, 1,
                                                                   ld, counter);
      c0 += 1;
    } else if (kk == 2) {
      c0 += spmm(std::integral_constant<int, 2>());
    } else if (kk <= 4) {
      c0 += spmm(std::integral_constant<int, 4>());
    } else {
      c0 += spmm(std::integral_constant<int, 8>());
    }
    HIP_CHECK(hipGetLastError());
  }
}

template <typename VT, typename XT>
static void launchSigmaPermT(const int64_t* rp, const int32_t* cols, const VT* vals, const double* diag,
                             const uint32_t* perm, const uint32_t* bPerm, int64_t n, const double* x, double* y, int k, int64_t ld,
                             const XScratch& xP, const SigmaConfig& cfg, unsigned* counter, hipStream_t s) {
  if (n <= 0 || k <= 0) return;
  if (cfg.persistent && !counter) throw HipError("persistent sigma needs a work counter");
  switch (cfg.chunk * 2 + (cfg.persistent ? 1 : 0)) {
    case 8: sigmaPermChunk<VT, XT, 4, false>(rp, cols, vals, diag, perm, bPerm, n, x, y, k, ld, xP, counter, s); break;
    case 9: sigmaPermChunk<VT, XT, 4, true>(rp, cols, vals, diag, perm, bPerm, n, x, y, k, ld, xP, counter, s); break;
    case 16: sigmaPermChunk<VT, XT, 8, false>(rp, cols, vals, diag, perm, bPerm, n, x, y, k, ld, xP, counter, s); break;
    case 17: sigmaPermChunk<VT, XT, 8, true>(rp, cols, vals, diag, perm, bPerm, n, x, y, k, ld, xP, counter, s); break;
    case 32: sigmaPermChunk<VT, XT, 16, false>(rp, cols, vals, diag, perm, bPerm, n, x, y, k, ld, xP, counter, s); break;
    case 33: sigmaPermChunk<VT, XT, 16, true>(rp, cols, vals, diag, perm, bPerm, n, x, y, k, ld, xP, counter, s); break;
    case 64: sigmaPermChunk<VT, XT, 32, false>(rp, cols, vals, diag, perm, bPerm, n, x, y, k, ld, xP, counter, s); break;
    case 65: sigmaPermChunk<VT, XT, 32, true>(rp, cols, vals, diag, perm, bPerm, n, x, y, k, ld, xP, counter, s); break;
    default: throw HipError("unsupported sigma configuration");
  }
}

void launchSigmaPermF64(const int64_t* rp, const int32_t* cols, const double* vals, const double* diag,
                        const uint32_t* perm, const uint32_t* bPerm, int64_t n, const double* x, double* y, int k,
                        int64_t ld, const XScratch& xP, const SigmaConfig& cfg, unsigned* counter, hipStream_t s) {
  launchSigmaPermT<double, double>(rp, cols, vals, diag, perm, bPerm, n, x, y, k, ld, xP, cfg, counter, s);
}

void launchSigmaPermF32(const int64_t* rp, const int32_t* cols, const float* vals, const double* diag,
                        const uint32_t* perm, const uint32_t* bPerm, int64_t n, const double* x, double* y, int k,
                        int64_t ld, const XScratch& xP, const SigmaConfig& cfg, unsigned* counter, hipStream_t s) {
  launchSigmaPermT<float, float>(rp, cols, vals, diag, perm, bPerm, n, x, y, k, ld, xP, cfg, counter, s);
}

void TiledHam::release() {
  for (DevMem* m : {&segBase, &segTile, &segCnt, &segStart, &segInst, &segTotal, &fill, &blockE, &blockI, &idx, &vals,
                    &lo, &inst})
    m->release();
  n = -1;
  hasLo = false;
}

size_t tiledHamBytes(int64_t n, int64_t nnz, bool f32) {
  const int64_t blocks = (n + kTileRows - 1) / kTileRows;
  // entries, one instruction per 32 entries at most on average plus one per
  // segment, segment tables (48 bytes) with at most one segment per 64
  // entries or per block
  const int64_t segs = nnz / 64 + blocks;
  return (size_t)nnz * (4 + (f32 ? 4 : 8)) + (size_t)(nnz / 32 + segs) * 8 + (size_t)segs * 48 +
         (size_t)(blocks + 1) * 24;
}

bool buildTiledHam(TiledHam& T, const int64_t* rowPtr, const int32_t* cols, const double* vals64,
                   const float* vals32, const uint32_t* perm, const uint32_t* pos, const uint32_t* bIdx,
                   const uint32_t* bPos, int64_t n, int64_t nnz, Workspace& ws, hipStream_t s, bool splitCols,
                   int stageMin, const float* valsLo) {
  T.n = -1;
  T.hasLo = false;
  if (n <= 0) return false;
  const bool split = pos ? bIdx != nullptr : splitCols;
  const int64_t nx = split ? 2 * n : n;
  const int64_t tiles = (nx + kTileCols - 1) / kTileCols;
  const int nwords = (int)((tiles + 31) / 32);
  if (nwords > kTileMaxWords) return false;
  const int64_t blocks = (n + kTileRows - 1) / kTileRows;
  if (blocks > 0x7fffffffLL) return false;
  const unsigned grid = (unsigned)blocks;
  T.segBase.reserve((size_t)(blocks + 1) * 8);
  int64_t* segBase = T.segBase.as<int64_t>();
  HIP_CHECK(hipMemsetAsync(segBase + blocks, 0, 8, s));
  kTileMark<<<grid, kBlock, 0, s>>>(rowPtr, cols, perm, pos, bIdx, bPos, n, nwords, segBase);
  HIP_CHECK(hipGetLastError());
  exclusiveScanI64(segBase, blocks + 1, ws, s);
  int64_t S = 0;
  HIP_CHECK(hipMemcpyAsync(&S, segBase + blocks, 8, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  const size_t sb = (size_t)std::max<int64_t>(S, 1);
  T.segTile.reserve(sb * 4);
  T.segCnt.reserve(sb * 16);
  T.segStart.reserve(sb * 4);
  T.segInst.reserve(sb * 4);
  T.segTotal.reserve(sb * 4);
  T.fill.reserve(sb * 16);
  T.blockE.reserve((size_t)(blocks + 1) * 8);
  T.blockI.reserve((size_t)(blocks + 1) * 8);
  HIP_CHECK(hipMemsetAsync(T.segCnt.p, 0, sb * 16, s));
  int64_t* blockE = T.blockE.as<int64_t>();
  int64_t* blockI = T.blockI.as<int64_t>();
  HIP_CHECK(hipMemsetAsync(blockE + blocks, 0, 8, s));
  HIP_CHECK(hipMemsetAsync(blockI + blocks, 0, 8, s));
  kTileCount<<<grid, kBlock, 0, s>>>(rowPtr, cols, perm, pos, bIdx, bPos, n, nwords, segBase,
                                     T.segTile.as<uint32_t>(), T.segCnt.as<uint32_t>(), T.segStart.as<uint32_t>(),
                                     T.segInst.as<uint32_t>(), T.segTotal.as<uint32_t>(), T.fill.as<uint32_t>(),
                                     blockE, blockI);
  HIP_CHECK(hipGetLastError());
  exclusiveScanI64(blockE, blocks + 1, ws, s);
  exclusiveScanI64(blockI, blocks + 1, ws, s);
  int64_t E = 0, I = 0;
  HIP_CHECK(hipMemcpyAsync(&E, blockE + blocks, 8, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(&I, blockI + blocks, 8, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  if (E != nnz) throw HipError("tiled Hamiltonian: entry count " + std::to_string(E) + " != " + std::to_string(nnz));
  const bool f32 = vals64 == nullptr;
  T.idx.reserve((size_t)std::max<int64_t>(E, 1) * 4);
  T.vals.reserve((size_t)std::max<int64_t>(E, 1) * (f32 ? 4 : 8));
  const bool withLo = f32 && valsLo;
  if (withLo) T.lo.reserve((size_t)std::max<int64_t>(E, 1) * 4);
  // kSigmaTiled reads up to kTileInstPad descriptors past a block's last
  // one: zero padding after the last block (count 0, tile 0)
  T.inst.reserve((size_t)(I + kTileInstPad) * 8);
  HIP_CHECK(hipMemsetAsync(T.inst.as<uint2>() + I, 0, (size_t)kTileInstPad * 8, s));
  if (f32)
    kTileFill<float><<<grid, kBlock, 0, s>>>(rowPtr, cols, vals32, perm, pos, bIdx, bPos, n, nwords, segBase,
                                             T.segTile.as<uint32_t>(), T.segStart.as<uint32_t>(),
                                             T.segInst.as<uint32_t>(),
                                             T.segTotal.as<uint32_t>(), T.fill.as<uint32_t>(), blockE, blockI,
                                             T.idx.as<uint32_t>(), T.vals.as<float>(), T.inst.as<uint2>(),
                                             stageMin, withLo ? valsLo : nullptr,
                                             withLo ? T.lo.as<float>() : nullptr);
  else
    kTileFill<double><<<grid, kBlock, 0, s>>>(rowPtr, cols, vals64, perm, pos, bIdx, bPos, n, nwords, segBase,
                                              T.segTile.as<uint32_t>(), T.segStart.as<uint32_t>(),
                                              T.segInst.as<uint32_t>(),
                                              T.segTotal.as<uint32_t>(), T.fill.as<uint32_t>(), blockE, blockI,
                                              T.idx.as<uint32_t>(), T.vals.as<double>(), T.inst.as<uint2>(),
                                              stageMin, nullptr, nullptr);
  HIP_CHECK(hipGetLastError());
  kTileRunBits<<<grid, kBlock, 0, s>>>(T.inst.as<uint2>(), blockI, blockE, T.idx.as<uint32_t>());
  HIP_CHECK(hipGetLastError());
  T.n = n;
  T.blocks = blocks;
  T.segments = S;
  T.nnz = E;
  T.instructions = I;
  T.f32 = f32;
  T.split = split;
  T.hasLo = withLo;
  return true;
}

template <typename VT, typename XT, bool PERSIST, bool STAGE>
static void sigmaTiledS(int step, unsigned grid, const TiledHam& T, const double* diag, const uint32_t* perm,
                        int64_t n, const XT* xs, int64_t nx, const double* x, double* y, unsigned* counter,
                        hipStream_t s) {
  if (step >= 16)
    kSigmaTiled<VT, XT, PERSIST, 16, STAGE><<<grid, kBlock, 0, s>>>(
        T.inst.as<uint2>(), T.blockI.as<int64_t>(), T.blockE.as<int64_t>(), T.idx.as<uint32_t>(), T.vals.as<VT>(),
        diag, perm, n, T.blocks, xs, nx, x, y, counter);
  else
    kSigmaTiled<VT, XT, PERSIST, 8, STAGE><<<grid, kBlock, 0, s>>>(
        T.inst.as<uint2>(), T.blockI.as<int64_t>(), T.blockE.as<int64_t>(), T.idx.as<uint32_t>(), T.vals.as<VT>(),
        diag, perm, n, T.blocks, xs, nx, x, y, counter);
}

template <typename VT, typename XT, bool PERSIST>
static void sigmaTiledU(int step, bool stage, unsigned grid, const TiledHam& T, const double* diag,
                        const uint32_t* perm, int64_t n, const XT* xs, int64_t nx, const double* x, double* y,
                        unsigned* counter, hipStream_t s) {
  if (stage)
    sigmaTiledS<VT, XT, PERSIST, true>(step, grid, T, diag, perm, n, xs, nx, x, y, counter, s);
  else
    sigmaTiledS<VT, XT, PERSIST, false>(step, grid, T, diag, perm, n, xs, nx, x, y, counter, s);
}

void launchSigmaTiled(const TiledHam& T, const double* diag, const uint32_t* perm, const uint32_t* bPerm,
                      const double* x, double* y, const XScratch& xP, bool persistent, unsigned* counter,
                      hipStream_t s, int step, bool stage) {
  const int64_t n = T.n;
  if (n <= 0) return;
  if (T.split != (bPerm != nullptr)) throw HipError("tiled sigma: split columns need the beta-major order");
  const int64_t nx = bPerm ? 2 * n : n;
  const unsigned grid =
      (unsigned)(persistent ? std::min<int64_t>(T.blocks, (int64_t)computeUnits() * 8) : T.blocks);
  if (persistent && !counter) throw HipError("persistent sigma needs a work counter");
  if (T.f32) {
    float* xs = (float*)xP.get((size_t)nx * 4);
    kGatherPerm<float, 1><<<gridFor(nx), kBlock, 0, s>>>(x, 1, n, perm, bPerm, n, xs, counter);
    if (persistent)
      sigmaTiledU<float, float, true>(step, stage, grid, T, diag, perm, n, xs, nx, x, y, counter, s);
    else
      sigmaTiledU<float, float, false>(step, stage, grid, T, diag, perm, n, xs, nx, x, y, counter, s);
  } else {
    double* xs = (double*)xP.get((size_t)nx * 8);
    kGatherPerm<double, 1><<<gridFor(nx), kBlock, 0, s>>>(x, 1, n, perm, bPerm, n, xs, counter);
    if (persistent)
      sigmaTiledU<double, double, true>(step, stage, grid, T, diag, perm, n, xs, nx, x, y, counter, s);
    else
      sigmaTiledU<double, double, false>(step, stage, grid, T, diag, perm, n, xs, nx, x, y, counter, s);
  }
  HIP_CHECK(hipGetLastError());
}

void launchSigmaTiledSym(const TiledHam& T, const double* diag, const uint32_t* perm, const uint32_t* bPerm,
                         const uint32_t* pos, const uint32_t* bPos, const double* x, double* y, const XScratch& xP,
                         unsigned long long* yT, const unsigned long long* bound, hipStream_t s, bool exact) {
  const int64_t n = T.n;
  if (n <= 0) return;
  if (T.split != (bPerm != nullptr) || T.split != (bPos != nullptr))
    throw HipError("tiled sigma: split columns need the beta-major order");
  if (exact && T.f32 && !T.hasLo) throw HipError("tiled sigma: the exact FP32 form needs the lo values");
  const int64_t nx = bPerm ? 2 * n : n;
  const unsigned grid = (unsigned)T.blocks;
  kFillZeroU64<<<gridFor(nx + 1), kBlock, 0, s>>>(yT, nx + 1);
  if (T.f32 && !exact) {
    float* xs = (float*)xP.get((size_t)nx * 4);
    kGatherPermMax<float><<<gridFor(nx, kBlock, 1024), kBlock, 0, s>>>(x, perm, bPerm, n, xs, yT + nx);
    kSigmaTiledSym<float, float, 8, false><<<grid, kBlock, 0, s>>>(
        T.inst.as<uint2>(), T.blockI.as<int64_t>(), T.blockE.as<int64_t>(), T.idx.as<uint32_t>(), T.vals.as<float>(),
        nullptr, n, nx, xs, yT, bound);
  } else {
    double* xs = (double*)xP.get((size_t)nx * 8);
    kGatherPermMax<double><<<gridFor(nx, kBlock, 1024), kBlock, 0, s>>>(x, perm, bPerm, n, xs, yT + nx);
    if (T.f32)  // the exact FP32 form: hi + lo values, x in FP64
      kSigmaTiledSym<float, double, 8, true><<<grid, kBlock, 0, s>>>(
          T.inst.as<uint2>(), T.blockI.as<int64_t>(), T.blockE.as<int64_t>(), T.idx.as<uint32_t>(),
          T.vals.as<float>(), T.lo.as<float>(), n, nx, xs, yT, bound);
    else
      kSigmaTiledSym<double, double, 8, false><<<grid, kBlock, 0, s>>>(
          T.inst.as<uint2>(), T.blockI.as<int64_t>(), T.blockE.as<int64_t>(), T.idx.as<uint32_t>(),
          T.vals.as<double>(), nullptr, n, nx, xs, yT, bound);
  }
  kSymFinish<<<gridFor(n), kBlock, 0, s>>>(diag, x, y, yT, bound, pos, bPos, n, nx);
  HIP_CHECK(hipGetLastError());
}

void launchSymRowBound(const int64_t* rowPtr, const int32_t* cols, const double* vals64, const float* vals32,
                       int64_t nOld, int64_t n, unsigned long long* acc, unsigned long long* bound, hipStream_t s) {
  if (n > nOld) kFillZeroU64<<<gridFor(n - nOld), kBlock, 0, s>>>(acc + nOld, n - nOld);
  kFillZeroU64<<<1, 64, 0, s>>>(bound, 1);
  if (n > nOld) {
    if (vals32)
      kRowAbsSum<float><<<gridFor(n - nOld, kWaves), kBlock, 0, s>>>(rowPtr, cols, vals32, nOld, n, acc);
    else
      kRowAbsSum<double><<<gridFor(n - nOld, kWaves), kBlock, 0, s>>>(rowPtr, cols, vals64, nOld, n, acc);
  }
  if (n > 0) kMaxU64<<<gridFor(n, kBlock, 1024), kBlock, 0, s>>>(acc, n, bound);
  kU64ScaledToDouble<<<1, 64, 0, s>>>(bound);
  HIP_CHECK(hipGetLastError());
}

void launchGram(const double* A, int ma, const double* B, int mb, int64_t n, int64_t ld, double* G, DevMem& partial,
                hipStream_t s) {
  const int nb = (int)std::min<int64_t>(kRedBlocks, (n + kBlock - 1) / kBlock);
  const int nbb = nb < 1 ? 1 : nb;
  if (ma <= 1 && mb <= 1) {
    partial.reserve((size_t)nbb * 1 * sizeof(double));
    kGramPartial<1, 1><<<nbb, kBlock, 0, s>>>(A, B, ma, mb, n, ld, partial.as<double>());
    kGramFinal<<<ma * mb, kBlock, 0, s>>>(partial.as<double>(), nbb, 1, 1, ma, mb, G);
  } else if (ma <= 1 && mb <= 8) {
    partial.reserve((size_t)nbb * 8 * sizeof(double));
    kGramPartial<1, 8><<<nbb, kBlock, 0, s>>>(A, B, ma, mb, n, ld, partial.as<double>());
    kGramFinal<<<ma * mb, kBlock, 0, s>>>(partial.as<double>(), nbb, 1, 8, ma, mb, G);
  } else if (ma <= 8 && mb <= 8) {
    partial.reserve((size_t)nbb * 64 * sizeof(double));
    kGramPartial<8, 8><<<nbb, kBlock, 0, s>>>(A, B, ma, mb, n, ld, partial.as<double>());
    kGramFinal<<<ma * mb, kBlock, 0, s>>>(partial.as<double>(), nbb, 8, 8, ma, mb, G);
  } else if (ma <= 12 && mb <= 12) {
    partial.reserve((size_t)nbb * 144 * sizeof(double));
    kGramPartial<12, 12><<<nbb, kBlock, 0, s>>>(A, B, ma, mb, n, ld, partial.as<double>());
    kGramFinal<<<ma * mb, kBlock, 0, s>>>(partial.as<double>(), nbb, 12, 12, ma, mb, G);
  } else {
    throw HipError("launchGram: subspace too large");
  }
  HIP_CHECK(hipGetLastError());
}

void launchGramTiled(const double* A, int ma, const double* B, int mb, int64_t n, int64_t ld, double* G,
                     DevMem& partial, hipStream_t s) {
  if (ma > 64 || mb > 64) throw HipError("launchGramTiled: block too large");
  if (ma <= 0 || mb <= 0) return;
  const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(kRedBlocks / 2, (n + kBlock - 1) / kBlock));
  const int tiles = ((ma + 7) / 8) * ((mb + 7) / 8);
  partial.reserve((size_t)tiles * nb * 64 * sizeof(double));
  kGramTile<<<dim3(nb, tiles), kBlock, 0, s>>>(A, ma, B, mb, n, ld, partial.as<double>());
  kGramTileFinal<<<ma * mb, kBlock, 0, s>>>(partial.as<double>(), nb, ma, mb, G);
  HIP_CHECK(hipGetLastError());
}

void launchGemmTS(const double* A, int ma, const double* W, int mb, int64_t n, int64_t lda, double* out,
                  int64_t ldo, double alpha, double beta, hipStream_t s) {
  if (ma > 64 || mb > 64) throw HipError("launchGemmTS: block too large");
  if (n <= 0 || mb <= 0) return;
  kGemmTS<<<gridFor(n, kBlock, 4096), kBlock, 0, s>>>(A, ma, W, mb, n, lda, out, ldo, alpha, beta);
  HIP_CHECK(hipGetLastError());
}

__global__ void kSubMatrix(const double* C, int ldc, int rows, int cols, double* out) {
  for (int q = threadIdx.x; q < rows * cols; q += blockDim.x) out[q] = C[(q / cols) * ldc + q % cols];
}

void launchSubMatrix(const double* C, int ldc, int rows, int cols, double* out, hipStream_t s) {
  if (rows <= 0 || cols <= 0) return;
  kSubMatrix<<<1, kBlock, 0, s>>>(C, ldc, rows, cols, out);
  HIP_CHECK(hipGetLastError());
}

void launchRotate(double* X, int m, int64_t n, int64_t ld, const double* V, hipStream_t s) {
  if (m > 16) throw HipError("launchRotate: subspace too large");
  kRotate<<<gridFor(n, kBlock, 4096), kBlock, 0, s>>>(X, m, n, ld, V);
  HIP_CHECK(hipGetLastError());
}

void launchResidual(const double* sig, const double* b, const double* e, int64_t n, double* r, double* out,
                    DevMem& partial, hipStream_t s) {
  const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(kRedBlocks, (n + kBlock - 1) / kBlock));
  partial.reserve((size_t)nb * sizeof(double));
  kResidual<<<nb, kBlock, 0, s>>>(sig, b, e, n, r, partial.as<double>());
  kFinalSum<<<1, kBlock, 0, s>>>(partial.as<double>(), nb, 1, 1, out);
  HIP_CHECK(hipGetLastError());
}

void launchPrecondition(double* r, const double* diag, const double* e, int64_t n, hipStream_t s) {
  kPrecondition<<<gridFor(n, kBlock, 8192), kBlock, 0, s>>>(r, diag, e, n);
  HIP_CHECK(hipGetLastError());
}

void launchDot2(const double* b, const double* r, int64_t n, double* out, DevMem& partial, hipStream_t s) {
  const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(kRedBlocks, (n + kBlock - 1) / kBlock));
  partial.reserve((size_t)nb * 2 * sizeof(double));
  kDot2<<<nb, kBlock, 0, s>>>(b, r, n, partial.as<double>());
  kFinalSum<<<2, kBlock, 0, s>>>(partial.as<double>(), nb, 2, 2, out);
  HIP_CHECK(hipGetLastError());
}

void launchOrthoStep(double* r, const double* b, const double* d, int64_t n, hipStream_t s) {
  kOrthoStep<<<gridFor(n, kBlock, 8192), kBlock, 0, s>>>(r, b, d, n);
  HIP_CHECK(hipGetLastError());
}

void launchNorm2(const double* x, int64_t n, double* out, DevMem& partial, hipStream_t s) {
  const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(kRedBlocks, (n + kBlock - 1) / kBlock));
  partial.reserve((size_t)nb * sizeof(double));
  kNorm2<<<nb, kBlock, 0, s>>>(x, n, partial.as<double>());
  kFinalSum<<<1, kBlock, 0, s>>>(partial.as<double>(), nb, 1, 1, out);
  HIP_CHECK(hipGetLastError());
}

void launchScaleInto(double* dst, const double* src, const double* nrm2, int64_t n, hipStream_t s) {
  kScaleInto<<<gridFor(n, kBlock, 8192), kBlock, 0, s>>>(dst, src, nrm2, n);
  HIP_CHECK(hipGetLastError());
}

void launchAxpy(double* x, const double* y, double alpha, int64_t n, hipStream_t s) {
  kAxpy<<<gridFor(n, kBlock, 8192), kBlock, 0, s>>>(x, y, alpha, n);
  HIP_CHECK(hipGetLastError());
}

void launchSymmetrizeUpper(const double* G, int m, double* A, hipStream_t s) {
  kSymUpper<<<1, 256, 0, s>>>(G, m, A);
  HIP_CHECK(hipGetLastError());
}

void launchJacobiRR(const double* A, int m, double* w, double* V, hipStream_t s) {
  if (m > kJacobiMax) throw HipError("launchJacobiRR: subspace too large");
  kJacobiRR<<<1, 64, 0, s>>>(A, m, w, V);
  HIP_CHECK(hipGetLastError());
}

void launchJacobi(const double* A, int m, double* w, double* V, hipStream_t s) {
  if (m > kJacobiMax) throw HipError("launchJacobi: subspace too large");
  kJacobi<<<1, 64, 0, s>>>(A, m, w, V);
  HIP_CHECK(hipGetLastError());
}

}  // namespace shci
