// Device-side data layout and determinant primitives for the SHCI
// variational step on CDNA4 (gfx950).
//
// Determinants are spin-orbital bitstrings held in W = 1 or 2 64-bit words
// (W = nso/64 + 1, the reference's EffDetLen, SHCI/SHCI.cpp:236); bit k is
// spin-orbital k, even = alpha, odd = beta (SHCI/Determinants.h:166-200).
// Every kernel that produces a Hamiltonian element follows the reference's
// summation order term by term with FP contraction off, so values are
// bit-identical to SHCI/Determinants.cpp:252-300,406-505.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SHCI_MAX_RESTRICT 4
#define SHCI_MAX_NSO 126  // 2 words, 63 spatial orbitals

// Read-only system tables resident in HBM for the whole run.
struct DevSystem {
  int norb;            // spatial orbitals
  int nso;             // spin orbitals
  int W;               // words per determinant
  int nelec;
  const double* h1;    // norb x norb one-electron integrals (spin-free block)
  const double* I2;    // packed (pq|rs), index A(A+1)/2+B over pair indices
  const double* direct;   // norb x norb (ii|jj)
  const double* exch;     // norb x norb (ij|ji)
  const double* tsame;    // norb^3: T[a][i][j] = (ai|jj) - (aj|ji)  (same-spin j)
  const double* topp;     // norb^3: T[a][i][j] = (ai|jj)            (opposite-spin j)
  const double* singles;  // nso x nso heat-bath single bound Singles(p,q)
  const float* hbVal[2];      // [0] opposite spin, [1] same spin
  const uint32_t* hbPair[2];  // packed (a | b << 16)
  const int64_t* hbStart[2];  // npair + 1
  double maxEntryAbs;
  // RAS restrictions (SHCI/OccRestrictions.cpp): per restriction, the
  // multiplicity of each spin-orbital in its orbital list.
  int nrestrict;
  int rmin[SHCI_MAX_RESTRICT], rmax[SHCI_MAX_RESTRICT];
  const int8_t* rmult;  // nrestrict x nso
  // time reversal (Treversal): the variational space holds standard forms
  int trev;
};

// ---------------------------------------------------------------------------
// time reversal (SHCI/Determinants.h:245-301): spin flip swaps the alpha and
// beta bit of every spatial orbital

__host__ __device__ __forceinline__ uint64_t flipWord(uint64_t r) {
  return ((r & 0x5555555555555555ull) << 1) + ((r & 0xAAAAAAAAAAAAAAAAull) >> 1);
}

template <int W>
__host__ __device__ __forceinline__ bool hasUnpaired(const uint64_t* d) {
  bool u = false;
#pragma unroll
  for (int w = 0; w < W; w++) u |= ((d[w] & 0x5555555555555555ull) << 1) != (d[w] & 0xAAAAAAAAAAAAAAAAull);
  return u;
}

// standard form: of a determinant and its flip, the one that is larger as
// unsigned words compared from the highest word down (Determinant::isStandard)
template <int W>
__host__ __device__ __forceinline__ void makeStandardForm(uint64_t* d) {
  if (!hasUnpaired<W>(d)) return;
  bool flip = false;
#pragma unroll
  for (int w = W - 1; w >= 0; w--) {
    const uint64_t f = flipWord(d[w]);
    if (d[w] != f) {
      flip = d[w] < f;
      break;
    }
  }
  if (flip) {
#pragma unroll
    for (int w = 0; w < W; w++) d[w] = flipWord(d[w]);
  }
}

// Determinant::parityOfFlipAlphaBeta: (-1)^(doubly occupied orbitals) times
// (-1)^(beta electrons)
template <int W>
__host__ __device__ __forceinline__ double flipParity(const uint64_t* d) {
  int paired = 0, beta = 0;
#pragma unroll
  for (int w = 0; w < W; w++) {
    paired += __builtin_popcountll(((d[w] & 0x5555555555555555ull) << 1) & d[w]);
    beta += __builtin_popcountll(d[w] & 0xAAAAAAAAAAAAAAAAull);
  }
  const double p = (paired & 1) ? -1.0 : 1.0;
  return (beta & 1) ? -p : p;
}

// ---------------------------------------------------------------------------
// bit helpers

__device__ __forceinline__ int popc64(uint64_t x) { return __popcll(x); }

// mask of bits [s, e) within one word, 0 <= s < e <= 64
__device__ __forceinline__ uint64_t maskRange(int s, int e) {
  uint64_t hi = (e >= 64) ? ~0ull : ((1ull << e) - 1ull);
  return hi & (~0ull << s);
}

// number of occupied spin-orbitals k with lo < k < hi (lo < hi)
template <int W>
__device__ __forceinline__ int countBetween(const uint64_t* d, int lo, int hi) {
  int c = 0;
#pragma unroll
  for (int w = 0; w < W; w++) {
    int s = lo + 1 - 64 * w, e = hi - 64 * w;
    if (s < 0) s = 0;
    if (e > 64) e = 64;
    if (s < e) c += popc64(d[w] & maskRange(s, e));
  }
  return c;
}

template <int W>
__device__ __forceinline__ bool occ(const uint64_t* d, int k) {
  return (d[k >> 6] >> (k & 63)) & 1ull;
}

// packed two-electron integral with the spin selection of twoInt::operator()
// (SHCI/integral.h:86-101); spin-orbital indices
__device__ __forceinline__ double i2so(const DevSystem& S, int i, int j, int k, int l) {
  if (!(((i ^ j) & 1) == 0 && ((k ^ l) & 1) == 0)) return 0.0;
  int I = i >> 1, J = j >> 1, K = k >> 1, L = l >> 1;
  int IJ = I > J ? I * (I + 1) / 2 + J : J * (J + 1) / 2 + I;
  int KL = K > L ? K * (K + 1) / 2 + L : L * (L + 1) / 2 + K;
  long A = IJ > KL ? IJ : KL, B = IJ > KL ? KL : IJ;
  return S.I2[A * (A + 1) / 2 + B];
}

// Determinant::Energy (SHCI/Determinants.cpp:252-300), coreE = 0
template <int W>
__device__ double detEnergy(const DevSystem& S, const uint64_t* d) {
  double e = 0.0;
  const int n = S.norb;
#pragma unroll
  for (int w = 0; w < W; w++) {
    uint64_t bi = d[w];
    while (bi) {
      int I = 64 * w + __ffsll((unsigned long long)bi) - 1;
      bi &= bi - 1;
      e = __dadd_rn(e, S.h1[(I >> 1) * n + (I >> 1)]);
      // j > i: remaining bits of this word, then later words
      uint64_t bj = bi;
      int wj = w;
      while (true) {
        while (bj) {
          int J = 64 * wj + __ffsll((unsigned long long)bj) - 1;
          bj &= bj - 1;
          e = __dadd_rn(e, S.direct[(I >> 1) * n + (J >> 1)]);
          if (((I ^ J) & 1) == 0) e = __dsub_rn(e, S.exch[(I >> 1) * n + (J >> 1)]);
        }
        wj++;
        if (wj >= W) break;
        bj = d[wj];
      }
    }
  }
  return e;
}

// detEnergy with the tables from elsewhere (e.g. LDS): hdiag[K] = h1(K,K),
// D / X = the direct / exchange tables (n x n); the same operations in the
// same order, so the result is identical
template <int W>
__device__ double detEnergyTab(const double* hdiag, const double* D, const double* X, int n, const uint64_t* d) {
  double e = 0.0;
#pragma unroll
  for (int w = 0; w < W; w++) {
    uint64_t bi = d[w];
    while (bi) {
      int I = 64 * w + __ffsll((unsigned long long)bi) - 1;
      bi &= bi - 1;
      e = __dadd_rn(e, hdiag[I >> 1]);
      uint64_t bj = bi;
      int wj = w;
      while (true) {
        while (bj) {
          int J = 64 * wj + __ffsll((unsigned long long)bj) - 1;
          bj &= bj - 1;
          e = __dadd_rn(e, D[(I >> 1) * n + (J >> 1)]);
          if (((I ^ J) & 1) == 0) e = __dsub_rn(e, X[(I >> 1) * n + (J >> 1)]);
        }
        wj++;
        if (wj >= W) break;
        bj = d[wj];
      }
    }
  }
  return e;
}

// member Hij_1Excite(a, i) on ket (SHCI/Determinants.cpp:472-505)
template <int W>
__device__ double hijSingle(const DevSystem& S, const uint64_t* ket, int a, int i) {
  int lo = a < i ? a : i, hi = a < i ? i : a;
  int nb = countBetween<W>(ket, lo, hi);
  const int n = S.norb;
  if (((a ^ i) & 1) != 0) return 0.0;  // spin-changing single: every term is 0
  double e = S.h1[(a >> 1) * n + (i >> 1)];
  const double* ts = S.tsame + ((size_t)(a >> 1) * n + (i >> 1)) * n;
  const double* to = S.topp + ((size_t)(a >> 1) * n + (i >> 1)) * n;
  const int sa = a & 1;
#pragma unroll
  for (int w = 0; w < W; w++) {
    uint64_t b = ket[w];
    while (b) {
      int j = 64 * w + __ffsll((unsigned long long)b) - 1;
      b &= b - 1;
      e = __dadd_rn(e, ((j & 1) == sa) ? ts[j >> 1] : to[j >> 1]);
    }
  }
  return (nb & 1) ? -e : e;
}

// Determinant::Hij_2Excite(i, j, a, b) on ket (SHCI/Determinants.cpp:406-434)
template <int W>
__device__ double hijDouble(const DevSystem& S, const uint64_t* ket, int i, int j, int a, int b) {
  int I = i < j ? i : j, J = i < j ? j : i, A = a < b ? a : b, B = a < b ? b : a;
  int c = countBetween<W>(ket, I < A ? I : A, I < A ? A : I);
  c += countBetween<W>(ket, J < B ? J : B, J < B ? B : J);
  if (A > J || B < I) c++;
  double v = __dsub_rn(i2so(S, A, I, B, J), i2so(S, A, J, B, I));
  return (c & 1) ? -v : v;
}

// Hij(bra, ket) for determinants one or two excitations apart
// (SHCI/Determinants.cpp:564-631); returns 0 otherwise.
template <int W>
__device__ double hijElement(const DevSystem& S, const uint64_t* bra, const uint64_t* ket) {
  int cre[2] = {0, 0}, des[2] = {0, 0}, nc = 0, nd = 0;
#pragma unroll
  for (int w = 0; w < W; w++) {
    uint64_t u = bra[w] ^ ket[w];
    uint64_t bb = u & bra[w], kk = u & ket[w];
    while (bb) {
      int p = 64 * w + __ffsll((unsigned long long)bb) - 1;
      bb &= bb - 1;
      if (nc < 2) cre[nc] = p;
      nc++;
    }
    while (kk) {
      int p = 64 * w + __ffsll((unsigned long long)kk) - 1;
      kk &= kk - 1;
      if (nd < 2) des[nd] = p;
      nd++;
    }
  }
  if (nc == 1) return hijSingle<W>(S, ket, cre[0], des[0]);
  if (nc == 2) return hijDouble<W>(S, ket, des[0], des[1], cre[0], cre[1]);
  return 0.0;
}

// determinant hash for the membership table
template <int W>
__device__ __forceinline__ uint64_t detHash(const uint64_t* d) {
  uint64_t h = 0x9E3779B97F4A7C15ull;
#pragma unroll
  for (int w = 0; w < W; w++) {
    uint64_t z = d[w] + h;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    h = z ^ (z >> 31);
  }
  return h;
}

// wavefront helpers (64 lanes)
__device__ __forceinline__ int laneId() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemaskLt() {
  int l = __lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}
