// Host launchers for the SHCI kernels (implemented in kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "prims.h"
#include "shci_common.h"

namespace shci {

// --- determinant primitives ---------------------------------------------------
void launchEnergy(const DevSystem& S, const uint64_t* dets, int64_t n, double* out, hipStream_t s);
void launchHijPairs(const DevSystem& S, const uint64_t* bra, const uint64_t* ket, int64_t n, double* out,
                    hipStream_t s);

// --- membership table (open addressing; slots hold determinant indices) -----------
void hashClear(int32_t* table, uint64_t slots, hipStream_t s);
void hashInsert(const DevSystem& S, const uint64_t* dets, int64_t begin, int64_t end, int32_t* table,
                uint64_t slots, hipStream_t s);

// --- screening (heat-bath selection of new determinants) --------------------------
// cmax = sqrt(sum_r c_r^2) for every determinant; ci column-major with
// leading dimension ld.
void launchCmax(const double* ci, int nroots, int64_t ld, int64_t n, double* cmax, hipStream_t s);
// Emits candidate determinants not present in the table into `out`
// (capacity `cap` determinants); *counter receives the total number found
// (which may exceed cap: the caller then grows `out` and runs again).
// Screens the source determinants [begin, end) of `dets` (the whole
// variational space is probed for membership).
// Deterministic EN-PT2 candidates of sources [0, n) (SHCI/SHCIgetdeterminants.cpp:40-171):
// with fillStart == nullptr the count pass writes counts[src]; otherwise the
// fill pass writes each source's candidates (determinants outD, numerators
// H_ai c_i outN) from fillStart[src] in the reference's generation order.
// Only candidates outside the membership table and in partition `part` of
// `nparts` are kept.
void launchPtGen(const DevSystem& S, const uint64_t* dets, int64_t n, const double* ci, double eps2,
                 const int32_t* table, uint64_t slots, uint32_t nparts, uint32_t part, int64_t* counts,
                 const int64_t* fillStart, uint64_t* outD, double* outN, hipStream_t s);
// Semistochastic PT candidates of the sampled sources src[0..nsrc) with
// weights w (count pass when fillStart == nullptr): numerators H_ai w, second
// moments H_ai^2 w (w nmc/(nmc-1) - c) and the epsilon2Large flag
// (SHCI/SHCIgetdeterminants.cpp:889-1046), in generation order
void launchPtGenStoch(const DevSystem& S, const uint64_t* dets, const uint32_t* src, const double* w, int64_t nsrc,
                      const double* ci, double eps2, double eps2Large, double nmc, const int32_t* table, uint64_t slots,
                      int64_t* counts, const int64_t* fillStart, uint64_t* outD, double* outN, double* outN2,
                      uint8_t* outP, hipStream_t s);
// out[0..1] (device) = the two energy sums of one semistochastic iteration
// over m candidates sorted by determinant (payload idx)
void launchPtStochReduce(const DevSystem& S, const uint64_t* keys, const uint32_t* idx, const double* num,
                         const double* num2, const uint8_t* pres, int64_t m, double e0, double nmc, DevMem& partial,
                         double* out, hipStream_t s);
// time reversal: flags + scan of the open-shell determinants among
// dets[i0, i1); with block, the block of the plain basis E for them (the
// i1 - i0 determinants, then the flipped copies of the open-shell ones) to
// be placed at E index base, and per determinant its E index, its partner's
// (-1 when closed-shell) and its flip parity.  Returns the number of
// open-shell determinants (synchronises).
int64_t launchTrevBasis(const uint64_t* dets, int64_t i0, int64_t i1, int W, DevMem& flags, uint64_t* block,
                        int64_t base, int64_t* eIdx, int64_t* partner, double* parity, Workspace& ws, hipStream_t s);
// xE = U x and y = U^T yE for k columns (U: adapted basis -> E)
void launchTrevExpand(const double* x, int64_t n, int k, int64_t ld, const int64_t* eIdx, const int64_t* partner,
                      const double* parity, double* xE, int64_t ldE, hipStream_t s);
void launchTrevContract(const double* yE, int64_t ldE, int64_t n, int k, const int64_t* eIdx, const int64_t* partner,
                        const double* parity, double* y, int64_t ld, hipStream_t s);
// the unpacked coefficients in E (Treversal sign trev)
void launchTrevUnpack(const double* ci, int64_t n, int nroots, int64_t ld, const int64_t* partner,
                      const double* parity, int trev, double* out, int64_t ldE, hipStream_t s);
void launchIota(uint32_t* v, int64_t n, hipStream_t s);
// RDMs of <bra| ... |ket> over the n determinants and their stored
// connections (full symmetric CSR rp / cols), accumulated into the spatial
// g1 (norb^2) and g2 (norb^4) and, when not null, the spin-orbital s1
// (nso^2) and pair-indexed s2 ((nso(nso+1)/2)^2).  The caller zeroes g1, s1
// and s2; g2 is overwritten.  scratch holds rdmScratchDoubles(norb, W)
// doubles.
size_t rdmScratchDoubles(int norb, int W);
void launchRdm(int norb, int W, const uint64_t* dets, const int64_t* rp, const int32_t* cols, const double* bra,
               const double* ket, int64_t n, double* g1, double* g2, double* s1, double* s2, double* scratch,
               hipStream_t s);
// m candidates sorted by determinant (payload idx into num): out[0..2]
// (device) = E_PT, psi1 norm, number of external determinants
void launchPtReduce(const DevSystem& S, const uint64_t* keys, const uint32_t* idx, const double* num, int64_t m,
                    double e0, DevMem& partial, DevMem& counts, DevMem& starts, Workspace& ws, double* out,
                    hipStream_t s);
void launchScreen(const DevSystem& S, const uint64_t* dets, int64_t begin, int64_t end, const double* cmax,
                  double eps1, int ncore, int nact, const int32_t* table, uint64_t slots, uint64_t* out,
                  unsigned long long* counter, int64_t cap, hipStream_t s);

// One double excitation of a source determinant proposed to the learned
// ranking: spin-orbitals ci, cj (occupied) -> a, b (empty), heat-bath value v.
struct CandRecord {
  uint32_t src;
  uint8_t ci, cj, a, b;
  float v;
};
// singles only (the learned form's exact singles)
void launchScreenSingles(const DevSystem& S, const uint64_t* dets, int64_t begin, int64_t end, const double* cmax,
                         double eps1, int ncore, int nact, const int32_t* table, uint64_t slots, uint64_t* out,
                         unsigned long long* counter, int64_t cap, hipStream_t s);
// doubles whose |v| >= eps / loosen (eps = eps1 / cmax), as records; *counter
// receives the total (may exceed cap)
void launchScreenRecords(const DevSystem& S, const uint64_t* dets, int64_t begin, int64_t end, const double* cmax,
                         double eps1, double loosen, int ncore, int nact, CandRecord* rec,
                         unsigned long long* counter, int64_t cap, hipStream_t s);

// features of records (row r: occupation bits, one-hot ci, cj, a, b, log10
// |cmax[src]|; din >= 5 nso + 1, zero padded)
void launchCandFeatures(const DevSystem& S, const CandRecord* rec, int64_t n, const uint64_t* dets,
                        const double* cmax, float* X, int din, hipStream_t s);
// t = log10(|cmax[src]| |v|)
void launchCandTargets(const CandRecord* rec, int64_t n, const double* cmax, float* t, hipStream_t s);
// exact test of the records with score >= tau; accepted determinants not in
// the table are appended to out (cap, counter as launchScreen); counts[0] +=
// evaluated, counts[1] += accepted
void launchCandExact(const DevSystem& S, const CandRecord* rec, int64_t n, const float* score, float tau,
                     const uint64_t* dets, const double* cmax, double eps1, const int32_t* table, uint64_t slots,
                     uint64_t* out, unsigned long long* counter, int64_t cap, unsigned long long* counts,
                     hipStream_t s);

// for sampled records idx[0..m): flags bit 0 = passes the exact test, bit 1 =
// score >= tau; vals = scores
void launchCandSample(const CandRecord* rec, const float* score, const int64_t* idx, int m, const double* cmax,
                      double eps1, float tau, int* flags, float* vals, hipStream_t s);

// out[i] = rec[idx[i]]
void launchGatherRecords(const CandRecord* rec, const int64_t* idx, int64_t m, CandRecord* out, hipStream_t s);

// --- string tables and the Hamiltonian -----------------------------------------------
struct StringTables {
  int64_t nA = 0, nB = 0, n = 0;
  DevMem alpha, beta;           // per det: alpha / beta string (u64)
  DevMem aIdx, bIdx;            // per det: index into sorted unique strings (u32)
  DevMem sortedA, sortedB;      // unique strings
  DevMem aStart, bStart;        // group offsets (int64, nA+1 / nB+1)
  DevMem aMajorDet, bMajorDet;  // det index per group member (u32)
  DevMem aMajorB, bMajorA;      // partner string per group member (u64)
  DevMem aSinglesStart, aSingles;  // alpha strings one excitation apart (int64 / u32)
  DevMem bSinglesStart, bSingles;  // beta strings one excitation apart (int64 / u32)
  DevMem aSortB, aSortDet;         // alpha-major groups with members by beta string: beta (u64), det (u32)
  DevMem key, val;              // scratch
};
void buildStringTables(const DevSystem& S, const uint64_t* dets, int64_t n, StringTables& T, Workspace& ws,
                       hipStream_t s);

// Count (rowCount) or fill (cols/vals from fillStart[I]) the off-diagonal
// connections of every row, all J != I one or two excitations apart, except
// that rows I < nOld only take J >= nOld (the columns a Hamiltonian built over
// the first nOld determinants lacks).  Values are
// Hij(bra = D_min(I,J), ket = D_max(I,J)).
// lower = half storage: only the connections J < I of each row (rows
// I < nOld get none)
void launchCsrCount(const DevSystem& S, const uint64_t* dets, const StringTables& T, int64_t nOld,
                    int64_t* rowCount, hipStream_t s, bool lower = false);
void launchCsrFill(const DevSystem& S, const uint64_t* dets, const StringTables& T, int64_t nOld,
                   const int64_t* fillStart, int32_t* cols, double* vals, hipStream_t s, bool lower = false);
// direct sigma without a stored Hamiltonian (HmultDirect, SHCI/Hmult.h:153-382):
// y[:, r] = diag .* x[:, r] + sum_J Hij(D_I, D_J) x[J, r]; y must not overlap x
void launchSigmaDirect(const DevSystem& S, const uint64_t* dets, const StringTables& T, const double* diag,
                       const double* x, double* y, int k, int64_t ld, hipStream_t s);
// fill with each value stored as FP32 hi = fp32(v) and lo = fp32(v - hi)
void launchCsrFillSplit(const DevSystem& S, const uint64_t* dets, const StringTables& T, int64_t nOld,
                        const int64_t* fillStart, int32_t* cols, float* hi, float* lo, hipStream_t s,
                        bool lower = false);
// half storage (each pair once, in row max(I, J)): row offsets of rows
// [nOld, n] from the exclusive scan cnt of the new rows' lengths, and sigma
// on the stored rows applying every stored element to both rows (FP64
// atomic adds for the transposed half, so in no fixed order; the tiled
// copy's launchSigmaTiledSym is the default; y must not overlap x)
void launchLowerRowPtr(const int64_t* cnt, int64_t nOld, int64_t n, int64_t* rowPtr, hipStream_t s);
void launchSigmaLowerF64(const int64_t* rowPtr, const int32_t* cols, const double* vals, const double* diag, int64_t n,
                         const double* x, double* y, int k, int64_t ld, hipStream_t s);
void launchSigmaLowerF32(const int64_t* rowPtr, const int32_t* cols, const float* vals, const float* lo,
                         const double* diag, int64_t n, const double* x, double* y, int k, int64_t ld, hipStream_t s);
// incremental rebuild helpers: cnt[I] += old row length (I < nOld);
// fs[I] = rowPtr[I] + old row length; old rows copied to the front of the
// new rows
void launchNewRowLengths(const int64_t* oldRowPtr, int64_t nOld, int64_t n, int64_t* cnt, hipStream_t s);
void launchFillStart(const int64_t* oldRowPtr, int64_t nOld, int64_t n, const int64_t* rowPtr, int64_t* fs,
                     hipStream_t s);
void launchCopyOldRowsF64(const int64_t* oldRowPtr, const int32_t* oldCols, const double* oldVals, int64_t nOld,
                          const int64_t* rowPtr, int32_t* cols, double* vals, hipStream_t s);
void launchCopyOldRowsF32(const int64_t* oldRowPtr, const int32_t* oldCols, const float* oldHi, const float* oldLo,
                          int64_t nOld, const int64_t* rowPtr, int32_t* cols, float* hi, float* lo, hipStream_t s);
void launchCastValues(const double* v, float* out, int64_t n, hipStream_t s);
// sort the columns of every row (ascending), values carried along
void launchSortRows(const int64_t* rowPtr, int64_t n, int32_t* cols, double* vals, hipStream_t s);

// --- sigma = H x ------------------------------------------------------------------------
// y[:,r] = diag .* x[:,r] + offdiag(H) x[:,r] for r < k (column-major, ld);
// y must not overlap x
void launchSigmaF64(const int64_t* rowPtr, const int32_t* cols, const double* vals, const double* diag,
                    int64_t n, const double* x, double* y, int k, int64_t ld, hipStream_t s);
// FP32 values (`lo` null) or hi + lo FP32 pairs; FP64 accumulation either way
void launchSigmaF32(const int64_t* rowPtr, const int32_t* cols, const float* vals, const float* lo,
                    const double* diag, int64_t n, const double* x, double* y, int k, int64_t ld, hipStream_t s);
// hi = fp32(v), lo = fp32(v - hi)
void launchSplitHiLo(const double* v, float* hi, float* lo, int64_t n, hipStream_t s);

// --- alpha-major relabelled copy of H for sigma -----------------------------------------
// perm[m] = determinant at alpha-major position m (StringTables::aMajorDet);
// pos = inverse of perm
void launchInversePerm(const uint32_t* perm, int64_t n, uint32_t* pos, hipStream_t s);
// len[m] = length of row perm[m]
void launchPermRowLengths(const int64_t* rowPtr, const uint32_t* perm, int64_t n, int64_t* len, hipStream_t s);
// row m of the copy = row perm[m] with columns pos[J]; with bIdx (per
// determinant beta string index) and bPos (position in the beta-major order,
// inverse of StringTables::bMajorDet), columns J with the row's beta string
// become n + bPos[J], for the sigma's gather from [x alpha-major | x
// beta-major] (n < 2^30)
void launchPermRowsF64(const int64_t* rowPtr, const int32_t* cols, const double* vals, const uint32_t* perm,
                       const uint32_t* pos, const uint32_t* bIdx, const uint32_t* bPos, int64_t n,
                       const int64_t* rpP, int32_t* colsP, double* valsP, hipStream_t s);
void launchPermRowsF32(const int64_t* rowPtr, const int32_t* cols, const float* vals, const uint32_t* perm,
                       const uint32_t* pos, const uint32_t* bIdx, const uint32_t* bPos, int64_t n,
                       const int64_t* rpP, int32_t* colsP, float* valsP, hipStream_t s);
// Launch shape of the relabelled sigma: rows per work item (4, 8, 16 or 32)
// and static (one workgroup per work item) or persistent (workgroups pull
// work items from an atomic counter).  The struct's default is 8 rows,
// persistent (7.6 ms against 8.3 ms for 16 rows, static, on the
// 1.2M-determinant benchmark matrix with every column alpha-major).
struct SigmaConfig {
  int chunk = 8;
  bool persistent = true;
};
// Scratch for the permuted x of the relabelled sigma: a buffer grown on
// demand (mem), or a fixed device region of fixedBytes (the allocation-free
// entry point; a larger request throws).
struct XScratch {
  DevMem* mem = nullptr;
  void* fixed = nullptr;
  size_t fixedBytes = 0;
  void* get(size_t bytes) const;
};
// bytes of permuted x the relabelled sigma of k vectors needs (the largest
// pass: 1, 2, 4 or 8 interleaved vectors), xBytes = sizeof of its x type
size_t sigmaPermScratchBytes(int64_t n, bool split, int k, size_t xBytes);
// y = H x through the copy; x, y in determinant order (column-major, ld),
// y must not overlap x; k > 1 runs as SpMM over up to 8 interleaved
// vectors at a time.  xP is scratch for x in alpha-major order followed, when
// bPerm (the beta-major determinant order) is given, by x in beta-major
// order, for columns split by launchPermRows (FP64, or FP32 for the
// FP32-value form); counter (device, 64 unsigned words) is needed by the
// persistent configuration and is cleared by the launch itself.
void launchSigmaPermF64(const int64_t* rp, const int32_t* cols, const double* vals, const double* diag,
                        const uint32_t* perm, const uint32_t* bPerm, int64_t n, const double* x, double* y, int k,
                        int64_t ld, const XScratch& xP, const SigmaConfig& cfg, unsigned* counter, hipStream_t s);
void launchSigmaPermF32(const int64_t* rp, const int32_t* cols, const float* vals, const double* diag,
                        const uint32_t* perm, const uint32_t* bPerm, int64_t n, const double* x, double* y, int k,
                        int64_t ld, const XScratch& xP, const SigmaConfig& cfg, unsigned* counter, hipStream_t s);
// The fixed scratch region of the allocation-free sigma on the current
// device: its device address (looked up once per device with
// hipGetSymbolAddress) and size.  Every caller on one device shares it.
void* fixedSigmaScratch(size_t* bytes);

// Tiled copy of the relabelled Hamiltonian for the one-vector sigma: the
// alpha-major rows in blocks of 256, each block's entries grouped by tile of
// 1,024 consecutive permuted-x values and cut into wave instructions of up
// to 64 entries of one tile (descriptor: entry offset within the block,
// tile << 7 | count).  Per entry: (row within block) << 16 | (column within
// tile), and the value (FP64, or the FP32 hi part).
struct TiledHam {
  int64_t n = -1, blocks = 0, segments = 0, nnz = 0, instructions = 0;
  bool f32 = false, split = false, hasLo = false;
  // lo: the FP32 values' remainders (lo = fp32(v - hi)), when built with them
  DevMem segBase, segTile, segCnt, segStart, segInst, segTotal, fill, blockE, blockI, idx, vals, lo, inst;
  void release();
};
// device bytes buildTiledHam needs for nnz entries over n rows (upper bound
// on the per-segment tables)
size_t tiledHamBytes(int64_t n, int64_t nnz, bool f32);
// Builds T from H in determinant order (rowPtr, cols, values: vals64 or, for
// the FP32 form, vals32) through the alpha-major relabelling (perm: position
// -> determinant, pos: determinant -> position) and, with bIdx / bPos given,
// the split columns (the beta-major copy of x after the alpha-major one);
// or, with perm, pos, bIdx and bPos null, from the relabelled row copy
// itself (rows and columns already permuted; splitCols says whether its
// columns reach into the beta-major copy).  Segments of at least stageMin
// instructions are marked as large (the staged sigma copies their x tile to
// LDS; the half-storage sigma sums their transposed products there).  With
// vals32, valsLo (the FP32 remainders) is copied beside them when given.
// Returns false, with nothing built, when the permuted x has more tiles than
// the build's bitmap holds.
bool buildTiledHam(TiledHam& T, const int64_t* rowPtr, const int32_t* cols, const double* vals64,
                   const float* vals32, const uint32_t* perm, const uint32_t* pos, const uint32_t* bIdx,
                   const uint32_t* bPos, int64_t n, int64_t nnz, Workspace& ws, hipStream_t s,
                   bool splitCols = false, int stageMin = 8, const float* valsLo = nullptr);
// y = H x through the tiled copy (one vector, x and y in determinant order;
// bPerm = the beta-major determinant order when the copy has split columns)
// (step: instructions per wave per step, 8 or 16; stage: each wave copies
// the x tile of a large segment into LDS and gathers from there)
void launchSigmaTiled(const TiledHam& T, const double* diag, const uint32_t* perm, const uint32_t* bPerm,
                      const double* x, double* y, const XScratch& xP, bool persistent, unsigned* counter,
                      hipStream_t s, int step = 8, bool stage = false);
// y = H x through a tiled copy of half-stored H (each pair once): the copy's
// entries applied as stored and transposed; pos / bPos: determinant ->
// alpha-major / beta-major position; yT: nx + 1 words of scratch (nx: the
// permuted x's length); bound: the row bound set by launchSymRowBound.  The
// off-diagonal sums are 64-bit fixed point, summed with integer atomics:
// the result does not depend on their order.  exact (FP32 values): hi + lo
// values and x in FP64 (the copy must hold lo).
void launchSigmaTiledSym(const TiledHam& T, const double* diag, const uint32_t* perm, const uint32_t* bPerm,
                         const uint32_t* pos, const uint32_t* bPos, const double* x, double* y, const XScratch& xP,
                         unsigned long long* yT, const unsigned long long* bound, hipStream_t s, bool exact = false);
// kSigmaTiledSym's bound (the largest off-diagonal row sum of |H|) into
// *bound, from the row sums in acc (n words): the rows [nOld, n) of a half
// grown from nOld rows (CSR rowPtr, cols, values: vals64 or vals32, as
// stored) are added to the sums of the first nOld rows (nOld = 0 for a new
// half).
void launchSymRowBound(const int64_t* rowPtr, const int32_t* cols, const double* vals64, const float* vals32,
                       int64_t nOld, int64_t n, unsigned long long* acc, unsigned long long* bound, hipStream_t s);

// --- Davidson vector kernels ----------------------------------------------------------
// G[i*mb + j] = A[:,i] . B[:,j] for i < ma, j < mb (columns of length n, ld)
void launchGram(const double* A, int ma, const double* B, int mb, int64_t n, int64_t ld, double* G,
                DevMem& partial, hipStream_t s);
// X[:, :m] = X[:, :m] * V (V row-major m x m on device)
void launchRotate(double* X, int m, int64_t n, int64_t ld, const double* V, hipStream_t s);
// out[i*cols + j] = C[i*ldc + j] for i < rows, j < cols (small matrices)
void launchSubMatrix(const double* C, int ldc, int rows, int cols, double* out, hipStream_t s);
// G[i*mb + j] = A[:,i] . B[:,j] for ma, mb <= 64 (8 x 8 tiles; fixed-order sums)
void launchGramTiled(const double* A, int ma, const double* B, int mb, int64_t n, int64_t ld, double* G,
                     DevMem& partial, hipStream_t s);
// out[:, j] = alpha * sum_l A[:, l] W[l*mb + j] + beta * out[:, j] for j < mb
// (W row-major ma x mb on device, ma, mb <= 64); out must not overlap A
void launchGemmTS(const double* A, int ma, const double* W, int mb, int64_t n, int64_t lda, double* out,
                  int64_t ldo, double alpha, double beta, hipStream_t s);
// r = s - e * b ; out[0] = ||r||^2
void launchResidual(const double* sig, const double* b, const double* e, int64_t n, double* r, double* out,
                    DevMem& partial, hipStream_t s);
// r_k /= (e - diag_k) with the 1e-12 guard
void launchPrecondition(double* r, const double* diag, const double* e, int64_t n, hipStream_t s);
// out[0] = b.r, out[1] = b.b
void launchDot2(const double* b, const double* r, int64_t n, double* out, DevMem& partial, hipStream_t s);
// r -= (ov/nn) b  with ov = d[0], nn = d[1]  (r - ov*b/nn elementwise)
void launchOrthoStep(double* r, const double* b, const double* d, int64_t n, hipStream_t s);
// out[0] = ||x||^2
void launchNorm2(const double* x, int64_t n, double* out, DevMem& partial, hipStream_t s);
// dst = src / sqrt(nrm2[0])
void launchScaleInto(double* dst, const double* src, const double* nrm2, int64_t n, hipStream_t s);
// x += alpha * y
void launchAxpy(double* x, const double* y, double alpha, int64_t n, hipStream_t s);
// A[i*m+j] = G[min(i,j)*m + max(i,j)]  (upper triangle mirrored)
void launchSymmetrizeUpper(const double* G, int m, double* A, hipStream_t s);
// symmetric eigen-decomposition of the m x m matrix (row-major) by cyclic
// Jacobi, eigenvalues ascending in w, eigenvectors in the columns of V
void launchJacobi(const double* A, int m, double* w, double* V, hipStream_t s);
// the same with the rotations of a sweep in round-robin order (disjoint pairs
// at once): eigenpairs equal to rounding, fewer barriers (block Davidson)
void launchJacobiRR(const double* A, int m, double* w, double* V, hipStream_t s);

}  // namespace shci
