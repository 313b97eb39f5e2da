// Device-resident state of one SHCI variational calculation.
#pragma once
#include <stdint.h>

#include <cstdio>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "host_io.h"
#include "kernels.h"
#include "mlp.h"

namespace shci {

struct IterRecord {
  int iter;
  double eps1;
  int64_t ndets;
  std::vector<double> e0;  // coreE added
  int davidsonIters;
  double seconds;       // cumulative wall time at the end of the iteration
  double tScreen = 0, tHam = 0, tDavidson = 0, tIter = 0;  // phase wall times (s)
  int64_t nnzLower = 0;
};

struct Engine {
  HostSystem hs;
  DevSystem S{};
  hipStream_t stream = nullptr;
  Workspace ws;
  int precision = 0;
  int ncore = 0, nact = 1000000;
  std::vector<Restriction> restrictions;
  std::vector<Restriction> restrictionsPT;  // the PT space's ("restrict", not "restrictv")
  // the C library's rand() sequence for the random Davidson guesses and the
  // PT alias sampling, seeded like the reference's srand(seed)
  LibcRand rng;

  // system tables
  DevMem dH1, dI2, dDirect, dExch, dTsame, dTopp, dSingles, dHbVal[2], dHbPair[2], dHbStart[2], dRmult, dRmultPt;

  // variational space
  int64_t n = 0;
  DevMem dets;                 // n x W
  DevMem table;                // membership slots (int32)
  uint64_t slots = 0;
  DevMem newDets;              // staged new determinants
  int64_t nNew = 0;
  DevMem cand;                 // screening candidates
  int64_t candCap = 0;
  DevMem counter;              // unsigned long long
  DevMem cmax;

  // Hamiltonian: diagonal + full symmetric off-diagonal CSR
  StringTables T;
  DevMem diag;
  int64_t nDiag = 0;
  DevMem rowPtr, cols, vals, vals32, valsLo;  // full symmetric off-diagonal CSR
  DevMem rowPtrNew, colsNew, valsNew, vals32New, valsLoNew, fillStart;  // incremental rebuild
  bool incremental = true;
  // alpha-major relabelled copy of H used by sigma (rows/columns permuted by
  // T.aMajorDet); valid when permRows == n
  bool relabel = true;
  DevMem permPos, permRp, permCols, permVals, permX;
  int64_t permRows = -1;
  // split columns: in the relabelled copy, columns whose determinant shares
  // the row's beta string point into a beta-major copy of x (n + position);
  // valid when splitRows == n
  bool split = true;
  DevMem bPos;
  int64_t splitRows = -1;
  // tiled copy for the one-vector sigma (blocks of 256 relabelled rows,
  // entries grouped by column tile); valid when tiledRows == n.  With it the
  // row-ordered relabelled copy is kept only for the block Davidson's SpMM.
  bool tiledOn = true;
  bool tiledFromH = false;  // build the tiled copy from H even when the row copy fits
  bool tiledStage = true;  // large segments gather x from an LDS copy of their tile
  int tiledStageMin = 16;   // instructions that make a segment large
  int tiledStep = 16;           // instructions per wave per step (8 or 16; 5.5 against 6.0 ms)
  bool tiledPersistent = false;  // one workgroup per block (6.4 ms) beat persistent ones (7.5 ms)
  TiledHam tiled;
  int64_t tiledRows = -1;
  int64_t nnzOff = 0, hamRows = -1;
  // half storage: each pair once, in row max(I, J) (the reference's
  // lower-triangular rows without the diagonal), grown in place by new rows
  // only.  storageMode 1 (SHCI_STORAGE=half, half_storage > 0) always, -1
  // (SHCI_STORAGE=full, half_storage < 0) never, 0: whenever sigma runs on
  // the relabelled tiled layout (the fastest form measured, for the
  // reference and the block Davidson, FP64 and FP32 values); and from the
  // build on which the two-triangle Hamiltonian does not fit
  // (lowerByMemory).  Sigma then adds the transposed half.
  bool lowerOnly = false;
  int storageMode = 0;
  bool lowerByMemory = false;
  bool wantLower() const;
  void buildLower(int64_t nOld);
  // with half storage, the tiled copy holds the stored half (built from it
  // when both fit) and sigma runs kSigmaTiledSym on it, its transposed sums
  // gathered in lowerYT (fixed point, permuted order; then max|x|); the FP32
  // form's exact sigma keeps the row kernel
  bool lowerTiled = true;
  // instructions that make a segment large (sigma on the 24-orbital
  // workload, 1: 3.42, 2: 3.22, 3: 3.10, 4: 3.09, 6: 3.28, 8: 3.59 ms; 16
  // instructions per wave per step instead of 8: 12% slower)
  int lowerStageMin = 4;
  // split columns in the half's tiled copy (sigma on the 24-orbital
  // workload: 3.08 ms with, 2.96 ms without, and a shorter build)
  bool lowerSplit = false;
  DevMem lowerYT;
  // the off-diagonal row sums of |H| (fixed point) over the first
  // lowerAbsRows rows' stored entries, grown with the half, and their
  // largest (kSigmaTiledSym's bound)
  DevMem lowerRowAbs, lowerBound;
  int64_t lowerAbsRows = -1, lowerAbsNnz = 0;
  bool lowerAbsF32 = false;
  void buildLowerTiled(int64_t nOld, int64_t nnzOld);

  // Davidson scratch
  DevMem dvB, dvS, dvR, dvScal, dvPart, dvG, dvW, dvV, dvX;
  // the reference's direct mode (`directdavidson`): no stored H, sigma from
  // the string tables (HmultDirect), and davidsonDirect's restart size,
  // iteration cap and progress lines (printed to davidsonPrint when set)
  bool direct = false;
  bool davidsonDirectRules = false;
  FILE* davidsonPrint = nullptr;
  double davidsonPrintStart = 0;

  // batching form: block Davidson whose new block's sigma is one SpMM,
  // iterations replayed from captured hipGraphs, sigma launch shape chosen
  // by timing on the current H
  int davidsonMode = 0;  // 0: the reference's single-vector Davidson, 1: block Davidson
  bool autotune = false;
  bool useGraphs = false;
  // [0]: one vector (4-row items, persistent: 6.71 ms against 6.93 ms for 8
  // rows on the 1.18M-determinant matrix with split columns), [1]: SpMM
  SigmaConfig sigmaCfg[2] = {SigmaConfig{4, true}, SigmaConfig{8, false}};
  int64_t tunedRows = 0;
  std::string tuneLog;
  DevMem sigmaCounter, tuneX, tuneY;
  DevMem bdV[2], bdS[2], bdT, bdX, bdHX, bdR, bdSmall;
  double* bdHost = nullptr;  // pinned: Ritz values and residual norms
  struct GraphEntry {
    hipGraph_t graph;
    hipGraphExec_t exec;
  };
  std::map<std::string, GraphEntry> graphs;
  int64_t graphN = -1;
  int64_t graphCaptures = 0, graphReplays = 0, blockMatvecs = 0;
  void clearGraphs();
  void tuneSigma();
  std::vector<double> davidsonBlock(double* dX0, int nr, double tol, int& numIter, bool exactSigma);

  // learned screening form: the plain screening of the first learnedLogIters
  // iterations logs (determinant, double excitation, log10 |c| |v|) samples
  // over the candidates whose heat-bath value passes eps / learnedLoosen;
  // the MLP trained on them predicts log10(|c| |v|) for the candidates of
  // later iterations and only those predicted above log10(eps1) -
  // learnedMargin are evaluated with the exact eps1 test; when a sample shows
  // that more than learnedMaxFraction would be, the iteration screens in
  // full.
  bool learned = false;
  int learnedLogIters = 2;
  double learnedLoosen = 10.0, learnedMargin = 0.8, learnedMaxFraction = 0.5;
  bool learnedControl = false;  // keep the untrained network (control runs)
  // gate: rank only when the candidate records per source times the
  // ranking cost per record (learnedNsPerRecord, an optimistic 2 ns) stay
  // below the exact screen's cost in the same iteration.  Checked before the
  // last logging iteration logs anything (records per source counted over a
  // few windows of sources; the first iteration starts from one
  // determinant); otherwise the form turns itself off (no more logging, no
  // training, full screening), recorded as one fallback record with
  // evaluated = 0
  bool learnedGate = false;
  double learnedNsPerRecord = 2.0;
  int64_t lastLogRecords = 0, lastLogSources = 0;
  bool learnedWorthIt(double exactScreenSeconds, int64_t sources, int iter);
  // records per source over a few windows of sources (counted, not kept)
  void learnedProbe(const double* dCmax, double eps1);
  int learnedHidden = 256;
  int64_t learnedMaxSamples = (int64_t)1 << 18;
  MlpModel mlp;
  bool mlpReady = false;
  std::vector<float> sampX, sampT;
  uint64_t learnedRng = 0x9e3779b97f4a7c15ull;
  DevMem recBuf, recScore, lcounts, featX, lIdx, lFlags, lVals;
  struct LearnedRecord {
    int iter = 0;
    bool fallback = false;
    int64_t candidates = 0, evaluated = 0, accepted = 0, samplePass = 0, sampleHit = 0, sampleSelected = 0,
            sampleTotal = 0;
    double seconds = 0;
  };
  std::vector<LearnedRecord> learnedRecs;
  float trainFirst = 0, trainLast = 0;
  double trainSeconds = 0;
  int64_t trainSamples = 0;
  int mlpDin() const { return (5 * hs.nso + 1 + 15) / 16 * 16; }
  void learnedLog(const double* dCmax, double eps1);
  void learnedTrain();
  int64_t screenLearned(const double* dCmax, double eps1, int iter);

  // deterministic EN-PT2 (SHCI/SHCIbasics.cpp:453-728) of root vector dCi
  // (device, n values) with variational energy e0 (no core energy): E_PT;
  // candidates are processed in hash partitions sized to the free memory
  DevMem ptCounts, ptKeys, ptNum, ptIdx, ptOut, ptPartial, ptStartCnt, ptStarts;
  double ptDeterministic(const double* dCi, double e0, double eps2, double* psi1Norm, int64_t* nExternal,
                         int* nPartitions = nullptr);
  DevSystem ptSystem() const;
  // semistochastic EN-PT2 (SHCI/SHCIbasics.cpp:55-452) of root vector dCi
  // (device, n values): the deterministic part at eps2Large, then sampled
  // iterations (C library rand(), as the reference) until the standard error
  // is below targetError or nPTiter iterations ran; log lines as the
  // reference's when log is set (coreE only shifts printed energies).
  // Returns the reference function's value (the average correction, or the
  // unaveraged sum when nPTiter runs out).
  DevMem ptN2, ptP, ptSrc, ptW;
  double ptSemistochastic(const double* dCi, double e0, double eps2, double eps2Large, int sampleN, int nPTiter,
                          double targetError, double coreE, int root, FILE* log, double startofCalc, double* error,
                          int* iters, std::vector<double>* perIter);

  // RDMs (SHCI/SHCIrdm.cpp:863-1100) of <bra| ... |ket> (device vectors of
  // n values) over the current determinants and their connections (built
  // when direct mode does not store them), copied to the host: the spatial
  // g1 (norb^2) and g2 (norb^4, g2[((P*n+Q)*n+R)*n+S] = the reference's
  // s2RDM(P*n+Q, R*n+S)) and, when asked, the spin-orbital s1 (nso^2) and s2
  // (pair-indexed, pair(p,q) = p(p+1)/2 + q)
  DevMem rdmG1, rdmG2, rdmS1, rdmS2, rdmScratch;
  void rdm(const double* dBra, const double* dKet, std::vector<double>& g1, std::vector<double>& g2,
           std::vector<double>* s1 = nullptr, std::vector<double>* s2 = nullptr);

  // time reversal (Treversal, SHCI/Determinants.cpp:32-121,
  // SHCI/SHCImakeHamiltonian.cpp:344-483): `dets` holds standard forms and
  // the Davidson works in the adapted basis; sigma = U^T H_E U with H_E the
  // plain Hamiltonian of E = dets + the flipped copies of the open-shell ones,
  // held by the sub-engine tE (same system tables and stream).  The
  // preconditioner keeps the plain diagonal energies, as the reference's.
  int trev = 0;
  std::unique_ptr<Engine> tE;
  DevMem trevFlags, trevEIdx, trevPartner, trevParity, trevX, trevY;
  int64_t trevN = -1;  // n the adapted basis was built for
  void buildTrev();
  void sigmaTrev(const double* dx, double* dy, int k, int64_t ld, bool exact);
  // unpackTrevState (SHCI/SHCIbasics.cpp:730-776): dets and ci become the
  // plain ones (Treversal sign sign); time reversal is off afterwards
  void unpackTrev(int sign);
  void setDetsDevice(const uint64_t* d, int64_t cnt);

  // last run
  std::vector<IterRecord> iters;
  std::vector<double> energies;
  std::vector<double> ptEnergies;  // E_var + E_PT per root when the run did the PT step
  DevMem ci;  // n x nroots, column-major, ld = n
  int nroots = 1;
  std::string err;
  double tFinalDavidson = 0;  // wall time of the last run's final tight Davidson (s)

  // optional timing of every sigma launch with HIP events on `stream`
  struct SigmaEvent {
    hipEvent_t start, stop;
    double bytes;
  };
  bool profileSigma = false;
  std::vector<SigmaEvent> sigmaEvents;
  size_t sigmaEventsUsed = 0;
  void resetSigmaProfile() { sigmaEventsUsed = 0; }
  // sum over recorded launches: calls, kernel milliseconds, algorithmic bytes
  void readSigmaProfile(int64_t& calls, double& ms, double& bytes);

  explicit Engine();
  ~Engine();
  bool ownsStream = true;
  // an engine on the system tables and stream of parent (owned by parent)
  void initShared(const Engine& parent);
  void load(const std::string& fcidump);  // throws
  void loadSystem(HostSystem&& sys);      // integrals already on the host; throws
  void uploadRestrictions();
  void setDets(const int64_t* h, int64_t cnt);
  void ensureTable(int64_t need);
  int64_t screen(const double* dCmax, double eps1);  // returns nNew
  void appendNew();
  void buildHamiltonian();
  void buildPermuted();
  void sigma(const double* dx, double* dy, int k, int64_t ld, bool exact);
  // sigma(dx, dy, k, ld, false) on the fixed device scratch (no allocation);
  // sigmaFixedScratchNeed = its bytes of permuted x (0: none needed)
  void sigmaFixed(const double* dx, double* dy, int k, int64_t ld);
  size_t sigmaFixedScratchNeed(int k) const;
  std::vector<double> davidson(double* dX0, int nr, int maxCopies, double tol, int& numIter, bool exactSigma);
  int64_t lowerNnz() const { return (lowerOnly ? nnzOff : nnzOff / 2) + n; }
};

}  // namespace shci
