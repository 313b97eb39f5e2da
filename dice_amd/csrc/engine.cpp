// Engine: device-resident SHCI variational space, Hamiltonian and Davidson.
#include "engine.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <cstdio>
#include <ctime>
#include <sys/time.h>

namespace shci {

namespace {
template <typename T>
void upload(DevMem& m, const std::vector<T>& v, hipStream_t s) {
  m.reserve(std::max<size_t>(v.size() * sizeof(T), 8));
  if (!v.empty()) HIP_CHECK(hipMemcpyAsync(m.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}
}  // namespace

Engine::Engine() {}

Engine::~Engine() {
  clearGraphs();
  if (bdHost) (void)hipHostFree(bdHost);
  for (SigmaEvent& e : sigmaEvents) {
    (void)hipEventDestroy(e.start);
    (void)hipEventDestroy(e.stop);
  }
  if (stream && ownsStream) (void)hipStreamDestroy(stream);
}

void Engine::initShared(const Engine& parent) {
  hs = parent.hs;
  S = parent.S;
  S.trev = 0;
  stream = parent.stream;
  ownsStream = false;
  precision = parent.precision;
  relabel = parent.relabel;
  split = parent.split;
  tiledOn = parent.tiledOn;
  tiledFromH = parent.tiledFromH;
  tiledStage = parent.tiledStage;
  tiledStageMin = parent.tiledStageMin;
  tiledStep = parent.tiledStep;
  tiledPersistent = parent.tiledPersistent;
  lowerOnly = parent.lowerOnly;
  lowerTiled = parent.lowerTiled;
  lowerStageMin = parent.lowerStageMin;
  lowerSplit = parent.lowerSplit;
  storageMode = parent.storageMode;
  lowerByMemory = parent.lowerByMemory;
  autotune = parent.autotune;
  davidsonMode = parent.davidsonMode;
  sigmaCfg[0] = parent.sigmaCfg[0];
  sigmaCfg[1] = parent.sigmaCfg[1];
  incremental = false;
  counter.reserve(64);
}

void Engine::setDetsDevice(const uint64_t* d, int64_t cnt) {
  const int W = hs.W;
  dets.reserve((size_t)std::max<int64_t>(cnt, 1) * W * 8, stream);
  if (cnt) HIP_CHECK(hipMemcpyAsync(dets.p, d, cnt * W * 8, hipMemcpyDeviceToDevice, stream));
  n = cnt;
  slots = 0;
  ensureTable(std::max<int64_t>(n, 512));
  nDiag = 0;
  hamRows = -1;
  permRows = -1;
  tiledRows = -1;
  nNew = 0;
  HIP_CHECK(hipStreamSynchronize(stream));
}

// The plain basis E grows in blocks, one per call: the determinants added
// since the last call, then the flipped copies of the open-shell ones among
// them, so the sub-engine's Hamiltonian is extended incrementally like the
// plain one (the E order is not the unpacked order; unpackTrev builds that).
void Engine::buildTrev() {
  const int W = hs.W;
  const bool fresh = !tE || trevN < 0 || trevN > n;
  const int64_t i0 = fresh ? 0 : trevN;
  if (fresh) {
    tE.reset(new Engine());
    tE->initShared(*this);
    tE->incremental = true;
  }
  tE->precision = precision;
  if (i0 < n) {
    const int64_t cnt = n - i0;
    const int64_t nu = launchTrevBasis(dets.as<uint64_t>(), i0, n, W, trevFlags, nullptr, 0, nullptr, nullptr, nullptr,
                                       ws, stream);
    const size_t keep = (size_t)std::max<int64_t>(n, 1) * 8;
    trevEIdx.reserve(keep, stream, true);
    trevPartner.reserve(keep, stream, true);
    trevParity.reserve(keep, stream, true);
    DevMem block;
    block.reserve((size_t)(cnt + nu) * W * 8, stream);
    const int64_t base = fresh ? 0 : tE->n;
    launchTrevBasis(dets.as<uint64_t>(), i0, n, W, trevFlags, block.as<uint64_t>(), base, trevEIdx.as<int64_t>(),
                    trevPartner.as<int64_t>(), trevParity.as<double>(), ws, stream);
    if (fresh) {
      tE->setDetsDevice(block.as<uint64_t>(), cnt + nu);
    } else {
      tE->newDets.reserve((size_t)(cnt + nu) * W * 8, stream);
      HIP_CHECK(hipMemcpyAsync(tE->newDets.p, block.p, (size_t)(cnt + nu) * W * 8, hipMemcpyDeviceToDevice, stream));
      tE->nNew = cnt + nu;
      tE->appendNew();
      HIP_CHECK(hipStreamSynchronize(stream));
    }
  }
  tE->buildHamiltonian();
  trevN = n;
}

void Engine::sigmaTrev(const double* dx, double* dy, int k, int64_t ld, bool exact) {
  const int64_t nE = tE->n;
  trevX.reserve((size_t)std::max<int64_t>(nE * k, 1) * 8, stream);
  trevY.reserve((size_t)std::max<int64_t>(nE * k, 1) * 8, stream);
  launchTrevExpand(dx, n, k, ld, trevEIdx.as<int64_t>(), trevPartner.as<int64_t>(), trevParity.as<double>(),
                   trevX.as<double>(), nE, stream);
  tE->sigma(trevX.as<double>(), trevY.as<double>(), k, nE, exact);
  launchTrevContract(trevY.as<double>(), nE, n, k, trevEIdx.as<int64_t>(), trevPartner.as<int64_t>(),
                     trevParity.as<double>(), dy, ld, stream);
}

void Engine::unpackTrev(int sign) {
  // the reference's order: the determinants, then the flipped copies of the
  // open-shell ones in determinant order
  const int W = hs.W;
  const int64_t nu = launchTrevBasis(dets.as<uint64_t>(), 0, n, W, trevFlags, nullptr, 0, nullptr, nullptr, nullptr,
                                     ws, stream);
  const int64_t nE = n + nu;
  DevMem detsU, eIdxU, partnerU, parityU, out;
  detsU.reserve((size_t)std::max<int64_t>(nE, 1) * W * 8, stream);
  eIdxU.reserve((size_t)std::max<int64_t>(n, 1) * 8, stream);
  partnerU.reserve((size_t)std::max<int64_t>(n, 1) * 8, stream);
  parityU.reserve((size_t)std::max<int64_t>(n, 1) * 8, stream);
  launchTrevBasis(dets.as<uint64_t>(), 0, n, W, trevFlags, detsU.as<uint64_t>(), 0, eIdxU.as<int64_t>(),
                  partnerU.as<int64_t>(), parityU.as<double>(), ws, stream);
  out.reserve((size_t)std::max<int64_t>(nE * nroots, 1) * 8, stream);
  launchTrevUnpack(ci.as<double>(), n, nroots, n, partnerU.as<int64_t>(), parityU.as<double>(), sign,
                   out.as<double>(), nE, stream);
  ci.swap(out);
  trev = 0;
  S.trev = 0;
  tE.reset();
  trevN = -1;
  setDetsDevice(detsU.as<uint64_t>(), nE);
}

void Engine::load(const std::string& fcidump) {
  std::string e;
  HostSystem sys;
  if (!readFcidump(fcidump, sys, e)) throw std::invalid_argument(e);
  loadSystem(std::move(sys));
}

void Engine::loadSystem(HostSystem&& sys) {
  hs = std::move(sys);
  buildScreeningTables(hs);
  if (!stream) HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  upload(dH1, hs.h1, stream);
  upload(dI2, hs.I2, stream);
  upload(dDirect, hs.direct, stream);
  upload(dExch, hs.exch, stream);
  upload(dTsame, hs.tsame, stream);
  upload(dTopp, hs.topp, stream);
  upload(dSingles, hs.singles, stream);
  for (int k = 0; k < 2; k++) {
    upload(dHbVal[k], hs.hbVal[k], stream);
    upload(dHbPair[k], hs.hbPair[k], stream);
    upload(dHbStart[k], hs.hbStart[k], stream);
  }
  S.norb = hs.norb;
  S.nso = hs.nso;
  S.W = hs.W;
  S.nelec = hs.nelec;
  S.h1 = dH1.as<double>();
  S.I2 = dI2.as<double>();
  S.direct = dDirect.as<double>();
  S.exch = dExch.as<double>();
  S.tsame = dTsame.as<double>();
  S.topp = dTopp.as<double>();
  S.singles = dSingles.as<double>();
  for (int k = 0; k < 2; k++) {
    S.hbVal[k] = dHbVal[k].as<float>();
    S.hbPair[k] = dHbPair[k].as<uint32_t>();
    S.hbStart[k] = dHbStart[k].as<int64_t>();
  }
  S.maxEntryAbs = std::fabs(hs.maxEntry);
  // SHCI_SIGMA_LAYOUT=external keeps sigma on H in determinant order
  const char* layout = std::getenv("SHCI_SIGMA_LAYOUT");
  relabel = !(layout && std::strcmp(layout, "external") == 0);
  // SHCI_SIGMA_LAYOUT=alpha keeps every relabelled column alpha-major (no
  // beta-major copy of x)
  split = !(layout && std::strcmp(layout, "alpha") == 0);
  // the one-vector sigma on the tiled copy of the relabelled Hamiltonian
  // (SHCI_SIGMA_LAYOUT=rows: on the row-ordered copy instead)
  tiledOn = !(layout && std::strcmp(layout, "rows") == 0);
  // SHCI_SIGMA_PERSISTENT=0/1: static or persistent workgroups for the
  // one-vector sigma (the autotuner, when on, decides per Hamiltonian)
  if (const char* p = std::getenv("SHCI_SIGMA_PERSISTENT")) sigmaCfg[0].persistent = std::atoi(p) != 0;
  // SHCI_SIGMA_TILE_STEP=8/16: instructions per wave per step of the tiled sigma
  if (const char* p = std::getenv("SHCI_SIGMA_TILE_STEP")) tiledStep = std::atoi(p) >= 16 ? 16 : 8;
  // SHCI_SIGMA_TILE_BUILD=direct: the tiled copy always built from H through
  // the permutation (as when the row copy does not fit beside it)
  if (const char* p = std::getenv("SHCI_SIGMA_TILE_BUILD")) tiledFromH = std::strcmp(p, "direct") == 0;
  // SHCI_SIGMA_TILE_STAGE=0/1: the tiled sigma gathers a large segment's x
  // from a copy of its tile in LDS (1) or from global memory (0)
  if (const char* p = std::getenv("SHCI_SIGMA_TILE_STAGE")) tiledStage = std::atoi(p) != 0;
  // SHCI_SIGMA_TILE_STAGE_MIN=k: segments of at least k instructions are staged
  if (const char* p = std::getenv("SHCI_SIGMA_TILE_STAGE_MIN")) tiledStageMin = std::max(1, std::atoi(p));
  // SHCI_SIGMA_TILE_PERSISTENT=1: persistent workgroups for the tiled sigma
  if (const char* p = std::getenv("SHCI_SIGMA_TILE_PERSISTENT")) tiledPersistent = std::atoi(p) != 0;
  // SHCI_STORAGE=half / full: half storage always / never (the default
  // chooses per form)
  const char* storage = std::getenv("SHCI_STORAGE");
  if (storage && std::strcmp(storage, "half") == 0) storageMode = 1;
  if (storage && std::strcmp(storage, "full") == 0) storageMode = -1;
  // SHCI_HALF_SIGMA=rows: half storage's sigma on the stored rows (no tiled
  // copy); SHCI_HALF_STAGE_MIN=k: segments of at least k instructions sum
  // their transposed products in LDS first (1: every segment)
  if (const char* p = std::getenv("SHCI_HALF_SIGMA")) lowerTiled = std::strcmp(p, "rows") != 0;
  if (const char* p = std::getenv("SHCI_HALF_STAGE_MIN")) lowerStageMin = std::max(1, std::atoi(p));
  // SHCI_HALF_SPLIT=1: split columns in the half's tiled copy too
  if (const char* p = std::getenv("SHCI_HALF_SPLIT")) lowerSplit = std::atoi(p) != 0;
  uploadRestrictions();
  counter.reserve(64);
  // the allocation-free sigma's fixed scratch: its address on this device
  (void)fixedSigmaScratch(nullptr);
  HIP_CHECK(hipStreamSynchronize(stream));
}

void Engine::uploadRestrictions() {
  if (restrictions.size() > SHCI_MAX_RESTRICT) throw std::invalid_argument("too many restrict keywords (max 4)");
  std::vector<int8_t> mult(std::max<size_t>(restrictions.size(), 1) * hs.nso, 0);
  S.nrestrict = (int)restrictions.size();
  for (size_t r = 0; r < restrictions.size(); r++) {
    S.rmin[r] = restrictions[r].minElec;
    S.rmax[r] = restrictions[r].maxElec;
    for (int o : restrictions[r].orbs)
      if (o >= 0 && o < hs.nso) mult[r * hs.nso + o]++;
  }
  upload(dRmult, mult, stream);
  S.rmult = dRmult.as<int8_t>();
  if (restrictionsPT.size() > SHCI_MAX_RESTRICT) throw std::invalid_argument("too many restrict keywords (max 4)");
  std::vector<int8_t> multPt(std::max<size_t>(restrictionsPT.size(), 1) * hs.nso, 0);
  for (size_t r = 0; r < restrictionsPT.size(); r++)
    for (int o : restrictionsPT[r].orbs)
      if (o >= 0 && o < hs.nso) multPt[r * hs.nso + o]++;
  upload(dRmultPt, multPt, stream);
  HIP_CHECK(hipStreamSynchronize(stream));
}

DevSystem Engine::ptSystem() const {
  DevSystem P = S;
  P.nrestrict = (int)restrictionsPT.size();
  for (size_t r = 0; r < restrictionsPT.size(); r++) {
    P.rmin[r] = restrictionsPT[r].minElec;
    P.rmax[r] = restrictionsPT[r].maxElec;
  }
  P.rmult = dRmultPt.as<int8_t>();
  return P;
}

double Engine::ptDeterministic(const double* dCi, double e0, double eps2, double* psi1Norm, int64_t* nExternal,
                               int* nPartitions) {
  const int W = hs.W;
  double pt = 0.0, psi = 0.0;
  int64_t next = 0;
  uint32_t nparts = 1;
  if (n > 0) {
    ensureTable(n);
    const DevSystem P = ptSystem();
    ptCounts.reserve((size_t)(n + 1) * 8);
    int64_t* cnt = ptCounts.as<int64_t>();
    // per-source candidate counts of one partition, scanned into offsets
    auto countPass = [&](uint32_t parts, uint32_t part) {
      HIP_CHECK(hipMemsetAsync(cnt + n, 0, 8, stream));
      launchPtGen(P, dets.as<uint64_t>(), n, dCi, eps2, table.as<int32_t>(), slots, parts, part, cnt, nullptr,
                  nullptr, nullptr, stream);
      exclusiveScanI64(cnt, n + 1, ws, stream);
      int64_t total = 0;
      HIP_CHECK(hipMemcpyAsync(&total, cnt + n, 8, hipMemcpyDeviceToHost, stream));
      HIP_CHECK(hipStreamSynchronize(stream));
      return total;
    };
    const int64_t total = countPass(1, 0);
    // keys and their sort scratch, numerator, 32-bit payload and its scratch
    const size_t perCand = (size_t)16 * W + 16;
    const size_t budget = std::min<size_t>(deviceFreeBytes() / 3, size_t(96) << 30);
    const int64_t maxPer = std::max<int64_t>(1 << 20, std::min<int64_t>((int64_t)(budget / perCand), INT32_MAX));
    nparts = (uint32_t)std::max<int64_t>(1, (total + maxPer - 1) / maxPer);
    // SHCI_PT_PARTITIONS: at least this many partitions
    if (const char* e = std::getenv("SHCI_PT_PARTITIONS")) nparts = std::max<uint32_t>(nparts, (uint32_t)std::atoi(e));
    ptOut.reserve(64);
    double host[3];
    for (uint32_t part = 0; part < nparts; part++) {
      const int64_t m = nparts == 1 ? total : countPass(nparts, part);
      if (m == 0) continue;
      if (m > INT32_MAX) throw std::invalid_argument("PT partition exceeds 2^31 candidates");
      ptKeys.reserve((size_t)m * W * 8);
      ptNum.reserve((size_t)m * 8);
      ptIdx.reserve((size_t)m * 4);
      launchPtGen(P, dets.as<uint64_t>(), n, dCi, eps2, table.as<int32_t>(), slots, nparts, part, nullptr, cnt,
                  ptKeys.as<uint64_t>(), ptNum.as<double>(), stream);
      launchIota(ptIdx.as<uint32_t>(), m, stream);
      radixSortKeys(ptKeys.as<uint64_t>(), ptIdx.as<uint32_t>(), m, W, hs.nso, true, ws, stream);
      launchPtReduce(P, ptKeys.as<uint64_t>(), ptIdx.as<uint32_t>(), ptNum.as<double>(), m, e0, ptPartial, ptStartCnt,
                     ptStarts, ws, ptOut.as<double>(), stream);
      HIP_CHECK(hipMemcpyAsync(host, ptOut.p, 3 * 8, hipMemcpyDeviceToHost, stream));
      HIP_CHECK(hipStreamSynchronize(stream));
      pt += host[0];
      psi += host[1];
      next += (int64_t)host[2];
    }
  }
  if (psi1Norm) *psi1Norm = psi;
  if (nExternal) *nExternal = next;
  if (nPartitions) *nPartitions = (int)nparts;
  return pt;
}

double Engine::ptSemistochastic(const double* dCi, double e0, double eps2, double eps2Large, int sampleN, int nPTiter,
                                double targetError, double coreE, int root, FILE* log, double startofCalc,
                                double* error, int* iters, std::vector<double>* perIter) {
  if (nPTiter == 0) return 0.0;
  auto now = [] {
    timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
  };
  const int W = hs.W;
  const double E0c = e0 + coreE;
  if (log) fprintf(log, "Performing (semi)stochastic PT for state: %3d\n", root);
  // the deterministic part always runs; only its heading depends on the
  // threshold (SHCI/SHCIbasics.cpp:69-75)
  if (log && eps2Large < 999) fprintf(log, "\n1/ Deterministic calculation with epsilon2=%g\n", eps2Large);
  const double eptLarge = ptDeterministic(dCi, e0, eps2Large, nullptr, nullptr);
  if (log) {  // DoPerturbativeDeterministic's lines
    fprintf(log, "Deterministic PT calculation converged\n");
    fprintf(log, "PTEnergy: %18.10f\n", E0c + eptLarge);
    fprintf(log, "Time(s):  %10.2f\n", now() - startofCalc);
    fprintf(log, "\n2/ Stochastic calculation with epsilon2=%g\n", eps2);
  }
  const int N = (int)n;
  std::vector<double> c(N);
  HIP_CHECK(hipMemcpy(c.data(), dCi, (size_t)N * 8, hipMemcpyDeviceToHost));
  const int Nmc = sampleN;
  double cumulative = 0.0;
  for (int i = 0; i < N; i++) cumulative += std::fabs(c[i]);
  // alias tables (SHCI/SHCIsampledeterminants.cpp:58-78)
  std::vector<int> alias(N, 0);
  std::vector<double> prob(N, 0.0);
  {
    std::vector<double> larger, smaller;
    for (int i = 0; i < N; i++) {
      prob[i] = std::fabs(c[i]) * N / cumulative;
      (prob[i] < 1.0 ? smaller : larger).push_back(i);
    }
    while (!larger.empty() && !smaller.empty()) {
      const int l = (int)larger.back(), sm = (int)smaller.back();
      larger.pop_back();
      smaller.pop_back();
      alias[sm] = l;
      prob[l] = prob[l] - (1.0 - prob[sm]);
      (prob[l] < 1.0 ? smaller : larger).push_back(l);
    }
  }
  std::vector<int> sample(Nmc, -1);
  std::vector<double> wts(Nmc, 0.0);
  std::vector<uint32_t> src32;
  ensureTable(n);
  const DevSystem P = ptSystem();
  ptCounts.reserve((size_t)(Nmc + 1) * 8);
  ptOut.reserve(64);
  int64_t* cnt = ptCounts.as<int64_t>();
  if (log) fprintf(log, "%6s  %18s  %5s %18s %10s  %10s\n", "Iter", "EPTcurrent", "State", "EPTavg", "Error", "Time(s)");
  double avg = 0.0, avg2 = 0.0, stddev = 0.0;
  int currentIter = 0;
  for (int iter = 0; iter < nPTiter; iter++) {
    // sample_N2_alias (SHCI/SHCIsampledeterminants.cpp:148-176)
    std::fill(sample.begin(), sample.end(), -1);
    int distinct = 0;
    for (int index = 0; index < Nmc; index++) {
      // floor(r / RAND_MAX * N) (SHCIsampledeterminants.cpp:160), clamped:
      // r == RAND_MAX would index one past the alias and probability tables
      int detIndex = (int)std::floor(1. * ((double)rng.next() / (LibcRand::kMax)) * N);
      if (detIndex >= N) detIndex = N - 1;
      const double r = ((double)rng.next() / (LibcRand::kMax));
      if (r >= prob[detIndex]) detIndex = alias[detIndex];
      auto it = std::find(sample.begin(), sample.end(), detIndex);
      if (it == sample.end()) {
        sample[distinct] = detIndex;
        wts[distinct] = cumulative * c[detIndex] / std::fabs(c[detIndex]);
        distinct++;
      } else {
        wts[std::distance(sample.begin(), it)] += cumulative * c[detIndex] / std::fabs(c[detIndex]);
      }
    }
    for (int i = 0; i < Nmc; i++) wts[i] /= Nmc;
    src32.assign(sample.begin(), sample.begin() + distinct);
    ptSrc.reserve((size_t)std::max(distinct, 1) * 4);
    ptW.reserve((size_t)std::max(distinct, 1) * 8);
    HIP_CHECK(hipMemcpyAsync(ptSrc.p, src32.data(), (size_t)distinct * 4, hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(ptW.p, wts.data(), (size_t)distinct * 8, hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemsetAsync(cnt + distinct, 0, 8, stream));
    launchPtGenStoch(P, dets.as<uint64_t>(), ptSrc.as<uint32_t>(), ptW.as<double>(), distinct, dCi, eps2, eps2Large,
                     (double)Nmc, table.as<int32_t>(), slots, cnt, nullptr, nullptr, nullptr, nullptr, nullptr,
                     stream);
    exclusiveScanI64(cnt, distinct + 1, ws, stream);
    int64_t m = 0;
    HIP_CHECK(hipMemcpyAsync(&m, cnt + distinct, 8, hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    if (m > INT32_MAX) throw std::invalid_argument("semistochastic PT: more than 2^31 candidates in one iteration");
    double host[2] = {0.0, 0.0};
    if (m > 0) {
      ptKeys.reserve((size_t)m * W * 8);
      ptNum.reserve((size_t)m * 8);
      ptN2.reserve((size_t)m * 8);
      ptP.reserve((size_t)m);
      ptIdx.reserve((size_t)m * 4);
      launchPtGenStoch(P, dets.as<uint64_t>(), ptSrc.as<uint32_t>(), ptW.as<double>(), distinct, dCi, eps2,
                       eps2Large, (double)Nmc, table.as<int32_t>(), slots, nullptr, cnt, ptKeys.as<uint64_t>(),
                       ptNum.as<double>(), ptN2.as<double>(), ptP.as<uint8_t>(), stream);
      launchIota(ptIdx.as<uint32_t>(), m, stream);
      radixSortKeys(ptKeys.as<uint64_t>(), ptIdx.as<uint32_t>(), m, W, hs.nso, true, ws, stream);
      launchPtStochReduce(P, ptKeys.as<uint64_t>(), ptIdx.as<uint32_t>(), ptNum.as<double>(), ptN2.as<double>(),
                          ptP.as<uint8_t>(), m, e0, (double)Nmc, ptPartial, ptOut.as<double>(), stream);
      HIP_CHECK(hipMemcpyAsync(host, ptOut.p, 16, hipMemcpyDeviceToHost, stream));
      HIP_CHECK(hipStreamSynchronize(stream));
    }
    // the running mean and standard error (SHCI/SHCIbasics.cpp:404-440)
    currentIter++;
    const double est = -host[0] + host[1] + eptLarge;
    if (perIter) perIter->push_back(est);
    avg += est;
    avg2 += std::pow(est, 2);
    stddev = currentIter < 5
                 ? 1e4
                 : std::pow((currentIter * avg2 - std::pow(avg, 2)) / currentIter / (currentIter - 1) / currentIter, 0.5);
    if (log) {
      if (currentIter < 5)
        fprintf(log, "%6d  %18.10f  %5d %18.10f %10s  %10.2f\n", currentIter, E0c + est, root, E0c + avg / currentIter,
                "--", now() - startofCalc);
      else
        fprintf(log, "%6d  %18.10f  %5d %18.10f %10.2e  %10.2f\n", currentIter, E0c + est, root,
                E0c + avg / currentIter, stddev, now() - startofCalc);
      fflush(log);
    }
    if (stddev < targetError) {
      avg /= currentIter;
      if (log) {
        fprintf(log, "Semistochastic PT calculation converged\n");
        fprintf(log, "PTEnergy: %18.10f +/- %8.2e\n", E0c + avg, stddev);
        fprintf(log, "Time(s):  %10.2f\n", now() - startofCalc);
      }
      break;
    }
  }
  if (error) *error = stddev;
  if (iters) *iters = currentIter;
  return avg;
}

void Engine::rdm(const double* dBra, const double* dKet, std::vector<double>& g1, std::vector<double>& g2,
                 std::vector<double>* s1, std::vector<double>* s2) {
  const int nso = hs.nso, norb = nso / 2;
  if (direct) {  // the connections are not stored in direct mode: build them
    direct = false;
    hamRows = -1;
    buildHamiltonian();
    direct = true;
  } else {
    buildHamiltonian();
  }
  const size_t n1 = (size_t)norb * norb, n2 = n1 * n1, ns1 = (size_t)nso * nso;
  const size_t npair = (size_t)nso * (nso + 1) / 2, ns2 = npair * npair;
  auto zeroed = [&](DevMem& m, size_t count) {
    m.reserve(std::max<size_t>(count, 1) * 8, stream);
    HIP_CHECK(hipMemsetAsync(m.p, 0, count * 8, stream));
    return m.as<double>();
  };
  double* g1d = zeroed(rdmG1, n1);
  rdmG2.reserve(n2 * 8, stream);
  double* g2d = rdmG2.as<double>();
  double* s1d = s1 ? zeroed(rdmS1, ns1) : nullptr;
  double* s2d = s2 ? zeroed(rdmS2, ns2) : nullptr;
  rdmScratch.reserve(rdmScratchDoubles(norb, hs.W) * 8, stream);
  launchRdm(norb, hs.W, dets.as<uint64_t>(), rowPtr.as<int64_t>(), cols.as<int32_t>(), dBra, dKet, n, g1d, g2d, s1d,
            s2d, rdmScratch.as<double>(), stream);
  g1.resize(n1);
  g2.resize(n2);
  HIP_CHECK(hipMemcpyAsync(g1.data(), g1d, n1 * 8, hipMemcpyDeviceToHost, stream));
  HIP_CHECK(hipMemcpyAsync(g2.data(), g2d, n2 * 8, hipMemcpyDeviceToHost, stream));
  if (s1) {
    s1->resize(ns1);
    HIP_CHECK(hipMemcpyAsync(s1->data(), s1d, ns1 * 8, hipMemcpyDeviceToHost, stream));
  }
  if (s2) {
    s2->resize(ns2);
    HIP_CHECK(hipMemcpyAsync(s2->data(), s2d, ns2 * 8, hipMemcpyDeviceToHost, stream));
  }
  HIP_CHECK(hipStreamSynchronize(stream));
}

void Engine::ensureTable(int64_t need) {
  uint64_t want = 1024;
  while (want < (uint64_t)need * 2) want <<= 1;
  if (want <= slots) return;
  slots = want;
  table.reserve(slots * sizeof(int32_t));
  hashClear(table.as<int32_t>(), slots, stream);
  hashInsert(S, dets.as<uint64_t>(), 0, n, table.as<int32_t>(), slots, stream);
}

void Engine::setDets(const int64_t* h, int64_t cnt) {
  const int W = hs.W;
  dets.reserve((size_t)std::max<int64_t>(cnt, 1) * W * 8, stream);
  if (cnt) HIP_CHECK(hipMemcpyAsync(dets.p, h, cnt * W * 8, hipMemcpyHostToDevice, stream));
  n = cnt;
  slots = 0;
  ensureTable(std::max<int64_t>(n, 512));
  nDiag = 0;
  hamRows = -1;
  permRows = -1;
  tiledRows = -1;
  nNew = 0;
  trevN = -1;
  HIP_CHECK(hipStreamSynchronize(stream));
}

// Screening of every held determinant, in chunks of source determinants so
// the candidate buffer stays within a fixed budget: each chunk's candidates
// are sorted and made unique, then merged into the running unique set.
int64_t Engine::screen(const double* dCmax, double eps1) {
  const int W = hs.W;
  const int64_t budget = (int64_t)1 << 27;  // candidates held at once
  if (candCap == 0) candCap = std::min<int64_t>(budget, std::max<int64_t>(1 << 16, 8 * n));
  nNew = 0;
  int64_t begin = 0, chunk = n;
  while (begin < n) {
    const int64_t end = std::min(n, begin + chunk);
    cand.reserve((size_t)candCap * W * 8);
    HIP_CHECK(hipMemsetAsync(counter.p, 0, 8, stream));
    launchScreen(S, dets.as<uint64_t>(), begin, end, dCmax, eps1, ncore, nact, table.as<int32_t>(), slots,
                 cand.as<uint64_t>(), counter.as<unsigned long long>(), candCap, stream);
    unsigned long long found = 0;
    HIP_CHECK(hipMemcpyAsync(&found, counter.p, 8, hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    if ((int64_t)found > candCap) {
      if (candCap < budget) {  // grow the buffer first
        candCap = std::min<int64_t>(budget, (int64_t)found + (int64_t)found / 4 + 1024);
        cand.release();
      } else {  // then shrink the chunk
        if (chunk == 1) throw std::invalid_argument("one determinant yields more candidates than the buffer");
        chunk = std::max<int64_t>(1, (int64_t)((double)(end - begin) * candCap / (double)found * 0.8));
      }
      continue;
    }
    int64_t m = (int64_t)found;
    radixSortKeys(cand.as<uint64_t>(), nullptr, m, W, hs.nso, true, ws, stream);
    m = uniqueKeys(cand.as<uint64_t>(), m, W, ws, stream);
    if (nNew == 0) {
      newDets.reserve((size_t)std::max<int64_t>(m, 1) * W * 8);
      if (m) HIP_CHECK(hipMemcpyAsync(newDets.p, cand.p, m * W * 8, hipMemcpyDeviceToDevice, stream));
      nNew = m;
    } else if (m) {
      newDets.reserve((size_t)(nNew + m) * W * 8, stream, true);
      HIP_CHECK(hipMemcpyAsync(newDets.as<uint64_t>() + nNew * W, cand.p, m * W * 8, hipMemcpyDeviceToDevice,
                               stream));
      radixSortKeys(newDets.as<uint64_t>(), nullptr, nNew + m, W, hs.nso, true, ws, stream);
      nNew = uniqueKeys(newDets.as<uint64_t>(), nNew + m, W, ws, stream);
    }
    begin = end;
  }
  HIP_CHECK(hipStreamSynchronize(stream));
  return nNew;
}

void Engine::appendNew() {
  const int W = hs.W;
  if (nNew == 0) return;
  dets.reserve((size_t)(n + nNew) * W * 8, stream, true);
  HIP_CHECK(hipMemcpyAsync(dets.as<uint64_t>() + n * W, newDets.p, nNew * W * 8, hipMemcpyDeviceToDevice, stream));
  int64_t old = n;
  n += nNew;
  nNew = 0;
  uint64_t oldSlots = slots;
  ensureTable(n);
  if (slots == oldSlots) hashInsert(S, dets.as<uint64_t>(), old, n, table.as<int32_t>(), slots, stream);
}

// gettimeofday seconds (the clock of the run's elapsed-time columns)
static double wallClock() {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + 1e-6 * tv.tv_usec;
}

static double wallSeconds() {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void swapMem(DevMem& a, DevMem& b) { a.swap(b); }

// Builds H over the held determinants.  When a Hamiltonian over the first
// hamRows determinants exists, only the connections that involve the new
// determinants are enumerated and evaluated (new rows in full, old rows only
// their new columns); the old rows are copied forward.
void Engine::buildHamiltonian() {
  // diagonal for determinants without one
  diag.reserve((size_t)std::max<int64_t>(n, 1) * 8, stream, true);
  if (nDiag < n) launchEnergy(S, dets.as<uint64_t>() + nDiag * hs.W, n - nDiag, diag.as<double>() + nDiag, stream);
  nDiag = n;
  if (hamRows == n) return;
  if (trev) {  // the adapted basis (also in direct mode: the E Hamiltonian is stored)
    buildTrev();
    hamRows = n;
    return;
  }
  if (direct) {  // HmultDirect: sigma evaluates H from the string tables
    buildStringTables(S, dets.as<uint64_t>(), n, T, ws, stream);
    for (DevMem* m : {&rowPtr, &cols, &vals, &vals32, &valsLo, &rowPtrNew, &colsNew, &valsNew, &vals32New,
                      &valsLoNew, &permCols, &permVals})
      m->release();
    nnzOff = 0;
    permRows = -1;
    tiledRows = -1;
    tiled.release();
    hamRows = n;
    HIP_CHECK(hipStreamSynchronize(stream));
    return;
  }
  if (wantLower() != lowerOnly) {  // the storage changes: rebuilt from scratch in the other one
    for (DevMem* m : {&cols, &vals, &vals32, &valsLo, &colsNew, &valsNew, &vals32New, &valsLoNew, &permCols,
                      &permVals, &rowPtrNew})
      m->release();
    tiled.release();
    permRows = -1;
    tiledRows = -1;
    splitRows = -1;
    nnzOff = 0;
    hamRows = -1;
    lowerOnly = !lowerOnly;
  }
  const int64_t nOld = (incremental && hamRows > 0 && hamRows < n) ? hamRows : 0;
  if (lowerOnly) {
    buildLower(nOld);
    return;
  }
  // SHCI_TIMING: after the build, print per sub-phase the host time spent
  // issuing it and the GPU time between its stream events (no added
  // synchronisation, so the timed build runs as an untimed one does)
  const bool timing = std::getenv("SHCI_TIMING") != nullptr;
  struct Mark {
    const char* what;
    double host;
    hipEvent_t ev;
  };
  std::vector<Mark> marks;
  auto mark = [&](const char* what) {
    if (!timing) return;
    Mark m{what, wallSeconds(), nullptr};
    HIP_CHECK(hipEventCreate(&m.ev));
    HIP_CHECK(hipEventRecord(m.ev, stream));
    marks.push_back(m);
  };
  mark("start");
  buildStringTables(S, dets.as<uint64_t>(), n, T, ws, stream);
  mark("strings");
  rowPtrNew.reserve((size_t)(n + 1) * 8);
  int64_t* rp = rowPtrNew.as<int64_t>();
  const int64_t* orp = rowPtr.as<int64_t>();
  HIP_CHECK(hipMemsetAsync(rp + n, 0, 8, stream));
  launchCsrCount(S, dets.as<uint64_t>(), T, nOld, rp, stream);
  launchNewRowLengths(orp, nOld, n, rp, stream);
  exclusiveScanI64(rp, n + 1, ws, stream);
  int64_t nnzNew = 0;
  HIP_CHECK(hipMemcpyAsync(&nnzNew, rp + n, 8, hipMemcpyDeviceToHost, stream));
  HIP_CHECK(hipStreamSynchronize(stream));
  if (std::getenv("SHCI_TIMING"))
    std::fprintf(stderr, "# build n=%lld: %lld off-diagonal elements (both triangles)\n", (long long)n,
                 (long long)nnzNew);
  mark("count");
  const bool f32 = precision == 1;
  // FP64: cols + values; FP32 form: cols + hi + lo (the FP32 remainder lets
  // the refinement sigma reproduce the FP64 values)
  const size_t perEntry = 12;
  {
    size_t freeB = 0, totalB = 0;
    freeB = deviceFreeBytes(&totalB);
    const size_t need = (size_t)nnzNew * perEntry;
    const size_t reusable = colsNew.bytes + valsNew.bytes + vals32New.bytes + valsLoNew.bytes;
    // the two-triangle Hamiltonian with a half-size slack for the next
    // build does not fit, while half storage would: switch to half storage
    // for good (rebuilt from scratch)
    const size_t have = freeB + reusable + cols.bytes + vals.bytes + vals32.bytes + valsLo.bytes + permCols.bytes +
                        permVals.bytes + tiled.idx.bytes + tiled.vals.bytes;
    if (need > freeB + reusable && (size_t)nnzNew / 2 * perEntry * 3 / 2 <= have) {
      for (DevMem* m : {&cols, &vals, &vals32, &valsLo, &colsNew, &valsNew, &vals32New, &valsLoNew, &permCols,
                        &permVals, &rowPtrNew})
        m->release();
      tiled.release();
      permRows = -1;
      tiledRows = -1;
      lowerOnly = true;
      lowerByMemory = true;
      hamRows = -1;
      buildLower(0);
      return;
    }
    if (need > freeB + reusable)
      throw HipError("out of memory: the Hamiltonian of " + std::to_string(n) + " determinants has " +
                     std::to_string(nnzNew) + " off-diagonal elements (" + std::to_string(need >> 30) +
                     " GiB needed beside the previous one), " + std::to_string(freeB >> 30) + " GiB free");
  }
  fillStart.reserve((size_t)std::max<int64_t>(n, 1) * 8);
  launchFillStart(orp, nOld, n, rp, fillStart.as<int64_t>(), stream);
  colsNew.reserve((size_t)std::max<int64_t>(nnzNew, 1) * 4);
  if (f32) {
    vals32New.reserve((size_t)std::max<int64_t>(nnzNew, 1) * 4);
    valsLoNew.reserve((size_t)std::max<int64_t>(nnzNew, 1) * 4);
    mark("reserve");
    launchCopyOldRowsF32(orp, cols.as<int32_t>(), vals32.as<float>(), valsLo.as<float>(), nOld, rp,
                         colsNew.as<int32_t>(), vals32New.as<float>(), valsLoNew.as<float>(), stream);
    mark("copy old");
    launchCsrFillSplit(S, dets.as<uint64_t>(), T, nOld, fillStart.as<int64_t>(), colsNew.as<int32_t>(),
                       vals32New.as<float>(), valsLoNew.as<float>(), stream);
    mark("fill");
    swapMem(vals32, vals32New);
    swapMem(valsLo, valsLoNew);
    vals.release();
    valsNew.release();
  } else {
    valsNew.reserve((size_t)std::max<int64_t>(nnzNew, 1) * 8);
    mark("reserve");
    launchCopyOldRowsF64(orp, cols.as<int32_t>(), vals.as<double>(), nOld, rp, colsNew.as<int32_t>(),
                         valsNew.as<double>(), stream);
    mark("copy old");
    launchCsrFill(S, dets.as<uint64_t>(), T, nOld, fillStart.as<int64_t>(), colsNew.as<int32_t>(),
                  valsNew.as<double>(), stream);
    mark("fill");
    swapMem(vals, valsNew);
    vals32.release();
    valsLo.release();
    vals32New.release();
    valsLoNew.release();
  }
  swapMem(rowPtr, rowPtrNew);
  swapMem(cols, colsNew);
  nnzOff = nnzNew;
  hamRows = n;
  // the previous H's buffers are kept for the next build unless memory is short
  {
    size_t freeB = 0, totalB = 0;
    freeB = deviceFreeBytes(&totalB);
    if (freeB < totalB / 4) {
      colsNew.release();
      valsNew.release();
      vals32New.release();
      valsLoNew.release();
    }
  }
  buildPermuted();
  mark("relabel");
  HIP_CHECK(hipStreamSynchronize(stream));
  if (timing) {
    const double end = wallSeconds();
    std::string line = "  build n=" + std::to_string(n) + " (host ms / gpu ms):";
    char buf[96];
    for (size_t k = 1; k < marks.size(); k++) {
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, marks[k - 1].ev, marks[k].ev));
      std::snprintf(buf, sizeof buf, " %s %.1f/%.1f", marks[k].what, 1e3 * (marks[k].host - marks[k - 1].host), ms);
      line += buf;
    }
    std::snprintf(buf, sizeof buf, " | sync wait %.1f ms, total %.1f ms", 1e3 * (end - marks.back().host),
                  1e3 * (end - marks.front().host));
    std::fprintf(stderr, "%s%s\n", line.c_str(), buf);
    for (Mark& m : marks) (void)hipEventDestroy(m.ev);
  }
}

bool Engine::wantLower() const {
  if (lowerByMemory || storageMode > 0) return true;
  if (storageMode < 0) return false;
  return relabel && tiledOn && lowerTiled;
}

// Half storage: the rows of the determinants added since the last build
// (all rows when nOld = 0), each with its connections J < I, appended in
// place to the stored ones (old rows never gain an element: new
// determinants have larger indices).
void Engine::buildLower(int64_t nOld) {
  const bool f32 = precision == 1;
  // the tiled copy goes first: the stored rows may have to move to grow
  tiled.release();
  tiledRows = -1;
  splitRows = -1;
  buildStringTables(S, dets.as<uint64_t>(), n, T, ws, stream);
  rowPtrNew.reserve((size_t)(n + 1) * 8);
  int64_t* cnt = rowPtrNew.as<int64_t>();
  HIP_CHECK(hipMemsetAsync(cnt + n, 0, 8, stream));
  launchCsrCount(S, dets.as<uint64_t>(), T, nOld, cnt, stream, true);
  exclusiveScanI64(cnt, n + 1, ws, stream);
  rowPtr.reserve((size_t)(n + 1) * 8, stream, nOld > 0);
  int64_t* rp = rowPtr.as<int64_t>();
  launchLowerRowPtr(cnt, nOld, n, rp, stream);
  int64_t total = 0;
  HIP_CHECK(hipMemcpyAsync(&total, rp + n, 8, hipMemcpyDeviceToHost, stream));
  HIP_CHECK(hipStreamSynchronize(stream));
  if (std::getenv("SHCI_TIMING"))
    std::fprintf(stderr, "# half-storage build n=%lld: %lld stored off-diagonal elements\n", (long long)n,
                 (long long)total);
  const int64_t nnzOld = nOld ? nnzOff : 0;
  const int64_t added = total - nnzOld;
  {
    size_t totalB = 0;
    const size_t freeB = deviceFreeBytes(&totalB);
    const size_t need = (size_t)added * 12;
    if (need > freeB)
      throw HipError("out of memory: the half-stored Hamiltonian of " + std::to_string(n) + " determinants has " +
                     std::to_string(total) + " off-diagonal elements (" + std::to_string(need >> 30) +
                     " GiB more needed), " + std::to_string(freeB >> 30) + " GiB free");
  }
  cols.reserve((size_t)std::max<int64_t>(total, 1) * 4, stream, nOld > 0);
  if (f32) {
    vals32.reserve((size_t)std::max<int64_t>(total, 1) * 4, stream, nOld > 0);
    valsLo.reserve((size_t)std::max<int64_t>(total, 1) * 4, stream, nOld > 0);
    launchCsrFillSplit(S, dets.as<uint64_t>(), T, nOld, rp, cols.as<int32_t>(), vals32.as<float>(),
                       valsLo.as<float>(), stream, true);
    vals.release();
  } else {
    vals.reserve((size_t)std::max<int64_t>(total, 1) * 8, stream, nOld > 0);
    launchCsrFill(S, dets.as<uint64_t>(), T, nOld, rp, cols.as<int32_t>(), vals.as<double>(), stream, true);
    vals32.release();
    valsLo.release();
  }
  rowPtrNew.release();
  nnzOff = total;
  hamRows = n;
  permRows = -1;
  buildLowerTiled(nOld, nnzOld);
  HIP_CHECK(hipStreamSynchronize(stream));
}

// The tiled copy of the stored half, built from it through the alpha-major
// relabelling (and the split columns when their x copy fits), when it fits
// beside it with the permuted x and the transposed sums (nOld, nnzOld: the
// rows and entries kept from the previous build).
void Engine::buildLowerTiled(int64_t nOld, int64_t nnzOld) {
  tiledRows = -1;
  splitRows = -1;
  if (!relabel || !tiledOn || !lowerTiled || n == 0) return;
  const bool f32 = precision == 1;
  // (the FP32 form's copy holds the lo values too, for its exact sigma)
  const size_t tiledCopy = tiledHamBytes(n, nnzOff, f32) + (f32 ? (size_t)nnzOff * 4 : 0);
  const size_t base = (size_t)n * 4 + (size_t)n * 64;
  // permuted x (FP64 also in the FP32 form's exact sigma) and the
  // accumulators, plus bPos with split columns
  const size_t xPlain = (size_t)n * 16 + 16, xSplit = 2 * (size_t)n * 16 + (size_t)n * 4 + 16;
  const size_t have = deviceFreeBytes() + permX.bytes + lowerYT.bytes + bPos.bytes;
  if (base + tiledCopy + xPlain > have) return;
  const uint32_t* perm = T.aMajorDet.as<uint32_t>();
  permPos.reserve((size_t)n * 4);
  launchInversePerm(perm, n, permPos.as<uint32_t>(), stream);
  const bool doSplit = split && lowerSplit && 2 * n < 0x7fffffffLL && base + tiledCopy + xSplit <= have;
  if (doSplit) {
    bPos.reserve((size_t)n * 4);
    launchInversePerm(T.bMajorDet.as<uint32_t>(), n, bPos.as<uint32_t>(), stream);
  }
  if (!buildTiledHam(tiled, rowPtr.as<int64_t>(), cols.as<int32_t>(), f32 ? nullptr : vals.as<double>(),
                     f32 ? vals32.as<float>() : nullptr, perm, permPos.as<uint32_t>(),
                     doSplit ? T.bIdx.as<uint32_t>() : nullptr, doSplit ? bPos.as<uint32_t>() : nullptr, n, nnzOff,
                     ws, stream, false, lowerStageMin, f32 ? valsLo.as<float>() : nullptr)) {
    tiled.release();
    return;
  }
  const int64_t nx = doSplit ? 2 * n : n;
  lowerYT.reserve((size_t)(nx + 1) * 8);
  // the row sums grow with the half: when they cover exactly the rows kept
  // from the previous build, only the added rows are summed
  const bool grow = nOld > 0 && lowerAbsRows == nOld && lowerAbsNnz == nnzOld && lowerAbsF32 == f32;
  const int64_t r0 = grow ? lowerAbsRows : 0;
  lowerRowAbs.reserve((size_t)n * 8, stream, grow);
  lowerBound.reserve(8);
  launchSymRowBound(rowPtr.as<int64_t>(), cols.as<int32_t>(), f32 ? nullptr : vals.as<double>(),
                    f32 ? vals32.as<float>() : nullptr, r0, n, lowerRowAbs.as<unsigned long long>(),
                    lowerBound.as<unsigned long long>(), stream);
  lowerAbsRows = n;
  lowerAbsNnz = nnzOff;
  lowerAbsF32 = f32;
  tiledRows = n;
  if (doSplit) splitRows = n;
}

// Alpha-major relabelled copies of H (FP64 values, or the FP32 hi part) for
// the sigma kernels: the row-ordered copy (SpMM of the block Davidson, or
// one vector without the tiled copy), and the tiled copy (one vector; built
// from the row-ordered copy, which is then released unless the SpMM needs
// it).  A copy that does not fit beside H is skipped: with no tiled copy the
// one-vector sigma runs on the row-ordered one, with neither on H itself.
void Engine::buildPermuted() {
  permRows = -1;
  splitRows = -1;
  tiledRows = -1;
  if (!relabel || n == 0) return;
  const bool f32 = precision == 1;
  // (the block Davidson keeps the row copy: its SpMM needs it, and its
  // solves, about half as many sigma calls as the reference Davidson's, do
  // not repay building the tiled copy as well)
  const bool wantTiled = tiledOn && davidsonMode != 1;
  // a copy (columns, values, row offsets, inverse permutation) with 64
  // bytes per determinant of slack, and the permuted x of the sigma calls:
  // up to 8 interleaved vectors in the block Davidson, in alpha-major order
  // and, with split columns, beta-major order too (plus bPos)
  const size_t xt = f32 ? 4 : 8, kmax = davidsonMode == 1 ? 8 : 1;
  const size_t rowsCopy = (size_t)nnzOff * (4 + (f32 ? 4 : 8)) + (size_t)(n + 1) * 8;
  const size_t tiledCopy = tiledHamBytes(n, nnzOff, f32);
  const size_t base = (size_t)n * 4 + (size_t)n * 64;
  const size_t xPlain = (size_t)n * kmax * xt, xSplit = 2 * (size_t)n * kmax * xt + (size_t)n * 4;
  size_t freeB = 0, totalB = 0;
  freeB = deviceFreeBytes(&totalB);
  const size_t have = freeB + permCols.bytes + permVals.bytes + permX.bytes + bPos.bytes + tiled.idx.bytes +
                      tiled.vals.bytes + tiled.inst.bytes;
  // the tiled copy is built from the row-ordered one when both fit beside H;
  // when only the tiled copy fits it is built from H directly through the
  // permutation (a slower build)
  const bool viaRows = wantTiled && !tiledFromH && base + rowsCopy + tiledCopy + xPlain <= have;
  const bool fromH = wantTiled && !viaRows && base + tiledCopy + xPlain <= have;
  if (!fromH && base + rowsCopy + xPlain > have) {
    permCols.release();
    permVals.release();
    tiled.release();
    return;
  }
  const bool buildTiled = viaRows || fromH;
  if (fromH) {  // no row copy this time: its memory goes to the tiled copy
    permCols.release();
    permVals.release();
  }
  const uint32_t* perm = T.aMajorDet.as<uint32_t>();
  permPos.reserve((size_t)n * 4);
  launchInversePerm(perm, n, permPos.as<uint32_t>(), stream);
  // split columns only when their x copy fits too
  const size_t used = base + (fromH ? 0 : rowsCopy) + (buildTiled ? tiledCopy : 0);
  const bool doSplit = split && 2 * n < 0x7fffffffLL && used + xSplit <= have;
  if (doSplit) {
    bPos.reserve((size_t)n * 4);
    launchInversePerm(T.bMajorDet.as<uint32_t>(), n, bPos.as<uint32_t>(), stream);
  }
  const uint32_t* bIdx = doSplit ? T.bIdx.as<uint32_t>() : nullptr;
  const uint32_t* bp = doSplit ? bPos.as<uint32_t>() : nullptr;
  if (fromH) {
    if (buildTiledHam(tiled, rowPtr.as<int64_t>(), cols.as<int32_t>(), f32 ? nullptr : vals.as<double>(),
                      f32 ? vals32.as<float>() : nullptr, perm, permPos.as<uint32_t>(), bIdx, bp, n, nnzOff, ws,
                      stream, false, tiledStageMin)) {
      tiledRows = n;
      if (doSplit) splitRows = n;
    } else {
      tiled.release();
    }
    return;
  }
  permRp.reserve((size_t)(n + 1) * 8);
  int64_t* rp = permRp.as<int64_t>();
  HIP_CHECK(hipMemsetAsync(rp + n, 0, 8, stream));
  launchPermRowLengths(rowPtr.as<int64_t>(), perm, n, rp, stream);
  exclusiveScanI64(rp, n + 1, ws, stream);
  permCols.reserve((size_t)std::max<int64_t>(nnzOff, 1) * 4);
  if (f32) {
    permVals.reserve((size_t)std::max<int64_t>(nnzOff, 1) * 4);
    launchPermRowsF32(rowPtr.as<int64_t>(), cols.as<int32_t>(), vals32.as<float>(), perm, permPos.as<uint32_t>(),
                      bIdx, bp, n, rp, permCols.as<int32_t>(), permVals.as<float>(), stream);
  } else {
    permVals.reserve((size_t)std::max<int64_t>(nnzOff, 1) * 8);
    launchPermRowsF64(rowPtr.as<int64_t>(), cols.as<int32_t>(), vals.as<double>(), perm, permPos.as<uint32_t>(),
                      bIdx, bp, n, rp, permCols.as<int32_t>(), permVals.as<double>(), stream);
  }
  permRows = n;
  if (doSplit) splitRows = n;
  if (buildTiled) {
    if (buildTiledHam(tiled, rp, permCols.as<int32_t>(), f32 ? nullptr : permVals.as<double>(),
                      f32 ? permVals.as<float>() : nullptr, nullptr, nullptr, nullptr, nullptr, n, nnzOff, ws,
                      stream, doSplit, tiledStageMin)) {
      tiledRows = n;
      // the row-ordered copy served only the build: its buffers are kept
      // for the next build (mapping them again costs more than the copy)
      // unless memory is short
      permRows = -1;
      if (deviceFreeBytes() < totalB / 4) {
        HIP_CHECK(hipStreamSynchronize(stream));
        permCols.release();
        permVals.release();
      }
    } else {
      tiled.release();
    }
  } else {
    tiled.release();
  }
  // re-tune when the space has doubled since the last tuning (small spaces
  // keep the default shape: their sigma is too short to time reliably)
  if (autotune && n >= 20000 && n >= 2 * tunedRows) tuneSigma();
}

// Times every relabelled-sigma launch shape (rows per work item 4 / 8 / 16 /
// 32, static or persistent) on the current H, for one vector and, in the block
// Davidson, for an SpMM of four, and keeps the fastest of each.
void Engine::tuneSigma() {
  const int ks[2] = {1, 4};
  tuneX.reserve((size_t)n * 4 * 8);
  tuneY.reserve((size_t)n * 4 * 8);
  HIP_CHECK(hipMemsetAsync(tuneX.p, 0, (size_t)n * 4 * 8, stream));
  hipEvent_t a, b;
  HIP_CHECK(hipEventCreate(&a));
  HIP_CHECK(hipEventCreate(&b));
  const bool saveProfile = profileSigma;
  profileSigma = false;
  tuneLog.clear();
  char line[160];
  for (int which = 0; which < (davidsonMode == 1 ? 2 : 1); which++) {
    // one vector goes to the tiled copy when it exists, several to the row
    // copy's SpMM only when that copy was kept: nothing to tune otherwise
    if ((which == 0 && tiledRows == n) || permRows != n) continue;
    const int k = ks[which];
    SigmaConfig best = sigmaCfg[which];
    float bestMs = 1e30f;
    for (int chunk : {4, 8, 16, 32})
      for (bool pers : {false, true}) {
        sigmaCfg[which].chunk = chunk;
        sigmaCfg[which].persistent = pers;
        sigma(tuneX.as<double>(), tuneY.as<double>(), k, n, false);
        HIP_CHECK(hipEventRecord(a, stream));
        for (int r = 0; r < 2; r++) sigma(tuneX.as<double>(), tuneY.as<double>(), k, n, false);
        HIP_CHECK(hipEventRecord(b, stream));
        HIP_CHECK(hipEventSynchronize(b));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, a, b));
        ms *= 0.5f;
        std::snprintf(line, sizeof line, "k=%d chunk=%d %s %.3f ms; ", k, chunk, pers ? "persistent" : "static", ms);
        tuneLog += line;
        if (ms < bestMs) {
          bestMs = ms;
          best = sigmaCfg[which];
        }
      }
    sigmaCfg[which] = best;
    std::snprintf(line, sizeof line, "k=%d chosen chunk=%d %s. ", k, best.chunk, best.persistent ? "persistent" : "static");
    tuneLog += line;
  }
  profileSigma = saveProfile;
  HIP_CHECK(hipEventDestroy(a));
  HIP_CHECK(hipEventDestroy(b));
  tunedRows = n;
  if (std::getenv("SHCI_TIMING")) std::fprintf(stderr, "# sigma autotune (n=%lld): %s\n", (long long)n, tuneLog.c_str());
}

void Engine::clearGraphs() {
  for (auto& kv : graphs) {
    (void)hipGraphExecDestroy(kv.second.exec);
    (void)hipGraphDestroy(kv.second.graph);
  }
  graphs.clear();
}

// Block Davidson (the batching form).  Every iteration: Rayleigh-Ritz on the
// subspace V (m <= maxM columns, S = H V), Ritz vectors and residuals of the
// nr lowest roots; each root whose residual norm is >= tol contributes a
// preconditioned correction (the reference's diagonal preconditioner,
// Davidson.cpp:85-106); the corrections are orthogonalised against V (two
// classical Gram-Schmidt passes) and among themselves, appended, and their
// sigma is one SpMM.  When the subspace is full it restarts from the
// 2 nr + 2 lowest Ritz vectors.  The iteration body after the host's choice
// of corrections (SpMM, Gram matrix, eigensolve, Ritz vectors, residuals,
// copy of Ritz values and residual norms to pinned host memory) is captured
// once per (subspace size, block size) as a hipGraph and replayed.
std::vector<double> Engine::davidsonBlock(double* dX0, int nr, double tol, int& numIter, bool exactSigma) {
  const int64_t N = n;
  // subspace size: min(32, max(6 nr, nr + 16)) unless SHCI_BLOCK_MAXM (nr + 2 .. 32) sets it
  int maxDefault = std::min(32, std::max(6 * nr, nr + 16));
  if (const char* e = std::getenv("SHCI_BLOCK_MAXM")) maxDefault = std::max(nr + 2, std::min(32, std::atoi(e)));
  const int maxM = (int)std::min<int64_t>(N, maxDefault);
  if (nr > 8 || nr > maxM) throw std::invalid_argument("block Davidson supports up to 8 roots");
  hipStream_t s = stream;
  if (graphN != N) {
    clearGraphs();
    graphN = N;
  }
  auto reserveKeep = [&](DevMem& m, size_t bytes) {
    if (m.bytes < bytes) {
      clearGraphs();  // captured graphs hold the old addresses
      m.reserve(bytes);
    }
  };
  for (int q = 0; q < 2; q++) {
    reserveKeep(bdV[q], (size_t)N * maxM * 8);
    reserveKeep(bdS[q], (size_t)N * maxM * 8);
  }
  reserveKeep(bdT, (size_t)N * nr * 8);
  reserveKeep(bdX, (size_t)N * nr * 8);
  reserveKeep(bdHX, (size_t)N * nr * 8);
  reserveKeep(bdR, (size_t)N * nr * 8);
  reserveKeep(bdSmall, (size_t)8 * 4096 * 8);
  reserveKeep(dvPart, (size_t)16 * 256 * 64 * 8);  // largest Gram: 16 tiles x 256 blocks x 8 x 8
  reserveKeep(dvScal, 64 * 8);
  // x in alpha- and beta-major order, up to 8 interleaved vectors
  reserveKeep(permX, (size_t)2 * N * 8 * (precision == 1 && !exactSigma ? 4 : 8));
  reserveKeep(sigmaCounter, 256);
  if (!bdHost) HIP_CHECK(hipHostMalloc((void**)&bdHost, 64 * 8, hipHostMallocDefault));
  double* sm = bdSmall.as<double>();
  double* G = sm;                 // m x m Gram matrix (upper triangle used)
  double* A = sm + 4096;          // symmetrised
  double* C = sm + 2 * 4096;      // eigenvectors (row-major, column j = vector j)
  double* w = sm + 3 * 4096;      // eigenvalues
  double* Cn = sm + 4 * 4096;     // first nr (or keep) columns of C
  double* coef = sm + 5 * 4096;   // V^T T
  double* nrm = sm + 6 * 4096;    // residual norms^2 (nr), then correction norms^2
  double* sc = dvScal.as<double>();
  int cur = 0;
  double* V = bdV[0].as<double>();
  double* S = bdS[0].as<double>();
  double* Tm = bdT.as<double>();
  double* X = bdX.as<double>();
  double* HX = bdHX.as<double>();
  double* R = bdR.as<double>();
  auto col = [&](double* M, int c) { return M + (int64_t)c * N; };
  auto sync = [&]() { HIP_CHECK(hipStreamSynchronize(s)); };
  // orthonormalise T[:, 0:b] against V[:, 0:m] and among themselves; append
  // the columns that keep more than 1e-8 of their norm (and > 1e-10);
  // returns how many were appended
  auto appendBlock = [&](int m, int b) {
    for (int j = 0; j < b; j++) launchNorm2(col(Tm, j), N, nrm + 16 + j, dvPart, s);
    if (m > 0)
      for (int pass = 0; pass < 2; pass++) {
        launchGramTiled(V, m, Tm, b, N, N, coef, dvPart, s);
        launchGemmTS(V, m, coef, b, N, N, Tm, N, -1.0, 1.0, s);
      }
    for (int j = 0; j < b; j++) {
      for (int i = 0; i < j; i++) {
        launchDot2(col(Tm, i), col(Tm, j), N, sc, dvPart, s);
        launchOrthoStep(col(Tm, j), col(Tm, i), sc, N, s);
      }
      launchNorm2(col(Tm, j), N, nrm + j, dvPart, s);
    }
    std::vector<double> h(b), h0(b);
    HIP_CHECK(hipMemcpyAsync(h.data(), nrm, b * 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(h0.data(), nrm + 16, b * 8, hipMemcpyDeviceToHost, s));
    sync();
    int added = 0;
    for (int j = 0; j < b; j++) {
      if (!(std::sqrt(h[j]) > 1e-10) || !(std::sqrt(h[j]) > 1e-8 * std::sqrt(h0[j]))) continue;
      launchScaleInto(col(V, m + added), col(Tm, j), nrm + j, N, s);
      added++;
    }
    return added;
  };
  // sigma of the new columns, Rayleigh-Ritz, Ritz vectors and residuals;
  // leaves Ritz values and residual norms^2 in bdHost[0:nr], bdHost[32:32+nr]
  auto body = [&](int m0, int b) {
    const int m = m0 + b;
    sigma(col(V, m0), col(S, m0), b, N, exactSigma);
    launchGramTiled(V, m, S, m, N, N, G, dvPart, s);
    launchSymmetrizeUpper(G, m, A, s);
    launchJacobiRR(A, m, w, C, s);
    launchSubMatrix(C, m, m, nr, Cn, s);
    launchGemmTS(V, m, Cn, nr, N, N, X, N, 1.0, 0.0, s);
    launchGemmTS(S, m, Cn, nr, N, N, HX, N, 1.0, 0.0, s);
    for (int j = 0; j < nr; j++) launchResidual(col(HX, j), col(X, j), w + j, N, col(R, j), nrm + j, dvPart, s);
    HIP_CHECK(hipMemcpyAsync(bdHost, w, nr * 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(bdHost + 32, nrm, nr * 8, hipMemcpyDeviceToHost, s));
  };
  auto runBody = [&](int m0, int b) {
    blockMatvecs += b;
    const bool graphOk = useGraphs && !profileSigma;  // per-launch events are not captured
    if (!graphOk) {
      body(m0, b);
      sync();
      return;
    }
    char key[96];
    std::snprintf(key, sizeof key, "%d/%d/%d/%d/%d/%d", m0, b, nr, cur, exactSigma ? 1 : 0, precision);
    auto it = graphs.find(key);
    if (it == graphs.end()) {
      GraphEntry e{};
      HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      body(m0, b);
      HIP_CHECK(hipStreamEndCapture(s, &e.graph));
      HIP_CHECK(hipGraphInstantiate(&e.exec, e.graph, nullptr, nullptr, 0));
      it = graphs.emplace(key, e).first;
      graphCaptures++;
    }
    HIP_CHECK(hipGraphLaunch(it->second.exec, s));
    graphReplays++;
    sync();
  };

  // initial block: the given vectors (random where zero, as the reference's
  // setRandom), orthonormalised
  auto randomColumn = [&](double* x) {
    std::vector<double> h(N);
    for (int64_t k = 0; k < N; k++) h[k] = -1.0 + 2.0 * double(rng.next()) / double(LibcRand::kMax);
    HIP_CHECK(hipMemcpyAsync(x, h.data(), N * 8, hipMemcpyHostToDevice, s));
    sync();
  };
  for (int i = 0; i < nr; i++) {
    HIP_CHECK(hipMemcpyAsync(col(Tm, i), dX0 + (int64_t)i * N, N * 8, hipMemcpyDeviceToDevice, s));
    launchNorm2(col(Tm, i), N, nrm + i, dvPart, s);
  }
  {
    std::vector<double> h(nr);
    HIP_CHECK(hipMemcpyAsync(h.data(), nrm, nr * 8, hipMemcpyDeviceToHost, s));
    sync();
    for (int i = 0; i < nr; i++)
      if (!(std::sqrt(h[i]) > 1e-10)) randomColumn(col(Tm, i));
  }
  int m = 0;
  {
    const int added = appendBlock(0, nr);
    if (added < nr) throw std::invalid_argument("block Davidson: linearly dependent initial vectors");
    runBody(0, added);
    m = added;
  }
  numIter = 0;
  std::vector<double> theta(nr);
  while (true) {
    numIter++;
    std::vector<int> unconv;
    for (int j = 0; j < nr; j++) {
      theta[j] = bdHost[j];
      if (!(std::sqrt(std::max(bdHost[32 + j], 0.0)) < tol)) unconv.push_back(j);
    }
    if (unconv.empty() || m >= N || numIter > 800 * nr) {
      for (int j = 0; j < nr; j++)
        HIP_CHECK(hipMemcpyAsync(dX0 + (int64_t)j * N, col(X, j), N * 8, hipMemcpyDeviceToDevice, s));
      sync();
      return theta;
    }
    const int b = (int)unconv.size();
    // corrections (before any restart: they use R and the Ritz values w)
    for (int t = 0; t < b; t++) {
      const int j = unconv[t];
      HIP_CHECK(hipMemcpyAsync(col(Tm, t), col(R, j), N * 8, hipMemcpyDeviceToDevice, s));
      launchPrecondition(col(Tm, t), diag.as<double>(), w + j, N, s);
    }
    if (m + b > maxM) {  // thick restart: keep the lowest Ritz vectors
      const int keep = std::min(m, std::min(maxM - b, 2 * nr + 2));
      launchSubMatrix(C, m, m, keep, Cn, s);
      double* V2 = bdV[1 - cur].as<double>();
      double* S2 = bdS[1 - cur].as<double>();
      launchGemmTS(V, m, Cn, keep, N, N, V2, N, 1.0, 0.0, s);
      launchGemmTS(S, m, Cn, keep, N, N, S2, N, 1.0, 0.0, s);
      cur = 1 - cur;
      V = V2;
      S = S2;
      m = keep;
    }
    const int added = appendBlock(m, b);
    if (added == 0) {  // no new direction: the subspace is exhausted
      for (int j = 0; j < nr; j++)
        HIP_CHECK(hipMemcpyAsync(dX0 + (int64_t)j * N, col(X, j), N * 8, hipMemcpyDeviceToDevice, s));
      sync();
      return theta;
    }
    runBody(m, added);
    m += added;
  }
}

void Engine::sigma(const double* dx, double* dy, int k, int64_t ld, bool exact) {
  const bool f32 = precision == 1;
  SigmaEvent* ev = nullptr;
  if (profileSigma) {
    if (sigmaEventsUsed == sigmaEvents.size()) {
      SigmaEvent e{};
      HIP_CHECK(hipEventCreate(&e.start));
      HIP_CHECK(hipEventCreate(&e.stop));
      sigmaEvents.push_back(e);
    }
    ev = &sigmaEvents[sigmaEventsUsed++];
    // algorithmic bytes: stored values + column indices, row offsets,
    // diagonal, x read once and y written once per vector
    ev->bytes = (double)nnzOff * ((f32 ? (exact ? 8 : 4) : 8) + 4) + (double)(n + 1) * 8 + (double)n * 8 +
                2.0 * k * n * 8;
    HIP_CHECK(hipEventRecord(ev->start, stream));
  }
  if (trev) {
    sigmaTrev(dx, dy, k, ld, exact);
    if (ev) HIP_CHECK(hipEventRecord(ev->stop, stream));
    return;
  }
  if (direct) {
    launchSigmaDirect(S, dets.as<uint64_t>(), T, diag.as<double>(), dx, dy, k, ld, stream);
    if (ev) HIP_CHECK(hipEventRecord(ev->stop, stream));
    return;
  }
  if (lowerOnly && tiledRows == n && (!(f32 && exact) || tiled.hasLo)) {
    const bool sp = splitRows == n;
    const XScratch xs{&permX};
    for (int r = 0; r < k; r++)
      launchSigmaTiledSym(tiled, diag.as<double>(), T.aMajorDet.as<uint32_t>(),
                          sp ? T.bMajorDet.as<uint32_t>() : nullptr, permPos.as<uint32_t>(),
                          sp ? bPos.as<uint32_t>() : nullptr, dx + r * ld, dy + r * ld, xs,
                          lowerYT.as<unsigned long long>(), lowerBound.as<unsigned long long>(), stream,
                          f32 && exact);
    if (ev) HIP_CHECK(hipEventRecord(ev->stop, stream));
    return;
  }
  if (lowerOnly) {
    if (f32)
      launchSigmaLowerF32(rowPtr.as<int64_t>(), cols.as<int32_t>(), vals32.as<float>(),
                          exact ? valsLo.as<float>() : nullptr, diag.as<double>(), n, dx, dy, k, ld, stream);
    else
      launchSigmaLowerF64(rowPtr.as<int64_t>(), cols.as<int32_t>(), vals.as<double>(), diag.as<double>(), n, dx, dy, k,
                          ld, stream);
    if (ev) HIP_CHECK(hipEventRecord(ev->stop, stream));
    return;
  }
  const bool perm = permRows == n && !(f32 && exact);
  const SigmaConfig& cfg = sigmaCfg[k > 1 ? 1 : 0];
  // the tiled copy: one vector, or each of k vectors when the row-ordered
  // copy for the SpMM is absent
  const bool tiledSigma = (k == 1 || !perm) && tiledRows == n && !(f32 && exact);
  if (perm || tiledSigma) sigmaCounter.reserve(256);
  const uint32_t* bPerm = splitRows == n ? T.bMajorDet.as<uint32_t>() : nullptr;
  const XScratch xs{&permX};
  if (tiledSigma)
    for (int r = 0; r < k; r++)
      launchSigmaTiled(tiled, diag.as<double>(), T.aMajorDet.as<uint32_t>(), bPerm, dx + r * ld, dy + r * ld, xs,
                       tiledPersistent, sigmaCounter.as<unsigned>(), stream, tiledStep, tiledStage);
  else if (perm && f32)
    launchSigmaPermF32(permRp.as<int64_t>(), permCols.as<int32_t>(), permVals.as<float>(), diag.as<double>(),
                       T.aMajorDet.as<uint32_t>(), bPerm, n, dx, dy, k, ld, xs, cfg, sigmaCounter.as<unsigned>(),
                       stream);
  else if (perm)
    launchSigmaPermF64(permRp.as<int64_t>(), permCols.as<int32_t>(), permVals.as<double>(), diag.as<double>(),
                       T.aMajorDet.as<uint32_t>(), bPerm, n, dx, dy, k, ld, xs, cfg, sigmaCounter.as<unsigned>(),
                       stream);
  else if (f32)
    launchSigmaF32(rowPtr.as<int64_t>(), cols.as<int32_t>(), vals32.as<float>(), exact ? valsLo.as<float>() : nullptr,
                   diag.as<double>(), n, dx, dy, k, ld, stream);
  else
    launchSigmaF64(rowPtr.as<int64_t>(), cols.as<int32_t>(), vals.as<double>(), diag.as<double>(), n, dx, dy, k, ld,
                   stream);
  if (ev) HIP_CHECK(hipEventRecord(ev->stop, stream));
}

// the allocation-free half-storage sigma on the tiled copy: the permuted x
// (rounded up to 256 bytes), then the transposed sums and max|x|
static size_t symXBytes(int64_t n, bool split, bool f32) {
  return ((size_t)(split ? 2 * n : n) * (f32 ? 4 : 8) + 255) & ~(size_t)255;
}

size_t Engine::sigmaFixedScratchNeed(int k) const {
  if (lowerOnly)
    return tiledRows == n ? symXBytes(n, splitRows == n, precision == 1) + (size_t)((splitRows == n ? 2 : 1) * n + 1) * 8
                          : 0;
  const bool tiledSigma = (k == 1 || permRows != n) && tiledRows == n;
  if (permRows != n && !tiledSigma) return 0;  // the unrelabelled sigma needs no scratch
  return sigmaPermScratchBytes(n, splitRows == n, tiledSigma ? 1 : k, precision == 1 ? 4 : 8);
}

// sigma as sigma(dx, dy, k, ld, false) with the fixed scratch region in place
// of the handle's buffers: no allocation, no profiling events
void Engine::sigmaFixed(const double* dx, double* dy, int k, int64_t ld) {
  if (trev || direct) throw std::invalid_argument("the allocation-free sigma does not run in time-reversal or direct mode");
  const bool f32 = precision == 1;
  if (lowerOnly && tiledRows == n) {
    size_t cap = 0;
    unsigned char* base = (unsigned char*)fixedSigmaScratch(&cap);
    const bool sp = splitRows == n;
    const size_t xb = symXBytes(n, sp, f32);
    const XScratch xs{nullptr, base + 256, xb};
    auto* yT = (unsigned long long*)(base + 256 + xb);
    for (int r = 0; r < k; r++)
      launchSigmaTiledSym(tiled, diag.as<double>(), T.aMajorDet.as<uint32_t>(),
                          sp ? T.bMajorDet.as<uint32_t>() : nullptr, permPos.as<uint32_t>(),
                          sp ? bPos.as<uint32_t>() : nullptr, dx + r * ld, dy + r * ld, xs, yT,
                          lowerBound.as<unsigned long long>(), stream);
    return;
  }
  if (lowerOnly) {
    if (f32)
      launchSigmaLowerF32(rowPtr.as<int64_t>(), cols.as<int32_t>(), vals32.as<float>(), nullptr, diag.as<double>(), n,
                          dx, dy, k, ld, stream);
    else
      launchSigmaLowerF64(rowPtr.as<int64_t>(), cols.as<int32_t>(), vals.as<double>(), diag.as<double>(), n, dx, dy, k,
                          ld, stream);
    return;
  }
  const bool tiledSigma = (k == 1 || permRows != n) && tiledRows == n;
  if (permRows != n && !tiledSigma) {
    if (f32)
      launchSigmaF32(rowPtr.as<int64_t>(), cols.as<int32_t>(), vals32.as<float>(), nullptr, diag.as<double>(), n, dx,
                     dy, k, ld, stream);
    else
      launchSigmaF64(rowPtr.as<int64_t>(), cols.as<int32_t>(), vals.as<double>(), diag.as<double>(), n, dx, dy, k,
                     ld, stream);
    return;
  }
  size_t cap = 0;
  unsigned char* base = (unsigned char*)fixedSigmaScratch(&cap);
  const XScratch xs{nullptr, base + 256, cap - 256};
  const SigmaConfig& cfg = sigmaCfg[k > 1 ? 1 : 0];
  const uint32_t* bPerm = splitRows == n ? T.bMajorDet.as<uint32_t>() : nullptr;
  unsigned* ctr = (unsigned*)base;
  if (tiledSigma)
    for (int r = 0; r < k; r++)
      launchSigmaTiled(tiled, diag.as<double>(), T.aMajorDet.as<uint32_t>(), bPerm, dx + r * ld, dy + r * ld, xs,
                       tiledPersistent, ctr, stream, tiledStep, tiledStage);
  else if (f32)
    launchSigmaPermF32(permRp.as<int64_t>(), permCols.as<int32_t>(), permVals.as<float>(), diag.as<double>(),
                       T.aMajorDet.as<uint32_t>(), bPerm, n, dx, dy, k, ld, xs, cfg, ctr, stream);
  else
    launchSigmaPermF64(permRp.as<int64_t>(), permCols.as<int32_t>(), permVals.as<double>(), diag.as<double>(),
                       T.aMajorDet.as<uint32_t>(), bPerm, n, dx, dy, k, ld, xs, cfg, ctr, stream);
}

void Engine::readSigmaProfile(int64_t& calls, double& ms, double& bytes) {
  HIP_CHECK(hipStreamSynchronize(stream));
  calls = (int64_t)sigmaEventsUsed;
  ms = 0;
  bytes = 0;
  for (size_t i = 0; i < sigmaEventsUsed; i++) {
    float t = 0;
    HIP_CHECK(hipEventElapsedTime(&t, sigmaEvents[i].start, sigmaEvents[i].stop));
    ms += t;
    bytes += sigmaEvents[i].bytes;
  }
}

// Davidson with the semantics of SHCI/Davidson.cpp:111-348 (see the oracle
// for the line-by-line statement); vectors stay on the device, the host
// keeps the control flow and the m x m subspace scalars.
std::vector<double> Engine::davidson(double* dX0, int nr, int maxCopies, double tol, int& numIter, bool exactSigma) {
  const int64_t N = n;
  std::vector<double> eroots;
  dvB.reserve((size_t)N * maxCopies * 8);
  dvS.reserve((size_t)N * maxCopies * 8);
  dvR.reserve((size_t)N * 8);
  dvScal.reserve(64 * 8);
  dvG.reserve(512 * 8);
  dvW.reserve(32 * 8);
  dvV.reserve(256 * 8);
  double* B = dvB.as<double>();
  double* Sg = dvS.as<double>();
  double* r = dvR.as<double>();
  double* sc = dvScal.as<double>();  // [0..1] dot2, [2] norm2, [3..] residual norms
  double* G = dvG.as<double>();
  double* Wv = dvW.as<double>();
  double* V = dvV.as<double>();
  hipStream_t s = stream;
  auto col = [&](double* M, int c) { return M + (int64_t)c * N; };
  auto readScalars = [&](double* dst, const double* src, int cnt) {
    HIP_CHECK(hipMemcpyAsync(dst, src, cnt * 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  };
  auto norm = [&](const double* x) {
    double v;
    launchNorm2(x, N, sc + 2, dvPart, s);
    readScalars(&v, sc + 2, 1);
    return std::sqrt(v);
  };
  auto setRandom = [&](double* x) {
    std::vector<double> h(N);
    for (int64_t k = 0; k < N; k++) h[k] = -1.0 + (1.0 - (-1.0)) * double(rng.next()) / double(LibcRand::kMax);
    HIP_CHECK(hipMemcpyAsync(x, h.data(), N * 8, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));
  };
  auto dotHost = [&](const double* a, const double* b) {
    double v;
    launchGram(a, 1, b, 1, N, N, sc + 2, dvPart, s);
    readScalars(&v, sc + 2, 1);
    return v;
  };
  // initial vectors
  for (int i = 0; i < nr; i++) {
    HIP_CHECK(hipMemcpyAsync(col(B, i), dX0 + (int64_t)i * N, N * 8, hipMemcpyDeviceToDevice, s));
    if (norm(dX0 + (int64_t)i * N) < 1.e-10) {
      setRandom(col(B, i));
      launchNorm2(col(B, i), N, sc + 2, dvPart, s);
      launchScaleInto(col(B, i), col(B, i), sc + 2, N, s);
    }
  }
  for (int i = 0; i < nr; i++) {
    for (int j = 0; j < i; j++) launchAxpy(col(B, i), col(B, j), -dotHost(col(B, j), col(B, i)), N, s);
    if (norm(col(B, i)) < 1e-8) setRandom(col(B, i));
    for (int j = 0; j < i; j++) launchAxpy(col(B, i), col(B, j), -dotHost(col(B, j), col(B, i)), N, s);
    launchNorm2(col(B, i), N, sc + 2, dvPart, s);
    launchScaleInto(col(B, i), col(B, i), sc + 2, N, s);
  }
  int sigmaSize = 0, bsize = nr, convergedRoot = 0;
  numIter = 0;
  std::vector<double> w(16), hres(16);
  while (true) {
    for (int i = sigmaSize; i < bsize; i++) {
      sigma(col(B, i), col(Sg, i), 1, N, exactSigma);
      sigmaSize++;
    }
    launchGram(B, bsize, Sg, bsize, N, N, G, dvPart, s);
    launchSymmetrizeUpper(G, bsize, G + 256, s);
    launchJacobi(G + 256, bsize, Wv, V, s);
    launchRotate(B, bsize, N, N, V, s);
    launchRotate(Sg, bsize, N, N, V, s);
    const int cr0 = convergedRoot;
    // residual norms of the roots below convergedRoot, then of convergedRoot
    for (int i = 0; i < cr0; i++) launchResidual(col(Sg, i), col(B, i), Wv + i, N, r, sc + 3 + i, dvPart, s);
    launchResidual(col(Sg, cr0), col(B, cr0), Wv + cr0, N, r, sc + 3 + cr0, dvPart, s);
    HIP_CHECK(hipMemcpyAsync(w.data(), Wv, bsize * 8, hipMemcpyDeviceToHost, s));
    readScalars(hres.data(), sc + 3, cr0 + 1);
    const double* ePtr = Wv + cr0;  // ei = eigenvalue of the convergedRoot on entry
    for (int i = 0; i < cr0; i++) {
      if (std::sqrt(hres[i]) > tol) {
        convergedRoot = i;
        if (davidsonPrint) std::fprintf(davidsonPrint, "going back to converged root %d\n", i);
        break;
      }
    }
    double error;
    if (convergedRoot != cr0) {
      launchResidual(col(Sg, convergedRoot), col(B, convergedRoot), ePtr, N, r, sc + 2, dvPart, s);
      double v;
      readScalars(&v, sc + 2, 1);
      error = std::sqrt(v);
    } else {
      error = std::sqrt(hres[cr0]);
    }
    if (numIter % 100 == 0 && numIter > 0 && std::getenv("SHCI_TIMING"))
      std::fprintf(stderr, "# davidson n=%lld iteration %d root %d error %.3g\n", (long long)N, numIter,
                   convergedRoot, error);
    if (davidsonPrint) {  // davidsonDirect's progress lines (SHCI/Davidson.cpp:526-530)
      if (numIter == 0) std::fprintf(davidsonPrint, "nIter  Root               Energy                Error\n");
      std::fprintf(davidsonPrint, "%5i  %4i   %18.10g   %18.10g  %10.2f\n", numIter, convergedRoot, w[cr0], error,
                   wallClock() - davidsonPrintStart);
    }
    numIter++;
    if (bsize == N) {
      for (int i = 0; i < nr; i++) {
        HIP_CHECK(hipMemcpyAsync(dX0 + (int64_t)i * N, col(B, i), N * 8, hipMemcpyDeviceToDevice, s));
        eroots.push_back(w[i]);
        if (davidsonPrint)
          std::fprintf(davidsonPrint, "#niter:%3d root:%3d -> Energy : %18.10g  \n", numIter, i, w[i]);
      }
      HIP_CHECK(hipStreamSynchronize(s));
      return eroots;
    }
    // the reference's davidson caps at 800 iterations per root (and gives up
    // at 2000); davidsonDirect at 400, with a message (Davidson.cpp:305-309,
    // 548-555)
    const int cap = davidsonDirectRules ? 400 * nr : 800 * nr;
    if (error < tol || numIter > cap) {
      if (davidsonDirectRules && numIter > cap) {
        char m[128];
        std::snprintf(m, sizeof m, "Davidson calculation did not converge for root %3d, #iter %5d",
                      convergedRoot + 1, numIter);
        err = m;
        return {};
      }
      if (numIter > 2000 * nr) {
        err = "Davidson calculation did not converge";
        return eroots;
      }
      convergedRoot++;
      if (davidsonPrint)
        std::fprintf(davidsonPrint, "#niter:%3d root:%3d -> Energy : %18.10g  \n", numIter, convergedRoot - 1,
                     w[cr0]);
      if (convergedRoot == nr) {
        for (int i = 0; i < nr; i++) {
          HIP_CHECK(hipMemcpyAsync(dX0 + (int64_t)i * N, col(B, i), N * 8, hipMemcpyDeviceToDevice, s));
          eroots.push_back(w[i]);
        }
        HIP_CHECK(hipStreamSynchronize(s));
        return eroots;
      }
    }
    launchPrecondition(r, diag.as<double>(), ePtr, N, s);
    for (int i = 0; i < bsize; i++) {
      launchDot2(col(B, i), r, N, sc, dvPart, s);
      launchOrthoStep(r, col(B, i), sc, N, s);
    }
    launchNorm2(r, N, sc + 2, dvPart, s);
    if (bsize < maxCopies) {
      launchScaleInto(col(B, bsize), r, sc + 2, N, s);
      bsize++;
    } else {
      // restart (davidsonDirect keeps min(nroots + 3, maxCopies) vectors)
      bsize = davidsonDirectRules ? std::min(nr + 3, maxCopies) : nr + 3;
      sigmaSize = bsize - 1;
      launchScaleInto(col(B, bsize - 1), r, sc + 2, N, s);
    }
  }
}

}  // namespace shci
