// Learned screening form: sample logging, training and ranked screening.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "engine.h"

namespace shci {

namespace {

uint64_t splitmix(uint64_t& x) {
  uint64_t z = (x += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

constexpr int64_t kRecordCap = (int64_t)1 << 25;   // records held at once
constexpr int64_t kScoreChunk = (int64_t)1 << 20;  // records per network pass
constexpr int kSample = 1 << 16;                   // records sampled for tau and recall

}  // namespace

// Logs training samples from the current (plain) iteration: the candidate
// records of 16 windows of 256 source determinants at random offsets (one
// window holding the whole space when it is no larger); from
// each window a random subset of records (features and targets
// log10(|c| |v|)) so that the iterations that log fill learnedMaxSamples.
void Engine::learnedLog(const double* dCmax, double eps1) {
  const int din = mlpDin();
  counter.reserve(64);
  const int64_t window = std::min<int64_t>(n, 256);
  // (a space no larger than one window is logged once)
  const int windows = n > window ? 16 : 1;
  const int64_t perWindow = std::max<int64_t>(1, learnedMaxSamples / windows / std::max(1, learnedLogIters));
  DevMem picked;
  picked.reserve((size_t)perWindow * sizeof(CandRecord));
  lIdx.reserve((size_t)perWindow * sizeof(int64_t));
  featX.reserve((size_t)perWindow * din * sizeof(float));
  lVals.reserve((size_t)perWindow * sizeof(float));
  std::vector<float> hx, ht;
  std::vector<int64_t> idx;
  lastLogRecords = 0;
  lastLogSources = 0;
  for (int w = 0; w < windows; w++) {
    const int64_t begin = n > window ? (int64_t)(splitmix(learnedRng) % (uint64_t)(n - window + 1)) : 0;
    // count the window's records, then hold them (the buffer grows to the
    // largest window seen, at most kRecordCap records)
    unsigned long long found = 0;
    HIP_CHECK(hipMemsetAsync(counter.p, 0, 8, stream));
    launchScreenRecords(S, dets.as<uint64_t>(), begin, begin + window, dCmax, eps1, learnedLoosen, ncore, nact,
                        nullptr, counter.as<unsigned long long>(), 0, stream);
    HIP_CHECK(hipMemcpyAsync(&found, counter.p, 8, hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    const int64_t total = std::min<int64_t>((int64_t)found, kRecordCap);
    if (total > 0) {
      recBuf.reserve((size_t)total * sizeof(CandRecord));
      HIP_CHECK(hipMemsetAsync(counter.p, 0, 8, stream));
      launchScreenRecords(S, dets.as<uint64_t>(), begin, begin + window, dCmax, eps1, learnedLoosen, ncore, nact,
                          recBuf.as<CandRecord>(), counter.as<unsigned long long>(), total, stream);
    }
    lastLogRecords += (int64_t)found;
    lastLogSources += std::min<int64_t>(window, n - begin);
    if (total == 0) continue;
    const int64_t m = std::min(total, perWindow);
    idx.resize(m);
    for (int64_t i = 0; i < m; i++) idx[i] = total <= perWindow ? i : (int64_t)(splitmix(learnedRng) % (uint64_t)total);
    HIP_CHECK(hipMemcpyAsync(lIdx.p, idx.data(), m * sizeof(int64_t), hipMemcpyHostToDevice, stream));
    launchGatherRecords(recBuf.as<CandRecord>(), lIdx.as<int64_t>(), m, picked.as<CandRecord>(), stream);
    launchCandFeatures(S, picked.as<CandRecord>(), m, dets.as<uint64_t>(), dCmax, featX.as<float>(), din, stream);
    launchCandTargets(picked.as<CandRecord>(), m, dCmax, lVals.as<float>(), stream);
    const size_t old = sampT.size();
    sampX.resize((old + m) * din);
    sampT.resize(old + m);
    HIP_CHECK(hipMemcpyAsync(sampX.data() + old * din, featX.p, (size_t)m * din * sizeof(float), hipMemcpyDeviceToHost,
                             stream));
    HIP_CHECK(hipMemcpyAsync(sampT.data() + old, lVals.p, (size_t)m * sizeof(float), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    trainSamples += total;
  }
}

void Engine::learnedProbe(const double* dCmax, double eps1) {
  counter.reserve(64);
  const int64_t window = std::min<int64_t>(n, 256);
  lastLogRecords = 0;
  lastLogSources = 0;
  for (int w = 0; w < 4; w++) {
    const int64_t begin = n > window ? (int64_t)(splitmix(learnedRng) % (uint64_t)(n - window + 1)) : 0;
    HIP_CHECK(hipMemsetAsync(counter.p, 0, 8, stream));
    // capacity 0: the kernel counts the records and writes none
    launchScreenRecords(S, dets.as<uint64_t>(), begin, begin + window, dCmax, eps1, learnedLoosen, ncore, nact,
                        nullptr, counter.as<unsigned long long>(), 0, stream);
    unsigned long long found = 0;
    HIP_CHECK(hipMemcpyAsync(&found, counter.p, 8, hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    lastLogRecords += (int64_t)found;
    lastLogSources += std::min<int64_t>(window, n - begin);
  }
}

bool Engine::learnedWorthIt(double exactScreenSeconds, int64_t sources, int iter) {
  if (lastLogSources <= 0 || sources <= 0) return false;
  const double perSource = (double)lastLogRecords / (double)lastLogSources;
  const double rankSeconds = perSource * (double)sources * learnedNsPerRecord * 1e-9;
  if (rankSeconds < exactScreenSeconds) return true;
  LearnedRecord rec;
  rec.iter = iter;
  rec.fallback = true;
  rec.candidates = (int64_t)(perSource * (double)sources);
  learnedRecs.push_back(rec);
  if (std::getenv("SHCI_TIMING"))
    std::fprintf(stderr,
                 "# learned: off (ranking %.3g records at %.1f ns each, %.3g s, against %.3g s of exact screening)\n",
                 perSource * (double)sources, learnedNsPerRecord, rankSeconds, exactScreenSeconds);
  return false;
}

void Engine::learnedTrain() {
  const double t0 = now();
  const int din = mlpDin();
  const int64_t m = (int64_t)sampT.size();
  if (m < 1024) {
    mlpReady = false;
    return;
  }
  mlp.init(din, learnedHidden, 20240601, stream);
  if (learnedControl) {  // control runs keep the untrained network
    trainSeconds = now() - t0;
    mlpReady = true;
    return;
  }
  DevMem X, t;
  X.reserve((size_t)m * din * sizeof(float));
  t.reserve((size_t)m * sizeof(float));
  HIP_CHECK(hipMemcpyAsync(X.p, sampX.data(), (size_t)m * din * sizeof(float), hipMemcpyHostToDevice, stream));
  HIP_CHECK(hipMemcpyAsync(t.p, sampT.data(), (size_t)m * sizeof(float), hipMemcpyHostToDevice, stream));
  mlp.train(X.as<float>(), t.as<float>(), m, 6, 2048, 1e-3f, 7, trainFirst, trainLast, stream);
  trainSeconds = now() - t0;
  mlpReady = true;
  if (std::getenv("SHCI_TIMING"))
    std::fprintf(stderr, "# learned: trained on %lld samples (%lld logged), loss %.4g -> %.4g, %.3f s\n",
                 (long long)m, (long long)trainSamples, trainFirst, trainLast, trainSeconds);
}

// Ranked screening: exact singles; doubles from the candidate records of
// each chunk of sources, scored by the network (a prediction of
// log10(|c| |v|)); the records predicted above log10(eps1) - learnedMargin
// are evaluated exactly.  A random sample of each chunk's records measures
// the evaluated fraction and the recall of the records that pass.  Returns the new determinant count, or -1 (the
// caller then screens in full) when that threshold would evaluate more than
// learnedMaxFraction of the first chunk.
int64_t Engine::screenLearned(const double* dCmax, double eps1, int iter) {
  const double t0 = now();
  const int W = hs.W;
  const int din = mlpDin();
  LearnedRecord rec;
  rec.iter = iter;
  recBuf.reserve((size_t)kRecordCap * sizeof(CandRecord));
  recScore.reserve((size_t)kRecordCap * sizeof(float));
  featX.reserve((size_t)kScoreChunk * din * sizeof(float));
  lcounts.reserve(64);
  lIdx.reserve((size_t)kSample * sizeof(int64_t));
  lFlags.reserve((size_t)kSample * sizeof(int));
  lVals.reserve((size_t)kSample * sizeof(float));
  counter.reserve(64);
  int64_t cap = std::max<int64_t>(candCap, (int64_t)1 << 22);
  unsigned long long emitted = 0;
  for (int attempt = 0; attempt < 2; attempt++) {
    cand.reserve((size_t)cap * W * 8);
    HIP_CHECK(hipMemsetAsync(lcounts.p, 0, 64, stream));
    HIP_CHECK(hipMemsetAsync(counter.p, 0, 8, stream));
    // singles, exact
    launchScreenSingles(S, dets.as<uint64_t>(), 0, n, dCmax, eps1, ncore, nact, table.as<int32_t>(), slots,
                        cand.as<uint64_t>(), counter.as<unsigned long long>(), cap, stream);
    // doubles, ranked, by chunks of sources
    int64_t begin = 0, chunk = n;
    rec.candidates = rec.samplePass = rec.sampleHit = rec.sampleSelected = rec.sampleTotal = 0;
    bool first = true;
    DevMem recCounter;
    recCounter.reserve(64);
    while (begin < n) {
      const int64_t end = std::min(n, begin + chunk);
      HIP_CHECK(hipMemsetAsync(recCounter.p, 0, 8, stream));
      launchScreenRecords(S, dets.as<uint64_t>(), begin, end, dCmax, eps1, learnedLoosen, ncore, nact,
                          recBuf.as<CandRecord>(), recCounter.as<unsigned long long>(), kRecordCap, stream);
      unsigned long long found = 0;
      HIP_CHECK(hipMemcpyAsync(&found, recCounter.p, 8, hipMemcpyDeviceToHost, stream));
      HIP_CHECK(hipStreamSynchronize(stream));
      if ((int64_t)found > kRecordCap) {
        if (end - begin == 1) throw std::invalid_argument("one determinant has more candidate records than the buffer");
        chunk = std::max<int64_t>(1, (int64_t)((double)(end - begin) * kRecordCap / (double)found * 0.8));
        continue;
      }
      const int64_t m = (int64_t)found;
      rec.candidates += m;
      for (int64_t r0 = 0; r0 < m; r0 += kScoreChunk) {
        const int64_t mm = std::min(kScoreChunk, m - r0);
        launchCandFeatures(S, recBuf.as<CandRecord>() + r0, mm, dets.as<uint64_t>(), dCmax, featX.as<float>(), din,
                           stream);
        mlp.predict(featX.as<float>(), mm, recScore.as<float>() + r0, stream);
      }
      if (m > 0) {
        // threshold: score quantile of a random sample of the chunk's records
        const int ms = (int)std::min<int64_t>(kSample, m);
        std::vector<int64_t> idx(ms);
        for (int i = 0; i < ms; i++) idx[i] = m <= kSample ? i : (int64_t)(splitmix(learnedRng) % (uint64_t)m);
        HIP_CHECK(hipMemcpyAsync(lIdx.p, idx.data(), ms * sizeof(int64_t), hipMemcpyHostToDevice, stream));
        launchCandSample(recBuf.as<CandRecord>(), recScore.as<float>(), lIdx.as<int64_t>(), ms, dCmax, eps1, 3.4e38f,
                         lFlags.as<int>(), lVals.as<float>(), stream);
        std::vector<float> sv(ms);
        HIP_CHECK(hipMemcpyAsync(sv.data(), lVals.p, ms * sizeof(float), hipMemcpyDeviceToHost, stream));
        std::vector<int> fl(ms);
        HIP_CHECK(hipMemcpyAsync(fl.data(), lFlags.p, ms * sizeof(int), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        // threshold: the network predicts log10(|c| |v|), which passes the
        // exact test at log10(eps1); records predicted above log10(eps1) -
        // learnedMargin are evaluated
        const float tau = (float)(std::log10(eps1) - learnedMargin);
        int64_t pass = 0, hit = 0, sel = 0;
        for (int i = 0; i < ms; i++) {
          pass += fl[i] & 1;
          hit += (fl[i] & 1) && sv[i] >= tau;
          sel += sv[i] >= tau;
        }
        rec.samplePass += pass;
        rec.sampleHit += hit;
        rec.sampleSelected += sel;
        rec.sampleTotal += ms;
        // the ranking pays off only when it keeps well under all candidates
        if (first && (double)sel > learnedMaxFraction * (double)ms) {
          rec.fallback = true;
          rec.seconds = now() - t0;
          learnedRecs.push_back(rec);
          return -1;
        }
        first = false;
        launchCandExact(S, recBuf.as<CandRecord>(), m, recScore.as<float>(), tau, dets.as<uint64_t>(), dCmax, eps1,
                        table.as<int32_t>(), slots, cand.as<uint64_t>(), counter.as<unsigned long long>(), cap,
                        lcounts.as<unsigned long long>(), stream);
      }
      begin = end;
    }
    HIP_CHECK(hipMemcpyAsync(&emitted, counter.p, 8, hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    if ((int64_t)emitted <= cap) break;
    cap = (int64_t)emitted + (int64_t)emitted / 4 + 1024;  // candidates overflowed: once more, larger
    cand.release();
  }
  unsigned long long counts[2] = {0, 0};
  HIP_CHECK(hipMemcpyAsync(counts, lcounts.p, 16, hipMemcpyDeviceToHost, stream));
  HIP_CHECK(hipStreamSynchronize(stream));
  rec.evaluated = (int64_t)counts[0];
  rec.accepted = (int64_t)counts[1];
  int64_t m = (int64_t)emitted;
  radixSortKeys(cand.as<uint64_t>(), nullptr, m, W, hs.nso, true, ws, stream);
  m = uniqueKeys(cand.as<uint64_t>(), m, W, ws, stream);
  newDets.reserve((size_t)std::max<int64_t>(m, 1) * W * 8);
  if (m) HIP_CHECK(hipMemcpyAsync(newDets.p, cand.p, m * W * 8, hipMemcpyDeviceToDevice, stream));
  nNew = m;
  HIP_CHECK(hipStreamSynchronize(stream));
  rec.seconds = now() - t0;
  learnedRecs.push_back(rec);
  if (std::getenv("SHCI_TIMING"))
    std::fprintf(stderr,
                 "# learned iter %d: %lld candidate records, %lld evaluated, %lld accepted, sampled recall %.4f, "
                 "%.3f s\n",
                 iter, (long long)rec.candidates, (long long)rec.evaluated, (long long)rec.accepted,
                 rec.samplePass ? (double)rec.sampleHit / rec.samplePass : 1.0, rec.seconds);
  return nNew;
}

}  // namespace shci
