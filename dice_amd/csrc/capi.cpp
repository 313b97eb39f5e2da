// C ABI (include/shci_gpu.h) and the variational-step driver.
#include <sys/time.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>

#include "../../include/shci_gpu.h"
#include "engine.h"
#include "mlp.h"

using namespace shci;

struct shci_handle {
  Engine eng;
  std::string err;
};

namespace {

thread_local std::string g_createError;

shci_status fail(shci_handle* h, shci_status st, const std::string& msg) {
  if (h)
    h->err = msg;
  else
    g_createError = msg;
  return st;
}

template <typename F>
shci_status guarded(shci_handle* h, F&& f) {
  try {
    return f();
  } catch (const std::bad_alloc& e) {
    return fail(h, SHCI_OOM, e.what());
  } catch (const HipError& e) {
    std::string m = e.what();
    // a failed allocation also sets HIP's per-thread last error; clear it so
    // the next launch check on this thread does not report it again
    (void)hipGetLastError();
    return fail(h, m.find("out of memory") != std::string::npos ? SHCI_OOM : SHCI_HIP_ERROR, m);
  } catch (const std::invalid_argument& e) {
    return fail(h, SHCI_BAD_INPUT, e.what());
  } catch (const std::exception& e) {
    return fail(h, SHCI_HIP_ERROR, e.what());
  }
}

double wallTime() {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + 1e-6 * tv.tv_usec;
}

bool getocc(const int64_t* d, int k) { return (((uint64_t)d[k / 64]) >> (k % 64)) & 1ull; }

std::string detString(const int64_t* d, int nso) {  // SHCI/Determinants.h:459-477
  std::string s;
  for (int i = 0; i < nso / 2; i++) {
    bool a = getocc(d, 2 * i), b = getocc(d, 2 * i + 1);
    s += (!a && !b) ? "0 " : (a && !b) ? "a " : (!a && b) ? "b " : "2 ";
    if ((i + 1) % 5 == 0) s += "  ";
  }
  return s;
}

double parityDiceToAlphaBeta(const int64_t* d, int nso) {  // SHCI/Determinants.cpp:691-706
  double par = 1.0;
  int norbs = nso / 2, nalpha = 0;
  for (int i = 0; i < norbs; i++) nalpha += getocc(d, 2 * i);
  for (int i = 0; i < norbs; i++) {
    if (getocc(d, 2 * (norbs - 1 - i) + 1)) {
      int before = 0;
      for (int j = 0; j < norbs - i - 1; j++) before += getocc(d, 2 * j);
      int after = nalpha - before - (getocc(d, 2 * (norbs - 1 - i)) ? 1 : 0);
      if (after % 2 == 1) par *= -1;
    }
  }
  return par;
}

void applyMethodOptions(Engine& E, const shci_run_options& opt) {
  E.profileSigma = opt.profile_sigma != 0;
  E.davidsonMode = opt.davidson_block ? 1 : 0;
  E.autotune = opt.autotune != 0;
  E.useGraphs = opt.hip_graphs != 0;
  // half_storage > 0 / < 0: half storage always / never, 0: the engine's
  // default (SHCI_STORAGE, else chosen per form); a change of storage takes
  // effect at the next build
  if (opt.half_storage) E.storageMode = opt.half_storage > 0 ? 1 : -1;
  E.learned = opt.learned != 0;
  // 1: gated (the ranking runs only where it is predicted to pay); 2: the
  // untrained control and 3: the trained network, both always ranking
  E.learnedGate = opt.learned == 1;
  if (E.direct != (opt.direct_sigma != 0)) {  // switching between stored and direct H: rebuild
    E.hamRows = -1;
    E.permRows = -1;
    E.tiledRows = -1;
  }
  E.direct = opt.direct_sigma != 0;
  E.learnedControl = opt.learned == 2;
  // SHCI_LEARNED_MARGIN: log10 margin below eps1 of the learned form's
  // evaluation threshold (default 0.8)
  if (const char* m = std::getenv("SHCI_LEARNED_MARGIN")) E.learnedMargin = std::atof(m);
  if (!E.autotune) {  // the shapes measured fastest on the benchmark matrices
    E.sigmaCfg[0] = SigmaConfig{4, true};
    E.sigmaCfg[1] = SigmaConfig{8, false};
    E.tunedRows = 0;
  }
}

// The variational step of `Dice input.dat` (SHCI/SHCI.cpp:157-446 and
// SHCIbasics::DoVariational, SHCI/SHCIbasics.cpp:783-1378) on the device.
// Checkpoint of the variational step in this build's own portable format
// (the role of SHCIbasics::writeVariationalResult / readVariationalResult,
// SHCI/SHCIbasics.cpp:1481-1575, whose files are Boost archives):
// <prefix>/shci_var.bkp, little-endian: char[8] "SHCIVAR1", int32 W,
// nroots, iter, converged, int64 n, double E0[nroots] (no core energy),
// int64 dets[n][W] (run order), double ci[nroots][n].
struct Checkpoint {
  int W = 0, nroots = 0, iter = 0, converged = 0;
  std::vector<double> E0;
  std::vector<int64_t> dets;
  std::vector<double> ci;
  int64_t n = 0;
};

bool writeCheckpoint(const std::string& path, const Checkpoint& c) {
  const std::string tmp = path + ".tmp";
  FILE* f = fopen(tmp.c_str(), "wb");
  if (!f) return false;
  const char magic[8] = {'S', 'H', 'C', 'I', 'V', 'A', 'R', '1'};
  const int32_t hdr[4] = {c.W, c.nroots, c.iter, c.converged};
  bool ok = fwrite(magic, 1, 8, f) == 8 && fwrite(hdr, 4, 4, f) == 4 && fwrite(&c.n, 8, 1, f) == 1 &&
            fwrite(c.E0.data(), 8, c.nroots, f) == (size_t)c.nroots &&
            fwrite(c.dets.data(), 8, c.n * c.W, f) == (size_t)(c.n * c.W) &&
            fwrite(c.ci.data(), 8, c.n * c.nroots, f) == (size_t)(c.n * c.nroots);
  ok = fclose(f) == 0 && ok;
  return ok && std::rename(tmp.c_str(), path.c_str()) == 0;
}

bool readCheckpoint(const std::string& path, Checkpoint& c, std::string& err) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) {
    err = "cannot open the checkpoint " + path;
    return false;
  }
  char magic[8];
  int32_t hdr[4];
  bool ok = fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "SHCIVAR1", 8) == 0 && fread(hdr, 4, 4, f) == 4 &&
            fread(&c.n, 8, 1, f) == 1;
  if (ok) {
    c.W = hdr[0];
    c.nroots = hdr[1];
    c.iter = hdr[2];
    c.converged = hdr[3];
    ok = c.W >= 1 && c.W <= 2 && c.nroots >= 1 && c.nroots <= 64 && c.n >= 1 && c.n < ((int64_t)1 << 32);
  }
  if (ok) {
    c.E0.resize(c.nroots);
    c.dets.resize(c.n * c.W);
    c.ci.resize(c.n * c.nroots);
    ok = fread(c.E0.data(), 8, c.nroots, f) == (size_t)c.nroots &&
         fread(c.dets.data(), 8, c.n * c.W, f) == (size_t)(c.n * c.W) &&
         fread(c.ci.data(), 8, c.n * c.nroots, f) == (size_t)(c.n * c.nroots);
  }
  fclose(f);
  if (!ok) err = "the checkpoint " + path + " is not a valid variational checkpoint";
  return ok;
}

// RDM text files: spatial1RDM / spin1RDM as save1RDM writes them
// (SHCI/SHCIrdm.cpp:355-400), spatialRDM / spinRDM as saveRDM
// (SHCI/SHCIrdm.cpp:477-526); the Boost archives beside them are not written.
void save1Rdm(const std::string& prefix, int r1, int r2, int norb, const std::vector<double>& g1,
              const std::vector<double>* s1) {
  char name[64];
  snprintf(name, sizeof name, "/spatial1RDM.%d.%d.txt", r1, r2);
  if (FILE* f = fopen((prefix + name).c_str(), "w")) {
    fprintf(f, "%d\n", norb);
    for (int p = 0; p < norb; p++)
      for (int q = 0; q < norb; q++) {
        const double v = g1[(size_t)p * norb + q];
        if (std::fabs(v) > 1.e-15) fprintf(f, "%3d   %3d   %16.12g\n", p, q, v);
      }
    fclose(f);
  }
  if (!s1) return;
  const int nso = 2 * norb;
  snprintf(name, sizeof name, "/spin1RDM.%d.%d.txt", r1, r2);
  if (FILE* f = fopen((prefix + name).c_str(), "w")) {
    fprintf(f, "%d\n", nso);
    for (int p = 0; p < nso; p++)
      for (int q = 0; q < nso; q++) {
        const double v = (*s1)[(size_t)p * nso + q];
        if (std::fabs(v) > 1.e-6) fprintf(f, "%3d   %3d   %10.8g\n", p, q, v);
      }
    fclose(f);
  }
}

void save2Rdm(const std::string& prefix, int r1, int r2, int norb, const std::vector<double>& g2,
              const std::vector<double>* s2) {
  char name[64];
  snprintf(name, sizeof name, "/spatialRDM.%d.%d.txt", r1, r2);
  const size_t n = norb;
  if (FILE* f = fopen((prefix + name).c_str(), "w")) {
    fprintf(f, "%d\n", norb);
    for (size_t p = 0; p < n; p++)
      for (size_t q = 0; q < n; q++)
        for (size_t r = 0; r < n; r++)
          for (size_t t = 0; t < n; t++) {
            const double v = g2[((p * n + q) * n + r) * n + t];
            if (std::fabs(v) > 1.e-15)
              fprintf(f, "%3d   %3d   %3d   %3d   %16.12g\n", (int)p, (int)q, (int)r, (int)t, v);
          }
    fclose(f);
  }
  if (!s2) return;
  const size_t nso = 2 * n, npair = nso * (nso + 1) / 2;
  snprintf(name, sizeof name, "/spinRDM.%d.%d.txt", r1, r2);
  if (FILE* f = fopen((prefix + name).c_str(), "w")) {
    for (size_t a = 1; a < nso; a++)
      for (size_t b = 0; b < a; b++)
        for (size_t c = 1; c < nso; c++)
          for (size_t d = 0; d < c; d++) {
            const double v = (*s2)[(a * (a + 1) / 2 + b) * npair + c * (c + 1) / 2 + d];
            if (std::fabs(v) > 1.e-6) fprintf(f, "%3d %3d %3d %3d %16.14g\n", (int)a, (int)b, (int)c, (int)d, v);
          }
    fclose(f);
  }
}

// ComputeEnergyFromSpatialRDM (SHCI/SHCIrdm.cpp:1188-1245): the one-body part
// from the 1-RDM that the 2-RDM implies, gamma(p,q) = sum_r G(p r, q r) /
// (nelec - 1)
double energyFromSpatialRdm(const HostSystem& hs, const std::vector<double>& g2, FILE* log) {
  const int n = hs.norb;
  double onebody = 0.0, twobody = 0.0;
  for (int p = 0; p < n; p++)
    for (int q = 0; q < n; q++) {
      double g = 0.0;
      for (int r = 0; r < n; r++) g += g2[(((size_t)p * n + r) * n + q) * n + r] / (1. * hs.nelec - 1.);
      onebody += hs.h1[(size_t)p * n + q] * g;
    }
  for (int p = 0; p < n; p++)
    for (int q = 0; q < n; q++)
      for (int r = 0; r < n; r++)
        for (int t = 0; t < n; t++)
          twobody += 0.5 * g2[(((size_t)p * n + q) * n + r) * n + t] * hs.i2(2 * p, 2 * r, 2 * q, 2 * t);
  const double e = hs.coreE + onebody + twobody;
  if (log) {
    fprintf(log, "E(one-body) from 2RDM: %18.10f\n", onebody);
    fprintf(log, "E(two-body) from 2RDM: %18.10f\n", twobody);
    fprintf(log, "E from 2RDM:           %18.10f\n", e);
  }
  return e;
}

// the RDM step at the end of the variational iterations
// (SHCI/SHCIbasics.cpp:1236-1340)
void rdmStep(Engine& E, const HostSchedule& schd, FILE* log) {
  const std::string& prefix = schd.prefix[0];
  const int norb = E.hs.norb, nroots = E.nroots;
  const double* ci = E.ci.as<double>();
  std::vector<double> g1, g2, s1, s2;
  if (schd.doOneRDM) {
    if (log) fprintf(log, "\nCalculating 1-RDM\n");
    for (int i = 0; i < nroots; i++) {
      const double* c = ci + (int64_t)i * E.n;
      E.rdm(c, c, g1, g2, schd.doSpinOneRDM ? &s1 : nullptr, nullptr);
      save1Rdm(prefix, i, i, norb, g1, schd.doSpinOneRDM ? &s1 : nullptr);
    }
  }
  if (!(schd.doRDM || schd.doTransitionRDM)) return;
  if (log) fprintf(log, "\nCalculating 2-RDM\n");
  std::vector<double>* sp2 = schd.doSpinRDM ? &s2 : nullptr;
  for (int i = 0; i < nroots; i++) {
    const double* ciI = ci + (int64_t)i * E.n;
    if (E.direct) {  // the reference's direct-mode branch: the 2-RDM only
      E.rdm(ciI, ciI, g1, g2, nullptr, sp2);
      energyFromSpatialRdm(E.hs, g2, log);
      save2Rdm(prefix, i, i, norb, g2, sp2);
      continue;
    }
    if (schd.doTransitionRDM)
      for (int j = 0; j < i; j++) {
        E.rdm(ciI, ci + (int64_t)j * E.n, g1, g2, schd.doSpinOneRDM ? &s1 : nullptr, sp2);
        save1Rdm(prefix, i, j, norb, g1, schd.doSpinOneRDM ? &s1 : nullptr);
        energyFromSpatialRdm(E.hs, g2, log);
        save2Rdm(prefix, i, j, norb, g2, sp2);
      }
    E.rdm(ciI, ciI, g1, g2, schd.doSpinOneRDM ? &s1 : nullptr, sp2);
    energyFromSpatialRdm(E.hs, g2, log);
    save2Rdm(prefix, i, i, norb, g2, sp2);
    save1Rdm(prefix, i, i, norb, g1, schd.doSpinOneRDM ? &s1 : nullptr);
  }
  if (log) fflush(log);
}

// The variational iterations (SHCIbasics::DoVariational,
// SHCI/SHCIbasics.cpp:783-1378) from the held determinants (diagonal built)
// with the coefficients already in E.ci (n x nroots, column-major): the
// reference's table on log, per-iteration records, checkpoints and restart as
// schd asks, the final tight Davidson and the RDM step; E.energies = the
// energies with the core energy added.  fromConverged: a restart found a
// converged checkpoint (the iterations were skipped).
shci_status runIterations(shci_handle* h, HostSchedule& schd, FILE* log, double startofCalc, double e0init,
                          bool& fromConverged) {
  Engine& E = h->eng;
  const HostSystem& hs = E.hs;
  const int W = hs.W;
  const int nroots = schd.nroots;
  if (log) {
    fprintf(log, "\n\n**************************************************************\n");
    fprintf(log, "VARIATIONAL STEP  \n");
    fprintf(log, "**************************************************************\n");
  }
  const double coreEbkp = hs.coreE;
  std::vector<double> E0(nroots, e0init);
  E.iters.clear();
  if (log) fprintf(log, "Iter Root       Eps1   #Var. Det.               Energy   #Davidson     Time(s)\n");
  DevMem X0;
  const std::string ckptPath = schd.prefix[0] + "/shci_var.bkp";
  auto saveCheckpoint = [&](int iter) {  // SHCI/SHCIbasics.cpp:1224-1227,1345-1352 (always "converged")
    Checkpoint c;
    c.W = W;
    c.nroots = nroots;
    c.iter = iter;
    c.converged = 1;
    c.n = E.n;
    c.E0 = E0;
    c.dets.resize(c.n * W);
    c.ci.resize(c.n * nroots);
    HIP_CHECK(hipMemcpy(c.dets.data(), E.dets.p, c.n * W * 8, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(c.ci.data(), E.ci.p, c.n * nroots * 8, hipMemcpyDeviceToHost));
    if (!writeCheckpoint(ckptPath, c) && log) fprintf(log, "could not write the checkpoint %s\n", ckptPath.c_str());
  };
  // restart from the checkpoint (SHCI/SHCIbasics.cpp:871-941)
  int iterstart = 0;
  fromConverged = false;
  if (schd.restart || schd.fullrestart) {
    Checkpoint c;
    std::string e;
    if (!readCheckpoint(ckptPath, c, e) || c.W != W || c.nroots != nroots) {
      if (e.empty()) e = "the checkpoint " + ckptPath + " does not match this system (words or roots)";
      if (log) fprintf(log, "%s\n", e.c_str());
      return fail(h, SHCI_BAD_INPUT, e);
    }
    E.setDets(c.dets.data(), c.n);
    E.buildHamiltonian();
    E.ci.reserve((size_t)c.n * nroots * 8);
    HIP_CHECK(hipMemcpy(E.ci.p, c.ci.data(), c.n * nroots * 8, hipMemcpyHostToDevice));
    E0 = c.E0;
    const int last = (int)schd.epsilon1.size() - 1;
    iterstart = c.iter;
    for (int i = 0; i < nroots; i++)
      if (log)
        fprintf(log, "%4d %4d  %10.2e  %10lld -   %18.10f  %10.2f\n", iterstart, i,
                schd.epsilon1[std::min(iterstart, last)], (long long)c.n, E0[i] + coreEbkp, wallTime() - startofCalc);
    iterstart = schd.fullrestart ? 0 : iterstart + 1;
    // (a full restart re-runs the schedule: the reference's test would read
    // epsilon1[-1] there)
    if (!schd.fullrestart && c.converged &&
        (iterstart - 1 > last || schd.epsilon1[iterstart - 1] <= schd.epsilon1[last])) {
      if (log) fprintf(log, "# restarting from a converged calculation, moving to perturbative part.!!\n");
      fromConverged = true;
      if (E.trev) E.unpackTrev(schd.Trev);  // SHCI/SHCIbasics.cpp:939
    }
  }
  for (int iter = iterstart; !fromConverged && iter < (int)schd.epsilon1.size(); iter++) {
    const double eps1 = schd.epsilon1[iter];
    const double t0 = wallTime();
    const int64_t nold = E.n;
    E.cmax.reserve((size_t)std::max<int64_t>(nold, 1) * 8);
    launchCmax(E.ci.as<double>(), nroots, nold, nold, E.cmax.as<double>(), E.stream);
    int64_t nnew = -1;
    if (E.learned && E.mlpReady && iter >= E.learnedLogIters) {
      nnew = E.screenLearned(E.cmax.as<double>(), eps1, iter);
      // once the ranking would not save evaluations, later iterations
      // screen in full without trying it again
      if (nnew < 0) E.mlpReady = false;
    }
    const double tScreenStart = wallTime();
    if (nnew < 0) nnew = E.screen(E.cmax.as<double>(), eps1);
    const double tExactScreen = wallTime() - tScreenStart;
    // gated form: before the last logging iteration logs anything (the
    // first starts from the reference determinant alone), rank only when
    // ranking this iteration's candidate records (counted over a few
    // windows of sources) is predicted to cost less than its exact screen
    if (E.learned && E.learnedGate && iter == E.learnedLogIters - 1) {
      E.learnedProbe(E.cmax.as<double>(), eps1);
      if (!E.learnedWorthIt(tExactScreen, nold, iter)) E.learned = false;
    }
    if (E.learned && iter < E.learnedLogIters) {
      E.learnedLog(E.cmax.as<double>(), eps1);
      if (iter == E.learnedLogIters - 1) E.learnedTrain();
    }
    if (std::getenv("SHCI_TIMING"))
      fprintf(stderr, "# iter %d: screening selected %lld new determinants\n", iter, (long long)nnew);
    E.appendNew();
    const int64_t N = E.n;
    if (E.direct && log) fprintf(log, "New size of determinant space %8lld\n", (long long)N);
    // X0 = [ci; 0]
    X0.reserve((size_t)N * nroots * 8);
    HIP_CHECK(hipMemsetAsync(X0.p, 0, (size_t)N * nroots * 8, E.stream));
    for (int r = 0; r < nroots; r++)
      HIP_CHECK(hipMemcpyAsync(X0.as<double>() + (int64_t)r * N, E.ci.as<double>() + (int64_t)r * nold, nold * 8,
                               hipMemcpyDeviceToDevice, E.stream));
    (void)nnew;
    HIP_CHECK(hipStreamSynchronize(E.stream));
    const double t1 = wallTime();
    E.buildHamiltonian();
    const double t2 = wallTime();
    const double prevE0 = iter == 0 ? -10.0 : E0[0];
    int numIter = 0;
    const int looseCopies = E.direct ? nroots + 2 : nroots + 4;
    E0 = E.davidsonMode == 1 ? E.davidsonBlock(X0.as<double>(), nroots, schd.davidsonTolLoose, numIter, false)
                             : E.davidson(X0.as<double>(), nroots, looseCopies, schd.davidsonTolLoose, numIter, false);
    if ((int)E0.size() != nroots) {
      if (log) fprintf(log, "%s\n", E.err.c_str());
      return fail(h, SHCI_NO_CONVERGENCE, E.err);
    }
    const double t3 = wallTime();
    const double t = t3 - startofCalc;
    if (log) {
      fprintf(log, "%4d %4d  %10.2e  %10lld", iter, 0, eps1, (long long)N);
      fprintf(log, "   %18.10f  %9d  %10.2f\n", E0[0] + coreEbkp, numIter, t);
      for (int i = 1; i < nroots; i++)
        fprintf(log, "%4d %4d  %10.2e  %10lld   %18.10f  %9d  %10.2f\n", iter, i, eps1, (long long)N,
                E0[i] + coreEbkp, numIter, t);
      if (nroots > 1) fprintf(log, "\n");
      fflush(log);
    }
    IterRecord rec{iter, eps1, N, {}, numIter, t};
    rec.tScreen = t1 - t0;
    rec.tHam = t2 - t1;
    rec.tDavidson = t3 - t2;
    rec.tIter = t3 - t0;
    rec.nnzLower = E.lowerNnz();
    if (std::getenv("SHCI_TIMING"))
      fprintf(stderr, "# iter %d: n=%lld nnz_lower=%lld screen %.3fs hamiltonian %.3fs davidson %.3fs (%d it)\n", iter,
              (long long)N, (long long)rec.nnzLower, rec.tScreen, rec.tHam, rec.tDavidson, numIter);
    for (double e : E0) rec.e0.push_back(e + coreEbkp);
    E.iters.push_back(rec);
    // ci = X0
    E.ci.reserve((size_t)N * nroots * 8);
    HIP_CHECK(hipMemcpyAsync(E.ci.p, X0.p, (size_t)N * nroots * 8, hipMemcpyDeviceToDevice, E.stream));
    if (std::fabs(E0[0] - prevE0) < schd.dE || iter == (int)schd.epsilon1.size() - 1) {
      if (log) fprintf(log, "Performing final tight davidson with tol: %g\n", schd.davidsonTol);
      const double tf0 = wallTime();
      if (E.direct) E.davidsonPrint = log;  // davidsonDirect(..., print = true)
      E0 = E.davidsonMode == 1 ? E.davidsonBlock(E.ci.as<double>(), nroots, schd.davidsonTol, numIter, true)
                               : E.davidson(E.ci.as<double>(), nroots, nroots + 4, schd.davidsonTol, numIter, true);
      E.davidsonPrint = nullptr;
      E.tFinalDavidson = wallTime() - tf0;
      if ((int)E0.size() != nroots) return fail(h, SHCI_NO_CONVERGENCE, E.err);
      if (log) fprintf(log, "Exiting variational iterations\n");
      if (schd.io) saveCheckpoint(iter);
      if (E.trev) E.unpackTrev(schd.Trev);  // SHCI/SHCIbasics.cpp:1229-1234
      rdmStep(E, schd, log);
      break;
    }
    if (schd.io) saveCheckpoint(iter);
  }
  if (log && !fromConverged) {
    fprintf(log, "\nVariational calculation result\n");
    fprintf(log, "%4s %18s  %10s\n", "Root", "Energy", "Time(s)");
  }
  for (size_t i = 0; i < E0.size(); i++) {
    E0[i] += coreEbkp;
    if (log && !fromConverged) fprintf(log, "%4d  %18.10f  %10.2f\n", (int)i, E0[i], wallTime() - startofCalc);
  }
  if (log && !fromConverged) fprintf(log, "\n");
  E.energies = E0;
  return SHCI_OK;
}

shci_status runVariational(shci_handle* h, const char* inputPath, FILE* log, const shci_run_options& opt) {
  Engine& E = h->eng;
  const int precision = opt.precision;
  applyMethodOptions(E, opt);
  E.learnedRecs.clear();
  E.sampX.clear();
  E.sampT.clear();
  E.trainSamples = 0;
  E.mlpReady = false;
  E.resetSigmaProfile();
  const double startofCalc = wallTime();
  HostSchedule schd;
  std::string err, echo;
  bool ok = readInputFile(inputPath, schd, err, &echo);
  if (log) {
    fprintf(log, "\n\n**************************************************************\n");
    fprintf(log, "INPUT FILE\n");
    fprintf(log, "**************************************************************\n%s", echo.c_str());
  }
  if (!ok) {
    if (log) fprintf(log, "%s\n", err.c_str());
    return fail(h, SHCI_BAD_INPUT, err);
  }
  if (schd.davidsonType == 2 || schd.doSOC || schd.searchLowest || schd.doLCC || schd.doResponse ||
      schd.doHigherRDM) {
    std::string m =
        "keyword outside the variational path implemented here (diskdavidson, dosoc, "
        "searchForLowestEnergyDet, dolcc, doresponse, DoThreeRDM, DoFourRDM)";
    if (log) fprintf(log, "%s\n", m.c_str());
    return fail(h, SHCI_UNSUPPORTED, m);
  }
  if (schd.doRDM && !schd.rdmUnrelaxed && !schd.stochastic && schd.nblocks == 1) {
    // SHCI/SHCI.cpp:759-806 would add the PT correction to the 2-RDM
    std::string m = "DoRDM with deterministic PT asks for the relaxed PT 2-RDM, which is not implemented here "
                    "(relaxedRDM keeps the variational RDMs)";
    if (log) fprintf(log, "%s\n", m.c_str());
    return fail(h, SHCI_UNSUPPORTED, m);
  }
  if (schd.occupied[0].size() % 2 != 0 && schd.Trev != 0) {  // SHCI/SHCI.cpp:189-195
    if (log) fprintf(log, "\nCannot use time reversal symmetry for odd electron system.\n");
    schd.Trev = 0;
  }
  if (schd.Trev != 0 && (opt.davidson_block || opt.learned)) {
    std::string m = "Treversal runs with the reference's Davidson and screening only (no block Davidson or "
                    "learned screening)";
    if (log) fprintf(log, "%s\n", m.c_str());
    return fail(h, SHCI_UNSUPPORTED, m);
  }
  E.rng.reseed((unsigned)(schd.randomSeed + 0));  // srand(seed + rank), SHCI/SHCI.cpp:199
  E.load(schd.integralFile);
  E.trev = schd.Trev;
  E.S.trev = schd.Trev != 0;
  E.precision = precision;
  // directdavidson: no stored Hamiltonian (HmultDirect) and davidsonDirect
  // (SHCI/SHCIbasics.cpp:1092,1167-1172,1206-1208)
  E.direct = schd.davidsonType == 1 || opt.direct_sigma != 0;
  E.davidsonDirectRules = schd.davidsonType == 1;
  E.davidsonPrint = nullptr;
  E.davidsonPrintStart = startofCalc;
  E.ncore = schd.ncore;
  E.nact = schd.nact;
  E.restrictions = schd.restrictions;
  E.restrictionsPT = schd.restrictionsPT;
  E.uploadRestrictions();
  E.ptEnergies.clear();
  const HostSystem& hs = E.hs;
  const int W = hs.W, nso = hs.nso;
  if ((int)schd.occupied[0].size() != hs.nelec) {
    std::string m = "The number of electrons given in the FCIDUMP should be equal to the nocc given in the shci input file.";
    if (log) fprintf(log, "%s\n", m.c_str());
    return fail(h, SHCI_BAD_INPUT, m);
  }
  if (log) {
    fprintf(log, "\n\n**************************************************************\n");
    fprintf(log, "SELECTING REFERENCE DETERMINANT(S)\n");
    fprintf(log, "**************************************************************\n");
  }
  // reference determinant(s): SHCI/SHCI.cpp:286-316
  const int64_t nref = (int64_t)schd.occupied.size();
  std::vector<int64_t> hdets(nref * W, 0);
  for (int64_t d = 0; d < nref; d++) {
    for (int o : schd.occupied[d]) {
      if (o < 0 || o >= nso || getocc(&hdets[d * W], o)) {
        std::string m = "orbital " + std::to_string(o) + " appears twice in input determinant number " +
                        std::to_string(d);
        if (log) fprintf(log, "%s\n", m.c_str());
        return fail(h, SHCI_BAD_INPUT, m);
      }
      hdets[d * W + o / 64] |= (int64_t)(1ull << (o % 64));
    }
    if (schd.Trev != 0) {  // SHCI/SHCI.cpp:302
      if (W == 1)
        makeStandardForm<1>(reinterpret_cast<uint64_t*>(&hdets[d * W]));
      else
        makeStandardForm<2>(reinterpret_cast<uint64_t*>(&hdets[d * W]));
    }
    for (int64_t i = 0; i < d; i++)
      if (std::equal(&hdets[d * W], &hdets[d * W] + W, &hdets[i * W])) {
        std::string m = "Determinant appears twice in the input determinant list.";
        if (log) fprintf(log, "%s\n", m.c_str());
        return fail(h, SHCI_BAD_INPUT, m);
      }
  }
  E.setDets(hdets.data(), nref);
  E.buildHamiltonian();  // diagonal of the references (and their H)
  std::vector<double> eref(nref);
  HIP_CHECK(hipMemcpy(eref.data(), E.diag.p, nref * 8, hipMemcpyDeviceToHost));
  int lowest = 0;
  double lowestE = 1.e12;
  for (int64_t d = 0; d < nref; d++) {
    double e = eref[d] + hs.coreE;
    if (log) fprintf(log, "%s Given Ref. Energy:    %18.10f\n", detString(&hdets[d * W], nso).c_str(), e);
    if (e < lowestE) {
      lowestE = e;
      lowest = (int)d;
    }
  }
  const int nroots = schd.nroots;
  E.nroots = nroots;
  // ci: N x nroots column-major on the device (ld = N)
  std::vector<double> hci(nref * nroots, 0.0);
  {
    double s = 0.0;
    for (int64_t j = 0; j < nref; j++) hci[j] = 1.0;
    hci[lowest] += (double)nref;
    for (int64_t j = 0; j < nref; j++) s += hci[j] * hci[j];
    s = std::sqrt(s);
    for (int64_t j = 0; j < nref; j++) hci[j] /= s;
  }
  E.ci.reserve(hci.size() * 8);
  HIP_CHECK(hipMemcpy(E.ci.p, hci.data(), hci.size() * 8, hipMemcpyHostToDevice));
  bool fromConverged = false;
  {
    const shci_status st = runIterations(h, schd, log, startofCalc, std::fabs(eref[0]), fromConverged);
    if (st != SHCI_OK) return st;
  }
  std::vector<double> E0 = E.energies;
  const double coreEbkp = hs.coreE;
  {  // SHCI/SHCI.cpp:437-446
    std::string efile = schd.prefix[0] + "/shci.e";
    FILE* f = fopen(efile.c_str(), "wb");
    if (f) {
      for (double e : E0) fwrite(&e, 1, sizeof(double), f);
      fclose(f);
    }
  }
  std::vector<int64_t> dd;
  std::vector<double> cc;
  const int64_t N = E.n;
  if (log || schd.writeBestDeterminants > 0) {
    dd.resize(N * W);
    cc.resize(N * nroots);
    HIP_CHECK(hipMemcpy(dd.data(), E.dets.p, N * W * 8, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(cc.data(), E.ci.p, N * nroots * 8, hipMemcpyDeviceToHost));
  }
  if (log && schd.printAllDeterminants) {  // SHCI/SHCI.cpp:449-462
    fprintf(log, "Printing all determinants\n");
    fprintf(log, "%4s %10s  Determinant string\n", "Det", "weight");
    for (int root = 0; root < nroots; root++) {
      fprintf(log, "State :%d\n", root);
      for (int64_t i = 0; i < N; i++)
        fprintf(log, "%18.8e  %s\n", cc[root * N + i] * parityDiceToAlphaBeta(&dd[i * W], nso),
                detString(&dd[i * W], nso).c_str());
    }
  }
  if (schd.writeBestDeterminants > 0) {  // SHCI/SHCI.cpp:464-505: dets.bin in the working directory
    const int num = (int)std::min<int64_t>(schd.writeBestDeterminants, N);
    const int nspatorbs = nso / 2;
    for (int root = 0; root < nroots; root++) {
      const std::string fname = root == 0 ? "dets.bin" : "dets_" + std::to_string(root) + ".bin";
      FILE* f = fopen(fname.c_str(), "wb");
      if (!f) continue;
      fwrite(&num, sizeof(int), 1, f);
      fwrite(&nspatorbs, sizeof(int), 1, f);
      // |c| descending; ties (unordered in the reference) in determinant order
      std::vector<int64_t> idx(N);
      for (int64_t k = 0; k < N; k++) idx[k] = k;
      const double* c = cc.data() + root * N;
      std::stable_sort(idx.begin(), idx.end(), [c](int64_t a, int64_t b) { return std::fabs(c[a]) > std::fabs(c[b]); });
      for (int i = 0; i < num; i++) {
        const int64_t m = idx[i];
        const double w = parityDiceToAlphaBeta(&dd[m * W], nso) * c[m];
        fwrite(&w, sizeof(double), 1, f);
        for (int k = 0; k < nspatorbs; k++) {
          const bool a = (dd[m * W + (2 * k) / 64] >> ((2 * k) % 64)) & 1;
          const bool b = (dd[m * W + (2 * k + 1) / 64] >> ((2 * k + 1) % 64)) & 1;
          const char o = a ? (b ? '2' : 'a') : (b ? 'b' : '0');
          fwrite(&o, 1, 1, f);
        }
      }
      fclose(f);
    }
  }
  if (log) {  // SHCI/SHCI.cpp:526-551
    fprintf(log, "Printing most important determinants\n");
    fprintf(log, "%4s %10s  Determinant string\n", "Det", "weight");
    for (int root = 0; root < nroots; root++) {
      fprintf(log, "State : %3d\n", root);
      std::vector<double> prev(cc.begin() + root * N, cc.begin() + (root + 1) * N);
      int num = std::max(6, schd.printBestDeterminants);
      for (int i = 0; i < std::min<int64_t>(num, N); i++) {
        int64_t m = 0;
        for (int64_t k = 1; k < N; k++)
          if (std::fabs((float)prev[m]) < std::fabs((float)prev[k])) m = k;
        double par = parityDiceToAlphaBeta(&dd[m * W], nso);
        fprintf(log, "%4d %18.10f  %s\n", i, prev[m] * par, detString(&dd[m * W], nso).c_str());
        prev[m] = 0.0;
      }
    }
    fflush(log);
  }
  // deterministic PT step (SHCI/SHCI.cpp:706-721): shci.e becomes E_var + E_PT
  if (!schd.stochastic && schd.nblocks == 1) {
    if (log) {  // log_pt, SHCI/SHCI.cpp:140-156
      fprintf(log, "\n\n**************************************************************\n");
      fprintf(log, "PERTURBATION THEORY STEP  \n");
      fprintf(log, "**************************************************************\n");
    }
    std::vector<double> ept(E0.size());
    for (size_t root = 0; root < E0.size(); root++) {
      const double pt = E.ptDeterministic(E.ci.as<double>() + (int64_t)root * E.n, E0[root] - coreEbkp,
                                          schd.epsilon2, nullptr, nullptr);
      ept[root] = E0[root] + pt;
      if (log) {  // SHCI/SHCIbasics.cpp:664-668
        fprintf(log, "Deterministic PT calculation converged\n");
        fprintf(log, "PTEnergy: %18.10f\n", ept[root]);
        fprintf(log, "Time(s):  %10.2f\n", wallTime() - startofCalc);
        fflush(log);
      }
    }
    E.ptEnergies = ept;
    std::string efile = schd.prefix[0] + "/shci.e";
    FILE* f = fopen(efile.c_str(), "wb");
    if (f) {
      for (double e : ept) fwrite(&e, 1, sizeof(double), f);
      fclose(f);
    }
  } else if (schd.SampleN != -1 && schd.nPTiter != 0) {  // semistochastic PT, SHCI/SHCI.cpp:722-741
    if (log) {
      fprintf(log, "\n\n**************************************************************\n");
      fprintf(log, "PERTURBATION THEORY STEP  \n");
      fprintf(log, "**************************************************************\n");
      if (schd.doRDM) fprintf(log, "(We cannot perform PT RDM with stochastic PT. Disabling RDM.)\n\n");
    }
    std::vector<double> ept(E0.size());
    for (size_t root = 0; root < E0.size(); root++)
      ept[root] = E0[root] + E.ptSemistochastic(E.ci.as<double>() + (int64_t)root * E.n, E0[root] - coreEbkp,
                                                schd.epsilon2, schd.epsilon2Large, schd.SampleN, schd.nPTiter,
                                                schd.targetError, coreEbkp, (int)root, log, startofCalc, nullptr,
                                                nullptr, nullptr);
    E.ptEnergies = ept;
    std::string efile = schd.prefix[0] + "/shci.e";
    FILE* f = fopen(efile.c_str(), "wb");
    if (f) {
      for (double e : ept) fwrite(&e, 1, sizeof(double), f);
      fclose(f);
    }
  }
  return SHCI_OK;
}

}  // namespace

extern "C" {

shci_status shci_create(const char* fcidump_path, shci_handle** out) {
  if (!out || !fcidump_path) return fail(nullptr, SHCI_BAD_INPUT, "null argument");
  *out = nullptr;
  std::unique_ptr<shci_handle> h;
  try {
    h.reset(new shci_handle());
  } catch (const std::exception& e) {
    return fail(nullptr, SHCI_HIP_ERROR, e.what());
  }
  shci_status st = guarded(h.get(), [&] {
    h->eng.load(fcidump_path);
    return SHCI_OK;
  });
  if (st != SHCI_OK) {
    g_createError = h->err;
    return st;
  }
  *out = h.release();
  return SHCI_OK;
}

shci_status shci_create_from_integrals(int norb, int nelec, const double* h1, const double* eri8, double core_energy,
                                      const int* orbsym, shci_handle** out) {
  if (!out) return fail(nullptr, SHCI_BAD_INPUT, "null argument");
  *out = nullptr;
  HostSystem sys;
  std::string err;
  if (!integralsFromArrays(norb, nelec, h1, eri8, core_energy, orbsym, sys, err))
    return fail(nullptr, SHCI_BAD_INPUT, err);
  std::unique_ptr<shci_handle> h;
  try {
    h.reset(new shci_handle());
  } catch (const std::exception& e) {
    return fail(nullptr, SHCI_HIP_ERROR, e.what());
  }
  shci_status st = guarded(h.get(), [&] {
    h->eng.loadSystem(std::move(sys));
    return SHCI_OK;
  });
  if (st != SHCI_OK) {
    g_createError = h->err;
    return st;
  }
  *out = h.release();
  return SHCI_OK;
}

shci_status shci_do_variational(shci_handle* h, const shci_var_params* p, const int64_t* dets, int64_t ndets,
                                const double* ci, double* energies) {
  if (!h) return fail(nullptr, SHCI_BAD_INPUT, "null handle");
  return guarded(h, [&] {
    if (!p || !dets || !energies) throw std::invalid_argument("null argument");
    if (p->nroots < 1 || p->n_iter < 1 || !p->eps1) throw std::invalid_argument("need nroots >= 1 and an eps1 list");
    if (ndets < 1) throw std::invalid_argument("need at least one determinant");
    if (!(p->davidson_tol > 0) || !(p->davidson_tol_loose > 0) || !(p->dE > 0))
      throw std::invalid_argument("tolerances must be positive");
    Engine& E = h->eng;
    if (E.trev) throw std::invalid_argument("time reversal runs through shci_run (Treversal keyword)");
    HostSchedule schd;
    schd.nroots = p->nroots;
    schd.epsilon1.assign(p->eps1, p->eps1 + p->n_iter);
    schd.davidsonTol = p->davidson_tol;
    schd.davidsonTolLoose = p->davidson_tol_loose;
    schd.dE = p->dE;
    schd.randomSeed = p->seed;
    schd.io = false;
    schd.prefix.assign(1, ".");
    E.rng.reseed(p->seed);
    E.ncore = p->ncore;
    E.nact = p->nact;
    E.learnedRecs.clear();
    E.sampX.clear();
    E.sampT.clear();
    E.trainSamples = 0;
    E.mlpReady = false;
    E.resetSigmaProfile();
    E.ptEnergies.clear();
    E.davidsonDirectRules = false;
    E.davidsonPrint = nullptr;
    E.setDets(dets, ndets);
    E.buildHamiltonian();
    const int nroots = p->nroots;
    E.nroots = nroots;
    std::vector<double> hci((size_t)ndets * nroots, 0.0);
    if (ci) {
      std::copy(ci, ci + hci.size(), hci.begin());
    } else {  // SHCI/SHCI.cpp:395-399: every determinant 1, the lowest-energy one + ndets, normalised
      std::vector<double> d(ndets);
      HIP_CHECK(hipMemcpy(d.data(), E.diag.p, (size_t)ndets * 8, hipMemcpyDeviceToHost));
      const int64_t lowest = std::min_element(d.begin(), d.end()) - d.begin();
      for (int64_t j = 0; j < ndets; j++) hci[j] = 1.0;
      hci[lowest] += (double)ndets;
      double s = 0.0;
      for (int64_t j = 0; j < ndets; j++) s += hci[j] * hci[j];
      s = std::sqrt(s);
      for (int64_t j = 0; j < ndets; j++) hci[j] /= s;
    }
    E.ci.reserve(hci.size() * 8);
    HIP_CHECK(hipMemcpy(E.ci.p, hci.data(), hci.size() * 8, hipMemcpyHostToDevice));
    double d0 = 0.0;
    HIP_CHECK(hipMemcpy(&d0, E.diag.p, 8, hipMemcpyDeviceToHost));
    bool fromConverged = false;
    const shci_status st = runIterations(h, schd, nullptr, wallTime(), std::fabs(d0), fromConverged);
    if (st != SHCI_OK) return st;
    for (int r = 0; r < nroots; r++) energies[r] = E.energies[r];
    return SHCI_OK;
  });
}

shci_status shci_destroy(shci_handle* h) {
  delete h;
  return SHCI_OK;
}

const char* shci_last_error(const shci_handle* h) { return h ? h->err.c_str() : g_createError.c_str(); }
int shci_norb(const shci_handle* h) { return h->eng.hs.norb; }
int shci_nelec(const shci_handle* h) { return h->eng.hs.nelec; }
int shci_det_words(const shci_handle* h) { return h->eng.hs.W; }
double shci_core_energy(const shci_handle* h) { return h->eng.hs.coreE; }

shci_status shci_set_precision(shci_handle* h, int precision) {
  if (precision != SHCI_PREC_FP64 && precision != SHCI_PREC_FP32)
    return fail(h, SHCI_BAD_INPUT, "unknown precision");
  h->eng.precision = precision;
  h->eng.hamRows = -1;
  h->eng.permRows = -1;
  h->eng.tiledRows = -1;
  return SHCI_OK;
}

shci_status shci_set_options(shci_handle* h, const shci_run_options* options) {
  if (!options) return fail(h, SHCI_BAD_INPUT, "options is NULL");
  if (options->precision != h->eng.precision) {
    shci_status st = shci_set_precision(h, options->precision);
    if (st != SHCI_OK) return st;
  }
  applyMethodOptions(h->eng, *options);
  return SHCI_OK;
}

shci_status shci_set_active_space(shci_handle* h, int ncore, int nact) {
  h->eng.ncore = ncore;
  h->eng.nact = nact;
  return SHCI_OK;
}

shci_status shci_add_restriction(shci_handle* h, int min_elec, int max_elec, const int* spatial_orbs, int n) {
  return guarded(h, [&] {
    Restriction r;
    r.minElec = min_elec;
    r.maxElec = max_elec;
    for (int i = 0; i < n; i++) {
      r.orbs.push_back(2 * spatial_orbs[i]);
      r.orbs.push_back(2 * spatial_orbs[i] + 1);
    }
    // as the input's "restrict": the variational and the PT space
    h->eng.restrictions.push_back(r);
    h->eng.restrictionsPT.push_back(r);
    h->eng.uploadRestrictions();
    return SHCI_OK;
  });
}

shci_status shci_det_energy(shci_handle* h, const int64_t* dets, int64_t n, double* energies) {
  return guarded(h, [&] {
    if (n <= 0) return SHCI_OK;
    Engine& E = h->eng;
    const int W = E.hs.W;
    DevMem d, o;
    d.reserve(n * W * 8);
    o.reserve(n * 8);
    HIP_CHECK(hipMemcpyAsync(d.p, dets, n * W * 8, hipMemcpyHostToDevice, E.stream));
    launchEnergy(E.S, d.as<uint64_t>(), n, o.as<double>(), E.stream);
    HIP_CHECK(hipMemcpyAsync(energies, o.p, n * 8, hipMemcpyDeviceToHost, E.stream));
    HIP_CHECK(hipStreamSynchronize(E.stream));
    return SHCI_OK;
  });
}

shci_status shci_hij(shci_handle* h, const int64_t* bra, const int64_t* ket, int64_t n, double* out) {
  return guarded(h, [&] {
    if (n <= 0) return SHCI_OK;
    Engine& E = h->eng;
    const int W = E.hs.W;
    DevMem b, k, o;
    b.reserve(n * W * 8);
    k.reserve(n * W * 8);
    o.reserve(n * 8);
    HIP_CHECK(hipMemcpyAsync(b.p, bra, n * W * 8, hipMemcpyHostToDevice, E.stream));
    HIP_CHECK(hipMemcpyAsync(k.p, ket, n * W * 8, hipMemcpyHostToDevice, E.stream));
    launchHijPairs(E.S, b.as<uint64_t>(), k.as<uint64_t>(), n, o.as<double>(), E.stream);
    HIP_CHECK(hipMemcpyAsync(out, o.p, n * 8, hipMemcpyDeviceToHost, E.stream));
    HIP_CHECK(hipStreamSynchronize(E.stream));
    return SHCI_OK;
  });
}

shci_status shci_set_dets(shci_handle* h, const int64_t* dets, int64_t n) {
  return guarded(h, [&] {
    if (n < 0) throw std::invalid_argument("negative determinant count");
    h->eng.setDets(dets, n);
    return SHCI_OK;
  });
}

shci_status shci_num_dets(const shci_handle* h, int64_t* n) {
  *n = h->eng.n;
  return SHCI_OK;
}

shci_status shci_get_dets(shci_handle* h, int64_t* out, int64_t cap) {
  return guarded(h, [&] {
    Engine& E = h->eng;
    int64_t m = std::min(cap, E.n);
    if (m > 0) HIP_CHECK(hipMemcpy(out, E.dets.p, m * E.hs.W * 8, hipMemcpyDeviceToHost));
    return SHCI_OK;
  });
}

shci_status shci_screen(shci_handle* h, const double* cmax, double eps1, int64_t* n_new) {
  return guarded(h, [&] {
    Engine& E = h->eng;
    E.cmax.reserve((size_t)std::max<int64_t>(E.n, 1) * 8);
    if (E.n) HIP_CHECK(hipMemcpyAsync(E.cmax.p, cmax, E.n * 8, hipMemcpyHostToDevice, E.stream));
    *n_new = E.screen(E.cmax.as<double>(), eps1);
    HIP_CHECK(hipStreamSynchronize(E.stream));
    return SHCI_OK;
  });
}

shci_status shci_get_new_dets(shci_handle* h, int64_t* out, int64_t cap) {
  return guarded(h, [&] {
    Engine& E = h->eng;
    int64_t m = std::min(cap, E.nNew);
    if (m > 0) HIP_CHECK(hipMemcpy(out, E.newDets.p, m * E.hs.W * 8, hipMemcpyDeviceToHost));
    return SHCI_OK;
  });
}

shci_status shci_append_new(shci_handle* h) {
  return guarded(h, [&] {
    h->eng.appendNew();
    HIP_CHECK(hipStreamSynchronize(h->eng.stream));
    return SHCI_OK;
  });
}

shci_status shci_build_hamiltonian(shci_handle* h, int64_t* nnz_lower) {
  return guarded(h, [&] {
    h->eng.buildHamiltonian();
    if (nnz_lower) *nnz_lower = h->eng.lowerNnz();
    return SHCI_OK;
  });
}

shci_status shci_get_hamiltonian_rows(shci_handle* h, int64_t row_begin, int64_t* rowptr, int32_t* cols,
                                      double* vals, int64_t cap) {
  return guarded(h, [&] {
    Engine& E = h->eng;
    if (E.hamRows != E.n) throw std::invalid_argument("Hamiltonian not built for the held determinants");
    const int64_t n = E.n;
    if (row_begin < 0 || row_begin > n) throw std::invalid_argument("row_begin out of range");
    std::vector<int64_t> rp(n + 1);
    std::vector<int32_t> c(E.nnzOff);
    std::vector<double> v(E.nnzOff), dg(n);
    HIP_CHECK(hipMemcpy(rp.data(), E.rowPtr.p, (n + 1) * 8, hipMemcpyDeviceToHost));
    if (E.nnzOff) {
      HIP_CHECK(hipMemcpy(c.data(), E.cols.p, E.nnzOff * 4, hipMemcpyDeviceToHost));
      if (E.precision == 1) {  // FP32 storage: report hi + lo
        std::vector<float> hi(E.nnzOff), lo(E.nnzOff);
        HIP_CHECK(hipMemcpy(hi.data(), E.vals32.p, E.nnzOff * 4, hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(lo.data(), E.valsLo.p, E.nnzOff * 4, hipMemcpyDeviceToHost));
        for (int64_t q = 0; q < E.nnzOff; q++) v[q] = (double)hi[q] + (double)lo[q];
      } else {
        HIP_CHECK(hipMemcpy(v.data(), E.vals.p, E.nnzOff * 8, hipMemcpyDeviceToHost));
      }
    }
    HIP_CHECK(hipMemcpy(dg.data(), E.diag.p, n * 8, hipMemcpyDeviceToHost));
    int64_t k = 0;
    rowptr[0] = 0;
    std::vector<std::pair<int32_t, double>> row;
    for (int64_t I = row_begin; I < n; I++) {
      row.clear();
      for (int64_t p = rp[I]; p < rp[I + 1]; p++)
        if (c[p] < I) row.emplace_back(c[p], v[p]);
      std::sort(row.begin(), row.end(), [](const std::pair<int32_t, double>& a, const std::pair<int32_t, double>& b) {
        return a.first < b.first;
      });
      if (k + 1 + (int64_t)row.size() > cap) throw std::invalid_argument("output capacity too small");
      cols[k] = (int32_t)I;
      vals[k] = dg[I];
      k++;
      for (auto& e : row) {
        cols[k] = e.first;
        vals[k] = e.second;
        k++;
      }
      rowptr[I - row_begin + 1] = k;
    }
    return SHCI_OK;
  });
}

shci_status shci_sigma_device(shci_handle* h, const double* d_x, double* d_y, int k, int64_t ld) {
  return guarded(h, [&] {
    Engine& E = h->eng;
    if (E.hamRows != E.n) throw std::invalid_argument("Hamiltonian not built for the held determinants");
    if (k < 1 || ld < E.n) throw std::invalid_argument("need k >= 1 and ld >= number of determinants");
    const char *xb = (const char*)d_x, *xe = xb + ((k - 1) * ld + E.n) * 8;
    const char *yb = (const char*)d_y, *ye = yb + ((k - 1) * ld + E.n) * 8;
    if (xb < ye && yb < xe) throw std::invalid_argument("d_y overlaps d_x");
    E.sigma(d_x, d_y, k, ld, false);
    return SHCI_OK;
  });
}

shci_status shci_sigma_device_noalloc(shci_handle* h, const double* d_x, double* d_y, int k, int64_t ld) {
  return guarded(h, [&] {
    Engine& E = h->eng;
    if (E.hamRows != E.n) throw std::invalid_argument("Hamiltonian not built for the held determinants");
    if (k < 1 || ld < E.n) throw std::invalid_argument("need k >= 1 and ld >= number of determinants");
    const char *xb = (const char*)d_x, *xe = xb + ((k - 1) * ld + E.n) * 8;
    const char *yb = (const char*)d_y, *ye = yb + ((k - 1) * ld + E.n) * 8;
    if (xb < ye && yb < xe) throw std::invalid_argument("d_y overlaps d_x");
    if (E.trev || E.direct)
      return fail(h, SHCI_UNSUPPORTED, "shci_sigma_device_noalloc: not available in time-reversal or direct mode");
    size_t cap = 0;
    (void)fixedSigmaScratch(&cap);
    const size_t need = E.sigmaFixedScratchNeed(k);
    if (need > cap - 256)
      return fail(h, SHCI_UNSUPPORTED,
                  "shci_sigma_device_noalloc: " + std::to_string(E.n) + " determinants x " + std::to_string(k) +
                      " vectors need " + std::to_string(need) + " bytes of scratch, more than its fixed " +
                      std::to_string(cap - 256) + " (use shci_sigma_device)");
    E.sigmaFixed(d_x, d_y, k, ld);
    return SHCI_OK;
  });
}

shci_status shci_sigma(shci_handle* h, const double* x, double* y, int k) {
  return guarded(h, [&] {
    Engine& E = h->eng;
    if (E.hamRows != E.n) throw std::invalid_argument("Hamiltonian not built for the held determinants");
    const int64_t n = E.n;
    DevMem dx, dy;
    dx.reserve(std::max<int64_t>(n * k, 1) * 8);
    dy.reserve(std::max<int64_t>(n * k, 1) * 8);
    HIP_CHECK(hipMemcpyAsync(dx.p, x, n * k * 8, hipMemcpyHostToDevice, E.stream));
    E.sigma(dx.as<double>(), dy.as<double>(), k, n, false);
    HIP_CHECK(hipMemcpyAsync(y, dy.p, n * k * 8, hipMemcpyDeviceToHost, E.stream));
    HIP_CHECK(hipStreamSynchronize(E.stream));
    return SHCI_OK;
  });
}

shci_status shci_davidson(shci_handle* h, double* x0, int nroots, int max_copies, double tol, double* energies,
                          int* iterations) {
  return guarded(h, [&] {
    Engine& E = h->eng;
    if (E.hamRows != E.n) throw std::invalid_argument("Hamiltonian not built for the held determinants");
    if (nroots < 1 || max_copies < nroots + 3 || max_copies > 12)
      throw std::invalid_argument("need nroots >= 1 and nroots + 3 <= max_copies <= 12");
    const int64_t n = E.n;
    DevMem dx;
    dx.reserve(n * nroots * 8);
    HIP_CHECK(hipMemcpyAsync(dx.p, x0, n * nroots * 8, hipMemcpyHostToDevice, E.stream));
    int it = 0;
    std::vector<double> e = E.davidsonMode == 1 ? E.davidsonBlock(dx.as<double>(), nroots, tol, it, false)
                                                : E.davidson(dx.as<double>(), nroots, max_copies, tol, it, false);
    if ((int)e.size() != nroots) return fail(h, SHCI_NO_CONVERGENCE, E.err);
    HIP_CHECK(hipMemcpyAsync(x0, dx.p, n * nroots * 8, hipMemcpyDeviceToHost, E.stream));
    HIP_CHECK(hipStreamSynchronize(E.stream));
    for (int r = 0; r < nroots; r++) energies[r] = e[r];
    if (iterations) *iterations = it;
    return SHCI_OK;
  });
}

shci_status shci_pt_deterministic(shci_handle* h, const double* ci, double e0, double eps2, double* e_pt,
                                  double* psi1_norm, int64_t* n_external) {
  return guarded(h, [&] {
    Engine& E = h->eng;
    if (!ci || !e_pt) throw std::invalid_argument("null argument");
    const int64_t n = E.n;
    DevMem dc;
    dc.reserve((size_t)std::max<int64_t>(n, 1) * 8);
    if (n) HIP_CHECK(hipMemcpyAsync(dc.p, ci, n * 8, hipMemcpyHostToDevice, E.stream));
    *e_pt = E.ptDeterministic(dc.as<double>(), e0, eps2, psi1_norm, n_external);
    return SHCI_OK;
  });
}

shci_status shci_rdm(shci_handle* h, const double* bra, const double* ket, double* one_rdm, double* two_rdm,
                     double* spin_one_rdm, double* spin_two_rdm) {
  return guarded(h, [&] {
    Engine& E = h->eng;
    if (!bra || !ket || !one_rdm || !two_rdm) throw std::invalid_argument("null argument");
    const int64_t n = E.n;
    if (n < 1) throw std::invalid_argument("no determinants held");
    DevMem db, dk;
    db.reserve((size_t)n * 8);
    HIP_CHECK(hipMemcpyAsync(db.p, bra, n * 8, hipMemcpyHostToDevice, E.stream));
    if (ket != bra) {  // (the same vector on both sides lets the kernels use hermiticity)
      dk.reserve((size_t)n * 8);
      HIP_CHECK(hipMemcpyAsync(dk.p, ket, n * 8, hipMemcpyHostToDevice, E.stream));
    }
    const double* dKet = ket != bra ? dk.as<double>() : db.as<double>();
    std::vector<double> g1, g2, s1, s2;
    E.rdm(db.as<double>(), dKet, g1, g2, spin_one_rdm ? &s1 : nullptr, spin_two_rdm ? &s2 : nullptr);
    std::copy(g1.begin(), g1.end(), one_rdm);
    std::copy(g2.begin(), g2.end(), two_rdm);
    if (spin_one_rdm) std::copy(s1.begin(), s1.end(), spin_one_rdm);
    if (spin_two_rdm) std::copy(s2.begin(), s2.end(), spin_two_rdm);
    return SHCI_OK;
  });
}

shci_status shci_pt_semistochastic(shci_handle* h, const double* ci, double e0, double eps2, double eps2_large,
                                   int sample_n, int n_iter, double target_error, unsigned seed, double* e_pt,
                                   double* per_iter, int cap, int* iterations, double* error) {
  return guarded(h, [&] {
    Engine& E = h->eng;
    if (!ci || !e_pt) throw std::invalid_argument("null argument");
    if (sample_n < 2) throw std::invalid_argument("sample_n must be at least 2");
    const int64_t n = E.n;
    DevMem dc;
    dc.reserve((size_t)std::max<int64_t>(n, 1) * 8);
    if (n) HIP_CHECK(hipMemcpyAsync(dc.p, ci, n * 8, hipMemcpyHostToDevice, E.stream));
    HIP_CHECK(hipStreamSynchronize(E.stream));
    E.rng.reseed(seed);
    std::vector<double> per;
    *e_pt = E.ptSemistochastic(dc.as<double>(), e0, eps2, eps2_large, sample_n, n_iter, target_error, 0.0, 0,
                               nullptr, 0.0, error, iterations, &per);
    for (int k = 0; per_iter && k < cap && k < (int)per.size(); k++) per_iter[k] = per[k];
    return SHCI_OK;
  });
}

shci_status shci_run(const char* input_path, const char* log_path, int precision, shci_handle** out) {
  shci_run_options opt;
  std::memset(&opt, 0, sizeof(opt));
  opt.precision = precision;
  return shci_run_ex(input_path, log_path, &opt, out);
}

shci_status shci_run_ex(const char* input_path, const char* log_path, const shci_run_options* options,
                        shci_handle** out) {
  if (!out || !input_path || !options) return fail(nullptr, SHCI_BAD_INPUT, "null argument");
  const shci_run_options opt = *options;
  *out = nullptr;
  shci_handle* h = nullptr;
  try {
    h = new shci_handle();
  } catch (const std::exception& e) {
    return fail(nullptr, SHCI_HIP_ERROR, e.what());
  }
  FILE* log = stdout;
  bool own = false;
  if (log_path && log_path[0]) {
    log = fopen(log_path, "w");
    own = log != nullptr;
    if (!log) log = stdout;
  }
  shci_status st = guarded(h, [&] { return runVariational(h, input_path, log, opt); });
  if (own) fclose(log);
  *out = h;
  return st;
}

int shci_run_iterations(const shci_handle* h) { return (int)h->eng.iters.size(); }

shci_status shci_run_iteration_phases(const shci_handle* h, int it, double* screen_s, double* hamiltonian_s,
                                      double* davidson_s, int64_t* nnz_lower) {
  if (it < 0 || it >= (int)h->eng.iters.size()) return SHCI_BAD_INPUT;
  const IterRecord& r = h->eng.iters[it];
  if (screen_s) *screen_s = r.tScreen;
  if (hamiltonian_s) *hamiltonian_s = r.tHam;
  if (davidson_s) *davidson_s = r.tDavidson;
  if (nnz_lower) *nnz_lower = r.nnzLower;
  return SHCI_OK;
}

double shci_run_final_davidson_seconds(const shci_handle* h) { return h->eng.tFinalDavidson; }

shci_status shci_profile_sigma_read(shci_handle* h, int64_t* calls, double* milliseconds, double* bytes) {
  return guarded(h, [&] {
    int64_t c = 0;
    double ms = 0, b = 0;
    if (h->eng.stream) h->eng.readSigmaProfile(c, ms, b);
    if (calls) *calls = c;
    if (milliseconds) *milliseconds = ms;
    if (bytes) *bytes = b;
    return SHCI_OK;
  });
}

shci_status shci_run_iteration(const shci_handle* h, int it, int64_t* ndets, double* energy0, int* davidson_iters,
                               double* seconds) {
  if (it < 0 || it >= (int)h->eng.iters.size()) return SHCI_BAD_INPUT;
  const IterRecord& r = h->eng.iters[it];
  if (ndets) *ndets = r.ndets;
  if (energy0) *energy0 = r.e0[0];
  if (davidson_iters) *davidson_iters = r.davidsonIters;
  if (seconds) *seconds = r.seconds;
  return SHCI_OK;
}

shci_status shci_run_energies(const shci_handle* h, double* energies, int* nroots) {
  for (size_t i = 0; i < h->eng.energies.size(); i++) energies[i] = h->eng.energies[i];
  if (nroots) *nroots = (int)h->eng.energies.size();
  return SHCI_OK;
}

shci_status shci_run_pt_energies(const shci_handle* h, double* energies, int* nroots) {
  for (size_t i = 0; i < h->eng.ptEnergies.size(); i++) energies[i] = h->eng.ptEnergies[i];
  if (nroots) *nroots = (int)h->eng.ptEnergies.size();
  return SHCI_OK;
}

shci_status shci_run_ci(shci_handle* h, double* out, int64_t cap) {
  return guarded(h, [&] {
    Engine& E = h->eng;
    int64_t m = std::min(cap, E.n * (int64_t)E.nroots);
    if (m > 0) HIP_CHECK(hipMemcpy(out, E.ci.p, m * 8, hipMemcpyDeviceToHost));
    return SHCI_OK;
  });
}

}  // extern "C"

// ---- candidate-ranking MLP (handle-free) ----------------------------------

namespace {
struct MlpCall {
  hipStream_t s = nullptr;
  shci::MlpModel model;
  shci::DevMem X, t, y;
  MlpCall(int din, int hidden) {
    if (din < 1 || hidden < 1 || hidden > 4096 || din > 65536) throw std::invalid_argument("bad MLP shape");
    HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    model.init(din, hidden, 0, s);
  }
  ~MlpCall() {
    if (s) (void)hipStreamDestroy(s);
  }
  void upload(shci::DevMem& m, const float* h, size_t count) {
    m.reserve(std::max<size_t>(count, 1) * sizeof(float));
    if (count) HIP_CHECK(hipMemcpyAsync(m.p, h, count * sizeof(float), hipMemcpyHostToDevice, s));
  }
};
}  // namespace

int64_t shci_mlp_param_count(int din, int hidden) { return (int64_t)shci::mlpLayout(din, hidden).total; }

shci_status shci_mlp_init(float* params, int din, int hidden, uint64_t seed) {
  return guarded(nullptr, [&] {
    MlpCall c(din, hidden);
    c.model.init(din, hidden, seed, c.s);
    c.model.getParams(params, c.s);
    return SHCI_OK;
  });
}

shci_status shci_mlp_forward(const float* params, int din, int hidden, const float* X, int64_t n, float* y) {
  return guarded(nullptr, [&] {
    if (n < 0) throw std::invalid_argument("n < 0");
    MlpCall c(din, hidden);
    c.model.setParams(params, c.s);
    c.upload(c.X, X, (size_t)n * din);
    c.y.reserve(std::max<int64_t>(n, 1) * sizeof(float));
    c.model.predict(c.X.as<float>(), n, c.y.as<float>(), c.s);
    if (n) HIP_CHECK(hipMemcpyAsync(y, c.y.p, n * sizeof(float), hipMemcpyDeviceToHost, c.s));
    HIP_CHECK(hipStreamSynchronize(c.s));
    return SHCI_OK;
  });
}

shci_status shci_mlp_loss_grad(const float* params, int din, int hidden, const float* X, const float* t, int64_t n,
                               float* loss, float* grad) {
  return guarded(nullptr, [&] {
    if (n < 1) throw std::invalid_argument("need n >= 1");
    MlpCall c(din, hidden);
    c.model.setParams(params, c.s);
    c.upload(c.X, X, (size_t)n * din);
    c.upload(c.t, t, (size_t)n);
    *loss = c.model.lossGrad(c.X.as<float>(), c.t.as<float>(), n, c.s);
    HIP_CHECK(hipMemcpyAsync(grad, c.model.G.p, c.model.L.total * sizeof(float), hipMemcpyDeviceToHost, c.s));
    HIP_CHECK(hipStreamSynchronize(c.s));
    return SHCI_OK;
  });
}

shci_status shci_mlp_train(float* params, int din, int hidden, const float* X, const float* t, int64_t n, int epochs,
                           int batch, float lr, uint64_t seed, float* first_loss, float* last_loss) {
  return guarded(nullptr, [&] {
    if (n < 1 || batch < 1 || epochs < 0) throw std::invalid_argument("need n >= 1, batch >= 1, epochs >= 0");
    MlpCall c(din, hidden);
    c.model.setParams(params, c.s);
    c.upload(c.X, X, (size_t)n * din);
    c.upload(c.t, t, (size_t)n);
    float f = 0.f, l = 0.f;
    c.model.train(c.X.as<float>(), c.t.as<float>(), n, epochs, batch, lr, seed, f, l, c.s);
    c.model.getParams(params, c.s);
    if (first_loss) *first_loss = f;
    if (last_loss) *last_loss = l;
    return SHCI_OK;
  });
}

int shci_run_learned_count(const shci_handle* h) { return h ? (int)h->eng.learnedRecs.size() : 0; }

shci_status shci_run_learned(const shci_handle* h, int k, int* iteration, int64_t* candidates, int64_t* evaluated,
                             int64_t* accepted, double* sampled_recall, double* sampled_fraction, int* fallback,
                             double* seconds) {
  if (!h || k < 0 || k >= (int)h->eng.learnedRecs.size()) return SHCI_BAD_INPUT;
  const Engine::LearnedRecord& r = h->eng.learnedRecs[k];
  if (iteration) *iteration = r.iter;
  if (candidates) *candidates = r.candidates;
  if (evaluated) *evaluated = r.evaluated;
  if (accepted) *accepted = r.accepted;
  if (sampled_recall) *sampled_recall = r.samplePass ? (double)r.sampleHit / (double)r.samplePass : 1.0;
  if (sampled_fraction) *sampled_fraction = r.sampleTotal ? (double)r.sampleSelected / (double)r.sampleTotal : 0.0;
  if (fallback) *fallback = r.fallback ? 1 : 0;
  if (seconds) *seconds = r.seconds;
  return SHCI_OK;
}

shci_status shci_run_learned_training(const shci_handle* h, int64_t* samples, float* first_loss, float* last_loss,
                                      double* seconds) {
  if (!h) return SHCI_BAD_INPUT;
  const Engine& E = h->eng;
  if (samples) *samples = (int64_t)E.sampT.size();
  if (first_loss) *first_loss = E.trainFirst;
  if (last_loss) *last_loss = E.trainLast;
  if (seconds) *seconds = E.trainSeconds;
  return SHCI_OK;
}
