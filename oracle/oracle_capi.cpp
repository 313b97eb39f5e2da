// C entry points of the CPU oracle, for the test suite and bench.py's
// cpu_baseline leg (loaded with ctypes).  TEST INFRASTRUCTURE ONLY.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "cpu_ref.h"

using namespace oref;

namespace {
struct Handle {
  System s;
  Schedule schd;  // default schedule (ncore 0, nact 1e6, no restrictions)
  SparseHam H;
  Helpers h;
  std::vector<Det> dets;
  std::string err;
};
thread_local std::string g_err;

Det loadDet(const int64_t* p, int W) {
  Det d;
  for (int i = 0; i < W; i++) d.r[i] = p[i];
  return d;
}
}  // namespace

extern "C" {

const char* oref_last_error() { return g_err.c_str(); }

void* oref_create(const char* fcidump) {
  std::unique_ptr<Handle> h(new Handle());
  if (!readIntegrals(fcidump, h->s, g_err)) return nullptr;
  buildHeatBath(h->s);
  return h.release();
}
void oref_destroy(void* p) { delete (Handle*)p; }

int oref_norb(void* p) { return ((Handle*)p)->s.norb; }
int oref_nelec(void* p) { return ((Handle*)p)->s.nelec; }
int oref_det_words(void* p) { return ((Handle*)p)->s.W; }
double oref_core_energy(void* p) { return ((Handle*)p)->s.coreE; }
double oref_max_entry(void* p) { return ((Handle*)p)->s.maxEntry; }

int64_t oref_i2_size(void* p) { return (int64_t)((Handle*)p)->s.I2.size(); }
void oref_get_integrals(void* p, double* I1, double* I2, double* direct, double* exchange) {
  System& s = ((Handle*)p)->s;
  if (I1) memcpy(I1, s.I1.data(), s.I1.size() * sizeof(double));
  if (I2) memcpy(I2, s.I2.data(), s.I2.size() * sizeof(double));
  if (direct) memcpy(direct, s.Direct.data(), s.Direct.size() * sizeof(double));
  if (exchange) memcpy(exchange, s.Exchange.data(), s.Exchange.size() * sizeof(double));
}

int64_t oref_hb_size(void* p, int same) {
  System& s = ((Handle*)p)->s;
  return (int64_t)(same ? s.hbSameVal.size() : s.hbOppVal.size());
}
void oref_hb_get(void* p, int same, float* vals, int16_t* pairs, int64_t* start) {
  System& s = ((Handle*)p)->s;
  const std::vector<float>& v = same ? s.hbSameVal : s.hbOppVal;
  const std::vector<int16_t>& pr = same ? s.hbSamePair : s.hbOppPair;
  const std::vector<size_t>& st = same ? s.hbSameStart : s.hbOppStart;
  memcpy(vals, v.data(), v.size() * sizeof(float));
  memcpy(pairs, pr.data(), pr.size() * sizeof(int16_t));
  for (size_t i = 0; i < st.size(); i++) start[i] = (int64_t)st[i];
}
void oref_singles(void* p, double* out) {
  System& s = ((Handle*)p)->s;
  memcpy(out, s.Singles.data(), s.Singles.size() * sizeof(double));
}

// screening restrictions (defaults: ncore 0, nact 1e6, none)
void oref_set_active(void* p, int ncore, int nact) {
  ((Handle*)p)->schd.ncore = ncore;
  ((Handle*)p)->schd.nact = nact;
}

// restriction applied to the PT space (the "restrict" keyword's PT half)
void oref_add_pt_restriction(void* p, int minElec, int maxElec, const int* spatial, int n) {
  Restriction r;
  r.minElec = minElec;
  r.maxElec = maxElec;
  for (int k = 0; k < n; k++) {
    r.orbs.push_back(2 * spatial[k]);
    r.orbs.push_back(2 * spatial[k] + 1);
  }
  ((Handle*)p)->schd.restrictionsPT.push_back(r);
}

// deterministic EN-PT2 of the n determinants with coefficients ci (one root)
// and variational energy e0 (no core energy)
double oref_pt_deterministic(void* p, const int64_t* dets, int64_t n, const double* ci, double e0, double eps2,
                             double* psi1Norm, int64_t* nExternal) {
  Handle* h = (Handle*)p;
  std::vector<Det> d(n);
  for (int64_t i = 0; i < n; i++) d[i] = loadDet(dets + i * h->s.W, h->s.W);
  size_t next = 0;
  const double pt = ptDeterministic(h->s, h->schd, d, ci, e0, eps2, psi1Norm, &next);
  if (nExternal) *nExternal = (int64_t)next;
  return pt;
}

// semistochastic EN-PT2 of one root (srand(seed) first, as the run's start
// does); perIter (cap entries) receives each iteration's estimate; returns
// the reference function's value
double oref_pt_semistochastic(void* p, const int64_t* dets, int64_t n, const double* ci, double e0, double eps2,
                              double eps2Large, int sampleN, int nPTiter, double targetError, unsigned seed,
                              double* perIter, int cap, int* iters, double* error) {
  Handle* h = (Handle*)p;
  std::vector<Det> d(n);
  for (int64_t i = 0; i < n; i++) d[i] = loadDet(dets + i * h->s.W, h->s.W);
  Schedule schd = h->schd;
  schd.epsilon2 = eps2;
  schd.epsilon2Large = eps2Large;
  schd.SampleN = sampleN;
  schd.nPTiter = nPTiter;
  schd.targetError = targetError;
  std::srand(seed);
  std::vector<double> est;
  const double r = ptSemistochastic(h->s, schd, d, ci, e0, 0.0, 0, nullptr, 0.0, error, iters, &est);
  for (int k = 0; k < (int)est.size() && k < cap; k++) perIter[k] = est[k];
  return r;
}

double oref_energy(void* p, const int64_t* det) {
  Handle* h = (Handle*)p;
  return energy(h->s, loadDet(det, h->s.W));
}

double oref_hij(void* p, const int64_t* bra, const int64_t* ket) {
  Handle* h = (Handle*)p;
  return hij(h->s, loadDet(bra, h->s.W), loadDet(ket, h->s.W));
}

// dets: n x W words; cmax: n; sorted: nsorted x W (ascending, signed order).
// Writes the sorted unique new candidates (up to cap) and returns their count.
int64_t oref_screen(void* p, const int64_t* dets, int64_t n, const double* cmax, double eps1,
                    const int64_t* sorted, int64_t nsorted, int64_t* out, int64_t cap) {
  Handle* h = (Handle*)p;
  int W = h->s.W;
  std::vector<Det> sd(nsorted);
  for (int64_t i = 0; i < nsorted; i++) sd[i] = loadDet(sorted + i * W, W);
  std::vector<Det> cand;
  for (int64_t i = 0; i < n; i++)
    screenDet(h->s, h->schd, loadDet(dets + i * W, W), eps1 / std::fabs(cmax[i]), sd.data(), sd.size(), cand);
  std::sort(cand.begin(), cand.end(), [W](const Det& a, const Det& b) { return detLess(a, b, W); });
  cand.erase(std::unique(cand.begin(), cand.end(), [W](const Det& a, const Det& b) { return detEq(a, b, W); }),
             cand.end());
  for (int64_t i = 0; i < (int64_t)cand.size() && i < cap; i++)
    for (int w = 0; w < W; w++) out[i * W + w] = cand[i].r[w];
  return (int64_t)cand.size();
}

// Append dets[nold:n] (nold = number already held) and build their H rows;
// returns the total number of stored elements.
int64_t oref_csr_append(void* p, const int64_t* dets, int64_t n) {
  Handle* h = (Handle*)p;
  int W = h->s.W;
  int start = (int)h->dets.size();
  for (int64_t i = start; i < n; i++) h->dets.push_back(loadDet(dets + i * W, W));
  populateHelpers(h->s, h->h, h->dets, start);
  makeHRows(h->s, h->h, h->dets, start, h->H);
  int64_t nnz = 0;
  for (auto& r : h->H.connections) nnz += (int64_t)r.size();
  return nnz;
}
// Rows in reference order: rowptr (n+1), cols, vals (as stored, diagonal first).
void oref_csr_get(void* p, int64_t* rowptr, int32_t* cols, double* vals) {
  Handle* h = (Handle*)p;
  int64_t k = 0;
  rowptr[0] = 0;
  for (size_t i = 0; i < h->H.connections.size(); i++) {
    for (size_t j = 0; j < h->H.connections[i].size(); j++) {
      cols[k] = h->H.connections[i][j];
      vals[k] = h->H.Helements[i][j];
      k++;
    }
    rowptr[i + 1] = k;
  }
}
void oref_sigma(void* p, const double* x, double* y) {
  Handle* h = (Handle*)p;
  hmult(h->H, x, y, h->H.connections.size());
}
// HmultDirect over the determinants appended with oref_csr_append
void oref_sigma_direct(void* p, const double* x, double* y) {
  Handle* h = (Handle*)p;
  hmultDirect(h->s, h->dets, h->H, x, y, h->H.connections.size());
}
// RDMs of <bra| ... |ket> over the determinants appended with oref_csr_append
// (zeroes the outputs first; s1 / s2 may be null).
void oref_rdm(void* p, const double* bra, const double* ket, double* g1, double* g2, double* s1, double* s2) {
  Handle* h = (Handle*)p;
  const size_t n = h->s.norb, nso = h->s.nso, npair = nso * (nso + 1) / 2;
  std::fill(g1, g1 + n * n, 0.0);
  std::fill(g2, g2 + n * n * n * n, 0.0);
  if (s1) std::fill(s1, s1 + nso * nso, 0.0);
  if (s2) std::fill(s2, s2 + npair * npair, 0.0);
  evaluateRdm(h->s, h->dets, h->H, bra, ket, g1, g2, s1, s2);
}
// Davidson on the stored H (x0: nroots columns of length n, in/out).
int oref_davidson(void* p, double* x0, int nroots, const double* diag, int maxCopies, double tol, double* eroots) {
  Handle* h = (Handle*)p;
  size_t n = h->H.connections.size();
  std::vector<std::vector<double>> x(nroots, std::vector<double>(n));
  for (int r = 0; r < nroots; r++) memcpy(x[r].data(), x0 + r * n, n * sizeof(double));
  std::vector<double> d(diag, diag + n);
  int numIter = 0;
  std::vector<double> e = davidson(h->H, x, d, maxCopies, tol, numIter, false, nullptr);
  for (int r = 0; r < nroots; r++) {
    memcpy(x0 + r * n, x[r].data(), n * sizeof(double));
    if (r < (int)e.size()) eroots[r] = e[r];
  }
  return numIter;
}

void oref_sym_eig(int m, const double* A, double* w, double* V) { symEig(m, A, w, V); }
int oref_alias_draw_index(int r, int n) { return aliasDrawIndex(r, n); }
int oref_rand_max() { return RAND_MAX; }

// ---- whole variational runs (input.dat driven) -----------------------------
void* oref_run(const char* input_path, const char* log_path, int record_iters) {
  RunResult* r = new RunResult();
  FILE* log = nullptr;
  if (log_path && log_path[0]) log = fopen(log_path, "w");
  runVariational(input_path, *r, log, record_iters != 0, true);
  if (log) fclose(log);
  return r;
}
int oref_run_status(void* p) { return ((RunResult*)p)->status; }
const char* oref_run_error(void* p) { return ((RunResult*)p)->error.c_str(); }
int oref_run_niter(void* p) { return (int)((RunResult*)p)->iters.size(); }
int oref_run_nroots(void* p) { return (int)((RunResult*)p)->energies.size(); }
void oref_run_iter(void* p, int it, int64_t* ndets, double* e0, int* ndav) {
  RunResult* r = (RunResult*)p;
  const IterRecord& rec = r->iters[it];
  *ndets = rec.ndets;
  for (size_t i = 0; i < rec.e0.size(); i++) e0[i] = rec.e0[i];
  *ndav = rec.numIter;
}
void oref_run_energies(void* p, double* e) {
  RunResult* r = (RunResult*)p;
  for (size_t i = 0; i < r->energies.size(); i++) e[i] = r->energies[i];
}
// E_var + E_PT per root when the run did the deterministic PT step; returns the count (0 otherwise)
int oref_run_pt(void* p, double* e) {
  RunResult* r = (RunResult*)p;
  for (size_t i = 0; i < r->ptEnergies.size(); i++) e[i] = r->ptEnergies[i];
  return (int)r->ptEnergies.size();
}
int64_t oref_run_ndets(void* p) { return (int64_t)((RunResult*)p)->dets.size(); }
int64_t oref_run_iter_ndets(void* p, int it) { return (int64_t)((RunResult*)p)->detsPerIter[it].size(); }
void oref_run_dets(void* p, int it, int64_t* out, int W) {
  RunResult* r = (RunResult*)p;
  const std::vector<Det>& d = it < 0 ? r->dets : r->detsPerIter[it];
  for (size_t i = 0; i < d.size(); i++)
    for (int w = 0; w < W; w++) out[i * W + w] = d[i].r[w];
}
void oref_run_ci(void* p, double* out) {
  RunResult* r = (RunResult*)p;
  size_t n = r->dets.size();
  for (size_t k = 0; k < r->ci.size(); k++) memcpy(out + k * n, r->ci[k].data(), n * sizeof(double));
}
void oref_run_free(void* p) { delete (RunResult*)p; }

}  // extern "C"
