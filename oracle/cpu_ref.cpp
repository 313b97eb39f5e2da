// cpu_ref — CPU ORACLE for the SHCI variational step (test infrastructure).
// See cpu_ref.h for scope, citations and the documented deviations.
#include "cpu_ref.h"

#ifdef _OPENMP
#include <omp.h>
#endif

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <sys/time.h>

namespace oref {

// Host threads for the loops the reference spreads over MPI ranks (rows of
// H, screened determinants, the sigma rows): ORACLE_THREADS (default 1, the
// serial reference order).  With more threads H and the determinant sets
// are unchanged; sigma sums its per-thread partial vectors in thread order,
// as the reference's MPI_Reduce sums per-rank partials.
static int oracleThreads() {
  const char* t = std::getenv("ORACLE_THREADS");
  int n = t ? std::atoi(t) : 1;
  return n < 1 ? 1 : n;
}

static double wallTime() {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + 1e-6 * tv.tv_usec;
}

// ---------------------------------------------------------------------------
// string helpers mirroring boost::trim / boost::split(token_compress_on) /
// boost::iequals as used by SHCI/integral.cpp:83-86 and SHCI/input.cpp:129-134
static std::string trim(const std::string& s) {
  size_t b = 0, e = s.size();
  while (b < e && std::isspace((unsigned char)s[b])) b++;
  while (e > b && std::isspace((unsigned char)s[e - 1])) e--;
  return s.substr(b, e - b);
}

static std::vector<std::string> split(const std::string& s, const char* delims) {
  // token_compress_on: adjacent delimiters form one separator; leading /
  // trailing delimiters still produce one empty token each.
  std::vector<std::string> out;
  std::string cur;
  size_t i = 0, n = s.size();
  auto isd = [&](char c) { return std::strchr(delims, c) != nullptr && c != 0; };
  while (i < n) {
    if (isd(s[i])) {
      out.push_back(cur);
      cur.clear();
      while (i < n && isd(s[i])) i++;
    } else {
      cur.push_back(s[i]);
      i++;
    }
  }
  out.push_back(cur);
  return out;
}

static bool iequals(const std::string& a, const std::string& b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); i++)
    if (std::tolower((unsigned char)a[i]) != std::tolower((unsigned char)b[i])) return false;
  return true;
}

static std::string prefixOf(const std::string& s, size_t n) { return s.substr(0, std::min(n, s.size())); }

// ---------------------------------------------------------------------------
// SHCI/integral.cpp:43-225
bool readIntegrals(const std::string& path, System& s, std::string& err) {
  std::ifstream dump(path.c_str());
  if (!dump.good()) {
    err = "Integral file " + path + " does not exist!";
    return false;
  }
  int norbs = -1, nelec = -1;
  bool ksym = false;
  std::string msg;
  while (!dump.eof()) {
    std::getline(dump, msg);
    msg = trim(msg);
    std::vector<std::string> tok = split(msg, ", \t=");
    if (tok.size() == 1 && (iequals(tok[0], "&END") || iequals(tok[0], "/"))) break;
    if (iequals(prefixOf(tok[0], 4), "&FCI")) {
      if (tok.size() > 2 && iequals(prefixOf(tok[1], 4), "NORB")) norbs = atoi(tok[2].c_str());
      if (tok.size() > 4 && iequals(prefixOf(tok[3], 5), "NELEC")) nelec = atoi(tok[4].c_str());
    } else if (iequals(prefixOf(tok[0], 4), "ISYM")) {
      continue;
    } else if (iequals(prefixOf(tok[0], 4), "KSYM")) {
      ksym = true;
    } else if (iequals(prefixOf(tok[0], 6), "ORBSYM")) {
      for (size_t i = 1; i < tok.size(); i++) s.irrep.push_back(atoi(tok[i].c_str()));
    } else {
      for (size_t i = 0; i < tok.size(); i++) s.irrep.push_back(atoi(tok[i].c_str()));
    }
  }
  if (norbs == -1 || nelec == -1) {
    err = "could not read the norbs or nelec";
    return false;
  }
  if (ksym) {
    err = "KSYM integrals are outside the restated path";
    return false;
  }
  if (norbs > 63) {
    err = "oracle supports at most 63 spatial orbitals";
    return false;
  }
  s.irrep.resize(norbs);
  s.norb = norbs;
  s.nso = 2 * norbs;
  s.nelec = nelec;
  s.W = s.nso / 64 + 1;  // SHCI/SHCI.cpp:236
  size_t npair = (size_t)norbs * (norbs + 1) / 2;
  s.I2.assign(npair * (npair + 1) / 2, 0.0);
  s.I1.assign((size_t)s.nso * s.nso, 0.0);
  s.coreE = 0.0;
  while (!dump.eof()) {
    std::getline(dump, msg);
    msg = trim(msg);
    std::vector<std::string> tok = split(msg, ", \t");
    if (tok.size() != 5) continue;
    double integral = atof(tok[0].c_str());
    int a = atoi(tok[1].c_str()), b = atoi(tok[2].c_str()), c = atoi(tok[3].c_str()),
        d = atoi(tok[4].c_str());
    if (a == b && b == c && c == d && d == 0) {
      s.coreE = integral;
    } else if (b == c && c == d && d == 0) {
      continue;
    } else if (c == d && d == 0) {
      s.I1[(size_t)(2 * (a - 1)) * s.nso + 2 * (b - 1)] = integral;
      s.I1[(size_t)(2 * (a - 1) + 1) * s.nso + 2 * (b - 1) + 1] = integral;
      s.I1[(size_t)(2 * (b - 1)) * s.nso + 2 * (a - 1)] = integral;
      s.I1[(size_t)(2 * (b - 1) + 1) * s.nso + 2 * (a - 1) + 1] = integral;
    } else {
      s.i2ref(2 * (a - 1), 2 * (b - 1), 2 * (c - 1), 2 * (d - 1)) = integral;
    }
  }
  // maxEntry: the element of largest magnitude, first one on ties (std::max_element)
  double best = s.I2.empty() ? 0.0 : s.I2[0];
  for (size_t k = 1; k < s.I2.size(); k++)
    if (std::fabs(best) < std::fabs(s.I2[k])) best = s.I2[k];
  s.maxEntry = best;
  s.Direct.assign((size_t)norbs * norbs, 0.0);
  s.Exchange.assign((size_t)norbs * norbs, 0.0);
  for (int i = 0; i < norbs; i++)
    for (int j = 0; j < norbs; j++) {
      s.Direct[(size_t)i * norbs + j] = s.i2(2 * i, 2 * i, 2 * j, 2 * j);
      s.Exchange[(size_t)i * norbs + j] = s.i2(2 * i, 2 * j, 2 * j, 2 * i);
    }
  return true;
}

// SHCI/integral.h:125-153 (twoIntHeatBath::constructClass, epsilon 1e-10 from
// SHCI/SHCI.cpp:250) and SHCI/integral.cpp:228-389 (flattening by reverse
// iteration of the |float|-keyed multimap).
void buildHeatBath(System& s) {
  const double eps = 1e-10;
  int n = s.norb;
  size_t npair = (size_t)n * (n + 1) / 2;
  struct Ent {
    float v;
    int16_t a, b;
  };
  s.hbSameStart.assign(npair + 1, 0);
  s.hbOppStart.assign(npair + 1, 0);
  s.hbSameVal.clear();
  s.hbOppVal.clear();
  s.hbSamePair.clear();
  s.hbOppPair.clear();
  size_t pairIter = 1;
  std::vector<Ent> same, opp;
  for (int i = 0; i < n; i++)
    for (int j = 0; j <= i; j++) {
      same.clear();
      opp.clear();
      for (int a = 0; a < n; a++)
        for (int b = 0; b < n; b++) {
          double vo = s.i2(2 * i, 2 * a, 2 * j, 2 * b);
          if (std::fabs(vo) > eps) opp.push_back(Ent{(float)vo, (int16_t)a, (int16_t)b});
          if (a >= b) {
            double vs = s.i2(2 * i, 2 * a, 2 * j, 2 * b) - s.i2(2 * i, 2 * b, 2 * j, 2 * a);
            if (std::fabs(vs) > eps) same.push_back(Ent{(float)vs, (int16_t)a, (int16_t)b});
          }
        }
      // multimap<float,...,compAbs>: ascending |float|, ties in insertion order;
      // read back with a reverse iterator.
      auto cmp = [](const Ent& x, const Ent& y) { return std::fabs(x.v) < std::fabs(y.v); };
      std::stable_sort(same.begin(), same.end(), cmp);
      std::stable_sort(opp.begin(), opp.end(), cmp);
      for (auto it = same.rbegin(); it != same.rend(); ++it) {
        s.hbSameVal.push_back(it->v);
        s.hbSamePair.push_back(it->a);
        s.hbSamePair.push_back(it->b);
      }
      for (auto it = opp.rbegin(); it != opp.rend(); ++it) {
        s.hbOppVal.push_back(it->v);
        s.hbOppPair.push_back(it->a);
        s.hbOppPair.push_back(it->b);
      }
      s.hbSameStart[pairIter] = s.hbSameVal.size();
      s.hbOppStart[pairIter] = s.hbOppVal.size();
      pairIter++;
    }
  int N = s.nso;
  s.Singles.assign((size_t)N * N, 0.0);
  for (int i = 0; i < N; i++)
    for (int a = 0; a < N; a++) {
      double v = std::fabs(s.i1(i, a));
      for (int j = 0; j < N; j++) {
        double t = s.i2(i, a, j, j) - s.i2(i, j, j, a);
        if (std::fabs(v) < std::fabs(t)) v = std::fabs(t);
      }
      s.Singles[(size_t)i * N + a] = v;
    }
}

// ---------------------------------------------------------------------------
// input.dat grammar: SHCI/input.cpp:35-404
bool readInput(const std::string& path, Schedule& schd, std::string& err, FILE* echo) {
  std::ifstream ifs(path.c_str());
  if (!ifs.good()) {
    err = "input file " + path + " does not exist";
    return false;
  }
  std::stringstream dump;
  dump << ifs.rdbuf();
  int maxiter = -1;
  std::vector<int> sweep_iter;
  std::vector<double> sweep_epsilon;
  int nocc = -1;
  schd.randomSeed = (unsigned int)wallTime();
  if (echo) {
    fprintf(echo, "\n\n**************************************************************\n");
    fprintf(echo, "INPUT FILE\n");
    fprintf(echo, "**************************************************************\n");
  }
  while (dump.good()) {
    std::string Line;
    std::getline(dump, Line);
    Line = trim(Line);
    if (echo) fprintf(echo, "%s\n", Line.c_str());
    std::vector<std::string> tok = split(Line, ", \t\n");
    std::string ArgName = tok[0];
    if (!ArgName.empty() && ArgName[0] == '#') continue;
    if (ArgName.empty()) continue;
    auto arg = [&](size_t k) -> std::string { return k < tok.size() ? tok[k] : std::string(); };
    if (iequals(ArgName, "nocc")) {
      nocc = atoi(arg(1).c_str());
      std::string L;
      std::getline(dump, L);
      L = trim(L);
      std::vector<std::string> t = split(L, ", \t");
      while (!iequals(t[0], "end")) {
        if ((int)t.size() != nocc) {
          err = "nocc: " + std::to_string(nocc) + " neq " + std::to_string(t.size());
          return false;
        }
        std::vector<int> occ(nocc);
        for (size_t i = 0; i < t.size(); i++) {
          occ[i] = atoi(t[i].c_str());
          if (echo) fprintf(echo, "%d ", occ[i]);
        }
        if (echo) fprintf(echo, "\n");
        schd.occupied.push_back(occ);
        if (!dump.good()) {
          err = "unterminated nocc block";
          return false;
        }
        std::getline(dump, L);
        L = trim(L);
        t = split(L, ", \t");
      }
    } else if (iequals(ArgName, "nact")) schd.nact = atoi(arg(1).c_str());
    else if (iequals(ArgName, "ncore")) schd.ncore = atoi(arg(1).c_str());
    else if (iequals(ArgName, "noio")) schd.io = false;
    else if (iequals(ArgName, "dolcc")) schd.doLCC = true;
    else if (iequals(ArgName, "io")) schd.io = true;
    else if (iequals(ArgName, "directdavidson")) schd.davidsonType = 1;
    else if (iequals(ArgName, "diskdavidson")) schd.davidsonType = 2;
    else if (iequals(ArgName, "relaxedRDM")) {}
    else if (iequals(ArgName, "num_thrds")) {}
    else if (iequals(ArgName, "outputlevel")) schd.outputlevel = atoi(arg(1).c_str());
    else if (iequals(ArgName, "extrapolate")) {}
    else if (iequals(ArgName, "dosoc")) schd.doSOC = true;
    else if (iequals(ArgName, "algorithm")) {}
    else if (iequals(ArgName, "doresponse")) {}
    else if (iequals(ArgName, "maxseniority")) {}
    else if (iequals(ArgName, "maxexcitation")) {}
    else if (iequals(ArgName, "SenioAndExc")) {}
    else if (iequals(ArgName, "dogtensor")) {}
    else if (iequals(ArgName, "targetError")) schd.targetError = atof(arg(1).c_str());
    else if (iequals(ArgName, "orbitals")) schd.integralFile = arg(1);
    else if (iequals(ArgName, "dosocqdpt")) {}
    else if (iequals(ArgName, "nptiter")) schd.nPTiter = atoi(arg(1).c_str());
    else if (iequals(ArgName, "epsilon2")) schd.epsilon2 = atof(arg(1).c_str());
    else if (iequals(ArgName, "socmultiplier")) {}
    else if (iequals(ArgName, "seed")) schd.randomSeed = (unsigned int)atof(arg(1).c_str());
    else if (iequals(ArgName, "nroots")) schd.nroots = (int)atof(arg(1).c_str());
    else if (iequals(ArgName, "epsilon2Large")) schd.epsilon2Large = atof(arg(1).c_str());
    else if (iequals(ArgName, "onlyperturbative")) {}
    else if (iequals(ArgName, "printbestdeterminants")) schd.printBestDeterminants = atoi(arg(1).c_str());
    else if (iequals(ArgName, "printalldeterminants")) schd.printAllDeterminants = true;
    else if (iequals(ArgName, "writebestdeterminants")) schd.writeBestDeterminants = atoi(arg(1).c_str());
    else if (iequals(ArgName, "dordm")) schd.DoRDM = true;
    else if (iequals(ArgName, "DoOneRDM")) schd.DoOneRDM = true;
    else if (iequals(ArgName, "DoSpinOneRDM")) {}
    else if (iequals(ArgName, "TransitionRDM")) {}
    else if (iequals(ArgName, "Treversal")) {
      schd.Trev = atoi(arg(1).c_str());
      if (!(schd.Trev == 0 || schd.Trev == 1 || schd.Trev == -1)) {
        err = "Treversal should be either 0, 1, or -1.";
        return false;
      }
    } else if (iequals(ArgName, "dospinrdm")) schd.DoRDM = true;
    else if (iequals(ArgName, "quasiq")) {}
    else if (iequals(ArgName, "nblocks")) schd.nblocks = (int)atof(arg(1).c_str());
    else if (iequals(ArgName, "restrict") || iequals(ArgName, "restrictv")) {  // input.cpp:267-306
      Restriction r;
      r.minElec = atoi(arg(1).c_str());
      r.maxElec = atoi(arg(2).c_str());
      for (size_t i = 3; i < tok.size(); i++) {
        r.orbs.push_back(2 * atoi(tok[i].c_str()));
        r.orbs.push_back(2 * atoi(tok[i].c_str()) + 1);
      }
      schd.restrictionsV.push_back(r);
      if (iequals(ArgName, "restrict")) schd.restrictionsPT.push_back(r);
    } else if (iequals(ArgName, "applyB")) {
      err = "applyB can only be used with ZDice";
      return false;
    } else if (iequals(ArgName, "cdfciTol")) {}
    else if (iequals(ArgName, "davidsonTol")) schd.davidsonTol = atof(arg(1).c_str());
    else if (iequals(ArgName, "davidsonTolLoose")) schd.davidsonTolLoose = atof(arg(1).c_str());
    else if (iequals(ArgName, "excitation")) {}
    else if (iequals(ArgName, "sampleN")) schd.SampleN = atoi(arg(1).c_str());
    else if (iequals(ArgName, "restart")) schd.restart = true;
    else if (iequals(ArgName, "nvirt")) {}
    else if (iequals(ArgName, "deterministic")) schd.stochastic = false;
    else if (iequals(ArgName, "fullrestart")) schd.fullrestart = true;
    else if (iequals(ArgName, "dE")) schd.dE = atof(arg(1).c_str());
    else if (iequals(ArgName, "eps")) {}
    else if (iequals(ArgName, "prefix")) schd.prefix.push_back(arg(1));
    else if (iequals(ArgName, "pointGroup")) {}
    else if (iequals(ArgName, "spin")) {}
    else if (iequals(ArgName, "irrep")) {}
    else if (iequals(ArgName, "searchForLowestEnergyDet")) schd.searchForLowestEnergyDet = true;
    else if (iequals(ArgName, "DoThreeRDM")) {}
    else if (iequals(ArgName, "DoFourRDM")) {}
    else if (iequals(ArgName, "schedule")) {
      std::getline(dump, Line);
      if (echo) fprintf(echo, "%s\n", Line.c_str());
      std::vector<std::string> st = split(Line, " \t");
      while (!iequals(st[0], "END")) {
        if (st[0].empty() || st[0][0] != '#') {
          sweep_iter.push_back(atoi(st[0].c_str()));
          sweep_epsilon.push_back(atof(st.size() > 1 ? st[1].c_str() : ""));
        }
        if (!dump.good()) {
          err = "unterminated schedule block";
          return false;
        }
        std::getline(dump, Line);
        if (echo) fprintf(echo, "%s\n", Line.c_str());
        st = split(Line, " \t");
      }
    } else if (iequals(ArgName, "maxiter")) maxiter = atoi(arg(1).c_str());
    else if (iequals(ArgName, "cdfciIter")) {}
    else if (iequals(ArgName, "cdfciOn")) schd.cdfci_on = atoi(arg(1).c_str());
    else if (iequals(ArgName, "reportInterval")) {}
    else if (iequals(ArgName, "z_threshold")) {}
    else if (iequals(ArgName, "cdfciSample")) {}
    else {
      err = "cannot read option " + ArgName;
      return false;
    }
  }
  if (sweep_iter.empty()) {
    err = "schedule keyword has to be included.";
    return false;
  }
  if (maxiter < sweep_iter.back()) {
    err = "maxiter should be greater than last entry of sweep_iter";
    return false;
  }
  if (nocc == -1) {
    err = "nocc keyword has to be included.";
    return false;
  }
  if (schd.davidsonType == 1) schd.davidsonTolLoose = 3.e-2;
  for (size_t i = 1; i < sweep_iter.size(); i++)
    for (int j = sweep_iter[i - 1]; j < sweep_iter[i]; j++) schd.epsilon1.push_back(sweep_epsilon[i - 1]);
  for (int j = sweep_iter.back(); j < maxiter; j++) schd.epsilon1.push_back(sweep_epsilon.back());
  if (schd.prefix.empty()) schd.prefix.push_back(".");
  return true;
}

// ---------------------------------------------------------------------------
// determinant primitives (SHCI/Determinants.h)
static inline int bitCount(int64_t x) { return __builtin_popcountll((uint64_t)x); }
static inline bool getocc(const Det& d, int i) { return ((d.r[i / 64] >> (i % 64)) & 1) != 0; }
static inline void setocc(Det& d, int i, bool occ) {
  int64_t one = 1;
  if (occ)
    d.r[i / 64] |= (one << (i % 64));
  else
    d.r[i / 64] &= ~(one << (i % 64));
}

bool detLess(const Det& a, const Det& b, int W) {  // Determinants.h:409-417 (signed words, high first)
  for (int i = W - 1; i >= 0; i--) {
    if (a.r[i] < b.r[i]) return true;
    if (a.r[i] > b.r[i]) return false;
  }
  return false;
}
bool detEq(const Det& a, const Det& b, int W) {
  for (int i = W - 1; i >= 0; i--)
    if (a.r[i] != b.r[i]) return false;
  return true;
}
int excitationDistance(const Det& a, const Det& b, int W) {  // Determinants.h:381-387
  int nd = 0;
  for (int i = 0; i < W; i++) nd += bitCount(a.r[i] ^ b.r[i]);
  return nd / 2;
}

// parity(start,end): sign of the occupied spin-orbitals strictly between
// start and end (Determinants.h:202-220)
static void parity(const Det& d, int start, int end, double& par) {
  uint64_t one = 1;
  uint64_t mask = (one << (start % 64)) - one;
  int nz = -bitCount(d.r[start / 64] & (int64_t)mask);
  for (int i = start / 64; i < end / 64; i++) nz += bitCount(d.r[i]);
  mask = (one << (end % 64)) - one;
  nz += bitCount(d.r[end / 64] & (int64_t)mask);
  par *= (-2. * (nz % 2) + 1);
  if (getocc(d, start)) par *= -1.;
}

static void parityDouble(Det d, int i, int j, int a, int b, double& sgn);

static void occList(const Det& d, int W, int* closed, int& nclosed) {
  nclosed = 0;
  for (int i = 0; i < W; i++) {
    uint64_t b = (uint64_t)d.r[i];
    while (b) {
      int pos = __builtin_ctzll(b);
      closed[nclosed++] = i * 64 + pos;
      b &= b - 1;
    }
  }
}

// Determinant::Energy, Determinants.cpp:252-300 (coreE = 0 during the var step)
double energy(const System& s, const Det& d) {
  int closed[128], nc;
  occList(d, s.W, closed, nc);
  double e = 0.0;
  for (int i = 0; i < nc; i++) {
    int I = closed[i];
    e += s.i1(I, I);
    for (int j = i + 1; j < nc; j++) {
      int J = closed[j];
      e += s.Direct[(size_t)(I / 2) * s.norb + J / 2];
      if ((I % 2) == (J % 2)) e -= s.Exchange[(size_t)(I / 2) * s.norb + J / 2];
    }
  }
  return e;
}

// Determinant::Hij_1Excite (member), Determinants.cpp:472-505
static double hij1Member(const System& s, const Det& ket, int a, int i) {
  double sgn = 1.0;
  parity(ket, std::min(a, i), std::max(a, i), sgn);
  double e = s.i1(a, i);
  int closed[128], nc;
  occList(ket, s.W, closed, nc);
  for (int k = 0; k < nc; k++) {
    int j = closed[k];
    e += (s.i2(a, i, j, j) - s.i2(a, j, j, i));
  }
  e *= sgn;
  return e;
}

// free Hij_1Excite, Determinants.cpp:437-469 (used by the screening)
static double hij1Free(const System& s, int a, int i, const int* closed, int nclosed) {
  double sgn = 1.0;
  double e = s.i1(a, i);
  for (int j = 0; j < nclosed; j++) {
    if (closed[j] > std::min(i, a) && closed[j] < std::max(i, a)) sgn *= -1.;
    e += (s.i2(a, i, closed[j], closed[j]) - s.i2(a, closed[j], closed[j], i));
  }
  return e * sgn;
}

// Determinant::Hij_2Excite, Determinants.cpp:406-434
static double hij2(const System& s, const Det& ket, int i, int j, int a, int b) {
  double sgn = 1.0;
  int I = std::min(i, j), J = std::max(i, j), A = std::min(a, b), B = std::max(a, b);
  parity(ket, std::min(I, A), std::max(I, A), sgn);
  parity(ket, std::min(J, B), std::max(J, B), sgn);
  if (A > J || B < I) sgn *= -1.;
  return sgn * (s.i2(A, I, B, J) - s.i2(A, J, B, I));
}

// Hij(bra, ket), Determinants.cpp:564-631
double hij(const System& s, const Det& bra, const Det& ket) {
  int cre[4], des[4], ncre = 0, ndes = 0;
  for (int i = 0; i < s.W; i++) {
    uint64_t u = (uint64_t)(bra.r[i] ^ ket.r[i]);
    uint64_t b = u & (uint64_t)bra.r[i];
    uint64_t k = u & (uint64_t)ket.r[i];
    while (b) {
      if (ncre < 4) cre[ncre] = __builtin_ctzll(b) + i * 64;
      ncre++;
      b &= b - 1;
    }
    while (k) {
      if (ndes < 4) des[ndes] = __builtin_ctzll(k) + i * 64;
      ndes++;
      k &= k - 1;
    }
  }
  if (ncre == 1) return hij1Member(s, ket, cre[0], des[0]);
  if (ncre == 2) return hij2(s, ket, des[0], des[1], cre[0], cre[1]);
  return 0.0;
}

// ---------------------------------------------------------------------------
// time reversal (Determinants.h:245-301, Determinants.cpp:32-121,
// SHCImakeHamiltonian.cpp:411-483)
static const uint64_t kEvenBits = 0x5555555555555555ull, kOddBits = 0xAAAAAAAAAAAAAAAAull;
static Det flipAlphaBeta(const Det& d, int W) {
  Det f;
  for (int i = 0; i < W; i++) {
    const uint64_t r = (uint64_t)d.r[i];
    f.r[i] = (int64_t)(((r & kEvenBits) << 1) + ((r & kOddBits) >> 1));
  }
  return f;
}
static bool hasUnpairedElectrons(const Det& d, int W) {
  for (int i = W - 1; i >= 0; i--) {
    const uint64_t r = (uint64_t)d.r[i];
    if (((r & kEvenBits) << 1) != (r & kOddBits)) return true;
  }
  return false;
}
// the reference compares each word with its flip as unsigned 64-bit values,
// from the highest word down
static bool isStandard(const Det& d, int W) {
  if (!hasUnpairedElectrons(d, W)) return true;
  for (int i = W - 1; i >= 0; i--) {
    const uint64_t r = (uint64_t)d.r[i], f = ((r & kEvenBits) << 1) + ((r & kOddBits) >> 1);
    if (r < f) return false;
    if (r > f) return true;
  }
  return true;
}
static void makeStandard(Det& d, int W) {
  if (!isStandard(d, W)) d = flipAlphaBeta(d, W);
}
static double parityOfFlipAlphaBeta(const Det& d, int W) {
  int numpaired = 0, nbeta = 0;
  for (int i = 0; i < W; i++) {
    const uint64_t r = (uint64_t)d.r[i];
    numpaired += bitCount((int64_t)(((r & kEvenBits) << 1) & r));
    nbeta += bitCount((int64_t)(r & kOddBits));
  }
  const double p = numpaired % 2 == 0 ? 1.0 : -1.0;
  return nbeta % 2 == 0 ? p : -p;
}
static bool connectedDets(const Det& a, const Det& b, int W) {  // excitation distance <= 2
  int nd = 0;
  for (int i = 0; i < W; i++) nd += bitCount(a.r[i] ^ b.r[i]);
  return nd <= 4;
}

// updateHijForTReversal(hij, dj, dk), Determinants.cpp:32-121: the element
// between the time-reversal adapted functions of dj and dk (as the reference
// writes it, without the Treversal sign, which enters only when the state is
// unpacked)
static void updateHijForTReversal(const System& s, double& h, const Det& dj, const Det& dk) {
  const int W = s.W;
  const bool uj = hasUnpairedElectrons(dj, W), uk = hasUnpairedElectrons(dk, W);
  if (!uj && uk) {
    const Det c = flipAlphaBeta(dk, W);
    if (!connectedDets(c, dj, W)) return;
    const double parity = parityOfFlipAlphaBeta(dk, W);
    const double hc = hij(s, dj, c);
    h = (h + parity * hc) / std::pow(2., 0.5);
  } else if (uj && !uk) {
    const Det c = flipAlphaBeta(dj, W);
    if (!connectedDets(c, dk, W)) return;
    const double parity = parityOfFlipAlphaBeta(dj, W);
    const double hc = hij(s, c, dk);
    h = (h + parity * hc) / std::pow(2., 0.5);
  } else if (uj && uk) {
    const Det ck = flipAlphaBeta(dk, W);
    if (!connectedDets(ck, dj, W)) return;
    const double parityk = parityOfFlipAlphaBeta(dk, W);
    const double hc = hij(s, dj, ck);
    h = h + parityk * hc;
  }
}

// fixForTreversal (SHCImakeHamiltonian.cpp:411-483) on the element hij =
// Hij(D_|J|, D_|I|) of a pair found through the helpers, where a negative
// index stands for the flipped copy of an open-shell determinant
static void fixForTreversal(const System& s, const std::vector<Det>& dets, int DetI, int DetJ, double& h) {
  const int W = s.W;
  if (DetI < 0) {
    h = 0.0;
    return;
  }
  const Det& dJ = dets[std::abs(DetJ) - 1];
  const Det& dI = dets[std::abs(DetI) - 1];
  const bool uJ = hasUnpairedElectrons(dJ, W), uI = hasUnpairedElectrons(dI, W);
  if (!uJ && !uI) return;
  if (!uJ || DetJ > 0) {
    updateHijForTReversal(s, h, dJ, dI);
    return;
  }
  if (connectedDets(dJ, dI, W)) {
    h = 0.0;  // the pair is also found through D_J itself
    return;
  }
  const Det fj = flipAlphaBeta(dJ, W);
  double parity = parityOfFlipAlphaBeta(dJ, W);
  h = parity * hij(s, fj, dI);
  if (uI) {
    const Det fi = flipAlphaBeta(dI, W);
    parity = parityOfFlipAlphaBeta(dI, W);
    h += parity * hij(s, dJ, fi);
    h /= 2.0;
  } else {
    h /= std::pow(2.0, 0.5);
  }
}

// unpackTrevState, SHCIbasics.cpp:730-776: every open-shell determinant's
// flipped partner is appended (in determinant order) with coefficient
// Treversal * parity * c / sqrt(2), and c becomes c / sqrt(2)
static void unpackTrevState(std::vector<Det>& dets, std::vector<std::vector<double>>& ci, int W, int trev) {
  const size_t oldLen = dets.size();
  std::vector<Det> add;
  std::vector<size_t> src;
  for (size_t i = 0; i < oldLen; i++)
    if (hasUnpairedElectrons(dets[i], W)) {
      add.push_back(flipAlphaBeta(dets[i], W));
      src.push_back(i);
    }
  dets.insert(dets.end(), add.begin(), add.end());
  for (auto& c : ci) {
    c.resize(dets.size());
    for (size_t k = 0; k < src.size(); k++) {
      const size_t i = src[k];
      const double v = c[i];
      c[i] = v / std::sqrt(2.0);
      c[oldLen + k] = trev * parityOfFlipAlphaBeta(dets[i], W) * v / std::sqrt(2.0);
    }
  }
}

// ---------------------------------------------------------------------------
// OccRestrictions.cpp
static bool oneElecAllowed(const Restriction& r, int cur, int i, int a) {
  int elec = cur;
  for (int x : r.orbs) {
    if (x == i) elec--;
    if (x == a) elec++;
  }
  return !(elec < r.minElec || elec > r.maxElec);
}
static bool twoElecAllowed(const Restriction& r, int cur, int i, int j, int a, int b) {
  int elec = cur;
  for (int x : r.orbs) {
    if (x == i || x == j) elec--;
    if (x == a || x == b) elec++;
  }
  return !(elec < r.minElec || elec > r.maxElec);
}

// getDeterminantsVariationalApprox, SHCIgetdeterminants.cpp:586-739
void screenDet(const System& s, const Schedule& schd, const Det& d, double epsilon,
               const Det* sorted, size_t nsorted, std::vector<Det>& out) {
  int norbs = s.nso;
  int closed[128], open[128], nclosed = 0, nopen = 0;
  for (int i = 0; i < norbs; i++) {
    if (getocc(d, i))
      closed[nclosed++] = i;
    else
      open[nopen++] = i;
  }
  std::vector<int> cur(schd.restrictionsV.size());
  for (size_t r = 0; r < schd.restrictionsV.size(); r++) {
    int c = 0;
    for (int k = 0; k < nclosed; k++)
      for (int x : schd.restrictionsV[r].orbs)
        if (closed[k] == x) c++;
    cur[r] = c;
  }
  int W = s.W;
  auto present = [&](const Det& di) {
    return std::binary_search(sorted, sorted + nsorted, di,
                              [W](const Det& x, const Det& y) { return detLess(x, y, W); });
  };
  for (int ia = 0; ia < nopen * nclosed; ia++) {
    int i = ia / nopen, a = ia % nopen;
    if (closed[i] / 2 < schd.ncore || open[a] / 2 >= schd.ncore + schd.nact) continue;
    bool ok = true;
    for (size_t r = 0; r < schd.restrictionsV.size(); r++)
      if (!oneElecAllowed(schd.restrictionsV[r], cur[r], closed[i], open[a])) ok = false;
    if (!ok) continue;
    double integral = s.Singles[(size_t)open[a] * s.nso + closed[i]];
    if (std::fabs(integral) > epsilon)
      if (closed[i] % 2 == open[a] % 2) integral = hij1Free(s, open[a], closed[i], closed, nclosed);
    if (std::fabs(integral) > epsilon) {
      Det di = d;
      setocc(di, open[a], true);
      setocc(di, closed[i], false);
      if (!present(di)) out.push_back(di);
    }
  }
  if (std::fabs(s.maxEntry) < epsilon) return;
  for (int ij = 0; ij < nclosed * nclosed; ij++) {
    int i = ij / nclosed, j = ij % nclosed;
    if (i <= j) continue;
    int I = closed[i] / 2, J = closed[j] / 2;
    int X = std::max(I, J), Y = std::min(I, J);
    if (closed[i] / 2 < schd.ncore || closed[j] / 2 < schd.ncore) continue;
    int pairIndex = X * (X + 1) / 2 + Y;
    bool sameSpin = closed[i] % 2 == closed[j] % 2;
    size_t start = sameSpin ? s.hbSameStart[pairIndex] : s.hbOppStart[pairIndex];
    size_t end = sameSpin ? s.hbSameStart[pairIndex + 1] : s.hbOppStart[pairIndex + 1];
    const float* integrals = sameSpin ? s.hbSameVal.data() : s.hbOppVal.data();
    const int16_t* orbIndices = sameSpin ? s.hbSamePair.data() : s.hbOppPair.data();
    for (size_t index = start; index < end; index++) {
      if (std::fabs(integrals[index]) < epsilon) break;
      int a = 2 * orbIndices[2 * index] + closed[i] % 2, b = 2 * orbIndices[2 * index + 1] + closed[j] % 2;
      if (a / 2 >= schd.ncore + schd.nact || b / 2 >= schd.ncore + schd.nact) continue;
      bool ok = true;
      for (size_t r = 0; r < schd.restrictionsV.size(); r++)
        if (!twoElecAllowed(schd.restrictionsV[r], cur[r], closed[i], closed[j], a, b)) ok = false;
      if (!ok) continue;
      if (!(getocc(d, a) || getocc(d, b))) {
        Det di = d;
        setocc(di, a, true);
        setocc(di, b, true);
        setocc(di, closed[i], false);
        setocc(di, closed[j], false);
        if (!present(di)) out.push_back(di);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// deterministic EN-PT2

// Determinant::parity(i, j, a, b, sgn), Determinants.cpp:332-357: sign of
// i -> a then j -> b applied to d
static void parityDouble(Det d, int i, int j, int a, int b, double& sgn) {
  parity(d, std::min(i, a), std::max(i, a), sgn);
  setocc(d, i, false);
  setocc(d, a, true);
  parity(d, std::min(j, b), std::max(j, b), sgn);
}

// getDeterminantsDeterministicPT, SHCIgetdeterminants.cpp:40-171: candidates
// of one source with numerator H_ai c_i (singles: Hij_1Excite, doubles: the
// float heat-bath value times the parity of the excitation)
static void ptCandidates(const System& s, const Schedule& schd, const Det& d, double epsilon, double ci1,
                         std::vector<Det>& dets, std::vector<double>& num, std::vector<char>* isDouble = nullptr) {
  int closed[128], open[128], nclosed = 0, nopen = 0;
  for (int i = 0; i < s.nso; i++) {
    if (getocc(d, i))
      closed[nclosed++] = i;
    else
      open[nopen++] = i;
  }
  std::vector<int> cur(schd.restrictionsPT.size());
  for (size_t r = 0; r < schd.restrictionsPT.size(); r++) {
    int c = 0;
    for (int k = 0; k < nclosed; k++)
      for (int x : schd.restrictionsPT[r].orbs)
        if (closed[k] == x) c++;
    cur[r] = c;
  }
  for (int ia = 0; ia < nopen * nclosed; ia++) {
    int i = ia / nopen, a = ia % nopen;
    bool ok = true;
    for (size_t r = 0; r < schd.restrictionsPT.size(); r++)
      if (!oneElecAllowed(schd.restrictionsPT[r], cur[r], closed[i], open[a])) ok = false;
    if (!ok) continue;
    // (across spins only the one-electron term survives: I1 times the parity)
    double integral = hij1Free(s, open[a], closed[i], closed, nclosed);
    if (std::fabs(integral) > epsilon) {
      Det di = d;
      setocc(di, open[a], true);
      setocc(di, closed[i], false);
      dets.push_back(di);
      num.push_back(integral * ci1);
      if (isDouble) isDouble->push_back(0);
    }
  }
  if (std::fabs(s.maxEntry) < epsilon) return;
  for (int ij = 0; ij < nclosed * nclosed; ij++) {
    int i = ij / nclosed, j = ij % nclosed;
    if (i <= j) continue;
    int I = closed[i] / 2, J = closed[j] / 2;
    int X = std::max(I, J), Y = std::min(I, J);
    int pairIndex = X * (X + 1) / 2 + Y;
    bool sameSpin = closed[i] % 2 == closed[j] % 2;
    size_t start = sameSpin ? s.hbSameStart[pairIndex] : s.hbOppStart[pairIndex];
    size_t end = sameSpin ? s.hbSameStart[pairIndex + 1] : s.hbOppStart[pairIndex + 1];
    const float* integrals = sameSpin ? s.hbSameVal.data() : s.hbOppVal.data();
    const int16_t* orbIndices = sameSpin ? s.hbSamePair.data() : s.hbOppPair.data();
    for (size_t index = start; index < end; index++) {
      if (std::fabs(integrals[index]) < epsilon) break;
      int a = 2 * orbIndices[2 * index] + closed[i] % 2, b = 2 * orbIndices[2 * index + 1] + closed[j] % 2;
      bool ok = true;
      for (size_t r = 0; r < schd.restrictionsPT.size(); r++)
        if (!twoElecAllowed(schd.restrictionsPT[r], cur[r], closed[i], closed[j], a, b)) ok = false;
      if (!ok) continue;
      if (!(getocc(d, a) || getocc(d, b))) {
        Det di = d;
        setocc(di, a, true);
        setocc(di, b, true);
        setocc(di, closed[i], false);
        setocc(di, closed[j], false);
        double sgn = 1.0;
        parityDouble(di, a, b, closed[i], closed[j], sgn);
        dets.push_back(di);
        num.push_back((double)integrals[index] * sgn * ci1);
        if (isDouble) isDouble->push_back(1);
      }
    }
  }
}

double ptDeterministic(const System& s, const Schedule& schd, const std::vector<Det>& dets, const double* ci,
                       double E0, double eps2, double* psi1Norm, size_t* nExternal) {
  const int W = s.W;
  const size_t N = dets.size();
  // sources in contiguous blocks per thread, concatenated in source order
  const int nt = std::max(1, std::min<int>(oracleThreads(), (int)std::max<size_t>(N, 1)));
  std::vector<std::vector<Det>> cd(nt);
  std::vector<std::vector<double>> cn(nt);
#pragma omp parallel for num_threads(nt) schedule(static, 1)
  for (int t = 0; t < nt; t++) {
    const size_t b0 = N * t / nt, b1 = N * (t + 1) / nt;
    for (size_t i = b0; i < b1; i++)
      ptCandidates(s, schd, dets[i], std::fabs(eps2 / ci[i]), ci[i], cd[t], cn[t]);
  }
  std::vector<Det> cand;
  std::vector<double> num;
  for (int t = 0; t < nt; t++) {
    cand.insert(cand.end(), cd[t].begin(), cd[t].end());
    num.insert(num.end(), cn[t].begin(), cn[t].end());
  }
  // MergeSortAndRemoveDuplicates (SHCISortMpiUtils.cpp:541-686): stable
  // sort, numerators of equal determinants summed in generation order
  std::vector<size_t> order(cand.size());
  for (size_t k = 0; k < order.size(); k++) order[k] = k;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t x, size_t y) { return detLess(cand[x], cand[y], W); });
  std::vector<Det> sortedVar = dets;
  std::sort(sortedVar.begin(), sortedVar.end(), [W](const Det& x, const Det& y) { return detLess(x, y, W); });
  double pt = 0.0, norm = 0.0;
  size_t next = 0;
  for (size_t k = 0; k < order.size();) {
    const Det& d = cand[order[k]];
    double sum = num[order[k]];
    size_t k1 = k + 1;
    while (k1 < order.size() && detEq(cand[order[k1]], d, W)) sum += num[order[k1++]];
    k = k1;
    // RemoveDetsPresentIn (SHCISortMpiUtils.cpp:750-860)
    if (std::binary_search(sortedVar.begin(), sortedVar.end(), d,
                           [W](const Det& x, const Det& y) { return detLess(x, y, W); }))
      continue;
    const double Ea = energy(s, d);
    pt += std::pow(std::fabs(sum), 2) / (E0 - Ea);
    norm += std::pow(std::fabs(sum / (E0 - Ea)), 2);
    next++;
  }
  if (psi1Norm) *psi1Norm = norm;
  if (nExternal) *nExternal = next;
  return pt;
}

// alias tables for sampling i with probability |c_i| / cumulative
// (SHCIsampledeterminants.cpp:58-78)
static void aliasSetup(const double* ci, int n, double cumulative, std::vector<int>& alias, std::vector<double>& prob) {
  alias.assign(n, 0);
  prob.assign(n, 0.0);
  std::vector<double> larger, smaller;  // (indices held as doubles, as there)
  for (int i = 0; i < n; i++) {
    prob[i] = std::fabs(ci[i]) * n / cumulative;
    if (prob[i] < 1.0)
      smaller.push_back(i);
    else
      larger.push_back(i);
  }
  while (larger.size() > 0 && smaller.size() > 0) {
    int l = (int)larger.back();
    larger.pop_back();
    int sm = (int)smaller.back();
    smaller.pop_back();
    alias[sm] = l;
    prob[l] = prob[l] - (1.0 - prob[sm]);
    if (prob[l] < 1.0)
      smaller.push_back(l);
    else
      larger.push_back(l);
  }
}

// The determinant index of one alias draw from a raw rand() value r
// (SHCIsampledeterminants.cpp:160): floor(r / RAND_MAX * n), clamped to
// n - 1.  The reference leaves r == RAND_MAX at index n, one past its alias
// and probability tables; the clamp keeps every other draw unchanged.
int aliasDrawIndex(int r, int n) {
  const int i = (int)std::floor(1. * ((double)r / (RAND_MAX)) * n);
  return i < n ? i : n - 1;
}

// sample_N2_alias (SHCIsampledeterminants.cpp:148-176): Nmc draws with
// replacement; distinct determinants in first-drawn order with weights
// cumulative * sign(c) * (times drawn) / Nmc; returns the distinct count
static int sampleAlias(const double* ci, int n, double cumulative, std::vector<int>& sample, std::vector<double>& wts,
                       const std::vector<int>& alias, const std::vector<double>& prob) {
  const int niter = (int)sample.size();
  int sampleIndex = 0;
  for (int index = 0; index < niter; index++) {
    int detIndex = aliasDrawIndex(rand(), n);
    double rand_no = ((double)rand() / (RAND_MAX));
    if (rand_no >= prob[detIndex]) detIndex = alias[detIndex];
    auto it = std::find(sample.begin(), sample.end(), detIndex);
    if (it == sample.end()) {
      sample[sampleIndex] = detIndex;
      wts[sampleIndex] = cumulative * ci[detIndex] / std::fabs(ci[detIndex]);
      sampleIndex++;
    } else {
      wts[std::distance(sample.begin(), it)] += cumulative * ci[detIndex] / std::fabs(ci[detIndex]);
    }
  }
  for (int i = 0; i < niter; i++) wts[i] /= niter;
  return sampleIndex;
}

// getDeterminantsStochastic2Epsilon (SHCIgetdeterminants.cpp:889-1046):
// the deterministic candidates of D with w = the sample weight in place of c
// in num1 = H_ai w, num2 = H_ai^2 w (w Nmc/(Nmc-1) - c), and present = the
// term passes epsLarge
static void ptStochCandidates(const System& s, const Schedule& schd, const Det& d, double eps, double epsLarge,
                              double w, double c, int Nmc, std::vector<Det>& dets, std::vector<double>& num1,
                              std::vector<double>& num2, std::vector<char>& present) {
  std::vector<Det> cd;
  std::vector<double> v;
  std::vector<char> dbl;
  ptCandidates(s, schd, d, eps, 1.0, cd, v, &dbl);  // v = H_ai (c = 1)
  const double Nmcd = 1. * Nmc;
  for (size_t k = 0; k < cd.size(); k++) {
    dets.push_back(cd[k]);
    num1.push_back(v[k] * w);
    // (a double's value is a float, squared in float arithmetic there)
    const double v2 = dbl[k] ? (double)((float)std::fabs(v[k]) * (float)std::fabs(v[k])) : v[k] * v[k];
    num2.push_back(v2 * w * (w * Nmcd / (Nmcd - 1) - c));
    present.push_back(std::fabs(v[k]) > epsLarge);
  }
}

double ptSemistochastic(const System& s, const Schedule& schd, const std::vector<Det>& dets, const double* ci,
                        double E0, double coreE, int root, FILE* log, double startofCalc, double* error, int* iters,
                        std::vector<double>* perIter) {
  if (schd.nPTiter == 0) return 0.0;
  const int W = s.W;
  const int DetsSize = (int)dets.size();
  const double E0c = E0 + coreE;  // the reference's E0 holds the core energy
  if (log) fprintf(log, "Performing (semi)stochastic PT for state: %3d\n", root);
  // the deterministic part always runs (only the message depends on the
  // threshold, SHCIbasics.cpp:69-75)
  if (log && schd.epsilon2Large < 999) fprintf(log, "\n1/ Deterministic calculation with epsilon2=%g\n", schd.epsilon2Large);
  const double EptLarge = ptDeterministic(s, schd, dets, ci, E0, schd.epsilon2Large);
  if (log) {  // DoPerturbativeDeterministic's own lines
    fprintf(log, "Deterministic PT calculation converged\n");
    fprintf(log, "PTEnergy: %18.10f\n", E0c + EptLarge);
    fprintf(log, "Time(s):  %10.2f\n", wallTime() - startofCalc);
    fprintf(log, "\n2/ Stochastic calculation with epsilon2=%g\n", schd.epsilon2);
  }
  std::vector<Det> sortedVar = dets;
  std::sort(sortedVar.begin(), sortedVar.end(), [W](const Det& x, const Det& y) { return detLess(x, y, W); });
  const int Nmc = schd.SampleN;
  double cumulative = 0.0;
  for (int i = 0; i < DetsSize; i++) cumulative += std::fabs(ci[i]);
  std::vector<int> alias;
  std::vector<double> prob;
  aliasSetup(ci, DetsSize, cumulative, alias, prob);
  std::vector<int> allSample(Nmc, -1);
  std::vector<double> allwts(Nmc, 0.);
  if (log) fprintf(log, "%6s  %18s  %5s %18s %10s  %10s\n", "Iter", "EPTcurrent", "State", "EPTavg", "Error", "Time(s)");
  double Avg = 0.0, Avg2 = 0.0, stddev = 0.0;
  int currentIter = 0;
  for (int iter = 0; iter < schd.nPTiter; iter++) {
    std::fill(allSample.begin(), allSample.end(), -1);
    const int distinct = sampleAlias(ci, DetsSize, cumulative, allSample, allwts, alias, prob);
    std::vector<Det> cand;
    std::vector<double> n1, n2;
    std::vector<char> pres;
    for (int k = 0; k < distinct; k++) {
      const int I = allSample[k];
      ptStochCandidates(s, schd, dets[I], schd.epsilon2 / std::fabs(ci[I]), schd.epsilon2Large / std::fabs(ci[I]),
                        allwts[k], ci[I], Nmc, cand, n1, n2, pres);
    }
    // MergeSort (stable), then one pass over the runs of equal determinants
    // that are not variational (SHCIbasics.cpp:264-385)
    std::vector<size_t> order(cand.size());
    for (size_t k = 0; k < order.size(); k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return detLess(cand[x], cand[y], W); });
    double energyEN = 0.0, energyENLarge = 0.0;
    for (size_t k = 0; k < order.size();) {
      const Det& d = cand[order[k]];
      double a1 = 0.0, a2 = 0.0, b1 = 0.0, b2 = 0.0;
      size_t k1 = k;
      for (; k1 < order.size() && detEq(cand[order[k1]], d, W); k1++) {
        a1 += n1[order[k1]];
        a2 += n2[order[k1]];
        if (pres[order[k1]]) {
          b1 += n1[order[k1]];
          b2 += n2[order[k1]];
        }
      }
      k = k1;
      if (std::binary_search(sortedVar.begin(), sortedVar.end(), d,
                             [W](const Det& x, const Det& y) { return detLess(x, y, W); }))
        continue;
      const double den = energy(s, d) - E0;
      energyEN += (std::pow(std::fabs(a1), 2) * Nmc / (Nmc - 1) - a2) / den;
      energyENLarge += (std::pow(std::fabs(b1), 2) * Nmc / (Nmc - 1) - b2) / den;
    }
    currentIter++;
    const double est = -energyEN + energyENLarge + EptLarge;
    if (perIter) perIter->push_back(est);
    Avg += est;
    Avg2 += std::pow(est, 2);
    stddev = currentIter < 5 ? 1e4
                             : std::pow((currentIter * Avg2 - std::pow(Avg, 2)) / currentIter / (currentIter - 1) /
                                            currentIter,
                                        0.5);
    if (log) {
      if (currentIter < 5)
        fprintf(log, "%6d  %18.10f  %5d %18.10f %10s  %10.2f\n", currentIter, E0c + est, root,
                E0c + Avg / currentIter, "--", wallTime() - startofCalc);
      else
        fprintf(log, "%6d  %18.10f  %5d %18.10f %10.2e  %10.2f\n", currentIter, E0c + est, root,
                E0c + Avg / currentIter, stddev, wallTime() - startofCalc);
    }
    if (stddev < schd.targetError) {
      Avg /= currentIter;
      if (log) {
        fprintf(log, "Semistochastic PT calculation converged\n");
        fprintf(log, "PTEnergy: %18.10f +/- %8.2e\n", E0c + Avg, stddev);
        fprintf(log, "Time(s):  %10.2f\n", wallTime() - startofCalc);
      }
      break;
    }
  }
  if (error) *error = stddev;
  if (iters) *iters = currentIter;
  return Avg;
}

// ---------------------------------------------------------------------------
// helpers: SHCImakeHamiltonian.cpp:184-408
static uint64_t alphaString(const Det& d, int W) {
  uint64_t s = 0;
  for (int i = 0; i < W; i++)
    for (int j = 0; j < 32; j++)
      if ((d.r[i] >> (2 * j)) & 1) s |= (uint64_t)1 << (i * 32 + j);
  return s;
}
static uint64_t betaString(const Det& d, int W) {
  uint64_t s = 0;
  for (int i = 0; i < W; i++)
    for (int j = 0; j < 32; j++)
      if ((d.r[i] >> (2 * j + 1)) & 1) s |= (uint64_t)1 << (i * 32 + j);
  return s;
}

static int findOrInsert(std::map<uint64_t, int>& N, uint64_t str, std::vector<std::vector<int>>& major,
                        std::vector<std::vector<int>>& majorDet, std::vector<std::vector<int>>& singles,
                        int norb) {
  auto it = N.find(str);
  if (it != N.end()) return it->second;
  int idx = (int)majorDet.size();
  N.insert(std::make_pair(str, idx));
  major.resize(idx + 1);
  majorDet.resize(idx + 1);
  singles.resize(idx + 1);
  for (int j = 0; j < norb; j++) {
    if (!((str >> j) & 1)) continue;
    for (int k = 0; k < norb; k++) {
      if ((str >> k) & 1) continue;
      uint64_t c = (str & ~((uint64_t)1 << j)) | ((uint64_t)1 << k);
      auto f = N.find(c);
      if (f != N.end()) {
        singles[idx].push_back(f->second);
        singles[f->second].push_back(idx);
      }
    }
  }
  return idx;
}

void populateHelpers(const System& s, Helpers& h, const std::vector<Det>& dets, int start) {
  for (int i = start; i < (int)dets.size(); i++) {
    uint64_t da = alphaString(dets[i], s.W), db = betaString(dets[i], s.W);
    int ib = findOrInsert(h.BetaN, db, h.BetaMajorToAlpha, h.BetaMajorToDet, h.SinglesFromBeta, s.norb);
    int ia = findOrInsert(h.AlphaN, da, h.AlphaMajorToBeta, h.AlphaMajorToDet, h.SinglesFromAlpha, s.norb);
    h.AlphaMajorToBeta[ia].push_back(ib);
    h.AlphaMajorToDet[ia].push_back(i + 1);
    h.BetaMajorToAlpha[ib].push_back(ia);
    h.BetaMajorToDet[ib].push_back(i + 1);
    if (s.trev != 0 && hasUnpairedElectrons(dets[i], s.W)) {
      // the flipped copy, under index -(i+1) (SHCImakeHamiltonian.cpp:344-351)
      int fb = findOrInsert(h.BetaN, da, h.BetaMajorToAlpha, h.BetaMajorToDet, h.SinglesFromBeta, s.norb);
      int fa = findOrInsert(h.AlphaN, db, h.AlphaMajorToBeta, h.AlphaMajorToDet, h.SinglesFromAlpha, s.norb);
      h.AlphaMajorToBeta[fa].push_back(fb);
      h.AlphaMajorToDet[fa].push_back(-(i + 1));
      h.BetaMajorToAlpha[fb].push_back(fa);
      h.BetaMajorToDet[fb].push_back(-(i + 1));
    }
  }
  auto sortMajor = [](std::vector<int>& key, std::vector<int>& det) {
    std::vector<std::pair<int, int>> p(key.size());
    for (size_t k = 0; k < key.size(); k++) p[k] = std::make_pair(key[k], det[k]);
    std::sort(p.begin(), p.end());
    for (size_t k = 0; k < key.size(); k++) {
      key[k] = p[k].first;
      det[k] = p[k].second;
    }
  };
  for (size_t i = 0; i < h.AlphaMajorToBeta.size(); i++) {
    sortMajor(h.AlphaMajorToBeta[i], h.AlphaMajorToDet[i]);
    std::sort(h.SinglesFromAlpha[i].begin(), h.SinglesFromAlpha[i].end());
  }
  for (size_t i = 0; i < h.BetaMajorToAlpha.size(); i++) {
    sortMajor(h.BetaMajorToAlpha[i], h.BetaMajorToDet[i]);
    std::sort(h.SinglesFromBeta[i].begin(), h.SinglesFromBeta[i].end());
  }
}

static int binarySearch(const int* arr, int l, int r, int x) {  // SHCISortMpiUtils.cpp:57-76
  while (r >= l) {
    int mid = l + (r - l) / 2;
    if (arr[mid] == x) return mid;
    if (arr[mid] > x)
      r = mid - 1;
    else
      l = mid + 1;
  }
  return -1;
}

// MakeHfromSMHelpers2, SHCImakeHamiltonian.cpp:486-715 (serial: nprocs = 1)
void makeHRows(const System& s, const Helpers& h, const std::vector<Det>& dets, int start, SparseHam& H) {
  int end = (int)dets.size();
  for (int k = start; k < end; k++) {
    H.connections.push_back(std::vector<int>(1, k));
    double e = energy(s, dets[k]);
    if (s.trev != 0) updateHijForTReversal(s, e, dets[k], dets[k]);
    H.Helements.push_back(std::vector<double>(1, e));
  }
  // (with time reversal, DetJ < 0 is the flipped copy of determinant -DetJ)
  auto add = [&](int DetI, int DetJ) {
    double v = hij(s, dets[std::abs(DetJ) - 1], dets[DetI - 1]);
    if (s.trev != 0) fixForTreversal(s, dets, DetI, DetJ, v);
    H.connections[DetI - 1].push_back(std::abs(DetJ) - 1);
    H.Helements[DetI - 1].push_back(v);
  };
  auto doubleApart = [&](int DetI, int DetJ) {
    Det di = dets[DetI - 1];
    if (DetJ < 0) di = flipAlphaBeta(di, s.W);
    return excitationDistance(dets[std::abs(DetJ) - 1], di, s.W) == 2;
  };
  // each determinant belongs to one alpha string, so the rows written by
  // different alpha strings are disjoint
  const int nt = oracleThreads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt) if (nt > 1)
  for (size_t i = 0; i < h.AlphaMajorToBeta.size(); i++) {
    for (size_t ii = 0; ii < h.AlphaMajorToBeta[i].size(); ii++) {
      int Astring = (int)i, Bstring = h.AlphaMajorToBeta[i][ii], DetI = h.AlphaMajorToDet[i][ii];
      if (DetI < 0 || DetI - 1 < start) continue;
      const std::vector<int>& bToA = h.BetaMajorToAlpha[Bstring];
      const std::vector<int>& bToDet = h.BetaMajorToDet[Bstring];
      // singles from Astring
      for (int Asingle : h.SinglesFromAlpha[Astring]) {
        int index = binarySearch(bToA.data(), 0, (int)bToA.size() - 1, Asingle);
        if (index != -1) {
          int DetJ = bToDet[index];
          if (std::abs(DetJ) >= DetI) continue;
          add(DetI, DetJ);
        }
      }
      // single alpha and single beta (merge walk)
      for (int Asingle : h.SinglesFromAlpha[Astring]) {
        const std::vector<int>& aToB = h.AlphaMajorToBeta[Asingle];
        const std::vector<int>& aToDet = h.AlphaMajorToDet[Asingle];
        int SearchStartIndex = 0, AlphaToBetaLen = (int)aToB.size();
        for (int Bsingle : h.SinglesFromBeta[Bstring]) {
          if (SearchStartIndex >= AlphaToBetaLen) break;
          int index = SearchStartIndex;
          for (; index < AlphaToBetaLen && aToB[index] < Bsingle; index++) {
          }
          SearchStartIndex = index;
          if (index < AlphaToBetaLen && aToB[index] == Bsingle) {
            int DetJ = aToDet[SearchStartIndex];
            if (std::abs(DetJ) >= DetI) continue;
            add(DetI, DetJ);
          }
        }
      }
      // singles from Bstring
      const std::vector<int>& aToB0 = h.AlphaMajorToBeta[Astring];
      const std::vector<int>& aToDet0 = h.AlphaMajorToDet[Astring];
      for (int Bsingle : h.SinglesFromBeta[Bstring]) {
        int index = binarySearch(aToB0.data(), 0, (int)aToB0.size() - 1, Bsingle);
        if (index != -1) {
          int DetJ = aToDet0[index];
          if (std::abs(DetJ) >= DetI) continue;
          add(DetI, DetJ);
        }
      }
      // double beta excitation
      for (size_t j = 0; j < aToDet0.size(); j++) {
        int DetJ = aToDet0[j];
        if (std::abs(DetJ) >= DetI) continue;
        if (doubleApart(DetI, DetJ)) add(DetI, DetJ);
      }
      // double alpha excitation
      for (size_t j = 0; j < bToDet.size(); j++) {
        int DetJ = bToDet[j];
        if (std::abs(DetJ) >= DetI) continue;
        if (doubleApart(DetI, DetJ)) add(DetI, DetJ);
      }
    }
  }
}

// Hmult2::operator(), Hmult.h:110-122 (serial branch)
void hmult(const SparseHam& H, const double* x, double* y, size_t n) {
  const int nt = oracleThreads();
  std::vector<std::vector<double>> part(nt, std::vector<double>(nt > 1 ? n : 0, 0.0));
  std::vector<double> yt(n, 0.0);
#pragma omp parallel num_threads(nt) if (nt > 1)
  {
#ifdef _OPENMP
    const int t = omp_get_thread_num();
#else
    const int t = 0;
#endif
    double* acc = nt > 1 ? part[t].data() : yt.data();
    // rows dealt round-robin, as the reference deals rows over ranks
    for (size_t i = (size_t)t; i < H.connections.size(); i += (size_t)nt) {
      const std::vector<int>& c = H.connections[i];
      const std::vector<double>& v = H.Helements[i];
      for (size_t j = 0; j < c.size(); j++) {
        double h = v[j];
        size_t J = (size_t)c[j];
        acc[i] += h * x[J];
        if (J != i) acc[J] += h * x[i];
      }
    }
  }
  if (nt > 1)
    for (int t = 0; t < nt; t++)
      for (size_t j = 0; j < n; j++) yt[j] += part[t][j];
  for (size_t j = 0; j < n; j++) y[j] = yt[j];
}

void hmultDirect(const System& s, const std::vector<Det>& dets, const SparseHam& H, const double* x, double* y,
                 size_t n) {
  std::vector<double> yt(n, 0.0);
  for (size_t i = 0; i < H.connections.size(); i++) {
    const std::vector<int>& c = H.connections[i];
    for (size_t j = 0; j < c.size(); j++) {
      const size_t J = (size_t)c[j];
      if (J == i) {
        yt[i] += energy(s, dets[i]) * x[i];
      } else {
        yt[i] += hij(s, dets[i], dets[J]) * x[J];
        yt[J] += hij(s, dets[J], dets[i]) * x[i];
      }
    }
  }
  for (size_t j = 0; j < n; j++) y[j] = yt[j];
}

// ---------------------------------------------------------------------------
// RDMs: SHCIrdm::EvaluateOneRDM / EvaluateRDM (SHCI/SHCIrdm.cpp:863-1100)
// over the stored rows (row i: i itself, then every connected J < i), each
// connection's excitation as Hij(Dets[J], Dets[i]) encodes it
// (Determinants.cpp:564-631: creation = in J not in i, destruction = in i
// not in J, both ascending), with the reference's sign rules and its
// spatial scatter populateSpatialRDM (SHCIrdm.cpp:846-860).
void evaluateRdm(const System& s, const std::vector<Det>& dets, const SparseHam& H, const double* bra,
                 const double* ket, double* g1, double* g2, double* s1, double* s2) {
  const int nso = s.nso, n = s.norb;
  const size_t nn = (size_t)n * n, npair = (size_t)nso * (nso + 1) / 2;
  auto spatial = [&](int i, int j, int k, int l, double v) {
    const size_t I = i / 2, J = j / 2, K = k / 2, L = l / 2;
    if (i % 2 == l % 2 && j % 2 == k % 2) {
      g2[(I * n + J) * nn + L * n + K] -= v;
      g2[(J * n + I) * nn + K * n + L] -= v;
    }
    if (i % 2 == k % 2 && l % 2 == j % 2) {
      g2[(I * n + J) * nn + K * n + L] += v;
      g2[(J * n + I) * nn + L * n + K] += v;
    }
  };
  auto spin2 = [&](size_t a, size_t b, size_t c, size_t d, double v) {  // twoRDM(a(a+1)/2+b, c(c+1)/2+d)
    if (s2) s2[(a * (a + 1) / 2 + b) * npair + c * (c + 1) / 2 + d] += v;
  };
  for (size_t i = 0; i < H.connections.size(); i++) {
    int closed[128], nc;
    occList(dets[i], s.W, closed, nc);
    const double cc = bra[i] * ket[i];
    for (int n1 = 0; n1 < nc; n1++) {  // <Di| Gamma |Di>
      const int o1 = closed[n1];
      g1[(size_t)(o1 / 2) * n + o1 / 2] += cc;
      if (s1) s1[(size_t)o1 * nso + o1] += cc;
      for (int n2 = 0; n2 < n1; n2++) {
        const int o2 = closed[n2];
        spin2(o1, o2, o1, o2, cc);
        spatial(o1, o2, o1, o2, cc);
      }
    }
    for (size_t jj = 1; jj < H.connections[i].size(); jj++) {
      const size_t J = (size_t)H.connections[i][jj];
      int cre[2] = {-1, -1}, des[2] = {-1, -1}, ncre = 0, ndes = 0;
      for (int w = 0; w < s.W; w++) {
        const uint64_t u = (uint64_t)(dets[J].r[w] ^ dets[i].r[w]);
        uint64_t b = u & (uint64_t)dets[J].r[w], k = u & (uint64_t)dets[i].r[w];
        for (; b; b &= b - 1) {
          if (ncre < 2) cre[ncre] = w * 64 + __builtin_ctzll(b);
          ncre++;
        }
        for (; k; k &= k - 1) {
          if (ndes < 2) des[ndes] = w * 64 + __builtin_ctzll(k);
          ndes++;
        }
      }
      const double vE = bra[J] * ket[i], vT = ket[J] * bra[i];
      if (ncre == 1) {
        const int c0 = cre[0], d0 = des[0];
        double sgn = 1.0;
        parity(dets[i], std::min(c0, d0), std::max(c0, d0), sgn);
        g1[(size_t)(c0 / 2) * n + d0 / 2] += sgn * vE;
        g1[(size_t)(d0 / 2) * n + c0 / 2] += sgn * vT;
        if (s1) {
          s1[(size_t)c0 * nso + d0] += sgn * vE;
          s1[(size_t)d0 * nso + c0] += sgn * vT;
        }
        for (int n1 = 0; n1 < nc; n1++) {
          const int cl = closed[n1];
          if (cl == d0 || cl == c0) continue;
          double sg = 1.0;
          parity(dets[i], std::min(d0, c0), std::max(d0, c0), sg);
          const int a = std::max(cl, c0), b = std::min(cl, c0), I = std::max(cl, d0), J2 = std::min(cl, d0);
          if (!((cl > c0 && cl > d0) || (cl < c0 && cl < d0))) sg *= -1.;
          spin2(a, b, I, J2, sg * vE);
          spin2(I, J2, a, b, sg * vT);
          spatial(a, b, I, J2, sg * vE);
          spatial(I, J2, a, b, sg * vT);
        }
      } else if (ncre == 2) {
        const int c0 = cre[0], c1 = cre[1], d0 = des[0], d1 = des[1];
        double sgn = 1.0;
        parityDouble(dets[i], d1, d0, c1, c0, sgn);
        spin2(c1, c0, d1, d0, sgn * vE);
        spin2(d1, d0, c1, c0, sgn * vT);
        spatial(c1, c0, d1, d0, sgn * vE);
        spatial(d1, d0, c1, c0, sgn * vT);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// cyclic Jacobi eigensolver for the small symmetric subspace matrix.
// A row-major m x m; w ascending; V row-major with eigenvectors in columns.
void symEig(int m, const double* A, double* w, double* V) {
  std::vector<double> a(A, A + (size_t)m * m);
  std::vector<double> v((size_t)m * m, 0.0);
  for (int i = 0; i < m; i++) v[(size_t)i * m + i] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0;
    for (int p = 0; p < m; p++)
      for (int q = p + 1; q < m; q++) off += std::fabs(a[(size_t)p * m + q]);
    if (off == 0.0) break;
    for (int p = 0; p < m; p++)
      for (int q = p + 1; q < m; q++) {
        double apq = a[(size_t)p * m + q];
        double app = a[(size_t)p * m + p], aqq = a[(size_t)q * m + q];
        double g = 100.0 * std::fabs(apq);
        if (sweep > 3 && std::fabs(app) + g == std::fabs(app) && std::fabs(aqq) + g == std::fabs(aqq)) {
          a[(size_t)p * m + q] = 0.0;
          a[(size_t)q * m + p] = 0.0;
          continue;
        }
        if (apq == 0.0) continue;
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
        for (int k = 0; k < m; k++) {  // columns p,q
          double akp = a[(size_t)k * m + p], akq = a[(size_t)k * m + q];
          a[(size_t)k * m + p] = c * akp - sn * akq;
          a[(size_t)k * m + q] = sn * akp + c * akq;
        }
        for (int k = 0; k < m; k++) {  // rows p,q
          double apk = a[(size_t)p * m + k], aqk = a[(size_t)q * m + k];
          a[(size_t)p * m + k] = c * apk - sn * aqk;
          a[(size_t)q * m + k] = sn * apk + c * aqk;
        }
        a[(size_t)p * m + q] = 0.0;
        a[(size_t)q * m + p] = 0.0;
        for (int k = 0; k < m; k++) {
          double vkp = v[(size_t)k * m + p], vkq = v[(size_t)k * m + q];
          v[(size_t)k * m + p] = c * vkp - sn * vkq;
          v[(size_t)k * m + q] = sn * vkp + c * vkq;
        }
      }
  }
  std::vector<int> order(m);
  for (int i = 0; i < m; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return a[(size_t)x * m + x] < a[(size_t)y * m + y]; });
  for (int j = 0; j < m; j++) {
    w[j] = a[(size_t)order[j] * m + order[j]];
    for (int k = 0; k < m; k++) V[(size_t)k * m + j] = v[(size_t)k * m + order[j]];
  }
}

static double dot(const double* a, const double* b, size_t n) {
  double s = 0.0;
  for (size_t i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}
static double norm(const double* a, size_t n) { return std::sqrt(dot(a, a, n)); }
static void setRandom(double* a, size_t n) {  // Eigen random(): -1 + 2*rand()/RAND_MAX
  for (size_t i = 0; i < n; i++) a[i] = -1.0 + (1.0 - (-1.0)) * double(std::rand()) / double(RAND_MAX);
}

// davidson, Davidson.cpp:111-348 (precondition :85-106)
std::vector<double> davidson(const SparseHam& H, std::vector<std::vector<double>>& x0,
                             const std::vector<double>& diag, int maxCopies, double tol, int& numIter,
                             bool print, FILE* out) {
  const size_t n = x0[0].size();
  return davidson([&](const double* x, double* y) { hmult(H, x, y, n); }, n, x0, diag, maxCopies, tol, numIter,
                  print, out);
}

std::vector<double> davidson(const std::function<void(const double*, double*)>& op, size_t N,
                             std::vector<std::vector<double>>& x0, const std::vector<double>& diag, int maxCopies,
                             double tol, int& numIter, bool print, FILE* out, bool directRules,
                             double startofCalc) {
  std::vector<double> eroots;
  int nroots = (int)x0.size();
  std::vector<double> b(N * maxCopies, 0.0), sigma(N * maxCopies, 0.0), r(N, 0.0);
  auto col = [&](std::vector<double>& M, int c) { return M.data() + (size_t)c * N; };
  for (int i = 0; i < nroots; i++) {
    std::copy(x0[i].begin(), x0[i].end(), col(b, i));
    if (norm(x0[i].data(), N) < 1.e-10) {
      setRandom(col(b, i), N);
      double nr = norm(col(b, i), N);
      for (size_t k = 0; k < N; k++) col(b, i)[k] /= nr;
    }
  }
  for (int i = 0; i < nroots; i++) {
    for (int j = 0; j < i; j++) {
      double ov = dot(col(b, j), col(b, i), N);
      for (size_t k = 0; k < N; k++) col(b, i)[k] -= ov * col(b, j)[k];
    }
    if (norm(col(b, i), N) < 1e-8) setRandom(col(b, i), N);
    for (int j = 0; j < i; j++) {
      double ov = dot(col(b, j), col(b, i), N);
      for (size_t k = 0; k < N; k++) col(b, i)[k] -= ov * col(b, j)[k];
    }
    double nr = norm(col(b, i), N);
    for (size_t k = 0; k < N; k++) col(b, i)[k] /= nr;
  }
  int sigmaSize = 0, bsize = nroots, convergedRoot = 0;
  numIter = 0;
  double ei = 0.0;
  std::vector<double> hsub, w, V, tmp;
  while (true) {
    for (int i = sigmaSize; i < bsize; i++) {
      op(col(b, i), col(sigma, i));
      sigmaSize++;
    }
    hsub.assign((size_t)bsize * bsize, 0.0);
    for (int i = 0; i < bsize; i++)
      for (int j = i; j < bsize; j++) {
        double v = dot(col(b, i), col(sigma, j), N);
        hsub[(size_t)i * bsize + j] = v;
        hsub[(size_t)j * bsize + i] = v;
      }
    w.assign(bsize, 0.0);
    V.assign((size_t)bsize * bsize, 0.0);
    symEig(bsize, hsub.data(), w.data(), V.data());
    tmp.assign(bsize, 0.0);
    for (size_t k = 0; k < N; k++) {  // Ritz rotation of b and sigma
      for (int j = 0; j < bsize; j++) {
        double s = 0.0;
        for (int l = 0; l < bsize; l++) s += b[(size_t)l * N + k] * V[(size_t)l * bsize + j];
        tmp[j] = s;
      }
      for (int j = 0; j < bsize; j++) b[(size_t)j * N + k] = tmp[j];
      for (int j = 0; j < bsize; j++) {
        double s = 0.0;
        for (int l = 0; l < bsize; l++) s += sigma[(size_t)l * N + k] * V[(size_t)l * bsize + j];
        tmp[j] = s;
      }
      for (int j = 0; j < bsize; j++) sigma[(size_t)j * N + k] = tmp[j];
    }
    ei = w[convergedRoot];
    for (int i = 0; i < convergedRoot; i++) {
      for (size_t k = 0; k < N; k++) r[k] = col(sigma, i)[k] - w[i] * col(b, i)[k];
      double error = norm(r.data(), N);
      if (error > tol) {
        convergedRoot = i;
        if (print && out) fprintf(out, "going back to converged root %d\n", i);
        continue;
      }
    }
    for (size_t k = 0; k < N; k++) r[k] = col(sigma, convergedRoot)[k] - ei * col(b, convergedRoot)[k];
    double error = norm(r.data(), N);
    if (directRules && print && out) {  // SHCI/Davidson.cpp:526-530
      if (numIter == 0) fprintf(out, "nIter  Root               Energy                Error\n");
      fprintf(out, "%5i  %4i   %18.10g   %18.10g  %10.2f\n", numIter, convergedRoot, ei, error,
              wallTime() - startofCalc);
    }
    numIter++;
    if ((size_t)bsize == N) {
      for (int i = 0; i < nroots; i++) {
        std::copy(col(b, i), col(b, i) + N, x0[i].begin());
        eroots.push_back(w[i]);
        if (directRules && print && out) fprintf(out, "#niter:%3d root:%3d -> Energy : %18.10g  \n", numIter, i, w[i]);
      }
      return eroots;
    }
    if (directRules && (error < tol || numIter > 400 * nroots)) {  // SHCI/Davidson.cpp:548-565
      if (numIter > 400 * nroots) {
        if (out) fprintf(out, "Davidson calculation did not converge for root %3d, #iter %5d\n", convergedRoot + 1, numIter);
        return {};
      }
      convergedRoot++;
      if (print && out) fprintf(out, "#niter:%3d root:%3d -> Energy : %18.10g  \n", numIter, convergedRoot - 1, ei);
      if (convergedRoot == nroots) {
        for (int i = 0; i < convergedRoot; i++) {
          std::copy(col(b, i), col(b, i) + N, x0[i].begin());
          eroots.push_back(w[i]);
        }
        return eroots;
      }
    } else if (!directRules && (error < tol || numIter > 800 * nroots)) {
      if (numIter > 2000 * nroots) {
        if (out) fprintf(out, "Davidson calculation did not converge for root %3d, #iter %5d\n", convergedRoot + 1, numIter);
        return eroots;
      }
      convergedRoot++;
      if (convergedRoot == nroots) {
        for (int i = 0; i < convergedRoot; i++) {
          std::copy(col(b, i), col(b, i) + N, x0[i].begin());
          eroots.push_back(w[i]);
        }
        return eroots;
      }
    }
    // precondition
    for (size_t k = 0; k < N; k++) {
      if (std::fabs(ei - diag[k]) > 1e-12)
        r[k] = r[k] / (ei - diag[k]);
      else
        r[k] = r[k] / (ei - diag[k] - 1.e-12);
    }
    for (int i = 0; i < bsize; i++) {
      double ov = dot(col(b, i), r.data(), N);
      double nn = dot(col(b, i), col(b, i), N);
      for (size_t k = 0; k < N; k++) r[k] = r[k] - ov * col(b, i)[k] / nn;
    }
    double nr = norm(r.data(), N);
    if (bsize < maxCopies) {
      for (size_t k = 0; k < N; k++) col(b, bsize)[k] = r[k] / nr;
      bsize++;
    } else {
      bsize = directRules ? std::min(nroots + 3, maxCopies) : nroots + 3;
      sigmaSize = bsize - 1;
      for (size_t k = 0; k < N; k++) col(b, bsize - 1)[k] = r[k] / nr;
    }
  }
}

// ---------------------------------------------------------------------------
static std::string detString(const Det& d, int nso) {  // Determinants.h:459-477
  std::string s;
  for (int i = 0; i < nso / 2; i++) {
    bool a = getocc(d, 2 * i), b = getocc(d, 2 * i + 1);
    if (!a && !b) s += "0 ";
    else if (a && !b) s += "a ";
    else if (!a && b) s += "b ";
    else s += "2 ";
    if ((i + 1) % 5 == 0) s += "  ";
  }
  return s;
}

// getParityForDiceToAlphaBeta, Determinants.cpp:691-706
static double parityDiceToAlphaBeta(const Det& d, int nso) {
  double par = 1.0;
  int nalpha = 0;
  for (int i = 0; i < nso; i += 2) nalpha += getocc(d, i);
  int norbs = nso / 2;
  for (int i = 0; i < norbs; i++) {
    if (getocc(d, 2 * (norbs - 1 - i) + 1)) {
      int before = 0;
      for (int j = 0; j < norbs - i - 1; j++)
        if (getocc(d, 2 * j)) before++;
      int after = nalpha - before;
      if (getocc(d, 2 * (norbs - 1 - i))) after--;
      if (after % 2 == 1) par *= -1;
    }
  }
  return par;
}

// DoVariational, SHCIbasics.cpp:783-1378, entered from main SHCI.cpp:157-446
void runVariational(const std::string& inputPath, RunResult& res, FILE* log, bool recordIters,
                    bool writeFiles) {
  double startofCalc = wallTime();
  Schedule schd;
  std::string err;
  if (!readInput(inputPath, schd, err, log)) {
    res.status = 1;
    res.error = err;
    if (log) fprintf(log, "%s\n", err.c_str());
    return;
  }
  if (schd.davidsonType == 2 || schd.doSOC || schd.searchForLowestEnergyDet || schd.doLCC) {
    res.status = 2;
    res.error = "keyword outside the restated variational path (disk/soc/search/lcc)";
    if (log) fprintf(log, "%s\n", res.error.c_str());
    return;
  }
  std::srand(schd.randomSeed + 0);
  System s;
  if (!readIntegrals(schd.integralFile, s, err)) {
    res.status = 1;
    res.error = err;
    if (log) fprintf(log, "%s\n", err.c_str());
    return;
  }
  if ((int)schd.occupied[0].size() != s.nelec) {
    res.status = 1;
    res.error = "The number of electrons given in the FCIDUMP should be equal to the nocc given in the shci input file.";
    if (log) fprintf(log, "%s\n", res.error.c_str());
    return;
  }
  buildHeatBath(s);
  if (schd.occupied[0].size() % 2 != 0 && schd.Trev != 0) {  // SHCI.cpp:189-195
    if (log) fprintf(log, "\nCannot use time reversal symmetry for odd electron system.\n");
    schd.Trev = 0;
  }
  s.trev = schd.Trev;
  int W = s.W;
  auto less = [W](const Det& x, const Det& y) { return detLess(x, y, W); };
  auto eq = [W](const Det& x, const Det& y) { return detEq(x, y, W); };

  if (log) {
    fprintf(log, "\n\n**************************************************************\n");
    fprintf(log, "SELECTING REFERENCE DETERMINANT(S)\n");
    fprintf(log, "**************************************************************\n");
  }
  std::vector<Det> Dets(schd.occupied.size());
  int lowestEnergyDet = 0;
  double lowestEnergy = 1.e12;
  for (size_t d = 0; d < schd.occupied.size(); d++) {
    for (int o : schd.occupied[d]) {
      if (o < 0 || o >= s.nso || getocc(Dets[d], o)) {
        res.status = 1;
        res.error = "orbital " + std::to_string(o) + " appears twice in input determinant number " + std::to_string(d);
        if (log) fprintf(log, "%s\n", res.error.c_str());
        return;
      }
      setocc(Dets[d], o, true);
    }
    if (s.trev != 0) makeStandard(Dets[d], W);
    for (size_t i = 0; i < d; i++)
      if (eq(Dets[d], Dets[i])) {
        res.status = 1;
        res.error = "Determinant appears twice in the input determinant list.";
        return;
      }
    double E = energy(s, Dets[d]) + s.coreE;
    if (log) fprintf(log, "%s Given Ref. Energy:    %18.10f\n", detString(Dets[d], s.nso).c_str(), E);
    if (E < lowestEnergy) {
      lowestEnergy = E;
      lowestEnergyDet = (int)d;
    }
  }
  int nroots = schd.nroots;
  std::vector<std::vector<double>> ci(nroots, std::vector<double>(Dets.size(), 0.0));
  {
    for (size_t j = 0; j < Dets.size(); j++) ci[0][j] = 1.0;
    ci[0][lowestEnergyDet] += (double)Dets.size();
    double nr = norm(ci[0].data(), Dets.size());
    for (size_t j = 0; j < Dets.size(); j++) ci[0][j] /= nr;
  }
  if (log) {
    fprintf(log, "\n\n**************************************************************\n");
    fprintf(log, "VARIATIONAL STEP  \n");
    fprintf(log, "**************************************************************\n");
  }
  // ---- DoVariational ----
  std::vector<Det> SortedDets = Dets;
  std::sort(SortedDets.begin(), SortedDets.end(), less);
  int DetsSize = (int)Dets.size();
  int SortedDetsSize = DetsSize;
  double coreEbkp = s.coreE;
  Helpers helpers;
  SparseHam H;
  std::vector<double> diag;
  double e0 = energy(s, Dets[0]);
  if (s.trev != 0) updateHijForTReversal(s, e0, Dets[0], Dets[0]);
  std::vector<double> E0(nroots, std::fabs(e0));
  populateHelpers(s, helpers, Dets, 0);
  makeHRows(s, helpers, Dets, 0, H);
  if (log) fprintf(log, "Iter Root       Eps1   #Var. Det.               Energy   #Davidson     Time(s)\n");
  // directdavidson: the same product with every element evaluated on the fly
  const bool direct = schd.davidsonType == 1;
  // (with time reversal the elements are those of the stored rows, which
  // HmultDirect recomputes with the same fixForTreversal)
  auto directOp = [&](const double* x, double* y) {
    if (s.trev != 0)
      hmult(H, x, y, (size_t)DetsSize);
    else
      hmultDirect(s, Dets, H, x, y, (size_t)DetsSize);
  };
  // checkpoint: the role of writeVariationalResult / readVariationalResult
  // (SHCIbasics.cpp:1481-1575) in the drop-in's own format: <prefix>/shci_var.bkp,
  // "SHCIVAR1", int32 W, nroots, iter, converged, int64 n, E0[nroots] (no
  // core energy), int64 dets[n][W], double ci[nroots][n]
  const std::string ckpt = schd.prefix[0] + "/shci_var.bkp";
  auto saveCheckpoint = [&](int iter) {  // SHCIbasics.cpp:1224-1227,1345-1352 (always "converged")
    if (!writeFiles || !schd.io) return;
    FILE* f = fopen(ckpt.c_str(), "wb");
    if (!f) return;
    const int32_t hdr[4] = {s.W, nroots, iter, 1};
    const int64_t n = DetsSize;
    fwrite("SHCIVAR1", 1, 8, f);
    fwrite(hdr, 4, 4, f);
    fwrite(&n, 8, 1, f);
    fwrite(E0.data(), 8, nroots, f);
    for (const Det& d : Dets) fwrite(d.r, 8, s.W, f);
    for (int r = 0; r < nroots; r++) fwrite(ci[r].data(), 8, n, f);
    fclose(f);
  };
  // restart (SHCIbasics.cpp:871-941)
  int iterstart = 0;
  bool fromConverged = false;
  if (schd.restart || schd.fullrestart) {
    FILE* f = fopen(ckpt.c_str(), "rb");
    char magic[8];
    int32_t hdr[4];
    int64_t n = 0;
    bool ok = f && fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "SHCIVAR1", 8) == 0 && fread(hdr, 4, 4, f) == 4 &&
              fread(&n, 8, 1, f) == 1 && hdr[0] == s.W && hdr[1] == nroots && n > 0;
    std::vector<double> e(nroots);
    std::vector<Det> d(ok ? n : 0);
    std::vector<std::vector<double>> c(nroots, std::vector<double>(ok ? n : 0));
    if (ok) {
      ok = fread(e.data(), 8, nroots, f) == (size_t)nroots;
      for (int64_t k = 0; ok && k < n; k++) ok = fread(d[k].r, 8, s.W, f) == (size_t)s.W;
      for (int r = 0; ok && r < nroots; r++) ok = fread(c[r].data(), 8, n, f) == (size_t)n;
    }
    if (f) fclose(f);
    if (!ok) {
      res.status = 1;
      res.error = "cannot read the checkpoint " + ckpt;
      if (log) fprintf(log, "%s\n", res.error.c_str());
      return;
    }
    Dets = d;
    ci = c;
    E0 = e;
    DetsSize = (int)n;
    SortedDets = Dets;
    std::sort(SortedDets.begin(), SortedDets.end(), less);
    SortedDetsSize = DetsSize;
    helpers = Helpers();
    H = SparseHam();
    populateHelpers(s, helpers, Dets, 0);
    makeHRows(s, helpers, Dets, 0, H);
    diag.resize(DetsSize);
    for (int k = 0; k < DetsSize; k++) diag[k] = energy(s, Dets[k]);
    const int last = (int)schd.epsilon1.size() - 1;
    iterstart = hdr[2];
    for (int i = 0; i < nroots; i++)
      if (log)
        fprintf(log, "%4d %4d  %10.2e  %10d -   %18.10f  %10.2f\n", iterstart, i,
                schd.epsilon1[std::min(iterstart, last)], DetsSize, E0[i] + coreEbkp, wallTime() - startofCalc);
    iterstart = schd.fullrestart ? 0 : iterstart + 1;
    if (!schd.fullrestart && hdr[3] &&
        (iterstart - 1 > last || schd.epsilon1[iterstart - 1] <= schd.epsilon1[last])) {
      if (log) fprintf(log, "# restarting from a converged calculation, moving to perturbative part.!!\n");
      fromConverged = true;
      if (s.trev != 0) unpackTrevState(Dets, ci, W, s.trev);
    }
  }
  for (int iter = iterstart; !fromConverged && iter < (int)schd.epsilon1.size(); iter++) {
    double epsilon1 = schd.epsilon1[iter];
    std::vector<double> cMax(DetsSize, 0.0);
    for (int j = 0; j < DetsSize; j++) {
      for (int i = 0; i < nroots; i++) cMax[j] += std::pow(std::fabs(ci[i][j]), 2);
      cMax[j] = std::pow(cMax[j], 0.5);
    }
    std::vector<Det> newDets;
    {
      const int nt = oracleThreads();
      std::vector<std::vector<Det>> found(nt);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt) if (nt > 1)
      for (int i = 0; i < SortedDetsSize; i++) {
#ifdef _OPENMP
        const int t = omp_get_thread_num();
#else
        const int t = 0;
#endif
        screenDet(s, schd, Dets[i], epsilon1 / std::fabs(cMax[i]), SortedDets.data(), SortedDetsSize, found[t]);
      }
      for (auto& f : found) newDets.insert(newDets.end(), f.begin(), f.end());
    }
    if (s.trev != 0)  // SHCIbasics.cpp:984-996
      for (Det& d : newDets) makeStandard(d, W);
    std::sort(newDets.begin(), newDets.end(), less);
    newDets.erase(std::unique(newDets.begin(), newDets.end(), eq), newDets.end());
    if (s.trev != 0)
      newDets.erase(std::remove_if(newDets.begin(), newDets.end(),
                                   [&](const Det& d) {
                                     return std::binary_search(SortedDets.begin(), SortedDets.begin() + SortedDetsSize,
                                                               d, less);
                                   }),
                    newDets.end());
    std::vector<std::vector<double>> X0(nroots, std::vector<double>(DetsSize + newDets.size(), 0.0));
    for (int i = 0; i < nroots; i++) std::copy(ci[i].begin(), ci[i].end(), X0[i].begin());
    Dets.insert(Dets.end(), newDets.begin(), newDets.end());
    DetsSize = (int)Dets.size();
    if (direct && log) fprintf(log, "New size of determinant space %8i\n", DetsSize);  // SHCIbasics.cpp:1083-1084
    populateHelpers(s, helpers, Dets, SortedDetsSize);
    makeHRows(s, helpers, Dets, SortedDetsSize, H);
    SortedDets = Dets;
    std::sort(SortedDets.begin(), SortedDets.end(), less);
    SortedDetsSize = DetsSize;
    size_t oldDiag = diag.size();
    diag.resize(DetsSize);
    for (size_t k = oldDiag; k < (size_t)DetsSize; k++) diag[k] = energy(s, Dets[k]);
    double prevE0 = E0[0];
    if (iter == 0) prevE0 = -10.0;
    int numIter = 0;
    if (direct)  // davidsonDirect(Hdirect, X0, diag, nroots + 2, ...), SHCIbasics.cpp:1167-1169
      E0 = davidson(directOp, (size_t)DetsSize, X0, diag, nroots + 2, schd.davidsonTolLoose, numIter,
                    schd.outputlevel > 0, log, true, startofCalc);
    else
      E0 = davidson(H, X0, diag, nroots + 4, schd.davidsonTolLoose, numIter, schd.outputlevel > 0, log);
    if ((int)E0.size() != nroots) {
      res.status = 3;
      res.error = "Davidson calculation did not converge";
      return;
    }
    double t = wallTime() - startofCalc;
    if (log) {
      fprintf(log, "%4d %4d  %10.2e  %10d", iter, 0, schd.epsilon1[iter], DetsSize);
      fprintf(log, "   %18.10f  %9d  %10.2f\n", E0[0] + coreEbkp, numIter, t);
      for (size_t i = 1; i < E0.size(); i++)
        fprintf(log, "%4d %4d  %10.2e  %10d   %18.10f  %9d  %10.2f\n", iter, (int)i, schd.epsilon1[iter], DetsSize,
                E0[i] + coreEbkp, numIter, t);
      if (E0.size() > 1) fprintf(log, "\n");
      fflush(log);
    }
    IterRecord rec;
    rec.iter = iter;
    rec.eps1 = epsilon1;
    rec.ndets = DetsSize;
    rec.numIter = numIter;
    rec.time = t;
    for (double e : E0) rec.e0.push_back(e + coreEbkp);
    res.iters.push_back(rec);
    if (recordIters) res.detsPerIter.push_back(Dets);
    for (int i = 0; i < nroots; i++) ci[i] = X0[i];
    if (std::fabs(E0[0] - prevE0) < schd.dE || iter == (int)schd.epsilon1.size() - 1) {
      if (log) fprintf(log, "Performing final tight davidson with tol: %g\n", schd.davidsonTol);
      if (direct)  // SHCIbasics.cpp:1206-1208
        E0 = davidson(directOp, (size_t)DetsSize, ci, diag, nroots + 4, schd.davidsonTol, numIter, true, log, true,
                      startofCalc);
      else
        E0 = davidson(H, ci, diag, nroots + 4, schd.davidsonTol, numIter, false, log);
      if ((int)E0.size() != nroots) {
        res.status = 3;
        res.error = "Davidson calculation did not converge";
        return;
      }
      if (log) fprintf(log, "Exiting variational iterations\n");
      saveCheckpoint(iter);
      if (s.trev != 0) unpackTrevState(Dets, ci, W, s.trev);
      break;
    }
    saveCheckpoint(iter);
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
  DetsSize = (int)Dets.size();  // (unpacked with time reversal)
  res.energies = E0;
  res.dets = Dets;
  res.ci = ci;
  res.coreE = coreEbkp;
  if (writeFiles) {
    std::string efile = schd.prefix[0] + "/shci.e";
    FILE* f = fopen(efile.c_str(), "wb");
    if (f) {
      for (double e : E0) fwrite(&e, 1, sizeof(double), f);
      fclose(f);
    }
  }
  if (log && schd.printAllDeterminants) {  // SHCI.cpp:449-462
    fprintf(log, "Printing all determinants\n");
    fprintf(log, "%4s %10s  Determinant string\n", "Det", "weight");
    for (int root = 0; root < nroots; root++) {
      fprintf(log, "State :%d\n", root);
      for (int i = 0; i < DetsSize; i++)
        fprintf(log, "%18.8e  %s\n", ci[root][i] * parityDiceToAlphaBeta(Dets[i], s.nso),
                detString(Dets[i], s.nso).c_str());
    }
  }
  if (writeFiles && schd.writeBestDeterminants > 0) {  // SHCI.cpp:464-505
    const int num = std::min(schd.writeBestDeterminants, DetsSize);
    const int nspatorbs = s.nso / 2;
    for (int root = 0; root < nroots; root++) {
      std::string fname = root == 0 ? "dets.bin" : "dets_" + std::to_string(root) + ".bin";
      FILE* f = fopen(fname.c_str(), "wb");
      if (!f) continue;
      fwrite(&num, sizeof(int), 1, f);
      fwrite(&nspatorbs, sizeof(int), 1, f);
      // |c| descending (ties, which the reference's sort leaves unordered,
      // in determinant order)
      std::vector<size_t> idx(DetsSize);
      for (size_t k = 0; k < idx.size(); k++) idx[k] = k;
      const std::vector<double>& c = ci[root];
      std::stable_sort(idx.begin(), idx.end(), [&c](size_t a, size_t b) { return std::fabs(c[a]) > std::fabs(c[b]); });
      for (int i = 0; i < num; i++) {
        const size_t m = idx[i];
        const double w = parityDiceToAlphaBeta(Dets[m], s.nso) * c[m];
        fwrite(&w, sizeof(double), 1, f);
        for (int k = 0; k < nspatorbs; k++) {
          const bool a = getocc(Dets[m], 2 * k), b = getocc(Dets[m], 2 * k + 1);
          const char o = a ? (b ? '2' : 'a') : (b ? 'b' : '0');
          fwrite(&o, 1, 1, f);
        }
      }
      fclose(f);
    }
  }
  if (log) {  // SHCI.cpp:526-551
    fprintf(log, "Printing most important determinants\n");
    fprintf(log, "%4s %10s  Determinant string\n", "Det", "weight");
    for (int root = 0; root < nroots; root++) {
      fprintf(log, "State : %3d\n", root);
      std::vector<double> prevci = ci[root];
      int num = std::max(6, schd.printBestDeterminants);
      for (int i = 0; i < std::min(num, DetsSize); i++) {
        int m = 0;
        for (int k = 1; k < DetsSize; k++)  // compAbs picks its float overload for doubles
          if (std::fabs((float)prevci[m]) < std::fabs((float)prevci[k])) m = k;
        double par = parityDiceToAlphaBeta(Dets[m], s.nso);
        fprintf(log, "%4d %18.10f  %s\n", i, prevci[m] * par, detString(Dets[m], s.nso).c_str());
        prevci[m] = 0.0;
      }
    }
    fflush(log);
  }
  // deterministic PT (SHCI/SHCI.cpp:706-721): shci.e becomes E_var + E_PT
  if (!schd.stochastic && schd.nblocks == 1) {
    if (log) {  // log_pt, SHCI.cpp:140-156
      fprintf(log, "\n\n**************************************************************\n");
      fprintf(log, "PERTURBATION THEORY STEP  \n");
      fprintf(log, "**************************************************************\n");
    }
    std::vector<double> ept(nroots);
    for (int root = 0; root < nroots; root++) {
      const double e0 = E0[root] - coreEbkp;
      const double pt = ptDeterministic(s, schd, Dets, ci[root].data(), e0, schd.epsilon2);
      ept[root] = E0[root] + pt;
      if (log) {  // SHCIbasics.cpp:664-668
        fprintf(log, "Deterministic PT calculation converged\n");
        fprintf(log, "PTEnergy: %18.10f\n", ept[root]);
        fprintf(log, "Time(s):  %10.2f\n", wallTime() - startofCalc);
        fflush(log);
      }
    }
    res.ptEnergies = ept;
    if (writeFiles) {
      std::string efile = schd.prefix[0] + "/shci.e";
      FILE* f = fopen(efile.c_str(), "wb");
      if (f) {
        for (double e : ept) fwrite(&e, 1, sizeof(double), f);
        fclose(f);
      }
    }
  } else if (schd.SampleN != -1 && schd.nPTiter != 0) {  // semistochastic PT, SHCI/SHCI.cpp:722-741
    if (log) {
      fprintf(log, "\n\n**************************************************************\n");
      fprintf(log, "PERTURBATION THEORY STEP  \n");
      fprintf(log, "**************************************************************\n");
    }
    std::vector<double> ept(nroots);
    for (int root = 0; root < nroots; root++)
      ept[root] = E0[root] + ptSemistochastic(s, schd, Dets, ci[root].data(), E0[root] - coreEbkp, coreEbkp, root,
                                              log, startofCalc);
    res.ptEnergies = ept;
    if (writeFiles) {
      std::string efile = schd.prefix[0] + "/shci.e";
      FILE* f = fopen(efile.c_str(), "wb");
      if (f) {
        for (double e : ept) fwrite(&e, 1, sizeof(double), f);
        fclose(f);
      }
    }
  }
}

}  // namespace oref
