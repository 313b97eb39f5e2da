// Host-side FCIDUMP / input.dat handling and screening-table construction.
#include "host_io.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <sys/time.h>

namespace shci {

namespace {

std::string stripWs(const std::string& s) {
  size_t b = s.find_first_not_of(" \t\r\n\f\v");
  if (b == std::string::npos) return std::string();
  size_t e = s.find_last_not_of(" \t\r\n\f\v");
  return s.substr(b, e - b + 1);
}

// Tokenise on any of `delims`, runs of delimiters collapsing into one
// separator; a leading or trailing run yields an empty token (the behaviour
// of boost::split with token_compress_on that the reference parsers use).
std::vector<std::string> tokens(const std::string& s, const char* delims) {
  std::vector<std::string> out(1);
  bool inRun = false;
  for (char c : s) {
    if (c != 0 && std::strchr(delims, c)) {
      if (!inRun) out.emplace_back();
      inRun = true;
    } else {
      out.back().push_back(c);
      inRun = false;
    }
  }
  return out;
}

bool same(const std::string& a, const char* b) {
  size_t n = std::strlen(b);
  if (a.size() != n) return false;
  for (size_t i = 0; i < n; i++)
    if (std::tolower((unsigned char)a[i]) != std::tolower((unsigned char)b[i])) return false;
  return true;
}
bool startsWith(const std::string& a, const char* b) {
  size_t n = std::strlen(b);
  return a.size() >= n && same(a.substr(0, n), b);
}

inline size_t pairIndex(int p, int q) { return p > q ? (size_t)p * (p + 1) / 2 + q : (size_t)q * (q + 1) / 2 + p; }
inline size_t packIndex(int p, int q, int r, int s) {
  size_t A = pairIndex(p, q), B = pairIndex(r, s);
  if (A < B) std::swap(A, B);
  return A * (A + 1) / 2 + B;
}

}  // namespace

double HostSystem::i2(int i, int j, int k, int l) const {
  if (((i ^ j) & 1) || ((k ^ l) & 1)) return 0.0;
  return I2[packIndex(i >> 1, j >> 1, k >> 1, l >> 1)];
}

// FCIDUMP: header lines until "&END" or "/", then "value p q r s" lines
// (format of SHCI/integral.cpp:43-225; writer lib/shcitools.c:165-234).
bool readFcidump(const std::string& path, HostSystem& s, std::string& err) {
  std::ifstream in(path.c_str());
  if (!in.good()) {
    err = "Integral file " + path + " does not exist!";
    return false;
  }
  int norb = -1, nelec = -1;
  bool ksym = false;
  std::string line;
  while (!in.eof()) {
    std::getline(in, line);
    std::vector<std::string> t = tokens(stripWs(line), ", \t=");
    if (t.size() == 1 && (same(t[0], "&END") || same(t[0], "/"))) break;
    if (startsWith(t[0], "&FCI")) {
      if (t.size() > 2 && startsWith(t[1], "NORB")) norb = std::atoi(t[2].c_str());
      if (t.size() > 4 && startsWith(t[3], "NELEC")) nelec = std::atoi(t[4].c_str());
    } else if (startsWith(t[0], "ISYM")) {
    } else if (startsWith(t[0], "KSYM")) {
      ksym = true;
    } else if (startsWith(t[0], "ORBSYM")) {
      for (size_t i = 1; i < t.size(); i++) s.irrep.push_back(std::atoi(t[i].c_str()));
    } else {
      for (auto& x : t) s.irrep.push_back(std::atoi(x.c_str()));
    }
  }
  if (norb == -1 || nelec == -1) {
    err = "could not read the norbs or nelec";
    return false;
  }
  if (ksym) {
    err = "KSYM (complex, k-point) integrals are not supported by this implementation";
    return false;
  }
  if (norb > 63) {
    err = "at most 63 spatial orbitals (126 spin-orbitals) are supported";
    return false;
  }
  s.irrep.resize(norb);
  s.norb = norb;
  s.nso = 2 * norb;
  s.nelec = nelec;
  s.W = s.nso / 64 + 1;
  size_t npair = (size_t)norb * (norb + 1) / 2;
  s.I2.assign(npair * (npair + 1) / 2, 0.0);
  s.h1.assign((size_t)norb * norb, 0.0);
  s.coreE = 0.0;
  while (!in.eof()) {
    std::getline(in, line);
    std::vector<std::string> t = tokens(stripWs(line), ", \t");
    if (t.size() != 5) continue;
    double v = std::atof(t[0].c_str());
    int p = std::atoi(t[1].c_str()), q = std::atoi(t[2].c_str()), r = std::atoi(t[3].c_str()),
        u = std::atoi(t[4].c_str());
    if (p == 0 && q == 0 && r == 0 && u == 0)
      s.coreE = v;
    else if (q == 0 && r == 0 && u == 0)
      continue;
    else if (r == 0 && u == 0) {
      s.h1[(size_t)(p - 1) * norb + (q - 1)] = v;
      s.h1[(size_t)(q - 1) * norb + (p - 1)] = v;
    } else {
      s.I2[packIndex(p - 1, q - 1, r - 1, u - 1)] = v;
    }
  }
  finishIntegrals(s);
  return true;
}

// maxEntry, Direct(i,j) = (ii|jj), Exchange(i,j) = (ij|ji) (SHCI/integral.cpp:181-205)
void finishIntegrals(HostSystem& s) {
  const int norb = s.norb;
  double best = s.I2.empty() ? 0.0 : s.I2[0];
  for (double v : s.I2)
    if (std::fabs(best) < std::fabs(v)) best = v;
  s.maxEntry = best;
  s.direct.assign((size_t)norb * norb, 0.0);
  s.exch.assign((size_t)norb * norb, 0.0);
  for (int i = 0; i < norb; i++)
    for (int j = 0; j < norb; j++) {
      s.direct[(size_t)i * norb + j] = s.I2[packIndex(i, i, j, j)];
      s.exch[(size_t)i * norb + j] = s.I2[packIndex(i, j, j, i)];
    }
}

bool integralsFromArrays(int norb, int nelec, const double* h1, const double* eri8, double coreE, const int* orbsym,
                         HostSystem& s, std::string& err) {
  if (norb < 1 || norb > 63) {
    err = "norb must be between 1 and 63 (at most 126 spin-orbitals are supported)";
    return false;
  }
  if (nelec < 1 || nelec > 2 * norb) {
    err = "nelec must be between 1 and 2 norb";
    return false;
  }
  if (!h1 || !eri8) {
    err = "h1 and eri8 are required";
    return false;
  }
  s = HostSystem();
  s.norb = norb;
  s.nso = 2 * norb;
  s.nelec = nelec;
  s.W = s.nso / 64 + 1;
  s.irrep.assign(norb, 1);
  if (orbsym)
    for (int i = 0; i < norb; i++) s.irrep[i] = orbsym[i];
  const size_t npair = (size_t)norb * (norb + 1) / 2;
  s.I2.assign(eri8, eri8 + npair * (npair + 1) / 2);
  s.h1.assign((size_t)norb * norb, 0.0);
  for (int p = 0; p < norb; p++)
    for (int q = 0; q <= p; q++) {  // the lower triangle, mirrored (a FCIDUMP lists each pair once)
      s.h1[(size_t)p * norb + q] = h1[(size_t)p * norb + q];
      s.h1[(size_t)q * norb + p] = h1[(size_t)p * norb + q];
    }
  s.coreE = coreE;
  finishIntegrals(s);
  return true;
}

// Heat-bath lists per spatial pair (i >= j): opposite spin {(a,b): (ia|jb)},
// same spin {(a >= b): (ia|jb) - (ib|ja)}, |v| > 1e-10, stored as float in
// descending |float| order with ties in reverse generation order
// (SHCI/integral.h:125-153, SHCI/integral.cpp:228-389); the single bound
// Singles(p,q) = max(|h_pq|, max_r |(pq|rr) - (pr|rq)|) over spin-orbitals r.
void buildScreeningTables(HostSystem& s) {
  const int n = s.norb, N = s.nso;
  const size_t npair = (size_t)n * (n + 1) / 2;
  struct Item {
    float v;
    uint32_t ab;
    int order;
  };
  for (int k = 0; k < 2; k++) {
    s.hbVal[k].clear();
    s.hbPair[k].clear();
    s.hbStart[k].assign(npair + 1, 0);
  }
  std::vector<Item> lst[2];
  size_t pi = 0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j <= i; j++, pi++) {
      lst[0].clear();
      lst[1].clear();
      int order = 0;
      for (int a = 0; a < n; a++)
        for (int b = 0; b < n; b++, order++) {
          double vab = s.I2[packIndex(i, a, j, b)];
          if (std::fabs(vab) > 1e-10) lst[0].push_back(Item{(float)vab, (uint32_t)a | ((uint32_t)b << 16), order});
          if (a >= b) {
            double d = vab - s.I2[packIndex(i, b, j, a)];
            if (std::fabs(d) > 1e-10) lst[1].push_back(Item{(float)d, (uint32_t)a | ((uint32_t)b << 16), order});
          }
        }
      for (int k = 0; k < 2; k++) {
        std::sort(lst[k].begin(), lst[k].end(), [](const Item& x, const Item& y) {
          float ax = std::fabs(x.v), ay = std::fabs(y.v);
          if (ax != ay) return ax > ay;
          return x.order > y.order;
        });
        for (const Item& it : lst[k]) {
          s.hbVal[k].push_back(it.v);
          s.hbPair[k].push_back(it.ab);
        }
        s.hbStart[k][pi + 1] = (int64_t)s.hbVal[k].size();
      }
    }
  s.singles.assign((size_t)N * N, 0.0);
  for (int p = 0; p < N; p++)
    for (int q = 0; q < N; q++) {
      double v = ((p ^ q) & 1) ? 0.0 : std::fabs(s.h1[(size_t)(p >> 1) * n + (q >> 1)]);
      for (int r = 0; r < N; r++) {
        double t = s.i2(p, q, r, r) - s.i2(p, r, r, q);
        if (std::fabs(v) < std::fabs(t)) v = std::fabs(t);
      }
      s.singles[(size_t)p * N + q] = v;
    }
  // single-excitation term tables: for a, i same spin and occupied j,
  // I2(a,i,j,j) - I2(a,j,j,i) equals tsame (j same spin) or topp (j opposite)
  s.tsame.assign((size_t)n * n * n, 0.0);
  s.topp.assign((size_t)n * n * n, 0.0);
  for (int a = 0; a < n; a++)
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        double v = s.I2[packIndex(a, i, j, j)];
        double x = s.I2[packIndex(a, j, j, i)];
        s.tsame[((size_t)a * n + i) * n + j] = v - x;
        s.topp[((size_t)a * n + i) * n + j] = v - 0.0;
      }
}

// input.dat keyword grammar of SHCI/input.cpp:35-404.  Keywords that select
// paths outside the variational step are accepted and recorded so the caller
// can refuse or ignore them; unknown keywords fail like the reference.
bool readInputFile(const std::string& path, HostSchedule& schd, std::string& err, std::string* echo) {
  std::ifstream ifs(path.c_str());
  if (!ifs.good()) {
    err = "input file " + path + " does not exist";
    return false;
  }
  std::stringstream dump;
  dump << ifs.rdbuf();
  int maxiter = -1, nocc = -1;
  std::vector<int> sweepIter;
  std::vector<double> sweepEps;
  {
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    schd.randomSeed = (unsigned int)(tv.tv_sec + 1e-6 * tv.tv_usec);
  }
  auto say = [&](const std::string& l) {
    if (echo) *echo += l + "\n";
  };
  while (dump.good()) {
    std::string line;
    std::getline(dump, line);
    line = stripWs(line);
    say(line);
    std::vector<std::string> t = tokens(line, ", \t\n");
    const std::string& key = t[0];
    if (key.empty() || key[0] == '#') continue;
    auto arg = [&](size_t k) { return k < t.size() ? t[k] : std::string(); };
    auto iarg = [&](size_t k) { return std::atoi(arg(k).c_str()); };
    auto farg = [&](size_t k) { return std::atof(arg(k).c_str()); };
    if (same(key, "nocc")) {
      nocc = iarg(1);
      std::string l;
      std::getline(dump, l);
      std::vector<std::string> o = tokens(stripWs(l), ", \t");
      while (!same(o[0], "end")) {
        if ((int)o.size() != nocc) {
          err = "nocc: " + std::to_string(nocc) + " neq " + std::to_string(o.size());
          return false;
        }
        std::vector<int> occ;
        std::string e;
        for (auto& x : o) {
          occ.push_back(std::atoi(x.c_str()));
          e += std::to_string(occ.back()) + " ";
        }
        say(e);
        schd.occupied.push_back(occ);
        if (!dump.good()) {
          err = "nocc block is not terminated by end";
          return false;
        }
        std::getline(dump, l);
        o = tokens(stripWs(l), ", \t");
      }
    } else if (same(key, "nact")) schd.nact = iarg(1);
    else if (same(key, "ncore")) schd.ncore = iarg(1);
    else if (same(key, "noio")) schd.io = false;
    else if (same(key, "io")) schd.io = true;
    else if (same(key, "dolcc")) schd.doLCC = true;
    else if (same(key, "directdavidson")) schd.davidsonType = 1;
    else if (same(key, "diskdavidson")) schd.davidsonType = 2;
    else if (same(key, "outputlevel")) schd.outputlevel = iarg(1);
    else if (same(key, "dosoc")) schd.doSOC = true;
    else if (same(key, "orbitals")) schd.integralFile = arg(1);
    else if (same(key, "nptiter")) schd.nPTiter = iarg(1);
    else if (same(key, "epsilon2")) schd.epsilon2 = farg(1);
    else if (same(key, "epsilon2Large")) schd.epsilon2Large = farg(1);
    else if (same(key, "targetError")) schd.targetError = farg(1);
    else if (same(key, "seed")) schd.randomSeed = (unsigned int)farg(1);
    else if (same(key, "nroots")) schd.nroots = (int)farg(1);
    else if (same(key, "printbestdeterminants")) schd.printBestDeterminants = iarg(1);
    else if (same(key, "printalldeterminants")) schd.printAllDeterminants = true;
    else if (same(key, "writebestdeterminants")) schd.writeBestDeterminants = iarg(1);
    else if (same(key, "dordm")) schd.doRDM = true;
    else if (same(key, "dospinrdm")) schd.doRDM = schd.doSpinRDM = true;
    else if (same(key, "DoOneRDM")) schd.doOneRDM = true;
    else if (same(key, "DoSpinOneRDM")) schd.doSpinOneRDM = true;
    else if (same(key, "TransitionRDM")) schd.doTransitionRDM = true;
    else if (same(key, "relaxedRDM")) schd.rdmUnrelaxed = true;  // (sic: SHCI/input.cpp:187-188)
    else if (same(key, "doresponse")) schd.doResponse = true;
    else if (same(key, "DoThreeRDM") || same(key, "DoFourRDM")) schd.doHigherRDM = true;
    else if (same(key, "Treversal")) {
      schd.Trev = iarg(1);
      if (!(schd.Trev == 0 || schd.Trev == 1 || schd.Trev == -1)) {
        err = "Treversal should be either 0, 1, or -1.";
        return false;
      }
    } else if (same(key, "restrict") || same(key, "restrictv")) {
      Restriction r;
      r.minElec = iarg(1);
      r.maxElec = iarg(2);
      for (size_t k = 3; k < t.size(); k++) {
        r.orbs.push_back(2 * std::atoi(t[k].c_str()));
        r.orbs.push_back(2 * std::atoi(t[k].c_str()) + 1);
      }
      schd.restrictions.push_back(r);
      if (same(key, "restrict")) schd.restrictionsPT.push_back(r);
    } else if (same(key, "nblocks")) {
      schd.nblocks = (int)std::atof(arg(1).c_str());
    } else if (same(key, "applyB")) {
      err = "applyB can only be used with ZDice";
      return false;
    } else if (same(key, "davidsonTol")) schd.davidsonTol = farg(1);
    else if (same(key, "davidsonTolLoose")) schd.davidsonTolLoose = farg(1);
    else if (same(key, "sampleN")) schd.SampleN = iarg(1);
    else if (same(key, "restart")) schd.restart = true;
    else if (same(key, "fullrestart")) schd.fullrestart = true;
    else if (same(key, "deterministic")) schd.stochastic = false;
    else if (same(key, "dE")) schd.dE = farg(1);
    else if (same(key, "prefix")) schd.prefix.push_back(arg(1));
    else if (same(key, "searchForLowestEnergyDet")) schd.searchLowest = true;
    else if (same(key, "cdfciOn")) schd.cdfciOn = iarg(1);
    else if (same(key, "schedule")) {
      std::getline(dump, line);
      say(line);
      std::vector<std::string> st = tokens(line, " \t");
      while (!same(st[0], "END")) {
        if (st[0].empty() || st[0][0] != '#') {
          sweepIter.push_back(std::atoi(st[0].c_str()));
          sweepEps.push_back(std::atof(st.size() > 1 ? st[1].c_str() : ""));
        }
        if (!dump.good()) {
          err = "schedule block is not terminated by end";
          return false;
        }
        std::getline(dump, line);
        say(line);
        st = tokens(line, " \t");
      }
    } else if (same(key, "maxiter")) maxiter = iarg(1);
    else if (same(key, "num_thrds") || same(key, "extrapolate") ||
             same(key, "algorithm") || same(key, "maxseniority") ||
             same(key, "maxexcitation") || same(key, "SenioAndExc") || same(key, "dogtensor") ||
             same(key, "dosocqdpt") || same(key, "socmultiplier") ||
             same(key, "onlyperturbative") || same(key, "quasiq") ||
             same(key, "cdfciTol") || same(key, "excitation") || same(key, "nvirt") || same(key, "eps") ||
             same(key, "pointGroup") || same(key, "spin") || same(key, "irrep") || same(key, "cdfciIter") || same(key, "reportInterval") ||
             same(key, "z_threshold") || same(key, "cdfciSample")) {
      // accepted; affects only steps after the variational one
    } else {
      err = "cannot read option " + key;
      return false;
    }
  }
  if (sweepIter.empty()) {
    err = "schedule keyword has to be included.";
    return false;
  }
  if (maxiter < sweepIter.back()) {
    err = "maxiter should be greater than last entry of sweep_iter";
    return false;
  }
  if (nocc == -1) {
    err = "nocc keyword has to be included.";
    return false;
  }
  if (schd.davidsonType == 1) schd.davidsonTolLoose = 3.e-2;
  for (size_t i = 1; i < sweepIter.size(); i++)
    for (int j = sweepIter[i - 1]; j < sweepIter[i]; j++) schd.epsilon1.push_back(sweepEps[i - 1]);
  for (int j = sweepIter.back(); j < maxiter; j++) schd.epsilon1.push_back(sweepEps.back());
  if (schd.prefix.empty()) schd.prefix.push_back(".");
  return true;
}

}  // namespace shci
