// Host-side input handling for the SHCI variational step: FCIDUMP reader,
// input.dat parser and the heat-bath screening tables.  These keep the
// reference's file formats and semantics so the executable is a drop-in for
// `Dice input.dat` on the variational path.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace shci {

// The C library's rand() sequence (glibc's additive feedback generator,
// TYPE_3: 31 words, separation 3, seeded by the Park-Miller minimal
// standard generator through Schrage's method and 310 discarded outputs),
// held privately so that nothing else in the process (the GPU runtime
// included) can advance it between the reference's srand(seed) and its
// draws (SHCI/SHCI.cpp:199; alias sampling, SHCI/SHCIsampledeterminants.cpp:148-176;
// random Davidson guesses, eigen/Eigen/src/Core/MathFunctions.h:825-834).
struct LibcRand {
  static const int kMax = 2147483647;  // RAND_MAX
  int32_t state[31];
  int f = 3, r = 0;
  explicit LibcRand(unsigned seed = 1) { reseed(seed); }
  void reseed(unsigned seed) {
    if (seed == 0) seed = 1;
    state[0] = (int32_t)seed;
    int32_t word = (int32_t)seed;
    for (int i = 1; i < 31; i++) {
      const long hi = word / 127773, lo = word % 127773;
      word = (int32_t)(16807 * lo - 2836 * hi);
      if (word < 0) word += 2147483647;
      state[i] = word;
    }
    f = 3;
    r = 0;
    for (int i = 0; i < 310; i++) (void)next();
  }
  int next() {
    const uint32_t val = (uint32_t)state[f] + (uint32_t)state[r];
    state[f] = (int32_t)val;
    f = f + 1 == 31 ? 0 : f + 1;
    r = r + 1 == 31 ? 0 : r + 1;
    return (int)(val >> 1);
  }
};

struct HostSystem {
  int norb = 0, nso = 0, W = 1, nelec = 0;
  double coreE = 0.0;
  std::vector<int> irrep;
  std::vector<double> h1;      // norb x norb
  std::vector<double> I2;      // packed pair-of-pairs triangle
  double maxEntry = 0.0;       // signed element of largest magnitude
  std::vector<double> direct, exch;  // norb x norb
  std::vector<double> tsame, topp;   // norb^3 single-excitation term tables
  std::vector<double> singles;       // nso x nso
  // heat-bath double-excitation lists, index 0 = opposite spin, 1 = same spin
  std::vector<float> hbVal[2];
  std::vector<uint32_t> hbPair[2];
  std::vector<int64_t> hbStart[2];

  double i2(int i, int j, int k, int l) const;  // spin-orbital accessor
};

struct Restriction {
  int minElec = 0, maxElec = 10000;
  std::vector<int> orbs;  // spin-orbitals
};

struct HostSchedule {
  std::vector<std::vector<int>> occupied;
  std::vector<double> epsilon1;
  double davidsonTol = 5e-5, davidsonTolLoose = 5e-5, dE = 1e-8, epsilon2 = 1e-8;
  double epsilon2Large = 1000.0, targetError = 1e-4;
  int nroots = 1;
  unsigned int randomSeed = 0;
  bool io = true;
  int ncore = 0, nact = 1000000;
  std::string integralFile = "FCIDUMP";
  std::vector<std::string> prefix;
  int SampleN = -1, nPTiter = 1000000;
  bool stochastic = true;
  int outputlevel = 0, printBestDeterminants = 0, writeBestDeterminants = 0;
  bool printAllDeterminants = false;
  std::vector<Restriction> restrictions;    // variational space ("restrict", "restrictv")
  std::vector<Restriction> restrictionsPT;  // PT space ("restrict")
  int nblocks = 1;
  int Trev = 0, davidsonType = 0;
  bool restart = false, fullrestart = false, doSOC = false, searchLowest = false, doLCC = false;
  bool doRDM = false, doOneRDM = false, doSpinRDM = false, doSpinOneRDM = false, doTransitionRDM = false;
  bool rdmUnrelaxed = false, doResponse = false, doHigherRDM = false;
  int cdfciOn = 1000;
  // extensions of this implementation (not reference keywords; set by the
  // executable's command line): Hamiltonian value precision and block size
  int precision = 0;
};

// Returns false and fills err on failure (messages follow the reference).
bool readFcidump(const std::string& path, HostSystem& s, std::string& err);
void buildScreeningTables(HostSystem& s);
// maxEntry, Direct and Exchange from I2
void finishIntegrals(HostSystem& s);
// The system from in-memory integrals: h1 = norb x norb row-major (its lower
// triangle is used, mirrored, as a FCIDUMP lists each pair once), eri8 = the
// 8-fold packed (pq|rs) in the reference's twoInt order (index A(A+1)/2 + B
// of pair indices A >= B, pair(p,q) = p(p+1)/2 + q for p >= q; PySCF's
// ao2mo.restore(8, ...) layout), orbsym optional (norb irreps, default 1).
bool integralsFromArrays(int norb, int nelec, const double* h1, const double* eri8, double coreE, const int* orbsym,
                         HostSystem& s, std::string& err);
bool readInputFile(const std::string& path, HostSchedule& schd, std::string& err, std::string* echo);

}  // namespace shci
