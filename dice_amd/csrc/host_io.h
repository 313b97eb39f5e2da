// Host-side input handling for the SHCI variational step: FCIDUMP reader,
// input.dat parser and the heat-bath screening tables.  These keep the
// reference's file formats and semantics so the executable is a drop-in for
// `Dice input.dat` on the variational path.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace shci {

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
bool readInputFile(const std::string& path, HostSchedule& schd, std::string& err, std::string* echo);

}  // namespace shci
