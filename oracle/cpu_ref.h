// cpu_ref — CPU ORACLE for the SHCI variational step.
//
// THIS IS TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, the smoke()
// entry and bench.py's cpu_baseline leg may load it, and only as the checker
// (or the timed CPU baseline).  The GPU product in dice_amd/ never links it.
//
// It restates, in plain C++17, the reference algorithm of sanshar/Dice's
// variational path (file:line citations are into the Dice source tree):
//   * FCIDUMP parsing and integral packing      SHCI/integral.cpp:43-225, SHCI/integral.h:47-102
//   * heat-bath tables (float, |v|-descending)  SHCI/integral.h:125-153, SHCI/integral.cpp:228-389
//   * determinant bit operations, parity        SHCI/Determinants.h:35-521
//   * Energy / Hij / Hij_1Excite / Hij_2Excite  SHCI/Determinants.cpp:252-300,406-505,564-631
//   * heat-bath screening                       SHCI/SHCIgetdeterminants.cpp:586-739
//   * string helpers and the sparse H rows      SHCI/SHCImakeHamiltonian.cpp:184-408,486-715
//   * sigma = H x                               SHCI/Hmult.h:52-150
//   * Davidson                                  SHCI/Davidson.cpp:85-348
//   * DoVariational driver                      SHCI/SHCIbasics.cpp:783-1378
//   * deterministic EN-PT2                      SHCI/SHCIbasics.cpp:453-728,
//                                               SHCI/SHCIgetdeterminants.cpp:40-171,
//                                               SHCI/SHCISortMpiUtils.cpp:541-690
//   * input.dat grammar                         SHCI/input.cpp:35-404
//   * RAS restrictions                          SHCI/OccRestrictions.cpp
// Deviations (documented, none changes a result beyond rounding):
//   * the m x m Rayleigh-Ritz eigenproblem uses cyclic Jacobi instead of
//     Eigen's SelfAdjointEigenSolver (same eigenpairs to ~1e-15);
//   * the PT energies of the external determinants are full Energy()
//     evaluations instead of EnergyAfterExcitation's update of the source's
//     energy (the same value up to rounding);
//   * half-determinant strings are held in one 64-bit word, so n <= 63
//     spatial orbitals (the reference allows 95); single-excitation string
//     probes stop at orbital n-1 (strings with higher bits never exist).
#pragma once
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <map>
#include <functional>
#include <string>
#include <vector>

namespace oref {

// Determinant: spin-orbital bitstring; bit k = spin-orbital k; even = alpha.
// Words beyond System::W are zero.  (SHCI/Determinants.h:166-200)
struct Det {
  int64_t r[2];
  Det() { r[0] = 0; r[1] = 0; }
};

struct Restriction {  // SHCI/OccRestrictions.h:37-72
  int minElec = 0, maxElec = 10000;
  std::vector<int> orbs;
  int currentElec = 0;
};

struct Schedule {  // subset of SHCI/input.h:35-194 that the variational path reads
  std::vector<std::vector<int>> occupied;
  std::vector<double> epsilon1;
  double davidsonTol = 5e-5, davidsonTolLoose = 5e-5, dE = 1e-8;
  double epsilon2 = 1e-8, epsilon2Large = 1000.0, targetError = 1e-4;
  int nroots = 1;
  unsigned int randomSeed = 0;
  bool io = true;
  int ncore = 0, nact = 1000000;
  std::string integralFile = "FCIDUMP";
  std::vector<std::string> prefix;
  int SampleN = -1, nPTiter = 1000000;
  bool stochastic = true;
  int outputlevel = 0;
  int printBestDeterminants = 0;
  bool printAllDeterminants = false;
  int writeBestDeterminants = 0;
  std::vector<Restriction> restrictionsV;
  std::vector<Restriction> restrictionsPT;  // "restrict" (not "restrictv") also restricts the PT space
  int nblocks = 1;
  int Trev = 0;
  int davidsonType = 0;  // 0 MEMORY, 1 DIRECT, 2 DISK
  bool restart = false, fullrestart = false, doSOC = false;
  bool searchForLowestEnergyDet = false, doLCC = false;
  bool DoRDM = false, DoOneRDM = false;
  int cdfci_on = 1000;
};

struct System {
  int norb = 0;  // spatial orbitals
  int nso = 0;   // spin orbitals
  int W = 1;     // 64-bit words per determinant (EffDetLen)
  int nelec = 0;
  double coreE = 0.0;
  int trev = 0;  // Treversal (Determinant::Trev): the basis is time-reversal adapted
  std::vector<int> irrep;
  std::vector<double> I1;  // nso x nso, row-major
  std::vector<double> I2;  // packed pair-of-pairs triangle
  double maxEntry = 0.0;
  std::vector<double> Direct, Exchange;  // norb x norb
  // heat-bath tables
  std::vector<float> hbSameVal, hbOppVal;
  std::vector<int16_t> hbSamePair, hbOppPair;
  std::vector<size_t> hbSameStart, hbOppStart;
  std::vector<double> Singles;  // nso x nso, Singles(p,q) at p*nso+q

  double i1(int i, int j) const { return I1[(size_t)i * nso + j]; }
  double i2(int i, int j, int k, int l) const {
    if (!((i % 2 == j % 2) && (k % 2 == l % 2))) return 0.0;
    int I = i / 2, J = j / 2, K = k / 2, L = l / 2;
    int IJ = std::max(I, J) * (std::max(I, J) + 1) / 2 + std::min(I, J);
    int KL = std::max(K, L) * (std::max(K, L) + 1) / 2 + std::min(K, L);
    size_t A = std::max(IJ, KL), B = std::min(IJ, KL);
    return I2[A * (A + 1) / 2 + B];
  }
  double& i2ref(int i, int j, int k, int l) {
    int I = i / 2, J = j / 2, K = k / 2, L = l / 2;
    int IJ = std::max(I, J) * (std::max(I, J) + 1) / 2 + std::min(I, J);
    int KL = std::max(K, L) * (std::max(K, L) + 1) / 2 + std::min(K, L);
    size_t A = std::max(IJ, KL), B = std::min(IJ, KL);
    return I2[A * (A + 1) / 2 + B];
  }
};

// sparse lower-triangular H, rows "diagonal first" (SHCI/SHCImakeHamiltonian.h:84-119)
struct SparseHam {
  std::vector<std::vector<int>> connections;
  std::vector<std::vector<double>> Helements;
};

// string helpers (SHCI/SHCImakeHamiltonian.h:44-82)
struct Helpers {
  std::map<uint64_t, int> AlphaN, BetaN;
  std::vector<std::vector<int>> AlphaMajorToBeta, AlphaMajorToDet;
  std::vector<std::vector<int>> BetaMajorToAlpha, BetaMajorToDet;
  std::vector<std::vector<int>> SinglesFromAlpha, SinglesFromBeta;
};

// --- building blocks -------------------------------------------------------
bool readIntegrals(const std::string& path, System& s, std::string& err);
void buildHeatBath(System& s);
bool readInput(const std::string& path, Schedule& schd, std::string& err,
               FILE* echo);

bool detLess(const Det& a, const Det& b, int W);
bool detEq(const Det& a, const Det& b, int W);
int excitationDistance(const Det& a, const Det& b, int W);
double energy(const System& s, const Det& d);
double hij(const System& s, const Det& bra, const Det& ket);
void screenDet(const System& s, const Schedule& schd, const Det& d,
               double epsilon, const Det* sorted, size_t nsorted,
               std::vector<Det>& out);
void populateHelpers(const System& s, Helpers& h, const std::vector<Det>& dets,
                     int start);
void makeHRows(const System& s, const Helpers& h, const std::vector<Det>& dets,
               int start, SparseHam& H);
void hmult(const SparseHam& H, const double* x, double* y, size_t n);
void symEig(int m, const double* A, double* w, double* V);
// HmultDirect (SHCI/Hmult.h:153-382): y = H x with every element evaluated
// on the fly as Hij(bra = D_I, ket = D_J) for row I; the connectivity is
// taken from the stored rows of H
void hmultDirect(const System& s, const std::vector<Det>& dets, const SparseHam& H, const double* x, double* y,
                 size_t n);
// SHCIrdm::EvaluateOneRDM / EvaluateRDM (SHCI/SHCIrdm.cpp:863-1100) of
// <bra| ... |ket> over the stored rows of H: g1 (norb^2), g2 (norb^4,
// g2[((P*n+Q)*n+R)*n+S] = s2RDM(P*n+Q, R*n+S)) and, when not null, the
// spin-orbital s1 (nso^2) and the reference's twoRDM s2 (npair^2, pair(p,q)
// = p(p+1)/2 + q); the caller zeroes them.
void evaluateRdm(const System& s, const std::vector<Det>& dets, const SparseHam& H, const double* bra,
                 const double* ket, double* g1, double* g2, double* s1, double* s2);
// davidson (SHCI/Davidson.cpp:111-348) on the product `op`; directRules
// selects davidsonDirect's restart size, iteration cap and progress lines
// (SHCI/Davidson.cpp:353-590)
std::vector<double> davidson(const std::function<void(const double*, double*)>& op, size_t N,
                             std::vector<std::vector<double>>& x0, const std::vector<double>& diag, int maxCopies,
                             double tol, int& numIter, bool print, FILE* out, bool directRules = false,
                             double startofCalc = 0.0);
std::vector<double> davidson(const SparseHam& H, std::vector<std::vector<double>>& x0,
                             const std::vector<double>& diag, int maxCopies,
                             double tol, int& numIter, bool print, FILE* out);

// Deterministic Epstein-Nesbet PT2 of one root (SHCI/SHCIbasics.cpp:453-728):
// every determinant outside `dets` reached from D_i by a single with
// |Hij| > eps2/|c_i| or a double whose heat-bath value passes eps2/|c_i|
// contributes (sum_i H_ai c_i)^2 / (E0 - E_a).  dets need not be sorted;
// E0 is the root's variational energy without the core energy.
// psi1Norm (optional) receives sum (num / (E0 - E_a))^2.
double ptDeterministic(const System& s, const Schedule& schd, const std::vector<Det>& dets, const double* ci,
                       double E0, double eps2, double* psi1Norm = nullptr, size_t* nExternal = nullptr);

// Semistochastic EN-PT2 of one root (SHCIbasics.cpp:55-452): the
// deterministic PT at epsilon2Large plus, per iteration, an unbiased estimate
// of the remainder down to epsilon2 from SampleN determinants drawn with
// probability |c_i| / sum |c| (alias method, SHCIsampledeterminants.cpp:37-181,
// with the C library's rand()); stops when the standard error is below
// targetError or after nPTiter iterations.  E0 without the core energy;
// coreE only shifts the printed energies.  Returns what the reference's
// function returns (the average energy correction, or the unaveraged sum
// when nPTiter runs out); *error / *iters receive the last standard error and
// the iteration count.
// index of one alias draw from a raw rand() value r: floor(r / RAND_MAX * n)
// clamped to n - 1 (SHCIsampledeterminants.cpp:160 without its r == RAND_MAX
// overrun)
int aliasDrawIndex(int r, int n);
double ptSemistochastic(const System& s, const Schedule& schd, const std::vector<Det>& dets, const double* ci,
                        double E0, double coreE, int root, FILE* log, double startofCalc, double* error = nullptr,
                        int* iters = nullptr, std::vector<double>* perIter = nullptr);

struct IterRecord {
  int iter;
  double eps1;
  int64_t ndets;
  std::vector<double> e0;
  int numIter;
  double time;
};

struct RunResult {
  int status = 0;
  std::string error;
  std::vector<IterRecord> iters;
  std::vector<Det> dets;
  std::vector<std::vector<double>> ci;
  std::vector<double> energies;
  std::vector<std::vector<Det>> detsPerIter;  // recorded when requested
  double coreE = 0.0;
  std::vector<double> ptEnergies;  // E_var + E_PT per root (written to shci.e) when PT ran
};

// Full variational step of the Dice executable (SHCI/SHCI.cpp:157-446 up to
// shci.e), then the deterministic PT step when the input asks for it
// (SHCI/SHCI.cpp:706-721).
void runVariational(const std::string& inputPath, RunResult& res, FILE* log,
                    bool recordIters, bool writeFiles);

}  // namespace oref
