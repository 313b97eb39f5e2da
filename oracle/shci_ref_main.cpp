// shci_ref: the CPU oracle as a `Dice input.dat` stand-in (variational step
// only).  TEST INFRASTRUCTURE / CPU BASELINE ONLY.  Usage: shci_ref [input.dat]
#include <cstdio>
#include <string>

#include "cpu_ref.h"

int main(int argc, char** argv) {
  std::string input = argc > 1 ? argv[1] : "input.dat";
  oref::RunResult r;
  oref::runVariational(input, r, stdout, false, true);
  return r.status;
}
