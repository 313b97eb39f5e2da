// shci_var: drop-in for `Dice input.dat` on the SHCI variational step.
//   shci_var [input.dat] [--precision fp64|fp32]
// Writes <prefix>/shci.e and the reference's log tables to stdout.
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/shci_gpu.h"

int main(int argc, char** argv) {
  std::string input = "input.dat";
  int precision = SHCI_PREC_FP64;
  for (int i = 1; i < argc; i++) {
    if (!std::strcmp(argv[i], "--precision") && i + 1 < argc) {
      std::string p = argv[++i];
      if (p == "fp32")
        precision = SHCI_PREC_FP32;
      else if (p == "fp64")
        precision = SHCI_PREC_FP64;
      else {
        fprintf(stderr, "unknown precision %s\n", p.c_str());
        return 1;
      }
    } else {
      input = argv[i];
    }
  }
  shci_handle* h = nullptr;
  shci_status st = shci_run(input.c_str(), nullptr, precision, &h);
  if (st != SHCI_OK) fprintf(stdout, "%s\n", shci_last_error(h));
  shci_destroy(h);
  return st == SHCI_OK ? 0 : (int)st;
}
