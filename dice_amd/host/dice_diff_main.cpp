// dice_diff: compares two DICe output files, as tools/DICe_Diff.cpp does for
// serial files, so the reference's regression golds can be checked against this
// library's output.
//   dice_diff <file_a (gold)> <file_b> [-t <relative tolerance>] [-f <floor>] [-n] [-v]
// -n compares numbers only; -v prints the differing lines.  Prints TEST PASSED
// or TEST FAILED; the exit status is the number of differing lines.
#include <cstdlib>
#include <iostream>
#include <string>

#include "dice_schema.hpp"

int main(int argc, char **argv) {
  if (argc < 3) {
    std::cerr << "usage: dice_diff <file_a> <file_b> [-t <tol>] [-f <floor>] [-n] [-v]\n";
    return 99;
  }
  double tol = 1.0e-6, floor = 0.0;
  bool use_floor = false, numerical_only = false, verbose = false;
  for (int i = 3; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "-t" && i + 1 < argc) tol = std::strtod(argv[++i], nullptr);
    else if (a == "-f" && i + 1 < argc) { floor = std::strtod(argv[++i], nullptr); use_floor = true; }
    else if (a == "-n") numerical_only = true;
    else if (a == "-v") verbose = true;
    else { std::cerr << "unknown option " << a << "\n"; return 99; }
  }
  std::string report;
  const int errors = dice_hip::diff_files(argv[1], argv[2], tol, floor, use_floor, numerical_only, &report);
  if (verbose && !report.empty()) std::cout << report;
  std::cout << (errors == 0 ? "TEST PASSED" : "TEST FAILED") << std::endl;
  return errors;
}
