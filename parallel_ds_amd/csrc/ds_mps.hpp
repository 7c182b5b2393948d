// MPS reader (host only): the model that the reference's Problem::readMps
// takes from COIN-OR Clp (common/problem.tpp:331-415), read by this library's
// own parser.
#pragma once
#include <string>
#include <vector>

namespace pds {

struct MpsModel {
  std::string name;
  int m = 0, n = 0;
  std::vector<int> col_ptr, row_idx;  // A (structurals) in CSC, rows in file order within a column
  std::vector<double> val;
  std::vector<double> costs, lower_bound, upper_bound;  // per column
  std::vector<int> bound_type;                          // BoundaryType order (ds_abi.h DS_*)
  std::vector<double> lower_range, upper_range;         // per row
  std::vector<int> range_type;
  double offset = 0.0;  // constant of the objective: minus the RHS entry of the objective row
  bool maximize = false;
};

// Parses a free- or fixed-format MPS file (sections NAME, OBJSENSE, ROWS,
// COLUMNS, RHS, RANGES, BOUNDS, ENDATA; names without blanks).  Throws
// std::runtime_error with the line number on malformed input.
void read_mps(const std::string& path, MpsModel& M);

}  // namespace pds
