// MPS reader.  Row bounds come from the row type, the RHS and the RANGES
// entry (the MPS convention that Clp's readMps applies); column bounds from
// BOUNDS, default [0, +inf).  Bound and range types then follow the
// reference's classification of the parsed bounds (problem.tpp:351-399):
// both infinite -> Free, only the upper finite -> Upper, only the lower ->
// Lower, equal -> Fixed, else Boxed, with |v| > 1e100 infinite
// (problem.hpp:31,83).  A model with more rows than columns is rejected, as
// checkConstraints does (problem.tpp:229-238).
#include "ds_mps.hpp"

#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

#include "ds_abi.h"

namespace pds {

namespace {

constexpr double MPS_INF = 1e30;  // |value| at or above this is infinite in MPS files
constexpr double BOUND_INF = 1e100;

bool isinf_bound(double v) { return v > BOUND_INF || v < -BOUND_INF; }

int classify(double lo, double up) {
  const bool li = isinf_bound(lo), ui = isinf_bound(up);
  if (li && ui) return DS_FREE;
  if (li) return DS_UPPER;
  if (ui) return DS_LOWER;
  if (lo == up) return DS_FIXED;
  return DS_BOXED;
}

double to_value(const std::string& tok, int line) {
  char* end = nullptr;
  const double v = std::strtod(tok.c_str(), &end);
  if (end == tok.c_str() || *end != '\0') throw std::runtime_error("MPS line " + std::to_string(line) + ": bad number '" + tok + "'");
  if (v >= MPS_INF) return INFINITY;
  if (v <= -MPS_INF) return -INFINITY;
  return v;
}

std::vector<std::string> split(const std::string& s) {
  std::vector<std::string> out;
  std::istringstream is(s);
  std::string t;
  while (is >> t) out.push_back(t);
  return out;
}

}  // namespace

void read_mps(const std::string& path, MpsModel& M) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open MPS file " + path);
  M = MpsModel();
  enum Sec { NONE, NAME, OBJSENSE, ROWS, COLUMNS, RHS, RANGES, BOUNDS, END } sec = NONE;
  std::string obj_row;
  std::unordered_map<std::string, int> row_of, col_of;
  std::vector<char> row_kind;  // 'E', 'L', 'G'
  std::vector<double> rhs, rng;
  std::vector<char> has_rng;
  std::unordered_map<std::string, char> ignored_rows;  // N rows after the first
  // COLUMNS entries by column, in file order
  std::vector<std::vector<std::pair<int, double>>> cols;
  std::string line;
  int ln = 0;
  auto fail = [&](const std::string& what) { throw std::runtime_error("MPS line " + std::to_string(ln) + ": " + what); };
  while (std::getline(in, line)) {
    ++ln;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.empty() || line[0] == '*') continue;
    const bool header = line[0] != ' ' && line[0] != '\t';
    std::vector<std::string> f = split(line);
    if (f.empty()) continue;
    if (header) {
      const std::string& h = f[0];
      if (h == "NAME") {
        sec = NAME;
        M.name = f.size() > 1 ? f[1] : "";
      } else if (h == "OBJSENSE") {
        sec = OBJSENSE;
        if (f.size() > 1) M.maximize = f[1] == "MAX" || f[1] == "MAXIMIZE";
      } else if (h == "ROWS") sec = ROWS;
      else if (h == "COLUMNS") sec = COLUMNS;
      else if (h == "RHS") sec = RHS;
      else if (h == "RANGES") sec = RANGES;
      else if (h == "BOUNDS") sec = BOUNDS;
      else if (h == "ENDATA") { sec = END; break; }
      else fail("unknown section '" + h + "'");
      continue;
    }
    switch (sec) {
      case OBJSENSE:
        M.maximize = f[0] == "MAX" || f[0] == "MAXIMIZE";
        break;
      case ROWS: {
        if (f.size() < 2) fail("ROWS entry needs a type and a name");
        const char t = f[0][0];
        if (t == 'N' || t == 'n') {
          if (obj_row.empty()) obj_row = f[1];
          else ignored_rows[f[1]] = 1;
        } else if (t == 'E' || t == 'L' || t == 'G' || t == 'e' || t == 'l' || t == 'g') {
          if (row_of.count(f[1])) fail("duplicate row '" + f[1] + "'");
          row_of[f[1]] = (int)row_kind.size();
          row_kind.push_back((char)std::toupper(t));
          rhs.push_back(0.0);
          rng.push_back(0.0);
          has_rng.push_back(0);
        } else {
          fail("unknown row type '" + f[0] + "'");
        }
        break;
      }
      case COLUMNS: {
        if (f.size() >= 3 && f[1] == "'MARKER'") break;  // integer markers: the LP relaxation is read
        if (f.size() != 3 && f.size() != 5) fail("COLUMNS entry needs a column and one or two (row, value) pairs");
        auto it = col_of.find(f[0]);
        int j;
        if (it == col_of.end()) {
          j = (int)cols.size();
          col_of[f[0]] = j;
          cols.emplace_back();
          M.costs.push_back(0.0);
        } else {
          j = it->second;
        }
        for (size_t k = 1; k + 1 < f.size(); k += 2) {
          const double v = to_value(f[k + 1], ln);
          if (f[k] == obj_row) {
            M.costs[j] = v;
          } else if (ignored_rows.count(f[k])) {
          } else {
            auto r = row_of.find(f[k]);
            if (r == row_of.end()) fail("unknown row '" + f[k] + "'");
            if (v != 0.0) cols[j].push_back({r->second, v});
          }
        }
        break;
      }
      case RHS:
      case RANGES: {
        // optional set name first: an odd count of fields after it means the name is present
        size_t k0 = (f.size() % 2 == 1) ? 1 : 0;
        if (f.size() - k0 < 2) fail("entry needs (row, value) pairs");
        for (size_t k = k0; k + 1 < f.size(); k += 2) {
          const double v = to_value(f[k + 1], ln);
          if (sec == RHS && f[k] == obj_row) {
            M.offset = -v;
            continue;
          }
          if (ignored_rows.count(f[k])) continue;
          auto r = row_of.find(f[k]);
          if (r == row_of.end()) fail("unknown row '" + f[k] + "'");
          if (sec == RHS) rhs[r->second] = v;
          else {
            rng[r->second] = v;
            has_rng[r->second] = 1;
          }
        }
        break;
      }
      case BOUNDS: {
        if (f.size() < 2) fail("BOUNDS entry needs a type and a column");
        const std::string& t = f[0];
        const bool novalue = t == "FR" || t == "MI" || t == "PL" || t == "BV";
        if (!novalue && f.size() < 3) fail("BOUNDS entry needs a value");
        // "type set col [value]" or "type col [value]" (no set name)
        std::string cname;
        std::string vtok;
        if (novalue) cname = f.size() >= 3 ? f[2] : f[1];
        else if (f.size() >= 4) { cname = f[2]; vtok = f[3]; }
        else { cname = f[1]; vtok = f[2]; }
        auto it = col_of.find(cname);
        if (it == col_of.end()) fail("unknown column '" + cname + "'");
        const int j = it->second;
        if ((int)M.lower_bound.size() < (int)cols.size()) {
          M.lower_bound.resize(cols.size(), 0.0);
          M.upper_bound.resize(cols.size(), INFINITY);
        }
        const double v = vtok.empty() ? 0.0 : to_value(vtok, ln);
        if (t == "UP" || t == "UI") {
          M.upper_bound[j] = v;
          if (v < 0.0 && M.lower_bound[j] == 0.0) M.lower_bound[j] = -INFINITY;  // MPS convention
        } else if (t == "LO" || t == "LI") {
          M.lower_bound[j] = v;
        } else if (t == "FX") {
          M.lower_bound[j] = M.upper_bound[j] = v;
        } else if (t == "FR") {
          M.lower_bound[j] = -INFINITY;
          M.upper_bound[j] = INFINITY;
        } else if (t == "MI") {
          M.lower_bound[j] = -INFINITY;
        } else if (t == "PL") {
          M.upper_bound[j] = INFINITY;
        } else if (t == "BV") {
          M.lower_bound[j] = 0.0;
          M.upper_bound[j] = 1.0;
        } else {
          fail("unknown bound type '" + t + "'");
        }
        break;
      }
      case NAME:
        break;
      default:
        fail("data before the first section");
    }
  }
  if (sec != END) throw std::runtime_error("MPS file " + path + " has no ENDATA");
  if (obj_row.empty()) throw std::runtime_error("MPS file " + path + " has no objective row");
  M.m = (int)row_kind.size();
  M.n = (int)cols.size();
  M.lower_bound.resize(M.n, 0.0);
  M.upper_bound.resize(M.n, INFINITY);
  if (M.maximize)
    for (double& c : M.costs) c = -c;
  M.col_ptr.assign(M.n + 1, 0);
  for (int j = 0; j < M.n; ++j) {
    M.col_ptr[j + 1] = M.col_ptr[j] + (int)cols[j].size();
    for (const auto& e : cols[j]) {
      M.row_idx.push_back(e.first);
      M.val.push_back(e.second);
    }
  }
  M.lower_range.resize(M.m);
  M.upper_range.resize(M.m);
  M.range_type.resize(M.m);
  for (int i = 0; i < M.m; ++i) {
    double lo, up;
    const double r = std::fabs(rng[i]);
    switch (row_kind[i]) {
      case 'E':
        lo = up = rhs[i];
        if (has_rng[i]) {
          if (rng[i] >= 0) up = rhs[i] + r;
          else lo = rhs[i] - r;
        }
        break;
      case 'L':
        lo = has_rng[i] ? rhs[i] - r : -INFINITY;
        up = rhs[i];
        break;
      default:  // 'G'
        lo = rhs[i];
        up = has_rng[i] ? rhs[i] + r : INFINITY;
        break;
    }
    M.lower_range[i] = lo;
    M.upper_range[i] = up;
    M.range_type[i] = classify(lo, up);
  }
  M.bound_type.resize(M.n);
  for (int j = 0; j < M.n; ++j) M.bound_type[j] = classify(M.lower_bound[j], M.upper_bound[j]);
  if (M.m > M.n) throw std::runtime_error("Incorrect constraint size: more rows than columns");
}

}  // namespace pds
