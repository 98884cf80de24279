// MPS files: the reader that feeds a model into the PDHG path, and a writer
// whose output that reader takes back unchanged.
//
// The reader accepts what the reference's reader accepts and builds the same
// model, field for field (src/anukulan/mps_reader.cpp:216-712): free or fixed
// format decided from the ROWS section, the first N row as the objective and
// later N rows dropped, RANGES applied as HiGHS applies them, every BOUNDS
// type with the HiGHS convention for a negative UP bound by default, integer
// MARKER sections, QUADOBJ / QMATRIX, the same warnings (at most 64) and the
// same error text, prefixed "source:line: ". The whole file is read into one
// buffer and tokens are views into it, so a line costs no allocation.
//
// The reference has no writer. This one writes free format with every value
// in the shortest form that reads back to the same double, so reading the
// file back gives the same model bit for bit, except
// where MPS itself cannot carry a model: free rows (both bounds infinite) are
// written as N rows, which the reader drops; a ranged row keeps both bounds
// only when one of them minus (or plus) the range reproduces the other
// exactly; finite bounds at or beyond 1e30 in magnitude read back as infinite.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

#include "lp.hpp"

namespace anuk {

// Column types in the reference's VarType order (model.hpp:12).
enum ColType : std::int32_t {
  kContinuous = 0,
  kInteger = 1,
  kSemiContinuous = 2,
  kSemiInteger = 3,
};

struct MpsOptions {
  // 0: decide from the ROWS section, 1: free, 2: fixed (mps_reader.hpp:13-19).
  int format = 0;
  // A negative UP bound on a column with no lower bound: false keeps the lower
  // bound at 0 (HiGHS), true sets it to -infinity for a continuous column
  // (CPLEX) (mps_reader.hpp:21-26).
  bool negative_up_minus_infinity = false;
  // BOUNDS values at or beyond this magnitude are infinite (mps_reader.hpp:29).
  double infinity_threshold = kFileInf;
};

// The model as the file states it (model.hpp:22-45): A x in [row_lower,
// row_upper], x in [col_lower, col_upper], objective c'x + offset (+ x'Qx/2).
struct MpsModel {
  std::string name;
  bool maximize = false;
  double objective_offset = 0.0;
  Csr a;
  Csr hessian;  // full symmetric, cols x cols; empty for an LP
  std::vector<double> objective, row_lower, row_upper, col_lower, col_upper;
  std::vector<std::int32_t> col_type;
  std::vector<std::string> row_names, col_names;
  std::vector<std::string> warnings;
  Int free_rows_dropped = 0;
  std::int64_t objective_nonzeros = 0;
  bool fixed_format = false;
  // A in triplet form, CSR order, for callers that take triplets.
  std::vector<Int> coo_row, coo_col;

  void fill_coo();
};

// A file the reader rejects; the message is "source:line: what".
struct MpsError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// A file that cannot be opened, read or written.
struct FileError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// The model's consistency checks (model.cpp:15-62): false, with the first
// problem in `error`, for mismatched lengths, NaN data, an empty bound
// interval or an asymmetric Hessian.
bool validate_model(const MpsModel& m, std::string* error);

MpsModel read_mps_text(std::string_view text, const std::string& source, const MpsOptions& opt);
MpsModel read_mps_file(const std::string& path, const MpsOptions& opt);

// What the writer needs; names and types may be empty (generated names R<i>
// and C<j>, all columns continuous).
struct MpsWriteModel {
  std::string name;
  Int rows = 0, cols = 0;
  const Int* start = nullptr;  // A as CSR
  const Int* index = nullptr;
  const double* value = nullptr;
  const double* objective = nullptr;
  const double* row_lower = nullptr;
  const double* row_upper = nullptr;
  const double* col_lower = nullptr;
  const double* col_upper = nullptr;
  bool maximize = false;
  double objective_offset = 0.0;
  const char* const* row_names = nullptr;
  const char* const* col_names = nullptr;
  const std::int32_t* col_type = nullptr;
};

// Throws std::invalid_argument for a model MPS cannot state (a name with
// white space, a duplicate name, NaN data).
std::string write_mps_text(const MpsWriteModel& m);

}  // namespace anuk
