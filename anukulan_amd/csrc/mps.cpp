#include "mps.hpp"

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <cstdio>
#include <unordered_set>

namespace anuk {
namespace {

// The C locale's white space, which is what the reference tokenises on
// (std::isspace, mps_reader.cpp:41-56).
inline bool blank(char c) {
  return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}

inline char up(char c) { return (c >= 'a' && c <= 'z') ? static_cast<char>(c - 'a' + 'A') : c; }

bool same_upper(std::string_view token, std::string_view word) {
  if (token.size() != word.size()) return false;
  for (std::size_t i = 0; i < token.size(); ++i)
    if (up(token[i]) != word[i]) return false;
  return true;
}

std::string upper_copy(std::string_view s) {
  std::string out(s);
  for (char& c : out) c = up(c);
  return out;
}

using Tokens = std::vector<std::string_view>;

void split_free(std::string_view line, Tokens* out) {
  out->clear();
  std::size_t i = 0;
  while (i < line.size()) {
    while (i < line.size() && blank(line[i])) ++i;
    if (i >= line.size()) break;
    const std::size_t b = i;
    while (i < line.size() && !blank(line[i])) ++i;
    out->push_back(line.substr(b, i - b));
  }
}

// Fixed format: six fields at 0-based column ranges [1,3) [4,12) [14,22)
// [24,36) [39,47) [49,end), each trimmed of spaces and tabs; empty fields are
// skipped (mps_reader.cpp:58-88).
void split_fixed(std::string_view line, Tokens* out) {
  static constexpr std::size_t kFrom[6] = {1, 4, 14, 24, 39, 49};
  static constexpr std::size_t kTo[6] = {3, 12, 22, 36, 47, std::string_view::npos};
  out->clear();
  for (int f = 0; f < 6; ++f) {
    if (kFrom[f] >= line.size()) break;
    const std::size_t to = std::min(kTo[f], line.size());
    std::size_t b = kFrom[f], e = to;
    while (b < e && (line[b] == ' ' || line[b] == '\t')) ++b;
    while (e > b && (line[e - 1] == ' ' || line[e - 1] == '\t')) --e;
    if (b < e) out->push_back(line.substr(b, e - b));
  }
}

// strtod after mapping a Fortran D exponent to E; trailing white space is
// allowed, anything else (or NaN) is not (mps_reader.cpp:91-106).
bool number(std::string_view token, double* out) {
  std::string s(token);
  for (char& c : s)
    if (c == 'd' || c == 'D') c = 'E';
  const char* b = s.c_str();
  char* e = nullptr;
  const double v = std::strtod(b, &e);
  if (e == b) return false;
  while (*e != '\0' && blank(*e)) ++e;
  if (*e != '\0' || std::isnan(v)) return false;
  *out = v;
  return true;
}

enum class Part { kNone, kName, kObjSense, kRows, kColumns, kRhs, kRanges, kBounds, kQuadObj, kQMatrix, kEndata };

Part part_of(std::string_view word) {
  const std::string w = upper_copy(word);
  if (w == "NAME") return Part::kName;
  if (w == "OBJSENSE" || w == "OBJSENS") return Part::kObjSense;
  if (w == "ROWS") return Part::kRows;
  if (w == "COLUMNS") return Part::kColumns;
  if (w == "RHS") return Part::kRhs;
  if (w == "RANGES") return Part::kRanges;
  if (w == "BOUNDS") return Part::kBounds;
  if (w == "QUADOBJ" || w == "QSECTION" || w == "QUADS") return Part::kQuadObj;
  if (w == "QMATRIX") return Part::kQMatrix;
  if (w == "ENDATA") return Part::kEndata;
  return Part::kNone;
}

bool is_max(std::string_view w) { return same_upper(w, "MAX") || same_upper(w, "MAXIMIZE"); }

// Names to indices: open addressing with linear probing at load <= 1/2,
// keys are views into the file buffer stored with their 64-bit FNV-1a hash.
// One probe usually touches one slot and the name itself, where a node-based
// map walks a bucket and a node first.
class NameTable {
 public:
  const Int* find(std::string_view k) const {
    if (slots_.empty()) return nullptr;
    const std::uint64_t h = hash(k);
    for (std::size_t i = h & mask_;; i = (i + 1) & mask_) {
      const Slot& s = slots_[i];
      if (s.p == nullptr) return nullptr;
      if (s.h == h && s.len == k.size() && std::memcmp(s.p, k.data(), k.size()) == 0) return &s.v;
    }
  }
  bool contains(std::string_view k) const { return find(k) != nullptr; }
  void insert(std::string_view k, Int v) {
    if (2 * (used_ + 1) > slots_.size()) grow();
    place(hash(k), k, v);
    ++used_;
  }

 private:
  struct Slot {
    std::uint64_t h = 0;
    const char* p = nullptr;  // null: empty slot
    std::size_t len = 0;
    Int v = 0;
  };
  static std::uint64_t hash(std::string_view k) {
    std::uint64_t h = 1469598103934665603ull;
    for (char c : k) h = (h ^ static_cast<unsigned char>(c)) * 1099511628211ull;
    return h ^ (h >> 29);
  }
  void place(std::uint64_t h, std::string_view k, Int v) {
    std::size_t i = h & mask_;
    while (slots_[i].p != nullptr) i = (i + 1) & mask_;
    // An empty name still needs a non-null pointer to mark the slot used.
    slots_[i] = Slot{h, k.data() != nullptr ? k.data() : "", k.size(), v};
  }
  void grow() {
    std::vector<Slot> old = std::move(slots_);
    slots_.assign(old.empty() ? 64 : old.size() * 2, Slot{});
    mask_ = slots_.size() - 1;
    for (const Slot& s : old)
      if (s.p != nullptr) place(s.h, std::string_view(s.p, s.len), s.v);
  }
  std::vector<Slot> slots_;
  std::size_t mask_ = 0;
  std::size_t used_ = 0;
};

constexpr Int kObjectiveRow = -1;
constexpr Int kDroppedRow = -2;
constexpr std::size_t kWarningCap = 64;

class Parser {
 public:
  Parser(std::string_view text, std::string source, const MpsOptions& opt)
      : text_(text), source_(std::move(source)), opt_(opt) {}

  MpsModel run();

 private:
  [[noreturn]] void fail(const std::string& what) const {
    throw MpsError(source_ + ":" + std::to_string(line_) + ": " + what);
  }
  void warn(std::string what) {
    if (m_.warnings.size() < kWarningCap) m_.warnings.push_back(std::move(what));
    else if (m_.warnings.size() == kWarningCap) m_.warnings.emplace_back("(further warnings suppressed)");
  }
  void tokens_of(std::string_view line) {
    if (fixed_) split_fixed(line, &tok_);
    else split_free(line, &tok_);
  }

  Int row_of(std::string_view name, bool* known) const {
    const Int* v = rows_.find(name);
    *known = v != nullptr;
    return v != nullptr ? *v : 0;
  }
  // Consecutive COLUMNS lines almost always name the same column, so the
  // last one looked up is checked before the map.
  Int column(std::string_view name) {
    if (last_col_ >= 0 && name == last_name_) return last_col_;
    const Int* v = cols_.find(name);
    last_name_ = name;
    if (v != nullptr) return last_col_ = *v;
    const Int j = static_cast<Int>(m_.col_names.size());
    last_col_ = j;
    cols_.insert(name, j);
    m_.col_names.emplace_back(name);
    m_.objective.push_back(0.0);
    m_.col_lower.push_back(0.0);
    m_.col_upper.push_back(kInf);
    m_.col_type.push_back(integer_ ? kInteger : kContinuous);
    has_lower_.push_back(0);
    has_upper_.push_back(0);
    return j;
  }

  void flush_rows();
  void on_row(const Tokens& t);
  void on_column(const Tokens& t);
  void on_rhs(const Tokens& t);
  void on_range(const Tokens& t);
  void on_bound(const Tokens& t);
  void on_quadratic(const Tokens& t, bool full);
  void finish();

  std::string_view text_;
  std::string source_;
  MpsOptions opt_;
  MpsModel m_;
  long long line_ = 0;
  bool fixed_ = false;
  bool objective_seen_ = false;
  bool integer_ = false;
  Tokens tok_;

  std::vector<std::string_view> held_rows_;
  std::vector<long long> held_lines_;

  NameTable rows_;
  NameTable cols_;
  std::string_view last_name_;
  Int last_col_ = -1;

  std::vector<char> sense_;  // 'E', 'L', 'G' per kept row
  std::vector<double> rhs_, range_;
  std::vector<char> has_rhs_, has_range_;
  std::vector<Entry> a_, q_;
  std::unordered_set<long long> q_seen_;
  bool q_both_ = false;
  bool q_full_ = false;
  std::vector<char> has_lower_, has_upper_;
  std::size_t negative_up_ = 0;
};

// ROWS lines are held until the section ends so that the format can be chosen
// from all of them: in free format every ROWS line is exactly "type name".
void Parser::flush_rows() {
  if (held_rows_.empty()) return;
  if (opt_.format == 2) {
    fixed_ = true;
  } else if (opt_.format == 1) {
    fixed_ = false;
  } else {
    fixed_ = false;
    Tokens t;
    for (std::string_view l : held_rows_) {
      split_free(l, &t);
      if (t.size() != 2) {
        fixed_ = true;
        break;
      }
    }
  }
  m_.fixed_format = fixed_;
  for (std::size_t i = 0; i < held_rows_.size(); ++i) {
    line_ = held_lines_[i];
    tokens_of(held_rows_[i]);
    on_row(tok_);
  }
  held_rows_.clear();
  held_lines_.clear();
}

void Parser::on_row(const Tokens& t) {
  if (t.size() < 2) fail("ROWS entry needs a type and a name");
  const std::string type = upper_copy(t[0]);
  const std::string_view name = t[1];
  if (rows_.contains(name)) {
    warn("duplicate row name \"" + std::string(name) + "\": later definition ignored");
    return;
  }
  if (type == "N") {
    // The first N row is the objective; later ones are free rows, dropped.
    if (!objective_seen_) {
      objective_seen_ = true;
      rows_.insert(name, kObjectiveRow);
    } else {
      rows_.insert(name, kDroppedRow);
      ++m_.free_rows_dropped;
    }
    return;
  }
  if (type != "L" && type != "G" && type != "E") fail("unknown row type \"" + std::string(t[0]) + "\"");
  rows_.insert(name, static_cast<Int>(sense_.size()));
  sense_.push_back(type[0]);
  rhs_.push_back(0.0);
  has_rhs_.push_back(0);
  range_.push_back(0.0);
  has_range_.push_back(0);
  m_.row_names.emplace_back(name);
}

void Parser::on_column(const Tokens& t) {
  // A MARKER line switches integer columns on or off (mps_reader.cpp:126-148).
  bool marker = false;
  for (std::string_view w : t)
    if (same_upper(w, "'MARKER'") || same_upper(w, "MARKER")) marker = true;
  if (marker) {
    bool known = false;
    for (std::string_view w : t) {
      if (same_upper(w, "'INTORG'") || same_upper(w, "INTORG")) {
        integer_ = true;
        known = true;
      } else if (same_upper(w, "'INTEND'") || same_upper(w, "INTEND")) {
        integer_ = false;
        known = true;
      }
    }
    if (!known) warn("MARKER line without INTORG/INTEND ignored");
    return;
  }
  if (t.size() < 3 || t.size() % 2 == 0)
    fail("COLUMNS entry must be a column name followed by name/value pairs");
  const Int j = column(t[0]);
  for (std::size_t k = 1; k + 1 < t.size(); k += 2) {
    bool known = false;
    const Int r = row_of(t[k], &known);
    if (!known) fail("unknown row name \"" + std::string(t[k]) + "\" in COLUMNS");
    double v = 0.0;
    if (!number(t[k + 1], &v)) fail("cannot parse value \"" + std::string(t[k + 1]) + "\" in COLUMNS");
    if (r == kDroppedRow) continue;
    if (r == kObjectiveRow) {
      m_.objective[uz(j)] += v;
      ++m_.objective_nonzeros;
    } else {
      a_.push_back(Entry{r, j, v});
    }
  }
}

// RHS and RANGES: an optional set name comes first exactly when the token
// count is odd.
void Parser::on_rhs(const Tokens& t) {
  const std::size_t first = t.size() % 2;
  if (t.size() < first + 2) fail("RHS entry must contain at least one name/value pair");
  for (std::size_t k = first; k + 1 < t.size(); k += 2) {
    bool known = false;
    const Int r = row_of(t[k], &known);
    if (!known) fail("unknown row name \"" + std::string(t[k]) + "\" in RHS");
    double v = 0.0;
    if (!number(t[k + 1], &v)) fail("cannot parse value \"" + std::string(t[k + 1]) + "\" in RHS");
    if (r == kDroppedRow) continue;
    if (r == kObjectiveRow) {
      m_.objective_offset = -v;  // the objective's RHS is its negated constant
      continue;
    }
    if (has_rhs_[uz(r)]) warn("duplicate RHS entry for row \"" + std::string(t[k]) + "\"");
    rhs_[uz(r)] = v;
    has_rhs_[uz(r)] = 1;
  }
}

void Parser::on_range(const Tokens& t) {
  const std::size_t first = t.size() % 2;
  if (t.size() < first + 2) fail("RANGES entry must contain at least one name/value pair");
  for (std::size_t k = first; k + 1 < t.size(); k += 2) {
    bool known = false;
    const Int r = row_of(t[k], &known);
    if (!known) fail("unknown row name \"" + std::string(t[k]) + "\" in RANGES");
    double v = 0.0;
    if (!number(t[k + 1], &v)) fail("cannot parse value \"" + std::string(t[k + 1]) + "\" in RANGES");
    if (r < 0) continue;
    if (has_range_[uz(r)]) warn("duplicate RANGES entry for row \"" + std::string(t[k]) + "\"");
    range_[uz(r)] = v;
    has_range_[uz(r)] = 1;
  }
}

void Parser::on_bound(const Tokens& t) {
  if (t.empty()) fail("empty BOUNDS entry");
  const std::string type = upper_copy(t[0]);
  const bool valued = type == "UP" || type == "LO" || type == "FX" || type == "LI" || type == "UI" ||
                      type == "SI" || type == "SC";
  const bool bare = type == "FR" || type == "MI" || type == "PL" || type == "BV";
  if (!valued && !bare) fail("unknown bound type \"" + std::string(t[0]) + "\"");

  // TYPE [set] column [value]: the token count places the column, except for
  // a valueless type with three tokens, where the second token is the column
  // if a column of that name already exists (mps_reader.cpp:422-447).
  std::size_t at = 0, val = 0;
  if (valued) {
    if (t.size() == 3) {
      at = 1;
      val = 2;
    } else if (t.size() == 4) {
      at = 2;
      val = 3;
    } else {
      fail("bound type " + type + " needs a column and a value");
    }
  } else if (t.size() == 2) {
    at = 1;
  } else if (t.size() == 3) {
    at = cols_.contains(t[1]) ? 1 : 2;
  } else if (t.size() == 4) {
    at = 2;
  } else {
    fail("bound type " + type + " needs a column name");
  }

  double v = 0.0;
  if (valued && !number(t[val], &v)) fail("cannot parse bound value \"" + std::string(t[val]) + "\"");
  if (valued && std::fabs(v) >= opt_.infinity_threshold) v = v > 0 ? kInf : -kInf;

  const std::size_t j = uz(column(t[at]));
  const bool lower = type == "LO" || type == "FX" || type == "MI" || type == "LI" || type == "BV" ||
                     type == "FR";
  const bool upper = type == "UP" || type == "FX" || type == "PL" || type == "UI" || type == "BV" ||
                     type == "FR" || type == "SI" || type == "SC";
  if ((lower && has_lower_[j]) || (upper && has_upper_[j])) {
    warn("duplicate bound for column \"" + std::string(t[at]) + "\": ignored");
    return;
  }

  if (type == "FR") {
    m_.col_lower[j] = -kInf;
    m_.col_upper[j] = kInf;
  } else if (type == "MI") {
    m_.col_lower[j] = -kInf;
  } else if (type == "PL") {
    m_.col_upper[j] = kInf;
  } else if (type == "BV") {
    m_.col_lower[j] = 0.0;
    m_.col_upper[j] = 1.0;
    m_.col_type[j] = kInteger;
  } else if (type == "LO") {
    m_.col_lower[j] = v;
  } else if (type == "UP") {
    m_.col_upper[j] = v;
    if (v < 0.0 && !has_lower_[j]) {
      ++negative_up_;
      if (opt_.negative_up_minus_infinity && m_.col_type[j] == kContinuous) {
        m_.col_lower[j] = -kInf;
        warn("negative UP bound on \"" + std::string(t[at]) +
             "\" with no lower bound: lower set to -infinity (CPLEX convention)");
      } else {
        warn("negative UP bound on \"" + std::string(t[at]) +
             "\" with no lower bound: lower kept at 0 (HiGHS convention)");
      }
    }
  } else if (type == "FX") {
    m_.col_lower[j] = v;
    m_.col_upper[j] = v;
  } else if (type == "LI") {
    m_.col_lower[j] = v;
    m_.col_type[j] = kInteger;
  } else if (type == "UI") {
    m_.col_upper[j] = v;
    m_.col_type[j] = kInteger;
  } else if (type == "SI") {
    m_.col_upper[j] = v;
    m_.col_type[j] = kSemiInteger;
  } else if (type == "SC") {
    m_.col_upper[j] = v;
    m_.col_type[j] = kSemiContinuous;
  }
  if (lower) has_lower_[j] = 1;
  if (upper) has_upper_[j] = 1;
}

void Parser::on_quadratic(const Tokens& t, bool full) {
  if (t.size() < 3 || t.size() % 2 == 0)
    fail("quadratic entry must be a column name followed by name/value pairs");
  if (full) q_full_ = true;
  const Int i = column(t[0]);
  for (std::size_t k = 1; k + 1 < t.size(); k += 2) {
    const Int j = column(t[k]);
    double v = 0.0;
    if (!number(t[k + 1], &v)) fail("cannot parse quadratic value \"" + std::string(t[k + 1]) + "\"");
    q_.push_back(Entry{i, j, v});
    if (i != j) {
      const long long key = (static_cast<long long>(i) << 32) + j;
      const long long mirror = (static_cast<long long>(j) << 32) + i;
      if (q_seen_.count(mirror) != 0) q_both_ = true;
      q_seen_.insert(key);
    }
  }
}

void Parser::finish() {
  const Int n = static_cast<Int>(m_.col_names.size());
  const Int rows = static_cast<Int>(sense_.size());
  m_.a = Csr::from_triplets(rows, n, std::move(a_));
  m_.row_lower.assign(uz(rows), 0.0);
  m_.row_upper.assign(uz(rows), 0.0);
  for (std::size_t i = 0; i < uz(rows); ++i) {
    const double b = rhs_[i];
    const char s = sense_[i];
    if (s == 'E') {
      m_.row_lower[i] = b;
      m_.row_upper[i] = b;
    } else if (s == 'L') {
      m_.row_lower[i] = -kInf;
      m_.row_upper[i] = b;
    } else {
      m_.row_lower[i] = b;
      m_.row_upper[i] = kInf;
    }
    if (!has_range_[i]) continue;
    // HiGHS's RANGES rule: the sign matters only on an E row.
    const double r = range_[i];
    if ((s == 'E' && r < 0.0) || s == 'L') m_.row_lower[i] = m_.row_upper[i] - std::fabs(r);
    else if ((s == 'E' && r > 0.0) || s == 'G') m_.row_upper[i] = m_.row_lower[i] + std::fabs(r);
  }

  if (!q_.empty()) {
    std::vector<Entry> q;
    if (!q_full_ && !q_both_) {
      q.reserve(q_.size() * 2);
      for (const Entry& e : q_) {
        q.push_back(e);
        if (e.row != e.col) q.push_back(Entry{e.col, e.row, e.value});
      }
    } else {
      q = std::move(q_);
    }
    m_.hessian = Csr::from_triplets(n, n, std::move(q));
  } else {
    m_.hessian = Csr::from_triplets(0, 0, {});
  }

  if (negative_up_ > 0)
    warn("negative UP bounds seen on " + std::to_string(negative_up_) + " column(s); reader is using the " +
         std::string(opt_.negative_up_minus_infinity ? "CPLEX" : "HiGHS") + " convention");

  std::string why;
  if (!validate_model(m_, &why)) fail("model failed validation: " + why);
  m_.fill_coo();
}

MpsModel Parser::run() {
  Part part = Part::kNone;
  std::size_t pos = 0;
  bool ended = false;
  // Lines as std::getline yields them: split at '\n', a final line without
  // one still counts, and one trailing '\r' is dropped.
  while (pos < text_.size() && !ended) {
    const std::size_t nl = text_.find('\n', pos);
    std::string_view line = text_.substr(pos, nl == std::string_view::npos ? std::string_view::npos : nl - pos);
    pos = nl == std::string_view::npos ? text_.size() : nl + 1;
    ++line_;
    if (!line.empty() && line.back() == '\r') line.remove_suffix(1);
    if (line.empty() || line[0] == '*') continue;
    const bool header = !blank(line[0]);
    if (!header && part == Part::kRows) {
      held_rows_.push_back(line);
      held_lines_.push_back(line_);
      continue;
    }
    if (header) split_free(line, &tok_);
    else tokens_of(line);
    if (tok_.empty()) continue;

    if (header) {
      const Part next = part_of(tok_[0]);
      if (next == Part::kNone) fail("unknown section \"" + std::string(tok_[0]) + "\"");
      const long long at = line_;
      if (part == Part::kRows) {
        const Tokens keep = tok_;
        flush_rows();
        tok_ = keep;
      }
      line_ = at;
      part = next;
      if (part == Part::kName && tok_.size() >= 2) m_.name = std::string(tok_[1]);
      if (part == Part::kObjSense && tok_.size() >= 2) m_.maximize = is_max(tok_[1]);
      if (part == Part::kEndata) ended = true;
      continue;
    }
    switch (part) {
      case Part::kObjSense: m_.maximize = is_max(tok_[0]); break;
      case Part::kColumns: on_column(tok_); break;
      case Part::kRhs: on_rhs(tok_); break;
      case Part::kRanges: on_range(tok_); break;
      case Part::kBounds: on_bound(tok_); break;
      case Part::kQuadObj: on_quadratic(tok_, false); break;
      case Part::kQMatrix: on_quadratic(tok_, true); break;
      case Part::kNone: fail("data line before any section header");
      case Part::kRows:
      case Part::kName:
      case Part::kEndata: break;
    }
  }
  flush_rows();
  if (!objective_seen_) warn("no N row found: the objective is taken to be zero");
  finish();
  return std::move(m_);
}

// ---- writer ---------------------------------------------------------------

void put_number(std::string* out, double v) {
  char buf[64];
  const auto r = std::to_chars(buf, buf + sizeof(buf), v);  // shortest form that reads back exactly
  out->append(buf, r.ptr);
}

void check_name(std::string_view name, const char* what) {
  if (name.empty()) throw std::invalid_argument(std::string(what) + " name is empty");
  for (char c : name)
    if (blank(c)) throw std::invalid_argument(std::string(what) + " name \"" + std::string(name) + "\" holds white space");
  if (same_upper(name, "MARKER") || same_upper(name, "'MARKER'"))
    throw std::invalid_argument(std::string(what) + " name \"" + std::string(name) + "\" reads as a MARKER line");
}

std::string unique_name(const std::unordered_set<std::string>& taken, const std::string& base) {
  std::string name = base;
  for (int k = 1; taken.count(name) != 0; ++k) name = base + "_" + std::to_string(k);
  return name;
}

}  // namespace

// The model's own consistency checks (model.cpp:15-62), in the same order.
bool validate_model(const MpsModel& m, std::string* error) {
  auto bad = [&](const std::string& why) {
    if (error) *error = why;
    return false;
  };
  std::string csr_error;
  if (!m.a.validate(&csr_error)) return bad(csr_error);
  const std::size_t n = uz(m.a.cols), rows = uz(m.a.rows);
  if (m.objective.size() != n) return bad("objective length does not match columns");
  if (m.col_lower.size() != n || m.col_upper.size() != n)
    return bad("column bound length does not match columns");
  if (m.col_type.size() != n) return bad("column type length does not match columns");
  if (m.row_lower.size() != rows || m.row_upper.size() != rows)
    return bad("row bound length does not match rows");
  if (!m.row_names.empty() && m.row_names.size() != rows) return bad("row name length does not match rows");
  if (!m.col_names.empty() && m.col_names.size() != n) return bad("column name length does not match columns");
  for (std::size_t j = 0; j < n; ++j) {
    if (std::isnan(m.col_lower[j]) || std::isnan(m.col_upper[j]))
      return bad("NaN column bound at column " + std::to_string(j));
    if (m.col_lower[j] > m.col_upper[j]) return bad("empty column bound interval at column " + std::to_string(j));
    if (std::isnan(m.objective[j])) return bad("NaN objective coefficient at column " + std::to_string(j));
  }
  for (std::size_t i = 0; i < rows; ++i) {
    if (std::isnan(m.row_lower[i]) || std::isnan(m.row_upper[i]))
      return bad("NaN row bound at row " + std::to_string(i));
    if (m.row_lower[i] > m.row_upper[i]) return bad("empty row bound interval at row " + std::to_string(i));
  }
  if (m.hessian.nnz() > 0) {
    if (m.hessian.rows != m.a.cols || m.hessian.cols != m.a.cols)
      return bad("hessian is not square in the number of columns");
    const Csr qt = m.hessian.transpose();
    if (qt.start != m.hessian.start || qt.index != m.hessian.index)
      return bad("hessian sparsity pattern is not symmetric");
    for (std::size_t k = 0; k < qt.value.size(); ++k)
      if (qt.value[k] != m.hessian.value[k]) return bad("hessian is not symmetric");
  }
  return true;
}

void MpsModel::fill_coo() {
  coo_row.resize(a.value.size());
  coo_col = a.index;
  for (Int r = 0; r < a.rows; ++r)
    std::fill(coo_row.begin() + a.start[uz(r)], coo_row.begin() + a.start[uz(r) + 1], r);
}

MpsModel read_mps_text(std::string_view text, const std::string& source, const MpsOptions& opt) {
  return Parser(text, source, opt).run();
}

MpsModel read_mps_file(const std::string& path, const MpsOptions& opt) {
  std::FILE* f = std::fopen(path.c_str(), "rb");
  if (f == nullptr) throw FileError("cannot open \"" + path + "\"");
  std::string text;
  char buf[1 << 16];
  for (std::size_t got; (got = std::fread(buf, 1, sizeof(buf), f)) > 0;) text.append(buf, got);
  const bool failed = std::ferror(f) != 0;
  std::fclose(f);
  if (failed) throw FileError("cannot read \"" + path + "\"");
  return read_mps_text(text, path, opt);
}

std::string write_mps_text(const MpsWriteModel& m) {
  if (m.rows < 0 || m.cols < 0) throw std::invalid_argument("negative model size");
  const std::size_t n = uz(m.cols), rows = uz(m.rows);
  std::vector<std::string> rname(rows), cname(n);
  std::unordered_set<std::string> rtaken, ctaken;
  for (std::size_t i = 0; i < rows; ++i) {
    rname[i] = m.row_names ? std::string(m.row_names[i]) : "R" + std::to_string(i);
    check_name(rname[i], "row");
    if (!rtaken.insert(rname[i]).second) throw std::invalid_argument("duplicate row name \"" + rname[i] + "\"");
  }
  for (std::size_t j = 0; j < n; ++j) {
    cname[j] = m.col_names ? std::string(m.col_names[j]) : "C" + std::to_string(j);
    check_name(cname[j], "column");
    if (!ctaken.insert(cname[j]).second) throw std::invalid_argument("duplicate column name \"" + cname[j] + "\"");
  }
  const std::string obj = unique_name(rtaken, "OBJ");
  // A valueless bound with a set name is read as "TYPE column" when the set
  // name is itself a column, so the set name must not be one.
  const std::string bnd = unique_name(ctaken, "BND");
  std::string title = m.name.empty() ? std::string("MODEL") : m.name;
  for (char& c : title)
    if (blank(c)) c = '_';

  for (std::size_t j = 0; j < n; ++j) {
    if (std::isnan(m.objective[j]) || std::isnan(m.col_lower[j]) || std::isnan(m.col_upper[j]))
      throw std::invalid_argument("NaN column data at column " + std::to_string(j));
    if (m.col_lower[j] > m.col_upper[j])
      throw std::invalid_argument("empty column bound interval at column " + std::to_string(j));
  }
  for (std::size_t i = 0; i < rows; ++i) {
    if (std::isnan(m.row_lower[i]) || std::isnan(m.row_upper[i]))
      throw std::invalid_argument("NaN row bound at row " + std::to_string(i));
    if (m.row_lower[i] > m.row_upper[i])
      throw std::invalid_argument("empty row bound interval at row " + std::to_string(i));
  }

  // Columns of A: a counting sort of the CSR by column keeps rows ascending.
  const Int nnz = m.start[rows];
  std::vector<Int> cstart(n + 1, 0), crow(uz(nnz));
  std::vector<double> cval(uz(nnz));
  for (Int k = 0; k < nnz; ++k) {
    if (std::isnan(m.value[k])) throw std::invalid_argument("NaN matrix value");
    ++cstart[uz(m.index[k]) + 1];
  }
  for (std::size_t j = 0; j < n; ++j) cstart[j + 1] += cstart[j];
  {
    std::vector<Int> fill(cstart.begin(), cstart.end() - 1);
    for (std::size_t r = 0; r < rows; ++r)
      for (Int k = m.start[r]; k < m.start[r + 1]; ++k) {
        const Int at = fill[uz(m.index[k])]++;
        crow[uz(at)] = static_cast<Int>(r);
        cval[uz(at)] = m.value[k];
      }
  }

  std::string out;
  out.reserve(64 * (uz(nnz) + n + rows) + 256);
  out += "NAME          " + title + "\n";
  if (m.maximize) out += "OBJSENSE\n    MAX\n";
  out += "ROWS\n N  " + obj + "\n";
  // Row kinds. A ranged row is written as L (upper, range) or G (lower,
  // range), whichever the reader turns back into both bounds exactly; a
  // range next to upper - lower is searched when neither does.
  std::vector<char> kind(rows);
  std::vector<double> rhs(rows, 0.0), rng(rows, 0.0);
  std::vector<char> ranged(rows, 0);
  for (std::size_t i = 0; i < rows; ++i) {
    const double lo = m.row_lower[i], hi = m.row_upper[i];
    if (lo == hi) {
      kind[i] = 'E';
      rhs[i] = lo;
    } else if (lo == -kInf && hi == kInf) {
      kind[i] = 'N';
    } else if (lo == -kInf) {
      kind[i] = 'L';
      rhs[i] = hi;
    } else if (hi == kInf) {
      kind[i] = 'G';
      rhs[i] = lo;
    } else {
      ranged[i] = 1;
      // Candidates: upper - lower, then its neighbours up to 8 ulps away.
      const double r0 = hi - lo;
      std::vector<double> cands{r0};
      double below = r0, above = r0;
      for (int s = 0; s < 8; ++s) {
        above = std::nextafter(above, kInf);
        below = std::nextafter(below, 0.0);
        cands.push_back(above);
        cands.push_back(below);
      }
      kind[i] = 'L', rhs[i] = hi, rng[i] = r0;
      for (double r : cands) {
        if (hi - r == lo) {
          kind[i] = 'L', rhs[i] = hi, rng[i] = r;
          break;
        }
        if (lo + r == hi) {
          kind[i] = 'G', rhs[i] = lo, rng[i] = r;
          break;
        }
      }
    }
    out += ' ';
    out += kind[i];
    out += "  " + rname[i] + "\n";
  }

  out += "COLUMNS\n";
  bool in_int = false;
  for (std::size_t j = 0; j < n; ++j) {
    const bool integer = m.col_type && m.col_type[j] == kInteger;
    if (integer != in_int) {
      out += integer ? "    MARKER                 'MARKER'                 'INTORG'\n"
                     : "    MARKER                 'MARKER'                 'INTEND'\n";
      in_int = integer;
    }
    int on_line = 0;
    auto pair = [&](const std::string& row, double v) {
      if (on_line == 0) out += "    " + cname[j];
      out += "  " + row + "  ";
      put_number(&out, v);
      if (++on_line == 2) {
        out += '\n';
        on_line = 0;
      }
    };
    const bool empty = cstart[j] == cstart[j + 1];
    if (m.objective[j] != 0.0 || empty) pair(obj, m.objective[j]);
    for (Int k = cstart[j]; k < cstart[j + 1]; ++k) pair(rname[uz(crow[uz(k)])], cval[uz(k)]);
    if (on_line != 0) out += '\n';
  }
  if (in_int) out += "    MARKER                 'MARKER'                 'INTEND'\n";

  out += "RHS\n";
  for (std::size_t i = 0; i < rows; ++i) {
    if (kind[i] == 'N' || (rhs[i] == 0.0 && !std::signbit(rhs[i]))) continue;
    out += "    RHS  " + rname[i] + "  ";
    put_number(&out, rhs[i]);
    out += '\n';
  }
  if (m.objective_offset != 0.0 || std::signbit(m.objective_offset)) {
    out += "    RHS  " + obj + "  ";
    put_number(&out, -m.objective_offset);
    out += '\n';
  }
  bool any_range = false;
  for (std::size_t i = 0; i < rows; ++i) {
    if (!ranged[i]) continue;
    if (!any_range) out += "RANGES\n";
    any_range = true;
    out += "    RNG  " + rname[i] + "  ";
    put_number(&out, rng[i]);
    out += '\n';
  }

  out += "BOUNDS\n";
  auto bound = [&](const char* type, std::size_t j, const double* v) {
    out += ' ';
    out += type;
    out += ' ' + bnd + "  " + cname[j];
    if (v) {
      out += "  ";
      put_number(&out, *v);
    }
    out += '\n';
  };
  for (std::size_t j = 0; j < n; ++j) {
    const double lo = m.col_lower[j], hi = m.col_upper[j];
    const std::int32_t type = m.col_type ? m.col_type[j] : kContinuous;
    const bool default_lo = lo == 0.0 && !std::signbit(lo);
    if (type == kSemiContinuous || type == kSemiInteger) {
      if (lo == -kInf) bound("MI", j, nullptr);
      else if (!default_lo) bound("LO", j, &lo);
      bound(type == kSemiContinuous ? "SC" : "SI", j, &hi);
    } else if (lo == hi) {
      bound("FX", j, &lo);
    } else if (lo == -kInf && hi == kInf) {
      bound("FR", j, nullptr);
    } else {
      if (lo == -kInf) bound("MI", j, nullptr);
      else if (!default_lo) bound("LO", j, &lo);
      if (hi != kInf) bound("UP", j, &hi);
    }
  }
  out += "ENDATA\n";
  return out;
}

}  // namespace anuk
