// MCLS host facade: error convention, shared-pointer alias, parameter list,
// and the host CSR matrix / vector types the traits are specialised for.
//
// Mirrors (paths relative to the MCLS source root):
//   MCLS::Assertion, MCLS_INSIST        src/MCLS_DBC.hpp:59-162
//   Teuchos::RCP / ParameterList usage  throughout src/ (RCP -> std::shared_ptr)
//   Tpetra::CrsMatrix insert/fillComplete semantics used by the tests
//     (tpetra/test/tstTpetraMCSASolverManager.cpp:83-116: duplicate inserts sum)
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <variant>
#include <vector>

namespace MCLS {

template <class T>
using RCP = std::shared_ptr<T>;

// src/MCLS_DBC.hpp:59-97: contract violations throw MCLS::Assertion.
class Assertion : public std::logic_error {
 public:
  explicit Assertion(const std::string& msg) : std::logic_error(msg) {}
};

#define MCLS_INSIST(c, msg)                                                        \
  do {                                                                             \
    if (!(c)) throw ::MCLS::Assertion(std::string("MCLS assertion: ") + (msg));    \
  } while (0)

// Typed key/value list standing in for Teuchos::ParameterList (get<T>,
// set<T>, isParameter).  Values keep the type they were set with; get<T>
// with another type throws, as Teuchos does.
class ParameterList {
 public:
  using Value = std::variant<int, int64_t, double, bool, std::string>;

  template <class T>
  ParameterList& set(const std::string& key, const T& v) {
    if constexpr (std::is_same_v<T, const char*> || std::is_array_v<T>)
      d_values[key] = std::string(v);
    else
      d_values[key] = Value(v);
    return *this;
  }
  ParameterList& set(const std::string& key, const char* v) {
    d_values[key] = std::string(v);
    return *this;
  }
  bool isParameter(const std::string& key) const { return d_values.count(key) != 0; }
  template <class T>
  T get(const std::string& key) const {
    auto it = d_values.find(key);
    MCLS_INSIST(it != d_values.end(), "parameter '" + key + "' not set");
    const T* p = std::get_if<T>(&it->second);
    MCLS_INSIST(p != nullptr, "parameter '" + key + "' has another type");
    return *p;
  }
  template <class T>
  T get(const std::string& key, const T& def) const {
    return isParameter(key) ? get<T>(key) : def;
  }
  const std::map<std::string, Value>& values() const { return d_values; }

 private:
  std::map<std::string, Value> d_values;
};

inline RCP<ParameterList> parameterList() { return std::make_shared<ParameterList>(); }

// Host vector (the build's Tpetra::Vector<double> counterpart).
class Vector {
 public:
  Vector() = default;
  explicit Vector(int64_t n, double v = 0.0) : d_data((size_t)n, v) {}
  int64_t size() const { return (int64_t)d_data.size(); }
  double* data() { return d_data.data(); }
  const double* data() const { return d_data.data(); }
  double& operator[](int64_t i) { return d_data[(size_t)i]; }
  const double& operator[](int64_t i) const { return d_data[(size_t)i]; }
  std::vector<double>& raw() { return d_data; }
  const std::vector<double>& raw() const { return d_data; }

 private:
  std::vector<double> d_data;
};

// Host CSR matrix (the build's Tpetra::CrsMatrix counterpart).  Rows are
// assembled with insertGlobalValues; fillComplete sorts each row and sums
// duplicate entries.  Scalar is double or float (stored values).
template <class Scalar = double>
class CrsMatrix {
 public:
  using scalar_type = Scalar;
  explicit CrsMatrix(int64_t n) : d_n(n), d_rows((size_t)n) {}
  CrsMatrix(int64_t n, std::vector<int64_t> rp, std::vector<int32_t> col, std::vector<Scalar> val)
      : d_n(n), d_rp(std::move(rp)), d_col(std::move(col)), d_val(std::move(val)), d_filled(true) {
    MCLS_INSIST((int64_t)d_rp.size() == n + 1, "row pointer length");
    MCLS_INSIST(d_rp.back() == (int64_t)d_col.size() && d_col.size() == d_val.size(), "CSR sizes");
  }

  void insertGlobalValues(int64_t row, const std::vector<int64_t>& cols, const std::vector<Scalar>& vals) {
    MCLS_INSIST(!d_filled, "matrix already fill-complete");
    MCLS_INSIST(row >= 0 && row < d_n && cols.size() == vals.size(), "bad insert");
    for (size_t k = 0; k < cols.size(); ++k) d_rows[(size_t)row].emplace_back(cols[k], vals[k]);
  }
  void fillComplete() {
    if (d_filled) return;
    d_rp.assign((size_t)d_n + 1, 0);
    d_col.clear();
    d_val.clear();
    for (int64_t i = 0; i < d_n; ++i) {
      auto& r = d_rows[(size_t)i];
      std::sort(r.begin(), r.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
      for (size_t k = 0; k < r.size(); ++k) {
        MCLS_INSIST(r[k].first >= 0 && r[k].first < d_n, "column out of range");
        if (!d_col.empty() && d_rp[(size_t)i] < (int64_t)d_col.size() && d_col.back() == r[k].first)
          d_val.back() += r[k].second;
        else {
          d_col.push_back((int32_t)r[k].first);
          d_val.push_back(r[k].second);
        }
      }
      d_rp[(size_t)i + 1] = (int64_t)d_col.size();
    }
    d_rows.clear();
    d_rows.shrink_to_fit();
    d_filled = true;
  }
  bool isFillComplete() const { return d_filled; }
  int64_t numRows() const { return d_n; }
  int64_t nnz() const { return d_filled ? d_rp.back() : 0; }
  const std::vector<int64_t>& rowPtr() const { return d_rp; }
  const std::vector<int32_t>& colInd() const { return d_col; }
  const std::vector<Scalar>& values() const { return d_val; }
  std::vector<Scalar>& valuesNonConst() { return d_val; }
  // replace the (fill-complete) contents with another CSR of the same rows
  void setCrs(std::vector<int64_t> rp, std::vector<int32_t> col, std::vector<Scalar> val) {
    MCLS_INSIST((int64_t)rp.size() == d_n + 1 && rp.back() == (int64_t)col.size() && col.size() == val.size(),
                "CSR sizes");
    d_rp = std::move(rp);
    d_col = std::move(col);
    d_val = std::move(val);
    d_rows.clear();
    d_filled = true;
  }
  int64_t maxRowEntries() const {
    int64_t m = 0;
    for (int64_t i = 0; i < d_n; ++i) m = std::max(m, d_rp[(size_t)i + 1] - d_rp[(size_t)i]);
    return m;
  }

 private:
  int64_t d_n;
  std::vector<std::vector<std::pair<int64_t, Scalar>>> d_rows;
  std::vector<int64_t> d_rp;
  std::vector<int32_t> d_col;
  std::vector<Scalar> d_val;
  bool d_filled = false;
};

}  // namespace MCLS
