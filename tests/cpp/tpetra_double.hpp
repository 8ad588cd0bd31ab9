// A test double of the part of Trilinos' Teuchos / Tpetra / Kokkos member
// API that mcls_amd/host/MCLS_TpetraAdapter.hpp calls, for one process and
// host memory, so that the adapter is compiled and exercised where Trilinos
// is not installed.  Names and member signatures follow Tpetra's (Trilinos
// 12); the behaviour is written from Tpetra's documented semantics, with one
// deliberate difference from a real one-process Tpetra matrix: the column
// map lists the columns in order of first appearance, not ascending, so the
// adapter's local-to-global column mapping is really tested.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <random>
#include <stdexcept>
#include <vector>

namespace Teuchos {

template <class T>
using RCP = std::shared_ptr<T>;
template <class T>
RCP<T> rcp(T* p) {
  return RCP<T>(p);
}

enum DataAccess { Copy, View };
enum ETransp { NO_TRANS, TRANS, CONJ_TRANS };

template <class T>
class ArrayView {
 public:
  ArrayView() = default;
  ArrayView(T* p, int n) : d_p(p), d_n(n) {}
  T& operator[](int i) const { return d_p[i]; }
  int size() const { return d_n; }
  T* getRawPtr() const { return d_p; }
  T* begin() const { return d_p; }
  T* end() const { return d_p + d_n; }

 private:
  T* d_p = nullptr;
  int d_n = 0;
};

// a non-owning view of a vector's entries (Tpetra's getData on a host node)
template <class T>
class ArrayRCP {
 public:
  ArrayRCP(T* p, size_t n) : d_p(p), d_n(n) {}
  T* getRawPtr() const { return d_p; }
  size_t size() const { return d_n; }
  T& operator[](size_t i) const { return d_p[i]; }

 private:
  T* d_p;
  size_t d_n;
};

template <class Ordinal>
class Comm {
 public:
  virtual ~Comm() = default;
  virtual int getRank() const = 0;
  virtual int getSize() const = 0;
};

template <class Ordinal>
class SerialComm : public Comm<Ordinal> {
 public:
  int getRank() const override { return 0; }
  int getSize() const override { return 1; }
};

}  // namespace Teuchos

namespace Tpetra {

// the global indices this process owns, in local order
template <class LO = int, class GO = long long, class NT = void>
class Map {
 public:
  // a contiguous map of num_global indices from index_base (one process)
  Map(size_t num_global, GO index_base, const Teuchos::RCP<const Teuchos::Comm<int>>& comm) : d_comm(comm) {
    d_gids.resize(num_global);
    for (size_t i = 0; i < num_global; ++i) d_gids[i] = index_base + (GO)i;
    index();
  }
  // the given global indices in the given order
  Map(const std::vector<GO>& gids, const Teuchos::RCP<const Teuchos::Comm<int>>& comm) : d_gids(gids), d_comm(comm) {
    index();
  }
  Teuchos::RCP<const Teuchos::Comm<int>> getComm() const { return d_comm; }
  size_t getNodeNumElements() const { return d_gids.size(); }
  size_t getGlobalNumElements() const { return d_gids.size(); }
  GO getMinGlobalIndex() const { return d_gids.empty() ? 0 : *std::min_element(d_gids.begin(), d_gids.end()); }
  bool isContiguous() const {
    for (size_t i = 1; i < d_gids.size(); ++i)
      if (d_gids[i] != d_gids[i - 1] + 1) return false;
    return true;
  }
  GO getGlobalElement(LO lid) const { return isNodeLocalElement(lid) ? d_gids[(size_t)lid] : (GO)-1; }
  LO getLocalElement(GO gid) const {
    auto it = d_lid.find(gid);
    return it == d_lid.end() ? (LO)-1 : it->second;
  }
  bool isNodeGlobalElement(GO gid) const { return d_lid.count(gid) != 0; }
  bool isNodeLocalElement(LO lid) const { return lid >= 0 && (size_t)lid < d_gids.size(); }
  Teuchos::ArrayView<const GO> getNodeElementList() const {
    return Teuchos::ArrayView<const GO>(d_gids.data(), (int)d_gids.size());
  }

 private:
  void index() {
    for (size_t i = 0; i < d_gids.size(); ++i) d_lid[d_gids[i]] = (LO)i;
  }
  std::vector<GO> d_gids;
  std::map<GO, LO> d_lid;
  Teuchos::RCP<const Teuchos::Comm<int>> d_comm;
};

template <class S = double, class LO = int, class GO = long long, class NT = void>
class Vector {
 public:
  using map_type = Map<LO, GO, NT>;
  explicit Vector(const Teuchos::RCP<const map_type>& map) : d_map(map), d_v(map->getNodeNumElements(), S(0)) {}
  Vector(const Vector& src, Teuchos::DataAccess) : d_map(src.d_map), d_v(src.d_v) {}
  Teuchos::RCP<const map_type> getMap() const { return d_map; }
  size_t getLocalLength() const { return d_v.size(); }
  size_t getGlobalLength() const { return d_v.size(); }
  Teuchos::ArrayRCP<const S> getData() const { return Teuchos::ArrayRCP<const S>(d_v.data(), d_v.size()); }
  Teuchos::ArrayRCP<S> getDataNonConst() { return Teuchos::ArrayRCP<S>(d_v.data(), d_v.size()); }
  void putScalar(S a) { std::fill(d_v.begin(), d_v.end(), a); }
  void randomize() {
    std::mt19937_64 g(0x7e7);
    std::uniform_real_distribution<S> u(-1, 1);
    for (auto& x : d_v) x = u(g);
  }
  S dot(const Vector& b) const {
    S s = 0;
    for (size_t i = 0; i < d_v.size(); ++i) s += d_v[i] * b.d_v[i];
    return s;
  }
  S norm1() const {
    S s = 0;
    for (S x : d_v) s += std::fabs(x);
    return s;
  }
  S norm2() const { return std::sqrt(dot(*this)); }
  S normInf() const {
    S s = 0;
    for (S x : d_v) s = std::max(s, std::fabs(x));
    return s;
  }
  S meanValue() const {
    S s = 0;
    for (S x : d_v) s += x;
    return d_v.empty() ? S(0) : s / (S)d_v.size();
  }
  void abs(const Vector& b) {
    for (size_t i = 0; i < d_v.size(); ++i) d_v[i] = std::fabs(b.d_v[i]);
  }
  void scale(S a) {
    for (auto& x : d_v) x *= a;
  }
  void scale(S a, const Vector& b) {
    for (size_t i = 0; i < d_v.size(); ++i) d_v[i] = a * b.d_v[i];
  }
  void reciprocal(const Vector& b) {
    for (size_t i = 0; i < d_v.size(); ++i) d_v[i] = S(1) / b.d_v[i];
  }
  // this = beta this + alpha a (this is not read when beta = 0)
  void update(S alpha, const Vector& a, S beta) {
    for (size_t i = 0; i < d_v.size(); ++i) d_v[i] = (beta == S(0) ? S(0) : beta * d_v[i]) + alpha * a.d_v[i];
  }
  // this = gamma this + alpha a + beta b
  void update(S alpha, const Vector& a, S beta, const Vector& b, S gamma) {
    for (size_t i = 0; i < d_v.size(); ++i)
      d_v[i] = (gamma == S(0) ? S(0) : gamma * d_v[i]) + alpha * a.d_v[i] + beta * b.d_v[i];
  }
  // this = scalar_this this + scalar_ab a .* b
  void elementWiseMultiply(S scalar_ab, const Vector& a, const Vector& b, S scalar_this) {
    for (size_t i = 0; i < d_v.size(); ++i) d_v[i] = scalar_this * d_v[i] + scalar_ab * a.d_v[i] * b.d_v[i];
  }
  void replaceGlobalValue(GO row, S a) { d_v[(size_t)d_map->getLocalElement(row)] = a; }
  void replaceLocalValue(LO row, S a) { d_v[(size_t)row] = a; }
  void sumIntoGlobalValue(GO row, S a) { d_v[(size_t)d_map->getLocalElement(row)] += a; }
  void sumIntoLocalValue(LO row, S a) { d_v[(size_t)row] += a; }

 private:
  Teuchos::RCP<const map_type> d_map;
  std::vector<S> d_v;
};

// Kokkos-shaped views of the local CRS arrays (getLocalMatrix())
template <class T>
struct HostView {
  const T* p = nullptr;
  const T* data() const { return p; }
};
template <class S, class LO>
struct LocalCrsMatrix {
  struct {
    HostView<size_t> row_map;
    HostView<LO> entries;
  } graph;
  HostView<S> values;
};

template <class S = double, class LO = int, class GO = long long, class NT = void>
class CrsMatrix {
 public:
  using map_type = Map<LO, GO, NT>;
  using vector_type = Vector<S, LO, GO, NT>;
  using local_matrix_type = LocalCrsMatrix<S, LO>;

  CrsMatrix(const Teuchos::RCP<const map_type>& row_map, size_t) : d_row_map(row_map), d_rows(row_map->getNodeNumElements()) {}
  void insertGlobalValues(GO row, const Teuchos::ArrayView<const GO>& cols, const Teuchos::ArrayView<const S>& vals) {
    if (d_filled) throw std::runtime_error("insertGlobalValues after fillComplete");
    auto& r = d_rows[(size_t)d_row_map->getLocalElement(row)];
    for (int k = 0; k < cols.size(); ++k) r.emplace_back(cols[k], vals[k]);
  }
  void fillComplete() { fillComplete(d_row_map, d_row_map); }
  // sums duplicates; local columns numbered in order of first appearance
  void fillComplete(const Teuchos::RCP<const map_type>& domain, const Teuchos::RCP<const map_type>& range) {
    d_domain = domain;
    d_range = range;
    std::vector<GO> colg;
    std::map<GO, LO> seen;
    for (auto& r : d_rows)
      for (auto& e : r)
        if (!seen.count(e.first)) {
          seen[e.first] = (LO)colg.size();
          colg.push_back(e.first);
        }
    d_col_map = std::make_shared<map_type>(colg, d_row_map->getComm());
    d_rp.assign(1, 0);
    d_lcol.clear();
    d_val.clear();
    for (auto& r : d_rows) {
      std::map<LO, S> row;  // sorted by local column, duplicates summed
      for (auto& e : r) row[seen[e.first]] += e.second;
      for (auto& e : row) {
        d_lcol.push_back(e.first);
        d_val.push_back(e.second);
      }
      d_rp.push_back(d_lcol.size());
    }
    d_filled = true;
  }
  bool isFillComplete() const { return d_filled; }
  Teuchos::RCP<const map_type> getRowMap() const { return d_row_map; }
  Teuchos::RCP<const map_type> getColMap() const { return d_col_map; }
  Teuchos::RCP<const map_type> getDomainMap() const { return d_domain; }
  Teuchos::RCP<const map_type> getRangeMap() const { return d_range; }
  Teuchos::RCP<const Teuchos::Comm<int>> getComm() const { return d_row_map->getComm(); }
  size_t getGlobalNumRows() const { return d_rows.size(); }
  size_t getNodeNumRows() const { return d_rows.size(); }
  size_t getNodeNumEntries() const { return d_lcol.size(); }
  size_t getGlobalMaxNumRowEntries() const {
    size_t m = 0;
    for (size_t i = 0; i + 1 < d_rp.size(); ++i) m = std::max(m, d_rp[i + 1] - d_rp[i]);
    return m;
  }
  size_t getNumEntriesInLocalRow(LO row) const { return d_rp[(size_t)row + 1] - d_rp[(size_t)row]; }
  size_t getNumEntriesInGlobalRow(GO row) const { return getNumEntriesInLocalRow(d_row_map->getLocalElement(row)); }
  void getGlobalRowCopy(GO row, const Teuchos::ArrayView<GO>& cols, const Teuchos::ArrayView<S>& vals,
                        size_t& n) const {
    const LO lr = d_row_map->getLocalElement(row);
    n = getNumEntriesInLocalRow(lr);
    for (size_t k = 0; k < n; ++k) {
      cols[(int)k] = d_col_map->getGlobalElement(d_lcol[d_rp[(size_t)lr] + k]);
      vals[(int)k] = d_val[d_rp[(size_t)lr] + k];
    }
  }
  void getLocalRowCopy(LO row, const Teuchos::ArrayView<LO>& cols, const Teuchos::ArrayView<S>& vals,
                       size_t& n) const {
    n = getNumEntriesInLocalRow(row);
    for (size_t k = 0; k < n; ++k) {
      cols[(int)k] = d_lcol[d_rp[(size_t)row] + k];
      vals[(int)k] = d_val[d_rp[(size_t)row] + k];
    }
  }
  void getLocalDiagCopy(vector_type& d) const {
    auto v = d.getDataNonConst();
    for (size_t i = 0; i + 1 < d_rp.size(); ++i) {
      v[i] = S(0);
      const GO gi = d_row_map->getGlobalElement((LO)i);
      for (size_t k = d_rp[i]; k < d_rp[i + 1]; ++k)
        if (d_col_map->getGlobalElement(d_lcol[k]) == gi) v[i] = d_val[k];
    }
  }
  void leftScale(const vector_type& x) {
    auto xv = x.getData();
    for (size_t i = 0; i + 1 < d_rp.size(); ++i)
      for (size_t k = d_rp[i]; k < d_rp[i + 1]; ++k) d_val[k] *= xv[i];
  }
  void rightScale(const vector_type& x) {
    auto xv = x.getData();
    for (size_t k = 0; k < d_lcol.size(); ++k)
      d_val[k] *= xv[(size_t)x.getMap()->getLocalElement(d_col_map->getGlobalElement(d_lcol[k]))];
  }
  bool hasTransposeApply() const { return true; }
  void apply(const vector_type& x, vector_type& y, Teuchos::ETransp mode = Teuchos::NO_TRANS) const {
    auto xv = x.getData();
    auto yv = y.getDataNonConst();
    auto col = [&](size_t k) { return (size_t)x.getMap()->getLocalElement(d_col_map->getGlobalElement(d_lcol[k])); };
    if (mode == Teuchos::NO_TRANS) {
      for (size_t i = 0; i + 1 < d_rp.size(); ++i) {
        S s = 0;
        for (size_t k = d_rp[i]; k < d_rp[i + 1]; ++k) s += d_val[k] * xv[col(k)];
        yv[i] = s;
      }
    } else {
      for (size_t j = 0; j < y.getLocalLength(); ++j) yv[j] = S(0);
      for (size_t i = 0; i + 1 < d_rp.size(); ++i)
        for (size_t k = d_rp[i]; k < d_rp[i + 1]; ++k) yv[col(k)] += d_val[k] * xv[i];
    }
  }
  local_matrix_type getLocalMatrix() const {
    local_matrix_type m;
    m.graph.row_map.p = d_rp.data();
    m.graph.entries.p = d_lcol.data();
    m.values.p = d_val.data();
    return m;
  }

 private:
  Teuchos::RCP<const map_type> d_row_map, d_col_map, d_domain, d_range;
  std::vector<std::vector<std::pair<GO, S>>> d_rows;
  std::vector<size_t> d_rp;
  std::vector<LO> d_lcol;
  std::vector<S> d_val;
  bool d_filled = false;
};

}  // namespace Tpetra
