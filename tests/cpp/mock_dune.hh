// mock_dune.hh -- the few Dune/eWoms members the ewg adapters use, for compiling and
// driving them without Dune (test infrastructure).
#pragma once

#include <cstddef>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace Dune {
template <class T, int n>
struct FieldVector {
    T v[n] = {};
    T &operator[](int i) { return v[i]; }
    const T &operator[](int i) const { return v[i]; }
};
template <class T, int r, int c>
struct FieldMatrix {
    FieldVector<T, c> row[r];
    FieldVector<T, c> &operator[](int i) { return row[i]; }
    const FieldVector<T, c> &operator[](int i) const { return row[i]; }
};
template <class B>
struct BlockVector {
    std::vector<B> data;
    void resize(size_t n) { data.resize(n); }
    size_t size() const { return data.size(); }
    B &operator[](size_t i) { return data[i]; }
    const B &operator[](size_t i) const { return data[i]; }
};
}  // namespace Dune

namespace Ewoms {
struct NumericalIssue : std::runtime_error {
    explicit NumericalIssue(const std::string &m) : std::runtime_error(m) {}
};

// the IstlSparseMatrixAdapter members the adapters use (istlsparsematrixadapter.hh:72-164)
struct MockSparseMatrixAdapter {
    using MatrixBlock = Dune::FieldMatrix<double, 2, 2>;
    std::vector<std::map<unsigned, MatrixBlock>> rows;
    void reserve(const std::vector<std::set<unsigned>> &pattern) {
        rows.assign(pattern.size(), {});
        for (size_t i = 0; i < pattern.size(); ++i)
            for (unsigned c : pattern[i]) rows[i][c] = MatrixBlock{};
    }
    void setBlock(unsigned r, unsigned c, const MatrixBlock &b) { rows.at(r).at(c) = b; }
    const MatrixBlock &block(unsigned r, unsigned c) const { return rows.at(r).at(c); }
};
}  // namespace Ewoms
