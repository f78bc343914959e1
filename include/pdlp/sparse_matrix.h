// Sparse constraint matrix in both orientations -- the reference's
// pdlp::SparseMatrix interface (include/pdlp/sparse_matrix.h:27-135 there).
// Storage: CSR (int64 offsets, int32 columns, double values) plus a CSC copy
// whose columns are sorted by row.  The engine copies both orientations to
// the GPU; the products below are the host-side utility callers use to build
// instances (the solver's own products run on the device).
#pragma once

#include "pdlp/parallel.h"

#include <cstddef>
#include <cstdint>
#include <vector>

namespace pdlp {

class SparseMatrix;

// Row / column chunks of roughly equal nonzero count for the host products.
struct SpmvPlan {
    std::vector<int> rowChunk;
    std::vector<int> columnChunk;
    int parts = 1;

    [[nodiscard]] int rowChunkCount() const noexcept { return static_cast<int>(rowChunk.size()) - 1; }
    [[nodiscard]] int columnChunkCount() const noexcept { return static_cast<int>(columnChunk.size()) - 1; }

    [[nodiscard]] static SpmvPlan build(const SparseMatrix& matrix, int parts, int chunksPerPart = 8);
};

struct MatrixTriplet {
    int row = 0;
    int column = 0;
    double value = 0.0;
};

class SparseMatrix {
public:
    using Offset = std::int64_t;

    SparseMatrix() = default;

    // Sorts by row then column, sums duplicates, drops exact zeros; throws
    // std::invalid_argument on out-of-range indices or non-finite values.
    static SparseMatrix fromTriplets(int rows, int columns, std::vector<MatrixTriplet> triplets);

    [[nodiscard]] int rows() const noexcept { return rows_; }
    [[nodiscard]] int columns() const noexcept { return columns_; }
    [[nodiscard]] std::size_t nonzeros() const noexcept { return csrValues_.size(); }

    [[nodiscard]] bool validate() const noexcept;

    // diag(rowScale) * A * diag(columnScale), same pattern.
    [[nodiscard]] SparseMatrix scaled(const std::vector<double>& rowScale,
                                      const std::vector<double>& columnScale) const;

    // result <- A x / A^T y (resized, fully overwritten).
    void multiply(const std::vector<double>& x, std::vector<double>& result, Executor* executor = nullptr,
                  const SpmvPlan* plan = nullptr) const;
    void transposeMultiply(const std::vector<double>& y, std::vector<double>& result,
                           Executor* executor = nullptr, const SpmvPlan* plan = nullptr) const;

    void multiplyRange(const double* x, double* result, int rowBegin, int rowEnd) const noexcept;
    void transposeMultiplyRange(const double* y, double* result, int columnBegin, int columnEnd) const noexcept;

    [[nodiscard]] const std::vector<Offset>& csrRowStart() const noexcept { return csrRowStart_; }
    [[nodiscard]] const std::vector<int>& csrColumnIndex() const noexcept { return csrColumnIndex_; }
    [[nodiscard]] const std::vector<double>& csrValues() const noexcept { return csrValues_; }
    [[nodiscard]] const std::vector<Offset>& cscColumnStart() const noexcept { return cscColumnStart_; }
    [[nodiscard]] const std::vector<int>& cscRowIndex() const noexcept { return cscRowIndex_; }
    [[nodiscard]] const std::vector<double>& cscValues() const noexcept { return cscValues_; }

private:
    int rows_ = 0;
    int columns_ = 0;
    std::vector<Offset> csrRowStart_;
    std::vector<int> csrColumnIndex_;
    std::vector<double> csrValues_;
    std::vector<Offset> cscColumnStart_;
    std::vector<int> cscRowIndex_;
    std::vector<double> cscValues_;
};

}  // namespace pdlp
