// The LP handed to the engine -- the reference's pdlp::CompiledLp
// (include/pdlp/compiled_lp.h:17-63 there):
//   minimize objective^T x + objectiveOffset
//   subject to rowLower <= A x <= rowUpper, variableLower <= x <= variableUpper
// Infinite bounds are IEEE infinities.
#pragma once

#include "pdlp/sparse_matrix.h"

#include <cmath>
#include <cstddef>
#include <stdexcept>
#include <vector>

namespace pdlp {

struct CompiledLp {
    SparseMatrix matrix;
    std::vector<double> objective;
    double objectiveOffset = 0.0;
    std::vector<double> rowLower;
    std::vector<double> rowUpper;
    std::vector<double> variableLower;
    std::vector<double> variableUpper;

    [[nodiscard]] int numRows() const noexcept { return matrix.rows(); }
    [[nodiscard]] int numColumns() const noexcept { return matrix.columns(); }

    // Throws std::invalid_argument with the reference's messages.
    void validate() const {
        const auto n = static_cast<std::size_t>(numColumns());
        const auto m = static_cast<std::size_t>(numRows());
        if (!matrix.validate()) throw std::invalid_argument("CompiledLp contains an invalid sparse matrix");
        if (objective.size() != n || variableLower.size() != n || variableUpper.size() != n)
            throw std::invalid_argument("CompiledLp column vector dimensions do not match A");
        if (rowLower.size() != m || rowUpper.size() != m)
            throw std::invalid_argument("CompiledLp row vector dimensions do not match A");
        if (!std::isfinite(objectiveOffset)) throw std::invalid_argument("Objective offset must be finite");
        for (std::size_t j = 0; j < n; ++j) {
            const bool bad = !std::isfinite(objective[j]) || std::isnan(variableLower[j]) ||
                             std::isnan(variableUpper[j]) || variableLower[j] > variableUpper[j];
            if (bad) throw std::invalid_argument("Invalid objective coefficient or variable bounds");
        }
        for (std::size_t i = 0; i < m; ++i) {
            const bool bad = std::isnan(rowLower[i]) || std::isnan(rowUpper[i]) || rowLower[i] > rowUpper[i];
            if (bad) throw std::invalid_argument("Invalid row bounds");
        }
    }
};

}  // namespace pdlp
