// Ruiz equilibration -- the reference's pdlp::ProblemScaling / RuizScaler
// (include/pdlp/scaling.h:25-63 there).  equilibrate() runs the passes on the
// GPU; the scale vectors are bitwise those of the reference.
//   Ahat = diag(rowScale) A diag(columnScale), x = columnScale .* xhat,
//   y = rowScale .* yhat
#pragma once

#include "pdlp/compiled_lp.h"

#include <vector>

namespace pdlp {

struct ProblemScaling {
    std::vector<double> rowScale;
    std::vector<double> columnScale;
    CompiledLp problem;

    void toOriginal(const std::vector<double>& scaledPrimal, const std::vector<double>& scaledDual,
                    std::vector<double>& primal, std::vector<double>& dual) const;
    void toScaled(const std::vector<double>& primal, const std::vector<double>& dual,
                  std::vector<double>& scaledPrimal, std::vector<double>& scaledDual) const;
};

class RuizScaler {
public:
    [[nodiscard]] static ProblemScaling equilibrate(const CompiledLp& problem, int iterations);
    // Ratio of the largest to the smallest nonzero row infinity norm.
    [[nodiscard]] static double conditionSpread(const SparseMatrix& matrix);
};

}  // namespace pdlp
