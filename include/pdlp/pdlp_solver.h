// Engine entry point -- the reference's pdlp::PdlpSolver
// (include/pdlp/pdlp_solver.h:9-15 there), running on an MI355X.
#pragma once

#include "pdlp/compiled_lp.h"
#include "pdlp/pdlp_options.h"
#include "pdlp/pdlp_result.h"

#include <vector>

namespace pdlp {

// Optional start point in the caller's coordinates (empty = cold start).
struct WarmStart {
    std::vector<double> primal;
    std::vector<double> rowDual;
};

class PdlpSolver {
public:
    // Never throws for a bad problem or a device error: they come back as
    // InvalidProblem / NumericalFailure with the reason in statusMessage.
    [[nodiscard]] PdlpResult solve(const CompiledLp& problem, const PdlpOptions& options = {}) const;

    // Same, started from `warmStart` (projected into the box and the dual cone).
    [[nodiscard]] PdlpResult solve(const CompiledLp& problem, const PdlpOptions& options,
                                   const WarmStart& warmStart) const;
};

}  // namespace pdlp
