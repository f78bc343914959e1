// Solve outcome -- the reference's pdlp::PdlpResult / PdlpStatus
// (include/pdlp/pdlp_result.h:10-64 there).
#pragma once

#include <cstdint>
#include <limits>
#include <string>
#include <vector>

namespace pdlp {

enum class PdlpStatus {
    Optimal,
    Infeasible,        // Farkas ray in dualRay
    Unbounded,         // improving ray in primalRay
    IterationLimit,
    TimeLimit,
    NumericalFailure,
    InvalidProblem
};

struct PdlpResult {
    PdlpStatus status = PdlpStatus::InvalidProblem;
    std::string statusMessage;

    std::vector<double> primal;
    std::vector<double> rowDual;
    std::vector<double> dualRay;
    std::vector<double> primalRay;

    double primalObjective = std::numeric_limits<double>::quiet_NaN();
    double dualObjective = -std::numeric_limits<double>::infinity();
    double primalResidual = std::numeric_limits<double>::infinity();
    double dualResidual = std::numeric_limits<double>::infinity();
    double relativeGap = std::numeric_limits<double>::infinity();

    std::int64_t iterations = 0;
    std::int64_t stepTrials = 0;   // includes rejected linesearch trials
    int restartCount = 0;
    double solveTimeSeconds = 0.0;

    double finalStepSize = 0.0;
    double staticStepBound = 0.0;
    double finalPrimalWeight = 1.0;
};

[[nodiscard]] const char* toString(PdlpStatus status) noexcept;

}  // namespace pdlp
