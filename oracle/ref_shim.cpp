// CPU ORACLE -- test infrastructure only.
//
// C ABI shim over the reference PDLP engine compiled from its own sources
// (oracle/Makefile builds it into oracle/_ref/libref_pdlp.so).  This file is
// ours; the reference sources stay where they lie and are only compiled.
// It lets the tests pin the restatement in pdlp_oracle.cpp against the real
// thing and lets bench.py time the real reference as the CPU baseline.

#include "oracle.h"

#include "pdlp/pdlp_solver.h"
#include "pdlp/preconditioner.h"
#include "pdlp/scaling.h"
#include "pdlp/termination.h"

#include <cstring>
#include <cstdio>
#include <vector>

namespace {

pdlp::CompiledLp toCompiled(const pdlpx_problem& p) {
    pdlp::CompiledLp lp;
    std::vector<pdlp::MatrixTriplet> t;
    t.reserve(static_cast<size_t>(p.nnz));
    for (int r = 0; r < p.rows; ++r)
        for (int64_t k = p.csr_row_start[r]; k < p.csr_row_start[r + 1]; ++k)
            t.push_back({r, p.csr_col[k], p.csr_val[k]});
    lp.matrix = pdlp::SparseMatrix::fromTriplets(p.rows, p.cols, std::move(t));
    lp.objective.assign(p.objective, p.objective + p.cols);
    lp.objectiveOffset = p.objective_offset;
    lp.rowLower.assign(p.row_lower, p.row_lower + p.rows);
    lp.rowUpper.assign(p.row_upper, p.row_upper + p.rows);
    lp.variableLower.assign(p.var_lower, p.var_lower + p.cols);
    lp.variableUpper.assign(p.var_upper, p.var_upper + p.cols);
    return lp;
}

pdlp::PdlpOptions toOptions(const pdlpx_options& o) {
    pdlp::PdlpOptions r;
    r.iterationLimit = o.iteration_limit;
    r.timeLimitSeconds = o.time_limit_seconds;
    r.primalTolerance = o.primal_tolerance;
    r.dualTolerance = o.dual_tolerance;
    r.gapTolerance = o.gap_tolerance;
    r.terminationCheckFrequency = o.termination_check_frequency;
    r.powerIterationCount = o.power_iteration_count;
    r.threadCount = o.thread_count;
    r.parallelNonzeroThreshold = o.parallel_nonzero_threshold;
    r.initialStepSafety = o.initial_step_safety;
    r.minimumStepScale = o.minimum_step_scale;
    r.maximumPrimalWeight = o.maximum_primal_weight;
    r.minimumPrimalWeight = o.minimum_primal_weight;
    r.useRuizScaling = o.use_ruiz_scaling != 0;
    r.ruizIterations = o.ruiz_iterations;
    r.detectInfeasibility = o.detect_infeasibility != 0;
    r.infeasibilityTolerance = o.infeasibility_tolerance;
    r.useDiagonalPreconditioning = o.use_diagonal_preconditioning != 0;
    r.useAdaptiveLinesearch = o.use_adaptive_linesearch != 0;
    r.maximumStepTrials = o.maximum_step_trials;
    r.useAveraging = o.use_averaging != 0;
    r.useRestarts = o.use_restarts != 0;
    r.useFeasibilityPolishing = o.use_feasibility_polishing != 0;
    r.minimumRestartIterations = o.minimum_restart_iterations;
    r.maximumRestartIterations = o.maximum_restart_iterations;
    r.restartSufficientDecay = o.restart_sufficient_decay;
    r.restartNecessaryDecay = o.restart_necessary_decay;
    r.restartArtificialFraction = o.restart_artificial_fraction;
    r.initialPrimalWeight = o.initial_primal_weight;
    r.primalWeightSmoothing = o.primal_weight_smoothing;
    r.polishingIterations = o.polishing_iterations;
    return r;
}

}  // namespace

extern "C" {

int ref_solve(const pdlpx_problem* problem, const pdlpx_options* options, pdlpx_summary* s,
              double* x_out, double* y_out, double* ray_out) {
    const pdlp::CompiledLp lp = toCompiled(*problem);
    const pdlp::PdlpResult r = pdlp::PdlpSolver{}.solve(lp, toOptions(*options));
    std::memset(s, 0, sizeof(*s));
    s->status = static_cast<int32_t>(r.status);
    std::snprintf(s->message, sizeof(s->message), "%s", r.statusMessage.c_str());
    s->primal_objective = r.primalObjective;
    s->dual_objective = r.dualObjective;
    s->primal_residual = r.primalResidual;
    s->dual_residual = r.dualResidual;
    s->relative_gap = r.relativeGap;
    s->iterations = r.iterations;
    s->step_trials = r.stepTrials;
    s->restart_count = r.restartCount;
    s->solve_time_seconds = r.solveTimeSeconds;
    s->final_step_size = r.finalStepSize;
    s->static_step_bound = r.staticStepBound;
    s->final_primal_weight = r.finalPrimalWeight;
    s->has_dual_ray = r.dualRay.empty() ? 0 : 1;
    s->has_primal_ray = r.primalRay.empty() ? 0 : 1;
    if (x_out && r.primal.size() == size_t(problem->cols))
        std::memcpy(x_out, r.primal.data(), r.primal.size() * sizeof(double));
    if (y_out && r.rowDual.size() == size_t(problem->rows))
        std::memcpy(y_out, r.rowDual.data(), r.rowDual.size() * sizeof(double));
    if (ray_out && !r.dualRay.empty())
        std::memcpy(ray_out, r.dualRay.data(), r.dualRay.size() * sizeof(double));
    if (ray_out && !r.primalRay.empty())
        std::memcpy(ray_out, r.primalRay.data(), r.primalRay.size() * sizeof(double));
    return 0;
}

void* ref_matrix_from_triplets(int32_t rows, int32_t cols, int64_t count, const int32_t* tr,
                               const int32_t* tc, const double* tv) {
    std::vector<pdlp::MatrixTriplet> t(static_cast<size_t>(count));
    for (int64_t k = 0; k < count; ++k) t[k] = {tr[k], tc[k], tv[k]};
    return new pdlp::SparseMatrix(pdlp::SparseMatrix::fromTriplets(rows, cols, std::move(t)));
}

int64_t ref_matrix_nnz(const void* h) {
    return static_cast<int64_t>(static_cast<const pdlp::SparseMatrix*>(h)->nonzeros());
}

void ref_matrix_export(const void* h, int64_t* rp, int32_t* ci, double* rv, int64_t* cp,
                       int32_t* ri, double* cv) {
    const auto& m = *static_cast<const pdlp::SparseMatrix*>(h);
    std::memcpy(rp, m.csrRowStart().data(), m.csrRowStart().size() * 8);
    std::memcpy(ci, m.csrColumnIndex().data(), m.csrColumnIndex().size() * 4);
    std::memcpy(rv, m.csrValues().data(), m.csrValues().size() * 8);
    std::memcpy(cp, m.cscColumnStart().data(), m.cscColumnStart().size() * 8);
    std::memcpy(ri, m.cscRowIndex().data(), m.cscRowIndex().size() * 4);
    std::memcpy(cv, m.cscValues().data(), m.cscValues().size() * 8);
}

void ref_matrix_free(void* h) { delete static_cast<pdlp::SparseMatrix*>(h); }

int ref_spmv(const pdlpx_problem* problem, int32_t transpose, const double* x, double* y) {
    const pdlp::CompiledLp lp = toCompiled(*problem);
    std::vector<double> in, out;
    if (transpose) {
        in.assign(x, x + problem->rows);
        lp.matrix.transposeMultiply(in, out);
    } else {
        in.assign(x, x + problem->cols);
        lp.matrix.multiply(in, out);
    }
    std::memcpy(y, out.data(), out.size() * sizeof(double));
    return 0;
}

int ref_scaling(const pdlpx_problem* problem, int32_t ruiz_iterations, int32_t use_pc,
                double* row_scale, double* col_scale, double* primal_scale, double* dual_scale) {
    const pdlp::CompiledLp lp = toCompiled(*problem);
    const pdlp::ProblemScaling sc = pdlp::RuizScaler::equilibrate(lp, ruiz_iterations);
    const pdlp::DiagonalPreconditioner d = pdlp::Preconditioner::compute(sc.problem, use_pc != 0);
    for (int i = 0; i < problem->rows; ++i) {
        if (row_scale) row_scale[i] = sc.rowScale[i];
        if (dual_scale) dual_scale[i] = d.dualScale[i];
    }
    for (int j = 0; j < problem->cols; ++j) {
        if (col_scale) col_scale[j] = sc.columnScale[j];
        if (primal_scale) primal_scale[j] = d.primalScale[j];
    }
    return 0;
}

int ref_evaluate(const pdlpx_problem* problem, const pdlpx_options* options, const double* x,
                 const double* y, double* metrics) {
    const pdlp::CompiledLp lp = toCompiled(*problem);
    pdlp::TerminationChecker checker(toOptions(*options), lp);
    const std::vector<double> xv(x, x + problem->cols), yv(y, y + problem->rows);
    const pdlp::CandidateMetrics m = checker.evaluate(xv, yv);
    const double out[8] = {m.primalObjective, m.dualObjective, m.primalResidual, m.dualResidual,
                           m.relativeGap, m.kktScore, m.dualObjectiveValid ? 1.0 : 0.0,
                           m.finite ? 1.0 : 0.0};
    std::memcpy(metrics, out, sizeof(out));
    return 0;
}

}  // extern "C"
