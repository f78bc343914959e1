// CPU ORACLE -- test infrastructure only (see pdlp_oracle.cpp).
//
// C ABI of the serial restatement of the reference PDLP (oracle/liboracle.so)
// and, with the ref_ prefix, of the shim over the reference compiled from its
// own sources (oracle/_ref/libref_pdlp.so, built by oracle/Makefile when the
// reference checkout is present).  Both reuse the engine's plain-data structs
// from include/pdlpx.h so one set of bindings drives all three.
#ifndef PDLP_ORACLE_H
#define PDLP_ORACLE_H

#include "../include/pdlpx.h"

#define ORACLE_TRACE_WIDTH 8   /* k, trials, eta, omega, kkt_cur, kkt_avg, kkt_best, restarts */

#ifdef __cplusplus
extern "C" {
#endif

int oracle_solve(const pdlpx_problem* problem, const pdlpx_options* options,
                 const double* warm_x, const double* warm_y, pdlpx_summary* summary,
                 double* x_out, double* y_out, double* ray_out);
int oracle_solve_traced(const pdlpx_problem* problem, const pdlpx_options* options,
                        const double* warm_x, const double* warm_y, pdlpx_summary* summary,
                        double* x_out, double* y_out, double* ray_out, double* trace,
                        int32_t trace_cap, int32_t* trace_len);
void* oracle_matrix_from_triplets(int32_t rows, int32_t cols, int64_t count, const int32_t* tr,
                                  const int32_t* tc, const double* tv);
int64_t oracle_matrix_nnz(const void* handle);
void oracle_matrix_export(const void* handle, int64_t* rp, int32_t* ci, double* rv, int64_t* cp,
                          int32_t* ri, double* cv);
void oracle_matrix_free(void* handle);
int oracle_spmv(const pdlpx_problem* problem, int32_t transpose, const double* x, double* y);
int oracle_scaling(const pdlpx_problem* problem, int32_t ruiz_iterations, int32_t use_pc,
                   double* row_scale, double* col_scale, double* primal_scale, double* dual_scale);
int oracle_evaluate(const pdlpx_problem* problem, const pdlpx_options* options, const double* x,
                    const double* y, double* metrics);
int oracle_step_judge(const pdlpx_options* options, double max_safe, double eta, double omega,
                      double pm, double dm, double inter, int64_t k, double* eta_out);
double oracle_reweight(const pdlpx_options* options, double omega, double dx, double dy);
int oracle_restart_choose(const pdlpx_options* options, double cur_kkt, double avg_kkt,
                          double baseline, double prev_score, int64_t since_restart,
                          int64_t total, double* score_out, double* prev_out);

/* Shim over the compiled reference (same meaning as the oracle_ functions). */
int ref_solve(const pdlpx_problem* problem, const pdlpx_options* options, pdlpx_summary* summary,
              double* x_out, double* y_out, double* ray_out);
void* ref_matrix_from_triplets(int32_t rows, int32_t cols, int64_t count, const int32_t* tr,
                               const int32_t* tc, const double* tv);
int64_t ref_matrix_nnz(const void* handle);
void ref_matrix_export(const void* handle, int64_t* rp, int32_t* ci, double* rv, int64_t* cp,
                       int32_t* ri, double* cv);
void ref_matrix_free(void* handle);
int ref_spmv(const pdlpx_problem* problem, int32_t transpose, const double* x, double* y);
int ref_scaling(const pdlpx_problem* problem, int32_t ruiz_iterations, int32_t use_pc,
                double* row_scale, double* col_scale, double* primal_scale, double* dual_scale);
int ref_evaluate(const pdlpx_problem* problem, const pdlpx_options* options, const double* x,
                 const double* y, double* metrics);

#ifdef __cplusplus
}
#endif

#endif
