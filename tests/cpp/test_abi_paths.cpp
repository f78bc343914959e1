// The C ABI paths of include/pdlpx.h that the drop-in suite does not reach,
// driven from C++: a batch with an empty and an invalid member in mixed
// precision (with compaction as instances finish), the autotuner,
// pdlpx_update_vectors with warm starts, row slicing of a batch, the
// profiling hooks, SpMV and scaling of a power-law LP, a Farkas ray, and the
// warm-start network's ABI (include/pdlpx_mlp.h).
// Checks return codes and statuses; numerical parity is the Python tests'
// job.  Built by optimisation_solver_amd/build.py (and, with the host code
// under AddressSanitizer, by tests/cpp/asan.sh); run from the repository root.
#include "pdlpx.h"
#include "pdlpx_mlp.h"

#include <cmath>
#include <cstdio>
#include <vector>

// Seeded instance generators (optimisation_solver_amd/csrc/lpgen.cpp).
extern "C" {
void* lpgen_box(int32_t m, int32_t n, int32_t per_col, uint64_t seed);
void* lpgen_zipf(int32_t m, int32_t n, int32_t literal, uint64_t seed);
void lpgen_sizes(const void* h, int32_t* m, int32_t* n, int64_t* nnz);
void lpgen_export(const void* h, int64_t* rp, int32_t* ci, double* rv, int64_t* cp, int32_t* ri, double* cv,
                  double* c, double* lo, double* hi, double* rlo, double* rhi, double* x0);
void lpgen_free(void* h);
}

struct Lp {
    int32_t m = 0, n = 0;
    int64_t z = 0;
    std::vector<int64_t> rp, cp;
    std::vector<int32_t> ci, ri;
    std::vector<double> rv, cv, c, lo, hi, rlo, rhi;
    pdlpx_problem prob() const {
        pdlpx_problem p{};
        p.rows = m;
        p.cols = n;
        p.nnz = z;
        p.csr_row_start = rp.data();
        p.csr_col = ci.data();
        p.csr_val = rv.data();
        p.csc_col_start = cp.data();
        p.csc_row = ri.data();
        p.csc_val = cv.data();
        p.objective = c.data();
        p.row_lower = rlo.data();
        p.row_upper = rhi.data();
        p.var_lower = lo.data();
        p.var_upper = hi.data();
        return p;
    }
};

static Lp take(void* h) {
    Lp l;
    lpgen_sizes(h, &l.m, &l.n, &l.z);
    l.rp.resize(l.m + 1);
    l.cp.resize(l.n + 1);
    l.ci.resize(l.z);
    l.ri.resize(l.z);
    l.rv.resize(l.z);
    l.cv.resize(l.z);
    l.c.resize(l.n);
    l.lo.resize(l.n);
    l.hi.resize(l.n);
    l.rlo.resize(l.m);
    l.rhi.resize(l.m);
    lpgen_export(h, l.rp.data(), l.ci.data(), l.rv.data(), l.cp.data(), l.ri.data(), l.cv.data(), l.c.data(),
                 l.lo.data(), l.hi.data(), l.rlo.data(), l.rhi.data(), nullptr);
    lpgen_free(h);
    return l;
}

static int fails = 0;
#define EXPECT(cond, what)                                  \
    do {                                                    \
        if (!(cond)) {                                      \
            std::printf("  FAIL %s\n", what);               \
            ++fails;                                        \
        } else {                                            \
            std::printf("  ok   %s\n", what);               \
        }                                                   \
        std::fflush(stdout);                                \
    } while (0)

int main() {
    pdlpx_options o;
    pdlpx_default_options(&o);
    o.primal_tolerance = o.dual_tolerance = o.gap_tolerance = 1e-6;

    // ---- batch: 6 box LPs, an empty LP and an invalid one, mixed precision
    std::vector<Lp> lps;
    for (int k = 0; k < 6; ++k) lps.push_back(take(lpgen_box(300 + 170 * k, 750 + 425 * k, 5, 77 + k)));
    Lp empty;
    empty.rp = {0};
    empty.cp = {0};
    lps.push_back(empty);
    Lp bad = lps[0];
    bad.rlo[3] = bad.rhi[3] + 1.0;   // crossed row bounds
    lps.push_back(bad);
    const int B = static_cast<int>(lps.size());
    std::vector<pdlpx_problem> probs;
    for (auto& l : lps) probs.push_back(l.prob());
    pdlpx_ctx* X = nullptr;
    EXPECT(pdlpx_create(&X, 0) == 0, "create");
    EXPECT(pdlpx_load_batch(X, B, probs.data(), PDLPX_MIXED) == 0, "load_batch mixed");
    pdlpx_tiling t{};
    EXPECT(pdlpx_autotune(X, PDLPX_FP64, 5, &t) == 0, "autotune");
    std::vector<pdlpx_summary> s(B);
    std::vector<std::vector<double>> xs(B), ys(B);
    std::vector<double*> xp(B), yp(B);
    for (int k = 0; k < B; ++k) {
        xs[k].assign(std::max(1, lps[k].n), 0.0);
        ys[k].assign(std::max(1, lps[k].m), 0.0);
        xp[k] = xs[k].data();
        yp[k] = ys[k].data();
    }
    o.precision = PDLPX_MIXED;
    EXPECT(pdlpx_solve_loaded(X, &o, nullptr, nullptr, s.data(), xp.data(), yp.data()) == 0, "solve_loaded");
    int opt = 0;
    for (int k = 0; k < 6; ++k) opt += s[k].status == PDLPX_OPTIMAL;
    EXPECT(opt == 6, "six box LPs optimal");
    EXPECT(s[7].status == PDLPX_INVALID_PROBLEM, "invalid member rejected");
    double a = 0, b = 0;
    EXPECT(pdlpx_time_trial(X, PDLPX_MIXED, 3, &a, &b) == 0 && a > 0 && b > 0, "time_trial mixed");
    EXPECT(pdlpx_time_gathers(X, PDLPX_MIXED, 3, &a, &b) == 0 && a > 0 && b > 0, "time_gathers mixed");
    // new costs for one member, warm starts for all, FP64 again
    std::vector<double> c2 = lps[2].c;
    for (auto& v : c2) v *= 1.01;
    EXPECT(pdlpx_update_vectors(X, 2, c2.data(), nullptr, nullptr) == 0, "update_vectors");
    std::vector<const double*> wx(B), wy(B);
    for (int k = 0; k < B; ++k) {
        wx[k] = k % 2 ? xs[k].data() : nullptr;
        wy[k] = k % 2 ? ys[k].data() : nullptr;
    }
    o.precision = PDLPX_FP64;
    EXPECT(pdlpx_solve_loaded(X, &o, wx.data(), wy.data(), s.data(), xp.data(), yp.data()) == 0,
           "solve_loaded warm");
    opt = 0;
    for (int k = 0; k < 6; ++k) opt += s[k].status == PDLPX_OPTIMAL;
    EXPECT(opt == 6, "six box LPs optimal again");
    // slicing on the batch
    EXPECT(pdlpx_set_slicing(X, 200) == 0, "set_slicing 200");
    int32_t rps = 0, rows = 0;
    EXPECT(pdlpx_get_slicing(X, &rps, &rows) == 0 && rps == 200 && rows > 0, "get_slicing");
    EXPECT(pdlpx_solve_loaded(X, &o, nullptr, nullptr, s.data(), xp.data(), yp.data()) == 0, "solve sliced");
    opt = 0;
    for (int k = 0; k < 6; ++k) opt += s[k].status == PDLPX_OPTIMAL;
    EXPECT(opt == 6, "sliced batch optimal");
    EXPECT(pdlpx_time_trial(X, PDLPX_FP64, 3, &a, &b) == 0, "time_trial sliced");
    EXPECT(pdlpx_time_gathers(X, PDLPX_FP64, 3, &a, &b) == 0, "time_gathers sliced");
    EXPECT(pdlpx_set_slicing(X, 0) == 0, "slicing off");
    double rate = 0;
    EXPECT(pdlpx_gather_rate(X, 1 << 20, 8, 1 << 20, 3, &rate) == 0 && rate > 0, "gather_rate");
    pdlpx_destroy(X);

    // ---- a single power-law LP: load, spmv, scaling, infeasible variant with a ray
    Lp z = take(lpgen_zipf(3000, 7000, 0, 5));
    pdlpx_problem pz = z.prob();
    EXPECT(pdlpx_create(&X, 0) == 0, "create 2");
    EXPECT(pdlpx_load(X, &pz, PDLPX_FP64) == 0, "load zipf");
    std::vector<double> xv(z.n, 1.0), yv(z.m), zv(z.n);
    EXPECT(pdlpx_spmv(X, 0, xv.data(), yv.data()) == 0, "spmv");
    EXPECT(pdlpx_spmv(X, 1, yv.data(), zv.data()) == 0, "spmv transpose");
    std::vector<double> rs(z.m), cs(z.n), ps(z.n), ds(z.m);
    EXPECT(pdlpx_scaling(X, 10, 1, rs.data(), cs.data(), ps.data(), ds.data()) == 0, "scaling");
    pdlpx_summary one{};
    std::vector<double> x1(z.n), y1(z.m), ray(std::max(z.m, z.n));
    EXPECT(pdlpx_solve(X, &o, nullptr, nullptr, &one, x1.data(), y1.data(), ray.data()) == 0 &&
               one.status == PDLPX_OPTIMAL,
           "zipf optimal");
    pdlpx_destroy(X);
    // infeasible: two rows that contradict each other on one column
    Lp inf;
    inf.m = 2;
    inf.n = 1;
    inf.z = 2;
    inf.rp = {0, 1, 2};
    inf.ci = {0, 0};
    inf.rv = {1.0, 1.0};
    inf.cp = {0, 2};
    inf.ri = {0, 1};
    inf.cv = {1.0, 1.0};
    inf.c = {1.0};
    inf.lo = {-INFINITY};
    inf.hi = {INFINITY};
    inf.rlo = {1.0, -INFINITY};
    inf.rhi = {INFINITY, -1.0};
    pdlpx_problem pi = inf.prob();
    EXPECT(pdlpx_create(&X, 0) == 0, "create 3");
    EXPECT(pdlpx_load(X, &pi, PDLPX_FP64) == 0, "load infeasible");
    std::vector<double> r2(2), x2(1), y2(2);
    EXPECT(pdlpx_solve(X, &o, nullptr, nullptr, &one, x2.data(), y2.data(), r2.data()) == 0 &&
               one.status == PDLPX_INFEASIBLE && one.has_dual_ray,
           "infeasible with ray");
    pdlpx_destroy(X);

    // ---- the warm-start network: 64 -> 128 -> 64
    pdlpx_mlp* N = nullptr;
    const int32_t bad_dims[3] = {64, 100, 64};
    EXPECT(pdlpx_mlp_create(&N, 0, 2, bad_dims, 128, 1) != 0 && N == nullptr, "network: 100-wide layer rejected");
    const int32_t dims[3] = {64, 128, 64};
    EXPECT(pdlpx_mlp_create(&N, 0, 2, dims, 128, 7) == 0 && N != nullptr, "network: create");
    std::vector<float> W0(128 * 64), b0(128), W0b(128 * 64), b0b(128);
    EXPECT(pdlpx_mlp_get_params(N, 0, W0.data(), b0.data()) == 0, "network: get_params");
    EXPECT(pdlpx_mlp_set_params(N, 0, W0.data(), b0.data()) == 0 &&
               pdlpx_mlp_get_params(N, 0, W0b.data(), b0b.data()) == 0 && W0 == W0b && b0 == b0b,
           "network: parameters round-trip");
    std::vector<float> x(256 * 64), tgt(256 * 64), y(5 * 64);
    for (size_t q = 0; q < x.size(); ++q) {
        x[q] = static_cast<float>(std::sin(0.37 * static_cast<double>(q)));
        tgt[q] = 0.5f * x[q];   // a learnable target: half the input
    }
    EXPECT(pdlpx_mlp_forward(N, 5, x.data(), y.data()) == 0 && std::isfinite(y[0]) && std::isfinite(y.back()),
           "network: forward of 5 rows (padded)");
    double loss = 0.0;
    std::vector<float> dW0(128 * 64), db0(128), dW1(64 * 128), db1(64);
    float* dW[2] = {dW0.data(), dW1.data()};
    float* db[2] = {db0.data(), db1.data()};
    EXPECT(pdlpx_mlp_gradients(N, 64, x.data(), tgt.data(), &loss, dW, db) == 0 && std::isfinite(loss) && loss > 0,
           "network: gradients of one batch");
    EXPECT(pdlpx_mlp_set_data(N, 256, x.data(), tgt.data()) == 0, "network: set_data");
    pdlpx_adam adam{};
    adam.learning_rate = 1e-3;
    adam.beta1 = 0.9;
    adam.beta2 = 0.999;
    adam.epsilon = 1e-8;
    adam.final_learning_rate = 1e-4;
    adam.weight_decay = 0.0;
    std::vector<double> eloss(20);
    double secs = 0.0, flops = 0.0;
    EXPECT(pdlpx_mlp_train(N, 20, 64, &adam, 3, eloss.data(), &secs, &flops) == 0 && eloss.back() < eloss[0] &&
               flops > 0,
           "network: 20 epochs of Adam lower the loss");
    std::vector<float> A(64 * 64, 0.0f), Bm(64 * 64), Cm(64 * 64);
    for (int i = 0; i < 64; ++i) A[i * 64 + i] = 1.0f;
    for (size_t q = 0; q < Bm.size(); ++q) Bm[q] = static_cast<float>(static_cast<int>(q % 17) - 8);
    double ms = 0.0;
    EXPECT(pdlpx_mlp_gemm(N, 64, 64, 64, A.data(), 0, Bm.data(), 0, Cm.data(), 2, &ms) == 0 && Cm == Bm && ms > 0,
           "network: GEMM with the identity is exact");
    pdlpx_mlp_destroy(N);
    std::printf("%d failure(s)\n", fails);
    return fails ? 1 : 0;
}
