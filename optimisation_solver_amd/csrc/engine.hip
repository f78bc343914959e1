// MI355X PDLP engine: the C ABI of include/pdlpx.h over hand-written gfx950
// kernels (kernels.hpp).
//
// One solve keeps the LP, the scaled working copy and every iterate resident
// in HBM.  The PDHG iteration runs as "trial slots" of four kernels --
// primal half, dual half, linesearch decision, averaging blend -- whose
// accept/reject and buffer swap are decided on the device (DevState), so a
// batch of up to one termination interval of iterations runs with no host
// round trip; the batch is captured once into a hipGraph and replayed.  The
// host only wakes at termination checks (every terminationCheckFrequency
// iterations, pdlp_solver.cpp:350-476) to read a few reduced scalars, decide
// restarts and best candidates, and launch the next batch.

#include "kernels.hpp"

#include "../../include/pdlpx.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <random>
#include <string>
#include <vector>

using namespace pdlpx_dev;

namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();
constexpr int kStreamBlocks = 512;   // grid of the streaming reduction kernels

struct Failure {
    int code;
    std::string what;
};

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        const hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                               \
            throw Failure{e_ == hipErrorOutOfMemory ? PDLPX_E_NOMEM : PDLPX_E_HIP,         \
                          std::string("HIP: ") + #expr + ": " + hipGetErrorString(e_)};     \
    } while (0)

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        if (count == n && p) return;
        release();
        const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        HIPCHK(hipMalloc(&p, bytes));
        n = count;
    }
    void upload(const T* src, size_t count, hipStream_t s) {
        alloc(count);
        if (count) HIPCHK(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

int gridFor(int64_t n, int threads = kThreads) {
    return static_cast<int>(std::max<int64_t>(1, (n + threads - 1) / threads));
}
int streamGrid(int64_t n) { return std::min(kStreamBlocks, gridFor(n)); }

// Row blocks of one orientation: at most kTileNnz nonzeros and kTileRows rows
// each; a longer row gets a block of its own.
std::vector<int32_t> partitionRows(const std::vector<int32_t>& ptr, int nrows) {
    std::vector<int32_t> blk;
    blk.push_back(0);
    int r = 0;
    while (r < nrows) {
        const int start = r;
        const int32_t base = ptr[r];
        if (ptr[r + 1] - base > kTileNnz) {
            blk.push_back(++r);
            continue;
        }
        while (r < nrows && r - start < kTileRows && ptr[r + 1] - base <= kTileNnz) ++r;
        blk.push_back(r);
    }
    if (blk.size() == 1) blk.push_back(0);   // at least one (empty) block
    return blk;
}

struct OrientBuf {
    DBuf<int32_t> ptr, idx, blk;
    DBuf<double> val;   // original values
    int32_t nblk = 0, nrows = 0;
    template <typename T>
    Orient<T> view(const T* values) const {
        return Orient<T>{ptr.p, idx.p, values, blk.p, nblk, nrows};
    }
};

template <typename T>
struct Working {   // scaled problem data in precision T
    DBuf<T> vr, vc;   // CSR / CSC values
    DBuf<T> c, lo, hi, tau, tau_inv, rlo, rhi, sig, sig_inv;
};

template <typename T>
struct IterBuf {
    DBuf<T> x[2], y[2], ax[2], avx, avy, anx, any;
    void alloc(int n, int m) {
        for (int k = 0; k < 2; ++k) {
            x[k].alloc(n);
            y[k].alloc(m);
            ax[k].alloc(m);
        }
        avx.alloc(n);
        avy.alloc(m);
        anx.alloc(n);
        any.alloc(m);
    }
    Iterates<T> view() {
        return Iterates<T>{{x[0].p, x[1].p}, {y[0].p, y[1].p}, {ax[0].p, ax[1].p}, avx.p, avy.p};
    }
};

struct Metrics {   // termination.h:13-22
    double pobj = std::numeric_limits<double>::quiet_NaN();
    double dobj = -kInf;
    double pres = kInf, dres = kInf, gap = kInf, kkt = kInf;
    bool dualValid = false, finite = false;
};

void setStatusText(pdlpx_summary* s, const std::string& text) {
    std::snprintf(s->message, sizeof(s->message), "%s", text.c_str());
}

}  // namespace

struct pdlpx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string error;

    // ---- loaded problem (caller's coordinates) ----
    bool loaded = false;
    bool valid = false;
    std::string invalid;
    int m = 0, n = 0;
    int64_t nnz = 0;
    double offset = 0.0;
    double pnorm = 1.0, dnorm = 1.0;   // termination normalisers (termination.cpp:15-61)
    OrientBuf A, AT;                   // CSR of A and CSR of A^T (= CSC of A)
    DBuf<double> c, lo, hi, rlo, rhi;

    // ---- per-solve work (sized at load) ----
    DBuf<double> rs, cs, rmax, cmax;
    Working<double> w64;
    Working<float> w32;
    IterBuf<double> it64;
    IterBuf<float> it32;
    DBuf<double> xo0, yo0, xo1, yo1, bestx, besty, rx, ry, proj, u, warmx, warmy;
    DBuf<double> part_p, part_d, part_i, part_rows, part_cols, part_s, part_s2, part_s3;
    DBuf<double> red;   // reduction results
    DBuf<DevState> st;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    // captured trial-slot graphs, keyed on precision and slot count
    struct GraphEntry {
        hipGraphExec_t exec = nullptr;
        int prec = -1;
        int slots = 0;
    };
    std::vector<GraphEntry> graphs;

    void dropGraphs() {
        for (auto& g : graphs)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        graphs.clear();
    }
};

namespace {

// ---------------------------------------------------------------------------
// Validation (compiled_lp.h:32-62, sparse_matrix.cpp:203-241).
// ---------------------------------------------------------------------------
std::string validateProblem(const pdlpx_problem& p) {
    if (p.rows < 0 || p.cols < 0 || p.nnz < 0) return "CompiledLp contains an invalid sparse matrix";
    if (p.nnz > 0 && (!p.csr_col || !p.csr_val || !p.csc_row || !p.csc_val))
        return "CompiledLp contains an invalid sparse matrix";
    if (!p.csr_row_start || !p.csc_col_start) return "CompiledLp contains an invalid sparse matrix";
    if (p.csr_row_start[0] != 0 || p.csc_col_start[0] != 0 || p.csr_row_start[p.rows] != p.nnz ||
        p.csc_col_start[p.cols] != p.nnz)
        return "CompiledLp contains an invalid sparse matrix";
    for (int i = 0; i < p.rows; ++i)
        if (p.csr_row_start[i + 1] < p.csr_row_start[i]) return "CompiledLp contains an invalid sparse matrix";
    for (int j = 0; j < p.cols; ++j)
        if (p.csc_col_start[j + 1] < p.csc_col_start[j]) return "CompiledLp contains an invalid sparse matrix";
    for (int64_t k = 0; k < p.nnz; ++k) {
        if (p.csr_col[k] < 0 || p.csr_col[k] >= p.cols || !std::isfinite(p.csr_val[k]))
            return "CompiledLp contains an invalid sparse matrix";
        if (p.csc_row[k] < 0 || p.csc_row[k] >= p.rows || !std::isfinite(p.csc_val[k]))
            return "CompiledLp contains an invalid sparse matrix";
    }
    if ((p.cols > 0 && (!p.objective || !p.var_lower || !p.var_upper)))
        return "CompiledLp column vector dimensions do not match A";
    if ((p.rows > 0 && (!p.row_lower || !p.row_upper))) return "CompiledLp row vector dimensions do not match A";
    if (!std::isfinite(p.objective_offset)) return "Objective offset must be finite";
    for (int j = 0; j < p.cols; ++j)
        if (!std::isfinite(p.objective[j]) || std::isnan(p.var_lower[j]) || std::isnan(p.var_upper[j]) ||
            p.var_lower[j] > p.var_upper[j])
            return "Invalid objective coefficient or variable bounds";
    for (int i = 0; i < p.rows; ++i)
        if (std::isnan(p.row_lower[i]) || std::isnan(p.row_upper[i]) || p.row_lower[i] > p.row_upper[i])
            return "Invalid row bounds";
    if (p.nnz >= (int64_t(1) << 31)) return "Matrix has too many nonzeros for 32-bit device offsets";
    return "";
}

bool optionsValid(const pdlpx_options& o) {   // pdlp_solver.cpp:127-133
    return !(o.iteration_limit < 0 || o.termination_check_frequency <= 0 || o.primal_tolerance < 0.0 ||
             o.dual_tolerance < 0.0 || o.gap_tolerance < 0.0 || o.initial_step_safety <= 0.0 ||
             o.minimum_primal_weight <= 0.0 || o.maximum_primal_weight < o.minimum_primal_weight);
}

void uploadOrient(OrientBuf& ob, int nrows, const int64_t* start, const int32_t* idx, const double* val,
                  int64_t nnz, hipStream_t s) {
    std::vector<int32_t> ptr(static_cast<size_t>(nrows) + 1);
    for (int r = 0; r <= nrows; ++r) ptr[r] = static_cast<int32_t>(start[r]);
    const std::vector<int32_t> blk = partitionRows(ptr, nrows);
    ob.nrows = nrows;
    ob.nblk = static_cast<int32_t>(blk.size()) - 1;
    ob.ptr.upload(ptr.data(), ptr.size(), s);
    ob.blk.upload(blk.data(), blk.size(), s);
    ob.idx.upload(idx, static_cast<size_t>(nnz), s);
    ob.val.upload(val, static_cast<size_t>(nnz), s);
    HIPCHK(hipStreamSynchronize(s));   // host staging vectors die here
}

// ===========================================================================
// One solve.
// ===========================================================================
class Run {
public:
    Run(pdlpx_ctx& c, const pdlpx_options& o) : X(c), o(o), s(c.stream) {}

    void go(const double* warmX, const double* warmY, pdlpx_summary* sum, double* xOut, double* yOut,
            double* rayOut);

    // Runs only the set-up (scaling and step scales) and waits for it.
    void setupOnly() {
        setup();
        HIPCHK(hipStreamSynchronize(s));
    }

private:
    pdlpx_ctx& X;
    const pdlpx_options& o;
    hipStream_t s;
    using Clock = std::chrono::steady_clock;
    Clock::time_point t0 = Clock::now();
    double since() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

    bool scaled = false;
    int prec = PDLPX_FP64;   // precision of the current phase
    DevState hs{};           // host mirror of the device state
    double maxSafe = 0.99;
    int64_t fp32Its = 0;
    int checks = 0;
    double loopSeconds = 0.0, setupSeconds = 0.0;

    // ------------------------------------------------------------------
    void pullState() {
        HIPCHK(hipMemcpyAsync(&hs, X.st.p, sizeof(DevState), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    void pushState() { HIPCHK(hipMemcpyAsync(X.st.p, &hs, sizeof(DevState), hipMemcpyHostToDevice, s)); }

    // Reduces `k` partial arrays into X.red[0..k) and returns them.
    std::vector<double> reduce(std::initializer_list<std::pair<const double*, int>> arrays) {
        ReduceJob job{};
        int k = 0;
        for (const auto& a : arrays) {
            job.src[k] = a.first;
            job.count[k] = a.second;
            job.dst[k] = X.red.p + k;
            ++k;
        }
        k_reduce<<<k, kReduceThreads, 0, s>>>(job);
        HIPCHK(hipGetLastError());
        std::vector<double> out(static_cast<size_t>(k));
        HIPCHK(hipMemcpyAsync(out.data(), X.red.p, sizeof(double) * k, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        return out;
    }

    template <typename T>
    Working<T>& work();
    template <typename T>
    IterBuf<T>& iters();

    template <typename T>
    ColData<T> cols() {
        Working<T>& w = work<T>();
        return ColData<T>{w.c.p, w.lo.p, w.hi.p, w.tau.p, w.tau_inv.p};
    }
    template <typename T>
    RowData<T> rows() {
        Working<T>& w = work<T>();
        return RowData<T>{w.rlo.p, w.rhi.p, w.sig.p, w.sig_inv.p};
    }

    // ------------------------------------------------------------------
    void setup();
    double powerIteration();

    template <typename T>
    void launchSlot() {
        Working<T>& w = work<T>();
        Iterates<T> it = iters<T>().view();
        k_primal<T><<<X.AT.nblk, kThreads, 0, s>>>(X.AT.view<T>(w.vc.p), cols<T>(), it, X.st.p, X.part_p.p);
        k_dual<T><<<X.A.nblk, kThreads, 0, s>>>(X.A.view<T>(w.vr.p), rows<T>(), it, X.st.p, X.part_d.p,
                                                 X.part_i.p);
        k_decide<<<1, kReduceThreads, 0, s>>>(X.st.p, X.part_p.p, X.AT.nblk, X.part_d.p, X.part_i.p,
                                               X.A.nblk);
        k_average<T><<<streamGrid(X.n + X.m) * 4, kThreads, 0, s>>>(X.st.p, it, X.n, X.m);
    }

    template <typename T>
    hipGraphExec_t graphFor(int slots) {
        const int key = sizeof(T) == 8 ? PDLPX_FP64 : PDLPX_FP32;
        for (auto& g : X.graphs)
            if (g.prec == key && g.slots == slots) return g.exec;
        hipGraph_t graph = nullptr;
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < slots; ++k) launchSlot<T>();
        HIPCHK(hipStreamEndCapture(s, &graph));
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPCHK(e);
        X.graphs.push_back({exec, key, slots});
        return exec;
    }

    // Runs trials until hs.iteration reaches `target` or the device flags a
    // failure.  The state mirror is pushed first and pulled at the end.
    template <typename T>
    void runTrials(int64_t target) {
        hs.target = target;
        hs.done = hs.iteration >= target ? 1 : 0;
        pushState();
        if (hs.done) return;
        while (true) {
            const int64_t left = target - hs.iteration;
            const int big = o.termination_check_frequency + 4;
            int slots = left + 4 > 16 ? big : 16;
            if (slots > big) slots = big;
            HIPCHK(hipGraphLaunch(graphFor<T>(slots), s));
            pullState();
            if (hs.done) break;
        }
    }

    template <typename T>
    void refreshActivity() {   // A x for the current iterate (pdhg_kernel.cpp:40-42)
        IterBuf<T>& I = iters<T>();
        k_spmv<T><<<X.A.nblk, kThreads, 0, s>>>(X.A.view<T>(work<T>().vr.p), I.x[hs.parity].p,
                                                 I.ax[hs.parity].p);
        HIPCHK(hipGetLastError());
    }

    Metrics finish(double viol, double support, double rowDrift, double obj, double inf,
                   double colDrift) const {   // termination.cpp:296-324
        Metrics m;
        m.pres = std::sqrt(viol) / X.pnorm;
        m.dres = std::sqrt(colDrift + rowDrift) / X.dnorm;
        m.pobj = X.offset + obj;
        m.dobj = X.offset + inf - support;
        m.finite = std::isfinite(m.pobj) && std::isfinite(m.dobj) && std::isfinite(m.pres) &&
                   std::isfinite(m.dres);
        if (!m.finite) {
            m.kkt = kInf;
            return m;
        }
        m.dualValid = true;
        m.gap = std::abs(m.pobj - m.dobj) / (1.0 + std::abs(m.pobj) + std::abs(m.dobj));
        m.kkt = std::max(std::max(m.pres, m.dres), m.gap);
        return m;
    }

    bool optimal(const Metrics& m) const {   // termination.cpp:327-333
        return m.finite && m.dualValid && m.pres <= o.primal_tolerance && m.dres <= o.dual_tolerance &&
               m.gap <= o.gap_tolerance;
    }

    // Scores the candidates already in original coordinates in xo0/yo0 and
    // xo1/yo1 (one fused row pass and one fused column pass).
    void scorePair(Metrics& m0, Metrics& m1) {
        k_term_rows<<<X.A.nblk, kThreads, 0, s>>>(X.A.view<double>(X.A.val.p), X.xo0.p, X.xo1.p, X.yo0.p,
                                                  X.yo1.p, X.rlo.p, X.rhi.p, X.part_rows.p);
        k_term_cols<<<X.AT.nblk, kThreads, 0, s>>>(X.AT.view<double>(X.AT.val.p), X.yo0.p, X.yo1.p, X.xo0.p,
                                                   X.xo1.p, X.c.p, X.lo.p, X.hi.p, X.part_cols.p);
        HIPCHK(hipGetLastError());
        const int nr = X.A.nblk, nc = X.AT.nblk;
        const double* pr = X.part_rows.p;
        const double* pc = X.part_cols.p;
        const auto v = reduce({{pr, nr}, {pr + nr, nr}, {pr + 2 * nr, nr}, {pr + 3 * nr, nr},
                               {pr + 4 * nr, nr}, {pr + 5 * nr, nr}, {pc, nc}, {pc + nc, nc},
                               {pc + 2 * nc, nc}, {pc + 3 * nc, nc}, {pc + 4 * nc, nc},
                               {pc + 5 * nc, nc}});
        m0 = finish(v[0], v[1], v[2], v[6], v[7], v[8]);
        m1 = finish(v[3], v[4], v[5], v[9], v[10], v[11]);
        ++checks;
    }

    template <typename T>
    void toOriginal(const T* x, const T* y, double* xo, double* yo) {
        k_to_original<T><<<streamGrid(X.n + X.m) * 4, kThreads, 0, s>>>(X.n, X.m, X.cs.p, X.rs.p, x, y, xo,
                                                                         yo);
        HIPCHK(hipGetLastError());
    }

    void copyD(double* dst, const double* src, size_t count) {
        if (count) HIPCHK(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToDevice, s));
    }

    template <typename T>
    void copyT(T* dst, const T* src, size_t count) {
        if (count) HIPCHK(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToDevice, s));
    }

    template <typename T>
    void zeroT(T* p, int64_t count) {
        if (count) k_fill<T><<<streamGrid(count), kThreads, 0, s>>>(count, p, T(0));
        HIPCHK(hipGetLastError());
    }

    // Ray tests (infeasibility.cpp:51-192) on the iterate's displacement since
    // the last restart.  Returns 1 = Farkas ray (ry), 2 = improving ray (rx).
    template <typename T>
    int rayTests() {
        IterBuf<T>& I = iters<T>();
        const int p = hs.parity;
        const int g = streamGrid(X.n + X.m);
        k_ray_dir<T><<<g, kThreads, 0, s>>>(X.n, X.m, X.cs.p, X.rs.p, I.x[p].p, I.anx.p, I.y[p].p, I.any.p,
                                             X.rx.p, X.ry.p, X.part_s.p);
        HIPCHK(hipGetLastError());
        // norms^2 into red[0] (x) and red[1] (y) for the next kernels
        ReduceJob job{};
        job.src[0] = X.part_s.p;
        job.count[0] = g;
        job.dst[0] = X.red.p + 16;
        job.src[1] = X.part_s.p + g;
        job.count[1] = g;
        job.dst[1] = X.red.p + 17;
        k_reduce<<<2, kReduceThreads, 0, s>>>(job);
        const double* norms = X.red.p + 16;
        const int gm = streamGrid(X.m), gn = streamGrid(X.n);
        k_farkas_rows<<<gm, kThreads, 0, s>>>(X.m, X.ry.p, norms, X.rlo.p, X.rhi.p, X.u.p, X.part_s2.p);
        k_farkas_cols<<<X.AT.nblk, kThreads, 0, s>>>(X.AT.view<double>(X.AT.val.p), X.u.p, X.lo.p, X.hi.p,
                                                     X.part_cols.p);
        k_ray_cols<<<gn, kThreads, 0, s>>>(X.n, X.rx.p, norms, X.c.p, X.lo.p, X.hi.p, X.proj.p, X.part_s3.p);
        k_ray_rows<<<X.A.nblk, kThreads, 0, s>>>(X.A.view<double>(X.A.val.p), X.proj.p, X.rlo.p, X.rhi.p,
                                                 X.part_rows.p);
        HIPCHK(hipGetLastError());
        const int nc = X.AT.nblk, nr = X.A.nblk;
        const auto v = reduce({{X.part_s2.p, gm}, {X.part_s2.p + gm, gm}, {X.part_cols.p, nc},
                               {X.part_cols.p + nc, nc}, {X.part_s3.p, gn}, {X.part_s3.p + gn, gn},
                               {X.part_rows.p, nr}, {X.red.p + 16, 1}, {X.red.p + 17, 1}});
        const double nx = std::sqrt(v[7]), ny = std::sqrt(v[8]);
        // Dual ray: value = box infimum - support, violation from both sides.
        if (X.m > 0 && ny > 1e-300 && std::isfinite(ny)) {
            const double value = v[3] - v[1];
            if (std::isfinite(value) && value > 0.0) {
                const double viol = std::sqrt(v[2] + v[0]);
                if (viol / value <= o.infeasibility_tolerance) return 1;
            }
        }
        if (X.n > 0 && nx > 1e-300 && std::isfinite(nx)) {
            const double rate = v[5];
            if (std::isfinite(rate) && rate < 0.0) {
                const double viol = std::sqrt(v[4] + v[6]);
                if (viol / (-rate) <= o.infeasibility_tolerance) return 2;
            }
        }
        return 0;
    }

    // Squared displacement of (x, y) from the restart anchors.
    template <typename T>
    std::pair<double, double> displacement(const T* x, const T* y) {
        IterBuf<T>& I = iters<T>();
        const int g = streamGrid(X.n + X.m);
        k_sqdist<T><<<g, kThreads, 0, s>>>(X.n, x, I.anx.p, X.m, y, I.any.p, X.part_s.p);
        HIPCHK(hipGetLastError());
        const auto v = reduce({{X.part_s.p, g}, {X.part_s.p + g, g}});
        return {v[0], v[1]};
    }

    // Step controller's primal-weight update (step_controller.cpp:80-98).
    void reweight(double dx, double dy) {
        if (dx <= 1e-16 || dy <= 1e-16 || !std::isfinite(dx) || !std::isfinite(dy)) return;
        const double target = std::clamp(dy / dx, 1e-12, 1e12);
        const double sm = std::clamp(o.primal_weight_smoothing, 0.0, 1.0);
        setWeight(std::exp(sm * std::log(target) + (1.0 - sm) * std::log(hs.omega)));
    }
    void setWeight(double w) {
        if (w > 0.0 && std::isfinite(w)) hs.omega = std::clamp(w, o.minimum_primal_weight, o.maximum_primal_weight);
    }

    void resetSteps(double safe) {   // StepController constructor (step_controller.cpp:17-28)
        hs.max_safe = safe;
        hs.min_step = o.minimum_step_scale * safe;
        hs.eta = std::clamp(o.initial_step_safety * safe, o.minimum_step_scale * safe, safe);
        hs.eta_trial = hs.eta;
        hs.omega = 1.0;
    }

    // Loads a working-coordinate start (from original-coordinate xo/yo) into
    // the precision-T iterate, refreshes A x, sets the anchors and clears the
    // average.
    template <typename T>
    void adoptOriginal(const double* xo, const double* yo) {
        IterBuf<T>& I = iters<T>();
        hs.parity = 0;
        k_to_working<T><<<streamGrid(X.n + X.m) * 4, kThreads, 0, s>>>(X.n, X.m, X.cs.p, X.rs.p, xo, yo,
                                                                        I.x[0].p, I.y[0].p);
        HIPCHK(hipGetLastError());
        refreshActivity<T>();
        copyT(I.anx.p, I.x[0].p, X.n);
        copyT(I.any.p, I.y[0].p, X.m);
        zeroT(I.avx.p, X.n);
        zeroT(I.avy.p, X.m);
        hs.avg_total = 0.0;
    }

    struct Best {
        Metrics m;
    } best;

    template <typename T>
    bool mainLoop(int& status, std::string& msg, int& restarts);
    template <typename T>
    void polish(int64_t& pIts, int64_t& pTrials, Metrics& pBest, bool& improved);
};

template <>
Working<double>& Run::work<double>() { return X.w64; }
template <>
Working<float>& Run::work<float>() { return X.w32; }
template <>
IterBuf<double>& Run::iters<double>() { return X.it64; }
template <>
IterBuf<float>& Run::iters<float>() { return X.it32; }

// Set-up on the device: Ruiz (scaling.cpp:55-136), the scaled problem, and
// Pock-Chambolle (preconditioner.cpp:31-80).
void Run::setup() {
    const int m = X.m, n = X.n;
    const int bt = 256;
    k_fill<double><<<gridFor(m), kThreads, 0, s>>>(m, X.rs.p, 1.0);
    k_fill<double><<<gridFor(n), kThreads, 0, s>>>(n, X.cs.p, 1.0);
    scaled = o.use_ruiz_scaling && o.ruiz_iterations > 0 && X.nnz > 0;
    if (scaled) {
        for (int pass = 0; pass < o.ruiz_iterations; ++pass) {
            k_ruiz_max<<<gridFor(m, bt), bt, 0, s>>>(m, X.A.ptr.p, X.A.idx.p, X.A.val.p, X.rs.p, X.cs.p, true,
                                                     X.rmax.p);
            k_ruiz_max<<<gridFor(n, bt), bt, 0, s>>>(n, X.AT.ptr.p, X.AT.idx.p, X.AT.val.p, X.cs.p, X.rs.p,
                                                     false, X.cmax.p);
            k_ruiz_update<<<gridFor(m, bt), bt, 0, s>>>(m, X.rs.p, X.rmax.p);
            k_ruiz_update<<<gridFor(n, bt), bt, 0, s>>>(n, X.cs.p, X.cmax.p);
        }
    }
    Working<double>& w = X.w64;
    const bool f32 = o.precision == PDLPX_FP32;
    float* v32r = f32 ? X.w32.vr.p : nullptr;
    float* v32c = f32 ? X.w32.vc.p : nullptr;
    k_scale_values<<<gridFor(m, bt), bt, 0, s>>>(m, X.A.ptr.p, X.A.idx.p, X.A.val.p, X.rs.p, X.cs.p, w.vr.p, v32r);
    k_scale_values<<<gridFor(n, bt), bt, 0, s>>>(n, X.AT.ptr.p, X.AT.idx.p, X.AT.val.p, X.cs.p, X.rs.p, w.vc.p,
                                                 v32c);
    k_scale_cols<<<gridFor(n, bt), bt, 0, s>>>(n, X.cs.p, X.c.p, X.lo.p, X.hi.p, w.c.p, w.lo.p, w.hi.p);
    k_scale_rows<<<gridFor(m, bt), bt, 0, s>>>(m, X.rs.p, X.rlo.p, X.rhi.p, w.rlo.p, w.rhi.p);
    if (o.use_diagonal_preconditioning) {
        k_abs_sums<<<gridFor(n, bt), bt, 0, s>>>(n, X.AT.ptr.p, w.vc.p, w.tau.p, w.tau_inv.p);
        k_abs_sums<<<gridFor(m, bt), bt, 0, s>>>(m, X.A.ptr.p, w.vr.p, w.sig.p, w.sig_inv.p);
    } else {
        for (auto* v : {&w.tau, &w.tau_inv}) k_fill<double><<<gridFor(n), kThreads, 0, s>>>(n, v->p, 1.0);
        for (auto* v : {&w.sig, &w.sig_inv}) k_fill<double><<<gridFor(m), kThreads, 0, s>>>(m, v->p, 1.0);
    }
    HIPCHK(hipGetLastError());
    if (f32) {
        Working<float>& h = X.w32;
        const int g = streamGrid(std::max(n, m)) * 4;
        for (auto pr : {std::make_pair(&h.c, &w.c), std::make_pair(&h.lo, &w.lo), std::make_pair(&h.hi, &w.hi),
                        std::make_pair(&h.tau, &w.tau), std::make_pair(&h.tau_inv, &w.tau_inv)})
            k_convert<float, double><<<g, kThreads, 0, s>>>(n, pr.second->p, pr.first->p);
        for (auto pr : {std::make_pair(&h.rlo, &w.rlo), std::make_pair(&h.rhi, &w.rhi),
                        std::make_pair(&h.sig, &w.sig), std::make_pair(&h.sig_inv, &w.sig_inv)})
            k_convert<float, double><<<g, kThreads, 0, s>>>(m, pr.second->p, pr.first->p);
        HIPCHK(hipGetLastError());
    }
    maxSafe = 0.99;
    if (!o.use_diagonal_preconditioning) {   // pdlp_solver.cpp:189-202
        const double norm = powerIteration();
        // powerIteration() stores convergence in hs.fail temporarily
        const bool conv = hs.fail == 0;
        hs.fail = 0;
        maxSafe = norm > 0.0 ? 0.99 / (norm * (conv ? 1.02 : 1.25)) : 1.0;
    }
}

// Power iteration on A^T A of the working matrix (preconditioner.cpp:82-140).
double Run::powerIteration() {
    hs.fail = 0;
    if (X.m == 0 || X.n == 0 || X.nnz == 0) return 0.0;
    std::mt19937 gen(1);
    std::uniform_real_distribution<double> dist(-1.0, 1.0);
    std::vector<double> x(static_cast<size_t>(X.n));
    for (double& e : x) e = dist(gen);
    double sq = 0.0;
    for (double e : x) sq += e * e;
    double norm = std::sqrt(sq);
    if (norm > 0.0) {
        const double inv = 1.0 / norm;
        for (double& e : x) e *= inv;
    } else {
        x.assign(x.size(), 1.0 / std::sqrt(static_cast<double>(x.size())));
    }
    double* dx = X.xo0.p;
    double* dy = X.yo0.p;
    double* dz = X.xo1.p;
    HIPCHK(hipMemcpyAsync(dx, x.data(), x.size() * 8, hipMemcpyHostToDevice, s));
    const int budget = std::max(o.power_iteration_count, 1);
    double prev = 0.0, est = 0.0;
    bool hit = false;
    const int gn = streamGrid(X.n);
    for (int it = 0; it < budget; ++it) {
        k_spmv<double><<<X.A.nblk, kThreads, 0, s>>>(X.A.view<double>(X.w64.vr.p), dx, dy);
        k_spmv<double><<<X.AT.nblk, kThreads, 0, s>>>(X.AT.view<double>(X.w64.vc.p), dy, dz);
        k_dot<<<gn, kThreads, 0, s>>>(X.n, dx, dz, X.part_s.p);
        k_dot<<<gn, kThreads, 0, s>>>(X.n, dz, dz, X.part_s.p + gn);
        HIPCHK(hipGetLastError());
        const auto v = reduce({{X.part_s.p, gn}, {X.part_s.p + gn, gn}});
        est = std::sqrt(std::max(v[0], 0.0));
        const double zn = std::sqrt(v[1]);
        if (!(zn > 0.0)) {
            hs.fail = 0;
            return est;
        }
        k_scale_vec<<<gridFor(X.n), kThreads, 0, s>>>(X.n, dz, 1.0 / zn);
        std::swap(dx, dz);
        if (it > 0 && est > 0.0 && std::abs(est - prev) <= 1e-8 * est) {
            hit = true;
            break;
        }
        prev = est;
    }
    hs.fail = hit ? 0 : 1;
    return est;
}

// The PDLP outer loop (pdlp_solver.cpp:277-477) in precision T.  Returns true
// when the solve is finished (status/msg set); false when it ran out of
// iterations or time (status/msg hold the limit) or, in the FP32 phase, when it
// decided to hand over to FP64.
template <typename T>
bool Run::mainLoop(int& status, std::string& msg, int& restarts) {
    IterBuf<T>& I = iters<T>();
    const bool f32phase = sizeof(T) == 4;
    const int freq = o.termination_check_frequency;
    double baseline = best.m.kkt;   // restart baseline
    double prevScore = kInf;        // RestartController::previousCandidateScore_
    int64_t sinceRestart = 0;
    int stall = 0;
    double lastBest = best.m.kkt;
    while (hs.iteration < o.iteration_limit) {
        if (o.time_limit_seconds > 0.0 && since() >= o.time_limit_seconds) {
            status = PDLPX_TIME_LIMIT;
            msg = "Time limit reached";
            return false;
        }
        const int64_t target = std::min<int64_t>((hs.iteration / freq + 1) * freq, o.iteration_limit);
        const int64_t before = hs.iteration;
        runTrials<T>(target);
        sinceRestart += hs.iteration - before;
        if (f32phase) fp32Its += hs.iteration - before;
        if (hs.fail == 1) {
            status = PDLPX_NUMERICAL_FAILURE;
            msg = "PDHG generated a non-finite iterate";
            return true;
        }
        if (hs.fail == 2) {
            status = PDLPX_NUMERICAL_FAILURE;
            msg = "Adaptive linesearch failed to accept a step";
            return true;
        }
        if (hs.iteration % freq != 0) continue;   // stopped at the iteration limit

        // ---- termination check (pdlp_solver.cpp:354-435) ----
        const int p = hs.parity;
        const bool haveAvg = o.use_averaging && hs.avg_total != 0.0;
        toOriginal<T>(I.x[p].p, I.y[p].p, X.xo0.p, X.yo0.p);
        if (haveAvg) toOriginal<T>(I.avx.p, I.avy.p, X.xo1.p, X.yo1.p);
        else {
            copyD(X.xo1.p, X.xo0.p, X.n);
            copyD(X.yo1.p, X.yo0.p, X.m);
        }
        Metrics cur, avg;
        scorePair(cur, avg);
        if (!haveAvg) avg = cur;
        if (cur.kkt < best.m.kkt) {
            copyD(X.bestx.p, X.xo0.p, X.n);
            copyD(X.besty.p, X.yo0.p, X.m);
            best.m = cur;
        }
        if (haveAvg && avg.kkt < best.m.kkt) {
            copyD(X.bestx.p, X.xo1.p, X.n);
            copyD(X.besty.p, X.yo1.p, X.m);
            best.m = avg;
        }
        if (o.detect_infeasibility) {
            const int ray = rayTests<T>();
            if (ray == 1) {
                status = PDLPX_INFEASIBLE;
                msg = "Primal infeasible: Farkas ray found";
                return true;
            }
            if (ray == 2) {
                status = PDLPX_UNBOUNDED;
                msg = "Unbounded: improving ray found";
                return true;
            }
        }
        if (optimal(best.m)) {
            status = PDLPX_OPTIMAL;
            msg = "Primal, dual, and gap tolerances satisfied";
            return true;
        }
        if (f32phase) {   // FP32 -> FP64 hand-off
            const double tol = std::max(o.primal_tolerance, std::max(o.dual_tolerance, o.gap_tolerance));
            stall = best.m.kkt < lastBest ? 0 : stall + 1;
            lastBest = std::min(lastBest, best.m.kkt);
            if (best.m.kkt <= o.fp32_handoff_factor * tol ||
                (o.fp32_stall_checks > 0 && stall >= o.fp32_stall_checks))
                return false;
        }

        // ---- restart (restart_controller.cpp:8-63, pdlp_solver.cpp:437-476) ----
        int pick = 0;   // 1 current, 2 average
        double score = kInf;
        if (o.use_restarts && sinceRestart >= o.minimum_restart_iterations) {
            const bool af = std::isfinite(avg.kkt), cf = std::isfinite(cur.kkt);
            if (af || cf) {
                int cand = 2;
                score = avg.kkt;
                if (!af || (cf && cur.kkt < avg.kkt)) {
                    cand = 1;
                    score = cur.kkt;
                }
                const double prev = prevScore;
                prevScore = score;
                const bool bf = std::isfinite(baseline);
                const auto art = static_cast<int64_t>(o.restart_artificial_fraction * static_cast<double>(hs.iteration));
                if (bf && score <= o.restart_sufficient_decay * baseline) pick = cand;
                else if (bf && score <= o.restart_necessary_decay * baseline && std::isfinite(prev) && score > prev)
                    pick = cand;
                else if (sinceRestart >= std::max<int64_t>(art, 1) || sinceRestart >= o.maximum_restart_iterations)
                    pick = cand;
            }
        }
        if (pick != 0) {
            const bool fromAvg = pick == 2 && haveAvg;
            const T* sx = fromAvg ? I.avx.p : I.x[p].p;
            const T* sy = fromAvg ? I.avy.p : I.y[p].p;
            const auto d = displacement<T>(sx, sy);
            reweight(std::sqrt(d.first), std::sqrt(d.second));
            if (fromAvg) {
                copyT(I.x[p].p, I.avx.p, X.n);
                copyT(I.y[p].p, I.avy.p, X.m);
                refreshActivity<T>();
            }
            copyT(I.anx.p, I.x[p].p, X.n);
            copyT(I.any.p, I.y[p].p, X.m);
            baseline = score;
            zeroT(I.avx.p, X.n);
            zeroT(I.avy.p, X.m);
            hs.avg_total = 0.0;
            prevScore = kInf;
            sinceRestart = 0;
            ++restarts;
        }
    }
    status = PDLPX_ITERATION_LIMIT;
    msg = "Iteration limit reached";
    return false;
}

// FeasibilityPolisher::polish (feasibility_polishing.cpp:13-145): a fresh step
// controller from the best point at half the last step, averaging, checks at
// the usual cadence, keep the best complete KKT score.
template <typename T>
void Run::polish(int64_t& pIts, int64_t& pTrials, Metrics& pBest, bool& improved) {
    IterBuf<T>& I = iters<T>();
    pIts = 0;
    pTrials = 0;
    improved = false;
    if (!o.use_feasibility_polishing || o.polishing_iterations <= 0) return;
    double remaining = 0.0;
    if (o.time_limit_seconds > 0.0) remaining = o.time_limit_seconds - since();
    const double lastEta = hs.eta;
    const double omega = hs.omega;
    // best (original) -> working start
    adoptOriginal<T>(X.bestx.p, X.besty.p);
    resetSteps(hs.max_safe);
    setWeight(omega);
    hs.eta_trial = 0.5 * lastEta;
    hs.iteration = 0;
    hs.attempts = 0;
    hs.fail = 0;
    const int64_t trials0 = hs.trials;
    const auto pt0 = Clock::now();
    const int freq = std::max(o.termination_check_frequency, 1);
    while (hs.iteration < o.polishing_iterations) {
        const int64_t target = std::min<int64_t>((hs.iteration / freq + 1) * freq, o.polishing_iterations);
        runTrials<T>(target);
        pIts = hs.iteration;
        pTrials = hs.trials - trials0;
        if (hs.fail) break;
        if (remaining > 0.0 &&
            std::chrono::duration<double>(Clock::now() - pt0).count() >= remaining)
            break;
        const int p = hs.parity;
        const bool haveAvg = o.use_averaging && hs.avg_total != 0.0;
        toOriginal<T>(I.x[p].p, I.y[p].p, X.xo0.p, X.yo0.p);
        if (haveAvg) toOriginal<T>(I.avx.p, I.avy.p, X.xo1.p, X.yo1.p);
        else {
            copyD(X.xo1.p, X.xo0.p, X.n);
            copyD(X.yo1.p, X.yo0.p, X.m);
        }
        Metrics cur, avg;
        scorePair(cur, avg);
        if (cur.kkt < pBest.kkt) {
            copyD(X.proj.p, X.xo0.p, X.n);   // polish-best kept in proj/u
            copyD(X.u.p, X.yo0.p, X.m);
            pBest = cur;
            improved = true;
        }
        if (haveAvg && avg.kkt < pBest.kkt) {
            copyD(X.proj.p, X.xo1.p, X.n);
            copyD(X.u.p, X.yo1.p, X.m);
            pBest = avg;
            improved = true;
        }
        if (optimal(pBest)) break;
    }
}

void Run::go(const double* warmX, const double* warmY, pdlpx_summary* sum, double* xOut, double* yOut,
             double* rayOut) {
    std::memset(sum, 0, sizeof(*sum));
    sum->final_primal_weight = 1.0;
    const int n = X.n, m = X.m;
    hipEvent_t e0 = X.ev0, e1 = X.ev1;
    HIPCHK(hipEventRecord(e0, s));
    setup();
    prec = o.precision == PDLPX_FP32 ? PDLPX_FP32 : PDLPX_FP64;

    // start point (working coordinates, FP64), original copy for scoring
    const double* dwx = nullptr;
    const double* dwy = nullptr;
    if (warmX && n) {
        HIPCHK(hipMemcpyAsync(X.warmx.p, warmX, sizeof(double) * n, hipMemcpyHostToDevice, s));
        dwx = X.warmx.p;
    }
    if (warmY && m) {
        HIPCHK(hipMemcpyAsync(X.warmy.p, warmY, sizeof(double) * m, hipMemcpyHostToDevice, s));
        dwy = X.warmy.p;
    }
    IterBuf<double>& I = X.it64;
    k_start<<<gridFor(n + m), kThreads, 0, s>>>(n, m, X.w64.lo.p, X.w64.hi.p, X.w64.rlo.p, X.w64.rhi.p, X.cs.p,
                                                X.rs.p, dwx, dwy, I.x[0].p, I.y[0].p);
    HIPCHK(hipGetLastError());
    std::memset(&hs, 0, sizeof(hs));
    resetSteps(maxSafe);
    setWeight(o.initial_primal_weight);
    hs.adaptive = o.use_adaptive_linesearch ? 1 : 0;
    hs.averaging = o.use_averaging ? 1 : 0;
    hs.max_trials = std::max(o.maximum_step_trials, 1);

    // score the start (pdlp_solver.cpp:239-264)
    toOriginal<double>(I.x[0].p, I.y[0].p, X.xo0.p, X.yo0.p);
    copyD(X.xo1.p, X.xo0.p, n);
    copyD(X.yo1.p, X.yo0.p, m);
    Metrics m0, m1;
    scorePair(m0, m1);
    copyD(X.bestx.p, X.xo0.p, n);
    copyD(X.besty.p, X.yo0.p, m);
    best.m = m0;
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipStreamSynchronize(s));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    setupSeconds = ms * 1e-3;

    int status = PDLPX_ITERATION_LIMIT;
    std::string msg = "Iteration limit reached";
    int restarts = 0;
    bool diag = false;
    int64_t extraIts = 0, extraTrials = 0;
    int rayKind = 0;

    auto fetchRay = [&](int kind) { rayKind = kind; };

    if (optimal(best.m)) {
        status = PDLPX_OPTIMAL;
        msg = "Initial point satisfies the requested tolerances";
    } else {
        diag = true;
        HIPCHK(hipEventRecord(e0, s));
        // FP64 working start -> precision of the first phase
        if (prec == PDLPX_FP32) {
            IterBuf<float>& F = X.it32;
            k_convert<float, double><<<streamGrid(n) * 4, kThreads, 0, s>>>(n, I.x[0].p, F.x[0].p);
            k_convert<float, double><<<streamGrid(m) * 4, kThreads, 0, s>>>(m, I.y[0].p, F.y[0].p);
            HIPCHK(hipGetLastError());
            refreshActivity<float>();
            copyT(F.anx.p, F.x[0].p, n);
            copyT(F.any.p, F.y[0].p, m);
            zeroT(F.avx.p, n);
            zeroT(F.avy.p, m);
        } else {
            refreshActivity<double>();
            copyT(I.anx.p, I.x[0].p, n);
            copyT(I.any.p, I.y[0].p, m);
            zeroT(I.avx.p, n);
            zeroT(I.avy.p, m);
        }
        bool finished = false;
        if (prec == PDLPX_FP32) {
            finished = mainLoop<float>(status, msg, restarts);
            if (finished && (status == PDLPX_INFEASIBLE || status == PDLPX_UNBOUNDED)) fetchRay(status);
            if (!finished && hs.iteration < o.iteration_limit &&
                !(status == PDLPX_TIME_LIMIT)) {
                // hand over: continue in FP64 from the best point, as after a restart
                const int64_t its = hs.iteration, trials = hs.trials;
                const double eta = hs.eta, omega = hs.omega;
                adoptOriginal<double>(X.bestx.p, X.besty.p);
                hs.iteration = its;
                hs.trials = trials;
                hs.eta = eta;
                hs.eta_trial = eta;
                hs.omega = omega;
                prec = PDLPX_FP64;
                finished = mainLoop<double>(status, msg, restarts);
            }
        } else {
            finished = mainLoop<double>(status, msg, restarts);
        }
        if (finished && (status == PDLPX_INFEASIBLE || status == PDLPX_UNBOUNDED)) fetchRay(status);
        const int64_t mainIts = hs.iteration, mainTrials = hs.trials;
        const double mainEta = hs.eta, mainOmega = hs.omega, mainSafe = hs.max_safe;
        HIPCHK(hipEventRecord(e1, s));
        HIPCHK(hipStreamSynchronize(s));
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        loopSeconds = ms * 1e-3;
        if (!finished && o.use_feasibility_polishing && n > 0) {
            const bool expired = o.time_limit_seconds > 0.0 && o.time_limit_seconds - since() <= 0.0;
            if (!expired) {
                int64_t pIts = 0, pTrials = 0;
                Metrics pBest = best.m;
                bool improved = false;
                if (prec == PDLPX_FP32) polish<float>(pIts, pTrials, pBest, improved);
                else polish<double>(pIts, pTrials, pBest, improved);
                if (pBest.kkt < best.m.kkt && improved) {
                    copyD(X.bestx.p, X.proj.p, n);
                    copyD(X.besty.p, X.u.p, m);
                    best.m = pBest;
                }
                if (optimal(best.m)) {
                    status = PDLPX_OPTIMAL;
                    msg = "Tolerances satisfied after refinement polishing";
                    extraIts = pIts;
                    extraTrials = pTrials;
                }
            }
        }
        hs.iteration = mainIts + extraIts;
        hs.trials = mainTrials + extraTrials;
        hs.eta = mainEta;
        hs.omega = mainOmega;
        hs.max_safe = mainSafe;
    }

    // results (pdlp_solver.cpp:43-68,267-272)
    if (n && xOut) HIPCHK(hipMemcpyAsync(xOut, X.bestx.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    if (m && yOut) HIPCHK(hipMemcpyAsync(yOut, X.besty.p, sizeof(double) * m, hipMemcpyDeviceToHost, s));
    if (rayOut && rayKind == PDLPX_INFEASIBLE && m)
        HIPCHK(hipMemcpyAsync(rayOut, X.ry.p, sizeof(double) * m, hipMemcpyDeviceToHost, s));
    if (rayOut && rayKind == PDLPX_UNBOUNDED && n)
        HIPCHK(hipMemcpyAsync(rayOut, X.rx.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    sum->status = status;
    setStatusText(sum, msg);
    sum->primal_objective = best.m.pobj;
    sum->dual_objective = best.m.dobj;
    sum->primal_residual = best.m.pres;
    sum->dual_residual = best.m.dres;
    sum->relative_gap = best.m.gap;
    sum->iterations = diag ? hs.iteration : 0;
    sum->step_trials = diag ? hs.trials : 0;
    sum->restart_count = restarts;
    sum->final_step_size = diag ? hs.eta : 0.0;
    sum->static_step_bound = diag ? hs.max_safe : 0.0;
    sum->final_primal_weight = diag ? hs.omega : 1.0;
    sum->has_dual_ray = rayKind == PDLPX_INFEASIBLE ? 1 : 0;
    sum->has_primal_ray = rayKind == PDLPX_UNBOUNDED ? 1 : 0;
    sum->fp32_iterations = fp32Its;
    sum->device_loop_seconds = loopSeconds;
    sum->setup_seconds = setupSeconds;
    sum->termination_checks = checks;
    sum->solve_time_seconds = since();
}

void allocWork(pdlpx_ctx& X) {
    const int n = X.n, m = X.m;
    const size_t nz = static_cast<size_t>(X.nnz);
    for (auto* b : {&X.rs, &X.rmax, &X.yo0, &X.yo1, &X.besty, &X.ry, &X.u, &X.warmy}) b->alloc(m);
    for (auto* b : {&X.cs, &X.cmax, &X.xo0, &X.xo1, &X.bestx, &X.rx, &X.proj, &X.warmx}) b->alloc(n);
    // power iteration reuses xo0 (n), yo0 (m), xo1 (n)
    Working<double>& w = X.w64;
    w.vr.alloc(nz);
    w.vc.alloc(nz);
    for (auto* b : {&w.c, &w.lo, &w.hi, &w.tau, &w.tau_inv}) b->alloc(n);
    for (auto* b : {&w.rlo, &w.rhi, &w.sig, &w.sig_inv}) b->alloc(m);
    X.it64.alloc(n, m);
    X.part_p.alloc(X.AT.nblk);
    X.part_d.alloc(X.A.nblk);
    X.part_i.alloc(X.A.nblk);
    X.part_rows.alloc(6 * size_t(X.A.nblk));
    X.part_cols.alloc(6 * size_t(X.AT.nblk));
    X.part_s.alloc(2 * kStreamBlocks);
    X.part_s2.alloc(2 * kStreamBlocks);
    X.part_s3.alloc(2 * kStreamBlocks);
    X.red.alloc(32);
    X.st.alloc(1);
}

void allocFp32(pdlpx_ctx& X) {
    const int n = X.n, m = X.m;
    const size_t nz = static_cast<size_t>(X.nnz);
    Working<float>& h = X.w32;
    h.vr.alloc(nz);
    h.vc.alloc(nz);
    for (auto* b : {&h.c, &h.lo, &h.hi, &h.tau, &h.tau_inv}) b->alloc(n);
    for (auto* b : {&h.rlo, &h.rhi, &h.sig, &h.sig_inv}) b->alloc(m);
    X.it32.alloc(n, m);
}

template <typename F>
int guarded(pdlpx_ctx* X, F&& body) {
    if (!X) return PDLPX_E_INVALID;
    try {
        HIPCHK(hipSetDevice(X->device));
        body();
        return PDLPX_OK;
    } catch (const Failure& f) {
        X->error = f.what;
        return f.code;
    } catch (const std::bad_alloc&) {
        X->error = "host allocation failed";
        return PDLPX_E_NOMEM;
    } catch (const std::exception& e) {
        X->error = e.what();
        return PDLPX_E_INVALID;
    }
}

}  // namespace

extern "C" {

void pdlpx_default_options(pdlpx_options* o) {   // pdlp_options.h:7-75
    std::memset(o, 0, sizeof(*o));
    o->iteration_limit = 200000;
    o->time_limit_seconds = 0.0;
    o->primal_tolerance = 1e-6;
    o->dual_tolerance = 1e-6;
    o->gap_tolerance = 1e-6;
    o->termination_check_frequency = 100;
    o->power_iteration_count = 100;
    o->thread_count = 0;
    o->parallel_nonzero_threshold = 20000;
    o->initial_step_safety = 0.90;
    o->minimum_step_scale = 1e-6;
    o->maximum_primal_weight = 1e6;
    o->minimum_primal_weight = 1e-6;
    o->use_ruiz_scaling = 1;
    o->ruiz_iterations = 10;
    o->detect_infeasibility = 1;
    o->infeasibility_tolerance = 1e-8;
    o->use_diagonal_preconditioning = 1;
    o->use_adaptive_linesearch = 1;
    o->maximum_step_trials = 60;
    o->use_averaging = 1;
    o->use_restarts = 1;
    o->use_feasibility_polishing = 1;
    o->minimum_restart_iterations = 64;
    o->maximum_restart_iterations = 100000;
    o->restart_sufficient_decay = 0.20;
    o->restart_necessary_decay = 0.80;
    o->restart_artificial_fraction = 0.36;
    o->initial_primal_weight = 1.0;
    o->primal_weight_smoothing = 0.50;
    o->polishing_iterations = 5000;
    o->precision = PDLPX_FP64;
    o->fp32_handoff_factor = 10.0;
    o->fp32_stall_checks = 20;
}

int pdlpx_create(pdlpx_ctx** out, int device) {
    if (!out) return PDLPX_E_INVALID;
    *out = nullptr;
    auto* X = new (std::nothrow) pdlpx_ctx();
    if (!X) return PDLPX_E_NOMEM;
    X->device = device;
    const int rc = guarded(X, [&] {
        int count = 0;
        HIPCHK(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) throw Failure{PDLPX_E_INVALID, "no such HIP device"};
        HIPCHK(hipStreamCreateWithFlags(&X->stream, hipStreamNonBlocking));
        HIPCHK(hipEventCreate(&X->ev0));
        HIPCHK(hipEventCreate(&X->ev1));
    });
    if (rc != PDLPX_OK) {
        delete X;
        return rc;
    }
    *out = X;
    return PDLPX_OK;
}

void pdlpx_destroy(pdlpx_ctx* X) {
    if (!X) return;
    (void)hipSetDevice(X->device);
    X->dropGraphs();
    if (X->stream) (void)hipStreamSynchronize(X->stream);
    if (X->ev0) (void)hipEventDestroy(X->ev0);
    if (X->ev1) (void)hipEventDestroy(X->ev1);
    if (X->stream) (void)hipStreamDestroy(X->stream);
    delete X;
}

const char* pdlpx_last_error(const pdlpx_ctx* X) { return X ? X->error.c_str() : "null context"; }

int pdlpx_load(pdlpx_ctx* X, const pdlpx_problem* p, int32_t precision) {
    if (!p) return PDLPX_E_INVALID;
    return guarded(X, [&] {
        X->dropGraphs();
        X->loaded = true;
        X->invalid = validateProblem(*p);
        X->valid = X->invalid.empty();
        if (!X->valid) return;
        hipStream_t s = X->stream;
        X->m = p->rows;
        X->n = p->cols;
        X->nnz = p->nnz;
        X->offset = p->objective_offset;
        uploadOrient(X->A, p->rows, p->csr_row_start, p->csr_col, p->csr_val, p->nnz, s);
        uploadOrient(X->AT, p->cols, p->csc_col_start, p->csc_row, p->csc_val, p->nnz, s);
        X->c.upload(p->objective, p->cols, s);
        X->lo.upload(p->var_lower, p->cols, s);
        X->hi.upload(p->var_upper, p->cols, s);
        X->rlo.upload(p->row_lower, p->rows, s);
        X->rhi.upload(p->row_upper, p->rows, s);
        // termination normalisers (termination.cpp:15-37,60-61), host order
        double sq = 0.0;
        for (int i = 0; i < p->rows; ++i) {
            if (std::isfinite(p->row_lower[i])) sq += p->row_lower[i] * p->row_lower[i];
            if (std::isfinite(p->row_upper[i]) && p->row_upper[i] != p->row_lower[i])
                sq += p->row_upper[i] * p->row_upper[i];
        }
        X->pnorm = 1.0 + std::sqrt(sq);
        double cq = 0.0;
        for (int j = 0; j < p->cols; ++j) cq += p->objective[j] * p->objective[j];
        X->dnorm = 1.0 + std::sqrt(cq);
        allocWork(*X);
        if (precision == PDLPX_FP32) allocFp32(*X);
        HIPCHK(hipStreamSynchronize(s));
    });
}

int pdlpx_solve(pdlpx_ctx* X, const pdlpx_options* options, const double* warm_x, const double* warm_y,
                pdlpx_summary* summary, double* x_out, double* y_out, double* ray_out) {
    if (!options || !summary) return PDLPX_E_INVALID;
    const auto t0 = std::chrono::steady_clock::now();
    return guarded(X, [&] {
        std::memset(summary, 0, sizeof(*summary));
        summary->final_primal_weight = 1.0;
        if (!X->loaded) throw Failure{PDLPX_E_STATE, "pdlpx_solve called before pdlpx_load"};
        auto reject = [&](const std::string& why) {
            summary->status = PDLPX_INVALID_PROBLEM;
            setStatusText(summary, why);
            summary->primal_objective = std::numeric_limits<double>::quiet_NaN();
            summary->dual_objective = -kInf;
            summary->primal_residual = summary->dual_residual = summary->relative_gap = kInf;
            summary->solve_time_seconds =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        };
        if (!X->valid) return reject(X->invalid);
        if (!optionsValid(*options)) return reject("Invalid PDLP options");
        if (options->precision == PDLPX_FP32 && !X->w32.vr.p) allocFp32(*X);
        Run run(*X, *options);
        run.go(warm_x, warm_y, summary, x_out, y_out, ray_out);
    });
}

int pdlpx_spmv(pdlpx_ctx* X, int32_t transpose, const double* x, double* y) {
    if (!x || !y) return PDLPX_E_INVALID;
    return guarded(X, [&] {
        if (!X->loaded || !X->valid) throw Failure{PDLPX_E_STATE, "no valid problem loaded"};
        hipStream_t s = X->stream;
        const int nin = transpose ? X->m : X->n;
        const int nout = transpose ? X->n : X->m;
        double* din = transpose ? X->yo0.p : X->xo0.p;
        double* dout = transpose ? X->xo1.p : X->yo1.p;
        if (nin) HIPCHK(hipMemcpyAsync(din, x, sizeof(double) * nin, hipMemcpyHostToDevice, s));
        const OrientBuf& M = transpose ? X->AT : X->A;
        if (nout) {
            k_spmv<double><<<M.nblk, kThreads, 0, s>>>(M.view<double>(M.val.p), din, dout);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(y, dout, sizeof(double) * nout, hipMemcpyDeviceToHost, s));
        }
        HIPCHK(hipStreamSynchronize(s));
    });
}

int pdlpx_spmv_device(pdlpx_ctx* X, int32_t transpose, const void* x, void* y) {
    if (!x || !y) return PDLPX_E_INVALID;
    return guarded(X, [&] {
        if (!X->loaded || !X->valid) throw Failure{PDLPX_E_STATE, "no valid problem loaded"};
        const OrientBuf& M = transpose ? X->AT : X->A;
        k_spmv<double><<<M.nblk, kThreads, 0, X->stream>>>(M.view<double>(M.val.p),
                                                            static_cast<const double*>(x),
                                                            static_cast<double*>(y));
        HIPCHK(hipGetLastError());
    });
}

int pdlpx_scaling(pdlpx_ctx* X, int32_t ruiz_iterations, int32_t use_pc, double* row_scale, double* col_scale,
                  double* primal_scale, double* dual_scale) {
    return guarded(X, [&] {
        if (!X->loaded || !X->valid) throw Failure{PDLPX_E_STATE, "no valid problem loaded"};
        pdlpx_options o;
        pdlpx_default_options(&o);
        o.ruiz_iterations = ruiz_iterations;
        o.use_ruiz_scaling = ruiz_iterations > 0;
        o.use_diagonal_preconditioning = use_pc;
        Run run(*X, o);
        run.setupOnly();
        hipStream_t s = X->stream;
        const int m = X->m, n = X->n;
        if (row_scale && m) HIPCHK(hipMemcpyAsync(row_scale, X->rs.p, 8 * m, hipMemcpyDeviceToHost, s));
        if (col_scale && n) HIPCHK(hipMemcpyAsync(col_scale, X->cs.p, 8 * n, hipMemcpyDeviceToHost, s));
        if (primal_scale && n) HIPCHK(hipMemcpyAsync(primal_scale, X->w64.tau.p, 8 * n, hipMemcpyDeviceToHost, s));
        if (dual_scale && m) HIPCHK(hipMemcpyAsync(dual_scale, X->w64.sig.p, 8 * m, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    });
}

int pdlpx_time_trial(pdlpx_ctx* X, int32_t precision, int32_t reps, double* ms_primal, double* ms_dual) {
    if (!ms_primal || !ms_dual || reps <= 0) return PDLPX_E_INVALID;
    return guarded(X, [&] {
        if (!X->loaded || !X->valid) throw Failure{PDLPX_E_STATE, "no valid problem loaded"};
        if (precision == PDLPX_FP32 && !X->w32.vr.p) throw Failure{PDLPX_E_STATE, "no FP32 solve has run"};
        hipStream_t s = X->stream;
        DevState saved{}, probe{};
        HIPCHK(hipMemcpyAsync(&saved, X->st.p, sizeof(DevState), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        probe = saved;
        probe.done = 0;
        HIPCHK(hipMemcpyAsync(X->st.p, &probe, sizeof(DevState), hipMemcpyHostToDevice, s));
        auto timeIt = [&](auto&& launch) {
            HIPCHK(hipEventRecord(X->ev0, s));
            for (int r = 0; r < reps; ++r) launch();
            HIPCHK(hipEventRecord(X->ev1, s));
            HIPCHK(hipEventSynchronize(X->ev1));
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, X->ev0, X->ev1));
            return static_cast<double>(ms) / reps;
        };
        if (precision == PDLPX_FP32) {
            Iterates<float> it = X->it32.view();
            const Working<float>& w = X->w32;
            ColData<float> cd{w.c.p, w.lo.p, w.hi.p, w.tau.p, w.tau_inv.p};
            RowData<float> rd{w.rlo.p, w.rhi.p, w.sig.p, w.sig_inv.p};
            *ms_primal = timeIt([&] {
                k_primal<float><<<X->AT.nblk, kThreads, 0, s>>>(X->AT.view<float>(w.vc.p), cd, it, X->st.p, X->part_p.p);
            });
            *ms_dual = timeIt([&] {
                k_dual<float><<<X->A.nblk, kThreads, 0, s>>>(X->A.view<float>(w.vr.p), rd, it, X->st.p, X->part_d.p,
                                                             X->part_i.p);
            });
        } else {
            Iterates<double> it = X->it64.view();
            const Working<double>& w = X->w64;
            ColData<double> cd{w.c.p, w.lo.p, w.hi.p, w.tau.p, w.tau_inv.p};
            RowData<double> rd{w.rlo.p, w.rhi.p, w.sig.p, w.sig_inv.p};
            *ms_primal = timeIt([&] {
                k_primal<double><<<X->AT.nblk, kThreads, 0, s>>>(X->AT.view<double>(w.vc.p), cd, it, X->st.p,
                                                                 X->part_p.p);
            });
            *ms_dual = timeIt([&] {
                k_dual<double><<<X->A.nblk, kThreads, 0, s>>>(X->A.view<double>(w.vr.p), rd, it, X->st.p,
                                                              X->part_d.p, X->part_i.p);
            });
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(X->st.p, &saved, sizeof(DevState), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
    });
}

int pdlpx_solve_batch(pdlpx_ctx* X, int32_t count, const pdlpx_problem* problems, const pdlpx_options* options,
                      pdlpx_summary* summaries, double* const* x_out, double* const* y_out) {
    if (count < 0 || (count > 0 && (!problems || !summaries || !options))) return PDLPX_E_INVALID;
    // Sequential reference implementation of the batch entry; the batched
    // persistent kernel replaces this.
    for (int k = 0; k < count; ++k) {
        int rc = pdlpx_load(X, &problems[k], options->precision);
        if (rc != PDLPX_OK) return rc;
        rc = pdlpx_solve(X, options, nullptr, nullptr, &summaries[k], x_out ? x_out[k] : nullptr,
                         y_out ? y_out[k] : nullptr, nullptr);
        if (rc != PDLPX_OK) return rc;
    }
    return PDLPX_OK;
}

}  // extern "C"
