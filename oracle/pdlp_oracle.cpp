// CPU ORACLE -- test infrastructure only.
//
// A serial C++17 restatement of the reference PDLP engine
// (pdlp_engine/src/*.cpp of Aryan-Kumar06/optimisation_solver).  It exists to
// check the MI355X engine: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and never as the thing measured or shipped.
//
// Every function reproduces the reference's floating-point operation order
// (four-accumulator sparse dot products, the Moreau dual prox, the k+2 step
// schedule, ...) so that, built with the reference's flags
// (-O3 -ffp-contract=fast, no -march), it matches the reference's serial
// (threadCount = 1) run bit for bit.  tests/test_oracle_pin.py checks that
// against oracle/_ref (the reference compiled from its own sources) and
// against the reference's closed-form test problems.
//
// The only additions are a warm-start input (projected into the box and the
// dual cone, the engine extension the reference lacks) and a per-check trace.

#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <random>
#include <string>
#include <utility>
#include <vector>

namespace {

using Vec = std::vector<double>;
constexpr double kInf = std::numeric_limits<double>::infinity();

// ---------------------------------------------------------------------------
// Matrix in both orientations (sparse_matrix.h:52-135).
// ---------------------------------------------------------------------------
struct Csx {
    int rows = 0;
    int cols = 0;
    std::vector<int64_t> rp;   // CSR row start
    std::vector<int32_t> ci;   // CSR column index
    Vec rv;                    // CSR values
    std::vector<int64_t> cp;   // CSC column start
    std::vector<int32_t> ri;   // CSC row index
    Vec cv;                    // CSC values
};

// COO -> CSR -> CSC.  sparse_matrix.cpp:76-201: counting sort by row, per-row
// std::sort by column, duplicates summed in sorted order, exact zeros dropped,
// CSC by a row-order scan so each column is sorted by row.
Csx buildFromTriplets(int rows, int cols, int64_t count, const int32_t* tr,
                      const int32_t* tc, const double* tv) {
    Csx m;
    m.rows = rows;
    m.cols = cols;
    std::vector<int64_t> start(static_cast<size_t>(rows) + 1, 0);
    for (int64_t t = 0; t < count; ++t) start[static_cast<size_t>(tr[t]) + 1]++;
    for (int r = 0; r < rows; ++r) start[r + 1] += start[r];
    std::vector<int32_t> sc(static_cast<size_t>(count));
    Vec sv(static_cast<size_t>(count));
    {
        std::vector<int64_t> fill(start);
        for (int64_t t = 0; t < count; ++t) {
            const int64_t at = fill[static_cast<size_t>(tr[t])]++;
            sc[at] = tc[t];
            sv[at] = tv[t];
        }
    }
    m.rp.assign(static_cast<size_t>(rows) + 1, 0);
    m.ci.resize(static_cast<size_t>(count));
    m.rv.resize(static_cast<size_t>(count));
    std::vector<std::pair<int, double>> line;
    int64_t out = 0;
    for (int r = 0; r < rows; ++r) {
        line.clear();
        for (int64_t k = start[r]; k < start[r + 1]; ++k) line.emplace_back(sc[k], sv[k]);
        std::sort(line.begin(), line.end(),
                  [](const std::pair<int, double>& a, const std::pair<int, double>& b) {
                      return a.first < b.first;
                  });
        size_t k = 0;
        while (k < line.size()) {
            const int col = line[k].first;
            double sum = line[k].second;
            for (++k; k < line.size() && line[k].first == col; ++k) sum += line[k].second;
            if (sum != 0.0) {
                m.ci[out] = col;
                m.rv[out] = sum;
                ++out;
            }
        }
        m.rp[r + 1] = out;
    }
    m.ci.resize(static_cast<size_t>(out));
    m.rv.resize(static_cast<size_t>(out));
    m.cp.assign(static_cast<size_t>(cols) + 1, 0);
    for (int64_t k = 0; k < out; ++k) m.cp[static_cast<size_t>(m.ci[k]) + 1]++;
    for (int c = 0; c < cols; ++c) m.cp[c + 1] += m.cp[c];
    m.ri.resize(static_cast<size_t>(out));
    m.cv.resize(static_cast<size_t>(out));
    std::vector<int64_t> fill(m.cp);
    for (int r = 0; r < rows; ++r) {
        for (int64_t k = m.rp[r]; k < m.rp[r + 1]; ++k) {
            const int64_t at = fill[static_cast<size_t>(m.ci[k])]++;
            m.ri[at] = r;
            m.cv[at] = m.rv[k];
        }
    }
    return m;
}

// Sparse dot product with four interleaved accumulators combined as
// (a0 + a1) + (a2 + a3): sparse_matrix.cpp:286-308, pdhg_kernel.cpp:75-90.
inline double dot4(const int64_t b, const int64_t e, const int32_t* idx, const double* val,
                   const double* v) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int64_t k = b;
    for (; k + 4 <= e; k += 4) {
        a0 += val[k] * v[idx[k]];
        a1 += val[k + 1] * v[idx[k + 1]];
        a2 += val[k + 2] * v[idx[k + 2]];
        a3 += val[k + 3] * v[idx[k + 3]];
    }
    for (; k < e; ++k) a0 += val[k] * v[idx[k]];
    return (a0 + a1) + (a2 + a3);
}

void csrTimes(const Csx& m, const double* x, double* y) {
    for (int r = 0; r < m.rows; ++r) y[r] = dot4(m.rp[r], m.rp[r + 1], m.ci.data(), m.rv.data(), x);
}
void cscTimes(const Csx& m, const double* y, double* z) {
    for (int c = 0; c < m.cols; ++c) z[c] = dot4(m.cp[c], m.cp[c + 1], m.ri.data(), m.cv.data(), y);
}

inline double clip(double v, double lo, double hi) { return std::max(lo, std::min(v, hi)); }

// ---------------------------------------------------------------------------
// LP (compiled_lp.h:17-63)
// ---------------------------------------------------------------------------
struct Lp {
    Csx a;
    Vec c, lo, hi, rlo, rhi;
    double offset = 0.0;
    int m() const { return a.rows; }
    int n() const { return a.cols; }
};

Lp lpFromProblem(const pdlpx_problem& p) {
    Lp lp;
    lp.a.rows = p.rows;
    lp.a.cols = p.cols;
    const size_t nz = static_cast<size_t>(p.nnz);
    lp.a.rp.assign(p.csr_row_start, p.csr_row_start + p.rows + 1);
    lp.a.ci.assign(p.csr_col, p.csr_col + nz);
    lp.a.rv.assign(p.csr_val, p.csr_val + nz);
    lp.a.cp.assign(p.csc_col_start, p.csc_col_start + p.cols + 1);
    lp.a.ri.assign(p.csc_row, p.csc_row + nz);
    lp.a.cv.assign(p.csc_val, p.csc_val + nz);
    lp.c.assign(p.objective, p.objective + p.cols);
    lp.lo.assign(p.var_lower, p.var_lower + p.cols);
    lp.hi.assign(p.var_upper, p.var_upper + p.cols);
    lp.rlo.assign(p.row_lower, p.row_lower + p.rows);
    lp.rhi.assign(p.row_upper, p.row_upper + p.rows);
    lp.offset = p.objective_offset;
    return lp;
}

// CompiledLp::validate (compiled_lp.h:32-62) and SparseMatrix::validate
// (sparse_matrix.cpp:203-241).  Returns an empty string when valid.
std::string validateLp(const Lp& lp) {
    const Csx& a = lp.a;
    const int64_t nz = static_cast<int64_t>(a.rv.size());
    bool ok = a.rows >= 0 && a.cols >= 0 && a.rp.size() == size_t(a.rows) + 1 &&
              a.cp.size() == size_t(a.cols) + 1 && a.ci.size() == a.rv.size() &&
              a.ri.size() == a.cv.size() && a.rv.size() == a.cv.size();
    if (ok) ok = a.rp.front() == 0 && a.cp.front() == 0 && a.rp.back() == nz && a.cp.back() == nz;
    for (size_t i = 1; ok && i < a.rp.size(); ++i) ok = a.rp[i] >= a.rp[i - 1];
    for (size_t j = 1; ok && j < a.cp.size(); ++j) ok = a.cp[j] >= a.cp[j - 1];
    for (size_t k = 0; ok && k < a.rv.size(); ++k)
        ok = a.ci[k] >= 0 && a.ci[k] < a.cols && std::isfinite(a.rv[k]);
    for (size_t k = 0; ok && k < a.cv.size(); ++k)
        ok = a.ri[k] >= 0 && a.ri[k] < a.rows && std::isfinite(a.cv[k]);
    if (!ok) return "CompiledLp contains an invalid sparse matrix";
    if (!std::isfinite(lp.offset)) return "Objective offset must be finite";
    for (int j = 0; j < lp.n(); ++j)
        if (!std::isfinite(lp.c[j]) || std::isnan(lp.lo[j]) || std::isnan(lp.hi[j]) || lp.lo[j] > lp.hi[j])
            return "Invalid objective coefficient or variable bounds";
    for (int i = 0; i < lp.m(); ++i)
        if (std::isnan(lp.rlo[i]) || std::isnan(lp.rhi[i]) || lp.rlo[i] > lp.rhi[i])
            return "Invalid row bounds";
    return "";
}

// ---------------------------------------------------------------------------
// Ruiz equilibration (scaling.cpp:55-136).
// ---------------------------------------------------------------------------
struct Scaling {
    Vec rs, cs;
    Lp lp;
};

void ruizScales(const Lp& lp, int passes, Vec& rs, Vec& cs) {
    const Csx& a = lp.a;
    rs.assign(static_cast<size_t>(lp.m()), 1.0);
    cs.assign(static_cast<size_t>(lp.n()), 1.0);
    Vec rmax(rs.size()), cmax(cs.size());
    for (int p = 0; p < std::max(passes, 0); ++p) {
        std::fill(rmax.begin(), rmax.end(), 0.0);
        std::fill(cmax.begin(), cmax.end(), 0.0);
        for (int r = 0; r < a.rows; ++r) {
            const double f = rs[r];
            double best = 0.0;
            for (int64_t k = a.rp[r]; k < a.rp[r + 1]; ++k) {
                const int c = a.ci[k];
                const double mag = std::abs(a.rv[k]) * f * cs[c];
                best = std::max(best, mag);
                cmax[c] = std::max(cmax[c], mag);
            }
            rmax[r] = best;
        }
        for (int r = 0; r < a.rows; ++r)
            if (rmax[r] > 1e-30) rs[r] /= std::sqrt(rmax[r]);
        for (int c = 0; c < a.cols; ++c)
            if (cmax[c] > 1e-30) cs[c] /= std::sqrt(cmax[c]);
    }
}

inline double scaleBound(double b, double f) { return std::isfinite(b) ? b * f : b; }

Scaling equilibrate(const Lp& lp, int passes) {
    Scaling s;
    ruizScales(lp, passes, s.rs, s.cs);
    s.lp = lp;
    Csx& a = s.lp.a;
    for (int r = 0; r < a.rows; ++r)
        for (int64_t k = a.rp[r]; k < a.rp[r + 1]; ++k) a.rv[k] *= s.rs[r] * s.cs[a.ci[k]];
    for (int c = 0; c < a.cols; ++c)
        for (int64_t k = a.cp[c]; k < a.cp[c + 1]; ++k) a.cv[k] *= s.cs[c] * s.rs[a.ri[k]];
    for (int j = 0; j < lp.n(); ++j) {
        const double f = s.cs[j];
        s.lp.c[j] = lp.c[j] * f;
        s.lp.lo[j] = scaleBound(lp.lo[j], 1.0 / f);
        s.lp.hi[j] = scaleBound(lp.hi[j], 1.0 / f);
    }
    for (int i = 0; i < lp.m(); ++i) {
        s.lp.rlo[i] = scaleBound(lp.rlo[i], s.rs[i]);
        s.lp.rhi[i] = scaleBound(lp.rhi[i], s.rs[i]);
    }
    return s;
}

// ---------------------------------------------------------------------------
// Pock-Chambolle diagonal steps (preconditioner.cpp:31-80) and the power
// iteration used without them (preconditioner.cpp:82-140).
// ---------------------------------------------------------------------------
struct Diag {
    Vec t, tInv, s, sInv;   // primal T_j, dual Sigma_i and reciprocals
};

Diag pockChambolle(const Lp& lp, bool enabled) {
    Diag d;
    d.t.assign(static_cast<size_t>(lp.n()), 1.0);
    d.tInv = d.t;
    d.s.assign(static_cast<size_t>(lp.m()), 1.0);
    d.sInv = d.s;
    if (!enabled) return d;
    const Csx& a = lp.a;
    Vec colSum(static_cast<size_t>(lp.n()), 0.0), rowSum(static_cast<size_t>(lp.m()), 0.0);
    for (int r = 0; r < a.rows; ++r) {
        double acc = 0.0;
        for (int64_t k = a.rp[r]; k < a.rp[r + 1]; ++k) {
            const double mag = std::abs(a.rv[k]);
            acc += mag;
            colSum[a.ci[k]] += mag;
        }
        rowSum[r] = acc;
    }
    for (int j = 0; j < lp.n(); ++j) {
        const double v = colSum[j] > 1e-12 ? 1.0 / colSum[j] : 1.0;
        d.t[j] = v;
        d.tInv[j] = 1.0 / v;
    }
    for (int i = 0; i < lp.m(); ++i) {
        const double v = rowSum[i] > 1e-12 ? 1.0 / rowSum[i] : 1.0;
        d.s[i] = v;
        d.sInv[i] = 1.0 / v;
    }
    return d;
}

double unitize(Vec& v) {
    double sq = 0.0;
    for (double e : v) sq += e * e;
    const double norm = std::sqrt(sq);
    if (norm > 0.0) {
        const double inv = 1.0 / norm;
        for (double& e : v) e *= inv;
    }
    return norm;
}

double spectralNorm(const Csx& a, int budgetIn, bool* converged) {
    *converged = true;
    if (a.rows == 0 || a.cols == 0 || a.rv.empty()) return 0.0;
    std::mt19937 gen(1);
    std::uniform_real_distribution<double> dist(-1.0, 1.0);
    Vec x(static_cast<size_t>(a.cols));
    for (double& e : x) e = dist(gen);
    if (unitize(x) == 0.0) x.assign(x.size(), 1.0 / std::sqrt(static_cast<double>(x.size())));
    Vec y(static_cast<size_t>(a.rows)), z(static_cast<size_t>(a.cols));
    double prev = 0.0, est = 0.0;
    bool hit = false;
    const int budget = std::max(budgetIn, 1);
    for (int it = 0; it < budget; ++it) {
        csrTimes(a, x.data(), y.data());
        cscTimes(a, y.data(), z.data());
        double ray = 0.0;
        for (size_t j = 0; j < x.size(); ++j) ray += x[j] * z[j];
        est = std::sqrt(std::max(ray, 0.0));
        if (unitize(z) == 0.0) return est;
        x.swap(z);
        if (it > 0 && est > 0.0 && std::abs(est - prev) <= 1e-8 * est) {
            hit = true;
            break;
        }
        prev = est;
    }
    *converged = hit;
    return est;
}

// ---------------------------------------------------------------------------
// Step size / primal weight policy (step_controller.cpp:17-98).
// ---------------------------------------------------------------------------
struct Steps {
    double eta = 1.0, omega = 1.0, maxSafe = 1.0;
    double minScale = 1e-6, wMin = 1e-6, wMax = 1e6, smoothing = 0.5;

    void init(const pdlpx_options& o, double safe) {
        maxSafe = safe;
        minScale = o.minimum_step_scale;
        wMin = o.minimum_primal_weight;
        wMax = o.maximum_primal_weight;
        smoothing = o.primal_weight_smoothing;
        eta = std::clamp(o.initial_step_safety * safe, o.minimum_step_scale * safe, safe);
        omega = 1.0;
    }
    void setWeight(double w) {
        if (w > 0.0 && std::isfinite(w)) omega = std::clamp(w, wMin, wMax);
    }
    // Returns whether the trial is accepted; always updates eta.
    bool judge(double pm, double dm, double inter, int64_t k) {
        const double mov = omega * pm + dm / omega;
        const double cpl = std::abs(inter);
        double limit = kInf;
        if (cpl > 0.0 && mov > 0.0) limit = mov / (2.0 * cpl);
        const double idx = static_cast<double>(k) + 2.0;
        const double shrink = 1.0 - std::pow(idx, -0.3);
        const double grow = 1.0 + std::pow(idx, -0.6);
        const bool ok = eta <= limit;
        double next = std::min(shrink * limit, grow * eta);
        next = std::clamp(next, minScale * maxSafe, 1e12);
        if (!std::isfinite(next) || next <= 0.0) next = eta;
        eta = next;
        return ok;
    }
    void reweight(double dx, double dy) {
        if (dx <= 1e-16 || dy <= 1e-16 || !std::isfinite(dx) || !std::isfinite(dy)) return;
        const double target = std::clamp(dy / dx, 1e-12, 1e12);
        const double s = std::clamp(smoothing, 0.0, 1.0);
        setWeight(std::exp(s * std::log(target) + (1.0 - s) * std::log(omega)));
    }
};

// ---------------------------------------------------------------------------
// One PDHG trial (pdhg_kernel.cpp:47-187) and commit (:279-284).
// ---------------------------------------------------------------------------
struct Iterate {
    Vec x, y, ax;          // x^k, y^k, A x^k (working coordinates)
    Vec xt, yt, axt;       // trial buffers
    int64_t k = 0;
};

struct TrialSums {
    double pm = 0.0, dm = 0.0, inter = 0.0;
    bool finite = true;
};

TrialSums trial(const Lp& w, const Diag& d, Iterate& s, double eta, double omega) {
    TrialSums out;
    const Csx& a = w.a;
    const double pmul = eta / omega;
    double pm = 0.0;
    for (int j = 0; j < a.cols; ++j) {
        const double g = w.c[j] + dot4(a.cp[j], a.cp[j + 1], a.ri.data(), a.cv.data(), s.y.data());
        const double was = s.x[j];
        const double now = clip(was - pmul * d.t[j] * g, w.lo[j], w.hi[j]);
        const double dx = now - was;
        s.xt[j] = now;
        pm += dx * dx * d.tInv[j];
    }
    const double dmul = eta * omega;
    double dm = 0.0, inter = 0.0;
    for (int i = 0; i < a.rows; ++i) {
        const double act = dot4(a.rp[i], a.rp[i + 1], a.ci.data(), a.rv.data(), s.xt.data());
        const double dact = act - s.ax[i];
        const double ext = act + dact;
        const double st = std::max(dmul * d.s[i], 1e-300);
        const double was = s.y[i];
        const double v = was + st * ext;
        const double now = std::max(v - st * w.rhi[i], 0.0) + std::min(v - st * w.rlo[i], 0.0);
        const double dy = now - was;
        s.yt[i] = now;
        s.axt[i] = act;
        dm += dy * dy * d.sInv[i];
        inter += dy * dact;
    }
    out.pm = pm;
    out.dm = dm;
    out.inter = inter;
    out.finite = std::isfinite(pm) && std::isfinite(dm) && std::isfinite(inter);
    return out;
}

void commit(Iterate& s) {
    s.x.swap(s.xt);
    s.y.swap(s.yt);
    s.ax.swap(s.axt);
    ++s.k;
}

// ---------------------------------------------------------------------------
// Weighted running average (iterate_average.cpp:14-95).
// ---------------------------------------------------------------------------
struct Average {
    Vec x, y;
    double total = 0.0;
    void add(const Vec& px, const Vec& py, double w) {
        if (!(w > 0.0)) return;
        const double nt = total + w;
        const double f = w / nt;
        for (size_t j = 0; j < x.size(); ++j) x[j] += f * (px[j] - x[j]);
        for (size_t i = 0; i < y.size(); ++i) y[i] += f * (py[i] - y[i]);
        total = nt;
    }
    bool empty() const { return total == 0.0; }
    void clear() {
        std::fill(x.begin(), x.end(), 0.0);
        std::fill(y.begin(), y.end(), 0.0);
        total = 0.0;
    }
};

// ---------------------------------------------------------------------------
// KKT scoring on the original problem (termination.cpp:15-333).
// ---------------------------------------------------------------------------
struct Metrics {
    double pobj = std::numeric_limits<double>::quiet_NaN();
    double dobj = -kInf;
    double pres = kInf, dres = kInf, gap = kInf, kkt = kInf;
    bool dualValid = false, finite = false;
};

struct Scorer {
    const Lp* lp = nullptr;
    double pnorm = 1.0, dnorm = 1.0;
    double tolP = 0, tolD = 0, tolG = 0;

    void bind(const Lp& problem, const pdlpx_options& o) {
        lp = &problem;
        double sq = 0.0;
        for (int i = 0; i < problem.m(); ++i) {
            if (std::isfinite(problem.rlo[i])) sq += problem.rlo[i] * problem.rlo[i];
            if (std::isfinite(problem.rhi[i]) && problem.rhi[i] != problem.rlo[i])
                sq += problem.rhi[i] * problem.rhi[i];
        }
        pnorm = 1.0 + std::sqrt(sq);
        double cq = 0.0;
        for (double v : problem.c) cq += v * v;
        dnorm = 1.0 + std::sqrt(cq);
        tolP = o.primal_tolerance;
        tolD = o.dual_tolerance;
        tolG = o.gap_tolerance;
    }

    Metrics score(const Vec& x, const Vec& y) const {
        const Lp& p = *lp;
        const Csx& a = p.a;
        double viol = 0.0, support = 0.0, rowDrift = 0.0;
        for (int i = 0; i < a.rows; ++i) {
            const double act = dot4(a.rp[i], a.rp[i + 1], a.ci.data(), a.rv.data(), x.data());
            const double l = p.rlo[i], u = p.rhi[i];
            const double e = act - clip(act, l, u);
            viol += e * e;
            double q = y[i];
            if (!std::isfinite(u)) q = std::min(q, 0.0);
            if (!std::isfinite(l)) q = std::max(q, 0.0);
            const double dr = y[i] - q;
            rowDrift += dr * dr;
            if (q > 0.0) support += q * u;
            else if (q < 0.0) support += q * l;
        }
        double obj = 0.0, inf = 0.0, colDrift = 0.0;
        for (int j = 0; j < a.cols; ++j) {
            const double rc = p.c[j] + dot4(a.cp[j], a.cp[j + 1], a.ri.data(), a.cv.data(), y.data());
            const double l = p.lo[j], u = p.hi[j];
            obj += p.c[j] * x[j];
            double q = rc;
            if (!std::isfinite(l)) q = std::min(q, 0.0);
            if (!std::isfinite(u)) q = std::max(q, 0.0);
            const double dr = rc - q;
            colDrift += dr * dr;
            if (q > 0.0) inf += q * l;
            else if (q < 0.0) inf += q * u;
        }
        return finish(viol, support, rowDrift, obj, inf, colDrift);
    }

    Metrics finish(double viol, double support, double rowDrift, double obj, double inf,
                   double colDrift) const {
        Metrics m;
        m.pres = std::sqrt(viol) / pnorm;
        m.dres = std::sqrt(colDrift + rowDrift) / dnorm;
        m.pobj = lp->offset + obj;
        m.dobj = lp->offset + inf - support;
        m.finite = std::isfinite(m.pobj) && std::isfinite(m.dobj) && std::isfinite(m.pres) &&
                   std::isfinite(m.dres);
        if (!m.finite) {
            m.dualValid = false;
            m.kkt = kInf;
            return m;
        }
        m.dualValid = true;
        m.gap = std::abs(m.pobj - m.dobj) / (1.0 + std::abs(m.pobj) + std::abs(m.dobj));
        m.kkt = std::max(std::max(m.pres, m.dres), m.gap);
        return m;
    }

    bool optimal(const Metrics& m) const {
        return m.finite && m.dualValid && m.pres <= tolP && m.dres <= tolD && m.gap <= tolG;
    }
};

// ---------------------------------------------------------------------------
// Restart policy (restart_controller.cpp:8-63).
// ---------------------------------------------------------------------------
enum class Pick { None, Current, Average };

struct Restarter {
    pdlpx_options o{};
    double prevScore = kInf;
    Pick choose(const Metrics& cur, const Metrics& avg, double baseline, int64_t since,
                int64_t total, double* scoreOut) {
        *scoreOut = kInf;
        if (!o.use_restarts || since < o.minimum_restart_iterations) return Pick::None;
        const bool af = std::isfinite(avg.kkt), cf = std::isfinite(cur.kkt);
        if (!af && !cf) return Pick::None;
        Pick cand = Pick::Average;
        double sc = avg.kkt;
        if (!af || (cf && cur.kkt < avg.kkt)) {
            cand = Pick::Current;
            sc = cur.kkt;
        }
        *scoreOut = sc;
        const double prev = prevScore;
        prevScore = sc;
        const bool bf = std::isfinite(baseline);
        if (bf && sc <= o.restart_sufficient_decay * baseline) return cand;
        if (bf && sc <= o.restart_necessary_decay * baseline && std::isfinite(prev) && sc > prev)
            return cand;
        const auto art = static_cast<int64_t>(o.restart_artificial_fraction * static_cast<double>(total));
        if (since >= std::max<int64_t>(art, 1) || since >= o.maximum_restart_iterations) return cand;
        return Pick::None;
    }
};

// ---------------------------------------------------------------------------
// Ray certificates (infeasibility.cpp:51-192).
// ---------------------------------------------------------------------------
bool unitInto(const Vec& src, Vec& dst) {
    double sq = 0.0;
    for (double v : src) sq += v * v;
    const double norm = std::sqrt(sq);
    if (!(norm > 1e-300) || !std::isfinite(norm)) return false;
    dst.resize(src.size());
    const double inv = 1.0 / norm;
    for (size_t i = 0; i < src.size(); ++i) dst[i] = src[i] * inv;
    return true;
}

bool farkas(const Lp& p, const Vec& dir, double tol) {
    if (dir.size() != size_t(p.m())) return false;
    Vec u;
    if (!unitInto(dir, u)) return false;
    Vec aty(static_cast<size_t>(p.n()));
    cscTimes(p.a, u.data(), aty.data());
    double boxInf = 0.0, viol = 0.0;
    for (int j = 0; j < p.n(); ++j) {
        const double rc = aty[j];
        double q = rc;
        if (!std::isfinite(p.lo[j])) q = std::min(q, 0.0);
        if (!std::isfinite(p.hi[j])) q = std::max(q, 0.0);
        const double dr = rc - q;
        viol += dr * dr;
        if (q > 0.0) boxInf += q * p.lo[j];
        else if (q < 0.0) boxInf += q * p.hi[j];
    }
    double support = 0.0;
    for (int i = 0; i < p.m(); ++i) {
        const double v = u[i];
        double q = v;
        if (!std::isfinite(p.rhi[i])) q = std::min(q, 0.0);
        if (!std::isfinite(p.rlo[i])) q = std::max(q, 0.0);
        const double dr = v - q;
        viol += dr * dr;
        if (q > 0.0) support += q * p.rhi[i];
        else if (q < 0.0) support += q * p.rlo[i];
    }
    const double value = boxInf - support;
    if (!std::isfinite(value)) return false;
    const double v = std::sqrt(viol);
    return value > 0.0 && v / value <= tol;
}

bool improvingRay(const Lp& p, const Vec& dir, double tol) {
    if (dir.size() != size_t(p.n())) return false;
    Vec u;
    if (!unitInto(dir, u)) return false;
    Vec proj(static_cast<size_t>(p.n())), act(static_cast<size_t>(p.m()));
    double viol = 0.0, rate = 0.0;
    for (int j = 0; j < p.n(); ++j) {
        const double v = u[j];
        double q = v;
        if (std::isfinite(p.hi[j])) q = std::min(q, 0.0);
        if (std::isfinite(p.lo[j])) q = std::max(q, 0.0);
        const double dr = v - q;
        viol += dr * dr;
        proj[j] = q;
        rate += p.c[j] * q;
    }
    csrTimes(p.a, proj.data(), act.data());
    for (int i = 0; i < p.m(); ++i) {
        const double v = act[i];
        double q = v;
        if (std::isfinite(p.rhi[i])) q = std::min(q, 0.0);
        if (std::isfinite(p.rlo[i])) q = std::max(q, 0.0);
        const double dr = v - q;
        viol += dr * dr;
    }
    if (!std::isfinite(rate)) return false;
    const double v = std::sqrt(viol);
    return rate < 0.0 && v / (-rate) <= tol;
}

double sqDist(const Vec& a, const Vec& b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) {
        const double d = a[i] - b[i];
        s += d * d;
    }
    return s;
}

double startValue(double lo, double hi) {   // pdlp_solver.cpp:30-41
    if (lo <= 0.0 && 0.0 <= hi) return 0.0;
    if (std::isfinite(lo)) return lo;
    if (std::isfinite(hi)) return hi;
    return 0.0;
}

using Clock = std::chrono::steady_clock;
double since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

// ---------------------------------------------------------------------------
// The solve driver (pdlp_solver.cpp:116-548) and polishing
// (feasibility_polishing.cpp:13-145).
// ---------------------------------------------------------------------------
struct Outcome {
    int status = PDLPX_INVALID_PROBLEM;
    std::string message;
    Vec x, y, dualRay, primalRay;
    Metrics metrics;
    int64_t iterations = 0, trials = 0;
    int restarts = 0;
    bool diag = false;
    double eta = 0, safe = 0, omega = 1;
    int checks = 0;
};

struct TraceSink {
    double* buf = nullptr;
    int cap = 0;
    int len = 0;
    void put(const double* row) {
        if (len < cap && buf) std::memcpy(buf + size_t(len) * ORACLE_TRACE_WIDTH, row,
                                          sizeof(double) * ORACLE_TRACE_WIDTH);
        ++len;
    }
};

Outcome runSolve(const Lp& orig, const pdlpx_options& o, const double* warmX, const double* warmY,
                 TraceSink* trace, Clock::time_point t0) {
    Outcome R;
    const int n = orig.n(), m = orig.m();

    Scaling sc;
    const bool scaled = o.use_ruiz_scaling && o.ruiz_iterations > 0 && !orig.a.rv.empty();
    if (scaled) sc = equilibrate(orig, o.ruiz_iterations);
    const Lp& w = scaled ? sc.lp : orig;
    auto toOrig = [&](const Vec& xs, const Vec& ys, Vec& xo, Vec& yo) {
        if (!scaled) {
            xo = xs;
            yo = ys;
            return;
        }
        xo.resize(xs.size());
        for (size_t j = 0; j < xs.size(); ++j) xo[j] = sc.cs[j] * xs[j];
        yo.resize(ys.size());
        for (size_t i = 0; i < ys.size(); ++i) yo[i] = sc.rs[i] * ys[i];
    };

    Iterate s;
    s.x.resize(static_cast<size_t>(n));
    s.y.assign(static_cast<size_t>(m), 0.0);
    for (int j = 0; j < n; ++j) s.x[j] = startValue(w.lo[j], w.hi[j]);
    // Warm start (engine extension): map to working coordinates, project the
    // primal into the box and the dual onto the cone the row bounds allow.
    if (warmX) {
        for (int j = 0; j < n; ++j) {
            const double v = scaled ? warmX[j] / sc.cs[j] : warmX[j];
            s.x[j] = clip(v, w.lo[j], w.hi[j]);
        }
    }
    if (warmY) {
        for (int i = 0; i < m; ++i) {
            double v = scaled ? warmY[i] / sc.rs[i] : warmY[i];
            if (!std::isfinite(w.rhi[i])) v = std::min(v, 0.0);
            if (!std::isfinite(w.rlo[i])) v = std::max(v, 0.0);
            s.y[i] = v;
        }
    }
    s.xt.assign(static_cast<size_t>(n), 0.0);
    s.yt.assign(static_cast<size_t>(m), 0.0);
    s.axt.assign(static_cast<size_t>(m), 0.0);

    const Diag d = pockChambolle(w, o.use_diagonal_preconditioning);
    double safe = 0.99;
    if (!o.use_diagonal_preconditioning) {
        bool conv = false;
        const double norm = spectralNorm(w.a, o.power_iteration_count, &conv);
        safe = norm > 0.0 ? 0.99 / (norm * (conv ? 1.02 : 1.25)) : 1.0;
    }
    Steps st;
    st.init(o, safe);
    st.setWeight(o.initial_primal_weight);
    s.ax.resize(static_cast<size_t>(m));
    csrTimes(w.a, s.x.data(), s.ax.data());

    Average av;
    av.x.assign(static_cast<size_t>(n), 0.0);
    av.y.assign(static_cast<size_t>(m), 0.0);
    Scorer chk;
    chk.bind(orig, o);
    Restarter rc;
    rc.o = o;

    Vec ox, oy;
    auto score = [&](const Vec& xs, const Vec& ys) {
        toOrig(xs, ys, ox, oy);
        return chk.score(ox, oy);
    };

    Metrics cur = score(s.x, s.y);
    Vec bestX = ox, bestY = oy;
    Metrics best = cur;
    double baseline = cur.kkt;
    int64_t sinceRestart = 0;
    int restarts = 0;
    Vec anchorX = s.x, anchorY = s.y;
    int64_t trials = 0;

    auto out = [&](int status, const char* msg, int64_t its, int64_t tr, bool diag) {
        R.status = status;
        R.message = msg;
        R.x = bestX;
        R.y = bestY;
        R.metrics = best;
        R.iterations = its;
        R.trials = tr;
        R.restarts = restarts;
        R.diag = diag;
        R.eta = st.eta;
        R.safe = st.maxSafe;
        R.omega = st.omega;
        return R;
    };

    if (chk.optimal(cur)) return out(PDLPX_OPTIMAL, "Initial point satisfies the requested tolerances", 0, 0, false);

    int limitStatus = PDLPX_ITERATION_LIMIT;
    const char* limitMsg = "Iteration limit reached";
    const int maxTrials = std::max(o.maximum_step_trials, 1);

    for (int64_t it = 0; it < o.iteration_limit; ++it) {
        if (o.time_limit_seconds > 0.0 && since(t0) >= o.time_limit_seconds) {
            limitStatus = PDLPX_TIME_LIMIT;
            limitMsg = "Time limit reached";
            break;
        }
        bool took = false;
        for (int a = 0; a < maxTrials; ++a) {
            const TrialSums t = trial(w, d, s, st.eta, st.omega);
            ++trials;
            if (!t.finite) return out(PDLPX_NUMERICAL_FAILURE, "PDHG generated a non-finite iterate", s.k, trials, true);
            const bool ok = !o.use_adaptive_linesearch || st.judge(t.pm, t.dm, t.inter, s.k);
            if (ok) {
                commit(s);
                took = true;
                break;
            }
        }
        if (!took) return out(PDLPX_NUMERICAL_FAILURE, "Adaptive linesearch failed to accept a step", s.k, trials, true);
        ++sinceRestart;
        if (o.use_averaging) av.add(s.x, s.y, st.eta);
        if (s.k % o.termination_check_frequency != 0) continue;

        ++R.checks;
        cur = score(s.x, s.y);
        if (cur.kkt < best.kkt) {
            bestX = ox;
            bestY = oy;
            best = cur;
        }
        Metrics avgM = cur;
        const bool haveAvg = o.use_averaging && !av.empty();
        if (haveAvg) {
            avgM = score(av.x, av.y);
            if (avgM.kkt < best.kkt) {
                bestX = ox;
                bestY = oy;
                best = avgM;
            }
        }
        if (trace) {
            double row[ORACLE_TRACE_WIDTH] = {static_cast<double>(s.k), static_cast<double>(trials),
                                              st.eta, st.omega, cur.kkt, avgM.kkt, best.kkt,
                                              static_cast<double>(restarts)};
            trace->put(row);
        }
        if (o.detect_infeasibility) {
            Vec dx(static_cast<size_t>(n)), dy(static_cast<size_t>(m));
            for (int j = 0; j < n; ++j) dx[j] = s.x[j] - anchorX[j];
            for (int i = 0; i < m; ++i) dy[i] = s.y[i] - anchorY[i];
            Vec colRay, rowRay;
            toOrig(dx, dy, colRay, rowRay);
            if (farkas(orig, rowRay, o.infeasibility_tolerance)) {
                out(PDLPX_INFEASIBLE, "Primal infeasible: Farkas ray found", s.k, trials, true);
                R.dualRay = rowRay;
                return R;
            }
            if (improvingRay(orig, colRay, o.infeasibility_tolerance)) {
                out(PDLPX_UNBOUNDED, "Unbounded: improving ray found", s.k, trials, true);
                R.primalRay = colRay;
                return R;
            }
        }
        if (chk.optimal(best)) return out(PDLPX_OPTIMAL, "Primal, dual, and gap tolerances satisfied", s.k, trials, true);

        double candScore = kInf;
        const Pick pick = rc.choose(cur, avgM, baseline, sinceRestart, s.k, &candScore);
        if (pick != Pick::None) {
            const bool fromAvg = pick == Pick::Average && haveAvg;
            const Vec& px = fromAvg ? av.x : s.x;
            const Vec& py = fromAvg ? av.y : s.y;
            st.reweight(std::sqrt(sqDist(px, anchorX)), std::sqrt(sqDist(py, anchorY)));
            if (fromAvg) {
                s.x = av.x;
                s.y = av.y;
                csrTimes(w.a, s.x.data(), s.ax.data());
            }
            anchorX = s.x;
            anchorY = s.y;
            baseline = candScore;
            av.clear();
            rc.prevScore = kInf;
            sinceRestart = 0;
            ++restarts;
        }
    }

    if (o.use_feasibility_polishing && !bestX.empty()) {
        double remaining = 0.0;
        if (o.time_limit_seconds > 0.0) {
            remaining = o.time_limit_seconds - since(t0);
            if (remaining <= 0.0) return out(limitStatus, limitMsg, s.k, trials, true);
        }
        // feasibility_polishing.cpp:26-145
        Metrics pBest = best;
        Vec pX = bestX, pY = bestY;
        int pIts = 0;
        int64_t pTrials = 0;
        if (o.polishing_iterations > 0) {
            double trialEta = st.eta * 0.5;
            const double trialOmega = st.omega;
            Iterate q;
            q.x.resize(static_cast<size_t>(n));
            q.y.resize(static_cast<size_t>(m));
            for (int j = 0; j < n; ++j) q.x[j] = scaled ? bestX[j] / sc.cs[j] : bestX[j];
            for (int i = 0; i < m; ++i) q.y[i] = scaled ? bestY[i] / sc.rs[i] : bestY[i];
            q.xt.assign(static_cast<size_t>(n), 0.0);
            q.yt.assign(static_cast<size_t>(m), 0.0);
            q.axt.assign(static_cast<size_t>(m), 0.0);
            q.ax.resize(static_cast<size_t>(m));
            csrTimes(w.a, q.x.data(), q.ax.data());
            Steps ps;
            ps.init(o, st.maxSafe);
            ps.setWeight(trialOmega);
            Average pav;
            pav.x.assign(static_cast<size_t>(n), 0.0);
            pav.y.assign(static_cast<size_t>(m), 0.0);
            const auto pt0 = Clock::now();
            const int freq = std::max(o.termination_check_frequency, 1);
            for (int it = 0; it < o.polishing_iterations; ++it) {
                bool took = false, broken = false;
                for (int a = 0; a < maxTrials; ++a) {
                    const TrialSums t = trial(w, d, q, trialEta, trialOmega);
                    ++pTrials;
                    if (!t.finite) {
                        broken = true;
                        break;
                    }
                    if (!o.use_adaptive_linesearch) {
                        commit(q);
                        took = true;
                        break;
                    }
                    const bool ok = ps.judge(t.pm, t.dm, t.inter, q.k);
                    trialEta = ps.eta;
                    if (ok) {
                        commit(q);
                        took = true;
                        break;
                    }
                }
                if (broken || !took) break;
                if (o.use_averaging) pav.add(q.x, q.y, trialEta);
                ++pIts;
                const bool last = it + 1 == o.polishing_iterations;
                if ((it + 1) % freq != 0 && !last) continue;
                if (remaining > 0.0 && since(pt0) >= remaining) break;
                const Metrics qc = score(q.x, q.y);
                if (qc.kkt < pBest.kkt) {
                    pX = ox;
                    pY = oy;
                    pBest = qc;
                }
                if (o.use_averaging && !pav.empty()) {
                    const Metrics qa = score(pav.x, pav.y);
                    if (qa.kkt < pBest.kkt) {
                        pX = ox;
                        pY = oy;
                        pBest = qa;
                    }
                }
                if (chk.optimal(pBest)) break;
            }
        }
        if (pBest.kkt < best.kkt) {
            bestX = pX;
            bestY = pY;
            best = pBest;
        }
        if (chk.optimal(best))
            return out(PDLPX_OPTIMAL, "Tolerances satisfied after refinement polishing", s.k + pIts,
                       trials + pTrials, true);
    }
    return out(limitStatus, limitMsg, s.k, trials, true);
}

bool optionsValid(const pdlpx_options& o) {   // pdlp_solver.cpp:127-133
    return !(o.iteration_limit < 0 || o.termination_check_frequency <= 0 || o.primal_tolerance < 0.0 ||
             o.dual_tolerance < 0.0 || o.gap_tolerance < 0.0 || o.initial_step_safety <= 0.0 ||
             o.minimum_primal_weight <= 0.0 || o.maximum_primal_weight < o.minimum_primal_weight);
}

void fillSummary(const Outcome& r, double seconds, pdlpx_summary* s) {
    std::memset(s, 0, sizeof(*s));
    s->status = r.status;
    std::snprintf(s->message, sizeof(s->message), "%s", r.message.c_str());
    s->primal_objective = r.metrics.pobj;
    s->dual_objective = r.metrics.dobj;
    s->primal_residual = r.metrics.pres;
    s->dual_residual = r.metrics.dres;
    s->relative_gap = r.metrics.gap;
    s->iterations = r.iterations;
    s->step_trials = r.trials;
    s->restart_count = r.restarts;
    s->solve_time_seconds = seconds;
    s->final_step_size = r.diag ? r.eta : 0.0;
    s->static_step_bound = r.diag ? r.safe : 0.0;
    s->final_primal_weight = r.diag ? r.omega : 1.0;
    s->has_dual_ray = r.dualRay.empty() ? 0 : 1;
    s->has_primal_ray = r.primalRay.empty() ? 0 : 1;
    s->termination_checks = r.checks;
}

}  // namespace

extern "C" {

int oracle_solve_traced(const pdlpx_problem* problem, const pdlpx_options* options,
                        const double* warm_x, const double* warm_y, pdlpx_summary* summary,
                        double* x_out, double* y_out, double* ray_out, double* trace,
                        int32_t trace_cap, int32_t* trace_len) {
    const auto t0 = Clock::now();
    const Lp lp = lpFromProblem(*problem);
    std::string why = validateLp(lp);
    Outcome r;
    if (!why.empty()) {
        r.message = why;
    } else if (!optionsValid(*options)) {
        r.message = "Invalid PDLP options";
    } else {
        TraceSink sink{trace, trace_cap, 0};
        r = runSolve(lp, *options, warm_x, warm_y, &sink, t0);
        if (trace_len) *trace_len = sink.len;
    }
    fillSummary(r, since(t0), summary);
    if (x_out && r.x.size() == size_t(problem->cols)) std::memcpy(x_out, r.x.data(), r.x.size() * 8);
    if (y_out && r.y.size() == size_t(problem->rows)) std::memcpy(y_out, r.y.data(), r.y.size() * 8);
    if (ray_out && !r.dualRay.empty()) std::memcpy(ray_out, r.dualRay.data(), r.dualRay.size() * 8);
    if (ray_out && !r.primalRay.empty()) std::memcpy(ray_out, r.primalRay.data(), r.primalRay.size() * 8);
    return 0;
}

int oracle_solve(const pdlpx_problem* problem, const pdlpx_options* options, const double* warm_x,
                 const double* warm_y, pdlpx_summary* summary, double* x_out, double* y_out,
                 double* ray_out) {
    return oracle_solve_traced(problem, options, warm_x, warm_y, summary, x_out, y_out, ray_out,
                               nullptr, 0, nullptr);
}

void* oracle_matrix_from_triplets(int32_t rows, int32_t cols, int64_t count, const int32_t* tr,
                                  const int32_t* tc, const double* tv) {
    return new Csx(buildFromTriplets(rows, cols, count, tr, tc, tv));
}

int64_t oracle_matrix_nnz(const void* handle) {
    return static_cast<int64_t>(static_cast<const Csx*>(handle)->rv.size());
}

void oracle_matrix_export(const void* handle, int64_t* rp, int32_t* ci, double* rv, int64_t* cp,
                          int32_t* ri, double* cv) {
    const Csx& m = *static_cast<const Csx*>(handle);
    std::copy(m.rp.begin(), m.rp.end(), rp);
    std::copy(m.ci.begin(), m.ci.end(), ci);
    std::copy(m.rv.begin(), m.rv.end(), rv);
    std::copy(m.cp.begin(), m.cp.end(), cp);
    std::copy(m.ri.begin(), m.ri.end(), ri);
    std::copy(m.cv.begin(), m.cv.end(), cv);
}

void oracle_matrix_free(void* handle) { delete static_cast<Csx*>(handle); }

int oracle_spmv(const pdlpx_problem* problem, int32_t transpose, const double* x, double* y) {
    const Lp lp = lpFromProblem(*problem);
    if (transpose) cscTimes(lp.a, x, y);
    else csrTimes(lp.a, x, y);
    return 0;
}

int oracle_scaling(const pdlpx_problem* problem, int32_t ruiz_iterations, int32_t use_pc,
                   double* row_scale, double* col_scale, double* primal_scale, double* dual_scale) {
    const Lp lp = lpFromProblem(*problem);
    const bool scaled = ruiz_iterations > 0 && !lp.a.rv.empty();
    Scaling sc;
    if (scaled) sc = equilibrate(lp, ruiz_iterations);
    const Lp& w = scaled ? sc.lp : lp;
    const Diag d = pockChambolle(w, use_pc != 0);
    for (int i = 0; i < lp.m(); ++i) {
        if (row_scale) row_scale[i] = scaled ? sc.rs[i] : 1.0;
        if (dual_scale) dual_scale[i] = d.s[i];
    }
    for (int j = 0; j < lp.n(); ++j) {
        if (col_scale) col_scale[j] = scaled ? sc.cs[j] : 1.0;
        if (primal_scale) primal_scale[j] = d.t[j];
    }
    return 0;
}

// metrics: pobj, dobj, pres, dres, gap, kkt, dualValid, finite
int oracle_evaluate(const pdlpx_problem* problem, const pdlpx_options* options, const double* x,
                    const double* y, double* metrics) {
    const Lp lp = lpFromProblem(*problem);
    Scorer s;
    s.bind(lp, *options);
    const Vec xv(x, x + lp.n()), yv(y, y + lp.m());
    const Metrics m = s.score(xv, yv);
    const double outv[8] = {m.pobj, m.dobj, m.pres, m.dres, m.gap, m.kkt, m.dualValid ? 1.0 : 0.0,
                            m.finite ? 1.0 : 0.0};
    std::memcpy(metrics, outv, sizeof(outv));
    return 0;
}

// One evaluateTrial call on a controller in the given state; returns accept.
int oracle_step_judge(const pdlpx_options* options, double max_safe, double eta, double omega,
                      double pm, double dm, double inter, int64_t k, double* eta_out) {
    Steps s;
    s.init(*options, max_safe);
    s.eta = eta;
    s.omega = omega;
    const bool ok = s.judge(pm, dm, inter, k);
    *eta_out = s.eta;
    return ok ? 1 : 0;
}

double oracle_reweight(const pdlpx_options* options, double omega, double dx, double dy) {
    Steps s;
    s.init(*options, 1.0);
    s.omega = omega;
    s.reweight(dx, dy);
    return s.omega;
}

// Restart decision: returns 0 none, 1 current, 2 average.
int oracle_restart_choose(const pdlpx_options* options, double cur_kkt, double avg_kkt,
                          double baseline, double prev_score, int64_t since_restart,
                          int64_t total, double* score_out, double* prev_out) {
    Restarter r;
    r.o = *options;
    r.prevScore = prev_score;
    Metrics c, a;
    c.kkt = cur_kkt;
    a.kkt = avg_kkt;
    const Pick p = r.choose(c, a, baseline, since_restart, total, score_out);
    *prev_out = r.prevScore;
    return p == Pick::None ? 0 : (p == Pick::Current ? 1 : 2);
}

}  // extern "C"
