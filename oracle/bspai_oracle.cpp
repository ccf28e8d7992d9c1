// CPU FP64 oracle for the SPAI-GMRES-IR path.  TEST INFRASTRUCTURE ONLY.
//
// Only the tests, the smoke check and the CPU-baseline leg of the benchmark load
// this library, and only as the checker.  The product (the HIP library under
// spai-gmres-ir_amd/) never links, loads or calls it.
//
// It restates the reference's single-precision-free CPU algorithm operation by
// operation, in the same floating-point order, so that its results are
// bit-identical to the reference's `cpu_fp64_ir`:
//   canonical CSR            src/cpu.cpp:48-103
//   SPAI construction        src/spai.cpp:15-374
//   left-preconditioned GMRES src/gmres.cpp:12-262
//   iterative refinement     src/ir.cpp:13-227
// Two changes make it usable at n = 2,097,152, neither changes a bit:
//   * the shadow I = {i : B(i,j) != 0 for some j in J} is formed as the union of
//     A's rows j in J (B = A^T D, so B(i,j) != 0 <=> a_ji != 0), instead of the
//     O(n) scan per fit of src/spai.cpp:110-117;
//   * the full fitted residual B*m - e_k (src/spai.cpp:147-155) equals the
//     reduced one bit-for-bit: rows outside I are +0 and hypot(s, +0) == s, and
//     inserting +-0 products into a sum that starts at +0 never changes it.
//   * GMRES forms the candidate x0 + V y and recomputes its residual only on
//     iterations whose acceptance test can pass (estimate <= tau, breakdown or
//     the last iteration, src/gmres.cpp:238-241); all other candidates are
//     overwritten unread, so every decision and the returned solution are
//     unchanged.  Histories (basis snapshots, per-iteration solutions) are not
//     retained.
// Parity is pinned by tests/test_oracle.py against the reference's own golden
// outputs (tests/golden/), against closed forms and against brute force.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

// Optional progress report of the long-running calls (the benchmark's CPU
// baseline prints it): phase 0 = SPAI columns done, 1 = refinement steps done
// (value = that step's inner iterations).  Set by oracle_set_progress.
typedef void (*OracleProgress)(int phase, std::int64_t done, std::int64_t value);
static OracleProgress g_progress = nullptr;

namespace {

using idx = std::int64_t;

struct Csr {
    idx n = 0;
    std::vector<idx> ptr, col;
    std::vector<double> val;
};

Csr wrap(idx n, const idx *ptr, const idx *col, const double *val) {
    Csr a;
    a.n = n;
    a.ptr.assign(ptr, ptr + n + 1);
    a.col.assign(col, col + ptr[n]);
    a.val.assign(val, val + ptr[n]);
    return a;
}

// y = A x in CSR order; nonfinite row -> overflow (src/cpu.cpp:80-94).
std::vector<double> multiply(const Csr &a, const std::vector<double> &x) {
    for (double v : x)
        if (!std::isfinite(v))
            throw std::invalid_argument("Vector must be finite");
    std::vector<double> y(a.n, 0.0);
    for (idx r = 0; r < a.n; ++r) {
        for (idx q = a.ptr[r]; q < a.ptr[r + 1]; ++q)
            y[r] += a.val[q] * x[a.col[q]];
        if (!std::isfinite(y[r]))
            throw std::overflow_error("Matrix-vector result is nonfinite");
    }
    return y;
}

double lookup(const Csr &a, idx row, idx c) {
    auto first = a.col.begin() + a.ptr[row], last = a.col.begin() + a.ptr[row + 1];
    auto it = std::lower_bound(first, last, c);
    return (it == last || *it != c) ? 0.0 : a.val[it - a.col.begin()];
}

// ---------------------------------------------------------------- SPAI
enum SpaiStatus {
    SPAI_SUCCESS,
    SPAI_NOT_ATTEMPTED,
    SPAI_INVALID_INPUT,
    SPAI_RANK_DEFICIENT,
    SPAI_NUMERICAL_FAILURE,
    SPAI_EMPTY_CANDIDATES,
    SPAI_ZERO_DENOMINATOR,
    SPAI_INVALID_SCORE,
    SPAI_NO_GROWTH,
    SPAI_ITERATION_LIMIT
};

double hypot_chain(const double *v, idx len) {
    double r = 0;
    for (idx i = 0; i < len; ++i)
        r = std::hypot(r, v[i]);
    return r;
}

// Householder least squares (src/spai.cpp:30-98); r is m x q row-major.
int least_squares(std::vector<double> r, std::vector<double> rhs, idx m, idx q,
                  std::vector<double> &sol, double &pivot_ratio) {
    if (m < q)
        return SPAI_RANK_DEFICIENT;
    double inf_norm = 0;
    for (idx i = 0; i < m; ++i) {
        double s = 0;
        for (idx j = 0; j < q; ++j)
            s += std::abs(r[i * q + j]);
        inf_norm = std::max(inf_norm, s);
    }
    double min_pivot = std::numeric_limits<double>::infinity();
    std::vector<double> v(m);
    for (idx j = 0; j < q; ++j) {
        const idx len = m - j;
        for (idx i = j; i < m; ++i)
            v[i - j] = r[i * q + j];
        const double length = hypot_chain(v.data(), len);
        if (!std::isfinite(length))
            return SPAI_NUMERICAL_FAILURE;
        if (length == 0)
            return SPAI_RANK_DEFICIENT;
        const double diagonal = -std::copysign(length, v[0]);
        for (idx i = 0; i < len; ++i)
            v[i] /= length;
        v[0] += std::copysign(1.0, v[0]);
        const double vnorm = hypot_chain(v.data(), len);
        for (idx i = 0; i < len; ++i)
            v[i] /= vnorm;
        for (idx c = j; c < q; ++c) {
            double dot = 0;
            for (idx i = j; i < m; ++i)
                dot += v[i - j] * r[i * q + c];
            for (idx i = j; i < m; ++i) {
                r[i * q + c] -= 2 * v[i - j] * dot;
                if (!std::isfinite(r[i * q + c]))
                    return SPAI_NUMERICAL_FAILURE;
            }
        }
        double dot = 0;
        for (idx i = j; i < m; ++i)
            dot += v[i - j] * rhs[i];
        for (idx i = j; i < m; ++i) {
            rhs[i] -= 2 * v[i - j] * dot;
            if (!std::isfinite(rhs[i]))
                return SPAI_NUMERICAL_FAILURE;
        }
        r[j * q + j] = diagonal;
        for (idx i = j + 1; i < m; ++i)
            r[i * q + j] = 0;
        min_pivot = std::min(min_pivot, std::abs(diagonal));
    }
    pivot_ratio = inf_norm == 0 ? 0 : min_pivot / inf_norm;
    sol.assign(q, 0);
    for (idx j = q; j-- > 0;) {
        double value = rhs[j];
        for (idx k = j + 1; k < q; ++k)
            value -= r[j * q + k] * sol[k];
        const double d = r[j * q + j];
        if (!std::isfinite(d))
            return SPAI_NUMERICAL_FAILURE;
        if (d == 0)
            return SPAI_RANK_DEFICIENT;
        sol[j] = value / d;
        if (!std::isfinite(sol[j]))
            return SPAI_NUMERICAL_FAILURE;
    }
    return SPAI_SUCCESS;
}

struct Score {
    int status;
    double score;
    bool clamped;
};

// src/spai.cpp:273-306
Score score_of(double rr, double dot, double den) {
    Score s{SPAI_NOT_ATTEMPTED, 0, false};
    if (!std::isfinite(rr) || !std::isfinite(dot) || !std::isfinite(den)) {
        s.status = SPAI_NUMERICAL_FAILURE;
        return s;
    }
    if (den == 0) {
        s.status = SPAI_ZERO_DENOMINATOR;
        return s;
    }
    if (rr < 0 || den < 0) {
        s.status = SPAI_INVALID_SCORE;
        return s;
    }
    const double quotient = dot * dot / den;
    double delta = rr - quotient;
    const double bound = 32 * std::ldexp(1.0, -53) * (rr + std::abs(quotient));
    if (!std::isfinite(quotient) || !std::isfinite(delta) || !std::isfinite(bound)) {
        s.status = SPAI_NUMERICAL_FAILURE;
        return s;
    }
    if (delta < 0) {
        if (-delta > bound) {
            s.status = SPAI_INVALID_SCORE;
            return s;
        }
        delta = 0;
        s.clamped = true;
    }
    s.score = std::sqrt(delta);
    s.status = SPAI_SUCCESS;
    return s;
}

struct Column {
    int status = SPAI_NOT_ATTEMPTED;
    std::string message;
    std::vector<idx> pattern;
    std::vector<double> coef;
    idx fits = 0, growths = 0, clamps = 0;
    double residual = std::numeric_limits<double>::quiet_NaN();
    double pivot_ratio = std::numeric_limits<double>::quiet_NaN();
    // Per-fit trace kept only when requested (small problems).
    std::vector<std::vector<idx>> trace_shadow, trace_candidates, trace_selected;
    std::vector<std::vector<double>> trace_scores;
};

// Adaptive column (src/spai.cpp:100-244).  A is the original CSR (its row j is
// B's column j), bt is B in CSR (its row i lists the j with a_ji != 0), d the
// row scaling.
Column construct_column(const Csr &a, const Csr &bt, const std::vector<double> &d, idx k,
                        double eps, idx alpha, idx beta, bool trace) {
    Column c;
    std::vector<idx> pattern(1, k), shadow, cand;
    std::vector<double> reduced, rhs, residual, scores;
    for (;;) {
        // Shadow: union of cols(A row j), j in J.
        shadow.clear();
        for (idx j : pattern)
            for (idx q = a.ptr[j]; q < a.ptr[j + 1]; ++q)
                shadow.push_back(a.col[q]);
        std::sort(shadow.begin(), shadow.end());
        shadow.erase(std::unique(shadow.begin(), shadow.end()), shadow.end());
        const idx m = (idx)shadow.size(), q = (idx)pattern.size();
        reduced.assign(m * q, 0.0);
        rhs.assign(m, 0.0);
        for (idx jj = 0; jj < q; ++jj) {
            const idx j = pattern[jj];
            // B(i, j) = a_ji * d_j for the stored a_ji (src/spai.cpp:338-348).
            idx p = 0;
            for (idx e = a.ptr[j]; e < a.ptr[j + 1]; ++e) {
                while (shadow[p] < a.col[e])
                    ++p;
                reduced[p * q + jj] = a.val[e] * d[j];
            }
        }
        for (idx i = 0; i < m; ++i)
            rhs[i] = shadow[i] == k ? 1.0 : 0.0;
        std::vector<double> sol;
        double pr = 0;
        const int fs = least_squares(reduced, rhs, m, q, sol, pr);
        if (fs != SPAI_SUCCESS) {
            c.status = fs;
            c.message = "Householder QR/triangular fit failed";
            return c;
        }
        residual.assign(m, 0.0);
        for (idx i = 0; i < m; ++i) {
            for (idx j = 0; j < q; ++j)
                residual[i] += reduced[i * q + j] * sol[j];
            residual[i] -= rhs[i];
        }
        for (idx i = 0; i < m; ++i)
            if (!std::isfinite(residual[i])) {
                c.status = SPAI_NUMERICAL_FAILURE;
                c.message = "Full fitted residual is nonfinite";
                return c;
            }
        const double rnorm = hypot_chain(residual.data(), m);
        if (!std::isfinite(rnorm)) {
            c.status = SPAI_NUMERICAL_FAILURE;
            c.message = "Residual norm is nonfinite";
            return c;
        }
        if (c.fits != 0)
            ++c.growths;
        ++c.fits;
        c.pattern = pattern;
        c.coef = sol;
        c.residual = rnorm;
        c.pivot_ratio = pr;
        if (trace)
            c.trace_shadow.push_back(shadow);
        if (rnorm <= eps) {
            c.status = SPAI_SUCCESS;
            return c;
        }
        if (c.growths == alpha) {
            c.status = SPAI_ITERATION_LIMIT;
            c.message = "Requested accuracy not reached within growth/refit budget";
            return c;
        }
        // Candidates: union of B rows i in I (k is in I), minus J, ascending.
        cand.clear();
        for (idx i : shadow)
            for (idx e = bt.ptr[i]; e < bt.ptr[i + 1]; ++e)
                cand.push_back(bt.col[e]);
        for (idx e = bt.ptr[k]; e < bt.ptr[k + 1]; ++e)
            cand.push_back(bt.col[e]);
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
        {
            std::vector<idx> kept;
            kept.reserve(cand.size());
            std::set_difference(cand.begin(), cand.end(), pattern.begin(), pattern.end(),
                                std::back_inserter(kept));
            cand.swap(kept);
        }
        if (trace)
            c.trace_candidates.push_back(cand);
        if (cand.empty()) {
            c.status = SPAI_EMPTY_CANDIDATES;
            c.message = "No candidate remains before convergence";
            return c;
        }
        scores.clear();
        const double rr = rnorm * rnorm;
        for (idx j : cand) {
            // dot/den over I in ascending order; only stored B(i,j) contribute.
            double dot = 0, den = 0;
            idx p = 0;
            for (idx e = a.ptr[j]; e < a.ptr[j + 1]; ++e) {
                const idx i = a.col[e];
                while (p < m && shadow[p] < i)
                    ++p;
                if (p < m && shadow[p] == i) {
                    const double value = a.val[e] * d[j];
                    dot += residual[p] * value;
                    den += value * value;
                }
            }
            const Score s = score_of(rr, dot, den);
            if (s.status != SPAI_SUCCESS) {
                c.status = s.status;
                c.message = "Candidate scoring failed for index " + std::to_string(j);
                return c;
            }
            if (s.clamped)
                ++c.clamps;
            scores.push_back(s.score);
        }
        const double mean = std::accumulate(scores.begin(), scores.end(), 0.0) /
                            static_cast<double>(scores.size());
        if (!std::isfinite(mean)) {
            c.status = SPAI_NUMERICAL_FAILURE;
            c.message = "Candidate mean score is nonfinite";
            return c;
        }
        std::vector<idx> order(cand.size());
        std::iota(order.begin(), order.end(), 0);
        std::sort(order.begin(), order.end(), [&](idx x, idx y) {
            return scores[x] < scores[y] || (scores[x] == scores[y] && cand[x] < cand[y]);
        });
        std::vector<idx> sel;
        for (idx o : order) {
            if ((idx)sel.size() == beta)
                break;
            if (scores[o] <= mean)
                sel.push_back(cand[o]);
        }
        if (trace) {
            c.trace_scores.push_back(scores);
            c.trace_selected.push_back(sel);
        }
        if (sel.empty()) {
            c.status = SPAI_NO_GROWTH;
            c.message = "No acceptable growth permitted before convergence";
            return c;
        }
        pattern.insert(pattern.end(), sel.begin(), sel.end());
        std::sort(pattern.begin(), pattern.end());
    }
}

struct Spai {
    int status = SPAI_INVALID_INPUT;
    std::string message;
    std::vector<double> scaling;
    std::vector<Column> cols;
    Csr p; // P = N^T D, canonical (zeros dropped)
};

// src/spai.cpp:314-374
Spai construct_spai(const Csr &a, double eps, idx alpha, idx beta, bool trace) {
    Spai r;
    r.p.n = a.n;
    r.p.ptr.assign(a.n + 1, 0);
    if (a.n == 0 || !std::isfinite(eps) || eps <= 0) {
        r.message = "Require positive square dimensions and positive finite epsilon";
        return r;
    }
    r.scaling.resize(a.n);
    for (idx row = 0; row < a.n; ++row) {
        if (lookup(a, row, row) == 0) {
            r.message = "Zero diagonal at row " + std::to_string(row);
            return r;
        }
        double mx = 0;
        for (idx q = a.ptr[row]; q < a.ptr[row + 1]; ++q)
            mx = std::max(mx, std::abs(a.val[q]));
        const double s = 1.0 / mx;
        if (!std::isfinite(s) || s == 0) {
            r.message = "Reciprocal row scaling outside binary64 range";
            return r;
        }
        r.scaling[row] = s;
    }
    for (idx row = 0; row < a.n; ++row)
        for (idx q = a.ptr[row]; q < a.ptr[row + 1]; ++q) {
            const double v = a.val[q] * r.scaling[row];
            if (!std::isfinite(v) || v == 0) {
                r.message = "Scaling changes nonzero structure/range";
                return r;
            }
        }
    // B in CSR: row i holds {(j, a_ji d_j)} ascending j (counting transpose).
    Csr bt;
    bt.n = a.n;
    bt.ptr.assign(a.n + 1, 0);
    for (idx q = 0; q < a.ptr[a.n]; ++q)
        ++bt.ptr[a.col[q] + 1];
    for (idx i = 0; i < a.n; ++i)
        bt.ptr[i + 1] += bt.ptr[i];
    bt.col.resize(a.ptr[a.n]);
    {
        std::vector<idx> fill(bt.ptr.begin(), bt.ptr.end() - 1);
        for (idx row = 0; row < a.n; ++row)
            for (idx q = a.ptr[row]; q < a.ptr[row + 1]; ++q)
                bt.col[fill[a.col[q]]++] = row;
    }
    r.status = SPAI_SUCCESS;
    r.cols.resize(a.n);
    for (idx k = 0; k < a.n; ++k) {
        if (g_progress && k > 0 && k % (1 << 19) == 0)
            g_progress(0, k, a.n);
        r.cols[k] = construct_column(a, bt, r.scaling, k, eps, alpha, beta, trace);
        const Column &c = r.cols[k];
        if (c.status != SPAI_SUCCESS && r.status == SPAI_SUCCESS) {
            r.status = c.status;
            r.message = "Column " + std::to_string(k) + ": " + c.message;
        }
        for (std::size_t i = 0; i < c.pattern.size(); ++i) {
            const double v = c.coef[i] * r.scaling[c.pattern[i]];
            if (!std::isfinite(v) || (c.coef[i] != 0 && v == 0)) {
                r.status = SPAI_NUMERICAL_FAILURE;
                r.message = "Preconditioner assembly outside binary64 range";
                return r;
            }
            if (v != 0) {
                r.p.col.push_back(c.pattern[i]);
                r.p.val.push_back(v);
            }
        }
        r.p.ptr[k + 1] = (idx)r.p.col.size();
    }
    return r;
}

// ---------------------------------------------------------------- GMRES
enum GmresStatus {
    G_SUCCESS,
    G_ZERO_RHS,
    G_EXACT_INITIAL,
    G_INVALID_INPUT,
    G_ZERO_PRECONDITIONED_RHS,
    G_BREAKDOWN,
    G_ITERATION_LIMIT,
    G_NONCONVERGENCE,
    G_NUMERICAL_FAILURE
};

double fin(double v, const char *what) {
    if (!std::isfinite(v))
        throw std::overflow_error(what);
    return v;
}
const char *kGm = "Nonfinite GMRES arithmetic";

double norm2_checked(const std::vector<double> &v) {
    double r = 0;
    for (double x : v)
        r = fin(std::hypot(r, x), kGm);
    return r;
}

std::vector<double> subtract(const std::vector<double> &a, const std::vector<double> &b) {
    std::vector<double> r(a.size());
    for (std::size_t i = 0; i < a.size(); ++i)
        r[i] = fin(a[i] - b[i], kGm);
    return r;
}

struct Gmres {
    int status = G_INVALID_INPUT;
    std::string message;
    bool converged = false;
    idx iterations = 0;
    std::vector<double> solution;
    double init_orig = NAN, init_rel = NAN, final_orig = NAN, final_rel = NAN;
    std::vector<double> estimates; // per iteration |g_{k+1}|/den
};

// src/gmres.cpp:108-262
Gmres gmres(const Csr &a, const Csr &p, const std::vector<double> &rhs, double tol, idx max_it,
            const std::vector<double> &x0) {
    Gmres res;
    const idx n = a.n;
    if (n == 0 || p.n != n || (idx)rhs.size() != n || (!x0.empty() && (idx)x0.size() != n) ||
        !std::isfinite(tol) || tol <= 0 || max_it > n) {
        res.message = "Invalid dimensions, initial guess, tolerance, or iteration cap";
        return res;
    }
    for (double v : rhs)
        if (!std::isfinite(v)) {
            res.message = "RHS must be finite";
            return res;
        }
    for (double v : x0)
        if (!std::isfinite(v)) {
            res.message = "Initial guess must be finite";
            return res;
        }
    const std::vector<double> initial = x0.empty() ? std::vector<double>(n, 0) : x0;
    res.solution = initial;
    const idx cap = max_it == 0 ? n : max_it;
    try {
        const std::vector<double> ax0 = multiply(a, initial);
        const std::vector<double> orig0 = subtract(rhs, ax0);
        res.init_orig = norm2_checked(orig0);
        res.final_orig = res.init_orig;
        if (res.init_orig == 0) {
            bool zero = std::all_of(rhs.begin(), rhs.end(), [](double v) { return v == 0; });
            res.status = zero ? G_ZERO_RHS : G_EXACT_INITIAL;
            res.converged = true;
            res.init_rel = res.final_rel = 0;
            return res;
        }
        const std::vector<double> prhs = multiply(p, rhs);
        const double den = norm2_checked(prhs);
        if (den == 0) {
            res.status = G_ZERO_PRECONDITIONED_RHS;
            res.message = "Zero preconditioned RHS with nonzero original initial residual";
            return res;
        }
        const std::vector<double> r0 = subtract(prhs, multiply(p, ax0));
        const double beta = norm2_checked(r0);
        res.init_rel = fin(beta / den, kGm);
        res.final_rel = res.init_rel;
        if (res.init_rel <= tol) {
            res.status = G_SUCCESS;
            res.converged = true;
            return res;
        }
        std::vector<std::vector<double>> basis;
        basis.emplace_back(n);
        for (idx i = 0; i < n; ++i)
            basis[0][i] = fin(r0[i] / beta, kGm);
        // Column-major upper Hessenberg, grown one column per iteration.
        std::vector<std::vector<double>> rot; // rot[k] has k+2 entries
        std::vector<double> cs, sn, g(1, beta);
        auto candidate = [&](idx count, std::vector<double> &out) -> bool {
            std::vector<double> y(count, 0);
            for (idx row = count; row-- > 0;) {
                if (rot[row][row] == 0)
                    return false;
                double value = g[row];
                for (idx c = row + 1; c < count; ++c)
                    value = fin(value - fin(rot[c][row] * y[c], kGm), kGm);
                y[row] = fin(value / rot[row][row], kGm);
            }
            out = initial;
            for (idx c = 0; c < count; ++c)
                for (idx r = 0; r < n; ++r)
                    out[r] = fin(out[r] + fin(basis[c][r] * y[c], kGm), kGm);
            return true;
        };
        auto check = [&](const std::vector<double> &x, double &orig, double &rel) {
            const std::vector<double> ax = multiply(a, x);
            orig = norm2_checked(subtract(rhs, ax));
            rel = fin(norm2_checked(subtract(prhs, multiply(p, ax))) / den, kGm);
        };
        for (idx it = 0; it < cap; ++it) {
            res.iterations = it + 1;
            std::vector<double> w = multiply(p, multiply(a, basis[it]));
            std::vector<double> h(it + 2, 0.0);
            for (idx c = 0; c <= it; ++c) {
                double dot = 0;
                const std::vector<double> &v = basis[c];
                for (idx r = 0; r < n; ++r)
                    dot = fin(dot + fin(v[r] * w[r], kGm), kGm);
                h[c] = dot;
                for (idx r = 0; r < n; ++r)
                    w[r] = fin(w[r] - fin(dot * v[r], kGm), kGm);
            }
            const double next = norm2_checked(w);
            h[it + 1] = next;
            const bool breakdown = next == 0;
            if (!breakdown) {
                for (double &x : w)
                    x = fin(x / next, kGm);
                basis.push_back(std::move(w));
            }
            for (idx r = 0; r < it; ++r) {
                const double up = h[r], lo = h[r + 1];
                h[r] = fin(cs[r] * up + sn[r] * lo, kGm);
                h[r + 1] = fin(-sn[r] * up + cs[r] * lo, kGm);
            }
            rot.push_back(h);
            const double upper = rot[it][it];
            const double radius = fin(std::hypot(upper, rot[it][it + 1]), kGm);
            bool defined = radius != 0;
            bool triggered = false;
            std::vector<double> cand;
            if (defined) {
                cs.push_back(fin(upper / radius, kGm));
                sn.push_back(fin(rot[it][it + 1] / radius, kGm));
                rot[it][it] = radius;
                rot[it][it + 1] = 0;
                const double old = g[it];
                g[it] = fin(cs[it] * old, kGm);
                g.push_back(fin(-sn[it] * old, kGm));
                const double est = fin(std::abs(g[it + 1]) / den, kGm);
                res.estimates.push_back(est);
                triggered = breakdown || est <= tol || it + 1 == cap;
                if (triggered)
                    defined = candidate(it + 1, cand);
            } else {
                cs.push_back(0);
                sn.push_back(0);
            }
            if (!defined) {
                // Recheck only the previous candidate (or the initial guess).
                std::vector<double> prev = initial;
                if (it > 0)
                    candidate(it, prev);
                res.solution = prev;
                double orig, rel;
                check(prev, orig, rel);
                res.final_orig = orig;
                res.final_rel = rel;
                const bool ok = rel <= tol;
                res.status = ok ? G_SUCCESS : G_BREAKDOWN;
                res.converged = ok;
                res.message = ok ? "Previously defined candidate verified at zero pivot"
                                 : "Zero Givens/triangular pivot";
                return res;
            }
            if (triggered) {
                double orig, rel;
                check(cand, orig, rel);
                res.final_orig = orig;
                res.final_rel = rel;
                res.solution = cand;
                if (rel <= tol) {
                    res.status = G_SUCCESS;
                    res.converged = true;
                    return res;
                }
                if (breakdown) {
                    res.status = G_BREAKDOWN;
                    res.message = "Arnoldi breakdown candidate failed recomputed residual criterion";
                    return res;
                }
            }
        }
        res.status = cap < n ? G_ITERATION_LIMIT : G_NONCONVERGENCE;
        res.message = "Iteration cap reached without verified convergence";
    } catch (const std::overflow_error &e) {
        res.status = G_NUMERICAL_FAILURE;
        res.message = e.what();
    }
    return res;
}

// ---------------------------------------------------------------- IR
enum IrStatus {
    IR_SUCCESS,
    IR_INVALID_INPUT,
    IR_CONSTRUCTION_FAILURE,
    IR_INNER_FAILURE,
    IR_STAGNATION,
    IR_ITERATION_LIMIT,
    IR_NUMERICAL_FAILURE
};

const char *kIr = "Nonfinite outer residual/backward-error arithmetic";

struct Diagnostics {
    std::vector<double> residual;
    double rinf = 0, r2 = 0, nbe = 0, cbe = 0;
};

double ratio(double num, double den) {
    if (den == 0)
        return num == 0 ? 0 : std::numeric_limits<double>::infinity();
    return num / den;
}

// src/ir.cpp:83-133 (forward-error part omitted: no reference vector here)
Diagnostics diagnose(const Csr &a, const std::vector<double> &b, const std::vector<double> &x) {
    for (double v : x)
        if (!std::isfinite(v))
            throw std::invalid_argument("Iterate must be finite");
    Diagnostics d;
    const std::vector<double> ax = multiply(a, x);
    d.residual.resize(a.n);
    double ainf = 0, xinf = 0, binf = 0;
    for (double v : x)
        xinf = std::max(xinf, std::abs(v));
    for (idx r = 0; r < a.n; ++r) {
        const double res = fin(b[r] - ax[r], kIr);
        d.residual[r] = res;
        d.rinf = std::max(d.rinf, std::abs(res));
        d.r2 = fin(std::hypot(d.r2, res), kIr);
        binf = std::max(binf, std::abs(b[r]));
        double rs = 0, cd = std::abs(b[r]);
        for (idx q = a.ptr[r]; q < a.ptr[r + 1]; ++q) {
            const double v = std::abs(a.val[q]);
            rs = fin(rs + v, kIr);
            cd = fin(cd + fin(v * std::abs(x[a.col[q]]), kIr), kIr);
        }
        ainf = std::max(ainf, rs);
        d.cbe = std::max(d.cbe, ratio(std::abs(res), cd));
    }
    const double den = fin(fin(ainf * xinf, kIr) + binf, kIr);
    d.nbe = ratio(d.rinf, den);
    return d;
}

struct Ir {
    int status = IR_INVALID_INPUT;
    std::string message;
    bool converged = false;
    idx outer_steps = 0, inner_attempts = 0;
    std::vector<double> solution;
    std::vector<Diagnostics> history; // residual vectors dropped after use
    std::vector<Gmres> inner;
    Spai spai;
    bool constructed = false;
};

const char *gmres_name(int s) {
    static const char *names[] = {"success",
                                  "zero_rhs",
                                  "exact_initial_solution",
                                  "invalid_input",
                                  "zero_preconditioned_rhs",
                                  "breakdown",
                                  "iteration_limit",
                                  "nonconvergence",
                                  "numerical_failure"};
    return names[s];
}
const char *spai_name(int s) {
    static const char *names[] = {"success",           "not_attempted",    "invalid_input",
                                  "rank_deficient",    "numerical_failure", "empty_candidates",
                                  "zero_denominator",  "invalid_score",     "no_growth",
                                  "iteration_limit"};
    return names[s];
}

bool valid(const Csr &a, const std::vector<double> &b, double tol, idx max_it, idx outer) {
    if (a.n == 0 || (idx)b.size() != a.n || outer > 20 || !std::isfinite(tol) || tol <= 0 ||
        max_it > a.n)
        return false;
    for (double v : b)
        if (!std::isfinite(v))
            return false;
    return true;
}

const double kTol = 10 * std::ldexp(1.0, -53); // src/ir.cpp:79-81

// src/ir.cpp:138-204
void refine(Ir &out, const Csr &a, const Csr &p, const std::vector<double> &b, double tol,
            idx max_it, idx outer, double relres_target) {
    try {
        out.solution = multiply(p, b);
        out.history.push_back(diagnose(a, b, out.solution));
        double bnorm = 0;
        for (double v : b)
            bnorm = std::hypot(bnorm, v);
        auto done = [&](const Diagnostics &d) {
            if (d.nbe > kTol)
                return false;
            return relres_target <= 0 || d.r2 <= relres_target * bnorm;
        };
        if (done(out.history.back())) {
            out.status = IR_SUCCESS;
            out.converged = true;
            return;
        }
        for (idx step = 0; step < outer; ++step) {
            const std::vector<double> r = out.history.back().residual;
            out.inner.push_back(gmres(a, p, r, tol, max_it, {}));
            ++out.inner_attempts;
            const Gmres &in = out.inner.back();
            if (!in.converged) {
                out.status = IR_INNER_FAILURE;
                out.message = "Correction GMRES failed: " + std::string(gmres_name(in.status)) +
                              ": " + in.message;
                return;
            }
            std::vector<double> next = out.solution;
            for (idx i = 0; i < a.n; ++i)
                next[i] = fin(next[i] + in.solution[i], kIr);
            out.history.back().residual.clear();
            out.history.push_back(diagnose(a, b, next));
            const bool unchanged = next == out.solution;
            out.solution.swap(next);
            ++out.outer_steps;
            if (g_progress)
                g_progress(1, out.outer_steps, (std::int64_t)in.iterations);
            if (done(out.history.back())) {
                out.status = IR_SUCCESS;
                out.converged = true;
                return;
            }
            if (unchanged) {
                out.status = IR_STAGNATION;
                out.message = "Unchanged outer iterate above original backward-error tolerance";
                return;
            }
        }
        out.status = IR_ITERATION_LIMIT;
        out.message = "Outer correction cap reached above original backward-error tolerance";
    } catch (const std::overflow_error &e) {
        out.status = IR_NUMERICAL_FAILURE;
        out.message = e.what();
    }
}

thread_local std::string g_error;

} // namespace

// ======================================================================= C API
extern "C" {
void oracle_set_progress(OracleProgress fn) {
    g_progress = fn;
}

struct OracleSpai {
    Spai s;
};
struct OracleIr {
    Ir r;
};

const char *oracle_last_error() {
    return g_error.c_str();
}

OracleSpai *oracle_spai_build(std::int64_t n, const std::int64_t *ptr, const std::int64_t *col,
                              const double *val, double eps, std::int64_t alpha,
                              std::int64_t beta, int trace) {
    try {
        auto *o = new OracleSpai;
        o->s = construct_spai(wrap(n, ptr, col, val), eps, alpha, beta, trace != 0);
        return o;
    } catch (const std::exception &e) {
        g_error = e.what();
        return nullptr;
    }
}
void oracle_spai_free(OracleSpai *o) {
    delete o;
}
int oracle_spai_status(const OracleSpai *o) {
    return o->s.status;
}
const char *oracle_spai_message(const OracleSpai *o) {
    return o->s.message.c_str();
}
std::int64_t oracle_spai_p_nnz(const OracleSpai *o) {
    return (std::int64_t)o->s.p.col.size();
}
void oracle_spai_p(const OracleSpai *o, std::int64_t *ptr, std::int64_t *col, double *val) {
    std::copy(o->s.p.ptr.begin(), o->s.p.ptr.end(), ptr);
    std::copy(o->s.p.col.begin(), o->s.p.col.end(), col);
    std::copy(o->s.p.val.begin(), o->s.p.val.end(), val);
}
void oracle_spai_scaling(const OracleSpai *o, double *out) {
    std::copy(o->s.scaling.begin(), o->s.scaling.end(), out);
}
// Per column: status, fits, growths, score clamps, final reduced residual, pivot ratio.
void oracle_spai_columns(const OracleSpai *o, std::int32_t *status, std::int32_t *fits,
                         std::int32_t *growths, std::int32_t *clamps, double *residual,
                         double *pivot) {
    for (std::size_t k = 0; k < o->s.cols.size(); ++k) {
        const Column &c = o->s.cols[k];
        status[k] = c.status;
        fits[k] = (std::int32_t)c.fits;
        growths[k] = (std::int32_t)c.growths;
        clamps[k] = (std::int32_t)c.clamps;
        residual[k] = c.residual;
        pivot[k] = c.pivot_ratio;
    }
}
// Column patterns (J) and coefficients, including exact-zero coefficients.
std::int64_t oracle_spai_pattern_nnz(const OracleSpai *o) {
    std::int64_t t = 0;
    for (const Column &c : o->s.cols)
        t += (std::int64_t)c.pattern.size();
    return t;
}
void oracle_spai_patterns(const OracleSpai *o, std::int64_t *ptr, std::int64_t *idx_out,
                          double *coef) {
    std::int64_t t = 0;
    ptr[0] = 0;
    for (std::size_t k = 0; k < o->s.cols.size(); ++k) {
        const Column &c = o->s.cols[k];
        for (std::size_t i = 0; i < c.pattern.size(); ++i) {
            idx_out[t] = c.pattern[i];
            coef[t] = c.coef[i];
            ++t;
        }
        ptr[k + 1] = t;
    }
}
// Trace of fit f of column k (requires trace=1): which = 0 shadow, 1 candidates,
// 2 selected.  Returns the length; copies when out != NULL.
std::int64_t oracle_spai_trace(const OracleSpai *o, std::int64_t k, std::int64_t f, int which,
                               std::int64_t *out) {
    const Column &c = o->s.cols[k];
    const std::vector<std::vector<idx>> &src =
        which == 0 ? c.trace_shadow : which == 1 ? c.trace_candidates : c.trace_selected;
    if (f < 0 || f >= (std::int64_t)src.size())
        return -1;
    if (out)
        std::copy(src[f].begin(), src[f].end(), out);
    return (std::int64_t)src[f].size();
}
std::int64_t oracle_spai_trace_scores(const OracleSpai *o, std::int64_t k, std::int64_t f,
                                      double *out) {
    const Column &c = o->s.cols[k];
    if (f < 0 || f >= (std::int64_t)c.trace_scores.size())
        return -1;
    if (out)
        std::copy(c.trace_scores[f].begin(), c.trace_scores[f].end(), out);
    return (std::int64_t)c.trace_scores[f].size();
}

// Columns `cols[0..m)` only (the scaling/validity pass still covers all rows):
// for bit-exact spot checks of large constructions.  Writes per requested
// column status/fits/growths and its pattern + coefficients into
// (ptr[m+1], idx[cap], coef[cap]); returns the total pattern size or -1.
std::int64_t oracle_spai_subset(std::int64_t n, const std::int64_t *ptr, const std::int64_t *col,
                                const double *val, double eps, std::int64_t alpha,
                                std::int64_t beta, std::int64_t m, const std::int64_t *cols,
                                std::int32_t *status, std::int32_t *fits, std::int32_t *growths,
                                std::int64_t *optr, std::int64_t *oidx, double *ocoef,
                                std::int64_t cap) {
    try {
        const Csr a = wrap(n, ptr, col, val);
        std::vector<double> d(n);
        for (idx row = 0; row < n; ++row) {
            double mx = 0;
            for (idx q = a.ptr[row]; q < a.ptr[row + 1]; ++q)
                mx = std::max(mx, std::abs(a.val[q]));
            d[row] = 1.0 / mx;
        }
        Csr bt;
        bt.n = n;
        bt.ptr.assign(n + 1, 0);
        for (idx q = 0; q < a.ptr[n]; ++q)
            ++bt.ptr[a.col[q] + 1];
        for (idx i = 0; i < n; ++i)
            bt.ptr[i + 1] += bt.ptr[i];
        bt.col.resize(a.ptr[n]);
        std::vector<idx> fill(bt.ptr.begin(), bt.ptr.end() - 1);
        for (idx row = 0; row < n; ++row)
            for (idx q = a.ptr[row]; q < a.ptr[row + 1]; ++q)
                bt.col[fill[a.col[q]]++] = row;
        std::int64_t t = 0;
        optr[0] = 0;
        for (std::int64_t i = 0; i < m; ++i) {
            const Column c = construct_column(a, bt, d, cols[i], eps, alpha, beta, false);
            status[i] = c.status;
            fits[i] = (std::int32_t)c.fits;
            growths[i] = (std::int32_t)c.growths;
            for (std::size_t j = 0; j < c.pattern.size(); ++j) {
                if (t >= cap)
                    return -1;
                oidx[t] = c.pattern[j];
                ocoef[t] = c.coef[j];
                ++t;
            }
            optr[i + 1] = t;
        }
        return t;
    } catch (const std::exception &e) {
        g_error = e.what();
        return -1;
    }
}

double oracle_score(double rr, double dot, double den, int *status, int *clamped) {
    const Score s = score_of(rr, dot, den);
    *status = s.status;
    *clamped = s.clamped;
    return s.score;
}

double oracle_hypot_chain(const double *v, std::int64_t len) {
    return hypot_chain(v, len);
}

// One GMRES solve.  Outputs: solution[n]; info[0]=status, [1]=converged,
// [2]=iterations; res[0..3] = initial original, initial relative, final
// original, final relative.
int oracle_gmres(std::int64_t n, const std::int64_t *aptr, const std::int64_t *acol,
                 const double *aval, const std::int64_t *pptr, const std::int64_t *pcol,
                 const double *pval, const double *rhs, const double *x0, double tol,
                 std::int64_t max_it, double *solution, std::int64_t *info, double *res) {
    try {
        const Csr a = wrap(n, aptr, acol, aval), p = wrap(n, pptr, pcol, pval);
        std::vector<double> b(rhs, rhs + n), init;
        if (x0)
            init.assign(x0, x0 + n);
        const Gmres g = gmres(a, p, b, tol, max_it, init);
        if (!g.solution.empty())
            std::copy(g.solution.begin(), g.solution.end(), solution);
        info[0] = g.status;
        info[1] = g.converged;
        info[2] = g.iterations;
        res[0] = g.init_orig;
        res[1] = g.init_rel;
        res[2] = g.final_orig;
        res[3] = g.final_rel;
        return 0;
    } catch (const std::exception &e) {
        g_error = e.what();
        return -1;
    }
}

// Full SPAI-GMRES-IR (p == NULL: construct SPAI first, src/ir.cpp:206-227) or
// refinement with a fixed P (src/ir.cpp:138-204).  relres_target > 0 adds the
// requirement ||b-Ax||_2 <= target*||b||_2 to the backward-error stop.
OracleIr *oracle_ir(std::int64_t n, const std::int64_t *aptr, const std::int64_t *acol,
                    const double *aval, const std::int64_t *pptr, const std::int64_t *pcol,
                    const double *pval, const double *rhs, double eps, std::int64_t alpha,
                    std::int64_t beta, double tol, std::int64_t max_it, std::int64_t outer,
                    double relres_target) {
    try {
        auto *o = new OracleIr;
        Ir &r = o->r;
        const Csr a = wrap(n, aptr, acol, aval);
        std::vector<double> b(rhs, rhs + n);
        if (!valid(a, b, tol, max_it, outer)) {
            r.message = "Invalid dimensions, finite inputs, inner options, or outer cap";
            return o;
        }
        if (!pptr) {
            r.constructed = true;
            r.spai = construct_spai(a, eps, alpha, beta, false);
            if (r.spai.status != SPAI_SUCCESS) {
                r.status = IR_CONSTRUCTION_FAILURE;
                r.message = "SPAI failed: " + std::string(spai_name(r.spai.status)) + ": " +
                            r.spai.message;
                return o;
            }
            refine(r, a, r.spai.p, b, tol, max_it, outer, relres_target);
        } else {
            refine(r, a, wrap(n, pptr, pcol, pval), b, tol, max_it, outer, relres_target);
        }
        return o;
    } catch (const std::exception &e) {
        g_error = e.what();
        return nullptr;
    }
}
void oracle_ir_free(OracleIr *o) {
    delete o;
}
// info: status, converged, outer_steps, inner_attempts, history length, spai status
void oracle_ir_info(const OracleIr *o, std::int64_t *info) {
    info[0] = o->r.status;
    info[1] = o->r.converged;
    info[2] = o->r.outer_steps;
    info[3] = o->r.inner_attempts;
    info[4] = (std::int64_t)o->r.history.size();
    info[5] = o->r.constructed ? o->r.spai.status : -1;
}
const char *oracle_ir_message(const OracleIr *o) {
    return o->r.message.c_str();
}
void oracle_ir_solution(const OracleIr *o, double *x) {
    std::copy(o->r.solution.begin(), o->r.solution.end(), x);
}
// Per history entry h: residual_inf, residual_2, nbe, cbe.
void oracle_ir_history(const OracleIr *o, double *out) {
    for (std::size_t h = 0; h < o->r.history.size(); ++h) {
        out[4 * h + 0] = o->r.history[h].rinf;
        out[4 * h + 1] = o->r.history[h].r2;
        out[4 * h + 2] = o->r.history[h].nbe;
        out[4 * h + 3] = o->r.history[h].cbe;
    }
}
// Per inner solve: status, converged, iterations.
void oracle_ir_inner(const OracleIr *o, std::int64_t *out) {
    for (std::size_t i = 0; i < o->r.inner.size(); ++i) {
        out[3 * i + 0] = o->r.inner[i].status;
        out[3 * i + 1] = o->r.inner[i].converged;
        out[3 * i + 2] = o->r.inner[i].iterations;
    }
}
std::int64_t oracle_ir_p_nnz(const OracleIr *o) {
    return (std::int64_t)o->r.spai.p.col.size();
}
void oracle_ir_p(const OracleIr *o, std::int64_t *ptr, std::int64_t *col, double *val) {
    std::copy(o->r.spai.p.ptr.begin(), o->r.spai.p.ptr.end(), ptr);
    std::copy(o->r.spai.p.col.begin(), o->r.spai.p.col.end(), col);
    std::copy(o->r.spai.p.val.begin(), o->r.spai.p.val.end(), val);
}

// Canonical CSR from coordinate entries (src/cpu.cpp:48-76): every entry's
// indices, then its value, are checked in input order; a stable sort by
// (row, column) keeps duplicates in input order; each run of duplicates is
// summed left to right, a non-finite partial sum rejected; zero sums are
// dropped.  Returns nullptr with the reference's message on error.
struct OracleCsr {
    std::vector<std::int64_t> ptr, col;
    std::vector<double> val;
};
OracleCsr *oracle_from_entries(std::int64_t rows, std::int64_t cols, std::int64_t count,
                               const std::int64_t *r, const std::int64_t *c, const double *v) {
    try {
        struct Entry {
            std::int64_t row, col;
            double value;
        };
        std::vector<Entry> e((std::size_t)count);
        for (std::int64_t i = 0; i < count; ++i) {
            if (r[i] < 0 || r[i] >= rows || c[i] < 0 || c[i] >= cols)
                throw std::invalid_argument("Entry index outside matrix dimensions");
            if (!std::isfinite(v[i]))
                throw std::invalid_argument("Matrix entries must be finite");
            e[(std::size_t)i] = Entry{r[i], c[i], v[i]};
        }
        std::stable_sort(e.begin(), e.end(), [](const Entry &x, const Entry &y) {
            return x.row != y.row ? x.row < y.row : x.col < y.col;
        });
        std::unique_ptr<OracleCsr> o(new OracleCsr);
        o->ptr.assign((std::size_t)rows + 1, 0);
        std::size_t i = 0;
        while (i < e.size()) {
            const Entry head = e[i];
            double sum = head.value;
            for (++i; i < e.size() && e[i].row == head.row && e[i].col == head.col; ++i) {
                sum += e[i].value;
                if (!std::isfinite(sum))
                    throw std::invalid_argument("Duplicate sum is nonfinite");
            }
            if (sum != 0.0) {
                o->col.push_back(head.col);
                o->val.push_back(sum);
                ++o->ptr[(std::size_t)head.row + 1];
            }
        }
        for (std::int64_t k = 0; k < rows; ++k)
            o->ptr[(std::size_t)k + 1] += o->ptr[(std::size_t)k];
        return o.release();
    } catch (const std::exception &ex) {
        g_error = ex.what();
        return nullptr;
    }
}
std::int64_t oracle_csr_nnz(const OracleCsr *o) {
    return (std::int64_t)o->col.size();
}
void oracle_csr_get(const OracleCsr *o, std::int64_t *ptr, std::int64_t *col, double *val) {
    std::copy(o->ptr.begin(), o->ptr.end(), ptr);
    std::copy(o->col.begin(), o->col.end(), col);
    std::copy(o->val.begin(), o->val.end(), val);
}
void oracle_csr_free(OracleCsr *o) {
    delete o;
}

} // extern "C"
