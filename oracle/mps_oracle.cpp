// mps_oracle.cpp — CPU FP64 restatement of wqc-core's MPS hot path.
//
// TEST INFRASTRUCTURE ONLY. This file is the correctness oracle and the CPU timing
// baseline for the HIP engine. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker. The product library never
// links or calls it.
//
// Every function cites the reference item it restates (paths relative to the
// reference repository root). Arithmetic mirrors Rust's num-complex: no fused
// multiply-add (build with -ffp-contract=off), products written out as
// (ac - bd, ad + bc), sums accumulated left to right from zero.
//
// Substitutions for third-party code the reference calls (nalgebra 0.33.3):
//  - SVD: the reference tries nalgebra's Golub-Kahan SVD, then its transpose, then its
//    own one-sided Jacobi SVD (src/tn/svd_guard.rs:91-139). Here the Jacobi SVD
//    (src/tn/svd_guard.rs:152-229) is used directly; it is the reference's own accepted
//    factorization and gives the same singular values and the same gauge-invariant
//    products U_k S_k V_k^H to rounding. Which phase each singular vector gets differs,
//    which the sqrt(sigma) split carries into site tensors but not into amplitudes.
//  - QR (thin_q, src/tn/mps.rs:1091-1101): Householder QR. Column phases of Q differ
//    from nalgebra's; the range finder's output Q.U_B.S.V_B^H does not depend on them.
//  - Complex matmul (src/tn/mps.rs:1055-1079): plain triple loop.
//
// Parity status per function is listed in the header comment of tests/test_oracle.py.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;

// ---------------------------------------------------------------------------------
// num-complex restatement
// ---------------------------------------------------------------------------------
struct C {
    double re, im;
};
inline C cx(double r, double i) { return C{r, i}; }
inline C operator+(C a, C b) { return C{a.re + b.re, a.im + b.im}; }
inline C operator-(C a, C b) { return C{a.re - b.re, a.im - b.im}; }
inline C operator-(C a) { return C{-a.re, -a.im}; }
inline C operator*(C a, C b) { return C{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline C operator*(C a, double k) { return C{a.re * k, a.im * k}; }
inline C operator/(C a, double k) { return C{a.re / k, a.im / k}; }
inline C& operator+=(C& a, C b) {
    a.re += b.re;
    a.im += b.im;
    return a;
}
inline C conj(C a) { return C{a.re, -a.im}; }
inline double norm_sqr(C a) { return a.re * a.re + a.im * a.im; }
inline double cnorm(C a) { return std::hypot(a.re, a.im); }
const C ZERO{0.0, 0.0};
const C ONE{1.0, 0.0};

// Rust f64::signum: +1 for +0.0 and positives, -1 for -0.0 and negatives.
inline double signum(double x) { return std::signbit(x) ? -1.0 : 1.0; }

// ---------------------------------------------------------------------------------
// gates (src/tn/gates.rs:10-112, src/tn/mps.rs:1173-1180)
// ---------------------------------------------------------------------------------
enum Kind {
    GX = 1, GY = 2, GZ = 3, GH = 4, GS = 5, GT = 6, GCNOT = 7, GCZ = 8, GCCNOT = 9,
    GRX = 10, GRY = 11, GRZ = 12, GMEASURE = 13, GRESET = 14, GIF = 15
};

struct OGate {
    int32_t kind, q0, q1, q2;
    double angle;
};

struct Mat2 {
    C m[2][2];
};
struct Mat4 {
    C m[2][2][2][2];  // [c_out][t_out][c_in][t_in]
};

Mat2 unary_matrix(const OGate& g) {
    Mat2 u{};
    const double fs2 = 0.70710678118654752440;  // std::f64::consts::FRAC_1_SQRT_2
    switch (g.kind) {
        case GH: {
            double k = 1.0 / std::sqrt(2.0);
            u.m[0][0] = cx(k, 0);
            u.m[0][1] = cx(k, 0);
            u.m[1][0] = cx(k, 0);
            u.m[1][1] = cx(-k, -0.0);
            break;
        }
        case GX:
            u.m[0][1] = ONE;
            u.m[1][0] = ONE;
            break;
        case GY:
            u.m[0][1] = cx(-0.0, -1.0);
            u.m[1][0] = cx(0.0, 1.0);
            break;
        case GZ:
            u.m[0][0] = ONE;
            u.m[1][1] = cx(-1.0, -0.0);
            break;
        case GS:
            u.m[0][0] = ONE;
            u.m[1][1] = cx(0.0, 1.0);
            break;
        case GT:
            u.m[0][0] = ONE;
            u.m[1][1] = cx(fs2, fs2);
            break;
        case GRX: {
            double s = std::sin(g.angle / 2.0), c = std::cos(g.angle / 2.0);
            u.m[0][0] = cx(c, 0);
            u.m[0][1] = cx(0, -s);
            u.m[1][0] = cx(0, -s);
            u.m[1][1] = cx(c, 0);
            break;
        }
        case GRY: {
            double s = std::sin(g.angle / 2.0), c = std::cos(g.angle / 2.0);
            u.m[0][0] = cx(c, 0);
            u.m[0][1] = cx(-s, 0);
            u.m[1][0] = cx(s, 0);
            u.m[1][1] = cx(c, 0);
            break;
        }
        case GRZ: {
            double s = std::sin(g.angle / 2.0), c = std::cos(g.angle / 2.0);
            u.m[0][0] = cx(c, -s);
            u.m[1][1] = cx(c, s);
            break;
        }
        default:
            break;
    }
    return u;
}

Mat2 tdg_matrix() {
    Mat2 u{};
    const double fs2 = 0.70710678118654752440;
    u.m[0][0] = ONE;
    u.m[1][1] = cx(fs2, -fs2);
    return u;
}

Mat4 cnot_matrix() {
    Mat4 u{};
    u.m[0][0][0][0] = ONE;
    u.m[0][1][0][1] = ONE;
    u.m[1][0][1][1] = ONE;
    u.m[1][1][1][0] = ONE;
    return u;
}

Mat4 cz_matrix() {
    Mat4 u{};
    u.m[0][0][0][0] = ONE;
    u.m[0][1][0][1] = ONE;
    u.m[1][0][1][0] = ONE;
    u.m[1][1][1][1] = cx(-1.0, -0.0);
    return u;
}

Mat4 swap_matrix() {
    Mat4 u{};
    u.m[0][0][0][0] = ONE;
    u.m[0][1][1][0] = ONE;
    u.m[1][0][0][1] = ONE;
    u.m[1][1][1][1] = ONE;
    return u;
}

Mat4 two_qubit_matrix(const OGate& g) { return g.kind == GCZ ? cz_matrix() : cnot_matrix(); }

// ---------------------------------------------------------------------------------
// dense column-major matrix (nalgebra DMatrix stand-in)
// ---------------------------------------------------------------------------------
struct Mat {
    size_t rows = 0, cols = 0;
    std::vector<C> d;  // column-major
    Mat() = default;
    Mat(size_t r, size_t c) : rows(r), cols(c), d(r * c, ZERO) {}
    C& operator()(size_t r, size_t c) { return d[c * rows + r]; }
    const C& operator()(size_t r, size_t c) const { return d[c * rows + r]; }
    C* col(size_t c) { return d.data() + c * rows; }
    const C* col(size_t c) const { return d.data() + c * rows; }
};

Mat adjoint_of(const Mat& a) {
    Mat o(a.cols, a.rows);
    for (size_t r = 0; r < a.rows; ++r)
        for (size_t c = 0; c < a.cols; ++c) o(c, r) = conj(a(r, c));
    return o;
}

struct Svd {
    Mat u;      // rows x k
    Mat v_t;    // k x cols
    std::vector<double> s;
    bool ok = false;
};

// One-sided Hestenes Jacobi SVD (src/tn/svd_guard.rs:152-229).
Svd jacobi_svd(const Mat& a) {
    if (a.rows < a.cols) {
        Svd t = jacobi_svd(adjoint_of(a));
        Svd o;
        o.u = adjoint_of(t.v_t);
        o.v_t = adjoint_of(t.u);
        o.s = t.s;
        o.ok = true;
        return o;
    }
    const size_t m = a.rows, n = a.cols;
    Mat w = a;
    Mat v(n, n);
    for (size_t i = 0; i < n; ++i) v(i, i) = ONE;
    for (int sweep = 0; sweep < 80; ++sweep) {
        bool rotated = false;
        for (size_t p = 0; p < n; ++p) {
            for (size_t q = p + 1; q < n; ++q) {
                const C* wp = w.col(p);
                const C* wq = w.col(q);
                double alpha = 0.0, beta = 0.0;
                for (size_t r = 0; r < m; ++r) alpha += norm_sqr(wp[r]);
                for (size_t r = 0; r < m; ++r) beta += norm_sqr(wq[r]);
                C gamma = ZERO;
                for (size_t r = 0; r < m; ++r) gamma += conj(wp[r]) * wq[r];
                double g = cnorm(gamma);
                if (g <= 1e-15 * std::sqrt(alpha * beta) || g == 0.0) continue;
                rotated = true;
                C phase = gamma / g;
                double zeta = (beta - alpha) / (2.0 * g);
                double t = signum(zeta) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                double c = 1.0 / std::sqrt(1.0 + t * t);
                double s = c * t;
                C pc = conj(phase);
                for (Mat* mat : {&w, &v}) {
                    C* cp = mat->col(p);
                    C* cq = mat->col(q);
                    for (size_t r = 0; r < mat->rows; ++r) {
                        C xp = cp[r];
                        C xq = cq[r] * pc;
                        cp[r] = xp * c - xq * s;
                        cq[r] = (xp * s + xq * c) * phase;
                    }
                }
            }
        }
        if (!rotated) break;
    }
    std::vector<std::pair<double, size_t>> order(n);
    for (size_t j = 0; j < n; ++j) {
        double acc = 0.0;
        const C* wj = w.col(j);
        for (size_t r = 0; r < m; ++r) acc += norm_sqr(wj[r]);
        order[j] = {std::sqrt(acc), j};
    }
    // Stable sort, descending (Rust sort_by with y.total_cmp(x)).
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<double, size_t>& x, const std::pair<double, size_t>& y) {
                         return x.first > y.first;
                     });
    Svd o;
    o.u = Mat(m, n);
    o.v_t = Mat(n, n);
    o.s.assign(n, 0.0);
    for (size_t k = 0; k < n; ++k) {
        double sigma = order[k].first;
        size_t j = order[k].second;
        o.s[k] = sigma;
        if (sigma > 0.0) {
            for (size_t r = 0; r < m; ++r) o.u(r, k) = w(r, j) / sigma;
        }
        for (size_t r = 0; r < n; ++r) o.v_t(k, r) = conj(v(r, j));
    }
    o.ok = true;
    return o;
}

// svd_reproduces (src/tn/svd_guard.rs:242-283).
bool svd_reproduces(const Mat& a, const Svd& svd) {
    if (!svd.ok) return false;
    const auto& s = svd.s;
    for (size_t i = 0; i < s.size(); ++i) {
        if (!std::isfinite(s[i]) || s[i] < 0.0) return false;
        if (i + 1 < s.size() && s[i] < s[i + 1]) return false;
    }
    double fro2 = 0.0;
    for (const C& c : a.d) fro2 += norm_sqr(c);
    double s2 = 0.0;
    for (double x : s) s2 += x * x;
    const double minpos = 2.2250738585072014e-308;
    double scale = std::max(std::sqrt(fro2), minpos);
    if (std::fabs(s2 - fro2) > 1e-9 * std::max(fro2, minpos) + 1e-28) return false;
    size_t k = s.size();
    if (svd.u.cols < k || svd.v_t.rows < k || svd.u.rows != a.rows || svd.v_t.cols != a.cols)
        return false;
    for (uint32_t probe = 0; probe < 2; ++probe) {
        std::vector<C> x(a.cols);
        for (size_t j = 0; j < a.cols; ++j) {
            double t = ((double)j + 1.0) * (0.754877666 + (double)probe * 0.56984029);
            x[j] = cx(std::sin(t), std::cos(t));
        }
        std::vector<C> direct(a.rows, ZERO);
        for (size_t c = 0; c < a.cols; ++c)
            for (size_t r = 0; r < a.rows; ++r) direct[r] += a(r, c) * x[c];
        std::vector<C> inner(k, ZERO);
        for (size_t c = 0; c < a.cols; ++c)
            for (size_t row = 0; row < k; ++row) inner[row] += svd.v_t(row, c) * x[c];
        for (size_t row = 0; row < k; ++row) inner[row] = inner[row] * s[row];
        std::vector<C> rebuilt(a.rows, ZERO);
        for (size_t c = 0; c < k; ++c)
            for (size_t r = 0; r < a.rows; ++r) rebuilt[r] += svd.u(r, c) * inner[c];
        double xn = 0.0;
        for (const C& v : x) xn += norm_sqr(v);
        xn = std::sqrt(xn);
        double err = 0.0;
        for (size_t r = 0; r < a.rows; ++r) err += norm_sqr(direct[r] - rebuilt[r]);
        err = std::sqrt(err);
        if (err > 1e-9 * scale * xn + 1e-14) return false;
    }
    return true;
}

struct Counters {
    uint64_t units = 0, exact = 0, rsvd_kept = 0, rsvd_fallbacks = 0, svd_failures = 0,
             truncated = 0;
};

// checked_svd (src/tn/svd_guard.rs:91-139) with the Jacobi SVD as the factorization.
Svd checked_svd(Mat mat, Counters* ctr) {
    double fro = 0.0;
    for (const C& c : mat.d) fro += norm_sqr(c);
    fro = std::sqrt(fro);
    double dust = fro * 1e-14;
    for (C& c : mat.d)
        if (cnorm(c) < dust) c = ZERO;
    Svd svd = jacobi_svd(mat);
    if (svd_reproduces(mat, svd)) return svd;
    if (ctr) ctr->svd_failures++;
    svd.ok = false;
    return svd;
}

// Truncation rule (src/tn/svd_guard.rs:12-76).
const double SINGULAR_CUTOFF = 1e-12;
const double CHI_REL_GAP = 1e-3;
const double CUTOFF_ABS_BAND = 1e-10;

size_t retained_chi(const std::vector<double>& v, size_t max_bond, size_t rows, size_t cols) {
    if (v.empty() || max_bond == 0 || rows == 0 || cols == 0) return 0;
    size_t rank = 0;
    for (double s : v)
        if (s > SINGULAR_CUTOFF) rank++;
    rank = std::max<size_t>(rank, 1);
    return std::min(std::min(rank, max_bond), std::min(rows, cols));
}

bool needs_f64_fallback(const std::vector<double>& v, size_t chi) {
    for (double s : v)
        if (!std::isfinite(s) || s < 0.0) return true;
    for (double s : v)
        if (s > 0.0 && std::fabs(s - SINGULAR_CUTOFF) <= CUTOFF_ABS_BAND) return true;
    if (chi > 0 && chi < v.size()) {
        double kept = v[chi - 1], dropped = v[chi];
        double scale = std::max(std::max(kept, dropped), 1e-30);
        if (std::fabs(kept - dropped) / scale < CHI_REL_GAP) return true;
    }
    return false;
}

// ---------------------------------------------------------------------------------
// sites and the two-site split (src/tn/mps.rs)
// ---------------------------------------------------------------------------------
struct Site {
    size_t l = 1, r = 1;
    std::vector<C> d;  // (l, 2, r) row-major
    C& at(size_t a, size_t s, size_t b) { return d[a * 2 * r + s * r + b]; }
    const C& at(size_t a, size_t s, size_t b) const { return d[a * 2 * r + s * r + b]; }
};

// theta (dl, 2, 2, dr) row-major == (2dl) x (2dr) with row a*2+s, col t*dr+b.
struct Theta {
    size_t dl = 0, dr = 0;
    std::vector<C> d;
    C& at(size_t a, size_t s, size_t t, size_t b) { return d[((a * 2 + s) * 2 + t) * dr + b]; }
    const C& at(size_t a, size_t s, size_t t, size_t b) const {
        return d[((a * 2 + s) * 2 + t) * dr + b];
    }
};

struct SplitError {
    int code = 0;  // 0 ok, 3 exec failed, 5 bond exceeded
    std::string msg;
};

// cpu_merge (src/tn/mps.rs:812-831).
Theta cpu_merge(const Site& left, const Site& right) {
    Theta th;
    th.dl = left.l;
    th.dr = right.r;
    size_t bond = std::min(left.r, right.l);
    th.d.assign(th.dl * 4 * th.dr, ZERO);
    for (size_t a = 0; a < th.dl; ++a)
        for (size_t b = 0; b < th.dr; ++b)
            for (size_t s = 0; s < 2; ++s)
                for (size_t t = 0; t < 2; ++t) {
                    C sum = ZERO;
                    for (size_t g = 0; g < bond; ++g) sum += left.at(a, s, g) * right.at(g, t, b);
                    th.at(a, s, t, b) = sum;
                }
    return th;
}

// apply_two_qubit_unitary (src/tn/mps.rs:833-866; inline copy :396-421).
Theta apply_two_qubit_unitary(const Theta& th, const Mat4& u, bool ctrl_is_first) {
    Theta o;
    o.dl = th.dl;
    o.dr = th.dr;
    o.d.assign(th.d.size(), ZERO);
    for (size_t a = 0; a < th.dl; ++a)
        for (size_t b = 0; b < th.dr; ++b)
            for (int so = 0; so < 2; ++so)
                for (int to = 0; to < 2; ++to) {
                    C sum = ZERO;
                    for (int si = 0; si < 2; ++si)
                        for (int ti = 0; ti < 2; ++ti) {
                            int co = ctrl_is_first ? so : to, tco = ctrl_is_first ? to : so;
                            int ci = ctrl_is_first ? si : ti, tci = ctrl_is_first ? ti : si;
                            sum += u.m[co][tco][ci][tci] * th.at(a, si, ti, b);
                        }
                    o.at(a, so, to, b) = sum;
                }
    return o;
}

// truncated_svd_split (src/tn/mps.rs:1124-1171) via svd_split_two_site (:1182-1201).
bool svd_split_two_site(const Theta& th, size_t max_bond, bool reject, Site& left, Site& right,
                        Counters* ctr, SplitError& err, double* discarded) {
    size_t dl = th.dl, dr = th.dr, rows = dl * 2, cols = 2 * dr;
    Mat mat(rows, cols);
    for (size_t l = 0; l < dl; ++l)
        for (size_t s = 0; s < 2; ++s)
            for (size_t t = 0; t < 2; ++t)
                for (size_t r = 0; r < dr; ++r) mat(l * 2 + s, t * dr + r) = th.at(l, s, t, r);
    Svd svd = checked_svd(mat, ctr);
    if (!svd.ok) {
        err.code = 3;
        err.msg = "SVD of a " + std::to_string(rows) + "x" + std::to_string(cols) +
                  " matrix failed its check";
        return false;
    }
    size_t chi = retained_chi(svd.s, max_bond, rows, cols);
    if (reject) {
        for (size_t i = chi; i < svd.s.size(); ++i)
            if (svd.s[i] > SINGULAR_CUTOFF) {
                err.code = 5;
                err.msg = "mps bond exceeded chi";
                return false;
            }
    }
    if (discarded) {
        double tot = 0, drop = 0;
        for (size_t i = 0; i < svd.s.size(); ++i) {
            tot += svd.s[i] * svd.s[i];
            if (i >= chi) drop += svd.s[i] * svd.s[i];
        }
        *discarded = tot > 0 ? drop / tot : 0.0;
    }
    if (ctr) {
        ctr->exact++;
        for (size_t i = chi; i < svd.s.size(); ++i)
            if (svd.s[i] > SINGULAR_CUTOFF) {
                ctr->truncated++;
                break;
            }
    }
    left.l = dl;
    left.r = chi;
    left.d.assign(dl * 2 * chi, ZERO);
    right.l = chi;
    right.r = dr;
    right.d.assign(chi * 2 * dr, ZERO);
    for (size_t m = 0; m < chi; ++m) {
        double s = std::sqrt(svd.s[m]);
        for (size_t l = 0; l < dl; ++l)
            for (size_t p = 0; p < 2; ++p) left.at(l, p, m) = svd.u(l * 2 + p, m) * s;
        for (size_t r = 0; r < dr; ++r)
            for (size_t t = 0; t < 2; ++t) right.at(m, t, r) = svd.v_t(m, t * dr + r) * s;
    }
    return true;
}

// rademacher / mix_seed (src/tn/mps.rs:1103-1122).
std::vector<C> rademacher(size_t cols, size_t ell, uint64_t salt) {
    uint64_t state = salt | 1ull;
    double scale = 1.0 / std::sqrt((double)cols);
    std::vector<C> om;
    om.reserve(cols * ell);
    for (size_t i = 0; i < cols * ell; ++i) {
        state = state * 6364136223846793005ull + 1ull;
        double re = (state & 0x100) == 0 ? scale : -scale;
        double im = (state & 0x200) == 0 ? scale : -scale;
        om.push_back(cx(re, im));
    }
    return om;
}

uint64_t mix_seed(size_t rows, size_t cols, size_t ell) {
    return ((uint64_t)rows * 0x9E37ull + (uint64_t)cols) * 0x9E37ull + (uint64_t)ell;
}

// matmul (src/tn/mps.rs:1055-1079): row-major (m x k) * (k x n).
std::vector<C> matmul(size_t m, size_t k, size_t n, const std::vector<C>& a,
                      const std::vector<C>& b) {
    std::vector<C> out(m * n, ZERO);
    for (size_t i = 0; i < m; ++i) {
        C* o = out.data() + i * n;
        for (size_t p = 0; p < k; ++p) {
            C av = a[i * k + p];
            const C* br = b.data() + p * n;
            for (size_t j = 0; j < n; ++j) o[j] += av * br[j];
        }
    }
    return out;
}

// adjoint (src/tn/mps.rs:1081-1089): row-major (rows x cols) -> (cols x rows).
std::vector<C> adjoint_rm(const std::vector<C>& a, size_t rows, size_t cols) {
    std::vector<C> out(cols * rows);
    for (size_t r = 0; r < rows; ++r)
        for (size_t c = 0; c < cols; ++c) out[c * rows + r] = conj(a[r * cols + c]);
    return out;
}

// Rank threshold of thin_q. 0 (default) is the literal floating-point reference: a
// reflector is skipped only for an exactly zero residual column, so the columns of Y
// that are dependent in exact arithmetic get reflectors built from rounding noise. A
// positive value tau skips the reflector when the residual norm is <= tau times the
// column's original norm, which is what the reference computes in exact arithmetic
// (a zero residual gives an identity reflector).
double g_qr_rank_tol = 0.0;

// thin_q (src/tn/mps.rs:1091-1101): Householder QR, Q is rows x ell row-major. When vs_out is
// given it receives the unit reflector vectors (vs[i] acts on rows i..rows-1; empty where the
// column took no reflector).
std::vector<C> thin_q(const std::vector<C>& y, size_t rows, size_t ell,
                      std::vector<std::vector<C>>* vs_out = nullptr) {
    Mat a(rows, ell);
    for (size_t r = 0; r < rows; ++r)
        for (size_t c = 0; c < ell; ++c) a(r, c) = y[r * ell + c];
    std::vector<double> cn(ell, 0.0);
    for (size_t c = 0; c < ell; ++c) {
        for (size_t r = 0; r < rows; ++r) cn[c] += norm_sqr(a(r, c));
        cn[c] = std::sqrt(cn[c]);
    }
    std::vector<std::vector<C>> vs(ell);
    for (size_t i = 0; i < ell && i < rows; ++i) {
        C* x = a.col(i);
        double alpha = 0.0;
        for (size_t r = i; r < rows; ++r) alpha += norm_sqr(x[r]);
        alpha = std::sqrt(alpha);
        std::vector<C> v(rows - i, ZERO);
        if (alpha == 0.0 || alpha <= g_qr_rank_tol * cn[i]) continue;
        double x0n = cnorm(x[i]);
        C ph = x0n > 0.0 ? x[i] / x0n : ONE;
        for (size_t r = i; r < rows; ++r) v[r - i] = x[r];
        v[0] += ph * alpha;
        double vn = 0.0;
        for (const C& e : v) vn += norm_sqr(e);
        vn = std::sqrt(vn);
        if (vn == 0.0) continue;
        for (C& e : v) e = e / vn;
        for (size_t c = i; c < ell; ++c) {
            C* col = a.col(c);
            C dot = ZERO;
            for (size_t r = i; r < rows; ++r) dot += conj(v[r - i]) * col[r];
            dot = dot * 2.0;
            for (size_t r = i; r < rows; ++r) col[r] = col[r] - v[r - i] * dot;
        }
        vs[i] = std::move(v);
    }
    Mat q(rows, ell);
    for (size_t i = 0; i < ell && i < rows; ++i) q(i, i) = ONE;
    for (size_t ii = std::min(ell, rows); ii-- > 0;) {
        const std::vector<C>& v = vs[ii];
        if (v.empty()) continue;
        for (size_t c = 0; c < ell; ++c) {
            C* col = q.col(c);
            C dot = ZERO;
            for (size_t r = ii; r < rows; ++r) dot += conj(v[r - ii]) * col[r];
            dot = dot * 2.0;
            for (size_t r = ii; r < rows; ++r) col[r] = col[r] - v[r - ii] * dot;
        }
    }
    std::vector<C> out(rows * ell);
    for (size_t r = 0; r < rows; ++r)
        for (size_t c = 0; c < ell; ++c) out[r * ell + c] = q(r, c);
    if (vs_out) *vs_out = std::move(vs);
    return out;
}

// rsvd_sites (src/tn/mps.rs:983-1053). Returns false for the reference's None.
bool rsvd_sites(const std::vector<C>& mat, size_t rows, size_t cols, size_t ell, size_t dl,
                size_t dr, size_t max_bond, Site& left, Site& right, Counters* ctr,
                double* discarded) {
    std::vector<C> omega = rademacher(cols, ell, mix_seed(rows, cols, ell));
    std::vector<C> y0 = matmul(rows, cols, ell, mat, omega);
    std::vector<C> ah = adjoint_rm(mat, rows, cols);
    std::vector<C> z = matmul(cols, rows, ell, ah, y0);
    std::vector<C> y = matmul(rows, cols, ell, mat, z);
    std::vector<C> q = thin_q(y, rows, ell);
    std::vector<C> qh = adjoint_rm(q, rows, ell);
    std::vector<C> b = matmul(ell, rows, cols, qh, mat);
    Mat bmat(ell, cols);
    for (size_t r = 0; r < ell; ++r)
        for (size_t c = 0; c < cols; ++c) bmat(r, c) = b[r * cols + c];
    Svd svd = checked_svd(bmat, ctr);
    if (!svd.ok) return false;
    std::vector<double> values = svd.s;
    for (double& s : values)
        if (s < SINGULAR_CUTOFF * 1e-2) s = 0.0;
    size_t chi = retained_chi(values, max_bond, rows, cols);
    if (chi == 0 || needs_f64_fallback(values, chi)) {
        if (ctr) ctr->rsvd_fallbacks++;
        return false;
    }
    if (ctr) {
        ctr->rsvd_kept++;
        for (size_t i = chi; i < values.size(); ++i)
            if (values[i] > SINGULAR_CUTOFF) {
                ctr->truncated++;
                break;
            }
    }
    if (discarded) {
        double tot = 0, drop = 0;
        for (size_t i = 0; i < values.size(); ++i) {
            tot += values[i] * values[i];
            if (i >= chi) drop += values[i] * values[i];
        }
        *discarded = tot > 0 ? drop / tot : 0.0;
    }
    left.l = dl;
    left.r = chi;
    left.d.assign(dl * 2 * chi, ZERO);
    right.l = chi;
    right.r = dr;
    right.d.assign(chi * 2 * dr, ZERO);
    for (size_t mode = 0; mode < chi; ++mode) {
        double scale = std::sqrt(svd.s[mode]);
        for (size_t r = 0; r < rows; ++r) {
            C acc = ZERO;
            for (size_t t = 0; t < ell; ++t) acc += q[r * ell + t] * svd.u(t, mode);
            left.at(r / 2, r % 2, mode) = acc * scale;
        }
        for (size_t c = 0; c < cols; ++c) right.at(mode, c / dr, c % dr) = svd.v_t(mode, c) * scale;
    }
    return true;
}

const size_t RSVD_MIN_SIDE = 160;
const size_t RSVD_OVERSAMPLE = 16;

// rsvd_worth (src/tn/mps.rs:929-935).
bool rsvd_worth(size_t rows, size_t cols, size_t ell) {
    unsigned __int128 f = (unsigned __int128)rows * cols * ell;
    unsigned __int128 side = RSVD_MIN_SIDE;
    return f >= side * side * side;
}

// split_theta (src/tn/mps.rs:937-981).
bool split_theta(const Theta& th, size_t max_bond, bool reject, Site& left, Site& right,
                 Counters* ctr, SplitError& err, double* discarded, int* path) {
    if (path) *path = 0;
    if (reject) return svd_split_two_site(th, max_bond, true, left, right, ctr, err, discarded);
    size_t dl = th.dl, dr = th.dr, rows = dl * 2, cols = 2 * dr;
    size_t ell = std::min(std::min(max_bond + RSVD_OVERSAMPLE, rows), cols);
    if (ell == 0 || ell >= std::min(rows, cols) || !rsvd_worth(rows, cols, ell))
        return svd_split_two_site(th, max_bond, false, left, right, ctr, err, discarded);
    if (path) *path = 1;
    if (rsvd_sites(th.d, rows, cols, ell, dl, dr, max_bond, left, right, ctr, discarded))
        return true;
    if (path) *path = 2;
    return svd_split_two_site(th, max_bond, false, left, right, ctr, err, discarded);
}

// ---------------------------------------------------------------------------------
// MpsState (src/tn/mps.rs:55-781)
// ---------------------------------------------------------------------------------
struct MpsState {
    std::vector<Site> sites;
    size_t n = 0;
    size_t max_bond = 0;
    bool reject_truncation = false;
    Counters ctr;
    double max_discarded = 0.0;
    std::vector<int> split_paths;  // per unit: 0 exact, 1 rsvd kept, 2 rsvd fallback
    std::vector<size_t> split_chis;
    bool record = false;

    // try_new_with_bond (src/tn/mps.rs:82-126), host-RAM check skipped as under cfg(test).
    bool init(size_t qn, size_t chi) {
        if (qn == 0) {
            g_err = "invalid qubit count 0";
            return false;
        }
        if (chi == 0) {
            g_err = "MPS max bond dimension must be > 0";
            return false;
        }
        n = qn;
        max_bond = chi;
        sites.assign(n, Site{});
        for (auto& s : sites) {
            s.l = 1;
            s.r = 1;
            s.d = {ONE, ZERO};
        }
        return true;
    }

    // apply_one_qubit (src/tn/mps.rs:333-351).
    void apply_one_qubit(size_t q, const Mat2& u) {
        Site& s = sites[q];
        Site up;
        up.l = s.l;
        up.r = s.r;
        up.d.assign(s.d.size(), ZERO);
        for (size_t a = 0; a < s.l; ++a)
            for (size_t b = 0; b < s.r; ++b) {
                C v0 = s.at(a, 0, b), v1 = s.at(a, 1, b);
                up.at(a, 0, b) = u.m[0][0] * v0 + u.m[0][1] * v1;
                up.at(a, 1, b) = u.m[1][0] * v0 + u.m[1][1] * v1;
            }
        s = std::move(up);
    }

    void record_split(int path, size_t chi, double disc) {
        if (disc > max_discarded) max_discarded = disc;
        if (record) {
            split_paths.push_back(path);
            split_chis.push_back(chi);
        }
    }

    // apply_two_qubit_at_positions (src/tn/mps.rs:377-429).
    int apply_two_qubit_at_positions(size_t p, size_t q, const Mat4& u, bool ctrl_is_first) {
        ctr.units++;
        Theta th = cpu_merge(sites[p], sites[q]);
        Theta up = apply_two_qubit_unitary(th, u, ctrl_is_first);
        Site l, r;
        SplitError err;
        double disc = 0.0;
        int path = 0;
        if (!split_theta(up, max_bond, reject_truncation, l, r, &ctr, err, &disc, &path)) {
            g_err = err.msg;
            return err.code;
        }
        record_split(path, l.r, disc);
        sites[p] = std::move(l);
        sites[q] = std::move(r);
        return 0;
    }

    int apply_swap_positions(size_t p, size_t q) {
        return apply_two_qubit_at_positions(p, q, swap_matrix(), true);
    }

    // apply_two_qubit_gate (src/tn/mps.rs:353-375).
    int apply_two_qubit_gate(size_t ctrl, size_t tgt, const Mat4& u) {
        size_t left = std::min(ctrl, tgt), right = std::max(ctrl, tgt);
        size_t span = right - left;
        int rc;
        if (span > 1)
            for (size_t pos = right - 1; pos >= left + 1; --pos)
                if ((rc = apply_swap_positions(pos, pos + 1))) return rc;
        if ((rc = apply_two_qubit_at_positions(left, left + 1, u, ctrl < tgt))) return rc;
        if (span > 1)
            for (size_t pos = left + 1; pos < right; ++pos)
                if ((rc = apply_swap_positions(pos, pos + 1))) return rc;
        return 0;
    }

    // apply_ccnot (src/tn/mps.rs:657-677).
    int apply_ccnot(size_t c1, size_t c2, size_t t) {
        Mat2 tdg = tdg_matrix();
        OGate gh{GH, 0, 0, 0, 0}, gt{GT, 0, 0, 0, 0};
        Mat4 cn = cnot_matrix();
        int rc;
        apply_one_qubit(t, unary_matrix(gh));
        if ((rc = apply_two_qubit_gate(c2, t, cn))) return rc;
        apply_one_qubit(t, tdg);
        if ((rc = apply_two_qubit_gate(c1, t, cn))) return rc;
        apply_one_qubit(t, unary_matrix(gt));
        if ((rc = apply_two_qubit_gate(c2, t, cn))) return rc;
        apply_one_qubit(t, tdg);
        if ((rc = apply_two_qubit_gate(c1, t, cn))) return rc;
        apply_one_qubit(t, unary_matrix(gt));
        apply_one_qubit(t, unary_matrix(gh));
        apply_one_qubit(c2, unary_matrix(gt));
        if ((rc = apply_two_qubit_gate(c1, c2, cn))) return rc;
        apply_one_qubit(c1, unary_matrix(gt));
        apply_one_qubit(c2, tdg);
        if ((rc = apply_two_qubit_gate(c1, c2, cn))) return rc;
        return 0;
    }

    bool check_gate(const OGate& g) {
        auto bad = [&](int q) { return q < 0 || (size_t)q >= n; };
        switch (g.kind) {
            case GH: case GX: case GY: case GZ: case GS: case GT: case GRX: case GRY: case GRZ:
            case GMEASURE: case GRESET:
                if (bad(g.q0)) { g_err = "qubit index out of bounds"; return false; }
                return true;
            case GCNOT: case GCZ:
                if (bad(g.q0) || bad(g.q1) || g.q0 == g.q1) { g_err = "bad two-qubit gate"; return false; }
                return true;
            case GCCNOT:
                if (bad(g.q0) || bad(g.q1) || bad(g.q2) || g.q0 == g.q1 || g.q0 == g.q2 || g.q1 == g.q2) {
                    g_err = "bad CCNOT";
                    return false;
                }
                return true;
            case GIF:
                return true;
            default:
                g_err = "unknown gate kind";
                return false;
        }
    }

    // apply_gate (src/tn/mps.rs:278-297).
    int apply_gate(const OGate& g) {
        if (!check_gate(g)) return 1;
        switch (g.kind) {
            case GH: case GX: case GY: case GZ: case GS: case GT: case GRX: case GRY: case GRZ:
                apply_one_qubit(g.q0, unary_matrix(g));
                return 0;
            case GCNOT: case GCZ:
                return apply_two_qubit_gate(g.q0, g.q1, two_qubit_matrix(g));
            case GCCNOT:
                return apply_ccnot(g.q0, g.q1, g.q2);
            default:
                return 0;
        }
    }

    // apply_adjacent_disjoint_gates (src/tn/mps.rs:520-565). Jobs are independent, so
    // they run in parallel; results equal the sequential order.
    int apply_adjacent_disjoint_gates(const OGate* gates, size_t count) {
        if (count < 2) {
            for (size_t i = 0; i < count; ++i)
                if (int rc = apply_gate(gates[i])) return rc;
            return 0;
        }
        struct Job {
            size_t p, q;
            bool cf;
            Mat4 u;
            Site l, r;
            int rc;
            std::string msg;
            int path;
            double disc;
            Counters c;
        };
        std::vector<Job> jobs(count);
        for (size_t i = 0; i < count; ++i) {
            const OGate& g = gates[i];
            if (g.kind != GCNOT && g.kind != GCZ) {
                g_err = "disjoint layer contains a gate that is not CNOT or CZ";
                return 3;
            }
            if (!check_gate(g)) return 1;
            size_t c = g.q0, t = g.q1;
            Job& j = jobs[i];
            j.p = std::min(c, t);
            j.q = std::max(c, t);
            j.cf = c < t;
            if (j.q != j.p + 1) {
                g_err = "disjoint layer gate is not adjacent";
                return 1;
            }
            j.u = two_qubit_matrix(g);
        }
#pragma omp parallel for schedule(dynamic, 1)
        for (size_t i = 0; i < count; ++i) {
            Job& j = jobs[i];
            Theta th = cpu_merge(sites[j.p], sites[j.q]);
            Theta up = apply_two_qubit_unitary(th, j.u, j.cf);
            SplitError err;
            j.disc = 0.0;
            bool ok = split_theta(up, max_bond, reject_truncation, j.l, j.r, &j.c, err, &j.disc,
                                  &j.path);
            j.rc = ok ? 0 : err.code;
            j.msg = err.msg;
        }
        for (size_t i = 0; i < count; ++i) {
            Job& j = jobs[i];
            ctr.units++;
            ctr.exact += j.c.exact;
            ctr.rsvd_kept += j.c.rsvd_kept;
            ctr.rsvd_fallbacks += j.c.rsvd_fallbacks;
            ctr.svd_failures += j.c.svd_failures;
            ctr.truncated += j.c.truncated;
            if (j.rc) {
                g_err = j.msg;
                return j.rc;
            }
            record_split(j.path, j.l.r, j.disc);
            sites[j.p] = std::move(j.l);
            sites[j.q] = std::move(j.r);
        }
        return 0;
    }

    // The apply loop of execute_with_trace_using (src/tn/trace.rs:71-161).
    int apply_circuit(const OGate* gates, size_t count) {
        size_t i = 0;
        while (i < count) {
            const OGate& g = gates[i];
            if (g.kind == GMEASURE || g.kind == GRESET || g.kind == GIF) {
                ++i;
                continue;
            }
            if (g.kind == GH) {  // consecutive_h_run (trace.rs:189-202)
                size_t run = 1;
                while (i + run < count && gates[i + run].kind == GH && gates[i + run].q0 == g.q0) run++;
                for (size_t k = 0; k < run; ++k)
                    if (int rc = apply_gate(gates[i + k])) return rc;
                i += run;
                continue;
            }
            if (g.kind == GRX || g.kind == GRY || g.kind == GRZ) {  // trace.rs:204-233
                size_t run = 1;
                double sum = g.angle;
                while (i + run < count && gates[i + run].kind == g.kind && gates[i + run].q0 == g.q0) {
                    sum += gates[i + run].angle;
                    run++;
                }
                const double tau = 6.283185307179586;
                double red = std::fmod(sum, tau);
                if (red < 0) red += tau;
                bool identity = std::fabs(red) < 1e-9 || std::fabs(tau - red) < 1e-9;
                if (identity) {
                    for (size_t k = 0; k < run; ++k)
                        if (int rc = apply_gate(gates[i + k])) return rc;
                } else {
                    OGate net = g;
                    net.angle = sum;
                    if (int rc = apply_gate(net)) return rc;
                }
                i += run;
                continue;
            }
            size_t layer = 0;
            if (n > 16) {  // adjacent_disjoint_len (trace.rs:247-268)
                std::vector<int> seen;
                size_t k = 0;
                for (; i + k < count; ++k) {
                    const OGate& h = gates[i + k];
                    if (h.kind != GCNOT && h.kind != GCZ) break;
                    int lo = std::min(h.q0, h.q1), hi = std::max(h.q0, h.q1);
                    if (hi - lo != 1) break;
                    if (std::find(seen.begin(), seen.end(), lo) != seen.end() ||
                        std::find(seen.begin(), seen.end(), hi) != seen.end())
                        break;
                    seen.push_back(lo);
                    seen.push_back(hi);
                }
                layer = k >= 2 ? k : 0;
            }
            if (layer >= 2) {
                if (int rc = apply_adjacent_disjoint_gates(gates + i, layer)) return rc;
                i += layer;
                continue;
            }
            if (int rc = apply_gate(g)) return rc;
            ++i;
        }
        return 0;
    }

    // amplitude_at_compact_zero (src/tn/mps.rs:262-276).
    C amplitude_zero() const {
        std::vector<C> v{ONE};
        for (const Site& s : sites) {
            std::vector<C> next(s.r, ZERO);
            size_t lim = std::min(s.l, v.size());
            for (size_t a = 0; a < lim; ++a)
                for (size_t b = 0; b < s.r; ++b) next[b] += v[a] * s.at(a, 0, b);
            v = std::move(next);
        }
        return v.empty() ? ZERO : v[0];
    }

    // contract_to_statevector (src/tn/mps.rs:215-260).
    int statevector(std::vector<C>& out) const {
        if (n > 20) {
            g_err = "full statevector projection requires qubit_count <= 20";
            return 3;
        }
        std::vector<C> st{ONE};
        for (size_t qi = 0; qi < n; ++qi) {
            const Site& s = sites[qi];
            size_t prefix = (size_t)1 << qi;
            if (st.size() != s.l * prefix) {
                g_err = "MPS bond layout mismatch during statevector projection";
                return 3;
            }
            std::vector<C> next(s.r * prefix * 2, ZERO);
            for (size_t a = 0; a < s.l; ++a)
                for (size_t bp = 0; bp < prefix; ++bp) {
                    C v = st[a * prefix + bp];
                    for (size_t ph = 0; ph < 2; ++ph)
                        for (size_t b = 0; b < s.r; ++b)
                            next[b * (prefix * 2) + bp + ph * prefix] += v * s.at(a, ph, b);
                }
            st = std::move(next);
        }
        out = std::move(st);
        return 0;
    }

    // to_dense_state / fill_dense_recursive (src/tn/mps.rs:742-780).
    void fill(size_t si, size_t bond_in, size_t prefix, C coeff, std::vector<C>& st) const {
        if (si == n) {
            st[prefix] += coeff;
            return;
        }
        const Site& s = sites[si];
        if (bond_in >= s.l) return;
        for (size_t ph = 0; ph < 2; ++ph)
            for (size_t b = 0; b < s.r; ++b) {
                C amp = s.at(bond_in, ph, b);
                if (norm_sqr(amp) < 1e-30) continue;
                fill(si + 1, b, prefix | (ph << si), coeff * amp, st);
            }
    }

    // left_environment / right_environment / site_amplitudes (src/tn/mps.rs:679-740).
    std::vector<C> left_env(size_t target) const {
        std::vector<C> v{ONE};
        for (size_t i = 0; i < target; ++i) {
            const Site& s = sites[i];
            std::vector<C> next(s.r, ZERO);
            size_t lim = std::min(s.l, v.size());
            for (size_t a = 0; a < lim; ++a)
                for (size_t ph = 0; ph < 2; ++ph)
                    for (size_t b = 0; b < s.r; ++b) next[b] += v[a] * s.at(a, ph, b);
            v = std::move(next);
        }
        return v;
    }
    std::vector<C> right_env(size_t target) const {
        std::vector<C> v{ONE};
        if (target + 1 >= n) return v;
        for (size_t i = n; i-- > target + 1;) {
            const Site& s = sites[i];
            std::vector<C> next(s.l, ZERO);
            size_t lim = std::min(s.r, v.size());
            for (size_t b = 0; b < lim; ++b)
                for (size_t ph = 0; ph < 2; ++ph)
                    for (size_t a = 0; a < s.l; ++a) next[a] += v[b] * s.at(a, ph, b);
            v = std::move(next);
        }
        return v;
    }
    void site_amplitudes(size_t t, C& a0, C& a1) const {
        std::vector<C> le = left_env(t), re = right_env(t);
        const Site& s = sites[t];
        a0 = ZERO;
        a1 = ZERO;
        for (size_t l = 0; l < le.size(); ++l)
            for (size_t r = 0; r < s.r; ++r) {
                C rv = r < re.size() ? re[r] : ONE;
                a0 += le[l] * s.at(l, 0, r) * rv;
                a1 += le[l] * s.at(l, 1, r) * rv;
            }
    }
};

// ---------------------------------------------------------------------------------
// sampling: rand 0.8.6 StdRng (= rand_chacha 0.3.1 ChaCha12Rng), seeded by
// rand_core 0.6 seed_from_u64 (PCG32 expansion); gen::<f64>() = (u64 >> 11) * 2^-53.
// ---------------------------------------------------------------------------------
inline uint32_t rotl(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

void chacha_block(const uint32_t key[8], uint64_t counter, uint64_t nonce, int rounds,
                  uint32_t out[16]) {
    uint32_t s[16] = {0x61707865, 0x3320646e, 0x79622d32, 0x6b206574, key[0], key[1], key[2],
                      key[3], key[4], key[5], key[6], key[7], (uint32_t)counter,
                      (uint32_t)(counter >> 32), (uint32_t)nonce, (uint32_t)(nonce >> 32)};
    uint32_t x[16];
    std::memcpy(x, s, sizeof x);
#define QR(a, b, c, d)             \
    x[a] += x[b];                  \
    x[d] = rotl(x[d] ^ x[a], 16);  \
    x[c] += x[d];                  \
    x[b] = rotl(x[b] ^ x[c], 12);  \
    x[a] += x[b];                  \
    x[d] = rotl(x[d] ^ x[a], 8);   \
    x[c] += x[d];                  \
    x[b] = rotl(x[b] ^ x[c], 7);
    for (int i = 0; i < rounds; i += 2) {
        QR(0, 4, 8, 12) QR(1, 5, 9, 13) QR(2, 6, 10, 14) QR(3, 7, 11, 15)
        QR(0, 5, 10, 15) QR(1, 6, 11, 12) QR(2, 7, 8, 13) QR(3, 4, 9, 14)
    }
#undef QR
    for (int i = 0; i < 16; ++i) out[i] = x[i] + s[i];
}

struct StdRng {
    uint32_t key[8];
    uint64_t counter = 0;
    uint32_t buf[64];  // rand_chacha fills 4 blocks per refill
    size_t index = 64;

    explicit StdRng(uint64_t state) {
        // rand_core 0.6 SeedableRng::seed_from_u64.
        const uint64_t MUL = 6364136223846793005ull, INC = 11634580027462260723ull;
        for (int i = 0; i < 8; ++i) {
            state = state * MUL + INC;
            uint32_t xorshifted = (uint32_t)(((state >> 18) ^ state) >> 27);
            uint32_t rot = (uint32_t)(state >> 59);
            key[i] = (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
        }
    }
    void refill() {
        for (int b = 0; b < 4; ++b) chacha_block(key, counter + b, 0, 12, buf + 16 * b);
        counter += 4;
        index = 0;
    }
    // rand_core BlockRng::next_u64.
    uint64_t next_u64() {
        if (index < 63) {
            uint64_t lo = buf[index], hi = buf[index + 1];
            index += 2;
            return (hi << 32) | lo;
        } else if (index >= 64) {
            refill();
            uint64_t lo = buf[0], hi = buf[1];
            index = 2;
            return (hi << 32) | lo;
        } else {
            uint64_t x = buf[63];
            refill();
            uint64_t y = buf[0];
            index = 1;
            return (y << 32) | x;
        }
    }
    double gen_f64() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }
};

inline double snap(double v) { return std::round(v * 1e6) / 1e6; }

}  // namespace

// =====================================================================================
// C API (ctypes)
// =====================================================================================
extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }

int oracle_num_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

void* oracle_new(size_t n, size_t chi) {
    auto* s = new MpsState();
    if (!s->init(n, chi)) {
        delete s;
        return nullptr;
    }
    return s;
}

void oracle_free(void* p) { delete static_cast<MpsState*>(p); }

void oracle_set_reject_truncation(void* p, int on) {
    static_cast<MpsState*>(p)->reject_truncation = on != 0;
}

void oracle_set_max_bond(void* p, size_t chi) { static_cast<MpsState*>(p)->max_bond = chi; }

void oracle_set_record(void* p, int on) { static_cast<MpsState*>(p)->record = on != 0; }

int oracle_apply_gate(void* p, const OGate* g) { return static_cast<MpsState*>(p)->apply_gate(*g); }

int oracle_apply_gates(void* p, const OGate* g, size_t count) {
    auto* s = static_cast<MpsState*>(p);
    for (size_t i = 0; i < count; ++i)
        if (int rc = s->apply_gate(g[i])) return rc;
    return 0;
}

int oracle_apply_circuit(void* p, const OGate* g, size_t count) {
    return static_cast<MpsState*>(p)->apply_circuit(g, count);
}

int oracle_apply_disjoint(void* p, const OGate* g, size_t count) {
    return static_cast<MpsState*>(p)->apply_adjacent_disjoint_gates(g, count);
}

int oracle_apply_unary_at(void* p, size_t site, const OGate* g) {
    auto* s = static_cast<MpsState*>(p);
    if (site >= s->n) return 1;
    s->apply_one_qubit(site, unary_matrix(*g));
    return 0;
}

int oracle_apply_pair_at(void* p, size_t site, int kind, int ctrl_on_left) {
    auto* s = static_cast<MpsState*>(p);
    if (site + 1 >= s->n) return 1;
    OGate g{kind, 0, 0, 0, 0};
    Mat4 u = kind == 0 ? swap_matrix() : two_qubit_matrix(g);
    return s->apply_two_qubit_at_positions(site, site + 1, u, ctrl_on_left != 0);
}

int oracle_amplitude_zero(void* p, double* out) {
    C a = static_cast<MpsState*>(p)->amplitude_zero();
    out[0] = a.re;
    out[1] = a.im;
    return 0;
}

size_t oracle_qubits(void* p) { return static_cast<MpsState*>(p)->n; }

int oracle_bond_dims(void* p, size_t* out, size_t len) {
    auto* s = static_cast<MpsState*>(p);
    for (size_t i = 0; i < s->n && i < len; ++i) out[i] = s->sites[i].r;
    return 0;
}

int oracle_get_site(void* p, size_t i, double* out, size_t len, size_t* l, size_t* r) {
    auto* s = static_cast<MpsState*>(p);
    if (i >= s->n) return 1;
    const Site& st = s->sites[i];
    *l = st.l;
    *r = st.r;
    if (out && len >= st.d.size()) std::memcpy(out, st.d.data(), st.d.size() * sizeof(C));
    return 0;
}

int oracle_set_site(void* p, size_t i, const double* data, size_t l, size_t r) {
    auto* s = static_cast<MpsState*>(p);
    if (i >= s->n) return 1;
    Site st;
    st.l = l;
    st.r = r;
    st.d.resize(l * 2 * r);
    std::memcpy(st.d.data(), data, st.d.size() * sizeof(C));
    s->sites[i] = std::move(st);
    return 0;
}

int oracle_statevector(void* p, double* out, size_t len) {
    std::vector<C> sv;
    int rc = static_cast<MpsState*>(p)->statevector(sv);
    if (rc) return rc;
    if (len < sv.size()) {
        g_err = "output too small";
        return 1;
    }
    std::memcpy(out, sv.data(), sv.size() * sizeof(C));
    return 0;
}

int oracle_dense_state(void* p, double* out, size_t len) {
    auto* s = static_cast<MpsState*>(p);
    if (s->n > 26) return 1;
    size_t dim = (size_t)1 << s->n;
    if (len < dim) return 1;
    std::vector<C> st(dim, ZERO);
    s->fill(0, 0, 0, ONE, st);
    std::memcpy(out, st.data(), dim * sizeof(C));
    return 0;
}

int oracle_site_amplitudes(void* p, size_t t, double* out) {
    auto* s = static_cast<MpsState*>(p);
    if (t >= s->n) return 1;
    C a0, a1;
    s->site_amplitudes(t, a0, a1);
    out[0] = a0.re;
    out[1] = a0.im;
    out[2] = a1.re;
    out[3] = a1.im;
    return 0;
}

// counters: units, exact, rsvd_kept, rsvd_fallbacks, svd_failures, truncated
int oracle_counters(void* p, uint64_t* out, size_t len) {
    auto* s = static_cast<MpsState*>(p);
    uint64_t v[6] = {s->ctr.units, s->ctr.exact, s->ctr.rsvd_kept, s->ctr.rsvd_fallbacks,
                     s->ctr.svd_failures, s->ctr.truncated};
    for (size_t i = 0; i < 6 && i < len; ++i) out[i] = v[i];
    return 0;
}

double oracle_max_discarded(void* p) { return static_cast<MpsState*>(p)->max_discarded; }

size_t oracle_split_log(void* p, int32_t* paths, size_t* chis, size_t len) {
    auto* s = static_cast<MpsState*>(p);
    size_t m = s->split_paths.size();
    for (size_t i = 0; i < m && i < len; ++i) {
        paths[i] = s->split_paths[i];
        chis[i] = s->split_chis[i];
    }
    return m;
}

// split_theta on an explicit theta (dl,2,2,dr). left/right receive (dl,2,chi), (chi,2,dr);
// capacity is dl*2*min(2dl,2dr) and min(2dl,2dr)*2*dr complex. Returns status;
// *chi_out the kept bond; *path_out 0 exact, 1 rsvd, 2 rsvd fallback to exact.
int oracle_split_theta(const double* theta, size_t dl, size_t dr, size_t max_bond, int reject,
                       double* left, double* right, size_t* chi_out, int* path_out,
                       double* discarded) {
    Theta th;
    th.dl = dl;
    th.dr = dr;
    th.d.resize(dl * 4 * dr);
    std::memcpy(th.d.data(), theta, th.d.size() * sizeof(C));
    Site l, r;
    SplitError err;
    Counters ctr;
    double disc = 0.0;
    if (!split_theta(th, max_bond, reject != 0, l, r, &ctr, err, &disc, path_out)) {
        g_err = err.msg;
        return err.code;
    }
    *chi_out = l.r;
    if (discarded) *discarded = disc;
    std::memcpy(left, l.d.data(), l.d.size() * sizeof(C));
    std::memcpy(right, r.d.data(), r.d.size() * sizeof(C));
    return 0;
}

// rsvd_sites on a row-major rows x cols matrix. Returns 1 for Some, 0 for None.
int oracle_rsvd_sites(const double* mat, size_t rows, size_t cols, size_t ell, size_t dl,
                      size_t dr, size_t max_bond, double* left, double* right, size_t* chi_out) {
    std::vector<C> m(rows * cols);
    std::memcpy(m.data(), mat, m.size() * sizeof(C));
    Site l, r;
    if (!rsvd_sites(m, rows, cols, ell, dl, dr, max_bond, l, r, nullptr, nullptr)) return 0;
    *chi_out = l.r;
    std::memcpy(left, l.d.data(), l.d.size() * sizeof(C));
    std::memcpy(right, r.d.data(), r.d.size() * sizeof(C));
    return 1;
}

// checked_svd of a row-major rows x cols matrix: u (rows x k), s (k), vt (k x cols),
// k = min(rows, cols), all row-major. Returns 0 on success.
int oracle_checked_svd(const double* a, size_t rows, size_t cols, double* u, double* s,
                       double* vt) {
    Mat m(rows, cols);
    const C* src = reinterpret_cast<const C*>(a);
    for (size_t r = 0; r < rows; ++r)
        for (size_t c = 0; c < cols; ++c) m(r, c) = src[r * cols + c];
    Svd svd = checked_svd(m, nullptr);
    if (!svd.ok) {
        g_err = "SVD failed its check";
        return 3;
    }
    size_t k = std::min(rows, cols);
    C* uo = reinterpret_cast<C*>(u);
    C* vo = reinterpret_cast<C*>(vt);
    for (size_t r = 0; r < rows; ++r)
        for (size_t c = 0; c < k; ++c) uo[r * k + c] = svd.u(r, c);
    for (size_t r = 0; r < k; ++r)
        for (size_t c = 0; c < cols; ++c) vo[r * cols + c] = svd.v_t(r, c);
    for (size_t i = 0; i < k; ++i) s[i] = svd.s[i];
    return 0;
}

size_t oracle_retained_chi(const double* v, size_t len, size_t max_bond, size_t rows, size_t cols) {
    return retained_chi(std::vector<double>(v, v + len), max_bond, rows, cols);
}

int oracle_needs_f64_fallback(const double* v, size_t len, size_t chi) {
    return needs_f64_fallback(std::vector<double>(v, v + len), chi) ? 1 : 0;
}

uint64_t oracle_mix_seed(size_t rows, size_t cols, size_t ell) { return mix_seed(rows, cols, ell); }

// 0 = literal floating-point thin_q (default); tau > 0 = exact-arithmetic rank rule.
void oracle_set_qr_rank_tol(double tau) { g_qr_rank_tol = tau; }

void oracle_rademacher(size_t cols, size_t ell, uint64_t salt, double* out) {
    std::vector<C> om = rademacher(cols, ell, salt);
    std::memcpy(out, om.data(), om.size() * sizeof(C));
}

int oracle_thin_q(const double* y, size_t rows, size_t ell, double* q) {
    std::vector<C> m(rows * ell);
    std::memcpy(m.data(), y, m.size() * sizeof(C));
    std::vector<C> out = thin_q(m, rows, ell);
    std::memcpy(q, out.data(), out.size() * sizeof(C));
    return 0;
}

// thin_q at rank threshold tau with its reflectors exported: q (rows x ell), v (rows x ell
// row-major, column i = reflector i, zero above row i and for skipped columns), mask[i] = 1
// where column i took a reflector.
int oracle_thin_q_reflectors(const double* y, size_t rows, size_t ell, double tau, double* q, double* v,
                             int* mask) {
    const double saved = g_qr_rank_tol;
    g_qr_rank_tol = tau;
    std::vector<C> m(rows * ell);
    std::memcpy(m.data(), y, m.size() * sizeof(C));
    std::vector<std::vector<C>> vs;
    std::vector<C> out = thin_q(m, rows, ell, &vs);
    g_qr_rank_tol = saved;
    std::memcpy(q, out.data(), out.size() * sizeof(C));
    C* vo = reinterpret_cast<C*>(v);
    for (size_t i = 0; i < rows * ell; ++i) vo[i] = ZERO;
    for (size_t i = 0; i < ell; ++i) {
        mask[i] = (i < vs.size() && !vs[i].empty()) ? 1 : 0;
        if (!mask[i]) continue;
        for (size_t r = i; r < rows; ++r) vo[r * ell + i] = vs[i][r - i];
    }
    return 0;
}

// ⟨ψ|P|ψ⟩ (src/expectation.rs:98-167): Qiskit order, rightmost char = qubit 0.
int oracle_pauli_expectation(const double* sv, size_t n, const char* label, double* out) {
    size_t len = std::strlen(label);
    if (len == 0 || len != n) {
        g_err = "bad pauli label";
        return 1;
    }
    const C* psi = reinterpret_cast<const C*>(sv);
    size_t dim = (size_t)1 << n;
    C sum = ZERO;
    for (size_t i = 0; i < dim; ++i) {
        size_t j = i;
        C phase = ONE;
        for (size_t q = 0; q < n; ++q) {
            char ch = label[n - 1 - q];
            size_t bit = (i >> q) & 1;
            if (ch == 'I') {
            } else if (ch == 'Z') {
                if (bit == 1) phase = phase * -1.0;
            } else if (ch == 'X') {
                j ^= (size_t)1 << q;
            } else if (ch == 'Y') {
                j ^= (size_t)1 << q;
                phase = phase * cx(0.0, bit == 0 ? 1.0 : -1.0);
            } else {
                g_err = "invalid pauli character";
                return 1;
            }
        }
        sum += conj(psi[j]) * phase * psi[i];
    }
    out[0] = sum.re;
    out[1] = sum.im;
    return 0;
}

// compute_outcome_probabilities + sample_counts_from_probabilities
// (src/sample.rs:85-189). counts has 2^cbits entries indexed by the classical register
// value (bit c = cbit c); BTreeMap order over Qiskit keys equals that numeric order.
int oracle_sample_counts(const double* sv, size_t n, const int32_t* mq, const int32_t* mc,
                         size_t nm, size_t cbits, uint64_t shots, uint64_t seed,
                         uint64_t* counts) {
    if (cbits == 0 || cbits > 24) {
        g_err = "classical_bit_count must be in 1..=24";
        return 1;
    }
    const C* psi = reinterpret_cast<const C*>(sv);
    size_t dim = (size_t)1 << n, cdim = (size_t)1 << cbits;
    std::vector<double> probs(cdim, 0.0);
    std::vector<char> present(cdim, 0);
    for (size_t basis = 0; basis < dim; ++basis) {
        double re = snap(psi[basis].re), im = snap(psi[basis].im);
        double pr = re * re + im * im;
        if (pr == 0.0) continue;
        size_t key = 0;
        for (size_t m = 0; m < nm; ++m) {
            size_t bit = (basis >> mq[m]) & 1;
            key = (key & ~((size_t)1 << mc[m])) | (bit << mc[m]);
        }
        probs[key] += pr;
        present[key] = 1;
    }
    for (size_t k = 0; k < cdim; ++k) counts[k] = 0;
    std::vector<std::pair<size_t, double>> cum;
    double total = 0.0;
    for (size_t k = 0; k < cdim; ++k)
        if (present[k]) total += probs[k];
    if (shots == 0 || total == 0.0) return 0;
    double acc = 0.0;
    for (size_t k = 0; k < cdim; ++k)
        if (present[k]) {
            acc += probs[k] / total;
            cum.push_back({k, acc});
        }
    StdRng rng(seed);
    for (uint64_t s = 0; s < shots; ++s) {
        double r = rng.gen_f64();
        size_t label = cum.back().first;
        for (auto& e : cum)
            if (r <= e.second) {
                label = e.first;
                break;
            }
        counts[label]++;
    }
    return 0;
}

void oracle_chacha_block(const uint32_t* key, uint64_t counter, uint64_t nonce, int rounds,
                         uint32_t* out) {
    chacha_block(key, counter, nonce, rounds, out);
}

void oracle_stdrng_f64(uint64_t seed, size_t count, double* out) {
    StdRng rng(seed);
    for (size_t i = 0; i < count; ++i) out[i] = rng.gen_f64();
}

}  // extern "C"
