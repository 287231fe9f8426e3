// qiw_oracle.cpp -- CPU float64 restatement of QInchworm.jl's quasi Monte Carlo
// diagram-evaluation path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle: the checker that the GPU path
// (libqiw) is compared against, and the timed CPU baseline.  It is never linked into,
// loaded by or called from the product.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may use it.
//
// It restates, in complex<double> arithmetic and with the reference's own loop order
// (per-sample depth-first traversal with prefix-cached chain products):
//   src/scrambled_sobol.jl   ScrambledSobolSeq ctor/next!/skip!        (orc_sobol_*)
//   src/qmc_integrate.jl     RootTransform, DoubleSimplexRootTransform,
//                            qmc_integral, contour_integral           (Transform, integrate)
//   src/diagrammatics.jl     generate_topologies, is_doubly_k_connected,
//                            get_topologies_at_order                   (topologies_at_order)
//   src/utility.jl           LazyMatrixProduct, IncrementalSpline     (LMP, IncSpline)
//   src/spline_gf.jl         SplineInterpolatedGF / IncSplineImaginaryTimeGF lookups
//   src/exact_atomic_ppgf.jl ExactAtomicPPGF
//   src/ppgf.jl              atomic_ppgf, set_ppgf!, normalize!, partition_function,
//                            initial_ppgf_derivative, density_matrix
//   src/topology_eval.jl     TopologyEvaluator and _traverse_configuration_tree!
//   src/inchworm.jl          inchworm_step, inchworm_step_bare, inchworm!, correlator_2p
//   src/randomization.jl     mean_std_from_randomization               (mean_std)
//   src/inchworm.jl          diff_inchworm_step!, diff_inchworm!      (diff_step, orc_diff_inchworm)
//
// Parity of each part is pinned by the tests in tests/test_oracle_*.py (golden vectors of
// the reference's test suite, SciPy, closed forms).  Grid-stored Green's functions (the
// bold P, QIW_GF_GRID_LINEAR pairs) are read with Keldysh's cell rule (grid_cell below),
// pinned by test/inchworm.h5 and test/topology_eval.h5.  The Sobol direction numbers come
// from the caller, so the oracle holds no table of its own.
//
// Parity unpinned: the rows of a grid-stored self-energy (interpolate_ppgf = false) read
// by the Σ * P convolution of diff_step (SigmaTable).  They follow set_ppgf! on an
// ImaginaryTimeGF (src/inchworm.jl:612-625), and no reference test covers them; the rows
// of the Σ_orders output of the same run are pinned by the order-1 closed form.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/qiw.h"

typedef std::complex<double> cd;
static const cd I(0.0, 1.0);

static std::string g_err;
static void set_err(const std::string &s) { g_err = s; }

// ------------------------------------------------------------------------------------
// Small dense complex matrices (column-major)
// ------------------------------------------------------------------------------------
struct CMat {
    int r = 0, c = 0;
    std::vector<cd> a;
    CMat() {}
    CMat(int r_, int c_) : r(r_), c(c_), a((size_t)r_ * c_, cd(0)) {}
    cd &operator()(int i, int j) { return a[(size_t)i + (size_t)j * r]; }
    const cd &operator()(int i, int j) const { return a[(size_t)i + (size_t)j * r]; }
    void zero() { std::fill(a.begin(), a.end(), cd(0)); }
};

// C = A * B  (C preallocated with right shape)
static inline void matmul(CMat &C, const CMat &A, const CMat &B) {
    C.r = A.r;
    C.c = B.c;
    C.a.resize((size_t)A.r * B.c);
    for (int j = 0; j < B.c; ++j)
        for (int i = 0; i < A.r; ++i) {
            cd s = 0;
            for (int k = 0; k < A.c; ++k) s += A(i, k) * B(k, j);
            C(i, j) = s;
        }
}

// ------------------------------------------------------------------------------------
// Instrumentation (algorithmic work, real FLOPs as in the work model of the path)
// ------------------------------------------------------------------------------------
static int64_t g_leaves = 0, g_diagram_evals = 0;
static double g_flops = 0;

// ------------------------------------------------------------------------------------
// Sobol sequence -- src/scrambled_sobol.jl:66-208
// ------------------------------------------------------------------------------------
extern "C" int orc_sobol_direction_numbers(int D, const uint32_t *a_tab, const uint32_t *minit_tab,
                                           int minit_stride, const uint8_t *ltm,
                                           const uint8_t *shift_bits, uint32_t *m_out,
                                           uint32_t *x_out) {
    // m = ones(UInt32, (D, 32))                                   scrambled_sobol.jl:69
    std::vector<uint32_t> m((size_t)D * 32, 1u);
    auto M = [&](int d, int j) -> uint32_t & { return m[(size_t)(d - 1) * 32 + (j - 1)]; };
    for (int d = 2; d <= D; ++d) {                                // :78-94
        uint32_t a = a_tab[d - 2];
        int deg = 0;
        while ((a >> (deg + 1)) != 0) ++deg;                      // floor(log2(a))
        for (int j = 1; j <= deg; ++j) M(d, j) = minit_tab[(size_t)(d - 2) * minit_stride + (j - 1)];
        for (int j = deg + 1; j <= 32; ++j) {
            uint32_t ac = a;
            M(d, j) = M(d, j - deg);
            for (int k = 0; k <= deg - 1; ++k) {
                M(d, j) = M(d, j) ^ (((ac & 1u) * M(d, j - deg + k)) << (deg - k));
                ac >>= 1;
            }
        }
    }
    for (int j = 1; j <= 32; ++j)                                 // :98-100
        for (int d = 1; d <= D; ++d) M(d, j) *= (uint32_t)(1ull << (32 - j));
    std::vector<uint32_t> x((size_t)D, 0u);
    if (ltm != nullptr) {                                         // LMS + shift :104-140
        for (int d = 0; d < D; ++d) {
            uint32_t s = 0;
            for (int j = 0; j < 32; ++j) s += (uint32_t)shift_bits[(size_t)d * 32 + j] << j;
            x[d] = s;
        }
        // ltm[d][p][q] (0-based); lower triangular with unit diagonal
        for (int d = 1; d <= D; ++d)
            for (int j = 1; j <= 32; ++j) {
                uint32_t mdj = M(d, j);
                uint64_t l = 1, t2 = 0;
                for (int p = 32; p >= 1; --p) {
                    uint32_t lsmdp = 0;
                    for (int q = 1; q <= 32; ++q) {
                        uint32_t bit = (q < p) ? ltm[((size_t)(d - 1) * 32 + (p - 1)) * 32 + (q - 1)]
                                                : (q == p ? 1u : 0u);
                        lsmdp += bit * (uint32_t)(1ull << (32 - q));
                    }
                    uint32_t t1 = 0;
                    for (int k = 1; k <= 32; ++k)
                        t1 += ((lsmdp >> (k - 1)) & 1u) * ((mdj >> (k - 1)) & 1u);
                    t1 = t1 % 2;
                    t2 = t2 + t1 * l;
                    l = 2 * l;
                }
                M(d, j) = (uint32_t)t2;
            }
    }
    std::memcpy(m_out, m.data(), m.size() * sizeof(uint32_t));
    std::memcpy(x_out, x.data(), x.size() * sizeof(uint32_t));
    return 0;
}

struct SobolSeq {  // ScrambledSobolSeq{D}
    int D;
    std::vector<uint32_t> m;  // [D][32]
    std::vector<uint32_t> x;  // [D]
    uint32_t n = 0;
    SobolSeq(int D_, const uint32_t *m_all, int m_dims, const uint32_t *shift) : D(D_) {
        if (m_all)
            m.assign(m_all, m_all + (size_t)D * 32);
        else
            m.assign((size_t)D * 32, 0u);
        (void)m_dims;
        x.assign((size_t)D, 0u);
        if (shift)
            for (int d = 0; d < D; ++d) x[d] = shift[d];
    }
    void next(double *out) {  // next!  :158-173
        if (n == 0) {
            n += 1;
            for (int d = 0; d < D; ++d) out[d] = std::ldexp((double)x[d], -32);
            return;
        }
        n += 1;
        uint32_t c = (uint32_t)__builtin_ctz(n - 1);
        for (int d = 0; d < D; ++d) {
            x[d] = x[d] ^ m[(size_t)d * 32 + c];
            out[d] = std::ldexp((double)x[d], -32);
        }
    }
    void skip(int64_t nn, bool exact) {  // skip!  :189-197
        if (nn <= 0) return;
        int64_t nskip = exact ? nn : (int64_t(1) << (int)std::floor(std::log2((double)(nn + 1))));
        std::vector<double> tmp((size_t)std::max(D, 1));
        for (int64_t i = 0; i < nskip; ++i) next(tmp.data());
    }
};

extern "C" int orc_sobol_points(int D, const uint32_t *m, const uint32_t *shift, int64_t skip,
                                int exact, int64_t npts, double *out) {
    SobolSeq s(D, m, D, shift);
    s.skip(skip, exact != 0);
    for (int64_t i = 0; i < npts; ++i) s.next(out + (size_t)i * D);
    return 0;
}

// ------------------------------------------------------------------------------------
// Scramble bits of randomized qMC.  The reference draws the LMS + shift bits of each
// ScrambledSobolSeq(D; scramble_rng) from one RNG (src/scrambled_sobol.jl:104-109); the
// ABI (qiw_rand) defines them by a counter-based generator instead: word w of dimension
// d of a stream is the low half of mix(mix(seed ^ mix(stream)) + 33 d + w), mix being
// SplitMix64's output step after its increment.  Word 0 is the shift, word 1 + p row p of
// the lower-triangular matrix (entries (p, q < p) in bits 31 - q, unit diagonal).
// ------------------------------------------------------------------------------------
static uint64_t rq_mix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static uint64_t rq_stream(int domain, int64_t key, int order, int n_after, int seq) {
    return ((uint64_t)domain << 60) | (((uint64_t)key & 0xFFFFFFFull) << 32) | ((uint64_t)order << 24) |
           ((uint64_t)n_after << 16) | (uint64_t)seq;
}
static uint32_t rq_word(uint64_t seed, uint64_t stream, int d, int w) {
    return (uint32_t)rq_mix(rq_mix(seed ^ rq_mix(stream)) + (uint64_t)(33 * d + w));
}

// LMS + shift scramble of direction numbers m [D][32] (src/scrambled_sobol.jl:111-140):
// bit 32 - p (1-based row p) of a scrambled number is the parity of row p AND the plain one.
extern "C" int orc_rq_scramble(uint64_t seed, uint64_t stream, int D, const uint32_t *m, uint32_t *m_out,
                               uint32_t *x_out) {
    for (int d = 0; d < D; ++d) {
        x_out[d] = rq_word(seed, stream, d, 0);
        uint32_t rows[32];
        for (int p = 0; p < 32; ++p) {
            const uint32_t below = p == 0 ? 0u : (0xFFFFFFFFu << (32 - p));
            rows[p] = (rq_word(seed, stream, d, 1 + p) & below) | (1u << (31 - p));
        }
        for (int j = 0; j < 32; ++j) {
            uint32_t t2 = 0;
            for (int p = 0; p < 32; ++p) t2 |= (uint32_t)(__builtin_popcount(rows[p] & m[d * 32 + j]) & 1) << (31 - p);
            m_out[d * 32 + j] = t2;
        }
    }
    return 0;
}

extern "C" uint64_t orc_rq_stream(int domain, int64_t key, int order, int n_after, int seq) {
    return rq_stream(domain, key, order, n_after, seq);
}

// ------------------------------------------------------------------------------------
// Domain transforms -- src/qmc_integrate.jl:46, :198-258, :363-463
// ------------------------------------------------------------------------------------
static double simplex_volume(int d, double edge) {  // _simplex_volume  :46
    if (d <= 0) return 1.0;
    double p = 1.0;
    for (int i = 1; i <= d; ++i) p *= edge / i;
    return p;
}

struct Transform {
    // Root (d_lesser = -1 marks a plain RootTransform) or DoubleSimplexRoot
    bool dbl = false;
    int d = 0, d_lesser = 0, d_greater = 0;
    double u_i = 0, u_w = 0, u_diff_wi = 0, u_diff_fw = 0, u_diff = 0;
    static Transform root(int d, double t_i, double t_f) {  // :220-227
        Transform t;
        t.dbl = false;
        t.d = d;
        t.u_i = t_i;
        t.u_diff = t_f - t_i;
        return t;
    }
    static Transform dsimplex(int d_before, int d_after, double t_i, double t_w, double t_f) {
        Transform t;  // :398-417
        t.dbl = true;
        t.d = d_before + d_after;
        t.d_lesser = d_before;
        t.d_greater = d_after;
        t.u_i = t_i;
        t.u_w = t_w;
        t.u_diff_wi = t_w - t_i;
        t.u_diff_fw = t_f - t_w;
        return t;
    }
    void apply(const double *x, double *u) const {
        if (!dbl) {  // make_trans_f(::RootTransform)  :235-245
            if (d == 0) return;
            u[0] = std::pow(x[0], 1.0 / d);
            for (int i = 2; i <= d; ++i) u[i - 1] = u[i - 2] * std::pow(x[i - 1], 1.0 / (d - i + 1));
            for (int i = 0; i < d; ++i) u[i] = u_i + u[i] * u_diff;
            return;
        }
        // make_trans_f(::DoubleSimplexRootTransform)  :425-449
        u[0] = std::pow(x[0], 1.0 / d_greater);
        for (int i = 2; i <= d_greater; ++i)
            u[i - 1] = u[i - 2] * std::pow(x[i - 1], 1.0 / (d_greater - i + 1));
        for (int i = 0; i < d_greater; ++i) u[i] *= u_diff_fw;
        for (int i = 0; i < d_greater; ++i) u[i] += u_w;
        if (d_lesser == 0) return;
        u[d_greater] = std::pow(x[d_greater], 1.0 / d_lesser);
        for (int i = d_greater + 2; i <= d; ++i) u[i - 1] = u[i - 2] * std::pow(x[i - 1], 1.0 / (d - i + 1));
        for (int i = d_greater; i < d; ++i) u[i] *= u_diff_wi;
        for (int i = d_greater; i < d; ++i) u[i] += u_i;
    }
    double jacobian() const {
        if (!dbl) return simplex_volume(d, u_diff);  // :254-258
        return simplex_volume(d_lesser, u_diff_wi) * simplex_volume(d_greater, u_diff_fw);  // :458-463
    }
};

extern "C" int orc_simplex_points(int d_before, int d_after, double t_i, double t_w, double t_f,
                                  const double *x, int64_t n, double *out, double *jac_out) {
    Transform t = (d_after == 0) ? Transform::root(d_before, t_i, t_f)
                                 : Transform::dsimplex(d_before, d_after, t_i, t_w, t_f);
    int d = d_before + d_after;
    for (int64_t i = 0; i < n; ++i) t.apply(x + (size_t)i * d, out + (size_t)i * d);
    if (jac_out) *jac_out = t.jacobian();
    return 0;
}

// ------------------------------------------------------------------------------------
// Topologies -- src/diagrammatics.jl:46-337
// ------------------------------------------------------------------------------------
struct Topology {
    int order;
    std::vector<std::pair<int, int>> pairs;
    int parity;
};

static void gen_topologies_impl(const Topology &partial, const std::vector<int> &unpaired,
                                std::vector<Topology> &out) {  // :288-303
    if (unpaired.empty()) {
        out.push_back(partial);
        return;
    }
    for (size_t i = 2; i <= unpaired.size(); ++i) {
        std::pair<int, int> p(unpaired[0], unpaired[i - 1]);
        std::vector<int> rest;
        for (size_t k = 1; k <= unpaired.size(); ++k)
            if (k != 1 && k != i) rest.push_back(unpaired[k - 1]);
        Topology t = partial;
        t.pairs.push_back(p);
        t.order = (int)t.pairs.size();
        t.parity = partial.parity * ((i % 2 == 0) ? 1 : -1);  // (-1)^i
        gen_topologies_impl(t, rest, out);
    }
}

static std::vector<Topology> generate_topologies(int n) {  // :309-312
    Topology empty{0, {}, 1};
    std::vector<int> v;
    for (int i = 1; i <= 2 * n; ++i) v.push_back(i);
    std::vector<Topology> out;
    gen_topologies_impl(empty, v, out);
    return out;
}

static bool iscrossing(std::pair<int, int> p1, std::pair<int, int> p2) {  // :108-114
    if (p1.first > p1.second) std::swap(p1.first, p1.second);
    if (p2.first > p2.second) std::swap(p2.first, p2.second);
    if (p1.first < p2.first && p2.first < p1.second && p1.second < p2.second) return true;
    if (p2.first < p1.first && p1.first < p2.second && p2.second < p1.second) return true;
    return false;
}

static bool pair_doubly_k(std::pair<int, int> p, int k) {  // :98-100
    return (p.first <= k && p.second > k) || (p.second <= k && p.first > k);
}

static bool is_doubly_k_connected(const Topology &t, int k) {  // :231-235, :142-153
    std::vector<std::pair<int, int>> stack, connected, disconnected;
    for (auto &p : t.pairs) (pair_doubly_k(p, k) ? stack : disconnected).push_back(p);
    while (!stack.empty()) {
        auto pc = stack.back();
        stack.pop_back();
        connected.push_back(pc);
        for (int i = (int)disconnected.size() - 1; i >= 0; --i)
            if (iscrossing(pc, disconnected[i])) {
                stack.push_back(disconnected[i]);
                disconnected.erase(disconnected.begin() + i);
            }
    }
    return disconnected.empty();
}

static std::vector<Topology> topologies_at_order(int order, int k, bool with_external_arc) {
    std::vector<Topology> tops = generate_topologies(order);  // :322-337
    if (k < 0) return tops;
    std::vector<Topology> out;
    for (auto &t : tops)
        if (is_doubly_k_connected(t, k)) out.push_back(t);
    if (with_external_arc)
        for (auto &t : out) t.parity *= (k % 2 == 0) ? 1 : -1;
    return out;
}

extern "C" int orc_topologies(int order, int k, int with_external_arc, int max_out,
                              int32_t *pairs_out, int32_t *parity_out) {
    auto tops = topologies_at_order(order, k, with_external_arc != 0);
    if (pairs_out) {
        for (int t = 0; t < (int)tops.size() && t < max_out; ++t) {
            for (int a = 0; a < order; ++a) {
                pairs_out[(size_t)t * 2 * order + 2 * a] = tops[t].pairs[a].first;
                pairs_out[(size_t)t * 2 * order + 2 * a + 1] = tops[t].pairs[a].second;
            }
            parity_out[t] = tops[t].parity;
        }
    }
    return (int)tops.size();
}

// ------------------------------------------------------------------------------------
// Block-sparse operators -- src/sector_block_matrix.jl
// ------------------------------------------------------------------------------------
struct SBM {  // s_i -> (s_f, mat); s_f = -1 when no block
    std::vector<int> s_f;
    std::vector<CMat> mat;
};

static bool load_sbm(const qiw_sbm &in, const std::vector<int> &dims, SBM &out) {
    int ns = (int)dims.size();
    out.s_f.assign(ns, -1);
    out.mat.assign(ns, CMat());
    size_t off = 0;
    for (int b = 0; b < in.n_blocks; ++b) {
        int si = in.s_i[b], sf = in.s_f[b];
        if (si < 0 || si >= ns || sf < 0 || sf >= ns) return false;
        if (out.s_f[si] != -1) {
            set_err("operator is not representable by a SectorBlockMatrix");
            return false;
        }
        CMat m(dims[sf], dims[si]);
        for (size_t e = 0; e < m.a.size(); ++e) m.a[e] = cd(in.data[2 * (off + e)], in.data[2 * (off + e) + 1]);
        off += m.a.size();
        out.s_f[si] = sf;
        out.mat[si] = m;
    }
    return true;
}

// ------------------------------------------------------------------------------------
// Interpolated Green's functions
// ------------------------------------------------------------------------------------
// Natural cubic spline through uniformly spaced values (BSpline(Cubic(Line(OnGrid())))
// with scale(.., LinRange(0, beta, n)), src/spline_gf.jl:194-202).  Stored as second
// derivatives (moments) of the interpolant.
struct CubicSpline {
    double h = 1;
    std::vector<cd> y, M;
    void build(const std::vector<cd> &vals, double beta) {
        int n = (int)vals.size();
        y = vals;
        h = beta / (n - 1);
        M.assign(n, cd(0));
        if (n < 3) return;
        // tridiagonal: M_{i-1} + 4 M_i + M_{i+1} = 6 (y_{i+1} - 2y_i + y_{i-1}) / h^2
        int m = n - 2;
        std::vector<double> cp(m);
        std::vector<cd> dp(m);
        for (int i = 0; i < m; ++i) {
            cd rhs = 6.0 * (y[i + 2] - 2.0 * y[i + 1] + y[i]) / (h * h);
            double b = 4.0, a = (i > 0) ? 1.0 : 0.0, c = (i < m - 1) ? 1.0 : 0.0;
            double denom = b - (i > 0 ? a * cp[i - 1] : 0.0);
            cp[i] = c / denom;
            dp[i] = (rhs - (i > 0 ? a * dp[i - 1] : cd(0))) / denom;
        }
        for (int i = m - 1; i >= 0; --i) M[i + 1] = dp[i] - (i < m - 1 ? cp[i] * M[i + 2] : cd(0));
    }
    cd eval(double z) const {
        int n = (int)y.size();
        double x = z / h;
        int j = (int)std::floor(x);
        if (j < 0) j = 0;
        if (j > n - 2) j = n - 2;
        double t = x - j;  // in [0,1]
        double A = 1.0 - t, B = t;
        return A * y[j] + B * y[j + 1] +
               ((A * A * A - A) * M[j] + (B * B * B - B) * M[j + 1]) * (h * h / 6.0);
    }
};

// IncrementalSpline -- src/utility.jl:106-157
struct IncSpline {
    double h = 1;  // step(knots)
    std::vector<cd> data, der_data;
    void init(double h_, cd val1, cd der1) {  // :120-128
        h = h_;
        data.assign(1, val1);
        der_data.assign(1, der1 * h);
    }
    void extend(cd val) {  // :136-140
        data.push_back(val);
        der_data.push_back(2.0 * (data[data.size() - 1] - data[data.size() - 2]) - der_data.back());
    }
    cd eval(double z) const {  // :147-157
        double x = 1 + z / h;
        int i = (int)std::floor(x);
        i = std::min(i, (int)data.size() - 1);
        double dx = x - i;
        cd c3 = der_data[i - 1] - 2.0 * data[i];
        return data[i - 1] * (1 - dx * dx) + data[i] * (1 - (1 - dx) * (1 - dx)) +
               c3 * (0.25 - (dx - 0.5) * (dx - 0.5));
    }
};

// A grid-stored imaginary-time Green's function evaluated at two contour times t1 >= t2:
// Keldysh.jl's interpolate! on a grid container, behind kd.interpolate!(x, P_s, t_f, t_i)
// (src/topology_eval.jl:368) and behind the call of an ImaginaryTimeGF pair propagator
// (src/topology_eval.jl:414).  The container stores g by index difference, G[i1, i2] =
// g(i1 - i2); the value is the bilinear interpolation over the grid cell (i1, i2) that holds
// (t1, t2), and inside a diagonal cell (i1 == i2) the interpolation over its lower triangle,
// which is linear in t1 - t2.  With a, b the fractions of t1, t2 in their intervals and
// d = i1 - i2 both read
//     g(d) + a (1 - b) (g(d+1) - g(d)) - (1 - a) b (g(d) - g(d-1)),   g(-1) = 2 g(0) - g(1).
// Pinned by the reference's golden /inchworm/P and /inchworm/g (test/inchworm.jl:180-212,
// a grid P) and /values of test/topology_eval.h5 (test/topology_eval.jl:37-155, a grid P
// and grid Delta), both reproduced to round-off; linear interpolation in t1 - t2 alone
// misses them by 2e-8 and 3e-4.
struct GridCell {
    int d;          // i1 - i2
    double wp, wm;  // a (1 - b), (1 - a) b
};
static inline GridCell grid_cell(double t1, double t2, double h, int n) {
    double x1 = t1 / h, x2 = t2 / h;
    int i1 = (int)std::floor(x1), i2 = (int)std::floor(x2);
    if (i1 < 0) i1 = 0;
    if (i2 < 0) i2 = 0;
    if (i1 > n - 2) i1 = n - 2;
    if (i2 > n - 2) i2 = n - 2;
    const double a = x1 - i1, b = x2 - i2;
    return GridCell{i1 - i2, a * (1.0 - b), (1.0 - a) * b};
}
// g at cell c from grid values vals(j), j = 0..n-1
template <class G>
static inline cd grid_bilinear(const GridCell &c, G vals) {
    const cd g0 = vals(c.d), gp = vals(c.d + 1);
    const cd dm = c.d > 0 ? g0 - vals(c.d - 1) : gp - g0;
    return g0 + c.wp * (gp - g0) - c.wm * dm;
}

// A scalar propagator of an InteractionPair: g(t_f, t_i)
// The fermionic DLR kernel e^{-w tau} / (1 + e^{-w beta}), 0 <= tau <= beta: the basis in
// which Lehmann.jl's dlr2tau sums a DLRImaginaryTimeGF's coefficients (called at
// src/keldysh_dlr.jl:174; Lehmann.jl Spectral.kernelFermiT).
static double dlr_kernel(double tau, double w, double beta) {
    return w >= 0.0 ? std::exp(-w * tau) / (1.0 + std::exp(-w * beta))
                    : std::exp(w * (beta - tau)) / (1.0 + std::exp(w * beta));
}

struct ScalarGF {
    int kind = QIW_GF_GRID_LINEAR;
    int xi = -1;
    double beta = 1, h = 1;
    std::vector<cd> vals;
    CubicSpline spl;
    std::vector<double> dlr_w;  // QIW_GF_DLR: frequencies and coefficients
    std::vector<cd> dlr_c;
    // DLRImaginaryTimeGF (src/keldysh_dlr.jl:172-200): the pole sum at the time difference,
    // or xi times it at beta + ref1 - ref2 when ref1 < ref2.
    cd dlr_eval(double ref1, double ref2) const {
        double d = ref1 - ref2, sign = 1.0;
        if (!(ref1 >= ref2)) {
            d += beta;
            sign = xi;
        }
        cd s = 0;
        for (size_t k = 0; k < dlr_w.size(); ++k) s += dlr_c[k] * dlr_kernel(d, dlr_w[k], beta);
        return sign * s;
    }
    void build(const double *v, int n, int kind_, int xi_, double beta_) {
        kind = kind_;
        xi = xi_;
        beta = beta_;
        h = beta / (n - 1);
        vals.resize(n);
        for (int j = 0; j < n; ++j) vals[j] = cd(v[2 * j], v[2 * j + 1]);
        if (kind == QIW_GF_CUBIC_SPLINE) spl.build(vals, beta);
    }
    // SplineInterpolatedGF (src/spline_gf.jl:210-219): the spline in ref1 - ref2, or
    // xi * spline(beta + ref1 - ref2) when ref1 < ref2.  A grid ImaginaryTimeGF: the cell
    // interpolation above (the evaluator calls it with ref1 >= ref2 only, after its ordering
    // clamp, src/topology_eval.jl:408-411).
    cd operator()(double ref1, double ref2) const {
        if (kind == QIW_GF_DLR) return dlr_eval(ref1, ref2);
        if (kind == QIW_GF_CUBIC_SPLINE)
            return ref1 >= ref2 ? spl.eval(ref1 - ref2) : double(xi) * spl.eval(beta + ref1 - ref2);
        if (!(ref1 >= ref2)) return cd(0);
        const int n = (int)vals.size();
        return grid_bilinear(grid_cell(ref1, ref2, h, n), [&](int j) { return vals[j]; });
    }
};

// ------------------------------------------------------------------------------------
// Model + propagator state
// ------------------------------------------------------------------------------------
struct Model {
    double beta;
    int ntau;
    double h;
    int ns;
    std::vector<int> dims;
    std::vector<std::vector<double>> E;  // per sector eigenvalues
    int ppgf_kind;
    int np;
    std::vector<std::pair<SBM, SBM>> pair_ops;  // (op_i, op_f)
    std::vector<ScalarGF> props;
    int ncorr;
    std::vector<std::pair<SBM, SBM>> corr_ops;  // (A, B)
    SBM identity;
    std::vector<std::vector<int>> attachable;  // subspace_attachable_pairs
    double lambda0;                            // log(Z_atomic)/beta
    std::vector<int> bt_off;
    int bt_size;
};

static bool load_model(const qiw_model *mi, Model &m) {
    m.beta = mi->beta;
    m.ntau = mi->n_tau;
    m.h = m.beta / (m.ntau - 1);
    m.ns = mi->n_sectors;
    m.dims.assign(mi->sector_dim, mi->sector_dim + m.ns);
    m.E.clear();
    size_t off = 0;
    for (int s = 0; s < m.ns; ++s) {
        m.E.emplace_back(mi->energies + off, mi->energies + off + m.dims[s]);
        off += m.dims[s];
    }
    m.ppgf_kind = mi->ppgf_kind;
    m.np = mi->n_pairs;
    m.pair_ops.resize(m.np);
    m.props.resize(m.np);
    for (int p = 0; p < m.np; ++p) {
        if (!load_sbm(mi->pair_op_i[p], m.dims, m.pair_ops[p].first)) return false;
        if (!load_sbm(mi->pair_op_f[p], m.dims, m.pair_ops[p].second)) return false;
        m.props[p].build(mi->pair_prop + (size_t)2 * p * m.ntau, m.ntau, mi->pair_prop_kind[p],
                         mi->pair_prop_xi[p], m.beta);
        if (mi->pair_prop_kind[p] == QIW_GF_DLR) {
            if (mi->dlr_rank < 1 || !mi->pair_dlr_omega || !mi->pair_dlr_coef) return false;
            for (int k = 0; k < mi->dlr_rank; ++k) {
                const size_t i = (size_t)p * mi->dlr_rank + k;
                m.props[p].dlr_w.push_back(mi->pair_dlr_omega[i]);
                m.props[p].dlr_c.push_back(cd(mi->pair_dlr_coef[2 * i], mi->pair_dlr_coef[2 * i + 1]));
            }
        }
    }
    m.ncorr = mi->n_corr;
    m.corr_ops.resize(m.ncorr);
    for (int c = 0; c < m.ncorr; ++c) {
        if (!load_sbm(mi->corr_A[c], m.dims, m.corr_ops[c].first)) return false;
        if (!load_sbm(mi->corr_B[c], m.dims, m.corr_ops[c].second)) return false;
    }
    m.identity.s_f.resize(m.ns);
    m.identity.mat.resize(m.ns);
    for (int s = 0; s < m.ns; ++s) {
        m.identity.s_f[s] = s;
        CMat id(m.dims[s], m.dims[s]);
        for (int i = 0; i < m.dims[s]; ++i) id(i, i) = 1.0;
        m.identity.mat[s] = id;
    }
    m.attachable.assign(m.ns, {});  // src/expansion.jl:180-183
    for (int s = 0; s < m.ns; ++s)
        for (int p = 0; p < m.np; ++p)
            if (m.pair_ops[p].first.s_f[s] >= 0) m.attachable[s].push_back(p);
    // lambda = log(Z)/beta, Z = sum exp(-beta E)  (src/exact_atomic_ppgf.jl:130-135)
    double emin = 1e300;
    for (auto &v : m.E)
        for (double e : v) emin = std::min(emin, e);
    double zs = 0;
    for (auto &v : m.E)
        for (double e : v) zs += std::exp(-m.beta * (e - emin));
    m.lambda0 = (std::log(zs) - m.beta * emin) / m.beta;
    m.bt_off.resize(m.ns);
    m.bt_size = 0;
    for (int s = 0; s < m.ns; ++s) {
        m.bt_off[s] = m.bt_size;
        m.bt_size += m.dims[s] * m.dims[s];
    }
    return true;
}

// The bold propagator expansion.P: grid table + (for INC_SPLINE) incremental splines.
struct PPGF {
    int kind;
    int n_known;                  // interpolant data length (INC_SPLINE)
    std::vector<std::vector<CMat>> tab;  // [ntau][ns]
    std::vector<std::vector<IncSpline>> spl;  // [ns][m*m] (INC_SPLINE)
};

static void atomic_ppgf(const Model &m, PPGF &P) {  // src/ppgf.jl:144-207, :710-720
    P.kind = m.ppgf_kind;
    P.tab.assign(m.ntau, std::vector<CMat>(m.ns));
    for (int j = 0; j < m.ntau; ++j) {
        double tau = j * m.h;
        for (int s = 0; s < m.ns; ++s) {
            CMat x(m.dims[s], m.dims[s]);
            for (int i = 0; i < m.dims[s]; ++i) x(i, i) = -I * std::exp(-tau * (m.E[s][i] + m.lambda0));
            P.tab[j][s] = x;
        }
    }
    P.n_known = 1;
    if (P.kind == QIW_GF_INC_SPLINE) {
        P.spl.assign(m.ns, {});
        for (int s = 0; s < m.ns; ++s) {
            int d = m.dims[s];
            P.spl[s].resize((size_t)d * d);
            for (int l = 0; l < d; ++l)
                for (int k = 0; k < d; ++k) {
                    cd der = (k == l) ? -I * (-m.E[s][k] - m.lambda0) : cd(0);
                    P.spl[s][k + l * d].init(m.h, P.tab[0][s](k, l), der);
                }
        }
    } else {
        P.n_known = m.ntau;
    }
}

// Bold P lookup, kd.interpolate!(x, P_s, t_f, t_i) (src/topology_eval.jl:368): the grid
// cell interpolation of the stored values P_s(tau_j) (grid_bilinear) for both container
// kinds.  For IncSplineImaginaryTimeGF Keldysh's generic interpolate! works on the wrapped
// grid function, the incremental splines are bookkeeping only (the quadratic
// IncrementalSpline of src/utility.jl:147-157 misses the golden first bold step by 1e-6).
static void ppgf_lookup(const Model &m, const PPGF &P, int s, double tf, double ti, CMat &out) {
    int d = m.dims[s];
    out.r = out.c = d;
    out.a.resize((size_t)d * d);
    if (!(tf >= ti)) {  // not reached after the ordering clamp; kept for completeness
        for (auto &v : out.a) v = cd(0);
        return;
    }
    const GridCell c = grid_cell(tf, ti, m.h, m.ntau);
    for (int e = 0; e < d * d; ++e) out.a[e] = grid_bilinear(c, [&](int j) { return P.tab[j][s].a[e]; });
}

// Exact atomic P0 (src/exact_atomic_ppgf.jl:109-118)
static void p0_lookup(const Model &m, int s, double tf, double ti, CMat &out) {
    int d = m.dims[s];
    out = CMat(d, d);
    cd dz = cd(0, -tf) - cd(0, -ti);  // z1.val - z2.val, val = -i tau
    for (int i = 0; i < d; ++i) out(i, i) = -I * std::exp(-I * dz * (m.E[s][i] + m.lambda0));
}

// set_ppgf! (src/ppgf.jl:474-504)
static void set_ppgf(const Model &m, PPGF &P, int jf, const std::vector<CMat> &val) {
    if (P.kind == QIW_GF_INC_SPLINE) {
        for (int s = 0; s < m.ns; ++s) {
            int d = m.dims[s];
            for (int e = 0; e < d * d; ++e) P.spl[s][e].extend(val[s].a[e]);
        }
        P.n_known = (int)P.spl[0][0].data.size();
        for (int s = 0; s < m.ns; ++s) P.tab[P.n_known - 1][s] = val[s];
        return;
    }
    for (int s = 0; s < m.ns; ++s) P.tab[jf][s] = val[s];
}

// normalize!(P, tau) (src/ppgf.jl:646-668, :677-695)
static void normalize_tau(const Model &m, PPGF &P, int jf) {
    double tau_f = jf * m.h;
    double pmax = -1e300;
    for (int s = 0; s < m.ns; ++s)
        for (int i = 0; i < m.dims[s]; ++i) pmax = std::max(pmax, -P.tab[jf][s](i, i).imag());
    double lam = std::log(pmax) / tau_f;
    if (P.kind == QIW_GF_INC_SPLINE) {
        int n = P.n_known;
        for (int s = 0; s < m.ns; ++s) {
            for (int j = 0; j < n; ++j) {
                cd f = std::exp(-I * cd(0, -j * m.h) * lam);
                for (auto &v : P.tab[j][s].a) v *= f;
            }
            int d = m.dims[s];
            for (int e = 0; e < d * d; ++e) {
                IncSpline &sp = P.spl[s][e];
                cd der0 = sp.der_data[0] / sp.h - I * lam * sp.data[0];
                sp.init(m.h, P.tab[0][s].a[e], der0);
                for (int j = 1; j < n; ++j) sp.extend(P.tab[j][s].a[e]);
            }
        }
        return;
    }
    for (int s = 0; s < m.ns; ++s)
        for (int j = 0; j < m.ntau; ++j) {
            cd f = std::exp(-I * cd(0, -j * m.h) * lam);
            for (auto &v : P.tab[j][s].a) v *= f;
        }
}

static cd partition_function(const Model &m, const PPGF &P) {  // src/ppgf.jl:611-617
    cd z = 0;
    for (int s = 0; s < m.ns; ++s) {
        cd tr = 0;
        for (int i = 0; i < m.dims[s]; ++i) tr += P.tab[m.ntau - 1][s](i, i);
        z += I * tr;
    }
    return z;
}

// ------------------------------------------------------------------------------------
// LazyMatrixProduct -- src/utility.jl:234-323
// ------------------------------------------------------------------------------------
struct LMP {
    std::vector<const CMat *> mats;
    std::vector<CMat> partial;
    int n_mats = 0, n_prods = 0;
    int ncols = 0;
    void init(int max_n, int ncols_) {
        mats.assign(max_n, nullptr);
        partial.assign(max_n, CMat());
        ncols = ncols_;
        n_mats = n_prods = 0;
    }
    void pushfirst(const CMat *A) { mats[n_mats++] = A; }
    void popfirst() {
        n_mats -= 1;
        n_prods = std::min(n_prods, n_mats);
    }
    const CMat &eval() {
        if (n_prods == 0) {
            partial[0] = *mats[0];
            n_prods = 1;
        }
        for (int n = n_prods + 1; n <= n_mats; ++n) {
            matmul(partial[n - 1], *mats[n - 1], partial[n - 2]);
            g_flops += 2.0 * mats[n - 1]->r * mats[n - 1]->c * partial[n - 2].c;
        }
        n_prods = n_mats;
        return partial[n_prods - 1];
    }
};

// ------------------------------------------------------------------------------------
// TopologyEvaluator -- src/topology_eval.jl:179-556
// ------------------------------------------------------------------------------------
enum NodeKind { PAIR = 1, IDENTITY = 2, INCH = 3, OPERATOR = 4 };
struct Node {
    NodeKind kind;
    int arc_index;       // arc (pair node) or corr pair index (operator node)
    int operator_index;  // 1 = head/op_i or A, 2 = tail/op_f or B
};
struct FixedNode {
    Node node;
    double time;
};

struct ESBM {  // EvolvingSectorBlockMatrix: s_i -> (s_f, mat0, matt)
    std::vector<int> s_f;
    std::vector<CMat> mat0, matt;
    void from(const SBM &b) {
        s_f = b.s_f;
        mat0 = b.mat;
        matt = b.mat;
    }
    void update(const std::vector<CMat> &p) {  // update_finite_time!  :162-166
        for (size_t si = 0; si < s_f.size(); ++si)
            if (s_f[si] >= 0) {
                matmul(matt[si], mat0[si], p[si]);
                g_flops += 2.0 * mat0[si].r * mat0[si].c * p[si].c;
            }
    }
};

struct Evaluator {
    const Model *m;
    const PPGF *P;
    bool use_bold;
    int order, n_nodes;
    std::vector<Node> conf;
    std::vector<double> times;
    std::vector<int> top_to_conf_pos, var_time_pos;
    std::vector<std::vector<CMat>> ppgf_mats;  // [n_nodes-1][ns]
    std::vector<std::vector<cd>> pair_ints;    // [order][np]
    std::vector<int> selected;
    std::vector<CMat> top_result;
    std::vector<LMP> lmp;
    // node_mats: fixed non-pair -> single ESBM; pair positions -> [np][2]
    std::vector<int> node_is_pair;
    std::vector<ESBM> node_single;
    std::vector<std::vector<ESBM>> node_pair;  // [pos][p*2 + idx]
    std::vector<CMat> result;
    int64_t leaves = 0;

    void build(const Model &mod, const PPGF *Pp, int order_, bool bold, const std::map<int, FixedNode> &fixed) {
        m = &mod;
        P = Pp;
        use_bold = bold;
        order = order_;
        int nfix = 0;
        for (auto &kv : fixed)
            if (kv.second.node.kind != PAIR) ++nfix;
        n_nodes = 2 * order + nfix;  // :254
        conf.assign(n_nodes + 1, Node{PAIR, -1, -1});
        times.assign(n_nodes + 1, 0.0);
        for (auto &kv : fixed) {
            conf[kv.first] = kv.second.node;
            times[kv.first] = kv.second.time;
        }
        top_to_conf_pos.clear();  // :269-272
        for (int pos = n_nodes; pos >= 1; --pos) {
            auto it = fixed.find(pos);
            if (!(it != fixed.end() && it->second.node.kind != PAIR)) top_to_conf_pos.push_back(pos);
        }
        var_time_pos.clear();  // :275
        for (int pos = n_nodes; pos >= 1; --pos)
            if (!fixed.count(pos)) var_time_pos.push_back(pos);
        ppgf_mats.assign(std::max(n_nodes - 1, 0), std::vector<CMat>(m->ns));
        pair_ints.assign(order, std::vector<cd>(m->np));
        selected.assign(order, -1);
        top_result.resize(m->ns);
        int maxd = *std::max_element(m->dims.begin(), m->dims.end());
        (void)maxd;
        lmp.resize(m->ns);
        for (int s = 0; s < m->ns; ++s) lmp[s].init(2 * n_nodes - 1, m->dims[s]);
        result.resize(m->ns);
        node_is_pair.assign(n_nodes + 1, 0);
        node_single.assign(n_nodes + 1, ESBM());
        node_pair.assign(n_nodes + 1, {});
        for (int pos = 1; pos <= n_nodes; ++pos) {  // :293-314
            auto it = fixed.find(pos);
            if (it != fixed.end() && it->second.node.kind != PAIR) {
                const Node &nd = it->second.node;
                if (nd.kind == OPERATOR)
                    node_single[pos].from(nd.operator_index == 1 ? m->corr_ops[nd.arc_index].first
                                                                 : m->corr_ops[nd.arc_index].second);
                else
                    node_single[pos].from(m->identity);
            } else {
                node_is_pair[pos] = 1;
                node_pair[pos].resize((size_t)m->np * 2);
                for (int p = 0; p < m->np; ++p) {
                    node_pair[pos][p * 2 + 0].from(m->pair_ops[p].first);
                    node_pair[pos][p * 2 + 1].from(m->pair_ops[p].second);
                }
            }
        }
    }

    void traverse(int pos, int s_i, int s_f, cd w) {  // :454-556
        if (pos > n_nodes) {
            if (s_i != s_f) throw std::runtime_error("off-diagonal");
            const CMat &prod = lmp[s_f].eval();
            CMat &tr = top_result[s_i];
            for (size_t e = 0; e < tr.a.size(); ++e) tr.a[e] += w * prod.a[e];
            g_flops += 2.0 * (double)tr.a.size();
            ++leaves;
            return;
        }
        const Node &node = conf[pos];
        if (node.kind == PAIR) {
            if (node.operator_index == 1) {
                for (int p : m->attachable[s_i]) {
                    selected[node.arc_index - 1] = p;
                    const ESBM &e = node_pair[pos][p * 2 + 0];
                    int s_next = e.s_f[s_i];
                    lmp[s_f].pushfirst(&e.matt[s_i]);
                    traverse(pos + 1, s_next, s_f, w);
                    lmp[s_f].popfirst();
                }
            } else {
                int p = selected[node.arc_index - 1];
                const ESBM &e = node_pair[pos][p * 2 + 1];
                if (e.s_f[s_i] >= 0) {
                    lmp[s_f].pushfirst(&e.matt[s_i]);
                    traverse(pos + 1, e.s_f[s_i], s_f, pair_ints[node.arc_index - 1][p] * w);
                    lmp[s_f].popfirst();
                }
            }
        } else if (node.kind == OPERATOR) {
            const ESBM &e = node_single[pos];
            if (e.s_f[s_i] >= 0) {
                lmp[s_f].pushfirst(&e.matt[s_i]);
                traverse(pos + 1, e.s_f[s_i], s_f, w);
                lmp[s_f].popfirst();
            }
        } else {
            if (pos != 1) lmp[s_f].pushfirst(&node_single[pos].matt[s_i]);
            traverse(pos + 1, s_i, s_f, w);
            if (pos != 1) lmp[s_f].popfirst();
        }
    }

    // functor (src/topology_eval.jl:350-437); vt = variable times in var_time_pos order
    const std::vector<CMat> &operator()(const std::vector<Topology> &tops, const double *vt) {
        for (size_t i = 0; i < var_time_pos.size(); ++i) times[var_time_pos[i]] = vt[i];
        for (int i = 1; i <= n_nodes - 1; ++i) {  // :357-374
            double ti = times[i], tf = times[i + 1];
            if (tf < ti) tf = ti;
            for (int s = 0; s < m->ns; ++s) {
                CMat &x = ppgf_mats[i - 1][s];
                if (use_bold)
                    ppgf_lookup(*m, *P, s, tf, ti, x);
                else
                    p0_lookup(*m, s, tf, ti, x);
                for (auto &v : x.a) v *= I;
            }
        }
        for (int pos = 2; pos <= n_nodes; ++pos) {  // :376-388
            if (!node_is_pair[pos])
                node_single[pos].update(ppgf_mats[pos - 2]);
            else
                for (auto &e : node_pair[pos]) e.update(ppgf_mats[pos - 2]);
        }
        for (int s = 0; s < m->ns; ++s) result[s] = CMat(m->dims[s], m->dims[s]);
        for (const Topology &top : tops) {  // :392-434
            for (int a = 0; a < (int)top.pairs.size(); ++a) {
                int pos_head = top_to_conf_pos[top.pairs[a].second - 1];
                int pos_tail = top_to_conf_pos[top.pairs[a].first - 1];
                conf[pos_head] = Node{PAIR, a + 1, 1};
                conf[pos_tail] = Node{PAIR, a + 1, 2};
                double ti = times[pos_head], tf = times[pos_tail];
                if (tf < ti) tf = ti;
                for (int p = 0; p < m->np; ++p) pair_ints[a][p] = I * m->props[p](tf, ti);
            }
            for (int s = 0; s < m->ns; ++s) top_result[s] = CMat(m->dims[s], m->dims[s]);
            for (int s0 = 0; s0 < m->ns; ++s0) traverse(1, s0, s0, cd(1.0));
            cd f = -I * double(top.parity) * ((top.order % 2 == 0) ? 1.0 : -1.0);
            for (int s = 0; s < m->ns; ++s)
                for (size_t e = 0; e < result[s].a.size(); ++e) result[s].a[e] += f * top_result[s].a[e];
        }
        return result;
    }
};

// ------------------------------------------------------------------------------------
// Integration -- src/qmc_integrate.jl:497-507, :597-612 with the sample range split into
// fixed chunks that are summed in order (deterministic for any thread count).
// ------------------------------------------------------------------------------------
struct SobolSource {
    const uint32_t *m;
    const uint32_t *shift;
    const double *pts = nullptr;  // caller-supplied points instead (scramble = 3): [npts][pstride]
    int pstride = 0;
    int64_t npts = 0;
};

static int g_threads = 1;

// RandomizationParams of a run (src/randomization.jl:46-57)
struct RandParams {
    int n_seqs = 1;
    bool scramble = false;
    bool given = false;   // the caller's sequences (scramble = 2 or 3)
    bool points = false;  // ... as points (scramble = 3)
    uint64_t seed = 0;
    double target_std = 0.0;
};
static RandParams rand_params(const qiw_run *run) {
    RandParams r;
    if (run->rand_params) {
        r.n_seqs = run->rand_params->N_seqs;
        r.scramble = run->rand_params->scramble != 0;
        r.given = run->rand_params->scramble == 2 || run->rand_params->scramble == 3;
        r.points = run->rand_params->scramble == 3;
        r.seed = run->rand_params->seed;
        r.target_std = run->rand_params->target_std;
    }
    return r;
}

// Caller-drawn sequences (qiw_rand.scramble = 2): the states (m, x) of the
// ScrambledSobolSeq objects the reference's loops construct with the caller's RNG
// (src/randomization.jl:94), consumed in construction order -- here, the order in which the
// oracle, following the reference's loops, asks for them.  Sequence i holds `stride` rows
// of which the consuming group uses the first 2k.
// The (domain, key, order, n_pts_after, s) of every consumed sequence is logged (the
// consumption order, for tests).
static struct {
    std::vector<uint32_t> m, x;
    int stride = 0;
    int64_t n = 0, next = 0;
    std::vector<int64_t> log;
} g_given;

extern "C" int orc_set_scrambles(int64_t n, int stride, const uint32_t *m, const uint32_t *x) {
    g_given.n = n;
    g_given.stride = stride;
    g_given.next = 0;
    g_given.log.clear();
    g_given.m.assign(m, m + (size_t)n * stride * 32);
    g_given.x.assign(x, x + (size_t)n * stride);
    return 0;
}
// Caller-supplied point sequences (qiw_rand.scramble = 3): the points that the sequences
// seq_type(D; scramble_rng = rng) of mean_std_from_randomization (src/randomization.jl:86-94)
// yield through next! (src/qmc_integrate.jl:597-612), consumed as under scramble = 2;
// point r of sequence i is x[(i * npts + r) * stride + d].
static struct {
    std::vector<double> x;
    int stride = 0;
    int64_t n = 0, npts = 0;
} g_points;

extern "C" int orc_set_points(int64_t n, int64_t npts, int stride, const double *x) {
    g_points.n = n;
    g_points.npts = npts;
    g_points.stride = stride;
    g_points.x.assign(x, x + (size_t)n * npts * stride);
    g_given.next = 0;
    g_given.log.clear();
    return 0;
}
extern "C" int64_t orc_scrambles_used(void) { return g_given.next; }
extern "C" int64_t orc_scrambles_log(int64_t cap, int64_t *out) {
    for (size_t i = 0; i < g_given.log.size() && (int64_t)i < 5 * cap; ++i) out[i] = g_given.log[i];
    return (int64_t)g_given.log.size() / 5;
}

// Direction numbers and shift of sequence `seq` of a group: the staged ones (rng =
// nothing), the staged direction numbers scrambled with the stream's bits, or the caller's
// next sequence.
struct SeqStore {
    std::vector<uint32_t> m, x;
};
static SobolSource sequence_source(const RandParams &rq, SobolSource staged, int dim, int domain, int64_t key,
                                   int order, int n_after, int seq, SeqStore &store) {
    if (rq.given) {
        const int64_t have = rq.points ? g_points.n : g_given.n;
        const int stride = rq.points ? g_points.stride : g_given.stride;
        if (g_given.next >= have || dim > stride) throw std::runtime_error("not enough caller-drawn sequences");
        const int64_t i = g_given.next++;
        for (int64_t v : {(int64_t)domain, key, (int64_t)order, (int64_t)n_after, (int64_t)seq}) g_given.log.push_back(v);
        if (rq.points)
            return SobolSource{nullptr, nullptr, g_points.x.data() + (size_t)i * g_points.npts * stride, stride,
                               g_points.npts};
        store.m.assign(g_given.m.begin() + i * g_given.stride * 32, g_given.m.begin() + i * g_given.stride * 32 + (size_t)dim * 32);
        store.x.assign(g_given.x.begin() + i * g_given.stride, g_given.x.begin() + i * g_given.stride + std::max(dim, 1));
        return SobolSource{store.m.data(), store.x.data()};
    }
    if (!rq.scramble) return staged;
    store.m.assign((size_t)dim * 32, 0u);
    store.x.assign((size_t)std::max(dim, 1), 0u);
    orc_rq_scramble(rq.seed, rq_stream(domain, key, order, n_after, seq), dim, staged.m, store.m.data(),
                    store.x.data());
    return SobolSource{store.m.data(), store.x.data()};
}

// mean_std_from_randomization (src/randomization.jl:86-100) on flattened samples: sample
// s = f(s) for s = 0, 1, ..., stopping after s >= 1 once the largest element of the
// standard deviation is <= target_std.  The mean is sum * (1/n) for sector block matrices
// (their `/` multiplies by the reciprocal, src/sector_block_matrix.jl:125) and sum / n for
// scalars; the variance is sum |x - sum/n|^2 / (n - 1) (Statistics.var,
// src/sector_block_matrix.jl:191-223), NaN for one sample.
template <class F>
static void mean_std(F f, const RandParams &rq, bool recip_mean, std::vector<cd> &mean, std::vector<double> &sd) {
    std::vector<std::vector<cd>> xs;
    auto stdev = [&](size_t n) {
        const size_t ne = xs[0].size();
        std::vector<double> out(ne);
        for (size_t e = 0; e < ne; ++e) {
            cd sum = 0;
            for (size_t s = 0; s < n; ++s) sum += xs[s][e];
            const cd mu = sum / (double)n;
            double v = 0;
            for (size_t s = 0; s < n; ++s) v += std::norm(xs[s][e] - mu);
            out[e] = std::sqrt(v / (double)(n - 1));
        }
        return out;
    };
    for (int s = 0; s < rq.n_seqs; ++s) {
        xs.push_back(f(s));
        if (s >= 1) {
            std::vector<double> st = stdev(xs.size());
            double mx = 0;
            for (double v : st) mx = std::max(mx, v);
            if (mx <= rq.target_std) break;
        }
    }
    const size_t n = xs.size(), ne = xs[0].size();
    mean.assign(ne, 0.0);
    for (size_t e = 0; e < ne; ++e) {
        cd sum = 0;
        for (size_t s = 0; s < n; ++s) sum += xs[s][e];
        mean[e] = recip_mean ? sum * (1.0 / (double)n) : sum / (double)n;
    }
    sd = stdev(n);
}

static std::vector<CMat> integrate_group(const Model &m, const PPGF *P, bool bold, int order,
                                         const std::map<int, FixedNode> &fixed, const Transform &tr,
                                         const std::vector<Topology> &tops, int64_t N, SobolSource sob,
                                         int64_t sample_first = 0, int64_t sample_count = -1, int n_var = -1) {
    int d = n_var >= 0 ? n_var : 2 * order;  // variable times (self-energy diagrams: 2k - 2)
    if (sample_count < 0) sample_count = N;
    if (sob.pts && sample_first + sample_count > sob.npts) throw std::runtime_error("too few caller-supplied points");
    double J = tr.jacobian();
    cd dir = 1.0;
    for (int i = 0; i < d; ++i) dir *= -I;  // prod(branch_direction) over the imaginary branch
    // chunks of up to 64 samples, smaller when N is small so that every thread has work
    const int64_t CH = std::max<int64_t>(1, std::min<int64_t>(64, sample_count / std::max(g_threads, 1)));
    int64_t nch = (sample_count + CH - 1) / CH;
    std::vector<std::vector<CMat>> partial(nch);
    int64_t leaves_total = 0;
    bool offdiag = false;
#pragma omp parallel num_threads(g_threads) reduction(+ : leaves_total)
    {
        Evaluator ev;
        ev.build(m, P, order, bold, fixed);
        std::vector<double> x(std::max(d, 1)), u(std::max(d, 1));
#pragma omp for schedule(dynamic, 1)
        for (int64_t c = 0; c < nch; ++c) {
            int64_t i0 = sample_first + c * CH;
            int64_t i1 = std::min(sample_first + sample_count, i0 + CH);
            SobolSeq seq(d, sob.pts ? nullptr : sob.m, d, sob.pts ? nullptr : sob.shift);
            if (!sob.pts) seq.skip(i0, true);
            std::vector<CMat> acc(m.ns);
            for (int s = 0; s < m.ns; ++s) acc[s] = CMat(m.dims[s], m.dims[s]);
            for (int64_t i = i0; i < i1; ++i) {
                if (sob.pts) {
                    for (int k = 0; k < d; ++k) x[k] = sob.pts[i * sob.pstride + k];
                } else {
                    seq.next(x.data());
                }
                tr.apply(x.data(), u.data());
                bool ok = true;
                for (int k = 0; k < d; ++k)
                    if (!(u[k] >= 0.0)) ok = false;  // :608
                if (!ok) continue;
                try {
                    const std::vector<CMat> &r = ev(tops, u.data());
                    for (int s = 0; s < m.ns; ++s)
                        for (size_t e = 0; e < acc[s].a.size(); ++e) acc[s].a[e] += (dir * r[s].a[e]) * J;
                } catch (...) {
                    offdiag = true;
                }
            }
            partial[c] = acc;
        }
        leaves_total += ev.leaves;
    }
    if (offdiag) throw std::runtime_error("off-diagonal");
    g_leaves += leaves_total;
    g_diagram_evals += sample_count * (int64_t)tops.size();
    std::vector<CMat> res(m.ns);
    for (int s = 0; s < m.ns; ++s) res[s] = CMat(m.dims[s], m.dims[s]);
    for (int64_t c = 0; c < nch; ++c)
        for (int s = 0; s < m.ns; ++s)
            for (size_t e = 0; e < res[s].a.size(); ++e) res[s].a[e] += partial[c][s].a[e];
    for (int s = 0; s < m.ns; ++s)
        for (auto &v : res[s].a) v *= (1.0 / (double)N);
    return res;
}

// ------------------------------------------------------------------------------------
// Drivers -- src/inchworm.jl
// ------------------------------------------------------------------------------------
struct TopData {  // TopologiesInputData  :76-87
    int order, n_pts_after;
    std::vector<Topology> tops;
};

static void bt_store(const Model &m, const std::vector<CMat> &v, double *out) {
    for (int s = 0; s < m.ns; ++s)
        for (size_t e = 0; e < v[s].a.size(); ++e) {
            out[2 * (m.bt_off[s] + e)] = v[s].a[e].real();
            out[2 * (m.bt_off[s] + e) + 1] = v[s].a[e].imag();
        }
}
static std::vector<CMat> bt_zero(const Model &m) {
    std::vector<CMat> v(m.ns);
    for (int s = 0; s < m.ns; ++s) v[s] = CMat(m.dims[s], m.dims[s]);
    return v;
}
static void bt_add(std::vector<CMat> &a, const std::vector<CMat> &b) {
    for (size_t s = 0; s < a.size(); ++s)
        for (size_t e = 0; e < a[s].a.size(); ++e) a[s].a[e] += b[s].a[e];
}
static std::vector<cd> bt_flat(const std::vector<CMat> &v) {
    std::vector<cd> out;
    for (auto &b : v) out.insert(out.end(), b.a.begin(), b.a.end());
    return out;
}

// One sampled group of a step under randomized qMC: per_order += mean, per_order_std +=
// deviation (complex with zero imaginary part; NaN + NaN i for one sequence).
template <class Integrate>
static void add_randomized(const Model &m, const RandParams &rq, SobolSource staged, int domain, int64_t key,
                           const TopData &d, Integrate integrate, std::vector<CMat> &acc, std::vector<CMat> &acc_std) {
    std::vector<cd> cached;
    auto f = [&](int s) -> std::vector<cd> {
        if (!rq.scramble && !cached.empty()) return cached;  // every sequence is the staged one
        SeqStore st;
        SobolSource src = sequence_source(rq, staged, 2 * d.order, domain, key, d.order, d.n_pts_after, s, st);
        std::vector<cd> v = bt_flat(integrate(src));
        if (!rq.scramble) cached = v;
        return v;
    };
    std::vector<cd> mean;
    std::vector<double> sd;
    mean_std(f, rq, true, mean, sd);
    size_t k = 0;
    for (int s = 0; s < m.ns; ++s)
        for (size_t e = 0; e < acc[s].a.size(); ++e, ++k) {
            acc[s].a[e] += mean[k];
            acc_std[s].a[e] += std::isnan(sd[k]) ? cd(sd[k], sd[k]) : cd(sd[k], 0.0);
        }
}

struct StepResult {
    std::vector<CMat> value;
    std::map<int, std::vector<CMat>> per_order, per_order_std;
};

enum { RQ_BARE = 1, RQ_BOLD = 2, RQ_CORR = 3, RQ_DIFF = 4 };  // stream domains of the scramble bits

static StepResult inchworm_step(const Model &m, const PPGF &P, int jw, int jf, const std::vector<TopData> &td,
                                int64_t N, SobolSource sob, const RandParams &rq) {  // :123-203
    double t_i = 0.0, t_w = jw * m.h, t_f = jf * m.h;
    StepResult r;
    for (auto &d : td) {
        r.per_order[d.order] = bt_zero(m);
        r.per_order_std[d.order] = bt_zero(m);
    }
    for (auto &d : td) {
        if (d.order == 0) {
            std::map<int, FixedNode> fixed{{1, {{IDENTITY, -1, -1}, t_i}},
                                           {2, {{INCH, -1, -1}, t_w}},
                                           {3, {{IDENTITY, -1, -1}, t_f}}};
            Evaluator ev;
            ev.build(m, &P, 0, true, fixed);
            bt_add(r.per_order[0], ev(d.tops, nullptr));
            g_diagram_evals += (int64_t)d.tops.size();
            g_leaves += ev.leaves;
        } else {
            if (N <= 0) continue;
            int d_after = d.n_pts_after, d_before = 2 * d.order - d.n_pts_after;
            std::map<int, FixedNode> fixed{{1, {{IDENTITY, -1, -1}, t_i}},
                                           {d_before + 2, {{INCH, -1, -1}, t_w}},
                                           {2 * d.order + 3, {{IDENTITY, -1, -1}, t_f}}};
            Transform tr = Transform::dsimplex(d_before, d_after, t_i, t_w, t_f);
            add_randomized(
                m, rq, sob, RQ_BOLD, jw, d,
                [&](SobolSource src) { return integrate_group(m, &P, true, d.order, fixed, tr, d.tops, N, src); },
                r.per_order[d.order], r.per_order_std[d.order]);
        }
    }
    r.value = bt_zero(m);
    for (auto &kv : r.per_order) bt_add(r.value, kv.second);
    return r;
}

static StepResult inchworm_step_bare(const Model &m, const std::vector<TopData> &td, int64_t N,
                                     SobolSource sob, const RandParams &rq) {  // :228-304
    double t_i = 0.0, t_f = m.h;
    StepResult r;
    for (auto &d : td) {
        r.per_order[d.order] = bt_zero(m);
        r.per_order_std[d.order] = bt_zero(m);
    }
    for (auto &d : td) {
        if (d.order == 0) {
            std::map<int, FixedNode> fixed{{1, {{IDENTITY, -1, -1}, t_i}}, {2, {{IDENTITY, -1, -1}, t_f}}};
            Evaluator ev;
            ev.build(m, nullptr, 0, false, fixed);
            r.per_order[0] = ev(d.tops, nullptr);
            g_diagram_evals += (int64_t)d.tops.size();
            g_leaves += ev.leaves;
        } else {
            if (N <= 0) continue;
            std::map<int, FixedNode> fixed{{1, {{IDENTITY, -1, -1}, t_i}},
                                           {2 * d.order + 2, {{IDENTITY, -1, -1}, t_f}}};
            Transform tr = Transform::root(2 * d.order, t_i, t_f);
            add_randomized(
                m, rq, sob, RQ_BARE, 0, d,
                [&](SobolSource src) { return integrate_group(m, nullptr, false, d.order, fixed, tr, d.tops, N, src); },
                r.per_order[d.order], r.per_order_std[d.order]);
        }
    }
    r.value = bt_zero(m);
    for (auto &kv : r.per_order) bt_add(r.value, kv.second);
    return r;
}

static std::vector<TopData> bold_top_data(const qiw_run *run, bool external_arc) {
    std::vector<TopData> td;  // :422-452, :957-981
    for (int oi = 0; oi < run->n_orders; ++oi) {
        int order = run->orders[oi];
        int lo = 0, hi = 0;
        if (order != 0) {
            lo = 1;
            hi = 2 * order - 1;
            if (run->n_pts_after_max > 0) hi = std::min(hi, (int)run->n_pts_after_max);
        }
        for (int n = lo; n <= hi; ++n) {
            auto tops = topologies_at_order(order, n, external_arc);
            if (!tops.empty()) td.push_back({order, n, tops});
        }
    }
    return td;
}

static std::vector<TopData> bare_top_data(const qiw_run *run) {
    std::vector<TopData> td;  // :373-392
    for (int oi = 0; oi < run->n_orders_bare; ++oi) {
        int order = run->orders_bare[oi];
        td.push_back({order, 2 * order, topologies_at_order(order, -1, false)});
    }
    return td;
}

// PPGF state exchange in the ABI's layout
static void ppgf_from_table(const Model &m, PPGF &P, int n_known, const double *tab, const double *der0) {
    atomic_ppgf(m, P);
    for (int j = 0; j < m.ntau; ++j)
        for (int s = 0; s < m.ns; ++s) {
            CMat &x = P.tab[j][s];
            for (size_t e = 0; e < x.a.size(); ++e)
                x.a[e] = cd(tab[2 * ((size_t)j * m.bt_size + m.bt_off[s] + e)],
                            tab[2 * ((size_t)j * m.bt_size + m.bt_off[s] + e) + 1]);
        }
    if (P.kind == QIW_GF_INC_SPLINE) {
        for (int s = 0; s < m.ns; ++s) {
            int d = m.dims[s];
            for (int e = 0; e < d * d; ++e) {
                cd d0(der0[2 * (m.bt_off[s] + e)], der0[2 * (m.bt_off[s] + e) + 1]);
                P.spl[s][e].init(m.h, P.tab[0][s].a[e], d0);
                for (int j = 1; j < n_known; ++j) P.spl[s][e].extend(P.tab[j][s].a[e]);
            }
        }
        P.n_known = n_known;
    }
}

static void ppgf_to_table(const Model &m, const PPGF &P, double *tab, double *der0) {
    if (tab)
        for (int j = 0; j < m.ntau; ++j) bt_store(m, P.tab[j], tab + (size_t)2 * j * m.bt_size);
    if (der0) {
        std::memset(der0, 0, sizeof(double) * 2 * m.bt_size);
        if (P.kind == QIW_GF_INC_SPLINE)
            for (int s = 0; s < m.ns; ++s)
                for (size_t e = 0; e < P.spl[s].size(); ++e) {
                    cd v = P.spl[s][e].der_data[0] / P.spl[s][e].h;
                    der0[2 * (m.bt_off[s] + e)] = v.real();
                    der0[2 * (m.bt_off[s] + e) + 1] = v.imag();
                }
    }
}

// ------------------------------------------------------------------------------------
// C API of the oracle
// ------------------------------------------------------------------------------------
extern "C" const char *orc_last_error() { return g_err.c_str(); }
extern "C" void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }
extern "C" void orc_reset_counters() {
    g_leaves = 0;
    g_diagram_evals = 0;
    g_flops = 0;
}
extern "C" void orc_get_counters(int64_t *leaves, int64_t *diagram_evals, double *flops) {
    *leaves = g_leaves;
    *diagram_evals = g_diagram_evals;
    *flops = g_flops;
}

static std::vector<int> all_orders(const qiw_run *run) {
    std::vector<int> v;
    for (int i = 0; i < run->n_orders; ++i) v.push_back(run->orders[i]);
    for (int i = 0; i < run->n_orders_bare; ++i) v.push_back(run->orders_bare[i]);
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    return v;
}

// inchworm! (src/inchworm.jl:332-498).  n_steps_max < 0: all n_tau-2 bold steps; else
// stop after that many bold steps (bounded CPU-baseline samples).
extern "C" int orc_inchworm(const qiw_model *mi, const qiw_run *run, const uint32_t *sob_m,
                            const uint32_t *sob_shift, int n_steps_max, double *P_out,
                            double *P_orders_out, double *P_orders_std_out) {
    Model m;
    if (!load_model(mi, m)) return QIW_E_NOT_SBM;
    if (!(run->N_samples == 0 || (run->N_samples & (run->N_samples - 1)) == 0)) {
        set_err("N_samples must be 0 or a power of 2");
        return QIW_E_NOT_POW2;
    }
    SobolSource sob{sob_m, sob_shift};
    std::vector<int> oall = all_orders(run);
    std::map<int, std::vector<std::vector<CMat>>> P_orders, P_orders_std;  // [order][j]
    for (int o : oall) {
        P_orders[o].assign(m.ntau, bt_zero(m));
        P_orders_std[o].assign(m.ntau, bt_zero(m));
    }
    PPGF P;
    atomic_ppgf(m, P);
    try {
        auto tdb = bare_top_data(run);
        StepResult r = inchworm_step_bare(m, tdb, run->N_samples, sob, rand_params(run));
        set_ppgf(m, P, 1, r.value);
        for (auto &kv : r.per_order) {
            P_orders[kv.first][1] = kv.second;
            P_orders_std[kv.first][1] = r.per_order_std[kv.first];
        }
        auto td = bold_top_data(run, false);
        int steps = 0;
        for (int n = 2; n <= m.ntau - 1; ++n) {  // :475-493
            if (n_steps_max >= 0 && steps >= n_steps_max) break;
            int jw = n - 1, jf = n;
            StepResult s = inchworm_step(m, P, jw, jf, td, run->N_samples, sob, rand_params(run));
            set_ppgf(m, P, jf, s.value);
            normalize_tau(m, P, jf);
            for (auto &kv : s.per_order) {
                P_orders[kv.first][jf] = kv.second;
                P_orders_std[kv.first][jf] = s.per_order_std[kv.first];
            }
            ++steps;
        }
    } catch (std::exception &e) {
        set_err("A block off-diagonal contribution to a pseudo-particle propagator detected");
        return QIW_E_OFFDIAG;
    }
    if (P_out) ppgf_to_table(m, P, P_out, nullptr);
    for (size_t oi = 0; oi < oall.size(); ++oi)
        for (int j = 0; j < m.ntau; ++j) {
            if (P_orders_out) bt_store(m, P_orders[oall[oi]][j], P_orders_out + (size_t)2 * (oi * m.ntau + j) * m.bt_size);
            if (P_orders_std_out)
                bt_store(m, P_orders_std[oall[oi]][j], P_orders_std_out + (size_t)2 * (oi * m.ntau + j) * m.bt_size);
        }
    return 0;
}

// One bold step on a given P state (inchworm_step with tau_w = grid[jw], tau_f = grid[jw+1]).
extern "C" int orc_inchworm_step(const qiw_model *mi, const qiw_run *run, const uint32_t *sob_m,
                                 const uint32_t *sob_shift, int n_known, const double *P_tab,
                                 const double *der0, int jw, double *value_out, double *orders_out) {
    Model m;
    if (!load_model(mi, m)) return QIW_E_NOT_SBM;
    PPGF P;
    ppgf_from_table(m, P, n_known, P_tab, der0);
    auto td = bold_top_data(run, false);
    try {
        StepResult r =
            inchworm_step(m, P, jw, jw + 1, td, run->N_samples, SobolSource{sob_m, sob_shift}, rand_params(run));
        if (value_out) bt_store(m, r.value, value_out);
        if (orders_out)
            for (int oi = 0; oi < run->n_orders; ++oi) {
                auto it = r.per_order.find(run->orders[oi]);
                std::vector<CMat> z = bt_zero(m);
                bt_store(m, it != r.per_order.end() ? it->second : z, orders_out + (size_t)2 * oi * m.bt_size);
            }
    } catch (std::exception &e) {
        set_err("A block off-diagonal contribution to a pseudo-particle propagator detected");
        return QIW_E_OFFDIAG;
    }
    return 0;
}

extern "C" int orc_inchworm_step_bare(const qiw_model *mi, const qiw_run *run, const uint32_t *sob_m,
                                      const uint32_t *sob_shift, double *value_out, double *orders_out) {
    Model m;
    if (!load_model(mi, m)) return QIW_E_NOT_SBM;
    auto td = bare_top_data(run);
    try {
        StepResult r = inchworm_step_bare(m, td, run->N_samples, SobolSource{sob_m, sob_shift}, rand_params(run));
        if (value_out) bt_store(m, r.value, value_out);
        if (orders_out)
            for (int oi = 0; oi < run->n_orders_bare; ++oi) bt_store(m, r.per_order[run->orders_bare[oi]], orders_out + (size_t)2 * oi * m.bt_size);
    } catch (std::exception &e) {
        set_err("A block off-diagonal contribution to a pseudo-particle propagator detected");
        return QIW_E_OFFDIAG;
    }
    return 0;
}

// correlator_2p at one tau index jA (src/inchworm.jl:805-890), without the final 1/Z:
// the value and the sum of the groups' deviations.  One set of scrambled sequences per
// operator pair serves all its tau rows (stream key = pair index).
static cd correlator_at(const Model &m, const PPGF &P, int pair_idx, int jA, const std::vector<TopData> &td,
                        int64_t N, SobolSource sob, const RandParams &rq, double *result_std) {
    double t_B = 0.0, t_A = jA * m.h, t_f = m.beta;
    cd result = 0;
    *result_std = 0.0;
    for (auto &d : td) {
        cd contrib = 0;
        if (d.order == 0) {
            std::map<int, FixedNode> fixed{{1, {{OPERATOR, pair_idx, 2}, t_B}},
                                           {2, {{OPERATOR, pair_idx, 1}, t_A}},
                                           {3, {{IDENTITY, -1, -1}, t_f}}};
            Evaluator ev;
            ev.build(m, &P, 0, true, fixed);
            const std::vector<CMat> &r = ev(d.tops, nullptr);
            for (int s = 0; s < m.ns; ++s)
                for (int i = 0; i < m.dims[s]; ++i) contrib += r[s](i, i);
            g_diagram_evals += (int64_t)d.tops.size();
            g_leaves += ev.leaves;
        } else {
            if (N <= 0) continue;
            int d_after = d.n_pts_after, d_before = 2 * d.order - d.n_pts_after;
            std::map<int, FixedNode> fixed{{1, {{OPERATOR, pair_idx, 2}, t_B}},
                                           {d_before + 2, {{OPERATOR, pair_idx, 1}, t_A}},
                                           {2 * d.order + 3, {{IDENTITY, -1, -1}, t_f}}};
            Transform tr = Transform::dsimplex(d_before, d_after, t_B, t_A, t_f);
            std::vector<cd> cached;
            auto f = [&](int q) -> std::vector<cd> {
                if (!rq.scramble && !cached.empty()) return cached;  // every sequence is the staged one
                SeqStore st;
                SobolSource src =
                    sequence_source(rq, sob, 2 * d.order, RQ_CORR, pair_idx, d.order, d.n_pts_after, q, st);
                std::vector<CMat> r = integrate_group(m, &P, true, d.order, fixed, tr, d.tops, N, src);
                cd tr_sum = 0;
                for (int s = 0; s < m.ns; ++s)
                    for (int i = 0; i < m.dims[s]; ++i) tr_sum += r[s](i, i);
                cached = {tr_sum};
                return cached;
            };
            std::vector<cd> mean;
            std::vector<double> sd;
            mean_std(f, rq, false, mean, sd);
            contrib = mean[0];
            *result_std += sd[0];
        }
        result += contrib;
    }
    return result;
}

// correlator_2p (src/inchworm.jl:918-1052).  tau_first/tau_count restrict the tau
// indices evaluated (bounded samples); the rest of corr_out is left untouched.
extern "C" int orc_correlator_2p_std(const qiw_model *mi, const qiw_run *run, const uint32_t *sob_m,
                                     const uint32_t *sob_shift, int n_known, const double *P_tab,
                                     const double *der0, int tau_first, int tau_count, double *corr_out,
                                     double *corr_std_out) {
    Model m;
    if (!load_model(mi, m)) return QIW_E_NOT_SBM;
    PPGF P;
    ppgf_from_table(m, P, n_known, P_tab, der0);
    auto td = bold_top_data(run, true);
    if (td.empty()) return 0;
    cd Z = partition_function(m, P);
    SobolSource sob{sob_m, sob_shift};
    if (tau_count < 0) tau_count = m.ntau;
    try {
        for (int c = 0; c < m.ncorr; ++c) {
            for (int j = tau_first; j < std::min(m.ntau, tau_first + tau_count); ++j) {
                cd v;
                double sd = 0.0;
                const RandParams rq = rand_params(run);
                if (j == 0) {
                    if (td[0].order != 0) continue;
                    std::vector<TopData> td0(td.begin(), td.begin() + 1);
                    v = correlator_at(m, P, c, 0, td0, run->N_samples, sob, rq, &sd);
                } else {
                    v = correlator_at(m, P, c, j, td, run->N_samples, sob, rq, &sd);
                }
                v = v / Z;
                const cd vs = sd / Z;  // (result, result_std) ./ Z
                corr_out[2 * ((size_t)c * m.ntau + j)] = v.real();
                corr_out[2 * ((size_t)c * m.ntau + j) + 1] = v.imag();
                if (corr_std_out) {
                    corr_std_out[2 * ((size_t)c * m.ntau + j)] = std::isnan(sd) ? sd : vs.real();
                    corr_std_out[2 * ((size_t)c * m.ntau + j) + 1] = std::isnan(sd) ? sd : vs.imag();
                }
            }
        }
    } catch (std::exception &e) {
        set_err("A block off-diagonal contribution to a pseudo-particle propagator detected");
        return QIW_E_OFFDIAG;
    }
    return 0;
}

extern "C" int orc_correlator_2p(const qiw_model *mi, const qiw_run *run, const uint32_t *sob_m,
                                 const uint32_t *sob_shift, int n_known, const double *P_tab,
                                 const double *der0, int tau_first, int tau_count, double *corr_out) {
    return orc_correlator_2p_std(mi, run, sob_m, sob_shift, n_known, P_tab, der0, tau_first, tau_count, corr_out,
                                 nullptr);
}

// ------------------------------------------------------------------------------------
// Differential inchworm -- src/inchworm.jl:528-778
// ------------------------------------------------------------------------------------
static std::vector<TopData> diff_top_data(const qiw_run *run) {  // :711-727
    std::vector<TopData> td;
    for (int oi = 0; oi < run->n_orders; ++oi) {
        const int order = run->orders[oi];
        if (order == 0) continue;  // no 0-th order self-energy
        auto tops = topologies_at_order(order, 1, false);
        if (!tops.empty()) td.push_back({order, 1, tops});
    }
    return td;
}

// The self-energy container Σ = zero(expansion.P).  set_ppgf! on a grid-stored Σ writes
// the row of tau_f_prev; on an IncSplineImaginaryTimeGF it appends (extend!, src/spline_gf.jl
// :392-402), so the k-th value lands on row k: one row later than tau_f_prev.  Σ(t, t')
// of a grid point pair reads the row of t - t' (time-translation-invariant storage).
struct SigmaTable {
    std::vector<std::vector<CMat>> tab;  // [n_tau][ns]
    int len = 1;                         // values held by the incremental container
    void store(const Model &m, int a, const std::vector<CMat> &v) {
        if (m.ppgf_kind == QIW_GF_INC_SPLINE) tab[len++] = v;
        else tab[a] = v;
    }
};

// One differential step (diff_inchworm_step!, :528-636) from tau_f_prev = grid[a] to
// grid[a + 1]: the self-energy at (tau_f_prev, tau_i) order by order, then
// P(tau_f) = P_prev + h (-i sum_j w_j Σ(tau_f_prev - tau_j) P(tau_j) - H P_prev) with the
// trapezoid weights w of kd.integrate over grid[0..a] and H = diag(E + lambda).  kd.integrate
// integrates over contour time, dt = -i dtau on the imaginary branch; the reference's golden
// /diff_inchworm/P pins both that factor and the row shift of the incremental Σ container
// (without either, the run misses it by 1e-3 or more).
static std::vector<CMat> diff_step(const Model &m, const PPGF &P, SigmaTable &S, int a, const std::vector<TopData> &td,
                                   int64_t N, SobolSource sob, const RandParams &rq,
                                   std::map<int, std::vector<CMat>> &per_order,
                                   std::map<int, std::vector<CMat>> &per_order_std) {
    const double t_i = 0.0, t_fp = a * m.h;
    per_order.clear();
    per_order_std.clear();
    for (auto &d : td) {
        per_order[d.order] = bt_zero(m);
        per_order_std[d.order] = bt_zero(m);
    }
    for (auto &d : td) {
        if (d.order == 1) {  // exact: no variable times
            std::map<int, FixedNode> fixed{{1, {{PAIR, -1, -1}, t_i}}, {2, {{PAIR, -1, -1}, t_fp}}};
            Evaluator ev;
            ev.build(m, &P, 1, true, fixed);
            per_order[1] = ev(d.tops, nullptr);
            g_diagram_evals += (int64_t)d.tops.size();
            g_leaves += ev.leaves;
            continue;
        }
        if (N <= 0) continue;
        const int dv = 2 * d.order - 2;
        std::map<int, FixedNode> fixed{{1, {{PAIR, -1, -1}, t_i}}, {dv + 2, {{PAIR, -1, -1}, t_fp}}};
        Transform tr = Transform::root(dv, t_i, t_fp);
        add_randomized(
            m, rq, sob, RQ_DIFF, a, d,
            [&](SobolSource src) { return integrate_group(m, &P, true, d.order, fixed, tr, d.tops, N, src, 0, -1, dv); },
            per_order[d.order], per_order_std[d.order]);
    }
    std::vector<CMat> sig = bt_zero(m);
    for (auto &kv : per_order) bt_add(sig, kv.second);
    S.store(m, a, sig);
    std::vector<CMat> out = bt_zero(m);
    for (int s = 0; s < m.ns; ++s) {
        const int dm = m.dims[s];
        const CMat &Pp = P.tab[a][s];
        CMat conv(dm, dm), tmp(dm, dm);
        for (int j = 0; j <= a && a > 0; ++j) {  // kd.integrate: trapezoid over grid[0..a]
            const double w = (j == 0 || j == a) ? 0.5 * m.h : m.h;
            matmul(tmp, S.tab[a - j][s], P.tab[j][s]);
            for (int e = 0; e < dm * dm; ++e) conv.a[e] += w * tmp.a[e];
        }
        for (int r = 0; r < dm; ++r)
            for (int c = 0; c < dm; ++c) {
                const cd rhs = -I * conv(r, c) - (m.E[s][r] + m.lambda0) * Pp(r, c);
                out[s](r, c) = Pp(r, c) + m.h * rhs;
            }
    }
    return out;
}

// diff_inchworm! (:666-778).  n_steps_max < 0: all n_tau - 1 steps.  Sigma_orders_out
// [n_orders][n_tau][bt] for run->orders (order 0: zero), written at tau_f of each step as
// the reference's set_ppgf!(Σ_orders[order], tau_i, tau_f, ...) does.
extern "C" int orc_diff_inchworm(const qiw_model *mi, const qiw_run *run, const uint32_t *sob_m,
                                 const uint32_t *sob_shift, int n_steps_max, double *P_out, double *Sigma_orders_out,
                                 double *Sigma_orders_std_out) {
    Model m;
    if (!load_model(mi, m)) return QIW_E_NOT_SBM;
    if (!(run->N_samples == 0 || (run->N_samples & (run->N_samples - 1)) == 0)) {
        set_err("N_samples must be 0 or a power of 2");
        return QIW_E_NOT_POW2;
    }
    SobolSource sob{sob_m, sob_shift};
    const RandParams rq = rand_params(run);
    PPGF P;
    atomic_ppgf(m, P);
    SigmaTable S;
    S.tab.assign(m.ntau, bt_zero(m));
    auto td = diff_top_data(run);
    const int no = run->n_orders;
    std::vector<std::vector<std::vector<CMat>>> so(no, std::vector<std::vector<CMat>>(m.ntau, bt_zero(m)));
    std::vector<std::vector<std::vector<CMat>>> ss = so;
    try {
        std::map<int, std::vector<CMat>> per, per_std;
        for (int a = 0; a + 1 < m.ntau; ++a) {  // :752-774, tau_f_prev = grid[a], tau_f = grid[a + 1]
            if (n_steps_max >= 0 && a >= n_steps_max) break;
            std::vector<CMat> Pn = diff_step(m, P, S, a, td, run->N_samples, sob, rq, per, per_std);
            set_ppgf(m, P, a + 1, Pn);
            for (int oi = 0; oi < no; ++oi) {
                auto it = per.find(run->orders[oi]);
                if (it == per.end()) continue;
                so[oi][a + 1] = it->second;
                ss[oi][a + 1] = per_std[run->orders[oi]];
            }
        }
    } catch (std::exception &e) {
        set_err("A block off-diagonal contribution to a pseudo-particle propagator detected");
        return QIW_E_OFFDIAG;
    }
    if (P_out) ppgf_to_table(m, P, P_out, nullptr);
    for (int oi = 0; oi < no; ++oi)
        for (int j = 0; j < m.ntau; ++j) {
            if (Sigma_orders_out) bt_store(m, so[oi][j], Sigma_orders_out + (size_t)2 * (oi * m.ntau + j) * m.bt_size);
            if (Sigma_orders_std_out)
                bt_store(m, ss[oi][j], Sigma_orders_std_out + (size_t)2 * (oi * m.ntau + j) * m.bt_size);
        }
    return 0;
}

// TopologyEvaluator at explicit variable times (unit tests of the evaluator).
// kind 0: bold inchworm nodes {1: Id(t_i), d_before+2: Inch(t_w), 2k+3: Id(t_f)};
// kind 1: bare {1: Id(t_i), 2k+2: Id(t_f)}.  The P state is the atomic one.
extern "C" int orc_eval_topologies(const qiw_model *mi, int kind, int order, int d_before, double t_i,
                                   double t_w, double t_f, int n_top, const int32_t *pairs,
                                   const int32_t *parity, const double *vt, double *out) {
    Model m;
    if (!load_model(mi, m)) return QIW_E_NOT_SBM;
    PPGF P;
    atomic_ppgf(m, P);
    std::vector<Topology> tops(n_top);
    for (int t = 0; t < n_top; ++t) {
        tops[t].order = order;
        tops[t].parity = parity[t];
        for (int a = 0; a < order; ++a)
            tops[t].pairs.push_back({pairs[(size_t)t * 2 * order + 2 * a], pairs[(size_t)t * 2 * order + 2 * a + 1]});
    }
    std::map<int, FixedNode> fixed;
    if (kind == 0) {
        fixed = {{1, {{IDENTITY, -1, -1}, t_i}},
                 {d_before + 2, {{INCH, -1, -1}, t_w}},
                 {2 * order + 3, {{IDENTITY, -1, -1}, t_f}}};
    } else {
        fixed = {{1, {{IDENTITY, -1, -1}, t_i}}, {2 * order + 2, {{IDENTITY, -1, -1}, t_f}}};
    }
    Evaluator ev;
    ev.build(m, &P, order, kind == 0, fixed);
    try {
        const std::vector<CMat> &r = ev(tops, vt);
        bt_store(m, r, out);
    } catch (std::exception &e) {
        set_err("A block off-diagonal contribution to a pseudo-particle propagator detected");
        return QIW_E_OFFDIAG;
    }
    return 0;
}

// A DLR lookup for unit tests: g(t1, t2) of the pole sum (w[k], c[k]) with statistics xi.
extern "C" int orc_dlr_eval(int rank, const double *w, const double *c, int xi, double beta, double t1, double t2,
                            double *out) {
    ScalarGF g;
    g.kind = QIW_GF_DLR;
    g.xi = xi;
    g.beta = beta;
    for (int k = 0; k < rank; ++k) {
        g.dlr_w.push_back(w[k]);
        g.dlr_c.push_back(cd(c[2 * k], c[2 * k + 1]));
    }
    const cd v = g(t1, t2);
    out[0] = v.real();
    out[1] = v.imag();
    return 0;
}

// Propagator lookups for unit tests: kind = QIW_GF_CUBIC_SPLINE / GRID_LINEAR on a
// scalar table; QIW_GF_INC_SPLINE builds an IncrementalSpline from vals (n_known
// values, der0) and evaluates it.
extern "C" int orc_gf_eval(int kind, int n, double beta, const double *vals, int n_known,
                           const double *der0, int xi, int npts, const double *t1, const double *t2,
                           double *out) {
    if (kind == QIW_GF_INC_SPLINE) {
        IncSpline sp;
        double h = beta / (n - 1);
        sp.init(h, cd(vals[0], vals[1]), cd(der0[0], der0[1]));
        for (int j = 1; j < n_known; ++j) sp.extend(cd(vals[2 * j], vals[2 * j + 1]));
        for (int i = 0; i < npts; ++i) {
            cd v = t1[i] >= t2[i] ? sp.eval(t1[i] - t2[i]) : double(xi) * sp.eval(beta + t1[i] - t2[i]);
            out[2 * i] = v.real();
            out[2 * i + 1] = v.imag();
        }
        return 0;
    }
    ScalarGF g;
    g.build(vals, n, kind, xi, beta);
    for (int i = 0; i < npts; ++i) {
        cd v = g(t1[i], t2[i]);
        out[2 * i] = v.real();
        out[2 * i + 1] = v.imag();
    }
    return 0;
}
