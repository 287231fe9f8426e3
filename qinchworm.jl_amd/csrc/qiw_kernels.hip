// qiw_kernels.hip -- gfx950 kernels of libqiw: Sobol points and simplex maps, the
// trie-walking diagram evaluator, and the device-side step bookkeeping.
//
// Work mapping of the evaluators: all lanes of a wavefront walk the same precompiled
// configuration trie (the depth-first traversal of src/topology_eval.jl:454-556, compiled on
// the host into program words), so control flow is uniform and program words are broadcast
// from one coalesced block load.  The flat scalar walk and the vector walk give each lane
// one qMC sample (64 consecutive samples of one (order, n_pts_after) group per wavefront)
// and keep the sample's hot operands in registers; the matrix-core and LDS block walks give
// the wavefront one sample at a time and spread its matrix products over the lanes.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/qiw.h"
#include "qiw_device.h"

// ---------------------------------------------------------------------------------------
// Sobol points (ScrambledSobolSeq next!, src/scrambled_sobol.jl:158-173) in closed form:
// x_i = x_0 xor (xor over the set bits j of gray(i) of m[:, j]), then ldexp(x, -32).
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t sobol_word(const uint32_t *__restrict__ m, uint32_t x0, int d,
                                                uint64_t i) {
    uint64_t g = i ^ (i >> 1);
    uint32_t x = x0;
    const uint32_t *md = m + d * 32;
#pragma unroll 4
    for (int j = 0; j < 32; ++j)
        if ((g >> j) & 1ull) x ^= md[j];
    return x;
}

__global__ void k_sobol_points(const uint32_t *__restrict__ m, const uint32_t *__restrict__ x0, int dim,
                               int64_t first, int64_t n, double *__restrict__ out) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    for (int d = 0; d < dim; ++d)
        out[r * dim + d] = ldexp((double)sobol_word(m, x0[d], d, (uint64_t)(first + r)), -32);
}

// Unscaled double-simplex coordinates u~ of DoubleSimplexRootTransform
// (src/qmc_integrate.jl:425-449): the first d_after coordinates form the simplex after
// tau_w, the remaining d_before the one before it.  RootTransform is d_before = 0.
// Stored [d][N] so that a wavefront of consecutive samples reads contiguous rows.
__device__ __forceinline__ void root_map(const double *x, int d_after, int d_before, double *u) {
    int d = d_after + d_before;
    if (d_after > 0) {
        u[0] = pow(x[0], 1.0 / d_after);
        for (int i = 1; i < d_after; ++i) u[i] = u[i - 1] * pow(x[i], 1.0 / (d_after - i));
    }
    if (d_before > 0) {
        u[d_after] = pow(x[d_after], 1.0 / d_before);
        for (int i = d_after + 1; i < d; ++i) u[i] = u[i - 1] * pow(x[i], 1.0 / (d - i));
    }
}

__global__ void k_simplex_coords(const uint32_t *__restrict__ m, const uint32_t *__restrict__ x0, int d_before,
                                 int d_after, int64_t n, double *__restrict__ out) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double x[2 * QIW_KMAX], u[2 * QIW_KMAX];
    int d = d_before + d_after;
    for (int k = 0; k < d; ++k) x[k] = ldexp((double)sobol_word(m, x0[k], k, (uint64_t)r), -32);
    root_map(x, d_after, d_before, u);
    for (int k = 0; k < d; ++k) out[(int64_t)k * n + r] = u[k];
}

// Self-energy groups (RootTransform of the 2k - 2 variable times on [tau_i, tau_f_prev],
// src/inchworm.jl:573-577): row 0 = 1 and row 2k - 1 = 0 pin the first and last pair
// positions to tau_f_prev and tau_i; rows 1..2k-2 hold the descending root-map coordinates.
__device__ __forceinline__ void sigma_row_values(const double *x, int order, double *u) {
    const int dv = 2 * order - 2;
    double r[2 * QIW_KMAX];
    root_map(x, 0, dv, r);
    u[0] = 1.0;
    for (int k = 0; k < dv; ++k) u[1 + k] = r[k];
    u[dv + 1] = 0.0;
}

__global__ void k_sigma_coords(const uint32_t *__restrict__ m, const uint32_t *__restrict__ x0, int order, int64_t n,
                               double *__restrict__ out) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double x[2 * QIW_KMAX], u[2 * QIW_KMAX];
    for (int k = 0; k < 2 * order - 2; ++k) x[k] = ldexp((double)sobol_word(m, x0[k], k, (uint64_t)r), -32);
    sigma_row_values(x, order, u);
    for (int k = 0; k < 2 * order; ++k) out[(int64_t)k * n + r] = u[k];
}

// Points of the unit cube for the learned warp: Sobol coordinates 0..dim-1 of sample r in
// row r (stride `stride`, unused coordinates 0).
__global__ void k_sobol_cube(const uint32_t *__restrict__ m, const uint32_t *__restrict__ x0, int dim, int stride,
                             int64_t n, double *__restrict__ out) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    for (int k = 0; k < stride; ++k)
        out[r * stride + k] = (k < dim) ? ldexp((double)sobol_word(m, x0[k], k, (uint64_t)r), -32) : 0.0;
}

// Warped points -> simplex coordinates (rows 0..d-1) and the per-sample weight
// exp(log |dz/dx|) (row d) of a weighted group.
__global__ void k_warp_coords(const double *__restrict__ z, int stride, const double *__restrict__ logj,
                              int d_before, int d_after, int64_t n, double *__restrict__ out) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double u[2 * QIW_KMAX];
    const int d = d_before + d_after;
    root_map(z + r * stride, d_after, d_before, u);
    for (int k = 0; k < d; ++k) out[(int64_t)k * n + r] = u[k];
    out[(int64_t)d * n + r] = exp(logj[r]);
}

// Full transform to contour times (tests of the simplex map).
__global__ void k_simplex_points(const uint32_t *__restrict__ m, const uint32_t *__restrict__ x0, int d_before,
                                 int d_after, double t_i, double t_w, double t_f, int64_t n,
                                 double *__restrict__ out) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double x[2 * QIW_KMAX], u[2 * QIW_KMAX];
    int d = d_before + d_after;
    for (int k = 0; k < d; ++k) x[k] = ldexp((double)sobol_word(m, x0[k], k, (uint64_t)r), -32);
    if (d_after == 0) {  // RootTransform on [t_i, t_f]
        root_map(x, d, 0, u);
        for (int k = 0; k < d; ++k) out[r * d + k] = t_i + u[k] * (t_f - t_i);
        return;
    }
    root_map(x, d_after, d_before, u);
    for (int k = 0; k < d_after; ++k) out[r * d + k] = u[k] * (t_f - t_w) + t_w;
    for (int k = d_after; k < d; ++k) out[r * d + k] = u[k] * (t_w - t_i) + t_i;
}

// ---------------------------------------------------------------------------------------
// Propagator lookups
// ---------------------------------------------------------------------------------------
// A grid-stored function of the time difference (the bold i*P table, a grid-stored i*Delta)
// at two contour times t1 >= t2, as Keldysh.jl's interpolate! evaluates a grid container
// (kd.interpolate! on expansion.P, src/topology_eval.jl:368; the call of an ImaginaryTimeGF
// pair propagator, :414): bilinear over the grid cell (i1, i2) holding (t1, t2), with the
// stored G[i1, i2] = g(i1 - i2), and linear in t1 - t2 inside a diagonal cell (i1 == i2).
// With a, b the fractions of t1, t2 in their grid intervals and d = i1 - i2 >= 0 both are
//     g(d) + a (1 - b) (g(d+1) - g(d)) - (1 - a) b (g(d) - g(d-1)),   g(-1) = 2 g(0) - g(1),
// i.e. row d plus wp times the forward and minus wm times the backward difference.
struct GridCell {
    int d;
    double wp, wm;  // a (1 - b), (1 - a) b
};
__device__ __forceinline__ GridCell grid_cell(double t1, double t2, double inv_h, int n_tau) {
    t1 = t1 > t2 ? t1 : t2;  // ordering clamp (src/topology_eval.jl:362-365, :408-411)
    const double x1 = t1 * inv_h, x2 = t2 * inv_h;
    int i1 = (int)floor(x1), i2 = (int)floor(x2);
    i1 = i1 < 0 ? 0 : (i1 > n_tau - 2 ? n_tau - 2 : i1);
    i2 = i2 < 0 ? 0 : (i2 > n_tau - 2 ? n_tau - 2 : i2);
    const double a = x1 - (double)i1, b = x2 - (double)i2;
    return GridCell{i1 - i2, a * (1.0 - b), (1.0 - a) * b};
}
// Value at cell c from g(d - 1) (any finite value when d = 0), g(d) and g(d + 1).
__device__ __forceinline__ double cell_value(const GridCell &c, double gm, double g0, double gp) {
    const double dm = c.d > 0 ? g0 - gm : gp - g0;
    return g0 + c.wp * (gp - g0) - c.wm * dm;
}

// Hybridization tables (QiwDeltaTables): per arc, the spline interval j and fraction u of
// the time difference, and the grid cell; a table is evaluated by its kind (grid_mask).
struct ArcPos {
    int j;
    double u;
    double z;  // the time difference itself (DLR tables)
    GridCell c;
};
__device__ __forceinline__ ArcPos arc_pos(double t_tail, double t_head, double inv_h, int n_tau) {
    ArcPos p;
    double z = t_tail - t_head;
    z = z > 0.0 ? z : 0.0;
    const double x = z * inv_h;
    int j = (int)floor(x);
    j = j < 0 ? 0 : (j > n_tau - 2 ? n_tau - 2 : j);
    p.j = j;
    p.u = x - (double)j;
    p.z = z;
    p.c = grid_cell(t_tail, t_head, inv_h, n_tau);
    return p;
}

// The fermionic kernel of the DLR, e^{-w tau} / (1 + e^{-w beta}) for 0 <= tau <= beta, in
// the form whose exponents are never positive.
__device__ __forceinline__ double dlr_kernel(double tau, double w, double beta) {
    return w >= 0.0 ? exp(-w * tau) / (1.0 + exp(-w * beta)) : exp(w * (beta - tau)) / (1.0 + exp(w * beta));
}
// i*Delta of table tab at arc position p: a natural cubic spline in the time difference
// (SplineInterpolatedGF, src/spline_gf.jl:194-219), coefficients (c0..c3) per interval, a
// grid container, (g(d), g(d+1) - g(d), g(d) - g(d-1), 0) per index d, or a DLR pole sum
// (keldysh_dlr.DLRImaginaryTimeGF, src/keldysh_dlr.jl:172-176).
// DLR = false leaves out the pole sums (a launch with no DLR table: fewer registers).
template <bool DLR = true>
__device__ __forceinline__ double delta_value(const QiwDeltaTables &dt, int tab, const ArcPos &p) {
    if (DLR && ((dt.dlr_mask >> tab) & 1)) {
        const double *c = dt.dlr + (int64_t)tab * dt.dlr_rank * 2;
        double s = 0.0;
        for (int k = 0; k < dt.dlr_rank; ++k) s = fma(c[2 * k + 1], dlr_kernel(p.z, c[2 * k], dt.beta), s);
        return s;
    }
    if ((dt.grid_mask >> tab) & 1) {
        const double *c = dt.coef + ((int64_t)tab * (dt.n_tau - 1) + p.c.d) * 4;
        return c[0] + p.c.wp * c[1] - p.c.wm * c[2];
    }
    const double *c = dt.coef + ((int64_t)tab * (dt.n_tau - 1) + p.j) * 4;
    return c[0] + p.u * (c[1] + p.u * (c[2] + p.u * c[3]));
}

// ---------------------------------------------------------------------------------------
// The diagram evaluator for models whose sectors are all one-dimensional.
// ---------------------------------------------------------------------------------------

// Program words reach the wavefront as 64-word blocks: lane l loads word base+l with one
// coalesced vector load and each word is broadcast with v_readlane.  A block is loaded
// when the previous one is used up, so the memory latency is paid once per 64 words.
struct ProgStream {
    const uint64_t *__restrict__ base;
    uint64_t blk;
    int idx, lane;
    __device__ __forceinline__ void init(const uint64_t *p, int l) {
        base = p;
        lane = l;
        blk = p[l];
        idx = 0;
    }
    __device__ __forceinline__ uint64_t next() {
        if (idx == QIW_WAVE) {
            base += QIW_WAVE;
            blk = base[lane];
            idx = 0;
        }
        const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)blk, idx);
        const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(blk >> 32), idx);
        ++idx;
        return ((uint64_t)hi << 32) | lo;
    }
};

// Register-resident state of one lane (one qMC sample).  Array indices are compile-time
// constants inside the walk, or uniform program fields read through the index mode.
// Register vectors indexed by a uniform (SGPR) index: the compiler reads them with
// s_set_gpr_idx_on / v_mov / s_set_gpr_idx_off, a few instructions per pick instead of a
// compare-and-select chain over every element.
template <typename T, int N>
struct RegVec;
template <typename T>
struct RegVec<T, 8> {
    typedef T type __attribute__((ext_vector_type(8)));
};
template <typename T>
struct RegVec<T, 16> {
    typedef T type __attribute__((ext_vector_type(16)));
};
template <typename T>
struct RegVec<T, 32> {
    typedef T type __attribute__((ext_vector_type(32)));
};
constexpr int regvec_size(int n) { return n <= 8 ? 8 : 16; }

// v[i] for a uniform i as a v_cndmask chain over registers.  The empty asm makes each
// element an opaque register value, so the select chain is not rewritten into an indexed
// load from a private (scratch) copy of the array.
template <int N>
__device__ __forceinline__ double pick(const double (&v)[N], int i) {
    double r = v[0];
#pragma unroll
    for (int q = 1; q < N; ++q) {
        double x = v[q];
        asm volatile("" : "+v"(x));
        r = (i == q) ? x : r;
    }
    return r;
}

// pick for any register array type
template <typename T, int N>
__device__ __forceinline__ T pickv(const T (&v)[N], int i) {
    T r = v[0];
#pragma unroll
    for (int q = 1; q < N; ++q) {
        T x = v[q];
        asm volatile("" : "+v"(x));
        r = (i == q) ? x : r;
    }
    return r;
}

// ---------------------------------------------------------------------------------------
// The flat walk of scalar programs (all sectors 1x1).
//
// One loop over 32-bit program words, no recursion: every word is one trie node with its
// level L stated in the word.  The running product `cur` flows down each chain of the
// trie; a node with several children saves it in the slot of its children's level (PUSH)
// and every child but the first restores it (POP).  The propagator values of the inner
// segments (levels 3..NM-1, read by every edge), the arc values and the outputs of the lane's
// sample stay in registers, read with uniform indices through the VGPR index mode.  LDS
// holds what is read once per topology, per restored branch or per initial sector: the
// sampled contour times, the saved products and the outputs.
// ---------------------------------------------------------------------------------------

// 32-bit program words in 64-word blocks: lane l loads word base+l of the block with one
// coalesced load and each word is broadcast with v_readlane.  The last word of every block
// is NEXT (the host compiler puts it there), so reading a word needs no bounds test.
struct ProgStream32 {
    const uint32_t *__restrict__ p;
    uint32_t blk;
    int idx, lane;
    __device__ __forceinline__ void init(const uint32_t *p0, int l) {
        p = p0;
        lane = l;
        blk = p0[l];
        idx = 0;
    }
    __device__ __forceinline__ uint32_t next() { return __builtin_amdgcn_readlane(blk, idx++); }
    __device__ __forceinline__ void next_block() {
        p += QIW_WAVE;
        blk = p[lane];
        idx = 0;
    }
};

template <int KM, int SM>
struct FlatShape {
    static constexpr int NM = 2 * KM + 3;
    static constexpr int NH = (NM - 3) * SM;   // inner entries (NM-1-L)*SM + s, L = 3..NM-1
    static constexpr int NC = (NH + 15) / 16;  // register vectors of 16 entries
    static_assert(NC >= 1 && NC <= 6, "inner propagator entries beyond 6 register vectors");
    // at most 32 entries: low and high words in one pair of 32-dword vectors
    static constexpr bool SPLIT = NH <= 32;
};

// The inner propagator entries of one lane live in up to six register vectors of 16 doubles
// (the largest VGPR tuple), declared as separate locals P0..P5: a uniform choice of vector
// then selects between register values and never becomes a dynamically indexed (scratch)
// address.  Entry e is element e % 16 of vector e / 16 (e uniform; constant in unrolled
// loops).  NC_ is the number of vectors in use.
//
// Up to 32 entries (SPLIT) are held instead as two vectors of 32 dwords, the low words of
// the entries in PL and their high words in PH: one uniform index then reads both halves of
// an entry from the same element of the two vectors (two indexed moves), where two vectors
// of 16 doubles need a read from each and a select.  Wider shapes keep the double vectors:
// the dword pairs cost them registers (and occupancy) that the select does not.
#define QIW_PIN_SET(NC_, e, x)                              \
    do {                                                    \
        const int c_ = (e) >> 4, i_ = (e)&15;               \
        if (F::SPLIT) {                                     \
            const unsigned long long b_ = (unsigned long long)__double_as_longlong(x); \
            PL[(e)&31] = (uint32_t)b_;                      \
            PH[(e)&31] = (uint32_t)(b_ >> 32);              \
        } else if (c_ == 0) P0[i_] = (x);                   \
        else if ((NC_) > 1 && c_ == 1) P1[i_] = (x);        \
        else if ((NC_) > 2 && c_ == 2) P2[i_] = (x);        \
        else if ((NC_) > 3 && c_ == 3) P3[i_] = (x);        \
        else if ((NC_) > 4 && c_ == 4) P4[i_] = (x);        \
        else if ((NC_) > 5) P5[i_] = (x);                   \
    } while (0)
// Entry of the EDGE word w (bits 0-7: twice the entry): vector (w >> 5) & 7, element
// (w >> 1) & 15.
#define QIW_PIN_GET(NC_, w, out)                            \
    do {                                                    \
        const int c_ = ((w) >> 5) & 7, i_ = ((w) >> 1) & 15; \
        if (F::SPLIT) {                                     \
            const int j_ = ((w) >> 1) & 31;                 \
            const uint32_t lo_ = PL[j_], hi_ = PH[j_];      \
            out = __longlong_as_double((long long)(((unsigned long long)hi_ << 32) | lo_)); \
        } else if ((NC_) == 1 || ((NC_) == 2 ? ((w) & 32) == 0 : c_ == 0)) out = P0[i_]; \
        else if ((NC_) == 2 || c_ == 1) out = P1[i_];       \
        else if ((NC_) == 3 || c_ == 2) out = P2[i_];       \
        else if ((NC_) == 4 || c_ == 3) out = P3[i_];       \
        else if ((NC_) == 5 || c_ == 4) out = P4[i_];       \
        else out = P5[i_];                                  \
    } while (0)

template <int KM, int SM, int TM, bool DLR>
__global__ __launch_bounds__(QIW_WAVE) void k_eval_flat_t(EvalArgs a) {
    using F = FlatShape<KM, SM>;
    constexpr int NM = F::NM, NC = F::NC;
    // arc values: slots arc * TM + table in one register vector of up to 16 doubles, and a
    // second one for slots 16..31 when KM * TM > 16 (the order-6 instantiation)
    constexpr int NW = regvec_size(KM * TM < 16 ? KM * TM : 16);
    constexpr bool W2 = KM * TM > 16;
    static_assert(KM * TM <= 32, "arc-value slots beyond 32");
    __shared__ double tl[2 * KM * QIW_WAVE];     // time of simplex coordinate j: tl[j * 64 + lane]
    __shared__ double sk[(NM - 3) * QIW_WAVE];   // saved product of level L >= 3: sk[(L - 3) * 64 + lane]
    __shared__ double so[SM * QIW_WAVE];         // output of initial sector s: so[s * 64 + lane]
    const int lane = threadIdx.x;
    const QiwTask task = a.tasks[blockIdx.x];
    const QiwGroup g = a.groups[task.group];
    const int ns = a.ns, n = g.n_nodes, k = g.order;
    const int ntab = a.dtab.n_tables;
    const double inv_h = 1.0 / a.h;

    double tau_i = 0.0, tau_w, tau_f;
    if (g.kind == QIW_G_BARE) {
        tau_w = 0.0;
        tau_f = a.h;
    } else if (g.kind == QIW_G_BOLD) {
        tau_w = a.step_w * a.h;
        tau_f = (a.step_w + 1) * a.h;
    } else if (g.kind == QIW_G_SIGMA) {  // tau_f_prev = grid[step_w]
        tau_w = a.step_w * a.h;
        tau_f = tau_w;
    } else {
        tau_w = task.tau_index * a.h;
        tau_f = a.beta;
    }
    const int64_t sample = (int64_t)task.first_sample + lane;
    const bool active = sample < g.n_samples;
    const double wt = g.weighted ? a.ut[g.ut_off + (int64_t)task.tau_index * a.ut_row_stride + (int64_t)(g.d_before + g.d_after) * g.n_samples +
                                        (active ? sample : 0)]
                                 : 1.0;

    // ---- contour times of the configuration positions (1-based) ----
    double t[NM + 1];
    const int pos_w = (g.kind == QIW_G_BARE) ? -1 : g.pos_w;
#pragma unroll
    for (int p = 0; p <= NM; ++p) t[p] = (p == 1) ? tau_i : (p == pos_w) ? tau_w : (p == n) ? tau_f : 0.0;
    if (g.sampled) {
        const int64_t N = g.n_samples;
        const double *u = a.ut + g.ut_off + (int64_t)task.tau_index * a.ut_row_stride + (active ? sample : 0);
        const double dfw = tau_f - tau_w, dwi = tau_w - tau_i;
#pragma unroll
        for (int j = 0; j < 2 * KM; ++j) {
            if (j < g.d_after) {
                const double v = u[j * N] * dfw + tau_w;
                tl[j * QIW_WAVE + lane] = v;
#pragma unroll
                for (int p = 1; p <= NM; ++p)
                    if (p == n - 1 - j) t[p] = v;
            } else if (j < g.d_after + g.d_before) {
                const double v = u[j * N] * dwi + tau_i;
                tl[j * QIW_WAVE + lane] = v;
                const int pos = g.pos_w - 1 - (j - g.d_after);
#pragma unroll
                for (int p = 1; p <= NM; ++p)
                    if (p == pos) t[p] = v;
            }
        }
    }

    // ---- propagator values i*P_s(t_L - t_{L-1}), L = 2..n (ordering clamp) ----
    typename RegVec<double, 16>::type P0 = 0.0, P1 = 0.0, P2 = 0.0, P3 = 0.0, P4 = 0.0, P5 = 0.0;
    typename RegVec<uint32_t, 32>::type PL = 0u, PH = 0u;  // SPLIT shapes
    double p_first[SM], p_last[SM];  // iP_s of segment 2 and of segment n
#pragma unroll
    for (int L = 2; L <= NM; ++L) {
        if (L <= n) {
            double dt = t[L] - t[L - 1];
            dt = dt > 0.0 ? dt : 0.0;
            const GridCell gc = grid_cell(t[L], t[L - 1], inv_h, a.n_tau);
            const double *r0 = a.ptab + (int64_t)gc.d * ns, *r1 = r0 + ns, *rm = gc.d > 0 ? r0 - ns : r0;
#pragma unroll
            for (int s = 0; s < SM; ++s) {
                double v = 0.0;
                if (s < ns)
                    v = (g.kind == QIW_G_BARE) ? exp(-dt * a.e_shift[s]) : cell_value(gc, rm[s], r0[s], r1[s]);
                if (L == 2) p_first[s] = v;
                if (L == n) p_last[s] = v;
                if (L >= 3 && L <= NM - 1) QIW_PIN_SET(NC, (NM - 1 - L) * SM + s, v);
            }
        }
    }

    typename RegVec<double, NW>::type W;
    typename RegVec<double, 16>::type Wh;  // slots 16..31 (unused unless W2)
#pragma unroll
    for (int q = 0; q < NW; ++q) W[q] = 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q) Wh[q] = 0.0;
#pragma unroll
    for (int s = 0; s < SM; ++s) so[s * QIW_WAVE + lane] = 0.0;
    double acc = 0.0, cur = 0.0, cur0 = 0.0;  // cur0: product of the root (level 2 restores it)

    // ---- walk the program: one word per iteration, until END ----
    const double *coefs = a.coefs + g.coef_off;
    const uint32_t *topo = a.topo + g.topo_off;
    ProgStream32 ps;
    ps.init(reinterpret_cast<const uint32_t *>(a.prog + g.prog_off + task.chunk), lane);
    uint32_t w = ps.next();
    for (;;) {
        // Runs of chain words (EDGE without POP, LEAF or PUSH) take the short inner loop; it
        // reads the arc values and the propagators but never writes them, so they stay in
        // place across its iterations.  The empty volatile asm statements keep the rare
        // cases as uniform branches: the compiler would otherwise evaluate them on every
        // word and select the result.
        while ((w & (0xF0000000u | QIW_FL_POP | QIW_FL_LEAF | QIW_FL_PUSH)) == 0) {
            const uint32_t wn = ps.next();  // the next word, read ahead (the block ends in NEXT)
            if (w & QIW_FL_TAIL) {
                asm volatile("");
                cur *= (W2 && (w & (16u << 8))) ? Wh[(int)((w >> 8) & 15)] : W[(int)((w >> 8) & (NW - 1))];
            }
            double pv;
            QIW_PIN_GET(NC, w, pv);
            cur *= pv;
            w = wn;
        }
        if (w < QIW_FL_OP) {  // EDGE with POP, LEAF or PUSH
            const int L = (int)((w >> 16) & 15);
            if (w & QIW_FL_POP) {  // level 2 restores the root's product, deeper levels a saved row
                asm volatile("");
                const double saved = sk[(L > 2 ? L - 3 : 0) * QIW_WAVE + lane];
                cur = (L == 2) ? cur0 : saved;
            }
            if (w & QIW_FL_TAIL) {
                asm volatile("");
                cur *= (W2 && (w & (16u << 8))) ? Wh[(int)((w >> 8) & 15)] : W[(int)((w >> 8) & (NW - 1))];
            }
            if (w & QIW_FL_LEAF) {
                // acc += c * product with c = +-1 (UNIT, a uniform select) or the group's
                // coefficient; one fused multiply-add either way (exact for c = +-1)
                double c = (w & QIW_FL_NEG) ? -1.0 : 1.0;
                if (!(w & QIW_FL_UNIT)) {
                    asm volatile("");
                    c = coefs[w & 0xFF];
                }
                acc = fma(c, cur, acc);
                w = ps.next();
                continue;
            }
            double pv;
            QIW_PIN_GET(NC, w, pv);
            cur *= pv;
            if (w & QIW_FL_PUSH) sk[(L - 2) * QIW_WAVE + lane] = cur;
            w = ps.next();
            continue;
        }
        const uint32_t op = w >> 28;
        if (op == QIW_FL_OP_NEXT) {
            ps.next_block();
        } else if (op == QIW_FL_OP_END) {
            break;
        } else if (op == QIW_FL_OP_ROOT) {  // position 1: cur = iP_2[s], or the leaf of n = 2
            if (w & QIW_FL_LEAF) {
                acc += (w & QIW_FL_UNIT) ? ((w & QIW_FL_NEG) ? -1.0 : 1.0) : coefs[w & 0xFF];
            } else {
                cur = pick(p_first, (int)(w & 0xFF));
                cur0 = cur;
            }
        } else if (op == QIW_FL_OP_TOPO) {
            const uint32_t *tw = topo + 2 * (w & 0xFFFFFF);
            const uint32_t lo = tw[0], hi = tw[1];
            // every slot is rewritten, so the old values die here: arcs beyond the order
            // read coordinate 0, tables beyond the model's count read its last table
#pragma unroll
            for (int ai = 0; ai < KM; ++ai) {
                const uint32_t b = ai >= k ? 0u : ai < 4 ? (lo >> (8 * ai)) : (hi >> (8 * (ai - 4)));
                const int head = (int)(b & 15), tail = (int)((b >> 4) & 15);
                const ArcPos ap = arc_pos(tl[tail * QIW_WAVE + lane], tl[head * QIW_WAVE + lane], inv_h, a.n_tau);
#pragma unroll
                for (int tb = 0; tb < TM; ++tb) {
                    const int tt = tb < ntab ? tb : (ntab > 0 ? ntab - 1 : 0);
                    const double v = delta_value<DLR>(a.dtab, tt, ap);
                    if (ai * TM + tb < 16)
                        W[(ai * TM + tb) & (NW - 1)] = v;
                    else
                        Wh[(ai * TM + tb) & 15] = v;
                }
            }
#pragma unroll
            for (int q = KM * TM; q < NW; ++q) W[q] = 0.0;
        } else {  // FLUSH: out[s0] += acc * iP_n[s0]
            const int s0 = (int)(w & 0xFF);
            so[s0 * QIW_WAVE + lane] += acc * pick(p_last, s0);
            acc = 0.0;
        }
        w = ps.next();
    }

    // ---- wavefront reduction of the per-s0 outputs ----
    double out[SM];
#pragma unroll
    for (int s = 0; s < SM; ++s) out[s] = so[s * QIW_WAVE + lane];
    if (a.per_sample && active) {
        double m = 0.0;
#pragma unroll
        for (int s = 0; s < SM; ++s) m += (s < ns) ? fabs(out[s]) : 0.0;
        a.per_sample[g.ps_off + sample] = m;
    }
#pragma unroll
    for (int s = 0; s < SM; ++s) {
        if (s < ns) {
            double v = active ? out[s] * wt : 0.0;
#pragma unroll
            for (int off = QIW_WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, QIW_WAVE);
            if (lane == 0) a.partial[(int64_t)task.out_index * ns + s] = v;
        }
    }
}

// ---------------------------------------------------------------------------------------
// The diagram evaluator for models with sector blocks larger than 1x1.
//
// One wavefront walks the trie for the 64 samples of its task one after the other; for
// each sample the lanes share the matrix work of every chain product: the node matrix of
// an edge is O_block * iP_{s_in}(t_L - t_{L-1}) (the reference's node_mats,
// src/topology_eval.jl:376-388) applied to the parent partial product, and lane e computes
// element e of the m_out x m_s0 result.  Per-sample state lives in LDS:
//   iP slots  [(L-2)*bt + BT offset]   interpolated propagator blocks of every segment
//   W slots   [arc*ntab + tab]         arc values i*Delta
//   stack     [L][mmax*mmax]           partial products (column-major, leading dim mmax)
//   X         [mmax*mmax]              i*P * parent, before the operator block
//   acc, out  accumulators of the current initial sector and of all sectors (BT layout)
// ---------------------------------------------------------------------------------------

__global__ __launch_bounds__(QIW_WAVE) void k_eval_block(BlockArgs b) {
    extern __shared__ double lds[];
    const EvalArgs &a = b.e;
    const int lane = threadIdx.x;
    const QiwTask task = a.tasks[blockIdx.x];
    const QiwGroup g = a.groups[task.group];
    const int n = g.n_nodes, k = g.order, bt = b.bt, mm = b.mmax * b.mmax, ld = b.mmax;
    const int ntab = a.dtab.n_tables;
    const double inv_h = 1.0 / a.h;
    double *ip = lds;
    double *wsl = ip + (n - 1) * bt;
    double *stk = wsl + k * ntab + 1;
    double *X = stk + (n + 1) * mm;
    double *acc = X + mm;
    double *out = acc + mm;
    for (int e = lane; e < bt; e += QIW_WAVE) out[e] = 0.0;

    double tau_i = 0.0, tau_w, tau_f;
    if (g.kind == QIW_G_BARE) {
        tau_w = 0.0;
        tau_f = a.h;
    } else if (g.kind == QIW_G_BOLD) {
        tau_w = a.step_w * a.h;
        tau_f = (a.step_w + 1) * a.h;
    } else if (g.kind == QIW_G_SIGMA) {  // tau_f_prev = grid[step_w]
        tau_w = a.step_w * a.h;
        tau_f = tau_w;
    } else {
        tau_w = task.tau_index * a.h;
        tau_f = a.beta;
    }
    const int pos_w = (g.kind == QIW_G_BARE) ? -1 : g.pos_w;
    const int n_here = g.sampled ? min(a.spt, g.n_samples - task.first_sample) : 1;
    for (int r = 0; r < n_here; ++r) {
        const int64_t sample = (int64_t)task.first_sample + r;
        // times of the configuration positions (identical in every lane)
        double t[QIW_NMAX + 1];
#pragma unroll
        for (int p = 0; p <= QIW_NMAX; ++p) t[p] = (p == 1) ? tau_i : (p == pos_w) ? tau_w : (p == n) ? tau_f : 0.0;
        if (g.sampled) {
            const int64_t N = g.n_samples;
            const double *u = a.ut + g.ut_off + (int64_t)task.tau_index * a.ut_row_stride + sample;
#pragma unroll
            for (int j = 0; j < 2 * QIW_KMAX; ++j) {
                if (j < g.d_after) {
                    const double v = u[j * N] * (tau_f - tau_w) + tau_w;
#pragma unroll
                    for (int p = 1; p <= QIW_NMAX; ++p)
                        if (p == n - 1 - j) t[p] = v;
                } else if (j < g.d_after + g.d_before) {
                    const double v = u[j * N] * (tau_w - tau_i) + tau_i;
                    const int pos = g.pos_w - 1 - (j - g.d_after);
#pragma unroll
                    for (int p = 1; p <= QIW_NMAX; ++p)
                        if (p == pos) t[p] = v;
                }
            }
        }
        // iP slots
#pragma unroll
        for (int L = 2; L <= QIW_NMAX; ++L) {
            if (L <= n) {
                double dt = t[L] - t[L - 1];
                dt = dt > 0.0 ? dt : 0.0;
                double *dst = ip + (L - 2) * bt;
                if (g.kind == QIW_G_BARE) {
                    // exact atomic propagator: diagonal exp(-dt (E + lambda))
                    for (int e = lane; e < bt; e += QIW_WAVE) dst[e] = 0.0;
                    for (int s = 0, o = 0; s < a.ns; ++s) {
                        const int m = b.dims[s];
                        for (int i = lane; i < m; i += QIW_WAVE) dst[b.bt_off[s] + i + i * m] = exp(-dt * a.e_shift[o + i]);
                        o += m;
                    }
                } else {
                    const GridCell gc = grid_cell(t[L], t[L - 1], inv_h, a.n_tau);
                    const double *r0 = a.ptab + (int64_t)gc.d * bt, *r1 = r0 + bt, *rm = gc.d > 0 ? r0 - bt : r0;
                    for (int e = lane; e < bt; e += QIW_WAVE) dst[e] = cell_value(gc, rm[e], r0[e], r1[e]);
                }
            }
        }
        __syncthreads();
        // walk
        int s0 = 0, m0 = 1;
        const uint64_t *__restrict__ prog = a.prog + g.prog_off;
        for (int pc = 0;; ++pc) {
            const uint64_t w = prog[pc];
            const int op = (int)((w >> 48) & 0xF);
            if (op == QIW_OP_EDGE) {
                const int L = (int)(w & 0xF);
                const int s_in = (int)((w >> 8) & 0xFF), s_out = (int)((w >> 40) & 0xFF);
                const int blk = (int)((w >> 24) & 0xFFFF);
                const int mi = b.dims[s_in], mo = b.dims[s_out];
                const double *par = stk + (L - 1) * mm;
                const double *P = ip + (L - 2) * bt + b.bt_off[s_in];
                const double wv = (w & QIW_F_TAIL) ? wsl[(int)((w >> 16) & 0xFF)] : 1.0;
                double *dst = stk + L * mm;
                if (blk == 0xFFFF) {  // identity / inch node: iP * parent
                    for (int e = lane; e < mi * m0; e += QIW_WAVE) {
                        const int i = e % mi, j = e / mi;
                        double v = 0.0;
                        for (int q = 0; q < mi; ++q) v += P[i + q * mi] * par[q + j * ld];
                        dst[i + j * ld] = v * wv;
                    }
                } else {
                    for (int e = lane; e < mi * m0; e += QIW_WAVE) {
                        const int i = e % mi, j = e / mi;
                        double v = 0.0;
                        for (int q = 0; q < mi; ++q) v += P[i + q * mi] * par[q + j * ld];
                        X[i + j * ld] = v;
                    }
                    __syncthreads();
                    const double *O = b.op_data + b.op_desc[3 * blk];
                    for (int e = lane; e < mo * m0; e += QIW_WAVE) {
                        const int i = e % mo, j = e / mo;
                        double v = 0.0;
                        for (int q = 0; q < mi; ++q) v += O[i + q * mo] * X[q + j * ld];
                        dst[i + j * ld] = v * wv;
                    }
                }
                __syncthreads();
                if (w & QIW_F_LEAF) {
                    const double c = __longlong_as_double((long long)prog[++pc]);
                    for (int e = lane; e < m0 * m0; e += QIW_WAVE) {
                        const int i = e % m0, j = e / m0;
                        acc[i + j * m0] += c * dst[i + j * ld];
                    }
                }
            } else if (op == QIW_OP_ROOT) {
                s0 = (int)((w >> 8) & 0xFF);
                const int s1 = (int)((w >> 40) & 0xFF);
                const int blk = (int)((w >> 24) & 0xFFFF);
                m0 = b.dims[s0];
                const int m1 = b.dims[s1];
                for (int e = lane; e < m1 * m0; e += QIW_WAVE) {
                    const int i = e % m1, j = e / m1;
                    stk[1 * mm + i + j * ld] =
                        (blk == 0xFFFF) ? (i == j ? 1.0 : 0.0) : b.op_data[b.op_desc[3 * blk] + i + j * m1];
                }
                for (int e = lane; e < m0 * m0; e += QIW_WAVE) acc[e] = 0.0;
                __syncthreads();
            } else if (op == QIW_OP_TOPO) {
                if (lane < k * ntab) {
                    const int ai = lane / ntab, tb = lane % ntab;
                    const int head = (int)((w >> (8 * ai)) & 0xF), tail = (int)((w >> (8 * ai + 4)) & 0xF);
                    double th = 0.0, tt = 0.0;
#pragma unroll
                    for (int p = 1; p <= QIW_NMAX; ++p) {
                        th = (p == head) ? t[p] : th;
                        tt = (p == tail) ? t[p] : tt;
                    }
                    wsl[lane] = delta_value(a.dtab, tb, arc_pos(tt, th, inv_h, a.n_tau));
                }
                __syncthreads();
            } else if (op == QIW_OP_FLUSH) {
                for (int e = lane; e < m0 * m0; e += QIW_WAVE) out[b.bt_off[s0] + e] += acc[e];
                __syncthreads();
            } else if (op == QIW_OP_END) {
                break;
            }
        }
    }
    for (int e = lane; e < bt; e += QIW_WAVE) a.partial[(int64_t)task.out_index * bt + e] = out[e];
}

// ---------------------------------------------------------------------------------------
// The diagram evaluator for models with small sector blocks (all m_s <= MB, MB <= 8).
//
// One lane = one qMC sample, as in the scalar walk.  A matrix chain X = O_n iP_n ... O_2
// iP_2 B is evaluated one column at a time: the program of an initial sector s0 is a
// segment that the wavefront replays once for every column j < m_s0, starting from the
// unit vector e_j (or column j of the correlator operator B).  Along the trie every lane
// keeps, per level, the vector y = iP_L[s_par] * x_par of its own sample in registers;
// a child edge applies its operator block O (uniform: read through the scalar cache) and
// passes x = O y one level down.  Products of the arc values i*Delta travel as a scalar
// weight beside the vector and meet the leaf coefficient at the leaf.  The last propagator
// iP_n[s0] is the same for every leaf of a column, so it is applied once, to the column
// sum, when the column ends.
//
// Propagator lookups read the grid table as cells: row d of a segment's grid cell with its
// forward and backward differences (see grid_cell), so one entry is one load (two in
// float64) and two FMAs at the lane's own cell.
// Vector entries beyond the current block size hold finite values that are never read.
//
// Value type V: float64, or float32 for the chain products (QIW_PREC_FP32: propagator
// cells, operator blocks, arc values and the per-level vectors in float32; the column
// sums, the last propagator and everything after them in float64).
// ---------------------------------------------------------------------------------------

// For every index j < n_tau - 1 and BT element e: T[j][e], the forward difference
// T[j+1][e] - T[j][e] and the backward difference T[j][e] - T[j-1][e] (for j = 0 the
// forward one, the cell value's convention g(-1) = 2 g(0) - g(1)); float64 as (value,
// forward) pairs plus the backward differences, float32 as packed 12-byte (value, forward,
// backward) cells.
__global__ void k_ptab_cells(const double *__restrict__ T, int n_tau, int bt, double2 *__restrict__ out,
                             double *__restrict__ outm, QiwCell32 *__restrict__ out32) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= (int64_t)(n_tau - 1) * bt) return;
    const double v0 = T[r], v1 = T[r + bt];
    const double dm = r >= bt ? v0 - T[r - bt] : v1 - v0;
    out[r] = make_double2(v0, v1 - v0);
    outm[r] = dm;
    out32[r] = QiwCell32{(float)v0, (float)(v1 - v0), (float)dm};
}

// Cell tables of the two value types: entry i at weights (wp, wm) = row + wp forward -
// wm backward.
template <typename V>
struct CellTab;
template <>
struct CellTab<double> {
    const double2 *__restrict__ p;
    const double *__restrict__ m;
    __device__ __forceinline__ double at(int64_t i, double wp, double wm) const {
        const double2 e = p[i];
        return fma(wp, e.y, fma(-wm, m[i], e.x));
    }
};
template <>
struct CellTab<float> {
    const QiwCell32 *__restrict__ p;  // one 12-byte load per entry
    __device__ __forceinline__ float at(int64_t i, float wp, float wm) const {
        const QiwCell32 e = p[i];
        return fma(wp, e.f, fma(-wm, e.b, e.v));
    }
};

// Read-only tables at wavefront-uniform addresses (operator blocks, atomic energies) are
// read through the constant address space, so the loads are scalar (one SMEM load for the
// wavefront, through the scalar cache) instead of vector loads that all 64 lanes repeat.
template <typename T>
using ConstPtr = const T __attribute__((address_space(4))) *;
template <typename T>
__device__ __forceinline__ ConstPtr<T> as_const(const T *p) {
    return (ConstPtr<T>)p;
}

// State of one lane (one qMC sample) of the vector walk, which carries NC columns of the
// initial sector's block at a time: every propagator entry it loads serves NC products.
template <typename V, int MB, int KM, int TM, int NC>
struct VecState {
    static constexpr int NM = 2 * KM + 3;
    int row[NM - 1];      // BT offset of the grid index d of segment L's cell (bold)
    V wp[NM - 1];         // cell weights of segment L (bold)
    V wm[NM - 1];
    double dtb[NM - 1];   // time difference (bare)
    V W[KM * TM];         // i*Delta_tab of the arcs, per arc and table
    double acc[NC][MB];   // columns col .. col + NC - 1 of the current initial sector
    // LDS of the wavefront: the configuration times t[p] at ts[p * 64 + lane], and the
    // propagated vectors of levels 2 .. 1 + YL at ys[((L - 2) * NC + c) * MB + i][lane],
    // which are reused by few children, so they need no registers across the recursion
    double *ts;
    V (*ys)[QIW_WAVE];
    int lane;
    ProgStream ps;
    CellTab<V> pt;
    CellTab<double> pt64;
    ConstPtr<double> ediag;
    ConstPtr<V> ops;
    int bt;
};

// y = iP_L[s] x for the m x m block of sector s at BT offset off, NC columns.
template <int L, typename V, int MB, int KM, int TM, int NC, bool BARE>
__device__ __forceinline__ void prop_apply(const VecState<V, MB, KM, TM, NC> &st, const V (&x)[NC][MB], int m,
                                           int off, V (&y)[NC][MB]) {
    if constexpr (BARE) {  // exact atomic propagator: diagonal exp(-dt (E + lambda))
        const double dt = st.dtb[L - 2];
#pragma unroll
        for (int i = 0; i < MB; ++i) {
            const V e = (i < m) ? (V)exp(-dt * st.ediag[off + i * (m + 1)]) : (V)0;
#pragma unroll
            for (int c = 0; c < NC; ++c) y[c][i] = e * x[c][i];
        }
    } else {
        const int64_t base = st.row[L - 2] + off;
        const V wp = st.wp[L - 2], wm = st.wm[L - 2];
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int i = 0; i < MB; ++i) y[c][i] = (V)0;
#pragma unroll
        for (int q = 0; q < MB; ++q) {
            if (q < m) {
#pragma unroll
                for (int i = 0; i < MB; ++i) {
                    const V pv = st.pt.at(base + q * m + i, wp, wm);
#pragma unroll
                    for (int c = 0; c < NC; ++c) y[c][i] = fma(pv, x[c][q], y[c][i]);
                }
            }
        }
    }
}

// y = iP_n[s0] x for the last segment (runtime n), at the end of a column (float64)
template <typename V, int MB, int KM, int TM, int NC, bool BARE>
__device__ __forceinline__ void last_prop(const VecState<V, MB, KM, TM, NC> &st, int n, const double (&x)[MB], int m,
                                          int off, double inv_h, int n_tau, double (&y)[MB]) {
    const double tn = st.ts[n * QIW_WAVE + st.lane], tp = st.ts[(n - 1) * QIW_WAVE + st.lane];
    if constexpr (BARE) {
        double dt = tn - tp;
        dt = dt > 0.0 ? dt : 0.0;
#pragma unroll
        for (int i = 0; i < MB; ++i) y[i] = (i < m) ? exp(-dt * st.ediag[off + i * (m + 1)]) * x[i] : 0.0;
    } else {
        const GridCell gc = grid_cell(tn, tp, inv_h, n_tau);
        const int64_t base = (int64_t)gc.d * st.bt + off;
#pragma unroll
        for (int i = 0; i < MB; ++i) y[i] = 0.0;
#pragma unroll
        for (int q = 0; q < MB; ++q) {
            if (q < m) {
                const double xq = x[q];
#pragma unroll
                for (int i = 0; i < MB; ++i) y[i] = fma(st.pt64.at(base + q * m + i, gc.wp, gc.wm), xq, y[i]);
            }
        }
    }
}

#ifndef QIW_VEC_YL
#define QIW_VEC_YL 4
#endif
template <int L, typename V, int MB, int KM, int TM, int NC, bool BARE>
__device__ __forceinline__ void vwalk(VecState<V, MB, KM, TM, NC> &st, int n, const V (&xp)[NC][MB], V wp, int mp,
                                      int offp, int nchild) {
    if constexpr (L <= 2 * KM + 3) {
        constexpr bool in_lds = L - 2 < QIW_VEC_YL;
        V y[NC][MB];
        if (L == n) {  // leaves: iP_n[s0] is applied to the column sums
#pragma unroll
            for (int c = 0; c < NC; ++c)
#pragma unroll
                for (int i = 0; i < MB; ++i) y[c][i] = xp[c][i];
        } else {
            prop_apply<L, V, MB, KM, TM, NC, BARE>(st, xp, mp, offp, y);
        }
        if constexpr (in_lds) {
#pragma unroll
            for (int c = 0; c < NC; ++c)
#pragma unroll
                for (int i = 0; i < MB; ++i) st.ys[((L - 2) * NC + c) * MB + i][st.lane] = y[c][i];
        }
        for (int ch = 0; ch < nchild; ++ch) {
            if constexpr (in_lds) {
#pragma unroll
                for (int c = 0; c < NC; ++c)
#pragma unroll
                    for (int i = 0; i < MB; ++i) y[c][i] = st.ys[((L - 2) * NC + c) * MB + i][st.lane];
            }
            const uint64_t w = st.ps.next();
            const V wv = (w & QIW_F_TAIL) ? wp * pickv(st.W, (int)((w >> 16) & 0xFF)) : wp;
            const int mo = (int)(w & 0x3F);
            const int blk = (int)((w >> 32) & 0xFFFF);
            V x[NC][MB];
            if (blk == 0xFFFF) {
#pragma unroll
                for (int c = 0; c < NC; ++c)
#pragma unroll
                    for (int i = 0; i < MB; ++i) x[c][i] = y[c][i];
            } else {
                const ConstPtr<V> O = st.ops + blk;  // m_o x m_p, column-major
#pragma unroll
                for (int c = 0; c < NC; ++c)
#pragma unroll
                    for (int i = 0; i < MB; ++i) x[c][i] = (V)0;
#pragma unroll
                for (int q = 0; q < MB; ++q) {
                    if (q < mp) {
#pragma unroll
                        for (int i = 0; i < MB; ++i) {
                            const V o = O[q * mo + i];
#pragma unroll
                            for (int c = 0; c < NC; ++c) x[c][i] = fma(o, y[c][q], x[c][i]);
                        }
                    }
                }
            }
            if (w & QIW_F_LEAF) {
                const double cw = (w & QIW_FV_UNIT) ? ((w & QIW_FV_NEG) ? -(double)wv : (double)wv)
                                                    : __longlong_as_double((long long)st.ps.next()) * (double)wv;
#pragma unroll
                for (int c = 0; c < NC; ++c)
#pragma unroll
                    for (int i = 0; i < MB; ++i) st.acc[c][i] = fma(cw, (double)x[c][i], st.acc[c][i]);
            } else if (L < n) {
                vwalk<L + 1, V, MB, KM, TM, NC, BARE>(st, n, x, wv, mo, (int)(w >> 52), (int)((w >> 24) & 0xFF));
            }
        }
    }
}

template <typename V, int MB, int KM, int TM, int NC, bool BARE>
__device__ __forceinline__ void vec_program(VecState<V, MB, KM, TM, NC> &st, const EvalArgs &a, const QiwGroup &g,
                                            double inv_h, bool active, double wt, double &absum,
                                            double *__restrict__ prow) {
    const int n = g.n_nodes, k = g.order;
    const int ntab = a.dtab.n_tables;
    const uint64_t *seg = nullptr;
    int s_m0 = 1, s_off0 = 0, col = 0;
    for (;;) {
        const uint64_t w = st.ps.next();
        const int op = (int)((w >> 48) & 0xF);
        if (op == QIW_OP_ROOT) {
            // position 1: e_col, or column col of the operator B (correlators), for the
            // columns col .. col + NC - 1 that exist (zero beyond)
            const int m1 = (int)(w & 0x3F);
            const int blk = (int)((w >> 32) & 0xFFFF);
            V x[NC][MB];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const bool live = col + c < s_m0;
                if (blk == 0xFFFF) {
#pragma unroll
                    for (int i = 0; i < MB; ++i) x[c][i] = (live && i == col + c) ? (V)1 : (V)0;
                } else {
#pragma unroll
                    for (int i = 0; i < MB; ++i) x[c][i] = live ? st.ops[blk + (col + c) * m1 + i] : (V)0;
                }
            }
            vwalk<2, V, MB, KM, TM, NC, BARE>(st, n, x, (V)1, m1, (int)(w >> 52), (int)((w >> 24) & 0xFF));
        } else if (op == QIW_OP_TOPO) {
#pragma unroll
            for (int ai = 0; ai < KM; ++ai) {
                const int head = (int)((w >> (8 * ai)) & 0xF), tail = (int)((w >> (8 * ai + 4)) & 0xF);
                if (ai < k) {
                    const double th = st.ts[head * QIW_WAVE + st.lane], tt = st.ts[tail * QIW_WAVE + st.lane];
                    const ArcPos ap = arc_pos(tt, th, inv_h, a.n_tau);
#pragma unroll
                    for (int tb = 0; tb < TM; ++tb)
                        if (tb < ntab) st.W[ai * TM + tb] = (V)delta_value(a.dtab, tb, ap);
                }
            }
        } else if (op == QIW_OP_SEG) {  // initial sector s0: m0 columns, BT offset
            s_m0 = (int)(w & 0x3F);
            s_off0 = (int)(w >> 52);
            col = 0;
            seg = st.ps.base + st.ps.idx;
#pragma unroll
            for (int c = 0; c < NC; ++c)
#pragma unroll
                for (int i = 0; i < MB; ++i) st.acc[c][i] = 0.0;
        } else if (op == QIW_OP_FLUSH) {  // columns col .. col + NC - 1 of the s0 block are complete
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                if (col + c < s_m0) {
                    double y[MB];
                    last_prop<V, MB, KM, TM, NC, BARE>(st, n, st.acc[c], s_m0, s_off0, inv_h, a.n_tau, y);
#pragma unroll
                    for (int i = 0; i < MB; ++i) {
                        if (i < s_m0) {
                            absum += fabs(y[i]);
                            double v = active ? y[i] * wt : 0.0;
#pragma unroll
                            for (int o = QIW_WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, QIW_WAVE);
                            if (threadIdx.x == 0) prow[s_off0 + (col + c) * s_m0 + i] = v;
                        }
                    }
                }
#pragma unroll
                for (int i = 0; i < MB; ++i) st.acc[c][i] = 0.0;
            }
            col += NC;
            if (col < s_m0) st.ps.init(seg, threadIdx.x);
        } else if (op == QIW_OP_END) {
            break;
        }
    }
}

#ifndef QIW_VEC_MINW
#define QIW_VEC_MINW 2
#endif
template <typename V, int MB, int KM, int TM, int NC>
__global__ __launch_bounds__(QIW_WAVE, QIW_VEC_MINW) void k_eval_vec_t(VecArgs va) {
    constexpr int NM = 2 * KM + 3;
    const EvalArgs &a = va.e;
    const int lane = threadIdx.x;
    const QiwTask task = a.tasks[blockIdx.x];
    const QiwGroup g = a.groups[task.group];
    const int n = g.n_nodes;
    const double inv_h = 1.0 / a.h;

    double tau_i = 0.0, tau_w, tau_f;
    if (g.kind == QIW_G_BARE) {
        tau_w = 0.0;
        tau_f = a.h;
    } else if (g.kind == QIW_G_BOLD) {
        tau_w = a.step_w * a.h;
        tau_f = (a.step_w + 1) * a.h;
    } else if (g.kind == QIW_G_SIGMA) {  // tau_f_prev = grid[step_w]
        tau_w = a.step_w * a.h;
        tau_f = tau_w;
    } else {
        tau_w = task.tau_index * a.h;
        tau_f = a.beta;
    }
    const int64_t sample = (int64_t)task.first_sample + lane;
    const bool active = sample < g.n_samples;

    double t[NM + 1];
    const int pos_w = (g.kind == QIW_G_BARE) ? -1 : g.pos_w;
#pragma unroll
    for (int p = 0; p <= NM; ++p) t[p] = (p == 1) ? tau_i : (p == pos_w) ? tau_w : (p == n) ? tau_f : 0.0;
    if (g.sampled) {
        const int64_t N = g.n_samples;
        const double *u = a.ut + g.ut_off + (int64_t)task.tau_index * a.ut_row_stride + (active ? sample : 0);
        const double dfw = tau_f - tau_w, dwi = tau_w - tau_i;
#pragma unroll
        for (int j = 0; j < 2 * KM; ++j) {
            if (j < g.d_after) {
                const double v = u[j * N] * dfw + tau_w;
#pragma unroll
                for (int p = 1; p <= NM; ++p)
                    if (p == n - 1 - j) t[p] = v;
            } else if (j < g.d_after + g.d_before) {
                const double v = u[j * N] * dwi + tau_i;
                const int pos = g.pos_w - 1 - (j - g.d_after);
#pragma unroll
                for (int p = 1; p <= NM; ++p)
                    if (p == pos) t[p] = v;
            }
        }
    }

    constexpr int YL = QIW_VEC_YL < NM - 1 ? QIW_VEC_YL : NM - 1;
    __shared__ double s_t[(NM + 1) * QIW_WAVE];
    __shared__ V s_y[(YL > 0 ? YL : 1) * NC * MB][QIW_WAVE];
#pragma unroll
    for (int p = 0; p <= NM; ++p) s_t[p * QIW_WAVE + lane] = t[p];
    VecState<V, MB, KM, TM, NC> st;
    st.ts = s_t;
    st.ys = s_y;
    st.lane = lane;
    if constexpr (sizeof(V) == 8) {
        st.pt.p = va.pt;
        st.pt.m = va.ptm;
        st.ops = as_const((const V *)va.ops);
    } else {
        st.pt.p = va.pt32;
        st.ops = as_const((const V *)va.ops32);
    }
    st.pt64.p = va.pt;
    st.pt64.m = va.ptm;
    st.ediag = as_const(va.ediag);
    st.bt = va.bt;
#pragma unroll
    for (int L = 2; L <= NM; ++L) {
        double dt = t[L] - t[L - 1];
        dt = dt > 0.0 ? dt : 0.0;
        st.dtb[L - 2] = dt;
        if (g.kind == QIW_G_BARE) {
            st.wp[L - 2] = (V)0;
            st.wm[L - 2] = (V)0;
            st.row[L - 2] = 0;
        } else {
            const GridCell gc = grid_cell(t[L], t[L - 1], inv_h, a.n_tau);
            st.wp[L - 2] = (V)gc.wp;
            st.wm[L - 2] = (V)gc.wm;
            st.row[L - 2] = gc.d * va.bt;
        }
    }
#pragma unroll
    for (int q = 0; q < KM * TM; ++q) st.W[q] = (V)0;
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int i = 0; i < MB; ++i) st.acc[c][i] = 0.0;

    const double wt = g.weighted ? a.ut[g.ut_off + (int64_t)task.tau_index * a.ut_row_stride +
                                        (int64_t)(g.d_before + g.d_after) * g.n_samples + (active ? sample : 0)]
                                 : 1.0;
    double absum = 0.0;
    double *prow = a.partial + (int64_t)task.out_index * va.bt;
    st.ps.init(a.prog + g.prog_off + task.chunk, lane);
    if (g.kind == QIW_G_BARE)
        vec_program<V, MB, KM, TM, NC, true>(st, a, g, inv_h, active, wt, absum, prow);
    else
        vec_program<V, MB, KM, TM, NC, false>(st, a, g, inv_h, active, wt, absum, prow);
    if (a.per_sample && active) a.per_sample[g.ps_off + sample] = absum;
}

// ---------------------------------------------------------------------------------------
// The diagram evaluator for wide sector blocks (up to 32) on the matrix cores.
//
// One wavefront evaluates one sample at a time (all lanes share the sample's times, so
// propagators and arc values are uniform).  Partial products X (m_L x m_s0, 16 columns of
// s0 per pass) stay in registers as 16x16 tiles of v_mfma_f32_16x16x4f32 accumulators: lane
// l = 16 q + c holds rows 16 ti + 4 q + r (r = 0..3) of column c.  The instruction takes
// row l / 16 of its 4-row B slice from lane l, so element s of every lane is, as it stands,
// the B slice of rows {16 tk + 4 q + s}; the A operand of that slice is A(16 ti + c,
// 16 tk + 4 q + s) in lane l, and the 16x16 result lands in the same layout.  A product
// C = A X therefore feeds the next product without moving X, and each lane reads its A
// values of a tile pair as four consecutive floats of a row-major A (one 16-byte load): the
// operator blocks (global memory) and the i*P blocks of the sample (LDS) are stored
// row-major with rows padded to a multiple of 4 floats.  Lanes past the block's rows or
// columns read four zeros.  The walk follows the vector programs (one segment per initial
// sector); products of arc values travel as a float64 weight, leaves accumulate into
// float64 accumulators, and the last propagator iP_n[s0] is applied once per segment.
// float32 products (the QIW_PREC_FP32 mode); sums over samples in float64.
// ---------------------------------------------------------------------------------------
typedef float mf4 __attribute__((ext_vector_type(4)));

template <int TR>
struct MTile {
    mf4 t[TR];  // row tiles of one 16-column pass
};

template <int KM, int TM, int TR>
struct MState {
    typename RegVec<double, regvec_size(KM * TM)>::type W;  // arc values of the topology
    double acc[TR][4];              // leaf sums, in the tile layout
    ProgStream ps;
    const float *__restrict__ pl;   // i*P blocks of the sample in LDS: [(L - 2) * btp + LDS offset]
    const float *__restrict__ ops;  // operator blocks, row-major, rows padded to 4 floats
    float *__restrict__ ystk;       // the walk's stack of saved levels (LDS)
    int btp;                        // floats per segment of i*P blocks in LDS
    int pl_zero, ops_zero;          // positions of four 0.0 in pl and in ops
    int dummy;                      // 64 floats of LDS that take the writes of absent rows
};

// C = A X for A (m_a x K, row-major with rows of round4(K) floats, at base + off) and X
// (K x 16 columns), both in the tile layout, for blocks up to 32.  One 16-byte load per
// tile pair, all issued before the first MFMA; a row tile of X takes all four of its slices
// (slices past K meet zeros in A).
__device__ __forceinline__ MTile<2> mmul(const float *__restrict__ base, int off, int m_a, int K, const MTile<2> &X,
                                         int zpos) {
    const int l = threadIdx.x & (QIW_WAVE - 1), q = l >> 4, c = l & 15;
    const int ldr = (K + 3) & ~3;
    const bool two_m = m_a > 16, two_k = K > 16;
    const mf4 z = mf4{0.f, 0.f, 0.f, 0.f};
    auto ld = [&](int tk, int ti) {
        const int row = 16 * ti + c, kk = 16 * tk + 4 * q;
        const int idx = (kk < K && row < m_a) ? off + row * ldr + kk : zpos;
        return *(const mf4 *)(base + (unsigned)idx);
    };
    const mf4 a00 = ld(0, 0);
    const mf4 a01 = two_m ? ld(0, 1) : z;
    const mf4 a10 = two_k ? ld(1, 0) : z;
    const mf4 a11 = (two_k && two_m) ? ld(1, 1) : z;
    MTile<2> C;
    C.t[0] = z;
    C.t[1] = z;
    if (!two_m) {
#pragma unroll
        for (int s = 0; s < 4; ++s) C.t[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a00[s], X.t[0][s], C.t[0], 0, 0, 0);
        if (two_k) {
#pragma unroll
            for (int s = 0; s < 4; ++s) C.t[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a10[s], X.t[1][s], C.t[0], 0, 0, 0);
        }
    } else {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            C.t[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a00[s], X.t[0][s], C.t[0], 0, 0, 0);
            C.t[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a01[s], X.t[0][s], C.t[1], 0, 0, 0);
        }
        if (two_k) {
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                C.t[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a10[s], X.t[1][s], C.t[0], 0, 0, 0);
                C.t[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a11[s], X.t[1][s], C.t[1], 0, 0, 0);
            }
        }
    }
    return C;
}

// One root of the trie (a ROOT word with X0 = O_1 or the identity on s0), walked
// depth-first without recursion.  Y (= iP_L X of the parent, shared by a node's children)
// of the current level stays in registers.  When the walk descends from a level that has
// children left, it pushes a frame on the LDS stack: a 16-byte header (children left,
// rows, level, the previous frame, the weight of the arc values above) and Y (row-major,
// mp rows of 16 columns); a level whose last child is being walked pushes nothing.  Back
// from a subtree, the walk pops the top frame, header and Y in one round trip, and resumes
// that level.  The loop carries only the current level in registers.
template <int KM, int TM>
__device__ __forceinline__ void mwalk(MState<KM, TM, 2> &st, int n, const MTile<2> &X0, int m1, int off1, int nch1) {
    const int l = threadIdx.x & (QIW_WAVE - 1), q = l >> 4, c = l & 15;
    int L = 2, mp = m1, rem = nch1, top = -1, next = 0;
    double wp = 1.0;
    MTile<2> Y = (n == 2) ? X0 : mmul(st.pl, off1, m1, m1, X0, st.pl_zero);
    for (;;) {
        if (rem == 0) {  // back to the level of the top frame
            if (top < 0) return;
            const int4 h = *(const int4 *)(st.ystk + top);
            const float *fr = st.ystk + top + 4;
            const int rows = (next - top - 4) >> 4;  // rows of the top frame (it ends at next)
            float y[2][4];
#pragma unroll
            for (int ti = 0; ti < 2; ++ti)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * ti + 4 * q + r;
                    y[ti][r] = fr[(row < rows ? row : 0) * 16 + c];
                }
            next = top;
            rem = h.x & 0xFF;
            mp = (h.x >> 8) & 0xFF;
            L = h.x >> 16;
            top = h.y;
            wp = __hiloint2double(h.w, h.z);
#pragma unroll
            for (int ti = 0; ti < 2; ++ti)
#pragma unroll
                for (int r = 0; r < 4; ++r) Y.t[ti][r] = (16 * ti + 4 * q + r < mp) ? y[ti][r] : 0.f;
            continue;
        }
        --rem;
        const uint64_t w = st.ps.next();
        const int mo = (int)(w & 0x3F);
        const int blk = (int)((w >> 32) & 0xFFFF);
        const MTile<2> X = (blk == 0xFFFF) ? Y : mmul(st.ops, blk, mo, mp, Y, st.ops_zero);
        const double wv = (w & QIW_F_TAIL) ? wp * st.W[(int)((w >> 16) & 0xFF)] : wp;
        if (w & QIW_F_LEAF) {
            const double cw = (w & QIW_FV_UNIT) ? ((w & QIW_FV_NEG) ? -wv : wv)
                                                : __longlong_as_double((long long)st.ps.next()) * wv;
#pragma unroll
            for (int r = 0; r < 4; ++r) st.acc[0][r] = fma(cw, (double)X.t[0][r], st.acc[0][r]);
            if (mo > 16) {
#pragma unroll
                for (int r = 0; r < 4; ++r) st.acc[1][r] = fma(cw, (double)X.t[1][r], st.acc[1][r]);
            }
        } else if (L < n) {  // an inner node: its children follow
            if (rem > 0) {
                float *fr = st.ystk + next;
                if (l == 0)
                    *(int4 *)fr = make_int4(rem | (mp << 8) | (L << 16), top, __double2loint(wp), __double2hiint(wp));
#pragma unroll
                for (int ti = 0; ti < 2; ++ti)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = 16 * ti + 4 * q + r;
                        st.ystk[row < mp ? next + 4 + row * 16 + c : st.dummy + l] = Y.t[ti][r];
                    }
                top = next;
                next += 4 + mp * 16;
            }
            ++L;
            wp = wv;
            mp = mo;
            rem = (int)((w >> 24) & 0xFF);
            Y = (L == n) ? X : mmul(st.pl, (L - 2) * st.btp + (int)(w >> 52), mo, mo, X, st.pl_zero);
        }
    }
}

// A workgroup of W wavefronts: tasks W b .. W b + W - 1, one per wavefront, are topology
// chunks of the same sample block (a task with chunk -1 only helps stage); they share the
// samples' i*P blocks, staged by all W wavefronts, and each walks its chunk.
template <int KM, int TM, int TR, int W>
__global__ __launch_bounds__(QIW_WAVE * W, 2) void k_eval_mfma_t(MfmaArgs ma) {
    constexpr int NM = 2 * KM + 3;
    static_assert(TR == 2, "tiles for blocks up to 32");
    // [(NM - 1) * btp] i*P blocks, four 0.0, then per wavefront 64 floats for the writes of
    // absent rows and the stack of its walk (wave_lds floats each)
    extern __shared__ __attribute__((aligned(16))) float mlds[];
    __shared__ double s_t[16];      // the times of the configuration positions 0..NM
    const EvalArgs &a = ma.e;
    const int wv = threadIdx.x / QIW_WAVE, lane = threadIdx.x & (QIW_WAVE - 1), q = lane >> 4, cl = lane & 15;
    const QiwTask task = a.tasks[blockIdx.x * W + wv];
    const bool walks = task.chunk >= 0;
    const QiwGroup g = a.groups[task.group];
    const int n = g.n_nodes, k = g.order, bt = ma.bt, btp = ma.btp;
    const int ntab = a.dtab.n_tables;
    const double inv_h = 1.0 / a.h;
    float *pl = mlds;
    MState<KM, TM, TR> st;
    st.pl = pl;
    st.ops = ma.ops;
    st.btp = btp;
    st.pl_zero = ((NM - 1) * btp + 3) & ~3;
    st.ops_zero = ma.ops_zero;
    st.ystk = pl + st.pl_zero + 4 + wv * ma.wave_lds + 64;  // past the 64 floats for absent rows
    st.dummy = -64;                                         // those 64 floats, relative to ystk
    if (threadIdx.x < 4) pl[st.pl_zero + threadIdx.x] = 0.f;
#pragma unroll
    for (int e = 0; e < regvec_size(KM * TM); ++e) st.W[e] = 0.0;

    double tau_i = 0.0, tau_w, tau_f;
    if (g.kind == QIW_G_BARE) {
        tau_w = 0.0;
        tau_f = a.h;
    } else if (g.kind == QIW_G_BOLD) {
        tau_w = a.step_w * a.h;
        tau_f = (a.step_w + 1) * a.h;
    } else if (g.kind == QIW_G_SIGMA) {  // tau_f_prev = grid[step_w]
        tau_w = a.step_w * a.h;
        tau_f = tau_w;
    } else {
        tau_w = task.tau_index * a.h;
        tau_f = a.beta;
    }
    double *prow = a.partial + (int64_t)(walks ? task.out_index : 0) * bt;
    if (walks)
        for (int e = lane; e < bt; e += QIW_WAVE) prow[e] = 0.0;
    __threadfence_block();
    __syncthreads();
    const int pos_w = (g.kind == QIW_G_BARE) ? -1 : g.pos_w;
    const int n_here = g.sampled ? min(ma.e.spt, g.n_samples - task.first_sample) : 1;
    for (int sr = 0; sr < n_here; ++sr) {
        const int64_t sample = (int64_t)task.first_sample + sr;
        // ---- times of the configuration positions (the same in every lane) ----
        double t[NM + 1];
#pragma unroll
        for (int p = 0; p <= NM; ++p) t[p] = (p == 1) ? tau_i : (p == pos_w) ? tau_w : (p == n) ? tau_f : 0.0;
        if (g.sampled) {
            const int64_t N = g.n_samples;
            const double *u = a.ut + g.ut_off + (int64_t)task.tau_index * a.ut_row_stride + sample;
#pragma unroll
            for (int j = 0; j < 2 * KM; ++j) {
                if (j < g.d_after) {
                    const double v = u[j * N] * (tau_f - tau_w) + tau_w;
#pragma unroll
                    for (int p = 1; p <= NM; ++p)
                        if (p == n - 1 - j) t[p] = v;
                } else if (j < g.d_after + g.d_before) {
                    const double v = u[j * N] * (tau_w - tau_i) + tau_i;
                    const int pos = g.pos_w - 1 - (j - g.d_after);
#pragma unroll
                    for (int p = 1; p <= NM; ++p)
                        if (p == pos) t[p] = v;
                }
            }
        }
        const double wt = g.weighted ? a.ut[g.ut_off + (int64_t)task.tau_index * a.ut_row_stride + (int64_t)(g.d_before + g.d_after) * g.n_samples + sample] : 1.0;
#pragma unroll
        for (int p = 0; p <= NM; ++p) s_t[p] = t[p];
        // ---- i*P blocks of every segment into LDS, row-major with padded rows (lanes over
        // the LDS positions; padding positions get 0.0) ----
        for (int L = 2; L <= n; ++L) {
            double dt = pick(t, L) - pick(t, L - 1);
            dt = dt > 0.0 ? dt : 0.0;
            float *dst = pl + (L - 2) * btp;
            if (g.kind == QIW_G_BARE) {
                for (int p = threadIdx.x; p < btp; p += QIW_WAVE * W) {
                    const int e = ma.pinv[p];
                    dst[p] = (e >= 0 && ma.emask[e]) ? (float)exp(-dt * ma.ediag[e]) : 0.f;
                }
            } else {
                const GridCell gc = grid_cell(pick(t, L), pick(t, L - 1), inv_h, a.n_tau);
                const double *r0 = a.ptab + (int64_t)gc.d * bt, *r1 = r0 + bt, *rm = gc.d > 0 ? r0 - bt : r0;
                for (int p = threadIdx.x; p < btp; p += QIW_WAVE * W) {
                    const int e = ma.pinv[p];
                    dst[p] = e >= 0 ? (float)cell_value(gc, rm[e], r0[e], r1[e]) : 0.f;
                }
            }
        }
        __syncthreads();
        // ---- walk the program: one segment per initial sector ----
        double mag = 0.0;
        int m0 = 1, off0 = 0, bto0 = 0, col0 = 0;
        st.ps.init(a.prog + g.prog_off + (walks ? task.chunk : 0), lane);
        const uint64_t *seg_base = st.ps.base;  // the stream just after the segment's SEG word
        int seg_idx = 0;
        while (walks) {  // until the chunk's END word
            const uint64_t w = st.ps.next();
            const int op = (int)((w >> 48) & 0xF);
            if (op == QIW_OP_ROOT) {
                const int m1 = (int)(w & 0x3F);
                const int blk = (int)((w >> 32) & 0xFFFF);
                const int ldr0 = (m0 + 3) & ~3;
                MTile<TR> X;
#pragma unroll
                for (int ti = 0; ti < TR; ++ti)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = 16 * ti + 4 * q + r, col = col0 + cl;
                        float v = 0.f;
                        if (blk == 0xFFFF)
                            v = (row == col && row < m0) ? 1.f : 0.f;
                        else if (row < m1 && col < m0)
                            v = st.ops[blk + row * ldr0 + col];
                        X.t[ti][r] = v;
                    }
                mwalk<KM, TM>(st, n, X, m1, (int)(w >> 52), (int)((w >> 24) & 0xFF));
            } else if (op == QIW_OP_TOPO) {
#pragma unroll
                for (int ai = 0; ai < KM; ++ai) {
                    const int head = (int)((w >> (8 * ai)) & 0xF), tail = (int)((w >> (8 * ai + 4)) & 0xF);
                    const ArcPos ap = arc_pos(s_t[tail], s_t[head], inv_h, a.n_tau);
#pragma unroll
                    for (int tb = 0; tb < TM; ++tb)
                        if (ai < k && tb < ntab) st.W[ai * TM + tb] = delta_value(a.dtab, tb, ap);
                }
            } else if (op == QIW_OP_SEG) {
                m0 = (int)(w & 0x3F);
                off0 = (int)(w >> 52);
                bto0 = ma.bto[(int)((w >> 8) & 0xFF)];
                col0 = 0;
                seg_base = st.ps.base;
                seg_idx = st.ps.idx;
#pragma unroll
                for (int ti = 0; ti < TR; ++ti)
#pragma unroll
                    for (int r = 0; r < 4; ++r) st.acc[ti][r] = 0.0;
            } else if (op == QIW_OP_FLUSH) {  // the s0 block of this sample: iP_n[s0] times the leaf sum
                MTile<TR> S;
#pragma unroll
                for (int ti = 0; ti < TR; ++ti)
#pragma unroll
                    for (int r = 0; r < 4; ++r) S.t[ti][r] = (float)st.acc[ti][r];
                const MTile<TR> R = mmul(pl, (n - 2) * btp + off0, m0, m0, S, st.pl_zero);
#pragma unroll
                for (int ti = 0; ti < TR; ++ti)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = 16 * ti + 4 * q + r, col = col0 + cl;
                        if (row < m0 && col < m0) {
                            const double v = (double)R.t[ti][r];
                            prow[bto0 + row + col * m0] += wt * v;
                            mag += fabs(v);
                        }
                    }
                if (col0 + 16 < m0) {  // the next 16 columns of s0: walk the segment again
                    col0 += 16;
                    st.ps.base = seg_base;
                    st.ps.blk = seg_base[lane];
                    st.ps.idx = seg_idx;
#pragma unroll
                    for (int ti = 0; ti < TR; ++ti)
#pragma unroll
                        for (int r = 0; r < 4; ++r) st.acc[ti][r] = 0.0;
                }
            } else if (op == QIW_OP_END) {
                break;
            }
        }
        if (a.per_sample && walks) {
#pragma unroll
            for (int o = QIW_WAVE / 2; o > 0; o >>= 1) mag += __shfl_xor(mag, o, QIW_WAVE);
            if (lane == 0) a.per_sample[g.ps_off + sample] = mag;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------
// Device-side step bookkeeping: reduce the partial sums of every group in fixed task order,
// apply the integration prefactors, form the order-resolved and total step results
// (inchworm_step, src/inchworm.jl:142-202), store i*P(tau_f) (set_ppgf!,
// src/ppgf.jl:474-504) and rescale (normalize!(P, tau_f), src/ppgf.jl:646-668).
// ---------------------------------------------------------------------------------------

__device__ double simplex_volume(int d, double edge) {  // _simplex_volume, src/qmc_integrate.jl:46
    double p = 1.0;
    for (int i = 1; i <= d; ++i) p *= edge / i;
    return p;
}

// Compensated (Kahan) summation of the qMC sums: the running sum carries the low-order part
// lost by each addition.
struct KahanSum {
    double s = 0.0, c = 0.0;
    __device__ __forceinline__ void add(double x) {
        const double y = x - c, t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

// Sum over the task rows of each group, one block per group (and per 256 elements when
// bt > 256).  The rows of a group are contiguous, so the block reads them as one
// contiguous range: thread t covers element t % bt of rows t / bt, t / bt + R, ... with R
// = 256 / bt rows per pass, and the R partial sums of each element are then added in a
// fixed order, so the result does not depend on scheduling.  Both stages are compensated.
__device__ void reduce_rows_block(const int32_t *__restrict__ first, const int32_t *__restrict__ count,
                                  const double *__restrict__ partial, int bt, double *__restrict__ gsum) {
    __shared__ double red[256], cmp[256];
    const int g = blockIdx.x, t = threadIdx.x;
    const int f0 = first[g], nr = count[g];
    const int R = bt <= 256 ? 256 / bt : 1;     // rows per pass
    const int e = bt <= 256 ? t % bt : blockIdx.y * 256 + t;
    const int r0 = bt <= 256 ? t / bt : 0;
    // four independent compensated sums over rows r0, r0 + R, ... (round robin), so four
    // loads are in flight at a time; combined in a fixed order
    KahanSum k4[4];
    if (r0 < R && e < bt) {
        const double *p = partial + (int64_t)(f0 + r0) * bt + e;
        const int64_t step = (int64_t)R * bt;
        int r = r0;
        for (; r + 3 * R < nr; r += 4 * R, p += 4 * step) {
            const double x0 = p[0], x1 = p[step], x2 = p[2 * step], x3 = p[3 * step];
            k4[0].add(x0);
            k4[1].add(x1);
            k4[2].add(x2);
            k4[3].add(x3);
        }
        for (int q = 0; r < nr; r += R, p += step, ++q) k4[q].add(p[0]);
    }
    KahanSum k;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        k.add(k4[q].s);
        k.add(-k4[q].c);
    }
    red[t] = k.s;
    cmp[t] = k.c;
    __syncthreads();
    if (t < (bt <= 256 ? bt : 256) && e < bt) {
        KahanSum v;
        for (int q = 0; q < R; ++q) {
            const int i = q * (bt <= 256 ? bt : 0) + t;
            v.add(red[i]);
            v.add(-cmp[i]);
        }
        gsum[(int64_t)g * bt + e] = v.s;
    }
}

// Step bookkeeping on the per-group sums gsum, run by one block of 256 threads.  gsum is
// read through a volatile pointer: other blocks of the same launch wrote it.  The group
// prefactors are computed once per group (in parallel), each thread then sums its
// elements over the groups in group order, per order slot in registers.
constexpr int kFinMaxGroups = 64;
constexpr int kFinMaxSlots = QIW_KMAX + 1;

// Prefactor of a group's sample sum: the sign (-1)^k of the order, the volume of the
// integration domain (RootTransform / DoubleSimplexRootTransform Jacobians) and 1/N.
__device__ double step_group_factor(const QiwGroup &g, const FinalizeArgs &a) {
    if (!g.sampled) return 1.0;
    const double tau_i = 0.0, tau_w = a.tau_w, tau_f = a.tau_f;
    // self-energy: RootTransform of 2k - 2 times on [tau_i, tau_f_prev = tau_w]; the group
    // sum times this factor is sigma with Σ = i sigma
    const double J = g.kind == QIW_G_SIGMA ? simplex_volume(2 * g.order - 2, tau_w - tau_i)
                     : a.bare ? simplex_volume(g.d_after, tau_f - tau_i)
                              : simplex_volume(g.d_before, tau_w - tau_i) * simplex_volume(g.d_after, tau_f - tau_w);
    return ((g.order & 1) ? -1.0 : 1.0) * J / (double)g.n_samples;
}

// normalize!(P, tau_f) after row jf was stored (src/ppgf.jl:646-668): lambda from the
// largest diagonal element of row jf, then every known row j scaled by exp(-tau_j lambda)
// (one exp per row).
__device__ void step_normalize(const FinalizeArgs &a) {
    const int tid = threadIdx.x, jf = a.jf;
    __threadfence_block();
    __syncthreads();
    __shared__ double lam_sh;
    if (tid == 0) {
        double pmax = -INFINITY;
        for (int i = 0; i < a.n_diag; ++i) pmax = fmax(pmax, a.ptab[(int64_t)jf * a.bt + a.diag_idx[i]]);
        lam_sh = log(pmax) / a.tau_f;
        a.lambda_out[jf] = lam_sh;
    }
    __syncthreads();
    const double lam = lam_sh;
    for (int j = tid; j < a.n_known; j += blockDim.x) {
        const double f = exp(-(j * a.h) * lam);
        double *row = a.ptab + (int64_t)j * a.bt;
        for (int e = 0; e < a.bt; ++e) row[e] *= f;
    }
}

__device__ void finalize_step_body(const FinalizeArgs &a) {
    const volatile double *gsum = a.gsum;
    const int tid = threadIdx.x;
    const int jf = a.jf;
    __shared__ double fac_sh[kFinMaxGroups];
    __shared__ int slot_sh[kFinMaxGroups];
    for (int gi = tid; gi < a.n_groups; gi += blockDim.x) {
        fac_sh[gi] = step_group_factor(a.groups[gi], a);
        slot_sh[gi] = a.group_order_slot[gi];
    }
    __syncthreads();
    for (int e = tid; e < a.bt; e += blockDim.x) {
        double acc[kFinMaxSlots];
#pragma unroll
        for (int o = 0; o < kFinMaxSlots; ++o) acc[o] = 0.0;
        for (int gi = 0; gi < a.n_groups; ++gi) {
            const double v = fac_sh[gi] * gsum[(int64_t)gi * a.bt + e];
            const int sl = slot_sh[gi];
#pragma unroll
            for (int o = 0; o < kFinMaxSlots; ++o) acc[o] += (o == sl) ? v : 0.0;
        }
        double tot = 0.0;
#pragma unroll
        for (int o = 0; o < kFinMaxSlots; ++o) {
            if (o < a.n_order_slots) {
                a.porders[((int64_t)o * a.n_tau + jf) * a.bt + e] = acc[o];
                tot += acc[o];
            }
        }
        a.ptab[(int64_t)a.jt * a.bt + e] = tot;
    }
    if (a.normalize) step_normalize(a);
}

// One launch per step: every block reduces its group's rows; the last block to finish
// (a counter every block increments after its sums are visible) runs the bookkeeping and
// resets the counter for the next step.  This saves a launch and its gap per step.
__global__ __launch_bounds__(256) void k_step_reduce_finalize(FinalizeArgs a) {
    reduce_rows_block(a.group_task_first, a.group_task_count, a.partial, a.bt, a.gsum);
    __threadfence();
    __syncthreads();
    __shared__ int last;
    if (threadIdx.x == 0) last = atomicAdd(a.done, 1u) == gridDim.x * gridDim.y - 1;
    __syncthreads();
    if (!last) return;
    if (threadIdx.x == 0) *a.done = 0u;
    finalize_step_body(a);
}

// Correlator bookkeeping (src/inchworm.jl:825-889): per (pair, tau) row, sum the groups'
// prefactor-weighted traces.  out[c][j] = sum_g (-1)^k J_g / N_g * R_g (real; the value
// is -i * out / Z).

// One block per tau row: for each group the block sums the diagonal elements of the
// group's task rows (items spread over the threads, then a fixed-order tree), scaled by
// the group's sign, simplex volume and 1/N.
__global__ __launch_bounds__(256) void k_finalize_corr(CorrFinalizeArgs a) {
    __shared__ double red[256];
    const int row = blockIdx.x, t = threadIdx.x;
    const int j = row % a.n_tau;
    const double tau_i = 0.0, tau_w = j * a.h, tau_f = a.beta;
    double tot = 0.0;
    for (int gi = 0; gi < a.n_groups; ++gi) {
        const int nt = a.row_task_count[row * a.n_groups + gi];
        if (nt == 0) continue;
        const int f0 = a.row_task_first[row * a.n_groups + gi];
        const int64_t items = (int64_t)nt * a.n_diag;
        KahanSum s;
        for (int64_t i = t; i < items; i += 256) {
            const int r = (int)(i / a.n_diag), e = (int)(i % a.n_diag);
            s.add(a.partial[(int64_t)(f0 + r) * a.bt + a.diag_idx[e]]);
        }
        red[t] = s.s - s.c;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (t < w) red[t] += red[t + w];
            __syncthreads();
        }
        if (t == 0) {
            const QiwGroup g = a.groups[gi];
            double fac = 1.0;
            if (g.sampled) {
                double J = simplex_volume(g.d_before, tau_w - tau_i) * simplex_volume(g.d_after, tau_f - tau_w);
                fac = ((g.order & 1) ? -1.0 : 1.0) * J / (double)g.n_samples;
            }
            tot += fac * red[0];
        }
        __syncthreads();
    }
    if (t == 0) a.out[row] = tot;
}

// ---------------------------------------------------------------------------------------
// Randomized qMC (mean_std_from_randomization, src/randomization.jl:86-100)
// ---------------------------------------------------------------------------------------

// Sample statistics of sequences 0..n-1 of one element, x(s) its value for sequence s:
// the sum in sequence order, the mean as Statistics.var takes it (sum / n) and the
// corrected variance sum |x - mean|^2 / (n - 1) (src/sector_block_matrix.jl:191-223).
// Every operation rounds on its own, as in the reference (no fused multiply-adds: a fused
// x - mean would give equal samples a deviation of the size of their rounding error).
template <class X>
__device__ double seq_sum(X x, int n) {
#pragma clang fp contract(off)
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += x(i);
    return s;
}
template <class X>
__device__ double seq_std(X x, int n) {
#pragma clang fp contract(off)
    const double mean = seq_sum(x, n) / n;
    double v = 0.0;
    for (int i = 0; i < n; ++i) {
        const double d = x(i) - mean;
        v += d * d;
    }
    return sqrt(v / (n - 1));
}

// Simplex coordinates of one step's randomized-qMC sequences; grid: 256-sample blocks x
// groups.  The block scrambles the staged direction numbers of its group's dimensions
// with the bits of the group's stream (LMS + digital shift, src/scrambled_sobol.jl:102-140:
// row p of the lower-triangular matrix holds entries (p, q < p) in bits 31 - q and 1 on
// the diagonal; bit 31 - p of a scrambled direction number is the parity of row p AND the
// plain one), then maps its samples as k_simplex_coords does.
__global__ __launch_bounds__(256) void k_rqmc_coords(const QiwGroup *__restrict__ groups, const uint32_t *__restrict__ m,
                                                     int64_t n, uint64_t seed, uint64_t domain, uint64_t key,
                                                     double *__restrict__ ut) {
    const QiwGroup g = groups[blockIdx.y];
    if (!g.sampled) return;
    const bool sigma = g.kind == QIW_G_SIGMA;
    const int d = sigma ? 2 * g.order - 2 : g.d_before + g.d_after;  // Sobol dimensions used
    __shared__ uint32_t rows[2 * QIW_KMAX][33];
    __shared__ uint32_t ms[2 * QIW_KMAX * 32];
    // self-energy groups are keyed like the reference's topology data, n_pts_after = 1
    const uint64_t stream = qiw_rq_stream(domain, key, g.order, sigma ? 1 : g.n_after, g.seq);
    for (int i = threadIdx.x; i < d * 33; i += blockDim.x) {
        const int k = i / 33, w = i % 33;
        uint32_t v = qiw_rq_word(seed, stream, k, w);
        if (w > 0) {
            const int p = w - 1;
            const uint32_t below = p == 0 ? 0u : (0xFFFFFFFFu << (32 - p));
            v = (v & below) | (1u << (31 - p));
        }
        rows[k][w] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < d * 32; i += blockDim.x) {
        const int k = i / 32, j = i % 32;
        const uint32_t mkj = m[k * 32 + j];
        uint32_t v = 0;
        for (int p = 0; p < 32; ++p) v |= (uint32_t)(__popc(rows[k][1 + p] & mkj) & 1) << (31 - p);
        ms[k * 32 + j] = v;
    }
    __syncthreads();
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ng = g.n_samples;  // <= n
    if (r >= ng) return;
    double x[2 * QIW_KMAX], u[2 * QIW_KMAX];
    for (int k = 0; k < d; ++k) x[k] = ldexp((double)sobol_word(ms, rows[k][0], k, (uint64_t)r), -32);
    double *out = ut + g.ut_off;
    if (sigma) {
        sigma_row_values(x, g.order, u);
        for (int k = 0; k < 2 * g.order; ++k) out[(int64_t)k * ng + r] = u[k];
        return;
    }
    root_map(x, g.d_after, g.d_before, u);
    for (int k = 0; k < d; ++k) out[(int64_t)k * ng + r] = u[k];
}

// Simplex coordinates from caller-drawn sequences (the states of the ScrambledSobolSeq
// objects the reference constructs with the caller's RNG, src/randomization.jl:94 and
// src/scrambled_sobol.jl:66-143); grid: 256-sample blocks x groups x rows.  Row r of group
// g maps sequence idx[r * n_groups + g] (gm: [.][stride][32] direction numbers, gx:
// [.][stride] shifts) as k_rqmc_coords maps a scrambled one, into ut + (row0 + r) *
// row_stride.
__global__ __launch_bounds__(256) void k_given_coords(const QiwGroup *__restrict__ groups,
                                                      const uint32_t *__restrict__ gm, const uint32_t *__restrict__ gx,
                                                      int stride, int64_t n, const int64_t *__restrict__ idx,
                                                      int row0, int64_t row_stride, double *__restrict__ ut) {
    const QiwGroup g = groups[blockIdx.y];
    const int64_t seq = idx[(int64_t)blockIdx.z * gridDim.y + blockIdx.y];
    if (!g.sampled) return;
    const bool sigma = g.kind == QIW_G_SIGMA;
    const int d = sigma ? 2 * g.order - 2 : g.d_before + g.d_after;  // Sobol dimensions used
    // the exact order-1 self-energy group uses no sequence (d = 0) but still has its row
    if (seq < 0 && d > 0) return;
    __shared__ uint32_t ms[2 * QIW_KMAX * 32];
    __shared__ uint32_t xs[2 * QIW_KMAX];
    for (int i = threadIdx.x; i < d * 32; i += blockDim.x) ms[i] = gm[(seq * stride + i / 32) * 32 + i % 32];
    for (int i = threadIdx.x; i < d; i += blockDim.x) xs[i] = gx[seq * stride + i];
    __syncthreads();
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ng = g.n_samples;  // <= n
    if (r >= ng) return;
    double x[2 * QIW_KMAX], u[2 * QIW_KMAX];
    for (int k = 0; k < d; ++k) x[k] = ldexp((double)sobol_word(ms, xs[k], k, (uint64_t)r), -32);
    double *out = ut + (int64_t)(row0 + blockIdx.z) * row_stride + g.ut_off;
    if (sigma) {
        sigma_row_values(x, g.order, u);
        for (int k = 0; k < 2 * g.order; ++k) out[(int64_t)k * ng + r] = u[k];
        return;
    }
    root_map(x, g.d_after, g.d_before, u);
    for (int k = 0; k < d; ++k) out[(int64_t)k * ng + r] = u[k];
}

// Simplex coordinates from caller-supplied points (the sequence plug point seq_type of
// mean_std_from_randomization, src/randomization.jl:86-99: any sequence the caller
// constructs as seq_type(D; scramble_rng) and draws with next!, e.g. RandomSeq,
// src/utility.jl:336-385).  Row r of group g maps the first n_samples points of sequence
// idx[r * n_groups + g] (pts: [.][npts][stride]) as k_given_coords maps Sobol points.
__global__ __launch_bounds__(256) void k_points_coords(const QiwGroup *__restrict__ groups,
                                                       const double *__restrict__ pts, int stride, int64_t npts,
                                                       const int64_t *__restrict__ idx, int row0, int64_t row_stride,
                                                       double *__restrict__ ut) {
    const QiwGroup g = groups[blockIdx.y];
    const int64_t seq = idx[(int64_t)blockIdx.z * gridDim.y + blockIdx.y];
    if (!g.sampled) return;
    const bool sigma = g.kind == QIW_G_SIGMA;
    const int d = sigma ? 2 * g.order - 2 : g.d_before + g.d_after;
    if (seq < 0 && d > 0) return;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ng = g.n_samples;  // <= npts
    if (r >= ng) return;
    double x[2 * QIW_KMAX], u[2 * QIW_KMAX];
    const double *p = pts + ((seq < 0 ? 0 : seq) * npts + r) * stride;
    for (int k = 0; k < d; ++k) x[k] = p[k];
    double *out = ut + (int64_t)(row0 + blockIdx.z) * row_stride + g.ut_off;
    if (sigma) {
        sigma_row_values(x, g.order, u);
        for (int k = 0; k < 2 * g.order; ++k) out[(int64_t)k * ng + r] = u[k];
        return;
    }
    root_map(x, g.d_after, g.d_before, u);
    for (int k = 0; k < d; ++k) out[(int64_t)k * ng + r] = u[k];
}

// Step bookkeeping with randomized qMC (inchworm_step, src/inchworm.jl:172-195), one
// block of 256 threads after all per-sequence sums are in gsum.  For each (order,
// n_pts_after) group, the prefactor-scaled sums of its sequences are the samples; the
// sequences are taken in order until, from the second on, the largest element of their
// standard deviation is <= target_std.  The group adds the mean (sum * (1/n), as
// SectorBlockMatrix / n) and the deviation over the sequences taken to its order's rows.
// Thread t owns elements t, t + 256, ... throughout, so each element's sums run in group
// order.
__device__ void finalize_step_rqmc_body(const FinalizeArgs &a) {
    const volatile double *gsum = a.gsum;
    const int tid = threadIdx.x, jf = a.jf, bt = a.bt;
    __shared__ double red[256];
    for (int o = 0; o < a.n_order_slots; ++o)
        for (int e = tid; e < bt; e += blockDim.x) {
            a.porders[((int64_t)o * a.n_tau + jf) * bt + e] = 0.0;
            a.porders_std[((int64_t)o * a.n_tau + jf) * bt + e] = 0.0;
        }
    for (int r = 0; r < a.n_real; ++r) {
        const int base = a.real_first[r], cnt = a.real_count[r];
        const QiwGroup g = a.groups[base];
        const double fac = step_group_factor(g, a);
        const int slot = a.group_order_slot[base];
        const bool exact = qiw_group_exact(g);
        const int nmax = exact ? 1 : a.n_seqs;
        int nuse = nmax;
        for (int n = 2; n <= nmax; ++n) {
            double loc = 0.0;
            for (int e = tid; e < bt; e += blockDim.x) {
                auto x = [&](int s) { return __dmul_rn(fac, gsum[(int64_t)(base + (cnt > 1 ? s : 0)) * bt + e]); };
                loc = fmax(loc, seq_std(x, n));
            }
            red[tid] = loc;
            __syncthreads();
            for (int w = 128; w > 0; w >>= 1) {
                if (tid < w) red[tid] = fmax(red[tid], red[tid + w]);
                __syncthreads();
            }
            const bool stop = red[0] <= a.target_std;
            __syncthreads();
            if (stop) {
                nuse = n;
                break;
            }
        }
        if (a.nuse_out && tid == 0) a.nuse_out[(int64_t)jf * a.n_real + r] = nuse;
        for (int e = tid; e < bt; e += blockDim.x) {
            auto x = [&](int s) { return __dmul_rn(fac, gsum[(int64_t)(base + (cnt > 1 ? s : 0)) * bt + e]); };
            const double mean = __dmul_rn(seq_sum(x, nuse), 1.0 / nuse);
            const double sd = exact ? 0.0 : (nuse > 1 ? seq_std(x, nuse) : __builtin_nan(""));
            a.porders[((int64_t)slot * a.n_tau + jf) * bt + e] += mean;  // after the group's mean is rounded
            a.porders_std[((int64_t)slot * a.n_tau + jf) * bt + e] += sd;
        }
    }
    for (int e = tid; e < bt; e += blockDim.x) {
        double tot = 0.0;
        for (int o = 0; o < a.n_order_slots; ++o) tot += a.porders[((int64_t)o * a.n_tau + jf) * bt + e];
        a.ptab[(int64_t)a.jt * bt + e] = tot;
    }
    if (a.normalize) step_normalize(a);
}

__global__ __launch_bounds__(256) void k_step_reduce_finalize_rqmc(FinalizeArgs a) {
    reduce_rows_block(a.group_task_first, a.group_task_count, a.partial, a.bt, a.gsum);
    __threadfence();
    __syncthreads();
    __shared__ int last;
    if (threadIdx.x == 0) last = atomicAdd(a.done, 1u) == gridDim.x * gridDim.y - 1;
    __syncthreads();
    if (!last) return;
    if (threadIdx.x == 0) *a.done = 0u;
    finalize_step_rqmc_body(a);
}

// Euler step of diff_inchworm_step! (src/inchworm.jl:603-635) after the self-energy row
// is stored: with p = i P and Σ = i sigma (real),
//   p(tau_{a+1}) = p(tau_a) + h (sum_j w_j sigma(tau_a - tau_j) p(tau_j) - H p(tau_a)),
// w the trapezoid weights of kd.integrate over grid[0..a] (the contour measure -i dtau
// turns Σ P into sigma p), H = diag(E + lambda).  Per element of each sector block.
__device__ void diff_euler(const FinalizeArgs &a) {
    __threadfence_block();
    __syncthreads();
    const int A = a.a, bt = a.bt;
    for (int e = threadIdx.x; e < bt; e += blockDim.x) {
        int s = 0;
        while (s + 1 < a.ns && a.bt_off[s + 1] <= e) ++s;
        const int m = a.dims[s], o = a.bt_off[s], r = (e - o) % m, c = (e - o) / m;
        double conv = 0.0;
        for (int j = 0; j <= A && A > 0; ++j) {
            const double w = (j == 0 || j == A) ? 0.5 * a.h : a.h;
            const double *S = a.ptab + (int64_t)(A - j) * bt + o;
            const double *Pj = a.P + (int64_t)j * bt + o;
            double t = 0.0;
            for (int l = 0; l < m; ++l) t += S[r + l * m] * Pj[l + c * m];
            conv += w * t;
        }
        const double pp = a.P[(int64_t)A * bt + e];
        a.P[(int64_t)(A + 1) * bt + e] = pp + a.h * (conv - a.hdiag[o + r * (m + 1)] * pp);
    }
}

// One launch per differential step: per-group row sums, then (last workgroup) the
// self-energy bookkeeping as for an inchworm step (ptab = sigma table, no normalization)
// and the Euler step.
__global__ __launch_bounds__(256) void k_diff_reduce_finalize(FinalizeArgs a, int rqmc) {
    reduce_rows_block(a.group_task_first, a.group_task_count, a.partial, a.bt, a.gsum);
    __threadfence();
    __syncthreads();
    __shared__ int last;
    if (threadIdx.x == 0) last = atomicAdd(a.done, 1u) == gridDim.x * gridDim.y - 1;
    __syncthreads();
    if (!last) return;
    if (threadIdx.x == 0) *a.done = 0u;
    if (rqmc)
        finalize_step_rqmc_body(a);
    else
        finalize_step_body(a);
    diff_euler(a);
}

// Correlator bookkeeping with randomized qMC (correlator_2p, src/inchworm.jl:857-885):
// one block per tau row; per group, the prefactor-scaled trace sums of its sequences are
// the samples (as in k_finalize_corr), reduced to the mean (sum / n) and the deviation
// with the same stopping rule; the row's value and deviation are their sums over groups.
__global__ __launch_bounds__(256) void k_finalize_corr_rqmc(CorrFinalizeArgs a) {
    __shared__ double red[256];
    __shared__ double xs[QIW_MAX_SEQS];
    const int row = blockIdx.x, t = threadIdx.x;
    const int j = row % a.n_tau;
    const double tau_i = 0.0, tau_w = j * a.h, tau_f = a.beta;
    double tot = 0.0, tot_std = 0.0;
    for (int r = 0; r < a.n_real; ++r) {
        const int base = a.real_first[r], cnt = a.real_count[r];
        if (a.row_task_count[row * a.n_groups + base] == 0) continue;  // tau_A = tau_B: order 0 only
        const QiwGroup g = a.groups[base];
        double fac = 1.0;
        if (g.sampled) {
            double J = simplex_volume(g.d_before, tau_w - tau_i) * simplex_volume(g.d_after, tau_f - tau_w);
            fac = ((g.order & 1) ? -1.0 : 1.0) * J / (double)g.n_samples;
        }
        for (int q = 0; q < cnt; ++q) {
            const int gi = base + q;
            const int nt = a.row_task_count[row * a.n_groups + gi];
            const int f0 = a.row_task_first[row * a.n_groups + gi];
            const int64_t items = (int64_t)nt * a.n_diag;
            KahanSum s;
            for (int64_t i = t; i < items; i += 256) {
                const int rr = (int)(i / a.n_diag), e = (int)(i % a.n_diag);
                s.add(a.partial[(int64_t)(f0 + rr) * a.bt + a.diag_idx[e]]);
            }
            red[t] = s.s - s.c;
            __syncthreads();
            for (int w = 128; w > 0; w >>= 1) {
                if (t < w) red[t] += red[t + w];
                __syncthreads();
            }
            if (t == 0) xs[q] = fac * red[0];
            __syncthreads();
        }
        if (t == 0) {
            auto x = [&](int s) { return xs[cnt > 1 ? s : 0]; };
            const int nmax = qiw_group_exact(g) ? 1 : a.n_seqs;
            int nuse = nmax;
            for (int n = 2; n <= nmax; ++n)
                if (seq_std(x, n) <= a.target_std) {
                    nuse = n;
                    break;
                }
            if (a.nuse_out) a.nuse_out[(int64_t)row * a.n_real + r] = nuse;
            tot += seq_sum(x, nuse) / nuse;
            tot_std += qiw_group_exact(g) ? 0.0 : (nuse > 1 ? seq_std(x, nuse) : __builtin_nan(""));
        }
        __syncthreads();
    }
    if (t == 0) {
        a.out[row] = tot;
        a.out_std[row] = tot_std;
    }
}

// Zero n 32-bit words (clears counters and output tables with a kernel instead of a memset
// node, so the stream work stays capturable).
__global__ void k_zero_words(uint32_t *__restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ---------------------------------------------------------------------------------------
// Host-callable launch wrappers (plain C++ signatures for qiw_host.cpp)
// ---------------------------------------------------------------------------------------
hipError_t qiw_launch_zero(hipStream_t st, void *p, size_t bytes) {
    const int64_t n = (int64_t)(bytes / 4);
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_words, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (uint32_t *)p, n);
    return hipGetLastError();
}
hipError_t qiw_launch_sobol_points(hipStream_t st, const uint32_t *m, const uint32_t *x0, int dim, int64_t first,
                                   int64_t n, double *out) {
    int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(k_sobol_points, dim3(blocks), dim3(256), 0, st, m, x0, dim, first, n, out);
    return hipGetLastError();
}

hipError_t qiw_launch_simplex_coords(hipStream_t st, const uint32_t *m, const uint32_t *x0, int d_before,
                                     int d_after, int64_t n, double *out) {
    int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(k_simplex_coords, dim3(blocks), dim3(256), 0, st, m, x0, d_before, d_after, n, out);
    return hipGetLastError();
}

hipError_t qiw_launch_sobol_cube(hipStream_t st, const uint32_t *m, const uint32_t *x0, int dim, int stride,
                                 int64_t n, double *out) {
    int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(k_sobol_cube, dim3(blocks), dim3(256), 0, st, m, x0, dim, stride, n, out);
    return hipGetLastError();
}

hipError_t qiw_launch_warp_coords(hipStream_t st, const double *z, int stride, const double *logj, int d_before,
                                  int d_after, int64_t n, double *out) {
    int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(k_warp_coords, dim3(blocks), dim3(256), 0, st, z, stride, logj, d_before, d_after, n, out);
    return hipGetLastError();
}

hipError_t qiw_launch_simplex_points(hipStream_t st, const uint32_t *m, const uint32_t *x0, int d_before,
                                     int d_after, double t_i, double t_w, double t_f, int64_t n, double *out) {
    int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(k_simplex_points, dim3(blocks), dim3(256), 0, st, m, x0, d_before, d_after, t_i, t_w,
                       t_f, n, out);
    return hipGetLastError();
}

// Register budget of the scalar walk: KM = largest order, SM = sectors, TM = tables.
// The instantiation (KM, SM, TM) is the one the host compiler laid the program out for:
// qiw_scalar_km / _sm / _tm of the launch's largest order, sectors and tables.
hipError_t qiw_launch_eval_flat(hipStream_t st, const EvalArgs &a, int n_tasks, int km, int sm, int tm) {
    if (km > QIW_KMAX || sm > 8 || tm > 4) return hipErrorInvalidValue;
    const int KM = qiw_scalar_km(km, sm, tm), SM = qiw_scalar_sm(km, sm, tm), TM = qiw_scalar_tm(km, sm, tm);
    // the pole sums of DLR tables cost registers: instantiations without them for the rest
    const bool dlr = a.dtab.dlr_mask != 0;
    if (KM == 4 && SM == 4 && TM == 2 && !dlr)
        hipLaunchKernelGGL((k_eval_flat_t<4, 4, 2, false>), dim3(n_tasks), dim3(QIW_WAVE), 0, st, a);
    else if (KM == 4 && SM == 4 && TM == 2)
        hipLaunchKernelGGL((k_eval_flat_t<4, 4, 2, true>), dim3(n_tasks), dim3(QIW_WAVE), 0, st, a);
    else if (KM == 4 && !dlr)
        hipLaunchKernelGGL((k_eval_flat_t<4, 8, 4, false>), dim3(n_tasks), dim3(QIW_WAVE), 0, st, a);
    else if (KM == 4)
        hipLaunchKernelGGL((k_eval_flat_t<4, 8, 4, true>), dim3(n_tasks), dim3(QIW_WAVE), 0, st, a);
    else
        hipLaunchKernelGGL((k_eval_flat_t<QIW_KMAX, 8, 4, true>), dim3(n_tasks), dim3(QIW_WAVE), 0, st, a);
    return hipGetLastError();
}

// Register budget of the vector walk: MB = largest block, KM = largest order, TM = tables;
// fp32 selects float32 chain products.
#ifndef QIW_VEC_NC32
#define QIW_VEC_NC32 2
#endif
// float32 chains carry QIW_VEC_NC32 columns per pass; float64 chains one (register budget).
hipError_t qiw_launch_eval_vec(hipStream_t st, const VecArgs &v, int n_tasks, int mb, int km, int tm, int fp32) {
    constexpr int NC = QIW_VEC_NC32;
    if (mb <= 6 && km <= 4 && tm <= 2) {
        if (fp32)
            hipLaunchKernelGGL((k_eval_vec_t<float, 6, 4, 2, NC>), dim3(n_tasks), dim3(QIW_WAVE), 0, st, v);
        else
            hipLaunchKernelGGL((k_eval_vec_t<double, 6, 4, 2, 1>), dim3(n_tasks), dim3(QIW_WAVE), 0, st, v);
    } else if (mb <= 6 && km <= 4 && tm <= 4) {
        if (fp32)
            hipLaunchKernelGGL((k_eval_vec_t<float, 6, 4, 4, NC>), dim3(n_tasks), dim3(QIW_WAVE), 0, st, v);
        else
            hipLaunchKernelGGL((k_eval_vec_t<double, 6, 4, 4, 1>), dim3(n_tasks), dim3(QIW_WAVE), 0, st, v);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t qiw_launch_ptab_cells(hipStream_t st, const double *T, int n_tau, int bt, double2 *out, double *outm,
                                 QiwCell32 *out32) {
    int64_t n = (int64_t)(n_tau - 1) * bt;
    hipLaunchKernelGGL(k_ptab_cells, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, T, n_tau, bt, out, outm,
                       out32);
    return hipGetLastError();
}

// Wide blocks on the matrix cores: orders up to 3 or 4, blocks up to 32, <= 2 tables.
hipError_t qiw_launch_eval_mfma(hipStream_t st, const MfmaArgs &m, int n_tasks, int km, int tm, size_t lds_bytes) {
    if (tm > 2 || n_tasks % QIW_MFMA_WAVES) return hipErrorInvalidValue;
    if (km <= 3)
        hipLaunchKernelGGL((k_eval_mfma_t<3, 2, 2, QIW_MFMA_WAVES>), dim3(n_tasks / QIW_MFMA_WAVES),
                           dim3(QIW_WAVE * QIW_MFMA_WAVES), lds_bytes, st, m);
    else if (km <= 4)
        hipLaunchKernelGGL((k_eval_mfma_t<4, 2, 2, QIW_MFMA_WAVES>), dim3(n_tasks / QIW_MFMA_WAVES),
                           dim3(QIW_WAVE * QIW_MFMA_WAVES), lds_bytes, st, m);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t qiw_launch_eval_block(hipStream_t st, const BlockArgs &b, int n_tasks, size_t lds_bytes) {
    hipLaunchKernelGGL(k_eval_block, dim3(n_tasks), dim3(QIW_WAVE), lds_bytes, st, b);
    return hipGetLastError();
}

hipError_t qiw_launch_finalize_step(hipStream_t st, const FinalizeArgs &a) {
    hipLaunchKernelGGL(k_step_reduce_finalize, dim3(a.n_groups, a.bt <= 256 ? 1 : (a.bt + 255) / 256), dim3(256), 0,
                       st, a);
    return hipGetLastError();
}

hipError_t qiw_launch_finalize_corr(hipStream_t st, const CorrFinalizeArgs &a) {
    hipLaunchKernelGGL(k_finalize_corr, dim3(a.n_rows), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t qiw_launch_finalize_step_rqmc(hipStream_t st, const FinalizeArgs &a) {
    if (a.n_seqs < 1 || a.n_seqs > QIW_MAX_SEQS) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_step_reduce_finalize_rqmc, dim3(a.n_groups, a.bt <= 256 ? 1 : (a.bt + 255) / 256),
                       dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t qiw_launch_finalize_diff(hipStream_t st, const FinalizeArgs &a, int rqmc) {
    if (a.n_seqs < 1 || a.n_seqs > QIW_MAX_SEQS || a.normalize) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_diff_reduce_finalize, dim3(a.n_groups, a.bt <= 256 ? 1 : (a.bt + 255) / 256), dim3(256), 0,
                       st, a, rqmc);
    return hipGetLastError();
}

hipError_t qiw_launch_sigma_coords(hipStream_t st, const uint32_t *m, const uint32_t *x0, int order, int64_t n,
                                   double *out) {
    if (order < 1 || order > QIW_KMAX || n < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_sigma_coords, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, m, x0, order, n, out);
    return hipGetLastError();
}

hipError_t qiw_launch_finalize_corr_rqmc(hipStream_t st, const CorrFinalizeArgs &a) {
    if (a.n_seqs < 1 || a.n_seqs > QIW_MAX_SEQS) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_finalize_corr_rqmc, dim3(a.n_rows), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t qiw_launch_given_coords(hipStream_t st, const QiwGroup *groups, int n_groups, int64_t n, const GivenSrc &src,
                                   const int64_t *idx, int n_rows, int row0, int64_t row_stride, double *ut) {
    if (n_groups < 1 || n < 1 || n_rows < 1 || src.stride < 1 || src.stride > 2 * QIW_KMAX) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((n + 255) / 256), (unsigned)n_groups, (unsigned)n_rows);
    if (src.pts) {
        if (src.npts < n) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_points_coords, grid, dim3(256), 0, st, groups, src.pts, src.stride, src.npts, idx, row0,
                           row_stride, ut);
    } else {
        hipLaunchKernelGGL(k_given_coords, grid, dim3(256), 0, st, groups, src.gm, src.gx, src.stride, n, idx, row0,
                           row_stride, ut);
    }
    return hipGetLastError();
}

hipError_t qiw_launch_rqmc_coords(hipStream_t st, const QiwGroup *groups, int n_groups, int max_dim, int64_t n,
                                  const uint32_t *m, uint64_t seed, int domain, int64_t key, double *ut) {
    if (max_dim > 2 * QIW_KMAX || n_groups < 1 || n < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_rqmc_coords, dim3((unsigned)((n + 255) / 256), (unsigned)n_groups), dim3(256), 0, st,
                       groups, m, n, seed, (uint64_t)domain, (uint64_t)key, ut);
    return hipGetLastError();
}
