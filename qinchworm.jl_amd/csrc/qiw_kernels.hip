// qiw_kernels.hip -- gfx950 kernels of libqiw: Sobol points and simplex maps, the
// trie-walking diagram evaluator, and the device-side step bookkeeping.
//
// Work mapping of the evaluator: one wavefront of 64 lanes = 64 consecutive qMC samples of
// one (order, n_pts_after) group.  All lanes walk the same precompiled configuration trie
// (the depth-first traversal of src/topology_eval.jl:454-556, flattened on the host), so
// control flow is uniform: program words are read through the scalar cache and only the
// per-sample values differ between lanes.  Per-lane operands (propagator values per
// segment and sector, hybridization values per arc, the partial-product stack) live in LDS
// as [slot][lane] columns, so every access is a conflict-free contiguous 512-byte row.
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
// i*P_s(dt) by linear interpolation of the grid table in the time difference
// (the evaluation of expansion.P behind kd.interpolate!, src/topology_eval.jl:368).
__device__ __forceinline__ double ptab_lookup(const double *__restrict__ T, int ld, int s, int ntau, double inv_h,
                                              double dt) {
    double x = dt * inv_h;
    int j = (int)floor(x);
    j = j < 0 ? 0 : (j > ntau - 2 ? ntau - 2 : j);
    double f = x - (double)j;
    return (1.0 - f) * T[j * ld + s] + f * T[(j + 1) * ld + s];
}

// i*Delta_tab(dt) from per-interval cubic coefficients (natural cubic spline,
// src/spline_gf.jl:194-219, or linear with c2 = c3 = 0).
__device__ __forceinline__ double delta_lookup(const QiwDeltaTables &dt, int tab, double z) {
    double x = z / dt.h;
    int j = (int)floor(x);
    j = j < 0 ? 0 : (j > dt.n_tau - 2 ? dt.n_tau - 2 : j);
    double t = x - (double)j;
    const double *c = dt.coef + ((int64_t)tab * (dt.n_tau - 1) + j) * 4;
    return c[0] + t * (c[1] + t * (c[2] + t * c[3]));
}

// ---------------------------------------------------------------------------------------
// The diagram evaluator for models whose sectors are all one-dimensional.
// ---------------------------------------------------------------------------------------

#define LDS(i) lds[(i) * QIW_WAVE + lane]

__global__ __launch_bounds__(QIW_WAVE) void k_eval_scalar(EvalArgs a) {
    extern __shared__ double lds[];
    const int lane = threadIdx.x;
    const QiwTask task = a.tasks[blockIdx.x];
    const QiwGroup g = a.groups[task.group];
    const int ns = a.ns, n = g.n_nodes, k = g.order;
    const int ntab = a.dtab.n_tables;
    const double inv_h = 1.0 / a.h;

    // slot layout: P slots | W slots | value stack [0..n] | per-s0 outputs
    const int p_base = 0, w_base = g.n_pslots, v_base = w_base + g.n_wslots, o_base = v_base + n + 1;

    // ---- contour times of the configuration positions (1-based) ----
    double tau_i, tau_w, tau_f;
    if (g.kind == QIW_G_BARE) {
        tau_i = 0.0;
        tau_w = 0.0;
        tau_f = a.h;
    } else if (g.kind == QIW_G_BOLD) {
        tau_i = 0.0;
        tau_w = a.step_w * a.h;
        tau_f = (a.step_w + 1) * a.h;
    } else {
        tau_i = 0.0;
        tau_w = task.tau_index * a.h;
        tau_f = a.beta;
    }
    const int64_t sample = (int64_t)task.first_sample + lane;
    const bool active = sample < g.n_samples;
    double t[QIW_NMAX + 1];
    const int pos_w = (g.kind == QIW_G_BARE) ? -1 : g.pos_w;
#pragma unroll
    for (int p = 0; p <= QIW_NMAX; ++p) t[p] = (p == 1) ? tau_i : (p == pos_w) ? tau_w : (p == n) ? tau_f : 0.0;
    if (g.sampled) {
        const int64_t N = g.n_samples;
        const double *u = a.ut + g.ut_off + (active ? sample : 0);
        const double dfw = tau_f - tau_w, dwi = tau_w - tau_i;
#pragma unroll
        for (int j = 0; j < 2 * QIW_KMAX; ++j) {
            if (j < g.d_after) {
                double v = u[j * N] * dfw + tau_w;
#pragma unroll
                for (int p = 1; p <= QIW_NMAX; ++p)
                    if (p == n - 1 - j) t[p] = v;
            } else if (j < g.d_after + g.d_before) {
                double v = u[j * N] * dwi + tau_i;
                const int pos = g.pos_w - 1 - (j - g.d_after);
#pragma unroll
                for (int p = 1; p <= QIW_NMAX; ++p)
                    if (p == pos) t[p] = v;
            }
        }
    }

    // ---- propagator slots: i*P_s(t_L - t_{L-1}) for L = 2..n (ordering clamp) ----
#pragma unroll
    for (int L = 2; L <= QIW_NMAX; ++L) {
        if (L <= n) {
            double dt = t[L] - t[L - 1];
            dt = dt > 0.0 ? dt : 0.0;
            for (int s = 0; s < ns; ++s) {
                double v = (g.kind == QIW_G_BARE) ? exp(-dt * a.e_shift[s])
                                                  : ptab_lookup(a.ptab, ns, s, a.n_tau, inv_h, dt);
                LDS(p_base + (L - 2) * ns + s) = v;
            }
        }
    }
    LDS(v_base + 1) = 1.0;
    for (int s = 0; s < ns; ++s) LDS(o_base + s) = 0.0;

    // ---- walk the program ----
    double cur = 1.0, acc = 0.0;
    const uint64_t *__restrict__ prog = a.prog + g.prog_off;
    for (int pc = 0;; ++pc) {
        const uint64_t w = prog[pc];
        const int op = (int)((w >> 48) & 0xF);
        if (op == QIW_OP_EDGE) {
            const int L = (int)(w & 0xF);
            double v = (w & QIW_F_RESTORE) ? LDS(v_base + L - 1) : cur;
            v *= LDS(p_base + (int)((w >> 8) & 0xFF));
            if (w & QIW_F_TAIL) v *= LDS(w_base + (int)((w >> 16) & 0xFF));
            cur = v;
            if (w & QIW_F_SAVE) LDS(v_base + L) = v;
            if (w & QIW_F_LEAF) acc += a.coefs[g.coef_off + (int)((w >> 24) & 0xFFFFFF)] * v;
        } else if (op == QIW_OP_TOPO) {
            const uint8_t *arc = a.arcs + g.topo_off + (int)((w >> 24) & 0xFFFFFF) * 2 * k;
            for (int ai = 0; ai < k; ++ai) {
                const int head = arc[2 * ai], tail = arc[2 * ai + 1];
                double th = 0.0, tt = 0.0;
#pragma unroll
                for (int p = 1; p <= QIW_NMAX; ++p) {
                    th = (p == head) ? t[p] : th;
                    tt = (p == tail) ? t[p] : tt;
                }
                double z = tt - th;
                z = z > 0.0 ? z : 0.0;
                for (int tb = 0; tb < ntab; ++tb) LDS(w_base + ai * ntab + tb) = delta_lookup(a.dtab, tb, z);
            }
        } else if (op == QIW_OP_FLUSH) {
            LDS(o_base + (int)((w >> 8) & 0xFF)) += acc;
            acc = 0.0;
        } else {
            break;
        }
    }

    // ---- wavefront reduction of the per-s0 outputs ----
    for (int s = 0; s < ns; ++s) {
        double v = active ? LDS(o_base + s) : 0.0;
#pragma unroll
        for (int off = QIW_WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, QIW_WAVE);
        if (lane == 0) a.partial[(int64_t)task.out_index * ns + s] = v;
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

__global__ __launch_bounds__(256) void k_finalize_step(FinalizeArgs a) {
    const int tid = threadIdx.x;
    const int jf = a.jf;
    const double tau_i = 0.0, tau_w = a.tau_w, tau_f = a.tau_f;
    for (int i = tid; i < a.n_order_slots * a.bt; i += blockDim.x) a.scratch[i] = 0.0;
    __syncthreads();
    // one thread per output element, groups in order (fixed summation order)
    for (int e = tid; e < a.bt; e += blockDim.x) {
        for (int gi = 0; gi < a.n_groups; ++gi) {
            const QiwGroup g = a.groups[gi];
            double s = 0.0;
            const int f0 = a.group_task_first[gi], nt = a.group_task_count[gi];
            for (int r = 0; r < nt; ++r) s += a.partial[(int64_t)(f0 + r) * a.bt + e];
            double fac = 1.0;
            if (g.sampled) {
                double J = a.bare ? simplex_volume(g.d_after, tau_f - tau_i)
                                  : simplex_volume(g.d_before, tau_w - tau_i) * simplex_volume(g.d_after, tau_f - tau_w);
                fac = ((g.order & 1) ? -1.0 : 1.0) * J / (double)g.n_samples;
            }
            a.scratch[a.group_order_slot[gi] * a.bt + e] += fac * s;
        }
        double tot = 0.0;
        for (int o = 0; o < a.n_order_slots; ++o) {
            double v = a.scratch[o * a.bt + e];
            a.porders[((int64_t)o * a.n_tau + jf) * a.bt + e] = v;
            tot += v;
        }
        a.ptab[(int64_t)jf * a.bt + e] = tot;
    }
    if (!a.normalize) return;
    __threadfence_block();
    __syncthreads();
    __shared__ double lam_sh;
    if (tid == 0) {
        double pmax = -INFINITY;
        for (int i = 0; i < a.n_diag; ++i) pmax = fmax(pmax, a.ptab[(int64_t)jf * a.bt + a.diag_idx[i]]);
        lam_sh = log(pmax) / tau_f;
        a.lambda_out[jf] = lam_sh;
    }
    __syncthreads();
    const double lam = lam_sh;
    const int rows = a.n_known;
    for (int i = tid; i < rows * a.bt; i += blockDim.x) {
        const int j = i / a.bt;
        a.ptab[i] *= exp(-(j * a.h) * lam);
    }
}

// Correlator bookkeeping (src/inchworm.jl:825-889): per (pair, tau) row, sum the groups'
// prefactor-weighted traces.  out[c][j] = sum_g (-1)^k J_g / N_g * R_g (real; the value
// is -i * out / Z).

__global__ __launch_bounds__(256) void k_finalize_corr(CorrFinalizeArgs a) {
    int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= a.n_rows) return;
    const int j = row % a.n_tau;
    const double tau_i = 0.0, tau_w = j * a.h, tau_f = a.beta;
    double tot = 0.0;
    for (int gi = 0; gi < a.n_groups; ++gi) {
        const int nt = a.row_task_count[row * a.n_groups + gi];
        if (nt == 0) continue;
        const int f0 = a.row_task_first[row * a.n_groups + gi];
        const QiwGroup g = a.groups[gi];
        double s = 0.0;
        for (int r = 0; r < nt; ++r)
            for (int e = 0; e < a.ns; ++e) s += a.partial[(int64_t)(f0 + r) * a.ns + e];
        double fac = 1.0;
        if (g.sampled) {
            double J = simplex_volume(g.d_before, tau_w - tau_i) * simplex_volume(g.d_after, tau_f - tau_w);
            fac = ((g.order & 1) ? -1.0 : 1.0) * J / (double)g.n_samples;
        }
        tot += fac * s;
    }
    a.out[row] = tot;
}

// ---------------------------------------------------------------------------------------
// Host-callable launch wrappers (plain C++ signatures for qiw_host.cpp)
// ---------------------------------------------------------------------------------------
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

hipError_t qiw_launch_simplex_points(hipStream_t st, const uint32_t *m, const uint32_t *x0, int d_before,
                                     int d_after, double t_i, double t_w, double t_f, int64_t n, double *out) {
    int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(k_simplex_points, dim3(blocks), dim3(256), 0, st, m, x0, d_before, d_after, t_i, t_w,
                       t_f, n, out);
    return hipGetLastError();
}

hipError_t qiw_launch_eval_scalar(hipStream_t st, const EvalArgs &a, int n_tasks) {
    size_t lds = (size_t)a.lds_slots * QIW_WAVE * sizeof(double);
    hipLaunchKernelGGL(k_eval_scalar, dim3(n_tasks), dim3(QIW_WAVE), lds, st, a);
    return hipGetLastError();
}

hipError_t qiw_launch_finalize_step(hipStream_t st, const FinalizeArgs &a) {
    hipLaunchKernelGGL(k_finalize_step, dim3(1), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t qiw_launch_finalize_corr(hipStream_t st, const CorrFinalizeArgs &a) {
    int blocks = (a.n_rows + 255) / 256;
    hipLaunchKernelGGL(k_finalize_corr, dim3(blocks), dim3(256), 0, st, a);
    return hipGetLastError();
}
