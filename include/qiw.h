/*
 * qiw.h -- C ABI of libqiw, the MI355X (gfx950) quasi Monte Carlo diagram-evaluation
 * engine behind the inchworm solver's `inchworm!` and `correlator_2p`.
 *
 * Every entry point takes plain pointers and sizes.  Complex numbers are interleaved
 * (re, im) float64 pairs, i.e. C99 `double _Complex` / Julia `ComplexF64`.
 * Sector indices are 0-based.  Matrices are column-major.  The library copies every
 * input during the call; outputs are written into caller-owned arrays before the call
 * returns (all entry points are synchronous).  No entry point throws; each returns
 * QIW_OK (0) or a negative QIW_E* code and leaves a message for qiw_last_error().
 * One context per host thread; a context is not re-entrant.
 *
 * Block-table layout ("BT layout") used for every block-diagonal quantity
 * (P(tau), step results, per-order contributions): sector s of dimension m_s occupies
 * the m_s*m_s complex slots starting at off_s = sum_{s'<s} m_{s'}^2, element (r, c) at
 * off_s + r + c*m_s.  A table over the imaginary-time grid stores n_tau such blocks of
 * sum_s m_s^2 complex numbers, tau index slowest.
 */
#ifndef QIW_H
#define QIW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QIW_ABI_VERSION 6

/* ---- return codes (every reference @assert on this path maps to one) ---------------- */
enum {
    QIW_OK = 0,
    QIW_E_ARG = -1,          /* malformed argument (null pointer, bad size)              */
    QIW_E_NOT_POW2 = -2,     /* N_samples is neither 0 nor a power of 2
                                (src/inchworm.jl:343, :930)                              */
    QIW_E_OFFDIAG = -3,      /* a diagram leaf ends in a sector other than the one it
                                started in (src/topology_eval.jl:462-464)                */
    QIW_E_NOT_SBM = -4,      /* operator has more than one block per column
                                (src/sector_block_matrix.jl:57-61)                       */
    QIW_E_DIM = -5,          /* Sobol dimension beyond the supplied table
                                (src/scrambled_sobol.jl:67)                              */
    QIW_E_STATE = -6,        /* call out of order (e.g. no model set)                   */
    QIW_E_DEVICE = -7,       /* HIP runtime failure                                      */
    QIW_E_LIMIT = -8,        /* size beyond a compiled-in limit of the GPU kernels      */
    QIW_E_NO_GPU = -9        /* no usable gfx950 device                                  */
};

/* ---- interpolation kinds of time-dependent quantities ------------------------------- */
enum {
    QIW_GF_GRID_LINEAR = 0,  /* Keldysh ImaginaryTimeGF stored on the grid, evaluated at
                                (t1, t2) as Keldysh's interpolate! does: bilinear over the
                                grid cell of (t1, t2) with G[i1, i2] = g(i1 - i2), linear
                                in t1 - t2 inside a diagonal cell                        */
    QIW_GF_CUBIC_SPLINE = 1, /* SplineInterpolatedGF: natural cubic spline through the
                                grid values (src/spline_gf.jl:194-242)                   */
    QIW_GF_INC_SPLINE = 2,   /* IncSplineImaginaryTimeGF: incremental quadratic spline
                                (src/spline_gf.jl:273-425, src/utility.jl:106-157)      */
    QIW_GF_DLR = 3           /* keldysh_dlr.DLRImaginaryTimeGF (src/keldysh_dlr.jl:156-
                                190): an exact pole sum over its DLR frequencies (pair
                                propagators only; qiw_model.pair_dlr_*)                  */
};

/* ---- arithmetic of the chain products ------------------------------------------------ */
enum {
    QIW_PREC_FP64 = 0,       /* all arithmetic in float64                                */
    QIW_PREC_FP32 = 1        /* block chain products in float32; Sobol-to-simplex map,
                                accumulation (compensated) and P bookkeeping in float64  */
};

/* One block-sparse operator, the reference's SectorBlockMatrix
 * (src/sector_block_matrix.jl:44): for every block, column sector s_i maps to row
 * sector s_f through an m_f x m_i matrix.  At most one block per column sector. */
typedef struct {
    int32_t n_blocks;
    const int32_t *s_i;      /* [n_blocks] column (initial) sector                       */
    const int32_t *s_f;      /* [n_blocks] row (final) sector                            */
    const double *data;      /* complex128, blocks concatenated in order, each
                                column-major m_f x m_i                                   */
} qiw_sbm;

/* The expansion problem, the reference's `Expansion` (src/expansion.jl:107-197) after
 * the caller has solved the local problem.  The imaginary-time grid is uniform,
 * tau_j = j * beta / (n_tau - 1), j = 0..n_tau-1 (kd.ImaginaryTimeGrid). */
typedef struct {
    double beta;
    int32_t n_tau;
    int32_t n_sectors;
    const int32_t *sector_dim;     /* [n_sectors] block sizes m_s                         */
    const double *energies;        /* [sum m_s] eigenvalues of H_loc per sector (float64),
                                      sector-major; defines the atomic propagators        */
    int32_t ppgf_kind;             /* QIW_GF_GRID_LINEAR (interpolate_ppgf=false) or
                                      QIW_GF_INC_SPLINE (interpolate_ppgf=true)           */
    int32_t n_pairs;               /* InteractionPairs (src/expansion.jl:79-90)           */
    const qiw_sbm *pair_op_i;      /* [n_pairs] operator_i (earlier, "head" end)          */
    const qiw_sbm *pair_op_f;      /* [n_pairs] operator_f (later, "tail" end)            */
    const double *pair_prop;       /* [n_pairs][n_tau] complex128: propagator values on
                                      the grid, g_p(tau_j)                                */
    const int32_t *pair_prop_kind; /* [n_pairs] QIW_GF_GRID_LINEAR, QIW_GF_CUBIC_SPLINE or
                                      QIW_GF_DLR                                           */
    const int32_t *pair_prop_xi;   /* [n_pairs] statistics sign (-1 fermionic, +1 bosonic),
                                      used when the time difference is negative           */
    int32_t n_corr;                /* correlator operator pairs (A, B), add_corr_operators! */
    const qiw_sbm *corr_A;         /* [n_corr] A, placed at tau                            */
    const qiw_sbm *corr_B;         /* [n_corr] B, placed at 0                              */
    /* Pair propagators of kind QIW_GF_DLR: g_p(tau) = sum_k c_pk K(tau, w_pk) for
     * 0 <= tau <= beta (xi_p g_p(beta + tau) for tau < 0, src/keldysh_dlr.jl:185-200),
     * K(tau, w) = e^{-w tau} / (1 + e^{-w beta}) the fermionic kernel of the DLR
     * coefficients that Lehmann.jl's dlr2tau sums (its Spectral.kernelFermiT); c_pk are
     * the GF's coefficients (G.mat.data), w_pk its grid's DLR frequencies (G.grid.dlr.ω).
     * The grid values pair_prop are not read for these pairs. */
    int32_t dlr_rank;              /* frequencies per DLR propagator, padded with zero
                                      coefficients (0: no DLR pair)                       */
    const double *pair_dlr_omega;  /* [n_pairs][dlr_rank] w_pk (NULL when dlr_rank = 0)    */
    const double *pair_dlr_coef;   /* [n_pairs][dlr_rank] complex128 c_pk                  */
} qiw_model;

/* Randomized qMC, RandomizationParams (src/randomization.jl:46-57) consumed by
 * mean_std_from_randomization (:86-100): every sampled (order, n_pts_after) group of a
 * step (a tau row of a correlator) is integrated with up to N_seqs sequences, stopping
 * after sequence s >= 2 once the largest element of the standard deviation over the
 * sequences so far is <= target_std; the group contributes the mean and (corrected)
 * standard deviation over the sequences used.
 *   scramble = 0  rng = nothing: every sequence is the staged one (qiw_set_sobol), so
 *                 the mean is its integral and the deviation 0 from the second on.
 *   scramble = 1  sequence s of a group is the unscrambled Sobol sequence after an
 *                 LMS + digital-shift scramble (src/scrambled_sobol.jl:102-140) whose
 *                 bits are drawn from a counter-based generator keyed by seed and
 *                 (step, order, n_pts_after, s), not from one RNG stream in evaluation
 *                 order, so that all sequences of a step are evaluated in one launch:
 *                 bit word w of dimension d is the low 32 bits of
 *                   mix(mix(seed ^ mix(stream)) + 33 d + w),
 *                 mix = SplitMix64's output function applied after adding
 *                 0x9E3779B97F4A7C15, stream = (domain << 60) | (key << 32) |
 *                 (order << 24) | (n_pts_after << 16) | s, domain 1 for the initial
 *                 step (key 0), 2 for the bold step at tau_w = grid[key], 3 for the
 *                 correlator of pair `key` (one set of sequences for all its tau rows),
 *                 4 for the differential step at tau_f_prev = grid[key] (n_pts_after 1).
 *                 Word 0 is the shift x (bit j = shift bit j); word 1 + p holds row p
 *                 of the lower-triangular matrix: its bits 31..32-p are the entries
 *                 (p, 0..p-1), the diagonal is 1.
 *   scramble = 2  rng given, target_std = 0: the sequences are the caller's own, staged
 *                 with qiw_set_scrambles -- the ScrambledSobolSeq(2k; scramble_rng = rng)
 *                 objects the reference constructs (src/randomization.jl:94), drawn by the
 *                 caller in the reference's order, so the points are the reference's bit for
 *                 bit whatever the RNG.  Every call consumes them from the first on, N_seqs
 *                 per sampled group (order > 0, N_samples > 0; the exact order-1
 *                 self-energy group of diff_inchworm! draws none), groups in the reference's
 *                 order (orders as given, n_pts_after ascending, groups without topologies
 *                 skipped):
 *                   qiw_inchworm      the initial step's groups (orders_bare), then every
 *                                     bold step's (orders), step by step;
 *                   qiw_inchworm_step(_bare)   the one step's groups;
 *                   qiw_correlator_2p(_std)    per operator pair, per tau index 1..n_tau-1,
 *                                     the groups (tau = 0 is exact);
 *                   qiw_diff_inchworm every step's self-energy groups, step by step.
 *                 A group whose integral vanishes identically takes min(2, N_seqs), where
 *                 the reference's stopping rule ends (qiw_scramble_dims lists the exact
 *                 dimensions).  A call with too few staged sequences, or with fewer
 *                 dimensions than 2k, returns QIW_E_ARG before it evaluates anything.
 *   scramble = 3  rng given, target_std = 0, any sequence type: the sequence plug point
 *                 seq_type of mean_std_from_randomization (src/randomization.jl:86-94,
 *                 src/inchworm.jl:129,233,339; e.g. RandomSeq, src/utility.jl:336-414).
 *                 The caller constructs seq_type(D; scramble_rng = rng) in the order of
 *                 scramble = 2, draws N_samples points of each with next!, and stages them
 *                 with qiw_set_points; they are consumed as under scramble = 2.
 * The GPU evaluates all N_seqs sequences and applies the stopping rule afterwards; the
 * result is that of the reference's sequential loop over the same sequences. */
typedef struct {
    int32_t N_seqs;                /* >= 1                                                */
    int32_t scramble;              /* 0, 1 or 2 (above)                                   */
    uint64_t seed;
    double target_std;
} qiw_rand;

/* Parameters of one inchworm!/correlator_2p run (keyword arguments of
 * src/inchworm.jl:332-339 and :1075-1081). */
typedef struct {
    int32_t n_orders;
    const int32_t *orders;         /* orders of a regular (bold) step                     */
    int32_t n_orders_bare;
    const int32_t *orders_bare;    /* orders of the initial (bare) step                   */
    int64_t N_samples;             /* qMC points per (order, n_pts_after) group, 0 or 2^n */
    int32_t n_pts_after_max;       /* <= 0 means unrestricted                             */
    int32_t precision;             /* QIW_PREC_*                                          */
    const qiw_rand *rand_params;   /* NULL: RandomizationParams() (N_seqs = 1, unscrambled) */
} qiw_run;

typedef struct qiw_ctx qiw_ctx;

/* ---- context --------------------------------------------------------------------------
 * Replaces the per-rank state of the MPI setup (src/mpi.jl:39-54): one context drives
 * one GPU. */
int qiw_create(qiw_ctx **out, int device);
void qiw_destroy(qiw_ctx *ctx);
const char *qiw_last_error(const qiw_ctx *ctx);
int qiw_abi_version(void);

/* ---- inputs ----------------------------------------------------------------------------
 * qiw_set_model: stages an Expansion (src/expansion.jl:147-197).  Resets the bold
 * propagator to the atomic one (ppgf.atomic_ppgf, src/ppgf.jl:144-207; for
 * QIW_GF_INC_SPLINE also initial_ppgf_derivative, src/ppgf.jl:710-720). */
int qiw_set_model(qiw_ctx *ctx, const qiw_model *model);

/* qiw_set_sobol: direction numbers of the (optionally scrambled) Sobol sequence, the
 * state of ScrambledSobolSeq (src/scrambled_sobol.jl:50-143) after its constructor:
 * m[d*32 + j] is m[d+1, j+1] of the reference (already multiplied by 2^(31-j) and, when
 * scrambled, already LMS-scrambled), shift[d] is the initial point x (NULL = zeros,
 * the unscrambled sequence).  Points of an order-k group use dimensions 0..2k-1. */
int qiw_set_sobol(qiw_ctx *ctx, int32_t dim, const uint32_t *m, const uint32_t *shift);

/* qiw_set_scrambles: the caller-drawn sequences of qiw_rand.scramble = 2, replacing the
 * ScrambledSobolSeq(D; scramble_rng = rng) constructions of mean_std_from_randomization
 * (src/randomization.jl:94, src/scrambled_sobol.jl:66-143): sequence i has the direction
 * numbers m[i][d][0..31] and the initial point shift[i][d] (NULL: zeros) of dimensions
 * d = 0 .. stride-1 (1 <= stride <= 12; a group of order k reads dimensions 0 .. 2k-1 of
 * its sequences).  n = 0 clears them.  They stay staged for later calls. */
int qiw_set_scrambles(qiw_ctx *ctx, int64_t n, int32_t stride, const uint32_t *m, const uint32_t *shift);

/* qiw_set_points: the caller-supplied sequences of qiw_rand.scramble = 3, replacing the
 * seq_type(D; scramble_rng = rng) constructions of mean_std_from_randomization
 * (src/randomization.jl:94) and the next! calls of contour_integral
 * (src/qmc_integrate.jl:597-612): point r of sequence i is x[(i * n_points + r) * stride
 * + d], d = 0 .. stride-1 (1 <= stride <= 12; a group of order k reads dimensions
 * 0 .. 2k-1, a call with N_samples reads points 0 .. N_samples-1, n_points >= N_samples),
 * every coordinate in [0, 1].  n = 0 clears them.  They stay staged for later calls. */
int qiw_set_points(qiw_ctx *ctx, int64_t n, int64_t n_points, int32_t stride, const double *x);

/* qiw_scramble_dims: the dimensions D = 2k of the sequences a call with `run` consumes
 * under scramble = 2 or 3, in the order it consumes them -- the order in which the reference
 * constructs ScrambledSobolSeq(D; scramble_rng = rng) objects, so a caller draws exactly
 * these from its RNG and stages them with qiw_set_scrambles.  N_seqs per sampled group,
 * except min(2, N_seqs) for a group whose integral vanishes identically (no contributing
 * diagram, or a simplex of zero volume: the correlator row tau_A = beta, the first step of
 * diff_inchworm!), where the reference's stopping rule ends after the second sequence.  A
 * call whose groups stop elsewhere (values identical across sequences for another reason)
 * returns QIW_E_ARG after evaluating.  Writes up to cap entries and returns the count, or
 * a negative error code.  No counterpart in the reference (its loops draw as they go). */
enum {
    QIW_CALL_INCHWORM = 0,       /* qiw_inchworm                                          */
    QIW_CALL_STEP = 1,           /* qiw_inchworm_step                                     */
    QIW_CALL_STEP_BARE = 2,      /* qiw_inchworm_step_bare                                */
    QIW_CALL_CORRELATOR = 3,     /* qiw_correlator_2p, qiw_correlator_2p_std              */
    QIW_CALL_DIFF = 4            /* qiw_diff_inchworm                                     */
};
int64_t qiw_scramble_dims(qiw_ctx *ctx, const qiw_run *run, int32_t call, int64_t cap, int32_t *dims);

/* qiw_set_ppgf: overwrite the bold propagator state (expansion.P) with a table in BT
 * layout over the grid.  n_known grid points are valid.  der0 (BT layout, one block) is
 * the tau-derivative at 0 of the incremental spline (QIW_GF_INC_SPLINE only, else NULL). */
int qiw_set_ppgf(qiw_ctx *ctx, int32_t n_known, const double *P, const double *der0);
/* qiw_get_ppgf: read expansion.P back (same layout); n_known_out may be NULL. */
int qiw_get_ppgf(qiw_ctx *ctx, int32_t *n_known_out, double *P, double *der0);

/* ---- drivers ---------------------------------------------------------------------------
 * qiw_inchworm: inchworm!(expansion, grid, orders, orders_bare, N_samples;
 * n_pts_after_max) (src/inchworm.jl:332-498).  The bare step and all n_tau-2 bold steps
 * run on the device.  Outputs (caller-owned, may be NULL):
 *   P_out            [n_tau][sum m^2]   expansion.P after the run
 *   P_orders_out     [n_orders_all][n_tau][sum m^2]  per-order contributions, orders_all
 *                    sorted ascending (union of orders and orders_bare)
 *   P_orders_std_out same shape: the standard deviation over the sequences, summed over
 *                    the order's groups (real part); NaN for sampled orders with
 *                    N_seqs = 1 (std of one sample), 0 for order 0 and for rows no step
 *                    wrote. */
int qiw_inchworm(qiw_ctx *ctx, const qiw_run *run, double *P_out, double *P_orders_out,
                 double *P_orders_std_out);

/* qiw_inchworm_step: one regular step inchworm_step(expansion, c, tau_i, tau_w, tau_f,
 * top_data) (src/inchworm.jl:123-203) with tau_i = grid[0], tau_w = grid[n_w],
 * tau_f = grid[n_w+1], on the current expansion.P; does not modify it.  Outputs
 * value [sum m^2] and per-order [n_orders][sum m^2] (orders as in run->orders). */
int qiw_inchworm_step(qiw_ctx *ctx, const qiw_run *run, int32_t n_w, double *value_out,
                      double *orders_out);

/* qiw_inchworm_step_noalloc: no counterpart in the reference.  For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): qiw_inchworm_step with the same arguments and results that calls no
 * hipMalloc/hipFree.  It needs what a previous qiw_inchworm_step with the same run, model
 * structure and Sobol tables left resident on the context (compiled programs, simplex
 * coordinates, model tables; QIW_E_STATE otherwise) and takes its per-call buffers
 * (per-task partial sums, the propagator table, the bookkeeping sums and counter) from a
 * static device array of QIW_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends
 * (QIW_E_LIMIT for a step that needs more).  Every call on one GPU shares that array and
 * clears what it needs cleared: calls that overlap on one GPU (several contexts or host
 * threads) race.  The learned warp is not available here (QIW_E_ARG). */
#define QIW_NOALLOC_SCRATCH_BYTES (64u << 20)
int qiw_inchworm_step_noalloc(qiw_ctx *ctx, const qiw_run *run, int32_t n_w, double *value_out,
                              double *orders_out);

/* qiw_inchworm_step_bare: the initial step inchworm_step_bare (src/inchworm.jl:228-304)
 * between grid[0] and grid[1], orders from run->orders_bare. */
int qiw_inchworm_step_bare(qiw_ctx *ctx, const qiw_run *run, double *value_out,
                           double *orders_out);

/* qiw_correlator_2p: correlator_2p(expansion, grid, orders, N_samples)
 * (src/inchworm.jl:918-1052, :1075-1084) for every (A, B) pair of the model, using the
 * current expansion.P.  corr_out [n_corr][n_tau] complex128: <A(tau_j) B(0)> / Z. */
int qiw_correlator_2p(qiw_ctx *ctx, const qiw_run *run, double *corr_out);

/* qiw_correlator_2p_std: correlator_2p(expansion, grid, orders, N_samples, RequestStdDev();
 * rand_params) (src/inchworm.jl:918-1052): corr_out as above and corr_std_out
 * [n_corr][n_tau] complex128, the sum over the groups of each row's standard deviation
 * over the sequences, divided by Z (NaN with N_seqs = 1, 0 at tau = 0). */
int qiw_correlator_2p_std(qiw_ctx *ctx, const qiw_run *run, double *corr_out, double *corr_std_out);

/* qiw_diff_inchworm: diff_inchworm!(expansion, grid, orders, N_samples; rand_params)
 * (src/inchworm.jl:666-778), the differential formulation: n_tau - 1 steps from the staged
 * atomic propagator, each evaluating the self-energy diagrams of run->orders (order 1
 * exact, higher orders by qMC over the 2k - 2 inner times, src/inchworm.jl:528-601) at
 * (tau_f_prev, tau_i), then P(tau_f) = P(tau_f_prev) + h (int Σ P - H P(tau_f_prev)) with
 * the trapezoid rule and H = E + lambda (:603-635).  run->orders_bare is ignored.
 * Outputs (caller-owned, may be NULL):
 *   P_out                [n_tau][sum m^2]   expansion.P after the run
 *   Sigma_orders_out     [n_orders][n_tau][sum m^2] per-order self-energy, orders as in
 *                        run->orders, stored at tau_f of each step as the reference does
 *   Sigma_orders_std_out same shape: standard deviations (real part); NaN for orders >= 2
 *                        with N_seqs = 1, 0 for order 1.
 * The self-energy container follows the propagator's kind: with QIW_GF_INC_SPLINE it is
 * incremental and step a's value lands one grid row later, as in the reference. */
int qiw_diff_inchworm(qiw_ctx *ctx, const qiw_run *run, double *P_out, double *Sigma_orders_out,
                      double *Sigma_orders_std_out);

/* ---- observables (src/ppgf.jl:611-636, :729-733) --------------------------------------
 * partition_function(P) = sum_s tr iP_s(beta); density_matrix after normalize!(P, beta):
 * rho_out [sum m^2] complex (BT layout).  Neither modifies the context. */
int qiw_partition_function(qiw_ctx *ctx, double *Z_out /* complex128 */);
int qiw_density_matrix(qiw_ctx *ctx, double *rho_out);

/* ---- component entry points, exposed for parity testing ------------------------------
 * qiw_sobol_points: points first..first+n-1 (0-based) of the staged sequence in
 * dimensions 0..dim-1, computed on the device; out [n][dim] float64
 * (ScrambledSobolSeq next!, src/scrambled_sobol.jl:158-173). */
int qiw_sobol_points(qiw_ctx *ctx, int32_t dim, int64_t first, int64_t n, double *out);

/* qiw_simplex_points: device Sobol points pushed through DoubleSimplexRootTransform
 * (src/qmc_integrate.jl:398-449) for d_before/d_after on [t_i, t_w, t_f]; with
 * d_after = 0 the RootTransform (src/qmc_integrate.jl:220-245) on [t_i, t_f].
 * out [n][d_before + d_after] float64 times, descending within a point. */
int qiw_simplex_points(qiw_ctx *ctx, int32_t d_before, int32_t d_after, double t_i,
                       double t_w, double t_f, int64_t n, double *out);

/* ---- measurement ---------------------------------------------------------------------
 * Statistics of the last qiw_inchworm / qiw_correlator_2p call on this context. */
typedef struct {
    double wall_s;              /* host wall time of the call                             */
    double kernel_ms;           /* summed device time of the evaluator launches (HIP events
                                   on the launch stream around each launch); a bold-step
                                   loop replayed from its hipGraph counts as a whole, its
                                   bookkeeping kernels included (qiw_set_graphs)          */
    int64_t n_kernel_launches;  /* evaluator launches                                     */
    int64_t diagram_evals;      /* (topology, sample) pairs evaluated                     */
    int64_t leaves;             /* (topology, sample, initial sector, pair assignment)
                                   leaves evaluated                                       */
    double flops;               /* algorithmic real FLOPs of the chain products           */
    int64_t graph_replays;      /* bold-step loops replayed from a hipGraph               */
} qiw_stats;
int qiw_get_stats(qiw_ctx *ctx, qiw_stats *out);

/* ---- work mapping and its tuner ----------------------------------------------------------
 * No counterpart in the reference (its work units are fixed by MPI rank ranges).  How
 * the evaluator's work is cut into workgroups:
 *   chunk_words       flat scalar programs (1x1 sectors): program words per topology chunk;
 *                     each (chunk, 64-sample block) is one workgroup.  0: 896
 *                     (640 for the self-energy batches of qiw_diff_inchworm).
 *   chunk_cost        vector programs (blocks <= 32): program words x replayed columns per
 *                     topology chunk.  0: 2^18.
 *   samples_per_task  walks that evaluate one sample per wavefront at a time (matrix-core
 *                     and LDS block walks): samples per workgroup, a power of 2 <= 64.
 *                     0: the largest that gives >= 4096 workgroups per launch.
 * The settings change only how work is split and summed (results agree to rounding).
 * qiw_set_tuning fixes them; qiw_autotune times candidate settings on one bold step of
 * `run` (tau_w = grid[n_tau/2], the current propagator) and keeps the fastest.  Both last
 * until the next qiw_set_model with a different model structure. */
typedef struct {
    int32_t chunk_words;
    int32_t samples_per_task;
    int64_t chunk_cost;
    int32_t n_tried;    /* out (qiw_autotune): candidate settings timed              */
    int32_t pad_;
    double best_ms;     /* out (qiw_autotune): evaluator launch time of the kept one */
} qiw_tuning;
int qiw_set_tuning(qiw_ctx *ctx, const qiw_tuning *t);
/* qiw_set_graphs: no counterpart in the reference.  With on = 1 (the default) the
 * bold-step loop of qiw_inchworm / qiw_inchworm_step on the flat scalar walk (all sectors
 * 1x1, no learned warp) is captured once in a hipGraph and replayed while every launch
 * argument is the same, so a repeated run issues one graph launch instead of two kernel
 * launches per step; results are identical.  on = 0 launches every kernel from the host. */
int qiw_set_graphs(qiw_ctx *ctx, int32_t on);
int qiw_get_tuning(qiw_ctx *ctx, qiw_tuning *out);
int qiw_autotune(qiw_ctx *ctx, const qiw_run *run, qiw_tuning *out);

/* ---- learned warp of the unit cube ----------------------------------------------------
 * No counterpart in the reference, which integrates with plain qMC
 * (src/qmc_integrate.jl:497-507).  A normalising flow of coupling layers with monotone
 * rational-quadratic splines on [0,1] (Durkan et al., NeurIPS 2019), conditioned on the
 * order, the (order, n_pts_after) group and the step fraction tau_w / beta, maps each
 * Sobol point x of a bold-step group to z = G(x); the group's integral becomes
 * mean_i f(z_i) |dG/dx|(x_i), which is unbiased for any parameters.  Training fits the
 * density of the warped points to |f| on samples of the plain method.
 *
 * Coordinates: double [n][8], a point of order k uses the first 2k.  Group index of the
 * bold group (k, n_after): (k-1)^2 + n_after - 1 (k <= 4).  Parameters: float32, per
 * layer W1 [25][H], b1 [H], W2 [H][H], b2 [H], W3 [H][4(3K+1)], b3 [4(3K+1)] row-major,
 * then the group embedding table [16][16]. */
typedef struct {
    int32_t n_layers;          /* coupling layers (4)                                     */
    int32_t n_bins;            /* spline bins K (8), <= 16                                */
    int32_t hidden;            /* conditioner width H (64), <= 128                        */
    int32_t iters;             /* Adam iterations of qiw_warp_train                       */
    int32_t batch;             /* samples per iteration                                   */
    int32_t n_train_steps;     /* bold steps whose integrands form the training set       */
    int64_t n_train;           /* plain qMC points per (training step, group), 2^n        */
    double lr;                 /* Adam step size                                          */
    uint64_t seed;             /* parameter initialisation and minibatch order            */
} qiw_warp_config;

/* qiw_warp_init: (re)initialise the flow: Glorot-uniform hidden layers, zero output layer,
 * so the untrained warp is the identity map. */
int qiw_warp_init(qiw_ctx *ctx, const qiw_warp_config *cfg);

/* qiw_warp_params: number of parameters (n_out, may be NULL), then copy them out (get,
 * may be NULL) and/or replace them (set, may be NULL; applied after the copy). */
int qiw_warp_params(qiw_ctx *ctx, int64_t *n_out, float *get, const float *set);

/* qiw_warp_apply: inverse = 0: the training direction z -> x with log q(z)
 * (logdet = log |dx/dz|); inverse = 1: sampling x -> z = G(x) with log |dz/dx|. */
int qiw_warp_apply(qiw_ctx *ctx, int32_t inverse, int64_t n, const double *pts, const int32_t *order,
                   const int32_t *group, const double *step, double *out, double *logdet);

/* qiw_warp_loss_grad: loss = -mean_i coef_i log q(z_i) and its gradient [n_params]. */
int qiw_warp_loss_grad(qiw_ctx *ctx, int64_t n, const double *z, const int32_t *order,
                       const int32_t *group, const double *step, const double *coef, double *loss,
                       float *grad);

/* qiw_warp_train: collect |integrand| of every bold group of run at n_train_steps steps
 * spread over the grid (plain qMC, the current expansion.P, e.g. after qiw_inchworm),
 * then run cfg->iters Adam iterations.  loss_history [iters] may be NULL. */
int qiw_warp_train(qiw_ctx *ctx, const qiw_run *run, double *loss_history);

/* qiw_warp_enable: when on, qiw_inchworm and qiw_inchworm_step use warped points for
 * bold groups of order <= 4 (bare step and correlators stay plain). */
int qiw_warp_enable(qiw_ctx *ctx, int32_t on);

#ifdef __cplusplus
}
#endif

#endif /* QIW_H */
