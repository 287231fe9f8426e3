// qiw_device.h -- data structures shared by the host driver (qiw_host.cpp) and the gfx950
// kernels (qiw_kernels.hip) of libqiw.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#define QIW_KMAX 6                      // largest expansion order of a compiled program
#define QIW_NMAX (2 * QIW_KMAX + 3)     // largest number of configuration positions
#define QIW_WAVE 64

// ---- program words ----------------------------------------------------------------------
// Block programs (k_eval_block) and vector programs carry 64-bit words; all formats use
// bits 48-51 for the op: 0 EDGE, 1 TOPO, 2 FLUSH, 3 END, 4 NOP, 5 ROOT, 6 SEG.
//
// Vector programs (k_eval_vec_t, blocks m_s <= MB): one segment per initial sector s0,
// SEG, then TOPO ROOT EDGE... per topology, then FLUSH (replayed for every column of s0):
// SEG    bits  0-5  : m_s0, bits 8-15 : s0, bits 52-63 : BT offset of the s0 block
// ROOT   bits  0-5  : dimension of the sector after position 1, bits 24-31 : children,
//        bits 32-47 : operator block B (offset into the operator data; 0xFFFF identity),
//        bits 52-63 : BT offset of that sector's block
// EDGE   bits  0-5  : dimension m_o of the sector after the node's operator,
//        bit 6 LEAF (coefficient word follows unless bit 9 UNIT: +1, or -1 with bit 8 NEG),
//        bit 7 TAIL (W slot bits 16-23), bits 24-31 : children,
//        bits 32-47 : operator block (offset into the operator data; 0xFFFF identity),
//        bits 52-63 : BT offset of the sector's block
// FLUSH  the column of s0 is complete
enum { QIW_OP_EDGE = 0, QIW_OP_TOPO = 1, QIW_OP_FLUSH = 2, QIW_OP_END = 3, QIW_OP_NOP = 4, QIW_OP_ROOT = 5,
       QIW_OP_SEG = 6 };
#define QIW_F_LEAF (1ull << 6)
#define QIW_F_TAIL (1ull << 7)
// Leaf coefficients of magnitude 1 carry no coefficient word: UNIT, with NEG for -1
// (vector programs: bits 8, 9)
#define QIW_FV_NEG (1ull << 8)
#define QIW_FV_UNIT (1ull << 9)

// Table stride of the W slots (slot = arc * stride + table) of the register walks: the
// table count TM of the kernel instantiation that runs the program.  The host compiler and
// the launchers choose it by the same rule, from the largest order km of the launch, the
// sector count sm and the model's table count tm.  The LDS block walk uses tm itself.
static inline int qiw_scalar_tm(int km, int sm, int tm) { return (km <= 4 && sm <= 4 && tm <= 2) ? 2 : 4; }
static inline int qiw_vec_tm(int tm) { return tm <= 2 ? 2 : 4; }

// Flat scalar programs (k_eval_flat_t, all sectors 1x1): 32-bit words, one per trie node in
// preorder, the leaf (the last position, always an only child) folded into its parent.  A
// running product flows down each chain; a node with several children saves it in the slot
// of its children's level (PUSH) and every child but the first restores it (POP).
// EDGE   (a node at level L >= 2, its configuration position)
//        bits  0-7  : twice the inner propagator entry multiplied after the node, entry
//                     (NM - 1 - (L + 1)) * SM + sector after the node's operator (NM = 2 KM + 3,
//                     SM: sector stride; qiw_scalar_km / _sm), or for LEAF without UNIT the
//                     index of the coefficient in the group's table
//        bits  8-15 : W slot of the arc value (TAIL)
//        bits 16-19 : level L
//        bit  20 TAIL : multiply the arc value
//        bit  21 LEAF : the only child is the last position: acc += coefficient * product
//        bit  22 PUSH : after the propagator, save the product for level L + 1
//        bit  23 POP  : first restore the product saved for level L (level 2: the root's)
//        bit  24 UNIT : coefficient +1, or -1 with bit 25 NEG
//        bits 28-31 : op, 0 for EDGE
// ROOT   op 4, position 1: product = iP_2[s], s in bits 0-7 (kept for the level-2 children);
//        with LEAF (two positions only): acc += coefficient (bits 0-7, UNIT, NEG as for EDGE)
// TOPO   op 1, bits 0-23 : topology index into the group's table of arc head/tail nibbles
//        (each arc end as its simplex coordinate j: position n - 1 - j for j < d_after,
//        pos_w - 1 - (j - d_after) for the coordinates before tau_w)
// FLUSH  op 2, bits 0-7 : s0 (out[s0] += acc * iP_n[s0]; acc = 0)
// END    op 3
// NEXT   op 5, the last word of every 64-word block: continue with the next block
#define QIW_FL_TAIL (1u << 20)
#define QIW_FL_LEAF (1u << 21)
#define QIW_FL_PUSH (1u << 22)
#define QIW_FL_POP (1u << 23)
#define QIW_FL_UNIT (1u << 24)
#define QIW_FL_NEG (1u << 25)
#define QIW_FL_OP (1u << 28)
enum { QIW_FL_OP_TOPO = 1, QIW_FL_OP_FLUSH = 2, QIW_FL_OP_END = 3, QIW_FL_OP_ROOT = 4, QIW_FL_OP_NEXT = 5 };

// Sector stride SM and largest order KM of the flat scalar walk instantiation (see
// qiw_scalar_tm); its inner propagator entries are (2*KM + 2 - L) * SM + s for level L,
// sector s.
static inline int qiw_scalar_sm(int km, int sm, int tm) { return qiw_scalar_tm(km, sm, tm) == 2 ? 4 : 8; }
static inline int qiw_scalar_km(int km, int sm, int tm) { return (km <= 4 && sm <= 8 && tm <= 4) ? 4 : QIW_KMAX; }

// Kinds of integration group (below).  A group is exact (no qMC sum, no standard
// deviation) at order 0 and for the order-1 self-energy, which has no variable times.
// SIGMA: a self-energy diagram of diff_inchworm_step! (src/inchworm.jl:557-601) laid out
// as a bold group whose 2k pair positions all lie before tau_w = tau_f_prev, with tau_f =
// tau_w: the first and last pair positions are pinned to tau_f_prev and tau_i by their
// coordinates (1 and 0), so the three identity segments are P(0) = -i and the chain is the
// self-energy's.
enum { QIW_G_BOLD = 0, QIW_G_BARE = 1, QIW_G_CORR = 2, QIW_G_SIGMA = 3 };

// A float32 cell of a grid-container table: value, forward and backward difference
// (12 bytes, loaded with one instruction).
struct QiwCell32 {
    float v, f, b;
};

// One (order, n_pts_after) group of diagrams, compiled.
struct QiwGroup {
    int32_t kind;          // QIW_G_*
    int32_t order;         // k
    int32_t n_after;       // points after tau_w (bold/corr); 2k for bare
    int32_t n_nodes;       // configuration positions
    int32_t d_before, d_after;
    int32_t pos_w;         // position of the inch / A operator node (0 for bare)
    int32_t n_pslots, n_wslots;
    int32_t prog_off;      // offset into the program word array
    int32_t topo_off;      // offset into the topology table: [n_top][2] words of head/tail nibbles
    int32_t coef_off;      // offset into the leaf coefficient table (scalar programs)
    int64_t ut_off;        // offset (in doubles) of this group's simplex coordinates
    int32_t n_samples;     // qMC points (1 for order 0: exact evaluation)
    int32_t sampled;       // 0 for order 0
    int32_t weighted;      // row d_before + d_after of the coordinates holds per-sample weights
    int32_t seq;           // randomized qMC: index of the group's sequence (0 otherwise)
    int64_t ps_off;        // offset of the group's samples in EvalArgs::per_sample
};

__host__ __device__ static inline bool qiw_group_exact(const QiwGroup &g) {
    return !g.sampled || (g.kind == QIW_G_SIGMA && g.order == 1);
}

// A unit of work of one launch: up to 64 samples of one group at one time configuration,
// over one chunk of the group's topologies (flat scalar programs; the whole group otherwise).
struct QiwTask {
    int32_t group;
    int32_t first_sample;
    int32_t tau_index;     // bold: n_w (tau_w = grid[n_w]); corr: index of tau_A; bare: 0
    int32_t out_index;     // row of the per-task partial-sum output
    int32_t chunk;         // offset of the chunk's program from the group's, in 64-bit words
    int32_t pad_;
};

// Interpolation tables of the hybridization functions, all stored as i*Delta (real).
struct QiwDeltaTables {
    int32_t n_tables;
    int32_t n_tau;
    double h, beta;
    int32_t grid_mask;     // bit t set: table t is a grid container (else a natural cubic spline)
    int32_t dlr_mask;      // bit t set: table t is a DLR pole sum (dlr below)
    int32_t dlr_rank;      // poles per DLR table
    const double *dlr;     // [n_tables][dlr_rank] (frequency w, coefficient of i*Delta)
    const int32_t *kind;   // [n_tables] QIW_GF_GRID_LINEAR / QIW_GF_CUBIC_SPLINE / QIW_GF_DLR
    // [n_tables][n_tau-1][4]: a spline table holds the cubic (c0..c3) of interval j in the
    // time difference; a grid table holds, for index d, (g(d), g(d+1) - g(d), g(d) - g(d-1), 0)
    // with g(-1) = 2 g(0) - g(1) (the grid cell interpolation of the kernels' grid_cell)
    const double *coef;
};

// Arguments of the diagram evaluators (k_eval_flat_t; the other walks embed them)
struct EvalArgs {
    const QiwGroup *groups;
    const QiwTask *tasks;
    const uint64_t *prog;
    const uint32_t *topo;      // per topology two words of arc head/tail nibbles
    const double *coefs;
    const double *ut;          // simplex coordinates of all groups
    int64_t ut_row_stride;     // coordinates of task row tau_index at ut + tau_index * stride (0: shared)
    const double *ptab;        // [n_tau][ns] i*P on the grid (bold groups)
    const double *e_shift;     // [ns] E_s + lambda (bare groups, exact atomic P0)
    QiwDeltaTables dtab;
    int ns;
    int n_tau;
    double h, beta;
    int lds_slots;             // LDS doubles per wavefront
    int step_w;                // tau_w index of a bold step
    int max_order;             // largest order of the launch's groups
    double *partial;           // [n_out][ns]
    double *per_sample;        // optional: sum of |integrand| over the outputs, per sample
    int spt;                   // samples per task (64 for the lane-per-sample walks)
};

// Arguments of the step reduction and bookkeeping kernel (k_step_reduce_finalize)
struct FinalizeArgs {
    const QiwGroup *groups;
    const int32_t *group_task_first;   // [n_groups] first out row
    const int32_t *group_task_count;   // [n_groups]
    const int32_t *group_order_slot;   // [n_groups] slot in the order-resolved output
    int n_groups;
    int n_order_slots;
    const double *partial;             // [rows][bt]
    int bt;                            // outputs per row (sum m^2)
    const int32_t *diag_idx;           // [n_diag] BT indices of diagonal elements
    int n_diag;
    double *ptab;                      // [n_tau][bt] i*P (differential step: the self-energy table)
    int jt;                            // row of ptab receiving the sum over orders (jf for inchworm steps)
    double *porders;                   // [n_order_slots][n_tau][bt] raw i*P contributions
    double *lambda_out;                // [n_tau] per-step normalization constants
    int n_tau;
    int n_known;                       // rows rescaled by normalize! (n_tau for grid P)
    double h;
    int jf;                            // row receiving the step result
    double tau_w, tau_f;               // step times (tau_i = 0)
    int bare;                          // RootTransform Jacobian (initial step)
    int normalize;                     // apply normalize!(P, tau_f)
    double *scratch;                   // [n_order_slots][bt]
    double *gsum;                      // [n_groups][bt] per-group sums of the partial rows
    unsigned *done;                    // blocks of the step launch that have finished (0 between launches)
    // randomized qMC (k_step_reduce_finalize_rqmc): the groups above are the sequences of
    // n_real (order, n_pts_after) groups, those of real group r at real_first[r] ..
    // real_first[r] + real_count[r] - 1 (one entry when every sequence is the same one)
    int n_real;
    const int32_t *real_first;
    const int32_t *real_count;
    int n_seqs;                        // N_seqs
    double target_std;
    double *porders_std;               // [n_order_slots][n_tau][bt] standard deviations
    int32_t *nuse_out;                 // optional [n_tau][n_real]: sequences used, row jf
    // differential step (k_diff_reduce_finalize): ptab is the table of sigma (Σ = i sigma,
    // rows by time difference), porders its order-resolved rows; the Euler step writes row
    // a + 1 of P from rows 0..a
    double *P;                         // [n_tau][bt] i*P
    const double *hdiag;               // [bt] E + lambda on the diagonal positions
    const int32_t *dims, *bt_off;      // [ns]
    int ns;
    int a;                             // index of tau_f_prev
};

struct CorrFinalizeArgs {
    const QiwGroup *groups;
    const int32_t *row_task_first;   // [n_rows * n_groups]
    const int32_t *row_task_count;
    int n_groups;
    int n_rows;                      // n_corr * n_tau
    int n_tau;
    const double *partial;           // [tasks][bt] per-task block sums
    int bt;
    const int32_t *diag_idx;         // BT indices of the diagonal elements (the trace)
    int n_diag;
    double h, beta;
    double *out;                     // [n_rows]
    // randomized qMC (k_finalize_corr_rqmc): real groups as in FinalizeArgs
    int n_real;
    const int32_t *real_first;
    const int32_t *real_count;
    int n_seqs;
    double target_std;
    double *out_std;                 // [n_rows]
    int32_t *nuse_out;               // optional [n_rows][n_real]: sequences used
};

// Arguments of the block-sector evaluator (k_eval_block)
struct BlockArgs {
    EvalArgs e;
    const double *op_data;     // operator blocks, column-major
    const int32_t *op_desc;    // [n_blocks][3] offset, m_f, m_i
    const int32_t *dims;       // [ns]
    const int32_t *bt_off;     // [ns]
    int bt, mmax;
};

// Arguments of the vector evaluator (k_eval_vec_t)
struct VecArgs {
    EvalArgs e;
    const double2 *pt;         // [n_tau-1][bt] (value, forward difference) of the i*P table (bold groups)
    const double *ptm;         // [n_tau-1][bt] backward differences (row 0: the forward one)
    const QiwCell32 *pt32;     // (value, forward, backward) in float32
    const double *ediag;       // [bt] E + lambda on the diagonal positions (bare groups)
    const double *ops;         // operator blocks, column-major, padded by 64 entries
    const float *ops32;        // the same in float32
    int bt;
};

// Arguments of the matrix-core evaluator (k_eval_mfma_t)
struct MfmaArgs {
    EvalArgs e;
    const float *ops;          // operator blocks, row-major with rows padded to 4 floats, float32
    const double *ediag;       // [bt] E + lambda on the diagonal positions (bare groups)
    const uint8_t *emask;      // [bt] 1 on the diagonal positions
    const int32_t *pinv;       // [btp] element of the bt layout at each LDS position of a
                               // segment's i*P blocks (row-major, padded rows), -1 for padding
    const int32_t *bto;        // [ns] offset of each sector's block in the bt layout
    int bt, btp;
    int ops_zero;              // position of four 0.0 in ops (16-byte aligned, past the blocks)
    int wave_lds;              // floats of LDS per wavefront (absent-row writes and walk stack)
};
// Wavefronts per workgroup of k_eval_mfma_t (tasks come in units of this many)
#define QIW_MFMA_WAVES 4

// ---- randomized qMC scrambles -------------------------------------------------------------
#define QIW_MAX_SEQS 64                 // largest N_seqs
// SplitMix64 output function after the increment (Steele, Lea, Flood, OOPSLA 2014).
__host__ __device__ static inline uint64_t qiw_mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// Stream of sequence `seq` of group (order, n_after): domain 1 initial step (key 0),
// 2 bold step (key = tau_w index), 3 correlator (key = operator pair), 4 differential step
// (key = tau_f_prev index, n_after = 1).
enum { QIW_RQ_BARE = 1, QIW_RQ_BOLD = 2, QIW_RQ_CORR = 3, QIW_RQ_DIFF = 4 };
__host__ __device__ static inline uint64_t qiw_rq_stream(uint64_t domain, uint64_t key, int order, int n_after,
                                                         int seq) {
    return (domain << 60) | ((key & 0xFFFFFFFull) << 32) | ((uint64_t)order << 24) | ((uint64_t)n_after << 16) |
           (uint64_t)seq;
}
// Bit word w of dimension d of a stream: 0 the shift, 1 + p row p of the scramble matrix.
__host__ __device__ static inline uint32_t qiw_rq_word(uint64_t seed, uint64_t stream, int d, int w) {
    return (uint32_t)qiw_mix64(qiw_mix64(seed ^ qiw_mix64(stream)) + (uint64_t)(33 * d + w));
}

// Launch wrappers (qiw_kernels.hip)
hipError_t qiw_launch_zero(hipStream_t st, void *p, size_t bytes);  // bytes: a multiple of 4
hipError_t qiw_launch_sobol_points(hipStream_t st, const uint32_t *m, const uint32_t *x0, int dim, int64_t first,
                                   int64_t n, double *out);
hipError_t qiw_launch_simplex_coords(hipStream_t st, const uint32_t *m, const uint32_t *x0, int d_before,
                                     int d_after, int64_t n, double *out);
hipError_t qiw_launch_sobol_cube(hipStream_t st, const uint32_t *m, const uint32_t *x0, int dim, int stride,
                                 int64_t n, double *out);
hipError_t qiw_launch_warp_coords(hipStream_t st, const double *z, int stride, const double *logj, int d_before,
                                  int d_after, int64_t n, double *out);
hipError_t qiw_launch_simplex_points(hipStream_t st, const uint32_t *m, const uint32_t *x0, int d_before,
                                     int d_after, double t_i, double t_w, double t_f, int64_t n, double *out);
hipError_t qiw_launch_eval_flat(hipStream_t st, const EvalArgs &a, int n_tasks, int km, int sm, int tm);
hipError_t qiw_launch_eval_block(hipStream_t st, const BlockArgs &b, int n_tasks, size_t lds_bytes);
hipError_t qiw_launch_eval_mfma(hipStream_t st, const MfmaArgs &m, int n_tasks, int km, int tm, size_t lds_bytes);
hipError_t qiw_launch_eval_vec(hipStream_t st, const VecArgs &v, int n_tasks, int mb, int km, int tm, int fp32);
hipError_t qiw_launch_ptab_cells(hipStream_t st, const double *T, int n_tau, int bt, double2 *out, double *outm,
                                 QiwCell32 *out32);
hipError_t qiw_launch_finalize_step(hipStream_t st, const FinalizeArgs &a);
hipError_t qiw_launch_finalize_corr(hipStream_t st, const CorrFinalizeArgs &a);
hipError_t qiw_launch_finalize_step_rqmc(hipStream_t st, const FinalizeArgs &a);
hipError_t qiw_launch_finalize_diff(hipStream_t st, const FinalizeArgs &a, int rqmc);
// Coordinates of a self-energy group: rows 0 and 2k-1 are 1 and 0 (tau_f_prev, tau_i),
// rows 1..2k-2 the RootTransform of Sobol dimensions 0..2k-3.
hipError_t qiw_launch_sigma_coords(hipStream_t st, const uint32_t *m, const uint32_t *x0, int order, int64_t n,
                                   double *out);
hipError_t qiw_launch_finalize_corr_rqmc(hipStream_t st, const CorrFinalizeArgs &a);
// Simplex coordinates of every sampled group's sequence for one step: the staged
// direction numbers m (dimensions 0..dim_max-1) LMS + shift scrambled with the bits of the
// group's stream (domain, key, order, n_after, seq).
// The caller's sequences of a call: the states of caller-drawn scrambled Sobol sequences
// (gm: [.][stride][32] direction numbers, gx: [.][stride] shifts), or, with pts set, the
// points of caller-supplied sequences (pts: [.][npts][stride]).
struct GivenSrc {
    const uint32_t *gm = nullptr, *gx = nullptr;
    const double *pts = nullptr;
    int stride = 0;
    int64_t npts = 0;
};
// coordinates from the caller's sequences: row r = 0..n_rows-1 of group g (at ut + (row0 +
// r) * row_stride) maps sequence idx[r * n_groups + g] (< 0: none)
hipError_t qiw_launch_given_coords(hipStream_t st, const QiwGroup *groups, int n_groups, int64_t n, const GivenSrc &src,
                                   const int64_t *idx, int n_rows, int row0, int64_t row_stride, double *ut);
hipError_t qiw_launch_rqmc_coords(hipStream_t st, const QiwGroup *groups, int n_groups, int max_dim, int64_t n,
                                  const uint32_t *m, uint64_t seed, int domain, int64_t key, double *ut);
