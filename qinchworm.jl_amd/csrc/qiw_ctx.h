// qiw_ctx.h -- internal state of a libqiw context, shared by the host translation units
// (qiw_host.cpp: model, programs and drivers; qiw_warp.cpp: the learned warp).
#pragma once
#include <hip/hip_runtime.h>

#include <complex>
#include <map>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/qiw.h"
#include "qiw_device.h"

namespace qiwi {

typedef std::complex<double> cd;

// Bytes of the static scratch array of qiw_inchworm_step_noalloc (one per GPU)
constexpr size_t kNoAllocScratchBytes = QIW_NOALLOC_SCRATCH_BYTES;

// ---------------------------------------------------------------------------------------
// Model in the representation the kernels use: i*P, i*Delta and the operator elements are
// real for every model the kernels accept.
// ---------------------------------------------------------------------------------------
struct ScalarOp {  // SectorBlockMatrix: s_i -> (s_f, value of a 1x1 block, block id)
    std::vector<int> sf;
    std::vector<double> val;
    std::vector<int> blk;   // index into HostModel::op_blocks
};

struct HostModel {
    double beta = 0, h = 0;
    int ntau = 0, ns = 0;
    std::vector<int> dims;
    std::vector<std::vector<double>> E;
    double lambda0 = 0;
    int ppgf_kind = QIW_GF_GRID_LINEAR;
    int np = 0, ncorr = 0;
    std::vector<ScalarOp> op_i, op_f, corr_A, corr_B;
    std::vector<int> pair_table;  // pair -> distinct hybridization table
    int ntab = 0;
    std::vector<double> dcoef;    // [ntab][ntau-1][4]
    std::vector<int32_t> dkind;
    int dlr_rank = 0;             // poles per DLR table
    std::vector<double> dlr;      // [ntab][dlr_rank] (w, coefficient of i*Delta)
    std::vector<std::vector<int>> attachable;
    int bt = 0;
    std::vector<int> bt_off;
    bool scalar = true;          // every sector is one-dimensional
    bool scalar_kernel = true;   // flat scalar programs for k_eval_flat_t (scalar model, <= 8 sectors, <= 4 tables)
    bool vec_ok = false;         // programs for k_eval_vec_t up to order 4 (blocks <= 6, <= 4 tables)
    bool mfma_ok = false;        // the same programs on k_eval_mfma_t (blocks <= 32, <= 2 tables)
    int mmax = 1;
    std::vector<double> op_data;    // all operator blocks, column-major, concatenated
    std::vector<int32_t> op_desc;   // [n_blocks][3]: offset, m_f, m_i
    // Layout of the matrix-core walk (blocks wider than 6): operator blocks row-major with
    // rows padded to a multiple of 4 floats (offsets 16-byte aligned, four zeros past the
    // last block), and the i*P blocks of one segment in LDS the same way.
    std::vector<float> ops_rm;
    std::vector<int32_t> op_rm_off;  // [n_blocks]
    std::vector<int32_t> lds_off;    // [ns] offset of each sector's block within a segment
    std::vector<int32_t> pinv;       // [btp] bt element at each LDS position, -1 for padding
    int btp = 0;
};


struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
};

struct PlanStore;  // compiled batches resident on the device, by buffer tag (qiw_host.cpp)
void plans_destroy(PlanStore *p);

struct StepGraph;  // the bold-step loop of an inchworm run captured in a hipGraph (qiw_host.cpp)
void step_graph_destroy(StepGraph *g);

struct WarpState;  // learned warp of the unit cube (qiw_warp.cpp)
void warp_destroy(WarpState *w);
bool warp_enabled(const WarpState *w);

// A bold-step group whose sample points the warp moves.
struct WarpGroup {
    int order, n_after, d_before, d_after, gid;
    int64_t off;  // ut offset (evaluation) or per-sample offset (training data)
};

}  // namespace qiwi

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
};

struct qiw_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool has_model = false;
    qiwi::HostModel m;
    std::string model_fp;      // bytes of the model content the compiled programs depend on
    bool vec_staged = false;   // operator blocks of the vector walk are on the device
    bool mfma_staged = false;  // operator blocks of the matrix-core walk are on the device
    // Sobol state
    int sob_dim = 0;
    std::vector<uint32_t> sob_m, sob_x;
    // caller-drawn sequences (qiw_set_scrambles), on the device as buffers given.m / given.x
    int64_t given_n = 0;
    int given_stride = 0;
    int64_t points_n = 0, points_npts = 0;  // caller-supplied point sequences (qiw_set_points)
    int points_stride = 0;
    // propagator state (host mirror of the device table), i*P in BT layout
    std::vector<double> ip;  // [ntau][bt]
    int n_known = 1;
    std::vector<qiwi::cd> der0;      // incremental-spline derivative dP/dtau at 0 (bookkeeping of normalize!)
    std::vector<double> lambdas;
    // device buffers
    std::map<std::string, qiwi::DevBuf> bufs;
    qiw_stats stats{};
    qiwi::WarpState *warp = nullptr;  // learned warp, or null
    qiwi::PlanStore *plans = nullptr;  // compiled batches kept between calls
    qiw_tuning tune{};                 // work mapping (qiw_set_tuning / qiw_autotune); 0 = default
    bool graphs = true;                // replay the bold-step loop from a hipGraph (qiw_set_graphs)
    qiwi::StepGraph *graph = nullptr;  // the captured loop and the launch arguments it holds
    std::vector<hipEvent_t> evpool;    // timing events of the eager launches, reused
    // qiw_inchworm_step_noalloc: this device's static scratch array (its device address),
    // the bytes handed out during the current call, and the flag that makes dev() serve
    // the per-call buffers from it and allocate nothing
    void *scratch = nullptr;
    size_t scratch_used = 0;
    bool no_alloc = false;
};

namespace qiwi {

// Records msg as the context's last error and returns code.
int fail(qiw_ctx *c, int code, const std::string &msg);

// Runs the body of a C entry point: a C++ exception (host memory exhausted, an index out of
// range) becomes a QIW_E_* code and a message instead of crossing the C ABI.
template <class F>
auto guarded(qiw_ctx *c, F &&body) -> decltype(body()) {
    try {
        return body();
    } catch (const std::bad_alloc &) {
        return c ? fail(c, QIW_E_LIMIT, "host memory exhausted") : QIW_E_LIMIT;
    } catch (const std::exception &e) {
        return c ? fail(c, QIW_E_ARG, std::string("internal error: ") + e.what()) : QIW_E_ARG;
    } catch (...) {
        return c ? fail(c, QIW_E_ARG, "internal error") : QIW_E_ARG;
    }
}

// Device buffer `name` of at least `bytes` bytes, kept by the context between calls.
void *dev(qiw_ctx *c, const std::string &name, size_t bytes, int *rc);

// Learned warp in the drivers (qiw_warp.cpp): unit-cube points and conditions of the
// weighted groups of a run (n samples each), then per bold step their warped simplex
// coordinates and weights, written into the groups' rows of ut.
int warp_prepare(qiw_ctx *c, const std::vector<WarpGroup> &gs, int64_t n);
int warp_points(qiw_ctx *c, double step_frac, double *ut);

// Training data (qiw_host.cpp): sum of |integrand| per sample of every bold group of run
// (order <= 4) at each step in `steps`, plain qMC with n points, current expansion.P.
// mag [steps][groups][n]; gs[i].off is unused.
int warp_collect(qiw_ctx *c, const qiw_run *run, const std::vector<int> &steps, int64_t n,
                 std::vector<WarpGroup> &gs, std::vector<double> &mag);

}  // namespace qiwi

#define HIPCHK(ctx, expr)                                                                  \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(ctx, QIW_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)
