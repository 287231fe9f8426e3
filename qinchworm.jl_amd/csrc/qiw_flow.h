// qiw_flow.h -- layout and launch wrappers of the learned warp of the unit cube
// (qiw_flow.hip).
//
// The warp is a normalising flow of coupling layers with monotone rational-quadratic
// splines on [0, 1] (Durkan et al., "Neural Spline Flows", NeurIPS 2019), conditioned on
// the expansion order, the integration group and the step fraction.  Layout:
//   coordinates  double [n][QIW_FLOW_D]; a sample of order k uses the first 2k
//   layer l      conditions on coordinates i with i % 2 == l % 2 (slot a <-> i = 2a + l % 2)
//                and transforms the others (slot b <-> i = 2b + 1 - l % 2)
//   features     float [n][QIW_FLOW_FEAT]: x_A (4 slots), one-hot(order - 1) (4), group
//                embedding (16), step fraction (1)
//   conditioner  tanh(f W1 + b1) -> tanh(. W2 + b2) -> . W3 + b3 -> 4 slots x (3K + 1)
//   parameters   float, per layer W1 [FEAT][H], b1 [H], W2 [H][H], b2 [H], W3 [H][4(3K+1)],
//                b3 [4(3K+1)] (row-major), then the embedding table [GROUPS][EMBED]
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#define QIW_FLOW_D 8
#define QIW_FLOW_SLOTS 4
#define QIW_FLOW_KMAX 4
#define QIW_FLOW_EMBED 16
#define QIW_FLOW_GROUPS 16
#define QIW_FLOW_FEAT (QIW_FLOW_SLOTS + QIW_FLOW_KMAX + QIW_FLOW_EMBED + 1)
#define QIW_FLOW_MAX_BINS 16
#define QIW_FLOW_MAX_HIDDEN 128
#define QIW_FLOW_MIN_W 1e-3
#define QIW_FLOW_MIN_D 1e-3

// Rows of the batched dense kernels per workgroup
#define QIW_FLOW_TILE 64
// Rows of one partial sum of the weight gradients
#define QIW_FLOW_CHUNK 512

hipError_t qiw_flow_features(hipStream_t st, int layer, const double *x, const int32_t *order,
                             const int32_t *group, const float *step, const float *emb, float *feat, int n);
hipError_t qiw_flow_dense(hipStream_t st, const float *in, const float *W, const float *b, float *out, int n,
                          int n_in, int n_out, int tanh_act);
hipError_t qiw_flow_rqs(hipStream_t st, int layer, const double *xin, double *xout, const float *raw,
                        const int32_t *order, double *logdet, int n, int bins, int inverse);
hipError_t qiw_flow_rqs_bwd(hipStream_t st, int layer, const double *xin, const float *raw, const int32_t *order,
                            const float *coef, double scale, const double *gx_out, double *gx_in, float *graw,
                            int n, int bins);
hipError_t qiw_flow_tanh_bwd(hipStream_t st, const float *dout, const float *out, float *dpre, int64_t count);
hipError_t qiw_flow_dense_bwd_data(hipStream_t st, const float *dpre, const float *W, float *din, int n, int n_in,
                                   int n_out);
hipError_t qiw_flow_dense_bwd_weight(hipStream_t st, const float *in, const float *dpre, float *partial,
                                     float *dW, float *db, int n, int n_in, int n_out);
hipError_t qiw_flow_feat_bwd(hipStream_t st, int layer, const float *dfeat, const int32_t *order,
                             const int32_t *group, double *gx, float *partial, float *demb, int n);
hipError_t qiw_flow_loss(hipStream_t st, const double *logq, const float *coef, double scale, double *loss, int n);
hipError_t qiw_flow_adam(hipStream_t st, float *p, const float *g, float *m, float *v, int64_t count, float lr,
                         float b1, float b2, float eps, float bc1, float bc2);
hipError_t qiw_flow_fill(hipStream_t st, float *p, int64_t count, float value);
hipError_t qiw_flow_fill_f64(hipStream_t st, double *p, int64_t count, double value);
