/* imau_net.h -- C ABI of the learned warm start of the DIVA solve in
 * libimauvel.so.
 *
 * No counterpart in the reference: IMAU-ICE starts every Picard loop from the
 * velocities of its previous call (src/ice_velocity_module.f90:249-374 reads
 * ice%u_vav_cx / v_vav_cy as the first iterate).  For independent solves this
 * library can supply that first iterate from a small convolutional network
 * instead: four resolution levels, 3x3 convolutions with 32 channels and ReLU,
 * 2x2 average pooling, nearest-neighbour upsampling with skip addition, BF16
 * weights and activations with FP32 accumulation on the matrix cores.  The
 * network reads ten per-cell features of the geometry and friction and
 * predicts the a-grid velocities; imau_net_warm_start maps them onto the
 * staggered grids, and the solve then runs with warm_visc_iters > 0.
 *
 * Tensors are device pointers.  An image of rows x cols pixels with C channels
 * is stored pixel-major with the channels innermost; the images of a level are
 * packed one after the other (imau_net_plan_pixels gives the level's total).
 * BF16 values are passed as uint16_t.  Convolution weights are FP32 or BF16
 * [tap][cout][cin], tap = (dr + 1) * 3 + (dc + 1); pixels outside an image
 * read as zero.  Every call is asynchronous on the context's stream;
 * imau_net_sync waits for it.  All return IMAU_OK or a negative status
 * (imau_vel_last_error has the text).
 */
#ifndef IMAU_NET_H
#define IMAU_NET_H

#include <stdint.h>

#include "imau_vel.h"

#ifdef __cplusplus
extern "C" {
#endif

#define IMAU_NET_NLEVELS 4
#define IMAU_NET_NFEATURES 16 /* input channels (10 used, 6 zero)       */
#define IMAU_NET_NOUT 16      /* channels of the last layer (2 used)     */
#define IMAU_NET_WIDTH 32     /* channels of every hidden layer          */

typedef struct imau_net_plan imau_net_plan;

/* layout of nimg images (rows[g] x cols[g] at level 0) at every level */
int imau_net_plan_create(imau_ctx *ctx, int32_t nimg, const int32_t *rows, const int32_t *cols,
                         imau_net_plan **plan);
/* one image per member of the problem: its a-grid, rows = nx, cols = ny */
int imau_net_plan_for_problem(imau_ctx *ctx, const imau_problem *P, imau_net_plan **plan);
void imau_net_plan_destroy(imau_net_plan *plan);
int64_t imau_net_plan_pixels(const imau_net_plan *plan, int32_t level);
/* sizes of the work buffers of imau_net_conv_grad_w (floats) and imau_net_loss (doubles) */
int64_t imau_net_grad_w_work(const imau_net_plan *plan, int32_t level, int32_t cin, int32_t cout);
int64_t imau_net_loss_work(const imau_net_plan *plan);

/* y = act(conv3x3(x, w) + bias) at `level`; act = ReLU if relu.  y is BF16
 * [pix][cout], or with y_f32 FP32 [pix][2] (channels 0 and 1).  (cin, cout) in
 * {(16,32), (32,32), (32,16)}; bias may be NULL (zero). */
int imau_net_conv(imau_ctx *ctx, const imau_net_plan *plan, int32_t level, int32_t cin, int32_t cout,
                  const uint16_t *x, const uint16_t *w, const float *bias, int32_t relu, void *y, int32_t y_f32);
/* dw[tap][co][ci] = sum_p g[p][co] x[p + tap][ci], db[co] = sum_p g[p][co] (FP32) */
int imau_net_conv_grad_w(imau_ctx *ctx, const imau_net_plan *plan, int32_t level, int32_t cin, int32_t cout,
                         const uint16_t *g, const uint16_t *x, float *dw, float *db, float *work);
/* level -> level+1: 2x2 average, zero outside the image (32 channels) */
int imau_net_pool(imau_ctx *ctx, const imau_net_plan *plan, int32_t level, const uint16_t *x, uint16_t *y);
/* y = a + up(b): a, y at level, b at level+1 (32 channels) */
int imau_net_upadd(imau_ctx *ctx, const imau_net_plan *plan, int32_t level, const uint16_t *a, const uint16_t *b,
                   uint16_t *y);
/* g = (a + pool^T(pc)) * (yact > 0) at level; a (level) or pc (level+1) may be NULL */
int imau_net_grad_pool(imau_ctx *ctx, const imau_net_plan *plan, int32_t level, const uint16_t *a,
                       const uint16_t *pc, const uint16_t *yact, uint16_t *g);
/* g = up^T(a) * (yact > 0): a at level, yact and g at level+1 */
int imau_net_grad_up(imau_ctx *ctx, const imau_net_plan *plan, int32_t level, const uint16_t *a,
                     const uint16_t *yact, uint16_t *g);
/* loss = scale * sum_p w[p] |out[p] - t[p]|^2 over level 0 (out, t FP32 [pix][2]);
 * g = d loss / d out as BF16 [pix][16] (channels 2..15 zero); *loss written
 * after the stream drains */
int imau_net_loss(imau_ctx *ctx, const imau_net_plan *plan, const float *out, const float *target,
                  const float *weight, float scale, uint16_t *g, double *work, double *loss);
/* Adam step `step` (1-based) over n FP32 parameters */
int imau_net_adam(imau_ctx *ctx, int64_t n, float *p, const float *g, float *m, float *v, float lr, float beta1,
                  float beta2, float eps, int32_t step);
/* FP32 weights [9][cout][cin] -> BF16 wf (same layout) and wt[tap][ci][co] =
 * w[8 - tap][co][ci], the weights of the data gradient (a conv with cin and cout swapped) */
int imau_net_pack(imau_ctx *ctx, int32_t cin, int32_t cout, const float *w, uint16_t *wf, uint16_t *wt);

/* network input of every member, BF16 [pix][16], from the problem's uploaded
 * geometry, friction and rate factor (plan from imau_net_plan_for_problem) */
int imau_net_features(imau_ctx *ctx, imau_problem *P, const imau_cfg *cfg, const imau_net_plan *plan,
                      uint16_t *x);
/* training targets from the problem's current (solved) velocities: FP32
 * [pix][2] asinh(u_a / 10 m/yr), asinh(v_a / 10 m/yr) and weights [pix] (1 on ice) */
int imau_net_targets(imau_ctx *ctx, imau_problem *P, const imau_net_plan *plan, float *target, float *weight);
/* overwrite the velocities of every member with grounded ice by the network
 * output (FP32 [pix][2], same scale as the targets) */
int imau_net_warm_start(imau_ctx *ctx, imau_problem *P, const imau_cfg *cfg, const imau_net_plan *plan,
                        const float *out);
int imau_net_sync(imau_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif /* IMAU_NET_H */
