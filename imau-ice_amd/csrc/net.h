// net.h -- device side of the warm-start network (net.hip): a four-level
// convolutional net over batches of images of uneven size.
//
// Activations are BF16, pixel-major with the channels innermost
// (x[pix * C + c]); an image of rows x cols pixels starts at pixel `off` of its
// level, pixel (r, c) is off + r * cols + c.  Level l+1 halves each extent
// (rounded up).  Work is split into tiles of IMAU_NET_TR x IMAU_NET_TC pixels,
// one wavefront each.  Convolution weights are [tap][cout][cin] with
// tap = (dr + 1) * 3 + (dc + 1); 3x3 convolutions read zeros outside the image.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#define IMAU_NET_LEVELS 4
#define IMAU_NET_TR 4
#define IMAU_NET_TC 16
#define IMAU_NET_FEATURES 16  // input channels (10 used, the rest zero)
#define IMAU_NET_OUT 16       // output channels of the last convolution (2 used)

namespace imau {
namespace net {

struct NImg {
  int rows, cols;
  long long off;
};
struct NTile {
  int img, r0, c0, pad;
};
struct LevelView {
  const NImg *img;
  const NTile *tiles;
  int ntiles;
  long long npix;
};

// y = act(conv3x3(x, w) + bias); act = ReLU when relu; y BF16 [pix][cout], or
// with yf32 FP32 [pix][2] (the first two channels).  (cin, cout) in
// {(16,32), (32,32), (32,16)}; false for any other pair.
bool conv(hipStream_t s, const LevelView &L, int cin, int cout, const uint16_t *x, const uint16_t *w,
          const float *bias, int relu, void *y, int yf32);
// weight and bias gradients of a convolution: dw[tap][co][ci] = sum_p g[p][co] x[p + tap][ci],
// db[co] = sum_p g[p][co]; part holds dw_blocks(L) * (9 cout cin + cout) floats
int dw_blocks(const LevelView &L);
bool conv_dw(hipStream_t s, const LevelView &L, int cin, int cout, const uint16_t *g, const uint16_t *x, float *dw,
             float *db, float *part);
// 2x2 average (zero outside the fine image) of a 32-channel fine level F into C
void pool(hipStream_t s, const LevelView &F, const LevelView &C, const uint16_t *x, uint16_t *y);
// y = a + nearest-neighbour upsampling of b (32 channels; a, y on F, b on C)
void upadd(hipStream_t s, const LevelView &F, const LevelView &C, const uint16_t *a, const uint16_t *b, uint16_t *y);
// backward of pool / skip into a ReLU output yact on F:
// g = (a + pool^T pc) * (yact > 0); a (on F) or pc (on C) may be null
void gpool(hipStream_t s, const LevelView &F, const LevelView &C, const uint16_t *a, const uint16_t *pc,
           const uint16_t *yact, uint16_t *g);
// backward of the upsampling into a ReLU output yact on C: g = up^T(a) * (yact > 0), a on F
void gup(hipStream_t s, const LevelView &F, const LevelView &C, const uint16_t *a, const uint16_t *yact, uint16_t *g);
// weighted squared error: partial sums of w (out - t)^2 per block (loss_blocks of them)
// and g = 2 scale w (out - t) as BF16 [pix][16] (channels 2.. zero)
int loss_blocks(const LevelView &L);
void loss(hipStream_t s, const LevelView &L, const float *out, const float *t, const float *w, float scale,
          uint16_t *g, double *partial);
// Adam step over n parameters (bias corrections bc1 = 1 - b1^k, bc2 = 1 - b2^k)
void adam(hipStream_t s, long long n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2,
          float eps, float bc1, float bc2);
// FP32 weights [9][cout][cin] -> BF16 forward weights (same layout) and the
// weights of the data gradient, wt[tap][ci][co] = w[8 - tap][co][ci]
void pack(hipStream_t s, int cin, int cout, const float *w, uint16_t *wf, uint16_t *wt);
void sum_partials(hipStream_t s, int nblk, int n, const float *part, float *dw, int nw, float *db);

}  // namespace net
}  // namespace imau
