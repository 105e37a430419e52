// net.hip -- kernels of the warm-start network: 3x3 convolutions on the
// matrix cores (v_mfma_f32_16x16x32_bf16 for 32 input channels,
// v_mfma_f32_16x16x16_bf16 for 16: BF16 operands, FP32 accumulation),
// their weight gradients, 2x2 pooling, upsampling with skip addition, the
// backward passes of those, a weighted squared-error loss and Adam.
//
// Convolution as a matrix product per tile: D[cout][pixel] = W[cout][9 cin] X[9 cin][pixel],
// one wavefront per tile of IMAU_NET_TR x IMAU_NET_TC pixels (IMAU_NET_TR
// column blocks of 16 pixels), the layer's weights held in registers as the A
// operands, the input pixels (channels contiguous) loaded straight from HBM as
// the B operands.  Operand layout of the 16x16x16 instruction, lane l:
// A[i][k] with i = l % 16, k = 4 (l / 16) + e;  B[k][j] with j = l % 16,
// k = 4 (l / 16) + e;  D[i][j] with j = l % 16, i = 4 (l / 16) + e  (e = 0..3).
#include <algorithm>

#include "net.h"

namespace {

using imau::net::LevelView;
using imau::net::NImg;
using imau::net::NTile;

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int TR = IMAU_NET_TR, TC = IMAU_NET_TC;

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// round to nearest, ties to even (NaN kept quiet)
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ s4 ld4(const uint16_t *p) { return *reinterpret_cast<const s4 *>(p); }
__device__ __forceinline__ void st4(uint16_t *p, s4 v) { *reinterpret_cast<s4 *>(p) = v; }

// ---------------------------------------------------------------------------
// convolution
// ---------------------------------------------------------------------------
template <int CIN, int COUT, bool OUTF>
__global__ void __launch_bounds__(256) k_conv(const NImg *__restrict__ img, const NTile *__restrict__ tiles,
                                              int ntiles, const uint16_t *__restrict__ x,
                                              const uint16_t *__restrict__ w, const float *__restrict__ bias,
                                              int relu, void *__restrict__ yv) {
  constexpr int NCH = CIN / 16, NCT = COUT / 16;
  const int lane = threadIdx.x & 63, lc = lane & 15, lk = lane >> 4;
  const int wid = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (wid >= ntiles) return;  // uniform per wavefront
  s4 wa[9][NCH][NCT];
#pragma unroll
  for (int tap = 0; tap < 9; ++tap)
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch)
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct)
        wa[tap][ch][ct] = ld4(w + ((size_t)(tap * COUT + ct * 16 + lc) * CIN + ch * 16 + 4 * lk));
  float bs[NCT][4];
#pragma unroll
  for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
    for (int e = 0; e < 4; ++e) bs[ct][e] = bias ? bias[ct * 16 + 4 * lk + e] : 0.0f;

  for (int tix = wid; tix < ntiles; tix += gridDim.x * 4) {
    const NTile t = tiles[tix];
    const NImg im = img[t.img];
    f4 acc[TR][NCT];
#pragma unroll
    for (int r = 0; r < TR; ++r)
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) acc[r][ct] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      const int dr = tap / 3 - 1, dc = tap % 3 - 1;
      const int c = t.c0 + lc + dc;
      const bool cok = c >= 0 && c < im.cols;
#pragma unroll
      for (int r = 0; r < TR; ++r) {
        const int rr = t.r0 + r + dr;
        const bool ok = cok && rr >= 0 && rr < im.rows;
        const uint16_t *px = x + (im.off + (long long)rr * im.cols + c) * CIN + 4 * lk;
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
          s4 b = s4{0, 0, 0, 0};
          if (ok) b = ld4(px + ch * 16);
#pragma unroll
          for (int ct = 0; ct < NCT; ++ct)
            acc[r][ct] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wa[tap][ch][ct], b, acc[r][ct], 0, 0, 0);
        }
      }
    }
    const int c = t.c0 + lc;
#pragma unroll
    for (int r = 0; r < TR; ++r) {
      const int rr = t.r0 + r;
      if (rr >= im.rows || c >= im.cols) continue;
      const long long pix = im.off + (long long)rr * im.cols + c;
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) {
        float v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          v[e] = acc[r][ct][e] + bs[ct][e];
          if (relu) v[e] = fmaxf(v[e], 0.0f);
        }
        if constexpr (OUTF) {
          if (ct == 0 && lk == 0) {
            float *y = static_cast<float *>(yv);
            y[2 * pix] = v[0];
            y[2 * pix + 1] = v[1];
          }
        } else {
          s4 o;
#pragma unroll
          for (int e = 0; e < 4; ++e) o[e] = (short)f2bf(v[e]);
          st4(static_cast<uint16_t *>(yv) + pix * COUT + ct * 16 + 4 * lk, o);
        }
      }
    }
  }
}

// The same convolution for 32 input channels on v_mfma_f32_16x16x32_bf16: one
// instruction takes all 32 channels of a tap (lane l holds k = 8 (l / 16) + e,
// e = 0..7, of its row / column), with 16-byte operand loads.
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ b8 ld8(const uint16_t *p) { return *reinterpret_cast<const b8 *>(p); }

template <int COUT, bool OUTF>
__global__ void __launch_bounds__(256) k_conv32(const NImg *__restrict__ img, const NTile *__restrict__ tiles,
                                                int ntiles, const uint16_t *__restrict__ x,
                                                const uint16_t *__restrict__ w, const float *__restrict__ bias,
                                                int relu, void *__restrict__ yv) {
  constexpr int CIN = 32, NCT = COUT / 16;
  // each wavefront stages its tile plus a one-pixel halo ((TR+2) x (TC+2)
  // pixels, 32 channels, rows padded to 40 channels against bank conflicts)
  // in LDS once, instead of nine overlapping global loads per tap
  constexpr int XR = TR + 2, XC = TC + 2, CP = 40, WSZ = XR * XC * CP;
  __shared__ __align__(16) uint16_t sx[4 * WSZ];
  const int lane = threadIdx.x & 63, lc = lane & 15, lk = lane >> 4, wv = threadIdx.x >> 6;
  uint16_t *my = sx + wv * WSZ;
  b8 wa[9][NCT];
#pragma unroll
  for (int tap = 0; tap < 9; ++tap)
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) wa[tap][ct] = ld8(w + ((size_t)(tap * COUT + ct * 16 + lc) * CIN + 8 * lk));
  float bs[NCT][4];
#pragma unroll
  for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
    for (int e = 0; e < 4; ++e) bs[ct][e] = bias ? bias[ct * 16 + 4 * lk + e] : 0.0f;

  // the same number of rounds for every wavefront of the block (barriers)
  for (int base = blockIdx.x * 4; base < ntiles; base += gridDim.x * 4) {
    const int tix = base + wv;
    const bool have = tix < ntiles;
    NTile t = NTile{0, 0, 0, 0};
    NImg im = NImg{0, 0, 0};
    if (have) {
      t = tiles[tix];
      im = img[t.img];
    }
    if (have)
      for (int q = lane; q < XR * XC * 4; q += 64) {
        const int px = q >> 2, ch = q & 3;
        const int hr = px / XC, hc = px - hr * XC;
        const int rr = t.r0 + hr - 1, cc = t.c0 + hc - 1;
        b8 v = {};
        if (rr >= 0 && rr < im.rows && cc >= 0 && cc < im.cols)
          v = ld8(x + (im.off + (long long)rr * im.cols + cc) * CIN + 8 * ch);
        *reinterpret_cast<b8 *>(my + px * CP + 8 * ch) = v;
      }
    __syncthreads();
    if (have) {
      f4 acc[TR][NCT];
#pragma unroll
      for (int r = 0; r < TR; ++r)
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) acc[r][ct] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int tap = 0; tap < 9; ++tap) {
        const int dr = tap / 3 - 1, dc = tap % 3 - 1;
#pragma unroll
        for (int r = 0; r < TR; ++r) {
          const b8 b = *reinterpret_cast<const b8 *>(my + ((r + 1 + dr) * XC + lc + 1 + dc) * CP + 8 * lk);
#pragma unroll
          for (int ct = 0; ct < NCT; ++ct)
            acc[r][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[tap][ct], b, acc[r][ct], 0, 0, 0);
        }
      }
      const int c = t.c0 + lc;
#pragma unroll
      for (int r = 0; r < TR; ++r) {
        const int rr = t.r0 + r;
        if (rr >= im.rows || c >= im.cols) continue;
        const long long pix = im.off + (long long)rr * im.cols + c;
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) {
          float v[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            v[e] = acc[r][ct][e] + bs[ct][e];
            if (relu) v[e] = fmaxf(v[e], 0.0f);
          }
          if constexpr (OUTF) {
            if (ct == 0 && lk == 0) {
              float *y = static_cast<float *>(yv);
              y[2 * pix] = v[0];
              y[2 * pix + 1] = v[1];
            }
          } else {
            s4 o;
#pragma unroll
            for (int e = 0; e < 4; ++e) o[e] = (short)f2bf(v[e]);
            st4(static_cast<uint16_t *>(yv) + pix * COUT + ct * 16 + 4 * lk, o);
          }
        }
      }
    }
    __syncthreads();
  }
}

// weight and bias gradients: per tile, the gradient g (64 pixels x COUT) and
// the input x with a one-pixel halo ((TR+2) x (TC+2) pixels x CIN) are staged in
// LDS; for each tap D[co][ci] += G^T[co][pixels] X_tap[pixels][ci] with K = 16
// pixels (one tile row) per instruction.  Each workgroup (one wavefront) sums
// its tiles in registers and writes one partial; sum_partials adds them in a
// fixed order.
template <int CIN, int COUT>
__global__ void __launch_bounds__(64) k_conv_dw(const NImg *__restrict__ img, const NTile *__restrict__ tiles,
                                                int ntiles, const uint16_t *__restrict__ g,
                                                const uint16_t *__restrict__ x, float *__restrict__ part) {
  constexpr int NCI = CIN / 16, NCO = COUT / 16;
  constexpr int XR = TR + 2, XC = TC + 2;
  constexpr int GP = COUT + 2, XP = CIN + 2;  // padded rows (LDS banks)
  __shared__ uint16_t sg[TR * TC * GP];
  __shared__ uint16_t sx[XR * XC * XP];
  const int lane = threadIdx.x, lc = lane & 15, lk = lane >> 4;
  f4 acc[9][NCO][NCI];
#pragma unroll
  for (int tap = 0; tap < 9; ++tap)
#pragma unroll
    for (int a = 0; a < NCO; ++a)
#pragma unroll
      for (int b = 0; b < NCI; ++b) acc[tap][a][b] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  float bacc = 0.0f;
  for (int tix = blockIdx.x; tix < ntiles; tix += gridDim.x) {
    const NTile t = tiles[tix];
    const NImg im = img[t.img];
    {
      const int rr = t.r0 + (lane >> 4), cc = t.c0 + (lane & 15);
      const bool ok = rr < im.rows && cc < im.cols;
      const uint16_t *src = g + (im.off + (long long)rr * im.cols + cc) * COUT;
#pragma unroll
      for (int k = 0; k < COUT; k += 4) {
        s4 v = s4{0, 0, 0, 0};
        if (ok) v = ld4(src + k);
#pragma unroll
        for (int e = 0; e < 4; ++e) sg[lane * GP + k + e] = (uint16_t)v[e];
      }
    }
    for (int q = lane; q < XR * XC; q += 64) {
      const int hr = q / XC, hc = q - hr * XC;
      const int rr = t.r0 + hr - 1, cc = t.c0 + hc - 1;
      const bool ok = rr >= 0 && rr < im.rows && cc >= 0 && cc < im.cols;
      const uint16_t *src = x + (im.off + (long long)rr * im.cols + cc) * CIN;
#pragma unroll
      for (int k = 0; k < CIN; k += 4) {
        s4 v = s4{0, 0, 0, 0};
        if (ok) v = ld4(src + k);
#pragma unroll
        for (int e = 0; e < 4; ++e) sx[q * XP + k + e] = (uint16_t)v[e];
      }
    }
    __syncthreads();
    if (lane < COUT)
      for (int p = 0; p < TR * TC; ++p) bacc += bf2f(sg[p * GP + lane]);
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      const int dr = tap / 3 - 1, dc = tap % 3 - 1;
#pragma unroll
      for (int r = 0; r < TR; ++r) {
        s4 a[NCO], b[NCI];
#pragma unroll
        for (int co = 0; co < NCO; ++co)
#pragma unroll
          for (int e = 0; e < 4; ++e) a[co][e] = (short)sg[(r * TC + 4 * lk + e) * GP + co * 16 + lc];
#pragma unroll
        for (int ci = 0; ci < NCI; ++ci)
#pragma unroll
          for (int e = 0; e < 4; ++e)
            b[ci][e] = (short)sx[((r + 1 + dr) * XC + 4 * lk + e + 1 + dc) * XP + ci * 16 + lc];
#pragma unroll
        for (int co = 0; co < NCO; ++co)
#pragma unroll
          for (int ci = 0; ci < NCI; ++ci)
            acc[tap][co][ci] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[co], b[ci], acc[tap][co][ci], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  constexpr int NW = 9 * COUT * CIN;
  float *P = part + (size_t)blockIdx.x * (NW + COUT);
#pragma unroll
  for (int tap = 0; tap < 9; ++tap)
#pragma unroll
    for (int co = 0; co < NCO; ++co)
#pragma unroll
      for (int ci = 0; ci < NCI; ++ci)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          P[(tap * COUT + co * 16 + 4 * lk + e) * CIN + ci * 16 + lc] = acc[tap][co][ci][e];
  if (lane < COUT) P[NW + lane] = bacc;
}

__global__ void __launch_bounds__(256) k_sum_partials(int nblk, int n, const float *__restrict__ part,
                                                      float *__restrict__ dw, int nw, float *__restrict__ db) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= n) return;
  float s = 0.0f;
  for (int b = 0; b < nblk; ++b) s += part[(size_t)b * n + q];
  if (q < nw) dw[q] = s;
  else db[q - nw] = s;
}

// ---------------------------------------------------------------------------
// pooling, upsampling and their backward passes (32 channels).  Two
// workgroups per tile, eight lanes per pixel (four channels each), so a
// wavefront reads and writes whole 64-byte channel vectors of 8 neighbouring
// pixels.
// ---------------------------------------------------------------------------
#define NET_PIXEL8(V)                                                 \
  const NTile t = V.tiles[blockIdx.x >> 1];                           \
  const int px = (int)(threadIdx.x >> 3), k = 4 * (int)(threadIdx.x & 7); \
  const int r = t.r0 + 2 * (int)(blockIdx.x & 1) + (px >> 4), c = t.c0 + (px & 15);

__global__ void __launch_bounds__(256) k_pool(LevelView F, LevelView C, const uint16_t *__restrict__ x,
                                              uint16_t *__restrict__ y) {
  NET_PIXEL8(C)
  const NImg cm = C.img[t.img], fm = F.img[t.img];
  if (r >= cm.rows || c >= cm.cols) return;
  const long long q = cm.off + (long long)r * cm.cols + c;
  float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int fr = 2 * r + a, fc = 2 * c + b;
      if (fr < fm.rows && fc < fm.cols) {
        const s4 v = ld4(x + (fm.off + (long long)fr * fm.cols + fc) * 32 + k);
#pragma unroll
        for (int e = 0; e < 4; ++e) s[e] += bf2f((uint16_t)v[e]);
      }
    }
  s4 o;
#pragma unroll
  for (int e = 0; e < 4; ++e) o[e] = (short)f2bf(0.25f * s[e]);
  st4(y + q * 32 + k, o);
}

__global__ void __launch_bounds__(256) k_upadd(LevelView F, LevelView C, const uint16_t *__restrict__ a,
                                               const uint16_t *__restrict__ b, uint16_t *__restrict__ y) {
  NET_PIXEL8(F)
  const NImg fm = F.img[t.img], cm = C.img[t.img];
  if (r >= fm.rows || c >= fm.cols) return;
  const long long p = fm.off + (long long)r * fm.cols + c;
  const long long q = cm.off + (long long)(r >> 1) * cm.cols + (c >> 1);
  const s4 va = ld4(a + p * 32 + k), vb = ld4(b + q * 32 + k);
  s4 o;
#pragma unroll
  for (int e = 0; e < 4; ++e) o[e] = (short)f2bf(bf2f((uint16_t)va[e]) + bf2f((uint16_t)vb[e]));
  st4(y + p * 32 + k, o);
}

__global__ void __launch_bounds__(256) k_gpool(LevelView F, LevelView C, const uint16_t *__restrict__ a,
                                               const uint16_t *__restrict__ pc, const uint16_t *__restrict__ yact,
                                               uint16_t *__restrict__ g) {
  NET_PIXEL8(F)
  const NImg fm = F.img[t.img];
  if (r >= fm.rows || c >= fm.cols) return;
  const long long p = fm.off + (long long)r * fm.cols + c;
  const s4 ya = ld4(yact + p * 32 + k);
  s4 va = s4{0, 0, 0, 0}, vc = s4{0, 0, 0, 0};
  if (a) va = ld4(a + p * 32 + k);
  if (pc) {
    const NImg cm = C.img[t.img];
    vc = ld4(pc + (cm.off + (long long)(r >> 1) * cm.cols + (c >> 1)) * 32 + k);
  }
  s4 o;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const float s = bf2f((uint16_t)va[e]) + 0.25f * bf2f((uint16_t)vc[e]);
    o[e] = (bf2f((uint16_t)ya[e]) > 0.0f) ? (short)f2bf(s) : (short)0;
  }
  st4(g + p * 32 + k, o);
}

__global__ void __launch_bounds__(256) k_gup(LevelView F, LevelView C, const uint16_t *__restrict__ a,
                                             const uint16_t *__restrict__ yact, uint16_t *__restrict__ g) {
  NET_PIXEL8(C)
  const NImg cm = C.img[t.img], fm = F.img[t.img];
  if (r >= cm.rows || c >= cm.cols) return;
  const long long q = cm.off + (long long)r * cm.cols + c;
  float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int da = 0; da < 2; ++da)
#pragma unroll
    for (int db = 0; db < 2; ++db) {
      const int fr = 2 * r + da, fc = 2 * c + db;
      if (fr < fm.rows && fc < fm.cols) {
        const s4 v = ld4(a + (fm.off + (long long)fr * fm.cols + fc) * 32 + k);
#pragma unroll
        for (int e = 0; e < 4; ++e) s[e] += bf2f((uint16_t)v[e]);
      }
    }
  const s4 ya = ld4(yact + q * 32 + k);
  s4 o;
#pragma unroll
  for (int e = 0; e < 4; ++e) o[e] = (bf2f((uint16_t)ya[e]) > 0.0f) ? (short)f2bf(s[e]) : (short)0;
  st4(g + q * 32 + k, o);
}

// ---------------------------------------------------------------------------
// loss, optimiser, weight packing
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_loss(LevelView L, const float *__restrict__ out, const float *__restrict__ tg,
                                              const float *__restrict__ wt, float scale, uint16_t *__restrict__ g,
                                              double *__restrict__ partial) {
  __shared__ double red[256];
  double acc = 0.0;
  const int lane = threadIdx.x & 63;
  const int wid = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (wid < L.ntiles) {
    const NTile t = L.tiles[wid];
    const NImg im = L.img[t.img];
    const int r = t.r0 + (lane >> 4), c = t.c0 + (lane & 15);
    if (r < im.rows && c < im.cols) {
      const long long p = im.off + (long long)r * im.cols + c;
      const float w = wt[p];
      const float d0 = out[2 * p] - tg[2 * p], d1 = out[2 * p + 1] - tg[2 * p + 1];
      acc = (double)w * ((double)d0 * d0 + (double)d1 * d1);
      const s4 z = s4{0, 0, 0, 0};
      s4 o = z;
      o[0] = (short)f2bf(2.0f * scale * w * d0);
      o[1] = (short)f2bf(2.0f * scale * w * d1);
      st4(g + p * IMAU_NET_OUT, o);
#pragma unroll
      for (int k = 4; k < IMAU_NET_OUT; k += 4) st4(g + p * IMAU_NET_OUT + k, z);
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_adam(long long n, float *__restrict__ p, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                                              float b2, float eps, float bc1, float bc2) {
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q >= n) return;
  const float gq = g[q];
  const float mq = b1 * m[q] + (1.0f - b1) * gq;
  const float vq = b2 * v[q] + (1.0f - b2) * gq * gq;
  m[q] = mq;
  v[q] = vq;
  p[q] -= lr * (mq / bc1) / (sqrtf(vq / bc2) + eps);
}

__global__ void __launch_bounds__(256) k_pack(int cin, int cout, const float *__restrict__ w, uint16_t *__restrict__ wf,
                                              uint16_t *__restrict__ wt) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= 9 * cout * cin) return;
  const int tap = q / (cout * cin), co = (q / cin) % cout, ci = q % cin;
  wf[q] = f2bf(w[q]);
  wt[(tap * cin + ci) * cout + co] = f2bf(w[((8 - tap) * cout + co) * cin + ci]);
}

inline int nblk4(int ntiles) { return (ntiles + 3) / 4; }

}  // namespace

namespace imau {
namespace net {

bool conv(hipStream_t s, const LevelView &L, int cin, int cout, const uint16_t *x, const uint16_t *w,
          const float *bias, int relu, void *y, int yf32) {
  if (L.ntiles == 0) return true;
  // persistent waves: at most 2048 workgroups, each wavefront strides over the tiles
  const int nb = std::min(nblk4(L.ntiles), 2048);
  if (cin == 16 && cout == 32 && !yf32)
    hipLaunchKernelGGL((k_conv<16, 32, false>), dim3(nb), dim3(256), 0, s, L.img, L.tiles, L.ntiles, x, w, bias, relu, y);
  else if (cin == 32 && cout == 32 && !yf32)
    hipLaunchKernelGGL((k_conv32<32, false>), dim3(nb), dim3(256), 0, s, L.img, L.tiles, L.ntiles, x, w, bias, relu, y);
  else if (cin == 32 && cout == 16 && yf32)
    hipLaunchKernelGGL((k_conv32<16, true>), dim3(nb), dim3(256), 0, s, L.img, L.tiles, L.ntiles, x, w, bias, relu, y);
  else if (cin == 32 && cout == 16 && !yf32)
    hipLaunchKernelGGL((k_conv32<16, false>), dim3(nb), dim3(256), 0, s, L.img, L.tiles, L.ntiles, x, w, bias, relu, y);
  else
    return false;
  return true;
}

int dw_blocks(const LevelView &L) { return std::max(1, std::min(L.ntiles, 1024)); }

bool conv_dw(hipStream_t s, const LevelView &L, int cin, int cout, const uint16_t *g, const uint16_t *x, float *dw,
             float *db, float *part) {
  const int nb = dw_blocks(L);
  const int n = 9 * cout * cin + cout;
  if (L.ntiles == 0) return false;
  if (cin == 16 && cout == 32)
    hipLaunchKernelGGL((k_conv_dw<16, 32>), dim3(nb), dim3(64), 0, s, L.img, L.tiles, L.ntiles, g, x, part);
  else if (cin == 32 && cout == 32)
    hipLaunchKernelGGL((k_conv_dw<32, 32>), dim3(nb), dim3(64), 0, s, L.img, L.tiles, L.ntiles, g, x, part);
  else if (cin == 32 && cout == 16)
    hipLaunchKernelGGL((k_conv_dw<32, 16>), dim3(nb), dim3(64), 0, s, L.img, L.tiles, L.ntiles, g, x, part);
  else
    return false;
  hipLaunchKernelGGL(k_sum_partials, dim3((n + 255) / 256), dim3(256), 0, s, nb, n, part, dw, 9 * cout * cin, db);
  return true;
}

void sum_partials(hipStream_t s, int nblk, int n, const float *part, float *dw, int nw, float *db) {
  hipLaunchKernelGGL(k_sum_partials, dim3((n + 255) / 256), dim3(256), 0, s, nblk, n, part, dw, nw, db);
}

// two workgroups of 256 threads (32 pixels x 8 lanes) per 64-pixel tile
void pool(hipStream_t s, const LevelView &F, const LevelView &C, const uint16_t *x, uint16_t *y) {
  if (C.ntiles) hipLaunchKernelGGL(k_pool, dim3(2 * C.ntiles), dim3(256), 0, s, F, C, x, y);
}
void upadd(hipStream_t s, const LevelView &F, const LevelView &C, const uint16_t *a, const uint16_t *b, uint16_t *y) {
  if (F.ntiles) hipLaunchKernelGGL(k_upadd, dim3(2 * F.ntiles), dim3(256), 0, s, F, C, a, b, y);
}
void gpool(hipStream_t s, const LevelView &F, const LevelView &C, const uint16_t *a, const uint16_t *pc,
           const uint16_t *yact, uint16_t *g) {
  if (F.ntiles) hipLaunchKernelGGL(k_gpool, dim3(2 * F.ntiles), dim3(256), 0, s, F, C, a, pc, yact, g);
}
void gup(hipStream_t s, const LevelView &F, const LevelView &C, const uint16_t *a, const uint16_t *yact, uint16_t *g) {
  if (C.ntiles) hipLaunchKernelGGL(k_gup, dim3(2 * C.ntiles), dim3(256), 0, s, F, C, a, yact, g);
}
int loss_blocks(const LevelView &L) { return std::max(1, nblk4(L.ntiles)); }
void loss(hipStream_t s, const LevelView &L, const float *out, const float *t, const float *w, float scale,
          uint16_t *g, double *partial) {
  hipLaunchKernelGGL(k_loss, dim3(loss_blocks(L)), dim3(256), 0, s, L, out, t, w, scale, g, partial);
}
void adam(hipStream_t s, long long n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2,
          float eps, float bc1, float bc2) {
  if (n > 0)
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, p, g, m, v, lr, b1, b2, eps,
                       bc1, bc2);
}
void pack(hipStream_t s, int cin, int cout, const float *w, uint16_t *wf, uint16_t *wt) {
  const int n = 9 * cin * cout;
  hipLaunchKernelGGL(k_pack, dim3((n + 255) / 256), dim3(256), 0, s, cin, cout, w, wf, wt);
}

}  // namespace net
}  // namespace imau
