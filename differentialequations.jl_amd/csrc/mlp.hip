// mlp.hip -- the dt0 / method-class / cost / never-stiff predictor: a 16 -> 64 -> 64 -> 64 -> 4
// GELU MLP with BF16 weights and activations and FP32 accumulation on the
// CDNA4 matrix cores (v_mfma_f32_16x16x16_bf16 / v_mfma_f32_16x16x32_bf16),
// trained on the GPU by the kernels below (forward + backward per tile,
// weight-gradient GEMMs over the batch, Adam).
//
// Layout trick: every layer is computed transposed, H^T = W^T . X^T, with
// the 16 samples of a tile as the MFMA N dimension.  A 16x16 f32 MFMA
// result leaves lane l holding rows 4*(l/16)..+3 of column l%16, which is
// exactly the lane/slot arrangement of the B operand of the next MFMA when
// K slots are numbered k = 4*(l/16)+j (and 16 + 4*(l/16) + j-4 for the
// second half of a 32-deep K chunk).  A and B use the same slot numbering,
// so the product pairs the same logical feature whatever the hardware's
// internal K order is.  No LDS transposes are needed in the forward or
// backward chain; the weight-gradient GEMMs read activations that the
// forward pass stored feature-major.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mideq.h"

namespace mideq {
namespace mlp {

constexpr int NIN = 16, NH = 64, NOUT = 4;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + NIN * NH;
constexpr int OFF_W2 = OFF_B1 + NH;
constexpr int OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH;
constexpr int OFF_B3 = OFF_W3 + NH * NH;
constexpr int OFF_W4 = OFF_B3 + NH;
constexpr int OFF_B4 = OFF_W4 + NH * NOUT;
constexpr int NPARAM = OFF_B4 + NOUT;
static_assert(NPARAM == MIDEQ_MLP_NPARAM, "parameter count");

typedef short v4s __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8b __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr float GELU_C = 0.7978845608028654f;
constexpr float GELU_A = 0.044715f;

__device__ __forceinline__ short f2bf(float x) {
  unsigned u = __float_as_uint(x);
  u = u + 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

__device__ __forceinline__ float gelu(float z) {
  float t = tanhf(GELU_C * (z + GELU_A * (z * z * z)));
  return 0.5f * z * (1.0f + t);
}

__device__ __forceinline__ float gelu_grad(float z) {
  float t = tanhf(GELU_C * (z + GELU_A * (z * z * z)));
  return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * GELU_C * (1.0f + 3.0f * GELU_A * z * z);
}

// logical feature of slot j (0..7) in a 32-deep K chunk for lane group g
__device__ __forceinline__ int kmap32(int g, int j) { return j < 4 ? 4 * g + j : 12 + 4 * g + j; }

__device__ __forceinline__ v4f mfma16(v4s a, v4s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mfma32(v8s a, v8s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8b, a),
                                                 __builtin_bit_cast(v8b, b), c, 0, 0, 0);
}

struct Frags {
  // forward: A = W^T (rows: out features, K: in features)
  v4s a1[4];
  v8s a2[4][2], a3[4][2], a4[2];
  float b1[4][4], b2[4][4], b3[4][4], b4[4];
};

struct BackFrags {
  // backward: A = W (rows: in features, K: out features)
  v4s t4[4];
  v8s t3[4][2], t2[4][2];
};

__device__ void load_frags(const short* wb, const float* w, Frags& F, int lane) {
  const int g = lane >> 4, col = lane & 15;
#pragma unroll
  for (int ob = 0; ob < 4; ob++) {
    int out = 16 * ob + col;
#pragma unroll
    for (int j = 0; j < 4; j++) F.a1[ob][j] = wb[OFF_W1 + (4 * g + j) * NH + out];
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int j = 0; j < 8; j++) {
        int in = 32 * c + kmap32(g, j);
        F.a2[ob][c][j] = wb[OFF_W2 + in * NH + out];
        F.a3[ob][c][j] = wb[OFF_W3 + in * NH + out];
      }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      int f = 16 * ob + 4 * g + r;
      F.b1[ob][r] = w[OFF_B1 + f];
      F.b2[ob][r] = w[OFF_B2 + f];
      F.b3[ob][r] = w[OFF_B3 + f];
    }
  }
#pragma unroll
  for (int c = 0; c < 2; c++)
#pragma unroll
    for (int j = 0; j < 8; j++) {
      int in = 32 * c + kmap32(g, j);
      F.a4[c][j] = col < NOUT ? wb[OFF_W4 + in * NOUT + col] : (short)0;
    }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    int f = 4 * g + r;
    F.b4[r] = f < NOUT ? w[OFF_B4 + f] : 0.0f;
  }
}

__device__ void load_back_frags(const short* wb, BackFrags& T, int lane) {
  const int g = lane >> 4, col = lane & 15;
#pragma unroll
  for (int ib = 0; ib < 4; ib++) {
    int in = 16 * ib + col;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      int out = 4 * g + j;
      T.t4[ib][j] = out < NOUT ? wb[OFF_W4 + in * NOUT + out] : (short)0;
    }
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int j = 0; j < 8; j++) {
        int out = 32 * c + kmap32(g, j);
        T.t3[ib][c][j] = wb[OFF_W3 + in * NH + out];
        T.t2[ib][c][j] = wb[OFF_W2 + in * NH + out];
      }
  }
}

// B operand of a 32-deep chunk from two consecutive 16-feature blocks
__device__ __forceinline__ v8s chunk(const v4s lo, const v4s hi) {
  v8s b;
  b[0] = lo[0]; b[1] = lo[1]; b[2] = lo[2]; b[3] = lo[3];
  b[4] = hi[0]; b[5] = hi[1]; b[6] = hi[2]; b[7] = hi[3];
  return b;
}

struct Tile {
  v4s x;          // normalised input, bf16 (features 4g..4g+3)
  v4f z[3][4];    // pre-activations of the hidden layers
  v4s h[3][4];    // hidden activations, bf16
  v4f y;          // outputs (rows 0..3 on lanes with g == 0)
};

__device__ void forward_tile(const Frags& F, const float* X, const float* mean,
                             const float* rstd, int64_t s, bool valid, int g, Tile& T) {
#pragma unroll
  for (int j = 0; j < 4; j++) {
    int f = 4 * g + j;
    float xv = valid ? X[s * NIN + f] : 0.0f;
    T.x[j] = valid ? f2bf((xv - mean[f]) * rstd[f]) : (short)0;
  }
  const v4f zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ob = 0; ob < 4; ob++) {
    v4f d = mfma16(F.a1[ob], T.x, zero);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      float z = d[r] + F.b1[ob][r];
      T.z[0][ob][r] = z;
      T.h[0][ob][r] = f2bf(gelu(z));
    }
  }
#pragma unroll
  for (int layer = 1; layer < 3; layer++) {
#pragma unroll
    for (int ob = 0; ob < 4; ob++) {
      v4f d = zero;
#pragma unroll
      for (int c = 0; c < 2; c++) {
        v8s a = layer == 1 ? F.a2[ob][c] : F.a3[ob][c];
        d = mfma32(a, chunk(T.h[layer - 1][2 * c], T.h[layer - 1][2 * c + 1]), d);
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        float z = d[r] + (layer == 1 ? F.b2[ob][r] : F.b3[ob][r]);
        T.z[layer][ob][r] = z;
        T.h[layer][ob][r] = f2bf(gelu(z));
      }
    }
  }
  v4f d = zero;
#pragma unroll
  for (int c = 0; c < 2; c++) d = mfma32(F.a4[c], chunk(T.h[2][2 * c], T.h[2][2 * c + 1]), d);
#pragma unroll
  for (int r = 0; r < 4; r++) T.y[r] = d[r] + F.b4[r];
}

// ---------------------------------------------------------------- inference
__global__ __launch_bounds__(256) void k_mlp_infer(const short* wb, const float* w,
                                                   const float* mean, const float* rstd,
                                                   const float* X, int64_t B, float* Y) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, col = lane & 15;
  Frags F;
  load_frags(wb, w, F, lane);
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int64_t ntiles = (B + 15) / 16;
  for (int64_t tile = wave; tile < ntiles; tile += nwaves) {
    int64_t s = tile * 16 + col;
    bool valid = s < B;
    Tile T;
    forward_tile(F, X, mean, rstd, s, valid, g, T);
    if (valid && g == 0) {
      Y[s * NOUT + 0] = T.y[0];
      Y[s * NOUT + 1] = T.y[1];
      Y[s * NOUT + 2] = T.y[2];
      Y[s * NOUT + 3] = T.y[3];
    }
  }
}

// ---------------------------------------------------------------- training
// Forward + backward of one 16-sample tile per wave.  Stores bf16 inputs of
// every layer (Hs) and bf16 pre-activation gradients (dZs) feature-major
// ([feature][Bp]) for the weight-gradient GEMMs, adds bias gradients and
// the loss with atomics.
struct TrainArgs {
  const short* wb;
  const float* w;
  const float* mean;
  const float* rstd;
  const float* X;
  const float* target;
  const float* label;
  const float* cost;   // may be null: no cost term
  const float* never;  // may be null: no never-stiff term
  int64_t B, Bp;
  float w_dt, w_cls, w_cost, w_never;
  short* hs0;   // [NIN][Bp]
  short* hs;    // 3 x [NH][Bp]: inputs of layers 2, 3, 4
  short* dzs;   // 3 x [NH][Bp]: dZ of layers 1, 2, 3
  short* dz4;   // [NOUT][Bp]
  float* grad;  // [NPARAM] (bias parts accumulated here)
  float* loss;  // [1]
};

__device__ __forceinline__ float sum16(float v) {
  // sum over the 16 lanes of a row group (same g)
  v += __shfl_xor(v, 1, 64);
  v += __shfl_xor(v, 2, 64);
  v += __shfl_xor(v, 4, 64);
  v += __shfl_xor(v, 8, 64);
  return v;
}

__global__ __launch_bounds__(64) void k_mlp_fwdbwd(TrainArgs a) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, col = lane & 15;
  const int64_t tile = blockIdx.x;
  const int64_t s = tile * 16 + col;
  const bool valid = s < a.B;
  Frags F;
  load_frags(a.wb, a.w, F, lane);
  Tile T;
  forward_tile(F, a.X, a.mean, a.rstd, s, valid, g, T);
  const int64_t Bp = a.Bp;
  // stored activations: layer inputs
#pragma unroll
  for (int j = 0; j < 4; j++) a.hs0[(int64_t)(4 * g + j) * Bp + s] = T.x[j];
#pragma unroll
  for (int l = 0; l < 3; l++)
#pragma unroll
    for (int ob = 0; ob < 4; ob++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        a.hs[(int64_t)l * NH * Bp + (int64_t)(16 * ob + 4 * g + r) * Bp + s] =
            valid ? T.h[l][ob][r] : (short)0;
  // output gradient (rows 0..3 on lanes g == 0)
  float dy0 = 0.f, dy1 = 0.f, dy2 = 0.f, dy3 = 0.f, lsum = 0.f;
  if (valid && g == 0) {
    float t = a.target[s], c = a.label[s];
    float e = T.y[0] - t;
    float y1 = T.y[1];
    float sig = 1.0f / (1.0f + expf(-y1));
    dy0 = 2.0f * e * a.w_dt / (float)a.B;
    dy1 = (sig - c) * a.w_cls / (float)a.B;
    float bce = fmaxf(y1, 0.f) - y1 * c + log1pf(expf(-fabsf(y1)));
    float ec = a.cost ? T.y[2] - a.cost[s] : 0.0f;
    dy2 = 2.0f * ec * a.w_cost / (float)a.B;
    float bn = 0.0f;
    if (a.never) {
      const float y3 = T.y[3], cn = a.never[s];
      const float sig3 = 1.0f / (1.0f + expf(-y3));
      dy3 = (sig3 - cn) * a.w_never / (float)a.B;
      bn = fmaxf(y3, 0.f) - y3 * cn + log1pf(expf(-fabsf(y3)));
    }
    lsum = (a.w_dt * e * e + a.w_cls * bce + a.w_cost * ec * ec + a.w_never * bn) / (float)a.B;
  }
  v4s d4 = {0, 0, 0, 0};
  if (g == 0) {
    d4[0] = f2bf(dy0);
    d4[1] = f2bf(dy1);
    d4[2] = f2bf(dy2);
    d4[3] = f2bf(dy3);
  }
  if (g == 0) {
    a.dz4[0 * Bp + s] = d4[0];
    a.dz4[1 * Bp + s] = d4[1];
    a.dz4[2 * Bp + s] = d4[2];
    a.dz4[3 * Bp + s] = d4[3];
  }
  {
    float b0 = sum16(dy0), b1 = sum16(dy1), b2 = sum16(dy2), b3 = sum16(dy3), ls = sum16(lsum);
    if (lane == 0) {
      atomicAdd(&a.grad[OFF_B4 + 0], b0);
      atomicAdd(&a.grad[OFF_B4 + 1], b1);
      atomicAdd(&a.grad[OFF_B4 + 2], b2);
      atomicAdd(&a.grad[OFF_B4 + 3], b3);
      atomicAdd(a.loss, ls);
    }
  }
  BackFrags W;
  load_back_frags(a.wb, W, lane);
  const v4f zero = {0.f, 0.f, 0.f, 0.f};
  // dH3 = W4 . dZ4 ; dZ3 = dH3 * gelu'(Z3)
  v4s dzb[4];
#pragma unroll
  for (int ib = 0; ib < 4; ib++) {
    v4f dh = mfma16(W.t4[ib], d4, zero);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      float dz = valid ? dh[r] * gelu_grad(T.z[2][ib][r]) : 0.0f;
      dzb[ib][r] = f2bf(dz);
      float bs = sum16(dz);
      if (col == 0) atomicAdd(&a.grad[OFF_B3 + 16 * ib + 4 * g + r], bs);
      a.dzs[(int64_t)2 * NH * Bp + (int64_t)(16 * ib + 4 * g + r) * Bp + s] = dzb[ib][r];
    }
  }
  // layers 2 and 1
#pragma unroll
  for (int l = 1; l >= 0; l--) {
    v4s nb[4];
#pragma unroll
    for (int ib = 0; ib < 4; ib++) {
      v4f dh = zero;
#pragma unroll
      for (int c = 0; c < 2; c++) {
        v8s A = (l == 1) ? W.t3[ib][c] : W.t2[ib][c];
        dh = mfma32(A, chunk(dzb[2 * c], dzb[2 * c + 1]), dh);
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        float dz = valid ? dh[r] * gelu_grad(T.z[l][ib][r]) : 0.0f;
        nb[ib][r] = f2bf(dz);
        float bs = sum16(dz);
        if (col == 0)
          atomicAdd(&a.grad[(l == 1 ? OFF_B2 : OFF_B1) + 16 * ib + 4 * g + r], bs);
        a.dzs[(int64_t)l * NH * Bp + (int64_t)(16 * ib + 4 * g + r) * Bp + s] = nb[ib][r];
      }
    }
#pragma unroll
    for (int ib = 0; ib < 4; ib++) dzb[ib] = nb[ib];
  }
}

// Weight gradients dW[in][out] = sum_s H[in][s] dZ[out][s], one wave per
// 16x16 block, K = samples in chunks of 32.
struct WGradArgs {
  const short* hs0;
  const short* hs;
  const short* dzs;
  const short* dz4;
  int64_t Bp;
  float* grad;
};

__global__ __launch_bounds__(64) void k_mlp_wgrad(WGradArgs a) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, col = lane & 15;
  // block list: layer 1: 1x4, layer 2: 4x4, layer 3: 4x4, layer 4: 4x1
  int bid = blockIdx.x;
  int layer, ib, ob;
  if (bid < 4) { layer = 0; ib = 0; ob = bid; }
  else if (bid < 20) { layer = 1; ib = (bid - 4) / 4; ob = (bid - 4) % 4; }
  else if (bid < 36) { layer = 2; ib = (bid - 20) / 4; ob = (bid - 20) % 4; }
  else { layer = 3; ib = bid - 36; ob = 0; }
  const int64_t Bp = a.Bp;
  const short* H = layer == 0 ? a.hs0 : a.hs + (int64_t)(layer - 1) * NH * Bp;
  const int in = 16 * ib + col;     // A row (lane's column index)
  const int out = 16 * ob + col;    // B column
  const int outdim = layer == 3 ? NOUT : NH;
  const short* Hrow = H + (int64_t)in * Bp;
  const short* Drow = nullptr;
  bool bvalid = true;
  if (layer == 3) {
    bvalid = out < NOUT;
    Drow = a.dz4 + (int64_t)(bvalid ? out : 0) * Bp;
  } else {
    Drow = a.dzs + (int64_t)layer * NH * Bp + (int64_t)out * Bp;
  }
  v4f acc = {0.f, 0.f, 0.f, 0.f};
  for (int64_t s0 = 0; s0 < Bp; s0 += 32) {
    v8s A, Bv;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      int64_t s = s0 + kmap32(g, j);
      A[j] = Hrow[s];
      Bv[j] = bvalid ? Drow[s] : (short)0;
    }
    acc = mfma32(A, Bv, acc);
  }
  // lane holds dW rows (in) 16*ib + 4g + r, column (out) 16*ob + col
  int off = layer == 0 ? OFF_W1 : layer == 1 ? OFF_W2 : layer == 2 ? OFF_W3 : OFF_W4;
  if (out < outdim || layer != 3) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      int inr = 16 * ib + 4 * g + r;
      if (layer == 3 && out >= NOUT) continue;
      a.grad[off + inr * outdim + out] = acc[r];
    }
  }
}

__global__ void k_mlp_zero(float* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

struct AdamArgs {
  float* w;
  short* wb;
  float* m;
  float* v;
  const float* grad;
  float lr, b1, b2, eps, c1, c2;  // c1 = 1 - b1^t, c2 = 1 - b2^t
};

__global__ void k_mlp_adam(AdamArgs a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NPARAM) return;
  float gr = a.grad[i];
  float m = a.b1 * a.m[i] + (1.0f - a.b1) * gr;
  float v = a.b2 * a.v[i] + (1.0f - a.b2) * gr * gr;
  a.m[i] = m;
  a.v[i] = v;
  float w = a.w[i] - a.lr * (m / a.c1) / (sqrtf(v / a.c2) + a.eps);
  a.w[i] = w;
  a.wb[i] = f2bf(w);
}

__global__ void k_mlp_tobf16(const float* w, short* wb, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) wb[i] = f2bf(w[i]);
}

// ------------------------------------------------------------ launch order
// Counting sort of the predictions into a launch order: predicted
// non-stiff class first, each class by decreasing predicted log2 cost in
// 1/64-octave buckets (log2 cost clamped to [0, 32)).  Chunks of ORD_CHUNK
// samples per block rank their samples in LDS, so the only global atomics
// are one per (block, non-empty bucket).  Order inside a bucket is arbitrary.
constexpr int ORD_NB = 2048;
constexpr int ORD_BINS = 2 * ORD_NB;
constexpr int ORD_CHUNK = 4096;

__device__ __forceinline__ int ord_key(const float* Y, int64_t s) {
  float cls = Y[s * NOUT + 1], c = Y[s * NOUT + 2];
  int b;
  if (!(c >= 0.0f))
    b = 0;
  else if (c >= 32.0f)
    b = ORD_NB - 1;
  else
    b = (int)(c * 64.0f);
  return (cls > 0.0f ? ORD_NB : 0) + (ORD_NB - 1 - b);
}

__global__ __launch_bounds__(256) void k_ord_hist(const float* Y, int64_t B, int* hist) {
  __shared__ int lh[ORD_BINS];
  for (int b = threadIdx.x; b < ORD_BINS; b += blockDim.x) lh[b] = 0;
  __syncthreads();
  const int64_t s0 = (int64_t)blockIdx.x * ORD_CHUNK;
  const int64_t s1 = s0 + ORD_CHUNK < B ? s0 + ORD_CHUNK : B;
  for (int64_t s = s0 + threadIdx.x; s < s1; s += blockDim.x) atomicAdd(&lh[ord_key(Y, s)], 1);
  __syncthreads();
  for (int b = threadIdx.x; b < ORD_BINS; b += blockDim.x)
    if (lh[b]) atomicAdd(&hist[b], lh[b]);
}

// exclusive scan of the ORD_BINS counts: one block of ORD_SCAN threads,
// ORD_BINS / ORD_SCAN consecutive bins per thread
constexpr int ORD_SCAN = 1024;
constexpr int ORD_PER = ORD_BINS / ORD_SCAN;
static_assert(ORD_BINS % ORD_SCAN == 0, "scan layout");

__global__ __launch_bounds__(ORD_SCAN) void k_ord_scan(const int* hist, int* offs) {
  __shared__ int v[ORD_SCAN];
  const int t = threadIdx.x;
  int loc[ORD_PER];
  int sum = 0;
#pragma unroll
  for (int k = 0; k < ORD_PER; k++) {
    loc[k] = hist[t * ORD_PER + k];
    sum += loc[k];
  }
  v[t] = sum;
  __syncthreads();
  for (int d = 1; d < ORD_SCAN; d <<= 1) {
    int x = t >= d ? v[t - d] : 0;
    __syncthreads();
    v[t] += x;
    __syncthreads();
  }
  int run = v[t] - sum;  // exclusive prefix of this thread's bins
#pragma unroll
  for (int k = 0; k < ORD_PER; k++) {
    offs[t * ORD_PER + k] = run;
    run += loc[k];
  }
}

__global__ __launch_bounds__(256) void k_ord_scatter(const float* Y, int64_t B, int* offs,
                                                     int64_t* order) {
  __shared__ int lh[ORD_BINS];
  __shared__ int base[ORD_BINS];
  for (int b = threadIdx.x; b < ORD_BINS; b += blockDim.x) lh[b] = 0;
  __syncthreads();
  const int64_t s0 = (int64_t)blockIdx.x * ORD_CHUNK;
  const int64_t s1 = s0 + ORD_CHUNK < B ? s0 + ORD_CHUNK : B;
  for (int64_t s = s0 + threadIdx.x; s < s1; s += blockDim.x) atomicAdd(&lh[ord_key(Y, s)], 1);
  __syncthreads();
  for (int b = threadIdx.x; b < ORD_BINS; b += blockDim.x) {
    base[b] = lh[b] ? atomicAdd(&offs[b], lh[b]) : 0;
    lh[b] = 0;
  }
  __syncthreads();
  for (int64_t s = s0 + threadIdx.x; s < s1; s += blockDim.x) {
    int k = ord_key(Y, s);
    int r = atomicAdd(&lh[k], 1);
    order[base[k] + r] = s;
  }
}

__global__ void k_ord_zero(int* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace mlp
}  // namespace mideq

using namespace mideq::mlp;

// ---------------------------------------------------------------- C ABI
struct mideq_mlp {
  int device = 0;
  float* w = nullptr;
  short* wb = nullptr;
  float* m = nullptr;
  float* v = nullptr;
  float* grad = nullptr;
  float* norm = nullptr;  // mean[16], rstd[16]
  float* loss = nullptr;
  short* scratch = nullptr;
  size_t scratch_bytes = 0;
  hipStream_t scratch_stream = nullptr;  // stream of the last call that used scratch
  int* ord = nullptr;  // launch-order histogram and offsets [2][ORD_BINS]
  int ncu = 256;
  int64_t step = 0;
  std::string err;
};

static int mlp_err(mideq_mlp* m, int code, const char* msg, hipError_t e = hipSuccess) {
  if (m) m->err = std::string(msg) + (e != hipSuccess ? std::string(": ") + hipGetErrorString(e) : "");
  return code;
}

extern "C" int mideq_mlp_create(mideq_mlp** out, int32_t device) {
  if (!out) return MIDEQ_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MIDEQ_E_NODEVICE;
  if (device < 0 || device >= ndev) return MIDEQ_E_ARG;
  if (hipSetDevice(device) != hipSuccess) return MIDEQ_E_HIP;
  mideq_mlp* m = new mideq_mlp();
  m->device = device;
  bool ok = hipMalloc((void**)&m->w, sizeof(float) * NPARAM) == hipSuccess &&
            hipMalloc((void**)&m->wb, sizeof(short) * NPARAM) == hipSuccess &&
            hipMalloc((void**)&m->m, sizeof(float) * NPARAM) == hipSuccess &&
            hipMalloc((void**)&m->v, sizeof(float) * NPARAM) == hipSuccess &&
            hipMalloc((void**)&m->grad, sizeof(float) * NPARAM) == hipSuccess &&
            hipMalloc((void**)&m->norm, sizeof(float) * 2 * NIN) == hipSuccess &&
            hipMalloc((void**)&m->loss, sizeof(float) * 4) == hipSuccess &&
            hipMalloc((void**)&m->ord, sizeof(int) * 2 * ORD_BINS) == hipSuccess;
  hipDeviceProp_t prop;
  if (ok && hipGetDeviceProperties(&prop, device) == hipSuccess) m->ncu = prop.multiProcessorCount;
  if (!ok) {
    mideq_mlp_destroy(m);
    return MIDEQ_E_NOMEM;
  }
  std::vector<float> zero(NPARAM, 0.f), norm(2 * NIN);
  for (int i = 0; i < NIN; i++) {
    norm[i] = 0.f;
    norm[NIN + i] = 1.f;
  }
  if (hipMemcpy(m->w, zero.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->m, zero.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->v, zero.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->norm, norm.data(), sizeof(float) * 2 * NIN, hipMemcpyHostToDevice) != hipSuccess) {
    mideq_mlp_destroy(m);
    return MIDEQ_E_HIP;
  }
  hipLaunchKernelGGL(k_mlp_tobf16, dim3((NPARAM + 255) / 256), dim3(256), 0, nullptr, m->w, m->wb, NPARAM);
  if (hipDeviceSynchronize() != hipSuccess) {
    mideq_mlp_destroy(m);
    return MIDEQ_E_HIP;
  }
  *out = m;
  return MIDEQ_OK;
}

extern "C" void mideq_mlp_destroy(mideq_mlp* m) {
  if (!m) return;
  hipError_t e = hipSetDevice(m->device);
  e = hipDeviceSynchronize();
  for (void* p : {(void*)m->w, (void*)m->wb, (void*)m->m, (void*)m->v, (void*)m->grad,
                  (void*)m->norm, (void*)m->loss, (void*)m->scratch, (void*)m->ord})
    if (p) e = hipFree(p);
  (void)e;
  delete m;
}

extern "C" const char* mideq_mlp_last_error(const mideq_mlp* m) {
  return m ? m->err.c_str() : "null mlp";
}

extern "C" int mideq_mlp_set_params(mideq_mlp* m, const float* params) {
  if (!m || !params) return MIDEQ_E_ARG;
  hipError_t e = hipSetDevice(m->device);
  if (e == hipSuccess) e = hipMemcpy(m->w, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice);
  if (e != hipSuccess) return mlp_err(m, MIDEQ_E_HIP, "set_params", e);
  std::vector<float> zero(NPARAM, 0.f);
  e = hipMemcpy(m->m, zero.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(m->v, zero.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice);
  if (e != hipSuccess) return mlp_err(m, MIDEQ_E_HIP, "set_params", e);
  m->step = 0;
  hipLaunchKernelGGL(k_mlp_tobf16, dim3((NPARAM + 255) / 256), dim3(256), 0, nullptr, m->w, m->wb, NPARAM);
  e = hipDeviceSynchronize();
  if (e != hipSuccess) return mlp_err(m, MIDEQ_E_HIP, "set_params", e);
  return MIDEQ_OK;
}

extern "C" int mideq_mlp_get_params(mideq_mlp* m, float* params) {
  if (!m || !params) return MIDEQ_E_ARG;
  hipError_t e = hipSetDevice(m->device);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(params, m->w, sizeof(float) * NPARAM, hipMemcpyDeviceToHost);
  return e == hipSuccess ? MIDEQ_OK : mlp_err(m, MIDEQ_E_HIP, "get_params", e);
}

extern "C" int mideq_mlp_get_grads(mideq_mlp* m, float* grads) {
  if (!m || !grads) return MIDEQ_E_ARG;
  hipError_t e = hipSetDevice(m->device);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(grads, m->grad, sizeof(float) * NPARAM, hipMemcpyDeviceToHost);
  return e == hipSuccess ? MIDEQ_OK : mlp_err(m, MIDEQ_E_HIP, "get_grads", e);
}

extern "C" int mideq_mlp_set_norm(mideq_mlp* m, const float* mean, const float* std_) {
  if (!m || !mean || !std_) return MIDEQ_E_ARG;
  float h[2 * NIN];
  for (int i = 0; i < NIN; i++) {
    if (!(std_[i] > 0)) return mlp_err(m, MIDEQ_E_ARG, "norm std must be > 0");
    h[i] = mean[i];
    h[NIN + i] = 1.0f / std_[i];
  }
  hipError_t e = hipSetDevice(m->device);
  if (e == hipSuccess) e = hipMemcpy(m->norm, h, sizeof h, hipMemcpyHostToDevice);
  return e == hipSuccess ? MIDEQ_OK : mlp_err(m, MIDEQ_E_HIP, "set_norm", e);
}

extern "C" int mideq_mlp_infer(mideq_mlp* m, const float* X, int64_t B, float* Y, void* stream) {
  if (!m || B < 0 || (B > 0 && (!X || !Y))) return MIDEQ_E_ARG;
  if (B == 0) return MIDEQ_OK;
  hipError_t e = hipSetDevice(m->device);
  if (e != hipSuccess) return mlp_err(m, MIDEQ_E_HIP, "infer", e);
  // every wave first loads the whole net into registers (~19 KB of
  // fragments), so keep the grid to 8 waves per CU and loop over tiles
  int64_t tiles = (B + 15) / 16;
  int64_t cap = (int64_t)m->ncu * 8;
  int64_t waves = tiles < cap ? tiles : cap;
  int blocks = (int)((waves + 3) / 4);
  hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(256), 0, (hipStream_t)stream, m->wb, m->w,
                     m->norm, m->norm + NIN, X, B, Y);
  e = hipGetLastError();
  return e == hipSuccess ? MIDEQ_OK : mlp_err(m, MIDEQ_E_HIP, "infer launch", e);
}

extern "C" int mideq_mlp_grad(mideq_mlp* m, const float* X, const float* target,
                              const float* label, const float* cost, const float* never,
                              int64_t B, float w_dt, float w_cls, float w_cost, float w_never,
                              float* loss_host, void* stream) {
  if (!m || B <= 0 || !X || !target || !label) return MIDEQ_E_ARG;
  hipError_t e = hipSetDevice(m->device);
  if (e != hipSuccess) return mlp_err(m, MIDEQ_E_HIP, "grad", e);
  hipStream_t s = (hipStream_t)stream;
  int64_t Bp = (B + 31) / 32 * 32;
  size_t need = sizeof(short) * (size_t)Bp * (NIN + 3 * NH + 3 * NH + NOUT);
  if (m->scratch_bytes < need) {
    // the previous scratch may still be read by the last call's kernels
    if (m->scratch) {
      e = hipStreamSynchronize(m->scratch_stream);
      e = hipStreamSynchronize(s);
      e = hipFree(m->scratch);
    }
    m->scratch = nullptr;
    m->scratch_bytes = 0;
    e = hipMalloc((void**)&m->scratch, need);
    if (e != hipSuccess) return mlp_err(m, MIDEQ_E_NOMEM, "grad scratch", e);
    m->scratch_bytes = need;
  }
  TrainArgs a{};
  a.wb = m->wb;
  a.w = m->w;
  a.mean = m->norm;
  a.rstd = m->norm + NIN;
  a.X = X;
  a.target = target;
  a.label = label;
  a.cost = cost;
  a.never = never;
  a.B = B;
  a.Bp = Bp;
  a.w_dt = w_dt;
  a.w_cls = w_cls;
  a.w_cost = cost ? w_cost : 0.0f;
  a.w_never = never ? w_never : 0.0f;
  m->scratch_stream = s;
  a.hs0 = m->scratch;
  a.hs = a.hs0 + (size_t)NIN * Bp;
  a.dzs = a.hs + (size_t)3 * NH * Bp;
  a.dz4 = a.dzs + (size_t)3 * NH * Bp;
  a.grad = m->grad;
  a.loss = m->loss;
  hipLaunchKernelGGL(k_mlp_zero, dim3((NPARAM + 255) / 256), dim3(256), 0, s, m->grad, NPARAM);
  hipLaunchKernelGGL(k_mlp_zero, dim3(1), dim3(64), 0, s, m->loss, 4);
  hipLaunchKernelGGL(k_mlp_fwdbwd, dim3((unsigned)(Bp / 16)), dim3(64), 0, s, a);
  WGradArgs g{a.hs0, a.hs, a.dzs, a.dz4, Bp, m->grad};
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(40), dim3(64), 0, s, g);
  e = hipGetLastError();
  if (e != hipSuccess) return mlp_err(m, MIDEQ_E_HIP, "grad launch", e);
  if (loss_host) {
    e = hipMemcpyAsync(loss_host, m->loss, sizeof(float), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return mlp_err(m, MIDEQ_E_HIP, "grad loss", e);
  }
  return MIDEQ_OK;
}

extern "C" int mideq_mlp_adam_step(mideq_mlp* m, float lr, float beta1, float beta2, float eps,
                                   void* stream) {
  if (!m) return MIDEQ_E_ARG;
  hipError_t e = hipSetDevice(m->device);
  if (e != hipSuccess) return mlp_err(m, MIDEQ_E_HIP, "adam", e);
  m->step += 1;
  AdamArgs a{m->w, m->wb, m->m, m->v, m->grad, lr, beta1, beta2, eps,
             (float)(1.0 - std::pow((double)beta1, (double)m->step)),
             (float)(1.0 - std::pow((double)beta2, (double)m->step))};
  hipLaunchKernelGGL(k_mlp_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, (hipStream_t)stream, a);
  e = hipGetLastError();
  return e == hipSuccess ? MIDEQ_OK : mlp_err(m, MIDEQ_E_HIP, "adam launch", e);
}

extern "C" int mideq_mlp_launch_order(mideq_mlp* m, const float* Y, int64_t B, int64_t* order,
                                      void* stream) {
  if (!m || B < 0 || (B > 0 && (!Y || !order))) return MIDEQ_E_ARG;
  if (B > (int64_t)0x7fffffff) return mlp_err(m, MIDEQ_E_ARG, "launch_order: B >= 2^31");
  if (B == 0) return MIDEQ_OK;
  hipError_t e = hipSetDevice(m->device);
  if (e != hipSuccess) return mlp_err(m, MIDEQ_E_HIP, "launch_order", e);
  hipStream_t s = (hipStream_t)stream;
  int* hist = m->ord;
  int* offs = m->ord + ORD_BINS;
  int nblk = (int)((B + ORD_CHUNK - 1) / ORD_CHUNK);
  hipLaunchKernelGGL(k_ord_zero, dim3((ORD_BINS + 255) / 256), dim3(256), 0, s, hist, ORD_BINS);
  hipLaunchKernelGGL(k_ord_hist, dim3(nblk), dim3(256), 0, s, Y, B, hist);
  hipLaunchKernelGGL(k_ord_scan, dim3(1), dim3(ORD_SCAN), 0, s, (const int*)hist, offs);
  hipLaunchKernelGGL(k_ord_scatter, dim3(nblk), dim3(256), 0, s, Y, B, offs, order);
  e = hipGetLastError();
  return e == hipSuccess ? MIDEQ_OK : mlp_err(m, MIDEQ_E_HIP, "launch_order launch", e);
}
