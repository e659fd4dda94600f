// mlp.hip -- the learned viscosity switch: a per-particle MLP 8 -> 64 -> 64 -> 1
// (SiLU, SiLU, sigmoid) whose output is the viscosity alpha in [0, 1].  It has
// no counterpart in the reference; it replaces the Cullen-Dehnen switch
// (CullenDehnen2010.py:218-331), which stays its fallback for features outside
// the training range.
//
// Arithmetic: BF16 weights, inputs and activations, FP32 accumulation on the
// matrix cores (v_mfma_f32_16x16x16_bf16).  One wave owns a tile of 16
// particles.  Every layer is computed transposed, Z^T = W Z_prev^T, with the
// weights as the A operand and the activations as the B operand: the
// accumulator layout of one layer (lane l holds rows 4(l/16)..+3 of column
// l%16) is then exactly the B-operand layout of the next, so activations stay
// in registers between layers.  The backward pass needs the activations in
// the A-operand layout for the weight gradients; those tiles go through LDS.
//
// Parameters (FP32 master copy, CRK_MLP_NPARAM floats):
//   W1[64][8] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3
// BF16 operand copy (crk_mlp_pack): W1p[64][16] (inputs 8..15 zero) | W2[64][64] | w3[64]
#include <cmath>

#include "pairs.hpp"

namespace crk {

constexpr int kIn = CRK_MLP_IN, kHid = CRK_MLP_HIDDEN;
constexpr int oW1 = 0, oB1 = oW1 + kHid * kIn, oW2 = oB1 + kHid, oB2 = oW2 + kHid * kHid, oW3 = oB2 + kHid,
              oB3 = oW3 + kHid, kNP = oB3 + 1;
static_assert(kNP == CRK_MLP_NPARAM, "parameter layout");
constexpr int qW1 = 0, qW2 = qW1 + kHid * 16, qW3 = qW2 + kHid * kHid, kNQ = qW3 + kHid;

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// round to nearest even
__host__ __device__ inline uint16_t to_bf16(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, sizeof u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float from_bf16(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ inline float bfr(float f) { return from_bf16(to_bf16(f)); }  // value rounded to bf16
__device__ inline float sigm(float z) { return 1.f / (1.f + expf(-z)); }

__device__ inline s4 pack4(float a, float b, float c, float d) {
  s4 r;
  r.x = (short)to_bf16(a);
  r.y = (short)to_bf16(b);
  r.z = (short)to_bf16(c);
  r.w = (short)to_bf16(d);
  return r;
}
__device__ inline s4 load4(const uint16_t* p) {  // 4 consecutive bf16
  const uint2 u = *(const uint2*)p;
  s4 r;
  r.x = (short)(u.x & 0xffffu);
  r.y = (short)(u.x >> 16);
  r.z = (short)(u.y & 0xffffu);
  r.w = (short)(u.y >> 16);
  return r;
}
__device__ inline f4 mfma(s4 a, s4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

// Weights of one wave, in A-operand layout (lane l: row l%16, k = 4(l/16)..+3).
struct WRegs {
  s4 w1[4];      // W1p tile mt
  s4 w2[4][4];   // W2 tile (mt = out block, kt = in block)
  float w3[4][4];  // w3[16 mt + 4 g + j] (bf16 values)
};
__device__ inline void load_weights(const uint16_t* __restrict__ q, int r, int g, WRegs& W) {
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    W.w1[mt] = load4(&q[qW1 + (16 * mt + r) * 16 + 4 * g]);
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) W.w2[mt][kt] = load4(&q[qW2 + (16 * mt + r) * kHid + 16 * kt + 4 * g]);
#pragma unroll
    for (int j = 0; j < 4; ++j) W.w3[mt][j] = from_bf16(q[qW3 + 16 * mt + 4 * g + j]);
  }
}

// Forward pass of one tile.  x4: this lane's B operand of layer 1 (features
// 4g..4g+3 of particle r; zero for g >= 2).  Outputs (lane l, particle l%16):
// z1, z2 pre-activations in accumulator layout (hidden 16 mt + 4g + j), the
// bf16 activations a1 (as B operands), a2 (bf16 values), z3.
struct Fwd {
  f4 z1[4], z2[4];
  s4 a1[4];
  float a2[4][4];
  float z3;
};
__device__ inline void forward(const WRegs& W, const float* __restrict__ b, s4 x4, Fwd& o, int g) {
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    f4 c = {0.f, 0.f, 0.f, 0.f};
    o.z1[mt] = mfma(W.w1[mt], x4, c);
    float a[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float z = o.z1[mt][j] + b[oB1 + 16 * mt + 4 * g + j];
      o.z1[mt][j] = z;
      a[j] = z * sigm(z);
    }
    o.a1[mt] = pack4(a[0], a[1], a[2], a[3]);
  }
  float part = 0.f;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    f4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) c = mfma(W.w2[mt][kt], o.a1[kt], c);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float z = c[j] + b[oB2 + 16 * mt + 4 * g + j];
      c[j] = z;
      const float a = bfr(z * sigm(z));
      o.a2[mt][j] = a;
      part += W.w3[mt][j] * a;
    }
    o.z2[mt] = c;
  }
  // sum over the four lane groups holding the same particle
  part += __shfl_xor(part, 16);
  part += __shfl_xor(part, 32);
  o.z3 = part + b[oB3];
}

// B operand of layer 1 for particle p of a [n][8] feature array (normalised):
// lanes of groups 0 and 1 hold features 4g..4g+3, groups 2 and 3 zeros.
__device__ inline s4 input4(const float* __restrict__ x, int32_t p, int32_t n, int g) {
  if (g >= 2 || p >= n) return pack4(0.f, 0.f, 0.f, 0.f);
  const float4 v = *(const float4*)&x[(int64_t)p * kIn + 4 * g];
  return pack4(v.x, v.y, v.z, v.w);
}

// alpha = scale sigmoid(MLP(x)) for n particles (times alt[p] when mul);
// where flag[p] != 0 the fallback value alt[p] is written instead.  Written
// to out[p] and, when out2 is not null, to out2[map[p]].
__global__ void __launch_bounds__(256) k_mlp_infer(const uint16_t* __restrict__ q, const float* __restrict__ b,
                                                   const float* __restrict__ x, int32_t n,
                                                   const uint8_t* __restrict__ flag, const float* alt,
                                                   float* out, const int32_t* __restrict__ map,
                                                   float* __restrict__ out2, float scale, int mul) {
  const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
  WRegs W;
  load_weights(q, r, g, W);
  const int32_t ntiles = (n + 15) / 16;
  const int32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  for (int32_t t = wave; t < ntiles; t += nwaves) {
    const int32_t p = 16 * t + r;
    Fwd f;
    forward(W, b, input4(x, p, n, g), f, g);
    if (g == 0 && p < n) {
      float a = scale * sigm(f.z3);
      if (mul) a = a * alt[p];
      if (flag && flag[p]) a = alt[p];
      out[p] = a;
      if (out2) out2[map[p]] = a;
    }
  }
}

// Forward and backward over a batch with targets: loss = mean (alpha - t)^2.
// Per-wave partial gradients (and loss) go to part[wave][kNP + 1]; summed by
// k_mlp_reduce in a fixed order (deterministic).
__global__ void __launch_bounds__(256) k_mlp_grad(const uint16_t* __restrict__ q, const float* __restrict__ b,
                                                  const float* __restrict__ x, const float* __restrict__ target,
                                                  int32_t n, float inv_n, float* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) uint16_t lds[4][3][16 * 64 + 16 * 16];  // per wave: dZ2^T, a1^T, dZ1^T, x^T
  const int l = threadIdx.x & 63, r = l & 15, g = l >> 4, w = threadIdx.x >> 6;
  uint16_t* tZ2 = lds[w][0];  // [64 hidden][16 particles]
  uint16_t* tA1 = lds[w][1];
  uint16_t* tZ1 = lds[w][2];
  uint16_t* tX = lds[w][0] + 16 * 64;  // [16 inputs][16 particles]
  WRegs W;
  load_weights(q, r, g, W);
  // W2 in the layout of the A operand of dA1^T = W2^T dZ2^T: A[m = i][k = o] = W2[o][i]
  s4 w2t[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) {
      const uint16_t* c = &q[qW2 + (16 * kt + 4 * g) * kHid + 16 * mt + r];
      s4 v;
      v.x = (short)c[0];
      v.y = (short)c[kHid];
      v.z = (short)c[2 * kHid];
      v.w = (short)c[3 * kHid];
      w2t[mt][kt] = v;
    }
  f4 gW2[4][4], gW1[4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    gW1[mt] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) gW2[mt][nt] = f4{0.f, 0.f, 0.f, 0.f};
  }
  float gb1[4][4], gb2[4][4], gw3[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int j = 0; j < 4; ++j) gb1[mt][j] = gb2[mt][j] = gw3[mt][j] = 0.f;
  float gb3 = 0.f, loss = 0.f;
  const int32_t ntiles = (n + 15) / 16;
  const int32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  for (int32_t t = wave; t < ntiles; t += nwaves) {
    const int32_t p = 16 * t + r;
    const bool valid = p < n;
    const s4 x4 = input4(x, p, n, g);
    Fwd f;
    forward(W, b, x4, f, g);
    const float al = sigm(f.z3);
    const float dz3 = valid ? 2.f * (al - target[p]) * inv_n * al * (1.f - al) : 0.f;
    if (g == 0 && valid) {
      const float e = al - target[p];
      loss += e * e * inv_n;
      gb3 += dz3;
    }
    // layer 3 and 2: dZ2^T = w3 dz3 * silu'(z2)
    s4 dz2b[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      float d[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        gw3[mt][j] += f.a2[mt][j] * dz3;
        const float z = f.z2[mt][j], s = sigm(z);
        d[j] = W.w3[mt][j] * dz3 * (s * (1.f + z * (1.f - s)));
        gb2[mt][j] += d[j];
      }
      dz2b[mt] = pack4(d[0], d[1], d[2], d[3]);
      // tiles in [hidden][particle] order for the transposed reads
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        tZ2[(16 * mt + 4 * g + j) * 16 + r] = (uint16_t)dz2b[mt][j];
        tA1[(16 * mt + 4 * g + j) * 16 + r] = (uint16_t)f.a1[mt][j];
      }
    }
    // dA1^T = W2^T dZ2^T, dZ1^T = dA1^T * silu'(z1)
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      f4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) c = mfma(w2t[mt][kt], dz2b[kt], c);
      float d[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float z = f.z1[mt][j], s = sigm(z);
        d[j] = c[j] * (s * (1.f + z * (1.f - s)));
        gb1[mt][j] += d[j];
      }
      const s4 db = pack4(d[0], d[1], d[2], d[3]);
#pragma unroll
      for (int j = 0; j < 4; ++j) tZ1[(16 * mt + 4 * g + j) * 16 + r] = (uint16_t)db[j];
    }
    // x^T tile [input][particle] (inputs 8..15 zero)
#pragma unroll
    for (int j = 0; j < 4; ++j) tX[(4 * g + j) * 16 + r] = (uint16_t)x4[j];
    // the tiles were written by other lanes of this wave
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // weight gradients: dW2[o][i] += sum_p dZ2[o][p] a1[i][p], dW1[o][i] += sum_p dZ1[o][p] x[i][p]
    // A operand (lane: row r, k = particles 4g..+3) and B operand (lane: column r, k = particles 4g..+3)
    s4 bA1[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) bA1[nt] = load4(&tA1[(16 * nt + r) * 16 + 4 * g]);
    const s4 bX = load4(&tX[r * 16 + 4 * g]);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      const s4 aZ2 = load4(&tZ2[(16 * mt + r) * 16 + 4 * g]);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) gW2[mt][nt] = mfma(aZ2, bA1[nt], gW2[mt][nt]);
      const s4 aZ1 = load4(&tZ1[(16 * mt + r) * 16 + 4 * g]);
      gW1[mt] = mfma(aZ1, bX, gW1[mt]);
    }
    // the next tile overwrites them
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  // per-wave partials
  float* P = part + (int64_t)wave * (kNP + 1);
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int o = 16 * mt + 4 * g + j;
      // accumulator layout: row o, column r
      if (r < kIn) P[oW1 + o * kIn + r] = gW1[mt][j];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) P[oW2 + o * kHid + 16 * nt + r] = gW2[mt][nt][j];
      // sums over the 16 particles (lanes r) of the row
      float s1 = gb1[mt][j], s2 = gb2[mt][j], s3 = gw3[mt][j];
#pragma unroll
      for (int m = 1; m < 16; m <<= 1) {
        s1 += __shfl_xor(s1, m);
        s2 += __shfl_xor(s2, m);
        s3 += __shfl_xor(s3, m);
      }
      if (r == 0) {
        P[oB1 + o] = s1;
        P[oB2 + o] = s2;
        P[oW3 + o] = s3;
      }
    }
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) {
    gb3 += __shfl_xor(gb3, m);
    loss += __shfl_xor(loss, m);
  }
  if (l == 0) {
    P[oB3] = gb3;
    P[kNP] = loss;
  }
}

__global__ void k_mlp_reduce(const float* __restrict__ part, int32_t nwaves, float* __restrict__ grad) {
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > kNP) return;
  float s = 0.f;
  for (int32_t w = 0; w < nwaves; ++w) s += part[(int64_t)w * (kNP + 1) + c];
  grad[c] = s;
}

__global__ void k_mlp_pack(const float* __restrict__ p, uint16_t* __restrict__ q) {
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= kNQ) return;
  float v = 0.f;
  if (c < qW2) {
    const int o = c / 16, i = c % 16;
    v = i < kIn ? p[oW1 + o * kIn + i] : 0.f;
  } else if (c < qW3) {
    v = p[oW2 + (c - qW2)];
  } else {
    v = p[oW3 + (c - qW3)];
  }
  q[c] = to_bf16(v);
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters
__global__ void k_mlp_adam(int32_t np, float* __restrict__ p, float* __restrict__ m1, float* __restrict__ m2,
                           const float* __restrict__ grad, float lr, float b1, float b2, float eps, float c1,
                           float c2) {
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= np) return;
  const float gr = grad[c];
  const float m = b1 * m1[c] + (1.f - b1) * gr;
  const float v = b2 * m2[c] + (1.f - b2) * gr * gr;
  m1[c] = m;
  m2[c] = v;
  p[c] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

inline unsigned mlp_blocks(int32_t n, int32_t cap) {
  const int32_t tiles = (n + 15) / 16;
  int32_t b = (tiles + 3) / 4;
  b = b < 1 ? 1 : b;
  return (unsigned)(b < cap ? b : cap);
}

int mlp_infer_launch(const CrkMlp* m, int32_t n, const float* x, const uint8_t* flag, const float* alt, float* out,
                     const int32_t* map, float* out2, hipStream_t s, bool raw) {
  const float scale = raw ? 1.f : m->scale;
  const int mul = raw ? 0 : (m->multiply != 0);
  CRK_REQUIRE(!mul || alt, "the multiplying switch needs the Cullen-Dehnen alpha");
  hipLaunchKernelGGL(k_mlp_infer, dim3(mlp_blocks(n, 4096)), dim3(256), 0, s, m->wbf16, m->params, x, n, flag, alt,
                     out, map, out2, scale, mul);
  CRK_LAUNCH_CHECK("k_mlp_infer");
  return CRK_OK;
}

}  // namespace crk

using namespace crk;

extern "C" {

int crk_mlp_pack(const float* params, uint16_t* wbf16, void* stream) {
  CRK_REQUIRE(params && wbf16, "crk_mlp_pack: null argument");
  hipLaunchKernelGGL(k_mlp_pack, dim3(nblocks(kNQ)), dim3(kBlock), 0, (hipStream_t)stream, params, wbf16);
  CRK_LAUNCH_CHECK("k_mlp_pack");
  return CRK_OK;
}

int crk_mlp_forward(const CrkMlp* m, int32_t n, const float* x, float* alpha, void* stream) {
  CRK_REQUIRE(m && m->params && m->wbf16 && x && alpha && n >= 0, "crk_mlp_forward: invalid arguments");
  if (n == 0) return CRK_OK;
  return mlp_infer_launch(m, n, x, nullptr, nullptr, alpha, nullptr, nullptr, (hipStream_t)stream, true);
}

int crk_mlp_grad(CrkContext* ctx, const CrkMlp* m, int32_t n, const float* x, const float* target, float* grad,
                 void* stream) {
  CRK_REQUIRE(ctx && m && m->params && m->wbf16 && x && target && grad && n > 0, "crk_mlp_grad: invalid arguments");
  hipStream_t s = (hipStream_t)stream;
  const unsigned nb = mlp_blocks(n, 256);
  const int32_t nwaves = (int32_t)nb * 4;
  int st = ensure(ctx->s_counts, sizeof(float) * (size_t)nwaves * (kNP + 1));
  if (st) return st;
  float* part = (float*)ctx->s_counts.p;
  hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(256), 0, s, m->wbf16, m->params, x, target, n, 1.f / (float)n, part);
  CRK_LAUNCH_CHECK("k_mlp_grad");
  hipLaunchKernelGGL(k_mlp_reduce, dim3(nblocks(kNP + 1)), dim3(kBlock), 0, s, part, nwaves, grad);
  CRK_LAUNCH_CHECK("k_mlp_reduce");
  return CRK_OK;
}

int crk_mlp_adam(int32_t n_param, float* params, float* m1, float* m2, const float* grad, float lr, float beta1,
                 float beta2, float eps, int32_t step, void* stream) {
  CRK_REQUIRE(params && m1 && m2 && grad && n_param > 0 && step >= 1, "crk_mlp_adam: invalid arguments");
  const float c1 = (float)(1.0 - std::pow((double)beta1, step));
  const float c2 = (float)(1.0 - std::pow((double)beta2, step));
  hipLaunchKernelGGL(k_mlp_adam, dim3(nblocks(n_param)), dim3(kBlock), 0, (hipStream_t)stream, n_param, params, m1,
                     m2, grad, lr, beta1, beta2, eps, c1, c2);
  CRK_LAUNCH_CHECK("k_mlp_adam");
  return CRK_OK;
}

}  // extern "C"
