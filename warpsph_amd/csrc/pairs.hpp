// pairs.hpp -- device machinery shared by the pair passes of one CRKSPH
// derivative evaluation: lane groups, compensated sums, kernel functions,
// the CRK-corrected gradient, the minimum image, and the launch dispatch.
//
// Work mapping: a group of G consecutive lanes of a wave64 owns one sorted
// particle row; lane g handles the row's pairs g, g + G, ... so pair arrays
// (neighbour ids, ap, av) stream coalesced.  Each lane keeps
// Kahan-compensated FP32 partial sums; the G partials are combined by a
// compensated FP32 butterfly (deterministic, every lane ends with the same
// value).
#pragma once

#include "common.hpp"

namespace crk {

// ------------------------------------------------------------ fast math
// v_rcp_f32 / v_sqrt_f32 (about 1 ulp, no denormal rescaling) for the pair
// terms; none of their operands (distances, supports, limiter ratios) is
// denormal in a physical configuration.
__device__ inline float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ inline float inv_or_zero(float r) { return r > 0.f ? frcp(r) : 0.f; }

// x_i - x_j with the minimum image (the neighbour search's convention: a
// periodic component beyond +-L/2 is shifted by L), branch free.
template <int D>
__device__ inline void pair_dx(const Geo& g, const float* xi, const float* xj, float* dx) {
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const float t = xi[d] - xj[d];
    float s = t > g.wrap[d] ? g.L[d] : 0.f;  // two selects (a nested one compiles to branches)
    s = t < -g.wrap[d] ? -g.L[d] : s;
    dx[d] = t - s;
  }
}

// D-vector store with one instruction for D = 2 (8-byte aligned rows)
template <int D>
__device__ inline void store_vec(float* p, const float* v) {
  if (D == 2) {
    *reinterpret_cast<float2*>(p) = make_float2(v[0], v[1]);
  } else {
#pragma unroll
    for (int a = 0; a < D; ++a) p[a] = v[a];
  }
}

// ------------------------------------------------------------ neighbour records
// Every per-particle field that the acceleration pass gathers from its
// neighbours, packed per sorted particle into NF floats (whole float4s): a
// neighbour then costs NF/4 16-byte loads from one or two cache lines
// instead of one load per field array.  The viscosity's density and sound
// speed enter only as Q = mu (qa + qb mu) with qa = -C_l rho c and
// qb = C_q rho, and the switch's divergence only as m sgn(div): the record
// holds those products.  The lighter passes gather a few
// narrow fields and stay on the SoA arrays (a wide record would touch a
// cache line per neighbour for a few bytes).
template <int D>
struct Rec;
template <>
struct Rec<1> {
  static constexpr int NF = 16, X = 0, H = 1, V = 2, VEL = 3, QA = 4, QB = 5, J = 6, A = 7, B = 8, GA = 9, GB = 10,
                       P = 11, M = 12, MSG = 13;
};
template <>
struct Rec<2> {
  static constexpr int NF = 24, X = 0, H = 2, V = 3, VEL = 4, QA = 6, QB = 7, J = 8, A = 12, B = 13, GA = 15,
                       GB = 17, P = 21, M = 22, MSG = 23;
};
template <>
struct Rec<3> {
  static constexpr int NF = 40, X = 0, H = 3, V = 4, VEL = 5, QA = 8, QB = 9, P = 10, M = 11, MSG = 12, A = 13,
                       B = 14, GA = 17, GB = 20, J = 29;
};
// the whole record of particle j (loads of fields the caller does not use
// are removed by the compiler)
template <int D>
__device__ inline void rec_load(const float4* __restrict__ rec, int32_t j, float* r) {
  const float4* p = rec + (int64_t)j * (Rec<D>::NF / 4);
#pragma unroll
  for (int c = 0; c < Rec<D>::NF / 4; ++c) {
    const float4 t = p[c];
    r[4 * c] = t.x;
    r[4 * c + 1] = t.y;
    r[4 * c + 2] = t.z;
    r[4 * c + 3] = t.w;
  }
}

// ------------------------------------------------------------ lane groups
struct Comp {  // compensated value s - c
  float s = 0.f, c = 0.f;
  __device__ inline void add(float y) {
    float yy = y - c;
    float t = s + yy;
    c = (t - s) - yy;
    s = t;
  }
};

// Sum of the G lanes' compensated partials s - c, in FP32: an xor butterfly
// whose step adds two partials (s1 - c1) + (s2 - c2) as s = s1 + s2 with the
// rounding error e of that addition recovered exactly (TwoSum, Knuth) and
// folded into the compensation, c = (c1 + c2) - e.  Both the sum and the
// exact error are symmetric in the two operands, so every lane of the group
// ends with the same value.
template <int G>
__device__ inline Comp group_comp(Comp a) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) {
    const float s2 = __shfl_xor(a.s, off, G), c2 = __shfl_xor(a.c, off, G);
    const float t = a.s + s2;
    const float bb = t - a.s;
    const float e = (a.s - (t - bb)) + (s2 - bb);
    a.c = (a.c + c2) - e;
    a.s = t;
  }
  return a;
}
template <int G>
__device__ inline float group_total(Comp a) {
  a = group_comp<G>(a);
  return a.s - a.c;
}
// the same total formed in FP64 (input of the FP64 moment solve)
template <int G>
__device__ inline double group_total_d(Comp a) {
  a = group_comp<G>(a);
  return (double)a.s - (double)a.c;
}

// A plain FP32 partial sum, for the sums of the support volume, the CRK
// moments and the velocity gradient: a lane adds ~10-40 terms, and against
// the FP64 oracle these passes' results are as accurate without
// compensation (the Owen sum, which sets h and through it the CFL step, and
// the acceleration and energy sums, whose terms cancel to a small net, stay
// compensated).  Lane partials are still combined with TwoSum.
struct Sum {
  float s = 0.f;
  // s + y, never contracted with the product that forms y, and s + a b as
  // one explicit FMA: every pass variant rounds the same way
  __device__ inline void add(float y) {
#pragma clang fp contract(off)
    s = s + y;
  }
  __device__ inline void fma(float a, float b) { s = fmaf(a, b, s); }
};
template <int G>
__device__ inline float group_total(Sum a) {
  Comp c;
  c.s = a.s;
  return group_total<G>(c);
}
template <int G>
__device__ inline double group_total_d(Sum a) {
  Comp c;
  c.s = a.s;
  return group_total_d<G>(c);
}

template <int G>
__device__ inline float group_max(float v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, G));
  return v;
}

struct Lane {
  int g;       // lane inside the group
  int32_t k;   // row (sorted particle)
  bool valid;
};
// Blocks are dispatched to the 8 XCDs round-robin (block b -> XCD b % 8).
// Renumber them so that every XCD works on one contiguous range of sorted
// rows: its L2 then holds one compact region of the domain.
__device__ inline int32_t xcd_block() {
  const int32_t nb = gridDim.x, b = blockIdx.x;
  const int32_t q = nb / 8, r = nb % 8, x = b % 8;
  return x * q + (x < r ? x : r) + b / 8;
}

template <int G>
__device__ inline Lane lane_of(int32_t n) {
  Lane l;
  l.g = threadIdx.x % G;
  l.k = xcd_block() * (int32_t)(blockDim.x / G) + (int32_t)(threadIdx.x / G);
  l.valid = l.k < n;
  if (!l.valid) l.k = n - 1;  // safe index for loads; results discarded
  return l;
}

// ------------------------------------------------------------ kernel functions
// k(q) and dk/dq for 0 <= q < 1 (Wendland C2, C4, C6 as K = 0, 1, 2).
template <int K, int D>
__device__ inline void kwd(float q, float& k, float& dk) {
  if (K == CRK_KERNEL_B7) { k = b7_k(q); dk = b7_dk(q); return; }
  float t = 1.f - q;
  float t2 = t * t;
  if (K == CRK_KERNEL_WENDLAND2) {
    if (D == 1) { k = t2 * t * (1.f + 3.f * q); dk = -12.f * q * t2; }
    else { float t3 = t2 * t; k = t3 * t * (1.f + 4.f * q); dk = -20.f * q * t3; }
  } else if (K == CRK_KERNEL_WENDLAND4) {
    float t4 = t2 * t2;
    if (D == 1) { k = t4 * t * (1.f + 5.f * q + 8.f * q * q); dk = -14.f * q * t4 * (1.f + 4.f * q); }
    else { float t5 = t4 * t; k = t5 * t * (1.f + 6.f * q + (35.f / 3.f) * q * q); dk = (-56.f / 3.f) * q * t5 * (1.f + 5.f * q); }
  } else {
    float t4 = t2 * t2;
    if (D == 1) { float t6 = t4 * t2; k = t6 * t * (1.f + 7.f * q + 19.f * q * q + 21.f * q * q * q); dk = -6.f * q * t6 * (3.f + 18.f * q + 35.f * q * q); }
    else { float t7 = t4 * t2 * t; k = t7 * t * (1.f + 8.f * q + 25.f * q * q + 32.f * q * q * q); dk = -22.f * q * t7 * (1.f + 7.f * q + 16.f * q * q); }
  }
}

// k(q) and (dk/dq) / q: every Wendland dk/dq carries a factor q, so the
// gradient's dW/dr / r needs no 1/r (and stays finite at r = 0, where it
// multiplies x_ij = 0)
template <int K, int D>
__device__ inline void kwq(float q, float& k, float& dkq) {
  if (K == CRK_KERNEL_B7) {
    k = b7_k(q);
    dkq = q < 0.25f ? b7_dkq_inner(q) : b7_dk_outer(q) * frcp(q);
    return;
  }
  float t = 1.f - q;
  float t2 = t * t;
  if (K == CRK_KERNEL_WENDLAND2) {
    if (D == 1) { k = t2 * t * (1.f + 3.f * q); dkq = -12.f * t2; }
    else { float t3 = t2 * t; k = t3 * t * (1.f + 4.f * q); dkq = -20.f * t3; }
  } else if (K == CRK_KERNEL_WENDLAND4) {
    float t4 = t2 * t2;
    if (D == 1) { k = t4 * t * (1.f + 5.f * q + 8.f * q * q); dkq = -14.f * t4 * (1.f + 4.f * q); }
    else { float t5 = t4 * t; k = t5 * t * (1.f + 6.f * q + (35.f / 3.f) * q * q); dkq = (-56.f / 3.f) * t5 * (1.f + 5.f * q); }
  } else {
    float t4 = t2 * t2;
    if (D == 1) { float t6 = t4 * t2; k = t6 * t * (1.f + 7.f * q + 19.f * q * q + 21.f * q * q * q); dkq = -6.f * t6 * (3.f + 18.f * q + 35.f * q * q); }
    else { float t7 = t4 * t2 * t; k = t7 * t * (1.f + 8.f * q + 25.f * q * q + 32.f * q * q * q); dkq = -22.f * t7 * (1.f + 7.f * q + 16.f * q * q); }
  }
}

template <int D>
__device__ inline float ipow(float x) {
  float r = x;
#pragma unroll
  for (int k = 1; k < D; ++k) r = r * x;
  return r;
}

// W(r, h) and dW/dr / r (gradient = fac * x_ab) with 1/h and 1/r given.
// q = r / h formed as r * (1/h) and rounded on its own: never contracted into
// what the kernel polynomial computes from it (1 - q, ...), so that every
// pass variant, whatever its control flow, evaluates the same kernel values
__device__ inline float qmul(float r, float invh) {
#pragma clang fp contract(off)
  return r * invh;
}

struct WdW {
  float W, fac;
};
template <int K, int D>
__device__ inline WdW wdw(float C, float r, float invh) {
  WdW o;
  float q = qmul(r, invh);
  if (K == CRK_KERNEL_B7) {
    if (!(q < 1.f)) { o.W = 0.f; o.fac = 0.f; return o; }
  } else {
    // the Wendland polynomials carry a factor (1 - q)^n: at q = 1 the value
    // and the derivative are exactly zero, so q is clamped instead of branched
    q = fminf(q, 1.f);
  }
  float k, dkq;
  kwq<K, D>(q, k, dkq);
  const float CihD = C * ipow<D>(invh);
  o.W = CihD * k;
  o.fac = CihD * dkq * (invh * invh);  // (dW/dr) / r = C h^-d (dk/dq / q) / h^2
  return o;
}

// CRK-corrected gradient d_g W^R_ab (Frontiere, Raskin & Owen 2017):
// gA (1 + B.x) W + A (B_g + gB^T x) W + A (1 + B.x) gradW
template <int D>
struct CorrD {
  float A, B[D], gA[D], gB[D * D];
};
template <int D>
__device__ inline CorrD<D> load_corr(const float* Ap, const float* Bp, const float* gAp, const float* gBp, int32_t k) {
  CorrD<D> c;
  c.A = Ap[k];
#pragma unroll
  for (int a = 0; a < D; ++a) { c.B[a] = Bp[(int64_t)k * D + a]; c.gA[a] = gAp[(int64_t)k * D + a]; }
#pragma unroll
  for (int a = 0; a < D * D; ++a) c.gB[a] = gBp[(int64_t)k * D * D + a];
  return c;
}
template <int D>
__device__ inline void grad_wr(const CorrD<D>& cr, const float* x, const WdW& w, float* out) {
  float bx = 1.f;
#pragma unroll
  for (int a = 0; a < D; ++a) bx = bx + cr.B[a] * x[a];
  float AW = cr.A * w.W;
  float ABf = cr.A * bx * w.fac;
  float bxW = bx * w.W;
#pragma unroll
  for (int gm = 0; gm < D; ++gm) {
    float gbx = cr.B[gm];
#pragma unroll
    for (int a = 0; a < D; ++a) gbx = gbx + cr.gB[a * D + gm] * x[a];
    out[gm] = cr.gA[gm] * bxW + AW * gbx + ABf * x[gm];
  }
}

// ------------------------------------------------------------ energy partition
// Fraction f_ij of the pair work dE = m_i dt (v_half_j - v_half_i).a_ij given
// to particle i, CRK energy scheme (compSPH/balance.py:126-154): s = P/rho^gamma;
// equal s -> 1/2; otherwise the pair work goes preferentially to the particle
// that moves s towards the other's.
__device__ inline float energy_fraction(float dE, float si, float sj) {
  float asi = fabsf(si), asj = fabsf(sj);
  float smin = fminf(asi, asj), smax = fmaxf(asi, asj);
  bool mB = (dE >= 0.f && si >= sj) || (dE < 0.f && si < sj);
  bool mC = (dE >= 0.f && si < sj) || (dE < 0.f && si >= sj);
  float ff = 0.5f;
  if (!(si - sj == 0.f)) {
    float rs = frcp(smin + smax + 1.0e-14f);
    if (mB) ff = smin * rs;
    else if (mC) ff = smax * rs;
  }
  return fmaxf(0.f, fminf(1.f, ff));
}
// f_ij from the stage record rec = {v_half (D components), ..., s} of i and j
template <int D>
__device__ inline float pair_fraction(const float4& ri, const float4& rj, float mdt, const float* a) {
  const float vi[3] = {ri.x, ri.y, ri.z}, vj[3] = {rj.x, rj.y, rj.z};
  float d = 0.f;
#pragma unroll
  for (int c = 0; c < D; ++c) d = d + (vj[c] - vi[c]) * a[c];
  return energy_fraction(mdt * d, ri.w, rj.w);
}

// ------------------------------------------------------------ rows
// Full Verlet rows of the bound adjacency.
struct Rows {
  const int32_t* off;
  const int32_t* nbr;
};
// In-support pairs of each row (|x_ij| < max(h_i, h_j)), compacted in row
// order into the slots [off[k], off[k] + cnt[k]) of nbr; the pair data (ap,
// av, f) live in the same slots.  About half of a Verlet row is skin; the
// kernel-weighted passes iterate only these pairs, so every lane of a group
// works on a contributing pair.
struct ARows {
  const int32_t* off;
  const int32_t* cnt;
  const int32_t* nbr;
};

inline Rows rows_of(const CrkContext* ctx) { return Rows{ctx->adj.row_offsets, ctx->adj.nbr}; }
inline ARows arows_of(const CrkContext* ctx) { return ARows{ctx->adj.row_offsets, ctx->pair_cnt, ctx->pair_nbr}; }
inline unsigned group_blocks(int32_t n, int G) { return nblocks((int64_t)n * G); }

template <class T>
inline T* F(Buf& b) { return (T*)b.p; }

int require_bound(CrkContext* ctx, const CrkConfig* cfg);

// the MLP of the learned switch (mlp.hip): alpha into out[p] (sorted) and
// out2[map[p]]; flag[p] selects alt[p] instead; raw: the sigmoid output
// alone (no scale, no Cullen-Dehnen factor)
int mlp_infer_launch(const CrkMlp* m, int32_t n, const float* x, const uint8_t* flag, const float* alt, float* out,
                     const int32_t* map, float* out2, hipStream_t s, bool raw);

}  // namespace crk

// ------------------------------------------------------------ dispatch
// over the kernel function (KK), the dimension (DD) and the group size (G)
#define CRK_G_CASE(GG, ...) case GG: { constexpr int G = GG; __VA_ARGS__; } break;
#define CRK_G_SWITCH(GS, ...)                                                                  \
  switch (GS) {                                                                                \
    CRK_G_CASE(1, __VA_ARGS__) CRK_G_CASE(2, __VA_ARGS__) CRK_G_CASE(4, __VA_ARGS__)           \
    CRK_G_CASE(8, __VA_ARGS__) CRK_G_CASE(16, __VA_ARGS__) CRK_G_CASE(32, __VA_ARGS__)         \
    CRK_G_CASE(64, __VA_ARGS__)                                                                \
    default: ::crk::set_error("bad group size %d", GS); return CRK_EINVAL;                     \
  }
#define CRK_D_CASE(DV, ...) case DV: { constexpr int DD = DV; __VA_ARGS__; } break;
#define CRK_DIM_G_DISPATCH(DIM, GS, ...)                                                       \
  switch (DIM) {                                                                               \
    CRK_D_CASE(1, CRK_G_SWITCH(GS, __VA_ARGS__))                                               \
    CRK_D_CASE(2, CRK_G_SWITCH(GS, __VA_ARGS__))                                               \
    CRK_D_CASE(3, CRK_G_SWITCH(GS, __VA_ARGS__))                                               \
    default: ::crk::set_error("dim must be 1..3"); return CRK_EINVAL;                          \
  }
#define CRK_K_CASE(KV, ...) case KV: { constexpr int KK = KV; __VA_ARGS__; } break;
#define CRK_KDG_DISPATCH(KERN, DIM, GS, ...)                                                   \
  switch (KERN) {                                                                              \
    CRK_K_CASE(0, CRK_DIM_G_DISPATCH(DIM, GS, __VA_ARGS__))                                    \
    CRK_K_CASE(1, CRK_DIM_G_DISPATCH(DIM, GS, __VA_ARGS__))                                    \
    CRK_K_CASE(2, CRK_DIM_G_DISPATCH(DIM, GS, __VA_ARGS__))                                    \
    CRK_K_CASE(3, CRK_DIM_G_DISPATCH(DIM, GS, __VA_ARGS__))                                    \
    default: ::crk::set_error("unknown kernel %d", KERN); return CRK_EINVAL;                   \
  }
#define CRK_DIM_DISPATCH(D, ...)                                                               \
  switch (D) {                                                                                 \
    CRK_D_CASE(1, __VA_ARGS__) CRK_D_CASE(2, __VA_ARGS__) CRK_D_CASE(3, __VA_ARGS__)           \
    default: ::crk::set_error("dim must be 1..3"); return CRK_EINVAL;                          \
  }
