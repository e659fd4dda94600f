// integrate.hip -- RK2 helpers, the compatible-energy finalize and the
// reductions of the CRKSPH step (CFL dt, FP64 conservation sums, min/max).
#include <cmath>
#include <cstring>

#include "pairs.hpp"
#include "tiles.hpp"

namespace crk {

__global__ void k_axpy(int64_t cnt, const float* __restrict__ y0, const float* __restrict__ d, float dt,
                       float* __restrict__ y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) y[i] = y0[i] + dt * d[i];
}

__global__ void k_combine(int64_t cnt, const float* __restrict__ y0, const float* __restrict__ d0,
                          const float* __restrict__ d1, float b0, float b1, float dt, float* __restrict__ y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) y[i] = y0[i] + dt * (b0 * d0[i] + b1 * d1[i]);
}

// the same with a step per particle: element e belongs to particle e / comps
__global__ void k_axpy_p(int64_t cnt, int32_t comps, const float* __restrict__ y0, const float* __restrict__ d,
                         const float* __restrict__ dtp, float* __restrict__ y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) y[i] = y0[i] + dtp[i / comps] * d[i];
}
__global__ void k_combine_p(int64_t cnt, int32_t comps, const float* __restrict__ y0, const float* __restrict__ d0,
                            const float* __restrict__ d1, float b0, float b1, const float* __restrict__ dtp,
                            float* __restrict__ y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) y[i] = y0[i] + dtp[i / comps] * (b0 * d0[i] + b1 * d1[i]);
}

__global__ void k_wrap(int32_t n, int D, Geo g, float lo0, float lo1, float lo2, float hi0, float hi1, float hi2,
                       float* __restrict__ x) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
  for (int d = 0; d < D; ++d) {
    if (!g.per[d]) continue;
    float& xx = x[(int64_t)i * D + d];
    if (xx >= hi[d]) xx = xx - g.L[d];
    else if (xx < lo[d]) xx = xx + g.L[d];
  }
}

// v_half = v0 + 0.5 * ((0 + b0 dvdt0) + b1 dvdt1) * dt  (multistep.py:63-68),
// written in an adjacency's sorted order
template <int D>
__global__ void k_vhalf_sorted(const int32_t* __restrict__ perm, int32_t n, const float* __restrict__ v0,
                               const float* __restrict__ a0, const float* __restrict__ a1, float b0, float b1,
                               DtArg dta, float* __restrict__ vhs) {
  int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int64_t i = perm[k];
  const float dt = dta.at(k);
#pragma unroll
  for (int d = 0; d < D; ++d) {
    float fa = 0.f + b0 * a0[i * D + d];
    fa = fa + b1 * a1[i * D + d];
    vhs[(int64_t)k * D + d] = v0[i * D + d] + 0.5f * fa * dt;
  }
}

// one stage's in-support pairs (slots [off[k], off[k] + cnt[k])), pair data
// and energy-partition records
struct CompIn {
  const int32_t *off, *cnt, *nbr;
  const float* pf;  // per slot: a = ap + av (D)
  const float4* rec;
  float b;
  const uint16_t* loc;  // tile-local neighbour index per slot (tiled adjacencies), or null
};

template <int G>
__device__ inline float group_total2(float s, float c) {
  Comp a;
  a.s = s;
  a.c = c;
  return group_total<G>(a);
}

__device__ inline void kahan_add(float& s, float& c, float x) {
  float y = x - c;
  float t = s + y;
  c = (t - s) - y;
  s = t;
}

// Neighbour source of the finalize: a stage's energy-partition records and
// the step's v_half through the sorted ids.
template <int D>
struct CompSrcG {
  const int32_t* __restrict__ nbr;
  const float4* __restrict__ rec;
  const float* __restrict__ vhs;
  __device__ inline int32_t id(int64_t q) const { return nbr[q]; }
  __device__ inline float4 r(int32_t j) const { return rec[j]; }
  __device__ inline float vh(int32_t j, int c) const { return vhs[(int64_t)j * D + c]; }
};

// The same from a tile's LDS copies (the stage's records and v_half of the
// particles the tile's rows reach) through the 16-bit local indices.
template <int D>
struct CompSrcT {
  const uint16_t* sid;  // the stage's 16-bit indices (from entry base)
  int32_t base;
  const float4* rec;  // LDS
  const float* vhs;   // LDS, D per staged particle
  __device__ inline int32_t id(int64_t q) const { return sid[q - base]; }
  __device__ inline float4 r(int32_t j) const { return rec[j]; }
  __device__ inline float vh(int32_t j, int c) const { return vhs[j * D + c]; }
};

// The sum over one stage's pairs of row k: b f_ij (v_half_j - v_half_i).a_ij
// with f_ij from the stage's records (the energy partition of crk_balance).
template <int D, int G, class Src>
__device__ inline void compat_sum(const CompIn& s, const Src& src, int32_t k, bool valid, int g, const float* vi,
                                  float mdt, float& sa, float& ca) {
  const int32_t b = s.off[k], e = valid ? b + s.cnt[k] : b;
  const float4 ri = s.rec[k];
  // software pipeline: the neighbour's records and the pair force one
  // iteration ahead, the neighbour's id two ahead
  int32_t jc = b + g < e ? src.id(b + g) : 0;
  int32_t jn = b + g + G < e ? src.id(b + g + G) : 0;
  float4 rn = src.r(jc);
  float vn[3], an[3];
#pragma unroll
  for (int c = 0; c < D; ++c) vn[c] = src.vh(jc, c);
  auto pf_load = [&](int64_t q, float* a) {
    if (D == 2) {
      const float2 p = reinterpret_cast<const float2*>(s.pf)[q];
      a[0] = p.x;
      a[1] = p.y;
    } else {
#pragma unroll
      for (int c = 0; c < D; ++c) a[c] = s.pf[q * D + c];
    }
  };
  if (b + g < e) pf_load(b + g, an);
  for (int32_t q = b + g; q < e; q += G) {
    const float4 rj = rn;
    float vj[3], a[3];
#pragma unroll
    for (int c = 0; c < D; ++c) {
      vj[c] = vn[c];
      a[c] = an[c];
    }
    if (q + G < e) {
      jc = jn;
      rn = src.r(jc);
#pragma unroll
      for (int c = 0; c < D; ++c) vn[c] = src.vh(jc, c);
      pf_load(q + G, an);
      if (q + 2 * G < e) jn = src.id(q + 2 * G);
    }
    float d0 = 0.f;
#pragma unroll
    for (int c = 0; c < D; ++c) d0 = d0 + (vj[c] - vi[c]) * a[c];
    const float f = pair_fraction<D>(ri, rj, mdt, a);
    kahan_add(sa, ca, s.b * f * d0);
  }
}

template <int D, int G, int NST, int MODE, class S0, class S1>
__device__ inline void compat_row(const Lane& L, const S0& src0, const S1& src1, const int32_t* __restrict__ perm,
                                  const float* __restrict__ vhs, const CompIn& s0, const CompIn& s1,
                                  const float* __restrict__ m_u, float* __restrict__ du0, const float* __restrict__ u0,
                                  DtArg dta, float* __restrict__ u, float* __restrict__ dudt_last) {
  const int g = L.g;
  const int32_t k = L.k;
  const bool valid = L.valid;
  float vi[3];
#pragma unroll
  for (int d = 0; d < D; ++d) vi[d] = vhs[(int64_t)k * D + d];
  const int32_t i = perm[k];
  const float dt = dta.at(k);
  const float mdt = m_u[i] * dt;
  float sa = 0.f, ca = 0.f, sb = 0.f, cb = 0.f;
  compat_sum<D, G>(s0, src0, k, valid, g, vi, mdt, sa, ca);
  if (NST == 2) compat_sum<D, G>(s1, src1, k, valid, g, vi, mdt, sb, cb);
  float ta = group_total2<G>(sa, ca);
  float tb = NST == 2 ? group_total2<G>(sb, cb) : 0.f;
  if (!valid || g != 0) return;
  if (MODE == 0) {
    du0[i] = 0.f + ta;
  } else {
    float du = MODE == 1 ? du0[i] + ta : (0.f + ta) + tb;
    float un = u0[i] + du * dt;
    u[i] = un;
    dudt_last[i] = (un - u0[i]) / dt;
  }
}

// Row sums of b_k f (v_half_j - v_half_i).(ap + av) over the in-support pairs
// of each row (multistep.py:70-94; the skin pairs of the reference's Verlet
// list carry ap = av = 0), G lanes per row.  NST = 2: both stages use the
// same sorted order and are summed in one sweep (separate sums, added in
// stage order).  MODE 0: first of two sweeps (writes du0); MODE 1: last
// sweep (u = u0 + (du0 + sum) dt, dudt = (u - u0)/dt); MODE 2: single sweep
// of both stages.
template <int D, int G, int NST, int MODE>
__global__ void __launch_bounds__(kBlock) k_compat(const int32_t* __restrict__ perm, int32_t n,
                                                   const float* __restrict__ vhs, CompIn s0, CompIn s1,
                                                   const float* __restrict__ m_u, float* __restrict__ du0,
                                                   const float* __restrict__ u0, DtArg dta, float* __restrict__ u,
                                                   float* __restrict__ dudt_last) {
  const Lane L = lane_of<G>(n);  // XCD-aware row order
  compat_row<D, G, NST, MODE>(L, CompSrcG<D>{s0.nbr, s0.rec, vhs}, CompSrcG<D>{s1.nbr, s1.rec, vhs}, perm, vhs, s0,
                              s1, m_u, du0, u0, dta, u, dudt_last);
}

// The single sweep of both stages (MODE 2, shared sorted order) on LDS
// tiles (d <= 2): one block per tile stages both stages' records and
// v_half of every particle the tile's rows reach; wide tiles gather.
template <int D, int G>
__global__ void __launch_bounds__(kTileThreadsMax) k_compat_t(const int32_t* __restrict__ tiles,
                                                              const int32_t* __restrict__ perm, int32_t n,
                                                              const float* __restrict__ vhs, CompIn s0, CompIn s1,
                                                              const float* __restrict__ m_u, float* __restrict__ du0,
                                                              const float* __restrict__ u0, DtArg dta,
                                                              float* __restrict__ u, float* __restrict__ dudt_last) {
  static_assert(D <= 2, "tiled finalize: d <= 2");
  __shared__ int32_t tab[kTileWords];
  __shared__ float4 sr0[kTileCap];
  __shared__ float4 sr1[kTileCap];
  __shared__ float svh[kTileCap * D];
  // (the pair indices stay in global memory here: staging both stages' in
  // LDS was measured slower, the LDS then limiting the pass to one block per CU)
  const int32_t t = xcd_block();
  struct PV {
    float4 a, b;
    float vh[D];
  };
  tile_stage_fn<2, 0, PV>(
      tiles, t, 0, 0, nullptr, tab, nullptr,
      [&](int32_t g) {
        PV v;
        v.a = s0.rec[g];
        v.b = s1.rec[g];
#pragma unroll
        for (int c = 0; c < D; ++c) v.vh[c] = vhs[(int64_t)g * D + c];
        return v;
      },
      [&](int32_t u, const PV& v) {
        sr0[u] = v.a;
        sr1[u] = v.b;
#pragma unroll
        for (int c = 0; c < D; ++c) svh[u * D + c] = v.vh[c];
      });
  const bool wide = tab[0] < 0;
  const int nr = tile_rounds<G>();
  for (int round = 0; round < nr; ++round) {
    const Lane L = tile_lane<G>(t, round, n);
    if (wide)
      compat_row<D, G, 2, 2>(L, CompSrcG<D>{s0.nbr, s0.rec, vhs}, CompSrcG<D>{s1.nbr, s1.rec, vhs}, perm, vhs, s0,
                             s1, m_u, du0, u0, dta, u, dudt_last);
    else
      compat_row<D, G, 2, 2>(L, CompSrcT<D>{s0.loc, 0, sr0, svh}, CompSrcT<D>{s1.loc, 0, sr1, svh}, perm, vhs, s0, s1,
                             m_u, du0, u0, dta, u, dudt_last);
  }
}

// ------------------------------------------------------------ reductions
__global__ void k_minmax_init(float* mm, int32_t nr) {
  for (int32_t r = threadIdx.x; r < nr; r += blockDim.x) {
    mm[2 * r] = 3.0e38f;
    mm[2 * r + 1] = -3.0e38f;
  }
}
// min/max of positive values via bit-pattern atomics (segmented grid-stride,
// one atomic per block): mm[2r], mm[2r+1] of run r
__global__ void k_minmax(const float* __restrict__ a, int32_t n, float* mm, RunsDev R) {
  float lo = 3.0e38f, hi = 0.f;
  const Span sp = block_span(R, n);
  for (int32_t i = sp.k; i < sp.e; i += sp.step) {
    lo = fminf(lo, a[i]);
    hi = fmaxf(hi, a[i]);
  }
  lo = block_reduce(lo, MinOp());
  hi = block_reduce(hi, MaxOp());
  if (threadIdx.x == 0) {
    atomic_min_pos(&mm[2 * sp.r], lo);
    atomicMax((int*)&mm[2 * sp.r + 1], __float_as_int(hi));
  }
}

// dt_i = cfl * hmin / (c_i xi), c_i = sqrt(|u| g (g-1))  (timestep/compressible.py:12-66)
// (per run: mm[2r] = min h of run r, out[r])
__global__ void k_cfl(const float* __restrict__ u, int32_t n, float gamma, float cfl, float xi,
                      const float* __restrict__ mm, float* out, RunsDev R) {
  float d = 3.0e38f;
  const Span sp = block_span(R, n);
  const float hmin = mm[2 * sp.r];
  for (int32_t i = sp.k; i < sp.e; i += sp.step) {
    float cs = sqrtf(fabsf(u[i]) * gamma * (gamma - 1.f));
    d = fminf(d, cfl * hmin / (cs * xi));
  }
  d = block_reduce(d, MinOp());
  if (threadIdx.x == 0) atomic_min_pos(&out[sp.r], d);
}
// the clamps of crk_cfl_dt on the device (FP64, as on the host):
// dt = min(max(raw, minDt), maxDt), limited to growth * prev when prev > 0
__global__ void k_dt_finish(int32_t nr, const float* __restrict__ raw, const float* __restrict__ prev, double mind,
                            double maxd, double growth, float* __restrict__ out) {
  for (int32_t r = threadIdx.x; r < nr; r += blockDim.x) {
    double dt = raw[r];
    dt = fmin(fmax(dt, mind), maxd);
    const double p = prev ? (double)prev[r] : -1.0;
    if (p > 0 && dt > p) dt = fmin(dt, growth * p);
    out[r] = (float)dt;
  }
}
__global__ void k_time_advance(const int32_t* __restrict__ flag, int32_t nr, const float* __restrict__ dt,
                               double* __restrict__ t) {
  if (*flag) return;
  for (int32_t r = threadIdx.x; r < nr; r += blockDim.x) t[r] += (double)dt[r];
}
// dtp[k] = dt of the run of particle k
__global__ void k_dt_broadcast(const int32_t* __restrict__ run, int32_t n, const float* __restrict__ dtr,
                               float* __restrict__ dtp) {
  int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dtp[k] = dtr[run[k]];
}

// Per-run extent: o[7r] = max h (bits of a non-negative float), o[7r+1+d]
// = min x_d and o[7r+4+d] = max x_d as order-preserving keys of signed floats.
__device__ inline uint32_t fkey(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__host__ inline float fkey_inv(uint32_t k) {
  const uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  float f;
  std::memcpy(&f, &b, sizeof f);
  return f;
}
__global__ void k_extent_init(uint32_t* o, int32_t nr) {
  for (int32_t r = threadIdx.x; r < nr; r += blockDim.x) {
    o[7 * r] = 0u;
    for (int d = 0; d < 3; ++d) {
      o[7 * r + 1 + d] = 0xffffffffu;
      o[7 * r + 4 + d] = 0u;
    }
  }
}
template <int D>
__global__ void k_extent(const float* __restrict__ x, const float* __restrict__ h, int32_t n, uint32_t* o,
                         RunsDev R) {
  float hm = 0.f, lo[3], hi[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    lo[d] = 3.0e38f;
    hi[d] = -3.0e38f;
  }
  const Span sp = block_span(R, n);
  for (int32_t i = sp.k; i < sp.e; i += sp.step) {
    hm = fmaxf(hm, h[i]);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      lo[d] = fminf(lo[d], x[(int64_t)i * D + d]);
      hi[d] = fmaxf(hi[d], x[(int64_t)i * D + d]);
    }
  }
  hm = block_reduce(hm, MaxOp());
#pragma unroll
  for (int d = 0; d < D; ++d) {
    lo[d] = block_reduce(lo[d], MinOp());
    hi[d] = block_reduce(hi[d], MaxOp());
  }
  if (threadIdx.x == 0) {
    atomicMax(&o[7 * sp.r], __float_as_uint(hm));
#pragma unroll
    for (int d = 0; d < D; ++d) {
      atomicMin(&o[7 * sp.r + 1 + d], fkey(lo[d]));
      atomicMax(&o[7 * sp.r + 4 + d], fkey(hi[d]));
    }
  }
}

// FP64 conservation partials: per block [m, m v (D), m u, m |v|^2 / 2]
template <int D>
__global__ void k_conserve(const float* __restrict__ m, const float* __restrict__ v, const float* __restrict__ u,
                           int32_t n, double* __restrict__ partial) {
  constexpr int K = D + 3;
  __shared__ double sh[K][kBlock];
  int t = threadIdx.x;
  int32_t i = blockIdx.x * blockDim.x + t;
  double q[K];
#pragma unroll
  for (int c = 0; c < K; ++c) q[c] = 0.0;
  if (i < n) {
    double mi = m[i];
    q[0] = mi;
    double v2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      double vd = v[(int64_t)i * D + d];
      q[1 + d] = mi * vd;
      v2 += vd * vd;
    }
    q[D + 1] = mi * (double)u[i];
    q[D + 2] = 0.5 * mi * v2;
  }
#pragma unroll
  for (int c = 0; c < K; ++c) sh[c][t] = q[c];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (t < s)
#pragma unroll
      for (int c = 0; c < K; ++c) sh[c][t] += sh[c][t + s];
    __syncthreads();
  }
  if (t == 0)
#pragma unroll
    for (int c = 0; c < K; ++c) partial[(int64_t)blockIdx.x * K + c] = sh[c][0];
}

template <int K>
__global__ void k_conserve_final(const double* __restrict__ partial, int32_t nb, double* __restrict__ out) {
  __shared__ double sh[K][kBlock];
  int t = threadIdx.x;
  double q[K];
#pragma unroll
  for (int c = 0; c < K; ++c) q[c] = 0.0;
  for (int32_t b = t; b < nb; b += kBlock)
#pragma unroll
    for (int c = 0; c < K; ++c) q[c] += partial[(int64_t)b * K + c];
#pragma unroll
  for (int c = 0; c < K; ++c) sh[c][t] = q[c];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (t < s)
#pragma unroll
      for (int c = 0; c < K; ++c) sh[c][t] += sh[c][t + s];
    __syncthreads();
  }
  if (t == 0)
#pragma unroll
    for (int c = 0; c < K; ++c) out[c] = sh[c][0];
}

// ------------------------------------------------------------ allocation-free conservation sums
// crk_conservation_noalloc keeps its per-block partial sums and its K
// results in this file-scope array instead of context memory: one copy per
// GPU, held until the process ends.  Every call writes each slot it reads
// (partials first, then the final reduction), so nothing needs clearing, but
// two calls that overlap on one GPU share it and race.
constexpr int32_t kNoAllocMaxN = 1 << 23;  // particles per call
constexpr size_t kNoAllocSlots = (size_t)(kNoAllocMaxN / kBlock) * 6 + 8;  // blocks x (dim + 3) + results
static __device__ __align__(256) double scratch[kNoAllocSlots];

// Resolve this device's address of `scratch` (the host-side name is only a
// stand-in): called once per context, i.e. once per device, with that device
// current.  Kernels receive ctx->noalloc_scratch, never the array's name.
int noalloc_setup(CrkContext* ctx) {
  void* p = nullptr;
  CRK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  ctx->noalloc_scratch = (double*)p;
  return CRK_OK;
}

// ------------------------------------------------------------ finalize launch
struct CompArgs {
  int32_t n;
  const float *v0, *u0, *dvdt0, *dvdt1;
  float b0, b1;
  const float* masses;
  DtArg dta;
  float *vhs, *du0, *u_out, *dudt_last;
  int32_t threads;  // tile block size (tiled sweep)
};

// One finalize sweep over the stages X (and Y when NST = 2) of adjacency A.
template <int D, int G, int NST, int MODE>
static int compat_sweep(const CompArgs& a, const CrkAdjacency* A, const CompIn& X, const CompIn& Y, hipStream_t s) {
  hipLaunchKernelGGL(k_vhalf_sorted<D>, dim3(nblocks(a.n)), dim3(kBlock), 0, s, A->perm, a.n, a.v0, a.dvdt0, a.dvdt1,
                     a.b0, a.b1, a.dta, a.vhs);
  CRK_LAUNCH_CHECK("k_vhalf_sorted");
  // both stages in one tiled sweep when their pairs carry tile-local indices
  if constexpr (D <= 2 && NST == 2 && MODE == 2) {
    if (A->tiles && A->nbr_loc && (A->n_edges == 0 || (X.loc && Y.loc))) {
      hipLaunchKernelGGL((k_compat_t<D, G>), dim3(tile_count(a.n)), dim3(a.threads), 0, s, A->tiles, A->perm, a.n,
                         a.vhs, X, Y, a.masses, a.du0, a.u0, a.dta, a.u_out, a.dudt_last);
      CRK_LAUNCH_CHECK("k_compat_t");
      return CRK_OK;
    }
  }
  hipLaunchKernelGGL((k_compat<D, G, NST, MODE>), dim3(nblocks((int64_t)a.n * G)), dim3(kBlock), 0, s, A->perm, a.n,
                     a.vhs, X, Y, a.masses, a.du0, a.u0, a.dta, a.u_out, a.dudt_last);
  CRK_LAUNCH_CHECK("k_compat");
  return CRK_OK;
}

template <int D, int G>
static int compat_launch(const CompArgs& a, const CrkAdjacency* adj0, const CompIn& c0, const CrkAdjacency* adj1,
                         const CompIn& c1, bool shared, hipStream_t s) {
  if (shared) return compat_sweep<D, G, 2, 2>(a, adj0, c0, c1, s);
  int st = compat_sweep<D, G, 1, 0>(a, adj0, c0, c0, s);
  if (st) return st;
  return compat_sweep<D, G, 1, 1>(a, adj1, c1, c1, s);
}

}  // namespace crk

using namespace crk;

extern "C" {

int crk_axpy(int64_t count, const float* y0, const float* d, float dt, float* y, void* stream) {
  CRK_REQUIRE(count >= 0 && (count == 0 || (y0 && d && y)), "crk_axpy: invalid arguments");
  if (count == 0) return CRK_OK;
  hipLaunchKernelGGL(k_axpy, dim3(nblocks(count)), dim3(kBlock), 0, (hipStream_t)stream, count, y0, d, dt, y);
  CRK_LAUNCH_CHECK("k_axpy");
  return CRK_OK;
}

int crk_rk_combine(int64_t count, const float* y0, const float* d0, const float* d1, float b0, float b1, float dt,
                   float* y, void* stream) {
  CRK_REQUIRE(count >= 0 && (count == 0 || (y0 && d0 && d1 && y)), "crk_rk_combine: invalid arguments");
  if (count == 0) return CRK_OK;
  hipLaunchKernelGGL(k_combine, dim3(nblocks(count)), dim3(kBlock), 0, (hipStream_t)stream, count, y0, d0, d1, b0,
                     b1, dt, y);
  CRK_LAUNCH_CHECK("k_combine");
  return CRK_OK;
}

int crk_axpy_p(int64_t n, int32_t comps, const float* y0, const float* d, const float* dtp, float* y,
               void* stream) {
  CRK_REQUIRE(n >= 0 && comps >= 1 && (n == 0 || (y0 && d && dtp && y)), "crk_axpy_p: invalid arguments");
  if (n == 0) return CRK_OK;
  const int64_t cnt = n * comps;
  hipLaunchKernelGGL(k_axpy_p, dim3(nblocks(cnt)), dim3(kBlock), 0, (hipStream_t)stream, cnt, comps, y0, d, dtp, y);
  CRK_LAUNCH_CHECK("k_axpy_p");
  return CRK_OK;
}

int crk_rk_combine_p(int64_t n, int32_t comps, const float* y0, const float* d0, const float* d1, float b0, float b1,
                     const float* dtp, float* y, void* stream) {
  CRK_REQUIRE(n >= 0 && comps >= 1 && (n == 0 || (y0 && d0 && d1 && dtp && y)), "crk_rk_combine_p: invalid arguments");
  if (n == 0) return CRK_OK;
  const int64_t cnt = n * comps;
  hipLaunchKernelGGL(k_combine_p, dim3(nblocks(cnt)), dim3(kBlock), 0, (hipStream_t)stream, cnt, comps, y0, d0, d1,
                     b0, b1, dtp, y);
  CRK_LAUNCH_CHECK("k_combine_p");
  return CRK_OK;
}

int crk_wrap_positions(const CrkConfig* cfg, int32_t n, float* x, void* stream) {
  CRK_REQUIRE(cfg && x && n > 0, "crk_wrap_positions: invalid arguments");
  Geo g = make_geo(*cfg);
  const CrkDomain& dm = cfg->domain;
  hipLaunchKernelGGL(k_wrap, dim3(nblocks(n)), dim3(kBlock), 0, (hipStream_t)stream, n, dm.dim, g, dm.min[0],
                     dm.min[1], dm.min[2], dm.max[0], dm.max[1], dm.max[2], x);
  CRK_LAUNCH_CHECK("k_wrap");
  return CRK_OK;
}

int crk_compatible_energy(CrkContext* ctx, const CrkConfig* cfg, int32_t n, float dt, const float* masses,
                          const float* v0, const float* u0, const float* dvdt0, const float* dvdt1, float b0,
                          float b1, const CrkAdjacency* adj0, const CrkEdges* e0, const CrkAdjacency* adj1,
                          const CrkEdges* e1, float* u_out, float* dudt_last, void* stream) {
  CRK_REQUIRE(ctx && cfg && masses && v0 && u0 && dvdt0 && dvdt1 && adj0 && e0 && adj1 && e1 && u_out && dudt_last,
              "crk_compatible_energy: null argument");
  CRK_REQUIRE(e0->rec && e1->rec, "crk_compatible_energy: energy-partition records missing (crk_balance)");
  CRK_REQUIRE(n > 0 && adj0->n == n && adj1->n == n, "crk_compatible_energy: particle count mismatch");
  CRK_REQUIRE(adj0->perm && adj1->perm && adj0->row_offsets && adj1->row_offsets && e0->count && e1->count,
              "crk_compatible_energy: adjacency/pair counts missing");
  CRK_REQUIRE(adj0->n_edges == 0 || (e0->nbr && e0->pf), "stage-0 pairs missing");
  CRK_REQUIRE(adj1->n_edges == 0 || (e1->nbr && e1->pf), "stage-1 pairs missing");
  int D = cfg->domain.dim;
  CRK_REQUIRE(D >= 1 && D <= 3, "dim must be 1..3");
  const DtArg dta = dt_arg(ctx, dt);
  CRK_REQUIRE(dt > 0.f || dta.p, "crk_compatible_energy: dt must be positive");
  hipStream_t s = (hipStream_t)stream;
  int st;
  if ((st = require_runs_n(ctx, n, "crk_compatible_energy"))) return st;
  if ((st = ensure(ctx->vhs, sizeof(float) * (size_t)n * D))) return st;
  if ((st = ensure(ctx->du0, sizeof(float) * (size_t)n))) return st;
  float* vhs = (float*)ctx->vhs.p;
  float* du0 = (float*)ctx->du0.p;
  // both stages in one sweep when they share the sorted order
  const bool shared = adj0->perm == adj1->perm;
  CompIn c0{adj0->row_offsets, e0->count, e0->nbr, e0->pf, (const float4*)e0->rec, b0, e0->loc};
  CompIn c1{adj1->row_offsets, e1->count, e1->nbr, e1->pf, (const float4*)e1->rec, b1, e1->loc};
  const int GS = ctx->group[CRK_PASS_COMPAT];
  const CompArgs ca{n, v0, u0, dvdt0, dvdt1, b0, b1, masses, dta, vhs, du0, u_out, dudt_last,
                    ctx->threads[CRK_PASS_COMPAT]};
#define RUNG(DD)                                                                                  \
  switch (GS) {                                                                                   \
    case 1: st = compat_launch<DD, 1>(ca, adj0, c0, adj1, c1, shared, s); break;                  \
    case 2: st = compat_launch<DD, 2>(ca, adj0, c0, adj1, c1, shared, s); break;                  \
    case 4: st = compat_launch<DD, 4>(ca, adj0, c0, adj1, c1, shared, s); break;                  \
    case 8: st = compat_launch<DD, 8>(ca, adj0, c0, adj1, c1, shared, s); break;                  \
    case 16: st = compat_launch<DD, 16>(ca, adj0, c0, adj1, c1, shared, s); break;                \
    case 32: st = compat_launch<DD, 32>(ca, adj0, c0, adj1, c1, shared, s); break;                \
    case 64: st = compat_launch<DD, 64>(ca, adj0, c0, adj1, c1, shared, s); break;                \
    default: set_error("bad group size %d", GS); return CRK_EINVAL;                               \
  }
  switch (D) {
    case 1: RUNG(1) break;
    case 2: RUNG(2) break;
    case 3: RUNG(3) break;
  }
#undef RUNG
  return st;
}

int crk_minmax(CrkContext* ctx, int32_t n, const float* values, float* mm, void* stream) {
  CRK_REQUIRE(ctx && values && mm && n > 0, "crk_minmax: invalid arguments");
  int st = require_runs_n(ctx, n, "crk_minmax");
  if (st) return st;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(kBlock), 0, s, mm, run_count(ctx));
  hipLaunchKernelGGL(k_minmax, dim3(span_blocks(ctx, reduce_blocks(n))), dim3(kBlock), 0, s, values, n, mm,
                     runs_of(ctx));
  CRK_LAUNCH_CHECK("k_minmax");
  return CRK_OK;
}

int crk_cfl_dt(CrkContext* ctx, const CrkConfig* cfg, int32_t n, const float* supports, const float* u,
               double prev_dt, double* dt_out, void* stream) {
  CRK_REQUIRE(ctx && cfg && supports && u && dt_out && n > 0, "crk_cfl_dt: invalid arguments");
  hipStream_t s = (hipStream_t)stream;
  int st = ensure(ctx->s_red, 64);
  if (st) return st;
  float* red = (float*)ctx->s_red.p;  // [0,1] min/max h, [2] dt
  hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(kBlock), 0, s, red, 1);
  hipLaunchKernelGGL(k_minmax, dim3(reduce_blocks(n)), dim3(kBlock), 0, s, supports, n, red, RunsDev{});
  CRK_LAUNCH_CHECK("k_minmax");
  if ((st = fill_f32(red + 2, 1, 3.0e38f, s))) return st;
  float xi = (float)kernel_xi(cfg->kernel, cfg->domain.dim);
  hipLaunchKernelGGL(k_cfl, dim3(reduce_blocks(n)), dim3(kBlock), 0, s, u, n, cfg->gamma, cfg->cflFactor, xi, red,
                     red + 2, RunsDev{});
  CRK_LAUNCH_CHECK("k_cfl");
  float* hostv = (float*)ctx->pinned.p;
  CRK_HIP(hipMemcpyAsync(hostv, red + 2, sizeof(float), hipMemcpyDeviceToHost, s));
  CRK_HIP(hipStreamSynchronize(s));
  double dt = hostv[0];
  dt = std::fmin(std::fmax(dt, (double)cfg->minDt), (double)cfg->maxDt);
  if (prev_dt > 0 && dt > prev_dt) dt = std::fmin(dt, (double)cfg->dtGrowthFactor * prev_dt);
  *dt_out = dt;
  return CRK_OK;
}

int crk_cfl_dt_runs(CrkContext* ctx, const CrkConfig* cfg, int32_t n, const float* supports, const float* u,
                    const double* prev_dt, double* dt_out, float* dt_particles, void* stream) {
  CRK_REQUIRE(ctx && cfg && supports && u && dt_out && n > 0, "crk_cfl_dt_runs: invalid arguments");
  CRK_REQUIRE(ctx->n_runs > 0, "crk_cfl_dt_runs: no runs bound (crk_bind_runs)");
  int st = require_runs_n(ctx, n, "crk_cfl_dt_runs");
  if (st) return st;
  hipStream_t s = (hipStream_t)stream;
  const int32_t R = ctx->n_runs;
  if ((st = ensure(ctx->s_red, sizeof(float) * 3 * (size_t)R))) return st;
  if ((st = ensure_pinned(ctx, sizeof(float) * (size_t)R))) return st;
  float* mm = (float*)ctx->s_red.p;  // [2r, 2r+1] min/max h of run r; [2R + r] dt of run r
  float* dtr = mm + 2 * R;
  const RunsDev Rd = runs_of(ctx);
  hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(kBlock), 0, s, mm, R);
  hipLaunchKernelGGL(k_minmax, dim3(ctx->n_chunks), dim3(kBlock), 0, s, supports, n, mm, Rd);
  CRK_LAUNCH_CHECK("k_minmax");
  if ((st = fill_f32(dtr, R, 3.0e38f, s))) return st;
  float xi = (float)kernel_xi(cfg->kernel, cfg->domain.dim);
  hipLaunchKernelGGL(k_cfl, dim3(ctx->n_chunks), dim3(kBlock), 0, s, u, n, cfg->gamma, cfg->cflFactor, xi, mm, dtr,
                     Rd);
  CRK_LAUNCH_CHECK("k_cfl");
  float* hostv = (float*)ctx->pinned.p;
  CRK_HIP(hipMemcpyAsync(hostv, dtr, sizeof(float) * (size_t)R, hipMemcpyDeviceToHost, s));
  CRK_HIP(hipStreamSynchronize(s));
  // the clamps of crk_cfl_dt, run by run
  for (int32_t r = 0; r < R; ++r) {
    double dt = hostv[r];
    dt = std::fmin(std::fmax(dt, (double)cfg->minDt), (double)cfg->maxDt);
    if (prev_dt && prev_dt[r] > 0 && dt > prev_dt[r]) dt = std::fmin(dt, (double)cfg->dtGrowthFactor * prev_dt[r]);
    dt_out[r] = dt;
    hostv[r] = (float)dt;
  }
  if (dt_particles) {
    CRK_HIP(hipMemcpyAsync(dtr, hostv, sizeof(float) * (size_t)R, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_dt_broadcast, dim3(nblocks(n)), dim3(kBlock), 0, s, Rd.run, n, dtr, dt_particles);
    CRK_LAUNCH_CHECK("k_dt_broadcast");
    CRK_HIP(hipStreamSynchronize(s));
  }
  return CRK_OK;
}

int crk_run_extents(CrkContext* ctx, const CrkConfig* cfg, int32_t n, const float* x, const float* h,
                    float* out, void* stream) {
  CRK_REQUIRE(ctx && cfg && x && h && out && n > 0, "crk_run_extents: invalid arguments");
  const int D = cfg->domain.dim;
  CRK_REQUIRE(D >= 1 && D <= 3, "dim must be 1..3");
  int st = require_runs_n(ctx, n, "crk_run_extents");
  if (st) return st;
  hipStream_t s = (hipStream_t)stream;
  const int32_t R = run_count(ctx);
  if ((st = ensure(ctx->s_red, sizeof(uint32_t) * 7 * (size_t)R))) return st;
  if ((st = ensure_pinned(ctx, sizeof(uint32_t) * 7 * (size_t)R))) return st;
  uint32_t* o = (uint32_t*)ctx->s_red.p;
  hipLaunchKernelGGL(k_extent_init, dim3(1), dim3(kBlock), 0, s, o, R);
  const unsigned nb = span_blocks(ctx, reduce_blocks(n));
  const RunsDev Rd = runs_of(ctx);
  CRK_DIM_DISPATCH(D, hipLaunchKernelGGL(k_extent<DD>, dim3(nb), dim3(kBlock), 0, s, x, h, n, o, Rd));
  CRK_LAUNCH_CHECK("k_extent");
  uint32_t* hostv = (uint32_t*)ctx->pinned.p;
  CRK_HIP(hipMemcpyAsync(hostv, o, sizeof(uint32_t) * 7 * (size_t)R, hipMemcpyDeviceToHost, s));
  CRK_HIP(hipStreamSynchronize(s));
  for (int32_t r = 0; r < R; ++r) {
    std::memcpy(&out[7 * r], &hostv[7 * r], sizeof(float));
    for (int d = 0; d < 3; ++d) {
      out[7 * r + 1 + d] = d < D ? fkey_inv(hostv[7 * r + 1 + d]) : 0.f;
      out[7 * r + 4 + d] = d < D ? fkey_inv(hostv[7 * r + 4 + d]) : 0.f;
    }
  }
  return CRK_OK;
}

int crk_cfl_dt_device(CrkContext* ctx, const CrkConfig* cfg, int32_t n, const float* supports, const float* u,
                      const float* prev_dt, float* dt_out, float* dt_particles, void* stream) {
  CRK_REQUIRE(ctx && cfg && supports && u && dt_out && dt_particles && n > 0, "crk_cfl_dt_device: invalid arguments");
  CRK_REQUIRE(ctx->n_runs > 0, "crk_cfl_dt_device: no runs bound (crk_bind_runs)");
  int st = require_runs_n(ctx, n, "crk_cfl_dt_device");
  if (st) return st;
  hipStream_t s = (hipStream_t)stream;
  const int32_t R = ctx->n_runs;
  if ((st = ensure(ctx->s_red, sizeof(float) * 3 * (size_t)R))) return st;
  float* mm = (float*)ctx->s_red.p;
  float* raw = mm + 2 * R;
  const RunsDev Rd = runs_of(ctx);
  hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(kBlock), 0, s, mm, R);
  hipLaunchKernelGGL(k_minmax, dim3(ctx->n_chunks), dim3(kBlock), 0, s, supports, n, mm, Rd);
  CRK_LAUNCH_CHECK("k_minmax");
  if ((st = fill_f32(raw, R, 3.0e38f, s))) return st;
  float xi = (float)kernel_xi(cfg->kernel, cfg->domain.dim);
  hipLaunchKernelGGL(k_cfl, dim3(ctx->n_chunks), dim3(kBlock), 0, s, u, n, cfg->gamma, cfg->cflFactor, xi, mm, raw,
                     Rd);
  CRK_LAUNCH_CHECK("k_cfl");
  hipLaunchKernelGGL(k_dt_finish, dim3(1), dim3(kBlock), 0, s, R, raw, prev_dt, (double)cfg->minDt,
                     (double)cfg->maxDt, (double)cfg->dtGrowthFactor, dt_out);
  CRK_LAUNCH_CHECK("k_dt_finish");
  hipLaunchKernelGGL(k_dt_broadcast, dim3(nblocks(n)), dim3(kBlock), 0, s, Rd.run, n, dt_out, dt_particles);
  CRK_LAUNCH_CHECK("k_dt_broadcast");
  return CRK_OK;
}

int crk_runs_advance_time(CrkContext* ctx, const float* dt_runs, double* t_runs, void* stream) {
  CRK_REQUIRE(ctx && dt_runs && t_runs, "crk_runs_advance_time: null argument");
  CRK_REQUIRE(ctx->n_runs > 0, "crk_runs_advance_time: no runs bound (crk_bind_runs)");
  hipLaunchKernelGGL(k_time_advance, dim3(1), dim3(kBlock), 0, (hipStream_t)stream, ctx->d_flag, ctx->n_runs,
                     dt_runs, t_runs);
  CRK_LAUNCH_CHECK("k_time_advance");
  return CRK_OK;
}

int crk_runs_dt(CrkContext* ctx, int32_t n, const double* dt_runs, float* dt_particles, void* stream) {
  CRK_REQUIRE(ctx && dt_runs && dt_particles, "crk_runs_dt: null argument");
  CRK_REQUIRE(ctx->n_runs > 0, "crk_runs_dt: no runs bound (crk_bind_runs)");
  int st = require_runs_n(ctx, n, "crk_runs_dt");
  if (st) return st;
  const int32_t R = ctx->n_runs;
  for (int32_t r = 0; r < R; ++r) CRK_REQUIRE(dt_runs[r] > 0, "crk_runs_dt: dt of run %d must be positive", r);
  hipStream_t s = (hipStream_t)stream;
  if ((st = ensure(ctx->s_red, sizeof(float) * 3 * (size_t)R))) return st;
  if ((st = ensure_pinned(ctx, sizeof(float) * (size_t)R))) return st;
  float* dtr = (float*)ctx->s_red.p + 2 * R;
  float* hostv = (float*)ctx->pinned.p;
  for (int32_t r = 0; r < R; ++r) hostv[r] = (float)dt_runs[r];
  CRK_HIP(hipMemcpyAsync(dtr, hostv, sizeof(float) * (size_t)R, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_dt_broadcast, dim3(nblocks(n)), dim3(kBlock), 0, s, run_ids(ctx), n, dtr, dt_particles);
  CRK_LAUNCH_CHECK("k_dt_broadcast");
  CRK_HIP(hipStreamSynchronize(s));
  return CRK_OK;
}

int crk_conservation(CrkContext* ctx, int32_t dim, int32_t n, const float* masses, const float* velocities,
                     const float* u, double* out, void* stream) {
  CRK_REQUIRE(ctx && masses && velocities && u && out && n > 0, "crk_conservation: invalid arguments");
  CRK_REQUIRE(dim >= 1 && dim <= 3, "dim must be 1..3");
  hipStream_t s = (hipStream_t)stream;
  int32_t nb = (int32_t)nblocks(n);
  int K = dim + 3;
  int st = ensure(ctx->s_counts, sizeof(double) * ((size_t)nb * K + 8));
  if (st) return st;
  double* partial = (double*)ctx->s_counts.p;
  double* dout = partial + (size_t)nb * K;
  switch (dim) {
    case 1:
      hipLaunchKernelGGL(k_conserve<1>, dim3(nb), dim3(kBlock), 0, s, masses, velocities, u, n, partial);
      hipLaunchKernelGGL(k_conserve_final<4>, dim3(1), dim3(kBlock), 0, s, partial, nb, dout);
      break;
    case 2:
      hipLaunchKernelGGL(k_conserve<2>, dim3(nb), dim3(kBlock), 0, s, masses, velocities, u, n, partial);
      hipLaunchKernelGGL(k_conserve_final<5>, dim3(1), dim3(kBlock), 0, s, partial, nb, dout);
      break;
    case 3:
      hipLaunchKernelGGL(k_conserve<3>, dim3(nb), dim3(kBlock), 0, s, masses, velocities, u, n, partial);
      hipLaunchKernelGGL(k_conserve_final<6>, dim3(1), dim3(kBlock), 0, s, partial, nb, dout);
      break;
  }
  CRK_LAUNCH_CHECK("k_conserve");
  double* hostv = (double*)ctx->pinned.p;
  CRK_HIP(hipMemcpyAsync(hostv, dout, sizeof(double) * K, hipMemcpyDeviceToHost, s));
  CRK_HIP(hipStreamSynchronize(s));
  for (int c = 0; c < K; ++c) out[c] = hostv[c];
  return CRK_OK;
}


int crk_conservation_noalloc(CrkContext* ctx, int32_t dim, int32_t n, const float* masses, const float* velocities,
                             const float* u, double* out_dev, void* stream) {
  CRK_REQUIRE(ctx && masses && velocities && u && out_dev && n > 0, "crk_conservation_noalloc: invalid arguments");
  CRK_REQUIRE(dim >= 1 && dim <= 3, "dim must be 1..3");
  CRK_REQUIRE(n <= kNoAllocMaxN, "crk_conservation_noalloc: n = %d exceeds the fixed scratch capacity of %d particles",
              n, kNoAllocMaxN);
  CRK_REQUIRE(ctx->noalloc_scratch, "crk_conservation_noalloc: scratch address not resolved for this context");
  int cur = -1;
  CRK_HIP(hipGetDevice(&cur));
  CRK_REQUIRE(cur == ctx->device, "crk_conservation_noalloc: the current device (%d) is not the context's (%d)", cur,
              ctx->device);
  hipStream_t s = (hipStream_t)stream;
  const int32_t nb = (int32_t)nblocks(n);
  double* partial = ctx->noalloc_scratch;  // [nb][dim + 3]
  // the same kernels and block count as crk_conservation: bit-identical sums
  switch (dim) {
    case 1:
      hipLaunchKernelGGL(k_conserve<1>, dim3(nb), dim3(kBlock), 0, s, masses, velocities, u, n, partial);
      hipLaunchKernelGGL(k_conserve_final<4>, dim3(1), dim3(kBlock), 0, s, partial, nb, out_dev);
      break;
    case 2:
      hipLaunchKernelGGL(k_conserve<2>, dim3(nb), dim3(kBlock), 0, s, masses, velocities, u, n, partial);
      hipLaunchKernelGGL(k_conserve_final<5>, dim3(1), dim3(kBlock), 0, s, partial, nb, out_dev);
      break;
    case 3:
      hipLaunchKernelGGL(k_conserve<3>, dim3(nb), dim3(kBlock), 0, s, masses, velocities, u, n, partial);
      hipLaunchKernelGGL(k_conserve_final<6>, dim3(1), dim3(kBlock), 0, s, partial, nb, out_dev);
      break;
  }
  CRK_LAUNCH_CHECK("k_conserve (noalloc)");
  return CRK_OK;
}

}  // extern "C"
