// forces.hip -- the pair passes after the CRK factors: velocity gradient
// (with the Cullen-Dehnen divergence), the Cullen-Dehnen switch, the
// acceleration / energy pass (optionally with the switch fused into the same
// sweep) and the energy partition (src/warpSPH/schemes/crkSPH.py:158-364).
#include <cmath>

#include "tiles.hpp"

namespace crk {

// ------------------------------------------------------------ velocity gradient
// J_i[a][g] = sum_j V_j (v_j - v_i)[a] d_g W^R_ij (scatter), fused with the
// Cullen-Dehnen divergence (super-symmetric support max(h_i, h_j), m/rho
// volumes).
struct VgIn {
  const float *x, *h, *v, *V, *m, *rho, *c, *A, *B, *gA, *gB;
  // neighbour records: geo = {x, h}; vg = {v, V, m / rho} (d < 3) or
  // {v, V} with mr = m / rho (d = 3)
  const float4 *geo4, *vg4;
  const float* mr;
};
struct VgOut {
  float *J, *divcd, *vsig, *drhodt_u;
};

// One row of the velocity-gradient pass from a neighbour source (geo = {x, h},
// aux = the {v, V, m / rho} record; d = 3 reads m / rho through the ids).
template <int D, int G, int K, bool CULLEN, class Src>
__device__ inline void velgrad_row(const Lane& L, const Src& src, ARows rows, Geo geo, float C, int32_t n,
                                   const VgIn& in, const int32_t* __restrict__ perm, const VgOut& o) {
  const int32_t k = L.k;
  // the row's own {x, h} and {v, ...} from its records
  const float4 gk = in.geo4[k], wk = in.vg4[k];
  const float xi[3] = {gk.x, gk.y, gk.z}, vi[3] = {wk.x, wk.y, wk.z};
  const float hi = gk.w;
  CorrD<D> cr = load_corr<D>(in.A, in.B, in.gA, in.gB, k);
  Sum acc[D * D];
  Sum dacc;
  const int32_t b = rows.off[k], e = L.valid ? b + rows.cnt[k] : b;
  // software pipeline: the neighbour's records one iteration ahead, its id two ahead
  int32_t jc = b + L.g < e ? src.id(b + L.g) : 0;
  int32_t jn = b + L.g + G < e ? src.id(b + L.g + G) : 0;
  float4 gn = src.geo(jc), wn = src.aux(jc);
  float mn = (D == 3 && CULLEN) ? in.mr[jc] : 0.f;
  for (int32_t q = b + L.g; q < e; q += G) {
    const float4 gj = gn, wj = wn;
    const float mrj = D < 3 ? wj.w : mn;
    if (q + G < e) {
      jc = jn;
      gn = src.geo(jc);
      wn = src.aux(jc);
      if (D == 3 && CULLEN) mn = in.mr[jc];
      if (q + 2 * G < e) jn = src.id(q + 2 * G);
    }
    float dx[3];
    const float xj[3] = {gj.x, gj.y, gj.z};
    pair_dx<D>(geo, xi, xj, dx);
    float r = fsqrt(dot<D>(dx, dx));
    const float vj[3] = {wj.x, wj.y, wj.z};
    float hj = gj.w;
    if (CULLEN) {
      float H = fmaxf(hi, hj);
      if (r < H) {
        WdW w = wdw<K, D>(C, r, frcp(H));
        float dt_ = 0.f;
#pragma unroll
        for (int a = 0; a < D; ++a) dt_ = dt_ + (vj[a] - vi[a]) * dx[a];
        dacc.fma(mrj, w.fac * dt_);
      }
    }
    if (!(r < hj)) continue;
    WdW w = wdw<K, D>(C, r, frcp(hj));
    float g[3];
    grad_wr<D>(cr, dx, w, g);
    float Vj = D < 3 ? wj.z : wj.w;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      float dv = Vj * (vj[a] - vi[a]);
#pragma unroll
      for (int gm = 0; gm < D; ++gm) acc[a * D + gm].fma(dv, g[gm]);
    }
  }
  float Jt[D * D];
#pragma unroll
  for (int a = 0; a < D * D; ++a) Jt[a] = group_total<G>(acc[a]);
  float dv = 0.f;
  if (CULLEN) dv = group_total<G>(dacc);
  if (!L.valid || L.g != 0) return;
  float tr = 0.f;
#pragma unroll
  for (int a = 0; a < D * D; ++a) o.J[(int64_t)k * D * D + a] = Jt[a];
#pragma unroll
  for (int a = 0; a < D; ++a) tr = tr + Jt[a * D + a];
  o.drhodt_u[perm[k]] = -tr * in.rho[k];
  if (CULLEN) o.divcd[k] = dv;
}

template <int D, int G, int K, bool CULLEN>
__global__ void __launch_bounds__(kBlock) k_velgrad(ARows rows, Geo geo, float C, int32_t n, VgIn in,
                                                    const int32_t* __restrict__ perm, VgOut o) {
  const Lane L = lane_of<G>(n);
  velgrad_row<D, G, K, CULLEN>(L, GlobalSrc<float4>{rows.nbr, in.geo4, in.vg4}, rows, geo, C, n, in, perm, o);
}

// One block per tile, the pairs' {x, h} and {v, V, m / rho} records staged
// in LDS (d <= 2).
template <int D, int G, int K, bool CULLEN>
__global__ void __launch_bounds__(kTileThreadsMax) k_velgrad_t(ARows rows, const uint16_t* __restrict__ ploc,
                                                      const int32_t* __restrict__ tiles, Geo geo, float C, int32_t n,
                                                      VgIn in, const int32_t* __restrict__ perm, VgOut o) {
  static_assert(D <= 2, "tiled velocity gradient: d <= 2");
  __shared__ int32_t tab[kTileWords];
  __shared__ float4 sg[kTileCap];
  __shared__ float4 sw[kTileCap];
  // (the pair indices stay in global memory here: staging them in LDS was
  // measured slower for this pass)
  const int32_t t = xcd_block();
  struct PV {
    float4 a, b;
  };
  tile_stage_fn<2, 0, PV>(
      tiles, t, 0, 0, nullptr, tab, nullptr, [&](int32_t g) { return PV{in.geo4[g], in.vg4[g]}; },
      [&](int32_t u, const PV& v) {
        sg[u] = v.a;
        sw[u] = v.b;
      });
  const bool wide = tab[0] < 0;
  const int nr = tile_rounds<G>();
  for (int round = 0; round < nr; ++round) {
    const Lane L = tile_lane<G>(t, round, n);
    if (wide)
      velgrad_row<D, G, K, CULLEN>(L, GlobalSrc<float4>{rows.nbr, in.geo4, in.vg4}, rows, geo, C, n, in, perm, o);
    else
      velgrad_row<D, G, K, CULLEN>(L, TileSrc<float4>{ploc, 0, rows.nbr, sg, sw}, rows, geo, C, n, in, perm, o);
  }
}

// ------------------------------------------------------------ Cullen-Dehnen
// CullenDehnen2010.py:218-331: R (gather interpolation of rho sign(div)), Xi,
// A = Xi max(0, -ddiv/dt), local alpha, and the alpha0 relaxation.
struct CdIn {
  const float *x, *h, *m, *rho, *divcd, *vsig, *divp;
};
struct CdOut {
  float *a0, *alpha, *Rcd, *alphas_u, *alpha0s_u;
};
struct CdParams {
  DtArg dt;
  float fkern, amin, amax, beta;
  int limitXi, evolve;
};

// the per-particle inputs of the switch's row update (loaded apart from it,
// so that a pass can issue the loads before its pair sweep)
struct CdRow {
  float rho, divcd, divp, vsig, a0, dt;
};
__device__ inline CdRow cd_row_load(const CdIn& in, const CdParams& p, const CdOut& o, int32_t k) {
  return CdRow{in.rho[k], in.divcd[k], in.divp[k], in.vsig[k], o.a0[k], p.dt.at(k)};
}

// per-row update from S = sum_j m_j sgn(div_j) W(x_ij, h_i)
__device__ inline void cullen_row(const CdRow& r, const CdParams& p, int32_t k, int32_t i, float S, float hi,
                                  const CdOut& o) {
  float Rv = S / r.rho;
  o.Rcd[k] = Rv;
  float dk_ = r.divcd;
  float t1 = 1.f - Rv;
  float t2 = t1 * t1;
  float nom = p.beta * (t2 * t2) * dk_;
  float Xi = p.limitXi ? (nom * nom + 1e-14f * hi) / (nom * nom + 0.f + 1e-14f * hi)
                       : (nom * nom) / (nom * nom + 0.f + 1e-14f * hi);
  const float dt = r.dt;
  float ddiv = (dk_ - r.divp) / dt;
  float md = -ddiv;
  float Ai = Xi * (md > 0.f ? md : 0.f);
  float vsg = fabsf(r.vsig);
  float scaling = hi * hi * Ai / (vsg * vsg + hi * hi * Ai + 1e-14f * hi);
  float al = p.amax * scaling;
  al = al < p.amin ? p.amin : (al > p.amax ? p.amax : al);
  float tau = hi * p.fkern / (0.1f * vsg);
  float a0k = r.a0;
  float adot = (al - a0k) / tau;
  float na = al > a0k ? al : a0k;
  na = al < na ? na + adot * dt : na;
  na = na < p.amin ? p.amin : (na > p.amax ? p.amax : na);
  o.alpha[k] = na;
  o.alphas_u[i] = na;
  if (p.evolve) { o.a0[k] = na; if (o.alpha0s_u) o.alpha0s_u[i] = na; }
}

template <int D, int G, int K>
__global__ void __launch_bounds__(kBlock) k_cullen(ARows rows, Geo geo, float C, int32_t n, CdIn in, CdParams p,
                                                   const int32_t* __restrict__ perm, CdOut o) {
  Lane L = lane_of<G>(n);
  const int32_t k = L.k;
  float xi[3];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = in.x[(int64_t)k * D + d];
  const float hi = in.h[k];
  const float ih = 1.f / hi;
  Comp acc;
  const int32_t b = rows.off[k], e = L.valid ? b + rows.cnt[k] : b;
  int32_t jn = b + L.g < e ? rows.nbr[b + L.g] : 0;  // neighbour ids loaded one iteration ahead
  for (int32_t q = b + L.g; q < e; q += G) {
    const int32_t j = jn;
    if (q + G < e) jn = rows.nbr[q + G];
    float dx[3];
    pair_dx<D>(geo, xi, &in.x[(int64_t)j * D], dx);
    float qq = qmul(fsqrt(dot<D>(dx, dx)), ih);
    if (!(qq < 1.f)) continue;
    float dj = in.divcd[j];
    float sg = dj > 0.f ? 1.f : (dj < 0.f ? -1.f : 0.f);
    float kk, dk;
    kwd<K, D>(qq, kk, dk);
    acc.add(in.m[j] * sg * (C * kk * ipow<D>(ih)));
  }
  float S = group_total<G>(acc);
  if (!L.valid || L.g != 0) return;
  cullen_row(cd_row_load(in, p, o, k), p, k, perm[k], S, hi, o);
}

// ------------------------------------------------------------ learned switch
// |grad rho|, |grad P| with the CRK gradient of the velocity gradient pass:
// grad f_i = sum_j V_j (f_j - f_i) grad W^R_ij over the in-support pairs
template <int D, int G, int K>
__global__ void __launch_bounds__(kBlock) k_gradnorm(ARows rows, Geo geo, float C, int32_t n, VgIn in,
                                                     const float* __restrict__ P, float2* __restrict__ out) {
  Lane L = lane_of<G>(n);
  const int32_t k = L.k;
  float xi[3];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = in.x[(int64_t)k * D + d];
  const float ri = in.rho[k], Pi = P[k];
  CorrD<D> cr = load_corr<D>(in.A, in.B, in.gA, in.gB, k);
  Comp gr[D], gp[D];
  const int32_t b = rows.off[k], e = L.valid ? b + rows.cnt[k] : b;
  for (int32_t q = b + L.g; q < e; q += G) {
    const int32_t j = rows.nbr[q];
    float dx[3];
    pair_dx<D>(geo, xi, &in.x[(int64_t)j * D], dx);
    const float r = fsqrt(dot<D>(dx, dx));
    const float hj = in.h[j];
    if (!(r < hj)) continue;
    WdW w = wdw<K, D>(C, r, frcp(hj));
    float g[3];
    grad_wr<D>(cr, dx, w, g);
    const float Vj = in.V[j];
    const float dr = Vj * (in.rho[j] - ri), dp = Vj * (P[j] - Pi);
#pragma unroll
    for (int a = 0; a < D; ++a) {
      gr[a].add(dr * g[a]);
      gp[a].add(dp * g[a]);
    }
  }
  float nr = 0.f, np = 0.f;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    const float u = group_total<G>(gr[a]), v = group_total<G>(gp[a]);
    nr += u * u;
    np += v * v;
  }
  if (L.valid && L.g == 0) out[k] = make_float2(sqrtf(nr), sqrtf(np));
}

// Features of the learned switch (dimensionless forms of div v, d(div v)/dt,
// |curl v|, h, c, |grad rho|, |grad P| and the neighbour count):
//   f0 = h div / c,  f1 = h^2 (div - div_prev) / (dt c^2),  f2 = h |curl v| / c,
//   f3 = h (rho / m)^(1/d),  f4 = c / v_sig,  f5 = h |grad rho| / rho,
//   f6 = h |grad P| / P,  f7 = (in-support neighbours) / n_H.
// div is the Cullen-Dehnen divergence and div_prev the previous one (the
// inputs of CullenDehnen2010.py:281).  raw: features in user order (perm);
// otherwise normalised (f - mu) / sigma in sorted order with flag = 1 where a
// raw feature leaves [lo, hi] or is not finite.
struct FeatIn {
  const float *h, *c, *rho, *m, *P, *divcd, *divp, *J, *vsig;
  const float2* gn;
  const int32_t* cnt;
};
template <int D>
__global__ void k_features(int32_t n, FeatIn in, DtArg dt, float nT, CrkMlp model, const int32_t* __restrict__ perm,
                           float* __restrict__ raw, float* __restrict__ xn, uint8_t* __restrict__ flag) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float h = in.h[k], c = in.c[k], rho = in.rho[k], P = in.P[k];
  const float ic = 1.f / c;
  const float* J = &in.J[(int64_t)k * D * D];
  float curl = 0.f;
  if (D == 2) curl = fabsf(J[2] - J[1]);
  if (D == 3) {
    const float a = J[7] - J[5], b = J[2] - J[6], d = J[3] - J[1];
    curl = sqrtf(a * a + b * b + d * d);
  }
  const float div = in.divcd[k];
  const float2 gn = in.gn[k];
  const float vs = in.vsig[k];
  float f[8];
  f[0] = h * div * ic;
  f[1] = h * h * (div - in.divp[k]) / dt.at(k) * ic * ic;
  f[2] = h * curl * ic;
  f[3] = h * powf(rho / in.m[k], 1.f / (float)D);
  f[4] = vs > 0.f ? c / vs : 1.f;
  f[5] = h * gn.x / rho;
  f[6] = h * gn.y / P;
  f[7] = (float)in.cnt[k] / nT;
  if (raw) {
    float* o = &raw[(int64_t)perm[k] * 8];
#pragma unroll
    for (int q = 0; q < 8; ++q) o[q] = f[q];
    return;
  }
  uint8_t out = 0;
  float* o = &xn[(int64_t)k * 8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    if (!(f[q] >= model.lo[q] && f[q] <= model.hi[q])) out = 1;  // also catches NaN
    o[q] = (f[q] - model.mu[q]) / model.sigma[q];
  }
  flag[k] = out;
}

// ------------------------------------------------------------ accel + dudt
struct LimParams {
  float eta_fold, inv_eta_fold, eta_crit;
  int crk, vl, off, on;
};

// limiter.py:11-16 (van Leer), :28-86 (computeVanLeer), :90-127 (crkLimiter)
// cI = J_i x_ij, cJ = J_j x_ij (shared with the velocity reconstruction)
template <int D>
__device__ inline float limiter_phi(const LimParams& lp, const float* dx, float r, float ihi, float ihj,
                                    const float* cI, const float* cJ) {
  float factor = 1.f;
  if (lp.crk) {
    float eta = r * fminf(ihi, ihj);
    const float t = (eta - lp.eta_crit) * lp.inv_eta_fold;  // (select, not a branch)
    const float f = __expf(-(t * t));
    factor = eta < lp.eta_crit ? f : 1.f;
  }
  float phi = 0.f;
  if (lp.vl) {
    // g_i = (J_i x/2) . x/2
    float gi = 0.f, gj = 0.f;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      gi = gi + cI[a] * dx[a];
      gj = gj + cJ[a] * dx[a];
    }
    gi = 0.25f * gi;
    gj = 0.25f * gj;
    // r_i = g_i / g_j, r_j = g_j / g_i (sgn(g)|g| = g) and min(r_i, r_j):
    // for slopes of one sign min(|g_i|, |g_j|) / max(|g_i|, |g_j|) (one
    // reciprocal instead of two); opposite signs or one zero slope give
    // min <= 0; two zero slopes give 0/0 -> 1 in both ratios
    const float agi = fabsf(gi), agj = fabsf(gj);
    const bool same = (gi > 0.f && gj > 0.f) || (gi < 0.f && gj < 0.f);
    const float rr0 = fminf(agi, agj) * frcp(fmaxf(agi, agj));  // (0 * inf when both are 0: not selected)
    float rr = same ? rr0 : 0.f;
    if (gi == 0.f && gj == 0.f) rr = 1.f;
    // (rr = 0 gives vl = 0 here too)
    const float q = 2.f * frcp(1.f + rr);
    const float vl = rr * q * q;
    phi = vl * factor;
  }
  if (lp.off) phi = 0.f;
  if (lp.on) phi = 1.f;
  phi = phi < 1.f ? phi : 1.f;
  phi = phi > 0.f ? phi : 0.f;
  return phi;
}

struct AdIn {
  const float *x, *h, *v, *m, *V, *rho, *P, *c, *J, *A, *B, *gA, *gB, *alpha;
  const float* drhodt_u;
};
struct AdOut {
  float *pf, *th, *dvdt_s, *dvdt_u, *dudt_u, *dEdt_u, *div_u;
};

// a D-vector of a slot (2D: one 8-byte store)
template <int D>
__device__ inline void store_slot(float* p, int64_t q, const float* v) {
  if (D == 2) {
    reinterpret_cast<float2*>(p)[q] = make_float2(v[0], v[1]);
  } else {
#pragma unroll
    for (int c = 0; c < D; ++c) p[q * D + c] = v[c];
  }
}

// The acceleration pass gathers nearly every per-particle field of its
// neighbours; one coalesced pass packs them into the neighbour record
// (Rec<D>) first, so that a neighbour costs NF/4 16-byte loads.
template <int D>
__global__ void k_pack_rec(int32_t n, AdIn in, const float* __restrict__ divcd, float Cl, float Cq,
                           float4* __restrict__ rec) {
  int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  float r[Rec<D>::NF];
#pragma unroll
  for (int c = 0; c < Rec<D>::NF; ++c) r[c] = 0.f;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    r[Rec<D>::X + d] = in.x[(int64_t)k * D + d];
    r[Rec<D>::VEL + d] = in.v[(int64_t)k * D + d];
    r[Rec<D>::B + d] = in.B[(int64_t)k * D + d];
    r[Rec<D>::GA + d] = in.gA[(int64_t)k * D + d];
  }
#pragma unroll
  for (int a = 0; a < D * D; ++a) {
    r[Rec<D>::J + a] = in.J[(int64_t)k * D * D + a];
    r[Rec<D>::GB + a] = in.gB[(int64_t)k * D * D + a];
  }
  r[Rec<D>::H] = in.h[k];
  r[Rec<D>::V] = in.V[k];
  const float rho = in.rho[k];
  r[Rec<D>::QA] = -Cl * rho * in.c[k];
  r[Rec<D>::QB] = Cq * rho;
  r[Rec<D>::A] = in.A[k];
  r[Rec<D>::P] = in.P[k];
  r[Rec<D>::M] = in.m[k];
  const float dv = divcd ? divcd[k] : 0.f;
  r[Rec<D>::MSG] = dv > 0.f ? in.m[k] : (dv < 0.f ? -in.m[k] : 0.f);
  float4* p = rec + (int64_t)k * (Rec<D>::NF / 4);
#pragma unroll
  for (int c = 0; c < Rec<D>::NF / 4; ++c) p[c] = make_float4(r[4 * c], r[4 * c + 1], r[4 * c + 2], r[4 * c + 3]);
}

// SW: the Cullen-Dehnen switch of the next step is evaluated in the same
// sweep (valid when the viscosity does not read alpha, i.e. without
// alpha-weighted Q): S_i = sum_j m_j sgn(div_j) W(x_ij, h_i) over r < h_i.
// Record sources of the acceleration pass: the packed records in global
// memory through the sorted ids, or a tile's LDS copy through the local ids.
struct RecSrcG {
  const int32_t* __restrict__ nbr;
  const float4* __restrict__ rec;
  __device__ inline int32_t id(int64_t q) const { return nbr[q]; }
  template <int D>
  __device__ inline void load(int32_t j, float* r) const { rec_load<D>(rec, j, r); }
};
struct RecSrcT {
  const uint16_t* sid;  // LDS copy of the tile's 16-bit indices from entry base
  int32_t base;
  const float4* rec;  // LDS
  __device__ inline int32_t id(int64_t q) const { return sid[q - base]; }
  template <int D>
  __device__ inline void load(int32_t j, float* r) const { rec_load<D>(rec, j, r); }
};

// One row of the acceleration pass (alpha-weighted Q reads alpha_j through
// the sorted ids: global source only).
// The row's record and CSR range, loaded by the caller (the tiled pass loads
// its first round's rows before staging the tile) or here.
template <int D>
struct RowPre {
  float r[Rec<D>::NF];
  int32_t off, cnt;
};
template <int D>
__device__ inline void row_pre_load(RowPre<D>& p, const float4* __restrict__ rec, const ARows& rows, int32_t k) {
  rec_load<D>(rec, k, p.r);
  p.off = rows.off[k];
  p.cnt = rows.cnt[k];
}

template <int D, int G, int K, bool SW, class Src>
__device__ inline void accel_row(const Lane& L, const Src& src, ARows rows, Geo geo, float C, const LimParams& lp,
                                 float inv_kscale, int alphaQ, int32_t n, const AdIn& in,
                                 const int32_t* __restrict__ perm, const AdOut& o, const CdIn& cin,
                                 const CdParams& cp, const CdOut& cout, const float4* __restrict__ rec,
                                 const RowPre<D>& pre) {
  const int32_t k = L.k;
  // the row's own fields from its record (NF/4 16-byte loads)
  const float* ri = pre.r;
  // the epilogue's per-row loads, issued before the sweep so that their
  // latency hides behind it
  const int32_t i = perm[k];
  const float drhodt_i = in.drhodt_u[i], rho_k = in.rho[k];
  CdRow cdr{};
  if (SW) cdr = cd_row_load(cin, cp, cout, k);
  float xi[3], vi[3], Ji[D * D];
#pragma unroll
  for (int d = 0; d < D; ++d) { xi[d] = ri[Rec<D>::X + d]; vi[d] = ri[Rec<D>::VEL + d]; }
#pragma unroll
  for (int a = 0; a < D * D; ++a) Ji[a] = ri[Rec<D>::J + a];
  const float hi = ri[Rec<D>::H], Pi = ri[Rec<D>::P], qai = ri[Rec<D>::QA], qbi = ri[Rec<D>::QB],
              Vi = ri[Rec<D>::V], mi = ri[Rec<D>::M];
  const float ihi = 1.f / hi;
  const float ali = alphaQ ? in.alpha[k] : 1.f;
  CorrD<D> cri;
  cri.A = ri[Rec<D>::A];
#pragma unroll
  for (int a = 0; a < D; ++a) { cri.B[a] = ri[Rec<D>::B + a]; cri.gA[a] = ri[Rec<D>::GA + a]; }
#pragma unroll
  for (int a = 0; a < D * D; ++a) cri.gB[a] = ri[Rec<D>::GB + a];
  const float smi = hi * inv_kscale;
  const float kscale = 1.f / inv_kscale;
  const float ismi = 1.f / smi;
  const float epsi = 1.0e-7f * smi * smi;
  const float ViOm = Vi / mi;
  Comp acc[D], accu, accs;
  const int32_t b = pre.off, e = L.valid ? b + pre.cnt : b;
  // neighbour ids loaded two iterations ahead (the record itself is loaded
  // at the start of its iteration: from global memory a prefetch costs
  // more in occupancy than it hides, and from LDS, two record buffers in
  // turn were measured no faster)
  int32_t jc = b + L.g < e ? src.id(b + L.g) : 0;
  int32_t jn = b + L.g + G < e ? src.id(b + L.g + G) : 0;
  for (int32_t q0 = b + L.g; q0 < e; q0 += G) {
    const int32_t j = jc;
    float rj[Rec<D>::NF];
    src.template load<D>(j, rj);
    // (branch free: past the row's end the id of its last slot is read and
    // not used)
    jc = jn;
    jn = src.id(min(q0 + 2 * G, e - 1));
    float dx[3];
    pair_dx<D>(geo, xi, &rj[Rec<D>::X], dx);
    const float r2 = dot<D>(dx, dx);
    const float r = fsqrt(r2);
    const float hj = rj[Rec<D>::H];
    const float ihj = frcp(hj);
    const float* vj = &rj[Rec<D>::VEL];
    const float* Jj = &rj[Rec<D>::J];
    float cI[3], cJ[3];
#pragma unroll
    for (int a = 0; a < D; ++a) {
      cI[a] = 0.f;
      cJ[a] = 0.f;
#pragma unroll
      for (int gm = 0; gm < D; ++gm) { cI[a] = cI[a] + Ji[a * D + gm] * dx[gm]; cJ[a] = cJ[a] + Jj[a * D + gm] * dx[gm]; }
    }
    float phi = limiter_phi<D>(lp, dx, r, ihi, ihj, cI, cJ);
    float phi2 = 0.5f * phi;
    float vdi[3], vdj[3], vdjU[3];
#pragma unroll
    for (int a = 0; a < D; ++a) {
      vdi[a] = vi[a] - phi2 * cI[a];
      vdj[a] = vj[a] + phi2 * cJ[a];   // accel.py:137-141
      vdjU[a] = vj[a] - phi2 * cJ[a];  // dudt.py:123 (J_j (-x_ij))
    }
    CorrD<D> crj;
    crj.A = rj[Rec<D>::A];
#pragma unroll
    for (int a = 0; a < D; ++a) { crj.B[a] = rj[Rec<D>::B + a]; crj.gA[a] = rj[Rec<D>::GA + a]; }
#pragma unroll
    for (int a = 0; a < D * D; ++a) crj.gB[a] = rj[Rec<D>::GB + a];
    WdW wj = wdw<K, D>(C, r, ihj);  // W(x_ij, h_j) for the i-side gradient
    WdW wi = wdw<K, D>(C, r, ihi);  // W(x_ji, h_i) for the j-side gradient
    if (SW) {
      // (a select, not a branch: a zero term would still move the compensation)
      Comp t = accs;
      t.add(rj[Rec<D>::MSG] * wi.W);
      if (wi.W != 0.f) accs = t;
    }
    float gi[3], gjr[3], dxn[3];
#pragma unroll
    for (int a = 0; a < D; ++a) dxn[a] = -dx[a];
    grad_wr<D>(cri, dx, wj, gi);
    grad_wr<D>(crj, dxn, wi, gjr);
    float gw[3];
#pragma unroll
    for (int a = 0; a < D; ++a) gw[a] = 0.5f * (gi[a] - gjr[a]);
    // mu_i = (vdot_i - vdot_j).eta_i / (eta_i.eta_i + 1e-7 s_i^2), eta_i = x / s_i
    //      = s_i (w.x) / (r^2 + 1e-7 s_i^4) ... evaluated in the reference's form
    const float smj = hj * inv_kscale;
    const float ismj = ihj * kscale;  // 1 / s_j without a second reciprocal
    const float rdi = frcp(r2 * ismi * ismi + epsi);
    const float rdj = frcp(r2 * ismj * ismj + 1.0e-7f * smj * smj);
    float wx = 0.f, wxU = 0.f;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      wx = wx + (vdi[a] - vdj[a]) * dx[a];
      wxU = wxU + (vdi[a] - vdjU[a]) * dx[a];
    }
    const float fi = ismi * rdi, fj = ismj * rdj;
    float mui = fminf(wx * fi, 0.f), muj = fminf(wx * fj, 0.f);
    float muiU = fminf(wxU * fi, 0.f), mujU = fminf(wxU * fj, 0.f);
    // Q = rho (-C_l c mu + C_q mu^2) (accel.py:202-206, dudt.py) as mu (qa + qb mu)
    const float qaj = rj[Rec<D>::QA], qbj = rj[Rec<D>::QB];
    float Qi = mui * (qai + qbi * mui);
    float Qj = muj * (qaj + qbj * muj);
    float QiU = muiU * (qai + qbi * muiU);
    float QjU = mujU * (qaj + qbj * mujU);
    if (alphaQ) {
      float wa = 0.5f * (ali + in.alpha[j]);
      Qi = Qi * wa; Qj = Qj * wa; QiU = QiU * wa; QjU = QjU * wa;
    }
    const float VV = ViOm * rj[Rec<D>::V];
    const float Pj = rj[Rec<D>::P];
    const float pP = -(Pi + Pj) * VV;
    const float pQ = -(Qi + Qj) * VV;
    // a_ij = ap + av = (pP + pQ) gradW (accel.py:228,241); pP and pQ have one
    // sign (P, Q >= 0), so the viscous part is the fraction pQ / (pP + pQ)
    const float pS = pP + pQ;
    float aa[3];
#pragma unroll
    for (int a = 0; a < D; ++a) {
      aa[a] = pP * gw[a] + pQ * gw[a];
      acc[a].add(aa[a]);
    }
    store_slot<D>(o.pf, q0, aa);
    o.th[q0] = pS != 0.f ? pQ * frcp(pS) : 0.f;
    float dt_ = 0.f;
#pragma unroll
    for (int a = 0; a < D; ++a) dt_ = dt_ + (vi[a] - vj[a]) * gw[a];
    accu.add((Pj + 0.5f * (QiU + QjU)) * dt_ * VV);
  }
  float dv[D];
#pragma unroll
  for (int a = 0; a < D; ++a) dv[a] = group_total<G>(acc[a]);
  float du = group_total<G>(accu);
  float S = SW ? group_total<G>(accs) : 0.f;
  if (!L.valid || L.g != 0) return;
  float ke = 0.f;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    o.dvdt_s[(int64_t)k * D + a] = dv[a];
    o.dvdt_u[(int64_t)i * D + a] = dv[a];
    ke = ke + vi[a] * dv[a];
  }
  o.dudt_u[i] = du;
  o.dEdt_u[i] = mi * ke + mi * du;
  o.div_u[i] = -drhodt_i / rho_k;
  if (SW) cullen_row(cdr, cp, k, i, S, hi, cout);
}

template <int D, int G, int K, bool SW>
__global__ void __launch_bounds__(kBlock) k_accel_dudt(ARows rows, Geo geo, float C, LimParams lp, float inv_kscale,
                                                       int alphaQ, int32_t n, AdIn in,
                                                       const int32_t* __restrict__ perm, AdOut o, CdIn cin,
                                                       CdParams cp, CdOut cout, const float4* __restrict__ rec) {
  const Lane L = lane_of<G>(n);
  RowPre<D> pre;
  row_pre_load<D>(pre, rec, rows, L.k);
  accel_row<D, G, K, SW>(L, RecSrcG{rows.nbr, rec}, rows, geo, C, lp, inv_kscale, alphaQ, n, in, perm, o,
                         cin, cp, cout, rec, pre);
}

// The acceleration pass on an LDS tile: the block's threads (1024 by
// default) share one tile's staged records (NF floats per particle, d <= 2),
// G lanes per row, the tile's rows in rounds of blockDim / G.
template <int D, int G, int K, bool SW>
__global__ void __launch_bounds__(kTileThreadsMax) k_accel_dudt_t(ARows rows, const uint16_t* __restrict__ ploc,
                                                                const int32_t* __restrict__ tiles, Geo geo, float C,
                                                                LimParams lp, float inv_kscale,
                                                                int32_t n, AdIn in, const int32_t* __restrict__ perm,
                                                                AdOut o, CdIn cin, CdParams cp, CdOut cout,
                                                                const float4* __restrict__ rec) {
  static_assert(D <= 2, "tiled acceleration pass: d <= 2");
  constexpr int C4 = Rec<D>::NF / 4;
  __shared__ int32_t tab[kTileWords];
  __shared__ float4 srec[kTileCap * C4];
  __shared__ __align__(16) uint16_t sid[kTileIds + 8];
  const int32_t t = xcd_block();
  // the first round's rows: their records and CSR ranges load while the
  // tile is staged
  RowPre<D> pre0;
  row_pre_load<D>(pre0, rec, rows, tile_lane<G>(t, 0, n).k);
  const int32_t base = tile_stage<C4>(tiles, t, rows.off, n, ploc, rec, tab, srec, sid);
  const bool wide = tab[0] < 0;
  const int nr = tile_rounds<G>();
  for (int round = 0; round < nr; ++round) {
    const Lane L = tile_lane<G>(t, round, n);
    if (round > 0) row_pre_load<D>(pre0, rec, rows, L.k);
    if (wide)
      accel_row<D, G, K, SW>(L, RecSrcG{rows.nbr, rec}, rows, geo, C, lp, inv_kscale, 0, n, in, perm, o, cin,
                             cp, cout, rec, pre0);
    else
      accel_row<D, G, K, SW>(L, RecSrcT{sid, base, srec}, rows, geo, C, lp, inv_kscale, 0, n, in, perm, o, cin,
                             cp, cout, rec, pre0);
  }
}

// ------------------------------------------------------------ balance
// computeCompSPHBalanceTermWarp, CRK energy scheme (compSPH/balance.py:63-154)
// with v_half = v + dt/2 dvdt (crkSPH.py:341).  The per-particle inputs of the
// partition are packed into one record per sorted particle,
// rec = {v_half (D components, zero padded), s = P/rho^gamma}; f_ij itself is
// a function of rec_i, rec_j, m_i dt and a_ij = ap + av and is evaluated
// where it is consumed (the compatible-energy finalize, the export).
template <int D>
__global__ void k_balance_rec(int32_t n, const float* __restrict__ v, const float* __restrict__ dvdt,
                              const float* __restrict__ sP, DtArg dt, float4* __restrict__ rec) {
  int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float hdt = 0.5f * dt.at(k);
  float vh[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int d = 0; d < D; ++d) vh[d] = v[(int64_t)k * D + d] + hdt * dvdt[(int64_t)k * D + d];
  rec[k] = make_float4(vh[0], vh[1], vh[2], sP[k]);
}

// f_ij over the pair slots (optional output of crk_balance)
template <int D, int G>
__global__ void __launch_bounds__(kBlock) k_balance(ARows rows, int32_t n, const float4* __restrict__ rec,
                                                    const float* __restrict__ m, DtArg dt,
                                                    const float* __restrict__ pf, float* __restrict__ f) {
  Lane L = lane_of<G>(n);
  const int32_t k = L.k;
  const float4 ri = rec[k];
  const float mdt = m[k] * dt.at(k);
  const int32_t b = rows.off[k], e = L.valid ? b + rows.cnt[k] : b;
  int32_t jn = b + L.g < e ? rows.nbr[b + L.g] : 0;  // neighbour ids loaded one iteration ahead
  for (int32_t q = b + L.g; q < e; q += G) {
    const int32_t j = jn;
    if (q + G < e) jn = rows.nbr[q + G];
    float a[3];
#pragma unroll
    for (int c = 0; c < D; ++c) a[c] = pf[(int64_t)q * D + c];
    f[q] = pair_fraction<D>(ri, rec[j], mdt, a);
  }
}

}  // namespace crk

using namespace crk;

static CdParams cd_params(const CrkContext* ctx, const CrkConfig* cfg, float dt) {
  const CrkSwitchParams& sp = cfg->switchParams;
  return CdParams{dt_arg(ctx, dt), (float)(1.0 / kernel_xi(cfg->kernel, cfg->domain.dim)), sp.alpha_min, sp.alpha_max,
                  sp.beta_xi, sp.limitXi, cfg->evolveAlpha0};
}
static CdIn cd_in(CrkContext* ctx) {
  return CdIn{F<float>(ctx->x), F<float>(ctx->h), F<float>(ctx->m), F<float>(ctx->rho), F<float>(ctx->divcd),
              F<float>(ctx->vsig), F<float>(ctx->divp)};
}
static CdOut cd_out(CrkContext* ctx, const CrkState* st) {
  return CdOut{F<float>(ctx->a0), F<float>(ctx->alpha), F<float>(ctx->Rcd), st->alphas, st->alpha0s};
}

extern "C" {

int crk_velocity_gradient(CrkContext* ctx, const CrkConfig* cfg, CrkUpdate* up, void* stream) {
  int r = require_bound(ctx, cfg);
  if (r) return r;
  CRK_REQUIRE(up && up->drhodt, "crk_velocity_gradient: update->drhodt missing");
  if (!ctx->have_factors) { set_error("crk_velocity_gradient before crk_factors_eos"); return CRK_ESTATE; }
  hipStream_t s = (hipStream_t)stream;
  int32_t n = ctx->n;
  Geo geo = make_geo(*cfg);
  float C = (float)kernel_C(cfg->kernel, cfg->domain.dim);
  bool cullen = cfg->switchParams.scheme == CRK_SWITCH_CULLEN_DEHNEN || cfg->switchParams.scheme == CRK_SWITCH_LEARNED;
  VgIn in{F<float>(ctx->x), F<float>(ctx->h), F<float>(ctx->v), F<float>(ctx->V), F<float>(ctx->m),
          F<float>(ctx->rho), F<float>(ctx->c), F<float>(ctx->A), F<float>(ctx->B), F<float>(ctx->gA),
          F<float>(ctx->gB), F<float4>(ctx->geo4), F<float4>(ctx->vg4), F<float>(ctx->mr)};
  VgOut out{F<float>(ctx->J), F<float>(ctx->divcd), F<float>(ctx->vsig), up->drhodt};
  int GS = ctx->group[CRK_PASS_VELGRAD];
  const int TB = ctx->threads[CRK_PASS_VELGRAD];
  if (use_tiles(ctx) && ctx->pair_loc) {
    if (cullen) {
      CRK_KDG12_DISPATCH(cfg->kernel, ctx->dim, GS,
          hipLaunchKernelGGL((k_velgrad_t<DD, G, KK, true>), dim3(tile_count(n)), dim3(TB), 0, s,
                             arows_of(ctx), ctx->pair_loc, ctx->adj.tiles, geo, C, n, in, ctx->adj.perm, out));
    } else {
      CRK_KDG12_DISPATCH(cfg->kernel, ctx->dim, GS,
          hipLaunchKernelGGL((k_velgrad_t<DD, G, KK, false>), dim3(tile_count(n)), dim3(TB), 0, s,
                             arows_of(ctx), ctx->pair_loc, ctx->adj.tiles, geo, C, n, in, ctx->adj.perm, out));
    }
  } else if (cullen) {
    CRK_KDG_DISPATCH(cfg->kernel, ctx->dim, GS,
        hipLaunchKernelGGL((k_velgrad<DD, G, KK, true>), dim3(group_blocks(n, G)), dim3(kBlock), 0, s,
                           arows_of(ctx), geo, C, n, in, ctx->adj.perm, out));
  } else {
    CRK_KDG_DISPATCH(cfg->kernel, ctx->dim, GS,
        hipLaunchKernelGGL((k_velgrad<DD, G, KK, false>), dim3(group_blocks(n, G)), dim3(kBlock), 0, s,
                           arows_of(ctx), geo, C, n, in, ctx->adj.perm, out));
  }
  CRK_LAUNCH_CHECK("k_velgrad");
  ctx->have_velgrad = true;
  ctx->have_divcd = cullen;
  return CRK_OK;
}

int crk_switch(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st, float dt, void* stream) {
  int r = require_bound(ctx, cfg);
  if (r) return r;
  CRK_REQUIRE(st && st->alphas, "crk_switch: state->alphas missing");
  CRK_REQUIRE(!cfg->evolveAlpha0 || st->alpha0s, "crk_switch: state->alpha0s missing");
  CRK_REQUIRE(dt > 0.f || dt_arg(ctx, dt).p, "crk_switch: dt must be positive");
  if (!ctx->have_velgrad) { set_error("crk_switch before crk_velocity_gradient"); return CRK_ESTATE; }
  if (cfg->switchParams.scheme != CRK_SWITCH_CULLEN_DEHNEN && cfg->switchParams.scheme != CRK_SWITCH_LEARNED) {
    set_error("crk_switch: switch scheme %d is not the Cullen-Dehnen switch", cfg->switchParams.scheme);
    return CRK_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  int32_t n = ctx->n;
  Geo geo = make_geo(*cfg);
  float C = (float)kernel_C(cfg->kernel, cfg->domain.dim);
  int GS = ctx->group[CRK_PASS_SWITCH];
  CRK_KDG_DISPATCH(cfg->kernel, ctx->dim, GS,
      hipLaunchKernelGGL((k_cullen<DD, G, KK>), dim3(group_blocks(n, G)), dim3(kBlock), 0, s, arows_of(ctx), geo, C,
                         n, cd_in(ctx), cd_params(ctx, cfg, dt), ctx->adj.perm, cd_out(ctx, st)));
  CRK_LAUNCH_CHECK("k_cullen");
  ctx->have_rcd = true;
  return CRK_OK;
}

static int accel_dudt(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st, CrkUpdate* up, CrkEdges* ed,
                      bool sw, float dt, void* stream) {
  int r = require_bound(ctx, cfg);
  if (r) return r;
  CRK_REQUIRE(st && st->divergence, "crk_accel_dudt: state->divergence missing");
  CRK_REQUIRE(up && up->dvdt && up->dudt && up->dEdt && up->drhodt, "crk_accel_dudt: update buffers missing");
  CRK_REQUIRE(ed && (ctx->adj.n_edges == 0 || (ed->pf && ed->th)), "crk_accel_dudt: pair-force buffers missing");
  CRK_REQUIRE(ed->count == ctx->pair_cnt && ed->nbr == ctx->pair_nbr,
              "crk_accel_dudt: edges differ from the pairs of crk_support");
  if (!ctx->have_velgrad) { set_error("crk_accel_dudt before crk_velocity_gradient"); return CRK_ESTATE; }
  hipStream_t s = (hipStream_t)stream;
  int32_t n = ctx->n;
  Geo geo = make_geo(*cfg);
  float C = (float)kernel_C(cfg->kernel, cfg->domain.dim);
  LimParams lp{cfg->crkViscosity.eta_fold, 1.f / cfg->crkViscosity.eta_fold, cfg->crkViscosity.eta_crit,
               cfg->crkViscosity.enableCRKLimiter, cfg->crkViscosity.enableVanLeerLimiter,
               cfg->crkViscosity.forceVanLeerOff, cfg->crkViscosity.forceVanLeerOn};
  float inv_kscale = (float)(1.0 / kernel_scale(cfg->kernel, cfg->domain.dim));
  AdIn in{F<float>(ctx->x), F<float>(ctx->h), F<float>(ctx->v), F<float>(ctx->m), F<float>(ctx->V),
          F<float>(ctx->rho), F<float>(ctx->P), F<float>(ctx->c), F<float>(ctx->J), F<float>(ctx->A),
          F<float>(ctx->B), F<float>(ctx->gA), F<float>(ctx->gB), F<float>(ctx->alpha), (const float*)up->drhodt};
  AdOut out{ed->pf, ed->th, F<float>(ctx->dvdt), up->dvdt, up->dudt, up->dEdt, st->divergence};
  // staged on tiles unless alpha-weighted Q reads alpha_j through the sorted ids
  // (or a group of fewer than 4 lanes would leave a round beyond the tile)
  const bool tiled = use_tiles(ctx) && ctx->pair_loc && !cfg->alphaWeightedQ && ctx->group[CRK_PASS_ACCEL] >= 4;
  int GS = ctx->group[CRK_PASS_ACCEL];
  const float* divcd = (cfg->switchParams.scheme == CRK_SWITCH_CULLEN_DEHNEN ||
                        cfg->switchParams.scheme == CRK_SWITCH_LEARNED) ? F<float>(ctx->divcd) : nullptr;
  CRK_DIM_DISPATCH(ctx->dim, hipLaunchKernelGGL(k_pack_rec<DD>, dim3(nblocks(n)), dim3(kBlock), 0, s, n, in, divcd,
                                                cfg->diffusion.C_l, cfg->diffusion.C_q,
                                                F<float4>(ctx->rec)));
  CRK_LAUNCH_CHECK("k_pack_rec");
  if (tiled) {
#define ACCEL_T(SWV, CI, CP, CO)                                                                                 \
    CRK_KDG12_DISPATCH(cfg->kernel, ctx->dim, GS,                                                              \
        hipLaunchKernelGGL((k_accel_dudt_t<DD, (G < 4 ? 4 : G), KK, SWV>), dim3(tile_count(n)),                 \
                           dim3(ctx->threads[CRK_PASS_ACCEL]), 0, s, arows_of(ctx), ctx->pair_loc,             \
                           ctx->adj.tiles, geo, C, lp,                                                          \
                           inv_kscale, n, in, ctx->adj.perm, out, CI, CP,                                       \
                           CO, F<float4>(ctx->rec)))
    if (sw) { ACCEL_T(true, cd_in(ctx), cd_params(ctx, cfg, dt), cd_out(ctx, st)); }
    else { ACCEL_T(false, CdIn{}, CdParams{}, CdOut{}); }
#undef ACCEL_T
  } else if (sw) {
    CRK_KDG_DISPATCH(cfg->kernel, ctx->dim, GS,
        hipLaunchKernelGGL((k_accel_dudt<DD, G, KK, true>), dim3(group_blocks(n, G)), dim3(kBlock), 0, s,
                           arows_of(ctx), geo, C, lp, inv_kscale,
                           cfg->alphaWeightedQ, n, in, ctx->adj.perm, out, cd_in(ctx), cd_params(ctx, cfg, dt),
                           cd_out(ctx, st), F<float4>(ctx->rec)));
  } else {
    CRK_KDG_DISPATCH(cfg->kernel, ctx->dim, GS,
        hipLaunchKernelGGL((k_accel_dudt<DD, G, KK, false>), dim3(group_blocks(n, G)), dim3(kBlock), 0, s,
                           arows_of(ctx), geo, C, lp, inv_kscale,
                           cfg->alphaWeightedQ, n, in, ctx->adj.perm, out, CdIn{}, CdParams{}, CdOut{},
                           F<float4>(ctx->rec)));
  }
  CRK_LAUNCH_CHECK("k_accel_dudt");
  ctx->have_accel = true;
  if (sw) ctx->have_rcd = true;
  return CRK_OK;
}

int crk_accel_dudt(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st, CrkUpdate* up, CrkEdges* ed,
                   void* stream) {
  return accel_dudt(ctx, cfg, st, up, ed, false, 0.f, stream);
}

int crk_switch_accel_dudt(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st, CrkUpdate* up, CrkEdges* ed,
                          float dt, void* stream) {
  int r = require_bound(ctx, cfg);
  if (r) return r;
  CRK_REQUIRE(cfg->switchParams.scheme == CRK_SWITCH_CULLEN_DEHNEN,
              "crk_switch_accel_dudt: switch scheme %d is not the Cullen-Dehnen switch", cfg->switchParams.scheme);
  CRK_REQUIRE(!cfg->alphaWeightedQ, "crk_switch_accel_dudt: alpha-weighted Q reads the switch; call crk_switch "
                                    "then crk_accel_dudt");
  CRK_REQUIRE(st && st->alphas, "crk_switch_accel_dudt: state->alphas missing");
  CRK_REQUIRE(!cfg->evolveAlpha0 || st->alpha0s, "crk_switch_accel_dudt: state->alpha0s missing");
  CRK_REQUIRE(dt > 0.f || dt_arg(ctx, dt).p, "crk_switch_accel_dudt: dt must be positive");
  return accel_dudt(ctx, cfg, st, up, ed, true, dt, stream);
}

static int features_launch(CrkContext* ctx, const CrkConfig* cfg, float dt, const CrkMlp* model, float* raw,
                           hipStream_t s) {
  if (!ctx->have_velgrad || !ctx->have_support) {
    set_error("learned switch features need crk_support, crk_factors_eos and crk_velocity_gradient");
    return CRK_ESTATE;
  }
  const int32_t n = ctx->n;
  const int D = ctx->dim;
  int st;
  if ((st = ensure(ctx->gnorm, sizeof(float2) * (size_t)n))) return st;
  Geo geo = make_geo(*cfg);
  float C = (float)kernel_C(cfg->kernel, D);
  VgIn in{F<float>(ctx->x), F<float>(ctx->h), F<float>(ctx->v), F<float>(ctx->V), F<float>(ctx->m),
          F<float>(ctx->rho), F<float>(ctx->c), F<float>(ctx->A), F<float>(ctx->B), F<float>(ctx->gA),
          F<float>(ctx->gB)};
  constexpr int GG = 8;
  switch (cfg->kernel) {
#define GN_K(KV)                                                                                                  \
  case KV:                                                                                                        \
    CRK_DIM_DISPATCH(D, hipLaunchKernelGGL((k_gradnorm<DD, GG, KV>), dim3(group_blocks(n, GG)), dim3(kBlock), 0, s, \
                                           arows_of(ctx), geo, C, n, in, F<float>(ctx->P), F<float2>(ctx->gnorm))); \
    break;
    GN_K(0) GN_K(1) GN_K(2) GN_K(3)
#undef GN_K
    default: set_error("unknown kernel %d", cfg->kernel); return CRK_EINVAL;
  }
  CRK_LAUNCH_CHECK("k_gradnorm");
  FeatIn fi{F<float>(ctx->h), F<float>(ctx->c), F<float>(ctx->rho), F<float>(ctx->m), F<float>(ctx->P),
            F<float>(ctx->divcd), F<float>(ctx->divp), F<float>(ctx->J), F<float>(ctx->vsig), F<float2>(ctx->gnorm),
            ctx->pair_cnt};
  CrkMlp mm = model ? *model : CrkMlp{};
  float* xn = nullptr;
  uint8_t* fl = nullptr;
  if (!raw) {
    if ((st = ensure(ctx->feat, sizeof(float) * 8 * (size_t)n))) return st;
    if ((st = ensure(ctx->fflag, (size_t)n))) return st;
    xn = F<float>(ctx->feat);
    fl = F<uint8_t>(ctx->fflag);
  }
  CRK_DIM_DISPATCH(D, hipLaunchKernelGGL(k_features<DD>, dim3(nblocks(n)), dim3(kBlock), 0, s, n, fi,
                                         dt_arg(ctx, dt), cfg->targetNeighbors, mm, ctx->adj.perm, raw, xn, fl));
  CRK_LAUNCH_CHECK("k_features");
  return CRK_OK;
}

int crk_learned_features(CrkContext* ctx, const CrkConfig* cfg, float dt, float* features_out, void* stream) {
  int r = require_bound(ctx, cfg);
  if (r) return r;
  CRK_REQUIRE(features_out, "crk_learned_features: null output");
  CRK_REQUIRE(dt > 0.f || dt_arg(ctx, dt).p, "crk_learned_features: dt must be positive");
  CRK_REQUIRE(cfg->switchParams.scheme == CRK_SWITCH_CULLEN_DEHNEN || cfg->switchParams.scheme == CRK_SWITCH_LEARNED,
              "crk_learned_features: the features need the Cullen-Dehnen divergence and signal speed");
  return features_launch(ctx, cfg, dt, nullptr, features_out, (hipStream_t)stream);
}

int crk_learned_switch(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st, float dt, const CrkMlp* model,
                       void* stream) {
  int r = require_bound(ctx, cfg);
  if (r) return r;
  CRK_REQUIRE(model && model->params && model->wbf16, "crk_learned_switch: model missing");
  for (int q = 0; q < 8; ++q)
    CRK_REQUIRE(model->sigma[q] > 0.f && model->lo[q] <= model->hi[q], "crk_learned_switch: bad normalisation");
  CRK_REQUIRE(cfg->switchParams.scheme == CRK_SWITCH_LEARNED, "crk_learned_switch: switch scheme is not learned");
  // the Cullen-Dehnen switch first: the fallback alpha (and alpha0 evolution)
  if ((r = crk_switch(ctx, cfg, st, dt, stream))) return r;
  hipStream_t s = (hipStream_t)stream;
  if ((r = features_launch(ctx, cfg, dt, model, nullptr, s))) return r;
  CRK_REQUIRE(model->scale > 0.f, "crk_learned_switch: model scale must be positive");
  return mlp_infer_launch(model, ctx->n, F<float>(ctx->feat), F<uint8_t>(ctx->fflag), F<float>(ctx->alpha),
                          F<float>(ctx->alpha), ctx->adj.perm, st->alphas, s, false);
}

int crk_balance(CrkContext* ctx, const CrkConfig* cfg, float dt, CrkEdges* ed, void* stream) {
  int r = require_bound(ctx, cfg);
  if (r) return r;
  CRK_REQUIRE(ed && ed->rec, "crk_balance: edges->rec missing");
  CRK_REQUIRE(!ed->f || ctx->adj.n_edges == 0 || ed->pf, "crk_balance: pair-force buffer missing");
  CRK_REQUIRE(ed->count == ctx->pair_cnt && ed->nbr == ctx->pair_nbr,
              "crk_balance: edges differ from the pairs of crk_support");
  if (!ctx->have_accel) { set_error("crk_balance before crk_accel_dudt"); return CRK_ESTATE; }
  hipStream_t s = (hipStream_t)stream;
  int32_t n = ctx->n;
  const DtArg dta = dt_arg(ctx, dt);
  float4* rec = (float4*)ed->rec;
  CRK_DIM_DISPATCH(ctx->dim, hipLaunchKernelGGL(k_balance_rec<DD>, dim3(nblocks(n)), dim3(kBlock), 0, s, n,
                                                F<float>(ctx->v), F<float>(ctx->dvdt), F<float>(ctx->sP), dta,
                                                rec));
  CRK_LAUNCH_CHECK("k_balance_rec");
  if (ed->f && ctx->adj.n_edges > 0) {
    int GS = ctx->group[CRK_PASS_BALANCE];
    CRK_DIM_G_DISPATCH(ctx->dim, GS,
        hipLaunchKernelGGL((k_balance<DD, G>), dim3(group_blocks(n, G)), dim3(kBlock), 0, s, arows_of(ctx), n, rec,
                           F<float>(ctx->m), dta, ed->pf, ed->f));
    CRK_LAUNCH_CHECK("k_balance");
  }
  return CRK_OK;
}

}  // extern "C"
