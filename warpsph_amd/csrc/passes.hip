// passes.hip -- the pair passes of one CRKSPH derivative evaluation
// (src/warpSPH/schemes/crkSPH.py:19-366), run on Morton-sorted copies of the
// state held by the context.
//
// Pass order inside a stage:
//   crk_stage_begin      gather x, v, h, m, u, alpha0, alpha, divergence to sorted order
//   crk_owen_iteration   psi0 -> n_h -> new h           (optimalSupportOwen.py:70-114)
//   crk_support          V, signal speed, in-support pairs compacted per row
//   crk_factors_eos      moments -> A, B, gradA, gradB (FP64 solve),
//                        rho = A (S0 + B.S1), EOS      (computeCRKFactors; idealGas.py)
//   crk_velocity_gradient J (CRK, scatter), drhodt, and the Cullen divergence
//                        and signal speed              (crkSPH.py:158-172; wp_vsig.py)
//   crk_switch           R, Xi, alpha update           (CullenDehnen2010.py:218-331)
//   crk_accel_dudt       dvdt, ap, av, dudt, dEdt      (crk/accel.py, crk/dudt.py)
//   crk_balance          f_ij                          (compSPH/balance.py)
#include <cmath>

#include "tiles.hpp"

namespace crk {

// ------------------------------------------------------------ gather
template <int D>
__global__ void k_gather(const int32_t* __restrict__ perm, int32_t n, const float* __restrict__ X,
                         const float* __restrict__ Vel, const float* __restrict__ H, const float* __restrict__ M,
                         const float* __restrict__ U, const float* __restrict__ A0, const float* __restrict__ DIV,
                         const float* __restrict__ AL, float* x, float* v, float* h, float* m, float* u, float* a0,
                         float* divp, float* alpha, float cfac, float4* __restrict__ geo4,
                         float4* __restrict__ vel4) {
  int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int32_t i = perm[k];
  float xs[3] = {0.f, 0.f, 0.f}, vs[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int d = 0; d < D; ++d) {
    xs[d] = X[(int64_t)i * D + d];
    vs[d] = Vel[(int64_t)i * D + d];
    x[(int64_t)k * D + d] = xs[d];
    v[(int64_t)k * D + d] = vs[d];
  }
  const float hk = H[i], uk = U[i];
  h[k] = hk;
  m[k] = M[i];
  u[k] = uk;
  geo4[k] = make_float4(xs[0], xs[1], xs[2], hk);
  vel4[k] = make_float4(vs[0], vs[1], vs[2], sqrtf(fabsf(uk) * cfac));
  a0[k] = A0 ? A0[i] : 1.f;
  divp[k] = DIV ? DIV[i] : 0.f;
  alpha[k] = AL ? AL[i] : 1.f;
}

// The gather of crk_stage_begin fused with the Verlet validity test of the
// bound adjacency (built at xb, hb) and the support range for Owen, per run
// r: red[4r] = max |x - xb|^2 (exact), red[4r+1] = max h / hb (exact),
// red[4r+2] = min hb; hmm[2r] = min h, hmm[2r+1] = max h.  Ensembles: one
// atomic per block and cell.  One run (part != null): each block writes its
// five values to part[8 b ..] and k_check_finish reduces them (no contended
// atomics, so the grid can be large).
template <int D>
__global__ void k_gather_check(const int32_t* __restrict__ perm, int32_t n, Geo geo, const float* __restrict__ X,
                               const float* __restrict__ Vel, const float* __restrict__ H,
                               const float* __restrict__ M, const float* __restrict__ U,
                               const float* __restrict__ A0, const float* __restrict__ DIV,
                               const float* __restrict__ AL, float* x, float* v, float* h, float* m, float* u,
                               float* a0, float* divp, float* alpha, float cfac, float4* __restrict__ geo4,
                               float4* __restrict__ vel4, const float* __restrict__ xb,
                               const float* __restrict__ hb, float* __restrict__ red, float* __restrict__ hmm,
                               RunsDev R, float* __restrict__ part) {
  float d2 = 0.f, g = 0.f, hbm = 3.0e38f, hlo = 3.0e38f, hhi = 0.f;
  const Span sp = block_span(R, n);
  for (int32_t k = sp.k; k < sp.e; k += sp.step) {
    const int32_t i = perm[k];
    float xs[3] = {0.f, 0.f, 0.f}, vs[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < D; ++d) {
      xs[d] = X[(int64_t)i * D + d];
      vs[d] = Vel[(int64_t)i * D + d];
      x[(int64_t)k * D + d] = xs[d];
      v[(int64_t)k * D + d] = vs[d];
    }
    const float hk = H[i], uk = U[i], hbk = hb[k];
    h[k] = hk;
    m[k] = M[i];
    u[k] = uk;
    geo4[k] = make_float4(xs[0], xs[1], xs[2], hk);
    vel4[k] = make_float4(vs[0], vs[1], vs[2], sqrtf(fabsf(uk) * cfac));
    a0[k] = A0 ? A0[i] : 1.f;
    divp[k] = DIV ? DIV[i] : 0.f;
    alpha[k] = AL ? AL[i] : 1.f;
    d2 = fmaxf(d2, exact_r2<D>(geo, xs, &xb[(int64_t)k * D]));
    g = fmaxf(g, __fdiv_rn(hk, hbk));
    hbm = fminf(hbm, hbk);
    hlo = fminf(hlo, hk);
    hhi = fmaxf(hhi, hk);
  }
  d2 = block_reduce(d2, MaxOp());
  g = block_reduce(g, MaxOp());
  hbm = block_reduce(hbm, MinOp());
  hlo = block_reduce(hlo, MinOp());
  hhi = block_reduce(hhi, MaxOp());
  if (part) {
    if (threadIdx.x == 0) {
      float* o = part + 8 * (int64_t)blockIdx.x;
      o[0] = d2;
      o[1] = g;
      o[2] = hbm;
      o[3] = hlo;
      o[4] = hhi;
    }
    return;
  }
  if (threadIdx.x == 0) {
    atomic_max_pos(&red[4 * sp.r], d2);
    atomic_max_pos(&red[4 * sp.r + 1], g);
    atomic_min_pos(&red[4 * sp.r + 2], hbm);
    atomic_min_pos(&hmm[2 * sp.r], hlo);
    atomic_max_pos(&hmm[2 * sp.r + 1], hhi);
  }
}
// the block values of a single-run checked gather into red[0..2], hmm[0..1]
__global__ void k_check_finish(const float* __restrict__ part, int32_t nb, float* red, float* hmm) {
  float d2 = 0.f, g = 0.f, hbm = 3.0e38f, hlo = 3.0e38f, hhi = 0.f;
  for (int32_t b = threadIdx.x; b < nb; b += blockDim.x) {
    const float* o = part + 8 * (int64_t)b;
    d2 = fmaxf(d2, o[0]);
    g = fmaxf(g, o[1]);
    hbm = fminf(hbm, o[2]);
    hlo = fminf(hlo, o[3]);
    hhi = fmaxf(hhi, o[4]);
  }
  d2 = block_reduce(d2, MaxOp());
  g = block_reduce(g, MaxOp());
  hbm = block_reduce(hbm, MinOp());
  hlo = block_reduce(hlo, MinOp());
  hhi = block_reduce(hhi, MaxOp());
  if (threadIdx.x == 0) {
    red[0] = d2;
    red[1] = g;
    red[2] = hbm;
    hmm[0] = hlo;
    hmm[1] = hhi;
  }
}
// grid of the single-run checked gather (block values reduced by
// k_check_finish): 1M particles in 2D, per stage, 66 us at 1024 blocks,
// 67 at 2048 or 4096 (81 us with 512 blocks of atomics)
constexpr unsigned kCheckPartBlocks = 1024u;
// grid of the checked gather: two blocks per CU.  The blocks' atomics on the
// same few cells serialise in L2, so fewer blocks with longer grid-stride
// loops are faster (1M particles in 2D, per call: 1024 blocks 90 us, 512
// blocks 81 us, 2048 blocks 112 us)
inline unsigned check_blocks(int32_t n) {
  unsigned b = nblocks(n);
  return b < 512u ? b : 512u;
}
// max h / hb of each run over the sorted particles (after the Owen update)
// into red[4r+3]
__global__ void k_hb_ratio(const float* __restrict__ h, const float* __restrict__ hb, int32_t n, float* red,
                           RunsDev R) {
  float g = 0.f;
  const Span sp = block_span(R, n);
  for (int32_t k = sp.k; k < sp.e; k += sp.step) g = fmaxf(g, __fdiv_rn(h[k], hb[k]));
  g = block_reduce(g, MaxOp());
  if (threadIdx.x == 0) atomic_max_pos(&red[4 * sp.r + 3], g);
}
__global__ void k_check_init(float* red, float* hmm, int32_t nr) {
  for (int32_t r = threadIdx.x; r < nr; r += blockDim.x) {
    red[4 * r] = 0.f;
    red[4 * r + 1] = 0.f;
    red[4 * r + 2] = 3.0e38f;
    red[4 * r + 3] = 0.f;  // max h / hb after Owen
    hmm[2 * r] = 3.0e38f;
    hmm[2 * r + 1] = 0.f;
  }
}

// ------------------------------------------------------------ support sweep
// One pass over the full adjacency after the Owen update: the CRK apparent
// volume V_i = 1 / sum_j W(x_ij, h_i); the in-support pairs compacted per row
// with group ballots; and the signal speed
// v_sig = max_j (0.5 (c_i + c_j) - mu_ij if mu_ij <= 0 else 0),
// mu_ij = (v_i - v_j).x_ij / (r_ij + 1e-14 h_i), over every entry
// (wp_vsig.py:55-86), c = sqrt(|u| gamma (gamma - 1)).

// The support sweep on an LDS tile.  The neighbour's {x, h} and {v, c}
// records and the tile's row offsets are staged in LDS and the entries are
// read through their 16-bit local indices only; each compacted pair keeps
// its local index (for the tiled passes after this one) and its sorted id
// (for every other reader of the pairs; the staged {x, h} records carry it).  A wide
// tile reads and writes the 32-bit ids.  Each round's row heads (entry
// range, the first two indices, the row's own records) are loaded during the
// round before.
struct SupHead {
  int32_t b, e, j0, j1;
  float4 gk, wk;
};
template <int G, class Src>
__device__ inline SupHead support_head(const Lane& L, const Src& src, const int32_t* soff, int32_t r0,
                                       const float4* __restrict__ geo4, const float4* __restrict__ vel4) {
  SupHead h;
  h.b = soff[L.k - r0];
  h.e = L.valid ? soff[L.k - r0 + 1] : h.b;
  h.j0 = h.b + L.g < h.e ? src.id(h.b + L.g) : 0;
  h.j1 = h.b + L.g + G < h.e ? src.id(h.b + L.g + G) : 0;
  h.gk = geo4[L.k];
  h.wk = vel4[L.k];
  return h;
}
template <int D, int G, int K, bool VS, class Src>
__device__ inline void support_row_t(const Lane& L, const Src& src, const SupHead& hd, Geo geo, float C,
                                     int32_t* __restrict__ anbr, uint16_t* __restrict__ aloc,
                                     int32_t* __restrict__ acnt, float* __restrict__ vsig,
                                     float* __restrict__ V, const float* __restrict__ m, float2* __restrict__ vm2) {
  const int32_t k = L.k;
  const float4 gk = hd.gk, wk = hd.wk;
  const float xi[3] = {gk.x, gk.y, gk.z}, vi[3] = {wk.x, wk.y, wk.z};
  const float hi = gk.w;
  const float ih = 1.f / hi;
  const float ci = wk.w;
  const float eps = 1.0e-14f * hi;
  const int lane = threadIdx.x & 63;
  const int gbase = lane & ~(G - 1);
  const unsigned long long gmask = G == 64 ? ~0ull : (((1ull << G) - 1ull) << gbase);
  const unsigned long long lt = (1ull << lane) - 1ull;
  const int32_t b = hd.b, e = hd.e;
  int32_t run = 0;
  float vs = 0.f;
  Sum vol;
  int32_t jc = hd.j0, jn = hd.j1;
  float4 gn = src.geo(jc), wn = VS ? src.aux(jc) : make_float4(0.f, 0.f, 0.f, 0.f);
  for (int32_t base = b; base < e; base += G) {
    const int32_t q = base + L.g;
    const bool inside = q < e;
    bool act = false;
    int32_t j = 0, jg = 0;
    if (inside) {
      j = jc;
      const float4 gj = gn;
      const float4 wj = wn;
      jg = aloc ? __float_as_int(gj.z) : j;  // staged records carry the sorted id
      if (q + G < e) {
        jc = jn;
        gn = src.geo(jc);
        if (VS) wn = src.aux(jc);
        if (q + 2 * G < e) jn = src.id(q + 2 * G);
      }
      const float xj[3] = {gj.x, gj.y, D == 3 ? gj.z : 0.f};
      float dx[3];
      pair_dx<D>(geo, xi, xj, dx);
      float r = fsqrt(dot<D>(dx, dx));
      if (VS) {
        const float vj[3] = {wj.x, wj.y, wj.z};
        float mu = 0.f;
#pragma unroll
        for (int a = 0; a < D; ++a) mu = mu + (vi[a] - vj[a]) * dx[a];
        mu = mu * frcp(r + eps);
        float vv = 0.5f * (ci + wj.w) - mu;
        if (mu > 0.f) vv = 0.f;
        vs = vv > vs ? vv : vs;
      }
      float qq = qmul(r, ih);
      if (qq < 1.f) {
        float kk, dk;
        kwd<K, D>(qq, kk, dk);
        vol.fma(C * kk, ipow<D>(ih));
      }
      act = r < fmaxf(hi, gj.w);
    }
    const unsigned long long mm = __ballot(act ? 1 : 0) & gmask;
    if (act) {
      const int32_t slot = b + run + __popcll(mm & lt);
      anbr[slot] = jg;
      if (aloc) aloc[slot] = (uint16_t)j;  // a staged tile: the local index too
    }
    run += __popcll(mm);
  }
  vs = group_max<G>(vs);
  float Sv = group_total<G>(vol);
  if (!L.valid || L.g != 0) return;
  acnt[k] = run;
  const float Vk = 1.f / Sv;
  V[k] = Vk;
  vm2[k] = make_float2(Vk, m[k]);
  if (VS) vsig[k] = vs;
}

template <int D, int G, int K, bool VS>
__global__ void __launch_bounds__(kTileThreadsMax) k_support_t(Rows rows, const uint16_t* __restrict__ nloc,
                                                      const int32_t* __restrict__ tiles, Geo geo, float C, int32_t n,
                                                      int32_t* __restrict__ anbr, uint16_t* __restrict__ aloc,
                                                      int32_t* __restrict__ acnt, float* __restrict__ vsig,
                                                      float* __restrict__ V, const float4* __restrict__ geo4,
                                                      const float4* __restrict__ vel4, const float* __restrict__ m,
                                                      float2* __restrict__ vm2, const int32_t* __restrict__ gate) {
  __shared__ int32_t tab[kTileWords];
  __shared__ int32_t soff[kTileRows + 1];
  __shared__ float4 sg[kTileCap];
  __shared__ float4 sw[VS ? kTileCap : 1];
  __shared__ __align__(16) uint16_t sid[kTileIds + 8];
  const int32_t t = xcd_block();
  const int32_t r0 = t * kTileRows;
  if (gate && *gate) {  // (the list failed its deferred test: no pairs)
    for (int w = threadIdx.x; w < kTileRows; w += blockDim.x)
      if (r0 + w < n) acnt[r0 + w] = 0;
    return;
  }
  for (int w = threadIdx.x; w <= kTileRows; w += blockDim.x) soff[w] = rows.off[min(r0 + w, n)];
  tile_table(tiles, t, tab);
  const bool wide = tab[0] < 0;
  const int32_t base = wide ? 0 : tile_ids_fill(nloc, rows.off[r0], rows.off[min(n, r0 + kTileRows)], sid);
  if (!wide) {
    // the staged {x, h} record carries the particle's sorted id in its unused
    // z component (d <= 2), for the compacted pairs' 32-bit ids (a staging
    // with several particles' loads in flight per thread, tile_stage_fn, was
    // measured slower for this pass)
    const int32_t U = tab[1];
    for (int32_t u = threadIdx.x; u < U; u += blockDim.x) {
      const int32_t g = tile_global(tab, u);
      float4 a = geo4[g];
      a.z = __int_as_float(g);
      sg[u] = a;
      if (VS) sw[u] = vel4[g];
    }
    __syncthreads();
  }
  auto sweep = [&](const auto& src, uint16_t* lout) {
    const int nr = tile_rounds<G>();
    Lane L = tile_lane<G>(t, 0, n);
    SupHead nxt = support_head<G>(L, src, soff, r0, geo4, vel4);
    for (int round = 0; round < nr; ++round) {
      const Lane Lc = L;
      const SupHead cur = nxt;
      if (round + 1 < nr) {
        L = tile_lane<G>(t, round + 1, n);
        nxt = support_head<G>(L, src, soff, r0, geo4, vel4);
      }
      support_row_t<D, G, K, VS>(Lc, src, cur, geo, C, anbr, lout, acnt, vsig, V, m, vm2);
    }
  };
  if (wide) sweep(GlobalSrc<float4>{rows.nbr, geo4, vel4}, nullptr);
  else sweep(TileSrc<float4>{sid, base, nullptr, sg, sw}, aloc);
}

// The support sweep without tiles: one row per lane group, the neighbours'
// records gathered through the sorted ids (the same row code as a wide tile).
template <int D, int G, int K, bool VS>
__global__ void __launch_bounds__(kBlock) k_support(Rows rows, Geo geo, float C, int32_t n,
                                                    int32_t* __restrict__ anbr, int32_t* __restrict__ acnt,
                                                    float* __restrict__ vsig, float* __restrict__ V,
                                                    const float4* __restrict__ geo4, const float4* __restrict__ vel4,
                                                    const float* __restrict__ m, float2* __restrict__ vm2,
                                                    const int32_t* __restrict__ gate) {
  const Lane L = lane_of<G>(n);
  if (gate && *gate) {  // (the list failed its deferred test: no pairs)
    if (L.valid && L.g == 0) acnt[L.k] = 0;
    return;
  }
  const GlobalSrc<float4> src{rows.nbr, geo4, vel4};
  SupHead hd;
  hd.b = rows.off[L.k];
  hd.e = L.valid ? rows.off[L.k + 1] : hd.b;
  hd.j0 = hd.b + L.g < hd.e ? src.id(hd.b + L.g) : 0;
  hd.j1 = hd.b + L.g + G < hd.e ? src.id(hd.b + L.g + G) : 0;
  hd.gk = geo4[L.k];
  hd.wk = vel4[L.k];
  support_row_t<D, G, K, VS>(L, src, hd, geo, C, anbr, nullptr, acnt, vsig, V, m, vm2);
}

// v_sig alone over the full rows (wp_vsig.py:55-86), for a state changed
// after the support sweep (crk_state_refresh)
template <int D, int G>
__global__ void __launch_bounds__(kBlock) k_vsig(Rows rows, Geo geo, int32_t n, const float4* __restrict__ geo4,
                                                 const float4* __restrict__ vel4, float* __restrict__ vsig) {
  Lane L = lane_of<G>(n);
  const int32_t k = L.k;
  const float4 gi = geo4[k], wi = vel4[k];
  const float xi[3] = {gi.x, gi.y, gi.z};
  const float vi[3] = {wi.x, wi.y, wi.z};
  const float eps = 1.0e-14f * gi.w;
  const int32_t b = rows.off[k], e = L.valid ? rows.off[k + 1] : b;
  float vs = 0.f;
  for (int32_t q = b + L.g; q < e; q += G) {
    const int32_t j = rows.nbr[q];
    const float4 gj = geo4[j], wj = vel4[j];
    const float xj[3] = {gj.x, gj.y, gj.z};
    float dx[3];
    pair_dx<D>(geo, xi, xj, dx);
    const float r = fsqrt(dot<D>(dx, dx));
    const float vj[3] = {wj.x, wj.y, wj.z};
    float mu = 0.f;
#pragma unroll
    for (int a = 0; a < D; ++a) mu = mu + (vi[a] - vj[a]) * dx[a];
    mu = mu * frcp(r + eps);
    float vv = 0.5f * (wi.w + wj.w) - mu;
    if (mu > 0.f) vv = 0.f;
    vs = vv > vs ? vv : vs;
  }
  vs = group_max<G>(vs);
  if (L.valid && L.g == 0) vsig[k] = vs;
}

// ------------------------------------------------------------ Owen
__device__ inline double lut_interp(const double* __restrict__ xs, const double* __restrict__ ys, int n, double x) {
  int lo = 0, hi = n;  // upper_bound
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (xs[mid] <= x) lo = mid + 1;
    else hi = mid;
  }
  int il = lo - 1;
  il = il < 0 ? 0 : (il > n - 2 ? n - 2 : il);
  int ir = il + 1;
  double a = (x - xs[il]) / (xs[ir] - xs[il]);
  a = a < 0.0 ? 0.0 : (a > 1.0 ? 1.0 : a);
  return ys[il] * (1.0 - a) + ys[ir] * a;
}

// FINAL: the last Owen iteration of a stage also computes, with the new h_i,
// the per-row parts of the support sweep (the CRK volume V_i and, with the
// Cullen-Dehnen switch, the signal speed) in a second sweep over the row,
// whose neighbour data the first sweep has just brought into the caches.
struct OwenFinal {
  const float *v, *u, *m;
  float cfac;
  float *V, *vsig;
  float2* vm2;
};

template <int D, int G, int K, bool FINAL>
__global__ void __launch_bounds__(kBlock) k_owen(Rows rows, Geo geo, float C, int32_t n,
                                                 const float* __restrict__ x, float* __restrict__ h,
                                                 const int32_t* __restrict__ perm, float* __restrict__ supports,
                                                 const double* __restrict__ lut_psiH, const double* __restrict__ lut_nh,
                                                 int nlut, float nhT, const float* __restrict__ hminmax,
                                                 float* __restrict__ ratio, float4* __restrict__ geo4,
                                                 OwenFinal fin, const int32_t* __restrict__ run,
                                                 const int32_t* __restrict__ gate) {
  if (gate && *gate) return;  // (the list failed its deferred test)
  Lane L = lane_of<G>(n);
  const int32_t k = L.k;
  const float4 gk = geo4[k];  // the row's {x, h} (h before this update)
  const float xi[3] = {gk.x, gk.y, gk.z};
  const float hi = gk.w;
  const float ih = 1.f / hi;
  Comp acc;
  const int32_t b = rows.off[k], e = L.valid ? rows.off[k + 1] : b;
  {
    // software pipeline: the neighbour's position one iteration ahead, its id two ahead
    int32_t jn = b + L.g + G < e ? rows.nbr[b + L.g + G] : 0;
    float xn[3];
    {
      const int32_t j0 = b + L.g < e ? rows.nbr[b + L.g] : 0;
#pragma unroll
      for (int d = 0; d < D; ++d) xn[d] = x[(int64_t)j0 * D + d];
    }
    for (int32_t q = b + L.g; q < e; q += G) {
      float xj[3];
#pragma unroll
      for (int d = 0; d < D; ++d) xj[d] = xn[d];
      if (q + G < e) {
#pragma unroll
        for (int d = 0; d < D; ++d) xn[d] = x[(int64_t)jn * D + d];
        if (q + 2 * G < e) jn = rows.nbr[q + 2 * G];
      }
      float dx[3];
      pair_dx<D>(geo, xi, xj, dx);
      float qq = qmul(fsqrt(dot<D>(dx, dx)), ih);
      if (!(qq < 1.f)) continue;
      float kk, dk;
      kwd<K, D>(qq, kk, dk);
      acc.add(C * fabsf(dk));
    }
  }
  float S = group_total<G>(acc);
  float rmax = 0.f;
  // every lane of the group evaluates the same new support (FINAL needs it)
  float ps = powf(S, 1.f / (float)D);
  float nh = (float)lut_interp(lut_psiH, lut_nh, nlut, (double)ps);
  float sr = nhT / nh;
  float a = sr >= 1.f ? 0.4f * (1.f + 1.f / (sr * sr * sr)) : 0.4f * (1.f + sr * sr);
  float hn = (1.f - a + a * sr) * hi;
  const int32_t rr = run_index(run, k);  // the support range of the particle's run
  float lo = hminmax[2 * rr] * 0.25f, up = hminmax[2 * rr + 1] * 4.0f;
  hn = hn < lo ? lo : (hn > up ? up : hn);
  if (L.valid && L.g == 0) {
    h[k] = hn;
    geo4[k].w = hn;
    supports[perm[k]] = hn;
    rmax = fabsf(hn / hi - 1.f);
  }
  if (FINAL) {
    // crk_support's per-row work with h_i = hn (V: gather W(x_ij, h_i) over
    // r < h_i; v_sig over every entry, wp_vsig.py:55-86)
    float vi[3];
#pragma unroll
    for (int d = 0; d < D; ++d) vi[d] = fin.v[(int64_t)k * D + d];
    const float ihn = 1.f / hn;
    const float ci = sqrtf(fabsf(fin.u[k]) * fin.cfac);
    const float eps = 1.0e-14f * hn;
    float vs = 0.f;
    Sum vol;
    int32_t jn = b + L.g < e ? rows.nbr[b + L.g] : 0;  // neighbour ids loaded one iteration ahead
    for (int32_t q = b + L.g; q < e; q += G) {
      const int32_t j = jn;
      if (q + G < e) jn = rows.nbr[q + G];
      float dx[3];
      pair_dx<D>(geo, xi, &x[(int64_t)j * D], dx);
      float r = fsqrt(dot<D>(dx, dx));
      if (fin.vsig) {
        float mu = 0.f;
#pragma unroll
        for (int c = 0; c < D; ++c) mu = mu + (vi[c] - fin.v[(int64_t)j * D + c]) * dx[c];
        mu = mu * frcp(r + eps);
        float vv = 0.5f * (ci + fsqrt(fabsf(fin.u[j]) * fin.cfac)) - mu;
        if (mu > 0.f) vv = 0.f;
        vs = vv > vs ? vv : vs;
      }
      float qq = qmul(r, ihn);
      if (qq < 1.f) {
        float kk, dk;
        kwd<K, D>(qq, kk, dk);
        vol.fma(C * kk, ipow<D>(ihn));
      }
    }
    vs = group_max<G>(vs);
    float Sv = group_total<G>(vol);
    if (L.valid && L.g == 0) {
      const float Vk = 1.f / Sv;
      fin.V[k] = Vk;
      fin.vm2[k] = make_float2(Vk, fin.m[k]);
      if (fin.vsig) fin.vsig[k] = vs;
    }
  }
  if (ratio) {  // max |h_new / h - 1|, one atomic per block
    rmax = block_reduce(rmax, MaxOp());
    if (threadIdx.x == 0) atomic_max_pos(ratio, rmax);
  }
}

// Owen's update on an LDS tile (the non-final sweep of k_owen, row by row):
// the row's psi0 over its Verlet entries from the source's {x} records.
template <int D, int G, int K, class Src>
__device__ inline float owen_row(const Lane& L, const Src& src, const int32_t* __restrict__ off, Geo geo, float C,
                                 float* __restrict__ h, const int32_t* __restrict__ perm,
                                 float* __restrict__ supports, const double* __restrict__ lut_psiH,
                                 const double* __restrict__ lut_nh, int nlut, float nhT,
                                 const float* __restrict__ hminmax, float4* __restrict__ geo4,
                                 const int32_t* __restrict__ run) {
  const int32_t k = L.k;
  const float4 gk = geo4[k];  // the row's {x, h} (h before this update)
  const float xi[3] = {gk.x, gk.y, gk.z};
  const float hi = gk.w;
  const float ih = 1.f / hi;
  Comp acc;
  const int32_t b = off[k], e = L.valid ? off[k + 1] : b;
  // the neighbour's record one iteration ahead, its id two ahead
  int32_t jn = b + L.g + G < e ? src.id(b + L.g + G) : 0;
  float4 gn = src.geo(b + L.g < e ? src.id(b + L.g) : 0);
  for (int32_t q = b + L.g; q < e; q += G) {
    const float4 gj = gn;
    if (q + G < e) {
      gn = src.geo(jn);
      if (q + 2 * G < e) jn = src.id(q + 2 * G);
    }
    const float xj[3] = {gj.x, gj.y, gj.z};
    float dx[3];
    pair_dx<D>(geo, xi, xj, dx);
    float qq = qmul(fsqrt(dot<D>(dx, dx)), ih);
    if (!(qq < 1.f)) continue;
    float kk, dk;
    kwd<K, D>(qq, kk, dk);
    acc.add(C * fabsf(dk));
  }
  float S = group_total<G>(acc);
  float ps = powf(S, 1.f / (float)D);
  float nh = (float)lut_interp(lut_psiH, lut_nh, nlut, (double)ps);
  float sr = nhT / nh;
  float a = sr >= 1.f ? 0.4f * (1.f + 1.f / (sr * sr * sr)) : 0.4f * (1.f + sr * sr);
  float hn = (1.f - a + a * sr) * hi;
  const int32_t rr = run_index(run, k);
  float lo = hminmax[2 * rr] * 0.25f, up = hminmax[2 * rr + 1] * 4.0f;
  hn = hn < lo ? lo : (hn > up ? up : hn);
  if (!L.valid || L.g != 0) return 0.f;
  h[k] = hn;
  geo4[k].w = hn;
  supports[perm[k]] = hn;
  return fabsf(hn / hi - 1.f);
}

// One block per tile: the tile's {x, h} records staged in LDS (the Owen
// update only reads the neighbours' positions, so the rows' own h writes to
// geo4 do not race with what another tile stages).
template <int D, int G, int K>
__global__ void __launch_bounds__(kTileThreadsMax) k_owen_t(Rows rows, const uint16_t* __restrict__ nloc,
                                                   const int32_t* __restrict__ tiles, Geo geo, float C, int32_t n,
                                                   float* __restrict__ h, const int32_t* __restrict__ perm,
                                                   float* __restrict__ supports, const double* __restrict__ lut_psiH,
                                                   const double* __restrict__ lut_nh, int nlut, float nhT,
                                                   const float* __restrict__ hminmax, float* __restrict__ ratio,
                                                   float4* __restrict__ geo4, const int32_t* __restrict__ run,
                                                   const int32_t* __restrict__ gate) {
  __shared__ int32_t tab[kTileWords];
  __shared__ float4 sg[kTileCap];
  __shared__ __align__(16) uint16_t sid[kTileIds + 8];
  if (gate && *gate) return;  // (the list failed its deferred test)
  const int32_t t = xcd_block();
  const int32_t base = tile_stage_fn<2, 2, float4>(
      tiles, t, rows.off[t * kTileRows], rows.off[min(n, (t + 1) * kTileRows)], nloc, tab, sid,
      [&](int32_t g) { return geo4[g]; }, [&](int32_t u, const float4& v) { sg[u] = v; });
  const bool wide = tab[0] < 0;
  float rmax = 0.f;
  const int nr = tile_rounds<G>();
  for (int round = 0; round < nr; ++round) {
    const Lane L = tile_lane<G>(t, round, n);
    const float r = wide
        ? owen_row<D, G, K>(L, GlobalSrc<float4>{rows.nbr, geo4, geo4}, rows.off, geo, C, h, perm, supports,
                            lut_psiH, lut_nh, nlut, nhT, hminmax, geo4, run)
        : owen_row<D, G, K>(L, TileSrc<float4>{sid, base, rows.nbr, sg, sg}, rows.off, geo, C, h, perm, supports,
                            lut_psiH, lut_nh, nlut, nhT, hminmax, geo4, run);
    rmax = fmaxf(rmax, r);
  }
  if (ratio) {  // max |h_new / h - 1|, one atomic per block
    rmax = block_reduce(rmax, MaxOp());
    if (threadIdx.x == 0) atomic_max_pos(ratio, rmax);
  }
}

// ------------------------------------------------------------ Monaghan support
// volumeToSupportHelper (utils/support.py:18-44): the radius holding n_T
// neighbours of volume V, in the reference's operation order
template <int D>
__device__ inline float support_of_volume(float V, float nT) {
  if (D == 1) return nT * V / 2.f;
  if (D == 2) return sqrtf(nT * V / 3.14159265f);
  return powf(nT * V / 3.14159265f * 3.f / 4.f, 1.f / 3.f);
}

// evaluateOptimalSupportMonaghan (optimalSupportMonaghan.py:20-108), one
// Newton iteration on F(h) = h - H(m / rho) (computeH, :10-18) over the
// Verlet row: rho_i = sum_j m_j W(x_ij, h_i) (Density, Gather) and
// Omega_i = 1 - sum_j (-h_i / (rho_i d)) m_j dW/dh(x_ij, h_i)
// (computeOmega, wp_omega.py:52-73, 143) with
// dW/dh = -C / h^(d+1) (d k(q) + q k'(q));  h' = h - F / (Omega + 1e-6),
// non-finite Omega, step or h' dropped as the reference's nan_to_num
// (:66-71), clamped to [hMin/4, 4 hMax] of the particle's run (:73).
// ratio: max |h' / (h + 1e-6) - 1| (:77-90).
template <int D, int G, int K>
__global__ void __launch_bounds__(kBlock) k_monaghan(Rows rows, Geo geo, float C, int32_t n,
                                                     const float* __restrict__ x, float* __restrict__ h,
                                                     const float* __restrict__ m, const int32_t* __restrict__ perm,
                                                     float* __restrict__ supports, float nT,
                                                     const float* __restrict__ hminmax, float* __restrict__ ratio,
                                                     float4* __restrict__ geo4, const int32_t* __restrict__ run,
                                                     const int32_t* __restrict__ gate) {
  if (gate && *gate) return;  // (the list failed its deferred test)
  Lane L = lane_of<G>(n);
  const int32_t k = L.k;
  float xi[3];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = x[(int64_t)k * D + d];
  const float hi = h[k];
  const float ih = 1.f / hi;
  Comp racc, wacc;
  const int32_t b = rows.off[k], e = L.valid ? rows.off[k + 1] : b;
  int32_t jn = b + L.g < e ? rows.nbr[b + L.g] : 0;  // neighbour ids loaded one iteration ahead
  for (int32_t q = b + L.g; q < e; q += G) {
    const int32_t j = jn;
    if (q + G < e) jn = rows.nbr[q + G];
    float dx[3];
    pair_dx<D>(geo, xi, &x[(int64_t)j * D], dx);
    const float qq = qmul(fsqrt(dot<D>(dx, dx)), ih);
    if (!(qq < 1.f)) continue;
    float kk, dk;
    kwd<K, D>(qq, kk, dk);
    const float cw = C * ipow<D>(ih);
    const float mj = m[j];
    racc.add(mj * (cw * kk));
    wacc.add(mj * (-cw * ih * ((float)D * kk + qq * dk)));
  }
  const float rho = group_total<G>(racc);
  const float S = group_total<G>(wacc);
  float rs = isfinite(rho) ? rho : 1.f;
  rs = rs < 1e-12f ? 1e-12f : rs;
  const float F = hi - support_of_volume<D>(m[k] / rs, nT);
  const float dh_drho = -hi / (rho * (float)D);
  float om = 1.f - dh_drho * S;
  om = isfinite(om) ? om : 0.f;
  float step = F / (om + 1e-6f);
  step = isfinite(step) ? step : 0.f;
  float hn = hi - step;
  hn = isfinite(hn) ? hn : hi;
  const int32_t rr = run_index(run, k);  // the support range of the particle's run
  const float lo = hminmax[2 * rr] * 0.25f, up = hminmax[2 * rr + 1] * 4.0f;
  hn = hn < lo ? lo : (hn > up ? up : hn);
  float rmax = 0.f;
  if (L.valid && L.g == 0) {
    h[k] = hn;
    geo4[k].w = hn;
    supports[perm[k]] = hn;
    rmax = fabsf(hn / (hi + 1e-6f) - 1.f);
  }
  if (ratio) {
    rmax = block_reduce(rmax, MaxOp());
    if (threadIdx.x == 0) atomic_max_pos(ratio, rmax);
  }
}

// Density by gather over the Verlet row, rho_i = sum_j m_j W(x_ij, h_i)
// (warpOperation Density with SupportScheme.Gather: the final pass of
// optimalSupportOwen.py:116-125 and the initial density of
// sedov/initial.py:83-93), written in user order.
template <int D, int G, int K>
__global__ void __launch_bounds__(kBlock) k_density(Rows rows, Geo geo, float C, int32_t n,
                                                    const float* __restrict__ x, const float* __restrict__ h,
                                                    const float* __restrict__ m, const int32_t* __restrict__ perm,
                                                    float* __restrict__ rho_u) {
  Lane L = lane_of<G>(n);
  const int32_t k = L.k;
  float xi[3];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = x[(int64_t)k * D + d];
  const float ih = 1.f / h[k];
  Comp acc;
  const int32_t b = rows.off[k], e = L.valid ? rows.off[k + 1] : b;
  for (int32_t q = b + L.g; q < e; q += G) {
    const int32_t j = rows.nbr[q];
    float dx[3];
    pair_dx<D>(geo, xi, &x[(int64_t)j * D], dx);
    const float qq = qmul(fsqrt(dot<D>(dx, dx)), ih);
    if (!(qq < 1.f)) continue;
    float kk, dk;
    kwd<K, D>(qq, kk, dk);
    acc.add(m[j] * (C * kk * ipow<D>(ih)));
  }
  const float rho = group_total<G>(acc);
  if (L.valid && L.g == 0) rho_u[perm[k]] = rho;
}

// ------------------------------------------------------------ CRK moments
// Linear CRK (Frontiere, Raskin & Owen 2017) with W_ij = W(x_ij, h_j):
//   m0 = sum V_j W, m1 = sum V_j x W, m2 = sum V_j x x W and x_i-gradients,
//   B = -m2^-1 m1, A = 1/(m0 + B.m1), grad A = -A^2 (dm0 + 2 B.dm1 + B.dm2.B),
//   grad B = -m2^-1 (dm1 + dm2 B); singular m2 -> A = 1/m0, B = 0.
// CRK density rho = A (S0 + B.S1), S0 = sum m_j W, S1 = sum m_j x W; then the
// ideal-gas EOS with u given (eos/idealGas.py:6-38) and s = P / rho^gamma
// (the entropy function the energy partition compares, compSPH/balance.py:136-137).
struct MomOut {
  float *A, *B, *gA, *gB, *rho, *P, *c, *sP;
  float *dens_u, *ent_u, *pres_u, *cs_u;
  int32_t* nsing;
  // neighbour records: {V, m} read; the velocity-gradient record written
  const float2* vm2;
  const float* v;
  float4* vg4;
  float* mr;
};

template <int D>
struct MomAcc {
  static constexpr int NM2 = D * (D + 1) / 2;
  Sum m0, m1[D], m2[NM2], dm0[D], dm1[D * D], dm2[NM2 * D], S0, S1[D];
  // one pair with r < h_j: W = W(x_ij, h_j) and its x_i-gradient
  template <int K>
  __device__ inline void add(float C, const float* dx, float r, float hj, float Vj, float mj) {
    WdW w = wdw<K, D>(C, r, frcp(hj));
    float g[3];
#pragma unroll
    for (int d = 0; d < D; ++d) g[d] = w.fac * dx[d];
    float VW = Vj * w.W;
    m0.add(VW);
#pragma unroll
    for (int a = 0; a < D; ++a) m1[a].fma(VW, dx[a]);
    {
      int p = 0;
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int bb = a; bb < D; ++bb, ++p) m2[p].fma(VW * dx[a], dx[bb]);
    }
#pragma unroll
    for (int gm = 0; gm < D; ++gm) dm0[gm].fma(Vj, g[gm]);
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int gm = 0; gm < D; ++gm) dm1[a * D + gm].fma(Vj, dx[a] * g[gm] + (a == gm ? w.W : 0.f));
    {
      int p = 0;
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int bb = a; bb < D; ++bb, ++p)
#pragma unroll
          for (int gm = 0; gm < D; ++gm) {
            float lin = (a == gm ? dx[bb] : 0.f) + (bb == gm ? dx[a] : 0.f);
            dm2[p * D + gm].fma(Vj, dx[a] * dx[bb] * g[gm] + lin * w.W);
          }
    }
    float mW = mj * w.W;
    S0.add(mW);
#pragma unroll
    for (int a = 0; a < D; ++a) S1[a].fma(mW, dx[a]);
  }
};

// COMPACT = false: iterate the pairs compacted by the support sweep.
// COMPACT = true: iterate the full Verlet rows and compact the pairs with
// |x_ij| < max(h_i, h_j) into anbr / acnt on the way (the support sweep's
// compaction, for when V and v_sig were computed by the final Owen sweep).
// One row of the moments pass from a neighbour source (geo = {x, h}, aux =
// {V, m}); COMPACT rows scan the full Verlet row through the global ids.
template <int D, int G, int K, bool COMPACT, class Src>
__device__ inline void moments_row(const Lane& L, const Src& src, ARows rows, Rows full, Geo geo, float C, int32_t n,
                                   const float* __restrict__ x, const float* __restrict__ h,
                                   const float* __restrict__ V, const float* __restrict__ m,
                                   const float* __restrict__ u, float gamma, const int32_t* __restrict__ perm,
                                   const MomOut& o, int32_t* __restrict__ anbr, int32_t* __restrict__ acnt,
                                   const float4* __restrict__ geo4) {
  const int32_t k = L.k;
  const float4 gk = geo4[k];
  const float xi[3] = {gk.x, gk.y, gk.z};
  MomAcc<D> acc;
  if (!COMPACT) {
    const int32_t b = rows.off[k], e = L.valid ? b + rows.cnt[k] : b;
    // software pipeline: the neighbour's records one iteration ahead, its id two ahead
    int32_t jc = b + L.g < e ? src.id(b + L.g) : 0;
    int32_t jn = b + L.g + G < e ? src.id(b + L.g + G) : 0;
    float4 gn = src.geo(jc);
    float2 vn = src.aux(jc);
    for (int32_t q = b + L.g; q < e; q += G) {
      const float4 gj = gn;
      const float2 vmj = vn;
      if (q + G < e) {
        jc = jn;
        gn = src.geo(jc);
        vn = src.aux(jc);
        if (q + 2 * G < e) jn = src.id(q + 2 * G);
      }
      float dx[3];
      const float xj[3] = {gj.x, gj.y, gj.z};
      pair_dx<D>(geo, xi, xj, dx);
      float r = fsqrt(dot<D>(dx, dx));
      float hj = gj.w;
      if (!(r < hj)) continue;
      acc.template add<K>(C, dx, r, hj, vmj.x, vmj.y);
    }
  } else {
    const float hi = h[k];
    const int lane = threadIdx.x & 63;
    const int gbase = lane & ~(G - 1);
    const unsigned long long gmask = G == 64 ? ~0ull : (((1ull << G) - 1ull) << gbase);
    const unsigned long long lt = (1ull << lane) - 1ull;
    const int32_t b = full.off[k], e = L.valid ? full.off[k + 1] : b;
    int32_t run = 0;
    int32_t jn = b + L.g < e ? full.nbr[b + L.g] : 0;  // neighbour ids loaded one iteration ahead
    for (int32_t base = b; base < e; base += G) {
      const int32_t q = base + L.g;
      bool act = false;
      int32_t j = 0;
      if (q < e) {
        j = jn;
        if (q + G < e) jn = full.nbr[q + G];
        float dx[3];
        const float4 gj = geo4[j];
        const float xj[3] = {gj.x, gj.y, gj.z};
        pair_dx<D>(geo, xi, xj, dx);
        float r = fsqrt(dot<D>(dx, dx));
        float hj = gj.w;
        act = r < fmaxf(hi, hj);
        if (r < hj) {
          const float2 vmj = o.vm2[j];
          acc.template add<K>(C, dx, r, hj, vmj.x, vmj.y);
        }
      }
      const unsigned long long bm = __ballot(act ? 1 : 0) & gmask;
      if (act) anbr[b + run + __popcll(bm & lt)] = j;
      run += __popcll(bm);
    }
    if (L.valid && L.g == 0) acnt[k] = run;
  }
  Sum &m0 = acc.m0, *m1 = acc.m1, *m2 = acc.m2, *dm0 = acc.dm0, *dm1 = acc.dm1, *dm2 = acc.dm2, &S0 = acc.S0,
       *S1 = acc.S1;
  // combine partial sums across the group
  double M0 = group_total_d<G>(m0), M1[3], M2[3][3], DM0[3], DM1[3][3], DM2[3][3][3];
#pragma unroll
  for (int a = 0; a < D; ++a) M1[a] = group_total_d<G>(m1[a]);
  {
    int p = 0;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int bb = a; bb < D; ++bb, ++p) { double t = group_total_d<G>(m2[p]); M2[a][bb] = t; M2[bb][a] = t; }
  }
#pragma unroll
  for (int gm = 0; gm < D; ++gm) DM0[gm] = group_total_d<G>(dm0[gm]);
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int gm = 0; gm < D; ++gm) DM1[a][gm] = group_total_d<G>(dm1[a * D + gm]);
  {
    int p = 0;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int bb = a; bb < D; ++bb, ++p)
#pragma unroll
        for (int gm = 0; gm < D; ++gm) { double t = group_total_d<G>(dm2[p * D + gm]); DM2[a][bb][gm] = t; DM2[bb][a][gm] = t; }
  }
  float s0 = group_total<G>(S0), s1[3];
#pragma unroll
  for (int a = 0; a < D; ++a) s1[a] = group_total<G>(S1[a]);
  if (!L.valid || L.g != 0) return;
  // FP64 solve
  double inv[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int bb = 0; bb < 3; ++bb) inv[a][bb] = 0.0;
  bool sing;
  if (D == 1) {
    sing = !(M2[0][0] > 0.0);
    if (!sing) inv[0][0] = 1.0 / M2[0][0];
  } else if (D == 2) {
    double det = M2[0][0] * M2[1][1] - M2[0][1] * M2[1][0];
    double tr = 0.5 * (M2[0][0] + M2[1][1]);
    sing = !(det > 1e-10 * tr * tr);
    if (!sing) {
      double id = 1.0 / det;
      inv[0][0] = M2[1][1] * id; inv[1][1] = M2[0][0] * id;
      inv[0][1] = -M2[0][1] * id; inv[1][0] = -M2[1][0] * id;
    }
  } else {
    double c00 = M2[1][1] * M2[2][2] - M2[1][2] * M2[2][1];
    double c01 = M2[1][2] * M2[2][0] - M2[1][0] * M2[2][2];
    double c02 = M2[1][0] * M2[2][1] - M2[1][1] * M2[2][0];
    double det = M2[0][0] * c00 + M2[0][1] * c01 + M2[0][2] * c02;
    double tr = (M2[0][0] + M2[1][1] + M2[2][2]) / 3.0;
    sing = !(det > 1e-10 * tr * tr * tr);
    if (!sing) {
      double id = 1.0 / det;
      inv[0][0] = c00 * id; inv[1][0] = c01 * id; inv[2][0] = c02 * id;
      inv[0][1] = (M2[0][2] * M2[2][1] - M2[0][1] * M2[2][2]) * id;
      inv[1][1] = (M2[0][0] * M2[2][2] - M2[0][2] * M2[2][0]) * id;
      inv[2][1] = (M2[0][1] * M2[2][0] - M2[0][0] * M2[2][1]) * id;
      inv[0][2] = (M2[0][1] * M2[1][2] - M2[0][2] * M2[1][1]) * id;
      inv[1][2] = (M2[0][2] * M2[1][0] - M2[0][0] * M2[1][2]) * id;
      inv[2][2] = (M2[0][0] * M2[1][1] - M2[0][1] * M2[1][0]) * id;
    }
  }
  double Ad, Bd[3] = {0, 0, 0}, gAd[3] = {0, 0, 0}, gBd[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int bb = 0; bb < 3; ++bb) gBd[a][bb] = 0.0;
  if (sing) {
    if (o.nsing) atomicAdd(o.nsing, 1);
    Ad = 1.0 / M0;
#pragma unroll
    for (int gm = 0; gm < D; ++gm) gAd[gm] = -DM0[gm] / (M0 * M0);
  } else {
#pragma unroll
    for (int a = 0; a < D; ++a) {
      double s = 0.0;
#pragma unroll
      for (int bb = 0; bb < D; ++bb) s += inv[a][bb] * M1[bb];
      Bd[a] = -s;
    }
    double den = M0;
#pragma unroll
    for (int a = 0; a < D; ++a) den += Bd[a] * M1[a];
    Ad = 1.0 / den;
#pragma unroll
    for (int gm = 0; gm < D; ++gm) {
      double dinvA = DM0[gm];
#pragma unroll
      for (int a = 0; a < D; ++a) dinvA += 2.0 * Bd[a] * DM1[a][gm];
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int bb = 0; bb < D; ++bb) dinvA += Bd[a] * DM2[a][bb][gm] * Bd[bb];
      gAd[gm] = -Ad * Ad * dinvA;
      double tmp[3];
#pragma unroll
      for (int bb = 0; bb < D; ++bb) {
        double t2 = DM1[bb][gm];
#pragma unroll
        for (int dd = 0; dd < D; ++dd) t2 += DM2[bb][dd][gm] * Bd[dd];
        tmp[bb] = t2;
      }
#pragma unroll
      for (int a = 0; a < D; ++a) {
        double s = 0.0;
#pragma unroll
        for (int bb = 0; bb < D; ++bb) s += inv[a][bb] * tmp[bb];
        gBd[a][gm] = -s;
      }
    }
  }
  float Af = (float)Ad;
  float Bf[3];
  o.A[k] = Af;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    Bf[a] = (float)Bd[a];
    o.B[(int64_t)k * D + a] = Bf[a];
    o.gA[(int64_t)k * D + a] = (float)gAd[a];
  }
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int gm = 0; gm < D; ++gm) o.gB[(int64_t)k * D * D + a * D + gm] = (float)gBd[a][gm];
  float bs = s0;
#pragma unroll
  for (int a = 0; a < D; ++a) bs = bs + Bf[a] * s1[a];
  float rh = Af * bs;
  o.rho[k] = rh;
  {
    // the velocity-gradient pass's neighbour record (m / rho as that pass
    // evaluates it)
    const float2 vmk = o.vm2[k];
    const float mrk = vmk.y * frcp(rh);
    float vk[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < D; ++d) vk[d] = o.v[(int64_t)k * D + d];
    if (D < 3) {
      o.vg4[k] = make_float4(vk[0], vk[1], vmk.x, mrk);
    } else {
      o.vg4[k] = make_float4(vk[0], vk[1], vk[2], vmk.x);
      o.mr[k] = mrk;
    }
  }
  float ui = u[k];
  float cc = sqrtf(fabsf(ui) * gamma * (gamma - 1.f));
  float P = (gamma - 1.f) * rh * ui;
  float s = ui * (gamma - 1.f) * powf(rh, 1.f - gamma);
  o.P[k] = P;
  o.c[k] = cc;
  o.sP[k] = P / powf(rh, gamma);
  int32_t i = perm[k];
  o.dens_u[i] = rh;
  o.pres_u[i] = P;
  o.cs_u[i] = cc;
  o.ent_u[i] = s;
}

template <int D, int G, int K, bool COMPACT>
__global__ void __launch_bounds__(kBlock) k_moments(ARows rows, Rows full, Geo geo, float C, int32_t n,
                                                    const float* __restrict__ x, const float* __restrict__ h,
                                                    const float* __restrict__ V, const float* __restrict__ m,
                                                    const float* __restrict__ u, float gamma,
                                                    const int32_t* __restrict__ perm, MomOut o,
                                                    int32_t* __restrict__ anbr, int32_t* __restrict__ acnt,
                                                    const float4* __restrict__ geo4) {
  const Lane L = lane_of<G>(n);
  moments_row<D, G, K, COMPACT>(L, GlobalSrc<float2>{rows.nbr, geo4, o.vm2}, rows, full, geo, C, n, x, h, V, m, u,
                                gamma, perm, o, anbr, acnt, geo4);
}

// One block per tile, the pairs' {x, h} and {V, m} records staged in LDS.
template <int D, int G, int K>
__global__ void __launch_bounds__(kTileThreadsMax) k_moments_t(ARows rows, const uint16_t* __restrict__ ploc,
                                                      const int32_t* __restrict__ tiles, Geo geo, float C, int32_t n,
                                                      const float* __restrict__ x, const float* __restrict__ h,
                                                      const float* __restrict__ V, const float* __restrict__ m,
                                                      const float* __restrict__ u, float gamma,
                                                      const int32_t* __restrict__ perm, MomOut o,
                                                      const float4* __restrict__ geo4) {
  __shared__ int32_t tab[kTileWords];
  __shared__ float4 sg[kTileCap];
  __shared__ float2 sv[kTileCap];
  // (the pair indices stay in global memory here: staging them in LDS was
  // measured slower for this pass, whose occupancy its registers limit)
  const int32_t t = xcd_block();
  struct PV {
    float4 a;
    float2 b;
  };
  const float2* vm2 = (const float2*)o.vm2;
  tile_stage_fn<2, 0, PV>(
      tiles, t, 0, 0, nullptr, tab, nullptr, [&](int32_t g) { return PV{geo4[g], vm2[g]}; },
      [&](int32_t u, const PV& v) {
        sg[u] = v.a;
        sv[u] = v.b;
      });
  const bool wide = tab[0] < 0;
  const int nr = tile_rounds<G>();
  for (int round = 0; round < nr; ++round) {
    const Lane L = tile_lane<G>(t, round, n);
    if (wide)
      moments_row<D, G, K, false>(L, GlobalSrc<float2>{rows.nbr, geo4, o.vm2}, rows, Rows{}, geo, C, n, x, h, V, m,
                                  u, gamma, perm, o, nullptr, nullptr, geo4);
    else
      moments_row<D, G, K, false>(L, TileSrc<float2>{ploc, 0, rows.nbr, sg, sv}, rows, Rows{}, geo, C, n, x, h, V, m,
                                  u, gamma, perm, o, nullptr, nullptr, geo4);
  }
}

// ------------------------------------------------------------ state refresh
// After the caller changed the state between the CRK factors and the rest of
// the stage (boundary conditions, crkSPH.py:94): the equation of state from
// the (possibly changed) densities and energies, as at the end of the moments
// pass, and the velocity-gradient record {v, V, m/rho}.
template <int D>
__global__ void k_refresh_eos(const int32_t* __restrict__ perm, int32_t n, const float* __restrict__ rho_user,
                              const float* __restrict__ m, const float* __restrict__ V, const float* __restrict__ u,
                              float gamma, MomOut o) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int32_t i = perm[k];
  const float rh = rho_user[i];
  o.rho[k] = rh;
  const float mrk = m[k] * frcp(rh);
  float vk[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int d = 0; d < D; ++d) vk[d] = o.v[(int64_t)k * D + d];
  if (D < 3) {
    o.vg4[k] = make_float4(vk[0], vk[1], V[k], mrk);
  } else {
    o.vg4[k] = make_float4(vk[0], vk[1], vk[2], V[k]);
    o.mr[k] = mrk;
  }
  const float ui = u[k];
  const float cc = sqrtf(fabsf(ui) * gamma * (gamma - 1.f));
  const float P = (gamma - 1.f) * rh * ui;
  const float s = ui * (gamma - 1.f) * powf(rh, 1.f - gamma);
  o.P[k] = P;
  o.c[k] = cc;
  o.sP[k] = P / powf(rh, gamma);
  o.pres_u[i] = P;
  o.cs_u[i] = cc;
  o.ent_u[i] = s;
}

// one user-order field of C components into sorted order
template <int C>
__global__ void k_gather_field(const float* __restrict__ src, const int32_t* __restrict__ perm, int32_t n,
                               float* __restrict__ dst) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int32_t i = perm[k];
#pragma unroll
  for (int c = 0; c < C; ++c) dst[(int64_t)k * C + c] = src[(int64_t)i * C + c];
}

// ------------------------------------------------------------ export
// Slot data of a row to adjacency order: the pairs are the row's edges with
// r < max(h_i, h_j) in row order, so one merge walk over the row matches
// them (neighbours are distinct within a row).  f_ij is evaluated from the
// stage's energy-partition records (or read from the slots when stored).
template <int D>
__global__ void k_edges_export(const int32_t* __restrict__ off, const int32_t* __restrict__ nbr,
                               const int32_t* __restrict__ cnt, const int32_t* __restrict__ pnbr, int32_t n,
                               const int32_t* __restrict__ perm, const float* __restrict__ m_u, DtArg dt,
                               const float* __restrict__ pf, const float* __restrict__ th,
                               const float4* __restrict__ rec, const float* __restrict__ f,
                               float* __restrict__ ap_o, float* __restrict__ av_o, float* __restrict__ f_o) {
  int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int32_t b = off[k], e = off[k + 1], c = cnt[k];
  const float mdt = (f_o && !f) ? m_u[perm[k]] * dt.at(k) : 0.f;
  const float4 ri = (f_o && !f) ? rec[k] : make_float4(0.f, 0.f, 0.f, 0.f);
  int32_t p = 0;
  for (int32_t q = b; q < e; ++q) {
    const bool hit = p < c && pnbr[b + p] == nbr[q];
    const int64_t sl = b + p;
    const float tv = hit ? th[sl] : 0.f;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      const float fa = hit ? pf[sl * D + a] : 0.f;
      if (ap_o) ap_o[(int64_t)q * D + a] = (1.f - tv) * fa;
      if (av_o) av_o[(int64_t)q * D + a] = tv * fa;
    }
    if (f_o) {
      float fv = 0.f;
      if (hit) {
        if (f) {
          fv = f[sl];
        } else {
          float a[3];
#pragma unroll
          for (int d = 0; d < D; ++d) a[d] = pf[sl * D + d];
          fv = pair_fraction<D>(ri, rec[nbr[q]], mdt, a);
        }
      }
      f_o[q] = fv;
    }
    p += hit ? 1 : 0;
  }
}

template <int C>
__global__ void k_scatter_field(const float* __restrict__ src, const int32_t* __restrict__ perm, int32_t n,
                                float* __restrict__ dst) {
  int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int32_t i = perm[k];
#pragma unroll
  for (int c = 0; c < C; ++c) dst[(int64_t)i * C + c] = src[(int64_t)k * C + c];
}


int require_bound(CrkContext* ctx, const CrkConfig* cfg) {
  if (!ctx || !cfg) { set_error("null context/config"); return CRK_EINVAL; }
  if (!ctx->bound) { set_error("no stage bound: call crk_stage_begin first"); return CRK_ESTATE; }
  if (cfg->domain.dim != ctx->dim) { set_error("config dimension %d differs from the bound stage (%d)", cfg->domain.dim, ctx->dim); return CRK_EINVAL; }
  return CRK_OK;
}


static int alloc_stage(CrkContext* ctx, int32_t n, int D) {
  size_t f = sizeof(float) * (size_t)n;
  Buf* s1[] = {&ctx->h, &ctx->m, &ctx->u, &ctx->a0, &ctx->divp, &ctx->V, &ctx->A, &ctx->rho, &ctx->P, &ctx->c,
               &ctx->divcd, &ctx->vsig, &ctx->Rcd, &ctx->du0, &ctx->alpha, &ctx->sP};
  for (Buf* b : s1) { int st = ensure(*b, f); if (st) return st; }
  Buf* sd[] = {&ctx->x, &ctx->v, &ctx->B, &ctx->gA, &ctx->dvdt, &ctx->vh};
  for (Buf* b : sd) { int st = ensure(*b, f * D); if (st) return st; }
  Buf* sdd[] = {&ctx->gB, &ctx->J};
  for (Buf* b : sdd) { int st = ensure(*b, f * D * D); if (st) return st; }
  { int st = ensure(ctx->rec, sizeof(float) * (size_t)n * (D == 1 ? Rec<1>::NF : D == 2 ? Rec<2>::NF : Rec<3>::NF));
    if (st) return st; }
  Buf* s4[] = {&ctx->geo4, &ctx->vel4, &ctx->vg4};
  for (Buf* b : s4) { int st = ensure(*b, f * 4); if (st) return st; }
  { int st = ensure(ctx->vm2, f * 2); if (st) return st; }
  { int st = ensure(ctx->mr, D == 3 ? f : 4); if (st) return st; }
  return CRK_OK;
}

}  // namespace crk

using namespace crk;

extern "C" {

static int validate_adj(const CrkAdjacency* adj, const CrkState* st) {
  CRK_REQUIRE(adj && st, "null adjacency/state");
  CRK_REQUIRE(adj->n > 0 && adj->n == st->n, "adjacency has %d particles, state %d", adj->n, st->n);
  CRK_REQUIRE(adj->perm && adj->row_offsets && (adj->n_edges == 0 || adj->nbr), "adjacency buffers missing");
  CRK_REQUIRE(st->positions && st->velocities && st->supports && st->masses && st->internalEnergies,
              "state buffers missing");
  return CRK_OK;
}

int crk_set_group_size(CrkContext* ctx, int32_t pass, int32_t group) {
  CRK_REQUIRE(ctx, "crk_set_group_size: null context");
  CRK_REQUIRE(pass >= 0 && pass < CRK_NPASS, "unknown pass %d", pass);
  CRK_REQUIRE(group == 1 || group == 2 || group == 4 || group == 8 || group == 16 || group == 32 || group == 64,
              "group size must be a power of two in [1, 64] (got %d)", group);
  ctx->group[pass] = group;
  return CRK_OK;
}

int crk_set_block_size(CrkContext* ctx, int32_t pass, int32_t threads) {
  CRK_REQUIRE(ctx, "crk_set_block_size: null context");
  CRK_REQUIRE(pass >= 0 && pass < CRK_NPASS, "unknown pass %d", pass);
  CRK_REQUIRE(valid_tile_threads(threads), "tile block size must be 256, 512 or 1024 (got %d)", threads);
  ctx->threads[pass] = threads;
  return CRK_OK;
}

int crk_stage_begin(CrkContext* ctx, const CrkConfig* cfg, const CrkAdjacency* adj, const CrkState* st, void* stream) {
  CRK_REQUIRE(ctx && cfg, "crk_stage_begin: null context/config");
  int r = validate_adj(adj, st);
  if (r) return r;
  int D = cfg->domain.dim;
  CRK_REQUIRE(D >= 1 && D <= 3, "dim must be 1..3");
  CRK_REQUIRE(cfg->kernel >= 0 && cfg->kernel <= 3, "unknown kernel %d", cfg->kernel);
  if ((r = require_runs_n(ctx, adj->n, "crk_stage_begin"))) return r;
  if ((r = alloc_stage(ctx, adj->n, D))) return r;
  ctx->n = adj->n;
  ctx->dim = D;
  ctx->adj = *adj;
  ctx->st = *st;
  ctx->bound = true;
  ctx->have_factors = ctx->have_velgrad = ctx->have_accel = ctx->have_support = ctx->have_volume = false;
  ctx->have_vsig = ctx->have_divcd = ctx->have_rcd = false;
  ctx->check_hb = nullptr;
  ctx->gate = nullptr;
  ctx->owen_checked = false;
  hipStream_t s = (hipStream_t)stream;
  int32_t n = adj->n;
  CRK_DIM_DISPATCH(D, hipLaunchKernelGGL(k_gather<DD>, dim3(nblocks(n)), dim3(kBlock), 0, s, adj->perm, n,
                                         (const float*)st->positions, (const float*)st->velocities,
                                         (const float*)st->supports, (const float*)st->masses,
                                         (const float*)st->internalEnergies, (const float*)st->alpha0s,
                                         (const float*)st->divergence, (const float*)st->alphas, F<float>(ctx->x),
                                         F<float>(ctx->v), F<float>(ctx->h), F<float>(ctx->m), F<float>(ctx->u),
                                         F<float>(ctx->a0), F<float>(ctx->divp), F<float>(ctx->alpha),
                                         cfg->gamma * (cfg->gamma - 1.f), F<float4>(ctx->geo4),
                                         F<float4>(ctx->vel4)));
  CRK_LAUNCH_CHECK("k_gather");
  return CRK_OK;
}

int crk_state_refresh(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st, void* stream) {
  int r = require_bound(ctx, cfg);
  if (r) return r;
  CRK_REQUIRE(st && st->n == ctx->n && st->positions && st->velocities && st->supports && st->masses &&
                  st->internalEnergies && st->densities && st->entropies && st->pressures && st->soundspeeds,
              "crk_state_refresh: state buffers missing or of another size");
  if (!ctx->have_factors) { set_error("crk_state_refresh before crk_factors_eos"); return CRK_ESTATE; }
  hipStream_t s = (hipStream_t)stream;
  const int32_t n = ctx->n;
  const int D = ctx->dim;
  ctx->st = *st;
  CRK_DIM_DISPATCH(D, hipLaunchKernelGGL(k_gather<DD>, dim3(nblocks(n)), dim3(kBlock), 0, s, ctx->adj.perm, n,
                                         (const float*)st->positions, (const float*)st->velocities,
                                         (const float*)st->supports, (const float*)st->masses,
                                         (const float*)st->internalEnergies, (const float*)st->alpha0s,
                                         (const float*)st->divergence, (const float*)st->alphas, F<float>(ctx->x),
                                         F<float>(ctx->v), F<float>(ctx->h), F<float>(ctx->m), F<float>(ctx->u),
                                         F<float>(ctx->a0), F<float>(ctx->divp), F<float>(ctx->alpha),
                                         cfg->gamma * (cfg->gamma - 1.f), F<float4>(ctx->geo4),
                                         F<float4>(ctx->vel4)));
  CRK_LAUNCH_CHECK("k_gather");
  MomOut mo{F<float>(ctx->A), F<float>(ctx->B), F<float>(ctx->gA), F<float>(ctx->gB), F<float>(ctx->rho),
            F<float>(ctx->P), F<float>(ctx->c), F<float>(ctx->sP), st->densities, st->entropies, st->pressures,
            st->soundspeeds, nullptr, F<float2>(ctx->vm2), F<float>(ctx->v), F<float4>(ctx->vg4),
            F<float>(ctx->mr)};
  CRK_DIM_DISPATCH(D, hipLaunchKernelGGL(k_refresh_eos<DD>, dim3(nblocks(n)), dim3(kBlock), 0, s, ctx->adj.perm, n,
                                         (const float*)st->densities, F<float>(ctx->m), F<float>(ctx->V),
                                         F<float>(ctx->u), cfg->gamma, mo));
  CRK_LAUNCH_CHECK("k_refresh_eos");
  if (ctx->have_vsig) {
    // the switch's signal speed, from the changed velocities and sound speeds
    // (the reference evaluates it inside the switch, CullenDehnen2010.py:291)
    Geo geo = make_geo(*cfg);
    const int GS = ctx->group[CRK_PASS_SUPPORT];
    CRK_DIM_G_DISPATCH(D, GS, hipLaunchKernelGGL((k_vsig<DD, G>), dim3(group_blocks(n, G)), dim3(kBlock), 0, s,
                                                 rows_of(ctx), geo, n, F<float4>(ctx->geo4), F<float4>(ctx->vel4),
                                                 F<float>(ctx->vsig)));
    CRK_LAUNCH_CHECK("k_vsig");
  }
  return CRK_OK;
}

int crk_update_refresh(CrkContext* ctx, const CrkUpdate* up, void* stream) {
  CRK_REQUIRE(ctx && up && up->dvdt, "crk_update_refresh: null argument");
  if (!ctx->bound || !ctx->have_accel) { set_error("crk_update_refresh before crk_accel_dudt"); return CRK_ESTATE; }
  hipStream_t s = (hipStream_t)stream;
  const int32_t n = ctx->n;
  switch (ctx->dim) {
    case 1: hipLaunchKernelGGL(k_gather_field<1>, dim3(nblocks(n)), dim3(kBlock), 0, s, up->dvdt, ctx->adj.perm, n, F<float>(ctx->dvdt)); break;
    case 2: hipLaunchKernelGGL(k_gather_field<2>, dim3(nblocks(n)), dim3(kBlock), 0, s, up->dvdt, ctx->adj.perm, n, F<float>(ctx->dvdt)); break;
    default: hipLaunchKernelGGL(k_gather_field<3>, dim3(nblocks(n)), dim3(kBlock), 0, s, up->dvdt, ctx->adj.perm, n, F<float>(ctx->dvdt)); break;
  }
  CRK_LAUNCH_CHECK("k_gather_field");
  return CRK_OK;
}

int crk_stage_rebind(CrkContext* ctx, const CrkConfig* cfg, const CrkAdjacency* adj, const CrkState* st, void* stream) {
  return crk_stage_begin(ctx, cfg, adj, st, stream);
}

int crk_stage_begin_checked(CrkContext* ctx, const CrkConfig* cfg, const CrkAdjacency* adj, const CrkState* st,
                            const float* xb, const float* hb, float* hminmax_dev, int32_t* valid, void* stream) {
  CRK_REQUIRE(ctx && cfg && xb && hb && hminmax_dev, "crk_stage_begin_checked: null argument");
  int r = validate_adj(adj, st);
  if (r) return r;
  int D = cfg->domain.dim;
  CRK_REQUIRE(D >= 1 && D <= 3, "dim must be 1..3");
  CRK_REQUIRE(cfg->kernel >= 0 && cfg->kernel <= 3, "unknown kernel %d", cfg->kernel);
  if ((r = require_runs_n(ctx, adj->n, "crk_stage_begin_checked"))) return r;
  if ((r = alloc_stage(ctx, adj->n, D))) return r;
  const int32_t nr = run_count(ctx);
  if ((r = ensure(ctx->s_chk, 4 * sizeof(float) * (size_t)nr))) return r;
  if ((r = ensure_pinned(ctx, 4 * sizeof(float) * (size_t)nr))) return r;
  ctx->n = adj->n;
  ctx->dim = D;
  ctx->adj = *adj;
  ctx->st = *st;
  ctx->bound = true;
  ctx->have_factors = ctx->have_velgrad = ctx->have_accel = ctx->have_support = ctx->have_volume = false;
  ctx->have_vsig = ctx->have_divcd = ctx->have_rcd = false;
  ctx->check_hb = hb;
  ctx->owen_checked = false;
  // deferred: the adaptive-support and support passes of this stage do
  // nothing once the flag is set
  ctx->gate = valid ? nullptr : ctx->d_flag;
  hipStream_t s = (hipStream_t)stream;
  int32_t n = adj->n;
  float* red = F<float>(ctx->s_chk);
  hipLaunchKernelGGL(k_check_init, dim3(1), dim3(kBlock), 0, s, red, hminmax_dev, nr);
  Geo geo = make_geo(*cfg);
  const RunsDev R = runs_of(ctx);
  // one run: block values reduced by a second kernel; ensembles: atomics
  unsigned nb = span_blocks(ctx, check_blocks(n));
  float* part = nullptr;
  if (ctx->n_runs == 0) {
    nb = nblocks(n) < kCheckPartBlocks ? (nblocks(n) > 0 ? nblocks(n) : 1u) : kCheckPartBlocks;
    if ((r = ensure(ctx->s_chk_part, 8 * sizeof(float) * (size_t)nb))) return r;
    part = F<float>(ctx->s_chk_part);
  }
  CRK_DIM_DISPATCH(D, hipLaunchKernelGGL(k_gather_check<DD>, dim3(nb), dim3(kBlock), 0, s, adj->perm, n, geo,
                                         (const float*)st->positions, (const float*)st->velocities,
                                         (const float*)st->supports, (const float*)st->masses,
                                         (const float*)st->internalEnergies, (const float*)st->alpha0s,
                                         (const float*)st->divergence, (const float*)st->alphas, F<float>(ctx->x),
                                         F<float>(ctx->v), F<float>(ctx->h), F<float>(ctx->m), F<float>(ctx->u),
                                         F<float>(ctx->a0), F<float>(ctx->divp), F<float>(ctx->alpha),
                                         cfg->gamma * (cfg->gamma - 1.f), F<float4>(ctx->geo4),
                                         F<float4>(ctx->vel4), xb, hb, red, hminmax_dev, R, part));
  CRK_LAUNCH_CHECK("k_gather_check");
  if (part) {
    hipLaunchKernelGGL(k_check_finish, dim3(1), dim3(kBlock), 0, s, part, (int32_t)nb, red, hminmax_dev);
    CRK_LAUNCH_CHECK("k_check_finish");
  }
  ctx->check_runs = nr;
  const double sv = cfg->verletScale;
  if (!valid) return verlet_flag_launch(ctx, red, 4, 0, 1, 2, nr, sv, s);  // deferred: into the device flag
  float* hostv = (float*)ctx->pinned.p;
  CRK_HIP(hipMemcpyAsync(hostv, red, 4 * sizeof(float) * (size_t)nr, hipMemcpyDeviceToHost, s));
  CRK_HIP(hipStreamSynchronize(s));
  // every run's list must hold (the lists of all runs are rebuilt together)
  int ok = 1;
  for (int32_t q = 0; q < nr; ++q) {
    const double g = hostv[4 * q + 1];
    ok = ok && g < sv && 2.0 * std::sqrt((double)hostv[4 * q]) <= (sv - g) * (double)hostv[4 * q + 2];
  }
  *valid = ok;
  return CRK_OK;
}

int crk_owen_valid(CrkContext* ctx, const CrkConfig* cfg, int32_t* valid, void* stream) {
  CRK_REQUIRE(ctx && cfg, "crk_owen_valid: null argument");
  if (!ctx->owen_checked) {
    set_error("crk_owen_valid needs crk_stage_begin_checked followed by crk_owen_iteration");
    return CRK_ESTATE;
  }
  const int32_t nr = run_count(ctx);
  if (ctx->check_runs != nr) {
    set_error("crk_owen_valid: the runs changed since crk_stage_begin_checked");
    return CRK_ESTATE;
  }
  hipStream_t s = (hipStream_t)stream;
  float* red = F<float>(ctx->s_chk);  // cleared by crk_stage_begin_checked
  hipLaunchKernelGGL(k_hb_ratio, dim3(span_blocks(ctx, check_blocks(ctx->n))), dim3(kBlock), 0, s, F<float>(ctx->h),
                     ctx->check_hb, ctx->n, red, runs_of(ctx));
  CRK_LAUNCH_CHECK("k_hb_ratio");
  const double sv = cfg->verletScale;
  // the positions' part of the test from the checked gather, h/hb after Owen
  if (!valid) {  // deferred: into the device flag, which then gates the support pass
    ctx->gate = ctx->d_flag;
    return verlet_flag_launch(ctx, red, 4, 0, 3, 2, nr, sv, s);
  }
  ctx->gate = nullptr;
  float* hostv = (float*)ctx->pinned.p;
  CRK_HIP(hipMemcpyAsync(hostv, red, 4 * sizeof(float) * (size_t)nr, hipMemcpyDeviceToHost, s));
  CRK_HIP(hipStreamSynchronize(s));
  int ok = 1;
  for (int32_t q = 0; q < nr; ++q) {
    const double g = hostv[4 * q + 3];
    ok = ok && g < sv && 2.0 * std::sqrt((double)hostv[4 * q]) <= (sv - g) * (double)hostv[4 * q + 2];
  }
  *valid = ok;
  return CRK_OK;
}

int crk_runs_failed(CrkContext* ctx, const CrkConfig* cfg, int32_t which, uint8_t* failed, void* stream) {
  CRK_REQUIRE(ctx && cfg && failed, "crk_runs_failed: null argument");
  CRK_REQUIRE(which == 0 || which == 1, "crk_runs_failed: which must be 0 (stage) or 1 (updated supports)");
  const int32_t nr = run_count(ctx);
  if (ctx->check_runs != nr || !ctx->check_hb) {
    set_error("crk_runs_failed needs crk_stage_begin_checked on the bound runs first");
    return CRK_ESTATE;
  }
  if (which == 1 && !ctx->owen_checked) {
    set_error("crk_runs_failed(1) needs crk_owen_valid after the adaptive-support pass");
    return CRK_ESTATE;
  }
  hipStream_t s = (hipStream_t)stream;
  int r = ensure_pinned(ctx, 4 * sizeof(float) * (size_t)nr);
  if (r) return r;
  float* hostv = (float*)ctx->pinned.p;
  CRK_HIP(hipMemcpyAsync(hostv, ctx->s_chk.p, 4 * sizeof(float) * (size_t)nr, hipMemcpyDeviceToHost, s));
  CRK_HIP(hipStreamSynchronize(s));
  const double sv = cfg->verletScale;
  for (int32_t q = 0; q < nr; ++q) {
    const double g = hostv[4 * q + (which == 0 ? 1 : 3)];
    const bool ok = g < sv && 2.0 * std::sqrt((double)hostv[4 * q]) <= (sv - g) * (double)hostv[4 * q + 2];
    failed[q] = ok ? 0 : 1;
  }
  return CRK_OK;
}

static int owen_launch(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st, const float* hminmax_dev,
                       float* max_ratio_dev, bool final_sweep, void* stream) {
  int r = require_bound(ctx, cfg);
  if (r) return r;
  CRK_REQUIRE(st && st->supports && hminmax_dev, "crk_owen_iteration: null argument");
  CRK_REQUIRE(ctx->nlut >= 2, "Owen table not set (crk_set_owen_lut)");
  hipStream_t s = (hipStream_t)stream;
  int32_t n = ctx->n;
  Geo geo = make_geo(*cfg);
  float C = (float)kernel_C(cfg->kernel, cfg->domain.dim);
  if (max_ratio_dev) { if ((r = fill_f32(max_ratio_dev, 1, 0.f, s))) return r; }
  ctx->have_support = false;
  ctx->have_volume = false;
  ctx->owen_checked = ctx->check_hb != nullptr;
  int GS = ctx->group[CRK_PASS_OWEN];
  if (final_sweep) {
    float* vs = (cfg->switchParams.scheme == CRK_SWITCH_CULLEN_DEHNEN || cfg->switchParams.scheme == CRK_SWITCH_LEARNED)
                    ? F<float>(ctx->vsig) : nullptr;
    OwenFinal fin{F<float>(ctx->v), F<float>(ctx->u), F<float>(ctx->m), cfg->gamma * (cfg->gamma - 1.f),
                  F<float>(ctx->V), vs, F<float2>(ctx->vm2)};
    CRK_KDG_DISPATCH(cfg->kernel, ctx->dim, GS,
        hipLaunchKernelGGL((k_owen<DD, G, KK, true>), dim3(group_blocks(n, G)), dim3(kBlock), 0, s, rows_of(ctx),
                           geo, C, n, F<float>(ctx->x), F<float>(ctx->h), ctx->adj.perm, st->supports,
                           F<double>(ctx->lut_psiH), F<double>(ctx->lut_nh), ctx->nlut, (float)ctx->nh_target,
                           hminmax_dev, max_ratio_dev, F<float4>(ctx->geo4), fin, run_ids(ctx), ctx->gate));
    ctx->have_volume = true;
    ctx->have_vsig = vs != nullptr;
  } else if (use_tiles(ctx)) {
    CRK_KDG12_DISPATCH(cfg->kernel, ctx->dim, GS,
        hipLaunchKernelGGL((k_owen_t<DD, G, KK>), dim3(tile_count(n)), dim3(ctx->threads[CRK_PASS_OWEN]), 0, s,
                           rows_of(ctx), ctx->adj.nbr_loc, ctx->adj.tiles, geo, C, n, F<float>(ctx->h), ctx->adj.perm,
                           st->supports, F<double>(ctx->lut_psiH), F<double>(ctx->lut_nh), ctx->nlut,
                           (float)ctx->nh_target, hminmax_dev, max_ratio_dev, F<float4>(ctx->geo4), run_ids(ctx),
                           ctx->gate));
  } else {
    CRK_KDG_DISPATCH(cfg->kernel, ctx->dim, GS,
        hipLaunchKernelGGL((k_owen<DD, G, KK, false>), dim3(group_blocks(n, G)), dim3(kBlock), 0, s, rows_of(ctx),
                           geo, C, n, F<float>(ctx->x), F<float>(ctx->h), ctx->adj.perm, st->supports,
                           F<double>(ctx->lut_psiH), F<double>(ctx->lut_nh), ctx->nlut, (float)ctx->nh_target,
                           hminmax_dev, max_ratio_dev, F<float4>(ctx->geo4), OwenFinal{}, run_ids(ctx), ctx->gate));
  }
  CRK_LAUNCH_CHECK("k_owen");
  return CRK_OK;
}

int crk_owen_iteration(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st, const float* hminmax_dev,
                       float* max_ratio_dev, void* stream) {
  return owen_launch(ctx, cfg, st, hminmax_dev, max_ratio_dev, false, stream);
}

int crk_owen_support(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st, const float* hminmax_dev,
                     float* max_ratio_dev, void* stream) {
  return owen_launch(ctx, cfg, st, hminmax_dev, max_ratio_dev, true, stream);
}

int crk_monaghan_iteration(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st, const float* hminmax_dev,
                           float* max_ratio_dev, void* stream) {
  int r = require_bound(ctx, cfg);
  if (r) return r;
  CRK_REQUIRE(st && st->supports && hminmax_dev, "crk_monaghan_iteration: null argument");
  hipStream_t s = (hipStream_t)stream;
  int32_t n = ctx->n;
  Geo geo = make_geo(*cfg);
  float C = (float)kernel_C(cfg->kernel, cfg->domain.dim);
  if (max_ratio_dev) { if ((r = fill_f32(max_ratio_dev, 1, 0.f, s))) return r; }
  ctx->have_support = false;
  ctx->have_volume = false;
  ctx->owen_checked = ctx->check_hb != nullptr;  // crk_owen_valid tests the new supports
  int GS = ctx->group[CRK_PASS_OWEN];
  CRK_KDG_DISPATCH(cfg->kernel, ctx->dim, GS,
      hipLaunchKernelGGL((k_monaghan<DD, G, KK>), dim3(group_blocks(n, G)), dim3(kBlock), 0, s, rows_of(ctx), geo, C,
                         n, F<float>(ctx->x), F<float>(ctx->h), F<float>(ctx->m), ctx->adj.perm, st->supports,
                         cfg->targetNeighbors, hminmax_dev, max_ratio_dev, F<float4>(ctx->geo4), run_ids(ctx),
                         ctx->gate));
  CRK_LAUNCH_CHECK("k_monaghan");
  return CRK_OK;
}

int crk_density_gather(CrkContext* ctx, const CrkConfig* cfg, float* rho_user, void* stream) {
  int r = require_bound(ctx, cfg);
  if (r) return r;
  CRK_REQUIRE(rho_user, "crk_density_gather: null output");
  hipStream_t s = (hipStream_t)stream;
  const int32_t n = ctx->n;
  Geo geo = make_geo(*cfg);
  float C = (float)kernel_C(cfg->kernel, cfg->domain.dim);
  int GS = ctx->group[CRK_PASS_OWEN];
  CRK_KDG_DISPATCH(cfg->kernel, ctx->dim, GS,
      hipLaunchKernelGGL((k_density<DD, G, KK>), dim3(group_blocks(n, G)), dim3(kBlock), 0, s, rows_of(ctx), geo, C, n,
                         F<float>(ctx->x), F<float>(ctx->h), F<float>(ctx->m), ctx->adj.perm, rho_user));
  CRK_LAUNCH_CHECK("k_density");
  return CRK_OK;
}

int crk_support(CrkContext* ctx, const CrkConfig* cfg, CrkEdges* ed, void* stream) {
  int r = require_bound(ctx, cfg);
  if (r) return r;
  CRK_REQUIRE(ed && ed->count && (ctx->adj.n_edges == 0 || ed->nbr), "crk_support: pair buffers (count, nbr) missing");
  hipStream_t s = (hipStream_t)stream;
  int32_t n = ctx->n;
  Geo geo = make_geo(*cfg);
  float C = (float)kernel_C(cfg->kernel, cfg->domain.dim);
  float cfac = cfg->gamma * (cfg->gamma - 1.f);
  float* vs = (cfg->switchParams.scheme == CRK_SWITCH_CULLEN_DEHNEN || cfg->switchParams.scheme == CRK_SWITCH_LEARNED)
                  ? F<float>(ctx->vsig) : nullptr;
  int GS = ctx->group[CRK_PASS_SUPPORT];
  const bool tiled = use_tiles(ctx);
  CRK_REQUIRE(!tiled || ctx->adj.n_edges == 0 || ed->loc, "crk_support: the adjacency has tiles: pair buffer loc missing");
  if (tiled) {
#define SUPPORT_T(VS)                                                                                            \
    CRK_KDG12_DISPATCH(cfg->kernel, ctx->dim, GS,                                                              \
        hipLaunchKernelGGL((k_support_t<DD, G, KK, VS>), dim3(tile_count(n)), dim3(ctx->threads[CRK_PASS_SUPPORT]), \
                           0, s, rows_of(ctx),                                                                  \
                           ctx->adj.nbr_loc, ctx->adj.tiles, geo, C, n, ed->nbr, ed->loc, ed->count, vs,       \
                           F<float>(ctx->V), F<float4>(ctx->geo4), F<float4>(ctx->vel4), F<float>(ctx->m),    \
                           F<float2>(ctx->vm2), ctx->gate))
    if (vs) { SUPPORT_T(true); } else { SUPPORT_T(false); }
#undef SUPPORT_T
  } else if (vs) {
    CRK_KDG_DISPATCH(cfg->kernel, ctx->dim, GS,
        hipLaunchKernelGGL((k_support<DD, G, KK, true>), dim3(group_blocks(n, G)), dim3(kBlock), 0, s, rows_of(ctx),
                           geo, C, n, ed->nbr, ed->count, vs, F<float>(ctx->V), F<float4>(ctx->geo4),
                           F<float4>(ctx->vel4), F<float>(ctx->m), F<float2>(ctx->vm2), ctx->gate));
  } else {
    CRK_KDG_DISPATCH(cfg->kernel, ctx->dim, GS,
        hipLaunchKernelGGL((k_support<DD, G, KK, false>), dim3(group_blocks(n, G)), dim3(kBlock), 0, s, rows_of(ctx),
                           geo, C, n, ed->nbr, ed->count, vs, F<float>(ctx->V), F<float4>(ctx->geo4),
                           F<float4>(ctx->vel4), F<float>(ctx->m), F<float2>(ctx->vm2), ctx->gate));
  }
  CRK_LAUNCH_CHECK("k_support");
  ctx->pair_cnt = ed->count;
  ctx->pair_nbr = ed->nbr;
  ctx->pair_loc = tiled ? ed->loc : nullptr;
  ctx->have_support = true;
  ctx->have_volume = true;
  ctx->have_vsig = vs != nullptr;
  return CRK_OK;
}

static int factors_launch(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st, CrkEdges* ed,
                          int32_t* n_singular_dev, void* stream) {
  CRK_REQUIRE(st && st->densities && st->entropies && st->pressures && st->soundspeeds,
              "crk_factors_eos: state output buffers missing");
  hipStream_t s = (hipStream_t)stream;
  int32_t n = ctx->n;
  Geo geo = make_geo(*cfg);
  float C = (float)kernel_C(cfg->kernel, cfg->domain.dim);
  MomOut mo{F<float>(ctx->A), F<float>(ctx->B), F<float>(ctx->gA), F<float>(ctx->gB), F<float>(ctx->rho),
            F<float>(ctx->P), F<float>(ctx->c), F<float>(ctx->sP), st->densities, st->entropies, st->pressures,
            st->soundspeeds, n_singular_dev, F<float2>(ctx->vm2), F<float>(ctx->v), F<float4>(ctx->vg4),
            F<float>(ctx->mr)};
  int GM = ctx->group[CRK_PASS_MOMENTS];
  if (ed) {
    CRK_KDG_DISPATCH(cfg->kernel, ctx->dim, GM,
        hipLaunchKernelGGL((k_moments<DD, G, KK, true>), dim3(group_blocks(n, G)), dim3(kBlock), 0, s, ARows{},
                           rows_of(ctx), geo, C, n, F<float>(ctx->x), F<float>(ctx->h), F<float>(ctx->V),
                           F<float>(ctx->m), F<float>(ctx->u), cfg->gamma, ctx->adj.perm, mo, ed->nbr, ed->count,
                           F<float4>(ctx->geo4)));
    ctx->pair_cnt = ed->count;
    ctx->pair_nbr = ed->nbr;
    ctx->pair_loc = nullptr;
    ctx->have_support = true;
  } else if (use_tiles(ctx) && ctx->pair_loc) {
    CRK_KDG12_DISPATCH(cfg->kernel, ctx->dim, GM,
        hipLaunchKernelGGL((k_moments_t<DD, G, KK>), dim3(tile_count(n)), dim3(ctx->threads[CRK_PASS_MOMENTS]), 0, s,
                           arows_of(ctx), ctx->pair_loc, ctx->adj.tiles, geo, C, n, F<float>(ctx->x), F<float>(ctx->h),
                           F<float>(ctx->V), F<float>(ctx->m), F<float>(ctx->u), cfg->gamma, ctx->adj.perm, mo,
                           F<float4>(ctx->geo4)));
  } else {
    CRK_KDG_DISPATCH(cfg->kernel, ctx->dim, GM,
        hipLaunchKernelGGL((k_moments<DD, G, KK, false>), dim3(group_blocks(n, G)), dim3(kBlock), 0, s,
                           arows_of(ctx), Rows{}, geo, C, n, F<float>(ctx->x), F<float>(ctx->h), F<float>(ctx->V),
                           F<float>(ctx->m), F<float>(ctx->u), cfg->gamma, ctx->adj.perm, mo, nullptr, nullptr,
                           F<float4>(ctx->geo4)));
  }
  CRK_LAUNCH_CHECK("k_moments");
  ctx->have_factors = true;
  return CRK_OK;
}

int crk_factors_eos(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st, int32_t* n_singular_dev, void* stream) {
  int r = require_bound(ctx, cfg);
  if (r) return r;
  if (!ctx->have_support) { set_error("crk_factors_eos before crk_support"); return CRK_ESTATE; }
  return factors_launch(ctx, cfg, st, nullptr, n_singular_dev, stream);
}

int crk_pairs_factors_eos(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st, CrkEdges* ed,
                          int32_t* n_singular_dev, void* stream) {
  int r = require_bound(ctx, cfg);
  if (r) return r;
  CRK_REQUIRE(ed && ed->count && (ctx->adj.n_edges == 0 || ed->nbr),
              "crk_pairs_factors_eos: pair buffers (count, nbr) missing");
  if (!ctx->have_volume) {
    set_error("crk_pairs_factors_eos needs V from crk_owen_support (or crk_support) on the bound adjacency");
    return CRK_ESTATE;
  }
  return factors_launch(ctx, cfg, st, ed, n_singular_dev, stream);
}

int crk_edges_export(CrkContext* ctx, int32_t dim, const CrkAdjacency* adj, const CrkEdges* ed,
                     const float* masses, float dt, float* ap_out, float* av_out, float* f_out, void* stream) {
  CRK_REQUIRE(ctx && adj && ed, "crk_edges_export: null argument");
  CRK_REQUIRE(dim >= 1 && dim <= 3, "dim must be 1..3");
  CRK_REQUIRE(adj->n > 0 && adj->row_offsets && adj->perm && ed->count, "crk_edges_export: adjacency/pairs missing");
  if (adj->n_edges == 0) return CRK_OK;
  CRK_REQUIRE(adj->nbr && ed->nbr, "crk_edges_export: neighbour lists missing");
  CRK_REQUIRE((!ap_out && !av_out) || (ed->pf && ed->th), "crk_edges_export: pair data missing");
  const DtArg dta = dt_arg(ctx, dt);
  CRK_REQUIRE(!f_out || ed->f || (ed->rec && masses && ed->pf && (dt > 0.f || dta.p)),
              "crk_edges_export: f needs edges->f, or edges->rec with masses, pair forces and dt > 0");
  int r = dta.p ? require_runs_n(ctx, adj->n, "crk_edges_export") : 0;
  if (r) return r;
  hipStream_t s = (hipStream_t)stream;
  int32_t n = adj->n;
  CRK_DIM_DISPATCH(dim, hipLaunchKernelGGL(k_edges_export<DD>, dim3(nblocks(n)), dim3(kBlock), 0, s, adj->row_offsets,
                                           adj->nbr, ed->count, ed->nbr, n, adj->perm, masses, dta, ed->pf, ed->th,
                                           (const float4*)ed->rec, ed->f, ap_out, av_out, f_out));
  CRK_LAUNCH_CHECK("k_edges_export");
  return CRK_OK;
}

int crk_export_field(CrkContext* ctx, int32_t field, float* out_user, void* stream) {
  CRK_REQUIRE(ctx && out_user, "crk_export_field: null argument");
  if (!ctx->bound) { set_error("no stage bound"); return CRK_ESTATE; }
  hipStream_t s = (hipStream_t)stream;
  int32_t n = ctx->n;
  int D = ctx->dim;
  const float* src = nullptr;
  int comps = 1;
  const bool have[9] = {ctx->have_volume, ctx->have_factors, ctx->have_factors, ctx->have_factors, ctx->have_factors,
                        ctx->have_velgrad, ctx->have_divcd, ctx->have_vsig, ctx->have_rcd};
  if (field >= 0 && field < 9 && !have[field]) {
    set_error("crk_export_field: field %d was not computed in the bound stage (v_sig, div and R only with the "
              "Cullen-Dehnen switch)", field);
    return CRK_ESTATE;
  }
  switch (field) {
    case 0: src = F<float>(ctx->V); break;
    case 1: src = F<float>(ctx->A); break;
    case 2: src = F<float>(ctx->B); comps = D; break;
    case 3: src = F<float>(ctx->gA); comps = D; break;
    case 4: src = F<float>(ctx->gB); comps = D * D; break;
    case 5: src = F<float>(ctx->J); comps = D * D; break;
    case 6: src = F<float>(ctx->divcd); break;
    case 7: src = F<float>(ctx->vsig); break;
    case 8: src = F<float>(ctx->Rcd); break;
    default: set_error("crk_export_field: unknown field %d", field); return CRK_EINVAL;
  }
  switch (comps) {
    case 1: hipLaunchKernelGGL(k_scatter_field<1>, dim3(nblocks(n)), dim3(kBlock), 0, s, src, ctx->adj.perm, n, out_user); break;
    case 2: hipLaunchKernelGGL(k_scatter_field<2>, dim3(nblocks(n)), dim3(kBlock), 0, s, src, ctx->adj.perm, n, out_user); break;
    case 3: hipLaunchKernelGGL(k_scatter_field<3>, dim3(nblocks(n)), dim3(kBlock), 0, s, src, ctx->adj.perm, n, out_user); break;
    case 4: hipLaunchKernelGGL(k_scatter_field<4>, dim3(nblocks(n)), dim3(kBlock), 0, s, src, ctx->adj.perm, n, out_user); break;
    case 9: hipLaunchKernelGGL(k_scatter_field<9>, dim3(nblocks(n)), dim3(kBlock), 0, s, src, ctx->adj.perm, n, out_user); break;
    default: set_error("bad component count"); return CRK_EINVAL;
  }
  CRK_LAUNCH_CHECK("k_scatter_field");
  return CRK_OK;
}

}  // extern "C"
