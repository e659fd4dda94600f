// passes.hip -- the pair passes of one CRKSPH derivative evaluation
// (src/warpSPH/schemes/crkSPH.py:19-366), run on Morton-sorted copies of the
// state held by the context.  One thread owns one sorted particle row; pair
// sums are Kahan-compensated FP32, the CRK moment solve is FP64.
//
// Pass order inside a stage:
//   crk_stage_begin      gather x, v, h, m, u, alpha0, divergence to sorted order
//   crk_owen_iteration   psi0 -> n_h -> new h           (optimalSupportOwen.py:70-114)
//   crk_factors_eos      V, then moments -> A, B, gradA, gradB (FP64 solve),
//                        rho = A (S0 + B.S1), EOS      (computeCRKFactors; idealGas.py)
//   crk_velocity_gradient J (CRK, scatter), drhodt, and the Cullen divergence
//                        and signal speed              (crkSPH.py:158-172; wp_vsig.py)
//   crk_switch           R, Xi, alpha update           (CullenDehnen2010.py:218-331)
//   crk_accel_dudt       dvdt, ap, av, dudt, dEdt      (crk/accel.py, crk/dudt.py)
//   crk_balance          f_ij                          (compSPH/balance.py)
#include <cmath>

#include "common.hpp"

namespace crk {

struct Rows {
  const int32_t* off;
  const int32_t* nbr;
};

// ------------------------------------------------------------ gather
template <int D>
__global__ void k_gather(const int32_t* __restrict__ perm, int32_t n, const float* __restrict__ X,
                         const float* __restrict__ Vel, const float* __restrict__ H, const float* __restrict__ M,
                         const float* __restrict__ U, const float* __restrict__ A0, const float* __restrict__ DIV,
                         const float* __restrict__ AL, float* x, float* v, float* h, float* m, float* u, float* a0,
                         float* divp, float* alpha) {
  int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int32_t i = perm[k];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    x[(int64_t)k * D + d] = X[(int64_t)i * D + d];
    v[(int64_t)k * D + d] = Vel[(int64_t)i * D + d];
  }
  h[k] = H[i];
  m[k] = M[i];
  u[k] = U[i];
  a0[k] = A0 ? A0[i] : 1.f;
  divp[k] = DIV ? DIV[i] : 0.f;
  alpha[k] = AL ? AL[i] : 1.f;
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

template <int D>
__global__ void k_owen(Rows rows, Geo geo, int kernel, float C, int32_t n, const float* __restrict__ x,
                       float* __restrict__ h, const int32_t* __restrict__ perm, float* __restrict__ supports,
                       const double* __restrict__ lut_psiH, const double* __restrict__ lut_nh, int nlut, float nhT,
                       const float* __restrict__ hminmax, float* __restrict__ ratio) {
  int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  float xi[3];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = x[(int64_t)k * D + d];
  float hi = h[k];
  Kahan acc;
  for (int32_t e = rows.off[k]; e < rows.off[k + 1]; ++e) {
    int32_t j = rows.nbr[e];
    float dx[3];
    min_image<D>(geo, xi, &x[(int64_t)j * D], dx);
    float r = sqrtf(dot<D>(dx, dx));
    float q = r / hi;
    if (!(q < 1.f)) continue;
    acc.add(C * fabsf(kder<float>(kernel, D, q)));
  }
  float ps = powf(acc.s, 1.f / (float)D);
  float nh = (float)lut_interp(lut_psiH, lut_nh, nlut, (double)ps);
  float sr = nhT / nh;
  float a = sr >= 1.f ? 0.4f * (1.f + 1.f / (sr * sr * sr)) : 0.4f * (1.f + sr * sr);
  float hn = (1.f - a + a * sr) * hi;
  float lo = hminmax[0] * 0.25f, up = hminmax[1] * 4.0f;
  hn = hn < lo ? lo : (hn > up ? up : hn);
  h[k] = hn;
  supports[perm[k]] = hn;
  if (ratio) atomic_max_pos(ratio, fabsf(hn / hi - 1.f));
}

// ------------------------------------------------------------ CRK volume
template <int D>
__global__ void k_volume(Rows rows, Geo geo, DevKernel<D> K, int32_t n, const float* __restrict__ x,
                         const float* __restrict__ h, float* __restrict__ V) {
  int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  float xi[3];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = x[(int64_t)k * D + d];
  float hi = h[k];
  Kahan acc;
  for (int32_t e = rows.off[k]; e < rows.off[k + 1]; ++e) {
    int32_t j = rows.nbr[e];
    float dx[3];
    min_image<D>(geo, xi, &x[(int64_t)j * D], dx);
    float r = sqrtf(dot<D>(dx, dx));
    if (!(r < hi)) continue;
    acc.add(K.W(r, hi));
  }
  V[k] = 1.f / acc.s;
}

// ------------------------------------------------------------ CRK moments
// Linear CRK (Frontiere, Raskin & Owen 2017) with W_ij = W(x_ij, h_j):
//   m0 = sum V_j W, m1 = sum V_j x W, m2 = sum V_j x x W and x_i-gradients,
//   B = -m2^-1 m1, A = 1/(m0 + B.m1), grad A = -A^2 (dm0 + 2 B.dm1 + B.dm2.B),
//   grad B = -m2^-1 (dm1 + dm2 B); singular m2 -> A = 1/m0, B = 0.
// CRK density rho = A (S0 + B.S1), S0 = sum m_j W, S1 = sum m_j x W; then the
// ideal-gas EOS with u given (eos/idealGas.py:6-38).
template <int D>
__global__ void __launch_bounds__(kBlock) k_moments(Rows rows, Geo geo, DevKernel<D> K, int32_t n,
                                                    const float* __restrict__ x, const float* __restrict__ h,
                                                    const float* __restrict__ V, const float* __restrict__ m,
                                                    const float* __restrict__ u, float gamma, float* __restrict__ Ao,
                                                    float* __restrict__ Bo, float* __restrict__ gAo,
                                                    float* __restrict__ gBo, float* __restrict__ rho,
                                                    float* __restrict__ Po, float* __restrict__ co,
                                                    const int32_t* __restrict__ perm, float* __restrict__ dens_u,
                                                    float* __restrict__ ent_u, float* __restrict__ pres_u,
                                                    float* __restrict__ cs_u, int32_t* __restrict__ nsing) {
  constexpr int NM2 = D * (D + 1) / 2;
  int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  float xi[3];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = x[(int64_t)k * D + d];
  Kahan m0, m1[D], m2[NM2], dm0[D], dm1[D * D], dm2[NM2 * D], S0, S1[D];
  for (int32_t e = rows.off[k]; e < rows.off[k + 1]; ++e) {
    int32_t j = rows.nbr[e];
    float dx[3];
    min_image<D>(geo, xi, &x[(int64_t)j * D], dx);
    float r = sqrtf(dot<D>(dx, dx));
    float hj = h[j];
    if (!(r < hj)) continue;
    float W = K.W(r, hj);
    float g[3];
    K.gradW(dx, r, hj, g);
    float Vj = V[j];
    float mj = m[j];
    m0.add(Vj * W);
#pragma unroll
    for (int a = 0; a < D; ++a) m1[a].add(Vj * dx[a] * W);
    {
      int p = 0;
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = a; b < D; ++b, ++p) m2[p].add(Vj * dx[a] * dx[b] * W);
    }
#pragma unroll
    for (int gm = 0; gm < D; ++gm) dm0[gm].add(Vj * g[gm]);
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int gm = 0; gm < D; ++gm) dm1[a * D + gm].add(Vj * (dx[a] * g[gm] + (a == gm ? W : 0.f)));
    {
      int p = 0;
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = a; b < D; ++b, ++p)
#pragma unroll
          for (int gm = 0; gm < D; ++gm) {
            float lin = (a == gm ? dx[b] : 0.f) + (b == gm ? dx[a] : 0.f);
            dm2[p * D + gm].add(Vj * (dx[a] * dx[b] * g[gm] + lin * W));
          }
    }
    S0.add(mj * W);
#pragma unroll
    for (int a = 0; a < D; ++a) S1[a].add(mj * dx[a] * W);
  }
  // FP64 solve
  double M0 = m0.s, M1[3], M2[3][3], DM0[3], DM1[3][3], DM2[3][3][3];
#pragma unroll
  for (int a = 0; a < D; ++a) M1[a] = m1[a].s;
  {
    int p = 0;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int b = a; b < D; ++b, ++p) { M2[a][b] = m2[p].s; M2[b][a] = m2[p].s; }
  }
#pragma unroll
  for (int gm = 0; gm < D; ++gm) DM0[gm] = dm0[gm].s;
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int gm = 0; gm < D; ++gm) DM1[a][gm] = dm1[a * D + gm].s;
  {
    int p = 0;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int b = a; b < D; ++b, ++p)
#pragma unroll
        for (int gm = 0; gm < D; ++gm) { DM2[a][b][gm] = dm2[p * D + gm].s; DM2[b][a][gm] = dm2[p * D + gm].s; }
  }
  double inv[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) inv[a][b] = 0.0;
  bool sing;
  if (D == 1) {
    sing = !(M2[0][0] > 0.0);
    if (!sing) inv[0][0] = 1.0 / M2[0][0];
  } else if (D == 2) {
    double det = M2[0][0] * M2[1][1] - M2[0][1] * M2[1][0];
    double tr = 0.5 * (M2[0][0] + M2[1][1]);
    sing = !(det > 1e-10 * tr * tr);
    if (!sing) {
      inv[0][0] = M2[1][1] / det; inv[1][1] = M2[0][0] / det;
      inv[0][1] = -M2[0][1] / det; inv[1][0] = -M2[1][0] / det;
    }
  } else {
    double c00 = M2[1][1] * M2[2][2] - M2[1][2] * M2[2][1];
    double c01 = M2[1][2] * M2[2][0] - M2[1][0] * M2[2][2];
    double c02 = M2[1][0] * M2[2][1] - M2[1][1] * M2[2][0];
    double det = M2[0][0] * c00 + M2[0][1] * c01 + M2[0][2] * c02;
    double tr = (M2[0][0] + M2[1][1] + M2[2][2]) / 3.0;
    sing = !(det > 1e-10 * tr * tr * tr);
    if (!sing) {
      inv[0][0] = c00 / det; inv[1][0] = c01 / det; inv[2][0] = c02 / det;
      inv[0][1] = (M2[0][2] * M2[2][1] - M2[0][1] * M2[2][2]) / det;
      inv[1][1] = (M2[0][0] * M2[2][2] - M2[0][2] * M2[2][0]) / det;
      inv[2][1] = (M2[0][1] * M2[2][0] - M2[0][0] * M2[2][1]) / det;
      inv[0][2] = (M2[0][1] * M2[1][2] - M2[0][2] * M2[1][1]) / det;
      inv[1][2] = (M2[0][2] * M2[1][0] - M2[0][0] * M2[1][2]) / det;
      inv[2][2] = (M2[0][0] * M2[1][1] - M2[0][1] * M2[1][0]) / det;
    }
  }
  double Ad, Bd[3] = {0, 0, 0}, gAd[3] = {0, 0, 0}, gBd[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) gBd[a][b] = 0.0;
  if (sing) {
    if (nsing) atomicAdd(nsing, 1);
    Ad = 1.0 / M0;
#pragma unroll
    for (int gm = 0; gm < D; ++gm) gAd[gm] = -DM0[gm] / (M0 * M0);
  } else {
#pragma unroll
    for (int a = 0; a < D; ++a) {
      double s = 0.0;
#pragma unroll
      for (int b = 0; b < D; ++b) s += inv[a][b] * M1[b];
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
        for (int b = 0; b < D; ++b) dinvA += Bd[a] * DM2[a][b][gm] * Bd[b];
      gAd[gm] = -Ad * Ad * dinvA;
      double tmp[3];
#pragma unroll
      for (int b = 0; b < D; ++b) {
        double t2 = DM1[b][gm];
#pragma unroll
        for (int dd = 0; dd < D; ++dd) t2 += DM2[b][dd][gm] * Bd[dd];
        tmp[b] = t2;
      }
#pragma unroll
      for (int a = 0; a < D; ++a) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < D; ++b) s += inv[a][b] * tmp[b];
        gBd[a][gm] = -s;
      }
    }
  }
  float Af = (float)Ad;
  float Bf[3];
  Ao[k] = Af;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    Bf[a] = (float)Bd[a];
    Bo[(int64_t)k * D + a] = Bf[a];
    gAo[(int64_t)k * D + a] = (float)gAd[a];
  }
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int gm = 0; gm < D; ++gm) gBo[(int64_t)k * D * D + a * D + gm] = (float)gBd[a][gm];
  float bs = S0.s;
#pragma unroll
  for (int a = 0; a < D; ++a) bs = bs + Bf[a] * S1[a].s;
  float rh = Af * bs;
  rho[k] = rh;
  // EOS
  float ui = u[k];
  float cc = sqrtf(fabsf(ui) * gamma * (gamma - 1.f));
  float P = (gamma - 1.f) * rh * ui;
  float s = ui * (gamma - 1.f) * powf(rh, 1.f - gamma);
  Po[k] = P;
  co[k] = cc;
  int32_t i = perm[k];
  dens_u[i] = rh;
  pres_u[i] = P;
  cs_u[i] = cc;
  ent_u[i] = s;
}

// ------------------------------------------------------------ CRK gradient
template <int D>
__device__ inline void grad_wr(const DevKernel<D>& K, const float* xab, float r, float hh, float Aa,
                               const float* Ba, const float* gAa, const float* gBa, float* out) {
  float W = K.W(r, hh);
  float g[3];
  K.gradW(xab, r, hh, g);
  float bx = 1.f;
#pragma unroll
  for (int a = 0; a < D; ++a) bx = bx + Ba[a] * xab[a];
#pragma unroll
  for (int gm = 0; gm < D; ++gm) {
    float gbx = Ba[gm];
#pragma unroll
    for (int a = 0; a < D; ++a) gbx = gbx + gBa[a * D + gm] * xab[a];
    out[gm] = gAa[gm] * bx * W + Aa * gbx * W + Aa * bx * g[gm];
  }
}

template <int D>
struct Corr {
  float A, B[D], gA[D], gB[D * D];
  __device__ inline void load(const float* Ap, const float* Bp, const float* gAp, const float* gBp, int32_t k) {
    A = Ap[k];
#pragma unroll
    for (int a = 0; a < D; ++a) { B[a] = Bp[(int64_t)k * D + a]; gA[a] = gAp[(int64_t)k * D + a]; }
#pragma unroll
    for (int a = 0; a < D * D; ++a) gB[a] = gBp[(int64_t)k * D * D + a];
  }
};

// velocity gradient J_i[a][g] = sum_j V_j (v_j - v_i)[a] d_g W^R_ij (scatter),
// fused with the Cullen-Dehnen divergence (super-symmetric support
// max(h_i, h_j), m/rho volumes) and the signal speed (every adjacency entry).
template <int D, bool CULLEN>
__global__ void __launch_bounds__(kBlock) k_velgrad(Rows rows, Geo geo, DevKernel<D> K, int32_t n,
                                                    const float* __restrict__ x, const float* __restrict__ h,
                                                    const float* __restrict__ v, const float* __restrict__ V,
                                                    const float* __restrict__ m, const float* __restrict__ rho,
                                                    const float* __restrict__ c, const float* __restrict__ Ap,
                                                    const float* __restrict__ Bp, const float* __restrict__ gAp,
                                                    const float* __restrict__ gBp, float* __restrict__ J,
                                                    float* __restrict__ divcd, float* __restrict__ vsig,
                                                    const int32_t* __restrict__ perm, float* __restrict__ drhodt_u) {
  int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  float xi[3], vi[3];
#pragma unroll
  for (int d = 0; d < D; ++d) { xi[d] = x[(int64_t)k * D + d]; vi[d] = v[(int64_t)k * D + d]; }
  float hi = h[k];
  float ci = c[k];
  Corr<D> cr;
  cr.load(Ap, Bp, gAp, gBp, k);
  Kahan acc[D * D];
  Kahan dacc;
  float vs = 0.f;
  for (int32_t e = rows.off[k]; e < rows.off[k + 1]; ++e) {
    int32_t j = rows.nbr[e];
    float dx[3];
    min_image<D>(geo, xi, &x[(int64_t)j * D], dx);
    float r = sqrtf(dot<D>(dx, dx));
    float vj[3];
#pragma unroll
    for (int d = 0; d < D; ++d) vj[d] = v[(int64_t)j * D + d];
    float hj = h[j];
    if (CULLEN) {
      float mu = 0.f;
#pragma unroll
      for (int a = 0; a < D; ++a) mu = mu + (vi[a] - vj[a]) * dx[a];
      mu = mu / (r + 1.0e-14f * hi);
      float cb = 0.5f * (ci + c[j]);
      float vv = cb - mu;
      if (mu > 0.f) vv = 0.f;
      vs = vv > vs ? vv : vs;
      float H = fmaxf(hi, hj);
      if (r < H) {
        float g[3];
        K.gradW(dx, r, H, g);
        float dt_ = 0.f;
#pragma unroll
        for (int a = 0; a < D; ++a) dt_ = dt_ + (vj[a] - vi[a]) * g[a];
        dacc.add(m[j] / rho[j] * dt_);
      }
    }
    if (!(r < hj)) continue;
    float g[3];
    grad_wr<D>(K, dx, r, hj, cr.A, cr.B, cr.gA, cr.gB, g);
    float Vj = V[j];
#pragma unroll
    for (int a = 0; a < D; ++a) {
      float dv = vj[a] - vi[a];
#pragma unroll
      for (int gm = 0; gm < D; ++gm) acc[a * D + gm].add(Vj * dv * g[gm]);
    }
  }
  float tr = 0.f;
#pragma unroll
  for (int a = 0; a < D * D; ++a) J[(int64_t)k * D * D + a] = acc[a].s;
#pragma unroll
  for (int a = 0; a < D; ++a) tr = tr + acc[a * D + a].s;
  drhodt_u[perm[k]] = -tr * rho[k];
  if (CULLEN) {
    divcd[k] = dacc.s;
    vsig[k] = vs;
  }
}

// Cullen & Dehnen (2010) as in CullenDehnen2010.py:218-331: R (gather
// interpolation of rho sign(div)), Xi, A = Xi max(0, -ddiv/dt), local alpha,
// and the alpha0 relaxation.
template <int D>
__global__ void k_cullen(Rows rows, Geo geo, DevKernel<D> K, int32_t n, const float* __restrict__ x,
                         const float* __restrict__ h, const float* __restrict__ m, const float* __restrict__ rho,
                         const float* __restrict__ divcd, const float* __restrict__ vsig,
                         const float* __restrict__ divp, float* __restrict__ a0, float* __restrict__ alpha,
                         float* __restrict__ Rcd, float dt, float fkern, float amin, float amax, float beta,
                         int limitXi, int evolve, const int32_t* __restrict__ perm, float* __restrict__ alphas_u,
                         float* __restrict__ alpha0s_u) {
  int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  float xi[3];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = x[(int64_t)k * D + d];
  float hi = h[k];
  Kahan acc;
  for (int32_t e = rows.off[k]; e < rows.off[k + 1]; ++e) {
    int32_t j = rows.nbr[e];
    float dx[3];
    min_image<D>(geo, xi, &x[(int64_t)j * D], dx);
    float r = sqrtf(dot<D>(dx, dx));
    if (!(r < hi)) continue;
    float dj = divcd[j];
    float sg = dj > 0.f ? 1.f : (dj < 0.f ? -1.f : 0.f);
    float term = rho[j] * sg;
    acc.add(m[j] / rho[j] * term * K.W(r, hi));
  }
  float Rv = 1.f / rho[k] * acc.s;
  Rcd[k] = Rv;
  float t1 = 1.f - Rv;
  float t2 = t1 * t1;
  float nom = beta * (t2 * t2) * divcd[k];
  float Xi = limitXi ? (nom * nom + 1e-14f * hi) / (nom * nom + 0.f + 1e-14f * hi)
                     : (nom * nom) / (nom * nom + 0.f + 1e-14f * hi);
  float ddiv = (divcd[k] - divp[k]) / dt;
  float md = -ddiv;
  float Ai = Xi * (md > 0.f ? md : 0.f);
  float vsg = fabsf(vsig[k]);
  float scaling = hi * hi * Ai / (vsg * vsg + hi * hi * Ai + 1e-14f * hi);
  float al = amax * scaling;
  al = al < amin ? amin : (al > amax ? amax : al);
  float tau = hi * fkern / (0.1f * vsg);
  float a0k = a0[k];
  float adot = (al - a0k) / tau;
  float na = al > a0k ? al : a0k;
  na = al < na ? na + adot * dt : na;
  na = na < amin ? amin : (na > amax ? amax : na);
  alpha[k] = na;
  int32_t i = perm[k];
  alphas_u[i] = na;
  if (evolve) { a0[k] = na; if (alpha0s_u) alpha0s_u[i] = na; }
}

// ------------------------------------------------------------ accel + dudt
struct LimParams {
  float eta_fold, eta_crit;
  int crk, vl, off, on;
};

template <int D>
__device__ inline float limiter_phi(const LimParams& lp, const float* dx, float hi, float hj, const float* Ji,
                                    const float* Jj) {
  float factor = 1.f;
  if (lp.crk) {
    float ei = 0.f, ej = 0.f;
#pragma unroll
    for (int a = 0; a < D; ++a) { float t1 = dx[a] / hi; float t2 = dx[a] / hj; ei = ei + t1 * t1; ej = ej + t2 * t2; }
    ei = sqrtf(ei);
    ej = sqrtf(ej);
    float eta = ei < ej ? ei : ej;
    if (eta < lp.eta_crit) { float t = (eta - lp.eta_crit) / lp.eta_fold; factor = expf(-(t * t)); }
  }
  float phi = 0.f;
  if (lp.vl) {
    float xt[3];
#pragma unroll
    for (int a = 0; a < D; ++a) xt[a] = 0.5f * dx[a];
    float gi = 0.f, gj = 0.f;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      float ci = 0.f, cj = 0.f;
#pragma unroll
      for (int gm = 0; gm < D; ++gm) { ci = ci + Ji[a * D + gm] * xt[gm]; cj = cj + Jj[a * D + gm] * xt[gm]; }
      gi = gi + ci * xt[a];
      gj = gj + cj * xt[a];
    }
    float sj = gj >= 0.f ? 1.f : -1.f;
    float si = gi >= 0.f ? 1.f : -1.f;
    float ri = gi / (sj * fabsf(gj));
    float rj = gj / (si * fabsf(gi));
    if (ri != ri) ri = 1.f;
    if (rj != rj) rj = 1.f;
    float rr = ri < rj ? ri : rj;
    float vl = 0.f;
    if (rr > 0.f) { float q = 2.f / (1.f + rr); vl = rr * q * q; }
    phi = vl * factor;
  }
  if (lp.off) phi = 0.f;
  if (lp.on) phi = 1.f;
  phi = phi < 1.f ? phi : 1.f;
  phi = phi > 0.f ? phi : 0.f;
  return phi;
}

template <int D>
__global__ void __launch_bounds__(kBlock) k_accel_dudt(
    Rows rows, Geo geo, DevKernel<D> K, LimParams lp, float kscale, float Cl, float Cq, int alphaQ, int32_t n,
    const float* __restrict__ x, const float* __restrict__ h, const float* __restrict__ v,
    const float* __restrict__ m, const float* __restrict__ V, const float* __restrict__ rho,
    const float* __restrict__ P, const float* __restrict__ c, const float* __restrict__ J,
    const float* __restrict__ Ap, const float* __restrict__ Bp, const float* __restrict__ gAp,
    const float* __restrict__ gBp, const float* __restrict__ alpha, float* __restrict__ ap,
    float* __restrict__ av, float* __restrict__ dvdt_s, const int32_t* __restrict__ perm,
    float* __restrict__ dvdt_u, float* __restrict__ dudt_u, float* __restrict__ dEdt_u,
    const float* __restrict__ drhodt_u, float* __restrict__ div_u) {
  int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  float xi[3], vi[3], Ji[D * D];
#pragma unroll
  for (int d = 0; d < D; ++d) { xi[d] = x[(int64_t)k * D + d]; vi[d] = v[(int64_t)k * D + d]; }
#pragma unroll
  for (int a = 0; a < D * D; ++a) Ji[a] = J[(int64_t)k * D * D + a];
  float hi = h[k], rhoi = rho[k], Pi = P[k], ci = c[k], Vi = V[k], mi = m[k];
  float ali = alphaQ ? alpha[k] : 1.f;
  Corr<D> cri;
  cri.load(Ap, Bp, gAp, gBp, k);
  float smi = hi / kscale;
  Kahan acc[D], accu;
  for (int32_t e = rows.off[k]; e < rows.off[k + 1]; ++e) {
    int32_t j = rows.nbr[e];
    float dx[3];
    min_image<D>(geo, xi, &x[(int64_t)j * D], dx);
    float r = sqrtf(dot<D>(dx, dx));
    float hj = h[j];
    if (!(r < fmaxf(hi, hj))) {
#pragma unroll
      for (int a = 0; a < D; ++a) { ap[(int64_t)e * D + a] = 0.f; av[(int64_t)e * D + a] = 0.f; }
      continue;
    }
    float vj[3], Jj[D * D];
#pragma unroll
    for (int d = 0; d < D; ++d) vj[d] = v[(int64_t)j * D + d];
#pragma unroll
    for (int a = 0; a < D * D; ++a) Jj[a] = J[(int64_t)j * D * D + a];
    float phi = limiter_phi<D>(lp, dx, hi, hj, Ji, Jj);
    float phi2 = phi / 2.f;
    float vdi[3], vdj[3], vdjU[3];
#pragma unroll
    for (int a = 0; a < D; ++a) {
      float cI = 0.f, cJ = 0.f;
#pragma unroll
      for (int gm = 0; gm < D; ++gm) { cI = cI + Ji[a * D + gm] * dx[gm]; cJ = cJ + Jj[a * D + gm] * dx[gm]; }
      vdi[a] = vi[a] - phi2 * cI;
      vdj[a] = vj[a] + phi2 * cJ;
      vdjU[a] = vj[a] - phi2 * cJ;
    }
    Corr<D> crj;
    crj.load(Ap, Bp, gAp, gBp, j);
    float gi[3], gjr[3], dxn[3];
#pragma unroll
    for (int a = 0; a < D; ++a) dxn[a] = -dx[a];
    grad_wr<D>(K, dx, r, hj, cri.A, cri.B, cri.gA, cri.gB, gi);
    grad_wr<D>(K, dxn, r, hi, crj.A, crj.B, crj.gA, crj.gB, gjr);
    float gw[3];
#pragma unroll
    for (int a = 0; a < D; ++a) gw[a] = 0.5f * (gi[a] - gjr[a]);
    float smj = hj / kscale;
    float ei[3], ej[3];
#pragma unroll
    for (int a = 0; a < D; ++a) { ei[a] = dx[a] / smi; ej[a] = dx[a] / smj; }
    float eii = dot<D>(ei, ei), ejj = dot<D>(ej, ej);
    float di = eii + 1.0e-7f * smi * smi, dj = ejj + 1.0e-7f * smj * smj;
    float ti = 0.f, tj = 0.f, tiU = 0.f, tjU = 0.f;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      float w = vdi[a] - vdj[a];
      float wU = vdi[a] - vdjU[a];
      ti = ti + w * ei[a];
      tj = tj + w * ej[a];
      tiU = tiU + wU * ei[a];
      tjU = tjU + wU * ej[a];
    }
    float mui = ti / di, muj = tj / dj, muiU = tiU / di, mujU = tjU / dj;
    mui = mui < 0.f ? mui : 0.f;
    muj = muj < 0.f ? muj : 0.f;
    muiU = muiU < 0.f ? muiU : 0.f;
    mujU = mujU < 0.f ? mujU : 0.f;
    float rhoj = rho[j], cj = c[j];
    float Qi = rhoi * (-Cl * ci * mui + Cq * mui * mui);
    float Qj = rhoj * (-Cl * cj * muj + Cq * muj * muj);
    float QiU = rhoi * (-Cl * ci * muiU + Cq * muiU * muiU);
    float QjU = rhoj * (-Cl * cj * mujU + Cq * mujU * mujU);
    if (alphaQ) {
      float wa = 0.5f * (ali + alpha[j]);
      Qi = Qi * wa; Qj = Qj * wa; QiU = QiU * wa; QjU = QjU * wa;
    }
    float Vj = V[j], Pj = P[j];
    float pP = -(Pi + Pj);
    float pQ = -(Qi + Qj);
#pragma unroll
    for (int a = 0; a < D; ++a) {
      float apv = pP * gw[a] * Vi * Vj / mi;
      float avv = pQ * Vj * gw[a] * Vi / mi;
      ap[(int64_t)e * D + a] = apv;
      av[(int64_t)e * D + a] = avv;
      acc[a].add(apv + avv);
    }
    float dt_ = 0.f;
#pragma unroll
    for (int a = 0; a < D; ++a) dt_ = dt_ + (vi[a] - vj[a]) * gw[a];
    float pT = Pj * dt_ * Vi * Vj / mi;
    float vT = 0.5f * (QiU + QjU) * dt_ * Vi * Vj / mi;
    accu.add(pT + vT);
  }
  int32_t i = perm[k];
  float ke = 0.f;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    dvdt_s[(int64_t)k * D + a] = acc[a].s;
    dvdt_u[(int64_t)i * D + a] = acc[a].s;
    ke = ke + vi[a] * acc[a].s;
  }
  float du = accu.s;
  dudt_u[i] = du;
  dEdt_u[i] = mi * ke + mi * du;
  div_u[i] = -drhodt_u[i] / rhoi;
}

// ------------------------------------------------------------ balance
template <int D>
__global__ void k_balance(Rows rows, int32_t n, const float* __restrict__ v, const float* __restrict__ dvdt,
                          const float* __restrict__ m, const float* __restrict__ rho, const float* __restrict__ P,
                          float gamma, float dt, float hdt, const float* __restrict__ ap,
                          const float* __restrict__ av, float* __restrict__ f) {
  int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  float vhi[3];
#pragma unroll
  for (int d = 0; d < D; ++d) vhi[d] = v[(int64_t)k * D + d] + hdt * dvdt[(int64_t)k * D + d];
  float si = P[k] / powf(rho[k], gamma);
  float mi = m[k];
  for (int32_t e = rows.off[k]; e < rows.off[k + 1]; ++e) {
    int32_t j = rows.nbr[e];
    float d_ = 0.f;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      float vhj = v[(int64_t)j * D + a] + hdt * dvdt[(int64_t)j * D + a];
      d_ = d_ + (vhj - vhi[a]) * (ap[(int64_t)e * D + a] + av[(int64_t)e * D + a]);
    }
    float dE = mi * d_ * dt;
    float sj = P[j] / powf(rho[j], gamma);
    float smin = fminf(fabsf(si), fabsf(sj));
    float smax = fmaxf(fabsf(si), fabsf(sj));
    bool mA = fabsf(si - sj) == 0.f;
    bool mB = (dE >= 0.f && si >= sj) || (dE < 0.f && si < sj);
    bool mC = (dE >= 0.f && si < sj) || (dE < 0.f && si >= sj);
    float ff = 0.5f;
    if (mA) ff = 0.5f;
    else if (mB) ff = smin / (smin + smax + 1.0e-14f);
    else if (mC) ff = smax / (smin + smax + 1.0e-14f);
    ff = fmaxf(0.f, fminf(1.f, ff));
    f[e] = ff;
  }
}

// ------------------------------------------------------------ export
template <int C>
__global__ void k_scatter_field(const float* __restrict__ src, const int32_t* __restrict__ perm, int32_t n,
                                float* __restrict__ dst) {
  int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int32_t i = perm[k];
#pragma unroll
  for (int c = 0; c < C; ++c) dst[(int64_t)i * C + c] = src[(int64_t)k * C + c];
}

static int require_bound(CrkContext* ctx, const CrkConfig* cfg) {
  if (!ctx || !cfg) { set_error("null context/config"); return CRK_EINVAL; }
  if (!ctx->bound) { set_error("no stage bound: call crk_stage_begin first"); return CRK_ESTATE; }
  if (cfg->domain.dim != ctx->dim) { set_error("config dimension %d differs from the bound stage (%d)", cfg->domain.dim, ctx->dim); return CRK_EINVAL; }
  return CRK_OK;
}

static int alloc_stage(CrkContext* ctx, int32_t n, int D) {
  size_t f = sizeof(float) * (size_t)n;
  Buf* s1[] = {&ctx->h, &ctx->m, &ctx->u, &ctx->a0, &ctx->divp, &ctx->V, &ctx->A, &ctx->rho, &ctx->P, &ctx->c,
               &ctx->divcd, &ctx->vsig, &ctx->Rcd, &ctx->du0, &ctx->alpha};
  for (Buf* b : s1) { int st = ensure(*b, f); if (st) return st; }
  Buf* sd[] = {&ctx->x, &ctx->v, &ctx->B, &ctx->gA, &ctx->dvdt, &ctx->vh};
  for (Buf* b : sd) { int st = ensure(*b, f * D); if (st) return st; }
  Buf* sdd[] = {&ctx->gB, &ctx->J};
  for (Buf* b : sdd) { int st = ensure(*b, f * D * D); if (st) return st; }
  return CRK_OK;
}

template <class T>
static T* F(Buf& b) { return (T*)b.p; }

static Rows rows_of(const CrkContext* ctx) { return Rows{ctx->adj.row_offsets, ctx->adj.nbr}; }

template <int D>
static DevKernel<D> devkernel(const CrkConfig* cfg) {
  DevKernel<D> K;
  K.kernel = cfg->kernel;
  K.C = (float)kernel_C(cfg->kernel, D);
  return K;
}

}  // namespace crk

using namespace crk;

#define CRK_DIM_DISPATCH(D, ...)                       \
  switch (D) {                                         \
    case 1: { constexpr int DD = 1; __VA_ARGS__; } break; \
    case 2: { constexpr int DD = 2; __VA_ARGS__; } break; \
    case 3: { constexpr int DD = 3; __VA_ARGS__; } break; \
    default: set_error("dim must be 1..3"); return CRK_EINVAL; \
  }

extern "C" {

static int validate_adj(const CrkAdjacency* adj, const CrkState* st) {
  CRK_REQUIRE(adj && st, "null adjacency/state");
  CRK_REQUIRE(adj->n > 0 && adj->n == st->n, "adjacency has %d particles, state %d", adj->n, st->n);
  CRK_REQUIRE(adj->perm && adj->row_offsets && (adj->n_edges == 0 || adj->nbr), "adjacency buffers missing");
  CRK_REQUIRE(st->positions && st->velocities && st->supports && st->masses && st->internalEnergies,
              "state buffers missing");
  return CRK_OK;
}

int crk_stage_begin(CrkContext* ctx, const CrkConfig* cfg, const CrkAdjacency* adj, const CrkState* st, void* stream) {
  CRK_REQUIRE(ctx && cfg, "crk_stage_begin: null context/config");
  int r = validate_adj(adj, st);
  if (r) return r;
  int D = cfg->domain.dim;
  CRK_REQUIRE(D >= 1 && D <= 3, "dim must be 1..3");
  CRK_REQUIRE(cfg->kernel >= 0 && cfg->kernel <= 2, "unknown kernel %d", cfg->kernel);
  if ((r = alloc_stage(ctx, adj->n, D))) return r;
  ctx->n = adj->n;
  ctx->dim = D;
  ctx->adj = *adj;
  ctx->st = *st;
  ctx->bound = true;
  ctx->have_factors = ctx->have_velgrad = ctx->have_accel = false;
  hipStream_t s = (hipStream_t)stream;
  int32_t n = adj->n;
  CRK_DIM_DISPATCH(D, hipLaunchKernelGGL(k_gather<DD>, dim3(nblocks(n)), dim3(kBlock), 0, s, adj->perm, n,
                                         (const float*)st->positions, (const float*)st->velocities,
                                         (const float*)st->supports, (const float*)st->masses,
                                         (const float*)st->internalEnergies, (const float*)st->alpha0s,
                                         (const float*)st->divergence, (const float*)st->alphas, F<float>(ctx->x),
                                         F<float>(ctx->v), F<float>(ctx->h), F<float>(ctx->m), F<float>(ctx->u),
                                         F<float>(ctx->a0), F<float>(ctx->divp), F<float>(ctx->alpha)));
  CRK_LAUNCH_CHECK("k_gather");
  return CRK_OK;
}

int crk_stage_rebind(CrkContext* ctx, const CrkConfig* cfg, const CrkAdjacency* adj, const CrkState* st, void* stream) {
  return crk_stage_begin(ctx, cfg, adj, st, stream);
}

int crk_owen_iteration(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st, const float* hminmax_dev,
                       float* max_ratio_dev, void* stream) {
  int r = require_bound(ctx, cfg);
  if (r) return r;
  CRK_REQUIRE(st && st->supports && hminmax_dev, "crk_owen_iteration: null argument");
  CRK_REQUIRE(ctx->nlut >= 2, "Owen table not set (crk_set_owen_lut)");
  hipStream_t s = (hipStream_t)stream;
  int32_t n = ctx->n;
  Geo geo = make_geo(*cfg);
  float C = (float)kernel_C(cfg->kernel, cfg->domain.dim);
  if (max_ratio_dev) { if ((r = fill_f32(max_ratio_dev, 1, 0.f, s))) return r; }
  CRK_DIM_DISPATCH(ctx->dim, hipLaunchKernelGGL(k_owen<DD>, dim3(nblocks(n)), dim3(kBlock), 0, s, rows_of(ctx), geo,
                                                cfg->kernel, C, n, F<float>(ctx->x), F<float>(ctx->h),
                                                ctx->adj.perm, st->supports, F<double>(ctx->lut_psiH),
                                                F<double>(ctx->lut_nh), ctx->nlut, (float)ctx->nh_target,
                                                hminmax_dev, max_ratio_dev));
  CRK_LAUNCH_CHECK("k_owen");
  return CRK_OK;
}

int crk_factors_eos(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st, int32_t* n_singular_dev, void* stream) {
  int r = require_bound(ctx, cfg);
  if (r) return r;
  CRK_REQUIRE(st && st->densities && st->entropies && st->pressures && st->soundspeeds,
              "crk_factors_eos: state output buffers missing");
  hipStream_t s = (hipStream_t)stream;
  int32_t n = ctx->n;
  Geo geo = make_geo(*cfg);
  CRK_DIM_DISPATCH(ctx->dim, {
    DevKernel<DD> K = devkernel<DD>(cfg);
    hipLaunchKernelGGL(k_volume<DD>, dim3(nblocks(n)), dim3(kBlock), 0, s, rows_of(ctx), geo, K, n, F<float>(ctx->x),
                       F<float>(ctx->h), F<float>(ctx->V));
    CRK_LAUNCH_CHECK("k_volume");
    hipLaunchKernelGGL(k_moments<DD>, dim3(nblocks(n)), dim3(kBlock), 0, s, rows_of(ctx), geo, K, n,
                       F<float>(ctx->x), F<float>(ctx->h), F<float>(ctx->V), F<float>(ctx->m), F<float>(ctx->u),
                       cfg->gamma, F<float>(ctx->A), F<float>(ctx->B), F<float>(ctx->gA), F<float>(ctx->gB),
                       F<float>(ctx->rho), F<float>(ctx->P), F<float>(ctx->c), ctx->adj.perm, st->densities,
                       st->entropies, st->pressures, st->soundspeeds, n_singular_dev);
    CRK_LAUNCH_CHECK("k_moments");
  });
  ctx->have_factors = true;
  return CRK_OK;
}

int crk_velocity_gradient(CrkContext* ctx, const CrkConfig* cfg, CrkUpdate* up, void* stream) {
  int r = require_bound(ctx, cfg);
  if (r) return r;
  CRK_REQUIRE(up && up->drhodt, "crk_velocity_gradient: update->drhodt missing");
  if (!ctx->have_factors) { set_error("crk_velocity_gradient before crk_factors_eos"); return CRK_ESTATE; }
  hipStream_t s = (hipStream_t)stream;
  int32_t n = ctx->n;
  Geo geo = make_geo(*cfg);
  bool cullen = cfg->switchParams.scheme == CRK_SWITCH_CULLEN_DEHNEN || cfg->switchParams.scheme == CRK_SWITCH_LEARNED;
  CRK_DIM_DISPATCH(ctx->dim, {
    DevKernel<DD> K = devkernel<DD>(cfg);
    if (cullen)
      hipLaunchKernelGGL((k_velgrad<DD, true>), dim3(nblocks(n)), dim3(kBlock), 0, s, rows_of(ctx), geo, K, n,
                         F<float>(ctx->x), F<float>(ctx->h), F<float>(ctx->v), F<float>(ctx->V), F<float>(ctx->m),
                         F<float>(ctx->rho), F<float>(ctx->c), F<float>(ctx->A), F<float>(ctx->B), F<float>(ctx->gA),
                         F<float>(ctx->gB), F<float>(ctx->J), F<float>(ctx->divcd), F<float>(ctx->vsig),
                         ctx->adj.perm, up->drhodt);
    else
      hipLaunchKernelGGL((k_velgrad<DD, false>), dim3(nblocks(n)), dim3(kBlock), 0, s, rows_of(ctx), geo, K, n,
                         F<float>(ctx->x), F<float>(ctx->h), F<float>(ctx->v), F<float>(ctx->V), F<float>(ctx->m),
                         F<float>(ctx->rho), F<float>(ctx->c), F<float>(ctx->A), F<float>(ctx->B), F<float>(ctx->gA),
                         F<float>(ctx->gB), F<float>(ctx->J), F<float>(ctx->divcd), F<float>(ctx->vsig),
                         ctx->adj.perm, up->drhodt);
  });
  CRK_LAUNCH_CHECK("k_velgrad");
  ctx->have_velgrad = true;
  return CRK_OK;
}

int crk_switch(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st, float dt, void* stream) {
  int r = require_bound(ctx, cfg);
  if (r) return r;
  CRK_REQUIRE(st && st->alphas, "crk_switch: state->alphas missing");
  CRK_REQUIRE(!cfg->evolveAlpha0 || st->alpha0s, "crk_switch: state->alpha0s missing");
  CRK_REQUIRE(dt > 0.f, "crk_switch: dt must be positive");
  if (!ctx->have_velgrad) { set_error("crk_switch before crk_velocity_gradient"); return CRK_ESTATE; }
  if (cfg->switchParams.scheme != CRK_SWITCH_CULLEN_DEHNEN) {
    set_error("crk_switch: switch scheme %d is not the Cullen-Dehnen switch", cfg->switchParams.scheme);
    return CRK_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  int32_t n = ctx->n;
  Geo geo = make_geo(*cfg);
  float fkern = (float)(1.0 / kernel_scale(cfg->kernel, cfg->domain.dim));
  const CrkSwitchParams& sp = cfg->switchParams;
  CRK_DIM_DISPATCH(ctx->dim, {
    DevKernel<DD> K = devkernel<DD>(cfg);
    hipLaunchKernelGGL(k_cullen<DD>, dim3(nblocks(n)), dim3(kBlock), 0, s, rows_of(ctx), geo, K, n, F<float>(ctx->x),
                       F<float>(ctx->h), F<float>(ctx->m), F<float>(ctx->rho), F<float>(ctx->divcd),
                       F<float>(ctx->vsig), F<float>(ctx->divp), F<float>(ctx->a0), F<float>(ctx->alpha),
                       F<float>(ctx->Rcd), dt, fkern, sp.alpha_min, sp.alpha_max, sp.beta_xi, sp.limitXi,
                       cfg->evolveAlpha0, ctx->adj.perm, st->alphas, st->alpha0s);
  });
  CRK_LAUNCH_CHECK("k_cullen");
  return CRK_OK;
}

int crk_accel_dudt(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st, CrkUpdate* up, CrkEdges* ed,
                   void* stream) {
  int r = require_bound(ctx, cfg);
  if (r) return r;
  CRK_REQUIRE(st && st->divergence, "crk_accel_dudt: state->divergence missing");
  CRK_REQUIRE(up && up->dvdt && up->dudt && up->dEdt && up->drhodt, "crk_accel_dudt: update buffers missing");
  CRK_REQUIRE(ed && (ctx->adj.n_edges == 0 || (ed->ap && ed->av)), "crk_accel_dudt: edge buffers missing");
  if (!ctx->have_velgrad) { set_error("crk_accel_dudt before crk_velocity_gradient"); return CRK_ESTATE; }
  hipStream_t s = (hipStream_t)stream;
  int32_t n = ctx->n;
  Geo geo = make_geo(*cfg);
  LimParams lp{cfg->crkViscosity.eta_fold, cfg->crkViscosity.eta_crit, cfg->crkViscosity.enableCRKLimiter,
               cfg->crkViscosity.enableVanLeerLimiter, cfg->crkViscosity.forceVanLeerOff,
               cfg->crkViscosity.forceVanLeerOn};
  float kscale = (float)kernel_scale(cfg->kernel, cfg->domain.dim);
  const float* alpha_s = F<float>(ctx->alpha);
  CRK_DIM_DISPATCH(ctx->dim, {
    DevKernel<DD> K = devkernel<DD>(cfg);
    hipLaunchKernelGGL(k_accel_dudt<DD>, dim3(nblocks(n)), dim3(kBlock), 0, s, rows_of(ctx), geo, K, lp, kscale,
                       cfg->diffusion.C_l, cfg->diffusion.C_q, cfg->alphaWeightedQ, n, F<float>(ctx->x),
                       F<float>(ctx->h), F<float>(ctx->v), F<float>(ctx->m), F<float>(ctx->V), F<float>(ctx->rho),
                       F<float>(ctx->P), F<float>(ctx->c), F<float>(ctx->J), F<float>(ctx->A), F<float>(ctx->B),
                       F<float>(ctx->gA), F<float>(ctx->gB), alpha_s, ed->ap, ed->av, F<float>(ctx->dvdt),
                       ctx->adj.perm, up->dvdt, up->dudt, up->dEdt, (const float*)up->drhodt, st->divergence);
  });
  CRK_LAUNCH_CHECK("k_accel_dudt");
  ctx->have_accel = true;
  return CRK_OK;
}

int crk_balance(CrkContext* ctx, const CrkConfig* cfg, float dt, CrkEdges* ed, void* stream) {
  int r = require_bound(ctx, cfg);
  if (r) return r;
  CRK_REQUIRE(ed && (ctx->adj.n_edges == 0 || (ed->ap && ed->av && ed->f)), "crk_balance: edge buffers missing");
  if (!ctx->have_accel) { set_error("crk_balance before crk_accel_dudt"); return CRK_ESTATE; }
  hipStream_t s = (hipStream_t)stream;
  int32_t n = ctx->n;
  float hdt = (float)(0.5 * (double)dt);
  CRK_DIM_DISPATCH(ctx->dim, hipLaunchKernelGGL(k_balance<DD>, dim3(nblocks(n)), dim3(kBlock), 0, s, rows_of(ctx), n,
                                                F<float>(ctx->v), F<float>(ctx->dvdt), F<float>(ctx->m),
                                                F<float>(ctx->rho), F<float>(ctx->P), cfg->gamma, dt, hdt, ed->ap,
                                                ed->av, ed->f));
  CRK_LAUNCH_CHECK("k_balance");
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
