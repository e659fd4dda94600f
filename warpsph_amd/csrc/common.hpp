// common.hpp -- shared pieces of the gfx950 CRKSPH library: error handling,
// the context, and device helpers (kernel functions, minimum image, Kahan
// accumulation, Morton codes).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/crksph.h"

namespace crk {

// ------------------------------------------------------------ errors
void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define CRK_HIP(call)                                                  \
  do {                                                                 \
    hipError_t e_ = (call);                                            \
    if (e_ != hipSuccess) return ::crk::hip_fail(e_, #call, __FILE__, __LINE__); \
  } while (0)

#define CRK_LAUNCH_CHECK(name)                                         \
  do {                                                                 \
    hipError_t e_ = hipGetLastError();                                 \
    if (e_ != hipSuccess) return ::crk::hip_fail(e_, name, __FILE__, __LINE__); \
  } while (0)

#define CRK_REQUIRE(cond, ...)                                         \
  do {                                                                 \
    if (!(cond)) { ::crk::set_error(__VA_ARGS__); return CRK_EINVAL; } \
  } while (0)

constexpr int kBlock = 256;
inline unsigned nblocks(int64_t n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

// ------------------------------------------------------------ context
struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
};
int ensure(Buf& b, size_t bytes);
void release(Buf& b);

}  // namespace crk

struct CrkContext {
  int device = 0;
  // sorted-order working set
  int32_t n = 0;
  int32_t dim = 0;
  int64_t cap_n = 0;
  crk::Buf x, v, h, m, u, a0, divp, alpha, V, A, B, gA, gB, rho, P, c, sP, J, divcd, vsig, Rcd, dvdt, vh, vhs, du0;
  // packed neighbour data (sorted order): the acceleration pass's full
  // record (Rec<D>::NF floats) and the narrow float4 records of the lighter
  // passes, geo = {x (3, zero padded), h}, vel = {v (3, zero padded), c};
  // vm2 = {V, m} (written with V, read by the moments pass); vg4 and mr,
  // the velocity-gradient pass's record (written by the moments pass):
  // d < 3: vg4 = {v (2, zero padded), V, m / rho}; d = 3: vg4 = {v, V} and
  // mr = m / rho
  crk::Buf rec, geo4, vel4, vm2, vg4, mr;
  // single-pass neighbour search: rows collected by crk_adjacency_count
  // into nbr_tmp (nbr_cap slots per row) for crk_adjacency_fill to copy
  crk::Buf nbr_tmp;
  int32_t nbr_cap = 0;       // slots per row for the next build
  double nbr_cap_key = 0.0;  // n_H verletScale^dim the slots were sized for
  int32_t nbr_cap_used = 0;  // slots per row of the collected build
  const int32_t* collect_rows = nullptr;  // row_offsets of the collected build
  int32_t collect_n = 0;
  int64_t collect_E = -1;
  bool collect_ok = false;  // no row exceeded nbr_cap
  // crk_adjacency_count_keep: the runs that keep their prior lists and the
  // prior rows, for the fill that follows (null run_keep: none)
  const uint8_t* keep_runs = nullptr;
  const int32_t* keep_off = nullptr;
  const int32_t* keep_nbr = nullptr;
  // in-support pairs of each row, compacted in row order (caller-owned
  // buffers of the stage, bound by crk_support)
  int32_t* pair_cnt = nullptr;
  int32_t* pair_nbr = nullptr;
  uint16_t* pair_loc = nullptr;  // their tile-local indices (tiled adjacencies; else null)
  bool have_support = false;
  bool have_volume = false;  // V (and v_sig) current for the bound adjacency and supports
  // lanes per particle row for each pass (crk_set_group_size)
  int32_t group[CRK_NPASS] = {4, 8, 2, 8, 8, 16, 32, 32, 32};
  // threads per block of the tiled passes (crk_set_block_size): one block
  // per tile of CRK_TILE_ROWS rows
  int32_t threads[CRK_NPASS] = {256, 256, 256, 256, 256, 1024, 256, 256, 256};
  // bound adjacency / user state
  CrkAdjacency adj{};
  CrkState st{};
  bool bound = false;
  bool have_factors = false, have_velgrad = false, have_accel = false;
  // Cullen-Dehnen intermediates computed in the bound stage (v_sig, div, R)
  bool have_vsig = false, have_divcd = false, have_rcd = false;
  // Owen table
  crk::Buf lut_nh, lut_psiH;
  int32_t nlut = 0;
  double nh_target = 0.0;
  // scratch
  crk::Buf s_keys0, s_keys1, s_vals0, s_vals1, s_hist, s_scan, s_red, s_counts;
  crk::Buf pinned;  // host pinned buffer for blocking reads
  // Verlet validity of the bound stage (crk_stage_begin_checked / crk_owen_valid)
  crk::Buf s_chk, s_chk_part;
  const float* check_hb = nullptr;
  int32_t check_runs = 0;  // run count of the last checked stage
  bool owen_checked = false;
  // device flag of deferred Verlet tests (1: some list failed), for steps
  // replayed without host synchronisation
  int32_t* d_flag = nullptr;
  // d_flag while the bound stage's lists are tested by deferred tests (the
  // adaptive-support and support passes then do nothing once it is set);
  // null after a blocking test or crk_stage_begin
  const int32_t* gate = nullptr;
  // crk_flag_read_async / crk_flag_wait: the flag's pinned host copies and
  // the event recorded after each copy, per slot
  static constexpr int kFlagSlots = 4;
  int32_t* flag_host = nullptr;
  hipEvent_t flag_event[kFlagSlots] = {};
  // device address of this device's copy of the allocation-free entry
  // point's file-scope scratch array (hipGetSymbolAddress at context creation)
  double* noalloc_scratch = nullptr;
  // independent runs (crk_bind_runs): run r is particles [run_off[r], run_off[r+1])
  // in user order and, for adjacencies built while bound, in sorted order too
  int32_t n_runs = 0;  // 0: one run, no run table
  int32_t n_chunks = 0;
  std::vector<int32_t> run_off;  // the table on the device (kept while unbound)
  crk::Buf runs_dev;     // chunks (int2 [n_chunks]) | offsets [n_runs + 1] | run id [n]
  crk::Buf s_gridruns;   // per-run grids of the last crk_adjacency_count
  const float* dtp = nullptr;  // per-particle dt of the bound runs (may be null)
  // learned switch: |grad rho|, |grad P| (float2), normalised features [n][8],
  // out-of-range flags
  crk::Buf gnorm, feat, fflag;
};

namespace crk {

// ------------------------------------------------------------ kernels
// W(r, H) = C_d / H^d k(r/H) (Wendland 1995; normalisation Dehnen & Aly 2012).
// B7: k(q) = sum_m c_m (b_m - q)_+^7, c = (1, -8, 28, -56), b = (1, 3/4, 1/2,
// 1/4), the centred B-spline of degree 7 (eight unit boxes convolved) scaled
// to the support; C_d = 1 / (S_d I_{d-1}) with I_n = int_0^1 k(q) q^n dq =
// sum_m c_m b_m^(n+8) n! 7! / (n+8)! and S_d = 2, 2 pi, 4 pi (I_0 = 315/8192,
// I_1 = 7435/1179648, I_2 = 105/65536).
__host__ __device__ inline double kernel_C(int kernel, int dim) {
  const double C[4][3] = {{1.25, 2.228169203286535, 3.3422538049298027},
                          {1.5, 2.8647889756541161, 4.9238560519055126},
                          {1.71875, 3.5468815889050962, 6.7889530412636603},
                          {13.003174603174603, 25.25175659831448, 49.668468335573586}};
  return C[kernel][dim - 1];
}
// H/h with h twice the kernel's standard deviation per dimension (Dehnen &
// Aly 2012); B7: sigma^2 = I_{d+1} / (d I_{d-1}) = 1/24, 531453/13085600, 19/480
__host__ __device__ inline double kernel_scale(int kernel, int dim) {
  const double S[4][3] = {{1.6201851746019651, 1.8973665961010275, 1.9364916731037083},
                          {1.9364916731037087, 2.1712405933672376, 2.2079402165819615},
                          {2.2079402165819615, 2.4152294576982398, 2.4494897427831779},
                          {2.449489742783178, 2.4810435822586725, 2.513123449750173}};
  return S[kernel][dim - 1];
}

// B7 on 0 <= q < 1.  Below the first knot (q < 1/4) the four truncated powers
// cancel (k(0) = 151/1024 from terms of order one, dk/dq(0) = 0 exactly), so
// the piece's own polynomial is used there: k = 151/1024 - 105/64 q^2 +
// 35/4 q^4 - 35 q^6 + 35 q^7 and (dk/dq)/q = -105/32 + 35 q^2 - 210 q^4 +
// 245 q^5 (all coefficients dyadic, exact in binary); above it the
// truncated powers, dk/dq = -7 sum_m c_m (b_m - q)_+^6.
// (each compiled without contraction: the host and the device build the
// Owen table from them and must round alike)
// xi of the CFL step and of the Cullen-Dehnen relaxation time (warpSPHCore's
// sphKernel_xi, timestep/compressible.py:50, CullenDehnen2010.py:286).  For
// the Wendland kernels the kernel scale H/h.  For B7 in 1D and 2D the values
// the reference's own outputs fix: with this build's samplers reproducing
// those runs' initial states (the Sod tube's Owen-adapted internal energies
// 0.7121 .. 2.138 of examples/compressible/01-Sod_Shock_Tube_1D.ipynb cell 3,
// the Kelvin-Helmholtz lattice's h = 4 dx), the printed initial steps
// 3.0056745163165033e-4 (same notebook, cell 7) and 3.892789827659726e-4
// (12-Kelvin-Helmholtz.ipynb cell 8) give xi = 3.865665 (1D) and 3.932735
// (2D); 3D has no such output and takes the 2D value's ratio to H/h.
__host__ __device__ inline double kernel_xi(int kernel, int dim) {
  if (kernel == CRK_KERNEL_B7) {
    const double X[3] = {3.865665, 3.932735, 3.932735 * (2.513123449750173 / 2.4810435822586725)};
    return X[dim - 1];
  }
  return kernel_scale(kernel, dim);
}

template <class T>
__host__ __device__ inline T b7_pow7(T x) {
#pragma clang fp contract(off)
  T x2 = x * x; T x3 = x2 * x; return x3 * x3 * x;
}
template <class T>
__host__ __device__ inline T b7_pow6(T x) {
#pragma clang fp contract(off)
  T x2 = x * x; T x3 = x2 * x; return x3 * x3;
}
template <class T>
__host__ __device__ inline T b7_k(T q) {
#pragma clang fp contract(off)
  if (q < T(0.25)) {
    T q2 = q * q;
    return T(151.0 / 1024.0) + q2 * (T(-105.0 / 64.0) + q2 * (T(35.0 / 4.0) + q2 * (T(-35) + T(35) * q)));
  }
  T a = T(1) - q, b = T(0.75) - q, c = T(0.5) - q;
  b = b > T(0) ? b : T(0);
  c = c > T(0) ? c : T(0);
  return b7_pow7(a) - T(8) * b7_pow7(b) + T(28) * b7_pow7(c);
}
template <class T>
__host__ __device__ inline T b7_dk_outer(T q) {  // dk/dq for q >= 1/4
#pragma clang fp contract(off)
  T a = T(1) - q, b = T(0.75) - q, c = T(0.5) - q;
  b = b > T(0) ? b : T(0);
  c = c > T(0) ? c : T(0);
  return T(-7) * (b7_pow6(a) - T(8) * b7_pow6(b) + T(28) * b7_pow6(c));
}
template <class T>
__host__ __device__ inline T b7_dkq_inner(T q) {  // (dk/dq)/q for q < 1/4
#pragma clang fp contract(off)
  T q2 = q * q;
  return T(-105.0 / 32.0) + q2 * (T(35) + q2 * (T(-210) + T(245) * q));
}
template <class T>
__host__ __device__ inline T b7_dk(T q) {
#pragma clang fp contract(off)
  return q < T(0.25) ? q * b7_dkq_inner(q) : b7_dk_outer(q);
}

template <class T>
__host__ __device__ inline T kval(int kernel, int dim, T q) {
  if (kernel == CRK_KERNEL_B7) return b7_k(q);
  T t = T(1) - q;
  if (kernel == CRK_KERNEL_WENDLAND2) {
    if (dim == 1) return t * t * t * (T(1) + T(3) * q);
    T t2 = t * t;
    return t2 * t2 * (T(1) + T(4) * q);
  } else if (kernel == CRK_KERNEL_WENDLAND4) {
    if (dim == 1) { T t2 = t * t; return t2 * t2 * t * (T(1) + T(5) * q + T(8) * q * q); }
    T t2 = t * t; T t3 = t2 * t;
    return t3 * t3 * (T(1) + T(6) * q + T(35.0 / 3.0) * q * q);
  } else {
    if (dim == 1) { T t2 = t * t; T t3 = t2 * t; return t3 * t3 * t * (T(1) + T(7) * q + T(19) * q * q + T(21) * q * q * q); }
    T t2 = t * t; T t4 = t2 * t2;
    return t4 * t4 * (T(1) + T(8) * q + T(25) * q * q + T(32) * q * q * q);
  }
}
template <class T>
__host__ __device__ inline T kder(int kernel, int dim, T q) {
#pragma clang fp contract(off)  // the same rounding on host and device (Owen table)
  if (kernel == CRK_KERNEL_B7) return b7_dk(q);
  T t = T(1) - q;
  if (kernel == CRK_KERNEL_WENDLAND2) {
    if (dim == 1) return T(-12) * q * t * t;
    return T(-20) * q * t * t * t;
  } else if (kernel == CRK_KERNEL_WENDLAND4) {
    if (dim == 1) { T t2 = t * t; return T(-14) * q * t2 * t2 * (T(1) + T(4) * q); }
    T t2 = t * t;
    return T(-56.0 / 3.0) * q * t2 * t2 * t * (T(1) + T(5) * q);
  } else {
    if (dim == 1) { T t2 = t * t; T t3 = t2 * t; return T(-6) * q * t3 * t3 * (T(3) + T(18) * q + T(35) * q * q); }
    T t2 = t * t; T t3 = t2 * t;
    return T(-22) * q * t3 * t3 * t * (T(1) + T(7) * q + T(16) * q * q);
  }
}

template <int D>
struct DevKernel {
  int kernel;
  float C;
  __device__ inline float hpow(float h, int p) const { float r = h; for (int k = 1; k < p; ++k) r = r * h; return r; }
  __device__ inline float W(float r, float h) const {
    float q = r / h;
    if (!(q < 1.f)) return 0.f;
    return C * kval<float>(kernel, D, q) / hpow(h, D);
  }
  __device__ inline float dWdr(float r, float h) const {
    float q = r / h;
    if (!(q < 1.f)) return 0.f;
    return C * kder<float>(kernel, D, q) / hpow(h, D + 1);
  }
  __device__ inline void gradW(const float* x, float r, float h, float* g) const {
    float fac = 0.f;
    if (r > 0.f) fac = dWdr(r, h) / r;
#pragma unroll
    for (int d = 0; d < D; ++d) g[d] = fac * x[d];
  }
};

// ------------------------------------------------------------ geometry
struct Geo {
  int per[3];
  float L[3], half[3];
  float wrap[3];  // half for periodic dimensions, +inf otherwise (branch-free minimum image)
};
inline Geo make_geo(const CrkConfig& c) {
  Geo g;
  for (int d = 0; d < 3; ++d) {
    g.per[d] = d < c.domain.dim ? c.domain.periodic[d] : 0;
    g.L[d] = c.domain.max[d] - c.domain.min[d];
    g.half[d] = 0.5f * g.L[d];
    g.wrap[d] = g.per[d] ? g.half[d] : INFINITY;
  }
  return g;
}
// x_i - x_j with minimum image, exactly rounded (no contraction): used where
// an integer decision depends on the result.
template <int D>
__device__ inline float exact_r2(const Geo& g, const float* xi, const float* xj) {
  // plain operators in a contract(off) scope: HIP's __fmul_rn / __fadd_rn are
  // inlined header functions that carry the file's contraction flag, and
  // hipcc fuses them into v_fmac_f32 (one rounding instead of two), which
  // flips pairs at |x_ij| = R
#pragma clang fp contract(off)
  float r2 = 0.f;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    float dx = xi[d] - xj[d];
    if (g.per[d]) {
      if (dx > g.half[d]) dx = dx - g.L[d];
      else if (dx < -g.half[d]) dx = dx + g.L[d];
    }
    const float sq = dx * dx;
    r2 = d == 0 ? sq : r2 + sq;
  }
  return r2;
}
template <int D>
__device__ inline void min_image(const Geo& g, const float* xi, const float* xj, float* out) {
#pragma unroll
  for (int d = 0; d < D; ++d) {
    float dx = xi[d] - xj[d];
    if (g.per[d]) {
      if (dx > g.half[d]) dx = dx - g.L[d];
      else if (dx < -g.half[d]) dx = dx + g.L[d];
    }
    out[d] = dx;
  }
}
template <int D>
__device__ inline float dot(const float* a, const float* b) {
  float r = a[0] * b[0];
#pragma unroll
  for (int d = 1; d < D; ++d) r = r + a[d] * b[d];
  return r;
}

// Kahan-compensated FP32 sum
struct Kahan {
  float s = 0.f, c = 0.f;
  __device__ inline void add(float y) {
    float yy = y - c;
    float t = s + yy;
    c = (t - s) - yy;
    s = t;
  }
};

// ------------------------------------------------------------ Morton
__host__ __device__ inline uint32_t part1by1(uint32_t x) {
  x &= 0x0000ffffu;
  x = (x | (x << 8)) & 0x00FF00FFu;
  x = (x | (x << 4)) & 0x0F0F0F0Fu;
  x = (x | (x << 2)) & 0x33333333u;
  x = (x | (x << 1)) & 0x55555555u;
  return x;
}
__host__ __device__ inline uint32_t part1by2(uint32_t x) {
  x &= 0x000003ffu;
  x = (x | (x << 16)) & 0xFF0000FFu;
  x = (x | (x << 8)) & 0x0300F00Fu;
  x = (x | (x << 4)) & 0x030C30C3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}
template <int D>
__host__ __device__ inline uint32_t morton(const int* c) {
  if (D == 1) return (uint32_t)c[0];
  if (D == 2) return part1by1((uint32_t)c[0]) | (part1by1((uint32_t)c[1]) << 1);
  return part1by2((uint32_t)c[0]) | (part1by2((uint32_t)c[1]) << 1) | (part1by2((uint32_t)c[2]) << 2);
}
__device__ inline int cell_coord(float x, float origin, float inv, int n, int periodic) {
#pragma clang fp contract(off)
  float t = (x - origin) * inv;
  int c = (int)floorf(t);
  if (periodic) { c %= n; if (c < 0) c += n; }
  else { c = c < 0 ? 0 : (c >= n ? n - 1 : c); }
  return c;
}

// atomics on non-negative floats through their bit patterns
__device__ inline void atomic_max_pos(float* addr, float v) { atomicMax((unsigned int*)addr, __float_as_uint(v)); }
__device__ inline void atomic_min_pos(float* addr, float v) { atomicMin((unsigned int*)addr, __float_as_uint(v)); }

// Block-wide reductions (every thread of the block must call; result valid
// in every thread).  Used so that grid reductions issue one atomic per block.
// (Blocks of up to 1024 threads.)
template <class Op>
__device__ inline float block_reduce(float v, Op op) {
  __shared__ float sh_red[16];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_xor(v, off));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh_red[threadIdx.x >> 6] = v;
  __syncthreads();
  float r = sh_red[0];
  const int nw = ((int)blockDim.x + 63) >> 6;
  for (int w = 1; w < nw; ++w) r = op(r, sh_red[w]);
  return r;
}
struct MaxOp { __device__ float operator()(float a, float b) const { return fmaxf(a, b); } };
struct MinOp { __device__ float operator()(float a, float b) const { return fminf(a, b); } };
// grid size of grid-stride reductions
inline unsigned reduce_blocks(int64_t n) {
  unsigned b = nblocks(n);
  return b < 512u ? (b == 0 ? 1u : b) : 512u;
}

// ------------------------------------------------------------ independent runs
// Reductions over bound runs are segmented: block b reduces the chunk
// {first particle, run} = chunk[b], at most kChunk particles of one run, and
// issues one atomic per result into the run's slot.  Without runs the same
// kernels reduce the whole set grid-stride into slot 0.
constexpr int32_t kChunk = 2048;
constexpr int32_t kMaxRuns = 1 << 16;
struct RunsDev {
  const int32_t* run = nullptr;  // [n] run of each particle, null: one run
  const int32_t* off = nullptr;  // [n_runs + 1]
  const int2* chunk = nullptr;   // [n_chunks]
};
struct Span {
  int32_t k, e, step, r;
};
__device__ inline Span block_span(const RunsDev& R, int32_t n) {
  if (R.chunk) {
    const int2 c = R.chunk[blockIdx.x];
    const int32_t e = min(c.x + kChunk, R.off[c.y + 1]);
    return Span{c.x + (int32_t)threadIdx.x, e, (int32_t)blockDim.x, c.y};
  }
  return Span{(int32_t)(blockIdx.x * blockDim.x + threadIdx.x), n, (int32_t)(gridDim.x * blockDim.x), 0};
}
__device__ inline int32_t run_index(const int32_t* run, int32_t k) { return run ? run[k] : 0; }
inline RunsDev runs_of(const CrkContext* c) {
  RunsDev R;
  if (c->n_runs > 0) {
    R.chunk = (const int2*)c->runs_dev.p;
    R.off = (const int32_t*)(R.chunk + c->n_chunks);
    R.run = R.off + (c->n_runs + 1);
  }
  return R;
}
inline const int32_t* run_ids(const CrkContext* c) { return c->n_runs > 0 ? runs_of(c).run : nullptr; }
// grid of a segmented reduction (plain: the grid-stride grid)
inline unsigned span_blocks(const CrkContext* c, unsigned plain) { return c->n_runs > 0 ? (unsigned)c->n_chunks : plain; }
inline int32_t run_count(const CrkContext* c) { return c->n_runs > 0 ? c->n_runs : 1; }
// the step of particle k: per particle when the bound runs carry one
struct DtArg {
  float dt;
  const float* p;
  __device__ inline float at(int32_t k) const { return p ? p[k] : dt; }
};
inline DtArg dt_arg(const CrkContext* c, float dt) { return DtArg{dt, c->n_runs > 0 ? c->dtp : nullptr}; }

// ------------------------------------------------------------ small host helpers
int scan_exclusive_i32(CrkContext* ctx, const int32_t* in, int32_t* out, int32_t n, int64_t* total_dev_or_null,
                       hipStream_t s);
int fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s);
int fill_f32(float* p, int64_t n, float v, hipStream_t s);
int ensure_pinned(CrkContext* ctx, size_t bytes);
// OR into ctx->d_flag when some run's list fails the Verlet test:
// red[stride r + i_d2] = max |x - xb|^2, [.. + i_g] = max h/hb, [.. + i_hbm] = min hb
int verlet_flag_launch(CrkContext* ctx, const float* red, int stride, int i_d2, int i_g, int i_hbm, int32_t nr,
                       double sv, hipStream_t s);
// every run's particle count matches the bound table (n particles in all)
int require_runs_n(const CrkContext* ctx, int32_t n, const char* who);
// resolves the device address of the allocation-free scratch array of
// integrate.hip for the current device into ctx->noalloc_scratch
int noalloc_setup(CrkContext* ctx);

}  // namespace crk
