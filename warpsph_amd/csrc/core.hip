// core.hip -- context, error reporting, scans and host utilities of the
// gfx950 CRKSPH library.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace crk {

static thread_local char g_err[1024] = {0};

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  set_error("HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
  return CRK_EHIP;
}

int ensure(Buf& b, size_t bytes) {
  if (bytes <= b.bytes && b.p) return CRK_OK;
  if (b.p) {
    CRK_HIP(hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
  }
  size_t alloc = bytes < 256 ? 256 : bytes;
  CRK_HIP(hipMalloc(&b.p, alloc));
  b.bytes = alloc;
  return CRK_OK;
}

void release(Buf& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

// ------------------------------------------------------------ fills
__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_f32(float* p, int64_t n, float v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
int fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s) {
  if (n <= 0) return CRK_OK;
  hipLaunchKernelGGL(k_fill_i32, dim3(nblocks(n)), dim3(kBlock), 0, s, p, n, v);
  CRK_LAUNCH_CHECK("k_fill_i32");
  return CRK_OK;
}
int fill_f32(float* p, int64_t n, float v, hipStream_t s) {
  if (n <= 0) return CRK_OK;
  hipLaunchKernelGGL(k_fill_f32, dim3(nblocks(n)), dim3(kBlock), 0, s, p, n, v);
  CRK_LAUNCH_CHECK("k_fill_f32");
  return CRK_OK;
}

// host pinned staging for blocking reads (and uploads of small tables);
// every user synchronises the stream before it returns, so it is idle here
int ensure_pinned(CrkContext* ctx, size_t bytes) {
  if (bytes <= ctx->pinned.bytes) return CRK_OK;
  void* p = nullptr;
  CRK_HIP(hipHostMalloc(&p, bytes, hipHostMallocDefault));
  if (ctx->pinned.p) CRK_HIP(hipHostFree(ctx->pinned.p));
  ctx->pinned.p = p;
  ctx->pinned.bytes = bytes;
  return CRK_OK;
}

int require_runs_n(const CrkContext* ctx, int32_t n, const char* who) {
  if (ctx->n_runs > 0 && ctx->run_off.back() != n) {
    set_error("%s: %d particles, but the bound runs hold %d", who, n, ctx->run_off.back());
    return CRK_EINVAL;
  }
  return CRK_OK;
}

__global__ void k_verlet_flag(const float* __restrict__ red, int stride, int i_d2, int i_g, int i_hbm, int32_t nr,
                              double sv, int32_t* flag) {
  for (int32_t r = threadIdx.x; r < nr; r += blockDim.x) {
    const double g = red[stride * r + i_g];
    const bool ok = g < sv && 2.0 * sqrt((double)red[stride * r + i_d2]) <= (sv - g) * (double)red[stride * r + i_hbm];
    if (!ok) atomicOr(flag, 1);
  }
}
int verlet_flag_launch(CrkContext* ctx, const float* red, int stride, int i_d2, int i_g, int i_hbm, int32_t nr,
                       double sv, hipStream_t s) {
  hipLaunchKernelGGL(k_verlet_flag, dim3(1), dim3(kBlock), 0, s, red, stride, i_d2, i_g, i_hbm, nr, sv, ctx->d_flag);
  CRK_LAUNCH_CHECK("k_verlet_flag");
  return CRK_OK;
}
__global__ void k_flag_set(int32_t* flag, int32_t v) { *flag = v; }
__global__ void k_copy_unless(const int32_t* __restrict__ flag, const uint32_t* __restrict__ src,
                              uint32_t* __restrict__ dst, int64_t n) {
  if (*flag) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// run id of every particle, one block per chunk
__global__ void k_run_ids(const int2* __restrict__ chunk, const int32_t* __restrict__ off, int32_t* __restrict__ run) {
  const int2 c = chunk[blockIdx.x];
  const int32_t e = min(c.x + kChunk, off[c.y + 1]);
  for (int32_t k = c.x + (int32_t)threadIdx.x; k < e; k += blockDim.x) run[k] = c.y;
}

// ------------------------------------------------------------ exclusive scan
// Tiles of 2048 (256 threads x 8 items) -> per-tile exclusive scan and tile
// sums; one block scans the tile sums (int64, overflow-checked); a last pass
// adds the tile offsets.
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

__global__ void k_scan_tiles(const int32_t* in, int32_t* out, int32_t n, int64_t* tile_sums) {
  __shared__ int32_t sh[kBlock];
  int t = threadIdx.x;
  int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)t * kScanItems;
  int32_t vals[kScanItems];
  int32_t sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int64_t idx = base + k;
    int32_t v = idx < n ? in[idx] : 0;
    vals[k] = sum;
    sum += v;
  }
  sh[t] = sum;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {
    int32_t x = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  int32_t excl = sh[t] - sum;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int64_t idx = base + k;
    if (idx < n) out[idx] = vals[k] + excl;
  }
  if (t == kBlock - 1) tile_sums[blockIdx.x] = (int64_t)sh[kBlock - 1];
}

__global__ void k_scan_single(int64_t* a, int32_t m, int64_t* total) {
  __shared__ int64_t sh[kBlock];
  int t = threadIdx.x;
  int32_t chunk = (m + kBlock - 1) / kBlock;
  int32_t lo = t * chunk, hi = min(m, lo + chunk);
  int64_t sum = 0;
  for (int32_t i = lo; i < hi; ++i) sum += a[i];
  sh[t] = sum;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {
    int64_t x = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  int64_t run = sh[t] - sum;
  for (int32_t i = lo; i < hi; ++i) {
    int64_t v = a[i];
    a[i] = run;
    run += v;
  }
  if (t == kBlock - 1 && total) *total = sh[kBlock - 1];
}

__global__ void k_scan_add(int32_t* out, int32_t n, const int64_t* tile_off) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (int32_t)((int64_t)out[i] + tile_off[i / kScanTile]);
}

int scan_exclusive_i32(CrkContext* ctx, const int32_t* in, int32_t* out, int32_t n, int64_t* total_dev,
                       hipStream_t s) {
  if (n <= 0) return CRK_OK;
  int32_t ntiles = (int32_t)((n + kScanTile - 1) / kScanTile);
  int st = ensure(ctx->s_scan, sizeof(int64_t) * (size_t)(ntiles + 1));
  if (st) return st;
  int64_t* tiles = (int64_t*)ctx->s_scan.p;
  hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(kBlock), 0, s, in, out, n, tiles);
  CRK_LAUNCH_CHECK("k_scan_tiles");
  hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kBlock), 0, s, tiles, ntiles, total_dev);
  CRK_LAUNCH_CHECK("k_scan_single");
  hipLaunchKernelGGL(k_scan_add, dim3(nblocks(n)), dim3(kBlock), 0, s, out, n, tiles);
  CRK_LAUNCH_CHECK("k_scan_add");
  return CRK_OK;
}

// ------------------------------------------------------------ Owen table
// computePsi (src/warpSPH/modules/adaptiveSupport/wp_psi.py:8-99), psiH
// column: psiH(n_h) = (sum_{eta = 1/n_h, 2/n_h, ... < 1} |k'(eta)| C_d * ring(eta))^(1/d)
// with ring = 2 (1D), 2 pi eta n_h (2D), 4 pi eta^2 n_h^2 (3D); FP64.
__host__ __device__ inline void owen_entry(int kernel, int dim, int nlut, double nmin, double nmax, int i, double* nh,
                                            double* psiH) {
#pragma clang fp contract(off)  // host and device round every operation alike
  const double C = kernel_C(kernel, dim);
  const double dn = (nmax - nmin) / ((double)nlut - 1.0);
  const double n_h = nmin + (double)i * dn;
  const double deta = 1.0 / n_h;
  double etar = deta, result = 0.0;
  while (etar < 1.0) {
    const double corr =
        dim == 1 ? 2.0 : (dim == 2 ? 2.0 * M_PI * etar / deta : 4.0 * M_PI * (etar * etar) / (deta * deta));
    const double val = fabs(kder<double>(kernel, dim, etar) * C / 1.0);
    result += val * corr;
    etar += deta;
  }
  nh[i] = n_h;
  psiH[i] = pow(result, 1.0 / (double)dim);
}
static void owen_table(int kernel, int dim, int nlut, double nmin, double nmax, double* nh, double* psiH) {
  for (int i = 0; i < nlut; ++i) owen_entry(kernel, dim, nlut, nmin, nmax, i, nh, psiH);
}
// the same table computed on the device, one entry per thread (FP64; pow may
// differ from the host's in the last bit)
__global__ void k_owen_table(int kernel, int dim, int nlut, double nmin, double nmax, double* nh, double* psiH) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nlut) owen_entry(kernel, dim, nlut, nmin, nmax, i, nh, psiH);
}

}  // namespace crk

using namespace crk;

extern "C" {

const char* crk_version(void) { return "crksph-hip 1 gfx950"; }

int crk_struct_sizes(int64_t* out, int32_t n) {
  CRK_REQUIRE(out && n >= 0, "crk_struct_sizes: invalid arguments");
  const int64_t sz[] = {sizeof(CrkDomain), sizeof(CrkDiffusionParams), sizeof(CrkViscosityParams),
                        sizeof(CrkSwitchParams), sizeof(CrkConfig), sizeof(CrkGrid), sizeof(CrkAdjacency),
                        sizeof(CrkState), sizeof(CrkUpdate), sizeof(CrkEdges), sizeof(CrkMlp)};
  for (int i = 0; i < n && i < (int)(sizeof(sz) / sizeof(sz[0])); ++i) out[i] = sz[i];
  return CRK_OK;
}

int crk_last_error(char* buf, size_t n) {
  if (!buf || n == 0) return CRK_EINVAL;
  std::strncpy(buf, g_err, n - 1);
  buf[n - 1] = 0;
  return CRK_OK;
}

int crk_ctx_create(CrkContext** out) {
  if (!out) { set_error("crk_ctx_create: null output"); return CRK_EINVAL; }
  CrkContext* c = new CrkContext();
  hipError_t e = hipGetDevice(&c->device);
  if (e != hipSuccess) { delete c; return hip_fail(e, "hipGetDevice", __FILE__, __LINE__); }
  void* p = nullptr;
  e = hipHostMalloc(&p, 4096, hipHostMallocDefault);
  if (e != hipSuccess) { delete c; return hip_fail(e, "hipHostMalloc", __FILE__, __LINE__); }
  c->pinned.p = p;
  c->pinned.bytes = 4096;
  e = hipMalloc((void**)&c->d_flag, 64);
  if (e != hipSuccess) { (void)hipHostFree(p); delete c; return hip_fail(e, "hipMalloc", __FILE__, __LINE__); }
  e = hipMemset(c->d_flag, 0, 64);
  if (e != hipSuccess) { (void)hipFree(c->d_flag); (void)hipHostFree(p); delete c; return hip_fail(e, "hipMemset", __FILE__, __LINE__); }
  void* fh = nullptr;
  e = hipHostMalloc(&fh, 64, hipHostMallocDefault);
  for (int q = 0; q < CrkContext::kFlagSlots && e == hipSuccess; ++q)
    e = hipEventCreateWithFlags(&c->flag_event[q], hipEventDisableTiming);
  if (e != hipSuccess || noalloc_setup(c) != CRK_OK) {
    if (e != hipSuccess) (void)hip_fail(e, "flag host copy / event", __FILE__, __LINE__);
    for (hipEvent_t ev : c->flag_event)
      if (ev) (void)hipEventDestroy(ev);
    if (fh) (void)hipHostFree(fh);
    (void)hipFree(c->d_flag);
    (void)hipHostFree(p);
    delete c;
    return CRK_EHIP;
  }
  c->flag_host = (int32_t*)fh;
  *out = c;
  return CRK_OK;
}

int crk_bind_runs(CrkContext* ctx, int32_t n_runs, const int32_t* offsets, const float* dt_particles,
                  void* stream) {
  CRK_REQUIRE(ctx, "crk_bind_runs: null context");
  CRK_REQUIRE(n_runs >= 0 && n_runs <= kMaxRuns, "crk_bind_runs: n_runs must be in [0, %d] (got %d)", kMaxRuns,
              n_runs);
  if (n_runs == 0) {
    ctx->n_runs = 0;  // the device table stays for a later bind of the same runs
    ctx->dtp = nullptr;
    return CRK_OK;
  }
  CRK_REQUIRE(offsets && offsets[0] == 0, "crk_bind_runs: offsets must start at 0");
  for (int32_t r = 0; r < n_runs; ++r)
    CRK_REQUIRE(offsets[r + 1] > offsets[r], "crk_bind_runs: run %d is empty or offsets decrease", r);
  const bool same = (int32_t)ctx->run_off.size() == n_runs + 1 &&
                    std::equal(ctx->run_off.begin(), ctx->run_off.end(), offsets);
  ctx->dtp = dt_particles;
  if (same) {  // no upload, no synchronisation (capturable)
    ctx->n_runs = n_runs;
    return CRK_OK;
  }
  const int32_t n = offsets[n_runs];
  std::vector<int32_t> table;  // chunks (pairs) | offsets
  for (int32_t r = 0; r < n_runs; ++r)
    for (int32_t k = offsets[r]; k < offsets[r + 1]; k += kChunk) {
      table.push_back(k);
      table.push_back(r);
    }
  const int32_t n_chunks = (int32_t)(table.size() / 2);
  table.insert(table.end(), offsets, offsets + n_runs + 1);
  int st;
  if ((st = ensure(ctx->runs_dev, sizeof(int32_t) * (table.size() + (size_t)n)))) return st;
  if ((st = ensure_pinned(ctx, std::max<size_t>(sizeof(int32_t) * table.size(), 16 * sizeof(float) * n_runs))))
    return st;
  hipStream_t s = (hipStream_t)stream;
  std::memcpy(ctx->pinned.p, table.data(), sizeof(int32_t) * table.size());
  CRK_HIP(hipMemcpyAsync(ctx->runs_dev.p, ctx->pinned.p, sizeof(int32_t) * table.size(), hipMemcpyHostToDevice, s));
  ctx->n_runs = n_runs;
  ctx->n_chunks = n_chunks;
  ctx->run_off.assign(offsets, offsets + n_runs + 1);
  RunsDev R = runs_of(ctx);
  hipLaunchKernelGGL(k_run_ids, dim3(n_chunks), dim3(kBlock), 0, s, R.chunk, R.off, (int32_t*)R.run);
  CRK_LAUNCH_CHECK("k_run_ids");
  CRK_HIP(hipStreamSynchronize(s));
  return CRK_OK;
}

int crk_flag_clear(CrkContext* ctx, void* stream) {
  CRK_REQUIRE(ctx, "crk_flag_clear: null context");
  hipLaunchKernelGGL(k_flag_set, dim3(1), dim3(1), 0, (hipStream_t)stream, ctx->d_flag, 0);
  CRK_LAUNCH_CHECK("k_flag_set");
  return CRK_OK;
}

int crk_flag_read(CrkContext* ctx, int32_t* flag, void* stream) {
  CRK_REQUIRE(ctx && flag, "crk_flag_read: null argument");
  hipStream_t s = (hipStream_t)stream;
  int32_t* h = (int32_t*)ctx->pinned.p;
  CRK_HIP(hipMemcpyAsync(h, ctx->d_flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  CRK_HIP(hipStreamSynchronize(s));
  *flag = *h;
  return CRK_OK;
}

int crk_flag_read_async(CrkContext* ctx, int32_t slot, void* stream) {
  CRK_REQUIRE(ctx, "crk_flag_read_async: null context");
  CRK_REQUIRE(slot >= 0 && slot < CrkContext::kFlagSlots, "crk_flag_read_async: slot %d not in [0, %d)", slot,
              CrkContext::kFlagSlots);
  hipStream_t s = (hipStream_t)stream;
  CRK_HIP(hipMemcpyAsync(ctx->flag_host + slot, ctx->d_flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  CRK_HIP(hipEventRecord(ctx->flag_event[slot], s));
  return CRK_OK;
}

int crk_flag_wait(CrkContext* ctx, int32_t slot, int32_t* flag) {
  CRK_REQUIRE(ctx && flag, "crk_flag_wait: null argument");
  CRK_REQUIRE(slot >= 0 && slot < CrkContext::kFlagSlots, "crk_flag_wait: slot %d not in [0, %d)", slot,
              CrkContext::kFlagSlots);
  CRK_HIP(hipEventSynchronize(ctx->flag_event[slot]));
  *flag = ((volatile int32_t*)ctx->flag_host)[slot];
  return CRK_OK;
}

int crk_copy_unless_flag(CrkContext* ctx, const void* src, void* dst, int64_t bytes, void* stream) {
  CRK_REQUIRE(ctx && src && dst && bytes >= 0 && bytes % 4 == 0, "crk_copy_unless_flag: invalid arguments");
  if (bytes == 0) return CRK_OK;
  const int64_t n = bytes / 4;
  unsigned nb = nblocks(n);
  nb = nb > 2048u ? 2048u : nb;
  hipLaunchKernelGGL(k_copy_unless, dim3(nb), dim3(kBlock), 0, (hipStream_t)stream, ctx->d_flag,
                     (const uint32_t*)src, (uint32_t*)dst, n);
  CRK_LAUNCH_CHECK("k_copy_unless");
  return CRK_OK;
}

int crk_run_ids(CrkContext* ctx, int32_t n, int32_t* run_out, void* stream) {
  CRK_REQUIRE(ctx && run_out, "crk_run_ids: null argument");
  CRK_REQUIRE(ctx->n_runs > 0, "crk_run_ids: no runs bound");
  int st = require_runs_n(ctx, n, "crk_run_ids");
  if (st) return st;
  CRK_HIP(hipMemcpyAsync(run_out, run_ids(ctx), sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToDevice,
                         (hipStream_t)stream));
  return CRK_OK;
}

int crk_ctx_destroy(CrkContext* c) {
  if (!c) return CRK_OK;
  (void)hipDeviceSynchronize();
  Buf* bufs[] = {&c->x, &c->v, &c->h, &c->m, &c->u, &c->a0, &c->divp, &c->alpha, &c->V, &c->A, &c->B, &c->gA, &c->gB,
                 &c->rho, &c->P, &c->c, &c->sP, &c->vhs, &c->J, &c->divcd, &c->vsig, &c->Rcd, &c->dvdt, &c->vh, &c->du0,
                 &c->rec, &c->geo4, &c->vel4, &c->vm2, &c->vg4, &c->mr,
                 &c->lut_nh, &c->lut_psiH, &c->s_keys0, &c->s_keys1, &c->s_vals0, &c->s_vals1, &c->s_hist,
                 &c->s_scan, &c->s_red, &c->s_counts, &c->s_chk, &c->s_chk_part, &c->runs_dev, &c->s_gridruns, &c->gnorm, &c->feat, &c->fflag};
  for (Buf* b : bufs) release(*b);
  if (c->pinned.p) (void)hipHostFree(c->pinned.p);
  if (c->d_flag) (void)hipFree(c->d_flag);
  for (hipEvent_t ev : c->flag_event)
    if (ev) (void)hipEventDestroy(ev);
  if (c->flag_host) (void)hipHostFree(c->flag_host);
  delete c;
  return CRK_OK;
}

int crk_kernel_constants(int32_t kernel, int32_t dim, double* C_d, double* kernel_scale_out) {
  CRK_REQUIRE(kernel >= 0 && kernel <= 3, "unknown kernel %d", kernel);
  CRK_REQUIRE(dim >= 1 && dim <= 3, "dim must be 1..3 (got %d)", dim);
  CRK_REQUIRE(C_d && kernel_scale_out, "null output");
  *C_d = kernel_C(kernel, dim);
  *kernel_scale_out = kernel_scale(kernel, dim);
  return CRK_OK;
}

int crk_kernel_xi(int32_t kernel, int32_t dim, double* xi) {
  CRK_REQUIRE(kernel >= 0 && kernel <= 3, "unknown kernel %d", kernel);
  CRK_REQUIRE(dim >= 1 && dim <= 3, "dim must be 1..3");
  CRK_REQUIRE(xi, "null output");
  *xi = kernel_xi(kernel, dim);
  return CRK_OK;
}

int crk_owen_lut(int32_t kernel, int32_t dim, int32_t nlut, double n_min, double n_max, double* n_h_out,
                 double* psiH_out) {
  CRK_REQUIRE(kernel >= 0 && kernel <= 3, "unknown kernel %d", kernel);
  CRK_REQUIRE(dim >= 1 && dim <= 3, "dim must be 1..3 (got %d)", dim);
  CRK_REQUIRE(nlut >= 2 && n_max > n_min && n_min > 0, "invalid table range");
  CRK_REQUIRE(n_h_out && psiH_out, "null output");
  owen_table(kernel, dim, nlut, n_min, n_max, n_h_out, psiH_out);
  return CRK_OK;
}

int crk_owen_lut_device(CrkContext* ctx, int32_t kernel, int32_t dim, int32_t nlut, double n_min, double n_max,
                        double nh_target, void* stream) {
  CRK_REQUIRE(ctx, "crk_owen_lut_device: null context");
  CRK_REQUIRE(kernel >= 0 && kernel <= 3, "unknown kernel %d", kernel);
  CRK_REQUIRE(dim >= 1 && dim <= 3, "dim must be 1..3 (got %d)", dim);
  CRK_REQUIRE(nlut >= 2 && n_max > n_min && n_min > 0, "invalid table range");
  int st = ensure(ctx->lut_nh, sizeof(double) * nlut);
  if (st) return st;
  if ((st = ensure(ctx->lut_psiH, sizeof(double) * nlut))) return st;
  hipLaunchKernelGGL(k_owen_table, dim3(nblocks(nlut)), dim3(kBlock), 0, (hipStream_t)stream, kernel, dim, nlut,
                     n_min, n_max, (double*)ctx->lut_nh.p, (double*)ctx->lut_psiH.p);
  CRK_LAUNCH_CHECK("k_owen_table");
  ctx->nlut = nlut;
  ctx->nh_target = nh_target;
  return CRK_OK;
}

int crk_owen_lut_read(CrkContext* ctx, double* n_h_out, double* psiH_out, void* stream) {
  CRK_REQUIRE(ctx && n_h_out && psiH_out && ctx->nlut >= 2, "crk_owen_lut_read: no table");
  hipStream_t s = (hipStream_t)stream;
  CRK_HIP(hipMemcpyAsync(n_h_out, ctx->lut_nh.p, sizeof(double) * ctx->nlut, hipMemcpyDeviceToHost, s));
  CRK_HIP(hipMemcpyAsync(psiH_out, ctx->lut_psiH.p, sizeof(double) * ctx->nlut, hipMemcpyDeviceToHost, s));
  CRK_HIP(hipStreamSynchronize(s));
  return CRK_OK;
}

int crk_set_owen_lut(CrkContext* ctx, int32_t nlut, const double* n_h, const double* psiH, double nh_target,
                     void* stream) {
  CRK_REQUIRE(ctx && n_h && psiH && nlut >= 2, "crk_set_owen_lut: invalid arguments");
  for (int i = 1; i < nlut; ++i)
    CRK_REQUIRE(psiH[i] > psiH[i - 1], "Owen table must be strictly increasing in psiH (index %d)", i);
  int st = ensure(ctx->lut_nh, sizeof(double) * nlut);
  if (st) return st;
  st = ensure(ctx->lut_psiH, sizeof(double) * nlut);
  if (st) return st;
  hipStream_t s = (hipStream_t)stream;
  CRK_HIP(hipMemcpyAsync(ctx->lut_nh.p, n_h, sizeof(double) * nlut, hipMemcpyHostToDevice, s));
  CRK_HIP(hipMemcpyAsync(ctx->lut_psiH.p, psiH, sizeof(double) * nlut, hipMemcpyHostToDevice, s));
  CRK_HIP(hipStreamSynchronize(s));
  ctx->nlut = nlut;
  ctx->nh_target = nh_target;
  return CRK_OK;
}

// Grid for a search radius R = verletScale * h_max: ncell = floor(extent / R')
// with R' = R (1 + 1e-4) + 1e-6 extent, so every cell is wider than R even
// after FP32 rounding of the key computation; capped so that the Morton table
// has at most 2^24 entries.
int crk_grid_plan(const CrkConfig* cfg, float h_max, const float* lo, const float* hi, CrkGrid* g) {
  CRK_REQUIRE(cfg && g, "crk_grid_plan: null argument");
  int D = cfg->domain.dim;
  CRK_REQUIRE(D >= 1 && D <= 3, "dim must be 1..3 (got %d)", D);
  CRK_REQUIRE(h_max > 0.f && std::isfinite(h_max), "supports must be positive and finite (max %g)", (double)h_max);
  std::memset(g, 0, sizeof(*g));
  g->dim = D;
  double R = (double)cfg->verletScale * (double)h_max;
  int cap = D == 1 ? (1 << 24) : (D == 2 ? 4096 : 256);
  int maxn = 1;
  for (int d = 0; d < D; ++d) {
    double ext, o;
    if (cfg->domain.periodic[d]) {
      ext = (double)cfg->domain.max[d] - (double)cfg->domain.min[d];
      o = cfg->domain.min[d];
      CRK_REQUIRE(R <= 0.5 * ext,
                  "search radius %g exceeds half the periodic extent %g in dimension %d", R, ext, d);
    } else {
      CRK_REQUIRE(lo && hi, "non-periodic dimension needs the particle bounding box");
      ext = (double)hi[d] - (double)lo[d];
      o = lo[d];
    }
    double Reff = R * (1.0 + 1e-4) + 1e-6 * ext;
    long n = ext > 0 ? (long)std::floor(ext / Reff) : 1;
    if (n < 1) n = 1;
    if (n > cap) n = cap;
    g->ncell[d] = (int)n;
    g->origin[d] = (float)o;
    g->inv_cell[d] = ext > 0 ? (float)((double)n / ext) : 0.f;
    if (n > maxn) maxn = (int)n;
  }
  int bits = 0;
  while ((1 << bits) < maxn) ++bits;
  g->bits = bits;
  g->table_size = (int64_t)1 << (D * bits);
  for (int d = D; d < 3; ++d) { g->ncell[d] = 1; g->origin[d] = 0.f; g->inv_cell[d] = 0.f; }
  return CRK_OK;
}

int crk_grid_plan_runs(const CrkConfig* cfg, int32_t n_runs, const float* extents, CrkGrid* grids,
                       int64_t* table_total) {
  CRK_REQUIRE(cfg && extents && grids && n_runs > 0, "crk_grid_plan_runs: invalid arguments");
  int64_t total = 0;
  for (int32_t r = 0; r < n_runs; ++r) {
    const float* e = extents + 7 * (int64_t)r;
    int st = crk_grid_plan(cfg, e[0], e + 1, e + 4, grids + r);
    if (st) return st;
    total += grids[r].table_size;
  }
  if (table_total) *table_total = total;
  return CRK_OK;
}

}  // extern "C"
