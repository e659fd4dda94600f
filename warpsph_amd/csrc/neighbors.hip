// neighbors.hip -- adjacency build of the CRKSPH step (replaces
// warpSPHCore.buildVerletList, called at src/warpSPH/schemes/crkSPH.py:54-58
// and src/warpSPH/modules/adaptiveSupport/optimalSupportOwen.py:74).
//
//   1. Morton key of each particle's cell (k_keys)
//   2. stable LSD radix sort of (key, index), 8-bit digits: LDS histograms per
//      4096-key tile, a device-wide exclusive scan, and a scatter that ranks
//      keys inside each wave64 with ballot match masks (k_rs_*)
//   3. cell start/end table over the Morton code space (k_cells)
//   4. CSR rows in sorted order: every j with |x_ij| < s * max(h_i, h_j),
//      found by scanning the 3^d neighbouring cells (k_count / k_fill).
// All integer decisions use exactly rounded FP32 arithmetic so that the CPU
// oracle reproduces them bit for bit.
#include "tiles.hpp"

namespace crk {

constexpr int kRsItems = 16;
constexpr int kRsTile = kBlock * kRsItems;  // 4096 keys per tile

struct CellScan {
  int cnt[3];
  int off[3][3];
};

static CellScan make_scan(const CrkGrid& g, const CrkConfig& cfg) {
  CellScan s;
  for (int d = 0; d < 3; ++d) {
    if (d >= g.dim) { s.cnt[d] = 1; s.off[d][0] = 0; continue; }
    int n = g.ncell[d];
    if (cfg.domain.periodic[d]) {
      if (n >= 3) { s.off[d][0] = -1; s.off[d][1] = 0; s.off[d][2] = 1; s.cnt[d] = 3; }
      else if (n == 2) { s.off[d][0] = 0; s.off[d][1] = 1; s.cnt[d] = 2; }
      else { s.off[d][0] = 0; s.cnt[d] = 1; }
    } else {
      s.off[d][0] = -1; s.off[d][1] = 0; s.off[d][2] = 1; s.cnt[d] = 3;
    }
  }
  return s;
}

struct GridDev {
  float origin[3], inv[3];
  int ncell[3];
  int per[3];
};
static GridDev make_griddev(const CrkGrid& g, const CrkConfig& cfg) {
  GridDev d;
  for (int k = 0; k < 3; ++k) {
    d.origin[k] = g.origin[k];
    d.inv[k] = g.inv_cell[k];
    d.ncell[k] = g.ncell[k] < 1 ? 1 : g.ncell[k];
    d.per[k] = k < g.dim ? cfg.domain.periodic[k] : 0;
  }
  return d;
}

// Independent runs in one adjacency (an ensemble): every run has its own
// grid (its own support range), and run r's cells occupy the key range
// [base, base + table size) after the cells of runs 0..r-1, so that the sort
// groups particles by run and then by cell and no run sees another's cells.
// The runs are contiguous in user order, so each keeps its index range in
// sorted order and one run-id array serves both orders.
struct GridRun {
  GridDev g;
  CellScan sc;
  uint32_t base;
};

template <int D>
__device__ inline void cell_of(const GridDev& g, const float* x, int* c) {
#pragma unroll
  for (int d = 0; d < 3; ++d) c[d] = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) c[d] = cell_coord(x[d], g.origin[d], g.inv[d], g.ncell[d], g.per[d]);
}

template <int D>
__global__ void k_keys(const float* __restrict__ x, int32_t n, GridDev g, const GridRun* __restrict__ runs,
                       const int32_t* __restrict__ run_of, uint32_t* __restrict__ keys, int32_t* __restrict__ vals) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float xi[3];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = x[(int64_t)i * D + d];
  int c[3];
  if (runs) {
    const GridRun& gr = runs[run_of[i]];
    cell_of<D>(gr.g, xi, c);
    keys[i] = gr.base + morton<D>(c);
  } else {
    cell_of<D>(g, xi, c);
    keys[i] = morton<D>(c);
  }
  vals[i] = i;
}

__global__ void k_rs_hist(const uint32_t* __restrict__ keys, int32_t n, int shift, int32_t* __restrict__ hist,
                          int32_t ntiles) {
  __shared__ int32_t cnt[256];
  int t = threadIdx.x;
  cnt[t] = 0;
  __syncthreads();
  int64_t base = (int64_t)blockIdx.x * kRsTile;
  for (int it = 0; it < kRsItems; ++it) {
    int64_t idx = base + it * kBlock + t;
    if (idx < n) atomicAdd(&cnt[(keys[idx] >> shift) & 255u], 1);
  }
  __syncthreads();
  hist[(int64_t)t * ntiles + blockIdx.x] = cnt[t];
}

__global__ void k_rs_scatter(const uint32_t* __restrict__ kin, const int32_t* __restrict__ vin,
                             uint32_t* __restrict__ kout, int32_t* __restrict__ vout, int32_t n, int shift,
                             const int32_t* __restrict__ hist, int32_t ntiles) {
  __shared__ int32_t running[256];
  __shared__ int32_t wcnt[kBlock / 64][256];
  const int t = threadIdx.x;
  const int w = t >> 6;
  const int lane = t & 63;
  running[t] = hist[(int64_t)t * ntiles + blockIdx.x];
#pragma unroll
  for (int q = 0; q < kBlock / 64; ++q) wcnt[q][t] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kRsTile;
  const unsigned long long lt = (1ull << lane) - 1ull;
  for (int it = 0; it < kRsItems; ++it) {
    int64_t idx = base + it * kBlock + t;
    bool valid = idx < n;
    uint32_t key = valid ? kin[idx] : 0u;
    int32_t val = valid ? vin[idx] : 0;
    uint32_t dig = (key >> shift) & 255u;
    unsigned long long peers = __ballot(valid ? 1 : 0);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      int bit = (dig >> b) & 1u;
      unsigned long long bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    int rank = __popcll(peers & lt);
    if (valid && rank == 0) wcnt[w][dig] = __popcll(peers);
    __syncthreads();
    if (valid) {
      int32_t pos = running[dig];
      for (int q = 0; q < w; ++q) pos += wcnt[q][dig];
      pos += rank;
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
    int32_t add = 0;
#pragma unroll
    for (int q = 0; q < kBlock / 64; ++q) { add += wcnt[q][t]; wcnt[q][t] = 0; }
    running[t] += add;
    __syncthreads();
  }
}

template <int D>
__global__ void k_post_sort(const int32_t* __restrict__ perm, const float* __restrict__ x, const float* __restrict__ h,
                            int32_t n, int32_t* __restrict__ rank, float* __restrict__ xb, float* __restrict__ hb) {
  int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int32_t i = perm[k];
  rank[i] = k;
#pragma unroll
  for (int d = 0; d < D; ++d) xb[(int64_t)k * D + d] = x[(int64_t)i * D + d];
  hb[k] = h[i];
}

__global__ void k_cells(const uint32_t* __restrict__ keys, int32_t n, int32_t* __restrict__ cs, int32_t* __restrict__ ce) {
  int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  uint32_t kk = keys[k];
  if (k == 0 || keys[k - 1] != kk) cs[kk] = k;
  if (k == n - 1 || keys[k + 1] != kk) ce[kk] = k + 1;
}

// Neighbour rows: for sorted particle k, every particle of the 3^D cells
// around its cell with |x_k - x_j| < s max(h_k, h_j) (exact arithmetic),
// cells in scan order and candidates ascending.
//   MODE 2 (crk_adjacency_count): counts each row and writes it to its own
//          cap slots of a scratch array (a row longer than cap sets
//          *overflow);
//   MODE 0: counts only (when the scratch would be too large);
//   MODE 1 (crk_adjacency_fill after an overflow or a count-only pass, or
//          for rows counted elsewhere): fills the rows at their offsets by
//          testing the candidates again.
// G lanes per row (G >= 3^D): lane t of the group first looks up cell t's
// particle range (all cells of the row in one round of loads instead of one
// dependent load per cell); then the lanes test consecutive candidates of
// each cell in turn, and the hits are compacted with a group ballot, so the
// row is written with coalesced stores in the same order as the
// one-thread-per-row scan (cells in scan order, candidates ascending).
// Runs of a bound ensemble that keep their prior lists through a rebuild
// (crk_adjacency_count_keep): run_keep[r] != 0, with the prior row offsets
// and entries (a kept run's sorted indices are unchanged).
struct KeepDev {
  const uint8_t* run_keep;
  const int32_t* off;
  const int32_t* nbr;
};

// A kept run's particles take the keys of the prior build, shifted by the
// change of the run's key base, so that the stable sort reproduces the
// prior order.
__global__ void k_keys_keep(int32_t n, const int32_t* __restrict__ run, const uint8_t* __restrict__ keep,
                            const int64_t* __restrict__ delta, const int32_t* __restrict__ prior_rank,
                            const uint32_t* __restrict__ prior_keys, uint32_t* __restrict__ k0) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t r = run[i];
  if (!keep[r]) return;
  k0[i] = (uint32_t)((int64_t)prior_keys[prior_rank[i]] + delta[r]);
}

// ... and keep the positions and supports their lists were built at
template <int D>
__global__ void k_keep_restore(int32_t n, const int32_t* __restrict__ run, const uint8_t* __restrict__ keep,
                               const float* __restrict__ prior_xb, const float* __restrict__ prior_hb,
                               float* __restrict__ xb, float* __restrict__ hb) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || !keep[run[k]]) return;
#pragma unroll
  for (int d = 0; d < D; ++d) xb[(int64_t)k * D + d] = prior_xb[(int64_t)k * D + d];
  hb[k] = prior_hb[k];
}

template <int D, int G, int MODE, bool RUNS>
__global__ void __launch_bounds__(kBlock) k_neighbors_g(GridDev g0, CellScan sc0, Geo geo, float s,
                                                        const float* __restrict__ xb, const float* __restrict__ hb,
                                                        const int32_t* __restrict__ cs, const int32_t* __restrict__ ce,
                                                        int32_t n, int32_t* __restrict__ counts_or_offsets,
                                                        int32_t* __restrict__ nbr, const GridRun* __restrict__ runs,
                                                        const int32_t* __restrict__ run_sorted, int32_t cap,
                                                        int32_t* __restrict__ overflow, KeepDev kp) {
  constexpr bool FILL = MODE == 1;
  static_assert(G >= (D == 1 ? 3 : (D == 2 ? 9 : 27)), "a group needs a lane per cell of the scan");
  static_assert(MODE >= 0 && MODE <= 2, "count (0), fill (1) or single pass (2)");
  const Lane L = lane_of<G>(n);
  if (!L.valid) return;  // whole groups only: every lane of a group shares its row
  const int32_t k = L.k;
  if (RUNS && kp.run_keep && kp.run_keep[run_sorted[k]]) {
    // a kept run's row: its prior length, or its prior entries (fill)
    const int32_t ob = kp.off[k], len = kp.off[k + 1] - ob;
    if (MODE == 1) {
      const int32_t b = counts_or_offsets[k];
      for (int32_t q = L.g; q < len; q += G) nbr[b + q] = kp.nbr[ob + q];
    } else if (L.g == 0) {
      counts_or_offsets[k] = len;
    }
    return;
  }
  float xi[3];
#pragma unroll
  for (int d = 0; d < D; ++d) xi[d] = xb[(int64_t)k * D + d];
  const float hi = hb[k];
  GridDev g = g0;
  CellScan sc = sc0;
  uint32_t base = 0;
  if (RUNS) {  // the run's own grid and key range
    const GridRun& gr = runs[run_sorted[k]];
    g = gr.g;
    sc = gr.sc;
    base = gr.base;
  }
  int c[3];
  cell_of<D>(g, xi, c);
  const int lane = threadIdx.x & 63;
  const int gbase = lane & ~(G - 1);
  const unsigned long long gmask = G == 64 ? ~0ull : (((1ull << G) - 1ull) << gbase);
  const unsigned long long lt = (1ull << lane) - 1ull;
  // this lane's cell of the scan: t = ox + cnt0 (oy + cnt1 oz)
  const int ncell = sc.cnt[0] * sc.cnt[1] * sc.cnt[2];
  int32_t cb0 = 0, ce0 = 0;
  if (L.g < ncell) {
    const int t = L.g;
    const int ox = t % sc.cnt[0], oy = (t / sc.cnt[0]) % sc.cnt[1], oz = t / (sc.cnt[0] * sc.cnt[1]);
    const int o[3] = {sc.off[0][ox], sc.off[1][oy], sc.off[2][oz]};
    int cc[3] = {0, 0, 0};
    bool ok = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      int q = c[d] + o[d];
      int nn = g.ncell[d];
      if (g.per[d]) { q %= nn; if (q < 0) q += nn; }
      else if (q < 0 || q >= nn) ok = false;
      cc[d] = q;
    }
    if (ok) {
      const uint32_t code = base + morton<D>(cc);
      cb0 = cs[code];
      ce0 = ce[code];
    }
  }
  int32_t run = 0;
  const int64_t out = FILL ? (int64_t)counts_or_offsets[k] : (MODE == 2 ? (int64_t)k * cap : 0);
  // the steps (cell t, candidates cb .. cb + G - 1) in scan order; each
  // step's candidate data is loaded one step ahead (t, cb and e are uniform
  // over the group, so are the loop trip counts)
  int t = 0;
  int32_t cb = __shfl(cb0, gbase), e = __shfl(ce0, gbase);
  while (cb >= e && ++t < ncell) { cb = __shfl(cb0, gbase + t); e = __shfl(ce0, gbase + t); }
  int32_t nk = cb + L.g;
  bool nval = t < ncell && nk < e;
  float nx[3] = {0.f, 0.f, 0.f}, nh = 0.f;
  if (nval) {
#pragma unroll
    for (int d = 0; d < D; ++d) nx[d] = xb[(int64_t)nk * D + d];
    nh = hb[nk];
  }
  while (t < ncell) {
    const int32_t kk = nk;
    const bool val = nval;
    float xj[3];
#pragma unroll
    for (int d = 0; d < D; ++d) xj[d] = nx[d];
    const float hj = nh;
    // the next step
    cb += G;
    while (cb >= e && ++t < ncell) { cb = __shfl(cb0, gbase + t); e = __shfl(ce0, gbase + t); }
    nk = cb + L.g;
    nval = t < ncell && nk < e;
    if (nval) {
#pragma unroll
      for (int d = 0; d < D; ++d) nx[d] = xb[(int64_t)nk * D + d];
      nh = hb[nk];
    }
    bool hit = false;
    if (val) {
      const float r2 = exact_r2<D>(geo, xi, xj);
      const float R = __fmul_rn(s, fmaxf(hi, hj));
      hit = r2 < __fmul_rn(R, R);
    }
    const unsigned long long m = __ballot(hit ? 1 : 0) & gmask;
    if (FILL && hit) nbr[out + run + __popcll(m & lt)] = kk;
    if (MODE == 2 && hit) {
      const int32_t q = run + __popcll(m & lt);
      if (q < cap) nbr[out + q] = kk;
    }
    run += __popcll(m);
  }
  if (MODE != 1 && L.g == 0) {
    counts_or_offsets[k] = run;
    if (MODE == 2 && run > cap) atomicOr(overflow, 1);
  }
}

// largest scratch of the single-pass search (rows x slots x 4 bytes)
constexpr double kCollectMaxBytes = 16.0 * (1ull << 30);

// crk_adjacency_fill after a single-pass count: row k's collected entries
// (cap slots per row) to its CSR offsets; G lanes per row
template <int G>
__global__ void __launch_bounds__(kBlock) k_rows_copy(const int32_t* __restrict__ tmp, int32_t cap,
                                                      const int32_t* __restrict__ off, int32_t n,
                                                      int32_t* __restrict__ nbr, const int32_t* __restrict__ run,
                                                      KeepDev kp) {
  const Lane L = lane_of<G>(n);
  if (!L.valid) return;
  const int32_t k = L.k;
  const int32_t b = off[k], e = off[k + 1];
  // (a kept run's row comes from the prior list)
  const int32_t* src = kp.run_keep && kp.run_keep[run[k]] ? kp.nbr + kp.off[k] : tmp + (int64_t)k * cap;
  for (int32_t q = b + L.g; q < e; q += G) nbr[q] = src[q - b];
}

__global__ void k_set_total(const int64_t* __restrict__ total, int32_t* __restrict__ dst) { *dst = (int32_t)*total; }
// lanes per row of the neighbour scan (candidates per row: ~9, ~180, ~830 in 1D/2D/3D)
template <int D>
constexpr int nbr_group() { return D == 1 ? 4 : (D == 2 ? 16 : 32); }

// validity of each run r: red[3r] = max |x - xb|^2, red[3r+1] = max h/hb,
// red[3r+2] = min hb (segmented grid-stride, one atomic per block)
template <int D>
__global__ void k_valid(Geo geo, const int32_t* __restrict__ perm, const float* __restrict__ xb,
                        const float* __restrict__ hb, const float* __restrict__ x, const float* __restrict__ h,
                        int32_t n, float* red, RunsDev R) {
  float d2 = 0.f, g = 0.f, hm = 3.0e38f;
  const Span sp = block_span(R, n);
  for (int32_t k = sp.k; k < sp.e; k += sp.step) {
    int32_t i = perm[k];
    float xi[3];
#pragma unroll
    for (int d = 0; d < D; ++d) xi[d] = x[(int64_t)i * D + d];
    d2 = fmaxf(d2, exact_r2<D>(geo, xi, &xb[(int64_t)k * D]));
    g = fmaxf(g, __fdiv_rn(h[i], hb[k]));
    hm = fminf(hm, hb[k]);
  }
  d2 = block_reduce(d2, MaxOp());
  g = block_reduce(g, MaxOp());
  hm = block_reduce(hm, MinOp());
  if (threadIdx.x == 0) {
    atomic_max_pos(&red[3 * sp.r], d2);
    atomic_max_pos(&red[3 * sp.r + 1], g);
    atomic_min_pos(&red[3 * sp.r + 2], hm);
  }
}

__global__ void k_valid_init(float* red, int32_t nr) {
  for (int32_t r = threadIdx.x; r < nr; r += blockDim.x) {
    red[3 * r] = 0.f;
    red[3 * r + 1] = 0.f;
    red[3 * r + 2] = 3.0e38f;
  }
}

__global__ void k_export(const int32_t* __restrict__ perm, const int32_t* __restrict__ off,
                         const int32_t* __restrict__ nbr, int32_t n, int64_t* __restrict__ iout,
                         int64_t* __restrict__ jout, int64_t* __restrict__ numn, int64_t* __restrict__ noff) {
  int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int32_t i = perm[k];
  int32_t b = off[k], e = off[k + 1];
  if (numn) numn[i] = e - b;
  if (noff) noff[i] = b;
  for (int32_t q = b; q < e; ++q) {
    if (iout) iout[q] = i;
    if (jout) jout[q] = perm[nbr[q]];
  }
}

// per-run grids of the bound runs on the device (ctx->s_gridruns); returns
// the total cell-table size
static int upload_gridruns(CrkContext* ctx, const CrkConfig* cfg, const CrkGrid* grids, int64_t* table,
                           hipStream_t s) {
  const int32_t R = ctx->n_runs;
  int st;
  if ((st = ensure(ctx->s_gridruns, sizeof(GridRun) * (size_t)R))) return st;
  if ((st = ensure_pinned(ctx, sizeof(GridRun) * (size_t)R))) return st;
  GridRun* h = (GridRun*)ctx->pinned.p;
  int64_t base = 0;
  for (int32_t r = 0; r < R; ++r) {
    h[r].g = make_griddev(grids[r], *cfg);
    h[r].sc = make_scan(grids[r], *cfg);
    h[r].base = (uint32_t)base;
    base += grids[r].table_size;
  }
  CRK_HIP(hipMemcpyAsync(ctx->s_gridruns.p, h, sizeof(GridRun) * (size_t)R, hipMemcpyHostToDevice, s));
  CRK_HIP(hipStreamSynchronize(s));
  *table = base;
  return CRK_OK;
}

static int key_bits(int64_t table) {
  int b = 0;
  while (((int64_t)1 << b) < table) ++b;
  return b;
}

// the kept runs of crk_adjacency_count_keep and the prior build they keep
struct KeepIn {
  const uint8_t* keep;
  const int64_t* delta;
  const int32_t* prior_rank;
  const uint32_t* prior_keys;
  const float* prior_xb;
  const float* prior_hb;
  const int32_t* prior_off;
  const int32_t* prior_nbr;
};

template <int D>
static int count_impl(CrkContext* ctx, const CrkConfig* cfg, const CrkGrid* grid, int32_t n, const float* x,
                      const float* h, uint32_t* keys, int32_t* perm, int32_t* rank, int32_t* cell_start,
                      int32_t* cell_end, int32_t* row_offsets, float* xb, float* hb, int64_t* n_edges,
                      hipStream_t s, const KeepIn* kin = nullptr) {
  GridDev g = make_griddev(*grid, *cfg);
  const bool runs = ctx->n_runs > 0;
  int64_t table = grid->table_size;
  int st;
  if (runs && (st = upload_gridruns(ctx, cfg, grid, &table, s))) return st;
  const GridRun* gr = runs ? (const GridRun*)ctx->s_gridruns.p : nullptr;
  const int32_t* run = run_ids(ctx);
  // keys and the identity permutation
  if ((st = ensure(ctx->s_keys0, sizeof(uint32_t) * (size_t)n))) return st;
  if ((st = ensure(ctx->s_keys1, sizeof(uint32_t) * (size_t)n))) return st;
  if ((st = ensure(ctx->s_vals0, sizeof(int32_t) * (size_t)n))) return st;
  if ((st = ensure(ctx->s_vals1, sizeof(int32_t) * (size_t)n))) return st;
  uint32_t* k0 = (uint32_t*)ctx->s_keys0.p;
  uint32_t* k1 = (uint32_t*)ctx->s_keys1.p;
  int32_t* v0 = (int32_t*)ctx->s_vals0.p;
  int32_t* v1 = (int32_t*)ctx->s_vals1.p;
  hipLaunchKernelGGL(k_keys<D>, dim3(nblocks(n)), dim3(kBlock), 0, s, x, n, g, gr, run, k0, v0);
  CRK_LAUNCH_CHECK("k_keys");
  if (kin) {
    hipLaunchKernelGGL(k_keys_keep, dim3(nblocks(n)), dim3(kBlock), 0, s, n, run, kin->keep, kin->delta,
                       kin->prior_rank, kin->prior_keys, k0);
    CRK_LAUNCH_CHECK("k_keys_keep");
  }
  int keybits = runs ? key_bits(table) : D * grid->bits;
  int npass = (keybits + 7) / 8;
  int32_t ntiles = (int32_t)((n + kRsTile - 1) / kRsTile);
  if ((st = ensure(ctx->s_hist, sizeof(int32_t) * 256 * (size_t)ntiles))) return st;
  int32_t* hist = (int32_t*)ctx->s_hist.p;
  for (int p = 0; p < npass; ++p) {
    int shift = 8 * p;
    hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(kBlock), 0, s, k0, n, shift, hist, ntiles);
    CRK_LAUNCH_CHECK("k_rs_hist");
    if ((st = scan_exclusive_i32(ctx, hist, hist, 256 * ntiles, nullptr, s))) return st;
    hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(kBlock), 0, s, k0, v0, k1, v1, n, shift, hist, ntiles);
    CRK_LAUNCH_CHECK("k_rs_scatter");
    std::swap(k0, k1);
    std::swap(v0, v1);
  }
  CRK_HIP(hipMemcpyAsync(keys, k0, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice, s));
  CRK_HIP(hipMemcpyAsync(perm, v0, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_post_sort<D>, dim3(nblocks(n)), dim3(kBlock), 0, s, perm, x, h, n, rank, xb, hb);
  CRK_LAUNCH_CHECK("k_post_sort");
  if (kin) {
    hipLaunchKernelGGL(k_keep_restore<D>, dim3(nblocks(n)), dim3(kBlock), 0, s, n, run, kin->keep, kin->prior_xb,
                       kin->prior_hb, xb, hb);
    CRK_LAUNCH_CHECK("k_keep_restore");
  }
  const KeepDev kp{kin ? kin->keep : nullptr, kin ? kin->prior_off : nullptr, kin ? kin->prior_nbr : nullptr};
  ctx->keep_runs = kp.run_keep;
  ctx->keep_off = kp.off;
  ctx->keep_nbr = kp.nbr;
  if ((st = fill_i32(cell_start, table, 0, s))) return st;
  if ((st = fill_i32(cell_end, table, 0, s))) return st;
  hipLaunchKernelGGL(k_cells, dim3(nblocks(n)), dim3(kBlock), 0, s, keys, n, cell_start, cell_end);
  CRK_LAUNCH_CHECK("k_cells");
  CellScan sc = make_scan(*grid, *cfg);
  Geo geo = make_geo(*cfg);
  constexpr int GN = nbr_group<D>();
  // single pass: rows collected into cap slots each (cap from the expected
  // row length, grown after an overflow; the overflowing build falls back to
  // the fill pass)
  {
    double sd = 1.0;
    for (int d = 0; d < D; ++d) sd *= (double)cfg->verletScale;
    const double key = (double)cfg->targetNeighbors * sd;  // expected entries per row
    if (ctx->nbr_cap <= 0 || key != ctx->nbr_cap_key) {
      ctx->nbr_cap = (int32_t)std::min(1024.0, std::max(32.0, std::ceil(2.5 * key / 32.0) * 32.0));
      ctx->nbr_cap_key = key;
    }
  }
  const int32_t cap = ctx->nbr_cap;
  // the scratch is bounded (kCollectMaxBytes): larger builds count only and
  // test the candidates again in the fill
  const bool collect = (double)n * (double)cap * sizeof(int32_t) <= kCollectMaxBytes;
  if (collect && (st = ensure(ctx->nbr_tmp, sizeof(int32_t) * (size_t)n * (size_t)cap))) return st;
  if ((st = ensure(ctx->s_red, 64))) return st;
  int64_t* total_dev = (int64_t*)ctx->s_red.p;
  int32_t* overflow_dev = (int32_t*)((char*)ctx->s_red.p + 8);
  if ((st = fill_i32(overflow_dev, 1, collect ? 0 : 1, s))) return st;
  int32_t* tmp = collect ? (int32_t*)ctx->nbr_tmp.p : nullptr;
  if (collect && runs)
    hipLaunchKernelGGL((k_neighbors_g<D, GN, 2, true>), dim3(group_blocks(n, GN)), dim3(kBlock), 0, s, g, sc, geo,
                       cfg->verletScale, xb, hb, cell_start, cell_end, n, row_offsets, tmp, gr, run, cap,
                       overflow_dev, kp);
  else if (collect)
    hipLaunchKernelGGL((k_neighbors_g<D, GN, 2, false>), dim3(group_blocks(n, GN)), dim3(kBlock), 0, s, g, sc, geo,
                       cfg->verletScale, xb, hb, cell_start, cell_end, n, row_offsets, tmp, gr, run, cap,
                       overflow_dev, kp);
  else if (runs)
    hipLaunchKernelGGL((k_neighbors_g<D, GN, 0, true>), dim3(group_blocks(n, GN)), dim3(kBlock), 0, s, g, sc, geo,
                       cfg->verletScale, xb, hb, cell_start, cell_end, n, row_offsets, tmp, gr, run, 0,
                       overflow_dev, kp);
  else
    hipLaunchKernelGGL((k_neighbors_g<D, GN, 0, false>), dim3(group_blocks(n, GN)), dim3(kBlock), 0, s, g, sc, geo,
                       cfg->verletScale, xb, hb, cell_start, cell_end, n, row_offsets, tmp, gr, run, 0,
                       overflow_dev, kp);
  CRK_LAUNCH_CHECK("k_neighbors(count)");
  if ((st = scan_exclusive_i32(ctx, row_offsets, row_offsets, n, total_dev, s))) return st;
  hipLaunchKernelGGL(k_set_total, dim3(1), dim3(1), 0, s, total_dev, row_offsets + n);
  CRK_LAUNCH_CHECK("k_set_total");
  if ((st = ensure_pinned(ctx, 16))) return st;
  int64_t* host = (int64_t*)ctx->pinned.p;
  CRK_HIP(hipMemcpyAsync(host, total_dev, 12, hipMemcpyDeviceToHost, s));  // total, overflow flag
  CRK_HIP(hipStreamSynchronize(s));
  const int64_t E = host[0];
  const int32_t overflow = ((const int32_t*)host)[2];
  if (E > (int64_t)INT32_MAX) {
    set_error("adjacency has %lld edges; more than 2^31-1 is not supported", (long long)E);
    return CRK_ECAPACITY;
  }
  ctx->collect_rows = row_offsets;
  ctx->collect_n = n;
  ctx->collect_E = E;
  ctx->collect_ok = overflow == 0;
  ctx->nbr_cap_used = cap;
  if (overflow && collect) ctx->nbr_cap = std::min<int32_t>(4096, cap + cap / 2 + 32);
  *n_edges = E;
  return CRK_OK;
}

// ------------------------------------------------------------ LDS tiles
// One block per tile: the distinct cells reached by the tile's entries (an
// LDS hash of their keys, inserted where the key changes along a thread's
// contiguous chunk of entries), sorted by their first sorted index and
// merged into runs; then each entry's local index j - start(cell) +
// offset(cell).
constexpr int kTileHash = 1024;   // hash slots (a power of two)
constexpr int kTileCells = 512;   // distinct cells a tile may reach
constexpr uint32_t kNoKey = 0xffffffffu;

__device__ inline int tile_hash(uint32_t key) { return (int)((key * 2654435761u) >> 22); }  // 10 bits

__global__ void __launch_bounds__(kBlock) k_tiles(int32_t n, const int32_t* __restrict__ off,
                                                  const int32_t* __restrict__ nbr, const uint32_t* __restrict__ keys,
                                                  const int32_t* __restrict__ cs, const int32_t* __restrict__ ce,
                                                  int32_t cap, int32_t* __restrict__ tiles,
                                                  uint16_t* __restrict__ loc) {
  __shared__ uint32_t hkey[kTileHash];
  __shared__ int32_t hadd[kTileHash];  // local index = j + hadd[slot of j's cell]
  __shared__ int32_t cslot[kTileCells], cstart[kTileCells], clen[kTileCells];
  __shared__ int32_t sslot[kTileCells], sstart[kTileCells], slen[kTileCells];
  __shared__ int32_t tab[kTileWords];
  __shared__ int32_t ncell, overflow;
  const int32_t t = blockIdx.x;
  const int32_t r0 = t * kTileRows, r1 = min(n, r0 + kTileRows);
  const int64_t e0 = off[r0], e1 = off[r1];
  for (int w = threadIdx.x; w < kTileHash; w += blockDim.x) hkey[w] = kNoKey;
  for (int w = threadIdx.x; w < kTileWords; w += blockDim.x) tab[w] = 0;
  if (threadIdx.x == 0) { ncell = 0; overflow = 0; }
  __syncthreads();
  // entries in coalesced strides, four per thread in flight; a lane inserts
  // its key only where it differs from the previous lane's (the entries of a
  // row run through a few cells in order)
  constexpr int kBatch = 4;
  for (int64_t q0 = e0; q0 < e1; q0 += kBatch * (int64_t)blockDim.x) {
    int32_t jj[kBatch];
    uint32_t kk[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int64_t q = q0 + (int64_t)u * blockDim.x + threadIdx.x;
      jj[u] = q < e1 ? nbr[q] : -1;
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) kk[u] = jj[u] >= 0 ? keys[jj[u]] : kNoKey;
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const uint32_t key = kk[u];
      const uint32_t up = __shfl_up(key, 1);
      if (key == kNoKey || ((threadIdx.x & 63) != 0 && up == key)) continue;
      int h = tile_hash(key);
      bool placed = false;
      for (int probe = 0; probe < kTileHash && !placed; ++probe) {
        const uint32_t old = atomicCAS(&hkey[h], kNoKey, key);
        if (old == kNoKey) {
          const int c = atomicAdd(&ncell, 1);
          if (c < kTileCells) cslot[c] = h;
          else overflow = 1;
          placed = true;
        } else if (old == key) {
          placed = true;
        } else {
          h = (h + 1) & (kTileHash - 1);
        }
      }
      if (!placed) overflow = 1;
    }
  }
  __syncthreads();
  const int32_t nc = min(ncell, kTileCells);
  if (!overflow) {
    for (int c = threadIdx.x; c < nc; c += blockDim.x) {
      const uint32_t key = hkey[cslot[c]];
      cstart[c] = cs[key];
      clen[c] = ce[key] - cs[key];
    }
  }
  __syncthreads();
  if (!overflow) {  // rank sort by first index (distinct: distinct non-empty cells)
    for (int c = threadIdx.x; c < nc; c += blockDim.x) {
      int rk = 0;
      for (int d = 0; d < nc; ++d) rk += cstart[d] < cstart[c] ? 1 : 0;
      sslot[rk] = cslot[c];
      sstart[rk] = cstart[c];
      slen[rk] = clen[c];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t nr = 0, base = 0;
    bool wide = overflow != 0;
    for (int c = 0; c < nc && !wide; ++c) {
      if (c == 0 || sstart[c] != sstart[c - 1] + slen[c - 1]) {
        if (nr == kTileRuns) { wide = true; break; }
        tab[4 + 2 * nr] = sstart[c];
        tab[5 + 2 * nr] = base;
        ++nr;
      }
      hadd[sslot[c]] = base - sstart[c];
      base += slen[c];
    }
    // the tiled passes stage the rows' entry indices too (tile_ids_fill);
    // the count is formed in 32 bits (row offsets are int32): hipcc
    // (ROCm 7.2) emitted the select of tab[0] on the wrong condition code
    // for the 64-bit form of this test
    const int32_t ids = off[r1] - (off[r0] & ~7);
    if (base > cap || ids > kTileIds) wide = true;
    if (wide) {
      tab[0] = -1;
      tab[1] = 0;
    } else {
      tab[0] = nr;
      tab[1] = base;
      tab[4 + 2 * nr] = 0;
      tab[5 + 2 * nr] = base;  // sentinel
    }
  }
  __syncthreads();
  for (int w = threadIdx.x; w < kTileWords; w += blockDim.x) tiles[(int64_t)t * kTileWords + w] = tab[w];
  if (tab[0] < 0) return;  // wide: the passes use the 32-bit ids
  for (int64_t q0 = e0; q0 < e1; q0 += kBatch * (int64_t)blockDim.x) {
    int32_t jj[kBatch];
    uint32_t kk[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int64_t q = q0 + (int64_t)u * blockDim.x + threadIdx.x;
      jj[u] = q < e1 ? nbr[q] : -1;
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) kk[u] = jj[u] >= 0 ? keys[jj[u]] : kNoKey;
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      if (jj[u] < 0) continue;
      int h = tile_hash(kk[u]);
      while (hkey[h] != kk[u]) h = (h + 1) & (kTileHash - 1);  // present: inserted above
      loc[q0 + (int64_t)u * blockDim.x + threadIdx.x] = (uint16_t)(jj[u] + hadd[h]);
    }
  }
}

}  // namespace crk

using namespace crk;

extern "C" {

int crk_adjacency_count(CrkContext* ctx, const CrkConfig* cfg, const CrkGrid* grid, int32_t n, const float* x,
                        const float* h, uint32_t* keys, int32_t* perm, int32_t* rank, int32_t* cell_start,
                        int32_t* cell_end, int32_t* row_offsets, float* xb, float* hb, int64_t* n_edges,
                        void* stream) {
  CRK_REQUIRE(ctx && cfg && grid && n_edges, "crk_adjacency_count: null argument");
  CRK_REQUIRE(n > 0, "crk_adjacency_count: n must be positive (got %d)", n);
  CRK_REQUIRE(x && h && keys && perm && rank && cell_start && cell_end && row_offsets && xb && hb,
              "crk_adjacency_count: null buffer");
  int st = require_runs_n(ctx, n, "crk_adjacency_count");
  if (st) return st;
  int64_t table = 0;
  for (int32_t r = 0; r < run_count(ctx); ++r) {
    CRK_REQUIRE(grid[r].dim == cfg->domain.dim, "grid/config dimension mismatch (grid %d)", r);
    CRK_REQUIRE(grid[r].table_size >= 1 && grid[r].table_size <= ((int64_t)1 << 30), "bad grid table size (grid %d)",
                r);
    table += grid[r].table_size;
  }
  CRK_REQUIRE(table <= ((int64_t)1 << 30), "the runs' cell tables hold %lld cells; at most 2^30 are supported",
              (long long)table);
  hipStream_t s = (hipStream_t)stream;
  switch (cfg->domain.dim) {
    case 1: return count_impl<1>(ctx, cfg, grid, n, x, h, keys, perm, rank, cell_start, cell_end, row_offsets, xb, hb, n_edges, s);
    case 2: return count_impl<2>(ctx, cfg, grid, n, x, h, keys, perm, rank, cell_start, cell_end, row_offsets, xb, hb, n_edges, s);
    case 3: return count_impl<3>(ctx, cfg, grid, n, x, h, keys, perm, rank, cell_start, cell_end, row_offsets, xb, hb, n_edges, s);
  }
  set_error("dim must be 1..3");
  return CRK_EINVAL;
}

int crk_adjacency_count_keep(CrkContext* ctx, const CrkConfig* cfg, const CrkGrid* grid, int32_t n, const float* x,
                             const float* h, const uint8_t* keep, const int64_t* delta, const CrkAdjacency* prior,
                             const uint32_t* prior_keys, const float* prior_xb, const float* prior_hb,
                             uint32_t* keys, int32_t* perm, int32_t* rank, int32_t* cell_start, int32_t* cell_end,
                             int32_t* row_offsets, float* xb, float* hb, int64_t* n_edges, void* stream) {
  CRK_REQUIRE(ctx && ctx->n_runs > 0, "crk_adjacency_count_keep: needs bound runs (crk_bind_runs)");
  CRK_REQUIRE(keep && delta && prior && prior_keys && prior_xb && prior_hb,
              "crk_adjacency_count_keep: null keep, delta or prior buffer");
  CRK_REQUIRE(prior->n == n && prior->rank && prior->row_offsets && (prior->n_edges == 0 || prior->nbr),
              "crk_adjacency_count_keep: the prior adjacency must hold the same %d particles", n);
  CRK_REQUIRE(ctx && cfg && grid && n_edges, "crk_adjacency_count_keep: null argument");
  CRK_REQUIRE(x && h && keys && perm && rank && cell_start && cell_end && row_offsets && xb && hb,
              "crk_adjacency_count_keep: null buffer");
  int st = require_runs_n(ctx, n, "crk_adjacency_count_keep");
  if (st) return st;
  int64_t table = 0;
  for (int32_t r = 0; r < ctx->n_runs; ++r) {
    CRK_REQUIRE(grid[r].dim == cfg->domain.dim, "grid/config dimension mismatch (grid %d)", r);
    CRK_REQUIRE(grid[r].table_size >= 1 && grid[r].table_size <= ((int64_t)1 << 30), "bad grid table size (grid %d)",
                r);
    table += grid[r].table_size;
  }
  CRK_REQUIRE(table <= ((int64_t)1 << 30), "the runs' cell tables hold %lld cells; at most 2^30 are supported",
              (long long)table);
  const KeepIn kin{keep, delta, prior->rank, prior_keys, prior_xb, prior_hb, prior->row_offsets, prior->nbr};
  hipStream_t s = (hipStream_t)stream;
  switch (cfg->domain.dim) {
    case 1: return count_impl<1>(ctx, cfg, grid, n, x, h, keys, perm, rank, cell_start, cell_end, row_offsets, xb, hb, n_edges, s, &kin);
    case 2: return count_impl<2>(ctx, cfg, grid, n, x, h, keys, perm, rank, cell_start, cell_end, row_offsets, xb, hb, n_edges, s, &kin);
    case 3: return count_impl<3>(ctx, cfg, grid, n, x, h, keys, perm, rank, cell_start, cell_end, row_offsets, xb, hb, n_edges, s, &kin);
  }
  set_error("dim must be 1..3");
  return CRK_EINVAL;
}

int crk_adjacency_fill(CrkContext* ctx, const CrkConfig* cfg, const CrkGrid* grid, int32_t n,
                       const int32_t* cell_start, const int32_t* cell_end, const int32_t* row_offsets,
                       const float* xb, const float* hb, int32_t* nbr, int64_t capacity, void* stream) {
  CRK_REQUIRE(ctx && cfg && grid && cell_start && cell_end && row_offsets && xb && hb, "crk_adjacency_fill: null argument");
  CRK_REQUIRE(n > 0, "n must be positive");
  hipStream_t s = (hipStream_t)stream;
  const bool collected = ctx->collect_rows == row_offsets && ctx->collect_n == n && ctx->collect_E >= 0;
  int32_t E = 0;
  if (collected) {
    E = (int32_t)ctx->collect_E;
  } else {
    CRK_HIP(hipMemcpyAsync(&E, row_offsets + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    CRK_HIP(hipStreamSynchronize(s));
  }
  if ((int64_t)E > capacity) {
    set_error("crk_adjacency_fill: %d edges exceed the capacity %lld", E, (long long)capacity);
    return CRK_ECAPACITY;
  }
  if (E > 0) CRK_REQUIRE(nbr, "crk_adjacency_fill: null neighbour buffer");
  int st = require_runs_n(ctx, n, "crk_adjacency_fill");
  if (st) return st;
  // the kept runs of a crk_adjacency_count_keep (for this fill only)
  const KeepDev kp{ctx->keep_runs, ctx->keep_off, ctx->keep_nbr};
  ctx->keep_runs = nullptr;
  ctx->keep_off = ctx->keep_nbr = nullptr;
  CRK_REQUIRE(!kp.run_keep || ctx->n_runs > 0, "crk_adjacency_fill: kept runs without bound runs");
  GridDev g = make_griddev(*grid, *cfg);
  CellScan sc = make_scan(*grid, *cfg);
  Geo geo = make_geo(*cfg);
  if (collected) {
    ctx->collect_rows = nullptr;  // one fill per count
    ctx->collect_E = -1;
    if (ctx->collect_ok) {
      if (E > 0) {
        hipLaunchKernelGGL(k_rows_copy<16>, dim3(group_blocks(n, 16)), dim3(kBlock), 0, s,
                           (const int32_t*)ctx->nbr_tmp.p, ctx->nbr_cap_used, row_offsets, n, nbr, run_ids(ctx),
                           kp);
        CRK_LAUNCH_CHECK("k_rows_copy");
      }
      return CRK_OK;
    }
  }
  // with runs bound: the per-run grids uploaded by crk_adjacency_count
  const GridRun* gr = ctx->n_runs > 0 ? (const GridRun*)ctx->s_gridruns.p : nullptr;
  const int32_t* run = run_ids(ctx);
  if (gr) {
    CRK_DIM_DISPATCH(cfg->domain.dim,
        hipLaunchKernelGGL((k_neighbors_g<DD, nbr_group<DD>(), 1, true>), dim3(group_blocks(n, nbr_group<DD>())),
                           dim3(kBlock), 0, s, g, sc, geo, cfg->verletScale, xb, hb, cell_start, cell_end, n,
                           (int32_t*)row_offsets, nbr, gr, run, 0, (int32_t*)nullptr, kp));
  } else {
    CRK_DIM_DISPATCH(cfg->domain.dim,
        hipLaunchKernelGGL((k_neighbors_g<DD, nbr_group<DD>(), 1, false>), dim3(group_blocks(n, nbr_group<DD>())),
                           dim3(kBlock), 0, s, g, sc, geo, cfg->verletScale, xb, hb, cell_start, cell_end, n,
                           (int32_t*)row_offsets, nbr, gr, run, 0, (int32_t*)nullptr, kp));
  }
  CRK_LAUNCH_CHECK("k_neighbors(fill)");
  return CRK_OK;
}

int crk_adjacency_valid(CrkContext* ctx, const CrkConfig* cfg, int32_t n, const int32_t* perm, const float* xb,
                        const float* hb, const float* x, const float* h, int32_t* valid, void* stream) {
  CRK_REQUIRE(ctx && cfg && perm && xb && hb && x && h, "crk_adjacency_valid: null argument");
  CRK_REQUIRE(n > 0, "n must be positive");
  hipStream_t s = (hipStream_t)stream;
  int st = require_runs_n(ctx, n, "crk_adjacency_valid");
  if (st) return st;
  const int32_t nr = run_count(ctx);
  if ((st = ensure(ctx->s_red, sizeof(float) * 3 * (size_t)nr))) return st;
  if ((st = ensure_pinned(ctx, sizeof(float) * 3 * (size_t)nr))) return st;
  float* red = (float*)ctx->s_red.p;
  hipLaunchKernelGGL(k_valid_init, dim3(1), dim3(kBlock), 0, s, red, nr);
  Geo geo = make_geo(*cfg);
  const RunsDev R = runs_of(ctx);
  const unsigned nb = span_blocks(ctx, reduce_blocks(n));
  CRK_DIM_DISPATCH(cfg->domain.dim, hipLaunchKernelGGL(k_valid<DD>, dim3(nb), dim3(kBlock), 0, s, geo, perm, xb, hb,
                                                       x, h, n, red, R));
  CRK_LAUNCH_CHECK("k_valid");
  if (!valid) return verlet_flag_launch(ctx, red, 3, 0, 1, 2, nr, (double)cfg->verletScale, s);  // deferred
  float* hostv = (float*)ctx->pinned.p;
  CRK_HIP(hipMemcpyAsync(hostv, red, 3 * sizeof(float) * (size_t)nr, hipMemcpyDeviceToHost, s));
  CRK_HIP(hipStreamSynchronize(s));
  // every run's list must hold (the lists of all runs are rebuilt together)
  const double sv = cfg->verletScale;
  int ok = 1;
  for (int32_t r = 0; r < nr; ++r) {
    const double g = hostv[3 * r + 1];
    ok = ok && g < sv && 2.0 * std::sqrt((double)hostv[3 * r]) <= (sv - g) * (double)hostv[3 * r + 2];
  }
  *valid = ok;
  return CRK_OK;
}

int crk_adjacency_export(CrkContext* ctx, int32_t n, const int32_t* perm, const int32_t* row_offsets,
                         const int32_t* nbr, int64_t n_edges, int64_t* i_out, int64_t* j_out, int64_t* num_neighbors,
                         int64_t* neighbor_offsets, void* stream) {
  CRK_REQUIRE(ctx && perm && row_offsets, "crk_adjacency_export: null argument");
  CRK_REQUIRE(n > 0, "n must be positive");
  CRK_REQUIRE(n_edges == 0 || nbr || (!i_out && !j_out), "crk_adjacency_export: null neighbour list");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_export, dim3(nblocks(n)), dim3(kBlock), 0, s, perm, row_offsets, nbr, n, i_out, j_out,
                     num_neighbors, neighbor_offsets);
  CRK_LAUNCH_CHECK("k_export");
  return CRK_OK;
}

int crk_adjacency_tiles(CrkContext* ctx, const CrkAdjacency* adj, const uint32_t* keys, const int32_t* cell_start,
                        const int32_t* cell_end, int32_t cap, int32_t* tiles, uint16_t* nbr_loc, void* stream) {
  CRK_REQUIRE(ctx && adj && keys && cell_start && cell_end && tiles && adj->row_offsets,
              "crk_adjacency_tiles: null argument");
  CRK_REQUIRE(adj->n > 0, "crk_adjacency_tiles: empty adjacency");
  CRK_REQUIRE(adj->n_edges == 0 || (adj->nbr && nbr_loc), "crk_adjacency_tiles: neighbour lists missing");
  CRK_REQUIRE(cap > 0 && cap <= kTileCap, "crk_adjacency_tiles: cap must be in [1, %d] (got %d)", kTileCap, cap);
  const unsigned nt = nblocks(adj->n, kTileRows);
  hipLaunchKernelGGL(k_tiles, dim3(nt), dim3(kBlock), 0, (hipStream_t)stream, adj->n, adj->row_offsets, adj->nbr,
                     keys, cell_start, cell_end, cap, tiles, nbr_loc);
  CRK_LAUNCH_CHECK("k_tiles");
  return CRK_OK;
}

}  // extern "C"
