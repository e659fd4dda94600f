// tiles.hpp -- LDS tiles of an adjacency.
//
// The sorted rows are grouped CRK_TILE_ROWS at a time (a tile).  Rows that are
// consecutive in Morton order sit in a compact patch of cells, so every
// neighbour of a tile's rows lies in a few contiguous ranges of sorted
// indices ("runs": the patch's cells, merged where one ends where the next
// begins).  crk_adjacency_tiles records each tile's runs and, for every
// Verlet entry, the neighbour's position in the concatenation of its tile's
// runs (a 16-bit local index).  A pair pass then copies the neighbour
// records of one tile into LDS with coalesced loads and reads them there,
// instead of gathering every neighbour through the texture path.
//
// Tile table, CRK_TILE_WORDS int32 per tile:
//   [0] number of runs, or -1 for a "wide" tile (more than cap particles or
//       CRK_TILE_MAXRUNS runs: its passes gather from global memory)
//   [1] U, particles in the runs
//   [4 + 2 r], [5 + 2 r]  first sorted index and local offset of run r;
//   run nruns is a sentinel with local offset U.
#pragma once

#include "pairs.hpp"

namespace crk {

constexpr int kTileRows = CRK_TILE_ROWS;
constexpr int kTileCap = CRK_TILE_CAP;
constexpr int kTileRuns = CRK_TILE_MAXRUNS;
constexpr int kTileWords = CRK_TILE_WORDS;
constexpr int kTileIds = CRK_TILE_IDS;

// Copy tile t's table into LDS (every thread of the block calls; synchronises;
// requires blockDim.x == kBlock).
__device__ inline void tile_table(const int32_t* __restrict__ tiles, int32_t t, int32_t* sh) {
  for (int w = threadIdx.x; w < kTileWords; w += blockDim.x) sh[w] = tiles[(int64_t)t * kTileWords + w];
  __syncthreads();
}

// Sorted index of local slot u (binary search over the runs' offsets).
__device__ inline int32_t tile_global(const int32_t* sh, int32_t u) {
  int lo = 0, hi = sh[0];  // run lo .. hi - 1 holds u
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (sh[5 + 2 * mid] <= u) lo = mid;
    else hi = mid;
  }
  return sh[4 + 2 * lo] + (u - sh[5 + 2 * lo]);
}

// A tile's rows are processed by one block of blockDim.x threads (256, 512
// or 1024, the pass's tile block size), G lanes per row, blockDim.x / G rows
// per round.
constexpr int kTileThreadsMax = 1024;
template <int G>
__device__ inline int tile_rounds() {
  return (kTileRows * G + (int)blockDim.x - 1) / (int)blockDim.x;
}
// Row k of round `round` of tile t.  Lanes past the tile's last row (a round
// wider than the tile) or past n are invalid and keep a row index inside
// the tile, so that their (unused) reads stay in bounds.
template <int G>
__device__ inline Lane tile_lane(int32_t t, int round, int32_t n) {
  Lane l;
  l.g = threadIdx.x % G;
  const int32_t lr = round * (int32_t)(blockDim.x / G) + (int32_t)(threadIdx.x / G);
  l.k = t * kTileRows + lr;
  l.valid = lr < kTileRows && l.k < n;
  if (!l.valid) l.k = min(t * kTileRows + min(lr, kTileRows - 1), n - 1);
  return l;
}

// Neighbour sources of the pair passes: the global records through the
// 32-bit neighbour ids, or the tile's LDS copy through the local indices.
// Each source serves the row's neighbours' {x, h} record (geo) and one
// pass-specific record (aux); gid(q) is the entry's sorted neighbour index.
template <class A>
struct GlobalSrc {
  const int32_t* __restrict__ nbr;
  const float4* __restrict__ g;
  const A* __restrict__ a;
  __device__ inline int32_t id(int64_t q) const { return nbr[q]; }
  __device__ inline int32_t gid(int64_t q) const { return nbr[q]; }
  __device__ inline float4 geo(int32_t j) const { return g[j]; }
  __device__ inline A aux(int32_t j) const { return a[j]; }
};
template <class A>
struct TileSrc {
  const uint16_t* sid;  // the 16-bit indices of entries base ..: an LDS copy
                        // (tile_ids_fill) or the global array with base 0
  int32_t base;
  const int32_t* __restrict__ nbr;
  const float4* g;  // LDS
  const A* a;       // LDS
  __device__ inline int32_t id(int64_t q) const { return sid[q - base]; }
  __device__ inline int32_t gid(int64_t q) const { return nbr[q]; }
  __device__ inline float4 geo(int32_t j) const { return g[j]; }
  __device__ inline A aux(int32_t j) const { return a[j]; }
};

// The 16-bit indices of a tile's entries, [off[r0], off[r1]) of its rows'
// CSR ranges, copied into LDS once per block with 16-byte loads (the entries
// of 256 consecutive rows are one contiguous range), so that a pair pass
// reads every index from LDS instead of one 2-byte global load per entry.
// The tile table admits a tile only when the range fits: at most kTileIds
// entries from its 8-aligned start (crk_adjacency_tiles), so that every
// tiled pass takes the same path for a tile.  Returns the index of the first
// staged entry (the base of TileSrc).  Synchronises.
__device__ inline int32_t tile_ids_fill(const uint16_t* __restrict__ loc, int32_t e0, int32_t e1, uint16_t* sids) {
  const int32_t start = e0 & ~7;
  const int32_t vend = max(e1 & ~7, start);
  const int32_t nv = (vend - start) >> 3;
  const uint4* src = reinterpret_cast<const uint4*>(loc + start);
  uint4* dst = reinterpret_cast<uint4*>(sids);
  for (int32_t w = threadIdx.x; w < nv; w += blockDim.x) dst[w] = src[w];
  for (int32_t q = vend + threadIdx.x; q < e1; q += blockDim.x) sids[q - start] = loc[q];
  __syncthreads();
  return start;
}

// tile_global with a fixed number of steps (the run holding u is the last
// with offset <= u; at most kTileRuns < 64 runs), so that it unrolls where
// several lookups are in flight
__device__ inline int32_t tile_global_unrolled(const int32_t* sh, int32_t u) {
  static_assert(kTileRuns < 64, "six search steps");
  const int32_t nruns = sh[0];
  int32_t lo = 0;
#pragma unroll
  for (int step = 32; step > 0; step >>= 1) {
    const int32_t c = lo + step;
    if (c < nruns && sh[5 + 2 * c] <= u) lo = c;
  }
  return sh[4 + 2 * lo] + (u - sh[5 + 2 * lo]);
}

// Stage tile t for a pass whose occupancy its LDS limits (one block per CU,
// registers to spare): the table, then the entries' 16-bit indices (as
// tile_ids_fill) and the records of the tile's particles (C4 float4 each)
// with all of a thread's global loads in flight at once.  A load-then-store
// loop would expose one memory latency per iteration, and with one block on
// the CU nothing else hides it.  Returns the base of the staged indices
// (wide tiles, tab[0] < 0, stage nothing).  Synchronises.
template <int C4, int KR = 8>
__device__ inline int32_t tile_stage(const int32_t* __restrict__ tiles, int32_t t, const int32_t* __restrict__ off,
                                     int32_t n, const uint16_t* __restrict__ loc, const float4* __restrict__ rec,
                                     int32_t* tab, float4* srec, uint16_t* sid) {
  const int32_t e0 = off[t * kTileRows], e1 = off[min(n, (t + 1) * kTileRows)];
  for (int w = threadIdx.x; w < kTileWords; w += blockDim.x) tab[w] = tiles[(int64_t)t * kTileWords + w];
  __syncthreads();
  if (tab[0] < 0) return 0;
  const int32_t nr = tab[1] * C4;
  const int32_t start = e0 & ~7;
  const int32_t vend = max(e1 & ~7, start);
  const int32_t nv = (vend - start) >> 3;
  const uint4* isrc = reinterpret_cast<const uint4*>(loc + start);
  uint4* idst = reinterpret_cast<uint4*>(sid);
  const int32_t bd = blockDim.x;
  constexpr int KI = 4;  // (KR = 8: one round at 1024 threads)
  // (zero-initialised and loaded per element, so that the staged values
  // stay in registers)
  for (int32_t w0 = threadIdx.x, r0 = threadIdx.x; w0 < nv || r0 < nr; w0 += KI * bd, r0 += KR * bd) {
    uint4 iv[KI];
    float4 rv[KR];
#pragma unroll
    for (int k = 0; k < KI; ++k) {
      iv[k] = make_uint4(0u, 0u, 0u, 0u);
      if (w0 + k * bd < nv) iv[k] = isrc[w0 + k * bd];
    }
#pragma unroll
    for (int k = 0; k < KR; ++k) {
      const int32_t e = r0 + k * bd;
      rv[k] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (e < nr) {
        const int32_t u = e / C4;
        rv[k] = rec[(int64_t)tile_global_unrolled(tab, u) * C4 + (e - u * C4)];
      }
    }
#pragma unroll
    for (int k = 0; k < KI; ++k)
      if (w0 + k * bd < nv) idst[w0 + k * bd] = iv[k];
#pragma unroll
    for (int k = 0; k < KR; ++k)
      if (r0 + k * bd < nr) srec[r0 + k * bd] = rv[k];
  }
  for (int32_t q = vend + threadIdx.x; q < e1; q += bd) sid[q - start] = loc[q];
  __syncthreads();
  return start;
}

// Stage tile t for a pass with per-particle records of type V: the table,
// the entries' 16-bit indices (as tile_ids_fill) and, for every particle u of
// the tile's runs, store(u, load(global index of u)), with up to KP
// particles' and KI index chunks' loads of a thread in flight at once (a
// load-then-store loop such as tile_ids_fill's exposes one memory latency
// per iteration).  KP and KI trade that latency against the
// registers of the pass's prologue.  Returns the base of the staged indices
// (wide tiles, tab[0] < 0, stage nothing).  KI = 0 stages no indices (the
// pass reads them from global memory; e0, e1 and loc are then unused).
// Synchronises.
template <int KP, int KI, class V, class Load, class Store>
__device__ inline int32_t tile_stage_fn(const int32_t* __restrict__ tiles, int32_t t, int32_t e0, int32_t e1,
                                        const uint16_t* __restrict__ loc, int32_t* tab, uint16_t* sid,
                                        const Load& load, const Store& store) {
  for (int w = threadIdx.x; w < kTileWords; w += blockDim.x) tab[w] = tiles[(int64_t)t * kTileWords + w];
  __syncthreads();
  if (tab[0] < 0) return 0;
  const int32_t U = tab[1];
  const int32_t start = e0 & ~7;
  const int32_t vend = max(e1 & ~7, start);
  const int32_t nv = (vend - start) >> 3;
  const uint4* isrc = reinterpret_cast<const uint4*>(loc + start);
  uint4* idst = reinterpret_cast<uint4*>(sid);
  const int32_t bd = blockDim.x;
  for (int32_t w0 = threadIdx.x, u0 = threadIdx.x; (KI > 0 && w0 < nv) || u0 < U; w0 += KI * bd, u0 += KP * bd) {
    uint4 iv[KI > 0 ? KI : 1];
    V pv[KP];
#pragma unroll
    for (int k = 0; k < KI; ++k) {
      iv[k] = make_uint4(0u, 0u, 0u, 0u);
      if (w0 + k * bd < nv) iv[k] = isrc[w0 + k * bd];
    }
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      pv[k] = V{};
      if (u0 + k * bd < U) pv[k] = load(tile_global_unrolled(tab, u0 + k * bd));
    }
#pragma unroll
    for (int k = 0; k < KI; ++k)
      if (w0 + k * bd < nv) idst[w0 + k * bd] = iv[k];
#pragma unroll
    for (int k = 0; k < KP; ++k)
      if (u0 + k * bd < U) store(u0 + k * bd, pv[k]);
  }
  if (KI > 0)
    for (int32_t q = vend + threadIdx.x; q < e1; q += bd) sid[q - start] = loc[q];
  __syncthreads();
  return start;
}

// the bound adjacency carries tiles the tiled passes can use (d <= 2)
inline bool use_tiles(const CrkContext* c) { return c->adj.tiles && c->adj.nbr_loc && c->dim <= 2; }
inline unsigned tile_count(int32_t n) { return nblocks(n, kTileRows); }

static_assert(kTileRows == kBlock, "the default tile block holds one thread per tile row");
// tile block size of a pass (crk_set_block_size): 256, 512 or 1024
inline bool valid_tile_threads(int32_t t) { return t == 256 || t == 512 || t == 1024; }

}  // namespace crk

// dispatch of the tiled kernels over the kernel function (KK), d in {1, 2}
// (DD) and the group size (G)
#define CRK_DIM12_G_DISPATCH(DIM, GS, ...)                                                     \
  switch (DIM) {                                                                               \
    CRK_D_CASE(1, CRK_G_SWITCH(GS, __VA_ARGS__))                                               \
    CRK_D_CASE(2, CRK_G_SWITCH(GS, __VA_ARGS__))                                               \
    default: ::crk::set_error("tiled passes: dim must be 1 or 2"); return CRK_EINVAL;         \
  }
#define CRK_KDG12_DISPATCH(KERN, DIM, GS, ...)                                                 \
  switch (KERN) {                                                                              \
    CRK_K_CASE(0, CRK_DIM12_G_DISPATCH(DIM, GS, __VA_ARGS__))                                  \
    CRK_K_CASE(1, CRK_DIM12_G_DISPATCH(DIM, GS, __VA_ARGS__))                                  \
    CRK_K_CASE(2, CRK_DIM12_G_DISPATCH(DIM, GS, __VA_ARGS__))                                  \
    CRK_K_CASE(3, CRK_DIM12_G_DISPATCH(DIM, GS, __VA_ARGS__))                                  \
    default: ::crk::set_error("unknown kernel %d", KERN); return CRK_EINVAL;                   \
  }
