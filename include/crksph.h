/*
 * crksph.h -- C-ABI of the MI355X-native CRKSPH compressible step.
 *
 * The library (warpsph_amd/lib/libcrksph_hip.so, gfx950 HIP) replaces the
 * device side of warpSPH's CRKSPH step: everything below
 *   crkSPH_step(system, dt, config, compParams)      src/warpSPH/schemes/crkSPH.py:19-366
 * including the warpSPHCore pieces it calls (buildVerletList, computeCRKFactors,
 * warpOperation, the SPH kernel functions) that are not part of the reference
 * repository.
 *
 * Conventions
 *  - All particle arrays are device pointers, FP32, row-major: scalars [n],
 *    vectors [n*dim], dim x dim tensors [n*dim*dim] (row = component).
 *  - "user order" is the caller's particle order (the order of CompSPHState).
 *    The library works internally in Morton-sorted order ("sorted order");
 *    an adjacency carries the permutation between the two.
 *  - Edge arrays (ap_ij, av_ij, f_ij) are stored in the adjacency's row order:
 *    row r (sorted particle r, user particle perm[r]) owns edges
 *    [row_offsets[r], row_offsets[r+1]).
 *  - Every call is asynchronous on the given stream unless documented as
 *    blocking.  Calls return CRK_OK (0) or a negative status; the message of
 *    the last failure on the calling thread is available from crk_last_error.
 *  - The caller owns every buffer passed in.  The library owns only scratch
 *    memory inside a CrkContext.  A context is not thread-safe; use one per
 *    host thread / stream.
 */
#ifndef CRKSPH_H
#define CRKSPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
enum {
  CRK_OK = 0,
  CRK_EINVAL = -1,     /* invalid argument (shape, null pointer, parameter)   */
  CRK_EHIP = -2,       /* HIP runtime error                                   */
  CRK_ECAPACITY = -3,  /* caller buffer too small (edge capacity)             */
  CRK_ESTATE = -4      /* call order violated (e.g. pass before gather)       */
};

/* --------------------------------------------------------------- enums
 * Values mirror the reference enums where one exists.                      */
enum { /* KernelFunctions subset (warpSPHCore; values are this library's) */
  CRK_KERNEL_WENDLAND2 = 0,
  CRK_KERNEL_WENDLAND4 = 1,
  CRK_KERNEL_WENDLAND6 = 2,
  CRK_KERNEL_B7 = 3     /* degree-7 B-spline, knots at r/H = 1/4, 1/2, 3/4
                           (the kernel of every reference compressible example,
                           e.g. examples/compressible/12-Kelvin-Helmholtz.ipynb) */
};
enum { /* EnergyScheme, src/warpSPH/enumTypes.py:5-11 */
  CRK_ENERGY_EQUALWORK = 0, CRK_ENERGY_PDV = 1, CRK_ENERGY_DIMINISHING = 2,
  CRK_ENERGY_MONOTONIC = 3, CRK_ENERGY_HYBRID = 4, CRK_ENERGY_CRK = 5
};
enum { /* AdaptiveSupportScheme, src/warpSPH/enumTypes.py:14-17 */
  CRK_SUPPORT_NONE = 0, CRK_SUPPORT_MONAGHAN = 1, CRK_SUPPORT_OWEN = 2
};
enum { /* ViscositySwitch, src/warpSPH/enumTypes.py:20-27 */
  CRK_SWITCH_CULLEN_DEHNEN = 2, CRK_SWITCH_NONE = 6,
  CRK_SWITCH_LEARNED = 100 /* no reference counterpart: per-particle MLP */
};

/* ------------------------------------------------------ parameter blocks */
#pragma pack(push, 4)
/* DomainDescription, src/warpSPH/utils/domain.py:7-17 */
typedef struct CrkDomain {
  int32_t dim;            /* 1, 2 or 3 */
  int32_t periodic[3];
  float   min[3];
  float   max[3];
} CrkDomain;

/* DiffusionParameters, src/warpSPH/configurations/moduleConfigurations/diffusionParameters.py:23-36 */
typedef struct CrkDiffusionParams {
  float   c_s, C_l, C_q, Cu_l, Cu_q, K, thermalConductivity;
  int32_t viscosityTerm, thermalConductivityTerm;
  int32_t scaleBeta, monaghanSwitch, correctXi;
} CrkDiffusionParams;

/* CRKViscosity, src/warpSPH/configurations/crkSPH.py:17-25 */
typedef struct CrkViscosityParams {
  float   eta_fold, eta_crit;
  int32_t enableCRKLimiter, enableVanLeerLimiter;
  int32_t forceVanLeerOff, forceVanLeerOn;
} CrkViscosityParams;

/* ViscositySwitchConfig, .../moduleConfigurations/viscositySwitchParameters.py:12-24 */
typedef struct CrkSwitchParams {
  int32_t scheme;         /* CRK_SWITCH_* */
  int32_t limitXi;
  int32_t correctVelocityGradient; /* must be 0 (M-matrix correction not built) */
  float   alpha_min, alpha_max, beta_c, beta_d, beta_xi;
} CrkSwitchParams;

/* Everything one derivative evaluation needs.  Mirrors SimulationConfig
 * (src/warpSPH/configurations/simulationConfig.py:16-44) and CRKSPHConfig
 * (src/warpSPH/configurations/crkSPH.py:71-78 + compressibleConfig.py:23-40). */
typedef struct CrkConfig {
  CrkDomain          domain;
  int32_t            kernel;            /* CRK_KERNEL_* */
  float              verletScale;       /* SimulationConfig.verletScale */
  float              targetNeighbors;   /* SimulationConfig.targetNeighbors (n_H) */
  float              cflFactor, minDt, maxDt, dtGrowthFactor;
  float              gamma;             /* CompressibleSPHConfig.gamma */
  int32_t            energyScheme;      /* CRK_ENERGY_* (only CRK is built) */
  int32_t            adaptiveSupportScheme;   /* CRK_SUPPORT_NONE / _MONAGHAN / _OWEN */
  int32_t            adaptiveSupportIterations;
  float              adaptiveSupportThreshold;
  int32_t            compatibleEnergy;
  CrkDiffusionParams diffusion;
  CrkViscosityParams crkViscosity;
  CrkSwitchParams    switchParams;
  /* Extensions with no reference counterpart (0 = reference behaviour):
   * alphaWeightedQ: scale Q_i, Q_j by (alpha_i + alpha_j)/2 in accel/dudt.
   * evolveAlpha0:   store the updated alpha0 back into the state.          */
  int32_t            alphaWeightedQ;
  int32_t            evolveAlpha0;
} CrkConfig;

/* Uniform cell grid over which keys are Morton-interleaved. */
typedef struct CrkGrid {
  int32_t dim;
  int32_t bits;           /* bits per dimension of the Morton code         */
  int32_t ncell[3];       /* cells per dimension                            */
  float   origin[3];
  float   inv_cell[3];    /* 1 / cell width, as used by the key function   */
  int64_t table_size;     /* 2^(dim*bits): entries of cell_start/cell_end  */
} CrkGrid;
#pragma pack(pop)

/* ------------------------------------------------------------ context */
typedef struct CrkContext CrkContext;

int  crk_ctx_create(CrkContext** ctx);
int  crk_ctx_destroy(CrkContext* ctx);
/* Copies the last error message of this thread into buf (always terminated). */
int  crk_last_error(char* buf, size_t n);
/* Library version string, e.g. "crksph-hip 1 gfx950". */
const char* crk_version(void);
/* ABI self-check: sizeof of CrkDomain, CrkDiffusionParams, CrkViscosityParams,
 * CrkSwitchParams, CrkConfig, CrkGrid, CrkAdjacency, CrkState, CrkUpdate,
 * CrkEdges, CrkMlp (in this order) written to out[0..n-1]. */
int crk_struct_sizes(int64_t* out, int32_t n);

/* ---------------------------------------------- independent runs (ensemble)
 * Many independent simulations stepped together: the caller concatenates
 * their particles (run r = user indices [offsets[r], offsets[r+1])) and
 * binds the run table; every pass then covers all runs in one launch.  The
 * reference steps one simulation per call (crkSPH.py:19-366); no pair ever
 * joins two runs, and while runs are bound:
 *  - crk_adjacency_count/fill take grid[n_runs] (each from crk_grid_plan with
 *    the run's own h_max and extent, see crk_run_extents); cell_start/end hold
 *    the sum of the table sizes, run r's cells after those of runs 0..r-1.
 *    Each run keeps its index range in sorted order.
 *  - reductions are per run: hminmax_dev holds 2 n_runs floats (run r at
 *    [2r, 2r+1]), crk_minmax writes 2 n_runs, and the Verlet tests
 *    (crk_adjacency_valid, crk_stage_begin_checked, crk_owen_valid) pass when
 *    every run's list holds (all lists are rebuilt together).  The Owen
 *    max_ratio is the maximum over the runs.
 *  - with dt_particles (device [n], each particle's step, see crk_runs_dt and
 *    crk_cfl_dt_runs) the scalar dt of crk_switch, crk_switch_accel_dudt,
 *    crk_balance, crk_compatible_energy and crk_edges_export is ignored.
 * Calls with a different particle count fail while runs are bound.
 * n_runs = 0 unbinds (one simulation).  Blocking when the table changes;
 * binding the same offsets again only updates dt_particles.
 * No counterpart in the reference: it serves the batched ensemble.          */
int crk_bind_runs(CrkContext* ctx, int32_t n_runs, const int32_t* offsets_host,
                  const float* dt_particles, void* stream);
/* Steps replayed without host synchronisation (graph capture): the Verlet
 * tests of crk_adjacency_valid, crk_stage_begin_checked and crk_owen_valid
 * with valid == NULL OR their result into a device flag of the context (1:
 * some list failed, the step must be redone after a rebuild) instead of
 * returning it; crk_flag_clear / crk_flag_read reset and read it (read is
 * blocking), crk_copy_unless_flag copies bytes (a multiple of 4) only while
 * the flag is clear, crk_runs_advance_time adds each run's dt to its time
 * only while the flag is clear.  All but crk_flag_read are graph-capturable. */
int crk_flag_clear(CrkContext* ctx, void* stream);
int crk_flag_read(CrkContext* ctx, int32_t* flag, void* stream);
/* Speculative stages: after a deferred crk_stage_begin_checked (or
 * crk_owen_valid), the adaptive-support pass (or the support sweep) may be
 * launched before the host knows the test's result; it does nothing once
 * the flag is set (the support sweep then leaves every row without pairs),
 * until the next blocking test or crk_stage_begin.  crk_flag_read_async
 * queues a copy of the flag to pinned host memory in slot 0..3;
 * crk_flag_wait waits for that slot's copy alone (not for the work queued
 * after it) and returns the flag as it was copied.
 * No counterpart in the reference (its Verlet test returns to Python). */
int crk_flag_read_async(CrkContext* ctx, int32_t slot, void* stream);
int crk_flag_wait(CrkContext* ctx, int32_t slot, int32_t* flag);
int crk_copy_unless_flag(CrkContext* ctx, const void* src, void* dst, int64_t bytes, void* stream);
int crk_runs_advance_time(CrkContext* ctx, const float* dt_runs, double* t_runs, void* stream);
/* crk_cfl_dt_runs on the device (no host synchronisation): dt_out[r] (float)
 * with the clamps and the growth limit of prev_dt[r] (device, may be NULL),
 * and each particle's dt. */
int crk_cfl_dt_device(CrkContext* ctx, const CrkConfig* cfg, int32_t n, const float* supports,
                      const float* internalEnergies, const float* prev_dt, float* dt_out,
                      float* dt_particles, void* stream);
/* Run id of every particle (device int32[n]) of the bound runs. */
int crk_run_ids(CrkContext* ctx, int32_t n, int32_t* run_out, void* stream);
/* Per-run extent for grid planning (blocking): out_host[7r..7r+6] = h_max,
 * lo[3], hi[3] of run r (lo/hi over the dimensions in use, 0 elsewhere).     */
int crk_run_extents(CrkContext* ctx, const CrkConfig* cfg, int32_t n, const float* x,
                    const float* h, float* out_host, void* stream);
/* dt_particles[i] = (float) dt_runs[run of i] (dt_runs on the host). */
int crk_runs_dt(CrkContext* ctx, int32_t n, const double* dt_runs, float* dt_particles,
                void* stream);
/* crk_cfl_dt of every bound run (blocking): dt_out[r] with the clamps and
 * the growth limit of prev_dt[r] (prev_dt may be NULL); dt_particles, when
 * not NULL, receives each particle's (float) dt.                           */
int crk_cfl_dt_runs(CrkContext* ctx, const CrkConfig* cfg, int32_t n, const float* supports,
                    const float* internalEnergies, const double* prev_dt, double* dt_out,
                    float* dt_particles, void* stream);

/* ------------------------------------------------------ host utilities */
/* Kernel normalisation C_d and kernel scale H/h (Dehnen & Aly 2012). */
int crk_kernel_constants(int32_t kernel, int32_t dim, double* C_d, double* kernel_scale);
/* xi of the CFL step and of the Cullen-Dehnen relaxation time, the value of
 * warpSPHCore's sphKernel_xi (timestep/compressible.py:50,
 * CullenDehnen2010.py:286): the kernel scale H/h for the Wendland kernels;
 * for B7 the values the reference's example outputs fix (1D, 2D). */
int crk_kernel_xi(int32_t kernel, int32_t dim, double* xi);

/* Owen psi_H look-up table on n_h in [n_min, n_max] (nlut entries).
 * Restates computePsi, src/warpSPH/modules/adaptiveSupport/wp_psi.py:8-99
 * (only the psiH column is used by the step). Host function, blocking.    */
int crk_owen_lut(int32_t kernel, int32_t dim, int32_t nlut, double n_min, double n_max,
                 double* n_h_out, double* psiH_out);
/* The same table computed on the GPU directly into the context (one thread
 * per entry, FP64; within an ulp of crk_owen_lut), and a blocking read of the
 * context's table. */
int crk_owen_lut_device(CrkContext* ctx, int32_t kernel, int32_t dim, int32_t nlut, double n_min,
                        double n_max, double nh_target, void* stream);
int crk_owen_lut_read(CrkContext* ctx, double* n_h_out, double* psiH_out, void* stream);
/* Upload an Owen table into the context (device copy). */
int crk_set_owen_lut(CrkContext* ctx, int32_t nlut, const double* n_h, const double* psiH,
                     double nh_target, void* stream);

/* Grid for a neighbour search of radius verletScale * h_max.  For non-periodic
 * dimensions lo/hi are the particle bounding box (ignored when periodic).
 * Host function.                                                            */
int crk_grid_plan(const CrkConfig* cfg, float h_max, const float* lo, const float* hi,
                  CrkGrid* grid);
/* crk_grid_plan of every run from the extents of crk_run_extents
 * (extents[7r] = h_max, [7r+1..7r+3] = lo, [7r+4..7r+6] = hi), in one call;
 * table_total (may be NULL) = the sum of the grids' table sizes.  Host
 * function.  No counterpart in the reference (it serves the ensemble). */
int crk_grid_plan_runs(const CrkConfig* cfg, int32_t n_runs, const float* extents, CrkGrid* grids,
                       int64_t* table_total);

/* ------------------------------------------------- (a) adjacency build
 * Replaces warpSPHCore.buildVerletList (called at crkSPH.py:54-58 and
 * optimalSupportOwen.py:74): Morton keys -> stable LSD radix sort -> cell
 * start/end table -> CSR rows of every j with |x_ij| < verletScale*max(h_i,h_j)
 * (minimum image).  Rows are stored in sorted order; inside a row the
 * neighbours follow the 3^dim cell scan order, then sorted order.
 *
 * Step 1 (blocking: returns the edge count):
 *   keys[n] (sorted), perm[n] (sorted->user), rank[n] (user->sorted),
 *   cell_start/cell_end[grid.table_size], row_offsets[n+1], xb[n*dim], hb[n]
 *   (positions and supports at build time, sorted order).  The candidates are
 *   tested once: the rows are also collected into context scratch (a fixed
 *   number of slots per row, about 2.5 n_H verletScale^dim; at most 16 GiB),
 *   which step 2 copies; a build with a longer row, or one too large for the
 *   scratch, tests them again in step 2 (after an overflow the next build
 *   gets more slots).                                                       */
int crk_adjacency_count(CrkContext* ctx, const CrkConfig* cfg, const CrkGrid* grid,
                        int32_t n, const float* x, const float* h,
                        uint32_t* keys, int32_t* perm, int32_t* rank,
                        int32_t* cell_start, int32_t* cell_end,
                        int32_t* row_offsets, float* xb, float* hb,
                        int64_t* n_edges, void* stream);
/* Step 2: fills nbr[n_edges] (sorted neighbour indices).  Non-blocking
 * after step 1 of the same row_offsets on this context.                   */
int crk_adjacency_fill(CrkContext* ctx, const CrkConfig* cfg, const CrkGrid* grid,
                       int32_t n, const int32_t* cell_start, const int32_t* cell_end,
                       const int32_t* row_offsets, const float* xb, const float* hb,
                       int32_t* nbr, int64_t capacity, void* stream);


/* Blocking validity test of an adjacency built at (xb, hb) for the current
 * (x, h): valid iff max h/hb < s and 2*max|x - xb| <= (s - max h/hb)*min hb. */
int crk_adjacency_valid(CrkContext* ctx, const CrkConfig* cfg, int32_t n,
                        const int32_t* perm, const float* xb, const float* hb,
                        const float* x, const float* h, int32_t* valid, void* stream);
/* Expand to the reference AdjacencyList layout (user indices, int64):
 * i[E], j[E] per edge, numNeighbors[n] and neighborOffsets[n] per user particle. */
int crk_adjacency_export(CrkContext* ctx, int32_t n, const int32_t* perm, const int32_t* row_offsets,
                         const int32_t* nbr, int64_t n_edges, int64_t* i_out, int64_t* j_out,
                         int64_t* num_neighbors, int64_t* neighbor_offsets, void* stream);

/* ------------------------------------------------------ per-stage passes
 * A stage binds the context to one adjacency and one user-order state; the
 * passes below then run in sorted order on context-owned copies.            */
typedef struct CrkAdjacency {
  int32_t        n;
  int64_t        n_edges;
  const int32_t* perm;
  const int32_t* rank;
  const int32_t* row_offsets;
  const int32_t* nbr;
  /* LDS tiles (crk_adjacency_tiles), both NULL when the adjacency has none:
   * tiles [ceil(n / CRK_TILE_ROWS) * CRK_TILE_WORDS], nbr_loc [E] the
   * tile-local index of each entry's neighbour */
  const int32_t*  tiles;
  const uint16_t* nbr_loc;
} CrkAdjacency;

/* Step 1 for a bound ensemble whose runs with keep[r] != 0 (device
 * uint8[n_runs]) keep their lists of `prior` (built on this context with
 * the same runs): a kept run's sort order, build positions xb/hb and rows
 * are carried over (its keys are prior_keys shifted by delta[r], device
 * int64[n_runs], the change of the run's key base when other runs' cell
 * tables change size; grid[r] of a kept run must be the prior build's),
 * the other runs are rebuilt as crk_adjacency_count does.  The
 * crk_adjacency_fill that follows copies the kept rows.  No counterpart in
 * the reference (one list per simulation there): each run of an ensemble
 * then rebuilds its list only when its own test fails, as it would alone. */
int crk_adjacency_count_keep(CrkContext* ctx, const CrkConfig* cfg, const CrkGrid* grid, int32_t n,
                             const float* x, const float* h, const uint8_t* keep, const int64_t* delta,
                             const CrkAdjacency* prior, const uint32_t* prior_keys,
                             const float* prior_xb, const float* prior_hb, uint32_t* keys,
                             int32_t* perm, int32_t* rank, int32_t* cell_start, int32_t* cell_end,
                             int32_t* row_offsets, float* xb, float* hb, int64_t* n_edges,
                             void* stream);
/* Which bound runs failed the last Verlet test (blocking): which = 0 the
 * checked stage gather's (crk_stage_begin_checked), 1 the updated
 * supports' (crk_owen_valid); failed_host[r] = 1 for each failed run.
 * No counterpart in the reference (it serves crk_adjacency_count_keep). */
int crk_runs_failed(CrkContext* ctx, const CrkConfig* cfg, int32_t which, uint8_t* failed_host,
                    void* stream);

/* LDS tiles: CRK_TILE_ROWS consecutive sorted rows share one LDS copy of the
 * records of every particle their entries reach, at most CRK_TILE_CAP
 * particles in at most CRK_TILE_MAXRUNS contiguous ranges of sorted indices,
 * and one LDS copy of their entries' 16-bit local indices, at most
 * CRK_TILE_IDS entries from the 8-aligned start of the rows' CSR range
 * (larger tiles are "wide": every pass gathers them from global memory).    */
#define CRK_TILE_ROWS 256
#define CRK_TILE_CAP 1152
#define CRK_TILE_MAXRUNS 40
#define CRK_TILE_IDS 20480
#define CRK_TILE_WORDS (4 + 2 * (CRK_TILE_MAXRUNS + 1))
/* LDS tiles of a filled adjacency (no counterpart in the reference; it serves
 * the tiled pair passes).  keys [n] (sorted order), cell_start/end as built
 * by crk_adjacency_count; cap <= CRK_TILE_CAP particles per tile (smaller
 * values make more tiles wide).  Writes tiles and nbr_loc (see CrkAdjacency);
 * asynchronous. */
int crk_adjacency_tiles(CrkContext* ctx, const CrkAdjacency* adj, const uint32_t* keys,
                        const int32_t* cell_start, const int32_t* cell_end, int32_t cap,
                        int32_t* tiles, uint16_t* nbr_loc, void* stream);

typedef struct CrkState {          /* user order, caller owned */
  int32_t n;
  float* positions;   /* [n*dim] */
  float* velocities;  /* [n*dim] */
  float* supports;    /* [n] */
  float* masses;
  float* densities;
  float* internalEnergies;
  float* entropies;
  float* pressures;
  float* soundspeeds;
  float* divergence;
  float* alpha0s;
  float* alphas;
} CrkState;

typedef struct CrkUpdate {         /* user order, caller owned: CompressibleSystemUpdate */
  float* dvdt;    /* [n*dim] */
  float* dudt;    /* [n] */
  float* drhodt;  /* [n] */
  float* dEdt;    /* [n] */
} CrkUpdate;

typedef struct CrkEdges {          /* one stage's interacting pairs, caller owned */
  int32_t* count; /* [n]     pairs of sorted row k with |x_ij| < max(h_i, h_j) */
  int32_t* nbr;   /* [E]     sorted neighbour j of each pair: row k's pairs fill the
                             slots row_offsets[k] .. row_offsets[k] + count[k] - 1
                             (adjacency row order kept); the other slots are unused */
  float*   pf;    /* [E*dim] a_ij = ap_ij + av_ij of each slot, the pairwise
                             acceleration (what the compatible-energy finalize
                             and the energy partition read)                    */
  float*   rec;   /* [n*4]   energy-partition record of sorted particle k:
                             v_half = v + dt/2 dvdt (dim components, zero padded
                             to 3) and s = P/rho^gamma (crk_balance)           */
  float*   f;     /* [E]     energy partition f_ij in the pair slots; optional:
                             NULL keeps it implicit (evaluated from rec where
                             it is used)                                       */
  uint16_t* loc;  /* [E]     tile-local index of each pair's neighbour (written
                             by crk_support when the adjacency has tiles)      */
  float*   th;    /* [E]     viscous fraction of each slot's pair force:
                             av_ij = th a_ij, ap_ij = (1 - th) a_ij (ap and av
                             are parallel, both along the pair's CRK
                             gradient, and of one sign: th in [0, 1])         */
} CrkEdges;

/* Gathers the state into sorted order (x, v, h, m, u, alpha0, divergence). */
int crk_stage_begin(CrkContext* ctx, const CrkConfig* cfg, const CrkAdjacency* adj,
                    const CrkState* st, void* stream);
/* crk_stage_begin fused with the Verlet validity test of the adjacency for the
 * current state (the test of crk_adjacency_valid, with xb, hb the positions
 * and supports the adjacency was built at) and with the support range
 * (hminmax_dev[0] = min h, [1] = max h, for crk_owen_iteration; per run
 * when runs are bound).  Blocking:
 * *valid = 1 when the adjacency holds every pair within the Verlet radius.
 * When *valid = 0 the caller rebuilds and calls this again.  After it,
 * crk_owen_iteration also records max h_new / h_b for crk_owen_valid.     */
int crk_stage_begin_checked(CrkContext* ctx, const CrkConfig* cfg, const CrkAdjacency* adj,
                            const CrkState* st, const float* xb, const float* hb,
                            float* hminmax_dev, int32_t* valid, void* stream);
/* Verlet validity after the Owen update of a checked stage (positions are
 * unchanged, so only the supports moved).  Blocking.                       */
int crk_owen_valid(CrkContext* ctx, const CrkConfig* cfg, int32_t* valid, void* stream);
/* Re-bind to a new adjacency mid-stage (after a rebuild), keeping h. */
int crk_stage_rebind(CrkContext* ctx, const CrkConfig* cfg, const CrkAdjacency* adj,
                     const CrkState* st, void* stream);
/* Boundary-condition hooks (modules/boundaryConditions/bcs.py:16-92) run on
 * the caller's user-order tensors between passes; these two calls hand their
 * changes back to the stage.
 * crk_state_refresh: after crk_factors_eos and enforceDirichlet
 * (crkSPH.py:94-101): re-reads the state (positions, velocities, supports,
 * masses, energies, alphas, divergence, densities), evaluates the
 * equation of state (idealGas.py:6-38) from the densities and energies,
 * writing st->entropies, pressures and soundspeeds, and (with the switch)
 * the signal speed from the new velocities and sound speeds.
 * crk_update_refresh: after crk_accel_dudt and computeForcing /
 * enforceUpdates (crkSPH.py:327-339): re-reads update->dvdt for the energy
 * partition (crk_balance).                                                 */
int crk_state_refresh(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st, void* stream);
int crk_update_refresh(CrkContext* ctx, const CrkUpdate* up, void* stream);
/* Owen adaptive support, one iteration (optimalSupportOwen.py:70-114):
 * psi0 pass over the bound adjacency, LUT inversion, computeNewSupport,
 * clamp to [hMin/4, 4 hMax].  hmin/hmax are the bounds of the supports at
 * the start of the Owen loop (device scalars, see crk_minmax_h).  Writes the
 * new supports (sorted copy and st->supports) and, when max_ratio_dev is not
 * NULL, max |h_new/h - 1| to it (device float).                            */
int crk_owen_iteration(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st,
                       const float* hminmax_dev, float* max_ratio_dev, void* stream);
/* The last Owen iteration of a stage fused with the per-row parts of
 * crk_support: after the new h_i, a second sweep over the same row computes
 * the CRK volume V_i = 1 / sum_j W(x_ij, h_i) and (Cullen-Dehnen switch) the
 * signal speed with the new h_i.  Follow it, when the adjacency stays valid
 * for the new supports, with crk_pairs_factors_eos (which compacts the pairs);
 * otherwise rebuild, crk_stage_begin and crk_support as usual.  Same results
 * as crk_owen_iteration + crk_support.                                      */
int crk_owen_support(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st,
                     const float* hminmax_dev, float* max_ratio_dev, void* stream);
/* Monaghan adaptive support, one Newton iteration
 * (evaluateOptimalSupportMonaghan, optimalSupportMonaghan.py:20-108, with
 * computeOmega, wp_omega.py:52-73, 143): rho_i = sum_j m_j W(x_ij, h_i),
 * F = h - volumeToSupportHelper(m/rho, n_H, d), Omega = 1 + h/(rho d)
 * sum_j m_j dW/dh, h' = h - F/(Omega + 1e-6), clamped to [hMin/4, 4 hMax].
 * Same arguments and outputs as crk_owen_iteration (max_ratio is
 * max |h'/(h + 1e-6) - 1|); crk_owen_valid tests the list afterwards.       */
int crk_monaghan_iteration(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st,
                           const float* hminmax_dev, float* max_ratio_dev, void* stream);
/* Density by gather over the bound adjacency, rho_i = sum_j m_j W(x_ij, h_i)
 * (warpOperation Density, SupportScheme.Gather; the final pass of
 * optimalSupportOwen.py:116-125, overwritten in the step and therefore not
 * run there; used by the initial-condition generators).  User order. */
int crk_density_gather(CrkContext* ctx, const CrkConfig* cfg, float* rho_user, void* stream);
/* hminmax_dev[0] = min supports, [1] = max supports (user array); per run
 * ([2r], [2r+1]) when runs are bound. */
int crk_minmax(CrkContext* ctx, int32_t n, const float* values, float* minmax_dev, void* stream);
/* Support sweep of the bound adjacency for the current positions and supports
 * (after the Owen update, before crk_factors_eos).  One pass over every edge:
 *  - the CRK apparent volume V_i = 1 / sum_j W(x_ij, h_i) (computeCRKFactors);
 *  - the pairs with |x_ij| < max(h_i, h_j) are compacted per row (row order
 *    kept) into edges->count / edges->nbr; every later pass of the stage, and
 *    the pair data (pf: ap, av; f) use these slots.  The other (Verlet skin) edges
 *    have ap = av = 0 (both CRK gradients vanish there) and are not stored;
 *    crk_edges_export expands to the full adjacency order;
 *  - with the Cullen-Dehnen switch, the signal speed
 *    max_j (c_ij - mu_ij) (wp_vsig.py:11-86) over every adjacency entry.
 * The context keeps edges->count and edges->nbr for the passes that follow. */
int crk_support(CrkContext* ctx, const CrkConfig* cfg, CrkEdges* edges, void* stream);
/* computeCRKFactors (warpSPHCore; called at crkSPH.py:61), after crk_support:
 * moments with FP64 solve -> A, B, gradA, gradB, CRK density (written to
 * st->densities), followed by idealGasEOS (eos/idealGas.py:6-38) writing
 * st->entropies, st->pressures, st->soundspeeds.  n_singular_dev (device
 * int, may be NULL) receives the number of particles that used the
 * moment-matrix fallback.                                                   */
int crk_factors_eos(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st,
                    int32_t* n_singular_dev, void* stream);
/* crk_factors_eos fused with crk_support's compaction: one sweep over the
 * full Verlet rows compacts the pairs with |x_ij| < max(h_i, h_j) into
 * edges->count / edges->nbr (as crk_support does) while accumulating the CRK
 * moments; V must be current (crk_owen_support or crk_support on the bound
 * adjacency).  The context keeps edges->count and edges->nbr for the passes
 * that follow.                                                              */
int crk_pairs_factors_eos(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st,
                          CrkEdges* edges, int32_t* n_singular_dev, void* stream);
/* Velocity gradient (crkSPH.py:158-172): J_i = sum_j V_j (v_j - v_i) (x) gradW^R_ij,
 * drhodt = -tr(J) rho.  Writes update->drhodt. */
int crk_velocity_gradient(CrkContext* ctx, const CrkConfig* cfg, CrkUpdate* up, void* stream);
/* Cullen-Dehnen switch (CullenDehnen2010.py:218-331 + wp_vsig.py:11-86):
 * writes st->alphas (and st->alpha0s when cfg->evolveAlpha0). */
int crk_switch(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st, float dt, void* stream);
/* computeCrkSPHAccelWarp + computeCrkSPHdudtWarp fused (crk/accel.py:15-250,
 * crk/dudt.py:15-229).  Writes update->dvdt, update->dudt, update->dEdt,
 * edges->pf (ap_ij, av_ij per pair slot), and st->divergence = -drhodt/rho
 * (crkSPH.py:324). */
int crk_accel_dudt(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st,
                   CrkUpdate* up, CrkEdges* edges, void* stream);
/* crk_switch followed by crk_accel_dudt in one sweep over the pairs: the
 * switch's kernel sum (CullenDehnen2010.py:218-331) is accumulated while the
 * accelerations are computed.  Same results as the two calls; only valid
 * when the viscosity does not read alpha (cfg->alphaWeightedQ == 0), since
 * the reference evaluates the switch first (crkSPH.py:195-200 before
 * :241-284).  Writes everything both calls write.                          */
int crk_switch_accel_dudt(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st,
                          CrkUpdate* up, CrkEdges* edges, float dt, void* stream);
/* computeCompSPHBalanceTermWarp, CRK energy scheme (compSPH/balance.py:63-154)
 * with v_half = v + dt/2 dvdt (crkSPH.py:341).  Writes edges->rec; f_ij is a
 * function of rec_i, rec_j, m_i dt and ap + av of the pair, evaluated by
 * crk_compatible_energy and crk_edges_export.  When edges->f is not NULL it
 * is also written for every pair slot.                                      */
int crk_balance(CrkContext* ctx, const CrkConfig* cfg, float dt, CrkEdges* edges, void* stream);
/* Launch shape: pass ids for crk_set_group_size.  G consecutive lanes share
 * one particle row (G in 1, 2, 4, ..., 64); chosen by the autotuner. */
enum {
  CRK_PASS_OWEN = 0, CRK_PASS_VOLUME = 1, CRK_PASS_MOMENTS = 2, CRK_PASS_VELGRAD = 3,
  CRK_PASS_SWITCH = 4, CRK_PASS_ACCEL = 5, CRK_PASS_BALANCE = 6, CRK_PASS_COMPAT = 7,
  CRK_PASS_SUPPORT = 8, CRK_NPASS = 9
};
int crk_set_group_size(CrkContext* ctx, int32_t pass, int32_t group);
/* Threads per block of a pass that runs on LDS tiles (one block per tile of
 * CRK_TILE_ROWS rows): 256, 512 or 1024; chosen by the autotuner with G.
 * Launch configuration only (no reference counterpart: the reference's Warp
 * launches one thread per particle, accel.py:477-502). */
int crk_set_block_size(CrkContext* ctx, int32_t pass, int32_t threads);

/* Pair data of one stage in the reference layout (ap_ij, av_ij, f_ij over the
 * adjacency edges, src/warpSPH/systems/compSPH.py:30-32): slot data of
 * *edges scattered to adjacency order, zero on the pairs out of support.
 * f_ij is read from edges->f when stored, else evaluated from edges->rec with
 * the (user-order) masses and the stage's dt.  Any output may be NULL.     */
int crk_edges_export(CrkContext* ctx, int32_t dim, const CrkAdjacency* adj, const CrkEdges* edges,
                     const float* masses, float dt, float* ap_out, float* av_out, float* f_out,
                     void* stream);

/* An intermediate of the bound stage in user order (for tests and
 * diagnostics): 0 V, 1 A, 2 B, 3 grad A, 4 grad B, 5 grad v, 6 Cullen-Dehnen
 * div, 7 v_sig, 8 R.  CRK_ESTATE for a field the stage has not computed
 * (6-8 exist only with the Cullen-Dehnen or learned switch).              */
int crk_export_field(CrkContext* ctx, int32_t field, float* out_user, void* stream);

/* ----------------------------------------- learned viscosity switch (MLP)
 * No counterpart in the reference: a per-particle MLP 8 -> 64 -> 64 -> 1
 * (SiLU, SiLU, sigmoid) gives the viscosity alpha in [0, 1] in place of the
 * Cullen-Dehnen switch (CullenDehnen2010.py:218-331), which stays the
 * fallback for particles whose features leave the training range.  BF16
 * weights and activations on the matrix cores, FP32 accumulation.
 * Parameters (FP32): W1[64][8] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3.  */
#define CRK_MLP_IN 8
#define CRK_MLP_HIDDEN 64
#define CRK_MLP_NPARAM (64 * 8 + 64 + 64 * 64 + 64 + 64 + 1)
#define CRK_MLP_NBF16 (64 * 16 + 64 * 64 + 64)
typedef struct CrkMlp {
  const float*    params;  /* device [CRK_MLP_NPARAM] */
  const uint16_t* wbf16;   /* device [CRK_MLP_NBF16], from crk_mlp_pack */
  float mu[8], sigma[8];   /* input normalisation x' = (f - mu) / sigma */
  float lo[8], hi[8];      /* raw feature range of the training data */
  /* the switch's alpha from the network output s = sigmoid(MLP(x')):
   * alpha = scale * s, times the Cullen-Dehnen alpha of the particle when
   * multiply != 0 (the network then rescales the plain switch; s = 1/scale
   * reproduces it) */
  float   scale;
  int32_t multiply;
} CrkMlp;
/* BF16 operand copy of the parameters. */
int crk_mlp_pack(const float* params, uint16_t* wbf16, void* stream);
/* alpha[p] = sigmoid(MLP(x[p])), x [n][8] normalised features. */
int crk_mlp_forward(const CrkMlp* model, int32_t n, const float* x, float* alpha, void* stream);
/* Gradient of mean_p (alpha_p - target_p)^2 over the batch into grad
 * [CRK_MLP_NPARAM + 1] (last entry: the loss).  Deterministic. */
int crk_mlp_grad(CrkContext* ctx, const CrkMlp* model, int32_t n, const float* x, const float* target,
                 float* grad, void* stream);
/* Raw features of the bound stage (after crk_velocity_gradient), user order
 * [n][8]: h div/c, h^2 (div - div_prev)/(dt c^2), h |curl v|/c, h (rho/m)^(1/d),
 * c/v_sig, h |grad rho|/rho, h |grad P|/P, neighbours/n_H (div: the
 * Cullen-Dehnen divergence).  Needs a Cullen-Dehnen or learned switch.     */
int crk_learned_features(CrkContext* ctx, const CrkConfig* cfg, float dt, float* features_out,
                         void* stream);
/* The learned switch of the bound stage (in place of crk_switch): the
 * Cullen-Dehnen switch, then alpha = scale sigmoid(MLP(features)) (times the
 * Cullen-Dehnen alpha when model->multiply) where every feature lies in
 * [lo, hi], the Cullen-Dehnen alpha elsewhere.  Writes st->alphas and the
 * stage's alpha (read by crk_accel_dudt with alpha-weighted Q).            */
int crk_learned_switch(CrkContext* ctx, const CrkConfig* cfg, const CrkState* st, float dt,
                       const CrkMlp* model, void* stream);
/* One Adam step (bias-corrected, step >= 1) on the FP32 parameters. */
int crk_mlp_adam(int32_t n_param, float* params, float* m1, float* m2, const float* grad, float lr,
                 float beta1, float beta2, float eps, int32_t step, void* stream);

/* ------------------------------------------------ integrator helpers
 * Heun RK2 (b = 1/2, 1/2).  All user order.                                 */
/* y = y0 + dt * d, elementwise over n*components (stage update). */
int crk_axpy(int64_t count, const float* y0, const float* d, float dt, float* y, void* stream);
/* y = y0 + dt * (b0*d0 + b1*d1) */
int crk_rk_combine(int64_t count, const float* y0, const float* d0, const float* d1,
                   float b0, float b1, float dt, float* y, void* stream);
/* The same with each particle's step dtp[n] (n particles of comps values). */
int crk_axpy_p(int64_t n, int32_t comps, const float* y0, const float* d, const float* dtp,
               float* y, void* stream);
int crk_rk_combine_p(int64_t n, int32_t comps, const float* y0, const float* d0, const float* d1,
                     float b0, float b1, const float* dtp, float* y, void* stream);
/* Wrap positions into the periodic domain. */
int crk_wrap_positions(const CrkConfig* cfg, int32_t n, float* x, void* stream);
/* Compatible-energy finalize (compSPH/multistep.py:52-94, systems/compSPH.py:87-129):
 * v_half = v0 + dt/2 (b0 dvdt0 + b1 dvdt1);
 * du_i = sum_k b_k sum_{pairs e of row i in stage k} f_e (v_half_j - v_half_i).(ap_e + av_e);
 * u = u0 + dt du; dudt_last = (u - u0)/dt.  Deterministic row sums.  f_e is
 * evaluated from the stage's records (crk_balance) with m_i dt; masses are
 * in user order.                                                           */
int crk_compatible_energy(CrkContext* ctx, const CrkConfig* cfg, int32_t n, float dt,
                          const float* masses, const float* v0, const float* u0,
                          const float* dvdt0, const float* dvdt1, float b0, float b1,
                          const CrkAdjacency* adj0, const CrkEdges* e0,
                          const CrkAdjacency* adj1, const CrkEdges* e1,
                          float* u_out, float* dudt_last, void* stream);

/* ------------------------------------------------ (c) reductions
 * computeTimestep (timestep/compressible.py:12-66): blocking, returns dt.
 * prev_dt <= 0 means no growth limit. */
int crk_cfl_dt(CrkContext* ctx, const CrkConfig* cfg, int32_t n, const float* supports,
               const float* internalEnergies, double prev_dt, double* dt_out, void* stream);
/* FP64 conservation sums: out[0] = sum m, out[1..dim] = sum m v,
 * out[dim+1] = sum m u, out[dim+2] = sum m |v|^2 / 2.  Blocking. */
int crk_conservation(CrkContext* ctx, int32_t dim, int32_t n, const float* masses,
                     const float* velocities, const float* internalEnergies,
                     double* out, void* stream);
/* crk_conservation without any allocation and without host synchronisation:
 * the same sums (bit-identical) written to out_dev, a device array of
 * dim + 3 doubles.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops).  Its scratch is one fixed file-scope device array per GPU
 * (n <= 2^23 = 8,388,608 particles, 1.5 MiB, held until the process ends);
 * larger n returns CRK_EINVAL.  Asynchronous on `stream`; calls that overlap
 * on one GPU share that array and race, so serialise them. */
int crk_conservation_noalloc(CrkContext* ctx, int32_t dim, int32_t n, const float* masses,
                             const float* velocities, const float* internalEnergies,
                             double* out_dev, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* CRKSPH_H */
