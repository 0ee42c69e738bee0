// Device runtime of the MI355X-native explicit step: HBM-resident scenario state, the
// HIP kernels of one step (forcing, momentum + CFL reduction, combined-uv averaging,
// continuity), graph-captured step loops, batched multi-scenario launches and the
// C ABI entry points of include/sfx.h that need a GPU.
//
// Layouts in HBM (structure of arrays, one allocation per field):
//  * cells: reference order (level-major, then column-major n fastest); zs and z_volume
//    in double, all other cell fields in R (float, or double for the FP64 twin).
//  * uv-points, "dense" scenarios (full rectangles): two planes of nmax*mmax slots,
//    slot c = u-point right of cell c, slot np+c = v-point above cell c, plus one dummy
//    slot (always zero) -- neighbours are implicit (+-1, +-nmax), no index tables.
//  * uv-points, "indexed" scenarios (quadtree, masked grids): reference order with the
//    ncuv combined points and the dummy, neighbour tables converted to 0-based slots.
//  * q and uv are double-buffered: a momentum sweep reads buffer `cur` and writes
//    `cur^1`, which replaces the reference's q0/uv0 copy (sfincs_momentum.f90:125-138).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "sfx_internal.hpp"
#include "sfx_physics.hpp"

using sfx::Consts;
using sfx::set_error;

namespace sfx {
bool pow_tables_device(int device, PowTables& out, std::string& err);
}

namespace {

#define HIPCHK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) {                                                                \
      set_error(std::string(#x) + ": " + hipGetErrorString(e_));                           \
      return SFX_ERR_HIP;                                                                  \
    }                                                                                      \
  } while (0)

constexpr int BLOCK = 256;
constexpr int WAVE = 64;

// Device arrays are reached through pointers stored in DevInst; used as generic pointers
// every access becomes a flat load with a 64-bit per-lane address, which the compiler
// must wait for one group at a time. G() wraps the pointer as a global-address-space
// array indexed by 32-bit element numbers (sfx::GA), so accesses become global loads and
// stores with a scalar base that can all be in flight together.
using sfx::GA;
using sfx::ldf4;
template <typename T> __device__ __forceinline__ GA<T> G(T* p) { return sfx::ga(p); }

// Device-resident step scalars of one scenario (t, dt, CFL reduction, forcing scalars).
struct Scalars {
  double t, tend;
  float dt, min_dt, dtchk;
  int32_t nt, active, error, error_pending, cur;
  int32_t mom_done;              // a momentum sweep wrote buffer cur^1 and its per-tile minima
  int32_t itbndlast, itsrclast, itwndlast, itprcplast;
  float tauwu, tauwv, netprcp;   // uniform meteo forcing of this step
  float factime, facrel, one_minus_factime;
  // gridded meteo: the next update time, the current window [meteo_t0, meteo_t1] and its
  // frame mwin, the step's time weight mtw (and 1 - mtw), the spin-up factor msm (mspin)
  float twindupd, meteo_t0, meteo_t1, mtw, momtw, msm;
  int32_t mwin, mspin;
  int32_t ahead;                 // the next step's fold, scalars and series are staged (StepImage)
  int32_t pad_;
};

// Step scalars and forcing series of one scenario as a forcing block works on them: the
// Scalars as they will be after the step's fold and forcing (s), whether the step runs
// (go), whether the fold consumed a momentum sweep (folded), and the step's position in
// the boundary series (itb0, itb1, tbfac).
struct StepImage {
  Scalars s;
  int32_t go, folded, itb0, itb1;
  float tbfac;
  int32_t pad_;
};
static_assert(sizeof(Scalars) % 4 == 0 && sizeof(StepImage) % 4 == 0, "copied as words");

struct DevInst {
  Consts c;
  int32_t dense, subgrid, fp64;
  int32_t np, npuv, ncuv, nslot, nmax, mmax, nl;
  int32_t tn, tm, ntn;           // dense continuity tile shape and tiles along n
  int32_t tn_sh, mtn_sh;         // log2 of tn and mtn (tiles hold 256 cells, so both are powers of two)
  int32_t sk, nstrip;            // dense momentum kernel: strip march with sk rows per wave (sk > 0), per-cell (0), per-point with -sk points per thread (-1, -2, -4)
  int32_t mtn, mtm, mntn;        // dense momentum, per-cell kernel: tile shape and tiles along n
  int32_t xchunk, xchunk_c;      // momentum / continuity work-list runs per XCD (0: one contiguous band per XCD)
  int32_t count;                 // per-branch counting on
  int32_t nmt;                   // momentum tiles of the current tile shape
  int32_t mrow_sh;               // per-point kernel: log2 of the cell rows of a tile (mtm * P)
  // Quiescent per-point tiles: mq[tile] holds one byte per wave, 1 when all the wave's points
  // were dry with their zeros held at the tile's last run; md[tile] is raised when a cell
  // level the tile reads may have changed since (continuity, forcing; host writes reset
  // mq). A tile with mq = 0x01010101 and md = 0 would do nothing and is skipped.
  uint32_t* mq;
  uint8_t* md;
  Scalars* sc;
  StepImage* snext;              // the next step computed ahead (valid while sc->ahead)
  float* qtsrc_nx;               // its discharges
  // CFL reduction: each wave of a momentum tile lowers blkmin[tile] to the minimum of its
  // points' time steps (an atomic min on the tile's own word, so no two tiles share an
  // address and no wave waits for the others of its block); the fold at the start of the
  // next step reduces the tiles and sets every word back to dtmax.
  float* blkmin;
  uint32_t* blkcnt;              // per tile, NCNT words: uv-points copied (kcuv not 1/6), dry, wet/full, partial, dry with zeros held, skipped
  uint32_t* cntstore;            // per continuity tile, NSTORE words: cells that stored zs, V, zs0, zsderv, zsmax; steady cells skipped
  // cells
  double* zs;
  double* vol;
  void* zs0;
  void* zsderv;
  void* zsmax;
  void* vmax;                    // compute_store_variables fields, allocated when switched on
  void* qmax;
  void* twet;
  // timestep_analysis, per uv slot (allocated when switched on): the required time step of
  // the last sweep (wet points), its running sum, the sweeps wet and those limiting
  float* tsa_req; float* tsa_sum; int32_t* tsa_wet; int32_t* tsa_lim;
  float* tsa_min;                // the last sweep's minimum time step
  float* qext;                   // external source per cell, m/s (use_qext)
  const float* qinf;             // infiltration rate per cell, m/s (qinffield; F_INFIL), or S in m (F_CNA)
  float* cumprcp; float* cuminf; // cumulative rain and infiltration per cell, m (F_CNA)
  // gridded meteo fields at the update times, [mnframes][np]: wind stress (F_MWIND),
  // atmospheric pressure (F_PATM), rain in m/s (F_MPRCP)
  const float* mtauu; const float* mtauv; const float* mpatm; const float* mprcp;
  int32_t mnframes;
  // FP64 volume balance of the interior cells (kcs 1). Each step adds to bal[0..3] the
  // volume that entered them through faces whose other side is not interior (boundary,
  // Neumann and inactive cells; summed by one block of the continuity launch), from point
  // sources as applied (the regular branch's bed clamp included), rain and qext; bal[4..5]
  // receive the interior volume and the storage left when they are summed; bal[6..6+NINF)
  // the uniform rain volume of interior cells less what they received (infiltration and
  // gridded rain), one slot per wave modulo NINF.
  double* bal;
  int32_t nbf;                   // faces read by an interior cell whose other side is not interior
  const int32_t* bf_slot;        // their q slots
  const double* bf_coef;         // volume per unit q and dt as the interior cell counts it (+- width)
  double rain_area;              // total area of the cells rain falls on (kcs 1)
  void* storage;                 // storage volume left per subgrid cell, m3 (use_storage_volume)
  uint8_t* kcs;                  // kcs (KCS_MASK) | KCS_SOURCE (point source) | KCS_STEADY (continuity)
  const float* zb;
  const int8_t* z_iref;
  const int32_t* zmd; const int32_t* zmu; const int32_t* znd; const int32_t* znu;
  const float* z_zmin; const float* z_zmax; const float* z_volmax; const float* z_dep;
  // uv slots
  void* q[2];
  void* uv[2];
  int8_t* kfuv;                  // bit 0: kfuv of the reference; bits 1-2: zero flags of the per-point kernel (KF_*)
  const int8_t* kcuv;
  const int8_t* mask_adv;
  const float* zbuvmx;
  const float* gn2uv;
  const uint8_t* uvf;            // indexed: iref | (type+1)<<4 | dir<<6
  const int32_t* nb;             // indexed: 10 tables of npuv (nm,nmu,u_nmd,u_nmu,u_ndm,u_num,v_ndm,v_nm,v_ndmu,v_nmu)
  const int32_t* cuv;            // indexed: 3 tables of ncuv (cuv, uv1, uv2)
  const float* uv_zmin;
  const float4* uv_s4;           // {zmax, havg_zmax, nrep_zmax, navg_w}
  const float* uv_fnfit;
  const float4* uv_lev;          // nslot*nl {havg, nrep, pwet, 0}
  // boundaries
  int32_t ngbnd, nkcuv2, nbnd, ntbnd;
  const int32_t* gb_cell; const int32_t* gb_ind1; const int32_t* gb_ind2; const float* gb_fac; const int32_t* gb_nmi;
  const int32_t* gb_rcell; const float* gb_rdzs;   // river cells (kcs 5): interior cell and level offset
  const float* zsi_bnd;                            // infragravity levels as zs_bnd (null: none)
  float* zsb; float* zsb0;
  const int32_t* k2_slot; const int32_t* k2_nmb; const int32_t* k2_nmi; const int32_t* k2_ib; const int8_t* k2_dir;
  const float* k2_zbuv;
  void* uvmean;
  const float* t_bnd; const float* zs_bnd;
  int32_t bnd_overlap;           // a Neumann copy and a boundary flux point share a cell
  int32_t n6;                    // Neumann boundary cells (kcs 6) ...
  const int32_t* gb6;            // ... as pairs {cell, interior cell}
  int32_t nflux;                 // kcuv 2/6 points whose boundary cell is not a Neumann cell ...
  const int32_t* k2_flux;        // ... as indices into the k2_* tables
  // sources and meteo; src_cell and qtsrc hold nsrc point sources, then each drainage
  // structure's intake and outfall
  int32_t nsrc, ntsrc, ntprcp, ntwnd;
  int32_t ndrn;                  // drainage structures
  float structure_relax;
  const int32_t* drn_type; const float* drn_par; const float* drn_dist;   // per structure: type, 6 params, distance
  int32_t* drn_status; float* drn_frac;    // gate state: 0 closed, 1 open, 2 closing, 3 opening; fraction open
  const int32_t* src_cell; const float* tsrc; const float* qsrc; float* qtsrc;
  const float* tprcpt; const float* tprcpv; const float* twnd; const float* wndmag; const float* wnddir;
  const float* cdval; const float* x73;
  sfx::PowTables pw;
  // weirs: per uv slot the structure index (-1: none), per structure {crest, Cd, length, 0};
  // null when the scenario has none
  const int32_t* struc_of; const float4* struc_par;
};

// Flags of a dry uv-point kept in its kfuv byte by the per-point and indexed momentum
// kernels (bit 0 is the reference's kfuv and every reader masks it): KF_ZERO1, the kernel
// stored q = uv = 0 at the point in the last step; KF_ZERO2, it did so in two consecutive
// steps, which write the two ping-pong buffers, so both hold zeros and a dry point needs
// no store. Every other write of kfuv (wet points, the other kernels, the host) stores 0
// or 1 and clears them.
constexpr int8_t KF_ZERO2 = 2, KF_ZERO1 = 4;

// The cell byte: kcs in the low bits, KCS_SOURCE when the cell holds a point source,
// KCS_STEADY kept by continuity (see cont_update).
constexpr uint8_t KCS_MASK = 0x3f, KCS_STEADY = 0x40, KCS_SOURCE = 0x80;

// Per-branch uv-point counters of a momentum tile: copied (kcuv not 1 or 6), dry with zeros
// stored, wet (regular) or full (subgrid), partial (subgrid), dry with zeros already held,
// points of a quiescent tile that was skipped.
constexpr int NCNT = 6;
// Continuity counted per tile: cells that stored zs, V, zs0, zsderv, zsmax (a cell stores a
// field only when its value changes), and steady cells skipped.
constexpr int NSTORE = 6;
// Volume balance: six sums, then NINF slots for the infiltrated volume (waves add to slot
// wave number mod NINF, so their double atomics spread over many addresses).
constexpr int NINF = 256, NBAL = 6 + NINF;

struct Item { int32_t inst, tile; };

// Work list of one launch. A direct list holds the tiles 0..n-1 of one scenario in the
// XCD order that xcd_order builds on the host: block b runs on XCD b mod 8, and runs of
// 2^run_sh consecutive tiles are dealt to the XCDs in turn (run_sh < 0: one contiguous
// band per XCD). Its blocks compute their tile from blockIdx, the kernel's descriptor
// argument points at that scenario's descriptor, and the step scalars come from the
// launch arguments, so a block's first memory reads are the descriptor fields and the
// step scalars, together. Other lists read one Item per block. The descriptor is always
// reached from the kernel's restrict-qualified argument, which lets the compiler keep
// its fields in scalar loads after the kernel's first stores.
struct WorkList {
  const Item* items;
  Scalars* sc;
  int32_t n, run_sh, direct;
  int32_t inst;                  // host side: the scenario of a direct list
};

// Dense tile shapes (cells along n x along m, BLOCK threads) of the per-cell momentum and
// continuity kernels, and strip-march lengths (rows per wave) of the strip momentum
// kernel, that the autotuner tries; the first of each is the default.
constexpr int CELL_SHAPES[][2] = {{64, 4}, {128, 2}, {256, 1}, {32, 8}, {16, 16}};
constexpr int N_CELL_SHAPES = sizeof(CELL_SHAPES) / sizeof(CELL_SHAPES[0]);

// log2 of a power of two (tile widths: tiles hold BLOCK = 256 cells).
inline int32_t ilog2(int32_t v) { int32_t k = 0; while ((1 << (k + 1)) <= v) ++k; return k; }
constexpr int STRIP_ROWS[] = {4, 8, 16};
constexpr int N_STRIP_ROWS = sizeof(STRIP_ROWS) / sizeof(STRIP_ROWS[0]);
constexpr int STRIP_COLS = WAVE - 2;             // columns owned by one wave

size_t strip_tiles(const DevInst& I, int sk) {
  const int nstrip = (I.nmax + STRIP_COLS - 1) / STRIP_COLS;
  const int rows = sk * (BLOCK / WAVE);
  return (size_t)nstrip * (size_t)((I.mmax + rows - 1) / rows);
}

size_t cell_tiles(const DevInst& I, int tn, int tm) {
  return (size_t)((I.nmax + tn - 1) / tn) * (size_t)((I.mmax + tm - 1) / tm);
}

// Momentum tiles of one scenario: per-cell tiles, per-point tile pairs (u and v points of
// mtn x (P mtm) cells, P = -sk) or strip-march blocks (dense), or BLOCK uv-points (indexed).
size_t mom_tiles(const DevInst& I) {
  if (!I.dense) return (size_t)((I.npuv + BLOCK - 1) / BLOCK);
  if (I.sk > 0) return strip_tiles(I, I.sk);
  return I.sk < 0 ? 2 * cell_tiles(I, I.mtn, I.mtm * -I.sk) : cell_tiles(I, I.mtn, I.mtm);
}

// Continuity tiles: tn x tm cells, or BLOCK cells (indexed).
size_t cont_tiles(const DevInst& I) {
  return I.dense ? cell_tiles(I, I.tn, I.tm) : (size_t)((I.np + BLOCK - 1) / BLOCK);
}

__device__ __forceinline__ float wave_min(float v) {
  for (int o = WAVE / 2; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, WAVE));
  return v;
}

// Minimum over the 64 lanes of a wave of positive floats, as their bit patterns (which
// order like the values), valid in lane 63 only; all lanes must be active. DPP row shifts
// reduce each row of 16 lanes into its last lane; row_bcast:15 then folds row 0 into
// row 1 and row 2 into row 3, row_bcast:31 lane 31 into rows 2-3. Lanes without a source
// keep their own value (bound_ctrl off, old = v), which the minimum ignores.
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_min(uint32_t v) {
  const uint32_t o = (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, CTRL, ROWS, 0xf, false);
  return o < v ? o : v;
}
__device__ __forceinline__ uint32_t wave_min_last(uint32_t v) {
  v = dpp_min<0x111, 0xf>(v);   // row_shr:1
  v = dpp_min<0x112, 0xf>(v);   // row_shr:2
  v = dpp_min<0x114, 0xf>(v);   // row_shr:4
  v = dpp_min<0x118, 0xf>(v);   // row_shr:8
  v = dpp_min<0x142, 0xa>(v);   // row_bcast:15
  v = dpp_min<0x143, 0xc>(v);   // row_bcast:31
  return v;
}

// Block-wide minimum over NT threads; the result is valid in every thread.
template <int NT = BLOCK>
__device__ __forceinline__ float block_min(float v) {
  __shared__ float red[NT / WAVE];
  v = wave_min(v);
  const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
  if (lane == 0) red[w] = v;
  __syncthreads();
  float m = red[0];
  for (int i = 1; i < NT / WAVE; ++i) m = fminf(m, red[i]);
  return m;
}

// Completes a momentum sweep (run by one whole block): min_dt = min(dtmax, per-tile
// minima) as in sfincs_momentum.f90:121,542-545, and the q/uv double buffer flips.
// Blocks of the per-scenario kernels (forcing, fold): one block covers a scenario's
// boundary points and tile minima.
constexpr int FBLOCK = 1024;

__device__ void fold_sweep(const DevInst& I) {
  const auto S = G(I.sc);
  __shared__ int sh_done;
  if (threadIdx.x == 0) sh_done = S->mom_done;
  __syncthreads();
  if (!sh_done) return;
  float v = I.c.dtmax;
  const auto p = G(I.blkmin);
  const int n = I.nmt;
  // chunks of 32 entries per thread, each chunk in one round trip (clamped indices
  // repeat the last entry, which leaves the minimum unchanged)
  for (int base = 0; base < n; base += 32 * FBLOCK) {
    float w[32];
#pragma unroll
    for (int j = 0; j < 32; ++j) w[j] = p[min(base + (int)threadIdx.x + j * FBLOCK, n - 1)];
#pragma unroll
    for (int j = 0; j < 32; ++j) v = fminf(v, w[j]);
#pragma unroll
    for (int j = 0; j < 32; ++j) {
      const int i = base + (int)threadIdx.x + j * FBLOCK;
      if (i < n) p[i] = I.c.dtmax;   // ready for the next sweep
    }
  }
  v = block_min<FBLOCK>(v);
  if (threadIdx.x == 0) { S->cur ^= 1; S->mom_done = 0; S->min_dt = v; }
  __syncthreads();
}

__global__ void __launch_bounds__(FBLOCK) k_fold(const DevInst* __restrict__ insts, const int32_t* __restrict__ ids) {
  const DevInst& I = insts[ids ? ids[blockIdx.x] : blockIdx.x];
  fold_sweep(I);
  if (threadIdx.x == 0) G(I.sc)->ahead = 0;   // a staged next step folded this sweep too
}

// Raises md of the per-point tiles (both planes) covering cells [nlo, nhi] x [mlo, mhi]
// and the row below and column to the left, whose points read these cells' levels.
__device__ __forceinline__ void mark_cells(const DevInst& I, int nlo, int nhi, int mlo, int mhi) {
  const int c0 = max(nlo - 1, 0) >> I.mtn_sh, c1 = min(nhi, I.nmax - 1) >> I.mtn_sh;
  const int r0 = max(mlo - 1, 0) >> I.mrow_sh, r1 = min(mhi, I.mmax - 1) >> I.mrow_sh;
  for (int r = r0; r <= r1; ++r)
    for (int c = c0; c <= c1; ++c) {
      const uint32_t t = 2u * (uint32_t)(r * I.mntn + c);
      G(I.md)[t] = 1;
      G(I.md)[t + 1] = 1;
    }
}

// A level set outside continuity (boundary cells in forcing).
__device__ __forceinline__ void mark_cell(const DevInst& I, int c) {
  if (!I.dense || I.sk >= 0) return;
  const int m = c / I.nmax, n = c - m * I.nmax;
  mark_cells(I, n, n, m, m);
}

// Continuity: when a level of this wave's cells (part of dense tile `tile`) changed, the
// first active lane marks the per-point tiles that read them.
__device__ __forceinline__ void mark_wave(const DevInst& I, int tile, bool changed) {
  const uint64_t act = __ballot(1);
  if (__ballot(changed) == 0) return;
  if ((int)(threadIdx.x % WAVE) != __ffsll((long long)act) - 1) return;
  const int tn = tile % I.ntn, tmi = tile / I.ntn;
  const int w0 = threadIdx.x & ~(WAVE - 1);
  const int nlo = tn * I.tn + (w0 & (I.tn - 1));
  const int nhi = nlo + min(WAVE, I.tn) - 1;
  const int mlo = tmi * I.tm + (w0 >> I.tn_sh);
  const int mhi = tmi * I.tm + ((w0 + WAVE - 1) >> I.tn_sh);
  mark_cells(I, nlo, nhi, mlo, mhi);
}

// ------------------------------------------------------------------------ gridded meteo
// The step's position in the gridded meteo: element offsets of the frames at the window's
// start (a) and end (b), the time weight w (omw = 1 - w) and the spin-up factor.
struct MeteoNow {
  uint32_t a, b;
  float w, omw, sm;
  int spin;
};
template <typename SP>
__device__ __forceinline__ MeteoNow meteo_now(const DevInst& I, SP S) {
  MeteoNow m;
  const int win = max((int)S->mwin, 0);
  m.a = (uint32_t)win * (uint32_t)I.np;
  m.b = m.a + (uint32_t)I.np;
  m.w = S->mtw; m.omw = S->momtw; m.sm = S->msm; m.spin = S->mspin;
  return m;
}
// update_meteo_forcing's pressure for one cell (sfincs_meteo.f90:1303-1307, 1362-1364): the
// value between its two frames, then the spin-up towards gapres.
__device__ __forceinline__ float meteo_patm(const DevInst& I, const MeteoNow& m, int c) {
  float v = G(I.mpatm)[m.a + (uint32_t)c] * m.omw + G(I.mpatm)[m.b + (uint32_t)c] * m.w;
  if (m.spin) v = v * m.sm + I.c.gapres * (1.0f - m.sm);
  return v;
}
// The meteo a wet uv-point reads (sfincs_momentum.f90:549-589): the wind stress of its cell
// nm, uniform (time series) or gridded, and the pressure difference between its two cells
// (real*4, as the reference forms it before scaling). Plain scalars rather than MeteoNow:
// the per-point kernels keep them in registers (64 VGPRs, 8 waves per SIMD).
template <typename R, typename SP>
__device__ __forceinline__ void uv_meteo(const DevInst& I, SP S, int nm, int nmu, sfx::UvIn<R>& in) {
  const uint32_t f = I.c.flags;
  if (!(f & sfx::F_MWIND)) { in.tauwu = (R)S->tauwu; in.tauwv = (R)S->tauwv; }
  if (!(f & (sfx::F_MWIND | sfx::F_PATM))) return;
  const uint32_t a = (uint32_t)max((int)S->mwin, 0) * (uint32_t)I.np, b = a + (uint32_t)I.np;
  const float w = S->mtw, omw = S->momtw, sm = S->msm;
  const bool spin = S->mspin != 0;
  if (f & sfx::F_MWIND) {
    float tu = G(I.mtauu)[a + (uint32_t)nm] * omw + G(I.mtauu)[b + (uint32_t)nm] * w;
    float tv = G(I.mtauv)[a + (uint32_t)nm] * omw + G(I.mtauv)[b + (uint32_t)nm] * w;
    if (spin) { tu = tu * sm; tv = tv * sm; }
    in.tauwu = (R)tu; in.tauwv = (R)tv;
  }
  if (f & sfx::F_PATM) {
    float p0 = G(I.mpatm)[a + (uint32_t)nm] * omw + G(I.mpatm)[b + (uint32_t)nm] * w;
    float p1 = G(I.mpatm)[a + (uint32_t)nmu] * omw + G(I.mpatm)[b + (uint32_t)nmu] * w;
    if (spin) { p0 = p0 * sm + I.c.gapres * (1.0f - sm); p1 = p1 * sm + I.c.gapres * (1.0f - sm); }
    in.dpatm = (R)(p0 - p1);
  }
}

// ----------------------------------------------------------------------------- forcing
// One block per scenario: step scalars, meteo series, boundary points/conditions/fluxes,
// discharge series (sfincs_lib.f90:379-550 with sfincs_meteo.f90:1423-1558,
// sfincs_boundaries.f90:682-1122, sfincs_discharges.f90:353-369).
__device__ float f_modulo(float a, float p) {
  float r = fmodf(a, p);
  if (r != 0.0f && ((r < 0.0f) != (p < 0.0f))) r += p;
  return r;
}

// Word-parallel copies of a scenario's scalars and step images (one word per thread).
template <typename T>
__device__ __forceinline__ void copy_words(T* dst, const T* src) {
  constexpr int NW = sizeof(T) / 4;
  if ((int)threadIdx.x < NW) ((uint32_t*)dst)[threadIdx.x] = ((const uint32_t*)src)[threadIdx.x];
}

// Fold of the tile minima into the image, leaving the minima in place (a step computed
// ahead of its turn leaves them to the launch that commits it, which resets them, and to
// a host-side fold, which re-reads them).
template <int NT>
__device__ void fold_image(const DevInst& I, StepImage& F) {
  if (F.s.mom_done) {
    float v = I.c.dtmax;
    const auto p = G(I.blkmin);
    const int n = I.nmt;
    for (int base = 0; base < n; base += 16 * NT) {
      float w[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) w[j] = p[min(base + (int)threadIdx.x + j * NT, n - 1)];
#pragma unroll
      for (int j = 0; j < 16; ++j) v = fminf(v, w[j]);
    }
    v = block_min<NT>(v);
    if (threadIdx.x == 0) { F.s.cur ^= 1; F.s.mom_done = 0; F.s.min_dt = v; F.folded = 1; }
  } else if (threadIdx.x == 0) {
    F.folded = 0;
  }
  __syncthreads();
}

__device__ void reset_tiles(const DevInst& I) {
  const auto p = G(I.blkmin);
  for (int i = threadIdx.x; i < I.nmt; i += blockDim.x) p[i] = I.c.dtmax;
}

// The step scalars (sfincs_lib.f90:379-392) and the forcing series at the new time
// (meteo, boundary series, discharges), on the image F after its fold; discharges go to
// qout. The series are skipped when the step does not run (F.go = 0).
__device__ void step_series(const DevInst& I, StepImage& F, float* qout, int mode) {
  const Consts& c = I.c;
  if (threadIdx.x == 0) {
    Scalars& sc = F.s;
    int go = 1;
    if (mode == 0) {
      if (sc.error_pending) { sc.error = 1; sc.error_pending = 0; sc.t = (double)c.t1 + 1.0; }
      go = (sc.t < sc.tend) && !sc.error;
      sc.active = go;
      if (go) {
        sc.nt += 1;
        sc.dt = c.alpha * sc.min_dt;
        sc.dtchk = c.alpha * sc.min_dt;
        sc.t = sc.t + (double)sc.dt;
        if (sc.dtchk < c.dtmin && sc.nt > 1) sc.error_pending = 1;
      }
    } else {
      sc.active = 1;
    }
    F.go = go;
  }
  __syncthreads();
  if (!F.go) return;
  // the forcing series at the new time, independent of each other: wave 0 time factors and
  // discharges, wave 1 wind, wave 2 rain, wave 3 boundary series (any block of 4+ waves)
  const double t = F.s.t;
  const float dt = F.s.dt;
  const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  if (wave == 0 && lane == 0) {
    F.s.factime = sfx::vmin<float>(dt / c.btfilter, 1.0f);
    F.s.one_minus_factime = 1.0f - F.s.factime;
    F.s.facrel = 1.0f - sfx::vmin<float>(dt / c.btrelax, 1.0f);
  }
  if (wave == 1 && lane == 0 && (c.flags & sfx::F_METEO3D)) {
    // gridded meteo: a new window at the first step with t >= twindupd (sfincs_lib.f90:493-503),
    // the step's time weight and spin-up factor (sfincs_meteo.f90:1270-1273, 1349-1352)
    if (t >= (double)F.s.twindupd) {
      F.s.meteo_t0 = F.s.twindupd;
      F.s.meteo_t1 = F.s.twindupd + c.dtwnd;
      F.s.twindupd = F.s.twindupd + c.dtwnd;
      F.s.mwin = min(F.s.mwin + 1, I.mnframes - 2);
    }
    const float tw = (float)((t - (double)F.s.meteo_t0) / (double)(F.s.meteo_t1 - F.s.meteo_t0));
    F.s.mtw = tw;
    F.s.momtw = 1.0f - tw;
    F.s.mspin = 0;
    F.s.msm = 1.0f;
    if ((c.flags & sfx::F_MSPIN) && t < (double)(c.tspinup - 1.0e-3f)) {
      F.s.msm = (float)((t - (double)c.t0) / (double)(c.tspinup - c.t0));
      F.s.mspin = 1;
    }
  }
  if (wave == 1 && lane == 0 && (c.flags & sfx::F_WIND) && I.ntwnd > 0) {
    for (int itw = max(F.s.itwndlast, 2); itw <= I.ntwnd; ++itw) {
      if ((double)G(I.twnd)[itw - 1] > t) {
        const float twfac = (float)((t - (double)G(I.twnd)[itw - 2]) / (double)(G(I.twnd)[itw - 1] - G(I.twnd)[itw - 2]));
        const float vmag = G(I.wndmag)[itw - 2] * (1.0f - twfac) + G(I.wndmag)[itw - 1] * twfac;
        float dr0 = f_modulo(G(I.wnddir)[itw - 2], 2 * c.pi);
        float dr1 = f_modulo(G(I.wnddir)[itw - 1], 2 * c.pi);
        if (dr1 > dr0 + c.pi) dr0 = dr0 + 2 * c.pi;
        else if (dr0 > dr1 + c.pi) dr1 = dr1 + 2 * c.pi;
        const float vdir = dr0 * (1.0f - twfac) + dr1 * twfac;
        const float cd = G(I.cdval)[min(max((int)(vmag * 10) + 1 - 1, 0), 999)];
        // cos and sin as the C library's cosf and sinf give them (exact tables)
        const float cv = sfx::libm_trig(vdir, 0, (const uint64_t*)I.pw.tcos, (const uint64_t*)I.pw.tsin);
        const float sv = sfx::libm_trig(vdir, 1, (const uint64_t*)I.pw.tcos, (const uint64_t*)I.pw.tsin);
        F.s.tauwu = vmag * vmag * cv * c.rhoa * cd / c.rhow;
        F.s.tauwv = vmag * vmag * sv * c.rhoa * cd / c.rhow;
        F.s.itwndlast = itw - 1;
        break;
      }
    }
  }
  if (wave == 2 && lane == 0 && (c.flags & sfx::F_PRECIP) && I.ntprcp > 0) {
    int itp;
    for (itp = F.s.itprcplast; itp <= I.ntprcp; ++itp)
      if ((double)G(I.tprcpt)[itp - 1] > t) break;
    if (itp > I.ntprcp) itp = I.ntprcp;
    if (itp < 2) itp = 2;   // series start at or before t0 (checked at initialisation)
    const float twfac = (float)((t - (double)G(I.tprcpt)[itp - 2]) / (double)(G(I.tprcpt)[itp - 1] - G(I.tprcpt)[itp - 2]));
    F.s.netprcp = (G(I.tprcpv)[itp - 2] * (1.0f - twfac) + G(I.tprcpv)[itp - 1] * twfac) / (float)(1000 * 3600);
    F.s.itprcplast = itp - 1;
  }
  if (wave == 3 && lane == 0 && I.nbnd > 0) {
    int itb0 = 1, itb1 = 1;
    float tb = 0.f;
    const int nt = I.ntbnd;
    if ((double)G(I.t_bnd)[0] > t - (double)1.0e-3f) { itb0 = 1; itb1 = 1; tb = G(I.t_bnd)[0]; }
    else if ((double)G(I.t_bnd)[nt - 1] < t + (double)1.0e-3f) { itb0 = nt; itb1 = nt; tb = G(I.t_bnd)[nt - 1]; }
    else {
      for (int itb = F.s.itbndlast; itb <= nt; ++itb)
        if ((double)G(I.t_bnd)[itb - 1] > t + (double)1.0e-6f) { itb0 = itb - 1; itb1 = itb; tb = (float)t; F.s.itbndlast = itb - 1; break; }
    }
    F.itb0 = itb0; F.itb1 = itb1;
    F.tbfac = (tb - G(I.t_bnd)[itb0 - 1]) / sfx::vmax<float>(G(I.t_bnd)[itb1 - 1] - G(I.t_bnd)[itb0 - 1], 1.0e-6f);
  }
  if (wave == 0 && I.nsrc > 0) {   // discharges: the series position once, the sources one lane each
    int itsel = 0;
    if (lane == 0) {
      for (int it = max(F.s.itsrclast, 2); it <= I.ntsrc; ++it)
        if ((double)G(I.tsrc)[it - 1] > t) { itsel = it; F.s.itsrclast = it - 1; break; }
    }
    itsel = __shfl(itsel, 0, WAVE);
    if (itsel) {
      const int it = itsel;
      for (int src = lane; src < I.nsrc; src += WAVE) {
        const float q0v = G(I.qsrc)[(it - 2) * I.nsrc + src], q1v = G(I.qsrc)[(it - 1) * I.nsrc + src];
        G(qout)[src] = (float)((double)q0v + (double)(q1v - q0v) * (t - (double)G(I.tsrc)[it - 2]) /
                                                  (double)(G(I.tsrc)[it - 1] - G(I.tsrc)[it - 2]));
      }
    }
  }
  __syncthreads();
}

// The fold, step scalars and forcing series of the next step, computed while the current
// step's continuity runs (they depend on the momentum sweep, not on the water levels):
// one BLOCK-thread block of the continuity launch per scenario, staged in snext /
// qtsrc_nx with sc->ahead set. No other step kernel reads these, so the continuity
// blocks beside it are undisturbed. The next k_forcing commits them; a host write of the
// scalars, a new momentum sweep or a host-side fold clears the mark, and that k_forcing
// then computes the step itself.
__device__ void forcing_ahead_block(const DevInst& I) {
  __shared__ StepImage F;
  copy_words(&F.s, I.sc);
  __syncthreads();
  if (F.s.ahead) return;   // staged already (nothing has run since)
  fold_image<BLOCK>(I, F);
  step_series(I, F, I.qtsrc_nx, 0);
  copy_words(I.snext, &F);
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) G(I.sc)->ahead = 1;
}

// One drainage structure (sfincs_discharges.f90:371-655): its discharge from the levels of
// its intake and outfall cells (pump, culvert, check valve, or a gate whose opening follows
// the intake level or set times, with the Bates et al. (2010) inertial formula), relaxed
// towards the last step's over structure_relax seconds and limited by the water the
// giving cell holds; the intake loses it and the outfall gains it as point sources. The
// arithmetic and its mix of f32 and f64 (zs, z_volume) are the reference's.
__device__ void drainage_update(const DevInst& I, int k, double t, float dt) {
  const int jin = I.nsrc + 2 * k, jout = jin + 1;
  const int nmin = G(I.src_cell)[jin], nmout = G(I.src_cell)[jout];
  const double zin = G(I.zs)[nmin], zout = G(I.zs)[nmout];
  const auto qt = G(I.qtsrc);
  const auto par = G(I.drn_par);
  const uint32_t b = 6u * (uint32_t)k;
  const int type = G(I.drn_type)[k];
  float qq = 0.0f;
  if (type == 1) {
    qq = par[b];
  } else if (type == 2 || type == 3) {
    if (zin > zout) qq = (float)((double)par[b] * sqrt(zin - zout));
    else qq = (float)((double)(-par[b]) * sqrt(zout - zin));
    if (type == 3) qq = sfx::vmax<float>(qq, 0.0f);   // a check valve passes intake -> outfall only
  } else {
    const float wdt = par[b], zsill = par[b + 1], mng = par[b + 2], p4 = par[b + 3], p5 = par[b + 4];
    const float tcls = par[b + 5];
    const float dzds = (float)((zout - zin) / (double)G(I.drn_dist)[k]);
    float frac = G(I.drn_frac)[k];
    const float hgate = (float)sfx::vmax<double>(sfx::vmax<double>(zin, zout) - (double)zsill, 0.0);
    const float dfrac = dt / tcls;
    const float qq0 = -qt[jin] / (wdt * sfx::vmax<float>(frac, 0.001f));
    int st = G(I.drn_status)[k];
    if (type == 4) {   // open while the intake level lies between zmin (p4) and zmax (p5)
      if (st == 0) { if (zin > (double)p4 && zin < (double)p5) st = 3; }
      else if (st == 1) { if (zin <= (double)p4 || zin >= (double)p5) st = 2; }
    } else {           // closes at t_close (p4), opens at t_open (p5)
      if (st == 0) { if (t >= (double)p5) st = 3; }
      else if (st == 1) { if (t >= (double)p4 && t < (double)p5) st = 2; }
    }
    if (st == 2) {
      frac = frac - dfrac;
      if (frac < 0.0f) { frac = 0.0f; st = 0; }
    } else if (st == 3) {
      frac = frac + dfrac;
      if (frac > 1.0f) { frac = 1.0f; st = 1; }
    }
    G(I.drn_frac)[k] = frac;
    G(I.drn_status)[k] = st;
    const float g = I.c.g;
    const float h73 = sfx::libm_pow_dev(hgate, 2, (const uint64_t*)I.pw.t13, (const uint64_t*)I.pw.t53,
                                        (const uint64_t*)I.pw.t73, (const uint64_t*)I.pw.t15);
    qq = (qq0 - g * hgate * dzds * dt) / (1.0f + g * (mng * mng) * dt * fabsf(qq0) / h73);
    qq = qq * wdt * frac;
  }
  const float rr = 1.0f / (I.structure_relax / dt);
  qq = rr * qq + (1.0f - rr) * -qt[jin];
  if (I.subgrid) {
    if (qq > 0.0f) qq = (float)sfx::vmin<double>((double)qq, sfx::vmax<double>(G(I.vol)[nmin], 0.0) / (double)dt);
    else qq = (float)sfx::vmax<double>((double)qq, -(sfx::vmax<double>(G(I.vol)[nmout], 0.0) / (double)dt));
  } else {
    const int ri = I.dense ? 0 : G(I.z_iref)[nmin] - 1, ro = I.dense ? 0 : G(I.z_iref)[nmout] - 1;
    if (qq > 0.0f)
      qq = (float)sfx::vmin<double>((double)qq, sfx::vmax<double>((zin - (double)G(I.zb)[nmin]) *
                                                                      (double)I.c.cell_area[ri], 0.0) / (double)dt);
    else
      qq = (float)sfx::vmax<double>((double)qq, -(sfx::vmax<double>((zout - (double)G(I.zb)[nmout]) *
                                                                         (double)I.c.cell_area[ro], 0.0) / (double)dt));
  }
  qt[jin] = -qq;
  qt[jout] = qq;
}

// The step's boundary-series position and uvmean filter factors, for the boundary pass.
template <typename R>
struct BndStep {
  int itb0, itb1;
  float tbfac;
  double t;
  R factime, facrel, one_minus_factime;
  MeteoNow m;   // for the pressure correction of water-level boundaries (F_PATM)
};

// One boundary flux point's inputs: everything update_boundary_fluxes reads from tables and
// levels before its arithmetic (the level of its boundary cell, the interior level, the uv
// tables), loaded by bnd_load without a store in between.
template <typename R>
struct BndPt {
  int ib, nmb, nmi, ip, dir, kb;
  float zsb1, zsb0, floorz, zbuv_or_zmin, zmaxu, havg_zmax, zb_nmi;
  R zsnmi, uvm;
  double vol_nmi;
  int kcuv;
};

// update_boundary_points / update_boundary_conditions (:728-917) for the point's boundary
// cell, and the point's other inputs (:941-1000).
template <typename R>
__device__ __forceinline__ void bnd_load(const DevInst& I, const BndStep<R>& B, int ib, BndPt<R>& p) {
  const Consts& c = I.c;
  p.ib = ib;
  p.nmb = G(I.k2_nmb)[ib];
  p.kb = G(I.kcs)[p.nmb] & KCS_MASK;
  const int indb = G(I.k2_ib)[ib];
  p.nmi = G(I.k2_nmi)[ib];
  p.ip = G(I.k2_slot)[ib];
  p.dir = G(I.k2_dir)[ib];
  p.zsnmi = (R)G(I.zs)[p.nmi];
  p.uvm = G((const R*)I.uvmean)[ib];
  p.kcuv = G(I.kcuv)[p.ip];
  p.floorz = I.subgrid ? G(I.z_zmin)[p.nmb] : G(I.zb)[p.nmb];
  if (I.subgrid) {
    const float4 s4 = ldf4(G(I.uv_s4), p.ip);
    p.zmaxu = s4.x; p.havg_zmax = s4.y;
    p.zbuv_or_zmin = G(I.uv_zmin)[p.ip];
    p.vol_nmi = G(I.vol)[p.nmi];
    p.zb_nmi = 0.0f;
  } else {
    p.zbuv_or_zmin = G(I.k2_zbuv)[ib];
    p.zb_nmi = G(I.zb)[p.nmi];
    p.zmaxu = p.havg_zmax = 0.0f;
    p.vol_nmi = 0.0;
  }
  if (p.kb == 2) {
    // update_boundary_points (:728-756) at the points this cell interpolates between
    auto series_at = [&](const float* ser, int k) -> float {
      const float za = G(ser)[(B.itb0 - 1) * I.nbnd + k], zb = G(ser)[(B.itb1 - 1) * I.nbnd + k];
      return za + (zb - za) * B.tbfac;
    };
    // update_boundary_conditions, water-level boundary (:789-917); the infragravity level
    // zig (bzi, :829-848) is interpolated as the water level
    double zst;
    float zig = 0.0f;
    if (I.nbnd > 1) {
      const float f = G(I.gb_fac)[indb];
      const int i1 = G(I.gb_ind1)[indb] - 1, i2 = G(I.gb_ind2)[indb] - 1;
      zst = (double)(series_at(I.zs_bnd, i1) * f + series_at(I.zs_bnd, i2) * (1.0f - f));
      if (I.zsi_bnd) zig = series_at(I.zsi_bnd, i1) * f + series_at(I.zsi_bnd, i2) * (1.0f - f);
    } else {
      zst = (double)series_at(I.zs_bnd, 0);
      if (I.zsi_bnd) zig = series_at(I.zsi_bnd, 0);
    }
    if ((c.flags & sfx::F_PATM) && c.pavbnd > 1.0f)   // barometric correction (:820-826)
      zst = zst + (double)((c.pavbnd - meteo_patm(I, B.m, p.nmb)) / (c.rhow * 9.81f));
    float zb0, zb1;
    if (B.t < (double)(c.tspinup - 1.0e-3f)) {
      const float smfac = (float)(1.0 - (B.t - (double)c.t0) / (double)(c.tspinup - c.t0));
      const float zs0act = (float)(zst + 0.0);
      zb0 = c.zini * smfac + zs0act * (1.0f - smfac);
      const float zs1act = (float)(zst + 0.0 + (double)zig);
      zb1 = c.zini * smfac + zs1act * (1.0f - smfac);
    } else {
      zb0 = (float)(zst + 0.0);
      zb1 = (float)(zst + 0.0 + (double)zig);
    }
    p.zsb0 = sfx::vmax<float>(zb0, p.floorz);
    p.zsb1 = sfx::vmax<float>(zb1, p.floorz);
  } else if (p.kb == 5) {
    // downstream river cell (:878-901): the interior level plus the slope offset, raised
    // to the bed, in f64 (the interior cell is an ordinary cell, which no boundary step writes)
    const double zst = sfx::vmax<double>(G(I.zs)[G(I.gb_rcell)[indb]] + (double)G(I.gb_rdzs)[indb], (double)p.floorz);
    p.zsb0 = (float)zst;
    p.zsb1 = p.zsb0;
  } else {
    p.zsb1 = G(I.zsb)[indb];
    p.zsb0 = G(I.zsb0)[indb];
  }
}

// update_boundary_fluxes (:941-1117), bndtype 1 (weakly reflective), from the loaded inputs.
template <typename R, typename QP, typename MP>
__device__ __forceinline__ void bnd_finish(const DevInst& I, const BndStep<R>& B, const BndPt<R>& p, QP q, QP uv,
                                           MP uvmean, MP zsmax) {
  const Consts& c = I.c;
  const int ip = p.ip, ib = p.ib, nmb = p.nmb;
  const R zsnmi = p.zsnmi;
  R zsnmb = p.zsb1;
  R zs0nmb = p.zsb0;
  R hnmb;
  const R huth = c.huthresh;
  if (I.subgrid) {
    const int nl = c.nl;
    const R zsuv = sfx::vmax<R>(zsnmb, zsnmi);
    const R zmaxu = p.zmaxu, zminu = p.zbuv_or_zmin;
    R depthuv;
    if (zsuv >= zmaxu - R(1.0e-4f)) {
      depthuv = (R)p.havg_zmax + zsuv;
    } else if (zsuv > zminu) {
      const R dzuv = (zmaxu - zminu) / (R)(nl - 1);
      const int iuv = (int)((zsuv - zminu) / dzuv) + 1;
      const R facint = (zsuv - (zminu + (R)(iuv - 1) * dzuv)) / dzuv;
      const auto lv = G(I.uv_lev) + (size_t)ip * nl;
      depthuv = (R)ldf4(lv, iuv - 1).x + ((R)ldf4(lv, iuv).x - (R)ldf4(lv, iuv - 1).x) * facint;
    } else {
      depthuv = R(0);
    }
    hnmb = sfx::vmax<R>(depthuv, huth);
    zsnmb = sfx::vmax<R>(zsnmb, (R)p.floorz);
    zs0nmb = sfx::vmax<R>(zs0nmb, (R)p.floorz);
  } else {
    hnmb = sfx::vmax<R>(R(0.5f) * (zsnmb + zsnmi) - (R)p.zbuv_or_zmin, huth);
    zsnmb = sfx::vmax<R>(zsnmb, (R)p.floorz);
    zs0nmb = sfx::vmax<R>(zs0nmb, (R)p.floorz);
  }
  R qb, um;   // the point's discharge and uvmean as the filter below reads them
  const R uvm = p.uvm;
  if (hnmb < huth + R(1.0e-6f) || p.kcuv == 3) {
    qb = R(0);
    um = R(0);
    q[ip] = R(0); uv[ip] = R(0);
  } else {
    const R g = c.g;
    const R ui = sfx::rsqrt_<R>(g / hnmb) * (zsnmb - zs0nmb);
    const int dir = p.dir;
    const R ub = (R)dir * (R(2) * ui - sfx::rsqrt_<R>(g / hnmb) * (zsnmi - zs0nmb));
    qb = ub * hnmb + uvm;
    if (I.subgrid) {
      if (p.vol_nmi <= 0.0) qb = (dir == 1) ? sfx::vmax<R>(qb, R(0)) : sfx::vmin<R>(qb, R(0));
      if (zsnmb - (R)p.floorz < huth) qb = (dir == 1) ? sfx::vmin<R>(qb, R(0)) : sfx::vmax<R>(qb, R(0));
    } else {
      if (zsnmi - (R)p.zb_nmi <= huth) qb = (dir == 1) ? sfx::vmax<R>(qb, R(0)) : sfx::vmin<R>(qb, R(0));
      if (zsnmb - (R)p.floorz < huth) qb = (dir == 1) ? sfx::vmin<R>(qb, R(0)) : sfx::vmax<R>(qb, R(0));
    }
    q[ip] = qb;
    uv[ip] = sfx::vmax<R>(sfx::vmin<R>(qb / hnmb, R(4)), R(-4));
    um = uvm;
  }
  if (c.btfilter >= -1.0e-6f) uvmean[ib] = B.factime * qb + B.facrel * B.one_minus_factime * um;
  else uvmean[ib] = R(0);
  G(I.zs)[nmb] = (double)p.zsb1;
  mark_cell(I, nmb);
  if (c.flags & sfx::F_ZSMAX) zsmax[nmb] = (R)sfx::vmax<double>((double)zsmax[nmb], (double)p.zsb1);
}

// One block per scenario: the step's fold, scalars and forcing series (committed from
// k_forcing_ahead, or computed here), then boundary points, conditions and fluxes
// (sfincs_lib.f90:379-550 with sfincs_meteo.f90:1423-1558, sfincs_boundaries.f90:682-1122,
// sfincs_discharges.f90:353-369).
template <typename R>
__global__ void __launch_bounds__(FBLOCK) k_forcing(const DevInst* __restrict__ insts, const int32_t* __restrict__ ids,
                                                   int mode) {
  const DevInst& I = insts[ids[blockIdx.x]];
  const auto S = G(I.sc);
  const Consts& c = I.c;
  __shared__ StepImage F;
  __shared__ int sh_ahead;
  if (threadIdx.x == 0) sh_ahead = mode == 0 && S->ahead;
  __syncthreads();
  if (sh_ahead) {
    copy_words(&F, I.snext);
    __syncthreads();
    if (F.go)
      for (int k = threadIdx.x; k < I.nsrc; k += FBLOCK) G(I.qtsrc)[k] = G(I.qtsrc_nx)[k];
    if (F.folded) reset_tiles(I);
  } else {
    fold_sweep(I);
    copy_words(&F.s, I.sc);
    __syncthreads();
    step_series(I, F, I.qtsrc, mode);
  }
  if (threadIdx.x == 0) F.s.ahead = 0;
  __syncthreads();
  copy_words(I.sc, &F.s);
  if (!F.go) return;
  const double t = F.s.t;
  // Boundary points, conditions and fluxes (:728-1117) in two passes. A flux point evaluates
  // the level of its water-level boundary cell itself (update_boundary_points and the kcs 2
  // branch of update_boundary_conditions, the same arithmetic), so only the Neumann copies
  // (kcs 6) run first, and the barrier after them is needed only when a copy reads a cell
  // whose level a flux point sets or sets a cell that a flux point reads (bnd_overlap, found
  // on the host).
  for (int k = threadIdx.x; k < I.n6; k += blockDim.x) {
    const int cell = G(I.gb6)[2 * k], inner = G(I.gb6)[2 * k + 1];
    G(I.zs)[cell] = G(I.zs)[inner];
    mark_cell(I, cell);
  }
  if (I.bnd_overlap) __syncthreads();
  // update_boundary_fluxes (:941-1117), bndtype 1 (weakly reflective), over the flux
  // points (kcuv 2 points whose boundary cell is not a Neumann cell). Each thread takes two
  // points and issues the loads of both before it finishes either, so the dependent chain
  // of table and level reads is paid once per two points.
  const int curb = __builtin_amdgcn_readfirstlane(F.s.cur);
  const auto q = G((R*)I.q[curb]);
  const auto uv = G((R*)I.uv[curb]);
  const auto uvmean = G((R*)I.uvmean);
  const auto zsmax = G((R*)I.zsmax);
  const BndStep<R> B{F.itb0, F.itb1, F.tbfac, t, (R)F.s.factime, (R)F.s.facrel, (R)F.s.one_minus_factime,
                     meteo_now(I, &F.s)};
  for (int k0 = threadIdx.x; k0 < I.nflux; k0 += 2 * FBLOCK) {
    const int k1 = k0 + FBLOCK;
    BndPt<R> a, b;
    bnd_load<R>(I, B, G(I.k2_flux)[k0], a);
    if (k1 < I.nflux) bnd_load<R>(I, B, G(I.k2_flux)[k1], b);
    bnd_finish<R>(I, B, a, q, uv, uvmean, zsmax);
    if (k1 < I.nflux) bnd_finish<R>(I, B, b, q, uv, uvmean, zsmax);
  }
  if (I.ndrn > 0) {   // update_discharges' drainage structures, after the boundaries (sfincs_lib.f90:530-550)
    __syncthreads();
    for (int k = threadIdx.x; k < I.ndrn; k += blockDim.x) drainage_update(I, k, t, F.s.dt);
  }
}

// ----------------------------------------------------------------------------- momentum
// Loads the wet/dry threshold of a uv-point: subgrid_uv_zmin or zbuvmx.
template <bool SUBGRID>
__device__ __forceinline__ float uv_floor(const DevInst& I, int slot) {
  return SUBGRID ? G(I.uv_zmin)[slot] : G(I.zbuvmx)[slot];
}

// Completes one active uv-point from its gathered inputs and its preloaded wet/dry
// threshold `floor` (uv_floor): dry points get q = uv = 0, wet points the full update.
// The update of a wet uv-point (zsu from uv_wet): its tables, the physics, the stores
// (kfuv only when `kf_old`, its kfuv byte as loaded, is not already 1; -1: not loaded).
template <typename R, bool SUBGRID, typename QP>
__device__ __forceinline__ void wet_uv(const DevInst& I, int slot, const sfx::UvIn<R>& in, sfx::UvTables& tb, R zsu,
                                       R dt, QP q1, QP uv1, R& mden, int& br, int kf_old = -1) {
  if (SUBGRID) {
    const float4 s4 = ldf4(G(I.uv_s4), slot);
    tb.zmax = s4.x; tb.havg_zmax = s4.y; tb.nrep_zmax = s4.z; tb.navg_w = s4.w; tb.fnfit = G(I.uv_fnfit)[slot];
  } else {
    tb.gn2uv = G(I.gn2uv)[slot];
  }
  const sfx::UvOut<R> o = sfx::uv_update<R, SUBGRID>(in, tb, I.c, zsu,
                                                     sfx::GRow<const float4>{G(I.uv_lev), SUBGRID ? (uint32_t)slot * I.c.nl : 0u},
                                                     G(I.x73), sfx::pow_tables_g(I.pw), dt);
  q1[slot] = o.q; uv1[slot] = o.uv;
  if (kf_old != 1) G(I.kfuv)[slot] = 1;
  // the step is the minimum of 1 / den over the points, which is 1 / (the largest den),
  // since a correctly rounded quotient falls as its divisor grows (NaN is skipped)
  if (o.den > mden) mden = o.den;
  if (I.c.flags & sfx::F_TSA) G(I.tsa_req)[slot] = (float)(R(1) / o.den);   // sfincs_momentum.f90:737-741
  br = o.branch;
}

template <typename R, bool SUBGRID, typename QP>
__device__ __forceinline__ void finish_uv(const DevInst& I, int slot, const sfx::UvIn<R>& in, float floor, R dt,
                                          QP q1, QP uv1, R& mden, int& br) {
  sfx::UvTables tb;
  if (SUBGRID) tb.zmin = floor;
  else tb.zbuvmx = floor;
  R zsu;
  if (!sfx::uv_wet<R, SUBGRID>(in, tb, I.c, zsu)) {
    q1[slot] = R(0); uv1[slot] = R(0); G(I.kfuv)[slot] = 0;
    br = 1;
    return;
  }
  wet_uv<R, SUBGRID>(I, slot, in, tb, zsu, dt, q1, uv1, mden, br);
}

// Dense scenarios, one thread per cell: the u-point on its right and the v-point above
// it. Every load that does not depend on a branch is issued before the first test, so a
// thread waits for memory once for its stencil and thresholds, once more for the tables
// of wet points. Neighbours follow the reference's tables for a full rectangle
// (sfincs_domain.f90:1005-1261): missing left/right/below/above faces read the dummy
// point (q = uv = 0), missing cross neighbours of u_ndm/u_num fall back to the point.
// compute_fluxes_over_structures (sfincs_structures.f90:592-697), which the reference runs
// after the momentum sweep of the step on the same levels, so it is fused here: at a weir
// point (kcuv 3) the broad-crested weir formula on the levels of its two cells (as real*4),
// submerged when the lower head exceeds 2/3 of the upper, times the crest length; uv = 0.
template <typename R>
__device__ __forceinline__ R weir_flux(const DevInst& I, int s, double zs_nm, double zs_nmu) {
  const float4 p = ldf4(G(I.struc_par), G(I.struc_of)[s]);
  const R zc = (R)p.x, cd = (R)p.y, len = (R)p.z;
  const R a = (R)zs_nm, b = (R)zs_nmu;
  if (a < zc && b < zc) return R(0);
  const bool fwd = a > b;
  const R h1 = (fwd ? a : b) - zc, h2 = (fwd ? b : a) - zc;
  R qs;
  if (h2 > (R)(2.0f / 3.0f) * h1)
    qs = cd * h2 * sfx::rsqrt_<R>((R)(2.0f * 9.81f) * (h1 - h2));
  else
    qs = cd * (R)1.7049f * sfx::pow_third<R>(h1, 3, sfx::pow_tables_g(I.pw));
  qs = qs * len;
  return fwd ? qs : -qs;
}

// Gathers the inputs of one uv-point. The neighbour values are read whether or not the
// scenario's physics uses them (uv_update reads them only under the flags that need
// them), so that the loads carry no branch.
template <typename R, bool SUBGRID, typename QP>
__device__ __forceinline__ void gather_uv(const DevInst& I, QP q0, QP u0, int s, int s_nmd, int s_nmu, int s_ndm,
                                          int s_num, int v_nm, int v_ndm, int v_nmu, int v_ndmu, int c_nm,
                                          int c_nmu, sfx::UvIn<R>& in) {
  in.zs_nm = G(I.zs)[c_nm]; in.zs_nmu = G(I.zs)[c_nmu];
  in.qx_nm = q0[s]; in.uu_nm = u0[s];
  in.qx_nmd = q0[s_nmd]; in.qx_nmu = q0[s_nmu];
  in.qy_nm = q0[v_nm]; in.qy_nmu = q0[v_nmu]; in.qy_ndm = q0[v_ndm]; in.qy_ndmu = q0[v_ndmu];
  in.uu_nmd = u0[s_nmd]; in.uu_nmu = u0[s_nmu]; in.uu_ndm = u0[s_ndm]; in.uu_num = u0[s_num];
  in.v_ndm = u0[v_ndm]; in.v_ndmu = u0[v_ndmu]; in.v_nm = u0[v_nm]; in.v_nmu = u0[v_nmu];
  in.kcuv_nmd = G(I.kcuv)[s_nmd]; in.kcuv_nmu = G(I.kcuv)[s_nmu];
  in.kfuv = G(I.kfuv)[s] & 1; in.mask_adv = G(I.mask_adv)[s];
  if (SUBGRID) {
    in.vol_nm = G(I.vol)[c_nm]; in.vol_nmu = G(I.vol)[c_nmu];
    in.zsderv_nm = G((const R*)I.zsderv)[c_nm]; in.zsderv_nmu = G((const R*)I.zsderv)[c_nmu];
  } else {
    in.zb_nm = G(I.zb)[c_nm]; in.zb_nmu = G(I.zb)[c_nmu];
  }
}

template <typename R, bool SUBGRID>
__device__ __forceinline__ void mom_cell(const DevInst& I, int tile, int* cnt, R& mden) {
  const auto S = G(I.sc);
  const int cur = __builtin_amdgcn_readfirstlane(S->cur);
  const R dt = (R)S->dt;
  const auto q0 = G((const R*)I.q[cur]);
  const auto u0 = G((const R*)I.uv[cur]);
  const auto q1 = G((R*)I.q[cur ^ 1]);
  const auto u1 = G((R*)I.uv[cur ^ 1]);
  const int nmax = I.nmax, mmax = I.mmax, np = I.np, dm = 2 * np;
  const int tn = tile % I.mntn, tmi = tile / I.mntn;
  const int n = tn * I.mtn + (threadIdx.x & (I.mtn - 1));
  const int m = tmi * I.mtm + (threadIdx.x >> I.mtn_sh);
  if (n >= nmax || m >= mmax) return;
  const int c = m * nmax + n;
  const bool hu = m < mmax - 1, hv = n < nmax - 1;
  // A point that does not exist (u on the last row, v on the last column) reads the dummy
  // slot and its own cell, and is not updated.
  sfx::UvIn<R> iu, iv;
  {   // u-point between c and c+nmax
    const int s = hu ? c : dm;
    gather_uv<R, SUBGRID>(I, q0, u0, s, hu && m >= 1 ? c - nmax : dm, m + 1 < mmax - 1 ? c + nmax : dm,
                          hu ? (n >= 1 ? c - 1 : c) : dm, hu ? (n < nmax - 1 ? c + 1 : c) : dm,
                          hu && hv ? np + c : dm, hu && n >= 1 ? np + c - 1 : dm, hu && hv ? np + c + nmax : dm,
                          hu && n >= 1 ? np + c + nmax - 1 : dm, c, hu ? c + nmax : c, iu);
  }
  {   // v-point between c and c+1
    const int s = hv ? np + c : dm;
    gather_uv<R, SUBGRID>(I, q0, u0, s, hv && n >= 1 ? np + c - 1 : dm, n + 1 < nmax - 1 ? np + c + 1 : dm,
                          hv ? (m >= 1 ? s - nmax : s) : dm, hv ? (m < mmax - 1 ? s + nmax : s) : dm,
                          hv && hu ? c : dm, hv && m >= 1 ? c - nmax : dm, hv && hu ? c + 1 : dm,
                          hv && m >= 1 ? c + 1 - nmax : dm, c, hv ? c + 1 : c, iv);
  }
  const int ku = G(I.kcuv)[hu ? c : dm], kv = G(I.kcuv)[hv ? np + c : dm];
  const float fu = uv_floor<SUBGRID>(I, hu ? c : dm), fv = uv_floor<SUBGRID>(I, hv ? np + c : dm);
  if (hu) {
    if (ku == 1 || ku == 6) {
      iu.idir = 0; iu.iref = 1; iu.itype = 0;
      uv_meteo<R>(I, S, c, c + nmax, iu);
      int br = 0;
      finish_uv<R, SUBGRID>(I, c, iu, fu, dt, q1, u1, mden, br);
      cnt[br]++;
    } else if (ku == 3) {
      q1[c] = weir_flux<R>(I, c, iu.zs_nm, iu.zs_nmu); u1[c] = R(0);
      cnt[0]++;
    } else {
      q1[c] = iu.qx_nm; u1[c] = iu.uu_nm;
      cnt[0]++;
    }
  } else {
    cnt[0]++;
  }
  if (hv) {
    const int s = np + c;
    if (kv == 1 || kv == 6) {
      iv.idir = 1; iv.iref = 1; iv.itype = 0;
      uv_meteo<R>(I, S, c, c + 1, iv);
      int br = 0;
      finish_uv<R, SUBGRID>(I, s, iv, fv, dt, q1, u1, mden, br);
      cnt[br]++;
    } else if (kv == 3) {
      q1[s] = weir_flux<R>(I, s, iv.zs_nm, iv.zs_nmu); u1[s] = R(0);
      cnt[0]++;
    } else {
      q1[s] = iv.qx_nm; u1[s] = iv.uu_nm;
      cnt[0]++;
    }
  } else {
    cnt[0]++;
  }
}

// The stencil of dense uv-point (n, m) of plane `vplane` beyond its branch inputs: the q
// and uv of its neighbours in both planes (missing ones read the dummy slot or the point
// itself, as the reference's tables for a full rectangle give them), the kcuv of its two
// neighbours along its direction, its advection mask and the cell values of its two cells.
template <typename R, bool SUBGRID, typename QC>
__device__ __forceinline__ void point_stencil(const DevInst& I, int n, int m, int vplane, int kf, QC q0, QC u0,
                                              sfx::UvIn<R>& in) {
  const int nmax = I.nmax, mmax = I.mmax, np = I.np, dm = 2 * np;
  const bool hu = m < mmax - 1, hv = n < nmax - 1;
  const int c = m * nmax + n;
  const int s = vplane ? np + c : c;
  const int c_nmu = vplane ? c + 1 : c + nmax;
  int s_nmd, s_nmu, s_ndm, s_num, v_nm, v_ndm, v_nmu, v_ndmu;
  if (!vplane) {
    s_nmd = m >= 1 ? c - nmax : dm; s_nmu = m + 1 < mmax - 1 ? c + nmax : dm;
    s_ndm = n >= 1 ? c - 1 : c; s_num = n < nmax - 1 ? c + 1 : c;
    v_nm = hv ? np + c : dm; v_ndm = n >= 1 ? np + c - 1 : dm;
    v_nmu = hv ? np + c + nmax : dm; v_ndmu = n >= 1 ? np + c + nmax - 1 : dm;
  } else {
    s_nmd = n >= 1 ? s - 1 : dm; s_nmu = n + 1 < nmax - 1 ? s + 1 : dm;
    s_ndm = m >= 1 ? s - nmax : s; s_num = m < mmax - 1 ? s + nmax : s;
    v_nm = hu ? c : dm; v_ndm = m >= 1 ? c - nmax : dm;
    v_nmu = hu ? c + 1 : dm; v_ndmu = m >= 1 ? c + 1 - nmax : dm;
  }
  in.qx_nm = q0[s]; in.uu_nm = u0[s];
  in.qx_nmd = q0[s_nmd]; in.qx_nmu = q0[s_nmu];
  in.qy_nm = q0[v_nm]; in.qy_nmu = q0[v_nmu]; in.qy_ndm = q0[v_ndm]; in.qy_ndmu = q0[v_ndmu];
  in.uu_nmd = u0[s_nmd]; in.uu_nmu = u0[s_nmu]; in.uu_ndm = u0[s_ndm]; in.uu_num = u0[s_num];
  in.v_ndm = u0[v_ndm]; in.v_ndmu = u0[v_ndmu]; in.v_nm = u0[v_nm]; in.v_nmu = u0[v_nmu];
  in.kcuv_nmd = G(I.kcuv)[s_nmd]; in.kcuv_nmu = G(I.kcuv)[s_nmu];
  in.kfuv = kf & 1; in.mask_adv = G(I.mask_adv)[s];
  if (SUBGRID) {
    in.vol_nm = G(I.vol)[c]; in.vol_nmu = G(I.vol)[c_nmu];
    in.zsderv_nm = G((const R*)I.zsderv)[c]; in.zsderv_nmu = G((const R*)I.zsderv)[c_nmu];
  } else {
    in.zb_nm = G(I.zb)[c]; in.zb_nmu = G(I.zb)[c_nmu];
  }
  in.idir = vplane; in.iref = 1; in.itype = 0;
}

// Dense scenarios, one thread per uv-point: tile 2t covers the u-points and tile 2t+1 the
// v-points of cell tile t (mtn x mtm cells), so both planes of a tile run next to each
// other on one XCD and share its cached cell data. A point first reads
// what decides its branch (kcuv, the two water levels, its wet/dry threshold); only wet
// points read the rest of their stencil, so waves without wet points finish after one
// round trip to memory.
// One uv-point of the per-point kernel from its preloaded branch inputs: kcuv `kc`, the
// kfuv byte `kf` (with the zero flags), the water levels of its two cells and its wet/dry
// threshold `fl`. Returns true when there is no point or it is dry with its zeros held.
template <typename R, bool SUBGRID, typename QC, typename QW>
__device__ __forceinline__ bool point_finish(const DevInst& I, GA<Scalars> S, int n, int m, int vplane, int kc, int kf,
                                             double zs_nm, double zs_nmu, float fl, R dt, QC q0, QC u0, QW q1,
                                             QW u1, int* cnt, R& mden) {
  const int nmax = I.nmax, mmax = I.mmax, np = I.np, dm = 2 * np;
  if (n >= nmax || m >= mmax) return true;
  const bool hu = m < mmax - 1, hv = n < nmax - 1;
  if (vplane ? !hv : !hu) { cnt[0]++; return true; }
  const int c = m * nmax + n;
  const int s = vplane ? np + c : c;
  const int c_nmu = vplane ? c + 1 : c + nmax;
  sfx::UvIn<R> in;
  in.zs_nm = zs_nm; in.zs_nmu = zs_nmu;
  sfx::UvTables tb;
  if (SUBGRID) tb.zmin = fl;
  else tb.zbuvmx = fl;
  if (kc == 3) {
    q1[s] = weir_flux<R>(I, s, zs_nm, zs_nmu); u1[s] = R(0);
    cnt[0]++;
    return false;
  }
  if (!(kc == 1 || kc == 6)) {
    q1[s] = q0[s]; u1[s] = u0[s];
    cnt[0]++;
    return false;
  }
  R zsu;
  if (!sfx::uv_wet<R, SUBGRID>(in, tb, I.c, zsu)) {
    if (kf & KF_ZERO2) { cnt[4]++; return true; }   // both buffers already hold the zeros
    q1[s] = R(0); u1[s] = R(0);
    G(I.kfuv)[s] = (kf & KF_ZERO1) ? KF_ZERO2 : KF_ZERO1;
    cnt[1]++;
    return false;
  }
  // the rest of the stencil of a wet point
  uv_meteo<R>(I, S, c, c_nmu, in);
  point_stencil<R, SUBGRID>(I, n, m, vplane, kf, q0, u0, in);
  int br = 0;
  wet_uv<R, SUBGRID>(I, s, in, tb, zsu, dt, q1, u1, mden, br, kf);
  cnt[br]++;
  return false;
}

// Dense scenarios, per-point kernel: tile 2t holds the u-points and tile 2t+1 the v-points
// of an mtn x (P mtm) block of cells; each thread updates P points, mtm rows apart. The
// geometry and branch inputs (kcuv, the two water levels, the wet/dry threshold) of all P
// points do not depend on the step scalars, so their loads are issued together before the
// scalars are read and the active test is made. Returns false if the scenario is not
// active (nothing is written).
template <typename R, bool SUBGRID, int P>
__device__ __forceinline__ bool mom_point(const DevInst& I, Scalars* Sp, int tile, int* cnt, R& mden, bool& held) {
  const auto S = G(Sp);
  const int nmax = I.nmax, mmax = I.mmax, np = I.np;
  const int vplane = tile & 1;
  const int t = tile >> 1;
  const int tn = t % I.mntn, tmi = t / I.mntn;
  const int n = tn * I.mtn + (threadIdx.x & (I.mtn - 1));
  const int m0 = tmi * I.mtm * P + (threadIdx.x >> I.mtn_sh);
  int kc[P], kf[P];
  float fl[P];
  double za[P], zb[P];
#pragma unroll
  for (int j = 0; j < P; ++j) {
    const int m = m0 + j * I.mtm;
    const bool has = n < nmax && m < mmax && (vplane ? n < nmax - 1 : m < mmax - 1);
    const int c = m * nmax + n;
    const int s = vplane ? np + c : c;
    kc[j] = 0; kf[j] = 0; fl[j] = 0.f; za[j] = 0.0; zb[j] = 0.0;
    if (has) {
      kc[j] = G(I.kcuv)[s]; kf[j] = G(I.kfuv)[s];
      za[j] = G(I.zs)[c]; zb[j] = G(I.zs)[vplane ? c + 1 : c + nmax];
      fl[j] = uv_floor<SUBGRID>(I, s);
    }
  }
  if (!S->active) return false;
  const int cur = __builtin_amdgcn_readfirstlane(S->cur);
  const R dt = (R)S->dt;
  const auto q0 = G((const R*)I.q[cur]);
  const auto u0 = G((const R*)I.uv[cur]);
  const auto q1 = G((R*)I.q[cur ^ 1]);
  const auto u1 = G((R*)I.uv[cur ^ 1]);
  held = true;
#pragma unroll
  for (int j = 0; j < P; ++j)
    held = point_finish<R, SUBGRID>(I, S, n, m0 + j * I.mtm, vplane, kc[j], kf[j], za[j], zb[j], fl[j], dt, q0, u0, q1,
                                    u1, cnt, mden) && held;
  return true;
}

// Points of a skipped (quiescent) per-point tile, for the counters: class 5, or copied (0)
// for the absent edge slots, as a run would count them.
template <int P>
__device__ __forceinline__ void count_tile_points(const DevInst& I, int tile, int* cnt) {
  const int nmax = I.nmax, mmax = I.mmax;
  const int vplane = tile & 1, t = tile >> 1;
  const int tn = t % I.mntn, tmi = t / I.mntn;
  const int n = tn * I.mtn + (threadIdx.x & (I.mtn - 1));
  const int m0 = tmi * I.mtm * P + (threadIdx.x >> I.mtn_sh);
#pragma unroll
  for (int j = 0; j < P; ++j) {
    const int m = m0 + j * I.mtm;
    if (n < nmax && m < mmax) ++cnt[(vplane ? n < nmax - 1 : m < mmax - 1) ? 5 : 0];
  }
}

// Dense scenarios, strip march. A wave covers columns n = n0-1 .. n0+62 of one strip
// (lane l -> n0+l-1) and marches over K rows m0 .. m0+K-1; lanes 1..62 own their column,
// lanes 0 and 63 only supply the neighbour values of the owned lanes. Each lane keeps rows
// m-1, m, m+1 (and the prefetched row m+2) of its column in registers and takes the n-1 /
// n+1 values from its neighbour lanes, so every element is read from memory once per
// strip, plus the two halo lanes and the halo rows at the ends of the march.
//
// Neighbours follow the reference's tables for a full rectangle (sfincs_domain.f90:
// 1005-1261): missing left/right/below/above faces read the dummy point (q = uv = 0),
// missing cross neighbours of u_ndm/u_num fall back to the point itself.
template <typename R>
struct Row {
  R uq, uu, vq, vu;     // q/uv of the u-point (n,m) and the v-point (n,m) (dummy where absent)
  int ku, kv;           // kcuv of the two points
  double z, vol;        // zs and z_volume of cell (n,m)
  R zd;                 // zsderv (subgrid) or zb (regular) of cell (n,m)
  R r_uq, r_uu, r_vq, r_vu, r_zd;   // from column n+1
  int r_kv;
  double r_z, r_vol;
  R l_uu, l_vq, l_vu;               // from column n-1
  int l_kv;
};

template <typename R, bool SUBGRID, typename QP>
__device__ __forceinline__ void strip_raw(const DevInst& I, QP q0, QP u0, int n, int m, bool colok, Row<R>& w,
                                          R dq, R du, int dk) {
  const int nmax = I.nmax, mmax = I.mmax, np = I.np;
  const bool ok = colok && m >= 0 && m < mmax;
  const int c = ok ? m * nmax + n : 0;
  const bool hu = ok && m < mmax - 1;
  const bool hv = ok && n < nmax - 1;
  w.uq = hu ? q0[c] : dq; w.uu = hu ? u0[c] : du; w.ku = hu ? (int)G(I.kcuv)[c] : dk;
  w.vq = hv ? q0[np + c] : dq; w.vu = hv ? u0[np + c] : du; w.kv = hv ? (int)G(I.kcuv)[np + c] : dk;
  w.z = ok ? G(I.zs)[c] : 0.0;
  if (SUBGRID) {
    w.vol = ok ? G(I.vol)[c] : 0.0;
    w.zd = ok ? G((const R*)I.zsderv)[c] : R(0);
  } else {
    w.vol = 0.0;
    w.zd = ok ? (R)G(I.zb)[c] : R(0);
  }
}

// Values of the neighbour columns; executed by every lane of the wave.
template <typename R, bool SUBGRID>
__device__ __forceinline__ void strip_nbrs(Row<R>& w) {
  w.r_uq = __shfl_down(w.uq, 1, WAVE); w.r_uu = __shfl_down(w.uu, 1, WAVE);
  w.r_vq = __shfl_down(w.vq, 1, WAVE); w.r_vu = __shfl_down(w.vu, 1, WAVE);
  w.r_kv = __shfl_down(w.kv, 1, WAVE);
  w.r_z = __shfl_down(w.z, 1, WAVE);
  w.r_vol = SUBGRID ? __shfl_down(w.vol, 1, WAVE) : 0.0;
  w.r_zd = __shfl_down(w.zd, 1, WAVE);
  w.l_uu = __shfl_up(w.uu, 1, WAVE); w.l_vq = __shfl_up(w.vq, 1, WAVE); w.l_vu = __shfl_up(w.vu, 1, WAVE);
  w.l_kv = __shfl_up(w.kv, 1, WAVE);
}

template <typename R, bool SUBGRID>
__device__ __forceinline__ void mom_strip(const DevInst& I, int tile, int* cnt, R& mden) {
  const auto S = G(I.sc);
  const int cur = __builtin_amdgcn_readfirstlane(S->cur);
  const R dt = (R)S->dt;
  const auto q0 = G((const R*)I.q[cur]);
  const auto u0 = G((const R*)I.uv[cur]);
  const auto q1 = G((R*)I.q[cur ^ 1]);
  const auto u1 = G((R*)I.uv[cur ^ 1]);
  const int nmax = I.nmax, mmax = I.mmax, np = I.np, dm = 2 * np;
  const int K = I.sk;
  const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
  const int strip = tile % I.nstrip, seg = tile / I.nstrip;
  const int n = strip * (WAVE - 2) + lane - 1;
  const bool colok = n >= 0 && n < nmax;
  const bool own = colok && lane >= 1 && lane <= WAVE - 2;
  const int m0 = (seg * (BLOCK / WAVE) + wv) * K;
  const int m1 = min(m0 + K, mmax);
  if (m0 >= mmax) return;                              // uniform over the wave
  const R dq = q0[dm], du = u0[dm];
  const int dk = G(I.kcuv)[dm];
  Row<R> A, B, C, D;
  strip_raw<R, SUBGRID>(I, q0, u0, n, m0 - 1, colok, A, dq, du, dk);
  strip_raw<R, SUBGRID>(I, q0, u0, n, m0, colok, B, dq, du, dk);
  strip_raw<R, SUBGRID>(I, q0, u0, n, m0 + 1, colok, C, dq, du, dk);
  strip_nbrs<R, SUBGRID>(A);
  strip_nbrs<R, SUBGRID>(B);
  for (int m = m0; m < m1; ++m) {
    strip_nbrs<R, SUBGRID>(C);
    strip_raw<R, SUBGRID>(I, q0, u0, n, m + 2, colok, D, dq, du, dk);   // prefetch
    if (own) {
      const int c = m * nmax + n;
      // u-point between cells (n,m) and (n,m+1)
      if (m < mmax - 1) {
        if (B.ku == 1 || B.ku == 6) {
          sfx::UvIn<R> in;
          in.zs_nm = B.z; in.zs_nmu = C.z;
          in.qx_nm = B.uq; in.uu_nm = B.uu;
          in.qx_nmd = A.uq; in.uu_nmd = A.uu; in.kcuv_nmd = A.ku;
          in.qx_nmu = C.uq; in.uu_nmu = C.uu; in.kcuv_nmu = C.ku;
          in.uu_ndm = n >= 1 ? B.l_uu : B.uu;
          in.uu_num = n < nmax - 1 ? B.r_uu : B.uu;
          in.qy_nm = B.vq; in.v_nm = B.vu;
          in.qy_ndm = B.l_vq; in.v_ndm = B.l_vu;
          in.qy_nmu = C.vq; in.v_nmu = C.vu;
          in.qy_ndmu = C.l_vq; in.v_ndmu = C.l_vu;
          in.kfuv = G(I.kfuv)[c] & 1; in.mask_adv = G(I.mask_adv)[c]; in.idir = 0; in.iref = 1; in.itype = 0;
          uv_meteo<R>(I, S, c, c + nmax, in);
          if (SUBGRID) { in.vol_nm = B.vol; in.vol_nmu = C.vol; in.zsderv_nm = B.zd; in.zsderv_nmu = C.zd; }
          else { in.zb_nm = (float)B.zd; in.zb_nmu = (float)C.zd; }
          int br = 0;
          finish_uv<R, SUBGRID>(I, c, in, uv_floor<SUBGRID>(I, c), dt, q1, u1, mden, br);
          cnt[br]++;
        } else if (B.ku == 3) {
          q1[c] = weir_flux<R>(I, c, B.z, C.z); u1[c] = R(0);
          cnt[0]++;
        } else {
          q1[c] = B.uq; u1[c] = B.uu;
          cnt[0]++;
        }
      } else {
        cnt[0]++;
      }
      // v-point between cells (n,m) and (n+1,m)
      if (n < nmax - 1) {
        const int s = np + c;
        if (B.kv == 1 || B.kv == 6) {
          sfx::UvIn<R> in;
          in.zs_nm = B.z; in.zs_nmu = B.r_z;
          in.qx_nm = B.vq; in.uu_nm = B.vu;
          in.qx_nmd = B.l_vq; in.uu_nmd = B.l_vu; in.kcuv_nmd = B.l_kv;
          in.qx_nmu = B.r_vq; in.uu_nmu = B.r_vu; in.kcuv_nmu = B.r_kv;
          in.uu_ndm = m >= 1 ? A.vu : B.vu;
          in.uu_num = m < mmax - 1 ? C.vu : B.vu;
          in.qy_nm = B.uq; in.v_nm = B.uu;
          in.qy_ndm = A.uq; in.v_ndm = A.uu;
          in.qy_nmu = B.r_uq; in.v_nmu = B.r_uu;
          in.qy_ndmu = A.r_uq; in.v_ndmu = A.r_uu;
          in.kfuv = G(I.kfuv)[s] & 1; in.mask_adv = G(I.mask_adv)[s]; in.idir = 1; in.iref = 1; in.itype = 0;
          uv_meteo<R>(I, S, c, c + 1, in);
          if (SUBGRID) { in.vol_nm = B.vol; in.vol_nmu = B.r_vol; in.zsderv_nm = B.zd; in.zsderv_nmu = B.r_zd; }
          else { in.zb_nm = (float)B.zd; in.zb_nmu = (float)B.r_zd; }
          int br = 0;
          finish_uv<R, SUBGRID>(I, s, in, uv_floor<SUBGRID>(I, s), dt, q1, u1, mden, br);
          cnt[br]++;
        } else if (B.kv == 3) {
          q1[s] = weir_flux<R>(I, s, B.z, B.r_z); u1[s] = R(0);
          cnt[0]++;
        } else {
          q1[s] = B.vq; u1[s] = B.vu;
          cnt[0]++;
        }
      } else {
        cnt[0]++;
      }
    }
    A = B; B = C; C = D;
  }
}

// Indexed scenarios: one thread per uv-point, neighbours through the reference's tables.
// As on dense meshes, the branch inputs come first and only wet points read the
// neighbour tables and values.
template <typename R, bool SUBGRID>
__device__ __forceinline__ void mom_idx(const DevInst& I, int tile, int* cnt, R& mden) {
  const auto S = G(I.sc);
  const int cur = __builtin_amdgcn_readfirstlane(S->cur);
  const R dt = (R)S->dt;
  const auto q0 = G((const R*)I.q[cur]);
  const auto u0 = G((const R*)I.uv[cur]);
  const auto q1 = G((R*)I.q[cur ^ 1]);
  const auto u1 = G((R*)I.uv[cur ^ 1]);
  const int ip = tile * BLOCK + threadIdx.x;
  const int npuv = I.npuv;
  if (ip >= npuv) return;
  const auto nb = G(I.nb);
  const int kc = G(I.kcuv)[ip], kf = G(I.kfuv)[ip];
  const int nm = nb[ip], nmu = nb[npuv + ip];
  sfx::UvIn<R> in;
  in.zs_nm = G(I.zs)[nm]; in.zs_nmu = G(I.zs)[nmu];
  sfx::UvTables tb;
  if (SUBGRID) tb.zmin = G(I.uv_zmin)[ip];
  else tb.zbuvmx = G(I.zbuvmx)[ip];
  if (kc == 3) {
    q1[ip] = weir_flux<R>(I, ip, in.zs_nm, in.zs_nmu); u1[ip] = R(0);
    cnt[0]++;
    return;
  }
  if (!(kc == 1 || kc == 6)) {
    q1[ip] = q0[ip]; u1[ip] = u0[ip];
    cnt[0]++;
    return;
  }
  R zsu;
  if (!sfx::uv_wet<R, SUBGRID>(in, tb, I.c, zsu)) {   // zero flags as in the per-point kernel
    if (kf & KF_ZERO2) { cnt[4]++; return; }
    q1[ip] = R(0); u1[ip] = R(0);
    G(I.kfuv)[ip] = (kf & KF_ZERO1) ? KF_ZERO2 : KF_ZERO1;
    cnt[1]++;
    return;
  }
  const uint8_t f = G(I.uvf)[ip];
  in.iref = (I.c.flags & sfx::F_QUADTREE) ? (f & 15) : 1;
  in.itype = (I.c.flags & sfx::F_QUADTREE) ? ((f >> 4) & 3) - 1 : 0;
  in.idir = (f >> 6) & 1;
  in.qx_nm = q0[ip]; in.uu_nm = u0[ip];
  if (I.c.flags & sfx::F_NBRS) {
    const int a = nb[2 * npuv + ip], b = nb[3 * npuv + ip], cdn = nb[4 * npuv + ip], d = nb[5 * npuv + ip];
    const int e = nb[6 * npuv + ip], f2 = nb[7 * npuv + ip], g = nb[8 * npuv + ip], h = nb[9 * npuv + ip];
    in.qx_nmd = q0[a]; in.qx_nmu = q0[b];
    in.qy_nm = q0[f2]; in.qy_nmu = q0[h]; in.qy_ndm = q0[e]; in.qy_ndmu = q0[g];
    in.uu_nmd = u0[a]; in.uu_nmu = u0[b]; in.uu_ndm = u0[cdn]; in.uu_num = u0[d];
    in.v_ndm = u0[e]; in.v_ndmu = u0[g]; in.v_nm = u0[f2]; in.v_nmu = u0[h];
    in.kcuv_nmd = G(I.kcuv)[a]; in.kcuv_nmu = G(I.kcuv)[b];
  } else {
    in.qx_nmd = in.qx_nmu = in.qy_nm = in.qy_nmu = in.qy_ndm = in.qy_ndmu = R(0);
    in.uu_nmd = in.uu_nmu = in.uu_ndm = in.uu_num = in.v_ndm = in.v_ndmu = in.v_nm = in.v_nmu = R(0);
    in.kcuv_nmd = in.kcuv_nmu = 0;
  }
  in.kfuv = kf & 1; in.mask_adv = G(I.mask_adv)[ip];
  uv_meteo<R>(I, S, nm, nmu, in);
  if (SUBGRID) {
    in.vol_nm = G(I.vol)[nm]; in.vol_nmu = G(I.vol)[nmu];
    in.zsderv_nm = G((const R*)I.zsderv)[nm]; in.zsderv_nmu = G((const R*)I.zsderv)[nmu];
  } else {
    in.zb_nm = G(I.zb)[nm]; in.zb_nmu = G(I.zb)[nmu];
  }
  int br = 0;
  wet_uv<R, SUBGRID>(I, ip, in, tb, zsu, dt, q1, u1, mden, br, kf);
  cnt[br]++;
}

// Wave epilogue of a momentum tile, without a block barrier: the wave's minimum time step
// lowers blkmin[tile] (NaN or non-positive values do not take part; the reduction starts
// at dtmax), the wave's per-branch counts are added to the tile's blkcnt slots, and the
// sweep-done flag is raised. Positive floats order like their bit patterns, so the
// minimum is an unsigned atomic min. Waves with no value below dtmax (every dry wave)
// skip the reduction.
template <typename R>
__device__ __forceinline__ void mom_epilogue(const DevInst& I, int tile, R mden, const int* cnt) {
  // the thread's step: min(dtmax, 1 / den) over its wet points (dtmax when it has none)
  float mdt = I.c.dtmax;
  if (mden > R(0)) mdt = fminf(mdt, (float)(R(1) / mden));
  const float v = mdt == mdt && mdt > 0.f ? mdt : I.c.dtmax;
  if (__ballot(v < I.c.dtmax)) {
    const uint32_t m = wave_min_last(__float_as_uint(v));
    if (threadIdx.x % WAVE == WAVE - 1) atomicMin((unsigned int*)&G(I.blkmin)[tile], m);
  }
  int cw[NCNT] = {0, 0, 0, 0, 0, 0};
  if (I.count)
    for (int k = 0; k < NCNT; ++k) {
      int v = cnt[k];
      for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
      cw[k] = v;
    }
  if (threadIdx.x % WAVE == 0) {
    if (tile == 0 && threadIdx.x == 0) { G(I.sc)->mom_done = 1; G(I.sc)->ahead = 0; }   // a new sweep: no staged step
    if (I.count)
      for (int k = 0; k < NCNT; ++k)
        if (cw[k]) atomicAdd((unsigned int*)&G(I.blkcnt)[NCNT * (uint32_t)tile + k], (unsigned int)cw[k]);
  }
}

// The (scenario, tile) of this block; false for the padding of a list. The scenario is an
// index into the launch's descriptor argument (0 for a direct list).
// b: the block's index in the list, grid: the list's (padded) length.
__device__ __forceinline__ bool work_item(const WorkList& L, uint32_t b, uint32_t grid, int& tile, int& inst) {
  if (L.direct) {
    const uint32_t x = b & 7u, k = b >> 3;
    uint32_t nat;
    if (L.run_sh < 0) nat = x * (grid >> 3) + k;
    else nat = ((((k >> L.run_sh) << 3) + x) << L.run_sh) + (k & ((1u << L.run_sh) - 1u));
    if (nat >= (uint32_t)L.n) return false;
    tile = (int)nat; inst = 0;
    return true;
  }
  // items are stored XCD-major (see xcd_major): the blocks of one XCD read consecutive items
  const Item it = L.items[(b & 7u) * (grid >> 3) + (b >> 3)];
  if (it.inst < 0) return false;
  tile = it.tile; inst = it.inst;
  return true;
}

// One launch per (layout, mesh kind): every block handles one (scenario, tile) item.
template <typename R, bool SUBGRID>
__global__ void __launch_bounds__(BLOCK) k_mom_cell(const DevInst* __restrict__ insts, WorkList L) {
  int tile, inst;
  if (!work_item(L, blockIdx.x, gridDim.x, tile, inst)) return;
  const DevInst& I = insts[inst];
  Scalars* Sp = L.direct ? L.sc : I.sc;
  if (!G(Sp)->active) return;
  R mden = R(0);
  int cnt[NCNT] = {0, 0, 0, 0, 0, 0};
  mom_cell<R, SUBGRID>(I, tile, cnt, mden);
  mom_epilogue<R>(I, tile, mden, cnt);
}

template <typename R, bool SUBGRID, int P>
__global__ void __launch_bounds__(BLOCK) k_mom_point(const DevInst* __restrict__ insts, WorkList L) {
  int tile, inst;
  if (!work_item(L, blockIdx.x, gridDim.x, tile, inst)) return;
  const DevInst& I = insts[inst];
  Scalars* Sp = L.direct ? L.sc : I.sc;
  R mden = R(0);
  int cnt[NCNT] = {0, 0, 0, 0, 0, 0};
  if (G(I.mq)[tile] == 0x01010101u && G(I.md)[tile] == 0) {   // quiescent: nothing would change
    if (!G(Sp)->active) return;
    if (I.count) count_tile_points<P>(I, tile, cnt);
    mom_epilogue<R>(I, tile, mden, cnt);
    return;
  }
  bool held;
  if (!mom_point<R, SUBGRID, P>(I, Sp, tile, cnt, mden, held)) return;   // scenario not active
  if (threadIdx.x == 0) G(I.md)[tile] = 0;
  const bool wave_held = __ballot(!held) == 0;
  if (threadIdx.x % WAVE == 0) G((uint8_t*)I.mq)[4 * (uint32_t)tile + threadIdx.x / WAVE] = wave_held ? 1 : 0;
  mom_epilogue<R>(I, tile, mden, cnt);
}

template <typename R, bool SUBGRID>
__global__ void __launch_bounds__(BLOCK) k_mom_strip(const DevInst* __restrict__ insts, WorkList L) {
  int tile, inst;
  if (!work_item(L, blockIdx.x, gridDim.x, tile, inst)) return;
  const DevInst& I = insts[inst];
  Scalars* Sp = L.direct ? L.sc : I.sc;
  if (!G(Sp)->active) return;
  R mden = R(0);
  int cnt[NCNT] = {0, 0, 0, 0, 0, 0};
  mom_strip<R, SUBGRID>(I, tile, cnt, mden);
  mom_epilogue<R>(I, tile, mden, cnt);
}

template <typename R, bool SUBGRID>
__global__ void __launch_bounds__(BLOCK) k_mom_idx(const DevInst* __restrict__ insts, WorkList L) {
  int tile, inst;
  if (!work_item(L, blockIdx.x, gridDim.x, tile, inst)) return;
  const DevInst& I = insts[inst];
  Scalars* Sp = L.direct ? L.sc : I.sc;
  if (!G(Sp)->active) return;
  R mden = R(0);
  int cnt[NCNT] = {0, 0, 0, 0, 0, 0};
  mom_idx<R, SUBGRID>(I, tile, cnt, mden);
  mom_epilogue<R>(I, tile, mden, cnt);
}

// Combined uv-points at refinement transitions (sfincs_momentum.f90:757-779).
template <typename R>
__global__ void __launch_bounds__(BLOCK) k_cuv(const DevInst* __restrict__ insts, WorkList L) {
  int tile, inst;
  if (!work_item(L, blockIdx.x, gridDim.x, tile, inst)) return;
  const DevInst& I = insts[inst];
  Scalars* Sp = L.direct ? L.sc : I.sc;
  if (!G(Sp)->active) return;
  const int j = tile * BLOCK + threadIdx.x;
  if (j >= I.ncuv) return;
  const int cur = G(Sp)->cur, cur1 = cur ^ 1;
  const auto q = G((R*)I.q[cur1]);
  const auto uv = G((R*)I.uv[cur1]);
  const int c = G(I.cuv)[j], a = G(I.cuv)[I.ncuv + j], b = G(I.cuv)[2 * I.ncuv + j];
  // The reference averages inside compute_fluxes, before compute_fluxes_over_structures
  // sets this step's weir discharges, so a weir point (kcuv 3) enters with the last
  // step's value: the other ping-pong buffer. Its uv is 0 in both.
  const auto q_last = G((const R*)I.q[cur]);
  const R qa = G(I.kcuv)[a] == 3 ? q_last[a] : q[a];
  const R qb = G(I.kcuv)[b] == 3 ? q_last[b] : q[b];
  q[c] = (qa + qb) / R(2);
  uv[c] = (uv[a] + uv[b]) / R(2);
}

// ----------------------------------------------------------------------------- continuity
// Bitwise equality: a cell stores a value only when its bits change (dry cells without
// inflow keep every value, so most of the grid writes nothing; storing identical bits
// would change nothing).
__device__ __forceinline__ bool same_bits(double a, double b) { return __double_as_longlong(a) == __double_as_longlong(b); }
__device__ __forceinline__ bool same_bits(float a, float b) { return __float_as_uint(a) == __float_as_uint(b); }

// The net rain of one cell, `dry` telling whether it held no water at the start of the
// step (volume not above zero on subgrid meshes, level not above the bed on regular ones;
// point sources come later, in continuity): gridded rain between its two frames, a
// negative value dropped on a dry cell, then spun up (update_meteo_forcing,
// sfincs_meteo.f90:1309-1330, 1370-1390); less the infiltration rate of a wet cell
// ('con' / 'c2d', update_infiltration_map, sfincs_infiltration.f90:632-663). Without
// either the step's uniform rain np_u. By how much an interior cell's rain volume differs
// from the uniform rain's goes to inf.
template <typename R>
__device__ __forceinline__ R cell_rain(const DevInst& I, int c, int kcs, int iref, R dt, R np_u, bool dry,
                                       const MeteoNow& m, double& inf) {
  const uint32_t f = I.c.flags;
  if (!(f & (sfx::F_MPRCP | sfx::F_INFIL | sfx::F_CNA))) return np_u;
  R np_c = np_u;
  if (f & sfx::F_MPRCP) {
    float p = G(I.mprcp)[m.a + (uint32_t)c] * m.omw + G(I.mprcp)[m.b + (uint32_t)c] * m.w;
    if (p < 0.0f && dry) p = 0.0f;
    if (m.spin) {
      p = p * m.sm;
      if (p < 0.0f && dry) p = 0.0f;
    }
    np_c = (R)p;
  }
  if ((f & sfx::F_INFIL) && !dry) np_c = np_c - (R)G(I.qinf)[c];
  if (f & sfx::F_CNA) {   // update_infiltration_map 'cna' (sfincs_infiltration.f90:678-722), wet or dry
    const float prcp = (float)np_c, S = G(I.qinf)[c], sf = I.c.sfacinf, dtf = (float)dt;
    const float P = G(I.cumprcp)[c] + prcp * dtf;   // update_meteo_forcing (sfincs_meteo.f90:1547-1549)
    const float cinf = G(I.cuminf)[c];
    float qinfmap;
    if (P > sf * S) {
      const float e = P - sf * S;
      const float Qq = e * e / (P + (1.0f - sf) * S);
      qinfmap = ((P - Qq) - cinf) / dtf;
    } else {
      qinfmap = prcp;
    }
    np_c = (R)(prcp - qinfmap);
    G(I.cumprcp)[c] = P;
    G(I.cuminf)[c] = cinf + qinfmap * dtf;
  }
  if (kcs == 1) inf = ((double)(np_u * dt) - (double)(np_c * dt)) * (double)I.c.cell_area[iref - 1];
  return np_c;
}

// Returns the fields it stored, one bit each in NSTORE order.
template <typename R, bool SUBGRID, bool RX>
__device__ __forceinline__ unsigned cont_cell(const DevInst& I, int c, uint8_t kraw, R q_md, R q_mu, R q_nd, R q_nu,
                                              int iref, R dt, R netprcp, const MeteoNow& mt, double& inf) {
  unsigned wrote = 0;
  const int kcs = kraw & KCS_MASK;
  double zs = G(I.zs)[c];
  const auto zsmaxp = G((R*)I.zsmax);
  R zsmax = zsmaxp[c];
  const double zs_in = zs;
  const R zsmax_in = zsmax;
  const R qx = (I.c.flags & sfx::F_QEXT) ? (R)G(I.qext)[c] : R(0);
  if (qx != R(0) && kcs == 1)   // qext volume of the balance (sparse: BMI coupling cells)
    atomicAdd(I.bal + 3, (double)(qx * dt) * (double)I.c.cell_area[iref - 1]);
  if (SUBGRID) {
    double vol = G(I.vol)[c];
    const double vol_in = vol;   // as loaded: a source alone changes the volume too
    if (RX) netprcp = cell_rain<R>(I, c, kcs, iref, dt, netprcp, vol <= 0.0, mt, inf);
    if (kraw & KCS_SOURCE) {   // point sources of this cell, in source order (:311-325)
      double applied = 0.0;
      for (int s = 0; s < I.nsrc + 2 * I.ndrn; ++s)
        if (G(I.src_cell)[s] == c) {
          const float qs = G(I.qtsrc)[s];
          if ((vol >= 0) || ((qs < 0.0f) && (vol >= 0))) {
            vol = vol + (double)((R)qs * dt);
            applied += (double)((R)qs * dt);
          }
        }
      if (applied != 0.0) atomicAdd(I.bal + 1, applied);
    }
    const auto zs0p = G((R*)I.zs0);
    const auto zsdp = G((R*)I.zsderv);
    R zs0 = zs0p[c], zsd = zsdp[c];
    const R zs0_in = zs0, zsd_in = zsd;
    const int nl = I.c.nl;
    R stor = (I.c.flags & sfx::F_STORAGE) ? G((R*)I.storage)[c] : R(0);
    const R stor_in = stor;
    sfx::cell_subgrid<R>(zs, vol, zs0, zsd, zsmax, kcs, q_md, q_mu, q_nd, q_nu, netprcp, qx, stor, iref, G(I.z_zmin)[c],
                         G(I.z_zmax)[c], G(I.z_volmax)[c], sfx::GRow<const float>{G(I.z_dep), (uint32_t)c * nl},
                         I.c, dt);
    if (!same_bits(vol, vol_in)) { G(I.vol)[c] = vol; wrote |= 2u; }
    if (!same_bits(stor, stor_in)) G((R*)I.storage)[c] = stor;
    if (kcs == 1 || kcs == 4) {
      if (!same_bits(zs0, zs0_in)) { zs0p[c] = zs0; wrote |= 4u; }
      if (!same_bits(zsd, zsd_in)) { zsdp[c] = zsd; wrote |= 8u; }
    }
  } else {
    if (RX) netprcp = cell_rain<R>(I, c, kcs, iref, dt, netprcp, zs <= (double)G(I.zb)[c], mt, inf);
    if (kraw & KCS_SOURCE) {   // (:87-106)
      const double before = zs;
      for (int s = 0; s < I.nsrc + 2 * I.ndrn; ++s)
        if (G(I.src_cell)[s] == c) {
          const R inc = (R)G(I.qtsrc)[s] * dt / (R)I.c.cell_area[iref - 1];
          zs = sfx::vmax<double>(zs + (double)inc, (double)G(I.zb)[c]);
        }
      if (zs != before) atomicAdd(I.bal + 1, (zs - before) * (double)I.c.cell_area[iref - 1]);
    }
    sfx::cell_regular<R>(zs, zsmax, kcs, q_md, q_mu, q_nd, q_nu, netprcp, qx, iref, I.c, dt);
  }
  if (!same_bits(zs, zs_in)) { G(I.zs)[c] = zs; wrote |= 1u; }
  if ((I.c.flags & sfx::F_ZSMAX) && !same_bits(zsmax, zsmax_in)) { zsmaxp[c] = zsmax; wrote |= 16u; }
  return wrote;
}

// One cell of continuity with its steady bit. A cell is quiet when it is an ordinary cell
// (kcs 1, no point source), its four face fluxes are zero, no rain falls and it has no
// infiltration rate, and no gridded rain falls on it: its update then depends on its own
// state only (not on dt). A quiet cell whose last update, also
// quiet, changed nothing (KCS_STEADY) would change nothing again, so it is skipped after
// reading its byte and faces; any other cell is updated and its bit set or cleared.
// Boundary cells, whose levels forcing sets, are never steady; a host write of cell
// state clears every bit.
template <typename R, bool SUBGRID, bool RX>
__device__ __forceinline__ unsigned cont_update(const DevInst& I, int c, R q_md, R q_mu, R q_nd, R q_nu, int iref,
                                                R dt, R netprcp, bool dry_sky, const MeteoNow& mt, double& inf) {
  const uint8_t kraw = G(I.kcs)[c];
  const bool quiet = (kraw & (KCS_MASK | KCS_SOURCE)) == 1 && dry_sky && q_md == R(0) && q_mu == R(0) &&
                     q_nd == R(0) && q_nu == R(0) && !(I.c.flags & (sfx::F_STORE | sfx::F_QEXT | sfx::F_CNA)) &&
                     (!RX || ((!(I.c.flags & sfx::F_INFIL) || G(I.qinf)[c] == 0.0f) &&
                              (!(I.c.flags & sfx::F_MPRCP) ||
                               (G(I.mprcp)[mt.a + (uint32_t)c] == 0.0f && G(I.mprcp)[mt.b + (uint32_t)c] == 0.0f))));
  if (quiet && (kraw & KCS_STEADY)) return 32u;   // skipped (counted as such)
  const unsigned w = cont_cell<R, SUBGRID, RX>(I, c, kraw, q_md, q_mu, q_nd, q_nu, iref, dt, netprcp, mt, inf);
  const bool steady = quiet && w == 0u;
  if (steady != ((kraw & KCS_STEADY) != 0))
    G(I.kcs)[c] = (uint8_t)((kraw & ~KCS_STEADY) | (steady ? KCS_STEADY : 0));
  return w;
}

// compute_store_variables (sfincs_continuity.f90:626-720) for one cell after its update,
// fused into continuity (the same faces, the level just computed): for kcs 1 or 4, vmax
// and qmax keep the largest magnitude of the face-averaged uv and q, and twet adds dt while
// the level exceeds the bed (regular) or the cell's lowest subgrid level by
// twet_threshold. A maximum is stored only when it grows.
template <typename R>
__device__ __forceinline__ void store_vars(const DevInst& I, int c, R q_md, R q_mu, R q_nd, R q_nu, R u_md, R u_mu,
                                           R u_nd, R u_nu, R dt) {
  const int kcs = G(I.kcs)[c] & KCS_MASK;
  if (kcs != 1 && kcs != 4) return;
  const uint32_t f = I.c.flags;
  if (f & sfx::F_VELMAX) {
    const R quz = (u_md + u_mu) / R(2), qvz = (u_nd + u_nu) / R(2);
    const R uvz = sfx::rsqrt_<R>(quz * quz + qvz * qvz);
    const auto vm = G((R*)I.vmax);
    if (uvz > vm[c]) vm[c] = uvz;
  }
  if (f & sfx::F_FLUXMAX) {
    const R quz = (q_md + q_mu) / R(2), qvz = (q_nd + q_nu) / R(2);
    const R qz = sfx::rsqrt_<R>(quz * quz + qvz * qvz);
    const auto qm = G((R*)I.qmax);
    if (qz > qm[c]) qm[c] = qz;
  }
  if (f & sfx::F_TWET) {
    const double zs = G(I.zs)[c];   // this thread's own update of the cell, just stored
    const double base = (f & sfx::F_SUBGRID) ? (double)G(I.z_zmin)[c] : (double)G(I.zb)[c];
    if (zs - base > (double)I.c.twet_threshold) {
      const auto tw = G((R*)I.twet);
      tw[c] = tw[c] + dt;
    }
  }
}

// Adds a wave's cells' differences from the uniform rain volume (infiltration, gridded rain)
// to its balance slot (every lane active).
__device__ __forceinline__ void add_infiltrated(const DevInst& I, int tile, double v) {
  for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  if (threadIdx.x % WAVE == 0 && v != 0.0)
    atomicAdd(I.bal + 6 + (int)(((uint32_t)tile * (BLOCK / WAVE) + threadIdx.x / WAVE) & (NINF - 1)), v);
}

// Adds a wave's continuity stores per field to its tile's counters (active lanes only).
__device__ __forceinline__ void count_stores(const DevInst& I, int tile, unsigned wrote) {
  const int first = __ffsll((long long)__ballot(1)) - 1;
  for (int k = 0; k < NSTORE; ++k) {
    const int n = __popcll(__ballot((wrote >> k) & 1u));
    if ((int)(threadIdx.x % WAVE) == first && n)
      atomicAdd((unsigned int*)&G(I.cntstore)[NSTORE * (uint32_t)tile + k], (unsigned int)n);
  }
}

// One block of a continuity launch per scenario: the volume this step's update moves
// through the faces between interior cells and cells that are not interior, from the same
// q buffer and dt as the cells, and the step's rain volume, added in FP64 to bal[0], bal[2].
template <typename R>
__device__ void balance_block(const DevInst& I) {
  const auto S = G(I.sc);
  if (!S->active) return;
  const int qb = __builtin_amdgcn_readfirstlane(S->mom_done ? (S->cur ^ 1) : S->cur);
  const auto q = G((const R*)I.q[qb]);
  const R dt = (R)S->dt;
  double acc = 0.0;
  for (int f = threadIdx.x; f < I.nbf; f += BLOCK) acc += G(I.bf_coef)[f] * (double)q[G(I.bf_slot)[f]];
  for (int o = WAVE / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, WAVE);
  __shared__ double red[BLOCK / WAVE];
  if (threadIdx.x % WAVE == 0) red[threadIdx.x / WAVE] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double sum = 0.0;
    for (int i = 0; i < BLOCK / WAVE; ++i) sum += red[i];
    const auto bal = G(I.bal);
    bal[0] = bal[0] + sum * (double)dt;
    if (I.c.flags & sfx::F_PRECIP) bal[2] = bal[2] + (double)((R)S->netprcp * dt) * I.rain_area;
  }
}

// The first `nahead` blocks compute the next step of scenario ahead_ids[b] of ahead_insts
// (forcing_ahead_block), the next `nbal` blocks the volume balance of scenario
// ahead_ids[b - nahead] (balance_block); the leading blocks are padded to a multiple of 8,
// so the cells keep their XCDs. RX: some scenario of the launch has infiltration or
// gridded rain (per-cell rain and its balance slots); the instantiation without it
// carries none of that code.
template <typename R, bool RX>
__global__ void __launch_bounds__(BLOCK) k_continuity(const DevInst* __restrict__ insts, WorkList L, int indexed,
                                                      int nahead, int nbal, const DevInst* __restrict__ ahead_insts,
                                                      const int32_t* __restrict__ ahead_ids) {
  const uint32_t lead = (uint32_t)(nahead + nbal + 7) / 8 * 8;
  if (blockIdx.x < lead) {
    if ((int)blockIdx.x < nahead) forcing_ahead_block(ahead_insts[ahead_ids[blockIdx.x]]);
    else if ((int)blockIdx.x < nahead + nbal) balance_block<R>(ahead_insts[ahead_ids[blockIdx.x - nahead]]);
    return;
  }
  int tile, inst;
  if (!work_item(L, blockIdx.x - lead, gridDim.x - lead, tile, inst)) return;
  const DevInst& I = insts[inst];
  Scalars* Sp = L.direct ? L.sc : I.sc;
  const auto S = G(Sp);
  if (!S->active) return;
  const int qb = __builtin_amdgcn_readfirstlane(S->mom_done ? (S->cur ^ 1) : S->cur);
  const auto q = G((const R*)I.q[qb]);
  const R dt = (R)S->dt;
  const R netprcp = (R)S->netprcp;
  const bool dry_sky = !(I.c.flags & sfx::F_PRECIP) || netprcp == R(0);
  double inf = 0.0;   // uniform rain volume less what this thread's cell receives
  MeteoNow mt{};
  if (RX) mt = meteo_now(I, S);
  if (!indexed) {
    const int nmax = I.nmax, mmax = I.mmax, np = I.np;
    const int tn = tile % I.ntn, tmi = tile / I.ntn;
    const int n = tn * I.tn + (threadIdx.x & (I.tn - 1));
    const int m = tmi * I.tm + (threadIdx.x >> I.tn_sh);
    if (n < nmax && m < mmax) {
    const int c = m * nmax + n;
    const R q_md = m >= 1 ? q[c - nmax] : R(0);
    const R q_mu = m < mmax - 1 ? q[c] : R(0);
    const R q_nd = n >= 1 ? q[np + c - 1] : R(0);
    const R q_nu = n < nmax - 1 ? q[np + c] : R(0);
    const unsigned w = I.subgrid ? cont_update<R, true, RX>(I, c, q_md, q_mu, q_nd, q_nu, 1, dt, netprcp, dry_sky, mt, inf)
                                 : cont_update<R, false, RX>(I, c, q_md, q_mu, q_nd, q_nu, 1, dt, netprcp, dry_sky, mt, inf);
    if (I.count) count_stores(I, tile, w);
    if (I.c.flags & sfx::F_STORE) {
      R u_md = R(0), u_mu = R(0), u_nd = R(0), u_nu = R(0);
      if (I.c.flags & sfx::F_VELMAX) {
        const auto uv = G((const R*)I.uv[qb]);
        u_md = m >= 1 ? uv[c - nmax] : R(0);
        u_mu = m < mmax - 1 ? uv[c] : R(0);
        u_nd = n >= 1 ? uv[np + c - 1] : R(0);
        u_nu = n < nmax - 1 ? uv[np + c] : R(0);
      }
      store_vars<R>(I, c, q_md, q_mu, q_nd, q_nu, u_md, u_mu, u_nd, u_nu, dt);
    }
    if (I.sk < 0) mark_wave(I, tile, (w & 1u) != 0);
    }
  } else {
    const int c = tile * BLOCK + threadIdx.x;
    if (c < I.np) {
    const R q_md = q[G(I.zmd)[c]], q_mu = q[G(I.zmu)[c]], q_nd = q[G(I.znd)[c]], q_nu = q[G(I.znu)[c]];
    const int iref = G(I.z_iref)[c];
    const unsigned w = I.subgrid ? cont_update<R, true, RX>(I, c, q_md, q_mu, q_nd, q_nu, iref, dt, netprcp, dry_sky, mt, inf)
                                 : cont_update<R, false, RX>(I, c, q_md, q_mu, q_nd, q_nu, iref, dt, netprcp, dry_sky, mt, inf);
    if (I.count) count_stores(I, tile, w);
    if (I.c.flags & sfx::F_STORE) {
      R u_md = R(0), u_mu = R(0), u_nd = R(0), u_nu = R(0);
      if (I.c.flags & sfx::F_VELMAX) {
        const auto uv = G((const R*)I.uv[qb]);
        u_md = uv[G(I.zmd)[c]]; u_mu = uv[G(I.zmu)[c]]; u_nd = uv[G(I.znd)[c]]; u_nu = uv[G(I.znu)[c]];
      }
      store_vars<R>(I, c, q_md, q_mu, q_nd, q_nu, u_md, u_mu, u_nd, u_nu, dt);
    }
    }
  }
  if (RX && (I.c.flags & (sfx::F_INFIL | sfx::F_MPRCP | sfx::F_CNA))) add_infiltrated(I, tile, inf);   // every lane of the block is here
}

// FP64 volume accumulator: sum over active cells of V (subgrid) or zs*A (regular).
__global__ void __launch_bounds__(BLOCK) k_volume(const DevInst* __restrict__ I0, double* out) {
  const DevInst& I = *I0;
  double acc = 0.0;
  for (int c = blockIdx.x * BLOCK + threadIdx.x; c < I.np; c += gridDim.x * BLOCK) {
    const int k = G(I.kcs)[c] & KCS_MASK;
    if (k == 0) continue;
    const int iref = I.dense ? 1 : G(I.z_iref)[c];
    if (I.subgrid) acc += G(I.vol)[c];
    else acc += G(I.zs)[c] * (double)I.c.cell_area[iref - 1];
  }
  for (int o = WAVE / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, WAVE);
  __shared__ double red[BLOCK / WAVE];
  if (threadIdx.x % WAVE == 0) red[threadIdx.x / WAVE] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int i = 0; i < BLOCK / WAVE; ++i) s += red[i];
    atomicAdd(out, s);
  }
}

__global__ void k_zero_f64(double* p) { *p = 0.0; }

// Volume balance sums: bal[4] = interior volume (V, or zs*A, over kcs 1 cells), bal[5] =
// storage volume left; both set to zero before the launch.
__global__ void __launch_bounds__(BLOCK) k_balance_volume(const DevInst* __restrict__ I0) {
  const DevInst& I = *I0;
  double v = 0.0, st = 0.0;
  for (int c = blockIdx.x * BLOCK + threadIdx.x; c < I.np; c += gridDim.x * BLOCK) {
    if ((G(I.kcs)[c] & KCS_MASK) != 1) continue;
    const int iref = I.dense ? 1 : G(I.z_iref)[c];
    v += I.subgrid ? G(I.vol)[c] : G(I.zs)[c] * (double)I.c.cell_area[iref - 1];
    if (I.c.flags & sfx::F_STORAGE)
      st += I.fp64 ? G((const double*)I.storage)[c] : (double)G((const float*)I.storage)[c];
  }
  for (int o = WAVE / 2; o > 0; o >>= 1) { v += __shfl_xor(v, o, WAVE); st += __shfl_xor(st, o, WAVE); }
  __shared__ double red[2][BLOCK / WAVE];
  if (threadIdx.x % WAVE == 0) { red[0][threadIdx.x / WAVE] = v; red[1][threadIdx.x / WAVE] = st; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0;
    for (int i = 0; i < BLOCK / WAVE; ++i) { a += red[0][i]; b += red[1][i]; }
    atomicAdd(I.bal + 4, a);
    if (I.c.flags & sfx::F_STORAGE) atomicAdd(I.bal + 5, b);
  }
}

// compute_zbuvmx (sfincs_domain.f90:1746-1762): zbuvmx = max(zb(nm), zb(nmu)) + huthresh for
// every uv-point from the bed levels on the device (the uv-points of a dense mesh are the
// slots whose right / upper cell exists).
__global__ void __launch_bounds__(BLOCK) k_zbuvmx(const DevInst* __restrict__ I0) {
  const DevInst& I = *I0;
  float* zbuvmx = (float*)I.zbuvmx;
  const float huth = I.c.huthresh;
  for (int s = blockIdx.x * BLOCK + threadIdx.x; s < (I.dense ? 2 * I.np : I.npuv); s += gridDim.x * BLOCK) {
    int nm, nmu;
    if (I.dense) {
      const int c = s < I.np ? s : s - I.np;
      const int n = c % I.nmax, m = c / I.nmax;
      if (s < I.np) { if (m >= I.mmax - 1) continue; nm = c; nmu = c + I.nmax; }
      else { if (n >= I.nmax - 1) continue; nm = c; nmu = c + 1; }
    } else {
      nm = G(I.nb)[s];
      nmu = G(I.nb)[(uint32_t)I.npuv + s];
    }
    zbuvmx[s] = sfx::vmax<float>(G(I.zb)[nm], G(I.zb)[nmu]) + huth;
  }
}

__global__ void k_zero_bal(double* p, int from, int to) {
  for (int k = from + (int)threadIdx.x; k < to; k += blockDim.x) p[k] = 0.0;
}

}  // namespace

// =============================================================================== host side
struct sfx_handle {
  int device = 0;
  bool tsa_paused = false;      // timestep analysis off while sweeps are only timed (autotune)
  int fp64 = 0;
  sfx_sizes sz{};
  DevInst hI{};
  DevInst* dI = nullptr;
  Scalars* dS = nullptr;
  std::vector<void*> allocs;
  std::vector<int32_t> ref2slot;   // reference uv index (0-based, incl. cuv and dummy) -> slot
  int32_t nslot = 0;
  hipStream_t stream = nullptr;
  struct sfx_batch* solo = nullptr;
  // tuned shapes: continuity tile, momentum per-cell tile, momentum strip rows (0: per-cell)
  int tile_tn = CELL_SHAPES[0][0], tile_tm = CELL_SHAPES[0][1];
  int mtile_tn = 32, mtile_tm = 8, strip_rows = -1;   // per-point kernel on 32 x 8 tiles
  int xchunk = 64, xchunk_c = 0;
  size_t nmt_cap = 0;            // capacity of blkmin (tiles of the largest tile count)
  size_t nct_cap = 0;            // continuity tiles of the largest tile count (cntstore)
  int in_batches = 0;            // user batches holding this scenario (tile shape is frozen)
  double* dvol = nullptr;
  template <typename T> T* upload(const T* src, size_t n);
  template <typename T> T* alloc(size_t n);
};

template <typename T> T* sfx_handle::alloc(size_t n) {
  void* p = nullptr;
  if (n == 0) n = 1;
  if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return nullptr;
  (void)hipMemset(p, 0, n * sizeof(T));
  allocs.push_back(p);
  return (T*)p;
}

template <typename T> T* sfx_handle::upload(const T* src, size_t n) {
  T* p = alloc<T>(n);
  if (p && n && src) (void)hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice);
  return p;
}

// Momentum work lists of a batch, one per (kernel, mesh kind): [cell sg, cell reg, strip sg,
// strip reg, idx sg, idx reg, point P=1 sg, reg, point P=2 sg, reg, point P=4 sg, reg].
constexpr int NLIST_M = 12;

// Momentum list of a scenario's kernel choice.
inline int mom_list(const DevInst& I) {
  int k;
  if (!I.dense) k = 4;
  else if (I.sk > 0) k = 2;
  else if (I.sk == 0) k = 0;
  else k = 6 + 2 * ilog2(-I.sk);
  return k + (I.subgrid ? 0 : 1);
}

struct sfx_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = true;
  std::vector<sfx_handle*> members;
  std::vector<double> tend;
  std::vector<char> live;        // members short of their end time; finished ones leave the work lists
  int n_live = 0;
  DevInst* dInsts = nullptr;
  size_t dInsts_cap = 0;
  Item* dItems = nullptr;
  size_t items_cap = 0;
  int32_t* dIds = nullptr;
  size_t ids_cap = 0;
  // launch descriptors (offsets into dItems)
  // work lists: momentum dense/indexed x subgrid/regular, combined points, continuity
  // [cell sg, cell reg, strip sg, strip reg, idx sg, idx reg, point sg, point reg]
  size_t off_m[NLIST_M] = {}, n_m[NLIST_M] = {};
  WorkList wl_m[NLIST_M] = {}, wl_cuv = {}, wl_cd = {}, wl_ci = {};   // launch arguments of the lists
  size_t off_cuv = 0, n_cuv = 0, off_cd = 0, n_cd = 0, off_ci = 0, n_ci = 0;
  bool dirty = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_steps = 0;
  int fp64 = 0;
};

namespace {

uint32_t build_flags(const sfx_model* M) {
  const sfx_params& p = M->p;
  uint32_t f = 0;
  if (p.advection) f |= sfx::F_ADVECTION;
  if (p.advection_scheme == 1) f |= sfx::F_UPW1;
  if (p.viscosity) f |= sfx::F_VISCOSITY;
  if (M->coriolis) f |= sfx::F_CORIOLIS;
  if (p.friction2d) f |= sfx::F_FRICTION2D;
  if (M->thetasmoothing) f |= sfx::F_THETA;
  if (p.wiggle_suppression) f |= sfx::F_WIGGLE;
  if (p.store_zsmax) f |= sfx::F_ZSMAX;
  if (p.store_velmax) f |= sfx::F_VELMAX;
  if (p.store_fluxmax) f |= sfx::F_FLUXMAX;
  if (p.store_twet) f |= sfx::F_TWET;
  if (p.timestep_analysis) f |= sfx::F_TSA;
  if (p.use_qext) f |= sfx::F_QEXT;
  if (M->subgrid && !M->storage_volume.empty()) f |= sfx::F_STORAGE;
  if (M->wind) f |= sfx::F_WIND;
  if (M->precip) f |= sfx::F_PRECIP;
  if (M->precip && !M->qinffield.empty()) f |= sfx::F_INFIL;
  if (M->precip && !M->scs_m.empty()) f |= sfx::F_CNA;
  if (M->meteo3d) f |= sfx::F_METEO3D;
  if (M->meteo3d && M->meteo_wind) f |= sfx::F_MWIND;
  if (M->meteo3d && M->patmos) f |= sfx::F_PATM;
  if (M->meteo3d && M->meteo_rain) f |= sfx::F_MPRCP;
  if (M->meteo3d && p.spinup_meteo) f |= sfx::F_MSPIN;
  if (p.h73table) f |= sfx::F_H73;
  if (M->use_quadtree) f |= sfx::F_QUADTREE;
  if (M->subgrid) f |= sfx::F_SUBGRID;
  if (p.advection || M->coriolis || p.viscosity || p.friction2d || M->thetasmoothing) f |= sfx::F_NBRS;
  return f;
}

int fill_consts(const sfx_model* M, Consts& c) {
  const sfx_params& p = M->p;
  if (M->nref > sfx::MAXREF) { set_error("more than 8 refinement levels"); return SFX_ERR_ARG; }
  std::memset(&c, 0, sizeof(c));
  c.g = p.g; c.huthresh = p.huthresh; c.huvmin = p.huvmin; c.advlim = p.advlim; c.slopelim = p.slopelim;
  c.uvlim = p.uvlim; c.nuviscfac = p.nuviscfac; c.fcorio = M->fcorio; c.theta = p.theta;
  c.wiggle_factor = p.wiggle_factor; c.wiggle_threshold = p.wiggle_threshold; c.dtmax = M->dtmax; c.dtmin = M->dtmin;
  c.alpha = p.alpha; c.rhoa = p.rhoa; c.rhow = p.rhow; c.btfilter = p.btfilter; c.btrelax = p.btrelax;
  c.t0 = p.t0; c.t1 = p.t1; c.tspinup = p.t0 + p.tspinup; c.zini = p.zsini; c.dx = M->dxrm[0]; c.dy = M->dyrm[0];
  c.pi = p.pi; c.twet_threshold = p.twet_threshold;
  c.dtwnd = p.dtwnd; c.gapres = p.gapres; c.pavbnd = p.pavbnd; c.sfacinf = p.sfacinf;
  for (int i = 0; i < M->nref; ++i) {
    c.dxrinv[i] = M->dxrinv[i]; c.dyrinv[i] = M->dyrinv[i]; c.dxr2inv[i] = M->dxr2inv[i]; c.dyr2inv[i] = M->dyr2inv[i];
    c.dxrinvc[i] = M->dxrinvc[i]; c.dyrinvc[i] = M->dyrinvc[i]; c.cell_area[i] = M->cell_area[i];
    c.dxrm[i] = M->dxrm[i]; c.dyrm[i] = M->dyrm[i]; c.nuvisc[i] = M->nuvisc[i];
  }
  c.flags = build_flags(M);
  c.nl = M->nl;
  return SFX_OK;
}

template <typename R> std::vector<R> to_r(const std::vector<float>& v) { return std::vector<R>(v.begin(), v.end()); }

template <typename R>
int upload_state(sfx_handle* H, const sfx_model* M) {
  DevInst& I = H->hI;
  const int np = M->np;
  // q / uv double buffers in slot layout
  std::vector<R> qs(H->nslot + 1, R(0)), uvs(H->nslot + 1, R(0));
  for (size_t k = 0; k < H->ref2slot.size(); ++k) {
    const int s = H->ref2slot[k];
    if (s >= 0 && k < M->q.size()) { qs[s] = M->q[k]; uvs[s] = M->uv[k]; }
  }
  for (int b = 0; b < 2; ++b) { I.q[b] = H->upload(qs.data(), qs.size()); I.uv[b] = H->upload(uvs.data(), uvs.size()); }
  auto zs0 = to_r<R>(M->zs0); auto zsd = to_r<R>(M->zsderv); auto zsm = to_r<R>(M->zsmax);
  I.zs0 = H->upload(zs0.data(), np); I.zsderv = H->upload(zsd.data(), np); I.zsmax = H->upload(zsm.data(), np);
  const uint32_t f = build_flags(M);
  if (f & sfx::F_VELMAX) { auto v = to_r<R>(M->vmax); I.vmax = H->upload(v.data(), np); }
  if (f & sfx::F_FLUXMAX) { auto v = to_r<R>(M->qmax); I.qmax = H->upload(v.data(), np); }
  if (f & sfx::F_TWET) { auto v = to_r<R>(M->twet); I.twet = H->upload(v.data(), np); }
  if (f & sfx::F_TSA) {   // initialize_timestep_analysis (sfincs_timestep_analysis.f90:10-22)
    const size_t ns = (size_t)H->nslot + 1;
    std::vector<float> req(ns, M->dtmax), zero(ns, 0.f);
    std::vector<int32_t> izero(ns, 0);
    I.tsa_req = H->upload(req.data(), ns); I.tsa_sum = H->upload(zero.data(), ns);
    I.tsa_wet = H->upload(izero.data(), ns); I.tsa_lim = H->upload(izero.data(), ns);
    I.tsa_min = H->upload(zero.data(), 1);
  }
  if (f & sfx::F_STORAGE) { auto v = to_r<R>(M->storage_volume); I.storage = H->upload(v.data(), np); }
  if (f & sfx::F_INFIL) I.qinf = H->upload(M->qinffield.data(), np);
  if (f & sfx::F_CNA) I.qinf = H->upload(M->scs_m.data(), np);
  if (f & sfx::F_CNA) {
    std::vector<float> z(np, 0.f);
    I.cumprcp = H->upload(z.data(), np);
    I.cuminf = H->upload(z.data(), np);
  }
  if (f & sfx::F_METEO3D) {
    I.mnframes = M->meteo_nframes;
    if (f & sfx::F_MWIND) { I.mtauu = H->upload(M->meteo_tauwu.data(), M->meteo_tauwu.size());
                            I.mtauv = H->upload(M->meteo_tauwv.data(), M->meteo_tauwv.size()); }
    if (f & sfx::F_PATM) I.mpatm = H->upload(M->meteo_patm.data(), M->meteo_patm.size());
    if (f & sfx::F_MPRCP) I.mprcp = H->upload(M->meteo_prcp.data(), M->meteo_prcp.size());
  }
  if (f & sfx::F_QEXT) {   // sfincs_domain.f90:2383-2389
    std::vector<float> z(np, 0.f);
    I.qext = H->upload(z.data(), np);
  }
  auto um = to_r<R>(M->uvmean);
  I.uvmean = H->upload(um.data(), um.size());
  return SFX_OK;
}

}  // namespace

// -------------------------------------------------------------------------- create / destroy
extern "C" int sfx_model_footprint(const sfx_model* M, size_t* bytes) {
  if (!M || !bytes) { set_error("sfx_model_footprint: null"); return SFX_ERR_ARG; }
  const size_t np = M->np, nslot = M->dense ? 2 * (size_t)M->np : (size_t)(M->npuv + M->ncuv);
  size_t b = np * (8 + 4 * 3 + 1) + nslot * (4 * 4 + 3);
  if (M->subgrid) b += np * (8 + 12 + 4 * (size_t)M->nl) + nslot * (4 + 16 + 4 + 16 * (size_t)M->nl);
  else b += np * 4 + nslot * 8;
  if (!M->dense) b += (size_t)M->npuv * (40 + 1) + np * 17 + (size_t)M->ncuv * 12;
  *bytes = b + (1 << 20);
  return SFX_OK;
}

extern "C" int sfx_create_ex(const sfx_model* M, int32_t device, int32_t fp64, sfx_handle** out);

extern "C" int sfx_create(const sfx_model* M, int32_t device, sfx_handle** out) {
  return sfx_create_ex(M, device, 0, out);
}

extern "C" int sfx_create_ex(const sfx_model* M, int32_t device, int32_t fp64, sfx_handle** out) {
  if (!M || !out || !M->initialized) { set_error("sfx_create: model not initialized"); return SFX_ERR_STATE; }
  if (M->p.bndtype != 1) { set_error("only bndtype=1 (weakly reflective) boundaries are supported"); return SFX_ERR_ARG; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { set_error("no HIP device available"); return SFX_ERR_HIP; }
  if (device < 0 || device >= ndev) { set_error("bad device index"); return SFX_ERR_ARG; }
  HIPCHK(hipSetDevice(device));
  sfx_handle* H = new sfx_handle();
  H->device = device;
  H->fp64 = fp64 ? 1 : 0;
  sfx_model_sizes(M, &H->sz);
  DevInst& I = H->hI;
  std::memset(&I, 0, sizeof(I));
  int rc = fill_consts(M, I.c);
  if (rc) { delete H; return rc; }
  const int np = M->np, npuv = M->npuv, ncuv = M->ncuv, nl = M->nl;
  I.dense = M->dense; I.subgrid = M->subgrid; I.fp64 = H->fp64;
  I.np = np; I.npuv = npuv; I.ncuv = ncuv; I.nmax = M->nmax; I.mmax = M->mmax; I.nl = nl;
  // slot mapping
  H->ref2slot.assign(npuv + ncuv + 1, -1);
  if (M->dense) {
    H->nslot = 2 * np;
    for (int ip = 0; ip < npuv; ++ip) {
      const int c = M->uv_index_z_nm[ip] - 1;
      H->ref2slot[ip] = M->uv_flags_dir[ip] == 0 ? c : np + c;
    }
    H->ref2slot[npuv + ncuv] = H->nslot;
  } else {
    H->nslot = npuv + ncuv;
    for (int k = 0; k < npuv + ncuv + 1; ++k) H->ref2slot[k] = k;
  }
  I.nslot = H->nslot;
  const int ns = H->nslot + 1;
  {   // kernels index every array with 32-bit byte offsets
    const size_t big = std::max({(size_t)ns * (size_t)std::max(nl, 1) * 16, (size_t)np * (size_t)std::max(nl, 1) * 4,
                                 (size_t)npuv * 40, (size_t)ns * 8, (size_t)np * 8});
    if (big >= ((size_t)1 << 32)) {
      set_error("sfx_create: a device array of this mesh would exceed 4 GiB; split the domain");
      delete H;
      return SFX_ERR_ARG;
    }
  }
  auto perm8 = [&](const std::vector<int8_t>& src, int8_t fill) {
    std::vector<int8_t> d(ns, fill);
    for (size_t k = 0; k < src.size() && k < H->ref2slot.size(); ++k) if (H->ref2slot[k] >= 0) d[H->ref2slot[k]] = src[k];
    return d;
  };
  auto permf = [&](const std::vector<float>& src) {
    std::vector<float> d(ns, 0.f);
    for (size_t k = 0; k < src.size() && k < H->ref2slot.size(); ++k) if (H->ref2slot[k] >= 0) d[H->ref2slot[k]] = src[k];
    return d;
  };
  // cells
  I.zs = H->upload(M->zs.data(), np);
  if (M->subgrid) I.vol = H->upload(M->z_volume.data(), np);
  {
    std::vector<uint8_t> k(np);
    for (int c = 0; c < np; ++c) k[c] = (uint8_t)M->kcs[c];
    for (int s = 0; s < M->nsrc + 2 * M->ndrn; ++s) k[M->nmindsrc[s] - 1] |= KCS_SOURCE;
    I.kcs = H->upload(k.data(), np);
  }
  if (!M->subgrid) I.zb = H->upload(M->zb.data(), np);
  if (M->subgrid) {
    I.z_zmin = H->upload(M->z_zmin.data(), np); I.z_zmax = H->upload(M->z_zmax.data(), np);
    I.z_volmax = H->upload(M->z_volmax.data(), np); I.z_dep = H->upload(M->z_dep.data(), (size_t)np * nl);
  }
  if (!M->dense) {
    I.z_iref = H->upload(M->z_flags_iref.data(), np);
    auto conv = [&](const std::vector<int32_t>& v) { std::vector<int32_t> d(v.size()); for (size_t i = 0; i < v.size(); ++i) d[i] = v[i] - 1; return d; };
    auto a = conv(M->z_index_uv_md), b = conv(M->z_index_uv_mu), c2 = conv(M->z_index_uv_nd), d = conv(M->z_index_uv_nu);
    I.zmd = H->upload(a.data(), np); I.zmu = H->upload(b.data(), np); I.znd = H->upload(c2.data(), np); I.znu = H->upload(d.data(), np);
    std::vector<int32_t> nb((size_t)10 * npuv);
    const std::vector<int32_t>* tabs[10] = {&M->uv_index_z_nm, &M->uv_index_z_nmu, &M->uv_index_u_nmd, &M->uv_index_u_nmu,
                                             &M->uv_index_u_ndm, &M->uv_index_u_num, &M->uv_index_v_ndm, &M->uv_index_v_nm,
                                             &M->uv_index_v_ndmu, &M->uv_index_v_nmu};
    for (int t = 0; t < 10; ++t)
      for (int i = 0; i < npuv; ++i) nb[(size_t)t * npuv + i] = (*tabs[t])[i] - 1;
    I.nb = H->upload(nb.data(), nb.size());
    std::vector<uint8_t> f(npuv);
    for (int i = 0; i < npuv; ++i)
      f[i] = (uint8_t)((M->uv_flags_iref[i] & 15) | (((M->uv_flags_type[i] + 1) & 3) << 4) | ((M->uv_flags_dir[i] & 1) << 6));
    I.uvf = H->upload(f.data(), npuv);
    std::vector<int32_t> cv((size_t)3 * ncuv);
    for (int j = 0; j < ncuv; ++j) {
      cv[j] = M->cuv_index_uv[j] - 1; cv[ncuv + j] = M->cuv_index_uv1[j] - 1; cv[2 * ncuv + j] = M->cuv_index_uv2[j] - 1;
    }
    I.cuv = H->upload(cv.data(), cv.size());
  }
  // uv slots
  {
    auto kc = perm8(M->kcuv, 0);
    I.kcuv = H->upload(kc.data(), ns);
    if (!M->structure_uv_index.empty()) {   // weir uv-points: slot -> structure
      std::vector<int32_t> so(ns, -1);
      std::vector<float4> sp(M->structure_uv_index.size());
      for (size_t k = 0; k < sp.size(); ++k) {
        so[H->ref2slot[M->structure_uv_index[k] - 1]] = (int32_t)k;
        sp[k] = make_float4(M->structure_height[k], M->structure_cd[k], M->structure_length[k], 0.f);
      }
      I.struc_of = H->upload(so.data(), ns);
      I.struc_par = H->upload(sp.data(), sp.size());
    }
    auto ma = perm8(M->mask_adv, 0);
    I.mask_adv = H->upload(ma.data(), ns);
    auto kf = perm8(M->kfuv, 0);
    I.kfuv = H->upload(kf.data(), ns);
  }
  if (!M->subgrid) {
    auto a = permf(M->zbuvmx), b = permf(M->gn2uv);
    I.zbuvmx = H->upload(a.data(), ns); I.gn2uv = H->upload(b.data(), ns);
  } else {
    auto zmn = permf(M->uv_zmin);
    I.uv_zmin = H->upload(zmn.data(), ns);
    std::vector<float4> s4(ns, make_float4(0, 0, 0, 0));
    std::vector<float> ff(ns, 0.f);
    std::vector<float4> lev((size_t)ns * nl, make_float4(0, 0, 0, 0));
    for (int ip = 0; ip < npuv; ++ip) {
      const int s = H->ref2slot[ip];
      s4[s] = make_float4(M->uv_zmax[ip], M->uv_havg_zmax[ip], M->uv_nrep_zmax[ip], M->uv_navg_w[ip]);
      ff[s] = M->uv_fnfit[ip];
      for (int l = 0; l < nl; ++l)
        lev[(size_t)s * nl + l] = make_float4(M->uv_havg[(size_t)ip * nl + l], M->uv_nrep[(size_t)ip * nl + l],
                                              M->uv_pwet[(size_t)ip * nl + l], 0.f);
    }
    I.uv_s4 = H->upload(s4.data(), ns); I.uv_fnfit = H->upload(ff.data(), ns);
    I.uv_lev = H->upload(lev.data(), lev.size());
  }
  // boundaries
  I.ngbnd = M->ngbnd; I.nkcuv2 = M->nkcuv2; I.nbnd = M->nbnd; I.ntbnd = M->ntbnd;
  {
    std::vector<int32_t> cell(M->ngbnd), nmi(M->ngbnd);
    for (int i = 0; i < M->ngbnd; ++i) { cell[i] = M->nmindbnd[i] - 1; nmi[i] = M->nmi_gbp[i] - 1; }
    I.gb_cell = H->upload(cell.data(), cell.size()); I.gb_nmi = H->upload(nmi.data(), nmi.size());
    I.gb_ind1 = H->upload(M->ind1_bnd_gbp.data(), M->ngbnd); I.gb_ind2 = H->upload(M->ind2_bnd_gbp.data(), M->ngbnd);
    I.gb_fac = H->upload(M->fac_bnd_gbp.data(), M->ngbnd);
    std::vector<int32_t> rcell(M->ngbnd, -1);
    std::vector<float> rdzs(M->ngbnd, 0.f);
    for (int i = 0; i < M->ngbnd; ++i)
      if (M->index_bdr_gbp[i] > 0) {
        rcell[i] = M->index_zsi_bdr[M->index_bdr_gbp[i] - 1] - 1;
        rdzs[i] = M->dzs_bdr[M->index_bdr_gbp[i] - 1];
      }
    I.gb_rcell = H->upload(rcell.data(), rcell.size()); I.gb_rdzs = H->upload(rdzs.data(), rdzs.size());
    I.zsb = H->upload(M->zsb.data(), M->ngbnd); I.zsb0 = H->upload(M->zsb0.data(), M->ngbnd);
    std::vector<int32_t> slot(M->nkcuv2), nmb(M->nkcuv2), nmi2(M->nkcuv2), ib(M->nkcuv2);
    std::vector<float> zbuv(M->nkcuv2, 0.f);
    for (int i = 0; i < M->nkcuv2; ++i) {
      slot[i] = H->ref2slot[M->index_kcuv2[i] - 1]; nmb[i] = M->nmbkcuv2[i] - 1; nmi2[i] = M->nmikcuv2[i] - 1;
      ib[i] = M->ibkcuv2[i] - 1;
      if (!M->subgrid) zbuv[i] = M->zbuv[M->index_kcuv2[i] - 1];
    }
    I.k2_slot = H->upload(slot.data(), slot.size()); I.k2_nmb = H->upload(nmb.data(), nmb.size());
    I.k2_nmi = H->upload(nmi2.data(), nmi2.size()); I.k2_ib = H->upload(ib.data(), ib.size());
    I.k2_dir = H->upload(M->ibuvdir.data(), M->ibuvdir.size()); I.k2_zbuv = H->upload(zbuv.data(), zbuv.size());
    I.t_bnd = H->upload(M->t_bnd.data(), M->t_bnd.size()); I.zs_bnd = H->upload(M->zs_bnd.data(), M->zs_bnd.size());
    I.zsi_bnd = M->zsi_bnd.empty() ? nullptr : H->upload(M->zsi_bnd.data(), M->zsi_bnd.size());
    // Neumann copies (kcs 6) run before the flux points only when they share a cell
    std::vector<uint8_t> fw(M->np, 0), fr(M->np, 0);
    for (int i = 0; i < M->nkcuv2; ++i)
      if (M->kcs[nmb[i]] != 6) {
        fw[nmb[i]] = 1; fr[nmi2[i]] = 1;
        if (ib[i] >= 0 && rcell[ib[i]] >= 0) fr[rcell[ib[i]]] = 1;   // a river cell's interior level
      }
    I.bnd_overlap = 0;
    for (int i = 0; i < M->ngbnd; ++i)
      if (M->kcs[cell[i]] == 6 && (fr[cell[i]] || (nmi[i] >= 0 && fw[nmi[i]]))) I.bnd_overlap = 1;
    std::vector<int32_t> g6;
    for (int i = 0; i < M->ngbnd; ++i)
      if (M->kcs[cell[i]] == 6) { g6.push_back(cell[i]); g6.push_back(nmi[i]); }
    I.n6 = (int32_t)(g6.size() / 2);
    I.gb6 = g6.empty() ? nullptr : H->upload(g6.data(), g6.size());
    std::vector<int32_t> fl;
    for (int i = 0; i < M->nkcuv2; ++i)
      if (M->kcs[nmb[i]] != 6) fl.push_back(i);
    I.nflux = (int32_t)fl.size();
    I.k2_flux = fl.empty() ? nullptr : H->upload(fl.data(), fl.size());
  }
  I.nsrc = M->nsrc; I.ntsrc = M->ntsrc; I.ntprcp = M->ntprcp; I.ntwnd = M->ntwnd;
  I.ndrn = M->ndrn; I.structure_relax = M->p.structure_relax;
  {
    const int nsd = M->nsrc + 2 * M->ndrn;
    std::vector<int32_t> sc(nsd);
    for (int s = 0; s < nsd; ++s) sc[s] = M->nmindsrc[s] - 1;
    I.src_cell = H->upload(sc.data(), sc.size()); I.tsrc = H->upload(M->tsrc.data(), M->tsrc.size());
    I.qsrc = H->upload(M->qsrc.data(), M->qsrc.size()); I.qtsrc = H->alloc<float>(std::max(1, nsd));
    I.drn_type = H->upload(M->drainage_type.data(), M->drainage_type.size());
    I.drn_par = H->upload(M->drainage_params.data(), M->drainage_params.size());
    I.drn_dist = H->upload(M->drainage_distance.data(), M->drainage_distance.size());
    std::vector<int32_t> st(M->ndrn, 1);           // open, fully (sfincs_discharges.f90:189-190)
    std::vector<float> fr(M->ndrn, 1.0f);
    I.drn_status = H->upload(st.data(), st.size());
    I.drn_frac = H->upload(fr.data(), fr.size());
    I.qtsrc_nx = H->alloc<float>(std::max(1, M->nsrc)); I.snext = H->alloc<StepImage>(1);
    I.tprcpt = H->upload(M->tprcpt.data(), M->tprcpt.size()); I.tprcpv = H->upload(M->tprcpv.data(), M->tprcpv.size());
    I.twnd = H->upload(M->twnd.data(), M->twnd.size()); I.wndmag = H->upload(M->wndmag.data(), M->wndmag.size());
    I.wnddir = H->upload(M->wnddir.data(), M->wnddir.size()); I.cdval = H->upload(M->cdval.data(), M->cdval.size());
    I.x73 = H->upload(M->x73.data(), M->x73.size());
    std::string perr;
    if (!sfx::pow_tables_device(device, I.pw, perr)) { set_error(perr); delete H; return SFX_ERR_STATE; }
  }
  rc = H->fp64 ? upload_state<double>(H, M) : upload_state<float>(H, M);
  // scalars
  Scalars s{};
  s.t = (double)M->p.t0; s.tend = (double)M->p.t1;
  s.dt = 1.0e-6f; s.min_dt = 1.0e-6f;   // sfincs_lib.f90:281-282
  s.itbndlast = 1; s.itsrclast = 1; s.itwndlast = 1; s.itprcplast = 1;
  s.twindupd = M->p.t0; s.mwin = -1; s.momtw = 1.0f; s.msm = 1.0f;   // sfincs_lib.f90:290
  H->dS = H->upload(&s, 1);
  I.sc = H->dS;
  // tiles
  I.tn = H->tile_tn; I.tm = H->tile_tm;
  I.ntn = M->dense ? (M->nmax + I.tn - 1) / I.tn : 0;
  I.sk = H->strip_rows;
  I.nstrip = M->dense ? (M->nmax + STRIP_COLS - 1) / STRIP_COLS : 0;
  I.mtn = H->mtile_tn; I.mtm = H->mtile_tm;
  I.xchunk = H->xchunk; I.xchunk_c = H->xchunk_c;
  I.mntn = M->dense ? (M->nmax + I.mtn - 1) / I.mtn : 0;
  I.tn_sh = ilog2(I.tn); I.mtn_sh = ilog2(I.mtn);
  I.mrow_sh = ilog2(I.mtm * (I.sk < 0 ? -I.sk : 1));
  I.nmt = (int32_t)mom_tiles(I);
  H->nmt_cap = I.nmt;
  if (M->dense) {
    for (int k = 0; k < N_STRIP_ROWS; ++k) H->nmt_cap = std::max(H->nmt_cap, strip_tiles(I, STRIP_ROWS[k]));
    for (int k = 0; k < N_CELL_SHAPES; ++k)
      H->nmt_cap = std::max(H->nmt_cap, 2 * cell_tiles(I, CELL_SHAPES[k][0], CELL_SHAPES[k][1]));
  }
  {
    std::vector<float> init(H->nmt_cap, I.c.dtmax);
    I.blkmin = H->upload(init.data(), init.size());
  }
  I.blkcnt = H->alloc<uint32_t>(NCNT * H->nmt_cap);
  I.mq = H->alloc<uint32_t>(H->nmt_cap);
  I.md = H->alloc<uint8_t>(H->nmt_cap);
  H->nct_cap = M->dense ? 0 : (size_t)((I.np + BLOCK - 1) / BLOCK);
  if (M->dense)
    for (int k = 0; k < N_CELL_SHAPES; ++k)
      H->nct_cap = std::max(H->nct_cap, cell_tiles(I, CELL_SHAPES[k][0], CELL_SHAPES[k][1]));
  I.cntstore = H->alloc<uint32_t>(NSTORE * std::max<size_t>(1, H->nct_cap));
  H->dvol = H->alloc<double>(1);
  {   // volume balance: faces of interior cells whose other side is not interior
    const std::vector<int32_t>* faces[4] = {&M->z_index_uv_md, &M->z_index_uv_mu, &M->z_index_uv_nd, &M->z_index_uv_nu};
    const int dummy = npuv + ncuv + 1;
    std::vector<int32_t> slot;
    std::vector<double> coef;
    double rain_area = 0.0;
    for (int c = 0; c < np; ++c) {
      if (M->kcs[c] != 1) continue;
      const int ir = M->z_flags_iref[c] - 1;
      const double area = (double)M->cell_area[ir];
      rain_area += area;
      for (int k = 0; k < 4; ++k) {
        const int f = (*faces[k])[c];
        if (f <= 0 || f == dummy || M->kcuv[f - 1] == 1) continue;
        // continuity counts md and nd faces in, mu and nu faces out (sfincs_continuity.f90:130-148, 337-352)
        double w;
        if (M->subgrid) w = k < 2 ? (double)(M->use_quadtree ? M->dyrm[ir] : I.c.dy) : (double)(M->use_quadtree ? M->dxrm[ir] : I.c.dx);
        else w = area * (double)(k < 2 ? M->dxrinv[ir] : M->dyrinv[ir]);
        slot.push_back(H->ref2slot[f - 1]);
        coef.push_back((k % 2 == 0) ? w : -w);
      }
    }
    I.nbf = (int32_t)slot.size();
    I.bf_slot = H->upload(slot.data(), slot.size());
    I.bf_coef = H->upload(coef.data(), coef.size());
    I.rain_area = rain_area;
    I.bal = H->alloc<double>(NBAL);
  }
  H->dI = H->upload(&I, 1);
  for (void* p : H->allocs) if (!p) { set_error("hipMalloc failed"); return SFX_ERR_MEMORY; }
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipStreamCreateWithFlags(&H->stream, hipStreamNonBlocking));
  *out = H;
  return rc;
}

extern "C" int sfx_batch_destroy(sfx_batch* b);

extern "C" int sfx_destroy(sfx_handle* H) {
  if (!H) return SFX_OK;
  hipSetDevice(H->device);
  if (H->solo) { sfx_batch_destroy(H->solo); H->solo = nullptr; }
  hipDeviceSynchronize();
  for (void* p : H->allocs) hipFree(p);
  if (H->stream) hipStreamDestroy(H->stream);
  delete H;
  return SFX_OK;
}

// -------------------------------------------------------------------------- state exchange
namespace {

int read_scalars(sfx_handle* H, Scalars& s) {
  HIPCHK(hipStreamSynchronize(H->stream));
  HIPCHK(hipMemcpy(&s, H->dS, sizeof(Scalars), hipMemcpyDeviceToHost));
  return SFX_OK;
}
// Host write of the step scalars; it also drops a next step computed ahead on the old ones.
int write_scalars(sfx_handle* H, const Scalars& s) {
  HIPCHK(hipStreamSynchronize(H->stream));
  Scalars w = s;
  w.ahead = 0;
  HIPCHK(hipMemcpy(H->dS, &w, sizeof(Scalars), hipMemcpyHostToDevice));
  return SFX_OK;
}
// Completes a pending momentum sweep (fold of the tile minima, buffer flip) so that
// `cur` holds the latest q/uv and min_dt the sweep's minimum.
int normalise(sfx_handle* H, Scalars& s) {
  int rc = read_scalars(H, s);
  if (rc || !s.mom_done) return rc;
  hipLaunchKernelGGL(k_fold, dim3(1), dim3(FBLOCK), 0, H->stream, H->dI, (const int32_t*)nullptr);
  HIPCHK(hipGetLastError());
  return read_scalars(H, s);
}

// 4-byte per-slot words (f32 or i32) of the npuv uv-points, in reference order.
int get_uv_words(sfx_handle* H, const void* src, void* dst, size_t nbytes) {
  const size_t n = H->sz.npuv;
  if (nbytes != n * 4) { set_error("size mismatch"); return SFX_ERR_ARG; }
  std::vector<uint32_t> buf((size_t)H->nslot + 1);
  HIPCHK(hipMemcpy(buf.data(), src, buf.size() * 4, hipMemcpyDeviceToHost));
  uint32_t* d = (uint32_t*)dst;
  for (size_t k = 0; k < n; ++k) d[k] = H->ref2slot[k] >= 0 ? buf[H->ref2slot[k]] : 0u;
  return SFX_OK;
}

template <typename R>
int get_uv_field(sfx_handle* H, void* src, void* dst, size_t nbytes, bool is_q) {
  (void)is_q;
  const size_t n = H->ref2slot.size();
  if (nbytes != n * sizeof(R) && nbytes != n * sizeof(float)) { set_error("size mismatch"); return SFX_ERR_ARG; }
  std::vector<R> buf(H->nslot + 1);
  HIPCHK(hipMemcpy(buf.data(), src, buf.size() * sizeof(R), hipMemcpyDeviceToHost));
  if (nbytes == n * sizeof(R)) {
    R* d = (R*)dst;
    for (size_t k = 0; k < n; ++k) d[k] = H->ref2slot[k] >= 0 ? buf[H->ref2slot[k]] : R(0);
  } else {
    float* d = (float*)dst;
    for (size_t k = 0; k < n; ++k) d[k] = H->ref2slot[k] >= 0 ? (float)buf[H->ref2slot[k]] : 0.f;
  }
  return SFX_OK;
}

template <typename R>
int set_uv_field(sfx_handle* H, void* dst, const void* src, size_t nbytes) {
  const size_t n = H->ref2slot.size();
  if (nbytes != n * sizeof(float) && nbytes != n * sizeof(R)) { set_error("size mismatch"); return SFX_ERR_ARG; }
  std::vector<R> buf(H->nslot + 1, R(0));
  HIPCHK(hipMemcpy(buf.data(), dst, buf.size() * sizeof(R), hipMemcpyDeviceToHost));
  for (size_t k = 0; k < n; ++k) {
    if (H->ref2slot[k] < 0) continue;
    buf[H->ref2slot[k]] = nbytes == n * sizeof(R) ? ((const R*)src)[k] : (R)((const float*)src)[k];
  }
  buf[H->nslot] = R(0);
  HIPCHK(hipMemcpy(dst, buf.data(), buf.size() * sizeof(R), hipMemcpyHostToDevice));
  return SFX_OK;
}

template <typename R>
int get_cell_r(sfx_handle* H, const void* src, void* dst, size_t nbytes, size_t n) {
  if (nbytes == n * sizeof(R)) { HIPCHK(hipMemcpy(dst, src, nbytes, hipMemcpyDeviceToHost)); return SFX_OK; }
  if (nbytes != n * sizeof(float)) { set_error("size mismatch"); return SFX_ERR_ARG; }
  std::vector<R> b(n);
  HIPCHK(hipMemcpy(b.data(), src, n * sizeof(R), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; ++i) ((float*)dst)[i] = (float)b[i];
  return SFX_OK;
}

template <typename R>
int set_cell_r(sfx_handle* H, void* dst, const void* src, size_t nbytes, size_t n) {
  (void)H;
  if (nbytes == n * sizeof(R)) { HIPCHK(hipMemcpy(dst, src, nbytes, hipMemcpyHostToDevice)); return SFX_OK; }
  if (nbytes != n * sizeof(float)) { set_error("size mismatch"); return SFX_ERR_ARG; }
  std::vector<R> b(n);
  for (size_t i = 0; i < n; ++i) b[i] = (R)((const float*)src)[i];
  HIPCHK(hipMemcpy(dst, b.data(), n * sizeof(R), hipMemcpyHostToDevice));
  return SFX_OK;
}

}  // namespace

extern "C" int sfx_get_state(sfx_handle* H, const char* name, void* dst, size_t nbytes) {
  if (!H || !name || !dst) { set_error("sfx_get_state: null"); return SFX_ERR_ARG; }
  HIPCHK(hipSetDevice(H->device));
  Scalars s;
  int rc = normalise(H, s);
  if (rc) return rc;
  const std::string k = name;
  const size_t np = H->hI.np;
  DevInst& I = H->hI;
  if (k == "zs") { if (nbytes != np * 8) { set_error("size mismatch"); return SFX_ERR_ARG; } HIPCHK(hipMemcpy(dst, I.zs, nbytes, hipMemcpyDeviceToHost)); return SFX_OK; }
  if (k == "zb") {
    if (!I.zb || nbytes != np * 4) { set_error("sfx_get_state: zb needs a regular (not subgrid) mesh and np f32 values"); return SFX_ERR_ARG; }
    HIPCHK(hipMemcpy(dst, I.zb, nbytes, hipMemcpyDeviceToHost));
    return SFX_OK;
  }
  if (k == "z_volume") {
    if (!I.vol || nbytes != np * 8) { set_error("no z_volume / size mismatch"); return SFX_ERR_ARG; }
    HIPCHK(hipMemcpy(dst, I.vol, nbytes, hipMemcpyDeviceToHost)); return SFX_OK;
  }
  if (k == "q" || k == "uv") {
    void* src = k == "q" ? I.q[s.cur] : I.uv[s.cur];
    return H->fp64 ? get_uv_field<double>(H, src, dst, nbytes, k == "q") : get_uv_field<float>(H, src, dst, nbytes, k == "q");
  }
  if (k == "kfuv") {
    const size_t n = H->sz.npuv;
    if (nbytes != n) { set_error("size mismatch"); return SFX_ERR_ARG; }
    std::vector<int8_t> b(H->nslot + 1);
    HIPCHK(hipMemcpy(b.data(), I.kfuv, b.size(), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) ((int8_t*)dst)[i] = (int8_t)(b[H->ref2slot[i]] & 1);
    return SFX_OK;
  }
  if (k == "zs0" || k == "zsderv" || k == "zsmax") {
    void* src = k == "zs0" ? I.zs0 : (k == "zsderv" ? I.zsderv : I.zsmax);
    return H->fp64 ? get_cell_r<double>(H, src, dst, nbytes, np) : get_cell_r<float>(H, src, dst, nbytes, np);
  }
  if (k == "tsa_sum" || k == "tsa_wet" || k == "tsa_lim") {   // timestep analysis, per uv-point
    if (!I.tsa_req) { set_error("sfx_get_state: " + k + ": timestep analysis is off"); return SFX_ERR_ARG; }
    const void* src = k == "tsa_sum" ? (const void*)I.tsa_sum : (k == "tsa_wet" ? (const void*)I.tsa_wet : (const void*)I.tsa_lim);
    return get_uv_words(H, src, dst, nbytes);
  }
  if (k == "storage_volume") {
    if (!I.storage) { set_error("sfx_get_state: no storage volume"); return SFX_ERR_ARG; }
    return H->fp64 ? get_cell_r<double>(H, I.storage, dst, nbytes, np) : get_cell_r<float>(H, I.storage, dst, nbytes, np);
  }
  if (k == "vmax" || k == "qmax" || k == "twet") {
    void* src = k == "vmax" ? I.vmax : (k == "qmax" ? I.qmax : I.twet);
    if (!src) { set_error("sfx_get_state: " + k + " is not stored (its store switch is off)"); return SFX_ERR_ARG; }
    return H->fp64 ? get_cell_r<double>(H, src, dst, nbytes, np) : get_cell_r<float>(H, src, dst, nbytes, np);
  }
  if (k == "uvmean") {
    const size_t n = H->sz.nkcuv2;
    return H->fp64 ? get_cell_r<double>(H, I.uvmean, dst, nbytes, n) : get_cell_r<float>(H, I.uvmean, dst, nbytes, n);
  }
  if (k == "cumprcp" || k == "cuminf") {   // curve-number cumulative rain and infiltration, f32 per cell
    const float* src = k == "cumprcp" ? I.cumprcp : I.cuminf;
    if (!src || nbytes != (size_t)np * 4) { set_error("sfx_get_state: " + k + " needs curve-number infiltration and np f32"); return SFX_ERR_ARG; }
    HIPCHK(hipMemcpy(dst, src, nbytes, hipMemcpyDeviceToHost));
    return SFX_OK;
  }
  if (k == "qtsrc") {   // the step's discharges: point sources, then each drainage intake and outfall
    const size_t n = (size_t)I.nsrc + 2 * (size_t)I.ndrn;
    if (n == 0 || nbytes != n * 4) { set_error("sfx_get_state: qtsrc needs (nsrc + 2 ndrn) f32 values"); return SFX_ERR_ARG; }
    HIPCHK(hipMemcpy(dst, I.qtsrc, nbytes, hipMemcpyDeviceToHost));
    return SFX_OK;
  }
  set_error("unknown state field " + k);
  return SFX_ERR_ARG;
}

namespace { int reset_tile_flags(sfx_handle* H); }

// A host write of cell state: no cell keeps its steady bit and no tile is quiescent.
int clear_steady(sfx_handle* H) {
  std::vector<uint8_t> b(H->hI.np);
  HIPCHK(hipMemcpy(b.data(), H->hI.kcs, b.size(), hipMemcpyDeviceToHost));
  for (auto& v : b) v = (uint8_t)(v & ~KCS_STEADY);
  HIPCHK(hipMemcpy(H->hI.kcs, b.data(), b.size(), hipMemcpyHostToDevice));
  return reset_tile_flags(H);
}

extern "C" int sfx_set_state(sfx_handle* H, const char* name, const void* src, size_t nbytes) {
  if (!H || !name || !src) { set_error("sfx_set_state: null"); return SFX_ERR_ARG; }
  HIPCHK(hipSetDevice(H->device));
  Scalars s;
  int rc = normalise(H, s);
  if (rc) return rc;
  const std::string k = name;
  const size_t np = H->hI.np;
  DevInst& I = H->hI;
  if (k == "zs") {
    if (nbytes != np * 8) { set_error("size mismatch"); return SFX_ERR_ARG; }
    HIPCHK(hipMemcpy(I.zs, src, nbytes, hipMemcpyHostToDevice));
    return clear_steady(H);
  }
  if (k == "z_volume") {
    if (!I.vol || nbytes != np * 8) { set_error("no z_volume / size mismatch"); return SFX_ERR_ARG; }
    HIPCHK(hipMemcpy(I.vol, src, nbytes, hipMemcpyHostToDevice));
    return clear_steady(H);
  }
  if (k == "zb") {   // bed level of a regular mesh, e.g. written by a coupler through the BMI
    if (!I.zb || nbytes != np * 4) { set_error("sfx_set_state: zb needs a regular (not subgrid) mesh and np f32 values"); return SFX_ERR_ARG; }
    HIPCHK(hipMemcpy((void*)I.zb, src, nbytes, hipMemcpyHostToDevice));
    return clear_steady(H);
  }
  if (k == "q" || k == "uv") {
    void* dst = k == "q" ? I.q[s.cur] : I.uv[s.cur];
    rc = H->fp64 ? set_uv_field<double>(H, dst, src, nbytes) : set_uv_field<float>(H, dst, src, nbytes);
    if (rc) return rc;
    // only the current buffer changed: the momentum kernels' zero flags no longer hold
    std::vector<int8_t> b(H->nslot + 1);
    HIPCHK(hipMemcpy(b.data(), I.kfuv, b.size(), hipMemcpyDeviceToHost));
    for (auto& v : b) v = (int8_t)(v & 1);
    HIPCHK(hipMemcpy(I.kfuv, b.data(), b.size(), hipMemcpyHostToDevice));
    return reset_tile_flags(H);
  }
  if (k == "kfuv") {
    const size_t n = H->sz.npuv;
    if (nbytes != n) { set_error("size mismatch"); return SFX_ERR_ARG; }
    std::vector<int8_t> b(H->nslot + 1, 0);
    HIPCHK(hipMemcpy(b.data(), I.kfuv, b.size(), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) b[H->ref2slot[i]] = ((const int8_t*)src)[i];
    HIPCHK(hipMemcpy(I.kfuv, b.data(), b.size(), hipMemcpyHostToDevice));
    return reset_tile_flags(H);
  }
  if (k == "zs0" || k == "zsderv" || k == "zsmax") {
    void* dst = k == "zs0" ? I.zs0 : (k == "zsderv" ? I.zsderv : I.zsmax);
    rc = H->fp64 ? set_cell_r<double>(H, dst, src, nbytes, np) : set_cell_r<float>(H, dst, src, nbytes, np);
    if (rc) return rc;
    return clear_steady(H);
  }
  if (k == "storage_volume") {
    if (!I.storage) { set_error("sfx_set_state: no storage volume"); return SFX_ERR_ARG; }
    rc = H->fp64 ? set_cell_r<double>(H, I.storage, src, nbytes, np) : set_cell_r<float>(H, I.storage, src, nbytes, np);
    if (rc) return rc;
    return clear_steady(H);
  }
  if (k == "qext") {   // external sources (m/s per cell), e.g. from a coupled model through the BMI
    if (!I.qext || nbytes != np * 4) { set_error("sfx_set_state: qext needs use_qext and np f32 values"); return SFX_ERR_ARG; }
    HIPCHK(hipMemcpy(I.qext, src, nbytes, hipMemcpyHostToDevice));
    return SFX_OK;
  }
  if (k == "vmax" || k == "qmax" || k == "twet") {   // e.g. the reset after each maximum output
    void* dst = k == "vmax" ? I.vmax : (k == "qmax" ? I.qmax : I.twet);
    if (!dst) { set_error("sfx_set_state: " + k + " is not stored (its store switch is off)"); return SFX_ERR_ARG; }
    return H->fp64 ? set_cell_r<double>(H, dst, src, nbytes, np) : set_cell_r<float>(H, dst, src, nbytes, np);
  }
  set_error("unknown state field " + k);
  return SFX_ERR_ARG;
}

// write_rst_file (sfincs_output.f90:657-729), type 1: the state is copied to the host
// and written there.
extern "C" int sfx_write_restart(sfx_handle* H, const char* path) {
  if (!H || !path) { set_error("sfx_write_restart: null"); return SFX_ERR_ARG; }
  const int32_t np = H->hI.np, nq = H->sz.npuv + H->sz.ncuv + 1, nu = H->sz.nkcuv2;
  std::vector<double> zs((size_t)np);
  std::vector<float> q((size_t)nq), uvm((size_t)std::max(nu, 1));
  int rc = sfx_get_state(H, "zs", zs.data(), zs.size() * 8);
  if (!rc) rc = sfx_get_state(H, "q", q.data(), q.size() * 4);
  if (!rc && nu > 0) rc = sfx_get_state(H, "uvmean", uvm.data(), (size_t)nu * 4);
  if (rc) return rc;
  return sfx::write_restart_arrays(path, zs.data(), np, q.data(), nq, uvm.data(), nu);
}

extern "C" int sfx_get_time(sfx_handle* H, double* t, float* dt, int64_t* nt) {
  if (!H) { set_error("null"); return SFX_ERR_ARG; }
  HIPCHK(hipSetDevice(H->device));
  Scalars s;
  int rc = read_scalars(H, s);
  if (rc) return rc;
  if (s.error_pending) { s.error = 1; s.t = (double)H->hI.c.t1 + 1.0; }
  if (t) *t = s.t;
  if (dt) *dt = s.dt;
  if (nt) *nt = s.nt;
  return SFX_OK;
}

// -------------------------------------------------------------------------- batches / launches
namespace {


// XCD-aware ordering: the hardware hands workgroup b to XCD b % 8. With chunk == 0, item
// b is the (b%8)-th eighth of the natural tile order: each XCD sweeps one contiguous band
// and the neighbours a tile reads stay in its L2. With chunk > 0 the natural order is cut
// into runs of `chunk` items dealt to the XCDs in turn (XCD x gets runs x, x+8, ...): a
// run still shares its neighbours in one L2, and spatially clustered work (wet regions
// in the momentum sweep) is spread over all XCDs instead of loading some of them more.
void xcd_order(std::vector<Item>& v, int chunk) {
  const size_t n = v.size();
  const size_t group = chunk > 0 ? 8 * (size_t)chunk : 8;
  const size_t padded = (n + group - 1) / group * group;
  std::vector<Item> nat(v);
  nat.resize(padded, Item{-1, 0});
  v.assign(padded, Item{-1, 0});
  if (chunk <= 0) {
    const size_t per = padded / 8;
    for (size_t b = 0; b < padded; ++b) v[b] = nat[(b % 8) * per + b / 8];
  } else {
    const size_t run = (size_t)chunk;
    for (size_t b = 0; b < padded; ++b) {
      const size_t x = b % 8, k = b / 8;
      v[b] = nat[((k / run) * 8 + x) * run + k % run];
    }
  }
}

// Storage order of a launch's items: item b (run by block b, on XCD b mod 8) at
// (b mod 8) n/8 + b/8, so the items of one XCD are contiguous and each XCD's L2 fetches
// only its own lines. n is a multiple of 8 (xcd_order pads).
void xcd_major(std::vector<Item>& v) {
  const size_t n = v.size(), per = n / 8;
  std::vector<Item> w(n);
  for (size_t b = 0; b < n; ++b) w[(b % 8) * per + b / 8] = v[b];
  v.swap(w);
}

int rebuild(sfx_batch* B) {
  HIPCHK(hipSetDevice(B->device));
  const size_t nm = B->members.size();
  std::vector<DevInst> insts(nm);
  std::vector<Item> mom[NLIST_M], cu, cd, ci;
  std::vector<int32_t> ids;
  int fp64 = -1;
  if (B->live.size() != nm) B->live.assign(nm, 1);
  for (size_t i = 0; i < nm; ++i) {
    sfx_handle* H = B->members[i];
    insts[i] = H->hI;
    if (fp64 < 0) fp64 = H->fp64;
    if (fp64 != H->fp64) { set_error("a batch cannot mix FP32 and FP64 members"); return SFX_ERR_ARG; }
    if (!B->live[i]) continue;
    ids.push_back((int32_t)i);
    const DevInst& I = H->hI;
    std::vector<Item>& mv = mom[mom_list(I)];
    const size_t tm = mom_tiles(I), tc = cont_tiles(I);
    for (size_t t = 0; t < tm; ++t) mv.push_back(Item{(int)i, (int)t});
    for (size_t t = 0; t < tc; ++t) (I.dense ? cd : ci).push_back(Item{(int)i, (int)t});
    if (!I.dense) {
      const int tq = (I.ncuv + BLOCK - 1) / BLOCK;
      for (int t = 0; t < tq; ++t) cu.push_back(Item{(int)i, t});
    }
  }
  B->fp64 = fp64 < 0 ? 0 : fp64;
  std::vector<Item> all;
  const int xchunk = nm ? B->members[0]->hI.xchunk : 0;
  const int xchunk_c = nm ? B->members[0]->hI.xchunk_c : 0;
  // A list filled by one member (the lists are appended member by member) with a run
  // length of 0 or a power of two is launched direct: its blocks compute their tiles.
  struct Src { int inst = -1; int32_t n = 0, run_sh = 0; };
  auto source = [](const std::vector<Item>& v, int chunk) {
    Src r;
    if (v.empty() || v.front().inst != v.back().inst) return r;
    if (chunk > 0 && (chunk & (chunk - 1))) return r;
    r.inst = v.front().inst; r.n = (int32_t)v.size(); r.run_sh = chunk > 0 ? ilog2(chunk) : -1;
    return r;
  };
  Src src_m[NLIST_M], src_cu = source(cu, 0), src_cd = source(cd, xchunk_c), src_ci = source(ci, xchunk_c);
  for (int k = 0; k < NLIST_M; ++k) {
    src_m[k] = source(mom[k], xchunk);
    if (mom[k].size()) { xcd_order(mom[k], xchunk); xcd_major(mom[k]); }
    B->off_m[k] = all.size(); all.insert(all.end(), mom[k].begin(), mom[k].end()); B->n_m[k] = mom[k].size();
  }
  if (cu.size()) { xcd_order(cu, 0); xcd_major(cu); }
  if (cd.size()) { xcd_order(cd, xchunk_c); xcd_major(cd); }
  if (ci.size()) { xcd_order(ci, xchunk_c); xcd_major(ci); }
  B->off_cuv = all.size(); all.insert(all.end(), cu.begin(), cu.end()); B->n_cuv = cu.size();
  B->off_cd = all.size(); all.insert(all.end(), cd.begin(), cd.end()); B->n_cd = cd.size();
  B->off_ci = all.size(); all.insert(all.end(), ci.begin(), ci.end()); B->n_ci = ci.size();
  if (B->dInsts_cap < nm) {
    if (B->dInsts) hipFree(B->dInsts);
    B->dInsts_cap = std::max<size_t>(nm, 8);
    HIPCHK(hipMalloc(&B->dInsts, B->dInsts_cap * sizeof(DevInst)));
  }
  if (B->items_cap < all.size()) {
    if (B->dItems) hipFree(B->dItems);
    B->items_cap = std::max<size_t>(all.size(), 64);
    HIPCHK(hipMalloc(&B->dItems, B->items_cap * sizeof(Item)));
  }
  if (B->ids_cap < nm) {
    if (B->dIds) hipFree(B->dIds);
    B->ids_cap = std::max<size_t>(nm, 8);
    HIPCHK(hipMalloc(&B->dIds, B->ids_cap * sizeof(int32_t)));
  }
  if (nm) HIPCHK(hipMemcpy(B->dInsts, insts.data(), nm * sizeof(DevInst), hipMemcpyHostToDevice));
  if (all.size()) HIPCHK(hipMemcpy(B->dItems, all.data(), all.size() * sizeof(Item), hipMemcpyHostToDevice));
  if (ids.size()) HIPCHK(hipMemcpy(B->dIds, ids.data(), ids.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  B->n_live = (int)ids.size();
  auto wl = [&](size_t off, const Src& r) {
    WorkList w{};
    w.items = B->dItems + off;
    if (r.inst >= 0) {
      w.direct = 1; w.inst = r.inst; w.sc = B->members[r.inst]->hI.sc; w.n = r.n; w.run_sh = r.run_sh;
    }
    return w;
  };
  for (int k = 0; k < NLIST_M; ++k) B->wl_m[k] = wl(B->off_m[k], src_m[k]);
  B->wl_cuv = wl(B->off_cuv, src_cu); B->wl_cd = wl(B->off_cd, src_cd); B->wl_ci = wl(B->off_ci, src_ci);
  if (B->exec) { hipGraphExecDestroy(B->exec); B->exec = nullptr; }
  if (B->graph) { hipGraphDestroy(B->graph); B->graph = nullptr; }
  B->graph_steps = 0;
  B->dirty = false;
  return SFX_OK;
}

// Descriptor argument of a launch: the scenario's own descriptor for a direct list.
inline const DevInst* base_of(const sfx_batch* B, const WorkList& w) {
  return w.direct ? B->dInsts + w.inst : B->dInsts;
}

// Launch phases of one step. what: bit 0 forcing, bit 1 momentum (+cuv), bit 2 continuity.
// timestep_analysis_update (sfincs_timestep_analysis.f90:26-74, called after compute_fluxes
// at sfincs_lib.f90:588) in two launches after a momentum sweep. First, per scenario, the
// sweep's minimum time step from its tile minima (read only: the next fold still uses
// them); then, per uv slot, the running sum of the wet points' required steps, the sweeps
// they were wet and those in which they limited the step (required <= min_dt + 1e-6).
__global__ void __launch_bounds__(FBLOCK) k_tsa_min(const DevInst* __restrict__ insts,
                                                   const int32_t* __restrict__ ids) {
  const DevInst& I = insts[ids[blockIdx.x]];
  if (!(I.c.flags & sfx::F_TSA) || !G(I.sc)->mom_done) return;
  float v = I.c.dtmax;
  for (int i = threadIdx.x; i < I.nmt; i += FBLOCK) v = fminf(v, G(I.blkmin)[i]);
  v = block_min<FBLOCK>(v);
  if (threadIdx.x == 0) G(I.tsa_min)[0] = v;
}

__global__ void __launch_bounds__(BLOCK) k_tsa_update(const DevInst* __restrict__ insts,
                                                     const int32_t* __restrict__ ids) {
  const DevInst& I = insts[ids[blockIdx.y]];
  if (!(I.c.flags & sfx::F_TSA) || !G(I.sc)->mom_done) return;
  const float lim = G(I.tsa_min)[0] + 1.0e-6f;
  for (int s = blockIdx.x * BLOCK + threadIdx.x; s < I.nslot; s += gridDim.x * BLOCK) {
    const int kc = G(I.kcuv)[s];
    if ((kc == 1 || kc == 6) && (G(I.kfuv)[s] & 1)) {
      const float r = G(I.tsa_req)[s];
      G(I.tsa_sum)[s] = G(I.tsa_sum)[s] + r;
      G(I.tsa_wet)[s] = G(I.tsa_wet)[s] + 1;
      if (r <= lim) G(I.tsa_lim)[s] = G(I.tsa_lim)[s] + 1;
    }
  }
}

template <typename R>
void launch_phases(sfx_batch* B, int what, int forcing_mode) {
  hipStream_t st = B->stream;
  const int nm = B->n_live;
  if ((what & 1) && nm) hipLaunchKernelGGL(k_forcing<R>, dim3(nm), dim3(FBLOCK), 0, st, B->dInsts, B->dIds, forcing_mode);
  if (what & 2) {
    if (B->n_m[0]) hipLaunchKernelGGL((k_mom_cell<R, true>), dim3(B->n_m[0]), dim3(BLOCK), 0, st, base_of(B, B->wl_m[0]), B->wl_m[0]);
    if (B->n_m[1]) hipLaunchKernelGGL((k_mom_cell<R, false>), dim3(B->n_m[1]), dim3(BLOCK), 0, st, base_of(B, B->wl_m[1]), B->wl_m[1]);
    if (B->n_m[2]) hipLaunchKernelGGL((k_mom_strip<R, true>), dim3(B->n_m[2]), dim3(BLOCK), 0, st, base_of(B, B->wl_m[2]), B->wl_m[2]);
    if (B->n_m[3]) hipLaunchKernelGGL((k_mom_strip<R, false>), dim3(B->n_m[3]), dim3(BLOCK), 0, st, base_of(B, B->wl_m[3]), B->wl_m[3]);
    if (B->n_m[4]) hipLaunchKernelGGL((k_mom_idx<R, true>), dim3(B->n_m[4]), dim3(BLOCK), 0, st, base_of(B, B->wl_m[4]), B->wl_m[4]);
    if (B->n_m[5]) hipLaunchKernelGGL((k_mom_idx<R, false>), dim3(B->n_m[5]), dim3(BLOCK), 0, st, base_of(B, B->wl_m[5]), B->wl_m[5]);
#define SFX_MOM_POINT(k, SG, P) \
    if (B->n_m[k]) hipLaunchKernelGGL((k_mom_point<R, SG, P>), dim3(B->n_m[k]), dim3(BLOCK), 0, st, base_of(B, B->wl_m[k]), B->wl_m[k]);
    SFX_MOM_POINT(6, true, 1) SFX_MOM_POINT(7, false, 1)
    SFX_MOM_POINT(8, true, 2) SFX_MOM_POINT(9, false, 2)
    SFX_MOM_POINT(10, true, 4) SFX_MOM_POINT(11, false, 4)
#undef SFX_MOM_POINT
    if (B->n_cuv) hipLaunchKernelGGL(k_cuv<R>, dim3(B->n_cuv), dim3(BLOCK), 0, st, base_of(B, B->wl_cuv), B->wl_cuv);
    int tsa_slots = 0;   // timestep analysis of the members that have it on (not while tuning)
    for (sfx_handle* H : B->members)
      if (H->hI.tsa_req && !H->tsa_paused) tsa_slots = std::max(tsa_slots, (int)H->hI.nslot);
    if (tsa_slots && nm) {
      hipLaunchKernelGGL(k_tsa_min, dim3(nm), dim3(FBLOCK), 0, st, B->dInsts, B->dIds);
      const int nb = std::min(1024, (tsa_slots + BLOCK - 1) / BLOCK);
      hipLaunchKernelGGL(k_tsa_update, dim3(nb, nm), dim3(BLOCK), 0, st, B->dInsts, B->dIds);
    }
  }
  // A whole step: the next step's fold, scalars and series (forcing_ahead_block) run as
  // the leading blocks of this step's first continuity launch, beside its cells.
  // The volume balance of every member rides in the same leading blocks.
  const int nahead = (what == 7 && forcing_mode == 0) ? nm : 0;
  const int nbal = nm;
  const int lead = (nahead + nbal + 7) / 8 * 8;   // keeps block b of the cells on XCD b mod 8
  if (what & 4) {
    bool rx = false;   // per-cell rain (infiltration, gridded rain) in some member
    for (sfx_handle* H : B->members) rx = rx || (H->hI.c.flags & (sfx::F_INFIL | sfx::F_MPRCP | sfx::F_CNA)) != 0;
    auto kc = rx ? k_continuity<R, true> : k_continuity<R, false>;
    if (B->n_cd)
      hipLaunchKernelGGL(kc, dim3(B->n_cd + lead), dim3(BLOCK), 0, st, base_of(B, B->wl_cd), B->wl_cd, 0,
                         nahead, nbal, B->dInsts, B->dIds);
    if (B->n_ci) {
      const int la = B->n_cd ? 0 : nahead, lb = B->n_cd ? 0 : nbal, ll = B->n_cd ? 0 : lead;
      hipLaunchKernelGGL(kc, dim3(B->n_ci + ll), dim3(BLOCK), 0, st, base_of(B, B->wl_ci), B->wl_ci, 1,
                         la, lb, B->dInsts, B->dIds);
    }
  }
}

void launch(sfx_batch* B, int what, int mode) {
  if (B->fp64) launch_phases<double>(B, what, mode);
  else launch_phases<float>(B, what, mode);
}

int ensure_graph(sfx_batch* B, int steps) {
  if (B->dirty) { int rc = rebuild(B); if (rc) return rc; }
  if (B->exec && B->graph_steps == steps) return SFX_OK;
  if (B->exec) { hipGraphExecDestroy(B->exec); B->exec = nullptr; }
  if (B->graph) { hipGraphDestroy(B->graph); B->graph = nullptr; }
  HIPCHK(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
  for (int s = 0; s < steps; ++s) launch(B, 7, 0);
  HIPCHK(hipStreamEndCapture(B->stream, &B->graph));
  HIPCHK(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
  B->graph_steps = steps;
  return SFX_OK;
}

sfx_batch* solo(sfx_handle* H) {
  if (!H->solo) {
    H->solo = new sfx_batch();
    H->solo->device = H->device;
    H->solo->stream = H->stream;
    H->solo->own_stream = false;
    H->solo->members.push_back(H);
    H->solo->tend.push_back(0.0);
    H->solo->dirty = true;
  }
  return H->solo;
}

}  // namespace

extern "C" int sfx_compute_fluxes(sfx_handle* H, float dt, float* min_dt) {
  if (!H) { set_error("null"); return SFX_ERR_ARG; }
  HIPCHK(hipSetDevice(H->device));
  Scalars s;
  int rc = normalise(H, s);
  if (rc) return rc;
  s.dt = dt; s.active = 1;
  if ((rc = write_scalars(H, s))) return rc;
  sfx_batch* B = solo(H);
  if (B->dirty && (rc = rebuild(B))) return rc;
  launch(B, 2, 0);
  HIPCHK(hipGetLastError());
  if ((rc = normalise(H, s))) return rc;
  if (min_dt) *min_dt = s.min_dt;
  return SFX_OK;
}

extern "C" int sfx_compute_water_levels(sfx_handle* H, double t, float dt) {
  if (!H) { set_error("null"); return SFX_ERR_ARG; }
  HIPCHK(hipSetDevice(H->device));
  Scalars s;
  int rc = read_scalars(H, s);
  if (rc) return rc;
  s.t = t; s.dt = dt; s.active = 1;
  if ((rc = write_scalars(H, s))) return rc;
  sfx_batch* B = solo(H);
  if (B->dirty && (rc = rebuild(B))) return rc;
  launch(B, 4, 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(H->stream));
  return SFX_OK;
}

extern "C" int sfx_update_forcing(sfx_handle* H, double t, float dt) {
  if (!H) { set_error("null"); return SFX_ERR_ARG; }
  HIPCHK(hipSetDevice(H->device));
  Scalars s;
  int rc = normalise(H, s);
  if (rc) return rc;
  s.t = t; s.dt = dt;
  if ((rc = write_scalars(H, s))) return rc;
  sfx_batch* B = solo(H);
  if (B->dirty && (rc = rebuild(B))) return rc;
  launch(B, 1, 1);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(H->stream));
  return SFX_OK;
}

namespace {
// The step loop of sfincs_update (sfincs_lib.f90:379-648): one step (single), or steps
// while t < tend, in graph launches of 32 steps whose steps past tend do nothing.
int update_loop(sfx_handle* H, int mode, double arg, sfx_step_stats* st) {
  HIPCHK(hipSetDevice(H->device));
  auto t_start = std::chrono::steady_clock::now();
  Scalars s;
  int rc = read_scalars(H, s);
  if (rc) return rc;
  if (s.error_pending) { s.error = 1; s.error_pending = 0; s.t = (double)H->hI.c.t1 + 1.0; }
  const int64_t nt0 = s.nt;
  bool single = false;
  if (mode == 1) s.tend = arg;                          // absolute end time
  else if (arg < -999.0) s.tend = (double)H->hI.c.t1;
  else if (arg <= 0.0) { s.tend = (double)H->hI.c.t1; single = true; }
  else s.tend = s.t + arg;
  if ((rc = write_scalars(H, s))) return rc;
  sfx_batch* B = solo(H);
  if (B->dirty && (rc = rebuild(B))) return rc;
  if (single) {
    launch(B, 7, 0);
    HIPCHK(hipGetLastError());
  } else {
    const int K = 32;
    if ((rc = ensure_graph(B, K))) return rc;
    for (;;) {
      HIPCHK(hipGraphLaunch(B->exec, B->stream));
      if ((rc = read_scalars(H, s))) return rc;
      if (!s.active || s.error || s.error_pending || !(s.t < s.tend)) break;
    }
  }
  if ((rc = normalise(H, s))) return rc;
  if (s.error_pending) { s.error = 1; s.error_pending = 0; s.t = (double)H->hI.c.t1 + 1.0; if ((rc = write_scalars(H, s))) return rc; }
  if (st) {
    st->t = s.t; st->dt = s.dt; st->min_dt = s.min_dt; st->nt = s.nt; st->error = s.error;
    st->cell_updates = (double)(s.nt - nt0) * (double)H->hI.np;
    st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  }
  return s.error ? SFX_UNSTABLE : SFX_OK;
}
}  // namespace

extern "C" int sfx_update(sfx_handle* H, double dtrange, sfx_step_stats* st) {
  if (!H) { set_error("null"); return SFX_ERR_ARG; }
  return update_loop(H, 0, dtrange, st);
}

extern "C" int sfx_update_to(sfx_handle* H, double t_end, sfx_step_stats* st) {
  if (!H) { set_error("null"); return SFX_ERR_ARG; }
  return update_loop(H, 1, t_end, st);
}

extern "C" int sfx_set_counting(sfx_handle* H, int32_t on) {
  if (!H) { set_error("null"); return SFX_ERR_ARG; }
  HIPCHK(hipSetDevice(H->device));
  H->hI.count = on ? 1 : 0;
  HIPCHK(hipStreamSynchronize(H->stream));
  HIPCHK(hipMemcpy(H->dI, &H->hI, sizeof(DevInst), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(H->hI.blkcnt, 0, NCNT * H->nmt_cap * sizeof(uint32_t)));
  HIPCHK(hipMemset(H->hI.cntstore, 0, NSTORE * std::max<size_t>(1, H->nct_cap) * sizeof(uint32_t)));
  if (H->solo) H->solo->dirty = true;
  return SFX_OK;
}

extern "C" int sfx_get_counters(sfx_handle* H, int64_t counts[6]) {
  if (!H || !counts) { set_error("null"); return SFX_ERR_ARG; }
  HIPCHK(hipSetDevice(H->device));
  HIPCHK(hipStreamSynchronize(H->stream));
  std::vector<uint32_t> c(NCNT * H->nmt_cap);
  HIPCHK(hipMemcpy(c.data(), H->hI.blkcnt, c.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
  for (int k = 0; k < NCNT; ++k) counts[k] = 0;
  for (size_t i = 0; i < c.size(); ++i) counts[i % NCNT] += c[i];
  return SFX_OK;
}

extern "C" int sfx_get_store_counters(sfx_handle* H, int64_t counts[6]) {
  if (!H || !counts) { set_error("null"); return SFX_ERR_ARG; }
  HIPCHK(hipSetDevice(H->device));
  HIPCHK(hipStreamSynchronize(H->stream));
  std::vector<uint32_t> c(NSTORE * std::max<size_t>(1, H->nct_cap));
  HIPCHK(hipMemcpy(c.data(), H->hI.cntstore, c.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
  for (int k = 0; k < NSTORE; ++k) counts[k] = 0;
  for (size_t i = 0; i < c.size(); ++i) counts[i % NSTORE] += c[i];
  return SFX_OK;
}

extern "C" int sfx_total_volume(sfx_handle* H, double* vol) {
  if (!H || !vol) { set_error("null"); return SFX_ERR_ARG; }
  HIPCHK(hipSetDevice(H->device));
  hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(1), 0, H->stream, H->dvol);
  const int nb = std::min(1024, (H->hI.np + BLOCK - 1) / BLOCK);
  hipLaunchKernelGGL(k_volume, dim3(nb), dim3(BLOCK), 0, H->stream, H->dI, H->dvol);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(H->stream));
  HIPCHK(hipMemcpy(vol, H->dvol, sizeof(double), hipMemcpyDeviceToHost));
  return SFX_OK;
}

extern "C" int sfx_get_balance(sfx_handle* H, double out[6]) {
  if (!H || !out) { set_error("null"); return SFX_ERR_ARG; }
  HIPCHK(hipSetDevice(H->device));
  hipLaunchKernelGGL(k_zero_bal, dim3(1), dim3(64), 0, H->stream, H->hI.bal, 4, 6);
  const int nb = std::min(1024, (H->hI.np + BLOCK - 1) / BLOCK);
  hipLaunchKernelGGL(k_balance_volume, dim3(nb), dim3(BLOCK), 0, H->stream, H->dI);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(H->stream));
  double b[NBAL];
  HIPCHK(hipMemcpy(b, H->hI.bal, sizeof(b), hipMemcpyDeviceToHost));
  double infiltrated = 0.0;
  for (int k = 0; k < NINF; ++k) infiltrated += b[6 + k];
  out[0] = b[4]; out[1] = b[5]; out[2] = b[0]; out[3] = b[1]; out[4] = b[2] - infiltrated; out[5] = b[3];
  return SFX_OK;
}

extern "C" int sfx_update_zbuvmx(sfx_handle* H) {
  if (!H) { set_error("null"); return SFX_ERR_ARG; }
  if (!H->hI.zb || !H->hI.zbuvmx) { set_error("sfx_update_zbuvmx: needs a regular (not subgrid) mesh"); return SFX_ERR_ARG; }
  HIPCHK(hipSetDevice(H->device));
  const int n = H->hI.dense ? 2 * H->hI.np : H->hI.npuv;
  const int nb = std::max(1, std::min(4096, (n + BLOCK - 1) / BLOCK));
  hipLaunchKernelGGL(k_zbuvmx, dim3(nb), dim3(BLOCK), 0, H->stream, H->dI);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(H->stream));
  return clear_steady(H);
}

extern "C" int sfx_reset_balance(sfx_handle* H) {
  if (!H) { set_error("null"); return SFX_ERR_ARG; }
  HIPCHK(hipSetDevice(H->device));
  hipLaunchKernelGGL(k_zero_bal, dim3(1), dim3(64), 0, H->stream, H->hI.bal, 0, NBAL);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(H->stream));
  return SFX_OK;
}

// ------------------------------------------------------- volume sum without allocation
// sfx_total_volume_noalloc: the volume sum of sfx_total_volume for callers that cannot
// allocate (graph capture, serving loops). Its partial sums live in `scratch`, one slot
// per block of VOL_CHUNK cells, and a one-wave kernel adds them in block order into the
// caller's device double, so the result is the same bits on every call. The array is
// allocated with the library image, one copy per GPU, and lives until the process ends;
// calls that overlap on one GPU share it and race.
constexpr int VOL_MAX_BLOCKS = 4096;
constexpr int VOL_CHUNK = BLOCK * 64;   // cells per block: at most 4096 * 16384 = 67,108,864 cells
static __device__ __align__(256) double scratch[VOL_MAX_BLOCKS];

namespace {
__global__ void __launch_bounds__(BLOCK) k_volume_partial(const DevInst* __restrict__ I0, double* part) {
  const DevInst& I = *I0;
  const int c0 = blockIdx.x * VOL_CHUNK;
  const int c1 = std::min(I.np, c0 + VOL_CHUNK);
  double acc = 0.0;
  for (int c = c0 + (int)threadIdx.x; c < c1; c += BLOCK) {
    const int k = G(I.kcs)[c] & KCS_MASK;
    if (k == 0) continue;
    const int iref = I.dense ? 1 : G(I.z_iref)[c];
    acc += I.subgrid ? G(I.vol)[c] : G(I.zs)[c] * (double)I.c.cell_area[iref - 1];
  }
  for (int o = WAVE / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, WAVE);
  __shared__ double red[BLOCK / WAVE];
  if (threadIdx.x % WAVE == 0) red[threadIdx.x / WAVE] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < BLOCK / WAVE; ++i) s += red[i];
    part[blockIdx.x] = s;
  }
}

__global__ void __launch_bounds__(WAVE) k_volume_final(const double* part, int nb, double* out) {
  double acc = 0.0;
  for (int b = threadIdx.x; b < nb; b += WAVE) acc += part[b];
  for (int o = WAVE / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, WAVE);
  if (threadIdx.x == 0) *out = acc;
}

// Device address of `scratch` per device, taken once with that device current.
constexpr int MAX_DEVICES = 64;
double* scratch_ptr[MAX_DEVICES] = {};
std::mutex scratch_mu;

int scratch_of(int device, double** out) {
  if (device < 0 || device >= MAX_DEVICES) { set_error("sfx_total_volume_noalloc: device index out of range"); return SFX_ERR_ARG; }
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[device]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    scratch_ptr[device] = (double*)p;
  }
  *out = scratch_ptr[device];
  return SFX_OK;
}
}  // namespace

extern "C" int sfx_total_volume_noalloc(sfx_handle* H, double* vol_device) {
  if (!H || !vol_device) { set_error("sfx_total_volume_noalloc: null argument"); return SFX_ERR_ARG; }
  const int nb = (H->hI.np + VOL_CHUNK - 1) / VOL_CHUNK;
  if (nb > VOL_MAX_BLOCKS) {
    set_error("sfx_total_volume_noalloc: the mesh has " + std::to_string(H->hI.np) + " cells, more than the " +
              std::to_string((long long)VOL_MAX_BLOCKS * VOL_CHUNK) + " its static scratch covers");
    return SFX_ERR_ARG;
  }
  HIPCHK(hipSetDevice(H->device));
  double* part = nullptr;
  int rc = scratch_of(H->device, &part);
  if (rc) return rc;
  if (nb > 0) hipLaunchKernelGGL(k_volume_partial, dim3(nb), dim3(BLOCK), 0, H->stream, H->dI, part);
  hipLaunchKernelGGL(k_volume_final, dim3(1), dim3(WAVE), 0, H->stream, part, nb, vol_device);
  HIPCHK(hipGetLastError());
  return SFX_OK;
}

// -------------------------------------------------------------------------- batch API
extern "C" int sfx_batch_create(int32_t device, sfx_batch** out) {
  if (!out) { set_error("null"); return SFX_ERR_ARG; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { set_error("no HIP device available"); return SFX_ERR_HIP; }
  HIPCHK(hipSetDevice(device));
  sfx_batch* B = new sfx_batch();
  B->device = device;
  HIPCHK(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
  *out = B;
  return SFX_OK;
}

extern "C" int sfx_batch_destroy(sfx_batch* B) {
  if (!B) return SFX_OK;
  hipSetDevice(B->device);
  if (B->stream) hipStreamSynchronize(B->stream);
  if (B->own_stream)
    for (auto* H : B->members) H->in_batches--;
  if (B->exec) hipGraphExecDestroy(B->exec);
  if (B->graph) hipGraphDestroy(B->graph);
  if (B->dInsts) hipFree(B->dInsts);
  if (B->dItems) hipFree(B->dItems);
  if (B->dIds) hipFree(B->dIds);
  if (B->own_stream && B->stream) hipStreamDestroy(B->stream);
  delete B;
  return SFX_OK;
}

extern "C" int sfx_batch_add(sfx_batch* B, sfx_handle* H, double t_end) {
  if (!B || !H) { set_error("null"); return SFX_ERR_ARG; }
  if (H->device != B->device) { set_error("member lives on another device"); return SFX_ERR_ARG; }
  for (auto* m : B->members) if (m == H) { set_error("already a member"); return SFX_ERR_ARG; }
  HIPCHK(hipSetDevice(B->device));
  Scalars s;
  int rc = normalise(H, s);
  if (rc) return rc;
  s.tend = t_end;
  if ((rc = write_scalars(H, s))) return rc;
  B->members.push_back(H);
  B->tend.push_back(t_end);
  B->live.push_back(1);
  B->dirty = true;
  H->in_batches++;
  return SFX_OK;
}

extern "C" int sfx_batch_remove(sfx_batch* B, sfx_handle* H) {
  if (!B || !H) { set_error("null"); return SFX_ERR_ARG; }
  HIPCHK(hipSetDevice(B->device));
  HIPCHK(hipStreamSynchronize(B->stream));
  for (size_t i = 0; i < B->members.size(); ++i)
    if (B->members[i] == H) {
      B->members.erase(B->members.begin() + i);
      B->tend.erase(B->tend.begin() + i);
      if (i < B->live.size()) B->live.erase(B->live.begin() + i);
      H->in_batches--;
      B->dirty = true;
      return SFX_OK;
    }
  set_error("not a member");
  return SFX_ERR_ARG;
}

extern "C" int sfx_batch_advance(sfx_batch* B, int32_t nsteps, int32_t* n_active) {
  if (!B) { set_error("null"); return SFX_ERR_ARG; }
  HIPCHK(hipSetDevice(B->device));
  int rc;
  if (B->members.empty()) { if (n_active) *n_active = 0; return SFX_OK; }
  if (B->dirty && (rc = rebuild(B))) return rc;
  if (B->n_live > 0) {
    const int K = std::max(1, std::min(nsteps, 64));
    if ((rc = ensure_graph(B, K))) return rc;
    for (int r = 0; r < nsteps / K; ++r) HIPCHK(hipGraphLaunch(B->exec, B->stream));
    for (int r = 0; r < nsteps % K; ++r) launch(B, 7, 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(B->stream));
  }
  int act = 0;
  for (size_t i = 0; i < B->members.size(); ++i) {
    Scalars s;
    HIPCHK(hipMemcpy(&s, B->members[i]->dS, sizeof(Scalars), hipMemcpyDeviceToHost));
    if (!s.error && !s.error_pending && s.t < s.tend) ++act;
    else if (B->live[i]) { B->live[i] = 0; B->dirty = true; }   // leaves the work lists
  }
  if (n_active) *n_active = act;
  return SFX_OK;
}

extern "C" int sfx_batch_done(sfx_batch* B, sfx_handle* H, int32_t* done) {
  if (!B || !H || !done) { set_error("null"); return SFX_ERR_ARG; }
  HIPCHK(hipSetDevice(B->device));
  HIPCHK(hipStreamSynchronize(B->stream));
  Scalars s;
  HIPCHK(hipMemcpy(&s, H->dS, sizeof(Scalars), hipMemcpyDeviceToHost));
  *done = (s.error || s.error_pending || !(s.t < s.tend)) ? 1 : 0;
  return SFX_OK;
}

// -------------------------------------------------------------------------- autotuner
extern "C" int sfx_get_tuning(sfx_handle* H, int32_t out[8]) {
  if (!H || !out) { set_error("null"); return SFX_ERR_ARG; }
  const DevInst& I = H->hI;
  out[0] = I.dense; out[1] = I.sk; out[2] = I.mtn; out[3] = I.mtm; out[4] = I.tn; out[5] = I.tm; out[6] = I.xchunk;
  out[7] = I.xchunk_c;
  return SFX_OK;
}

namespace {
// Mean duration (ms) of `trials` launches of the phases in `what` on the handle's stream.
// Median launch time of the phases `what` over `trials` launches (after one warm-up).
int time_phases(sfx_handle* H, int what, int trials, hipEvent_t e0, hipEvent_t e1, float& median) {
  sfx_batch* B = solo(H);
  B->dirty = true;
  int rc = rebuild(B);
  if (rc) return rc;
  launch(B, what, 0);   // warm-up
  HIPCHK(hipStreamSynchronize(H->stream));
  std::vector<float> t(trials);
  for (int k = 0; k < trials; ++k) {
    HIPCHK(hipEventRecord(e0, H->stream));
    launch(B, what, 0);
    HIPCHK(hipEventRecord(e1, H->stream));
    HIPCHK(hipEventSynchronize(e1));
    HIPCHK(hipEventElapsedTime(&t[k], e0, e1));
  }
  std::sort(t.begin(), t.end());
  median = t[trials / 2];
  return SFX_OK;
}

// No per-point tile is quiescent (after host writes of state and tile-shape changes).
int reset_tile_flags(sfx_handle* H) {
  HIPCHK(hipMemset(H->hI.mq, 0, H->nmt_cap * sizeof(uint32_t)));
  return SFX_OK;
}

// Tile minima back to dtmax (after sweeps that were not folded, e.g. timing runs), and no
// tile quiescent.
int reset_blkmin(sfx_handle* H) {
  std::vector<float> init(H->nmt_cap, H->hI.c.dtmax);
  HIPCHK(hipMemcpy(H->hI.blkmin, init.data(), init.size() * sizeof(float), hipMemcpyHostToDevice));
  return reset_tile_flags(H);
}

void set_shapes(sfx_handle* H, int sk, int mtn, int mtm, int tn, int tm) {
  DevInst& I = H->hI;
  I.sk = sk; I.mtn = mtn; I.mtm = mtm; I.tn = tn; I.tm = tm;
  I.mntn = (I.nmax + mtn - 1) / mtn;
  I.ntn = (I.nmax + tn - 1) / tn;
  I.tn_sh = ilog2(tn); I.mtn_sh = ilog2(mtn);
  I.mrow_sh = ilog2(mtm * (sk < 0 ? -sk : 1));
  I.nmt = (int32_t)mom_tiles(I);
}
}  // namespace

extern "C" int sfx_set_tuning(sfx_handle* H, const int32_t in[5]) {
  if (!H || !in) { set_error("null"); return SFX_ERR_ARG; }
  const int sk = in[0], mtn = in[1], mtm = in[2], tn = in[3], tm = in[4];
  const bool sk_ok = sk == -1 || sk == -2 || sk == -4 || (sk >= 0 && sk <= 256);
  if (!sk_ok || mtn <= 0 || mtm <= 0 || mtn * mtm != BLOCK || tn <= 0 || tm <= 0 || tn * tm != BLOCK) {
    set_error("sfx_set_tuning: the momentum kernel must be -1, -2, -4 (per-point), 0 (per-cell) or 1..256 strip rows, "
              "and tiles must hold 256 cells");
    return SFX_ERR_ARG;
  }
  if (H->in_batches) { set_error("sfx_set_tuning: set shapes before adding the scenario to a batch"); return SFX_ERR_STATE; }
  if (!H->hI.dense) return SFX_OK;
  HIPCHK(hipSetDevice(H->device));
  Scalars s;
  int rc = normalise(H, s);
  if (rc) return rc;
  const DevInst keep = H->hI;
  set_shapes(H, sk, mtn, mtm, tn, tm);
  if (mom_tiles(H->hI) > H->nmt_cap) {
    H->hI = keep;
    set_error("sfx_set_tuning: more momentum tiles than the handle holds");
    return SFX_ERR_ARG;
  }
  HIPCHK(hipMemcpy(H->dI, &H->hI, sizeof(DevInst), hipMemcpyHostToDevice));
  H->tile_tn = tn; H->tile_tm = tm; H->strip_rows = sk; H->mtile_tn = mtn; H->mtile_tm = mtm;
  if ((rc = reset_blkmin(H))) return rc;
  if (H->solo) H->solo->dirty = true;
  return SFX_OK;
}

extern "C" int sfx_set_work_order(sfx_handle* H, int32_t mom_run, int32_t cont_run) {
  if (!H) { set_error("null"); return SFX_ERR_ARG; }
  if (mom_run < 0 || cont_run < 0 || mom_run > (1 << 20) || cont_run > (1 << 20)) {
    set_error("sfx_set_work_order: runs per XCD must lie in 0..2^20 (0: one contiguous band per XCD)");
    return SFX_ERR_ARG;
  }
  if (H->in_batches) { set_error("sfx_set_work_order: set the order before adding the scenario to a batch"); return SFX_ERR_STATE; }
  HIPCHK(hipSetDevice(H->device));
  H->hI.xchunk = mom_run; H->hI.xchunk_c = cont_run;
  H->xchunk = mom_run; H->xchunk_c = cont_run;
  HIPCHK(hipMemcpy(H->dI, &H->hI, sizeof(DevInst), hipMemcpyHostToDevice));
  if (H->solo) H->solo->dirty = true;
  return SFX_OK;
}

extern "C" int sfx_autotune(sfx_handle* H, int32_t trials) {
  if (!H) { set_error("null"); return SFX_ERR_ARG; }
  HIPCHK(hipSetDevice(H->device));
  if (H->in_batches) { set_error("sfx_autotune: tune a scenario before adding it to a batch"); return SFX_ERR_STATE; }
  const bool dense = H->hI.dense;   // indexed meshes have one kernel shape; only the work order is tuned
  trials = std::max(1, (int)trials);
  struct Pause { sfx_handle* h; ~Pause() { h->tsa_paused = false; } } pause{H};
  H->tsa_paused = true;   // timed sweeps are not steps of the run
  Scalars s0;
  int rc = normalise(H, s0);
  if (rc) return rc;
  // momentum and continuity write state: time on a snapshot and restore it afterwards
  std::vector<std::string> fields = {"zs", "q", "uv", "kfuv", "zs0", "zsderv", "zsmax"};
  if (H->hI.vol) fields.push_back("z_volume");
  if (H->hI.vmax) fields.push_back("vmax");
  if (H->hI.qmax) fields.push_back("qmax");
  if (H->hI.twet) fields.push_back("twet");
  if (H->hI.storage) fields.push_back("storage_volume");
  std::map<std::string, std::vector<char>> snap;
  for (auto& f : fields) {
    size_t n = 0;
    if (f == "zs" || f == "z_volume") n = (size_t)H->hI.np * 8;
    else if (f == "q" || f == "uv") n = H->ref2slot.size() * (H->fp64 ? 8 : 4);
    else if (f == "kfuv") n = H->sz.npuv;
    else n = (size_t)H->hI.np * (H->fp64 ? 8 : 4);
    snap[f].resize(n);
    if ((rc = sfx_get_state(H, f.c_str(), snap[f].data(), n))) return rc;
  }
  double bal_keep[NBAL];   // timed sweeps add to the volume balance: restored afterwards
  HIPCHK(hipMemcpy(bal_keep, H->hI.bal, sizeof(bal_keep), hipMemcpyDeviceToHost));
  Scalars s = s0;
  s.active = 1; s.dt = std::max(s0.dt, 1.0e-3f);
  if ((rc = write_scalars(H, s))) return rc;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const int tn0 = H->hI.tn, tm0 = H->hI.tm;
  // momentum: per-cell tiles, per-point tiles with 1, 2 or 4 points per thread, strip
  // marches. Each trial times momentum with the continuity sweep after it (the continuity
  // tiles stay fixed meanwhile), so that the per-point kernel's quiescent tiles are the ones
  // a step leaves, not those of repeated sweeps over unchanged levels.
  float best = 1e30f;
  int bsk = H->hI.sk, bmtn = H->hI.mtn, bmtm = H->hI.mtm;
  const int kinds[] = {0, -1, -2, -4};
  constexpr int NK = sizeof(kinds) / sizeof(kinds[0]);
  for (int k = 0; dense && k < NK * N_CELL_SHAPES + N_STRIP_ROWS && !rc; ++k) {
    const bool tiled = k < NK * N_CELL_SHAPES;
    const int sk = tiled ? kinds[k / N_CELL_SHAPES] : STRIP_ROWS[k - NK * N_CELL_SHAPES];
    const int mtn = tiled ? CELL_SHAPES[k % N_CELL_SHAPES][0] : bmtn;
    const int mtm = tiled ? CELL_SHAPES[k % N_CELL_SHAPES][1] : bmtm;
    set_shapes(H, sk, mtn, mtm, tn0, tm0);
    float ms = 0.f;
    rc = time_phases(H, 6, trials, e0, e1, ms);
    if (!rc && ms < best) { best = ms; bsk = sk; bmtn = mtn; bmtm = mtm; }
  }
  // momentum work-list order (runs of tiles per XCD)
  set_shapes(H, bsk, bmtn, bmtm, tn0, tm0);
  const int chunks[] = {64, 0, 16, 256, 1024};
  best = 1e30f;
  int bx = H->hI.xchunk;
  for (int k = 0; k < 5 && !rc; ++k) {
    H->hI.xchunk = chunks[k];
    float ms = 0.f;
    rc = time_phases(H, 6, trials, e0, e1, ms);
    if (!rc && ms < best) { best = ms; bx = chunks[k]; }
  }
  H->hI.xchunk = bx;
  // continuity tiles, then their work-list order
  best = 1e30f;
  int btn = tn0, btm = tm0;
  for (int k = 0; dense && k < N_CELL_SHAPES && !rc; ++k) {
    set_shapes(H, bsk, bmtn, bmtm, CELL_SHAPES[k][0], CELL_SHAPES[k][1]);
    float ms = 0.f;
    rc = time_phases(H, 4, trials, e0, e1, ms);
    if (!rc && ms < best) { best = ms; btn = CELL_SHAPES[k][0]; btm = CELL_SHAPES[k][1]; }
  }
  set_shapes(H, bsk, bmtn, bmtm, btn, btm);
  best = 1e30f;
  int bxc = H->hI.xchunk_c;
  for (int k = 0; k < 5 && !rc; ++k) {
    H->hI.xchunk_c = chunks[k];
    float ms = 0.f;
    rc = time_phases(H, 4, trials, e0, e1, ms);
    if (!rc && ms < best) { best = ms; bxc = chunks[k]; }
  }
  H->hI.xchunk_c = bxc;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc) return rc;
  HIPCHK(hipMemcpy(H->dI, &H->hI, sizeof(DevInst), hipMemcpyHostToDevice));
  H->tile_tn = H->hI.tn; H->tile_tm = H->hI.tm; H->strip_rows = H->hI.sk;
  H->mtile_tn = H->hI.mtn; H->mtile_tm = H->hI.mtm; H->xchunk = H->hI.xchunk; H->xchunk_c = H->hI.xchunk_c;
  if ((rc = write_scalars(H, s0))) return rc;
  if ((rc = reset_blkmin(H))) return rc;
  for (auto& f : fields)
    if ((rc = sfx_set_state(H, f.c_str(), snap[f].data(), snap[f].size()))) return rc;
  HIPCHK(hipMemcpy(H->hI.bal, bal_keep, sizeof(bal_keep), hipMemcpyHostToDevice));
  if (H->solo) H->solo->dirty = true;
  return SFX_OK;
}

// -------------------------------------------------------------------------- profiling
// Runs nsteps ordinary (non-graph) steps of one handle with HIP events around each phase
// on the stream the kernels are launched on; ms[0..2] receive the mean duration of the
// forcing, momentum (+combined points) and continuity launches.
extern "C" int sfx_profile_steps(sfx_handle* H, int32_t nsteps, double ms[3]) {
  if (!H || !ms || nsteps < 1) { set_error("sfx_profile_steps: bad args"); return SFX_ERR_ARG; }
  HIPCHK(hipSetDevice(H->device));
  sfx_batch* B = solo(H);
  int rc;
  if (B->dirty && (rc = rebuild(B))) return rc;
  Scalars s;
  if ((rc = read_scalars(H, s))) return rc;
  s.tend = 1.0e30;
  if ((rc = write_scalars(H, s))) return rc;
  hipEvent_t ev[4];
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  double acc[3] = {0, 0, 0};
  for (int i = 0; i < nsteps; ++i) {
    HIPCHK(hipEventRecord(ev[0], B->stream));
    launch(B, 1, 0);
    HIPCHK(hipEventRecord(ev[1], B->stream));
    launch(B, 2, 0);
    HIPCHK(hipEventRecord(ev[2], B->stream));
    launch(B, 4, 0);
    HIPCHK(hipEventRecord(ev[3], B->stream));
    HIPCHK(hipEventSynchronize(ev[3]));
    for (int k = 0; k < 3; ++k) {
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
      acc[k] += t;
    }
  }
  for (auto& e : ev) hipEventDestroy(e);
  for (int k = 0; k < 3; ++k) ms[k] = acc[k] / nsteps;
  return SFX_OK;
}
