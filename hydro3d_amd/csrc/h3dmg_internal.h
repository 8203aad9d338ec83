// h3dmg_internal.h -- device data layout and kernel launchers (not part of
// the C-ABI).
//
// Layout in HBM.  Every level of every block stores its cells COLOUR-SPLIT:
// the red cells ((i+j+k) even, 0-based interior indices -- MOD(i+j+k,2)==0 in
// the reference's 1-based padded indices, src_v70/mgsolver.for:556) and the
// black cells live in two separate arrays.  Cell (i,j,k) of colour c sits at
//     ((k*nj + j)*H + (i >> 1))            H = row pitch >= ceil(ni/2)
// so one colour pass streams whole cache lines of one array (3 words per cell
// per red+black sweep instead of 6 for interleaved colours).
//
// There is no ghost storage.  Every ghost value the reference would hold after
// mgbound + exchangepp (mgsolver.for:614-615, exchangepp.for:2830-2832) is
// the value of one interior cell, its IMAGE: the mirror (first interior) at a
// wall and the wrap (opposite interior) at a periodic face.  Kernels read
// images directly ("virtual ghosts").  The one place where the reference's
// ghosts are stale -- the black half-sweep reads ghosts from before the red
// half-sweep, because exchangepp runs once per full sweep -- is reproduced by
// double buffering: sweep s reads buffer b and writes buffer 1-b, and ghost
// reads always go to buffer b.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace h3dmg {

constexpr int TILE_X = 64;  // one wavefront along the contiguous axis
constexpr int TILE_Y = 4;   // four wavefronts per workgroup
constexpr int NTHREADS = TILE_X * TILE_Y;
constexpr int TAIL_THREADS = 512;  // V-cycle tail: one workgroup per block
constexpr int TAIL_U = 4;           // cells per thread with loads in flight together
constexpr int TAIL_MAX_LEVELS = 16;

struct LevelDev {
  int ni, nj, nk;  // interior cells
  int H;           // colour row pitch (elements), even, >= ceil(ni/2)
  int per;         // bit a set: axis a periodic (a = 0 x, 1 y, 2 z)
  int lpr;         // smoother lanes per row (power of two <= 64)
  int lpr1;        // restriction lanes per coarse row segment (two coarse cells per lane)
  int blk;         // block index
  int finer;       // descriptor index of level g-1 (or -1)
  int coarser;     // descriptor index of level g+1 (or -1)
  // same-level neighbour across face f (west, east, south, north, bottom,
  // top): its descriptor index, or -1 (a wall, or a periodic axis that wraps
  // onto this block: `per`).  A face is a wall when it has no neighbour and
  // its axis is not periodic.
  int nbr[6];
  double cx, cy, cz;  // 1/h_g^2 (coeff, mgsolver.for:54-60,185-191)
  void *phi[2][2];    // [buffer][colour]
  void *rhs[2];       // [colour]
  // variable-density operator (coeff under L_LSM, mgsolver.for:72-184), or
  // null: per-cell FP64 coefficients aw, ae, as, an, ab, at, ap, each as two
  // colour arrays of lsm_half doubles laid out like phi (array 2*q + colour)
  double *lsm;
  unsigned long long lsm_half;
  // line-TDMA tables of a constant-coefficient level, or null: per pass
  // (j-, i-, k-lines) and wall-state pair of the line (16), a(p) and
  // 1/(ap - alo*a(p-1)) at every position p of the line
  double *tdma;
};

// Finest-level FP64 state of one block (pressure, right-hand side).
struct FineDev {
  int ni, nj, nk, H, per, blk;
  int nbr[6];  // neighbouring BLOCK across face f (the level-1 descriptors), or -1
  int lpr;  // residual lanes per row segment (two colour-array entries per lane)
  double cx, cy, cz;
  double *p[2];  // [colour]
  double *f[2];  // [colour]
  double dx, dy, dz;     // spacing (the variable-density projection's volume factors)
  const double *dens;    // variable-density operator: density, interior (nk, nj, ni), or null
};

// A workgroup's work item: descriptor index, the tile origin, for the
// plane-marching kernels the end plane z1 (exclusive), and the tile kind of
// the smoother's tile lists.
struct Tile {
  int lev;
  int x0, y0, z0;
  int z1;
  int kind;
};

struct RedPartial {
  double ss;    // sum of squares
  double amax;  // max |value|
};

// ---- launchers (h3dmg_kernels.hip) --------------------------------------
// T = storage type of the level (double or float).
// one colour half-sweep of the red-black smoother over a tile list whose
// tiles are all `depth` (1, 2 or 4) planes deep: kinds 0 (march tile, y/z
// interior), 3 (march tile touching a y/z face), or, on levels one cell wide
// in x (depth 1 only), 1 (general boundary march) and 2 (x faces)
// A launch whose tiles all belong to one descriptor (lev; lev2: its finer
// level for a restriction, its coarser for a prolongation; blk its block):
// the kernel then reads the tile, the descriptors and the block's flag at
// once instead of one after the other.  lev = -1: tiles of several.
struct OneDesc {
  int lev = -1, lev2 = -1, blk = -1;
};
template <class T>
void launch_rb_sweep(hipStream_t s, const Tile *tiles, int ntiles, int depth, const LevelDev *levs,
                     const int *active, int colour, int src, int zero, OneDesc one = OneDesc{});
// one full red-black sweep (both colour halves) of small levels in one
// launch: tiles of FUSED_X x FUSED_Y x FUSED_Z cells (Tile::x0, y0, z0 in
// cells), kind 4
constexpr int FUSED_X = 32, FUSED_Y = 8, FUSED_Z = 4;
template <class T>
void launch_rb_fused(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                     const int *active, int src, int zero);
// restriction into the coarse levels of a tile list whose tiles are all
// `depth` (1, 2 or 4) coarse planes deep
// (depth 0: natural tiles of the coarse level and the general kernel, for
// blocks with neighbours).  zbuf >= 0 (depth > 0 only): the coarse level is
// relaxed next from zero; the restriction also writes the red half of that
// first sweep (each red cell rhs/ap, its neighbours being zero) into buffer
// zbuf, and the sweep starts with its black half.
template <class Tf, class Tc>
void launch_restrict(hipStream_t s, const Tile *tiles, int ntiles, int depth, const LevelDev *levs,
                     const int *active, int src_f, OneDesc one = OneDesc{}, int zbuf = -1);
// gen != 0: coarse ghosts read through the neighbour table (blocks with
// neighbours)
template <class Tf, class Tc>
void launch_prolong(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                    const int *active, int src_f, int src_c, int gen, OneDesc one = OneDesc{});
template <class T>
void launch_fill_zero(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                      const int *active, int buf);
// tiles of the finest level, all `depth` (1, 2, 4 or 8) planes deep, or
// depth 0 for sets that hold general (kind 1) tiles:
// Tile::lev is the BLOCK index b; fine[b] and levs[b] (level-1 descriptors
// come first) describe it.  use_p = 0 gives r = f (the norm of the
// right-hand side).
// zbuf >= 0 (the fast path, depth > 0 and use_p): also the red half of the
// level-1 relaxation from zero (each red cell rhs/ap) into buffer zbuf
template <class T>
void launch_residual(hipStream_t s, const Tile *tiles, int ntiles, int depth, const FineDev *fine,
                     const LevelDev *levs, const int *active, RedPartial *partials, int use_p,
                     int zbuf = -1);
template <class T>
void launch_update(hipStream_t s, const Tile *tiles, int ntiles, const FineDev *fine,
                   const LevelDev *levs, const int *active, int buf);
void launch_finalize(hipStream_t s, const int *tile_begin, int nblocks, const RedPartial *partials,
                     RedPartial *out);
// one workgroup per tile; Tile::lev = coarsest descriptor, Tile::x0 = block
void launch_coarse_exact(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                         const int *active, const double *const *eig_q,
                         const double *const *eig_l, double *const *scratch, int buf);

// levels >= the tail level of every block in one launch, one workgroup per
// block, every level in LDS (Tile::lev = the block's descriptor at the tail
// level); every level ends in buffer (nu1+nu2)&1.  vtail_lds_bytes gives the
// LDS a block needs from descriptor `top` down (host copies of the
// descriptors).
void launch_vtail(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                  const int *active, const double *const *eig_q, const double *const *eig_l,
                  int nu1, int nu2, int exact, int mixed, size_t lds_bytes);
size_t vtail_lds_bytes(const LevelDev *levs_host, int top, bool mixed, bool exact);
// permit `lds_bytes` of dynamic LDS for the tail kernel
hipError_t vtail_allow_lds(size_t lds_bytes);

// ---- variable-density operator (L_LSM) ----------------------------------
// dens_mg of level `lev` (interior, natural layout) from the finer level's:
// 0.125 * (sum of the 8 children, j fastest, then i, k: mgsolver.for:87-118)
void launch_lsm_dens_restrict(hipStream_t s, const double *fine, int fni, int fnj, int fnk,
                              double *coarse);
// coefficients of one level of one block from its dens_mg into L.lsm
// (mgsolver.for:121-184,194-210); dx, dy, dz the block's finest spacing
void launch_lsm_coef(hipStream_t s, const LevelDev *lev_dev, const double *dens, int ni, int nj,
                     int nk, double dx, double dy, double dz, int walls);
// one colour half-sweep / the fine residual / the restriction with the
// per-cell coefficients (every level of a variable-density handle is FP64)
void launch_lsm_sweep(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                      const int *active, int colour, int src, int zero);
void launch_lsm_residual(hipStream_t s, const Tile *tiles, int ntiles, const FineDev *fine,
                         const LevelDev *levs, const int *active, RedPartial *partials);
void launch_lsm_restrict(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                         const int *active, int src_f);

// ---- lexicographic GS and line-TDMA smoothers (mg_itrsch = 1, 2) ---------
// Tiles: GS tiles of GS_TILE^3 cells (Tile::x0, y0, z0), one launch per
// tile-diagonal (the tiles whose tile indices sum to D, every block); TDMA
// tiles of tdma_group_lines(len) lines of one pass (Tile::x0 = first line),
// a line of len cells, one launch per line diagonal d (pass 0: j-lines,
// i+k = d; 1: i-lines, j+k = d; 2: k-lines, i+j = d).  A sweep reads buffer
// src and writes 1-src; every level is FP64.
constexpr int GS_TILE = 8;
// 16^3 GS tiles on constant-coefficient levels: a tile-diagonal of a 64^3
// sweep holds 10 launches instead of 22; the tile, its halo and its rhs in
// dynamic LDS
constexpr int GS_BIG_TILE = 16;
constexpr size_t GS_BIG_LDS =
    sizeof(double) * ((GS_BIG_TILE + 2) * (GS_BIG_TILE + 2) * (GS_BIG_TILE + 2) +
                      GS_BIG_TILE * GS_BIG_TILE * GS_BIG_TILE);
// permits GS_BIG_LDS of dynamic LDS for the 16^3 GS kernel (once per process)
hipError_t gs_allow_lds();
constexpr int TDMA_CELLS = 1024;    // line cells one TDMA workgroup stages in LDS
constexpr int TDMA_MAX_LEN = 1024;  // longest line (cells along one axis)
__host__ __device__ inline int tdma_group_lines(int len) {
  int lw = 1;
  while (lw < NTHREADS && 2 * lw * len <= TDMA_CELLS) lw *= 2;
  return lw;
}
void launch_gs_tiles(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                     const int *active, int src, int zero, int tile);
void launch_tdma_diag(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                      const int *active, int pass, int d, int src, int zero);
// a whole TDMA sweep of a level, one workgroup of TDMA_SWEEP_THREADS per
// block (Tile::lev), every line of a diagonal and the previous diagonal's
// solution in LDS: tdma_sweep_lds_bytes(level) of dynamic LDS (the largest
// over the launch), TDMA_SWEEP_THREADS >= the lines of a diagonal; used
// where every pass has at most TDMA_SWEEP_CELLS cells per diagonal and the
// LDS fits the device (64^3 levels: 130 KB)
constexpr int TDMA_SWEEP_THREADS = 1024;
constexpr int TDMA_SWEEP_CPT = 2;  // cells per thread in each load batch
constexpr int TDMA_SWEEP_CELLS = 4096;
// pipelined: constant coefficients (the pipelined whole-sweep kernel and
// its larger LDS layout; the host decides, before the level's tables exist)
void launch_tdma_sweep(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                       const int *active, int src, int zero, size_t lds_bytes, bool pipelined);
size_t tdma_sweep_lds_bytes(const LevelDev &L, bool pipelined);
// permits `lds_bytes` of dynamic LDS for the whole-sweep TDMA kernel
hipError_t tdma_sweep_allow_lds(size_t lds_bytes);
// fill LevelDev::tdma of every descriptor that has one (32*(ni+nj+nk) doubles)
void launch_tdma_tables(hipStream_t s, const LevelDev *levs, int ndesc);

// velocity projection (calvel per axis over `faces` faces; calmas writes
// su (optional) and one partial max per workgroup, returns the group count)
constexpr int CALMAS_MAX_GROUPS = 1024;
void launch_calvel(hipStream_t s, const FineDev *fd, int axis, const double *us, double *u,
                   double dt_alfa, double h, size_t faces);
// blocks with neighbours: calvel of block b (fine = every block's
// descriptor) on the faces the block computes, and the copy of a near face
// from the far face of the block across it (dims of the face: nu x nv)
void launch_calvel_nb(hipStream_t s, const FineDev *fine, int b, int axis, const double *us,
                      double *u, double dt_alfa, double h, size_t faces);
void launch_face_copy(hipStream_t s, int axis, const double *src, int n_src, double *dst,
                      int n_dst, int nu, int nv);
// p of block b with its ghost layer (dense padded, ghosts resolved)
void launch_unpack_fine_padded(hipStream_t s, const FineDev *fine, int b, double *dst, int ni,
                               int nj, int nk);
int launch_calmas(hipStream_t s, const FineDev *fd, const double *u, const double *v,
                  const double *w, double *su, int to_fine, double dt, double dx, double dy,
                  double dz, RedPartial *partials, size_t cells);

// ---- the solve loop's stopping rule on the device ------------------------
// h3dmg_solve keeps one cycle queued ahead of the host: after each cycle one
// workgroup applies the loop's rule (per block REL_L2 against its tolerance,
// or the reference's global rmax criterion; min_cycles; a non-finite norm
// stops every block) to that cycle's statistics, writes the next cycle's
// block flags, and copies the outcome to host memory.
struct DecideBlock {  // per block, set when the solve starts
  double fnorm;       // ||f||_2
  double tol;         // its REL_L2 tolerance
  double vol;         // dx*dy*dz
  int enabled;
  int pad;
};
struct DecideState {  // per block, carried across cycles
  double rel;         // relative residual after the last cycle it ran
  double rmax;        // max |r| * rfac * vol after that cycle
};
struct DecideCtl {
  int it;        // cycles run
  int all_done;  // no block active after cycle it
  int bad;       // a non-finite residual norm after cycle it
  int pad;
  double worst;  // max rel over enabled blocks (REF: max rmax)
};
struct DecideParams {
  int nb, ref, connected, min_cycles;
  double tol, rfac;
};
// ring slot of one cycle in host memory: DecideCtl, then nb DecideState, nb
// RedPartial and nb int (the flags for the next cycle); slot it & 1
__host__ __device__ inline size_t decide_slot_bytes(int nb) {
  return (sizeof(DecideCtl) + (size_t)nb * (sizeof(DecideState) + sizeof(RedPartial) + sizeof(int)) +
          255) & ~(size_t)255;
}
// launch_finalize, then the rule in the last of the nblocks workgroups to
// finish (prm in device memory: set when a solve starts, the captured graph
// fixed); *arrived counts the workgroups (zero before the first launch; the
// last workgroup resets it)
void launch_finalize_decide(hipStream_t s, const int *tile_begin, int nblocks,
                            const RedPartial *partials, RedPartial *out, const DecideBlock *blocks,
                            DecideState *state, DecideCtl *ctl, int *active, char *ring,
                            const DecideParams *prm, unsigned *arrived);

// *out = max over n partials of RedPartial::amax (one workgroup)
void launch_partials_max(hipStream_t s, const RedPartial *parts, int n, double *out);
// workgroups launch_calmas uses for `cells` cells
inline int calmas_groups(size_t cells) {
  return (int)((cells + 255) / 256 < (size_t)CALMAS_MAX_GROUPS ? (cells + 255) / 256
                                                                : (size_t)CALMAS_MAX_GROUPS);
}

// calmas' 4th-order branch (press.for:17-37) reads faces -1 and n+1 of
// every axis.  For axis a, side s (0: face -1, 1: face n+1), entry 2a+s:
// n > 0: src is the dense face array of the block across that face (itself
// on an axis it spans periodically) with n cells along a, whose face n-1
// (s = 0) or face 1 (s = 1) is the value -- what exchange(1..3) puts in the
// second ghost layer; n == 0: src is this block's plane of host values at
// that face (a wall), indexed like a face of the dense arrays with the
// face coordinate dropped.
struct Outer4 {
  const double *src[6];
  int n[6];
};
// calmas with the 4th-order divergence: same outputs as launch_calmas
int launch_calmas4(hipStream_t s, const FineDev *fd, const double *u, const double *v,
                   const double *w, double *su, int to_fine, double dt, double dx, double dy,
                   double dz, RedPartial *partials, size_t cells, const Outer4 &outer);

// layout conversion: dense interior (i fastest) <-> colour split
void launch_pack_fine(hipStream_t s, const FineDev *fd, int which, const double *src, int ni,
                      int nj, int nk);
void launch_unpack_fine(hipStream_t s, const FineDev *fd, int which, double *dst, int ni, int nj,
                        int nk);
template <class T>
void launch_pack_level(hipStream_t s, const LevelDev *ld, int which, int buf, const double *src,
                       int ni, int nj, int nk);
template <class T>
void launch_unpack_level(hipStream_t s, const LevelDev *levs, int d, int which, int buf, double *dst,
                         int ni, int nj, int nk);
void launch_remove_mean(hipStream_t s, const FineDev *fd, int ni, int nj, int nk, double *scratch);

}  // namespace h3dmg
