// h3dmg_runtime.hip -- host runtime behind the C-ABI of include/h3dmg.h.
//
// Owns the device arena (every level of every block), the tile tables the
// kernels read, the V-cycle schedule of src_v70/mgsolver.for:226-372, the
// pressure-form outer loop of src_v70/newsolv_mg.for:19-59 and the hipGraph
// that replays one cycle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/h3dmg.h"
#include "h3dmg_internal.h"

using namespace h3dmg;

namespace {

thread_local std::string g_last_error;
// Serialises graph capture with the calls that allocate, free or copy
// synchronously (handle creation and destruction, tile uploads): handles may
// be driven from several host threads at once.
std::mutex g_device_mutex;

h3dmg_status fail(h3dmg_status st, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return st;
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(H3DMG_ERR_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                                  \
  } while (0)

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace

// ---- scratch of the allocation-free projection ----------------------------
// calmas' per-workgroup partial maxima of h3dmg_project_device_noalloc live
// in this file-scope device array (one copy per GPU, held until the process
// ends) instead of a handle's arena.  Its name is a host-side stand-in, not a
// device address: host code never passes `scratch` to a kernel or a copy.
// Each device's address is looked up once, with that device current
// (hipGetSymbolAddress), and kept in g_scratch_ptr.
constexpr int NOALLOC_PARTIALS = CALMAS_MAX_GROUPS;  // 16 KiB
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) RedPartial scratch[NOALLOC_PARTIALS];

namespace {
std::mutex g_scratch_mutex;
RedPartial *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// the device address of `scratch` on device `dev`
h3dmg_status scratch_on(int dev, RedPartial **out) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
    return fail(H3DMG_ERR_UNSUPPORTED, "device %d: beyond the %d devices the scratch table holds",
                dev, NOALLOC_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    int cur = -1;
    HIP_TRY(hipGetDevice(&cur));
    if (cur != dev) HIP_TRY(hipSetDevice(dev));
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != dev) HIP_TRY(hipSetDevice(cur));
    if (e != hipSuccess)
      return fail(H3DMG_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                  hipGetErrorString(e));
    if (!ptr) return fail(H3DMG_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[dev] = (RedPartial *)ptr;
  }
  *out = g_scratch_ptr[dev];
  return H3DMG_OK;
}

struct TileSet {
  std::vector<Tile> h;
  Tile *d = nullptr;
  Tile *owned = nullptr;  // smoother sets: separately allocated (re-tunable)
  int depth = 1;          // smoother sets: planes of every tile (1, 2 or 4)
  bool fused = false;     // smoother sets: fused red-black tiles (one launch per sweep)
  const TileSet *nat = nullptr;  // smoother sets: the group's natural tiles (variable density)
  int ndiag[3] = {0, 0, 0};      // GS / TDMA sets: launches per sweep (pass)
  int pass_begin[4] = {0, 0, 0, 0};  // TDMA sets: tiles of pass q at [begin[q], begin[q+1])
  std::vector<int> diag_begin;       // GS sets: tiles of tile-diagonal D at [begin[D], begin[D+1])
  int gs_tile = GS_TILE;             // GS sets: tile edge (GS_TILE or GS_BIG_TILE)
  size_t tdma_lds = 0;  // TDMA sets: > 0 = whole sweeps in one workgroup per block (LDS bytes)
  bool tdma_pipe = false;  // TDMA whole sweeps: the pipelined kernel (constant coefficients)
  OneDesc one;             // every tile on one descriptor: passed to the kernels
  int n() const { return (int)h.size(); }
};

// ts.one: the descriptor of every tile of the set, if they share one, with
// its finer (rel < 0, restrictions) or coarser (rel > 0, prolongations)
// level and its block
void set_one(TileSet &ts, const std::vector<LevelDev> &levs, int rel) {
  ts.one = OneDesc{};
  if (ts.h.empty()) return;
  const int lev = ts.h[0].lev;
  for (const Tile &t : ts.h)
    if (t.lev != lev) return;
  ts.one.lev = lev;
  ts.one.blk = levs[lev].blk;
  ts.one.lev2 = rel < 0 ? levs[lev].finer : rel > 0 ? levs[lev].coarser : -1;
}

// Append the tiles covering an (nx, ny, nz) index space of descriptor `lev`.
void add_tiles(TileSet &ts, int lev, int nx, int ny, int nz) {
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; y += TILE_Y)
      for (int x = 0; x < nx; x += TILE_X) ts.h.push_back(Tile{lev, x, y, z, z + 1});
}

// Smoother lanes per row: two entries per lane, a power of two so that a
// wavefront holds 64/lpr whole rows when a row is short.
int lanes_per_row(int H) {
  int l = 1;
  while (l < 64 && 2 * l < H) l *= 2;
  return l;
}

// Planes per workgroup of one plane-marching launch (every tile of a launch
// has the same depth, the kernels are instantiated per depth): `want` when
// > 0 (the autotuner's choice), else enough for ~`target` workgroups over
// `cols` = sum of tile columns x planes; capped at `cap`, then the largest
// power of two not above it that divides every block's nk in the launch.
int launch_depth(long cols, long target, int want, int cap, const std::vector<int> &nks) {
  long kb = want > 0 ? want : std::max<long>(1, cols / target);
  kb = std::min<long>(kb, cap);
  int d = 1;
  while (2 * d <= kb) d *= 2;
  for (int nk : nks)
    while (nk % d) d >>= 1;
  return d;
}

// Smoother tiles of level descriptor `lev`, in plane-chunk order (the XCD
// remap then hands every XCD a contiguous slab): per chunk of kb planes the
// march tiles (2*lpr entries x 4*(64/lpr) rows x kb planes), kind 0 when
// every y/z neighbour of the tile is inside the block, kind 3 otherwise; both
// update the x-face cells of their rows.  Levels one cell wide in x use the
// general boundary tiles (kind 1) and separate x-face tiles (kind 2, 256
// rows of one plane each).
int march_tiles_xy(const LevelDev &L) {
  const int xs = 2 * L.lpr, ys = TILE_Y * (64 / L.lpr);
  return ((L.H + xs - 1) / xs) * ((L.nj + ys - 1) / ys);
}

// kb planes per tile; kb divides L.nk; `general`: kinds 1 and 2 (every
// ghost through the neighbour table: blocks with neighbours)
void add_tiles_march(TileSet &ts, int lev, const LevelDev &L, int kb, bool general) {
  const int xs = 2 * L.lpr, ys = TILE_Y * (64 / L.lpr);
  const int nx = L.H, ny = L.nj, nz = L.nk;
  const bool faces_in_tiles = L.ni >= 2 && !general;
  for (int z = 0; z < nz; z += kb) {
    const int z1 = std::min(z + kb, nz);
    if (!faces_in_tiles)
      for (int zz = z; zz < z1; ++zz)
        for (int y = 0; y < ny; y += NTHREADS) ts.h.push_back(Tile{lev, y, zz, 0, 0, 2});
    for (int y = 0; y < ny; y += ys)
      for (int x = 0; x < nx; x += xs) {
        const bool inner = y >= 1 && y + ys <= ny - 1 && z >= 1 && z1 <= nz - 1;
        const int kind = faces_in_tiles ? (inner ? 0 : 3) : 1;
        ts.h.push_back(Tile{lev, x, y, z, z1, kind});
      }
  }
}

// Restriction tiles of coarse level descriptor `lev` (k_restrict): cpl*lpr1
// coarse cells (cpl per lane: 2 from an FP64 fine level, 1 from an FP32 one)
// x 4*(64/lpr1) coarse rows x kb coarse planes; kind 0 when every
// fine row and plane the tile reads is inside the block.  Deeper columns than
// the smoother's (~2048 workgroups): a column re-reads two fine planes.
int restrict_tiles_xy(const LevelDev &C, int cpl) {
  const int xs = cpl * C.lpr1, ys = TILE_Y * (64 / C.lpr1);
  return ((C.ni + xs - 1) / xs) * ((C.nj + ys - 1) / ys);
}

// kb coarse planes per tile (kb divides C.nk); kind 3 for tiles touching a
// y/z face
void add_tiles_restrict(TileSet &ts, int lev, const LevelDev &C, int kb, int cpl) {
  const int xs = cpl * C.lpr1, ys = TILE_Y * (64 / C.lpr1);
  const int nx = C.ni, ny = C.nj, nz = C.nk;
  for (int z = 0; z < nz; z += kb) {
    const int z1 = z + kb;
    for (int y = 0; y < ny; y += ys)
      for (int x = 0; x < nx; x += xs) {
        const bool inner = y >= 1 && y + ys <= ny - 1 && z >= 1 && z1 <= nz - 1;
        ts.h.push_back(Tile{lev, x, y, z, z1, inner ? 0 : 3});
      }
  }
}

// One depth for a restriction launch over the coarse descriptors `lev`:
// ~2048 workgroups, at most 4 coarse planes.
int restrict_depth(const std::vector<LevelDev> &levs, const std::vector<int> &lev, int cpl) {
  long cols = 0;
  std::vector<int> nks;
  for (int d : lev) {
    cols += (long)restrict_tiles_xy(levs[d], cpl) * levs[d].nk;
    nks.push_back(levs[d].nk);
  }
  return launch_depth(cols, 2048, 0, 4, nks);
}

// Prolongation tiles of fine level descriptor `lev` (k_prolong): fine entries
// x0.. (two per thread, 2*TILE_X per tile) x coarse row pairs m = y0..
// (TILE_Y per tile) x one coarse plane pair n = z0, with m in [0, nj/2] and n
// in [0, nk/2].
void add_tiles_prolong(TileSet &ts, int lev, const LevelDev &F) {
  const int ncj = F.nj / 2, nck = F.nk / 2, nfx = F.ni / 2;
  for (int n = 0; n <= nck; ++n)
    for (int m = 0; m <= ncj; m += TILE_Y)
      for (int x = 0; x < nfx; x += 2 * TILE_X) ts.h.push_back(Tile{lev, x, m, n, n + 1, 0});
}

struct BlockInfo {
  h3dmg_block spec;
  int L;
  int per;               // periodic axes bitmask
  std::vector<int> desc; // descriptor index of level g (1-based, desc[0] unused)
  double fnorm = 0.0;    // ||f||_2
  int cycles = 0;
  double rel = 0.0;
  double rmax = 0.0;
  double tol = 0.0;      // REL_L2 tolerance of this block (0: the handle's)
  bool enabled = true;   // disabled blocks keep their state and are skipped
};

}  // namespace

struct h3dmg_handle {
  h3dmg_params prm{};
  int device = 0;  // the HIP device the handle's memory lives on
  // blocks joined by shared faces (H3DMG_BC_NEIGHBOUR): one domain, every
  // kernel on its general path (ghosts through the neighbour table)
  bool connected = false;
  int nb = 0;
  int Lmax = 0;
  std::vector<BlockInfo> blk;

  hipStream_t stream = nullptr;
  bool own_stream = false;

  char *arena = nullptr;
  size_t arena_bytes = 0;

  std::vector<LevelDev> levs;  // level-major: all blocks' level 1 first
  std::vector<int> lev_is64;   // storage type of each descriptor
  LevelDev *d_levs = nullptr;
  std::vector<FineDev> fine;
  FineDev *d_fine = nullptr;
  std::vector<int> active;
  int *d_active = nullptr;

  // tile tables, indexed by level g (1..Lmax)
  std::vector<TileSet> relax_nc, relax_c;   // colour space; L_b > g / L_b == g
  std::vector<TileSet> restr_nc, restr_c;   // coarse natural at g; L_b > g / == g
  std::vector<TileSet> prol_nc, prol_c;     // fine natural at g; L_b > g+1 / == g+1
  std::vector<TileSet> nat_nc, nat_c;       // natural at g (zero fill)
  std::vector<TileSet> coarse_list;         // blocks whose coarsest level is g
  TileSet fine_tiles;                       // residual tiles of level 1, lev = block
  TileSet update_tiles;                     // p += pp tiles, lev = block
  std::vector<int> tile_begin;              // per block range in fine_tiles
  int *d_tile_begin = nullptr;
  RedPartial *d_partials = nullptr;
  RedPartial *d_stats = nullptr;
  RedPartial *h_stats = nullptr;  // pinned
  RedPartial *d_proj = nullptr;   // calmas partial maxima, block b at proj_begin[b]
  std::vector<int> proj_begin;    // per block range in d_proj (nb + 1 entries)
  int *d_proj_begin = nullptr;

  // velocity-form outer loop (h3dmg_velocity_begin / _step): per block the
  // device arrays u*, v*, w* and u, v, w, and dt*alfapr
  struct VelBlock {
    const double *us = nullptr, *vs = nullptr, *ws = nullptr;
    double *u = nullptr, *v = nullptr, *w = nullptr;
  };
  std::vector<VelBlock> vel;
  // handle-owned velocity arrays (h3dmg_set_velocity from host arrays):
  // per block u*, v*, w*, u, v, w, dense, in one allocation
  double *vel_store = nullptr;
  // 9 offsets (doubles) per block: u*, v*, w*, u, v, w, and per axis the two
  // outer face planes (faces -1 and n+1: the second ghost faces of calmas'
  // 4th-order branch where the axis ends in a wall), west plane first
  std::vector<size_t> vel_off;
  std::vector<char> vel_set;     // block's u*, v*, w* uploaded
  std::vector<char> vel_outer;   // block's outer face planes uploaded
  // calmas' branch (the reference's `differencing`): 1 second order
  // (press.for:48-63), 2 fourth order (press.for:17-37)
  int differencing = 1;
  double vel_dta = 0.0, vel_dt = 0.0;  // dt*alfapr and dt of the captured step
  bool vel_ready = false;
  hipGraph_t vgraph = nullptr;
  hipGraphExec_t vgexec = nullptr;

  // exact coarse solve
  double **d_eig_q = nullptr, **d_eig_l = nullptr, **d_scratch = nullptr;

  double *d_stage = nullptr;  // dense FP64 staging (largest padded level 1)
  size_t stage_elems = 0;

  std::vector<int> march_kb;  // planes per smoother workgroup per level (0 = auto)
  long fused_max = 0;  // levels whose blocks have at most this many cells: fused sweeps

  // variable-density operator (h3dmg_set_density): per level descriptor 14
  // coefficient arrays and dens_mg in one allocation
  bool lsm = false;
  double *lsm_store = nullptr;
  std::vector<size_t> lsm_dens;  // per descriptor: offset (doubles) of its dens_mg
  std::vector<char> lsm_set;     // block's density uploaded
  int tail_req = 0;           // requested tail level (0 = auto)
  int tail_lt = 1 << 30;      // levels >= tail_lt run in the single-workgroup tail kernel
  TileSet tail_tiles;         // one tile per block with L_b >= tail_lt
  size_t tail_lds = 0;        // dynamic LDS of the tail launch
  size_t lds_budget = 0;      // dynamic LDS a workgroup may use on this device

  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // h3dmg_solve's queued-ahead loop (made on its first graph-mode solve):
  // the cycle graph with the device-side stopping rule, the rule's state
  // (DecideBlock[nb], DecideState[nb], DecideCtl), the two-slot ring in
  // pinned host memory it writes each cycle's outcome to, and an event per
  // slot
  hipGraph_t graph_dec = nullptr;
  hipGraphExec_t gexec_dec = nullptr;
  char *d_dec = nullptr;
  char *h_ring = nullptr, *d_ring = nullptr;
  hipEvent_t ev_dec[2] = {nullptr, nullptr};
  bool rhs_ready = false;
  int pp_buf = 0;  // level-1 buffer that holds pp after the last V-cycle
  unsigned long long bytes = 0;

  // a handle built in a pool (h3dmg_create_pooled): the block regions are
  // the caller's pieces of the pool, the global region and the smoother
  // tables pieces the handle takes and gives back; nothing is allocated
  h3dmg_pool *pool = nullptr;
  std::vector<size_t> blk_off;  // pool offset of each block's region
  size_t glob_off = 0;
  bool glob_taken = false;
  int hs_slot = -1;             // first of the handle's entries in the pool's pinned stats
  // stepped solves (h3dmg_solve_begin / _step): per block running flag and
  // outcome
  std::vector<char> running;
  std::vector<int> outcome;
};

constexpr size_t POOL_STAGE_BYTES = (size_t)32 << 20;

// One device allocation carved into pieces (first fit, 256-byte aligned),
// and a pinned array of per-block statistics handed out in runs: the memory
// of handles whose blocks join and leave between V-cycles.
struct h3dmg_pool {
  int device = 0;
  char *base = nullptr;
  size_t bytes = 0;
  std::map<size_t, size_t> free_;  // offset -> size
  std::map<size_t, size_t> used_;  // offset -> size
  RedPartial *h_stats = nullptr;   // pinned, max_blocks entries
  std::vector<char> hs_used;
  // pinned staging of the tables a pooled handle uploads (builds hold the
  // device mutex, so one at a time; reset once the build's copies are done)
  char *h_stage = nullptr;
  size_t stage_bytes = 0, stage_used = 0;
  int live = 0;  // handles built in the pool and not yet destroyed
  std::mutex mu;
};

namespace {

bool is_mixed(const h3dmg_handle *h) { return h->prm.precision == H3DMG_MIXED; }

// ---- pool pieces ----
h3dmg_status pool_take(h3dmg_pool *pool, size_t bytes, size_t *off) {
  std::lock_guard<std::mutex> lock(pool->mu);
  const size_t need = align_up(std::max<size_t>(bytes, 1), 256);
  for (auto it = pool->free_.begin(); it != pool->free_.end(); ++it) {
    if (it->second < need) continue;
    const size_t o = it->first, rest = it->second - need;
    pool->free_.erase(it);
    if (rest) pool->free_[o + need] = rest;
    pool->used_[o] = need;
    *off = o;
    return H3DMG_OK;
  }
  return fail(H3DMG_ERR_NOMEM, "pool: no free piece of %zu bytes", need);
}

h3dmg_status pool_give(h3dmg_pool *pool, size_t off) {
  std::lock_guard<std::mutex> lock(pool->mu);
  auto u = pool->used_.find(off);
  if (u == pool->used_.end()) return fail(H3DMG_ERR_ARG, "pool: no piece at offset %zu", off);
  size_t o = off, n = u->second;
  pool->used_.erase(u);
  auto nx = pool->free_.lower_bound(o);
  if (nx != pool->free_.end() && nx->first == o + n) {
    n += nx->second;
    nx = pool->free_.erase(nx);
  }
  if (nx != pool->free_.begin()) {
    auto pv = std::prev(nx);
    if (pv->first + pv->second == o) {
      o = pv->first;
      n += pv->second;
      pool->free_.erase(pv);
    }
  }
  pool->free_[o] = n;
  return H3DMG_OK;
}

// bytes of the piece that starts at `off` (0: none does)
size_t pool_size_at(h3dmg_pool *pool, size_t off) {
  std::lock_guard<std::mutex> lock(pool->mu);
  auto u = pool->used_.find(off);
  return u == pool->used_.end() ? 0 : u->second;
}

// Host-to-device copy on the handle's stream.  A pooled handle copies
// through the pool's pinned staging (a true asynchronous DMA, no pageable
// bounce); when the staging is full it waits for the copies in flight and
// starts over.  The caller holds the device mutex.
hipError_t h2d(h3dmg_handle *h, void *dst, const void *src, size_t n) {
  h3dmg_pool *pool = h->pool;
  if (!pool || !pool->h_stage || n > pool->stage_bytes)
    return hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, h->stream);
  if (pool->stage_used + n > pool->stage_bytes) {
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return e;
    pool->stage_used = 0;
  }
  char *p = pool->h_stage + pool->stage_used;
  std::memcpy(p, src, n);
  pool->stage_used = align_up(pool->stage_used + n, 256);
  return hipMemcpyAsync(dst, p, n, hipMemcpyHostToDevice, h->stream);
}

// the handle's stream has drained: its staged copies are done
hipError_t drain(h3dmg_handle *h) {
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess && h->pool) h->pool->stage_used = 0;
  return e;
}

// device memory of a handle's separately held tables: the pool's or its own
h3dmg_status dev_alloc(h3dmg_handle *h, size_t bytes, void **p) {
  if (h->pool) {
    size_t off = 0;
    h3dmg_status st = pool_take(h->pool, bytes, &off);
    if (st) return st;
    *p = h->pool->base + off;
    return H3DMG_OK;
  }
  HIP_TRY(hipMalloc(p, bytes));
  return H3DMG_OK;
}

void dev_free(h3dmg_handle *h, void *p) {
  if (!p) return;
  if (h->pool)
    (void)pool_give(h->pool, (size_t)((char *)p - h->pool->base));
  else
    (void)hipFree(p);
}

// red_done: the red half of the first sweep (from zero) was written by the
// restriction into this level (launch_restrict's zbuf)
template <class T>
void relax_group(h3dmg_handle *h, const TileSet &ts, int nsweep, int &buf, bool &zero,
                 bool red_done = false) {
  for (int s = 0; s < nsweep; ++s) {
    if (h->prm.smoother == H3DMG_SMOOTH_GS)  // tile wavefront (every level FP64)
      for (int d = 0; d + 1 < (int)ts.diag_begin.size(); ++d)
        launch_gs_tiles(h->stream, ts.d + ts.diag_begin[d], ts.diag_begin[d + 1] - ts.diag_begin[d],
                        h->d_levs, h->d_active, buf, zero ? 1 : 0, ts.gs_tile);
    else if (h->prm.smoother == H3DMG_SMOOTH_TDMA && ts.tdma_lds)  // small levels
      launch_tdma_sweep(h->stream, ts.d, ts.n(), h->d_levs, h->d_active, buf, zero ? 1 : 0,
                        ts.tdma_lds, ts.tdma_pipe);
    else if (h->prm.smoother == H3DMG_SMOOTH_TDMA)  // j-, i-, k-line diagonals
      for (int pass = 0; pass < 3; ++pass)
        for (int d = 0; d < ts.ndiag[pass]; ++d)
          launch_tdma_diag(h->stream, ts.d + ts.pass_begin[pass],
                           ts.pass_begin[pass + 1] - ts.pass_begin[pass], h->d_levs, h->d_active,
                           pass, d, buf, zero ? 1 : 0);
    else if (h->lsm)  // per-cell coefficients (every level FP64)
      for (int colour = 0; colour < 2; ++colour)
        launch_lsm_sweep(h->stream, ts.nat->d, ts.nat->n(), h->d_levs, h->d_active, colour, buf,
                         zero ? 1 : 0);
    else if (ts.fused)
      launch_rb_fused<T>(h->stream, ts.d, ts.n(), h->d_levs, h->d_active, buf, zero ? 1 : 0);
    else
      for (int colour = (s == 0 && red_done) ? 1 : 0; colour < 2; ++colour)
        launch_rb_sweep<T>(h->stream, ts.d, ts.n(), ts.depth, h->d_levs, h->d_active, colour,
                           buf, zero ? 1 : 0, ts.one);
    buf ^= 1;
    zero = false;
  }
}

// (Re)build and upload the smoother tiles of level g for both block groups.
h3dmg_status upload_relax(h3dmg_handle *h, int g, bool sync = true);
h3dmg_status gen_relax(h3dmg_handle *h, int g);
// (Re)choose the tail level and upload its tile list.
h3dmg_status upload_tail(h3dmg_handle *h, bool sync = true);
h3dmg_status gen_tail(h3dmg_handle *h);

void relax_any(h3dmg_handle *h, bool is64, const TileSet &ts, int nsweep, int &buf, bool &zero,
               bool red_done = false) {
  if (ts.n() == 0) {
    // keep the bookkeeping identical for empty groups
    if (nsweep > 0) {
      buf ^= (nsweep & 1);
      zero = false;
    }
    return;
  }
  if (is64)
    relax_group<double>(h, ts, nsweep, buf, zero, red_done);
  else
    relax_group<float>(h, ts, nsweep, buf, zero, red_done);
}

void fill_any(h3dmg_handle *h, bool is64, const TileSet &ts, int buf) {
  if (is64)
    launch_fill_zero<double>(h->stream, ts.d, ts.n(), h->d_levs, h->d_active, buf);
  else
    launch_fill_zero<float>(h->stream, ts.d, ts.n(), h->d_levs, h->d_active, buf);
}

// The residual's fast path (enqueue_update_residual) also writes the red half
// of the level-1 relaxation from zero (each red cell rhs/ap, into buffer 1,
// where the first half-sweep from buffer 0 puts it): an FP64 level 1 (in
// mixed precision the FP32 residual slows by more than the half-sweep costs),
// the red-black smoother, constant coefficients, no neighbours, nu1 > 0, and
// nu1 + nu2 even (the correction pp then lies in buffer 0 and stays
// untouched).
bool red1_write(const h3dmg_handle *h) {
  const h3dmg_params &P = h->prm;
  return !is_mixed(h) && P.smoother == H3DMG_SMOOTH_RBGS && !h->lsm && !h->connected &&
         P.nu1 > 0 && ((P.nu1 + P.nu2) & 1) == 0 && h->fine_tiles.depth > 0;
}

// Enqueue one V-cycle (mgkcyc, mgsolver.for:272-352) on the level-1 rhs that
// is already in place.  red1: that rhs came from the residual with the red
// half of the first level-1 sweep (red1_write), so the sweep starts with its
// black half.  Per level the schedule is relax(g, nu1) on the way
// down and relax(g, nu2) after mgcorr on the way up; at a block's coarsest
// level relax(L, nu1) + relax(L, nu2) (the reference) or the exact solve.
// Sweep counts per level are nu1 + nu2 for every block, so the buffer that
// holds a level at mgcorr time is the same for all blocks.
void enqueue_vcycle(h3dmg_handle *h, bool red1 = false) {
  const int nu1 = h->prm.nu1, nu2 = h->prm.nu2, Lm = h->Lmax;
  const bool mixed = is_mixed(h);
  const bool exact = h->prm.coarse == H3DMG_COARSE_EXACT;
  std::vector<int> buf_nc(Lm + 2, 0), buf_c(Lm + 2, 0);
  std::vector<char> z_nc(Lm + 2, 1), z_c(Lm + 2, 1);
  auto nc_is64 = [&](int) { return !mixed; };

  // levels >= Lt of the blocks that reach them run in one tail launch
  const int Lt = h->tail_tiles.n() ? h->tail_lt : Lm + 1;
  const int gdown = std::min(Lm, Lt);
  for (int g = 1; g <= gdown; ++g) {
    // the restriction into g writes the red half of g's first pre-smoothing
    // sweep (from zero) when that sweep is a split red-black one
    // (for level 1: the residual wrote it, red1)
    const bool red_by_restr =
        (g >= 2 ? h->prm.smoother == H3DMG_SMOOTH_RBGS && !h->lsm && !h->connected &&
                      h->restr_nc[g].n() && h->restr_nc[g].depth > 0
                : red1 && red1_write(h)) &&
        g < Lt && g < Lm && h->prm.nu1 > 0 && h->relax_nc[g].n() && !h->relax_nc[g].fused &&
        z_nc[g];
    // likewise for the blocks whose coarsest level g is smoothed (not solved)
    const bool red_c_by_restr = g >= 2 && g < Lt && !(exact && g > 1) && h->prm.nu1 > 0 &&
                                h->prm.smoother == H3DMG_SMOOTH_RBGS && !h->lsm &&
                                !h->connected && h->restr_c[g].n() &&
                                h->restr_c[g].depth > 0 && h->relax_c[g].n() &&
                                !h->relax_c[g].fused;
    if (g >= 2) {
      // mgrestr(g): residual of level g-1 (non-coarsest there) into rhs_g
      const bool f64 = nc_is64(g - 1);
      if (z_nc[g - 1]) {
        fill_any(h, f64, h->nat_nc[g - 1], buf_nc[g - 1]);
        z_nc[g - 1] = 0;
      }
      if (h->lsm) {
        for (const TileSet *ts : {&h->restr_nc[g], &h->restr_c[g]})
          launch_lsm_restrict(h->stream, ts->d, ts->n(), h->d_levs, h->d_active, buf_nc[g - 1]);
      } else if (h->restr_nc[g].n()) {
        const int zbuf = red_by_restr ? 1 - buf_nc[g] : -1;
        if (f64)
          launch_restrict<double, double>(h->stream, h->restr_nc[g].d, h->restr_nc[g].n(),
                                          h->restr_nc[g].depth, h->d_levs, h->d_active,
                                          buf_nc[g - 1], h->restr_nc[g].one, zbuf);
        else
          launch_restrict<float, float>(h->stream, h->restr_nc[g].d, h->restr_nc[g].n(),
                                        h->restr_nc[g].depth, h->d_levs, h->d_active,
                                        buf_nc[g - 1], h->restr_nc[g].one, zbuf);
      }
      if (!h->lsm && h->restr_c[g].n()) {
        const int zbuf = red_c_by_restr ? 1 : -1;  // the coarsest level's sweeps start in buffer 0
        if (f64)
          launch_restrict<double, double>(h->stream, h->restr_c[g].d, h->restr_c[g].n(),
                                          h->restr_c[g].depth, h->d_levs, h->d_active,
                                          buf_nc[g - 1], h->restr_c[g].one, zbuf);
        else
          launch_restrict<float, double>(h->stream, h->restr_c[g].d, h->restr_c[g].n(),
                                         h->restr_c[g].depth, h->d_levs, h->d_active,
                                         buf_nc[g - 1], h->restr_c[g].one, zbuf);
      }
    }
    if (g == Lt) {
      launch_vtail(h->stream, h->tail_tiles.d, h->tail_tiles.n(), h->d_levs, h->d_active,
                   (const double *const *)h->d_eig_q, (const double *const *)h->d_eig_l, nu1,
                   nu2, exact ? 1 : 0, mixed ? 1 : 0, h->tail_lds);
      buf_nc[g] = buf_c[g] = (nu1 + nu2) & 1;
      z_nc[g] = z_c[g] = 0;
      break;
    }
    // blocks whose coarsest level is g
    if (h->relax_c[g].n()) {
      bool z = true;
      int b = 0;
      if (exact && g > 1) {
        const int out = (nu1 + nu2) & 1;
        launch_coarse_exact(h->stream, h->coarse_list[g].d, h->coarse_list[g].n(), h->d_levs,
                            h->d_active, (const double *const *)h->d_eig_q,
                            (const double *const *)h->d_eig_l, h->d_scratch, out);
        b = out;
        z = false;
      } else {
        relax_any(h, true, h->relax_c[g], nu1, b, z, red_c_by_restr);
        relax_any(h, true, h->relax_c[g], nu2, b, z);
        if (z) {  // nu1 = nu2 = 0
          fill_any(h, true, h->nat_c[g], b);
          z = false;
        }
      }
      buf_c[g] = b;
      z_c[g] = 0;
    }
    // pre-smoothing of the blocks that go on coarsening
    if (g < Lm && h->relax_nc[g].n()) {
      bool z = z_nc[g];
      relax_any(h, nc_is64(g), h->relax_nc[g], nu1, buf_nc[g], z, red_by_restr);
      z_nc[g] = z;
    }
  }
  for (int g = gdown - 1; g >= 1; --g) {
    // mgcorr(g) for blocks with L_b > g
    if (h->prol_nc[g].n() || h->prol_c[g].n()) {
      const bool f64 = nc_is64(g);
      if (z_nc[g]) {
        fill_any(h, f64, h->nat_nc[g], buf_nc[g]);
        z_nc[g] = 0;
      }
      if (h->prol_nc[g].n()) {
        if (z_nc[g + 1]) {
          fill_any(h, nc_is64(g + 1), h->nat_nc[g + 1], buf_nc[g + 1]);
          z_nc[g + 1] = 0;
        }
        if (f64)
          launch_prolong<double, double>(h->stream, h->prol_nc[g].d, h->prol_nc[g].n(), h->d_levs,
                                         h->d_active, buf_nc[g], buf_nc[g + 1], h->connected,
                                         h->prol_nc[g].one);
        else
          launch_prolong<float, float>(h->stream, h->prol_nc[g].d, h->prol_nc[g].n(), h->d_levs,
                                       h->d_active, buf_nc[g], buf_nc[g + 1], h->connected,
                                       h->prol_nc[g].one);
      }
      if (h->prol_c[g].n()) {
        if (f64)
          launch_prolong<double, double>(h->stream, h->prol_c[g].d, h->prol_c[g].n(), h->d_levs,
                                         h->d_active, buf_nc[g], buf_c[g + 1], h->connected,
                                         h->prol_c[g].one);
        else
          launch_prolong<float, double>(h->stream, h->prol_c[g].d, h->prol_c[g].n(), h->d_levs,
                                        h->d_active, buf_nc[g], buf_c[g + 1], h->connected,
                                        h->prol_c[g].one);
      }
      bool z = false;
      relax_any(h, f64, h->relax_nc[g], nu2, buf_nc[g], z);
    }
  }
  // the level-1 correction pp: blocks with L_b == 1 hold it in buf_c[1]
  h->pp_buf = (h->relax_c[1].n() && !h->relax_nc[1].n()) ? buf_c[1] : buf_nc[1];
}

// p += pp, r = f - A p, per-block ||r||, max|r| (one outer iteration tail)
// finalize = false: the caller reduces the partials itself
void enqueue_update_residual(h3dmg_handle *h, bool update, bool finalize = true) {
  const bool f64 = !is_mixed(h);
  const TileSet &ft = h->fine_tiles, &ut = h->update_tiles;
  if (update) {
    if (f64)
      launch_update<double>(h->stream, ut.d, ut.n(), h->d_fine, h->d_levs, h->d_active, h->pp_buf);
    else
      launch_update<float>(h->stream, ut.d, ut.n(), h->d_fine, h->d_levs, h->d_active, h->pp_buf);
  }
  if (h->lsm)
    launch_lsm_residual(h->stream, ft.d, ft.n(), h->d_fine, h->d_levs, h->d_active,
                        h->d_partials);
  else if (f64)
    launch_residual<double>(h->stream, ft.d, ft.n(), ft.depth, h->d_fine, h->d_levs, h->d_active,
                            h->d_partials, 1, red1_write(h) ? 1 : -1);
  else
    launch_residual<float>(h->stream, ft.d, ft.n(), ft.depth, h->d_fine, h->d_levs, h->d_active,
                           h->d_partials, 1, red1_write(h) ? 1 : -1);
  if (finalize) launch_finalize(h->stream, h->d_tile_begin, h->nb, h->d_partials, h->d_stats);
}

h3dmg_status read_stats(h3dmg_handle *h) {
  HIP_TRY(hipMemcpyAsync(h->h_stats, h->d_stats, sizeof(RedPartial) * h->nb,
                         hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return H3DMG_OK;
}

// a variable-density handle needs the density of every enabled block
h3dmg_status check_density(const h3dmg_handle *h) {
  if (!h->lsm) return H3DMG_OK;
  for (int b = 0; b < h->nb; ++b)
    if (h->blk[b].enabled && !h->lsm_set[b])
      return fail(H3DMG_ERR_ARG, "block %d: no density (h3dmg_set_density)", b);
  return H3DMG_OK;
}

// every enabled block active
void reset_active(h3dmg_handle *h) {
  for (int b = 0; b < h->nb; ++b) h->active[b] = h->blk[b].enabled ? 1 : 0;
}

h3dmg_status push_active(h3dmg_handle *h) {
  HIP_TRY(hipMemcpyAsync(h->d_active, h->active.data(), sizeof(int) * h->nb,
                         hipMemcpyHostToDevice, h->stream));
  return H3DMG_OK;
}

// Closed-form eigenpairs of the 1-D operator c*(second difference) with
// Neumann (mirror ghost, zero wall coefficient) or periodic ends, index 0 =
// the constant mode with eigenvalue exactly 0.  Q is row-major Q[x*n + m].
void eig1d(int n, bool periodic, double c, std::vector<double> &Q, std::vector<double> &lam) {
  Q.assign((size_t)n * n, 0.0);
  lam.assign(n, 0.0);
  const double pi = 3.14159265358979323846;
  if (!periodic) {
    for (int m = 0; m < n; ++m) {
      const double s = std::sin(pi * m / (2.0 * n));
      lam[m] = m == 0 ? 0.0 : -4.0 * c * s * s;
      const double nrm = m == 0 ? std::sqrt(1.0 / n) : std::sqrt(2.0 / n);
      for (int x = 0; x < n; ++x) Q[(size_t)x * n + m] = nrm * std::cos(pi * m * (x + 0.5) / n);
    }
  } else {
    int col = 0;
    for (int x = 0; x < n; ++x) Q[(size_t)x * n + 0] = std::sqrt(1.0 / n);
    lam[0] = 0.0;
    col = 1;
    for (int m = 1; 2 * m < n; ++m) {
      const double s = std::sin(pi * m / n);
      const double l = -4.0 * c * s * s;
      for (int x = 0; x < n; ++x) {
        Q[(size_t)x * n + col] = std::sqrt(2.0 / n) * std::cos(2.0 * pi * m * x / n);
        Q[(size_t)x * n + col + 1] = std::sqrt(2.0 / n) * std::sin(2.0 * pi * m * x / n);
      }
      lam[col] = l;
      lam[col + 1] = l;
      col += 2;
    }
    if (n % 2 == 0) {
      for (int x = 0; x < n; ++x) Q[(size_t)x * n + col] = (x % 2 ? -1.0 : 1.0) * std::sqrt(1.0 / n);
      lam[col] = -4.0 * c;
      col += 1;
    }
  }
}

// Host mirror of the device's ghost rule (resolve_cell): the block and
// interior cell whose value the ghost at padded position c of block b holds
// after exchangepp, with the axes resolved in the order `ord` (the device
// and the reference use x, y, z).  Walls mirror first; then each axis moves
// into the block across the face (H3DMG_BC_NEIGHBOUR) or wraps (periodic).
int host_resolve(const h3dmg_block *blocks, int b, int c[3], const int ord[3]) {
  auto n_of = [&](int q, int a) { return a == 0 ? blocks[q].ni : a == 1 ? blocks[q].nj : blocks[q].nk; };
  for (int a = 0; a < 3; ++a) {
    if (c[a] < 0 && blocks[b].bc[2 * a] == H3DMG_BC_WALL) c[a] = 0;
    if (c[a] >= n_of(b, a) && blocks[b].bc[2 * a + 1] == H3DMG_BC_WALL) c[a] = n_of(b, a) - 1;
  }
  for (int q = 0; q < 3; ++q) {
    const int a = ord[q];
    const int n = n_of(b, a);
    if (c[a] < 0) {
      const int k = blocks[b].bc[2 * a];
      if (k == H3DMG_BC_NEIGHBOUR) {
        b = blocks[b].nbr[2 * a];
        c[a] += n_of(b, a);
      } else {
        c[a] = k == H3DMG_BC_PERIODIC ? c[a] + n : 0;
      }
    } else if (c[a] >= n) {
      const int k = blocks[b].bc[2 * a + 1];
      if (k == H3DMG_BC_NEIGHBOUR) {
        b = blocks[b].nbr[2 * a + 1];
        c[a] -= n;
      } else {
        c[a] = k == H3DMG_BC_PERIODIC ? c[a] - n : n - 1;
      }
    }
  }
  return b;
}

// Blocks joined by shared faces (H3DMG_BC_NEIGHBOUR): each shared face names
// its partner on the opposite face of the partner, with the same extent,
// spacing and level count; and the blocks form a box -- for every block and
// every edge or corner ghost, resolving the axes in any order reaches the
// same cell, which holds exactly where each edge or corner between shared
// faces has its diagonal block (the reference's ghosts there are stale
// otherwise: exchangepp.for:44-1105 copies whole faces in the order x, y, z).
h3dmg_status validate_neighbours(const h3dmg_block *blocks, int nblocks, const h3dmg_params &P) {
  bool any = false;
  for (int b = 0; b < nblocks; ++b)
    for (int f = 0; f < 6; ++f) {
      if (blocks[b].bc[f] != H3DMG_BC_NEIGHBOUR) continue;
      any = true;
      const int a = f >> 1, o = f ^ 1;
      const int q = blocks[b].nbr[f];
      if (q < 0 || q >= nblocks)
        return fail(H3DMG_ERR_ARG, "block %d face %d: neighbour %d out of range", b, f, q);
      if (blocks[q].bc[o] != H3DMG_BC_NEIGHBOUR || blocks[q].nbr[o] != b)
        return fail(H3DMG_ERR_ARG, "block %d face %d: block %d does not name it back on face %d",
                    b, f, q, o);
      if (blocks[b].bc[f ^ 1] == H3DMG_BC_PERIODIC)
        return fail(H3DMG_ERR_ARG, "block %d: axis %d both periodic onto itself and shared", b, a);
      const int nb[3] = {blocks[b].ni, blocks[b].nj, blocks[b].nk};
      const int nq[3] = {blocks[q].ni, blocks[q].nj, blocks[q].nk};
      for (int e = 0; e < 3; ++e)
        if (e != a && nb[e] != nq[e])
          return fail(H3DMG_ERR_UNSUPPORTED,
                      "blocks %d and %d share face %d with different extents along axis %d", b, q,
                      f, e);
      if (blocks[q].dx != blocks[b].dx || blocks[q].dy != blocks[b].dy ||
          blocks[q].dz != blocks[b].dz || blocks[q].nlevels != blocks[b].nlevels)
        return fail(H3DMG_ERR_UNSUPPORTED,
                    "blocks %d and %d: neighbours with different spacing or level count (LMR "
                    "refinement)", b, q);
    }
  if (!any) return H3DMG_OK;
  if (P.smoother != H3DMG_SMOOTH_RBGS)
    return fail(H3DMG_ERR_UNSUPPORTED, "blocks with neighbours: red-black smoother only");
  if (P.coarse != H3DMG_COARSE_SMOOTH)
    return fail(H3DMG_ERR_UNSUPPORTED,
                "blocks with neighbours: the coarsest level is smoothed (H3DMG_COARSE_SMOOTH)");
  static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  for (int b = 0; b < nblocks; ++b) {
    const int n[3] = {blocks[b].ni, blocks[b].nj, blocks[b].nk};
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int d[3] = {dx, dy, dz};
          if ((dx != 0) + (dy != 0) + (dz != 0) < 2) continue;
          int first_b = -1, first[3] = {0, 0, 0};
          for (const auto &ord : perms) {
            int c[3];
            for (int a = 0; a < 3; ++a) c[a] = d[a] < 0 ? -1 : d[a] > 0 ? n[a] : 0;
            const int r = host_resolve(blocks, b, c, ord);
            if (first_b < 0) {
              first_b = r;
              for (int a = 0; a < 3; ++a) first[a] = c[a];
            } else if (r != first_b || c[0] != first[0] || c[1] != first[1] || c[2] != first[2]) {
              return fail(H3DMG_ERR_UNSUPPORTED,
                          "block %d: the edge/corner ghost towards (%d,%d,%d) has no diagonal "
                          "block (the blocks do not form a box)",
                          b, dx, dy, dz);
            }
          }
        }
  }
  return H3DMG_OK;
}

// Host-side checks of the block specifications, run before any device call:
// the level rule of localparameters.for:202-208 (every axis divisible by
// 2^(ngrid-1)), sizes, spacings and periodic pairs.
h3dmg_status validate_blocks(const h3dmg_block *blocks, int nblocks, const h3dmg_params &P) {
  for (int b = 0; b < nblocks; ++b) {
    const int n[3] = {blocks[b].ni, blocks[b].nj, blocks[b].nk};
    const int L = blocks[b].nlevels;
    if (L < 1 || L > 16) return fail(H3DMG_ERR_ARG, "block %d: nlevels %d out of range", b, L);
    for (int a = 0; a < 3; ++a) {
      if (n[a] < 1) return fail(H3DMG_ERR_ARG, "block %d: n[%d] = %d", b, a, n[a]);
      if (n[a] % (1 << (L - 1)) != 0)
        return fail(H3DMG_ERR_LEVELS,
                    "block %d: %d cells along axis %d not divisible by 2^(nlevels-1) = %d", b,
                    n[a], a, 1 << (L - 1));
    }
    for (int f = 0; f < 6; ++f) {
      const int k = blocks[b].bc[f];
      if (k != H3DMG_BC_WALL && k != H3DMG_BC_PERIODIC && k != H3DMG_BC_NEIGHBOUR)
        return fail(H3DMG_ERR_ARG, "block %d: face %d has boundary kind %d", b, f, k);
    }
    for (int a = 0; a < 3; ++a) {
      const int lo = blocks[b].bc[2 * a], hi = blocks[b].bc[2 * a + 1];
      if ((lo == H3DMG_BC_PERIODIC) != (hi == H3DMG_BC_PERIODIC))
        return fail(H3DMG_ERR_ARG, "block %d: axis %d periodic on one face only", b, a);
    }
    if (!(blocks[b].dx > 0 && blocks[b].dy > 0 && blocks[b].dz > 0))
      return fail(H3DMG_ERR_ARG, "block %d: non-positive spacing", b);
    if (P.precision == H3DMG_MIXED && L < 2)
      return fail(H3DMG_ERR_ARG, "block %d: mixed precision needs nlevels >= 2", b);
    if (P.smoother == H3DMG_SMOOTH_TDMA &&
        (n[0] > TDMA_MAX_LEN || n[1] > TDMA_MAX_LEN || n[2] > TDMA_MAX_LEN))
      return fail(H3DMG_ERR_UNSUPPORTED, "line TDMA: block %d has lines over %d cells", b,
                  TDMA_MAX_LEN);
  }
  return validate_neighbours(blocks, nblocks, P);
}

// Lays out and uploads a handle over `blocks`.  `fresh` (pooled handles):
// blocks whose region starts new (zero p and f, the exact solve's tables
// uploaded); the others carry theirs over.  With `region_bytes` set, only
// the per-block region sizes are computed, and nothing is allocated.
h3dmg_status build(h3dmg_handle *h, const h3dmg_block *blocks, int nblocks,
                   const int *fresh = nullptr, size_t *region_bytes = nullptr) {
  h->nb = nblocks;
  h->blk.resize(nblocks);
  h->Lmax = 0;
  for (int b = 0; b < nblocks; ++b) {
    BlockInfo &B = h->blk[b];
    B.spec = blocks[b];
    B.L = blocks[b].nlevels;
    B.per = 0;
    for (int a = 0; a < 3; ++a)
      if (blocks[b].bc[2 * a] == H3DMG_BC_PERIODIC) B.per |= 1 << a;
    for (int f = 0; f < 6; ++f)
      if (blocks[b].bc[f] == H3DMG_BC_NEIGHBOUR) h->connected = true;
    h->Lmax = std::max(h->Lmax, B.L);
  }
  const bool exact = h->prm.coarse == H3DMG_COARSE_EXACT;

  // ---- descriptors (level-major) and per-block regions ----
  // Everything one block owns (its levels, finest p and f, the exact coarse
  // solve's tables, the TDMA tables) lies in one contiguous region of the
  // block, at offsets boff[b] relative to the region; the region sizes are
  // h3dmg_block_bytes'.  The tables shared by the handle's blocks follow in
  // a global region (offsets `off`).
  std::vector<size_t> boff(nblocks, 0);
  struct Pending {
    int d, b;
    size_t phi[2][2], rhs[2];
  };
  std::vector<Pending> pend;
  for (int g = 1; g <= h->Lmax; ++g) {
    for (int b = 0; b < nblocks; ++b) {
      BlockInfo &B = h->blk[b];
      if (B.desc.empty()) B.desc.assign(B.L + 1, -1);
      if (g > B.L) continue;
      LevelDev L{};
      L.ni = B.spec.ni >> (g - 1);
      L.nj = B.spec.nj >> (g - 1);
      L.nk = B.spec.nk >> (g - 1);
      L.H = (((L.ni + 1) / 2) + 1) & ~1;  // even: aligned 2-entry accesses
      L.lpr = lanes_per_row(L.H);
      // restriction lanes: two coarse cells per lane from an FP64 fine level
      L.lpr1 = lanes_per_row(is_mixed(h) ? 2 * L.ni : L.ni);
      L.per = B.per;
      L.blk = b;
      const double s = (double)(1 << (g - 1));
      const double ndx = B.spec.dx * s, ndy = B.spec.dy * s, ndz = B.spec.dz * s;
      L.cx = 1.0 / (ndx * ndx);
      L.cy = 1.0 / (ndy * ndy);
      L.cz = 1.0 / (ndz * ndz);
      const bool is64 = !is_mixed(h) || g == B.L;
      const size_t w = is64 ? 8 : 4;
      const size_t half = (size_t)L.H * L.nj * L.nk * w;
      // kernels address a level's six arrays as 32-bit byte offsets from
      // its first one
      if (6 * align_up(half, 256) >= ((size_t)1 << 32))
        return fail(H3DMG_ERR_UNSUPPORTED,
                    "block %d level %d: %d x %d x %d cells exceed the 4 GiB level span", b, g,
                    L.ni, L.nj, L.nk);
      Pending P{};
      P.d = (int)h->levs.size();
      P.b = b;
      for (int bb = 0; bb < 2; ++bb)
        for (int c = 0; c < 2; ++c) {
          P.phi[bb][c] = boff[b];
          boff[b] = align_up(boff[b] + half, 256);
        }
      for (int c = 0; c < 2; ++c) {
        P.rhs[c] = boff[b];
        boff[b] = align_up(boff[b] + half, 256);
      }
      B.desc[g] = P.d;
      h->levs.push_back(L);
      h->lev_is64.push_back(is64 ? 1 : 0);
      pend.push_back(P);
    }
  }
  // level-1 descriptors must be indices 0..nb-1
  for (int b = 0; b < nblocks; ++b)
    if (h->blk[b].desc[1] != b) return fail(H3DMG_ERR_ARG, "internal: level-1 order");
  for (int b = 0; b < nblocks; ++b) {
    BlockInfo &B = h->blk[b];
    for (int g = 1; g <= B.L; ++g) {
      LevelDev &L = h->levs[B.desc[g]];
      L.finer = g > 1 ? B.desc[g - 1] : -1;
      L.coarser = g < B.L ? B.desc[g + 1] : -1;
      // neighbours share the level count, so level g exists on both sides
      for (int f = 0; f < 6; ++f)
        L.nbr[f] = B.spec.bc[f] == H3DMG_BC_NEIGHBOUR ? h->blk[B.spec.nbr[f]].desc[g] : -1;
    }
  }
  // finest-level FP64 p and f
  std::vector<size_t> foff(nblocks * 4);
  for (int b = 0; b < nblocks; ++b) {
    const LevelDev &L = h->levs[b];
    const size_t half = (size_t)L.H * L.nj * L.nk * 8;
    for (int q = 0; q < 4; ++q) {
      foff[4 * b + q] = boff[b];
      boff[b] = align_up(boff[b] + half, 256);
    }
  }
  // exact coarse solve tables
  std::vector<size_t> eoff;  // per block: Q0,Q1,Q2,l0,l1,l2,S0,S1
  std::vector<std::vector<double>> edata;
  if (exact) {
    eoff.resize((size_t)nblocks * 8);
    for (int b = 0; b < nblocks; ++b) {
      const LevelDev &L = h->levs[h->blk[b].desc[h->blk[b].L]];
      const int nn[3] = {L.ni, L.nj, L.nk};
      const double cc[3] = {L.cx, L.cy, L.cz};
      for (int a = 0; a < 3; ++a) {
        std::vector<double> Q, lam;
        eig1d(nn[a], (L.per >> a) & 1, cc[a], Q, lam);
        eoff[8 * b + a] = boff[b];
        boff[b] = align_up(boff[b] + Q.size() * 8, 256);
        edata.push_back(Q);
        eoff[8 * b + 3 + a] = boff[b];
        boff[b] = align_up(boff[b] + lam.size() * 8, 256);
        edata.push_back(lam);
      }
      const size_t n3 = (size_t)nn[0] * nn[1] * nn[2];
      eoff[8 * b + 6] = boff[b];
      boff[b] = align_up(boff[b] + n3 * 8, 256);
      eoff[8 * b + 7] = boff[b];
      boff[b] = align_up(boff[b] + n3 * 8, 256);
    }
  }
  // line-TDMA tables: per descriptor 32*(ni+nj+nk) doubles
  std::vector<size_t> tdma_off(h->levs.size(), 0);
  if (h->prm.smoother == H3DMG_SMOOTH_TDMA)
    for (size_t d = 0; d < h->levs.size(); ++d) {
      const LevelDev &L = h->levs[d];
      tdma_off[d] = boff[L.blk];
      boff[L.blk] = align_up(boff[L.blk] + (size_t)32 * (L.ni + L.nj + L.nk) * 8, 256);
    }
  if (region_bytes) {
    for (int b = 0; b < nblocks; ++b) region_bytes[b] = boff[b];
    return H3DMG_OK;
  }
  // staging: largest padded level-1 array (a pooled handle takes device
  // arrays only, and has none)
  size_t stage = 0;
  for (int b = 0; b < nblocks && !h->pool; ++b)
    stage = std::max(stage, (size_t)(h->blk[b].spec.ni + 2) * (h->blk[b].spec.nj + 2) *
                                (h->blk[b].spec.nk + 2));
  h->stage_elems = stage;
  size_t off = 0;
  const size_t stage_off = off;
  off = align_up(off + stage * 8, 256);

  // ---- tile tables ----
  const int Lm = h->Lmax;
  h->relax_nc.assign(Lm + 2, {});
  h->relax_c.assign(Lm + 2, {});
  h->restr_nc.assign(Lm + 2, {});
  h->restr_c.assign(Lm + 2, {});
  h->prol_nc.assign(Lm + 2, {});
  h->prol_c.assign(Lm + 2, {});
  h->nat_nc.assign(Lm + 2, {});
  h->nat_c.assign(Lm + 2, {});
  h->coarse_list.assign(Lm + 2, {});
  for (int b = 0; b < nblocks; ++b)
    h->coarse_list[h->blk[b].L].h.push_back(Tile{h->blk[b].desc[h->blk[b].L], b, 0, 0});
  for (int g = 1; g <= Lm; ++g)
    for (int b = 0; b < nblocks; ++b) {
      const BlockInfo &B = h->blk[b];
      if (g > B.L) continue;
      const int d = B.desc[g];
      const LevelDev &L = h->levs[d];
      if (g < B.L) {
        add_tiles(h->nat_nc[g], d, L.ni, L.nj, L.nk);
        add_tiles_prolong(g + 1 < B.L ? h->prol_nc[g] : h->prol_c[g], d, L);
      } else {
        add_tiles(h->nat_c[g], d, L.ni, L.nj, L.nk);
      }
    }
  // restriction tiles into level g >= 2, one depth per launch group
  for (int g = 2; g <= Lm; ++g)
    for (int grp = 0; grp < 2; ++grp) {
      std::vector<int> members;
      for (int b = 0; b < nblocks; ++b) {
        const BlockInfo &B = h->blk[b];
        if (g <= B.L && (grp == 1) == (g == B.L)) members.push_back(B.desc[g]);
      }
      if (members.empty()) continue;
      TileSet &ts = grp == 0 ? h->restr_nc[g] : h->restr_c[g];
      if (h->connected) {  // natural coarse tiles, the general kernel (depth 0)
        ts.depth = 0;
        for (int d : members) add_tiles(ts, d, h->levs[d].ni, h->levs[d].nj, h->levs[d].nk);
        continue;
      }
      const int cpl = is_mixed(h) ? 1 : 2;
      ts.depth = restrict_depth(h->levs, members, cpl);
      for (int d : members) add_tiles_restrict(ts, d, h->levs[d], ts.depth, cpl);
    }
  // finest level: residual tiles (two colour-array entries of both colours
  // per lane, lpr lanes per row segment, kb planes; kind 0 = y/z interior)
  // and p += pp tiles (1024 entries each)
  bool fine_general = false;  // a block one cell wide in x: general tiles
  // One depth for every block (one launch): ~2048 workgroups over all
  // blocks, at most 8 planes -- the tiles are software-pipelined (the loads
  // of plane k+1 are in flight while plane k is computed), so depth costs
  // little occupancy and the plane window is re-read (kb+2)/kb times.
  h->tile_begin.assign(nblocks + 1, 0);
  // a tile spans 2*lpr colour entries of TILE_Y*64/lpr rows (the residual
  // gives each lane two entries of both colours)
  auto fine_xy = [&](const LevelDev &L, int &xs, int &ys) {
    const int lpr = lanes_per_row(L.H);
    xs = 2 * lpr;
    ys = TILE_Y * (64 / lpr);
    return ((L.H + xs - 1) / xs) * ((L.nj + ys - 1) / ys);
  };
  {
    long cols = 0;
    std::vector<int> nks;
    for (int b = 0; b < nblocks; ++b) {
      int xs, ys;
      cols += (long)fine_xy(h->levs[b], xs, ys) * h->levs[b].nk;
      nks.push_back(h->levs[b].nk);
      fine_general = fine_general || h->levs[b].ni < 2 || h->connected;
    }
    h->fine_tiles.depth = launch_depth(cols, 2048, 0, 8, nks);
  }
  for (int b = 0; b < nblocks; ++b) {
    const LevelDev &L = h->levs[b];
    h->tile_begin[b] = h->fine_tiles.n();
    int xs, ys;
    fine_xy(L, xs, ys);
    const int kb = h->fine_tiles.depth;
    for (int z = 0; z < L.nk; z += kb) {
      const int z1 = z + kb;
      for (int y = 0; y < L.nj; y += ys)
        for (int x = 0; x < L.H; x += xs) {
          const bool inner = y >= 1 && y + ys <= L.nj - 1 && z >= 1 && z1 <= L.nk - 1;
          const int kind = (L.ni < 2 || h->connected) ? 1 : inner ? 0 : 3;
          h->fine_tiles.h.push_back(Tile{b, x, y, z, z1, kind});
        }
    }
    const long entries = (long)L.H * L.nj * L.nk;
    for (long e = 0; e < entries; e += 4 * NTHREADS)
      h->update_tiles.h.push_back(Tile{b, (int)e, 0, 0, 0, 0});
  }
  h->tile_begin[nblocks] = h->fine_tiles.n();
  if (fine_general) h->fine_tiles.depth = 0;  // every launch on the general path

  for (int g = 1; g <= Lm; ++g) {
    set_one(h->restr_nc[g], h->levs, -1);
    set_one(h->restr_c[g], h->levs, -1);
    set_one(h->prol_nc[g], h->levs, 1);
    set_one(h->prol_c[g], h->levs, 1);
  }
  // the smoother and tail tables (later re-tuning gives them pieces of
  // their own, place_owned)
  h->march_kb.assign(Lm + 1, 0);
  for (int g = 1; g <= Lm; ++g) {
    h3dmg_status st = gen_relax(h, g);
    if (st != H3DMG_OK) return st;
  }
  {
    // the tail kernel runs the red-black smoother on independent blocks only
    if (h->prm.smoother != H3DMG_SMOOTH_RBGS || h->connected) h->tail_req = Lm + 1;
    h3dmg_status st = gen_tail(h);
    if (st != H3DMG_OK) return st;
  }
  std::vector<TileSet *> all;
  for (int g = 1; g <= Lm; ++g)
    for (auto *v : {&h->restr_nc, &h->restr_c, &h->prol_nc, &h->prol_c, &h->nat_nc, &h->nat_c,
                    &h->coarse_list, &h->relax_nc, &h->relax_c})
      all.push_back(&(*v)[g]);
  all.push_back(&h->fine_tiles);
  all.push_back(&h->update_tiles);
  all.push_back(&h->tail_tiles);
  std::vector<size_t> toff(all.size());
  for (size_t q = 0; q < all.size(); ++q) {
    toff[q] = off;
    off = align_up(off + all[q]->h.size() * sizeof(Tile), 256);
  }
  const size_t levs_off = off;
  off = align_up(off + h->levs.size() * sizeof(LevelDev), 256);
  const size_t fine_off = off;
  off = align_up(off + nblocks * sizeof(FineDev), 256);
  const size_t act_off = off;
  off = align_up(off + nblocks * sizeof(int), 256);
  const size_t tb_off = off;
  off = align_up(off + (nblocks + 1) * sizeof(int), 256);
  const size_t part_off = off;
  off = align_up(off + std::max<size_t>(1, h->fine_tiles.h.size()) * sizeof(RedPartial), 256);
  const size_t stats_off = off;
  off = align_up(off + nblocks * sizeof(RedPartial), 256);
  // calmas partials: one range per block (its workgroup count)
  h->proj_begin.assign(nblocks + 1, 0);
  for (int b = 0; b < nblocks; ++b) {
    const BlockInfo &B = h->blk[b];
    const size_t cells = (size_t)B.spec.ni * B.spec.nj * B.spec.nk;
    h->proj_begin[b + 1] =
        h->proj_begin[b] + (int)std::min<size_t>((cells + 255) / 256, CALMAS_MAX_GROUPS);
  }
  const size_t proj_off = off;
  off = align_up(off + (size_t)h->proj_begin[nblocks] * sizeof(RedPartial), 256);
  const size_t pb_off = off;
  off = align_up(off + (nblocks + 1) * sizeof(int), 256);
  const size_t eptr_off = off;
  off = align_up(off + (size_t)nblocks * 8 * sizeof(double *), 256);

  // block regions and the global region: in one allocation, or (a handle
  // in a pool) at the caller's pool offsets plus a piece of the pool
  std::vector<char *> R(nblocks);
  char *A = nullptr;  // the global region
  if (h->pool) {
    for (int b = 0; b < nblocks; ++b) {
      const size_t have = pool_size_at(h->pool, h->blk_off[b]);
      if (have < boff[b])
        return fail(H3DMG_ERR_ARG, "block %d: pool piece at %zu holds %zu bytes, needs %zu", b,
                    h->blk_off[b], have, boff[b]);
      R[b] = h->pool->base + h->blk_off[b];
    }
    h3dmg_status st = pool_take(h->pool, off, &h->glob_off);
    if (st) return st;
    h->glob_taken = true;
    A = h->pool->base + h->glob_off;
    h->bytes = off;
  } else {
    size_t total = 0;
    std::vector<size_t> rbase(nblocks);
    for (int b = 0; b < nblocks; ++b) {
      rbase[b] = total;
      total = align_up(total + boff[b], 256);
    }
    const size_t gbase = total;
    total += off;
    h->arena_bytes = total;
    hipError_t e = hipMalloc((void **)&h->arena, total);
    if (e != hipSuccess)
      return fail(H3DMG_ERR_NOMEM, "hipMalloc(%zu bytes) failed: %s", total,
                  hipGetErrorString(e));
    h->bytes = total;
    for (int b = 0; b < nblocks; ++b) R[b] = h->arena + rbase[b];
    A = h->arena + gbase;
  }

  for (const Pending &P : pend) {
    LevelDev &L = h->levs[P.d];
    for (int bb = 0; bb < 2; ++bb)
      for (int c = 0; c < 2; ++c) L.phi[bb][c] = R[P.b] + P.phi[bb][c];
    for (int c = 0; c < 2; ++c) L.rhs[c] = R[P.b] + P.rhs[c];
    if (h->prm.smoother == H3DMG_SMOOTH_TDMA) L.tdma = (double *)(R[P.b] + tdma_off[P.d]);
  }
  h->fine.resize(nblocks);
  for (int b = 0; b < nblocks; ++b) {
    const LevelDev &L = h->levs[b];
    FineDev F{};
    F.ni = L.ni;
    F.nj = L.nj;
    F.nk = L.nk;
    F.H = L.H;
    F.per = L.per;
    for (int f = 0; f < 6; ++f) F.nbr[f] = L.nbr[f];  // level-1 descriptor d == block d
    F.lpr = lanes_per_row(L.H);  // two colour entries per lane
    F.blk = b;
    F.cx = L.cx;
    F.cy = L.cy;
    F.cz = L.cz;
    F.p[0] = (double *)(R[b] + foff[4 * b + 0]);
    F.p[1] = (double *)(R[b] + foff[4 * b + 1]);
    F.f[0] = (double *)(R[b] + foff[4 * b + 2]);
    F.f[1] = (double *)(R[b] + foff[4 * b + 3]);
    h->fine[b] = F;
  }
  h->d_stage = stage ? (double *)(A + stage_off) : nullptr;
  // Everything the global region holds that the host knows goes up as one
  // image (the tables, descriptors, flags and ranges; the partials and
  // statistics in between are written by the kernels before they are read)
  const size_t img0 = all.empty() ? levs_off : toff[0];
  std::vector<char> img(off - img0);
  auto put = [&](size_t at, const void *src, size_t n) {
    if (n) std::memcpy(img.data() + (at - img0), src, n);
  };
  for (size_t q = 0; q < all.size(); ++q) {
    all[q]->d = (Tile *)(A + toff[q]);
    all[q]->owned = nullptr;
    put(toff[q], all[q]->h.data(), all[q]->h.size() * sizeof(Tile));
  }
  h->d_levs = (LevelDev *)(A + levs_off);
  put(levs_off, h->levs.data(), h->levs.size() * sizeof(LevelDev));
  h->d_fine = (FineDev *)(A + fine_off);
  put(fine_off, h->fine.data(), nblocks * sizeof(FineDev));
  h->active.assign(nblocks, 1);
  h->d_active = (int *)(A + act_off);
  put(act_off, h->active.data(), nblocks * sizeof(int));
  h->d_tile_begin = (int *)(A + tb_off);
  put(tb_off, h->tile_begin.data(), (nblocks + 1) * sizeof(int));
  h->d_partials = (RedPartial *)(A + part_off);
  h->d_stats = (RedPartial *)(A + stats_off);
  h->d_proj = (RedPartial *)(A + proj_off);
  h->d_proj_begin = (int *)(A + pb_off);
  put(pb_off, h->proj_begin.data(), (nblocks + 1) * sizeof(int));
  h->vel.assign(nblocks, {});
  if (h->pool) {
    // a run of the pool's pinned statistics entries
    std::lock_guard<std::mutex> lock(h->pool->mu);
    const int cap = (int)h->pool->hs_used.size();
    int first = -1;
    for (int i = 0, run = 0; i < cap && first < 0; ++i) {
      run = h->pool->hs_used[i] ? 0 : run + 1;
      if (run == nblocks) first = i - nblocks + 1;
    }
    if (first < 0)
      return fail(H3DMG_ERR_NOMEM, "pool: no run of %d free block entries (of %d)", nblocks, cap);
    for (int b = 0; b < nblocks; ++b) h->pool->hs_used[first + b] = 1;
    h->hs_slot = first;
    h->h_stats = h->pool->h_stats + first;
  } else {
    HIP_TRY(hipHostMalloc((void **)&h->h_stats, nblocks * sizeof(RedPartial),
                          hipHostMallocDefault));
  }
  if (exact) {
    std::vector<double *> ptrs((size_t)nblocks * 8);
    size_t ed = 0;
    for (int b = 0; b < nblocks; ++b) {
      // the block's eigen tables (Q and lambda per axis, consecutive in its
      // region) in one copy; a block carried over into a pooled handle keeps
      // its tables
      const bool upload = fresh == nullptr || fresh[b];
      const size_t e0 = eoff[8 * b];
      std::vector<char> eimg(upload ? eoff[8 * b + 6] - e0 : 0);
      for (int a = 0; a < 3; ++a) {
        ptrs[8 * b + a] = (double *)(R[b] + eoff[8 * b + a]);
        if (upload)
          std::memcpy(eimg.data() + (eoff[8 * b + a] - e0), edata[ed].data(), edata[ed].size() * 8);
        ++ed;
        ptrs[8 * b + 3 + a] = (double *)(R[b] + eoff[8 * b + 3 + a]);
        if (upload)
          std::memcpy(eimg.data() + (eoff[8 * b + 3 + a] - e0), edata[ed].data(),
                      edata[ed].size() * 8);
        ++ed;
      }
      if (upload) HIP_TRY(h2d(h, R[b] + e0, eimg.data(), eimg.size()));
      if (upload && !h->pool) HIP_TRY(hipStreamSynchronize(h->stream));  // eimg goes out of scope
      ptrs[8 * b + 6] = (double *)(R[b] + eoff[8 * b + 6]);
      ptrs[8 * b + 7] = (double *)(R[b] + eoff[8 * b + 7]);
    }
    // the kernel indexes q/l tables with stride 3 and scratch with stride 2
    std::vector<double *> q3(3 * nblocks), l3(3 * nblocks), s2(2 * nblocks);
    for (int b = 0; b < nblocks; ++b) {
      for (int a = 0; a < 3; ++a) {
        q3[3 * b + a] = ptrs[8 * b + a];
        l3[3 * b + a] = ptrs[8 * b + 3 + a];
      }
      s2[2 * b] = ptrs[8 * b + 6];
      s2[2 * b + 1] = ptrs[8 * b + 7];
    }
    double **base = (double **)(A + eptr_off);
    h->d_eig_q = base;
    h->d_eig_l = base + 3 * nblocks;
    h->d_scratch = base + 6 * nblocks;
    put(eptr_off, q3.data(), q3.size() * sizeof(double *));
    put(eptr_off + 3 * nblocks * sizeof(double *), l3.data(), l3.size() * sizeof(double *));
    put(eptr_off + 6 * nblocks * sizeof(double *), s2.data(), s2.size() * sizeof(double *));
  }
  HIP_TRY(h2d(h, A + img0, img.data(), img.size()));
  if (h->prm.smoother == H3DMG_SMOOTH_TDMA) {  // after the descriptors are up
    launch_tdma_tables(h->stream, h->d_levs, (int)h->levs.size());
    HIP_TRY(hipGetLastError());
  }
  // p = 0, f = 0 (a block carried over into a pooled handle keeps its own)
  for (int b = 0; b < nblocks; ++b) {
    if (fresh && !fresh[b]) continue;
    const LevelDev &L = h->levs[b];
    const size_t half = (size_t)L.H * L.nj * L.nk * 8;
    for (int q = 0; q < 4; ++q) HIP_TRY(hipMemsetAsync(R[b] + foff[4 * b + q], 0, half, h->stream));
  }
  HIP_TRY(drain(h));  // img goes out of scope
  return H3DMG_OK;
}

// whether the whole-sweep TDMA kernel may use `budget` bytes of dynamic
// LDS (process-wide attribute: set once, to the device's budget)
bool tdma_sweep_ok(size_t budget) {
  static std::once_flag once;
  static bool ok = false;
  std::call_once(once, [budget]() { ok = tdma_sweep_allow_lds(budget) == hipSuccess; });
  return ok;
}

// whether the 16^3 GS kernel may use its dynamic LDS (the attribute is
// process-wide: set once, to what the kernel needs)
bool gs_big_ok() {
  static std::once_flag once;
  static bool ok = false;
  std::call_once(once, []() { ok = gs_allow_lds() == hipSuccess; });
  return ok;
}

// (Re)build the smoother tiles of level g on the host (both block groups)
h3dmg_status gen_relax(h3dmg_handle *h, int g) {
  for (int grp = 0; grp < 2; ++grp) {
    TileSet &ts = grp == 0 ? h->relax_nc[g] : h->relax_c[g];
    ts.h.clear();
    ts.nat = grp == 0 ? &h->nat_nc[g] : &h->nat_c[g];
    // one depth for the whole launch: the requested planes, or enough for
    // ~4096 workgroups over the group's blocks; at most 4, a power of two
    // dividing every block's nk; 1 where a level is one cell wide in x
    auto member = [&](const BlockInfo &B) { return B.L >= g && (grp == 1) == (B.L == g); };
    long cols = 0;
    bool general = false;
    std::vector<int> nks;
    for (const BlockInfo &B : h->blk)
      if (member(B)) {
        const LevelDev &L = h->levs[B.desc[g]];
        cols += (long)march_tiles_xy(L) * L.nk;
        nks.push_back(L.nk);
        general = general || L.ni < 2 || h->connected;
      }
    // small levels: one fused launch per sweep (latency-bound, the red halo
    // recomputed per tile costs less than a second launch)
    long most = 0;
    for (const BlockInfo &B : h->blk)
      if (member(B)) {
        const LevelDev &L = h->levs[B.desc[g]];
        most = std::max(most, (long)L.ni * L.nj * L.nk);
      }
    ts.fused = most > 0 && most <= h->fused_max && h->prm.smoother == H3DMG_SMOOTH_RBGS &&
               !h->connected;
    ts.ndiag[0] = ts.ndiag[1] = ts.ndiag[2] = 0;
    if (h->prm.smoother != H3DMG_SMOOTH_RBGS) {
      // wavefront smoothers: GS tiles of GS_TILE^3 cells sorted by
      // tile-diagonal (kind 5); TDMA tiles of tdma_group_lines(len) lines,
      // pass by pass (kind 6)
      ts.depth = 1;
      ts.diag_begin.clear();
      if (h->prm.smoother == H3DMG_SMOOTH_GS) {
        // 16^3 tiles where every block of the group is at least 16 cells
        // on each axis and the operator has constant coefficients (8^3
        // tiles stage the variable-density coefficients as well)
        bool big = !h->lsm && gs_big_ok();
        for (const BlockInfo &B : h->blk)
          if (member(B)) {
            const LevelDev &L = h->levs[B.desc[g]];
            big = big && std::min(L.ni, std::min(L.nj, L.nk)) >= GS_BIG_TILE;
          }
        const int gt = big ? GS_BIG_TILE : GS_TILE;
        ts.gs_tile = gt;
        auto tiles_of = [gt](int n) { return (n + gt - 1) / gt; };
        int nd = 0;
        for (const BlockInfo &B : h->blk)
          if (member(B)) {
            const LevelDev &L = h->levs[B.desc[g]];
            nd = std::max(nd, tiles_of(L.ni) + tiles_of(L.nj) + tiles_of(L.nk) - 2);
          }
        ts.ndiag[0] = nd;
        for (int D = 0; D < nd; ++D) {
          ts.diag_begin.push_back(ts.n());
          for (const BlockInfo &B : h->blk)
            if (member(B)) {
              const int d = B.desc[g];
              const LevelDev &L = h->levs[d];
              for (int K = 0; K < tiles_of(L.nk); ++K)
                for (int J = 0; J < tiles_of(L.nj); ++J) {
                  const int I = D - K - J;
                  if (I >= 0 && I < tiles_of(L.ni))
                    ts.h.push_back(Tile{d, I * gt, J * gt, K * gt, K * gt + 1, 5});
                }
            }
        }
        ts.diag_begin.push_back(ts.n());
      } else {
        // small levels: whole sweeps in one workgroup per block
        bool small = true;
        size_t lds = 0;
        for (const BlockInfo &B : h->blk)
          if (member(B)) {
            const LevelDev &L = h->levs[B.desc[g]];
            // cells per diagonal: ni*nj (j-lines), nj*ni (i-lines), ni*nk (k-lines)
            // (the pipelined passes keep whole waves free for the next
            // diagonal's loads: at most half the workgroup runs lines)
            small = small && std::max(L.ni, L.nj) <= TDMA_SWEEP_THREADS / 2 &&
                    (long)L.ni * L.nj <= TDMA_SWEEP_CELLS && (long)L.ni * L.nk <= TDMA_SWEEP_CELLS;
            lds = std::max(lds, tdma_sweep_lds_bytes(L, !h->lsm));
          }
        small = small && lds <= h->lds_budget && tdma_sweep_ok(h->lds_budget);
        ts.tdma_lds = small ? lds : 0;
        ts.tdma_pipe = !h->lsm;
        if (small)
          for (const BlockInfo &B : h->blk)
            if (member(B)) ts.h.push_back(Tile{B.desc[g], 0, 0, 0, 1, 7});
        for (int pass = 0; pass < 3 && !small; ++pass) {
          ts.pass_begin[pass] = ts.n();
          for (const BlockInfo &B : h->blk)
            if (member(B)) {
              const int d = B.desc[g];
              const LevelDev &L = h->levs[d];
              const int nu = pass == 1 ? L.nj : L.ni, nv = pass == 2 ? L.nj : L.nk;
              const int len = pass == 0 ? L.nj : pass == 1 ? L.ni : L.nk;
              const int lw = tdma_group_lines(len);
              for (int x = 0; x < nu; x += lw) ts.h.push_back(Tile{d, x, 0, 0, 1, 6});
              ts.ndiag[pass] = std::max(ts.ndiag[pass], nu + nv - 1);
            }
        }
        ts.pass_begin[3] = ts.n();
      }
    } else if (ts.fused) {
      ts.depth = FUSED_Z;
      for (const BlockInfo &B : h->blk)
        if (member(B)) {
          const LevelDev &L = h->levs[B.desc[g]];
          for (int z = 0; z < L.nk; z += FUSED_Z)
            for (int y = 0; y < L.nj; y += FUSED_Y)
              for (int x = 0; x < L.ni; x += FUSED_X)
                ts.h.push_back(Tile{B.desc[g], x, y, z, std::min(z + FUSED_Z, L.nk), 4});
        }
    } else {
      const int kb = general ? 1 : launch_depth(cols, 4096, h->march_kb[g], 4, nks);
      ts.depth = kb;
      for (int b = 0; b < h->nb; ++b) {
        const BlockInfo &B = h->blk[b];
        if (!member(B)) continue;
        add_tiles_march(ts, B.desc[g], h->levs[B.desc[g]], kb, h->connected);
      }
    }
    for (const Tile &t : ts.h)
      if (t.kind != 2 && t.kind != 4 && t.z1 - t.z0 != ts.depth)
        return fail(H3DMG_ERR_ARG, "internal: tile depth");
    set_one(ts, h->levs, 0);
  }
  return H3DMG_OK;
}

// Place a table set in a piece of its own (replacing the one it had) and
// upload it.  sync = false: nothing is replaced and the caller drains the
// stream (a fresh handle's build).
h3dmg_status place_owned(h3dmg_handle *h, TileSet &ts, bool sync) {
  Tile *d = nullptr;
  if (!ts.h.empty()) {
    h3dmg_status st = dev_alloc(h, ts.h.size() * sizeof(Tile), (void **)&d);
    if (st) return st;
    HIP_TRY(h2d(h, d, ts.h.data(), ts.h.size() * sizeof(Tile)));
  }
  if (sync || ts.owned) HIP_TRY(drain(h));
  dev_free(h, ts.owned);
  ts.owned = d;
  ts.d = d;
  return H3DMG_OK;
}

h3dmg_status upload_relax(h3dmg_handle *h, int g, bool sync) {
  h3dmg_status st = gen_relax(h, g);
  if (st) return st;
  if ((st = place_owned(h, h->relax_nc[g], sync))) return st;
  return place_owned(h, h->relax_c[g], sync);
}

// LDS the tail needs when it starts at level g (max over the blocks that
// reach g; 0 when none does)
size_t tail_lds_at(const h3dmg_handle *h, int g) {
  size_t need = 0;
  for (const BlockInfo &B : h->blk)
    if (B.L >= g)
      need = std::max(need, vtail_lds_bytes(h->levs.data(), B.desc[g], is_mixed(h),
                                            h->prm.coarse == H3DMG_COARSE_EXACT));
  return need;
}

// Automatic tail level: the first level from which every block's remaining
// levels fit in one workgroup's LDS; none when no level does.
// The tail's one workgroup per block is bound by its CU's VALU issue, so it
// pays for the exact solves of the blocks' coarsest levels (all of them in
// one launch, without their levels' global round trips), not for levels
// above them, and not for a smoothed coarsest level: with the exact solve,
// the first level from the shallowest block's coarsest down whose tail fits
// in LDS; otherwise off.  Measured on MI355X (single 32^3 to 256^3 blocks
// and batches of 2 to 48 blocks of mixed sizes, both precisions).
int auto_tail_level(const h3dmg_handle *h) {
  if (h->prm.coarse != H3DMG_COARSE_EXACT) return h->Lmax + 1;
  int t0 = h->Lmax;
  for (const BlockInfo &B : h->blk) t0 = std::min(t0, B.L);
  for (int g = std::max(t0, 1); g <= h->Lmax; ++g) {
    const size_t need = tail_lds_at(h, g);
    if (need > 0 && need <= h->lds_budget) return g;
  }
  return h->Lmax + 1;
}

h3dmg_status gen_tail(h3dmg_handle *h) {
  h->tail_lt = h->tail_req > 0 ? h->tail_req : auto_tail_level(h);
  h->tail_lds = h->tail_lt <= h->Lmax ? tail_lds_at(h, h->tail_lt) : 0;
  if (h->tail_lds > h->lds_budget) {
    const int lt = h->tail_lt;
    h->tail_req = 0;
    h->tail_lt = h->Lmax + 1;
    h->tail_lds = 0;
    return fail(H3DMG_ERR_ARG, "tail level %d needs %zu bytes of LDS (at most %zu here)", lt,
                tail_lds_at(h, lt), h->lds_budget);
  }
  // The permitted LDS is a process-wide attribute of the kernel: set once,
  // to the device's whole budget, so that handles of different tail sizes
  // coexist and no later handle changes it while another thread launches.
  if (h->tail_lds > 65536) {
    static std::once_flag once;
    static hipError_t once_err = hipSuccess;
    const size_t budget = h->lds_budget;
    std::call_once(once, [budget]() { once_err = vtail_allow_lds(budget); });
    HIP_TRY(once_err);
  }
  TileSet &ts = h->tail_tiles;
  ts.h.clear();
  if (h->tail_lt <= h->Lmax)
    for (int b = 0; b < h->nb; ++b)
      if (h->blk[b].L >= h->tail_lt) ts.h.push_back(Tile{h->blk[b].desc[h->tail_lt], b, 0, 0, 0, 0});
  return H3DMG_OK;
}

h3dmg_status upload_tail(h3dmg_handle *h, bool sync) {
  h3dmg_status st = gen_tail(h);
  if (st) return st;
  return place_owned(h, h->tail_tiles, sync);
}

h3dmg_status check_block(const h3dmg_handle *h, int b) {
  if (!h) return fail(H3DMG_ERR_ARG, "null handle");
  if (b < 0 || b >= h->nb) return fail(H3DMG_ERR_ARG, "block index %d out of range", b);
  return H3DMG_OK;
}

void drop_graph(h3dmg_handle *h) {
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->gexec = nullptr;
  h->graph = nullptr;
  if (h->gexec_dec) (void)hipGraphExecDestroy(h->gexec_dec);
  if (h->graph_dec) (void)hipGraphDestroy(h->graph_dec);
  h->gexec_dec = nullptr;
  h->graph_dec = nullptr;
  if (h->vgexec) (void)hipGraphExecDestroy(h->vgexec);
  if (h->vgraph) (void)hipGraphDestroy(h->vgraph);
  h->vgexec = nullptr;
  h->vgraph = nullptr;
}

// calvel + calmas (press.for:77-144, 2-75) of every enabled block with the
// current pressure: u, v, w from u*, v*, w*; su into the blocks' f arrays
// (the next V-cycle's rhs); per-block max |div u| dx dy dz into d_stats.
// Where calmas' 4th-order stencil finds faces -1 and n+1 of block b (the
// second ghost faces exchange(1..3) fills, press.for:17-37 and
// exchange_bc/exchange_smlvl): the block across the face -- the neighbour,
// or the block itself on an axis it spans periodically -- or, at a wall,
// the host values uploaded by h3dmg_set_velocity.
Outer4 outer_faces(const h3dmg_handle *h, int b) {
  const BlockInfo &B = h->blk[b];
  const size_t *o = &h->vel_off[9 * (size_t)b];
  Outer4 r{};
  for (int a = 0; a < 3; ++a) {
    const size_t plane = a == 0 ? (size_t)B.spec.nj * B.spec.nk
                         : a == 1 ? (size_t)B.spec.ni * B.spec.nk
                                  : (size_t)B.spec.ni * B.spec.nj;
    for (int side = 0; side < 2; ++side) {
      const int f = 2 * a + side, k = B.spec.bc[f];
      if (k == H3DMG_BC_WALL) {
        r.src[f] = h->vel_store + o[6 + a] + side * plane;
        r.n[f] = 0;
        continue;
      }
      const int q = k == H3DMG_BC_NEIGHBOUR ? B.spec.nbr[f] : b;
      const h3dmg_block &Q = h->blk[q].spec;
      const h3dmg_handle::VelBlock &V = h->vel[q];
      r.src[f] = a == 0 ? V.u : a == 1 ? V.v : V.w;
      r.n[f] = a == 0 ? Q.ni : a == 1 ? Q.nj : Q.nk;
    }
  }
  return r;
}

// calmas of block b into its rhs arrays, in the handle's branch
void enqueue_calmas(h3dmg_handle *h, int b) {
  const BlockInfo &B = h->blk[b];
  const h3dmg_handle::VelBlock &V = h->vel[b];
  const size_t cells = (size_t)B.spec.ni * B.spec.nj * B.spec.nk;
  if (h->differencing == 2)
    launch_calmas4(h->stream, h->d_fine + b, V.u, V.v, V.w, nullptr, 1, h->prm.dt, B.spec.dx,
                   B.spec.dy, B.spec.dz, h->d_proj + h->proj_begin[b], cells, outer_faces(h, b));
  else
    launch_calmas(h->stream, h->d_fine + b, V.u, V.v, V.w, nullptr, 1, h->prm.dt, B.spec.dx,
                  B.spec.dy, B.spec.dz, h->d_proj + h->proj_begin[b], cells);
}

void enqueue_project(h3dmg_handle *h) {
  if (h->connected) {
    // calvel on the faces each block computes, then exchange(1..3) of the
    // near faces from the blocks across them (press.for:77-144), then calmas
    for (int b = 0; b < h->nb; ++b) {
      const BlockInfo &B = h->blk[b];
      const h3dmg_handle::VelBlock &V = h->vel[b];
      const size_t ni = B.spec.ni, nj = B.spec.nj, nk = B.spec.nk;
      launch_calvel_nb(h->stream, h->d_fine, b, 0, V.us, V.u, h->vel_dta, B.spec.dx,
                       (ni + 1) * nj * nk);
      launch_calvel_nb(h->stream, h->d_fine, b, 1, V.vs, V.v, h->vel_dta, B.spec.dy,
                       ni * (nj + 1) * nk);
      launch_calvel_nb(h->stream, h->d_fine, b, 2, V.ws, V.w, h->vel_dta, B.spec.dz,
                       ni * nj * (nk + 1));
    }
    for (int b = 0; b < h->nb; ++b) {
      const BlockInfo &B = h->blk[b];
      const int n[3] = {B.spec.ni, B.spec.nj, B.spec.nk};
      for (int a = 0; a < 3; ++a) {
        const int k = B.spec.bc[2 * a];
        if (k == H3DMG_BC_WALL) continue;
        const int q = k == H3DMG_BC_NEIGHBOUR ? B.spec.nbr[2 * a] : b;
        const h3dmg_block &Q = h->blk[q].spec;
        const int nq[3] = {Q.ni, Q.nj, Q.nk};
        double *dst = a == 0 ? h->vel[b].u : a == 1 ? h->vel[b].v : h->vel[b].w;
        const double *src = a == 0 ? h->vel[q].u : a == 1 ? h->vel[q].v : h->vel[q].w;
        const int nu = a == 0 ? n[1] : n[0], nv = a == 2 ? n[1] : n[2];
        launch_face_copy(h->stream, a, src, nq[a], dst, n[a], nu, nv);
      }
    }
    for (int b = 0; b < h->nb; ++b) enqueue_calmas(h, b);
    launch_finalize(h->stream, h->d_proj_begin, h->nb, h->d_proj, h->d_stats);
    return;
  }
  for (int b = 0; b < h->nb; ++b) {
    const BlockInfo &B = h->blk[b];
    if (!B.enabled) continue;
    const h3dmg_handle::VelBlock &V = h->vel[b];
    const size_t ni = B.spec.ni, nj = B.spec.nj, nk = B.spec.nk;
    const FineDev *fd = h->d_fine + b;
    launch_calvel(h->stream, fd, 0, V.us, V.u, h->vel_dta, B.spec.dx, (ni + 1) * nj * nk);
    launch_calvel(h->stream, fd, 1, V.vs, V.v, h->vel_dta, B.spec.dy, ni * (nj + 1) * nk);
    launch_calvel(h->stream, fd, 2, V.ws, V.w, h->vel_dta, B.spec.dz, ni * nj * (nk + 1));
    enqueue_calmas(h, b);
  }
  launch_finalize(h->stream, h->d_proj_begin, h->nb, h->d_proj, h->d_stats);
}

// max over the enabled blocks of calmas' rmax (the reference's MPI MAX over
// blocks and ranks, press.for:72-73, for the blocks of this handle)
double projected_rmax(const h3dmg_handle *h) {
  double m = 0.0;
  for (int b = 0; b < h->nb; ++b)
    if (h->blk[b].enabled) m = std::max(m, h->h_stats[b].amax);
  return m;
}

// One iteration of newsolv_mg's loop (newsolv_mg.for:24-55): mgkcyc from su
// (level-1 rhs <- f, pp = 0), p = p + (pp - 0), calvel + calmas.
void enqueue_velocity_step(h3dmg_handle *h) {
  const bool f64 = !is_mixed(h);
  const TileSet &ft = h->fine_tiles, &ut = h->update_tiles;
  if (f64)
    launch_residual<double>(h->stream, ft.d, ft.n(), ft.depth, h->d_fine, h->d_levs, h->d_active,
                            h->d_partials, 0);
  else
    launch_residual<float>(h->stream, ft.d, ft.n(), ft.depth, h->d_fine, h->d_levs, h->d_active,
                           h->d_partials, 0);
  enqueue_vcycle(h);
  if (f64)
    launch_update<double>(h->stream, ut.d, ut.n(), h->d_fine, h->d_levs, h->d_active, h->pp_buf);
  else
    launch_update<float>(h->stream, ut.d, ut.n(), h->d_fine, h->d_levs, h->d_active, h->pp_buf);
  enqueue_project(h);
}

h3dmg_status run_velocity_step(h3dmg_handle *h) {
  if (h->prm.use_graph) {
    if (!h->vgexec) {
      std::lock_guard<std::mutex> lock(g_device_mutex);
      HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
      enqueue_velocity_step(h);
      hipGraph_t g = nullptr;
      HIP_TRY(hipStreamEndCapture(h->stream, &g));
      h->vgraph = g;
      HIP_TRY(hipGraphInstantiate(&h->vgexec, h->vgraph, nullptr, nullptr, 0));
    }
    HIP_TRY(hipGraphLaunch(h->vgexec, h->stream));
  } else {
    enqueue_velocity_step(h);
  }
  HIP_TRY(hipGetLastError());
  return H3DMG_OK;
}

// One outer iteration: V-cycle on the current level-1 rhs, p += pp, new
// residual and per-block statistics.
h3dmg_status run_cycle(h3dmg_handle *h) {
  if (h->prm.use_graph) {
    if (!h->gexec) {
      std::lock_guard<std::mutex> lock(g_device_mutex);
      HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
      enqueue_vcycle(h);
      enqueue_update_residual(h, true);
      hipGraph_t g = nullptr;
      HIP_TRY(hipStreamEndCapture(h->stream, &g));
      h->graph = g;
      HIP_TRY(hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0));
    }
    HIP_TRY(hipGraphLaunch(h->gexec, h->stream));
  } else {
    enqueue_vcycle(h);
    enqueue_update_residual(h, true);
  }
  HIP_TRY(hipGetLastError());
  return H3DMG_OK;
}

}  // namespace

// ===========================================================================
extern "C" {

void h3dmg_default_params(h3dmg_params *p) {
  if (!p) return;
  p->smoother = H3DMG_SMOOTH_RBGS;
  p->nu1 = 2;
  p->nu2 = 2;
  p->maxcy = 60;
  p->criterion = H3DMG_CRIT_REL_L2;
  p->tol = 1e-8;
  p->dt = 1e-3;
  p->precision = H3DMG_FP64;
  p->coarse = H3DMG_COARSE_SMOOTH;
  p->min_cycles = 1;
  p->use_graph = 1;
}

const char *h3dmg_last_error(void) { return g_last_error.c_str(); }

// Frees a handle; the caller holds g_device_mutex.
static void destroy_locked(h3dmg_handle *h) {
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  drop_graph(h);
  for (auto *v : {&h->relax_nc, &h->relax_c})
    for (TileSet &ts : *v) dev_free(h, ts.owned);
  dev_free(h, h->tail_tiles.owned);
  dev_free(h, h->d_dec);
  if (h->h_ring) (void)hipHostFree(h->h_ring);
  for (hipEvent_t e : h->ev_dec)
    if (e) (void)hipEventDestroy(e);
  if (h->pool) {
    if (h->glob_taken) (void)pool_give(h->pool, h->glob_off);
    std::lock_guard<std::mutex> lock(h->pool->mu);
    if (h->hs_slot >= 0)
      for (int b = 0; b < h->nb; ++b) h->pool->hs_used[h->hs_slot + b] = 0;
    --h->pool->live;
    h->h_stats = nullptr;  // the pool's
  }
  if (h->h_stats) (void)hipHostFree(h->h_stats);
  if (h->vel_store) (void)hipFree(h->vel_store);
  if (h->lsm_store) (void)hipFree(h->lsm_store);
  if (h->arena) (void)hipFree(h->arena);
  if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

}  // extern "C"

namespace {

h3dmg_status check_params(const h3dmg_params &P) {
  if (P.smoother != H3DMG_SMOOTH_GS && P.smoother != H3DMG_SMOOTH_TDMA &&
      P.smoother != H3DMG_SMOOTH_RBGS)
    return fail(H3DMG_ERR_ARG, "unknown smoother %d", P.smoother);
  if (P.smoother != H3DMG_SMOOTH_RBGS && P.precision != H3DMG_FP64)
    return fail(H3DMG_ERR_UNSUPPORTED, "the GS and line-TDMA smoothers run in FP64");
  if (P.nu1 < 0 || P.nu2 < 0 || P.nu1 + P.nu2 < 1 || P.maxcy < 0 ||
      (P.precision != H3DMG_FP64 && P.precision != H3DMG_MIXED) ||
      (P.coarse != H3DMG_COARSE_SMOOTH && P.coarse != H3DMG_COARSE_EXACT) ||
      (P.criterion != H3DMG_CRIT_REL_L2 && P.criterion != H3DMG_CRIT_REF_RMAX))
    return fail(H3DMG_ERR_ARG, "invalid parameters");
  return H3DMG_OK;
}

// h3dmg_create and h3dmg_create_pooled; the caller holds g_device_mutex
h3dmg_status create_locked(const h3dmg_block *blocks, int nblocks, const h3dmg_params *prm,
                           h3dmg_pool *pool, const unsigned long long *offsets,
                           const int *fresh, void *stream, h3dmg_handle **out) {
  if (!out) return fail(H3DMG_ERR_ARG, "null output handle pointer");
  *out = nullptr;
  if (!blocks || nblocks < 1) return fail(H3DMG_ERR_ARG, "no blocks");
  h3dmg_params P;
  if (prm)
    P = *prm;
  else
    h3dmg_default_params(&P);
  if (h3dmg_status st = check_params(P)) return st;
  if (h3dmg_status st = validate_blocks(blocks, nblocks, P)) return st;
  if (pool) {
    if (!offsets) return fail(H3DMG_ERR_ARG, "null block offsets");
    for (int b = 0; b < nblocks; ++b)
      for (int f = 0; f < 6; ++f)
        if (blocks[b].bc[f] == H3DMG_BC_NEIGHBOUR)
          return fail(H3DMG_ERR_UNSUPPORTED,
                      "a pooled handle holds independent blocks (block %d has a neighbour)", b);
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(H3DMG_ERR_DEVICE, "no HIP device");
  if (pool && dev != pool->device)
    return fail(H3DMG_ERR_ARG, "the pool lives on device %d, the current device is %d",
                pool->device, dev);
  h3dmg_handle *h = new h3dmg_handle();
  h->prm = P;
  h->device = dev;
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    lds = 65536;
  // the tail kernel's static LDS (its level descriptors) comes on top
  h->lds_budget = lds > 8192 ? (size_t)lds - 8192 : 0;
  if (stream) {
    h->stream = (hipStream_t)stream;
  } else {
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
      delete h;
      return fail(H3DMG_ERR_DEVICE, "hipStreamCreate failed");
    }
    h->own_stream = true;
  }
  if (pool) {
    h->pool = pool;
    h->blk_off.assign(offsets, offsets + nblocks);
    std::lock_guard<std::mutex> lock(pool->mu);
    ++pool->live;  // destroy_locked gives it back, on failure too
  }
  h3dmg_status st = build(h, blocks, nblocks, fresh);
  if (st != H3DMG_OK) {
    destroy_locked(h);
    return st;
  }
  h->running.assign(nblocks, 0);
  h->outcome.assign(nblocks, -1);
  *out = h;
  return H3DMG_OK;
}

}  // namespace

extern "C" {

h3dmg_status h3dmg_create(const h3dmg_block *blocks, int nblocks, const h3dmg_params *prm,
                          h3dmg_handle **out) {
  std::lock_guard<std::mutex> lock(g_device_mutex);
  return create_locked(blocks, nblocks, prm, nullptr, nullptr, nullptr, nullptr, out);
}

h3dmg_status h3dmg_pool_create(unsigned long long bytes, int max_blocks, h3dmg_pool **out) {
  std::lock_guard<std::mutex> lock(g_device_mutex);
  if (!out) return fail(H3DMG_ERR_ARG, "null output pool pointer");
  *out = nullptr;
  if (bytes < 256 || max_blocks < 1) return fail(H3DMG_ERR_ARG, "empty pool");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(H3DMG_ERR_DEVICE, "no HIP device");
  h3dmg_pool *p = new h3dmg_pool();
  p->device = dev;
  p->bytes = align_up(bytes, 256);
  hipError_t e = hipMalloc((void **)&p->base, p->bytes);
  if (e != hipSuccess) {
    delete p;
    return fail(H3DMG_ERR_NOMEM, "hipMalloc(%llu bytes) failed: %s", bytes, hipGetErrorString(e));
  }
  e = hipHostMalloc((void **)&p->h_stats, (size_t)max_blocks * sizeof(RedPartial),
                    hipHostMallocDefault);
  if (e != hipSuccess) {
    (void)hipFree(p->base);
    delete p;
    return fail(H3DMG_ERR_NOMEM, "hipHostMalloc failed: %s", hipGetErrorString(e));
  }
  p->hs_used.assign(max_blocks, 0);
  p->free_[0] = p->bytes;
  // staging of the handles' table uploads (optional: without it they copy
  // from pageable memory)
  if (hipHostMalloc((void **)&p->h_stage, POOL_STAGE_BYTES, hipHostMallocDefault) == hipSuccess)
    p->stage_bytes = POOL_STAGE_BYTES;
  else
    p->h_stage = nullptr;
  *out = p;
  return H3DMG_OK;
}

void h3dmg_pool_destroy(h3dmg_pool *pool) {
  if (!pool) return;
  std::lock_guard<std::mutex> lock(g_device_mutex);
  {
    std::lock_guard<std::mutex> plock(pool->mu);
    if (pool->live > 0) {  // its memory is still in use: keep it
      fail(H3DMG_ERR_ARG, "pool destroyed with %d live handles: not freed", pool->live);
      return;
    }
  }
  (void)hipDeviceSynchronize();
  (void)hipFree(pool->base);
  (void)hipHostFree(pool->h_stats);
  if (pool->h_stage) (void)hipHostFree(pool->h_stage);
  delete pool;
}

h3dmg_status h3dmg_pool_alloc(h3dmg_pool *pool, unsigned long long bytes,
                              unsigned long long *offset) {
  if (!pool || !offset) return fail(H3DMG_ERR_ARG, "null pool or offset");
  size_t off = 0;
  h3dmg_status st = pool_take(pool, bytes, &off);
  if (st) return st;
  *offset = off;
  return H3DMG_OK;
}

h3dmg_status h3dmg_pool_free(h3dmg_pool *pool, unsigned long long offset) {
  if (!pool) return fail(H3DMG_ERR_ARG, "null pool");
  return pool_give(pool, offset);
}

unsigned long long h3dmg_pool_free_bytes(h3dmg_pool *pool) {
  if (!pool) return 0;
  std::lock_guard<std::mutex> lock(pool->mu);
  size_t n = 0;
  for (const auto &f : pool->free_) n += f.second;
  return n;
}

unsigned long long h3dmg_pool_largest_free(h3dmg_pool *pool) {
  if (!pool) return 0;
  std::lock_guard<std::mutex> lock(pool->mu);
  size_t n = 0;
  for (const auto &f : pool->free_) n = std::max(n, f.second);
  return n;
}

unsigned long long h3dmg_block_bytes(const h3dmg_block *blk, const h3dmg_params *prm) {
  if (!blk) {
    fail(H3DMG_ERR_ARG, "null block");
    return 0;
  }
  h3dmg_params P;
  if (prm)
    P = *prm;
  else
    h3dmg_default_params(&P);
  if (check_params(P) || validate_blocks(blk, 1, P)) return 0;
  h3dmg_handle tmp;
  tmp.prm = P;
  size_t bytes = 0;
  if (build(&tmp, blk, 1, nullptr, &bytes)) return 0;
  return bytes;
}

h3dmg_status h3dmg_create_pooled(h3dmg_pool *pool, const h3dmg_block *blocks, int nblocks,
                                 const h3dmg_params *prm, const unsigned long long *offsets,
                                 const int *fresh, void *stream, h3dmg_handle **out) {
  if (!pool) return fail(H3DMG_ERR_ARG, "null pool");
  std::lock_guard<std::mutex> lock(g_device_mutex);
  return create_locked(blocks, nblocks, prm, pool, offsets, fresh, stream, out);
}

void h3dmg_destroy(h3dmg_handle *h) {
  if (!h) return;
  std::lock_guard<std::mutex> lock(g_device_mutex);
  destroy_locked(h);
}
unsigned long long h3dmg_device_bytes(const h3dmg_handle *h) { return h ? h->bytes : 0; }

h3dmg_status h3dmg_set_stream(h3dmg_handle *h, void *stream) {
  if (!h) return fail(H3DMG_ERR_ARG, "null handle");
  HIP_TRY(hipStreamSynchronize(h->stream));
  drop_graph(h);
  if (h->own_stream) (void)hipStreamDestroy(h->stream);
  if (stream) {
    h->stream = (hipStream_t)stream;
    h->own_stream = false;
  } else {
    HIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    h->own_stream = true;
  }
  return H3DMG_OK;
}

static h3dmg_status no_stage() {
  return fail(H3DMG_ERR_UNSUPPORTED, "a pooled handle takes device arrays only");
}

// host Fortran array (i fastest, leading dims ld1, ld2, interior at off) ->
// dense interior staging on the device
static h3dmg_status host_to_stage(h3dmg_handle *h, int b, const double *src, int ld1, int ld2,
                                  int off) {
  if (!h->d_stage) return no_stage();
  const h3dmg_block &S = h->blk[b].spec;
  if (!src || ld1 < S.ni + off || ld2 < S.nj + off || off < 0)
    return fail(H3DMG_ERR_ARG, "bad host array description");
  std::vector<double> dense((size_t)S.ni * S.nj * S.nk);
  for (int k = 0; k < S.nk; ++k)
    for (int j = 0; j < S.nj; ++j)
      std::memcpy(&dense[((size_t)k * S.nj + j) * S.ni],
                  src + (size_t)off + (size_t)ld1 * ((size_t)(j + off) + (size_t)ld2 * (k + off)),
                  sizeof(double) * S.ni);
  HIP_TRY(hipMemcpyAsync(h->d_stage, dense.data(), dense.size() * 8, hipMemcpyHostToDevice,
                         h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return H3DMG_OK;
}

// padded dense (n+2)^3 staging -> host Fortran array incl. the ghost layer
static h3dmg_status stage_padded_to_host(h3dmg_handle *h, int b, double *dst, int ld1, int ld2,
                                         int off) {
  const h3dmg_block &S = h->blk[b].spec;
  if (!dst || off < 1 || ld1 < S.ni + off + 1 || ld2 < S.nj + off + 1)
    return fail(H3DMG_ERR_ARG, "bad host array description (needs one ghost layer)");
  const size_t pi = S.ni + 2, pj = S.nj + 2, pk = S.nk + 2;
  std::vector<double> dense(pi * pj * pk);
  HIP_TRY(hipMemcpyAsync(dense.data(), h->d_stage, dense.size() * 8, hipMemcpyDeviceToHost,
                         h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  for (size_t k = 0; k < pk; ++k)
    for (size_t j = 0; j < pj; ++j)
      std::memcpy(dst + (size_t)(off - 1) +
                      (size_t)ld1 * ((j + off - 1) + (size_t)ld2 * (k + off - 1)),
                  &dense[(k * pj + j) * pi], sizeof(double) * pi);
  return H3DMG_OK;
}

h3dmg_status h3dmg_set_rhs_device(h3dmg_handle *h, int b, const double *d_su) {
  h3dmg_status st = check_block(h, b);
  if (st) return st;
  if (!d_su) return fail(H3DMG_ERR_ARG, "null rhs");
  const FineDev &F = h->fine[b];
  launch_pack_fine(h->stream, h->d_fine + b, 1, d_su, F.ni, F.nj, F.nk);
  // ||f||: residual with use_p = 0 over this block only
  std::vector<int> act(h->nb, 0);
  act[b] = 1;
  HIP_TRY(hipMemcpyAsync(h->d_active, act.data(), sizeof(int) * h->nb, hipMemcpyHostToDevice,
                         h->stream));
  const TileSet &ft = h->fine_tiles;
  if (is_mixed(h))
    launch_residual<float>(h->stream, ft.d, ft.n(), ft.depth, h->d_fine, h->d_levs, h->d_active,
                           h->d_partials, 0);
  else
    launch_residual<double>(h->stream, ft.d, ft.n(), ft.depth, h->d_fine, h->d_levs, h->d_active,
                            h->d_partials, 0);
  launch_finalize(h->stream, h->d_tile_begin, h->nb, h->d_partials, h->d_stats);
  HIP_TRY(hipGetLastError());
  st = read_stats(h);
  if (st) return st;
  h->blk[b].fnorm = std::sqrt(h->h_stats[b].ss);
  st = push_active(h);
  if (st) return st;
  h->rhs_ready = true;
  return H3DMG_OK;
}

h3dmg_status h3dmg_set_rhs(h3dmg_handle *h, int b, const double *su, int ld1, int ld2, int off) {
  h3dmg_status st = check_block(h, b);
  if (st) return st;
  st = host_to_stage(h, b, su, ld1, ld2, off);
  if (st) return st;
  return h3dmg_set_rhs_device(h, b, h->d_stage);
}

h3dmg_status h3dmg_set_p_device(h3dmg_handle *h, int b, const double *d_p) {
  h3dmg_status st = check_block(h, b);
  if (st) return st;
  if (!d_p) return fail(H3DMG_ERR_ARG, "null p");
  const FineDev &F = h->fine[b];
  launch_pack_fine(h->stream, h->d_fine + b, 0, d_p, F.ni, F.nj, F.nk);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  return H3DMG_OK;
}

h3dmg_status h3dmg_get_p_device(h3dmg_handle *h, int b, double *d_p) {
  h3dmg_status st = check_block(h, b);
  if (st) return st;
  if (!d_p) return fail(H3DMG_ERR_ARG, "null p");
  const FineDev &F = h->fine[b];
  launch_unpack_fine(h->stream, h->d_fine + b, 0, d_p, F.ni, F.nj, F.nk);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  return H3DMG_OK;
}

h3dmg_status h3dmg_set_p(h3dmg_handle *h, int b, const double *p, int ld1, int ld2, int off) {
  h3dmg_status st = check_block(h, b);
  if (st) return st;
  st = host_to_stage(h, b, p, ld1, ld2, off);
  if (st) return st;
  return h3dmg_set_p_device(h, b, h->d_stage);
}

h3dmg_status h3dmg_get_p(h3dmg_handle *h, int b, double *p, int ld1, int ld2, int off) {
  h3dmg_status st = check_block(h, b);
  if (st) return st;
  if (!h->d_stage) return no_stage();
  // the interior and the ghost layer (the cells pbound and exchange(4)
  // copy into it: mirror, wrap or the neighbour block's cell)
  const FineDev &F = h->fine[b];
  launch_unpack_fine_padded(h->stream, h->d_fine, b, h->d_stage, F.ni, F.nj, F.nk);
  HIP_TRY(hipGetLastError());
  return stage_padded_to_host(h, b, p, ld1, ld2, off);
}

h3dmg_status h3dmg_get_su_device(h3dmg_handle *h, int b, double *d_su) {
  h3dmg_status st = check_block(h, b);
  if (st) return st;
  if (!d_su) return fail(H3DMG_ERR_ARG, "null su");
  const FineDev &F = h->fine[b];
  launch_unpack_fine(h->stream, h->d_fine + b, 1, d_su, F.ni, F.nj, F.nk);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  return H3DMG_OK;
}

h3dmg_status h3dmg_vcycle(h3dmg_handle *h) {
  if (!h) return fail(H3DMG_ERR_ARG, "null handle");
  if (!h->rhs_ready) return fail(H3DMG_ERR_ARG, "rhs not set");
  if (h3dmg_status st = check_density(h)) return st;
  // level-1 rhs <- f (su), pp = 0
  reset_active(h);
  h3dmg_status st = push_active(h);
  if (st) return st;
  const TileSet &ft = h->fine_tiles;
  if (is_mixed(h))
    launch_residual<float>(h->stream, ft.d, ft.n(), ft.depth, h->d_fine, h->d_levs, h->d_active,
                           h->d_partials, 0);
  else
    launch_residual<double>(h->stream, ft.d, ft.n(), ft.depth, h->d_fine, h->d_levs, h->d_active,
                            h->d_partials, 0);
  enqueue_vcycle(h);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  return H3DMG_OK;
}

h3dmg_status h3dmg_get_pp(h3dmg_handle *h, int b, double *pp, int ld1, int ld2, int off) {
  h3dmg_status st = check_block(h, b);
  if (st) return st;
  if (!h->d_stage) return no_stage();
  const LevelDev &L = h->levs[b];
  if (h->lev_is64[b])
    launch_unpack_level<double>(h->stream, h->d_levs, b, 0, h->pp_buf, h->d_stage, L.ni, L.nj,
                                L.nk);
  else
    launch_unpack_level<float>(h->stream, h->d_levs, b, 0, h->pp_buf, h->d_stage, L.ni, L.nj,
                               L.nk);
  HIP_TRY(hipGetLastError());
  return stage_padded_to_host(h, b, pp, ld1, ld2, off);
}

namespace {

// the queued-ahead loop's device state (d_dec: DecideParams, DecideCtl,
// DecideBlock[nb], DecideState[nb], each part 256-byte aligned), ring, events
// and cycle graph
size_t dec_part(int part, int nb) {
  const size_t a = 256, p0 = 0, p1 = p0 + ((sizeof(DecideParams) + a - 1) & ~(a - 1));
  const size_t p2 = p1 + ((sizeof(DecideCtl) + a - 1) & ~(a - 1));
  const size_t p3 = p2 + ((sizeof(DecideBlock) * nb + a - 1) & ~(a - 1));
  const size_t p4 = p3 + ((sizeof(DecideState) * nb + a - 1) & ~(a - 1));
  const size_t offs[5] = {p0, p1, p2, p3, p4};
  return offs[part];
}

h3dmg_status decide_setup(h3dmg_handle *h) {
  const int nb = h->nb;
  std::lock_guard<std::mutex> lock(g_device_mutex);  // allocations, capture
  if (!h->d_dec) {
    void *p = nullptr;
    if (h3dmg_status st = dev_alloc(h, dec_part(4, nb), &p)) return st;
    h->d_dec = (char *)p;
  }
  if (!h->h_ring) {
    HIP_TRY(hipHostMalloc((void **)&h->h_ring, 2 * decide_slot_bytes(nb),
                          hipHostMallocMapped | hipHostMallocCoherent));
    HIP_TRY(hipHostGetDevicePointer((void **)&h->d_ring, h->h_ring, 0));
  }
  for (hipEvent_t &e : h->ev_dec)
    if (!e) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  if (!h->gexec_dec) {
    HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
    // every level-1 rhs of h3dmg_solve comes from the residual: red1
    enqueue_vcycle(h, true);
    enqueue_update_residual(h, true, false);
    // the partials' reduction and the stopping rule in one launch; its
    // arrival counter shares the 256-byte slot of DecideCtl, a line apart
    launch_finalize_decide(h->stream, h->d_tile_begin, nb, h->d_partials, h->d_stats,
                           (const DecideBlock *)(h->d_dec + dec_part(2, nb)),
                           (DecideState *)(h->d_dec + dec_part(3, nb)),
                           (DecideCtl *)(h->d_dec + dec_part(1, nb)), h->d_active, h->d_ring,
                           (const DecideParams *)(h->d_dec + dec_part(0, nb)),
                           (unsigned *)(h->d_dec + dec_part(1, nb) + 128));
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamEndCapture(h->stream, &g));
    h->graph_dec = g;
    HIP_TRY(hipGraphInstantiate(&h->gexec_dec, h->graph_dec, nullptr, nullptr, 0));
  }
  return H3DMG_OK;
}

// Cycle n in flight: is it likely the last one?  Each running block's
// relative residual (REF: the global rmax) after cycle n, extrapolated from
// its last two values at the rate between them, within 1.5x of its
// tolerance (a wrong "last" costs one wait for the host, a wrong "not last"
// one cycle of launches that do nothing).
bool likely_last(const h3dmg_handle *h, int n, const std::vector<double> &r1,
                 const std::vector<double> &r0, double g1, double g0) {
  const h3dmg_params &P = h->prm;
  const bool ref = P.criterion == H3DMG_CRIT_REF_RMAX;
  if (n < P.min_cycles || (ref && n < 2) || n < 3) return false;
  if (ref) return g0 > 0 && g1 * (g1 / g0) <= 1.5 * P.tol;
  for (int b = 0; b < h->nb; ++b) {
    if (!h->active[b]) continue;
    const double tol = h->blk[b].tol > 0 ? h->blk[b].tol : P.tol;
    if (!(r0[b] > 0) || r1[b] * (r1[b] / r0[b]) > 1.5 * tol) return false;
  }
  return true;
}

// h3dmg_solve's cycles with the next cycle queued before the host reads the
// last one's outcome: the stopping rule runs on the device (k_finalize_decide) and
// writes the flags the queued cycle reads, so the GPU does not wait for the
// host between cycles.  A cycle is queued ahead unless the residual history
// says the one in flight is likely the last; one queued after the last runs
// with every block inactive.  The host's bookkeeping (per block rel, rmax,
// cycles; the history; h_stats) is the device's outcome of each cycle.
h3dmg_status solve_queued(h3dmg_handle *h, double *hist, int hist_cap, int &it, double &worst,
                          bool &all_done, bool &bad) {
  const h3dmg_params &P = h->prm;
  const int nb = h->nb;
  if (h3dmg_status st = decide_setup(h)) return st;
  // the rule's inputs and the state after the initial residual
  std::vector<char> img(dec_part(4, nb), 0);
  DecideParams &dp = *(DecideParams *)(img.data() + dec_part(0, nb));
  dp.nb = nb;
  dp.ref = P.criterion == H3DMG_CRIT_REF_RMAX;
  dp.connected = h->connected ? 1 : 0;
  dp.min_cycles = P.min_cycles;
  dp.tol = P.tol;
  dp.rfac = P.dt;
  DecideBlock *db = (DecideBlock *)(img.data() + dec_part(2, nb));
  DecideState *ds = (DecideState *)(img.data() + dec_part(3, nb));
  for (int b = 0; b < nb; ++b) {
    const BlockInfo &B = h->blk[b];
    db[b].fnorm = B.fnorm;
    db[b].tol = B.tol > 0 ? B.tol : P.tol;
    db[b].vol = B.spec.dx * B.spec.dy * B.spec.dz;
    db[b].enabled = B.enabled ? 1 : 0;
    ds[b].rel = B.rel;
    ds[b].rmax = B.rmax;
  }
  {
    std::lock_guard<std::mutex> lock(g_device_mutex);  // a pool's staging
    HIP_TRY(h2d(h, h->d_dec, img.data(), img.size()));
  }
  if (h3dmg_status st = push_active(h)) return st;
  std::vector<double> r1(nb), r0(nb, 0.0);
  for (int b = 0; b < nb; ++b) r1[b] = h->blk[b].rel;
  double g1 = worst, g0 = 0.0;
  int queued = 0;
  auto queue = [&]() -> h3dmg_status {
    ++queued;
    HIP_TRY(hipGraphLaunch(h->gexec_dec, h->stream));
    HIP_TRY(hipEventRecord(h->ev_dec[queued & 1], h->stream));
    return H3DMG_OK;
  };
  if (h3dmg_status st = queue()) return st;
  const size_t slot = decide_slot_bytes(nb);
  while (true) {
    const int n = it + 1;  // the cycle whose outcome is read next
    if (queued == n && queued < P.maxcy && !likely_last(h, n, r1, r0, g1, g0))
      if (h3dmg_status st = queue()) return st;
    HIP_TRY(hipEventSynchronize(h->ev_dec[n & 1]));
    const char *sl = h->h_ring + (size_t)(n & 1) * slot;
    const DecideCtl c = *(const DecideCtl *)sl;
    const DecideState *rs = (const DecideState *)(sl + sizeof(DecideCtl));
    const RedPartial *rp = (const RedPartial *)(rs + nb);
    const int *ra = (const int *)(rp + nb);
    if (c.it != n) return fail(H3DMG_ERR_DEVICE, "cycle %d outcome reads cycle %d", n, c.it);
    it = n;
    for (int b = 0; b < nb; ++b) {
      BlockInfo &B = h->blk[b];
      if (h->active[b]) {
        r0[b] = r1[b];
        B.rel = rs[b].rel;
        B.rmax = rs[b].rmax;
        B.cycles = it;
        r1[b] = B.rel;
      }
      h->h_stats[b] = rp[b];
      h->active[b] = ra[b];
    }
    worst = c.worst;
    g0 = g1;
    g1 = worst;
    all_done = c.all_done != 0;
    bad = c.bad != 0;
    if (hist && it < hist_cap) hist[it] = worst;
    if (all_done || bad || it >= P.maxcy) break;
    if (queued == it)
      if (h3dmg_status st = queue()) return st;
  }
  return H3DMG_OK;
}

}  // namespace

h3dmg_status h3dmg_solve(h3dmg_handle *h, double *hist, int hist_cap, int *ncycles,
                         double *final_res) {
  if (!h) return fail(H3DMG_ERR_ARG, "null handle");
  if (!h->rhs_ready) return fail(H3DMG_ERR_ARG, "rhs not set");
  if (h3dmg_status st = check_density(h)) return st;
  const h3dmg_params &P = h->prm;
  const bool ref = P.criterion == H3DMG_CRIT_REF_RMAX;
  const double rfac_base = P.dt;
  reset_active(h);
  h3dmg_status st = push_active(h);
  if (st) return st;
  for (auto &B : h->blk) {
    B.cycles = 0;
    B.rel = 0.0;
  }
  enqueue_update_residual(h, false);
  HIP_TRY(hipGetLastError());
  st = read_stats(h);
  if (st) return st;

  auto measure = [&](int it, double &worst, bool &all_done, bool &bad) {
    worst = 0.0;
    all_done = true;
    bad = false;
    double gmax = 0.0;
    // blocks joined by shared faces are one domain: one relative residual
    double grel = 0.0;
    if (h->connected) {
      double ss = 0.0, ff = 0.0;
      for (int b = 0; b < h->nb; ++b)
        if (h->blk[b].enabled) {
          ss += h->h_stats[b].ss;
          ff += h->blk[b].fnorm * h->blk[b].fnorm;
        }
      grel = ff > 0 ? std::sqrt(ss) / std::sqrt(ff) : std::sqrt(ss);
    }
    for (int b = 0; b < h->nb; ++b) {
      BlockInfo &B = h->blk[b];
      if (!B.enabled) continue;
      const double nrm = std::sqrt(h->h_stats[b].ss);
      const double vol = B.spec.dx * B.spec.dy * B.spec.dz;
      if (h->active[b]) {
        B.rel = h->connected ? grel : B.fnorm > 0 ? nrm / B.fnorm : nrm;
        B.rmax = h->h_stats[b].amax * rfac_base * vol;
        B.cycles = it;
      }
      if (!std::isfinite(nrm)) bad = true;
      worst = std::max(worst, B.rel);
      gmax = std::max(gmax, B.rmax);
    }
    for (int b = 0; b < h->nb; ++b) {
      BlockInfo &B = h->blk[b];
      bool done;
      if (!B.enabled)
        done = true;
      else if (ref)
        done = it >= std::max(P.min_cycles, 2) && gmax < P.tol;  // newsolv_mg.for:57, global
      else
        done = it >= P.min_cycles && B.rel <= (B.tol > 0 ? B.tol : P.tol);
      if (done) h->active[b] = 0;
      if (h->active[b]) all_done = false;
    }
    if (ref) worst = gmax;
  };

  double worst;
  bool all_done, bad;
  measure(0, worst, all_done, bad);
  if (hist && hist_cap > 0) hist[0] = worst;
  if (P.min_cycles > 0 || ref) {
    reset_active(h);
    all_done = std::none_of(h->active.begin(), h->active.end(), [](int a) { return a != 0; });
  }
  int it = 0;
  // (a pooled handle allocates nothing: it keeps the host's loop)
  if (P.use_graph && !h->pool && !all_done && P.maxcy > 0 && !bad) {
    st = solve_queued(h, hist, hist_cap, it, worst, all_done, bad);
    if (st) return st;
  }
  while (!all_done && it < P.maxcy && !bad) {
    st = push_active(h);
    if (st) return st;
    st = run_cycle(h);
    if (st) return st;
    st = read_stats(h);
    if (st) return st;
    ++it;
    measure(it, worst, all_done, bad);
    if (hist && it < hist_cap) hist[it] = worst;
  }
  if (ncycles) *ncycles = it;
  if (final_res) *final_res = worst;
  if (bad) return fail(H3DMG_ERR_DIVERGED, "non-finite residual after %d cycles", it);
  if (ref && worst > 10.0) return fail(H3DMG_ERR_DIVERGED, "rmax %g > 10", worst);
  if (!all_done) return fail(H3DMG_ERR_NOT_CONVERGED, "not converged in %d cycles (%g)", it, worst);
  return H3DMG_OK;
}

h3dmg_status h3dmg_set_block_tol(h3dmg_handle *h, int b, double tol) {
  h3dmg_status st = check_block(h, b);
  if (st) return st;
  if (!(tol >= 0.0)) return fail(H3DMG_ERR_ARG, "negative tolerance");
  if (h->connected)
    return fail(H3DMG_ERR_UNSUPPORTED, "blocks with neighbours stop together (one tolerance)");
  h->blk[b].tol = tol;
  return H3DMG_OK;
}

h3dmg_status h3dmg_set_block_enabled(h3dmg_handle *h, int b, int enabled) {
  h3dmg_status st = check_block(h, b);
  if (st) return st;
  if (h->blk[b].enabled == (enabled != 0)) return H3DMG_OK;
  if (h->connected)
    return fail(H3DMG_ERR_UNSUPPORTED, "blocks with neighbours are solved together");
  h->blk[b].enabled = enabled != 0;
  if (!enabled) h->running[b] = 0;
  // the velocity loop's block set (and its captured step, which projects
  // the blocks chosen at h3dmg_velocity_begin) is no longer this one
  if (h->vel_ready || h->vgexec) {
    std::lock_guard<std::mutex> lock(g_device_mutex);
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (h->vgexec) (void)hipGraphExecDestroy(h->vgexec);
    if (h->vgraph) (void)hipGraphDestroy(h->vgraph);
    h->vgexec = nullptr;
    h->vgraph = nullptr;
    h->vel_ready = false;
  }
  return H3DMG_OK;
}

h3dmg_status h3dmg_get_block_state(const h3dmg_handle *h, int b, h3dmg_block_state *out) {
  h3dmg_status st = check_block(h, b);
  if (st) return st;
  if (!out) return fail(H3DMG_ERR_ARG, "null state");
  const BlockInfo &B = h->blk[b];
  out->running = h->running[b];
  out->cycles = B.cycles;
  out->status = h->outcome[b];
  out->enabled = B.enabled ? 1 : 0;
  out->fnorm = B.fnorm;
  out->rel = B.rel;
  out->tol = B.tol;
  return H3DMG_OK;
}

h3dmg_status h3dmg_set_block_state(h3dmg_handle *h, int b, const h3dmg_block_state *in) {
  h3dmg_status st = check_block(h, b);
  if (st) return st;
  if (!in) return fail(H3DMG_ERR_ARG, "null state");
  if (!(in->fnorm >= 0.0) || !(in->tol >= 0.0) || in->cycles < 0)
    return fail(H3DMG_ERR_ARG, "invalid block state");
  if (h->connected) return fail(H3DMG_ERR_UNSUPPORTED, "blocks with neighbours are solved together");
  BlockInfo &B = h->blk[b];
  B.cycles = in->cycles;
  B.fnorm = in->fnorm;
  B.rel = in->rel;
  B.tol = in->tol;
  B.enabled = in->enabled != 0;
  h->running[b] = (in->running && B.enabled) ? 1 : 0;
  h->outcome[b] = in->status;
  h->rhs_ready = true;
  return H3DMG_OK;
}

namespace {
// tolerance and stopping rule of one block's stepped solve (the REL_L2 rule
// of h3dmg_solve): returns the outcome once it stops, -1 while it runs
int stepped_outcome(const h3dmg_handle *h, int b, double nrm) {
  const h3dmg_params &P = h->prm;
  const BlockInfo &B = h->blk[b];
  if (!std::isfinite(nrm)) return H3DMG_ERR_DIVERGED;
  if (B.cycles >= P.min_cycles && B.rel <= (B.tol > 0 ? B.tol : P.tol)) return H3DMG_OK;
  if (B.cycles >= P.maxcy) return H3DMG_ERR_NOT_CONVERGED;
  return -1;
}

h3dmg_status check_stepped(const h3dmg_handle *h) {
  if (!h) return fail(H3DMG_ERR_ARG, "null handle");
  if (h->prm.criterion != H3DMG_CRIT_REL_L2)
    return fail(H3DMG_ERR_UNSUPPORTED, "stepped solves stop each block on its own (REL_L2 only)");
  if (h->connected)
    return fail(H3DMG_ERR_UNSUPPORTED, "blocks with neighbours are solved together (h3dmg_solve)");
  return check_density(h);
}
}  // namespace

h3dmg_status h3dmg_solve_begin(h3dmg_handle *h, const int *blocks, int n) {
  h3dmg_status st = check_stepped(h);
  if (st) return st;
  if (!blocks || n < 1) return fail(H3DMG_ERR_ARG, "no blocks to begin");
  std::vector<char> pick(h->nb, 0);
  for (int i = 0; i < n; ++i) {
    const int b = blocks[i];
    if ((st = check_block(h, b))) return st;
    if (!h->blk[b].enabled) return fail(H3DMG_ERR_ARG, "block %d is disabled", b);
    if (pick[b]) return fail(H3DMG_ERR_ARG, "block %d listed twice", b);
    pick[b] = 1;
  }
  // r = f - A p of the listed blocks only: the running blocks' residuals
  // (the next V-cycle's right-hand sides) stay as they are
  for (int b = 0; b < h->nb; ++b) h->active[b] = pick[b];
  if ((st = push_active(h))) return st;
  enqueue_update_residual(h, false);
  HIP_TRY(hipGetLastError());
  if ((st = read_stats(h))) return st;
  for (int b = 0; b < h->nb; ++b) {
    if (!pick[b]) continue;
    BlockInfo &B = h->blk[b];
    const double nrm = std::sqrt(h->h_stats[b].ss);
    B.cycles = 0;
    B.rel = B.fnorm > 0 ? nrm / B.fnorm : nrm;
    B.rmax = 0.0;
    h->outcome[b] = stepped_outcome(h, b, nrm);
    h->running[b] = h->outcome[b] < 0;
  }
  return H3DMG_OK;
}

h3dmg_status h3dmg_solve_step(h3dmg_handle *h, int *nrunning) {
  h3dmg_status st = check_stepped(h);
  if (st) return st;
  int count = 0;
  for (int b = 0; b < h->nb; ++b) {
    h->active[b] = h->running[b] && h->blk[b].enabled;
    count += h->active[b];
  }
  if (count) {
    if ((st = push_active(h))) return st;
    if ((st = run_cycle(h))) return st;
    if ((st = read_stats(h))) return st;
    count = 0;
    for (int b = 0; b < h->nb; ++b) {
      if (!h->active[b]) continue;
      BlockInfo &B = h->blk[b];
      const double nrm = std::sqrt(h->h_stats[b].ss);
      B.cycles += 1;
      B.rel = B.fnorm > 0 ? nrm / B.fnorm : nrm;
      h->outcome[b] = stepped_outcome(h, b, nrm);
      h->running[b] = h->outcome[b] < 0;
      count += h->running[b];
    }
  }
  if (nrunning) *nrunning = count;
  return H3DMG_OK;
}

h3dmg_status h3dmg_project_device(h3dmg_handle *h, int b, const double *d_ustar,
                                  const double *d_vstar, const double *d_wstar, double *d_u,
                                  double *d_v, double *d_w, double *d_su, double alfapr,
                                  double *rmax) {
  h3dmg_status st = check_block(h, b);
  if (st) return st;
  if (h->connected)
    return fail(H3DMG_ERR_UNSUPPORTED,
                "one block of blocks with neighbours: use the velocity loop (its faces are "
                "shared)");
  if (!d_ustar || !d_vstar || !d_wstar || !d_u || !d_v || !d_w)
    return fail(H3DMG_ERR_ARG, "null velocity array");
  if (h->differencing == 2)
    return fail(H3DMG_ERR_UNSUPPORTED,
                "differencing = 2: the projection of device arrays has no outer faces (use the "
                "velocity loop with h3dmg_set_velocity)");
  const FineDev &F = h->fine[b];
  const BlockInfo &B = h->blk[b];
  const double dt = h->prm.dt;
  if (!(dt > 0)) return fail(H3DMG_ERR_ARG, "projection needs params.dt > 0");
  const double dta = dt * alfapr;
  const size_t ni = F.ni, nj = F.nj, nk = F.nk;
  launch_calvel(h->stream, h->d_fine + b, 0, d_ustar, d_u, dta, B.spec.dx, (ni + 1) * nj * nk);
  launch_calvel(h->stream, h->d_fine + b, 1, d_vstar, d_v, dta, B.spec.dy, ni * (nj + 1) * nk);
  launch_calvel(h->stream, h->d_fine + b, 2, d_wstar, d_w, dta, B.spec.dz, ni * nj * (nk + 1));
  RedPartial *parts = h->d_proj + h->proj_begin[b];
  const int groups = launch_calmas(h->stream, h->d_fine + b, d_u, d_v, d_w, d_su, 0, dt, B.spec.dx,
                                   B.spec.dy, B.spec.dz, parts, ni * nj * nk);
  HIP_TRY(hipGetLastError());
  if (rmax) {
    std::vector<RedPartial> part(groups);
    HIP_TRY(hipMemcpyAsync(part.data(), parts, groups * sizeof(RedPartial),
                           hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    double m = 0.0;
    for (const RedPartial &r : part) m = std::max(m, r.amax);
    *rmax = m;
  } else {
    HIP_TRY(hipStreamSynchronize(h->stream));
  }
  return H3DMG_OK;
}

h3dmg_status h3dmg_project_device_noalloc(h3dmg_handle *h, int b, const double *d_ustar,
                                          const double *d_vstar, const double *d_wstar,
                                          double *d_u, double *d_v, double *d_w, double *d_su,
                                          double alfapr, double *d_rmax) {
  h3dmg_status st = check_block(h, b);
  if (st) return st;
  if (h->connected)
    return fail(H3DMG_ERR_UNSUPPORTED,
                "one block of blocks with neighbours: use the velocity loop (its faces are "
                "shared)");
  if (!d_ustar || !d_vstar || !d_wstar || !d_u || !d_v || !d_w)
    return fail(H3DMG_ERR_ARG, "null velocity array");
  if (h->differencing == 2)
    return fail(H3DMG_ERR_UNSUPPORTED,
                "differencing = 2: the projection of device arrays has no outer faces (use the "
                "velocity loop with h3dmg_set_velocity)");
  const FineDev &F = h->fine[b];
  const BlockInfo &B = h->blk[b];
  const double dt = h->prm.dt;
  if (!(dt > 0)) return fail(H3DMG_ERR_ARG, "projection needs params.dt > 0");
  const size_t ni = F.ni, nj = F.nj, nk = F.nk;
  const int groups = calmas_groups(ni * nj * nk);
  if (groups > NOALLOC_PARTIALS)
    return fail(H3DMG_ERR_ARG, "block %d: %d calmas partials exceed the %d the scratch holds", b,
                groups, NOALLOC_PARTIALS);
  RedPartial *parts = nullptr;
  st = scratch_on(h->device, &parts);
  if (st) return st;
  const double dta = dt * alfapr;
  launch_calvel(h->stream, h->d_fine + b, 0, d_ustar, d_u, dta, B.spec.dx, (ni + 1) * nj * nk);
  launch_calvel(h->stream, h->d_fine + b, 1, d_vstar, d_v, dta, B.spec.dy, ni * (nj + 1) * nk);
  launch_calvel(h->stream, h->d_fine + b, 2, d_wstar, d_w, dta, B.spec.dz, ni * nj * (nk + 1));
  // every one of the `groups` partials is written before it is read: nothing
  // in the scratch needs clearing
  launch_calmas(h->stream, h->d_fine + b, d_u, d_v, d_w, d_su, 0, dt, B.spec.dx, B.spec.dy,
                B.spec.dz, parts, ni * nj * nk);
  if (d_rmax) launch_partials_max(h->stream, parts, groups, d_rmax);
  HIP_TRY(hipGetLastError());
  return H3DMG_OK;
}

// Face arrays in the reference's layout (press.for:90-131): u(i,j,k) is the
// face between cells i and i+1, so dense face f (between 0-based cells f-1
// and f) is Fortran index isp-1+f, 0-based off-1+f.
namespace {
h3dmg_status alloc_velocity(h3dmg_handle *h) {
  if (h->vel_store) return H3DMG_OK;
  if (h->pool) return fail(H3DMG_ERR_UNSUPPORTED, "a pooled handle has no velocity loop");
  h->vel_off.assign(9 * (size_t)h->nb, 0);
  size_t off = 0;
  for (int b = 0; b < h->nb; ++b) {
    const h3dmg_block &S = h->blk[b].spec;
    const size_t n[3] = {(size_t)(S.ni + 1) * S.nj * S.nk, (size_t)S.ni * (S.nj + 1) * S.nk,
                         (size_t)S.ni * S.nj * (S.nk + 1)};
    const size_t planes[3] = {2 * (size_t)S.nj * S.nk, 2 * (size_t)S.ni * S.nk,
                              2 * (size_t)S.ni * S.nj};
    for (int q = 0; q < 9; ++q) {
      h->vel_off[9 * b + q] = off;
      off += ((q < 6 ? n[q % 3] : planes[q - 6]) + 31) & ~(size_t)31;
    }
  }
  hipError_t e = hipMalloc((void **)&h->vel_store, off * sizeof(double));
  if (e != hipSuccess) {
    h->vel_store = nullptr;
    return fail(H3DMG_ERR_NOMEM, "velocity arrays (%zu bytes): %s", off * 8, hipGetErrorString(e));
  }
  h->bytes += off * sizeof(double);
  h->vel_set.assign(h->nb, 0);
  h->vel_outer.assign(h->nb, 0);
  return H3DMG_OK;
}

// 0-based flat index of Fortran element (off+x, off+y, off+z) shifted by one
// face on `axis` (x, y, z interior-relative, face f maps to f-1 on its axis)
inline size_t face_index(int axis, int f, int a, int c, int ld1, int ld2, int off) {
  int x, y, z;
  if (axis == 0) {
    x = off - 1 + f; y = off + a; z = off + c;
  } else if (axis == 1) {
    x = off + a; y = off - 1 + f; z = off + c;
  } else {
    x = off + a; y = off + c; z = off - 1 + f;
  }
  return (size_t)x + (size_t)ld1 * ((size_t)y + (size_t)ld2 * (size_t)z);
}

// dense index of face f of `axis` with the other two coordinates (a, c) in
// (i, j, k) order minus the face axis
inline size_t dense_face(int axis, int f, int a, int c, const h3dmg_block &S) {
  if (axis == 0) return ((size_t)c * S.nj + a) * (S.ni + 1) + f;        // (k, j, f)
  if (axis == 1) return ((size_t)c * (S.nj + 1) + f) * S.ni + a;        // (k, f, i)
  return ((size_t)f * S.nj + c) * S.ni + a;                             // (f, j, i)
}

h3dmg_status check_face_layout(const h3dmg_block &S, int ld1, int ld2, int off) {
  if (off < 1 || ld1 < S.ni + off + 1 || ld2 < S.nj + off + 1)
    return fail(H3DMG_ERR_ARG, "bad host face array description (needs one ghost layer)");
  return H3DMG_OK;
}
}  // namespace

h3dmg_status h3dmg_set_velocity(h3dmg_handle *h, int b, const double *ustar, const double *vstar,
                                const double *wstar, const double *u, const double *v,
                                const double *w, int ld1, int ld2, int off) {
  h3dmg_status st = check_block(h, b);
  if (st) return st;
  if (!ustar || !vstar || !wstar) return fail(H3DMG_ERR_ARG, "null u*, v* or w*");
  const h3dmg_block &S = h->blk[b].spec;
  st = check_face_layout(S, ld1, ld2, off);
  if (st) return st;
  st = alloc_velocity(h);
  if (st) return st;
  const double *star[3] = {ustar, vstar, wstar}, *cur[3] = {u, v, w};
  const int n[3] = {S.ni, S.nj, S.nk};
  for (int axis = 0; axis < 3; ++axis) {
    const bool per = (h->blk[b].per >> axis) & 1;
    // the two axes that are not `axis`, in (i, j, k) order
    const int na = axis == 0 ? S.nj : S.ni, nc = axis == 2 ? S.nj : S.nk;
    std::vector<double> dense((size_t)(n[axis] + 1) * na * nc);
    for (int c = 0; c < nc; ++c)
      for (int a = 0; a < na; ++a)
        for (int f = 0; f <= n[axis]; ++f) {
          // faces calvel corrects take u*, the others keep the current u
          const bool corrected = (f > 0 && f < n[axis]) || per;
          const double *src = corrected || !cur[axis] ? star[axis] : cur[axis];
          dense[dense_face(axis, f, a, c, S)] = src[face_index(axis, f, a, c, ld1, ld2, off)];
        }
    HIP_TRY(hipMemcpyAsync(h->vel_store + h->vel_off[9 * b + axis], dense.data(),
                           dense.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (off >= 2) {
      // faces -1 and n+1 (the host's current u there, as calmas reads it)
      std::vector<double> planes(2 * (size_t)na * nc);
      const double *src = cur[axis] ? cur[axis] : star[axis];
      for (int side = 0; side < 2; ++side)
        for (int c = 0; c < nc; ++c)
          for (int a = 0; a < na; ++a)
            planes[((size_t)side * nc + c) * na + a] =
                src[face_index(axis, side ? n[axis] + 1 : -1, a, c, ld1, ld2, off)];
      HIP_TRY(hipMemcpyAsync(h->vel_store + h->vel_off[9 * b + 6 + axis], planes.data(),
                             planes.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
      HIP_TRY(hipStreamSynchronize(h->stream));
    }
  }
  h->vel_set[b] = 1;
  h->vel_outer[b] = off >= 2 ? 1 : 0;
  return H3DMG_OK;
}

namespace {
// u, v, w of block b into host Fortran arrays: every face of the interior
// rows, or (calvel_only) just the faces calvel computes -- isu..ieu and the
// like (press.for:90-131 with localparameters.for:106-192): faces 1..n of an
// axis whose far face is open (a neighbour or a periodic wrap), 1..n-1
// against a far wall; face 0 is never calvel's (exchange(1..3) fills it).
h3dmg_status velocity_to_host(h3dmg_handle *h, int b, double *u, double *v, double *w, int ld1,
                              int ld2, int off, bool calvel_only) {
  h3dmg_status st = check_block(h, b);
  if (st) return st;
  if (!u || !v || !w) return fail(H3DMG_ERR_ARG, "null u, v or w");
  const h3dmg_block &S = h->blk[b].spec;
  st = check_face_layout(S, ld1, ld2, off);
  if (st) return st;
  if (!h->vel_store || !h->vel_set[b])
    return fail(H3DMG_ERR_ARG, "block %d: no velocity on the device", b);
  double *dst[3] = {u, v, w};
  const int n[3] = {S.ni, S.nj, S.nk};
  for (int axis = 0; axis < 3; ++axis) {
    const int na = axis == 0 ? S.nj : S.ni, nc = axis == 2 ? S.nj : S.nk;
    const bool far_open = S.bc[2 * axis + 1] != H3DMG_BC_WALL;
    const int f0 = calvel_only ? 1 : 0;
    const int f1 = calvel_only && !far_open ? n[axis] - 1 : n[axis];
    std::vector<double> dense((size_t)(n[axis] + 1) * na * nc);
    HIP_TRY(hipMemcpyAsync(dense.data(), h->vel_store + h->vel_off[9 * b + 3 + axis],
                           dense.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (int c = 0; c < nc; ++c)
      for (int a = 0; a < na; ++a)
        for (int f = f0; f <= f1; ++f)
          dst[axis][face_index(axis, f, a, c, ld1, ld2, off)] = dense[dense_face(axis, f, a, c, S)];
  }
  return H3DMG_OK;
}
}  // namespace

h3dmg_status h3dmg_get_velocity(h3dmg_handle *h, int b, double *u, double *v, double *w, int ld1,
                                int ld2, int off) {
  return velocity_to_host(h, b, u, v, w, ld1, ld2, off, false);
}

h3dmg_status h3dmg_get_velocity_calvel(h3dmg_handle *h, int b, double *u, double *v, double *w,
                                       int ld1, int ld2, int off) {
  return velocity_to_host(h, b, u, v, w, ld1, ld2, off, true);
}

h3dmg_status h3dmg_get_su(h3dmg_handle *h, int b, double *su, int ld1, int ld2, int off) {
  h3dmg_status st = check_block(h, b);
  if (st) return st;
  if (!h->d_stage) return no_stage();
  const h3dmg_block &S = h->blk[b].spec;
  if (!su || off < 0 || ld1 < S.ni + off || ld2 < S.nj + off)
    return fail(H3DMG_ERR_ARG, "bad host array description");
  const FineDev &F = h->fine[b];
  launch_unpack_fine(h->stream, h->d_fine + b, 1, h->d_stage, F.ni, F.nj, F.nk);
  HIP_TRY(hipGetLastError());
  std::vector<double> dense((size_t)S.ni * S.nj * S.nk);
  HIP_TRY(hipMemcpyAsync(dense.data(), h->d_stage, dense.size() * 8, hipMemcpyDeviceToHost,
                         h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  for (int k = 0; k < S.nk; ++k)
    for (int j = 0; j < S.nj; ++j)
      std::memcpy(su + (size_t)off + (size_t)ld1 * ((size_t)(j + off) + (size_t)ld2 * (k + off)),
                  &dense[((size_t)k * S.nj + j) * S.ni], sizeof(double) * S.ni);
  return H3DMG_OK;
}

// Variable-density operator: coefficient arrays and dens_mg for every level
// descriptor, allocated on the first h3dmg_set_density*.
namespace {
h3dmg_status enable_lsm(h3dmg_handle *h) {
  if (h->lsm) return H3DMG_OK;
  if (h->pool)
    return fail(H3DMG_ERR_UNSUPPORTED, "a pooled handle has no variable-density operator");
  if (h->connected)
    return fail(H3DMG_ERR_UNSUPPORTED,
                "the variable-density operator on blocks with neighbours is not built");
  if (is_mixed(h) || h->prm.coarse != H3DMG_COARSE_SMOOTH)
    return fail(H3DMG_ERR_UNSUPPORTED,
                "the variable-density operator runs in FP64 with the reference coarsest level");
  if (h->differencing == 2)
    return fail(H3DMG_ERR_UNSUPPORTED,
                "the variable-density operator with differencing = 2 (the reference's 4th-order "
                "calmas writes su, not sup)");
  std::lock_guard<std::mutex> lock(g_device_mutex);
  HIP_TRY(hipStreamSynchronize(h->stream));
  drop_graph(h);
  size_t off = 0;
  h->lsm_dens.assign(h->levs.size(), 0);
  std::vector<size_t> coef_off(h->levs.size());
  for (size_t d = 0; d < h->levs.size(); ++d) {
    const LevelDev &L = h->levs[d];
    const size_t half = (size_t)L.H * L.nj * L.nk;
    coef_off[d] = off;
    off += (14 * half + 31) & ~(size_t)31;
    h->lsm_dens[d] = off;
    off += ((size_t)L.ni * L.nj * L.nk + 31) & ~(size_t)31;
  }
  hipError_t e = hipMalloc((void **)&h->lsm_store, off * sizeof(double));
  if (e != hipSuccess) {
    h->lsm_store = nullptr;
    return fail(H3DMG_ERR_NOMEM, "density arrays (%zu bytes): %s", off * 8, hipGetErrorString(e));
  }
  h->bytes += off * sizeof(double);
  for (size_t d = 0; d < h->levs.size(); ++d) {
    LevelDev &L = h->levs[d];
    L.lsm = h->lsm_store + coef_off[d];
    L.lsm_half = (unsigned long long)L.H * L.nj * L.nk;
  }
  HIP_TRY(hipMemcpy(h->d_levs, h->levs.data(), h->levs.size() * sizeof(LevelDev),
                    hipMemcpyHostToDevice));
  for (int b = 0; b < h->nb; ++b) {
    FineDev &F = h->fine[b];
    F.dens = h->lsm_store + h->lsm_dens[h->blk[b].desc[1]];
    F.dx = h->blk[b].spec.dx;
    F.dy = h->blk[b].spec.dy;
    F.dz = h->blk[b].spec.dz;
  }
  HIP_TRY(hipMemcpy(h->d_fine, h->fine.data(), h->nb * sizeof(FineDev), hipMemcpyHostToDevice));
  h->lsm_set.assign(h->nb, 0);
  h->lsm = true;
  // the fused sweep and the LDS tail hold no coefficients: both off
  h->fused_max = 0;
  for (int g = 1; g <= h->Lmax; ++g) {
    h3dmg_status st = upload_relax(h, g);
    if (st) return st;
  }
  h->tail_req = h->Lmax + 1;
  return upload_tail(h);
}

// dens_mg of every level of block b from its level-1 density (already in
// place), then the coefficients of every level (coeff, mgsolver.for:72-210)
h3dmg_status build_lsm_block(h3dmg_handle *h, int b) {
  const BlockInfo &B = h->blk[b];
  for (int g = 2; g <= B.L; ++g) {
    const int df = B.desc[g - 1], dc = B.desc[g];
    const LevelDev &Lf = h->levs[df];
    launch_lsm_dens_restrict(h->stream, h->lsm_store + h->lsm_dens[df], Lf.ni, Lf.nj, Lf.nk,
                             h->lsm_store + h->lsm_dens[dc]);
  }
  int walls = 0;
  for (int f = 0; f < 6; ++f)
    if (B.spec.bc[f] != H3DMG_BC_PERIODIC) walls |= 1 << f;
  for (int g = 1; g <= B.L; ++g) {
    const int d = B.desc[g];
    const LevelDev &L = h->levs[d];
    launch_lsm_coef(h->stream, h->d_levs + d, h->lsm_store + h->lsm_dens[d], L.ni, L.nj, L.nk,
                    B.spec.dx, B.spec.dy, B.spec.dz, walls);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->lsm_set[b] = 1;
  return H3DMG_OK;
}
}  // namespace

h3dmg_status h3dmg_set_density_device(h3dmg_handle *h, int b, const double *d_dens) {
  h3dmg_status st = check_block(h, b);
  if (st) return st;
  if (!d_dens) return fail(H3DMG_ERR_ARG, "null density");
  st = enable_lsm(h);
  if (st) return st;
  const h3dmg_block &S = h->blk[b].spec;
  HIP_TRY(hipMemcpyAsync(h->lsm_store + h->lsm_dens[h->blk[b].desc[1]], d_dens,
                         (size_t)S.ni * S.nj * S.nk * sizeof(double), hipMemcpyDeviceToDevice,
                         h->stream));
  return build_lsm_block(h, b);
}

h3dmg_status h3dmg_set_density(h3dmg_handle *h, int b, const double *dens, int ld1, int ld2,
                               int off) {
  h3dmg_status st = check_block(h, b);
  if (st) return st;
  st = host_to_stage(h, b, dens, ld1, ld2, off);
  if (st) return st;
  return h3dmg_set_density_device(h, b, h->d_stage);
}

h3dmg_status h3dmg_set_differencing(h3dmg_handle *h, int differencing) {
  if (!h) return fail(H3DMG_ERR_ARG, "null handle");
  if (differencing != 1 && differencing != 2)
    return fail(H3DMG_ERR_ARG, "differencing %d (1: 2nd order, 2: 4th order)", differencing);
  if (differencing == 2 && h->lsm)
    return fail(H3DMG_ERR_UNSUPPORTED,
                "differencing = 2 with the variable-density operator (the reference's 4th-order "
                "calmas writes su, not sup)");
  if (h->differencing == differencing) return H3DMG_OK;
  h->differencing = differencing;
  // the captured velocity step launches the other calmas
  std::lock_guard<std::mutex> lock(g_device_mutex);
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (h->vgexec) (void)hipGraphExecDestroy(h->vgexec);
  if (h->vgraph) (void)hipGraphDestroy(h->vgraph);
  h->vgexec = nullptr;
  h->vgraph = nullptr;
  h->vel_ready = false;
  return H3DMG_OK;
}

h3dmg_status h3dmg_set_dt(h3dmg_handle *h, double dt) {
  if (!h) return fail(H3DMG_ERR_ARG, "null handle");
  if (!(dt > 0)) return fail(H3DMG_ERR_ARG, "dt must be > 0");
  h->prm.dt = dt;  // the velocity step's graph is re-captured on the next begin
  return H3DMG_OK;
}

h3dmg_status h3dmg_velocity_begin(h3dmg_handle *h, const double *const *d_ustar,
                                  const double *const *d_vstar, const double *const *d_wstar,
                                  double *const *d_u, double *const *d_v, double *const *d_w,
                                  double alfapr, double *rmax) {
  if (!h) return fail(H3DMG_ERR_ARG, "null handle");
  const bool own = !d_ustar && !d_vstar && !d_wstar && !d_u && !d_v && !d_w;
  if (!own && (!d_ustar || !d_vstar || !d_wstar || !d_u || !d_v || !d_w))
    return fail(H3DMG_ERR_ARG, "null velocity array list");
  if (h->pool) return fail(H3DMG_ERR_UNSUPPORTED, "a pooled handle has no velocity loop");
  if (!(h->prm.dt > 0)) return fail(H3DMG_ERR_ARG, "the velocity form needs params.dt > 0");
  if (h3dmg_status st = check_density(h)) return st;
  if (h->differencing == 2) {
    // the 4th-order stencil reads faces -1 and n+1: at walls they are host
    // values that only h3dmg_set_velocity brings in
    if (!own)
      return fail(H3DMG_ERR_UNSUPPORTED,
                  "differencing = 2 takes the velocity from h3dmg_set_velocity (host arrays with "
                  "two ghost layers)");
    for (int b = 0; b < h->nb; ++b)
      if (h->blk[b].enabled && (!h->vel_store || !h->vel_outer[b]))
        return fail(H3DMG_ERR_ARG,
                    "block %d: differencing = 2 needs faces -1 and n+1 (host arrays with off >= 2)",
                    b);
  }
  bool same = h->vel_ready && h->vel_dta == h->prm.dt * alfapr && h->vel_dt == h->prm.dt;
  for (int b = 0; b < h->nb; ++b) {
    if (!h->blk[b].enabled) continue;
    h3dmg_handle::VelBlock V;
    if (own) {
      if (!h->vel_store || !h->vel_set[b])
        return fail(H3DMG_ERR_ARG, "block %d: h3dmg_set_velocity has not run", b);
      const size_t *o = &h->vel_off[9 * b];
      double *S = h->vel_store;
      V = h3dmg_handle::VelBlock{S + o[0], S + o[1], S + o[2], S + o[3], S + o[4], S + o[5]};
    } else {
      if (!d_ustar[b] || !d_vstar[b] || !d_wstar[b] || !d_u[b] || !d_v[b] || !d_w[b])
        return fail(H3DMG_ERR_ARG, "block %d: null velocity array", b);
      V = h3dmg_handle::VelBlock{d_ustar[b], d_vstar[b], d_wstar[b], d_u[b], d_v[b], d_w[b]};
    }
    const h3dmg_handle::VelBlock &O = h->vel[b];
    same = same && O.us == V.us && O.vs == V.vs && O.ws == V.ws && O.u == V.u && O.v == V.v &&
           O.w == V.w;
    h->vel[b] = V;
  }
  if (!same) {
    std::lock_guard<std::mutex> lock(g_device_mutex);
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (h->vgexec) (void)hipGraphExecDestroy(h->vgexec);
    if (h->vgraph) (void)hipGraphDestroy(h->vgraph);
    h->vgexec = nullptr;
    h->vgraph = nullptr;
  }
  h->vel_dta = h->prm.dt * alfapr;
  h->vel_dt = h->prm.dt;
  h->vel_ready = true;
  reset_active(h);
  h3dmg_status st = push_active(h);
  if (st) return st;
  enqueue_project(h);
  HIP_TRY(hipGetLastError());
  st = read_stats(h);
  if (st) return st;
  if (rmax) *rmax = projected_rmax(h);
  h->rhs_ready = true;  // f holds su
  return H3DMG_OK;
}

h3dmg_status h3dmg_velocity_step(h3dmg_handle *h, double *rmax) {
  if (!h) return fail(H3DMG_ERR_ARG, "null handle");
  if (!h->vel_ready)
    return fail(H3DMG_ERR_ARG,
                "h3dmg_velocity_begin has not run since the handle's enabled blocks last "
                "changed");
  for (int b = 0; b < h->nb; ++b)  // every projected block has its arrays
    if (h->blk[b].enabled && (!h->vel[b].us || !h->vel[b].u))
      return fail(H3DMG_ERR_ARG, "block %d: no velocity arrays", b);
  reset_active(h);
  h3dmg_status st = push_active(h);
  if (st) return st;
  st = run_velocity_step(h);
  if (st) return st;
  st = read_stats(h);
  if (st) return st;
  if (rmax) *rmax = projected_rmax(h);
  return H3DMG_OK;
}

h3dmg_status h3dmg_solve_velocity(h3dmg_handle *h, const double *const *d_ustar,
                                  const double *const *d_vstar, const double *const *d_wstar,
                                  double *const *d_u, double *const *d_v, double *const *d_w,
                                  double alfapr, double *hist, int hist_cap, int *ncycles,
                                  double *rmax) {
  if (!h) return fail(H3DMG_ERR_ARG, "null handle");
  double r = 0.0;
  h3dmg_status st =
      h3dmg_velocity_begin(h, d_ustar, d_vstar, d_wstar, d_u, d_v, d_w, alfapr, &r);
  if (st) return st;
  if (hist && hist_cap > 0) hist[0] = r;
  int it = 0;
  bool done = false;
  while (it < h->prm.maxcy) {
    st = h3dmg_velocity_step(h, &r);
    if (st) return st;
    ++it;
    if (hist && it < hist_cap) hist[it] = r;
    if (!std::isfinite(r)) break;
    if (r < h->prm.tol && it > 1) {  // newsolv_mg.for:57
      done = true;
      break;
    }
  }
  if (ncycles) *ncycles = it;
  if (rmax) *rmax = r;
  if (!std::isfinite(r)) return fail(H3DMG_ERR_DIVERGED, "non-finite rmax after %d cycles", it);
  if (r > 10.0) return fail(H3DMG_ERR_DIVERGED, "rmax %g > 10", r);  // newsolv_mg.for:65-75
  if (!done) return fail(H3DMG_ERR_NOT_CONVERGED, "not converged in %d cycles (rmax %g)", it, r);
  return H3DMG_OK;
}

h3dmg_status h3dmg_block_stats(const h3dmg_handle *h, int *cycles, double *rel_res) {
  if (!h) return fail(H3DMG_ERR_ARG, "null handle");
  for (int b = 0; b < h->nb; ++b) {
    if (cycles) cycles[b] = h->blk[b].cycles;
    if (rel_res) rel_res[b] = h->blk[b].rel;
  }
  return H3DMG_OK;
}

h3dmg_status h3dmg_remove_mean(h3dmg_handle *h) {
  if (!h) return fail(H3DMG_ERR_ARG, "null handle");
  if (h->connected)
    return fail(H3DMG_ERR_UNSUPPORTED, "the mean of blocks with neighbours is not per block");
  for (int b = 0; b < h->nb; ++b) {
    const FineDev &F = h->fine[b];
    launch_remove_mean(h->stream, h->d_fine + b, F.ni, F.nj, F.nk, nullptr);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  return H3DMG_OK;
}

// ---- test seams ----------------------------------------------------------
h3dmg_status h3dmg_dbg_put(h3dmg_handle *h, int b, int level, int which, const double *d_src) {
  h3dmg_status st = check_block(h, b);
  if (st) return st;
  if (level < 1 || level > h->blk[b].L || !d_src) return fail(H3DMG_ERR_ARG, "bad level/ptr");
  const int d = h->blk[b].desc[level];
  const LevelDev &L = h->levs[d];
  if (h->lev_is64[d])
    launch_pack_level<double>(h->stream, h->d_levs + d, which, 0, d_src, L.ni, L.nj, L.nk);
  else
    launch_pack_level<float>(h->stream, h->d_levs + d, which, 0, d_src, L.ni, L.nj, L.nk);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  return H3DMG_OK;
}

h3dmg_status h3dmg_dbg_get(h3dmg_handle *h, int b, int level, int which, double *d_dst) {
  h3dmg_status st = check_block(h, b);
  if (st) return st;
  if (level < 1 || level > h->blk[b].L || !d_dst) return fail(H3DMG_ERR_ARG, "bad level/ptr");
  const int d = h->blk[b].desc[level];
  const LevelDev &L = h->levs[d];
  if (h->lev_is64[d])
    launch_unpack_level<double>(h->stream, h->d_levs, d, which, 0, d_dst, L.ni, L.nj, L.nk);
  else
    launch_unpack_level<float>(h->stream, h->d_levs, d, which, 0, d_dst, L.ni, L.nj, L.nk);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  return H3DMG_OK;
}

// The seams keep every level's data in buffer 0: an odd sweep count is
// followed by a copy back (a second pass of the smoother would change it).
h3dmg_status h3dmg_dbg_relax(h3dmg_handle *h, int level, int nsweep) {
  if (!h || level < 1 || level > h->Lmax || nsweep < 0) return fail(H3DMG_ERR_ARG, "bad level");
  std::fill(h->active.begin(), h->active.end(), 1);
  h3dmg_status st = push_active(h);
  if (st) return st;
  for (int grp = 0; grp < 2; ++grp) {
    const TileSet &ts = grp == 0 ? h->relax_nc[level] : h->relax_c[level];
    if (!ts.n()) continue;
    const bool is64 = !is_mixed(h) || grp == 1;
    int buf = 0;
    bool z = false;
    relax_any(h, is64, ts, nsweep, buf, z);
    if (buf == 1) {
      // copy buffer 1 -> 0 for every block of this group
      for (int b = 0; b < h->nb; ++b) {
        if (h->blk[b].L < level) continue;
        if ((grp == 1) != (h->blk[b].L == level)) continue;
        const int d = h->blk[b].desc[level];
        const LevelDev &L = h->levs[d];
        const size_t half = (size_t)L.H * L.nj * L.nk * (h->lev_is64[d] ? 8 : 4);
        for (int c = 0; c < 2; ++c)
          HIP_TRY(hipMemcpyAsync(L.phi[0][c], L.phi[1][c], half, hipMemcpyDeviceToDevice,
                                 h->stream));
      }
    }
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  return H3DMG_OK;
}

h3dmg_status h3dmg_dbg_restrict(h3dmg_handle *h, int level) {
  if (!h || level < 2 || level > h->Lmax) return fail(H3DMG_ERR_ARG, "bad level");
  std::fill(h->active.begin(), h->active.end(), 1);
  h3dmg_status st = push_active(h);
  if (st) return st;
  const bool f64 = !is_mixed(h);
  if (h->lsm) {
    for (const TileSet *ts : {&h->restr_nc[level], &h->restr_c[level]})
      launch_lsm_restrict(h->stream, ts->d, ts->n(), h->d_levs, h->d_active, 0);
  } else if (h->restr_nc[level].n()) {
    if (f64)
      launch_restrict<double, double>(h->stream, h->restr_nc[level].d, h->restr_nc[level].n(),
                                      h->restr_nc[level].depth, h->d_levs, h->d_active, 0);
    else
      launch_restrict<float, float>(h->stream, h->restr_nc[level].d, h->restr_nc[level].n(),
                                    h->restr_nc[level].depth, h->d_levs, h->d_active, 0);
  }
  if (!h->lsm && h->restr_c[level].n()) {
    if (f64)
      launch_restrict<double, double>(h->stream, h->restr_c[level].d, h->restr_c[level].n(),
                                      h->restr_c[level].depth, h->d_levs, h->d_active, 0);
    else
      launch_restrict<float, double>(h->stream, h->restr_c[level].d, h->restr_c[level].n(),
                                     h->restr_c[level].depth, h->d_levs, h->d_active, 0);
  }
  // mgrestr zeroes the coarse phi (mgsolver.for:655-662)
  if (h->nat_nc[level].n()) fill_any(h, f64, h->nat_nc[level], 0);
  if (h->nat_c[level].n()) fill_any(h, true, h->nat_c[level], 0);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  return H3DMG_OK;
}

h3dmg_status h3dmg_dbg_prolong(h3dmg_handle *h, int level) {
  if (!h || level < 1 || level >= h->Lmax) return fail(H3DMG_ERR_ARG, "bad level");
  std::fill(h->active.begin(), h->active.end(), 1);
  h3dmg_status st = push_active(h);
  if (st) return st;
  const bool f64 = !is_mixed(h);
  if (h->prol_nc[level].n()) {
    if (f64)
      launch_prolong<double, double>(h->stream, h->prol_nc[level].d, h->prol_nc[level].n(),
                                     h->d_levs, h->d_active, 0, 0, h->connected);
    else
      launch_prolong<float, float>(h->stream, h->prol_nc[level].d, h->prol_nc[level].n(),
                                   h->d_levs, h->d_active, 0, 0, h->connected);
  }
  if (h->prol_c[level].n()) {
    if (f64)
      launch_prolong<double, double>(h->stream, h->prol_c[level].d, h->prol_c[level].n(),
                                     h->d_levs, h->d_active, 0, 0, h->connected);
    else
      launch_prolong<float, double>(h->stream, h->prol_c[level].d, h->prol_c[level].n(),
                                    h->d_levs, h->d_active, 0, 0, h->connected);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  return H3DMG_OK;
}

h3dmg_status h3dmg_set_march(h3dmg_handle *h, int level, int planes) {
  if (!h || level < 1 || level > h->Lmax) return fail(H3DMG_ERR_ARG, "bad level");
  std::lock_guard<std::mutex> lock(g_device_mutex);
  HIP_TRY(hipStreamSynchronize(h->stream));
  drop_graph(h);
  h->march_kb[level] = planes > 0 ? planes : 0;
  return upload_relax(h, level);
}

h3dmg_status h3dmg_set_lanes(h3dmg_handle *h, int level, int lanes) {
  if (!h || level < 1 || level > h->Lmax) return fail(H3DMG_ERR_ARG, "bad level");
  if (lanes < 0 || lanes > 64 || (lanes & (lanes - 1)))
    return fail(H3DMG_ERR_ARG, "lanes per row: 0 or a power of two up to 64");
  std::lock_guard<std::mutex> lock(g_device_mutex);
  HIP_TRY(hipStreamSynchronize(h->stream));
  drop_graph(h);
  for (const BlockInfo &B : h->blk) {
    if (B.L < level) continue;
    LevelDev &L = h->levs[B.desc[level]];
    const int natural = lanes_per_row(L.H);
    L.lpr = lanes ? std::min(lanes, natural) : natural;
  }
  HIP_TRY(hipMemcpyAsync(h->d_levs, h->levs.data(), h->levs.size() * sizeof(LevelDev),
                         hipMemcpyHostToDevice, h->stream));
  return upload_relax(h, level);  // synchronises the stream
}

int h3dmg_get_lanes(const h3dmg_handle *h, int level) {
  if (!h || level < 1 || level > h->Lmax) return -1;
  for (const BlockInfo &B : h->blk)
    if (B.L >= level) return h->levs[B.desc[level]].lpr;
  return -1;
}

h3dmg_status h3dmg_set_tail(h3dmg_handle *h, int level) {
  if (!h || level < 0) return fail(H3DMG_ERR_ARG, "bad tail level");
  if (h->lsm && level <= h->Lmax)
    return fail(H3DMG_ERR_UNSUPPORTED, "the tail kernel holds no variable-density coefficients");
  if (h->prm.smoother != H3DMG_SMOOTH_RBGS && level <= h->Lmax)
    return fail(H3DMG_ERR_UNSUPPORTED, "the tail kernel runs the red-black smoother only");
  if (h->connected && level <= h->Lmax)
    return fail(H3DMG_ERR_UNSUPPORTED, "the tail kernel runs independent blocks only");
  std::lock_guard<std::mutex> lock(g_device_mutex);
  HIP_TRY(hipStreamSynchronize(h->stream));
  drop_graph(h);
  h->tail_req = level;
  return upload_tail(h);
}

int h3dmg_get_tail(const h3dmg_handle *h) {
  if (!h) return -1;
  return h->tail_tiles.n() ? h->tail_lt : h->Lmax + 1;
}

int h3dmg_get_march(const h3dmg_handle *h, int level) {
  if (!h || level < 1 || level > h->Lmax) return -1;
  for (const TileSet *ts : {&h->relax_nc[level], &h->relax_c[level]})
    for (const Tile &t : ts->h)
      if (t.kind != 2) return t.kind >= 4 ? 0 : t.z1 - t.z0;
  return 0;
}

h3dmg_status h3dmg_set_fused(h3dmg_handle *h, long long max_cells) {
  if (!h || max_cells < 0) return fail(H3DMG_ERR_ARG, "bad fused threshold");
  if (h->lsm && max_cells > 0)
    return fail(H3DMG_ERR_UNSUPPORTED, "fused sweeps hold no variable-density coefficients");
  if (h->prm.smoother != H3DMG_SMOOTH_RBGS && max_cells > 0)
    return fail(H3DMG_ERR_UNSUPPORTED, "fused sweeps run the red-black smoother only");
  if (h->connected && max_cells > 0)
    return fail(H3DMG_ERR_UNSUPPORTED, "fused sweeps run independent blocks only");
  std::lock_guard<std::mutex> lock(g_device_mutex);
  HIP_TRY(hipStreamSynchronize(h->stream));
  drop_graph(h);
  h->fused_max = (long)max_cells;
  for (int g = 1; g <= h->Lmax; ++g) {
    h3dmg_status st = upload_relax(h, g);
    if (st) return st;
  }
  return H3DMG_OK;
}

long long h3dmg_get_fused(const h3dmg_handle *h) { return h ? h->fused_max : -1; }

}  // extern "C"
