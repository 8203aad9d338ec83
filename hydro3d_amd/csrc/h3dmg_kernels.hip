// h3dmg_kernels.hip -- gfx950 kernels of the pressure-Poisson multigrid.
//
// All kernels are HBM-streaming 7-point stencils over the colour-split level
// arrays described in h3dmg_internal.h.  A launch covers every (active) block
// at one level: workgroups read their (block, tile) from a tile table, so one
// launch serves a single 256^3 block and 48 uneven blocks alike.
//
// Arithmetic is written in the order of src_v70/mgsolver.for and the library
// is built with -ffp-contract=off, so the FP64 path is bitwise the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "h3dmg_internal.h"

namespace h3dmg {

namespace {

__device__ __forceinline__ size_t cidx(int H, int nj, int i, int j, int k) {
  return ((size_t)k * (size_t)nj + (size_t)j) * (size_t)H + (size_t)(i >> 1);
}

// image of a 0-based coordinate that may lie one cell outside [0, n)
__device__ __forceinline__ int image(int x, int n, bool periodic) {
  if (x < 0) return periodic ? n - 1 : 0;
  if (x >= n) return periodic ? 0 : n - 1;
  return x;
}

template <class T>
struct Coef {
  T aw, ae, as, an, ab, at, ap;
};

// Face f (0 west, 1 east, 2 south, 3 north, 4 bottom, 5 top) of a level or
// finest-level descriptor is a wall when its axis is not periodic and no
// neighbour block lies across it.
template <class LD>
__device__ __forceinline__ bool wall_face(const LD &L, int f) {
  return !((L.per >> (f >> 1)) & 1) && L.nbr[f] < 0;
}

// coeff, mgsolver.for:185-210: 1/h^2 towards every face, 0 towards a
// physical (non-periodic, neighbourless) face, ap = -1*(sum) summed aw..at
// left to right.
template <class T>
__device__ __forceinline__ Coef<T> cell_coef(const LevelDev &L, int i, int j, int k) {
  const T cx = (T)L.cx, cy = (T)L.cy, cz = (T)L.cz;
  Coef<T> c;
  c.aw = (i == 0 && wall_face(L, 0)) ? T(0) : cx;
  c.ae = (i == L.ni - 1 && wall_face(L, 1)) ? T(0) : cx;
  c.as = (j == 0 && wall_face(L, 2)) ? T(0) : cy;
  c.an = (j == L.nj - 1 && wall_face(L, 3)) ? T(0) : cy;
  c.ab = (k == 0 && wall_face(L, 4)) ? T(0) : cz;
  c.at = (k == L.nk - 1 && wall_face(L, 5)) ? T(0) : cz;
  c.ap = T(-1) * (c.aw + c.ae + c.as + c.an + c.ab + c.at);
  return c;
}

__device__ __forceinline__ Coef<double> fine_coef(const FineDev &F, int i, int j, int k) {
  Coef<double> c;
  c.aw = (i == 0 && wall_face(F, 0)) ? 0.0 : F.cx;
  c.ae = (i == F.ni - 1 && wall_face(F, 1)) ? 0.0 : F.cx;
  c.as = (j == 0 && wall_face(F, 2)) ? 0.0 : F.cy;
  c.an = (j == F.nj - 1 && wall_face(F, 3)) ? 0.0 : F.cy;
  c.ab = (k == 0 && wall_face(F, 4)) ? 0.0 : F.cz;
  c.at = (k == F.nk - 1 && wall_face(F, 5)) ? 0.0 : F.cz;
  c.ap = -1.0 * (c.aw + c.ae + c.as + c.an + c.ab + c.at);
  return c;
}

// The cell whose value a ghost holds after exchangepp (exchangepp.for:44-
// 1105, then mgbound and exchbc_mgpp, :2830-2832): (x, y, z) is a padded
// position in [-1, n] of descriptor d of `tab`; on return it is an interior
// position of the returned descriptor.  Walls mirror first, in the block
// itself (mgbound runs over whole faces after the face syncs); then along x,
// y and z in turn a coordinate outside the block moves into the neighbour
// across that face, or wraps around a periodic axis.  In a box of blocks
// this is the value the reference's sequence of whole-face copies, corner
// and edge messages and mirrors leaves in every face, edge and corner ghost.
template <class LD>
__device__ __forceinline__ int resolve_cell(const LD *__restrict__ tab, int d, int &x, int &y,
                                            int &z) {
  const LD *P = &tab[d];
  if (x < 0 && wall_face(*P, 0)) x = 0;
  if (x >= P->ni && wall_face(*P, 1)) x = P->ni - 1;
  if (y < 0 && wall_face(*P, 2)) y = 0;
  if (y >= P->nj && wall_face(*P, 3)) y = P->nj - 1;
  if (z < 0 && wall_face(*P, 4)) z = 0;
  if (z >= P->nk && wall_face(*P, 5)) z = P->nk - 1;
  int c[3] = {x, y, z};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int n = a == 0 ? P->ni : a == 1 ? P->nj : P->nk;
    if (c[a] < 0) {
      if (P->nbr[2 * a] >= 0) {
        d = P->nbr[2 * a];
        P = &tab[d];
        c[a] += a == 0 ? P->ni : a == 1 ? P->nj : P->nk;
      } else {
        c[a] = ((P->per >> a) & 1) ? c[a] + n : 0;
      }
    } else if (c[a] >= n) {
      if (P->nbr[2 * a + 1] >= 0) {
        d = P->nbr[2 * a + 1];
        P = &tab[d];
        c[a] -= n;
      } else {
        c[a] = ((P->per >> a) & 1) ? c[a] - n : n - 1;
      }
    }
  }
  // A position at most one cell outside the block is now inside block d.
  // Lanes that lie beyond their level (their results are discarded) may
  // pass positions further out; clamping keeps their reads inside block d.
  x = min(max(c[0], 0), P->ni - 1);
  y = min(max(c[1], 0), P->nj - 1);
  z = min(max(c[2], 0), P->nk - 1);
  return d;
}

// value of any cell at padded position (x,y,z) in [-1, n] of level
// descriptor d, buffer `buf`, ghosts resolved to the cell they hold
template <class T>
__device__ __forceinline__ T level_value(const LevelDev *__restrict__ levs, int d, int buf, int x,
                                         int y, int z) {
  d = resolve_cell(levs, d, x, y, z);
  const LevelDev &L = levs[d];
  const T *a = (const T *)L.phi[buf][(x + y + z) & 1];
  return a[cidx(L.H, L.nj, x, y, z)];
}

// the same for the finest-level FP64 pressure of block b
__device__ __forceinline__ double fine_value(const FineDev *__restrict__ fine, int b, int x, int y,
                                             int z) {
  b = resolve_cell(fine, b, x, y, z);
  const FineDev &F = fine[b];
  return F.p[(x + y + z) & 1][cidx(F.H, F.nj, x, y, z)];
}

// ---------------------------------------------------------------------------
// One colour half-sweep of the red-black smoother (mgsolver.for:553-605):
//   phi = (rhs - (aw*W + ae*E + as*S + an*N + ab*B + at*T)) / ap
// colour 0 = red, 1 = black.  A sweep reads buffer `src` and writes `1-src`:
//   red:   interior neighbours (black) from black[src]
//   black: interior neighbours (red)   from red[1-src] (just written)
//   ghost neighbours from buffer src, resolved to their image (the
//   pre-sweep state the reference's ghosts hold until exchangepp at the end
//   of the sweep, mgsolver.for:614-615).
// zero != 0: the level holds zeros in buffer src (the first sweep after
// mgrestr zeroed phi, mgsolver.for:655-662, or pp = 0 at the finest level).
//
// Three tile kinds share a half-sweep (one launch, k_rb_sweep): march tiles
// whose cells have all y/z neighbours inside the block (no ghost logic at
// all), the other march tiles, and x-face tiles for the cells on the two x
// faces (i = 0, ni-1), which the march tiles skip.
//
// Work layout of the march kernels.  A lane owns two adjacent entries m0,
// m0+1 of one row of the colour arrays; L.lpr lanes cover a row segment and a
// wavefront holds 64/L.lpr rows (coarse levels pack several short rows into
// one wavefront).  Each lane walks the planes [t.z0, t.z1).  For a cell
// i = 2m + s of colour c (s = (c+j+k)&1) entry m of the other colour in the
// same row is its east (s = 0) or west (s = 1) neighbour and entry m of planes
// k-1, k+1 its bottom and top neighbours: that "own column" pair streams once
// through a register window; the remaining x-neighbour is the lane's other
// entry or the adjacent lane's (DPP wavefront shift); y-neighbours are loaded.
template <class T>
struct Vec2;
template <>
struct Vec2<float> {
  typedef float __attribute__((ext_vector_type(2))) type;
};
template <>
struct Vec2<double> {
  typedef double __attribute__((ext_vector_type(2))) type;
};

// Workgroups are dispatched round-robin over the 8 XCDs (blockIdx % 8);
// give each XCD a contiguous range of tiles so that neighbouring rows and
// planes are read through the same L2.
__device__ __forceinline__ int xcd_remap(int b, int n) {
  const int per = n >> 3, rem = n & 7;
  const int x = b & 7, q = b >> 3;
  return x * per + min(x, rem) + q;
}

// Division by ap: the FP64 path divides, as the reference does (bitwise
// parity); the FP32 path multiplies by the reciprocal (one rounding more,
// within the FP32 V-cycle's own error).
__device__ __forceinline__ double div_ap(double num, double ap, double /*inv_ap*/) {
  return num / ap;
}
__device__ __forceinline__ float div_ap(float num, float /*ap*/, float inv_ap) {
  return num * inv_ap;
}

// DPP wavefront shifts by one lane (GFX9 wave_shr:1 = 0x138, wave_shl:1 =
// 0x130): lane l receives lane l-1's (shr) or l+1's (shl) value; the end lane
// receives 0.  Every lane of the wavefront must be active at the call.
__device__ __forceinline__ int dpp_i32(int v, int ctrl_shr) {
  return ctrl_shr ? __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false)
                  : __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, false);
}
__device__ __forceinline__ float wave_shift(float v, int shr) {
  return __int_as_float(dpp_i32(__float_as_int(v), shr));
}
__device__ __forceinline__ double wave_shift(double v, int shr) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_i32((int)(b & 0xffffffffll), shr);
  const int hi = dpp_i32((int)(b >> 32), shr);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Global-address-space views of the level arrays (the descriptor holds
// generic pointers; loads through these compile to global_load with a
// scalar base and a 32-bit lane offset).
template <class T>
using gmem = __attribute__((address_space(1))) T;

template <class T, bool INNER>
__device__ __forceinline__ void march_tile(const Tile &t, const LevelDev *__restrict__ levs,
                                           int colour, int src, int zero) {
  const LevelDev &L = levs[t.lev];
  typedef typename Vec2<T>::type v2;
  const int lpr = L.lpr;
  const int lane = (int)threadIdx.x;
  const int rl = lane & (lpr - 1);
  const int m0 = t.x0 + 2 * rl;
  const int j = t.y0 + (int)threadIdx.y * (64 / lpr) + lane / lpr;
  const int ni = L.ni, nj = L.nj, nk = L.nk, H = L.H;
  const bool col_ok = m0 < H && j < nj;  // H even: m0 + 1 < H as well
  const unsigned row = (unsigned)H, plane = (unsigned)H * (unsigned)nj;
  const gmem<T> *O = (const gmem<T> *)(colour == 0 ? L.phi[src][1] : L.phi[1 - src][0]);
  const gmem<T> *R = (const gmem<T> *)L.rhs[colour];
  gmem<T> *D = (gmem<T> *)L.phi[1 - src][colour];
  const bool zero_int = zero && colour == 0;  // uniform
  const bool zero_ghost = zero != 0;
  // lanes outside the level read a valid entry (clamped) and store nothing
  const int jc = min(j, nj - 1);
  const unsigned cbase = (unsigned)jc * row + (unsigned)min(m0, H - 2);
  const v2 vz = {T(0), T(0)};
  // tiles that do not start/end a row: the end lanes fetch entries m0-1 and
  // m0+2 from the neighbouring tile's range (uniform test, then one lane)
  const bool tile_left = t.x0 > 0, tile_right = t.x0 + 2 * lpr < H;
  const T cx = (T)L.cx, cy = (T)L.cy, cz = (T)L.cz;
  const T ap_in = T(-1) * (cx + cx + cy + cy + cz + cz);
  const T inv_in = T(1) / ap_in;

  // byte offsets in 32 bits (arrays < 4 GiB): scalar base + lane offset
  auto ld2 = [&](const gmem<T> *a, unsigned off) -> v2 {
    return *(const gmem<v2> *)((const gmem<char> *)a + off * (unsigned)sizeof(T));
  };
  auto ld1 = [&](const gmem<T> *a, unsigned off) -> T {
    return *(const gmem<T> *)((const gmem<char> *)a + off * (unsigned)sizeof(T));
  };
  auto own = [&](int k) -> v2 {
    if (zero_int || (!INNER && (unsigned)k >= (unsigned)nk)) return vz;
    return ld2(O, (unsigned)k * plane + cbase);
  };
  auto ghost = [&](int x, int y, int z) -> T {
    return zero_ghost ? T(0) : level_value<T>(levs, t.lev, src, x, y, z);
  };

  int k = t.z0;
  const int k1 = t.z1;
  v2 o_m = own(k - 1), o_c = own(k);
  for (; k < k1; ++k) {
    const v2 o_p = own(k + 1);
    const unsigned q = (unsigned)k * plane + cbase;
    v2 S2 = vz, N2 = vz;
    if (!zero_int) {
      // jc: the lane's row, clamped like cbase (lanes beyond the level load
      // in-bounds rows; their results are not stored)
      if (INNER || jc > 0) S2 = ld2(O, q - row);
      if (INNER || jc + 1 < nj) N2 = ld2(O, q + row);
    }
    const v2 R2 = ld2(R, q);
    T lf = wave_shift(o_c.y, 1);  // entry m0-1: lane-1's second entry
    T rt = wave_shift(o_c.x, 0);  // entry m0+2: lane+1's first entry
    if (tile_left && !zero_int && rl == 0) lf = ld1(O, q - 1);
    if (tile_right && !zero_int && rl == lpr - 1) rt = ld1(O, q + 2);
    const int s = (colour + j + k) & 1;
    // x-neighbours of the cells i0 = 2*m0 + s and i1 = i0 + 2
    const T W0 = s ? o_c.x : lf;
    const T E0 = s ? o_c.y : o_c.x;
    const T W1 = s ? o_c.y : o_c.x;
    const T E1 = s ? rt : o_c.y;
    const int i0 = 2 * m0 + s, i1 = i0 + 2;
    T out0, out1;
    if (INNER) {
      const T sum0 = cx * W0 + cx * E0 + cy * S2.x + cy * N2.x + cz * o_m.x + cz * o_p.x;
      const T sum1 = cx * W1 + cx * E1 + cy * S2.y + cy * N2.y + cz * o_m.y + cz * o_p.y;
      out0 = div_ap(R2.x - sum0, ap_in, inv_in);
      out1 = div_ap(R2.y - sum1, ap_in, inv_in);
    } else {
      const T Wv[2] = {W0, W1}, Ev[2] = {E0, E1};
      T o[2];
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int i = e ? i1 : i0;
        const T S = (jc > 0) ? (e ? S2.y : S2.x) : ghost(i, -1, k);
        const T N = (jc + 1 < nj) ? (e ? N2.y : N2.x) : ghost(i, nj, k);
        const T B = (k > 0) ? (e ? o_m.y : o_m.x) : ghost(i, jc, -1);
        const T Tt = (k + 1 < nk) ? (e ? o_p.y : o_p.x) : ghost(i, jc, nk);
        const Coef<T> c = cell_coef<T>(L, i, jc, k);
        const T sum = c.aw * Wv[e] + c.ae * Ev[e] + c.as * S + c.an * N + c.ab * B + c.at * Tt;
        o[e] = div_ap((e ? R2.y : R2.x) - sum, c.ap, T(1) / c.ap);
      }
      out0 = o[0];
      out1 = o[1];
    }
    // cells on the x faces belong to the x-face tiles of the same launch
    const bool w0 = col_ok && i0 > 0 && i0 < ni - 1, w1 = col_ok && i1 < ni - 1;
    gmem<char> *dq = (gmem<char> *)D + q * (unsigned)sizeof(T);
    if (w0 && w1) {
      const v2 o = {out0, out1};
      *(gmem<v2> *)dq = o;
    } else {
      if (w0) *(gmem<T> *)dq = out0;
      if (w1) *(gmem<T> *)(dq + sizeof(T)) = out1;
    }
    o_m = o_c;
    o_c = o_p;
  }
}

// March tile of KB planes with every load issued up front (one memory round
// trip per tile): the own-column window (KB+2 planes), rows j-1 and j+1, the
// rhs and one x-slot entry per plane, then the KB updates, then the stores.
//
// The tile also updates the cells on its block's faces.  A ghost neighbour
// takes the pre-sweep (buffer src) value of its image -- the cell on the
// opposite face across a periodic face, the cell itself across a wall (whose
// coefficient is 0) -- which is the exchange/boundary state the reference's
// ghosts hold during a sweep.  Every array of a level lies in one arena
// allocation, so each load is the level's base (scalar) plus a 32-bit lane
// offset that may select any of the arrays: the image rows and entries of
// face lanes cost no branch.
//   x faces: lane 0 of a row segment (entry m0 = 0) holds cell 0, the lane
//     with m0 = H-2 holds cell ni-1; their x-slot loads the x-ghost image.
//   y/z faces (BND tiles only): rows 0 and nj-1 read their image rows;
//     chunks at plane 0 or nk-1 read the image plane into the window.
// The arithmetic of every cell is cell_coef's and the sum of mgsolver.for's
// relaxation, as in march_tile.
template <class T, int KB, bool ZERO_INT, bool BND>
__device__ __forceinline__ void march_tile_fast(const Tile &t, const LevelDev &L, int colour,
                                                int src, int zero, bool on) {
  typedef typename Vec2<T>::type v2;
  constexpr unsigned SZ = sizeof(T);
  const int lpr = L.lpr;
  const int lane = (int)threadIdx.x;
  const int rl = lane & (lpr - 1);
  const int m0 = t.x0 + 2 * rl;
  const int j = t.y0 + (int)threadIdx.y * (64 / lpr) + lane / lpr;
  const int ni = L.ni, nj = L.nj, nk = L.nk, H = L.H;
  const bool col_ok = on && m0 < H && j < nj;  // on: the block's flag gates the stores
  const int jc = min(j, nj - 1);
  const unsigned row = (unsigned)H * SZ, plane = row * (unsigned)nj;  // bytes
  // byte offsets of the level's arrays from its arena base (phi[0][0])
  const char *base = (const char *)L.phi[0][0];
  auto aoff = [&](const void *a) -> unsigned { return (unsigned)((const char *)a - base); };
  const unsigned aO = aoff(colour == 0 ? L.phi[src][1] : L.phi[1 - src][0]);
  const unsigned aOs = aoff(L.phi[src][1 - colour]);  // other colour, pre-sweep
  const unsigned aPs = aoff(L.phi[src][colour]);      // own colour, pre-sweep
  const unsigned aR = aoff(L.rhs[colour]);
  const unsigned aD = aoff(L.phi[1 - src][colour]);
  const gmem<char> *Bs = (const gmem<char> *)base;
  gmem<char> *Bd = (gmem<char> *)base;
  const bool px = L.per & 1, py = L.per & 2, pz = L.per & 4;
  const bool zg = zero != 0;  // ghosts read the all-zero pre-sweep state
  // the image of a ghost across a periodic face of extent n has the other
  // colour when n is even; a wall's image (the cell itself) has its own
  const unsigned aGX = (px && !(ni & 1)) ? aOs : aPs;
  const unsigned aGY = (py && !(nj & 1)) ? aOs : aPs;
  const unsigned aGZ = (pz && !(nk & 1)) ? aOs : aPs;
  const unsigned jrow = (unsigned)jc * row;
  const unsigned cell = jrow + (unsigned)min(m0, H - 2) * SZ;
  const int me = (ni - 1) >> 1;  // entry of cell ni-1; its pair starts at H-2
  const bool tile_left = t.x0 > 0, tile_right = t.x0 + 2 * lpr < H;

  // x slots: entry m0-1 for lane 0 of a row segment (neighbouring tile or the
  // west ghost image), entry m0+2 for the last lane (neighbouring tile) or
  // the east ghost image for the lane of cell ni-1; one load serves both
  // unless a lane has both roles (one-lane rows, a last tile of one pair)
  const bool role_r = tile_right ? rl == lpr - 1 : m0 == H - 2;
  const unsigned off_l =
      tile_left ? aO + cell - SZ : aGX + jrow + (unsigned)(px ? me : 0) * SZ;
  const unsigned off_r =
      tile_right ? aO + cell + 2 * SZ : aGX + jrow + (unsigned)(px ? 0 : me) * SZ;
  const bool dual = lpr == 1 || (tile_left && t.x0 == H - 2);
  const unsigned off_x = rl == 0 ? off_l : (role_r ? off_r : aO + cell);
  const unsigned off_x2 = role_r ? off_r : aO + cell;
  const bool gx_l = rl == 0 && !tile_left, gx_r = role_r && !tile_right;
  const bool zx_1 = zg && (rl == 0 ? gx_l : gx_r), zx_2 = zg && gx_r;

  // y rows: image rows on the y faces of BND tiles
  unsigned off_s = aO + cell - row, off_n = aO + cell + row;
  bool zs = false, zn = false;
  T as = (T)L.cy, an = (T)L.cy;
  if (BND) {
    if (jc == 0) {
      off_s = py ? aGY + cell + (unsigned)(nj - 1) * row : aPs + cell;
      zs = zg;
      if (!py) as = T(0);
    }
    if (jc == nj - 1) {
      off_n = py ? aGY + cell - (unsigned)(nj - 1) * row : aPs + cell;
      zn = zg;
      if (!py) an = T(0);
    }
  }
  const T cx = (T)L.cx, cy = (T)L.cy, cz = (T)L.cz;
  const T ap_in = T(-1) * (cx + cx + cy + cy + cz + cz);
  const T inv_in = T(1) / ap_in;
  const T ap_x1 = T(-1) * (cx + cy + cy + cz + cz);  // one x coefficient 0 (wall face)
  const T inv_x1 = T(1) / ap_x1;
  const v2 vz = {T(0), T(0)};

  auto ld2 = [&](unsigned off) -> v2 { return *(const gmem<v2> *)(Bs + off); };
  auto ld1 = [&](unsigned off) -> T { return *(const gmem<T> *)(Bs + off); };

  {
    const int k0 = t.z0;
    const unsigned p0 = (unsigned)k0 * plane;
    v2 own[KB + 2], S2[KB], N2[KB], R2[KB];
    T x1[KB], x2[KB];
    if (!ZERO_INT) {
#pragma unroll
      for (int p = 0; p < KB + 2; ++p) {
        const int k = k0 - 1 + p;
        if (!BND || (k >= 0 && k < nk))
          own[p] = ld2(aO + cell + (unsigned)k * plane);
        else if (zg)
          own[p] = vz;
        else if (k < 0)
          own[p] = ld2((pz ? aGZ + (unsigned)(nk - 1) * plane : aPs) + cell);
        else
          own[p] = ld2((pz ? aGZ : aPs + (unsigned)(nk - 1) * plane) + cell);
      }
#pragma unroll
      for (int p = 0; p < KB; ++p) {
        const unsigned pp = p0 + (unsigned)p * plane;
        S2[p] = ld2(off_s + pp);
        N2[p] = ld2(off_n + pp);
        // x slots: read by the lanes at a row segment's ends only
        x1[p] = (rl == 0 || role_r) ? ld1(off_x + pp) : T(0);
        x2[p] = dual ? ld1(off_x2 + pp) : T(0);
      }
#pragma unroll
      for (int p = 0; p < KB; ++p) {
        if (BND && zs) S2[p] = vz;
        if (BND && zn) N2[p] = vz;
        if (zx_1) x1[p] = T(0);
        if (zx_2) x2[p] = T(0);
      }
    } else {
#pragma unroll
      for (int p = 0; p < KB + 2; ++p) own[p] = vz;
#pragma unroll
      for (int p = 0; p < KB; ++p) {
        S2[p] = N2[p] = vz;
        x1[p] = x2[p] = T(0);
      }
    }
#pragma unroll
    for (int p = 0; p < KB; ++p) R2[p] = ld2(aR + cell + p0 + (unsigned)p * plane);
#pragma unroll
    for (int p = 0; p < KB; ++p) {
      const int k = k0 + p;
      const v2 o_m = own[p], o_c = own[p + 1], o_p = own[p + 2];
      T lf = wave_shift(o_c.y, 1);  // entry m0-1: lane-1's second entry
      T rt = wave_shift(o_c.x, 0);  // entry m0+2: lane+1's first entry
      const T gr = dual && rl == 0 ? x2[p] : x1[p];
      if (rl == 0) lf = x1[p];
      if (role_r) rt = gr;
      const int s = (colour + j + k) & 1;
      const int i0 = 2 * m0 + s, i1 = i0 + 2;
      const bool fw0 = m0 == 0 && s == 0;  // cell 0
      const bool fe0 = i0 == ni - 1, fe1 = i1 == ni - 1;
      const T W0 = s ? o_c.x : lf;
      const T E0 = fe0 ? gr : (s ? o_c.y : o_c.x);
      const T W1 = s ? o_c.y : o_c.x;
      const T E1 = fe1 ? gr : (s ? rt : o_c.y);
      const T aw0 = (!px && fw0) ? T(0) : cx;
      const T ae0 = (!px && fe0) ? T(0) : cx;
      const T ae1 = (!px && fe1) ? T(0) : cx;
      T ab = cz, at = cz;
      if (BND) {
        if (!pz && k == 0) ab = T(0);
        if (!pz && k == nk - 1) at = T(0);
      }
      const T sum0 = aw0 * W0 + ae0 * E0 + as * S2[p].x + an * N2[p].x + ab * o_m.x + at * o_p.x;
      const T sum1 = cx * W1 + ae1 * E1 + as * S2[p].y + an * N2[p].y + ab * o_m.y + at * o_p.y;
      if (BND) {
        const T ap0 = T(-1) * (aw0 + ae0 + as + an + ab + at);
        const T ap1 = T(-1) * (cx + ae1 + as + an + ab + at);
        R2[p].x = div_ap(R2[p].x - sum0, ap0, T(1) / ap0);
        R2[p].y = div_ap(R2[p].y - sum1, ap1, T(1) / ap1);
      } else {
        const bool x0f = !px && (fw0 || fe0), x1f = !px && fe1;
        R2[p].x = div_ap(R2[p].x - sum0, x0f ? ap_x1 : ap_in, x0f ? inv_x1 : inv_in);
        R2[p].y = div_ap(R2[p].y - sum1, x1f ? ap_x1 : ap_in, x1f ? inv_x1 : inv_in);
      }
    }
#pragma unroll
    for (int p = 0; p < KB; ++p) {
      const int s = (colour + j + k0 + p) & 1;
      const int i0 = 2 * m0 + s, i1 = i0 + 2;
      const bool w0 = col_ok && i0 < ni, w1 = col_ok && i1 < ni;
      gmem<char> *dq = Bd + aD + cell + p0 + (unsigned)p * plane;
      if (w0 && w1) {
        *(gmem<v2> *)dq = R2[p];
      } else {
        if (w0) *(gmem<T> *)dq = R2[p].x;
        if (w1) *(gmem<T> *)(dq + SZ) = R2[p].y;
      }
    }
  }
}

// The cells of colour `colour` on the two x faces (i = 0 and i = ni-1), with
// the full ghost treatment.  The tile covers 256 rows j = t.x0 + ... of plane
// k = t.y0; one thread per row does both faces.
template <class T>
__device__ __forceinline__ void xface_tile(const Tile &t, const LevelDev *__restrict__ levs,
                                           int colour, int src, int zero) {
  const LevelDev &L = levs[t.lev];
  const int j = t.x0 + (int)threadIdx.x + TILE_X * (int)threadIdx.y;
  const int k = t.y0;
  if (j >= L.nj || k >= L.nk) return;
  const T *nb_int = colour == 0 ? (const T *)L.phi[src][1] : (const T *)L.phi[1 - src][0];
  const bool zero_int = zero && colour == 0;
  const bool zero_ghost = zero != 0;
  auto nb = [&](int x, int y, int z) -> T {
    const bool ghost = (unsigned)x >= (unsigned)L.ni || (unsigned)y >= (unsigned)L.nj ||
                       (unsigned)z >= (unsigned)L.nk;
    if (!ghost) return zero_int ? T(0) : nb_int[cidx(L.H, L.nj, x, y, z)];
    return zero_ghost ? T(0) : level_value<T>(levs, t.lev, src, x, y, z);
  };
  for (int side = 0; side < 2; ++side) {
    const int i = side ? L.ni - 1 : 0;
    if (side && i == 0) break;
    if (((i + j + k) & 1) != colour) continue;
    const Coef<T> c = cell_coef<T>(L, i, j, k);
    const T W = nb(i - 1, j, k), E = nb(i + 1, j, k);
    const T S = nb(i, j - 1, k), N = nb(i, j + 1, k);
    const T B = nb(i, j, k - 1), Tt = nb(i, j, k + 1);
    const size_t q = cidx(L.H, L.nj, i, j, k);
    const T r = ((const T *)L.rhs[colour])[q];
    const T sum = c.aw * W + c.ae * E + c.as * S + c.an * N + c.ab * B + c.at * Tt;
    ((T *)L.phi[1 - src][colour])[q] = div_ap(r - sum, c.ap, T(1) / c.ap);
  }
}

// One launch per colour half-sweep over a level.  Every tile of the list is
// KB planes deep: a march tile (kind 0: y/z interior, kind 3: touching a y/z
// face) or, on levels one cell wide in x (KB = 1 launches only), a general
// boundary march tile (kind 1) or an x-face tile (kind 2); the tiles write
// disjoint cells.  One instantiation per depth keeps each one's register
// count to what its own tiles need.
template <class T, int KB>
__device__ __forceinline__ void rb_dispatch(const Tile &t, const LevelDev &L,
                                            const LevelDev *__restrict__ levs, int colour, int src,
                                            int zero, bool on);

// In the one-descriptor kernels, one empty asm statement takes the tile,
// the block's flag and every descriptor field the kernel reads: all of them
// are loaded before it, together (scalar loads complete out of order, so any
// earlier use would wait for every load issued before it).

// Kernel arguments are scalars only (no aggregate) and the grid size is
// passed in: every argument then arrives preloaded in SGPRs instead of being
// read from the argument buffer.
// ONE: every tile of the launch has descriptor one_lev of block one_blk.  The
// tile, the descriptor and the block's flag are then read together, and the
// flag masks the stores instead of ending inactive workgroups early, so that
// the data loads wait for none of them but the tile and the descriptor
// (small levels are bound by this chain of latencies).
template <class T, int KB, bool ONE>
__global__ void __launch_bounds__(NTHREADS)
    k_rb_sweep(const Tile *__restrict__ tiles, const LevelDev *__restrict__ levs,
               const int *__restrict__ active, int colour, int src, int zero, int one_lev,
               int one_blk, int ntiles) {
  const Tile t = tiles[xcd_remap((int)blockIdx.x, ntiles)];
  if (ONE) {
    const LevelDev &L = levs[one_lev];
    const int on = active[one_blk];
    asm volatile("" ::"s"(t.x0), "s"(t.y0), "s"(t.z0), "s"(t.kind), "s"(on), "s"(L.ni), "s"(L.nj),
                 "s"(L.nk), "s"(L.H), "s"(L.per), "s"(L.lpr), "s"(L.cx), "s"(L.cy), "s"(L.cz),
                 "s"(L.phi[0][0]), "s"(L.phi[0][1]), "s"(L.phi[1][0]), "s"(L.phi[1][1]),
                 "s"(L.rhs[0]), "s"(L.rhs[1]));
    rb_dispatch<T, KB>(t, L, levs, colour, src, zero, on != 0);
  } else {
    const LevelDev &L = levs[t.lev];
    if (!active[L.blk]) return;
    rb_dispatch<T, KB>(t, L, levs, colour, src, zero, true);
  }
}

template <class T, int KB>
__device__ __forceinline__ void rb_dispatch(const Tile &t, const LevelDev &L,
                                            const LevelDev *__restrict__ levs, int colour, int src,
                                            int zero, bool on) {
  // the interior neighbours are zeros in the red half of a zero-state sweep
  const bool zint = zero && colour == 0;
  if (t.kind == 0) {
    zint ? march_tile_fast<T, KB, true, false>(t, L, colour, src, zero, on)
         : march_tile_fast<T, KB, false, false>(t, L, colour, src, zero, on);
  } else if (t.kind == 3) {
    zint ? march_tile_fast<T, KB, true, true>(t, L, colour, src, zero, on)
         : march_tile_fast<T, KB, false, true>(t, L, colour, src, zero, on);
  } else if (KB == 1 && on) {
    if (t.kind == 1)
      march_tile<T, false>(t, levs, colour, src, zero);
    else
      xface_tile<T>(t, levs, colour, src, zero);
  }
}

// ---------------------------------------------------------------------------
// Fused red-black sweep for small levels: both colour halves of one sweep
// in one launch (the split sweep costs two launches, and small levels are
// bound by launch and memory latency, not by bandwidth).  A tile is FX x FY
// x FZ cells at (x0, y0, z0) in cell coordinates.
//   Phase 1: the new red value of every red cell of the block inside the
//   tile grown by one cell on each side (the tile's own red cells and the
//   halo its black cells need), into LDS; the tile's own red cells also go
//   to buffer 1-src.  A red cell reads black neighbours from buffer src and
//   ghosts as images in buffer src, as the split red half-sweep does.
//   Phase 2: the tile's black cells from the red values in LDS (neighbours
//   inside the block) and ghost images in buffer src (pre-sweep, as the
//   split black half-sweep reads them), into buffer 1-src.
// Halo red cells are also computed by the neighbouring tiles with the same
// arithmetic, so every red value a black cell uses is bitwise the one the
// split sweep stores.  zero != 0: the level holds zeros in buffer src.
constexpr int FX = FUSED_X, FY = FUSED_Y, FZ = FUSED_Z;
constexpr int FEX = FX + 2, FEY = FY + 2, FEZ = FZ + 2;
constexpr int F_RED = FEX / 2 * FEY * FEZ;   // red positions of the grown tile
constexpr int F_BLACK = FX / 2 * FY * FZ;    // black cells of the tile
constexpr int F_RED_PER = (F_RED + NTHREADS - 1) / NTHREADS;
constexpr int F_BLACK_PER = F_BLACK / NTHREADS;

template <class T>
__global__ void __launch_bounds__(NTHREADS)
    k_rb_fused(const Tile *__restrict__ tiles, const LevelDev *__restrict__ levs,
               const int *__restrict__ active, int src, int zero) {
  const Tile t = tiles[xcd_remap((int)blockIdx.x, (int)gridDim.x)];
  const LevelDev &L = levs[t.lev];
  if (!active[L.blk]) return;
  __shared__ T red[FEZ * FEY * FEX];
  const int tid = (int)(threadIdx.x + blockDim.x * threadIdx.y);
  const int ni = L.ni, nj = L.nj, nk = L.nk, H = L.H;
  const bool px = L.per & 1, py = L.per & 2, pz = L.per & 4;
  // Addresses are 32-bit byte offsets from the level's first array (all six
  // arrays of a level lie in one arena allocation), through a global-
  // address-space base: a generic pointer's loads would also count against
  // the LDS wait counter.
  constexpr unsigned SZ = sizeof(T);
  const char *base = (const char *)L.phi[0][0];
  auto aoff = [&](const void *a) -> unsigned { return (unsigned)((const char *)a - base); };
  const unsigned aP0 = aoff(L.phi[src][0]), dP = aoff(L.phi[src][1]) - aP0;  // pre-sweep
  const unsigned aR0 = aoff(L.rhs[0]), aR1 = aoff(L.rhs[1]);
  const unsigned aD0 = aoff(L.phi[1 - src][0]), aD1 = aoff(L.phi[1 - src][1]);
  const gmem<char> *Bs = (const gmem<char> *)base;
  gmem<char> *Bd = (gmem<char> *)base;
  const unsigned row = (unsigned)H * SZ, plane = row * (unsigned)nj;
  auto inside = [&](int x, int y, int z) {
    return (unsigned)x < (unsigned)ni && (unsigned)y < (unsigned)nj && (unsigned)z < (unsigned)nk;
  };
  // byte offset of cell (x, y, z) inside its colour array
  auto at = [&](int x, int y, int z) -> unsigned {
    return (unsigned)z * plane + (unsigned)y * row + (unsigned)(x >> 1) * SZ;
  };
  // pre-sweep value of the image of (x, y, z) (the cell itself inside the
  // block; positions up to two cells outside map into the block)
  auto pre = [&](int x, int y, int z) -> T {
    return *(const gmem<T> *)(Bs + aP0 + (unsigned)((x + y + z) & 1) * dP + at(x, y, z));
  };
  auto prex = [&](int x, int y, int z) { return pre(image(x, ni, px), y, z); };
  auto prey = [&](int x, int y, int z) { return pre(x, image(y, nj, py), z); };
  auto prez = [&](int x, int y, int z) { return pre(x, y, image(z, nk, pz)); };
  const bool zl = zero != 0;  // pre-sweep state all zero

  // Every global load first.  Black cells of the tile: rhs and the images
  // of their six neighbours (used where a neighbour is a ghost).
  int bi[F_BLACK_PER], bj[F_BLACK_PER], bk[F_BLACK_PER];
  T brhs[F_BLACK_PER], bg[F_BLACK_PER][6];
#pragma unroll
  for (int q = 0; q < F_BLACK_PER; ++q) {
    const int b = tid + q * NTHREADS;
    const int row = b / (FX / 2), m = b % (FX / 2);
    const int j = t.y0 + row % FY, k = t.z0 + row / FY;
    const int i = t.x0 + (((t.x0 + j + k) & 1) ^ 1) + 2 * m;
    bi[q] = i;
    bj[q] = j;
    bk[q] = k;
    // clamped into the block: cells of a clipped tile load a valid entry
    const int ic = min(i, ni - 1), jc = min(j, nj - 1), kc = min(k, nk - 1);
    brhs[q] = *(const gmem<T> *)(Bs + aR1 + at(ic, jc, kc));
    bg[q][0] = prex(ic - 1, jc, kc);
    bg[q][1] = prex(ic + 1, jc, kc);
    bg[q][2] = prey(ic, jc - 1, kc);
    bg[q][3] = prey(ic, jc + 1, kc);
    bg[q][4] = prez(ic, jc, kc - 1);
    bg[q][5] = prez(ic, jc, kc + 1);
  }
  // red cells of the grown tile: rhs and six (black) neighbours, all
  // pre-sweep images
  int ri[F_RED_PER], rj[F_RED_PER], rk[F_RED_PER], re[F_RED_PER];
  T rrhs[F_RED_PER], rn[F_RED_PER][6];
#pragma unroll
  for (int q = 0; q < F_RED_PER; ++q) {
    const int r = min(tid + q * NTHREADS, F_RED - 1);
    const int row = r / (FEX / 2), m = r % (FEX / 2);
    const int ey = row % FEY, ez = row / FEY;
    const int j = t.y0 - 1 + ey, k = t.z0 - 1 + ez;
    const int ex = ((t.x0 - 1 + j + k) & 1) + 2 * m;  // (i + j + k) even
    const int i = t.x0 - 1 + ex;
    ri[q] = i;
    rj[q] = j;
    rk[q] = k;
    re[q] = (ez * FEY + ey) * FEX + ex;
    // positions of the grown tile outside the block load their image (unused)
    const int ic = image(i, ni, px), jc = image(j, nj, py), kc = image(k, nk, pz);
    rrhs[q] = *(const gmem<T> *)(Bs + aR0 + at(ic, jc, kc));
    rn[q][0] = prex(ic - 1, jc, kc);
    rn[q][1] = prex(ic + 1, jc, kc);
    rn[q][2] = prey(ic, jc - 1, kc);
    rn[q][3] = prey(ic, jc + 1, kc);
    rn[q][4] = prez(ic, jc, kc - 1);
    rn[q][5] = prez(ic, jc, kc + 1);
  }

  // phase 1: red cells of the grown tile that lie in the block, into LDS
  // (the tile's own red cells go to global memory after phase 2, so that
  // the barrier waits for LDS only)
  T rval[F_RED_PER];
#pragma unroll
  for (int q = 0; q < F_RED_PER; ++q) {
    const int i = ri[q], j = rj[q], k = rk[q];
    rval[q] = T(0);
    if (tid + q * NTHREADS >= F_RED || !inside(i, j, k)) continue;
    const Coef<T> c = cell_coef<T>(L, i, j, k);
    T v[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) v[d] = zl ? T(0) : rn[q][d];
    const T sum = c.aw * v[0] + c.ae * v[1] + c.as * v[2] + c.an * v[3] + c.ab * v[4] +
                  c.at * v[5];
    rval[q] = div_ap(rrhs[q] - sum, c.ap, T(1) / c.ap);
    red[re[q]] = rval[q];
  }
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): this wave's LDS writes done
  __builtin_amdgcn_s_barrier();

  // phase 2: the tile's black cells; neighbours in the block from LDS, ghost
  // neighbours their pre-sweep images
#pragma unroll
  for (int q = 0; q < F_BLACK_PER; ++q) {
    const int i = bi[q], j = bj[q], k = bk[q];
    if (!inside(i, j, k)) continue;
    const int di[6] = {-1, 1, 0, 0, 0, 0}, dj[6] = {0, 0, -1, 1, 0, 0}, dk[6] = {0, 0, 0, 0, -1, 1};
    T v[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
      const int x = i + di[d], y = j + dj[d], z = k + dk[d];
      v[d] = inside(x, y, z) ? red[((z - t.z0 + 1) * FEY + (y - t.y0 + 1)) * FEX + (x - t.x0 + 1)]
                             : (zl ? T(0) : bg[q][d]);
    }
    const Coef<T> c = cell_coef<T>(L, i, j, k);
    const T sum = c.aw * v[0] + c.ae * v[1] + c.as * v[2] + c.an * v[3] + c.ab * v[4] +
                  c.at * v[5];
    *(gmem<T> *)(Bd + aD1 + at(i, j, k)) = div_ap(brhs[q] - sum, c.ap, T(1) / c.ap);
  }
#pragma unroll
  for (int q = 0; q < F_RED_PER; ++q) {
    const int i = ri[q], j = rj[q], k = rk[q];
    if (tid + q * NTHREADS < F_RED && i >= t.x0 && i < t.x0 + FX && i < ni && j >= t.y0 &&
        j < t.y0 + FY && j < nj && k >= t.z0 && k < t.z0 + FZ && k < nk)
      *(gmem<T> *)(Bd + aD0 + at(i, j, k)) = rval[q];
  }
}

// ---------------------------------------------------------------------------
// mgrestr (mgsolver.for:693-756): fine residual
//   resf = rhs - (aw*W + ae*E + as*S + an*N + ab*B + at*T + ap*P)
// and coarse rhs = 0.125 * (sum of the 8 children, i fastest, then j, k).
// The coarse phi is not touched: the next relax of the coarse level runs with
// zero = 1, which is the zeroing of mgsolver.for:655-662.
//
// The first half-sweep of a relaxation from zero, for a red coarse cell: every
// neighbour is zero, so the sweep's sum is exactly zero and the cell becomes
// div_ap(rhs - 0, ap, 1/ap) with cell_coef's ap, which the restriction writes
// with the rhs (zbuf >= 0: into buffer zbuf), in the sweep's arithmetic.
template <class Tc>
__device__ __forceinline__ void zero_state_red(const LevelDev &C, int zbuf, int ic, int jc, int kc,
                                               unsigned q, Tc rhs) {
  if (zbuf < 0 || ((ic + jc + kc) & 1)) return;
  const Coef<Tc> c = cell_coef<Tc>(C, ic, jc, kc);
  ((gmem<Tc> *)C.phi[zbuf][0])[q] = div_ap(rhs - Tc(0), c.ap, Tc(1) / c.ap);
}

// Restriction tile of an FP64 fine level, KBC coarse planes (2*KBC fine
// planes): a rolling window
// of fine planes k-1, k, k+1 of rows 2jc and 2jc+1.  A lane holds two coarse
// cells l, l+1 of a coarse row: fine entries l, l+1 of both colours (cells
// 2l .. 2l+3), each array one vector load.  Neighbours are addressed by the
// coordinates of their image as 32-bit byte offsets from the fine level's
// arena base, so tiles touching a y/z face (BND) run the same code; their
// coefficients carry cell_coef's zeros towards walls.  Each coarse cell's
// eight residuals arrive in the reference's summation order (a, b of row
// 2jc, a, b of row 2jc+1; plane 2kc, then 2kc+1).  C.lpr1 lanes cover a
// coarse row segment; the tile's `lev` is the coarse descriptor.
template <class Tf, class Tc, int KBC, bool BND>
__device__ __forceinline__ void restrict_tile_two(const Tile &t, const LevelDev &C,
                                                  const LevelDev &F, int src_f, bool on,
                                                  int zbuf) {
  typedef typename Vec2<Tf>::type v2;
  constexpr unsigned SZ = sizeof(Tf);
  constexpr int NF = 2 * KBC;
  const int lpr = C.lpr1;
  const int lane = (int)threadIdx.x;
  const int rl = lane & (lpr - 1);
  const int l = t.x0 + 2 * rl;  // coarse cells l, l+1 = fine entries l, l+1
  const int jc = t.y0 + (int)threadIdx.y * (64 / lpr) + lane / lpr;
  const int ni = F.ni, nj = F.nj, nk = F.nk, H = F.H;
  // on: the block's flag gates the stores
  const bool ok0 = on && l < C.ni && jc < C.nj, ok1 = on && l + 1 < C.ni && jc < C.nj;
  const bool px = F.per & 1, py = F.per & 2, pz = F.per & 4;
  const unsigned row = (unsigned)H * SZ, plane = row * (unsigned)nj;  // bytes
  const int j0 = 2 * min(jc, C.nj - 1);  // even: a has colour (r + k) & 1 in row j0 + r
  // lanes past the row's last pair read that pair (nothing of theirs is stored)
  const unsigned ent = (unsigned)min(l, H - 2) * SZ;
  const char *base = (const char *)F.phi[0][0];
  auto aoff = [&](const void *a) -> unsigned { return (unsigned)((const char *)a - base); };
  const unsigned aP0 = aoff(F.phi[src_f][0]), dP = aoff(F.phi[src_f][1]) - aP0;
  const unsigned aR0 = aoff(F.rhs[0]), dR = aoff(F.rhs[1]) - aR0;
  const gmem<char> *Bs = (const gmem<char> *)base;
  auto ld = [&](unsigned off) -> Tf { return *(const gmem<Tf> *)(Bs + off); };
  auto ld2 = [&](unsigned off) -> v2 { return *(const gmem<v2> *)(Bs + off); };
  // colour array (parity & 1) of phi and of rhs, computed (not selected)
  auto pcol = [=](int par) -> unsigned { return aP0 + (unsigned)(par & 1) * dP; };
  auto rcol = [=](int par) -> unsigned { return aR0 + (unsigned)(par & 1) * dR; };
  const int x0 = 2 * l;  // fine cells x0 (a0), x0+1 (b0), x0+2 (a1), x0+3 (b1)
  // rows j0-1 and j0+2, ghosts resolved to their image row
  const int ys = BND ? image(j0 - 1, nj, py) : j0 - 1;
  const int yn = BND ? image(j0 + 2, nj, py) : j0 + 2;
  const unsigned o_r0 = (unsigned)j0 * row + ent, o_r1 = o_r0 + row;
  const unsigned o_s = (unsigned)ys * row + ent, o_n = (unsigned)yn * row + ent;
  // x neighbours no lane of the segment holds: fine cell x0-1 for the first
  // lane (the west ghost's image or the previous tile's last entry); for the
  // lane holding the last coarse cell (or lying past it) the east ghost's
  // image, for the segment's last lane otherwise the next tile's first entry
  const bool w_edge = rl == 0;
  const bool e_last = l + 1 >= C.ni - 1;
  const bool e_edge = e_last || rl == lpr - 1;
  const int xw = image(x0 - 1, ni, px);
  const int xe = e_last ? image(ni, ni, px) : x0 + 4;
  const unsigned w_ent = (unsigned)(xw >> 1) * SZ, e_ent = (unsigned)(xe >> 1) * SZ;
  const int w_par = xw & 1, e_par = xe & 1;
  const int x1_par = w_edge ? w_par : e_par;
  const unsigned x1_ent = w_edge ? w_ent : e_ent;
  const Tf cx = (Tf)F.cx, cy = (Tf)F.cy, cz = (Tf)F.cz;
  const Tf aw_a0 = (x0 == 0 && !px) ? Tf(0) : cx;
  const Tf ae_a0 = (x0 == ni - 1 && !px) ? Tf(0) : cx;
  const Tf ae_b0 = (x0 + 1 == ni - 1 && !px) ? Tf(0) : cx;
  const Tf ae_a1 = (x0 + 2 == ni - 1 && !px) ? Tf(0) : cx;
  const Tf ae_b1 = (x0 + 3 == ni - 1 && !px) ? Tf(0) : cx;
  const Tf as0 = (BND && !py && j0 == 0) ? Tf(0) : cy;       // row j0 towards j0-1
  const Tf an1 = (BND && !py && j0 + 1 == nj - 1) ? Tf(0) : cy;  // row j0+1 towards j0+2

  struct PlaneSet {
    v2 fa0, fb0, fa1, fb1;  // rhs of the a and b pairs in rows j0, j0+1
    v2 sa, sb, na, nb;      // rows j0-1 and j0+2
    v2 wa0, wb0, wa1, wb1;  // window plane k+1 of rows j0, j0+1
    Tf we0, we1, ee0, ee1;  // x-edge values of rows j0, j0+1
  };
  auto next_plane = [&](int k) -> int { return BND ? image(k + 1, nk, pz) : k + 1; };
  auto load_set = [&](int k) -> PlaneSet {
    PlaneSet q;
    const unsigned o = (unsigned)k * plane;
    const int c0 = k & 1;  // colour of a in row j0 (j0 even)
    q.fa0 = ld2(rcol(c0) + o + o_r0);
    q.fb0 = ld2(rcol(c0 + 1) + o + o_r0);
    q.fa1 = ld2(rcol(c0 + 1) + o + o_r1);
    q.fb1 = ld2(rcol(c0) + o + o_r1);
    q.sa = ld2(pcol(ys + k) + o + o_s);
    q.sb = ld2(pcol(ys + k + 1) + o + o_s);
    q.na = ld2(pcol(yn + k) + o + o_n);
    q.nb = ld2(pcol(yn + k + 1) + o + o_n);
    const int kp = next_plane(k);
    const unsigned op = (unsigned)kp * plane;
    q.wa0 = ld2(pcol(kp) + op + o_r0);
    q.wb0 = ld2(pcol(kp + 1) + op + o_r0);
    q.wa1 = ld2(pcol(kp + 1) + op + o_r1);
    q.wb1 = ld2(pcol(kp) + op + o_r1);
    // x-edge values: only the lanes at a row segment's ends load them
    const unsigned r0 = (unsigned)j0 * row;
    Tf x10 = Tf(0), x11 = Tf(0), x20 = Tf(0), x21 = Tf(0);
    if (w_edge || e_edge) {
      x10 = ld(pcol(x1_par + k) + o + r0 + x1_ent);
      x11 = ld(pcol(x1_par + k + 1) + o + r0 + row + x1_ent);
    }
    if (w_edge && e_edge) {
      x20 = ld(pcol(e_par + k) + o + r0 + e_ent);
      x21 = ld(pcol(e_par + k + 1) + o + r0 + row + e_ent);
    }
    q.we0 = x10;
    q.we1 = x11;
    q.ee0 = w_edge ? x20 : x10;
    q.ee1 = w_edge ? x21 : x11;
    return q;
  };
  const int k0 = 2 * t.z0;
  v2 ma0, mb0, ma1, mb1, ca0, cb0, ca1, cb1;  // window planes k-1 and k
  {
    const int km = BND ? image(k0 - 1, nk, pz) : k0 - 1;
    const unsigned om = (unsigned)km * plane, oc = (unsigned)k0 * plane;
    ma0 = ld2(pcol(km) + om + o_r0);
    mb0 = ld2(pcol(km + 1) + om + o_r0);
    ma1 = ld2(pcol(km + 1) + om + o_r1);
    mb1 = ld2(pcol(km) + om + o_r1);
    ca0 = ld2(pcol(k0) + oc + o_r0);
    cb0 = ld2(pcol(k0 + 1) + oc + o_r0);
    ca1 = ld2(pcol(k0 + 1) + oc + o_r1);
    cb1 = ld2(pcol(k0) + oc + o_r1);
  }
  // the four cells of one fine row: A/B its a and b pairs, S/N the pairs of
  // the rows below and above, M/W planes k-1 and k+1, F the rhs; adds the
  // residuals of cells a0, b0 to acc0 and a1, b1 to acc1
  auto row_cells = [&](const v2 &A, const v2 &B, const v2 &Sa, const v2 &Sb, const v2 &Na,
                       const v2 &Nb, const v2 &Ma, const v2 &Mb, const v2 &Wa, const v2 &Wb,
                       const v2 &Fa, const v2 &Fb, Tf we, Tf ee, Tf as, Tf an, Tf ab, Tf at,
                       bool first, Tc &acc0, Tc &acc1) {
    // cell x0-1 is lane-1's b1, cell x0+4 is lane+1's a0
    const Tf lf = wave_shift(B.y, 1), rt = wave_shift(A.x, 0);
    const Tf W_a0 = w_edge ? we : lf;
    const Tf E_b0 = l + 1 < C.ni ? A.y : ee;
    const Tf E_b1 = (l + 2 < C.ni && !e_edge) ? rt : ee;
    const Tf ap_a0 = Tf(-1) * (aw_a0 + ae_a0 + as + an + ab + at);
    const Tf ap_b0 = Tf(-1) * (cx + ae_b0 + as + an + ab + at);
    const Tf ap_a1 = Tf(-1) * (cx + ae_a1 + as + an + ab + at);
    const Tf ap_b1 = Tf(-1) * (cx + ae_b1 + as + an + ab + at);
    const Tf r_a0 = Fa.x - (aw_a0 * W_a0 + ae_a0 * B.x + as * Sa.x + an * Na.x + ab * Ma.x +
                            at * Wa.x + ap_a0 * A.x);
    const Tf r_b0 = Fb.x - (cx * A.x + ae_b0 * E_b0 + as * Sb.x + an * Nb.x + ab * Mb.x +
                            at * Wb.x + ap_b0 * B.x);
    const Tf r_a1 = Fa.y - (cx * B.x + ae_a1 * B.y + as * Sa.y + an * Na.y + ab * Ma.y +
                            at * Wa.y + ap_a1 * A.y);
    const Tf r_b1 = Fb.y - (cx * A.y + ae_b1 * E_b1 + as * Sb.y + an * Nb.y + ab * Mb.y +
                            at * Wb.y + ap_b1 * B.y);
    acc0 = first ? (Tc)r_a0 : acc0 + (Tc)r_a0;
    acc0 = acc0 + (Tc)r_b0;
    acc1 = first ? (Tc)r_a1 : acc1 + (Tc)r_a1;
    acc1 = acc1 + (Tc)r_b1;
  };
  Tc acc0 = Tc(0), acc1 = Tc(0);
#pragma unroll
  for (int p = 0; p < NF; ++p) {
    const int k = k0 + p;
    const PlaneSet cur = load_set(k);
    const Tf ab = (BND && !pz && k == 0) ? Tf(0) : cz;
    const Tf at = (BND && !pz && k == nk - 1) ? Tf(0) : cz;
    // row j0: S from row j0-1, N = row j0+1 of the window
    row_cells(ca0, cb0, cur.sa, cur.sb, ca1, cb1, ma0, mb0, cur.wa0, cur.wb0, cur.fa0, cur.fb0,
              cur.we0, cur.ee0, as0, cy, ab, at, !(p & 1), acc0, acc1);
    // row j0+1: S = row j0 of the window, N from row j0+2
    row_cells(ca1, cb1, ca0, cb0, cur.na, cur.nb, ma1, mb1, cur.wa1, cur.wb1, cur.fa1, cur.fb1,
              cur.we1, cur.ee1, cy, an1, ab, at, false, acc0, acc1);
    if (p & 1) {
      // coarse cells l, l+1 (l even): entry l/2 of the two colour arrays
      const int kc = k >> 1;
      const unsigned q =
          ((unsigned)kc * (unsigned)C.nj + (unsigned)jc) * (unsigned)C.H + (unsigned)(l >> 1);
      const int c0 = (l + jc + kc) & 1;
      const Tc v0 = Tc(0.125) * acc0, v1 = Tc(0.125) * acc1;
      if (ok0) {
        ((gmem<Tc> *)C.rhs[c0])[q] = v0;
        zero_state_red<Tc>(C, zbuf, l, jc, kc, q, v0);
      }
      if (ok1) {
        ((gmem<Tc> *)C.rhs[c0 ^ 1])[q] = v1;
        zero_state_red<Tc>(C, zbuf, l + 1, jc, kc, q, v1);
      }
    }
    ma0 = ca0;
    mb0 = cb0;
    ma1 = ca1;
    mb1 = cb1;
    ca0 = cur.wa0;
    cb0 = cur.wb0;
    ca1 = cur.wa1;
    cb1 = cur.wb1;
  }
}

// Restriction tile of an FP32 fine level (mixed precision), KBC coarse planes
// (2*KBC fine planes), software-
// pipelined over the fine planes: a rolling window of fine planes k-1, k, k+1
// of rows 2jc and 2jc+1, with the loads of plane k+1 (rhs of both rows, rows
// 2jc-1 and 2jc+2, x-edge values and window plane k+2) issued before plane k
// is computed.  Neighbours are addressed by the coordinates of their image
// as 32-bit byte offsets from the fine level's arena base, so tiles touching
// a y/z face (BND) run the same code; their coefficients carry cell_coef's
// zeros towards walls.  The eight residuals arrive in the reference's
// summation order (a, b of row 2jc, a, b of row 2jc+1; plane 2kc, then
// 2kc+1).  One coarse cell per lane (FP32 fine entries: four-byte loads; two
// cells per lane cost more occupancy than the loads they save here); C.lpr1
// lanes cover a coarse row; the tile's `lev` is the coarse descriptor.
template <class Tf, class Tc, int KBC, bool BND>
__device__ __forceinline__ void restrict_tile_one(const Tile &t, const LevelDev &C,
                                                  const LevelDev &F, int src_f, bool on,
                                                  int zbuf) {
  constexpr unsigned SZ = sizeof(Tf);
  constexpr int NF = 2 * KBC;
  const int lpr = C.lpr1;
  const int lane = (int)threadIdx.x;
  const int rl = lane & (lpr - 1);
  const int l = t.x0 + rl;
  const int jc = t.y0 + (int)threadIdx.y * (64 / lpr) + lane / lpr;
  const int ni = F.ni, nj = F.nj, nk = F.nk, H = F.H;
  const bool ok = on && l < C.ni && jc < C.nj;  // on: the block's flag gates the stores
  const bool px = F.per & 1, py = F.per & 2, pz = F.per & 4;
  const unsigned row = (unsigned)H * SZ, plane = row * (unsigned)nj;  // bytes
  const int j0 = 2 * min(jc, C.nj - 1);  // even: a has colour (r + k) & 1 in row j0 + r
  const unsigned ent = (unsigned)min(l, C.ni - 1) * SZ;
  const char *base = (const char *)F.phi[0][0];
  auto aoff = [&](const void *a) -> unsigned { return (unsigned)((const char *)a - base); };
  const unsigned aP0 = aoff(F.phi[src_f][0]), dP = aoff(F.phi[src_f][1]) - aP0;
  const unsigned aR0 = aoff(F.rhs[0]), dR = aoff(F.rhs[1]) - aR0;
  const gmem<char> *Bs = (const gmem<char> *)base;
  auto ld = [&](unsigned off) -> Tf { return *(const gmem<Tf> *)(Bs + off); };
  // colour array (parity & 1) of phi and of rhs, computed (not selected)
  auto pcol = [=](int par) -> unsigned { return aP0 + (unsigned)(par & 1) * dP; };
  auto rcol = [=](int par) -> unsigned { return aR0 + (unsigned)(par & 1) * dR; };
  const int ia = 2 * l, ib = ia + 1;
  // rows j0-1 and j0+2, ghosts resolved to their image row
  const int ys = BND ? image(j0 - 1, nj, py) : j0 - 1;
  const int yn = BND ? image(j0 + 2, nj, py) : j0 + 2;
  const unsigned o_r0 = (unsigned)j0 * row + ent, o_r1 = o_r0 + row;
  const unsigned o_s = (unsigned)ys * row + ent, o_n = (unsigned)yn * row + ent;
  // x neighbours not held by the adjacent lane: x-ghost image or the next
  // tile's entry
  const bool w_edge = ia == 0 || (rl == 0 && l > 0);
  const bool e_edge = ib + 1 >= ni || (rl == lpr - 1 && l + 1 < C.ni);
  const int xw = image(ia - 1, ni, px), xe = image(ib + 1, ni, px);
  const unsigned w_ent = w_edge ? (unsigned)(xw >> 1) * SZ : ent;
  const unsigned e_ent = e_edge ? (unsigned)(xe >> 1) * SZ : ent;
  const int w_par = w_edge ? (xw & 1) : 0, e_par = e_edge ? (xe & 1) : 0;
  const int x1_par = w_edge ? w_par : e_par;
  const unsigned x1_ent = w_edge ? w_ent : e_ent;
  const Tf cx = (Tf)F.cx, cy = (Tf)F.cy, cz = (Tf)F.cz;
  const Tf aw_a = (ia == 0 && !px) ? Tf(0) : cx;
  const Tf ae_a = (ia == ni - 1 && !px) ? Tf(0) : cx;
  const Tf ae_b = (ib == ni - 1 && !px) ? Tf(0) : cx;
  const Tf as0 = (BND && !py && j0 == 0) ? Tf(0) : cy;       // row j0 towards j0-1
  const Tf an1 = (BND && !py && j0 + 1 == nj - 1) ? Tf(0) : cy;  // row j0+1 towards j0+2

  struct PlaneSet {
    Tf fa0, fb0, fa1, fb1;  // rhs of a, b in rows j0, j0+1
    Tf sa, sb, na, nb;      // rows j0-1 and j0+2
    Tf wa0, wb0, wa1, wb1;  // window plane k+1 of rows j0, j0+1
    Tf we0, we1, ee0, ee1;  // x-edge values of rows j0, j0+1
  };
  auto next_plane = [&](int k) -> int { return BND ? image(k + 1, nk, pz) : k + 1; };
  auto load_set = [&](int k) -> PlaneSet {
    PlaneSet q;
    const unsigned o = (unsigned)k * plane;
    const int c0 = k & 1;  // colour of a in row j0 (j0 even)
    q.fa0 = ld(rcol(c0) + o + o_r0);
    q.fb0 = ld(rcol(c0 + 1) + o + o_r0);
    q.fa1 = ld(rcol(c0 + 1) + o + o_r1);
    q.fb1 = ld(rcol(c0) + o + o_r1);
    q.sa = ld(pcol(ys + k) + o + o_s);
    q.sb = ld(pcol(ys + k + 1) + o + o_s);
    q.na = ld(pcol(yn + k) + o + o_n);
    q.nb = ld(pcol(yn + k + 1) + o + o_n);
    const int kp = next_plane(k);
    const unsigned op = (unsigned)kp * plane;
    q.wa0 = ld(pcol(kp) + op + o_r0);
    q.wb0 = ld(pcol(kp + 1) + op + o_r0);
    q.wa1 = ld(pcol(kp + 1) + op + o_r1);
    q.wb1 = ld(pcol(kp) + op + o_r1);
    // x-edge values: only the lanes at a row segment's ends use them (the
    // others read no x-edge value at all)
    const unsigned r0 = (unsigned)j0 * row;
    Tf x10 = Tf(0), x11 = Tf(0), x20 = Tf(0), x21 = Tf(0);
    if (w_edge || e_edge) {
      x10 = ld(pcol(x1_par + k) + o + r0 + x1_ent);
      x11 = ld(pcol(x1_par + k + 1) + o + r0 + row + x1_ent);
    }
    if (w_edge && e_edge) {
      x20 = ld(pcol(e_par + k) + o + r0 + e_ent);
      x21 = ld(pcol(e_par + k + 1) + o + r0 + row + e_ent);
    }
    q.we0 = x10;
    q.we1 = x11;
    q.ee0 = w_edge ? x20 : x10;
    q.ee1 = w_edge ? x21 : x11;
    return q;
  };
  const int k0 = 2 * t.z0;
  Tf ma0, mb0, ma1, mb1, ca0, cb0, ca1, cb1;  // window planes k-1 and k
  {
    const int km = BND ? image(k0 - 1, nk, pz) : k0 - 1;
    const unsigned om = (unsigned)km * plane, oc = (unsigned)k0 * plane;
    ma0 = ld(pcol(km) + om + o_r0);
    mb0 = ld(pcol(km + 1) + om + o_r0);
    ma1 = ld(pcol(km + 1) + om + o_r1);
    mb1 = ld(pcol(km) + om + o_r1);
    ca0 = ld(pcol(k0) + oc + o_r0);
    cb0 = ld(pcol(k0 + 1) + oc + o_r0);
    ca1 = ld(pcol(k0 + 1) + oc + o_r1);
    cb1 = ld(pcol(k0) + oc + o_r1);
  }
  PlaneSet cur = load_set(k0);
  Tc acc = Tc(0);
#pragma unroll
  for (int p = 0; p < NF; ++p) {
    const int k = k0 + p;
    PlaneSet nxt;
    if (p + 1 < NF) nxt = load_set(k + 1);
    const Tf ab = (BND && !pz && k == 0) ? Tf(0) : cz;
    const Tf at = (BND && !pz && k == nk - 1) ? Tf(0) : cz;
    // row j0 (r = 0): S from row j0-1, N = row j0+1 of the window
    {
      const Tf Pa = ca0, Pb = cb0;
      // cell 2l-1 is lane l-1's b, cell 2l+2 is lane l+1's a
      const Tf lf = wave_shift(Pb, 1), rt = wave_shift(Pa, 0);
      const Tf Wa = w_edge ? cur.we0 : lf;
      const Tf Ea = ib < ni ? Pb : cur.ee0;
      const Tf Eb = e_edge ? cur.ee0 : rt;
      const Tf ap_a = Tf(-1) * (aw_a + ae_a + as0 + cy + ab + at);
      const Tf ap_b = Tf(-1) * (cx + ae_b + as0 + cy + ab + at);
      const Tf ra = cur.fa0 - (aw_a * Wa + ae_a * Ea + as0 * cur.sa + cy * ca1 + ab * ma0 +
                               at * cur.wa0 + ap_a * Pa);
      const Tf rb = cur.fb0 - (cx * Pa + ae_b * Eb + as0 * cur.sb + cy * cb1 + ab * mb0 +
                               at * cur.wb0 + ap_b * Pb);
      acc = (p & 1) ? acc + (Tc)ra : (Tc)ra;
      acc = acc + (Tc)rb;
    }
    // row j0+1 (r = 1): S = row j0 of the window, N from row j0+2
    {
      const Tf Pa = ca1, Pb = cb1;
      const Tf lf = wave_shift(Pb, 1), rt = wave_shift(Pa, 0);
      const Tf Wa = w_edge ? cur.we1 : lf;
      const Tf Ea = ib < ni ? Pb : cur.ee1;
      const Tf Eb = e_edge ? cur.ee1 : rt;
      const Tf ap_a = Tf(-1) * (aw_a + ae_a + cy + an1 + ab + at);
      const Tf ap_b = Tf(-1) * (cx + ae_b + cy + an1 + ab + at);
      const Tf ra = cur.fa1 - (aw_a * Wa + ae_a * Ea + cy * ca0 + an1 * cur.na + ab * ma1 +
                               at * cur.wa1 + ap_a * Pa);
      const Tf rb = cur.fb1 - (cx * Pa + ae_b * Eb + cy * cb0 + an1 * cur.nb + ab * mb1 +
                               at * cur.wb1 + ap_b * Pb);
      acc = acc + (Tc)ra;
      acc = acc + (Tc)rb;
    }
    if (p & 1) {
      const int kc = k >> 1;
      if (ok) {
        gmem<Tc> *D = (gmem<Tc> *)C.rhs[(l + jc + kc) & 1];
        const unsigned q =
            ((unsigned)kc * (unsigned)C.nj + (unsigned)jc) * (unsigned)C.H + (unsigned)(l >> 1);
        const Tc v = Tc(0.125) * acc;
        D[q] = v;
        zero_state_red<Tc>(C, zbuf, l, jc, kc, q, v);
      }
    }
    ma0 = ca0;
    mb0 = cb0;
    ma1 = ca1;
    mb1 = cb1;
    ca0 = cur.wa0;
    cb0 = cur.wb0;
    ca1 = cur.wa1;
    cb1 = cur.wb1;
    if (p + 1 < NF) cur = nxt;
  }
}

// The restriction tile of the fine level's storage type: two coarse cells per
// lane on an FP64 fine level (16-byte loads), one on an FP32 one; the tile
// geometry follows (restriction tiles span 2*lpr1 or lpr1 coarse cells).
template <class Tf, class Tc, int KBC, bool BND>
__device__ __forceinline__ void restrict_tile_fast(const Tile &t, const LevelDev &C,
                                                   const LevelDev &F, int src_f, bool on,
                                                   int zbuf) {
  if constexpr (sizeof(Tf) == 8)
    restrict_tile_two<Tf, Tc, KBC, BND>(t, C, F, src_f, on, zbuf);
  else
    restrict_tile_one<Tf, Tc, KBC, BND>(t, C, F, src_f, on, zbuf);
}

// One launch per coarse level: tiles of KBC coarse planes, kind 0 when every
// fine row and plane the tile reads is inside the block, kind 3 otherwise.
// ONE, scalar arguments: as k_rb_sweep (one_lev the coarse descriptor,
// one_lev2 its finer level).
template <class Tf, class Tc, int KBC, bool ONE>
__global__ void __launch_bounds__(NTHREADS)
    k_restrict(const Tile *__restrict__ tiles, const LevelDev *__restrict__ levs,
               const int *__restrict__ active, int src_f, int one_lev, int one_lev2, int one_blk,
               int ntiles, int zbuf) {
  const Tile t = tiles[xcd_remap((int)blockIdx.x, ntiles)];
  if (ONE) {
    const LevelDev &C = levs[one_lev], &F = levs[one_lev2];
    const int on = active[one_blk];
    asm volatile("" ::"s"(t.x0), "s"(t.y0), "s"(t.z0), "s"(t.kind), "s"(on), "s"(C.ni), "s"(C.nj),
                 "s"(C.H), "s"(C.lpr1), "s"(C.rhs[0]), "s"(C.rhs[1]), "s"(F.ni), "s"(F.nj),
                 "s"(F.nk), "s"(F.H), "s"(F.per), "s"(F.cx), "s"(F.cy), "s"(F.cz),
                 "s"(F.phi[0][0]), "s"(F.phi[0][1]), "s"(F.phi[1][0]), "s"(F.phi[1][1]),
                 "s"(F.rhs[0]), "s"(F.rhs[1]));
    if (t.kind == 0)
      restrict_tile_fast<Tf, Tc, KBC, false>(t, C, F, src_f, on != 0, zbuf);
    else
      restrict_tile_fast<Tf, Tc, KBC, true>(t, C, F, src_f, on != 0, zbuf);
    return;
  }
  const LevelDev &C = levs[t.lev];
  if (!active[C.blk]) return;
  const LevelDev &F = levs[C.finer];
  if (t.kind == 0)
    restrict_tile_fast<Tf, Tc, KBC, false>(t, C, F, src_f, true, zbuf);
  else
    restrict_tile_fast<Tf, Tc, KBC, true>(t, C, F, src_f, true, zbuf);
}

// Restriction for blocks with neighbours: one thread per coarse cell of
// natural tiles (TILE_X x TILE_Y cells of one coarse plane); the fine
// residual of each of the eight children, its neighbours read through
// resolve_cell where they are ghosts, in restrict_tile_fast's arithmetic:
// r = f - (aw*W + ae*E + as*S + an*N + ab*B + at*T + ap*P), the children
// summed i fastest, then j, then k, times 0.125 (mgsolver.for:693-756).
template <class Tf, class Tc>
__global__ void __launch_bounds__(NTHREADS)
    k_restrict_gen(const Tile *__restrict__ tiles, const LevelDev *__restrict__ levs,
                   const int *__restrict__ active, int src_f) {
  const Tile t = tiles[blockIdx.x];
  const LevelDev &C = levs[t.lev];
  if (!active[C.blk]) return;
  const int l = t.x0 + (int)threadIdx.x, jc = t.y0 + (int)threadIdx.y, kc = t.z0;
  if (l >= C.ni || jc >= C.nj || kc >= C.nk) return;
  const int fd = C.finer;
  const LevelDev &F = levs[fd];
  auto val = [&](int x, int y, int z) -> Tf {
    if ((unsigned)x < (unsigned)F.ni && (unsigned)y < (unsigned)F.nj &&
        (unsigned)z < (unsigned)F.nk)
      return ((const Tf *)F.phi[src_f][(x + y + z) & 1])[cidx(F.H, F.nj, x, y, z)];
    return level_value<Tf>(levs, fd, src_f, x, y, z);
  };
  Tc acc = Tc(0);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int i = 2 * l + (q & 1), j = 2 * jc + ((q >> 1) & 1), k = 2 * kc + (q >> 2);
    const Coef<Tf> c = cell_coef<Tf>(F, i, j, k);
    const Tf f = ((const Tf *)F.rhs[(i + j + k) & 1])[cidx(F.H, F.nj, i, j, k)];
    const Tf r = f - (c.aw * val(i - 1, j, k) + c.ae * val(i + 1, j, k) + c.as * val(i, j - 1, k) +
                      c.an * val(i, j + 1, k) + c.ab * val(i, j, k - 1) + c.at * val(i, j, k + 1) +
                      c.ap * val(i, j, k));
    acc = q == 0 ? (Tc)r : acc + (Tc)r;
  }
  ((Tc *)C.rhs[(l + jc + kc) & 1])[cidx(C.H, C.nj, l, jc, kc)] = Tc(0.125) * acc;
}

// ---------------------------------------------------------------------------
// mgcorr (mgsolver.for:828-896): fine += trilinear interpolation of coarse.
// Fine padded index ip = 2*icp + d with icp the coarse padded index; the
// weight of coarse corner (oi,oj,ok) is prod (3 if o==d else 1); terms are
// summed in the order c, c+j, c+k, c+k+j, c+i, c+i+j, c+i+k, c+i+j+k and the
// sum divided by 64.  Ghost corners are virtual (images of coarse interior).
//
// A thread owns fine entry l of both colours (cells 2l and 2l+1) in fine rows
// 2m-1, 2m and planes 2n-1, 2n (coarse row m and plane n in padded numbering):
// those eight fine cells share the twelve coarse corners x in {l-1, l, l+1},
// y in {m-1, m}, z in {n-1, n} (interior numbering), which are loaded once.
// Tile::x0 = first l, y0 = first m (TILE_Y per tile), z0 = n.  The fine
// level is updated in place (each fine cell by exactly one thread).
// GEN: blocks with neighbours -- the coarse ghost corners (edges and corners
// included) are read through resolve_cell, possibly in another block.
// ONE, scalar arguments: as k_rb_sweep (one_lev the fine descriptor,
// one_lev2 its coarser level).
template <class Tf, class Tc, bool GEN, bool ONE>
__global__ void __launch_bounds__(NTHREADS)
    k_prolong(const Tile *__restrict__ tiles, const LevelDev *__restrict__ levs,
              const int *__restrict__ active, int src_f, int src_c, int one_lev, int one_lev2,
              int one_blk, int ntiles) {
  const Tile t = tiles[xcd_remap((int)blockIdx.x, ntiles)];
  const LevelDev &F = levs[ONE ? one_lev : t.lev];
  const int on = ONE ? active[one_blk] : 1;
  if (!ONE && !active[F.blk]) return;
  const LevelDev &C = levs[ONE ? one_lev2 : F.coarser];
  if (ONE)
    asm volatile("" ::"s"(t.x0), "s"(t.y0), "s"(t.z0), "s"(on), "s"(F.ni), "s"(F.nj), "s"(F.nk),
                 "s"(F.H), "s"(F.phi[0][0]), "s"(F.phi[0][1]), "s"(F.phi[1][0]), "s"(F.phi[1][1]),
                 "s"(C.ni), "s"(C.nj), "s"(C.nk), "s"(C.H), "s"(C.per), "s"(C.phi[0][0]),
                 "s"(C.phi[0][1]), "s"(C.phi[1][0]), "s"(C.phi[1][1]));
  // a thread holds fine entries l, l+1 (l even): cells 2l .. 2l+3 of the
  // 2x2 rows and planes around coarse row pair m, plane pair n
  typedef typename Vec2<Tf>::type v2;
  const int l = t.x0 + 2 * (int)threadIdx.x;
  const int m = t.y0 + (int)threadIdx.y;
  const int n = t.z0;
  if (2 * l >= F.ni || m > C.nj || n > C.nk) return;
  const bool two = 2 * l + 2 < F.ni;  // entry l+1 holds cells (F.ni is even)
  const bool px = C.per & 1, py = C.per & 2, pz = C.per & 4;
  const gmem<Tc> *C0 = (const gmem<Tc> *)C.phi[src_c][0];
  const gmem<Tc> *C1 = (const gmem<Tc> *)C.phi[src_c][1];
  // coarse corner coordinates x = l-1 .. l+2 (interior numbering, ghosts
  // resolved; the last is read only when entry l+1 holds cells)
  int xc[4], yc[2], zc[2];
  for (int a = 0; a < 4; ++a) xc[a] = image(min(l - 1 + a, two ? l + 2 : l + 1), C.ni, px);
  for (int a = 0; a < 2; ++a) {
    yc[a] = image(m - 1 + a, C.nj, py);
    zc[a] = image(n - 1 + a, C.nk, pz);
  }
  Tc cv[4][2][2];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        if (GEN) {
          cv[a][b][c] = level_value<Tc>(levs, F.coarser, src_c, min(l - 1 + a, two ? l + 2 : l + 1),
                                        m - 1 + b, n - 1 + c);
        } else {
          const unsigned off = ((unsigned)zc[c] * (unsigned)C.nj + (unsigned)yc[b]) *
                                   (unsigned)C.H + (unsigned)(xc[a] >> 1);
          cv[a][b][c] = ((xc[a] + yc[b] + zc[c]) & 1) ? C1[off] : C0[off];
        }
      }
  gmem<Tf> *F0 = (gmem<Tf> *)F.phi[src_f][0];
  gmem<Tf> *F1 = (gmem<Tf> *)F.phi[src_f][1];
#pragma unroll
  for (int dk = 0; dk < 2; ++dk) {
    const int k = 2 * n - 1 + dk;
    if (k < 0 || k >= F.nk) continue;
#pragma unroll
    for (int dj = 0; dj < 2; ++dj) {
      const int j = 2 * m - 1 + dj;
      if (j < 0 || j >= F.nj) continue;
      const unsigned fq = ((unsigned)k * (unsigned)F.nj + (unsigned)j) * (unsigned)F.H + (unsigned)l;
      // the correction of cell 2(l+ee)+e: corners x = l+ee-1+e, l+ee+e with
      // di = 1-e, summed in the corner order c, c+j, c+k, c+k+j, c+i, ...
      Tc cr[2][2];  // [e][ee]
#pragma unroll
      for (int e = 0; e < 2; ++e)
#pragma unroll
        for (int ee = 0; ee < 2; ++ee) {
          const int di = 1 - e;
          Tc s;
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const int oi = q >> 2, oj = q & 1, ok = (q >> 1) & 1;
            const Tc w = (oi == di ? Tc(3) : Tc(1)) * (oj == dj ? Tc(3) : Tc(1)) *
                         (ok == dk ? Tc(3) : Tc(1));
            const Tc v = cv[ee + e + oi][oj][ok];
            s = (q == 0) ? w * v : s + w * v;
          }
          cr[e][ee] = s;
        }
      // cells 2l+e and 2l+2+e share a colour array: entries l, l+1
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        gmem<Tf> *fa = ((2 * l + e + j + k) & 1) ? F1 : F0;
        if (two) {
          v2 v = *(gmem<v2> *)(fa + fq);
          v.x = v.x + (Tf)(cr[e][0] / Tc(64));
          v.y = v.y + (Tf)(cr[e][1] / Tc(64));
          if (on) *(gmem<v2> *)(fa + fq) = v;
        } else {
          const Tf nv = fa[fq] + (Tf)(cr[e][0] / Tc(64));
          if (on) fa[fq] = nv;
        }
      }
    }
  }
}

template <class T>
__global__ void __launch_bounds__(NTHREADS)
    k_fill_zero(const Tile *__restrict__ tiles, const LevelDev *__restrict__ levs,
                const int *__restrict__ active, int buf) {
  const Tile t = tiles[blockIdx.x];
  const LevelDev &L = levs[t.lev];
  if (!active[L.blk]) return;
  const int i = t.x0 + (int)threadIdx.x;
  const int j = t.y0 + (int)threadIdx.y;
  const int k = t.z0;
  if (i >= L.ni || j >= L.nj || k >= L.nk) return;
  ((T *)L.phi[buf][(i + j + k) & 1])[cidx(L.H, L.nj, i, j, k)] = T(0);
}

// 64-lane wavefront + LDS reduction in a fixed order (deterministic).
__device__ __forceinline__ void block_reduce(double &ss, double &am) {
  for (int off = 32; off > 0; off >>= 1) {
    ss += __shfl_xor(ss, off, 64);
    am = fmax(am, __shfl_xor(am, off, 64));
  }
  __shared__ double s_ss[NTHREADS / 64], s_am[NTHREADS / 64];
  const int tid = (int)(threadIdx.x + blockDim.x * threadIdx.y);
  const int w = tid >> 6;
  if ((tid & 63) == 0) {
    s_ss[w] = ss;
    s_am[w] = am;
  }
  __syncthreads();
  if (tid == 0) {
    double a = s_ss[0], b = s_am[0];
    for (int q = 1; q < NTHREADS / 64; ++q) {
      a += s_ss[q];
      b = fmax(b, s_am[q]);
    }
    ss = a;
    am = b;
  }
}

// ---------------------------------------------------------------------------
// Finest-level residual r = f - A p in FP64 (the rhs of the next V-cycle; in
// the reference it is su = div(u* - dt grad p)/dt, press.for:2-144), stored
// into the level-1 rhs in the level's precision, with per-tile partial sum of
// r^2 and max |r| for the REL_L2 and REF_RMAX criteria.
//
// A lane holds entry l of both colour arrays of one row: the two adjacent
// cells a = 2l (colour ca = (j+k)&1) and b = 2l+1.  Planes k-1..k+1 of both
// entries stream through a register window, x-neighbours come from the lane's
// other cell or the adjacent lanes (DPP shift), rows j-1, j+1 are loaded.
// A residual tile spans 2*F.lpr colour entries of a row (64/F.lpr rows per
// wavefront; the fast kernel gives each lane two entries); this path walks
// it as two halves of F.lpr entries, one entry per lane (x0: the half's
// first entry), over the planes [t.z0, t.z1).  Tile kind 0: every row/plane
// of the tile has its y/z neighbours inside the block; kind 1: the general
// path with ghost images.  use_p = 0 gives r = f (the norm of the
// right-hand side).
template <class T, bool INNER>
__device__ __forceinline__ void residual_half(const Tile &t, int x0, const FineDev *__restrict__ fine,
                                              const LevelDev &L, int use_p, double &ss,
                                              double &am) {
  const FineDev &F = fine[t.lev];
  const int lpr = F.lpr;
  const int lane = (int)threadIdx.x;
  const int rl = lane & (lpr - 1);
  const int l = x0 + rl;
  const int j = t.y0 + (int)threadIdx.y * (64 / lpr) + lane / lpr;
  const int ni = F.ni, nj = F.nj, nk = F.nk, H = F.H;
  const bool ok = 2 * l < ni && j < nj;
  const bool b_ok = 2 * l + 1 < ni && j < nj;
  const unsigned row = (unsigned)H, plane = (unsigned)H * (unsigned)nj;
  // jc: the lane's row clamped like cbase -- lanes beyond the last row load
  // in-bounds rows and resolve in-bounds ghosts (their results are not
  // stored); for every stored cell jc == j
  const int jc = min(j, nj - 1);
  const unsigned cbase = (unsigned)jc * row + (unsigned)min(l, H - 1);
  const gmem<double> *P0 = (const gmem<double> *)F.p[0], *P1 = (const gmem<double> *)F.p[1];
  const gmem<double> *F0 = (const gmem<double> *)F.f[0], *F1 = (const gmem<double> *)F.f[1];
  gmem<T> *R0 = (gmem<T> *)L.rhs[0], *R1 = (gmem<T> *)L.rhs[1];
  const bool tile_left = l > 0 && rl == 0;             // entry l-1 in another tile
  const bool tile_right = rl == lpr - 1 && l + 1 < H;  // entry l+1 in another tile
  const bool px = F.per & 1;
  const double cx = F.cx, cy = F.cy, cz = F.cz;
  auto ldp = [&](int c, unsigned off) -> double { return c ? P1[off] : P0[off]; };
  auto own = [&](int k, int c) -> double {
    if (!INNER && (unsigned)k >= (unsigned)nk) return 0.0;
    return ldp(c, (unsigned)k * plane + cbase);
  };

  const int k1 = t.z1;
  int k = t.z0;
  double w0m = 0, w1m = 0, w0c = 0, w1c = 0;
  if (use_p) {
    w0m = own(k - 1, 0);
    w1m = own(k - 1, 1);
    w0c = own(k, 0);
    w1c = own(k, 1);
  }
  for (; k < k1; ++k) {
    const unsigned q = (unsigned)k * plane + cbase;
    const int ca = (j + k) & 1, cb = ca ^ 1;
    double ra = ca ? F1[q] : F0[q];
    double rb = cb ? F1[q] : F0[q];
    double w0p = 0, w1p = 0;
    if (use_p) {
      w0p = own(k + 1, 0);
      w1p = own(k + 1, 1);
      const double Pa = ca ? w1c : w0c, Pb = ca ? w0c : w1c;
      // x-neighbours: cell 2l-1 is lane-1's b, cell 2l+2 is lane+1's a
      double lf = wave_shift(Pb, 1), rt = wave_shift(Pa, 0);
      if (tile_left) lf = ldp(cb, q - 1);
      if (tile_right) rt = ldp(ca, q + 1);
      const int ia = 2 * l, ib = ia + 1;
      const double Wa = ia == 0 ? fine_value(fine, t.lev, -1, jc, k) : lf;
      const double Ea = ib < ni ? Pb : fine_value(fine, t.lev, ni, jc, k);
      const double Wb = Pa;
      const double Eb = ib + 1 < ni ? rt : fine_value(fine, t.lev, ni, jc, k);
      // y- and z-neighbours of a (colour cb there) and b (colour ca)
      double Sa, Na, Sb, Nb;
      if (INNER || jc > 0) {
        Sa = ldp(cb, q - row);
        Sb = ldp(ca, q - row);
      } else {
        Sa = fine_value(fine, t.lev, ia, -1, k);
        Sb = fine_value(fine, t.lev, ib, -1, k);
      }
      if (INNER || jc + 1 < nj) {
        Na = ldp(cb, q + row);
        Nb = ldp(ca, q + row);
      } else {
        Na = fine_value(fine, t.lev, ia, nj, k);
        Nb = fine_value(fine, t.lev, ib, nj, k);
      }
      const double Ba = (INNER || k > 0) ? (cb ? w1m : w0m) : fine_value(fine, t.lev, ia, jc, -1);
      const double Bb = (INNER || k > 0) ? (ca ? w1m : w0m) : fine_value(fine, t.lev, ib, jc, -1);
      const double Ta = (INNER || k + 1 < nk) ? (cb ? w1p : w0p) : fine_value(fine, t.lev, ia, jc, nk);
      const double Tb = (INNER || k + 1 < nk) ? (ca ? w1p : w0p) : fine_value(fine, t.lev, ib, jc, nk);
      Coef<double> c_a, c_b;
      if (INNER) {
        c_a.as = c_a.an = c_b.as = c_b.an = cy;
        c_a.ab = c_a.at = c_b.ab = c_b.at = cz;
        c_a.aw = (ia == 0 && !px) ? 0.0 : cx;
        c_a.ae = (ia == ni - 1 && !px) ? 0.0 : cx;
        c_b.aw = cx;
        c_b.ae = (ib == ni - 1 && !px) ? 0.0 : cx;
        c_a.ap = -1.0 * (c_a.aw + c_a.ae + c_a.as + c_a.an + c_a.ab + c_a.at);
        c_b.ap = -1.0 * (c_b.aw + c_b.ae + c_b.as + c_b.an + c_b.ab + c_b.at);
      } else {
        c_a = fine_coef(F, ia, j, k);
        c_b = fine_coef(F, ib, j, k);
      }
      ra = ra - (c_a.aw * Wa + c_a.ae * Ea + c_a.as * Sa + c_a.an * Na + c_a.ab * Ba +
                 c_a.at * Ta + c_a.ap * Pa);
      rb = rb - (c_b.aw * Wb + c_b.ae * Eb + c_b.as * Sb + c_b.an * Nb + c_b.ab * Bb +
                 c_b.at * Tb + c_b.ap * Pb);
    }
    if (ok) {
      (ca ? R1 : R0)[q] = (T)ra;
      ss += ra * ra;
      am = fmax(am, fabs(ra));
    }
    if (b_ok) {
      (cb ? R1 : R0)[q] = (T)rb;
      ss += rb * rb;
      am = fmax(am, fabs(rb));
    }
    w0m = w0c;
    w1m = w1c;
    w0c = w0p;
    w1c = w1p;
  }
}

template <class T, bool INNER>
__device__ __forceinline__ void residual_tile(const Tile &t, const FineDev *__restrict__ fine,
                                              const LevelDev &L, int use_p, double &ss,
                                              double &am) {
  const int lpr = fine[t.lev].lpr;
  residual_half<T, INNER>(t, t.x0, fine, L, use_p, ss, am);
  residual_half<T, INNER>(t, t.x0 + lpr, fine, L, use_p, ss, am);
}

// Finest-level residual tile of KB planes, software-pipelined: a rolling
// window of planes k-1, k, k+1 of both colours, with the loads of plane k+1
// (its f, rows j-1/j+1, x-edge values and window plane k+2) issued before
// plane k is computed), two colour entries per lane.  Every neighbour is
// addressed by the coordinates of its image (a ghost row, plane or x position resolved once per lane or per
// plane), as a 32-bit byte offset from the block's first fine array, so
// boundary tiles (BND) run the same code as interior ones; only their
// coefficients differ (cell_coef's zeros towards walls).  The arithmetic of
// every cell is residual_tile's.
template <class T, int KB, bool BND>
__device__ __forceinline__ void residual_tile_fast(const Tile &t, const FineDev &F,
                                                   const LevelDev &L, double &ss, double &am,
                                                   int zbuf) {
  typedef typename Vec2<double>::type d2;
  typedef typename Vec2<T>::type t2;
  constexpr unsigned SZ = sizeof(double);
  // a lane holds entries l0, l0+1 of both colour arrays of its row: cells
  // x0 (a0), x0+1 (b0), x0+2 (a1), x0+3 (b1); a0/a1 share a colour array, as
  // do b0/b1, so each array is one 16-byte load
  const int lpr = F.lpr;
  const int lane = (int)threadIdx.x;
  const int rl = lane & (lpr - 1);
  const int l0 = t.x0 + 2 * rl;
  const int j = t.y0 + (int)threadIdx.y * (64 / lpr) + lane / lpr;
  const int ni = F.ni, nj = F.nj, nk = F.nk, H = F.H;
  const int x0 = 2 * l0;
  const bool row_ok = j < nj;
  const bool ok_a0 = row_ok && x0 < ni, ok_b0 = row_ok && x0 + 1 < ni;
  const bool ok_a1 = row_ok && x0 + 2 < ni, ok_b1 = row_ok && x0 + 3 < ni;
  const int jc = min(j, nj - 1);
  const bool px = F.per & 1, py = F.per & 2, pz = F.per & 4;
  const unsigned row = (unsigned)H * SZ, plane = row * (unsigned)nj;  // bytes
  const char *base = (const char *)F.p[0];
  const unsigned aP1 = (unsigned)((const char *)F.p[1] - base);
  const unsigned aF0 = (unsigned)((const char *)F.f[0] - base);
  const unsigned aF1 = (unsigned)((const char *)F.f[1] - base);
  const gmem<char> *Bs = (const gmem<char> *)base;
  gmem<T> *R[2] = {(gmem<T> *)L.rhs[0], (gmem<T> *)L.rhs[1]};
  auto ld = [&](unsigned off) -> double { return *(const gmem<double> *)(Bs + off); };
  auto ld2 = [&](unsigned off) -> d2 { return *(const gmem<d2> *)(Bs + off); };
  // p of the cells of parity `par` (x + y + k) in a row: colour array par & 1
  auto pcol = [&](int par) -> unsigned { return (par & 1) ? aP1 : 0u; };
  // lanes past the row's last pair read that pair (nothing of theirs is stored)
  const unsigned ent = (unsigned)min(l0, H - 2) * SZ;
  // rows of the y neighbours, ghosts resolved to their image row
  const int ys = BND ? image(jc - 1, nj, py) : jc - 1;
  const int yn = BND ? image(jc + 1, nj, py) : jc + 1;
  const unsigned o_c = (unsigned)jc * row + ent;
  const unsigned o_s = (unsigned)ys * row + ent, o_n = (unsigned)yn * row + ent;
  // x neighbours no lane of the row segment holds: cell x0-1 for the
  // segment's first lane (the west ghost's image, or the previous tile's last
  // entry); for the lane holding cell ni-1 (or lying past it) the east
  // ghost's image, for the segment's last lane otherwise the next tile's
  // first entry
  const bool w_edge = rl == 0;
  const bool e_last = x0 + 3 >= ni - 1;
  const bool e_edge = e_last || rl == lpr - 1;
  const int xw = image(x0 - 1, ni, px);
  const int xe = e_last ? image(ni, ni, px) : x0 + 4;
  const unsigned w_ent = (unsigned)jc * row + (unsigned)(xw >> 1) * SZ;
  const unsigned e_ent = (unsigned)jc * row + (unsigned)(xe >> 1) * SZ;
  const int w_par = xw & 1, e_par = xe & 1;
  const int x1_par = w_edge ? w_par : e_par;
  const unsigned x1_ent = w_edge ? w_ent : e_ent;
  const double cx = F.cx, cy = F.cy, cz = F.cz;
  const double aw_a0 = (x0 == 0 && !px) ? 0.0 : cx;
  const double ae_a0 = (x0 == ni - 1 && !px) ? 0.0 : cx;
  const double ae_b0 = (x0 + 1 == ni - 1 && !px) ? 0.0 : cx;
  const double ae_a1 = (x0 + 2 == ni - 1 && !px) ? 0.0 : cx;
  const double ae_b1 = (x0 + 3 == ni - 1 && !px) ? 0.0 : cx;
  const double as = (BND && !py && jc == 0) ? 0.0 : cy;
  const double an = (BND && !py && jc == nj - 1) ? 0.0 : cy;

  struct PlaneSet {
    d2 fa, fb, sa, sb, na, nb, ap1, bp1;  // ap1/bp1: window plane k+1
    double we, ee;
  };
  // the window plane after k: plane k+1 or its image
  auto next_plane = [&](int k) -> int { return BND ? image(k + 1, nk, pz) : k + 1; };
  auto load_set = [&](int k) -> PlaneSet {
    PlaneSet q;
    const unsigned o = (unsigned)k * plane;
    const int ca = (jc + k) & 1;
    q.fa = ld2((ca ? aF1 : aF0) + o + o_c);
    q.fb = ld2((ca ? aF0 : aF1) + o + o_c);
    q.sa = ld2(pcol(ys + k) + o + o_s);
    q.sb = ld2(pcol(ys + k + 1) + o + o_s);
    q.na = ld2(pcol(yn + k) + o + o_n);
    q.nb = ld2(pcol(yn + k + 1) + o + o_n);
    const int kp = next_plane(k);
    q.ap1 = ld2(pcol(jc + kp) + (unsigned)kp * plane + o_c);
    q.bp1 = ld2(pcol(jc + kp + 1) + (unsigned)kp * plane + o_c);
    // x-edge values: only the lanes at a row segment's ends load them
    double x1 = 0.0, x2 = 0.0;
    if (w_edge || e_edge) x1 = ld(pcol(x1_par + jc + k) + o + x1_ent);
    if (w_edge && e_edge) x2 = ld(pcol(e_par + jc + k) + o + e_ent);
    q.we = x1;
    q.ee = w_edge ? x2 : x1;
    return q;
  };
  const int k0 = t.z0;
  // window: the a and b pairs of planes k0-1 and k0
  d2 am1, bm1, A, B;
  {
    const int km = BND ? image(k0 - 1, nk, pz) : k0 - 1;
    am1 = ld2(pcol(jc + km) + (unsigned)km * plane + o_c);
    bm1 = ld2(pcol(jc + km + 1) + (unsigned)km * plane + o_c);
    A = ld2(pcol(jc + k0) + (unsigned)k0 * plane + o_c);
    B = ld2(pcol(jc + k0 + 1) + (unsigned)k0 * plane + o_c);
  }
#pragma unroll
  for (int p = 0; p < KB; ++p) {
    const int k = k0 + p;
    const PlaneSet cur = load_set(k);
    // cell x0-1 is lane-1's b1, cell x0+4 is lane+1's a0
    const double lf = wave_shift(B.y, 1), rt = wave_shift(A.x, 0);
    const double W_a0 = w_edge ? cur.we : lf;
    const double E_a0 = x0 + 1 < ni ? B.x : cur.ee;
    const double E_b0 = x0 + 2 < ni ? A.y : cur.ee;
    const double E_a1 = x0 + 3 < ni ? B.y : cur.ee;
    const double E_b1 = (x0 + 4 < ni && !e_edge) ? rt : cur.ee;
    const double ab = (BND && !pz && k == 0) ? 0.0 : cz;
    const double at = (BND && !pz && k == nk - 1) ? 0.0 : cz;
    const double ap_a0 = -1.0 * (aw_a0 + ae_a0 + as + an + ab + at);
    const double ap_b0 = -1.0 * (cx + ae_b0 + as + an + ab + at);
    const double ap_a1 = -1.0 * (cx + ae_a1 + as + an + ab + at);
    const double ap_b1 = -1.0 * (cx + ae_b1 + as + an + ab + at);
    const double r_a0 = cur.fa.x - (aw_a0 * W_a0 + ae_a0 * E_a0 + as * cur.sa.x + an * cur.na.x +
                                    ab * am1.x + at * cur.ap1.x + ap_a0 * A.x);
    const double r_b0 = cur.fb.x - (cx * A.x + ae_b0 * E_b0 + as * cur.sb.x + an * cur.nb.x +
                                    ab * bm1.x + at * cur.bp1.x + ap_b0 * B.x);
    const double r_a1 = cur.fa.y - (cx * B.x + ae_a1 * E_a1 + as * cur.sa.y + an * cur.na.y +
                                    ab * am1.y + at * cur.ap1.y + ap_a1 * A.y);
    const double r_b1 = cur.fb.y - (cx * A.y + ae_b1 * E_b1 + as * cur.sb.y + an * cur.nb.y +
                                    ab * bm1.y + at * cur.bp1.y + ap_b1 * B.y);
    const int ca = (jc + k) & 1, cb = ca ^ 1;
    const unsigned q = ((unsigned)k * plane + o_c) / SZ;  // even: a pair's first entry
    // (FP64 level 1 only: the host asks for it there, and the FP32 kernel
    // stays free of it)
    if (sizeof(T) == 8 && zbuf >= 0) {
      // the level-1 relaxation from zero starts with the red cells: rhs/ap
      // (zero_state_red's arithmetic) into buffer zbuf; the red pair is a0,
      // a1 or b0, b1
      const bool red_a = ca == 0;
      const T r0 = (T)(red_a ? r_a0 : r_b0), r1 = (T)(red_a ? r_a1 : r_b1);
      const int xr = x0 + (red_a ? 0 : 1);
      const bool w0 = red_a ? ok_a0 : ok_b0, w1 = red_a ? ok_a1 : ok_b1;
      const Coef<T> c0 = cell_coef<T>(L, xr, jc, k), c1 = cell_coef<T>(L, xr + 2, jc, k);
      t2 v;
      v.x = div_ap(r0 - T(0), c0.ap, T(1) / c0.ap);
      v.y = div_ap(r1 - T(0), c1.ap, T(1) / c1.ap);
      gmem<T> *P = (gmem<T> *)L.phi[zbuf][0];
      if (w1)
        *(gmem<t2> *)(P + q) = v;
      else if (w0)
        P[q] = v.x;
    }
    if (ok_a1) {
      t2 v;
      v.x = (T)r_a0;
      v.y = (T)r_a1;
      *(gmem<t2> *)(R[ca] + q) = v;
    } else if (ok_a0) {
      R[ca][q] = (T)r_a0;
    }
    if (ok_b1) {
      t2 v;
      v.x = (T)r_b0;
      v.y = (T)r_b1;
      *(gmem<t2> *)(R[cb] + q) = v;
    } else if (ok_b0) {
      R[cb][q] = (T)r_b0;
    }
    if (ok_a0) {
      ss += r_a0 * r_a0;
      am = fmax(am, fabs(r_a0));
    }
    if (ok_b0) {
      ss += r_b0 * r_b0;
      am = fmax(am, fabs(r_b0));
    }
    if (ok_a1) {
      ss += r_a1 * r_a1;
      am = fmax(am, fabs(r_a1));
    }
    if (ok_b1) {
      ss += r_b1 * r_b1;
      am = fmax(am, fabs(r_b1));
    }
    am1 = A;
    bm1 = B;
    A = cur.ap1;
    B = cur.bp1;
  }
}

// One launch over the finest level of every block.  k_residual: tiles of KB
// planes (kind 0 y/z interior, kind 3 touching a y/z face); k_residual_gen:
// the general path for r = f (use_p = 0) and for blocks one cell wide in x
// (kind 1 tiles), in a kernel of its own so that its registers do not cost
// the main kernel occupancy.
template <class T, int KB>
__global__ void __launch_bounds__(NTHREADS)
    k_residual(const Tile *__restrict__ tiles, const FineDev *__restrict__ fine,
               const LevelDev *__restrict__ levs, const int *__restrict__ active,
               RedPartial *__restrict__ partials, int zbuf) {
  // XCD-contiguous tile ranges: rows j-1, j+1 of a tile are in the same L2
  const int ti = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  const Tile t = tiles[ti];
  const FineDev &F = fine[t.lev];
  const LevelDev &L = levs[t.lev];
  double ss = 0.0, am = 0.0;
  if (active[F.blk]) {
    if (t.kind == 0)
      residual_tile_fast<T, KB, false>(t, F, L, ss, am, zbuf);
    else
      residual_tile_fast<T, KB, true>(t, F, L, ss, am, zbuf);
  }
  block_reduce(ss, am);
  if (threadIdx.x == 0 && threadIdx.y == 0) partials[ti] = RedPartial{ss, am};
}

template <class T>
__global__ void __launch_bounds__(NTHREADS)
    k_residual_gen(const Tile *__restrict__ tiles, const FineDev *__restrict__ fine,
                   const LevelDev *__restrict__ levs, const int *__restrict__ active,
                   RedPartial *__restrict__ partials, int use_p) {
  const int ti = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  const Tile t = tiles[ti];
  const FineDev &F = fine[t.lev];
  const LevelDev &L = levs[t.lev];
  double ss = 0.0, am = 0.0;
  if (active[F.blk]) {
    if (t.kind == 0)
      residual_tile<T, true>(t, fine, L, use_p, ss, am);
    else
      residual_tile<T, false>(t, fine, L, use_p, ss, am);
  }
  block_reduce(ss, am);
  if (threadIdx.x == 0 && threadIdx.y == 0) partials[ti] = RedPartial{ss, am};
}

// p = p + (pp - ppref) with ppref = 0 (newsolv_mg.for:39-49), streamed over
// the colour arrays of one block (pad entries included: they are never read
// as cells); pp is the level-1 correction in buffer `buf`.  Tile::x0 is the
// first entry, 4 entries per thread.
template <class T>
__global__ void __launch_bounds__(NTHREADS)
    k_update(const Tile *__restrict__ tiles, const FineDev *__restrict__ fine,
             const LevelDev *__restrict__ levs, const int *__restrict__ active, int buf) {
  typedef typename Vec2<T>::type v2t;
  typedef typename Vec2<double>::type v2d;
  const Tile t = tiles[blockIdx.x];
  const FineDev &F = fine[t.lev];
  const LevelDev &L = levs[t.lev];
  if (!active[F.blk]) return;
  const unsigned n = (unsigned)F.H * (unsigned)F.nj * (unsigned)F.nk;
  const int tid = (int)(threadIdx.x + blockDim.x * threadIdx.y);
  for (int c = 0; c < 2; ++c) {
    gmem<double> *P = (gmem<double> *)F.p[c];
    const gmem<T> *Q = (const gmem<T> *)L.phi[buf][c];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const unsigned e = (unsigned)t.x0 + 2u * (unsigned)(tid + h * NTHREADS);
      if (e >= n) continue;
      const v2t pp = *(const gmem<v2t> *)(Q + e);
      v2d p = *(gmem<v2d> *)(P + e);
      p.x = p.x + ((double)pp.x - 0.0);
      p.y = p.y + ((double)pp.y - 0.0);
      *(gmem<v2d> *)(P + e) = p;
    }
  }
}

// per-block reduction of the partials of k_residual, fixed order
constexpr int FIN_THREADS = 1024;
__device__ __forceinline__ void finalize_body(const int *__restrict__ tile_begin,
                                              const RedPartial *__restrict__ partials,
                                              RedPartial *__restrict__ out) {
  const int b = blockIdx.x;
  const int t0 = tile_begin[b], t1 = tile_begin[b + 1];
  double ss = 0.0, am = 0.0;
  for (int q = t0 + (int)threadIdx.x; q < t1; q += FIN_THREADS) {
    ss += partials[q].ss;
    am = fmax(am, partials[q].amax);
  }
  __shared__ double s_ss[FIN_THREADS], s_am[FIN_THREADS];
  s_ss[threadIdx.x] = ss;
  s_am[threadIdx.x] = am;
  __syncthreads();
  for (int w = FIN_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s_ss[threadIdx.x] += s_ss[threadIdx.x + w];
      s_am[threadIdx.x] = fmax(s_am[threadIdx.x], s_am[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[b] = RedPartial{s_ss[0], s_am[0]};
}

__global__ void __launch_bounds__(FIN_THREADS)
    k_finalize(const int *__restrict__ tile_begin, const RedPartial *__restrict__ partials,
               RedPartial *__restrict__ out) {
  finalize_body(tile_begin, partials, out);
}

// The solve loop's stopping rule after cycle ctl->it + 1, as h3dmg_solve's
// host loop applies it (newsolv_mg.for:57-59 for the reference criterion),
// in the same order and IEEE operations: per block ||r|| = sqrt(ss), rel =
// ||r|| / ||f|| (one global rel for blocks joined by faces), rmax = amax *
// rfac * vol; done when (REF) it >= max(min_cycles, 2) and the global rmax <
// tol, else it >= min_cycles and rel <= the block's tolerance.  One thread:
// a handful of blocks, serial in block order like the host's sums.
__device__ void decide_body(const DecideBlock *__restrict__ blocks, DecideState *__restrict__ state,
                            DecideCtl *__restrict__ ctl, const RedPartial *__restrict__ stats,
                            int *__restrict__ active, char *__restrict__ ring,
                            const DecideParams *__restrict__ prm_in) {
  const DecideParams prm = *prm_in;
  const int nb = prm.nb;
  const int it = ctl->it + 1;
  double grel = 0.0;
  if (prm.connected) {
    double ss = 0.0, ff = 0.0;
    for (int b = 0; b < nb; ++b)
      if (blocks[b].enabled) {
        ss += stats[b].ss;
        ff += blocks[b].fnorm * blocks[b].fnorm;
      }
    grel = ff > 0 ? sqrt(ss) / sqrt(ff) : sqrt(ss);
  }
  double worst = 0.0, gmax = 0.0;
  int bad = 0;
  for (int b = 0; b < nb; ++b) {
    const DecideBlock &B = blocks[b];
    if (!B.enabled) continue;
    const double nrm = sqrt(stats[b].ss);
    if (active[b]) {
      state[b].rel = prm.connected ? grel : B.fnorm > 0 ? nrm / B.fnorm : nrm;
      state[b].rmax = stats[b].amax * prm.rfac * B.vol;
    }
    if (!isfinite(nrm)) bad = 1;
    worst = worst < state[b].rel ? state[b].rel : worst;
    gmax = gmax < state[b].rmax ? state[b].rmax : gmax;
  }
  int all_done = 1;
  for (int b = 0; b < nb; ++b) {
    const DecideBlock &B = blocks[b];
    bool done;
    if (!B.enabled)
      done = true;
    else if (prm.ref)
      done = it >= max(prm.min_cycles, 2) && gmax < prm.tol;
    else
      done = it >= prm.min_cycles && state[b].rel <= B.tol;
    // a non-finite norm ends the loop: a cycle queued behind this one does
    // nothing
    if (done || bad) active[b] = 0;
    if (active[b]) all_done = 0;
  }
  const DecideCtl c{it, all_done, bad, 0, prm.ref ? gmax : worst};
  *ctl = c;
  char *slot = ring + (size_t)(it & 1) * decide_slot_bytes(nb);
  *(DecideCtl *)slot = c;
  DecideState *rs = (DecideState *)(slot + sizeof(DecideCtl));
  RedPartial *rp = (RedPartial *)(rs + nb);
  int *ra = (int *)(rp + nb);
  for (int b = 0; b < nb; ++b) {
    rs[b] = state[b];
    rp[b] = stats[b];
    ra[b] = active[b];
  }
}

// k_finalize followed by the stopping rule in one launch: each block's
// workgroup writes its statistics, and the last workgroup to finish (one
// counter, reset by that workgroup for the next launch) applies the rule to
// all of them.
__global__ void __launch_bounds__(FIN_THREADS)
    k_finalize_decide(const int *__restrict__ tile_begin, const RedPartial *__restrict__ partials,
                      RedPartial *__restrict__ out, const DecideBlock *__restrict__ blocks,
                      DecideState *__restrict__ state, DecideCtl *__restrict__ ctl,
                      int *__restrict__ active, char *__restrict__ ring,
                      const DecideParams *__restrict__ prm, unsigned *__restrict__ arrived) {
  finalize_body(tile_begin, partials, out);
  if (threadIdx.x != 0) return;
  __threadfence();  // this block's statistics before the count
  const unsigned n = gridDim.x;
  if (atomicAdd(arrived, 1u) != n - 1) return;
  __threadfence();  // every block's statistics after it
  *arrived = 0u;
  decide_body(blocks, state, ctl, out, active, ring, prm);
}
// ---------------------------------------------------------------------------
// Exact FP64 solve of the coarsest level of one block (extension, no
// reference counterpart): A = Dx (+) Dy (+) Dz with Dx the 1-D Neumann or
// periodic second difference; with Q_a the orthonormal eigenvectors of D_a
// (columns) and l_a its eigenvalues, u = Q (Q^T f ./ (lx+ly+lz)) with the
// zero (constant) mode dropped -- the null-space projection.  One workgroup
// per block; scratch holds two n^3 arrays.  The body is shared by
// k_coarse_exact (arrays in global memory) and the V-cycle tail kernel
// (arrays in LDS, S = lds address space): the same arithmetic in the same
// order.  Every thread of the workgroup calls it.
template <class T>
using lds = __attribute__((address_space(3))) T;

template <bool IN_LDS>
struct Space;
template <>
struct Space<false> {
  template <class T>
  using ptr = T *;
  template <class T>
  static __device__ __forceinline__ T *cast(void *p) {
    return (T *)p;
  }
};
template <>
struct Space<true> {
  template <class T>
  using ptr = lds<T> *;
  template <class T>
  static __device__ __forceinline__ lds<T> *cast(void *p) {
    return (lds<T> *)p;
  }
};

__device__ __forceinline__ int cidx32(int H, int nj, int i, int j, int k) {
  return (k * nj + j) * H + (i >> 1);
}

template <bool IN_LDS>
__device__ void coarse_exact_body(const LevelDev &L, const double *Q0g, const double *Q1g,
                                  const double *Q2g, const double *l0g, const double *l1g,
                                  const double *l2g, double *Ag, double *Bg, int buf) {
  typedef Space<IN_LDS> S;
  typedef typename S::template ptr<double> dp;
  const dp Q0 = S::template cast<double>((void *)Q0g), Q1 = S::template cast<double>((void *)Q1g),
           Q2 = S::template cast<double>((void *)Q2g);
  const dp l0 = S::template cast<double>((void *)l0g), l1 = S::template cast<double>((void *)l1g),
           l2 = S::template cast<double>((void *)l2g);
  const dp A = S::template cast<double>(Ag), Bs = S::template cast<double>(Bg);
  const dp R0 = S::template cast<double>(L.rhs[0]), R1 = S::template cast<double>(L.rhs[1]);
  const dp P0 = S::template cast<double>(L.phi[buf][0]), P1 = S::template cast<double>(L.phi[buf][1]);
  const int n0 = L.ni, n1 = L.nj, n2 = L.nk;
  const int plane = n0 * n1;
  const int n = plane * n2;
  const int tid = (int)threadIdx.x;
  const int nthr = (int)blockDim.x;
  // gather rhs (colour split) into natural layout A[i + n0*(j + n1*k)]
  for (int q = tid; q < n; q += nthr) {
    const int i = q % n0, j = (q / n0) % n1, k = q / plane;
    const int e = cidx32(L.H, L.nj, i, j, k);
    A[q] = ((i + j + k) & 1) ? R1[e] : R0[e];
  }
  __syncthreads();
  // forward transforms: B = Q0^T along i, A = Q1^T along j, B = Q2^T along k
  for (int q = tid; q < n; q += nthr) {
    const int i = q % n0;
    const int base = q - i;
    double s = 0.0;
    for (int x = 0; x < n0; ++x) s += Q0[x * n0 + i] * A[base + x];
    Bs[q] = s;
  }
  __syncthreads();
  for (int q = tid; q < n; q += nthr) {
    const int j = (q / n0) % n1;
    const int base = q - j * n0;
    double s = 0.0;
    for (int y = 0; y < n1; ++y) s += Q1[y * n1 + j] * Bs[base + y * n0];
    A[q] = s;
  }
  __syncthreads();
  for (int q = tid; q < n; q += nthr) {
    const int k = q / plane;
    const int base = q - k * plane;
    double s = 0.0;
    for (int z = 0; z < n2; ++z) s += Q2[z * n2 + k] * A[base + z * plane];
    const int i = q % n0, j = (q / n0) % n1;
    const double lam = l0[i] + l1[j] + l2[k];
    // mode (0,0,0) is the constant (null-space) mode of every axis
    Bs[q] = ((i | j | k) == 0) ? 0.0 : s / lam;
  }
  __syncthreads();
  // inverse transforms: A = Q2 along k, B = Q1 along j, phi = Q0 along i
  for (int q = tid; q < n; q += nthr) {
    const int k = q / plane;
    const int base = q - k * plane;
    double s = 0.0;
    for (int z = 0; z < n2; ++z) s += Q2[k * n2 + z] * Bs[base + z * plane];
    A[q] = s;
  }
  __syncthreads();
  for (int q = tid; q < n; q += nthr) {
    const int j = (q / n0) % n1;
    const int base = q - j * n0;
    double s = 0.0;
    for (int y = 0; y < n1; ++y) s += Q1[j * n1 + y] * A[base + y * n0];
    Bs[q] = s;
  }
  __syncthreads();
  for (int q = tid; q < n; q += nthr) {
    const int i = q % n0, j = (q / n0) % n1, k = q / plane;
    const int base = q - i;
    double s = 0.0;
    for (int x = 0; x < n0; ++x) s += Q0[i * n0 + x] * Bs[base + x];
    const int e = cidx32(L.H, L.nj, i, j, k);
    if ((i + j + k) & 1)
      P1[e] = s;
    else
      P0[e] = s;
  }
  __syncthreads();
}

__global__ void __launch_bounds__(256)
    k_coarse_exact(const Tile *__restrict__ tiles, const LevelDev *__restrict__ levs,
                   const int *__restrict__ active, const double *const *__restrict__ eig_q,
                   const double *const *__restrict__ eig_l, double *const *__restrict__ scratch,
                   int buf) {
  const Tile t = tiles[blockIdx.x];  // lev = coarsest descriptor, x0 = block
  if (!active[t.x0]) return;
  const int b = t.x0;
  coarse_exact_body<false>(levs[t.lev], eig_q[3 * b], eig_q[3 * b + 1], eig_q[3 * b + 2],
                           eig_l[3 * b], eig_l[3 * b + 1], eig_l[3 * b + 2], scratch[2 * b],
                           scratch[2 * b + 1], buf);
}

// ---------------------------------------------------------------------------
// V-cycle tail: one workgroup per block runs every level from the tail level
// Lt down to the block's coarsest and back up (mgkcyc, mgsolver.for:272-352,
// restricted to levels >= Lt) entirely in LDS: pre-smoothing, restriction,
// the coarsest treatment, prolongation and post-smoothing, separated by
// workgroup barriers.  Small levels are latency-bound: one launch, LDS
// latency and on-chip barriers replace ~3 full-chip launches per level.
//
// The tail levels are copied into LDS with their global layout (colour
// split, pitch H); descriptors whose pointers point into LDS feed per-cell
// functions that compute every cell with the same expression, in the same
// order, as the per-level kernels (bitwise-identical results).  In: the top
// level's rhs (global).  Out: the top level's phi in buffer (nu1+nu2)&1
// (global), the buffer every level ends in on the host side.  The tile's
// `lev` is the block's level-Lt descriptor.  Levels are float in a
// mixed-precision handle except the coarsest (double).

// value of cell (x,y,z) in [-1, n] of an LDS-resident level, buffer `buf`,
// ghosts resolved to their image
template <class T>
__device__ __forceinline__ T lds_value(const LevelDev &L, int buf, int x, int y, int z) {
  x = image(x, L.ni, L.per & 1);
  y = image(y, L.nj, L.per & 2);
  z = image(z, L.nk, L.per & 4);
  const lds<T> *a = (const lds<T> *)L.phi[buf][(x + y + z) & 1];
  return a[cidx32(L.H, L.nj, x, y, z)];
}

// One colour half-sweep over a whole level, cell-parallel; the loads of
// TAIL_U cells per thread are issued before any of their stores.  A red cell
// reads its black neighbours before the sweep (buffer src), a black cell its
// red neighbours after it (buffer 1-src); ghosts read their image in buffer
// src; with zero != 0 buffer src holds zeros.  Coefficients and arithmetic
// are cell_coef's and the per-level kernels'.  Only the coordinate a
// neighbour differs in can leave the level, so each neighbour resolves one
// image and tests one ghost condition, and entries are split into (entry,
// row, plane) by float reciprocals with one exact correction (n < 2^24).
template <class T>
__device__ __forceinline__ void tail_half_sweep(const LevelDev &Ls, int colour, int src, int zero) {
  // the descriptor lives in LDS: its fields once, in registers
  const int H = Ls.H, nj = Ls.nj, nk = Ls.nk, ni = Ls.ni, per = Ls.per;
  const bool px = per & 1, py = per & 2, pz = per & 4;
  const bool wx0 = wall_face(Ls, 0), wx1 = wall_face(Ls, 1), wy0 = wall_face(Ls, 2),
             wy1 = wall_face(Ls, 3), wz0 = wall_face(Ls, 4), wz1 = wall_face(Ls, 5);
  const T cx = (T)Ls.cx, cy = (T)Ls.cy, cz = (T)Ls.cz;
  const int n = H * nj * nk, plane = H * nj;
  const lds<T> *nb_int =
      colour == 0 ? (const lds<T> *)Ls.phi[src][1] : (const lds<T> *)Ls.phi[1 - src][0];
  const lds<T> *g0 = (const lds<T> *)Ls.phi[src][0], *g1 = (const lds<T> *)Ls.phi[src][1];
  const lds<T> *R = (const lds<T> *)Ls.rhs[colour];
  lds<T> *D = (lds<T> *)Ls.phi[1 - src][colour];
  const bool zero_int = zero && colour == 0, zero_ghost = zero != 0;
  const float invH = 1.0f / (float)H, invJ = 1.0f / (float)nj;
  // a neighbour: interior from nb_int, a ghost from its image in buffer src
  // (colour array by the image's parity); zeros in the zero state
  auto nbv = [&](bool ghost, int par, int entry) -> T {
    const lds<T> *a = ghost ? ((par & 1) ? g1 : g0) : nb_int;
    const T v = a[entry];
    return (ghost ? zero_ghost : zero_int) ? T(0) : v;
  };
  for (int base = (int)threadIdx.x; base < n; base += TAIL_THREADS * TAIL_U) {
    T v[TAIL_U][7];
    int ci[TAIL_U], cj[TAIL_U], ck[TAIL_U];
    bool ok[TAIL_U];
#pragma unroll
    for (int u = 0; u < TAIL_U; ++u) {
      const int e0 = base + u * TAIL_THREADS;
      const int e = min(e0, n - 1);  // lanes past the end read a valid entry
      int r = (int)((float)e * invH);
      int m = e - r * H;
      if (m < 0) {
        --r;
        m += H;
      } else if (m >= H) {
        ++r;
        m -= H;
      }
      int k = (int)((float)r * invJ);
      int j = r - k * nj;
      if (j < 0) {
        --k;
        j += nj;
      } else if (j >= nj) {
        ++k;
        j -= nj;
      }
      const int i0 = 2 * m + ((colour + j + k) & 1);
      const int i = min(i0, ni - 1);
      ok[u] = e0 < n && i0 < ni;
      ci[u] = i;
      cj[u] = j;
      ck[u] = k;
      const int rb = (k * nj + j) * H;  // the row's first entry
      const int xw = image(i - 1, ni, px), xe = image(i + 1, ni, px);
      const int ys = image(j - 1, nj, py), yn = image(j + 1, nj, py);
      const int zb = image(k - 1, nk, pz), zt = image(k + 1, nk, pz);
      v[u][0] = nbv(i == 0, xw + j + k, rb + (xw >> 1));
      v[u][1] = nbv(i == ni - 1, xe + j + k, rb + (xe >> 1));
      v[u][2] = nbv(j == 0, i + ys + k, rb + (ys - j) * H + (i >> 1));
      v[u][3] = nbv(j == nj - 1, i + yn + k, rb + (yn - j) * H + (i >> 1));
      v[u][4] = nbv(k == 0, i + j + zb, rb + (zb - k) * plane + (i >> 1));
      v[u][5] = nbv(k == nk - 1, i + j + zt, rb + (zt - k) * plane + (i >> 1));
      v[u][6] = R[e];
    }
#pragma unroll
    for (int u = 0; u < TAIL_U; ++u) {
      if (!ok[u]) continue;
      const int i = ci[u], j = cj[u], k = ck[u];
      // cell_coef's coefficients: zero towards a wall
      const T aw = (i == 0 && wx0) ? T(0) : cx, ae = (i == ni - 1 && wx1) ? T(0) : cx;
      const T as = (j == 0 && wy0) ? T(0) : cy, an = (j == nj - 1 && wy1) ? T(0) : cy;
      const T ab = (k == 0 && wz0) ? T(0) : cz, at = (k == nk - 1 && wz1) ? T(0) : cz;
      const T ap = T(-1) * (aw + ae + as + an + ab + at);
      const T sum = aw * v[u][0] + ae * v[u][1] + as * v[u][2] + an * v[u][3] + ab * v[u][4] +
                    at * v[u][5];
      D[base + u * TAIL_THREADS] = div_ap(v[u][6] - sum, ap, T(1) / ap);
    }
  }
  __syncthreads();
}

template <class T>
__device__ __forceinline__ void tail_fill_zero(const LevelDev &L, int buf) {
  const int n = L.H * L.nj * L.nk;
  lds<T> *a = (lds<T> *)L.phi[buf][0], *b = (lds<T> *)L.phi[buf][1];
  for (int e = (int)threadIdx.x; e < n; e += TAIL_THREADS) {
    a[e] = T(0);
    b[e] = T(0);
  }
  __syncthreads();
}

// nsweep sweeps from buffer `buf` (flag `zero`: buffer holds zeros)
template <class T>
__device__ __forceinline__ void tail_relax(const LevelDev &L, int nsweep, int &buf, bool &zero) {
  for (int sw = 0; sw < nsweep; ++sw) {
    tail_half_sweep<T>(L, 0, buf, zero ? 1 : 0);
    tail_half_sweep<T>(L, 1, buf, zero ? 1 : 0);
    buf ^= 1;
    zero = false;
  }
}

// mgrestr for one coarse cell per thread (the arithmetic of restrict_tile)
template <class Tf, class Tc>
__device__ __forceinline__ void tail_restrict(const LevelDev &F, const LevelDev &C, int src_f) {
  const int n = C.ni * C.nj * C.nk;
  const lds<Tf> *P0 = (const lds<Tf> *)F.phi[src_f][0], *P1 = (const lds<Tf> *)F.phi[src_f][1];
  const lds<Tf> *F0 = (const lds<Tf> *)F.rhs[0], *F1 = (const lds<Tf> *)F.rhs[1];
  lds<Tc> *C0 = (lds<Tc> *)C.rhs[0], *C1 = (lds<Tc> *)C.rhs[1];
  for (int q0 = (int)threadIdx.x; q0 < n; q0 += TAIL_THREADS) {
    const int ic = q0 % C.ni, jc = (q0 / C.ni) % C.nj, kc = q0 / (C.ni * C.nj);
    Tc acc = Tc(0);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int i = 2 * ic + (q & 1), j = 2 * jc + ((q >> 1) & 1), k = 2 * kc + (q >> 2);
      const Coef<Tf> c = cell_coef<Tf>(F, i, j, k);
      const Tf W = lds_value<Tf>(F, src_f, i - 1, j, k);
      const Tf E = lds_value<Tf>(F, src_f, i + 1, j, k);
      const Tf S = lds_value<Tf>(F, src_f, i, j - 1, k);
      const Tf N = lds_value<Tf>(F, src_f, i, j + 1, k);
      const Tf B = lds_value<Tf>(F, src_f, i, j, k - 1);
      const Tf Tt = lds_value<Tf>(F, src_f, i, j, k + 1);
      const int fq = cidx32(F.H, F.nj, i, j, k);
      const bool fc = (i + j + k) & 1;
      const Tf P = fc ? P1[fq] : P0[fq];
      const Tf r = fc ? F1[fq] : F0[fq];
      const Tf res =
          r - (c.aw * W + c.ae * E + c.as * S + c.an * N + c.ab * B + c.at * Tt + c.ap * P);
      acc = q == 0 ? (Tc)res : acc + (Tc)res;
    }
    const int cq = cidx32(C.H, C.nj, ic, jc, kc);
    if ((ic + jc + kc) & 1)
      C1[cq] = Tc(0.125) * acc;
    else
      C0[cq] = Tc(0.125) * acc;
  }
  __syncthreads();
}

// mgcorr for one fine cell per thread (the arithmetic of k_prolong): fine
// cell (i,j,k) is cell e = i&1 of entry l = i>>1, coarse row/plane pair
// m = (j+1)>>1, n = (k+1)>>1 with offsets dj = (j+1)&1, dk = (k+1)&1
template <class Tf, class Tc>
__device__ __forceinline__ void tail_prolong(const LevelDev &F, const LevelDev &C, int src_f,
                                             int src_c) {
  const int n = F.ni * F.nj * F.nk;
  const bool px = C.per & 1, py = C.per & 2, pz = C.per & 4;
  const lds<Tc> *V0 = (const lds<Tc> *)C.phi[src_c][0], *V1 = (const lds<Tc> *)C.phi[src_c][1];
  lds<Tf> *D0 = (lds<Tf> *)F.phi[src_f][0], *D1 = (lds<Tf> *)F.phi[src_f][1];
  for (int q0 = (int)threadIdx.x; q0 < n; q0 += TAIL_THREADS) {
    const int i = q0 % F.ni, j = (q0 / F.ni) % F.nj, k = q0 / (F.ni * F.nj);
    const int l = i >> 1, e = i & 1, di = 1 - e;
    const int m = (j + 1) >> 1, dj = (j + 1) & 1;
    const int nn = (k + 1) >> 1, dk = (k + 1) & 1;
    Tc s = Tc(0);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int oi = q >> 2, oj = q & 1, ok = (q >> 1) & 1;
      const int x = image(l - 1 + e + oi, C.ni, px);
      const int y = image(m - 1 + oj, C.nj, py);
      const int z = image(nn - 1 + ok, C.nk, pz);
      const Tc w = (oi == di ? Tc(3) : Tc(1)) * (oj == dj ? Tc(3) : Tc(1)) *
                   (ok == dk ? Tc(3) : Tc(1));
      const int cq = cidx32(C.H, C.nj, x, y, z);
      const Tc v = ((x + y + z) & 1) ? V1[cq] : V0[cq];
      s = (q == 0) ? w * v : s + w * v;
    }
    lds<Tf> *fa = ((i + j + k) & 1) ? D1 : D0;
    const int fq = cidx32(F.H, F.nj, i, j, k);
    fa[fq] = fa[fq] + (Tf)(s / Tc(64));
  }
  __syncthreads();
}

template <class D, class S>
__device__ __forceinline__ void tail_copy(D *dst, const S *src, int n) {
  for (int e = (int)threadIdx.x; e < n; e += TAIL_THREADS) dst[e] = src[e];
}

template <bool MIXED>
__global__ void __launch_bounds__(TAIL_THREADS)
    k_vtail(const Tile *__restrict__ tiles, const LevelDev *__restrict__ levs,
            const int *__restrict__ active, const double *const *__restrict__ eig_q,
            const double *const *__restrict__ eig_l, int nu1, int nu2, int exact) {
  extern __shared__ __align__(16) char smem[];
  // descriptors of the tail's levels with their arrays in LDS: level q of the
  // tail is lv[q] (written by thread 0, read by all)
  __shared__ LevelDev lv[TAIL_MAX_LEVELS];
  __shared__ int s_nl;
  __shared__ size_t s_off;
  const Tile t = tiles[blockIdx.x];
  const int top = t.lev;
  const int blk = levs[top].blk;
  if (!active[blk]) return;
  const int fin = (nu1 + nu2) & 1;
  if (threadIdx.x == 0) {
    int nl = 0;
    size_t off = 0;
    for (int d = top; d >= 0 && nl < TAIL_MAX_LEVELS; d = levs[d].coarser) {
      LevelDev L = levs[d];
      const bool w8 = !MIXED || L.coarser < 0;
      const size_t half = (size_t)L.H * L.nj * L.nk * (w8 ? 8 : 4);
      for (int bb = 0; bb < 2; ++bb)
        for (int c = 0; c < 2; ++c) {
          L.phi[bb][c] = smem + off;
          off += (half + 15) & ~(size_t)15;
        }
      for (int c = 0; c < 2; ++c) {
        L.rhs[c] = smem + off;
        off += (half + 15) & ~(size_t)15;
      }
      lv[nl++] = L;
    }
    s_nl = nl;
    s_off = off;
  }
  __syncthreads();
  const int last = s_nl - 1;
  size_t off = s_off;
  auto is64 = [&](int q) { return !MIXED || lv[q].coarser < 0; };
  const bool do_exact = exact && lv[last].finer >= 0;  // as the host: not for a 1-level block
  // the exact solve's tables and scratch follow the levels
  const double *Q[3], *lam[3];
  double *A = nullptr, *Bs = nullptr;
  if (do_exact) {
    const LevelDev &L = lv[last];
    const int na[3] = {L.ni, L.nj, L.nk};
    for (int a = 0; a < 3; ++a) {
      double *q = (double *)(smem + off);
      off += (size_t)na[a] * na[a] * 8;
      double *l = (double *)(smem + off);
      off += ((size_t)na[a] * 8 + 15) & ~(size_t)15;
      tail_copy((lds<double> *)q, eig_q[3 * blk + a], na[a] * na[a]);
      tail_copy((lds<double> *)l, eig_l[3 * blk + a], na[a]);
      Q[a] = q;
      lam[a] = l;
    }
    const size_t n3 = (size_t)na[0] * na[1] * na[2];
    A = (double *)(smem + off);
    off += n3 * 8;
    Bs = (double *)(smem + off);
  }
  // top level's rhs into LDS
  {
    const LevelDev &G = levs[top];
    const int n = G.H * G.nj * G.nk;
    for (int c = 0; c < 2; ++c) {
      if (is64(0))
        tail_copy((lds<double> *)lv[0].rhs[c], (const double *)G.rhs[c], n);
      else
        tail_copy((lds<float> *)lv[0].rhs[c], (const float *)G.rhs[c], n);
    }
    __syncthreads();
  }
  // down: pre-smoothing and restriction
  for (int q = 0; q < last; ++q) {
    const LevelDev &L = lv[q];
    const LevelDev &C = lv[q + 1];
    int buf = 0;
    bool zero = true;
    if (is64(q)) {
      tail_relax<double>(L, nu1, buf, zero);
      if (zero) tail_fill_zero<double>(L, buf);
      tail_restrict<double, double>(L, C, buf);
    } else {
      tail_relax<float>(L, nu1, buf, zero);
      if (zero) tail_fill_zero<float>(L, buf);
      if (is64(q + 1))
        tail_restrict<float, double>(L, C, buf);
      else
        tail_restrict<float, float>(L, C, buf);
    }
  }
  // coarsest level of the block (always FP64)
  if (do_exact) {
    coarse_exact_body<true>(lv[last], Q[0], Q[1], Q[2], lam[0], lam[1], lam[2], A, Bs, fin);
  } else {
    int buf = 0;
    bool zero = true;
    tail_relax<double>(lv[last], nu1, buf, zero);
    tail_relax<double>(lv[last], nu2, buf, zero);
    if (zero) tail_fill_zero<double>(lv[last], buf);
  }
  // up: prolongation and post-smoothing
  for (int q = last - 1; q >= 0; --q) {
    const LevelDev &F = lv[q];
    const LevelDev &C = lv[q + 1];
    const int fb = nu1 & 1;
    int buf = fb;
    bool zero = false;
    if (is64(q)) {
      tail_prolong<double, double>(F, C, fb, fin);
      tail_relax<double>(F, nu2, buf, zero);
    } else {
      if (is64(q + 1))
        tail_prolong<float, double>(F, C, fb, fin);
      else
        tail_prolong<float, float>(F, C, fb, fin);
      tail_relax<float>(F, nu2, buf, zero);
    }
  }
  // top level's phi (buffer fin) back to global
  {
    const LevelDev &G = levs[top];
    const int n = G.H * G.nj * G.nk;
    for (int c = 0; c < 2; ++c) {
      if (is64(0))
        tail_copy((double *)G.phi[fin][c], (const lds<double> *)lv[0].phi[fin][c], n);
      else
        tail_copy((float *)G.phi[fin][c], (const lds<float> *)lv[0].phi[fin][c], n);
    }
  }
}

// ---------------------------------------------------------------------------
// Variable-density operator (coeff under L_LSM, mgsolver.for:72-184).  The
// reference builds, per level, dens_mg (the density of the level's cells:
// the input at level 1, the 8-child average below) and from it per-cell
// coefficients in finite-volume form with the FINEST spacing at every level:
//   aw = dy*dz/dx/dens(i-1/2),  dens(i-1/2) = 0.5*(dens(i) + dens(i-1)), or
//   dens(i) on the block's first cell; likewise ae, as, an, ab, at; zero
//   towards a wall; ap = -(aw+ae+as+an+ab+at).
// Kernels here are per cell (one thread per cell or coarse cell) with the
// coefficients read from FP64 arrays laid out like phi.

__device__ __forceinline__ const double *lsm_coef(const LevelDev &L, int q, int colour) {
  return L.lsm + (size_t)(2 * q + colour) * L.lsm_half;
}

// dens_mg of a coarse level: 0.125 * (c(j0 i0 k0) + c(j1 i0 k0) + c(j0 i1 k0)
// + c(j1 i1 k0) + c(j0 i0 k1) + ... ), the reference's j-fastest packing
// order (mgsolver.for:99-118); arrays interior, i fastest
__global__ void k_lsm_dens_restrict(const double *__restrict__ fine, int fni, int fnj, int fnk,
                                    double *__restrict__ coarse) {
  const int cni = fni / 2, cnj = fnj / 2, cnk = fnk / 2;
  const size_t n = (size_t)cni * cnj * cnk;
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int i = (int)(q % cni), j = (int)((q / cni) % cnj), k = (int)(q / ((size_t)cni * cnj));
  auto f = [&](int di, int dj, int dk) {
    return fine[((size_t)(2 * k + dk) * fnj + (2 * j + dj)) * fni + (2 * i + di)];
  };
  coarse[q] = 0.125 * (f(0, 0, 0) + f(0, 1, 0) + f(1, 0, 0) + f(1, 1, 0) + f(0, 0, 1) +
                       f(0, 1, 1) + f(1, 0, 1) + f(1, 1, 1));
}

// coefficients of every cell of one level of one block; walls: bit a*2 (+1)
// set = the low (high) face of axis a is a wall (coefficient 0)
__global__ void k_lsm_coef(const LevelDev *__restrict__ lev_dev, const double *__restrict__ dens,
                           int ni, int nj, int nk, double dx, double dy, double dz, int walls) {
  const size_t n = (size_t)ni * nj * nk;
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const LevelDev &L = *lev_dev;
  const int i = (int)(q % ni), j = (int)((q / ni) % nj), k = (int)(q / ((size_t)ni * nj));
  const size_t si = 1, sj = (size_t)ni, sk = (size_t)ni * nj;
  const double d = dens[q];
  // face densities (mgsolver.for:130-164): the cell's own on the block's
  // first/last cell of an axis, else the mean with the neighbour (own first)
  const double dim = i == 0 ? d : 0.5 * (d + dens[q - si]);
  const double dip = i == ni - 1 ? d : 0.5 * (d + dens[q + si]);
  const double djm = j == 0 ? d : 0.5 * (d + dens[q - sj]);
  const double djp = j == nj - 1 ? d : 0.5 * (d + dens[q + sj]);
  const double dkm = k == 0 ? d : 0.5 * (d + dens[q - sk]);
  const double dkp = k == nk - 1 ? d : 0.5 * (d + dens[q + sk]);
  double c[7];
  c[0] = dy * dz / dx / dim;
  c[1] = dy * dz / dx / dip;
  c[2] = dx * dz / dy / djm;
  c[3] = dx * dz / dy / djp;
  c[4] = dx * dy / dz / dkm;
  c[5] = dx * dy / dz / dkp;
  // walls (mgsolver.for:194-205)
  if ((walls & 1) && i == 0) c[0] = 0.0;
  if ((walls & 2) && i == ni - 1) c[1] = 0.0;
  if ((walls & 4) && j == 0) c[2] = 0.0;
  if ((walls & 8) && j == nj - 1) c[3] = 0.0;
  if ((walls & 16) && k == 0) c[4] = 0.0;
  if ((walls & 32) && k == nk - 1) c[5] = 0.0;
  c[6] = -1.0 * (c[0] + c[1] + c[2] + c[3] + c[4] + c[5]);
  const int colour = (i + j + k) & 1;
  const size_t e = cidx(L.H, L.nj, i, j, k);
  for (int a = 0; a < 7; ++a) L.lsm[(size_t)(2 * a + colour) * L.lsm_half + e] = c[a];
}

// one colour half-sweep (mgsolver.for:553-605 with per-cell coefficients):
// natural tiles (TILE_X x TILE_Y cells of one plane), threads of the other
// colour idle; ghosts and the stale-ghost buffer rule as in march_tile
__global__ void __launch_bounds__(NTHREADS)
    k_lsm_sweep(const Tile *__restrict__ tiles, const LevelDev *__restrict__ levs,
                const int *__restrict__ active, int colour, int src, int zero) {
  const Tile t = tiles[blockIdx.x];
  const LevelDev &L = levs[t.lev];
  if (!active[L.blk]) return;
  const int i = t.x0 + (int)threadIdx.x, j = t.y0 + (int)threadIdx.y, k = t.z0;
  if (i >= L.ni || j >= L.nj || k >= L.nk || ((i + j + k) & 1) != colour) return;
  const double *O = (const double *)(colour == 0 ? L.phi[src][1] : L.phi[1 - src][0]);
  const bool zero_int = zero && colour == 0, zero_ghost = zero != 0;
  auto nb = [&](int x, int y, int z) -> double {
    const bool ghost = (unsigned)x >= (unsigned)L.ni || (unsigned)y >= (unsigned)L.nj ||
                       (unsigned)z >= (unsigned)L.nk;
    if (!ghost) return zero_int ? 0.0 : O[cidx(L.H, L.nj, x, y, z)];
    return zero_ghost ? 0.0 : level_value<double>(levs, t.lev, src, x, y, z);
  };
  const size_t e = cidx(L.H, L.nj, i, j, k);
  double c[7];
  for (int a = 0; a < 7; ++a) c[a] = lsm_coef(L, a, colour)[e];
  const double W = nb(i - 1, j, k), E = nb(i + 1, j, k);
  const double S = nb(i, j - 1, k), N = nb(i, j + 1, k);
  const double B = nb(i, j, k - 1), T = nb(i, j, k + 1);
  const double r = ((const double *)L.rhs[colour])[e];
  const double sum = c[0] * W + c[1] * E + c[2] * S + c[3] * N + c[4] * B + c[5] * T;
  ((double *)L.phi[1 - src][colour])[e] = (r - sum) / c[6];
}

// finest-level residual r = f - A p with per-cell coefficients into the
// level-1 rhs, with per-tile partial sum of r^2 and max |r| (the residual
// tiles' work layout: lane l holds cells 2l and 2l+1 of one row)
__global__ void __launch_bounds__(NTHREADS)
    k_lsm_residual(const Tile *__restrict__ tiles, const FineDev *__restrict__ fine,
                   const LevelDev *__restrict__ levs, const int *__restrict__ active,
                   RedPartial *__restrict__ partials) {
  const int ti = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  const Tile t = tiles[ti];
  const FineDev &F = fine[t.lev];
  const LevelDev &L = levs[t.lev];
  double ss = 0.0, am = 0.0;
  if (active[F.blk]) {
    const int lpr = F.lpr;
    const int lane = (int)threadIdx.x;
    const int j = t.y0 + (int)threadIdx.y * (64 / lpr) + lane / lpr;
    // the tile's two halves of lpr entries, one entry per lane
    for (int half = 0; half < 2; ++half)
    for (int k = t.z0; k < t.z1; ++k)
      for (int s = 0; s < 2; ++s) {
        const int l = t.x0 + half * lpr + (lane & (lpr - 1));
        const int i = 2 * l + s;
        if (i >= F.ni || j >= F.nj) continue;
        const int c = (i + j + k) & 1;
        const size_t e = cidx(F.H, F.nj, i, j, k);
        auto nb = [&](int x, int y, int z) -> double { return fine_value(fine, t.lev, x, y, z); };
        double cf[7];
        for (int a = 0; a < 7; ++a) cf[a] = lsm_coef(L, a, c)[e];
        const double P = F.p[c][e];
        const double r = F.f[c][e] - (cf[0] * nb(i - 1, j, k) + cf[1] * nb(i + 1, j, k) +
                                      cf[2] * nb(i, j - 1, k) + cf[3] * nb(i, j + 1, k) +
                                      cf[4] * nb(i, j, k - 1) + cf[5] * nb(i, j, k + 1) +
                                      cf[6] * P);
        ((double *)L.rhs[c])[e] = r;
        ss += r * r;
        am = fmax(am, fabs(r));
      }
  }
  block_reduce(ss, am);
  if (threadIdx.x == 0 && threadIdx.y == 0) partials[ti] = RedPartial{ss, am};
}

// restriction (mgrestr, mgsolver.for:693-756) with per-cell coefficients:
// a thread per coarse cell of the restriction tile's layout; the eight fine
// residuals summed i fastest, then j, k, times 0.125
__global__ void __launch_bounds__(NTHREADS)
    k_lsm_restrict(const Tile *__restrict__ tiles, const LevelDev *__restrict__ levs,
                   const int *__restrict__ active, int src_f) {
  const Tile t = tiles[xcd_remap((int)blockIdx.x, (int)gridDim.x)];
  const LevelDev &C = levs[t.lev];
  if (!active[C.blk]) return;
  const LevelDev &F = levs[C.finer];
  const int lpr = C.lpr1;
  const int lane = (int)threadIdx.x;
  const int jc = t.y0 + (int)threadIdx.y * (64 / lpr) + lane / lpr;
  auto res = [&](int i, int j, int k) -> double {
    const int c = (i + j + k) & 1;
    const size_t e = cidx(F.H, F.nj, i, j, k);
    double cf[7];
    for (int a = 0; a < 7; ++a) cf[a] = lsm_coef(F, a, c)[e];
    auto nb = [&](int x, int y, int z) { return level_value<double>(levs, C.finer, src_f, x, y, z); };
    return ((const double *)F.rhs[c])[e] -
           (cf[0] * nb(i - 1, j, k) + cf[1] * nb(i + 1, j, k) + cf[2] * nb(i, j - 1, k) +
            cf[3] * nb(i, j + 1, k) + cf[4] * nb(i, j, k - 1) + cf[5] * nb(i, j, k + 1) +
            cf[6] * nb(i, j, k));
  };
  // the tile's two halves of lpr coarse cells, one cell per lane
  for (int half = 0; half < 2; ++half) {
  const int l = t.x0 + half * lpr + (lane & (lpr - 1));
  if (l >= C.ni || jc >= C.nj) continue;
  for (int kc = t.z0; kc < t.z1; ++kc) {
    const int i0 = 2 * l, j0 = 2 * jc, k0 = 2 * kc;
    double acc = res(i0, j0, k0);
    acc = acc + res(i0 + 1, j0, k0);
    acc = acc + res(i0, j0 + 1, k0);
    acc = acc + res(i0 + 1, j0 + 1, k0);
    acc = acc + res(i0, j0, k0 + 1);
    acc = acc + res(i0 + 1, j0, k0 + 1);
    acc = acc + res(i0, j0 + 1, k0 + 1);
    acc = acc + res(i0 + 1, j0 + 1, k0 + 1);
    ((double *)C.rhs[(l + jc + kc) & 1])[cidx(C.H, C.nj, l, jc, kc)] = 0.125 * acc;
  }
  }
}

// ---------------------------------------------------------------------------
// Velocity projection of the outer loop (press.for:77-144 calvel and 2-75
// calmas, 2nd-order branch) on the device.  Face arrays are dense, i
// fastest: u (nk, nj, ni+1) with u[.., f] the face between cells f-1 and f,
// v (nk, nj+1, ni), w (nk+1, nj, ni); cell arrays (nk, nj, ni).
//
// calvel: face f of axis a between cells f-1 and f,
//     u = ustar - dt*alfapr*(p(f) - p(f-1))/h
// on the faces the reference updates (isu..ieu, localparameters.for:
// 106-192): the interior faces and, on a periodic axis, the wrap face (its
// second copy f = 0 is what exchange(1) gives it); wall-normal faces keep
// ustar.  Pressure ghosts are images, as after pbound/exchange(4).
__global__ void __launch_bounds__(256)
    k_calvel(const FineDev *__restrict__ fd, int axis, const double *__restrict__ us,
             double *__restrict__ u, double dt_alfa, double h) {
  const FineDev &F = *fd;
  const int n[3] = {F.ni, F.nj, F.nk};
  int dim[3] = {F.ni, F.nj, F.nk};
  dim[axis] += 1;  // faces along `axis`
  const size_t total = (size_t)dim[0] * dim[1] * dim[2];
  const bool per = (F.per >> axis) & 1;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (size_t)gridDim.x * blockDim.x) {
    int c[3];
    c[0] = (int)(q % dim[0]);
    c[1] = (int)((q / dim[0]) % dim[1]);
    c[2] = (int)(q / ((size_t)dim[0] * dim[1]));
    const int f = c[axis];
    const bool upd = (f > 0 && f < n[axis]) || per;
    double v = us[q];
    if (upd) {
      int hi[3] = {c[0], c[1], c[2]}, lo[3] = {c[0], c[1], c[2]};
      hi[axis] = f == n[axis] ? 0 : f;        // wrap: cell 0 (the ghost image)
      lo[axis] = f == 0 ? n[axis] - 1 : f - 1;
      const double ph = F.p[(hi[0] + hi[1] + hi[2]) & 1][cidx(F.H, F.nj, hi[0], hi[1], hi[2])];
      const double pl = F.p[(lo[0] + lo[1] + lo[2]) & 1][cidx(F.H, F.nj, lo[0], lo[1], lo[2])];
      if (F.dens) {
        // L_LSM: divided by the face density (press.for:93-96)
        const double dl = F.dens[((size_t)lo[2] * F.nj + lo[1]) * F.ni + lo[0]];
        const double dh = F.dens[((size_t)hi[2] * F.nj + hi[1]) * F.ni + hi[0]];
        v = v - dt_alfa * (ph - pl) / h / (0.5 * (dl + dh));
      } else {
        v = v - dt_alfa * (ph - pl) / h;
      }
    }
    u[q] = v;
  }
}

// calvel for blocks with neighbours (the same update): face f of `axis` is
// computed for 0 < f < n and, when a neighbour or a periodic wrap lies
// across the block's far face, for f = n (the reference's isu..ieu,
// localparameters.for:106-192); the near face f = 0 is not written here --
// it is the neighbour's far face, which exchange(1..3) copies
// (k_face_copy); wall-normal faces keep u*.  Pressure ghosts are read
// through resolve_cell (pbound/exchange(4)).
__global__ void __launch_bounds__(256)
    k_calvel_nb(const FineDev *__restrict__ fine, int b, int axis, const double *__restrict__ us,
                double *__restrict__ u, double dt_alfa, double h) {
  const FineDev &F = fine[b];
  const int n[3] = {F.ni, F.nj, F.nk};
  int dim[3] = {F.ni, F.nj, F.nk};
  dim[axis] += 1;
  const size_t total = (size_t)dim[0] * dim[1] * dim[2];
  const bool far_open = !wall_face(F, 2 * axis + 1), near_open = !wall_face(F, 2 * axis);
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (size_t)gridDim.x * blockDim.x) {
    int c[3];
    c[0] = (int)(q % dim[0]);
    c[1] = (int)((q / dim[0]) % dim[1]);
    c[2] = (int)(q / ((size_t)dim[0] * dim[1]));
    const int f = c[axis];
    if (f == 0 && near_open) continue;
    const bool upd = (f > 0 && f < n[axis]) || (f == n[axis] && far_open);
    double v = us[q];
    if (upd) {
      int lo[3] = {c[0], c[1], c[2]};
      lo[axis] = f - 1;
      const double ph = fine_value(fine, b, c[0], c[1], c[2]);
      const double pl = fine_value(fine, b, lo[0], lo[1], lo[2]);
      v = v - dt_alfa * (ph - pl) / h;
    }
    u[q] = v;
  }
}

// exchange(1..3) of one face: the near face (f = 0) of `axis` in dst takes
// the far face (f = n_src) of the block across it (a neighbour, or the block
// itself on a periodic axis); both blocks share the face's extent (nu, nv =
// the two other axes' cell counts, in dst's order).
__global__ void __launch_bounds__(256)
    k_face_copy(int axis, const double *__restrict__ src, int n_src, double *__restrict__ dst,
                int n_dst, int nu, int nv) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)nu * nv) return;
  const size_t a = q % nu, c = q / nu;  // u faster than v
  // face arrays are (nk, nj, ni) with the face axis extended by one
  size_t qs, qd;
  if (axis == 0) {  // (c = k, a = j): u[(k*nj + j)*(ni+1) + f]
    qs = (c * nu + a) * (size_t)(n_src + 1) + n_src;
    qd = (c * nu + a) * (size_t)(n_dst + 1);
  } else if (axis == 1) {  // (c = k, a = i): v[(k*(nj+1) + f)*ni + i]
    qs = (c * (size_t)(n_src + 1) + n_src) * nu + a;
    qd = (c * (size_t)(n_dst + 1)) * nu + a;
  } else {  // (c = j, a = i): w[(f*nj + j)*ni + i]
    qs = ((size_t)n_src * nv + c) * nu + a;
    qd = c * nu + a;
  }
  dst[qd] = src[qs];
}

// calmas: su = ((du/dx + dv/dy) + dw/dz)/dt per cell, and per workgroup the
// max of |fact*dx*dy*dz| (rmax before its MPI_ALLREDUCE, press.for:59-73).
// su goes to the dense array `su` (when not null) or, with to_fine, into the
// block's colour-split f arrays (the next V-cycle's right-hand side).
__global__ void __launch_bounds__(256)
    k_calmas(const FineDev *__restrict__ fd, const double *__restrict__ u,
             const double *__restrict__ v, const double *__restrict__ w, double *__restrict__ su,
             int to_fine, double dt, double dx, double dy, double dz,
             RedPartial *__restrict__ partials) {
  const FineDev &F = *fd;
  const int ni = F.ni, nj = F.nj, nk = F.nk;
  const size_t total = (size_t)ni * nj * nk;
  double am = 0.0, ss = 0.0;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(q % ni), j = (int)((q / ni) % nj), k = (int)(q / ((size_t)ni * nj));
    const size_t qu = ((size_t)k * nj + j) * (ni + 1) + i;
    const size_t qv = ((size_t)k * (nj + 1) + j) * ni + i;
    const size_t qw = ((size_t)k * nj + j) * ni + i;
    const double fact = (u[qu + 1] - u[qu]) / dx + (v[qv + ni] - v[qv]) / dy +
                        (w[qw + (size_t)ni * nj] - w[qw]) / dz;
    // L_LSM: sup = div(u) dx dy dz / dt (press.for:53-55)
    const double rhs = F.dens ? fact * dx * dy * dz / dt : fact / dt;
    if (to_fine)
      F.f[(i + j + k) & 1][cidx(F.H, F.nj, i, j, k)] = rhs;
    else if (su)
      su[q] = rhs;
    const double fact1 = fact * dx * dy * dz;
    am = fmax(am, fabs(fact1));
  }
  // fixed-order workgroup reduction (block_reduce needs 256 = NTHREADS threads)
  block_reduce(ss, am);
  if (threadIdx.x == 0) partials[blockIdx.x] = RedPartial{0.0, am};
}

// max of the `amax` fields of n partials into *out (one workgroup; a max is
// exact in any order)
__global__ void __launch_bounds__(256)
    k_partials_max(const RedPartial *__restrict__ parts, int n, double *__restrict__ out) {
  double ss = 0.0, am = 0.0;
  for (int q = (int)threadIdx.x; q < n; q += 256) am = fmax(am, parts[q].amax);
  block_reduce(ss, am);
  if (threadIdx.x == 0) *out = am;
}

// calmas, differencing = 2 (press.for:17-37): per axis
//   (((-u(i+1) + 27 u(i)) - 27 u(i-1)) + u(i-2)) / (24 dx)
// in the reference's order (u(i) the east face of cell i: dense faces
// c+2, c+1, c, c-1 of cell c), the three axes summed x + y, then + z;
// su = fact/dt, rmax from |fact*dx*dy*dz| as in k_calmas.  Faces -1 and
// n+1 come from Outer4.
__global__ void __launch_bounds__(256)
    k_calmas4(const FineDev *__restrict__ fd, const double *__restrict__ u,
              const double *__restrict__ v, const double *__restrict__ w, double *__restrict__ su,
              int to_fine, double dt, double dx, double dy, double dz,
              RedPartial *__restrict__ partials, Outer4 o) {
  const FineDev &F = *fd;
  const int ni = F.ni, nj = F.nj, nk = F.nk;
  const size_t total = (size_t)ni * nj * nk;
  // face f of axis 0 at (j, k): own array for 0..ni, else the outer source
  auto uface = [&](int f, int j, int k) -> double {
    if (f >= 0 && f <= ni) return u[((size_t)k * nj + j) * (ni + 1) + f];
    const int s = f < 0 ? 0 : 1, m = o.n[s];
    if (m == 0) return o.src[s][(size_t)k * nj + j];
    return o.src[s][((size_t)k * nj + j) * (m + 1) + (s ? 1 : m - 1)];
  };
  auto vface = [&](int f, int i, int k) -> double {
    if (f >= 0 && f <= nj) return v[((size_t)k * (nj + 1) + f) * ni + i];
    const int s = f < 0 ? 0 : 1, m = o.n[2 + s];
    if (m == 0) return o.src[2 + s][(size_t)k * ni + i];
    return o.src[2 + s][((size_t)k * (m + 1) + (s ? 1 : m - 1)) * ni + i];
  };
  auto wface = [&](int f, int i, int j) -> double {
    if (f >= 0 && f <= nk) return w[((size_t)f * nj + j) * ni + i];
    const int s = f < 0 ? 0 : 1, m = o.n[4 + s];
    if (m == 0) return o.src[4 + s][(size_t)j * ni + i];
    return o.src[4 + s][((size_t)(s ? 1 : m - 1) * nj + j) * ni + i];
  };
  double am = 0.0, ss = 0.0;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(q % ni), j = (int)((q / ni) % nj), k = (int)(q / ((size_t)ni * nj));
    const double dux = ((-uface(i + 2, j, k) + 27.0 * uface(i + 1, j, k)) - 27.0 * uface(i, j, k)) +
                       uface(i - 1, j, k);
    const double dvy = ((-vface(j + 2, i, k) + 27.0 * vface(j + 1, i, k)) - 27.0 * vface(j, i, k)) +
                       vface(j - 1, i, k);
    const double dwz = ((-wface(k + 2, i, j) + 27.0 * wface(k + 1, i, j)) - 27.0 * wface(k, i, j)) +
                       wface(k - 1, i, j);
    const double fact = (dux / (24.0 * dx) + dvy / (24.0 * dy)) + dwz / (24.0 * dz);
    const double rhs = fact / dt;
    if (to_fine)
      F.f[(i + j + k) & 1][cidx(F.H, F.nj, i, j, k)] = rhs;
    else if (su)
      su[q] = rhs;
    const double fact1 = fact * dx * dy * dz;
    am = fmax(am, fabs(fact1));
  }
  block_reduce(ss, am);
  if (threadIdx.x == 0) partials[blockIdx.x] = RedPartial{0.0, am};
}

// ---- layout conversion ---------------------------------------------------
__global__ void k_pack_fine(const FineDev *__restrict__ fd, int which,
                            const double *__restrict__ src, int ni, int nj, int nk) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n = (size_t)ni * nj * nk;
  if (q >= n) return;
  const int i = (int)(q % ni), j = (int)((q / ni) % nj), k = (int)(q / ((size_t)ni * nj));
  const FineDev &F = *fd;
  double *dst = which == 0 ? F.p[(i + j + k) & 1] : F.f[(i + j + k) & 1];
  dst[cidx(F.H, F.nj, i, j, k)] = src[q];
}

__global__ void k_unpack_fine(const FineDev *__restrict__ fd, int which, double *__restrict__ dst,
                              int ni, int nj, int nk) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n = (size_t)ni * nj * nk;
  if (q >= n) return;
  const int i = (int)(q % ni), j = (int)((q / ni) % nj), k = (int)(q / ((size_t)ni * nj));
  const FineDev &F = *fd;
  const double *src = which == 0 ? F.p[(i + j + k) & 1] : F.f[(i + j + k) & 1];
  dst[q] = src[cidx(F.H, F.nj, i, j, k)];
}

// p of block b with its one-cell ghost layer, dense (nk+2, nj+2, ni+2), i
// fastest; ghosts are the cells resolve_cell names (pbound / exchange(4))
__global__ void k_unpack_fine_padded(const FineDev *__restrict__ fine, int b,
                                     double *__restrict__ dst) {
  const FineDev &F = fine[b];
  const int pi = F.ni + 2, pj = F.nj + 2, pk = F.nk + 2;
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)pi * pj * pk) return;
  const int x = (int)(q % pi) - 1, y = (int)((q / pi) % pj) - 1, z = (int)(q / ((size_t)pi * pj)) - 1;
  dst[q] = fine_value(fine, b, x, y, z);
}

// which 0: phi (dense padded (n+2)^3 <-> buffer buf), 1: rhs (dense interior)
template <class T>
__global__ void k_pack_level(const LevelDev *__restrict__ ld, int which, int buf,
                             const double *__restrict__ src, int ni, int nj, int nk) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n = (size_t)ni * nj * nk;
  if (q >= n) return;
  const int i = (int)(q % ni), j = (int)((q / ni) % nj), k = (int)(q / ((size_t)ni * nj));
  const LevelDev &L = *ld;
  const size_t c = cidx(L.H, L.nj, i, j, k);
  const int col = (i + j + k) & 1;
  if (which == 0) {
    const size_t pi = (size_t)ni + 2, pj = (size_t)nj + 2;
    ((T *)L.phi[buf][col])[c] = (T)src[(size_t)(i + 1) + pi * ((size_t)(j + 1) + pj * (size_t)(k + 1))];
  } else {
    ((T *)L.rhs[col])[c] = (T)src[q];
  }
}

template <class T>
__global__ void k_unpack_level(const LevelDev *__restrict__ levs, int d, int which, int buf,
                               double *__restrict__ dst, int ni, int nj, int nk) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const LevelDev &L = levs[d];
  if (which == 0) {
    const int pi = ni + 2, pj = nj + 2, pk = nk + 2;
    const size_t n = (size_t)pi * pj * pk;
    if (q >= n) return;
    const int x = (int)(q % pi) - 1, y = (int)((q / pi) % pj) - 1, z = (int)(q / ((size_t)pi * pj)) - 1;
    dst[q] = (double)level_value<T>(levs, d, buf, x, y, z);
  } else {
    const size_t n = (size_t)ni * nj * nk;
    if (q >= n) return;
    const int i = (int)(q % ni), j = (int)((q / ni) % nj), k = (int)(q / ((size_t)ni * nj));
    dst[q] = (double)((const T *)L.rhs[(i + j + k) & 1])[cidx(L.H, L.nj, i, j, k)];
  }
}

// mean removal of p (FP64): one workgroup, fixed-order reduction
__global__ void __launch_bounds__(256)
    k_remove_mean(const FineDev *__restrict__ fd, int ni, int nj, int nk) {
  const FineDev &F = *fd;
  const size_t n = (size_t)ni * nj * nk;
  double s = 0.0;
  for (size_t q = threadIdx.x; q < n; q += 256) {
    const int i = (int)(q % ni), j = (int)((q / ni) % nj), k = (int)(q / ((size_t)ni * nj));
    s += F.p[(i + j + k) & 1][cidx(F.H, F.nj, i, j, k)];
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  const double mean = sh[0] / (double)n;
  for (size_t q = threadIdx.x; q < n; q += 256) {
    const int i = (int)(q % ni), j = (int)((q / ni) % nj), k = (int)(q / ((size_t)ni * nj));
    double &v = F.p[(i + j + k) & 1][cidx(F.H, F.nj, i, j, k)];
    v = v - mean;
  }
}

inline dim3 tile_block() { return dim3(TILE_X, TILE_Y, 1); }


// ---------------------------------------------------------------------------
// Lexicographic Gauss-Seidel (mgsolver.for:407-432) and line TDMA
// (mgsolver.for:434-549), mg_itrsch = 1 and 2.
//
// The reference updates in place, in loop order.  A GS cell (loop k, j, i)
// reads its W, S and B neighbours after their update and E, N, T before it,
// so the cells of the hyperplane i+j+k = d depend on plane d-1 (updated) and
// plane d+1 (not yet) only: a hyperplane wavefront performs exactly the
// reference's arithmetic (k_gs_tile runs it over 8^3 tiles).  The TDMA
// passes are line-wise GS over the other two axes: j-lines (loop k, i) on
// diagonals i+k, i-lines (loop k, j) on j+k, k-lines (loop i, j) on i+j.  A
// sweep reads buffer src and writes 1-src:
//   GS and the j-line pass: values not yet updated from src, updated ones
//     from 1-src;
//   i-line and k-line passes: in place in 1-src;
//   ghosts, every pass: their images in src -- the pre-sweep state the
//     reference's ghosts keep until mgbound + exchangepp at the end of the
//     sweep (mgsolver.for:614-615).
// zero != 0: buffer src holds zeros.

// element e of a level array, read as global memory (no flat access)
__device__ __forceinline__ double gload(const void *base, size_t e) {
  return ((const gmem<double> *)base)[e];
}

// coefficients of one cell: the level's constant operator, or the
// variable-density arrays
__device__ __forceinline__ Coef<double> smoother_coef(const LevelDev &L, int i, int j, int k) {
  if (!L.lsm) return cell_coef<double>(L, i, j, k);
  const int colour = (i + j + k) & 1;
  const size_t e = cidx(L.H, L.nj, i, j, k);
  Coef<double> c;
  c.aw = gload(lsm_coef(L, 0, colour), e);
  c.ae = gload(lsm_coef(L, 1, colour), e);
  c.as = gload(lsm_coef(L, 2, colour), e);
  c.an = gload(lsm_coef(L, 3, colour), e);
  c.ab = gload(lsm_coef(L, 4, colour), e);
  c.at = gload(lsm_coef(L, 5, colour), e);
  c.ap = gload(lsm_coef(L, 6, colour), e);
  return c;
}

__device__ __forceinline__ double cell_at(const LevelDev &L, int buf, int i, int j, int k) {
  return gload(L.phi[buf][(i + j + k) & 1], cidx(L.H, L.nj, i, j, k));
}

// ghost or not-yet-updated value in buffer src (ghosts resolved to their
// image)
__device__ __forceinline__ double before_sweep(const LevelDev &L, int src, int zero, int x, int y,
                                               int z) {
  if (zero) return 0.0;
  x = image(x, L.ni, L.per & 1);
  y = image(y, L.nj, L.per & 2);
  z = image(z, L.nk, L.per & 4);
  return cell_at(L, src, x, y, z);
}

// One GS tile of GT^3 cells (Tile::x0, y0, z0): the tiles of one
// tile-diagonal I+J+K = D depend only on tiles of D-1 (updated) and D+1
// (not yet), as cells of a hyperplane do, so one launch per tile-diagonal
// repeats the reference's loop order, whatever the tile size.  The tile and
// its six face halos go to LDS -- the tile's cells and its east/north/top
// halo as before the sweep (src), its west/south/bottom halo as updated
// (dst, finished by earlier tile-diagonals), ghost positions as images of
// src -- and the wavefront walks the tile's own hyperplanes in place:
//   phi = (rhs - (aw*W + ae*E + as*S + an*N + ab*B + at*T)) / ap
// Thread (lx, ly) owns lz = h - lx - ly of plane h.  8^3 tiles (one
// wavefront, static LDS, variable-density coefficients staged too) and 16^3
// tiles (four wavefronts, constant coefficients, 79 KB of dynamic LDS:
// fewer tile-diagonals, so fewer launches per sweep).
template <int GT>
__device__ __forceinline__ void gs_tile_run(const Tile &t, const LevelDev &L, int src, int zero,
                                            double *v, double *rl, double *cl) {
  constexpr int GP = GT + 2;  // with the halo
  const int dst = 1 - src;
  const int tid = (int)threadIdx.x;
  const bool var = cl != nullptr && L.lsm != nullptr;
  // every load of the tile is issued before the first LDS store, with no
  // branch between them (clamped positions load an in-range cell whose
  // value is not stored or never read)
  constexpr int NT = GT * GT;
  constexpr int NR = GT * GT * GT / NT, NV = (GP * GP * GP + NT - 1) / NT;
  const int ni0 = L.ni, nj0 = L.nj, nk0 = L.nk, H0 = L.H;
  const bool px0 = L.per & 1, py0 = L.per & 2, pz0 = L.per & 4;
  // the level's phi arrays as byte offsets from src's colour 0
  const gmem<char> *const s0 = (const gmem<char> *)L.phi[src][0];
  const long long o_s1 = (const char *)L.phi[src][1] - (const char *)L.phi[src][0];
  const long long o_d0 = (const char *)L.phi[dst][0] - (const char *)L.phi[src][0];
  const long long o_d1 = (const char *)L.phi[dst][1] - (const char *)L.phi[src][0];
  double rr[NR], vv[NV];
#pragma unroll
  for (int it = 0; it < NR; ++it) {
    const int q = tid + it * NT;
    const int x = min(t.x0 + q % GT, ni0 - 1), y = min(t.y0 + (q / GT) % GT, nj0 - 1);
    const int z = min(t.z0 + q / (GT * GT), nk0 - 1);
    rr[it] = gload(L.rhs[(x + y + z) & 1], cidx(H0, nj0, x, y, z));
  }
#pragma unroll
  for (int it = 0; it < NV; ++it) {
    const int q = min(tid + it * NT, GP * GP * GP - 1);
    const int lz = q / (GP * GP) - 1, ly = (q / GP) % GP - 1, lx = q % GP - 1;
    int x = t.x0 + lx, y = t.y0 + ly, z = t.z0 + lz;
    const bool ghost = (unsigned)x >= (unsigned)ni0 || (unsigned)y >= (unsigned)nj0 ||
                       (unsigned)z >= (unsigned)nk0;
    // positions more than one cell outside the level (halo edges, corners
    // and the far side of a partial tile) are clamped; their values are
    // never read
    x = image(max(min(x, ni0), -1), ni0, px0);
    y = image(max(min(y, nj0), -1), nj0, py0);
    z = image(max(min(z, nk0), -1), nk0, pz0);
    // interior cells of the low-side halo: updated (dst); the tile, its
    // high-side halo and ghost images: before the sweep (src)
    const bool from_dst = !ghost && (lx < 0 || ly < 0 || lz < 0);
    const bool odd = (x + y + z) & 1;
    const long long o = from_dst ? (odd ? o_d1 : o_d0) : (odd ? o_s1 : 0);
    const double val =
        *(const gmem<double> *)(s0 + o + 8 * (((long long)z * nj0 + y) * H0 + (x >> 1)));
    vv[it] = (zero && !from_dst) ? 0.0 : val;
  }
#pragma unroll
  for (int it = 0; it < NR; ++it) rl[tid + it * NT] = rr[it];
#pragma unroll
  for (int it = 0; it < NV; ++it)
    if (tid + it * NT < GP * GP * GP) v[tid + it * NT] = vv[it];
  if (var)
    for (int q = tid; q < GT * GT * GT; q += NT) {
      const int x = t.x0 + q % GT, y = t.y0 + (q / GT) % GT, z = t.z0 + q / (GT * GT);
      if (x >= L.ni || y >= L.nj || z >= L.nk) continue;
      const size_t e = cidx(L.H, L.nj, x, y, z);
      for (int a = 0; a < 7; ++a) cl[a * GT * GT * GT + q] = gload(lsm_coef(L, a, (x + y + z) & 1), e);
    }
  __syncthreads();
  const int lx = tid % GT, ly = tid / GT;
  const int i = t.x0 + lx, j = t.y0 + ly;
  // the level's fields in registers: a barrier would make the compiler
  // reload them from memory on every plane
  const int ni = L.ni, nj = L.nj, nk = L.nk;
  const double cx = L.cx, cy = L.cy, cz = L.cz;
  const bool px = L.per & 1, py = L.per & 2, pz = L.per & 4;
  // the planes touch LDS only: a barrier then waits for no global store
  for (int h = 0; h < 3 * GT - 2; ++h) {
    const int lz = h - lx - ly, k = t.z0 + lz;
    if (lz >= 0 && lz < GT && i < ni && j < nj && k < nk) {
      const int c0 = ((lz + 1) * GP + ly + 1) * GP + lx + 1;
      const int q = (lz * GT + ly) * GT + lx;
      Coef<double> c;
      if (var) {
        constexpr int S = GT * GT * GT;
        c.aw = cl[q], c.ae = cl[S + q], c.as = cl[2 * S + q], c.an = cl[3 * S + q];
        c.ab = cl[4 * S + q], c.at = cl[5 * S + q], c.ap = cl[6 * S + q];
      } else {  // cell_coef
        c.aw = (!px && i == 0) ? 0.0 : cx;
        c.ae = (!px && i == ni - 1) ? 0.0 : cx;
        c.as = (!py && j == 0) ? 0.0 : cy;
        c.an = (!py && j == nj - 1) ? 0.0 : cy;
        c.ab = (!pz && k == 0) ? 0.0 : cz;
        c.at = (!pz && k == nk - 1) ? 0.0 : cz;
        c.ap = -1.0 * (c.aw + c.ae + c.as + c.an + c.ab + c.at);
      }
      const double W = v[c0 - 1], E = v[c0 + 1];
      const double S = v[c0 - GP], N = v[c0 + GP];
      const double B = v[c0 - GP * GP], T = v[c0 + GP * GP];
      const double r = rl[q];
      const double sum = c.aw * W + c.ae * E + c.as * S + c.an * N + c.ab * B + c.at * T;
      v[c0] = (r - sum) / c.ap;
    }
    __syncthreads();
  }
  // the tile's column (lx, ly) to dst
  gmem<double> *const out0 = (gmem<double> *)L.phi[dst][0];
  gmem<double> *const out1 = (gmem<double> *)L.phi[dst][1];
  if (i < ni && j < nj)
    for (int lz = 0; lz < GT && t.z0 + lz < nk; ++lz) {
      const int k = t.z0 + lz;
      (((i + j + k) & 1) ? out1 : out0)[cidx(L.H, L.nj, i, j, k)] =
          v[((lz + 1) * GP + ly + 1) * GP + lx + 1];
    }
}

__global__ void __launch_bounds__(GS_TILE * GS_TILE)
    k_gs_tile(const Tile *__restrict__ tiles, const LevelDev *__restrict__ levs,
              const int *__restrict__ active, int src, int zero) {
  constexpr int GP = GS_TILE + 2;
  __shared__ double v[GP * GP * GP];                    // phi, tile + face halos
  __shared__ double rl[GS_TILE * GS_TILE * GS_TILE];      // rhs of the tile
  __shared__ double cl[7 * GS_TILE * GS_TILE * GS_TILE];  // variable-density coefficients
  const Tile t = tiles[blockIdx.x];
  const LevelDev &L = levs[t.lev];
  if (!active[L.blk]) return;
  gs_tile_run<GS_TILE>(t, L, src, zero, v, rl, cl);
}

// 16^3 tiles, constant coefficients only (the host never launches it on a
// variable-density level): GS_BIG_LDS bytes of dynamic LDS
__global__ void __launch_bounds__(GS_BIG_TILE * GS_BIG_TILE)
    k_gs_tile_big(const Tile *__restrict__ tiles, const LevelDev *__restrict__ levs,
                  const int *__restrict__ active, int src, int zero) {
  extern __shared__ double gs_lds[];
  constexpr int GP = GS_BIG_TILE + 2;
  const Tile t = tiles[blockIdx.x];
  const LevelDev &L = levs[t.lev];
  if (!active[L.blk] || L.lsm) return;
  gs_tile_run<GS_BIG_TILE>(t, L, src, zero, gs_lds, gs_lds + GP * GP * GP, nullptr);
}

// positions of one line whose coefficients are loaded together
constexpr int TDMA_CHUNK = 8;

// cell (i, j, k) of position p on the line (u, v) of pass PASS
template <int PASS>
__device__ __forceinline__ void tdma_cell(int u, int v, int p, int &i, int &j, int &k) {
  if (PASS == 0) {
    i = u, j = p, k = v;
  } else if (PASS == 1) {
    i = p, j = u, k = v;
  } else {
    i = u, j = v, k = p;
  }
}

// Wall state of a cross coordinate c of a line (n cells on its axis): bit 0
// at the low wall, bit 1 at the high wall, 0 on a periodic axis.  Lines of
// one pass and state pair have the same coefficients along their length
// (cell_coef), so the same Thomas a(p) and 1/(ap - alo*a(p-1)).
__device__ __forceinline__ int wall_state(int c, int n, bool per) {
  return per ? 0 : (c == 0 ? 1 : 0) | (c == n - 1 ? 2 : 0);
}
// a coordinate of state s, or -1 when none has it
__device__ __forceinline__ int wall_rep(int s, int n, bool per) {
  if (per) return s == 0 ? 0 : -1;
  if (s == 0) return n >= 3 ? 1 : -1;
  if (s == 1 || s == 2) return n >= 2 ? (s == 1 ? 0 : n - 1) : -1;
  return n == 1 ? 0 : -1;
}

// geometry of the lines of pass PASS: lines (u, v), len cells each
template <int PASS>
struct TdmaGeom {
  int nu, nv, len;
  bool pu, pv;  // u / v axis periodic
  __device__ explicit TdmaGeom(const LevelDev &L) {
    nu = PASS == 1 ? L.nj : L.ni;
    nv = PASS == 2 ? L.nj : L.nk;
    len = PASS == 0 ? L.nj : PASS == 1 ? L.ni : L.nk;
    pu = PASS == 1 ? (L.per & 2) : (L.per & 1);
    pv = PASS == 2 ? (L.per & 2) : (L.per & 4);
  }
};

// (a(p), 1/(ap - alo*a(p-1))) of pass PASS, state pair `var`, position p
// of a constant-coefficient level: L.tdma + tdma_table_offset + 2*(var*len+p)
template <int PASS>
__device__ __forceinline__ size_t tdma_table_offset(const LevelDev &L) {
  return PASS == 0 ? 0 : PASS == 1 ? (size_t)32 * L.nj : (size_t)32 * (L.nj + L.ni);
}

template <int PASS>
__device__ __forceinline__ void tdma_table_fill(const LevelDev &L, int var) {
  const TdmaGeom<PASS> G(L);
  const int ru = wall_rep(var & 3, G.nu, G.pu), rv = wall_rep(var >> 2, G.nv, G.pv);
  if (ru < 0 || rv < 0) return;
  gmem<double> *tab = (gmem<double> *)(L.tdma + tdma_table_offset<PASS>(L) + (size_t)2 * var * G.len);
  double aprev = 0.0;
  for (int p = 0; p < G.len; ++p) {
    int i, j, k;
    tdma_cell<PASS>(ru, rv, p, i, j, k);
    const Coef<double> c = cell_coef<double>(L, i, j, k);
    const double alo = PASS == 0 ? c.as : PASS == 1 ? c.aw : c.ab;
    const double ahi = PASS == 0 ? c.an : PASS == 1 ? c.ae : c.at;
    const double apr = 1.0 / (c.ap - alo * aprev);
    aprev = ahi * apr;
    tab[2 * p] = aprev;
    tab[2 * p + 1] = apr;
  }
}

// a thread per (descriptor, pass, state pair)
__global__ void __launch_bounds__(256)
    k_tdma_tables(const LevelDev *__restrict__ levs, int ndesc) {
  const int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (q >= ndesc * 48) return;
  const LevelDev &L = levs[q / 48];
  if (!L.tdma) return;
  const int pass = (q / 16) % 3, var = q % 16;
  if (pass == 0)
    tdma_table_fill<0>(L, var);
  else if (pass == 1)
    tdma_table_fill<1>(L, var);
  else
    tdma_table_fill<2>(L, var);
}

// One diagonal d of one TDMA pass for the lines u0 .. u0+lw-1 (all threads
// of the workgroup call it; it holds one barrier), in two phases:
//  A (every thread, a cell at a time): the right-hand side of the Thomas
//    recurrence without its last term,
//      rs = rhs - x1*X1 - x2*X2 - x3*X3 - x4*X4,
//    with the cross terms in the reference's order -- j-lines ae*E, aw*W,
//    at*T, ab*B (mgsolver.for:459-466); i-lines an*N, as*S, at*T, ab*B
//    (:496-503); k-lines an*N, as*S, ae*E, aw*W (:533-540) -- into LDS at
//    xs[p*lwp + line].  The neighbours on the previous diagonal (W, B; S, B;
//    W, S) are updated ones.  On a constant-coefficient level the cell's
//    a(p) and 1/(ap - alo*a(p-1)) also go to LDS, from its line's table
//    (the same values the recurrence gives: a does not depend on the data).
//  B (a thread per line): the reference's recurrence (a(0) = 0, c(0) = the
//    ghost),
//      apr = 1/(ap - alo*a(p-1)),  a(p) = ahi*apr,
//      c(p) = (rs - alo*c(p-1))*apr,
//      phi(p) = c(p) - a(p)*phi(p+1),  phi(n) = the ghost,
//    (alo, ahi) = (as, an), (aw, ae), (ab, at); c overwrites rs in LDS.
//    Operands are read a chunk of positions ahead of the dependent chain.
// The subtraction of alo*c(p-1) is the last one of the reference's
// left-to-right sum, so splitting it off changes no rounding.  Lines of one
// diagonal never neighbour each other: phase A reads only lines of the
// diagonals before and after, which nothing writes meanwhile.
template <int PASS, int CPT, int NT = NTHREADS, bool PREV = false>
__device__ __forceinline__ void tdma_diagonal(const LevelDev &L, int u0, int lw, int lwp, int d,
                                              int src, int zero, double *xs, double *av,
                                              double *rv, int tid, double *prev = nullptr) {
  const TdmaGeom<PASS> G(L);
  const int nu = G.nu, nv = G.nv, len = G.len;
  const int ulo = max(u0, d - nv + 1), uhi = min(min(u0 + lw, nu), d + 1);
  const int dst = 1 - src;
  const bool tab = L.lsm == nullptr;
  const int ni = L.ni, nj = L.nj, nk = L.nk, H = L.H;
  const bool px = L.per & 1, py = L.per & 2, pz = L.per & 4;
  // the four phi arrays of the level as byte offsets from src's colour 0
  const gmem<char> *const s0 = (const gmem<char> *)L.phi[src][0];
  const long long o_s1 = (const char *)L.phi[src][1] - (const char *)L.phi[src][0];
  const long long o_d0 = (const char *)L.phi[dst][0] - (const char *)L.phi[src][0];
  const long long o_d1 = (const char *)L.phi[dst][1] - (const char *)L.phi[src][0];
  // value of a cross neighbour, with no branch: ghosts are images in src;
  // interior neighbours updated ones and (after the first pass) not-yet-
  // updated ones in dst, the first pass's others in src; src reads are 0
  // when src holds zeros
  auto cross = [=](int x, int y, int z, bool updated) -> double {
    const bool ghost = (unsigned)x >= (unsigned)ni || (unsigned)y >= (unsigned)nj ||
                       (unsigned)z >= (unsigned)nk;
    x = image(x, ni, px);
    y = image(y, nj, py);
    z = image(z, nk, pz);
    const bool from_dst = !ghost && (updated || PASS > 0);
    const bool odd = (x + y + z) & 1;
    const long long o = from_dst ? (odd ? o_d1 : o_d0) : (odd ? o_s1 : 0);
    const double val = *(const gmem<double> *)(s0 + o + 8 * (((long long)z * nj + y) * H + (x >> 1)));
    return (zero && !from_dst) ? 0.0 : val;
  };
  // the updated neighbours of line (u, v) lie on lines (u-1, v) and (u, v-1)
  // of diagonal d-1, at the same position: with `prev` (the previous
  // diagonal's solution in LDS, line u at prev[u*len + p]) they are read
  // from there -- the values the previous diagonal stored in dst -- instead
  // of from global memory
  auto upd = [=](int u, int v, int p, bool along_u, int x, int y, int z) -> double {
    if constexpr (PREV) {
      if (along_u ? u > 0 : v > 0) return prev[(along_u ? u - 1 : u) * len + p];
    }
    return cross(x, y, z, true);
  };
  // phase A: batches of CPT cells per thread, every load of a batch issued
  // before its first LDS store (positions past the end repeat the last cell
  // and are not stored)
  const int ncell = (uhi - ulo) * len;
  // q / len through a float reciprocal: exact while q < 2^12 * len
  const float inv_len = 1.0f / (float)len;
  for (int q0 = 0; q0 < ncell; q0 += CPT * NT) {
    double rsv[CPT], atv[CPT], rtv[CPT];
    int slot[CPT];
#pragma unroll
    for (int it = 0; it < CPT; ++it) {
      const int q = min(q0 + tid + it * NT, ncell - 1);
      const int lq = (int)(((float)q + 0.5f) * inv_len);
      const int l = ulo - u0 + lq, p = q - lq * len;
      const int u = u0 + l, v = d - u;
      int i, j, k;
      tdma_cell<PASS>(u, v, p, i, j, k);
      const Coef<double> c = smoother_coef(L, i, j, k);
      const double r = gload(L.rhs[(i + j + k) & 1], cidx(H, nj, i, j, k));
      double rs;
      if (PASS == 0)  // W on line (u-1, v), B on (u, v-1)
        rs = r - c.ae * cross(i + 1, j, k, false) - c.aw * upd(u, v, p, true, i - 1, j, k) -
             c.at * cross(i, j, k + 1, false) - c.ab * upd(u, v, p, false, i, j, k - 1);
      else if (PASS == 1)  // S on (u-1, v), B on (u, v-1)
        rs = r - c.an * cross(i, j + 1, k, false) - c.as * upd(u, v, p, true, i, j - 1, k) -
             c.at * cross(i, j, k + 1, false) - c.ab * upd(u, v, p, false, i, j, k - 1);
      else  // S on (u, v-1), W on (u-1, v)
        rs = r - c.an * cross(i, j + 1, k, false) - c.as * upd(u, v, p, false, i, j - 1, k) -
             c.ae * cross(i + 1, j, k, false) - c.aw * upd(u, v, p, true, i - 1, j, k);
      rsv[it] = rs;
      slot[it] = p * lwp + l;
      atv[it] = rtv[it] = 0.0;
      if (tab) {
        const int var = wall_state(u, nu, G.pu) | (wall_state(v, nv, G.pv) << 2);
        const size_t t = tdma_table_offset<PASS>(L) + (size_t)2 * (var * len + p);
        atv[it] = gload(L.tdma, t);
        rtv[it] = gload(L.tdma, t + 1);
      }
    }
#pragma unroll
    for (int it = 0; it < CPT; ++it)
      if (q0 + tid + it * NT < ncell) {
        xs[slot[it]] = rsv[it];
        if (tab) {
          av[slot[it]] = atv[it];
          rv[slot[it]] = rtv[it];
        }
      }
  }
  __syncthreads();
  const int l = tid, u = u0 + l, v = d - u;
  if (l >= lw || u < ulo || u >= uhi) return;
  int i, j, k;
  tdma_cell<PASS>(u, v, -1, i, j, k);
  double cprev = before_sweep(L, src, zero, i, j, k);
  if (tab) {
    // constant coefficients: alo is 1/h^2 of the line's axis, 0 at its
    // first cell next to a wall (cell_coef)
    const double cax = PASS == 0 ? L.cy : PASS == 1 ? L.cx : L.cz;
    const bool pax = PASS == 0 ? (L.per & 2) : PASS == 1 ? (L.per & 1) : (L.per & 4);
    for (int p0 = 0; p0 < len; p0 += TDMA_CHUNK) {
      // a chunk's operands, read with no branch in between (positions past
      // the line's end read its last cell and are not used)
      double rs[TDMA_CHUNK], apr[TDMA_CHUNK];
#pragma unroll
      for (int q = 0; q < TDMA_CHUNK; ++q) {
        const int p = min(p0 + q, len - 1);
        rs[q] = xs[p * lwp + l];
        apr[q] = rv[p * lwp + l];
      }
#pragma unroll
      for (int q = 0; q < TDMA_CHUNK; ++q) {
        const int p = p0 + q;
        if (p >= len) break;
        const double alo = (p == 0 && !pax) ? 0.0 : cax;
        cprev = (rs[q] - alo * cprev) * apr[q];
        xs[p * lwp + l] = cprev;
      }
    }
  } else {
    double aprev = 0.0;
    for (int p0 = 0; p0 < len; p0 += TDMA_CHUNK) {
      double rs[TDMA_CHUNK], ap[TDMA_CHUNK], alo[TDMA_CHUNK], ahi[TDMA_CHUNK];
#pragma unroll
      for (int q = 0; q < TDMA_CHUNK; ++q) {
        const int p = min(p0 + q, len - 1);
        tdma_cell<PASS>(u, v, p, i, j, k);
        const Coef<double> c = smoother_coef(L, i, j, k);
        rs[q] = xs[p * lwp + l];
        ap[q] = c.ap;
        alo[q] = PASS == 0 ? c.as : PASS == 1 ? c.aw : c.ab;
        ahi[q] = PASS == 0 ? c.an : PASS == 1 ? c.ae : c.at;
      }
#pragma unroll
      for (int q = 0; q < TDMA_CHUNK; ++q) {
        const int p = p0 + q;
        if (p >= len) break;
        const double a_r = 1.0 / (ap[q] - alo[q] * aprev);
        aprev = ahi[q] * a_r;
        cprev = (rs[q] - alo[q] * cprev) * a_r;
        xs[p * lwp + l] = cprev;
        av[p * lwp + l] = aprev;
      }
    }
  }
  tdma_cell<PASS>(u, v, len, i, j, k);
  double next = before_sweep(L, src, zero, i, j, k);
  // colour-split index of position p: base + p*step (j- and k-lines) or
  // base + p/2 (i-lines); colour (u+v+p) & 1.  The level's fields stay in
  // registers across the stores.
  gmem<double> *const out0 = (gmem<double> *)L.phi[dst][0];
  gmem<double> *const out1 = (gmem<double> *)L.phi[dst][1];
  const int base = PASS == 0 ? v * nj * H + (u >> 1) : PASS == 1 ? (v * nj + u) * H : v * H + (u >> 1);
  const int step = PASS == 0 ? H : PASS == 1 ? 0 : nj * H;
  for (int p1 = len - 1; p1 >= 0; p1 -= TDMA_CHUNK) {
    double cv[TDMA_CHUNK], aq[TDMA_CHUNK];
#pragma unroll
    for (int q = 0; q < TDMA_CHUNK; ++q) {
      const int p = max(p1 - q, 0);
      cv[q] = xs[p * lwp + l];
      aq[q] = av[p * lwp + l];
    }
#pragma unroll
    for (int q = 0; q < TDMA_CHUNK; ++q) {
      const int p = p1 - q;
      if (p < 0) break;
      next = cv[q] - aq[q] * next;
      const int e = PASS == 1 ? base + (p >> 1) : base + p * step;
      (((u + v + p) & 1) ? out1 : out0)[e] = next;
      if (PREV) prev[u * len + p] = next;
    }
  }
}

// one diagonal of one pass over groups of tdma_group_lines(len) lines
// (Tile::x0 = the group's first line)
template <int PASS>
__global__ void __launch_bounds__(NTHREADS)
    k_tdma_diag(const Tile *__restrict__ tiles, const LevelDev *__restrict__ levs,
                const int *__restrict__ active, int d, int src, int zero) {
  __shared__ double xs[2 * TDMA_CELLS];  // rs, then c
  __shared__ double av[2 * TDMA_CELLS];  // a
  __shared__ double rv[2 * TDMA_CELLS];  // 1/(ap - alo*a(p-1)) (constant coefficients)
  const Tile t = tiles[blockIdx.x];
  const LevelDev &L = levs[t.lev];
  if (!active[L.blk]) return;
  const TdmaGeom<PASS> G(L);
  const int lw = tdma_group_lines(G.len), lwp = lw > 1 ? lw + 1 : 1;
  const int u0 = t.x0;
  if (max(u0, d - G.nv + 1) >= min(min(u0 + lw, G.nu), d + 1)) return;  // no line of d
  tdma_diagonal<PASS, TDMA_CELLS / NTHREADS>(L, u0, lw, lwp, d, src, zero, xs, av, rv,
                                            (int)threadIdx.x + TILE_X * (int)threadIdx.y);
}

// a whole TDMA sweep (three passes, every diagonal) of one small level in
// one workgroup per block: every line of a diagonal in dynamic LDS,
// 3 * (nu+1) * len doubles (tdma_sweep_lds_bytes)
template <int PASS>
__device__ __forceinline__ void tdma_pass_in_group(const LevelDev &L, int src, int zero, double *buf, int tid) {
  const TdmaGeom<PASS> G(L);
  const int lwp = G.nu + 1;
  const size_t n = (size_t)lwp * G.len;
  for (int d = 0; d < G.nu + G.nv - 1; ++d) {
    tdma_diagonal<PASS, TDMA_SWEEP_CPT, TDMA_SWEEP_THREADS, true>(
        L, 0, G.nu, lwp, d, src, zero, buf, buf + n, buf + 2 * n, tid, buf + 3 * n);
    __syncthreads();  // the diagonal's phi before the next one reads it
  }
}

// One TDMA pass of a constant-coefficient level in one workgroup, pipelined
// over the diagonals.  Diagonal d's right-hand sides need the solution of
// diagonal d-1 only through its two updated cross terms, so phase A splits:
//   A0 (everything that does not depend on diagonal d-1: the rhs, the two
//      not-updated cross terms, ghost images, the lines' end ghosts) runs
//      for diagonal d+1 on the waves that phase B does not use, while
//   B (the Thomas recurrence of diagonal d, one thread per line) runs on
//      the first waves;
//   A1 (the updated cross terms from the previous diagonal's solution in
//      LDS, and the last subtractions) runs for every cell between them.
// The reference's left-to-right sum r - x1*X1 - x2*X2 - x3*X3 - x4*X4 is
// kept: A0 forms t1 = r - x1*X1 and the product x3*X3, A1 finishes
// ((t1 - x2*X2) - x3*X3) - x4*X4.  X2 and X4 are the updated neighbours
// (pass 0: W on line (u-1, v) and B on (u, v-1); pass 1: S on (u-1, v), B
// on (u, v-1); pass 2: S on (u, v-1), W on (u-1, v)); a neighbour that is a
// ghost (its line coordinate 0) is an image before the sweep, taken by A0.
// The Thomas tables of the pass (a(p), 1/(ap - alo*a(p-1)) per wall state)
// sit in LDS for the whole pass.  LDS, doubles (tdma_sweep_lds_bytes):
// t1, x3*X3, rs/c: 3*(nu+1)*len; tables 32*len; previous diagonal nu*len;
// ghosts of the first line 2*len; line-end ghosts 2 x 2*nu.
template <int PASS>
__device__ __forceinline__ void tdma_pass_pipelined(const LevelDev &L, int src, int zero,
                                                    double *buf, int tid) {
  const TdmaGeom<PASS> G(L);
  const int nu = G.nu, nv = G.nv, len = G.len, lwp = nu + 1;
  const size_t n = (size_t)lwp * len;
  double *const t1 = buf, *const p3 = buf + n, *const xs = buf + 2 * n;
  double *const T = buf + 3 * n;
  double *const prev = T + (size_t)32 * len;
  double *const gA = prev + (size_t)nu * len;  // term 2's ghost values (line coordinate 0)
  double *const gB = gA + len;                 // term 4's ghost values
  double *const cg = gB + len;                 // [2][2*nu]: c(0) and phi(len) ghosts per line
  constexpr bool U2 = PASS != 2;               // term 2 lies along u (else along v)
  const int nb = 64 * ((nu + 63) / 64);        // phase-B threads (whole waves)
  const int na = TDMA_SWEEP_THREADS - nb;      // A0 threads
  const int dst = 1 - src;
  const int ni = L.ni, nj = L.nj, nk = L.nk, H = L.H;
  const bool px = L.per & 1, py = L.per & 2, pz = L.per & 4;
  const gmem<char> *const s0 = (const gmem<char> *)L.phi[src][0];
  const long long o_s1 = (const char *)L.phi[src][1] - (const char *)L.phi[src][0];
  const long long o_d0 = (const char *)L.phi[dst][0] - (const char *)L.phi[src][0];
  const long long o_d1 = (const char *)L.phi[dst][1] - (const char *)L.phi[src][0];
  // as tdma_diagonal's: ghosts are images in src; interior neighbours
  // updated ones and (after the first pass) not-yet-updated ones in dst
  auto cross = [=](int x, int y, int z, bool updated) -> double {
    const bool ghost = (unsigned)x >= (unsigned)ni || (unsigned)y >= (unsigned)nj ||
                       (unsigned)z >= (unsigned)nk;
    x = image(x, ni, px);
    y = image(y, nj, py);
    z = image(z, nk, pz);
    const bool from_dst = !ghost && (updated || PASS > 0);
    const bool odd = (x + y + z) & 1;
    const long long o = from_dst ? (odd ? o_d1 : o_d0) : (odd ? o_s1 : 0);
    const double val = *(const gmem<double> *)(s0 + o + 8 * (((long long)z * nj + y) * H + (x >> 1)));
    return (zero && !from_dst) ? 0.0 : val;
  };
  const float inv_len = 1.0f / (float)len;  // q / len, exact while q < 2^12 * len
  const int nd = nu + nv - 1;

  // A0 of diagonal d on threads [t0, t0 + cnt): CPT cells a batch, every
  // load of a batch issued before its first LDS store; then the "extra"
  // items: the ghost terms of the lines whose term-2 or term-4 coordinate is
  // 0 (one line each: items [0, 2*len)) and the lines' end ghosts c(0),
  // phi(len) (items [2*len, 2*len + 2*lines)).
  auto a0 = [=, &L](int d, int t0, int cnt) {
    constexpr int CPT = 5;
    const int ulo = max(0, d - nv + 1), uhi = min(nu, d + 1);
    const int ncell = (uhi - ulo) * len;
    // term 2's ghost line: u == 0 (passes 0, 1) or v == 0 (pass 2); term
    // 4's: v == 0 (passes 0, 1) or u == 0 (pass 2)
    const int ua = U2 ? 0 : d, ub = U2 ? d : 0;  // the line of each, if on d
    const bool has_a = ua >= ulo && ua < uhi, has_b = ub >= ulo && ub < uhi;
    const int nextra = 2 * len + 2 * (uhi - ulo);
    double *const cgd = cg + (d & 1) * 2 * nu;
    // item e: its value and its LDS index from buf (-1: nothing to store)
    struct Item {
      double v;
      int at;
    };
    auto extra = [=, &L](int e) -> Item {
      int i, j, k;
      if (e < 2 * len) {
        const bool second = e >= len;
        const int p = second ? e - len : e;
        if (second ? !has_b : !has_a) return Item{0.0, -1};
        const int u = second ? ub : ua, v = d - u;
        tdma_cell<PASS>(u, v, p, i, j, k);
        const int at = (int)(((second ? gB : gA) + p) - buf);
        if (PASS == 0)
          return Item{second ? cross(i, j, k - 1, true) : cross(i - 1, j, k, true), at};
        if (PASS == 1)
          return Item{second ? cross(i, j, k - 1, true) : cross(i, j - 1, k, true), at};
        return Item{second ? cross(i - 1, j, k, true) : cross(i, j - 1, k, true), at};
      }
      const int q = e - 2 * len, u = ulo + (q >> 1), v = d - u;
      tdma_cell<PASS>(u, v, (q & 1) ? len : -1, i, j, k);
      return Item{before_sweep(L, src, zero, i, j, k), (int)((cgd + 2 * u + (q & 1)) - buf)};
    };
    for (int q0 = 0; q0 < ncell; q0 += CPT * cnt) {
      double tv[CPT], pv[CPT];
      int slot[CPT];
#pragma unroll
      for (int it = 0; it < CPT; ++it) {
        const int q = min(q0 + t0 + it * cnt, ncell - 1);
        const int lq = (int)(((float)q + 0.5f) * inv_len);
        const int p = q - lq * len, u = ulo + lq, v = d - u;
        int i, j, k;
        tdma_cell<PASS>(u, v, p, i, j, k);
        const Coef<double> c = cell_coef<double>(L, i, j, k);
        const double r = gload(L.rhs[(i + j + k) & 1], cidx(H, nj, i, j, k));
        if (PASS == 0) {
          tv[it] = r - c.ae * cross(i + 1, j, k, false);
          pv[it] = c.at * cross(i, j, k + 1, false);
        } else if (PASS == 1) {
          tv[it] = r - c.an * cross(i, j + 1, k, false);
          pv[it] = c.at * cross(i, j, k + 1, false);
        } else {
          tv[it] = r - c.an * cross(i, j + 1, k, false);
          pv[it] = c.ae * cross(i + 1, j, k, false);
        }
        slot[it] = p * lwp + u;
      }
#pragma unroll
      for (int it = 0; it < CPT; ++it)
        if (q0 + t0 + it * cnt < ncell) {
          t1[slot[it]] = tv[it];
          p3[slot[it]] = pv[it];
        }
    }
    for (int e = t0; e < nextra; e += cnt) {
      const Item x = extra(e);
      if (x.at >= 0) buf[x.at] = x.v;
    }
  };

  // the pass's Thomas tables into LDS, and A0 of diagonal 0
  for (int q = tid; q < 16 * len; q += TDMA_SWEEP_THREADS) {
    const size_t t = tdma_table_offset<PASS>(L) + (size_t)2 * q;
    T[2 * q] = gload(L.tdma, t);
    T[2 * q + 1] = gload(L.tdma, t + 1);
  }
  a0(0, tid, TDMA_SWEEP_THREADS);
  __syncthreads();
  const double cax = PASS == 0 ? L.cy : PASS == 1 ? L.cx : L.cz;
  const bool pax = PASS == 0 ? (L.per & 2) : PASS == 1 ? (L.per & 1) : (L.per & 4);
  gmem<double> *const out0 = (gmem<double> *)L.phi[dst][0];
  gmem<double> *const out1 = (gmem<double> *)L.phi[dst][1];
  for (int d = 0; d < nd; ++d) {
    const int ulo = max(0, d - nv + 1), uhi = min(nu, d + 1);
    // A1: the updated cross terms and the last subtractions, every cell,
    // four a batch (every LDS read of a batch before its first store)
    const int ncell = (uhi - ulo) * len;
    for (int q0 = 0; q0 < ncell; q0 += 4 * TDMA_SWEEP_THREADS) {
      double rs[4];
      int sl[4];
#pragma unroll
      for (int it = 0; it < 4; ++it) {
        const int q = min(q0 + tid + it * TDMA_SWEEP_THREADS, ncell - 1);
        const int lq = (int)(((float)q + 0.5f) * inv_len);
        const int p = q - lq * len, u = ulo + lq, v = d - u;
        int i, j, k;
        tdma_cell<PASS>(u, v, p, i, j, k);
        const Coef<double> c = cell_coef<double>(L, i, j, k);
        const double x2 = PASS == 0 ? c.aw : c.as;
        const double x4 = PASS == 2 ? c.aw : c.ab;
        const bool in2 = U2 ? u > 0 : v > 0, in4 = U2 ? v > 0 : u > 0;
        const double X2 = in2 ? prev[(U2 ? u - 1 : u) * len + p] : gA[p];
        const double X4 = in4 ? prev[(U2 ? u : u - 1) * len + p] : gB[p];
        sl[it] = p * lwp + u;
        rs[it] = ((t1[sl[it]] - x2 * X2) - p3[sl[it]]) - x4 * X4;
      }
#pragma unroll
      for (int it = 0; it < 4; ++it)
        if (q0 + tid + it * TDMA_SWEEP_THREADS < ncell) xs[sl[it]] = rs[it];
    }
    __syncthreads();
    if (tid < nb) {
      // B: the reference's recurrence, one thread per line
      const int u = tid, v = d - u;
      if (u >= ulo && u < uhi) {
        const double *const cgd = cg + (d & 1) * 2 * nu;
        const int var = wall_state(u, nu, G.pu) | (wall_state(v, nv, G.pv) << 2);
        const double *const Tv = T + (size_t)2 * var * len;
        double cprev = cgd[2 * u];
        for (int p0 = 0; p0 < len; p0 += TDMA_CHUNK) {
          double rs[TDMA_CHUNK], apr[TDMA_CHUNK];
#pragma unroll
          for (int q = 0; q < TDMA_CHUNK; ++q) {
            const int p = min(p0 + q, len - 1);
            rs[q] = xs[p * lwp + u];
            apr[q] = Tv[2 * p + 1];
          }
#pragma unroll
          for (int q = 0; q < TDMA_CHUNK; ++q) {
            const int p = p0 + q;
            if (p >= len) break;
            const double alo = (p == 0 && !pax) ? 0.0 : cax;
            cprev = (rs[q] - alo * cprev) * apr[q];
            xs[p * lwp + u] = cprev;
          }
        }
        double next = cgd[2 * u + 1];
        const int base =
            PASS == 0 ? v * nj * H + (u >> 1) : PASS == 1 ? (v * nj + u) * H : v * H + (u >> 1);
        const int step = PASS == 0 ? H : PASS == 1 ? 0 : nj * H;
        for (int p1 = len - 1; p1 >= 0; p1 -= TDMA_CHUNK) {
          double cv[TDMA_CHUNK], aq[TDMA_CHUNK];
#pragma unroll
          for (int q = 0; q < TDMA_CHUNK; ++q) {
            const int p = max(p1 - q, 0);
            cv[q] = xs[p * lwp + u];
            aq[q] = Tv[2 * p];
          }
#pragma unroll
          for (int q = 0; q < TDMA_CHUNK; ++q) {
            const int p = p1 - q;
            if (p < 0) break;
            next = cv[q] - aq[q] * next;
            const int e = PASS == 1 ? base + (p >> 1) : base + p * step;
            (((u + v + p) & 1) ? out1 : out0)[e] = next;
            prev[u * len + p] = next;
          }
        }
      }
    } else if (d + 1 < nd) {
      a0(d + 1, tid - nb, na);
    }
    __syncthreads();  // prev and the next diagonal's A0 before A1 reads them
  }
}

__global__ void __launch_bounds__(TDMA_SWEEP_THREADS)
    k_tdma_sweep(const Tile *__restrict__ tiles, const LevelDev *__restrict__ levs,
                 const int *__restrict__ active, int src, int zero) {
  extern __shared__ double line_lds[];
  const Tile t = tiles[blockIdx.x];
  const LevelDev &L = levs[t.lev];
  if (!active[L.blk]) return;
  const int tid = (int)threadIdx.x + TILE_X * (int)threadIdx.y;
  tdma_pass_in_group<0>(L, src, zero, line_lds, tid);
  tdma_pass_in_group<1>(L, src, zero, line_lds, tid);
  tdma_pass_in_group<2>(L, src, zero, line_lds, tid);
}

// the same sweep on a constant-coefficient level, passes pipelined over the
// diagonals (a kernel of its own: its registers do not cost the other one)
__global__ void __launch_bounds__(TDMA_SWEEP_THREADS)
    k_tdma_sweep_pipe(const Tile *__restrict__ tiles, const LevelDev *__restrict__ levs,
                      const int *__restrict__ active, int src, int zero) {
  extern __shared__ double line_lds[];
  const Tile t = tiles[blockIdx.x];
  const LevelDev &L = levs[t.lev];
  if (!active[L.blk] || L.lsm != nullptr || L.tdma == nullptr) return;
  const int tid = (int)threadIdx.x + TILE_X * (int)threadIdx.y;
  tdma_pass_pipelined<0>(L, src, zero, line_lds, tid);
  tdma_pass_pipelined<1>(L, src, zero, line_lds, tid);
  tdma_pass_pipelined<2>(L, src, zero, line_lds, tid);
}

}  // namespace

// ---- launchers -----------------------------------------------------------
template <class T>
void launch_rb_sweep(hipStream_t s, const Tile *tiles, int ntiles, int depth, const LevelDev *levs,
                     const int *active, int colour, int src, int zero, OneDesc one) {
  if (ntiles <= 0) return;
  auto k = one.lev >= 0 ? (depth == 4   ? k_rb_sweep<T, 4, true>
                           : depth == 2 ? k_rb_sweep<T, 2, true>
                                        : k_rb_sweep<T, 1, true>)
                        : (depth == 4   ? k_rb_sweep<T, 4, false>
                           : depth == 2 ? k_rb_sweep<T, 2, false>
                                        : k_rb_sweep<T, 1, false>);
  hipLaunchKernelGGL(k, dim3(ntiles), tile_block(), 0, s, tiles, levs, active, colour, src, zero,
                     one.lev, one.blk, ntiles);
}

void launch_lsm_dens_restrict(hipStream_t s, const double *fine, int fni, int fnj, int fnk,
                              double *coarse) {
  const size_t n = (size_t)(fni / 2) * (fnj / 2) * (fnk / 2);
  if (!n) return;
  hipLaunchKernelGGL(k_lsm_dens_restrict, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, fine,
                     fni, fnj, fnk, coarse);
}

void launch_lsm_coef(hipStream_t s, const LevelDev *lev_dev, const double *dens, int ni, int nj,
                     int nk, double dx, double dy, double dz, int walls) {
  const size_t n = (size_t)ni * nj * nk;
  hipLaunchKernelGGL(k_lsm_coef, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, lev_dev, dens,
                     ni, nj, nk, dx, dy, dz, walls);
}

void launch_lsm_sweep(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                      const int *active, int colour, int src, int zero) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_lsm_sweep, dim3(ntiles), tile_block(), 0, s, tiles, levs, active, colour,
                     src, zero);
}

void launch_gs_tiles(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                     const int *active, int src, int zero, int tile) {
  if (ntiles <= 0) return;
  if (tile == GS_BIG_TILE)
    hipLaunchKernelGGL(k_gs_tile_big, dim3(ntiles), dim3(GS_BIG_TILE * GS_BIG_TILE), GS_BIG_LDS,
                       s, tiles, levs, active, src, zero);
  else
    hipLaunchKernelGGL(k_gs_tile, dim3(ntiles), dim3(GS_TILE * GS_TILE), 0, s, tiles, levs,
                       active, src, zero);
}

hipError_t gs_allow_lds() {
  return hipFuncSetAttribute((const void *)k_gs_tile_big,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)GS_BIG_LDS);
}

void launch_tdma_diag(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                      const int *active, int pass, int d, int src, int zero) {
  if (ntiles <= 0) return;
  auto k = pass == 0 ? k_tdma_diag<0> : pass == 1 ? k_tdma_diag<1> : k_tdma_diag<2>;
  hipLaunchKernelGGL(k, dim3(ntiles), tile_block(), 0, s, tiles, levs, active, d, src, zero);
}

void launch_tdma_sweep(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                       const int *active, int src, int zero, size_t lds_bytes, bool pipelined) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(pipelined ? k_tdma_sweep_pipe : k_tdma_sweep, dim3(ntiles),
                     dim3(TILE_X, TDMA_SWEEP_THREADS / TILE_X), lds_bytes, s, tiles, levs, active,
                     src, zero);
}

hipError_t tdma_sweep_allow_lds(size_t lds_bytes) {
  hipError_t e = hipFuncSetAttribute((const void *)k_tdma_sweep,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void *)k_tdma_sweep_pipe,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
}

void launch_tdma_tables(hipStream_t s, const LevelDev *levs, int ndesc) {
  if (ndesc <= 0) return;
  hipLaunchKernelGGL(k_tdma_tables, dim3((unsigned)((ndesc * 48 + 255) / 256)), dim3(256), 0, s,
                     levs, ndesc);
}

size_t tdma_sweep_lds_bytes(const LevelDev &L, bool pipelined) {
  // per pass (nu lines of len cells): rs/c, a and 1/(ap - alo*a) (or, for
  // the pipelined passes, t1, x3*X3 and rs/c) for (nu+1)*len positions, the
  // previous diagonal's solution, nu*len; pipelined passes also the tables
  // (32*len), the first lines' ghosts (2*len) and the line-end ghosts (4*nu)
  size_t most = 0;
  const int geo[3][2] = {{L.ni, L.nj}, {L.nj, L.ni}, {L.ni, L.nk}};  // (nu, len) per pass
  for (const auto &g : geo) {
    const size_t nu = (size_t)g[0], len = (size_t)g[1];
    size_t need = 3 * (nu + 1) * len + nu * len;
    if (pipelined) need += 32 * len + 2 * len + 4 * nu;
    most = std::max(most, need);
  }
  return most * sizeof(double);
}

void launch_lsm_residual(hipStream_t s, const Tile *tiles, int ntiles, const FineDev *fine,
                         const LevelDev *levs, const int *active, RedPartial *partials) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_lsm_residual, dim3(ntiles), tile_block(), 0, s, tiles, fine, levs, active,
                     partials);
}

void launch_lsm_restrict(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                         const int *active, int src_f) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_lsm_restrict, dim3(ntiles), tile_block(), 0, s, tiles, levs, active, src_f);
}

template <class T>
void launch_rb_fused(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                     const int *active, int src, int zero) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_rb_fused<T>, dim3(ntiles), tile_block(), 0, s, tiles, levs, active, src,
                     zero);
}

template <class Tf, class Tc>
void launch_restrict(hipStream_t s, const Tile *tiles, int ntiles, int depth, const LevelDev *levs,
                     const int *active, int src_f, OneDesc one, int zbuf) {
  if (ntiles <= 0) return;
  if (depth == 0) {  // natural tiles of the coarse level, blocks with neighbours
    hipLaunchKernelGGL((k_restrict_gen<Tf, Tc>), dim3(ntiles), tile_block(), 0, s, tiles, levs,
                       active, src_f);
    return;
  }
  auto k = one.lev >= 0 ? (depth == 4   ? k_restrict<Tf, Tc, 4, true>
                           : depth == 2 ? k_restrict<Tf, Tc, 2, true>
                                        : k_restrict<Tf, Tc, 1, true>)
                        : (depth == 4   ? k_restrict<Tf, Tc, 4, false>
                           : depth == 2 ? k_restrict<Tf, Tc, 2, false>
                                        : k_restrict<Tf, Tc, 1, false>);
  hipLaunchKernelGGL(k, dim3(ntiles), tile_block(), 0, s, tiles, levs, active, src_f, one.lev,
                     one.lev2, one.blk, ntiles, zbuf);
}

template <class Tf, class Tc>
void launch_prolong(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                    const int *active, int src_f, int src_c, int gen, OneDesc one) {
  if (ntiles <= 0) return;
  if (gen)
    hipLaunchKernelGGL((k_prolong<Tf, Tc, true, false>), dim3(ntiles), tile_block(), 0, s, tiles,
                       levs, active, src_f, src_c, -1, -1, -1, ntiles);
  else if (one.lev >= 0)
    hipLaunchKernelGGL((k_prolong<Tf, Tc, false, true>), dim3(ntiles), tile_block(), 0, s, tiles,
                       levs, active, src_f, src_c, one.lev, one.lev2, one.blk, ntiles);
  else
    hipLaunchKernelGGL((k_prolong<Tf, Tc, false, false>), dim3(ntiles), tile_block(), 0, s, tiles,
                       levs, active, src_f, src_c, -1, -1, -1, ntiles);
}

template <class T>
void launch_fill_zero(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                      const int *active, int buf) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_fill_zero<T>, dim3(ntiles), tile_block(), 0, s, tiles, levs, active, buf);
}

template <class T>
void launch_residual(hipStream_t s, const Tile *tiles, int ntiles, int depth, const FineDev *fine,
                     const LevelDev *levs, const int *active, RedPartial *partials, int use_p,
                     int zbuf) {
  if (ntiles <= 0) return;
  if (!use_p || depth <= 0) {
    hipLaunchKernelGGL(k_residual_gen<T>, dim3(ntiles), tile_block(), 0, s, tiles, fine, levs,
                       active, partials, use_p);
    return;
  }
  auto k = depth == 8   ? k_residual<T, 8>
           : depth == 4 ? k_residual<T, 4>
           : depth == 2 ? k_residual<T, 2>
                        : k_residual<T, 1>;
  hipLaunchKernelGGL(k, dim3(ntiles), tile_block(), 0, s, tiles, fine, levs, active, partials,
                     zbuf);
}

template <class T>
void launch_update(hipStream_t s, const Tile *tiles, int ntiles, const FineDev *fine,
                   const LevelDev *levs, const int *active, int buf) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_update<T>, dim3(ntiles), tile_block(), 0, s, tiles, fine, levs, active,
                     buf);
}

void launch_finalize_decide(hipStream_t s, const int *tile_begin, int nblocks,
                            const RedPartial *partials, RedPartial *out, const DecideBlock *blocks,
                            DecideState *state, DecideCtl *ctl, int *active, char *ring,
                            const DecideParams *prm, unsigned *arrived) {
  hipLaunchKernelGGL(k_finalize_decide, dim3(nblocks), dim3(FIN_THREADS), 0, s, tile_begin,
                     partials, out, blocks, state, ctl, active, ring, prm, arrived);
}

void launch_finalize(hipStream_t s, const int *tile_begin, int nblocks, const RedPartial *partials,
                     RedPartial *out) {
  hipLaunchKernelGGL(k_finalize, dim3(nblocks), dim3(FIN_THREADS), 0, s, tile_begin, partials, out);
}

void launch_coarse_exact(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                         const int *active, const double *const *eig_q,
                         const double *const *eig_l, double *const *scratch, int buf) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_coarse_exact, dim3(ntiles), dim3(256), 0, s, tiles, levs, active, eig_q,
                     eig_l, scratch, buf);
}

void launch_vtail(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                  const int *active, const double *const *eig_q, const double *const *eig_l,
                  int nu1, int nu2, int exact, int mixed, size_t lds_bytes) {
  if (ntiles <= 0) return;
  if (mixed)
    hipLaunchKernelGGL(k_vtail<true>, dim3(ntiles), dim3(TAIL_THREADS), lds_bytes, s, tiles, levs,
                       active, eig_q, eig_l, nu1, nu2, exact);
  else
    hipLaunchKernelGGL(k_vtail<false>, dim3(ntiles), dim3(TAIL_THREADS), lds_bytes, s, tiles, levs,
                       active, eig_q, eig_l, nu1, nu2, exact);
}

hipError_t vtail_allow_lds(size_t lds_bytes) {
  hipError_t e = hipFuncSetAttribute((const void *)k_vtail<true>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void *)k_vtail<false>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
}

size_t vtail_lds_bytes(const LevelDev *levs_host, int top, bool mixed, bool exact) {
  size_t off = 0;
  int last = top, nl = 0;
  for (int d = top; d >= 0 && nl < TAIL_MAX_LEVELS; d = levs_host[d].coarser, ++nl) {
    const LevelDev &L = levs_host[d];
    const bool is64 = !mixed || L.coarser < 0;
    const size_t half = (size_t)L.H * L.nj * L.nk * (is64 ? 8 : 4);
    off += 6 * ((half + 15) & ~(size_t)15);
    last = d;
  }
  const LevelDev &C = levs_host[last];
  if (exact && C.finer >= 0) {
    const int na[3] = {C.ni, C.nj, C.nk};
    for (int a = 0; a < 3; ++a) off += (size_t)na[a] * na[a] * 8 + (((size_t)na[a] * 8 + 15) & ~(size_t)15);
    off += 2 * (size_t)na[0] * na[1] * na[2] * 8;
  }
  return off;
}

void launch_calvel(hipStream_t s, const FineDev *fd, int axis, const double *us, double *u,
                   double dt_alfa, double h, size_t faces) {
  const unsigned grid = (unsigned)std::min<size_t>((faces + 255) / 256, 4096);
  hipLaunchKernelGGL(k_calvel, dim3(grid), dim3(256), 0, s, fd, axis, us, u, dt_alfa, h);
}

void launch_calvel_nb(hipStream_t s, const FineDev *fine, int b, int axis, const double *us,
                      double *u, double dt_alfa, double h, size_t faces) {
  const unsigned grid = (unsigned)std::min<size_t>((faces + 255) / 256, 4096);
  hipLaunchKernelGGL(k_calvel_nb, dim3(grid), dim3(256), 0, s, fine, b, axis, us, u, dt_alfa, h);
}

void launch_face_copy(hipStream_t s, int axis, const double *src, int n_src, double *dst,
                      int n_dst, int nu, int nv) {
  const size_t n = (size_t)nu * nv;
  hipLaunchKernelGGL(k_face_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, axis, src,
                     n_src, dst, n_dst, nu, nv);
}

void launch_unpack_fine_padded(hipStream_t s, const FineDev *fine, int b, double *dst, int ni,
                               int nj, int nk) {
  const size_t n = (size_t)(ni + 2) * (nj + 2) * (nk + 2);
  hipLaunchKernelGGL(k_unpack_fine_padded, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     fine, b, dst);
}

int launch_calmas(hipStream_t s, const FineDev *fd, const double *u, const double *v,
                  const double *w, double *su, int to_fine, double dt, double dx, double dy,
                  double dz, RedPartial *partials, size_t cells) {
  const int grid = calmas_groups(cells);
  hipLaunchKernelGGL(k_calmas, dim3(grid), dim3(256), 0, s, fd, u, v, w, su, to_fine, dt, dx, dy,
                     dz, partials);
  return grid;
}

void launch_partials_max(hipStream_t s, const RedPartial *parts, int n, double *out) {
  hipLaunchKernelGGL(k_partials_max, dim3(1), dim3(256), 0, s, parts, n, out);
}

int launch_calmas4(hipStream_t s, const FineDev *fd, const double *u, const double *v,
                   const double *w, double *su, int to_fine, double dt, double dx, double dy,
                   double dz, RedPartial *partials, size_t cells, const Outer4 &outer) {
  const int grid = calmas_groups(cells);
  hipLaunchKernelGGL(k_calmas4, dim3(grid), dim3(256), 0, s, fd, u, v, w, su, to_fine, dt, dx, dy,
                     dz, partials, outer);
  return grid;
}

void launch_pack_fine(hipStream_t s, const FineDev *fd, int which, const double *src, int ni,
                      int nj, int nk) {
  const size_t n = (size_t)ni * nj * nk;
  hipLaunchKernelGGL(k_pack_fine, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, fd, which,
                     src, ni, nj, nk);
}

void launch_unpack_fine(hipStream_t s, const FineDev *fd, int which, double *dst, int ni, int nj,
                        int nk) {
  const size_t n = (size_t)ni * nj * nk;
  hipLaunchKernelGGL(k_unpack_fine, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, fd,
                     which, dst, ni, nj, nk);
}

template <class T>
void launch_pack_level(hipStream_t s, const LevelDev *ld, int which, int buf, const double *src,
                       int ni, int nj, int nk) {
  const size_t n = (size_t)ni * nj * nk;
  hipLaunchKernelGGL(k_pack_level<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ld,
                     which, buf, src, ni, nj, nk);
}

template <class T>
void launch_unpack_level(hipStream_t s, const LevelDev *levs, int d, int which, int buf, double *dst,
                         int ni, int nj, int nk) {
  const size_t n = (size_t)(ni + 2) * (nj + 2) * (nk + 2);
  hipLaunchKernelGGL(k_unpack_level<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, levs, d,
                     which, buf, dst, ni, nj, nk);
}

void launch_remove_mean(hipStream_t s, const FineDev *fd, int ni, int nj, int nk,
                        double * /*scratch*/) {
  hipLaunchKernelGGL(k_remove_mean, dim3(1), dim3(256), 0, s, fd, ni, nj, nk);
}

// explicit instantiations
template void launch_rb_fused<double>(hipStream_t, const Tile *, int, const LevelDev *,
                                      const int *, int, int);
template void launch_rb_fused<float>(hipStream_t, const Tile *, int, const LevelDev *, const int *,
                                     int, int);
template void launch_rb_sweep<double>(hipStream_t, const Tile *, int, int, const LevelDev *,
                                      const int *, int, int, int, OneDesc);
template void launch_rb_sweep<float>(hipStream_t, const Tile *, int, int, const LevelDev *,
                                     const int *, int, int, int, OneDesc);
template void launch_restrict<double, double>(hipStream_t, const Tile *, int, int, const LevelDev *,
                                              const int *, int, OneDesc, int);
template void launch_restrict<float, float>(hipStream_t, const Tile *, int, int, const LevelDev *,
                                            const int *, int, OneDesc, int);
template void launch_restrict<float, double>(hipStream_t, const Tile *, int, int, const LevelDev *,
                                             const int *, int, OneDesc, int);
template void launch_prolong<double, double>(hipStream_t, const Tile *, int, const LevelDev *,
                                             const int *, int, int, int, OneDesc);
template void launch_prolong<float, float>(hipStream_t, const Tile *, int, const LevelDev *,
                                           const int *, int, int, int, OneDesc);
template void launch_prolong<float, double>(hipStream_t, const Tile *, int, const LevelDev *,
                                            const int *, int, int, int, OneDesc);
template void launch_fill_zero<double>(hipStream_t, const Tile *, int, const LevelDev *,
                                       const int *, int);
template void launch_fill_zero<float>(hipStream_t, const Tile *, int, const LevelDev *,
                                      const int *, int);
template void launch_residual<double>(hipStream_t, const Tile *, int, int, const FineDev *,
                                      const LevelDev *, const int *, RedPartial *, int, int);
template void launch_residual<float>(hipStream_t, const Tile *, int, int, const FineDev *,
                                     const LevelDev *, const int *, RedPartial *, int, int);
template void launch_update<double>(hipStream_t, const Tile *, int, const FineDev *,
                                    const LevelDev *, const int *, int);
template void launch_update<float>(hipStream_t, const Tile *, int, const FineDev *,
                                   const LevelDev *, const int *, int);
template void launch_pack_level<double>(hipStream_t, const LevelDev *, int, int, const double *,
                                        int, int, int);
template void launch_pack_level<float>(hipStream_t, const LevelDev *, int, int, const double *,
                                       int, int, int);
template void launch_unpack_level<double>(hipStream_t, const LevelDev *, int, int, int, double *, int,
                                          int, int);
template void launch_unpack_level<float>(hipStream_t, const LevelDev *, int, int, int, double *, int,
                                         int, int);

}  // namespace h3dmg
