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

// coeff, mgsolver.for:185-210: 1/h^2 towards every face, 0 towards a
// non-periodic physical face, ap = -1*(sum) summed aw..at left to right.
template <class T>
__device__ __forceinline__ Coef<T> cell_coef(const LevelDev &L, int i, int j, int k) {
  const T cx = (T)L.cx, cy = (T)L.cy, cz = (T)L.cz;
  const bool px = L.per & 1, py = L.per & 2, pz = L.per & 4;
  Coef<T> c;
  c.aw = (!px && i == 0) ? T(0) : cx;
  c.ae = (!px && i == L.ni - 1) ? T(0) : cx;
  c.as = (!py && j == 0) ? T(0) : cy;
  c.an = (!py && j == L.nj - 1) ? T(0) : cy;
  c.ab = (!pz && k == 0) ? T(0) : cz;
  c.at = (!pz && k == L.nk - 1) ? T(0) : cz;
  c.ap = T(-1) * (c.aw + c.ae + c.as + c.an + c.ab + c.at);
  return c;
}

__device__ __forceinline__ Coef<double> fine_coef(const FineDev &F, int i, int j, int k) {
  const bool px = F.per & 1, py = F.per & 2, pz = F.per & 4;
  Coef<double> c;
  c.aw = (!px && i == 0) ? 0.0 : F.cx;
  c.ae = (!px && i == F.ni - 1) ? 0.0 : F.cx;
  c.as = (!py && j == 0) ? 0.0 : F.cy;
  c.an = (!py && j == F.nj - 1) ? 0.0 : F.cy;
  c.ab = (!pz && k == 0) ? 0.0 : F.cz;
  c.at = (!pz && k == F.nk - 1) ? 0.0 : F.cz;
  c.ap = -1.0 * (c.aw + c.ae + c.as + c.an + c.ab + c.at);
  return c;
}

// value of any cell at padded position (x,y,z) in [-1, n] of a level,
// buffer `buf`, ghosts resolved to their image
template <class T>
__device__ __forceinline__ T level_value(const LevelDev &L, int buf, int x, int y, int z) {
  x = image(x, L.ni, L.per & 1);
  y = image(y, L.nj, L.per & 2);
  z = image(z, L.nk, L.per & 4);
  const T *a = (const T *)L.phi[buf][(x + y + z) & 1];
  return a[cidx(L.H, L.nj, x, y, z)];
}

__device__ __forceinline__ double fine_value(const FineDev &F, int x, int y, int z) {
  x = image(x, F.ni, F.per & 1);
  y = image(y, F.nj, F.per & 2);
  z = image(z, F.nk, F.per & 4);
  return F.p[(x + y + z) & 1][cidx(F.H, F.nj, x, y, z)];
}

// ---------------------------------------------------------------------------
// One colour half-sweep of the red-black smoother (mgsolver.for:553-605).
// colour 0 = red, 1 = black.  Sweep reads buffer `src`, writes `1-src`:
//   red:   interior neighbours (black) from black[src]
//   black: interior neighbours (red)   from red[1-src] (just written)
//   ghost neighbours from buffer src (the pre-sweep state the reference's
//   ghosts hold until exchangepp at the end of the sweep).
// zero != 0: the level holds zeros in buffer src (first sweep after mgrestr
// zeroed phi, mgsolver.for:655-662, or pp = 0 at the finest level).
template <class T>
__global__ void __launch_bounds__(NTHREADS)
    k_rb_pass(const Tile *__restrict__ tiles, const LevelDev *__restrict__ levs,
              const int *__restrict__ active, int colour, int src, int zero) {
  const Tile t = tiles[blockIdx.x];
  const LevelDev &L = levs[t.lev];
  if (!active[L.blk]) return;
  const int m = t.x0 + (int)threadIdx.x;
  const int j = t.y0 + (int)threadIdx.y;
  const int k = t.z0;
  if (m >= L.H || j >= L.nj || k >= L.nk) return;
  const int s = (colour + j + k) & 1;
  const int i = 2 * m + s;
  if (i >= L.ni) return;

  const T *nb_int = colour == 0 ? (const T *)L.phi[src][1] : (const T *)L.phi[1 - src][0];
  const bool zero_int = zero && colour == 0;
  const bool zero_ghost = zero != 0;
  auto nb = [&](int x, int y, int z) -> T {
    const bool ghost = (unsigned)x >= (unsigned)L.ni || (unsigned)y >= (unsigned)L.nj ||
                       (unsigned)z >= (unsigned)L.nk;
    if (!ghost) return zero_int ? T(0) : nb_int[cidx(L.H, L.nj, x, y, z)];
    return zero_ghost ? T(0) : level_value<T>(L, src, x, y, z);
  };
  const Coef<T> c = cell_coef<T>(L, i, j, k);
  const T W = nb(i - 1, j, k), E = nb(i + 1, j, k);
  const T S = nb(i, j - 1, k), N = nb(i, j + 1, k);
  const T B = nb(i, j, k - 1), Tp = nb(i, j, k + 1);
  const size_t q = cidx(L.H, L.nj, i, j, k);
  const T r = ((const T *)L.rhs[colour])[q];
  const T sum = c.aw * W + c.ae * E + c.as * S + c.an * N + c.ab * B + c.at * Tp;
  ((T *)L.phi[1 - src][colour])[q] = (r - sum) / c.ap;
}

// ---------------------------------------------------------------------------
// mgrestr (mgsolver.for:693-756): fine residual
//   resf = rhs - (aw*W + ae*E + as*S + an*N + ab*B + at*T + ap*P)
// and coarse rhs = 0.125 * (sum of the 8 children, i fastest, then j, k).
// One thread per coarse cell; the tile's `lev` is the coarse descriptor.
// The coarse phi is not touched: the next relax of the coarse level runs with
// zero = 1, which is the zeroing of mgsolver.for:655-662.
template <class Tf, class Tc>
__global__ void __launch_bounds__(NTHREADS)
    k_restrict(const Tile *__restrict__ tiles, const LevelDev *__restrict__ levs,
               const int *__restrict__ active, int src_f) {
  const Tile t = tiles[blockIdx.x];
  const LevelDev &C = levs[t.lev];
  if (!active[C.blk]) return;
  const LevelDev &F = levs[C.finer];
  const int ic = t.x0 + (int)threadIdx.x;
  const int jc = t.y0 + (int)threadIdx.y;
  const int kc = t.z0;
  if (ic >= C.ni || jc >= C.nj || kc >= C.nk) return;

  Tf res[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int i = 2 * ic + (q & 1), j = 2 * jc + ((q >> 1) & 1), k = 2 * kc + (q >> 2);
    const Coef<Tf> c = cell_coef<Tf>(F, i, j, k);
    const Tf W = level_value<Tf>(F, src_f, i - 1, j, k);
    const Tf E = level_value<Tf>(F, src_f, i + 1, j, k);
    const Tf S = level_value<Tf>(F, src_f, i, j - 1, k);
    const Tf N = level_value<Tf>(F, src_f, i, j + 1, k);
    const Tf B = level_value<Tf>(F, src_f, i, j, k - 1);
    const Tf Tt = level_value<Tf>(F, src_f, i, j, k + 1);
    const size_t fq = cidx(F.H, F.nj, i, j, k);
    const int fc = (i + j + k) & 1;
    const Tf P = ((const Tf *)F.phi[src_f][fc])[fq];
    const Tf r = ((const Tf *)F.rhs[fc])[fq];
    res[q] = r - (c.aw * W + c.ae * E + c.as * S + c.an * N + c.ab * B + c.at * Tt + c.ap * P);
  }
  Tc acc = (Tc)res[0];
#pragma unroll
  for (int q = 1; q < 8; ++q) acc = acc + (Tc)res[q];
  ((Tc *)C.rhs[(ic + jc + kc) & 1])[cidx(C.H, C.nj, ic, jc, kc)] = Tc(0.125) * acc;
}

// ---------------------------------------------------------------------------
// mgcorr (mgsolver.for:828-896): fine += trilinear interpolation of coarse.
// Fine padded index ip = 2*icp + d with icp the coarse padded index; the
// weight of coarse corner (oi,oj,ok) is prod (3 if o==d else 1); terms are
// summed in the order c, c+j, c+k, c+k+j, c+i, c+i+j, c+i+k, c+i+j+k and the
// sum divided by 64.  Ghost corners are virtual (images of coarse interior).
// One thread per fine interior cell; the fine level is updated in place.
template <class Tf, class Tc>
__global__ void __launch_bounds__(NTHREADS)
    k_prolong(const Tile *__restrict__ tiles, const LevelDev *__restrict__ levs,
              const int *__restrict__ active, int src_f, int src_c) {
  const Tile t = tiles[blockIdx.x];
  const LevelDev &F = levs[t.lev];
  if (!active[F.blk]) return;
  const LevelDev &C = levs[F.coarser];
  const int i = t.x0 + (int)threadIdx.x;
  const int j = t.y0 + (int)threadIdx.y;
  const int k = t.z0;
  if (i >= F.ni || j >= F.nj || k >= F.nk) return;
  // padded fine (i+1) = 2*icp + di; coarse interior of padded icp is icp-1
  const int di = (i + 1) & 1, dj = (j + 1) & 1, dk = (k + 1) & 1;
  const int ci = ((i + 1) >> 1) - 1, cj = ((j + 1) >> 1) - 1, ck = ((k + 1) >> 1) - 1;
  const int ord[8][3] = {{0, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1},
                         {1, 0, 0}, {1, 1, 0}, {1, 0, 1}, {1, 1, 1}};
  Tc s = Tc(0);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int oi = ord[q][0], oj = ord[q][1], ok = ord[q][2];
    const Tc w = (oi == di ? Tc(3) : Tc(1)) * (oj == dj ? Tc(3) : Tc(1)) * (ok == dk ? Tc(3) : Tc(1));
    const Tc v = level_value<Tc>(C, src_c, ci + oi, cj + oj, ck + ok);
    s = (q == 0) ? w * v : s + w * v;
  }
  const size_t fq = cidx(F.H, F.nj, i, j, k);
  Tf *fa = (Tf *)F.phi[src_f][(i + j + k) & 1];
  fa[fq] = fa[fq] + (Tf)(s / Tc(64));
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
// into the level-1 rhs in the level's precision, with per-workgroup partial
// sum of r^2 and max |r| for the REL_L2 and REF_RMAX criteria.
template <class T>
__global__ void __launch_bounds__(NTHREADS)
    k_residual(const Tile *__restrict__ tiles, const FineDev *__restrict__ fine,
               const LevelDev *__restrict__ levs, const int *__restrict__ active,
               RedPartial *__restrict__ partials, int use_p) {
  const Tile t = tiles[blockIdx.x];
  const FineDev &F = fine[t.lev];
  const LevelDev &L = levs[t.lev];
  const int i = t.x0 + (int)threadIdx.x;
  const int j = t.y0 + (int)threadIdx.y;
  const int k = t.z0;
  double ss = 0.0, am = 0.0;
  if (active[F.blk] && i < F.ni && j < F.nj && k < F.nk) {
    const int col = (i + j + k) & 1;
    const size_t q = cidx(F.H, F.nj, i, j, k);
    double r = F.f[col][q];
    if (use_p) {
      const Coef<double> c = fine_coef(F, i, j, k);
      const double W = fine_value(F, i - 1, j, k), E = fine_value(F, i + 1, j, k);
      const double S = fine_value(F, i, j - 1, k), N = fine_value(F, i, j + 1, k);
      const double B = fine_value(F, i, j, k - 1), Tt = fine_value(F, i, j, k + 1);
      const double P = F.p[col][q];
      r = r - (c.aw * W + c.ae * E + c.as * S + c.an * N + c.ab * B + c.at * Tt + c.ap * P);
    }
    ((T *)L.rhs[col])[q] = (T)r;
    ss = r * r;
    am = fabs(r);
  }
  block_reduce(ss, am);
  if (threadIdx.x == 0 && threadIdx.y == 0) partials[blockIdx.x] = RedPartial{ss, am};
}

// p = p + (pp - ppref) with ppref = 0 (newsolv_mg.for:39-49); pp is the
// level-1 correction in buffer `buf`.
template <class T>
__global__ void __launch_bounds__(NTHREADS)
    k_update(const Tile *__restrict__ tiles, const FineDev *__restrict__ fine,
             const LevelDev *__restrict__ levs, const int *__restrict__ active, int buf) {
  const Tile t = tiles[blockIdx.x];
  const FineDev &F = fine[t.lev];
  const LevelDev &L = levs[t.lev];
  if (!active[F.blk]) return;
  const int i = t.x0 + (int)threadIdx.x;
  const int j = t.y0 + (int)threadIdx.y;
  const int k = t.z0;
  if (i >= F.ni || j >= F.nj || k >= F.nk) return;
  const int col = (i + j + k) & 1;
  const size_t q = cidx(F.H, F.nj, i, j, k);
  const double pp = (double)((const T *)L.phi[buf][col])[q];
  F.p[col][q] = F.p[col][q] + (pp - 0.0);
}

// per-block reduction of the partials of k_residual, fixed order
__global__ void __launch_bounds__(256)
    k_finalize(const int *__restrict__ tile_begin, const RedPartial *__restrict__ partials,
               RedPartial *__restrict__ out) {
  const int b = blockIdx.x;
  const int t0 = tile_begin[b], t1 = tile_begin[b + 1];
  double ss = 0.0, am = 0.0;
  for (int q = t0 + (int)threadIdx.x; q < t1; q += 256) {
    ss += partials[q].ss;
    am = fmax(am, partials[q].amax);
  }
  __shared__ double s_ss[256], s_am[256];
  s_ss[threadIdx.x] = ss;
  s_am[threadIdx.x] = am;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s_ss[threadIdx.x] += s_ss[threadIdx.x + w];
      s_am[threadIdx.x] = fmax(s_am[threadIdx.x], s_am[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[b] = RedPartial{s_ss[0], s_am[0]};
}

// ---------------------------------------------------------------------------
// Exact FP64 solve of the coarsest level of one block (extension, no
// reference counterpart): A = Dx (+) Dy (+) Dz with Dx the 1-D Neumann or
// periodic second difference; with Q_a the orthonormal eigenvectors of D_a
// (columns) and l_a its eigenvalues, u = Q (Q^T f ./ (lx+ly+lz)) with the
// zero (constant) mode dropped -- the null-space projection.  One workgroup
// per block; scratch holds two n^3 arrays.
__global__ void __launch_bounds__(256)
    k_coarse_exact(const Tile *__restrict__ tiles, const LevelDev *__restrict__ levs,
                   const int *__restrict__ active, const double *const *__restrict__ eig_q,
                   const double *const *__restrict__ eig_l, double *const *__restrict__ scratch,
                   int buf) {
  const Tile t = tiles[blockIdx.x];  // lev = coarsest descriptor, x0 = block
  const int b = t.x0;
  if (!active[b]) return;
  const LevelDev &L = levs[t.lev];
  const int n0 = L.ni, n1 = L.nj, n2 = L.nk;
  const size_t n = (size_t)n0 * n1 * n2;
  const double *Q0 = eig_q[3 * b + 0], *Q1 = eig_q[3 * b + 1], *Q2 = eig_q[3 * b + 2];
  const double *l0 = eig_l[3 * b + 0], *l1 = eig_l[3 * b + 1], *l2 = eig_l[3 * b + 2];
  double *A = scratch[2 * b], *Bs = scratch[2 * b + 1];
  const int tid = threadIdx.x;
  // gather rhs (colour split) into natural layout A[i + n0*(j + n1*k)]
  for (size_t q = tid; q < n; q += 256) {
    const int i = (int)(q % n0), j = (int)((q / n0) % n1), k = (int)(q / ((size_t)n0 * n1));
    A[q] = ((const double *)L.rhs[(i + j + k) & 1])[cidx(L.H, L.nj, i, j, k)];
  }
  __syncthreads();
  // forward transforms: B = Q0^T along i, A = Q1^T along j, B = Q2^T along k
  for (size_t q = tid; q < n; q += 256) {
    const int i = (int)(q % n0);
    const size_t base = q - i;
    double s = 0.0;
    for (int x = 0; x < n0; ++x) s += Q0[(size_t)x * n0 + i] * A[base + x];
    Bs[q] = s;
  }
  __syncthreads();
  for (size_t q = tid; q < n; q += 256) {
    const int i = (int)(q % n0), j = (int)((q / n0) % n1);
    const size_t base = q - (size_t)j * n0;
    double s = 0.0;
    for (int y = 0; y < n1; ++y) s += Q1[(size_t)y * n1 + j] * Bs[base + (size_t)y * n0];
    A[q] = s;
    (void)i;
  }
  __syncthreads();
  for (size_t q = tid; q < n; q += 256) {
    const size_t plane = (size_t)n0 * n1;
    const int k = (int)(q / plane);
    const size_t base = q - (size_t)k * plane;
    double s = 0.0;
    for (int z = 0; z < n2; ++z) s += Q2[(size_t)z * n2 + k] * A[base + (size_t)z * plane];
    const int i = (int)(q % n0), j = (int)((q / n0) % n1);
    const double lam = l0[i] + l1[j] + l2[k];
    // mode (0,0,0) is the constant (null-space) mode of every axis
    Bs[q] = ((i | j | k) == 0) ? 0.0 : s / lam;
  }
  __syncthreads();
  // inverse transforms: A = Q2 along k, B = Q1 along j, phi = Q0 along i
  for (size_t q = tid; q < n; q += 256) {
    const size_t plane = (size_t)n0 * n1;
    const int k = (int)(q / plane);
    const size_t base = q - (size_t)k * plane;
    double s = 0.0;
    for (int z = 0; z < n2; ++z) s += Q2[(size_t)k * n2 + z] * Bs[base + (size_t)z * plane];
    A[q] = s;
  }
  __syncthreads();
  for (size_t q = tid; q < n; q += 256) {
    const int j = (int)((q / n0) % n1);
    const size_t base = q - (size_t)j * n0;
    double s = 0.0;
    for (int y = 0; y < n1; ++y) s += Q1[(size_t)j * n1 + y] * A[base + (size_t)y * n0];
    Bs[q] = s;
  }
  __syncthreads();
  for (size_t q = tid; q < n; q += 256) {
    const int i = (int)(q % n0), j = (int)((q / n0) % n1), k = (int)(q / ((size_t)n0 * n1));
    const size_t base = q - i;
    double s = 0.0;
    for (int x = 0; x < n0; ++x) s += Q0[(size_t)i * n0 + x] * Bs[base + x];
    ((double *)L.phi[buf][(i + j + k) & 1])[cidx(L.H, L.nj, i, j, k)] = s;
  }
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
__global__ void k_unpack_level(const LevelDev *__restrict__ ld, int which, int buf,
                               double *__restrict__ dst, int ni, int nj, int nk) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const LevelDev &L = *ld;
  if (which == 0) {
    const int pi = ni + 2, pj = nj + 2, pk = nk + 2;
    const size_t n = (size_t)pi * pj * pk;
    if (q >= n) return;
    const int x = (int)(q % pi) - 1, y = (int)((q / pi) % pj) - 1, z = (int)(q / ((size_t)pi * pj)) - 1;
    dst[q] = (double)level_value<T>(L, buf, x, y, z);
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

}  // namespace

// ---- launchers -----------------------------------------------------------
template <class T>
void launch_rb_pass(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                    const int *active, int colour, int src, int zero) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_rb_pass<T>, dim3(ntiles), tile_block(), 0, s, tiles, levs, active, colour,
                     src, zero);
}

template <class Tf, class Tc>
void launch_restrict(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                     const int *active, int src_f) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL((k_restrict<Tf, Tc>), dim3(ntiles), tile_block(), 0, s, tiles, levs, active,
                     src_f);
}

template <class Tf, class Tc>
void launch_prolong(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                    const int *active, int src_f, int src_c) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL((k_prolong<Tf, Tc>), dim3(ntiles), tile_block(), 0, s, tiles, levs, active,
                     src_f, src_c);
}

template <class T>
void launch_fill_zero(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                      const int *active, int buf) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_fill_zero<T>, dim3(ntiles), tile_block(), 0, s, tiles, levs, active, buf);
}

template <class T>
void launch_residual(hipStream_t s, const Tile *tiles, int ntiles, const FineDev *fine,
                     const LevelDev *levs, const int *active, RedPartial *partials, int use_p) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_residual<T>, dim3(ntiles), tile_block(), 0, s, tiles, fine, levs, active,
                     partials, use_p);
}

template <class T>
void launch_update(hipStream_t s, const Tile *tiles, int ntiles, const FineDev *fine,
                   const LevelDev *levs, const int *active, int buf) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_update<T>, dim3(ntiles), tile_block(), 0, s, tiles, fine, levs, active,
                     buf);
}

void launch_finalize(hipStream_t s, const int *tile_begin, int nblocks,
                     const RedPartial *partials, RedPartial *out) {
  hipLaunchKernelGGL(k_finalize, dim3(nblocks), dim3(256), 0, s, tile_begin, partials, out);
}

void launch_coarse_exact(hipStream_t s, const Tile *tiles, int ntiles, const LevelDev *levs,
                         const int *active, const double *const *eig_q,
                         const double *const *eig_l, double *const *scratch, int buf) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_coarse_exact, dim3(ntiles), dim3(256), 0, s, tiles, levs, active, eig_q,
                     eig_l, scratch, buf);
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
void launch_unpack_level(hipStream_t s, const LevelDev *ld, int which, int buf, double *dst,
                         int ni, int nj, int nk) {
  const size_t n = (size_t)(ni + 2) * (nj + 2) * (nk + 2);
  hipLaunchKernelGGL(k_unpack_level<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ld,
                     which, buf, dst, ni, nj, nk);
}

void launch_remove_mean(hipStream_t s, const FineDev *fd, int ni, int nj, int nk,
                        double * /*scratch*/) {
  hipLaunchKernelGGL(k_remove_mean, dim3(1), dim3(256), 0, s, fd, ni, nj, nk);
}

// explicit instantiations
template void launch_rb_pass<double>(hipStream_t, const Tile *, int, const LevelDev *, const int *,
                                     int, int, int);
template void launch_rb_pass<float>(hipStream_t, const Tile *, int, const LevelDev *, const int *,
                                    int, int, int);
template void launch_restrict<double, double>(hipStream_t, const Tile *, int, const LevelDev *,
                                              const int *, int);
template void launch_restrict<float, float>(hipStream_t, const Tile *, int, const LevelDev *,
                                            const int *, int);
template void launch_restrict<float, double>(hipStream_t, const Tile *, int, const LevelDev *,
                                             const int *, int);
template void launch_prolong<double, double>(hipStream_t, const Tile *, int, const LevelDev *,
                                             const int *, int, int);
template void launch_prolong<float, float>(hipStream_t, const Tile *, int, const LevelDev *,
                                           const int *, int, int);
template void launch_prolong<float, double>(hipStream_t, const Tile *, int, const LevelDev *,
                                            const int *, int, int);
template void launch_fill_zero<double>(hipStream_t, const Tile *, int, const LevelDev *,
                                       const int *, int);
template void launch_fill_zero<float>(hipStream_t, const Tile *, int, const LevelDev *,
                                      const int *, int);
template void launch_residual<double>(hipStream_t, const Tile *, int, const FineDev *,
                                      const LevelDev *, const int *, RedPartial *, int);
template void launch_residual<float>(hipStream_t, const Tile *, int, const FineDev *,
                                     const LevelDev *, const int *, RedPartial *, int);
template void launch_update<double>(hipStream_t, const Tile *, int, const FineDev *,
                                    const LevelDev *, const int *, int);
template void launch_update<float>(hipStream_t, const Tile *, int, const FineDev *,
                                   const LevelDev *, const int *, int);
template void launch_pack_level<double>(hipStream_t, const LevelDev *, int, int, const double *,
                                        int, int, int);
template void launch_pack_level<float>(hipStream_t, const LevelDev *, int, int, const double *,
                                       int, int, int);
template void launch_unpack_level<double>(hipStream_t, const LevelDev *, int, int, double *, int,
                                          int, int);
template void launch_unpack_level<float>(hipStream_t, const LevelDev *, int, int, double *, int,
                                         int, int);

}  // namespace h3dmg
