/* h3dmg_oracle.c -- CPU restatement of the Hydro3D multigrid (test oracle).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py as the checker; the product never links it.
 *
 * Every routine cites the reference lines it follows.  Arithmetic is written
 * in the reference's evaluation order and built with -ffp-contract=off, so
 * results are bitwise those of the reference compiled the same way.
 */
#include "h3dmg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(i, j, k, pi, pj) ((size_t)(i) + (size_t)(pi) * ((size_t)(j) + (size_t)(pj) * (size_t)(k)))

void or_level_dims(const or_block *b, int g, int nl[3]) {
  /* ni = (iep-isp+1)/2**(gl-1) (+2 for the padded size), mgsolver.for:398-400 */
  for (int a = 0; a < 3; ++a) nl[a] = b->n[a] >> (g - 1);
}

/* Level-g coefficients of one interior cell (0-based interior i,j,k).
 * coeff, mgsolver.for:54-60 (ndx = dx*2**(gl-1)), 185-191 (1/dxx),
 * 194-205 (zero towards a non-periodic physical face), 207-210 (ap). */
typedef struct {
  double aw, ae, as, an, ab, at, ap;
} cof7;

typedef struct {
  double cx, cy, cz; /* 1/ndx**2 ... */
} levcoef;

static levcoef level_coef(const or_block *b, int g) {
  levcoef c;
  double s = (double)(1 << (g - 1));
  double ndx = b->h[0] * s, ndy = b->h[1] * s, ndz = b->h[2] * s;
  double dxx = ndx * ndx, dyy = ndy * ndy, dzz = ndz * ndz;
  c.cx = 1.0 / dxx;
  c.cy = 1.0 / dyy;
  c.cz = 1.0 / dzz;
  return c;
}

static inline cof7 cell_cof(const or_block *b, const levcoef *c, const int nl[3], int i, int j,
                            int k) {
  cof7 r;
  r.aw = c->cx;
  r.ae = c->cx;
  r.as = c->cy;
  r.an = c->cy;
  r.ab = c->cz;
  r.at = c->cz;
  if (b->bc[0] != OR_BC_PERIODIC && i == 0) r.aw = 0.0;
  if (b->bc[1] != OR_BC_PERIODIC && i == nl[0] - 1) r.ae = 0.0;
  if (b->bc[2] != OR_BC_PERIODIC && j == 0) r.as = 0.0;
  if (b->bc[3] != OR_BC_PERIODIC && j == nl[1] - 1) r.an = 0.0;
  if (b->bc[4] != OR_BC_PERIODIC && k == 0) r.ab = 0.0;
  if (b->bc[5] != OR_BC_PERIODIC && k == nl[2] - 1) r.at = 0.0;
  r.ap = -1.0 * (r.aw + r.ae + r.as + r.an + r.ab + r.at);
  return r;
}

/* mgbound (mgsolver.for:912-997): mirror ghost = first interior, over full
 * padded faces, x then y then z, on faces that are not periodic. */
void or_bound(const or_block *b, int g, double *phi) {
  int nl[3];
  or_level_dims(b, g, nl);
  const int pi = nl[0] + 2, pj = nl[1] + 2, pk = nl[2] + 2;
  if (b->bc[0] != OR_BC_PERIODIC)
    for (int k = 0; k < pk; ++k)
      for (int j = 0; j < pj; ++j) phi[IDX(0, j, k, pi, pj)] = phi[IDX(1, j, k, pi, pj)];
  if (b->bc[1] != OR_BC_PERIODIC)
    for (int k = 0; k < pk; ++k)
      for (int j = 0; j < pj; ++j) phi[IDX(pi - 1, j, k, pi, pj)] = phi[IDX(pi - 2, j, k, pi, pj)];
  if (b->bc[2] != OR_BC_PERIODIC)
    for (int k = 0; k < pk; ++k)
      for (int i = 0; i < pi; ++i) phi[IDX(i, 0, k, pi, pj)] = phi[IDX(i, 1, k, pi, pj)];
  if (b->bc[3] != OR_BC_PERIODIC)
    for (int k = 0; k < pk; ++k)
      for (int i = 0; i < pi; ++i) phi[IDX(i, pj - 1, k, pi, pj)] = phi[IDX(i, pj - 2, k, pi, pj)];
  if (b->bc[4] != OR_BC_PERIODIC)
    for (int j = 0; j < pj; ++j)
      for (int i = 0; i < pi; ++i) phi[IDX(i, j, 0, pi, pj)] = phi[IDX(i, j, 1, pi, pj)];
  if (b->bc[5] != OR_BC_PERIODIC)
    for (int j = 0; j < pj; ++j)
      for (int i = 0; i < pi; ++i) phi[IDX(i, j, pk - 1, pi, pj)] = phi[IDX(i, j, pk - 2, pi, pj)];
}

/* exchbc_mgpp with one block spanning each periodic axis
 * (exchangepp.for:2872-2881 west, 2922-2931 east, then y and z likewise):
 * ghost(1) = phi(n-1), ghost(n) = phi(2) in 1-based padded indices. */
static void periodic_wrap(const or_block *b, int g, double *phi) {
  int nl[3];
  or_level_dims(b, g, nl);
  const int pi = nl[0] + 2, pj = nl[1] + 2, pk = nl[2] + 2;
  if (b->bc[0] == OR_BC_PERIODIC || b->bc[1] == OR_BC_PERIODIC) {
    if (b->bc[0] == OR_BC_PERIODIC)
      for (int k = 0; k < pk; ++k)
        for (int j = 0; j < pj; ++j) phi[IDX(0, j, k, pi, pj)] = phi[IDX(pi - 2, j, k, pi, pj)];
    if (b->bc[1] == OR_BC_PERIODIC)
      for (int k = 0; k < pk; ++k)
        for (int j = 0; j < pj; ++j) phi[IDX(pi - 1, j, k, pi, pj)] = phi[IDX(1, j, k, pi, pj)];
  }
  if (b->bc[2] == OR_BC_PERIODIC || b->bc[3] == OR_BC_PERIODIC) {
    if (b->bc[2] == OR_BC_PERIODIC)
      for (int k = 0; k < pk; ++k)
        for (int i = 0; i < pi; ++i) phi[IDX(i, 0, k, pi, pj)] = phi[IDX(i, pj - 2, k, pi, pj)];
    if (b->bc[3] == OR_BC_PERIODIC)
      for (int k = 0; k < pk; ++k)
        for (int i = 0; i < pi; ++i) phi[IDX(i, pj - 1, k, pi, pj)] = phi[IDX(i, 1, k, pi, pj)];
  }
  if (b->bc[4] == OR_BC_PERIODIC || b->bc[5] == OR_BC_PERIODIC) {
    if (b->bc[4] == OR_BC_PERIODIC)
      for (int j = 0; j < pj; ++j)
        for (int i = 0; i < pi; ++i) phi[IDX(i, j, 0, pi, pj)] = phi[IDX(i, j, pk - 2, pi, pj)];
    if (b->bc[5] == OR_BC_PERIODIC)
      for (int j = 0; j < pj; ++j)
        for (int i = 0; i < pi; ++i) phi[IDX(i, j, pk - 1, pi, pj)] = phi[IDX(i, j, 1, pi, pj)];
  }
}

/* exchangepp(g) for a block whose six face neighbours are all physical:
 * the face/corner/edge sync loops find nothing to do, then mgbound
 * (exchangepp.for:2830) and exchbc_mgpp when PERIODIC (:2832). */
void or_exchange(const or_block *b, int g, double *phi) {
  or_bound(b, g, phi);
  int per = 0;
  for (int f = 0; f < 6; ++f) per |= (b->bc[f] == OR_BC_PERIODIC);
  if (per) periodic_wrap(b, g, phi);
}

/* 7-point neighbour sum in the reference order
 * (aw*W + ae*E + as*S + an*N + ab*B + at*T), mgsolver.for:567-573 */
static inline double nb_sum(const cof7 *c, const double *phi, size_t p, size_t si, size_t sj,
                            size_t sk) {
  return c->aw * phi[p - si] + c->ae * phi[p + si] + c->as * phi[p - sj] + c->an * phi[p + sj] +
         c->ab * phi[p - sk] + c->at * phi[p + sk];
}

/* red-black sweep, mgsolver.for:551-605: red = MOD(i+j+k,2)==0 in 1-based
 * padded indices, i.e. (i+j+k) even in 0-based interior indices here. */
void or_relax_rb(const or_block *b, int g, int nsweep, double *phi, const double *rhs) {
  int nl[3];
  or_level_dims(b, g, nl);
  const levcoef lc = level_coef(b, g);
  const size_t pi = (size_t)nl[0] + 2, pj = (size_t)nl[1] + 2;
  const size_t si = 1, sj = pi, sk = pi * pj;
  for (int it = 0; it < nsweep; ++it) {
    for (int pass = 0; pass < 2; ++pass) { /* pass 0 = red, 1 = black */
      for (int k = 0; k < nl[2]; ++k)
        for (int j = 0; j < nl[1]; ++j)
          for (int i = 0; i < nl[0]; ++i) {
            /* 1-based padded index = 0-based interior index + 2 */
            if (((i + j + k) & 1) != pass) continue;
            const cof7 c = cell_cof(b, &lc, nl, i, j, k);
            const size_t p = IDX(i + 1, j + 1, k + 1, pi, pj);
            const double r = rhs[(size_t)i + (size_t)nl[0] * ((size_t)j + (size_t)nl[1] * k)];
            phi[p] = (r - nb_sum(&c, phi, p, si, sj, sk)) / c.ap;
          }
    }
    or_bound(b, g, phi);    /* mgsolver.for:614 */
    or_exchange(b, g, phi); /* mgsolver.for:615 */
  }
}

/* lexicographic Gauss-Seidel, mgsolver.for:407-432 (k, j, i loop order) */
void or_relax_gs(const or_block *b, int g, int nsweep, double *phi, const double *rhs) {
  int nl[3];
  or_level_dims(b, g, nl);
  const levcoef lc = level_coef(b, g);
  const size_t pi = (size_t)nl[0] + 2, pj = (size_t)nl[1] + 2;
  const size_t si = 1, sj = pi, sk = pi * pj;
  for (int it = 0; it < nsweep; ++it) {
    for (int k = 0; k < nl[2]; ++k)
      for (int j = 0; j < nl[1]; ++j)
        for (int i = 0; i < nl[0]; ++i) {
          const cof7 c = cell_cof(b, &lc, nl, i, j, k);
          const size_t p = IDX(i + 1, j + 1, k + 1, pi, pj);
          const double r = rhs[(size_t)i + (size_t)nl[0] * ((size_t)j + (size_t)nl[1] * k)];
          phi[p] = (r - nb_sum(&c, phi, p, si, sj, sk)) / c.ap;
        }
    or_bound(b, g, phi);
    or_exchange(b, g, phi);
  }
}

/* line TDMA along j, then i, then k lines, mgsolver.for:434-549.
 * a(1) is never written and stays 0; c(1) is the ghost value. */
void or_relax_tdma(const or_block *b, int g, int nsweep, double *phi, const double *rhs) {
  int nl[3];
  or_level_dims(b, g, nl);
  const levcoef lc = level_coef(b, g);
  const size_t pi = (size_t)nl[0] + 2, pj = (size_t)nl[1] + 2;
  const size_t si = 1, sj = pi, sk = pi * pj;
  int nmax = nl[0] > nl[1] ? nl[0] : nl[1];
  if (nl[2] > nmax) nmax = nl[2];
  double *a = (double *)calloc((size_t)nmax + 2, sizeof(double));
  double *cc = (double *)calloc((size_t)nmax + 2, sizeof(double));
#define RHS(i, j, k) rhs[(size_t)(i) + (size_t)nl[0] * ((size_t)(j) + (size_t)nl[1] * (k))]
  for (int it = 0; it < nsweep; ++it) {
    /* J-lines (mgsolver.for:440-475) */
    for (int k = 0; k < nl[2]; ++k)
      for (int i = 0; i < nl[0]; ++i) {
        cc[0] = phi[IDX(i + 1, 0, k + 1, pi, pj)];
        for (int j = 0; j < nl[1]; ++j) {
          const cof7 c = cell_cof(b, &lc, nl, i, j, k);
          const size_t p = IDX(i + 1, j + 1, k + 1, pi, pj);
          const double apr = 1.0 / (c.ap - c.as * a[j]);
          a[j + 1] = c.an * apr;
          cc[j + 1] = (RHS(i, j, k) - c.ae * phi[p + si] - c.aw * phi[p - si] - c.at * phi[p + sk] -
                       c.ab * phi[p - sk] - c.as * cc[j]) *
                      apr;
        }
        for (int j = nl[1] - 1; j >= 0; --j) {
          const size_t p = IDX(i + 1, j + 1, k + 1, pi, pj);
          phi[p] = (cc[j + 1] - a[j + 1] * phi[p + sj]);
        }
      }
    /* I-lines (mgsolver.for:478-512) */
    for (int k = 0; k < nl[2]; ++k)
      for (int j = 0; j < nl[1]; ++j) {
        cc[0] = phi[IDX(0, j + 1, k + 1, pi, pj)];
        for (int i = 0; i < nl[0]; ++i) {
          const cof7 c = cell_cof(b, &lc, nl, i, j, k);
          const size_t p = IDX(i + 1, j + 1, k + 1, pi, pj);
          const double apr = 1.0 / (c.ap - c.aw * a[i]);
          a[i + 1] = c.ae * apr;
          cc[i + 1] = (RHS(i, j, k) - c.an * phi[p + sj] - c.as * phi[p - sj] - c.at * phi[p + sk] -
                       c.ab * phi[p - sk] - c.aw * cc[i]) *
                      apr;
        }
        for (int i = nl[0] - 1; i >= 0; --i) {
          const size_t p = IDX(i + 1, j + 1, k + 1, pi, pj);
          phi[p] = (cc[i + 1] - a[i + 1] * phi[p + si]);
        }
      }
    /* K-lines (mgsolver.for:515-549) */
    for (int i = 0; i < nl[0]; ++i)
      for (int j = 0; j < nl[1]; ++j) {
        cc[0] = phi[IDX(i + 1, j + 1, 0, pi, pj)];
        for (int k = 0; k < nl[2]; ++k) {
          const cof7 c = cell_cof(b, &lc, nl, i, j, k);
          const size_t p = IDX(i + 1, j + 1, k + 1, pi, pj);
          const double apr = 1.0 / (c.ap - c.ab * a[k]);
          a[k + 1] = c.at * apr;
          cc[k + 1] = (RHS(i, j, k) - c.an * phi[p + sj] - c.as * phi[p - sj] - c.ae * phi[p + si] -
                       c.aw * phi[p - si] - c.ab * cc[k]) *
                      apr;
        }
        for (int k = nl[2] - 1; k >= 0; --k) {
          const size_t p = IDX(i + 1, j + 1, k + 1, pi, pj);
          phi[p] = (cc[k + 1] - a[k + 1] * phi[p + sk]);
        }
      }
    or_bound(b, g, phi);
    or_exchange(b, g, phi);
  }
#undef RHS
  free(a);
  free(cc);
}

/* resf = rhs - (aw*W + ae*E + as*S + an*N + ab*B + at*T + ap*P),
 * mgsolver.for:707-714 */
void or_residual(const or_block *b, int g, const double *phi, const double *rhs, double *res) {
  int nl[3];
  or_level_dims(b, g, nl);
  const levcoef lc = level_coef(b, g);
  const size_t pi = (size_t)nl[0] + 2, pj = (size_t)nl[1] + 2;
  const size_t si = 1, sj = pi, sk = pi * pj;
  for (int k = 0; k < nl[2]; ++k)
    for (int j = 0; j < nl[1]; ++j)
      for (int i = 0; i < nl[0]; ++i) {
        const cof7 c = cell_cof(b, &lc, nl, i, j, k);
        const size_t p = IDX(i + 1, j + 1, k + 1, pi, pj);
        const size_t q = (size_t)i + (size_t)nl[0] * ((size_t)j + (size_t)nl[1] * k);
        res[q] = rhs[q] - (c.aw * phi[p - si] + c.ae * phi[p + si] + c.as * phi[p - sj] +
                           c.an * phi[p + sj] + c.ab * phi[p - sk] + c.at * phi[p + sk] +
                           c.ap * phi[p]);
      }
}

void or_apply(const or_block *b, const double *p, double *ap) {
  int nl[3];
  or_level_dims(b, 1, nl);
  const levcoef lc = level_coef(b, 1);
  const size_t pi = (size_t)nl[0] + 2, pj = (size_t)nl[1] + 2;
  const size_t si = 1, sj = pi, sk = pi * pj;
  for (int k = 0; k < nl[2]; ++k)
    for (int j = 0; j < nl[1]; ++j)
      for (int i = 0; i < nl[0]; ++i) {
        const cof7 c = cell_cof(b, &lc, nl, i, j, k);
        const size_t q = IDX(i + 1, j + 1, k + 1, pi, pj);
        ap[(size_t)i + (size_t)nl[0] * ((size_t)j + (size_t)nl[1] * k)] =
            c.aw * p[q - si] + c.ae * p[q + si] + c.as * p[q - sj] + c.an * p[q + sj] +
            c.ab * p[q - sk] + c.at * p[q + sk] + c.ap * p[q];
      }
}

/* mgrestr(g), mgsolver.for:655-672 (zero coarse phi incl. ghosts and coarse
 * rhs), 693-717 (fine residual), 741-756 (0.125 x the 8 children, i fastest,
 * then j, then k). */
void or_restrict(const or_block *b, int g, const double *phi_f, const double *rhs_f,
                 double *phi_c, double *rhs_c) {
  int nc[3], nf[3];
  or_level_dims(b, g, nc);
  or_level_dims(b, g - 1, nf);
  const size_t ncp = (size_t)(nc[0] + 2) * (nc[1] + 2) * (nc[2] + 2);
  memset(phi_c, 0, ncp * sizeof(double));
  double *res = (double *)malloc((size_t)nf[0] * nf[1] * nf[2] * sizeof(double));
  or_residual(b, g - 1, phi_f, rhs_f, res);
#define R(i, j, k) res[(size_t)(i) + (size_t)nf[0] * ((size_t)(j) + (size_t)nf[1] * (k))]
  for (int k = 0; k < nc[2]; ++k)
    for (int j = 0; j < nc[1]; ++j)
      for (int i = 0; i < nc[0]; ++i) {
        const int fi = 2 * i, fj = 2 * j, fk = 2 * k;
        rhs_c[(size_t)i + (size_t)nc[0] * ((size_t)j + (size_t)nc[1] * k)] =
            0.125 * (R(fi, fj, fk) + R(fi + 1, fj, fk) + R(fi, fj + 1, fk) + R(fi + 1, fj + 1, fk) +
                     R(fi, fj, fk + 1) + R(fi + 1, fj, fk + 1) + R(fi, fj + 1, fk + 1) +
                     R(fi + 1, fj + 1, fk + 1));
      }
#undef R
  free(res);
}

/* mgcorr(g), mgsolver.for:828-896 then mgbound (:907).  For coarse padded
 * (ic,jc,kc) in [0, nc] (ghosts included) the fine padded children are
 * (2ic+di, 2jc+dj, 2kc+dk); the weight of coarse corner (oi,oj,ok) is
 * prod_a (3 if o_a == d_a else 1); terms are summed in the order
 * c, c+j, c+k, c+k+j, c+i, c+i+j, c+i+k, c+i+j+k and divided by 64. */
void or_prolong(const or_block *b, int g, double *phi_f, const double *phi_c) {
  int nc[3], nf[3];
  or_level_dims(b, g + 1, nc);
  or_level_dims(b, g, nf);
  const size_t fpi = (size_t)nf[0] + 2, fpj = (size_t)nf[1] + 2;
  const size_t cpi = (size_t)nc[0] + 2, cpj = (size_t)nc[1] + 2;
  static const int ord[8][3] = {{0, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1},
                                {1, 0, 0}, {1, 1, 0}, {1, 0, 1}, {1, 1, 1}};
  for (int kc = 0; kc <= nc[2]; ++kc)
    for (int ic = 0; ic <= nc[0]; ++ic)
      for (int jc = 0; jc <= nc[1]; ++jc) {
        double cv[8];
        for (int t = 0; t < 8; ++t)
          cv[t] = phi_c[IDX(ic + ord[t][0], jc + ord[t][1], kc + ord[t][2], cpi, cpj)];
        for (int dk = 0; dk < 2; ++dk)
          for (int dj = 0; dj < 2; ++dj)
            for (int di = 0; di < 2; ++di) {
              double s = 0.0;
              for (int t = 0; t < 8; ++t) {
                const double w = (ord[t][0] == di ? 3.0 : 1.0) * (ord[t][1] == dj ? 3.0 : 1.0) *
                                 (ord[t][2] == dk ? 3.0 : 1.0);
                s = (t == 0) ? w * cv[t] : s + w * cv[t];
              }
              const size_t f = IDX(2 * ic + di, 2 * jc + dj, 2 * kc + dk, fpi, fpj);
              phi_f[f] = phi_f[f] + s / 64.0;
            }
      }
  or_bound(b, g, phi_f);
}

typedef void (*relax_fn)(const or_block *, int, int, double *, const double *);

/* mgkcyc, mgsolver.for:226-372 (kcycle = 1: kount==2 ends a level visit). */
int or_vcycle(const or_block *b, int scheme, int nu1, int nu2, const double *su, double *pp) {
  relax_fn relax = scheme == 1 ? or_relax_gs : scheme == 2 ? or_relax_tdma : or_relax_rb;
  const int L = b->nlev;
  if (L < 1 || L > 30) return -1;
  for (int a = 0; a < 3; ++a)
    if (b->n[a] % (1 << (L - 1)) != 0) return -2; /* localparameters.for:202-208 */
  double **phi = (double **)calloc((size_t)L + 1, sizeof(double *));
  double **rhs = (double **)calloc((size_t)L + 1, sizeof(double *));
  for (int g = 1; g <= L; ++g) {
    int nl[3];
    or_level_dims(b, g, nl);
    const size_t np = (size_t)(nl[0] + 2) * (nl[1] + 2) * (nl[2] + 2);
    phi[g] = (double *)calloc(np, sizeof(double));
    rhs[g] = (double *)calloc((size_t)nl[0] * nl[1] * nl[2], sizeof(double));
  }
  {
    int nl[3];
    or_level_dims(b, 1, nl);
    memcpy(phi[1], pp, (size_t)(nl[0] + 2) * (nl[1] + 2) * (nl[2] + 2) * sizeof(double));
    memcpy(rhs[1], su, (size_t)nl[0] * nl[1] * nl[2] * sizeof(double));
  }
  int *kount = (int *)calloc((size_t)L + 2, sizeof(int));
  relax(b, 1, nu1, phi[1], rhs[1]);
  if (L > 1) {
    or_restrict(b, 2, phi[1], rhs[1], phi[2], rhs[2]);
    int g = 2, nrel = nu1;
    while (g != 1) {
      kount[g] += 1;
      relax(b, g, nrel, phi[g], rhs[g]);
      if (kount[g] == 2) {
        or_prolong(b, g - 1, phi[g - 1], phi[g]);
        kount[g] = 0;
        g -= 1;
        nrel = nu2;
      } else if (g < L) {
        or_restrict(b, g + 1, phi[g], rhs[g], phi[g + 1], rhs[g + 1]);
        g += 1;
        nrel = nu1;
      } else {
        relax(b, g, nu2, phi[g], rhs[g]);
        or_prolong(b, g - 1, phi[g - 1], phi[g]);
        nrel = nu2;
        g = L - 1;
      }
    }
  }
  relax(b, 1, nu2, phi[1], rhs[1]);
  {
    int nl[3];
    or_level_dims(b, 1, nl);
    memcpy(pp, phi[1], (size_t)(nl[0] + 2) * (nl[1] + 2) * (nl[2] + 2) * sizeof(double));
  }
  for (int g = 1; g <= L; ++g) {
    free(phi[g]);
    free(rhs[g]);
  }
  free(phi);
  free(rhs);
  free(kount);
  return 0;
}

static double l2(const double *x, size_t n) {
  double s = 0.0;
  for (size_t q = 0; q < n; ++q) s += x[q] * x[q];
  return sqrt(s);
}

/* Pressure-form counterpart of newsolv_mg.for:19-59: the rhs of each V-cycle
 * is r = f - A p (what calvel+calmas produce in exact arithmetic), pp starts
 * at zero, p = p + (pp - 0) over the interior and the ghost layer
 * (newsolv_mg.for:39-49), then the p ghosts are re-imaged (pbound +
 * exchange(4) for one block). */
int or_solve(const or_block *b, int scheme, int nu1, int nu2, const double *f, double *p,
             int ncycles, double tol, double *hist) {
  int nl[3];
  or_level_dims(b, 1, nl);
  const size_t n = (size_t)nl[0] * nl[1] * nl[2];
  const size_t np = (size_t)(nl[0] + 2) * (nl[1] + 2) * (nl[2] + 2);
  double *r = (double *)malloc(n * sizeof(double));
  double *pp = (double *)malloc(np * sizeof(double));
  or_exchange(b, 1, p);
  or_residual(b, 1, p, f, r);
  hist[0] = l2(r, n);
  const double r0 = hist[0];
  int it = 0;
  int rc = 0;
  for (it = 1; it <= ncycles; ++it) {
    memset(pp, 0, np * sizeof(double));
    rc = or_vcycle(b, scheme, nu1, nu2, r, pp);
    if (rc != 0) break;
    for (size_t q = 0; q < np; ++q) p[q] = (p[q] + (pp[q] - 0.0));
    or_exchange(b, 1, p);
    or_residual(b, 1, p, f, r);
    hist[it] = l2(r, n);
    if (tol > 0.0 && hist[it] <= tol * r0) break;
  }
  free(r);
  free(pp);
  if (rc != 0) return rc;
  return it > ncycles ? ncycles : it;
}
