/* h3dmg_oracle.h -- CPU restatement of the Hydro3D pressure-Poisson multigrid.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP solver:
 * it is loaded by tests/, by __graft_entry__.smoke() and by the CPU-baseline
 * leg of bench.py, never by the product path.
 *
 * It restates, in plain C with the reference's arithmetic order, the routines
 * of src_v70/mgsolver.for (coeff, mgkcyc, mgrelax, mgrestr, mgcorr, mgbound)
 * and the single-block branches of src_v70/exchangepp.for (exchangepp with
 * no same-level neighbour, exchbc_mgpp with one block spanning a periodic
 * axis), plus a pressure-form outer loop around mgkcyc.
 *
 * Layout (differs from the reference's packed j-fastest `cof`, same values):
 * every level array is i-fastest.  phi is padded with one ghost layer,
 * (n_i+2)*(n_j+2)*(n_k+2) doubles, ghost at padded index 0 and n+1; rhs and
 * residual arrays are interior, n_i*n_j*n_k doubles.
 */
#ifndef H3DMG_ORACLE_H
#define H3DMG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* boundary kind per face, order west, east, south, north, bottom, top */
#define OR_BC_WALL 0     /* any reference bc code other than 5 (Neumann) */
#define OR_BC_PERIODIC 1 /* reference bc code 5 */

typedef struct {
  int n[3];      /* interior cells at the finest level (i, j, k) */
  double h[3];   /* dx, dy, dz at the finest level */
  int bc[6];     /* OR_BC_* per face */
  int nlev;      /* number of multigrid levels (ngrid) */
} or_block;

/* interior size of level g (1-based) along each axis */
void or_level_dims(const or_block *b, int g, int nl[3]);

/* mgrelax(g, nsweep) with the red-black scheme, including the halo update
 * after every sweep (mgsolver.for:374-621 with mg_itrsch = 3) */
void or_relax_rb(const or_block *b, int g, int nsweep, double *phi, const double *rhs);
/* mgrelax with lexicographic Gauss-Seidel (mg_itrsch = 1) */
void or_relax_gs(const or_block *b, int g, int nsweep, double *phi, const double *rhs);
/* mgrelax with line TDMA along j, i, k lines (mg_itrsch = 2) */
void or_relax_tdma(const or_block *b, int g, int nsweep, double *phi, const double *rhs);
/* mgbound(g) (mgsolver.for:912-997) */
void or_bound(const or_block *b, int g, double *phi);
/* exchangepp(g) for a block without neighbours (exchangepp.for:2830-2832,
 * exchbc_mgpp 2837-3453) */
void or_exchange(const or_block *b, int g, double *phi);
/* fine residual of mgrestr (mgsolver.for:693-717), interior sized */
void or_residual(const or_block *b, int g, const double *phi, const double *rhs, double *res);
/* mgrestr(g): g is the coarse level (mgsolver.for:623-766) */
void or_restrict(const or_block *b, int g, const double *phi_f, const double *rhs_f,
                 double *phi_c, double *rhs_c);
/* mgcorr(g): g is the fine level (mgsolver.for:768-910) */
void or_prolong(const or_block *b, int g, double *phi_f, const double *phi_c);
/* mgkcyc (mgsolver.for:226-372); scheme 1/2/3 = GS/TDMA/RB.
 * su: interior rhs; pp: padded level-1 array, in (initial guess) and out. */
int or_vcycle(const or_block *b, int scheme, int nu1, int nu2, const double *su, double *pp);
/* Pressure-form outer loop: r = f - A p, V-cycle on r from zero, p += pp.
 * hist[0..ncycles] receives ||f - A p_k||_2; stops after `ncycles` cycles or
 * when hist[k] <= tol * hist[0] (tol <= 0 disables).  p is padded, in/out.
 * Returns the number of cycles run or a negative error. */
int or_solve(const or_block *b, int scheme, int nu1, int nu2, const double *f, double *p,
             int ncycles, double tol, double *hist);
/* ||A p - f|| helpers: A p on the interior of a padded p at level 1 */
void or_apply(const or_block *b, const double *p, double *ap);

#ifdef __cplusplus
}
#endif
#endif
