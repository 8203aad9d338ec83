/* h3dmg.h -- C-ABI of the MI355X pressure-Poisson multigrid for Hydro3D.
 *
 * The reference exposes the solve as argument-less Fortran subroutines that
 * read and write module state (src_v70/module_vars.for, module_multidata.for):
 *
 *   newsolv_mg  src_v70/newsolv_mg.for:2-93   outer projection loop, called
 *                                            from src_v70/flosol.for:215
 *   coeff       src_v70/mgsolver.for:2-224    per-level 7-point coefficients
 *   mgkcyc      src_v70/mgsolver.for:226-372  one V-cycle, pp <- V(su)
 *
 * This header replaces that seam with explicit handles, plain pointers and
 * sizes (no Fortran module state, no torch types).  The Fortran binding is
 * hydro3d_amd/fortran/h3dmg_mod.f90 (ISO_C_BINDING) and the drop-in
 * `newsolv_mg` is hydro3d_amd/fortran/newsolv_mg_h3d.f90.
 *
 * Array conventions at this boundary:
 *   host arrays are Fortran column-major (i fastest).  A "Fortran array" is
 *   described by its leading dimensions ld1, ld2 and the 0-based offset `off`
 *   of the first interior cell on every axis (off = pl in the reference's
 *   p(ttc_i,ttc_j,ttc_k) arrays, src_v70/localparameters.for:106-143).
 *   Device-pointer entry points take dense interior arrays n_i*n_j*n_k
 *   (i fastest) that already live in HBM.
 *
 * Threading: a process may hold any number of handles and drive them from
 * different host threads (each handle launches on its own stream); calls on
 * one handle must not overlap.  All calls are synchronous with respect to
 * the host, except h3dmg_project_device_noalloc, which only enqueues work on
 * the handle's stream.
 */
#ifndef H3DMG_H
#define H3DMG_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  H3DMG_OK = 0,
  H3DMG_ERR_ARG = 1,           /* bad argument (null pointer, bad size, ...) */
  H3DMG_ERR_LEVELS = 2,        /* n not divisible by 2^(nlevels-1)
                                  (src_v70/localparameters.for:202-208)     */
  H3DMG_ERR_DEVICE = 3,        /* HIP runtime error                          */
  H3DMG_ERR_NOMEM = 4,         /* device allocation refused                  */
  H3DMG_ERR_NOT_CONVERGED = 5, /* maxcy cycles without meeting the criterion
                                  (newsolv_mg.for:61-62 prints and goes on)  */
  H3DMG_ERR_DIVERGED = 6,      /* residual non-finite or rmax > 10
                                  (newsolv_mg.for:65-75 stops the run)       */
  H3DMG_ERR_UNSUPPORTED = 7,   /* option not available on this build         */
  H3DMG_ERR_BUSY = 8           /* serving: no capacity for the request now   */
} h3dmg_status;

/* boundary kind per face (west, east, south, north, bottom, top).  The
 * reference treats bc code 5 as periodic and every other code as a Neumann
 * wall for the pressure (src_v70/mgsolver.for:194-205).  H3DMG_BC_NEIGHBOUR:
 * the face is shared with another block of the same handle (h3dmg_block.nbr),
 * the reference's same-level neighbour (iprev..knext, alloc_dom.for:524-690)
 * or, on a periodic axis spanning several blocks, its periodic partner
 * (per_ip..per_kn, alloc_dom.for:1040-1215). */
enum { H3DMG_BC_WALL = 0, H3DMG_BC_PERIODIC = 1, H3DMG_BC_NEIGHBOUR = 2 };

/* smoother (mg_itrsch, src_v70/control.cin line 10; mgrelax,
 * src_v70/mgsolver.for:407-605).  GS and TDMA run in FP64 on the
 * level-by-level path (no mixed precision, LDS tail or fused sweeps): one
 * launch per diagonal of 8^3 tiles (GS) or per line diagonal of each TDMA
 * pass, bitwise the reference's in-place loop order. */
enum { H3DMG_SMOOTH_GS = 1, H3DMG_SMOOTH_TDMA = 2, H3DMG_SMOOTH_RBGS = 3 };

/* stopping criterion */
enum {
  H3DMG_CRIT_REL_L2 = 0, /* ||f - A p||_2 / ||f||_2 <= tol                  */
  H3DMG_CRIT_REF_RMAX = 1 /* reference: rmax < eps and iter > 1, with
                             rmax = max|div(u)| dx dy dz (press.for:66-73)  */
};

/* arithmetic of the V-cycle */
enum {
  H3DMG_FP64 = 0, /* plain: every level in FP64 (the reference's arithmetic) */
  H3DMG_MIXED = 1 /* FP32 smoothing and transfers inside FP64 iterative
                     refinement; residual, norms, update and the coarsest
                     level in FP64                                            */
};

/* treatment of the coarsest level */
enum {
  H3DMG_COARSE_SMOOTH = 0, /* reference: nu1 + nu2 sweeps, no exact solve
                              (src_v70/mgsolver.for:305,335)                 */
  H3DMG_COARSE_EXACT = 1   /* FP64 direct solve in the discrete eigenbasis,
                              null-space (mean) mode removed                 */
};

typedef struct {
  int ni, nj, nk;     /* interior cells at the finest level */
  double dx, dy, dz;  /* spacing at the finest level */
  int nlevels;        /* multigrid levels (ngrid) */
  int bc[6];          /* H3DMG_BC_* per face; a periodic axis must have both
                         faces periodic */
  int nbr[6];         /* for bc[f] == H3DMG_BC_NEIGHBOUR: index of the block
                         across face f (read only for such faces) */
} h3dmg_block;

typedef struct {
  int smoother;      /* H3DMG_SMOOTH_*                              */
  int nu1, nu2;      /* irestr, iproln (control.cin line 11)        */
  int maxcy;         /* cycle cap (control.cin line 11)             */
  int criterion;     /* H3DMG_CRIT_*                                */
  double tol;        /* REL_L2 tolerance or reference eps           */
  double dt;         /* time step, used by the REF_RMAX criterion   */
  int precision;     /* H3DMG_FP64 | H3DMG_MIXED                    */
  int coarse;        /* H3DMG_COARSE_*                              */
  int min_cycles;    /* cycles always run (reference: 2)            */
  int use_graph;     /* capture one V-cycle in a hipGraph           */
} h3dmg_params;

typedef struct h3dmg_handle h3dmg_handle;

/* Fill `p` with the defaults: RBGS, V(2,2), maxcy 60, REL_L2 1e-8, FP64,
 * reference coarsest level, min_cycles 1, graphs on. */
void h3dmg_default_params(h3dmg_params *p);

/* Build the hierarchy of every block in device memory (the role of coeff,
 * mgsolver.for:2-224, and of the cof allocation, localparameters.for:225-239).
 *
 * Blocks without H3DMG_BC_NEIGHBOUR faces are independent solves.  Blocks
 * joined by H3DMG_BC_NEIGHBOUR faces form one domain, as the reference's
 * multi-block layout (one rank's blocks): every ghost a smoother, residual,
 * restriction or prolongation reads is the value exchangepp leaves in it
 * (exchangepp.for:2-2835, exchbc_mgpp :2837-3453) -- the neighbour's cell
 * from before the sweep, edge and corner ghosts included -- with no copies:
 * kernels read the neighbour's arrays.  Requirements (H3DMG_ERR_ARG or
 * H3DMG_ERR_UNSUPPORTED otherwise): the two blocks of a shared face name
 * each other on opposite faces, have the same spacing and level count and
 * the same extent along the face; the blocks form a box, every edge or
 * corner where two or three shared faces meet having its diagonal block
 * (the reference leaves such ghosts stale otherwise); red-black smoother,
 * H3DMG_COARSE_SMOOTH, constant coefficients.  Such a handle stops all its
 * blocks together: REL_L2 on the norms summed over the blocks, REF_RMAX on
 * the maximum. */
h3dmg_status h3dmg_create(const h3dmg_block *blocks, int nblocks, const h3dmg_params *prm,
                          h3dmg_handle **out);
void h3dmg_destroy(h3dmg_handle *h);
const char *h3dmg_last_error(void);

/* Device memory the handle holds, in bytes. */
unsigned long long h3dmg_device_bytes(const h3dmg_handle *h);

/* Right-hand side su of block `b` (dom(ib)%su, mgsolver.for:262) from a host
 * Fortran array; interior cells only.  Also records ||f||_2. */
h3dmg_status h3dmg_set_rhs(h3dmg_handle *h, int b, const double *su, int ld1, int ld2, int off);
/* Same from a dense device interior array. */
h3dmg_status h3dmg_set_rhs_device(h3dmg_handle *h, int b, const double *d_su);

/* Pressure of block b (dom(ib)%p): warm start in, solution out.  Host
 * Fortran arrays; `get` writes the interior and the one-cell ghost layer
 * (images: mirror at walls, wrap at periodic faces). */
h3dmg_status h3dmg_set_p(h3dmg_handle *h, int b, const double *p, int ld1, int ld2, int off);
h3dmg_status h3dmg_get_p(h3dmg_handle *h, int b, double *p, int ld1, int ld2, int off);
h3dmg_status h3dmg_set_p_device(h3dmg_handle *h, int b, const double *d_p);
h3dmg_status h3dmg_get_p_device(h3dmg_handle *h, int b, double *d_p);

/* One mgkcyc on the current rhs with pp = 0 (mgsolver.for:226-372); the
 * correction is then read with h3dmg_get_pp (padded, ghosts included, as the
 * reference copies cof(phi_1) back into pp at mgsolver.for:354-367). */
h3dmg_status h3dmg_vcycle(h3dmg_handle *h);
h3dmg_status h3dmg_get_pp(h3dmg_handle *h, int b, double *pp, int ld1, int ld2, int off);

/* Pressure-form outer loop (newsolv_mg.for:19-59 with su = f - A p):
 * cycles until every block meets the criterion (a converged block stops
 * cycling) or maxcy.  hist (may be null) receives, per cycle k = 0..ncycles,
 * the largest relative residual over the blocks; ncycles and final_res may be
 * null.  Returns H3DMG_ERR_NOT_CONVERGED when maxcy is reached (p holds the
 * last iterate, as the reference goes on).  With use_graph (not pooled) the
 * stopping rule runs on the device and the next cycle is queued before the
 * host reads the last one's outcome: on return, one more cycle in which no
 * block is active may still be running on the handle's stream (it changes
 * nothing the API returns; later calls on the stream follow it).  The first
 * such solve allocates a small pinned host buffer and device state. */
h3dmg_status h3dmg_solve(h3dmg_handle *h, double *hist, int hist_cap, int *ncycles,
                         double *final_res);

/* Velocity projection of block b with its current pressure, on the device
 * (calvel, src_v70/press.for:77-144, then calmas, press.for:2-75, 2nd-order
 * branch): u = u* - dt*alfapr*grad p on the faces the reference corrects
 * (interior faces and periodic wrap faces; wall-normal faces keep u*),
 * su = div(u)/dt, and *rmax = max|div(u)| dx dy dz (may be null).  Dense
 * device arrays, i fastest: u*, u (nk, nj, ni+1) with face f between cells
 * f-1 and f; v*, v (nk, nj+1, ni); w*, w (nk+1, nj, ni); su (nk, nj, ni) or
 * null.  dt is params.dt.  Bitwise the reference's arithmetic. */
h3dmg_status h3dmg_project_device(h3dmg_handle *h, int b, const double *d_ustar,
                                  const double *d_vstar, const double *d_wstar, double *d_u,
                                  double *d_v, double *d_w, double *d_su, double alfapr,
                                  double *rmax);

/* h3dmg_project_device without any allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops).  No
 * counterpart in the reference.  Same arrays and the same arithmetic as
 * h3dmg_project_device, but it only enqueues work on the handle's stream
 * (no host synchronisation, no hipMalloc, no host buffers) and writes rmax
 * into the device scalar d_rmax (may be null).  calmas' partial maxima go to
 * a 16 KiB static device array per GPU instead of the handle: at most 1024
 * partials, one per calmas workgroup (calmas caps its grid there; a block
 * that would need more returns H3DMG_ERR_ARG).  Every call on one GPU shares
 * that array: calls that overlap on one GPU (two streams, two threads) race
 * and must be serialised by the caller. */
h3dmg_status h3dmg_project_device_noalloc(h3dmg_handle *h, int b, const double *d_ustar,
                                          const double *d_vstar, const double *d_wstar,
                                          double *d_u, double *d_v, double *d_w, double *d_su,
                                          double alfapr, double *d_rmax);

/* The velocity form of the outer loop on the device, newsolv_mg
 * (src_v70/newsolv_mg.for:17-59) for every enabled block of the handle:
 *   calvel; calmas;
 *   repeat: mgkcyc (pp from su, zero guess); p = p + (pp - 0);
 *           calvel; calmas;            until rmax < eps and iter > 1
 * The arrays are lists of nblocks device pointers (layouts as in
 * h3dmg_project_device; entries of disabled blocks may be null), dt is
 * params.dt and eps params.tol.  h3dmg_velocity_begin runs the first
 * calvel + calmas with the current pressure (su goes to the blocks' rhs) and
 * returns rmax, the max over the handle's blocks (the host takes the MPI MAX
 * over ranks); each h3dmg_velocity_step runs one iteration and returns the
 * new rmax.  h3dmg_solve_velocity runs the whole loop on one rank: hist[0]
 * is the first rmax, hist[i] the rmax after iteration i; the status is
 * H3DMG_ERR_NOT_CONVERGED after maxcy iterations (newsolv_mg.for:61-62) and
 * H3DMG_ERR_DIVERGED for rmax > 10 (newsolv_mg.for:65-75).  In FP64 every
 * step is the reference's arithmetic, so rmax and p are bitwise the
 * reference loop's.  The final su is h3dmg_get_su_device's. */
h3dmg_status h3dmg_velocity_begin(h3dmg_handle *h, const double *const *d_ustar,
                                  const double *const *d_vstar, const double *const *d_wstar,
                                  double *const *d_u, double *const *d_v, double *const *d_w,
                                  double alfapr, double *rmax);
h3dmg_status h3dmg_velocity_step(h3dmg_handle *h, double *rmax);
h3dmg_status h3dmg_solve_velocity(h3dmg_handle *h, const double *const *d_ustar,
                                  const double *const *d_vstar, const double *const *d_wstar,
                                  double *const *d_u, double *const *d_v, double *const *d_w,
                                  double alfapr, double *hist, int hist_cap, int *ncycles,
                                  double *rmax);
/* Density of block b (dom(ib)%dens, interior cells) switches the handle to
 * the reference's variable-density operator (L_LSM): coeff's LSM branch
 * (src_v70/mgsolver.for:72-184) builds dens_mg and density-weighted
 * coefficients for every level, calvel divides by the face density
 * (press.for:93-131) and calmas writes sup = div(u) dx dy dz / dt
 * (press.for:53-55).  Call it for every block, and again whenever the density
 * changes (the reference rebuilds the coefficients every solve,
 * newsolv_mg.for:13-15).  Needs H3DMG_FP64 and H3DMG_COARSE_SMOOTH; the LDS
 * tail and fused sweeps are turned off.  Host Fortran array as h3dmg_set_rhs
 * takes it, or a dense device array (nk, nj, ni). */
h3dmg_status h3dmg_set_density(h3dmg_handle *h, int b, const double *dens, int ld1, int ld2,
                               int off);
h3dmg_status h3dmg_set_density_device(h3dmg_handle *h, int b, const double *d_dens);

/* calmas' branch in the velocity loop, the reference's `differencing`
 * (control.cin): 1 the 2nd-order divergence (press.for:48-63, the default),
 * 2 the 4th-order one (press.for:17-37),
 *   ((-u(i+1) + 27 u(i) - 27 u(i-1) + u(i-2)) / (24 dx) + ... for v, w),
 * which reads the second ghost faces u(isp-2), u(iep+1) (and v, w alike):
 * across a shared face or a periodic wrap the current face of the block
 * there (what exchange(1..3) copies), at a wall the host's value, which
 * h3dmg_set_velocity takes from its arrays (they then need off >= 2).
 * Differencing 2 needs the velocity of h3dmg_set_velocity, refuses the
 * variable-density operator (whose sup the reference's 4th-order branch
 * does not write), and makes h3dmg_project_device* return
 * H3DMG_ERR_UNSUPPORTED. */
h3dmg_status h3dmg_set_differencing(h3dmg_handle *h, int differencing);

/* The time step of the next solve (params.dt): the reference's dt changes
 * from step to step (checkdt.for:117-137). */
h3dmg_status h3dmg_set_dt(h3dmg_handle *h, double dt);

/* su of block b as the rhs arrays hold it (after h3dmg_set_rhs*, or the
 * last calmas of the velocity loop), into a dense device array (nk, nj, ni). */
h3dmg_status h3dmg_get_su_device(h3dmg_handle *h, int b, double *d_su);

/* Host Fortran face arrays for the velocity loop (dom(ib)%ustar, vstar,
 * wstar and u, v, w; the layout of p: ld1, ld2 and `off` >= 1, u(i,j,k) the
 * face between cells i and i+1, press.for:90-131).  set copies u*, v*, w*
 * into arrays the handle owns (allocated on first use), taking the current
 * u, v, w (may be null: then u*) on the faces calvel does not correct;
 * h3dmg_velocity_begin/h3dmg_solve_velocity called with six null lists use
 * them.  get writes u, v, w of every face of the interior rows back (the
 * remaining ghost layers are the host's exchange(1..3), as after calvel).
 * h3dmg_get_su writes su's interior into a host Fortran array. */
h3dmg_status h3dmg_set_velocity(h3dmg_handle *h, int b, const double *ustar, const double *vstar,
                                const double *wstar, const double *u, const double *v,
                                const double *w, int ld1, int ld2, int off);
h3dmg_status h3dmg_get_velocity(h3dmg_handle *h, int b, double *u, double *v, double *w, int ld1,
                                int ld2, int off);
h3dmg_status h3dmg_get_su(h3dmg_handle *h, int b, double *su, int ld1, int ld2, int off);
/* The faces calvel itself computes (press.for:90-131 over isu..ieu, jsv..jev,
 * ksw..kew, localparameters.for:106-192: faces 1..n of an axis, 1..n-1
 * against a wall at the far end), the rest of the host arrays untouched --
 * the host state after the reference's calvel loops and before its
 * exchange(1..3).  A host that runs exchange(1..3) after each such copy
 * reproduces the ghost layers the reference's loop leaves. */
h3dmg_status h3dmg_get_velocity_calvel(h3dmg_handle *h, int b, double *u, double *v, double *w,
                                       int ld1, int ld2, int off);

/* Per-block results of the last solve: cycles run and final relative
 * residual (arrays of nblocks). */
h3dmg_status h3dmg_block_stats(const h3dmg_handle *h, int *cycles, double *rel_res);

/* Per-block controls for batches of unrelated solves (no reference
 * counterpart; the serving layer's batches):
 *   set_block_tol      REL_L2 tolerance of block b (0: the handle's tol);
 *   set_block_enabled  0 removes block b from h3dmg_solve / h3dmg_vcycle
 *                      (its hierarchy, rhs and p stay resident), 1 restores it.
 * A solve over enabled blocks gives each of them the result it gets alone. */
h3dmg_status h3dmg_set_block_tol(h3dmg_handle *h, int b, double tol);
h3dmg_status h3dmg_set_block_enabled(h3dmg_handle *h, int b, int enabled);

/* ---- continuous batching (no reference counterpart: the serving layer) ---
 * Solves that join and leave a running batch between V-cycles.  The memory
 * of such batches comes from a pool: one device allocation made up front,
 * cut into pieces.  Each block owns a region of the pool (its levels, p, f
 * and coarse-solve tables; h3dmg_block_bytes long, taken with
 * h3dmg_pool_alloc) that outlives the handles built over it: when the set
 * of blocks changes, the caller builds a new handle over the new set with
 * h3dmg_create_pooled, naming each block's region, and destroys the old one;
 * the blocks carried over keep their state in their regions, and their
 * per-block counters move with h3dmg_get/set_block_state.  A pooled handle
 * allocates no device or pinned memory: its tables take pieces of the pool
 * and return them on h3dmg_destroy (H3DMG_ERR_NOMEM when the pool is full).
 * A pooled handle holds independent blocks only (no H3DMG_BC_NEIGHBOUR
 * faces), and refuses the velocity loop and the variable-density operator
 * (H3DMG_ERR_UNSUPPORTED).  One pool serves several handles at once, on
 * several streams and host threads. */
typedef struct h3dmg_pool h3dmg_pool;
/* `bytes` of device memory and `max_blocks` pinned per-block statistics
 * entries (the blocks of all the pool's live handles together). */
h3dmg_status h3dmg_pool_create(unsigned long long bytes, int max_blocks, h3dmg_pool **out);
/* Waits for the device and frees the pool.  While handles built in it are
 * alive it frees nothing (h3dmg_last_error says so): destroy them first. */
void h3dmg_pool_destroy(h3dmg_pool *pool);
/* A piece of at least `bytes` (256-byte aligned offset from the pool's start). */
h3dmg_status h3dmg_pool_alloc(h3dmg_pool *pool, unsigned long long bytes,
                              unsigned long long *offset);
h3dmg_status h3dmg_pool_free(h3dmg_pool *pool, unsigned long long offset);
unsigned long long h3dmg_pool_free_bytes(h3dmg_pool *pool);
unsigned long long h3dmg_pool_largest_free(h3dmg_pool *pool);
/* Bytes of one block's region under `prm` (0 for an invalid block or
 * parameters; h3dmg_last_error says why).  Host-only. */
unsigned long long h3dmg_block_bytes(const h3dmg_block *blk, const h3dmg_params *prm);
/* A handle over `blocks` whose regions start at `offsets` (pieces from
 * h3dmg_pool_alloc of at least h3dmg_block_bytes).  fresh[b] != 0 (or fresh
 * null): block b's region starts new (p = 0, f = 0, tables uploaded);
 * fresh[b] == 0: the region holds block b from an earlier handle with the
 * same block and parameters, and keeps its p, f and residual.  `stream`: the
 * hipStream_t to launch on (null: a new stream of the handle's own). */
h3dmg_status h3dmg_create_pooled(h3dmg_pool *pool, const h3dmg_block *blocks, int nblocks,
                                 const h3dmg_params *prm, const unsigned long long *offsets,
                                 const int *fresh, void *stream, h3dmg_handle **out);

/* Per-block counters of the stepped solve, carried from one handle to the
 * next: running (1 while its solve goes on), cycles, status (the outcome of
 * its last stepped solve: H3DMG_OK, H3DMG_ERR_NOT_CONVERGED at maxcy,
 * H3DMG_ERR_DIVERGED for a non-finite residual; -1 while running or before
 * the first), enabled, fnorm (||f||_2 of its rhs), rel (last relative
 * residual) and tol (0: the handle's). */
typedef struct {
  int running, cycles, status, enabled;
  double fnorm, rel, tol;
} h3dmg_block_state;
h3dmg_status h3dmg_get_block_state(const h3dmg_handle *h, int b, h3dmg_block_state *out);
h3dmg_status h3dmg_set_block_state(h3dmg_handle *h, int b, const h3dmg_block_state *in);

/* h3dmg_solve one V-cycle at a time, each block on its own clock (REL_L2
 * only; independent blocks).  h3dmg_solve_begin starts the listed blocks
 * (r = f - A p with their current p, as h3dmg_solve starts) while others
 * may be running; h3dmg_solve_step runs one outer iteration (V-cycle,
 * p += pp, new residual) of every running block and stops each block by
 * h3dmg_solve's rule (cycles >= min_cycles and rel <= tol: H3DMG_OK; maxcy
 * cycles: H3DMG_ERR_NOT_CONVERGED; non-finite: H3DMG_ERR_DIVERGED, in the
 * block's state).  *nrunning (may be null) gets the blocks still running.
 * A block's p after its stepped solve is bitwise the one h3dmg_solve gives
 * it alone, whichever blocks ran beside it. */
h3dmg_status h3dmg_solve_begin(h3dmg_handle *h, const int *blocks, int n);
h3dmg_status h3dmg_solve_step(h3dmg_handle *h, int *nrunning);

/* Remove the mean of p over the interior of every block (FP64). */
h3dmg_status h3dmg_remove_mean(h3dmg_handle *h);

/* Stream the handle launches on (a hipStream_t); null = the handle's own, a
 * non-blocking stream that does not wait for work on other streams.  Device
 * pointers passed to the *_device entry points must be ready on the handle's
 * stream when the call is made: synchronize the producer's stream first, or
 * hand the producer's stream to h3dmg_set_stream. */
h3dmg_status h3dmg_set_stream(h3dmg_handle *h, void *stream);

/* ---- tuning (used by the autotuner, hydro3d_amd/autotune.py) -------------
 * Planes each smoother workgroup walks at `level` (every block): 1, 2 or 4
 * (larger requests use 4; the depth in use is the largest power of two not
 * above the request that divides every block's nk at that level); 0 restores
 * the built-in choice.  Results do not depend on it.  Drops a captured graph.
 * h3dmg_get_march returns the depth in use (or -1 for a bad level). */
h3dmg_status h3dmg_set_march(h3dmg_handle *h, int level, int planes);
int h3dmg_get_march(const h3dmg_handle *h, int level);
/* Lanes of a wavefront that cover one row segment of the red-black
 * smoother's tiles at `level` (a power of two up to 64; the tile is 2*lanes
 * colour entries by 4*64/lanes rows): a lower count cuts rows into several
 * tiles and stacks more rows per tile.  Capped at the count that covers a
 * whole row; 0 restores it.  Results do not depend on it.  Drops a captured
 * graph.  h3dmg_get_lanes returns the count in use (-1 for a bad level). */
h3dmg_status h3dmg_set_lanes(h3dmg_handle *h, int level, int lanes);
int h3dmg_get_lanes(const h3dmg_handle *h, int level);
/* Levels whose blocks have at most `max_cells` cells run each red-black
 * sweep as one fused launch (both colour halves, the red halo recomputed
 * per tile) instead of two; 0 (the default) turns fusion off.  Results are
 * identical either way.  Drops a captured graph.  h3dmg_get_march returns 0
 * for a fused level; h3dmg_get_fused returns the threshold in use.  Red-black
 * smoother only (H3DMG_ERR_UNSUPPORTED otherwise). */
h3dmg_status h3dmg_set_fused(h3dmg_handle *h, long long max_cells);
long long h3dmg_get_fused(const h3dmg_handle *h);
/* Levels >= `level` of every block run as one workgroup per block in a
 * single launch with every level in LDS (the V-cycle tail); 0 restores the
 * automatic choice (with H3DMG_COARSE_EXACT the first level, from the
 * shallowest block's coarsest down, whose remaining levels fit in one
 * workgroup's LDS: the tail then runs the blocks' exact coarse solves;
 * otherwise off), a level above every block's coarsest turns the tail off; a level that does not fit returns H3DMG_ERR_ARG.  Results are
 * identical either way.  Drops a captured graph.  h3dmg_get_tail returns the
 * level in use (nlevels_max + 1 when off).  Red-black smoother only: with GS
 * or TDMA the tail is off and a level that would turn it on returns
 * H3DMG_ERR_UNSUPPORTED. */
h3dmg_status h3dmg_set_tail(h3dmg_handle *h, int level);
int h3dmg_get_tail(const h3dmg_handle *h);

/* ---- multi-block layout files (host only) -------------------------------
 * infodom.cin and mdmap.cin as read_infodom / read_mdmap read them
 * (src_v70/alloc_dom.for:284-482; infodom types 1 and 2, mdmap types 1 and
 * 2), with the reference's face-neighbour search (alloc_dom.for:524-690)
 * and periodic-partner search (alloc_dom.for:1040-1215).  Errors carry
 * their text in h3dmg_layout_error(). */
typedef struct h3dmg_layout h3dmg_layout;
h3dmg_status h3dmg_layout_read(const char *infodom_path, const char *mdmap_path,
                               h3dmg_layout **out);
void h3dmg_layout_destroy(h3dmg_layout *L);
const char *h3dmg_layout_error(void);
int h3dmg_layout_ndomains(const h3dmg_layout *L);
int h3dmg_layout_nranks(const h3dmg_layout *L);
/* domain d: its rank (dom_ad), refinement (rdiv) and coordinates */
h3dmg_status h3dmg_layout_domain(const h3dmg_layout *L, int d, int *rank, int *rdiv, double *lo,
                                 double *hi);
/* domain d's neighbours (iprev, inext, jprev, jnext, kprev, knext; -1 none)
 * and periodic partners (per_ip, per_in, per_jp, per_jn, per_kp, per_kn),
 * given the six global boundary codes of control.cin (5 = periodic); either
 * output may be null */
h3dmg_status h3dmg_layout_neighbours(const h3dmg_layout *L, int d, const int *bc_codes,
                                     int *prev_next, int *partner);
/* The blocks of one rank as an h3dmg_create table (localparameters.for:
 * 80-104: cells = extent / the global spacing gdx[3]): shared faces become
 * H3DMG_BC_NEIGHBOUR with local block indices, a periodic axis spanned by
 * one block H3DMG_BC_PERIODIC, the rest walls.  dom_ids (may be null)
 * receives each block's domain number.  With blocks == null only *nblocks
 * is set.  A neighbour on another rank or rdiv > 1 is refused
 * (H3DMG_ERR_UNSUPPORTED). */
h3dmg_status h3dmg_layout_blocks(const h3dmg_layout *L, int rank, const double *gdx,
                                 const int *bc_codes, int nlevels, h3dmg_block *blocks, int cap,
                                 int *nblocks, int *dom_ids);

/* ---- tecbin restart files (host only) ------------------------------------
 * The binary restart file of one block: written by tecbin
 * (src_v70/post.for:324-375), read back for a warm start by the restart
 * branch of src_v70/initial.for:715-747.  Fortran unformatted sequential
 * records (4-byte length, payload, 4-byte length): the dimensions
 * ttc_i, ttc_j, ttc_k; pl; inind, jnind, knind (-1 where the block's east,
 * north or top face is a non-periodic wall); then one record of 23 doubles
 * per cell (k slowest, i fastest): x(i), y(j), z(k) and the 20 fields below
 * at (i, j, k).  Field arrays are Fortran (ttc_i, ttc_j, ttc_k), ghost
 * layers included; x, y, z are the cell coordinates along each axis. */
enum {
  H3DMG_TECBIN_P = 0, H3DMG_TECBIN_PM, H3DMG_TECBIN_PPM, H3DMG_TECBIN_VIS, H3DMG_TECBIN_VISM,
  H3DMG_TECBIN_U, H3DMG_TECBIN_UM, H3DMG_TECBIN_UUM, H3DMG_TECBIN_V, H3DMG_TECBIN_VM,
  H3DMG_TECBIN_VVM, H3DMG_TECBIN_W, H3DMG_TECBIN_WM, H3DMG_TECBIN_WWM, H3DMG_TECBIN_UVM,
  H3DMG_TECBIN_UWM, H3DMG_TECBIN_VWM, H3DMG_TECBIN_KSGS, H3DMG_TECBIN_EPS, H3DMG_TECBIN_EPSM,
  H3DMG_TECBIN_FIELDS
};
typedef struct {
  int tt[3];  /* ttc_i, ttc_j, ttc_k */
  int pl;     /* ghost layers */
  int ind[3]; /* inind, jnind, knind */
} h3dmg_tecbin_header;
h3dmg_status h3dmg_tecbin_write(const char *path, const h3dmg_tecbin_header *hd, const double *x,
                                const double *y, const double *z, const double *const *fields);
h3dmg_status h3dmg_tecbin_read_header(const char *path, h3dmg_tecbin_header *hd);
/* pl >= 0: refuse a file with another ghost-layer count (initial.for:723-728
 * stops); x, y, z, fields and any fields[q] may be null (not read out) */
h3dmg_status h3dmg_tecbin_read(const char *path, int pl, double *x, double *y, double *z,
                               double *const *fields);
const char *h3dmg_tecbin_error(void);

/* ---- test seams: one operator at a time on level arrays ------------------
 * Level arrays cross this seam as dense device arrays in the layout of the
 * oracle: phi padded (n+2)^3 i-fastest, rhs interior.  Ghost cells of an
 * input phi must already be images of the interior (the state every level is
 * in after exchangepp); output ghosts are written as images. */
enum { H3DMG_ARR_PHI = 0, H3DMG_ARR_RHS = 1 };
h3dmg_status h3dmg_dbg_put(h3dmg_handle *h, int b, int level, int which, const double *d_src);
h3dmg_status h3dmg_dbg_get(h3dmg_handle *h, int b, int level, int which, double *d_dst);
/* mgrelax(level, nsweep) on every block */
h3dmg_status h3dmg_dbg_relax(h3dmg_handle *h, int level, int nsweep);
/* mgrestr(level): level is the coarse level */
h3dmg_status h3dmg_dbg_restrict(h3dmg_handle *h, int level);
/* mgcorr(level): level is the fine level */
h3dmg_status h3dmg_dbg_prolong(h3dmg_handle *h, int level);

#ifdef __cplusplus
}
#endif
#endif /* H3DMG_H */
