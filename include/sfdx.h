/*
 * sfdx.h -- C ABI of the MI355X-native SFD_EXCMG forward solve (libsfdx.so).
 *
 * All arrays are caller-owned host memory; the context owns device memory and
 * keeps no host pointer across calls.  Return value 0 = success, > 0 = error
 * (SFDX_ERR_*), < 0 = warning with a valid result (SFDX_WARN_*).  One host
 * thread per context; calls on one context are not re-entrant.
 *
 * Layouts follow the reference exactly (sfd_excmg/src):
 *   grid widths a(nx), b(ny), c(nz) in metres, nz including the nair air layers
 *   on top (file_opt.f90:40-52);
 *   sigma[comp*ncell + ie], comp = xx,yy,zz,xy,xz,yz, ie = (k*ny + j)*nx + i
 *   (Fortran column-major sigma(ncell,6), file_opt.f90:56-71);
 *   edge vectors in the reference edge numbering [Ex | Ey | Ez] with i fastest;
 *   ex[f*2*nl + pol*nl + e] (Fortran ex(2*nl, nf), fdfe_main.f90:1577-1589),
 *   pol 0 = xy (Ex = 1 at the top), pol 1 = yx.
 * Complex numbers are interleaved (re, im) doubles, matching Fortran
 * complex(c_double_complex) and C double _Complex.
 */
#ifndef SFDX_H
#define SFDX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFDX_OK 0
#define SFDX_ERR_ARG 1          /* bad argument / null pointer */
#define SFDX_ERR_MESH 2         /* dimensions not divisible for the EXCMG levels */
#define SFDX_ERR_NAN 3          /* NaN/Inf in inputs or iterates */
#define SFDX_ERR_HIP 4          /* HIP runtime error (message in sfdx_last_error) */
#define SFDX_ERR_DIRECT 5       /* coarse direct solver unavailable or failed */
#define SFDX_ERR_BREAKDOWN 6    /* Krylov breakdown */
#define SFDX_ERR_UNSUPPORTED 7  /* option not supported */
#define SFDX_WARN_NOCONV (-1)   /* maxit reached; last iterate returned */

#define SFDX_SOLVER_BICGSTAB 1  /* solvers.f90:522-719 */
#define SFDX_SOLVER_GPBICG 3    /* solvers.f90:137-315 */
#define SFDX_SOLVER_COCR 5      /* COCR (Sogabe & Zhang 2007) on the symmetric form */

#define SFDX_PREC_FP64 0
#define SFDX_PREC_MIXED 1       /* FP64 Krylov and operator, complex64 preconditioner */
#define SFDX_PREC_MIXED_IR 2    /* complex64 inner Krylov solves inside FP64 refinement.
                                   Experimental: with resistive air (sigma ~1e-8 S/m) the
                                   complex64 product K(G phi) cancels to noise, the inner
                                   solves diverge and the solve returns SFDX_WARN_NOCONV */
#define SFDX_PREC_MIXED_SPLIT 3 /* complex128 solution and residual; COCR directions, their
                                   operator products and the multigrid V-cycle in complex64,
                                   the directions in split edge + nodal form (v_e + (i/wmu)
                                   G v_psi) so that the curl part never acts on gradients.
                                   With precond SFDX_PRECOND_MG and solver COCR on the fine
                                   level; otherwise as SFDX_PREC_MIXED */

#define SFDX_PRECOND_SSOR 0
#define SFDX_PRECOND_HYBRID 1  /* SSOR + G (G^T A G)^-1 G^T, G^T A G = -i w mu P */
#define SFDX_PRECOND_MG 2      /* V-cycle over the EXCMG grid hierarchy, hybrid smoothing */
#define SFDX_PRECOND_AUTO 3    /* per frequency: with mg_merge_ratio > 0 the V-cycle on each Krylov
                                  level unless the skin depth in the most conductive earth cell is
                                  below half that level's smallest cell width; with
                                  mg_merge_ratio = 0 the V-cycle only where its first
                                  coarse grid resolves that skin depth, the skin depth at the median
                                  earth conductivity is within the mesh's horizontal extent and no
                                  cell is more than 200 times longer than wide; the hybrid
                                  preconditioner elsewhere (sfdx_stats.mg_freqs counts the
                                  frequencies whose fine level ran the V-cycle) */

#define SFDX_NORM_REFERENCE 0   /* ||b-Ax||/||b||, b incl. Dirichlet rows (solvers.f90:615-625) */
#define SFDX_NORM_INTERIOR 1    /* interior rows only */

#define SFDX_MAX_LEVELS 8

typedef struct sfdx_ctx sfdx_ctx;

typedef struct {
    int nx, ny, nz;   /* cells; nz counts the air layers */
    int nair;         /* air layers on top */
    const double *a, *b, *c;
} sfdx_grid;

typedef struct {
    int level;          /* EXCMG grid levels (>= 3), .fwd header (file_opt.f90:21) */
    int use_excmg;      /* 1: doublemode_acc, 0: doublemode_nor (global_para.f90:19) */
    int solver;         /* SFDX_SOLVER_* */
    double eps;         /* fine-level tolerance */
    int corr_intv;      /* divergence correction interval, <= 0 disables (global_para.f90:8-9) */
    double corr_tol;    /* PCG tolerance of the correction (global_para.f90:10) */
    double ssor_omega;  /* SSOR relaxation (solvers.f90:14) */
    int precision;      /* SFDX_PREC_* */
    int norm;           /* SFDX_NORM_* */
    int maxit_plain;    /* 5000, fdfe_main.f90:993 */
    int maxit_excmg;    /* 1000, fdfe_main.f90:1414 */
    double inner_tol;   /* mixed precision: relative tolerance of each inner solve */
    int max_refine;     /* mixed precision: maximum refinement steps */
    int verbose;
    int dc_restart;     /* 1: restart the Krylov recurrences after each correction;
                           0: recompute r and keep the search directions (the
                           reference keeps p, v, rho, alpha, omega: solvers.f90:630-646) */
    int dc_at_start;    /* correct the (prolongated) initial guess before iterating */
    int inner_maxit;    /* mixed precision: iteration cap of one inner solve */
    int precond;        /* SFDX_PRECOND_*: 0 multicolour SSOR (the reference's SSOR,
                           re-ordered), 1 SSOR + auxiliary nodal correction,
                           2 geometric multigrid V-cycle with hybrid smoothing,
                           3 (default) 2 or 1 per frequency and level */
    double sigma_floor32; /* complex64 edge smoother only: diagonal conductivities
                             below this floor (S/m) are raised to it, bounding the
                             amplification of near-null air modes; 0 = off */
    int mg_levels;      /* multigrid: grids in the hierarchy counted from the fine
                           grid (0 = as many as the dimensions allow, at least 2).
                           Fewer grids than EXCMG levels end the V-cycle above the
                           coarsest EXCMG level; Krylov levels below its coarsest
                           grid then use the hybrid preconditioner */
    int mg_coarse_pairs; /* multigrid: symmetric smoothing pairs on the coarsest grid */
    double restart_drop; /* SFDX_PREC_MIXED: restart the FP64 Krylov recurrences from
                            the true residual whenever the residual has dropped by
                            this factor since the last restart (0 = never).  Default
                            1e-3: with the hybrid preconditioner it converged at all 16
                            sounding frequencies (1e-3..1e3 Hz) of the 128x128x96
                            model, where 1e-5 and 1e-6 drifted for 8000 iterations at
                            398 Hz; the multigrid V-cycle at 0.1 Hz is fastest at 1e-5 */
    int use_graph;      /* 1: replay the steady-state COCR iteration as a captured HIP
                           graph (only without periodic divergence correction) */
    int nodal_smoother; /* nodal part of the hybrid preconditioner and of the multigrid
                           smoother: 0 8-colour SSOR / Gauss-Seidel on P = G^T M G, 1
                           (default) the two-step damped Jacobi polynomial (one pass over
                           P instead of 8 dependent colour sweeps per smoothing step) */
    int coarse_solver;  /* EXCMG levels 1-2: 1 (default) Krylov solves on the device from
                           a zero guess to eps*10^(i-L); 0 host direct solve (MKL
                           PARDISO, the reference's u1u2).  At 128x128x96 the two host
                           solves cost 0.7 s of a 1.7 s forward solve, the device
                           ones 0.065 s */
    int full_background;  /* 0: the reference's 1-D boundary (diagonal of cell column (1,1),
                             cross-polar boundary components 0; fdfe_main.f90:938-942,
                             470-494); 1: full tensor of that column, with the cross-polar
                             components of the 1-D field on the boundary (Ez stays 0) */
    int spmv_tile;      /* SpMV kernel of the Krylov levels: 0 (default) in complex128 the
                           cell-plane march (32x8 node columns, planes loaded straight into
                           LDS), in complex64 the k-marching 32x8 tile with prefetch; then
                           the k-marching tiles 1 32x8 without prefetch, 2 32x4 with, 3 32x4
                           without, 4 32x8 with prefetch; -1 autotune: time the five once per
                           grid shape and precision on the device and keep the fastest
                           (cached in the context).  The kernel sets the grouping of the
                           fused reductions, so a fixed choice keeps repeated runs
                           bit-identical */
    int ssor_cross_mass;  /* edge SSOR inside the hybrid preconditioner and the multigrid
                             smoother: 0 (default) K plus the lumped (diagonal) mass -- the
                             cross-orientation mass couplings of anisotropic cells are left out,
                             so each colour sweep reads one conductivity component instead of
                             three; 1: the full operator.  The plain SSOR preconditioner (the
                             reference's) always uses the full operator, and so do the
                             multigrid residuals and every operator application */
    double surface_sigma; /* sfdx_fwd_response and sfdx_adjoint: 0 (default) take the station
                             fields on the flat datum, node plane nair (fdfe_main.f90:597-640);
                             > 0: topography carved into sigma -- the surface of each cell column
                             is the top node plane of its first cell, from the top, whose sigma_xx
                             exceeds this value in S/m (as sfdx_surface_planes; the reference's
                             topography branch, fdfe_main.f90:604-614, 656-666) */
    int pol_streams;      /* Krylov levels (and the adjoint solves of a one-frequency sfdx_adjoint):
                             2 (default) solve the two polarisations at the same time,
                             the second on a helper HIP stream with its own work vectors (the
                             level data is shared and only read; results are bit-identical to
                             the sequential solves); 0 or 1: one after the other on the context's
                             stream.  sfdx_solve_batch runs its items with 1 */
    double mg_merge_ratio; /* grids of the multigrid V-cycle: r >= 1 (default 3) builds them per
                              Krylov level, merging two consecutive cells of an axis only where
                              both are at most r times the grid's smallest width (never across
                              the air-earth interface), so the long cells of stretched padding
                              stay unmerged until the short ones have grown to their size;
                              mg_levels then caps the number of grids (0: down to a grid of at
                              most 2048 cells, up to 12 grids); a larger coarsest grid gets a
                              smoothing pair per 2048 cells (at least mg_coarse_pairs).  0: the
                              EXCMG levels and 2:1 coarsenings of the fine grid */
} sfdx_opts;

typedef struct {
    int status;
    double total_seconds;
    double fine_seconds;            /* fine-level Krylov time, all frequencies and both polarisations */
    double fine_bytes;              /* algorithmic HBM bytes of the fine-level Krylov kernels (stencil,
                                       vector updates) and of an SSOR or hybrid preconditioner's edge
                                       SSOR; the multigrid V-cycle is not counted */
    int iters[SFDX_MAX_LEVELS][2];  /* last frequency: Krylov iterations per level and polarisation */
    double relres[SFDX_MAX_LEVELS][2];  /* final relative residual of each Krylov solve, recomputed from
                                           the returned iterate (b - A x), not the recurrence residual */
    int dc_calls[SFDX_MAX_LEVELS][2];
    int pcg_iters[SFDX_MAX_LEVELS][2];
    int refine_steps[2];            /* fine level: restarts (SFDX_PREC_MIXED) or refinement steps */
    int fine_iters_total;           /* summed over frequencies and polarisations */
    int fallback_fp64;              /* solves (any level, frequency, polarisation) that switched from
                                       the complex64 to the complex128 preconditioner after repeated
                                       restarts without progress */
    int solver_switches;            /* COCR solves that stagnated (restarts without progress) and
                                       were continued by GPBiCG from the current iterate */
    int spmv_tile;                  /* SpMV tile used on the fine level (complex128 operator) */
    int mg_freqs;                   /* frequencies whose fine level ran the multigrid V-cycle */
} sfdx_stats;

/* Context: device selection, stream, workspace. */
int sfdx_create(int device, sfdx_ctx **out);
int sfdx_destroy(sfdx_ctx *ctx);
const char *sfdx_last_error(void);
void sfdx_default_opts(sfdx_opts *o);

/* Replaces fwd3d::fwd_solver (sfd_excmg/src/fdfe_main.f90:1565-1594): all
 * frequencies, both polarisations; ex caller-allocated with 2*nl*nf entries.
 * With several frequencies and pol_streams = 2 (the default) the frequencies
 * run as concurrent items, as sfdx_solve_batch over this one model (identical
 * fields; st then sums the items' times); pol_streams 0 or 1 solves them one
 * after the other on the context's stream. */
int sfdx_fwd_solve(sfdx_ctx *ctx, const sfdx_grid *g, const double *sigma, int nf, const double *freq,
                   const sfdx_opts *o, double *ex, sfdx_stats *st);

/* Replaces fwd3d::post_process (sfd_excmg/src/fdfe_main.f90:549-716) and
 * arho_interp2 (:718-790).  recv(nrec,3) row-major (x,y,z); results and tipper
 * (nrec*nf, 4) row-major.  out_type 1: Re/Im Zxy, Re/Im Zyx; 2: rho_xy, rho_yx,
 * phi_xy, phi_yx.  rd = 2 reproduces the reference's station sampling
 * (arho_interp2 called with rd=2), rd = 1 samples at the station's own cell. */
int sfdx_post_process(const sfdx_grid *g, int nf, const double *freq, int nrec, const double *recv,
                      const double *ex, int out_type, int rd, double *results, double *tipper);

/* Replaces the topography branch of post_process (sfd_excmg/src/fdfe_main.f90:
 * 549-716 with topo present, lines 604-614 and 656-666): as sfdx_post_process,
 * but the surface of cell column (i,j) is node plane surf[i + nx*j] (0-based,
 * counted from the top of the air, 0 <= surf < nz); surf = NULL is the flat
 * datum nair.  The fields there are averaged to the cell centre and the z
 * difference spans the cell below the surface plane.  The reference selects
 * the first cell k1 whose centre lies below the ground and uses node plane k1
 * (the bottom of that cell) with the flat-datum spacing c(nair+1); here the
 * top of that cell, k1-1, and its own thickness are used, so that a flat
 * ground reproduces the flat-datum result. */
int sfdx_post_process_topo(const sfdx_grid *g, const int *surf, int nf, const double *freq, int nrec,
                           const double *recv, const double *ex, int out_type, int rd, double *results, double *tipper);

/* Surface node plane of every cell column of a model with topography carved
 * into sigma (cells above the ground set to air): surf[i + nx*j] = top node
 * plane of the column's first cell, from the top, whose sigma_xx exceeds
 * air_max (S/m).  Returns SFDX_ERR_ARG if a column is air throughout.  Takes
 * the place of the topography elevations the reference reads
 * (file_opt.f90:97-128) and compares with cell centres (fdfe_main.f90:606-611). */
int sfdx_surface_planes(const sfdx_grid *g, const double *sigma, double air_max, int *surf);

/* Replaces mesh_coarsen (sfd_excmg/src/fdfe_main.f90:18-84): widths summed over
 * 2^k fine cells, sigma the volume-weighted average over each 2^k x 2^k x 2^k
 * block (with the fine-cell index taken inside the block; the reference's
 * index runs one layer and one row too far).  Outputs a_c(nx/2^k),
 * b_c(ny/2^k), c_c(nz/2^k), sigma_c(ncell/8^k, 6) in the sigma layout above. */
int sfdx_mesh_coarsen(const sfdx_grid *g, const double *sigma, int k, double *a_c, double *b_c, double *c_c,
                      double *sigma_c);

/* The same coarsening with sigma averaged on the device (k_mesh_coarsen, one
 * coarse cell per thread, the host's order of operations: bit-identical to
 * sfdx_mesh_coarsen).  sfdx_fwd_solve coarsens its levels this way from the
 * resident fine sigma.  Replaces mesh_coarsen (sfd_excmg/src/fdfe_main.f90:
 * 18-84) on the device. */
int sfdx_mesh_coarsen_device(sfdx_ctx *ctx, const sfdx_grid *g, const double *sigma, int k, double *a_c,
                             double *b_c, double *c_c, double *sigma_c);

/* Replaces mt1d_anl (sfd_excmg/src/model_selec.f90:266-401): conv(nz,6)
 * column-major; ex, hx: 2*nz+2 complex; imp: 4 complex. */
int sfdx_mt1d_anl(int nz, const double *conv, const double *c, double freq, double *ex, double *hx, double *imp);

/* One model of a batched solve: grid, conductivity sigma(ncell,6), its nf
 * frequencies and the caller-allocated ex(2*nl, nf) it receives (same layouts
 * as sfdx_fwd_solve).  level > 0 overrides opts.level for this model. */
typedef struct {
    sfdx_grid grid;
    const double *sigma;
    int nf;
    const double *freq;
    double *ex;
    int level;
} sfdx_job;

/* Batched forward solves: every (model, frequency) pair of `jobs` (both
 * polarisations each) is an independent work item; items run concurrently on
 * `nstreams` HIP streams (0 = 6, at most 16), longest first (larger grids, then
 * higher frequencies), each stream with
 * its own workspace, so small uneven models fill the GPU together.  Results are
 * identical to calling sfdx_fwd_solve per model.  st (optional) receives one
 * sfdx_stats per job (times and iterations summed over its frequencies).
 * Replaces the serial frequency loop of fwd_solver (fdfe_main.f90:1580-1592)
 * repeated over models; the reference has no batched entry point. */
int sfdx_solve_batch(sfdx_ctx *ctx, int njobs, const sfdx_job *jobs, const sfdx_opts *o, int nstreams,
                     sfdx_stats *st);

/* Forward solve and station responses without moving the fields: after each
 * frequency's fine solve the surface E and H (cell centres of node plane nair)
 * are evaluated on the device and only they are copied to the host, where the
 * station interpolation and Z / rho_a / phi / tipper of post_process follow
 * (same results, out_type and rd as sfdx_post_process).  Replaces the
 * fwd_solver + post_process pair of the driver (fdfe_main.f90:1565-1594,
 * 549-716) when the fields themselves are not needed. */
int sfdx_fwd_response(sfdx_ctx *ctx, const sfdx_grid *g, const double *sigma, int nf, const double *freq,
                      const sfdx_opts *o, int nrec, const double *recv, int out_type, int rd, double *results,
                      double *tipper, sfdx_stats *st);

/* Data misfit of the fields ex(2*nl, nf) (host only, FP64):
 *   misfit = 1/2 sum_{f, station, xy|yx} (log10 rho_a - log10 rho_obs)^2 + (phi - phi_obs)^2
 * with phi in radians and rho_a, phi computed as post_process does (rd as
 * there); obs holds nf*nrec rows (rho_xy, rho_yx, phi_xy, phi_yx) in ohm-m and
 * degrees, i.e. post_process out_type 2 output of the observed data.  If w is
 * non-null it receives the field gradient (2*nl*nf complex) with
 * d misfit = Re(w^T d ex).  No counterpart in the reference (inversion support). */
int sfdx_misfit(const sfdx_grid *g, int nf, const double *freq, int nrec, const double *recv, int rd,
                const double *obs, const double *ex, double *misfit, double *w);

/* Misfit and its gradient with respect to log(s_c), s_c a scale factor on the
 * conductivity tensor of every cell c (for isotropic sigma: d/d log sigma_c).
 * Per frequency: forward solve (as sfdx_fwd_solve), field gradient
 * (sfdx_misfit), adjoint solves A_s^T lam = w with the forward kernels (A_s is
 * complex symmetric) and homogeneous Dirichlet data, then
 * grad_c += Re(i omega mu lam^T M_c e) summed over both polarisations, M_c the
 * cell's corner-lumped edge mass.  The frequencies run as independent items
 * on up to 6 HIP streams at once (higher frequencies first; each keeps its
 * fields and adjoint solutions on the device, 4*nl complex128 per frequency),
 * and the gradient is summed afterwards in frequency and polarisation order,
 * so the result does not depend on the schedule.  The Dirichlet data of the 1-D background are held fixed
 * (their dependence on the corner column, fdfe_main.f90:938-942, is not
 * differentiated).  grad: ncell doubles; ex (optional): the forward fields
 * 2*nl*nf; st (optional): forward-solve statistics.  No counterpart in the
 * reference. */
int sfdx_adjoint(sfdx_ctx *ctx, const sfdx_grid *g, const double *sigma, int nf, const double *freq, int nrec,
                 const double *recv, int rd, const double *obs, const sfdx_opts *o, double *misfit, double *grad,
                 double *ex, sfdx_stats *st);

/* Direct solve of one grid level for both polarisations on the host (MKL
 * PARDISO, complex symmetric): replaces doublemode_nor with mid = 1, 2
 * (fdfe_main.f90:997-998, 1028-1029) -> pardiso_spd (solvers.f90:765-875).
 * u: 2*nl complex, xy then yx, Dirichlet values included. */
int sfdx_direct_solve(const sfdx_grid *g, const double *sigma, double freq, double *u);

/* Kernel-level entry points (parity tests and benchmarks).
 * y = A_s x: the dual-volume scaled operator of fd_assemb (fdfe_main.f90:226-396,
 * rows times ve, lines 252/301/350) on interior edges, 0 on boundary edges;
 * replaces zamux (div_correc.f90:431-449) on the dcfast CSR. */
int sfdx_apply_operator(sfdx_ctx *ctx, const sfdx_grid *g, const double *sigma, double freq, int precision,
                        const double *x, double *y);
/* Allocation-free form of sfdx_apply_operator (complex128): y = A_s x on
 * caller-owned DEVICE memory.  No counterpart in the reference; it is for
 * callers that cannot allocate device memory -- code under HIP graph capture,
 * tight serving loops.  g->a, g->b, g->c (widths, nx/ny/nz doubles), sigma
 * (sigma[comp*ncell + ie], 6*ncell doubles), x and y (nl complex doubles)
 * are device pointers on the context's GPU; the call only enqueues work on
 * `stream` (a hipStream_t; NULL = the context's stream) and returns without
 * waiting.  It calls no allocator: the reciprocal widths live in a fixed
 * library scratch array of 3 x 4096 doubles per GPU (so nx, ny, nz <= 4096,
 * else SFDX_ERR_ARG), whose address is fetched by sfdx_create.  Every call on
 * one GPU shares that array, so calls that overlap on one GPU (other streams,
 * other contexts or host threads) race and must be serialised by the caller. */
int sfdx_apply_operator_noalloc(sfdx_ctx *ctx, const sfdx_grid *g, const double *sigma, double freq, const double *x,
                                double *y, void *stream);
/* t = r - A_s (z + (i/(omega mu)) G psi) on interior edges, 0 on boundary
 * edges: the residual of a correction held in the split edge + nodal form of
 * the multigrid V-cycle preconditioner, with G the node-to-edge gradient.  It
 * is formed as r - K z + i omega mu M (z + (i/(omega mu)) G psi) (K the curl
 * part, M the mass part of A_s; K G = 0), so the curl part never acts on
 * G psi.  r, z, t: nl complex; psi: (nx+1)(ny+1)(nz+1) complex node values.
 * precision SFDX_PREC_FP64 complex128, else complex64.  No counterpart in the
 * reference (whose solvers have no multigrid); exposed for parity tests. */
int sfdx_split_residual(sfdx_ctx *ctx, const sfdx_grid *g, const double *sigma, double freq, int precision,
                        const double *r, const double *z, const double *psi, double *t);
/* y = M^-1 x, multicolour SSOR of A_s (replaces ssorfac + lusol,
 * div_correc.f90:171-220, solvers.f90:721-762); x must vanish on boundary edges. */
int sfdx_ssor_apply(sfdx_ctx *ctx, const sfdx_grid *g, const double *sigma, double freq, double omega,
                    int precision, const double *x, double *y);
/* y = M_h^-1 x, the hybrid (auxiliary-space) preconditioner of the device
 * solvers with their default edge SSOR (lumped mass, ssor_cross_mass = 0):
 * SSOR_e(x) + (i/(omega mu)) G N_P(G^T x), with N_P the two-step Jacobi
 * polynomial (nodal 1) or the 8-colour SSOR (nodal 0) on P = G^T M_sigma G.
 * No counterpart in the reference (whose preconditioner is the SSOR alone,
 * div_correc.f90:171-220); exposed so that callers and tests can apply the
 * exact operator the solvers use.  precision SFDX_PREC_MIXED: the complex64
 * edge SSOR with the complex128 nodal part (nodal must be 1).  x must vanish
 * on boundary edges. */
int sfdx_hybrid_apply(sfdx_ctx *ctx, const sfdx_grid *g, const double *sigma, double freq, double omega,
                      int nodal, int precision, const double *x, double *y);
/* In-place divergence correction of an edge field (replaces div_correc,
 * div_correc.f90:11-147).  pcg_iters may be NULL. */
int sfdx_div_correct(sfdx_ctx *ctx, const sfdx_grid *g, const double *sigma, double *x, double corr_tol,
                     int *pcg_iters);
/* EXCMG initial guess on grid g from the two coarser levels (replaces
 * extint_v, excmg.f90:9-77): u0 on g/4, u1 on g/2, out on g. */
int sfdx_extint_v(sfdx_ctx *ctx, const sfdx_grid *g, const double *u0, const double *u1, double *out);
/* Times the fine-level kernels on the context's stream with HIP events,
 * averaged over reps: ms[0] SpMV, ms[1] SSOR apply (12 colour sweeps),
 * ms[2] COCR vector update, ms[3] the same bytes as ms[4] moved by a
 * bandwidth-tuned streaming copy (several 16-byte loads in flight per thread,
 * non-temporal stores: the attainable streaming rate), ms[4] plain copy of
 * one edge vector, ms[5] hybrid preconditioner apply, ms[6] SpMV in the
 * one-row-per-thread gather form, ms[7..10] the SpMV variants 0-3 (0: the
 * default, in complex128 the cell-plane march, in complex64 the k-marching
 * tile of 32x8 columns with prefetch; 1: 32x8 without prefetch; 2, 3: 32x4
 * with and without), ms[11] the k-marching 32x8 tile over 4-plane chunks,
 * ms[12] the k-marching 32x8 tile with prefetch (variant 4), ms[13] the
 * multigrid V-cycle's split residual r - A (z + (i/wmu) G psi).  ms holds 14
 * doubles.
 * No counterpart in the reference (measurement only). */
int sfdx_bench_kernels(sfdx_ctx *ctx, const sfdx_grid *g, const double *sigma, double freq, int precision,
                       int reps, double *ms);

#ifdef __cplusplus
}
#endif
#endif
