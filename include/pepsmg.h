/*
 * pepsmg.h -- C-ABI of libpepsmg.so, the MI355X implementation of PEPS_MG's
 * geometric-multigrid path (3-D cell-centred 7-point Poisson operator,
 * homogeneous Dirichlet boundaries, 1 process owning the whole grid).
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and
 * returns an int status (PEPSMG_OK = 0, negative on error; the message is
 * available from pepsmg_last_error()).  Host-pointer entry points take the
 * reference's padded Fortran arrays a(0:nx+1,0:ny+1,0:nz+1) (x fastest,
 * contiguous, as allocated by geometry_subdomain_create,
 * src/geometry.f90:415-421) and copy them to and from the GPU; the *_device
 * variants take device pointers in the same padded layout and a HIP stream.
 *
 * Ghost cells of solution arrays must be zero (homogeneous Dirichlet, the
 * only boundary condition the reference's driver sets up); the library
 * checks this and returns PEPSMG_ERR_GHOST otherwise.
 *
 * Matrices: the reference passes its AoS coefficient array
 * coeff(0:6,nx,ny,nz) (src/matrix.f90:27-30).  The GPU path is matrix-free:
 * it recomputes the seven coefficients from the grid metrics with the
 * reference formula (src/matrix.f90:78-88).  When a coeff pointer is given
 * it is compared with that recomputation and PEPSMG_ERR_MATRIX is returned
 * if any coefficient differs.  NULL skips the comparison.
 */
#ifndef PEPSMG_H
#define PEPSMG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PEPSMG_OK = 0,
  PEPSMG_ERR_ARG = -1,        /* invalid argument or size                     */
  PEPSMG_ERR_LEVELS = -2,     /* hierarchy rules of multigrid_create violated */
  PEPSMG_ERR_GHOST = -3,      /* non-zero ghost cell in a solution array      */
  PEPSMG_ERR_MATRIX = -4,     /* coeff differs from the metric rediscretisation */
  PEPSMG_ERR_HIP = -5,        /* HIP runtime error                            */
  PEPSMG_ERR_NOMEM = -6,      /* device allocation failed                     */
  PEPSMG_ERR_NOTCONV = -7,    /* solver stopped at its cap above tolerance (solver entry points only) */
  PEPSMG_ERR_DIVERGED = -8    /* a V-cycle solve's residual norm stopped being finite */
};

/* One subdomain's metrics (type subdomain, src/geometry.f90:59-118): cell
 * counts and the 1-D arrays dxm/dxg (cell size, centre spacing) of each
 * axis, each of length n+2 including the two ghost entries.  Every entry
 * point that takes a grid returns PEPSMG_ERR_ARG unless dxm(1..n) and
 * dxg(1..n+1) are positive and finite (the coefficients divide by them).
 * The reference's stretched-axis formula sets only cells 1..n/2 and their
 * mirrors (src/geometry.f90:190-197): a stretched axis needs an even n. */
typedef struct pepsmg_grid {
  int32_t nx, ny, nz;
  const double *dxm, *dxg;
  const double *dym, *dyg;
  const double *dzm, *dzg;
} pepsmg_grid;

/* Solver options with no counterpart in the reference (its V-cycle always
 * uses RBGS in FP64 with every level launched as a full-grid loop). */
typedef struct pepsmg_options {
  int32_t smoother;        /* 0: red/black Gauss-Seidel/SOR (reference), 1: CG iterator */
  int32_t precision;       /* 0: FP64; 1: FP32 V-cycles inside FP64 iterative refinement */
  int32_t coarse_switch;   /* levels with at most this many cells run inside one
                              persistent workgroup (0 = off); default 4096 (16^3) */
  int32_t inner_cycles;    /* precision 1: FP32 V-cycles per refinement step (default 2) */
  int32_t max_refine;      /* precision 1: cap on refinement steps (default 30) */
  int32_t fused_smoother;  /* 1 (default): single-pass out-of-place red+black kernel, loads issued one
                              plane ahead; 2: the same with loads two planes ahead; 0: two colour sweeps */
  int32_t tile_y;          /* fused smoother tile height 8, 16 or 32 cells (0: per level, 32 on levels
                              of 256 rows or more, 8 below); blocks have
                              max(256, tile_x * tile_y / 4) threads; 32-row tiles load one plane ahead */
  int32_t verbose;         /* 1: print the reference's [MG] cycle lines */
  int32_t tile_x;          /* FP32 fused smoother tile width 64 or 128 cells (0: 128 on levels of 256
                              rows or more, 64 below); FP64 uses 64 */
  int32_t inner_nu;        /* precision 1: smoothing steps of the FP32 V-cycles (default 2; 0: maxiteration) */
  double inner_omega;      /* precision 1: SOR factor of the FP32 V-cycles' red/black smoothing, in (0, 2)
                              (default 1.35; 0: the caller's omega_sor); the FP64 residual and update,
                              which decide convergence, do not use it */
  int32_t inner_nu_post;   /* precision 1: post-smoothing steps of the FP32 V-cycles (default 0: the same
                              as inner_nu, which then counts pre-smoothing steps only) */
} pepsmg_options;

/* Per-solve statistics (the reference only prints these, src/multigrid.f90:1972). */
typedef struct pepsmg_stats {
  int32_t cycles;          /* V-cycles run (FP64) or FP32 V-cycles run (precision 1) */
  int32_t refine_steps;    /* precision 1: refinement steps                      */
  int32_t converged;       /* 1 if rho < tolerance                               */
  int32_t history_cap;     /* capacity of history / coarse_iters (caller-owned)  */
  double res0;             /* |b - A x0|^2                                       */
  double rho;              /* final sqrt(|b - A x|^2 / res0)                     */
  double *history;         /* rho after each V-cycle (FP64) or refinement step   */
  int32_t *coarse_iters;   /* coarsest-level solver iterations per V-cycle       */
  double device_ms;        /* device time of the solve (events on the solve stream) */
} pepsmg_stats;

typedef struct pepsmg_ctx pepsmg_ctx;

/* Last error message of the calling thread ("" if none). */
const char *pepsmg_last_error(void);

/* Default options into *opt. */
void pepsmg_default_options(pepsmg_options *opt);

/* multigrid_create (src/multigrid.f90:75-159): builds the level hierarchy
 * (sizes halve per axis while even; per-level metrics as
 * multigrid_subdomain_CGS_make_grid, :421-495) and allocates device memory.
 * aggr_method 0/1/2 = CGS/CGA/CGPSA; on one process the three give the same
 * values, and the reference's checks on nlevel, CGA_level and level counts
 * (:101-107, :120-124, :191-195, :307-314, :530-534, :545-552, :726-733,
 * :1186-1193) are applied and reported as PEPSMG_ERR_LEVELS. */
int pepsmg_multigrid_create(const pepsmg_grid *grid, int32_t nlevel, int32_t ncycle,
                            int32_t aggr_method, int32_t single_aggr_level,
                            const int32_t adaptive_aggr_level[3], pepsmg_ctx **out);

int pepsmg_set_options(pepsmg_ctx *ctx, const pepsmg_options *opt);

/* multigrid_solve_vcycle (src/multigrid.f90:1864-1894, CGS cycle
 * :1907-1976): V(nu,nu) cycles from the initial guess in sol until
 * sqrt(|r|^2/|r0|^2) < tolerance or ncycle cycles.  On return sol holds the
 * solution and rsd the final residual b - A sol (ghosts zero).  Host arrays. */
int pepsmg_multigrid_solve_vcycle(pepsmg_ctx *ctx, double *sol, double *rsd, const double *coeff,
                                  const double *rhs, int32_t maxiteration, double tolerance,
                                  double omega_sor, pepsmg_stats *stats);

/* Same on device arrays (padded layout) on the given HIP stream (NULL: the
 * context's own stream, ordered after the work already queued on the default
 * stream).  Returns after the stream has finished.  rsd may be NULL, or rhs
 * itself (it is written after the solve). */
int pepsmg_multigrid_solve_vcycle_device(pepsmg_ctx *ctx, double *d_sol, double *d_rsd,
                                         const double *d_rhs, int32_t maxiteration,
                                         double tolerance, double omega_sor, void *stream,
                                         pepsmg_stats *stats);

/* multigrid_destroy (src/multigrid.f90:1320-1334). */
void pepsmg_multigrid_destroy(pepsmg_ctx *ctx);

/* rbgs_iterator_poisson_matrix (src/rbgs_poisson_matrix.f90:191-279):
 * maxiteration red/black iterations in place on sol.  Host arrays. */
int pepsmg_rbgs_iterator(const pepsmg_grid *grid, double *sol, const double *coeff,
                         const double *rhs, int32_t maxiteration, double omega);

/* No counterpart in the reference (its iterator is FP64 only): the red/black
 * iterator of pepsmg_rbgs_iterator in FP32 storage and arithmetic, i.e. the
 * inner smoother of the mixed-precision path (precision 1), launch for launch
 * as the solver runs it (the same tiles, alternating z-march directions).
 * sol and rhs are the padded FP64 host arrays: they are rounded to FP32 on
 * the way in and the FP32 result is widened into sol.  The coefficients are
 * the FP64 metric rediscretisation rounded to FP32; there is no coeff check.
 * For callers that want the FP32 smoother on its own, and to test it
 * directly rather than through the outer refinement that corrects it. */
int pepsmg_rbgs_iterator_fp32(const pepsmg_grid *grid, double *sol, const double *rhs,
                              int32_t maxiteration, double omega);

/* rbgs_solver_poisson_matrix (src/rbgs_poisson_matrix.f90:47-179): iterates
 * until sqrt(|r|^2/|r0|^2) <= tolerance or maxiteration iterations; *iters
 * receives the reference's loop variable on exit, *rel the last ratio. */
int pepsmg_rbgs_solver(const pepsmg_grid *grid, double *sol, const double *coeff,
                       const double *rhs, int32_t maxiteration, double tolerance, double omega,
                       int32_t *iters, double *rel);

/* cg_iterator_poisson_matrix (src/cg_poisson_matrix.f90:157-241). */
int pepsmg_cg_iterator(const pepsmg_grid *grid, double *sol, const double *coeff,
                       const double *rhs, int32_t maxiteration);

/* cg_solver_poisson_matrix (src/cg_poisson_matrix.f90:47-146): stops when
 * sqrt(|r|^2/|r0|^2) < tolerance. */
int pepsmg_cg_solver(const pepsmg_grid *grid, double *sol, const double *coeff,
                     const double *rhs, int32_t maxiteration, double tolerance, int32_t *iters,
                     double *rel);

/* mv_mul_poisson_matrix (src/poisson_matrix_operator.f90:110-161):
 * y = A x on the interior, y's ghosts set to zero. */
int pepsmg_mv_mul(const pepsmg_grid *grid, double *y, const double *coeff, const double *x);

/* vv_dot_3d_matrix (src/poisson_matrix_operator.f90:40-99): interior dot
 * product of two padded arrays. */
int pepsmg_vv_dot(int32_t nx, int32_t ny, int32_t nz, const double *x, const double *y,
                  double *result);

/* No counterpart in the reference: pepsmg_vv_dot for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  x and y are padded device arrays
 * (0:nx+1,0:ny+1,0:nz+1); the interior dot product is written to the device
 * double *d_result on `stream` (NULL: the default stream), without waiting
 * for it.  The sum is bit-identical to pepsmg_vv_dot's for the same arrays.
 * Allocates nothing: its partial sums live in a static 8 KiB device array
 * per GPU (1024 doubles, enough for any grid of at most 2^31-1 rows of
 * cells; larger grids are rejected with PEPSMG_ERR_ARG).  Every call on a
 * GPU shares that array, so calls that overlap on one GPU (other streams,
 * other host threads) race. */
int pepsmg_vv_dot_device_noalloc(int32_t nx, int32_t ny, int32_t nz, const double *d_x,
                                 const double *d_y, double *d_result, void *stream);

/* Measurement hook (no counterpart in the reference): launches the
 * fine-level smoother kernel `reps` times on the context's level-0 arrays
 * (precision 0: FP64, 1: FP32) on the context's stream, timing them with
 * HIP events on that stream.  *avg_ms receives the mean duration of one
 * launch, *bytes the algorithmic HBM bytes of one launch (each input read
 * once, each output written once), *kernel_name the kernel's symbol (a
 * per-thread buffer, valid until the calling thread's next call).  As in a
 * solve, a launch that follows an upward one marches down (the symbol's
 * k_rb_fused_down twin), so the launches alternate between the two. */
int pepsmg_bench_smoother(pepsmg_ctx *ctx, int32_t precision, int32_t reps, double omega,
                          double *avg_ms, double *bytes, const char **kernel_name);

/* Measurement hook (no counterpart in the reference): enable = 1 makes every
 * following solve on ctx bracket the plain red/black smoother launches on its
 * finest level (FP64; in precision 1 the FP32 smoother) with HIP events on
 * the solve stream, and clears the totals; 0 stops.  Up to 256 launches per
 * solve are timed, both upward and downward ones (k_rb_fused_down). */
int pepsmg_kernel_timing(pepsmg_ctx *ctx, int32_t enable);

/* Totals since the last pepsmg_kernel_timing(ctx, 1) (kept when timing is
 * switched off): summed duration of the timed launches (ms), their number,
 * and the kernel's symbol (valid until timing is next switched on). */
int pepsmg_kernel_timing_read(const pepsmg_ctx *ctx, double *total_ms, int32_t *launches,
                              const char **kernel_name);

/* Measurement hook (no counterpart in the reference): the bandwidth a
 * kernel with the smoother's access pattern but no stencil reaches, i.e.
 * y = x/2 + b over n doubles (two streams read, one written, 16 bytes per
 * lane), launched `reps` times on the default stream after `reps` warm-up
 * launches.  Allocates and frees its three arrays.  *avg_ms: mean launch
 * time; *bytes: 3 n 8. */
int pepsmg_bench_stream(int64_t n, int32_t reps, double *avg_ms, double *bytes);

/* Synthetic workload input (no counterpart in the reference, which has no
 * random right-hand side): writes into a padded device array the field
 *   f(i,j,k) = scale_rand * (2u-1) + scale_sin * sin(pi x) sin(pi y) sin(pi z)
 * where u is the 53-bit uniform of SplitMix64(seed ^ lin),
 * lin = (i-1) + nx*((j-1) + ny*(k-1)), and x,y,z are the cell centres
 * xg(i), yg(j), zg(k) (host arrays of length n+2); ghosts are set to zero. */
int pepsmg_fill_rhs_device(int32_t nx, int32_t ny, int32_t nz, double *d_out, uint64_t seed,
                           double scale_rand, double scale_sin, const double *xg,
                           const double *yg, const double *zg, void *stream);

/* The same field into a padded host array (the random part is bit-identical
 * to the device version; sin comes from the host math library). */
int pepsmg_fill_rhs_host(int32_t nx, int32_t ny, int32_t nz, double *out, uint64_t seed,
                         double scale_rand, double scale_sin, const double *xg, const double *yg,
                         const double *zg);

/* ---- Batched solves (no counterpart in the reference, which solves one
 * grid per call).  A batch holds nprob independent problems, problem p with
 * its own grid (grids[p]), number of coarse levels (nlevel[p], the
 * multigrid_create rule of src/multigrid.f90:307-314 with aggregation off)
 * and nrhs[p] right-hand sides; each (problem, right-hand side) pair is an
 * item, numbered problem by problem.  Every V-cycle operation runs as one
 * launch over all items and a solve replays one captured hipGraph per
 * V-cycle (FP64) or refinement step (precision 1) until every item has
 * rho < tolerance.  Options as for a single solve; the batch needs the
 * red/black smoother, the fused smoother and the persistent coarse kernel
 * (coarse_switch > 0, and each item must have a coarse level that fits it). */
typedef struct pepsmg_batch pepsmg_batch;

int pepsmg_batch_create(int32_t nprob, const pepsmg_grid *grids, const int32_t *nlevel, const int32_t *nrhs,
                        int32_t ncycle, pepsmg_batch **out);

int pepsmg_batch_set_options(pepsmg_batch *batch, const pepsmg_options *opt);

/* Number of items (sum of nrhs). */
int32_t pepsmg_batch_size(const pepsmg_batch *batch);

/* Fused-smoother tile shape per problem (arrays of nprob; 0 or NULL: the
 * options' tile_x / tile_y).  Items of different shapes run in separate
 * launches; the autotuner picks the shape per grid size. */
int pepsmg_batch_set_tiles(pepsmg_batch *batch, const int32_t *tile_x, const int32_t *tile_y);

/* Solves every item from its initial guess.  d_sol[i] / d_rhs[i]: device
 * arrays of item i in the padded layout (ghosts zero), overwritten /
 * read.  use_graph 1 replays the captured graph, 0 launches the same
 * sequence directly.  stream NULL: the batch's own stream, ordered after the
 * work already queued on the default stream.  stats: NULL or an array of batch_size entries
 * (history as for a single solve; device_ms is the time of the whole batch). */
int pepsmg_batch_solve_device(pepsmg_batch *batch, double *const *d_sol, const double *const *d_rhs,
                              int32_t maxiteration, double tolerance, double omega_sor, int32_t use_graph,
                              void *stream, pepsmg_stats *stats);

void pepsmg_batch_destroy(pepsmg_batch *batch);

#ifdef __cplusplus
}
#endif

#endif /* PEPSMG_H */
