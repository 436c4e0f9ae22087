/*
 * pepsmg_oracle.c -- CPU restatement of PEPS_MG's geometric-multigrid path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: the parity tests,
 * smoke() and the cpu_baseline leg of bench.py load it; the product
 * (peps_mg_amd / libpepsmg.so) never links, loads or calls it.
 *
 * It follows the reference's loop nests and operation order so that, built
 * with -ffp-contract=off, every stencil, transfer and smoother value is the
 * one the reference computes (1 MPI rank, 1 OpenMP thread).  Each function
 * cites the reference routine it restates (paths relative to the reference
 * root).  Dot products are summed sequentially in k,j,i order, the order of
 * the reference's single-thread reduction.  OpenMP is used only across
 * colour-independent or elementwise loops, so results do not depend on the
 * thread count.
 *
 * Storage: every grid array is the reference's padded (0:nx+1,0:ny+1,0:nz+1)
 * Fortran array, x fastest.  The coefficient matrix is the reference's AoS
 * coeff(0:6,nx,ny,nz).
 *
 * Pinning (see tests/test_oracle_*.py): metrics and coefficients by closed
 * forms; mat-vec against an independent Kronecker-sum Laplacian; the
 * red/black sweep against the textbook SOR matrix splitting of the permuted
 * system; restriction/prolongation against Kronecker products of 1-D
 * transfer matrices and their exactness properties; the coarse solver
 * against dense solves and its stopping rule; CG against textbook CG; the
 * V-cycle against a matrix-algebra V-cycle built from those pinned pieces,
 * and its fixed point against the dense solution.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORA_MAXLEV 32

#define PIDX(i, j, k, nx, ny) \
  ((size_t)(i) + (size_t)((nx) + 2) * ((size_t)(j) + (size_t)((ny) + 2) * (size_t)(k)))
#define CIDX(m, i, j, k, nx, ny) \
  ((size_t)(m) + 7u * ((size_t)((i) - 1) + (size_t)(nx) * ((size_t)((j) - 1) + (size_t)(ny) * (size_t)((k) - 1))))

/* ------------------------------------------------------------------ */
/* Geometry: geometry_domain_create, src/geometry.f90:139-281          */
/* One axis at a time: uniform (a == 1) or symmetric geometric stretch. */
/* ------------------------------------------------------------------ */
void ora_axis_create(int n, double o, double l, double a, double *dm, double *dg, double *g)
{
  int i;
  if (a == 1.0) {                                   /* geometry.f90:177-189 */
    dm[0] = 0.0;
    dg[0] = 0.0;
    g[0] = o;
    for (i = 1; i <= n; i++) {
      dm[i] = l / (double)n;
      dg[i] = 0.5 * (dm[i] + dm[i - 1]);
      g[i] = g[i - 1] + dg[i];
    }
    dm[n + 1] = 0.0;
    dg[n + 1] = 0.5 * (dm[n + 1] + dm[n]);
    g[n + 1] = g[n] + dg[n + 1];
  } else {                                          /* geometry.f90:190-209 */
    dm[0] = 0.0;
    dm[1] = 0.5 * l * (a - 1.0) / (pow(a, (double)(n / 2)) - 1.0);
    for (i = 2; i <= n / 2; i++) {
      dm[i] = dm[i - 1] * a;
      dm[n - i + 1] = dm[i];
    }
    dm[n] = dm[1];
    dm[n + 1] = 0.0;
    dg[0] = 0.0;
    g[0] = o;
    for (i = 1; i <= n; i++) {
      dg[i] = 0.5 * (dm[i] + dm[i - 1]);
      g[i] = g[i - 1] + dg[i];
    }
    dg[n + 1] = 0.5 * (dm[n + 1] + dm[n]);
    g[n + 1] = g[n] + dg[n + 1];
  }
}

/* ------------------------------------------------------------------ */
/* matrix_heptadiagonal_create, src/matrix.f90:39-94                   */
/* ------------------------------------------------------------------ */
void ora_matrix_create(int nx, int ny, int nz,
                       const double *dxm, const double *dxg,
                       const double *dym, const double *dyg,
                       const double *dzm, const double *dzg, double *coeff)
{
  int i, j, k;
  for (k = 1; k <= nz; k++) {
    double dzmp2i = 1.0 / (dzm[k] * dzg[k]);
    double dzmn2i = 1.0 / (dzm[k] * dzg[k + 1]);
    for (j = 1; j <= ny; j++) {
      double dymp2i = 1.0 / (dym[j] * dyg[j]);
      double dymn2i = 1.0 / (dym[j] * dyg[j + 1]);
      for (i = 1; i <= nx; i++) {
        double dxmp2i = 1.0 / (dxm[i] * dxg[i]);
        double dxmn2i = 1.0 / (dxm[i] * dxg[i + 1]);
        double *c = coeff + CIDX(0, i, j, k, nx, ny);
        c[0] = -(dxmp2i + dxmn2i + dymp2i + dymn2i + dzmp2i + dzmn2i);
        c[1] = dxmp2i;
        c[2] = dxmn2i;
        c[3] = dymp2i;
        c[4] = dymn2i;
        c[5] = dzmp2i;
        c[6] = dzmn2i;
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* vv_dot_3d_matrix, src/poisson_matrix_operator.f90:40-99              */
/* Sequential k,j,i order (the single-thread reduction order).          */
/* ------------------------------------------------------------------ */
double ora_vv_dot(const double *x, const double *y, int nx, int ny, int nz)
{
  double s = 0.0;
  int i, j, k;
  for (k = 1; k <= nz; k++)
    for (j = 1; j <= ny; j++) {
      size_t b = PIDX(0, j, k, nx, ny);
      for (i = 1; i <= nx; i++) s = s + x[b + i] * y[b + i];
    }
  return s;
}

/* ------------------------------------------------------------------ */
/* mv_mul_poisson_matrix, src/poisson_matrix_operator.f90:110-161       */
/* y = 0 on the whole padded array, then y = A x on the interior.        */
/* ------------------------------------------------------------------ */
void ora_mv_mul(double *y, const double *coeff, const double *x, int nx, int ny, int nz)
{
  size_t tot = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
  size_t sy = (size_t)(nx + 2), sz = (size_t)(nx + 2) * (ny + 2);
  int k;
  memset(y, 0, tot * sizeof(double));
#pragma omp parallel for schedule(static)
  for (k = 1; k <= nz; k++) {
    int i, j;
    for (j = 1; j <= ny; j++)
      for (i = 1; i <= nx; i++) {
        size_t p = PIDX(i, j, k, nx, ny);
        const double *c = coeff + CIDX(0, i, j, k, nx, ny);
        y[p] = c[0] * x[p] + c[1] * x[p - 1] + c[2] * x[p + 1] + c[3] * x[p - sy] +
               c[4] * x[p + sy] + c[5] * x[p - sz] + c[6] * x[p + sz];
      }
  }
}

/* multigrid_residual, src/multigrid.f90:1782-1807: rsd = rhs - A x, ghosts 0 */
void ora_residual(double *rsd, const double *coeff, const double *x, const double *rhs,
                  int nx, int ny, int nz)
{
  int k;
  ora_mv_mul(rsd, coeff, x, nx, ny, nz);
#pragma omp parallel for schedule(static)
  for (k = 1; k <= nz; k++) {
    int i, j;
    for (j = 1; j <= ny; j++)
      for (i = 1; i <= nx; i++) {
        size_t p = PIDX(i, j, k, nx, ny);
        rsd[p] = rhs[p] - rsd[p];
      }
  }
}

/* ------------------------------------------------------------------ */
/* One red/black colour sweep: src/rbgs_poisson_matrix.f90:239-253 and  */
/* :257-271.  colour 0 starts at ista = 1 + mod(j+k,2) (cells with       */
/* i+j+k odd, 1-based), colour 1 at ista = 1 + mod(j+k+1,2).  One rank:  */
/* offset = 0 (:230-232).                                                */
/* ------------------------------------------------------------------ */
static void ora_rb_colour(double *sol, const double *coeff, const double *rhs,
                          int nx, int ny, int nz, double omega, int colour)
{
  size_t sy = (size_t)(nx + 2), sz = (size_t)(nx + 2) * (ny + 2);
  int k;
#pragma omp parallel for schedule(static)
  for (k = 1; k <= nz; k++) {
    int i, j;
    for (j = 1; j <= ny; j++) {
      int ista = 1 + ((j + k + colour) % 2);
      for (i = ista; i <= nx; i += 2) {
        size_t p = PIDX(i, j, k, nx, ny);
        const double *c = coeff + CIDX(0, i, j, k, nx, ny);
        double temp = c[1] * sol[p - 1] + c[2] * sol[p + 1] + c[3] * sol[p - sy] +
                      c[4] * sol[p + sy] + c[5] * sol[p - sz] + c[6] * sol[p + sz];
        sol[p] = omega * (rhs[p] - temp) / c[0] + (1.0 - omega) * sol[p];
      }
    }
  }
}

/* rbgs_iterator_poisson_matrix, src/rbgs_poisson_matrix.f90:191-279 */
void ora_rbgs_iterator(double *sol, const double *coeff, const double *rhs,
                       int nx, int ny, int nz, int maxiteration, double omega)
{
  int it;
  for (it = 0; it < maxiteration; it++) {
    ora_rb_colour(sol, coeff, rhs, nx, ny, nz, omega, 0);
    ora_rb_colour(sol, coeff, rhs, nx, ny, nz, omega, 1);
  }
}

/* rbgs_solver_poisson_matrix, src/rbgs_poisson_matrix.f90:47-179.
 * Returns the Fortran loop variable `iter` on exit (:109, :167-169):
 * the 0-based iteration at which sqrt(|r|^2/|r0|^2) <= tol held, or
 * maxiteration if it never did.  *rel_out gets the last relative residual. */
int ora_rbgs_solver(double *sol, const double *coeff, const double *rhs,
                    int nx, int ny, int nz, int maxiteration, double tolerance,
                    double omega, double *rel_out)
{
  size_t tot = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
  double *rsd = (double *)calloc(tot, sizeof(double));
  double rsd0tol, rsd_norm;
  int iter;
  ora_residual(rsd, coeff, sol, rhs, nx, ny, nz);                 /* :73-81 */
  rsd0tol = ora_vv_dot(rsd, rsd, nx, ny, nz);                      /* :82 */
  rsd_norm = rsd0tol;
  for (iter = 0; iter < maxiteration; iter++) {
    ora_rb_colour(sol, coeff, rhs, nx, ny, nz, omega, 0);         /* :120-134 */
    ora_rb_colour(sol, coeff, rhs, nx, ny, nz, omega, 1);         /* :138-152 */
    ora_residual(rsd, coeff, sol, rhs, nx, ny, nz);               /* :154-164 */
    rsd_norm = ora_vv_dot(rsd, rsd, nx, ny, nz);                   /* :166 */
    if (sqrt(rsd_norm / rsd0tol) <= tolerance) break;             /* :167 */
  }
  if (rel_out) *rel_out = sqrt(rsd_norm / rsd0tol);
  free(rsd);
  return iter;
}

/* ------------------------------------------------------------------ */
/* Conjugate gradient: src/cg_poisson_matrix.f90                        */
/* ------------------------------------------------------------------ */
static void ora_cg_core(double *sol, const double *coeff, const double *rhs,
                        int nx, int ny, int nz, int maxiteration, double tolerance,
                        int use_tol, int *iters, double *rel_out)
{
  size_t tot = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
  double *p = (double *)calloc(tot, sizeof(double));
  double *ax = (double *)calloc(tot, sizeof(double));
  double *ap = (double *)calloc(tot, sizeof(double));
  double *rsd = (double *)calloc(tot, sizeof(double));
  double alpha, beta, temp1 = 0.0, temp2 = 0.0, rsd0tol = 0.0;
  int iter, k;
  ora_mv_mul(ax, coeff, sol, nx, ny, nz);                          /* :189 / :81 */
#pragma omp parallel for schedule(static)
  for (k = 1; k <= nz; k++) {
    int i, j;
    for (j = 1; j <= ny; j++)
      for (i = 1; i <= nx; i++) {
        size_t q = PIDX(i, j, k, nx, ny);
        rsd[q] = rhs[q] - ax[q];
        p[q] = rsd[q];
      }
  }
  if (use_tol) rsd0tol = ora_vv_dot(rsd, rsd, nx, ny, nz);        /* :94 */
  for (iter = 0; iter < maxiteration; iter++) {
    temp1 = ora_vv_dot(rsd, rsd, nx, ny, nz);                      /* :204 / :106 */
    ora_mv_mul(ap, coeff, p, nx, ny, nz);                          /* :206 */
    temp2 = ora_vv_dot(ap, p, nx, ny, nz);                         /* :207 */
    alpha = temp1 / temp2;                                         /* :209 */
#pragma omp parallel for schedule(static)
    for (k = 1; k <= nz; k++) {                                    /* :211-219 */
      int i, j;
      for (j = 1; j <= ny; j++)
        for (i = 1; i <= nx; i++) {
          size_t q = PIDX(i, j, k, nx, ny);
          sol[q] = sol[q] + alpha * p[q];
          rsd[q] = rsd[q] - alpha * ap[q];
        }
    }
    temp2 = ora_vv_dot(rsd, rsd, nx, ny, nz);                      /* :221 */
    if (use_tol && sqrt(temp2 / rsd0tol) < tolerance) break;       /* :126 */
    beta = temp2 / temp1;                                          /* :223 */
#pragma omp parallel for schedule(static)
    for (k = 1; k <= nz; k++) {                                    /* :225-232 */
      int i, j;
      for (j = 1; j <= ny; j++)
        for (i = 1; i <= nx; i++) {
          size_t q = PIDX(i, j, k, nx, ny);
          p[q] = rsd[q] + beta * p[q];
        }
    }
  }
  if (iters) *iters = iter;
  if (rel_out) *rel_out = use_tol ? sqrt(temp2 / rsd0tol) : sqrt(temp2);
  free(p);
  free(ax);
  free(ap);
  free(rsd);
}

/* cg_iterator_poisson_matrix, src/cg_poisson_matrix.f90:157-241 */
void ora_cg_iterator(double *sol, const double *coeff, const double *rhs,
                     int nx, int ny, int nz, int maxiteration)
{
  ora_cg_core(sol, coeff, rhs, nx, ny, nz, maxiteration, 0.0, 0, NULL, NULL);
}

/* cg_solver_poisson_matrix, src/cg_poisson_matrix.f90:47-146 */
int ora_cg_solver(double *sol, const double *coeff, const double *rhs,
                  int nx, int ny, int nz, int maxiteration, double tolerance, double *rel_out)
{
  int it = 0;
  ora_cg_core(sol, coeff, rhs, nx, ny, nz, maxiteration, tolerance, 1, &it, rel_out);
  return it;
}

/* ------------------------------------------------------------------ */
/* Multigrid hierarchy: src/multigrid.f90                               */
/* ------------------------------------------------------------------ */
typedef struct {
  int nx, ny, nz;
  double *dxm, *dym, *dzm, *dxg, *dyg, *dzg, *xg, *yg, *zg;
  double *x, *b, *r, *coeff;
} ora_level;

typedef struct {
  int n_levels, n_vcycles, aggr;
  int lvc[3];           /* lv_gdm_coarsest_{x,y,z} */
  int smoother;         /* 0: RBGS iterator (reference V-cycle), 1: CG iterator */
  ora_level lv[ORA_MAXLEV + 1];   /* lv[0] = fine level (metrics copied) */
} ora_mg;

/* multigrid_subdomain_CGS_make_grid, src/multigrid.f90:421-495 (1 rank). */
static void ora_make_grid(double *dm, double *dg, double *g, int n, const double *dmf,
                          const double *dgf, const double *gf, double o, int lv_cur, int lv_coarsest)
{
  int i;
  memset(dm, 0, (size_t)(n + 2) * sizeof(double));
  memset(dg, 0, (size_t)(n + 2) * sizeof(double));
  memset(g, 0, (size_t)(n + 2) * sizeof(double));
  if (lv_cur <= lv_coarsest) {
    for (i = 1; i <= n; i++) dm[i] = dmf[2 * i - 1] + dmf[2 * i];
    dm[0] = dmf[0];
    dm[n + 1] = dmf[2 * n + 1];
    g[0] = o - 0.5 * dm[0];
    for (i = 1; i <= n; i++) {
      dg[i] = 0.5 * (dm[i] + dm[i - 1]);
      g[i] = g[i - 1] + dg[i];
    }
    dg[n + 1] = 0.5 * (dm[n + 1] + dm[n]);
    g[n + 1] = g[n] + dg[n + 1];
  } else {
    memcpy(dm, dmf, (size_t)(n + 2) * sizeof(double));
    memcpy(dg, dgf, (size_t)(n + 2) * sizeof(double));
    memcpy(g, gf, (size_t)(n + 2) * sizeof(double));
  }
}

void ora_mg_destroy(ora_mg *mg);

/* multigrid_create (src/multigrid.f90:75-159) with the 1-rank level rules of
 * multigrid_subdomain_create_CGS (:167-402), _CGA (:503-838) and _CGPSA
 * (:998-1288).  Returns 0 or a negative error code:
 *  -1 nlevel <= 1 (:101-107)       -2 bad aggregation method (:120-124)
 *  -3 CGS with CGA_level != 0 (:191-195)   -4 CGA with CGA_level == 0 (:530-534)
 *  -5 odd grid at or before the CGA level (:545-552, :605-612)
 *  -6 too many levels (:307-314, :726-733)  -7 CGPSA level count mismatch (:1186-1193)
 *  -8 more levels than this restatement supports */
int ora_mg_create(ora_mg **out, int nx, int ny, int nz,
                  const double *dxm, const double *dxg, const double *xg,
                  const double *dym, const double *dyg, const double *yg,
                  const double *dzm, const double *dzg, const double *zg,
                  double ox, double oy, double oz,
                  int nlevel, int ncycle, int aggr, int cga_level, int smoother)
{
  ora_mg *mg;
  int l, n[3], lvc[3] = {0, 0, 0}, ax, stopped = 0;
  *out = NULL;
  if (nlevel <= 1) return -1;
  if (aggr < 0 || aggr > 2) return -2;
  if (aggr == 0 && cga_level != 0) return -3;
  if (aggr == 1 && cga_level == 0) return -4;
  if (nlevel > ORA_MAXLEV) return -8;
  mg = (ora_mg *)calloc(1, sizeof(ora_mg));
  mg->n_levels = nlevel;
  mg->n_vcycles = ncycle;
  mg->aggr = aggr;
  mg->smoother = smoother;
  n[0] = nx; n[1] = ny; n[2] = nz;
  for (l = 1; l <= nlevel && !stopped; l++) {
    for (ax = 0; ax < 3; ax++) {
      if (n[ax] % 2 == 0) {
        n[ax] /= 2;
        lvc[ax] = l;
      } else if (aggr == 1 && l <= cga_level) {
        ora_mg_destroy(mg);
        return -5;
      }
    }
    mg->lv[l].nx = n[0]; mg->lv[l].ny = n[1]; mg->lv[l].nz = n[2];
    /* CGA / CGPSA stop creating levels once every axis is odd (:688-691, :1151-1154) */
    if (aggr != 0 && ((long)n[0] * n[1] * n[2]) % 2 == 1) stopped = 1;
  }
  {
    int lvmax = lvc[0] > lvc[1] ? lvc[0] : lvc[1];
    lvmax = lvmax > lvc[2] ? lvmax : lvc[2];
    if (aggr == 2 ? (nlevel != lvmax) : (nlevel > lvmax)) {
      ora_mg_destroy(mg);
      return aggr == 2 ? -7 : -6;
    }
  }
  mg->lvc[0] = lvc[0]; mg->lvc[1] = lvc[1]; mg->lvc[2] = lvc[2];

  /* level 0: the caller's subdomain metrics */
  {
    ora_level *f = &mg->lv[0];
    f->nx = nx; f->ny = ny; f->nz = nz;
    f->dxm = malloc((nx + 2) * sizeof(double)); memcpy(f->dxm, dxm, (nx + 2) * sizeof(double));
    f->dxg = malloc((nx + 2) * sizeof(double)); memcpy(f->dxg, dxg, (nx + 2) * sizeof(double));
    f->xg = malloc((nx + 2) * sizeof(double)); memcpy(f->xg, xg, (nx + 2) * sizeof(double));
    f->dym = malloc((ny + 2) * sizeof(double)); memcpy(f->dym, dym, (ny + 2) * sizeof(double));
    f->dyg = malloc((ny + 2) * sizeof(double)); memcpy(f->dyg, dyg, (ny + 2) * sizeof(double));
    f->yg = malloc((ny + 2) * sizeof(double)); memcpy(f->yg, yg, (ny + 2) * sizeof(double));
    f->dzm = malloc((nz + 2) * sizeof(double)); memcpy(f->dzm, dzm, (nz + 2) * sizeof(double));
    f->dzg = malloc((nz + 2) * sizeof(double)); memcpy(f->dzg, dzg, (nz + 2) * sizeof(double));
    f->zg = malloc((nz + 2) * sizeof(double)); memcpy(f->zg, zg, (nz + 2) * sizeof(double));
  }
  /* per-level metrics (:318-400) and rediscretised matrices (:141-143) */
  for (l = 1; l <= nlevel; l++) {
    ora_level *c = &mg->lv[l], *f = &mg->lv[l - 1];
    size_t tot = (size_t)(c->nx + 2) * (c->ny + 2) * (c->nz + 2);
    c->dxm = malloc((c->nx + 2) * sizeof(double)); c->dxg = malloc((c->nx + 2) * sizeof(double));
    c->xg = malloc((c->nx + 2) * sizeof(double));
    c->dym = malloc((c->ny + 2) * sizeof(double)); c->dyg = malloc((c->ny + 2) * sizeof(double));
    c->yg = malloc((c->ny + 2) * sizeof(double));
    c->dzm = malloc((c->nz + 2) * sizeof(double)); c->dzg = malloc((c->nz + 2) * sizeof(double));
    c->zg = malloc((c->nz + 2) * sizeof(double));
    ora_make_grid(c->dxm, c->dxg, c->xg, c->nx, f->dxm, f->dxg, f->xg, ox, l, lvc[0]);
    ora_make_grid(c->dym, c->dyg, c->yg, c->ny, f->dym, f->dyg, f->yg, oy, l, lvc[1]);
    ora_make_grid(c->dzm, c->dzg, c->zg, c->nz, f->dzm, f->dzg, f->zg, oz, l, lvc[2]);
    c->x = calloc(tot, sizeof(double));
    c->b = calloc(tot, sizeof(double));
    c->r = calloc(tot, sizeof(double));
    c->coeff = malloc((size_t)7 * c->nx * c->ny * c->nz * sizeof(double));
    ora_matrix_create(c->nx, c->ny, c->nz, c->dxm, c->dxg, c->dym, c->dyg, c->dzm, c->dzg, c->coeff);
  }
  *out = mg;
  return 0;
}

/* multigrid_destroy, src/multigrid.f90:1320-1334 */
void ora_mg_destroy(ora_mg *mg)
{
  int l;
  if (!mg) return;
  for (l = 0; l <= ORA_MAXLEV; l++) {
    ora_level *c = &mg->lv[l];
    free(c->dxm); free(c->dxg); free(c->xg);
    free(c->dym); free(c->dyg); free(c->yg);
    free(c->dzm); free(c->dzg); free(c->zg);
    free(c->x); free(c->b); free(c->r); free(c->coeff);
  }
  free(mg);
}

int ora_mg_level_dims(const ora_mg *mg, int l, int *dims)
{
  if (l < 0 || l > mg->n_levels) return -1;
  dims[0] = mg->lv[l].nx; dims[1] = mg->lv[l].ny; dims[2] = mg->lv[l].nz;
  return 0;
}

/* Copies level-l metrics (dxm,dxg for x,y,z concatenated, each n+2) out. */
int ora_mg_level_metrics(const ora_mg *mg, int l, double *out)
{
  const ora_level *c;
  if (l < 0 || l > mg->n_levels) return -1;
  c = &mg->lv[l];
  memcpy(out, c->dxm, (c->nx + 2) * sizeof(double)); out += c->nx + 2;
  memcpy(out, c->dxg, (c->nx + 2) * sizeof(double)); out += c->nx + 2;
  memcpy(out, c->dym, (c->ny + 2) * sizeof(double)); out += c->ny + 2;
  memcpy(out, c->dyg, (c->ny + 2) * sizeof(double)); out += c->ny + 2;
  memcpy(out, c->dzm, (c->nz + 2) * sizeof(double)); out += c->nz + 2;
  memcpy(out, c->dzg, (c->nz + 2) * sizeof(double));
  return 0;
}

void ora_mg_lvc(const ora_mg *mg, int *lvc)
{
  lvc[0] = mg->lvc[0]; lvc[1] = mg->lvc[1]; lvc[2] = mg->lvc[2];
}

/* multigrid_restriction, src/multigrid.f90:1345-1504 (aggregation_type 0
 * sizes; on 1 rank the CGA/CGPSA sizes are the same).  val_c is zeroed
 * whole (:1366), then the volume-weighted 8-child average (:1460-1494). */
void ora_mg_restriction(const ora_mg *mg, double *val_c, const double *val_f, int level)
{
  const ora_level *f = &mg->lv[level], *c = &mg->lv[level + 1];
  int nfx = f->nx, nfy = f->ny, nfz = f->nz, ncx = c->nx, ncy = c->ny, ncz = c->nz;
  double *dxf = calloc(nfx + 2, sizeof(double)), *dyf = calloc(nfy + 2, sizeof(double)),
         *dzf = calloc(nfz + 2, sizeof(double));
  int is, io, js, jo, ks, ko, i, k;
  memset(val_c, 0, (size_t)(ncx + 2) * (ncy + 2) * (ncz + 2) * sizeof(double));
  if (level >= mg->lvc[0]) {                                         /* :1415-1428 */
    memcpy(dxf, f->dxm, (nfx + 2) * sizeof(double)); is = 1; io = 0;
  } else {
    for (i = 1; i <= nfx; i++) dxf[i] = c->dxm[(i - 1) / 2 + 1] - f->dxm[i];
    is = 2; io = 1;
  }
  if (level >= mg->lvc[1]) {                                         /* :1430-1443 */
    memcpy(dyf, f->dym, (nfy + 2) * sizeof(double)); js = 1; jo = 0;
  } else {
    for (i = 1; i <= nfy; i++) dyf[i] = c->dym[(i - 1) / 2 + 1] - f->dym[i];
    js = 2; jo = 1;
  }
  if (level >= mg->lvc[2]) {                                         /* :1445-1458 */
    memcpy(dzf, f->dzm, (nfz + 2) * sizeof(double)); ks = 1; ko = 0;
  } else {
    for (i = 1; i <= nfz; i++) dzf[i] = c->dzm[(i - 1) / 2 + 1] - f->dzm[i];
    ks = 2; ko = 1;
  }
#pragma omp parallel for schedule(static)
  for (k = 1; k <= ncz; k++) {
    int kp = k * ks, kz = kp - ko, j, ii;
    for (j = 1; j <= ncy; j++) {
      int jp = j * js, jz = jp - jo;
      for (ii = 1; ii <= ncx; ii++) {
        int ip = ii * is, iz = ip - io;
        double v111 = dxf[iz] * dyf[jz] * dzf[kz];
        double v211 = dxf[ip] * dyf[jz] * dzf[kz];
        double v121 = dxf[iz] * dyf[jp] * dzf[kz];
        double v221 = dxf[ip] * dyf[jp] * dzf[kz];
        double v112 = dxf[iz] * dyf[jz] * dzf[kp];
        double v212 = dxf[ip] * dyf[jz] * dzf[kp];
        double v122 = dxf[iz] * dyf[jp] * dzf[kp];
        double v222 = dxf[ip] * dyf[jp] * dzf[kp];
        double vol_c = v111 + v211 + v121 + v221 + v112 + v212 + v122 + v222;
        val_c[PIDX(ii, j, k, ncx, ncy)] =
            (v111 * val_f[PIDX(ip, jp, kp, nfx, nfy)] + v211 * val_f[PIDX(iz, jp, kp, nfx, nfy)] +
             v121 * val_f[PIDX(ip, jz, kp, nfx, nfy)] + v221 * val_f[PIDX(iz, jz, kp, nfx, nfy)] +
             v112 * val_f[PIDX(ip, jp, kz, nfx, nfy)] + v212 * val_f[PIDX(iz, jp, kz, nfx, nfy)] +
             v122 * val_f[PIDX(ip, jz, kz, nfx, nfy)] + v222 * val_f[PIDX(iz, jz, kz, nfx, nfy)]) /
            vol_c;
      }
    }
  }
  free(dxf); free(dyf); free(dzf);
}

/* multigrid_prolongation_linear_on_nonuniform_grid,
 * src/multigrid.f90:1516-1771.  Writes the interior of val_f only. */
void ora_mg_prolongation(const ora_mg *mg, double *val_f, const double *val_c, int level)
{
  const ora_level *f = &mg->lv[level], *c = &mg->lv[level + 1];
  int nfx = f->nx, nfy = f->ny, ncx = c->nx, ncy = c->ny, ncz = c->nz;
  double *dxp = calloc(f->nx + 2, sizeof(double)), *dyp = calloc(f->ny + 2, sizeof(double)),
         *dzp = calloc(f->nz + 2, sizeof(double));
  double *dxn = calloc(f->nx + 2, sizeof(double)), *dyn = calloc(f->ny + 2, sizeof(double)),
         *dzn = calloc(f->nz + 2, sizeof(double));
  int is, io, js, jo, ks, ko, i, k;
#define ORA_PW(LVC, STR, OFF, DP, DN, DMC, DMF, NC, NF)                 \
  if (level >= (LVC)) {                                                 \
    STR = 1; OFF = 0;                                                   \
    memcpy(DP, DMC, ((NC) + 2) * sizeof(double));                      \
    memcpy(DN, DMC, ((NC) + 2) * sizeof(double));                      \
  } else {                                                              \
    STR = 2; OFF = 1;                                                   \
    for (i = 1; i <= (NC); i++) {                                       \
      int pf = 2 * i;                                                   \
      DP[pf - 1] = 1.0 * DMF[pf - 2] + 0.5 * DMF[pf - 1];                \
      DP[pf] = 0.5 * DMF[pf - 1];                                       \
      DN[pf - 1] = 0.5 * DMF[pf];                                       \
      DN[pf] = 1.0 * DMF[pf + 1] + 0.5 * DMF[pf];                        \
    }                                                                   \
  }
  ORA_PW(mg->lvc[0], is, io, dxp, dxn, c->dxm, f->dxm, ncx, nfx)         /* :1594-1611 */
  ORA_PW(mg->lvc[1], js, jo, dyp, dyn, c->dym, f->dym, ncy, nfy)         /* :1613-1630 */
  ORA_PW(mg->lvc[2], ks, ko, dzp, dzn, c->dzm, f->dzm, ncz, f->nz)       /* :1632-1649 */
#undef ORA_PW
#define VC(a, b, cc) val_c[PIDX(a, b, cc, ncx, ncy)]
#define VF(a, b, cc) val_f[PIDX(a, b, cc, nfx, nfy)]
#pragma omp parallel for schedule(static)
  for (k = 1; k <= ncz; k++) {                                           /* :1655-1758 */
    int kp = k * ks, kz = kp - ko, km = k - ko, kpc = k + ko, j, ii;
    for (j = 1; j <= ncy; j++) {
      int jp = j * js, jz = jp - jo, jm = j - jo, jpc = j + jo;
      for (ii = 1; ii <= ncx; ii++) {
        int ip = ii * is, iz = ip - io, im = ii - io, ipc = ii + io;
        VF(iz, jz, kz) = (dxp[iz] * dyp[jz] * dzp[kz] * VC(ii, j, k) + dxp[ip] * dyp[jz] * dzp[kz] * VC(im, j, k) +
                          dxp[iz] * dyp[jp] * dzp[kz] * VC(ii, jm, k) + dxp[iz] * dyp[jz] * dzp[kp] * VC(ii, j, km) +
                          dxp[ip] * dyp[jp] * dzp[kz] * VC(im, jm, k) + dxp[ip] * dyp[jz] * dzp[kp] * VC(im, j, km) +
                          dxp[iz] * dyp[jp] * dzp[kp] * VC(ii, jm, km) + dxp[ip] * dyp[jp] * dzp[kp] * VC(im, jm, km)) /
                         ((dxp[iz] + dxp[ip]) * (dyp[jz] + dyp[jp]) * (dzp[kz] + dzp[kp]));
        VF(ip, jz, kz) = (dxn[ip] * dyp[jz] * dzp[kz] * VC(ii, j, k) + dxn[iz] * dyp[jz] * dzp[kz] * VC(ipc, j, k) +
                          dxn[ip] * dyp[jp] * dzp[kz] * VC(ii, jm, k) + dxn[ip] * dyp[jz] * dzp[kp] * VC(ii, j, km) +
                          dxn[iz] * dyp[jp] * dzp[kz] * VC(ipc, jm, k) + dxn[iz] * dyp[jz] * dzp[kp] * VC(ipc, j, km) +
                          dxn[ip] * dyp[jp] * dzp[kp] * VC(ii, jm, km) + dxn[iz] * dyp[jp] * dzp[kp] * VC(ipc, jm, km)) /
                         ((dxn[iz] + dxn[ip]) * (dyp[jz] + dyp[jp]) * (dzp[kz] + dzp[kp]));
        VF(iz, jp, kz) = (dxp[iz] * dyn[jp] * dzp[kz] * VC(ii, j, k) + dxp[ip] * dyn[jp] * dzp[kz] * VC(im, j, k) +
                          dxp[iz] * dyn[jz] * dzp[kz] * VC(ii, jpc, k) + dxp[iz] * dyn[jp] * dzp[kp] * VC(ii, j, km) +
                          dxp[ip] * dyn[jz] * dzp[kz] * VC(im, jpc, k) + dxp[ip] * dyn[jp] * dzp[kp] * VC(im, j, km) +
                          dxp[iz] * dyn[jz] * dzp[kp] * VC(ii, jpc, km) + dxp[ip] * dyn[jz] * dzp[kp] * VC(im, jpc, km)) /
                         ((dxp[iz] + dxp[ip]) * (dyn[jz] + dyn[jp]) * (dzp[kz] + dzp[kp]));
        VF(ip, jp, kz) = (dxn[ip] * dyn[jp] * dzp[kz] * VC(ii, j, k) + dxn[iz] * dyn[jp] * dzp[kz] * VC(ipc, j, k) +
                          dxn[ip] * dyn[jz] * dzp[kz] * VC(ii, jpc, k) + dxn[ip] * dyn[jp] * dzp[kp] * VC(ii, j, km) +
                          dxn[iz] * dyn[jz] * dzp[kz] * VC(ipc, jpc, k) + dxn[iz] * dyn[jp] * dzp[kp] * VC(ipc, j, km) +
                          dxn[ip] * dyn[jz] * dzp[kp] * VC(ii, jpc, km) + dxn[iz] * dyn[jz] * dzp[kp] * VC(ipc, jpc, km)) /
                         ((dxn[iz] + dxn[ip]) * (dyn[jz] + dyn[jp]) * (dzp[kz] + dzp[kp]));
        VF(iz, jz, kp) = (dxp[iz] * dyp[jz] * dzn[kp] * VC(ii, j, k) + dxp[ip] * dyp[jz] * dzn[kp] * VC(im, j, k) +
                          dxp[iz] * dyp[jp] * dzn[kp] * VC(ii, jm, k) + dxp[iz] * dyp[jz] * dzn[kz] * VC(ii, j, kpc) +
                          dxp[ip] * dyp[jp] * dzn[kp] * VC(im, jm, k) + dxp[ip] * dyp[jz] * dzn[kz] * VC(im, j, kpc) +
                          dxp[iz] * dyp[jp] * dzn[kz] * VC(ii, jm, kpc) + dxp[ip] * dyp[jp] * dzn[kz] * VC(im, jm, kpc)) /
                         ((dxp[iz] + dxp[ip]) * (dyp[jz] + dyp[jp]) * (dzn[kz] + dzn[kp]));
        VF(ip, jz, kp) = (dxn[ip] * dyp[jz] * dzn[kp] * VC(ii, j, k) + dxn[iz] * dyp[jz] * dzn[kp] * VC(ipc, j, k) +
                          dxn[ip] * dyp[jp] * dzn[kp] * VC(ii, jm, k) + dxn[ip] * dyp[jz] * dzn[kz] * VC(ii, j, kpc) +
                          dxn[iz] * dyp[jp] * dzn[kp] * VC(ipc, jm, k) + dxn[iz] * dyp[jz] * dzn[kz] * VC(ipc, j, kpc) +
                          dxn[ip] * dyp[jp] * dzn[kz] * VC(ii, jm, kpc) + dxn[iz] * dyp[jp] * dzn[kz] * VC(ipc, jm, kpc)) /
                         ((dxn[iz] + dxn[ip]) * (dyp[jz] + dyp[jp]) * (dzn[kz] + dzn[kp]));
        VF(iz, jp, kp) = (dxp[iz] * dyn[jp] * dzn[kp] * VC(ii, j, k) + dxp[ip] * dyn[jp] * dzn[kp] * VC(im, j, k) +
                          dxp[iz] * dyn[jz] * dzn[kp] * VC(ii, jpc, k) + dxp[iz] * dyn[jp] * dzn[kz] * VC(ii, j, kpc) +
                          dxp[ip] * dyn[jz] * dzn[kp] * VC(im, jpc, k) + dxp[ip] * dyn[jp] * dzn[kz] * VC(im, j, kpc) +
                          dxp[iz] * dyn[jz] * dzn[kz] * VC(ii, jpc, kpc) + dxp[ip] * dyn[jz] * dzn[kz] * VC(im, jpc, kpc)) /
                         ((dxp[iz] + dxp[ip]) * (dyn[jz] + dyn[jp]) * (dzn[kz] + dzn[kp]));
        VF(ip, jp, kp) = (dxn[ip] * dyn[jp] * dzn[kp] * VC(ii, j, k) + dxn[iz] * dyn[jp] * dzn[kp] * VC(ipc, j, k) +
                          dxn[ip] * dyn[jz] * dzn[kp] * VC(ii, jpc, k) + dxn[ip] * dyn[jp] * dzn[kz] * VC(ii, j, kpc) +
                          dxn[iz] * dyn[jz] * dzn[kp] * VC(ipc, jpc, k) + dxn[iz] * dyn[jp] * dzn[kz] * VC(ipc, j, kpc) +
                          dxn[ip] * dyn[jz] * dzn[kz] * VC(ii, jpc, kpc) + dxn[iz] * dyn[jz] * dzn[kz] * VC(ipc, jpc, kpc)) /
                         ((dxn[iz] + dxn[ip]) * (dyn[jz] + dyn[jp]) * (dzn[kz] + dzn[kp]));
      }
    }
  }
#undef VC
#undef VF
  free(dxp); free(dyp); free(dzp); free(dxn); free(dyn); free(dzn);
}

/* multigrid_solve_coarset_level, src/multigrid.f90:1819-1851 */
static int ora_mg_solve_coarsest(ora_mg *mg, double tolerance, double omega)
{
  ora_level *c = &mg->lv[mg->n_levels];
  size_t tot = (size_t)(c->nx + 2) * (c->ny + 2) * (c->nz + 2);
  memset(c->x, 0, tot * sizeof(double));
  if (c->nx * c->ny * c->nz == 1) {
    c->x[PIDX(1, 1, 1, 1, 1)] = c->b[PIDX(1, 1, 1, 1, 1)] / c->coeff[0];
    return 0;
  }
  return ora_rbgs_solver(c->x, c->coeff, c->b, c->nx, c->ny, c->nz, 1000, tolerance, omega, NULL);
}

static void ora_smooth(const ora_mg *mg, double *x, const double *coeff, const double *b,
                       int nx, int ny, int nz, int nu, double omega)
{
  if (mg->smoother == 1)
    ora_cg_iterator(x, coeff, b, nx, ny, nz, nu);
  else
    ora_rbgs_iterator(x, coeff, b, nx, ny, nz, nu, omega);
}

static void ora_axpy_padded(double *x, const double *r, size_t tot)
{
  size_t q;
#pragma omp parallel for schedule(static)
  for (q = 0; q < tot; q++) x[q] = x[q] + r[q];
}

/* multigrid_CGS_vcycle_solver, src/multigrid.f90:1907-1976 (on 1 rank the
 * CGA :1990-2173 and CGPSA :2187-2718 variants compute the same values).
 * coeff0 is the caller's fine matrix (a_poisson).  hist[c] gets the
 * full-precision relative residual after cycle c+1, coarse_it[c] the coarse
 * solver's iteration count.  Returns the number of cycles run. */
int ora_mg_solve_vcycle(ora_mg *mg, double *sol, double *rsd, const double *coeff0,
                        const double *rhs, int maxiteration, double tolerance, double omega,
                        double *hist, int *coarse_it, double *res0_out)
{
  ora_level *f = &mg->lv[0];
  int nx = f->nx, ny = f->ny, nz = f->nz, L = mg->n_levels, cyc, l, ncyc = 0;
  size_t tot0 = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
  double res0tol, rsd_val;
  ora_residual(rsd, coeff0, sol, rhs, nx, ny, nz);                            /* :1924 */
  res0tol = ora_vv_dot(rsd, rsd, nx, ny, nz);                                   /* :1925 */
  if (res0_out) *res0_out = res0tol;
  for (cyc = 1; cyc <= mg->n_vcycles; cyc++) {
    int cit;
    ora_smooth(mg, sol, coeff0, rhs, nx, ny, nz, maxiteration, omega);        /* :1931 */
    ora_residual(rsd, coeff0, sol, rhs, nx, ny, nz);                          /* :1932 */
    ora_mg_restriction(mg, mg->lv[1].b, rsd, 0);                               /* :1933 */
    for (l = 1; l <= L - 1; l++) {                                             /* :1936-1941 */
      ora_level *c = &mg->lv[l];
      size_t tot = (size_t)(c->nx + 2) * (c->ny + 2) * (c->nz + 2);
      memset(c->x, 0, tot * sizeof(double));
      ora_smooth(mg, c->x, c->coeff, c->b, c->nx, c->ny, c->nz, maxiteration, omega);
      ora_residual(c->r, c->coeff, c->x, c->b, c->nx, c->ny, c->nz);
      ora_mg_restriction(mg, mg->lv[l + 1].b, c->r, l);
    }
    cit = ora_mg_solve_coarsest(mg, tolerance, omega);                         /* :1944-1949 */
    if (coarse_it) coarse_it[cyc - 1] = cit;
    for (l = L - 1; l >= 1; l--) {                                             /* :1957-1962 */
      ora_level *c = &mg->lv[l];
      size_t tot = (size_t)(c->nx + 2) * (c->ny + 2) * (c->nz + 2);
      ora_mg_prolongation(mg, c->r, mg->lv[l + 1].x, l);
      ora_axpy_padded(c->x, c->r, tot);
      ora_smooth(mg, c->x, c->coeff, c->b, c->nx, c->ny, c->nz, maxiteration, omega);
    }
    ora_mg_prolongation(mg, rsd, mg->lv[1].x, 0);                              /* :1965 */
    ora_axpy_padded(sol, rsd, tot0);                                           /* :1966 */
    ora_smooth(mg, sol, coeff0, rhs, nx, ny, nz, maxiteration, omega);        /* :1967 */
    ora_residual(rsd, coeff0, sol, rhs, nx, ny, nz);                          /* :1970 */
    rsd_val = ora_vv_dot(rsd, rsd, nx, ny, nz);                                /* :1971 */
    ncyc = cyc;
    if (hist) hist[cyc - 1] = sqrt(rsd_val / res0tol);
    if (sqrt(rsd_val / res0tol) < tolerance) break;                           /* :1973 */
  }
  return ncyc;
}

/* Synthetic right-hand side used by every workload of this build (the
 * reference has no RNG): SplitMix64(seed ^ linear index) -> 53-bit uniform
 * u in [0,1) -> 2u-1, over the padded array's interior, linear index
 * (i-1) + nx*((j-1) + ny*(k-1)).  Returned as raw uniforms; the caller adds
 * the manufactured part and flips the sign. */
static uint64_t ora_splitmix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void ora_uniform_field(double *out, int nx, int ny, int nz, uint64_t seed)
{
  int i, j, k;
  for (k = 1; k <= nz; k++)
    for (j = 1; j <= ny; j++)
      for (i = 1; i <= nx; i++) {
        uint64_t lin = (uint64_t)(i - 1) + (uint64_t)nx * ((uint64_t)(j - 1) + (uint64_t)ny * (uint64_t)(k - 1));
        uint64_t z = ora_splitmix64(seed ^ lin);
        double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
        out[PIDX(i, j, k, nx, ny)] = 2.0 * u - 1.0;
      }
}
