/*
 * oracle.c -- CPU FP64 restatement of libcfd2lcs's structured-grid FTLE path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  See oracle.h for the scope and the
 * map to the reference sources.  One rank, double precision, rectilinear
 * grids, resolution 0.  Particles are kept in origin-node order: the
 * reference reorders them on every exchange (comms_m.f90:1135-1141) but all
 * per-particle work is independent and results are keyed by the origin node
 * (no0), so the order does not change any result.
 *
 * Index conventions: Fortran index i in [1-ng, ni+ng] with ng = 1 is stored at
 * array position i, i.e. ghosted arrays are (ni+2)*(nj+2)*(nk+2), i fastest.
 * Particle p (0-based) is the node lcs_ijk2l(i,j,k)-1 (data_m.f90:392-396).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---- constants shared with the reference headers (cfd2lcs_inc_double.h:27-65) */
enum { LCS_INTERNAL = 0, LCS_INFLOW = 1, LCS_OUTFLOW = 2, LCS_WALL = 3,
       LCS_SLIP = 4, LCS_2D = 5, LCS_MASK = 6 };
enum { FTLE_FWD = 0, FTLE_BKWD = 1, LP_TRACER = 2 };
enum { EULER = 0, TRAPEZOIDAL = 1, RK2 = 2, RK3 = 3, RK4 = 4 };
enum { NEAREST_NBR = 0, LINEAR = 1, QUADRATIC = 2, CUBIC = 3, TSE = 5,
       TSE_LIMIT = 6 };
enum { LCS_FALSE = 0, LCS_TRUE = 1 };
/* data_m.f90:110-119 */
enum { DIR_FWD = 1, DIR_BKWD = -1 };
enum { LP_IB = -1, LP_UNKNOWN = -2, LP_RECYCLE = -3, LP_STICK = -4 };
#define NMAX_STRUCT 1000

/* 27-neighbour order: self, 6 faces, then edges/corners (data_m.f90:123-152).
 * The order fixes the tie-breaks of the nearest-node search. */
static const int NBR[27][3] = {
    {0, 0, 0},   {-1, 0, 0},  {0, -1, 0}, {0, 0, -1}, {1, 0, 0},
    {0, 1, 0},   {0, 0, 1},   {-1, -1, -1}, {0, -1, -1}, {1, -1, -1},
    {-1, 0, -1}, {1, 0, -1},  {-1, 1, -1}, {0, 1, -1}, {1, 1, -1},
    {-1, -1, 0}, {1, -1, 0},  {-1, 1, 0}, {1, 1, 0},   {-1, -1, 1},
    {0, -1, 1},  {1, -1, 1},  {-1, 0, 1}, {1, 0, 1},   {-1, 1, 1},
    {0, 1, 1},   {1, 1, 1}};

#define DI(i, j, k) (((i) + 1) + 3 * ((j) + 1) + 9 * ((k) + 1))

static int imax3(int a, int b, int c) { int m = a > b ? a : b; return m > c ? m : c; }
static int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
/* Fortran max(min(r,1),0) with IEEE minNum/maxNum semantics: a NaN ratio
 * (first update, t_n == t_np1, lp_motion_m.f90:170) gives 1. */
static double clamp01(double r) { return fmax(fmin(r, 1.0), 0.0); }
/* Fortran NINT: round half away from zero */
static int fnint(double v) { return (int)(v >= 0 ? floor(v + 0.5) : -floor(-v + 0.5)); }

/* ======================================================================
 * Grid (sgrid_t, data_m.f90:279-312) on one rank
 * ====================================================================== */
typedef struct {
  int n[3];      /* ni, nj, nk */
  int N[3];      /* ghosted extents n+2 */
  long nt;       /* ghosted size */
  long np;       /* interior size */
  double *g[3];  /* ghosted coordinates (with ghosts/fakes) */
  int *flag;     /* ghosted bcflag */
  double *bn[3]; /* bcnorm (interior valid) */
  double *dl[3]; /* 1/delta (interior valid), compute_delta_xyz */
  int per[3];    /* periodic_i/j/k */
  double ps[27][3]; /* periodic_shift(i,j,k,1:3) */
  double bb[6];  /* bounding box (ghost inclusive) */
  int rect;
  double *lsw;   /* non-rectilinear grids: face least-squares gradient weights,
                    [(f * 3 + d) * nt + l] for face f = NBR 1..6, interior l */
  double pmin[3][3], pmax[3][3]; /* minloc / maxloc of x, y, z over the ghosted
                                    arrays as 1-based array positions */
} ogrid;

#define GID(g, i, j, k) \
  ((long)(i) + (long)(g)->N[0] * ((long)(j) + (long)(g)->N[1] * (long)(k)))
#define PIDG(g, i, j, k) \
  ((long)((i) - 1) + (long)(g)->n[0] * ((long)((j) - 1) + (long)(g)->n[1] * (long)((k) - 1)))

/* A grid with its interpolation lines and spline work arrays: the flow grid
 * (scfd%sgrid), or the grid of a diagnostic's particle set when that is not
 * the flow grid (resolution != 0, new_sgrid_from_sgrid, sgrid_m.f90:445-761;
 * auxiliary grids, create_aux_grid, lcs_m.f90:227-435). */
typedef struct {
  ogrid g;
  double *xl[3];  /* coordinate lines with 2-D overrides (lp_tracking_m.f90:362-378) */
  double *bst[3]; /* spline knots */
  double *bsc[3]; /* spline coefficients */
} ofg;

/* single-rank SELF_COMM test for direction (i,j,k) (comms_m.f90:258-283) */
static int self_comm(const ogrid *g, int i, int j, int k) {
  if (i == 0 && j == 0 && k == 0) return 0;
  if (i != 0 && !g->per[0]) return 0;
  if (j != 0 && !g->per[1]) return 0;
  if (k != 0 && !g->per[2]) return 0;
  return 1;
}

/* exchange_sdata on one rank (comms_m.f90:679-869): every ghost region whose
 * direction is SELF_COMM receives the interior layer on the opposite side.
 * face_only restricts to the 6 face regions (FACE_CONNECT, ncomm = 7).
 * For grid coordinates (translate = component 0,1,2) the pack adds
 * periodic_shift of the pack direction; the z component reads
 * periodic_shift(i,k,k,3) (comms_m.f90:750). */
static void exchange(const ogrid *g, double *f, int translate,
                     double ps[27][3], int face_only) {
  const int ni = g->n[0], nj = g->n[1], nk = g->n[2];
  for (int k = 0; k <= nk + 1; k++)
    for (int j = 0; j <= nj + 1; j++)
      for (int i = 0; i <= ni + 1; i++) {
        int ei = i < 1 ? -1 : (i > ni ? 1 : 0);
        int ej = j < 1 ? -1 : (j > nj ? 1 : 0);
        int ek = k < 1 ? -1 : (k > nk ? 1 : 0);
        if (ei == 0 && ej == 0 && ek == 0) continue;
        if (face_only && (abs(ei) + abs(ej) + abs(ek)) != 1) continue;
        if (!self_comm(g, ei, ej, ek)) continue;
        int si = ei < 0 ? ni : (ei > 0 ? 1 : i);
        int sj = ej < 0 ? nj : (ej > 0 ? 1 : j);
        int sk = ek < 0 ? nk : (ek > 0 ? 1 : k);
        double v = f[GID(g, si, sj, sk)];
        if (translate >= 0) {
          int di = -ei, dj = -ej, dk = -ek; /* pack direction */
          if (translate == 2)
            v = v + ps[DI(di, dk, dk)][2];
          else
            v = v + ps[DI(di, dj, dk)][translate];
        }
        f[GID(g, i, j, k)] = v;
      }
}

/* check_rectilinear (sgrid_m.f90:868-949), including the stale loop index
 * of the z test (sgrid_m.f90:927-932): the z slice is 1:k with k left over
 * from the y loop. */
static int check_rectilinear(const ogrid *g) {
  const int ni = g->n[0], nj = g->n[1], nk = g->n[2];
  const double TOL = 1e-4;
  double big = 0;
  for (int d = 0; d < 3; d++) {
    double mx = -DBL_MAX, mn = DBL_MAX;
    for (long l = 0; l < g->nt; l++) {
      if (g->g[d][l] > mx) mx = g->g[d][l];
      if (g->g[d][l] < mn) mn = g->g[d][l];
    }
    if (mx - mn > big) big = mx - mn;
  }
  int ox = 1, oy = 1, oz = 1;
  for (int k = 1; k <= nk && ox; k++)
    for (int j = 1; j <= nj && ox; j++) {
      double ymx = -DBL_MAX, ymn = DBL_MAX, zmx = -DBL_MAX, zmn = DBL_MAX;
      for (int i = 1; i <= ni; i++) {
        long l = GID(g, i, j, k);
        ymx = fmax(ymx, g->g[1][l]); ymn = fmin(ymn, g->g[1][l]);
        zmx = fmax(zmx, g->g[2][l]); zmn = fmin(zmn, g->g[2][l]);
      }
      if (fabs(ymx - ymn) / big > TOL || fabs(zmx - zmn) / big > TOL) ox = 0;
    }
  int kstale = nk + 1;
  for (int k = 1; k <= nk; k++) {
    int brk = 0;
    for (int i = 1; i <= ni; i++) {
      double xmx = -DBL_MAX, xmn = DBL_MAX, zmx = -DBL_MAX, zmn = DBL_MAX;
      for (int j = 1; j <= nj; j++) {
        long l = GID(g, i, j, k);
        xmx = fmax(xmx, g->g[0][l]); xmn = fmin(xmn, g->g[0][l]);
        zmx = fmax(zmx, g->g[2][l]); zmn = fmin(zmn, g->g[2][l]);
      }
      if (fabs(xmx - xmn) / big > TOL || fabs(zmx - zmn) / big > TOL) {
        oy = 0; brk = 1; break;
      }
    }
    if (brk) { kstale = k; break; }
  }
  for (int j = 1; j <= nj && oz; j++)
    for (int i = 1; i <= ni && oz; i++) {
      double xmx = -DBL_MAX, xmn = DBL_MAX, ymx = -DBL_MAX, ymn = DBL_MAX;
      for (int k = 1; k <= kstale; k++) {
        long l = GID(g, i, j, k);
        xmx = fmax(xmx, g->g[0][l]); xmn = fmin(xmn, g->g[0][l]);
        ymx = fmax(ymx, g->g[1][l]); ymn = fmin(ymn, g->g[1][l]);
      }
      if (fabs(xmx - xmn) / big > TOL || fabs(ymx - ymn) / big > TOL) oz = 0;
    }
  return ox && oy && oz;
}

static void cross_unit(const double a[3], const double b[3], double out[3]) {
  /* geometry_m.f90:12-30 with unitvector */
  double c0 = a[1] * b[2] - a[2] * b[1];
  double c1 = a[2] * b[0] - a[0] * b[2];
  double c2 = a[0] * b[1] - a[1] * b[0];
  double mag = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  out[0] = c0 / mag; out[1] = c1 / mag; out[2] = c2 / mag;
}

/* init_sgrid on one rank (sgrid_m.f90:16-444) */
static int grid_init(ogrid *g, const int n[3], const double *x,
                     const double *y, const double *z, const int *flag) {
  memset(g, 0, sizeof(*g));
  for (int d = 0; d < 3; d++) { g->n[d] = n[d]; g->N[d] = n[d] + 2; }
  const int ni = n[0], nj = n[1], nk = n[2];
  if (ni < 1 || nj < 1 || nk < 1) return 1; /* ng > n (sgrid_m.f90:85-89) */
  g->nt = (long)g->N[0] * g->N[1] * g->N[2];
  g->np = (long)ni * nj * nk;
  for (int d = 0; d < 3; d++) {
    g->g[d] = calloc(g->nt, sizeof(double));
    g->bn[d] = calloc(g->nt, sizeof(double));
    g->dl[d] = calloc(g->nt, sizeof(double));
  }
  g->flag = calloc(g->nt, sizeof(int));
  /* periodicity (sgrid_m.f90:107-137) */
  int perlo[3] = {0, 0, 0}, perhi[3] = {0, 0, 0};
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        int f = flag[(i - 1) + (long)ni * ((j - 1) + (long)nj * (k - 1))];
        if (f != LCS_INTERNAL) continue;
        if (i == 1) perlo[0] = 1;
        if (j == 1) perlo[1] = 1;
        if (k == 1) perlo[2] = 1;
        if (i == ni) perhi[0] = 1;
        if (j == nj) perhi[1] = 1;
        if (k == nk) perhi[2] = 1;
      }
  for (int d = 0; d < 3; d++) {
    if (perlo[d] != perhi[d]) return 2;
    g->per[d] = perlo[d] && n[d] != 1;
  }
  /* interior */
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        long u = (i - 1) + (long)ni * ((j - 1) + (long)nj * (k - 1));
        long l = GID(g, i, j, k);
        g->g[0][l] = x[u]; g->g[1][l] = y[u]; g->g[2][l] = z[u];
        g->flag[l] = flag[u];
      }
  /* fake coordinates by mirror extrapolation (sgrid_m.f90:181-214) */
  for (int k = 0; k <= nk + 1; k++)
    for (int j = 0; j <= nj + 1; j++)
      for (int i = 0; i <= ni + 1; i++) {
        int ib = iclamp(i, 1, ni), jb = iclamp(j, 1, nj), kb = iclamp(k, 1, nk);
        if (i == ib && j == jb && k == kb) continue;
        int iib = iclamp(ib + (ib - i), 1, ni), jib = iclamp(jb + (jb - j), 1, nj),
            kib = iclamp(kb + (kb - k), 1, nk);
        long lb = GID(g, ib, jb, kb), lib = GID(g, iib, jib, kib), l = GID(g, i, j, k);
        for (int d = 0; d < 3; d++) {
          double delta = g->g[d][lb] - g->g[d][lib];
          g->g[d][l] = g->g[d][lb] + delta;
        }
        g->flag[l] = g->flag[lb];
      }
  /* 2-D flags (sgrid_m.f90:217-255) */
  if (ni == 1)
    for (int k = 0; k <= nk + 1; k++)
      for (int j = 0; j <= nj + 1; j++) {
        int f1 = g->flag[GID(g, 1, j, k)];
        int v = (f1 == LCS_INTERNAL) ? LCS_2D : f1;
        g->flag[GID(g, 0, j, k)] = v; g->flag[GID(g, 2, j, k)] = v;
      }
  if (nj == 1)
    for (int k = 0; k <= nk + 1; k++)
      for (int i = 0; i <= ni + 1; i++) {
        int f1 = g->flag[GID(g, i, 1, k)];
        int v = (f1 == LCS_INTERNAL) ? LCS_2D : f1;
        g->flag[GID(g, i, 0, k)] = v; g->flag[GID(g, i, 2, k)] = v;
      }
  if (nk == 1)
    for (int j = 0; j <= nj + 1; j++)
      for (int i = 0; i <= ni + 1; i++) {
        int f1 = g->flag[GID(g, i, j, 1)];
        int v = (f1 == LCS_INTERNAL) ? LCS_2D : f1;
        g->flag[GID(g, i, j, 0)] = v; g->flag[GID(g, i, j, 2)] = v;
      }
  /* periodic shifts from two exchanges (sgrid_m.f90:257-297) */
  double *fake[3];
  double zero[27][3];
  memset(zero, 0, sizeof(zero));
  for (int d = 0; d < 3; d++) {
    fake[d] = malloc(g->nt * sizeof(double));
    memcpy(fake[d], g->g[d], g->nt * sizeof(double));
    exchange(g, g->g[d], d, zero, 0);
  }
  for (int k = -1; k <= 1; k++)
    for (int j = -1; j <= 1; j++)
      for (int i = -1; i <= 1; i++) {
        int ii = i == -1 ? 0 : (i == 0 ? 1 : ni + 1);
        int jj = j == -1 ? 0 : (j == 0 ? 1 : nj + 1);
        int kk = k == -1 ? 0 : (k == 0 ? 1 : nk + 1);
        long l = GID(g, ii, jj, kk);
        for (int d = 0; d < 3; d++) g->ps[DI(i, j, k)][d] = g->g[d][l] - fake[d][l];
      }
  for (int d = 0; d < 3; d++) exchange(g, g->g[d], d, g->ps, 0);
  for (int d = 0; d < 3; d++) free(fake[d]);
  /* 2-D out-of-plane offsets (sgrid_m.f90:304-359) */
  double rng = 0;
  for (int d = 0; d < 3; d++) {
    double mx = -DBL_MAX, mn = DBL_MAX;
    for (int k = 1; k <= nk; k++)
      for (int j = 1; j <= nj; j++)
        for (int i = 1; i <= ni; i++) {
          double v = g->g[d][GID(g, i, j, k)];
          if (v > mx) mx = v;
          if (v < mn) mn = v;
        }
    if (mx - mn > rng) rng = mx - mn;
  }
  double scaling = rng / (double)imax3(ni, nj, nk);
  for (int k = 0; k <= nk + 1; k++)
    for (int j = 0; j <= nj + 1; j++)
      for (int i = 0; i <= ni + 1; i++) {
        int im1 = i - 1 < 0 ? 0 : i - 1, ip1 = i + 1 > ni + 1 ? ni + 1 : i + 1;
        int jm1 = j - 1 < 0 ? 0 : j - 1, jp1 = j + 1 > nj + 1 ? nj + 1 : j + 1;
        int km1 = k - 1 < 0 ? 0 : k - 1, kp1 = k + 1 > nk + 1 ? nk + 1 : k + 1;
        double delta[3] = {0, 0, 0}, v1[3], v2[3], c[3];
        if (ni == 1 && i != 1) {
          for (int d = 0; d < 3; d++) {
            v1[d] = g->g[d][GID(g, 1, jp1, 1)] - g->g[d][GID(g, 1, jm1, 1)];
            v2[d] = g->g[d][GID(g, 1, 1, kp1)] - g->g[d][GID(g, 1, 1, km1)];
          }
          cross_unit(v1, v2, c);
          for (int d = 0; d < 3; d++) delta[d] = (double)(i - 1) * c[d];
        }
        if (nj == 1 && j != 1) {
          for (int d = 0; d < 3; d++) {
            v1[d] = g->g[d][GID(g, ip1, 1, 1)] - g->g[d][GID(g, im1, 1, 1)];
            v2[d] = g->g[d][GID(g, 1, 1, kp1)] - g->g[d][GID(g, 1, 1, km1)];
          }
          cross_unit(v1, v2, c);
          for (int d = 0; d < 3; d++) delta[d] = (double)(j - 1) * c[d];
        }
        if (nk == 1 && k != 1) {
          for (int d = 0; d < 3; d++) {
            v1[d] = g->g[d][GID(g, ip1, 1, 1)] - g->g[d][GID(g, im1, 1, 1)];
            v2[d] = g->g[d][GID(g, 1, jp1, 1)] - g->g[d][GID(g, 1, jm1, 1)];
          }
          cross_unit(v1, v2, c);
          for (int d = 0; d < 3; d++) delta[d] = (double)(k - 1) * c[d];
        }
        long l = GID(g, i, j, k);
        for (int d = 0; d < 3; d++) g->g[d][l] = g->g[d][l] + delta[d] * scaling;
      }
  /* bounding box incl. ghosts (sgrid_m.f90:362-376) */
  for (int d = 0; d < 3; d++) {
    double mx = -DBL_MAX, mn = DBL_MAX;
    for (long l = 0; l < g->nt; l++) {
      if (g->g[d][l] > mx) mx = g->g[d][l];
      if (g->g[d][l] < mn) mn = g->g[d][l];
    }
    g->bb[d] = mn; g->bb[3 + d] = mx;
  }
  g->rect = check_rectilinear(g);
  /* minloc / maxloc of each coordinate over the ghosted arrays, first
   * occurrence in array order, as 1-based positions (comms_m.f90:1625-1632) */
  for (int d = 0; d < 3; d++) {
    long amin = 0, amax = 0;
    for (long l = 1; l < g->nt; l++) {
      if (g->g[d][l] < g->g[d][amin]) amin = l;
      if (g->g[d][l] > g->g[d][amax]) amax = l;
    }
    const long a2[2] = {amin, amax};
    for (int m = 0; m < 2; m++) {
      const long l = a2[m];
      double pos[3] = {(double)(l % g->N[0] + 1), (double)((l / g->N[0]) % g->N[1] + 1),
                       (double)(l / ((long)g->N[0] * g->N[1]) + 1)};
      for (int e = 0; e < 3; e++) (m ? g->pmax : g->pmin)[d][e] = pos[e];
    }
  }
  /* bcnorm: normalised least-squares gradient of min(flag,1), 26 nbrs
   * (sgrid_m.f90:383-419, gradient_m.f90:114-217, 297-335) */
  {
    double *bc = malloc(g->nt * sizeof(double));
    for (long l = 0; l < g->nt; l++) {
      int f = g->flag[l];
      bc[l] = (f == LCS_2D) ? (double)LCS_INTERNAL : (double)(f < 1 ? f : 1);
    }
    exchange(g, bc, -1, NULL, 0);
    for (int k = 1; k <= nk; k++)
      for (int j = 1; j <= nj; j++)
        for (int i = 1; i <= ni; i++) {
          long l = GID(g, i, j, k);
          double sxx = 0, syy = 0, szz = 0, sxy = 0, sxz = 0, syz = 0, dx[3];
          for (int nb = 1; nb < 27; nb++) {
            long m = GID(g, i + NBR[nb][0], j + NBR[nb][1], k + NBR[nb][2]);
            for (int d = 0; d < 3; d++) dx[d] = g->g[d][m] - g->g[d][l];
            double idw = 1.0 / (dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
            sxx = sxx + idw * (dx[0] * dx[0]);
            syy = syy + idw * (dx[1] * dx[1]);
            szz = szz + idw * (dx[2] * dx[2]);
            sxy = sxy + idw * dx[0] * dx[1];
            sxz = sxz + idw * dx[0] * dx[2];
            syz = syz + idw * dx[1] * dx[2];
          }
          double sden = 2.0 * sxy * sxz * syz + sxx * syy * szz - sxx * (syz * syz) -
                        syy * (sxz * sxz) - szz * (sxy * sxy);
          double gxv = 0, gyv = 0, gzv = 0;
          for (int nb = 1; nb < 27; nb++) {
            long m = GID(g, i + NBR[nb][0], j + NBR[nb][1], k + NBR[nb][2]);
            for (int d = 0; d < 3; d++) dx[d] = g->g[d][m] - g->g[d][l];
            double idw = 1.0 / (dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
            double wx = idw * ((syy * szz - syz * syz) * dx[0] + (sxz * syz - sxy * szz) * dx[1] +
                               (sxy * syz - sxz * syy) * dx[2]) / sden;
            double wy = idw * ((sxz * syz - sxy * szz) * dx[0] + (sxx * szz - sxz * sxz) * dx[1] +
                               (sxy * sxz - syz * sxx) * dx[2]) / sden;
            double wz = idw * ((sxy * syz - sxz * syy) * dx[0] + (sxy * sxz - syz * sxx) * dx[1] +
                               (sxx * syy - sxy * sxy) * dx[2]) / sden;
            double t = bc[m] - bc[l];
            gxv = gxv + t * wx; gyv = gyv + t * wy; gzv = gzv + t * wz;
          }
          double mag = gxv * gxv + gyv * gyv + gzv * gzv;
          if (mag <= 0.0) {
            g->bn[0][l] = 0; g->bn[1][l] = 0; g->bn[2][l] = 0;
          } else {
            g->bn[0][l] = gxv / sqrt(mag); g->bn[1][l] = gyv / sqrt(mag);
            g->bn[2][l] = gzv / sqrt(mag);
          }
        }
    free(bc);
  }
  /* least-squares gradient weights of the face neighbours for a
   * non-rectilinear grid (sgrid_m.f90:421-424, compute_lsgw with
   * full_conn = .false., gradient_m.f90:114-217): inverse-distance weights
   * and the closed-form solution of the 3x3 normal equations */
  if (!g->rect) {
    g->lsw = calloc((size_t)18 * g->nt, sizeof(double));
    for (int k = 1; k <= nk; k++)
      for (int j = 1; j <= nj; j++)
        for (int i = 1; i <= ni; i++) {
          long l = GID(g, i, j, k);
          double sxx = 0, syy = 0, szz = 0, sxy = 0, sxz = 0, syz = 0, dx[3];
          for (int nb = 1; nb <= 6; nb++) {
            long m = GID(g, i + NBR[nb][0], j + NBR[nb][1], k + NBR[nb][2]);
            for (int d = 0; d < 3; d++) dx[d] = g->g[d][m] - g->g[d][l];
            double idw = 1.0 / (dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
            sxx = sxx + idw * (dx[0] * dx[0]);
            syy = syy + idw * (dx[1] * dx[1]);
            szz = szz + idw * (dx[2] * dx[2]);
            sxy = sxy + idw * dx[0] * dx[1];
            sxz = sxz + idw * dx[0] * dx[2];
            syz = syz + idw * dx[1] * dx[2];
          }
          double sden = 2.0 * sxy * sxz * syz + sxx * syy * szz - sxx * (syz * syz) -
                        syy * (sxz * sxz) - szz * (sxy * sxy);
          for (int nb = 1; nb <= 6; nb++) {
            long m = GID(g, i + NBR[nb][0], j + NBR[nb][1], k + NBR[nb][2]);
            for (int d = 0; d < 3; d++) dx[d] = g->g[d][m] - g->g[d][l];
            double idw = 1.0 / (dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
            double *w = g->lsw + (size_t)(nb - 1) * 3 * g->nt;
            w[l] = idw * ((syy * szz - syz * syz) * dx[0] + (sxz * syz - sxy * szz) * dx[1] +
                          (sxy * syz - sxz * syy) * dx[2]) / sden;
            w[g->nt + l] = idw * ((sxz * syz - sxy * szz) * dx[0] + (sxx * szz - sxz * sxz) * dx[1] +
                                  (sxy * sxz - syz * sxx) * dx[2]) / sden;
            w[2 * g->nt + l] = idw * ((sxy * syz - sxz * syy) * dx[0] + (sxy * sxz - syz * sxx) * dx[1] +
                                      (sxx * syy - sxy * sxy) * dx[2]) / sden;
          }
        }
  }
  /* 1/delta over the 3x3x3 neighbourhood (sgrid_m.f90:951-1014) */
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        long l = GID(g, i, j, k);
        for (int d = 0; d < 3; d++) {
          double mx = -DBL_MAX, mn = DBL_MAX;
          for (int c = k - 1; c <= k + 1; c++)
            for (int b = j - 1; b <= j + 1; b++)
              for (int a = i - 1; a <= i + 1; a++) {
                double v = g->g[d][GID(g, a, b, c)];
                if (v > mx) mx = v;
                if (v < mn) mn = v;
              }
          double dd = 0.5 * fabs(mx - mn);
          g->dl[d][l] = dd > 0.0 ? 1.0 / dd : 0.0;
        }
      }
  return 0;
}

static void grid_free(ogrid *g) {
  for (int d = 0; d < 3; d++) { free(g->g[d]); free(g->bn[d]); free(g->dl[d]); }
  free(g->flag);
  free(g->lsw);
}

/* grad_sr1_ls (gradient_m.f90:241-295) at interior node l of a
 * non-rectilinear grid: sum over the face neighbours, in NBR order, of
 * (F(nbr) - F(l)) times their weights; G[d] = dF/dx_d */
static void grad_ls(const ogrid *g, const double *F, long l, const long nbr[6], double G[3]) {
  G[0] = 0.0; G[1] = 0.0; G[2] = 0.0;
  for (int f = 0; f < 6; f++) {
    const double t = F[nbr[f]] - F[l];
    const double *w = g->lsw + (size_t)f * 3 * g->nt;
    G[0] = G[0] + t * w[l];
    G[1] = G[1] + t * w[g->nt + l];
    G[2] = G[2] + t * w[2 * g->nt + l];
  }
}

/* set_velocity_bc (sgrid_m.f90:810-866) on an exchanged field */
static int set_velocity_bc(const ogrid *g, double *f[3]) {
  const int ni = g->n[0], nj = g->n[1], nk = g->n[2];
  int err = 0;
  for (int k = 0; k <= nk + 1; k++)
    for (int j = 0; j <= nj + 1; j++)
      for (int i = 0; i <= ni + 1; i++) {
        long l = GID(g, i, j, k);
        int ib = (i >= 1 && j >= 1 && k >= 1 && i <= ni && j <= nj && k <= nk);
        switch (g->flag[l]) {
        case LCS_INTERNAL: break;
        case LCS_MASK:
          for (int d = 0; d < 3; d++) f[d][l] = 0.0;
          break;
        case LCS_WALL:
          if (ib) break;
          for (int d = 0; d < 3; d++) f[d][l] = 0.0;
          break;
        case LCS_SLIP: case LCS_INFLOW: case LCS_OUTFLOW: case LCS_2D: {
          if (ib) break;
          long m = GID(g, iclamp(i, 1, ni), iclamp(j, 1, nj), iclamp(k, 1, nk));
          for (int d = 0; d < 3; d++) f[d][l] = f[d][m];
          break;
        }
        default: err = 1;
        }
      }
  return err;
}

/* set_flowmap_bc (flowmap_m.f90:325-392) */
static int set_flowmap_bc(const ogrid *g, double *f[3]) {
  const int ni = g->n[0], nj = g->n[1], nk = g->n[2];
  int err = 0;
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        long l = GID(g, i, j, k);
        int fl = g->flag[l];
        if (fl == LCS_MASK || fl == LCS_INFLOW || fl == LCS_OUTFLOW)
          for (int d = 0; d < 3; d++) f[d][l] = 0.0;
      }
  for (int k = 0; k <= nk + 1; k++)
    for (int j = 0; j <= nj + 1; j++)
      for (int i = 0; i <= ni + 1; i++) {
        if (i >= 1 && j >= 1 && k >= 1 && i <= ni && j <= nj && k <= nk) continue;
        long l = GID(g, i, j, k);
        switch (g->flag[l]) {
        case LCS_INTERNAL: break;
        case LCS_WALL: case LCS_MASK: case LCS_INFLOW: case LCS_OUTFLOW:
          for (int d = 0; d < 3; d++) f[d][l] = 0.0;
          break;
        case LCS_SLIP: case LCS_2D: {
          long m = GID(g, iclamp(i, 1, ni), iclamp(j, 1, nj), iclamp(k, 1, nk));
          for (int d = 0; d < 3; d++) f[d][l] = f[d][m];
          break;
        }
        default: err = 1;
        }
      }
  return err;
}

/* ======================================================================
 * B-splines: restatement of the 3-D tensor interpolation the reference
 * vendors (bspline_sub_module.f90).  Algorithms are de Boor's (A Practical
 * Guide to Splines): knot placement, BSPLVB basis recurrence, banded LU
 * without pivoting, and the BVALUE evaluation recurrence.
 * ====================================================================== */

/* dbknot (bspline_sub_module.f90:1737-1785): k knots at x(1); k knots at
 * x(n)+0.1(x(n)-x(n-1)); interior knots at data points (even k) or at
 * midpoints (odd k).  1-based indices in comments, 0-based arrays here. */
void oc_bspline_knots(const double *x, int n, int k, double *t) {
  double rnot = x[n - 1] + 0.1 * (x[n - 1] - x[n - 2]);
  for (int j = 1; j <= k; j++) {
    t[j - 1] = x[0];
    t[n + j - 1] = rnot;
  }
  if (k % 2 == 1) {
    int i = (k - 1) / 2 - k;
    for (int j = k + 1; j <= n; j++) {
      int ipj = i + j;
      t[j - 1] = 0.5 * (x[ipj - 1] + x[ipj]);
    }
  } else {
    int i = k / 2 - k;
    for (int j = k + 1; j <= n; j++) t[j - 1] = x[i + j - 1];
  }
}

/* dintrv (bspline_sub_module.f90:2558-2634): largest ileft (1-based) in
 * [1,lxt] with xt(ileft) <= x; mflag -1 below, +1 at/after xt(lxt).  The
 * result does not depend on the search start, so a bisection is used. */
static int dintrv(const double *xt, int lxt, double x, int *mflag) {
  if (x >= xt[lxt - 1]) { *mflag = 1; return lxt; }
  if (lxt <= 1 || x < xt[0]) { *mflag = -1; return 1; }
  int lo = 1, hi = lxt; /* xt(lo) <= x < xt(hi) */
  while (hi - lo > 1) {
    int mid = (lo + hi) / 2;
    if (x >= xt[mid - 1]) lo = mid; else hi = mid;
  }
  *mflag = 0;
  return lo;
}

/* dbspvn with index=1, jhigh=k (bspline_sub_module.f90:2288-2370):
 * the k B-splines of order k nonzero at x, t(left) <= x <= t(left+1). */
static void bsplvb(const double *t, int k, double x, int left, double *v) {
  double dp[8], dm[8];
  v[0] = 1.0;
  for (int j = 1; j < k; j++) {
    dp[j - 1] = t[left + j - 1] - x;      /* t(left+j) - x */
    dm[j - 1] = x - t[left + 1 - j - 1];  /* x - t(left+1-j) */
    double vmprev = 0.0;
    for (int l = 1; l <= j; l++) {
      int jp1ml = j + 1 - l;
      double vm = v[l - 1] / (dp[l - 1] + dm[jp1ml - 1]);
      v[l - 1] = vm * dp[l - 1] + vmprev;
      vmprev = vm * dm[jp1ml - 1];
    }
    v[j] = vmprev;
  }
}

/* band storage: a(r,c) (0-based) at w[(r - c + mu) + (ml+mu+1)*c] */
#define WB(w, nb, mu, r, c) (w)[((r) - (c) + (mu)) + (long)(nb) * (c)]

/* dbintk matrix build + dbnfac (bspline_sub_module.f90:1896-2184) */
static int bs_factor(const double *x, const double *t, int n, int k, double *w) {
  const int ml = k - 1, mu = k - 1, nb = 2 * k - 1;
  memset(w, 0, sizeof(double) * (size_t)nb * n);
  for (int i = 0; i + 1 < n; i++)
    if (x[i] >= x[i + 1]) return 102;
  int left = k; /* 1-based */
  double v[8];
  for (int i = 1; i <= n; i++) {
    double xi = x[i - 1];
    int ilp1mx = (i + k < n + 1) ? i + k : n + 1;
    if (left < i) left = i;
    if (xi < t[left - 1]) return 103;
    int found = 0;
    for (;;) {
      if (xi < t[left]) { found = 1; break; } /* t(left+1) */
      left = left + 1;
      if (left >= ilp1mx) break;
    }
    if (!found) {
      left = left - 1;
      if (xi > t[left]) return 103;
    }
    bsplvb(t, k, xi, left, v);
    for (int j = 1; j <= k; j++) {
      int col = left - k + j; /* 1-based column */
      WB(w, nb, mu, i - 1, col - 1) = v[j - 1];
    }
  }
  /* dbnfac: LU without pivoting */
  for (int i = 0; i + 1 < n; i++) {
    double pivot = WB(w, nb, mu, i, i);
    if (pivot == 0.0) return 104;
    int jmax = ml < n - 1 - i ? ml : n - 1 - i;
    for (int j = 1; j <= jmax; j++) WB(w, nb, mu, i + j, i) /= pivot;
    int kmax = mu < n - 1 - i ? mu : n - 1 - i;
    for (int kk = 1; kk <= kmax; kk++) {
      int ipk = i + kk;
      double factor = WB(w, nb, mu, i, ipk);
      for (int j = 1; j <= jmax; j++)
        WB(w, nb, mu, i + j, ipk) = WB(w, nb, mu, i + j, ipk) - WB(w, nb, mu, i + j, i) * factor;
    }
  }
  if (WB(w, nb, mu, n - 1, n - 1) == 0.0) return 104;
  return 0;
}

/* dbnslv (bspline_sub_module.f90:2202-2259) */
static void bs_solve(const double *w, int n, int k, double *b, long stride) {
  const int ml = k - 1, mu = k - 1, nb = 2 * k - 1;
  if (n != 1) {
    for (int i = 0; i + 1 < n; i++) {
      int jmax = ml < n - 1 - i ? ml : n - 1 - i;
      for (int j = 1; j <= jmax; j++)
        b[(i + j) * stride] = b[(i + j) * stride] - b[i * stride] * WB(w, nb, mu, i + j, i);
    }
    for (int i = n - 1; i >= 1; i--) {
      b[i * stride] = b[i * stride] / WB(w, nb, mu, i, i);
      int jmax = mu < i ? mu : i;
      for (int j = 1; j <= jmax; j++)
        b[(i - j) * stride] = b[(i - j) * stride] - b[i * stride] * WB(w, nb, mu, i - j, i);
    }
  }
  b[0] = b[0] / WB(w, nb, mu, 0, 0);
}

/* db3ink (bspline_sub_module.f90:451-533) with knots chosen by dbknot and
 * the three dbtpcf passes x, y, z.  f and coef are nx*ny*nz, x fastest. */
int oc_bspline_fit3(const double *x, int nx, const double *y, int ny,
                    const double *z, int nz, int kx, int ky, int kz,
                    const double *f, double *tx, double *ty, double *tz,
                    double *coef) {
  oc_bspline_knots(x, nx, kx, tx);
  oc_bspline_knots(y, ny, ky, ty);
  oc_bspline_knots(z, nz, kz, tz);
  int nmax = imax3(nx, ny, nz), kmax = imax3(kx, ky, kz);
  double *w = malloc(sizeof(double) * (size_t)(2 * kmax - 1) * nmax);
  long nxy = (long)nx * ny, ntot = nxy * nz;
  memcpy(coef, f, sizeof(double) * ntot);
  int err;
  if ((err = bs_factor(x, tx, nx, kx, w))) { free(w); return err; }
  #pragma omp parallel for schedule(static)
  for (long c = 0; c < (long)ny * nz; c++) bs_solve(w, nx, kx, coef + c * nx, 1);
  if ((err = bs_factor(y, ty, ny, ky, w))) { free(w); return err; }
  #pragma omp parallel for schedule(static)
  for (long c = 0; c < (long)nx * nz; c++) {
    long i = c % nx, kk = c / nx;
    bs_solve(w, ny, ky, coef + i + kk * nxy, nx);
  }
  if ((err = bs_factor(z, tz, nz, kz, w))) { free(w); return err; }
  #pragma omp parallel for schedule(static)
  for (long c = 0; c < nxy; c++) bs_solve(w, nz, kz, coef + c, nxy);
  free(w);
  return 0;
}

/* dbvalu with ideriv = 0 (bspline_sub_module.f90:2397-2533): de Boor's
 * evaluation of sum a(j) B_j(x) with knots t(1..n+k). */
static double dbvalu(const double *t, const double *a, long astride, int n,
                     int k, double x, int *iflag) {
  int mflag;
  int i = dintrv(t, n + 1, x, &mflag);
  if (x < t[k - 1]) { *iflag = 404; return 0.0; }
  if (mflag != 0) {
    if (x > t[i - 1]) { *iflag = 405; return 0.0; }
    for (;;) {
      if (i == k) { *iflag = 406; return 0.0; }
      i = i - 1;
      if (x != t[i - 1]) break;
    }
  }
  double aj[8], dp[8], dm[8];
  for (int j = 1; j <= k; j++) aj[j - 1] = a[(long)(i - k + j - 1) * astride];
  for (int j = 1; j <= k; j++) {
    dp[j - 1] = t[i + j - 1] - x;
    dm[j - 1] = x - t[i + 1 - j - 1];
  }
  for (int j = 1; j <= k - 1; j++) {
    int kmj = k - j, ilo = kmj;
    for (int jj = 1; jj <= kmj; jj++) {
      aj[jj - 1] = (aj[jj] * dm[ilo - 1] + aj[jj - 1] * dp[jj - 1]) / (dm[ilo - 1] + dp[jj - 1]);
      ilo = ilo - 1;
    }
  }
  *iflag = 0;
  return aj[0];
}

/* db3val (bspline_sub_module.f90:574-656): 0 outside the knot range and at
 * y or z equal to the last knot. */
double oc_bspline_eval3(double xv, double yv, double zv, const double *tx,
                        const double *ty, const double *tz, int nx, int ny,
                        int nz, int kx, int ky, int kz, const double *coef,
                        int *iflag) {
  if (xv < tx[0] || xv > tx[nx + kx - 1]) { *iflag = 1; return 0.0; }
  if (yv < ty[0] || yv > ty[ny + ky - 1]) { *iflag = 2; return 0.0; }
  if (zv < tz[0] || zv > tz[nz + kz - 1]) { *iflag = 3; return 0.0; }
  int my, mz;
  *iflag = -1;
  int lefty = dintrv(ty, ny + ky, yv, &my);
  if (my != 0) return 0.0;
  int leftz = dintrv(tz, nz + kz, zv, &mz);
  if (mz != 0) return 0.0;
  double temp1[8][8], temp2[8];
  long nxy = (long)nx * ny;
  int kcolz = leftz - kz;
  for (int k = 1; k <= kz; k++) {
    kcolz = kcolz + 1;
    int kcoly = lefty - ky;
    for (int j = 1; j <= ky; j++) {
      kcoly = kcoly + 1;
      temp1[k - 1][j - 1] = dbvalu(tx, coef + (long)(kcoly - 1) * nx + (long)(kcolz - 1) * nxy, 1,
                                   nx, kx, xv, iflag);
      if (*iflag != 0) return 0.0;
    }
  }
  int kcoly = lefty - ky + 1;
  for (int k = 1; k <= kz; k++) {
    temp2[k - 1] = dbvalu(ty + kcoly - 1, temp1[k - 1], 1, ky, ky, yv, iflag);
    if (*iflag != 0) return 0.0;
  }
  int kcz = leftz - kz + 1;
  return dbvalu(tz + kcz - 1, temp2, 1, kz, kz, zv, iflag);
}

/* ======================================================================
 * Eigen-solve: cyclic Jacobi on a symmetric 3x3 (stands in for LAPACK
 * DSYEV('V','U') at lcs_m.f90:160-177; only the largest eigenvalue is used).
 * ====================================================================== */
static void eig_sym3_all(const double *ain, double w[3]);

double oc_eig_sym3_max(const double *ain) {
  double w[3];
  eig_sym3_all(ain, w);
  double m = w[0];
  if (w[1] > m) m = w[1];
  if (w[2] > m) m = w[2];
  return m;
}

/* all three eigenvalues (unordered) */
static void eig_sym3_all(const double *ain, double w[3]) {
  double a[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) a[i][j] = 0.5 * (ain[3 * i + j] + ain[3 * j + i]);
  for (int sweep = 0; sweep < 60; sweep++) {
    double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    double dia = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
    if (off == 0.0 || off <= 1e-36 * dia) break;
    for (int p = 0; p < 2; p++)
      for (int q = p + 1; q < 3; q++) {
        double apq = a[p][q];
        if (apq == 0.0) continue;
        double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
        double t;
        if (fabs(theta) > 1e150)
          t = 0.5 / theta;
        else
          t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        /* A' = J^T A J on rows/cols p,q */
        for (int r = 0; r < 3; r++) {
          double arp = a[r][p], arq = a[r][q];
          a[r][p] = c * arp - s * arq;
          a[r][q] = s * arp + c * arq;
        }
        for (int r = 0; r < 3; r++) {
          double apr = a[p][r], aqr = a[q][r];
          a[p][r] = c * apr - s * aqr;
          a[q][r] = s * apr + c * aqr;
        }
        a[p][q] = 0.0; a[q][p] = 0.0;
      }
  }
  w[0] = a[0][0];
  w[1] = a[1][1];
  w[2] = a[2][2];
}

/* lcs_m.f90:137-145: cg = gradfm^T gradfm, g[3*i+j] = d fm_i / d x_j */
void oc_cauchy_green(const double *g, double *c) {
  const double xx = g[0], xy = g[1], xz = g[2], yx = g[3], yy = g[4], yz = g[5],
               zx = g[6], zy = g[7], zz = g[8];
  c[0] = xx * xx + yx * yx + zx * zx;
  c[1] = xx * xy + yx * yy + zx * zy;
  c[2] = xx * xz + yx * yz + zx * zz;
  c[3] = xy * xx + yy * yx + zy * zx;
  c[4] = xy * xy + yy * yy + zy * zy;
  c[5] = xy * xz + yy * yz + zy * zz;
  c[6] = xz * xx + yz * yx + zz * zx;
  c[7] = xz * xy + yz * yy + zz * zy;
  c[8] = xz * xz + yz * yz + zz * zz;
}

/* ======================================================================
 * Library state (scfd_t, lp_t, lcs_t: data_m.f90:317-377)
 * ====================================================================== */
/* One particle set (lp_t, data_m.f90:317-337) with the diagnostic state
 * that goes with it.  A diagnostic owns one set on its grid and, with
 * AUX_GRID, six more on the shifted auxiliary grids (lcs_m.f90:51-53). */
typedef struct odiag_ {
  int id, type, dir;
  double T, h, dt_factor;
  char label[132];
  ofg *L;          /* lcs%sgrid / lp%sgrid: the flow grid or the set's own */
  int own_grid;    /* L was allocated for this set */
  long np;         /* particles = nodes of L */
  /* particles in origin order */
  double *xp[3], *dx[3], *up[3];
  int *no[3], *ns[3], *pflag;
  double *fm[3];   /* lp%fm (ghosted, on L) */
  int nstep, nslot;
  double **slot;   /* nslot * 3 interior fields */
  int *slot_ok;
  double *comp[3]; /* composed T flowmap (interior) */
  double *ftle;
  int noutput;
  int naux;                /* 6 with AUX_GRID, else 0 */
  struct odiag_ *auxs[6];  /* lpX0, lpY0, lpZ0, lpX1, lpY1, lpZ1 */
} odiag;

struct ocfd {
  ofg F;                  /* the flow grid (scfd%sgrid) */
  double t_n, t_np1;
  int first_call;
  double *un[3], *unp1[3];
  int ndiag;
  odiag *d[NMAX_STRUCT];
  int integrator, interpolator, incompressible, error;
  int aux_grid;           /* AUX_GRID (data_m.f90:39) */
  double cfl_max;
  int nsub_last;
  long long integ_fwd, integ_bkwd;
  int warn_opt, warn_par; /* SAVEd warn flags (cfd2lcs.f90:444,632) */
  double *work[3];        /* blended velocity */
  long oob;               /* out-of-range spline evaluations */
};

#define AUX_GRID_SCALING 0.2 /* data_m.f90:92 */

/* ---------------- interpolation (interp_s2u_r1, lp_tracking_m.f90:279-662) */
typedef struct {
  int interp, kx, ky, kz; /* spline orders */
  const double *t[3];
  const double *c[3];
} ointerp;

/* LINEAR (lp_tracking_m.f90:342-474).  The bracket follows the tracked node;
 * an index that would leave the ghosted range (the reference reads out of
 * bounds there) is shifted back inside. */
static void interp_linear(const ofg *G, double *const f[3], const double p[3],
                          const int nd[3], double out[3]) {
  const ogrid *g = &G->g;
  int i0[3], i1[3];
  double tt[3], mt[3];
  for (int d = 0; d < 3; d++) {
    const double *xl = G->xl[d];
    int n0 = nd[d];
    if (n0 < 0) n0 = 0;
    if (n0 > g->n[d] + 1) n0 = g->n[d] + 1;
    if (p[d] > xl[n0]) { i0[d] = n0; i1[d] = n0 + 1; }
    else { i0[d] = n0 - 1; i1[d] = n0; }
    if (i0[d] < 0) { i0[d] = 0; i1[d] = 1; }
    if (i1[d] > g->n[d] + 1) { i1[d] = g->n[d] + 1; i0[d] = g->n[d]; }
    double x0 = xl[i0[d]], x1 = xl[i1[d]];
    tt[d] = clamp01((p[d] - x0) / (x1 - x0));
    mt[d] = 1.0 - tt[d];
  }
  for (int c = 0; c < 3; c++) {
    const double *F = f[c];
    double f0 = F[GID(g, i0[0], i0[1], i0[2])], f1 = F[GID(g, i1[0], i0[1], i0[2])];
    double f2 = F[GID(g, i0[0], i1[1], i0[2])], f3 = F[GID(g, i1[0], i1[1], i0[2])];
    double f4 = F[GID(g, i0[0], i0[1], i1[2])], f5 = F[GID(g, i1[0], i0[1], i1[2])];
    double f6 = F[GID(g, i0[0], i1[1], i1[2])], f7 = F[GID(g, i1[0], i1[1], i1[2])];
    out[c] = mt[0] * mt[1] * mt[2] * f0 + tt[0] * mt[1] * mt[2] * f1 + mt[0] * tt[1] * mt[2] * f2 +
             tt[0] * tt[1] * mt[2] * f3 + mt[0] * mt[1] * tt[2] * f4 + tt[0] * mt[1] * tt[2] * f5 +
             mt[0] * tt[1] * tt[2] * f6 + tt[0] * tt[1] * tt[2] * f7;
  }
}

/* spline orders for QUADRATIC/CUBIC (lp_tracking_m.f90:478-508) */
static void spline_orders(const ogrid *g, int interp, int k[3]) {
  int order = interp == QUADRATIC ? 2 : 3;
  for (int d = 0; d < 3; d++) {
    int kk = order + 1;
    if (kk > g->n[d] + 2 - 1) kk = g->n[d] + 2 - 1;
    if (kk < 2) kk = 2;
    k[d] = kk;
  }
}

/* fit (db3ink) the three components of a ghosted field into G->bsc */
static int spline_fit(ofg *G, double *const f[3], int interp) {
  const ogrid *g = &G->g;
  int k[3];
  spline_orders(g, interp, k);
  for (int c = 0; c < 3; c++) {
    int e = oc_bspline_fit3(G->xl[0], g->N[0], G->xl[1], g->N[1], G->xl[2], g->N[2], k[0], k[1],
                            k[2], f[c], G->bst[0], G->bst[1], G->bst[2], G->bsc[c]);
    if (e) return e;
  }
  return 0;
}

static void spline_eval(ocfd *s, const ofg *G, int interp, const double p[3], double out[3]) {
  const ogrid *g = &G->g;
  int k[3], fl;
  spline_orders(g, interp, k);
  for (int c = 0; c < 3; c++) {
    out[c] = oc_bspline_eval3(p[0], p[1], p[2], G->bst[0], G->bst[1], G->bst[2], g->N[0], g->N[1],
                              g->N[2], k[0], k[1], k[2], G->bsc[c], &fl);
    if (fl > 0) {
      #pragma omp atomic
      s->oob++;
    }
  }
}

/* interp_s2u_r1 TSE / TSE_LIMIT (lp_tracking_m.f90:526-654): the value at
 * the node plus grad(f) at the node dotted with (q - x_node), the gradient
 * being grad_sr1's central difference over the ghosted grid coordinates
 * (gradient_m.f90:11-66, columns of one-node axes zeroed); TSE_LIMIT clamps
 * to the min/max of f over the 3x3x3 neighbourhood (594-643). */
static void interp_tse(const ofg *G, double *const f[3], const double q[3], const int n3[3], int limit,
                       double o[3]) {
  const ogrid *g = &G->g;
  const int i = iclamp(n3[0], 1, g->n[0]), j = iclamp(n3[1], 1, g->n[1]), k = iclamp(n3[2], 1, g->n[2]);
  const long l = GID(g, i, j, k);
  const long xp = GID(g, i + 1, j, k), xm = GID(g, i - 1, j, k);
  const long yp = GID(g, i, j + 1, k), ym = GID(g, i, j - 1, k);
  const long zp = GID(g, i, j, k + 1), zm = GID(g, i, j, k - 1);
  const double dx = q[0] - g->g[0][l], dy = q[1] - g->g[1][l], dz = q[2] - g->g[2][l];
  const long nbr[6] = {xm, ym, zm, xp, yp, zp};
  for (int c = 0; c < 3; c++) {
    const double *F = f[c];
    double gx, gy, gz;
    if (g->rect) {
      gx = (F[xp] - F[xm]) / (g->g[0][xp] - g->g[0][xm]);
      gy = (F[yp] - F[ym]) / (g->g[1][yp] - g->g[1][ym]);
      gz = (F[zp] - F[zm]) / (g->g[2][zp] - g->g[2][zm]);
      if (g->n[0] == 1) gx = 0.0;
      if (g->n[1] == 1) gy = 0.0;
      if (g->n[2] == 1) gz = 0.0;
    } else {
      double G3[3];
      grad_ls(g, F, l, nbr, G3);
      gx = G3[0]; gy = G3[1]; gz = G3[2];
    }
    double v = F[l] + dx * gx + dy * gy + dz * gz;
    if (limit) {
      double mn = F[l], mx = F[l];
      for (int c3 = -1; c3 <= 1; c3++)
        for (int b = -1; b <= 1; b++)
          for (int a = -1; a <= 1; a++) {
            const double w = F[GID(g, i + a, j + b, k + c3)];
            if (w < mn) mn = w;
            if (w > mx) mx = w;
          }
      if (v < mn) v = mn;
      if (v > mx) v = mx;
    }
    o[c] = v;
  }
}

/* Interpolate field f (on grid G) at the np particle positions pos using
 * node array nd.  For spline interpolators the field is fitted once (as the
 * reference does on every call, lp_tracking_m.f90:509-511). */
static int interp_field(ocfd *s, ofg *G, long np, double *const f[3], double *const pos[3],
                        int *const nd[3], double *const out[3]) {
  const ogrid *g = &G->g;
  int it = s->interpolator;
  /* non-rectilinear grids interpolate with TSE or TSE_LIMIT
   * (lp_tracking_m.f90:319-328) */
  if (!g->rect && it != TSE) it = TSE_LIMIT;
  if (it == QUADRATIC || it == CUBIC) {
    int e = spline_fit(G, f, it);
    if (e) return e;
  }
  if (it != NEAREST_NBR && it != LINEAR && it != QUADRATIC && it != CUBIC && it != TSE && it != TSE_LIMIT) return 1;
  #pragma omp parallel for schedule(static)
  for (long p = 0; p < np; p++) {
    double q[3] = {pos[0][p], pos[1][p], pos[2][p]}, o[3];
    int n3[3] = {nd[0][p], nd[1][p], nd[2][p]};
    if (it == NEAREST_NBR) {
      for (int c = 0; c < 3; c++) {
        int a = iclamp(n3[0], 0, g->n[0] + 1), b = iclamp(n3[1], 0, g->n[1] + 1),
            cc = iclamp(n3[2], 0, g->n[2] + 1);
        o[c] = f[c][GID(g, a, b, cc)];
      }
    } else if (it == LINEAR) {
      interp_linear(G, f, q, n3, o);
    } else if (it == TSE || it == TSE_LIMIT) {
      interp_tse(G, f, q, n3, it == TSE_LIMIT, o);
    } else {
      spline_eval(s, G, it, q, o);
    }
    for (int c = 0; c < 3; c++) out[c][p] = o[c];
  }
  return 0;
}

/* ---------------- particle operations */

/* project_lp2plane's move of one point q onto the data plane through the
 * node n3 of a 2-D grid g (lp_m.f90:450-536) */
static void project_point(const ogrid *g, const int n3[3], double q[3]) {
  const int ni = g->n[0], nj = g->n[1], nk = g->n[2];
  {
    int i = iclamp(n3[0], 1, ni), j = iclamp(n3[1], 1, nj), k = iclamp(n3[2], 1, nk);
    double v1[3], v2[3], nrm[3];
    for (int d = 0; d < 3; d++) {
      const double *G = g->g[d];
      if (ni == 1) {
        v1[d] = G[GID(g, i, j + 1, k)] - G[GID(g, i, j - 1, k)];
        v2[d] = G[GID(g, i, j, k + 1)] - G[GID(g, i, j, k - 1)];
      }
      if (nj == 1) {
        v1[d] = G[GID(g, i + 1, j, k)] - G[GID(g, i - 1, j, k)];
        v2[d] = G[GID(g, i, j, k + 1)] - G[GID(g, i, j, k - 1)];
      }
      if (nk == 1) {
        v1[d] = G[GID(g, i + 1, j, k)] - G[GID(g, i - 1, j, k)];
        v2[d] = G[GID(g, i, j + 1, k)] - G[GID(g, i, j - 1, k)];
      }
    }
    nrm[0] = v1[1] * v2[2] - v1[2] * v2[1];
    nrm[1] = v1[2] * v2[0] - v1[0] * v2[2];
    nrm[2] = v1[0] * v2[1] - v1[1] * v2[0];
    double mag = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    for (int d = 0; d < 3; d++) nrm[d] = nrm[d] / mag;
    long l = GID(g, i, j, k);
    double d2p[3];
    for (int d = 0; d < 3; d++) d2p[d] = q[d] - g->g[d][l];
    double nm = d2p[0] * nrm[0] + d2p[1] * nrm[1] + d2p[2] * nrm[2];
    if (nm < 0.0) {
      nm = -nm;
      for (int d = 0; d < 3; d++) nrm[d] = -nrm[d];
    }
    for (int d = 0; d < 3; d++) q[d] = q[d] - nm * nrm[d];
  }
}

/* project_lp2plane (lp_m.f90:418-538) of a particle set against grid g with
 * node array nd; both node arrays get index 1 along the one-node axis */
static void project_lp2plane(const ogrid *g, odiag *dg, int *const nd[3]) {
  const int ni = g->n[0], nj = g->n[1], nk = g->n[2];
  if (ni > 1 && nj > 1 && nk > 1) return;
  long np = dg->np;
  if (ni == 1) for (long p = 0; p < np; p++) { dg->no[0][p] = 1; dg->ns[0][p] = 1; }
  if (nj == 1) for (long p = 0; p < np; p++) { dg->no[1][p] = 1; dg->ns[1][p] = 1; }
  if (nk == 1) for (long p = 0; p < np; p++) { dg->no[2][p] = 1; dg->ns[2][p] = 1; }
  #pragma omp parallel for schedule(static)
  for (long p = 0; p < np; p++) {
    const int n3[3] = {nd[0][p], nd[1][p], nd[2][p]};
    double q[3] = {dg->xp[0][p], dg->xp[1][p], dg->xp[2][p]};
    project_point(g, n3, q);
    for (int d = 0; d < 3; d++) dg->xp[d][p] = q[d];
  }
}

/* set_lp_bc (lp_tracking_m.f90:185-277) */
static int set_lp_bc(const ogrid *g, odiag *dg, int *const nd[3]) {
  const int ni = g->n[0], nj = g->n[1], nk = g->n[2];
  int err = 0;
  #pragma omp parallel for schedule(static) reduction(| : err)
  for (long p = 0; p < dg->np; p++) {
    int i = iclamp(nd[0][p], 1, ni), j = iclamp(nd[1][p], 1, nj), k = iclamp(nd[2][p], 1, nk);
    nd[0][p] = i; nd[1][p] = j; nd[2][p] = k;
    dg->pflag[p] = LP_IB;
    long l = GID(g, i, j, k);
    int interior = (i > 1 && j > 1 && k > 1 && i < ni && j < nj && k < nk);
    switch (g->flag[l]) {
    case LCS_INTERNAL: case LCS_2D: break;
    case LCS_SLIP: case LCS_WALL:
      if (interior) break;
      /* fall through */
    case LCS_MASK: {
      double p2no[3], nrm[3];
      for (int d = 0; d < 3; d++) { p2no[d] = dg->xp[d][p] - g->g[d][l]; nrm[d] = g->bn[d][l]; }
      double dist = p2no[0] * nrm[0] + p2no[1] * nrm[1] + p2no[2] * nrm[2];
      if (dist > 0.0)
        for (int d = 0; d < 3; d++) dg->xp[d][p] = dg->xp[d][p] - nrm[d] * dist;
      break;
    }
    case LCS_INFLOW: case LCS_OUTFLOW:
      if (interior) break;
      dg->pflag[p] = LP_STICK;
      for (int d = 0; d < 3; d++) dg->xp[d][p] = g->g[d][l];
      break;
    default: err |= 1;
    }
  }
  return err;
}

/* one pass of the 27-neighbour nearest-node search (lp_tracking_m.f90:72-109);
 * returns 1 if the node moved */
static int track_pass(const ogrid *g, const double q[3], int n3[3]) {
  const int ni = g->n[0], nj = g->n[1], nk = g->n[2];
  double rsq_min = DBL_MAX;
  int off[3] = {0, 0, 0};
  for (int nb = 0; nb < 27; nb++) {
    if (ni == 1 && NBR[nb][0] != 0) continue;
    if (nj == 1 && NBR[nb][1] != 0) continue;
    if (nk == 1 && NBR[nb][2] != 0) continue;
    int i = iclamp(n3[0] + NBR[nb][0], 0, ni + 1);
    int j = iclamp(n3[1] + NBR[nb][1], 0, nj + 1);
    int k = iclamp(n3[2] + NBR[nb][2], 0, nk + 1);
    long l = GID(g, i, j, k);
    double a = g->g[0][l] - q[0], b = g->g[1][l] - q[1], c = g->g[2][l] - q[2];
    double rsq = a * a + b * b + c * c;
    if (rsq < 0.99 * rsq_min) {
      rsq_min = rsq;
      off[0] = NBR[nb][0]; off[1] = NBR[nb][1]; off[2] = NBR[nb][2];
    }
  }
  n3[0] += off[0]; n3[1] += off[1]; n3[2] += off[2];
  return off[0] != 0 || off[1] != 0 || off[2] != 0;
}

/* exchange_lpdata on one rank (comms_m.f90:872-1189) for tracking on grid
 * g: a particle whose tracked node n3 left [1,n] along periodic directions
 * is shifted by g's periodic_shift; lp%no is wrapped by g's extents and
 * lp%no_scfd by the flow grid's (comms_m.f90:1045-1073). */
static void wrap_particle(const ogrid *g, const ogrid *gf, odiag *dg, long p, const int n3[3]) {
  int c[3];
  for (int d = 0; d < 3; d++) c[d] = n3[d] < 1 ? -1 : (n3[d] > g->n[d] ? 1 : 0);
  if (c[0] == 0 && c[1] == 0 && c[2] == 0) return;
  if (!self_comm(g, c[0], c[1], c[2])) return;
  const double *sh = g->ps[DI(c[0], c[1], c[2])];
  for (int d = 0; d < 3; d++) {
    dg->xp[d][p] = dg->xp[d][p] + sh[d];
    int v = dg->no[d][p];
    if (v > g->n[d]) v -= g->n[d];
    if (v < 1) v += g->n[d];
    dg->no[d][p] = v;
    v = dg->ns[d][p];
    if (v > gf->n[d]) v -= gf->n[d];
    if (v < 1) v += gf->n[d];
    dg->ns[d][p] = v;
  }
  dg->pflag[p] = LP_IB;
}

/* reset_lp (lp_m.f90:228-294): particles on the nodes of the set's grid, the
 * flow-grid node guessed as nint(real(i*n_flow)/real(n)) (default REAL) */
static void reset_lp(ocfd *s, odiag *dg) {
  const ogrid *g = &dg->L->g, *gf = &s->F.g;
  for (int k = 1; k <= g->n[2]; k++)
    for (int j = 1; j <= g->n[1]; j++)
      for (int i = 1; i <= g->n[0]; i++) {
        long p = PIDG(g, i, j, k), l = GID(g, i, j, k);
        int ijk[3] = {i, j, k};
        for (int d = 0; d < 3; d++) {
          dg->xp[d][p] = g->g[d][l];
          dg->up[d][p] = 0; dg->dx[d][p] = 0;
          dg->no[d][p] = ijk[d];
          float r = (float)(ijk[d] * gf->n[d]) / (float)g->n[d];
          dg->ns[d][p] = iclamp(fnint((double)r), 1, gf->n[d]);
        }
        dg->pflag[p] = LP_UNKNOWN;
      }
  for (int d = 0; d < 3; d++) memset(dg->fm[d], 0, sizeof(double) * g->nt);
}

/* recursive track_lp2node of a reset set onto the flow grid
 * (lp_tracking_m.f90:16-183), with exchange_lpdata after every pass for
 * forward sets.  Passes repeat until no node moves; more than
 * 2*max(n) passes is an error.  A particle that did not move stays unmoved
 * in later passes, so each particle is run to its own fixed point.  On the
 * flow grid itself every particle sits on its node and nothing moves. */
static int track_to_flow(ocfd *s, odiag *dg) {
  const ogrid *gf = &s->F.g;
  const int max_track = 2 * imax3(gf->n[0], gf->n[1], gf->n[2]);
  int err = 0;
  #pragma omp parallel for schedule(static) reduction(| : err)
  for (long p = 0; p < dg->np; p++) {
    for (int pass = 1;; pass++) {
      double q[3] = {dg->xp[0][p], dg->xp[1][p], dg->xp[2][p]};
      int n3[3] = {dg->ns[0][p], dg->ns[1][p], dg->ns[2][p]};
      int moved = track_pass(gf, q, n3);
      for (int d = 0; d < 3; d++) dg->ns[d][p] = n3[d];
      if (pass > max_track) { err |= 1; break; }
      if (dg->dir == DIR_FWD) wrap_particle(gf, gf, dg, p, n3);
      if (!moved) break;
    }
  }
  return err;
}

/* set_dt (lp_motion_m.f90:56-127) */
static double set_dt(ocfd *s, odiag *dg, int *nsub) {
  const ogrid *g = &s->F.g;
  double dt_f = s->t_np1 - s->t_n;
  double mx = -DBL_MAX;
  for (int k = 1; k <= g->n[2]; k++)
    for (int j = 1; j <= g->n[1]; j++)
      for (int i = 1; i <= g->n[0]; i++) {
        long l = GID(g, i, j, k);
        double c = 0.0;
        c = c + fabs(s->unp1[0][l] * g->dl[0][l]);
        c = c + fabs(s->unp1[1][l] * g->dl[1][l]);
        c = c + fabs(s->unp1[2][l] * g->dl[2][l]);
        c = c * dt_f;
        if (c > mx) mx = c;
      }
  int n = (int)ceil(mx / (s->cfl_max * dg->dt_factor));
  if (n < 1) n = 1;
  *nsub = n;
  double dt = dt_f / (double)n;
  if (dg->dir == DIR_BKWD) dt = -1.0 * dt;
  if (fabs(dt) > 0.0) {
    if (dg->dir == DIR_FWD) s->integ_fwd += (long long)dg->np * n;
    if (dg->dir == DIR_BKWD) s->integ_bkwd += (long long)dg->np * n;
  }
  s->nsub_last = n;
  return dt;
}

/* blended velocity u = (1-ct) u_n + ct u_np1 over the ghosted grid */
static void blend(ocfd *s, double t, double c_dt) {
  double ct = clamp01((t + c_dt - s->t_n) / (s->t_np1 - s->t_n));
  long nt = s->F.g.nt;
  for (int d = 0; d < 3; d++) {
    double *w = s->work[d];
    const double *a = s->un[d], *b = s->unp1[d];
    #pragma omp parallel for schedule(static)
    for (long l = 0; l < nt; l++) w[l] = (1.0 - ct) * a[l] + ct * b[l];
  }
}

/* integrate_lp (lp_motion_m.f90:129-360) */
static int integrate_lp(ocfd *s, odiag *dg, double *t, double dt) {
  const long np = dg->np;
  const double ONEHALF = 1.0 / 2.0, ONETHIRD = 1.0 / 3.0, TWOTHIRD = 2.0 / 3.0,
               ONESIXTH = 1.0 / 6.0, TWO = 2.0;
  double *xn[3], *kk[3], *xnp1[3];
  for (int d = 0; d < 3; d++) {
    xn[d] = malloc(np * sizeof(double));
    kk[d] = malloc(np * sizeof(double));
    xnp1[d] = malloc(np * sizeof(double));
    memcpy(xn[d], dg->xp[d], np * sizeof(double));
    memcpy(xnp1[d], xn[d], np * sizeof(double));
  }
  int err = 0;
  double *const *u = s->work;
#define INTERP() err |= interp_field(s, &s->F, np, u, dg->xp, dg->ns, dg->up)
#define PLOOP for (long p = 0; p < np; p++) for (int d = 0; d < 3; d++)
  switch (s->integrator) {
  case EULER:
    blend(s, *t, 0.0); INTERP();
    PLOOP dg->xp[d][p] = dg->xp[d][p] + dt * dg->up[d][p];
    break;
  case TRAPEZOIDAL:
    blend(s, *t, ONEHALF * dt); INTERP();
    PLOOP dg->xp[d][p] = dg->xp[d][p] + dt * dg->up[d][p];
    break;
  case RK2:
    blend(s, *t, 0.0); INTERP();
    PLOOP dg->xp[d][p] = dg->xp[d][p] + ONEHALF * dt * dg->up[d][p];
    blend(s, *t, ONEHALF * dt); INTERP();
    PLOOP dg->xp[d][p] = xn[d][p] + dt * dg->up[d][p];
    break;
  case RK3: /* k3 at xn + 2 k2: the xn - k1 assignment is overwritten (lp_motion_m.f90:239-247) */
    blend(s, *t, 0.0); INTERP();
    PLOOP { kk[d][p] = dt * dg->up[d][p]; xnp1[d][p] = xnp1[d][p] + ONESIXTH * kk[d][p]; }
    blend(s, *t, ONEHALF * dt);
    PLOOP dg->xp[d][p] = xn[d][p] + ONEHALF * kk[d][p];
    INTERP();
    PLOOP {
      dg->xp[d][p] = xn[d][p] - kk[d][p];
      kk[d][p] = dt * dg->up[d][p];
      dg->xp[d][p] = xn[d][p] + TWO * kk[d][p];
      xnp1[d][p] = xnp1[d][p] + TWOTHIRD * kk[d][p];
    }
    blend(s, *t, dt); INTERP();
    PLOOP { kk[d][p] = dt * dg->up[d][p]; dg->xp[d][p] = xnp1[d][p] + ONESIXTH * kk[d][p]; }
    break;
  case RK4:
    blend(s, *t, 0.0); INTERP();
    PLOOP { kk[d][p] = dt * dg->up[d][p]; xnp1[d][p] = xnp1[d][p] + ONESIXTH * kk[d][p]; }
    blend(s, *t, ONEHALF * dt);
    PLOOP dg->xp[d][p] = xn[d][p] + ONEHALF * kk[d][p];
    INTERP();
    PLOOP { kk[d][p] = dt * dg->up[d][p]; xnp1[d][p] = xnp1[d][p] + ONETHIRD * kk[d][p]; }
    blend(s, *t, ONEHALF * dt);
    PLOOP dg->xp[d][p] = xn[d][p] + ONEHALF * kk[d][p];
    INTERP();
    PLOOP { kk[d][p] = dt * dg->up[d][p]; xnp1[d][p] = xnp1[d][p] + ONETHIRD * kk[d][p]; }
    blend(s, *t, dt);
    PLOOP dg->xp[d][p] = xn[d][p] + kk[d][p];
    INTERP();
    PLOOP { kk[d][p] = dt * dg->up[d][p]; dg->xp[d][p] = xnp1[d][p] + ONESIXTH * kk[d][p]; }
    break;
  default: err = 1;
  }
#undef INTERP
#undef PLOOP
  *t = *t + dt;
  /* the flow grid for the BC (lp_motion_m.f90:333-337) */
  project_lp2plane(&s->F.g, dg, dg->ns);
  err |= set_lp_bc(&s->F.g, dg, dg->ns);
  for (long p = 0; p < np; p++) {
    if (dg->pflag[p] == LP_STICK)
      for (int d = 0; d < 3; d++) dg->xp[d][p] = xn[d][p];
    for (int d = 0; d < 3; d++) dg->dx[d][p] = dg->dx[d][p] + (dg->xp[d][p] - xn[d][p]);
  }
  for (int d = 0; d < 3; d++) { free(xn[d]); free(kk[d]); free(xnp1[d]); }
  return err;
}

/* update_lp (lp_motion_m.f90:15-54): forward particles */
static int update_lp(ocfd *s, odiag *dg) {
  const ogrid *g = &s->F.g;
  int nsub, err = 0;
  double dt = set_dt(s, dg, &nsub);
  double t = s->t_n;
  for (int sc = 0; sc < nsub; sc++) {
    err |= integrate_lp(s, dg, &t, dt);
    /* track_lp2node (non-recursive) + exchange_lpdata on no_scfd */
    #pragma omp parallel for schedule(static)
    for (long p = 0; p < dg->np; p++) {
      double q[3] = {dg->xp[0][p], dg->xp[1][p], dg->xp[2][p]};
      int n3[3] = {dg->ns[0][p], dg->ns[1][p], dg->ns[2][p]};
      track_pass(g, q, n3);
      for (int d = 0; d < 3; d++) dg->ns[d][p] = n3[d];
      wrap_particle(g, g, dg, p, n3);
    }
  }
  return err;
}

static void fm_ghost_fill(ocfd *s, const ogrid *g, double *f[3]) {
  for (int d = 0; d < 3; d++) exchange(g, f[d], -1, NULL, 0);
  if (set_flowmap_bc(g, f)) s->error = 1;
}

/* update_flowmap_sl (flowmap_m.f90:17-104): backward semi-Lagrangian step.
 * Velocity, BC and mask on the flow grid through lp%no_scfd; the flowmap on
 * the set's grid through lp%no. */
static int update_flowmap_sl(ocfd *s, odiag *dg) {
  const ogrid *g = &dg->L->g, *gf = &s->F.g;
  const long np = dg->np;
  int nsub, err = 0;
  double *lpg[3], *fmp[3], *mask = malloc(np * sizeof(double));
  for (int d = 0; d < 3; d++) {
    lpg[d] = malloc(np * sizeof(double));
    fmp[d] = malloc(np * sizeof(double));
    memcpy(lpg[d], dg->xp[d], np * sizeof(double));
  }
  double dt = set_dt(s, dg, &nsub);
  for (long p = 0; p < np; p++)
    mask[p] = gf->flag[GID(gf, dg->ns[0][p], dg->ns[1][p], dg->ns[2][p])] != LCS_MASK ? 1.0 : 0.0;
  double t = s->t_np1;
  for (int sc = 0; sc < nsub; sc++) {
    err |= integrate_lp(s, dg, &t, dt);
    err |= interp_field(s, dg->L, np, dg->fm, dg->xp, dg->no, fmp);
    for (long p = 0; p < np; p++) {
      long l = GID(g, dg->no[0][p], dg->no[1][p], dg->no[2][p]);
      for (int d = 0; d < 3; d++)
        dg->fm[d][l] = (fmp[d][p] + (dg->xp[d][p] - lpg[d][p])) * mask[p];
    }
    fm_ghost_fill(s, g, dg->fm);
    for (int d = 0; d < 3; d++) memcpy(dg->xp[d], lpg[d], np * sizeof(double));
  }
  for (int d = 0; d < 3; d++) { free(lpg[d]); free(fmp[d]); }
  free(mask);
  return err;
}

/* exchange_lp_alltoall on one rank (comms_m.f90:1378-1709) + recursive
 * track_lp2node with exchange_lpdata after every pass (the direction is set
 * to FWD during reconstruction, flowmap_m.f90:211). */
static int relocate_and_track(ocfd *s, odiag *dg, long p) {
  const ogrid *g = &dg->L->g;
  const double *bb = g->bb;
#define INBB(x, y, z) ((x) >= bb[0] && (x) <= bb[3] && (y) >= bb[1] && (y) <= bb[4] && (z) >= bb[2] && (z) <= bb[5])
  double x = dg->xp[0][p], y = dg->xp[1][p], z = dg->xp[2][p];
  if (!INBB(x, y, z)) {
    int done = 0;
    for (int k = -1; k <= 1 && !done; k++)
      for (int j = -1; j <= 1 && !done; j++)
        for (int i = -1; i <= 1 && !done; i++) {
          const double *sh = g->ps[DI(i, j, k)];
          if (INBB(x - sh[0], y - sh[1], z - sh[2])) {
            dg->xp[0][p] = x - sh[0]; dg->xp[1][p] = y - sh[1]; dg->xp[2][p] = z - sh[2];
            done = 1;
          }
        }
  }
#undef INBB
  double stu[3];
  for (int d = 0; d < 3; d++) {
    double sv = (dg->xp[d][p] - bb[d]) / (bb[3 + d] - bb[d]);
    stu[d] = fmax(fmin(sv, 1.0), 0.0);
  }
  for (int d = 0; d < 3; d++) {
    int v;
    if (g->rect) {
      v = fnint((double)(float)g->n[d] * stu[d]);
    } else {
      /* non-rectilinear guess from the positions of the coordinate extrema
       * (comms_m.f90:1666-1681) */
      const double third = 1.0 / 3.0;
      v = fnint(third * (g->pmin[0][d] + g->pmin[1][d] + g->pmin[2][d])) +
          fnint(third * (stu[0] * (g->pmax[0][d] - g->pmin[0][d]) + stu[1] * (g->pmax[1][d] - g->pmin[1][d]) +
                         stu[2] * (g->pmax[2][d] - g->pmin[2][d])));
    }
    dg->no[d][p] = iclamp(v, 1, g->n[d]);
    dg->up[d][p] = 0.0;
  }
  dg->pflag[p] = LP_UNKNOWN;
  int max_track = 2 * imax3(g->n[0], g->n[1], g->n[2]);
  for (int pass = 1;; pass++) {
    double q[3] = {dg->xp[0][p], dg->xp[1][p], dg->xp[2][p]};
    int n3[3] = {dg->no[0][p], dg->no[1][p], dg->no[2][p]};
    int moved = track_pass(g, q, n3);
    for (int d = 0; d < 3; d++) dg->no[d][p] = n3[d];
    if (pass > max_track) return 1;
    wrap_particle(g, &s->F.g, dg, p, n3);
    if (!moved) break;
  }
  return 0;
}

/* reconstruct_flowmap (flowmap_m.f90:158-323) followed by exchange_lpmap
 * (comms_m.f90:1191-1376) and the flowmap BC */
static int reconstruct_flowmap(ocfd *s, odiag *dg) {
  const ogrid *g = &dg->L->g;
  const long np = dg->np;
  int err = 0;
  reset_lp(s, dg);
  int step = fnint(fmod(s->t_np1, dg->T) / dg->h);
  int nstep = dg->nstep, inc;
  if (dg->dir == DIR_FWD) { inc = 1; step = step + inc; } else inc = -1;
  for (int it = 0; it < nstep; it++) {
    if (step >= nstep) step = 0;
    if (step < 0) step = nstep - 1;
    if (!dg->slot_ok[step]) { step += inc; continue; }
    /* read the slot interior into lp%fm, then exchange + BC */
    for (int k = 1; k <= g->n[2]; k++)
      for (int j = 1; j <= g->n[1]; j++)
        for (int i = 1; i <= g->n[0]; i++)
          for (int d = 0; d < 3; d++)
            dg->fm[d][GID(g, i, j, k)] = dg->slot[3 * step + d][PIDG(g, i, j, k)];
    fm_ghost_fill(s, g, dg->fm);
    project_lp2plane(g, dg, dg->no);
    err |= set_lp_bc(g, dg, dg->no);
    err |= interp_field(s, dg->L, np, dg->fm, dg->xp, dg->no, dg->up);
    int terr = 0;
    #pragma omp parallel for schedule(static) reduction(| : terr)
    for (long p = 0; p < np; p++) {
      for (int d = 0; d < 3; d++) {
        dg->xp[d][p] = dg->xp[d][p] + dg->up[d][p];
        dg->dx[d][p] = dg->dx[d][p] + dg->up[d][p];
      }
      terr |= relocate_and_track(s, dg, p);
    }
    err |= terr;
    step += inc;
  }
  /* exchange_lpmap: fm(origin) = dx */
  for (int k = 1; k <= g->n[2]; k++)
    for (int j = 1; j <= g->n[1]; j++)
      for (int i = 1; i <= g->n[0]; i++)
        for (int d = 0; d < 3; d++)
          dg->fm[d][GID(g, i, j, k)] = dg->dx[d][PIDG(g, i, j, k)];
  fm_ghost_fill(s, g, dg->fm);
  return err;
}

/* compute_ftle (lcs_m.f90:66-193).  The gradient of the composed
 * displacement is grad_sr1 (gradient_m.f90:11-66) on the diagnostic's grid,
 * or with AUX_GRID the difference of the X1/X0, Y1/Y0, Z1/Z0 auxiliary
 * flowmaps over the difference of their node coordinates, a column left at
 * zero where that difference is not above 100*tiny (lcs_m.f90:95-126). */
static void compute_ftle(ocfd *s, odiag *dg) {
  const ogrid *g = &dg->L->g;
  const int ni = g->n[0], nj = g->n[1], nk = g->n[2];
  if (!dg->naux)
    for (int d = 0; d < 3; d++) exchange(g, dg->fm[d], -1, NULL, 1);
  double *const *F = dg->fm;
  double *const *X = g->g;
  const double SMALL = 100.0 * DBL_MIN;
  #pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        double G[9];
        if (dg->naux) {
          const long l = GID(g, i, j, k);
          for (int a = 0; a < 9; a++) G[a] = 0.0;
          for (int ax = 0; ax < 3; ax++) {
            const odiag *A0 = dg->auxs[ax], *A1 = dg->auxs[3 + ax];
            const double den = A1->L->g.g[ax][l] - A0->L->g.g[ax][l];
            if (fabs(den) > SMALL)
              for (int c = 0; c < 3; c++) G[3 * c + ax] = (A1->fm[c][l] - A0->fm[c][l]) / den;
          }
        } else {
          long ip = GID(g, i + 1, j, k), im = GID(g, i - 1, j, k);
          long jp = GID(g, i, j + 1, k), jm = GID(g, i, j - 1, k);
          long kp = GID(g, i, j, k + 1), km = GID(g, i, j, k - 1);
          if (!g->rect) {
            /* grad_sr1_ls on a non-rectilinear grid (lcs_m.f90:131-133) */
            const long nbr[6] = {im, jm, km, ip, jp, kp};
            for (int c = 0; c < 3; c++) grad_ls(g, F[c], GID(g, i, j, k), nbr, &G[3 * c]);
          } else {
            for (int c = 0; c < 3; c++) {
              G[3 * c + 0] = (F[c][ip] - F[c][im]) / (X[0][ip] - X[0][im]);
              G[3 * c + 1] = (F[c][jp] - F[c][jm]) / (X[1][jp] - X[1][jm]);
              G[3 * c + 2] = (F[c][kp] - F[c][km]) / (X[2][kp] - X[2][km]);
            }
            if (ni == 1) { G[0] = 0; G[3] = 0; G[6] = 0; }
            if (nj == 1) { G[1] = 0; G[4] = 0; G[7] = 0; }
            if (nk == 1) { G[2] = 0; G[5] = 0; G[8] = 0; }
          }
        }
        double C[9];
        oc_cauchy_green(G, C);
        double lam = oc_eig_sym3_max(C);
        double r = log(fmax(sqrt(lam), DBL_MIN)) / dg->T;
        if (s->incompressible) r = fmax(r, 0.0);
        long p = PIDG(g, i, j, k);
        dg->ftle[p] = r;
        for (int c = 0; c < 3; c++) dg->comp[c][p] = F[c][GID(g, i, j, k)];
      }
  dg->noutput++;
}

/* write_flowmap_substep (flowmap_m.f90:106-156): slot step =
 * nint(mod(t_np1,T)/h), not wrapped (slot nstep is written but never read) */
static void write_slot(ocfd *s, odiag *dg) {
  const ogrid *g = &dg->L->g;
  int step = fnint(fmod(s->t_np1, dg->T) / dg->h);
  if (step < 0 || step >= dg->nslot) { s->error = 1; return; }
  for (int k = 1; k <= g->n[2]; k++)
    for (int j = 1; j <= g->n[1]; j++)
      for (int i = 1; i <= g->n[0]; i++)
        for (int d = 0; d < 3; d++)
          dg->slot[3 * step + d][PIDG(g, i, j, k)] = dg->fm[d][GID(g, i, j, k)];
  dg->slot_ok[step] = 1;
}

/* ======================================================================
 * API
 * ====================================================================== */
/* a grid bundle from interior coordinate arrays (init_sgrid, sgrid_m.f90:16-444) */
static int fg_init(ofg *G, const int n[3], const double *x, const double *y, const double *z,
                   const int *flag) {
  memset(G, 0, sizeof(*G));
  if (grid_init(&G->g, n, x, y, z, flag)) return 1;
  const ogrid *g = &G->g;
  for (int d = 0; d < 3; d++) {
    G->bsc[d] = calloc(g->nt, sizeof(double));
    G->bst[d] = calloc(g->N[d] + 8, sizeof(double));
    G->xl[d] = calloc(g->N[d], sizeof(double));
    for (int a = 0; a < g->N[d]; a++) {
      int i = d == 0 ? a : 1, j = d == 1 ? a : 1, k = d == 2 ? a : 1;
      G->xl[d][a] = g->g[d][GID(g, i, j, k)];
    }
    if (g->n[d] == 1) { G->xl[d][0] = G->xl[d][1] - 1.0; G->xl[d][2] = G->xl[d][1] + 1.0; }
  }
  return 0;
}

static void fg_free(ofg *G) {
  if (!G->g.g[0]) return;
  for (int d = 0; d < 3; d++) { free(G->bsc[d]); free(G->bst[d]); free(G->xl[d]); }
  grid_free(&G->g);
}

ocfd *oc_init(const int n[3], const double *x, const double *y,
              const double *z, const int *flag) {
  ocfd *s = calloc(1, sizeof(ocfd));
  s->integrator = RK2;
  s->interpolator = LINEAR;
  s->cfl_max = 0.5;
  s->first_call = 1;
  if (fg_init(&s->F, n, x, y, z, flag)) { s->error = 1; return s; }
  const ogrid *g = &s->F.g;
  for (int d = 0; d < 3; d++) {
    s->un[d] = calloc(g->nt, sizeof(double));
    s->unp1[d] = calloc(g->nt, sizeof(double));
    s->work[d] = calloc(g->nt, sizeof(double));
  }
  return s;
}

static void set_free(odiag *dg) {
  for (int d = 0; d < 3; d++) {
    free(dg->xp[d]); free(dg->dx[d]); free(dg->up[d]); free(dg->no[d]); free(dg->ns[d]);
    free(dg->fm[d]); free(dg->comp[d]);
  }
  for (int k = 0; k < 3 * dg->nslot; k++) free(dg->slot[k]);
  free(dg->slot); free(dg->slot_ok); free(dg->pflag); free(dg->ftle);
  for (int a = 0; a < dg->naux; a++) set_free(dg->auxs[a]);
  if (dg->own_grid) { fg_free(dg->L); free(dg->L); }
  free(dg);
}

void oc_free(ocfd *s) {
  if (!s) return;
  for (int i = 0; i < s->ndiag; i++) set_free(s->d[i]);
  if (s->F.g.g[0]) {
    for (int d = 0; d < 3; d++) { free(s->un[d]); free(s->unp1[d]); free(s->work[d]); }
    fg_free(&s->F);
  }
  free(s);
}

int oc_error(const ocfd *s) { return s->error; }

/* new_sgrid_from_sgrid (sgrid_m.f90:445-761) on one rank for a rectilinear
 * grid g.  res > 0 inserts res points into every cell, including the wrap
 * cell after the last node of a periodic axis but not after the last node of
 * a non-periodic one, at x(i) + (x(i+1) - x(i))/(res+1) * r with the flag of
 * node i; res < 0 keeps the nodes whose 1-based index is a multiple of
 * |res|+1.  An axis with one node is kept.  |res| >= n along an axis with
 * n > 1 resets res to 0 (the flow grid itself).  Returns the effective res;
 * the interior arrays are allocated here. */
static int refine_grid(const ogrid *g, int res, int nn[3], double *xyz[3], int **flag) {
  for (int d = 0; d < 3; d++)
    if (g->n[d] > 1 && abs(res) >= g->n[d]) res = 0;
  if (res == 0) return 0;
  int *map[3], *sub[3];
  for (int d = 0; d < 3; d++) {
    const int n = g->n[d];
    int m;
    if (n == 1) m = 1;
    else if (res > 0) m = g->per[d] ? n * (res + 1) : n * (res + 1) - res;
    else m = n / (abs(res) + 1);
    nn[d] = m;
    map[d] = malloc(sizeof(int) * m);  /* source node of each new index */
    sub[d] = malloc(sizeof(int) * m);  /* inserted point number r */
    for (int a = 0; a < m; a++) {
      if (n == 1) { map[d][a] = 1; sub[d][a] = 0; }
      else if (res > 0) { map[d][a] = 1 + a / (res + 1); sub[d][a] = a % (res + 1); }
      else { map[d][a] = (a + 1) * (abs(res) + 1); sub[d][a] = 0; }
    }
  }
  const long np = (long)nn[0] * nn[1] * nn[2];
  for (int d = 0; d < 3; d++) xyz[d] = malloc(sizeof(double) * np);
  *flag = malloc(sizeof(int) * np);
  for (int k = 0; k < nn[2]; k++)
    for (int j = 0; j < nn[1]; j++)
      for (int i = 0; i < nn[0]; i++) {
        const long p = (long)i + (long)nn[0] * ((long)j + (long)nn[1] * k);
        const int src[3] = {map[0][i], map[1][j], map[2][k]}, r[3] = {sub[0][i], sub[1][j], sub[2][k]};
        const long l = GID(g, src[0], src[1], src[2]);
        for (int d = 0; d < 3; d++) {
          if (res > 0) {
            int up[3] = {src[0], src[1], src[2]};
            up[d] += 1;
            const double dx = (g->g[d][GID(g, up[0], up[1], up[2])] - g->g[d][l]) / (double)(res + 1);
            xyz[d][p] = g->g[d][l] + dx * (double)r[d];
          } else {
            xyz[d][p] = g->g[d][l];
          }
        }
        (*flag)[p] = g->flag[l];
      }
  for (int d = 0; d < 3; d++) { free(map[d]); free(sub[d]); }
  return res;
}

/* create_aux_grid (lcs_m.f90:227-435): node coordinates of auxiliary grid
 * a = 0..5 (X0, Y0, Z0, X1, Y1, Z1) of grid g, shifted along axis a % 3 by
 * -/+ aux_delta, aux_delta being AUX_GRID_SCALING times the mean distance to
 * the in-range face neighbours (faces in NBR order), then projected onto
 * the data plane of a 2-D grid.  The shift starts from the node coordinate
 * (the reference starts from the particle, which the tracking onto the flow
 * grid leaves on its node except where a forward particle of a refined
 * periodic grid is wrapped by a period). */
static void aux_coords(const ogrid *g, int a, double *xyz[3]) {
  const int ni = g->n[0], nj = g->n[1], nk = g->n[2];
  const int flat = ni == 1 || nj == 1 || nk == 1;
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        const long l = GID(g, i, j, k), p = PIDG(g, i, j, k);
        double nbrsum = 0.0, distsum = 0.0;
        for (int nb = 1; nb <= 6; nb++) {
          const int ii = i + NBR[nb][0], jj = j + NBR[nb][1], kk = k + NBR[nb][2];
          if (ii > 0 && jj > 0 && kk > 0 && ii <= ni && jj <= nj && kk <= nk) {
            const long m = GID(g, ii, jj, kk);
            nbrsum = nbrsum + 1.0;
            const double dx = g->g[0][l] - g->g[0][m], dy = g->g[1][l] - g->g[1][m],
                         dz = g->g[2][l] - g->g[2][m];
            distsum = distsum + sqrt(dx * dx + dy * dy + dz * dz);
          }
        }
        const double delta = distsum / nbrsum * AUX_GRID_SCALING;
        double q[3] = {g->g[0][l], g->g[1][l], g->g[2][l]};
        const int ax = a % 3;
        q[ax] = a < 3 ? q[ax] - delta : q[ax] + delta;
        if (flat) {
          const int n3[3] = {i, j, k};
          project_point(g, n3, q);
        }
        for (int d = 0; d < 3; d++) xyz[d][p] = q[d];
      }
}

/* one particle set on grid L (init_lp, lp_m.f90:13-114): seeded, tracked
 * onto the flow grid (init_ftle / create_aux_grid) */
static odiag *new_set(ocfd *s, ofg *L, int own, int lcs_type, double T, double h, double dt_factor,
                      const char *label) {
  const ogrid *g = &L->g;
  odiag *dg = calloc(1, sizeof(odiag));
  dg->type = lcs_type;
  dg->dir = lcs_type == FTLE_FWD ? DIR_FWD : DIR_BKWD;
  dg->T = T; dg->h = h;
  dg->dt_factor = dt_factor;
  dg->L = L;
  dg->own_grid = own;
  dg->np = g->np;
  snprintf(dg->label, sizeof(dg->label), "%s", label ? label : "");
  for (int d = 0; d < 3; d++) {
    dg->xp[d] = calloc(g->np, sizeof(double));
    dg->dx[d] = calloc(g->np, sizeof(double));
    dg->up[d] = calloc(g->np, sizeof(double));
    dg->no[d] = calloc(g->np, sizeof(int));
    dg->ns[d] = calloc(g->np, sizeof(int));
    dg->fm[d] = calloc(g->nt, sizeof(double));
    dg->comp[d] = calloc(g->np, sizeof(double));
  }
  dg->pflag = calloc(g->np, sizeof(int));
  dg->ftle = calloc(g->np, sizeof(double));
  dg->nstep = fnint(T / h);
  dg->nslot = dg->nstep + 1;
  dg->slot = calloc(3 * dg->nslot, sizeof(double *));
  for (int k = 0; k < 3 * dg->nslot; k++) dg->slot[k] = calloc(g->np, sizeof(double));
  dg->slot_ok = calloc(dg->nslot, sizeof(int));
  reset_lp(s, dg);
  if (track_to_flow(s, dg)) s->error = 1;
  return dg;
}

int oc_diagnostic_init(ocfd *s, int lcs_type, int resolution, double T,
                       double h, const char *label) {
  if (s->error) return 0;
  if (s->ndiag >= NMAX_STRUCT) { s->error = 1; return 0; }
  if (lcs_type != FTLE_FWD && lcs_type != FTLE_BKWD) { s->error = 1; return 0; }
  /* the diagnostic's grid (cfd2lcs.f90:298-310) */
  ofg *L = &s->F;
  int own = 0;
  /* new_sgrid_from_sgrid's trilinear refinement of a non-rectilinear grid
   * (sgrid_m.f90:666-683) is not restated */
  if (resolution != 0 && !s->F.g.rect) { s->error = 1; return 0; }
  if (resolution != 0) {
    int nn[3], *fl = NULL;
    double *xyz[3] = {NULL, NULL, NULL};
    resolution = refine_grid(&s->F.g, resolution, nn, xyz, &fl);
    if (resolution != 0) {
      L = calloc(1, sizeof(ofg));
      own = 1;
      const int e = fg_init(L, nn, xyz[0], xyz[1], xyz[2], fl);
      for (int d = 0; d < 3; d++) free(xyz[d]);
      free(fl);
      if (e) { free(L); s->error = 1; return 0; }
    }
  }
  const int dir = lcs_type == FTLE_FWD ? DIR_FWD : DIR_BKWD;
  const double dtf = dir == DIR_BKWD ? 1.0 / (double)(1 + resolution > 1 ? 1 + resolution : 1) : 1.0;
  char lab[140];
  snprintf(lab, sizeof(lab), "%s", label ? label : "");
  odiag *dg = new_set(s, L, own, lcs_type, T, h, dtf, lab);
  s->d[s->ndiag++] = dg;
  dg->id = s->ndiag;
  if (s->aux_grid) {
    const ogrid *g = &L->g;
    for (int a = 0; a < 6; a++) {
      double *xyz[3];
      int *fl = malloc(sizeof(int) * g->np);
      for (int d = 0; d < 3; d++) xyz[d] = malloc(sizeof(double) * g->np);
      aux_coords(g, a, xyz);
      for (int k = 1; k <= g->n[2]; k++)
        for (int j = 1; j <= g->n[1]; j++)
          for (int i = 1; i <= g->n[0]; i++) fl[PIDG(g, i, j, k)] = g->flag[GID(g, i, j, k)];
      ofg *A = calloc(1, sizeof(ofg));
      const int e = fg_init(A, g->n, xyz[0], xyz[1], xyz[2], fl);
      for (int d = 0; d < 3; d++) free(xyz[d]);
      free(fl);
      if (e) { free(A); s->error = 1; return 0; }
      dg->auxs[a] = new_set(s, A, 1, lcs_type, T, h, dtf, lab);
      dg->naux = a + 1;
    }
  }
  return s->error ? 0 : dg->id;
}

void oc_set_option(ocfd *s, const char *option, int val) {
  if (s->error) return;
  int warn = 0;
  if (!strcmp(option, "SYNCTIMER") || !strcmp(option, "DEBUG") ||
      !strcmp(option, "VELOCITY_INVARIANTS") || !strcmp(option, "WRITE_FLOWMAP") ||
      !strcmp(option, "WRITE_VELOCITY") || !strcmp(option, "WRITE_BCFLAG") ||
      !strcmp(option, "AUX_GRID")) {
    if (val != LCS_TRUE && val != LCS_FALSE) warn = 1;
    if (!strcmp(option, "AUX_GRID")) s->aux_grid = val == LCS_TRUE;
    if (!strcmp(option, "VELOCITY_INVARIANTS") && val == LCS_TRUE) warn = 1;
  } else if (!strcmp(option, "INCOMPRESSIBLE")) {
    if (val == LCS_TRUE) s->incompressible = 1;
    else if (val == LCS_FALSE) s->incompressible = 0;
    else warn = 1;
  } else if (!strcmp(option, "UPDATE_FREQ")) {
  } else if (!strcmp(option, "INTERPOLATOR")) {
    if (val != NEAREST_NBR && val != LINEAR && val != QUADRATIC && val != CUBIC && val != TSE &&
        val != TSE_LIMIT)
      warn = 1;
    s->interpolator = val;
  } else if (!strcmp(option, "INTEGRATOR")) {
    if (val < EULER || val > RK4) warn = 1;
    s->integrator = val;
  } else {
    warn = 1;
  }
  if (warn) s->warn_opt = 1;
  if (s->warn_opt) s->error = 1;
}

void oc_set_param(ocfd *s, const char *param, double val) {
  if (s->error) return;
  int warn = 0;
  if (!strcmp(param, "CFL")) s->cfl_max = val;
  else if (!strcmp(param, "TRACER_INJECT_X") || !strcmp(param, "TRACER_INJECT_Y") ||
           !strcmp(param, "TRACER_INJECT_Z") || !strcmp(param, "TRACER_INJECT_RADIUS")) {
  } else warn = 1;
  if (warn) s->warn_par = 1;
  if (s->warn_par) s->error = 1;
}

/* the sets of diagnostic dg: its own, then its auxiliary sets (the order of
 * lp_c, init_ftle + create_aux_grid) */
static int nsets(const odiag *dg) { return 1 + dg->naux; }
static odiag *set_at(odiag *dg, int a) { return a == 0 ? dg : dg->auxs[a - 1]; }

void oc_update(ocfd *s, const int n[3], const double *u, const double *v,
               const double *w, double time) {
  if (s->error) return;
  ogrid *g = &s->F.g;
  if (n[0] != g->n[0] || n[1] != g->n[1] || n[2] != g->n[2]) { s->error = 1; return; }
  if (s->first_call) { s->t_n = time; s->t_np1 = time; s->first_call = 0; }
  else { s->t_n = s->t_np1; s->t_np1 = time; }
  for (int d = 0; d < 3; d++) memcpy(s->un[d], s->unp1[d], g->nt * sizeof(double));
  const double *src[3] = {u, v, w};
  for (int k = 1; k <= g->n[2]; k++)
    for (int j = 1; j <= g->n[1]; j++)
      for (int i = 1; i <= g->n[0]; i++)
        for (int d = 0; d < 3; d++) s->unp1[d][GID(g, i, j, k)] = src[d][PIDG(g, i, j, k)];
  for (int d = 0; d < 3; d++) exchange(g, s->unp1[d], -1, NULL, 0);
  if (set_velocity_bc(g, s->unp1)) { s->error = 1; return; }
  for (int i = 0; i < s->ndiag; i++)
    for (int a = 0; a < nsets(s->d[i]); a++) {
      odiag *dg = set_at(s->d[i], a);
      int e = dg->dir == DIR_FWD ? update_lp(s, dg) : update_flowmap_sl(s, dg);
      if (e) { s->error = 1; return; }
    }
  for (int i = 0; i < s->ndiag; i++) {
    odiag *main = s->d[i];
    if ((long)(s->t_np1 / main->h) == (long)(s->t_n / main->h)) continue;
    for (int a = 0; a < nsets(main); a++) {
      odiag *dg = set_at(main, a);
      const ogrid *L = &dg->L->g;
      if (dg->dir == DIR_FWD) {
        for (int k = 1; k <= L->n[2]; k++)
          for (int j = 1; j <= L->n[1]; j++)
            for (int ii = 1; ii <= L->n[0]; ii++)
              for (int d = 0; d < 3; d++)
                dg->fm[d][GID(L, ii, j, k)] = dg->dx[d][PIDG(L, ii, j, k)];
      }
      write_slot(s, dg);
      if (s->error) return;
    }
    for (int a = 0; a < nsets(main); a++)
      if (reconstruct_flowmap(s, set_at(main, a))) { s->error = 1; return; }
    compute_ftle(s, main);
    for (int a = 0; a < nsets(main); a++) {
      odiag *dg = set_at(main, a);
      reset_lp(s, dg);
      if (track_to_flow(s, dg)) { s->error = 1; return; }
    }
  }
}

static odiag *get_diag(const ocfd *s, int handle) {
  if (handle < 1 || handle > s->ndiag) return NULL;
  return s->d[handle - 1];
}

int oc_get_flowmap(const ocfd *s, int handle, double *fx, double *fy, double *fz) {
  odiag *dg = get_diag(s, handle);
  if (!dg) return 1;
  const ogrid *g = &dg->L->g;
  double *o[3] = {fx, fy, fz};
  for (int k = 1; k <= g->n[2]; k++)
    for (int j = 1; j <= g->n[1]; j++)
      for (int i = 1; i <= g->n[0]; i++)
        for (int d = 0; d < 3; d++) {
          long p = PIDG(g, i, j, k);
          o[d][p] = dg->dir == DIR_FWD ? dg->dx[d][p] : dg->fm[d][GID(g, i, j, k)];
        }
  return 0;
}

int oc_get_composed(const ocfd *s, int handle, double *fx, double *fy, double *fz) {
  odiag *dg = get_diag(s, handle);
  if (!dg) return 1;
  memcpy(fx, dg->comp[0], dg->np * sizeof(double));
  memcpy(fy, dg->comp[1], dg->np * sizeof(double));
  memcpy(fz, dg->comp[2], dg->np * sizeof(double));
  return 0;
}

int oc_get_ftle(const ocfd *s, int handle, double *ftle) {
  odiag *dg = get_diag(s, handle);
  if (!dg) return 1;
  memcpy(ftle, dg->ftle, dg->np * sizeof(double));
  return 0;
}

int oc_get_lcs_grid(const ocfd *s, int handle, int set, int n[3], double *x, double *y, double *z) {
  odiag *dg = get_diag(s, handle);
  if (!dg || set < 0 || set > dg->naux) return 1;
  dg = set_at(dg, set);
  const ogrid *g = &dg->L->g;
  for (int d = 0; d < 3; d++) n[d] = g->n[d];
  double *o[3] = {x, y, z};
  for (int k = 1; k <= g->n[2]; k++)
    for (int j = 1; j <= g->n[1]; j++)
      for (int i = 1; i <= g->n[0]; i++)
        for (int d = 0; d < 3; d++)
          if (o[d]) o[d][PIDG(g, i, j, k)] = g->g[d][GID(g, i, j, k)];
  return 0;
}

int oc_get_noutput(const ocfd *s, int handle) {
  odiag *dg = get_diag(s, handle);
  return dg ? dg->noutput : -1;
}

int oc_get_nsub(const ocfd *s) { return s->nsub_last; }

long long oc_get_integrations(const ocfd *s, int direction) {
  return direction == DIR_FWD ? s->integ_fwd : s->integ_bkwd;
}

int oc_get_ghost_grid(const ocfd *s, double *gx, double *gy, double *gz,
                      int *flag, double *bnx, double *bny, double *bnz) {
  const ogrid *g = &s->F.g;
  if (!g->g[0]) return 1;
  size_t b = g->nt * sizeof(double);
  if (gx) memcpy(gx, g->g[0], b);
  if (gy) memcpy(gy, g->g[1], b);
  if (gz) memcpy(gz, g->g[2], b);
  if (flag) memcpy(flag, g->flag, g->nt * sizeof(int));
  if (bnx) memcpy(bnx, g->bn[0], b);
  if (bny) memcpy(bny, g->bn[1], b);
  if (bnz) memcpy(bnz, g->bn[2], b);
  return 0;
}

int oc_get_periodic_shift(const ocfd *s, double *ps, double *bb) {
  if (ps) memcpy(ps, s->F.g.ps, sizeof(s->F.g.ps));
  if (bb) memcpy(bb, s->F.g.bb, sizeof(s->F.g.bb));
  return 0;
}

/* ======================================================================
 * compute_invariants (invariants_m.f90:15-198) for one velocity field at T0
 * (lp%ugrid; u, v, w interior arrays or NULL for zero).  ugrid's ghosts are
 * zero (reset_lp, lp_m.f90:288-290) except where the face exchange
 * (invariants_m.f90:37) copies the periodic image; grad_sr1
 * (gradient_m.f90:11-66) then gives G[c][d] = d u_c / d x_d.  With
 * S = G + G^T (S_ab = d_a u_b + d_b u_a) and W_ab = d_a u_b - d_b u_a:
 *   QS = -tr(SS)/8, RS = -tr(SSS) * c24 with c24 the REAL(4) value of 1./24.
 *   Q  = QS - tr(WW)/8
 *   R  = RS - (tr(WWS) + W_23 W_31 (S_13 - S_12))/8: the reference's sum
 *        (invariants_m.f90:105-116) has S13 where tr(WWS) has S12 in the
 *        W_23 W_31 term
 *   outputs Q>0, RS>0, the D class of (Q^3 + 27/4 R^2, R), sqrt(-(Q-QS)/QS)
 *   and [middle eigenvalue of SS+WW with its (3,3) entry zero < 0]
 *   (invariants_m.f90:164-194).
 * out: Q, C, D, H, L2 (np each).
 * ====================================================================== */
static void mat3_mul(const double a[3][3], const double b[3][3], double c[3][3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double v = 0.0;
      for (int k = 0; k < 3; k++) v += a[i][k] * b[k][j];
      c[i][j] = v;
    }
}

void oc_velocity_invariants(const ocfd *s, const double *u, const double *v, const double *w, double *out) {
  const ogrid *g = &s->F.g;
  const long nt = g->nt, np = g->np;
  double *ug[3];
  const double *src[3] = {u, v, w};
  for (int c = 0; c < 3; c++) {
    ug[c] = (double *)calloc((size_t)nt, sizeof(double));
    if (!src[c]) continue;
    for (int k = 1; k <= g->n[2]; k++)
      for (int j = 1; j <= g->n[1]; j++)
        for (int i = 1; i <= g->n[0]; i++)
          ug[c][GID(g, i, j, k)] = src[c][(i - 1) + (long)g->n[0] * ((j - 1) + (long)g->n[1] * (k - 1))];
    /* face ghosts: periodic images */
    for (int k = 1; k <= g->n[2]; k++)
      for (int j = 1; j <= g->n[1]; j++) {
        if (g->per[0]) {
          ug[c][GID(g, 0, j, k)] = ug[c][GID(g, g->n[0], j, k)];
          ug[c][GID(g, g->n[0] + 1, j, k)] = ug[c][GID(g, 1, j, k)];
        }
      }
    for (int k = 1; k <= g->n[2]; k++)
      for (int i = 1; i <= g->n[0]; i++) {
        if (g->per[1]) {
          ug[c][GID(g, i, 0, k)] = ug[c][GID(g, i, g->n[1], k)];
          ug[c][GID(g, i, g->n[1] + 1, k)] = ug[c][GID(g, i, 1, k)];
        }
      }
    for (int j = 1; j <= g->n[1]; j++)
      for (int i = 1; i <= g->n[0]; i++) {
        if (g->per[2]) {
          ug[c][GID(g, i, j, 0)] = ug[c][GID(g, i, j, g->n[2])];
          ug[c][GID(g, i, j, g->n[2] + 1)] = ug[c][GID(g, i, j, 1)];
        }
      }
  }
  const double c24 = (double)(1.0f / 24.0f);
  for (int k = 1; k <= g->n[2]; k++)
    for (int j = 1; j <= g->n[1]; j++)
      for (int i = 1; i <= g->n[0]; i++) {
        const long p = (i - 1) + (long)g->n[0] * ((j - 1) + (long)g->n[1] * (k - 1));
        const long nb[3][2] = {{GID(g, i + 1, j, k), GID(g, i - 1, j, k)},
                               {GID(g, i, j + 1, k), GID(g, i, j - 1, k)},
                               {GID(g, i, j, k + 1), GID(g, i, j, k - 1)}};
        double G[3][3];
        for (int c = 0; c < 3; c++)
          for (int d = 0; d < 3; d++) {
            const double den = g->g[d][nb[d][0]] - g->g[d][nb[d][1]];
            G[c][d] = g->n[d] == 1 ? 0.0 : (ug[c][nb[d][0]] - ug[c][nb[d][1]]) / den;
          }
        double S[3][3], W[3][3], SS[3][3], WW[3][3], SSS[3][3], WWS[3][3];
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) {
            S[a][b] = G[b][a] + G[a][b];
            W[a][b] = G[b][a] - G[a][b];
          }
        mat3_mul(S, S, SS);
        mat3_mul(W, W, WW);
        mat3_mul(SS, S, SSS);
        mat3_mul(WW, S, WWS);
        const double trSS = SS[0][0] + SS[1][1] + SS[2][2];
        const double trWW = WW[0][0] + WW[1][1] + WW[2][2];
        const double trSSS = SSS[0][0] + SSS[1][1] + SSS[2][2];
        const double trWWS = WWS[0][0] + WWS[1][1] + WWS[2][2] + W[1][2] * W[2][0] * (S[0][2] - S[0][1]);
        const double qs = -0.125 * trSS, rs = -c24 * trSSS;
        const double q = qs - 0.125 * trWW, r = rs - 0.125 * trWWS;
        const double dd = q * q * q + 6.75 * (r * r);
        double cls;
        if (dd > 0.0) cls = r > 0.0 ? 1.0 : 2.0;
        else cls = r > 0.0 ? 3.0 : 4.0;
        double A[9] = {SS[0][0] + WW[0][0], SS[0][1] + WW[0][1], SS[0][2] + WW[0][2],
                       SS[0][1] + WW[0][1], SS[1][1] + WW[1][1], SS[1][2] + WW[1][2],
                       SS[0][2] + WW[0][2], SS[1][2] + WW[1][2], 0.0};
        double ev[3];
        eig_sym3_all(A, ev);
        /* sort3 */
        for (int a = 0; a < 2; a++)
          for (int b = 0; b < 2 - a; b++)
            if (ev[b] > ev[b + 1]) { double t = ev[b]; ev[b] = ev[b + 1]; ev[b + 1] = t; }
        out[p] = q > 0.0 ? 1.0 : 0.0;
        out[np + p] = rs > 0.0 ? 1.0 : 0.0;
        out[2 * np + p] = cls;
        out[3 * np + p] = sqrt(-(q - qs) / qs);
        out[4 * np + p] = ev[1] < 0.0 ? 1.0 : 0.0;
      }
  for (int c = 0; c < 3; c++) free(ug[c]);
}

/* ======================================================================
 * Single-point probes of the particle operations on the flow grid, for the
 * known-answer tests of tracking, particle BCs and interpolation.
 * ====================================================================== */

/* track_lp2node (lp_tracking_m.f90:72-109) of one point from node n3: one
 * pass, or (recursive) passes until the node stays, as reconstruct_flowmap's
 * recursive tracking does (without the periodic exchange).  Returns the
 * number of passes, or -1 after more than 2*max(n) passes. */
int oc_track_point(const ocfd *s, const double q[3], int n3[3], int recursive) {
  const ogrid *g = &s->F.g;
  const int max_track = 2 * imax3(g->n[0], g->n[1], g->n[2]);
  for (int pass = 1;; pass++) {
    int moved = track_pass(g, q, n3);
    if (!recursive || !moved) return pass;
    if (pass > max_track) return -1;
  }
}

/* set_lp_bc (lp_tracking_m.f90:185-277) of one point q tracked to node n3
 * (clamped in place); returns 1 for LP_STICK, -1 for an unknown flag */
int oc_lp_bc_point(const ocfd *s, double q[3], int n3[3]) {
  odiag one;
  memset(&one, 0, sizeof(one));
  double x0 = q[0], x1 = q[1], x2 = q[2];
  int a = n3[0], b = n3[1], c = n3[2], fl = 0;
  double *xp[3] = {&x0, &x1, &x2};
  int *nd[3] = {&a, &b, &c};
  for (int d = 0; d < 3; d++) one.xp[d] = xp[d];
  one.pflag = &fl;
  one.np = 1;
  int e = set_lp_bc(&s->F.g, &one, nd);
  q[0] = x0; q[1] = x1; q[2] = x2;
  n3[0] = a; n3[1] = b; n3[2] = c;
  if (e) return -1;
  return fl == LP_STICK ? 1 : 0;
}

/* interp_s2u_r1 (lp_tracking_m.f90:279-662) with interpolator interp of an
 * interior field (u, v, w: np each, Fortran order) at npt points q (3*npt,
 * point-major) from nodes n3 (3*npt), after the ghost exchange and
 * set_velocity_bc that cfd2lcs_update applies to a velocity field. */
int oc_interp_points(ocfd *s, int interp, const double *u, const double *v, const double *w, long npt,
                     const double *q, const int *n3, double *out) {
  const ogrid *g = &s->F.g;
  double *f[3], *pos[3], *res[3];
  int *nd[3];
  const double *src[3] = {u, v, w};
  for (int d = 0; d < 3; d++) {
    f[d] = calloc(g->nt, sizeof(double));
    pos[d] = malloc(npt * sizeof(double));
    res[d] = malloc(npt * sizeof(double));
    nd[d] = malloc(npt * sizeof(int));
    for (long p = 0; p < npt; p++) { pos[d][p] = q[3 * p + d]; nd[d][p] = n3[3 * p + d]; }
  }
  for (int k = 1; k <= g->n[2]; k++)
    for (int j = 1; j <= g->n[1]; j++)
      for (int i = 1; i <= g->n[0]; i++)
        for (int d = 0; d < 3; d++) f[d][GID(g, i, j, k)] = src[d][PIDG(g, i, j, k)];
  for (int d = 0; d < 3; d++) exchange(g, f[d], -1, NULL, 0);
  int e = set_velocity_bc(g, f);
  const int save = s->interpolator;
  s->interpolator = interp;
  if (!e) e = interp_field(s, &s->F, npt, f, pos, nd, res);
  s->interpolator = save;
  for (long p = 0; p < npt; p++)
    for (int d = 0; d < 3; d++) out[3 * p + d] = res[d][p];
  for (int d = 0; d < 3; d++) { free(f[d]); free(pos[d]); free(res[d]); free(nd[d]); }
  return e;
}
