/*
 * oracle.h -- CPU FP64 restatement of libcfd2lcs's structured-grid FTLE path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by the
 * parity tests under tests/, by __graft_entry__.smoke() and by the CPU-baseline
 * leg of bench.py.  It is never linked into, called by, or shipped with the
 * product library (libcfd2lcs_amd/).
 *
 * It restates, on one rank and in double precision, what the reference does:
 *   grid set-up ............ src/sgrid_m.f90:16-444, 810-866, 951-1014
 *   ghost exchange ......... src/comms_m.f90:679-869 (single-rank SELF_COMM)
 *   particle motion ........ src/lp_motion_m.f90:15-360
 *   tracking / BC / interp . src/lp_tracking_m.f90:16-662
 *   B-splines .............. src/bspline/bspline_sub_module.f90:451-656,1737-2634
 *   SL flowmap, composition  src/flowmap_m.f90:17-392
 *   FTLE ................... src/lcs_m.f90:66-193, src/gradient_m.f90:11-217
 *   driver / options ....... src/cfd2lcs.f90:2-682
 * Each function below cites the reference lines it follows.
 */
#ifndef CFD2LCS_ORACLE_H
#define CFD2LCS_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct ocfd ocfd;

/* cfd2lcs_init (cfd2lcs.f90:2-53).  Arrays are ni*nj*nk, Fortran order. */
ocfd *oc_init(const int n[3], const double *x, const double *y,
              const double *z, const int *flag);
void oc_free(ocfd *s);
/* cfd2lcs_diagnostic_init (cfd2lcs.f90:252-358); returns the 1-based handle */
int oc_diagnostic_init(ocfd *s, int lcs_type, int resolution, double T,
                       double h, const char *label);
/* cfd2lcs_set_option / cfd2lcs_set_param (cfd2lcs.f90:436-682) */
void oc_set_option(ocfd *s, const char *option, int val);
void oc_set_param(ocfd *s, const char *param, double val);
/* cfd2lcs_update (cfd2lcs.f90:55-250) */
void oc_update(ocfd *s, const int n[3], const double *u, const double *v,
               const double *w, double time);
int oc_error(const ocfd *s);

/* ---- inspection (interior, Fortran order, ni*nj*nk each) ---- */
/* current h-interval flowmap: BKWD = lp%fm, FWD = lp%dx by origin node */
int oc_get_flowmap(const ocfd *s, int handle, double *fx, double *fy,
                   double *fz);
/* T-flowmap of the latest composition and the FTLE computed from it */
int oc_get_composed(const ocfd *s, int handle, double *fx, double *fy,
                    double *fz);
int oc_get_ftle(const ocfd *s, int handle, double *ftle);
/* the grid of a diagnostic's particle set (set 0: its own; 1..6: the
 * AUX_GRID sets X0, Y0, Z0, X1, Y1, Z1): extents and interior coordinates */
int oc_get_lcs_grid(const ocfd *s, int handle, int set, int n[3], double *x, double *y, double *z);
int oc_get_noutput(const ocfd *s, int handle);
/* CFL sub-cycle count of the latest set_dt call and cumulative counters */
int oc_get_nsub(const ocfd *s);
long long oc_get_integrations(const ocfd *s, int direction);
/* ghosted grid access: (ni+2)*(nj+2)*(nk+2) each */
int oc_get_ghost_grid(const ocfd *s, double *gx, double *gy, double *gz,
                      int *flag, double *bnx, double *bny, double *bnz);
int oc_get_periodic_shift(const ocfd *s, double *ps /* 27*3 */, double *bb);

/* ---- stand-alone kernels, exposed for the known-answer tests ---- */
/* dbknot: knots for order k on n abscissae; t has n+k entries */
void oc_bspline_knots(const double *x, int n, int k, double *t);
/* dbintk + dbtpcf in 3-D (db3ink): coefficients on nx*ny*nz data */
int oc_bspline_fit3(const double *x, int nx, const double *y, int ny,
                    const double *z, int nz, int kx, int ky, int kz,
                    const double *f, double *tx, double *ty, double *tz,
                    double *coef);
/* db3val with idx=idy=idz=0 */
double oc_bspline_eval3(double xv, double yv, double zv, const double *tx,
                        const double *ty, const double *tz, int nx, int ny,
                        int nz, int kx, int ky, int kz, const double *coef,
                        int *iflag);
/* largest eigenvalue of a symmetric 3x3 (row-major a[9]) */
double oc_eig_sym3_max(const double *a);
/* C = G^T G for a 3x3 gradient g (row-major, g[3*i+j] = d fm_i / d x_j) */
void oc_cauchy_green(const double *g, double *c);
/* compute_invariants of an interior velocity field (u, v, w: np each, or
 * NULL for zero); out: Q, C, D, H, L2, np each */
void oc_velocity_invariants(const ocfd *s, const double *u, const double *v, const double *w, double *out);

/* ---- single-point probes on the flow grid (known-answer tests) ---- */
/* track_lp2node from node n3 (updated): one pass, or passes until the node
 * stays (recursive); returns the passes taken, -1 past 2*max(n) passes */
int oc_track_point(const ocfd *s, const double q[3], int n3[3], int recursive);
/* set_lp_bc of point q (updated) tracked to node n3 (clamped in place);
 * 1 = LP_STICK, 0 = in bounds, -1 = unknown flag */
int oc_lp_bc_point(const ocfd *s, double q[3], int n3[3]);
/* interp_s2u_r1 with interpolator interp of an interior field after the
 * ghost exchange and set_velocity_bc, at npt points q[3*p+d] from nodes
 * n3[3*p+d]; out[3*p+c] */
int oc_interp_points(ocfd *s, int interp, const double *u, const double *v, const double *w, long npt,
                     const double *q, const int *n3, double *out);

#ifdef __cplusplus
}
#endif
#endif
