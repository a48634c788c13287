/*
 * cfd2lcs_api.h -- flags and entry points of the MI355X-native libcfd2lcs
 * FTLE path.  Include through cfd2lcs_inc_dp.h or cfd2lcs_inc_sp.h, which
 * define lcsdata_t (double / float) exactly as the reference headers do.
 *
 * Part 1 is the reference interface, unchanged: a CFD solver built against
 * the reference's cfd2lcs_inc.h links against this library without source
 * changes.  Part 2 lists the additive entry points of this implementation;
 * none of them alters the behaviour of part 1.
 */
#ifndef CFD2LCS_API_H
#define CFD2LCS_API_H

#if defined(__has_include)
#if __has_include(<mpi.h>) && !defined(CFD2LCS_NO_MPI)
#include <mpi.h>
#define CFD2LCS_HAVE_MPI_H 1
#endif
#endif
#ifndef CFD2LCS_HAVE_MPI_H
/* No MPI on this host: the communicator is accepted and ignored (one GPU is
 * one rank).  int matches MPICH's MPI_Comm; any handle type passed in a
 * register is accepted by the library, which never dereferences it. */
typedef int MPI_Comm;
#endif

/* The string namelength (cfd2lcs_inc_double.h:22) */
#define LCS_NAMELEN 128

/* Boundary condition flags (cfd2lcs_inc_double.h:27-33) */
#define LCS_INTERNAL 0
#define LCS_INFLOW 1
#define LCS_OUTFLOW 2
#define LCS_WALL 3
#define LCS_SLIP 4
#define LCS_2D 5
#define LCS_MASK 6

/* LCS diagnostics (cfd2lcs_inc_double.h:38-40) */
#define FTLE_FWD 0
#define FTLE_BKWD 1
#define LP_TRACER 2

/* Integration methods (cfd2lcs_inc_double.h:45-49) */
#define EULER 0
#define TRAPEZOIDAL 1
#define RK2 2
#define RK3 3
#define RK4 4

/* Interpolation methods (cfd2lcs_inc_double.h:54-59) */
#define NEAREST_NBR 0
#define LINEAR 1
#define QUADRATIC 2
#define CUBIC 3
#define TSE 5
#define TSE_LIMIT 6

/* True/False indicators (cfd2lcs_inc_double.h:64-65) */
#define LCS_FALSE 0
#define LCS_TRUE 1

/* Values of the additive "PRECISION" / "VELOCITY_STORAGE" options */
#define LCS_FP64 8
#define LCS_FP32 4
#define LCS_FP16 2
#define LCS_BF16 3

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------
 * Part 1: the reference C API (cfd2lcs_inc_double.h:72-110,
 * implemented in the reference by cfd2lcs_c.c:67-142 -> cfd2lcs.f90)
 * --------------------------------------------------------------------- */

/* replaces cfd2lcs_init_c (cfd2lcs_c.c:67-83 -> cfd2lcs.f90:2-53) */
void cfd2lcs_init_c(MPI_Comm usercomm, int n[3], int offset[3], lcsdata_t *x,
                    lcsdata_t *y, lcsdata_t *z, int *flag);

/* replaces cfd2lcs_diagnostic_init_c (cfd2lcs_c.c:86-99 -> cfd2lcs.f90:252-358) */
int cfd2lcs_diagnostic_init_c(int lcs_type, int resolution, lcsdata_t t,
                              lcsdata_t h, char label[]);

/* replaces cfd2lcs_update_c (cfd2lcs_c.c:102-115 -> cfd2lcs.f90:55-250) */
void cfd2lcs_update_c(int n[3], lcsdata_t *u, lcsdata_t *v, lcsdata_t *w,
                      lcsdata_t time);

/* replaces cfd2lcs_diagnostic_destroy_c (cfd2lcs_c.c:118-122 -> cfd2lcs.f90:360-386) */
void cfd2lcs_diagnostic_destroy_c(int lcs_handle);

/* replaces cfd2lcs_set_option_c (cfd2lcs_c.c:125-132 -> cfd2lcs.f90:436-622) */
void cfd2lcs_set_option_c(char option[], int val);

/* replaces cfd2lcs_set_param_c (cfd2lcs_c.c:135-142 -> cfd2lcs.f90:624-682) */
void cfd2lcs_set_param_c(char option[], lcsdata_t val);

/* ---------------------------------------------------------------------
 * Part 2: additive entry points (no counterpart in the reference C shim)
 * --------------------------------------------------------------------- */

/* C wrapper of the Fortran-only cfd2lcs_finalize (cfd2lcs.f90:388-434).
 * Releases all device memory and resets the library, including the sticky
 * error flag and the first-update state, so that init may be called again. */
void cfd2lcs_finalize_c(void);

/* Sticky error flag (CFD2LCS_ERROR, data_m.f90:381): 0 when healthy. */
int cfd2lcs_error_c(void);

/* Output files.  As the reference does (write_lcs, io_nohdf5_m.f90:47-332),
 * every FTLE output writes ./cfd2lcs_output/<label>_<index>.dat, a Tecplot
 * ASCII file with X, Y, Z, FM-X..Z (option WRITE_FLOWMAP), U-X..Z
 * (WRITE_VELOCITY), FLAG (WRITE_BCFLAG) and FTLE in BLOCK packing, ES14.6,
 * ten values per line.  Files are written by a background thread from host
 * copies of the results (each under <name>.part until complete);
 * cfd2lcs_update_c returns without waiting for them.  This call waits until
 * every file requested so far is on disk; it returns 0, or -1 when a write
 * failed (the error flag is then set).  cfd2lcs_finalize_c waits as well. */
int cfd2lcs_flush_output_c(void);

/* The number formatting of those files: n values in the Fortran edit
 * descriptor ESw.d (ES14.6 for data, ES18.4 for SOLUTIONTIME), each exactly
 * w characters, concatenated into out (n*w bytes, no terminator).  Returns 0,
 * or -1 for unsupported w, d.  Host-only; for tools that compare or produce
 * such files. */
int cfd2lcs_format_es_c(const double *v, long long n, int w, int d, char *out);

/* Same as cfd2lcs_update_c, but u, v, w are device (HBM) pointers of n[0]*
 * n[1]*n[2] lcsdata_t each, Fortran order; no host staging is involved. */
void cfd2lcs_update_device_c(int n[3], const lcsdata_t *d_u,
                             const lcsdata_t *d_v, const lcsdata_t *d_w,
                             lcsdata_t time);

/* Results kept in HBM by the library; copied to host arrays of one value
 * per node of the diagnostic's grid (the flow grid's n[0]*n[1]*n[2] unless
 * resolution != 0, see cfd2lcs_get_lcs_grid_c), Fortran order.  Return 0 on
 * success.
 *   flowmap : displacement over the current h interval (the BKWD
 *             semi-Lagrangian field lp%fm, or the FWD particles' dx by
 *             origin node)
 *   composed: T-flowmap of the latest composition (what write_lcs writes
 *             as the flowmap, io_nohdf5_m.f90:183-265)
 *   ftle    : FTLE field of the latest composition (lcs_m.f90:66-193) */
int cfd2lcs_get_flowmap_c(int lcs_handle, lcsdata_t *fx, lcsdata_t *fy,
                          lcsdata_t *fz);
int cfd2lcs_get_composed_c(int lcs_handle, lcsdata_t *fx, lcsdata_t *fy,
                           lcsdata_t *fz);
int cfd2lcs_get_ftle_c(int lcs_handle, lcsdata_t *ftle);
/* The grid of a diagnostic (lcs%sgrid): the flow grid, or with
 * resolution != 0 the refined / coarsened grid of new_sgrid_from_sgrid
 * (sgrid_m.f90:445-761), whose extents size the arrays of the three getters
 * above.  Fills n[3] and, when x, y, z are not NULL, the node coordinates
 * (n[0]*n[1]*n[2] each, Fortran order).  Returns 0 on success.  No
 * counterpart in the reference, which keeps the grid internal. */
int cfd2lcs_get_lcs_grid_c(int lcs_handle, int n[3], lcsdata_t *x, lcsdata_t *y, lcsdata_t *z);
/* device pointer (double, n[0]*n[1]*n[2], Fortran order) to the FTLE field;
 * valid until the next update or finalize */
const double *cfd2lcs_get_ftle_device_c(int lcs_handle);
/* number of FTLE fields produced so far by a diagnostic, or -1 */
int cfd2lcs_get_noutput_c(int lcs_handle);

/* Counters (cfd2lcs_info, cfd2lcs.f90:708-810): fills out[8] with
 *   [0] n_subcycle of the latest set_dt
 *   [1] cumulative forward particle integrations (integrations_fwd_c)
 *   [2] cumulative backward integrations (integrations_bkwd_c)
 *   [3] seconds of GPU time in forward advection   (cpu_fwd)
 *   [4] seconds of GPU time in backward advection  (cpu_bkwd)
 *   [5] seconds of GPU time in composition          (cpu_reconstruct)
 *   [6] seconds of GPU time in gradient/CG/FTLE     (cpu_ftle)
 *   [7] spline evaluations outside the knot range (db3val returns 0) */
void cfd2lcs_get_stats_c(double out[8]);

/* GPU time (seconds) and launch counts accumulated per category, measured
 * with HIP events on the library's stream; returns the number of categories
 * filled (at most n, 16 in this build):
 *   0 forward advection   1 backward advection   2 composition   3 FTLE
 * and, when the option "KERNEL_TIMING" is LCS_TRUE, per kernel:
 *   8 forward particle step   9 semi-Lagrangian step   10 spline fit
 *   11 composition kernel    12 FTLE kernel           13 ghost/ingest/CFL
 *   14 time blend of the stage velocities  15 forward particle sort
 *   16 composition sort (between the per-slot composition launches) */
int cfd2lcs_get_timers_c(double *seconds, long long *counts, int n);

/* Host-side grid set-up, for inspection: lines[N0+N1+N2] (ghosted
 * coordinate lines), flag[(n0+2)(n1+2)(n2+2)], bn[3*n0*n1*n2] (boundary
 * normals), ps[81] (periodic shifts), bb[6] (bounding box).  Any pointer may
 * be NULL.  Returns 0 on success. */
int cfd2lcs_debug_grid_c(double *lines, int *flag, double *bn, double *ps,
                         double *bb);

/* Adjoint of the backward (semi-Lagrangian) FTLE over one window.  No
 * counterpart in the reference; the forward/backward pair plays the role of a
 * torch.autograd.Function for a scalar loss on the FTLE field.
 *
 * forward: the BKWD diagnostic's flowmap with h = T over nsnap velocity
 *   snapshots (device arrays of nsnap * n0*n1*n2 lcsdata_t each, snapshot s
 *   at offset s*n0*n1*n2, Fortran order) at increasing times[nsnap] (host),
 *   T = times[nsnap-1] - times[0], with the grid, INTEGRATOR, INTERPOLATOR,
 *   CFL and INCOMPRESSIBLE of the library; writes the FTLE (device double,
 *   n0*n1*n2) to d_ftle and keeps a flowmap checkpoint every
 *   checkpoint_every steps.  The snapshot arrays must stay valid until the
 *   backward call.  Returns 0 on success.
 * backward: given dL/dFTLE (device double, n0*n1*n2), writes dL/du, dL/dv,
 *   dL/dw for every snapshot (device, same layout as the snapshots),
 *   recomputing each segment from its checkpoint.  Returns 0 on success.
 * release: frees the tape.  stats: out[0] forward seconds, out[1] backward
 *   seconds, out[2] tape bytes allocated, out[3] total sub-cycles. */
int cfd2lcs_adjoint_forward_c(int n[3], int nsnap, const lcsdata_t *d_u,
                              const lcsdata_t *d_v, const lcsdata_t *d_w,
                              const lcsdata_t *times, int checkpoint_every,
                              double *d_ftle);
int cfd2lcs_adjoint_backward_c(const double *d_dftle, lcsdata_t *d_gu,
                               lcsdata_t *d_gv, lcsdata_t *d_gw);
void cfd2lcs_adjoint_release_c(void);
void cfd2lcs_adjoint_stats_c(double out[4]);

/* Allocation-free variant of cfd2lcs_adjoint_forward_c.  It has no
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under stream capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same inputs and FTLE output as
 * cfd2lcs_adjoint_forward_c, but it keeps no tape (no backward call can
 * follow it) and it calls no hipMalloc/hipFree: every buffer it needs is
 * carved from a fixed 64 MiB device workspace that the library holds on each
 * GPU from load until the process ends.  Grids whose buffers do not fit
 * (about 74^3 in FP32 with cubic B-splines, half that in FP64) fail with a
 * message.  Calls that overlap on one GPU share the workspace and race.
 * cfd2lcs_init_c must have run.  Returns 0 on success. */
int cfd2lcs_adjoint_forward_noalloc_c(int n[3], int nsnap, const lcsdata_t *d_u,
                                      const lcsdata_t *d_v, const lcsdata_t *d_w,
                                      const lcsdata_t *times, double *d_ftle);

/* ---------------------------------------------------------------------
 * Part 3: Fortran-callable entry points.  A Fortran host calls the
 * reference's external subroutines cfd2lcs_init, cfd2lcs_diagnostic_init,
 * ... (cfd2lcs.f90), which a Fortran compiler links as the lower-case,
 * underscore-suffixed names below with every argument by reference and the
 * hidden character lengths last (the same names the reference's C shim
 * binds, cfd2lcs_c.c:16-58).
 * --------------------------------------------------------------------- */
#include <stddef.h>
/* replaces cfd2lcs_init (cfd2lcs.f90:2-53) */
void cfd2lcs_init_(int *usercomm, int *n, int *offset, lcsdata_t *x,
                   lcsdata_t *y, lcsdata_t *z, int *flag);
/* replaces cfd2lcs_diagnostic_init (cfd2lcs.f90:252-358) */
void cfd2lcs_diagnostic_init_(int *lcs_handle, int *lcs_type, int *resolution,
                              lcsdata_t *t, lcsdata_t *h, char *label,
                              size_t namelen);
/* replaces cfd2lcs_update (cfd2lcs.f90:55-250) */
void cfd2lcs_update_(int *n, lcsdata_t *u, lcsdata_t *v, lcsdata_t *w,
                     lcsdata_t *time);
/* replaces cfd2lcs_diagnostic_destroy (cfd2lcs.f90:360-386) */
void cfd2lcs_diagnostic_destroy_(int *lcs_handle);
/* replaces cfd2lcs_set_option (cfd2lcs.f90:436-622) */
void cfd2lcs_set_option_(char *option, int *val, size_t namelen);
/* replaces cfd2lcs_set_param (cfd2lcs.f90:624-682) */
void cfd2lcs_set_param_(char *param, lcsdata_t *val, size_t namelen);
/* replaces cfd2lcs_finalize (cfd2lcs.f90:388-434) */
void cfd2lcs_finalize_(void);

#ifdef __cplusplus
}
#endif
#endif
