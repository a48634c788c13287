// lcs_device.h -- device-side data descriptors and launcher interface of the
// MI355X FTLE path.  The grid is one GPU's whole domain with a one-cell ghost
// shell (the reference's ng = 1, data_m.f90:106); fields are SoA, i fastest:
//   ghosted index  l = i + N0*(j + N1*k),  i in [0, n0+1]   (Fortran i)
//   interior index p = (i-1) + n0*((j-1) + n1*(k-1))        (lcs_ijk2l - 1)
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace lcs {

enum BcFlag { kInternal = 0, kInflow = 1, kOutflow = 2, kWall = 3, kSlip = 4, k2D = 5, kMask = 6 };
enum Integrator { kEuler = 0, kTrapezoidal = 1, kRK2 = 2, kRK3 = 3, kRK4 = 4 };
enum Interpolator { kNearest = 0, kLinear = 1, kQuadratic = 2, kCubic = 3, kTSE = 5, kTSELimit = 6 };

// device error bits (folded into the sticky error flag on the host)
enum DevErr { kErrBcFlag = 1, kErrTrack = 2, kErrSpline = 4, kErrRange = 8 };

// Grid description passed by value to every kernel.
template <typename T>
struct GridDev {
  int n[3];          // interior extents
  int N[3];          // ghosted extents (n + 2)
  long long nt, np;  // ghosted / interior sizes
  const T* gl[3];    // ghosted coordinate lines (grid%x(:,1,1) ...), N[d] each
  const T* xl[3];    // interpolation lines, 2-D override (lp_tracking_m.f90:362-378)
  const int8_t* flag;  // ghosted bcflag
  const T* bn[3];    // bcnorm at interior nodes (np)
  const T* dl[3];    // 1/delta at interior nodes (np)
  int per[3];        // periodic axes
  int flat;          // the axis with n == 1, or -1
  int max_track;     // 2*max(gn) (lp_tracking_m.f90:127)
  T ps[27][3];       // periodic_shift(i,j,k,:) at (i+1)+3(j+1)+9(k+1)
  // ps[.][d] depends on the direction's axis-d index only (every rectilinear
  // grid): the first periodic image of a point is then found axis by axis
  int ps_sep;
  T bb[6];           // ghost-inclusive bounding box
  int kord[3];       // spline orders
  const T* kn[3];    // spline knots, N[d] + kord[d] each
  const T* band[3];  // banded LU of the collocation matrices, (2k-1)*N[d]
  const T* bpoly[3]; // per knot interval: basis polynomials in the normalised
                     // interval coordinate and 1/length, (N[d]+1)*(k*k+1)
  // non-rectilinear grids (rect == 0): full ghosted coordinates (nt each),
  // face least-squares gradient weights (18 x np, [(f*3+d)*np + p]) and the
  // 1-based array positions of the coordinate extrema (comms_m.f90:1625-1632)
  int rect;
  const T* gx[3];
  const T* lsw;
  T pmin[3][3], pmax[3][3];
};

template <typename T>
struct F3 {
  T* c[3];
};
template <typename T>
struct CF3 {
  const T* c[3];
};

struct DevStatus {
  unsigned long long cfl_bits;  // max CFL over interior nodes (as ordered bits)
  unsigned long long umax_bits; // max |velocity component| over interior nodes (ordered bits)
  unsigned long long oob;       // spline evaluations outside the knot range
  int err;                      // DevErr bits
  int pad;
};

// Storage of the gathered stage velocities.  kStoreNative keeps the compute
// type, one array per component.  The 16-bit formats (FP32 compute only)
// pack the three components of a node into one 8-byte record (u, v, w, 0),
// so a node's velocity is a single aligned load and a 4-node spline row two
// 16-byte loads (against three for the FP32 component arrays); values are
// widened to FP32 in registers.  kStoreF32 (FP64 compute only) keeps one
// FP32 array per component: the gathers move FP32 rows and the sums, the
// positions and the flowmap stay FP64.
enum Storage { kStoreNative = 0, kStoreF16 = 1, kStoreBF16 = 2, kStoreF32 = 3 };
struct half_t { uint16_t u; };  // IEEE binary16
struct bf16_t { uint16_t u; };  // bfloat16 (upper half of a binary32)
struct __attribute__((aligned(8))) half4_t { half_t e[4]; };
struct __attribute__((aligned(8))) bf164_t { bf16_t e[4]; };
// bytes of one stage velocity field (all three components) over nt nodes
inline size_t stage_field_bytes(int storage, size_t native, size_t nt) {
  return storage == kStoreNative ? 3 * native * nt : (storage == kStoreF32 ? 3 * sizeof(float) * nt : 8 * nt);
}

// Per-subcycle integration parameters (lp_motion_m.f90:129-360).
template <typename T>
struct StageParams {
  int integ;   // Integrator
  T dt;        // signed sub-cycle step
  // velocity of each stage, u = (1-ct) u_n + ct u_np1 at the stage's ct
  // (lp_motion_m.f90:170-316): grid values for NEAREST/LINEAR, spline
  // coefficients for QUADRATIC/CUBIC; element type given by `storage`
  const void* v[4][3];
  int storage;
  // 16-bit storage holds u / vscale (a power of two chosen per update so
  // that the stored values and coefficients stay in the format's normal
  // range); gathered stage velocities are multiplied back by it
  T vscale;
  // grid values of u_n and u_np1 and the first stage's ct, for the
  // semi-Lagrangian step whose first stage starts on a node
  CF3<T> un, unp1;
  T ct0;
  // B-spline steps stage the basis tables of their intervals in LDS
  int lds;
};

// The forward particle set in one storage order: position, displacement since
// the last reset, tracked node (lp%xp, lp%dx, lp%no, data_m.f90:317-337) and
// the interior index of the node the particle started from (lp%no0).
// With compensated updates, xl and dl carry the rounding errors of x and d
// (x + xl and d + dl are the unevaluated sums).
template <typename T>
struct ParticleSet {
  T* x[3];
  T* d[3];
  T* xl[3];
  T* dl[3];
  int16_t* n[3];
  int* org;
};

// ---- launchers (kernels.hip, particles.hip); all asynchronous on `stream` ----
// zero `bytes` bytes at p (a fill kernel: usable where a memset node is not)
hipError_t launch_zero(void* p, size_t bytes, hipStream_t stream);
template <typename U, typename T>
void launch_ingest(const GridDev<T>& g, F3<T> dst, const U* u, const U* v, const U* w,
                   hipStream_t stream);
template <typename T>
void launch_vel_ghost(const GridDev<T>& g, F3<T> u, CF3<T> uold, DevStatus* st, hipStream_t stream);
template <typename T>
void launch_cfl(const GridDev<T>& g, F3<T> u, DevStatus* st, hipStream_t stream);
// flowmap fields hold E = T or, under FP64 compute, FP32 (FLOWMAP_STORAGE)
template <typename T, typename E>
void launch_fm_ghost(const GridDev<T>& g, F3<E> fm, CF3<E> fmold, DevStatus* st, hipStream_t stream);
// coef = spline coefficients of the data in src[0..2] (element type ES), or
// of coef itself (in place) when src is null; the solve runs in T
template <typename T, typename E, typename ES = E>
void launch_spline_fit(const GridDev<T>& g, F3<E> coef, int ncomp, hipStream_t stream,
                       const ES* const* src = nullptr);
// dst = src widened to T (n values)
template <typename T>
void launch_widen(const float* src, T* dst, long long n, hipStream_t stream);
// out (element type per `storage`) = ((1-ct) a + ct b) * inv, inv a power
// of two (1 for native storage); a 16-bit value out of range sets kErrRange
template <typename T, typename A>
void launch_blend(const GridDev<T>& g, void* const out[3], int storage, CF3<A> a, CF3<A> b, T ct, T inv,
                  DevStatus* st, hipStream_t stream);
// up to kBlendBatch blends of the same two fields in one pass, each as
// launch_blend forms it (packed 16-bit records: out[q][0])
constexpr int kBlendBatch = 8;
template <typename T>
struct BlendBatch {
  void* out[kBlendBatch][3];
  T ct[kBlendBatch];
  T inv;
  int n;
};
template <typename T, typename A>
void launch_blend_batch(const GridDev<T>& g, const BlendBatch<T>& bb, int storage, CF3<A> a, CF3<A> b,
                        DevStatus* st, hipStream_t stream);
// tiled: FP32 B-spline steps on 3-D grids stage the gathered coefficients
// in LDS (bit-identical results)
template <typename T, typename FM>
void launch_sl_step(const GridDev<T>& g, int interp, const StageParams<T>& sp, CF3<FM> fmold,
                    F3<FM> fmnew, DevStatus* st, hipStream_t stream, int tiled = 0);
// compensated: two-sum updates of position and displacement (uses xl, dl);
// np particles
template <typename T>
void launch_fwd_step(const GridDev<T>& g, int interp, const StageParams<T>& sp, ParticleSet<T> ps, int np,
                     int compensated, DevStatus* st, hipStream_t stream);
template <typename T>
void launch_fwd_reset(const GridDev<T>& g, ParticleSet<T> ps, hipStream_t stream);

// ---- diagnostic grids other than the flow grid g (resolution != 0, AUX_GRID):
// particles on the nodes of lg, velocity / particle BCs / tracking on g
// tracked flow node of every lg node (3 x lg.np int16, component-major),
// from reset_lp's guess by recursive tracking without periodic wrap
template <typename T>
void launch_track_nodes(const GridDev<T>& g, const GridDev<T>& lg, int16_t* nodes, DevStatus* st,
                        hipStream_t stream);
// forward particles reset onto the lg nodes, tracked onto g with the
// periodic wrap after every pass
template <typename T>
void launch_fwd_place(const GridDev<T>& g, const GridDev<T>& lg, ParticleSet<T> ps, DevStatus* st,
                      hipStream_t stream);
// one semi-Lagrangian sub-cycle of the flowmap on lg; fresh bit 0: the stage
// velocities and the particle BC use the tracked nodes as they are (else
// clamped to the interior), bit 1: the MASK test does
template <typename T, typename FM>
void launch_sl_step2(const GridDev<T>& g, const GridDev<T>& lg, const int16_t* nodes, int fresh, int interp,
                     const StageParams<T>& sp, CF3<FM> fmold, F3<FM> fmnew, DevStatus* st, hipStream_t stream);
// ghosted field on g from an interior one: interior copied, periodic ghosts
// from their images, other ghosts zero (exchange_sdata of a zeroed field)
template <typename T>
void launch_expand_periodic(const GridDev<T>& g, const T* interior, F3<T> dst, hipStream_t stream);
// field (values or spline coefficients on g) interpolated at the lg nodes
// through their clamped tracked nodes, into out (3 x lg.np)
template <typename T>
void launch_field_at_nodes(const GridDev<T>& g, const GridDev<T>& lg, const int16_t* nodes, int interp,
                           CF3<T> field, T* out, hipStream_t stream);
// FTLE from the six AUX_GRID sets: aux[a] holds set a's composed
// displacement by origin node (3 x np), line[a] the coordinate line of set
// a's grid along axis a % 3 (ghosted, as GridDev::gl)
template <typename T>
struct AuxSets {
  const T* disp[6][3];
  const T* line[6];
};
template <typename T>
void launch_ftle_aux(const GridDev<T>& lg, const AuxSets<T>& aux, double small, double T_int, int incompressible,
                     double tiny, double* ftle, hipStream_t stream);
// device temporary bytes of the particle sort for np particles with keys
// below nkeys (0 on error)
size_t particle_sort_temp_bytes(long long np, long long nkeys);
// sorts set a by tracked node into set b; keys and vals hold 2*np entries
// (np particles; keys from their tracked nodes on g)
template <typename T>
int launch_particle_sort(const GridDev<T>& g, ParticleSet<T> a, ParticleSet<T> b, long long np, unsigned* keys,
                         unsigned* vals, void* temp, size_t temp_bytes, hipStream_t stream);
// dst[org[q]] = src[q] (+ lo[q] when lo is given) for q < np: per-particle
// values back to origin order
template <typename T>
void launch_unpermute(long long np, const int* org, CF3<T> src, const T* const lo[3], F3<T> dst,
                      hipStream_t stream);
template <typename T>
void launch_field_from_origin(const GridDev<T>& g, CF3<T> src, F3<T> dst, hipStream_t stream);
template <typename T>
void launch_compose(const GridDev<T>& g, int interp, const T* slots, long long slot_stride,
                    const int* order, int nslot, F3<T> dxout, DevStatus* st, hipStream_t stream);
// composition state of one particle between slots: split position, tracked
// node, accumulated displacement, origin (48 bytes in FP32, 96 in FP64)
template <typename T>
struct __attribute__((aligned(16))) CompRec {
  T qb[3], off[3], dx[3];
  int16_t nd[3];
  int16_t pad;
  int org;
};
// one slot of the re-ordered composition: particles from in (through perm
// when given, else in place; first: seeded on the nodes) through the slot
// field f, written to out with their brick keys and positions (keys may be
// null), or, on the last slot, their displacement scattered by origin
template <typename T>
void launch_compose_slot(const GridDev<T>& g, int interp, CF3<T> f, const CompRec<T>* in, const unsigned* perm,
                         CompRec<T>* out, unsigned* keys, unsigned* vals, int first, int last, F3<T> dxout,
                         DevStatus* st, hipStream_t stream);
// radix sort of np (brick key, position) pairs held in double buffers of 2 np
// values each; returns the positions in key order (null on failure)
size_t compose_sort_temp_bytes(long long np, long long nkeys);
const unsigned* launch_compose_sort(unsigned* keys, unsigned* vals, long long np, long long nkeys, void* temp,
                                    size_t temp_bytes, hipStream_t stream);
// velocity invariants Q, C, D, H, L2 (5 * np) of the interior velocity u
// (3 * np, or null for zero) at T0
template <typename T>
void launch_invariants(const GridDev<T>& g, const T* u, T* out, hipStream_t stream);
template <typename T>
void launch_ftle(const GridDev<T>& g, CF3<T> fm, double T_int, int incompressible,
                 double tiny, double* ftle, F3<T> comp, hipStream_t stream);
template <typename T, typename U>
void launch_interior_to_user(const GridDev<T>& g, CF3<T> src, int ghosted, U* a, U* b, U* c,
                             hipStream_t stream);

// ---- adjoint launchers (adjoint.hip) ----
// tile-store scratch of the reverse sub-cycle scatter: tiles (of
// sl_adjoint_tile_bytes bytes) and a flag that the scatter sets to seq when a
// point's support leaves its window (seq differs between calls, and the flag
// starts at a value no seq takes)
template <typename T>
struct TileStore {
  T* tiles;
  int* flag;
  int seq;
};
// one semi-Lagrangian sub-cycle in reverse; stage s >= 2 scatters into
// adj_v[sbuf[s]]; the first stage's nodal adjoint goes to dun / dunp1; ovf[4]
// (zero on entry and exit) take the rare scatters outside a node's window
template <typename T>
bool launch_sl_adjoint(const GridDev<T>& g, int interp, const StageParams<T>& sp, CF3<T> fmold, CF3<T> lam,
                       F3<T> adj_fm, F3<T> const adj_v[3], F3<T> const ovf[4], const int sbuf[4], int nbuf,
                       F3<T> dun, F3<T> dunp1, const TileStore<T>* ts, const T bct[3], F3<T> dcn,
                       F3<T> dcnp1, hipStream_t stream);
// bytes of the tile-store scratch launch_sl_adjoint uses when given one.
// ts != nullptr: one launch over all bricks plus a gather pass that assigns
// adj_fm and adds the stage-buffer adjoints straight into dcn / dcnp1
// (weights 1 - bct[b], bct[b]); adj_v is not touched and it returns true.
// ts == nullptr: eight colour launches that add their tiles into adj_fm and
// adj_v in place (the caller unblends adj_v); it returns false.
template <typename T>
size_t sl_adjoint_tile_bytes(const GridDev<T>& g, int interp, int nbuf);
// an += (1-ct) b, anp1 += ct b, b = 0
template <typename T>
void launch_adj_unblend(const GridDev<T>& g, F3<T> b, T ct, F3<T> an, F3<T> anp1, hipStream_t stream);
// transpose of launch_spline_fit (three components, in place)
template <typename T>
void launch_spline_fit_T(const GridDev<T>& g, F3<T> a, hipStream_t stream);
// ghost adjoints back to their interior sources; mode 0 velocity rules, 1
// flowmap rules; interior nodes forced to zero drop their adjoint first
template <typename T>
void launch_fold_ghost(const GridDev<T>& g, F3<T> adj, int mode, hipStream_t stream);
template <typename T>
void launch_ftle_adjoint(const GridDev<T>& g, CF3<T> fm, double Tint, int incomp, double tiny, const double* dftle,
                         F3<T> adj, hipStream_t stream);
template <typename T, typename U>
void launch_adj_to_user(const GridDev<T>& g, CF3<T> adj, U* a, U* b, U* c, int accumulate, hipStream_t stream);
template <typename T>
void launch_interior(const GridDev<T>& g, CF3<T> src, F3<T> dst, hipStream_t stream);

}  // namespace lcs
