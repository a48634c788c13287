// kernels.hip -- HIP kernels of the MI355X FTLE path (gfx950).
//
// Stage (a)  semi-Lagrangian backward flowmap step  k_sl_step
//            forward particle step                  k_fwd_step
// Stage (b)  composition of the T/h sub-intervals   k_compose
// Stage (c)  gradient -> Cauchy-Green -> FTLE       k_ftle
// plus velocity ingest / ghost shell / CFL reduction and the tensor B-spline
// fit (three banded axis solves).  Reference semantics are cited per kernel.
#include "device_common.h"

namespace lcs {

// ---------------------------------------------------------------------------
// Velocity ingest, ghost shell, CFL (cfd2lcs.f90:112-117, sgrid_m.f90:810-866,
// lp_motion_m.f90:86-101)
// ---------------------------------------------------------------------------
template <typename U, typename T>
__global__ void __launch_bounds__(kBlock) k_ingest(GridDev<T> g, F3<T> dst, const U* __restrict__ u,
                                                   const U* __restrict__ v, const U* __restrict__ w) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.np) return;
  int i, j, k;
  decode_interior(g, p, i, j, k);
  const int l = gid(g, i, j, k);
  dst.c[0][l] = (T)u[p];
  dst.c[1][l] = (T)v[p];
  dst.c[2][l] = (T)w[p];
}

// ghost cells of u_np1: periodic copies (exchange_sdata) then
// set_velocity_bc; an INTERNAL ghost without a periodic source keeps the
// value it had (the previous u_np1, now held by uold).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_vel_ghost(GridDev<T> g, F3<T> u, CF3<T> uold, DevStatus* st) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int i, j, k;
  if (t >= ghost_count(g) || !decode_ghost(g, t, i, j, k)) return;
  const int l = gid(g, i, j, k);
  const int f = g.flag[l];
  const int ei = i < 1 ? -1 : (i > g.n[0] ? 1 : 0), ej = j < 1 ? -1 : (j > g.n[1] ? 1 : 0),
            ek = k < 1 ? -1 : (k > g.n[2] ? 1 : 0);
  switch (f) {
    case kInternal:
      if (self_comm(g, ei, ej, ek)) {
        const int s = gid(g, ei < 0 ? g.n[0] : (ei > 0 ? 1 : i), ej < 0 ? g.n[1] : (ej > 0 ? 1 : j),
                          ek < 0 ? g.n[2] : (ek > 0 ? 1 : k));
        for (int c = 0; c < 3; c++) u.c[c][l] = u.c[c][s];
      } else {
        for (int c = 0; c < 3; c++) u.c[c][l] = uold.c[c][l];
      }
      break;
    case kMask:
    case kWall:
      for (int c = 0; c < 3; c++) u.c[c][l] = T(0);
      break;
    case kSlip:
    case kInflow:
    case kOutflow:
    case k2D: {
      const int s = gid(g, iclamp(i, 1, g.n[0]), iclamp(j, 1, g.n[1]), iclamp(k, 1, g.n[2]));
      for (int c = 0; c < 3; c++) u.c[c][l] = u.c[c][s];
      break;
    }
    default:
      atomicOr(&st->err, (int)kErrBcFlag);
  }
}

// MASK nodes get zero velocity; cfl = sum_d |u_d / delta_d| over interior
// nodes (lp_motion_m.f90:90-98).  Grid-stride loop, block max folded into
// st->cfl_bits (cfl >= 0, so the IEEE bit pattern orders as an unsigned
// integer); the largest |component| goes to st->umax_bits the same way.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_cfl(GridDev<T> g, F3<T> u, DevStatus* st) {
  __shared__ double smax[2][kBlock / 64];
  double c = 0.0, um = 0.0;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < g.np;
       p += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    decode_interior(g, (int)p, i, j, k);
    const int l = gid(g, i, j, k);
    T a = u.c[0][l], b = u.c[1][l], d = u.c[2][l];
    if (g.flag[l] == kMask) {
      u.c[0][l] = T(0); u.c[1][l] = T(0); u.c[2][l] = T(0);
      a = b = d = T(0);
    }
    T s = T(0);
    s = s + fabs(a * g.dl[0][p]);
    s = s + fabs(b * g.dl[1][p]);
    s = s + fabs(d * g.dl[2][p]);
    c = fmax(c, (double)s);
    um = fmax(um, (double)fmax(fabs(a), fmax(fabs(b), fabs(d))));
  }
  for (int o = 32; o > 0; o >>= 1) {
    c = fmax(c, __shfl_down(c, o, 64));
    um = fmax(um, __shfl_down(um, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    smax[0][threadIdx.x >> 6] = c;
    smax[1][threadIdx.x >> 6] = um;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = smax[0][0], n = smax[1][0];
    for (int w = 1; w < kBlock / 64; w++) {
      m = fmax(m, smax[0][w]);
      n = fmax(n, smax[1][w]);
    }
    atomicMax(&st->cfl_bits, (unsigned long long)__double_as_longlong(m));
    atomicMax(&st->umax_bits, (unsigned long long)__double_as_longlong(n));
  }
}

// ghost cells of a flowmap: periodic copies then set_flowmap_bc's ghost
// pass (flowmap_m.f90:362-390).  Interior zeroing of MASK/INFLOW/OUTFLOW
// nodes is done by the producer.
template <typename T, typename E>
__global__ void __launch_bounds__(kBlock) k_fm_ghost(GridDev<T> g, F3<E> fm, CF3<E> fmold, DevStatus* st) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int i, j, k;
  if (t >= ghost_count(g) || !decode_ghost(g, t, i, j, k)) return;
  const int l = gid(g, i, j, k);
  const int f = g.flag[l];
  const int ei = i < 1 ? -1 : (i > g.n[0] ? 1 : 0), ej = j < 1 ? -1 : (j > g.n[1] ? 1 : 0),
            ek = k < 1 ? -1 : (k > g.n[2] ? 1 : 0);
  switch (f) {
    case kInternal:
      if (self_comm(g, ei, ej, ek)) {
        const int s = gid(g, ei < 0 ? g.n[0] : (ei > 0 ? 1 : i), ej < 0 ? g.n[1] : (ej > 0 ? 1 : j),
                          ek < 0 ? g.n[2] : (ek > 0 ? 1 : k));
        for (int c = 0; c < 3; c++) fm.c[c][l] = fm.c[c][s];
      } else {
        for (int c = 0; c < 3; c++) fm.c[c][l] = fmold.c[c][l];
      }
      break;
    case kWall:
    case kMask:
    case kInflow:
    case kOutflow:
      for (int c = 0; c < 3; c++) fm.c[c][l] = E(0);
      break;
    case kSlip:
    case k2D: {
      const int s = gid(g, iclamp(i, 1, g.n[0]), iclamp(j, 1, g.n[1]), iclamp(k, 1, g.n[2]));
      for (int c = 0; c < 3; c++) fm.c[c][l] = fm.c[c][s];
      break;
    }
    default:
      atomicOr(&st->err, (int)kErrBcFlag);
  }
}

// Ghosted flowmap from per-origin values: exchange_lpmap (fm(origin) = dx,
// comms_m.f90:1340-1359), then exchange_sdata + set_flowmap_bc.  Periodic
// ghosts copy the raw value, interior MASK/INFLOW/OUTFLOW nodes become 0,
// INTERNAL ghosts without a periodic source hold the reset value 0.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_field_from_origin(GridDev<T> g, CF3<T> src, F3<T> dst) {
  const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= g.nt) return;
  const int i = (int)(l % g.N[0]);
  const int r = (int)(l / g.N[0]);
  const int j = r % g.N[1], k = r / g.N[1];
  const int f = g.flag[l];
  const bool ib = i >= 1 && j >= 1 && k >= 1 && i <= g.n[0] && j <= g.n[1] && k <= g.n[2];
  if (ib) {
    const int p = pid(g, i, j, k);
    const bool z = (f == kMask || f == kInflow || f == kOutflow);
    for (int c = 0; c < 3; c++) dst.c[c][l] = z ? T(0) : src.c[c][p];
    return;
  }
  const int ei = i < 1 ? -1 : (i > g.n[0] ? 1 : 0), ej = j < 1 ? -1 : (j > g.n[1] ? 1 : 0),
            ek = k < 1 ? -1 : (k > g.n[2] ? 1 : 0);
  int sp = -1;
  if (f == kInternal) {
    if (self_comm(g, ei, ej, ek))
      sp = pid(g, ei < 0 ? g.n[0] : (ei > 0 ? 1 : i), ej < 0 ? g.n[1] : (ej > 0 ? 1 : j),
               ek < 0 ? g.n[2] : (ek > 0 ? 1 : k));
  } else if (f == kSlip || f == k2D) {
    sp = pid(g, iclamp(i, 1, g.n[0]), iclamp(j, 1, g.n[1]), iclamp(k, 1, g.n[2]));
  }
  for (int c = 0; c < 3; c++) dst.c[c][l] = sp >= 0 ? src.c[c][sp] : T(0);
}

// ---------------------------------------------------------------------------
// Stage (a): backward semi-Lagrangian sub-cycle (flowmap_m.f90:17-104).
// One thread per grid node: back-trace from the node with the blended
// velocity (bracket of the origin node, lp_motion_m.f90:171-327), particle
// BC, old flowmap at the departure point, displacement form write.
// ---------------------------------------------------------------------------
template <typename T, typename S, int FIX = -1, class I, typename FM, class H = NoHook>
__device__ __forceinline__ void sl_node(const GridDev<T>& g, const StageParams<T>& sp, const I& interp,
                                        const CF3<FM>& fmold, const F3<FM>& fmnew, int i, int j, int k,
                                        DevStatus* st, bool active = true, const H& hook = H{}) {
  constexpr bool NR = I::kNonRect;
  int nd[3] = {i, j, k};
  T x0[3];
  node_pos<NR>(g, i, j, k, x0);
  T xp[3] = {x0[0], x0[1], x0[2]};
  unsigned long long oob = 0;
  int err = 0;
  const int l = gid(g, i, j, k);
  // first stage at the node itself: the interpolant there is the node's
  // time-blended value (exactly for NEAREST/LINEAR, to rounding for the
  // interpolating splines), so it is read instead of gathered
  auto first = [&](T up[3]) {
#pragma clang fp contract(off)
    const T w0 = T(1) - sp.ct0;
#pragma unroll
    for (int c = 0; c < 3; c++) up[c] = w0 * sp.un.c[c][l] + sp.ct0 * sp.unp1.c[c][l];
  };
  // the departure point as x0 plus the summed stage increments: the
  // displacement keeps the increments' precision instead of that of a
  // difference of two rounded |x|-sized positions
  T inc[3], xr[3];
  integrate_inc<S, FIX>(g, sp, interp, first, nd, x0, inc, oob, (const T*)nullptr, hook);
#pragma unroll
  for (int d = 0; d < 3; d++) {
    xr[d] = x0[d] + inc[d];
    xp[d] = xr[d];
  }
  project_flat<NR>(g, nd, xp);
  if (set_lp_bc<NR>(g, nd, xp, err)) {
    xp[0] = x0[0]; xp[1] = x0[1]; xp[2] = x0[2];
  }
  int no[3] = {i, j, k};
  T fmp[3];
  // the old flowmap at the departure point, located from the node and the
  // increment unless a boundary condition moved the point (then from the
  // moved point plus a zero offset, which interpolates exactly as the
  // point itself: (q - x) + 0 = q - x)
  const bool split = xp[0] == xr[0] && xp[1] == xr[1] && xp[2] == xr[2];
  const Split<T> sq{{split ? x0[0] : xp[0], split ? x0[1] : xp[1], split ? x0[2] : xp[2]},
                    {split ? inc[0] : T(0), split ? inc[1] : T(0), split ? inc[2] : T(0)}};
  hook(std::integral_constant<int, 4>{});  // before the flowmap gather
  interp(g, xp, no, fmold, fmp, oob, &sq);
  const int f = g.flag[l];
  const T mask = (f != kMask) ? T(1) : T(0);
  const bool zero = (f == kMask || f == kInflow || f == kOutflow);
  if (!active) return;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    // a boundary projection or stick moved the point: its displacement
    const T disp = xp[d] == xr[d] ? inc[d] : xp[d] - x0[d];
    T v = (fmp[d] + disp) * mask;
    fmnew.c[d][l] = zero ? FM(0) : FM(v);
  }
  flush_status(st, err, oob);
}

// Waves per SIMD the semi-Lagrangian step asks for, where they fit with no
// extra spilling and keep more row loads in flight: FP32 fields 5 (96
// VGPRs; 6 spills and runs 20 % slower), FP16 records 6 (80 VGPRs).  The
// 3-D quadratic spline, BF16 records and non-rectilinear grids would spill
// and keep the compiler's choice.
template <class I>
struct QuadraticSpline : std::false_type {};
// FP32 stage fields under FP64 compute
template <typename T, typename S>
struct MixedF32 : std::integral_constant<bool, std::is_same<T, double>::value && std::is_same<S, float>::value> {};
template <typename T>
struct QuadraticSpline<IntSpline<T, 3, 3>> : std::true_type {};
template <typename T, typename S, class I>
constexpr int sl_min_waves() {
  return (std::is_same<T, float>::value && std::is_same<S, T>::value && !I::kNonRect && !QuadraticSpline<I>::value)
             ? 5
             : ((std::is_same<S, half4_t>::value && !I::kNonRect && !QuadraticSpline<I>::value)
                    ? 6
                    : ((MixedF32<T, S>::value && !I::kNonRect && !QuadraticSpline<I>::value) ? 4 : 1));
}
template <typename T, typename S, class I, typename FM>
__global__ void __launch_bounds__(kBlock, (sl_min_waves<T, S, I>())) k_sl_step(GridDev<T> g, StageParams<T> sp, CF3<FM> fmold, F3<FM> fmnew,
                                                    DevStatus* st) {
  const int p = xcd_thread_index();
  // 16-bit stage records and FP32 fields under FP64 compute: the basis
  // tables are a large share of what the step reads, and staging them in LDS
  // pays (2.21 -> 1.80 ms and 5.24 -> 3.47 ms at 256^3); with fields in the
  // compute type (FP32 or FP64) the staging costs more than it saves
  if constexpr (SplineInterp<I>::value && (Packed<S>::value || MixedF32<T, S>::value)) {
    constexpr int KX = I::kKX, KZ = I::kKZ;
    __shared__ T tbuf[spline_tables_lds<KX, KZ>(kTabX, kTabY, kTabY)];
    __shared__ int tred[6];
    const bool active = p < g.np;
    int nd[3] = {1, 1, 1};
    if (active) decode_interior_slab(g, p, nd[0], nd[1], nd[2]);
    I interp{};
    if (sp.lds) {
      const int cap[3] = {kTabX, kTabY, kTabY};
      stage_spline_tables<T, KX, KZ>(g, nd, active, (lptr<T>)tbuf, cap, (lptr<int>)tred, interp.lds);
      interp.use_lds = true;
    }
    if (!active) return;
    sl_node<T, S>(g, sp, interp, fmold, fmnew, nd[0], nd[1], nd[2], st);
  } else {
    const I interp{};
    if (p >= g.np) return;
    int i, j, k;
    decode_interior_slab(g, p, i, j, k);
    sl_node<T, S>(g, sp, interp, fmold, fmnew, i, j, k, st);
  }
}

// ---------------------------------------------------------------------------
// Stage (a) with the brick's coefficient tiles in LDS (cubic B-spline, 3-D
// grid, FP32): a workgroup owns a brick of kBrickX x kBrickY x kBrickZ nodes.
// The fields the step gathers from are staged into one LDS tile in the order
// the scheme first needs them (RK4: the half-step stage field, the
// full-step stage field, the old flowmap's coefficients), each from two
// cells below to two cells above the brick, with coalesced loads.  While a
// stage computes from the tile, the next field's loads are already in flight
// into registers; at the next field's first use the workgroup synchronises,
// writes them into the tile and synchronises again.  One 3-component tile
// (21 KB) leaves room for four or more workgroups per CU, which hide those
// synchronisations.  A point within half a cell of its node (every stage and
// departure point under CFL <= 0.5) finds its whole 4 x 4 x 4 support in the
// tile and sums it with LDS reads at compile-time offsets from one address;
// any other point (or field) gathers from global memory.  A wave is one
// z-layer of the brick, two rows of 32 nodes, so each half-wave's reads at
// one stencil offset touch 32 consecutive words: 32 distinct banks.  The basis
// tables and knots of the brick's intervals are staged too.  The sums run in
// spline_sum's order, so results equal k_sl_step's bit for bit; what changes
// is where the bytes come from: the texture path moves about 20 staged
// values per node and field instead of about 200 gathered ones.
// ---------------------------------------------------------------------------
constexpr int kBrickX = 32, kBrickY = 2, kBrickZ = 4;

// a point whose support leaves the tile (or a field not staged): the
// global-memory sums of spline_sum, kept out of line so that the kernel's
// interpolation sites share one copy of this rare path (weights passed as
// scalars, so that nothing goes through the stack)
template <typename T>
struct Sum3 {
  T v[3];
};
template <typename T, typename E>
__device__ __noinline__ Sum3<T> spline_sum3_global(const E* __restrict__ c0, const E* __restrict__ c1,
                                                   const E* __restrict__ c2, int base, int sy, int sz, T x0, T x1,
                                                   T x2, T x3, T y0, T y1, T y2, T y3, T z0, T z1, T z2, T z3) {
  Sum3<T> r;
  const T wx[4] = {x0, x1, x2, x3}, wy[4] = {y0, y1, y2, y3}, wz[4] = {z0, z1, z2, z3};
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const E* p = (c == 0 ? c0 : (c == 1 ? c1 : c2)) + base;
    T acc = T(0);
#pragma unroll
    for (int cz = 0; cz < 4; cz++) {
      T accy = T(0);
#pragma unroll
      for (int by = 0; by < 4; by++) {
        const E* row = p + by * sy + cz * sz;
        T accx = T(0);
#pragma unroll
        for (int a = 0; a < 4; a++) accx += wx[a] * T(row[a]);
        accy += wy[by] * accx;
      }
      acc += wz[cz] * accy;
    }
    r.v[c] = acc;
  }
  return r;
}

// adapts the brick kernel's interpolation lambda to the interpolant
// interface sl_node and integrate_inc call
template <class F>
struct BrickInterp {
  static constexpr bool kNonRect = false;
  const F& f;
  template <typename T, typename S>
  __device__ __forceinline__ void operator()(const GridDev<T>& g, const T q[3], const int nd[3], const CF3<S>& fld,
                                             T out[3], unsigned long long& oob, const Split<T>* sq = nullptr) const {
    f(g, q, nd, fld, out, oob, sq);
  }
};

// LDS rows of the staged basis tables per axis: a brick's nodes plus the
// reach of points within a cell of them
constexpr int kBrickTabX = kBrickX + 8, kBrickTabY = kBrickY + 8, kBrickTabZ = kBrickZ + 8;

// Tiles of the fields' element type E: the compute type T (FP32 or FP64),
// or FP32 under FP64 compute (FP64 arithmetic needs about twice the
// registers: two waves per SIMD)
template <typename T, typename E, int BX, int BY, int BZ, int INTEG>
__global__ void __launch_bounds__(BX * BY * BZ, sizeof(T) == 8 ? 2 : 4)
    k_sl_brick(GridDev<T> g, StageParams<T> sp, CF3<E> fmold, F3<E> fmnew, DevStatus* st, int nbx, int nby) {
  // stencil reach: a cubic point's support starts 2 nodes below its node
  // at the lowest and ends 2 above at the highest
  constexpr int TX = BX + 4, TY = BY + 4, TZ = BZ + 4, TILE = TX * TY * TZ;
  constexpr int NT = BX * BY * BZ, PER = (TILE + NT - 1) / NT;
  // fields in the order the scheme first gathers them: RK4 stage 2 (its
  // stage 3 reads the same half-step field), stage 4, then the flowmap;
  // RK3 stages 2 and 3; RK2 stage 2
  constexpr int NSEQ = INTEG == kRK2 ? 2 : 3;
  __shared__ E tile[3 * TILE];
  __shared__ T tbuf[spline_tables_lds<4, 4>(kBrickTabX, kBrickTabY, kBrickTabZ)];
  const int b = xcd_brick(blockIdx.x, gridDim.x);
  const int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
  const int t = threadIdx.x;
  const int o0 = bx * BX - 1, o1 = by * BY - 1, o2 = bz * BZ - 1;  // ghosted index of tile cell (0, 0, 0)
  constexpr int A1 = NSEQ == 3 ? (INTEG == kRK4 ? 3 : 2) : 1;
  const E* f0[3] = {(const E*)sp.v[1][0], (const E*)sp.v[1][1], (const E*)sp.v[1][2]};
  const E* f1[3] = {NSEQ == 3 ? (const E*)sp.v[A1][0] : fmold.c[0], NSEQ == 3 ? (const E*)sp.v[A1][1] : fmold.c[1],
                    NSEQ == 3 ? (const E*)sp.v[A1][2] : fmold.c[2]};
  const E* f2[3] = {fmold.c[0], fmold.c[1], fmold.c[2]};
  // component-0 pointer of field q of the sequence (its identity)

  // this thread's tile cells; cells outside the ghosted grid hold zero (no
  // point's support reaches them)
  int gidx[PER];
  bool cell[PER];
#pragma unroll
  for (int it = 0; it < PER; it++) {
    const int e = t + it * NT;
    const int x = e % TX, y = (e / TX) % TY, z = e / (TX * TY);
    const int gx = o0 + x, gy = o1 + y, gz = o2 + z;
    cell[it] = e < TILE && (unsigned)gx < (unsigned)g.N[0] && (unsigned)gy < (unsigned)g.N[1] &&
               (unsigned)gz < (unsigned)g.N[2];
    gidx[it] = cell[it] ? gx + g.N[0] * (gy + g.N[1] * gz) : 0;
  }
  E pre[PER][3];
  auto prefetch = [&](const E* a, const E* b2, const E* c2) __attribute__((always_inline)) {
#pragma unroll
    for (int it = 0; it < PER; it++) {
      pre[it][0] = a[gidx[it]];
      pre[it][1] = b2[gidx[it]];
      pre[it][2] = c2[gidx[it]];
    }
  };
  prefetch(f0[0], f0[1], f0[2]);
  // basis tables of the brick's intervals (visible after the first field's
  // synchronisation, before any point is located)
  SplineLds<T> lds;
  const int lo[3] = {bx * BX + 1, by * BY + 1, bz * BZ + 1};
  const int hi[3] = {min(lo[0] + BX - 1, g.n[0]), min(lo[1] + BY - 1, g.n[1]), min(lo[2] + BZ - 1, g.n[2])};
  const int cap[3] = {kBrickTabX, kBrickTabY, kBrickTabZ};
  copy_spline_tables<T, 4, 4>(g, lo, hi, (lptr<T>)tbuf, cap, lds);
  // The fields change at fixed points of the step, the same for every
  // thread: the hook runs before each gather with the stage number as a
  // type (1..3 the scheme's stages, 4 the flowmap), so which field is staged
  // where is fixed at compile time, and every thread reaches the
  // synchronisations.  RK4's stage 3 reads stage 2's half-step field.
  const void* cur_key = nullptr;  // component-0 pointer of the staged field
  auto hook = [&](auto K) __attribute__((always_inline)) {
    constexpr int k = decltype(K)::value;
    constexpr int need = k == 4 ? NSEQ - 1 : (INTEG == kRK4 ? (k <= 2 ? 0 : 1) : k - 1);
    if constexpr (!(INTEG == kRK4 && k == 2)) {
      __syncthreads();  // every thread is done with the previous field
#pragma unroll
      for (int it = 0; it < PER; it++) {
        const int e = t + it * NT;
        if (it < PER - 1 || e < TILE) {
#pragma unroll
          for (int c = 0; c < 3; c++) tile[c * TILE + e] = cell[it] ? pre[it][c] : E(0);
        }
      }
      __syncthreads();
      if constexpr (need == 0) prefetch(f1[0], f1[1], f1[2]);
      if constexpr (need == 1 && NSEQ == 3) prefetch(f2[0], f2[1], f2[2]);
      cur_key = need == 0 ? (const void*)f0[0] : (need == 1 ? (const void*)f1[0] : (const void*)f2[0]);
    }
  };
  auto interp = [&](const GridDev<T>& gg, const T q[3], const int nd[3], const CF3<E>& f, T out[3],
                    unsigned long long& oob, const Split<T>* sq) __attribute__((always_inline)) {
    const void* k0 = f.c[0];
    const void* kst = cur_key;
    SplineW<T, 4, 4> s;
    const int r = spline_setup(gg, q, nd, s, sq, &lds);
    if (r) {
      if (r == 1) oob++;
      out[0] = out[1] = out[2] = T(0);
      return;
    }
    const unsigned lx = (unsigned)(s.b3[0] - o0), ly = (unsigned)(s.b3[1] - o1), lz = (unsigned)(s.b3[2] - o2);
    if (k0 == kst && lx <= (unsigned)(TX - 4) && ly <= (unsigned)(TY - 4) &&
        lz <= (unsigned)(TZ - 4)) {
      const lcptr<E> tb = (lcptr<E>)tile + ((int)lz * TY + (int)ly) * TX + (int)lx;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        T acc = T(0);
#pragma unroll
        for (int cz = 0; cz < 4; cz++) {
          T accy = T(0);
#pragma unroll
          for (int by2 = 0; by2 < 4; by2++) {
            T accx = T(0);
#pragma unroll
            for (int a = 0; a < 4; a++) accx += s.wx[a] * T(tb[c * TILE + (cz * TY + by2) * TX + a]);
            accy += s.wy[by2] * accx;
          }
          acc += s.wz[cz] * accy;
        }
        out[c] = acc;
      }
    } else {
      const Sum3<T> r3 = spline_sum3_global<T, E>(f.c[0], f.c[1], f.c[2], s.base, gg.N[0], gg.N[0] * gg.N[1], s.wx[0],
                                               s.wx[1], s.wx[2], s.wx[3], s.wy[0], s.wy[1], s.wy[2], s.wy[3], s.wz[0],
                                               s.wz[1], s.wz[2], s.wz[3]);
#pragma unroll
      for (int c = 0; c < 3; c++) out[c] = r3.v[c];
    }
  };
  const BrickInterp<decltype(interp)> bi{interp};
  // every thread runs the node (clamped into the grid) so that each reaches
  // the field synchronisations; only nodes of the grid write
  const int i = lo[0] + t % BX, j = lo[1] + (t / BX) % BY, k = lo[2] + t / (BX * BY);
  const bool active = i <= g.n[0] && j <= g.n[1] && k <= g.n[2];
  sl_node<T, E, INTEG>(g, sp, bi, fmold, fmnew, min(i, g.n[0]), min(j, g.n[1]), min(k, g.n[2]), st, active, hook);
}

// time blend u = (1-ct) u_n + ct u_np1 over the ghosted grid, in the
// reference's expression order (lp_motion_m.f90:171-173).  16-bit storage
// keeps the blend times inv, a power of two (exact), so that values and
// spline coefficients sit in the format's normal range; a value that would
// still overflow FP16 sets kErrRange instead of turning into inf.
template <typename T, typename S, typename A>
__global__ void __launch_bounds__(kBlock) k_blend(F3<S> out, CF3<A> a, CF3<A> b, T ct, T inv, long long nt,
                                                  DevStatus* st) {
#pragma clang fp contract(off)
  const T w0 = T(1) - ct;
  bool range = false;
  for (long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x; l < nt; l += (long long)gridDim.x * blockDim.x) {
    if constexpr (Packed<S>::value) {
      using E = typename Packed<S>::E;
      S r;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const T x = (w0 * T(a.c[c][l]) + ct * T(b.c[c][l])) * inv;
        if constexpr (std::is_same<E, half_t>::value) range |= fabs(x) > T(65504);
        r.e[c] = narrow<E>(x);
      }
      r.e[3] = E{0};
      out.c[0][l] = r;
    } else {
#pragma unroll
      for (int c = 0; c < 3; c++) out.c[c][l] = narrow<S>(w0 * T(a.c[c][l]) + ct * T(b.c[c][l]));
    }
  }
  if (range) atomicOr(&st->err, (int)kErrRange);
}

// ---------------------------------------------------------------------------
// Stage (a'): forward particle sub-cycle (update_lp, lp_motion_m.f90:15-54):
// integrate, BC, stick, dx accumulation, one tracking pass, periodic wrap.
// ---------------------------------------------------------------------------
// FP16 records on a rectilinear grid: 4 waves per SIMD fit without spilling
// (114-120 VGPRs) and keep more row loads in flight; the other variants keep
// the compiler's register budget (FP32 and BF16 would spill)
template <typename T, typename S, class I>
constexpr int fwd_min_waves() {
  return ((std::is_same<S, half4_t>::value || MixedF32<T, S>::value) && !I::kNonRect) ? 4 : 1;
}
// One forward particle sub-cycle (update_lp): integrate, BC, stick, dx
// accumulation, one tracking pass, periodic wrap.
template <typename T, typename S, bool COMP, class I>
__device__ __forceinline__ void fwd_particle(const GridDev<T>& g, const StageParams<T>& sp, const I& interp,
                                             const ParticleSet<T>& ps, int p, DevStatus* st) {
  constexpr bool NR = I::kNonRect;
  T q[3] = {ps.x[0][p], ps.x[1][p], ps.x[2][p]};
  int nd[3] = {ps.n[0][p], ps.n[1][p], ps.n[2][p]};
  const T xn[3] = {q[0], q[1], q[2]};
  unsigned long long oob = 0;
  int err = 0;
  // compensated positions: the low parts, which every sample point carries
  T xnl[3] = {T(0), T(0), T(0)};
  if constexpr (COMP) {
#pragma unroll
    for (int d = 0; d < 3; d++) xnl[d] = ps.xl[d][p];
  }
  const Split<T> s0{{xn[0], xn[1], xn[2]}, {xnl[0], xnl[1], xnl[2]}};
  auto first = [&](T up[3]) {
    interp(g, xn, nd, stage<S>(sp, 0), up, oob, COMP ? &s0 : nullptr);
    stage_scale<S>(sp, up);
  };
  if constexpr (!COMP) {
    // displacement from the summed stage increments (see sl_node)
    T inc[3], qr[3];
    integrate_inc<S>(g, sp, interp, first, nd, xn, inc, oob);
#pragma unroll
    for (int d = 0; d < 3; d++) {
      qr[d] = xn[d] + inc[d];
      q[d] = qr[d];
    }
    project_flat<NR>(g, nd, q);
    if (set_lp_bc<NR>(g, nd, q, err)) {
      q[0] = xn[0]; q[1] = xn[1]; q[2] = xn[2];
    }
#pragma unroll
    for (int d = 0; d < 3; d++) ps.d[d][p] = ps.d[d][p] + (q[d] == qr[d] ? inc[d] : q[d] - xn[d]);
    track_pass<NR>(g, q, nd);
    wrap(g, nd, q);
  } else {
    // position and displacement as unevaluated sums hi + lo, advanced with
    // two-sum; BC moves and stick reset the low part of what they change
    T ql[3], inc[3];
    integrate_inc<S>(g, sp, interp, first, nd, xn, inc, oob, xnl);
#pragma unroll
    for (int d = 0; d < 3; d++) two_sum(xn[d], inc[d] + xnl[d], q[d], ql[d]);
    const T qb[3] = {q[0], q[1], q[2]};
    project_flat<NR>(g, nd, q);
    bool stick = set_lp_bc<NR>(g, nd, q, err);
#pragma unroll
    for (int d = 0; d < 3; d++) {
      if (stick) {
        q[d] = xn[d];
        ql[d] = xnl[d];
      } else if (q[d] != qb[d]) {
        ql[d] = T(0);
      }
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
      T step;
      {
#pragma clang fp contract(off)
        step = (q[d] - xn[d]) + (ql[d] - xnl[d]);
      }
      T s, e;
      two_sum(ps.d[d][p], step + ps.dl[d][p], s, e);
      ps.d[d][p] = s;
      ps.dl[d][p] = e;
    }
    track_pass<NR>(g, q, nd);
    const T qw[3] = {q[0], q[1], q[2]};
    T sh[3];
    wrap(g, nd, q, sh);
#pragma unroll
    for (int d = 0; d < 3; d++) {
      // the periodic shift's rounding goes into the low part
      T s, e;
      two_sum(qw[d], sh[d], s, e);
      ps.xl[d][p] = ql[d] + e;
    }
  }
#pragma unroll
  for (int d = 0; d < 3; d++) ps.x[d][p] = q[d];
  ps.n[0][p] = (int16_t)nd[0]; ps.n[1][p] = (int16_t)nd[1]; ps.n[2][p] = (int16_t)nd[2];
  flush_status(st, err, oob);
}

template <typename T, typename S, class I, bool COMP>
__global__ void __launch_bounds__(kBlock, (fwd_min_waves<T, S, I>())) k_fwd_step(GridDev<T> g, StageParams<T> sp, ParticleSet<T> ps, int np,
                                                     DevStatus* st) {
  const int p = xcd_thread_index();
  I interp{};
  if constexpr (SplineInterp<I>::value) {
    // the basis tables of the intervals the workgroup's particles are in
    constexpr int KX = I::kKX, KZ = I::kKZ;
    __shared__ T tbuf[spline_tables_lds<KX, KZ>(kTabX, kTabY, kTabY)];
    __shared__ int tred[6];
    if (sp.lds) {
      const bool active = p < np;
      int n0[3] = {1, 1, 1};
      if (active) {
        n0[0] = ps.n[0][p];
        n0[1] = ps.n[1][p];
        n0[2] = ps.n[2][p];
      }
      const int cap[3] = {kTabX, kTabY, kTabY};
      stage_spline_tables<T, KX, KZ>(g, n0, active, (lptr<T>)tbuf, cap, (lptr<int>)tred, interp.lds);
      interp.use_lds = true;
    }
  }
  if (p >= np) return;
  fwd_particle<T, S, COMP>(g, sp, interp, ps, p, st);
}

// reset_lp for forward particles (lp_m.f90:228-294, resolution 0)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_fwd_reset(GridDev<T> g, ParticleSet<T> ps) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.np) return;
  int i, j, k;
  decode_interior(g, p, i, j, k);
  T x[3];
  if (g.rect) node_pos<false>(g, i, j, k, x);
  else node_pos<true>(g, i, j, k, x);
  ps.x[0][p] = x[0]; ps.x[1][p] = x[1]; ps.x[2][p] = x[2];
  ps.d[0][p] = T(0); ps.d[1][p] = T(0); ps.d[2][p] = T(0);
  ps.xl[0][p] = T(0); ps.xl[1][p] = T(0); ps.xl[2][p] = T(0);
  ps.dl[0][p] = T(0); ps.dl[1][p] = T(0); ps.dl[2][p] = T(0);
  ps.n[0][p] = (int16_t)i; ps.n[1][p] = (int16_t)j; ps.n[2][p] = (int16_t)k;
  ps.org[p] = p;
}

// ---------------------------------------------------------------------------
// Diagnostic grids other than the flow grid: resolution != 0
// (new_sgrid_from_sgrid, sgrid_m.f90:445-761) and the AUX_GRID sets
// (create_aux_grid, lcs_m.f90:227-435).  Particles sit on the nodes of the
// diagnostic grid lg; their velocity, particle BCs and tracking use the flow
// grid g through the tracked flow node (lp%no_scfd), the flowmap and its
// interpolation the lg node (lp%no).
// ---------------------------------------------------------------------------
// reset_lp's flow-node guess nint(real(i*n_flow)/real(n)) in default REAL,
// clamped to [1, n_flow] (lp_m.f90:274-277)
__device__ __forceinline__ int guess_node(int i, int nf, int nl) {
  const double r = (double)((float)(i * nf) / (float)nl);
  const int v = (int)(r >= 0.0 ? floor(r + 0.5) : -floor(-r + 0.5));
  return iclamp(v, 1, nf);
}

// recursive track_lp2node onto g (lp_tracking_m.f90:16-183) of the particle
// at q from the node nd: passes until the node stays, each followed for
// forward sets by exchange_lpdata's periodic wrap (the shifts summed into
// sh).  A particle that did not move stays unmoved in later passes, so each
// runs to its own fixed point; false after more than 2 max(n) passes.
template <typename T>
__device__ bool track_recursive(const GridDev<T>& g, bool fwd, T q[3], int nd[3], T sh[3]) {
  for (int pass = 1;; pass++) {
    const bool moved = track_pass(g, q, nd);
    if (pass > g.max_track) return false;
    if (fwd) {
      T w[3];
      wrap(g, nd, q, w);
#pragma unroll
      for (int d = 0; d < 3; d++) sh[d] = sh[d] + w[d];
    }
    if (!moved) return true;
  }
}

// init_lp / reset_lp + track_lp2node of a backward set (no exchange): the
// tracked flow node of every lg node, which may be a ghost node when the lg
// node is nearer to a periodic ghost than to any interior node
template <typename T>
__global__ void __launch_bounds__(kBlock) k_track_nodes(GridDev<T> g, GridDev<T> lg, int16_t* __restrict__ nodes,
                                                        DevStatus* st) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= lg.np) return;
  int i, j, k;
  decode_interior(lg, p, i, j, k);
  T q[3] = {lg.gl[0][i], lg.gl[1][j], lg.gl[2][k]}, sh[3] = {T(0), T(0), T(0)};
  int nd[3] = {guess_node(i, g.n[0], lg.n[0]), guess_node(j, g.n[1], lg.n[1]), guess_node(k, g.n[2], lg.n[2])};
  if (!track_recursive(g, false, q, nd, sh)) atomicOr(&st->err, (int)kErrTrack);
  nodes[p] = (int16_t)nd[0];
  nodes[lg.np + p] = (int16_t)nd[1];
  nodes[2LL * lg.np + p] = (int16_t)nd[2];
}

// reset_lp + track_lp2node of a forward set: particles on the lg nodes,
// tracked onto g with exchange_lpdata's wrap after every pass
template <typename T>
__global__ void __launch_bounds__(kBlock) k_fwd_place(GridDev<T> g, GridDev<T> lg, ParticleSet<T> ps,
                                                      DevStatus* st) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= lg.np) return;
  int i, j, k;
  decode_interior(lg, p, i, j, k);
  const T x0[3] = {lg.gl[0][i], lg.gl[1][j], lg.gl[2][k]};
  T q[3] = {x0[0], x0[1], x0[2]}, sh[3] = {T(0), T(0), T(0)};
  int nd[3] = {guess_node(i, g.n[0], lg.n[0]), guess_node(j, g.n[1], lg.n[1]), guess_node(k, g.n[2], lg.n[2])};
  if (!track_recursive(g, true, q, nd, sh)) atomicOr(&st->err, (int)kErrTrack);
#pragma unroll
  for (int d = 0; d < 3; d++) {
    ps.x[d][p] = q[d];
    ps.d[d][p] = T(0);
    ps.dl[d][p] = T(0);
    // the low part of the node coordinate plus its periodic shifts
    T hi, lo;
    two_sum(x0[d], sh[d], hi, lo);
    ps.xl[d][p] = hi == q[d] ? lo : T(0);
    ps.n[d][p] = (int16_t)nd[d];
  }
  ps.org[p] = p;
}

// update_flowmap_sl (flowmap_m.f90:17-104) for one node of lg: the stages
// interpolate the velocity on g through the tracked flow node, the BC and
// the MASK test use it too (set_lp_bc clamps lp%no_scfd to the interior in
// place, so the tracked node is used as it is only until the first
// sub-cycle after a reset: bits of fresh), and the old flowmap is
// interpolated on lg through the node itself.  The flowmap BC of lg zeroes
// MASK/INFLOW/OUTFLOW nodes of lg (set_flowmap_bc, flowmap_m.f90:345-360).
template <typename T, typename S, class I, typename FM>
__global__ void __launch_bounds__(kBlock) k_sl_step2(GridDev<T> g, GridDev<T> lg, const int16_t* __restrict__ nodes,
                                                     int fresh, StageParams<T> sp, CF3<FM> fmold, F3<FM> fmnew,
                                                     DevStatus* st) {
  const int w = xcd_thread_index();
  if (w >= lg.np) return;
  int i, j, k;
  decode_interior_slab(lg, w, i, j, k);
  const int p = pid(lg, i, j, k);
  const I interp{};
  const T x0[3] = {lg.gl[0][i], lg.gl[1][j], lg.gl[2][k]};
  int nd[3], nm[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const int t = nodes[(long long)d * lg.np + p];
    const int c = iclamp(t, 1, g.n[d]);
    nd[d] = (fresh & 1) ? t : c;
    nm[d] = (fresh & 2) ? t : c;
  }
  const T mask = g.flag[gid(g, nm[0], nm[1], nm[2])] != kMask ? T(1) : T(0);
  unsigned long long oob = 0;
  int err = 0;
  auto first = [&](T up[3]) {
    interp(g, x0, nd, stage<S>(sp, 0), up, oob);
    stage_scale<S>(sp, up);
  };
  T inc[3], xr[3], xp[3];
  integrate_inc<S>(g, sp, interp, first, nd, x0, inc, oob);
#pragma unroll
  for (int d = 0; d < 3; d++) {
    xr[d] = x0[d] + inc[d];
    xp[d] = xr[d];
  }
  project_flat(g, nd, xp);
  if (set_lp_bc(g, nd, xp, err)) {
    xp[0] = x0[0]; xp[1] = x0[1]; xp[2] = x0[2];
  }
  int no[3] = {i, j, k};
  T fmp[3];
  // as in sl_node: a point a boundary condition moved is the moved point
  // plus a zero offset
  const bool split = xp[0] == xr[0] && xp[1] == xr[1] && xp[2] == xr[2];
  const Split<T> sq{{split ? x0[0] : xp[0], split ? x0[1] : xp[1], split ? x0[2] : xp[2]},
                    {split ? inc[0] : T(0), split ? inc[1] : T(0), split ? inc[2] : T(0)}};
  interp(lg, xp, no, fmold, fmp, oob, &sq);
  const int l = gid(lg, i, j, k);
  const int f = lg.flag[l];
  const bool zero = (f == kMask || f == kInflow || f == kOutflow);
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const T disp = xp[d] == xr[d] ? inc[d] : xp[d] - x0[d];
    const T v = (fmp[d] + disp) * mask;
    fmnew.c[d][l] = zero ? FM(0) : FM(v);
  }
  flush_status(st, err, oob);
}

// a field read back from its interior values (structured_io + exchange_sdata
// of a zeroed sr1, flowmap_m.f90:228-232): periodic ghosts are the images,
// every other ghost zero
template <typename T>
__global__ void __launch_bounds__(kBlock) k_expand_periodic(GridDev<T> g, const T* __restrict__ src, F3<T> dst) {
  const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= g.nt) return;
  const int i = (int)(l % g.N[0]), j = (int)((l / g.N[0]) % g.N[1]), k = (int)(l / ((long long)g.N[0] * g.N[1]));
  const int ei = i < 1 ? -1 : (i > g.n[0] ? 1 : 0), ej = j < 1 ? -1 : (j > g.n[1] ? 1 : 0),
            ek = k < 1 ? -1 : (k > g.n[2] ? 1 : 0);
  const bool ghost = ei || ej || ek;
  if (ghost && !self_comm(g, ei, ej, ek)) {
#pragma unroll
    for (int c = 0; c < 3; c++) dst.c[c][l] = T(0);
    return;
  }
  const int si = ei < 0 ? g.n[0] : (ei > 0 ? 1 : i), sj = ej < 0 ? g.n[1] : (ej > 0 ? 1 : j),
            sk = ek < 0 ? g.n[2] : (ek > 0 ? 1 : k);
  const long long q = pid(g, si, sj, sk);
#pragma unroll
  for (int c = 0; c < 3; c++) dst.c[c][l] = src[(long long)c * g.np + q];
}

// interp_s2u_r1 of a field on g at the lg nodes through their clamped
// tracked flow nodes (the velocity at T0 of a diagnostic grid,
// flowmap_m.f90:234-241)
template <typename T, class I>
__global__ void __launch_bounds__(kBlock) k_field_at_nodes(GridDev<T> g, GridDev<T> lg,
                                                           const int16_t* __restrict__ nodes, CF3<T> f,
                                                           T* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= lg.np) return;
  int i, j, k;
  decode_interior(lg, p, i, j, k);
  const T q[3] = {lg.gl[0][i], lg.gl[1][j], lg.gl[2][k]};
  int nd[3];
#pragma unroll
  for (int d = 0; d < 3; d++) nd[d] = iclamp((int)nodes[(long long)d * lg.np + p], 1, g.n[d]);
  unsigned long long oob = 0;
  T o[3];
  I{}(g, q, nd, f, o, oob);
#pragma unroll
  for (int c = 0; c < 3; c++) out[(long long)c * lg.np + p] = o[c];
}

// ---------------------------------------------------------------------------
// Stage (b): composition of the sub-interval flowmaps
// (reconstruct_flowmap, flowmap_m.f90:158-323).  One thread per origin node
// follows its trajectory through the slots in order: BC, interpolate the slot
// displacement, relocate (exchange_lp_alltoall, comms_m.f90:1378-1709: box /
// periodic image / node guess) and track recursively with a periodic wrap
// after each pass (the direction is FWD during reconstruction).
// ---------------------------------------------------------------------------
// One slot of the composition for one particle: the body of
// reconstruct_flowmap's slot loop (flowmap_m.f90:262-308) -- BC, the slot's
// displacement at the particle, relocation into the bounding box or its
// first periodic image (exchange_lp_alltoall, comms_m.f90:1378-1709), the
// node guess and recursive re-tracking.  q is sq.qb + sq.off on entry and
// on return.
template <bool NR, typename T>
__device__ __forceinline__ void compose_bc(const GridDev<T>& g, T q[3], Split<T>& sq, int nd[3], int& err) {
  const T qr[3] = {q[0], q[1], q[2]};
  project_flat<NR>(g, nd, q);
  set_lp_bc<NR>(g, nd, q, err);
#pragma unroll
  for (int d = 0; d < 3; d++)
    if (q[d] != qr[d]) {
      sq.qb[d] = q[d];
      sq.off[d] = T(0);
    }
}

template <bool NR, typename T>
__device__ __forceinline__ void compose_after(const GridDev<T>& g, const T up[3], T q[3], Split<T>& sq, int nd[3],
                                              T dxx[3], int& err) {
#pragma unroll
  for (int d = 0; d < 3; d++) {
    sq.off[d] = sq.off[d] + up[d];
    q[d] = sq.qb[d] + sq.off[d];
    dxx[d] = dxx[d] + up[d];
  }
  // relocation to the bounding box or its first periodic image
  const T* bb = g.bb;
  bool in = q[0] >= bb[0] && q[0] <= bb[3] && q[1] >= bb[1] && q[1] <= bb[4] && q[2] >= bb[2] && q[2] <= bb[5];
  if (!in && g.ps_sep) {
    // the shifts are separable: an image is inside the box iff each axis'
    // shifted coordinate is, and the first direction (i fastest, then j, k)
    // with an inside image takes the first inside index of every axis
    int first[3];
    bool any = true;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      first[d] = -2;
#pragma unroll
      for (int a = -1; a <= 1; a++) {
        const int c = d == 0 ? (a + 1) + 12 : (d == 1 ? 3 * (a + 1) + 10 : 9 * (a + 1) + 4);
        const T y = q[d] - g.ps[c][d];
        if (first[d] == -2 && y >= bb[d] && y <= bb[3 + d]) first[d] = a;
      }
      any = any && first[d] != -2;
    }
    if (any) {
      const int c = (first[0] + 1) + 3 * (first[1] + 1) + 9 * (first[2] + 1);
      const T y[3] = {q[0] - g.ps[c][0], q[1] - g.ps[c][1], q[2] - g.ps[c][2]};
#pragma unroll
      for (int d = 0; d < 3; d++)
        if (y[d] != q[d]) {
          sq.qb[d] = y[d];
          sq.off[d] = T(0);
        }
      q[0] = y[0]; q[1] = y[1]; q[2] = y[2];
    }
  } else if (!in) {
    for (int c = 0; c < 27 && !in; c++) {
      const T y0 = q[0] - g.ps[c][0], y1 = q[1] - g.ps[c][1], y2 = q[2] - g.ps[c][2];
      if (y0 >= bb[0] && y0 <= bb[3] && y1 >= bb[1] && y1 <= bb[4] && y2 >= bb[2] && y2 <= bb[5]) {
        const T y[3] = {y0, y1, y2};
#pragma unroll
        for (int d = 0; d < 3; d++)
          if (y[d] != q[d]) {
            sq.qb[d] = y[d];
            sq.off[d] = T(0);
          }
        q[0] = y0; q[1] = y1; q[2] = y2;
        in = true;
      }
    }
  }
#pragma unroll
  for (int d = 0; d < 3; d++) {
    T sv = clamp01((q[d] - bb[d]) / (bb[3 + d] - bb[d]));
    T v = (T)(float)g.n[d] * sv;
    nd[d] = iclamp((int)round(v), 1, g.n[d]);
  }
  if constexpr (NR) {
    // the node guess of a non-rectilinear grid from the positions of the
    // coordinate extrema (comms_m.f90:1666-1681)
    T stu[3];
#pragma unroll
    for (int d = 0; d < 3; d++) stu[d] = clamp01((q[d] - bb[d]) / (bb[3 + d] - bb[d]));
    const T third = T(1) / T(3);
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const T a = third * (g.pmin[0][d] + g.pmin[1][d] + g.pmin[2][d]);
      const T b = third * (stu[0] * (g.pmax[0][d] - g.pmin[0][d]) + stu[1] * (g.pmax[1][d] - g.pmin[1][d]) +
                           stu[2] * (g.pmax[2][d] - g.pmin[2][d]));
      nd[d] = iclamp((int)round(a) + (int)round(b), 1, g.n[d]);
    }
  }
  for (int pass = 1;; pass++) {
    bool moved = track_pass<NR>(g, q, nd);
    if (pass > g.max_track) {
      err |= kErrTrack;
      break;
    }
    T sh[3];
    wrap(g, nd, q, sh);
#pragma unroll
    for (int d = 0; d < 3; d++)
      if (sh[d] != T(0)) {
        sq.qb[d] = q[d];
        sq.off[d] = T(0);
      }
    if (!moved) break;
  }
}

// One slot of the composition for one particle: the body of
// reconstruct_flowmap's slot loop (flowmap_m.f90:262-308) -- BC, the slot's
// displacement at the particle, relocation into the bounding box or its
// first periodic image (exchange_lp_alltoall, comms_m.f90:1378-1709), the
// node guess and recursive re-tracking.  q is sq.qb + sq.off on entry and
// on return.
template <typename T, class I>
__device__ __forceinline__ void compose_slot(const GridDev<T>& g, const CF3<T>& f, T q[3], Split<T>& sq, int nd[3],
                                             T dxx[3], unsigned long long& oob, int& err) {
  constexpr bool NR = I::kNonRect;
  compose_bc<NR>(g, q, sq, nd, err);
  T up[3];
  I{}(g, q, nd, f, up, oob, &sq);
  compose_after<NR>(g, up, q, sq, nd, dxx, err);
}

// The whole composition in one launch: one thread follows one origin node
// through every slot.
template <typename T, class I>
__global__ void __launch_bounds__(kBlock) k_compose(GridDev<T> g, const T* __restrict__ slots,
                                                    long long stride, const int* __restrict__ order, int nslot,
                                                    F3<T> dxout, DevStatus* st) {
  constexpr bool NR = I::kNonRect;
  const int p = xcd_thread_index();
  if (p >= g.np) return;
  int i, j, k;
  decode_interior(g, p, i, j, k);
  int nd[3] = {i, j, k};
  T q[3];
  node_pos<NR>(g, i, j, k, q);
  T dxx[3] = {T(0), T(0), T(0)};
  // the particle as base + offset: the base is its node coordinate (exact)
  // until a boundary condition or a periodic relocation moves it, and the
  // offset accumulates the slot displacements, so that interpolation sees
  // the position without the rounding of an |x|-sized sum
  Split<T> sq{{q[0], q[1], q[2]}, {T(0), T(0), T(0)}};
  unsigned long long oob = 0;
  int err = 0;
  for (int s = 0; s < nslot; s++) {
    const T* base = slots + (long long)order[s] * 3 * stride;
    const CF3<T> f{{base, base + stride, base + 2 * stride}};
    compose_slot<T, I>(g, f, q, sq, nd, dxx, oob, err);
  }
  dxout.c[0][p] = dxx[0];
  dxout.c[1][p] = dxx[1];
  dxout.c[2][p] = dxx[2];
  flush_status(st, err, oob);
}

// The composition one slot per launch.  Between slots the particles are
// re-ordered by the brick (8^3 nodes) of their tracked node, so that the 64
// particles of a wavefront gather one slot's interpolation stencils from a
// few neighbouring bricks instead of from wherever their trajectories have
// dispersed to.  Each particle's arithmetic is exactly that of k_compose (the
// state between slots is its split position, node, displacement and origin),
// so the order changes no result.
// the first slot's particle order: whole bricks of 8 x 8 x 4 nodes (one
// workgroup each) where the grid divides into them, else slab order
template <typename T>
__device__ __forceinline__ void decode_comp_first(const GridDev<T>& g, int t, int& i, int& j, int& k) {
  if ((g.n[0] & 7) == 0 && (g.n[1] & 7) == 0 && (g.n[2] & 3) == 0) {
    const int b = t >> 8, l = t & 255, nbx = g.n[0] >> 3, nby = g.n[1] >> 3;
    i = (b % nbx) * 8 + (l & 7) + 1;
    j = ((b / nbx) % nby) * 8 + ((l >> 3) & 7) + 1;
    k = (b / (nbx * nby)) * 4 + (l >> 6) + 1;
  } else {
    decode_interior_slab(g, t, i, j, k);
  }
}

template <typename T>
__device__ __forceinline__ unsigned comp_key(const GridDev<T>& g, const int nd[3]) {
  // bricks of 8 x 8 x 4 nodes: about one workgroup's worth of particles
  const int i = (iclamp(nd[0], 1, g.n[0]) - 1) >> 3, j = (iclamp(nd[1], 1, g.n[1]) - 1) >> 3,
            k = (iclamp(nd[2], 1, g.n[2]) - 1) >> 2;
  const int nbx = (g.n[0] + 7) >> 3, nby = (g.n[1] + 7) >> 3;
  return (unsigned)((k * nby + j) * nbx + i);
}

template <typename T, class I>
__global__ void __launch_bounds__(kBlock) k_compose_slot(GridDev<T> g, CF3<T> f, const CompRec<T>* __restrict__ in,
                                                         const unsigned* __restrict__ perm, CompRec<T>* __restrict__ out,
                                                         unsigned* __restrict__ keys, unsigned* __restrict__ vals,
                                                         int first, int last, F3<T> dxout, DevStatus* st) {
  constexpr bool NR = I::kNonRect;
  const int t = xcd_thread_index();
  if (t >= g.np) return;
  T q[3], dxx[3];
  int nd[3], org;
  Split<T> sq;
  if (first) {
    int i, j, k;
    decode_comp_first(g, t, i, j, k);
    nd[0] = i; nd[1] = j; nd[2] = k;
    node_pos<NR>(g, i, j, k, q);
#pragma unroll
    for (int d = 0; d < 3; d++) {
      sq.qb[d] = q[d];
      sq.off[d] = T(0);
      dxx[d] = T(0);
    }
    org = pid(g, i, j, k);
  } else {
    const CompRec<T> r = in[perm ? perm[t] : (unsigned)t];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      sq.qb[d] = r.qb[d];
      sq.off[d] = r.off[d];
      q[d] = r.qb[d] + r.off[d];
      dxx[d] = r.dx[d];
      nd[d] = r.nd[d];
    }
    org = r.org;
  }
  unsigned long long oob = 0;
  int err = 0;
  compose_slot<T, I>(g, f, q, sq, nd, dxx, oob, err);
  if (last) {
#pragma unroll
    for (int d = 0; d < 3; d++) dxout.c[d][org] = dxx[d];
  } else {
    CompRec<T> r;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      r.qb[d] = sq.qb[d];
      r.off[d] = sq.off[d];
      r.dx[d] = dxx[d];
      r.nd[d] = (int16_t)nd[d];
    }
    r.pad = 0;
    r.org = org;
    out[t] = r;
    if (keys) {
      keys[t] = comp_key(g, nd);
      vals[t] = (unsigned)t;
    }
  }
  flush_status(st, err, oob);
}

// ---------------------------------------------------------------------------
// Stage (c): central gradient of the displacement flowmap (grad_sr1,
// gradient_m.f90:11-66, with the periodic face refresh of lcs_m.f90:128),
// C = G^T G (no identity term, lcs_m.f90:137-145), largest eigenvalue by
// cyclic Jacobi in FP64, ftle = log(max(sqrt(lmax), tiny))/T
// (lcs_m.f90:168-185).  Gradient, tensor and eigen-solve run in FP64 for
// both storage precisions.
// ---------------------------------------------------------------------------
// C = G^T G (no identity term), its largest eigenvalue, the FTLE
__device__ __forceinline__ double ftle_of_grad(const double G[3][3], double Tint, int incomp, double tiny) {
  const double cxx = G[0][0] * G[0][0] + G[1][0] * G[1][0] + G[2][0] * G[2][0];
  const double cxy = G[0][0] * G[0][1] + G[1][0] * G[1][1] + G[2][0] * G[2][1];
  const double cxz = G[0][0] * G[0][2] + G[1][0] * G[1][2] + G[2][0] * G[2][2];
  const double cyy = G[0][1] * G[0][1] + G[1][1] * G[1][1] + G[2][1] * G[2][1];
  const double cyz = G[0][1] * G[0][2] + G[1][1] * G[1][2] + G[2][1] * G[2][2];
  const double czz = G[0][2] * G[0][2] + G[1][2] * G[1][2] + G[2][2] * G[2][2];
  const double lam = eig_sym3_max(cxx, cxy, cxz, cyy, cyz, czz);
  double r = log(fmax(sqrt(lam), tiny)) / Tint;
  if (incomp) r = fmax(r, 0.0);
  return r;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_ftle(GridDev<T> g, CF3<T> fm, double Tint, int incomp, double tiny,
                                                 double* __restrict__ ftle, F3<T> comp) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.np) return;
  int i, j, k;
  decode_interior(g, p, i, j, k);
  const int l = gid(g, i, j, k);
  const int ip = (i == g.n[0] && g.per[0]) ? gid(g, 1, j, k) : gid(g, i + 1, j, k);
  const int im = (i == 1 && g.per[0]) ? gid(g, g.n[0], j, k) : gid(g, i - 1, j, k);
  const int jp = (j == g.n[1] && g.per[1]) ? gid(g, i, 1, k) : gid(g, i, j + 1, k);
  const int jm = (j == 1 && g.per[1]) ? gid(g, i, g.n[1], k) : gid(g, i, j - 1, k);
  const int kp = (k == g.n[2] && g.per[2]) ? gid(g, i, j, 1) : gid(g, i, j, k + 1);
  const int km = (k == 1 && g.per[2]) ? gid(g, i, j, g.n[2]) : gid(g, i, j, k - 1);
  double G[3][3];
  if (!g.rect) {
    // grad_sr1_ls on a non-rectilinear grid (lcs_m.f90:131-133,
    // gradient_m.f90:241-295): face neighbours in NBR order
    const int nb[6] = {im, jm, km, ip, jp, kp};
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const T* F = fm.c[c];
      G[c][0] = G[c][1] = G[c][2] = 0.0;
#pragma unroll
      for (int f = 0; f < 6; f++) {
        const double t = (double)F[nb[f]] - (double)F[l];
        const T* w = g.lsw + (long long)f * 3 * g.np;
        G[c][0] = G[c][0] + t * (double)w[p];
        G[c][1] = G[c][1] + t * (double)w[g.np + p];
        G[c][2] = G[c][2] + t * (double)w[2 * g.np + p];
      }
      comp.c[c][p] = F[l];
    }
  } else {
    const double hx = (double)g.gl[0][i + 1] - (double)g.gl[0][i - 1];
    const double hy = (double)g.gl[1][j + 1] - (double)g.gl[1][j - 1];
    const double hz = (double)g.gl[2][k + 1] - (double)g.gl[2][k - 1];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const T* F = fm.c[c];
      G[c][0] = ((double)F[ip] - (double)F[im]) / hx;
      G[c][1] = ((double)F[jp] - (double)F[jm]) / hy;
      G[c][2] = ((double)F[kp] - (double)F[km]) / hz;
      comp.c[c][p] = F[l];
    }
    if (g.n[0] == 1) { G[0][0] = 0; G[1][0] = 0; G[2][0] = 0; }
    if (g.n[1] == 1) { G[0][1] = 0; G[1][1] = 0; G[2][1] = 0; }
    if (g.n[2] == 1) { G[0][2] = 0; G[1][2] = 0; G[2][2] = 0; }
  }
  ftle[p] = ftle_of_grad(G, Tint, incomp, tiny);
}

// compute_ftle with AUX_GRID (lcs_m.f90:95-126): column d of the gradient is
// the difference of the d1 and d0 sets' composed displacements over the
// difference of their node coordinates along d, left at zero where that
// difference is not above small = 100 tiny(1.0) (per node of the main grid;
// the six sets share its extents)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_ftle_aux(GridDev<T> lg, AuxSets<T> aux, double small, double Tint,
                                                     int incomp, double tiny, double* __restrict__ ftle) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= lg.np) return;
  int ix[3];
  decode_interior(lg, p, ix[0], ix[1], ix[2]);
  double G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double den = (double)aux.line[3 + d][ix[d]] - (double)aux.line[d][ix[d]];
    if (fabs(den) > small) {
#pragma unroll
      for (int c = 0; c < 3; c++) G[c][d] = ((double)aux.disp[3 + d][c][p] - (double)aux.disp[d][c][p]) / den;
    }
  }
  ftle[p] = ftle_of_grad(G, Tint, incomp, tiny);
}

// ---------------------------------------------------------------------------
// Tensor B-spline fit (db3ink -> dbtpcf x3, bspline_sub_module.f90:451-533,
// 1801-1856): for every line of the ghosted block along one axis, solve the
// banded collocation system with the LU factors computed on the host
// (dbintk/dbnfac).  Forward and backward substitution in the order of
// dbnslv (bspline_sub_module.f90:2202-2259), in place.
// ---------------------------------------------------------------------------
// One line of the banded solve through the accessors ld(r) and st(r, v) (row
// r of the line, held in the element type of the array, computed in T).  The rows a recurrence step needs are independent of the recurrence
// itself, so they are loaded kFitPf rows ahead (double-buffered) and the
// dependent chain runs on registers only; the arithmetic and its order are
// those of the unpipelined loop.
constexpr int kFitPf = 8;
// Recurrence steps of A x = b with A = L U in band storage (element (r, c) at
// (r - c + K - 1) + (2K - 1) c): forward with the unit lower factor, backward
// with the upper one.  TR solves A^T x = b instead (the adjoint of the fit):
// forward with U^T (dividing by the diagonal), backward with the unit L^T.
template <typename T, int K, bool TR>
__device__ __forceinline__ T fwd_elim(int r, T v, const T* win, const T* __restrict__ W) {
  constexpr int M = K - 1, NB = 2 * K - 1;
#pragma unroll
  for (int jj = M; jj >= 1; jj--)
    if (r - jj >= 0) v = v - win[jj - 1] * (TR ? W[(M - jj) + NB * r] : W[(jj + M) + NB * (r - jj)]);
  if (TR) v = v / W[M + NB * r];
  return v;
}
template <typename T, int K, bool TR>
__device__ __forceinline__ T bwd_elim(int r, int n, T v, const T* win, const T* __restrict__ W) {
  constexpr int M = K - 1, NB = 2 * K - 1;
#pragma unroll
  for (int jj = M; jj >= 1; jj--)
    if (r + jj <= n - 1) v = v - win[jj - 1] * (TR ? W[(jj + M) + NB * r] : W[(M - jj) + NB * (r + jj)]);
  if (!TR) v = v / W[M + NB * r];
  return v;
}

// The prefetched rows stay in the element type E until used, so that their
// loads are still in flight while the recurrence runs.
template <typename T, int K, bool TR, typename E, class Ld, class St>
__device__ __forceinline__ void band_solve(Ld ld, St st, const T* __restrict__ W, int n) {
  constexpr int M = K - 1, PF = kFitPf;
  T win[M > 0 ? M : 1];  // forward: win[0] = b(r-1), ...; backward: win[0] = b(r+1), ...
#pragma unroll
  for (int m = 0; m < M; m++) win[m] = T(0);
  E cur[PF], nxt[PF];
#pragma unroll
  for (int u = 0; u < PF; u++) cur[u] = (u < n) ? ld(u) : E(0);
  for (int r0 = 0; r0 < n; r0 += PF) {
#pragma unroll
    for (int u = 0; u < PF; u++) nxt[u] = (r0 + PF + u < n) ? ld(r0 + PF + u) : E(0);
#pragma unroll
    for (int u = 0; u < PF; u++) {
      const int r = r0 + u;
      if (r < n) {
        const T v = fwd_elim<T, K, TR>(r, T(cur[u]), win, W);
#pragma unroll
        for (int m = M - 1; m >= 1; m--) win[m] = win[m - 1];
        win[0] = v;
        st(r, v);
      }
    }
#pragma unroll
    for (int u = 0; u < PF; u++) cur[u] = nxt[u];
  }
#pragma unroll
  for (int m = 0; m < M; m++) win[m] = T(0);
#pragma unroll
  for (int u = 0; u < PF; u++) cur[u] = (n - 1 - u >= 0) ? ld(n - 1 - u) : E(0);
  for (int r0 = n - 1; r0 >= 0; r0 -= PF) {
#pragma unroll
    for (int u = 0; u < PF; u++) nxt[u] = (r0 - PF - u >= 0) ? ld(r0 - PF - u) : E(0);
#pragma unroll
    for (int u = 0; u < PF; u++) {
      const int r = r0 - u;
      if (r >= 0) {
        const T v = bwd_elim<T, K, TR>(r, n, T(cur[u]), win, W);
#pragma unroll
        for (int m = M - 1; m >= 1; m--) win[m] = win[m - 1];
        win[0] = v;
        st(r, v);
      }
    }
#pragma unroll
    for (int u = 0; u < PF; u++) cur[u] = nxt[u];
  }
}

template <typename T, int K, bool TR, typename E>
__global__ void __launch_bounds__(kBlock) k_fit_axis(F3<E> a, const T* __restrict__ W, int n, int axis, int N0,
                                                     int N1, int nlines) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  const int comp = blockIdx.y;
  if (id >= nlines) return;
  long long start, stride;
  if (axis == 0) {
    start = (long long)N0 * id; stride = 1;
  } else if (axis == 1) {
    const int i = id % N0, kk = id / N0;
    start = i + (long long)N0 * N1 * kk; stride = N0;
  } else {
    start = id; stride = (long long)N0 * N1;
  }
  E* b = a.c[comp] + start;
  band_solve<T, K, TR, E>([&](int r) -> E { return b[r * stride]; }, [&](int r, T v) { b[r * stride] = E(v); }, W,
                          n);
}

// x-axis pass of the fit: the lines are contiguous rows.  One wavefront
// (block of 64) owns 64 rows and streams them in chunks of kXC columns: each
// chunk is loaded with row-contiguous accesses, transposed through a small
// LDS tile so that lane r runs the recurrence of row r, and stored back the
// same way; the next chunk's loads are in flight during the recurrence.  The
// forward sweep leaves its result in place and the backward sweep reads it
// back, as the strided passes do.  Same arithmetic, in the same order, as
// band_solve.  With src, the forward sweep reads the data from src instead
// of a (a copy of the data into a fused with the fit); E is the element type
// of the arrays, the arithmetic is in T.
constexpr int kXC = 32;
template <typename T, int K, bool TR, typename E, typename ES>
__global__ void __launch_bounds__(64) k_fit_x(F3<E> a, CF3<ES> src, const T* __restrict__ W, int n, int nlines) {
  __shared__ T tile[64][kXC + 1];
  constexpr int M = K - 1;
  const int lane = threadIdx.x;
  const long long line0 = (long long)blockIdx.x * 64;
  const int nl = (int)min((long long)64, nlines - line0);
  E* base = a.c[blockIdx.y] + line0 * n;
  const int nch = (n + kXC - 1) / kXC;
  // raw elements (when source and result share a type): converted when
  // written to LDS, so that the next chunk's loads stay in flight during
  // the recurrence
  using B = typename std::conditional<std::is_same<E, ES>::value, E, T>::type;
  B buf[kXC];
  // element it*64+lane of a chunk: row (it*64+lane)/kXC, column (it*64+lane)%kXC
  auto load = [&](const auto* in, int ch) {
#pragma unroll
    for (int it = 0; it < kXC; it++) {
      const int e = it * 64 + lane, row = e / kXC, col = ch * kXC + e % kXC;
      buf[it] = (row < nl && col < n) ? B(in[(long long)row * n + col]) : B(0);
    }
  };
  auto to_lds = [&]() {
#pragma unroll
    for (int it = 0; it < kXC; it++) {
      const int e = it * 64 + lane;
      tile[e / kXC][e % kXC] = T(buf[it]);
    }
  };
  auto store = [&](int ch) {
#pragma unroll
    for (int it = 0; it < kXC; it++) {
      const int e = it * 64 + lane, row = e / kXC, col = ch * kXC + e % kXC;
      if (row < nl && col < n) base[(long long)row * n + col] = E(tile[e / kXC][e % kXC]);
    }
  };
  T win[M > 0 ? M : 1];
  // forward substitution (unit lower factor)
#pragma unroll
  for (int m = 0; m < M; m++) win[m] = T(0);
  const ES* from = src.c[0] ? src.c[blockIdx.y] + line0 * n : nullptr;
  auto load_fwd = [&](int ch) {
    if constexpr (std::is_same<E, ES>::value) load(from ? from : (const ES*)base, ch);
    else load(from, ch);
  };
  load_fwd(0);
  for (int ch = 0; ch < nch; ch++) {
    to_lds();
    if (ch + 1 < nch) load_fwd(ch + 1);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kXC; u++) {
      const int r = ch * kXC + u;
      if (r < n) {
        const T v = fwd_elim<T, K, TR>(r, tile[lane][u], win, W);
#pragma unroll
        for (int m = M - 1; m >= 1; m--) win[m] = win[m - 1];
        win[0] = v;
        tile[lane][u] = v;
      }
    }
    __syncthreads();
    store(ch);
    __syncthreads();
  }
  // backward substitution (upper factor), last chunk first
#pragma unroll
  for (int m = 0; m < M; m++) win[m] = T(0);
  load((const E*)base, nch - 1);
  for (int ch = nch - 1; ch >= 0; ch--) {
    to_lds();
    if (ch > 0) load((const E*)base, ch - 1);
    __syncthreads();
#pragma unroll
    for (int u = kXC - 1; u >= 0; u--) {
      const int r = ch * kXC + u;
      if (r < n) {
        const T v = bwd_elim<T, K, TR>(r, n, tile[lane][u], win, W);
#pragma unroll
        for (int m = M - 1; m >= 1; m--) win[m] = win[m - 1];
        win[0] = v;
        tile[lane][u] = v;
      }
    }
    __syncthreads();
    store(ch);
    __syncthreads();
  }
}

// interior (or per-origin) values to a user-typed array, Fortran order
template <typename T, typename U>
__global__ void __launch_bounds__(kBlock) k_to_user(GridDev<T> g, CF3<T> s, int ghosted, U* a, U* b, U* c) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.np) return;
  int l = p;
  if (ghosted) {
    int i, j, k;
    decode_interior(g, p, i, j, k);
    l = gid(g, i, j, k);
  }
  a[p] = (U)s.c[0][l];
  b[p] = (U)s.c[1][l];
  c[p] = (U)s.c[2][l];
}

// ===========================================================================
// launchers
// ===========================================================================
static inline dim3 grid_for(long long n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

template <typename U, typename T>
void launch_ingest(const GridDev<T>& g, F3<T> dst, const U* u, const U* v, const U* w, hipStream_t s) {
  hipLaunchKernelGGL((k_ingest<U, T>), grid_for(g.np), dim3(kBlock), 0, s, g, dst, u, v, w);
}

template <typename T>
void launch_vel_ghost_impl(const GridDev<T>& g, F3<T> u, CF3<T> uold, DevStatus* st, hipStream_t s) {
  long long ng = 2LL * g.N[0] * g.N[1] + (2LL * g.N[0] + 2LL * g.n[1]) * g.n[2];
  hipLaunchKernelGGL((k_vel_ghost<T>), grid_for(ng), dim3(kBlock), 0, s, g, u, uold, st);
}

template <typename T>
void launch_cfl(const GridDev<T>& g, F3<T> u, DevStatus* st, hipStream_t s) {
  const long long nb = std::min<long long>((g.np + kBlock - 1) / kBlock, 2048);
  hipLaunchKernelGGL((k_cfl<T>), dim3((unsigned)nb), dim3(kBlock), 0, s, g, u, st);
}

template <typename T, typename E>
void launch_fm_ghost_impl(const GridDev<T>& g, F3<E> fm, CF3<E> fmold, DevStatus* st, hipStream_t s) {
  long long ng = 2LL * g.N[0] * g.N[1] + (2LL * g.N[0] + 2LL * g.n[1]) * g.n[2];
  hipLaunchKernelGGL((k_fm_ghost<T, E>), grid_for(ng), dim3(kBlock), 0, s, g, fm, fmold, st);
}

template <typename T>
void launch_field_from_origin(const GridDev<T>& g, CF3<T> src, F3<T> dst, hipStream_t s) {
  hipLaunchKernelGGL((k_field_from_origin<T>), grid_for(g.nt), dim3(kBlock), 0, s, g, src, dst);
}

template <typename T, bool TR, typename E, typename ES>
static void fit_pass(const GridDev<T>& g, F3<E> coef, CF3<ES> src, int ncomp, int axis, hipStream_t s) {
  const int n = g.N[axis];
  const int nlines = (int)(g.nt / n);
  if (axis == 0) {
    dim3 gx((nlines + 63) / 64, ncomp);
    switch (g.kord[0]) {
      case 2: hipLaunchKernelGGL((k_fit_x<T, 2, TR, E, ES>), gx, dim3(64), 0, s, coef, src, g.band[0], n, nlines); break;
      case 3: hipLaunchKernelGGL((k_fit_x<T, 3, TR, E, ES>), gx, dim3(64), 0, s, coef, src, g.band[0], n, nlines); break;
      default: hipLaunchKernelGGL((k_fit_x<T, 4, TR, E, ES>), gx, dim3(64), 0, s, coef, src, g.band[0], n, nlines); break;
    }
    return;
  }
  dim3 grid((nlines + kBlock - 1) / kBlock, ncomp);
  switch (g.kord[axis]) {
    case 2: hipLaunchKernelGGL((k_fit_axis<T, 2, TR, E>), grid, dim3(kBlock), 0, s, coef, g.band[axis], n, axis, g.N[0], g.N[1], nlines); break;
    case 3: hipLaunchKernelGGL((k_fit_axis<T, 3, TR, E>), grid, dim3(kBlock), 0, s, coef, g.band[axis], n, axis, g.N[0], g.N[1], nlines); break;
    default: hipLaunchKernelGGL((k_fit_axis<T, 4, TR, E>), grid, dim3(kBlock), 0, s, coef, g.band[axis], n, axis, g.N[0], g.N[1], nlines); break;
  }
}

// coef = the coefficients of the data in src (or in coef itself, in place,
// when src is null); the element type E of both is T or, under FP64
// compute, FP32 (the solve runs in T)
template <typename T, typename E, typename ES>
void launch_spline_fit(const GridDev<T>& g, F3<E> coef, int ncomp, hipStream_t s, const ES* const* src) {
  const CF3<ES> in{{src ? src[0] : nullptr, src ? src[1] : nullptr, src ? src[2] : nullptr}};
  fit_pass<T, false, E, ES>(g, coef, in, ncomp, 0, s);
  const CF3<E> none{{nullptr, nullptr, nullptr}};
  for (int axis = 1; axis < 3; axis++) fit_pass<T, false, E, E>(g, coef, none, ncomp, axis, s);
}

// transpose of launch_spline_fit (three components, in place): A^T solves in
// the reverse order of the passes
template <typename T>
void launch_spline_fit_T(const GridDev<T>& g, F3<T> a, hipStream_t s) {
  for (int axis = 2; axis >= 0; axis--) fit_pass<T, true, T, T>(g, a, CF3<T>{{nullptr, nullptr, nullptr}}, 3, axis, s);
}

// dispatch over interpolator and spline orders (KX = orders of x and y, KZ of z)
template <typename T, class F>
bool dispatch_spline(const GridDev<T>& g, F&& f) {
  const int kx = g.kord[0], kz = g.kord[2];
  if (g.kord[1] != kx) return false;
  if (kx == 4 && kz == 4) { f(std::integral_constant<int, 4>{}, std::integral_constant<int, 4>{}); return true; }
  if (kx == 4 && kz == 2) { f(std::integral_constant<int, 4>{}, std::integral_constant<int, 2>{}); return true; }
  if (kx == 3 && kz == 3) { f(std::integral_constant<int, 3>{}, std::integral_constant<int, 3>{}); return true; }
  if (kx == 3 && kz == 2) { f(std::integral_constant<int, 3>{}, std::integral_constant<int, 2>{}); return true; }
  return false;
}

// zero fill: 16-byte words, then the tail bytes
__global__ void __launch_bounds__(kBlock) k_zero(unsigned char* p, size_t bytes) {
  const size_t nw = bytes / 16;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < nw; w += stride)
    reinterpret_cast<uint4*>(p)[w] = make_uint4(0u, 0u, 0u, 0u);
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < bytes - nw * 16) p[nw * 16 + t] = 0;
}

hipError_t launch_zero(void* p, size_t bytes, hipStream_t s) {
  if (!bytes) return hipSuccess;
  if ((reinterpret_cast<uintptr_t>(p) & 15) != 0) return hipErrorInvalidValue;
  const size_t nw = bytes / 16;
  const unsigned nb = (unsigned)std::max<size_t>(1, std::min<size_t>((nw + kBlock - 1) / kBlock, 4096));
  hipLaunchKernelGGL(k_zero, dim3(nb), dim3(kBlock), 0, s, (unsigned char*)p, bytes);
  return hipGetLastError();
}

template <typename T, typename A>
void launch_blend(const GridDev<T>& g, void* const out[3], int storage, CF3<A> a, CF3<A> b, T ct, T inv,
                  DevStatus* st, hipStream_t s) {
  const long long nb = std::min<long long>((g.nt + kBlock - 1) / kBlock, 8192);
  auto go = [&](auto tag) {
    using S = decltype(tag);
    F3<S> o{{(S*)out[0], (S*)out[1], (S*)out[2]}};
    hipLaunchKernelGGL((k_blend<T, S, A>), dim3((unsigned)nb), dim3(kBlock), 0, s, o, a, b, ct, inv, g.nt, st);
  };
  if constexpr (std::is_same<T, float>::value) {
    if (storage == kStoreF16) return go(half4_t{});
    if (storage == kStoreBF16) return go(bf164_t{});
  } else {
    if (storage == kStoreF32) return go(float{});
  }
  go(T{});
}

// The blends of several weights from one read of the two fields: each
// output is k_blend's expression in k_blend's order, so its values are the
// same; the fields are read once per batch instead of once per blend.
template <typename T, typename S, typename A>
__global__ void __launch_bounds__(kBlock) k_blend_batch(BlendBatch<T> bb, CF3<A> a, CF3<A> b, long long nt,
                                                        DevStatus* st) {
#pragma clang fp contract(off)
  bool range = false;
  for (long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x; l < nt; l += (long long)gridDim.x * blockDim.x) {
    T va[3], vb[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      va[c] = T(a.c[c][l]);
      vb[c] = T(b.c[c][l]);
    }
#pragma unroll
    for (int q = 0; q < kBlendBatch; q++) {
      if (q < bb.n) {
        const T ct = bb.ct[q], w0 = T(1) - ct;
        if constexpr (Packed<S>::value) {
          using E = typename Packed<S>::E;
          S r;
#pragma unroll
          for (int c = 0; c < 3; c++) {
            const T x = (w0 * va[c] + ct * vb[c]) * bb.inv;
            if constexpr (std::is_same<E, half_t>::value) range |= fabs(x) > T(65504);
            r.e[c] = narrow<E>(x);
          }
          r.e[3] = E{0};
          ((S*)bb.out[q][0])[l] = r;
        } else {
#pragma unroll
          for (int c = 0; c < 3; c++) ((S*)bb.out[q][c])[l] = narrow<S>(w0 * va[c] + ct * vb[c]);
        }
      }
    }
  }
  if (range) atomicOr(&st->err, (int)kErrRange);
}

template <typename T, typename A>
void launch_blend_batch(const GridDev<T>& g, const BlendBatch<T>& bb, int storage, CF3<A> a, CF3<A> b,
                        DevStatus* st, hipStream_t s) {
  const long long nb = std::min<long long>((g.nt + kBlock - 1) / kBlock, 8192);
  auto go = [&](auto tag) {
    using S = decltype(tag);
    hipLaunchKernelGGL((k_blend_batch<T, S, A>), dim3((unsigned)nb), dim3(kBlock), 0, s, bb, a, b, g.nt, st);
  };
  if constexpr (std::is_same<T, float>::value) {
    if (storage == kStoreF16) return go(half4_t{});
    if (storage == kStoreBF16) return go(bf164_t{});
  } else {
    if (storage == kStoreF32) return go(float{});
  }
  go(T{});
}

// storage dispatch: the element type of the stage velocity fields
template <typename T, class F>
void with_storage(int storage, F&& f) {
  if constexpr (std::is_same<T, float>::value) {
    if (storage == kStoreF16) return f(half4_t{});
    if (storage == kStoreBF16) return f(bf164_t{});
  } else {
    if (storage == kStoreF32) return f(float{});
  }
  f(T{});
}

// A non-rectilinear grid interpolates with TSE, or TSE_LIMIT for every other
// choice, with least-squares gradients (lp_tracking_m.f90:319-328): f gets
// the interpolator; false on a rectilinear grid.
template <typename T, class F>
bool dispatch_nonrect(const GridDev<T>& g, int interp, F&& f) {
  if (g.rect) return false;
  if (interp == kTSE) f(IntTSELS<T, false>{});
  else f(IntTSELS<T, true>{});
  return true;
}

// with_storage and dispatch_nonrect together: f(storage tag, interpolator)
// on a non-rectilinear grid, false (nothing run) on a rectilinear one
template <typename T, class F>
bool with_storage_nonrect(const GridDev<T>& g, int interp, int storage, F&& f) {
  if (g.rect) return false;
  with_storage<T>(storage, [&](auto tag) { dispatch_nonrect(g, interp, [&](auto in) { f(tag, in); }); });
  return true;
}

template <typename T, typename FM>
void launch_sl_step(const GridDev<T>& g, int interp, const StageParams<T>& sp, CF3<FM> fmold, F3<FM> fmnew,
                    DevStatus* st, hipStream_t s, int tiled) {
  const dim3 grid = grid_for(g.np), blk(kBlock);
  if (with_storage_nonrect(g, interp, sp.storage, [&](auto tag, auto in) {
        using S = decltype(tag);
        using I = decltype(in);
        hipLaunchKernelGGL((k_sl_step<T, S, I, FM>), grid, blk, 0, s, g, sp, fmold, fmnew, st);
      }))
    return;
  // cubic step on a 3-D grid: LDS bricks with fields in the compute type
  // (native stage storage), or FP32 stage fields and flowmap under FP64
  // compute
  if constexpr (std::is_same<FM, float>::value || std::is_same<FM, T>::value) {
    const int want = std::is_same<FM, T>::value ? kStoreNative : kStoreF32;
    if (tiled && interp == kCubic && sp.storage == want && g.flat < 0 && g.kord[0] == 4 && g.kord[1] == 4 &&
        g.kord[2] == 4) {
      constexpr int BX = kBrickX, BY = kBrickY, BZ = kBrickZ;
      const int nbx = (g.n[0] + BX - 1) / BX, nby = (g.n[1] + BY - 1) / BY, nbz = (g.n[2] + BZ - 1) / BZ;
      const dim3 tgrid((unsigned)(nbx * nby * nbz)), tblk(BX * BY * BZ);
      auto go = [&](auto K) {
        hipLaunchKernelGGL((k_sl_brick<T, FM, BX, BY, BZ, decltype(K)::value>), tgrid, tblk, 0, s, g, sp, fmold,
                           fmnew, st, nbx, nby);
      };
      // the schemes with stages to gather; Euler and trapezoidal read the
      // node only and take the gather kernel
      if (sp.integ == kRK4) { go(std::integral_constant<int, kRK4>{}); return; }
      if (sp.integ == kRK3) { go(std::integral_constant<int, kRK3>{}); return; }
      if (sp.integ == kRK2) { go(std::integral_constant<int, kRK2>{}); return; }
    }
  }
  with_storage<T>(sp.storage, [&](auto tag) {
    using S = decltype(tag);
    if (interp == kLinear) {
      hipLaunchKernelGGL((k_sl_step<T, S, IntLinear<T>, FM>), grid, blk, 0, s, g, sp, fmold, fmnew, st);
    } else if (interp == kNearest) {
      hipLaunchKernelGGL((k_sl_step<T, S, IntNearest<T>, FM>), grid, blk, 0, s, g, sp, fmold, fmnew, st);
    } else if (interp == kTSE) {
      hipLaunchKernelGGL((k_sl_step<T, S, IntTSE<T, false>, FM>), grid, blk, 0, s, g, sp, fmold, fmnew, st);
    } else if (interp == kTSELimit) {
      hipLaunchKernelGGL((k_sl_step<T, S, IntTSE<T, true>, FM>), grid, blk, 0, s, g, sp, fmold, fmnew, st);
    } else {
      dispatch_spline(g, [&](auto KX, auto KZ) {
        constexpr int kx = decltype(KX)::value, kz = decltype(KZ)::value;
        hipLaunchKernelGGL((k_sl_step<T, S, IntSpline<T, kx, kz>, FM>), grid, blk, 0, s, g, sp, fmold, fmnew, st);
      });
    }
  });
}

template <typename T>
void launch_fwd_step(const GridDev<T>& g, int interp, const StageParams<T>& sp, ParticleSet<T> ps, int np, int comp,
                     DevStatus* st, hipStream_t s) {
  const dim3 grid = grid_for(np), blk(kBlock);
  with_storage<T>(sp.storage, [&](auto tag) {
    using S = decltype(tag);
    auto go = [&](auto C) {
      constexpr bool cm = decltype(C)::value;
      if (dispatch_nonrect(g, interp, [&](auto in) {
            using I = decltype(in);
            hipLaunchKernelGGL((k_fwd_step<T, S, I, cm>), grid, blk, 0, s, g, sp, ps, np, st);
          })) {
      } else if (interp == kLinear) {
        hipLaunchKernelGGL((k_fwd_step<T, S, IntLinear<T>, cm>), grid, blk, 0, s, g, sp, ps, np, st);
      } else if (interp == kNearest) {
        hipLaunchKernelGGL((k_fwd_step<T, S, IntNearest<T>, cm>), grid, blk, 0, s, g, sp, ps, np, st);
      } else if (interp == kTSE) {
        hipLaunchKernelGGL((k_fwd_step<T, S, IntTSE<T, false>, cm>), grid, blk, 0, s, g, sp, ps, np, st);
      } else if (interp == kTSELimit) {
        hipLaunchKernelGGL((k_fwd_step<T, S, IntTSE<T, true>, cm>), grid, blk, 0, s, g, sp, ps, np, st);
      } else {
        dispatch_spline(g, [&](auto KX, auto KZ) {
          constexpr int kx = decltype(KX)::value, kz = decltype(KZ)::value;
          hipLaunchKernelGGL((k_fwd_step<T, S, IntSpline<T, kx, kz>, cm>), grid, blk, 0, s, g, sp, ps, np, st);
        });
      }
    };
    if (comp) go(std::true_type{});
    else go(std::false_type{});
  });
}

template <typename T>
void launch_fwd_reset(const GridDev<T>& g, ParticleSet<T> ps, hipStream_t s) {
  hipLaunchKernelGGL((k_fwd_reset<T>), grid_for(g.np), dim3(kBlock), 0, s, g, ps);
}

template <typename T>
void launch_track_nodes(const GridDev<T>& g, const GridDev<T>& lg, int16_t* nodes, DevStatus* st, hipStream_t s) {
  hipLaunchKernelGGL((k_track_nodes<T>), grid_for(lg.np), dim3(kBlock), 0, s, g, lg, nodes, st);
}

template <typename T>
void launch_fwd_place(const GridDev<T>& g, const GridDev<T>& lg, ParticleSet<T> ps, DevStatus* st, hipStream_t s) {
  hipLaunchKernelGGL((k_fwd_place<T>), grid_for(lg.np), dim3(kBlock), 0, s, g, lg, ps, st);
}

template <typename T, typename FM>
void launch_sl_step2(const GridDev<T>& g, const GridDev<T>& lg, const int16_t* nodes, int fresh, int interp,
                     const StageParams<T>& sp, CF3<FM> fmold, F3<FM> fmnew, DevStatus* st, hipStream_t s) {
  const dim3 grid = grid_for(lg.np), blk(kBlock);
  with_storage<T>(sp.storage, [&](auto tag) {
    using S = decltype(tag);
    if (interp == kLinear) {
      hipLaunchKernelGGL((k_sl_step2<T, S, IntLinear<T>, FM>), grid, blk, 0, s, g, lg, nodes, fresh, sp, fmold, fmnew,
                         st);
    } else if (interp == kNearest) {
      hipLaunchKernelGGL((k_sl_step2<T, S, IntNearest<T>, FM>), grid, blk, 0, s, g, lg, nodes, fresh, sp, fmold, fmnew,
                         st);
    } else if (interp == kTSE) {
      hipLaunchKernelGGL((k_sl_step2<T, S, IntTSE<T, false>, FM>), grid, blk, 0, s, g, lg, nodes, fresh, sp, fmold,
                         fmnew, st);
    } else if (interp == kTSELimit) {
      hipLaunchKernelGGL((k_sl_step2<T, S, IntTSE<T, true>, FM>), grid, blk, 0, s, g, lg, nodes, fresh, sp, fmold,
                         fmnew, st);
    } else {
      dispatch_spline(g, [&](auto KX, auto KZ) {
        constexpr int kx = decltype(KX)::value, kz = decltype(KZ)::value;
        hipLaunchKernelGGL((k_sl_step2<T, S, IntSpline<T, kx, kz>, FM>), grid, blk, 0, s, g, lg, nodes, fresh, sp,
                           fmold, fmnew, st);
      });
    }
  });
}

template <typename T>
void launch_expand_periodic(const GridDev<T>& g, const T* interior, F3<T> dst, hipStream_t s) {
  hipLaunchKernelGGL((k_expand_periodic<T>), grid_for(g.nt), dim3(kBlock), 0, s, g, interior, dst);
}

template <typename T>
void launch_field_at_nodes(const GridDev<T>& g, const GridDev<T>& lg, const int16_t* nodes, int interp, CF3<T> f,
                           T* out, hipStream_t s) {
  const dim3 grid = grid_for(lg.np), blk(kBlock);
  if (interp == kLinear) {
    hipLaunchKernelGGL((k_field_at_nodes<T, IntLinear<T>>), grid, blk, 0, s, g, lg, nodes, f, out);
  } else if (interp == kNearest) {
    hipLaunchKernelGGL((k_field_at_nodes<T, IntNearest<T>>), grid, blk, 0, s, g, lg, nodes, f, out);
  } else if (interp == kTSE) {
    hipLaunchKernelGGL((k_field_at_nodes<T, IntTSE<T, false>>), grid, blk, 0, s, g, lg, nodes, f, out);
  } else if (interp == kTSELimit) {
    hipLaunchKernelGGL((k_field_at_nodes<T, IntTSE<T, true>>), grid, blk, 0, s, g, lg, nodes, f, out);
  } else {
    dispatch_spline(g, [&](auto KX, auto KZ) {
      constexpr int kx = decltype(KX)::value, kz = decltype(KZ)::value;
      hipLaunchKernelGGL((k_field_at_nodes<T, IntSpline<T, kx, kz>>), grid, blk, 0, s, g, lg, nodes, f, out);
    });
  }
}

template <typename T>
void launch_ftle_aux(const GridDev<T>& lg, const AuxSets<T>& aux, double small, double Tint, int incomp,
                     double tiny, double* ftle, hipStream_t s) {
  hipLaunchKernelGGL((k_ftle_aux<T>), grid_for(lg.np), dim3(kBlock), 0, s, lg, aux, small, Tint, incomp, tiny,
                     ftle);
}

template <typename T>
void launch_compose(const GridDev<T>& g, int interp, const T* slots, long long stride, const int* order, int nslot,
                    F3<T> dxout, DevStatus* st, hipStream_t s) {
  const dim3 grid = grid_for(g.np), blk(kBlock);
  if (dispatch_nonrect(g, interp, [&](auto in) {
        using I = decltype(in);
        hipLaunchKernelGGL((k_compose<T, I>), grid, blk, 0, s, g, slots, stride, order, nslot, dxout, st);
      })) {
  } else if (interp == kLinear) {
    hipLaunchKernelGGL((k_compose<T, IntLinear<T>>), grid, blk, 0, s, g, slots, stride, order, nslot, dxout, st);
  } else if (interp == kNearest) {
    hipLaunchKernelGGL((k_compose<T, IntNearest<T>>), grid, blk, 0, s, g, slots, stride, order, nslot, dxout, st);
  } else if (interp == kTSE) {
    hipLaunchKernelGGL((k_compose<T, IntTSE<T, false>>), grid, blk, 0, s, g, slots, stride, order, nslot, dxout,
                       st);
  } else if (interp == kTSELimit) {
    hipLaunchKernelGGL((k_compose<T, IntTSE<T, true>>), grid, blk, 0, s, g, slots, stride, order, nslot, dxout,
                       st);
  } else {
    dispatch_spline(g, [&](auto KX, auto KZ) {
      constexpr int kx = decltype(KX)::value, kz = decltype(KZ)::value;
      hipLaunchKernelGGL((k_compose<T, IntSpline<T, kx, kz>>), grid, blk, 0, s, g, slots, stride, order, nslot,
                         dxout, st);
    });
  }
}

// The slot kernel for the B-spline interpolants on a rectilinear grid: the
// workgroup's particles, sorted by brick, need the coefficients of a small
// box of nodes, which the workgroup loads into LDS once (coalesced along x)
// and every particle then sums its stencil from there instead of issuing
// its K^2 KZ scattered row loads through the texture path.  The box is the
// union of the particles' stencils; a workgroup whose box exceeds the LDS
// capacity (particles spread over many bricks) gathers from global memory.
// The sums run in spline_sum's order: results equal k_compose_slot's.
template <typename T>
constexpr int comp_tile_cap() {
  return sizeof(T) == 4 ? 2048 : 1024;
}

template <typename T, int KX, int KZ>
__global__ void __launch_bounds__(kBlock) k_compose_slot_tiled(GridDev<T> g, CF3<T> f, const CompRec<T>* __restrict__ in,
                                                               const unsigned* __restrict__ perm,
                                                               CompRec<T>* __restrict__ out, unsigned* __restrict__ keys,
                                                               unsigned* __restrict__ vals, int first, int last,
                                                               F3<T> dxout, DevStatus* st) {
  constexpr int CAP = comp_tile_cap<T>();
  __shared__ T tile[3 * CAP];
  __shared__ int box[6];
  const int t = xcd_thread_index();
  const bool active = t < g.np;
  if (threadIdx.x < 3) {
    box[threadIdx.x] = INT_MAX;
    box[3 + threadIdx.x] = INT_MIN;
  }
  __syncthreads();
  T q[3] = {T(0), T(0), T(0)}, dxx[3] = {T(0), T(0), T(0)};
  int nd[3] = {1, 1, 1}, org = 0;
  Split<T> sq{{T(0), T(0), T(0)}, {T(0), T(0), T(0)}};
  unsigned long long oob = 0;
  int err = 0;
  SplineW<T, KX, KZ> w;
  int r = 3;  // 3: no particle
  if (active) {
    if (first) {
      int i, j, k;
      decode_comp_first(g, t, i, j, k);
      nd[0] = i; nd[1] = j; nd[2] = k;
      node_pos<false>(g, i, j, k, q);
#pragma unroll
      for (int d = 0; d < 3; d++) sq.qb[d] = q[d];
      org = pid(g, i, j, k);
    } else {
      const CompRec<T> rr = in[perm ? perm[t] : (unsigned)t];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        sq.qb[d] = rr.qb[d];
        sq.off[d] = rr.off[d];
        q[d] = rr.qb[d] + rr.off[d];
        dxx[d] = rr.dx[d];
        nd[d] = rr.nd[d];
      }
      org = rr.org;
    }
    compose_bc<false>(g, q, sq, nd, err);
    r = spline_setup(g, q, nd, w, &sq);
  }
  // the box of coefficient indices the workgroup's stencils start in
  int lo[3], hi[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    lo[d] = r == 0 ? w.b3[d] : INT_MAX;
    hi[d] = r == 0 ? w.b3[d] : INT_MIN;
  }
  // wave reductions (every lane of the wave takes part)
#pragma unroll
  for (int d = 0; d < 3; d++) {
    lo[d] = __reduce_min_sync(~0ull, lo[d]);
    hi[d] = __reduce_max_sync(~0ull, hi[d]);
  }
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
      atomicMin(&box[d], lo[d]);
      atomicMax(&box[3 + d], hi[d]);
    }
  }
  __syncthreads();
  const int b0 = box[0], b1 = box[1], b2 = box[2];
  const int ex = box[3] - b0 + KX, ey = box[4] - b1 + KX, ez = box[5] - b2 + KZ;
  const bool staged = box[3] >= b0 && (long long)ex * ey * ez <= CAP;
  if (staged) {
    const int vol = ex * ey * ez, pl = ex * ey;
    const float rex = 1.0f / (float)ex, rpl = 1.0f / (float)pl;
    for (int e = threadIdx.x; e < vol; e += blockDim.x) {
      const int z = small_div(e, pl, rpl), r = e - z * pl, y = small_div(r, ex, rex), x = r - y * ex;
      const int gi = (b0 + x) + g.N[0] * ((b1 + y) + g.N[1] * (b2 + z));
#pragma unroll
      for (int c = 0; c < 3; c++) tile[c * CAP + e] = f.c[c][gi];
    }
  }
  __syncthreads();
  if (!active) return;
  T up[3];
  if (r) {
    if (r == 1) oob++;
    up[0] = up[1] = up[2] = T(0);
  } else if (staged) {
    const int ox = w.b3[0] - b0, oy = w.b3[1] - b1, oz = w.b3[2] - b2;
    const int pl = ex * ey, o0 = (oz * ey + oy) * ex + ox;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const T* tc = tile + c * CAP + o0;
      T acc = T(0);
#pragma unroll
      for (int cz = 0; cz < KZ; cz++) {
        T accy = T(0);
#pragma unroll
        for (int by = 0; by < KX; by++) {
          const T* row = tc + cz * pl + by * ex;
          T accx = T(0);
#pragma unroll
          for (int a = 0; a < KX; a++) accx += w.wx[a] * row[a];
          accy += w.wy[by] * accx;
        }
        acc += w.wz[cz] * accy;
      }
      up[c] = acc;
    }
  } else {
#pragma unroll
    for (int c = 0; c < 3; c++) up[c] = spline_sum(g, w, f.c[c]);
  }
  compose_after<false>(g, up, q, sq, nd, dxx, err);
  if (last) {
#pragma unroll
    for (int d = 0; d < 3; d++) dxout.c[d][org] = dxx[d];
  } else {
    CompRec<T> rr;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      rr.qb[d] = sq.qb[d];
      rr.off[d] = sq.off[d];
      rr.dx[d] = dxx[d];
      rr.nd[d] = (int16_t)nd[d];
    }
    rr.pad = 0;
    rr.org = org;
    out[t] = rr;
    if (keys) {
      keys[t] = comp_key(g, nd);
      vals[t] = (unsigned)t;
    }
  }
  flush_status(st, err, oob);
}

template <typename T>
void launch_compose_slot(const GridDev<T>& g, int interp, CF3<T> f, const CompRec<T>* in, const unsigned* perm,
                         CompRec<T>* out, unsigned* keys, unsigned* vals, int first, int last, F3<T> dxout,
                         DevStatus* st, hipStream_t s) {
  const dim3 grid = grid_for(g.np), blk(kBlock);
  auto go = [&](auto in_) {
    using I = decltype(in_);
    hipLaunchKernelGGL((k_compose_slot<T, I>), grid, blk, 0, s, g, f, in, perm, out, keys, vals, first, last, dxout,
                       st);
  };
  if (dispatch_nonrect(g, interp, go)) {
  } else if (interp == kLinear) {
    go(IntLinear<T>{});
  } else if (interp == kNearest) {
    go(IntNearest<T>{});
  } else if (interp == kTSE) {
    go(IntTSE<T, false>{});
  } else if (interp == kTSELimit) {
    go(IntTSE<T, true>{});
  } else {
    dispatch_spline(g, [&](auto KX, auto KZ) {
      constexpr int kx = decltype(KX)::value, kz = decltype(KZ)::value;
      hipLaunchKernelGGL((k_compose_slot_tiled<T, kx, kz>), grid, blk, 0, s, g, f, in, perm, out, keys, vals, first,
                         last, dxout, st);
    });
  }
}

// ---------------------------------------------------------------------------
// Velocity invariants (compute_invariants, invariants_m.f90:15-198) of the
// velocity at T0 (lp%ugrid), one node per thread.  u holds the interior
// velocity (np per component), or is null for the all-zero field the
// reference keeps when no velocity slot was found.  Face neighbours outside
// the interior are the periodic image on a periodic axis and zero otherwise
// (ugrid's ghosts after reset_lp and the face exchange); the gradient is
// grad_sr1's central difference over the ghosted coordinate lines, with the
// column of a one-node axis zeroed.  With S = grad u + grad u^T and
// W = (grad u)^T - grad u, out holds per node (np each):
//   Q  = [QS + QW > 0],  QS = -tr(S S)/8, QW = -tr(W W)/8
//   C  = [RS > 0],       RS = -tr(S S S) * (1/24 as a REAL(4) literal)
//   D  = 1..4 from the signs of D = Q^3 + 27/4 R^2 and R,
//        R = RS - (sum_ijk W_ij W_jk S_ki + W_23 W_31 (S_13 - S_12)) / 8
//        (the reference's seventh row multiplies W_23 W_31 by S_13)
//   H  = sqrt(-QW / QS)
//   L2 = [middle eigenvalue of S S + W W < 0], the (3,3) entry taken as 0
//        as the reference fills it (invariants_m.f90:173-175)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock) k_invariants(GridDev<T> g, const T* __restrict__ u, T* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.np) return;
  int ix[3];
  decode_interior(g, p, ix[0], ix[1], ix[2]);
  const int sp[3] = {1, g.n[0], g.n[0] * g.n[1]};
  T du[3][3];  // du[c][d] = d u_c / d x_d
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const int a = ix[d];
    int pp = -1, pm = -1;
    if (a < g.n[d]) pp = p + sp[d];
    else if (g.per[d]) pp = p - (g.n[d] - 1) * sp[d];
    if (a > 1) pm = p - sp[d];
    else if (g.per[d]) pm = p + (g.n[d] - 1) * sp[d];
    const T den = g.gl[d][a + 1] - g.gl[d][a - 1];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const T* F = u ? u + (size_t)c * g.np : nullptr;
      const T vp = (F && pp >= 0) ? F[pp] : T(0);
      const T vm = (F && pm >= 0) ? F[pm] : T(0);
      du[c][d] = g.n[d] == 1 ? T(0) : (vp - vm) / den;
    }
  }
  T S[3][3], W[3][3];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) {
      S[a][b] = du[b][a] + du[a][b];
      W[a][b] = du[b][a] - du[a][b];
    }
  const T eighth = T(0.125), c24 = T(1.0f / 24.0f), coef = T(6.75);
  T ss = 0, ww = 0, sss = 0, wws = 0;
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) {
      ss += S[a][b] * S[b][a];
      ww += W[a][b] * W[b][a];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        sss += S[a][b] * S[b][c] * S[c][a];
        wws += W[a][b] * W[b][c] * S[c][a];
      }
    }
  wws += W[1][2] * W[2][0] * (S[0][2] - S[0][1]);
  const T qs = -eighth * ss;
  const T rs = -c24 * sss;
  const T q = qs - eighth * ww;
  const T r = rs - eighth * wws;
  const T qw = q - qs;
  const T dd = q * q * q + coef * (r * r);
  T cls;
  if (dd > T(0)) cls = r > T(0) ? T(1) : T(2);
  else cls = r > T(0) ? T(3) : T(4);
  // S S + W W, upper part, (3,3) = 0
  T A[3][3];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) {
      T v = 0;
#pragma unroll
      for (int c = 0; c < 3; c++) v += S[a][c] * S[c][b] + W[a][c] * W[c][b];
      A[a][b] = v;
    }
  double w[3];
  eig_sym3(A[0][0], A[0][1], A[0][2], A[1][1], A[1][2], 0.0, w);
  const double lmid = w[0] + w[1] + w[2] - fmax(w[0], fmax(w[1], w[2])) - fmin(w[0], fmin(w[1], w[2]));
  const size_t np = (size_t)g.np;
  out[p] = q > T(0) ? T(1) : T(0);
  out[np + p] = rs > T(0) ? T(1) : T(0);
  out[2 * np + p] = cls;
  out[3 * np + p] = sqrt(-qw / qs);
  out[4 * np + p] = lmid < 0.0 ? T(1) : T(0);
}

template <typename T>
void launch_invariants(const GridDev<T>& g, const T* u, T* out, hipStream_t s) {
  hipLaunchKernelGGL((k_invariants<T>), grid_for(g.np), dim3(kBlock), 0, s, g, u, out);
}

template <typename T>
void launch_ftle(const GridDev<T>& g, CF3<T> fm, double T_int, int incompressible, double tiny, double* ftle,
                 F3<T> comp, hipStream_t s) {
  hipLaunchKernelGGL((k_ftle<T>), grid_for(g.np), dim3(kBlock), 0, s, g, fm, T_int, incompressible, tiny, ftle, comp);
}

template <typename T, typename U>
void launch_interior_to_user(const GridDev<T>& g, CF3<T> src, int ghosted, U* a, U* b, U* c, hipStream_t s) {
  hipLaunchKernelGGL((k_to_user<T, U>), grid_for(g.np), dim3(kBlock), 0, s, g, src, ghosted, a, b, c);
}

template <typename T>
void launch_vel_ghost(const GridDev<T>& g, F3<T> u, CF3<T> uold, DevStatus* st, hipStream_t s) {
  launch_vel_ghost_impl(g, u, uold, st, s);
}
template <typename T, typename E>
void launch_fm_ghost(const GridDev<T>& g, F3<E> fm, CF3<E> fmold, DevStatus* st, hipStream_t s) {
  launch_fm_ghost_impl(g, fm, fmold, st, s);
}

// dst = src widened to T (n values)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_widen(const float* __restrict__ src, T* __restrict__ dst, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = T(src[i]);
}
template <typename T>
void launch_widen(const float* src, T* dst, long long n, hipStream_t s) {
  const long long nb = std::min<long long>((n + kBlock - 1) / kBlock, 8192);
  hipLaunchKernelGGL((k_widen<T>), dim3((unsigned)std::max<long long>(nb, 1)), dim3(kBlock), 0, s, src, dst, n);
}

// explicit instantiations
#define INST_T(T)                                                                                             \
  template void launch_vel_ghost<T>(const GridDev<T>&, F3<T>, CF3<T>, DevStatus*, hipStream_t);               \
  template void launch_cfl<T>(const GridDev<T>&, F3<T>, DevStatus*, hipStream_t);                             \
  template void launch_fm_ghost<T, T>(const GridDev<T>&, F3<T>, CF3<T>, DevStatus*, hipStream_t);             \
  template void launch_field_from_origin<T>(const GridDev<T>&, CF3<T>, F3<T>, hipStream_t);                   \
  template void launch_spline_fit<T, T, T>(const GridDev<T>&, F3<T>, int, hipStream_t, const T* const*);      \
  template void launch_spline_fit_T<T>(const GridDev<T>&, F3<T>, hipStream_t);                                \
  template void launch_blend<T, T>(const GridDev<T>&, void* const*, int, CF3<T>, CF3<T>, T, T, DevStatus*, hipStream_t);  \
  template void launch_blend_batch<T, T>(const GridDev<T>&, const BlendBatch<T>&, int, CF3<T>, CF3<T>, DevStatus*, \
                                         hipStream_t);                                                         \
  template void launch_sl_step<T, T>(const GridDev<T>&, int, const StageParams<T>&, CF3<T>, F3<T>, DevStatus*, \
                                  hipStream_t, int);                                                          \
  template void launch_fwd_step<T>(const GridDev<T>&, int, const StageParams<T>&, ParticleSet<T>, int, int, \
                                   DevStatus*, hipStream_t);                                                   \
  template void launch_fwd_reset<T>(const GridDev<T>&, ParticleSet<T>, hipStream_t);                          \
  template void launch_track_nodes<T>(const GridDev<T>&, const GridDev<T>&, int16_t*, DevStatus*, hipStream_t); \
  template void launch_fwd_place<T>(const GridDev<T>&, const GridDev<T>&, ParticleSet<T>, DevStatus*,          \
                                    hipStream_t);                                                             \
  template void launch_sl_step2<T, T>(const GridDev<T>&, const GridDev<T>&, const int16_t*, int, int,          \
                                   const StageParams<T>&, CF3<T>, F3<T>, DevStatus*, hipStream_t);            \
  template void launch_expand_periodic<T>(const GridDev<T>&, const T*, F3<T>, hipStream_t);                    \
  template void launch_field_at_nodes<T>(const GridDev<T>&, const GridDev<T>&, const int16_t*, int, CF3<T>,   \
                                         T*, hipStream_t);                                                    \
  template void launch_ftle_aux<T>(const GridDev<T>&, const AuxSets<T>&, double, double, int, double, double*, \
                                   hipStream_t);                                                              \
  template void launch_compose<T>(const GridDev<T>&, int, const T*, long long, const int*, int, F3<T>,       \
                                  DevStatus*, hipStream_t);                                                   \
  template void launch_compose_slot<T>(const GridDev<T>&, int, CF3<T>, const CompRec<T>*, const unsigned*,     \
                                       CompRec<T>*, unsigned*, unsigned*, int, int, F3<T>, DevStatus*,        \
                                       hipStream_t);                                                          \
  template void launch_ftle<T>(const GridDev<T>&, CF3<T>, double, int, double, double*, F3<T>, hipStream_t); \
  template void launch_invariants<T>(const GridDev<T>&, const T*, T*, hipStream_t);                           \
  template void launch_ingest<float, T>(const GridDev<T>&, F3<T>, const float*, const float*, const float*,  \
                                        hipStream_t);                                                         \
  template void launch_ingest<double, T>(const GridDev<T>&, F3<T>, const double*, const double*,             \
                                         const double*, hipStream_t);                                         \
  template void launch_interior_to_user<T, float>(const GridDev<T>&, CF3<T>, int, float*, float*, float*,    \
                                                  hipStream_t);                                               \
  template void launch_interior_to_user<T, double>(const GridDev<T>&, CF3<T>, int, double*, double*,         \
                                                   double*, hipStream_t);

INST_T(float)
INST_T(double)
// FP64 compute over FP32 flowmap storage
template void launch_fm_ghost<double, float>(const GridDev<double>&, F3<float>, CF3<float>, DevStatus*, hipStream_t);
template void launch_spline_fit<double, float, float>(const GridDev<double>&, F3<float>, int, hipStream_t,
                                                      const float* const*);
template void launch_spline_fit<double, float, double>(const GridDev<double>&, F3<float>, int, hipStream_t,
                                                       const double* const*);
template void launch_blend<double, float>(const GridDev<double>&, void* const*, int, CF3<float>, CF3<float>, double,
                                          double, DevStatus*, hipStream_t);
template void launch_blend_batch<double, float>(const GridDev<double>&, const BlendBatch<double>&, int, CF3<float>,
                                                CF3<float>, DevStatus*, hipStream_t);
template void launch_sl_step<double, float>(const GridDev<double>&, int, const StageParams<double>&, CF3<float>,
                                            F3<float>, DevStatus*, hipStream_t, int);
template void launch_sl_step2<double, float>(const GridDev<double>&, const GridDev<double>&, const int16_t*, int, int,
                                             const StageParams<double>&, CF3<float>, F3<float>, DevStatus*,
                                             hipStream_t);
template void launch_widen<double>(const float*, double*, long long, hipStream_t);
template void launch_widen<float>(const float*, float*, long long, hipStream_t);

}  // namespace lcs
