// device_common.h -- device helpers shared by the forward kernels
// (kernels.hip) and the adjoint kernels (adjoint.hip): index maps, the
// XCD-aware block mapping, interpolation (linear and tensor B-spline),
// storage conversion, the integrators' stages, particle BCs, tracking,
// periodic wrap and the 3x3 symmetric eigen-solve.  Reference semantics are
// cited per function.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cstdint>

#include "lcs_device.h"

namespace lcs {

// 27-neighbour order: self, faces, edges/corners (data_m.f90:123-152)
static __constant__ int8_t c_nbr[27][3] = {
    {0, 0, 0},   {-1, 0, 0}, {0, -1, 0}, {0, 0, -1},  {1, 0, 0},   {0, 1, 0},  {0, 0, 1},
    {-1, -1, -1}, {0, -1, -1}, {1, -1, -1}, {-1, 0, -1}, {1, 0, -1}, {-1, 1, -1}, {0, 1, -1},
    {1, 1, -1},  {-1, -1, 0}, {1, -1, 0},  {-1, 1, 0},  {1, 1, 0},   {-1, -1, 1}, {0, -1, 1},
    {1, -1, 1},  {-1, 0, 1},  {1, 0, 1},   {-1, 1, 1},  {0, 1, 1},   {1, 1, 1}};

constexpr int kBlock = 256;

template <typename T>
__device__ __forceinline__ T dmin(T a, T b) { return fmin(a, b); }
template <typename T>
__device__ __forceinline__ T dmax(T a, T b) { return fmax(a, b); }
// Fortran max(min(r,1),0); a NaN ratio (first update) gives 1 as with minNum
template <typename T>
__device__ __forceinline__ T clamp01(T r) { return dmax(dmin(r, T(1)), T(0)); }
__device__ __forceinline__ int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

template <typename T>
__device__ __forceinline__ int gid(const GridDev<T>& g, int i, int j, int k) {
  return i + g.N[0] * (j + g.N[1] * k);
}
template <typename T>
__device__ __forceinline__ void decode_interior(const GridDev<T>& g, int p, int& i, int& j, int& k) {
  i = p % g.n[0] + 1;
  int r = p / g.n[0];
  j = r % g.n[1] + 1;
  k = r / g.n[1] + 1;
}
// floor(e / d) for 0 <= e < 2^20 and 0 < d < 2^20, from rd = 1/d estimated
// in FP32 and corrected by one step either way (a few instructions where an
// integer division by a run-time divisor takes dozens)
__device__ __forceinline__ int small_div(int e, int d, float rd) {
  int q = (int)((float)e * rd);
  if (q * d > e) q--;
  if ((q + 1) * d <= e) q++;
  return q;
}

// The same bijection in slab order: blocks of kSlabJ rows of j, each swept
// through every k, i fastest.  The nodes one XCD has in flight then lie in a
// slab of kSlabJ rows by a few planes, whose stencil rows (cubic: four
// planes and rows around each) mostly stay in that XCD's 4 MB L2 instead of
// being fetched again for every plane a whole-plane sweep passes.  32 rows
// measured best at 256^3 (FP32 and FP64; 16 and 64 rows within 2 %).
constexpr int kSlabJ = 32;
template <typename T>
__device__ __forceinline__ void decode_interior_slab(const GridDev<T>& g, int p, int& i, int& j, int& k) {
  const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
  const int full = n1 / kSlabJ;
  const int slab = kSlabJ * n0 * n2;  // nodes of a full slab
  int jb, rows, r;
  if (p < full * slab) {
    jb = p / slab;
    r = p - jb * slab;
    rows = kSlabJ;
  } else {
    jb = full;
    r = p - full * slab;
    rows = n1 - full * kSlabJ;
  }
  const int plane = rows * n0;
  k = r / plane + 1;
  r -= (k - 1) * plane;
  j = jb * kSlabJ + r / n0 + 1;
  i = r % n0 + 1;
}
// inverse of decode_interior_slab: the slab-order index of interior node
// (i, j, k) of an n0 x n1 x n2 grid
__device__ __forceinline__ int encode_interior_slab(int n0, int n1, int n2, int i, int j, int k) {
  const int full = n1 / kSlabJ;
  const int jb = (j - 1) / kSlabJ, jj = (j - 1) - jb * kSlabJ;
  const int rows = jb < full ? kSlabJ : n1 - full * kSlabJ;
  return jb * kSlabJ * n0 * n2 + ((k - 1) * rows + jj) * n0 + (i - 1);
}
template <typename T>
__device__ __forceinline__ int pid(const GridDev<T>& g, int i, int j, int k) {
  return (i - 1) + g.n[0] * ((j - 1) + g.n[1] * (k - 1));
}
template <typename T>
__device__ __forceinline__ bool self_comm(const GridDev<T>& g, int ei, int ej, int ek) {
  if (ei == 0 && ej == 0 && ek == 0) return false;
  return (ei == 0 || g.per[0]) && (ej == 0 || g.per[1]) && (ek == 0 || g.per[2]);
}
// ghost shell enumeration: thread t -> ghost cell (i,j,k)
template <typename T>
__device__ __forceinline__ bool decode_ghost(const GridDev<T>& g, long long t, int& i, int& j, int& k) {
  const long long plane = (long long)g.N[0] * g.N[1];
  if (t < plane) { k = 0; i = (int)(t % g.N[0]); j = (int)(t / g.N[0]); return true; }
  t -= plane;
  if (t < plane) { k = g.N[2] - 1; i = (int)(t % g.N[0]); j = (int)(t / g.N[0]); return true; }
  t -= plane;
  const long long per_k = 2LL * g.N[0] + 2LL * g.n[1];
  if (t >= per_k * g.n[2]) return false;
  k = 1 + (int)(t / per_k);
  int r = (int)(t % per_k);
  if (r < g.N[0]) { j = 0; i = r; return true; }
  r -= g.N[0];
  if (r < g.N[0]) { j = g.N[1] - 1; i = r; return true; }
  r -= g.N[0];
  j = 1 + r / 2;
  i = (r & 1) ? g.N[0] - 1 : 0;
  return true;
}
template <typename T>
__device__ __forceinline__ long long ghost_count(const GridDev<T>& g) {
  return 2LL * g.N[0] * g.N[1] + (2LL * g.N[0] + 2LL * g.n[1]) * g.n[2];
}

// Workgroups are dispatched round-robin over the 8 XCDs; this bijection gives
// each XCD a contiguous range of work items so that neighbouring items share
// that XCD's L2.
__device__ __forceinline__ int xcd_brick(int w, int nb) {
  const int xcd = w & 7, idx = w >> 3, q = nb >> 3, r = nb & 7;
  return xcd * q + min(xcd, r) + idx;
}
__device__ __forceinline__ int xcd_thread_index() {
  return xcd_brick(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
}

// ---------------------------------------------------------------------------
// Interpolation (interp_s2u_r1, lp_tracking_m.f90:279-662)
// ---------------------------------------------------------------------------
template <typename T>
struct LinW {
  int base;
  T t[3], mt[3];
};

// A point known as a grid coordinate plus an offset (q = qb + off before
// rounding): the interpolants form their local differences q - x_a as
// (qb - x_a) + off, which is exact or nearly so for x_a a neighbouring grid
// coordinate, instead of from the rounded |q|-sized position.
template <typename T>
struct Split {
  T qb[3], off[3];
};

// LINEAR bracket relative to the tracked node, local coordinate clamped to
// [0,1] (lp_tracking_m.f90:381-396,438-448).  A bracket that would leave the
// ghosted range is shifted inside.
template <typename T>
__device__ __forceinline__ void lin_setup(const GridDev<T>& g, const T q[3], const int nd[3], LinW<T>& w,
                                          const Split<T>* sq = nullptr) {
  int i0[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const T* xl = g.xl[d];
    int n0 = iclamp(nd[d], 0, g.n[d] + 1);
    int a = (q[d] > xl[n0]) ? n0 : n0 - 1;
    a = iclamp(a, 0, g.n[d]);
    T x0 = xl[a], x1 = xl[a + 1];
    const T dq = sq ? (sq->qb[d] - x0) + sq->off[d] : q[d] - x0;
    T tt = clamp01(dq / (x1 - x0));
    w.t[d] = tt;
    w.mt[d] = T(1) - tt;
    i0[d] = a;
  }
  w.base = i0[0] + g.N[0] * (i0[1] + g.N[1] * i0[2]);
}

// trilinear sum in the reference's term order (lp_tracking_m.f90:449-457)
template <typename T>
__device__ __forceinline__ T trilin(const LinW<T>& w, T f0, T f1, T f2, T f3, T f4, T f5, T f6, T f7) {
  return w.mt[0] * w.mt[1] * w.mt[2] * f0 + w.t[0] * w.mt[1] * w.mt[2] * f1 +
         w.mt[0] * w.t[1] * w.mt[2] * f2 + w.t[0] * w.t[1] * w.mt[2] * f3 +
         w.mt[0] * w.mt[1] * w.t[2] * f4 + w.t[0] * w.mt[1] * w.t[2] * f5 +
         w.mt[0] * w.t[1] * w.t[2] * f6 + w.t[0] * w.t[1] * w.t[2] * f7;
}

// B-spline basis of order K at x (bspline_sub_module.f90: dintrv 2558-2634,
// dbspvn 2288-2370).  The interval index is the largest L in [K, n] with
// t(L) <= x (dintrv; dbvalu's left limit at x = t(n+1)), found by probing from
// a hint next to the tracked node; the K nonzero basis values come from the
// interval's polynomial table (exact piecewise form of the Cox-de Boor
// recurrence, evaluated by Horner in the normalised interval coordinate).
// Returns 0 = ok, 1 = outside the knot range (db3val returns 0 and reports),
// 2 = y/z at the last knot (db3val returns 0 silently,
// bspline_sub_module.f90:634-636).
// LDS pointers (address space 3): data a workgroup stages once and its lanes
// read many times
template <typename T>
using lcptr = const __attribute__((address_space(3))) T*;
template <typename T>
using lptr = __attribute__((address_space(3))) T*;

// A workgroup's LDS copy of rows [r0, r0 + nr) of one axis's per-interval
// basis table (K*K+1 values per row) and of the knots t(r0-1) .. t(r0+nr)
// (0-based), covering the intervals its lanes' points fall in; nr = 0:
// nothing staged.  The B-spline set-up reads a row or knot from here when it
// is staged and from global memory otherwise, so staging changes no value.
template <typename T>
struct AxisLds {
  lcptr<T> poly = nullptr;
  lcptr<T> knot = nullptr;
  int r0 = 0, nr = 0;
};
template <typename T>
struct SplineLds {
  AxisLds<T> ax[3];
};

// interval of x: the largest L in [K, n] with t(L) <= x, probing from a hint
// (0 = ok, 1 = outside the knot range, 2 = y/z at the last knot)
template <typename T, int K>
__device__ __forceinline__ int bs_locate(const T* __restrict__ t, int n, T x, bool xaxis, int hint, int& left,
                                         const AxisLds<T>* al = nullptr) {
  const int L = n + K;
  if (!(x >= t[0]) || x > t[L - 1]) return 1;
  const bool st = al && al->nr > 0;
  const int klo = st ? al->r0 - 1 : 0, khi = st ? al->r0 + al->nr : -1;
  auto kt = [&](int a) -> T { return (a >= klo && a <= khi) ? al->knot[a - klo] : t[a]; };
  if (x >= t[L - 1]) {
    if (!xaxis) return 2;
    left = n;
  } else {
    left = iclamp(hint, K, n);
    while (left < n && x >= kt(left)) left++;
    while (left > K && x < kt(left - 1)) left--;
  }
  return 0;
}

// the K nonzero basis values at x in interval `left`, from the interval's
// polynomial table (exact piecewise form of the Cox-de Boor recurrence,
// Horner in the normalised interval coordinate)
template <typename T, int K>
__device__ __forceinline__ void bs_weights(const T* __restrict__ t, const T* __restrict__ poly, int left, T x, T v[K],
                                           bool hx = false, T xb = T(0), T xo = T(0),
                                           const AxisLds<T>* al = nullptr) {
  auto eval = [&](auto P, T tl) {
    // hx: x is given as base coordinate xb plus offset xo (Split)
    const T s = (hx ? (xb - tl) + xo : x - tl) * P[K * K];
#pragma unroll
    for (int a = 0; a < K; a++) {
      T w = P[a * K + K - 1];
#pragma unroll
      for (int m = K - 2; m >= 0; m--) w = w * s + P[a * K + m];
      v[a] = w;
    }
  };
  if (al && al->nr > 0 && left >= al->r0 && left < al->r0 + al->nr)
    eval(al->poly + (left - al->r0) * (K * K + 1), al->knot[left - al->r0]);
  else
    eval(poly + left * (K * K + 1), t[left - 1]);
}

// B-spline basis of order K at x (bspline_sub_module.f90: dintrv 2558-2634,
// dbspvn 2288-2370).  The interval index is the largest L in [K, n] with
// t(L) <= x (dintrv; dbvalu's left limit at x = t(n+1)), found by probing from
// a hint next to the tracked node; the K nonzero basis values come from the
// interval's polynomial table.  Returns 0 = ok, 1 = outside the knot range
// (db3val returns 0 and reports), 2 = y/z at the last knot (db3val returns 0
// silently, bspline_sub_module.f90:634-636).
template <typename T, int K>
__device__ __forceinline__ int bs_setup(const T* __restrict__ t, const T* __restrict__ poly, int n, T x, bool xaxis,
                                        int hint, int& base, T v[K], bool hx = false, T xb = T(0), T xo = T(0),
                                        const AxisLds<T>* al = nullptr) {
  int left = 0;
  const int r = bs_locate<T, K>(t, n, x, xaxis, hint, left, al);
  if (r) return r;
  bs_weights<T, K>(t, poly, left, x, v, hx, xb, xo, al);
  base = left - K;
  return 0;
}

// ---- stored element -> compute type ----
template <typename T>
__device__ __forceinline__ T widen(T x) { return x; }
__device__ __forceinline__ float widen(half_t h) { return __half2float(__ushort_as_half(h.u)); }
__device__ __forceinline__ float widen(bf16_t h) { return __uint_as_float((unsigned)h.u << 16); }
// compute type -> stored element, round to nearest even
template <typename S, typename T>
__device__ __forceinline__ S narrow(T x) { return S(x); }
template <>
__device__ __forceinline__ half_t narrow<half_t, float>(float x) { return half_t{__half_as_ushort(__float2half_rn(x))}; }
template <>
__device__ __forceinline__ bf16_t narrow<bf16_t, float>(float x) {
  unsigned u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return bf16_t{(uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u))};
  u += 0x7fffu + ((u >> 16) & 1u);
  return bf16_t{(uint16_t)(u >> 16)};
}
// packed 16-bit records: the element type and its widening
template <typename S> struct Packed { static constexpr bool value = false; };
template <> struct Packed<half4_t> { static constexpr bool value = true; using E = half_t; };
template <> struct Packed<bf164_t> { static constexpr bool value = true; using E = bf16_t; };
template <typename P>
__device__ __forceinline__ void widen3(const P& r, float v[3]) {
#pragma unroll
  for (int c = 0; c < 3; c++) v[c] = widen(r.e[c]);
}
// K consecutive packed records (one x row of a spline stencil), 8 K bytes
template <int K, typename P>
__device__ __forceinline__ void load_row3(const P* __restrict__ row, float v[K][3]) {
  struct __attribute__((aligned(8))) R { P e[K]; } r;
  __builtin_memcpy(&r, row, sizeof(R));
#pragma unroll
  for (int a = 0; a < K; a++) widen3(r.e[a], v[a]);
}

// K consecutive stored elements (one x row of a spline stencil); 16-bit rows
// are fetched as one unaligned 2K-byte load
template <typename T, int K, typename S>
__device__ __forceinline__ void load_row(const S* __restrict__ row, T v[K]) {
  if constexpr (sizeof(S) >= 4) {
#pragma unroll
    for (int a = 0; a < K; a++) v[a] = T(row[a]);
  } else {
    struct __attribute__((packed, aligned(2))) R { S e[K]; } r;
    __builtin_memcpy(&r, row, sizeof(R));
#pragma unroll
    for (int a = 0; a < K; a++) v[a] = T(widen(r.e[a]));
  }
}

template <typename T, int KX, int KZ>
struct SplineW {
  int base;       // flattened ghosted index of the first coefficient
  int b3[3];      // first coefficient index per axis
  T wx[KX], wy[KX], wz[KZ];
};

// nd: node next to q (tracked node), used as the interval hint; sl: the
// workgroup's staged tables, if any
template <typename T, int KX, int KZ>
__device__ __forceinline__ int spline_setup(const GridDev<T>& g, const T q[3], const int nd[3],
                                            SplineW<T, KX, KZ>& s, const Split<T>* sq = nullptr,
                                            const SplineLds<T>* sl = nullptr) {
  int bx, by, bz;
  const bool h = sq != nullptr;
  T xb[3] = {T(0), T(0), T(0)}, xo[3] = {T(0), T(0), T(0)};
  if (h) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
      xb[d] = sq->qb[d];
      xo[d] = sq->off[d];
    }
  }
  int rx = bs_setup<T, KX>(g.kn[0], g.bpoly[0], g.N[0], q[0], true, nd[0] + (KX + 2) / 2, bx, s.wx, h, xb[0], xo[0],
                           sl ? &sl->ax[0] : nullptr);
  int ry = bs_setup<T, KX>(g.kn[1], g.bpoly[1], g.N[1], q[1], false, nd[1] + (KX + 2) / 2, by, s.wy, h, xb[1], xo[1],
                           sl ? &sl->ax[1] : nullptr);
  int rz = bs_setup<T, KZ>(g.kn[2], g.bpoly[2], g.N[2], q[2], false, nd[2] + (KZ + 2) / 2, bz, s.wz, h, xb[2], xo[2],
                           sl ? &sl->ax[2] : nullptr);
  if (rx == 1 || ry == 1 || rz == 1) return 1;
  if (ry || rz) return 2;
  s.base = bx + g.N[0] * (by + g.N[1] * bz);
  s.b3[0] = bx; s.b3[1] = by; s.b3[2] = bz;
  return 0;
}

// Copies into LDS the basis-table rows and knots of the intervals of points
// within a cell of nodes [lo[d], hi[d]] (hints node + (K+2)/2, at most two
// intervals away), at most cap[d] rows per axis; no synchronisation.
template <typename T, int KX, int KZ>
__device__ __forceinline__ void copy_spline_tables(const GridDev<T>& g, const int lo[3], const int hi[3], lptr<T> buf,
                                                   const int cap[3], SplineLds<T>& sl) {
  size_t off = 0;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const int K = d < 2 ? KX : KZ, RW = K * K + 1;
    AxisLds<T>& al = sl.ax[d];
    int r0 = lo[d] + (K + 2) / 2 - 2, r1 = hi[d] + (K + 2) / 2 + 2;
    r0 = max(r0, K);
    r1 = min(r1, g.N[d]);
    int nr = r1 - r0 + 1;
    nr = nr < 0 ? 0 : (nr > cap[d] ? cap[d] : nr);
    lptr<T> pb = buf + off;
    lptr<T> kb = buf + off + (size_t)cap[d] * RW;
    off += (size_t)cap[d] * RW + cap[d] + 2;
    al.poly = pb;
    al.knot = kb;
    al.r0 = r0;
    al.nr = nr;
    const T* P = g.bpoly[d] + (size_t)r0 * RW;
    for (int e = threadIdx.x; e < nr * RW; e += blockDim.x) pb[e] = P[e];
    const T* t = g.kn[d] + (r0 - 1);
    for (int e = threadIdx.x; e < nr + 2; e += blockDim.x) kb[e] = t[e];
  }
}

// Stages into LDS the basis-table rows and knots of the intervals a
// workgroup's points can fall in: nodes [lo[d], hi[d]] (reduced over the
// workgroup's lanes, inactive lanes passing INT_MAX / INT_MIN) give hints
// node + (K+2)/2, and points within a cell of their node fall at most two
// intervals from the hint.  Axis capacities cap[d] rows; every thread of the
// block calls it (it synchronises).
template <typename T, int KX, int KZ>
__device__ __forceinline__ void stage_spline_tables(const GridDev<T>& g, const int nd[3], bool active, lptr<T> buf,
                                                    const int cap[3], lptr<int> red, SplineLds<T>& sl) {
  if (threadIdx.x < 3) {
    red[threadIdx.x] = INT_MAX;
    red[3 + threadIdx.x] = INT_MIN;
  }
  __syncthreads();
  int lo[3], hi[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    lo[d] = active ? nd[d] : INT_MAX;
    hi[d] = active ? nd[d] : INT_MIN;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
      lo[d] = min(lo[d], __shfl_xor(lo[d], o, 64));
      hi[d] = max(hi[d], __shfl_xor(hi[d], o, 64));
    }
  }
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
      atomicMin((int*)&red[d], lo[d]);
      atomicMax((int*)&red[3 + d], hi[d]);
    }
  }
  __syncthreads();
  const int lo3[3] = {red[0], red[1], red[2]}, hi3[3] = {red[3], red[4], red[5]};
  copy_spline_tables<T, KX, KZ>(g, lo3, hi3, buf, cap, sl);
  __syncthreads();
}
// LDS values stage_spline_tables needs for capacities cap (rows per axis)
template <int KX, int KZ>
constexpr int spline_tables_lds(int cx, int cy, int cz) {
  return cx * (KX * KX + 1) + cx + 2 + cy * (KX * KX + 1) + cy + 2 + cz * (KZ * KZ + 1) + cz + 2;
}

template <typename T, int KX, int KZ, typename S>
__device__ __forceinline__ T spline_sum(const GridDev<T>& g, const SplineW<T, KX, KZ>& s, const S* __restrict__ c) {
  const int sy = g.N[0], sz = g.N[0] * g.N[1];
  T acc = T(0);
#pragma unroll
  for (int cz = 0; cz < KZ; cz++) {
    T accy = T(0);
#pragma unroll
    for (int by = 0; by < KX; by++) {
      T row[KX];
      load_row<T, KX>(c + s.base + by * sy + cz * sz, row);
      T accx = T(0);
#pragma unroll
      for (int a = 0; a < KX; a++) accx += s.wx[a] * row[a];
      accy += s.wy[by] * accx;
    }
    acc += s.wz[cz] * accy;
  }
  return acc;
}

// ---- interpolants of a vector field f at q, bracket/hint from node nd.  For
// the velocity, f is the stage's time-blended field (values for NEAREST and
// LINEAR, spline coefficients for QUADRATIC/CUBIC); for the flowmap, the
// (fitted) flowmap. ----
template <typename T>
struct IntNearest {
  static constexpr bool kNonRect = false;
  template <typename S>
  __device__ void operator()(const GridDev<T>& g, const T q[3], const int nd[3], const CF3<S>& f, T out[3],
                             unsigned long long&, const Split<T>* = nullptr) const {
    int l = gid(g, iclamp(nd[0], 0, g.n[0] + 1), iclamp(nd[1], 0, g.n[1] + 1), iclamp(nd[2], 0, g.n[2] + 1));
    if constexpr (Packed<S>::value) {
      float v[3];
      widen3(f.c[0][l], v);
#pragma unroll
      for (int c = 0; c < 3; c++) out[c] = T(v[c]);
    } else {
#pragma unroll
      for (int c = 0; c < 3; c++) out[c] = T(widen(f.c[c][l]));
    }
  }
};

// TSE / TSE_LIMIT (lp_tracking_m.f90:526-654): first-order Taylor expansion
// about the node, phi_no + grad(phi)_no . (x - x_no), with the central
// difference gradient of grad_sr1 (gradient_m.f90:11-66: ghosted
// coordinates, the column of an axis with one node zeroed); LIMIT clamps the
// result to the min/max of phi over the node's 3x3x3 neighbourhood.  The
// gradient is formed at the node on the fly (7 or 27 gathers per component)
// instead of as a 9-component field.
template <typename T, bool LIMIT>
struct IntTSE {
  static constexpr bool kNonRect = false;
  template <typename S>
  __device__ static void at(const CF3<S>& f, int l, T v[3]) {
    if constexpr (Packed<S>::value) {
      float w[3];
      widen3(f.c[0][l], w);
#pragma unroll
      for (int c = 0; c < 3; c++) v[c] = T(w[c]);
    } else {
#pragma unroll
      for (int c = 0; c < 3; c++) v[c] = T(widen(f.c[c][l]));
    }
  }
  template <typename S>
  __device__ void operator()(const GridDev<T>& g, const T q[3], const int nd[3], const CF3<S>& f, T out[3],
                             unsigned long long&, const Split<T>* sq = nullptr) const {
    const int ix[3] = {iclamp(nd[0], 1, g.n[0]), iclamp(nd[1], 1, g.n[1]), iclamp(nd[2], 1, g.n[2])};
    const int st[3] = {1, g.N[0], g.N[0] * g.N[1]};
    const int l = gid(g, ix[0], ix[1], ix[2]);
    T r[3];
    at(f, l, r);
#pragma unroll
    for (int c = 0; c < 3; c++) out[c] = r[c];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      if (g.n[d] == 1) continue;  // zeroed gradient column
      const T* G = g.gl[d];
      const T den = G[ix[d] + 1] - G[ix[d] - 1];
      const T dl = sq ? (sq->qb[d] - G[ix[d]]) + sq->off[d] : q[d] - G[ix[d]];
      T a[3], b[3];
      at(f, l + st[d], a);
      at(f, l - st[d], b);
#pragma unroll
      for (int c = 0; c < 3; c++) out[c] = out[c] + dl * ((a[c] - b[c]) / den);
    }
    if constexpr (LIMIT) {
      T mn[3] = {r[0], r[1], r[2]}, mx[3] = {r[0], r[1], r[2]};
      for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
#pragma unroll
          for (int dx = -1; dx <= 1; dx++) {
            T v[3];
            at(f, l + dx + dy * st[1] + dz * st[2], v);
#pragma unroll
            for (int c = 0; c < 3; c++) {
              mn[c] = fmin(mn[c], v[c]);
              mx[c] = fmax(mx[c], v[c]);
            }
          }
#pragma unroll
      for (int c = 0; c < 3; c++) out[c] = fmin(fmax(out[c], mn[c]), mx[c]);
    }
  }
};

// TSE / TSE_LIMIT on a non-rectilinear grid, where every interpolation is
// one of them (lp_tracking_m.f90:319-328): the gradient at the node is
// grad_sr1_ls (gradient_m.f90:241-295), the sum over the face neighbours in
// NBR order of (phi(nbr) - phi(node)) times the node's least-squares
// weights, and x - x_no uses the full node coordinates.
template <typename T, bool LIMIT>
struct IntTSELS {
  static constexpr bool kNonRect = true;
  template <typename S>
  __device__ void operator()(const GridDev<T>& g, const T q[3], const int nd[3], const CF3<S>& f, T out[3],
                             unsigned long long&, const Split<T>* sq = nullptr) const {
    const int ix[3] = {iclamp(nd[0], 1, g.n[0]), iclamp(nd[1], 1, g.n[1]), iclamp(nd[2], 1, g.n[2])};
    const int st[3] = {1, g.N[0], g.N[0] * g.N[1]};
    const int l = gid(g, ix[0], ix[1], ix[2]);
    const long long p = pid(g, ix[0], ix[1], ix[2]);
    T r[3];
    IntTSE<T, LIMIT>::at(f, l, r);
    T G[3][3] = {{T(0), T(0), T(0)}, {T(0), T(0), T(0)}, {T(0), T(0), T(0)}};
    const int off[6] = {-st[0], -st[1], -st[2], st[0], st[1], st[2]};
#pragma unroll
    for (int fc = 0; fc < 6; fc++) {
      T a[3];
      IntTSE<T, LIMIT>::at(f, l + off[fc], a);
      const T* w = g.lsw + (long long)fc * 3 * g.np;
      const T w3[3] = {w[p], w[g.np + p], w[2 * g.np + p]};
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const T t = a[c] - r[c];
#pragma unroll
        for (int d = 0; d < 3; d++) G[c][d] = G[c][d] + t * w3[d];
      }
    }
    T dl[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const T xn = g.gx[d][l];
      dl[d] = sq ? (sq->qb[d] - xn) + sq->off[d] : q[d] - xn;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) out[c] = r[c] + dl[0] * G[c][0] + dl[1] * G[c][1] + dl[2] * G[c][2];
    if constexpr (LIMIT) {
      T mn[3] = {r[0], r[1], r[2]}, mx[3] = {r[0], r[1], r[2]};
      for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
#pragma unroll
          for (int dx = -1; dx <= 1; dx++) {
            T v[3];
            IntTSE<T, LIMIT>::at(f, l + dx + dy * st[1] + dz * st[2], v);
#pragma unroll
            for (int c = 0; c < 3; c++) {
              mn[c] = fmin(mn[c], v[c]);
              mx[c] = fmax(mx[c], v[c]);
            }
          }
#pragma unroll
      for (int c = 0; c < 3; c++) out[c] = fmin(fmax(out[c], mn[c]), mx[c]);
    }
  }
};

template <typename T>
struct IntLinear {
  static constexpr bool kNonRect = false;
  template <typename S>
  __device__ void operator()(const GridDev<T>& g, const T q[3], const int nd[3], const CF3<S>& f, T out[3],
                             unsigned long long&, const Split<T>* sq = nullptr) const {
    LinW<T> w;
    lin_setup(g, q, nd, w, sq);
    const int b = w.base, sy = g.N[0], sz = g.N[0] * g.N[1];
    if constexpr (Packed<S>::value) {
      const S* A = f.c[0];
      float v[8][3];
      const int o[8] = {b, b + 1, b + sy, b + sy + 1, b + sz, b + sz + 1, b + sz + sy, b + sz + sy + 1};
#pragma unroll
      for (int e = 0; e < 8; e++) widen3(A[o[e]], v[e]);
#pragma unroll
      for (int c = 0; c < 3; c++)
        out[c] = trilin(w, T(v[0][c]), T(v[1][c]), T(v[2][c]), T(v[3][c]), T(v[4][c]), T(v[5][c]), T(v[6][c]),
                        T(v[7][c]));
    } else {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const S* A = f.c[c];
        out[c] = trilin(w, T(widen(A[b])), T(widen(A[b + 1])), T(widen(A[b + sy])), T(widen(A[b + sy + 1])),
                        T(widen(A[b + sz])), T(widen(A[b + sz + 1])), T(widen(A[b + sz + sy])),
                        T(widen(A[b + sz + sy + 1])));
      }
    }
  }
};

template <typename T, int KX, int KZ>
struct IntSpline {
  static constexpr bool kNonRect = false;
  static constexpr int kKX = KX, kKZ = KZ;
  SplineLds<T> lds;      // the workgroup's staged basis tables
  bool use_lds = false;  // read them (else every table read goes to global memory)
  template <typename S>
  __device__ void operator()(const GridDev<T>& g, const T q[3], const int nd[3], const CF3<S>& f, T out[3],
                             unsigned long long& oob, const Split<T>* sq = nullptr) const {
    SplineW<T, KX, KZ> s;
    int r = spline_setup(g, q, nd, s, sq, use_lds ? &lds : nullptr);
    if (r) {
      if (r == 1) oob++;
      out[0] = out[1] = out[2] = T(0);
      return;
    }
    if constexpr (Packed<S>::value) {
      // all three components from the packed rows, same sum order per
      // component as spline_sum
      const int sy = g.N[0], sz = g.N[0] * g.N[1];
      T acc[3] = {T(0), T(0), T(0)};
#pragma unroll
      for (int cz = 0; cz < KZ; cz++) {
        T accy[3] = {T(0), T(0), T(0)};
#pragma unroll
        for (int by = 0; by < KX; by++) {
          float row[KX][3];
          load_row3<KX>(f.c[0] + s.base + by * sy + cz * sz, row);
#pragma unroll
          for (int c = 0; c < 3; c++) {
            T accx = T(0);
#pragma unroll
            for (int a = 0; a < KX; a++) accx += s.wx[a] * T(row[a][c]);
            accy[c] += s.wy[by] * accx;
          }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) acc[c] += s.wz[cz] * accy[c];
      }
#pragma unroll
      for (int c = 0; c < 3; c++) out[c] = acc[c];
    } else {
#pragma unroll
      for (int c = 0; c < 3; c++) out[c] = spline_sum(g, s, f.c[c]);
    }
  }
};

// the B-spline interpolants (whose kernels stage basis tables in LDS)
template <class I>
struct SplineInterp : std::false_type {};
template <typename T, int KX, int KZ>
struct SplineInterp<IntSpline<T, KX, KZ>> : std::true_type {};
// LDS rows per axis of the staged basis tables: x covers a workgroup's 256
// nodes of one row plus the stencil reach, y and z a few rows
constexpr int kTabX = 264, kTabY = 24;

// ---------------------------------------------------------------------------
// Particle operations
// ---------------------------------------------------------------------------

// stage velocity field i of a step (time-blended values or coefficients)
template <typename S, typename T>
__device__ __forceinline__ CF3<S> stage(const StageParams<T>& sp, int i) {
  return CF3<S>{{(const S*)sp.v[i][0], (const S*)sp.v[i][1], (const S*)sp.v[i][2]}};
}

// a velocity gathered from a 16-bit stage field, back to its scale
template <typename S, typename T>
__device__ __forceinline__ void stage_scale(const StageParams<T>& sp, T up[3]) {
  if constexpr (Packed<S>::value) {
#pragma unroll
    for (int d = 0; d < 3; d++) up[d] = up[d] * sp.vscale;
  }
}

// integrate_lp stages (lp_motion_m.f90:129-328) in increment form:
// inc = x_new - xn summed from the stage terms alone, so that no term is
// rounded against |xn|.  RK3 evaluates k3 at xn + 2 k2 as the reference does
// (lp_motion_m.f90:239-247).  The first stage of every scheme samples the
// velocity at the start position xn, which first(up) supplies (an
// interpolation, or the nodal value when xn is a node).
// FIX >= 0 fixes the scheme at compile time (the kernel then holds one
// scheme's stages only).  hook(std::integral_constant<int, k>) runs before
// the gather of stage field k (1..3).
struct NoHook {
  template <class K>
  __device__ __forceinline__ void operator()(K) const {}
};
template <typename S, int FIX = -1, typename T, class V, class V0, class H = NoHook>
__device__ __forceinline__ void integrate_inc(const GridDev<T>& g, const StageParams<T>& sp, const V& vel,
                                              const V0& first, const int nd[3], const T xn[3], T inc[3],
                                              unsigned long long& oob, const T* xnl = nullptr,
                                              const H& hook = H{}) {
  // no contraction of the stage combinations: their products feed more than
  // one sum, and whether the compiler fuses them would otherwise depend on
  // the kernel around them (every kernel must round a step alike)
#pragma clang fp contract(off)
  using K1 = std::integral_constant<int, 1>;
  using K2 = std::integral_constant<int, 2>;
  using K3 = std::integral_constant<int, 3>;
  const T HALF = T(1) / T(2), THIRD = T(1) / T(3), TWOTHIRD = T(2) / T(3), SIXTH = T(1) / T(6), TWO = T(2);
  const T dt = sp.dt;
  T up[3], k[3], xp[3];
  // stage point xn + c k, handed to the interpolant as xn plus an offset
  // (plus xnl, the low part of a compensated position)
  Split<T> sq;
  auto at = [&](T c, const T kk[3]) {
#pragma clang fp contract(off)
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const T o = c * kk[d];
      xp[d] = xn[d] + o;
      sq.qb[d] = xn[d];
      sq.off[d] = xnl ? xnl[d] + o : o;
    }
  };
  switch (FIX >= 0 ? FIX : sp.integ) {
    case kEuler:
    case kTrapezoidal:
      first(up);
#pragma unroll
      for (int d = 0; d < 3; d++) inc[d] = dt * up[d];
      break;
    case kRK2:
      first(up);
#pragma unroll
      for (int d = 0; d < 3; d++) k[d] = dt * up[d];
      at(HALF, k);
      hook(K1{});
      vel(g, xp, nd, stage<S>(sp, 1), up, oob, &sq);
      stage_scale<S>(sp, up);
#pragma unroll
      for (int d = 0; d < 3; d++) inc[d] = dt * up[d];
      break;
    case kRK3:
      first(up);
#pragma unroll
      for (int d = 0; d < 3; d++) {
        k[d] = dt * up[d];
        inc[d] = SIXTH * k[d];
      }
      at(HALF, k);
      hook(K1{});
      vel(g, xp, nd, stage<S>(sp, 1), up, oob, &sq);
      stage_scale<S>(sp, up);
#pragma unroll
      for (int d = 0; d < 3; d++) {
        k[d] = dt * up[d];
        inc[d] = inc[d] + TWOTHIRD * k[d];
      }
      at(TWO, k);
      hook(K2{});
      vel(g, xp, nd, stage<S>(sp, 2), up, oob, &sq);
      stage_scale<S>(sp, up);
#pragma unroll
      for (int d = 0; d < 3; d++) inc[d] = inc[d] + SIXTH * (dt * up[d]);
      break;
    case kRK4:
    default:
      first(up);
#pragma unroll
      for (int d = 0; d < 3; d++) {
        k[d] = dt * up[d];
        inc[d] = SIXTH * k[d];
      }
      at(HALF, k);
      hook(K1{});
      vel(g, xp, nd, stage<S>(sp, 1), up, oob, &sq);
      stage_scale<S>(sp, up);
#pragma unroll
      for (int d = 0; d < 3; d++) {
        k[d] = dt * up[d];
        inc[d] = inc[d] + THIRD * k[d];
      }
      at(HALF, k);
      hook(K2{});
      vel(g, xp, nd, stage<S>(sp, 2), up, oob, &sq);
      stage_scale<S>(sp, up);
#pragma unroll
      for (int d = 0; d < 3; d++) {
        k[d] = dt * up[d];
        inc[d] = inc[d] + THIRD * k[d];
      }
      at(T(1), k);
      hook(K3{});
      vel(g, xp, nd, stage<S>(sp, 3), up, oob, &sq);
      stage_scale<S>(sp, up);
#pragma unroll
      for (int d = 0; d < 3; d++) inc[d] = inc[d] + SIXTH * (dt * up[d]);
      break;
  }
}

// error-free transformation a + b = s + e (Knuth two-sum)
template <typename T>
__device__ __forceinline__ void two_sum(T a, T b, T& s, T& e) {
#pragma clang fp contract(off)
  s = a + b;
  const T bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// project_lp2plane on a flat separable grid (lp_m.f90:418-538): the plane
// normal is the collapsed axis, the node index on it becomes 1 and the
// particle's coordinate loses its offset from the node plane.
template <bool NR = false, typename T>
__device__ __forceinline__ void project_flat(const GridDev<T>& g, int nd[3], T q[3]) {
  if (g.flat < 0) return;
  const int f = g.flat;
  // axis f by unrolled selects: a runtime index would put nd and q in
  // scratch memory
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (d == f) {
      nd[d] = 1;
      if constexpr (!NR) q[d] = q[d] - (q[d] - g.gl[d][1]);
    }
  }
  if constexpr (NR) {
    // project_lp2plane (lp_m.f90:450-536): the plane through the node with
    // the normal v1 x v2 of the in-plane central differences
    const int i = iclamp(nd[0], 1, g.n[0]), j = iclamp(nd[1], 1, g.n[1]), k = iclamp(nd[2], 1, g.n[2]);
    const int a0 = f == 0 ? 1 : 0, a1 = f == 2 ? 1 : 2;
    int u0[3] = {i, j, k}, u1[3] = {i, j, k}, w0[3] = {i, j, k}, w1[3] = {i, j, k};
#pragma unroll
    for (int d = 0; d < 3; d++) {
      if (d == a0) {
        u0[d] -= 1;
        u1[d] += 1;
      }
      if (d == a1) {
        w0[d] -= 1;
        w1[d] += 1;
      }
    }
    T v1[3], v2[3], x[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      v1[d] = g.gx[d][gid(g, u1[0], u1[1], u1[2])] - g.gx[d][gid(g, u0[0], u0[1], u0[2])];
      v2[d] = g.gx[d][gid(g, w1[0], w1[1], w1[2])] - g.gx[d][gid(g, w0[0], w0[1], w0[2])];
      x[d] = g.gx[d][gid(g, i, j, k)];
    }
    T nrm[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2], v1[0] * v2[1] - v1[1] * v2[0]};
    const T mag = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
#pragma unroll
    for (int d = 0; d < 3; d++) nrm[d] = nrm[d] / mag;
    T nm = (q[0] - x[0]) * nrm[0] + (q[1] - x[1]) * nrm[1] + (q[2] - x[2]) * nrm[2];
    if (nm < T(0)) {
      nm = -nm;
#pragma unroll
      for (int d = 0; d < 3; d++) nrm[d] = -nrm[d];
    }
#pragma unroll
    for (int d = 0; d < 3; d++) q[d] = q[d] - nm * nrm[d];
  }
}

// node coordinates: the coordinate lines of a rectilinear grid, the full
// coordinates of a non-rectilinear one (NR)
template <bool NR, typename T>
__device__ __forceinline__ void node_pos(const GridDev<T>& g, int i, int j, int k, T x[3]) {
  if constexpr (NR) {
    const int l = gid(g, i, j, k);
    x[0] = g.gx[0][l]; x[1] = g.gx[1][l]; x[2] = g.gx[2][l];
  } else {
    x[0] = g.gl[0][i]; x[1] = g.gl[1][j]; x[2] = g.gl[2][k];
  }
}

// set_lp_bc (lp_tracking_m.f90:185-277); returns true for LP_STICK
template <bool NR = false, typename T>
__device__ __forceinline__ bool set_lp_bc(const GridDev<T>& g, int nd[3], T q[3], int& err) {
  const int i = iclamp(nd[0], 1, g.n[0]), j = iclamp(nd[1], 1, g.n[1]), k = iclamp(nd[2], 1, g.n[2]);
  nd[0] = i; nd[1] = j; nd[2] = k;
  const int f = g.flag[gid(g, i, j, k)];
  const bool inner = (i > 1 && j > 1 && k > 1 && i < g.n[0] && j < g.n[1] && k < g.n[2]);
  switch (f) {
    case kInternal:
    case k2D:
      return false;
    case kSlip:
    case kWall:
      if (inner) return false;
      // fall through
    case kMask: {
      const int p = pid(g, i, j, k);
      T nrm[3] = {g.bn[0][p], g.bn[1][p], g.bn[2][p]};
      T xn[3];
      node_pos<NR>(g, i, j, k, xn);
      T p2no[3] = {q[0] - xn[0], q[1] - xn[1], q[2] - xn[2]};
      T dist = p2no[0] * nrm[0] + p2no[1] * nrm[1] + p2no[2] * nrm[2];
      if (dist > T(0)) {
#pragma unroll
        for (int d = 0; d < 3; d++) q[d] = q[d] - nrm[d] * dist;
      }
      return false;
    }
    case kInflow:
    case kOutflow:
      if (inner) return false;
      node_pos<NR>(g, i, j, k, q);
      return true;
    default:
      err |= kErrBcFlag;
      return false;
  }
}

// One pass of the 27-neighbour nearest-node search with the 0.99
// hysteresis (lp_tracking_m.f90:72-109), in the reference's neighbour order.
// The three candidate coordinates per axis are loaded once; contraction is
// off so the squared distances round exactly as the CPU does.
template <bool NR = false, typename T>
__device__ __forceinline__ bool track_pass(const GridDev<T>& g, const T q[3], int nd[3]) {
#pragma clang fp contract(off)
  T sq[3][3];
  if constexpr (!NR) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
#pragma unroll
      for (int o = 0; o < 3; o++) {
        const T c = g.gl[d][iclamp(nd[d] + o - 1, 0, g.n[d] + 1)] - q[d];
        sq[d][o] = c * c;
      }
    }
  }
  const bool fx = g.n[0] == 1, fy = g.n[1] == 1, fz = g.n[2] == 1;
  const T huge = sizeof(T) == 8 ? T(DBL_MAX) : T(FLT_MAX);
  T rmin = huge;
  // the chosen neighbour as one code (di + 1) + 3 (dj + 1) + 9 (dk + 1): one
  // select per candidate instead of three
  int code = 13;
  constexpr int8_t nb[27][3] = {
      {0, 0, 0},   {-1, 0, 0}, {0, -1, 0}, {0, 0, -1},  {1, 0, 0},   {0, 1, 0},  {0, 0, 1},
      {-1, -1, -1}, {0, -1, -1}, {1, -1, -1}, {-1, 0, -1}, {1, 0, -1}, {-1, 1, -1}, {0, 1, -1},
      {1, 1, -1},  {-1, -1, 0}, {1, -1, 0},  {-1, 1, 0},  {1, 1, 0},   {-1, -1, 1}, {0, -1, 1},
      {1, -1, 1},  {-1, 0, 1},  {1, 0, 1},   {-1, 1, 1},  {0, 1, 1},   {1, 1, 1}};
#pragma unroll
  for (int b = 0; b < 27; b++) {
    const int di = nb[b][0], dj = nb[b][1], dk = nb[b][2];
    if ((fx && di) || (fy && dj) || (fz && dk)) continue;
    T rsq;
    if constexpr (NR) {
      // full coordinates of the neighbour (indices clamped to the ghost layer)
      const int l = gid(g, iclamp(nd[0] + di, 0, g.n[0] + 1), iclamp(nd[1] + dj, 0, g.n[1] + 1),
                        iclamp(nd[2] + dk, 0, g.n[2] + 1));
      const T a = g.gx[0][l] - q[0], b = g.gx[1][l] - q[1], c = g.gx[2][l] - q[2];
      rsq = (a * a + b * b) + c * c;
    } else {
      rsq = (sq[0][di + 1] + sq[1][dj + 1]) + sq[2][dk + 1];
    }
    if (rsq < T(0.99) * rmin) {
      rmin = rsq;
      code = (di + 1) + 3 * (dj + 1) + 9 * (dk + 1);
    }
  }
  const int ok = code / 9, r = code - 9 * ok, oj = r / 3, oi = r - 3 * oj;
  nd[0] += oi - 1; nd[1] += oj - 1; nd[2] += ok - 1;
  return code != 13;
}

// exchange_lpdata on one rank (comms_m.f90:872-1189): periodic wrap of a
// particle whose tracked node left [1,n].
template <typename T>
__device__ __forceinline__ void wrap(const GridDev<T>& g, int nd[3], T q[3], T sh[3] = nullptr) {
  int c[3];
#pragma unroll
  for (int d = 0; d < 3; d++) c[d] = nd[d] < 1 ? -1 : (nd[d] > g.n[d] ? 1 : 0);
  if (sh) sh[0] = sh[1] = sh[2] = T(0);
  if (!self_comm(g, c[0], c[1], c[2])) return;
  const int di = (c[0] + 1) + 3 * (c[1] + 1) + 9 * (c[2] + 1);
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (sh) sh[d] = g.ps[di][d];
    q[d] = q[d] + g.ps[di][d];
    int v = nd[d];
    if (v > g.n[d]) v -= g.n[d];
    if (v < 1) v += g.n[d];
    nd[d] = v;
  }
}

__device__ __forceinline__ void flush_status(DevStatus* st, int err, unsigned long long oob) {
  if (err) atomicOr(&st->err, err);
  if (oob) atomicAdd(&st->oob, oob);
}

// Eigenvalues (unordered) of a symmetric 3x3 matrix by cyclic Jacobi
// rotations in FP64 (the reference calls LAPACK DSYEV, lcs_m.f90:160-177,
// invariants_m.f90:178-193).
__device__ inline void eig_sym3(double a00, double a01, double a02, double a11, double a12, double a22,
                                double w[3]) {
  double a[3][3] = {{a00, a01, a02}, {a01, a11, a12}, {a02, a12, a22}};
  for (int sweep = 0; sweep < 60; sweep++) {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    const double dia = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
    if (off == 0.0 || off <= 1e-36 * dia) break;
#pragma unroll
    for (int pq = 0; pq < 3; pq++) {
      const int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2;
      const double apq = a[p][q];
      if (apq == 0.0) continue;
      const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
      double t;
      if (fabs(theta) > 1e150)
        t = 0.5 / theta;
      else
        t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
      const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
      for (int r = 0; r < 3; r++) {
        const double arp = a[r][p], arq = a[r][q];
        a[r][p] = c * arp - s * arq;
        a[r][q] = s * arp + c * arq;
      }
#pragma unroll
      for (int r = 0; r < 3; r++) {
        const double apr = a[p][r], aqr = a[q][r];
        a[p][r] = c * apr - s * aqr;
        a[q][r] = s * apr + c * aqr;
      }
      a[p][q] = 0.0;
      a[q][p] = 0.0;
    }
  }
  w[0] = a[0][0];
  w[1] = a[1][1];
  w[2] = a[2][2];
}

// largest eigenvalue
__device__ inline double eig_sym3_max(double a00, double a01, double a02, double a11, double a12, double a22) {
  double w[3];
  eig_sym3(a00, a01, a02, a11, a12, a22, w);
  return fmax(w[0], fmax(w[1], w[2]));
}


}  // namespace lcs
