// adjoint.hip -- reverse-mode derivative of the backward semi-Lagrangian FTLE
// pipeline (no counterpart in the reference).
//
// Forward map over one window (the BKWD diagnostic with h = T):
//   fm_0 = 0;  per sub-cycle  fm_new(x) = mask (I[fm_old](x_dep(x)) + x_dep - x)
//   with x_dep from the integrator's stages through the time-blended stage
//   velocities (flowmap_m.f90:17-104, lp_motion_m.f90:129-360), ghosts by the
//   flowmap BC, and FTLE = log(max(sqrt(lmax(G^T G)), tiny)) / T with G the
//   central-difference gradient of fm (lcs_m.f90:66-193).
// The kernels below run one sub-cycle backwards: they recompute the stage
// positions from the velocity, propagate the adjoint of x_dep back through
// the stages with the interpolants' Jacobians, and scatter the transposed
// interpolation weights into the adjoints of the old flowmap and of the stage
// velocity fields (global atomics).  Transposed banded solves carry
// coefficient-space adjoints back to grid values, and ghost folds carry ghost
// adjoints back to the interior nodes they were copied from.
#include "device_common.h"

namespace lcs {

// ---- B-spline basis values and first derivatives (same interval search and
// polynomial tables as bs_setup) ----
template <typename T, int K>
__device__ __forceinline__ int bs_setup_d(const T* __restrict__ t, const T* __restrict__ poly, int n, T x,
                                          bool xaxis, int hint, int& base, T v[K], T dv[K]) {
  const int L = n + K;
  if (!(x >= t[0]) || x > t[L - 1]) return 1;
  int left;
  if (x >= t[L - 1]) {
    if (!xaxis) return 2;
    left = n;
  } else {
    left = iclamp(hint, K, n);
    while (left < n && x >= t[left]) left++;
    while (left > K && x < t[left - 1]) left--;
  }
  const T* P = poly + left * (K * K + 1);
  const T ih = P[K * K];
  const T s = (x - t[left - 1]) * ih;
#pragma unroll
  for (int a = 0; a < K; a++) {
    T w = P[a * K + K - 1];
    T d = T(K - 1) * P[a * K + K - 1];
#pragma unroll
    for (int m = K - 2; m >= 0; m--) {
      w = w * s + P[a * K + m];
      if (m >= 1) d = d * s + T(m) * P[a * K + m];
    }
    v[a] = w;
    dv[a] = d * ih;
  }
  base = left - K;
  return 0;
}

// lin_setup plus d t_d / d q_d on the chosen bracket (zero where the local
// coordinate is clamped)
template <typename T>
__device__ __forceinline__ void lin_setup_d(const GridDev<T>& g, const T q[3], const int nd[3], LinW<T>& w,
                                            T dtq[3]) {
  int i0[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const T* xl = g.xl[d];
    int n0 = iclamp(nd[d], 0, g.n[d] + 1);
    int a = (q[d] > xl[n0]) ? n0 : n0 - 1;
    a = iclamp(a, 0, g.n[d]);
    const T x0 = xl[a], x1 = xl[a + 1];
    const T raw = (q[d] - x0) / (x1 - x0);
    const T tt = clamp01(raw);
    w.t[d] = tt;
    w.mt[d] = T(1) - tt;
    dtq[d] = (raw >= T(0) && raw <= T(1)) ? T(1) / (x1 - x0) : T(0);
    i0[d] = a;
  }
  w.base = i0[0] + g.N[0] * (i0[1] + g.N[1] * i0[2]);
}

// relaxed floating-point adds through address-space-qualified pointers, so
// that they compile to global_atomic_add / ds_add rather than flat atomics
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
using lptr = __attribute__((address_space(3))) T*;
template <typename T>
__device__ __forceinline__ void atomic_add(T* p, T v) {
  __hip_atomic_fetch_add((gptr<T>)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}


// ---- interpolants with value and Jacobian J[c][d] = d f_c / d q_d ----
// Each adjoint interpolant also exposes its per-axis support: K_d weights
// starting at coefficient base b_d, and a window of W_d cells starting at
// node + B_d that holds every support of a point within about half a cell of
// its node (the SL stage points under CFL <= 0.5).
template <typename T>
struct AdjNearest {
  static constexpr int KX = 1, KY = 1, KZ = 1, WX = 1, WY = 1, WZ = 1, BX = 0, BY = 0, BZ = 0;
  __device__ __forceinline__ void eval(const GridDev<T>& g, const T q[3], const int nd[3], const CF3<T>& f, T out[3],
                                       T J[3][3]) const {
    const int l = gid(g, iclamp(nd[0], 0, g.n[0] + 1), iclamp(nd[1], 0, g.n[1] + 1), iclamp(nd[2], 0, g.n[2] + 1));
#pragma unroll
    for (int c = 0; c < 3; c++) {
      out[c] = f.c[c][l];
      J[c][0] = J[c][1] = J[c][2] = T(0);
    }
  }
  __device__ __forceinline__ int weights(const GridDev<T>& g, const T q[3], const int nd[3], int b[3], T wx[1],
                                         T wy[1], T wz[1]) const {
    b[0] = iclamp(nd[0], 0, g.n[0] + 1);
    b[1] = iclamp(nd[1], 0, g.n[1] + 1);
    b[2] = iclamp(nd[2], 0, g.n[2] + 1);
    wx[0] = wy[0] = wz[0] = T(1);
    return 0;
  }
};

template <typename T>
struct AdjLinear {
  static constexpr int KX = 2, KY = 2, KZ = 2, WX = 3, WY = 3, WZ = 3, BX = -1, BY = -1, BZ = -1;
  __device__ __forceinline__ int weights(const GridDev<T>& g, const T q[3], const int nd[3], int b[3], T wx[2],
                                         T wy[2], T wz[2]) const {
    LinW<T> w;
    T dt[3];
    lin_setup_d(g, q, nd, w, dt);
    b[0] = w.base % g.N[0];
    b[1] = (w.base / g.N[0]) % g.N[1];
    b[2] = w.base / (g.N[0] * g.N[1]);
    wx[0] = w.mt[0]; wx[1] = w.t[0];
    wy[0] = w.mt[1]; wy[1] = w.t[1];
    wz[0] = w.mt[2]; wz[1] = w.t[2];
    return 0;
  }
  __device__ __forceinline__ void eval(const GridDev<T>& g, const T q[3], const int nd[3], const CF3<T>& f, T out[3],
                       T J[3][3]) const {
    LinW<T> w;
    T dt[3];
    lin_setup_d(g, q, nd, w, dt);
    const int b = w.base, sy = g.N[0], sz = g.N[0] * g.N[1];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const T* A = f.c[c];
      const T f0 = A[b], f1 = A[b + 1], f2 = A[b + sy], f3 = A[b + sy + 1], f4 = A[b + sz], f5 = A[b + sz + 1],
              f6 = A[b + sz + sy], f7 = A[b + sz + sy + 1];
      out[c] = trilin(w, f0, f1, f2, f3, f4, f5, f6, f7);
      // d/dt_x, d/dt_y, d/dt_z of the trilinear form
      const T ax = w.mt[1] * w.mt[2] * (f1 - f0) + w.t[1] * w.mt[2] * (f3 - f2) + w.mt[1] * w.t[2] * (f5 - f4) +
                   w.t[1] * w.t[2] * (f7 - f6);
      const T ay = w.mt[0] * w.mt[2] * (f2 - f0) + w.t[0] * w.mt[2] * (f3 - f1) + w.mt[0] * w.t[2] * (f6 - f4) +
                   w.t[0] * w.t[2] * (f7 - f5);
      const T az = w.mt[0] * w.mt[1] * (f4 - f0) + w.t[0] * w.mt[1] * (f5 - f1) + w.mt[0] * w.t[1] * (f6 - f2) +
                   w.t[0] * w.t[1] * (f7 - f3);
      J[c][0] = ax * dt[0];
      J[c][1] = ay * dt[1];
      J[c][2] = az * dt[2];
    }
  }
};

template <typename T, int KX_, int KZ_>
struct AdjSpline {
  static constexpr int KX = KX_, KY = KX_, KZ = KZ_;
  static constexpr int WX = KX + 1, WY = KX + 1, WZ = KZ + 1;
  static constexpr int BX = (KX % 2 == 0) ? -KX / 2 : -(KX + 1) / 2, BY = BX;
  static constexpr int BZ = (KZ % 2 == 0) ? -KZ / 2 : -(KZ + 1) / 2;
  struct W {
    int base;
    T wx[KX], wy[KX], wz[KZ], dx[KX], dy[KX], dz[KZ];
  };
  __device__ __forceinline__ int weights(const GridDev<T>& g, const T q[3], const int nd[3], int b[3], T wx[KX],
                                         T wy[KY], T wz[KZ]) const {
    W s;
    if (setup(g, q, nd, s)) return 1;
    b[0] = s.base % g.N[0];
    b[1] = (s.base / g.N[0]) % g.N[1];
    b[2] = s.base / (g.N[0] * g.N[1]);
#pragma unroll
    for (int a = 0; a < KX; a++) { wx[a] = s.wx[a]; wy[a] = s.wy[a]; }
#pragma unroll
    for (int a = 0; a < KZ; a++) wz[a] = s.wz[a];
    return 0;
  }
  __device__ __forceinline__ int setup(const GridDev<T>& g, const T q[3], const int nd[3], W& s) const {
    int bx, by, bz;
    const int rx = bs_setup_d<T, KX>(g.kn[0], g.bpoly[0], g.N[0], q[0], true, nd[0] + (KX + 2) / 2, bx, s.wx, s.dx);
    const int ry = bs_setup_d<T, KX>(g.kn[1], g.bpoly[1], g.N[1], q[1], false, nd[1] + (KX + 2) / 2, by, s.wy, s.dy);
    const int rz = bs_setup_d<T, KZ>(g.kn[2], g.bpoly[2], g.N[2], q[2], false, nd[2] + (KZ + 2) / 2, bz, s.wz, s.dz);
    if (rx || ry || rz) return 1;
    s.base = bx + g.N[0] * (by + g.N[1] * bz);
    return 0;
  }
  __device__ __forceinline__ void eval(const GridDev<T>& g, const T q[3], const int nd[3], const CF3<T>& f, T out[3],
                       T J[3][3]) const {
    W s;
    if (setup(g, q, nd, s)) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        out[c] = T(0);
        J[c][0] = J[c][1] = J[c][2] = T(0);
      }
      return;
    }
    const int sy = g.N[0], sz = g.N[0] * g.N[1];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      T v = 0, gx = 0, gy = 0, gz = 0;
#pragma unroll
      for (int cz = 0; cz < KZ; cz++) {
        T vy = 0, gxy = 0, gyy = 0;
#pragma unroll
        for (int by = 0; by < KX; by++) {
          T row[KX];
          load_row<T, KX>(f.c[c] + s.base + by * sy + cz * sz, row);
          T sx = 0, sdx = 0;
#pragma unroll
          for (int a = 0; a < KX; a++) {
            sx += s.wx[a] * row[a];
            sdx += s.dx[a] * row[a];
          }
          vy += s.wy[by] * sx;
          gxy += s.wy[by] * sdx;
          gyy += s.dy[by] * sx;
        }
        v += s.wz[cz] * vy;
        gx += s.wz[cz] * gxy;
        gy += s.wz[cz] * gyy;
        gz += s.dz[cz] * vy;
      }
      out[c] = v;
      J[c][0] = gx;
      J[c][1] = gy;
      J[c][2] = gz;
    }
  }
};

// dx_new(wrt q) of set_lp_bc: identity, a projection I - n n^T (MASK /
// boundary WALL / SLIP node with the particle outside), or zero (stick).
// Mirrors set_lp_bc's branches on the same node and position.
template <typename T>
__device__ __forceinline__ int lp_bc_kind(const GridDev<T>& g, const int nd0[3], const T q[3], T nrm[3]) {
  const int i = iclamp(nd0[0], 1, g.n[0]), j = iclamp(nd0[1], 1, g.n[1]), k = iclamp(nd0[2], 1, g.n[2]);
  const int f = g.flag[gid(g, i, j, k)];
  const bool inner = (i > 1 && j > 1 && k > 1 && i < g.n[0] && j < g.n[1] && k < g.n[2]);
  switch (f) {
    case kSlip:
    case kWall:
      if (inner) return 0;
      // fall through
    case kMask: {
      const int p = pid(g, i, j, k);
      nrm[0] = g.bn[0][p]; nrm[1] = g.bn[1][p]; nrm[2] = g.bn[2][p];
      const T dist = (q[0] - g.gl[0][i]) * nrm[0] + (q[1] - g.gl[1][j]) * nrm[1] + (q[2] - g.gl[2][k]) * nrm[2];
      return dist > T(0) ? 1 : 0;
    }
    case kInflow:
    case kOutflow:
      return inner ? 0 : 2;
    default:
      return 0;
  }
}

// ---------------------------------------------------------------------------
// Brick-tiled transposed scatter.  One wavefront owns a brick of BRX x BRY x
// BRZ nodes and an LDS tile that covers their windows.  Lane n scatters a
// point's K_x K_y K_z weights into the W_x W_y W_z window that starts at its
// own node + (B_x, B_y, B_z), with zeros outside the point's support; for a
// given window offset the 64 lanes then update 64 distinct LDS cells, so
// plain read-modify-writes in program order need no atomics (and sum
// deterministically; LDS float atomic adds measured four times slower here).
// A point whose support leaves its window adds to an overflow field with
// global atomics instead.  The finished tile reaches HBM
// either (tile store) as a plain store to the brick's own slot, which a
// gather pass then sums per cell, or (colours) with plain read-modify-writes
// of the fields, the bricks processed in 2 x 2 x 2 colours whose tiles never
// overlap, one launch per colour.
// ---------------------------------------------------------------------------
// cells per component of a tile stored as eight 4 x 4 x 4 octants
constexpr int kOct = 512;

struct AdjLaunch {
  int nbx, nby, nbz;  // bricks along each axis
  int cx, cy, cz;     // colour: brick parity along each axis
  int gx, gy;         // bricks of this colour along x and y
};

// A DPP quad permute of a value: lane q of each group of four receives the
// value of lane perm[q] (quad_perm control word)
template <int CTRL>
__device__ __forceinline__ float quad_perm(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double quad_perm(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// lane q receives from lane (q - d) & 3
template <int D, typename T>
__device__ __forceinline__ T from_minus(T v) {
  if constexpr (D == 1) return quad_perm<0x93>(v);  // [3,0,1,2]
  else if constexpr (D == 2) return quad_perm<0x4E>(v);  // [2,3,0,1]
  else return quad_perm<0x39>(v);  // [1,2,3,0]
}

// Transposed interpolation of one wave's brick into its LDS tile.  Every lane
// owns the window of WX x WY x WZ cells that starts at its node (zero weight
// outside its point's true support); a point whose support leaves the window
// goes to the overflow field with global atomics.  At one window offset the
// 64 lanes touch 64 distinct cells, and LDS performs one wave's operations in
// order, so the only ordering the code needs is between offsets, which a
// compiler barrier provides; each offset costs one LDS round trip.
// With 4-wide bricks (3-D) the four x-lanes of a brick row are a DPP quad:
// their WX offsets along x are folded by quad permutes into two cells per
// lane (q and q + 4 of the 8-cell tile row), so that one (oz, oy) offset row
// costs one round trip for all of x.
template <typename T, class A, int BRX, int BRY, int BRZ>
struct WindowScatter {
  static constexpr int TX = BRX + A::WX - 1, TY = BRY + A::WY - 1, TZ = BRZ + A::WZ - 1;
  static constexpr int TILE = TX * TY * TZ;
  static constexpr bool kQuad = BRX == 4 && TX <= 8;
  // tile-store layout: octants (4 x 4 x 4 bricks) or row-major, and the
  // values stored per brick and phase
  static constexpr bool kOctStore = BRX == 4 && BRY == 4 && BRZ == 4 && TX <= 8 && TY <= 8 && TZ <= 8;
  static constexpr int kStore = kOctStore ? 3 * kOct : 3 * TILE;
  // LDS layout of a component: row-major [TZ][TY][TX].  (A bank-swizzled
  // layout with a 128-cell z pitch, conflict-free for the quad accesses,
  // measured slower: the kernel is not bound by LDS bank conflicts.)
  static constexpr int LTILE = TILE;
  __device__ static __forceinline__ int lidx(int x, int y, int z) { return (z * TY + y) * TX + x; }
  lptr<T> tile;   // [3][TZ][TY][TX]
  int lx, ly, lz; // lane's node in the brick
  F3<T> ovf;
  int* oflag;  // set to oseq when a point overflows (tile-store mode)
  int oseq;
  int N0, N1;

  __device__ __forceinline__ void overflow(const int b[3], const T* wx, const T* wy, const T* wz,
                                           const T wv[3]) const {
#pragma unroll
    for (int c = 0; c < A::KZ; c++)
#pragma unroll
      for (int bb = 0; bb < A::KY; bb++)
#pragma unroll
        for (int a = 0; a < A::KX; a++) {
          const T w = wx[a] * wy[bb] * wz[c];
          const int idx = (b[0] + a) + N0 * ((b[1] + bb) + N1 * (b[2] + c));
#pragma unroll
          for (int m = 0; m < 3; m++) atomic_add(ovf.c[m] + idx, wv[m] * w);
        }
    if (oflag) __hip_atomic_store(oflag, oseq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }

  // NP points of one lane (use[p] false: no contribution); their
  // contributions are summed in registers before the tile is touched.  Every
  // lane of the wave must call this (the quad permutes read all lanes).
  template <int NP>
  __device__ __forceinline__ void multi(const A& ad, const GridDev<T>& g, const T (*qs)[3], const bool* use,
                                        const int nd[3], const T (*wvs)[3]) const {
    T ex[NP][A::WX], ey[NP][A::WY], ez[NP][A::WZ];
    bool any = false;
#pragma unroll
    for (int p = 0; p < NP; p++) {
      int b[3] = {0, 0, 0};
      T wx[A::KX] = {}, wy[A::KY] = {}, wz[A::KZ] = {};
      int sx = 0, sy = 0, sz = 0;
      bool ok = use[p] && !ad.weights(g, qs[p], nd, b, wx, wy, wz);
      if (ok) {
        sx = b[0] - (nd[0] + A::BX);
        sy = b[1] - (nd[1] + A::BY);
        sz = b[2] - (nd[2] + A::BZ);
        if (sx < 0 || sx > A::WX - A::KX || sy < 0 || sy > A::WY - A::KY || sz < 0 || sz > A::WZ - A::KZ) {
          overflow(b, wx, wy, wz, wvs[p]);
          ok = false;
        }
      }
      any = any || ok;
#pragma unroll
      for (int o = 0; o < A::WX; o++) {
        T v = T(0);
#pragma unroll
        for (int a = 0; a < A::KX; a++) v = (ok && o - a == sx) ? wx[a] : v;
        ex[p][o] = v;
      }
#pragma unroll
      for (int o = 0; o < A::WY; o++) {
        T v = T(0);
#pragma unroll
        for (int a = 0; a < A::KY; a++) v = (ok && o - a == sy) ? wy[a] : v;
        ey[p][o] = v;
      }
#pragma unroll
      for (int o = 0; o < A::WZ; o++) {
        T v = T(0);
#pragma unroll
        for (int a = 0; a < A::KZ; a++) v = (ok && o - a == sz) ? wz[a] : v;
        ez[p][o] = v;
      }
    }
    if constexpr (kQuad) {
      if (!__any(any)) return;  // wave-uniform
      const int q = lx;
      T mk[4], mh[4];
#pragma unroll
      for (int d = 0; d < 4; d++) {
        mk[d] = q >= d ? T(1) : T(0);
        mh[d] = T(1) - mk[d];
      }
      // window rows (oy, oz) where every lane's weight is zero add nothing:
      // the wave skips them (in smooth flow the nodes of a brick mostly sit
      // at the same offset in their windows, leaving 4 x 4 of the rows)
      bool ry[A::WY], rz[A::WZ];
#pragma unroll
      for (int o = 0; o < A::WY; o++) {
        bool nz = false;
#pragma unroll
        for (int p = 0; p < NP; p++) nz = nz || ey[p][o] != T(0);
        ry[o] = __any(nz);
      }
#pragma unroll
      for (int o = 0; o < A::WZ; o++) {
        bool nz = false;
#pragma unroll
        for (int p = 0; p < NP; p++) nz = nz || ez[p][o] != T(0);
        rz[o] = __any(nz);
      }
#pragma unroll
      for (int oz = 0; oz < A::WZ; oz++)
#pragma unroll
        for (int oy = 0; oy < A::WY; oy++) {
          if (!rz[oz] || !ry[oy]) continue;  // wave-uniform
          T lo[3], hi[3];
#pragma unroll
          for (int m = 0; m < 3; m++) {
            T v[A::WX];
#pragma unroll
            for (int o = 0; o < A::WX; o++) {
              T a = T(0);
#pragma unroll
              for (int p = 0; p < NP; p++) a += ex[p][o] * (ey[p][oy] * ez[p][oz] * wvs[p][m]);
              v[o] = a;
            }
            // lane q receives offset d of lane (q - d) mod 4: for q >= d
            // that is cell q (lo), otherwise cell q + 4 (hi) of the row; the
            // 0/1 factors mk[d] = (q >= d) select the cell without branches
            T l = v[0], h = T(0);
            if constexpr (A::WX > 4) h = v[4];
            if constexpr (A::WX > 1) {
              const T r = from_minus<1>(v[1]);
              l = fma(mk[1], r, l);
              h = fma(mh[1], r, h);
            }
            if constexpr (A::WX > 2) {
              const T r = from_minus<2>(v[2]);
              l = fma(mk[2], r, l);
              h = fma(mh[2], r, h);
            }
            if constexpr (A::WX > 3) {
              const T r = from_minus<3>(v[3]);
              l = fma(mk[3], r, l);
              h = fma(mh[3], r, h);
            }
            lo[m] = l;
            hi[m] = h;
          }
          const int rb = lidx(q, ly + oy, lz + oz);
          const bool has_hi = q + 4 < TX;
          T r0 = tile[rb], r1 = tile[LTILE + rb], r2 = tile[2 * LTILE + rb];
          T s0 = T(0), s1 = T(0), s2 = T(0);
          if (has_hi) {
            s0 = tile[rb + 4];
            s1 = tile[LTILE + rb + 4];
            s2 = tile[2 * LTILE + rb + 4];
          }
          tile[rb] = r0 + lo[0];
          tile[LTILE + rb] = r1 + lo[1];
          tile[2 * LTILE + rb] = r2 + lo[2];
          if (has_hi) {
            tile[rb + 4] = s0 + hi[0];
            tile[LTILE + rb + 4] = s1 + hi[1];
            tile[2 * LTILE + rb + 4] = s2 + hi[2];
          }
          __asm__ volatile("" ::: "memory");
        }
    } else {
      if (!any) return;
#pragma unroll
      for (int oz = 0; oz < A::WZ; oz++)
#pragma unroll
        for (int oy = 0; oy < A::WY; oy++) {
          T wyz[NP][3];
#pragma unroll
          for (int p = 0; p < NP; p++) {
            const T w = ey[p][oy] * ez[p][oz];
#pragma unroll
            for (int m = 0; m < 3; m++) wyz[p][m] = w * wvs[p][m];
          }
          const int row = lidx(lx, ly + oy, lz + oz);
#pragma unroll
          for (int ox = 0; ox < A::WX; ox++) {
            T add[3];
#pragma unroll
            for (int m = 0; m < 3; m++) {
              T a = T(0);
#pragma unroll
              for (int p = 0; p < NP; p++) a += ex[p][ox] * wyz[p][m];
              add[m] = a;
            }
            const int c = row + ox;
            const T r0 = tile[c], r1 = tile[LTILE + c], r2 = tile[2 * LTILE + c];
            tile[c] = r0 + add[0];
            tile[LTILE + c] = r1 + add[1];
            tile[2 * LTILE + c] = r2 + add[2];
            __asm__ volatile("" ::: "memory");
          }
        }
    }
  }
};

// stages of an integrator (the switch in k_sl_adjoint)
__device__ __forceinline__ int n_stages(int integ) {
  return (integ == kEuler || integ == kTrapezoidal) ? 1 : (integ == kRK2 ? 2 : (integ == kRK3 ? 3 : 4));
}

// One backward semi-Lagrangian sub-cycle in reverse for the bricks of one
// colour.  lam: adjoint of the new flowmap at the interior nodes (ghosted
// indexing).  Each lane recomputes its node's stages and propagates the
// adjoint of the departure point back through them; then, phase by phase,
// the wave scatters the transposed interpolation weights of one destination
// into its tile and flushes it.  Phase 0: the old flowmap (the field the
// forward kernel read: values or spline coefficients); phase 1 + b: stage
// velocity buffer b, which receives stage s >= 2 when sbuf[s] == b.  The
// first stage's nodal adjoint goes to dun / dunp1, split by ct0.
// TS (tile store): every brick of the grid in one launch, each writing its
// whole tile to its own slot tiles[phase][brick][component][TILE], which
// k_adj_gather then adds into the fields; otherwise the bricks of one colour
// read-modify-write their tiles into the fields directly.
template <typename T, class A, int BRX, int BRY, int BRZ, bool TS>
__global__ void __launch_bounds__(64) k_sl_adjoint(GridDev<T> g, AdjLaunch al, StageParams<T> sp, CF3<T> fmold,
                                                   CF3<T> lam, F3<T> adj_fm, F3<T> adj_v0, F3<T> adj_v1,
                                                   F3<T> adj_v2, F3<T> ovf_fm, F3<T> ovf_v0, F3<T> ovf_v1,
                                                   F3<T> ovf_v2, int4 sbuf, int nbuf, F3<T> dun, F3<T> dunp1,
                                                   TileStore<T> tsd) {
  using WS = WindowScatter<T, A, BRX, BRY, BRZ>;
  __shared__ T tile_mem[3 * WS::LTILE];
  const A ad{};
  // tile store: each XCD takes a contiguous run of bricks (workgroups are
  // dealt round-robin over the 8 XCDs), so bricks that share stencil rows
  // share that XCD's L2
  const int b = TS ? xcd_brick(blockIdx.x, gridDim.x) : blockIdx.x;
  int ibx, iby, ibz;
  if constexpr (TS) {
    ibx = b % al.nbx;
    iby = (b / al.nbx) % al.nby;
    ibz = b / (al.nbx * al.nby);
  } else {
    ibx = 2 * (b % al.gx) + al.cx;
    iby = 2 * ((b / al.gx) % al.gy) + al.cy;
    ibz = 2 * (b / (al.gx * al.gy)) + al.cz;
  }
  const int t = threadIdx.x;
  WS ws_;
  ws_.tile = (lptr<T>)tile_mem;
  ws_.lx = t % BRX;
  ws_.ly = (t / BRX) % BRY;
  ws_.lz = t / (BRX * BRY);
  ws_.N0 = g.N[0];
  ws_.N1 = g.N[1];
  ws_.oflag = TS ? tsd.flag : nullptr;
  ws_.oseq = tsd.seq;
  const int i = ibx * BRX + ws_.lx + 1, j = iby * BRY + ws_.ly + 1, k = ibz * BRZ + ws_.lz + 1;
  // tile origin (ghosted index of window cell 0 of the brick's first node)
  const int ox = ibx * BRX + 1 + A::BX, oy = iby * BRY + 1 + A::BY, oz = ibz * BRZ + 1 + A::BZ;

  bool active = i <= g.n[0] && j <= g.n[1] && k <= g.n[2];
  int l = 0;
  T gl3[3] = {0, 0, 0};
  if (active) {
    l = gid(g, i, j, k);
    const int f = g.flag[l];
    gl3[0] = lam.c[0][l]; gl3[1] = lam.c[1][l]; gl3[2] = lam.c[2][l];
    if (f == kMask || f == kInflow || f == kOutflow || (gl3[0] == T(0) && gl3[1] == T(0) && gl3[2] == T(0)))
      active = false;
  }
  const int nd[3] = {i, j, k};
  T xs[4][3] = {}, ws[4][3] = {}, xb[3] = {};
  int ns = 1;
  bool stuck = false;
  if (active) {
    const T x0[3] = {g.gl[0][i], g.gl[1][j], g.gl[2][k]};
    const T dt = sp.dt;
    const T HALF = T(1) / T(2), THIRD = T(1) / T(3), TWOTHIRD = T(2) / T(3), SIXTH = T(1) / T(6), TWO = T(2);
    // ---- recompute the stages (same expressions as integrate_inc(); the
    // departure point is x0 plus the summed increments) ----
    T Js[4][3][3], up[3], kk[3], x1[3], xp[3];
    CF3<T> vf[4];
#pragma unroll
    for (int a = 0; a < 4; a++) vf[a] = CF3<T>{{(const T*)sp.v[a][0], (const T*)sp.v[a][1], (const T*)sp.v[a][2]}};
    {
      const T w0 = T(1) - sp.ct0;
#pragma unroll
      for (int c = 0; c < 3; c++) up[c] = w0 * sp.un.c[c][l] + sp.ct0 * sp.unp1.c[c][l];
    }
    // Every integrator here has the form x_s = x0 + a_s dt u(x_{s-1}),
    // xp = x0 + sum_s b_s dt u(x_s) (RK2: a = 1/2, b = (0, 1); RK3: a =
    // (1/2, 2), b = (1/6, 2/3, 1/6); RK4: a = (1/2, 1/2, 1), b = (1/6, 1/3,
    // 1/3, 1/6); Euler and trapezoidal: one stage), with the increments
    // summed in the order integrate_inc() uses, so one loop holds every
    // stage evaluation.
    T a_[4] = {0, 0, 0, 0}, b_[4] = {0, 0, 0, 0};
    switch (sp.integ) {
      case kEuler:
      case kTrapezoidal:
        ns = 1; b_[0] = T(1);
        break;
      case kRK2:
        ns = 2; a_[1] = HALF; b_[1] = T(1);
        break;
      case kRK3:
        ns = 3; a_[1] = HALF; a_[2] = TWO; b_[0] = SIXTH; b_[1] = TWOTHIRD; b_[2] = SIXTH;
        break;
      case kRK4:
      default:
        ns = 4; a_[1] = HALF; a_[2] = HALF; a_[3] = T(1);
        b_[0] = SIXTH; b_[1] = THIRD; b_[2] = THIRD; b_[3] = SIXTH;
        break;
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
      kk[d] = dt * up[d];
      x1[d] = b_[0] * kk[d];
    }
#pragma unroll
    for (int st = 1; st < 4; st++) {
      if (st < ns) {  // the same for every lane
#pragma unroll
        for (int d = 0; d < 3; d++) xs[st][d] = x0[d] + a_[st] * kk[d];
        ad.eval(g, xs[st], nd, vf[st], up, Js[st]);
#pragma unroll
        for (int d = 0; d < 3; d++) {
          kk[d] = dt * up[d];
          x1[d] = x1[d] + b_[st] * kk[d];
        }
      }
    }
#pragma unroll
    for (int d = 0; d < 3; d++) xp[d] = x0[d] + x1[d];
    // ---- BC of the departure point ----
    int ndb[3] = {i, j, k};
    xb[0] = xp[0]; xb[1] = xp[1]; xb[2] = xp[2];
    project_flat(g, ndb, xb);
    T nrm[3] = {0, 0, 0};
    const int kind = lp_bc_kind(g, ndb, xb, nrm);
    int err = 0;
    if (set_lp_bc(g, ndb, xb, err)) {
      xb[0] = x0[0]; xb[1] = x0[1]; xb[2] = x0[2];
    }
    // ---- reverse: fm_new = F(xb) + xb - x0 ----
    T Fv[3], JF[3][3];
    ad.eval(g, xb, nd, fmold, Fv, JF);
    T xbar[3];
#pragma unroll
    for (int d = 0; d < 3; d++) xbar[d] = gl3[d] + JF[0][d] * gl3[0] + JF[1][d] * gl3[1] + JF[2][d] * gl3[2];
    stuck = kind == 2;
    if (kind == 1) {
      const T dn = xbar[0] * nrm[0] + xbar[1] * nrm[1] + xbar[2] * nrm[2];
#pragma unroll
      for (int d = 0; d < 3; d++) xbar[d] -= nrm[d] * dn;
    }
    // the flat axis by unrolled selects (a run-time index would put xbar in
    // scratch memory)
#pragma unroll
    for (int d = 0; d < 3; d++)
      if (d == g.flat) xbar[d] = T(0);
    T kb[4][3];
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int d = 0; d < 3; d++) kb[s][d] = stuck ? T(0) : b_[s] * xbar[d];
#pragma unroll
    for (int s = 3; s >= 1; s--) {
      if (s < ns) {
#pragma unroll
        for (int d = 0; d < 3; d++) ws[s][d] = dt * kb[s][d];
#pragma unroll
        for (int d = 0; d < 3; d++)
          kb[s - 1][d] += a_[s] * (Js[s][0][d] * ws[s][0] + Js[s][1][d] * ws[s][1] + Js[s][2][d] * ws[s][2]);
      }
    }
    const T w0 = T(1) - sp.ct0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const T kbar = dt * kb[0][c];
      dun.c[c][l] += w0 * kbar;
      dunp1.c[c][l] += sp.ct0 * kbar;
    }
  }
  // ---- scatter phases through the wave's tile ----
  for (int ph = 0; ph <= nbuf; ph++) {
    for (int e = t; e < 3 * WS::LTILE; e += 64) ws_.tile[e] = T(0);
    __syncthreads();
    const F3<T> dst = ph == 0 ? adj_fm : (ph == 1 ? adj_v0 : (ph == 2 ? adj_v1 : adj_v2));
    ws_.ovf = ph == 0 ? ovf_fm : (ph == 1 ? ovf_v0 : (ph == 2 ? ovf_v1 : ovf_v2));
    // the points of this phase, in one window pass: phase 0 the departure
    // point with the flowmap adjoint, phase 1 + b the stages (1..3) that
    // feed buffer b with their velocity adjoints; which points they are is
    // the same for every lane of the launch, and every lane calls the
    // scatter (inactive lanes contribute nothing)
    int sel[3] = {0, 0, 0}, np = 0;
    if (ph == 0) {
      np = 1;
    } else {
      const int nsu = n_stages(sp.integ);
#pragma unroll
      for (int st = 1; st < 4; st++) {
        const int sb_ = st == 1 ? sbuf.y : (st == 2 ? sbuf.z : sbuf.w);
        if (st < nsu && sb_ == ph - 1) {
          if (np == 0) sel[0] = st;
          else if (np == 1) sel[1] = st;
          else sel[2] = st;
          np++;
        }
      }
    }
    T q[3][3], w[3][3];
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
      for (int d = 0; d < 3; d++) {
        const int sp_ = sel[p];
        q[p][d] = sp_ == 0 ? xb[d] : (sp_ == 1 ? xs[1][d] : (sp_ == 2 ? xs[2][d] : xs[3][d]));
        w[p][d] = sp_ == 0 ? gl3[d] : (sp_ == 1 ? ws[1][d] : (sp_ == 2 ? ws[2][d] : ws[3][d]));
      }
    const bool on = ph == 0 ? active : (active && !stuck);
    const bool use[3] = {on, on, on};
    // (three points, which needs all stages to share one buffer, take a
    // two-point and a one-point pass: fewer instantiations, less code)
    if (np == 1) {
      ws_.template multi<1>(ad, g, q, use, nd, w);
    } else if (np >= 2) {
      ws_.template multi<2>(ad, g, q, use, nd, w);
      if (np == 3) ws_.template multi<1>(ad, g, q + 2, use + 2, nd, w + 2);
    }
    __syncthreads();
    if constexpr (TS) {
      // the whole tile, zeros included, to this brick's slot (plain stores)
      T* out = tsd.tiles + ((size_t)ph * al.nbx * al.nby * al.nbz + (size_t)b) * WS::kStore;
      if constexpr (WS::kOctStore) {
        // store e: component e / 512, octant (e / 64) % 8, lane position e % 64
        for (int e = t; e < 3 * kOct; e += 64) {
          const int m = e >> 9, o = (e >> 6) & 7;
          const int rx = (o & 1) * 4 + (t & 3), ry = ((o >> 1) & 1) * 4 + ((t >> 2) & 3), rz = (o >> 2) * 4 + (t >> 4);
          if (rx < WS::TX && ry < WS::TY && rz < WS::TZ)
            __builtin_nontemporal_store(ws_.tile[m * WS::LTILE + WS::lidx(rx, ry, rz)], out + e);
        }
      } else {
        for (int e = t; e < 3 * WS::TILE; e += 64) {
          const int m = e / WS::TILE, r = e - m * WS::TILE;
          const int rx = r % WS::TX, ry = (r / WS::TX) % WS::TY, rz = r / (WS::TX * WS::TY);
          __builtin_nontemporal_store(ws_.tile[m * WS::LTILE + WS::lidx(rx, ry, rz)], out + e);
        }
      }
      __syncthreads();
      continue;
    }
    // flush: batches of 8 independent read-modify-writes per lane, so that
    // the global loads of a batch are in flight together
    constexpr int kFlushBatch = 8;
    for (int e0 = t; e0 < 3 * WS::TILE; e0 += 64 * kFlushBatch) {
      T v[kFlushBatch], o[kFlushBatch];
      T* d[kFlushBatch];
#pragma unroll
      for (int q = 0; q < kFlushBatch; q++) {
        const int e = e0 + 64 * q;
        d[q] = nullptr;
        v[q] = T(0);
        if (e < 3 * WS::TILE) {
          const int m = e / WS::TILE, r = e - m * WS::TILE;
          const int rx = r % WS::TX, ry = (r / WS::TX) % WS::TY, rz = r / (WS::TX * WS::TY);
          v[q] = ws_.tile[m * WS::LTILE + WS::lidx(rx, ry, rz)];
          const int cx = ox + rx, cy = oy + ry, cz = oz + rz;
          if (v[q] != T(0) && cx >= 0 && cy >= 0 && cz >= 0 && cx < g.N[0] && cy < g.N[1] && cz < g.N[2])
            d[q] = (m == 0 ? dst.c[0] : (m == 1 ? dst.c[1] : dst.c[2])) + cx + g.N[0] * (cy + g.N[1] * cz);
        }
      }
#pragma unroll
      for (int q = 0; q < kFlushBatch; q++) o[q] = d[q] ? *d[q] : T(0);
#pragma unroll
      for (int q = 0; q < kFlushBatch; q++)
        if (d[q]) *d[q] = o[q] + v[q];
    }
    __syncthreads();
  }
}

// Destinations of the tile-store gather.  Phase 0 is the adjoint of the old
// flowmap (assigned); phase 1 + b is stage velocity buffer b, the blend
// (1 - c[b]) u_n + c[b] u_np1, whose adjoint goes straight to dn / dnp1 (the
// transpose of the blend).  The overflow fields are added and cleared when
// the scatter flagged an overflow (*flag == seq); otherwise they are zero.
template <typename T>
struct GatherOut {
  F3<T> fm, dn, dnp1;
  F3<T> ovf[4];
  T c[3];
  const int* flag;
  int seq;
};

template <typename T>
__device__ __forceinline__ void gather_store(const GatherOut<T>& o, long long l, int nph, T (&sum)[4][3]) {
  const bool ovf_on = *o.flag == o.seq;
#pragma unroll
  for (int ph = 0; ph < 4; ph++)
    if (ovf_on && ph < nph) {
#pragma unroll
      for (int m = 0; m < 3; m++) {
        const T ov = o.ovf[ph].c[m][l];
        if (ov != T(0)) {
          sum[ph][m] += ov;
          o.ovf[ph].c[m][l] = T(0);
        }
      }
    }
#pragma unroll
  for (int m = 0; m < 3; m++) o.fm.c[m][l] = sum[0][m];
  if (nph > 1) {
#pragma unroll
    for (int m = 0; m < 3; m++) {
      T an = T(0), anp1 = T(0);
#pragma unroll
      for (int b = 0; b < 3; b++)
        if (b + 1 < nph) {
          an += (T(1) - o.c[b]) * sum[b + 1][m];
          anp1 += o.c[b] * sum[b + 1][m];
        }
      if (an != T(0)) o.dn.c[m][l] += an;
      if (anp1 != T(0)) o.dnp1.c[m][l] += anp1;
    }
  }
}

// Tile-store gather, any brick shape: thread per ghosted cell; it adds the
// tiles of the bricks whose tile covers the cell (brick b's tile starts at
// b BR + 1 + B along each axis and spans TL cells, so at most ceil(TL / BR)
// bricks per axis).  Tiles are stored row-major [3][TZ][TY][TX].
__device__ __forceinline__ int floordiv(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }
template <typename T, int BRX, int BRY, int BRZ, int TX, int TY, int TZ, int BX, int BY, int BZ>
__global__ void __launch_bounds__(kBlock) k_adj_gather(GridDev<T> g, AdjLaunch al, const T* __restrict__ tiles,
                                                       int nph, GatherOut<T> out) {
  constexpr int TILE = TX * TY * TZ;
  constexpr int MX = (TX + BRX - 1) / BRX, MY = (TY + BRY - 1) / BRY, MZ = (TZ + BRZ - 1) / BRZ;
  const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= g.nt) return;
  const int cx = (int)(l % g.N[0]), cy = (int)((l / g.N[0]) % g.N[1]), cz = (int)(l / ((long long)g.N[0] * g.N[1]));
  // b BR + 1 + B <= c < b BR + 1 + B + TL
  const int bx0 = floordiv(cx - BX - TX + BRX - 1, BRX), bx1 = min(floordiv(cx - 1 - BX, BRX), al.nbx - 1);
  const int by0 = floordiv(cy - BY - TY + BRY - 1, BRY), by1 = min(floordiv(cy - 1 - BY, BRY), al.nby - 1);
  const int bz0 = floordiv(cz - BZ - TZ + BRZ - 1, BRZ), bz1 = min(floordiv(cz - 1 - BZ, BRZ), al.nbz - 1);
  const size_t nbricks = (size_t)al.nbx * al.nby * al.nbz;
  T sum[4][3] = {};
#pragma unroll
  for (int ph = 0; ph < 4; ph++)
    if (ph < nph) {
      const T* tph = tiles + (size_t)ph * nbricks * 3 * TILE;
#pragma unroll
      for (int az = 0; az < MZ; az++)
#pragma unroll
        for (int ay = 0; ay < MY; ay++)
#pragma unroll
          for (int ax = 0; ax < MX; ax++) {
            const int bx = bx0 + ax, by = by0 + ay, bz = bz0 + az;
            if (bx >= 0 && by >= 0 && bz >= 0 && bx <= bx1 && by <= by1 && bz <= bz1) {
              const int rx = cx - (bx * BRX + 1 + BX), ry = cy - (by * BRY + 1 + BY), rz = cz - (bz * BRZ + 1 + BZ);
              const T* tl =
                  tph + (bx + (size_t)al.nbx * (by + (size_t)al.nby * bz)) * 3 * TILE + (rz * TY + ry) * TX + rx;
#pragma unroll
              for (int m = 0; m < 3; m++) sum[ph][m] += tl[m * TILE];
            }
          }
    }
  gather_store(out, l, nph, sum);
}

// Tile-store gather for 4 x 4 x 4 bricks whose tiles fit 8 x 8 x 8 cells,
// stored as eight 4 x 4 x 4 octants per component ([3][oz][oy][ox][64]).
// Wave w handles the 4 x 4 x 4 block of cells C, c = 4 C + 1 + B + lane
// offset along each axis: brick C covers it with octant 0 and brick C - 1
// with octant 1 along that axis, at the lane's own position in the octant,
// so every wave reads each (brick, component) as one contiguous 256-byte run.
template <typename T, int TX, int TY, int TZ, int BX, int BY, int BZ>
__global__ void __launch_bounds__(256) k_adj_gather_oct(GridDev<T> g, AdjLaunch al, const T* __restrict__ tiles,
                                                        int nph, GatherOut<T> out) {
  static_assert(TX <= 8 && TY <= 8 && TZ <= 8, "octant tiles span at most 8 cells per axis");
  const int lane = threadIdx.x & 63;
  const long long cb = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int C0x = floordiv(-1 - BX, 4), C0y = floordiv(-1 - BY, 4), C0z = floordiv(-1 - BZ, 4);
  const int ncx = floordiv(g.N[0] - 2 - BX, 4) - C0x + 1, ncy = floordiv(g.N[1] - 2 - BY, 4) - C0y + 1,
            ncz = floordiv(g.N[2] - 2 - BZ, 4) - C0z + 1;
  if (cb >= (long long)ncx * ncy * ncz) return;
  const int Cx = C0x + (int)(cb % ncx), Cy = C0y + (int)((cb / ncx) % ncy), Cz = C0z + (int)(cb / ((long long)ncx * ncy));
  const int lx = lane & 3, ly = (lane >> 2) & 3, lz = lane >> 4;
  const int cx = 4 * Cx + 1 + BX + lx, cy = 4 * Cy + 1 + BY + ly, cz = 4 * Cz + 1 + BZ + lz;
  if (cx < 0 || cy < 0 || cz < 0 || cx >= g.N[0] || cy >= g.N[1] || cz >= g.N[2]) return;
  const long long l = cx + (long long)g.N[0] * (cy + (long long)g.N[1] * cz);
  const size_t nbricks = (size_t)al.nbx * al.nby * al.nbz;
  // the eight (brick, octant) slots: every load is issued (a slot outside
  // the grid of bricks reads brick 0, a padded cell reads its unused slot)
  // and the ones that do not cover the cell are dropped by a select
  size_t off[8];
  bool ok[8];
#pragma unroll
  for (int o = 0; o < 8; o++) {
    const int sx = o & 1, sy = (o >> 1) & 1, sz = o >> 2;
    const int bx = Cx - sx, by = Cy - sy, bz = Cz - sz;
    ok[o] = bx >= 0 && by >= 0 && bz >= 0 && bx < al.nbx && by < al.nby && bz < al.nbz && lx + 4 * sx < TX &&
            ly + 4 * sy < TY && lz + 4 * sz < TZ;
    off[o] = (ok[o] ? (bx + (size_t)al.nbx * (by + (size_t)al.nby * bz)) * 3 * kOct : 0) + o * 64 + lane;
  }
  T sum[4][3] = {};
#pragma unroll
  for (int ph = 0; ph < 4; ph++)
    if (ph < nph) {
      const T* tph = tiles + (size_t)ph * nbricks * 3 * kOct;
      T v[8][3];
#pragma unroll
      for (int o = 0; o < 8; o++)
#pragma unroll
        for (int m = 0; m < 3; m++) v[o][m] = __builtin_nontemporal_load(tph + off[o] + m * kOct);
#pragma unroll
      for (int o = 0; o < 8; o++)
#pragma unroll
        for (int m = 0; m < 3; m++) sum[ph][m] += ok[o] ? v[o][m] : T(0);
    }
  gather_store(out, l, nph, sum);
}

// dst += ovf, ovf = 0
template <typename T>
__global__ void __launch_bounds__(kBlock) k_merge_overflow(F3<T> dst, F3<T> ovf, long long nt) {
  for (long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x; l < nt; l += (long long)gridDim.x * blockDim.x) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const T v = ovf.c[c][l];
      if (v != T(0)) {
        dst.c[c][l] += v;
        ovf.c[c][l] = T(0);
      }
    }
  }
}

// adj_n += (1-ct) b, adj_np1 += ct b, b = 0 (transpose of the stage blend)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_adj_unblend(F3<T> b, T ct, F3<T> an, F3<T> anp1, long long nt) {
  const T w0 = T(1) - ct;
  for (long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x; l < nt; l += (long long)gridDim.x * blockDim.x) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const T v = b.c[c][l];
      if (v != T(0)) {
        an.c[c][l] += w0 * v;
        anp1.c[c][l] += ct * v;
        b.c[c][l] = T(0);
      }
    }
  }
}

// Ghost folds: every ghost whose forward value was copied from an interior
// node hands its adjoint back to that node; ghosts that were set to zero (or
// held an earlier value) drop theirs.  mode 0: velocity rules (k_vel_ghost),
// 1: flowmap rules (k_fm_ghost / k_field_from_origin).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_fold_ghost(GridDev<T> g, F3<T> adj, int mode) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int i, j, k;
  if (t >= ghost_count(g) || !decode_ghost(g, t, i, j, k)) return;
  const int l = gid(g, i, j, k);
  const int f = g.flag[l];
  const int ei = i < 1 ? -1 : (i > g.n[0] ? 1 : 0), ej = j < 1 ? -1 : (j > g.n[1] ? 1 : 0),
            ek = k < 1 ? -1 : (k > g.n[2] ? 1 : 0);
  int s = -1;
  if (f == kInternal) {
    if (self_comm(g, ei, ej, ek))
      s = gid(g, ei < 0 ? g.n[0] : (ei > 0 ? 1 : i), ej < 0 ? g.n[1] : (ej > 0 ? 1 : j),
              ek < 0 ? g.n[2] : (ek > 0 ? 1 : k));
  } else if (f == kSlip || f == k2D || (mode == 0 && (f == kInflow || f == kOutflow))) {
    s = gid(g, iclamp(i, 1, g.n[0]), iclamp(j, 1, g.n[1]), iclamp(k, 1, g.n[2]));
  }
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const T v = adj.c[c][l];
    adj.c[c][l] = T(0);
    if (s >= 0 && v != T(0)) atomic_add(adj.c[c] + s, v);
  }
}

// interior nodes whose forward value is forced to zero: MASK for the velocity
// (k_cfl), MASK / INFLOW / OUTFLOW for a flowmap built from origins
template <typename T>
__global__ void __launch_bounds__(kBlock) k_zero_forced(GridDev<T> g, F3<T> adj, int mode) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.np) return;
  int i, j, k;
  decode_interior(g, p, i, j, k);
  const int l = gid(g, i, j, k);
  const int f = g.flag[l];
  const bool z = mode == 0 ? (f == kMask) : (f == kMask || f == kInflow || f == kOutflow);
  if (z) {
#pragma unroll
    for (int c = 0; c < 3; c++) adj.c[c][l] = T(0);
  }
}

// eigenvector of the largest eigenvalue lam of symmetric C (FP64): the
// largest cross product of two rows of C - lam I
__device__ inline void eigvec3(const double C[3][3], double lam, double v[3]) {
  const double r[3][3] = {{C[0][0] - lam, C[0][1], C[0][2]}, {C[1][0], C[1][1] - lam, C[1][2]},
                          {C[2][0], C[2][1], C[2][2] - lam}};
  double best = -1.0;
  v[0] = 1; v[1] = 0; v[2] = 0;
  for (int a = 0; a < 3; a++)
    for (int b = a + 1; b < 3; b++) {
      const double c0 = r[a][1] * r[b][2] - r[a][2] * r[b][1];
      const double c1 = r[a][2] * r[b][0] - r[a][0] * r[b][2];
      const double c2 = r[a][0] * r[b][1] - r[a][1] * r[b][0];
      const double nn = c0 * c0 + c1 * c1 + c2 * c2;
      if (nn > best) {
        best = nn;
        v[0] = c0; v[1] = c1; v[2] = c2;
      }
    }
  const double nn = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (nn > 0) {
    v[0] /= nn; v[1] /= nn; v[2] /= nn;
  } else {
    v[0] = 1; v[1] = 0; v[2] = 0;
  }
}

// Adjoint of k_ftle: dL/dftle -> dL/dfm (ghosted, atomic scatter to the six
// stencil neighbours).  ftle = log(max(sqrt(lam), tiny))/T, lam = lmax(G^T G):
// d ftle / d lam = 1/(2 T lam), d lam / d G = 2 G v v^T.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_ftle_adjoint(GridDev<T> g, CF3<T> fm, double Tint, int incomp,
                                                         double tiny, const double* __restrict__ dftle,
                                                         F3<T> adj) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.np) return;
  const double fb = dftle[p];
  if (fb == 0.0) return;
  int i, j, k;
  decode_interior(g, p, i, j, k);
  const int ip = (i == g.n[0] && g.per[0]) ? gid(g, 1, j, k) : gid(g, i + 1, j, k);
  const int im = (i == 1 && g.per[0]) ? gid(g, g.n[0], j, k) : gid(g, i - 1, j, k);
  const int jp = (j == g.n[1] && g.per[1]) ? gid(g, i, 1, k) : gid(g, i, j + 1, k);
  const int jm = (j == 1 && g.per[1]) ? gid(g, i, g.n[1], k) : gid(g, i, j - 1, k);
  const int kp = (k == g.n[2] && g.per[2]) ? gid(g, i, j, 1) : gid(g, i, j, k + 1);
  const int km = (k == 1 && g.per[2]) ? gid(g, i, j, g.n[2]) : gid(g, i, j, k - 1);
  const double h[3] = {(double)g.gl[0][i + 1] - (double)g.gl[0][i - 1],
                       (double)g.gl[1][j + 1] - (double)g.gl[1][j - 1],
                       (double)g.gl[2][k + 1] - (double)g.gl[2][k - 1]};
  double G[3][3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const T* F = fm.c[c];
    G[c][0] = ((double)F[ip] - (double)F[im]) / h[0];
    G[c][1] = ((double)F[jp] - (double)F[jm]) / h[1];
    G[c][2] = ((double)F[kp] - (double)F[km]) / h[2];
  }
  bool live[3] = {g.n[0] != 1, g.n[1] != 1, g.n[2] != 1};
#pragma unroll
  for (int d = 0; d < 3; d++)
    if (!live[d]) G[0][d] = G[1][d] = G[2][d] = 0;
  double C[3][3];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) C[a][b] = G[0][a] * G[0][b] + G[1][a] * G[1][b] + G[2][a] * G[2][b];
  const double lam = eig_sym3_max(C[0][0], C[0][1], C[0][2], C[1][1], C[1][2], C[2][2]);
  if (!(sqrt(lam) > tiny)) return;
  const double r = log(sqrt(lam)) / Tint;
  if (incomp && r < 0.0) return;
  double v[3];
  eigvec3(C, lam, v);
  const double lb = fb / (2.0 * Tint * lam);
  const int pl[3] = {ip, jp, kp}, mi[3] = {im, jm, km};
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const double gv = G[c][0] * v[0] + G[c][1] * v[1] + G[c][2] * v[2];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      if (!live[d]) continue;
      const double gb = 2.0 * lb * gv * v[d] / h[d];
      atomic_add(adj.c[c] + pl[d], (T)gb);
      atomic_add(adj.c[c] + mi[d], (T)(-gb));
    }
  }
}

// out[s*np + p] (+)= adj interior (ghosted index), user layout
template <typename T, typename U>
__global__ void __launch_bounds__(kBlock) k_adj_to_user(GridDev<T> g, CF3<T> adj, U* a, U* b, U* c, int accumulate) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.np) return;
  int i, j, k;
  decode_interior(g, p, i, j, k);
  const int l = gid(g, i, j, k);
  if (accumulate) {
    a[p] += (U)adj.c[0][l];
    b[p] += (U)adj.c[1][l];
    c[p] += (U)adj.c[2][l];
  } else {
    a[p] = (U)adj.c[0][l];
    b[p] = (U)adj.c[1][l];
    c[p] = (U)adj.c[2][l];
  }
}

// interior values of a field (ghosted index) to origin order (np)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_interior(GridDev<T> g, CF3<T> src, F3<T> dst) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.np) return;
  int i, j, k;
  decode_interior(g, p, i, j, k);
  const int l = gid(g, i, j, k);
#pragma unroll
  for (int c = 0; c < 3; c++) dst.c[c][p] = src.c[c][l];
}

// ===========================================================================
// launchers
// ===========================================================================
static inline dim3 agrid(long long n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

template <typename T, class A, int BRX, int BRY, int BRZ>
static size_t tile_store_bytes(const GridDev<T>& g, int nbuf) {
  using WS = WindowScatter<T, A, BRX, BRY, BRZ>;
  const size_t nbricks = (size_t)((g.n[0] + BRX - 1) / BRX) * ((g.n[1] + BRY - 1) / BRY) * ((g.n[2] + BRZ - 1) / BRZ);
  return (size_t)(nbuf + 1) * nbricks * WS::kStore * sizeof(T);
}

template <typename T, class A, int BRX, int BRY, int BRZ>
static bool sl_adjoint_colours(const GridDev<T>& g, const StageParams<T>& sp, CF3<T> fmold, CF3<T> lam,
                               F3<T> adj_fm, F3<T> const adj_v[3], F3<T> const ovf[4], const int sbuf[4], int nbuf,
                               F3<T> dun, F3<T> dunp1, const TileStore<T>* ts, const T bct[3], F3<T> dcn,
                               F3<T> dcnp1, hipStream_t s) {
  AdjLaunch al;
  al.nbx = (g.n[0] + BRX - 1) / BRX;
  al.nby = (g.n[1] + BRY - 1) / BRY;
  al.nbz = (g.n[2] + BRZ - 1) / BRZ;
  const int4 sb = make_int4(sbuf[0], sbuf[1], sbuf[2], sbuf[3]);
  if (ts) {
    // one launch over every brick, then one gather pass (which also merges
    // the overflow fields)
    using WS = WindowScatter<T, A, BRX, BRY, BRZ>;
    al.cx = al.cy = al.cz = 0;
    al.gx = al.nbx;
    al.gy = al.nby;
    hipLaunchKernelGGL((k_sl_adjoint<T, A, BRX, BRY, BRZ, true>), dim3((unsigned)(al.nbx * al.nby * al.nbz)),
                       dim3(64), 0, s, g, al, sp, fmold, lam, adj_fm, adj_v[0], adj_v[1], adj_v[2], ovf[0], ovf[1],
                       ovf[2], ovf[3], sb, nbuf, dun, dunp1, *ts);
    GatherOut<T> go;
    go.fm = adj_fm;
    go.dn = dcn;
    go.dnp1 = dcnp1;
    for (int p = 0; p < 4; p++) go.ovf[p] = ovf[p];
    for (int b = 0; b < 3; b++) go.c[b] = bct[b];
    go.flag = ts->flag;
    go.seq = ts->seq;
    if constexpr (WS::kOctStore) {
      auto cblocks = [](int N, int B) {
        auto fd = [](int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); };
        return (long long)(fd(N - 2 - B, 4) - fd(-1 - B, 4) + 1);
      };
      const long long ncb = cblocks(g.N[0], A::BX) * cblocks(g.N[1], A::BY) * cblocks(g.N[2], A::BZ);
      hipLaunchKernelGGL((k_adj_gather_oct<T, WS::TX, WS::TY, WS::TZ, A::BX, A::BY, A::BZ>),
                         dim3((unsigned)((ncb + 3) / 4)), dim3(256), 0, s, g, al, (const T*)ts->tiles, nbuf + 1,
                         go);
    } else {
      hipLaunchKernelGGL((k_adj_gather<T, BRX, BRY, BRZ, WS::TX, WS::TY, WS::TZ, A::BX, A::BY, A::BZ>), agrid(g.nt),
                         dim3(kBlock), 0, s, g, al, (const T*)ts->tiles, nbuf + 1, go);
    }
    return true;
  }
  for (int c = 0; c < 8; c++) {
    al.cx = c & 1; al.cy = (c >> 1) & 1; al.cz = c >> 2;
    al.gx = (al.nbx - al.cx + 1) / 2;
    al.gy = (al.nby - al.cy + 1) / 2;
    const int gz = (al.nbz - al.cz + 1) / 2;
    if (al.gx <= 0 || al.gy <= 0 || gz <= 0) continue;
    hipLaunchKernelGGL((k_sl_adjoint<T, A, BRX, BRY, BRZ, false>), dim3((unsigned)(al.gx * al.gy * gz)), dim3(64), 0,
                       s, g, al, sp, fmold, lam, adj_fm, adj_v[0], adj_v[1], adj_v[2], ovf[0], ovf[1], ovf[2], ovf[3],
                       sb, nbuf, dun, dunp1, TileStore<T>{nullptr, nullptr, 0});
  }
  const long long nb = std::min<long long>((g.nt + kBlock - 1) / kBlock, 4096);
  F3<T> dsts[4] = {adj_fm, adj_v[0], adj_v[1], adj_v[2]};
  for (int ph = 0; ph <= nbuf; ph++)
    hipLaunchKernelGGL((k_merge_overflow<T>), dim3((unsigned)nb), dim3(kBlock), 0, s, dsts[ph], ovf[ph], g.nt);
  return false;
}

template <typename T>
bool launch_sl_adjoint(const GridDev<T>& g, int interp, const StageParams<T>& sp, CF3<T> fmold, CF3<T> lam,
                       F3<T> adj_fm, F3<T> const adj_v[3], F3<T> const ovf[4], const int sbuf[4], int nbuf,
                       F3<T> dun, F3<T> dunp1, const TileStore<T>* ts, const T bct[3], F3<T> dcn, F3<T> dcnp1,
                       hipStream_t s) {
  const bool flat = g.n[2] == 1;
  auto go = [&](auto tag) -> bool {
    using A = decltype(tag);
    if (flat)
      return sl_adjoint_colours<T, A, 8, 8, 1>(g, sp, fmold, lam, adj_fm, adj_v, ovf, sbuf, nbuf, dun, dunp1, ts,
                                               bct, dcn, dcnp1, s);
    return sl_adjoint_colours<T, A, 4, 4, 4>(g, sp, fmold, lam, adj_fm, adj_v, ovf, sbuf, nbuf, dun, dunp1, ts,
                                             bct, dcn, dcnp1, s);
  };
  if (interp == kNearest) return go(AdjNearest<T>{});
  if (interp == kLinear) return go(AdjLinear<T>{});
  const int kx = g.kord[0], kz = g.kord[2];
  if (kx == 4 && kz == 4) return go(AdjSpline<T, 4, 4>{});
  if (kx == 4 && kz == 2) return go(AdjSpline<T, 4, 2>{});
  if (kx == 3 && kz == 3) return go(AdjSpline<T, 3, 3>{});
  return go(AdjSpline<T, 3, 2>{});
}

template <typename T>
size_t sl_adjoint_tile_bytes(const GridDev<T>& g, int interp, int nbuf) {
  const bool flat = g.n[2] == 1;
  auto go = [&](auto tag) -> size_t {
    using A = decltype(tag);
    return flat ? tile_store_bytes<T, A, 8, 8, 1>(g, nbuf) : tile_store_bytes<T, A, 4, 4, 4>(g, nbuf);
  };
  if (interp == kNearest) return go(AdjNearest<T>{});
  if (interp == kLinear) return go(AdjLinear<T>{});
  const int kx = g.kord[0], kz = g.kord[2];
  if (kx == 4 && kz == 4) return go(AdjSpline<T, 4, 4>{});
  if (kx == 4 && kz == 2) return go(AdjSpline<T, 4, 2>{});
  if (kx == 3 && kz == 3) return go(AdjSpline<T, 3, 3>{});
  return go(AdjSpline<T, 3, 2>{});
}

template <typename T>
void launch_adj_unblend(const GridDev<T>& g, F3<T> b, T ct, F3<T> an, F3<T> anp1, hipStream_t s) {
  const long long nb = std::min<long long>((g.nt + kBlock - 1) / kBlock, 8192);
  hipLaunchKernelGGL((k_adj_unblend<T>), dim3((unsigned)nb), dim3(kBlock), 0, s, b, ct, an, anp1, g.nt);
}

template <typename T>
void launch_fold_ghost(const GridDev<T>& g, F3<T> adj, int mode, hipStream_t s) {
  hipLaunchKernelGGL((k_zero_forced<T>), agrid(g.np), dim3(kBlock), 0, s, g, adj, mode);
  const long long ng = 2LL * g.N[0] * g.N[1] + (2LL * g.N[0] + 2LL * g.n[1]) * g.n[2];
  hipLaunchKernelGGL((k_fold_ghost<T>), agrid(ng), dim3(kBlock), 0, s, g, adj, mode);
}

template <typename T>
void launch_ftle_adjoint(const GridDev<T>& g, CF3<T> fm, double Tint, int incomp, double tiny, const double* dftle,
                         F3<T> adj, hipStream_t s) {
  hipLaunchKernelGGL((k_ftle_adjoint<T>), agrid(g.np), dim3(kBlock), 0, s, g, fm, Tint, incomp, tiny, dftle, adj);
}

template <typename T, typename U>
void launch_adj_to_user(const GridDev<T>& g, CF3<T> adj, U* a, U* b, U* c, int accumulate, hipStream_t s) {
  hipLaunchKernelGGL((k_adj_to_user<T, U>), agrid(g.np), dim3(kBlock), 0, s, g, adj, a, b, c, accumulate);
}

template <typename T>
void launch_interior(const GridDev<T>& g, CF3<T> src, F3<T> dst, hipStream_t s) {
  hipLaunchKernelGGL((k_interior<T>), agrid(g.np), dim3(kBlock), 0, s, g, src, dst);
}

#define INST_ADJ(T, U)                                                                                          \
  template bool launch_sl_adjoint<T>(const GridDev<T>&, int, const StageParams<T>&, CF3<T>, CF3<T>, F3<T>,      \
                                     F3<T> const*, F3<T> const*, const int*, int, F3<T>, F3<T>,                 \
                                     const TileStore<T>*, const T*,                                             \
                                     F3<T>, F3<T>, hipStream_t);                                                \
  template size_t sl_adjoint_tile_bytes<T>(const GridDev<T>&, int, int);                                        \
  template void launch_adj_unblend<T>(const GridDev<T>&, F3<T>, T, F3<T>, F3<T>, hipStream_t);                  \
  template void launch_fold_ghost<T>(const GridDev<T>&, F3<T>, int, hipStream_t);                               \
  template void launch_ftle_adjoint<T>(const GridDev<T>&, CF3<T>, double, int, double, const double*, F3<T>,    \
                                       hipStream_t);                                                            \
  template void launch_adj_to_user<T, U>(const GridDev<T>&, CF3<T>, U*, U*, U*, int, hipStream_t);              \
  template void launch_interior<T>(const GridDev<T>&, CF3<T>, F3<T>, hipStream_t);

INST_ADJ(float, float)
INST_ADJ(double, double)
template void launch_adj_to_user<float, double>(const GridDev<float>&, CF3<float>, double*, double*, double*, int,
                                                hipStream_t);
template void launch_adj_to_user<double, float>(const GridDev<double>&, CF3<double>, float*, float*, float*, int,
                                                hipStream_t);

}  // namespace lcs
