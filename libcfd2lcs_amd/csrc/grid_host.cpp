// grid_host.cpp -- host set-up of the one-GPU structured grid.
//
// Restates init_sgrid (sgrid_m.f90:16-444) for a single rank: ghost and fake
// coordinates, periodicity, 2-D flags and offsets, periodic shifts, bounding
// box, rectilinearity, the boundary normal (normalised least-squares gradient
// of min(flag,1), gradient_m.f90:114-217,297-335) and 1/delta for the CFL
// number (sgrid_m.f90:951-1014).  The MPI halo exchange of the reference
// becomes the index wrap of a single rank (comms_m.f90:679-869 with every
// neighbour SELF_COMM or NO_COMM).
#include "grid_host.h"

#include <algorithm>
#include <cfloat>
#include <cmath>

namespace lcs {

namespace {

const int kNbr[27][3] = {{0, 0, 0},   {-1, 0, 0}, {0, -1, 0},  {0, 0, -1},  {1, 0, 0},   {0, 1, 0},  {0, 0, 1},
                         {-1, -1, -1}, {0, -1, -1}, {1, -1, -1}, {-1, 0, -1}, {1, 0, -1}, {-1, 1, -1}, {0, 1, -1},
                         {1, 1, -1},  {-1, -1, 0}, {1, -1, 0},  {-1, 1, 0},  {1, 1, 0},   {-1, -1, 1}, {0, -1, 1},
                         {1, -1, 1},  {-1, 0, 1},  {1, 0, 1},   {-1, 1, 1},  {0, 1, 1},   {1, 1, 1}};

inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
inline int dir_index(int i, int j, int k) { return (i + 1) + 3 * (j + 1) + 9 * (k + 1); }

struct Shell {
  const HostGrid& g;
  bool self(int ei, int ej, int ek) const {
    if (!ei && !ej && !ek) return false;
    return (!ei || g.per[0]) && (!ej || g.per[1]) && (!ek || g.per[2]);
  }
  // single-rank exchange of one ghosted field; translate = coordinate
  // component (0,1,2) or -1; face_only = FACE_CONNECT
  void exchange(std::vector<double>& f, int translate, const double ps[27][3]) const {
    const int ni = g.n[0], nj = g.n[1], nk = g.n[2];
    for (int k = 0; k <= nk + 1; k++)
      for (int j = 0; j <= nj + 1; j++)
        for (int i = 0; i <= ni + 1; i++) {
          const int ei = i < 1 ? -1 : (i > ni ? 1 : 0), ej = j < 1 ? -1 : (j > nj ? 1 : 0),
                    ek = k < 1 ? -1 : (k > nk ? 1 : 0);
          if (!self(ei, ej, ek)) continue;
          const int si = ei < 0 ? ni : (ei > 0 ? 1 : i), sj = ej < 0 ? nj : (ej > 0 ? 1 : j),
                    sk = ek < 0 ? nk : (ek > 0 ? 1 : k);
          double v = f[g.id(si, sj, sk)];
          if (translate >= 0) {
            // pack direction = -(ghost side); the z shift is read at
            // periodic_shift(i,k,k,3) (comms_m.f90:750)
            const int di = -ei, dj = -ej, dk = -ek;
            v = v + (translate == 2 ? ps[dir_index(di, dk, dk)][2] : ps[dir_index(di, dj, dk)][translate]);
          }
          f[g.id(i, j, k)] = v;
        }
  }
};

void cross_unit(const double a[3], const double b[3], double out[3]) {
  const double c0 = a[1] * b[2] - a[2] * b[1];
  const double c1 = a[2] * b[0] - a[0] * b[2];
  const double c2 = a[0] * b[1] - a[1] * b[0];
  const double mag = std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  out[0] = c0 / mag;
  out[1] = c1 / mag;
  out[2] = c2 / mag;
}

}  // namespace

std::string HostGrid::build(const int nn[3], const double* x, const double* y, const double* z, const int* fl,
                            double edge_tol, double node_tol) {
  for (int d = 0; d < 3; d++) {
    n[d] = nn[d];
    N[d] = nn[d] + 2;
  }
  const int ni = n[0], nj = n[1], nk = n[2];
  if (ni < 1 || nj < 1 || nk < 1) return "ng > ni, nj or nk (sgrid_m.f90:85-89)";
  if (ni > 32000 || nj > 32000 || nk > 32000) return "grid extent above 32000 nodes per axis";
  nt = (long long)N[0] * N[1] * N[2];
  np = (long long)ni * nj * nk;
  if (nt >= (1LL << 31)) return "ghosted grid above 2^31 nodes";
  int nflat = (ni == 1) + (nj == 1) + (nk == 1);
  if (nflat > 1) return "more than one axis with a single node is not supported";
  for (int d = 0; d < 3; d++) {
    g[d].assign(nt, 0.0);
    bn[d].assign(np, 0.0);
    dl[d].assign(np, 0.0);
  }
  flag.assign(nt, 0);
  // periodicity: an INTERNAL node on a boundary face (sgrid_m.f90:107-137)
  bool lo[3] = {false, false, false}, hi[3] = {false, false, false};
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        const long long u = pid(i, j, k);
        const int f = fl[u];
        if (f < 0 || f > 6) return "unknown bcflag value";
        if (f != 0) continue;
        if (i == 1) lo[0] = true;
        if (j == 1) lo[1] = true;
        if (k == 1) lo[2] = true;
        if (i == ni) hi[0] = true;
        if (j == nj) hi[1] = true;
        if (k == nk) hi[2] = true;
      }
  for (int d = 0; d < 3; d++) {
    if (lo[d] != hi[d]) return "periodicity does not match (check flag values passed to cfd2lcs_init)";
    per[d] = lo[d] && n[d] != 1;
  }
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        const long long u = pid(i, j, k), l = id(i, j, k);
        g[0][l] = x[u];
        g[1][l] = y[u];
        g[2][l] = z[u];
        flag[l] = (int8_t)fl[u];
      }
  // fake nodes by mirror extrapolation from the nearest boundary node
  for (int k = 0; k <= nk + 1; k++)
    for (int j = 0; j <= nj + 1; j++)
      for (int i = 0; i <= ni + 1; i++) {
        const int ib = clampi(i, 1, ni), jb = clampi(j, 1, nj), kb = clampi(k, 1, nk);
        if (i == ib && j == jb && k == kb) continue;
        const long long lb = id(ib, jb, kb);
        const long long lib = id(clampi(2 * ib - i, 1, ni), clampi(2 * jb - j, 1, nj), clampi(2 * kb - k, 1, nk));
        const long long l = id(i, j, k);
        for (int d = 0; d < 3; d++) g[d][l] = g[d][lb] + (g[d][lb] - g[d][lib]);
        flag[l] = flag[lb];
      }
  // 2-D: ghost layers of the single-node axis get LCS_2D where the node is INTERNAL
  auto set2d = [&](long long l1, long long la, long long lb2) {
    const int8_t v = flag[l1] == 0 ? (int8_t)5 : flag[l1];
    flag[la] = v;
    flag[lb2] = v;
  };
  if (ni == 1)
    for (int k = 0; k <= nk + 1; k++)
      for (int j = 0; j <= nj + 1; j++) set2d(id(1, j, k), id(0, j, k), id(2, j, k));
  if (nj == 1)
    for (int k = 0; k <= nk + 1; k++)
      for (int i = 0; i <= ni + 1; i++) set2d(id(i, 1, k), id(i, 0, k), id(i, 2, k));
  if (nk == 1)
    for (int j = 0; j <= nj + 1; j++)
      for (int i = 0; i <= ni + 1; i++) set2d(id(i, j, 1), id(i, j, 0), id(i, j, 2));
  // periodic shifts: exchange without shift, difference to the fakes,
  // exchange again with the shift
  Shell sh{*this};
  double zero[27][3] = {};
  std::vector<double> fake[3];
  for (int d = 0; d < 3; d++) {
    fake[d] = g[d];
    sh.exchange(g[d], d, zero);
  }
  for (int k = -1; k <= 1; k++)
    for (int j = -1; j <= 1; j++)
      for (int i = -1; i <= 1; i++) {
        const int ii = i < 0 ? 0 : (i == 0 ? 1 : ni + 1), jj = j < 0 ? 0 : (j == 0 ? 1 : nj + 1),
                  kk = k < 0 ? 0 : (k == 0 ? 1 : nk + 1);
        const long long l = id(ii, jj, kk);
        for (int d = 0; d < 3; d++) ps[dir_index(i, j, k)][d] = g[d][l] - fake[d][l];
      }
  for (int d = 0; d < 3; d++) sh.exchange(g[d], d, ps);
  // 2-D out-of-plane offset of the ghost layers by the unit normal times the
  // largest extent over max(n) (sgrid_m.f90:304-359)
  double range = 0;
  for (int d = 0; d < 3; d++) {
    double mx = -DBL_MAX, mn = DBL_MAX;
    for (int k = 1; k <= nk; k++)
      for (int j = 1; j <= nj; j++)
        for (int i = 1; i <= ni; i++) {
          const double v = g[d][id(i, j, k)];
          mx = std::max(mx, v);
          mn = std::min(mn, v);
        }
    range = std::max(range, mx - mn);
  }
  const double scaling = range / (double)std::max(ni, std::max(nj, nk));
  for (int k = 0; k <= nk + 1; k++)
    for (int j = 0; j <= nj + 1; j++)
      for (int i = 0; i <= ni + 1; i++) {
        const int im1 = std::max(i - 1, 0), ip1 = std::min(i + 1, ni + 1);
        const int jm1 = std::max(j - 1, 0), jp1 = std::min(j + 1, nj + 1);
        const int km1 = std::max(k - 1, 0), kp1 = std::min(k + 1, nk + 1);
        double delta[3] = {0, 0, 0}, v1[3], v2[3], c[3];
        int along = -1, off = 0;
        if (ni == 1 && i != 1) {
          for (int d = 0; d < 3; d++) {
            v1[d] = g[d][id(1, jp1, 1)] - g[d][id(1, jm1, 1)];
            v2[d] = g[d][id(1, 1, kp1)] - g[d][id(1, 1, km1)];
          }
          along = 0; off = i - 1;
        }
        if (nj == 1 && j != 1) {
          for (int d = 0; d < 3; d++) {
            v1[d] = g[d][id(ip1, 1, 1)] - g[d][id(im1, 1, 1)];
            v2[d] = g[d][id(1, 1, kp1)] - g[d][id(1, 1, km1)];
          }
          along = 1; off = j - 1;
        }
        if (nk == 1 && k != 1) {
          for (int d = 0; d < 3; d++) {
            v1[d] = g[d][id(ip1, 1, 1)] - g[d][id(im1, 1, 1)];
            v2[d] = g[d][id(1, jp1, 1)] - g[d][id(1, jm1, 1)];
          }
          along = 2; off = k - 1;
        }
        if (along >= 0) {
          cross_unit(v1, v2, c);
          for (int d = 0; d < 3; d++) delta[d] = (double)off * c[d];
        }
        const long long l = id(i, j, k);
        for (int d = 0; d < 3; d++) g[d][l] = g[d][l] + delta[d] * scaling;
      }
  for (int d = 0; d < 3; d++) {
    bb[d] = *std::min_element(g[d].begin(), g[d].end());
    bb[3 + d] = *std::max_element(g[d].begin(), g[d].end());
  }
  // rectilinear test (sgrid_m.f90:868-949): lines of constant j,k keep y,z
  // within 1e-4 of the largest extent, etc.  The z test spans k = 1..nk+1
  // when the y test passes (its loop index is reused, sgrid_m.f90:927-932).
  {
    const double TOL = 1e-4;
    double big = 0;
    for (int d = 0; d < 3; d++)
      big = std::max(big, *std::max_element(g[d].begin(), g[d].end()) - *std::min_element(g[d].begin(), g[d].end()));
    auto spread = [&](int d, int i0, int i1, int j0, int j1, int k0, int k1) {
      double mx = -DBL_MAX, mn = DBL_MAX;
      for (int k = k0; k <= k1; k++)
        for (int j = j0; j <= j1; j++)
          for (int i = i0; i <= i1; i++) {
            mx = std::max(mx, g[d][id(i, j, k)]);
            mn = std::min(mn, g[d][id(i, j, k)]);
          }
      return std::fabs(mx - mn) / big;
    };
    bool ox = true, oy = true, oz = true;
    for (int k = 1; k <= nk && ox; k++)
      for (int j = 1; j <= nj && ox; j++)
        if (spread(1, 1, ni, j, j, k, k) > TOL || spread(2, 1, ni, j, j, k, k) > TOL) ox = false;
    int kst = nk + 1;
    for (int k = 1; k <= nk && oy; k++)
      for (int i = 1; i <= ni && oy; i++)
        if (spread(0, i, i, 1, nj, k, k) > TOL || spread(2, i, i, 1, nj, k, k) > TOL) {
          oy = false;
          kst = k;
        }
    for (int j = 1; j <= nj && oz; j++)
      for (int i = 1; i <= ni && oz; i++)
        if (spread(0, i, i, j, j, 1, kst) > TOL || spread(1, i, i, j, j, 1, kst) > TOL) oz = false;
    rect = ox && oy && oz;
  }
  // separable coordinates: the rectilinear GPU path stores one line per axis
  // (a non-rectilinear grid keeps its full coordinates; the lines through
  // the first row are then unused)
  separable = rect;
  for (int d = 0; d < 3; d++) {
    line[d].resize(N[d]);
    for (int a = 0; a < N[d]; a++) line[d][a] = g[d][id(d == 0 ? a : 1, d == 1 ? a : 1, d == 2 ? a : 1)];
  }
  // Interior and face-ghost nodes must lie on the lines (within node_tol
  // times the extent: exactly for user grids).  Ghost edges
  // and corners carry the rounding of the periodic shifts (one shift per
  // neighbour direction), and where a periodic axis meets a wall the edge
  // keeps its mirror value, built from the spacing at the other end of the
  // axis; both differ from the lines by the rounding of the input coordinates
  // only, and the kernels read face ghosts only.
  double ext = 0;
  for (int d = 0; d < 3; d++) ext = std::max(ext, std::fabs(bb[3 + d] - bb[d]));
  const double tol = edge_tol * std::max(ext, 1.0), ntol = node_tol * std::max(ext, 1.0);
  for (int k = 0; k <= nk + 1 && separable; k++)
    for (int j = 0; j <= nj + 1 && separable; j++)
      for (int i = 0; i <= ni + 1; i++) {
        const long long l = id(i, j, k);
        const int ghosts = (i < 1 || i > ni) + (j < 1 || j > nj) + (k < 1 || k > nk);
        const double e = std::max(std::fabs(g[0][l] - line[0][i]),
                                  std::max(std::fabs(g[1][l] - line[1][j]), std::fabs(g[2][l] - line[2][k])));
        if (ghosts <= 1 ? e > ntol : e > tol) {
          separable = false;
          break;
        }
      }
  if (rect && !separable)
    return "the grid coordinates are not separable (x(i), y(j), z(k)) although rectilinear within 1e-4; "
           "not supported by this build";
  for (int d = 0; d < 3; d++) {
    xl[d] = line[d];
    if (n[d] == 1) {
      xl[d][0] = xl[d][1] - 1.0;
      xl[d][2] = xl[d][1] + 1.0;
    }
  }
  flat = ni == 1 ? 0 : (nj == 1 ? 1 : (nk == 1 ? 2 : -1));
  // boundary normal
  {
    std::vector<double> bc(nt);
    for (long long l = 0; l < nt; l++) bc[l] = flag[l] == 5 ? 0.0 : (double)std::min<int>(flag[l], 1);
    sh.exchange(bc, -1, zero);
    for (int k = 1; k <= nk; k++)
      for (int j = 1; j <= nj; j++)
        for (int i = 1; i <= ni; i++) {
          const long long l = id(i, j, k);
          double sxx = 0, syy = 0, szz = 0, sxy = 0, sxz = 0, syz = 0, dx[3];
          for (int nb = 1; nb < 27; nb++) {
            const long long m = id(i + kNbr[nb][0], j + kNbr[nb][1], k + kNbr[nb][2]);
            for (int d = 0; d < 3; d++) dx[d] = g[d][m] - g[d][l];
            const double idw = 1.0 / (dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
            sxx = sxx + idw * (dx[0] * dx[0]);
            syy = syy + idw * (dx[1] * dx[1]);
            szz = szz + idw * (dx[2] * dx[2]);
            sxy = sxy + idw * dx[0] * dx[1];
            sxz = sxz + idw * dx[0] * dx[2];
            syz = syz + idw * dx[1] * dx[2];
          }
          const double den =
              2.0 * sxy * sxz * syz + sxx * syy * szz - sxx * (syz * syz) - syy * (sxz * sxz) - szz * (sxy * sxy);
          double gv[3] = {0, 0, 0};
          for (int nb = 1; nb < 27; nb++) {
            const long long m = id(i + kNbr[nb][0], j + kNbr[nb][1], k + kNbr[nb][2]);
            for (int d = 0; d < 3; d++) dx[d] = g[d][m] - g[d][l];
            const double idw = 1.0 / (dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
            const double wx =
                idw * ((syy * szz - syz * syz) * dx[0] + (sxz * syz - sxy * szz) * dx[1] + (sxy * syz - sxz * syy) * dx[2]) / den;
            const double wy =
                idw * ((sxz * syz - sxy * szz) * dx[0] + (sxx * szz - sxz * sxz) * dx[1] + (sxy * sxz - syz * sxx) * dx[2]) / den;
            const double wz =
                idw * ((sxy * syz - sxz * syy) * dx[0] + (sxy * sxz - syz * sxx) * dx[1] + (sxx * syy - sxy * sxy) * dx[2]) / den;
            const double t = bc[m] - bc[l];
            gv[0] = gv[0] + t * wx;
            gv[1] = gv[1] + t * wy;
            gv[2] = gv[2] + t * wz;
          }
          const double mag = gv[0] * gv[0] + gv[1] * gv[1] + gv[2] * gv[2];
          const long long p = pid(i, j, k);
          for (int d = 0; d < 3; d++) bn[d][p] = mag <= 0.0 ? 0.0 : gv[d] / std::sqrt(mag);
        }
  }
  // non-rectilinear grids: face least-squares gradient weights
  // (compute_lsgw with full_conn = .false., sgrid_m.f90:421-424,
  // gradient_m.f90:114-217), [(f * 3 + d) * np + p] for the faces in NBR
  // order, and the positions of the coordinate extrema for the node guess
  // of a relocated particle (comms_m.f90:1625-1632)
  if (!rect) {
    static const int face[6][3] = {{-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    lsw.assign((size_t)18 * np, 0.0);
    for (int k = 1; k <= nk; k++)
      for (int j = 1; j <= nj; j++)
        for (int i = 1; i <= ni; i++) {
          const long long l = id(i, j, k), p = pid(i, j, k);
          double sxx = 0, syy = 0, szz = 0, sxy = 0, sxz = 0, syz = 0, dx[3];
          for (int f = 0; f < 6; f++) {
            const long long m = id(i + face[f][0], j + face[f][1], k + face[f][2]);
            for (int d = 0; d < 3; d++) dx[d] = g[d][m] - g[d][l];
            const double idw = 1.0 / (dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
            sxx = sxx + idw * (dx[0] * dx[0]);
            syy = syy + idw * (dx[1] * dx[1]);
            szz = szz + idw * (dx[2] * dx[2]);
            sxy = sxy + idw * dx[0] * dx[1];
            sxz = sxz + idw * dx[0] * dx[2];
            syz = syz + idw * dx[1] * dx[2];
          }
          const double den =
              2.0 * sxy * sxz * syz + sxx * syy * szz - sxx * (syz * syz) - syy * (sxz * sxz) - szz * (sxy * sxy);
          for (int f = 0; f < 6; f++) {
            const long long m = id(i + face[f][0], j + face[f][1], k + face[f][2]);
            for (int d = 0; d < 3; d++) dx[d] = g[d][m] - g[d][l];
            const double idw = 1.0 / (dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
            double* w = &lsw[(size_t)f * 3 * np];
            w[p] = idw * ((syy * szz - syz * syz) * dx[0] + (sxz * syz - sxy * szz) * dx[1] +
                          (sxy * syz - sxz * syy) * dx[2]) / den;
            w[np + p] = idw * ((sxz * syz - sxy * szz) * dx[0] + (sxx * szz - sxz * sxz) * dx[1] +
                               (sxy * sxz - syz * sxx) * dx[2]) / den;
            w[2 * np + p] = idw * ((sxy * syz - sxz * syy) * dx[0] + (sxy * sxz - syz * sxx) * dx[1] +
                                   (sxx * syy - sxy * sxy) * dx[2]) / den;
          }
        }
    for (int d = 0; d < 3; d++) {
      long long amin = 0, amax = 0;
      for (long long l = 1; l < nt; l++) {
        if (g[d][l] < g[d][amin]) amin = l;
        if (g[d][l] > g[d][amax]) amax = l;
      }
      const long long a2[2] = {amin, amax};
      for (int m = 0; m < 2; m++) {
        const long long l = a2[m];
        const double pos[3] = {(double)(l % N[0] + 1), (double)((l / N[0]) % N[1] + 1),
                               (double)(l / ((long long)N[0] * N[1]) + 1)};
        for (int e = 0; e < 3; e++) (m ? pmax : pmin)[d][e] = pos[e];
      }
    }
  }
  // 1/delta over the 3x3x3 neighbourhood
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        const long long p = pid(i, j, k);
        for (int d = 0; d < 3; d++) {
          double mx = -DBL_MAX, mn = DBL_MAX;
          for (int c = k - 1; c <= k + 1; c++)
            for (int b = j - 1; b <= j + 1; b++)
              for (int a = i - 1; a <= i + 1; a++) {
                mx = std::max(mx, g[d][id(a, b, c)]);
                mn = std::min(mn, g[d][id(a, b, c)]);
              }
          const double dd = 0.5 * std::fabs(mx - mn);
          dl[d][p] = dd > 0.0 ? 1.0 / dd : 0.0;
        }
      }
  return "";
}

// Knot placement of dbknot (bspline_sub_module.f90:1737-1785) and the
// collocation matrix of dbintk (1896-2036), factored without pivoting as
// dbnfac does (2098-2184).  B-spline values by de Boor's recurrence.
std::string spline_axis(const std::vector<double>& x, int k, SplineAxis& out) {
  const int n = (int)x.size();
  if (n < 2 || k < 2 || k > 4 || n < k) return "spline order/abscissae out of range";
  out.k = k;
  out.knots.assign(n + k, 0.0);
  std::vector<double>& t = out.knots;
  const double rnot = x[n - 1] + 0.1 * (x[n - 1] - x[n - 2]);
  for (int j = 0; j < k; j++) {
    t[j] = x[0];
    t[n + j] = rnot;
  }
  if (k % 2) {
    for (int j = k + 1; j <= n; j++) {
      const int m = j + (k - 1) / 2 - k;  // 1-based data index
      t[j - 1] = 0.5 * (x[m - 1] + x[m]);
    }
  } else {
    for (int j = k + 1; j <= n; j++) t[j - 1] = x[j + k / 2 - k - 1];
  }
  for (int i = 0; i + 1 < n; i++)
    if (!(x[i] < x[i + 1])) return "spline abscissae not strictly increasing";
  const int mu = k - 1, nb = 2 * k - 1;
  out.band.assign((size_t)nb * n, 0.0);
  auto A = [&](int r, int c) -> double& { return out.band[(r - c + mu) + (size_t)nb * c]; };
  int left = k;  // 1-based
  for (int i = 1; i <= n; i++) {
    const double xi = x[i - 1];
    const int cap = std::min(i + k, n + 1);
    left = std::max(left, i);
    if (xi < t[left - 1]) return "spline collocation matrix singular";
    bool found = false;
    for (;;) {
      if (xi < t[left]) {
        found = true;
        break;
      }
      left++;
      if (left >= cap) break;
    }
    if (!found) {
      left--;
      if (xi > t[left]) return "spline collocation matrix singular";
    }
    double v[4], dp[4], dm[4];
    v[0] = 1.0;
    for (int j = 1; j < k; j++) {
      dp[j - 1] = t[left + j - 1] - xi;
      dm[j - 1] = xi - t[left - j];
      double prev = 0.0;
      for (int l = 1; l <= j; l++) {
        const double vm = v[l - 1] / (dp[l - 1] + dm[j - l]);
        v[l - 1] = vm * dp[l - 1] + prev;
        prev = vm * dm[j - l];
      }
      v[j] = prev;
    }
    for (int j = 1; j <= k; j++) A(i - 1, left - k + j - 1) = v[j - 1];
  }
  for (int i = 0; i + 1 < n; i++) {
    const double pivot = A(i, i);
    if (pivot == 0.0) return "spline collocation matrix has a zero pivot";
    const int jmax = std::min(mu, n - 1 - i);
    for (int j = 1; j <= jmax; j++) A(i + j, i) /= pivot;
    for (int kk = 1; kk <= jmax; kk++) {
      const double f = A(i, i + kk);
      for (int j = 1; j <= jmax; j++) A(i + j, i + kk) = A(i + j, i + kk) - A(i + j, i) * f;
    }
  }
  if (A(n - 1, n - 1) == 0.0) return "spline collocation matrix has a zero pivot";
  // piecewise-polynomial form of the basis on every knot interval, by the
  // Cox-de Boor recurrence carried out on polynomials in s
  const int rowlen = k * k + 1;
  out.poly.assign((size_t)(n + 1) * rowlen, 0.0);
  for (int L = k; L <= n; L++) {
    const double tl = t[L - 1], hL = t[L] - t[L - 1];
    if (!(hL > 0)) return "empty knot interval";
    // B[i] for i = L-k+1..L (index i - (L-k+1)), each a polynomial of degree < k
    std::vector<std::vector<double>> B(k, std::vector<double>(k, 0.0));
    B[k - 1][0] = 1.0;  // order 1: B_L = 1 on [t(L), t(L+1))
    for (int ord = 2; ord <= k; ord++) {
      std::vector<std::vector<double>> Nb(k, std::vector<double>(k, 0.0));
      for (int i = L - ord + 1; i <= L; i++) {
        const int ii = i - (L - k + 1);
        // omega_i(x) = (x - t_i)/(t_{i+ord-1} - t_i), x = tl + hL s
        auto lin = [&](int i0, int span, double& c0, double& c1) {
          const double den = t[i0 + span - 1] - t[i0 - 1];
          if (den == 0.0) { c0 = 0.0; c1 = 0.0; return; }
          c0 = (tl - t[i0 - 1]) / den;
          c1 = hL / den;
        };
        double a0, a1, b0, b1;
        lin(i, ord - 1, a0, a1);      // omega_{i,ord}
        lin(i + 1, ord - 1, b0, b1);  // omega_{i+1,ord}
        std::vector<double>& r = Nb[ii];
        if (i >= L - ord + 2) {  // B_{i,ord-1} nonzero on this interval
          const std::vector<double>& P = B[ii];
          for (int m = 0; m < k; m++) {
            if (P[m] == 0.0) continue;
            r[m] += a0 * P[m];
            if (m + 1 < k) r[m + 1] += a1 * P[m];
          }
        }
        if (i + 1 <= L) {  // (1 - omega_{i+1}) B_{i+1,ord-1}
          const std::vector<double>& P = B[ii + 1];
          for (int m = 0; m < k; m++) {
            if (P[m] == 0.0) continue;
            r[m] += (1.0 - b0) * P[m];
            if (m + 1 < k) r[m + 1] += -b1 * P[m];
          }
        }
      }
      B.swap(Nb);
    }
    double* row = &out.poly[(size_t)L * rowlen];
    for (int a = 0; a < k; a++)
      for (int m = 0; m < k; m++) row[a * k + m] = B[a][m];
    row[k * k] = 1.0 / hL;
  }
  return "";
}

// new_sgrid_from_sgrid on one rank, rectilinear case (sgrid_m.f90:484-620
// counts, 635-665 coordinates, 693-740 removal).  The arithmetic runs in R,
// the precision of the library's reals (LCSRP).
template <typename R>
int refine_grid(const HostGrid& src, int res, int nn[3], std::vector<double> xyz[3], std::vector<int>& flag) {
  for (int d = 0; d < 3; d++)
    if (src.n[d] > 1 && std::abs(res) >= src.n[d]) res = 0;
  if (res == 0) return 0;
  std::vector<int> map[3], sub[3];
  for (int d = 0; d < 3; d++) {
    const int n = src.n[d];
    int m;
    if (n == 1) m = 1;
    else if (res > 0) m = src.per[d] ? n * (res + 1) : n * (res + 1) - res;
    else m = n / (std::abs(res) + 1);
    nn[d] = m;
    map[d].resize(m);
    sub[d].resize(m);
    for (int a = 0; a < m; a++) {
      if (n == 1) { map[d][a] = 1; sub[d][a] = 0; }
      else if (res > 0) { map[d][a] = 1 + a / (res + 1); sub[d][a] = a % (res + 1); }
      else { map[d][a] = (a + 1) * (std::abs(res) + 1); sub[d][a] = 0; }
    }
  }
  // per axis: the new coordinate line (the grid is rectilinear)
  std::vector<R> lin[3];
  for (int d = 0; d < 3; d++) {
    lin[d].resize(nn[d]);
    for (int a = 0; a < nn[d]; a++) {
      const int ii = map[d][a];
      const R x0 = (R)src.line[d][ii];
      if (res > 0) {
        const R dx = ((R)src.line[d][ii + 1] - x0) / (R)(res + 1);
        lin[d][a] = x0 + dx * (R)sub[d][a];
      } else {
        lin[d][a] = x0;
      }
    }
  }
  const long long np = (long long)nn[0] * nn[1] * nn[2];
  for (int d = 0; d < 3; d++) xyz[d].resize(np);
  flag.resize(np);
  for (int k = 0; k < nn[2]; k++)
    for (int j = 0; j < nn[1]; j++)
      for (int i = 0; i < nn[0]; i++) {
        const long long p = i + (long long)nn[0] * (j + (long long)nn[1] * k);
        xyz[0][p] = (double)lin[0][i];
        xyz[1][p] = (double)lin[1][j];
        xyz[2][p] = (double)lin[2][k];
        flag[p] = src.flag[src.id(map[0][i], map[1][j], map[2][k])];
      }
  return res;
}

// create_aux_grid (lcs_m.f90:264-300 offsets, 317-334 shifts, 339-344
// projection): aux_delta is AUX_GRID_SCALING (0.2, data_m.f90:92) times the
// mean distance to the in-range face neighbours, in NBR_OFFSET order; set a
// moves every node by -/+ aux_delta along axis a % 3 (X0, Y0, Z0, X1, Y1, Z1),
// and on a 2-D grid the node is put back onto the data plane.  The shift
// starts from the node coordinate.
template <typename R>
void aux_grid_coords(const HostGrid& src, int a, std::vector<double> xyz[3]) {
  static const int face[6][3] = {{-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  const int ni = src.n[0], nj = src.n[1], nk = src.n[2];
  for (int d = 0; d < 3; d++) xyz[d].resize(src.np);
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        const long long l = src.id(i, j, k), p = src.pid(i, j, k);
        R nbrsum = 0, distsum = 0;
        for (int f = 0; f < 6; f++) {
          const int ii = i + face[f][0], jj = j + face[f][1], kk = k + face[f][2];
          if (ii > 0 && jj > 0 && kk > 0 && ii <= ni && jj <= nj && kk <= nk) {
            const long long m = src.id(ii, jj, kk);
            nbrsum = nbrsum + (R)1;
            const R dx = (R)src.g[0][l] - (R)src.g[0][m], dy = (R)src.g[1][l] - (R)src.g[1][m],
                    dz = (R)src.g[2][l] - (R)src.g[2][m];
            distsum = distsum + std::sqrt(dx * dx + dy * dy + dz * dz);
          }
        }
        const R delta = distsum / nbrsum * (R)0.2;
        R q[3] = {(R)src.g[0][l], (R)src.g[1][l], (R)src.g[2][l]};
        const int ax = a % 3;
        q[ax] = a < 3 ? q[ax] - delta : q[ax] + delta;
        if (src.flat >= 0) {
          const int f = src.flat;
          const R plane = (R)src.g[f][l];
          q[f] = q[f] - (q[f] - plane);
        }
        for (int d = 0; d < 3; d++) xyz[d][p] = (double)q[d];
      }
}

template int refine_grid<float>(const HostGrid&, int, int*, std::vector<double>*, std::vector<int>&);
template int refine_grid<double>(const HostGrid&, int, int*, std::vector<double>*, std::vector<int>&);
template void aux_grid_coords<float>(const HostGrid&, int, std::vector<double>*);
template void aux_grid_coords<double>(const HostGrid&, int, std::vector<double>*);

}  // namespace lcs
