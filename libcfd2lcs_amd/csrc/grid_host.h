// grid_host.h -- one-time host set-up of the structured grid (the reference's
// init_sgrid on one rank) and of the B-spline collocation factors.  Runs once
// at cfd2lcs_init / on the first spline use; the results are uploaded to HBM.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace lcs {

struct HostGrid {
  int n[3] = {0, 0, 0};
  int N[3] = {0, 0, 0};
  long long nt = 0, np = 0;
  std::vector<double> g[3];      // ghosted coordinates with ghost/fake values
  std::vector<int8_t> flag;      // ghosted bcflag
  std::vector<double> bn[3];     // bcnorm at interior nodes
  std::vector<double> dl[3];     // 1/delta at interior nodes
  int per[3] = {0, 0, 0};
  double ps[27][3] = {};
  double bb[6] = {};
  bool rect = false;
  bool separable = false;        // g[d] depends on the d-th index only
  int flat = -1;                 // axis with n == 1
  std::vector<double> line[3];   // ghosted coordinate lines
  std::vector<double> xl[3];     // interpolation lines (2-D override)
  // non-rectilinear grids: face least-squares weights (18 per interior
  // node, [(f * 3 + d) * np + p]) and the 1-based array positions of the
  // coordinate extrema (pmin[d][e]: component e of minloc of coordinate d)
  std::vector<double> lsw;
  double pmin[3][3] = {}, pmax[3][3] = {};

  long long id(int i, int j, int k) const { return i + (long long)N[0] * (j + (long long)N[1] * k); }
  long long pid(int i, int j, int k) const {
    return (i - 1) + (long long)n[0] * ((j - 1) + (long long)n[1] * (k - 1));
  }
  // Builds the grid from the user's interior arrays (Fortran order).
  // edge_tol: relative tolerance (times the domain extent) for ghost edges and
  // corners against the coordinate lines; it follows the precision of the
  // user's coordinates.  node_tol: the same for interior and face-ghost nodes
  // (0: exactly on the lines; a few ulp for generated grids whose nodes carry
  // rounding, e.g. AUX_GRID shifts, which the reference's rectilinear test
  // accepts within 1e-4, sgrid_m.f90:875).  Returns an empty string on
  // success, else the reason.
  std::string build(const int nn[3], const double* x, const double* y, const double* z, const int* fl,
                    double edge_tol = 1e-9, double node_tol = 0.0);
};

// Knots (dbknot rule) and the banded LU factors of the collocation matrix
// (dbintk + dbnfac) of an order-k interpolating spline on abscissae x.
struct SplineAxis {
  int k = 0;
  std::vector<double> knots;  // n + k
  std::vector<double> band;   // (2k-1) * n, a(r,c) at (r - c + k - 1) + (2k-1)*c
  // per knot interval L = k..n (1-based, t(L) <= x < t(L+1)): the k nonzero
  // B-splines B_{L-k+1..L} as polynomials in s = (x - t(L)) / (t(L+1) - t(L)),
  // row L at poly[L*(k*k+1)]: coefficients [a*k + m] of s^m, then 1/(t(L+1)-t(L))
  std::vector<double> poly;
};
std::string spline_axis(const std::vector<double>& x, int k, SplineAxis& out);

// The grid of a diagnostic with resolution res != 0 (new_sgrid_from_sgrid):
// interior coordinates (Fortran order) and flags.  Returns the effective
// res, 0 when |res| >= n along an axis with n > 1 (the flow grid is used).
// R is the precision of the library's reals.
template <typename R>
int refine_grid(const HostGrid& src, int res, int nn[3], std::vector<double> xyz[3], std::vector<int>& flag);
// Interior coordinates of AUX_GRID set a = 0..5 (X0, Y0, Z0, X1, Y1, Z1) of
// grid src (create_aux_grid); its flags are those of src.
template <typename R>
void aux_grid_coords(const HostGrid& src, int a, std::vector<double> xyz[3]);

}  // namespace lcs
