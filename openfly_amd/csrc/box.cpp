// Host construction of the periodic box and its cell grid.
//
// Box selection follows box.hpp:63-76 (orthorhombic when the off-diagonal norm < 1e-4, otherwise an
// upper-triangular triclinic basis, triclinic.hpp:134-158); the grid follows hypergrid.hpp:43-53 with
// TriGrid's column-sum extents (triclinic.hpp:94) and per-axis plane normals (triclinic.hpp:94-115).
#include <cmath>
#include <limits>

#include "common.hpp"

namespace flyg {

  BoxDesc make_box(double const* b, int const* periodic) {
    BoxDesc box{};
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) box.B[r][c] = b[3 * r + c];
    for (int k = 0; k < 3; ++k) box.per[k] = periodic[k] ? 1 : 0;
    double skew = 0;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c)
        if (r != c) skew += box.B[r][c] * box.B[r][c];
    box.ortho = std::sqrt(skew) < 1e-4 ? 1 : 0;
    if (box.ortho) {
      for (int k = 0; k < 3; ++k) {
        box.ext[k] = box.B[k][k];
        require(box.ext[k] > 0, "Orthorhombic extents must be positive");
        box.inv_ext[k] = 1.0 / box.ext[k];
      }
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
          if (r != c) box.B[r][c] = 0;
      return box;
    }
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < r; ++c) require(std::abs(box.B[r][c]) < 1e-5, "Basis must be an upper triangular matrix.");
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        if (r > c) box.B[r][c] = 0;
        require(box.B[r][c] >= 0, "Basis elements must be positive");
      }
    for (int k = 0; k < 3; ++k) require(box.B[k][k] > 1e-5, "Diagonal elements must be non-zero");
    double a = box.B[0][0], bb = box.B[0][1], c = box.B[0][2], d = box.B[1][1], e = box.B[1][2], f = box.B[2][2];
    box.Binv[0][0] = 1 / a;
    box.Binv[1][1] = 1 / d;
    box.Binv[2][2] = 1 / f;
    box.Binv[0][1] = -bb / (a * d);
    box.Binv[1][2] = -e / (d * f);
    box.Binv[0][2] = (bb * e - c * d) / (a * d * f);
    for (int i = 0; i < 3; ++i) {
      int j = (i + 1) % 3, k = (i + 2) % 3;
      double bj[3] = {box.B[0][j], box.B[1][j], box.B[2][j]};
      double bk[3] = {box.B[0][k], box.B[1][k], box.B[2][k]};
      double n[3] = {bj[1] * bk[2] - bj[2] * bk[1], bj[2] * bk[0] - bj[0] * bk[2], bj[0] * bk[1] - bj[1] * bk[0]};
      double len = std::sqrt((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
      for (double& x : n) x /= len;
      double w = (n[0] * box.B[0][i] + n[1] * box.B[1][i]) + n[2] * box.B[2][i];
      for (int r = 0; r < 3; ++r) box.hyper[r][i] = w < 0 ? -n[r] : n[r];
    }
    return box;
  }

  double box_min_width(BoxDesc const& b) {
    if (b.ortho) return std::fmin(std::fmin(b.ext[0], b.ext[1]), b.ext[2]);
    double m = std::numeric_limits<double>::max();
    for (int i = 0; i < 3; ++i) {
      double w = (b.B[0][i] * b.hyper[0][i] + b.B[1][i] * b.hyper[1][i]) + b.B[2][i] * b.hyper[2][i];
      m = std::fmin(m, w);
    }
    return m;
  }

  GridDesc make_grid(BoxDesc const& box, double r_cut) {
    GridDesc g{};
    require(r_cut > 0, "r_cut is negative");
    g.rc = r_cut;
    for (int k = 0; k < 3; ++k) {
      g.gext[k] = box.ortho ? box.ext[k] : (box.B[0][k] + box.B[1][k]) + box.B[2][k];
      require(g.gext[k] > r_cut, "r_cut is too big for the box");
    }
    if (!box.ortho) {
      for (int i = 0; i < 3; ++i) {
        double w = (box.B[0][i] * box.hyper[0][i] + box.B[1][i] * box.hyper[1][i]) + box.B[2][i] * box.hyper[2][i];
        require(std::abs(w) > r_cut, "Box axis too small for this r_cut");
      }
    }
    int p = 1;
    for (int k = 0; k < 3; ++k) {
      g.shape[k] = 2 + static_cast<int>(g.gext[k] / r_cut);
      g.cell[k] = g.gext[k] / std::floor(g.gext[k] / r_cut);
      g.inv_cell[k] = 1.0 / g.cell[k];
      g.prod[k] = p;
      p *= g.shape[k];
    }
    g.ncells = p;
    return g;
  }

}  // namespace flyg
