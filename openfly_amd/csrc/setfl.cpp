// setfl / eam.fs reader and natural cubic spline tables (host side of the potential handle).
//
// File layout as read by the reference DataEAM ctor (data.cpp:67-153): three comment lines; a line with
// the species count and names; "Nrho drho Nr dr rcut"; for each species a "Z mass ..." line, Nrho values
// of F, then Nr values of each density table rho_{i,j} (stored at phi[i + n*j]); then for every pair
// j <= i, Nr values of r*V(r) -- divided by r = k*dr with V[0] := V[1] -- stored at V[i + n*j].
// Values are read five per line (read_chunked, data.cpp:48-65).
//
// Splines: a natural cubic spline on the uniform grid x_k = k*dx, padded with five copies of the last
// value (spline.cpp:27-103), i.e. the textbook construction (Burden & Faires, Numerical Analysis,
// Algorithm 3.4) specialised to equal intervals.
#include <cmath>
#include <fstream>
#include <sstream>

#include "eam_host.hpp"

namespace flyg {

  std::vector<double4> spline_coefficients(std::vector<double> y, double dx) {
    double const last = y.back();
    for (int k = 0; k < 5; ++k) y.push_back(last);
    std::size_t const n = y.size() - 1;   // number of intervals
    std::vector<double> alpha(n + 1, 0.0), l(n + 1), mu(n + 1), z(n + 1), c(n + 1), b(n), d(n);
    for (std::size_t i = 1; i < n; ++i) alpha[i] = 3 * (y[i + 1] - 2 * y[i] + y[i - 1]) / dx;
    l[0] = 1;
    mu[0] = 0;
    z[0] = 0;
    for (std::size_t i = 1; i < n; ++i) {
      l[i] = dx * (4 - mu[i - 1]);
      mu[i] = dx / l[i];
      z[i] = (alpha[i] - dx * z[i - 1]) / l[i];
    }
    l[n] = 1;
    z[n] = 0;
    c[n] = 0;
    for (std::size_t j = n; j-- > 0;) {
      c[j] = z[j] - mu[j] * c[j + 1];
      b[j] = (y[j + 1] - y[j]) / dx - dx * (c[j + 1] + 2 * c[j]) / 3;
      d[j] = (c[j + 1] - c[j]) / (3 * dx);
    }
    std::vector<double4> out(n);
    for (std::size_t i = 0; i < n; ++i) out[i] = make_double4(y[i], b[i], c[i], d[i]);
    return out;
  }

  namespace {
    std::istringstream line_of(std::istream& in) {
      std::string s;
      if (!std::getline(in, s)) throw Error(FLYG_INVALID_ARG, "File terminated too soon");
      return std::istringstream(s);
    }
    std::vector<double> values(std::istream& in, std::size_t count) {
      std::vector<double> v;
      v.reserve(count);
      std::istringstream ln;
      for (std::size_t i = 0; i < count; ++i) {
        if (i % 5 == 0) ln = line_of(in);
        double t = 0;
        ln >> t;
        v.push_back(t);
      }
      return v;
    }
  }  // namespace

  SetflData read_setfl(std::string const& path) {
    std::ifstream in(path);
    require(in.good(), "Could not open eam file " + path);
    return read_setfl_stream(in);
  }

  SetflData read_setfl_stream(std::istream& in) {
    SetflData D;
    for (int i = 0; i < 3; ++i) line_of(in);
    {
      auto ln = line_of(in);
      int n = 0;
      ln >> n;
      require(n > 0 && n <= MAX_TYPES, "unsupported species count in setfl file");
      D.ntypes = n;
      for (int i = 0; i < n; ++i) {
        std::string name;
        ln >> name;
        D.names.push_back(name);
      }
    }
    std::size_t nrho = 0, nr = 0;
    line_of(in) >> nrho >> D.drho >> nr >> D.dr >> D.rcut;
    require(nrho > 1 && nr > 1 && D.drho > 0 && D.dr > 0 && D.rcut > 0, "bad tabulation line in setfl file");
    int const n = D.ntypes;
    D.F.resize(n);
    D.phi.resize(n * n);
    D.V.resize(n * n);
    D.mass.resize(n);
    for (int i = 0; i < n; ++i) {
      int z = 0;
      line_of(in) >> z >> D.mass[i];
      D.F[i] = values(in, nrho);
      for (int j = 0; j < n; ++j) D.phi[i + n * j] = values(in, nr);
    }
    for (int i = 0; i < n; ++i)
      for (int j = 0; j <= i; ++j) {
        std::vector<double> rv = values(in, nr);
        for (std::size_t k = 0; k < rv.size(); ++k) rv[k] /= D.dr * static_cast<double>(k);
        rv[0] = rv[1];
        D.V[i + n * j] = std::move(rv);
      }
    return D;
  }

}  // namespace flyg
