// Environment self-symmetries and Master::append_syms (catalogue.cpp:33-64, find.cpp:1279-1303) on bcc
// Fe, checked against the point groups the lattice fixes: a perfect crystal's environment has the 48
// operations of O_h (signed permutation matrices), an atom beside a vacancy the 6 of C3v about the
// vacancy direction, a randomly distorted one only the identity; a <111> hop expands to the 8 <111>
// directions, a <100> hop to the 6 <100> ones, and under C3v the hop into the vacancy stays alone while a
// hop along another <111> expands to its 3-member orbit.  Prints one line per case; exit status 0 on success.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "libfly/gpu/facade.hpp"

using namespace fly;

static int failures = 0;
#define REQUIRE(c)                                                           \
  do {                                                                       \
    if (!(c)) {                                                              \
      std::fprintf(stderr, "REQUIRE failed: %s (line %d)\n", #c, __LINE__); \
      ++failures;                                                            \
    }                                                                        \
  } while (0)

constexpr double A = 2.855324;
constexpr int CELLS = 7;

struct Cell {
  system::Box box;
  system::SoA<Position, Frozen, TypeID> atoms;
};

static Cell lattice(bool vacancy, double noise) {
  std::vector<Vec> lat;
  std::mt19937_64 rng(7);
  std::normal_distribution<double> nd(0.0, noise > 0 ? noise : 1.0);
  for (int k = 0; k < CELLS; ++k)
    for (int j = 0; j < CELLS; ++j)
      for (int i = 0; i < CELLS; ++i)
        for (double off : {0.0, 0.5}) lat.push_back(Vec{(i + off) * A, (j + off) * A, (k + off) * A});
  if (vacancy) lat.erase(lat.begin() + 1);   // the site at a/2 (1, 1, 1) next to atom 0
  Mat basis = Mat::Zero();
  for (int i = 0; i < 3; ++i) basis(i, i) = A * CELLS;
  Cell c{system::Box(basis, Arr<bool>::Constant(true)),
         system::SoA<Position, Frozen, TypeID>(static_cast<std::int64_t>(lat.size()))};
  for (std::int64_t i = 0; i < c.atoms.size(); ++i) {
    Vec p = lat[static_cast<std::size_t>(i)];
    if (noise > 0) p = Vec{p[0] + nd(rng), p[1] + nd(rng), p[2] + nd(rng)};
    c.atoms(r_, i) = p;
    c.atoms(id_, i) = 0;
    c.atoms(fzn_, i) = false;
  }
  return c;
}

// Atoms within r of atom 0 (minimum image), ascending: the atoms a mechanism runs over.
static std::vector<std::int64_t> near0(Cell const& c, double r) {
  std::vector<std::int64_t> out;
  double const L = A * CELLS;
  for (std::int64_t i = 0; i < c.atoms.size(); ++i) {
    double d2 = 0;
    for (int k = 0; k < 3; ++k) {
      double d = c.atoms(r_, i)[k] - c.atoms(r_, 0)[k];
      d -= L * std::round(d / L);
      d2 += d * d;
    }
    if (d2 < r * r) out.push_back(i);
  }
  return out;
}

// A mechanism that moves atom 0 only: saddle half way along `dir`, forward minimum at `dir`.
static saddle::Master::Mech hop(std::size_t n, std::size_t k0, Vec dir) {
  saddle::Master::Mech m;
  m.delta_sp.assign(3 * n, 0.0);
  m.delta_fwd.assign(3 * n, 0.0);
  m.axis.assign(3 * n, 0.0);
  double norm = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
  for (int c = 0; c < 3; ++c) {
    m.delta_sp[3 * k0 + c] = 0.5 * dir[c];
    m.delta_fwd[3 * k0 + c] = dir[c];
    m.axis[3 * k0 + c] = dir[c] / norm;
  }
  return m;
}

static bool signed_permutation(Mat const& O) {
  for (int r = 0; r < 3; ++r) {
    int ones = 0;
    for (int c = 0; c < 3; ++c) {
      double v = std::abs(O(r, c));
      if (std::abs(v - 1.0) < 1e-9) ++ones;
      else if (v > 1e-9) return false;
    }
    if (ones != 1) return false;
  }
  return true;
}

static std::size_t expand(std::vector<env::SelfSymmetry> const& syms, saddle::Master::Mech const& m) {
  std::vector<saddle::Master::Mech> mechs{m};
  return saddle::Master::append_syms(syms, m, mechs, 0.1);
}

int main() {
  double const h = 0.5 * A;
  {
    Cell c = lattice(false, 0.0);
    auto atoms = near0(c, 6.0);
    auto syms = env::site_symmetries(c.box, c.atoms, 0, atoms);
    bool all_signed = true;
    for (auto const& s : syms) all_signed = all_signed && signed_permutation(s.O);
    std::size_t k0 = 0;   // atom 0 is the first of `atoms`
    std::size_t n111 = expand(syms, hop(atoms.size(), k0, Vec{h, h, h}));
    std::size_t n100 = expand(syms, hop(atoms.size(), k0, Vec{A, 0, 0}));
    std::printf("perfect: %zu symmetries, signed permutations %d, <111> hop +%zu, <100> hop +%zu\n", syms.size(),
                all_signed ? 1 : 0, n111, n100);
    REQUIRE(syms.size() == 48);
    REQUIRE(all_signed);
    REQUIRE(n111 == 7);
    REQUIRE(n100 == 5);
  }
  {
    Cell c = lattice(true, 0.0);
    auto atoms = near0(c, 6.0);
    auto syms = env::site_symmetries(c.box, c.atoms, 0, atoms);
    std::size_t into = expand(syms, hop(atoms.size(), 0, Vec{h, h, h}));
    std::size_t other = expand(syms, hop(atoms.size(), 0, Vec{h, h, -h}));
    std::printf("vacancy neighbour: %zu symmetries, hop into the vacancy +%zu, hop along [1 1 -1] +%zu\n", syms.size(),
                into, other);
    REQUIRE(syms.size() == 6);
    REQUIRE(into == 0);
    REQUIRE(other == 2);
  }
  {
    Cell c = lattice(false, 0.2);
    auto atoms = near0(c, 6.0);
    auto syms = env::site_symmetries(c.box, c.atoms, 0, atoms);
    std::size_t n111 = expand(syms, hop(atoms.size(), 0, Vec{h, h, h}));
    bool identity = syms.size() == 1;
    for (std::size_t i = 0; identity && i < syms[0].perm.size(); ++i) identity = syms[0].perm[i] == i;
    std::printf("distorted: %zu symmetries (identity %d), <111> hop +%zu\n", syms.size(), identity ? 1 : 0, n111);
    REQUIRE(identity);
    REQUIRE(n111 == 0);
  }
  if (failures) return 1;
  std::printf("symmetry ok\n");
  return 0;
}
