// Drop-in check: reference-style code (same includes, namespaces and calls as the reference's own EAM
// known-answer test, tests/potential/EAM/eam.cpp:36-147, and its dimer/minimiser call shapes) compiled
// against the GPU facade and linked to libflygpu.so.  Exit status 0 on success.
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include <unistd.h>

#include "libfly/gpu/facade.hpp"

using namespace fly;

static int failures = 0;
#define REQUIRE(c)                                                  \
  do {                                                              \
    if (!(c)) {                                                     \
      std::fprintf(stderr, "REQUIRE failed: %s (line %d)\n", #c, __LINE__); \
      ++failures;                                                   \
    }                                                               \
  } while (0)

struct Cell {
  system::Box box;
  system::TypeMap<> map;
  system::SoA<Position, Frozen, TypeID, PotentialGradient> atoms;
};

// make_super: 7^3 bcc at the T=300 K lattice parameter, site 1 erased.
static Cell make_super(bool erase = true) {
  double T = 300.0;
  double V = 11.64012 + T * (9.37798e-5 + T * (3.643134e-7 + T * (1.851593e-10 + T * 5.669148e-14)));
  double a = std::pow(2 * V, 1.0 / 3);
  std::vector<Vec> lat;
  for (int k = 0; k < 7; ++k)
    for (int j = 0; j < 7; ++j)
      for (int i = 0; i < 7; ++i)
        for (double off : {0.0, 0.5}) lat.push_back(Vec{(i + off) * a, (j + off) * a, (k + off) * a});
  if (erase) lat.erase(lat.begin() + 1);
  Mat basis = Mat::Zero();
  for (int i = 0; i < 3; ++i) basis(i, i) = a * 7;
  Cell c{system::Box(basis, Arr<bool>::Constant(true)), system::TypeMap<>(2),
         system::SoA<Position, Frozen, TypeID, PotentialGradient>(static_cast<std::int64_t>(lat.size()))};
  c.map.set(0, "Fe");
  c.map.set(1, "H");
  for (std::int64_t i = 0; i < c.atoms.size(); ++i) {
    c.atoms(r_, i) = lat[static_cast<std::size_t>(i)];
    c.atoms(id_, i) = 0;
    c.atoms(fzn_, i) = false;
  }
  return c;
}

int main(int argc, char** argv) {
  char const* path = argc > 1 ? argv[1] : "tests/golden/wen.eam.fs";

  // ---- EAM compute (tests/potential/EAM/eam.cpp:99-147) ----
  {
    auto data = std::make_shared<potential::DataEAM>(potential::DataEAM::Options{}, std::ifstream{path});
    Cell cell = make_super();
    auto pot = std::make_unique<potential::EAM>(cell.map, data);
    neigh::List nl(cell.box, pot->r_cut());
    nl.rebuild(cell.atoms, 4);
    double E0 = pot->energy(cell.atoms, nl, 4);
    std::printf("E0 = %.13f\n", E0);
    REQUIRE(std::abs(E0 - -2748.5339306065985) < 1e-10 * 2748.5339306065985);
    pot->gradient(cell.atoms, cell.atoms, nl, 4);
    Vec head = cell.atoms(g_, 0);
    REQUIRE(gnorm(head - Vec::Constant(-0.19182439085483005)) < 1e-10);

    // for_neighbours: 685 atoms, full lists, every real pair symmetric.
    std::int64_t count = 0;
    nl.for_neighbours(0, pot->r_cut(), [&](std::int64_t, double r, Vec const&) {
      REQUIRE(r < pot->r_cut());
      ++count;
    });
    REQUIRE(count > 0);

    system::Hessian H;
    // A small Hessian through the facade would be 3N HVPs; checked on the GPU tests instead.
    (void)H;
  }

  // ---- LBFGS + Dimer on a vacancy (call shapes of find.cpp:983-1009 / skmc.hpp:247-262) ----
  {
    auto data = std::make_shared<potential::DataEAM>(potential::DataEAM::Options{}, std::ifstream{path});
    Cell cell = make_super();
    potential::Generic pot{potential::EAM(cell.map, data)};
    minimise::LBFGS minimiser({}, cell.box);
    system::SoA<Position, PotentialGradient> relaxed(cell.atoms.size());
    bool failed = minimiser.minimise(relaxed, cell.atoms, pot, 1);
    REQUIRE(!failed);
    double g2 = 0;
    for (std::size_t k = 0; k < relaxed[g_].size(); ++k) g2 += relaxed[g_][k] * relaxed[g_][k];
    REQUIRE(std::sqrt(g2) < 1e-6);
    // LBFGS::Options::fout (lbfgs.cpp:75-77): the same relaxation writing one frame per iteration, the
    // first the input, the last the relaxed positions.
    {
      std::string const traj = "/tmp/facade_lbfgs_" + std::to_string(static_cast<long>(::getpid())) + ".gsd";
      std::uint64_t frames = 0;
      {
        io::BinaryFile file(traj, io::create);
        minimise::LBFGS::Options lo;
        lo.fout = &file;
        minimise::LBFGS traced(lo, cell.box);
        system::SoA<Position, PotentialGradient> again(cell.atoms.size());
        REQUIRE(!traced.minimise(again, cell.atoms, pot, 1));
        frames = file.n_frames();
        REQUIRE(frames >= 2);
        system::SoA<Position> first(cell.atoms.size()), last(cell.atoms.size());
        file.read_to(0, r_, first);
        file.read_to(frames - 1, r_, last);
        bool same = true;
        for (std::size_t k = 0; k < first[r_].size(); ++k) {
          same = same && first[r_][k] == cell.atoms[r_][k] && last[r_][k] == again[r_][k] &&
                 again[r_][k] == relaxed[r_][k];
        }
        REQUIRE(same);
      }
      std::remove(traj.c_str());
      std::printf("minimiser trajectory (Options::fout): %llu frames\n", static_cast<unsigned long long>(frames));
    }

    // Start: nudge the atom next to the vacancy towards it, axis along the hop.
    system::SoA<Position, Axis, TypeID, Frozen> in(cell.atoms.size());
    for (std::int64_t i = 0; i < cell.atoms.size(); ++i) {
      in(r_, i) = Vec(relaxed(r_, i));
      in(ax_, i) = Vec::Zero();
      in(id_, i) = 0;
      in(fzn_, i) = false;
    }
    // atom 0 sits at the origin; the vacancy (erased site 1) is at a(1/2,1/2,1/2).
    Vec start = in(r_, 0);
    in(r_, 0) = start + Vec::Constant(0.4);
    in(ax_, 0) = Vec::Constant(1.0 / std::sqrt(3.0));
    saddle::Dimer dimer({}, {}, cell.box);
    system::SoA<Position, Axis> out(cell.atoms.size());
    system::SoA<Position> in_min(cell.atoms.size());
    for (std::int64_t i = 0; i < cell.atoms.size(); ++i) in_min(r_, i) = Vec(relaxed(r_, i));
    auto exit = dimer.find_sp(out, in, in_min, pot, {}, 0.0, 0, 1);
    std::printf("dimer exit %d steps %d grads %ld\n", static_cast<int>(exit), dimer.last_stats().steps,
                static_cast<long>(dimer.last_stats().grads));
    REQUIRE(exit == saddle::Dimer::success);
    // Options::fout (dimer.cpp:88-90): the same search writing its trajectory to a GSD file, one frame of
    // centre positions per iteration, from the start to the saddle.
    {
      std::string const traj = "/tmp/facade_dimer_" + std::to_string(static_cast<long>(::getpid())) + ".gsd";
      std::uint64_t frames = 0;
      {
        io::BinaryFile file(traj, io::create);
        saddle::Dimer::Options dopt;
        dopt.fout = &file;
        saddle::Dimer traced(dopt, {}, cell.box);
        system::SoA<Position, Axis> out2(cell.atoms.size());
        REQUIRE(traced.find_sp(out2, in, in_min, pot, {}, 0.0, 0, 1) == exit);
        frames = file.n_frames();
        REQUIRE(frames == static_cast<std::uint64_t>(traced.last_stats().steps) + 1);
        system::SoA<Position> first(cell.atoms.size()), last(cell.atoms.size());
        file.read_to(0, r_, first);
        file.read_to(frames - 1, r_, last);
        bool same = true;
        for (std::size_t k = 0; k < first[r_].size(); ++k) {
          same = same && first[r_][k] == in[r_][k] && last[r_][k] == out2[r_][k] && out2[r_][k] == out[r_][k];
        }
        REQUIRE(same);
      }
      std::remove(traj.c_str());
      std::printf("dimer trajectory (Options::fout): %llu frames\n", static_cast<unsigned long long>(frames));
    }
    neigh::List nl(cell.box, pot.r_cut());
    system::SoA<Position, TypeID, Frozen> sp(cell.atoms.size());
    for (std::int64_t i = 0; i < cell.atoms.size(); ++i) {
      sp(r_, i) = Vec(out(r_, i));
      sp(id_, i) = 0;
      sp(fzn_, i) = false;
    }
    nl.rebuild(sp, 1);
    double Esp = pot.energy(sp, nl, 1);
    system::SoA<Position, TypeID, Frozen> mn(cell.atoms.size());
    for (std::int64_t i = 0; i < cell.atoms.size(); ++i) {
      mn(r_, i) = Vec(relaxed(r_, i));
      mn(id_, i) = 0;
      mn(fzn_, i) = false;
    }
    nl.rebuild(mn, 1);
    double Emin = pot.energy(mn, nl, 1);
    std::printf("vacancy hop barrier %.6f eV\n", Esp - Emin);
    REQUIRE(Esp - Emin > 0.5 && Esp - Emin < 0.8);

    // Rotor::eff_gradient at the saddle (call shape of dimer.cpp:64): negative curvature, |g_eff| = |g|
    // (with count_frozen > 0 there is no translation projection, so g_eff is an exact reflection), and the
    // list is back where it was (same gradient after the call).
    saddle::Rotor rotor({});
    neigh::List rnl(cell.box, rotor.r_cut(pot));
    rnl.rebuild(sp, 1);
    system::SoA<PotentialGradient> g_before(cell.atoms.size()), g_after(cell.atoms.size()), eff(cell.atoms.size());
    pot.gradient(g_before, sp, rnl, 1);
    system::SoA<Axis> axis(cell.atoms.size());
    for (std::int64_t i = 0; i < cell.atoms.size(); ++i) axis(ax_, i) = Vec(out(ax_, i));
    double curv = rotor.eff_gradient(eff, axis, sp, pot, rnl, 1, 1);
    pot.gradient(g_after, sp, rnl, 1);
    double ge = 0, gb = 0, dmax = 0;
    for (std::size_t k = 0; k < eff[g_].size(); ++k) {
      ge += eff[g_][k] * eff[g_][k];
      gb += g_before[g_][k] * g_before[g_][k];
      dmax = std::max(dmax, std::abs(g_after[g_][k] - g_before[g_][k]));
    }
    std::printf("rotor curvature %.6f, |g_eff| %.3e, |g| %.3e\n", curv, std::sqrt(ge), std::sqrt(gb));
    REQUIRE(curv < 0);
    REQUIRE(std::abs(std::sqrt(ge) - std::sqrt(gb)) < 1e-9 + 1e-6 * std::sqrt(gb));
    REQUIRE(dmax < 1e-10);   // positions return to x up to the rounding of x + d - d
  }

  // ---- Master::find_batch on a vacancy (find.cpp:684-755): 16 searches at a vacancy neighbour, the site's
  //      cache growing with the converged saddles, pathways with a THz-range prefactor ----
  {
    auto data = std::make_shared<potential::DataEAM>(potential::DataEAM::Options{}, std::ifstream{path});
    Cell cell = make_super();
    potential::Generic pot{potential::EAM(cell.map, data)};
    minimise::LBFGS minimiser({}, cell.box);
    system::SoA<Position, PotentialGradient> relaxed(cell.atoms.size());
    REQUIRE(!minimiser.minimise(relaxed, cell.atoms, pot, 1));
    system::SoA<Position, TypeID, Frozen> in(cell.atoms.size());
    for (std::int64_t i = 0; i < cell.atoms.size(); ++i) {
      in(r_, i) = Vec(relaxed(r_, i));
      in(id_, i) = 0;
      in(fzn_, i) = false;
    }
    saddle::Master::Options mopt;
    mopt.batch_size = 16;
    mopt.active_radius = 6.0;
    mopt.seed = 12345;
    minimise::LBFGS::Options lo;
    lo.f2norm = 1e-5;
    minimise::LBFGS relaxer(lo, cell.box);
    saddle::Dimer::Options dopt;
    dopt.f2norm = 1e-4;
    saddle::Dimer dimer(dopt, {}, cell.box);
    system::TypeMap<Mass> mass_map(2);
    saddle::Master master(mopt, cell.box, mass_map, pot, relaxer, dimer);
    std::vector<system::SoA<Position>> cache;
    std::vector<saddle::Master::Mech> mechs;
    auto found = master.find_batch(0, 0, in, cache, &mechs);
    int converged = 0, pathways = 0, fresh = 0;
    for (auto const& f : found) {
      if (f.exit == saddle::Dimer::success) ++converged;
      if (f.mech.code == 0) {
        ++pathways;
        fresh += f.is_new;
        REQUIRE(f.kinetic_pre > 1e11 && f.kinetic_pre < 1e15);
      }
    }
    std::printf("Master::find_batch: %d of %zu converged, %d pathways (%d distinct), cache %zu\n", converged,
                found.size(), pathways, fresh, cache.size());
    // Symmetric copies (Master::append_syms, find.cpp:1279-1303): atom 0 sits beside the vacancy, so its
    // environment keeps C3v's six operations about the vacancy direction.
    auto syms = env::site_symmetries(cell.box, in, 0, master.site_atoms(0, in));
    std::vector<saddle::Master::Mech> all = mechs;
    std::size_t copies = 0;
    for (auto const& m : mechs) copies += master.append_syms(syms, m, all);
    std::printf("append_syms: %zu symmetries, %zu distinct pathways + %zu symmetric copies\n", syms.size(),
                mechs.size(), copies);
    REQUIRE(syms.size() == 6);
    REQUIRE(all.size() == mechs.size() + copies);
    REQUIRE(found.size() == 16);
    REQUIRE(static_cast<int>(cache.size()) == converged);
    REQUIRE(converged >= 1);
    REQUIRE(static_cast<int>(mechs.size()) == fresh && fresh <= pathways);
    // Every pathway is known to the site's list afterwards: a second look finds none new.
    for (auto const& m : mechs) REQUIRE(!master.is_new_mech(m, mechs));
  }

  // ---- Spline::OutOfBounds surfaces as the reference's exception type ----
  {
    auto data = std::make_shared<potential::DataEAM>(potential::DataEAM::Options{}, std::ifstream{path});
    system::TypeMap<> map(2);
    map.set(0, "Fe");
    map.set(1, "H");
    potential::EAM pot(map, data);
    Mat basis = Mat::Zero();
    for (int i = 0; i < 3; ++i) basis(i, i) = 30.0;
    system::SoA<Position, TypeID, Frozen, PotentialGradient> a(40);
    for (int k = 0; k < 40; ++k) {
      a(r_, k) = Vec{10.0 + 0.05 * k, 10.0, 10.0};
      a(id_, k) = 0;
      a(fzn_, k) = false;
    }
    neigh::List nl(system::Box(basis, Arr<bool>::Constant(true)), pot.r_cut());
    nl.rebuild(a, 1);
    bool caught = false;
    try {
      pot.gradient(a, a, nl, 1);
    } catch (Spline::OutOfBounds const&) {
      caught = true;
    }
    REQUIRE(caught);
  }

  std::printf(failures ? "FAILED (%d)\n" : "facade ok\n", failures);
  return failures ? 1 : 0;
}
