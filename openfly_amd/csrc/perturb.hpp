// Seeded start states of saddle searches: Master::perturb (find.cpp:1147-1214) on Xoshiro256**
// (random.hpp:96-161), host side.
#pragma once

#include <cstdint>
#include <functional>
#include <vector>

#include "common.hpp"

namespace flyg {

  // Xoshiro256** (Blackman & Vigna, "Scrambled linear pseudorandom number generators", 2021), the
  // generator of the reference's fly::Xoshiro.
  struct Xoshiro256 {
    uint64_t s[4];
    uint64_t next();
    double uniform();   // [0, 1) from the top 53 bits
  };

  // Gaussian N(mean, sd) by the Marsaglia polar method (Marsaglia & Bray, SIAM Review 6 (1964) 260),
  // each sampler keeping the second variate of a pair for its next draw.
  struct Gaussian {
    double mean, sd;
    bool have = false;
    double saved = 0;
    double operator()(Xoshiro256& g);
  };

  // Master::perturb over an orthorhombic box: sd ~ N(stddev, stddev / 3) redrawn until positive; every
  // non-frozen atom j with |r_j - r_centre| < r_pert (each image, in ascending (j, r) order) moves by
  // N(0, sd) (1 - r / r_pert) per component and gets N(0, 1) axis components; the centre moves by N(0, sd)
  // and gets N(0, 1) axis components; the axis is normalised.  Sparse result: the displaced atoms in
  // ascending order, their new positions and axis components (3 each; every other atom keeps its position
  // and a zero axis).  `cand` (optional) lists the atoms to consider, a superset of those within r_pert
  // (ascending); null considers all n.  prng advances; returns sd.
  struct Kick {
    std::vector<long> atoms;
    std::vector<double> pos, axis;
    double sd = 0;
  };
  void perturb(BoxDesc const& box, double const* pos, long n, long centre, std::vector<long> const* cand,
               std::function<bool(long)> const& frozen, double stddev, double r_pert, Xoshiro256& prng, Kick& out);

}  // namespace flyg
