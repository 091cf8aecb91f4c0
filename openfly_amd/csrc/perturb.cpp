// Master::perturb (find.cpp:1147-1214) on the host: seeded search starts for the batched and served
// searches, and the re-perturbation after a spline-out-of-bounds failure (find.cpp:706-723).
#include "perturb.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <utility>

// Every product and sum below rounds as written (no fused multiply-add), so the draws and distances are
// reproducible on any host.
#pragma clang fp contract(off)

namespace flyg {

  uint64_t Xoshiro256::next() {
    auto rotl = [](uint64_t x, int k) { return (x << k) | (x >> (64 - k)); };
    uint64_t const result = rotl(s[1] * 5, 7) * 9;
    uint64_t const t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }

  double Xoshiro256::uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }

  double Gaussian::operator()(Xoshiro256& g) {
    if (have) {
      have = false;
      return mean + sd * saved;
    }
    double u, v, q;
    do {
      u = 2 * g.uniform() - 1;
      v = 2 * g.uniform() - 1;
      q = u * u + v * v;
    } while (q >= 1 || q == 0);
    double const m = std::sqrt(-2 * std::log(q) / q);
    saved = v * m;
    have = true;
    return mean + sd * (u * m);
  }

  void perturb(BoxDesc const& box, double const* pos, long n, long centre, std::vector<long> const* cand,
               std::function<bool(long)> const& frozen, double stddev, double r_pert, Xoshiro256& prng, Kick& out) {
    require(box.ortho, "perturb needs an orthorhombic box");
    require(centre >= 0 && centre < n, "centre out of range");
    require(!frozen(centre), "perturbation centred on a frozen atom");
    require(stddev > 0 && r_pert > 0, "stddev and r_pert must be positive");
    require(!(prng.s[0] == 0 && prng.s[1] == 0 && prng.s[2] == 0 && prng.s[3] == 0), "all-zero generator state");

    // Canonical images as the neighbour list holds them (x - L floor(x / L)); an image of atom j adds
    // +-L per periodic axis to its canonical coordinate before the difference is taken.
    auto canon = [&](double x, int k) { return x - box.ext[k] * std::floor(x * box.inv_ext[k]); };
    double const cc[3] = {canon(pos[3 * centre], 0), canon(pos[3 * centre + 1], 1), canon(pos[3 * centre + 2], 2)};
    double const rp2 = r_pert * r_pert;
    std::vector<std::pair<long, double>> nb;
    auto consider = [&](long j) {
      double d[3][3];
      int nd[3];
      for (int k = 0; k < 3; ++k) {
        double const cj = canon(pos[3 * j + k], k);
        nd[k] = 0;
        double const im[3] = {cj, cj + box.ext[k], cj - box.ext[k]};
        int const m = box.per[k] ? 3 : 1;
        for (int q = 0; q < m; ++q) {
          double const dk = im[q] - cc[k];
          if (dk * dk < rp2) d[k][nd[k]++] = dk;
        }
        if (nd[k] == 0) return;
      }
      for (int a = 0; a < nd[0]; ++a)
        for (int b = 0; b < nd[1]; ++b)
          for (int c = 0; c < nd[2]; ++c) {
            double const r2 = (d[0][a] * d[0][a] + d[1][b] * d[1][b]) + d[2][c] * d[2][c];
            bool const self = j == centre && d[0][a] == 0 && d[1][b] == 0 && d[2][c] == 0;
            if (r2 < rp2 && !self) nb.emplace_back(j, std::sqrt(r2));
          }
    };
    if (cand) {
      for (long j : *cand) {
        require(j >= 0 && j < n, "candidate index out of range");
        consider(j);
      }
    } else {
      for (long j = 0; j < n; ++j) consider(j);
    }
    std::sort(nb.begin(), nb.end());

    Gaussian normal{0., 1.};
    Gaussian prime{stddev, stddev / 3.0};
    double sd = -1;
    while (sd < 0) sd = prime(prng);
    Gaussian gauss{0, sd};
    // Displacements and axis components accumulate per atom (an atom seen through two images gets both).
    std::vector<std::pair<long, std::array<double, 6>>> acc;
    auto slot = [&](long j) -> std::array<double, 6>& {
      if (acc.empty() || acc.back().first != j) acc.push_back({j, {0, 0, 0, 0, 0, 0}});
      return acc.back().second;
    };
    for (auto const& [j, r] : nb) {
      if (frozen(j)) continue;
      auto& a = slot(j);
      for (int k = 0; k < 3; ++k) a[k] += gauss(prng) * (1. - r / r_pert);
      for (int k = 0; k < 3; ++k) a[3 + k] += normal(prng);
    }
    double cdisp[6];
    for (int k = 0; k < 3; ++k) cdisp[k] = gauss(prng);
    for (int k = 0; k < 3; ++k) cdisp[3 + k] = normal(prng);
    auto it = std::lower_bound(acc.begin(), acc.end(), centre,
                               [](auto const& e, long v) { return e.first < v; });
    if (it == acc.end() || it->first != centre) it = acc.insert(it, {centre, {0, 0, 0, 0, 0, 0}});
    for (int k = 0; k < 6; ++k) it->second[k] += cdisp[k];

    out.atoms.clear();
    out.pos.clear();
    out.axis.clear();
    out.sd = sd;
    // |axis| summed in index order over the non-zero entries (the zeros would not change it).
    double nrm = 0;
    for (auto const& e : acc)
      for (int k = 0; k < 3; ++k) nrm += e.second[3 + k] * e.second[3 + k];
    nrm = std::sqrt(nrm);
    for (auto const& e : acc) {
      out.atoms.push_back(e.first);
      for (int k = 0; k < 3; ++k) out.pos.push_back(pos[3 * e.first + k] + e.second[k]);
      for (int k = 0; k < 3; ++k) out.axis.push_back(e.second[3 + k] / nrm);
    }
  }

}  // namespace flyg
