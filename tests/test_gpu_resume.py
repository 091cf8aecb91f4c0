"""A batched search that leaves its region continues in the grown region from the dimer state it had
(positions, axis, the translation L-BFGS memory, trust radius and step count) instead of starting over.

Bar: on the 16,020-atom defect system (bcc Fe 20^3 cells, 20 vacancies, 40 interstitials, relaxed) with
96 perturbed starts, radii U(4, 9) A, FP64: with a tight initial skin some searches leave their region and
are resumed (stats.resumes = launches - 1); each of them ends as the same search does with a skin wide
enough to run in one launch: same exit, step count within 2, barrier within 1e-6 eV.  The regions differ
(so the row order and the rounding of the sums do), the dimer state carried over does not.
"""
import numpy as np
import pytest

import openfly_amd as fa
from openfly_amd import systems
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def defects(ctx):
    s = systems.fe_defects_20(3)
    pot = fa.EAM(ctx, H.WEN)
    mn, _, failed, _ = fa.LBFGS(ctx, fa.Box(s.basis), f2norm=1e-4, iter_max=5000).minimise(
        s.pos, s.types, np.zeros(s.n, np.uint8), pot)
    assert not failed
    s.pos = mn
    return s, pot


def run(ctx, s, pot, skin, centres, radii, prng):
    b = fa.Batch(ctx, pot, fa.Box(s.basis), s.pos, s.types, dimer=dict(f2norm=1e-3), precision=fa.FP64,
                 list_skin=skin, block_threads=512, lanes_per_atom=8)
    ids = b.submit_perturbed(centres, radii, prng.copy())
    b.run()
    return [b.result(i) for i in ids]


def test_region_exit_resumes(ctx, defects):
    s, pot = defects
    rng = np.random.default_rng(33)
    n = 96
    centres = rng.choice(systems.defect_adjacent_atoms(s), n, replace=True)
    radii = rng.uniform(4.0, 9.0, n)
    prng = np.random.default_rng(34).integers(1, 2 ** 63, size=(n, 4), dtype=np.uint64)
    print(f"case resume: {n} searches, tight skin vs wide skin", flush=True)
    tight = run(ctx, s, pot, 0.05, centres, radii, prng)
    wide = run(ctx, s, pot, 6.0, centres, radii, prng)
    compared = 0
    for k, (a, b) in enumerate(zip(tight, wide)):
        if a["launches"] == 1 or b["launches"] != 1:
            continue
        print(f"  search {k}: tight exit {a['exit']} steps {a['steps']} launches {a['launches']} resumes "
              f"{a['resumes']} barrier {a['barrier']:.8f} | wide exit {b['exit']} steps {b['steps']} barrier "
              f"{b['barrier']:.8f}")
        assert a["resumes"] == a["launches"] - 1
        assert a["exit"] == b["exit"]
        assert abs(a["steps"] - b["steps"]) <= 2
        if a["exit"] == 0:
            assert abs(a["barrier"] - b["barrier"]) < 1e-6
        compared += 1
    assert compared >= 1
