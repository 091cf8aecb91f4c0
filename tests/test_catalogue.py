"""Local-environment catalogue (env::Catalogue::rebuild, catalogue.cpp:66-200), pinned by lattice symmetry.

A perfect bcc crystal has one environment; a translated copy or one rotated by a cubic symmetry
operation adds none; a vacancy adds exactly its neighbour shells within r_env = 5.2 A (first to fifth:
a sqrt(3)/2, a, a sqrt(2), a sqrt(11)/2, a sqrt(3)); one frozen atom (another colour) adds its own
environment and those same five shells around it; displacements well inside delta_max add nothing and a
0.5 A displacement adds environments.  The host and torch fingerprint paths give the same new atoms.
A dumped and reloaded catalogue answers as the original.
"""
import numpy as np
import pytest

from openfly_amd import systems
from openfly_amd.catalogue import Catalogue

A = 2.855324
CELLS = 6


@pytest.fixture(scope="module")
def crystal():
    pos = systems.bcc_sites(A, CELLS)
    basis = np.diag([CELLS * A] * 3)
    n = len(pos)
    return pos, basis, np.zeros(n, np.uint32), np.zeros(n, np.uint8)


def shells_of(pos, basis, centre_pos, idx):
    L = np.diag(basis)
    d = pos[idx] - centre_pos
    d -= L * np.round(d / L)
    return np.round(np.linalg.norm(d, axis=1), 3)


def test_perfect_crystal_and_symmetry(crystal):
    pos, basis, t, f = crystal
    c = Catalogue()
    assert c.rebuild(pos, t, f, basis) == [0]
    assert c.rebuild(pos, t, f, basis) == []
    assert c.rebuild(pos + np.array([0.3, -1.1, 2.0]), t, f, basis) == []
    rot = pos[:, [1, 0, 2]].copy()          # 90 degree rotation about z (a cubic symmetry)
    rot[:, 0] = CELLS * A - rot[:, 0]
    assert c.rebuild(rot, t, f, basis) == []
    assert c.size == 1


def test_vacancy_adds_its_neighbour_shells(crystal):
    pos, basis, t, f = crystal
    c = Catalogue()
    c.rebuild(pos, t, f, basis)
    vac = np.delete(pos, 0, axis=0)
    new = c.rebuild(vac, t[1:], f[1:], basis)
    expect = np.round(A * np.array([np.sqrt(3) / 2, 1.0, np.sqrt(2), np.sqrt(11) / 2, np.sqrt(3)]), 3)
    assert len(new) == 5
    assert np.array_equal(np.sort(shells_of(vac, basis, pos[0], new)), expect)
    assert len(Catalogue().rebuild(vac, t[1:], f[1:], basis)) == 6   # with the bulk environment


def test_frozen_atom_is_another_colour(crystal):
    pos, basis, t, f = crystal
    c = Catalogue()
    c.rebuild(pos, t, f, basis)
    fz = f.copy()
    fz[7] = 1
    new = c.rebuild(pos, t, fz, basis)
    assert 7 in new
    others = [i for i in new if i != 7]
    expect = np.round(A * np.array([np.sqrt(3) / 2, 1.0, np.sqrt(2), np.sqrt(11) / 2, np.sqrt(3)]), 3)
    assert len(new) == 6
    assert np.array_equal(np.sort(shells_of(pos, basis, pos[7], others)), expect)


def test_tolerance(crystal):
    pos, basis, t, f = crystal
    c = Catalogue()
    c.rebuild(pos, t, f, basis)
    rng = np.random.default_rng(3)
    assert c.rebuild(pos + rng.normal(0, 0.005, pos.shape), t, f, basis) == []
    moved = pos.copy()
    moved[11] += np.array([0.5, 0.0, 0.0])
    new = c.rebuild(moved, t, f, basis)
    assert 11 in new and len(new) > 1


def test_host_and_torch_paths_agree():
    torch = pytest.importorskip("torch")
    s = systems.fe_defects_20(3)
    a = Catalogue().rebuild(s.pos, s.types, s.frozen, s.basis)
    b = Catalogue().rebuild(s.pos, s.types, s.frozen, s.basis, device=torch.device("cpu"))
    assert a == b and len(a) > 1


def test_atom_subset_finds_the_changed_environments(crystal):
    # A kMC step changes environments near the hop only: looking at the atoms within 6 A of a new vacancy
    # finds the same five shells as the whole cell does.
    pos, basis, t, f = crystal
    c = Catalogue()
    c.rebuild(pos, t, f, basis)
    vac = np.delete(pos, 0, axis=0)
    L = np.diag(basis)
    d = vac - pos[0]
    d -= L * np.round(d / L)
    near = np.nonzero(np.linalg.norm(d, axis=1) < 6.0)[0]
    new = c.rebuild(vac, t[1:], f[1:], basis, atoms=near)
    assert set(new) <= set(near.tolist()) and len(new) == 5
    assert c.rebuild(vac, t[1:], f[1:], basis) == []


def test_dump_and_load(crystal, tmp_path):
    # Catalogue::dump / Catalogue(opt, fin) (catalogue.hpp:314-348): a reloaded catalogue knows the same
    # environments (same size, same new atoms on a later system, none on the systems it has seen) and
    # refuses contradicting options and files that are not such an archive.
    from openfly_amd.flygpu import FlyError
    pos, basis, t, f = crystal
    c = Catalogue()
    c.rebuild(pos, t, f, basis)
    vac = np.delete(pos, 0, axis=0)
    c.rebuild(vac, t[1:], f[1:], basis)
    path = tmp_path / "cat.npz"
    c.dump(path)
    d = Catalogue.load(path)
    assert d.size == c.size == 6
    assert d.rebuild(vac, t[1:], f[1:], basis) == [] and d.rebuild(pos, t, f, basis) == []
    rng = np.random.default_rng(3)
    moved = vac + rng.normal(0, 0.3, vac.shape) * (np.arange(len(vac)) < 40)[:, None]
    new_d = d.rebuild(moved, t[1:], f[1:], basis)
    new_c = c.rebuild(moved, t[1:], f[1:], basis)
    assert new_d == new_c and len(new_d) > 0
    assert d.size == c.size
    with pytest.raises(FlyError, match="contradicting .r_env"):
        Catalogue.load(path, r_env=5.0)
    with pytest.raises(FlyError, match="contradicting .delta_max"):
        Catalogue.load(path, delta_max=0.2)
    bad = tmp_path / "bad.npz"
    bad.write_bytes(b"not an archive")
    with pytest.raises(FlyError, match="bad binary"):
        Catalogue.load(bad)
