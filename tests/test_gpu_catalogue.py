"""The catalogue's device path (fingerprints and matching on the GPU, environments from the library's
neighbour list exported at r_env) finds the same new environments as its host path (k-d tree, numpy)."""
import numpy as np
import pytest

import openfly_amd as fa
from openfly_amd import systems
from openfly_amd.catalogue import Catalogue

pytestmark = pytest.mark.gpu


def test_device_path_equals_host_path(ctx):
    torch = pytest.importorskip("torch")
    s = systems.fe_defects_20(3)
    print(f"case catalogue: {s.n} atoms", flush=True)
    host = Catalogue()
    a = host.rebuild(s.pos, s.types, s.frozen, s.basis)
    dev = Catalogue()
    nl = fa.List(ctx, fa.Box(s.basis), dev.r_env)
    nl.rebuild(s.pos)
    off, j, _, dr = nl.pairs(dev.r_env)
    b = dev.rebuild(s.pos, s.types, s.frozen, s.basis, pairs=(off, j, dr), device=torch.device("cuda", 0))
    print(f"  new environments: host {len(a)}, device {len(b)}")
    assert a == b and len(a) > 1
    # a second look at the same system finds nothing new
    assert dev.rebuild(s.pos, s.types, s.frozen, s.basis, pairs=(off, j, dr), device=torch.device("cuda", 0)) == []
