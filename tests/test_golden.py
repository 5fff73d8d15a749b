"""The oracle against its committed golden fixtures (tests/golden/, written
by tests/golden/make_golden.py): small meshes stored whole, larger
configurations as hashes of the refined ids and of the FP64 positions
rounded to 1e-9. A change of the oracle's numbering, order or arithmetic
shows up here before it can move the target of the GPU parity tests."""
import json
import os

import numpy as np
import pytest

from slak_amd import meshgen as g
from tests.golden import make_golden as mk

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(HERE, f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def host_mesh(f):
    ptr = f["in_ptr"].astype(np.int32)
    return g.HostMesh(int(f["in_type"][0]), ptr, f["in_ids"].astype(np.int32), f["in_vals"].astype(np.int32),
                      f["in_verts"].astype(np.float32))


@pytest.mark.parametrize("name", sorted(mk.SMALL))
def test_small_fixture(oracle, name):
    f = load(name)
    make, levels = mk.SMALL[name]
    m = make()
    # the generator still makes the stored input
    np.testing.assert_array_equal(m.face_ptr(), f["in_ptr"])
    np.testing.assert_array_equal(m.ids, f["in_ids"])
    np.testing.assert_array_equal(m.verts, f["in_verts"])
    for tag, scheme in (("lak", 0), ("slak", 1)):
        r = oracle.subdivide(host_mesh(f), int(f["levels"][0]), scheme=scheme, precision=64)
        np.testing.assert_array_equal(r.ids, f[f"{tag}_ids"])
        np.testing.assert_array_equal(r.P, f[f"{tag}_P"])
        if scheme == 0:
            np.testing.assert_array_equal(r.ptr, f["lak_ptr"])
            np.testing.assert_array_equal(r.vals, f["lak_vals"])


def manifest():
    with open(os.path.join(HERE, "manifest.json")) as fh:
        return json.load(fh)


@pytest.mark.parametrize("name", sorted(mk.LARGE))
def test_manifest_entry(oracle, name):
    want = manifest()[name]
    make, levels = mk.LARGE[name]
    assert want["levels"] == levels
    assert make().digest() == want["input_digest"]
    got = mk.manifest_entry(make, levels)
    assert got == want
