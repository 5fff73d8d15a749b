"""The HIP path against the committed golden fixtures (tests/golden/): the
small meshes' refined ids, LAK form and positions (within 1e-5 of the
bounding-box diagonal of the stored FP64 positions), and the larger
configurations' refined-id hashes, both schemes."""
import hashlib
import json
import os

import numpy as np
import pytest

from slak_amd import capi, meshgen as g
from slak_amd.mesh import subdivide_host
from tests.golden import make_golden as mk

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
POS_TOL = 1e-5


def load(name):
    with np.load(os.path.join(HERE, f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", sorted(mk.SMALL))
def test_gpu_matches_small_fixture(name):
    f = load(name)
    m = g.HostMesh(int(f["in_type"][0]), f["in_ptr"].astype(np.int32), f["in_ids"].astype(np.int32),
                   f["in_vals"].astype(np.int32), f["in_verts"].astype(np.float32))
    levels = int(f["levels"][0])
    for tag, scheme in (("lak", capi.SLAK_SCHEME_LAK), ("slak", capi.SLAK_SCHEME_SLAK)):
        out, _ = subdivide_host(m, levels, scheme)
        np.testing.assert_array_equal(out.ids, f[f"{tag}_ids"])
        if tag == "lak":
            np.testing.assert_array_equal(out.ptr, f["lak_ptr"])
            np.testing.assert_array_equal(out.vals, f["lak_vals"])
        err = float(np.abs(out.verts.astype(np.float64) - f[f"{tag}_P"]).max()) / m.bbox_diag()
        assert err <= POS_TOL, err


@pytest.mark.parametrize("name", sorted(mk.LARGE))
def test_gpu_topology_matches_manifest(name):
    with open(os.path.join(HERE, "manifest.json")) as fh:
        want = json.load(fh)[name]
    make, levels = mk.LARGE[name]
    m = make()
    for tag, scheme in (("lak", capi.SLAK_SCHEME_LAK), ("slak", capi.SLAK_SCHEME_SLAK)):
        out, _ = subdivide_host(m, levels, scheme)
        assert out.nfaces == want[tag]["nf"] and out.nverts == want[tag]["nv"]
        assert hashlib.sha256(np.ascontiguousarray(out.ids).tobytes()).hexdigest() == want[tag]["ids_sha256"]
