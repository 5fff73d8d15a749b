"""Seeded generators: determinism, counts and manifoldness of every config."""
import numpy as np

from slak_amd import meshgen as g


def test_splitmix_stream_is_counter_based():
    a = g.draws(7, 3, 1, 10)
    b = g.draws(7, 3, 1, 20)
    np.testing.assert_array_equal(a, b[:10])
    assert not np.array_equal(g.draws(7, 3, 1, 5), g.draws(7, 4, 1, 5))
    u = g.uniforms(1, 0, 0, 100000)
    assert 0.0 <= u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01
    z = g.gaussians(1, 0, 0, 100000)
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1.0) < 0.02


def test_configs_deterministic_and_sized():
    c1 = g.config1()
    assert (c1.nfaces, c1.nverts, c1.nnz) == (512, 512, 2048)
    assert c1.digest() == g.config1().digest()
    c2 = g.config2()
    assert c2.type == g.POLY and c2.nverts == 256 * 128
    orders = np.bincount(np.diff(c2.ptr))
    assert orders[5] == 256  # one pentagon per grid column
    assert c2.digest() == g.config2().digest()


def _directed_edges(h):
    ptr, ids = h.face_ptr(), h.ids
    he = set()
    for f in range(h.nfaces):
        fv = ids[ptr[f]:ptr[f + 1]]
        for k in range(len(fv)):
            e = (int(fv[k]), int(fv[(k + 1) % len(fv)]))
            assert e not in he
            he.add(e)
    return he


def test_mixed_torus_is_closed_manifold():
    h = g.torus_mixed(24, 12, 0.3, pentagons=True, seed=9)
    he = _directed_edges(h)
    assert all((w, v) in he for v, w in he)  # closed: every half-edge has a twin
    val = np.bincount(h.ids, minlength=h.nverts)
    assert val.min() >= 3 and val.max() <= 8


def test_batch_and_request_generators():
    spec = g.config4(n_meshes=32, seed=4)
    assert len(spec.meshes) == 32 and all(1 <= l <= 4 for l in spec.levels)
    for m in spec.meshes[:8]:
        _directed_edges(m)
    m, lvl = g.request_mesh(5, 3)
    assert 1 <= lvl <= 3 and 100 * 0.5 <= m.nfaces


def test_random_poly_torus_is_a_manifold_with_isolated_holes():
    for seed in range(5):
        m = g.random_poly_torus(15, 10, seed=seed, p_merge=0.2, p_hole=0.7)
        orders = np.diff(m.ptr)
        assert set(orders.tolist()) <= {3, 4, 6}
        faces = [m.ids[m.ptr[f]:m.ptr[f + 1]].tolist() for f in range(m.nfaces)]
        he = set()
        for fv in faces:
            for v, w in zip(fv, fv[1:] + fv[:1]):
                assert (v, w) not in he, "directed edge used twice"
                he.add((v, w))
        boundary = [(v, w) for v, w in he if (w, v) not in he]
        starts = [v for v, _ in boundary]
        assert len(starts) == len(set(starts)), "a vertex on two boundary loops"
        assert np.bincount(m.ids, minlength=m.nverts).min() >= 2
    assert g.random_poly_torus(15, 10, seed=3).nfaces == g.random_poly_torus(15, 10, seed=3).nfaces
