"""The procedural benchmark scene and the BVH, checked on the host.

The traversal runs on the host with the same code the trace kernel runs
(rpd_scene_query), so the BVH is checked against the exhaustive loop over
every primitive before any GPU run: identical hit, distance, primitive and
sphere flag for closest-hit and occlusion queries.
"""
import ctypes as C

import numpy as np
import pytest

import rpd_engine as eng


def _rays(rng, n, spread=4.0):
    o = rng.uniform(-spread, spread, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # a quarter aimed at the object, some axis-aligned (zero direction components)
    aim = rng.random(n) < 0.25
    tgt = rng.uniform(-0.5, 0.5, (n, 3))
    d[aim] = (tgt[aim] - o[aim]) / np.linalg.norm(tgt[aim] - o[aim], axis=1, keepdims=True)
    ax = rng.random(n) < 0.05
    d[ax] = 0.0
    d[ax, rng.integers(0, 3, ax.sum())] = rng.choice([-1.0, 1.0], ax.sum())
    return o, d.astype(np.float32)


@pytest.mark.parametrize("level,spheres", [(3, 40), (4, 64), (5, 8)])
def test_bvh_matches_exhaustive_loop(level, spheres):
    pc = eng.procedural_config(subdivision_level=level, n_spheres=spheres)
    sc = eng.HostScene(pc)
    try:
        nodes, depth = sc.bvh_info()
        assert nodes > 0 and 1 <= depth <= 16
        rng = np.random.default_rng(level)
        n = 20000
        o, d = _rays(rng, n)
        tmax = np.where(rng.random(n) < 0.5, np.inf, rng.uniform(0.1, 6.0, n)).astype(np.float32)
        occ = (rng.random(n) < 0.4).astype(np.int32)
        a = sc.query(o, d, tmax, occ, use_bvh=True)
        b = sc.query(o, d, tmax, occ, use_bvh=False)
        np.testing.assert_array_equal(a[0], b[0], err_msg="hit")   # occlusion queries report only this
        cl = occ == 0
        for x, y, name in zip(a[1:], b[1:], ("t", "index", "is_sphere")):
            np.testing.assert_array_equal(x[cl], y[cl], err_msg=name)
        assert a[0].mean() > 0.2   # the rays do hit things
        assert (a[3][a[0] == 1] == 1).any() or spheres == 0
    finally:
        sc.close()


def test_procedural_scene_structure():
    pc = eng.procedural_config(subdivision_level=3, n_spheres=16)
    tris, sph, mats, lights, env, n = eng.procedural_arrays(pc)
    nt, ns, nm, nl = n
    assert nt == 20 * 4 ** 3 + 8 and ns == 16 and nl == 4 and nm == 2 + 16
    v = np.array([[list(t.v0), list(t.v1), list(t.v2)] for t in tris[:nt - 8]], np.float32)
    # watertight: every edge of the closed icosphere is shared by exactly two faces, reversed
    edges = {}
    for f in v:
        for k in range(3):
            a, b = tuple(f[k]), tuple(f[(k + 1) % 3])
            edges[(a, b)] = edges.get((a, b), 0) + 1
    assert all(c == 1 for c in edges.values())
    assert all((b, a) in edges for (a, b) in edges)
    # stored normals are unit and point away from the centre
    nrm = np.array([list(t.normal) for t in tris[:nt - 8]], np.float32)
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1.0, atol=1e-5)
    assert (np.einsum("ij,ij->i", nrm, v.mean(axis=1)) > 0).all()
    r = np.linalg.norm(v.reshape(-1, 3), axis=1)
    assert r.min() > 1 - pc.displacement - 1e-5 and r.max() < 1 + pc.displacement + 1e-5 and r.std() > 0.01
    # sphere materials: GGX, roughness and albedo in the stated ranges
    for s in sph[:ns]:
        m = mats[s.material_id]
        assert m.kind == 1 and m.distribution == 0 and 0.05 <= m.roughness <= 0.8
        assert all(0.1 <= a <= 0.9 for a in m.albedo)
    # light quads face the object and are linked to their lights
    for t in tris[nt - 8:]:
        assert 0 <= t.light_id < 4 and max(t.emission) > 0
        c = (np.array(t.v0) + np.array(t.v1) + np.array(t.v2)) / 3
        assert np.dot(np.array(t.normal), -c) > 0
    assert max(env) > 0
    # same configuration, same bits
    t2 = eng.procedural_arrays(pc)[0]
    assert bytes(tris) == bytes(t2)


def test_full_size_procedural_scene_counts():
    pc = eng.procedural_config()
    assert pc.seed == 7 and pc.subdivision_level == 8 and pc.n_spheres == 256
    n = [C.c_int() for _ in range(4)]
    eng.check(eng.lib().rpd_procedural_scene_arrays(C.byref(pc), None, C.byref(n[0]), None, C.byref(n[1]), None,
                                                    C.byref(n[2]), None, C.byref(n[3]), None))
    assert n[0].value == 20 * 4 ** 8 + 8 and n[1].value == 256 and n[3].value == 4
