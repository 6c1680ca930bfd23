"""The reference's built-in scenes (path_tracer/scenes/*.h) and OBJ meshes,
checked on the host.

* Structure as the reference builds each scene: primitive, material and light
  counts, emitter links, light quads whose triangles span their rect light,
  unit normals, cornell walls facing the room.
* cornell_box: the engine's arrays rendered by the oracle equal the oracle's
  own cornell restatement bit for bit (two independent restatements of
  scenes/cornell_box.h).
* Every primitive scene renders in the oracle from the engine's arrays.
* The mesh scenes place the bunny as scenes/bunny.h and bunny_dof.h say (floor
  at y = 1e-3, the target height, centred on the placement); they need the
  reference's resources/models/bunny.obj and are skipped without it.
* OBJ parsing: face formats, negative indices, polygons, errors.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_bind as ob
import rpd_engine as eng

REF_ROOT = "/root/reference"
HAVE_BUNNY = os.path.exists(os.path.join(REF_ROOT, "resources", "models", "bunny.obj"))

# (triangles, spheres, materials, lights) as scenes/*.h add them
COUNTS = {
    "cornell_box": (32, 0, 5, 1),          # light 2 + walls 10 + two boxes 10 each
    "cornell_spheres": (12, 3, 7, 1),      # light 2 + walls 10
    "veach_mis": (10, 4, 6, 4),            # ground 2 + four light quads
    "material_showcase": (8, 9, 12, 3),    # ground 2 + wall 2 + two light quads; key, fill, rim
    "caustics": (12, 4, 5, 3),             # ground 2 + walls and ceiling 8 + light quad; two point lights
    "debug": (5, 0, 3, 0),
}


def _v(a):
    return np.array(a[:], np.float64)


def _arrays(name):
    tris, sph, mats, lights, env, n = eng.builtin_scene_arrays(name)
    return tris, sph, mats, lights, env, n


@pytest.mark.parametrize("name", sorted(COUNTS))
def test_scene_structure(name):
    tris, sph, mats, lights, env, n = _arrays(name)
    assert tuple(n) == COUNTS[name]
    for k in range(n[0]):
        t = tris[k]
        assert abs(np.linalg.norm(_v(t.normal)) - 1.0) < 1e-6
        assert 0 <= t.material_id < n[2]
        em = _v(t.emission)
        if em.max() > 0:                                   # emitters link to a light of the same radiance
            assert 0 <= t.light_id < n[3]
            assert np.array_equal(_v(lights[t.light_id].emission), em)
        else:
            assert t.light_id == -1
    for k in range(n[1]):
        assert sph[k].radius > 0 and 0 <= sph[k].material_id < n[2] and sph[k].light_id == -1
    # every rect light is spanned by exactly its two linked triangles: corners o, o+u, o+u+v, o+v
    for li in range(n[3]):
        L = lights[li]
        if L.kind != eng.RPD_LIGHT_RECT:
            continue
        o, u, w = _v(L.p0), _v(L.p1), _v(L.p2)
        corners = {tuple(np.round(c, 5)) for c in (o, o + u, o + u + w, o + w)}
        linked = [tris[k] for k in range(n[0]) if tris[k].light_id == li]
        assert len(linked) == 2
        verts = {tuple(np.round(_v(getattr(t, f)), 5)) for t in linked for f in ("v0", "v1", "v2")}
        assert verts == corners


def test_cornell_walls_face_the_room():
    for name in ("cornell_box", "cornell_spheres"):
        tris, sph, mats, lights, env, n = _arrays(name)
        center = np.array([0.278, 0.2744, 0.2796])
        for k in range(n[0]):
            t = tris[k]
            mid = (_v(t.v0) + _v(t.v1) + _v(t.v2)) / 3
            if name == "cornell_spheres" or k < 12:        # the boxes of cornell_box face away from their own centre
                assert np.dot(_v(t.normal), center - mid) > 0, (name, k)


def test_engine_cornell_arrays_render_like_the_oracle_cornell():
    rc = ob.config_c0(16, 12)
    own = ob.OracleJob(rc, 99, nthreads=1, keep_records=True)
    arr = ob.OracleJob(rc, 99, nthreads=1, keep_records=True, scene=eng.builtin_scene_arrays("cornell_box"))
    counts = np.full(16 * 12, 4, np.int32)
    own.run_pass(counts, statmc=True)
    arr.run_pass(counts, statmc=True)
    np.testing.assert_array_equal(own.get("records"), arr.get("records"))


@pytest.mark.parametrize("name,pos,look", [
    ("cornell_spheres", (0.278, 0.2744, -0.8), (0, 0, 1)),
    ("veach_mis", (0.0, 2.0, 8.0), (0, -0.2, -1)),
    ("material_showcase", (0.0, 3.0, -9.0), (0, -0.3, 1)),
    ("caustics", (0.0, 3.0, -9.0), (0, -0.2, 1)),
    ("debug", (0.0, 0.0, 1.0), (0, -0.3, -1)),
])
def test_primitive_scenes_render_in_the_oracle(name, pos, look):
    rc = ob.config_c0(12, 10)
    rc.cam_pos, rc.cam_dir = pos, look
    oj = ob.OracleJob(rc, 5, nthreads=1, keep_records=True, scene=eng.builtin_scene_arrays(name))
    oj.run_pass(np.full(12 * 10, 4, np.int32), statmc=True)
    rec = oj.get("records")
    assert len(rec) == 12 * 10 * 4
    assert np.isfinite(rec[:, 1:4]).all()
    if name != "debug":                                    # debug has no lights and a black environment
        assert rec[:, 1:4].max() > 0


def test_unknown_scene_message():
    with pytest.raises(eng.EngineError) as e:
        eng.builtin_scene_arrays("teapot")
    assert "Unknown scene 'teapot'" in str(e.value) and "bunny_dof" in str(e.value)


def _bbox(tris, lo, hi):
    v = np.array([[getattr(tris[k], f)[:] for f in ("v0", "v1", "v2")] for k in range(lo, hi)]).reshape(-1, 3)
    return v.min(0), v.max(0)


@pytest.mark.skipif(not HAVE_BUNNY, reason="needs the reference's resources/models/bunny.obj")
def test_bunny_scenes_place_the_mesh(monkeypatch):
    monkeypatch.chdir(REF_ROOT)
    tris, sph, mats, lights, env, n = eng.builtin_scene_arrays("bunny")
    assert tuple(n) == (12 + 4968, 0, 7, 1)
    lo, hi = _bbox(tris, 12, n[0])
    assert hi[1] - lo[1] == pytest.approx(0.32, abs=2e-6) and lo[1] == pytest.approx(1e-3, abs=1e-6)
    assert (lo[0] + hi[0]) / 2 == pytest.approx(0.28, abs=1e-6) and (lo[2] + hi[2]) / 2 == pytest.approx(0.25, abs=1e-6)
    assert all(tris[k].material_id == 6 for k in range(12, n[0]))       # the Gold microfacet
    assert mats[6].kind == eng.RPD_MAT_MICROFACET
    tris, sph, mats, lights, env, n = eng.builtin_scene_arrays("bunny_dof")
    assert tuple(n) == (12 + 5 * 4968, 0, 14, 1)
    heights = [0.14, 0.12, 0.15, 0.13, 0.11]
    for i, h in enumerate(heights):
        lo, hi = _bbox(tris, 12 + i * 4968, 12 + (i + 1) * 4968)
        assert hi[1] - lo[1] == pytest.approx(1.2 * h, abs=2e-6) and lo[1] == pytest.approx(1e-3, abs=1e-6)


@pytest.mark.skipif(not HAVE_BUNNY, reason="needs the reference's resources/models/bunny.obj")
def test_bunny_obj_counts():
    tris, n = eng.obj_triangles(os.path.join(REF_ROOT, "resources", "models", "bunny.obj"), 3)
    assert n == 4968 and tris[0].material_id == 3


def test_obj_formats_and_errors(tmp_path):
    p = tmp_path / "m.obj"
    p.write_text("# comment\nmtllib x.mtl\no thing\n"
                 "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvn 0 0 1\n"
                 "f 1 2 3\n"                 # plain
                 "f 1/1 3/1 4/1\n"           # with texture index
                 "f -4//1 -3//1 -2//1 -1//1\n"   # negative, normal index, a quad -> two triangles
                 "s off\nusemtl m\n")
    tris, n = eng.obj_triangles(p, 7)
    assert n == 4
    want = [((0, 0, 0), (1, 0, 0), (1, 1, 0)), ((0, 0, 0), (1, 1, 0), (0, 1, 0)),
            ((0, 0, 0), (1, 0, 0), (1, 1, 0)), ((0, 0, 0), (1, 1, 0), (0, 1, 0))]
    for k, (a, b, c) in enumerate(want):
        assert tuple(tris[k].v0) == a and tuple(tris[k].v1) == b and tuple(tris[k].v2) == c
        assert tuple(tris[k].normal) == (0.0, 0.0, 1.0) and tris[k].material_id == 7
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0 0\nv 1 0 0\nf 1 2 5\n")
    with pytest.raises(eng.EngineError, match="bad face index"):
        eng.obj_triangles(bad)
    with pytest.raises(eng.EngineError, match="cannot open"):
        eng.obj_triangles(tmp_path / "missing.obj")
