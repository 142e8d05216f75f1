"""Host side of the boundary: loaders (reference semantics and error behaviour),
the C-ABI symbol set, scene assembly."""
import ctypes as C
import json
import os
import re

import numpy as np
import pytest

import crystal2_amd as c2
from crystal2_amd import _abi
from conftest import GOLDEN, ROOT


def test_abi_exports_every_declared_symbol(c2lib):
    hdr = open(os.path.join(ROOT, "include", "crystal2.h")).read()
    names = set(re.findall(r"\b(c2_[a-z0-9_]+)\s*\(", hdr))
    assert names
    for n in names:
        assert hasattr(c2lib, n), n
    assert names == {s[0] for s in _abi.SIGNATURES}


def test_scene_load_order_and_lights():
    d = os.path.join(GOLDEN, "scenes", "zoo")
    sc = c2.Scene.load(os.path.join(d, "scene.json"), d)
    desc = sc.desc()
    # point light first (object order), then 2 light triangles, then the sphere light
    kinds = [desc.lights[i].type for i in range(desc.num_lights)]
    assert kinds == [1, 0, 0, 0]
    ents = [desc.lights[i].entity for i in range(1, desc.num_lights)]
    assert all(desc.entity_light[e] == i + 1 for i, e in enumerate(ents))
    # transparent curtain: last two entities have no material
    assert desc.entity_material[desc.num_entities - 1] == -1


def _write(tmp_path, text, name="scene.json"):
    p = tmp_path / name
    p.write_bytes(text if isinstance(text, bytes) else text.encode())
    return str(p)


BASE = {"Camera": {"Type": "Normal", "Position": [0, 0, 3], "LookAt": [0, 0, 0], "Up": [0, 1, 0], "FOV": 40},
        "Scene": {"Textures": {}, "Objects": [
            {"Type": "Geometry", "Shape": {"Type": "Sphere", "Position": [0, 0, 0], "Radius": 0.5},
             "Material": {"Type": "Diffuse", "Color": [0.5, 0.5, 0.5]}}]}}


def test_json_bom_and_numbers(tmp_path):
    s = json.dumps(BASE).replace('"Radius": 0.5', '"Radius": 5e-1')
    p = _write(tmp_path, b"\xef\xbb\xbf" + s.encode())
    sc = c2.Scene.load(p)
    d = sc.desc()
    assert d.num_entities == 1 and d.spheres[3] == np.float32(0.5)


@pytest.mark.parametrize("mutate,code", [
    (lambda s: s["Scene"].pop("Textures"), _abi.C2_OK - 2),
    (lambda s: s["Scene"]["Objects"][0].update({"AreaLight": None}), -2),
    (lambda s: s["Scene"]["Objects"][0]["Shape"].update({"Type": "Cone"}), -2),
    (lambda s: s["Scene"]["Objects"][0]["Material"].update({"Type": "Velvet"}), -2),
    (lambda s: (s["Scene"]["Objects"][0]["Shape"].update({"Type": "Disk"}),
                s["Scene"]["Objects"][0]["Shape"].pop("Radius")), -2),
    (lambda s: s["Scene"]["Objects"][0]["Material"].update({"Type": "Plastic"}), -2),  # no Kd / Ks
    (lambda s: s["Scene"]["Objects"][0]["Material"].update({"Type": "OrenNayar", "Kd": "wood"}), -6),
    (lambda s: s["Scene"]["Objects"][0]["Material"].update({"Type": "Micro", "Roughness": [0.1, 0.1],
                                                            "IOR": 1.5, "NDF": "Blinn"}), -2),
    (lambda s: s.update({"Camera": {"Type": "Fisheye"}}), -2),
])
def test_scene_errors(tmp_path, mutate, code):
    s = json.loads(json.dumps(BASE))
    mutate(s)
    p = _write(tmp_path, json.dumps(s))
    with pytest.raises(c2.Crystal2Error) as e:
        c2.Scene.load(p)
    assert e.value.code == code


def test_disk_record(tmp_path):
    """Disk shares the sphere record: centre, radius, local-to-world, world-to-local (its transpose)."""
    s = json.loads(json.dumps(BASE))
    s["Scene"]["Objects"][0]["Shape"] = {"Type": "Disk", "Position": [1, 2, 3], "Radius": 0.25,
                                         "Rotation": [90, 0, 0]}
    sc = c2.Scene.load(_write(tmp_path, json.dumps(s)))   # the scene owns the arrays the desc points to
    d = sc.desc()
    assert d.num_entities == 1 and d.entity_shape_kind[0] == _abi.C2_SHAPE_DISK
    rec = np.ctypeslib.as_array(d.spheres, shape=(22,))
    assert np.array_equal(rec[:4], np.float32([1, 2, 3, 0.25]))
    l2w = rec[4:13].reshape(3, 3)       # column-major: row c = column c of the matrix
    w2l = rec[13:22].reshape(3, 3)
    assert np.array_equal(w2l, l2w.T)
    # rotation of 90 degrees about x takes the local normal (0, 1, 0) to (0, 0, 1)
    assert np.allclose(l2w[1], [0, 0, 1], atol=1e-6)


def test_malformed_json(tmp_path):
    p = _write(tmp_path, '{"Camera": {"Type": "Normal", }')
    with pytest.raises(c2.Crystal2Error) as e:
        c2.Scene.load(p)
    assert e.value.code == -2 and "JSON 1:" in str(e.value)
    with pytest.raises(c2.Crystal2Error) as e:
        c2.Scene.load(str(tmp_path / "missing.json"))
    assert e.value.code == -3


def test_obj_without_normals_rejected(tmp_path):
    (tmp_path / "m.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    s = json.loads(json.dumps(BASE))
    s["Scene"]["Objects"] = [{"Type": "Geometry", "Shape": {"Type": "TriangleMesh", "ObjFile": "m.obj",
                                                            "Translation": [0, 0, 0], "Scale": [1, 1, 1],
                                                            "Rotation": [0, 0, 0]}}]
    p = _write(tmp_path, json.dumps(s))
    with pytest.raises(c2.Crystal2Error) as e:
        c2.Scene.load(p, str(tmp_path))
    assert e.value.code == -2 and "vn" in str(e.value)


def test_mesh_transform_matches_glm_order(tmp_path):
    """Translation/Scale through the glm-order transform: positions exact, normals renormalised."""
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    n = np.array([[0, 0, 1]] * 3, np.float32)
    from crystal2_amd import scenes
    scenes.write_obj(str(tmp_path / "t.obj"), v, n, np.array([[0, 1, 2]]))
    s = json.loads(json.dumps(BASE))
    s["Scene"]["Objects"] = [{"Type": "Geometry", "Shape": {"Type": "TriangleMesh", "ObjFile": "t.obj",
                                                            "Translation": [1, 2, 3], "Scale": [2, 2, 2],
                                                            "Rotation": [0, 0, 0]},
                              "Material": {"Type": "Diffuse", "Color": [1, 1, 1]}}]
    sc = c2.Scene.load(_write(tmp_path, json.dumps(s)), str(tmp_path))
    d = sc.desc()
    tri = np.ctypeslib.as_array(d.triangles, shape=(24,)).reshape(3, 8)
    assert np.array_equal(tri[:, :3], v * 2 + np.array([1, 2, 3], np.float32))
    assert np.allclose(tri[:, 3:6], n)


def test_config_load(tmp_path):
    from crystal2_amd import scenes
    p = str(tmp_path / "c.json")
    scenes.write_config(p, 320, 200, 16, 5)
    cfg = c2.Config.load(p)
    assert (cfg.width, cfg.height, cfg.spp, cfg.max_depth, cfg.sampler) == (320, 200, 16, 5, "Halton")
