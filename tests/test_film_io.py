"""Film output (the reference's result.hdr, and PFM) and the binary scene cache."""
import ctypes
import os

import numpy as np
import pytest

import crystal2_amd as c2
from crystal2_amd import _abi
from conftest import GOLDEN

HDR = os.path.join(GOLDEN, "hdr")


@pytest.mark.parametrize("name", ["zoo", "narrow", "runs"])
def test_hdr_bytes_equal_stbi_write_hdr(tmp_path, name):
    """stbi_write_hdr's output (written through oracle/_ref) byte for byte."""
    film = np.load(os.path.join(HDR, name + ".npy"))
    out = str(tmp_path / "result.hdr")
    c2.write_hdr(out, film)
    assert open(out, "rb").read() == open(os.path.join(HDR, name + ".hdr"), "rb").read()


def _read_pfm(path):
    raw = open(path, "rb").read()
    head, rest = raw.split(b"\n", 1)
    dims, rest = rest.split(b"\n", 1)
    scale, data = rest.split(b"\n", 1)
    w, h = map(int, dims.split())
    assert head == b"PF" and float(scale) < 0
    return np.frombuffer(data, "<f4").reshape(h, w, 3)[::-1]


def test_pfm_round_trip(tmp_path):
    film = np.load(os.path.join(GOLDEN, "film_zoo.npy"))
    out = str(tmp_path / "film.pfm")
    c2.write_pfm(out, film)
    assert np.array_equal(_read_pfm(out), film)


def _desc_arrays(sc):
    d = sc.desc()
    ne, nt, ns = d.num_entities, d.num_triangles, d.num_spheres
    arr = lambda p, n, dt: np.ctypeslib.as_array(p, shape=(n,)).astype(dt) if n else np.zeros(0, dt)
    out = [arr(d.entity_shape_kind, ne, np.int32), arr(d.entity_shape_index, ne, np.int32),
           arr(d.entity_material, ne, np.int32), arr(d.entity_light, ne, np.int32),
           arr(d.triangles, nt * 24, np.float32), arr(d.spheres, ns * 22, np.float32),
           bytes(C_bytes(d.materials, d.num_materials * ctypes.sizeof(_abi.c2_material))),
           bytes(C_bytes(d.lights, d.num_lights * ctypes.sizeof(_abi.c2_light))),
           bytes(d.camera), list(d.skybox)]
    for t in range(d.num_textures):
        tx = d.textures[t]
        out.append((tx.width, tx.height, bytes(np.ctypeslib.as_array(tx.rgb, shape=(tx.width * tx.height * 3,)))))
    return out


def C_bytes(ptr, n):
    return ctypes.cast(ptr, ctypes.POINTER(ctypes.c_uint8 * max(n, 1))).contents[:n]


@pytest.mark.parametrize("name", ["zoo", "textured", "disks"])
def test_scene_cache_round_trip(tmp_path, name):
    d = os.path.join(GOLDEN, "scenes", name)
    sc = c2.Scene.load(os.path.join(d, "scene.json"), d)
    p = str(tmp_path / "scene.c2s")
    sc.save(p)
    back = c2.Scene.load_binary(p)
    a, b = _desc_arrays(sc), _desc_arrays(back)
    assert len(a) == len(b)
    for x, y in zip(a, b):
        if isinstance(x, np.ndarray):
            assert np.array_equal(x, y)
        else:
            assert x == y


def test_scene_cache_errors(tmp_path):
    p = tmp_path / "bad.c2s"
    p.write_bytes(b"NOTASCENE")
    with pytest.raises(c2.Crystal2Error) as e:
        c2.Scene.load_binary(str(p))
    assert e.value.code == -2
    d = os.path.join(GOLDEN, "scenes", "zoo")
    sc = c2.Scene.load(os.path.join(d, "scene.json"), d)
    good = tmp_path / "good.c2s"
    sc.save(str(good))
    p.write_bytes(good.read_bytes()[:-100])  # truncated
    with pytest.raises(c2.Crystal2Error) as e:
        c2.Scene.load_binary(str(p))
    assert e.value.code == -2
    with pytest.raises(c2.Crystal2Error) as e:
        c2.Scene.load_binary(str(tmp_path / "missing.c2s"))
    assert e.value.code == -3
