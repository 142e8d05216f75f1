"""c2_accel_intersect_noalloc: the ray cast staged in the library's fixed device array
(no device allocation) returns exactly what c2_accel_intersect returns, on two calls in a
row (the second with fewer rays, so stale contents from the first must not leak), for the
BVH and the KD-tree, and rejects more rays than the array holds."""
import ctypes as C

import numpy as np
import pytest

import crystal2_amd as c2
from crystal2_amd import scenes

pytestmark = pytest.mark.gpu

MAX_RAYS = 1 << 20  # C2_NOALLOC_MAX_RAYS


def _call(lib, h, rays):
    n = len(rays)
    prim = np.full(n, 12345, np.int32)
    t = np.full(n, -7.0, np.float32)
    rc = lib.c2_accel_intersect_noalloc(h, rays.ctypes.data_as(C.POINTER(C.c_float)), n,
                                        prim.ctypes.data_as(C.POINTER(C.c_int32)),
                                        t.ctypes.data_as(C.POINTER(C.c_float)))
    return rc, prim, t


@pytest.mark.parametrize("accel", ["bvh", "kdtree"])
def test_noalloc_intersect_equals_intersect(gpu_available, accel):
    print(f"case noalloc intersect {accel}", flush=True)
    lib = c2.load_library()
    pos, nrm, faces = scenes.displaced_icosphere(5 if accel == "bvh" else 3, seed=7)
    sc = c2.Scene.create()
    m = sc.add_material(0, [0.5, 0.5, 0.5])
    sc.add_mesh(pos, nrm, faces, m)
    acc = c2.GpuBVH(sc, accel=accel)
    for n, seed in ((MAX_RAYS if accel == "bvh" else 200000, 11), (3001, 12)):
        rays = np.ascontiguousarray(scenes.icosphere_rays(n, pos, seed=seed), np.float32)
        rc, prim, t = _call(lib, acc._h, rays)
        assert rc == 0, lib.c2_last_error()
        ids, tt = acc.intersect(rays)
        assert np.array_equal(prim, ids)
        assert np.array_equal(t, tt)
        assert 0.05 < (prim >= 0).mean() < 0.95


def test_noalloc_limits_and_empty_scene(gpu_available):
    print("case noalloc limits", flush=True)
    lib = c2.load_library()
    sc = c2.Scene.create()
    acc = c2.GpuBVH(sc)
    rays = np.zeros((4, 8), np.float32)
    rays[:, 3] = 1
    rays[:, 7] = np.inf
    rc, prim, t = _call(lib, acc._h, rays)
    assert rc == 0 and (prim == -1).all() and np.isinf(t).all()
    one = np.zeros(8, np.float32)
    p = np.zeros(1, np.int32)
    tt = np.zeros(1, np.float32)
    rc = lib.c2_accel_intersect_noalloc(acc._h, one.ctypes.data_as(C.POINTER(C.c_float)), MAX_RAYS + 1,
                                        p.ctypes.data_as(C.POINTER(C.c_int32)), tt.ctypes.data_as(C.POINTER(C.c_float)))
    assert rc == -1 and b"staging maximum" in lib.c2_last_error()
