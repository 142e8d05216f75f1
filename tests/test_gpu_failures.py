"""Failure handling through the C-ABI on the GPU: rays with NaN, zero or infinite
directions, degenerate triangles, an allocation the device cannot hold, and empty
scenes cast from a non-default stream. Every kernel here still exits normally: the
inputs are ones the kernels' own arithmetic turns into misses or exact reference
results (the reference's Triangle::Intersect rejects NaN solutions,
src/Shapes/Triangle.cpp:32-81; its slab test ignores NaN, src/Core/Geometry.h:136-157)."""
import ctypes as C

import numpy as np
import pytest

import crystal2_amd as c2
import pyoracle
from crystal2_amd import scenes

pytestmark = pytest.mark.gpu


def _box_scene(degenerate=False):
    sc = c2.Scene.create()
    m = sc.add_material(0, [0.6, 0.6, 0.6])
    pos, nrm, faces = scenes.displaced_icosphere(2, seed=3)
    sc.add_mesh(pos, nrm, faces, m)
    if degenerate:
        # zero-area triangles: repeated vertices, collinear vertices, a point
        p = np.array([[0.2, 0.1, 0.0], [0.2, 0.1, 0.0], [0.5, 0.4, 0.1],
                      [-0.3, 0.0, 0.0], [0.0, 0.0, 0.0], [0.3, 0.0, 0.0],
                      [0.1, 0.1, 0.1], [0.1, 0.1, 0.1], [0.1, 0.1, 0.1]], np.float32)
        n = np.tile(np.array([[0, 1, 0]], np.float32), (9, 1))
        sc.add_mesh(p, n, np.arange(9).reshape(3, 3), m)
    sc.add_sphere([0.0, 0.0, 0.0], 0.5, m)
    return sc


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 8), np.float32)
    r[:, :3] = rng.uniform(-2, 2, (n, 3))
    d = rng.normal(size=(n, 3))
    r[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r[:, 7] = np.inf
    return r


def test_nan_zero_and_infinite_directions(gpu_available):
    print("case failure: nan/zero/inf directions", flush=True)
    sc = _box_scene()
    r = _rays(4096, 1)
    r[0::5, 3] = np.nan
    r[1::5, 3:6] = 0
    r[2::5, 4] = np.inf
    r[3::5, 0] = np.nan
    bvh = c2.GpuBVH(sc)
    ids, t = bvh.intersect(r)
    o = pyoracle.OracleScene(sc, tie=pyoracle.TIE_LOWEST)
    oi, ot = o.intersect(r)
    assert np.array_equal(ids, oi)
    assert np.array_equal(t[oi >= 0], ot[oi >= 0])
    assert (ids[0::5] == -1).all() and (ids[3::5] == -1).all()
    sr = r.copy()
    sr[:, 7] = 1.0
    assert np.array_equal(bvh.intersect_test(sr), o.intersect_test(sr))


def test_degenerate_triangles(gpu_available):
    print("case failure: degenerate triangles", flush=True)
    sc = _box_scene(degenerate=True)
    r = _rays(4096, 2)
    bvh = c2.GpuBVH(sc)
    ids, t = bvh.intersect(r)
    o = pyoracle.OracleScene(sc, tie=pyoracle.TIE_LOWEST)
    oi, ot = o.intersect(r)
    assert np.array_equal(ids, oi)
    assert np.array_equal(t[oi >= 0], ot[oi >= 0])
    sr = r.copy()
    sr[:, 7] = 1.0
    assert np.array_equal(bvh.intersect_test(sr), o.intersect_test(sr))
    sc.set_camera([0, 0, 3], [0, 0, 0], [0, 1, 0], 40)
    lm = sc.add_material(0, [0.5, 0.5, 0.5])
    lp = np.array([[-1, 2, -1], [1, 2, -1], [1, 2, 1], [-1, 2, 1]], np.float32)
    ln = np.tile(np.array([[0, -1, 0]], np.float32), (4, 1))
    sc.add_mesh(lp, ln, np.array([[0, 1, 2], [0, 2, 3]]), lm, lemit=[8, 8, 8])
    pt = c2.WavefrontPathIntegrator(sc, 48, 48, 2, 5)
    pt.render()
    pt.sync()
    g = pt.film()
    pt.close()
    f, _ = pyoracle.OracleScene(sc, tie=pyoracle.TIE_LOWEST).render(48, 48, 2, 5)
    assert np.array_equal(g, f)


def test_allocation_failure_is_nomem_and_recoverable(gpu_available):
    """2^42 rays need 128 TiB of device memory: C2_ERR_NOMEM before any input is read;
    the next call on the same handle works (no error is left behind)."""
    print("case failure: allocation", flush=True)
    lib = c2.load_library()
    sc = _box_scene()
    bvh = c2.GpuBVH(sc)
    r = _rays(16, 3)
    prim = np.empty(16, np.int32)
    t = np.empty(16, np.float32)
    rc = lib.c2_accel_intersect(bvh._h, r.ctypes.data_as(C.POINTER(C.c_float)), 1 << 42,
                                prim.ctypes.data_as(C.POINTER(C.c_int32)), t.ctypes.data_as(C.POINTER(C.c_float)))
    assert rc == -5  # C2_ERR_NOMEM
    assert b"hipMalloc" in lib.c2_last_error()
    ids, _ = bvh.intersect(r)
    oi, _ = pyoracle.OracleScene(sc, tie=pyoracle.TIE_LOWEST).intersect(r)
    assert np.array_equal(ids, oi)
    pt = c2.WavefrontPathIntegrator(sc, 16, 16, 1, 2)  # a renderer after the failure
    pt.render()
    pt.sync()
    pt.close()


def test_empty_scene_device_entry_points_on_a_stream(gpu_available):
    """An empty scene: prim -1 and t +inf, no occluder, written by kernels on the caller's
    (non-default) stream; buffers start as garbage so every element must be written."""
    import torch
    print("case failure: empty scene device entry points", flush=True)
    sc = c2.Scene.create()
    bvh = c2.GpuBVH(sc)
    n = 5000
    rays = torch.from_numpy(_rays(n, 4)).cuda()
    prim = torch.full((n,), 7, dtype=torch.int32, device="cuda")
    t = torch.full((n,), 3.0, dtype=torch.float32, device="cuda")
    hit = torch.full((n,), 9, dtype=torch.uint8, device="cuda")
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    bvh.intersect_device(rays.data_ptr(), n, prim.data_ptr(), t.data_ptr(), s.cuda_stream)
    bvh.intersect_test_device(rays.data_ptr(), 0, n, hit.data_ptr(), s.cuda_stream)
    s.synchronize()
    assert (prim.cpu() == -1).all()
    assert torch.isinf(t.cpu()).all() and (t.cpu() > 0).all()
    assert (hit.cpu() == 0).all()
