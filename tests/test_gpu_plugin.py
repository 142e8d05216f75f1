"""The reference renderer with the GPU plugins registered: oracle/_ref/crystal2_ref_gpu is the
reference's own sources, with one branch added to Accelerator::GetAccelerator ("GPUBVH") and
to Integrator::LoadIntegrator ("WavefrontPathTracer"), linked with libcrystal2.so
(crystal2_amd/plugin/crystal2_gpu.cpp). A config that names the plugins renders through
Scene::CreateScene and Integrator::LoadIntegrator (src/Core/RayTracer.cpp:11-22) what the
reference renders; the reference's own films and hits are the fixtures."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

BIN = os.path.join(ROOT, "oracle", "_ref", "crystal2_ref_gpu")
SCN = os.path.join(GOLDEN, "scenes")

needs_bin = pytest.mark.skipif(not os.path.exists(BIN), reason="plugin build absent (make -C oracle refgpu)")


def _config(tmp_path, scene, integrator=None, accel=None):
    with open(os.path.join(SCN, scene, "config.json")) as f:
        cfg = json.load(f)
    if integrator:
        cfg["Integrator"]["Type"] = integrator
    if accel:
        cfg["Acceleration"] = accel
    p = os.path.join(str(tmp_path), "config.json")
    with open(p, "w") as f:
        json.dump(cfg, f)
    return p, cfg


def _run(args, scene, timeout=300):
    # OBJ paths in the scene files are relative to the working directory, as the reference reads them
    return subprocess.run([BIN, *args], cwd=os.path.join(SCN, scene), capture_output=True, text=True,
                          timeout=timeout)


def _render(tmp_path, scene, integrator=None, accel=None):
    cfgp, cfg = _config(tmp_path, scene, integrator, accel)
    out = os.path.join(str(tmp_path), "film.f32")
    r = _run(["render", cfgp, os.path.join(SCN, scene, "scene.json"), out], scene)
    assert r.returncode == 0, r.stdout + r.stderr
    return np.fromfile(out, np.float32).reshape(cfg["Height"], cfg["Width"], 3), r


@needs_bin
def test_reference_path_unchanged_by_the_registration():
    """The patched factories still build the reference's own BVH and PathTracer."""
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        film, _ = _render(d, "cornell")
    assert np.array_equal(film, np.load(os.path.join(GOLDEN, "film_cornell.npy")))


@needs_bin
def test_plugins_fail_loudly_without_a_gpu(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    cfgp, _ = _config(tmp_path, "cornell", "WavefrontPathTracer", "GPUBVH")
    r = _run(["render", cfgp, os.path.join(SCN, "cornell", "scene.json"), os.path.join(str(tmp_path), "f.f32")],
             "cornell")
    assert r.returncode == 4 and "crystal2 GPU:" in r.stderr, r.stderr


@needs_bin
@pytest.mark.gpu
@pytest.mark.parametrize("accel,golden", [("GPUBVH", "film_cornell.npy"), ("BVH", "film_cornell.npy"),
                                          ("KDTree", "film_cornell_kd.npy")])
def test_wavefront_path_tracer_plugin_renders_the_reference_film(gpu_available, tmp_path, accel, golden):
    """"WavefrontPathTracer" through Integrator::LoadIntegrator on the Cornell box: the
    library's own film of that scene bit for bit (the clang build's argument order; with
    "KDTree" the device KD-tree), and the reference PathTracer's film within the relMSE of
    1e-4 that the library's films keep from it (the device evaluates transcendentals in
    double and rounds them, libm moves some pixels in the last bit)."""
    import crystal2_amd as c2
    from conftest import relmse
    print(f"case plugin WavefrontPathTracer accel={accel}", flush=True)
    film, r = _render(tmp_path, "cornell", "WavefrontPathTracer", accel)
    assert "Loaded Integrator: WavefrontPathTracer" in r.stdout
    d = os.path.join(SCN, "cornell")
    pt = c2.WavefrontPathIntegrator(c2.Scene.load(os.path.join(d, "scene.json"), d), 64, 64, 4, 5, arg_order="ltr",
                                    accel="kdtree" if accel == "KDTree" else "bvh")
    pt.render()
    pt.sync()
    own = pt.film()
    pt.close()
    assert np.array_equal(film, own)
    assert relmse(film, np.load(os.path.join(GOLDEN, golden))) <= 1e-4


@needs_bin
@pytest.mark.gpu
def test_reference_path_tracer_over_the_gpubvh_plugin(gpu_available, tmp_path):
    """The reference's own CPU PathTracer with "GPUBVH" as the scene's accelerator: every
    closest hit and every shadow ray (with the light's shape ignored) goes through the
    plugin, and the film equals the reference's."""
    print("case plugin GPUBVH under the reference PathTracer", flush=True)
    film, r = _render(tmp_path, "cornell", "PathTracer", "GPUBVH")
    assert "Created Acceleration Structure: GPUBVH" in r.stdout, r.stdout
    assert np.array_equal(film, np.load(os.path.join(GOLDEN, "film_cornell.npy")))


@needs_bin
@pytest.mark.gpu
def test_gpubvh_plugin_hits_vs_reference_golden(gpu_available, tmp_path):
    """Scene::Intersect / IntersectTest through "GPUBVH" on the zoo's 4,096 golden rays: the
    entities the reference's BVH returns (>= 99.99%, as the library's own hit test allows for
    ties), their distances bit for bit (the entity's own Intersect fills them), and the
    any-hit answers."""
    print("case plugin GPUBVH hits zoo", flush=True)
    g = np.load(os.path.join(GOLDEN, "hits_zoo.npz"))
    rays = g["rays"].astype(np.float32)
    n = rays.shape[0]
    rp = os.path.join(str(tmp_path), "rays.bin")
    with open(rp, "wb") as f:
        f.write(np.int64(n).tobytes())
        f.write(rays.tobytes())
    cfgp, _ = _config(tmp_path, "zoo", None, "GPUBVH")
    out = os.path.join(str(tmp_path), "hits.bin")
    r = _run(["hits", cfgp, os.path.join(SCN, "zoo", "scene.json"), rp, out], "zoo")
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(out, np.uint8)
    ids = raw[:4 * n].view(np.int32)
    t = raw[4 * n:8 * n].view(np.float32)
    anyh = raw[4 * n + 4 * n + 24 * n:]
    assert anyh.size == n
    agree = ids == g["ids"]
    assert agree.mean() >= 0.9999
    hit = agree & (ids >= 0)
    assert np.array_equal(t[hit], g["t"][hit])
    assert np.array_equal(anyh, g["any"])
