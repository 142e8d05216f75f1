"""Batched scenes: every film of a batch render (direct launches and a replayed
hipGraph, the wavefront schedule and the persistent kernel) is bit-identical to the
same scene rendered alone with the same options. The items differ in size, spp, light
count (2, 8, a point light) and materials, so the views share launches with different
trip counts."""
import os

import numpy as np
import pytest

import crystal2_amd as c2
from crystal2_amd import scenes
from conftest import GOLDEN

pytestmark = pytest.mark.gpu
DEPTH = 6


def _load(name):
    d = os.path.join(GOLDEN, "scenes", name)
    return c2.Scene.load(os.path.join(d, "scene.json"), d)


def _items():
    return [
        (_load("cornell"), 64, 64, 4),
        (_load("zoo"), 64, 64, 8),
        (scenes.build_spheres_scene(12, 2, seed=5), 80, 45, 6),
        (_load("zoo_nophong"), 48, 40, 3),
        (scenes.build_spheres_scene(30, 3, seed=9), 64, 64, 2),
    ]


def _alone(sc, w, h, spp, film="f32"):
    pt = c2.WavefrontPathIntegrator(sc, w, h, spp, DEPTH, arg_order="rtl", film=film)
    pt.render()
    pt.sync()
    f = pt.film()
    pt.close()
    return f


def test_batch_films_equal_single_renders(gpu_available):
    items = _items()
    ref = [_alone(*it) for it in items]
    b = c2.BatchRenderer(items, DEPTH, arg_order="rtl")
    for use_graph in (False, True, True):  # the second graph call replays the captured graph
        print(f"case batch use_graph={use_graph}", flush=True)
        b.render(use_graph)
        b.sync()
        for i, r in enumerate(ref):
            assert np.array_equal(b.film(i), r), f"item {i} graph={use_graph}"
    st = b.stats()
    assert st["samples"] == 3 * sum(w * h * s for _, w, h, s in items)
    assert st["graph_nodes"] > 0
    b.close()


def test_batch_autotune_keeps_results(gpu_available):
    print("case batch autotune", flush=True)
    items = _items()[:3]
    ref = [_alone(*it, film="f64") for it in items]
    b = c2.BatchRenderer(items, DEPTH, arg_order="rtl", film="f64")
    t = b.autotune()
    assert t["candidates"] >= 3 and t["pass_ms"] > 0
    assert t["bvh_width"] in (4, 8) and t["pass_ms_width8"] > 0
    b.render(True)
    b.sync()
    for i, r in enumerate(ref):
        assert np.array_equal(b.film(i), r)
    b.close()


@pytest.mark.parametrize("width", [8, 4])
def test_batch_bvh_width(gpu_available, width):
    """Both BVH widths of a batch (the wavefront traverses the 8-wide or the 4-wide tree)
    render the single-scene films, in direct launches and the captured graph."""
    print(f"case batch bvh width {width}", flush=True)
    items = _items()
    ref = [_alone(*it) for it in items]
    b = c2.BatchRenderer(items, DEPTH, arg_order="rtl")
    b.configure("wavefront", bvh_width=width)
    for use_graph in (False, True):
        b.render(use_graph)
        b.sync()
        for i, r in enumerate(ref):
            assert np.array_equal(b.film(i), r), f"item {i} width={width} graph={use_graph}"
    b.close()


def test_batch_small_chunks(gpu_available):
    """Chunks smaller than a pass split views across launches."""
    print("case batch chunks", flush=True)
    items = _items()[:3]
    ref = [_alone(*it) for it in items]
    b = c2.BatchRenderer(items, DEPTH, arg_order="rtl", paths_in_flight=1500, queue_blocks=64)
    b.render(False)
    b.sync()
    for i, r in enumerate(ref):
        assert np.array_equal(b.film(i), r)
    b.close()


def test_batch_direct_lighting(gpu_available):
    print("case batch direct lighting", flush=True)
    items = _items()[:3]
    ref = []
    for sc, w, h, spp in items:
        pt = c2.WavefrontPathIntegrator(sc, w, h, spp, DEPTH, arg_order="rtl", integrator="direct")
        pt.render()
        pt.sync()
        ref.append(pt.film())
        pt.close()
    b = c2.BatchRenderer(items, DEPTH, arg_order="rtl", integrator="direct")
    b.render(True)
    b.sync()
    for i, r in enumerate(ref):
        assert np.array_equal(b.film(i), r)
    b.close()


@pytest.mark.parametrize("block,waves", [(128, 0), (256, 0), (512, 0), (256, 4), (128, 6), (512, 8)])
def test_persistent_films_equal_single_renders(gpu_available, block, waves):
    """The persistent kernel (one launch; lanes take pixels from a device counter and
    trace each path to its end) gives every film of the wavefront bit for bit, and traces
    the same extension, shadow and MIS rays."""
    items = _items()
    ref = [_alone(*it) for it in items]
    w = c2.BatchRenderer(items, DEPTH, arg_order="rtl")
    w.render(False)
    w.sync()
    ws = w.stats()
    w.close()
    b = c2.BatchRenderer(items, DEPTH, arg_order="rtl")
    b.configure("persistent", block_threads=block, waves_per_simd=waves)
    for rep, use_graph in enumerate((False, True, True)):
        print(f"case persistent block={block} waves={waves} use_graph={use_graph}", flush=True)
        b.render(use_graph)
        b.sync()
        for i, r in enumerate(ref):
            assert np.array_equal(b.film(i), r), f"item {i} block={block} graph={use_graph}"
        if rep == 0:
            st = b.stats()
            assert st["schedule"] == "persistent" and st["block_threads"] == block
            assert st["waves_per_simd"] == waves
            for k in ("extension_rays", "shadow_rays", "mis_rays"):
                assert st[k] == ws[k], k
    st = b.stats()
    assert 0 < st["graph_nodes"] <= 4  # film reset, counter reset, the persistent launch
    b.close()


@pytest.mark.parametrize("film", ["f64", "kahan"])
def test_persistent_film_modes_and_direct_lighting(gpu_available, film):
    items = _items()[:3]
    for integrator in ("path", "direct"):
        print(f"case persistent film={film} integrator={integrator}", flush=True)
        ref = []
        for sc, w, h, spp in items:
            pt = c2.WavefrontPathIntegrator(sc, w, h, spp, DEPTH, arg_order="rtl", integrator=integrator, film=film)
            pt.render()
            pt.sync()
            ref.append(pt.film())
            pt.close()
        b = c2.BatchRenderer(items, DEPTH, arg_order="rtl", integrator=integrator, film=film)
        b.configure("persistent", block_threads=256)
        b.render(True)
        b.sync()
        for i, r in enumerate(ref):
            assert np.array_equal(b.film(i), r), f"item {i}"
        b.close()


def test_persistent_refuses_the_kdtree(gpu_available):
    print("case persistent kdtree refused", flush=True)
    b = c2.BatchRenderer(_items()[:2], DEPTH, arg_order="rtl", accel="kdtree")
    with pytest.raises(c2.Crystal2Error) as e:
        b.configure("persistent", block_threads=256)
    assert e.value.code == -6  # C2_ERR_UNSUPPORTED
    b.configure("wavefront", block_threads=100)  # the wavefront ignores the block size
    b.close()
    b = c2.BatchRenderer(_items()[:2], DEPTH, arg_order="rtl")
    for bs, wv in ((100, 0), (256, 3)):
        with pytest.raises(c2.Crystal2Error) as e:
            b.configure("persistent", block_threads=bs, waves_per_simd=wv)
        assert e.value.code == -1  # C2_ERR_INVALID
    b.close()


def test_batch_64_scene_workload(gpu_available):
    """The batched-scenes workload at its configured size: 64 sphere-field scenes with
    triangle counts log-uniform in 1k-200k, resolutions in {128, 256, 512}^2 and spp in
    {4, 16, 64} (198 M samples), depth 8, autotuned (schedule, launch shape, BVH width),
    captured into a hipGraph and replayed twice. Every film is bit-identical to the scene
    rendered alone."""
    print("case batch 64-scene workload", flush=True)
    specs = scenes.batch_specs(64, 6400)
    items = scenes.build_batch(specs)
    ref = []
    for sc, w, h, spp in items:
        pt = c2.WavefrontPathIntegrator(sc, w, h, spp, 8, arg_order="rtl")
        pt.render()
        pt.sync()
        ref.append(pt.film())
        pt.close()
    b = c2.BatchRenderer(items, 8, arg_order="rtl")
    t = b.autotune()
    print(f"autotune: {t}", flush=True)
    # 9 wavefront launch shapes, the other BVH width, 12 persistent launch shapes
    assert t["candidates"] == 22 and t["render_ms_wavefront"] > 0 and t["render_ms_persistent"] > 0
    assert t["bvh_width"] in (4, 8) and t["pass_ms_width8"] > 0
    for rep in range(2):
        b.render(True)
        b.sync()
        bad = [i for i, r in enumerate(ref) if not np.array_equal(b.film(i), r)]
        assert not bad, f"replay {rep}: films {bad} differ"
    # the schedule the autotuner did not keep gives the same films
    other = "wavefront" if t["schedule"] == "persistent" else "persistent"
    print(f"case batch 64-scene workload, {other}", flush=True)
    b.configure(other, block_threads=t["block_threads"], waves_per_simd=t["waves_per_simd"])
    b.render(True)
    b.sync()
    bad = [i for i, r in enumerate(ref) if not np.array_equal(b.film(i), r)]
    assert not bad, f"{other}: films {bad} differ"
    b.close()
