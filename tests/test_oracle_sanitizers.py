"""The CPU restatement (oracle/c2_oracle.cpp) and the host scene loader built with
ThreadSanitizer, and with AddressSanitizer + UndefinedBehaviorSanitizer, rendering golden
scenes on several threads (tests/oracle_san.cpp). Each instrumented film must also equal the
uninstrumented oracle's film bit for bit: the instrumented builds run the same float
operations (-ffp-contract=off, no fast-math)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from crystal2_amd import Scene
from oracle import pyoracle

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "crystal2_amd", "csrc")
SOURCES = [os.path.join(HERE, "oracle_san.cpp"), os.path.join(ROOT, "oracle", "c2_oracle.cpp")] + [
    os.path.join(CSRC, f) for f in ("c2_host.cpp", "c2_image.cpp", "c2_bvh.cpp", "c2_kdtree.cpp", "c2_film.cpp")]
SANITIZERS = {"tsan": ["-fsanitize=thread"],
              "asan_ubsan": ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"]}
ENV = {"tsan": {"TSAN_OPTIONS": "halt_on_error=1:exitcode=66"},
       "asan_ubsan": {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=1", "UBSAN_OPTIONS": "print_stacktrace=1"}}
# (scene, width, height, spp, depth, accel, integrator): the BVH and the KD-tree, both
# integrators, textures and a skybox, every material of the zoo
CASES = [("cornell", 48, 48, 4, 5, pyoracle.ACCEL_BVH, "path"),
         ("zoo", 40, 40, 4, 6, pyoracle.ACCEL_KDTREE, "path"),
         ("textured", 40, 40, 4, 6, pyoracle.ACCEL_BVH, "direct"),
         ("zoo_extra", 40, 40, 2, 6, pyoracle.ACCEL_BVH, "path")]
THREADS = 4


@pytest.fixture(scope="module", params=sorted(SANITIZERS))
def san_exe(request, tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("san") / request.param)
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off", "-pthread",
                    *SANITIZERS[request.param], "-I", os.path.join(ROOT, "include"), *SOURCES, "-o", exe],
                   check=True)
    return request.param, exe


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_oracle_under_sanitizers(san_exe, case, tmp_path):
    kind, exe = san_exe
    name, w, h, spp, depth, accel, integ = case
    sdir = os.path.join(HERE, "golden", "scenes", name)
    out = str(tmp_path / "film.raw")
    r = subprocess.run([exe, os.path.join(sdir, "scene.json"), sdir, str(w), str(h), str(spp), str(depth), str(accel),
                        str(pyoracle.INTEGRATORS[integ]), str(THREADS), out],
                       capture_output=True, text=True, env=dict(os.environ, **ENV[kind]), timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert "WARNING: ThreadSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    film = np.fromfile(out, np.float32).reshape(h, w, 3)
    scene = Scene.load(os.path.join(sdir, "scene.json"), sdir)
    ref, _ = pyoracle.OracleScene(scene, accel=accel, integrator=integ).render(w, h, spp, depth, threads=THREADS)
    assert np.array_equal(film, ref)
    assert float(film.sum()) > 0
