"""Writes the golden fixtures under tests/golden/ from the reference renderer.

Runs oracle/_ref/crystal2_ref (the reference sources from /root/reference,
compiled by oracle/Makefile) in THIS container. The fixtures are data only:
scene files (JSON + OBJ written by crystal2_amd.scenes) and the reference's
outputs on them. The reference binary is built with clang, which evaluates
SampleDirection(Get1D(), Get2D()) left to right, so every film here is an
arg_order="ltr" film.

    python tests/golden/make_golden.py          # every fixture
    python tests/golden/make_golden.py direct   # DirectLighting films only
    python tests/golden/make_golden.py extra    # Oren-Nayar / Plastic / Beckmann scene only
    python tests/golden/make_golden.py disk     # Disk-shape scene only
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from crystal2_amd import scenes  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "crystal2_ref")
SCN = os.path.join(HERE, "scenes")


def ref(*args, cwd):
    out = subprocess.run([REF, *args], cwd=cwd, check=True, capture_output=True, text=True)
    return out.stdout


def film(name, w, h, config="config.json", tag=""):
    d = os.path.join(SCN, name)
    ref("render", config, "scene.json", "film.f32", cwd=d)
    f = np.fromfile(os.path.join(d, "film.f32"), np.float32).reshape(h, w, 3)
    os.remove(os.path.join(d, "film.f32"))
    np.save(os.path.join(HERE, f"film_{name}{tag}.npy"), f)


def direct_lighting():
    """DirectLightingIntegrator films of the existing golden scenes (config_dl.json
    beside each scene.json); leaves every other fixture untouched."""
    if not os.path.exists(REF):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    for name, spp in (("cornell", 4), ("zoo", 8), ("zoo_nophong", 8)):
        scenes.write_config(os.path.join(SCN, name, "config_dl.json"), 64, 64, spp, 6, integrator="DirectLight")
        film(name, 64, 64, config="config_dl.json", tag="_dl")
    print("DirectLighting films written to", HERE)


def hits(name, rays):
    d = os.path.join(SCN, name)
    rp = os.path.join(d, "rays.bin")
    with open(rp, "wb") as fh:
        np.array([len(rays)], np.int64).tofile(fh)
        rays.astype(np.float32).tofile(fh)
    ref("hits", "config.json", "scene.json", "rays.bin", "hits.bin", cwd=d)
    n = len(rays)
    raw = open(os.path.join(d, "hits.bin"), "rb").read()
    off = 0
    ids = np.frombuffer(raw, np.int32, n, off); off += 4 * n
    t = np.frombuffer(raw, np.float32, n, off); off += 4 * n
    off += 4 * n * 6  # hit position and normal (not kept)
    anyh = np.frombuffer(raw, np.uint8, n, off)
    os.remove(rp)
    os.remove(os.path.join(d, "hits.bin"))
    np.savez_compressed(os.path.join(HERE, f"hits_{name}.npz"), rays=rays.astype(np.float32), ids=ids, t=t,
                        any=anyh)


def main():
    if not os.path.exists(REF):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    scenes.cornell(os.path.join(SCN, "cornell"), width=64, height=64, spp=4, depth=5)
    scenes.material_zoo(os.path.join(SCN, "zoo"), width=64, height=64, spp=8, depth=6)
    scenes.material_zoo(os.path.join(SCN, "zoo_nophong"), width=64, height=64, spp=8, depth=6, phong=False)
    # icosphere mesh (level 3) for closest / any hit
    pos, nrm, faces = scenes.displaced_icosphere(3, seed=7)
    d = os.path.join(SCN, "icosphere")
    os.makedirs(d, exist_ok=True)
    scenes.write_obj(os.path.join(d, "mesh.obj"), pos, nrm, faces)
    import json
    json.dump({"Camera": {"Type": "Normal", "Position": [0, 0, 4], "LookAt": [0, 0, 0], "Up": [0, 1, 0], "FOV": 40},
               "Scene": {"Textures": {}, "Objects": [{"Type": "Geometry", "Shape": {
                   "Type": "TriangleMesh", "ObjFile": "mesh.obj", "Translation": [0, 0, 0], "Scale": [1, 1, 1],
                   "Rotation": [0, 0, 0]}, "Material": {"Type": "Diffuse", "Color": [0.5, 0.5, 0.5]}}]}},
              open(os.path.join(d, "scene.json"), "w"), indent=1)
    scenes.write_config(os.path.join(d, "config.json"), 64, 64, 1, 1)
    for name in ("cornell", "zoo", "zoo_nophong"):
        w = h = 64
        film(name, w, h)
    rays = scenes.icosphere_rays(4096, pos, seed=11)
    rays[:, 7] = 1.5  # finite tmax for the any-hit queries
    hits("icosphere", rays)
    # rays into the zoo scene from inside the box (triangles + analytic spheres)
    rng = np.random.default_rng(5)
    zr = np.zeros((4096, 8), np.float32)
    zr[:, 0:3] = rng.uniform(-0.9, 0.9, (4096, 3))
    v = rng.normal(size=(4096, 3))
    zr[:, 3:6] = v / np.linalg.norm(v, axis=1, keepdims=True)
    zr[:, 6] = 0
    zr[:, 7] = rng.uniform(0.1, 2.0, 4096)
    hits("zoo", zr)
    # Halton values: (x, y, k) queries on a 256x256 and a 1920x1080 film, dims 0..99
    for (W, H) in ((256, 256), (1920, 1080), (64, 64)):
        q = np.stack([rng.integers(0, W, 512), rng.integers(0, H, 512), rng.integers(0, 64, 512)], 1).astype(np.int32)
        qp = os.path.join(HERE, "q.bin")
        with open(qp, "wb") as fh:
            np.array([len(q)], np.int64).tofile(fh)
            q.tofile(fh)
        op = os.path.join(HERE, "h.bin")
        subprocess.run([REF, "halton", str(W), str(H), "64", qp, "100", op], check=True)
        raw = open(op, "rb").read()
        idx = np.frombuffer(raw, np.int32, len(q))
        vals = np.frombuffer(raw, np.float32, len(q) * 100, 4 * len(q)).reshape(len(q), 100)
        np.savez_compressed(os.path.join(HERE, f"halton_{W}x{H}.npz"), queries=q, index=idx, values=vals)
        os.remove(qp)
        os.remove(op)
    print("golden fixtures written to", HERE)


def extra_materials():
    """The zoo with Oren-Nayar, Plastic and Beckmann spheres: the reference's films for
    both integrators; leaves every other fixture untouched."""
    if not os.path.exists(REF):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    d = os.path.join(SCN, "zoo_extra")
    scenes.material_zoo(d, width=64, height=64, spp=8, depth=6, extra=True)
    scenes.write_config(os.path.join(d, "config_dl.json"), 64, 64, 8, 6, integrator="DirectLight")
    film("zoo_extra", 64, 64)
    film("zoo_extra", 64, 64, config="config_dl.json", tag="_dl")
    print("extra-material films written to", HERE)


def disks():
    """The Disk-shape scene: the reference's films for both integrators and its closest /
    any hits for rays from inside the box; leaves every other fixture untouched."""
    if not os.path.exists(REF):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    d = os.path.join(SCN, "disks")
    scenes.disk_scene(d, width=64, height=64, spp=8, depth=6)
    scenes.write_config(os.path.join(d, "config_dl.json"), 64, 64, 8, 6, integrator="DirectLight")
    film("disks", 64, 64)
    film("disks", 64, 64, config="config_dl.json", tag="_dl")
    rng = np.random.default_rng(17)
    zr = np.zeros((4096, 8), np.float32)
    zr[:, 0:3] = rng.uniform(-0.9, 0.9, (4096, 3))
    v = rng.normal(size=(4096, 3))
    zr[:, 3:6] = v / np.linalg.norm(v, axis=1, keepdims=True)
    zr[:, 6] = 0
    zr[:, 7] = rng.uniform(0.1, 2.0, 4096)
    hits("disks", zr)
    print("disk films and hits written to", HERE)


if __name__ == "__main__":
    if sys.argv[1:] == ["direct"]:
        direct_lighting()
    elif sys.argv[1:] == ["extra"]:
        extra_materials()
    elif sys.argv[1:] == ["disk"]:
        disks()
    else:
        main()
