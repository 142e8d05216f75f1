"""Synthetic scenes in Crystal2's on-disk formats (scene/config JSON + OBJ).

The reference ships no loadable scene (its OBJ models are git-ignored and its
`"AreaLight": null` scenes fail to load), so every workload is generated here:

* ``cornell``   -- the 256x256, 16 spp, depth-5 Cornell box (10 wall triangles,
  a 2-triangle area-light quad, one diffuse and one mirror sphere).
* ``icosphere`` -- the ray-cast microbenchmark mesh: an icosphere subdivided
  ``level`` times (20 * 4**level triangles), vertex radius
  ``1 + 0.1 * value_noise(seed)``, and its seeded ray set.
* ``spheres_scene`` -- ground quad + N sphere meshes + 4 area-light quads with
  randomly drawn materials (diffuse, GGX, anisotropic GGX, glass, rough glass).

Files follow the reference schema: config keys per
``src/Loaders/ConfigLoader.cpp:12-16`` and ``src/Integrators/Integrator.cpp:19-35``;
scene keys per ``src/Core/Scene.cpp:23-199`` and ``src/Materials/Material.cpp:22-126``.
OBJ files are written with per-vertex normals (``vn``), which the reference's
loader requires (``src/Loaders/ObjLoader.cpp:122-126``); float coordinates are
written with Python's round-trip ``repr`` so ``%lf`` parsing followed by the
double->float conversion restores every float32 exactly.
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

__all__ = [
    "write_obj", "write_config", "cornell", "icosphere_mesh", "displaced_icosphere",
    "icosphere_rays", "spheres_scene", "SceneFiles", "disk_scene",
]


def _f(x) -> str:
    return repr(float(np.float32(x)))


def write_obj(path: str, positions: np.ndarray, normals: np.ndarray, faces: np.ndarray, uvs=None) -> None:
    """Write an indexed triangle OBJ (``v``, ``vn``, ``f a//a b//b c//c``, 1-based); with
    ``uvs`` also ``vt`` lines and ``f a/a/a ...`` faces."""
    positions = np.asarray(positions, dtype=np.float32)
    normals = np.asarray(normals, dtype=np.float32)
    faces = np.asarray(faces, dtype=np.int64)
    lines = []
    for p in positions:
        lines.append(f"v {_f(p[0])} {_f(p[1])} {_f(p[2])}")
    for n in normals:
        lines.append(f"vn {_f(n[0])} {_f(n[1])} {_f(n[2])}")
    if uvs is not None:
        for t in np.asarray(uvs, dtype=np.float32):
            lines.append(f"vt {_f(t[0])} {_f(t[1])}")
    for a, b, c in faces + 1:
        if uvs is not None:
            lines.append(f"f {a}/{a}/{a} {b}/{b}/{b} {c}/{c}/{c}")
        else:
            lines.append(f"f {a}//{a} {b}//{b} {c}//{c}")
    with open(path, "w") as fh:
        fh.write("\n".join(lines))
        fh.write("\n")


def write_config(path: str, width: int, height: int, spp: int, depth: int, threads: int = 8,
                 accel: str = "BVH", integrator: str = "PathTracer", sampler: str = "Halton") -> None:
    cfg = {
        "Width": int(width), "Height": int(height), "SamplesPerPixel": int(spp),
        "Threads": int(threads), "Acceleration": accel,
        "Integrator": {"Type": integrator, "Depth": int(depth),
                       "Sampler": {"Type": sampler, "Seed": 0}},
    }
    with open(path, "w") as fh:
        json.dump(cfg, fh, indent=1)


@dataclass
class SceneFiles:
    directory: str
    scene_json: str
    config_json: str
    extra: dict = field(default_factory=dict)


def _mesh_object(obj_name: str, material: dict | None, lemit=None) -> dict:
    o = {"Type": "Geometry",
         "Shape": {"Type": "TriangleMesh", "ObjFile": obj_name,
                   "Translation": [0, 0, 0], "Scale": [1, 1, 1], "Rotation": [0, 0, 0]}}
    if material is not None:
        o["Material"] = material
    if lemit is not None:
        o["AreaLight"] = {"Type": "Diffused", "Lemit": [float(v) for v in lemit], "Samples": 1}
    return o


def _quad(corner, u, v, normal):
    """Two triangles spanning corner, corner+u, corner+u+v, corner+v, normals = normal."""
    c = np.asarray(corner, np.float64)
    u = np.asarray(u, np.float64)
    v = np.asarray(v, np.float64)
    pos = np.array([c, c + u, c + u + v, c + v], dtype=np.float32)
    nrm = np.tile(np.asarray(normal, np.float32), (4, 1))
    faces = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int64)
    # orient faces so the geometric normal agrees with the shading normal
    e1 = pos[1] - pos[0]
    e2 = pos[2] - pos[0]
    if np.dot(np.cross(e1, e2), normal) < 0:
        faces = faces[:, ::-1].copy()
    return pos, nrm, faces


def cornell(directory: str, width: int = 256, height: int = 256, spp: int = 16, depth: int = 5,
            threads: int = 8, accel: str = "BVH") -> SceneFiles:
    """The Cornell-box workload: box [-1,1]^3 open towards +z, camera on +z.

    Walls: floor/ceiling/back white (0.73), left red (0.65,0.05,0.05), right
    green (0.12,0.45,0.15). Light: 0.5x0.5 quad just below the ceiling facing
    down, Lemit 15, diffuse 0.73. Spheres r=0.3: diffuse white and mirror.
    Pinhole camera, vertical FOV 40 degrees.
    """
    os.makedirs(directory, exist_ok=True)
    white = {"Type": "Diffuse", "Color": [0.73, 0.73, 0.73]}
    red = {"Type": "Diffuse", "Color": [0.65, 0.05, 0.05]}
    green = {"Type": "Diffuse", "Color": [0.12, 0.45, 0.15]}
    walls = {
        "floor.obj": (_quad([-1, -1, -1], [2, 0, 0], [0, 0, 2], [0, 1, 0]), white),
        "ceiling.obj": (_quad([-1, 1, -1], [2, 0, 0], [0, 0, 2], [0, -1, 0]), white),
        "back.obj": (_quad([-1, -1, -1], [2, 0, 0], [0, 2, 0], [0, 0, 1]), white),
        "left.obj": (_quad([-1, -1, -1], [0, 0, 2], [0, 2, 0], [1, 0, 0]), red),
        "right.obj": (_quad([1, -1, -1], [0, 0, 2], [0, 2, 0], [-1, 0, 0]), green),
    }
    objects = []
    for name, ((p, n, f), mat) in walls.items():
        write_obj(os.path.join(directory, name), p, n, f)
        objects.append(_mesh_object(name, mat))
    lp, ln, lf = _quad([-0.25, 0.998, -0.25], [0.5, 0, 0], [0, 0, 0.5], [0, -1, 0])
    write_obj(os.path.join(directory, "light.obj"), lp, ln, lf)
    objects.append(_mesh_object("light.obj", white, lemit=[15, 15, 15]))
    objects.append({"Type": "Geometry",
                    "Shape": {"Type": "Sphere", "Position": [-0.4, -0.7, -0.3], "Radius": 0.3},
                    "Material": {"Type": "Diffuse", "Color": [0.73, 0.73, 0.73]}})
    objects.append({"Type": "Geometry",
                    "Shape": {"Type": "Sphere", "Position": [0.4, -0.7, 0.2], "Radius": 0.3},
                    "Material": {"Type": "Mirror", "Color": [0.95, 0.95, 0.95]}})
    scene = {"Camera": {"Type": "Normal", "Position": [0, 0, 3.7], "LookAt": [0, 0, 0],
                        "Up": [0, 1, 0], "FOV": 40},
             "Scene": {"Textures": {}, "Objects": objects}}
    sj = os.path.join(directory, "scene.json")
    with open(sj, "w") as fh:
        json.dump(scene, fh, indent=1)
    cj = os.path.join(directory, "config.json")
    write_config(cj, width, height, spp, depth, threads=threads, accel=accel)
    return SceneFiles(directory, sj, cj)


def material_zoo(directory: str, width: int = 64, height: int = 64, spp: int = 8, depth: int = 6,
                 threads: int = 8, accel: str = "BVH", phong: bool = True, extra: bool = False) -> SceneFiles:
    """Every supported material and light type in one small scene (parity fixture).

    Cornell-style box; ceiling area-light quad; an emissive analytic sphere (sphere
    area light); a point light; analytic spheres of specular glass, rough glass
    (GGX 0.2), isotropic GGX, anisotropic GGX, Blinn-Phong and mirror; a glass
    icosphere mesh; and a material-less (transparent) triangle pair. `extra` swaps the
    six spheres for the remaining materials: Oren-Nayar (two sigmas), Plastic
    (FresnelBlend) with GGX and anisotropic Beckmann, and Micro with isotropic and
    anisotropic Beckmann.
    """
    os.makedirs(directory, exist_ok=True)
    white = {"Type": "Diffuse", "Color": [0.73, 0.73, 0.73]}
    objects = [{"Type": "Light", "LightType": "Point", "Position": [0.5, 0.6, 0.5], "Flux": [4, 4, 4]}]
    walls = {
        "floor.obj": (_quad([-1, -1, -1], [2, 0, 0], [0, 0, 2], [0, 1, 0]), white),
        "ceiling.obj": (_quad([-1, 1, -1], [2, 0, 0], [0, 0, 2], [0, -1, 0]), white),
        "back.obj": (_quad([-1, -1, -1], [2, 0, 0], [0, 2, 0], [0, 0, 1]),
                     {"Type": "Diffuse", "Color": [0.4, 0.5, 0.7]}),
        "left.obj": (_quad([-1, -1, -1], [0, 0, 2], [0, 2, 0], [1, 0, 0]),
                     {"Type": "Diffuse", "Color": [0.65, 0.05, 0.05]}),
        "right.obj": (_quad([1, -1, -1], [0, 0, 2], [0, 2, 0], [-1, 0, 0]),
                      {"Type": "Diffuse", "Color": [0.12, 0.45, 0.15]}),
    }
    for name, ((p, n, f), mat) in walls.items():
        write_obj(os.path.join(directory, name), p, n, f)
        objects.append(_mesh_object(name, mat))
    lp, ln, lf = _quad([-0.3, 0.998, -0.3], [0.6, 0, 0], [0, 0, 0.6], [0, -1, 0])
    write_obj(os.path.join(directory, "light.obj"), lp, ln, lf)
    objects.append(_mesh_object("light.obj", white, lemit=[10, 10, 10]))
    objects.append({"Type": "Geometry", "Shape": {"Type": "Sphere", "Position": [-0.7, 0.6, -0.6], "Radius": 0.08},
                    "Material": {"Type": "Diffuse", "Color": [0.9, 0.9, 0.9]},
                    "AreaLight": {"Type": "Diffused", "Lemit": [6, 4, 2], "Samples": 1}})
    spheres = [
        ([-0.55, -0.72, -0.35], 0.27, {"Type": "Glass", "Color": [1, 1, 1], "Roughness": [0, 0], "IOR": 1.5}),
        ([0.0, -0.72, -0.45], 0.27, {"Type": "Glass", "Color": [0.95, 0.95, 1], "Roughness": [0.2, 0.2], "IOR": 1.5}),
        ([0.55, -0.72, -0.35], 0.27, {"Type": "Micro", "Color": [0.9, 0.6, 0.3], "Roughness": [0.3, 0.3],
                                      "IOR": 1.5, "NDF": "GGX"}),
        ([-0.45, -0.2, 0.25], 0.22, {"Type": "Micro", "Color": [0.6, 0.8, 0.9], "Roughness": [0.1, 0.4],
                                     "IOR": 1.5, "NDF": "GGX"}),
        ([0.45, -0.2, 0.25], 0.22, {"Type": "Phong", "K": 20, "Color": [0.8, 0.8, 0.5]} if phong
         else {"Type": "Diffuse", "Color": [0.8, 0.8, 0.5]}),
        ([0.0, 0.3, -0.5], 0.2, {"Type": "Mirror", "Color": [0.95, 0.95, 0.95]}),
    ]
    if extra:
        spheres = [
            ([-0.55, -0.72, -0.35], 0.27, {"Type": "OrenNayar", "Color": [0.8, 0.6, 0.4], "Sigma": 0.5}),
            ([0.0, -0.72, -0.45], 0.27, {"Type": "Plastic", "Kd": [0.5, 0.1, 0.1], "Ks": [0.4, 0.4, 0.4],
                                         "Roughness": [0.2, 0.2], "NDF": "GGX"}),
            ([0.55, -0.72, -0.35], 0.27, {"Type": "Micro", "Color": [0.9, 0.6, 0.3], "Roughness": [0.3, 0.3],
                                          "IOR": 1.5, "NDF": "Beckmann"}),
            ([-0.45, -0.2, 0.25], 0.22, {"Type": "Micro", "Color": [0.6, 0.8, 0.9], "Roughness": [0.1, 0.4],
                                         "IOR": 1.5, "NDF": "Beckmann"}),
            ([0.45, -0.2, 0.25], 0.22, {"Type": "Plastic", "Kd": [0.1, 0.3, 0.6], "Ks": [0.3, 0.3, 0.3],
                                        "Roughness": [0.15, 0.35], "NDF": "Beckmann"}),
            ([0.0, 0.3, -0.5], 0.2, {"Type": "OrenNayar", "Color": [0.9, 0.9, 0.9], "Sigma": 1.2}),
        ]
    for pos, r, mat in spheres:
        objects.append({"Type": "Geometry", "Shape": {"Type": "Sphere", "Position": pos, "Radius": r},
                        "Material": mat})
    v, f = icosphere_mesh(2)
    write_obj(os.path.join(directory, "glassball.obj"), v * 0.22 + np.array([0.0, -0.78, 0.35]), v, f)
    objects.append(_mesh_object("glassball.obj", {"Type": "Glass", "Color": [1, 1, 1], "Roughness": [0, 0],
                                                  "IOR": 1.5}))
    cp, cn, cf = _quad([-0.9, -1.0, 0.6], [0.5, 0, 0], [0, 0.5, 0], [0, 0, 1])
    write_obj(os.path.join(directory, "curtain.obj"), cp, cn, cf)
    objects.append(_mesh_object("curtain.obj", None))
    scene = {"Camera": {"Type": "Normal", "Position": [0, 0, 3.7], "LookAt": [0, 0, 0], "Up": [0, 1, 0],
                        "FOV": 40},
             "Scene": {"Textures": {}, "Objects": objects}}
    sj = os.path.join(directory, "scene.json")
    with open(sj, "w") as fh:
        json.dump(scene, fh, indent=1)
    cj = os.path.join(directory, "config.json")
    write_config(cj, width, height, spp, depth, threads=threads, accel=accel)
    return SceneFiles(directory, sj, cj)


def disk_scene(directory: str, width: int = 64, height: int = 64, spp: int = 8, depth: int = 6,
               threads: int = 8, accel: str = "BVH") -> SceneFiles:
    """Disks in every role (parity fixture for the Disk shape).

    Cornell-style box lit by a disk area light under the ceiling (rotated 180 degrees about x,
    so it faces down) and a small tilted emissive disk seen partly from behind; disks with a
    mirror (on the floor, no Rotation key), diffuse (upright), glass, GGX and Oren-Nayar
    material at several rotations; one diffuse sphere and a point light.
    """
    os.makedirs(directory, exist_ok=True)
    white = {"Type": "Diffuse", "Color": [0.73, 0.73, 0.73]}
    objects = [{"Type": "Light", "LightType": "Point", "Position": [-0.5, 0.5, 0.6], "Flux": [2, 2, 2]}]
    walls = {
        "floor.obj": (_quad([-1, -1, -1], [2, 0, 0], [0, 0, 2], [0, 1, 0]), white),
        "ceiling.obj": (_quad([-1, 1, -1], [2, 0, 0], [0, 0, 2], [0, -1, 0]), white),
        "back.obj": (_quad([-1, -1, -1], [2, 0, 0], [0, 2, 0], [0, 0, 1]), white),
        "left.obj": (_quad([-1, -1, -1], [0, 0, 2], [0, 2, 0], [1, 0, 0]),
                     {"Type": "Diffuse", "Color": [0.65, 0.05, 0.05]}),
        "right.obj": (_quad([1, -1, -1], [0, 0, 2], [0, 2, 0], [-1, 0, 0]),
                      {"Type": "Diffuse", "Color": [0.12, 0.45, 0.15]}),
    }
    for name, ((p, n, f), mat) in walls.items():
        write_obj(os.path.join(directory, name), p, n, f)
        objects.append(_mesh_object(name, mat))

    def disk(pos, r, rot, mat, lemit=None):
        sh = {"Type": "Disk", "Position": pos, "Radius": r}
        if rot is not None:
            sh["Rotation"] = rot
        o = {"Type": "Geometry", "Shape": sh, "Material": mat}
        if lemit is not None:
            o["AreaLight"] = {"Type": "Diffused", "Lemit": lemit, "Samples": 1}
        objects.append(o)

    disk([0.0, 0.995, 0.0], 0.35, [180, 0, 0], white, lemit=[8, 8, 8])
    disk([0.62, 0.35, -0.55], 0.15, [45, 30, 0], white, lemit=[3, 5, 8])
    disk([-0.35, -0.99, -0.1], 0.45, None, {"Type": "Mirror", "Color": [0.95, 0.95, 0.95]})
    disk([0.45, -0.55, -0.3], 0.3, [90, 0, 20], {"Type": "Diffuse", "Color": [0.8, 0.7, 0.3]})
    disk([-0.45, 0.05, 0.2], 0.25, [60, 0, 0], {"Type": "Glass", "Color": [1, 1, 1], "Roughness": [0, 0],
                                                "IOR": 1.5})
    disk([0.0, -0.3, -0.75], 0.3, [-30, 45, 0], {"Type": "Micro", "Color": [0.9, 0.6, 0.3],
                                                 "Roughness": [0.2, 0.2], "IOR": 1.5, "NDF": "GGX"})
    disk([0.5, 0.3, 0.3], 0.2, [0, 0, 120], {"Type": "OrenNayar", "Color": [0.6, 0.8, 0.9], "Sigma": 0.7})
    objects.append({"Type": "Geometry", "Shape": {"Type": "Sphere", "Position": [0.05, -0.8, 0.35], "Radius": 0.2},
                    "Material": {"Type": "Diffuse", "Color": [0.9, 0.9, 0.9]}})
    scene = {"Camera": {"Type": "Normal", "Position": [0, 0, 3.7], "LookAt": [0, 0, 0], "Up": [0, 1, 0],
                        "FOV": 40},
             "Scene": {"Textures": {}, "Objects": objects}}
    sj = os.path.join(directory, "scene.json")
    with open(sj, "w") as fh:
        json.dump(scene, fh, indent=1)
    cj = os.path.join(directory, "config.json")
    write_config(cj, width, height, spp, depth, threads=threads, accel=accel)
    return SceneFiles(directory, sj, cj)


# ----------------------------------------------------------------------------------------------
# icosphere + value noise (ray-cast microbenchmark)
# ----------------------------------------------------------------------------------------------

def icosphere_mesh(level: int):
    """Unit icosphere: (vertices float64 [V,3], faces int64 [20*4**level, 3]).

    Faces are wound counter-clockwise seen from outside. Midpoints are shared
    between neighbouring faces, so V = 10 * 4**level + 2.
    """
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0),
             (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    v = np.array(verts, dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    for _ in range(level):
        nf = f.shape[0]
        edges = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)
        key = np.sort(edges, axis=1)
        uniq, inv = np.unique(key[:, 0] * (len(v) + 1) + key[:, 1], return_inverse=True)
        a = uniq // (len(v) + 1)
        b = uniq % (len(v) + 1)
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mid], axis=0)
        m01 = base + inv[:nf]
        m12 = base + inv[nf:2 * nf]
        m20 = base + inv[2 * nf:]
        f = np.concatenate([
            np.stack([f[:, 0], m01, m20], 1), np.stack([f[:, 1], m12, m01], 1),
            np.stack([f[:, 2], m20, m12], 1), np.stack([m01, m12, m20], 1)], axis=0)
    return v, f


def _hash3(ix, iy, iz, seed):
    h = (ix.astype(np.uint64) * np.uint64(0x8DA6B343) ^ iy.astype(np.uint64) * np.uint64(0xD8163841)
         ^ iz.astype(np.uint64) * np.uint64(0xCB1AB31F) ^ np.uint64(seed) * np.uint64(0x9E3779B1))
    h &= np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(13)
    h = (h * np.uint64(0x5BD1E995)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(15)
    return (h & np.uint64(0xFFFFFF)).astype(np.float64) / float(1 << 24)


def value_noise(p: np.ndarray, seed: int, freq: float = 4.0) -> np.ndarray:
    """Trilinear lattice value noise in [0,1): integer-hash lattice, smoothstep fade."""
    q = p * freq
    i = np.floor(q)
    fr = q - i
    w = fr * fr * (3 - 2 * fr)
    ii = i.astype(np.int64)
    out = np.zeros(len(p))
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                c = _hash3(ii[:, 0] + dx, ii[:, 1] + dy, ii[:, 2] + dz, seed)
                wx = w[:, 0] if dx else 1 - w[:, 0]
                wy = w[:, 1] if dy else 1 - w[:, 1]
                wz = w[:, 2] if dz else 1 - w[:, 2]
                out += c * wx * wy * wz
    return out


def displaced_icosphere(level: int, seed: int = 7):
    """Icosphere with vertex radius 1 + 0.1 * value_noise(seed); float32 positions/normals."""
    v, f = icosphere_mesh(level)
    r = 1.0 + 0.1 * value_noise(v, seed)
    pos = (v * r[:, None]).astype(np.float32)
    # area-weighted vertex normals of the displaced surface
    p64 = pos.astype(np.float64)
    fn = np.cross(p64[f[:, 1]] - p64[f[:, 0]], p64[f[:, 2]] - p64[f[:, 0]])
    vn = np.zeros_like(p64)
    for k in range(3):
        np.add.at(vn, f[:, k], fn)
    vn /= np.linalg.norm(vn, axis=1, keepdims=True)
    return pos, vn.astype(np.float32), f


def icosphere_rays(n: int, pos: np.ndarray, seed: int = 11) -> np.ndarray:
    """n rays: origins uniform in 1.5x the bounding box (about its centre), directions
    uniform on S^2 (z = 1 - 2u, phi = 2 pi v). Returns float32 [n, 8] =
    (o.xyz, d.xyz, tmin=0, tmax=inf), generated once on the host."""
    rng = np.random.default_rng(seed)
    lo = pos.min(0).astype(np.float64)
    hi = pos.max(0).astype(np.float64)
    c = (lo + hi) / 2
    h = (hi - lo) / 2 * 1.5
    o = c + (rng.random((n, 3)) * 2 - 1) * h
    z = 1 - 2 * rng.random(n)
    phi = 2 * np.pi * rng.random(n)
    s = np.sqrt(np.maximum(0, 1 - z * z))
    d = np.stack([s * np.cos(phi), s * np.sin(phi), z], 1)
    out = np.zeros((n, 8), np.float32)
    out[:, 0:3] = o
    out[:, 3:6] = d
    out[:, 6] = 0
    out[:, 7] = np.inf
    return out


# ----------------------------------------------------------------------------------------------
# ground + sphere meshes + area lights (full path-tracing workload)
# ----------------------------------------------------------------------------------------------

def spheres_scene_objects(n_spheres: int = 400, level: int = 4, seed: int = 2024,
                          extent: float = 12.0):
    """Draw the randomized sphere-field scene. Returns a dict of python lists that
    both the JSON writer and the in-memory scene builder consume:

    ``meshes``: list of (positions f32[V,3], normals f32[V,3], faces i64[F,3], material, lemit|None)
    ``camera``: dict in the reference's camera schema.
    Materials are drawn uniformly from diffuse (albedo U(0.1,0.9)), GGX
    (roughness U(0.05,0.5)), anisotropic GGX (alpha_x, alpha_y U(0.05,0.5)),
    glass (ior 1.5) and microfacet glass (roughness 0.2).
    """
    rng = np.random.default_rng(seed)
    unit_v, faces = icosphere_mesh(level)
    meshes = []
    g = extent
    gp, gn, gf = _quad([-g, 0, -g], [0, 0, 2 * g], [2 * g, 0, 0], [0, 1, 0])
    meshes.append((gp, gn, gf, {"Type": "Diffuse", "Color": [0.5, 0.5, 0.5]}, None))
    for _ in range(n_spheres):
        r = rng.uniform(0.2, 1.0)
        cx, cz = rng.uniform(-extent * 0.9, extent * 0.9, 2)
        kind = int(rng.integers(0, 5))
        if kind == 0:
            mat = {"Type": "Diffuse", "Color": [float(x) for x in rng.uniform(0.1, 0.9, 3)]}
        elif kind == 1:
            a = float(rng.uniform(0.05, 0.5))
            mat = {"Type": "Micro", "Color": [float(x) for x in rng.uniform(0.3, 0.95, 3)],
                   "Roughness": [a, a], "IOR": 1.5, "NDF": "GGX"}
        elif kind == 2:
            ax, ay = (float(x) for x in rng.uniform(0.05, 0.5, 2))
            mat = {"Type": "Micro", "Color": [float(x) for x in rng.uniform(0.3, 0.95, 3)],
                   "Roughness": [ax, ay], "IOR": 1.5, "NDF": "GGX"}
        elif kind == 3:
            mat = {"Type": "Glass", "Color": [1.0, 1.0, 1.0], "Roughness": [0, 0], "IOR": 1.5}
        else:
            mat = {"Type": "Glass", "Color": [1.0, 1.0, 1.0], "Roughness": [0.2, 0.2], "IOR": 1.5}
        pos = (unit_v * r + np.array([cx, r, cz])).astype(np.float32)
        nrm = unit_v.astype(np.float32)
        meshes.append((pos, nrm, faces, mat, None))
    for _ in range(4):
        lx, lz = rng.uniform(-extent * 0.6, extent * 0.6, 2)
        ly = rng.uniform(6.0, 9.0)
        s = rng.uniform(1.0, 2.5)
        le = float(rng.uniform(5, 20))
        lp, ln, lf = _quad([lx - s / 2, ly, lz - s / 2], [s, 0, 0], [0, 0, s], [0, -1, 0])
        meshes.append((lp, ln, lf, None, [le, le, le]))
    camera = {"Type": "Normal", "Position": [0.0, 7.0, 22.0], "LookAt": [0.0, 0.5, 0.0],
              "Up": [0, 1, 0], "FOV": 45}
    return {"meshes": meshes, "camera": camera}


def spheres_scene(directory: str, width: int = 1920, height: int = 1080, spp: int = 64,
                  depth: int = 8, n_spheres: int = 400, level: int = 4, seed: int = 2024,
                  threads: int = 8) -> SceneFiles:
    """Write the sphere-field scene in reference JSON + OBJ form (one OBJ per mesh)."""
    os.makedirs(directory, exist_ok=True)
    d = spheres_scene_objects(n_spheres, level, seed)
    objects = []
    for k, (p, n, f, mat, le) in enumerate(d["meshes"]):
        name = f"mesh{k:04d}.obj"
        write_obj(os.path.join(directory, name), p, n, f)
        objects.append(_mesh_object(name, mat, lemit=le))
    scene = {"Camera": d["camera"], "Scene": {"Textures": {}, "Objects": objects}}
    sj = os.path.join(directory, "scene.json")
    with open(sj, "w") as fh:
        json.dump(scene, fh, indent=1)
    cj = os.path.join(directory, "config.json")
    write_config(cj, width, height, spp, depth, threads=threads)
    return SceneFiles(directory, sj, cj)


def build_spheres_scene(n_spheres: int = 400, level: int = 4, seed: int = 2024, extent: float = 12.0):
    """The sphere-field scene as an in-memory Scene (same draws as spheres_scene)."""
    import crystal2_amd as c2
    d = spheres_scene_objects(n_spheres, level, seed, extent)
    sc = c2.Scene.create()
    for (p, n, f, mat, le) in d["meshes"]:
        if mat is None:
            m = -1
        else:
            mt = {"Diffuse": 0, "Mirror": 1, "Glass": 2, "Micro": 3, "Phong": 4}[mat["Type"]]
            m = sc.add_material(mt, mat["Color"], mat.get("Roughness", (0, 0)), mat.get("IOR", 1.0), mat.get("K", 0))
        sc.add_mesh(p, n, f, m, lemit=le)
    cam = d["camera"]
    sc.set_camera(cam["Position"], cam["LookAt"], cam["Up"], cam["FOV"])
    return sc


def batch_specs(n: int = 64, seed: int = 6400):
    """The batched-scenes workload: n sphere-field scenes with triangle counts
    log-uniform in [1k, 200k], square resolutions uniform in {128, 256, 512} and
    spp uniform in {4, 16, 64}. Spheres are icospheres of 320, 1280 or 5120
    triangles (level 2 below 10k triangles, 3 below 50k, else 4). Returns a list of
    dicts (n_spheres, level, seed, width, height, spp, triangles)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        target = float(np.exp(rng.uniform(np.log(1e3), np.log(2e5))))
        level = 2 if target < 1e4 else (3 if target < 5e4 else 4)
        per = 20 * 4 ** level
        ns = max(1, int(round((target - 10) / per)))
        res = int(rng.choice([128, 256, 512]))
        spp = int(rng.choice([4, 16, 64]))
        out.append({"n_spheres": ns, "level": level, "seed": seed + 1 + i, "width": res, "height": res,
                    "spp": spp, "triangles": ns * per + 10})
    return out


def build_batch(specs):
    """[(Scene, width, height, spp)] for batch_specs entries."""
    return [(build_spheres_scene(s["n_spheres"], s["level"], s["seed"]), s["width"], s["height"], s["spp"])
            for s in specs]
