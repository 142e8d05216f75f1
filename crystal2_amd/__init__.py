"""crystal2_amd -- MI355X (gfx950) wavefront path tracer for Crystal2's PathTracer + MIS path.

Python convenience layer over the C-ABI in ``include/crystal2.h`` (the product is
the native library ``crystal2_amd/_lib/libcrystal2.so``; this module only moves
arrays and handles). Names follow the reference's plugin interfaces:

* :class:`Scene`       -- ``Scene::CreateScene`` from the reference JSON/OBJ formats,
                          or built from arrays.
* :class:`GpuBVH`      -- ``Accelerator`` (``Intersect`` / ``IntersectTest``) on the GPU.
* :class:`WavefrontPathIntegrator` -- ``Integrator::Render`` on the GPU.

Every call fails loudly (``Crystal2Error``) when the native library is missing
or the device reports an error; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _abi
from ._abi import Crystal2Error, check

__all__ = ["Scene", "GpuBVH", "WavefrontPathIntegrator", "RadianceCache", "Config", "Crystal2Error",
           "halton_index", "halton_value", "load_library", "load_image", "write_hdr", "write_pfm"]


def load_library():
    return _abi.load()


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def write_hdr(path: str, film: np.ndarray):
    """The reference's result.hdr (stbi_write_hdr) for an (h, w, 3) float32 film."""
    f = np.ascontiguousarray(film, np.float32)
    check(load_library().c2_film_write_hdr(path.encode(), f.shape[1], f.shape[0], _fp(f)))


def write_pfm(path: str, film: np.ndarray):
    """Lossless little-endian PFM of an (h, w, 3) float32 film."""
    f = np.ascontiguousarray(film, np.float32)
    check(load_library().c2_film_write_pfm(path.encode(), f.shape[1], f.shape[0], _fp(f)))


def load_image(path: str) -> np.ndarray:
    """stbi_load(path, 3) as the reference's ImageTexels calls it: (height, width, 3) uint8."""
    lib = load_library()
    w, h = C.c_int32(), C.c_int32()
    buf = C.POINTER(C.c_uint8)()
    check(lib.c2_image_load(path.encode(), C.byref(w), C.byref(h), C.byref(buf)))
    try:
        return np.ctypeslib.as_array(buf, shape=(h.value * w.value * 3,)).reshape(h.value, w.value, 3).copy()
    finally:
        lib.c2_image_free(buf)


@dataclass
class Config:
    """Render configuration, the reference's config JSON (ConfigLoader.cpp:12-16, Integrator.cpp:19-35)."""
    width: int
    height: int
    spp: int
    threads: int
    max_depth: int
    accel: str = "BVH"
    integrator: str = "PathTracer"
    sampler: str = "Halton"

    @classmethod
    def load(cls, path: str) -> "Config":
        lib = load_library()
        c = _abi.c2_config()
        check(lib.c2_config_load(path.encode(), C.byref(c)))
        return cls(c.width, c.height, c.spp, c.threads, c.max_depth, c.accel.decode(), c.integrator.decode(),
                   c.sampler.decode())


class Scene:
    """Owns a host scene (entities, triangles, spheres, materials, lights, camera)."""

    def __init__(self, handle):
        self._lib = load_library()
        self._h = handle

    @classmethod
    def load(cls, scene_json: str, base_dir: str | None = None) -> "Scene":
        lib = load_library()
        h = C.c_void_p()
        check(lib.c2_scene_load(scene_json.encode(), base_dir.encode() if base_dir else None, C.byref(h)))
        return cls(h)

    @classmethod
    def create(cls) -> "Scene":
        lib = load_library()
        h = C.c_void_p()
        check(lib.c2_scene_create(C.byref(h)))
        return cls(h)

    @classmethod
    def load_binary(cls, path: str) -> "Scene":
        """A scene written by :meth:`save` (the binary scene cache)."""
        lib = load_library()
        h = C.c_void_p()
        check(lib.c2_scene_load_binary(path.encode(), C.byref(h)))
        return cls(h)

    def save(self, path: str):
        check(self._lib.c2_scene_save(self._h, path.encode()))

    def add_material(self, mtype: int, color, roughness=(0.0, 0.0), ior=1.0, phong_k=0, ndf=0, ks=(0.0, 0.0, 0.0),
                     sigma=0.0, kd_texture: int = 0) -> int:
        """c2_material fields: `color` is the albedo (Plastic: Kd), `ks` Plastic's Ks,
        `sigma` OrenNayar's Sigma, `ndf` 0 = GGX, 1 = Beckmann (Micro, Plastic),
        `kd_texture` a texture id from :meth:`add_texture` (Diffuse, Phong; 0 = `color`)."""
        m = _abi.c2_material()
        m.type = mtype
        m.ndf = ndf
        m.phong_k = phong_k
        m.kd_texture = int(kd_texture)
        m.color[:] = [float(x) for x in color]
        m.roughness[:] = [float(x) for x in roughness]
        m.ior = float(ior)
        m.ks[:] = [float(x) for x in ks]
        m.sigma = float(sigma)
        idx = C.c_int32()
        check(self._lib.c2_scene_add_material(self._h, C.byref(m), C.byref(idx)))
        return idx.value

    def add_triangles(self, vertex_data: np.ndarray, material: int, lemit=None, transform=None):
        """vertex_data: float32 [ntri, 3, 8] (pos, normal, uv) -- unshared corners."""
        vd = np.ascontiguousarray(vertex_data, dtype=np.float32).reshape(-1, 3, 8)
        le = None if lemit is None else np.asarray(lemit, np.float32)
        tr = None if transform is None else np.ascontiguousarray(transform, np.float32).reshape(16)
        check(self._lib.c2_scene_add_triangles(self._h, _fp(vd), vd.shape[0], None if tr is None else _fp(tr),
                                               material, None if le is None else _fp(le)))

    def add_mesh(self, positions, normals, faces, material: int, lemit=None, uv=None):
        p = np.asarray(positions, np.float32)
        n = np.asarray(normals, np.float32)
        f = np.asarray(faces, np.int64)
        vd = np.zeros((len(f), 3, 8), np.float32)
        vd[:, :, 0:3] = p[f]
        vd[:, :, 3:6] = n[f]
        if uv is not None:
            vd[:, :, 6:8] = np.asarray(uv, np.float32)[f]
        self.add_triangles(vd, material, lemit)

    def add_sphere(self, center, radius: float, material: int, lemit=None, rotation=(0.0, 0.0, 0.0)):
        c = np.asarray(center, np.float32)
        r = np.asarray(rotation, np.float32)
        le = None if lemit is None else np.asarray(lemit, np.float32)
        check(self._lib.c2_scene_add_sphere(self._h, _fp(c), float(radius), _fp(r), material,
                                            None if le is None else _fp(le)))

    def add_disk(self, center, radius: float, material: int, lemit=None, rotation=(0.0, 0.0, 0.0)):
        """Disk (Disk.cpp): facing local +y before the x, y, z rotations (degrees)."""
        c = np.asarray(center, np.float32)
        r = np.asarray(rotation, np.float32)
        le = None if lemit is None else np.asarray(lemit, np.float32)
        check(self._lib.c2_scene_add_disk(self._h, _fp(c), float(radius), _fp(r), material,
                                          None if le is None else _fp(le)))

    def add_point_light(self, position, flux):
        p = np.asarray(position, np.float32)
        f = np.asarray(flux, np.float32)
        check(self._lib.c2_scene_add_point_light(self._h, _fp(p), _fp(f)))

    def add_texture(self, rgb: np.ndarray) -> int:
        """ImageTexels from an (height, width, 3) uint8 image, rows top-down; returns its id."""
        a = np.ascontiguousarray(rgb, dtype=np.uint8)
        if a.ndim != 3 or a.shape[2] != 3:
            raise ValueError("texture must be (height, width, 3) uint8")
        tid = C.c_int32()
        check(self._lib.c2_scene_add_texture(self._h, a.shape[1], a.shape[0],
                                             a.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(tid)))
        return tid.value

    def set_skybox(self, face_ids):
        """Cubemap faces (posx, negx, posy, negy, posz, negz) as texture ids; all 0 removes it."""
        ids = np.asarray(face_ids, np.int32)
        check(self._lib.c2_scene_set_skybox(self._h, _ip(ids)))

    def set_camera(self, position, look_at, up, fov_degrees: float):
        cam = _abi.c2_camera()
        cam.position[:] = [float(x) for x in position]
        cam.look_at[:] = [float(x) for x in look_at]
        cam.up[:] = [float(x) for x in up]
        # Camera::CreateCamera: FOV / 180.f * pi<float>() evaluated in double, then float
        cam.fov = float(np.float32(float(fov_degrees) / float(np.float32(180.0)) * float(np.float32(np.pi))))
        check(self._lib.c2_scene_set_camera(self._h, C.byref(cam)))

    def desc(self) -> _abi.c2_scene_desc:
        d = _abi.c2_scene_desc()
        check(self._lib.c2_scene_get_desc(self._h, C.byref(d)))
        return d

    @property
    def num_entities(self) -> int:
        return self.desc().num_entities

    def close(self):
        if self._h:
            self._lib.c2_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuBVH:
    """Accelerator on the GPU (src/Accelerators/Accelerator.h:8-22): Intersect / IntersectTest."""

    ACCELS = {"bvh": _abi.C2_ACCEL_BVH, "BVH": _abi.C2_ACCEL_BVH, "GPUBVH": _abi.C2_ACCEL_BVH,
              "kdtree": _abi.C2_ACCEL_KDTREE, "KDTree": _abi.C2_ACCEL_KDTREE,
              "bvh8": _abi.C2_ACCEL_BVH8, "GPUBVH8": _abi.C2_ACCEL_BVH8}

    def __init__(self, scene: Scene, accel: str = "bvh"):
        """``accel``: "bvh" (the GPU BVH, 4 wide), "bvh8" (the same tree 8 wide) or "kdtree"
        (the reference's KD-tree, "KDTree")."""
        self._lib = load_library()
        self._scene = scene
        self._h = C.c_void_p()
        d = scene.desc()
        check(self._lib.c2_accel_create(C.byref(d), self.ACCELS[accel], C.byref(self._h)))

    def stats(self) -> dict:
        s = _abi.c2_accel_stats()
        check(self._lib.c2_accel_get_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def intersect(self, rays: np.ndarray):
        """rays float32 [n, 8] -> (entity int32 [n] (-1 miss), t float32 [n])."""
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = r.shape[0]
        prim = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        check(self._lib.c2_accel_intersect(self._h, _fp(r), n, _ip(prim), _fp(t)))
        return prim, t

    def intersect_test(self, rays: np.ndarray, ignore: np.ndarray | None = None):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = r.shape[0]
        hit = np.empty(n, np.uint8)
        ig = None if ignore is None else np.ascontiguousarray(ignore, np.int32)
        check(self._lib.c2_accel_intersect_test(self._h, _fp(r), None if ig is None else _ip(ig), n,
                                                hit.ctypes.data_as(C.POINTER(C.c_uint8))))
        return hit.astype(bool)

    def intersect_device(self, rays_ptr: int, n: int, prim_ptr: int, t_ptr: int, stream: int = 0):
        check(self._lib.c2_accel_intersect_device(self._h, rays_ptr, n, prim_ptr, t_ptr, stream or None))

    def intersect_test_device(self, rays_ptr: int, ignore_ptr: int, n: int, hit_ptr: int, stream: int = 0):
        check(self._lib.c2_accel_intersect_test_device(self._h, rays_ptr, ignore_ptr or None, n, hit_ptr,
                                                       stream or None))

    def count_visits(self, rays: np.ndarray, any_hit: bool = False):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        a, b = C.c_double(), C.c_double()
        check(self._lib.c2_accel_count_visits(self._h, _fp(r), r.shape[0], int(any_hit), C.byref(a), C.byref(b)))
        return a.value, b.value

    def close(self):
        if self._h:
            self._lib.c2_accel_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class WavefrontPathIntegrator:
    """Integrator::Render (PathTracer + MIS, Halton) on the GPU.

    ``arg_order``: order in which the reference's compiler evaluated
    ``SampleDirection(sampler->Get1D(), sampler->Get2D(), ...)``
    (PathTracingIntegrator.cpp:79): "rtl" (GCC, MSVC) or "ltr" (clang).
    ``film``: "f32" (reference FrameBuffer), "kahan" or "f64".
    ``integrator``: "path" (PathTracingIntegrator) or "direct" (DirectLightingIntegrator,
    DirectLightingIntegrator.cpp:22-156, on the same kernels).
    ``accel``: "bvh", "bvh8" or "kdtree" (config "Acceleration": the reference's KD-tree).
    """

    FILMS = {"f32": _abi.C2_FILM_F32, "kahan": _abi.C2_FILM_F32_KAHAN, "f64": _abi.C2_FILM_F64}
    INTEGRATORS = {"path": _abi.C2_INTEGRATOR_PATH, "direct": _abi.C2_INTEGRATOR_DIRECT,
                   "PathTracer": _abi.C2_INTEGRATOR_PATH, "DirectLight": _abi.C2_INTEGRATOR_DIRECT}

    def __init__(self, scene: Scene, width: int, height: int, spp: int, max_depth: int, arg_order: str = "rtl",
                 film: str = "f32", paths_in_flight: int = 0, window=None, queue_blocks: int = 0,
                 integrator: str = "path", accel: str = "bvh"):
        self._lib = load_library()
        self.width, self.height, self.spp, self.max_depth = width, height, spp, max_depth
        o = _abi.c2_render_opts()
        o.accel = GpuBVH.ACCELS[accel]
        o.width, o.height, o.spp, o.max_depth = width, height, spp, max_depth
        o.arg_order = _abi.C2_ARGORDER_LTR if arg_order == "ltr" else _abi.C2_ARGORDER_RTL
        o.film_mode = self.FILMS[film]
        o.paths_in_flight = paths_in_flight
        o.queue_blocks = queue_blocks
        o.integrator = self.INTEGRATORS[integrator]
        if window is not None:
            o.win_x0, o.win_y0, o.win_w, o.win_h = (int(v) for v in window)
        self.window = window
        self._h = C.c_void_p()
        d = scene.desc()
        check(self._lib.c2_pt_create(C.byref(d), C.byref(o), C.byref(self._h)))

    def render(self, sample_begin: int = 0, sample_count: int | None = None, stream: int = 0):
        n = self.spp - sample_begin if sample_count is None else sample_count
        check(self._lib.c2_pt_render(self._h, sample_begin, n, stream or None))

    def sync(self):
        check(self._lib.c2_pt_sync(self._h))

    def reset(self):
        check(self._lib.c2_pt_reset(self._h))

    def film(self) -> np.ndarray:
        out = np.empty((self.height, self.width, 3), np.float32)
        check(self._lib.c2_pt_get_film(self._h, _fp(out)))
        return out

    def set_cache(self, cache: "RadianceCache | None", cache_bounce: int = 2, train_ratio: int = 64):
        """Terminate camera paths at `cache_bounce` into the cache; train it online."""
        self._cache = cache
        check(self._lib.c2_pt_set_cache(self._h, cache._h if cache else None, cache_bounce, train_ratio))

    def cache_times(self) -> dict:
        t = _abi.c2_cache_times()
        check(self._lib.c2_pt_get_cache_times(self._h, C.byref(t)))
        return {"infer_ms": t.infer_ms, "train_ms": t.train_ms, "collect_ms": t.collect_ms}

    def set_profiling(self, enable: bool = True, kernels=None):
        """Time kernels with HIP events: all classes, or only those named in ``kernels``
        (names of :data:`_abi.KERNEL_CLASSES`)."""
        if kernels is None:
            check(self._lib.c2_pt_set_profiling(self._h, int(enable)))
            return
        mask = 0
        for k in kernels:
            mask |= 1 << _abi.KERNEL_CLASSES.index(k)
        check(self._lib.c2_pt_profile_kernels(self._h, mask if enable else 0))

    def kernel_times(self) -> dict:
        k = _abi.c2_kernel_times()
        check(self._lib.c2_pt_get_kernel_times(self._h, C.byref(k)))
        return {name: {"ms": k.ms[i], "launches": k.launches[i]} for i, name in enumerate(_abi.KERNEL_CLASSES)}

    def capture_rays(self, sample: int, bounce: int, kind: str = "shadow", n: int = 1 << 16):
        rays = np.zeros((n, 8), np.float32)
        ign = np.zeros(n, np.int32)
        w = C.c_int64()
        check(self._lib.c2_pt_capture_rays(self._h, sample, bounce, 1 if kind == "shadow" else 0, _fp(rays), _ip(ign),
                                           n, C.byref(w)))
        return rays[:w.value], ign[:w.value]

    def stats(self) -> dict:
        s = _abi.c2_pt_stats()
        check(self._lib.c2_pt_get_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def close(self):
        if self._h:
            self._lib.c2_pt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchRenderer:
    """Many scenes rendered together (no counterpart in the reference, which renders
    one scene per run): sample pass k traces every pixel of every scene with
    spp > k in shared launches; ``render(use_graph=True)`` replays the whole render
    as one captured hipGraph. Each film equals the scene's own
    ``WavefrontPathIntegrator(scene, w, h, spp, depth).render()`` bit for bit.

    ``items``: sequence of (Scene, width, height, spp).
    """

    def __init__(self, items, max_depth: int, arg_order: str = "rtl", film: str = "f32",
                 paths_in_flight: int = 0, queue_blocks: int = 0, integrator: str = "path", accel: str = "bvh"):
        self._lib = load_library()
        self.items = [(sc, int(w), int(h), int(spp)) for sc, w, h, spp in items]
        self._descs = [sc.desc() for sc, _, _, _ in self.items]  # kept alive for the call
        arr = (_abi.c2_batch_item * len(self.items))()
        for i, ((sc, w, h, spp), d) in enumerate(zip(self.items, self._descs)):
            arr[i].scene = C.cast(C.pointer(d), C.c_void_p)
            arr[i].width, arr[i].height, arr[i].spp = w, h, spp
        o = _abi.c2_render_opts()
        o.max_depth = max_depth
        o.arg_order = _abi.C2_ARGORDER_LTR if arg_order == "ltr" else _abi.C2_ARGORDER_RTL
        o.film_mode = WavefrontPathIntegrator.FILMS[film]
        o.paths_in_flight = paths_in_flight
        o.queue_blocks = queue_blocks
        o.integrator = WavefrontPathIntegrator.INTEGRATORS[integrator]
        o.accel = GpuBVH.ACCELS[accel]
        self._h = C.c_void_p()
        check(self._lib.c2_batch_create(arr, len(self.items), C.byref(o), C.byref(self._h)))

    SCHEDULES = {"wavefront": _abi.C2_SCHED_WAVEFRONT, "persistent": _abi.C2_SCHED_PERSISTENT}

    def autotune(self) -> dict:
        """Tunes both schedules (the wavefront's also over BVH widths 4 and 8) and keeps the
        faster (``c2_batch_autotune``)."""
        t = _abi.c2_batch_tuning()
        check(self._lib.c2_batch_autotune(self._h, C.byref(t)))
        out = {k: getattr(t, k) for k, _ in t._fields_}
        out["schedule"] = {v: k for k, v in self.SCHEDULES.items()}[out["schedule"]]
        return out

    def configure(self, schedule: str = "wavefront", block_threads: int = 256, waves_per_simd: int = 0,
                  paths_in_flight: int = 0, queue_blocks: int = 0, bvh_width: int = 0):
        """Sets the schedule ("wavefront" or "persistent"), launch shape and, for the
        wavefront, the BVH width (4 or 8; 0 keeps it) explicitly."""
        t = _abi.c2_batch_tuning()
        t.schedule = self.SCHEDULES[schedule]
        t.block_threads = block_threads
        t.waves_per_simd = waves_per_simd
        t.paths_in_flight = paths_in_flight
        t.queue_blocks = queue_blocks
        t.bvh_width = bvh_width
        check(self._lib.c2_batch_configure(self._h, C.byref(t)))

    def render(self, use_graph: bool = True):
        check(self._lib.c2_batch_render(self._h, int(use_graph)))

    def sync(self):
        check(self._lib.c2_batch_sync(self._h))

    def film(self, i: int) -> np.ndarray:
        _, w, h, _ = self.items[i]
        out = np.empty((h, w, 3), np.float32)
        check(self._lib.c2_batch_get_film(self._h, i, _fp(out)))
        return out

    def stats(self) -> dict:
        s = _abi.c2_batch_stats()
        check(self._lib.c2_batch_get_stats(self._h, C.byref(s)))
        out = {k: getattr(s, k) for k, _ in s._fields_}
        out["schedule"] = {v: k for k, v in self.SCHEDULES.items()}[out["schedule"]]
        return out

    def close(self):
        if self._h:
            self._lib.c2_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RadianceCache:
    """Online-trained MLP radiance cache (FP16 MFMA forward/backward, FP32 Adam)."""

    PARAMS = _abi.NRC_PARAMS

    def __init__(self, bbox_min, bbox_max, learning_rate=1e-2, batch_capacity=1 << 16, seed=1,
                 beta1=0.9, beta2=0.99, epsilon=1e-8):
        self._lib = load_library()
        o = _abi.c2_nrc_opts()
        o.learning_rate, o.beta1, o.beta2, o.epsilon = learning_rate, beta1, beta2, epsilon
        o.seed, o.batch_capacity = seed, batch_capacity
        o.bbox_min[:] = [float(v) for v in bbox_min]
        o.bbox_max[:] = [float(v) for v in bbox_max]
        self.opts = o
        self._h = C.c_void_p()
        check(self._lib.c2_nrc_create(C.byref(o), C.byref(self._h)))

    def params(self) -> np.ndarray:
        p = np.empty(self.PARAMS, np.float32)
        check(self._lib.c2_nrc_get_params(self._h, _fp(p)))
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        check(self._lib.c2_nrc_set_params(self._h, _fp(p)))

    def infer(self, feats: np.ndarray) -> np.ndarray:
        f = np.ascontiguousarray(feats, np.float32).reshape(-1, 16)
        out = np.empty((len(f), 3), np.float32)
        check(self._lib.c2_nrc_infer(self._h, _fp(f), len(f), _fp(out)))
        return out

    def train_step(self, feats: np.ndarray, targets: np.ndarray, apply: bool = True):
        f = np.ascontiguousarray(feats, np.float32).reshape(-1, 16)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        g = np.empty(self.PARAMS, np.float32)
        loss = C.c_float()
        check(self._lib.c2_nrc_train_step(self._h, _fp(f), _fp(t), len(f), int(apply), _fp(g), C.byref(loss)))
        return g, loss.value

    def stats(self) -> dict:
        st = _abi.c2_nrc_stats()
        check(self._lib.c2_nrc_get_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if self._h:
            self._lib.c2_nrc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def halton_index(width: int, height: int, x: int, y: int, k: int) -> int:
    lib = load_library()
    out = C.c_int32()
    check(lib.c2_halton_index(width, height, x, y, k, C.byref(out)))
    return out.value


def halton_value(width: int, height: int, index: int, dim: int) -> float:
    lib = load_library()
    out = C.c_float()
    check(lib.c2_halton_value(width, height, index, dim, C.byref(out)))
    return out.value


def stream_copy(nbytes: int, reps: int = 3) -> float:
    """Plain device copy over `nbytes` (read once, written once per launch);
    returns the mean launch time in ms. The known-size kernel for HBM counters."""
    lib = load_library()
    out = C.c_double()
    check(lib.c2_util_stream_copy(nbytes, reps, C.byref(out)))
    return out.value
