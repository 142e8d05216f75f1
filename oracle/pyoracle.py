"""ctypes binding of the CPU oracle (c2_oracle.h) -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker; never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBS = {"portable": os.path.join(HERE, "_build", "libc2oracle.so"),
        "libm": os.path.join(HERE, "_build", "libc2oracle_libm.so")}
REF_BIN = os.path.join(HERE, "_ref", "crystal2_ref")          # clang build: arguments left to right
REF_BIN_GCC = os.path.join(HERE, "_ref", "crystal2_ref_gcc")  # GCC build: arguments right to left

ACCEL_BVH, ACCEL_BRUTE, ACCEL_KDTREE = 0, 1, 2
TIE_FIRST, TIE_LOWEST = 0, 1
INTEGRATORS = {"path": 0, "direct": 1}  # PathTracingIntegrator, DirectLightingIntegrator

_loaded = {}


def build():
    import subprocess
    subprocess.run(["make", "-C", HERE, "oracle"], check=True, stdout=subprocess.DEVNULL)


def lib(variant: str = "portable"):
    if variant in _loaded:
        return _loaded[variant]
    path = LIBS[variant]
    if not os.path.exists(path):
        build()
    from crystal2_amd import _abi
    L = C.CDLL(path)
    L.orc_scene_create.argtypes = [C.POINTER(_abi.c2_scene_desc), C.c_int, C.c_int, C.POINTER(C.c_void_p)]
    L.orc_scene_destroy.argtypes = [C.c_void_p]
    L.orc_scene_set_integrator.argtypes = [C.c_void_p, C.c_int]
    L.orc_scene_set_reference_draws.argtypes = [C.c_void_p, C.c_int, C.c_uint64]
    L.orc_last_error.restype = C.c_char_p
    L.orc_render.argtypes = [C.c_void_p] + [C.c_int] * 8 + [C.POINTER(C.c_float), C.POINTER(C.c_double)]
    L.orc_render_samples.argtypes = [C.c_void_p] + [C.c_int] * 6 + [C.POINTER(C.c_float)]
    L.orc_intersect.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_int64, C.POINTER(C.c_int32),
                                C.POINTER(C.c_float)]
    L.orc_intersect_test.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_int32), C.c_int64,
                                     C.POINTER(C.c_uint8)]
    L.orc_halton.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int32), C.c_int64, C.c_int, C.POINTER(C.c_int32),
                             C.POINTER(C.c_float)]
    _loaded[variant] = L
    return L


class OracleScene:
    def __init__(self, scene, accel: int = ACCEL_BVH, tie: int = TIE_FIRST, variant: str = "portable",
                 integrator: str = "path", reference_draws: bool = False, skip: int = 0):
        """reference_draws: Halton dims >= 100 and Blinn-Phong redraws from the
        reference's own generators in one-thread render order (render with
        threads=1 only); the default is the hash streams the GPU uses. skip:
        draws of both generators discarded first (another point of the streams)."""
        self._L = lib(variant)
        self._scene = scene  # keeps the product-side arrays alive
        d = scene.desc()
        self._h = C.c_void_p()
        rc = self._L.orc_scene_create(C.byref(d), accel, tie, C.byref(self._h))
        if rc != 0:
            raise RuntimeError(self._L.orc_last_error().decode())
        if self._L.orc_scene_set_integrator(self._h, INTEGRATORS[integrator]) != 0:
            raise RuntimeError(self._L.orc_last_error().decode())
        self.reference_draws = bool(reference_draws)
        self._L.orc_scene_set_reference_draws(self._h, int(self.reference_draws), int(skip))

    def render(self, width, height, spp, depth, arg_order="rtl", threads=None, y_range=None):
        threads = 1 if self.reference_draws else (threads or min(16, os.cpu_count() or 1))
        y0, y1 = y_range if y_range else (0, height)
        film = np.zeros((height, width, 3), np.float32)
        secs = C.c_double()
        rc = self._L.orc_render(self._h, width, height, spp, depth, 1 if arg_order == "ltr" else 0, threads, y0, y1,
                                film.ctypes.data_as(C.POINTER(C.c_float)), C.byref(secs))
        if rc != 0:
            raise RuntimeError(self._L.orc_last_error().decode())
        return film, secs.value

    def render_samples(self, width, height, spp, depth, arg_order="rtl", threads=None):
        threads = threads or min(16, os.cpu_count() or 1)
        out = np.zeros((height, width, spp, 3), np.float32)
        rc = self._L.orc_render_samples(self._h, width, height, spp, depth, 1 if arg_order == "ltr" else 0, threads,
                                        out.ctypes.data_as(C.POINTER(C.c_float)))
        if rc != 0:
            raise RuntimeError(self._L.orc_last_error().decode())
        return out

    def intersect(self, rays):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        prim = np.empty(len(r), np.int32)
        t = np.empty(len(r), np.float32)
        self._L.orc_intersect(self._h, r.ctypes.data_as(C.POINTER(C.c_float)), len(r),
                              prim.ctypes.data_as(C.POINTER(C.c_int32)), t.ctypes.data_as(C.POINTER(C.c_float)))
        return prim, t

    def intersect_test(self, rays, ignore=None):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hit = np.empty(len(r), np.uint8)
        ig = None if ignore is None else np.ascontiguousarray(ignore, np.int32)
        self._L.orc_intersect_test(self._h, r.ctypes.data_as(C.POINTER(C.c_float)),
                                   None if ig is None else ig.ctypes.data_as(C.POINTER(C.c_int32)), len(r),
                                   hit.ctypes.data_as(C.POINTER(C.c_uint8)))
        return hit.astype(bool)

    def __del__(self):
        try:
            self._L.orc_scene_destroy(self._h)
        except Exception:
            pass


def halton(width, height, queries, ndims, variant="portable"):
    L = lib(variant)
    q = np.ascontiguousarray(queries, np.int32).reshape(-1, 3)
    idx = np.empty(len(q), np.int32)
    vals = np.empty((len(q), ndims), np.float32)
    L.orc_halton(width, height, q.ctypes.data_as(C.POINTER(C.c_int32)), len(q), ndims,
                 idx.ctypes.data_as(C.POINTER(C.c_int32)), vals.ctypes.data_as(C.POINTER(C.c_float)))
    return idx, vals
