"""ctypes bindings of libff_amd.so (include/ffk.h) and a thin host-side API.

The shared library is the product: every compute call below goes to a HIP
kernel on the GPU. There is no CPU fallback; if the library is missing or no
GPU is present, the calls raise.

Device buffers are torch tensors on the current HIP device (torch supplies
memory and the stream; it does no compute here).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libff_amd.so")

CAST_NAIVE, CAST_GRID = 0, 1
FRAME_REFERENCE, FRAME_ORTHONORMAL = 0, 1
F32, BF16, F16 = 0, 1, 2

_vp = C.c_void_p
_i = C.POINTER(C.c_int)
_f = C.POINTER(C.c_float)
_ll = C.c_longlong
_ull = C.c_ulonglong

# name -> (restype, argtypes); every function declared in include/ffk.h
SIGNATURES = {
    "ffk_last_error": (C.c_char_p, []),
    "ffk_device_info": (C.c_int, [_i, _i, _i, _i, _i, C.c_char_p, C.c_int]),
    "ffk_scene_create": (C.c_int, [C.POINTER(_vp)]),
    "ffk_scene_destroy": (None, [_vp]),
    "ffk_scene_set_faces": (C.c_int, [_vp, _f, C.c_int, _vp]),
    "ffk_scene_set_meshes": (C.c_int, [_vp, _i, C.c_int, _vp]),
    "ffk_scene_set_emission": (C.c_int, [_vp, _i, _i, C.c_int, C.POINTER(_ll), _vp]),
    "ffk_scene_build_grid": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp]),
    "ffk_scene_grid_info": (C.c_int, [_vp, _i, _f, _f, C.POINTER(_ll)]),
    "ffk_scene_counts": (C.c_int, [_vp, _i, _i, C.POINTER(_ll)]),
    "ffk_scene_layout": (C.c_int, [_vp, _f, _i, _i, _i, _i]),
    "ffk_scene_grid_cells": (C.c_int, [_vp, _i]),
    "ffk_autotune": (C.c_int, [_vp, _ll, _i, _i, _i, _vp]),
    "ffk_cast": (C.c_int, [_vp, C.c_int, _vp, _ll, _vp, _vp, _vp, _vp]),
    "ffk_emitted_counts": (C.c_int, [_f, C.c_int, C.c_int, C.c_float, _i, _i]),
    "ffk_area_weights": (C.c_int, [_f, C.c_int, _f, _f]),
    "ffk_emit_malley": (C.c_int, [_vp, _ull, _vp, _vp]),
    "ffk_emit_parallel": (C.c_int, [_f, _f, C.c_float, C.c_float, _ll, _ull, _vp, _vp]),
    "ffk_tally_zero": (C.c_int, [_vp, _vp, _vp]),
    "ffk_tally": (C.c_int, [_vp, _vp, _ll, _vp, _vp]),
    "ffk_normalise": (C.c_int, [_vp, _vp, _vp, _vp]),
    "ffk_ff_fused": (C.c_int, [_vp, C.c_int, _ull, _vp, _vp, _vp, _vp]),
    "ffk_set_tuning": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "ffk_ff_needs_tally": (C.c_int, [_vp]),
    "ffk_set_refill": (C.c_int, [_vp, C.c_int, C.c_int]),
    "ffk_get_refill": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "ffk_batch_set_refill": (C.c_int, [_vp, C.c_int, C.c_int]),
    "ffk_set_walk": (C.c_int, [_vp, C.c_int]),
    "ffk_batch_set_walk": (C.c_int, [_vp, C.c_int]),
    "ffk_batch_tally_flags": (C.c_int, [_vp, _i]),
    "ffk_batch_create": (C.c_int, [C.POINTER(_vp), C.POINTER(_vp), C.c_int, _vp]),
    "ffk_batch_destroy": (None, [_vp]),
    "ffk_batch_run": (C.c_int, [_vp, C.POINTER(_ull), C.POINTER(_vp), C.POINTER(_vp), C.c_int, _vp, _vp]),
    "ffk_batch_set_tuning": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "ffk_batch_get_tuning": (C.c_int, [_vp, _i, _i, _i]),
    "ffk_batch_n_scenes": (C.c_int, [_vp]),
    "ffk_batch_set_lds_limit": (C.c_int, [_vp, C.c_int]),
    "ffk_batch_set_concurrent": (C.c_int, [_vp, C.c_int]),
    "ffk_batch_set_order": (C.c_int, [_vp, C.c_int]),
    "ffk_batch_classes": (C.c_int, [_vp, _i, C.POINTER(_ll), _i, C.c_int]),
    "ffk_batch_signature": (C.c_int, [_vp, C.POINTER(_ull)]),
    "ffk_batch_set_class_tuning": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "ffk_batch_get_class_tuning": (C.c_int, [_vp, C.c_int, _i, _i, _i, _i]),
    "ffk_batch_run_class": (C.c_int, [_vp, C.c_int, _ll, _vp]),
    "ffk_batch_autotune": (C.c_int, [_vp, C.POINTER(_ull), C.POINTER(_vp), C.POINTER(_vp), _ll, _vp]),
    "ffk_tune_cache_path": (C.c_int, [C.c_char_p, C.c_int]),
    "ffk_tune_cache_reset": (None, []),
    "ffk_radiance": (C.c_int, [_vp, _ull, _f, C.POINTER(C.c_double), _vp, _vp]),
    "ffk_radiance_step": (C.c_int, [_vp, _f, _f, C.c_int, C.c_int, _ull, _f, C.POINTER(C.c_double), _vp, _vp]),
    "ffk_radiance_weights": (C.c_int, [_f, C.c_int, _i, C.c_int, _f, _f, _f, _f]),
    "ffk_radiance_powers": (C.c_int, [_f, C.c_int, _i, C.c_int, _f, _f, _i, _i, _f, _f]),
    "ffk_bsphere": (C.c_int, [_f, C.c_int, _f, _f]),
    "ffk_sun": (C.c_int, [_vp, _f, _f, C.c_float, C.c_float, C.c_int, _ll, _ull, C.POINTER(_ull), _vp, _vp]),
    "ffk_sun_deposit": (C.c_int, [C.c_int, C.POINTER(_ull), _f, C.c_float, C.POINTER(C.c_double)]),
    "ffk_convert_f": (C.c_int, [_vp, _ll, C.c_int, _vp, _vp]),
    "ffk_pack_exchange": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp]),
    "ffk_sbff": (C.c_int, [C.c_int, _vp, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ffk_sbff_noalloc": (C.c_int, [C.c_int, _vp, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ffk_conduction": (C.c_int, [C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ffk_adams_step": (C.c_int, [C.c_int, C.c_int, C.c_double, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ffk_adams_step_f": (C.c_int, [C.c_int, C.c_int, C.c_double, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ffk_radiance_step_dev": (C.c_int, [_vp, _f, _f, C.c_int, C.c_int, _ull, _vp, _vp, _vp]),
    "ffk_radiance_status": (C.c_int, [_vp]),
    "ffk_sun_count_dev": (C.c_int, [_vp, _f, _f, C.c_float, C.c_float, C.c_int, _ll, _ull, _vp, _vp]),
    "ffk_sun_deposit_dev": (C.c_int, [C.c_int, _vp, _vp, C.c_float, _vp, _vp]),
    "ffk_conduction_dev": (C.c_int, [C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ffk_sbff_dev": (C.c_int, [C.c_int, _vp, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                               _vp]),
    "ffk_heat_sources_dev": (C.c_int, [C.c_int, _vp, _vp, _vp, _vp, _vp]),
    "ffk_step_begin_dev": (C.c_int, [C.c_int, _vp, _vp, _vp, _vp, _vp]),
    "ffk_conduction_graph": (C.c_int, [_f, C.c_int, _i, C.c_int, _f, _f, _i, _i, C.POINTER(C.c_double), C.c_int,
                                       _i]),
    "ffk_thermal_create": (C.c_int, [C.POINTER(_vp), C.c_int, _vp, _vp, _vp, _vp]),
    "ffk_thermal_destroy": (None, [_vp]),
    "ffk_thermal_set_exchange": (C.c_int, [_vp, _vp, C.c_int, _vp]),
    "ffk_thermal_set_tuning": (C.c_int, [_vp, C.c_int, C.c_int]),
    "ffk_thermal_set_conduction": (C.c_int, [_vp, _i, _i, C.POINTER(C.c_double)]),
    "ffk_thermal_set_sources": (C.c_int, [_vp, C.c_int, _i, C.POINTER(C.c_double)]),
    "ffk_thermal_run": (C.c_int, [_vp, _vp, C.c_int, C.c_double, _vp, _vp, C.c_int, _vp]),
}

_lib = None


class FFKError(RuntimeError):
    pass


def lib():
    """Load libff_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FFKError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        # torch ships its own HIP runtime under the same soname as the one
        # this library links: load torch's first, so that one runtime (and
        # one device context) serves both in the process.
        import torch  # noqa: F401

        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc < 0:
        msg = lib().ffk_last_error().decode(errors="replace")
        raise FFKError(f"{what} failed ({rc}): {msg}")
    return rc


def _np(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _ptr(t):
    """Device pointer of a torch tensor (or None)."""
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream(stream=None):
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def require_gpu():
    import torch

    if not torch.cuda.is_available():
        raise FFKError("no HIP device: the form-factor kernels run on the GPU only")


# ------------------------------------------------------------ host helpers --

def area_weights(faces):
    """Reference two-sided area weights and float total (host arithmetic)."""
    faces = _np(faces, np.float32).reshape(-1, 9)
    w = np.zeros(2 * len(faces), np.float32)
    total = C.c_float()
    check(lib().ffk_area_weights(faces.ctypes.data_as(_f), len(faces), w.ctypes.data_as(_f),
                                 C.byref(total)), "ffk_area_weights")
    return w, np.float32(total.value)


def emitted_counts(weights, n_rays, total):
    w = _np(weights, np.float32)
    nf = len(w) // 2
    fr = np.zeros(nf, np.int32)
    rr = np.zeros(nf, np.int32)
    check(lib().ffk_emitted_counts(w.ctypes.data_as(_f), nf, int(n_rays), float(total),
                                   fr.ctypes.data_as(_i), rr.ctypes.data_as(_i)), "ffk_emitted_counts")
    return fr, rr


def bsphere(faces):
    """math::triangles_bsphere (host): (centre float32[3], radius float32)."""
    faces = _np(faces, np.float32).reshape(-1, 9)
    c = np.zeros(3, np.float32)
    r = C.c_float()
    check(lib().ffk_bsphere(faces.ctypes.data_as(_f), len(faces), c.ctypes.data_as(_f), C.byref(r)), "ffk_bsphere")
    return c, np.float32(r.value)


def radiance_weights(faces, meshes, eps2, T):
    """Reference radiance weights (2 per face) and float total (host)."""
    faces = _np(faces, np.float32).reshape(-1, 9)
    meshes = _np(meshes, np.int32).reshape(-1, 3)
    w = np.zeros(2 * len(faces), np.float32)
    total = C.c_float()
    check(lib().ffk_radiance_weights(faces.ctypes.data_as(_f), len(faces), meshes.ctypes.data_as(_i), len(meshes),
                                     _np(eps2, np.float32).ctypes.data_as(_f), _np(T, np.float32).ctypes.data_as(_f),
                                     w.ctypes.data_as(_f), C.byref(total)), "ffk_radiance_weights")
    return w, np.float32(total.value)


def radiance_powers(faces, meshes, eps2, T, front, rear):
    """Per-ray powers (2 per face) and per-mesh emission (host, float)."""
    faces = _np(faces, np.float32).reshape(-1, 9)
    meshes = _np(meshes, np.int32).reshape(-1, 3)
    p = np.zeros(2 * len(faces), np.float32)
    em = np.zeros(len(meshes), np.float32)
    check(lib().ffk_radiance_powers(faces.ctypes.data_as(_f), len(faces), meshes.ctypes.data_as(_i), len(meshes),
                                    _np(eps2, np.float32).ctypes.data_as(_f), _np(T, np.float32).ctypes.data_as(_f),
                                    _np(front, np.int32).ctypes.data_as(_i), _np(rear, np.int32).ctypes.data_as(_i),
                                    p.ctypes.data_as(_f), em.ctypes.data_as(_f)), "ffk_radiance_powers")
    return p, em


def device_info():
    n_cu, clk, mclk, bus, l2 = (C.c_int() for _ in range(5))
    name = C.create_string_buffer(256)
    check(lib().ffk_device_info(C.byref(n_cu), C.byref(clk), C.byref(mclk), C.byref(bus),
                                C.byref(l2), name, 256), "ffk_device_info")
    return {"n_cu": n_cu.value, "clock_khz": clk.value, "mem_clock_khz": mclk.value,
            "bus_width_bits": bus.value, "l2_bytes": l2.value, "name": name.value.decode()}


# ------------------------------------------------------------------ scene --

class Scene:
    """A device-resident scene (ffk_scene) with its emission layout and grid."""

    def __init__(self, faces, meshes=None):
        require_gpu()
        self.faces = _np(faces, np.float32).reshape(-1, 9)
        if meshes is None:  # one mesh per face
            n = len(self.faces)
            meshes = np.stack([np.arange(n), np.ones(n), np.zeros(n)], 1)
        self.meshes = _np(meshes, np.int32).reshape(-1, 3)
        h = C.c_void_p()
        check(lib().ffk_scene_create(C.byref(h)), "ffk_scene_create")
        self.h = h
        check(lib().ffk_scene_set_faces(self.h, self.faces.ctypes.data_as(_f), len(self.faces),
                                        _stream()), "ffk_scene_set_faces")
        check(lib().ffk_scene_set_meshes(self.h, self.meshes.ctypes.data_as(_i), len(self.meshes),
                                         _stream()), "ffk_scene_set_meshes")
        self._front = self._rear = None
        self._n_rays, self._frame = 0, FRAME_REFERENCE
        self._stale = False  # the library set the layout (radiance_step): read it back on access

    def _layout(self):
        if self._stale:
            self._sync_layout()
        return self

    @property
    def front(self):
        return self._layout()._front

    @property
    def rear(self):
        return self._layout()._rear

    @property
    def n_rays(self):
        return self._layout()._n_rays

    @property
    def frame(self):
        return self._layout()._frame

    @property
    def n_faces(self):
        return len(self.faces)

    @property
    def n_meshes(self):
        return len(self.meshes)

    def set_emission(self, n_rays=None, front=None, rear=None, weights=None, total=None,
                     frame=FRAME_REFERENCE):
        """Emission layout from explicit counts, or from n_rays and weights
        (default: the reference's two-sided area weights)."""
        if front is None:
            if weights is None:
                weights, total = area_weights(self.faces)
            front, rear = emitted_counts(weights, n_rays, total)
        fr, rr = _np(front, np.int32), _np(rear, np.int32)
        n = C.c_longlong()
        check(lib().ffk_scene_set_emission(self.h, fr.ctypes.data_as(_i), rr.ctypes.data_as(_i), frame, C.byref(n),
                                           _stream()), "ffk_scene_set_emission")
        self._front, self._rear, self._n_rays, self._frame, self._stale = fr, rr, n.value, frame, False
        return self._n_rays

    def _sync_layout(self):
        """Reads back the emission layout a library call set (front and rear
        counts, total rays, frame)."""
        fr = np.zeros(self.n_faces, np.int32)
        rr = np.zeros(self.n_faces, np.int32)
        frame = C.c_int()
        check(lib().ffk_scene_layout(self.h, None, None, fr.ctypes.data_as(_i), rr.ctypes.data_as(_i),
                                     C.byref(frame)), "ffk_scene_layout")
        n = C.c_longlong()
        check(lib().ffk_scene_counts(self.h, None, None, C.byref(n)), "ffk_scene_counts")
        self._front, self._rear, self._n_rays, self._frame, self._stale = fr, rr, n.value, frame.value, False

    def sun_counts(self, direction, center, distance, radius, n_rays, seed, aperture=0, stats=None):
        """Hits per mesh side of n_rays parallel rays: uint64[n_meshes, 2] (front, rear)."""
        counts = np.zeros(2 * self.n_meshes, np.uint64)
        check(lib().ffk_sun(self.h, _np(direction, np.float32).ctypes.data_as(_f),
                            _np(center, np.float32).ctypes.data_as(_f), float(distance), float(radius), int(aperture),
                            int(n_rays), seed, counts.ctypes.data_as(C.POINTER(_ull)), _ptr(stats), _stream()),
              "ffk_sun")
        return counts.reshape(-1, 2)

    def sun(self, absorb2, power, direction, n_rays, seed, aperture=0, stats=None):
        """One parallel_rays_cpu::calculate with the reference's aperture
        placement: absorbed power per mesh (float64, exact sum)."""
        c, r0 = bsphere(self.faces)
        radius = np.float32(r0 * np.float32(1.1))
        distance = np.float32(np.float32(1.1) + radius)
        counts = self.sun_counts(direction, c, distance, radius, n_rays, seed, aperture, stats)
        ray_power = np.float32(np.float32(power) * (np.float32(np.pi) * radius * radius) / np.float32(n_rays))
        out = np.zeros(self.n_meshes, np.float64)
        check(lib().ffk_sun_deposit(self.n_meshes, np.ascontiguousarray(counts.reshape(-1)).ctypes.data_as(
            C.POINTER(_ull)), _np(absorb2, np.float32).ctypes.data_as(_f), float(ray_power),
            out.ctypes.data_as(C.POINTER(C.c_double))), "ffk_sun_deposit")
        return out

    def radiance_step(self, eps2, T, n_rays, seed, frame=FRAME_REFERENCE, stats=None):
        """One radiance_equation::calculate (ffk_radiance_step): layout from
        the radiance weights, then the device pass. Returns (emission float32,
        absorption float64)."""
        e2 = _np(eps2, np.float32).reshape(-1)
        t = _np(T, np.float32).reshape(-1)
        assert len(e2) == 2 * self.n_meshes and len(t) == self.n_meshes
        em = np.zeros(self.n_meshes, np.float32)
        ab = np.zeros(self.n_meshes, np.float64)
        check(lib().ffk_radiance_step(self.h, e2.ctypes.data_as(_f), t.ctypes.data_as(_f), int(n_rays), int(frame),
                                      seed, em.ctypes.data_as(_f), ab.ctypes.data_as(C.POINTER(C.c_double)),
                                      _ptr(stats), _stream()), "ffk_radiance_step")
        self._stale = True
        return em, ab

    def radiance(self, seed, ray_power, stats=None):
        """Absorbed power per mesh (float64) for the current emission layout;
        ray_power: float32[2 * n_faces], power of one front / rear ray."""
        rp = _np(ray_power, np.float32)
        assert len(rp) == 2 * self.n_faces
        out = np.zeros(self.n_meshes, np.float64)
        check(lib().ffk_radiance(self.h, seed, rp.ctypes.data_as(_f), out.ctypes.data_as(C.POINTER(C.c_double)),
                                 _ptr(stats), _stream()), "ffk_radiance")
        return out

    def rays_per_mesh(self):
        """Rays emitted by each mesh (front + rear over its faces): the row
        divisor of F (cpuFactorsCalculator/cpu_system.cpp:193-195)."""
        per_face = self.front.astype(np.int64) + self.rear.astype(np.int64)
        first, count = self.meshes[:, 0], self.meshes[:, 1]
        csum = np.concatenate([[0], np.cumsum(per_face)])
        return csum[first + count] - csum[first]

    def build_grid(self, dims=(0, 0, 0)):
        check(lib().ffk_scene_build_grid(self.h, int(dims[0]), int(dims[1]), int(dims[2]), _stream()),
              "ffk_scene_build_grid")
        return self.grid_info()

    def grid_cells(self):
        """The grid's cell words as int32[nz, ny, nx, 2] (ffk_scene_grid_cells)."""
        dims = self.grid_info()["dims"]
        out = np.zeros((dims[2], dims[1], dims[0], 2), np.int32)
        check(lib().ffk_scene_grid_cells(self.h, out.ctypes.data_as(_i)), "ffk_scene_grid_cells")
        return out

    def grid_info(self):
        dims = (C.c_int * 3)()
        lo = (C.c_float * 3)()
        hi = (C.c_float * 3)()
        nr = C.c_longlong()
        check(lib().ffk_scene_grid_info(self.h, dims, lo, hi, C.byref(nr)), "ffk_scene_grid_info")
        return {"dims": tuple(dims), "lo": tuple(lo), "hi": tuple(hi), "n_refs": nr.value}

    def autotune(self, sample_rays=1 << 22):
        """Grid resolution, block size, walk mode and ray scheduling (refill
        threshold, pool ratio) from timed trials (cached per scene signature).
        Returns {"dims", "block", "walk", "refill", "pool", "cached"}."""
        dims = (C.c_int * 3)()
        block = C.c_int()
        walk = C.c_int()
        rc = check(lib().ffk_autotune(self.h, int(sample_rays), dims, C.byref(block), C.byref(walk), _stream()),
                   "ffk_autotune")
        refill, pool = self.get_refill()
        return {"dims": tuple(dims), "block": block.value, "walk": walk.value, "refill": refill, "pool": pool,
                "cached": rc == 1}

    def set_tuning(self, block=0, rays_per_tile=0, hist=-1):
        check(lib().ffk_set_tuning(self.h, block, rays_per_tile, hist), "ffk_set_tuning")

    def set_refill(self, lanes, pool=2):
        """Fused-kernel ray scheduling: refill threshold (idle lanes of 64,
        0 = one ray per lane) and pooled-test ratio (0 = per-lane tests)."""
        check(lib().ffk_set_refill(self.h, int(lanes), int(pool)), "ffk_set_refill")

    def get_refill(self):
        """(refill threshold, pooled-test ratio) currently set on the scene."""
        lanes, pool = C.c_int(), C.c_int()
        check(lib().ffk_get_refill(self.h, C.byref(lanes), C.byref(pool)), "ffk_get_refill")
        return lanes.value, pool.value

    def set_walk(self, mode):
        """Walk mode of the pooled schedule (0..3, default 3): bit 0 = always the
        jump code for empty cells (else a per-wavefront choice), bit 1 = pooled
        tests dealt in whole chunks of 64."""
        check(lib().ffk_set_walk(self.h, int(mode)), "ffk_set_walk")

    # -- device calls (torch tensors in, torch tensors out) --
    def cast(self, rays, method=CAST_GRID, points=True, stats=None):
        import torch

        n = rays.shape[0]
        idx = torch.empty(n, dtype=torch.int32, device=rays.device)
        pts = torch.zeros((n, 3), dtype=torch.float32, device=rays.device) if points else None
        check(lib().ffk_cast(self.h, method, _ptr(rays), n, _ptr(idx), _ptr(pts), _ptr(stats), _stream()),
              "ffk_cast")
        return idx, pts

    def emit_malley(self, seed):
        import torch

        rays = torch.empty((self.n_rays, 6), dtype=torch.float32, device="cuda")
        check(lib().ffk_emit_malley(self.h, seed, _ptr(rays), _stream()), "ffk_emit_malley")
        return rays

    def tally(self, hit_idx, tally=None):
        import torch

        N = self.n_meshes
        if tally is None:
            tally = torch.empty((N, N), dtype=torch.int64, device="cuda")
            check(lib().ffk_tally_zero(self.h, _ptr(tally), _stream()), "ffk_tally_zero")
        check(lib().ffk_tally(self.h, _ptr(hit_idx), hit_idx.shape[0], _ptr(tally), _stream()), "ffk_tally")
        return tally

    def normalise(self, tally, F=None):
        import torch

        N = self.n_meshes
        if F is None:
            F = torch.empty((N, N), dtype=torch.float32, device="cuda")
        check(lib().ffk_normalise(self.h, _ptr(tally), _ptr(F), _stream()), "ffk_normalise")
        return F

    def ff_fused(self, seed, F=None, tally=None, stats=None, need_tally=None):
        """Fused emit -> cast -> tally -> normalise. Allocates the uint64
        tally only when some row spans several tiles."""
        import torch

        N = self.n_meshes
        if F is None:
            F = torch.empty((N, N), dtype=torch.float32, device="cuda")
        if need_tally is None:
            need_tally = bool(lib().ffk_ff_needs_tally(self.h))
        if tally is None and need_tally:
            tally = torch.empty((N, N), dtype=torch.int64, device="cuda")
        check(lib().ffk_ff_fused(self.h, CAST_GRID, seed, _ptr(tally), _ptr(F), _ptr(stats), _stream()),
              "ffk_ff_fused")
        return F

    def close(self):
        if getattr(self, "h", None):
            lib().ffk_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def emit_parallel(direction, center, distance, radius, n_rays, seed):
    import torch

    rays = torch.empty((n_rays, 6), dtype=torch.float32, device="cuda")
    d = _np(direction, np.float32)
    c = _np(center, np.float32)
    check(lib().ffk_emit_parallel(d.ctypes.data_as(_f), c.ctypes.data_as(_f), float(distance),
                                  float(radius), n_rays, seed, _ptr(rays), _stream()), "ffk_emit_parallel")
    return rays


class Batch:
    """Many scenes through one persistent (optionally graph-replayed) launch."""

    def __init__(self, scenes):
        self.scenes = list(scenes)
        arr = (C.c_void_p * len(self.scenes))(*[s.h.value for s in self.scenes])
        h = C.c_void_p()
        check(lib().ffk_batch_create(C.byref(h), arr, len(self.scenes), _stream()), "ffk_batch_create")
        self.h = h

    def set_tuning(self, block=0, rays_per_tile=0, n_blocks=0):
        check(lib().ffk_batch_set_tuning(self.h, block, rays_per_tile, n_blocks), "ffk_batch_set_tuning")

    def tuning(self):
        """(block, rays_per_tile, n_blocks) of the batch (0 = automatic)."""
        v = [C.c_int() for _ in range(3)]
        check(lib().ffk_batch_get_tuning(self.h, *(C.byref(x) for x in v)), "ffk_batch_get_tuning")
        return tuple(x.value for x in v)

    def set_refill(self, lanes, pool=2):
        check(lib().ffk_batch_set_refill(self.h, int(lanes), int(pool)), "ffk_batch_set_refill")

    def set_concurrent(self, on):
        """Class launches on forked streams (True) or one after another (False)."""
        check(lib().ffk_batch_set_concurrent(self.h, 1 if on else 0), "ffk_batch_set_concurrent")

    def set_order(self, order):
        """Tile queue: 0 longest first across scenes, 1 scene by scene."""
        check(lib().ffk_batch_set_order(self.h, int(order)), "ffk_batch_set_order")

    def set_lds_limit(self, nbytes):
        """Scenes whose row histogram needs more LDS use the global tally (0 = default)."""
        check(lib().ffk_batch_set_lds_limit(self.h, int(nbytes)), "ffk_batch_set_lds_limit")

    def set_walk(self, mode):
        check(lib().ffk_batch_set_walk(self.h, int(mode)), "ffk_batch_set_walk")

    def tally_flags(self):
        """Scenes that need a uint64 tally buffer in run()."""
        flags = np.zeros(len(self.scenes), np.int32)
        check(lib().ffk_batch_tally_flags(self.h, flags.ctypes.data_as(_i)), "ffk_batch_tally_flags")
        return flags.astype(bool)

    def make_tallies(self):
        """Tally buffers for the scenes that need one (None elsewhere)."""
        import torch

        return [torch.empty((sc.n_meshes,) * 2, dtype=torch.int64, device="cuda") if need else None
                for sc, need in zip(self.scenes, self.tally_flags())]

    def classes(self):
        """[(class id, rays, tiles)] in launch order."""
        n = lib().ffk_batch_classes(self.h, None, None, None, 0)
        ids, tiles = np.zeros(n, np.int32), np.zeros(n, np.int32)
        rays = np.zeros(n, np.int64)
        check(lib().ffk_batch_classes(self.h, ids.ctypes.data_as(_i), rays.ctypes.data_as(C.POINTER(_ll)),
                                      tiles.ctypes.data_as(_i), n), "ffk_batch_classes")
        return [(int(a), int(b), int(c)) for a, b, c in zip(ids, rays, tiles)]

    def class_tuning(self, class_id):
        v = [C.c_int() for _ in range(4)]
        check(lib().ffk_batch_get_class_tuning(self.h, int(class_id), *[C.byref(x) for x in v]),
              "ffk_batch_get_class_tuning")
        return dict(zip(("block", "refill", "pool", "walk"), (x.value for x in v)))

    def set_class_tuning(self, class_id, block=-1, refill=-1, pool=-1, walk=-1):
        check(lib().ffk_batch_set_class_tuning(self.h, int(class_id), int(block), int(refill), int(pool), int(walk)),
              "ffk_batch_set_class_tuning")

    def autotune(self, seeds, Fs, tallies=None, sample_rays=1 << 28):
        """Per-class tuning (ffk_batch_autotune); returns {class id: settings}
        and whether they came from a cache."""
        n = len(self.scenes)
        sd = (C.c_ulonglong * n)(*[int(x) for x in seeds])
        fp = (C.c_void_p * n)(*[F.data_ptr() for F in Fs])
        tp = None if tallies is None else (C.c_void_p * n)(*[(t.data_ptr() if t is not None else 0)
                                                            for t in tallies])
        rc = check(lib().ffk_batch_autotune(self.h, sd, tp, fp, int(sample_rays), _stream()), "ffk_batch_autotune")
        return {cid: self.class_tuning(cid) for cid, _, _ in self.classes()}, rc == 1

    def run(self, seeds, Fs, tallies=None, use_graph=True, stats=None):
        n = len(self.scenes)
        sd = (C.c_ulonglong * n)(*[int(x) for x in seeds])
        fp = (C.c_void_p * n)(*[F.data_ptr() for F in Fs])
        tp = None
        if tallies is not None:
            tp = (C.c_void_p * n)(*[(t.data_ptr() if t is not None else 0) for t in tallies])
        check(lib().ffk_batch_run(self.h, sd, tp, fp, 1 if use_graph else 0, _ptr(stats), _stream()),
              "ffk_batch_run")

    def close(self):
        if getattr(self, "h", None):
            lib().ffk_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------- thermal --

def conduction_graph(faces, meshes, conductivity, thickness):
    """Symmetric CSR (row_ptr, cols, G) of mesh-to-mesh conductances
    (host arithmetic of the reference's conduction formula)."""
    faces = _np(faces, np.float32).reshape(-1, 9)
    meshes = _np(meshes, np.int32).reshape(-1, 3)
    k = _np(conductivity, np.float32)
    t = _np(thickness, np.float32)
    nnz = C.c_int()
    L = lib()
    check(L.ffk_conduction_graph(faces.ctypes.data_as(_f), len(faces), meshes.ctypes.data_as(_i), len(meshes),
                                 k.ctypes.data_as(_f), t.ctypes.data_as(_f), None, None, None, 0, C.byref(nnz)),
          "ffk_conduction_graph")
    row = np.zeros(len(meshes) + 1, np.int32)
    cols = np.zeros(max(1, nnz.value), np.int32)
    G = np.zeros(max(1, nnz.value), np.float64)
    check(L.ffk_conduction_graph(faces.ctypes.data_as(_f), len(faces), meshes.ctypes.data_as(_i), len(meshes),
                                 k.ctypes.data_as(_f), t.ctypes.data_as(_f), row.ctypes.data_as(_i),
                                 cols.ctypes.data_as(_i), G.ctypes.data_as(C.POINTER(C.c_double)), len(G),
                                 C.byref(nnz)), "ffk_conduction_graph")
    return row, cols[:nnz.value], G[:nnz.value]


def convert_f(F, dtype):
    """F (float32 device tensor) in the storage type (F32, BF16 or F16)."""
    import torch

    out = torch.empty(F.shape, dtype={F32: torch.float32, BF16: torch.bfloat16, F16: torch.float16}[dtype],
                      device=F.device)
    check(lib().ffk_convert_f(_ptr(F), F.numel(), dtype, _ptr(out), _stream()), "ffk_convert_f")
    return out


def pack_exchange(F, dtype, row_mul=None):
    """Exchange storage of F (N x N float32 device tensor): returns
    (St, row_scale) with St[j, m] = dtype(F[m, j] * mul_m) (transposed) and
    row_scale = 1 / mul (see ffk_pack_exchange; row_mul None = detect
    count-valued rows)."""
    import torch

    n = F.shape[0]
    assert F.shape == (n, n) and F.dtype == torch.float32 and F.is_contiguous()
    out = torch.empty(F.shape, dtype={F32: torch.float32, BF16: torch.bfloat16, F16: torch.float16}[dtype],
                      device=F.device)
    scale = torch.empty(n, dtype=torch.float64, device=F.device)
    mul = None if row_mul is None else torch.as_tensor(np.asarray(row_mul, np.float64), device=F.device)
    check(lib().ffk_pack_exchange(_ptr(F), n, dtype, _ptr(mul), _ptr(out), _ptr(scale), _stream()),
          "ffk_pack_exchange")
    return out, scale


class Thermal:
    """Device-resident transient run: SB_FF exchange + conduction + sources
    with the FP64 Adams update."""

    def __init__(self, C_, area, eps_front, eps_rear):
        require_gpu()
        self.n = len(C_)
        self._keep = [np.ascontiguousarray(x, np.float64) for x in (C_, area, eps_front, eps_rear)]
        h = C.c_void_p()
        check(lib().ffk_thermal_create(C.byref(h), self.n, *[x.ctypes.data_as(_vp) for x in self._keep]),
              "ffk_thermal_create")
        self.h = h
        self.F = None

    def set_exchange(self, F_dev, dtype, row_scale=None):
        self.F = F_dev  # keep the device buffer alive
        check(lib().ffk_thermal_set_exchange(self.h, _ptr(F_dev), dtype, _ptr(row_scale)),
              "ffk_thermal_set_exchange")

    def set_tuning(self, unroll=0, max_blocks=0):
        check(lib().ffk_thermal_set_tuning(self.h, unroll, max_blocks), "ffk_thermal_set_tuning")

    def set_conduction(self, row, cols, G):
        self._csr = (np.ascontiguousarray(row, np.int32), np.ascontiguousarray(cols, np.int32),
                     np.ascontiguousarray(G, np.float64))
        r, c, g = self._csr
        check(lib().ffk_thermal_set_conduction(self.h, r.ctypes.data_as(_i), c.ctypes.data_as(_i),
                                               g.ctypes.data_as(C.POINTER(C.c_double))), "ffk_thermal_set_conduction")

    def set_sources(self, mesh, power):
        m = np.ascontiguousarray(mesh, np.int32)
        p = np.ascontiguousarray(power, np.float64)
        check(lib().ffk_thermal_set_sources(self.h, len(m), m.ctypes.data_as(_i),
                                            p.ctypes.data_as(C.POINTER(C.c_double))), "ffk_thermal_set_sources")

    def run(self, n_steps, dt, T0=None, balance=False, use_graph=True):
        T_out = np.zeros(self.n, np.float64)
        bal = np.zeros(2 * max(1, n_steps), np.float64) if balance else None
        t0 = None if T0 is None else np.ascontiguousarray(T0, np.float64)
        check(lib().ffk_thermal_run(self.h, None if t0 is None else t0.ctypes.data_as(_vp), n_steps, dt,
                                    T_out.ctypes.data_as(_vp), None if bal is None else bal.ctypes.data_as(_vp),
                                    1 if use_graph else 0, _stream()), "ffk_thermal_run")
        return (T_out, bal.reshape(-1, 2)[:n_steps]) if balance else T_out

    def close(self):
        if getattr(self, "h", None):
            lib().ffk_thermal_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tune_cache_path():
    """The autotune cache file ('' when disabled)."""
    buf = C.create_string_buffer(4096)
    lib().ffk_tune_cache_path(buf, 4096)
    return buf.value.decode()
