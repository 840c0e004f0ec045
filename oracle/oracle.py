"""ctypes bindings for the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads oracle/liboracle.so (this project's restatement of the reference CPU
path) and, when it has been built, oracle/_ref/libffref.so (the reference's
own CPU sources compiled in place). Only tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg import this module; the product package never
does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# FF_ORACLE_SO names another build of the same sources (the sanitizer build
# of tests/test_sanitizers.py); default: the one make -C oracle writes here.
ORACLE_SO = os.environ.get("FF_ORACLE_SO") or os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libffref.so")

RNG_MT19937, RNG_PHILOX = 0, 1
FRAME_REFERENCE, FRAME_ORTHONORMAL = 0, 1

_f = C.POINTER(C.c_float)
_i = C.POINTER(C.c_int)
_d = C.POINTER(C.c_double)
_u32 = C.POINTER(C.c_uint32)


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _sig(lib, name, res, args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = args
    return fn


_lib = None
_ref = None


def build():
    """Compile liboracle.so (and the reference driver if the reference tree exists)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "ref", "-j8"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} is not built (make -C oracle)")
        L = C.CDLL(ORACLE_SO)
        _sig(L, "or_triangle_intersect", C.c_int, [_f, _f, _f])
        _sig(L, "or_triangle_area", C.c_float, [_f])
        _sig(L, "or_rotate_towards", None, [_f, _f, _f])
        _sig(L, "or_mat_vec", None, [_f, _f, _f])
        _sig(L, "or_philox4x32_10", None, [_u32, _u32, _u32])
        _sig(L, "or_sincos_turns", None, [C.c_float, _f, _f])
        _sig(L, "or_naive_cast", None, [_f, C.c_int, _f, C.c_long, _i, _f, C.c_int])
        _sig(L, "or_zgrid_build", C.c_void_p, [_f, C.c_int])
        _sig(L, "or_zgrid_free", None, [C.c_void_p])
        _sig(L, "or_zgrid_info", None, [C.c_void_p, _i, _i, _f, _f, _f, _i])
        _sig(L, "or_zgrid_cells", None, [C.c_void_p, _i, _i, _i])
        _sig(L, "or_zgrid_cast", None, [C.c_void_p, _f, C.c_long, _i, _f, C.c_int])
        _sig(L, "or_grid_traverse", C.c_int, [_f, _f, C.c_int, C.c_int, _f, C.c_int, _i, C.c_int])
        _sig(L, "or_grid_rasterize", C.c_int, [_f, _f, C.c_int, C.c_int, _f, _i, C.c_int])
        _sig(L, "or_ray_intersect_segment", C.c_int, [_f, _f, _f])
        _sig(L, "or_emitted_counts", None, [_f, C.c_int, C.c_int, C.c_float, _i, _i])
        _sig(L, "or_area_weights", C.c_float, [_f, C.c_int, _f])
        _sig(L, "or_malley_count", C.c_long, [_f, C.c_int, C.c_int, C.c_float])
        _sig(L, "or_malley_emit", C.c_long,
             [_f, C.c_int, _f, C.c_float, C.c_int, C.c_int, C.c_uint64, C.c_int, _f])
        _sig(L, "or_parallel_emit", None,
             [_f, _f, C.c_float, C.c_float, C.c_int, C.c_int, C.c_uint64, _f])
        _sig(L, "or_ff_tally", C.c_int, [_i, C.c_int, C.c_int, _i, _i, C.c_long, _i, _f])
        _sig(L, "or_ff_calculate", C.c_int,
             [_f, C.c_int, _i, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_int, C.c_int, C.c_int, _f])
        _sig(L, "or_parallel_emit2", None,
             [_f, _f, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int, C.c_uint64, _f])
        _sig(L, "or_bsphere", None, [_f, C.c_int, _f, _f])
        _sig(L, "or_sun", None,
             [_f, C.c_int, _i, C.c_int, _f, C.c_float, _f, C.c_int, C.c_int, C.c_uint64, C.c_int, C.c_int,
              C.c_int, _f, _d])
        _sig(L, "or_radiance_weights", C.c_float, [_f, C.c_int, _i, C.c_int, _f, _f, _f])
        _sig(L, "or_radiance_walk", None,
             [_f, C.c_int, _i, C.c_int, _f, _f, _i, _i, C.c_long, _i, _f, _f, _d])
        _sig(L, "or_radiance", C.c_int,
             [_f, C.c_int, _i, C.c_int, _f, _f, C.c_int, C.c_int, C.c_uint64, C.c_int, C.c_int, C.c_int,
              _f, _f, _d])
        _sig(L, "or_sbff_f64", None, [C.c_int, _d, _d, _d, _d, _d, _d, _d])
        _sig(L, "or_adams_combine_f64", None, [C.c_int, C.c_int, C.c_double, _d, _d, _d])
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    """The compiled reference driver (only where oracle/_ref was built)."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"{REF_SO} is not built (make -C oracle ref)")
        R = C.CDLL(REF_SO)
        _sig(R, "ffref_triangle_intersect", C.c_int, [_f, _f, _f])
        _sig(R, "ffref_triangle_area", C.c_float, [_f])
        _sig(R, "ffref_rotate_towards", None, [_f, _f, _f])
        _sig(R, "ffref_mat_vec", None, [_f, _f, _f])
        _sig(R, "ffref_axis_rotation", None, [C.c_float, C.c_float, C.c_float, _f])
        _sig(R, "ffref_ray_intersect_segment", C.c_int, [_f, _f, _f])
        _sig(R, "ffref_grid_traverse", C.c_int, [_f, _f, C.c_int, C.c_int, _f, C.c_int, _i, C.c_int])
        _sig(R, "ffref_grid_rasterize", C.c_int, [_f, _f, C.c_int, C.c_int, _f, _i, C.c_int])
        _sig(R, "ffref_zgrid_info", C.c_int,
             [_f, C.c_int, _i, _i, _f, _f, _f, _i, _i, _i, C.c_int])
        _sig(R, "ffref_cast", C.c_int, [C.c_int, _f, C.c_int, _f, C.c_long, _i, _f, C.c_int])
        _sig(R, "ffref_emitted_counts", None, [_f, C.c_int, C.c_int, C.c_float, _i, _i])
        _sig(R, "ffref_malley", C.c_long, [_f, C.c_int, _f, C.c_float, C.c_int, _f, _i, C.c_long])
        _sig(R, "ffref_parallel_rays", None,
             [_f, _f, C.c_float, C.c_float, C.c_int, _f, C.c_int, _f])
        _sig(R, "ffref_ff_calculate", C.c_int, [_f, C.c_int, _i, C.c_int, C.c_int, C.c_int, _f])
        _sig(R, "ffref_ff_prepare_code", C.c_int, [_f, C.c_int, _i, C.c_int])
        _sig(R, "ffref_adams", C.c_int,
             [_f, C.c_int, _i, C.c_int, _f, _f, _f, _f, C.c_float, C.c_int, _f])
        _sig(R, "ffref_radiance", C.c_int, [_f, C.c_int, _i, C.c_int, _f, _f, C.c_int, _f, _f])
        _sig(R, "ffref_bsphere", None, [_f, C.c_int, _f, _f])
        _sig(R, "ffref_sun", C.c_int, [_f, C.c_int, _i, C.c_int, _f, C.c_float, _f, C.c_int, _f])
        _sig(R, "ffref_heat_source", C.c_int, [_i, _f, C.c_int, _f, _f])
        _sig(R, "ffref_ff_timed", C.c_int, [_f, C.c_int, _i, C.c_int, C.c_int, C.c_int, _f, _d])
        _sig(R, "ffref_perftest_scene", None, [C.c_int, C.c_float, C.c_int, _f, _f])
        _ref = R
    return _ref


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---------------------------------------------------------------- wrappers --

def naive_cast(faces, rays, n_threads=1):
    faces, rays = f32(faces).reshape(-1, 9), f32(rays).reshape(-1, 6)
    n = len(rays)
    idx = np.full(n, -1, np.int32)
    pts = np.zeros((n, 3), np.float32)
    lib().or_naive_cast(_p(faces, _f), len(faces), _p(rays, _f), n, _p(idx, _i), _p(pts, _f), n_threads)
    return idx, pts


class ZGrid:
    def __init__(self, faces):
        self.faces = f32(faces).reshape(-1, 9)
        self.h = lib().or_zgrid_build(_p(self.faces, _f), len(self.faces))

    def info(self):
        nx, ny, nr = C.c_int(), C.c_int(), C.c_int()
        base, side, size = (np.zeros(3, np.float32) for _ in range(3))
        lib().or_zgrid_info(self.h, C.byref(nx), C.byref(ny), _p(base, _f), _p(side, _f),
                            _p(size, _f), C.byref(nr))
        return nx.value, ny.value, base, side, size, nr.value

    def cells(self):
        nx, ny, _, _, _, nr = self.info()
        counts = np.zeros(nx * ny, np.int32)
        offsets = np.zeros(nx * ny, np.int32)
        refs = np.zeros(max(nr, 1), np.int32)
        lib().or_zgrid_cells(self.h, _p(counts, _i), _p(offsets, _i), _p(refs, _i))
        return counts, offsets, refs[:nr]

    def cast(self, rays, n_threads=1):
        rays = f32(rays).reshape(-1, 6)
        n = len(rays)
        idx = np.full(n, -1, np.int32)
        pts = np.zeros((n, 3), np.float32)
        lib().or_zgrid_cast(self.h, _p(rays, _f), n, _p(idx, _i), _p(pts, _f), n_threads)
        return idx, pts

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_zgrid_free(self.h)
            self.h = None


def zgrid_cast(faces, rays, n_threads=1):
    return ZGrid(faces).cast(rays, n_threads)


def triangle_intersect(ray6, tri9):
    pt = np.zeros(3, np.float32)
    code = lib().or_triangle_intersect(_p(f32(ray6), _f), _p(f32(tri9), _f), _p(pt, _f))
    return code, pt


def rotate_towards(frm, to):
    m = np.zeros(9, np.float32)
    lib().or_rotate_towards(_p(f32(frm), _f), _p(f32(to), _f), _p(m, _f))
    return m.reshape(3, 3)


def mat_vec(m, v):
    out = np.zeros(3, np.float32)
    lib().or_mat_vec(_p(f32(m).reshape(9), _f), _p(f32(v), _f), _p(out, _f))
    return out


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib().or_philox4x32_10(_p(c, _u32), _p(k, _u32), _p(o, _u32))
    return o


def sincos_turns(u):
    s, c = C.c_float(), C.c_float()
    lib().or_sincos_turns(float(u), C.byref(s), C.byref(c))
    return s.value, c.value


def area_weights(faces):
    faces = f32(faces).reshape(-1, 9)
    w = np.zeros(2 * len(faces), np.float32)
    total = lib().or_area_weights(_p(faces, _f), len(faces), _p(w, _f))
    return w, np.float32(total)


def emitted_counts(weights, n_rays, total):
    w = f32(weights)
    nf = len(w) // 2
    fr = np.zeros(nf, np.int32)
    rr = np.zeros(nf, np.int32)
    lib().or_emitted_counts(_p(w, _f), nf, int(n_rays), float(total), _p(fr, _i), _p(rr, _i))
    return fr, rr


def malley_emit(faces, weights, total, n_rays, rng=RNG_PHILOX, seed=0, frame=FRAME_REFERENCE):
    faces, w = f32(faces).reshape(-1, 9), f32(weights)
    n = lib().or_malley_count(_p(w, _f), len(faces), int(n_rays), float(total))
    rays = np.zeros((max(n, 1), 6), np.float32)
    lib().or_malley_emit(_p(faces, _f), len(faces), _p(w, _f), float(total), int(n_rays), rng,
                         seed, frame, _p(rays, _f))
    return rays[:n]


def parallel_emit(direction, center, distance, radius, n_rays, rng=RNG_PHILOX, seed=0):
    rays = np.zeros((n_rays, 6), np.float32)
    lib().or_parallel_emit(_p(f32(direction), _f), _p(f32(center), _f), float(distance),
                           float(radius), int(n_rays), rng, seed, _p(rays, _f))
    return rays


def ff_tally(meshes, n_faces, front, rear, hit_idx):
    meshes = i32(meshes).reshape(-1, 3)
    nm = len(meshes)
    F = np.zeros((nm, nm), np.float32)
    lib().or_ff_tally(_p(meshes, _i), nm, int(n_faces), _p(i32(front), _i), _p(i32(rear), _i),
                      len(hit_idx), _p(i32(hit_idx), _i), _p(F, _f))
    return F


def ff_calculate(faces, meshes, n_rays, rng=RNG_PHILOX, seed=0, frame=FRAME_REFERENCE,
                 caster=0, n_threads=1):
    faces, meshes = f32(faces).reshape(-1, 9), i32(meshes).reshape(-1, 3)
    nm = len(meshes)
    F = np.zeros((nm, nm), np.float32)
    r = lib().or_ff_calculate(_p(faces, _f), len(faces), _p(meshes, _i), nm, int(n_rays), rng,
                              seed, frame, caster, n_threads, _p(F, _f))
    return r, F


def sbff_f64(F, T, area, eps_front, eps_rear):
    n = len(T)
    F = np.ascontiguousarray(F, np.float64)
    args = [np.ascontiguousarray(x, np.float64) for x in (T, area, eps_front, eps_rear)]
    em = np.zeros(n)
    ab = np.zeros(n)
    lib().or_sbff_f64(n, _p(F, _d), *[_p(a, _d) for a in args], _p(em, _d), _p(ab, _d))
    return em, ab


def parallel_emit2(direction, center, distance, radius, aperture, n_rays, rng=RNG_PHILOX, seed=0):
    """Parallel rays from a disk (aperture 0) or a square of half side radius (1)."""
    rays = np.zeros((max(n_rays, 1), 6), np.float32)
    lib().or_parallel_emit2(_p(f32(direction), _f), _p(f32(center), _f), float(distance), float(radius),
                            int(aperture), int(n_rays), rng, seed, _p(rays, _f))
    return rays[:n_rays]


def bsphere(faces):
    """math::triangles_bsphere: (centre, radius)."""
    faces = f32(faces).reshape(-1, 9)
    c = np.zeros(3, np.float32)
    r = np.zeros(1, np.float32)
    lib().or_bsphere(_p(faces, _f), len(faces), _p(c, _f), _p(r, _f))
    return c, np.float32(r[0])


def sun(faces, meshes, absorb2, power, direction, n_rays, rng=RNG_PHILOX, seed=0, aperture=0, caster=0,
        n_threads=1):
    """parallel_rays_cpu::calculate: (absorption float ray order, absorption double)."""
    faces, meshes = f32(faces).reshape(-1, 9), i32(meshes).reshape(-1, 3)
    nm = len(meshes)
    ab = np.zeros(nm, np.float32)
    ab64 = np.zeros(nm, np.float64)
    lib().or_sun(_p(faces, _f), len(faces), _p(meshes, _i), nm, _p(f32(absorb2), _f), float(power),
                 _p(f32(direction), _f), int(n_rays), rng, seed, int(aperture), caster, n_threads, _p(ab, _f),
                 _p(ab64, _d))
    return ab, ab64


def radiance_weights(faces, meshes, eps2, T):
    """radiance_cpu.cpp:96-122 weights (2 per face) and their float total."""
    faces, meshes = f32(faces).reshape(-1, 9), i32(meshes).reshape(-1, 3)
    w = np.zeros(2 * len(faces), np.float32)
    total = lib().or_radiance_weights(_p(faces, _f), len(faces), _p(meshes, _i), len(meshes),
                                      _p(f32(eps2), _f), _p(f32(T), _f), _p(w, _f))
    return w, np.float32(total)


def radiance_walk(faces, meshes, eps2, T, front, rear, hit_idx):
    """radiance_cpu.cpp:132-190 over given hits: (emission, absorption in
    float ray order, absorption in double)."""
    faces, meshes = f32(faces).reshape(-1, 9), i32(meshes).reshape(-1, 3)
    nm = len(meshes)
    em = np.zeros(nm, np.float32)
    ab = np.zeros(nm, np.float32)
    ab64 = np.zeros(nm, np.float64)
    hit = i32(hit_idx)
    lib().or_radiance_walk(_p(faces, _f), len(faces), _p(meshes, _i), nm, _p(f32(eps2), _f), _p(f32(T), _f),
                           _p(i32(front), _i), _p(i32(rear), _i), len(hit), _p(hit, _i), _p(em, _f), _p(ab, _f),
                           _p(ab64, _d))
    return em, ab, ab64


def radiance(faces, meshes, eps2, T, n_rays, rng=RNG_PHILOX, seed=0, frame=FRAME_REFERENCE, caster=0,
             n_threads=1):
    """The whole radiance_equation::calculate (radiance_cpu.cpp:132-190)."""
    faces, meshes = f32(faces).reshape(-1, 9), i32(meshes).reshape(-1, 3)
    nm = len(meshes)
    em = np.zeros(nm, np.float32)
    ab = np.zeros(nm, np.float32)
    ab64 = np.zeros(nm, np.float64)
    lib().or_radiance(_p(faces, _f), len(faces), _p(meshes, _i), nm, _p(f32(eps2), _f), _p(f32(T), _f),
                      int(n_rays), rng, seed, frame, caster, n_threads, _p(em, _f), _p(ab, _f), _p(ab64, _d))
    return em, ab, ab64


def adams_combine_f64(n_step, dt, T_n, E_hist):
    T_n = np.ascontiguousarray(T_n, np.float64)
    E_hist = np.ascontiguousarray(E_hist, np.float64)
    out = np.zeros_like(T_n)
    lib().or_adams_combine_f64(len(T_n), n_step, dt, _p(T_n, _d), _p(E_hist, _d), _p(out, _d))
    return out


# -------------------------------------------------- compiled reference ----

def ref_cast(kind, faces, rays, n_threads=1):
    faces, rays = f32(faces).reshape(-1, 9), f32(rays).reshape(-1, 6)
    n = len(rays)
    idx = np.full(n, -1, np.int32)
    pts = np.zeros((n, 3), np.float32)
    rc = ref().ffref_cast(kind, _p(faces, _f), len(faces), _p(rays, _f), n, _p(idx, _i), _p(pts, _f),
                          n_threads)
    if rc < 0:
        raise RuntimeError(f"reference caster returned {rc}")
    return idx, pts


def ref_ff_timed(faces, meshes, n_rays, caster=1):
    """One reference FORM_FACTORS_CPU calculate() (single-threaded, mt19937
    Malley emission, caster 0 = naive_cpu, 1 = zgrid_cpu): F and its time
    split (ray generation, casting, tally + normalisation) in seconds."""
    faces, meshes = f32(faces).reshape(-1, 9), i32(meshes).reshape(-1, 3)
    nm = len(meshes)
    F = np.zeros((nm, nm), np.float32)
    t = np.zeros(3, np.float64)
    rc = ref().ffref_ff_timed(_p(faces, _f), len(faces), _p(meshes, _i), nm, int(n_rays), caster, _p(F, _f),
                              _p(t, _d))
    if rc < 0:
        raise RuntimeError(f"reference form-factor calculator returned {rc}")
    return F, t


def ref_perftest_scene(n_faces=38000, radius=20.0, n_rays=None):
    """perfTest's confetti scene and collapsing rays from the reference's own
    spherical generator: (faces float32[n, 9], rays float32[m, 6])."""
    n_rays = 100 * n_faces if n_rays is None else n_rays
    faces = np.zeros((n_faces, 9), np.float32)
    rays = np.zeros((n_rays, 6), np.float32)
    ref().ffref_perftest_scene(int(n_faces), float(radius), int(n_rays), _p(faces, _f), _p(rays, _f))
    return faces, rays
