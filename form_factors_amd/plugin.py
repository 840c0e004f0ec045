"""ctypes view of include/ff_plugin.h: the reference's plugin systems
(ray caster, emission, form factors, thermal equation, thermal solution)
implemented in libff_amd.so. The structs are byte-compatible with the
reference's types, so tests drive the drop-in path exactly as a caller of
the reference would."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import ffk

RAY_CASTER_NAIVE_HIP, RAY_CASTER_GRID_HIP = 2, 4
EMISSION_MALLEY_HIP, EMISSION_PARALLEL_RAYS_HIP = 3, 4
FORM_FACTORS_HIP = 2
THERMAL_EQUATION_SB_FF_HIP, THERMAL_EQUATION_RADIANCE_HIP = 6, 7
THERMAL_EQUATION_CONDUCTIVE_HIP, THERMAL_EQUATION_HEAT_SOURCE_HIP, THERMAL_EQUATION_DISTANT_SOURCE_HIP = 8, 9, 10
THERMAL_SOLUTION_HIP_ADAMS = 2

RAY_CASTER_ERROR = 1
EMISSION_ERROR = 200
FORM_FACTORS_EMPTY_SCENE, FORM_FACTORS_SCENE_TOO_SMALL = 101, 102


class Vec3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Face(C.Structure):
    _fields_ = [("points", Vec3 * 3)]


class Ray(C.Structure):
    _fields_ = [("origin", Vec3), ("direction", Vec3)]


class RcTask(C.Structure):
    _fields_ = [("n_tasks", C.c_int), ("ray", C.POINTER(Ray)), ("hit_face", C.POINTER(C.POINTER(Face))),
                ("hit_point", C.POINTER(Vec3))]


class RcScene(C.Structure):
    _fields_ = [("n_faces", C.c_int), ("faces", C.POINTER(Face))]


class Mesh(C.Structure):
    _fields_ = [("first_idx", C.c_int), ("n_faces", C.c_int), ("padding0", C.c_int)]


class FfScene(C.Structure):
    _fields_ = [("n_faces", C.c_int), ("faces", C.POINTER(Face)), ("n_meshes", C.c_int),
                ("meshes", C.POINTER(Mesh))]


class FfTask(C.Structure):
    _fields_ = [("n_rays", C.c_int), ("form_factors", C.POINTER(C.c_float))]


class Material(C.Structure):
    _fields_ = [("density", C.c_float), ("heat_capacity", C.c_float), ("thermal_conductivity", C.c_float),
                ("thickness", C.c_float), ("front", C.c_float * 5), ("rear", C.c_float * 5),
                ("name", C.c_char * 128)]


class SubjectMesh(C.Structure):
    _fields_ = [("first_idx", C.c_int), ("n_faces", C.c_int), ("material_idx", C.c_int)]


class SubjectScene(C.Structure):
    _fields_ = [("n_faces", C.c_int), ("faces", C.POINTER(Face)), ("n_meshes", C.c_int),
                ("meshes", C.POINTER(SubjectMesh)), ("n_materials", C.c_int), ("materials", C.POINTER(Material))]


class TeTask(C.Structure):
    _fields_ = [("temperatures", C.POINTER(C.c_float)), ("emission", C.POINTER(C.c_float)),
                ("absorption", C.POINTER(C.c_float))]


class HeatSource(C.Structure):
    _fields_ = [("mesh_idx", C.c_int), ("power", C.c_float)]


class Source(C.Structure):
    _fields_ = [("power", C.c_float), ("direction", Vec3)]


class TsTask(C.Structure):
    _fields_ = [("n_step", C.c_int), ("time_delta", C.c_float), ("temperatures", C.POINTER(C.c_float))]


# thermal_equation::system_methods_t, for equations written in Python (tests)
EQ_INIT = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p)
EQ_SHUTDOWN = C.CFUNCTYPE(C.c_int, C.c_void_p)
EQ_SET_SCENE = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(SubjectScene))
EQ_CALCULATE = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(TeTask))


class EqMethods(C.Structure):
    _fields_ = [("init", EQ_INIT), ("shutdown", EQ_SHUTDOWN), ("set_scene", EQ_SET_SCENE),
                ("calculate", EQ_CALCULATE)]


class EqSystem(C.Structure):
    _fields_ = [("methods", C.POINTER(EqMethods))]


_P = C.c_void_p
SIGNATURES = {
    "ffp_ray_caster_create": (_P, [C.c_int]),
    "ffp_ray_caster_set_scene": (C.c_int, [_P, C.POINTER(RcScene)]),
    "ffp_ray_caster_prepare": (C.c_int, [_P]),
    "ffp_ray_caster_cast": (C.c_int, [_P, C.POINTER(RcTask)]),
    "ffp_ray_caster_free": (None, [_P]),
    "ffp_emission_create": (_P, [C.c_int, _P]),
    "ffp_emission_set_scene": (C.c_int, [_P, C.POINTER(RcScene)]),
    "ffp_emission_calculate_malley": (C.c_int, [_P, C.c_int, C.c_float, C.POINTER(C.c_float),
                                                C.POINTER(C.POINTER(RcTask))]),
    "ffp_emission_calculate_parallel": (C.c_int, [_P, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                                  C.c_float, C.c_float, C.POINTER(C.POINTER(RcTask))]),
    "ffp_emission_set_seed": (C.c_int, [_P, C.c_ulonglong, C.c_int]),
    "ffp_emission_free": (None, [_P]),
    "ffp_task_free": (None, [C.POINTER(RcTask)]),
    "ffp_form_factors_create": (_P, [C.c_int, _P]),
    "ffp_form_factors_set_scene": (C.c_int, [_P, C.POINTER(FfScene)]),
    "ffp_form_factors_prepare": (C.c_int, [_P]),
    "ffp_form_factors_calculate": (C.c_int, [_P, C.POINTER(FfTask)]),
    "ffp_form_factors_free": (None, [_P]),
    "ffp_thermal_equation_create": (_P, [C.c_int, _P, C.c_int, C.c_int]),
    "ffp_heat_source_create": (_P, [C.c_int, _P]),
    "ffp_distant_source_create": (_P, [_P, C.c_int, _P]),
    "ffp_thermal_equation_set_scene": (C.c_int, [_P, C.POINTER(SubjectScene)]),
    "ffp_thermal_equation_calculate": (C.c_int, [_P, C.POINTER(TeTask)]),
    "ffp_thermal_equation_free": (None, [_P]),
    "ffp_thermal_solution_create": (_P, [C.c_int, C.c_int, C.POINTER(_P)]),
    "ffp_thermal_solution_set_scene": (C.c_int, [_P, C.POINTER(SubjectScene), C.POINTER(C.c_float)]),
    "ffp_thermal_solution_calculate": (C.c_int, [_P, C.POINTER(TsTask)]),
    "ffp_thermal_solution_set_device_loop": (C.c_int, [_P, C.c_int]),
    "ffp_thermal_solution_free": (None, [_P]),
}

_done = False


def lib():
    global _done
    L = ffk.lib()
    if not _done:
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _done = True
    return L


def faces_array(faces):
    """numpy (n, 9) float32 -> ctypes Face array (kept alive by the caller)."""
    f = np.ascontiguousarray(faces, np.float32).reshape(-1, 9)
    arr = (Face * max(1, len(f)))()
    C.memmove(arr, f.ctypes.data, f.nbytes)
    return arr


def rc_scene(face_arr, n):
    return RcScene(n, C.cast(face_arr, C.POINTER(Face)))


def rays_array(rays):
    r = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    arr = (Ray * max(1, len(r)))()
    C.memmove(arr, r.ctypes.data, r.nbytes)
    return arr


def rc_task(ray_arr, n):
    hit_face = (C.POINTER(Face) * max(1, n))()
    hit_point = (Vec3 * max(1, n))()
    t = RcTask(n, C.cast(ray_arr, C.POINTER(Ray)), C.cast(hit_face, C.POINTER(C.POINTER(Face))),
               C.cast(hit_point, C.POINTER(Vec3)))
    return t, hit_face, hit_point


def hit_indices(task, face_arr):
    """hit_face pointers -> face indices (-1 for NULL)."""
    base = C.addressof(face_arr)
    out = np.empty(task.n_tasks, np.int32)
    for i in range(task.n_tasks):
        p = C.cast(task.hit_face[i], C.c_void_p).value
        out[i] = -1 if not p else (p - base) // C.sizeof(Face)
    return out


def task_rays(task):
    n = task.n_tasks
    buf = (C.c_float * (6 * n)).from_address(C.addressof(task.ray.contents))
    return np.frombuffer(buf, np.float32).reshape(n, 6).copy()


def task_points(task):
    n = task.n_tasks
    buf = (C.c_float * (3 * n)).from_address(C.addressof(task.hit_point.contents))
    return np.frombuffer(buf, np.float32).reshape(n, 3).copy()


def black_body():
    """subject/objects.cpp:177-188: shell 1,1,1,1; both sides absorbance 1, emissivity 1."""
    m = Material()
    m.density = m.heat_capacity = m.thermal_conductivity = m.thickness = 1.0
    for side in (m.front, m.rear):
        side[0], side[1], side[2], side[3], side[4] = 0.0, 0.0, 1.0, 0.0, 1.0
    m.name = b"black_body"
    return m
