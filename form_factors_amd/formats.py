"""Scene, material and frame files through libff_amd.so (include/ff_io.h):
Wavefront OBJ scenes with thermal MTL materials, .task frame files, the
OBJ-like scene and frame export, and the CSV export of F
(import_export/obj_import.cpp, obj_export.cpp, csv_export.cpp)."""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

from . import plugin as P

FILE_ERROR = 21
FORMAT_ERROR = 22
MATERIAL_NOT_DEFINED = 23
MATERIAL_NOT_ENOUGH_PARAMETERS = 24
MATERIAL_INVALID_PARAMETER_VALUE = 25
CSV_FILE_ERROR = 31

# shell 4, front 5, rear 5: the column order of `materials` below
PROPERTIES = ("shell.density", "shell.heat_capacity", "shell.thermal_conductivity", "shell.thickness",
              "front.specular_reflectance", "front.diffuse_reflectance", "front.absorbance",
              "front.transmittance", "front.emissivity", "rear.specular_reflectance", "rear.diffuse_reflectance",
              "rear.absorbance", "rear.transmittance", "rear.emissivity")


class FormatError(RuntimeError):
    """A negative OBJ_IMPORT_* / CSV_EXPORT_* code; .code holds it."""

    def __init__(self, code, what):
        super().__init__(f"{what} failed with code {code}")
        self.code = code


class HeatSourceParams(C.Structure):
    _fields_ = [("n_sources", C.c_int), ("sources", C.POINTER(P.HeatSource))]


_SP = C.POINTER(P.SubjectScene)
SIGNATURES = {
    "ffio_obj_import": (C.c_int, [C.c_char_p, C.POINTER(_SP)]),
    "ffio_obj_import_ex": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(_SP)]),
    "ffio_scene_free": (None, [_SP]),
    "ffio_task_import": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(P.TsTask), C.POINTER(HeatSourceParams),
                                   C.POINTER(C.POINTER(P.Source))]),
    "ffio_obj_export_scene": (C.c_int, [C.c_void_p, _SP]),
    "ffio_obj_export_task": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(P.TsTask)]),
    "ffio_obj_export_task_binary": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(P.TsTask)]),
    "ffio_csv_export": (C.c_int, [C.c_char_p, C.POINTER(P.FfScene), C.POINTER(P.FfTask)]),
}

_done = False
_libc = None


def lib():
    global _done
    L = P.lib()
    if not _done:
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _done = True
    return L


def libc():
    global _libc
    if _libc is None:
        _libc = C.CDLL(None)
        _libc.fopen.restype = C.c_void_p
        _libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
        _libc.fdopen.restype = C.c_void_p
        _libc.fdopen.argtypes = [C.c_int, C.c_char_p]
        _libc.fclose.argtypes = [C.c_void_p]
        _libc.fflush.argtypes = [C.c_void_p]
        _libc.free.argtypes = [C.c_void_p]
    return _libc


class _File:
    """A C FILE* for the duration of a with-block."""

    def __init__(self, path, mode):
        self.path, self.mode = os.fsencode(path), mode.encode()

    def __enter__(self):
        self.f = libc().fopen(self.path, self.mode)
        if not self.f:
            raise FormatError(-FILE_ERROR, f"fopen({self.path!r})")
        return self.f

    def __exit__(self, *exc):
        libc().fclose(self.f)


INLINE_MATERIALS = 1


def obj_import(path, inline_materials=False):
    """obj_import::scene. Returns {"faces": (n, 9) float32, "meshes": (m, 3)
    int32 (first face, face count, material), "materials": (k, 14) float32
    in PROPERTIES order, "names": [k material names]}. inline_materials also
    reads material blocks written inside the OBJ file (as obj_export::scene
    writes them)."""
    sp = _SP()
    r = lib().ffio_obj_import_ex(os.fsencode(path), INLINE_MATERIALS if inline_materials else 0, C.byref(sp))
    if r != 0:
        raise FormatError(r, f"ffio_obj_import({path})")
    try:
        s = sp.contents
        faces = np.ctypeslib.as_array(C.cast(s.faces, C.POINTER(C.c_float)), (s.n_faces * 9,)).reshape(-1, 9).copy() \
            if s.n_faces else np.zeros((0, 9), np.float32)
        meshes = np.ctypeslib.as_array(C.cast(s.meshes, C.POINTER(C.c_int)), (s.n_meshes * 3,)).reshape(-1, 3).copy()
        mats = np.empty((s.n_materials, 14), np.float32)
        names = []
        for k in range(s.n_materials):
            m = s.materials[k]
            mats[k, :4] = (m.density, m.heat_capacity, m.thermal_conductivity, m.thickness)
            mats[k, 4:9] = list(m.front)
            mats[k, 9:] = list(m.rear)
            names.append(m.name.decode("latin-1"))
        return {"faces": faces, "meshes": meshes, "materials": mats, "names": names}
    finally:
        lib().ffio_scene_free(sp)


def task_import(path, n_meshes, n_calls=1, step=0.0):
    """obj_import::task called n_calls times on one open file (the step is
    `step` until a "step" line sets it). Returns {"frames": [(count, step,
    temperatures[:count])], "heat_sources": [(mesh, power)],
    "distant_source": (power, x, y, z) or None}."""
    t = P.TsTask(0, float(step), None)
    hs = HeatSourceParams(0, None)
    ds = C.POINTER(P.Source)()
    frames = []
    try:
        with _File(path, "r") as f:
            for _ in range(n_calls):
                r = lib().ffio_task_import(f, n_meshes, C.byref(t), C.byref(hs), C.byref(ds))
                if r < 0:
                    raise FormatError(r, f"ffio_task_import({path})")
                temps = np.array([t.temperatures[i] for i in range(r)], np.float32)
                frames.append((r, t.time_delta, temps))
        sources = [(hs.sources[i].mesh_idx, hs.sources[i].power) for i in range(hs.n_sources)]
        dst = None
        if ds:
            d = ds.contents
            dst = (d.power, d.direction.x, d.direction.y, d.direction.z)
        return {"frames": frames, "heat_sources": sources, "distant_source": dst}
    finally:
        libc().free(C.cast(t.temperatures, C.c_void_p))
        libc().free(C.cast(hs.sources, C.c_void_p))
        libc().free(C.cast(ds, C.c_void_p))


def _subject_scene(faces, meshes, materials, names):
    faces = np.ascontiguousarray(faces, np.float32).reshape(-1, 9)
    meshes = np.ascontiguousarray(meshes, np.int32).reshape(-1, 3)
    materials = np.ascontiguousarray(materials, np.float32).reshape(-1, 14)
    fa = P.faces_array(faces)
    ma = (P.SubjectMesh * max(1, len(meshes)))()
    for i, (a, b, c) in enumerate(meshes):
        ma[i] = P.SubjectMesh(int(a), int(b), int(c))
    mt = (P.Material * max(1, len(materials)))()
    for k in range(len(materials)):
        v = materials[k]
        mt[k].density, mt[k].heat_capacity, mt[k].thermal_conductivity, mt[k].thickness = map(float, v[:4])
        for j in range(5):
            mt[k].front[j] = float(v[4 + j])
            mt[k].rear[j] = float(v[9 + j])
        mt[k].name = names[k].encode("latin-1")[:127]
    s = P.SubjectScene(len(faces), C.cast(fa, C.POINTER(P.Face)), len(meshes), C.cast(ma, C.POINTER(P.SubjectMesh)),
                       len(materials), C.cast(mt, C.POINTER(P.Material)))
    return s, (fa, ma, mt)


def obj_export_scene(path, faces, meshes, materials, names):
    """obj_export::scene into `path`."""
    s, keep = _subject_scene(faces, meshes, materials, names)
    with _File(path, "w") as f:
        r = lib().ffio_obj_export_scene(f, C.byref(s))
    if r != 0:
        raise FormatError(r, "ffio_obj_export_scene")


def task_export(path, frames, binary=False):
    """obj_export::task (or task_binary) for each (n_step, step, temperatures)
    of `frames`, appended to `path` in order."""
    fn = lib().ffio_obj_export_task_binary if binary else lib().ffio_obj_export_task
    with _File(path, "w") as f:
        for n_step, dt, temps in frames:
            T = np.ascontiguousarray(temps, np.float32)
            t = P.TsTask(int(n_step), float(dt), T.ctypes.data_as(C.POINTER(C.c_float)))
            r = fn(f, len(T), C.byref(t))
            if r != 0:
                raise FormatError(r, "ffio_obj_export_task")


class FrameWriter:
    """Frames written one at a time as obj_export::task (or task_binary)
    writes them, to a file or, for path "-", to standard output (the
    controller's result argument)."""

    def __init__(self, path, binary=False):
        self.fn = lib().ffio_obj_export_task_binary if binary else lib().ffio_obj_export_task
        if path == "-":
            sys.stdout.flush()
            self.f = libc().fdopen(os.dup(1), b"w")
        else:
            self.f = libc().fopen(os.fsencode(path), b"w")
        if not self.f:
            raise FormatError(-FILE_ERROR, f"cannot open {path!r} for the frames")

    def write(self, n_step, dt, temps):
        T = np.ascontiguousarray(temps, np.float32)
        t = P.TsTask(int(n_step), float(dt), T.ctypes.data_as(C.POINTER(C.c_float)))
        r = self.fn(self.f, len(T), C.byref(t))
        if r != 0:
            raise FormatError(r, "ffio_obj_export_task")
        libc().fflush(self.f)

    def close(self):
        if self.f:
            libc().fclose(self.f)
            self.f = None


def csv_export(path, F):
    """csv_export::export_csv of an N x N form-factor matrix."""
    F = np.ascontiguousarray(F, np.float32)
    n = F.shape[0]
    scene = P.FfScene(0, None, n, None)
    task = P.FfTask(0, F.ctypes.data_as(C.POINTER(C.c_float)))
    r = lib().ffio_csv_export(os.fsencode(path), C.byref(scene), C.byref(task))
    if r != 0:
        raise FormatError(r, f"ffio_csv_export({path})")
