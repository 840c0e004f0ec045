"""Writes tests/golden/io.npz: what the reference's own import/export code
(import_export/obj_import.cpp, obj_export.cpp, csv_export.cpp, compiled from
the reference sources into oracle/_ref by `make -C oracle ref`) returns for
the model files under tests/golden/models: scenes, materials, task frames,
re-exported scene text, frame export text and CSV text.

The reference is not run on the inputs where its behaviour is undefined
(face indices outside the vertex list, usemtl before any group); io.npz
records no reference result for those.

Run from the repository root:  python tests/golden/make_io_golden.py
"""
import ctypes as C
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
MODELS = os.path.join(ROOT, "tests", "golden", "models")
OUT = os.path.join(ROOT, "tests", "golden", "io.npz")

OBJ_CASES = ["parallel_planes.obj", "perpendicular_planes.obj", "test_material.obj", "unknown_material.obj",
             "tgrid_10x10.obj", "grid_10x10.obj", "confetti_2k.obj", "patch_sphere_2k.obj",
             "edge/formats.obj", "edge/crlf.obj", "edge/crlf_lib.obj", "edge/nogroup.obj", "edge/bad_vertex.obj", "edge/bad_face.obj",
             "edge/missing_mtl.obj", "edge/bad_value.obj", "edge/bad_newmtl.obj", "edge/trailing_space.obj",
             "edge/no_lib.obj", "edge/bad_usemtl.obj", "edge/usemtl_b.obj"]
# (file, n_meshes, calls)
TASK_CASES = [("parallel_planes.task", 2, 1), ("grid_10x10_heat.task", 100, 1), ("tgrid_10x10_heat.task", 200, 1),
              ("confetti_2k.task", 2000, 1), ("patch_sphere_2k_sun.task", 2000, 1), ("edge/frames.task", 3, 3),
              ("edge/frames.task", 2, 4), ("edge/bad_step.task", 1, 1), ("edge/bad_tmprt.task", 1, 1)]
REEXPORT = ["test_material.obj", "tgrid_10x10.obj", "edge/formats.obj", "edge/nogroup.obj"]


def key(name):
    return name.replace("/", "__").replace(".", "_")


def main():
    so = os.path.join(ROOT, "oracle", "_ref", "libffref.so")
    if not os.path.exists(so):
        sys.exit("build the reference first: make -C oracle ref")
    L = C.CDLL(so)
    P = C.c_void_p
    L.ffref_obj_import.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   C.POINTER(P)]
    L.ffref_obj_copy.argtypes = [P, P, P, P, P]
    L.ffref_obj_free.argtypes = [P]
    L.ffref_task_import.argtypes = [C.c_char_p, C.c_int, C.c_int, P, P, P, P, P, P, C.c_int, P, P]
    L.ffref_obj_reexport.argtypes = [C.c_char_p, C.c_char_p]
    L.ffref_task_export.argtypes = [C.c_char_p, C.c_int, C.c_float, P, C.c_int, C.c_int]
    L.ffref_csv_export.argtypes = [C.c_char_p, P, C.c_int]
    out = {}
    for name in OBJ_CASES:
        nf, nm, nk, h = C.c_int(), C.c_int(), C.c_int(), P()
        r = L.ffref_obj_import(os.path.join(MODELS, name).encode(), C.byref(nf), C.byref(nm), C.byref(nk),
                               C.byref(h))
        k = key(name)
        out[f"obj_{k}_code"] = np.int32(r)
        if r == 0:
            faces = np.zeros((nf.value, 9), np.float32)
            meshes = np.zeros((nm.value, 3), np.int32)
            props = np.zeros((nk.value, 14), np.float32)
            names = np.zeros((nk.value, 128), np.uint8)
            L.ffref_obj_copy(h, faces.ctypes.data, meshes.ctypes.data, props.ctypes.data, names.ctypes.data)
            out[f"obj_{k}_faces"], out[f"obj_{k}_meshes"] = faces, meshes
            out[f"obj_{k}_props"], out[f"obj_{k}_names"] = props, names
        if h.value:
            L.ffref_obj_free(h)
    for name, n, calls in TASK_CASES:
        codes = np.zeros(calls, np.int32)
        dts = np.zeros(calls, np.float32)
        temps = np.zeros((calls, n), np.float32)
        n_src = C.c_int()
        src_mesh = np.zeros(256, np.int32)
        src_power = np.zeros(256, np.float32)
        dst = np.zeros(4, np.float32)
        has = C.c_int()
        r = L.ffref_task_import(os.path.join(MODELS, name).encode(), n, calls, codes.ctypes.data, dts.ctypes.data,
                                temps.ctypes.data, C.byref(n_src), src_mesh.ctypes.data, src_power.ctypes.data, 256,
                                dst.ctypes.data, C.byref(has))
        assert r == 0
        k = f"task_{key(name)}_{n}_{calls}"
        out[k + "_codes"], out[k + "_dts"], out[k + "_temps"] = codes, dts, temps
        out[k + "_src"] = np.stack([src_mesh[:n_src.value].astype(np.float32), src_power[:n_src.value]], 1)
        out[k + "_dst"] = dst if has.value else np.zeros(0, np.float32)
    with tempfile.TemporaryDirectory() as tmp:
        for name in REEXPORT:
            path = os.path.join(tmp, "scene.txt")
            assert L.ffref_obj_reexport(os.path.join(MODELS, name).encode(), path.encode()) == 0
            out[f"reexport_{key(name)}"] = np.frombuffer(open(path, "rb").read(), np.uint8)
        rng = np.random.default_rng(4)
        temps = np.concatenate([[300.0, 0.1, 1e-7, 123456.789, -5.5], rng.uniform(100, 400, 20)]).astype(np.float32)
        out["frame_temps"] = temps
        for binary in (0, 1):
            path = os.path.join(tmp, "frame.txt")
            assert L.ffref_task_export(path.encode(), 7, np.float32(0.01), temps.ctypes.data, len(temps), binary) == 0
            out[f"frame_export_{binary}"] = np.frombuffer(open(path, "rb").read(), np.uint8)
        F = np.array([[0.0, 1.0, 0.5, 1e-9], [0.123456789, 2.5e-5, 7.0, 1e12], [-0.0, 3.4e38, 1.5e-45, 0.333333343],
                      [100.25, 0.1, 0.2, 0.3]], np.float32)
        out["csv_F"] = F
        path = os.path.join(tmp, "f.csv")
        assert L.ffref_csv_export(path.encode(), F.ctypes.data, 4) == 0
        out["csv_text"] = np.frombuffer(open(path, "rb").read(), np.uint8)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}: {len(out)} arrays")


if __name__ == "__main__":
    main()
