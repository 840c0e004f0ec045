"""The reference controller's thermal run (controller/controller.cpp:40-242)
over the HIP plugin systems: an OBJ scene and a .task file in, frames of
mesh temperatures out. Radiance (Malley rays over the grid caster, fused on
the device), conduction, heat sources and the distant source are summed by
the FP64 Adams solver, one step per frame, exactly in the controller's
order; every call goes through the C ABI of include/ff_plugin.h and
include/ff_io.h."""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import formats
from . import plugin as P


def run(scene_path, task_path, n_rays=1_000_000, n_steps=100, result_path=None, binary=False, seed=0,
        inline_materials=True, distant_rays=10_000, scene=None):
    """controller.cpp:40-242. Returns {"frames": [(n_step, step, T)],
    "seconds": wall time of set-up + steps (the controller's timer),
    "n_meshes", "n_faces"}. result_path: the frames are also written there
    as the controller writes them (obj_export::task / task_binary). scene:
    an already imported scene (formats.obj_import's dict) instead of
    scene_path."""
    L = P.lib()
    formats.lib()
    caster = L.ffp_ray_caster_create(P.RAY_CASTER_GRID_HIP)
    interior = L.ffp_emission_create(P.EMISSION_MALLEY_HIP, caster)
    exterior = L.ffp_emission_create(P.EMISSION_PARALLEL_RAYS_HIP, caster)
    if not (caster and interior and exterior):
        raise RuntimeError("failed to create the caster or the emitters")
    L.ffp_emission_set_seed(interior, seed, 0)
    L.ffp_emission_set_seed(exterior, seed + 1, 0)
    systems = []
    try:
        sc = scene if scene is not None else formats.obj_import(scene_path, inline_materials=inline_materials)
        scene, keep = formats._subject_scene(sc["faces"], sc["meshes"], sc["materials"], sc["names"])
        n = len(sc["meshes"])
        task_in = formats.task_import(task_path, n, 1, step=0.1)  # controller.cpp:129-146
        count, dt, temps = task_in["frames"][0]
        T = np.full(n, 300.0, np.float32)  # controller.cpp:148-154: missing temperatures are 300 K
        T[:count] = temps[:count]
        Tc = (C.c_float * n)(*T.tolist())
        task = P.TsTask(0, float(dt), C.cast(Tc, C.POINTER(C.c_float)))
        srcs = task_in["heat_sources"]
        hs = (P.HeatSource * max(1, len(srcs)))(*[P.HeatSource(int(m), float(p)) for m, p in srcs])
        dsrc = None
        if task_in["distant_source"] is not None:
            pw, x, y, z = task_in["distant_source"]
            dsrc = P.Source(pw, P.Vec3(x, y, z))
        eqs = [L.ffp_thermal_equation_create(P.THERMAL_EQUATION_RADIANCE_HIP, interior, int(n_rays), 0),
               L.ffp_thermal_equation_create(P.THERMAL_EQUATION_CONDUCTIVE_HIP, None, 0, 0),
               L.ffp_heat_source_create(len(srcs), C.cast(hs, C.c_void_p)),
               L.ffp_distant_source_create(exterior, int(distant_rays),
                                           C.cast(C.pointer(dsrc), C.c_void_p) if dsrc is not None else None)]
        systems += [("eq", e) for e in eqs if e]
        if not all(eqs):
            raise RuntimeError("failed to create a thermal equation")
        arr = (C.c_void_p * 4)(*eqs)
        sol = L.ffp_thermal_solution_create(P.THERMAL_SOLUTION_HIP_ADAMS, 4, arr)
        if not sol:
            raise RuntimeError("failed to create the thermal solution")
        systems.append(("sol", sol))
        frames = [(task.n_step, task.time_delta, np.frombuffer(Tc, np.float32).copy())]
        t0 = time.perf_counter()
        r = L.ffp_thermal_solution_set_scene(sol, C.byref(scene), Tc)
        if r < 0:
            raise RuntimeError(f"thermal solution set_scene failed ({r})")
        for step in range(n_steps):
            r = L.ffp_thermal_solution_calculate(sol, C.byref(task))
            if r < 0:
                raise RuntimeError(f"thermal solution step {step} failed ({r})")
            frames.append((task.n_step, task.time_delta, np.frombuffer(Tc, np.float32).copy()))
        seconds = time.perf_counter() - t0
        if result_path is not None:
            formats.task_export(result_path, frames, binary=binary)
        return {"frames": frames, "seconds": seconds, "n_meshes": n, "n_faces": len(sc["faces"])}
    finally:
        for kind, s in reversed(systems):
            (L.ffp_thermal_solution_free if kind == "sol" else L.ffp_thermal_equation_free)(s)
        L.ffp_emission_free(exterior)
        L.ffp_emission_free(interior)
        L.ffp_ray_caster_free(caster)
