"""The reference controller's thermal run (controller/controller.cpp:35-242)
over the HIP plugin systems: an OBJ scene and a .task file in, frames of
mesh temperatures out. Radiance (Malley rays over the grid caster, fused on
the device), conduction, heat sources and the distant source are summed by
the FP64 Adams solver, one step per frame, exactly in the controller's
order; every call goes through the C ABI of include/ff_plugin.h and
include/ff_io.h. Over the grid caster each step is device-resident: the
equations' emission and absorption stay on the device up to the FP64
update, and one copy brings the new temperatures back for the frame.

Command line (the controller's positional arguments, controller.cpp:35-96):

    python -m form_factors_amd.controller <scene.obj> <task> [result|-] [rays]
        [steps] [cpu|cuda|naive_cpu|naive_cuda] [binary 0|1] [--inline-materials]

The caster names select the HIP casters in the slots of the reference's:
cpu and cuda the grid caster, naive_cpu and naive_cuda the brute force.
Frames go to the result file (standard output for "-") as each step ends;
the run ends with "#Done in <seconds>s." as the controller prints it.
--inline-materials (not in the reference) also reads material blocks
written inside the OBJ file."""
from __future__ import annotations

import ctypes as C
import os
import sys
import time

import numpy as np

from . import formats
from . import plugin as P

CASTERS = {"cpu": P.RAY_CASTER_GRID_HIP, "cuda": P.RAY_CASTER_GRID_HIP,
           "naive_cpu": P.RAY_CASTER_NAIVE_HIP, "naive_cuda": P.RAY_CASTER_NAIVE_HIP}
RAY_CASTER_NOT_SUPPORTED = 2


def run(scene_path, task_path, n_rays=1_000_000, n_steps=100, result_path=None, binary=False, seed=0,
        inline_materials=True, distant_rays=10_000, scene=None, caster_type=P.RAY_CASTER_GRID_HIP,
        frame_sink=None, device_loop=True):
    """controller.cpp:98-242. Returns {"frames": [(n_step, step, T)],
    "seconds": wall time of set-up + steps (the controller's timer),
    "n_meshes", "n_faces", "task_temperatures"}. result_path: the frames are
    also written there as the controller writes them (obj_export::task /
    task_binary). scene: an already imported scene (formats.obj_import's
    dict) instead of scene_path. frame_sink(n_step, step, T) is called with
    the initial frame and after every step, as the controller writes them.
    device_loop: each step keeps the equations' emission and absorption in
    device memory up to the FP64 update (the grid caster; the brute-force
    caster always runs the equations on host arrays); False runs every step
    on host arrays, with the same temperatures bit for bit; a sequence of
    booleans sets it step by step."""
    L = P.lib()
    formats.lib()
    caster = L.ffp_ray_caster_create(int(caster_type))
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
        per_step = not isinstance(device_loop, (bool, int))
        if not per_step and L.ffp_thermal_solution_set_device_loop(sol, 1 if device_loop else 0) < 0:
            raise RuntimeError("thermal solution: device loop setting refused")
        frames = [(task.n_step, task.time_delta, np.frombuffer(Tc, np.float32).copy())]
        if frame_sink is not None:
            frame_sink(*frames[-1])
        t0 = time.perf_counter()
        r = L.ffp_thermal_solution_set_scene(sol, C.byref(scene), Tc)
        if r < 0:
            raise RuntimeError(f"thermal solution set_scene failed ({r})")
        for step in range(n_steps):
            if per_step and L.ffp_thermal_solution_set_device_loop(sol, 1 if device_loop[step] else 0) < 0:
                raise RuntimeError("thermal solution: device loop setting refused")
            r = L.ffp_thermal_solution_calculate(sol, C.byref(task))
            if r < 0:
                raise RuntimeError(f"thermal solution step {step} failed ({r})")
            frames.append((task.n_step, task.time_delta, np.frombuffer(Tc, np.float32).copy()))
            if frame_sink is not None:
                frame_sink(*frames[-1])
        seconds = time.perf_counter() - t0
        if result_path is not None:
            formats.task_export(result_path, frames, binary=binary)
        return {"frames": frames, "seconds": seconds, "n_meshes": n, "n_faces": len(sc["faces"]),
                "task_temperatures": int(count)}
    finally:
        for kind, s in reversed(systems):
            (L.ffp_thermal_solution_free if kind == "sol" else L.ffp_thermal_equation_free)(s)
        L.ffp_emission_free(exterior)
        L.ffp_emission_free(interior)
        L.ffp_ray_caster_free(caster)


def _atoi(text):
    """C atoi: optional sign and leading digits, else 0 (the controller's parsing)."""
    t = text.strip()
    k = 1 if t[:1] in "+-" else 0
    while k < len(t) and t[k].isdigit():
        k += 1
    try:
        return int(t[:k])
    except ValueError:
        return 0


def parse_args(argv):
    """The controller's positional arguments (controller.cpp:40-96):
    (options dict, 0), (None, 1) for a usage error, or (None, negative code)
    for an unknown caster."""
    inline = "--inline-materials" in argv
    args = [a for a in argv if a != "--inline-materials"]
    if len(args) < 2:
        return None, 1
    opts = {"scene": args[0], "task": args[1], "result": "-", "n_rays": 1_000_000, "n_steps": 100,
            "caster": "cpu", "caster_type": CASTERS["cpu"], "binary": False, "inline_materials": inline}
    if len(args) > 2:
        opts["result"] = args[2]
    if len(args) > 3:
        opts["n_rays"] = _atoi(args[3])
    if len(args) > 4:
        opts["n_steps"] = _atoi(args[4])
    if len(args) > 5:
        if args[5] not in CASTERS:
            return None, -RAY_CASTER_NOT_SUPPORTED
        opts["caster"], opts["caster_type"] = args[5], CASTERS[args[5]]
    if len(args) > 6:
        opts["binary"] = args[6] == "1"
    return opts, 0


USAGE = ("usage: python -m form_factors_amd.controller <scene.obj> <task> [result file, - for stdout] "
         "[rays per step, 1000000] [steps, 100] [caster: cpu|cuda|naive_cpu|naive_cuda, cpu] [binary frames 0|1] "
         "[--inline-materials]")


def main(argv=None):
    opts, code = parse_args(sys.argv[1:] if argv is None else list(argv))
    if opts is None:
        if code == 1:
            print(USAGE)
        else:
            print("controller: the caster must be one of cpu, cuda, naive_cpu, naive_cuda", file=sys.stderr)
        return code
    try:
        sc = formats.obj_import(opts["scene"], inline_materials=opts["inline_materials"])
    except formats.FormatError as e:
        print(f"controller: cannot read the scene {opts['scene']} ({e.code})", file=sys.stderr)
        return e.code
    if not os.access(opts["task"], os.R_OK):  # controller.cpp:129-136
        print(f"controller: cannot open the task {opts['task']}", file=sys.stderr)
        return -1
    writer = None
    try:
        writer = formats.FrameWriter(opts["result"], binary=opts["binary"])
        out = run(None, opts["task"], n_rays=opts["n_rays"], n_steps=opts["n_steps"], scene=sc,
                  caster_type=opts["caster_type"], frame_sink=writer.write)
    except formats.FormatError as e:
        print(f"controller: cannot read the task {opts['task']} ({e.code})", file=sys.stderr)
        return e.code if e.code < 0 else -1
    finally:
        if writer is not None:
            writer.close()
    if out["task_temperatures"] < out["n_meshes"]:
        print(f"controller: the task sets {out['task_temperatures']} of {out['n_meshes']} temperatures; "
              "the others started at 300 K", file=sys.stderr)
    print(f"#Done in {out['seconds']:f}s.\n", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
