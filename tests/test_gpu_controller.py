"""The reference controller's thermal run (controller/controller.cpp:40-242)
end to end on the GPU through the plugin and file ABIs: OBJ scene and .task
in, radiance + conduction + heat sources + distant source summed by the
FP64 Adams solver, frames out."""
import os

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

MODELS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models")


def confetti_scene(heat_capacity=None):
    from form_factors_amd import formats

    sc = formats.obj_import(os.path.join(MODELS, "confetti_2k.obj"), inline_materials=True)
    if heat_capacity is not None:
        sc["materials"][:, 1] = heat_capacity
    return sc


def test_confetti_2k_matches_the_reference_frames(tmp_path):
    """models/confetti_2k.frames is the reference controller's own output for
    this scene: 2000 black-body confetti at 300 K, 100 steps of 0.01 s, 1e6
    radiance rays per step. It was recorded with unit heat capacity (the
    black body of subject/objects.cpp; the model file now says 20): with
    that, the mean mesh temperature of every frame agrees within 0.15 K
    (Monte Carlo noise of the mean over 2000 meshes is a few hundredths of a
    kelvin; measured 0.036 K) and the spread across meshes within 0.35 K."""
    from form_factors_amd import controller, formats

    g = golden("controller_confetti_2k.npz")
    out = controller.run(None, os.path.join(MODELS, "confetti_2k.task"), n_rays=1_000_000, n_steps=100, seed=0,
                         scene=confetti_scene(1.0), result_path=str(tmp_path / "run.frames"))
    assert [f[0] for f in out["frames"]] == list(range(101))
    assert all(abs(f[1] - float(g["step"])) == 0 for f in out["frames"])
    T = np.array([f[2] for f in out["frames"]])
    assert np.array_equal(T[0], g["first"])
    dm = np.abs(T.mean(1) - g["mean"])
    ds = np.abs(T.std(1) - g["std"])
    print(f"controller confetti_2k: {out['seconds']:.3f} s for 100 steps (reference {float(g['seconds']):.1f} s); "
          f"max |mean diff| {dm.max():.4f} K, max |std diff| {ds.max():.4f} K", flush=True)
    assert dm.max() < 0.15 and ds.max() < 0.35
    # the frames file reads back as the controller wrote it
    back = formats.task_import(str(tmp_path / "run.frames"), 2000, 101)
    assert [f[0] for f in back["frames"]] == [2000] * 101
    assert np.allclose(np.array([f[2] for f in back["frames"]]), T, atol=5e-7 * 300)


def test_heat_capacity_comes_from_the_scene():
    """Same seed, same first step: the radiance rays and the net powers are
    identical, so the model's own heat capacity (20) cools each mesh 20 times
    less than unit heat capacity in the first step."""
    from form_factors_amd import controller

    task = os.path.join(MODELS, "confetti_2k.task")
    a = controller.run(None, task, n_rays=200_000, n_steps=1, seed=5, scene=confetti_scene(1.0))
    b = controller.run(None, task, n_rays=200_000, n_steps=1, seed=5, scene=confetti_scene())
    da = 300.0 - a["frames"][1][2].astype(np.float64)
    db = 300.0 - b["frames"][1][2].astype(np.float64)
    assert da.mean() > 1.0
    assert np.abs(20.0 * db - da).max() < 20 * 3e-5 * 300


def test_sun_task_energy_balance(oracle):
    """patch_sphere_2k_sun.task: emissivity 0 (no radiance), conduction
    between patches, a 30 W/m^2 distant source along -x. After the first
    (Euler) step the stored power sum_m C_m dT_m / dt equals the absorbed
    sunlight -- conduction only moves heat between meshes -- and the
    absorbed sunlight is the oracle's sum over the same parallel rays."""
    from form_factors_amd import controller, formats

    sc = formats.obj_import(os.path.join(MODELS, "patch_sphere_2k.obj"), inline_materials=True)
    task = os.path.join(MODELS, "patch_sphere_2k_sun.task")
    t = formats.task_import(task, len(sc["meshes"]), 1, step=0.1)
    seed = 9
    out = controller.run(None, task, n_rays=100_000, n_steps=1, seed=seed, scene=sc, distant_rays=10_000)
    T0, T1 = out["frames"][0][2].astype(np.float64), out["frames"][1][2].astype(np.float64)
    dt = float(out["frames"][1][1])
    faces = sc["faces"].reshape(-1, 3, 3)
    area = 0.5 * np.linalg.norm(np.cross(faces[:, 1] - faces[:, 0], faces[:, 2] - faces[:, 0]), axis=1)
    m = sc["meshes"]
    mat = sc["materials"][m[:, 2]]
    C_m = np.array([area[a:a + k].sum() for a, k, _ in m]) * mat[:, 0] * mat[:, 1] * mat[:, 3]
    stored = (C_m * (T1 - T0)).sum() / dt
    power, dx, dy, dz = t["distant_source"]
    absorb = np.stack([mat[:, 6], mat[:, 11]], 1).astype(np.float32)
    key = (seed + 1) % (1 << 64)  # the exterior emitter's first key
    _, ab64 = oracle.sun(sc["faces"], m, absorb, power, np.array([dx, dy, dz], np.float32), 10_000, seed=key,
                         n_threads=8)
    print(f"sun step: stored {stored:.3f} W, absorbed {ab64.sum():.3f} W", flush=True)
    assert ab64.sum() > 0
    assert abs(stored - ab64.sum()) <= 2e-3 * ab64.sum()


def unpacked(name, tmp_path):
    """A gzipped model file under tests/golden/models written out into tmp_path."""
    import gzip

    out = tmp_path / name[:-3]
    out.write_bytes(gzip.decompress(open(os.path.join(MODELS, name), "rb").read()))
    return str(out)


def test_patch_sphere_20k_sun_end_to_end(tmp_path, oracle):
    """The reference's models/patch_sphere_20k.obj (20,000 one-face black
    body meshes, inline material) with patch_sphere_20k_sun.task (three
    meshes at 400 K, the rest at 10 K, 30 W/m^2 along -x), read from the
    files and run by the controller for three Adams steps (AB1, AB2, AB3)
    at the controller's 1e6 radiance rays per step. Each step is recomputed
    from the frames: the oracle's radiance (the interior emitter's k-th key
    and frame 0) and sunlight (the exterior emitter's k-th key) at the
    frame's temperatures give E = (absorbed - emitted) / C, combined by the
    oracle's Adams update from its own FP64 state. Every mesh agrees to the
    float rounding of the powers and of the frame, except meshes that an
    edge-grazing ray can move (a different nearest face within the casters'
    1e-6 barycentric margin: at most 1e-5 of the rays, each moving one ray's
    power between two meshes)."""
    from form_factors_amd import controller, formats

    obj = unpacked("patch_sphere_20k.obj.gz", tmp_path)
    task = unpacked("patch_sphere_20k_sun.task.gz", tmp_path)
    seed, n_rays, sun_rays, n_steps = 4, 1_000_000, 10_000, 3
    print("controller patch_sphere_20k sun task, 3 steps", flush=True)
    out = controller.run(obj, task, n_rays=n_rays, n_steps=n_steps, seed=seed, distant_rays=sun_rays,
                         result_path=str(tmp_path / "run.frames"))
    assert out["n_meshes"] == 20000 and out["n_faces"] == 20000 and out["task_temperatures"] == 20000
    assert [f[0] for f in out["frames"]] == list(range(n_steps + 1))
    frames = np.array([f[2] for f in out["frames"]])
    assert np.isfinite(frames).all()
    sc = formats.obj_import(obj, inline_materials=True)
    t = formats.task_import(task, 20000, 1, step=0.1)
    assert np.array_equal(frames[0], t["frames"][0][2])
    dt = float(np.float32(t["frames"][0][1]))
    assert all(float(f[1]) == dt for f in out["frames"])
    faces, m = sc["faces"], sc["meshes"]
    mat = sc["materials"][m[:, 2]].astype(np.float64)
    tri = faces.reshape(-1, 3, 3).astype(np.float64)
    area = 0.5 * np.linalg.norm(np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=1)
    C = np.array([area[a:a + k].sum() for a, k, _ in m]) * mat[:, 0] * mat[:, 1] * mat[:, 3]
    eps = np.stack([mat[:, 8], mat[:, 13]], 1).astype(np.float32)
    absorb = np.stack([mat[:, 6], mat[:, 11]], 1).astype(np.float32)
    power, dx, dy, dz = t["distant_source"]
    golden_step = 0x9E3779B97F4A7C15
    E = np.zeros((5, 20000))
    T_o = frames[0].astype(np.float64)
    scale = np.zeros(20000)
    for k in range(n_steps):
        em, _, ab = oracle.radiance(faces, m, eps, frames[k], n_rays, rng=oracle.RNG_PHILOX,
                                    seed=(seed + golden_step * k) % (1 << 64), frame=0, caster=1, n_threads=8)
        _, sun = oracle.sun(faces, m, absorb, power, np.array([dx, dy, dz], np.float32), sun_rays,
                            seed=(seed + 1 + golden_step * k) % (1 << 64), n_threads=8)
        E[k % 5] = (ab + sun - em) / C
        T_o = oracle.adams_combine_f64(k, dt, T_o, E)
        scale += 4.0 * (ab + sun + em) / C
        got = frames[k + 1].astype(np.float64)
        err = np.abs(got - T_o)
        tol = np.spacing(np.float32(T_o)).astype(np.float64) + 1e-6 * dt * scale
        off = np.flatnonzero(err > tol)
        print(f"step {k + 1}: max |T - oracle| {err.max():.3e} K, {len(off)} meshes beyond rounding; "
              f"sun {sun.sum():.1f} W, radiance absorbed {ab.sum():.1f} W of {em.sum():.1f} W", flush=True)
        assert sun.sum() > 0 and ab.sum() > 0
        assert len(off) <= 2e-5 * n_rays
        assert (err[off] <= tol[off] + 4.0 * dt * em.max() / C[off]).all()
    # sunlit 10 K meshes warm; the 400 K meshes cool
    lit = np.flatnonzero(sun > 0)
    assert len(lit) > 1000 and (frames[-1][lit[lit >= 3]] > 10.0).all()
    assert (frames[-1][:3] < 400.0).all()
    back = formats.task_import(str(tmp_path / "run.frames"), 20000, n_steps + 1)
    assert np.allclose(np.array([f[2] for f in back["frames"]]), frames, rtol=1e-6)
    print(f"controller patch_sphere_20k: {out['seconds']:.3f} s for {n_steps} steps", flush=True)


@pytest.mark.parametrize("obj,task,n_rays,n_steps", [
    ("confetti_2k.obj", "confetti_2k.task", 200_000, 6),           # radiance + conduction
    ("patch_sphere_2k.obj", "patch_sphere_2k_sun.task", 100_000, 6),  # distant source, emissivity 0
    ("grid_10x10.obj", "grid_10x10_heat.task", 50_000, 8),         # heat source and sink
    ("tgrid_10x10.obj", "tgrid_10x10_heat.task", 50_000, 8),       # sinks, materials from the .mtl
])
def test_device_loop_equals_host_arrays(obj, task, n_rays, n_steps):
    """The Adams solution's device-resident step (the equations' emission and
    absorption kept in device memory up to the FP64 update) against the same
    run with every equation on host arrays, as the reference's solutions call
    them: the same rays, the same float sums in the same order, so every
    frame is identical bit for bit."""
    from form_factors_amd import controller, formats

    sc = formats.obj_import(os.path.join(MODELS, obj), inline_materials=True)
    path = os.path.join(MODELS, task)
    print(f"device loop vs host arrays: {obj} {task}", flush=True)
    runs = [controller.run(None, path, n_rays=n_rays, n_steps=n_steps, seed=11, scene=sc, distant_rays=20_000,
                           device_loop=dev) for dev in (True, False)]
    Td, Th = (np.array([f[2] for f in r["frames"]]) for r in runs)
    assert [f[0] for f in runs[0]["frames"]] == [f[0] for f in runs[1]["frames"]] == list(range(n_steps + 1))
    assert np.isfinite(Td).all() and not np.array_equal(Td[0], Td[-1])
    assert np.array_equal(Td.view(np.uint32), Th.view(np.uint32))
    print(f"  {n_steps} steps: device loop {runs[0]['seconds']:.4f} s, host arrays {runs[1]['seconds']:.4f} s",
          flush=True)


def test_device_loop_switched_mid_run():
    """Switching between the device-resident step and host arrays from one
    step to the next (the solution's FP64 state, the Adams history and the
    emitters' ray keys carry over either way): the frames are those of a
    run that never switched. Radiance, conduction and the distant source."""
    from form_factors_amd import controller, formats

    sc = formats.obj_import(os.path.join(MODELS, "patch_sphere_2k.obj"), inline_materials=True)
    sc["materials"][:, 8] = 0.9  # front emissivity: radiance emits as well as the sun heating
    path = os.path.join(MODELS, "patch_sphere_2k_sun.task")
    plan = [True, False, False, True, False, True, True, False]
    print("device loop switched step by step: patch_sphere_2k + sun", flush=True)
    runs = [controller.run(None, path, n_rays=100_000, n_steps=len(plan), seed=3, scene=sc, distant_rays=20_000,
                           device_loop=dl) for dl in (True, plan)]
    Ta, Tb = (np.array([f[2] for f in r["frames"]]) for r in runs)
    assert not np.array_equal(Ta[0], Ta[-1])
    assert np.array_equal(Ta.view(np.uint32), Tb.view(np.uint32))


@pytest.mark.parametrize("caster", ["cuda", "naive_cpu"])
def test_command_line_run(tmp_path, caster):
    """python -m form_factors_amd.controller with the controller's positional
    arguments: frames written to the result file as the steps end, the
    initial frame first, '#Done in' on standard output; with '-' the frames
    go to standard output ahead of that line."""
    import subprocess
    import sys

    from form_factors_amd import formats

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    scene = os.path.join(MODELS, "parallel_planes.obj")
    task = os.path.join(MODELS, "parallel_planes.task")
    out = tmp_path / "frames.txt"
    print(f"controller command line, caster {caster}", flush=True)
    r = subprocess.run([sys.executable, "-m", "form_factors_amd.controller", scene, task, str(out), "20000", "3",
                        caster, "0"], capture_output=True, text=True, cwd=root, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.startswith("#Done in ") and r.stdout.endswith("s.\n\n")
    text = out.read_text()
    assert text.count("newfrm frame_") == 4 and text.startswith("newfrm frame_0\n")
    frames = formats.task_import(str(out), 2, 4)["frames"]
    assert len(frames) == 4 and all(len(f[2]) == 2 for f in frames)
    r2 = subprocess.run([sys.executable, "-m", "form_factors_amd.controller", scene, task, "-", "20000", "2", caster],
                        capture_output=True, text=True, cwd=root, timeout=240)
    assert r2.returncode == 0, r2.stderr[-2000:]
    assert r2.stdout.count("newfrm frame_") == 3 and r2.stdout.rstrip().split("\n")[-1].startswith("#Done in ")
    bad = subprocess.run([sys.executable, "-m", "form_factors_amd.controller", scene, task, "-", "1", "1", "opencl"],
                         capture_output=True, text=True, cwd=root, timeout=240)
    assert bad.returncode == 254  # -RAY_CASTER_NOT_SUPPORTED as an exit status
