"""Throughput of the Monte Carlo form-factor path on one MI355X.

Workload (default): closed unit icosphere, subdivision 5 = 20,480 triangles
with inward normals, one mesh per triangle, n_rays = 2.048e8 (about 10k
Malley rays per triangle, two-sided), uniform grid 64^3, orthonormal emission
frame, output the 20,480 x 20,480 FP32 form-factor matrix in HBM.

One step = one full form-factor calculation on the device: Malley emission
from the counter-based RNG, grid casting, exact integer tallies, row
normalisation (the fused kernel). The scene, grid and emission layout are
built before the timed region; inputs are resident in HBM.

Prints one JSON line with a live roofline for the fused kernel and the
reference CPU caster timed on a bounded prefix of the identical ray set.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic FP32 work per unit (counted on the device per launch):
FLOP_EMIT = 40      # per ray: point 12, hemisphere 6 + 4 transcendental, rotation 15, ray 9 (RNG excluded)
FLOP_TEST = 51      # per triangle test with the per-face terms precomputed
FLOP_STEP = 6       # per grid cell stepped (3-D DDA)
PEAK_FP32 = 157.3   # MI355X FP32 vector/matrix peak, TFLOP/s
PEAK_FP32_UNPACKED = 78.6  # FP32 without packed (two-wide) instructions
PEAK_HBM = 8000.0   # GB/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--subdiv", type=int, default=5)
    ap.add_argument("--rays-per-face", type=int, default=10_000)
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--cpu-rays", type=int, default=1 << 22, help="prefix of the ray set timed on the CPU")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-thermal", action="store_true")
    ap.add_argument("--no-batch-set", action="store_true")
    ap.add_argument("--no-confetti", action="store_true")
    ap.add_argument("--no-radiance", action="store_true")
    ap.add_argument("--no-controller", action="store_true")
    ap.add_argument("--no-plates", action="store_true")
    ap.add_argument("--only", choices=["plates", "batch_set", "confetti", "radiance", "controller", "thermal"],
                    help="run this one workload alone (no headline line): for profiling")
    ap.add_argument("--leg-steps", type=int, default=3, help="timed launches of the confetti workload")
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--thermal-steps", type=int, default=10_000)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--rays-per-tile", type=int, default=0)
    ap.add_argument("--hist", type=int, default=-1)
    ap.add_argument("--refill", type=int, default=-1, help="fused-kernel refill threshold (lanes; -1 = default)")
    ap.add_argument("--pool", type=int, default=-1, help="pooled-test ratio (0 = per-lane tests; -1 = default)")
    ap.add_argument("--walk", type=int, default=-1,
                    help="walk mode 0..7 (bit 0: jump for every empty cell, bit 1: whole chunks of pooled tests, "
                         "bit 2: face terms loaded in two stages); "
                         "-1 = one probe launch of each, the fastest kept")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="run one untimed device copy of this size (HBM counter calibration under rocprofv3)")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles(name):
    """HBM bytes per launch of a kernel from the committed PMC summary
    (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the factors a plain copy
    of known size gives), if present."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def ff_roofline(n_rays, n_tests, n_cells, seconds, f_bytes):
    """Roofline of a form-factor pass from its device-counted work: algorithmic
    FP32 FLOPs (emission, reference tests, grid steps) against the FP32 peak,
    and the algorithmic HBM bytes (the F rows, written once) against the HBM
    peak."""
    flops = FLOP_EMIT * n_rays + FLOP_TEST * n_tests + FLOP_STEP * n_cells
    tf = flops / seconds / 1e12
    gbs = f_bytes / seconds / 1e9
    return {"bound": "valu", "engine": "FP32 VALU", "achieved": round(tf, 3), "peak": PEAK_FP32, "unit": "TFLOP/s",
            "frac": round(tf / PEAK_FP32, 4), "frac_unpacked_fp32": round(tf / PEAK_FP32_UNPACKED, 4),
            "hbm_GB_per_s": round(gbs, 1), "hbm_frac": round(gbs / PEAK_HBM, 5),
            "tests_per_ray": round(n_tests / n_rays, 3), "cells_per_ray": round(n_cells / n_rays, 3)}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(faces, meshes, rays_host, gpu_idx, n_threads, c1):
    """The reference CPU path on the host cores, on bounded samples of the
    same workloads (compiled from the reference sources when available, else
    the oracle's restatement):

    * zgrid_cpu cast of a prefix of the identical C2 ray set on n_threads
      threads through one caster (the reference casters are re-entrant on
      disjoint slices), with hit-index mismatches against the GPU -- `value`;
    * the same cast on one thread, and naive_cpu on a 2^16-ray prefix;
    * one reference FORM_FACTORS_CPU calculate() on C2's scene at a reduced
      ray count, single-threaded as the reference runs, split into ray
      generation, casting and tally + normalisation (cpu_system.cpp:149-201);
    * C1 (400 triangles, 4e6 rays, zgrid) in full, single-threaded.
    """
    from oracle import oracle as O

    kind = "reference" if O.ref_available() else "port"

    def cast(k, rays, threads):
        t0 = time.perf_counter()
        if kind == "reference":
            idx, _ = O.ref_cast(k, faces, rays, n_threads=threads)
        elif k == 1:
            idx, _ = O.zgrid_cast(faces, rays, n_threads=threads)
        else:
            idx, _ = O.naive_cast(faces, rays, n_threads=threads)
        return idx, time.perf_counter() - t0

    n = len(rays_host)
    idx, dt = cast(1, rays_host, n_threads)
    mism = int((idx != gpu_idx).sum())
    n1 = min(n, 1 << 18)
    idx1, dt1 = cast(1, rays_host[:n1], 1)
    mism += int((idx1 != gpu_idx[:n1]).sum())
    nn = min(n, 1 << 16)
    idxn, dtn = cast(0, rays_host[:nn], n_threads)
    mism_naive = int((idxn != gpu_idx[:nn]).sum())
    out = {"value": round(n / dt / 1e6, 4), "unit": "Mrays/s", "cores": n_threads, "kind": kind,
           "sample": f"zgrid_cpu cast of the first {n} rays of the identical C2 ray set on {n_threads} threads "
                     f"({dt:.1f} s; {mism} hit-index mismatches vs the GPU)",
           "cpu_model": cpu_model(), "host_cpus_visible": os.cpu_count(),
           "threads_note": f"{n_threads} threads (a one-GPU box has a 16-thread CPU share), each casting a "
                           "disjoint slice through one caster",
           "zgrid_1_thread": {"value": round(n1 / dt1 / 1e6, 5), "rays": n1, "seconds": round(dt1, 2)},
           "naive": {"value": round(nn / dtn / 1e6, 5), "rays": nn, "threads": n_threads, "seconds": round(dtn, 2),
                     "mismatches_vs_gpu": mism_naive}}
    if kind == "reference":
        n_req = 1 << 18
        w, total = O.area_weights(faces)
        fr, rr = O.emitted_counts(w, n_req, total)
        n_em = int(fr.sum() + rr.sum())
        _, t3 = O.ref_ff_timed(faces, meshes, n_req, caster=1)
        out["ff_calculate_split_1_thread"] = {
            "scene": f"C2 icosphere, {len(faces)} triangles, {len(meshes)} meshes", "rays": n_em,
            "emission_s": round(float(t3[0]), 4), "cast_s": round(float(t3[1]), 3),
            "tally_normalise_s": round(float(t3[2]), 3),
            "emission_Mrays_per_s": round(n_em / t3[0] / 1e6, 3), "cast_Mrays_per_s": round(n_em / t3[1] / 1e6, 5),
            "note": "the normalisation walks the whole N^2 F once per call: a fixed cost next to the rays"}
        c1_faces, c1_meshes = c1
        _, t1 = O.ref_ff_timed(c1_faces, c1_meshes, 4_000_000, caster=1)
        out["c1_full_1_thread"] = {
            "scene": "C1 two 10x10 plates, 400 triangles", "rays": 4_000_000,
            "seconds": round(float(t1.sum()), 3), "value": round(4.0 / t1.sum(), 4),
            "emission_s": round(float(t1[0]), 3), "cast_s": round(float(t1[1]), 3),
            "tally_normalise_s": round(float(t1[2]), 3)}
    return out, mism


def perftest_crosscheck(ffk, n_threads):
    """perfTest's caster comparison (perfTest/perfTest.cpp:155-255) on its own
    workload: 38,000 confetti triangles in a ball of radius 20 and 3.8e6 rays
    collapsing toward the centre (from the reference's spherical generator
    when the compiled reference is present). The HIP grid caster (the ZGRID
    slot: scene set, grid built, rays cast, timed as perfTest times it), the
    HIP brute force (the NAIVE slot) and the reference's zgrid_cpu caster on
    n_threads threads cast the same rays; CalculateError's report (mismatched
    hit faces; average, max and total hit-point distance) compares each with
    the HIP grid caster, as perfTest compares with its CUDA zgrid caster."""
    import torch

    from form_factors_amd import crosscheck
    from oracle import oracle as O

    if O.ref_available():
        faces, rays = O.ref_perftest_scene(38000, 20.0)
        gen = "reference spherical generator (subject/spherial_generator.cpp)"
    else:
        faces, rays = crosscheck.perftest_scene(38000, 20.0)
        gen = "restated generator (form_factors_amd/crosscheck.py)"
    d_rays = torch.from_numpy(rays).cuda()
    torch.cuda.synchronize()

    def gpu_cast(method):
        t0 = time.perf_counter()
        sc = ffk.Scene(faces)
        if method == ffk.CAST_GRID:  # the brute force needs no grid
            sc.build_grid()
        idx, pts = sc.cast(d_rays, method=method)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        t1 = time.perf_counter()
        sc.cast(d_rays, method=method)
        torch.cuda.synchronize()
        cast_only = time.perf_counter() - t1
        sc.close()
        return idx.cpu().numpy(), pts.cpu().numpy(), el, cast_only

    gpu_cast(ffk.CAST_GRID)  # warm-up (perfTest warms up on a 32-face scene)
    gi, gp, g_s, g_cast = gpu_cast(ffk.CAST_GRID)
    ni, np_, n_s, n_cast = gpu_cast(ffk.CAST_NAIVE)
    t0 = time.perf_counter()
    if O.ref_available():
        ci, cp = O.ref_cast(1, faces, rays, n_threads=n_threads)
    else:
        ci, cp = O.zgrid_cast(faces, rays, n_threads=n_threads)
    c_s = time.perf_counter() - t0
    e_naive = crosscheck.error_stats(ni, np_, gi, gp)
    e_cpu = crosscheck.error_stats(ci, cp, gi, gp)
    n = len(rays)
    log(f"perfTest: grid HIP {g_s:.3f} s ({g_cast * 1e3:.2f} ms cast), naive HIP {n_s:.3f} s, "
        f"zgrid_cpu {c_s:.2f} s on {n_threads} threads; mismatches naive {e_naive['mismatch']}, "
        f"cpu {e_cpu['mismatch']}; max distance naive {e_naive['max_distance']:.3g}, cpu {e_cpu['max_distance']:.3g}")
    return {"workload": f"perfTest confetti: {len(faces)} triangles in a ball of radius 20, {n} collapsing rays, "
                        f"{gen}", "rays": n, "hits": int((gi >= 0).sum()),
            "grid_hip": {"seconds": round(g_s, 4), "cast_seconds": round(g_cast, 5),
                         "Mrays_per_s": round(n / g_cast / 1e6, 1)},
            "naive_hip": {"seconds": round(n_s, 4), "cast_seconds": round(n_cast, 5),
                          "Mrays_per_s": round(n / n_cast / 1e6, 1), "error_vs_grid_hip": e_naive,
                          "naive_over_grid_time": round(n_s / g_s, 3)},
            "zgrid_cpu": {"kind": "reference" if O.ref_available() else "port", "threads": n_threads,
                          "seconds": round(c_s, 3), "Mrays_per_s": round(n / c_s / 1e6, 3),
                          "error_vs_grid_hip": e_cpu, "cpu_over_gpu_time": round(c_s / g_s, 1)}}


def batch_set_leg(ffk, scenes, n_scenes):
    """Uneven scene set: n_scenes scenes, triangle counts log-uniform in
    [200, 50000], a per-scene mix of sphere / cube shell / confetti and
    1k-20k Malley rays per triangle (scene set seed 0). Plain = one fused
    launch per scene, back to back; batched = one graph-replayed persistent
    launch per LDS class for all scenes. The timed passes use seeds the
    warm-up did not (the batched graph is kept and only the seeds are copied
    ahead of the replay), and must give every scene the same F bits, compared
    in full on the device."""
    import torch

    t0 = time.perf_counter()
    specs = scenes.batch_scene_set(n_scenes, seed=0)
    scs, seeds = [], []
    for faces, meshes, rpt, sd in specs:
        sc = ffk.Scene(faces, meshes)
        sc.set_emission(n_rays=rpt * len(faces), frame=ffk.FRAME_ORTHONORMAL)
        sc.build_grid()
        scs.append(sc)
        seeds.append(sd)
    n_rays = sum(sc.n_rays for sc in scs)
    Fs = [torch.empty((sc.n_meshes,) * 2, dtype=torch.float32, device="cuda") for sc in scs]
    plain_tallies = [torch.empty((sc.n_meshes,) * 2, dtype=torch.int64, device="cuda")
                     if ffk.lib().ffk_ff_needs_tally(sc.h) else None for sc in scs]
    torch.cuda.synchronize()
    log(f"scene set: {len(scs)} scenes, {sum(len(sc.faces) for sc in scs)} triangles, {n_rays} rays, "
        f"F {sum(sc.n_meshes ** 2 for sc in scs) * 4 / 1e9:.1f} GB, set-up {time.perf_counter() - t0:.1f} s")

    def plain_pass(sds):
        for sc, sd, F, T in zip(scs, sds, Fs, plain_tallies):
            sc.ff_fused(sd, F=F, tally=T, need_tally=T is not None)

    timed_seeds = [sd + 1 for sd in seeds]

    stream = torch.cuda.current_stream()

    def timed(fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t = time.perf_counter()
        e0.record(stream)
        fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return time.perf_counter() - t, e0.elapsed_time(e1) / 1e3

    plain_pass(seeds)  # warm-up
    wall_p, ev_p = timed(lambda: plain_pass(timed_seeds))
    F_plain = [F.clone() for F in Fs]
    del plain_tallies
    torch.cuda.empty_cache()
    log(f"scene set plain: {n_rays / wall_p / 1e6:.1f} Mrays/s ({wall_p:.3f} s)")
    b = ffk.Batch(scs)
    tallies = b.make_tallies()
    for F in Fs:
        F.fill_(-1.0)
    b.run(seeds, Fs, tallies=tallies, use_graph=True)  # warm-up: plan and capture
    wall_u, _ = timed(lambda: b.run(timed_seeds, Fs, tallies=tallies, use_graph=True))
    log(f"scene set batched, default settings: {n_rays / wall_u / 1e6:.1f} Mrays/s ({wall_u:.3f} s)")
    t0 = time.perf_counter()
    tuned, cached = b.autotune(seeds, Fs, tallies=tallies, sample_rays=1 << 28)
    t_tune = time.perf_counter() - t0
    rpt_tuned = b.tuning()[1]
    log(f"batch autotune ({t_tune:.1f} s{', cached' if cached else ''}): {tuned}, rays per tile {rpt_tuned or 'auto'}")
    b.run(seeds, Fs, tallies=tallies, use_graph=True)  # re-capture with the tuned settings
    wall_b, ev_b = timed(lambda: b.run(timed_seeds, Fs, tallies=tallies, use_graph=True))
    same = all(torch.equal(F.view(torch.int32), G.view(torch.int32)) for F, G in zip(Fs, F_plain))
    del F_plain
    log(f"scene set batched: {n_rays / wall_b / 1e6:.1f} Mrays/s ({wall_b:.3f} s), F identical: {same}")
    # device-counted work of the set (one counting run, untimed)
    stats = torch.zeros(2, dtype=torch.int64, device="cuda")
    b.run(seeds, Fs, tallies=tallies, use_graph=False, stats=stats)
    torch.cuda.synchronize()
    n_tests, n_cells = (int(x) for x in stats.cpu())
    roof = ff_roofline(n_rays, n_tests, n_cells, ev_b, sum(sc.n_meshes ** 2 for sc in scs) * 4)
    b.close()
    del Fs, tallies
    torch.cuda.empty_cache()
    return {"workload": f"{len(scs)} scenes, triangles log-uniform in [200, 50000] ({sum(len(sc.faces) for sc in scs)} "
                        f"total), sphere/cube/confetti mixes, 1k-20k rays per triangle, {n_rays} rays",
            "n_rays": n_rays,
            "plain": {"value": round(n_rays / wall_p / 1e6, 2), "seconds": round(wall_p, 4),
                      "events_seconds": round(ev_p, 4), "path": "one fused launch per scene"},
            "batched": {"value": round(n_rays / wall_b / 1e6, 2), "seconds": round(wall_b, 4),
                        "events_seconds": round(ev_b, 4),
                        "path": "persistent tile-queue launches per LDS class (scene by scene, scenes of more "
                                "than 24,576 meshes in the global tally), one hipGraph, autotuned per class",
                        "default_settings_value": round(n_rays / wall_u / 1e6, 2),
                        "class_tuning": {str(k): v for k, v in tuned.items()},
                        "rays_per_tile": rpt_tuned or "auto (at most 65535)",
                        "tuning_seconds": round(t_tune, 2), "tuning_cached": cached},
            "speedup": round(wall_p / wall_b, 4), "F_bits_identical": same,
            "F_compare": "every scene's F, all N^2 cells, bitwise on the device", "unit": "Mrays/s",
            "roofline_batched": roof}


def confetti_leg(ffk, scenes, steps):
    """2000 confetti triangles (centroids uniform in the unit cube, edges in
    [0.02, 0.1], random orientation) inside a 2028-triangle cube shell, one
    mesh per triangle, 50k Malley rays per triangle; the fused kernel on the
    default grid and after the autotuner; then 1e7 parallel rays from a fixed
    direction through a 2 x 2 square aperture and through the reference's
    disk aperture."""
    import torch

    faces, meshes = scenes.confetti_scene(2000, cells=13, seed=1)
    sc = ffk.Scene(faces, meshes)
    n_rays = sc.set_emission(n_rays=50_000 * len(faces), frame=ffk.FRAME_ORTHONORMAL)
    sc.build_grid()
    N = len(meshes)
    F = torch.empty((N, N), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()

    def run(k):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        sc.ff_fused(0xC3, F=F)  # warm-up
        torch.cuda.synchronize()
        e0.record(stream)
        for i in range(k):
            sc.ff_fused(0xC3 + i, F=F)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3 / k

    g0 = sc.grid_info()["dims"]
    t_default = run(steps)
    F_default = F.clone()
    t0 = time.perf_counter()
    tuned = sc.autotune(sample_rays=1 << 22)
    t_tune = time.perf_counter() - t0
    t_tuned = run(steps)
    same = torch.equal(F.view(torch.int32), F_default.view(torch.int32))
    del F_default
    stats = torch.zeros(2, dtype=torch.int64, device="cuda")
    sc.ff_fused(0xC3, F=F, stats=stats)  # device-counted work (untimed)
    torch.cuda.synchronize()
    n_tests, n_cells = (int(x) for x in stats.cpu())
    roof = ff_roofline(n_rays, n_tests, n_cells, t_tuned, N * N * 4)
    log(f"confetti: default grid {g0} {n_rays / t_default / 1e6:.1f} Mrays/s, tuned {tuned} "
        f"{n_rays / t_tuned / 1e6:.1f} Mrays/s (tuning {t_tune:.1f} s)")
    c, r0 = ffk.bsphere(faces)
    radius = np.float32(r0 * np.float32(1.1))
    d = np.array([0.36, -0.48, -0.8], np.float32)
    d /= np.linalg.norm(d)
    sun = {}
    for name, ap, rad in (("square_2x2", 1, np.float32(1.0)), ("disk_reference", 0, radius)):
        sc.sun_counts(d, c, np.float32(1.1) + radius, rad, 10_000_000, 5, aperture=ap)  # warm-up
        torch.cuda.synchronize()
        t = time.perf_counter()
        counts = sc.sun_counts(d, c, np.float32(1.1) + radius, rad, 10_000_000, 5, aperture=ap)
        el = time.perf_counter() - t
        sun[name] = {"rays": 10_000_000, "value": round(1e7 / el / 1e6, 2), "hits": int(counts.sum()),
                     "seconds": round(el, 4)}
        log(f"sun {name}: {1e7 / el / 1e6:.1f} Mrays/s, {int(counts.sum())} hits")
    del F
    torch.cuda.empty_cache()
    return {"workload": f"{len(faces)} triangles (2000 confetti + cube shell), {n_rays} Malley rays, F {N}x{N}",
            "n_rays": n_rays, "unit": "Mrays/s",
            "default_grid": {"dims": list(g0), "value": round(n_rays / t_default / 1e6, 2)},
            "autotuned": {"dims": list(tuned["dims"]), "block": tuned["block"], "walk": tuned["walk"],
                          "refill": tuned["refill"], "pool": tuned["pool"],
                          "value": round(n_rays / t_tuned / 1e6, 2), "tuning_seconds": round(t_tune, 2),
                          "roofline": roof},
            "F_bits_identical": same, "sun": sun}


def plates_leg(ffk, scenes, steps):
    """Parallel plates: two unit squares at distance 1, each a 10 x 10 grid of
    200 triangles (400 triangles, one mesh each), 10k Malley rays per triangle
    (4e6 rays), grid 16^3, the 400 x 400 F; the plate-to-plate form factor
    against the analytic 0.19982 (twice the reference's two-sided value)."""
    import torch

    faces, meshes = scenes.parallel_plates(10)
    sc = ffk.Scene(faces, meshes)
    n_rays = sc.set_emission(n_rays=10_000 * len(faces), frame=ffk.FRAME_ORTHONORMAL)
    sc.build_grid((16, 16, 16))
    N = len(meshes)
    F = torch.empty((N, N), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    for k in range(3):
        sc.ff_fused(0xC1 + k, F=F)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for k in range(steps):
        sc.ff_fused(0xC1 + k, F=F)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    ev = e0.elapsed_time(e1) / 1e3 / steps
    Fh = F.double().cpu().numpy()
    plate = float(2 * Fh[:200, 200:].sum() / 200)
    stats = torch.zeros(2, dtype=torch.int64, device="cuda")
    sc.ff_fused(0xC1, F=F, stats=stats)
    torch.cuda.synchronize()
    n_tests, n_cells = (int(x) for x in stats.cpu())
    roof = ff_roofline(n_rays, n_tests, n_cells, ev, N * N * 4)
    log(f"plates C1: {n_rays / wall / 1e6:.1f} Mrays/s ({wall * 1e3:.3f} ms/step, kernel {ev * 1e3:.3f} ms), "
        f"2 F(plate->plate) = {plate:.4f}")
    sc.close()
    return {"workload": f"two 10x10-cell unit plates at distance 1, {len(faces)} triangles, {n_rays} Malley rays, "
                        "grid 16^3, F 400x400", "n_rays": n_rays, "unit": "Mrays/s",
            "value": round(n_rays / wall / 1e6, 2), "ms_per_step": round(wall * 1e3, 4),
            "kernel_ms": round(ev * 1e3, 4), "roofline": roof,
            "plate_to_plate_x2": round(plate, 5), "analytic": 0.19982}


def radiance_leg(ffk, scenes):
    """radiance_equation::calculate per step on the two 100 x 100-cell plates
    (40,000 triangles, 20,000 meshes, temperatures 250-350 K): radiance
    weights T * eps * A, the emission layout, ray powers, then one device pass
    that emits, casts and deposits every ray (64-bit fixed-point sums). At the
    reference controller's 1e6 rays per step and at 1e8."""
    import torch

    prob = scenes.thermal_plates(cells=100)
    n = len(prob["meshes"])
    rng = np.random.default_rng(0x5AD)
    T = rng.uniform(250.0, 350.0, n).astype(np.float32)
    eps = np.stack([prob["eps_front"], prob["eps_rear"]], 1).astype(np.float32)
    sc = ffk.Scene(prob["faces"], prob["meshes"])
    sc.build_grid()
    out = {"workload": f"two 100x100-cell plates, {len(prob['faces'])} triangles, {n} meshes, T in [250, 350] K, "
                       "one radiance_equation step (weights, layout, powers, emit-cast-deposit pass)",
           "unit": "Mrays/s"}
    stream = torch.cuda.current_stream()
    for n_rays, reps in ((1_000_000, 20), (100_000_000, 3)):
        sc.set_walk(3)
        sc.radiance_step(eps, T, n_rays, 1)  # warm-up
        torch.cuda.synchronize()
        # Walk mode (the same absorption bits in every mode, checked): one
        # untimed pass of each at this ray count, the fastest kept, as the
        # C2 leg does for the fused kernel.
        w, total = ffk.radiance_weights(sc.faces, sc.meshes, eps, T)
        front, rear = ffk.emitted_counts(w, n_rays, total)
        p, _ = ffk.radiance_powers(sc.faces, sc.meshes, eps, T, front, rear)
        walk_ms, ref = {}, None
        for mode in range(8):
            sc.set_walk(mode)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            got = sc.radiance(5, p)
            e1.record(stream)
            torch.cuda.synchronize()
            walk_ms[mode] = e0.elapsed_time(e1)
            ref = got if ref is None else ref
            assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), f"radiance walk mode {mode} differs"
        walk = min(walk_ms, key=walk_ms.get)
        sc.set_walk(walk)
        t0 = time.perf_counter()
        for k in range(reps):
            em, ab = sc.radiance_step(eps, T, n_rays, 2 + k)
        step_s = (time.perf_counter() - t0) / reps
        cast = int(front.sum() + rear.sum())
        t0 = time.perf_counter()
        for k in range(reps):
            sc.radiance(2 + k, p)
        dep_s = (time.perf_counter() - t0) / reps
        frac = float(ab.sum() / em.astype(np.float64).sum())
        # device-counted work of the pass (one untimed counting run) and the
        # pass's event-timed duration on the launch stream
        stats = torch.zeros(2, dtype=torch.int64, device="cuda")
        sc.radiance(2, p, stats=stats)
        torch.cuda.synchronize()
        n_tests, n_cells = (int(x) for x in stats.cpu())
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for k in range(reps):
            sc.radiance(2 + k, p)
        e1.record(stream)
        torch.cuda.synchronize()
        pass_s = e0.elapsed_time(e1) / 1e3 / reps
        roof = ff_roofline(cast, n_tests, n_cells, pass_s, 0)
        roof["kernel"] = "k_radiance_pooled"
        roof["timing"] = "HIP events around the pass on its stream (walk-table upload and sums download included)"
        out[f"rays_{n_rays}"] = {"rays_cast": cast, "ms_per_step": round(step_s * 1e3, 3),
                                 "value": round(cast / step_s / 1e6, 2),
                                 "deposit_ms": round(dep_s * 1e3, 3), "deposit_value": round(cast / dep_s / 1e6, 2),
                                 "absorbed_over_emitted": round(frac, 5), "roofline": roof, "walk_mode": walk,
                                 "walk_probe_ms": {str(k): round(v, 3) for k, v in walk_ms.items()}}
        log(f"radiance {n_rays} rays: {step_s * 1e3:.3f} ms/step ({cast / step_s / 1e6:.1f} Mrays/s), "
            f"deposit pass {dep_s * 1e3:.3f} ms ({cast / dep_s / 1e6:.1f} Mrays/s), absorbed/emitted {frac:.4f}, "
            f"walk mode {walk} (" + " / ".join(f"{walk_ms[k]:.3f}" for k in sorted(walk_ms)) + " ms)")
    sc.close()
    return out


def controller_leg():
    """The reference controller's thermal run on its confetti_2k model (2000
    meshes, 100 Adams steps of 0.01 s, 1e6 radiance rays per step; the
    reference recorded 94.84 s on a GTX 780 Ti in models/confetti_2k.frames)
    through the plugin and file ABIs: OBJ and .task import, HIP radiance,
    conduction, heat-source and distant-source equations, FP64 Adams, frame
    export. Timed as the controller times it (set-up of the solution and the
    steps). Unit heat capacity, as in the recorded run; the frames' mean
    temperatures are compared with the recorded ones."""
    from form_factors_amd import controller, formats

    models = os.path.join(ROOT, "tests", "golden", "models")
    g = np.load(os.path.join(ROOT, "tests", "golden", "controller_confetti_2k.npz"))
    sc = formats.obj_import(os.path.join(models, "confetti_2k.obj"), inline_materials=True)
    sc["materials"][:, 1] = 1.0
    task = os.path.join(models, "confetti_2k.task")
    controller.run(None, task, n_rays=1_000_000, n_steps=5, seed=99, scene=sc)  # warm-up
    runs = []
    for seed in range(3):
        t0 = time.perf_counter()
        out = controller.run(None, task, n_rays=1_000_000, n_steps=100, seed=seed, scene=sc)
        wall = time.perf_counter() - t0
        T = np.array([f[2] for f in out["frames"]])
        runs.append((out["seconds"], wall, float(np.abs(T.mean(1) - g["mean"]).max())))
    best = min(r[0] for r in runs)
    ref = float(g["seconds"])
    log(f"controller confetti_2k: {best:.3f} s for 100 steps (reference {ref:.2f} s, {ref / best:.0f}x); "
        f"max |mean T - recorded| {max(r[2] for r in runs):.4f} K")
    return {"workload": "reference controller run, confetti_2k model (2000 meshes), 100 Adams steps dt=0.01 s, "
                        "1e6 radiance rays per step, plugin ABI end to end",
            "seconds": round(best, 4), "seconds_all": [round(r[0], 4) for r in runs],
            "wall_with_import_seconds": round(min(r[1] for r in runs), 4),
            "reference_seconds": ref, "reference_hardware": "GeForce GTX 780 Ti (recorded in the model's frames file)",
            "speedup_vs_recorded": round(ref / best, 1),
            "max_abs_mean_T_diff_K": round(max(r[2] for r in runs), 4),
            "patch_sphere_20k": controller_20k_leg()}


def controller_20k_leg():
    """The reference's largest model, models/patch_sphere_20k.obj (20,000
    one-face black-body meshes) with patch_sphere_20k_sun.task (a 30 W/m^2
    distant source and 1e4 sun rays per step), read from the files and run
    by the controller for 100 Adams steps at 1e6 radiance rays per step.
    The reference recorded no time for this model. Timed with the solution's
    device-resident step (the default) and, once, with every equation on
    host arrays as the reference's solutions call them (the same frames bit
    for bit), to show what the device-resident step saves."""
    import gzip
    import tempfile

    from form_factors_amd import controller

    models = os.path.join(ROOT, "tests", "golden", "models")
    with tempfile.TemporaryDirectory() as tmp:
        paths = []
        for name in ("patch_sphere_20k.obj", "patch_sphere_20k_sun.task"):
            paths.append(os.path.join(tmp, name))
            with open(os.path.join(models, name + ".gz"), "rb") as f, open(paths[-1], "wb") as o:
                o.write(gzip.decompress(f.read()))
        controller.run(paths[0], paths[1], n_rays=1_000_000, n_steps=5, seed=99)  # warm-up
        runs = []
        for seed in range(2):
            t0 = time.perf_counter()
            out = controller.run(paths[0], paths[1], n_rays=1_000_000, n_steps=100, seed=seed)
            runs.append((out["seconds"], time.perf_counter() - t0))
        host = controller.run(paths[0], paths[1], n_rays=1_000_000, n_steps=100, seed=1, device_loop=False)
    T = out["frames"][-1][2]
    same = all(np.array_equal(a[2], b[2]) for a, b in zip(out["frames"], host["frames"]))
    best = min(r[0] for r in runs)
    log(f"controller patch_sphere_20k sun: {best:.3f} s for 100 steps ({best * 10:.2f} ms/step; host arrays "
        f"{host['seconds'] * 10:.2f} ms/step, same frames: {same}), T after 1 s: {T.min():.2f}..{T.max():.2f} K")
    return {"workload": "reference controller run, patch_sphere_20k model (20,000 meshes) with the sun task, "
                        "100 Adams steps dt=0.01 s, 1e6 radiance rays + 1e4 sun rays per step",
            "seconds": round(best, 4), "seconds_all": [round(r[0], 4) for r in runs],
            "step": "device-resident (equations' arrays on the device up to the FP64 update)",
            "host_arrays_seconds": round(host["seconds"], 4), "host_arrays_same_frames": bool(same),
            "wall_with_import_seconds": round(min(r[1] for r in runs), 4),
            "T_range_after_100_steps_K": [round(float(T.min()), 3), round(float(T.max()), 3)]}


ADAMS_U = np.array([[1.0, 0, 0, 0, 0], [3 / 2, -1 / 2, 0, 0, 0], [23 / 12, -4 / 3, 5 / 12, 0, 0],
                    [55 / 24, -59 / 24, 37 / 24, -3 / 8, 0],
                    [1901 / 720, -1387 / 360, 109 / 30, -637 / 360, 251 / 720]])


def fp64_thermal_reference(F, prob, row, cols, G, T0, n_steps, dt):
    """All-FP64 transient run on the device in torch float64 (a separate code
    path from the library's kernels): q = F^T E with E = (eps_f + eps_r)
    sigma T^4 A (thermal_equation/form_factors.cpp:109-128), CSR conduction,
    heat sources, Adams AB1..AB5 (thermal_solution/cpu_adams.cpp:136-182).
    F is the float32 device matrix, promoted to FP64."""
    import torch

    dev = F.device
    n = F.shape[0]
    d = lambda a: torch.as_tensor(np.asarray(a, np.float64), device=dev)  # noqa: E731
    F64 = F.to(torch.float64)
    A, Cap, ef, er = d(prob["area"]), d(prob["C"]), d(prob["eps_front"]), d(prob["eps_rear"])
    src = torch.zeros(n, dtype=torch.float64, device=dev)
    src.index_add_(0, torch.as_tensor(np.asarray(prob["src_mesh"], np.int64), device=dev), d(prob["src_power"]))
    rows = torch.as_tensor(np.repeat(np.arange(n), np.diff(np.asarray(row))), dtype=torch.int64, device=dev)
    cols_t = torch.as_tensor(np.asarray(cols, np.int64), device=dev)
    G_t = d(G)
    E = torch.zeros((5, n), dtype=torch.float64, device=dev)
    T = d(T0).clone()
    for step in range(n_steps):
        power = 5.670400e-8 * T ** 4 * A
        em = ef * power + er * power
        flow = torch.zeros(n, dtype=torch.float64, device=dev)
        flow.index_add_(0, rows, G_t * (T[cols_t] - T[rows]))
        E[step % 5] = (torch.mv(F64.t(), em) - em + flow + src) / Cap
        k = min(step, 4)
        acc = torch.zeros(n, dtype=torch.float64, device=dev)
        for j in range(k + 1):
            acc += ADAMS_U[k, j] * E[(step - j) % 5]
        T = T + dt * acc
        if step % 2000 == 1999:
            log(f"fp64 reference: step {step + 1}")
    out = T.cpu().numpy()
    del F64
    torch.cuda.empty_cache()
    return out


def thermal_counts_above_256(ffk, scenes, n_steps):
    """200 meshes (two 10 x 10-cell plates) with 1e6 Malley rays per mesh, so
    F cells hold counts in the thousands: BF16 (8 significant bits) and FP16
    (11) round the stored counts. Max relative T error of an n_steps run
    against the FP64 reference, per storage type."""
    prob = scenes.thermal_plates(cells=10)
    n = len(prob["meshes"])
    sc = ffk.Scene(prob["faces"], prob["meshes"])
    sc.set_emission(n_rays=500_000 * len(prob["faces"]), frame=ffk.FRAME_ORTHONORMAL)
    sc.build_grid()
    F = sc.ff_fused(0xC6)
    rays = sc.rays_per_mesh().astype(np.float64)
    csr = ffk.conduction_graph(prob["faces"], prob["meshes"], prob["conductivity"], prob["thickness"])
    T0 = np.full(n, 300.0)
    ref = fp64_thermal_reference(F, prob, *csr, T0, n_steps, 0.1)
    out = {"workload": f"two 10x10-cell plates, {n} meshes, 1e6 rays per mesh, {n_steps} Adams steps",
           "max_count": int(round(float((F.double().cpu().numpy() * rays[:, None]).max())))}
    for name, dt in (("f32", ffk.F32), ("bf16", ffk.BF16), ("f16", ffk.F16)):
        th = ffk.Thermal(prob["C"], prob["area"], prob["eps_front"], prob["eps_rear"])
        S, scale = ffk.pack_exchange(F, dt, row_mul=rays)
        th.set_exchange(S, dt, scale)
        th.set_conduction(*csr)
        th.set_sources(prob["src_mesh"], prob["src_power"])
        T = th.run(n_steps, 0.1, T0=T0)
        out[f"max_rel_T_err_vs_fp64_{name}"] = float(np.abs(T - ref).max() / np.abs(ref).max())
        th.close()
    log(f"thermal counts > 256: {out}")
    sc.close()
    return out


def thermal_leg(ffk, scenes, n_steps):
    """Two 100 x 100-cell plates (40,000 triangles, 20,000 meshes): F from the
    fused kernel, then n_steps Adams steps of dt = 0.1 s with the exchange
    q = F^T E read from BF16 and FP16 copies of F. Each step reads F once:
    N^2 * 2 B + 3 N * 8 B of algorithmic HBM traffic."""
    import torch

    prob = scenes.thermal_plates(cells=100)
    n = len(prob["meshes"])
    sc = ffk.Scene(prob["faces"], prob["meshes"])
    sc.set_emission(n_rays=500 * len(prob["faces"]), frame=ffk.FRAME_ORTHONORMAL)
    sc.build_grid()
    F = sc.ff_fused(0xC5)
    rays = sc.rays_per_mesh().astype(np.float64)  # F rows are counts / rays: stored as exact counts
    csr = ffk.conduction_graph(prob["faces"], prob["meshes"], prob["conductivity"], prob["thickness"])
    T0 = np.full(n, 300.0)
    out = {"workload": f"two 100x100-cell plates, {len(prob['faces'])} triangles, {n} meshes, "
                       f"{n_steps} Adams steps dt=0.1 s, conduction CSR nnz={len(csr[2])}, 1 kW source",
           "n_steps": n_steps}
    runs = {}
    stream = torch.cuda.current_stream()
    for name, dt in (("f32", ffk.F32), ("bf16", ffk.BF16), ("f16", ffk.F16)):
        th = ffk.Thermal(prob["C"], prob["area"], prob["eps_front"], prob["eps_rear"])
        S, scale = ffk.pack_exchange(F, dt, row_mul=rays)
        th.set_exchange(S, dt, scale)
        th.set_conduction(*csr)
        th.set_sources(prob["src_mesh"], prob["src_power"])
        th.run(20, 0.1, T0=T0)  # warm-up: graph capture and first touch
        torch.cuda.synchronize()
        # The T0 upload, start-up kernels and the final copy are inside the
        # events too: per-step figures are conservative.
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        T = th.run(n_steps, 0.1, T0=T0)
        e1.record(stream)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        ev_ms = e0.elapsed_time(e1) / n_steps
        runs[name] = T
        elem = {"f32": 4, "bf16": 2, "f16": 2}[name]
        step_bytes = n * n * elem + 3 * n * 8
        gbs = step_bytes / (ev_ms * 1e-3) / 1e9
        out[name] = {"ms_per_step": round(el * 1e3 / n_steps, 4), "kernel_ms_per_step": round(ev_ms, 4),
                     "GB_per_s": round(gbs, 1), "frac_hbm": round(gbs / PEAK_HBM, 4),
                     "algorithmic_bytes_per_step": step_bytes}
        log(f"thermal {name}: {el * 1e3 / n_steps:.4f} ms/step wall, {ev_ms:.4f} ms/step events, {gbs:.0f} GB/s")
        th.close()
        del th, S
        torch.cuda.empty_cache()
    out["roofline"] = {"bound": "hbm", "kernel": "k_step (BF16 storage)", "achieved": out["bf16"]["GB_per_s"],
                       "peak": PEAK_HBM, "unit": "GB/s", "frac": out["bf16"]["frac_hbm"],
                       "traffic": traffic_from_profiles("k_step_bf16_plates20k_traffic.json")}
    t0 = time.perf_counter()
    ref = fp64_thermal_reference(F, prob, *csr, T0, n_steps, 0.1)
    log(f"thermal fp64 reference: {time.perf_counter() - t0:.1f} s")
    for name in ("f32", "bf16", "f16"):
        out[name]["max_rel_T_err_vs_fp64"] = float(np.abs(runs[name] - ref).max() / np.abs(ref).max())
    out["note"] = ("T error is against an all-FP64 run (torch float64 on the device: q = F^T E from F promoted "
                   "to FP64, CSR conduction, sources, Adams) from the same F over the same steps")
    out["T_range_K"] = [float(ref.min()), float(ref.max())]
    del F
    torch.cuda.empty_cache()
    out["counts_above_256"] = thermal_counts_above_256(ffk, scenes, n_steps)
    return out


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1).", file=sys.stderr)
        sys.exit(2)
    import torch

    from form_factors_amd import ffk, scenes

    torch.cuda.set_device(0)
    info = ffk.device_info()
    log(f"device: {info}")
    if args.only:
        legs = {"plates": lambda: plates_leg(ffk, scenes, 20),
                "batch_set": lambda: batch_set_leg(ffk, scenes, args.batch_scenes),
                "confetti": lambda: confetti_leg(ffk, scenes, args.leg_steps),
                "radiance": lambda: radiance_leg(ffk, scenes), "controller": controller_leg,
                "thermal": lambda: thermal_leg(ffk, scenes, args.thermal_steps)}
        print(json.dumps({"only": args.only, args.only: legs[args.only]()}), flush=True)
        return
    faces = scenes.icosphere(args.subdiv, inward=True)
    meshes = scenes.per_face_meshes(len(faces))
    n_req = args.rays_per_face * len(faces)
    sc = ffk.Scene(faces, meshes)
    n_rays = sc.set_emission(n_rays=n_req, frame=ffk.FRAME_ORTHONORMAL)
    g = sc.build_grid((args.grid,) * 3)
    sc.set_tuning(block=args.block, rays_per_tile=args.rays_per_tile, hist=args.hist)
    if args.refill >= 0 or args.pool >= 0:
        r0, p0 = sc.get_refill()
        sc.set_refill(r0 if args.refill < 0 else args.refill, p0 if args.pool < 0 else args.pool)
    if args.walk >= 0:
        sc.set_walk(args.walk)
    N = len(meshes)
    log(f"scene: {len(faces)} faces, {N} meshes, {n_rays} rays, grid {g['dims']} refs {g['n_refs']}")
    F = torch.empty((N, N), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    seed = 0x5EED
    if args.calib_copy_mib:
        src = torch.ones(args.calib_copy_mib << 18, dtype=torch.float32, device="cuda")
        dst = src.clone()
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calib_copy_mib} MiB read + {args.calib_copy_mib} MiB written")
        del src, dst

    # Device-counted work for the roofline (one untimed launch with counters).
    stats = torch.zeros(2, dtype=torch.int64, device="cuda")
    sc.ff_fused(seed, F=F, need_tally=False, stats=stats)
    torch.cuda.synchronize()
    n_tests, n_cells = (int(x) for x in stats.cpu())
    flops = FLOP_EMIT * n_rays + FLOP_TEST * n_tests + FLOP_STEP * n_cells
    log(f"work per step: {n_tests / n_rays:.2f} tests/ray, {n_cells / n_rays:.2f} cells/ray, {flops / 1e12:.3f} TFLOP")

    # Walk mode (same hits either way): unless given, one untimed launch of
    # each, the faster kept -- what ffk_autotune does for the walk mode,
    # without changing the grid this configuration fixes.
    walk_times = {}
    if args.walk < 0:
        for m in range(8):
            sc.set_walk(m)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            sc.ff_fused(seed, F=F, need_tally=False)
            e1.record(stream)
            torch.cuda.synchronize()
            walk_times[m] = e0.elapsed_time(e1)
        walk = min(walk_times, key=walk_times.get)
        sc.set_walk(walk)
        log(f"walk mode {walk} (one launch each: " + " / ".join(f"{walk_times[m]:.2f}" for m in sorted(walk_times)) + " ms)")
    else:
        walk = args.walk
    # Ray schedule (same hits in every setting): unless given, two untimed
    # launches of each of the autotuner's candidates, the fastest kept.
    sched_times = {}
    if args.refill < 0 and args.pool < 0:
        for r, p in ((32, 2), (24, 3), (16, 4), (40, 3), (40, 4)):
            sc.set_refill(r, p)
            best = 3e38
            for _ in range(2):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                sc.ff_fused(seed, F=F, need_tally=False)
                e1.record(stream)
                torch.cuda.synchronize()
                best = min(best, e0.elapsed_time(e1))
            sched_times[f"{r}/{p}"] = best
        pick = min(sched_times, key=sched_times.get)
        sc.set_refill(*(int(x) for x in pick.split("/")))
        log(f"ray schedule {pick} (" + " / ".join(f"{k} {v:.2f}" for k, v in sched_times.items()) + " ms)")
    schedule = "%d/%d" % sc.get_refill()
    for _ in range(args.warmup):
        sc.ff_fused(seed, F=F, need_tally=False)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        sc.ff_fused(seed + k, F=F, need_tally=False)
        ev[k][1].record(stream)
        log(f"step {k} done")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    ms_per_step = wall * 1e3 / args.steps
    value = n_rays / (ms_per_step * 1e-3) / 1e6

    # Analytic check of the last F: closed sphere, F_ij = A_j / (2 A_total)
    # for the two-sided convention (rear rays leave through the inside too).
    Fh = F[:64].double().cpu().numpy()
    w, _ = ffk.area_weights(faces)
    A = w[0::2].astype(np.float64)
    row_sums = Fh.sum(1)
    # Groups of 256 consecutive target faces: the group's count is
    # Binomial(rays of the row, sum of A_j / (2 A_total)) with about 60
    # expected hits, where a 3-sigma band holds 99.7% of the groups.
    rays_row = np.asarray(sc.rays_per_mesh()[:64], np.float64)
    p = A / (2.0 * A.sum())
    n_g = len(p) // 256
    pg = p[:n_g * 256].reshape(n_g, 256).sum(1)
    Fg = Fh[:, :n_g * 256].reshape(64, n_g, 256).sum(2)
    sigma = np.sqrt(pg[None, :] * (1 - pg[None, :]) / rays_row[:, None])
    outside = float((np.abs(Fg - pg[None, :]) > 3 * sigma).mean())
    sphere_ok = bool(np.abs(row_sums - 0.5).max() < 0.02 and outside < 0.01)

    result = {
        "metric": "Mrays/s form-factor MC (20k-tri sphere), batched vs plain; % HBM/FP32 peak",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"closed icosphere interior, {len(faces)} triangles (inward), one mesh per "
                               f"triangle, {n_rays} Malley rays (~{args.rays_per_face}/triangle), "
                               f"grid {args.grid}^3, plain fused emit-cast-tally kernel, F {N}x{N} fp32",
                   "n_faces": len(faces), "n_rays": n_rays, "grid": list(g["dims"]), "parallelism": "single-gpu",
                   "walk_mode": walk, "walk_probe_ms": {str(k): round(v, 3) for k, v in walk_times.items()},
                   "ray_schedule": schedule, "schedule_probe_ms": {k: round(v, 3) for k, v in sched_times.items()}},
        "roofline": {"bound": "valu", "achieved": round(flops / (kern_ms * 1e-3) / 1e12, 3), "peak": PEAK_FP32,
                     "unit": "TFLOP/s", "frac": round(flops / (kern_ms * 1e-3) / 1e12 / PEAK_FP32, 4),
                     "traffic": traffic_from_profiles("k_ff_fused_icosphere20k_traffic.json"),
                     "engine": "FP32 vector ALU: traversal and intersection have no matrix operation; peak is "
                               "the MI355X FP32 vector rate (157.3 TFLOP/s with packed FP32, 78.6 without)",
                     "frac_unpacked_fp32": round(flops / (kern_ms * 1e-3) / 1e12 / PEAK_FP32_UNPACKED, 4),
                     "kernel": "k_ff_fused", "kernel_ms": round(kern_ms, 3),
                     "work": {"tests_per_ray": round(n_tests / n_rays, 3), "cells_per_ray": round(n_cells / n_rays, 3),
                              "flop_per_launch": flops}},
        "check": {"sphere_row_sum_max_dev": float(np.abs(row_sums - 0.5).max()),
                  "sphere_groups_outside_3sigma": outside, "ok": sphere_ok},
        "device": info["name"],
    }

    log(f"plain: {value:.1f} Mrays/s, {ms_per_step:.2f} ms/step (kernel {kern_ms:.2f} ms)")
    if not args.no_batched:
        # The same scene through the persistent, graph-replayed launch.
        try:
            b = ffk.Batch([sc])
            b.set_walk(walk)
            b.set_refill(*sc.get_refill())
            b.run([seed], [F], use_graph=True)
            torch.cuda.synchronize()
            reps = max(2, args.steps // 2)
            tb = time.perf_counter()
            for k in range(reps):
                b.run([seed + 1 + k], [F], use_graph=True)
            torch.cuda.synchronize()
            bt = (time.perf_counter() - tb) / reps
            result["batched"] = {"value": round(n_rays / bt / 1e6, 2), "ms_per_step": round(bt * 1e3, 3),
                                 "path": "persistent tile-queue kernel, hipGraph replay"}
            log(f"batched: {n_rays / bt / 1e6:.1f} Mrays/s")
            b.close()
        except Exception as e:  # reported, never silently replaced
            result["batched"] = {"error": str(e)}

    del F
    torch.cuda.empty_cache()
    if not args.no_plates:
        result["plates"] = plates_leg(ffk, scenes, 20)
    if not args.no_batch_set:
        result["batch_set"] = batch_set_leg(ffk, scenes, args.batch_scenes)
    if not args.no_confetti:
        result["confetti"] = confetti_leg(ffk, scenes, args.leg_steps)
    if not args.no_radiance:
        result["radiance"] = radiance_leg(ffk, scenes)
    if not args.no_controller:
        result["controller"] = controller_leg()
    if not args.no_thermal:
        result["thermal"] = thermal_leg(ffk, scenes, args.thermal_steps)

    if not args.no_cpu:
        n_cpu = min(args.cpu_rays, n_rays)
        rays = torch.empty((n_rays, 6), dtype=torch.float32, device="cuda")
        ffk.lib().ffk_emit_malley(sc.h, seed, ffk._ptr(rays), ffk._stream())
        idx, _ = sc.cast(rays[:n_cpu], method=ffk.CAST_GRID, points=False)
        torch.cuda.synchronize()
        rays_h = rays[:n_cpu].cpu().numpy()
        idx_h = idx.cpu().numpy()
        del rays
        # the CPU share of a one-GPU box is 16 threads, whatever the host shows
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"cpu baseline on {n_cpu} rays, {threads} threads ...")
        result["cpu_baseline"], mism = cpu_baseline(faces, meshes, rays_h, idx_h, threads,
                                                    scenes.parallel_plates(10))
        result["check"]["cpu_prefix_mismatches"] = mism
        result["cpu_baseline"]["perftest"] = perftest_crosscheck(ffk, threads)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
