"""Batched form factors: many scenes of uneven size through one persistent
launch (optionally a replayed hipGraph) must give, scene by scene, the same
F bits as the plain single-scene launch and as the oracle."""
import numpy as np
import pytest

from conftest import bits

pytestmark = pytest.mark.gpu


def build_scenes(gpu, sizes, rays_per_tri, seed0=100):
    from form_factors_amd import scenes

    out = []
    for k, n in enumerate(sizes):
        faces, meshes = scenes.mixed_scene(n, seed0 + k)
        sc = gpu.Scene(faces, meshes)
        sc.set_emission(n_rays=rays_per_tri[k] * len(faces), frame=k % 2)
        sc.build_grid((0, 0, 0))
        out.append(sc)
    return out


def test_batched_equals_plain_and_oracle(gpu, oracle):
    import torch

    sizes = [40, 300, 120, 900, 60]
    rpt = [400, 100, 300, 60, 1000]
    scs = build_scenes(gpu, sizes, rpt)
    seeds = [11 + 7 * k for k in range(len(scs))]
    plain = [sc.ff_fused(s).cpu().numpy() for sc, s in zip(scs, seeds)]
    for sc, s, F in zip(scs, seeds, plain):
        r, want = oracle.ff_calculate(sc.faces, sc.meshes, int(rpt[scs.index(sc)] * len(sc.faces)), seed=s,
                                      frame=scs.index(sc) % 2, n_threads=8)
        assert np.array_equal(bits(F), bits(want))
    Fs = [torch.full((sc.n_meshes, sc.n_meshes), -1.0, device="cuda") for sc in scs]
    tallies = [torch.empty((sc.n_meshes, sc.n_meshes), dtype=torch.int64, device="cuda") for sc in scs]
    b = gpu.Batch(scs)
    for use_graph in (False, True):
        for rep in range(3):
            for F in Fs:
                F.fill_(-1.0)
            b.run(seeds, Fs, tallies=tallies, use_graph=use_graph)
            torch.cuda.synchronize()
            # eager work between graph replays
            junk = torch.randn(1000, device="cuda").sum().item()
            assert junk == junk
            for F, want in zip(Fs, plain):
                assert np.array_equal(bits(F.cpu().numpy()), bits(want)), (use_graph, rep)
    b.close()


def test_scenes_on_two_streams(gpu):
    """Fused launches of different scenes queued back to back on two streams
    keep their own arguments: each F equals its single-stream result."""
    import torch

    scs = build_scenes(gpu, [300, 200, 500, 100], [300, 400, 200, 600], seed0=21)
    want = [sc.ff_fused(40 + k).cpu().numpy() for k, sc in enumerate(scs)]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    Fs = [torch.full((sc.n_meshes,) * 2, -1.0, device="cuda") for sc in scs]
    torch.cuda.synchronize()
    for rep in range(3):
        for k, sc in enumerate(scs):
            with torch.cuda.stream(streams[k % 2]):
                sc.ff_fused(40 + k, F=Fs[k])
        torch.cuda.synchronize()
        for F, w in zip(Fs, want):
            assert np.array_equal(bits(F.cpu().numpy()), bits(w)), rep
        for F in Fs:
            F.fill_(-1.0)
        torch.cuda.synchronize()


def test_batched_tuning_does_not_change_bits(gpu):
    import torch

    scs = build_scenes(gpu, [200, 500], [200, 100], seed0=7)
    seeds = [3, 4]
    plain = [sc.ff_fused(s).cpu().numpy() for sc, s in zip(scs, seeds)]
    b = gpu.Batch(scs)
    for block, rays_per_tile, n_blocks, refill, pool, walk in [
            (256, 1024, 0, 0, 0, 0), (512, 4096, 64, 16, 0, 0), (1024, 0, 7, 0, 0, 0), (512, 0, 0, 32, 0, 0),
            (512, 0, 0, 16, 4, 0), (256, 4096, 9, 8, 1, 0), (1024, 0, 0, 16, 8, 0), (512, 0, 0, 16, 4, 1),
            (256, 4096, 9, 8, 1, 1), (512, 0, 0, 32, 2, 3), (256, 4096, 9, 8, 1, 2)]:
        b.set_tuning(block, rays_per_tile, n_blocks)
        b.set_refill(refill, pool)
        b.set_walk(walk)
        Fs = [torch.empty((sc.n_meshes,) * 2, device="cuda") for sc in scs]
        tallies = [torch.empty((sc.n_meshes,) * 2, dtype=torch.int64, device="cuda") for sc in scs]
        b.run(seeds, Fs, tallies=tallies, use_graph=True)
        torch.cuda.synchronize()
        for F, want in zip(Fs, plain):
            assert np.array_equal(bits(F.cpu().numpy()), bits(want)), (block, rays_per_tile, n_blocks)
    b.close()


def test_batch_scene_set_batched_equals_plain(gpu):
    """The batched workload shape (uneven scenes: log-uniform triangle counts,
    per-scene ray budgets), reduced in size: one persistent graph-replayed
    launch gives every scene the bits of its own plain launch."""
    import torch

    from form_factors_amd import scenes

    specs = scenes.batch_scene_set(12, seed=4, tri_range=(200, 4000), ray_range=(50, 400))
    scs = []
    for k, (faces, meshes, rpt, _) in enumerate(specs):
        sc = gpu.Scene(faces, meshes)
        sc.set_emission(n_rays=rpt * len(faces), frame=gpu.FRAME_ORTHONORMAL)
        sc.build_grid()
        scs.append(sc)
    seeds = [s for (_, _, _, s) in specs]
    plain = [sc.ff_fused(sd).cpu().numpy() for sc, sd in zip(scs, seeds)]
    b = gpu.Batch(scs)
    Fs = [torch.full((sc.n_meshes,) * 2, -1.0, device="cuda") for sc in scs]
    tallies = [torch.empty((sc.n_meshes,) * 2, dtype=torch.int64, device="cuda") for sc in scs]
    for rep in range(2):
        b.run(seeds, Fs, tallies=tallies, use_graph=True)
        torch.cuda.synchronize()
        for k, (F, want) in enumerate(zip(Fs, plain)):
            assert np.array_equal(bits(F.cpu().numpy()), bits(want)), (rep, k)
    b.close()


def test_graph_replay_with_new_seeds_each_run(gpu):
    """A captured batch replayed with different seeds on every run (the seeds
    are copied ahead of each replay; the graph is kept) gives each run the
    bits of plain launches with that run's seeds, with eager work between."""
    import torch

    scs = build_scenes(gpu, [150, 400, 90], [300, 100, 500], seed0=33)
    Fs = [torch.empty((sc.n_meshes,) * 2, device="cuda") for sc in scs]
    tallies = [torch.empty((sc.n_meshes,) * 2, dtype=torch.int64, device="cuda") for sc in scs]
    b = gpu.Batch(scs)
    for rep in range(6):
        seeds = [1000 * rep + 17 * k + 1 for k in range(len(scs))]
        print(f"batch replay {rep} seeds {seeds}", flush=True)
        b.run(seeds, Fs, tallies=tallies, use_graph=True)
        got = [F.cpu().numpy() for F in Fs]  # synchronises
        junk = torch.randn(4096, device="cuda").sum().item()
        assert junk == junk
        for sc, sd, F in zip(scs, seeds, got):
            assert np.array_equal(bits(F), bits(sc.ff_fused(sd).cpu().numpy())), (rep, sd)
    # several replays queued back to back before one synchronisation: the last one's seeds win
    for rep in range(5):
        b.run([7 + rep, 8 + rep, 9 + rep], Fs, tallies=tallies, use_graph=True)
    torch.cuda.synchronize()
    for sc, sd, F in zip(scs, [11, 12, 13], Fs):
        assert np.array_equal(bits(F.cpu().numpy()), bits(sc.ff_fused(sd).cpu().numpy()))
    b.close()


def test_batch_refuses_a_rebuilt_scene(gpu):
    """A batch holds its scenes' device tables as they were at creation: after
    a scene's grid is rebuilt or its emission set again, run and set_tuning
    raise before launching anything; a new batch runs."""
    import torch

    scs = build_scenes(gpu, [120, 200], [200, 200], seed0=5)
    Fs = [torch.empty((sc.n_meshes,) * 2, device="cuda") for sc in scs]
    tallies = [torch.empty((sc.n_meshes,) * 2, dtype=torch.int64, device="cuda") for sc in scs]
    for rebuild in ("grid", "emission"):
        b = gpu.Batch(scs)
        b.run([1, 2], Fs, tallies=tallies, use_graph=True)
        torch.cuda.synchronize()
        if rebuild == "grid":
            scs[1].build_grid((5, 5, 5))
        else:
            scs[0].set_emission(n_rays=150 * len(scs[0].faces), frame=gpu.FRAME_ORTHONORMAL)
        for call in (lambda: b.run([1, 2], Fs, tallies=tallies, use_graph=True),
                     lambda: b.run([1, 2], Fs, tallies=tallies, use_graph=False),
                     lambda: b.set_tuning(512, 0, 0)):
            with pytest.raises(RuntimeError, match="rebuilt"):
                call()
        b.close()
        b = gpu.Batch(scs)
        b.run([1, 2], Fs, tallies=tallies, use_graph=False)
        torch.cuda.synchronize()
        for sc, sd, F in zip(scs, [1, 2], Fs):
            assert np.array_equal(bits(F.cpu().numpy()), bits(sc.ff_fused(sd).cpu().numpy()))
        b.close()


def test_batch_autotune_keeps_bits_and_caches_on_disk(gpu, tmp_path, monkeypatch):
    """Per-class tuning (walk, ray scheduling, block) and the batch-wide tile
    size change no F bit; the result is written to the tuning cache file and,
    once the in-memory results are dropped, read back from it for a new batch
    of the same scenes (the same class settings and tile size). The
    single-scene autotuner shares the file."""
    import torch

    cache = tmp_path / "autotune.txt"
    monkeypatch.setenv("FFK_TUNE_CACHE", str(cache))
    gpu.lib().ffk_tune_cache_reset()
    assert gpu.tune_cache_path() == str(cache)
    scs = build_scenes(gpu, [300, 900, 120, 2500], [200, 60, 400, 30], seed0=71)
    seeds = [5, 6, 7, 8]
    plain = [sc.ff_fused(s).cpu().numpy() for sc, s in zip(scs, seeds)]
    b = gpu.Batch(scs)
    Fs = [torch.full((sc.n_meshes,) * 2, -1.0, device="cuda") for sc in scs]
    tallies = b.make_tallies()
    print("batch autotune", flush=True)
    tuned, cached = b.autotune(seeds, Fs, tallies, sample_rays=1 << 18)
    assert not cached and len(tuned) == len(b.classes())
    torch.cuda.synchronize()
    for F, want in zip(Fs, plain):  # the autotuner's own last run
        assert np.array_equal(bits(F.cpu().numpy()), bits(want))
    for F in Fs:
        F.fill_(-1.0)
    b.run(seeds, Fs, tallies=tallies, use_graph=True)
    torch.cuda.synchronize()
    for F, want in zip(Fs, plain):
        assert np.array_equal(bits(F.cpu().numpy()), bits(want))
    tiling = b.tuning()
    b.close()
    lines = cache.read_text().splitlines()
    assert len(lines) == 1 and lines[0].startswith("batch ")
    gpu.lib().ffk_tune_cache_reset()
    b2 = gpu.Batch(scs)
    tuned2, cached2 = b2.autotune(seeds, Fs, b2.make_tallies(), sample_rays=1 << 18)
    assert cached2 and tuned2 == tuned and b2.tuning() == tiling
    b2.close()
    # the single-scene tuner: timed once, then read back from the file
    sc = scs[1]
    t1 = sc.autotune(sample_rays=1 << 16)
    assert not t1["cached"] and sum(l.startswith("scene ") for l in cache.read_text().splitlines()) == 1
    gpu.lib().ffk_tune_cache_reset()
    t2 = sc.autotune(sample_rays=1 << 16)
    assert t2["cached"] and t2["dims"] == t1["dims"] and t2["block"] == t1["block"]
