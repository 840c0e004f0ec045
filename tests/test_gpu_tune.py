"""The autotuner: the chosen grid resolution and block size never change the
form factors (hit indices are independent of the grid), and a second call on
the same scene is served from the cache."""
import numpy as np
import pytest

from tests.conftest import bits

pytestmark = pytest.mark.gpu


def test_autotune_keeps_bits_and_caches(gpu, oracle):
    from form_factors_amd import scenes

    faces, meshes = scenes.confetti_scene(300, cells=4, seed=6)
    sc = gpu.Scene(faces, meshes)
    n = sc.set_emission(n_rays=400 * len(faces), frame=gpu.FRAME_ORTHONORMAL)
    sc.build_grid()
    before = sc.ff_fused(21).cpu().numpy()
    t = sc.autotune(sample_rays=50_000)
    print("tuned", t, flush=True)
    assert not t["cached"] and min(t["dims"]) >= 1 and t["block"] in (0, 256, 512, 1024) and t["walk"] in range(8)
    assert (t["refill"], t["pool"]) in ((32, 2), (24, 3), (16, 4), (40, 3), (40, 4))
    assert sc.get_refill() == (t["refill"], t["pool"])
    assert sc.grid_info()["dims"] == t["dims"]
    after = sc.ff_fused(21).cpu().numpy()
    assert np.array_equal(bits(after), bits(before))
    r, want = oracle.ff_calculate(faces, meshes, 400 * len(faces), seed=21, frame=1, n_threads=8)
    assert np.array_equal(bits(after), bits(want))
    # the layout was restored: same ray count, same emission
    assert sc.n_rays == n
    sc2 = gpu.Scene(faces, meshes)
    sc2.set_emission(n_rays=400 * len(faces), frame=gpu.FRAME_ORTHONORMAL)
    t2 = sc2.autotune(sample_rays=50_000)
    assert t2["cached"] and t2["dims"] == t["dims"] and t2["block"] == t["block"] and t2["walk"] == t["walk"]
    assert (t2["refill"], t2["pool"]) == (t["refill"], t["pool"]) and sc2.get_refill() == (t["refill"], t["pool"])
    assert np.array_equal(bits(sc2.ff_fused(21).cpu().numpy()), bits(before))
