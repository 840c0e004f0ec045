"""The uniform grid's cell words (ffk_scene_grid_cells) against their
definition: occupied cells hold the exclusive scan of the reference counts;
an empty cell holds, for each direction octant, the size k (1..15) of the
largest box [c, c + s*(k-1)], clipped to the grid, that holds no reference.
The expected sizes come from a summed-volume table of the occupancy (box
counts in O(1)), independent of the device's relaxation; the vertices of
every face lie in occupied cells."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _box_sizes(occ):
    """k[o, z, y, x] by brute force over box sizes 1..15 (0 where occupied)."""
    nz, ny, nx = occ.shape
    sv = np.zeros((nz + 1, ny + 1, nx + 1), np.int64)
    sv[1:, 1:, 1:] = occ.astype(np.int64).cumsum(0).cumsum(1).cumsum(2)

    def count(z0, z1, y0, y1, x0, x1):  # occupied cells in [z0, z1) x [y0, y1) x [x0, x1)
        return (sv[z1, y1, x1] - sv[z0, y1, x1] - sv[z1, y0, x1] - sv[z1, y1, x0]
                + sv[z0, y0, x1] + sv[z0, y1, x0] + sv[z1, y0, x0] - sv[z0, y0, x0])

    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    out = np.zeros((8, nz, ny, nx), np.int64)
    for o in range(8):
        sx, sy, sz = (-1 if o & 1 else 1), (-1 if o & 2 else 1), (-1 if o & 4 else 1)
        k = np.zeros_like(x)
        alive = ~occ
        for size in range(1, 16):
            def span(c, s, n):
                a = c if s > 0 else np.maximum(c - (size - 1), 0)
                b = np.minimum(c + (size - 1), n - 1) + 1 if s > 0 else c + 1
                return a, b
            x0, x1 = span(x, sx, nx)
            y0, y1 = span(y, sy, ny)
            z0, z1 = span(z, sz, nz)
            empty = count(z0, z1, y0, y1, x0, x1) == 0
            alive &= empty
            k = np.where(alive, size, k)
        out[o] = k
    return out


def _cases():
    from form_factors_amd import scenes

    faces, _ = scenes.confetti_scene(300, cells=5, seed=4)
    out = [("confetti_auto", faces, (0, 0, 0))]
    out.append(("sphere3_skewed", scenes.icosphere(3, inward=True), (10, 17, 7)))
    faces, _ = scenes.parallel_plates(cells=6, gap=1.0)
    out.append(("plates_40", faces, (40, 40, 40)))
    out.append(("sphere2_1", scenes.icosphere(2, inward=True), (1, 1, 1)))
    return out


@pytest.mark.parametrize("case", range(4))
def test_grid_cells_match_definition(gpu, case):
    name, faces, dims = _cases()[case]
    faces = np.asarray(faces, np.float32).reshape(-1, 9)
    print(f"grid cells {name} dims {dims}", flush=True)
    sc = gpu.Scene(faces)
    info = sc.build_grid(dims)
    cells = sc.grid_cells()
    first, cnt = cells[..., 0].astype(np.int64), cells[..., 1].astype(np.int64)
    occ = cnt > 0
    # occupied cells: exclusive scan of the counts, in cell order; counts sum to the references
    flat_c = cnt.reshape(-1)
    starts = np.concatenate([[0], np.cumsum(flat_c)[:-1]])
    assert np.array_equal(first.reshape(-1)[flat_c > 0], starts[flat_c > 0])
    assert flat_c.sum() == info["n_refs"]
    # empty cells: the per-octant box sizes
    want = _box_sizes(occ)
    words = cells[..., 0].astype(np.int64) & 0xFFFFFFFF
    for o in range(8):
        got = (words >> (4 * o)) & 15
        assert np.array_equal(np.where(occ, 0, got), want[o]), (name, o)
    # every vertex lies in an occupied cell (faces are binned into every cell they touch, grown by a pad)
    lo, hi = np.array(info["lo"], np.float64), np.array(info["hi"], np.float64)
    n = np.array(info["dims"])
    v = faces.reshape(-1, 3).astype(np.float64)
    c = np.clip(np.floor((v - lo) / ((hi - lo) / n)).astype(int), 0, n - 1)
    assert occ[c[:, 2], c[:, 1], c[:, 0]].all()
    sc.close()
