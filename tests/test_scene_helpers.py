"""Host-side scene helpers used by the thermal tests and the bench."""
import numpy as np

from form_factors_amd import scenes


def test_thermal_plates_patch_fits_small_plates():
    for cells in (3, 6, 12, 100):
        p = scenes.thermal_plates(cells=cells)
        n = len(p["meshes"])
        assert n == 2 * cells * cells
        assert p["src_mesh"].min() >= 0 and p["src_mesh"].max() < cells * cells
        assert np.isclose(p["src_power"].sum(), 1000.0)
