"""Host-side set-up of the thermal path (no kernel launch): the mesh
conduction operator against the reference's conductive equation."""
import numpy as np
import pytest

from conftest import golden


@pytest.mark.parametrize("name", ["plate4", "plate10", "box12", "two_plates6"])
def test_conduction_operator_matches_reference(name):
    from form_factors_amd import ffk

    g = golden("conduction.npz")
    faces, meshes = g[f"{name}_faces"], g[f"{name}_meshes"]
    row, cols, G = ffk.conduction_graph(faces, meshes, g[f"{name}_k"], g[f"{name}_thickness"])
    n = len(meshes)
    T = g[f"{name}_T"].astype(np.float64)
    # flow_m = sum_k G_mk (T_k - T_m); the reference splits each pair's flow
    # into absorption on one side and emission on the other, so compare nets
    flow = np.zeros(n)
    for m in range(n):
        for j in range(row[m], row[m + 1]):
            flow[m] += G[j] * (T[cols[j]] - T[m])
    net_ref = g[f"{name}_absorption"].astype(np.float64) - g[f"{name}_emission"].astype(np.float64)
    scale = np.abs(g[f"{name}_absorption"]).max() + 1e-30
    assert np.abs(flow - net_ref).max() <= 1e-5 * scale, (name, np.abs(flow - net_ref).max(), scale)
    # the reference's split: per neighbour, inflow is absorbed, outflow emitted
    ab = np.zeros(n)
    em = np.zeros(n)
    for m in range(n):
        for j in range(row[m], row[m + 1]):
            f = G[j] * (T[cols[j]] - T[m])
            if f > 0:
                ab[m] += f
            else:
                em[m] -= f
    assert np.abs(ab - g[f"{name}_absorption"]).max() <= 1e-5 * scale
    assert np.abs(em - g[f"{name}_emission"]).max() <= 1e-5 * scale
    # symmetric operator, energy conserving: sum of flows is zero
    dense = np.zeros((n, n))
    for m in range(n):
        dense[m, cols[row[m]:row[m + 1]]] = G[row[m]:row[m + 1]]
    assert np.allclose(dense, dense.T)
    assert abs(flow.sum()) <= 1e-9 * np.abs(flow).sum() + 1e-12


def test_conduction_operator_grid_neighbours():
    # a 10 x 10 plate: interior cells have 4 edge neighbours (diagonals share a vertex only)
    from form_factors_amd import ffk, scenes

    faces = scenes.plane_grid_faces(10, 10, 10, 10)
    meshes = scenes.plane_grid_meshes(10, 10)
    row, cols, G = ffk.conduction_graph(faces, meshes, np.full(100, 237.0), np.full(100, 0.001))
    deg = np.diff(row)
    assert deg.min() == 2 and deg.max() == 4 and (deg == 4).sum() == 64
    # uniform cells: G = k t edge / (d_l + d_r) = 237 * 0.001 * 1 / (0.5 + 0.5)
    assert np.allclose(G, 0.237, rtol=1e-5)
