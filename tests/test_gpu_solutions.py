"""The reference's own thermal-solution scenarios (test/conductive_solution.cpp,
test/radiance_solution.cpp) through the plugin ABI of ff_plugin.h: the HIP
conduction and radiance equations under the FP64 HIP Adams solution, with
the reference's scenes, materials, steps and acceptance bounds. Each runs
with the solution's device-resident step and with the equations on host
arrays; the two give the same temperatures bit for bit."""
import ctypes as C

import numpy as np
import pytest

from test_gpu_plugin import subject_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def P(gpu):
    from form_factors_amd import plugin

    plugin.lib()
    return plugin


def box():
    """subject::box() (subject/objects.cpp:29-50): the unit box's twelve
    faces with their normals unified to the first face's, which points out
    of the box (triangle_flip_normal swaps the second and third point)."""
    from form_factors_amd import scenes

    f = scenes.ff_box()[0].reshape(-1, 3, 3).copy()
    d = f.astype(np.float64)
    n = np.cross(d[:, 1] - d[:, 0], d[:, 2] - d[:, 0])
    inward = (n * (d.mean(1) - 0.5)).sum(1) < 0
    f[inward] = f[inward][:, [0, 2, 1]]
    return f.reshape(-1, 9)


def material_al(P, thickness):
    """subject::material_Al (subject/objects.cpp:166-196): aluminium shell,
    black optics with emissivity 0.09."""
    m = P.black_body()
    m.density, m.heat_capacity, m.thermal_conductivity, m.thickness = 2.6989e3, 903.0, 237.0, thickness
    m.front[4] = m.rear[4] = 0.09
    m.name = b"alliminuim"
    return m


class Solution:
    """A HIP Adams solution over HIP equations, as the reference's fixtures
    build theirs (thermal_solution::system_create over the equations)."""

    def __init__(self, P, faces, meshes, materials, T0, equations, device_loop, n_rays=1000,
                 caster_type=None):
        self.P, L = P, P.lib()
        self.rc = L.ffp_ray_caster_create(caster_type or P.RAY_CASTER_GRID_HIP)
        self.em = L.ffp_emission_create(P.EMISSION_MALLEY_HIP, self.rc)
        assert self.rc and self.em
        L.ffp_emission_set_seed(self.em, 7, 0)
        self.eqs = []
        for kind in equations:
            if kind == "radiance":
                e = L.ffp_thermal_equation_create(P.THERMAL_EQUATION_RADIANCE_HIP, self.em, n_rays, 0)
            else:
                e = L.ffp_thermal_equation_create(P.THERMAL_EQUATION_CONDUCTIVE_HIP, None, 0, 0)
            assert e
            self.eqs.append(e)
        arr = (C.c_void_p * len(self.eqs))(*self.eqs)
        self.sol = L.ffp_thermal_solution_create(P.THERMAL_SOLUTION_HIP_ADAMS, len(self.eqs), arr)
        assert self.sol and L.ffp_thermal_solution_set_device_loop(self.sol, int(device_loop)) == 0
        self.sc, self.keep = subject_scene(P, faces, meshes, materials)
        self.T = (C.c_float * len(T0))(*[float(t) for t in T0])
        assert L.ffp_thermal_solution_set_scene(self.sol, C.byref(self.sc), self.T) == 0
        self.task = P.TsTask(0, 0.0, C.cast(self.T, C.POINTER(C.c_float)))

    def step(self, dt):
        self.task.time_delta = dt
        assert self.P.lib().ffp_thermal_solution_calculate(self.sol, C.byref(self.task)) == 0
        return np.frombuffer(self.T, np.float32).copy()

    def close(self):
        L = self.P.lib()
        L.ffp_thermal_solution_free(self.sol)
        for e in self.eqs:
            L.ffp_thermal_equation_free(e)
        L.ffp_emission_free(self.em)
        L.ffp_ray_caster_free(self.rc)


def both(run):
    """run(device_loop) -> list of temperature frames, for both step forms,
    which must agree bit for bit."""
    a, b = run(True), run(False)
    assert len(a) == len(b)
    assert all(np.array_equal(x.view(np.uint32), y.view(np.uint32)) for x, y in zip(a, b))
    return a


def box_meshes():
    return np.array([[2 * m, 2, 0] for m in range(6)], np.int32)  # six sides of two faces


def test_conductive_box_preserves_energy(P):
    """ConductiveSolution.BoxPreserveEnergy: the box's six sides (aluminium,
    0.02 thick), one at 300 K and five at 0 K, 50 steps of 1 s: the sum of
    the temperatures (equal heat capacities) stays within 0.1 %."""
    def run(dev):
        s = Solution(P, box(), box_meshes(), [material_al(P, 0.02)], [300, 0, 0, 0, 0, 0], ["conduction"], dev)
        frames = [s.step(1.0) for _ in range(50)]
        s.close()
        return frames
    frames = both(run)
    ratio = frames[-1].astype(np.float64).sum() / 300.0
    print(f"conductive box: energy ratio {ratio:.6f}, T {frames[-1]}", flush=True)
    assert abs(ratio - 1.0) < 1e-3


def test_conductive_box_reaches_balance(P):
    """ConductiveSolution.ReachBalance: steps of 30 s until the heated side
    and the opposite one differ by at most 1 K, or 500 steps: energy kept
    within 0.1 % (the reference's bounds)."""
    def run(dev):
        s = Solution(P, box(), box_meshes(), [material_al(P, 0.02)], [300, 0, 0, 0, 0, 0], ["conduction"], dev)
        frames = [s.step(30.0)]
        while s.task.n_step < 500 and abs(frames[-1][5] - frames[-1][0]) > 1:
            frames.append(s.step(30.0))
        s.close()
        return frames
    frames = both(run)
    ratio = frames[-1].astype(np.float64).sum() / 300.0
    print(f"conductive box balance: {len(frames)} steps, |T5 - T0| {abs(frames[-1][5] - frames[-1][0]):.3f} K, "
          f"energy ratio {ratio:.6f}", flush=True)
    assert len(frames) <= 500
    assert abs(ratio - 1.0) < 1e-3


def radiance_material(P):
    m = P.black_body()  # radiance_solution.cpp: black body whose front (outer) side does not emit
    m.front[4] = 0.0
    return m


@pytest.mark.parametrize("caster", ["grid", "naive"])
def test_radiance_box_preserves_energy(P, caster):
    """RadianceSolution.BoxPreserveEnergy (the direct method, 1000 rays per
    step): one black box at 20 K radiating from its inner side onto itself,
    50 steps of 0.1 s: every ray lands on the box, so T stays within 0.1 %."""
    ct = P.RAY_CASTER_GRID_HIP if caster == "grid" else P.RAY_CASTER_NAIVE_HIP

    def run(dev):
        s = Solution(P, box(), np.array([[0, 12, 0]], np.int32), [radiance_material(P)], [20.0], ["radiance"], dev,
                     caster_type=ct)
        frames = [s.step(0.1) for _ in range(50)]
        s.close()
        return frames
    frames = both(run)
    print(f"radiance box ({caster}): T {frames[-1][0]:.6f} K", flush=True)
    assert abs(frames[-1][0] / 20.0 - 1.0) < 1e-3


def nested_boxes():
    """radiance_solution.cpp's fixture: the box, and a copy scaled by 1/3
    about the box's centre with its normals flipped."""
    outer = box().reshape(-1, 3, 3)
    factor = np.float32(1.0 / 3)
    shift = np.float32(0.5) - factor / np.float32(2)
    inner = (outer * factor + shift).astype(np.float32)[:, [0, 2, 1]]
    return np.concatenate([outer, inner]).reshape(-1, 9)


@pytest.mark.parametrize("caster", ["grid", "naive"])
def test_radiance_nested_boxes_preserve_energy(P, caster):
    """RadianceSolution.NestedBoxesPreserveEnergy: outer box at 20 K, inner
    box (a ninth of the area, so a ninth of the heat capacity) at 300 K, 100
    steps of 0.1 s: 9 T_outer + T_inner stays within 0.1 %."""
    ct = P.RAY_CASTER_GRID_HIP if caster == "grid" else P.RAY_CASTER_NAIVE_HIP
    meshes = np.array([[0, 12, 0], [12, 12, 0]], np.int32)

    def run(dev):
        s = Solution(P, nested_boxes(), meshes, [radiance_material(P)], [20.0, 300.0], ["radiance"], dev,
                     caster_type=ct)
        frames = [s.step(0.1) for _ in range(100)]
        s.close()
        return frames
    frames = both(run)
    e0 = 20.0 * 9 + 300.0
    e = float(frames[-1][0]) * 9 + float(frames[-1][1])
    print(f"nested boxes ({caster}): T {frames[-1]}, energy ratio {e / e0:.6f}", flush=True)
    assert frames[-1][1] < 300.0 and frames[-1][0] > 20.0  # heat moved from the inner box to the outer one
    assert abs(e / e0 - 1.0) < 1e-3


@pytest.mark.parametrize("dtype", [0, 1, 2])
def test_sbff_conduction_sources_device_equals_host(P, dtype):
    """A C5-shaped solution through the plugin ABI: SB_FF exchange (F in
    FP32, BF16 or FP16 on the device) with conduction and a heat source and
    sink, on two facing 4 x 4-cell plates (aluminium shell, emissivity 0.8),
    30 steps of 1 s. The device-resident step equals the host-array step bit
    for bit; both plates radiate to open space, the hot one faster, and the
    source and the sink warm and cool their own meshes."""
    from form_factors_amd import scenes

    L = P.lib()
    faces, meshes = scenes.two_plates_thermal(cells=4, gap=0.5)
    n = len(meshes)
    mat = material_al(P, 0.001)
    mat.front[4] = mat.rear[4] = 0.8
    sm = meshes.copy()
    sm[:, 2] = 0
    hs = (P.HeatSource * 2)(P.HeatSource(0, 50.0), P.HeatSource(n - 1, -50.0))

    def run(dev):
        rc = L.ffp_ray_caster_create(P.RAY_CASTER_GRID_HIP)
        em = L.ffp_emission_create(P.EMISSION_MALLEY_HIP, rc)
        L.ffp_emission_set_seed(em, 5, 0)
        ff = L.ffp_form_factors_create(P.FORM_FACTORS_HIP, em)
        eqs = [L.ffp_thermal_equation_create(P.THERMAL_EQUATION_SB_FF_HIP, ff, 200_000, dtype),
               L.ffp_thermal_equation_create(P.THERMAL_EQUATION_CONDUCTIVE_HIP, None, 0, 0),
               L.ffp_heat_source_create(2, C.cast(hs, C.c_void_p))]
        assert all(eqs)
        arr = (C.c_void_p * 3)(*eqs)
        sol = L.ffp_thermal_solution_create(P.THERMAL_SOLUTION_HIP_ADAMS, 3, arr)
        assert sol and L.ffp_thermal_solution_set_device_loop(sol, int(dev)) == 0
        sc, keep = subject_scene(P, faces, sm, [mat])
        T = (C.c_float * n)(*([400.0] * (n // 2) + [300.0] * (n // 2)))
        assert L.ffp_thermal_solution_set_scene(sol, C.byref(sc), T) == 0
        task = P.TsTask(0, 1.0, C.cast(T, C.POINTER(C.c_float)))
        frames = []
        for _ in range(30):
            assert L.ffp_thermal_solution_calculate(sol, C.byref(task)) == 0
            frames.append(np.frombuffer(T, np.float32).copy())
        L.ffp_thermal_solution_free(sol)
        for e in eqs:
            L.ffp_thermal_equation_free(e)
        L.ffp_form_factors_free(ff)
        L.ffp_emission_free(em)
        L.ffp_ray_caster_free(rc)
        return frames

    frames = both(run)
    T = frames[-1]
    print(f"SB_FF + conduction + sources (dtype {dtype}): hot plate {T[:n // 2].mean():.3f} K, "
          f"cold plate {T[n // 2:].mean():.3f} K, source mesh {T[0]:.3f} K, sink mesh {T[-1]:.3f} K", flush=True)
    assert 400.0 - T[1:n // 2].mean() > 300.0 - T[n // 2:-1].mean() > 0
    assert T[0] > T[1:n // 2].mean() and T[-1] < T[n // 2:-1].mean()
