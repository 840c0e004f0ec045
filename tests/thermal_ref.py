"""All-FP64 restatement of the thermal step on the device (torch float64), the
reference for full-size transient runs: 20,000 meshes x 10,000 steps is out
of reach of the CPU restatement (a 3.2 GB FP64 GEMV per step) but takes
seconds on the GPU. It applies the same equations as fp64_run in
tests/test_gpu_thermal.py, whose agreement with it is tested at small size:

* SB_FF exchange q = F^T E with E_m = (eps_f + eps_r) sigma T^4 A
  (thermal_equation/form_factors.cpp:109-128), F promoted to FP64;
* conduction through the CSR operator (row, cols, G): flow_i = sum G (T_j - T_i);
* heat sources (thermal_equation/heat_source_cpu.cpp:48-61);
* Adams-Bashforth AB1..AB5 start-up with a 5-slot ring
  (thermal_solution/cpu_adams.cpp:136-182).
"""
import numpy as np

SIGMA = 5.670400e-8
U = np.array([[1.0, 0, 0, 0, 0], [3 / 2, -1 / 2, 0, 0, 0], [23 / 12, -4 / 3, 5 / 12, 0, 0],
              [55 / 24, -59 / 24, 37 / 24, -3 / 8, 0],
              [1901 / 720, -1387 / 360, 109 / 30, -637 / 360, 251 / 720]])


def fp64_run_device(F, prob, row, cols, G, T0, n_steps, dt):
    """F: N x N float32 device tensor (source-major, F[m, j]). Returns T after
    n_steps as a float64 numpy array."""
    import torch

    dev = F.device
    n = F.shape[0]
    d = lambda a: torch.as_tensor(np.asarray(a, np.float64), device=dev)  # noqa: E731
    F64 = F.to(torch.float64)
    A, Cap = d(prob["area"]), d(prob["C"])
    ef, er = d(prob["eps_front"]), d(prob["eps_rear"])
    src = torch.zeros(n, dtype=torch.float64, device=dev)
    src.index_add_(0, torch.as_tensor(np.asarray(prob["src_mesh"], np.int64), device=dev), d(prob["src_power"]))
    rows = torch.as_tensor(np.repeat(np.arange(n), np.diff(np.asarray(row))), dtype=torch.int64, device=dev)
    cols_t = torch.as_tensor(np.asarray(cols, np.int64), device=dev)
    G_t = d(G)
    E = torch.zeros((5, n), dtype=torch.float64, device=dev)
    T = d(T0).clone()
    for step in range(n_steps):
        power = SIGMA * T ** 4 * A
        em = ef * power + er * power
        ab = torch.mv(F64.t(), em)
        flow = torch.zeros(n, dtype=torch.float64, device=dev)
        flow.index_add_(0, rows, G_t * (T[cols_t] - T[rows]))
        E[step % 5] = (ab - em + flow + src) / Cap
        k = min(step, 4)
        acc = torch.zeros(n, dtype=torch.float64, device=dev)
        for l in range(k + 1):
            acc += U[k, l] * E[(step - l) % 5]
        T = T + dt * acc
    out = T.cpu().numpy()
    del F64
    return out
