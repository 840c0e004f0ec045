"""ffk_sbff_noalloc (the allocation-free SB_FF exchange) against ffk_sbff
through the C ABI: identical emission and absorption bits on two calls in a
row, the second with fewer meshes (the static scratch keeps the first call's
values past the second's n, which must not leak into it); no device memory
is taken by the calls; a size above the scratch limit is refused before any
launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(gpu, n, seed):
    import torch

    rng = np.random.default_rng(seed)
    n_rays = rng.integers(500, 5000, n).astype(np.float32)
    counts = rng.poisson(1.2, (n, n)).astype(np.float32)
    F = (counts * (np.float32(1.0) / n_rays[:, None])).astype(np.float32)
    Fd = torch.from_numpy(F).cuda()
    S, scale = gpu.pack_exchange(Fd, gpu.BF16, row_mul=n_rays)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()  # noqa: E731
    return dict(S=S, scale=scale, T=dev(rng.uniform(200, 450, n)), A=dev(rng.uniform(0.2, 2.0, n)),
                ef=dev(rng.uniform(0.1, 1, n)), er=dev(rng.uniform(0.1, 1, n)))


def _outputs(n):
    import torch

    return (torch.full((n,), 1.5, dtype=torch.float64, device="cuda"),
            torch.full((n,), 2.5, dtype=torch.float64, device="cuda"))


def _call(gpu, P, n, work):
    p = gpu._ptr
    em, ab = _outputs(n)
    args = (n, p(P["S"]), gpu.BF16, p(P["scale"]), p(P["T"]), p(P["A"]), p(P["ef"]), p(P["er"]), p(em), p(ab))
    if work is None:
        gpu.check(gpu.lib().ffk_sbff_noalloc(*args, gpu._stream()), "ffk_sbff_noalloc")
    else:
        gpu.check(gpu.lib().ffk_sbff(*args, p(work), gpu._stream()), "ffk_sbff")
    return em, ab


def test_noalloc_equals_sbff_on_two_calls(gpu):
    import torch

    from conftest import bits

    for n, seed in ((3000, 1), (700, 2)):
        print(f"sbff_noalloc n={n}", flush=True)
        P = _problem(gpu, n, seed)
        work = torch.empty(n, dtype=torch.float32, device="cuda")
        em0, ab0 = _call(gpu, P, n, work)
        em1, ab1 = _call(gpu, P, n, None)
        torch.cuda.synchronize()
        assert np.array_equal(bits(em1.cpu().numpy()), bits(em0.cpu().numpy()))
        assert np.array_equal(bits(ab1.cpu().numpy()), bits(ab0.cpu().numpy()))
        assert ab1.min().item() > 2.5  # accumulated onto the caller's arrays


def test_noalloc_takes_no_device_memory(gpu):
    import torch

    P = _problem(gpu, 1500, 3)
    _call(gpu, P, 1500, None)  # first call on this device: the scratch address is looked up
    em, ab = _outputs(1500)
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    p = gpu._ptr
    for _ in range(20):
        gpu.check(gpu.lib().ffk_sbff_noalloc(1500, p(P["S"]), gpu.BF16, p(P["scale"]), p(P["T"]), p(P["A"]),
                                             p(P["ef"]), p(P["er"]), p(em), p(ab), gpu._stream()),
                  "ffk_sbff_noalloc")
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0


def test_noalloc_refuses_sizes_beyond_the_scratch(gpu):
    limit = 1 << 20
    p = gpu._ptr
    rc = gpu.lib().ffk_sbff_noalloc(limit + 1, p(None), gpu.BF16, None, None, None, None, None, None, None,
                                    gpu._stream())
    assert rc == -4  # FFK_ERR_LIMIT, before any launch
    assert b"static scratch" in gpu.lib().ffk_last_error()
    rc = gpu.lib().ffk_sbff_noalloc(0, None, gpu.BF16, None, None, None, None, None, None, None, gpu._stream())
    assert rc == -1
