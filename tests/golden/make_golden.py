"""Write the golden vectors under tests/golden/ from the COMPILED REFERENCE.

Run in a container that has the reference tree (oracle/Makefile target `ref`
builds oracle/_ref/libffref.so from it). The fixtures are data only: inputs
and the reference's outputs for them. They pin the CPU oracle on machines
without the reference, and through it the HIP kernels.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from form_factors_amd import scenes  # noqa: E402
from oracle import oracle as O  # noqa: E402

_f = O._f
_i = O._i


def P(a, t=_f):
    return a.ctypes.data_as(t)


def save(name, **arrays):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **arrays)
    print(f"{name}: {os.path.getsize(path) / 1024:.1f} KiB")


def ref_areas(faces):
    R = O.ref()
    return np.array([R.ffref_triangle_area(P(np.ascontiguousarray(f))) for f in faces], np.float32)


def float_total(areas):
    t = np.float32(0)
    for a in areas:
        t = np.float32(t + np.float32(2) * a)
    return t


def weights_of(faces):
    a = ref_areas(faces)
    w = np.repeat(a, 2).astype(np.float32)
    return w, float_total(a)


def ref_malley(faces, n_rays, max_rays=None):
    R = O.ref()
    w, total = weights_of(faces)
    n = O.lib().or_malley_count(P(w), len(faces), n_rays, float(total))
    m = n if max_rays is None else min(n, max_rays)
    rays = np.zeros((m, 6), np.float32)
    hit = np.zeros(m, np.int32)
    got = R.ffref_malley(P(faces), len(faces), P(w), float(total), n_rays, P(rays), P(hit, _i), m)
    assert got == n, (got, n)
    return rays, hit, w, total


def gen_intersect(rng):
    R = O.ref()
    n = 6000
    tris = rng.uniform(-1, 1, (n, 9)).astype(np.float32)
    # a slice of degenerate and tiny faces (the 1e-8 threshold, math/triangle.cpp:38)
    k = n // 10
    s = np.float32(0.01) * rng.uniform(0.5, 1.5, k).astype(np.float32)
    tiny = np.zeros((k, 9), np.float32)
    tiny[:, 3] = s
    tiny[:, 7] = s
    tris[:k] = tiny
    tris[k:2 * k, 6:9] = tris[k:2 * k, 3:6]  # two equal vertices
    p = tris.reshape(-1, 3, 3)
    bary = rng.dirichlet([1, 1, 1], n).astype(np.float32)
    bary[::7] = rng.uniform(-0.2, 1.2, (len(bary[::7]), 3)).astype(np.float32)  # edges / outside
    target = np.einsum("ni,nij->nj", bary, p).astype(np.float32)
    origin = (target + rng.normal(0, 1, (n, 3))).astype(np.float32)
    end = (origin + (target - origin) * rng.uniform(0.1, 2, (n, 1))).astype(np.float32)
    rays = np.concatenate([origin, end], 1).astype(np.float32)
    rays[::11, 3:] = rays[::11, :3] + (rays[::11, 3:] - rays[::11, :3])[:, ::-1]  # other directions
    codes = np.zeros(n, np.int32)
    pts = np.zeros((n, 3), np.float32)
    for i in range(n):
        codes[i] = R.ffref_triangle_intersect(P(rays[i]), P(tris[i]), P(pts[i]))
    save("intersect.npz", rays=rays, tris=tris, codes=codes, points=pts)


def gen_rotate(rng):
    R = O.ref()
    to = []
    to += [(0, 0, 1), (0, 0, -1), (1, 0, 0), (0, 1, 0), (10, -1, 3), (0, 0, 2), (0, 0, -0.5), (1e-20, 0, 1)]
    for scale in (1e-4, 1e-3, 1e-2, 1, 10):
        to += list(rng.standard_normal((200, 3)) * scale)
    to = np.array(to, np.float32)
    frm = np.zeros_like(to)
    frm[:, 2] = 1
    frm[-100:] = rng.standard_normal((100, 3)).astype(np.float32)
    mats = np.zeros((len(to), 9), np.float32)
    vec = rng.standard_normal((len(to), 3)).astype(np.float32)
    mv = np.zeros((len(to), 3), np.float32)
    for i in range(len(to)):
        R.ffref_rotate_towards(P(frm[i]), P(to[i]), P(mats[i]))
        R.ffref_mat_vec(P(mats[i]), P(vec[i]), P(mv[i]))
    axis = np.zeros(9, np.float32)
    R.ffref_axis_rotation(np.float32(np.pi / 4), np.float32(np.pi / 4), np.float32(np.pi / 4), P(axis))
    save("rotate.npz", frm=frm, to=to, mats=mats, vec=vec, mat_vec=mv, axis_rotation_pi4=axis)


def gen_emitted(rng):
    R = O.ref()
    out = {}
    cases = [("plates10", scenes.parallel_plates(10)[0], [4_000_000, 40_000, 123_457]),
             ("sphere3", scenes.icosphere(3), [2_000_000, 40_000, 999]),
             ("confetti", scenes.confetti(500, rng), [1_000_000, 77_777])]
    for name, faces, ns in cases:
        w, total = weights_of(faces)
        out[f"{name}_weights"] = w
        out[f"{name}_total"] = np.float32(total)
        for n in ns:
            fr = np.zeros(len(faces), np.int32)
            rr = np.zeros(len(faces), np.int32)
            R.ffref_emitted_counts(P(w), len(faces), n, float(total), P(fr, _i), P(rr, _i))
            out[f"{name}_{n}_front"] = fr
            out[f"{name}_{n}_rear"] = rr
    save("emitted.npz", **out)


def cast_scene_list(rng):
    plates = scenes.parallel_plates(10)[0]
    sphere = scenes.icosphere(3)
    conf = np.concatenate([scenes.confetti(300, rng), scenes.cube_shell(3)]).astype(np.float32)
    return [("plates", plates, 40_000), ("sphere3", sphere, 16_000), ("confetti", conf, 16_000)]


def gen_casts(rng):
    R = O.ref()
    for name, faces, n_rays in cast_scene_list(rng):
        rays, hit_emit, w, total = ref_malley(faces, n_rays)
        nai_i, nai_p = O.ref_cast(0, faces, rays)
        zg_i, zg_p = O.ref_cast(1, faces, rays)
        assert np.array_equal(nai_i, hit_emit)
        same_pts = np.array_equal(nai_p[nai_i >= 0].view(np.uint32), zg_p[nai_i >= 0].view(np.uint32))
        save(f"cast_{name}.npz", faces=faces, rays=rays, weights=w, total=np.float32(total),
             n_rays=np.int64(n_rays), naive_idx=nai_i, naive_pts=nai_p, zgrid_idx=zg_i,
             zgrid_pts_equal=np.bool_(same_pts), zgrid_pts=(np.zeros((0, 3), np.float32) if same_pts else zg_p))
        # the 2-D grid the zgrid caster builds for this scene
        nx, ny = C.c_int(), C.c_int()
        base, side, size = (np.zeros(3, np.float32) for _ in range(3))
        R.ffref_zgrid_info(P(faces), len(faces), C.byref(nx), C.byref(ny), P(base), P(side), P(size),
                           None, None, None, 0)
        counts = np.zeros(nx.value * ny.value, np.int32)
        offsets = np.zeros_like(counts)
        refs = np.zeros(4 * len(faces) * 64, np.int32)
        nref = R.ffref_zgrid_info(P(faces), len(faces), C.byref(nx), C.byref(ny), P(base), P(side), P(size),
                                  P(counts, _i), P(offsets, _i), P(refs, _i), len(refs))
        save(f"zgrid_{name}.npz", nx=np.int32(nx.value), ny=np.int32(ny.value), base=base, side=side,
             size=size, counts=counts, offsets=offsets, refs=refs[:nref])


def gen_grid_kat(rng):
    """The cases of test/grid.cpp:79-183 plus random walks, as the reference
    visits them (ordered cell lists)."""
    R = O.ref()
    base = np.zeros(3, np.float32)
    size = np.array([1.5, 1.5, 0], np.float32)
    rays = [
        (1.25, 1.25, 0, 1.5, 1.5, 0, 0), (0, .7, 0, 2, .8, 0, 0), (.7, .1, 0, .7, .2, 0, 0),
        (.3, .4, 0, .4, .5, 0, 0), (1.4, 1.49, 0, 1.3, 1.4, 0, 0), (1.1, 0, 0, 1, .1, 0, 0),
        (-.5, 2, 0, .25, 1.5, 0, 0), (-.5, 2, 0, 0, 1.25, 0, 0), (-.5, -.5, 0, 0, .25, 0, 0),
        (-.5, -.5, 0, .25, 0, 0, 0), (2, 2, 0, 1.25, 1.5, 0, 0), (2, 2, 0, 1.5, 1.25, 0, 0),
        (2, -.5, 0, 1.5, .25, 0, 0), (2, -.5, 0, 1.25, 0, 0, 0),
        (.75, 1.25, 0, .25, .25, 0, 1), (.55, .95, 0, .45, .55, 0, 1), (.55, .95, 0, .95, .55, 0, 1)]
    for _ in range(200):
        o = rng.uniform(-0.5, 2.0, 2)
        d = rng.uniform(-0.5, 2.0, 2)
        rays.append((o[0], o[1], 0, d[0], d[1], 0, int(rng.integers(0, 2))))
    rays = np.array(rays, np.float32)
    cells = np.full((len(rays), 64, 2), -1, np.int32)
    counts = np.zeros(len(rays), np.int32)
    for i, r in enumerate(rays):
        buf = np.zeros(128, np.int32)
        n = R.ffref_grid_traverse(P(base), P(size), 3, 3, P(np.ascontiguousarray(r[:6])), int(r[6]), P(buf, _i), 64)
        counts[i] = n
        cells[i, :n] = buf[:2 * n].reshape(-1, 2)
    tris = [(.7, .7, 0, .8, .8, 0, .8, .7, 0), (.75, 1.25, 0, .25, .25, 0, 1.25, .25, 0),
            (.25, .75, 0, 1.25, .76, 0, 1.25, .74, 0)]
    for _ in range(100):
        tris.append(tuple(rng.uniform(0, 1.5, 9) * np.array([1, 1, 0] * 3)))
    tris = np.array(tris, np.float32)
    rcells = np.full((len(tris), 16, 2), -1, np.int32)
    rcounts = np.zeros(len(tris), np.int32)
    for i, t in enumerate(tris):
        buf = np.zeros(32, np.int32)
        n = R.ffref_grid_rasterize(P(base), P(size), 3, 3, P(np.ascontiguousarray(t)), P(buf, _i), 16)
        rcounts[i] = n
        rcells[i, :n] = buf[:2 * n].reshape(-1, 2)
    seg = np.array([[1, 1, 0, 2, 1, 0], [0, 2, 0, 1, .9, 0]], np.float32)
    ray = np.array([[.5, 2, 0, 1, 1.5, 0], [0, 2, 0, 1, .9, 0]], np.float32)
    sv = np.zeros(2, np.float32)
    hits = np.zeros(2, np.int32)
    for i in range(2):
        s = C.c_float()
        hits[i] = R.ffref_ray_intersect_segment(P(ray[i]), P(seg[i]), C.byref(s))
        sv[i] = s.value
    save("grid_kat.npz", base=base, size=size, rays=rays, cells=cells, counts=counts, tris=tris,
         rcells=rcells, rcounts=rcounts, seg_p=ray, seg_q=seg, seg_hit=hits, seg_s=sv)


def gen_ff(rng):
    R = O.ref()
    axis = np.zeros(9, np.float32)
    R.ffref_axis_rotation(np.float32(np.pi / 4), np.float32(np.pi / 4), np.float32(np.pi / 4), P(axis))
    axis = axis.reshape(3, 3)
    cases = [("parallel_c1", *scenes.ff_parallel_planes(1.0), 40_000, 0),
             ("parallel_c1_rot", *scenes.ff_parallel_planes(1.0, axis), 40_000, 0),
             ("perpendicular", *scenes.ff_perpendicular_planes(), 40_000, 0),
             ("box", *scenes.ff_box(), 40_000, 0),
             ("plates5", *scenes.parallel_plates(5), 200_000, 1),
             ("sphere2", scenes.icosphere(2), scenes.per_face_meshes(320), 160_000, 1)]
    out = {}
    for name, faces, meshes, n_rays, caster in cases:
        F = np.zeros((len(meshes), len(meshes)), np.float32)
        rc = R.ffref_ff_calculate(P(faces), len(faces), P(meshes, _i), len(meshes), n_rays, caster, P(F))
        assert rc == 0, (name, rc)
        out[f"{name}_faces"] = faces
        out[f"{name}_meshes"] = meshes
        out[f"{name}_n_rays"] = np.int64(n_rays)
        out[f"{name}_caster"] = np.int32(caster)
        out[f"{name}_F"] = F
    # error codes of prepare(): zero-area scene, no meshes
    z = np.zeros((2, 9), np.float32)
    out["code_zero_area"] = np.int32(R.ffref_ff_prepare_code(P(z), 2, P(np.array([[0, 2, 0]], np.int32), _i), 1))
    out["code_no_meshes"] = np.int32(R.ffref_ff_prepare_code(P(cases[0][1]), 4, P(np.zeros(3, np.int32), _i), 0))
    out["axis_rotation"] = axis
    save("ff.npz", **out)


def gen_adams(rng):
    R = O.ref()
    out = {}
    face = np.array([[0, 0, 0, 1.41421, 0, 0, 0, 1.41421, 0]], np.float32)
    for name, shell in [("unit", (1, 1, 1)), ("thick2", (1, 1, 2)), ("rho2", (2, 1, 1)), ("c2", (1, 2, 1))]:
        T = np.zeros((11, 1), np.float32)
        rc = R.ffref_adams(P(face), 1, P(np.array([[0, 1, 0]], np.int32), _i), 1,
                           P(np.array([shell], np.float32)), P(np.array([300], np.float32)),
                           P(np.array([2], np.float32)), P(np.array([1], np.float32)), 1.0, 11, P(T))
        assert rc == 0
        out[f"{name}_shell"] = np.array(shell, np.float32)
        out[f"{name}_T"] = T
    n = 8
    faces = scenes.plane_grid_faces(1, 1, 2, 2)
    meshes = scenes.plane_grid_meshes(2, 2)
    shell = rng.uniform(0.5, 2, (4, 3)).astype(np.float32)
    T0 = rng.uniform(250, 350, 4).astype(np.float32)
    e = rng.uniform(0, 3, 4).astype(np.float32)
    a = rng.uniform(0, 3, 4).astype(np.float32)
    T = np.zeros((20, 4), np.float32)
    rc = R.ffref_adams(P(faces), len(faces), P(meshes, _i), 4, P(shell), P(T0), P(e), P(a), 0.1, 20, P(T))
    assert rc == 0
    out.update(multi_faces=faces, multi_meshes=meshes, multi_shell=shell, multi_T0=T0, multi_e=e, multi_a=a,
               multi_dt=np.float32(0.1), multi_T=T)
    del n
    save("adams.npz", **out)


def gen_radiance(rng):
    R = O.ref()
    faces, meshes = scenes.ff_box()
    meshes = scenes.per_face_meshes(12)
    eps = rng.uniform(0.2, 1.0, (12, 2)).astype(np.float32)
    T = rng.uniform(250, 400, 12).astype(np.float32)
    em = np.zeros(12, np.float32)
    ab = np.zeros(12, np.float32)
    rc = R.ffref_radiance(P(faces), 12, P(meshes, _i), 12, P(eps), P(T), 20_000, P(em), P(ab))
    assert rc == 0
    save("radiance.npz", faces=faces, meshes=meshes, eps=eps, T=T, n_rays=np.int32(20_000), emission=em,
         absorption=ab)
    # multi-face meshes listed out of face order (the walk attribution), on an
    # 80-triangle icosphere with inward normals
    faces = scenes.icosphere(1, inward=True)
    counts = [3, 1, 2, 4, 2, 3, 1, 4] * 4  # 80 faces
    firsts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    r2 = np.random.default_rng(7)  # own stream: the fixtures after this one keep their inputs
    order = r2.permutation(len(counts))
    meshes = np.stack([firsts[order], np.array(counts)[order], np.zeros(len(counts), int)], 1).astype(np.int32)
    nm = len(meshes)
    eps = r2.uniform(0.2, 1.0, (nm, 2)).astype(np.float32)
    T = r2.uniform(250, 400, nm).astype(np.float32)
    em = np.zeros(nm, np.float32)
    ab = np.zeros(nm, np.float32)
    rc = R.ffref_radiance(P(faces), len(faces), P(meshes, _i), nm, P(eps), P(T), 30_000, P(em), P(ab))
    assert rc == 0
    save("radiance_multi.npz", faces=faces, meshes=meshes, eps=eps, T=T, n_rays=np.int32(30_000), emission=em,
         absorption=ab)


def gen_sun():
    """parallel_rays_cpu (thermal) and triangles_bsphere; own random stream."""
    R = O.ref()
    r2 = np.random.default_rng(8)
    out = {}
    for k, (faces, _) in enumerate([scenes.confetti_scene(150, cells=2, seed=3), scenes.ff_box(),
                                    (scenes.icosphere(2), None)]):
        faces = np.ascontiguousarray(faces, np.float32)
        c = np.zeros(3, np.float32)
        r = np.zeros(1, np.float32)
        R.ffref_bsphere(P(faces), len(faces), P(c), P(r))
        out[f"bs{k}_faces"], out[f"bs{k}_center"], out[f"bs{k}_radius"] = faces, c, r
    faces, meshes = scenes.confetti_scene(150, cells=2, seed=3)
    faces = np.ascontiguousarray(faces, np.float32)
    nm = len(meshes)
    absorb = r2.uniform(0.1, 1.0, (nm, 2)).astype(np.float32)
    d = r2.normal(size=3)
    d = (d / np.linalg.norm(d)).astype(np.float32)
    ab = np.zeros(nm, np.float32)
    rc = R.ffref_sun(P(faces), len(faces), P(meshes, _i), nm, P(absorb), np.float32(1361.0), P(d), 20_000, P(ab))
    assert rc == 0
    out.update(faces=faces, meshes=meshes, absorb=absorb, power=np.float32(1361.0), direction=d,
               n_rays=np.int32(20_000), absorption=ab)
    mesh_idx = np.array([0, 3, 3, 7], np.int32)
    power = np.array([5.0, -2.0, 1.5, -0.25], np.float32)
    em = np.linspace(1, 2, 8).astype(np.float32)
    ab = np.linspace(3, 4, 8).astype(np.float32)
    em0, ab0 = em.copy(), ab.copy()
    assert R.ffref_heat_source(P(mesh_idx, _i), P(power), 4, P(em), P(ab)) == 0
    out.update(hs_mesh=mesh_idx, hs_power=power, hs_emission0=em0, hs_absorption0=ab0, hs_emission=em,
               hs_absorption=ab)
    save("sun.npz", **out)


def gen_parallel(rng):
    R = O.ref()
    faces = scenes.icosphere(1)
    d = np.array([0.3, -0.5, -0.8], np.float32)
    c = np.array([0.1, 0.2, 0.3], np.float32)
    rays = np.zeros((5000, 6), np.float32)
    R.ffref_parallel_rays(P(d), P(c), 2.5, 1.3, 5000, P(faces), len(faces), P(rays))
    save("parallel.npz", direction=d, center=c, distance=np.float32(2.5), radius=np.float32(1.3), rays=rays)


def gen_conduction(rng):
    """conductive_equation (thermal_equation/conductive_cpu.cpp) on scenes
    whose meshes share edges: the reference's emission/absorption arrays."""
    R = O.ref()
    R.ffref_conduction.restype = C.c_int
    R.ffref_conduction.argtypes = [_f, C.c_int, _i, C.c_int, _f, _f, _f, _f, _f]
    out = {}
    box, _ = scenes.ff_box()
    plate = scenes.plane_grid_faces(10, 10, 10, 10)  # cells of 1 unit, fuzzy vertex matches
    two = scenes.two_plates_thermal(cells=6, gap=5.0, size=10.0)[0]
    cases = [("plate4", scenes.plane_grid_faces(1, 1, 4, 4), scenes.plane_grid_meshes(4, 4)),
             ("plate10", plate, scenes.plane_grid_meshes(10, 10)),
             ("box12", box, scenes.per_face_meshes(12)),
             ("two_plates6", two, scenes.two_plates_thermal(cells=6, gap=5.0, size=10.0)[1])]
    for name, faces, meshes in cases:
        faces = np.ascontiguousarray(faces, np.float32)
        meshes = np.ascontiguousarray(meshes, np.int32)
        n = len(meshes)
        k = rng.uniform(50, 300, n).astype(np.float32)
        th = rng.uniform(0.001, 0.01, n).astype(np.float32)
        T = rng.uniform(250, 400, n).astype(np.float32)
        em = np.zeros(n, np.float32)
        ab = np.zeros(n, np.float32)
        rc = R.ffref_conduction(P(faces), len(faces), P(meshes, _i), n, P(k), P(th), P(T), P(em), P(ab))
        assert rc == 0, (name, rc)
        out.update({f"{name}_faces": faces, f"{name}_meshes": meshes, f"{name}_k": k, f"{name}_thickness": th,
                    f"{name}_T": T, f"{name}_emission": em, f"{name}_absorption": ab})
    save("conduction.npz", **out)


def gen_sbff():
    """form_factors_equation (thermal_equation/form_factors.cpp:77-128): the
    reference computes F in set_scene, then one calculate() adds emission and
    the absorption q = F^T E. Meshes of unequal area (so F is not symmetric)
    with unequal emissivities and temperatures; own random stream."""
    R = O.ref()
    R.ffref_sbff.restype = C.c_int
    R.ffref_sbff.argtypes = [_f, C.c_int, _i, C.c_int, _f, _f, C.c_int, _f, _f, _f, _f]
    r2 = np.random.default_rng(11)
    big = scenes.plane_grid_faces(1, 1, 3, 3)
    small = scenes.translate(scenes.plane_grid_faces(0.6, 0.6, 2, 2), (0.15, 0.3, 0.7))
    plates = np.concatenate([big, small]).astype(np.float32)
    plate_meshes = np.concatenate([scenes.plane_grid_meshes(3, 3),
                                   scenes.plane_grid_meshes(2, 2) + np.array([18, 0, 0], np.int32)])
    sph = scenes.icosphere(1, inward=True)
    counts = [1, 3, 2, 4, 2, 1, 3, 4] * 4  # 80 faces in 32 meshes of 1-4 faces, listed out of face order
    firsts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    order = r2.permutation(len(counts))
    sph_meshes = np.stack([firsts[order], np.array(counts)[order], np.zeros(len(counts), int)], 1)
    out = {}
    for name, faces, meshes, n_rays in (("plates", plates, plate_meshes, 40_000),
                                        ("sphere", sph, sph_meshes, 30_000)):
        faces = np.ascontiguousarray(faces, np.float32)
        meshes = np.ascontiguousarray(meshes, np.int32)
        nm = len(meshes)
        eps = r2.uniform(0.1, 1.0, (nm, 2)).astype(np.float32)
        T = r2.uniform(200, 450, nm).astype(np.float32)
        F = np.zeros((nm, nm), np.float32)
        A = np.zeros(nm, np.float32)
        em = np.zeros(nm, np.float32)
        ab = np.zeros(nm, np.float32)
        rc = R.ffref_sbff(P(faces), len(faces), P(meshes, _i), nm, P(eps), P(T), n_rays, P(F), P(A), P(em), P(ab))
        assert rc == 0, (name, rc)
        # F_out is the F the equation used: its float accumulation in the
        # equation's own order reproduces the reference's absorption bits
        sigma = np.float32(5.670400e-8)
        acc = np.zeros(nm, np.float32)
        for m in range(nm):
            power = np.float32(np.float32(sigma * np.float32(T[m] * T[m] * T[m] * T[m])) * A[m])
            e = np.float32(np.float32(eps[m, 0] * power) + np.float32(eps[m, 1] * power))
            acc = (acc + (e * F[m]).astype(np.float32)).astype(np.float32)
        assert np.array_equal(acc.view(np.uint32), ab.view(np.uint32)), name
        out.update({f"{name}_faces": faces, f"{name}_meshes": meshes, f"{name}_eps": eps, f"{name}_T": T,
                    f"{name}_n_rays": np.int32(n_rays), f"{name}_F": F, f"{name}_area": A,
                    f"{name}_emission": em, f"{name}_absorption": ab})
    save("sbff.npz", **out)


def main(only=()):
    if not O.ref_available():
        O.build()
    if only:  # python tests/golden/make_golden.py sbff ...: generators with their own streams
        for name in only:
            globals()["gen_" + name]()
        return
    rng = np.random.default_rng(20261203)
    gen_intersect(rng)
    gen_rotate(rng)
    gen_emitted(rng)
    gen_casts(rng)
    gen_grid_kat(rng)
    gen_ff(rng)
    gen_adams(rng)
    gen_radiance(rng)
    gen_parallel(rng)
    gen_conduction(rng)
    gen_sun()
    gen_sbff()


if __name__ == "__main__":
    main(sys.argv[1:])
