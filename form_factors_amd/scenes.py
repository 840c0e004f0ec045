"""Synthetic scenes for the form-factor workloads (host-side numpy).

Faces are float32 arrays of shape (n, 9) (p0, p1, p2); meshes are int32
(first_idx, n_faces, pad) triples, as in form_factors/system.h:33-38.
"""
from __future__ import annotations

import numpy as np

f32 = np.float32


def plane_grid_faces(width, height, cells_x, cells_y):
    """subject/objects.cpp:52-77 in float32: cell (i, j) -> faces
    (A, B, D) and (C, D, B) with +z normals."""
    width, height = f32(width), f32(height)
    step_x = f32(width / f32(cells_x))
    step_y = f32(height / f32(cells_y))
    faces = np.zeros((cells_x * cells_y * 2, 9), f32)
    for i in range(cells_x):
        for j in range(cells_y):
            x = f32(step_x * f32(i))
            y = f32(step_y * f32(j))
            A = (x, y, f32(0))
            B = (f32(x + step_x), y, f32(0))
            C = (f32(x + step_x), f32(y + step_y), f32(0))
            D = (x, f32(y + step_y), f32(0))
            k = (i * cells_y + j) * 2
            faces[k] = A + B + D
            faces[k + 1] = C + D + B
    return faces


def plane_grid_meshes(cells_x, cells_y):
    """subject/objects.cpp:123-138: one mesh of 2 faces per cell."""
    n = cells_x * cells_y
    idx = np.arange(n, dtype=np.int32)
    return np.stack([idx * 2, np.full(n, 2, np.int32), np.zeros(n, np.int32)], 1)


def per_face_meshes(n_faces):
    idx = np.arange(n_faces, dtype=np.int32)
    return np.stack([idx, np.ones(n_faces, np.int32), np.zeros(n_faces, np.int32)], 1)


def translate(faces, d):
    out = np.array(faces, f32).reshape(-1, 3, 3)
    out = out + np.asarray(d, f32)
    return out.reshape(-1, 9).astype(f32)


def parallel_plates(cells=10, gap=1.0):
    """Two unit squares, each cells x cells x 2 triangles, facing each other
    at distance `gap` (both with +z normals); one mesh per triangle."""
    a = plane_grid_faces(1, 1, cells, cells)
    b = translate(a, (0, 0, gap))
    faces = np.concatenate([a, b]).astype(f32)
    return faces, per_face_meshes(len(faces))


def icosphere(subdiv=5, radius=1.0, inward=True):
    """Unit icosphere: 20 * 4^subdiv triangles, vertices on the sphere."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    tris = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
            (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
            (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                p = verts[a] + verts[b]
                verts.append(p / np.linalg.norm(p))
                cache[key] = len(verts) - 1
            return cache[key]

        nt = []
        for a, b, c in tris:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nt += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        tris = nt
    V = np.array(verts) * radius
    T = np.array(tris)
    faces = V[T].reshape(-1, 9)
    # orient: outward winding has normal . centroid > 0
    p = faces.reshape(-1, 3, 3)
    n = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    c = p.mean(1)
    flip = (np.einsum("ij,ij->i", n, c) > 0) == inward
    q = p.copy()
    q[flip, 1], q[flip, 2] = p[flip, 2], p[flip, 1]
    return q.reshape(-1, 9).astype(f32)


def random_rotations(rng, n):
    """Uniform random rotations (QR of Gaussian matrices)."""
    A = rng.standard_normal((n, 3, 3))
    Q, R = np.linalg.qr(A)
    Q = Q * np.sign(np.diagonal(R, axis1=1, axis2=2))[:, None, :]
    det = np.linalg.det(Q)
    Q[det < 0, :, 0] *= -1
    return Q


def confetti(n, rng, lo=0.0, hi=1.0, edge_lo=0.02, edge_hi=0.1):
    """n triangles, centroids uniform in [lo,hi]^3, side lengths uniform in
    [edge_lo, edge_hi] (resampled until they form a triangle), orientation
    uniform on the sphere."""
    out = np.zeros((n, 3, 3))
    k = 0
    while k < n:
        s = rng.uniform(edge_lo, edge_hi, (2 * (n - k) + 8, 3))
        ok = (s[:, 0] < s[:, 1] + s[:, 2]) & (s[:, 1] < s[:, 0] + s[:, 2]) & (s[:, 2] < s[:, 0] + s[:, 1])
        s = s[ok][: n - k]
        a, b, c = s[:, 0], s[:, 1], s[:, 2]
        # p0 = 0, p1 = (a, 0), p2 from the law of cosines
        x = (a * a + c * c - b * b) / (2 * a)
        y = np.sqrt(np.maximum(c * c - x * x, 0.0))
        tri = np.zeros((len(s), 3, 3))
        tri[:, 1, 0] = a
        tri[:, 2, 0] = x
        tri[:, 2, 1] = y
        tri -= tri.mean(1, keepdims=True)
        out[k:k + len(s)] = tri
        k += len(s)
    R = random_rotations(rng, n)
    out = np.einsum("nij,nkj->nki", R, out)
    out += rng.uniform(lo, hi, (n, 1, 3))
    return out.reshape(-1, 9).astype(f32)


def cube_shell(cells=13, lo=-0.05, hi=1.05, inward=True):
    """Closed axis-aligned cube [lo,hi]^3, each face cells x cells x 2 triangles."""
    base = plane_grid_faces(1, 1, cells, cells).reshape(-1, 3, 3).astype(np.float64)
    L = hi - lo
    faces = []
    for axis in range(3):
        for side in (0, 1):
            p = base.copy()
            q = np.zeros_like(p)
            u, v = [a for a in range(3) if a != axis]
            q[..., u] = lo + L * p[..., 0]
            q[..., v] = lo + L * p[..., 1]
            q[..., axis] = lo + L * side
            n = np.cross(q[:, 1] - q[:, 0], q[:, 2] - q[:, 0])
            want = (1.0 if side == 0 else -1.0) * (1.0 if inward else -1.0)
            flip = np.sign(n[:, axis]) != want
            r = q.copy()
            r[flip, 1], r[flip, 2] = q[flip, 2], q[flip, 1]
            faces.append(r)
    return np.concatenate(faces).reshape(-1, 9).astype(f32)


def confetti_scene(n_confetti=2000, cells=13, seed=0):
    rng = np.random.default_rng(seed)
    faces = np.concatenate([confetti(n_confetti, rng), cube_shell(cells)]).astype(f32)
    return faces, per_face_meshes(len(faces))


def mixed_scene(n_tris, seed):
    """A random mix of a sphere, a cube shell and confetti with about n_tris
    triangles (for batched runs)."""
    rng = np.random.default_rng(seed)
    parts = []
    w = rng.dirichlet([1.0, 1.0, 1.0])
    n_sph, n_cube, n_conf = (w * n_tris).astype(int)
    if n_sph >= 80:
        sub = int(np.clip(np.floor(np.log(max(n_sph, 20) / 20) / np.log(4)), 0, 5))
        c = rng.uniform(0.3, 0.7, 3)
        parts.append(translate(icosphere(sub, radius=rng.uniform(0.1, 0.3), inward=False), c))
    if n_cube >= 12:
        cells = int(max(1, np.sqrt(n_cube / 12)))
        parts.append(cube_shell(cells, -0.05, 1.05))
    n_left = max(1, n_tris - sum(len(p) for p in parts))
    parts.append(confetti(n_left, rng))
    faces = np.concatenate(parts).astype(f32)
    return faces, per_face_meshes(len(faces))


def batch_scene_set(n_scenes=64, seed=0, tri_range=(200, 50_000), ray_range=(1_000, 20_000)):
    """Independent scenes of uneven size: triangle counts log-uniform in
    tri_range, each a random mix of a sphere, a cube shell and confetti from
    its own seed, with a per-scene Malley ray budget uniform in ray_range rays
    per triangle. Returns [(faces, meshes, rays_per_triangle, scene_seed)]."""
    rng = np.random.default_rng(seed)
    n_tris = np.exp(rng.uniform(np.log(tri_range[0]), np.log(tri_range[1]), n_scenes)).astype(int)
    rays = rng.integers(ray_range[0], ray_range[1] + 1, n_scenes)
    out = []
    for k in range(n_scenes):
        scene_seed = int(seed) * 1000 + k
        faces, meshes = mixed_scene(int(n_tris[k]), scene_seed)
        out.append((faces, meshes, int(rays[k]), scene_seed))
    return out


def two_plates_thermal(cells=100, gap=0.5, size=1.0):
    """Two size x size plates, cells x cells x 2 triangles each, one mesh per
    cell (2 faces), the second plate `gap` above the first."""
    a = plane_grid_faces(size, size, cells, cells)
    b = translate(a, (0, 0, gap))
    faces = np.concatenate([a, b]).astype(f32)
    ma = plane_grid_meshes(cells, cells)
    mb = ma.copy()
    mb[:, 0] += len(a)
    return faces, np.concatenate([ma, mb]).astype(np.int32)


def thermal_plates(cells=100, scale=10.0, gap=0.5, emissivity=0.8, thickness=0.001, power=1000.0,
                   patch=10):
    """The mixed-precision thermal workload: a heated plate facing a parallel
    plate, cells x cells cells of two triangles each (one mesh per cell),
    aluminium shell (subject/objects.cpp:167-197: density 2698.9 kg/m^3,
    heat capacity 903 J/(kg K), conductivity 237 W/(m K)), emissivity 0.8 on
    both sides, `power` watts spread over a patch x patch block of cells at
    the heated plate's centre.

    The plates are unit plates scaled by `scale` (10 m, gap 5 m by default):
    form factors are scale invariant, and at 1 m a 100 x 100 grid's triangles
    fall under the reference's 1e-8 degeneracy threshold (they would never be
    hit), while a 1 cm aluminium cell makes the explicit Adams step unstable
    at dt = 0.1 s.

    Returns a dict of faces, meshes, per-mesh area / heat capacity C /
    emissivities / conductivity / thickness, and the source meshes and powers.
    """
    from . import ffk

    faces, meshes = two_plates_thermal(cells=cells, gap=gap * scale, size=scale)
    n = len(meshes)
    w, _ = ffk.area_weights(faces)
    fa = w[0::2]
    area = np.zeros(n, np.float32)
    for m in range(n):  # subject/system.cpp:48-56: float sum per mesh
        a = np.float32(0)
        for f in range(meshes[m, 0], meshes[m, 0] + meshes[m, 1]):
            a = np.float32(a + fa[f])
        area[m] = a
    rho, cap, k = 2.6989e3, 903.0, 237.0
    C = area.astype(np.float64) * thickness * rho * cap
    patch = min(patch, cells)
    lo = (cells - patch) // 2
    src = np.array([i * cells + j for i in range(lo, lo + patch) for j in range(lo, lo + patch)], np.int32)
    return {"faces": faces, "meshes": meshes, "area": area.astype(np.float64), "C": C,
            "eps_front": np.full(n, emissivity), "eps_rear": np.full(n, emissivity),
            "conductivity": np.full(n, k, np.float32), "thickness": np.full(n, thickness, np.float32),
            "src_mesh": src, "src_power": np.full(len(src), power / len(src))}


# ------------------------------------------------ reference test scenes --
# test/form_factors.cpp:41-187 and test/ray_casting.cpp:55-107.

def _face(a, b, c, off=(0, 0, 0)):
    o = np.asarray(off, f32)
    return np.concatenate([np.asarray(a, f32) + o, np.asarray(b, f32) + o, np.asarray(c, f32) + o])


def ff_parallel_planes(c, rotation=None):
    faces = np.stack([
        _face((0, 0, 0), (1, 0, 0), (0, 1, 0)), _face((1, 0, 0), (1, 1, 0), (0, 1, 0)),
        _face((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, c)), _face((1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, c))])
    if rotation is not None:
        faces = rotate_faces(faces, rotation)
    return faces.astype(f32), np.array([[0, 2, 0], [2, 2, 0]], np.int32)


def ff_perpendicular_planes(rotation=None):
    faces = np.stack([
        _face((0, 0, 0), (1, 0, 0), (0, 1, 0)), _face((1, 0, 0), (1, 1, 0), (0, 1, 0)),
        _face((0, 0, 0), (0, 1, 0), (0, 0, 1)), _face((0, 1, 0), (0, 0, 1), (0, 1, 1))])
    if rotation is not None:
        faces = rotate_faces(faces, rotation)
    return faces.astype(f32), np.array([[0, 2, 0], [2, 2, 0]], np.int32)


def ff_box(rotation=None):
    v = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    idx = {0: (0, 1, 2), 1: (0, 3, 2), 8: (0, 1, 5), 9: (0, 4, 5), 4: (0, 3, 7), 5: (0, 4, 7), 6: (1, 2, 6),
           7: (1, 5, 6), 2: (3, 2, 6), 3: (3, 7, 6), 10: (4, 5, 6), 11: (4, 7, 6)}
    faces = np.stack([_face(v[idx[k][0]], v[idx[k][1]], v[idx[k][2]]) for k in range(12)])
    if rotation is not None:
        faces = rotate_faces(faces, rotation)
    return faces.astype(f32), np.array([[0, 12, 0]], np.int32)


def rotate_faces(faces, m):
    """Apply mat33 * vec3 with the reference's column access
    (math/mat.h:30-37) to every vertex, in float32 order."""
    m = np.asarray(m, f32).reshape(3, 3)
    p = np.asarray(faces, f32).reshape(-1, 3)
    x = (m[0, 0] * p[:, 0]) + (m[1, 0] * p[:, 1]) + (m[2, 0] * p[:, 2])
    y = (m[0, 1] * p[:, 0]) + (m[1, 1] * p[:, 1]) + (m[2, 1] * p[:, 2])
    z = (m[0, 2] * p[:, 0]) + (m[1, 2] * p[:, 1]) + (m[2, 2] * p[:, 2])
    return np.stack([x, y, z], 1).astype(f32).reshape(-1, 9)
