"""Cross-check of two casters on one ray set, as the reference's perfTest does
(perfTest/perfTest.cpp:42-153): its confetti workload and its error report.

* `perftest_scene` restates perfTest's scene and rays (MakeConfettiScene,
  MakeCollapsingRays, :42-83): each triangle is three points of the unit
  sphere moved by a point of the ball of `radius`; each ray ends on the
  sphere of radius 2 + radius and starts 10 % further out, so it runs toward
  the centre. Points on the sphere follow the reference's spherical
  generator (z uniform in [-1, 1), azimuth uniform; ball points scale a
  sphere point by u^0.3333, subject/spherial_generator.cpp:28-46), drawn from
  numpy's generator instead of std::mt19937.
* `error_stats` is CalculateError (:114-140): rays whose hit faces differ
  count as mismatches; for rays that hit the same face the distance between
  the two hit points is summed (float, in ray order) and its maximum kept;
  the average divides the sum by all rays.
"""
from __future__ import annotations

import numpy as np

f32 = np.float32


def _sphere_points(rng, n):
    theta = rng.uniform(0, 2 * np.pi, n).astype(f32)
    u = rng.uniform(-1, 1, n).astype(f32)
    c = np.sqrt(f32(1) - u * u)
    return np.stack([c * np.cos(theta), c * np.sin(theta), u], 1).astype(f32)


def perftest_scene(n_faces=38000, radius=20.0, n_rays=None, seed=0):
    """(faces float32[n_faces, 9], rays float32[n_rays, 6]); n_rays defaults
    to 100 per face, as in perfTest."""
    n_rays = 100 * n_faces if n_rays is None else n_rays
    rng = np.random.default_rng(seed)
    pts = _sphere_points(rng, 3 * n_faces).reshape(n_faces, 3, 3)
    ball = _sphere_points(rng, n_faces) * np.power(rng.uniform(0, 1, n_faces), 0.3333).astype(f32)[:, None]
    faces = (pts + (ball * f32(radius))[:, None, :]).astype(f32).reshape(n_faces, 9)
    end = _sphere_points(rng, n_rays) * f32(2 + radius)
    origin = (end * f32(1.1)).astype(f32)
    return faces, np.concatenate([origin, end], 1).astype(f32)


def error_stats(idx_ref, pts_ref, idx, pts):
    """CalculateError of perfTest between a reference cast (idx_ref, pts_ref)
    and another (idx, pts): hit face indices (-1 = miss) and hit points."""
    idx_ref, idx = np.asarray(idx_ref), np.asarray(idx)
    n = len(idx_ref)
    same = idx_ref == idx
    both = same & (idx >= 0)
    d = np.asarray(pts_ref, f32)[both] - np.asarray(pts, f32)[both]
    dist = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]).astype(f32)
    total = np.cumsum(dist, dtype=f32)[-1] if len(dist) else f32(0)  # sequential float sum, in ray order
    return {"mismatch": int(n - same.sum()), "total_distance": float(total),
            "max_distance": float(dist.max()) if len(dist) else 0.0,
            "average_distance": float(f32(total) / f32(n)) if n else 0.0, "rays": n}
