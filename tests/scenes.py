"""Small analytic test scenes built in numpy (host arrays in the layout of
wf_scene_create: vertices [nv, 6] = position + normal, indices [nt, 3])."""
from __future__ import annotations

import numpy as np


def uv_sphere(radius=10.0, n_lat=48, n_lon=96, center=(0.0, 0.0, 0.0)):
    """A closed UV sphere: two pole vertices, n_lat - 1 rings of n_lon
    vertices, pole fans and split quads.  Vertex normals are radial, so the
    interpolated normal at any point of a triangle points along that point's
    position relative to the centre (the interpolation is linear in the
    vertices, which all lie at the same radius)."""
    c = np.asarray(center, np.float64)
    th = np.pi * np.arange(1, n_lat) / n_lat
    ph = 2 * np.pi * np.arange(n_lon) / n_lon
    T, P = np.meshgrid(th, ph, indexing="ij")
    ring = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1).reshape(-1, 3)
    dirs = np.concatenate([[[0, 1, 0]], ring, [[0, -1, 0]]])
    top, bot = 0, len(dirs) - 1
    idx = []
    r0 = lambda i: 1 + i * n_lon  # noqa: E731  first vertex of ring i
    for j in range(n_lon):
        k = (j + 1) % n_lon
        idx.append([top, r0(0) + j, r0(0) + k])
        for i in range(n_lat - 2):
            a, b = r0(i) + j, r0(i) + k
            a2, b2 = r0(i + 1) + j, r0(i + 1) + k
            idx += [[a, a2, b2], [a, b2, b]]
        idx.append([r0(n_lat - 2) + j, bot, r0(n_lat - 2) + k])
    pos = (c + radius * dirs).astype(np.float32)
    nrm = dirs.astype(np.float32)
    return np.concatenate([pos, nrm], 1), np.array(idx, np.uint32)


def tri_arrays(vertices, indices):
    """(tri_verts [nt, 9], tri_normals [nt, 9]) in triangle order."""
    tv = vertices[:, :3][indices.reshape(-1)].reshape(-1, 9).astype(np.float32)
    tn = vertices[:, 3:][indices.reshape(-1)].reshape(-1, 9).astype(np.float32)
    return tv, tn


def furnace_value(albedo, depth):
    """Pixel value inside a closed sphere lit from its centre: every path
    vertex sees the light head-on (N.L = 1) and unoccluded, so the
    per-bounce estimator sums a^b over vertices b = 0..depth."""
    a = np.asarray(albedo, np.float64)
    return sum(a ** b for b in range(depth + 1))


def furnace_outcomes(albedo, background, depth):
    """Every value a pixel of the lit sphere can take: the closed path
    (furnace_value), or a path whose bounce k (1..depth) leaves through a
    facet and sees the background after k vertices: sum_{b<k} a^b + a^k bg.
    (A bounce drawn about the interpolated, radial shading normal can point
    below the flat facet's plane; it then exits the convex hull.)"""
    a = np.asarray(albedo, np.float64)
    bg = np.asarray(background, np.float64)
    out = [furnace_value(a, depth)]
    for k in range(1, depth + 1):
        out.append(sum(a ** b for b in range(k)) + a ** k * bg)
    return np.array(out)


def match_outcomes(img, outcomes, rtol=1e-6):
    """Index of the outcome each pixel equals (rtol per channel), -1 for
    none."""
    px = np.asarray(img, np.float64)[:, None, :3]
    ok = np.all(np.abs(px - outcomes[None]) <= rtol * np.abs(outcomes[None]), axis=2)
    return np.where(ok.any(1), ok.argmax(1), -1)
