"""Seeded synthetic shapes for the benchmark workloads and tests (host NumPy, float64).

The reference ships no generators or data files; these are this package's own.
"""
from __future__ import annotations

import numpy as np


def icosphere(subdiv: int):
    """Unit icosphere: 10*4^s + 2 vertices, 20*4^s outward counter-clockwise faces."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0),
         (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
         (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
         (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
         (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    faces = np.array(f, dtype=np.int64)
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            idx = cache.get(key)
            if idx is None:
                p = verts[a] + verts[b]
                verts.append(p / np.linalg.norm(p))
                idx = len(verts) - 1
                cache[key] = idx
            return idx

        nf = np.empty((4 * len(faces), 3), dtype=np.int64)
        for i, (a, b, c) in enumerate(faces):
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf[4 * i:4 * i + 4] = [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts)
    # Orient faces outward: the normal of each face must point away from the origin.
    n = np.cross(V[faces[:, 1]] - V[faces[:, 0]], V[faces[:, 2]] - V[faces[:, 0]])
    flip = (n * V[faces].mean(axis=1)).sum(axis=1) < 0
    faces[flip] = faces[flip][:, [0, 2, 1]]
    return V, faces


def bump_target(V, scale=(1.2, 0.9, 1.0), height=0.2, width=0.3):
    """Template vertices plus a radial Gaussian bump centred at +z (evaluated on the template
    coordinates), then scaled per axis: y = S (x + h exp(-|x-e_z|^2/(2 w^2)) x/|x|)."""
    ez = np.array([0.0, 0.0, 1.0])
    r2 = ((V - ez) ** 2).sum(axis=1)
    b = height * np.exp(-r2 / (2 * width ** 2))
    Y = V + b[:, None] * V / np.linalg.norm(V, axis=1)[:, None]
    return Y * np.asarray(scale)[None, :]


def unit_sphere_points(rng, n):
    p = rng.standard_normal((n, 3))
    return p / np.linalg.norm(p, axis=1)[:, None]


def rbf_field(rng, V, n_rbf=20, width=0.4, amp=0.1):
    """x + sum_j alpha_j exp(-|x - c_j|^2 / (2 width^2)), c_j uniform on the sphere,
    alpha_j ~ N(0, amp^2 I)."""
    C = unit_sphere_points(rng, n_rbf)
    A = amp * rng.standard_normal((n_rbf, V.shape[1]))
    d2 = ((V[:, None, :] - C[None, :, :]) ** 2).sum(axis=2)
    return V + np.exp(-d2 / (2 * width ** 2)) @ A


def random_rotation(rng):
    """Haar-distributed rotation (QR of a Gaussian matrix with sign fix)."""
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))[None, :]
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def circle(n, radius=1.0):
    th = 2 * np.pi * np.arange(n) / n
    return np.stack([radius * np.cos(th), radius * np.sin(th)], axis=1), th


def star_curve(n, lobes, amp, noise=0.0, rng=None):
    """r(theta) = 1 + amp cos(lobes theta) (+ N(0, noise^2) radial noise) at n equal angles."""
    th = 2 * np.pi * np.arange(n) / n
    r = 1.0 + amp * np.cos(lobes * th)
    if noise > 0:
        r = r + noise * rng.standard_normal(n)
    return np.stack([r * np.cos(th), r * np.sin(th)], axis=1)


def ball_points(rng, n, d=3):
    """Uniform in the unit ball: direction uniform, radius U^(1/d)."""
    p = rng.standard_normal((n, d))
    p /= np.linalg.norm(p, axis=1)[:, None]
    return p * rng.random(n)[:, None] ** (1.0 / d)


def closed_curve_edges(n):
    e = np.empty((n, 2), dtype=np.int64)
    e[:, 0] = np.arange(n)
    e[:, 1] = (np.arange(n) + 1) % n
    return e


# Workload definitions -------------------------------------------------------------------
def surface_pair_config1():
    """Icosphere s=3 -> scaled and bumped copy; sigma_V=0.5, sigma_W=0.3, sigma_err=0.1, T=10."""
    V, F = icosphere(3)
    return dict(template=(F, V), target=(F, bump_target(V)), sigmaKernel=0.5, sigmaDist=0.3,
                sigmaError=0.1, T=10, maxIter=50)


def surface_pair_config2(seed=2):
    """Icosphere s=6 -> seeded smooth RBF field; sigma_V=0.25, sigma_W=0.1, sigma_err=0.1, T=20."""
    rng = np.random.default_rng(seed)
    V, F = icosphere(6)
    return dict(template=(F, V), target=(F, rbf_field(rng, V)), sigmaKernel=0.25, sigmaDist=0.1,
                sigmaError=0.1, T=20, maxIter=100)


def served_surface_requests(n=512, seed=4, subdivs=(3, 4, 5, 6)):
    """Registration pairs of the served-surfaces workload: subdivision drawn so that the vertex
    count is log-uniform over the range, random rotation, per-pair RBF target, sigma_V ~ U[0.2,0.5],
    sigma_W = 0.1, sigma_err = 0.1, T = 10, 30 CG iterations."""
    rng = np.random.default_rng(seed)
    meshes = {s: icosphere(s) for s in subdivs}
    out = []
    for i in range(n):
        s = int(rng.choice(subdivs))
        V, F = meshes[s]
        R = random_rotation(rng)
        Vr = V @ R.T
        out.append(dict(template=(F, Vr), target=(F, rbf_field(rng, Vr)),
                        sigmaKernel=float(rng.uniform(0.2, 0.5)), sigmaDist=0.1, sigmaError=0.1,
                        T=10, maxIter=30, subdiv=s))
    return out
