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


def _log_uniform_int(rng, lo, hi):
    return int(round(float(np.exp(rng.uniform(np.log(lo), np.log(hi))))))


def landmark_momenta_scale(n, sigma, d=3):
    """Momentum scale of the landmark targets: 0.1 / sqrt(1 + n sigma^d), so that the summed
    kernel-weighted momenta (about n sigma^d points within one kernel width of uniform points
    in the unit ball) displace points by about 0.1 whatever n and sigma."""
    return 0.1 / np.sqrt(1.0 + n * sigma ** d)


def mixed_service_requests(n=2000, seed=5, shoot=None, rate=200.0, burst_factor=5.0,
                           burst_prob=0.2):
    """The mixed-service stream: 40% 2D closed curves (256-8192 points, log-uniform; star
    targets r = 1 + A cos(k theta), k in 3..8, A ~ U[0.05, 0.3]), 40% surfaces (icosphere
    subdivision 3-7, uniform in subdivision = log-uniform in vertex count; random rotation;
    RBF targets as for the large surface pair), 20% 3D landmark sets (1,000-200,000 points
    uniform in the unit ball, log-uniform; the target is x_T of one seeded geodesic shooting
    from momenta N(0, s^2), s = landmark_momenta_scale).  Per request: Gaussian or Laplacian
    (order 3) deformation kernel with sigma ~ U[0.1, 0.5], T in {10, 20, 40}, currents kernel
    width 0.1, sigma_err 0.1, 30 CG iterations.  Arrivals: Poisson at ``rate`` with each gap
    shortened ``burst_factor`` times with probability ``burst_prob`` (bursts).

    Request i is drawn from its own generator seeded (seed, i), so the first k requests of
    any longer stream are the stream of length k (and ``mixed_service_request`` rebuilds one).
    shoot(x0, a0, kernel_name, sigma, T) -> x_T builds the landmark targets; the default uses
    this package's GPU geodesic shooting (pointEvolution.landmarkEPDiff).  Returns a list of
    dicts with the Request fields plus 'kind', 'size' and 'arrival'."""
    meshes = {}
    out = []
    t = 0.0
    for i in range(n):
        req = _mixed_request(seed, i, shoot, rate, burst_factor, burst_prob, meshes)
        t += req.pop("gap")
        req["arrival"] = t
        out.append(req)
    return out


def mixed_service_request(i, seed=5, shoot=None, rate=200.0, burst_factor=5.0, burst_prob=0.2):
    """Request i of the mixed-service stream alone (same fields and values)."""
    t = 0.0
    for j in range(i):
        rng = np.random.default_rng((seed, j))
        gap = rng.exponential(1.0 / rate)
        t += gap / burst_factor if rng.random() < burst_prob else gap
    req = _mixed_request(seed, i, shoot, rate, burst_factor, burst_prob, {})
    req["arrival"] = t + req.pop("gap")
    return req


def _mixed_request(seed, i, shoot, rate, burst_factor, burst_prob, meshes):
    if shoot is None:
        shoot = _gpu_shoot
    rng = np.random.default_rng((seed, i))
    gap = rng.exponential(1.0 / rate)
    gap = gap / burst_factor if rng.random() < burst_prob else gap
    u = rng.random()
    kernel = 'gauss' if rng.random() < 0.5 else 'laplacian'
    sigma = float(rng.uniform(0.1, 0.5))
    T = int(rng.choice([10, 20, 40]))
    common = dict(sigmaKernel=sigma, sigmaDist=0.1, sigmaError=0.1, T=T, maxIter=30,
                  typeKernel=kernel, order=10 if kernel == 'gauss' else 3, gap=gap)
    if u < 0.4:
        npts = _log_uniform_int(rng, 256, 8192)
        X, _ = circle(npts)
        Y = star_curve(npts, int(rng.integers(3, 9)), float(rng.uniform(0.05, 0.3)))
        E = closed_curve_edges(npts)
        return dict(kind='curve', template=(E, X), target=(E, Y), size=npts, **common)
    if u < 0.8:
        s = int(rng.integers(3, 8))
        if s not in meshes:
            meshes[s] = icosphere(s)
        V, F = meshes[s]
        R = random_rotation(rng)
        Vr = V @ R.T
        return dict(kind='surface', template=(F, Vr), target=(F, rbf_field(rng, Vr)),
                    size=V.shape[0], subdiv=s, **common)
    npts = _log_uniform_int(rng, 1000, 200000)
    X = ball_points(rng, npts)
    a0 = landmark_momenta_scale(npts, sigma) * rng.standard_normal(X.shape)
    Y = shoot(X, a0, kernel, sigma, T)
    return dict(kind='landmark', template=(None, X), target=(None, Y), size=npts, **common)


def _gpu_shoot(x0, a0, kernel_name, sigma, T):
    from . import kernelFunctions as kfun
    from . import pointEvolution as pe
    K = kfun.Kernel(name=kernel_name, sigma=sigma, order=10 if kernel_name == 'gauss' else 3,
                    precision='fp64')
    xt, _ = pe.landmarkEPDiff(T, x0, a0, K)
    return np.asarray(xt[-1])
