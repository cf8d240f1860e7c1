"""Pins the FP64 CPU oracle (oracle/) before it is trusted as the checker.

Each check is independent of the oracle's own formulas: the reference's compiled Fortran
(golden fixtures, and live where oracle/_ref is built), brute force with SciPy distances,
closed forms, finite differences of the objective, and invariants of the flow.
"""
import numpy as np
import pytest
from scipy.spatial.distance import cdist

from oracle import lddmm_oracle as orc
from registration_amd import synthetic

F32 = lambda v: float(np.float32(v))


# ---------------------------------------------------------------- reference golden vectors
def test_oracle_matches_reference_golden(golden):
    for name, c in golden.items():
        d, ny, nx, m, na, order, sigma = c["params"]
        order = int(order)
        f = orc.apply_k(c["y"], c["x"], c["w"], sigma, order)
        np.testing.assert_allclose(f, c["applyk"], rtol=1e-12, atol=1e-13 * np.abs(c["applyk"]).max(),
                                   err_msg=name)
        if "applykdifft" in c:
            g = orc.apply_diff_kt(c["y"], c["x"], c["a1"], c["a2"], sigma, order)
            np.testing.assert_allclose(g, c["applykdifft"], rtol=1e-11,
                                       atol=1e-12 * np.abs(c["applykdifft"]).max(), err_msg=name)
        if "applykdiff11and12" in c:
            h = orc.apply_ddiff_k11and12(c["y"], c["dd_a1"], c["dd_a2"], c["dd_p"], sigma, order)
            np.testing.assert_allclose(h, c["applykdiff11and12"], rtol=1e-11,
                                       atol=1e-12 * np.abs(c["applykdiff11and12"]).max(),
                                       err_msg=name)


def test_oracle_matrix_sums_match_reference_golden(golden_mat):
    for name, c in golden_mat.items():
        d, ny, nx, m, order, sigma = c["params"]
        f = orc.apply_k_mat(c["y"], c["x"], c["beta"], sigma, int(order))
        np.testing.assert_allclose(f, c["applykmat"], rtol=1e-12,
                                   atol=1e-13 * np.abs(c["applykmat"]).max(), err_msg=name)
        g = orc.apply_diff_k_mat(c["y"], c["x"], c["beta"][..., 0], sigma, int(order))
        np.testing.assert_allclose(g, c["applykdiffmat"], rtol=1e-11,
                                   atol=1e-12 * np.abs(c["applykdiffmat"]).max(), err_msg=name)


def test_matrix_sums_reduce_to_vector_sums():
    """beta[k, l] = w[l] turns applyKmat into applyK; beta[k, l] = <a1_k, a2_l> turns
    applyDiffKmat into applyDiffKT (one pair)."""
    rng = np.random.default_rng(8)
    y, x = rng.standard_normal((20, 3)), rng.standard_normal((25, 3))
    w = rng.standard_normal((25, 2))
    for order in (10, 2):
        np.testing.assert_allclose(orc.apply_k_mat(y, x, np.broadcast_to(w, (20, 25, 2)), 0.7, order),
                                   orc.apply_k(y, x, w, 0.7, order), rtol=1e-13, atol=1e-14)
        a1, a2 = rng.standard_normal((20, 3)), rng.standard_normal((25, 3))
        np.testing.assert_allclose(orc.apply_diff_k_mat(y, x, a1 @ a2.T, 0.7, order),
                                   orc.apply_diff_kt(y, x, a1[None], a2[None], 0.7, order),
                                   rtol=1e-11, atol=1e-12)


@pytest.mark.skipif(not orc.ref_available(), reason="oracle/_ref not built here")
@pytest.mark.parametrize("order,sigma", [(10, 0.5), (1, 0.3), (2, 0.6), (3, 0.4), (4, 0.8)])
def test_oracle_matches_reference_live(order, sigma):
    rng = np.random.default_rng(order)
    y = rng.standard_normal((41, 3))
    x = rng.standard_normal((57, 3))
    a1 = rng.standard_normal((3, 41, 3))
    a2 = rng.standard_normal((3, 57, 3))
    w = rng.standard_normal((57, 3))
    for fn, args in ((orc.apply_k, (y, x, w)), (orc.apply_diff_kt, (y, x, a1, a2))):
        got = fn(*args, sigma, order)
        ref = fn(*args, sigma, order, sums="reference")
        np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-12 * np.abs(ref).max())


# ---------------------------------------------------------------- brute force and closed forms
def test_gaussian_brute_force():
    rng = np.random.default_rng(1)
    y, x, w = rng.standard_normal((30, 3)), rng.standard_normal((25, 3)), rng.standard_normal((25, 2))
    sig = 0.8
    Kmat = np.exp(-cdist(y, x, 'sqeuclidean') / (2 * sig ** 2))
    np.testing.assert_allclose(orc.apply_k(y, x, w, sig, 10), Kmat @ w, rtol=1e-12)


@pytest.mark.parametrize("order", [0, 1, 2, 3, 4])
def test_laplacian_two_point_closed_form(order):
    # Two points at distance r: f_0 = w_0 + K(r) w_1 with K = P_order(u) e^{-u}, u = r/sigma,
    # P built from the float32-rounded literals of the reference's table.
    P = {0: [1], 1: [1, 1], 2: [1, 1, F32(1 / 3)], 3: [1, 1, F32(0.4), F32(1 / 15)],
         4: [1, 1, F32(3 / 7), F32(2 / 21), F32(1 / 105)]}[order]
    sig, r = 0.7, 0.45
    x = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
    w = np.array([[1.0], [2.0]])
    u = r / sig
    K = sum(c * u ** j for j, c in enumerate(P)) * np.exp(-u)
    f = orc.apply_k(x, x, w, sig, order)
    assert abs(f[0, 0] - (1.0 + 2.0 * K)) < 1e-14
    # The exact-double table differs from the float32 one: the oracle must hold the latter.
    if order >= 2:
        Pexact = {2: [1, 1, 1 / 3], 3: [1, 1, 0.4, 1 / 15], 4: [1, 1, 3 / 7, 2 / 21, 1 / 105]}[order]
        Kx = sum(c * u ** j for j, c in enumerate(Pexact)) * np.exp(-u)
        assert abs(f[0, 0] - (1.0 + 2.0 * Kx)) > 1e-10


def test_cutoff_semantics():
    # Below the float32 literal 1e-8 the kernel is exactly 1; a Gaussian just above it is not.
    sig = 1.0
    for ut, expect_one in ((9.0e-9, True), (1.2e-8, False)):
        r = np.sqrt(ut) * sig
        x = np.array([[0.0, 0.0], [r, 0.0]])
        w = np.array([[0.0], [1.0]])
        f = orc.apply_k(x[:1], x, w, sig, 10)[0, 0]
        assert (f == 1.0) == expect_one
    # Laplacian: the cut-off applies to r/sigma
    x = np.array([[0.0, 0.0], [5e-9, 0.0]])
    f = orc.apply_k(x[:1], x, np.array([[0.0], [1.0]]), 1.0, 2)[0, 0]
    assert f == 1.0


@pytest.mark.parametrize("order,sigma", [(10, 0.6), (1, 0.5), (2, 0.7), (3, 0.6), (4, 0.9)])
def test_diff_kt_is_gradient_of_kernel_energy(order, sigma):
    # E(y) = sum_k <a1_k, (K(y,x) a2)_k>  =>  dE/dy_k = applyDiffKT(x, [a1], [a2], firstVar=y).
    rng = np.random.default_rng(order + 7)
    y, x = rng.standard_normal((9, 3)), rng.standard_normal((11, 3))
    a1, a2 = rng.standard_normal((9, 3)), rng.standard_normal((11, 3))
    g = orc.apply_diff_kt(y, x, a1[None], a2[None], sigma, order)
    E = lambda yy: (a1 * orc.apply_k(yy, x, a2, sigma, order)).sum()
    h = 1e-6
    fd = np.zeros_like(y)
    for k in range(9):
        for c in range(3):
            e = np.zeros_like(y)
            e[k, c] = h
            fd[k, c] = (E(y + e) - E(y - e)) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=2e-6, atol=2e-7 * np.abs(fd).max())


@pytest.mark.parametrize("order,sigma", [(10, 0.6), (2, 0.7), (3, 0.6), (4, 0.9)])
def test_ddiff_k11and12_is_directional_derivative(order, sigma):
    # applyDDiffK11and12(x, a1, a2, p) = d/de applyDiffKT(x + e p, [a1], [a2]) at e = 0.
    rng = np.random.default_rng(order + 11)
    x = rng.standard_normal((12, 3))
    a1, a2, p = (rng.standard_normal((12, 3)) for _ in range(3))
    f = orc.apply_ddiff_k11and12(x, a1, a2, p, sigma, order)
    G = lambda xx: orc.apply_diff_kt(xx, xx, a1[None], a2[None], sigma, order)
    h = 1e-6
    fd = (G(x + h * p) - G(x - h * p)) / (2 * h)
    np.testing.assert_allclose(f, fd, rtol=2e-5, atol=2e-6 * np.abs(fd).max())


def test_invalid_orders_rejected():
    x = np.zeros((2, 3))
    with pytest.raises(ValueError):
        orc.apply_diff_kt(x, x, np.zeros((1, 2, 3)), np.zeros((1, 2, 3)), 1.0, 0)
    with pytest.raises(ValueError):
        orc.apply_ddiff_k11and12(x, x, x, x, 1.0, 1)


# ---------------------------------------------------------------- flow, adjoint, EPDiff
def test_flow_invariants():
    K = orc.Kernel('gauss', sigma=0.5)
    rng = np.random.default_rng(3)
    x0 = rng.standard_normal((20, 3))
    xt = orc.landmarkDirectEvolutionEuler(x0, np.zeros((5, 20, 3)), K)
    assert np.array_equal(xt[-1], x0)
    # One point: K(x,x) = 1, so x_T = x_0 + dt sum_k a_k (a straight polyline).
    at = rng.standard_normal((4, 1, 3))
    xt = orc.landmarkDirectEvolutionEuler(x0[:1], at, K)
    np.testing.assert_allclose(xt[-1], x0[:1] + at.sum(axis=0) / 4, rtol=1e-14)


def test_covector_is_adjoint_of_objective():
    V, F = synthetic.icosphere(1)
    Y = synthetic.bump_target(V)
    o = orc.SurfaceMatching(orc.Surface(FV=(F, V)), orc.Surface(FV=(F, Y)),
                            orc.Kernel('gauss', sigma=0.5), orc.Kernel('gauss', sigma=0.3),
                            sigmaError=0.1, timeStep=0.25)
    rng = np.random.default_rng(5)
    o.at = 0.05 * rng.standard_normal(o.at.shape)
    o.objectiveFun()
    o.endOfIteration()
    g = o.getGradient(1.0)
    d = orc.Direction()
    d.diff = rng.standard_normal(o.at.shape)
    h = 1e-5
    fd = (o.updateTry(d, -h) - o.updateTry(d, h)) / (2 * h)
    [gd] = o.dotProduct(g, [d])
    assert abs(fd - gd) < 1e-6 * abs(fd)


def test_epdiff_conserves_hamiltonian():
    K = orc.Kernel('gauss', sigma=0.4)
    rng = np.random.default_rng(8)
    x0 = synthetic.ball_points(rng, 40)
    a0 = 0.05 * rng.standard_normal((40, 3))
    H = lambda x, a: 0.5 * (a * K.applyK(x, a)).sum()
    errs = []
    for T in (10, 20, 40):
        xt, at = orc.landmarkEPDiff(T, x0, a0, K)
        errs.append(abs(H(xt[T - 1], at[T - 1]) - H(x0, a0)))
    assert errs[1] < 0.7 * errs[0] and errs[2] < 0.7 * errs[1]  # O(dt) drift


# ---------------------------------------------------------------- currents
def test_currents_norm_identities():
    V, F = synthetic.icosphere(1)
    K = orc.Kernel('gauss', sigma=0.3)
    S = orc.Surface(FV=(F, V))
    T = orc.Surface(FV=(F, synthetic.bump_target(V)))
    assert abs(orc.surf_currentNormDef(S, S, K) + orc.surf_currentNorm0(S, K)) < 1e-12 * orc.surf_currentNorm0(S, K)
    assert orc.surf_currentNormDef(S, T, K) + orc.surf_currentNorm0(T, K) > 0


def test_currents_single_triangle_closed_form():
    K = orc.Kernel('gauss', sigma=0.5)
    A = orc.Surface(FV=(np.array([[0, 1, 2]]), np.array([[0., 0, 0], [1, 0, 0], [0, 1, 0]])))
    B = orc.Surface(FV=(np.array([[0, 1, 2]]), np.array([[0., 0, 0.3], [1, 0, 0.3], [0, 2, 0.5]])))
    n1, n2 = A.surfel[0], B.surfel[0]
    k12 = np.exp(-((A.centers[0] - B.centers[0]) ** 2).sum() / (2 * 0.25))
    expect = n1 @ n1 - 2 * k12 * (n1 @ n2)
    assert abs(orc.surf_currentNormDef(A, B, K) - expect) < 1e-14


@pytest.mark.parametrize("kind", ["surface", "curve"])
def test_currents_gradient_fd(kind):
    rng = np.random.default_rng(9)
    if kind == "surface":
        V, F = synthetic.icosphere(1)
        mk = lambda X: orc.Surface(FV=(F, X))
        norm = lambda a, b, K: orc.surf_currentNormDef(a, b, K) + orc.surf_currentNorm0(b, K)
        grad = orc.surf_currentNormGradient
        Y = synthetic.bump_target(V)
    else:
        V, _ = synthetic.circle(24)
        F = synthetic.closed_curve_edges(24)
        mk = lambda X: orc.Curve(FV=(F, X))
        norm = lambda a, b, K: orc.curve_currentNormDef(a, b, K) + orc.curve_currentNorm0(b, K)
        grad = orc.curve_currentNormGradient
        Y = synthetic.star_curve(24, 5, 0.2)
    K = orc.Kernel('gauss', sigma=0.4)
    X = V + 0.05 * rng.standard_normal(V.shape)
    g = grad(mk(X), mk(Y), K)
    h = 1e-6
    for (i, c) in [(0, 0), (3, 1), (7, 0), (10, 1)]:
        e = np.zeros_like(X)
        e[i, c] = h
        fd = (norm(mk(X + e), mk(Y), K) - norm(mk(X - e), mk(Y), K)) / (2 * h)
        assert abs(fd - g[i, c]) < 1e-6 * max(1.0, abs(fd))


# ---------------------------------------------------------------- measure data term
def _measure_brute(c1, a1, c2, a2, sigma):
    """|mu1 - mu2|^2 for mu = sum_k a_k delta_{c_k} in the Gaussian RKHS, by a double loop."""
    def dot(ca, aa, cb, ab):
        return sum(aa[i] * ab[j] * np.exp(-((ca[i] - cb[j]) ** 2).sum() / (2 * sigma ** 2))
                   for i in range(len(aa)) for j in range(len(ab)))
    return dot(c1, a1, c1, a1) - 2 * dot(c1, a1, c2, a2) + dot(c2, a2, c2, a2)


def test_measure_norm_brute_force_and_invariances():
    V, F = synthetic.icosphere(1)
    K = orc.Kernel('gauss', sigma=0.35)
    S = orc.Surface(FV=(F, V))
    T = orc.Surface(FV=(F, synthetic.bump_target(V)))
    a = lambda s: np.sqrt((s.surfel ** 2).sum(axis=1) + 1e-10)
    full = orc.surf_measureNormDef(S, T, K) + orc.surf_measureNorm0(T, K)
    brute = _measure_brute(S.centers, a(S), T.centers, a(T), 0.35)
    assert abs(full - brute) < 1e-12 * orc.surf_measureNorm0(T, K)
    assert abs(orc.surf_measureNormDef(S, S, K) + orc.surf_measureNorm0(S, K)) < 1e-12 * orc.surf_measureNorm0(S, K)
    # The measure ignores orientation (currents do not): flipping every face of T keeps it.
    Tf = orc.Surface(surf=T)
    Tf.flipFaces()
    assert abs(orc.surf_measureNormDef(S, Tf, K) - orc.surf_measureNormDef(S, T, K)) < 1e-12 * abs(full)
    assert abs(orc.surf_currentNormDef(S, Tf, K) - orc.surf_currentNormDef(S, T, K)) > 1e-3
    # Rigid motions of both shapes leave it unchanged.
    R = synthetic.random_rotation(np.random.default_rng(2))
    Sr, Tr = orc.Surface(FV=(F, S.vertices @ R.T + 0.3)), orc.Surface(FV=(F, T.vertices @ R.T + 0.3))
    assert abs(orc.surf_measureNormDef(Sr, Tr, K) - orc.surf_measureNormDef(S, T, K)) < 1e-11 * abs(full)


def test_curve_measure_norm0_has_no_offset():
    """curves.measureNorm0 weights are |lenel| while measureNormDef's are sqrt(|lenel|^2 +
    1e-10): the two-edge closed forms tell them apart."""
    K = orc.Kernel('gauss', sigma=0.5)
    C = orc.Curve(FV=(np.array([[0, 1]]), np.array([[0., 0], [1e-5, 0]])))
    l = 1e-5
    assert abs(orc.curve_measureNorm0(C, K) - l * l) < 1e-24
    assert abs(orc.curve_measureNormDef(C, C, K) - (-(l * l + 1e-10))) < 1e-24


@pytest.mark.parametrize("kind", ["surface", "curve"])
def test_measure_gradient_fd(kind):
    rng = np.random.default_rng(19)
    if kind == "surface":
        V, F = synthetic.icosphere(1)
        mk = lambda X: orc.Surface(FV=(F, X))
        norm = orc.surf_measureNormDef
        grad = orc.surf_measureNormGradient
        Y = synthetic.bump_target(V)
    else:
        V, _ = synthetic.circle(24)
        F = synthetic.closed_curve_edges(24)
        mk = lambda X: orc.Curve(FV=(F, X))
        norm = orc.curve_measureNormDef
        grad = orc.curve_measureNormGradient
        Y = synthetic.star_curve(24, 5, 0.2)
    K = orc.Kernel('gauss', sigma=0.4)
    X = V + 0.05 * rng.standard_normal(V.shape)
    g = grad(mk(X), mk(Y), K)
    h = 1e-6
    for (i, c) in [(0, 0), (3, 1), (7, 0), (10, 1), (5, V.shape[1] - 1)]:
        e = np.zeros_like(X)
        e[i, c] = h
        fd = (norm(mk(X + e), mk(Y), K) - norm(mk(X - e), mk(Y), K)) / (2 * h)
        assert abs(fd - g[i, c]) < 1e-6 * max(1.0, abs(fd)), (i, c, fd, g[i, c])


def _varifold_brute(c1, n1, c2, n2, sigma):
    """|V1 - V2|^2 for varifolds sum_k a_k delta_{(c_k, u_k)} with the kernel
    K(c, c') (1 + <u, u'>^2), a = sqrt(|n|^2 + 1e-10), u = n / a, by a double loop."""
    def parts(n):
        a = np.sqrt((n ** 2).sum(axis=1) + 1e-10)
        return a, n / a[:, None]

    def dot(ca, na, cb, nb):
        aa, ua = parts(na)
        ab, ub = parts(nb)
        return sum(aa[i] * ab[j] * (1 + (ua[i] @ ub[j]) ** 2)
                   * np.exp(-((ca[i] - cb[j]) ** 2).sum() / (2 * sigma ** 2))
                   for i in range(len(aa)) for j in range(len(ab)))
    return dot(c1, n1, c1, n1) - 2 * dot(c1, n1, c2, n2) + dot(c2, n2, c2, n2)


def test_varifold_norm_brute_force_and_invariances():
    V, F = synthetic.icosphere(1)
    K = orc.Kernel('gauss', sigma=0.35)
    S = orc.Surface(FV=(F, V))
    T = orc.Surface(FV=(F, synthetic.bump_target(V)))
    full = orc.surf_varifoldNormDef(S, T, K) + orc.surf_varifoldNorm0(T, K)
    brute = _varifold_brute(S.centers, S.surfel, T.centers, T.surfel, 0.35)
    n0 = orc.surf_varifoldNorm0(T, K)
    assert abs(full - brute) < 1e-12 * n0
    assert abs(orc.surf_varifoldNormDef(S, S, K) + orc.surf_varifoldNorm0(S, K)) < 1e-12 * n0
    # orientation-blind (g^2), and differs from the measure term (which ignores normals)
    Tf = orc.Surface(surf=T)
    Tf.flipFaces()
    assert abs(orc.surf_varifoldNormDef(S, Tf, K) - orc.surf_varifoldNormDef(S, T, K)) < 1e-12 * n0
    assert abs(orc.surf_varifoldNorm0(T, K) - 2 * orc.surf_measureNorm0(T, K)) > 1e-3 * n0
    # coplanar faces (every g = 1 up to the 1e-10 offset in a): twice the measure norm
    Vp = np.array([[0., 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [2, 0.5, 0]])
    Fp = np.array([[0, 1, 2], [1, 3, 2], [1, 4, 3]])
    P = orc.Surface(FV=(Fp, Vp))
    m0 = orc.surf_measureNorm0(P, K)
    assert abs(orc.surf_varifoldNorm0(P, K) - 2 * m0) < 1e-9 * m0


def test_varifold_gradient_fd():
    rng = np.random.default_rng(23)
    V, F = synthetic.icosphere(1)
    mk = lambda X: orc.Surface(FV=(F, X))
    Y = synthetic.bump_target(V)
    K = orc.Kernel('gauss', sigma=0.4)
    X = V + 0.05 * rng.standard_normal(V.shape)
    g = orc.surf_varifoldNormGradient(mk(X), mk(Y), K)
    h = 1e-6
    for (i, c) in [(0, 0), (3, 1), (7, 2), (10, 1), (25, 0)]:
        e = np.zeros_like(X)
        e[i, c] = h
        fd = (orc.surf_varifoldNormDef(mk(X + e), mk(Y), K)
              - orc.surf_varifoldNormDef(mk(X - e), mk(Y), K)) / (2 * h)
        assert abs(fd - g[i, c]) < 1e-6 * max(1.0, abs(fd)), (i, c, fd, g[i, c])


def test_measure_matching_reduces_objective():
    V, F = synthetic.icosphere(1)
    Kd, Kw = orc.Kernel('gauss', sigma=0.5), orc.Kernel('gauss', sigma=0.3)
    S, T = orc.Surface(FV=(F, V)), orc.Surface(FV=(F, synthetic.bump_target(V)))
    o = orc.SurfaceMatching(S, T, Kd, Kw, sigmaError=0.1, timeStep=0.25, maxIter=4,
                            errorType='measure')
    e0 = o.objectiveFun()
    o.optimizeMatching()
    assert o.obj < 0.5 * e0


# ---------------------------------------------------------------- landmarks
def test_landmark_data_term_closed_form_and_fd():
    rng = np.random.default_rng(13)
    X, Y = rng.standard_normal((30, 3)), rng.standard_normal((30, 3))
    A, B = orc.PointSet(X), orc.PointSet(Y)
    assert orc.landmark_normDef(A, A, None) == 0.0 and orc.landmark_norm0(B, None) == 0.0
    one = orc.landmark_normDef(orc.PointSet(X[:1]), orc.PointSet(Y[:1]), None)
    assert abs(one - np.linalg.norm(X[0] - Y[0]) ** 2) < 1e-14 * one
    g = orc.landmark_normGradient(A, B, None)
    h = 1e-6
    for (i, c) in [(0, 0), (5, 2), (29, 1)]:
        e = np.zeros_like(X)
        e[i, c] = h
        fd = (orc.landmark_normDef(orc.PointSet(X + e), B, None)
              - orc.landmark_normDef(orc.PointSet(X - e), B, None)) / (2 * h)
        assert abs(fd - g[i, c]) < 1e-7 * max(1.0, abs(fd))


def test_landmark_matching_gradient_is_adjoint():
    """Directional derivative of the landmark objective (central FD) equals the RKHS product
    of getGradient with the direction: pins the sign and the 2/sigma^2 of p_T."""
    rng = np.random.default_rng(14)
    X = synthetic.ball_points(rng, 60)
    Y = X + 0.1 * rng.standard_normal(X.shape)
    o = orc.LandmarkMatching(X, Y, orc.Kernel('gauss', sigma=0.4), sigmaError=0.2, timeStep=0.2)
    o.at = 0.05 * rng.standard_normal(o.at.shape)
    o.objectiveFun()
    o.endOfIteration()
    g = o.getGradient(1.0)
    d = orc.Direction()
    d.diff = rng.standard_normal(o.at.shape)
    h = 1e-5
    fd = (o.updateTry(d, -h) - o.updateTry(d, h)) / (2 * h)
    [gd] = o.dotProduct(g, [d])
    assert abs(fd - gd) < 1e-6 * abs(fd)


def test_landmark_matching_reduces_objective():
    rng = np.random.default_rng(15)
    X = synthetic.ball_points(rng, 50)
    K = orc.Kernel('gauss', sigma=0.4)
    xt, _ = orc.landmarkEPDiff(5, X, 0.3 * rng.standard_normal(X.shape), K)
    o = orc.LandmarkMatching(X, xt[-1], K, sigmaError=0.1, timeStep=0.2, maxIter=10)
    h = o.optimizeMatching()
    assert len(h) >= 3 and h[-1] < 0.2 * h[0]


# ---------------------------------------------------------------- optimiser
class _Quadratic:
    """f(v) = 0.5 v^T A v - b^T v with an SPD A: the optimiser must reach A^{-1} b."""

    def __init__(self, n=6, seed=0):
        rng = np.random.default_rng(seed)
        M = rng.standard_normal((n, n))
        self.A = M @ M.T + n * np.eye(n)
        self.b = rng.standard_normal(n)
        self.v = np.zeros(n)
        self.obj = None

    def getVariable(self):
        return self.v

    def objectiveFun(self):
        return 0.5 * self.v @ self.A @ self.v - self.b @ self.v

    def updateTry(self, d, eps, objRef=None):
        vt = self.v - eps * d
        o = 0.5 * vt @ self.A @ vt - self.b @ vt
        if objRef is None or o < objRef:
            self.vTry, self.objTry = vt, o
        return o

    def acceptVarTry(self):
        self.v = self.vTry

    def getGradient(self, coeff=1.0):
        return (self.A @ self.v - self.b) / coeff


def test_cg_restatement_minimises_quadratic():
    q = _Quadratic()
    q.gradEps = 1e-9
    orc.cg(q, maxIter=500, epsInit=0.01)
    np.testing.assert_allclose(q.v, np.linalg.solve(q.A, q.b), rtol=1e-3, atol=1e-4)


# ---------------------------------------------------------------- divergence and Jacobian flow
@pytest.mark.parametrize("order,sigma,tol", [(10, 0.6, 1e-8), (1, 0.5, 1e-6), (3, 0.4, 1e-6)])
def test_divergence_is_trace_of_velocity_gradient(order, sigma, tol):
    """applyDivergence = sum_c d v_c / d y_c of v(y) = K(y, x) a, by central FD of applyK
    (Laplacian: the dense exact-double polynomials differ from applyK's float32 tables at the
    1e-8 level, hence the looser bound)."""
    rng = np.random.default_rng(order)
    x, a = rng.standard_normal((30, 3)) * 0.5, rng.standard_normal((30, 3))
    y = rng.standard_normal((12, 3)) * 0.5
    K = orc.Kernel('gauss' if order > 4 else 'laplacian', sigma=sigma, order=order)
    div = K.applyDivergence(x, a, firstVar=y)
    h = 1e-5
    fd = np.zeros(len(y))
    for c in range(3):
        e = np.zeros(3)
        e[c] = h
        fd += (K.applyK(x, a, firstVar=y + e)[:, c] - K.applyK(x, a, firstVar=y - e)[:, c]) / (2 * h)
    np.testing.assert_allclose(div, fd, rtol=tol, atol=tol * np.abs(fd).max())
    # at the sources themselves the self term drops out (y_k - x_k = 0) whatever the diagonal
    d0 = K.applyDivergence(x, a)
    d1 = np.array([K.applyDivergence(np.delete(x, k, 0), np.delete(a, k, 0), firstVar=x[k:k + 1])[0]
                   for k in range(len(x))])
    np.testing.assert_allclose(d0, d1, rtol=1e-12, atol=1e-13 * np.abs(d0).max())


def test_jacobian_flow_tracks_log_det_of_the_flow_map():
    """J_T = sum_t dt div v_t approximates log det of the discrete flow map's Jacobian
    (measured by FD on a tracked point set) to O(dt)."""
    rng = np.random.default_rng(4)
    x0 = rng.standard_normal((20, 3)) * 0.4
    T = 40
    at = np.tile(0.3 * rng.standard_normal((1, 20, 3)), (T, 1, 1))
    K = orc.Kernel('gauss', sigma=0.5)
    y0 = rng.standard_normal((5, 3)) * 0.3
    xt, yt, Jt = orc.landmarkDirectEvolutionEuler(x0, at, K, withPointSet=y0, withJacobian=True)
    assert Jt.shape == (T + 1, 5) and np.all(Jt[0] == 0)
    h = 1e-6
    for i in range(5):
        Jm = np.zeros((3, 3))
        for c in range(3):
            e = np.zeros_like(y0)
            e[i, c] = h
            yp = orc.landmarkDirectEvolutionEuler(x0, at, K, withPointSet=y0 + e)[1][-1, i]
            ym = orc.landmarkDirectEvolutionEuler(x0, at, K, withPointSet=y0 - e)[1][-1, i]
            Jm[:, c] = (yp - ym) / (2 * h)
        logdet = np.log(np.linalg.det(Jm))
        assert abs(Jt[-1, i] - logdet) < 0.05 * max(abs(logdet), 0.05), (i, Jt[-1, i], logdet)
    # without a point set the Jacobian is taken at the moving points
    xt2, Jx = orc.landmarkDirectEvolutionEuler(x0, at, K, withJacobian=True)
    assert Jx.shape == (T + 1, 20) and np.allclose(xt2, xt)


# ---------------------------------------------------------------- applyDiffK family, dense DDiffK11 / 12
def test_diff_k_family_matches_reference_golden(golden_mat):
    for name, c in golden_mat.items():
        d, ny, nx, m, order, sigma = c["params"]
        for form in (1, 2, 3):
            f = orc.apply_diff_k(c["y"], c["dk_a1"], c["dk_a2"], sigma, int(order), form)
            ref = c[f"applykdiff_{form}"]
            np.testing.assert_allclose(f, ref, rtol=1e-11, atol=1e-12 * np.abs(ref).max(),
                                       err_msg=f"{name} form {form}")


def test_diff_k_forms_add_up_and_match_diff_kt():
    """applyDiffK1and2 = applyDiffK + applyDiffK2, and applyDiffK with a1 = e_c equals
    applyDiffKT's c-th column form by symmetry of the pair sum."""
    rng = np.random.default_rng(3)
    x, a1, a2 = rng.standard_normal((40, 3)), rng.standard_normal((40, 3)), rng.standard_normal((40, 3))
    for order in (10, 2):
        f1, f2, f3 = (orc.apply_diff_k(x, a1, a2, 0.6, order, f) for f in (1, 2, 3))
        np.testing.assert_allclose(f3, f1 + f2, rtol=1e-12, atol=1e-13 * np.abs(f3).max())
        # applyDiffK row k, component c: sum_l 2K' <dx, a1_k> a2_l[c]; applyDiffKT with
        # a1' = a2[:, c] (sources) and a1_k as the target weight gives the same contraction
        for c in range(3):
            w = np.zeros((40, 3))
            w[:, 0] = a2[:, c]
            g = orc.apply_diff_kt(x, x, a1[None, :, :1] * 0 + 1, w[None, :, :1], 0.6, order)
            np.testing.assert_allclose((g * a1).sum(axis=1), f1[:, c], rtol=1e-11,
                                       atol=1e-12 * np.abs(f1).max())


@pytest.mark.parametrize("order,sigma,tol", [(10, 0.7, 1e-12), (3, 0.5, 1e-6), (4, 0.6, 1e-6)])
def test_dense_ddiff_k11_k12(order, sigma, tol):
    """The dense siblings add up to applyDDiffK11and12 (Fortran) and are the target / source
    derivatives of applyDiffKT along p (central FD)."""
    rng = np.random.default_rng(order)
    x, n, a, p = (rng.standard_normal((30, 3)) * s for s in (0.5, 1, 1, 1))
    K = orc.Kernel('gauss' if order > 4 else 'laplacian', sigma=sigma, order=order)
    k11, k12 = K.applyDDiffK11(x, n, a, p), K.applyDDiffK12(x, n, a, p)
    both = K.applyDDiffK11and12(x, n, a, p)
    np.testing.assert_allclose(k11 + k12, both, rtol=tol, atol=tol * np.abs(both).max())
    h = 1e-6
    fd11 = (K.applyDiffKT(x, n[None], a[None], firstVar=x + h * p)
            - K.applyDiffKT(x, n[None], a[None], firstVar=x - h * p)) / (2 * h)
    fd12 = (K.applyDiffKT(x + h * p, n[None], a[None], firstVar=x)
            - K.applyDiffKT(x - h * p, n[None], a[None], firstVar=x)) / (2 * h)
    np.testing.assert_allclose(k11, fd11, rtol=1e-6, atol=1e-6 * np.abs(fd11).max())
    np.testing.assert_allclose(k12, fd12, rtol=1e-6, atol=1e-6 * np.abs(fd12).max())


def test_dense_kernel_matrices_reproduce_the_sums():
    """The dense K matrix times the weights is applyK (Gaussian: the Fortran's cut-off only
    matters within 1e-4 sigma); the dense K' matrix gives applyDivergence's reference form."""
    rng = np.random.default_rng(6)
    x, a, y = rng.standard_normal((50, 3)), rng.standard_normal((50, 3)), rng.standard_normal((20, 3))
    K = orc.kernel_matrix(y, x, 0.6, 10, False)
    np.testing.assert_allclose(K @ a, orc.apply_k(y, x, a, 0.6, 10), rtol=1e-13, atol=1e-14)
    Ks = orc.kernel_matrix(None, x, 0.6, 3, True)
    np.testing.assert_allclose(Ks @ a, orc.apply_k(x, x, a, 0.6, 3), rtol=1e-6, atol=1e-7)
    assert np.allclose(np.diag(Ks), 1.0) and np.allclose(Ks, Ks.T)


# ---------------------------------------------------------------- second-order flow
def _second_order_case(seed=5, N=18, T=8):
    rng = np.random.default_rng(seed)
    x0 = rng.standard_normal((N, 3)) * 0.5
    a0 = 0.3 * rng.standard_normal((N, 3))
    rhot = 0.2 * rng.standard_normal((T, N, 3))
    return x0, a0, rhot


@pytest.mark.parametrize("name,order", [("gauss", 10), ("laplacian", 3)])
def test_second_order_gradient_is_the_adjoint(name, order):
    """secondOrderGradient differentiates 1/2 <a0, K(x0) a0> + sum_t dt cw |rho_t|^2 / 2
    - <px1[0], x_{T/2}> - <px1[1], x_T> (the objective of secondOrderMatching.py:204-212 with
    linear end-point terms): da0 is the Euclidean gradient in a0, drhot the one in rho_t
    divided by dt.  Zero affine parts take the reference's affine branch, as its driver
    does."""
    K = orc.Kernel(name, sigma=0.6, order=order)
    x0, a0, rhot = _second_order_case()
    T, N = rhot.shape[0], x0.shape[0]
    rng = np.random.default_rng(9)
    px1 = [rng.standard_normal((N, 3)), rng.standard_normal((N, 3))]
    pa1 = [np.zeros((N, 3)), np.zeros((N, 3))]
    cw = 0.7
    aff = (np.zeros((T, 3, 3)), np.zeros((T, 3)))

    def J(a0_, rho_):
        xt, at = orc.secondOrderEvolution(x0, a0_, rho_, K, affine=aff)
        return (0.5 * (a0_ * K.applyK(x0, a0_)).sum() + sum(cw * (r ** 2).sum() / 2 for r in rho_) / T
                - (px1[0] * xt[T // 2]).sum() - (px1[1] * xt[T]).sum())

    da0, drhot, dA, db, xt, at = orc.secondOrderGradient(x0, a0, rhot, px1, pa1, K, affine=aff,
                                                         controlWeight=cw)
    h = 1e-6
    d = rng.standard_normal(a0.shape)
    fd = (J(a0 + h * d, rhot) - J(a0 - h * d, rhot)) / (2 * h)
    assert abs(fd - (da0 * d).sum()) < 1e-6 * max(1.0, abs(fd))
    dr = rng.standard_normal(rhot.shape)
    fd = (J(a0, rhot + h * dr) - J(a0, rhot - h * dr)) / (2 * h)
    assert abs(fd - (drhot * dr).sum() / T) < 1e-6 * max(1.0, abs(fd))
    # without affine parts the reference leaves drhot at zero
    assert not orc.secondOrderGradient(x0, a0, rhot, px1, pa1, K)[1].any()


def test_second_order_hamiltonian_drives_the_evolution():
    """dH/dpx = K(x) a and dH/dpa = -applyDiffKT(x,[a],[a]) + rho: the evolution's velocities,
    by FD of the Hamiltonian."""
    K = orc.Kernel('gauss', sigma=0.5)
    x0, a0, rhot = _second_order_case(N=10)
    rng = np.random.default_rng(2)
    px, pa = rng.standard_normal(x0.shape), rng.standard_normal(x0.shape)
    h = 1e-6
    for c, v in ((0, K.applyK(x0, a0)),
                 (1, -K.applyDiffKT(x0, a0[None], a0[None]) + rhot[0])):
        e = rng.standard_normal(x0.shape)
        args = lambda s: (px + s * e, pa) if c == 0 else (px, pa + s * e)
        fd = (orc.secondOrderHamiltonian(x0, a0, rhot[0], *args(h), K)
              - orc.secondOrderHamiltonian(x0, a0, rhot[0], *args(-h), K)) / (2 * h)
        assert abs(fd - (v * e).sum()) < 1e-6 * max(1.0, abs(fd))


# ---------------------------------------------------------------- global affine part
def test_affine_bases_are_orthonormal_and_span_their_groups():
    for d in (2, 3):
        for aff, dim in (('affine', d * d + d), ('similitude', d * (d - 1) // 2 + 1 + d),
                         ('euclidean', d * (d - 1) // 2 + d), ('translation', d)):
            B = orc.AffineBasis(d, aff)
            assert B.affineDim == dim
            np.testing.assert_allclose(B.basis.T @ B.basis, np.eye(dim), atol=1e-15)
            for k in B.rotComp:   # rotations: antisymmetric A, no translation
                A = B.basis[:d * d, k].reshape(d, d)
                assert np.allclose(A, -A.T) and not B.basis[d * d:, k].any()
            for k in B.affComp:   # the symmetric trace-free part
                A = B.basis[:d * d, k].reshape(d, d)
                assert np.allclose(A, A.T) and abs(np.trace(A)) < 1e-15
        assert orc.AffineBasis(d, 'none').affineDim == 0


@pytest.mark.parametrize("kind", ["surface", "curve"])
def test_affine_matching_gradient_is_the_derivative(kind):
    """The affine part of getGradient is the derivative of the objective in Afft under the
    drivers' own scalings (surface: grd.aff coeffAff coeff; curve: grd.aff affineWeight coeff;
    the penalty's dt already carries the 1/T), by central FD of updateTry."""
    rng = np.random.default_rng(4)
    if kind == "surface":
        V, F = synthetic.icosphere(1)
        o = orc.SurfaceMatching(orc.Surface(FV=(F, V)), orc.Surface(FV=(F, synthetic.bump_target(V))),
                                orc.Kernel('gauss', sigma=0.5), orc.Kernel('gauss', sigma=0.3),
                                sigmaError=0.1, timeStep=0.25, affine='similitude', affineWeight=2.0)
    else:
        X, _ = synthetic.circle(30)
        E = synthetic.closed_curve_edges(30)
        o = orc.CurveMatching(orc.Curve(FV=(E, X)), orc.Curve(FV=(E, 1.2 * synthetic.star_curve(30, 4, 0.2))),
                              orc.Kernel('gauss', sigma=0.4), orc.Kernel('gauss', sigma=0.2),
                              sigmaError=0.1, timeStep=0.25, affine='affine', affineWeight=2.0)
    o.at = 0.05 * rng.standard_normal(o.at.shape)
    o.Afft = 0.1 * rng.standard_normal(o.Afft.shape)
    o.objectiveFun()
    o.endOfIteration()
    g = o.getGradient(1.0)
    d = orc.Direction()
    d.diff = np.zeros_like(o.at)
    d.aff = rng.standard_normal(o.Afft.shape)
    h = 1e-6
    fd = (o.updateTry(d, -h) - o.updateTry(d, h)) / (2 * h)
    if kind == "surface":
        gd = (g.aff * d.aff).sum() * o.coeffAff
    else:
        gd = (g.aff * d.aff * o.affineWeight.reshape(1, -1)).sum()
    assert abs(fd - gd) < 1e-6 * max(1.0, abs(fd)), (fd, gd)


def test_affine_surface_matching_recovers_a_scaling():
    """A pure scaling of the template is matched mostly by the similitude part."""
    V, F = synthetic.icosphere(1)
    o = orc.SurfaceMatching(orc.Surface(FV=(F, V)), orc.Surface(FV=(F, 1.3 * V)),
                            orc.Kernel('gauss', sigma=0.5), orc.Kernel('gauss', sigma=0.3),
                            sigmaError=0.1, timeStep=0.25, maxIter=10, affine='similitude')
    e0 = o.objectiveFun()
    o.optimizeMatching()
    assert o.obj < 0.2 * e0
    B = o.affB
    assert np.abs(o.Afft[:, B.simComp]).sum() > 0
