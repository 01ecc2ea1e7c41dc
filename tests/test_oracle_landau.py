"""Pins of the oracle's Landau matching (restatement of solveEigenSystem,
src/LandauMatch.cpp:145-331, with GSL's gsl_eigen_nonsymmv restated).

Independent references: numpy.linalg.eig (LAPACK dgeev: balancing,
Hessenberg, Francis QR, unit-norm eigenvectors -- the same published
algorithm family), synthetic tensors with a known Landau frame, and the 2x2
closed form e = (a-c)/2 + sqrt((a+c)^2/4 - b^2)."""
import numpy as np

import oracle as O


def test_eig4_matches_numpy():
    rng = np.random.default_rng(10)
    for _ in range(500):
        A = rng.normal(size=(4, 4))
        wr, wi, V, ok = O.eig4(A)
        assert ok
        ref = np.sort_complex(np.linalg.eigvals(A))
        got = np.sort_complex(wr + 1j * wi)
        assert np.abs(ref - got).max() < 1e-10
        for k in range(4):
            if wi[k] == 0.0:
                v = V[:, k]
                assert abs(np.linalg.norm(v) - 1.0) < 1e-12
                assert np.abs(A @ v - wr[k] * v).max() < 1e-10


def _boost(u):
    """Lorentz boost matrix taking (1,0,0,0) to u (orthonormal frame)."""
    g = u[0]
    v = u[1:]
    L = np.eye(4)
    L[0, 0] = g
    L[0, 1:] = v
    L[1:, 0] = v
    vv = v @ v
    if vv > 0:
        L[1:, 1:] += np.outer(v, v) * (g - 1) / vv
    return L


def _synthetic(rng, tau, gamma_max=8.0, two_d=False):
    """T^{mu nu} (Milne components) with a known Landau frame: rest-frame
    diag(e, P1, P2, P3) rotated and boosted."""
    e = rng.uniform(0.5, 50.0)
    P = rng.uniform(0.0, 1.0, 3) * e / 3 * rng.uniform(0.2, 1.0)
    if two_d:
        P[2] = 0.0
    R = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    if two_d:
        th = rng.uniform(0, 2 * np.pi)
        R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    Srest = np.zeros((4, 4))
    Srest[0, 0] = e
    Srest[1:, 1:] = R @ np.diag(P) @ R.T
    rap = rng.uniform(0, np.arccosh(gamma_max))
    n = rng.normal(size=3)
    if two_d:
        n[2] = 0.0
    n /= np.linalg.norm(n)
    u = np.concatenate([[np.cosh(rap)], np.sinh(rap) * n])
    L = _boost(u)
    A = L @ Srest @ L.T  # orthonormal-frame T
    S = np.diag([1.0, 1.0, 1.0, 1.0 / tau])
    T = S @ A @ S  # Milne: T^{eta .} = A^{3 .}/tau
    t10 = np.array([T[0, 0], T[0, 1], T[0, 2], T[0, 3], T[1, 1], T[1, 2], T[1, 3], T[2, 2], T[2, 3], T[3, 3]])
    u_milne = u.copy()
    u_milne[3] /= tau
    return t10, e, u_milne


def test_landau_recovers_known_frame():
    rng = np.random.default_rng(11)
    tau = np.float32(0.9)
    for two_d in (False, True):
        Ts, es, us = zip(*[_synthetic(rng, float(tau), two_d=two_d) for _ in range(300)])
        e, u = O.landau(np.array(Ts), tau)
        us = np.array(us)
        assert np.abs(e - np.array(es)).max() / np.max(es) < 1e-9
        # the synthetic tensor carries its own rounding; the frame is
        # conditioned like gamma^2, so compare relative to u^tau
        assert (np.abs(u - us).max(axis=1) / us[:, 0]).max() < 1e-6
        # and against LAPACK's eigenvectors of the same float matrix
        for t, ue in zip(Ts, u):
            A = np.array([[t[0], t[1], t[2], t[3]], [t[1], t[4], t[5], t[6]],
                          [t[2], t[5], t[7], t[8]], [t[3], t[6], t[8], t[9]]])
            M = A * np.array([1.0, -1.0, -1.0, -float(tau) * float(tau)])[None, :]
            w, V = np.linalg.eig(M)
            k = int(np.argmax(w.real))
            v = V[:, k].real
            ml = v[0] ** 2 - (v[1] ** 2 + v[2] ** 2 + float(tau) ** 2 * v[3] ** 2)
            ref = v / np.sqrt(ml) * np.sign(v[0])
            assert np.abs(ue - ref).max() < 1e-11 * ref[0]


def test_landau_2x2_closed_form():
    rng = np.random.default_rng(12)
    a = rng.uniform(1, 10, 200)
    c = a * rng.uniform(0.1, 0.9, 200)
    b = (a + c) / 2 * rng.uniform(-0.95, 0.95, 200)
    T = np.zeros((200, 10))
    T[:, 0], T[:, 1], T[:, 4] = a, b, c
    e, u = O.landau(T, np.float32(1.0))
    e_ref = (a - c) / 2 + np.sqrt((a + c) ** 2 / 4 - b ** 2)
    assert np.abs(e - e_ref).max() < 1e-12 * a.max()
    assert np.abs(u[:, 1] / u[:, 0] - (a - e_ref) / b).max() < 1e-9


def test_landau_fallback_rules():
    # zero tensor: no positive eigenvalue -> e = 1e-3, u = (1,0,0,0) (:283-291)
    e, u = O.landau(np.zeros((1, 10)), np.float32(0.6))
    assert e[0] == 1e-3 and np.all(u[0] == [1, 0, 0, 0])
    # gamma above 100 is rejected (:243)
    rng = np.random.default_rng(13)
    t10, _, _ = _synthetic(rng, 0.6, gamma_max=1.0)
    S = np.zeros((4, 4))
    S[0, 0], S[1, 1], S[2, 2], S[3, 3] = 10.0, 3.0, 3.0, 3.0
    L = _boost(np.array([150.0, np.sqrt(150.0 ** 2 - 1), 0, 0]))
    A = L @ S @ L.T
    t10 = np.array([A[0, 0], A[0, 1], A[0, 2], A[0, 3], A[1, 1], A[1, 2], A[1, 3], A[2, 2], A[2, 3], A[3, 3]])
    e, u = O.landau(t10[None], np.float32(1.0))
    assert e[0] == 1e-3 and u[0, 0] == 1.0
