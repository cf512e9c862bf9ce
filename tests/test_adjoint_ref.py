"""Pins of the FP64 host adjoint (oracle/adjoint_ref.cpp): its reverse sweep against central finite differences of its
own double-precision forward simulator, for every material parameter and for initial velocities, across windows that
carry tangential history over; and the rotation statements alone against finite differences in both of their branches
(series for small angles, closed form otherwise)."""
import ctypes as C

import numpy as np
import pytest

import pyoracle
from deme_amd import scenes


def _loss_weights(n, seed=0):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, 3)), rng.normal(size=(n, 3)) * 1e-3


@pytest.fixture(scope="module")
def problem():
    n, steps, window = 60, 60, 20
    sc = scenes.adjoint_drop(n=n, steps=steps, window=window, seed=7)
    A = pyoracle.AdjointFP64(sc, steps, window)
    wX, wV = _loss_weights(n)
    gV, gM, X, V = A.gradient(wX, wV)
    return A, wX, wV, gV, gM, X, V


def _L(wX, wV, X, V):
    return float((wX * X).sum() + (wV * V).sum())


def test_forward_moves_and_touches(problem):
    A, _, _, _, _, X, V = problem
    X2, V2 = A.forward()
    assert np.array_equal(X, X2) and np.array_equal(V, V2)  # the gradient run's forward is the plain forward
    assert np.abs(X - A.X0).max() > 1e-4  # spheres moved by more than their sub-voxel noise


@pytest.mark.parametrize("prop", ["E", "nu", "CoR", "mu"])
def test_material_gradients_match_central_differences(problem, prop):
    A, wX, wV, _, gM, _, _ = problem
    k = ("E", "nu", "CoR", "mu").index(prop)
    base = A.mats[prop]
    nonzero = 0
    for m in range(len(base)):
        h = 1e-4 * base[m]
        p, q = base.copy(), base.copy()
        p[m] += h
        q[m] -= h
        fd = (_L(wX, wV, *A.forward(**{prop: p})) - _L(wX, wV, *A.forward(**{prop: q}))) / (2 * h)
        scale = max(abs(fd), abs(gM[m, k]))
        if scale == 0.0:
            continue
        nonzero += 1
        assert abs(fd - gM[m, k]) <= 1e-5 * scale, (prop, m, fd, gM[m, k])
    assert nonzero >= 2


def test_velocity_gradients_match_central_differences(problem):
    A, wX, wV, gV, _, _, _ = problem
    h = 1e-5  # m/s; truncation of the central difference stays below 1e-6 relative
    for o in (0, 2, 4, 7, 9, 11, 30, 59):
        for d in range(3):
            p, q = A.V0.copy(), A.V0.copy()
            p[o, d] += h
            q[o, d] -= h
            fd = (_L(wX, wV, *A.forward(V0=p)) - _L(wX, wV, *A.forward(V0=q))) / (2 * h)
            assert abs(fd - gV[o, d]) <= 1e-6 * max(abs(fd), 1e-3), (o, d, fd, gV[o, d])


def test_free_flight_gradient_is_the_closed_form():
    # one sphere far from everything: X_T = X_0 + T h v_0 + g-terms, so dX_T/dv_0 = T h exactly
    sc = scenes.spheres_scene([[0.0, 0.0, 0.0]], 0.002, (0.2, 0.2, 0.2), planes=False, vel=[[0.1, -0.2, 0.3]],
                              h=1e-5)
    A = pyoracle.AdjointFP64(sc, 50, 10)
    wX = np.array([[1.0, 2.0, -3.0]])
    gV, gM, _, _ = A.gradient(wX, None)
    assert np.allclose(gV, wX * 50 * float(sc.params.h), rtol=1e-12, atol=0)  # h as the simulator holds it
    assert np.all(gM == 0.0)


@pytest.mark.parametrize("scale", [1e-4, 3e-3, 0.3, 2.0])
def test_rotation_reverse_matches_central_differences(scale):
    lib = pyoracle.load()
    lib.ora_adj_rot.argtypes = [C.c_void_p] * 6
    rng = np.random.default_rng(int(scale * 1e4))
    th = rng.normal(size=3)
    th *= scale / np.linalg.norm(th)
    u, lY = rng.normal(size=3), rng.normal(size=3)

    def rot(t, v):
        y, a, b = np.zeros(3), np.zeros(3), np.zeros(3)
        t, v = np.ascontiguousarray(t), np.ascontiguousarray(v)
        lib.ora_adj_rot(t.ctypes.data, v.ctypes.data, lY.ctypes.data, y.ctypes.data, a.ctypes.data, b.ctypes.data)
        return y, a, b

    y, lU, lTh = rot(th, u)
    # forward against Rodrigues' formula written with the rotation matrix
    phi = np.linalg.norm(th)
    k = th / phi
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    R = np.eye(3) + np.sin(phi) * K + (1 - np.cos(phi)) * K @ K
    assert np.allclose(y, R @ u, rtol=0, atol=1e-12)
    assert np.allclose(lU, R.T @ lY, rtol=0, atol=1e-12)  # the adjoint of a rotation is its transpose
    eps = 1e-6 * max(scale, 1e-2)
    fd = np.zeros(3)
    for d in range(3):
        e = np.zeros(3)
        e[d] = eps
        fd[d] = (lY @ rot(th + e, u)[0] - lY @ rot(th - e, u)[0]) / (2 * eps)
    assert np.allclose(lTh, fd, rtol=1e-6, atol=1e-9 * max(1.0, np.abs(fd).max()))


def test_fp64_forward_follows_the_oracle_trajectory():
    # the differentiated model is the simulator: the FP64 forward stays on the oracle's float trajectory (reference
    # force model and extended-Taylor integration, contact list rebuilt every window with the same drift margin)
    n, steps, window = 200, 60, 20
    sc = scenes.adjoint_drop(n=n, steps=steps, window=window, seed=7)
    A = pyoracle.AdjointFP64(sc, steps, window)
    X, V = A.forward()
    o = pyoracle.OracleSystem(sc)
    o.step(steps, window)
    assert o.counts().nContacts > n // 4
    st = o.get_state()
    Xo = pyoracle.positions_fp64(sc.params, st["voxelID"], st["locX"], st["locY"], st["locZ"])[:n]
    Vo = np.stack([st["vX"], st["vY"], st["vZ"]], 1)[:n]
    assert np.abs(X - A.X0).max() > 1e-4
    assert np.abs(X - Xo).max() <= 1e-8  # m; the spheres move by ~1 mm
    assert np.abs(V - Vo).max() <= 1e-5  # m/s; speeds up to ~1.5 m/s
