"""The tracking restatements (oracle.camera_step_grad / adam6 /
apply_camera_step / photometric of crates/gs-track/src/lib.rs:276-329 and
photometric.wgsl) pinned by the FP64 oracle: the Riemannian rotation
gradient is the derivative of the loss along the right perturbation the step
applies (central differences of the FP64 render), and a small step along
-gradient lowers the loss."""

import numpy as np

import oracle
from microscene import make_scene, make_weights


def _loss(scene, cam, w):
    return float((oracle.render(scene, cam)["color"] * w).sum())


def test_riemannian_gradient_is_the_derivative_along_the_applied_perturbation():
    """The right perturbation R' = R exp([delta e_k]x) that apply_camera_step
    uses (delta = -step) is q' = q * (sin(delta/2) e_k, cos(delta/2)): the
    central difference of the FP64 render loss along it equals the FP64
    quaternion gradient (FD-pinned in test_oracle.py) contracted with
    q * (e_k / 2, 0) -- the quantity the next test shows camera_step_grad
    computes from dL/dR."""
    for seed, n, deg, size in ((11, 12, 0, 24), (37, 10, 1, 24)):
        scene, cam = make_scene(seed, n, deg, size)
        w = make_weights(seed ^ 0xABCDEF, size * size).reshape(size, size, 3)
        q = cam.quat / np.linalg.norm(cam.quat)
        cam.quat = q
        g = oracle.gradients(scene, cam, w)
        h = 1e-6
        fd = []
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            ang = h
            ep = np.append(e / ang * np.sin(ang / 2), np.cos(ang / 2))
            em = np.append(-e / ang * np.sin(ang / 2), np.cos(ang / 2))
            from splattar_amd.scenes import quat_mul
            cp, cm = cam.copy(), cam.copy()
            cp.quat = quat_mul(q, ep)
            cm.quat = quat_mul(q, em)
            fd.append((_loss(scene, cp, w) - _loss(scene, cm, w)) / (2 * h))
        # dL/dq of the oracle -> the same directional derivative: d/dt L(q exp(t e_k / 2))
        # = dL/dq . (q * (e_k / 2, 0)).
        for k in range(3):
            e = np.zeros(4)
            e[k] = 0.5
            dq = quat_mul(q, e)
            want = float(np.dot(g.cam_quat, dq))
            assert abs(fd[k] - want) <= 1e-7 + 1e-4 * max(abs(fd[k]), abs(want)), (seed, k, fd[k], want)


def test_camera_step_grad_matches_the_quaternion_chain():
    """camera_step_grad(dL/dR, q)[k] = dL/dq . (q * (e_k / 2, 0)) for any
    dL/dR (the chain of math.rs quat_grad restricted to unit q)."""
    rng = np.random.default_rng(4)
    from splattar_amd.scenes import quat_mul
    for _ in range(20):
        q = rng.normal(0, 1, 4)
        q /= np.linalg.norm(q)
        G = rng.normal(0, 1, (3, 3))  # dL/dR[j][k]
        gq = oracle.quat_grad(q, G)
        grad_cam = np.zeros(12, np.float32)
        grad_cam[:9] = G.T.reshape(9)  # column-major
        got = oracle.camera_step_grad(grad_cam, q.astype(np.float32))
        for k in range(3):
            e = np.zeros(4)
            e[k] = 0.5
            want = float(np.dot(gq, quat_mul(q, e)))
            assert abs(got[k] - want) <= 2e-5 * (1 + abs(want)), (k, got[k], want)


def test_adam6_first_step_and_photometric_closed_forms():
    m, v = np.zeros(6, np.float32), np.zeros(6, np.float32)
    g = np.array([1e-3, -2e-3, 0, 5.0, -1e-6, 0], np.float32)
    s = oracle.adam6(m, v, 1, g, 2e-3, 6e-3)
    # First Adam step: lr g / (|g| + 1e-10) (bias-corrected m = g, sqrt(v) = |g|), 0 where g = 0.
    want = [2e-3 * 1e-3 / (1e-3 + 1e-10), -2e-3 * 2e-3 / (2e-3 + 1e-10), 0, 6e-3 * 5 / (5 + 1e-10),
            -6e-3 * 1e-6 / (1e-6 + 1e-10), 0]
    np.testing.assert_allclose(s, want, rtol=2e-5, atol=1e-12)  # f32 bias corrections
    r = np.zeros((4, 4), np.float32)
    r[:, :3] = [[0.5, 0.5, 0.5], [0.2, 0.3, 0.4], [0.0, 1.0, 0.5], [0.7, 0.7, 0.7]]
    r[:, 3] = [1.0, 0.6, 0.2, 0.0]
    t = np.zeros((4, 4), np.float32)
    t[:, :3] = [[0.5, 0.5, 0.50005], [0.1, 0.3, 0.6], [0.0, 0.0, 0.0], [0.7, 0.7, 0.7]]
    dl, res, cov = oracle.photometric(r, t)
    assert abs(res - (0.00005 + 0.1 + 0.2 + 1.0 + 0.5) / 12.0) < 1e-7 and cov == 0.5
    # softened sign: |d| = 5e-5 < 1e-4 gives -0.5 / 3P; saturated ones +-1 / 3P; d = 0 gives 0.
    d0 = float(np.float32(0.5) - np.float32(0.50005))
    np.testing.assert_allclose(dl[0, :3], [0, 0, d0 / 1e-4 / 12], rtol=1e-5, atol=1e-9)
    np.testing.assert_allclose(dl[1, :3], [1 / 12, 0, -1 / 12], atol=1e-7)
    assert np.all(dl[:, 3] == 0)
