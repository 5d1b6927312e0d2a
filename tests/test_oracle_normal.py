"""The FP64 normal-consistency loss (oracle.normal_loss, a restatement of
first_attempt/crates/gs-train/src/shaders/normal_loss.wgsl:52-145) pinned by
central differences, a closed form and invariances."""

import numpy as np
import pytest

import oracle

FOCAL = 40.0


def _inputs(W=13, H=11, seed=3):
    rng = np.random.default_rng(seed)
    ys, xs = np.mgrid[0:H, 0:W]
    alpha = rng.uniform(0.4, 1.0, (H, W))
    t = 2.0 + 0.3 * np.sin(xs / 3.0) + 0.2 * np.cos(ys / 4.0) + rng.normal(0, 0.01, (H, W))
    n = np.stack([rng.normal(0, 0.3, (H, W)), rng.normal(0, 0.3, (H, W)), np.ones((H, W))], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    color4 = np.zeros((H, W, 4))
    color4[..., 3] = alpha
    aux4 = np.zeros((H, W, 4))
    aux4[..., 1] = alpha * t
    normal4 = np.zeros((H, W, 4))
    normal4[..., :3] = n * alpha[..., None] * rng.uniform(0.5, 1.5, (H, W, 1))
    return color4, aux4, normal4


def _loss(c4, a4, n4, lam):
    return lam * oracle.normal_loss(c4, a4, n4, FOCAL, lam)[0].sum()


def test_normal_loss_gradients_match_central_differences():
    """dL/d(depth_acc) and dL/d(normal image), alpha and the orientation
    sign held (the reference detaches both), against central differences
    of the loss (h = 1e-6)."""
    c4, a4, n4 = _inputs()
    lam = 0.7 / (13 * 11)
    _, dn, dd = oracle.normal_loss(c4, a4, n4, FOCAL, lam)
    rng = np.random.default_rng(1)
    h = 1e-6
    for _ in range(40):
        y, x = int(rng.integers(0, 11)), int(rng.integers(0, 13))
        ap, am = a4.copy(), a4.copy()
        ap[y, x, 1] += h
        am[y, x, 1] -= h
        fd = (_loss(c4, ap, n4, lam) - _loss(c4, am, n4, lam)) / (2 * h)
        assert abs(fd - dd[y, x]) <= 1e-9 + 1e-5 * max(abs(fd), abs(dd[y, x])), (y, x, fd, dd[y, x])
        k = int(rng.integers(0, 3))
        np_, nm = n4.copy(), n4.copy()
        np_[y, x, k] += h
        nm[y, x, k] -= h
        fd = (_loss(c4, a4, np_, lam) - _loss(c4, a4, nm, lam)) / (2 * h)
        assert abs(fd - dn[y, x, k]) <= 1e-9 + 1e-5 * max(abs(fd), abs(dn[y, x, k])), (y, x, k)
    assert np.abs(dd).max() > 1e-4 and np.abs(dn).max() > 1e-4  # non-vacuous


def test_planar_depth_with_its_own_normal_has_zero_loss_and_gradient():
    """Depths of a tilted plane n.X = d seen by each pixel ray and a normal
    image equal to the camera-facing plane normal: every depth normal equals
    the plane normal, so the loss and both gradients vanish."""
    W, H = 12, 9
    nrm = np.array([0.2, -0.3, 1.0])
    nrm /= np.linalg.norm(nrm)
    d = -2.5  # plane n.X = d in front of the camera (X.z < 0)
    ys, xs = np.mgrid[0:H, 0:W]
    rays = np.stack([(xs + 0.5 - W / 2) / FOCAL, -(ys + 0.5 - H / 2) / FOCAL, -np.ones((H, W))], -1)
    t = d / (rays @ nrm)
    assert np.all(t > 0)
    alpha = np.full((H, W), 0.8)
    c4 = np.zeros((H, W, 4))
    c4[..., 3] = alpha
    a4 = np.zeros((H, W, 4))
    a4[..., 1] = alpha * t
    facing = nrm if np.dot(nrm, rays[0, 0] * t[0, 0]) < 0 else -nrm
    n4 = np.zeros((H, W, 4))
    n4[..., :3] = facing * 0.8
    lm, dn, dd = oracle.normal_loss(c4, a4, n4, FOCAL, 1.0)
    assert np.abs(lm).max() < 1e-12
    assert np.abs(dn).max() < 1e-12 and np.abs(dd).max() < 1e-10
    # The same plane with a wrong normal image: loss 1 - cos(angle) per pixel.
    n4[..., :3] = np.array([0.0, 0.0, 1.0]) * 0.8
    lm2, _, _ = oracle.normal_loss(c4, a4, n4, FOCAL, 1.0)
    want = (1.0 - facing[2]) * 0.8
    assert np.allclose(lm2[:-1, :-1], want, atol=1e-12) and np.all(lm2[-1, :] == 0) and np.all(lm2[:, -1] == 0)


def test_normal_image_scale_and_masking():
    """The rendered normal enters normalized (a positive per-pixel scale
    changes nothing); a pixel with alpha < 0.2 switches off the three
    stencils that read it; lambda = 0 gives zeros everywhere."""
    c4, a4, n4 = _inputs(seed=5)
    lm, _, _ = oracle.normal_loss(c4, a4, n4, FOCAL, 0.01)
    n4s = n4.copy()
    n4s[..., :3] *= np.random.default_rng(2).uniform(0.3, 3.0, n4.shape[:2] + (1,))
    lm_s, _, _ = oracle.normal_loss(c4, a4, n4s, FOCAL, 0.01)
    np.testing.assert_allclose(lm_s, lm, rtol=1e-12, atol=1e-15)
    c4m = c4.copy()
    c4m[5, 6, 3] = 0.1
    lmm, _, _ = oracle.normal_loss(c4m, a4, n4, FOCAL, 0.01)
    for (y, x) in ((5, 6), (5, 5), (4, 6)):
        assert lmm[y, x] == 0.0 and lm[y, x] != 0.0
    diff = lmm != lm
    diff[5, 6] = diff[5, 5] = diff[4, 6] = False
    assert not diff.any()
    z = oracle.normal_loss(c4, a4, n4, FOCAL, 0.0)
    assert all(np.all(v == 0.0) for v in z)
