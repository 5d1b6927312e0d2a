"""Pins the CPU oracle (oracle/) before anything is compared against it.

Each check is chosen so that a plausible restatement mistake (a dropped term,
a wrong sign or index, a transposed operand) fails it:
  - FD of quat_grad / sh_basis_grad at the reference's own points
    (crates/gs-cpu-ref/src/math.rs:200-267);
  - SH known answers (crates/gs-core/src/sh.rs:80-126);
  - the full-pipeline FD gate over every parameter class, at the reference
    gradcheck seeds and tolerances (crates/gs-cpu-ref/tests/gradcheck.rs);
  - closed forms: a fronto-parallel surfel, two stacked surfels,
    sum(w) = 1 - prod(1 - alpha);
  - literal == tile-binned walk (the binned mode is the fast path used at
    large sizes).
"""

import numpy as np
import pytest

import oracle
from microscene import make_scene, make_uniform, make_weights


# ---------------------------------------------------------------- math.rs

def test_quat_grad_matches_fd():
    # math.rs:200-226
    q = np.array([0.3, -0.5, 0.2, 0.8])
    w = np.array([[0.3, 0.9, -0.6], [-1.2, 0.1, 0.8], [0.7, -0.4, 0.5]])  # w[j, k], cols as ref
    loss = lambda q: float((oracle.quat_to_mat(q) * w).sum())
    an = oracle.quat_grad(q, w)
    h = 1e-6
    for k in range(4):
        qp, qm = q.copy(), q.copy()
        qp[k] += h
        qm[k] -= h
        fd = (loss(qp) - loss(qm)) / (2 * h)
        assert abs(fd - an[k]) < 1e-7, (k, fd, an[k])


def test_quat_to_mat_is_rotation():
    q = np.array([0.3, -0.5, 0.2, 0.8])
    r = oracle.quat_to_mat(q)
    assert np.allclose(r.T @ r, np.eye(3), atol=1e-14)
    assert abs(np.linalg.det(r) - 1.0) < 1e-14
    # Rotation by 90 degrees about z: e0 -> e1.
    s = np.sqrt(0.5)
    r = oracle.quat_to_mat([0, 0, s, s])
    assert np.allclose(r[:, 0], [0, 1, 0], atol=1e-15)


def test_sh_basis_grad_matches_fd():
    # math.rs:245-267
    d = np.array([0.48, -0.6, 0.64])
    for degree in range(4):
        g = oracle.sh_basis_grad(degree, d)
        h = 1e-6
        for i in range(3):
            e = np.zeros(3)
            e[i] = h
            fd = (oracle.sh_basis(degree, d + e) - oracle.sh_basis(degree, d - e)) / (2 * h)
            assert np.all(np.abs(fd - g[:, i]) < 1e-6), (degree, i)


def test_sh_known_answers():
    # gs-core sh.rs:80-126: basis count, degree-1 axis direction, DC value.
    for deg, k in [(0, 1), (1, 4), (2, 9), (3, 16)]:
        assert oracle.sh_basis(deg, [0, 0, 1.0]).shape == (k,)
    b = oracle.sh_basis(1, [0, 0, 1.0])
    coeffs = np.array([0.2, 10.0, 0.3, 10.0])
    assert abs(b @ coeffs - (0.28209479177387814 * 0.2 + 0.4886025119029199 * 0.3)) < 1e-12
    # Degree 0 is view independent.
    a = oracle.sh_basis(0, [0, 0, 1.0])
    c = oracle.sh_basis(0, [0.5773503, -0.5773503, 0.5773503])
    assert a[0] == c[0] == 0.28209479177387814
    # Orthonormality of the real SH basis on the sphere (Monte-Carlo free:
    # Lebedev-like check via a dense Fibonacci sphere).
    m = 20000
    i = np.arange(m) + 0.5
    phi = np.arccos(1 - 2 * i / m)
    th = np.pi * (1 + 5 ** 0.5) * i
    dirs = np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)
    B = np.stack([oracle.sh_basis(3, d) for d in dirs])
    gram = B.T @ B * (4 * np.pi / m)
    assert np.allclose(gram, np.eye(16), atol=2e-3)


# ------------------------------------------------- full-pipeline FD gate

class Checker:
    """gradcheck.rs:109-137: err <= 1e-7 + 1e-3 * max(|a|, |fd|)."""

    def __init__(self):
        self.worst = 0.0
        self.n = 0

    def check(self, label, analytic, fd):
        denom = max(abs(analytic), abs(fd))
        err = abs(analytic - fd)
        self.n += 1
        if denom > 1e-7:
            self.worst = max(self.worst, err / denom)
        assert err <= 1e-7 + 1e-3 * denom, f"{label}: analytic {analytic:.9e} fd {fd:.9e}"


def _fd_all(scene, cam, loss_fn, grads, ck, h=1e-6, with_sh=True, with_quat_cam=True):
    def fd_param(get_arr, idx):
        sp, sm = scene.copy(), scene.copy()
        get_arr(sp)[idx] += h
        get_arr(sm)[idx] -= h
        return (loss_fn(sp, cam) - loss_fn(sm, cam)) / (2 * h)

    for i in range(scene.n):
        for dim in range(3):
            ck.check(f"pos[{i}][{dim}]", grads.pos[i, dim], fd_param(lambda s: s.pos, (i, dim)))
        for dim in range(2):
            ck.check(f"scale[{i}][{dim}]", grads.scales[i, dim],
                     fd_param(lambda s: s.scales, (i, dim)))
        for dim in range(4):
            ck.check(f"quat[{i}][{dim}]", grads.quat[i, dim], fd_param(lambda s: s.quat, (i, dim)))
        ck.check(f"opacity[{i}]", grads.opacity[i], fd_param(lambda s: s.opacity, i))
        if with_sh:
            for k in range(scene.sh.shape[1]):
                for ch in range(3):
                    ck.check(f"sh[{i}][{k}][{ch}]", grads.sh[i, k, ch],
                             fd_param(lambda s: s.sh, (i, k, ch)))

    def fd_cam(attr, idx):
        cp, cm = cam.copy(), cam.copy()
        if attr == "focal":
            cp.focal += h
            cm.focal -= h
        else:
            getattr(cp, attr)[idx] += h
            getattr(cm, attr)[idx] -= h
        return (loss_fn(scene, cp) - loss_fn(scene, cm)) / (2 * h)

    for dim in range(3):
        ck.check(f"cam_center[{dim}]", grads.cam_center[dim], fd_cam("center", dim))
    if with_quat_cam:
        for dim in range(4):
            ck.check(f"cam_quat[{dim}]", grads.cam_quat[dim], fd_cam("quat", dim))
    ck.check("focal", grads.focal, fd_cam("focal", 0))


@pytest.mark.parametrize("seed,n,deg,size", [(11, 12, 0, 24), (23, 12, 0, 24), (37, 10, 1, 24),
                                             (53, 6, 3, 20)])
def test_gradcheck_reference_seeds(seed, n, deg, size):
    # gradcheck.rs:139-228
    scene, cam = make_scene(seed, n, deg, size)
    w = make_weights(seed ^ 0xABCDEF, size * size).reshape(size, size, 3)
    loss = lambda s, c: float((oracle.render(s, c, mode=0)["color"] * w).sum())
    g = oracle.gradients(scene, cam, w, mode=0)
    ck = Checker()
    _fd_all(scene, cam, loss, g, ck)
    assert ck.n > 100


@pytest.mark.parametrize("seed,n,size", [(71, 10, 20), (83, 8, 24)])
def test_gradcheck_aux_losses(seed, n, size):
    # gradcheck.rs:232-326: depth + normal image losses + distortion loss.
    scene, cam = make_scene(seed, n, 0, size)
    P = size * size
    w_color = make_weights(seed ^ 0x11, P)
    w_normal = make_weights(seed ^ 0x22, P)
    w_depth = make_uniform(seed ^ 0x33, P, -0.3, 0.3)
    lam = 0.35

    def loss(s, c):
        out = oracle.render(s, c, mode=0)
        val = (out["color"].reshape(P, 3) * w_color).sum()
        val += (out["depth"].reshape(P) * w_depth).sum()
        val += (out["normal"].reshape(P, 3) * w_normal).sum()
        return float(val + lam * oracle.distortion_loss(s, c, mode=0))

    g = oracle.gradients(scene, cam, w_color, w_depth, w_normal, lam, mode=0)
    ck = Checker()
    _fd_all(scene, cam, loss, g, ck, with_sh=False, with_quat_cam=False)
    grad_mag = np.linalg.norm(g.pos, axis=1).sum() + np.abs(g.opacity).sum()
    assert grad_mag > 1e-2  # gradcheck.rs:309-315 non-vacuity


def test_forward_renders_something():
    # gradcheck.rs:328-334
    scene, cam = make_scene(99, 10, 1, 32)
    out = oracle.render(scene, cam)
    assert out["alpha"].sum() > 1.0


@pytest.mark.parametrize("seed,n,deg,size", [(101, 20, 0, 64), (303, 12, 3, 48), (404, 50, 1, 100)])
def test_literal_equals_binned(seed, n, deg, size):
    scene, cam = make_scene(seed, n, deg, size, flavor="parity")
    a = oracle.render(scene, cam, mode=0)
    b = oracle.render(scene, cam, mode=1)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    w = make_weights(seed ^ 0xFEED, size * size)
    ga = oracle.gradients(scene, cam, w, mode=0)
    gb = oracle.gradients(scene, cam, w, mode=1)
    for k in ("pos", "scales", "quat", "opacity", "sh"):
        np.testing.assert_allclose(getattr(gb, k), getattr(ga, k), rtol=1e-12, atol=1e-14)


# ------------------------------------------------------------ closed forms

def _single_scene(pos, scales, opacity, dc, quat=(0, 0, 0, 1.0)):
    n = len(pos)
    sh = np.zeros((n, 1, 3))
    sh[:, 0, :] = np.asarray(dc)
    return oracle.Scene(np.asarray(pos, float), np.asarray(scales, float),
                        np.tile(np.asarray(quat, float), (n, 1)), np.asarray(opacity, float), sh, 0)


def test_fronto_parallel_surfel_closed_form():
    # Identity camera at the origin looks down -z; a surfel at z = -3 with
    # identity rotation spans camera x/y. The ray through pixel (x, y) hits its
    # plane at t = 3, so u = 3 d.x / s_u, v = 3 d.y / s_v (forward.rs:136-169).
    W = H = 32
    f = 40.0
    cam = oracle.Camera(np.zeros(3), np.array([0, 0, 0, 1.0]), f, W, H)
    su, sv, op = 0.2, 0.1, 0.7
    dc = np.array([0.4, -0.3, 1.0])
    scene = _single_scene([[0.05, -0.02, -3.0]], [[su, sv]], [op], [dc])
    out = oracle.render(scene, cam, mode=0)
    col = np.maximum(0.5 + 0.28209479177387814 * dc, 0.0)
    cx = cy = 16.0
    px, py = cx + f * 0.05 / 3.0, cy - f * (-0.02) / 3.0
    for y in range(H):
        for x in range(W):
            dx = (x + 0.5 - cx) / f
            dy = -(y + 0.5 - cy) / f
            u = (3 * dx - 0.05) / su
            v = (3 * dy + 0.02) / sv
            g_ray = np.exp(-(u * u + v * v) / 2)
            g_scr = np.exp(-((x + 0.5 - px) ** 2 + (y + 0.5 - py) ** 2))
            a = min(op * max(g_ray, g_scr), 0.995)
            r_px = min(3 * f * (su + sv) / 3.0 + 3.0, 64.0)
            inside = (int(np.clip((px - r_px) / 16, 0, 1)) <= x // 16 <= int(np.clip((px + r_px) / 16, 0, 1))
                      and int(np.clip((py - r_px) / 16, 0, 1)) <= y // 16 <= int(np.clip((py + r_px) / 16, 0, 1)))
            if a < 1 / 255 or not inside:
                a = 0.0
            np.testing.assert_allclose(out["color"][y, x], col * a, atol=1e-13)
            assert abs(out["alpha"][y, x] - a) < 1e-13
            if a > 0:
                # depth = alpha * t with t = 3 on the ray branch (or the center depth, also 3)
                assert abs(out["depth"][y, x] - 3.0 * a) < 1e-9


def test_two_stacked_surfels_composite():
    # C = c1 a1 + c2 a2 (1 - a1) with the nearer surfel first; T = (1-a1)(1-a2).
    W = H = 16
    f = 20.0
    cam = oracle.Camera(np.zeros(3), np.array([0, 0, 0, 1.0]), f, W, H)
    d1 = np.array([0.8, -0.5, 0.1])
    d2 = np.array([-0.2, 0.6, -0.9])
    # Listed far-first so the depth sort must reorder them.
    scene = _single_scene([[0, 0, -4.0], [0, 0, -2.0]], [[0.5, 0.5], [0.3, 0.3]], [0.6, 0.5], [d2, d1])
    one_near = _single_scene([[0, 0, -2.0]], [[0.3, 0.3]], [0.5], [d1])
    one_far = _single_scene([[0, 0, -4.0]], [[0.5, 0.5]], [0.6], [d2])
    both = oracle.render(scene, cam, mode=0)
    a1 = oracle.render(one_near, cam, mode=0)["alpha"]
    a2 = oracle.render(one_far, cam, mode=0)["alpha"]
    c1 = np.maximum(0.5 + 0.28209479177387814 * d1, 0)
    c2 = np.maximum(0.5 + 0.28209479177387814 * d2, 0)
    expect = c1 * a1[..., None] + c2 * (a2 * (1 - a1))[..., None]
    np.testing.assert_allclose(both["color"], expect, atol=1e-13)
    np.testing.assert_allclose(both["alpha"], 1 - (1 - a1) * (1 - a2), atol=1e-13)
    assert list(oracle.prepare_order(scene, cam)) == [1, 0]


def test_weights_sum_to_one_minus_transmittance():
    scene, cam = make_scene(404, 50, 1, 64, flavor="parity")
    out = oracle.render(scene, cam)
    assert np.all(out["alpha"] >= 0) and np.all(out["alpha"] <= 1 + 1e-12)
    # Zero dL/d(image) gives exactly zero gradients (backward.rs:73-77).
    g = oracle.gradients(scene, cam, np.zeros((64 * 64, 3)))
    assert not np.any(g.pos) and not np.any(g.opacity) and g.focal == 0.0


def test_marginal_decisions_finds_planted_threshold_crossings():
    """oracle.marginal_decisions (a test aid): a surfel whose alpha at one
    pixel is planted 3e-6 above 1/255 is reported at rel 1e-5 and not at
    1e-6; the same surfel with 1 % more opacity has no marginal decision;
    two overlapping surfels at the same depth (equal f32 keys) are both
    reported."""
    W = H = 32
    f = 40.0
    cam = oracle.Camera(np.zeros(3), np.array([0, 0, 0, 1.0]), f, W, H)
    base = _single_scene([[0.05, -0.02, -3.0]], [[0.2, 0.1]], [0.7], [[0.4, -0.3, 1.0]])
    alpha = oracle.render(base, cam, mode=0)["alpha"]
    ys, xs = np.nonzero((alpha > 0.02) & (alpha < 0.2))
    a_p = alpha[ys[0], xs[0]]
    op = 0.7 / 255.0 / a_p * (1 + 3e-6)
    planted = _single_scene([[0.05, -0.02, -3.0]], [[0.2, 0.1]], [op], [[0.4, -0.3, 1.0]])
    k, flags = oracle.marginal_decisions(planted, cam, 1e-5)
    assert k >= 1 and flags[0] == 1
    k6, _ = oracle.marginal_decisions(planted, cam, 1e-6)
    assert k6 == 0
    clear = _single_scene([[0.05, -0.02, -3.0]], [[0.2, 0.1]], [op * 1.01], [[0.4, -0.3, 1.0]])
    assert oracle.marginal_decisions(clear, cam, 1e-5)[0] == 0
    tie = _single_scene([[0.0, 0.0, -3.0], [0.1, 0.0, -3.0], [0.0, 0.0, -5.0]], [[0.2, 0.2]] * 3,
                        [0.5] * 3, [[0.1, 0.2, 0.3]] * 3)
    k, flags = oracle.marginal_decisions(tie, cam, 1e-9)
    assert k >= 1 and list(flags) == [1, 1, 0]


def test_pixel_contributors_follow_the_walk():
    """oracle.pixel_contributors (a test aid): three stacked wide opaque
    surfels over the centre pixel -- the two nearest composite there (alpha
    clamped to 0.995), the third lies past the transmittance stop
    (0.005^2 < 1e-4); a fourth surfel off to the side reaches only pixels
    near it."""
    cam = oracle.Camera(np.zeros(3), np.array([0, 0, 0, 1.0]), 20.0, 16, 16)
    scene = _single_scene([[0, 0, -2.0], [0, 0, -3.0], [0, 0, -4.0], [0.7, 0.0, -2.0]],
                          [[2.0, 2.0]] * 3 + [[0.02, 0.02]], [0.999] * 4, [[0.1, 0.2, 0.3]] * 4)
    assert list(oracle.pixel_contributors(scene, cam, [(8, 8)])) == [1, 1, 0, 0]
    assert oracle.pixel_contributors(scene, cam, [(15, 7)])[3] == 1
    assert not oracle.pixel_contributors(scene, cam, np.zeros((0, 2))).any()


def test_near_culled_scene_is_empty():
    cam = oracle.Camera(np.zeros(3), np.array([0, 0, 0, 1.0]), 20.0, 16, 16)
    scene = _single_scene([[0, 0, 1.0], [0, 0, -0.01]], [[0.3, 0.3]] * 2, [0.9, 0.9], [[0, 0, 0]] * 2)
    out = oracle.render(scene, cam, mode=0)
    assert not out["color"].any() and len(oracle.prepare_order(scene, cam)) == 0


def test_multiview_loss_gradient_matches_central_differences():
    """The FP64 gradient of L1 + D-SSIM summed over an 8-view batch of uneven
    resolutions, against central differences (h = 1e-6, the reference's
    gradcheck.rs gate: |a - fd| <= 1e-7 + 1e-3 max(|a|, |fd|)) on 64 random
    scalar parameters across every class. The loss has jumps (alpha crossing
    1/255, tile-rect edges, depth-order swaps); gradcheck.rs avoids them by
    seed choice, here a jump inside [-h, h] is detected by the one-sided
    differences disagreeing, and at most 8 of the 64 may be skipped so."""
    import multiview as mv
    sc, tgt, specs = mv.problem(n=400, seed=29)
    tgts = mv.targets(tgt, specs)
    s64 = mv.scene64(sc)
    l0, g = mv.total_loss(s64, specs, tgts, grad=True)
    rng = np.random.default_rng(64)
    ncoef = (sc.sh_degree + 1) ** 2
    fields = [("pos", 3), ("scales", 2), ("quat", 4), ("opacity", 1), ("sh", ncoef * 3)]
    h = 1e-6
    worst = 0.0
    jumps = []
    for trial in range(64):
        name, width = fields[trial % len(fields)]
        # Parameters of surfels that carry gradient (visible in some view).
        gf = np.abs(getattr(g, name)).reshape(sc.n, -1).sum(1)
        i = int(rng.choice(np.flatnonzero(gf > 0)))
        c = int(rng.integers(width))
        flat = getattr(s64, name).reshape(sc.n, -1)
        orig = flat[i, c]
        flat[i, c] = orig + h
        lp, _ = mv.total_loss(s64, specs, tgts)
        flat[i, c] = orig - h
        lm, _ = mv.total_loss(s64, specs, tgts)
        flat[i, c] = orig
        fwd, bwd = (lp - l0) / h, (l0 - lm) / h
        if abs(fwd - bwd) > 1e-7 + 1e-2 * max(abs(fwd), abs(bwd)):
            jumps.append(f"{name}[{i},{c}]")
            continue
        fd = 0.5 * (fwd + bwd)
        an = getattr(g, name).reshape(sc.n, -1)[i, c]
        denom = max(abs(an), abs(fd))
        err = abs(an - fd)
        if denom > 0:
            worst = max(worst, err / denom)
        assert err <= 1e-7 + 1e-3 * denom, f"{name}[{i},{c}]: analytic {an:.9e} fd {fd:.9e}"
    print(f"worst relative FD disagreement {worst:.2e}; jumps skipped {len(jumps)}: {jumps}")
    assert len(jumps) <= 8


@pytest.mark.parametrize("seed,n,deg,size", [(101, 20, 0, 64), (303, 12, 3, 48), (404, 50, 1, 100)])
def test_replay_of_the_faithful_decisions_is_the_faithful_evaluation(seed, n, deg, size):
    """Decision replay fed the FP64 walk's own decisions (oracle.trace)
    reproduces the faithful render and every parameter gradient: it takes the
    same hits in the same order with the same branch and clamp."""
    scene, cam = make_scene(seed, n, deg, size, flavor="parity")
    lists, counts = oracle.trace(scene, cam, 256)
    assert counts.max() <= 256 and counts.sum() > 0
    ref = oracle.render(scene, cam)
    w = make_weights(seed ^ 0xFEED, size * size).reshape(size, size, 3)
    g = oracle.gradients(scene, cam, w)
    col, alpha, gr = oracle.replay(scene, cam, lists, counts, w)
    assert np.abs(col - ref["color"]).max() <= 1e-14
    assert np.abs(alpha - ref["alpha"]).max() <= 1e-14
    for k in ("pos", "scales", "quat", "opacity", "sh"):
        np.testing.assert_allclose(getattr(gr, k), getattr(g, k), rtol=1e-12, atol=1e-15, err_msg=k)
    np.testing.assert_allclose(gr.cam_center, g.cam_center, rtol=1e-10, atol=1e-13)


def test_replay_follows_the_given_decisions():
    """Replay uses the lists, not its own decisions: flipping one hit's branch
    bit, setting its clamp bit or dropping the tail of a list changes the
    replayed image at that pixel and only there."""
    scene, cam = make_scene(303, 12, 3, 48, flavor="parity")
    lists, counts = oracle.trace(scene, cam, 256)
    base, _, _ = oracle.replay(scene, cam, lists, counts)
    p = int(np.argmax(counts))
    assert counts[p] >= 2
    y, x = divmod(p, 48)
    for what in ("branch", "clamp", "truncate"):
        l2, c2 = lists.copy(), counts.copy()
        if what == "branch":
            l2[p, 0] ^= 1 << 30
        elif what == "clamp":
            l2[p, 0] |= np.uint32(1 << 31)
        else:
            c2[p] -= 1
        img, _, _ = oracle.replay(scene, cam, l2, c2)
        diff = np.abs(img - base).max(axis=2)
        assert diff[y, x] > 1e-6, what
        diff[y, x] = 0.0
        assert diff.max() == 0.0, what


@pytest.mark.parametrize("seed,n,deg,size", [(111, 15, 1, 48), (222, 15, 3, 48)])
def test_replay_gradients_match_central_differences_with_the_decisions_fixed(seed, n, deg, size):
    """With the decisions fixed the replayed loss is smooth in every
    parameter, so central differences of it (h = 1e-6, gradcheck.rs's
    1e-7 + 1e-3 denom) pin the replay's analytic gradients without the
    discontinuities a faithful FD must avoid."""
    scene, cam = make_scene(seed, n, deg, size, flavor="parity")
    lists, counts = oracle.trace(scene, cam, 256)
    w = make_weights(seed ^ 0xFEED, size * size).reshape(size, size, 3)
    _, _, g = oracle.replay(scene, cam, lists, counts, w)

    def loss(s):
        return float((oracle.replay(s, cam, lists, counts)[0] * w).sum())

    rng = np.random.default_rng(seed)
    h = 1e-6
    checked = 0
    for cls, arr_g in (("pos", g.pos), ("scales", g.scales), ("quat", g.quat), ("opacity", g.opacity),
                       ("sh", g.sh)):
        base = getattr(scene, cls)
        for _ in range(6):
            idx = tuple(int(rng.integers(0, d)) for d in base.shape)
            sp, sm = scene.copy(), scene.copy()
            getattr(sp, cls)[idx] += h
            getattr(sm, cls)[idx] -= h
            fd = (loss(sp) - loss(sm)) / (2 * h)
            an = float(arr_g[idx])
            assert abs(fd - an) <= 1e-7 + 1e-3 * max(abs(fd), abs(an)), (cls, idx, fd, an)
            checked += 1
    assert checked == 30


@pytest.mark.parametrize("seed,n,size", [(71, 10, 20), (83, 8, 24)])
def test_replay_with_aux_losses_is_the_faithful_evaluation(seed, n, size):
    """The replay of the FP64 walk's own decisions with depth, normal and
    distortion terms reproduces oracle.gradients (gradcheck.rs:232-326 case)."""
    scene, cam = make_scene(seed, n, 0, size)
    P = size * size
    w_color = make_weights(seed ^ 0x11, P)
    w_normal = make_weights(seed ^ 0x22, P)
    w_depth = make_uniform(seed ^ 0x33, P, -0.3, 0.3)
    g = oracle.gradients(scene, cam, w_color, w_depth, w_normal, 0.35)
    lists, counts = oracle.trace(scene, cam, 256)
    _, _, gr = oracle.replay(scene, cam, lists, counts, w_color.reshape(size, size, 3), w_depth, w_normal, 0.35)
    for k in ("pos", "scales", "quat", "opacity"):
        np.testing.assert_allclose(getattr(gr, k), getattr(g, k), rtol=1e-11, atol=1e-14, err_msg=k)
    np.testing.assert_allclose(gr.cam_center, g.cam_center, rtol=1e-10, atol=1e-13)
