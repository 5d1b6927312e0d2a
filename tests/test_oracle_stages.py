"""Pins the oracle's integer mirror, L1 + D-SSIM and Adam.

- scan / emit / stable sort / ranges against independent numpy
  restatements, on the reference's property-test case lists
  (crates/gs-wgpu/tests/prefix_sum.rs:57-73, sort.rs:27-113,
  crates/gs-kernels/tests/binning.rs:118-158);
- the fp32 prep mirror against an FP64 evaluation (rects/keys may only
  differ where f32 rounding straddles a boundary) and the oracle's own
  FP64 prepare order;
- SSIM: FD in FP64, SSIM(x, x) = 1, blur self-adjointness, the window
  constants against the closed-form Gaussian;
- Adam: the reference's own CPU loop (gs-train/tests/adam.rs:37-98).
"""

import numpy as np
import pytest

import oracle
from microscene import make_scene, xorshift32_stream


# -------------------------------------------------------------- scan / sort

@pytest.mark.parametrize("n", [1, 2, 255, 1024, 1025, 4096, 100_003])
def test_exclusive_scan(n):
    vals = xorshift32_stream(0x5EED0001 ^ n, n) % 16
    got, total = oracle.exclusive_scan(vals)
    ref = np.concatenate([[0], np.cumsum(vals.astype(np.uint64))[:-1]]).astype(np.uint32)
    assert np.array_equal(got, ref)
    assert total == int(vals.astype(np.uint64).sum() % (1 << 32))


def test_exclusive_scan_wraps_mod_2_32():
    vals = np.full(5, 0xFFFFFFFF, dtype=np.uint32)
    got, total = oracle.exclusive_scan(vals)
    assert list(got) == [0, 0xFFFFFFFF, 0xFFFFFFFE, 0xFFFFFFFD, 0xFFFFFFFC]
    assert total == 0xFFFFFFFB


@pytest.mark.parametrize("n", [0, 1, 255, 256, 1024, 4097, 100_000])
def test_stable_sort_matches_numpy(n):
    # sort.rs:33-40: heavy duplicates, payload = original index.
    seed = (0x12345678 ^ ((n * 2654435761) & 0xFFFFFFFF)) & 0xFFFFFFFF
    keys32 = xorshift32_stream(seed, n) % 8192
    keys = (keys32.astype(np.uint64) << 32) | (keys32.astype(np.uint64) * 7 % 5)
    pay = np.arange(n, dtype=np.uint32)
    k, p = oracle.stable_sort(keys, pay)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order]) and np.array_equal(p, pay[order])


def _items(num, seed, tiles=64):
    # binning.rs:123-141
    r = xorshift32_stream(seed, num * 5).reshape(num, 5)
    x0 = r[:, 0] % tiles
    y0 = r[:, 1] % tiles
    w = r[:, 2] % 4
    h = r[:, 3] % 4
    rects = np.stack([x0, y0, np.minimum(x0 + w, tiles - 1), np.minimum(y0 + h, tiles - 1)], 1)
    return rects.astype(np.uint32), (r[:, 4] % 32).astype(np.uint32)


def _brute_binning(rects, depths, tiles_x, num_tiles):
    ent = []
    for i, (x0, y0, x1, y1) in enumerate(rects):
        if x1 < x0 or y1 < y0:
            continue
        for ty in range(y0, y1 + 1):
            for tx in range(x0, x1 + 1):
                ent.append((ty * tiles_x + tx, int(depths[i]), len(ent), i))
    ent.sort(key=lambda t: (t[0], t[1], t[2]))
    ranges = np.zeros((num_tiles, 2), dtype=np.uint32)
    for j, (t, _, _, _) in enumerate(ent):
        if j == 0 or ent[j - 1][0] != t:
            ranges[t, 0] = j
        ranges[t, 1] = j + 1
    return [e[3] for e in ent], ranges


@pytest.mark.parametrize("num,seed", [(1, 7), (100, 11), (5000, 13)])
def test_binning_pipeline_matches_brute_force(num, seed):
    rects, depths = _items(num, seed)
    keys, pay, e = oracle.emit_keys(rects, depths, 64)
    assert e == len(keys)
    k, p = oracle.stable_sort(keys, pay)
    ranges = oracle.tile_ranges(k, 64 * 64)
    want_p, want_r = _brute_binning(rects, depths, 64, 64 * 64)
    assert list(p) == want_p
    assert np.array_equal(ranges, want_r)


def test_binning_degenerate_cases():
    # binning.rs:145-157: a culled item; 500 items stacked in one tile.
    keys, pay, e = oracle.emit_keys(np.array([[5, 5, 4, 4]], np.uint32), np.array([1], np.uint32), 64)
    assert e == 0
    rects = np.tile(np.array([[10, 10, 10, 10]], np.uint32), (500, 1))
    depths = (np.arange(500) % 3).astype(np.uint32)
    keys, pay, e = oracle.emit_keys(rects, depths, 64)
    k, p = oracle.stable_sort(keys, pay)
    r = oracle.tile_ranges(k, 64 * 64)
    assert tuple(r[10 * 64 + 10]) == (0, 500) and r.sum() == 500
    want = np.argsort(depths, kind="stable")
    assert np.array_equal(p, want)


def test_emit_keys_respects_max_entries():
    rects, depths = _items(100, 11)
    full_k, full_p, e = oracle.emit_keys(rects, depths, 64)
    k, p, e2 = oracle.emit_keys(rects, depths, 64, max_entries=50)
    assert e2 == e and len(k) == 50 and np.array_equal(k, full_k[:50])


# ---------------------------------------------------------- fp32 prep mirror

def test_prep_f32_agrees_with_fp64_prepare():
    scene, cam = make_scene(404, 50, 1, 100, flavor="parity")
    rot = oracle.quat_to_mat(cam.quat)
    c32 = oracle.cam32(rot, cam.center, cam.focal, cam.width, cam.height)
    pos4 = np.concatenate([scene.pos, np.ones((scene.n, 1))], 1)
    rects, keys, counts = oracle.prep_f32(pos4, scene.scales, scene.quat, c32)
    # Depth keys = f32 bits of -c.z, ordered like the FP64 prepare() order.
    order = oracle.prepare_order(scene, cam)
    live = keys != 0
    assert live.sum() == len(order)
    lk = keys[order]
    assert np.all(lk[:-1] <= lk[1:])
    # Depth from the key vs FP64 depth.
    c = (scene.pos - cam.center) @ rot
    depth64 = -c[:, 2]
    depth32 = keys.view(np.float32)
    np.testing.assert_allclose(depth32[live], depth64[live], rtol=1e-6)
    # Counts are the rect areas.
    r = rects.astype(np.int64)
    area = np.where((r[:, 2] >= r[:, 0]) & (r[:, 3] >= r[:, 1]),
                    (r[:, 2] - r[:, 0] + 1) * (r[:, 3] - r[:, 1] + 1), 0)
    assert np.array_equal(area, counts)
    assert counts.sum() > 0


@pytest.mark.parametrize("seed,n,size", [(404, 50, 100), (101, 400, 256), (7, 3000, 512)])
def test_prep_f32_rects_agree_with_fp64_rects(seed, n, size):
    """The fp32 mirror's tile rects (the HIP projection's, bit-exact) against
    FP64 forward.rs:104-123 rects: equal except where the FP64 bound
    (p -+ r) / 16 lies within f32 rounding (1e-5 relative) of an integer tile
    edge, so truncation may go either way; those are counted and must be
    rare. Off-screen and near-culled surfels agree exactly (an empty rect in
    the mirror)."""
    scene, cam = make_scene(seed, n, 1, size, flavor="parity")
    rot = oracle.quat_to_mat(cam.quat)
    c32 = oracle.cam32(rot, cam.center, cam.focal, cam.width, cam.height)
    pos4 = np.concatenate([scene.pos, np.ones((scene.n, 1))], 1)
    rects, keys, counts = oracle.prep_f32(pos4, scene.scales, scene.quat, c32)
    r64, on64, cul64 = oracle.prepare_rects(scene, cam)
    live = (rects[:, 0] <= rects[:, 2]) & (rects[:, 1] <= rects[:, 3]) & (keys != 0)
    # FP64 bounds before truncation, for the straddle check.
    c = (scene.pos - cam.center) @ rot
    w = -c[:, 2]
    R = [oracle.quat_to_mat(q) for q in scene.quat]
    tu = np.array([rot.T @ (Ri[:, 0] * s[0]) for Ri, s in zip(R, scene.scales)])
    tv = np.array([rot.T @ (Ri[:, 1] * s[1]) for Ri, s in zip(R, scene.scales)])
    r_px = np.minimum(3.0 * cam.focal * (np.linalg.norm(tu, axis=1) + np.linalg.norm(tv, axis=1)) / w + 3.0, 64.0)
    px = cam.width * 0.5 + cam.focal * c[:, 0] / w
    py = cam.height * 0.5 - cam.focal * c[:, 1] / w
    bounds = np.stack([(px - r_px) / 16, (py - r_px) / 16, (px + r_px) / 16, (py + r_px) / 16], 1)
    near_edge = np.abs(bounds - np.round(bounds)) <= 1e-5 * np.maximum(np.abs(bounds), 1.0)
    both = live & on64 & ~cul64
    diff = np.any(rects.astype(np.int64) != r64, axis=1) & both
    straddle = np.any(near_edge, axis=1)
    print(f"  {int(both.sum())} on-screen surfels; rects differing {int(diff.sum())}, all at a tile edge")
    assert not np.any(diff & ~straddle)
    assert diff.sum() <= max(2, both.sum() // 1000)
    # On-screen / culled status agrees (an empty mirror rect = not on screen or culled).
    assert np.array_equal(live, on64 & ~cul64)


# ------------------------------------------------------------------- SSIM

def test_gauss_window_is_normalized_gaussian():
    k = np.arange(-5, 6)
    g = np.exp(-k * k / (2 * 1.5 ** 2))
    g /= g.sum()
    img = np.zeros((11, 11))
    img[5, 5] = 1.0
    b = oracle.blur(img)
    # ssim.wgsl:32-34 holds rounded taps (they sum to 0.99998634, and tap 3
    # is 1.9e-4 low): the oracle keeps the reference's literals, so this pins
    # them against the closed form to 2e-4 relative (a swapped digit fails).
    np.testing.assert_allclose(b, np.outer(g, g), rtol=4e-4, atol=1e-12)
    assert abs(b.sum() - 0.99998634 ** 2) < 1e-7


def test_blur_is_self_adjoint():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((19, 23))
    y = rng.standard_normal((19, 23))
    assert abs((oracle.blur(x) * y).sum() - (x * oracle.blur(y)).sum()) < 1e-12


def test_ssim_of_identical_images_is_one():
    rng = np.random.default_rng(1)
    x = rng.random((20, 17, 3))
    loss, dl, l1, ss = oracle.l1_ssim(x, x, lam=0.2, dssim_scale=0.5)
    assert abs(ss - 1.0) < 1e-12 and l1 == 0.0 and abs(loss) < 1e-12
    # At x == y, sign(0) = 0 and dSSIM/dx = 0: the gradient vanishes.
    assert np.abs(dl).max() < 1e-12


@pytest.mark.parametrize("scale", [0.5, 1.0])
def test_ssim_gradient_matches_fd(scale):
    rng = np.random.default_rng(2)
    H, W = 14, 16
    x = rng.random((H, W, 3))
    y = np.clip(x + rng.normal(0, 0.2, x.shape), 0, 1)
    lam = 0.2
    loss, dl, _, _ = oracle.l1_ssim(x, y, lam, scale)
    h = 1e-6
    idx = rng.choice(H * W * 3, 60, replace=False)
    for k in idx:
        i, j, c = np.unravel_index(k, x.shape)
        if abs(x[i, j, c] - y[i, j, c]) < 1e-4:
            continue  # |x-y| kink
        xp, xm = x.copy(), x.copy()
        xp[i, j, c] += h
        xm[i, j, c] -= h
        fd = (oracle.l1_ssim(xp, y, lam, scale, grad=False)[0] -
              oracle.l1_ssim(xm, y, lam, scale, grad=False)[0]) / (2 * h)
        assert abs(fd - dl[i, j, c]) <= 1e-9 + 1e-5 * abs(fd), (k, fd, dl[i, j, c])


def test_loss_decomposition():
    rng = np.random.default_rng(3)
    x = rng.random((12, 13, 3))
    y = rng.random((12, 13, 3))
    for lam, sc in [(0.2, 0.5), (0.2, 1.0), (0.0, 1.0)]:
        loss, _, l1, ss = oracle.l1_ssim(x, y, lam, sc, grad=False)
        assert abs(l1 - np.abs(x - y).mean()) < 1e-14
        assert abs(loss - ((1 - lam) * l1 + lam * sc * (1 - ss))) < 1e-14


# ------------------------------------------------------------------- Adam

def _ref_xorshift(state):
    # gs-train/tests/adam.rs:17-24: f32 in [-1, 1).
    x = state[0]
    m = (1 << 64) - 1
    x ^= (x << 13) & m
    x ^= x >> 7
    x ^= (x << 17) & m
    state[0] = x
    return np.float32((x >> 11) / float(1 << 53) * 2.0 - 1.0)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_adam_matches_reference_cpu_loop(mode):
    # gs-train/tests/adam.rs:37-98 (N=1000, 5 steps, lr 0.01, |d| < 1e-5).
    N, STEPS, LR = 1000, 5, np.float32(0.01)
    st = [0xADA ^ mode]
    init = np.array([_ref_xorshift(st) * np.float32(0.8) for _ in range(N)], np.float32)
    grads = [np.array([_ref_xorshift(st) for _ in range(N)], np.float32) for _ in range(STEPS)]
    f = np.float32

    def dact(x):
        if mode == 1:
            return np.exp(x)
        if mode == 2:
            s = f(1) / (f(1) + np.exp(-x))
            return s * (f(1) - s)
        return np.ones_like(x)

    p = init.copy()
    m = np.zeros(N, np.float32)
    v = np.zeros(N, np.float32)
    for t, g in enumerate(grads, start=1):
        gg = g * dact(p)
        m = f(0.9) * m + f(0.1) * gg
        v = f(0.999) * v + f(0.001) * gg * gg
        mhat = m / (f(1) - f(0.9) ** t)
        vhat = v / (f(1) - f(0.999) ** t)
        p = p - LR * mhat / (np.sqrt(vhat) + f(1e-8))
    raw = init.copy()
    mo = np.zeros(N, np.float32)
    vo = np.zeros(N, np.float32)
    for t, g in enumerate(grads, start=1):
        act = oracle.adam_step_f32(raw, g, mo, vo, float(LR), mode, t)
    assert np.abs(raw - p).max() < 1e-5
    want_act = {0: raw, 1: np.exp(raw), 2: 1 / (1 + np.exp(-raw))}[mode]
    np.testing.assert_allclose(act, want_act, rtol=1e-6)


def _ssim_sum(w, h, a, b):
    img = np.full((h, w, 3), a)
    tgt = np.full((h, w, 3), b)
    _, _, _, ss = oracle.l1_ssim(img, tgt, 0.2, 1.0)
    return ss * w * h  # per-channel sum (all channels equal)


@pytest.mark.parametrize("a,b", [(0.5, 0.3), (0.9, 0.1), (0.2, 0.2), (0.05, 0.7)])
def test_ssim_constant_images_closed_form(a, b):
    """Constant images x = a, y = b: at a pixel 5 or more from every edge the
    window sums to S = (sum of the 11 taps)^2, so mu_x = a S, E[x^2] = a^2 S,
    E[xy] = a b S and SSIM = (2abS^2 + C1)(2abS(1-S) + C2) /
    (((a^2+b^2)S^2 + C1)((a^2+b^2)S(1-S) + C2)) with C1 = 1e-4, C2 = 9e-4
    (ssim.wgsl:29-34). A pixel's SSIM depends only on its distances to the
    edges, so the mixed second difference of the image sums over sizes
    (W, H), (W+1, H), (W, H+1), (W+1, H+1) is exactly one interior pixel."""
    taps = np.array([0.26601818, 0.21300237, 0.10936069, 0.03599398, 0.00759866, 0.00102838])
    s1 = taps[0] + 2 * taps[1:].sum()
    S = s1 * s1
    C1, C2 = 1e-4, 9e-4
    want = ((2 * a * b * S * S + C1) * (2 * a * b * S * (1 - S) + C2) /
            (((a * a + b * b) * S * S + C1) * ((a * a + b * b) * S * (1 - S) + C2)))
    W, H = 17, 19
    got = (_ssim_sum(W + 1, H + 1, a, b) - _ssim_sum(W + 1, H, a, b) - _ssim_sum(W, H + 1, a, b)
           + _ssim_sum(W, H, a, b))
    assert abs(got - want) < 1e-10, (got, want)
    # Swapped constants move the value well outside that bound.
    alt = ((2 * a * b * S * S + C2) * (2 * a * b * S * (1 - S) + C1) /
           (((a * a + b * b) * S * S + C2) * ((a * a + b * b) * S * (1 - S) + C1)))
    if a != b:
        assert abs(alt - want) > 1e-6
