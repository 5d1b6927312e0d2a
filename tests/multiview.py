"""Multi-view training problems and their FP64 loss and gradient: the sum
over views of L1 + D-SSIM (each view's mean over its own pixels), as the
batched trainer step sums per-view gradients (trainer.rs:1139-1332).

problem(): 8 views of uneven resolutions (the shape of the 8-view,
360p..1080p sphere-camera configuration, scaled down 10x so the FP64 oracle
evaluates it in seconds).
frames_problem(): 16 frames of a moving camera (SE(3) walk, 1 cm / 0.5 deg
per frame) over a growing scene, frame k rendering the first
base + k * 10 % base surfels (the streaming-frame configuration, 720p and
500k base surfels, scaled down 10x in size and ~800x in surfels).
"""

import numpy as np

import oracle

VIEWS = 8
LAM = 0.2
DSSIM_SCALE = 0.5


def problem(n=1200, deg=3, seed=17):
    from splattar_amd import scenes
    sc = scenes.make_scene(n, deg, -3.4, 0.4, seed)
    tgt = scenes.jittered(sc, 0.01, seed ^ 1)
    rng = np.random.default_rng(seed + 5)
    hs = rng.integers(36, 109, VIEWS)
    sizes = [(int(round(h * 16 / 9)), int(h)) for h in hs]
    specs = scenes.sphere_views(VIEWS, seed + 9, sizes)
    return sc, tgt, specs


def frames_problem(base=500, frames=16, size=(128, 72), seed=29):
    from splattar_amd import scenes
    counts = scenes.frame_surfel_counts(frames, base)
    sc = scenes.make_scene(counts[-1], 3, -3.4, 0.4, seed)
    tgt = scenes.jittered(sc, 0.01, seed ^ 1)
    specs = scenes.walk_views(frames, seed + 3, size)
    return sc, tgt, specs, counts


def subset(s64, m):
    """The first m surfels of an FP64 scene."""
    return oracle.Scene(s64.pos[:m], s64.scales[:m], s64.quat[:m], s64.opacity[:m], s64.sh[:m],
                        s64.sh_degree)


def scene64(sc):
    n = sc.n
    ncoef = (sc.sh_degree + 1) ** 2
    return oracle.Scene(sc.positions.astype(np.float64), sc.scales.astype(np.float64),
                        sc.quats.astype(np.float64), sc.opacities.astype(np.float64),
                        sc.sh.astype(np.float64).reshape(n, ncoef, 3), sc.sh_degree)


def camera64(sp):
    return oracle.Camera(np.asarray(sp.center, np.float32).astype(np.float64),
                         np.asarray(sp.quat, np.float32).astype(np.float64),
                         float(np.float32(sp.focal)), sp.width, sp.height)


def targets(tgt, specs, counts=None, offset=0.0):
    """FP64 renders of the target scene (view k: its first counts[k]
    surfels) plus `offset` per channel, rounded to f32 (what the GPU holds).

    An offset above every rendered value (2.0; colours here stay below 1.6)
    keeps sign(render - target) of the L1 term at -1 for every pixel, so
    f32 and FP64 cannot take opposite signs where the two agree to ~1e-7,
    while the target's structure (what SSIM sees) is unchanged."""
    s = scene64(tgt)
    counts = counts or [s.pos.shape[0]] * len(specs)
    return [(oracle.render(subset(s, m), camera64(sp))["color"] + offset).astype(np.float32)
            for sp, m in zip(specs, counts)]


def l1_sign_margin(s64, specs, tgts, counts=None):
    """Largest render - target over every pixel and channel (FP64): below
    zero by a margin means the L1 sign is the same for any rounding."""
    n = s64.pos.shape[0]
    counts = counts or [n] * len(specs)
    return max(float((oracle.render(subset(s64, m), camera64(sp))["color"] - t).max())
               for sp, t, m in zip(specs, tgts, counts))


def marginal_mask(s64, specs, counts=None, rel=1e-4):
    """Surfels in a forward decision within relative band rel of a threshold
    in some view (oracle.marginal_decisions), and the number of such
    decisions. 1e-4 is several times the relative error of the f32 path's
    alpha (its screen-branch exponent carries the f32 rounding of the
    projected centre)."""
    n = s64.pos.shape[0]
    counts = counts or [n] * len(specs)
    mask = np.zeros(n, bool)
    total = 0
    for sp, m in zip(specs, counts):
        k, f = oracle.marginal_decisions(subset(s64, m), camera64(sp), rel)
        total += k
        mask[:m] |= f.astype(bool)
    return mask, total


def total_loss(s64, specs, tgts, grad=False, counts=None):
    """Summed loss and FP64 gradient; view k sees the first counts[k]
    surfels (the others get no gradient from it)."""
    total = 0.0
    acc = None
    n = s64.pos.shape[0]
    counts = counts or [n] * len(specs)
    for sp, t, m in zip(specs, tgts, counts):
        cam = camera64(sp)
        sv = s64 if m == n else subset(s64, m)
        img = oracle.render(sv, cam)["color"]
        loss, dl, _, _ = oracle.l1_ssim(img, t.astype(np.float64), LAM, DSSIM_SCALE, grad)
        total += loss
        if grad:
            g = oracle.gradients(sv, cam, dl.reshape(-1, 3))
            if acc is None:
                acc = oracle.Gradients(np.zeros((n, 3)), np.zeros((n, 2)), np.zeros((n, 4)), np.zeros(n),
                                       np.zeros(s64.sh.shape))
            for f in ("pos", "scales", "quat", "opacity", "sh"):
                getattr(acc, f)[:m] += getattr(g, f)
    return total, acc


def rendered_scene(tr, sc):
    """The parameters a trainer renders with: its activated buffers hold
    exp / sigmoid of the raw log-scales and opacity logits (evaluated on the
    device), a few ulp from the host arrays it was created from."""
    from splattar_amd import scenes
    ncoef = (sc.sh_degree + 1) ** 2
    return scenes.SceneArrays(tr.positions.read()[:, :3].copy(), tr.scales.read().copy(),
                              tr.quats.read().copy(), tr.opacities.read().copy(),
                              tr.sh.read()[:, :ncoef * 3].copy(), sc.sh_degree)


def trace_views(gpu, sc, specs, counts=None, cap=160):
    """Per view, the contribution lists of the GPU forward with the
    trainer's binning (TIGHT) over the whole image: [(lists, counts, image)]."""
    import splattar_amd as sa
    n = sc.n
    counts = counts or [n] * len(specs)
    out = []
    for sp, m in zip(specs, counts):
        r = sa.Rasterizer(gpu, n, sp.width, sp.height, n * 64)
        try:
            r.upload_scene(sc.positions, sc.scales, sc.quats, sc.opacities, sc.sh)
            r.set_binning(sa._lib.GS_BINNING_TIGHT)
            r.set_trace(0, 0, sp.width, sp.height, cap)
            r.forward(sa.Camera(np.asarray(sp.center, np.float32), np.asarray(sp.quat, np.float32),
                                float(np.float32(sp.focal)), sc.sh_degree), m)
            lists, cnt = r.trace()
            img = r.out_color.read()[..., :3].astype(np.float64)
        finally:
            r.close()
        assert cnt.max() <= cap, f"a pixel composites {cnt.max()} surfels, above the trace cap {cap}"
        out.append((lists, cnt, img))
    return out


def _greedy_ill(moves, g_ref, rel):
    """Per class k, the fewest, largest movers whose exclusion leaves the rest
    moving by at most rel |g_k| (sqrt of the remaining squared moves)."""
    n = next(iter(moves.values())).size
    ill = np.zeros(n, bool)
    for k, mv in moves.items():
        budget = rel * np.linalg.norm(getattr(g_ref, k))
        order = np.argsort(-mv)
        # rest[m] = sqrt of the squared moves of all but the m largest.
        rest = np.sqrt(np.concatenate([np.cumsum((mv[order] ** 2)[::-1])[::-1], [0.0]]))
        m = int(np.argmax(rest <= budget))
        ill[order[:m]] = True
    return ill


def replay_views(s64, specs, tgts, traces, draws=2, rel=1e-5, seed=0):
    """The summed multi-view FP64 gradient along the GPU's decisions
    (trace_views): each view's image replayed, its L1 + D-SSIM gradient taken
    at the replayed image, the parameter gradient replayed and summed. With
    it the conditioning mask: scene and per-view loss gradients scaled by
    1 +- 2^-24 (random signs, one scene draw for all views), the sum replayed
    again, and per class the fewest, largest movers excluded so that the rest
    moves by at most rel normwise (what f32 inputs determine to rel).
    Returns (gradients, ill, replayed images)."""
    n = s64.n
    acc, dls, imgs = None, [], []
    for sp, t, (lists, cnt, _) in zip(specs, tgts, traces):
        cam = camera64(sp)
        img, _, _ = oracle.replay(s64, cam, lists, cnt)
        _, dl, _, _ = oracle.l1_ssim(img, t.astype(np.float64), LAM, DSSIM_SCALE)
        _, _, g = oracle.replay(s64, cam, lists, cnt, dl)
        dls.append(dl)
        imgs.append(img)
        acc = g if acc is None else _add(acc, g)
    rng = np.random.default_rng(seed)
    moves = {k: np.zeros(n) for k in ("pos", "scales", "quat", "opacity", "sh")}
    for _ in range(draws):
        sp64 = s64.copy()
        for arr in (sp64.pos, sp64.scales, sp64.quat, sp64.opacity, sp64.sh):
            arr *= 1.0 + rng.choice([-1.0, 1.0], arr.shape) * 2.0 ** -24
        acc2 = None
        for sp, dl, (lists, cnt, _) in zip(specs, dls, traces):
            dlj = dl * (1.0 + rng.choice([-1.0, 1.0], dl.shape) * 2.0 ** -24)
            _, _, g2 = oracle.replay(sp64, camera64(sp), lists, cnt, dlj)
            acc2 = g2 if acc2 is None else _add(acc2, g2)
        for k in moves:
            d = getattr(acc2, k).reshape(n, -1) - getattr(acc, k).reshape(n, -1)
            moves[k] = np.maximum(moves[k], np.linalg.norm(d, axis=1))
    return acc, _greedy_ill(moves, acc, rel), imgs


def _add(a, b):
    for f in ("pos", "scales", "quat", "opacity", "sh"):
        setattr(a, f, getattr(a, f) + getattr(b, f))
    return a


def replay_conditioning(s64, cam, lists, counts, dl, g_ref, img_ref, draws=2, rel=1e-5, seed=0,
                        dl_ddepth=None, dl_dnormal=None, lambda_dist=0.0):
    """Conditioning of the replayed FP64 evaluation: every input -- the scene
    and the per-pixel loss gradients, whose one-ulp perturbation stands for
    the rounding of each pixel's term in an f32 sum over pixels -- is scaled
    by 1 +- 2^-24 (random signs) and replayed along the same decision lists
    (oracle.replay), `draws` times. Returns (ill, pixel_sens):
    - ill[i]: surfel i is among the fewest, largest movers of some class k
      whose exclusion leaves that class moving by at most rel normwise
      (sqrt of the remaining squared moves <= rel |g_k|): what remains is
      determined by f32 inputs to rel, whatever evaluates it;
    - pixel_sens (H, W): the largest change of the pixel's colour."""
    n = s64.n
    rng = np.random.default_rng(seed)
    sens = np.zeros(img_ref.shape[:2])
    moves = {k: np.zeros(n) for k in ("pos", "scales", "quat", "opacity", "sh")}
    for _ in range(draws):
        sp = s64.copy()
        for arr in (sp.pos, sp.scales, sp.quat, sp.opacity, sp.sh):
            arr *= 1.0 + rng.choice([-1.0, 1.0], arr.shape) * 2.0 ** -24
        def jig(a):
            return None if a is None else np.asarray(a) * (1.0 + rng.choice([-1.0, 1.0], np.shape(a)) * 2.0 ** -24)
        img2, _, g2 = oracle.replay(sp, cam, lists, counts, jig(dl), jig(dl_ddepth), jig(dl_dnormal), lambda_dist)
        sens = np.maximum(sens, np.abs(img2 - img_ref).max(axis=2))
        for k in moves:
            a = getattr(g_ref, k).reshape(n, -1)
            b = getattr(g2, k).reshape(n, -1)
            moves[k] = np.maximum(moves[k], np.linalg.norm(b - a, axis=1))
    return _greedy_ill(moves, g_ref, rel), sens
