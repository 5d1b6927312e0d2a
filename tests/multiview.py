"""A multi-view training problem of uneven resolutions (the shape of the
8-view, 360p..1080p sphere-camera configuration, scaled down 10x so the FP64
oracle evaluates it in seconds), and its FP64 loss and gradient: the sum over
views of L1 + D-SSIM (each view's mean over its own pixels), as the batched
trainer step sums per-view gradients (trainer.rs:1139-1332).
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


def targets(tgt, specs):
    """FP64 renders of the target scene, rounded to f32 (what the GPU holds)."""
    s = scene64(tgt)
    return [oracle.render(s, camera64(sp))["color"].astype(np.float32) for sp in specs]


def total_loss(s64, specs, tgts, grad=False):
    total = 0.0
    acc = None
    for sp, t in zip(specs, tgts):
        cam = camera64(sp)
        img = oracle.render(s64, cam)["color"]
        loss, dl, _, _ = oracle.l1_ssim(img, t.astype(np.float64), LAM, DSSIM_SCALE, grad)
        total += loss
        if grad:
            g = oracle.gradients(s64, cam, dl.reshape(-1, 3))
            if acc is None:
                acc = g
            else:
                for f in ("pos", "scales", "quat", "opacity", "sh"):
                    setattr(acc, f, getattr(acc, f) + getattr(g, f))
    return total, acc
