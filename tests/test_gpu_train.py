"""GPU parity of the training-side stages: L1 + D-SSIM loss, fused Adam and
the whole training step in its plain, batched and graph forms.

Loss: against the FP64 oracle restatement of ssim.wgsl (gradient relative
error <= 1e-4 normwise; loss values to 1e-6 relative).
Adam: the reference's own test (first_attempt/crates/gs-train/tests/
adam.rs:37-98): N = 1000, 5 steps, lr 0.01, three activation modes, |d| < 1e-5.
Step: one full step equals the oracle's composition render -> loss ->
gradients -> activation chain -> Adam (f32); batched and graph steps equal
the plain step (same math, different launch structure).
"""

import numpy as np
import pytest

import oracle
from conftest import announce, assert_steps_close

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("W,H,scale", [(37, 29, 0.5), (64, 48, 1.0), (200, 113, 0.5)])
def test_l1_ssim_matches_fp64_oracle(gpu, W, H, scale):
    import splattar_amd as sa
    rng = np.random.default_rng(W * H)
    x = rng.random((H, W, 3)).astype(np.float32)
    y = np.clip(x + rng.normal(0, 0.15, x.shape), 0, 1).astype(np.float32)
    img = gpu.alloc(np.float32, (H, W, 4))
    dl = gpu.alloc(np.float32, (H, W, 4))
    x4 = np.zeros((H, W, 4), np.float32)
    x4[..., :3] = x
    y4 = np.zeros((H, W, 4), np.float32)
    y4[..., :3] = y
    img.write(x4)
    loss = sa.SsimLoss(gpu, W, H, 0.2, img, dl, dssim_scale=scale)
    loss.target.write(y4)
    announce(f"l1_ssim {W}x{H} scale={scale}")
    loss.encode()
    l1, dssim = loss.read_losses()
    got = dl.read()
    ref_loss, ref_dl, ref_l1, ref_ss = oracle.l1_ssim(x.astype(np.float64), y.astype(np.float64), 0.2,
                                                      scale)
    assert abs(l1 - ref_l1) <= 1e-6 * ref_l1
    assert abs(dssim - scale * (1 - ref_ss)) <= 1e-5 * abs(scale * (1 - ref_ss))
    nrm = np.linalg.norm(got[..., :3] - ref_dl) / np.linalg.norm(ref_dl)
    print(f"loss grad normwise rel {nrm:.2e}")
    assert nrm < 1e-4
    assert np.all(got[..., 3] == 0.0)
    loss.close()
    img.free()
    dl.free()


@pytest.mark.parametrize("mode", [0, 1])
def test_trainer_loss_with_target_statistics_equals_the_loss_op(gpu, mode):
    """The trainer's loss reads each target's window mean and second moment
    from a table computed once at creation (target_stats_kernel); the loss op
    (gs_loss, ssim.rs:25-298), whose target the caller may rewrite at any
    time, blurs all five statistics itself. Both sum the same taps in the
    same order, so on the image the trainer renders the two agree exactly:
    L1 and D-SSIM of every view, plain (one launch per view) and batched
    (one launch for all), on sizes that are not multiples of the 74 x 64
    strip (partial strips and bands at the right and bottom edges)."""
    import splattar_amd as sa
    from splattar_amd import scenes
    deg = 2
    sc = scenes.make_scene(2500, deg, -3.4, 0.4, 41)
    tgt = scenes.jittered(sc, 0.01, 42)
    specs = scenes.sphere_views(3, 47, [(150, 97), (96, 64), (37, 29)])
    targets = _render_targets(gpu, tgt, specs, deg)
    t = _trainer(gpu, sc, specs, targets, deg, 3)
    imgs = [t.render(v) for v in range(3)]
    announce(f"trainer loss vs loss op, mode {mode}")
    t.step([0, 1, 2], mode=t.BATCHED if mode == 1 else t.PLAIN)
    got = t.last_losses(3)
    t.close()
    for v, sp in enumerate(specs):
        W, H = sp.width, sp.height
        img = gpu.alloc(np.float32, (H, W, 4))
        dl = gpu.alloc(np.float32, (H, W, 4))
        img.write(imgs[v])
        loss = sa.SsimLoss(gpu, W, H, 0.2, img, dl, dssim_scale=0.5)
        loss.target.write(targets[v])
        loss.encode()
        l1, dssim = loss.read_losses()
        loss.close()
        img.free()
        dl.free()
        print(f"view {v} {W}x{H}: trainer ({got[v, 0]:.9f}, {got[v, 1]:.9f}) op ({l1:.9f}, {dssim:.9f})")
        assert got[v, 0] == l1 and got[v, 1] == dssim


def _ref_xorshift(state):
    x = state[0]
    m = (1 << 64) - 1
    x ^= (x << 13) & m
    x ^= x >> 7
    x ^= (x << 17) & m
    state[0] = x
    return np.float32((x >> 11) / float(1 << 53) * 2.0 - 1.0)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_adam_matches_reference(gpu, mode):
    import splattar_amd as sa
    N, STEPS, LR = 1000, 5, 0.01
    st = [0xADA ^ mode]
    init = np.array([_ref_xorshift(st) * np.float32(0.8) for _ in range(N)], np.float32)
    grads = [np.array([_ref_xorshift(st) for _ in range(N)], np.float32) for _ in range(STEPS)]
    gbuf = gpu.alloc(np.float32, N)
    act = gpu.alloc(np.float32, N)
    opt = sa.Optimizer(gpu)
    opt.add_class("test", N, mode, LR, gbuf, act)
    opt.raw("test").write(init)
    raw = init.copy()
    m = np.zeros(N, np.float32)
    v = np.zeros(N, np.float32)
    announce(f"adam mode={mode}")
    for t, g in enumerate(grads, start=1):
        gbuf.write(g)
        opt.encode_step()
        want_act = oracle.adam_step_f32(raw, g, m, v, LR, mode, t)
    got = opt.raw("test").read()
    assert np.abs(got - raw).max() < 1e-5
    np.testing.assert_allclose(act.read(), want_act, rtol=2e-6, atol=1e-7)
    assert opt.step_count() == STEPS
    opt.close()
    gbuf.free()
    act.free()


def test_optimizer_snapshot_and_restore(gpu):
    """Optimizer::snapshot_layout / encode_snapshot / encode_restore
    (optim.rs:148-165): every class's raw parameters copied out and back in
    add order; restore leaves the activated copies until encode_activate (the
    reference trainer's adaptive-stop restore, trainer.rs:1891-1893)."""
    import splattar_amd as sa
    N = 1000
    opt = sa.Optimizer(gpu)
    g1, a1 = gpu.alloc(np.float32, 3 * N), gpu.alloc(np.float32, 3 * N)
    g2, a2 = gpu.alloc(np.float32, N), gpu.alloc(np.float32, N)
    opt.add_class("pos", 3 * N, opt.IDENTITY, 0.01, g1, a1)
    opt.add_class("opacity", N, opt.SIGMOID, 0.01, g2, a2)
    assert opt.snapshot_layout() == [("pos", 12 * N), ("opacity", 4 * N)]
    rng = np.random.default_rng(5)
    r1 = rng.normal(size=3 * N).astype(np.float32)
    r2 = rng.normal(size=N).astype(np.float32)
    opt.raw("pos").write(r1)
    opt.raw("opacity").write(r2)
    snap = [gpu.alloc(np.float32, 3 * N), gpu.alloc(np.float32, N)]
    announce("optimizer snapshot / restore")
    opt.encode_snapshot(snap)
    g1.write(rng.normal(size=3 * N).astype(np.float32))
    g2.write(rng.normal(size=N).astype(np.float32))
    opt.encode_step()
    stepped = a2.read().copy()
    assert not np.array_equal(opt.raw("pos").read(), r1)
    opt.encode_restore(snap)
    assert np.array_equal(opt.raw("pos").read(), r1)
    assert np.array_equal(opt.raw("opacity").read(), r2)
    assert np.array_equal(a2.read(), stepped)  # restore touches raw only
    opt.encode_activate()
    assert np.array_equal(a1.read(), r1)
    np.testing.assert_allclose(a2.read(), 1.0 / (1.0 + np.exp(-r2.astype(np.float64))), rtol=2e-6)
    with pytest.raises(sa.GsError):
        opt.encode_snapshot(snap[:1])
    opt.close()
    for b in (g1, a1, g2, a2, *snap):
        b.free()


# ---------------------------------------------------------------- full step

def _small_training_problem(n=3000, views=3, size=(96, 64), deg=2, seed=3):
    from splattar_amd import scenes
    sc = scenes.make_scene(n, deg, -3.4, 0.4, seed)
    tgt = scenes.jittered(sc, 0.01, seed ^ 1)
    specs = scenes.sphere_views(views, seed + 7, [size])
    return sc, tgt, specs


def _render_targets(gpu, tgt, specs, deg):
    import splattar_amd as sa
    out = []
    for sp in specs:
        r = sa.Rasterizer(gpu, tgt.n, sp.width, sp.height, tgt.n * 64)
        r.upload_scene(tgt.positions, tgt.scales, tgt.quats, tgt.opacities, tgt.sh)
        r.forward(sa.Camera(sp.center, sp.quat, sp.focal, deg), tgt.n)
        out.append(r.out_color.read().copy())
        r.close()
    return out


def _trainer(gpu, sc, specs, targets, deg, views_per_step):
    import splattar_amd as sa
    cfg = sa.TrainConfig(lam=0.2, dssim_scale=0.5, lr_pos=1e-3, lr_scales=1e-3, lr_quat=1e-3,
                         lr_opacity=1e-3, lr_sh=1e-3, max_views_per_step=views_per_step)
    views = [(sa.Camera(sp.center, sp.quat, sp.focal, deg), t) for sp, t in zip(specs, targets)]
    return sa.Trainer(gpu, cfg, sc.as_init(), views)


def _params(t):
    return dict(pos=t.positions.read()[:, :3].copy(), scales=t.scales.read().copy(),
                quats=t.quats.read().copy(), opac=t.opacities.read().copy(), sh=t.sh.read().copy())


@pytest.mark.parametrize("mode", [0, 2])
def test_train_step_matches_oracle_composition(gpu, mode):
    """One step over 2 views == oracle render/loss/gradients/Adam, plain
    (geometry backward and Adam as separate kernels) and graph (fused)."""
    deg = 2
    sc, tgt, specs = _small_training_problem(n=1500, views=2, size=(64, 48), deg=deg)
    announce("train step vs oracle: targets")
    targets = _render_targets(gpu, tgt, specs, deg)
    t = _trainer(gpu, sc, specs, targets, deg, 2)
    announce(f"train step vs oracle: mode {mode}")
    t.step([0, 1], mode=mode)
    got = _params(t)
    losses = t.last_losses(2)
    t.close()
    # Oracle composition on the same f32 inputs.
    n = sc.n
    ncoef = (deg + 1) ** 2
    s64 = oracle.Scene(sc.positions.astype(np.float64), sc.scales.astype(np.float64),
                       sc.quats.astype(np.float64), sc.opacities.astype(np.float64),
                       sc.sh.astype(np.float64).reshape(n, ncoef, 3), deg)
    acc = None
    for k, sp in enumerate(specs):
        cam = oracle.Camera(np.asarray(sp.center, np.float32).astype(np.float64),
                            np.asarray(sp.quat, np.float32).astype(np.float64),
                            float(np.float32(sp.focal)), sp.width, sp.height)
        img = oracle.render(s64, cam)["color"]
        loss, dl, l1, ss = oracle.l1_ssim(img, targets[k][..., :3].astype(np.float64), 0.2, 0.5)
        assert abs(losses[k, 0] - l1) < 1e-5 * max(l1, 1e-3)
        # D-SSIM from the trainer's precomputed target statistics.
        dssim = 0.5 * (1.0 - ss)
        assert abs(losses[k, 1] - dssim) < 1e-5 * max(dssim, 1e-3)
        g = oracle.gradients(s64, cam, dl.reshape(-1, 3))
        if acc is None:
            acc = g
        else:
            for f in ("pos", "scales", "quat", "opacity", "sh"):
                setattr(acc, f, getattr(acc, f) + getattr(g, f))
    # Activation chains + Adam (t = 1) in f32 (adam.wgsl:44-74).
    lr = 1e-3

    def adam(raw, grad, mode):
        raw = raw.astype(np.float32).ravel().copy()
        m = np.zeros_like(raw)
        v = np.zeros_like(raw)
        act = oracle.adam_step_f32(raw, grad.astype(np.float32).ravel(), m, v, lr, mode, 1)
        return act

    want_pos = adam(sc.positions, acc.pos, 0).reshape(n, 3)
    o = np.clip(sc.opacities, 1e-4, 1 - 1e-4)
    want_op = adam(np.log(o / (1 - o)), acc.opacity, 2)
    want_sc = adam(np.log(np.maximum(sc.scales, 1e-6)), acc.scales, 1).reshape(n, 2)
    want_q = adam(sc.quats, acc.quat, 0).reshape(n, 4)
    want_sh = adam(sc.sh, acc.sh.reshape(n, -1), 0).reshape(n, -1)
    # Adam's first step moves every parameter with a non-negligible gradient by
    # ~lr; sign agreement is the sharp check, the value check is to 2e-6.
    for label, got_v, want_v in (("pos", got["pos"], want_pos), ("opac", got["opac"], want_op),
                                 ("scales", got["scales"], want_sc), ("quats", got["quats"], want_q),
                                 ("sh", got["sh"][:, :ncoef * 3], want_sh)):
        err = np.abs(got_v - want_v)
        frac = float((err > 2e-6 + 1e-5 * np.abs(want_v)).mean())
        print(f"  step {label}: max err {err.max():.2e}, frac off {frac:.4f}")
        assert frac < 0.01, label


def test_batched_and_graph_steps_equal_plain(gpu):
    deg = 2
    sc, tgt, specs = _small_training_problem(deg=deg)
    announce("batched vs plain: targets")
    targets = _render_targets(gpu, tgt, specs, deg)
    results = {}
    for mode in (0, 1, 2):
        t = _trainer(gpu, sc, specs, targets, deg, 3)
        announce(f"batched vs plain: mode {mode}")
        for it in range(3):
            t.step([it % 3, (it + 1) % 3, (it + 2) % 3], mode=mode)
        results[mode] = (_params(t), t.last_losses(3), t.entry_count())
        t.close()
    p0, l0, e0 = results[0]
    for mode in (1, 2):
        p, l, e = results[mode]
        assert e == e0
        np.testing.assert_allclose(l, l0, rtol=1e-5)
        # Same math; float atomics make the order of gradient sums differ.
        assert_steps_close(p, p0, f"mode {mode}")


def test_fused_geometry_adam_applies_adam_to_the_summed_gradients(gpu):
    """Batched steps run the geometry backward and Adam as one kernel. With
    the gradients kept, two graph steps equal the oracle's f32 Adam
    (adam.wgsl:44-74 with the optim.rs bias corrections) applied to the kept
    gradients on the identity classes (position, quaternion, SH); log-scales
    and opacity logits move by the Adam step of their chained gradients
    (dact = s, o(1 - o)); keeping the gradients changes nothing else."""
    deg = 2
    lr = 1e-3
    sc, tgt, specs = _small_training_problem(deg=deg)
    targets = _render_targets(gpu, tgt, specs, deg)
    t = _trainer(gpu, sc, specs, targets, deg, 3)
    t.set_keep_grads(True)
    p0 = _params(t)
    raw = dict(pos=t.positions.read().copy(), quats=t.quats.read().copy(), sh=t.sh.read().copy())
    m = {k: np.zeros_like(v) for k, v in raw.items()}
    v = {k: np.zeros_like(x) for k, x in raw.items()}
    for it in range(2):
        announce(f"fused geometry + adam: step {it}")
        t.step([0, 1, 2], mode=t.GRAPH)
        g = dict(pos=t.grad_pos.read().copy(), quats=t.grad_quat.read().copy(), sh=t.grad_sh.read().copy())
        assert np.all(g["pos"][:, 3] == 0.0)
        for k in raw:
            r, gg = raw[k].reshape(-1), g[k].reshape(-1)
            mm, vv = m[k].reshape(-1), v[k].reshape(-1)
            oracle.adam_step_f32(r, gg, mm, vv, lr, 0, it + 1)
        if it == 0:
            gs, go = t.grad_scales.read().copy(), t.grad_opacity.read().copy()
            p1 = _params(t)
    got = dict(pos=t.positions.read(), quats=t.quats.read(), sh=t.sh.read())
    for k in raw:
        err = np.abs(got[k] - raw[k]).max()
        print(f"  {k}: max |fused - host Adam| {err:.2e}")
        np.testing.assert_allclose(got[k], raw[k], rtol=0, atol=4e-7, err_msg=k)
    # First step (t = 1): raw moves by -lr * g / (|g| + eps) up to f32 rounding.
    gr = gs.astype(np.float64) * p0["scales"]
    want = -lr * gr / (np.abs(gr) + 1e-8)
    np.testing.assert_allclose(np.log(p1["scales"] / p0["scales"].astype(np.float64)), want, rtol=0,
                               atol=1e-6)
    o0, o1 = p0["opac"].astype(np.float64), p1["opac"].astype(np.float64)
    mid = (o0 > 0.01) & (o0 < 0.99)
    gr = go.astype(np.float64) * o0 * (1 - o0)
    want = -lr * gr / (np.abs(gr) + 1e-8)
    logit = lambda o: np.log(o / (1 - o))
    np.testing.assert_allclose((logit(o1) - logit(o0))[mid], want[mid], rtol=0, atol=1e-5)
    final = _params(t)
    t.close()
    # The same two steps without kept gradients (float atomics aside).
    t2 = _trainer(gpu, sc, specs, targets, deg, 3)
    for it in range(2):
        t2.step([0, 1, 2], mode=t2.GRAPH)
    q = _params(t2)
    for k in q:
        np.testing.assert_allclose(q[k], final[k], rtol=1e-4, atol=2e-6, err_msg=k)
    t2.close()


SCHEDULES = [(8, 8), (8, 4), (8, 0), (4, 8), (4, 4), (4, 0)]


def test_every_schedule_gives_the_same_step(gpu):
    """The autotuner's candidates (digit width x raster scheduling) change
    only launch shapes: three graph steps give the same parameters and
    losses under each, and autotune picks one of them."""
    deg = 2
    sc, tgt, specs = _small_training_problem(deg=deg)
    targets = _render_targets(gpu, tgt, specs, deg)
    results = {}
    for sched in SCHEDULES:
        t = _trainer(gpu, sc, specs, targets, deg, 3)
        t.set_schedule(*sched)
        announce(f"schedule digit={sched[0]} blocks_per_cu={sched[1]}")
        for it in range(3):
            t.step([it % 3, (it + 1) % 3, (it + 2) % 3], mode=t.GRAPH)
        results[sched] = (_params(t), t.last_losses(3), t.entry_count())
        t.close()
    p0, l0, e0 = results[SCHEDULES[0]]
    for sched in SCHEDULES[1:]:
        p, l, e = results[sched]
        assert e == e0
        np.testing.assert_allclose(l, l0, rtol=1e-5)
        assert_steps_close(p, p0, f"{sched}")
    t = _trainer(gpu, sc, specs, targets, deg, 3)
    t.step([0, 1, 2], mode=t.GRAPH)
    before = _params(t)
    announce("autotune")
    chosen = t.autotune([0, 1, 2], reps=1)
    # Tuning restores the training state it found.
    after = _params(t)
    for k in before:
        np.testing.assert_array_equal(after[k], before[k], err_msg=k)
    assert (chosen["digit_bits"], chosen["blocks_per_cu"]) in SCHEDULES
    assert (chosen["tile_size"], chosen["fwd_batch"], chosen["bwd_batch"]) in TILINGS
    t.step([0, 1, 2], mode=t.GRAPH)
    assert np.isfinite(t.last_losses(3)).all()
    t.close()


TILINGS = [(16, 256, 64), (16, 256, 128), (16, 128, 64), (32, 256, 128), (32, 512, 256), (8, 128, 128),
           (8, 64, 64)]


def test_overlapped_step_equals_in_order_step(gpu):
    """The overlapped batched step (each half's loss on a second stream beside
    the other half's raster pass) against the same step in stream order:
    three graph steps and three plain batched steps over 4 views of two
    sizes give the same parameters and losses up to float-atomic order, and
    every graph replay agrees with the eager overlapped step."""
    import splattar_amd as sa
    deg = 2
    sc, tgt, specs = _small_training_problem(n=3000, views=4, size=(96, 64), deg=deg, seed=5)
    specs[1].width, specs[1].height = 80, 48
    specs[3].width, specs[3].height = 80, 48
    targets = _render_targets(gpu, tgt, specs, deg)
    results = {}
    for overlap in (True, False):
        for mode in (sa.Trainer.GRAPH, sa.Trainer.BATCHED):
            t = _trainer(gpu, sc, specs, targets, deg, 4)
            t.set_overlap(overlap)
            announce(f"overlap={overlap} mode={mode}")
            per_step = []
            for it in range(3):
                t.step([(it + j) % 4 for j in range(4)], mode=mode)
                per_step.append((_params(t), t.last_losses(4)))
            results[(overlap, mode)] = per_step
            t.close()
    ref = results[(False, sa.Trainer.BATCHED)]
    for key, steps in results.items():
        for it, ((p, l), (p0, l0)) in enumerate(zip(steps, ref)):
            np.testing.assert_allclose(l, l0, rtol=1e-5, err_msg=f"{key} step {it}")
            assert_steps_close(p, p0, f"{key} step {it}")


def test_every_work_tiling_gives_the_same_images_and_step(gpu):
    """Work tiles of 8, 16 and 32 px and every LDS batch size: the per-pixel
    compositing sequence is the same (coverage stays the 16-px rects of the
    forward model, culling is exact), so renders are bit-identical; three
    graph steps give the same parameters and losses up to float-atomic
    order."""
    deg = 2
    sc, tgt, specs = _small_training_problem(n=4000, deg=deg, size=(150, 90))
    targets = _render_targets(gpu, tgt, specs, deg)
    results = {}
    for tl in TILINGS:
        t = _trainer(gpu, sc, specs, targets, deg, 3)
        t.set_tiling(*tl)
        announce(f"tiling tile={tl[0]} fwd_batch={tl[1]} bwd_batch={tl[2]}")
        imgs = [t.render(v) for v in range(3)]
        for it in range(3):
            t.step([it % 3, (it + 1) % 3, (it + 2) % 3], mode=t.GRAPH)
        results[tl] = (imgs, _params(t), t.last_losses(3))
        t.close()
    i0, p0, l0 = results[TILINGS[0]]
    for tl in TILINGS[1:]:
        imgs, p, l = results[tl]
        for v in range(3):
            np.testing.assert_array_equal(imgs[v][..., :3], i0[v][..., :3], err_msg=f"{tl} view {v}")
            np.testing.assert_array_equal(imgs[v][..., 3], i0[v][..., 3], err_msg=f"{tl} view {v} alpha")
        np.testing.assert_allclose(l, l0, rtol=1e-5)
        assert_steps_close(p, p0, f"{tl}")


def test_training_reduces_loss(gpu):
    deg = 1
    sc, tgt, specs = _small_training_problem(n=2000, views=3, size=(64, 64), deg=deg, seed=11)
    targets = _render_targets(gpu, tgt, specs, deg)
    t = _trainer(gpu, sc, specs, targets, deg, 3)
    announce("training loop 60 steps")
    first = None
    for it in range(60):
        t.step([0, 1, 2], mode=t.BATCHED)
        if it == 0:
            first = t.last_losses(3)[:, 0].mean()
    last = t.last_losses(3)[:, 0].mean()
    print(f"L1 {first:.5f} -> {last:.5f}")
    assert last < 0.7 * first
    t.close()


def _replayed_reference(gpu, tr, sc, specs, tgts, counts=None):
    """FP64 summed gradient of a multi-view step along the GPU's discrete
    decisions, and the surfels to leave out of the comparison.

    The trainer's rendered parameters (its activated buffers) are traced per
    view with the trainer's binning and replayed in FP64 (multiview.
    replay_views); the L1 sign is pinned by the target offset (checked
    here). Masked: the surfels whose replayed gradient is ill-conditioned
    (moves by more than 1e-5 of its class norm under one-ulp perturbations of
    the inputs with the decisions held)."""
    import multiview as mv
    sr = mv.rendered_scene(tr, sc)
    s64 = mv.scene64(sr)
    margin = mv.l1_sign_margin(s64, specs, tgts, counts)
    assert margin < -0.1, f"L1 sign not pinned (max render - target {margin})"
    traces = mv.trace_views(gpu, sr, specs, counts)
    ref, ill, imgs = mv.replay_views(s64, specs, tgts, traces)
    err = max(float(np.abs(g - r).max()) for (_, _, g), r in zip(traces, imgs))
    print(f"  GPU images vs replayed FP64: max {err:.2e}; surfels ill-conditioned {int(ill.sum())} of {ill.size}")
    assert err <= 1e-4 and ill.mean() < 0.1
    return ref, ill


def _check_step_grads(tr, sc, ref, mask):
    """Per class normwise <= 1e-4 over the unmasked surfels."""
    ncoef = (sc.sh_degree + 1) ** 2
    got = dict(pos=tr.grad_pos.read()[:, :3], scales=tr.grad_scales.read(), quat=tr.grad_quat.read(),
               opacity=tr.grad_opacity.read(), sh=tr.grad_sh.read()[:, :ncoef * 3])
    for k in ("pos", "scales", "quat", "opacity", "sh"):
        c = getattr(ref, k).reshape(got[k].shape)[~mask]
        nrm = float(np.linalg.norm(got[k][~mask] - c) / np.linalg.norm(c))
        print(f"  {k}: normwise rel {nrm:.2e} against the replayed FP64")
        assert nrm <= 1e-4, k


@pytest.mark.parametrize("tiling", [(16, 256, 64), (32, 256, 128), (8, 64, 64)])
def test_growing_scene_frame_batch_matches_fp64(gpu, tiling):
    """The streaming-frame configuration scaled down: 16 frames of a camera
    on an SE(3) walk (1 cm / 0.5 deg per frame) over a scene that grows 10 %
    per frame (frame k renders its first 500 + 50 k surfels), all frames in
    one captured batched step, at 8-, 16- and 32-px work tiles. Learning
    rates 0; gradients summed over frames against FP64 along the GPU's
    decisions, per class normwise <= 1e-4 over the well-conditioned surfels
    (see _replayed_reference)."""
    import multiview as mv
    import splattar_amd as sa
    sc, tgt, specs, counts = mv.frames_problem()
    tgts = mv.targets(tgt, specs, counts, offset=2.0)
    cfg = sa.TrainConfig(lam=mv.LAM, dssim_scale=mv.DSSIM_SCALE, lr_pos=0.0, lr_scales=0.0,
                         lr_quat=0.0, lr_opacity=0.0, lr_sh=0.0, max_views_per_step=len(specs))
    views = []
    for sp, t, m in zip(specs, tgts, counts):
        t4 = np.zeros((sp.height, sp.width, 4), np.float32)
        t4[..., :3] = t
        views.append((sa.Camera(sp.center, sp.quat, sp.focal, sc.sh_degree), t4, m))
    tr = sa.Trainer(gpu, cfg, sc.as_init(), views)
    announce(f"growing-scene 16 frames: traces and FP64 replay, counts {counts[0]}..{counts[-1]}")
    ref, mask = _replayed_reference(gpu, tr, sc, specs, tgts, counts)
    tr.set_keep_grads(True)
    tr.set_tiling(*tiling)
    announce(f"growing-scene 16-frame batch, tiling {tiling}")
    tr.step(list(range(len(specs))), mode=tr.GRAPH)
    _check_step_grads(tr, sc, ref, mask)
    tr.close()


def test_streaming_frame_batch_full_size(gpu):
    """The streaming-frame configuration at full size: 16 frames of
    1280x720 (SE(3) walk), 500k surfels growing 10 % per frame (1.25M), all
    frames in one graph launch with the autotuned schedule. With learning
    rates 0 (the parameters stay put), the graph step's gradients equal the
    plain per-frame step's (same math, per-frame launches, default
    schedule) per class normwise to 1e-5 (float-atomic order), losses to
    1e-6; no frame overflows its entry capacity."""
    import splattar_amd as sa
    from splattar_amd import scenes
    counts = scenes.frame_surfel_counts(16, 500_000)
    n = counts[-1]
    sc = scenes.config_scene(4, n)
    tgt = scenes.jittered(sc, 0.01, (scenes.CONFIG_SEED + 4) ^ 1)
    specs = scenes.walk_views(16, scenes.CONFIG_SEED + 4, (1280, 720))
    announce("streaming batch full size: targets")
    r = sa.Rasterizer(gpu, n, 1280, 720, 40_000_000)
    r.upload_scene(tgt.positions, tgt.scales, tgt.quats, tgt.opacities, tgt.sh)
    targets = []
    for sp, m in zip(specs, counts):
        r.forward(sa.Camera(sp.center, sp.quat, sp.focal, 3), m)
        targets.append(r.out_color.read().copy())
    r.close()
    cfg = sa.TrainConfig(lam=0.2, dssim_scale=1.0, lr_pos=0.0, lr_scales=0.0, lr_quat=0.0, lr_opacity=0.0,
                         lr_sh=0.0, max_views_per_step=16, entries_per_view=12_000_000)
    views = [(sa.Camera(sp.center, sp.quat, sp.focal, 3), t, m) for sp, t, m in zip(specs, targets, counts)]
    ids = list(range(16))
    results = {}
    for mode in (sa.Trainer.GRAPH, sa.Trainer.PLAIN):
        t = sa.Trainer(gpu, cfg, sc.as_init(), views)
        t.set_keep_grads(True)
        if mode == sa.Trainer.GRAPH:
            announce("streaming batch full size: autotune")
            chosen = t.autotune(ids, reps=1)
            print(f"  autotune chose {chosen}")
        announce(f"streaming batch full size: mode {mode}")
        t.step(ids, mode=mode)
        ents = t.entry_count()  # raises on overflow
        grads = dict(pos=t.grad_pos.read()[:, :3].copy(), scales=t.grad_scales.read().copy(),
                     quat=t.grad_quat.read().copy(), opacity=t.grad_opacity.read().copy(),
                     sh=t.grad_sh.read().copy())
        results[mode] = (grads, t.last_losses(16), ents)
        t.close()
    (gg, lg, eg), (gp, lp, ep) = results[sa.Trainer.GRAPH], results[sa.Trainer.PLAIN]
    print(f"  entries per step: graph {eg}, plain {ep}")
    np.testing.assert_allclose(lg, lp, rtol=1e-6)
    for k in gg:
        nrm = float(np.linalg.norm(gg[k] - gp[k]) / np.linalg.norm(gp[k]))
        print(f"  {k}: graph vs plain normwise {nrm:.2e}")
        assert nrm <= 1e-5, k


def test_uneven_multiview_batch_matches_fp64(gpu):
    """One captured batched step over 8 views of uneven resolution (one launch
    per stage for all views, gradients summed per surfel) against the FP64
    sum of per-view gradients along the GPU's decisions; learning rates 0
    leave the parameters at the state the gradient is taken at. Per class
    normwise <= 1e-4 over the well-conditioned surfels (see
    _replayed_reference)."""
    import multiview as mv
    import splattar_amd as sa
    sc, tgt, specs = mv.problem(n=1200, seed=17)
    tgts = mv.targets(tgt, specs, offset=2.0)
    cfg = sa.TrainConfig(lam=mv.LAM, dssim_scale=mv.DSSIM_SCALE, lr_pos=0.0, lr_scales=0.0,
                         lr_quat=0.0, lr_opacity=0.0, lr_sh=0.0, max_views_per_step=mv.VIEWS)
    views = []
    for sp, t in zip(specs, tgts):
        t4 = np.zeros((sp.height, sp.width, 4), np.float32)
        t4[..., :3] = t
        views.append((sa.Camera(sp.center, sp.quat, sp.focal, sc.sh_degree), t4))
    tr = sa.Trainer(gpu, cfg, sc.as_init(), views)
    announce("uneven 8-view batch: traces and FP64 replay")
    ref, mask = _replayed_reference(gpu, tr, sc, specs, tgts)
    tr.set_keep_grads(True)
    announce(f"uneven 8-view batch: sizes {[(sp.width, sp.height) for sp in specs]}")
    tr.step(list(range(mv.VIEWS)), mode=tr.GRAPH)
    _check_step_grads(tr, sc, ref, mask)
    tr.close()
