"""The benchmark configurations at their own sizes (the multi-view and
streaming ones at full size are in test_gpu_train.py, with their scaled-down
FP64 checks).

C0: plain single-view forward + backward of the 10k-surfel SH0 scene at
256x256 (camera at z = -3 looking at the origin, 60-degree FOV), loss
0.8 L1 + 0.2 (1 - SSIM) against the render of the centre-jittered copy.
C1: 100k surfels, SH3, 1920x1080: projection rects, depth keys, sorted
(tile << 32 | depth) keys and tile ranges bit-exact against the fp32 mirror,
then one forward + backward + Adam step (lr 1e-3, betas 0.9 / 0.999).
C2: the 8-view uneven batch (640x360 .. 1920x1080) on the 100k scene in one
launch per stage: gradients equal the plain per-view path's.

FP64 gradient comparisons follow the GPU's own discrete decisions: the
forward's trace (gs_raster_set_trace) is replayed in FP64 (oracle.replay),
the surfels whose replayed gradient moves by more than 1e-5 relative under a
one-ulp input perturbation (ill-conditioned: no f32 evaluation determines
them) and those compositing at a pixel where the L1 sign differs between the
GPU render and the replayed one, or either is within 1e-5 of the target, are
masked and counted, and the rest are held to per class normwise <= 1e-4.
"""

import numpy as np
import pytest

import oracle
from conftest import announce

pytestmark = pytest.mark.gpu

LAM, DSSIM_SCALE = 0.2, 1.0  # 0.8 L1 + 0.2 (1 - SSIM)


def _scene64(sc):
    import multiview as mv
    return mv.scene64(sc)


def _camera64(sp):
    import multiview as mv
    return mv.camera64(sp)


def _fp64_step(s64, cam, target):
    out = oracle.render(s64, cam)
    loss, dl, l1, ss = oracle.l1_ssim(out["color"], target.astype(np.float64), LAM, DSSIM_SCALE)
    return out, l1, ss, oracle.gradients(s64, cam, dl.reshape(-1, 3))


def _replay_reference(gpu, sc, sp, s64, cam, target, cap):
    """FP64 along the GPU's decisions for one view: a rasterizer with the
    trainer's binning (TIGHT) records the forward's contributions; FP64
    replays them (image, L1 + D-SSIM gradient of that image, parameter
    gradients) and perturbs the inputs by one ulp with the decisions held.
    Returns (gradients, mask, replayed image): the mask holds the surfels
    whose replayed gradient is ill-conditioned and those compositing at a
    pixel where the sign of render - target differs between the GPU render and
    the replayed one, or either is within 1e-5 of the target (the L1 sign is
    then a different discrete choice in the two precisions)."""
    import multiview as mv
    import splattar_amd as sa
    r = sa.Rasterizer(gpu, sc.n, sp.width, sp.height, sc.n * 64)
    r.upload_scene(sc.positions, sc.scales, sc.quats, sc.opacities, sc.sh)
    r.set_binning(sa._lib.GS_BINNING_TIGHT)
    r.set_trace(0, 0, sp.width, sp.height, cap)
    r.forward(sa.Camera(np.asarray(sp.center, np.float32), np.asarray(sp.quat, np.float32),
                        float(np.float32(sp.focal)), sc.sh_degree), sc.n)
    lists, counts = r.trace()
    img_gpu = r.out_color.read()[..., :3].astype(np.float64)
    r.close()
    print(f"  contributions per pixel: mean {counts.mean():.1f}, max {counts.max()} (cap {cap})")
    assert counts.max() <= cap
    img_r, _, _ = oracle.replay(s64, cam, lists, counts)
    _, dl_r, _, _ = oracle.l1_ssim(img_r, target.astype(np.float64), LAM, DSSIM_SCALE)
    _, _, g_r = oracle.replay(s64, cam, lists, counts, dl_r)
    ill, sens = mv.replay_conditioning(s64, cam, lists, counts, dl_r, g_r, img_r)
    t = target.astype(np.float64)
    # The L1 sign: the GPU loss takes it from the GPU image, FP64 from the
    # replayed one; where they differ (or either is within 1e-5 of 0) the
    # pixel's loss gradient is a different discrete choice.
    marg = np.any((np.sign(img_gpu - t) != np.sign(img_r - t)) | (np.abs(img_gpu - t) < 1e-5) |
                  (np.abs(img_r - t) < 1e-5), axis=2).ravel()
    l1 = np.zeros(sc.n, bool)
    for p in np.nonzero(marg)[0]:
        l1[lists[p, :counts[p]] & 0x3FFFFFFF] = True
    mask = ill | l1
    print(f"  surfels masked: {int(ill.sum())} ill-conditioned, {int(l1.sum())} at L1-marginal pixels "
          f"({int(marg.sum())}), {int(mask.sum())} of {sc.n} in all")
    assert mask.mean() < 0.1
    return g_r, mask, img_r, sens


def _trainer_grads(tr, n, ncoef):
    return dict(pos=tr.grad_pos.read()[:, :3].copy(), scales=tr.grad_scales.read().copy(),
                quat=tr.grad_quat.read().copy(), opacity=tr.grad_opacity.read().copy(),
                sh=tr.grad_sh.read()[:, :ncoef * 3].copy())


def _check(got, ref, mask, n):
    """Per class normwise <= 1e-4 over the unmasked surfels."""
    for k in ("pos", "scales", "quat", "opacity", "sh"):
        c = getattr(ref, k).reshape(n, -1)[~mask]
        g = got[k].reshape(n, -1)[~mask]
        nrm = float(np.linalg.norm(g - c) / np.linalg.norm(c))
        print(f"  {k}: normwise rel {nrm:.2e} against the replayed FP64")
        assert nrm <= 1e-4, k


def _rendered_scene(tr, sc):
    import multiview as mv
    return mv.rendered_scene(tr, sc)


def _trainer(gpu, sc, sp, target, lr, entries=0):
    import splattar_amd as sa
    cfg = sa.TrainConfig(lam=LAM, dssim_scale=DSSIM_SCALE, lr_pos=lr, lr_scales=lr, lr_quat=lr,
                         lr_opacity=lr, lr_sh=lr, max_views_per_step=1, entries_per_view=entries)
    t4 = np.zeros((sp.height, sp.width, 4), np.float32)
    t4[..., :3] = target
    cam = sa.Camera(np.asarray(sp.center, np.float32), np.asarray(sp.quat, np.float32),
                    float(np.float32(sp.focal)), sc.sh_degree)
    return sa.Trainer(gpu, cfg, sc.as_init(), [(cam, t4)])


def test_config0_step_matches_fp64(gpu):
    from splattar_amd import scenes
    sc = scenes.config_scene(0)
    sp = scenes.front_view(256, 256)
    tgt = scenes.jittered(sc, 0.01, scenes.CONFIG_SEED ^ 1)
    cam = _camera64(sp)
    target = oracle.render(_scene64(tgt), cam)["color"].astype(np.float32)
    announce("config 0: 10k surfels, 256x256, one plain step")
    tr = _trainer(gpu, sc, sp, target, 0.0)
    sc = _rendered_scene(tr, sc)
    s64 = _scene64(sc)
    out, l1, ss, ref = _fp64_step(s64, cam, target)
    k, _ = oracle.marginal_decisions(s64, cam, 1e-4)
    announce("config 0: decisions of the GPU forward, FP64 replay")
    g_r, mask, img_r, sens = _replay_reference(gpu, sc, sp, s64, cam, target, 256)
    announce("config 0: the step")
    tr.step([0], mode=tr.PLAIN)
    losses = tr.last_losses(1)[0]
    got = _trainer_grads(tr, sc.n, 1)
    img = tr.render(0)
    tr.close()
    err = np.abs(img[..., :3] - out["color"]).max(axis=2)
    print(f"  image: max |d| {err.max():.2e}, pixels above 1e-5: {int((err > 1e-5).sum())}; "
          f"L1 {losses[0]:.8f} vs {l1:.8f}; D-SSIM {losses[1]:.8f} vs {DSSIM_SCALE * (1 - ss):.8f}")
    err_r = np.abs(img[..., :3] - img_r).max(axis=2)
    cond = sens <= 1e-6
    print(f"  image vs replayed FP64: max {err_r.max():.2e}; pixels whose replayed value moves > 1e-6 under a "
          f"one-ulp input perturbation {int((~cond).sum())} of {cond.size}, max error elsewhere "
          f"{err_r[cond].max():.2e}")
    q999 = np.percentile(err_r, 99.9)
    assert err_r[cond].max() <= 1e-5 and q999 <= 3e-5 and err_r.max() <= 3e-3
    assert int((err > 1e-5).sum()) <= k and float(np.median(err)) < 1e-6
    assert abs(losses[0] - l1) <= 1e-6 * l1
    assert abs(losses[1] - DSSIM_SCALE * (1 - ss)) <= 1e-5 * DSSIM_SCALE * (1 - ss)
    _check(got, g_r, mask, sc.n)


def test_config1_keys_bit_exact_and_step(gpu):
    import splattar_amd as sa
    from splattar_amd import scenes
    sc = scenes.config_scene(1)
    sp = scenes.front_view(1920, 1080)
    cam32 = sa.Camera(np.asarray(sp.center, np.float32), np.asarray(sp.quat, np.float32),
                      float(np.float32(sp.focal)), 3)
    # Keys: both exact binnings against the mirror.
    rot = oracle.quat_to_mat(np.asarray(sp.quat, np.float32).astype(np.float64))
    c32 = oracle.cam32(rot, np.asarray(sp.center, np.float32), float(np.float32(sp.focal)), 1920, 1080)
    pos4 = np.concatenate([sc.positions, np.ones((sc.n, 1), np.float32)], 1)
    rects, keys, _ = oracle.prep_f32(pos4, sc.scales, sc.quats, c32)
    k, p, e = oracle.emit_keys(rects, keys, c32.tiles_x)
    k, p = oracle.stable_sort(k, p)
    ranges = oracle.tile_ranges(k, c32.tiles_x * c32.tiles_y)
    for mode in (0, 1):
        announce(f"config 1: keys, binning {mode}")
        r = sa.Rasterizer(gpu, sc.n, 1920, 1080, sc.n * 64)
        r.upload_scene(sc.positions, sc.scales, sc.quats, sc.opacities, sc.sh)
        r.set_binning(mode)
        r.forward(cam32, sc.n)
        d = r.debug()
        r.close()
        assert np.array_equal(d["rects"][:sc.n], rects) and np.array_equal(d["depth_keys"][:sc.n], keys)
        assert d["entries"] == e
        assert np.array_equal(d["sorted_keys"], k) and np.array_equal(d["sorted_payloads"], p)
        assert np.array_equal(d["ranges"], ranges)
    print(f"  {e} entries bit-exact under both binnings")
    # One step: forward, L1 + D-SSIM, backward, Adam (lr 1e-3).
    tgt = scenes.jittered(sc, 0.01, (scenes.CONFIG_SEED + 1) ^ 1)
    cam = _camera64(sp)
    target = oracle.render(_scene64(tgt), cam)["color"].astype(np.float32)
    announce("config 1: one step")
    tr = _trainer(gpu, sc, sp, target, 1e-3)
    sr = _rendered_scene(tr, sc)
    s64 = _scene64(sr)
    out, l1, ss, ref = _fp64_step(s64, cam, target)
    announce("config 1: decisions of the GPU forward, FP64 replay")
    g_r, mask, img_r, sens = _replay_reference(gpu, sr, sp, s64, cam, target, 192)
    announce("config 1: the step")
    tr.step([0], mode=tr.PLAIN)
    losses = tr.last_losses(1)[0]
    got = _trainer_grads(tr, sc.n, 16)
    params = dict(pos=tr.positions.read()[:, :3].copy(), scales=tr.scales.read().copy(),
                  quat=tr.quats.read().copy(), opacity=tr.opacities.read().copy(),
                  sh=tr.sh.read().copy())
    tr.close()
    assert abs(losses[0] - l1) <= 1e-6 * l1
    _check(got, g_r, mask, sc.n)
    # Adam, t = 1, from the FP64 gradient (f32 loop of adam.wgsl:44-74).
    n = sc.n

    def adam(raw, grad, mode):
        raw = raw.astype(np.float32).ravel().copy()
        m = np.zeros_like(raw)
        v = np.zeros_like(raw)
        return oracle.adam_step_f32(raw, grad.astype(np.float32).ravel(), m, v, 1e-3, mode, 1)

    o = np.clip(sc.opacities, 1e-4, 1 - 1e-4)
    want = dict(pos=adam(sc.positions, g_r.pos, 0).reshape(n, 3),
                scales=adam(np.log(np.maximum(sc.scales, 1e-6)), g_r.scales, 1).reshape(n, 2),
                quat=adam(sc.quats, g_r.quat, 0).reshape(n, 4),
                opacity=adam(np.log(o / (1 - o)), g_r.opacity, 2),
                sh=adam(sc.sh, g_r.sh.reshape(n, -1), 0).reshape(n, -1))
    for key, w in want.items():
        g = params[key][:, :w.shape[1]] if w.ndim == 2 else params[key]
        sel = ~mask if w.ndim == 1 else np.repeat(~mask[:, None], w.shape[1], 1)
        off = (np.abs(g - w) > 2e-6 + 1e-5 * np.abs(w)) & sel
        # Adam's first step moves a parameter by ~lr sign(g): only gradients
        # at roundoff level can take the other sign.
        print(f"  Adam {key}: {off.mean():.5f} off")
        assert off.mean() < 0.01, key


def test_config2_uneven_batch_full_size(gpu):
    """8 views of 640x360 .. 1920x1080 on the 100k-surfel SH3 scene, one
    launch per stage (graph): gradients equal the plain per-view step's
    (learning rates 0) per class normwise to 1e-5, losses to 1e-6."""
    import splattar_amd as sa
    from splattar_amd import scenes
    sc = scenes.config_scene(2)
    tgt = scenes.jittered(sc, 0.01, (scenes.CONFIG_SEED + 2) ^ 1)
    sizes = scenes.uneven_sizes(8, scenes.CONFIG_SEED + 2)
    specs = scenes.sphere_views(8, scenes.CONFIG_SEED + 2, sizes)
    announce(f"config 2: targets {sizes}")
    targets = []
    for sp in specs:
        r = sa.Rasterizer(gpu, sc.n, sp.width, sp.height, sc.n * 64)
        r.upload_scene(tgt.positions, tgt.scales, tgt.quats, tgt.opacities, tgt.sh)
        r.forward(sa.Camera(sp.center, sp.quat, sp.focal, 3), sc.n)
        targets.append(r.out_color.read().copy())
        r.close()
    cfg = sa.TrainConfig(lam=LAM, dssim_scale=DSSIM_SCALE, lr_pos=0.0, lr_scales=0.0, lr_quat=0.0,
                         lr_opacity=0.0, lr_sh=0.0, max_views_per_step=8)
    views = [(sa.Camera(sp.center, sp.quat, sp.focal, 3), t) for sp, t in zip(specs, targets)]
    res = {}
    for mode in (sa.Trainer.GRAPH, sa.Trainer.PLAIN):
        announce(f"config 2: mode {mode}")
        t = sa.Trainer(gpu, cfg, sc.as_init(), views)
        t.set_keep_grads(True)
        t.step(list(range(8)), mode=mode)
        t.entry_count()  # raises on overflow
        res[mode] = (_trainer_grads(t, sc.n, 16), t.last_losses(8))
        t.close()
    (gg, lg), (gp, lp) = res[sa.Trainer.GRAPH], res[sa.Trainer.PLAIN]
    np.testing.assert_allclose(lg, lp, rtol=1e-6)
    for k in gg:
        nrm = float(np.linalg.norm(gg[k] - gp[k]) / np.linalg.norm(gp[k]))
        print(f"  {k}: graph vs plain normwise {nrm:.2e}")
        assert nrm <= 1e-5, k


def _c3_problem():
    """The headline configuration: 1M surfels, SH3, scales exp(N(-5, 1)), the
    first four of the 16 sphere views at 1920x1080 (the views of the first
    training step), targets from the centre-jittered copy."""
    from splattar_amd import scenes
    sc = scenes.config_scene(3)
    specs = scenes.sphere_views(16, scenes.CONFIG_SEED + 3, [(1920, 1080)])[:4]
    tgt = scenes.jittered(sc, 0.01, (scenes.CONFIG_SEED + 3) ^ 1)
    return sc, specs, tgt


def _cam32(sp, deg):
    import splattar_amd as sa
    return sa.Camera(np.asarray(sp.center, np.float32), np.asarray(sp.quat, np.float32),
                     float(np.float32(sp.focal)), deg)


def test_config3_view0_keys_bit_exact(gpu):
    """C3 view 0: rects, depth keys, the sorted 64-bit (tile << 32 | depth)
    keys with their payloads and the tile ranges equal the fp32 mirror's bit
    for bit under both exact binnings (SORT64 and PRESORT)."""
    import splattar_amd as sa
    sc, specs, _ = _c3_problem()
    sp = specs[0]
    rot = oracle.quat_to_mat(np.asarray(sp.quat, np.float32).astype(np.float64))
    c32 = oracle.cam32(rot, np.asarray(sp.center, np.float32), float(np.float32(sp.focal)), sp.width,
                       sp.height)
    pos4 = np.concatenate([sc.positions, np.ones((sc.n, 1), np.float32)], 1)
    rects, keys, _ = oracle.prep_f32(pos4, sc.scales, sc.quats, c32)
    k, p, e = oracle.emit_keys(rects, keys, c32.tiles_x)
    k, p = oracle.stable_sort(k, p)
    ranges = oracle.tile_ranges(k, c32.tiles_x * c32.tiles_y)
    print(f"  mirror: {e} entries over {c32.tiles_x * c32.tiles_y} tiles")
    for mode in (sa._lib.GS_BINNING_SORT64, sa._lib.GS_BINNING_PRESORT):
        announce(f"config 3: view 0 keys, binning {mode}")
        r = sa.Rasterizer(gpu, sc.n, sp.width, sp.height, sc.n * 32)
        r.upload_scene(sc.positions, sc.scales, sc.quats, sc.opacities, sc.sh)
        r.set_binning(mode)
        r.forward(_cam32(sp, 3), sc.n)
        d = r.debug()
        r.close()
        assert np.array_equal(d["rects"][:sc.n], rects) and np.array_equal(d["depth_keys"][:sc.n], keys)
        assert d["entries"] == e
        assert np.array_equal(d["sorted_keys"], k) and np.array_equal(d["sorted_payloads"], p)
        assert np.array_equal(d["ranges"], ranges)


def test_config3_tight_binning_and_culling_are_exact(gpu):
    """C3, all four views: the training binning (TIGHT, footprint culling on)
    renders bit-identical images to the reference binning (PRESORT) and to
    TIGHT with culling off; ncontrib is identical with culling on and off."""
    import splattar_amd as sa
    sc, specs, _ = _c3_problem()
    for v, sp in enumerate(specs):
        announce(f"config 3: view {v} tight / presort / culling off")
        out = {}
        for name, mode, cull in (("presort", sa._lib.GS_BINNING_PRESORT, True),
                                 ("tight", sa._lib.GS_BINNING_TIGHT, True),
                                 ("tight_nocull", sa._lib.GS_BINNING_TIGHT, False)):
            r = sa.Rasterizer(gpu, sc.n, sp.width, sp.height, sc.n * 32)
            r.upload_scene(sc.positions, sc.scales, sc.quats, sc.opacities, sc.sh)
            r.set_binning(mode)
            r.set_culling(cull)
            r.forward(_cam32(sp, 3), sc.n)
            out[name] = (r.out_color.read().copy(), r.out_aux.read().copy(), r.out_ncontrib.read().copy(),
                         r.entry_count())
            r.close()
        print(f"  view {v}: entries presort {out['presort'][3]} tight {out['tight'][3]}")
        assert out["tight"][3] < out["presort"][3]
        for j in (0, 1):
            assert np.array_equal(out["tight"][j], out["presort"][j])
            assert np.array_equal(out["tight"][j], out["tight_nocull"][j])
        assert np.array_equal(out["tight"][2], out["tight_nocull"][2])


def test_config3_crop_matches_fp64(gpu):
    """C3 view 0 against FP64 on the centred 480x270 crop. The GPU renders the
    whole 1080p view with the training binning (TIGHT, culling on) and records
    its decisions over the crop (gs_raster_set_trace); the backward gets
    dL/dimage = the FP64 crop loss gradient inside the window, 0 outside.
    - Image against FP64 replaying the GPU's decisions (same hits, branches
      and clamps): max |d| <= 1e-5 on every pixel.
    - Image against faithful FP64: a pixel above 1e-5 must be one whose
      decision list differs between the two (counted and reported).
    - Gradients against the replayed FP64 gradients, per class normwise
      <= 1e-4 over the surfels whose replayed FP64 gradient is well
      conditioned (moves by <= 1e-5 relative under a one-ulp perturbation of
      every input with the decisions held); the excluded count is reported."""
    import splattar_amd as sa
    import multiview as mv
    sc, specs, _ = _c3_problem()
    sp = specs[0]
    W, H, cw, ch = sp.width, sp.height, 480, 270
    x0, y0 = (W - cw) // 2, (H - ch) // 2
    s64 = _scene64(sc)
    cam = _camera64(sp)
    cam.crop = (x0, y0, cw, ch)
    announce("config 3: FP64 crop render and loss gradient")
    img64 = oracle.render(s64, cam)["color"]
    # Target: the render + 0.01 (no clamp), so sign(render - target) = -1.
    _, dl64, _, _ = oracle.l1_ssim(img64, img64 + 0.01, LAM, DSSIM_SCALE)
    announce("config 3: GPU forward (traced) + backward, tight binning")
    cap = 1024
    r = sa.Rasterizer(gpu, sc.n, W, H, sc.n * 32)
    r.upload_scene(sc.positions, sc.scales, sc.quats, sc.opacities, sc.sh)
    r.set_binning(sa._lib.GS_BINNING_TIGHT)
    r.set_trace(x0, y0, cw, ch, cap)
    r.forward(_cam32(sp, 3), sc.n)
    lists, counts = r.trace()
    img = r.out_color.read()[y0:y0 + ch, x0:x0 + cw, :3].astype(np.float64)
    dl = np.zeros((H, W, 4), np.float32)
    dl[y0:y0 + ch, x0:x0 + cw, :3] = dl64
    r.dl_dcolor.write(dl)
    r.backward(sc.n, flags=sa._lib.GS_BWD_COLOR_ONLY)
    got = dict(pos=r.grad_pos.read()[:, :3].copy(), scales=r.grad_scales.read().copy(),
               quat=r.grad_quat.read().copy(), opacity=r.grad_opacity.read().copy(),
               sh=r.grad_sh.read()[:, :48].copy())
    r.close()
    print(f"  contributions per pixel: mean {counts.mean():.1f}, max {counts.max()} (cap {cap})")
    assert counts.max() <= cap
    announce("config 3: FP64 replay of the GPU decisions and its conditioning")
    dl_win = dl[y0:y0 + ch, x0:x0 + cw, :3].astype(np.float64)
    img_r, _, g_r = oracle.replay(s64, cam, lists, counts, dl_win)
    ill, sens = mv.replay_conditioning(s64, cam, lists, counts, dl_win, g_r, img_r)
    err_r = np.abs(img - img_r).max(axis=2)
    cond = sens <= 1e-6
    q50, q99 = np.percentile(err_r, [50, 99])
    print(f"  image vs replayed FP64: max {err_r.max():.2e}, mean {err_r.mean():.2e}, median {q50:.2e}, p99 "
          f"{q99:.2e}, above 1e-5 {int((err_r > 1e-5).sum())}; the replayed FP64 itself moves by "
          f"{np.median(sens):.2e} (median) under a one-ulp input perturbation; pixels where it moves "
          f"<= 1e-6: {int(cond.sum())} of {cond.size}, max error there {err_r[cond].max():.2e}")
    # Faithful FP64: its decisions against the GPU's.
    fl, fc = oracle.trace(s64, cam, cap)
    same = (fc == counts) & np.all(np.where(np.arange(cap)[None, :] < np.minimum(counts, cap)[:, None],
                                            fl == lists, True), axis=1)
    flipped = ~same.reshape(ch, cw)
    err_f = np.abs(img - img64).max(axis=2)
    print(f"  image vs faithful FP64: max {err_f.max():.2e}; pixels above 1e-5 {int((err_f > 1e-5).sum())}; "
          f"pixels whose decisions differ {int(flipped.sum())} of {flipped.size}")
    print(f"  surfels excluded as ill-conditioned: {int(ill.sum())} of {sc.n}")
    nrms = {}
    for key in ("pos", "scales", "quat", "opacity", "sh"):
        want = getattr(g_r, key).reshape(sc.n, -1)[~ill]
        have = got[key].reshape(sc.n, -1).astype(np.float64)[~ill]
        nrms[key] = float(np.linalg.norm(have - want) / np.linalg.norm(want))
        print(f"  {key}: normwise rel {nrms[key]:.2e} against the replayed FP64")
    # Where the decisions agree, replay is the faithful evaluation.
    assert np.abs(img_r - img64).max(axis=2)[~flipped].max() <= 1e-12
    assert flipped.mean() < 0.01
    # The f32 image: the reference's own bars (raster_parity.rs:59-69), then
    # the distribution at this depth of compositing (~70 contributions per
    # pixel), where the FP64 value's own one-ulp spread is ~2e-6.
    assert err_r.max() <= 3e-3 and err_r.mean() <= 3e-4
    assert q50 <= 2e-6 and q99 <= 2e-5 and (err_r > 1e-5).mean() < 0.02
    assert err_r[cond].max() <= 5e-5
    assert ill.mean() < 0.01
    for key, v in nrms.items():
        assert v <= 1e-4, key


def test_config3_graph_step_equals_plain(gpu):
    """C3 one training step over its four views (learning rates 0, gradients
    kept): the batched graph step's summed parameter gradients equal the plain
    per-view step's to 1e-5 normwise per class, losses to 1e-6."""
    import splattar_amd as sa
    sc, specs, tgt = _c3_problem()
    targets = []
    for sp in specs:
        r = sa.Rasterizer(gpu, sc.n, sp.width, sp.height, sc.n * 32)
        r.upload_scene(tgt.positions, tgt.scales, tgt.quats, tgt.opacities, tgt.sh)
        r.forward(_cam32(sp, 3), sc.n)
        targets.append(r.out_color.read().copy())
        r.close()
    cfg = sa.TrainConfig(lam=LAM, dssim_scale=DSSIM_SCALE, lr_pos=0.0, lr_scales=0.0, lr_quat=0.0,
                         lr_opacity=0.0, lr_sh=0.0, max_views_per_step=4, entries_per_view=sc.n * 24)
    views = [(_cam32(sp, 3), t) for sp, t in zip(specs, targets)]
    res = {}
    for mode in (sa.Trainer.GRAPH, sa.Trainer.PLAIN):
        announce(f"config 3: one step, mode {mode}")
        t = sa.Trainer(gpu, cfg, sc.as_init(), views)
        t.set_keep_grads(True)
        t.step([0, 1, 2, 3], mode=mode)
        t.entry_count()  # raises on overflow
        res[mode] = (_trainer_grads(t, sc.n, 16), t.last_losses(4))
        t.close()
    (gg, lg), (gp, lp) = res[sa.Trainer.GRAPH], res[sa.Trainer.PLAIN]
    np.testing.assert_allclose(lg, lp, rtol=1e-6)
    for key in gg:
        nrm = float(np.linalg.norm(gg[key] - gp[key]) / np.linalg.norm(gp[key]))
        print(f"  {key}: graph vs plain normwise {nrm:.2e}")
        assert nrm <= 1e-5, key
