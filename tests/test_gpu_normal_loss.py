"""The normal-consistency loss kernel (gs_normal_loss_*, the reference's
NormalLoss, first_attempt/crates/gs-train/src/shaders/normal_loss.wgsl:52-145)
against the FP64 oracle (oracle.normal_loss, pinned by central differences
and closed forms in test_oracle_normal.py), directly on the same f32 inputs
and end to end through the rasterizer's backward."""

import numpy as np
import pytest

import oracle
from conftest import announce
from microscene import make_scene, scene_from_f32, camera_from_f32

pytestmark = pytest.mark.gpu


def _buffers(gpu, H, W):
    return {k: gpu.alloc(np.float32, (H, W, 4)) for k in ("color", "aux", "normal", "dln", "dlc")}


def _inputs(W, H, focal, seed):
    rng = np.random.default_rng(seed)
    ys, xs = np.mgrid[0:H, 0:W]
    alpha = rng.uniform(0.05, 1.0, (H, W))  # some pixels below alpha_min = 0.2
    t = 2.0 + 0.3 * np.sin(xs / 7.0) + 0.2 * np.cos(ys / 5.0) + rng.normal(0, 0.002, (H, W))
    n = np.stack([rng.normal(0, 0.3, (H, W)), rng.normal(0, 0.3, (H, W)), np.ones((H, W))], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    c4 = np.zeros((H, W, 4), np.float32)
    c4[..., :3] = rng.uniform(0, 1, (H, W, 3))
    c4[..., 3] = alpha
    a4 = np.zeros((H, W, 4), np.float32)
    a4[..., 0] = rng.uniform(0, 1, (H, W))
    a4[..., 1] = alpha * t
    n4 = np.zeros((H, W, 4), np.float32)
    n4[..., :3] = n * alpha[..., None]
    dlc = rng.uniform(-1, 1, (H, W, 4)).astype(np.float32)
    return c4, a4, n4, dlc


@pytest.mark.parametrize("W,H,lam", [(93, 67, 0.8), (64, 48, 0.0)])
def test_normal_loss_matches_fp64_on_the_same_inputs(gpu, W, H, lam):
    """Ragged 93x67 (tiles cut at both borders) and lambda = 0: loss map,
    dl_dnormal and dl_dcolor.w against FP64 on the GPU's f32 inputs; the rgb
    of dl_dcolor is left bit-identical; the loss read back equals
    lambda/(W H) x the FP64 loss-map sum."""
    import splattar_amd as sa
    focal = 70.0
    c4, a4, n4, dlc = _inputs(W, H, focal, 11)
    b = _buffers(gpu, H, W)
    b["color"].write(c4)
    b["aux"].write(a4)
    b["normal"].write(n4)
    b["dlc"].write(dlc)
    b["dln"].write(np.full((H, W, 4), 7.0, np.float32))
    announce(f"normal loss {W}x{H} lambda {lam}")
    nl = sa.NormalLoss(gpu, W, H, b["color"], b["aux"], b["normal"], b["dln"], b["dlc"])
    nl.set_lambda(lam, focal)
    nl.encode()
    loss = nl.read_loss()
    lm = nl.loss_map.read()[..., 0].astype(np.float64)
    dln = b["dln"].read()
    dlc_out = b["dlc"].read()
    nl.close()
    for v in b.values():
        v.free()
    kl = np.float32(lam) / np.float32(W * H)
    rlm, rdn, rdd = oracle.normal_loss(c4, a4, n4, float(np.float32(focal)), float(kl))
    assert np.array_equal(dlc_out[..., :3], dlc[..., :3])
    assert np.all(dln[..., 3] == 0.0)
    if lam == 0.0:
        assert np.all(dln == 0.0) and np.all(dlc_out[..., 3] == 0.0) and np.all(lm == 0.0) and loss == 0.0
        return
    assert (rlm > 0).mean() > 0.3  # non-vacuous
    err_lm = np.abs(lm - rlm).max()
    nrm_dn = np.linalg.norm(dln[..., :3] - rdn) / np.linalg.norm(rdn)
    nrm_dd = np.linalg.norm(dlc_out[..., 3] - rdd) / np.linalg.norm(rdd)
    print(f"  loss map max |d| {err_lm:.2e}; dl_dnormal normwise {nrm_dn:.2e}; dl_ddepth normwise {nrm_dd:.2e}; "
          f"loss {loss:.8e} vs {float(kl) * rlm.sum():.8e}")
    assert err_lm <= 1e-5
    assert nrm_dn <= 1e-5 and nrm_dd <= 1e-4
    # Where FP64 masks a stencil, f32 does too (no stray values).
    assert np.all(dln[..., :3][rdn == 0.0] == 0.0)
    assert abs(loss - float(kl) * rlm.sum()) <= 1e-5 * abs(loss)


@pytest.mark.parametrize("seed,n,deg,size", [(222, 15, 1, 48), (444, 12, 0, 40)])
def test_normal_loss_end_to_end_gradients_match_fp64(gpu, seed, n, deg, size):
    """Forward (depth and normal outputs), normal loss, backward with the
    depth and normal terms, all on the GPU, against FP64: the normal loss on
    the GPU's own forward outputs (oracle.normal_loss; forward differences of
    depth would amplify the f32-vs-FP64 render difference into the loss
    gradient) chained through the backward along the GPU's decisions
    (oracle.replay of the forward's trace with dl_ddepth, dl_dnormal).
    Per-class parameter gradients normwise <= 1e-4; the pixels whose
    decisions differ from the faithful FP64 walk are counted."""
    import splattar_amd as sa
    scene, cam = make_scene(seed, n, deg, size, flavor="parity")
    s32, c32 = scene_from_f32(scene), camera_from_f32(cam)
    lam = 30.0
    announce(f"normal loss end to end seed={seed}")
    r = sa.Rasterizer(gpu, n, size, size, max(n * 64, 1024))
    r.upload_scene(scene.pos.astype(np.float32), scene.scales.astype(np.float32), scene.quat.astype(np.float32),
                   scene.opacity.astype(np.float32), scene.sh.reshape(n, -1).astype(np.float32))
    r.dl_dcolor.write(np.zeros((size, size, 4), np.float32))
    cam32 = sa.Camera(cam.center.astype(np.float32), cam.quat.astype(np.float32), float(np.float32(cam.focal)), deg)
    r.set_trace(0, 0, size, size, 64)
    r.forward(cam32, n)
    lists, counts = r.trace()
    c4, a4, n4 = (b.read().astype(np.float64) for b in (r.out_color, r.out_aux, r.out_normal))
    nl = sa.NormalLoss(gpu, size, size, r.out_color, r.out_aux, r.out_normal, r.dl_dnormal, r.dl_dcolor)
    nl.set_lambda(lam, float(np.float32(cam.focal)))
    nl.encode()
    loss = nl.read_loss()
    r.backward(n, flags=0)
    ncoef = (deg + 1) ** 2
    got = dict(pos=r.grad_pos.read()[:, :3], scales=r.grad_scales.read(), quat=r.grad_quat.read(),
               opacity=r.grad_opacity.read(), sh=r.grad_sh.read()[:, :ncoef * 3].reshape(n, ncoef, 3))
    nl.close()
    r.close()
    kl = float(np.float32(lam) / np.float32(size * size))
    lm, dn, dd = oracle.normal_loss(c4, a4, n4, float(c32.focal), kl)
    faithful = oracle.gradients(s32, c32, np.zeros((size * size, 3)), dd.reshape(-1), dn.reshape(-1, 3))
    # The same FP64 chain along the GPU's decisions (the forward's trace).
    assert counts.max() <= 64
    _, _, ref = oracle.replay(s32, c32, lists, counts, np.zeros((size, size, 3)), dd, dn)
    fl, fc = oracle.trace(s32, c32, 64)
    flips = int((~((fc == counts) & np.all(fl == lists, axis=1))).sum())
    print(f"  loss {loss:.8e} vs FP64 {kl * lm.sum():.8e}; {int((lm > 0).sum())} stencils; pixels whose "
          f"decisions differ from FP64's: {flips}; against faithful FP64: " +
          ", ".join(f"{k} {np.linalg.norm(got[k] - getattr(faithful, k)) / np.linalg.norm(getattr(faithful, k)):.2e}"
                    for k in ("pos", "scales", "quat", "opacity")))
    assert (lm > 0).sum() > 50 and flips <= 4
    assert abs(loss - kl * lm.sum()) <= 1e-5 * abs(kl * lm.sum())
    # The normal loss does not reach the colours: SH gradients are zero.
    assert np.abs(got["sh"]).max() == 0.0 and np.abs(ref.sh).max() == 0.0
    for k in ("pos", "scales", "quat", "opacity"):
        c = getattr(ref, k)
        nrm = float(np.linalg.norm(got[k] - c) / np.linalg.norm(c))
        print(f"  {k}: normwise rel {nrm:.2e}")
        assert nrm <= 1e-4, k
