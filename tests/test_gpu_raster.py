"""GPU rasterizer forward / backward parity against the FP64 oracle.

Restates crates/gs-kernels/tests/raster_parity.rs: the same seeds, scene
generator (tests/common/mod.rs) and loss weights; the oracle is evaluated on
the f32-rounded inputs the GPU sees. Tolerances are written per assertion:
  * image: max |gpu - fp64| <= 1e-5 except at isolated decision flips
    (alpha vs 1/255, alpha clamp, T < 1e-4, branch choice, N5 of the
    design notes), which are counted and bounded by the reference's own
    3e-3 per-channel limit (raster_parity.rs:59-69);
  * gradients: per class normwise relative error <= 1e-4, and per element
    err <= 1e-5 + 2e-3 max(|a|,|b|) (raster_parity.rs:101-113, 219). Where
    the FP64 gradient itself moves by more under unit-roundoff (2^-24)
    perturbations of the inputs, no f32 evaluation can be held to 1e-4: the
    normwise bound then becomes 4x that measured sensitivity (max of three
    perturbation trials, printed beside the error);
    aux losses 6e-3 / camera 3e-2 per element (raster_parity.rs:251).
Projection rects / depth keys, sorted 64-bit keys and tile ranges are
bit-exact against the fp32 mirror.
"""

import numpy as np
import pytest

import oracle
from conftest import announce
from microscene import XorShift64, make_scene, scene_from_f32, camera_from_f32

pytestmark = pytest.mark.gpu


def _gpu_inputs(scene):
    n = scene.n
    return (scene.pos.astype(np.float32), scene.scales.astype(np.float32),
            scene.quat.astype(np.float32), scene.opacity.astype(np.float32),
            scene.sh.reshape(n, -1).astype(np.float32))


def _camera(cam, deg):
    import splattar_amd as sa
    return sa.Camera(cam.center.astype(np.float32), cam.quat.astype(np.float32), float(np.float32(cam.focal)),
                     deg)


def _rasterizer(gpu, scene, cam, max_entries=None):
    import splattar_amd as sa
    n = scene.n
    r = sa.Rasterizer(gpu, n, cam.width, cam.height, max_entries or max(n * 64, 1024))
    r.upload_scene(*_gpu_inputs(scene))
    return r


def _forward_case(gpu, seed, n, deg, size, flavor="parity", binning=None, trace=False):
    scene, cam = make_scene(seed, n, deg, size, flavor=flavor)
    s32, c32 = scene_from_f32(scene), camera_from_f32(cam)
    announce(f"forward seed={seed} n={n} deg={deg} size={size} binning={binning}")
    r = _rasterizer(gpu, scene, cam)
    if binning is not None:
        r.set_binning(binning)
    if trace:
        r.set_trace(0, 0, size, size, 256)
    r.forward(_camera(cam, deg), n)
    if trace:
        cam.trace = r.trace()
    color = r.out_color.read()
    aux = r.out_aux.read()
    normal = r.out_normal.read()
    dbg = r.debug()
    ref = oracle.render(s32, c32)
    r.close()
    return scene, cam, color, aux, normal, dbg, ref


def _image_stats(got, want):
    d = np.abs(got - want)
    return float(d.max()), float(d.mean()), int((d > 1e-5).sum())


def _flipped_pixels(s32, c32, lists, counts, cap=256):
    """Pixels whose GPU contribution list differs from the faithful FP64
    walk's (a discrete decision taken the other way)."""
    fl, fc = oracle.trace(s32, c32, cap)
    same = (fc == counts) & np.all(np.where(np.arange(cap)[None, :] < np.minimum(counts, cap)[:, None],
                                            fl == lists, True), axis=1)
    return ~same.reshape(c32.height, c32.width)


@pytest.mark.parametrize("seed,n,deg,size", [(101, 20, 0, 64), (202, 20, 1, 64), (303, 12, 3, 48),
                                             (404, 50, 1, 128)])
def test_forward_matches_fp64_oracle(gpu, seed, n, deg, size):
    """Every pixel within 1e-5 of FP64 (colour, alpha, normal; depth sum w t
    in scene units, relative to the pixel's mean depth), except at pixels
    whose decision list differs
    from the FP64 walk's; against FP64 replaying the GPU's decisions, every
    pixel's colour and alpha within 1e-5."""
    scene, cam, color, aux, normal, dbg, ref = _forward_case(gpu, seed, n, deg, size, trace=True)
    s32, c32 = scene_from_f32(scene), camera_from_f32(cam)
    lists, counts = cam.trace
    flipped = _flipped_pixels(s32, c32, lists, counts)
    img_r, alpha_r, _ = oracle.replay(s32, c32, lists, counts)
    mx, mean, over = _image_stats(color[..., :3], ref["color"])
    rmx, _, _ = _image_stats(color[..., :3], img_r)
    ramx, _, _ = _image_stats(color[..., 3], alpha_r)
    ok = ~flipped
    errs = {"color": np.abs(color[..., :3] - ref["color"]).max(axis=2),
            "alpha": np.abs(color[..., 3] - ref["alpha"]),
            "depth": np.abs(aux[..., 1] - ref["depth"]) /
                     np.maximum(1.0, np.abs(ref["depth"]) / np.maximum(ref["alpha"], 1e-12)),
            "normal": np.abs(normal[..., :3] - ref["normal"]).max(axis=2)}
    print(f"seed {seed}: color max {mx:.2e} mean {mean:.2e} px>1e-5 {over}; decision flips {int(flipped.sum())}; "
          f"vs replay color {rmx:.2e} alpha {ramx:.2e}; at agreeing pixels " +
          ", ".join(f"{k} {v[ok].max():.2e}" for k, v in errs.items()))
    assert counts.max() <= 256
    assert rmx <= 1e-5 and ramx <= 1e-5
    for k, v in errs.items():
        assert v[ok].max() <= 1e-5, k
    assert flipped.sum() <= 2 and mean < 3e-6
    # T_final + sum(w) = 1 (the compositing invariant) up to f32 rounding.
    assert np.abs(aux[..., 0] + color[..., 3] - 1.0).max() < 2e-6


@pytest.mark.parametrize("binning", [0, 1])
def test_projection_and_binning_bit_exact(gpu, binning):
    scene, cam, color, aux, normal, dbg, ref = _forward_case(gpu, 404, 50, 1, 128, binning=binning)
    rot = oracle.quat_to_mat(cam.quat.astype(np.float32).astype(np.float64))
    c32 = oracle.cam32(rot, cam.center.astype(np.float32), cam.focal, cam.width, cam.height)
    pos4 = np.concatenate([scene.pos, np.ones((scene.n, 1))], 1).astype(np.float32)
    rects, keys, counts = oracle.prep_f32(pos4, scene.scales.astype(np.float32),
                                          scene.quat.astype(np.float32), c32)
    assert np.array_equal(dbg["rects"][:scene.n], rects)
    assert np.array_equal(dbg["depth_keys"][:scene.n], keys)
    k, p, e = oracle.emit_keys(rects, keys, c32.tiles_x)
    k, p = oracle.stable_sort(k, p)
    assert dbg["entries"] == e
    assert np.array_equal(dbg["sorted_keys"], k)
    assert np.array_equal(dbg["sorted_payloads"], p)
    assert np.array_equal(dbg["ranges"], oracle.tile_ranges(k, c32.tiles_x * c32.tiles_y))


@pytest.mark.parametrize("digit_bits", [8, 4])
def test_binning_modes_agree_on_config0_scene(gpu, digit_bits):
    """Presort and 64-bit binning give identical keys, payloads, ranges and
    images on the 10k-surfel 256x256 benchmark scene, equal to the mirror."""
    import splattar_amd as sa
    from splattar_amd import scenes
    sc = scenes.config_scene(0)
    sp = scenes.front_view(256, 256)
    cam = sa.Camera(sp.center.astype(np.float32), sp.quat.astype(np.float32), float(np.float32(sp.focal)), 0)
    outs = {}
    for mode in (0, 1):
        announce(f"config0 binning={mode} digit={digit_bits}")
        r = sa.Rasterizer(gpu, sc.n, 256, 256, sc.n * 64)
        r.upload_scene(sc.positions, sc.scales, sc.quats, sc.opacities, sc.sh)
        r.set_binning(mode)
        r.set_digit_bits(digit_bits)
        r.forward(cam, sc.n)
        outs[mode] = (r.debug(), r.out_color.read(), r.out_ncontrib.read())
        r.close()
    (d0, c0, n0), (d1, c1, n1) = outs[0], outs[1]
    for k in ("sorted_keys", "sorted_payloads", "ranges", "rects", "depth_keys"):
        assert np.array_equal(d0[k], d1[k]), k
    assert np.array_equal(c0, c1) and np.array_equal(n0, n1)
    rot = oracle.quat_to_mat(cam.quat.astype(np.float64))
    c32 = oracle.cam32(rot, cam.center, cam.focal, 256, 256)
    pos4 = np.concatenate([sc.positions, np.ones((sc.n, 1), np.float32)], 1)
    rects, keys, _ = oracle.prep_f32(pos4, sc.scales, sc.quats, c32)
    k, p, e = oracle.emit_keys(rects, keys, c32.tiles_x)
    k, p = oracle.stable_sort(k, p)
    assert d1["entries"] == e and np.array_equal(d1["sorted_keys"], k)
    assert np.array_equal(d1["sorted_payloads"], p)
    assert np.array_equal(d1["ranges"], oracle.tile_ranges(k, c32.tiles_x * c32.tiles_y))


def _tight_scene(kind):
    """(label, n, size, positions, scales, quats, opacities, sh, deg, camera spec)."""
    from splattar_amd import scenes
    if kind == "config0":
        sc, sp = scenes.config_scene(0), scenes.front_view(256, 256)
    elif kind == "heavy_tail":
        # The 1M-surfel benchmark distribution (log-scales N(-5, 1), SH3),
        # thinned to 60k surfels on a 320x240 sphere view.
        sc = scenes.config_scene(3, 60_000)
        sp = scenes.sphere_views(1, scenes.CONFIG_SEED + 3, [(320, 240)])[0]
    else:
        scene, cam = _dense_scene(4000, 96, 2, 5)
        from splattar_amd.scenes import SceneArrays, ViewSpec
        f = np.float32
        sc = SceneArrays(scene.pos.astype(f), scene.scales.astype(f), scene.quat.astype(f),
                         scene.opacity.astype(f), scene.sh.reshape(scene.n, -1).astype(f), 2)
        sp = ViewSpec(96, 96, cam.center, cam.quat, cam.focal)
    return sc, sp


def _run_binning(gpu, sc, sp, mode, dl, cull=True):
    import splattar_amd as sa
    cam = sa.Camera(np.asarray(sp.center, np.float32), np.asarray(sp.quat, np.float32),
                    float(np.float32(sp.focal)), sc.sh_degree)
    r = sa.Rasterizer(gpu, sc.n, sp.width, sp.height, sc.n * 64)
    r.upload_scene(sc.positions, sc.scales, sc.quats, sc.opacities, sc.sh)
    r.set_binning(mode)
    r.set_culling(cull)
    r.dl_dcolor.write(dl)
    r.forward(cam, sc.n)
    out = dict(dbg=r.debug(), color=r.out_color.read(), aux=r.out_aux.read(),
               normal=r.out_normal.read(), ncontrib=r.out_ncontrib.read())
    r.backward(sc.n, flags=0)
    out.update(pos=r.grad_pos.read()[:, :3], scales=r.grad_scales.read(), quat=r.grad_quat.read(),
               opacity=r.grad_opacity.read(), sh=r.grad_sh.read())
    r.close()
    return out


@pytest.mark.parametrize("kind", ["config0", "heavy_tail", "dense"])
def test_tight_binning_equals_presort(gpu, kind):
    """Tight binning drops only entries that reach no pixel: its tile lists
    are order-preserving sub-lists of the reference's, its rects lie inside
    the reference rects, the images are bit-identical and the gradients
    equal up to the order of float atomics."""
    import splattar_amd as sa
    sc, sp = _tight_scene(kind)
    rng = np.random.default_rng(17)
    dl = np.zeros((sp.height, sp.width, 4), np.float32)
    dl[..., :3] = rng.uniform(-1, 1, (sp.height, sp.width, 3))
    announce(f"tight vs presort {kind} n={sc.n} {sp.width}x{sp.height}")
    ref = _run_binning(gpu, sc, sp, sa._lib.GS_BINNING_PRESORT, dl)
    got = _run_binning(gpu, sc, sp, sa._lib.GS_BINNING_TIGHT, dl)
    dr, dt = ref["dbg"], got["dbg"]
    print(f"{kind}: entries presort {dr['entries']} tight {dt['entries']}")
    assert 0 < dt["entries"] < dr["entries"]
    # Rects: inside the reference rect (or empty).
    rr, rt = dr["rects"][:sc.n].astype(np.int64), dt["rects"][:sc.n].astype(np.int64)
    live = (rt[:, 0] <= rt[:, 2]) & (rt[:, 1] <= rt[:, 3])
    assert np.all(rt[live, 0] >= rr[live, 0]) and np.all(rt[live, 2] <= rr[live, 2])
    assert np.all(rt[live, 1] >= rr[live, 1]) and np.all(rt[live, 3] <= rr[live, 3])
    # Per tile, the tight list is a sub-list of the reference list in order.
    def tile_of(d):
        ln = (d["ranges"][:, 1] - d["ranges"][:, 0]).astype(np.int64)
        return np.repeat(np.arange(len(ln), dtype=np.int64), ln)
    tr, tt = tile_of(dr), tile_of(dt)
    kr = tr * sc.n + dr["sorted_payloads"].astype(np.int64)
    kt = tt * sc.n + dt["sorted_payloads"].astype(np.int64)
    order = np.argsort(kr, kind="stable")
    at = np.searchsorted(kr[order], kt)
    assert np.all(at < len(kr)) and np.array_equal(kr[order][at], kt)
    pos_in_ref = order[at]
    same_tile = tt[1:] == tt[:-1]
    assert np.all(np.diff(pos_in_ref)[same_tile] > 0)
    # Images bit-identical (ncontrib counts list positions and may differ).
    for k in ("color", "aux", "normal"):
        assert np.array_equal(got[k], ref[k]), k
    for k in ("pos", "scales", "quat", "opacity", "sh"):
        g, c = got[k].astype(np.float64), ref[k].astype(np.float64)
        nrm = np.linalg.norm(g - c) / max(np.linalg.norm(c), 1e-30)
        print(f"  {k}: normwise rel {nrm:.2e}")
        assert nrm < 1e-5, k


@pytest.mark.parametrize("kind", ["config0", "heavy_tail", "dense"])
def test_footprint_culling_is_exact(gpu, kind):
    """The per-quadrant footprint culling (box, screen disk, ellipse) skips
    only records that reach no pixel of a wave's 8x8 block: images and
    ncontrib are bit-identical with it off, gradients equal up to the order
    of float atomics."""
    import splattar_amd as sa
    sc, sp = _tight_scene(kind)
    rng = np.random.default_rng(23)
    dl = np.zeros((sp.height, sp.width, 4), np.float32)
    dl[..., :3] = rng.uniform(-1, 1, (sp.height, sp.width, 3))
    announce(f"culling on/off {kind} n={sc.n} {sp.width}x{sp.height}")
    ref = _run_binning(gpu, sc, sp, sa._lib.GS_BINNING_TIGHT, dl, cull=False)
    got = _run_binning(gpu, sc, sp, sa._lib.GS_BINNING_TIGHT, dl, cull=True)
    for k in ("color", "aux", "normal", "ncontrib"):
        assert np.array_equal(got[k], ref[k]), k
    for k in ("pos", "scales", "quat", "opacity", "sh"):
        g, c = got[k].astype(np.float64), ref[k].astype(np.float64)
        nrm = np.linalg.norm(g - c) / max(np.linalg.norm(c), 1e-30)
        print(f"  {k}: normwise rel {nrm:.2e}")
        assert nrm < 1e-5, k


def _weights(seed, size, aux):
    # raster_parity.rs:129-161 (rng = seed ^ 0xfeed).
    rng = XorShift64(seed ^ 0xFEED)
    P = size * size
    w = np.array([[rng.uniform(-1, 1) for _ in range(3)] for _ in range(P)])
    wd = np.array([rng.uniform(-0.3, 0.3) if aux else 0.0 for _ in range(P)])
    wn = np.array([[rng.uniform(-0.5, 0.5) for _ in range(3)] if aux else [0.0, 0.0, 0.0]
                   for _ in range(P)])
    return w, wd, wn


_CLASSES = ("pos", "scales", "quat", "opacity", "sh")


def _grad_case(gpu, seed, n, deg, size, aux, flags=0):
    scene, cam = make_scene(seed, n, deg, size, flavor="parity")
    s32, c32 = scene_from_f32(scene), camera_from_f32(cam)
    w, wd, wn = _weights(seed, size, aux)
    w32, wd32, wn32 = w.astype(np.float32), wd.astype(np.float32), wn.astype(np.float32)
    lam = 0.35 if aux else 0.0
    ref = oracle.gradients(s32, c32, w32.astype(np.float64), wd32.astype(np.float64),
                           wn32.astype(np.float64), lam)
    announce(f"backward seed={seed} n={n} deg={deg} size={size} aux={aux} flags={flags}")
    r = _rasterizer(gpu, scene, cam)
    dl = np.zeros((size, size, 4), np.float32)
    dl[..., :3] = w32.reshape(size, size, 3)
    dl[..., 3] = wd32.reshape(size, size)
    dn = np.zeros((size, size, 4), np.float32)
    dn[..., :3] = wn32.reshape(size, size, 3)
    r.dl_dcolor.write(dl)
    r.dl_dnormal.write(dn)
    r.set_lambda_dist(lam)
    r.set_trace(0, 0, size, size, 256)
    r.forward(_camera(cam, deg), n)
    lists, counts = r.trace()
    r.backward(n, flags=flags)
    # The same FP64 chain along the GPU's decisions, and its conditioning.
    import multiview as mv
    args = (w32.astype(np.float64).reshape(size, size, 3), wd32.astype(np.float64), wn32.astype(np.float64), lam)
    img_r, _, rep = oracle.replay(s32, c32, lists, counts, *args)
    ill, _ = mv.replay_conditioning(s32, c32, lists, counts, args[0], rep, img_r, draws=3, dl_ddepth=args[1],
                                    dl_dnormal=args[2], lambda_dist=lam)
    ref.replay, ref.ill = rep, ill
    ref.flipped = _flipped_pixels(s32, c32, lists, counts)
    ncoef = (deg + 1) ** 2
    got = dict(pos=r.grad_pos.read()[:, :3], scales=r.grad_scales.read(), quat=r.grad_quat.read(),
               opacity=r.grad_opacity.read(), sh=r.grad_sh.read()[:, :ncoef * 3].reshape(n, ncoef, 3),
               cam=r.grad_cam.read())
    r.close()
    return scene, cam, ref, got


def _check_class(label, gpu_v, cpu_v, rel_tol, norm_tol=1e-4):
    g = np.asarray(gpu_v, np.float64).ravel()
    c = np.asarray(cpu_v, np.float64).ravel()
    err = np.abs(g - c)
    denom = np.maximum(np.abs(g), np.abs(c))
    bad = err > 1e-5 + rel_tol * denom
    assert not bad.any(), f"{label}: {bad.sum()} elements off, worst {err.max():.3e}"
    cn = np.linalg.norm(c)
    nrm = np.linalg.norm(g - c) / cn if cn > 0 else np.linalg.norm(g)
    print(f"  {label}: normwise rel {nrm:.2e}")
    if nrm > 0.3 * norm_tol:
        for i in np.argsort(-err)[:4]:
            print(f"    [{i}] gpu {g[i]: .6e} fp64 {c[i]: .6e} err {err[i]:.2e} (|c|max {np.abs(c).max():.3e})")
    assert nrm <= norm_tol, f"{label}: normwise relative error {nrm:.3e}"
    return nrm


def _check_all(scene, cam, ref, got, rel_tol, norm_tol=1e-4):
    """Against FP64 replaying the GPU's decisions (ref.replay): per class
    normwise <= norm_tol and per element <= 1e-5 + rel_tol max(|a|, |b|)
    (raster_parity.rs:101-113) over the surfels whose replayed gradient is
    well conditioned (ref.ill excluded, counted); camera gradients (dL/dC,
    dL/df, dL/dq through the host quaternion chain) normwise <= norm_tol.
    The faithful FP64 gradients and the pixels whose decisions differ are
    reported."""
    rep, ill = ref.replay, ref.ill
    print(f"  decision flips {int(ref.flipped.sum())}; ill-conditioned surfels {int(ill.sum())} of {ill.size}")
    assert ill.mean() <= 0.2
    for k in _CLASSES:
        n = ill.size
        fa = np.linalg.norm(np.asarray(got[k], np.float64).reshape(n, -1) - getattr(ref, k).reshape(n, -1))
        print(f"  {k}: against faithful FP64 {fa / max(np.linalg.norm(getattr(ref, k)), 1e-30):.2e}")
        _check_class(k, np.asarray(got[k]).reshape(n, -1)[~ill], getattr(rep, k).reshape(n, -1)[~ill], rel_tol,
                     norm_tol)
    gc = got["cam"]
    _check_class("cam_center", gc[9:12], rep.cam_center, rel_tol, norm_tol)
    _check_class("focal", gc[12:13], [rep.focal], rel_tol, norm_tol)
    dl_dr = gc[:9].reshape(3, 3).T.astype(np.float64)  # column-major -> dl_dr[j, k]
    q = oracle.quat_grad(cam.quat.astype(np.float32).astype(np.float64), dl_dr)
    _check_class("cam_quat", q, rep.cam_quat, rel_tol, norm_tol)


@pytest.mark.parametrize("seed,n,deg,size", [(111, 15, 0, 48), (222, 15, 1, 48), (333, 10, 3, 40)])
def test_backward_matches_fp64_oracle(gpu, seed, n, deg, size):
    scene, cam, ref, got = _grad_case(gpu, seed, n, deg, size, aux=False)
    _check_all(scene, cam, ref, got, 2e-3)


@pytest.mark.parametrize("seed,n,deg,size", [(444, 12, 0, 40), (555, 12, 1, 40)])
def test_backward_aux_losses_match_fp64_oracle(gpu, seed, n, deg, size):
    scene, cam, ref, got = _grad_case(gpu, seed, n, deg, size, aux=True)
    _check_all(scene, cam, ref, got, 2e-3)


def test_color_only_kernel_matches_full_kernel(gpu):
    """The colour-only backward (one wave per 16-px tile, four pixels per
    lane) against the full kernel (two waves, depth / normal terms zero):
    the same sums in another order, so per class normwise <= 1e-6 and every
    element within 1e-6 of the class's largest magnitude."""
    import splattar_amd as sa
    _, _, ref, full = _grad_case(gpu, 222, 15, 1, 48, aux=False)
    _, _, _, fast = _grad_case(gpu, 222, 15, 1, 48, aux=False, flags=sa._lib.GS_BWD_COLOR_ONLY)
    for k in ("pos", "scales", "quat", "opacity", "sh"):
        f, g = np.asarray(full[k], np.float64), np.asarray(fast[k], np.float64)
        nrm = float(np.linalg.norm(g - f) / np.linalg.norm(f))
        print(f"  {k}: colour-only vs full normwise {nrm:.2e}")
        assert nrm <= 1e-6, k
        assert np.abs(g - f).max() <= 1e-6 * np.abs(f).max(), k


def test_backward_accumulates(gpu):
    import splattar_amd as sa
    scene, cam = make_scene(111, 15, 0, 48, flavor="parity")
    w, _, _ = _weights(111, 48, False)
    r = _rasterizer(gpu, scene, cam)
    dl = np.zeros((48, 48, 4), np.float32)
    dl[..., :3] = w.reshape(48, 48, 3)
    r.dl_dcolor.write(dl)
    announce("backward accumulate")
    r.forward(_camera(cam, 0), 15)
    r.backward(15, flags=sa._lib.GS_BWD_COLOR_ONLY)
    one = r.grad_pos.read().copy()
    r.backward(15, flags=sa._lib.GS_BWD_COLOR_ONLY | sa._lib.GS_BWD_ACCUMULATE)
    two = r.grad_pos.read()
    np.testing.assert_allclose(two, 2 * one, rtol=1e-5, atol=1e-8)
    r.close()


def _dense_scene(n, size, deg, seed):
    """Many small surfels: tile lists longer than one 256-record LDS batch,
    early termination, off-screen and behind-camera surfels."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1, 1, (n, 3))
    pos[: n // 50, 2] = -3.5  # behind the camera at z = -3 (culled)
    scales = np.exp(rng.normal(-3.3, 0.5, (n, 2)))
    q = rng.normal(size=(n, 4))
    op = 1 / (1 + np.exp(-rng.normal(0.5, 1.0, n)))
    ncoef = (deg + 1) ** 2
    sh = rng.normal(0, 0.3, (n, ncoef, 3))
    scene = oracle.Scene(pos, scales, q, op, sh, deg)
    from splattar_amd.scenes import look_at, focal_for
    eye = np.array([0.0, 0.0, -3.0])
    cam = oracle.Camera(eye, look_at(eye), focal_for(size), size, size)
    return scene, cam


def test_dense_scene_forward_backward(gpu):
    n, size, deg = 4000, 96, 2
    scene, cam = _dense_scene(n, size, deg, 5)
    s32, c32 = scene_from_f32(scene), camera_from_f32(cam)
    ref = oracle.render(s32, c32)
    rng = np.random.default_rng(9)
    w = rng.uniform(-1, 1, (size * size, 3)).astype(np.float32)
    refg = oracle.gradients(s32, c32, w.astype(np.float64))
    announce(f"dense forward/backward n={n} size={size}")
    r = _rasterizer(gpu, scene, cam)
    dl = np.zeros((size, size, 4), np.float32)
    dl[..., :3] = w.reshape(size, size, 3)
    r.dl_dcolor.write(dl)
    r.forward(_camera(cam, deg), n)
    color = r.out_color.read()
    ranges = r.debug()["ranges"]
    assert (ranges[:, 1] - ranges[:, 0]).max() > 256  # multi-batch tiles exercised
    mx, mean, over = _image_stats(color[..., :3], ref["color"])
    print(f"dense: color max {mx:.2e} mean {mean:.2e} px>1e-5 {over}")
    assert mean < 2e-5 and over <= color.size // 200
    r.backward(n, flags=0)
    gpos = r.grad_pos.read()[:, :3]
    gop = r.grad_opacity.read()
    for label, g, c in (("pos", gpos, refg.pos), ("opacity", gop, refg.opacity)):
        nrm = np.linalg.norm(g - c) / np.linalg.norm(c)
        print(f"  dense {label}: normwise rel {nrm:.2e}")
        assert nrm < 2e-3
    r.close()


def test_overflow_reported(gpu):
    import splattar_amd as sa
    scene, cam = make_scene(404, 50, 1, 128, flavor="parity")
    r = _rasterizer(gpu, scene, cam, max_entries=16)
    announce("forward overflow")
    r.forward(_camera(cam, 1), 50)
    with pytest.raises(sa.GsError) as ei:
        r.entry_count()
    assert ei.value.status == 5
    r.close()


def test_invalid_arguments_rejected(gpu):
    import splattar_amd as sa
    scene, cam = make_scene(101, 20, 0, 64, flavor="parity")
    r = _rasterizer(gpu, scene, cam)
    with pytest.raises(sa.GsError):
        r.forward(_camera(cam, 0), 21)          # beyond capacity
    bad = _camera(cam, 0)
    bad.focal = 0.0
    with pytest.raises(sa.GsError):
        r.forward(bad, 20)
    with pytest.raises(sa.GsError):
        r.backward(20)                           # backward before any forward
    r.close()


def test_batched_views_equal_single_view_renders(gpu):
    """gs_raster_forward_batch / backward_batch: three views of different
    sizes and surfel prefixes in one batch give bit-identical images to three
    single-view renders, and the batch backward equals the single-view
    backwards accumulated over the views (float-atomic order aside), with
    per-view camera gradients."""
    import splattar_amd as sa
    from splattar_amd import scenes
    deg, n = 1, 3000
    sc = scenes.make_scene(n, deg, -3.4, 0.4, 21)
    specs = scenes.sphere_views(3, 8, [(96, 64), (64, 48), (80, 40)])
    counts = [n, 2200, 2700]
    W, H = 96, 64  # the batch rasterizer's per-view maximum
    rng = np.random.default_rng(3)
    dls = [rng.normal(0, 1e-3, (sp.height * sp.width, 4)).astype(np.float32) for sp in specs]
    for d in dls:
        d[:, 3] = 0.0
    cams = [sa.Camera(sp.center, sp.quat, sp.focal, deg) for sp in specs]
    flags = sa._lib.GS_BWD_COLOR_ONLY
    # Single-view renders (a rasterizer of each view's size) and backwards.
    imgs, ncs, cam_g = [], [], []
    want = None
    for k, (sp, cam, m) in enumerate(zip(specs, cams, counts)):
        announce(f"batched views: single view {k}")
        single = sa.Rasterizer(gpu, n, sp.width, sp.height, n * 64)
        single.upload_scene(sc.positions, sc.scales, sc.quats, sc.opacities, sc.sh)
        single.forward(cam, m)
        imgs.append(single.out_color.read().reshape(-1, 4).copy())
        ncs.append(single.out_ncontrib.read().reshape(-1).copy())
        single.dl_dcolor.write(dls[k].reshape(sp.height, sp.width, 4))
        single.backward(m, flags=flags)
        g = dict(pos=single.grad_pos.read().copy(), sh=single.grad_sh.read().copy(),
                 quat=single.grad_quat.read().copy(), opacity=single.grad_opacity.read().copy())
        want = g if want is None else {c: want[c] + g[c] for c in g}
        cam_g.append(single.grad_cam.read().copy())
        single.close()
    # One batch.
    b = sa.Rasterizer(gpu, n, W, H, n * 64, max_views=3)
    b.upload_scene(sc.positions, sc.scales, sc.quats, sc.opacities, sc.sh)
    announce("batched views: batch forward")
    b.forward_batch([(cam, sp.width, sp.height, m) for sp, cam, m in zip(specs, cams, counts)])
    out = b.out_color.read()
    nc = b.out_ncontrib.read()
    off = 0
    for k, sp in enumerate(specs):
        P = sp.width * sp.height
        assert np.array_equal(out[off:off + P], imgs[k]), f"view {k} image"
        assert np.array_equal(nc[off:off + P], ncs[k]), f"view {k} ncontrib"
        off += P
    b.dl_dcolor.write(np.concatenate(dls + [np.zeros((3 * H * W - off, 4), np.float32)]))
    announce("batched views: batch backward")
    b.backward_batch(flags=flags)
    got = dict(pos=b.grad_pos.read(), sh=b.grad_sh.read(), quat=b.grad_quat.read(),
               opacity=b.grad_opacity.read())
    for k in want:
        nrm = float(np.linalg.norm(got[k] - want[k]) / max(np.linalg.norm(want[k]), 1e-30))
        print(f"  {k}: batch vs accumulated single views normwise {nrm:.2e}")
        assert nrm < 1e-5, k
    # Per-view camera gradients: the batch keeps view j's 16 slots apart.
    gc = b.grad_cam.read()
    for k in range(3):
        scale = max(float(np.abs(cam_g[k]).max()), 1e-12)
        assert float(np.abs(gc[k] - cam_g[k]).max()) <= 1e-4 * scale, f"view {k} camera gradient"
    b.close()


@pytest.mark.parametrize("case", ["behind_camera", "zero_surfels"])
def test_empty_frame_renders_background_and_zero_gradients(gpu, case):
    """A frame with no visible surfel (all behind the camera, or a forward
    over 0 surfels): zero entries, every pixel black with T = 1 and no
    contributor (the oracle's empty render, forward.rs:207-246), and a
    backward that leaves every gradient exactly 0."""
    n, size, deg = 500, 72, 1
    scene, cam = _dense_scene(n, size, deg, 21)
    scene.pos[:, 2] = -3.5  # behind the eye at z = -3
    s32, c32 = scene_from_f32(scene), camera_from_f32(cam)
    ref = oracle.render(s32, c32)
    assert not ref["color"].any()
    m = n if case == "behind_camera" else 0
    announce(f"empty frame: {case}")
    r = _rasterizer(gpu, scene, cam)
    dl = np.zeros((size, size, 4), np.float32)
    dl[..., :3] = np.random.default_rng(3).uniform(-1, 1, (size, size, 3))
    r.dl_dcolor.write(dl)
    r.grad_cam.write(np.full(16, 7.0, np.float32))  # must be overwritten
    r.forward(_camera(cam, deg), m)
    assert r.entry_count() == 0
    assert not r.out_color.read().any()
    assert np.all(r.out_aux.read()[..., 0] == 1.0)
    assert not r.out_ncontrib.read().any()
    r.backward(m, flags=0)
    for g in (r.grad_pos, r.grad_scales, r.grad_quat, r.grad_opacity, r.grad_sh, r.grad_cam):
        assert not g.read().any()
    r.close()
