"""Finite differences at the batched-differentiation configuration: 8 views
of uneven size (640x360 .. 1920x1080) on the 100k-surfel SH3 scene, loss
sum over views of 0.8 L1 + 0.2 (1 - SSIM), 64 random scalar parameters,
central differences at eps = 1e-3 (and eps / 2) in the raw parameter
(position, log-scale, quaternion, opacity logit, SH coefficient), the GPU
rendering in fp32 and the loss of its images accumulated in FP64.

A surfel only reaches the tiles of its rect, so the loss changes only inside
the union of its rects over the perturbed evaluations (the fp32 mirror gives
them exactly); the D-SSIM change reaches 5 px further. Each evaluation is a
GPU forward traced over that window, and the loss difference is the FP64
L1 + D-SSIM of the window plus a 10-px margin (pixels whose SSIM window
misses the changed region cancel between the two sides).

fp32 at eps = 1e-3 crosses discrete decisions (the alpha >= 1/255 cut-off
ring moves by ~0.3 px for a position step), and a position step of 1e-3 is
~15 % of a surfel's size (scales exp(N(-5, 0.5))), far from the asymptotic
regime of a central difference. So the checks are:
1. every parameter: the GPU's fp32 differences (eps, eps / 2) against FP64
   replays of the GPU's own decisions at the same perturbed points, within
   twice the first-order effect of the images' measured f32 errors
   (sum |dL/dI| |I_gpu - I_fp64|, over both sides);
2. every parameter: the GPU's analytic gradient against the FP64 central
   difference along the base decisions at the reference gradcheck's step
   h = 1e-6 (gradcheck.rs:118), to its tolerance 1e-3 max(|a|, |fd|)
   (gradcheck.rs:124-136) plus an absolute floor of 1e-4 of the RMS analytic
   gradient over the 64 parameters (the loss is a mean over pixels, so
   gradients are ~1e-6 and gradcheck's fixed 1e-7 would be loose); the
   FP64 analytic gradient itself to 1e-5. Not compared: parameters of a
   surfel compositing at a pixel whose render is within 1e-5 of the target
   (an L1 kink, no two-sided derivative), and, for the GPU's value only,
   those whose FP64 gradient moves by more than 1e-4 of itself under one-ulp
   input perturbations (counted);
3. the non-position parameters whose perturbations take no other decision
   in any view (counted): the GPU's analytic gradient against the
   Richardson extrapolation of its own fp32 differences, within the f32
   bound of 1. plus 1e-3 relative.
"""

import numpy as np
import pytest

import oracle
from conftest import announce

pytestmark = pytest.mark.gpu

LAM, DSSIM_SCALE = 0.2, 1.0
EPS = 1e-3
H64 = 1e-6
CAP = 192
CLASSES = ("pos", "scales", "quat", "opacity", "sh")
WIDTH = {"pos": 3, "scales": 2, "quat": 4, "opacity": 1, "sh": 48}


def _raw(cls, x):
    """Raw (optimizer) value of an activated scalar."""
    x = np.float64(x)
    if cls == "scales":
        return np.log(x)
    if cls == "opacity":
        return np.log(x / (1.0 - x))
    return x


def _act(cls, r):
    if cls == "scales":
        return np.exp(r)
    if cls == "opacity":
        return 1.0 / (1.0 + np.exp(-r))
    return r


class _View:
    def __init__(self, gpu, sc, sp, target):
        import splattar_amd as sa
        self.sp, self.W, self.H = sp, sp.width, sp.height
        self.target = target[..., :3].astype(np.float64)
        self.cam = sa.Camera(np.asarray(sp.center, np.float32), np.asarray(sp.quat, np.float32),
                             float(np.float32(sp.focal)), sc.sh_degree)
        self.cam64 = oracle.Camera(np.asarray(sp.center, np.float32).astype(np.float64),
                                   np.asarray(sp.quat, np.float32).astype(np.float64),
                                   float(np.float32(sp.focal)), sp.width, sp.height)
        rot = oracle.quat_to_mat(np.asarray(sp.quat, np.float32).astype(np.float64))
        self.c32 = oracle.cam32(rot, np.asarray(sp.center, np.float32), float(np.float32(sp.focal)),
                                sp.width, sp.height)
        self.r = sa.Rasterizer(gpu, sc.n, sp.width, sp.height, sc.n * 64)
        self.r.upload_scene(sc.positions, sc.scales, sc.quats, sc.opacities, sc.sh)
        self.r.set_binning(sa._lib.GS_BINNING_TIGHT)

    def rect(self, pos, scales, quat):
        """The fp32 mirror's tile rect of one surfel (x0, y0, x1, y1 inclusive),
        None when it reaches no tile."""
        pos4 = np.concatenate([np.asarray(pos, np.float32).reshape(1, 3), np.ones((1, 1), np.float32)], 1)
        rects, _, counts = oracle.prep_f32(pos4, np.asarray(scales, np.float32).reshape(1, 2),
                                           np.asarray(quat, np.float32).reshape(1, 4), self.c32)
        x0, y0, x1, y1 = (int(v) for v in rects[0])
        return None if counts[0] == 0 or x1 < x0 or y1 < y0 else (x0, y0, x1, y1)

    def forward(self, n, window):
        """GPU forward traced over window (x0, y0, w, h): (lists, counts, crop
        image rows (y0c..y1c, all columns) as f64, crop box)."""
        import splattar_amd as sa
        x0, y0, w, h = window
        self.r.set_trace(x0, y0, w, h, CAP)
        self.r.forward(self.cam, n)
        lists, counts = self.r.trace()
        assert counts.max() <= CAP
        cx0, cy0 = max(x0 - 10, 0), max(y0 - 10, 0)
        cx1, cy1 = min(x0 + w + 10, self.W), min(y0 + h + 10, self.H)
        rows = sa.DeviceBuffer(self.r.ctx, self.r.out_color.ptr + cy0 * self.W * 16, np.float32,
                               (cy1 - cy0, self.W, 4)).read()
        img = rows[:, cx0:cx1, :3].astype(np.float64)
        return lists, counts, img, (cx0, cy0, cx1, cy1)

    def loss(self, img, box):
        """The view's loss change contribution of a crop: its FP64 L1 + D-SSIM
        (mean over the crop) scaled to the view's pixel count."""
        cx0, cy0, cx1, cy1 = box
        t = self.target[cy0:cy1, cx0:cx1]
        val = oracle.l1_ssim(img, t, LAM, DSSIM_SCALE, grad=False)[0]
        return val * (cx1 - cx0) * (cy1 - cy0) / (self.W * self.H)


def _sub_replay(s64, cam, lists, counts, cls, i, c, v, dl=None, rng=None):
    """oracle.replay of the lists over the surfels they name only (indices
    remapped), with parameter (cls, i, c) of surfel i set to v. With dl
    (window, 3) also returns the parameter's FP64 gradient (activated); rng
    scales every input (the surfels and dl) by 1 +- 2^-24 first."""
    live = np.arange(lists.shape[1])[None, :] < counts[:, None]
    ids = lists & 0x3FFFFFFF
    used = np.unique(np.concatenate([ids[live], np.array([i], np.uint32)]))
    remap = np.searchsorted(used, ids).astype(np.uint32)
    lists2 = np.where(live, (lists & 0xC0000000) | remap, 0).astype(np.uint32)
    sub = oracle.Scene(s64.pos[used].copy(), s64.scales[used].copy(), s64.quat[used].copy(),
                       s64.opacity[used].copy(), s64.sh[used].copy(), s64.sh_degree)
    j = int(np.searchsorted(used, i))
    arr = {"pos": sub.pos, "scales": sub.scales, "quat": sub.quat, "opacity": sub.opacity[:, None],
           "sh": sub.sh.reshape(sub.n, -1)}
    if cls == "opacity":
        sub.opacity[j] = v
    else:
        arr[cls][j, c] = v
    if rng is not None:
        for a in (sub.pos, sub.scales, sub.quat, sub.opacity, sub.sh):
            a *= 1.0 + rng.choice([-1.0, 1.0], a.shape) * 2.0 ** -24
        dl = dl * (1.0 + rng.choice([-1.0, 1.0], dl.shape) * 2.0 ** -24)
    img, _, g = oracle.replay(sub, cam, lists2, counts, dl)
    if dl is None:
        return img
    ga = {"pos": g.pos, "scales": g.scales, "quat": g.quat, "opacity": g.opacity[:, None],
          "sh": g.sh.reshape(sub.n, -1)}[cls]
    return img, float(ga[j, c])


def _same_decisions(la, ca, lb, cb):
    if not np.array_equal(ca, cb):
        return False
    live = np.arange(la.shape[1])[None, :] < ca[:, None]
    return np.array_equal(np.where(live, la, 0), np.where(live, lb, 0))


def test_config2_finite_differences(gpu):
    import multiview as mv
    import splattar_amd as sa
    from splattar_amd import scenes
    sc = scenes.config_scene(2)
    n = sc.n
    tgt = scenes.jittered(sc, 0.01, (scenes.CONFIG_SEED + 2) ^ 1)
    sizes = scenes.uneven_sizes(8, scenes.CONFIG_SEED + 2)
    specs = scenes.sphere_views(8, scenes.CONFIG_SEED + 2, sizes)
    announce(f"FD: targets {sizes}")
    targets = []
    for sp in specs:
        r = sa.Rasterizer(gpu, n, sp.width, sp.height, n * 64)
        r.upload_scene(tgt.positions, tgt.scales, tgt.quats, tgt.opacities, tgt.sh)
        r.forward(sa.Camera(np.asarray(sp.center, np.float32), np.asarray(sp.quat, np.float32),
                            float(np.float32(sp.focal)), 3), n)
        targets.append(r.out_color.read().copy())
        r.close()
    views = [_View(gpu, sc, sp, t) for sp, t in zip(specs, targets)]
    announce("FD: analytic gradient (op API, per view, summed)")
    grad = {k: np.zeros((n, WIDTH[k])) for k in CLASSES}
    for v in views:
        loss = sa.SsimLoss(gpu, v.W, v.H, LAM, v.r.out_color, v.r.dl_dcolor, DSSIM_SCALE)
        t4 = np.zeros((v.H, v.W, 4), np.float32)
        t4[..., :3] = v.target
        loss.target.write(t4)
        v.r.forward(v.cam, n)
        loss.encode()
        v.r.backward(n)
        grad["pos"] += v.r.grad_pos.read()[:, :3]
        grad["scales"] += v.r.grad_scales.read()
        grad["quat"] += v.r.grad_quat.read()
        grad["opacity"] += v.r.grad_opacity.read()[:, None]
        grad["sh"] += v.r.grad_sh.read()
        loss.close()
    # 64 random scalars of surfels that reach some view's image.
    rng = np.random.default_rng(0xFD64)
    live = np.nonzero(grad["opacity"][:, 0] != 0.0)[0]
    params = [(CLASSES[rng.integers(5)], int(rng.choice(live))) for _ in range(64)]
    params = [(cls, i, int(rng.integers(WIDTH[cls]))) for cls, i in params]
    act0 = {"pos": sc.positions, "scales": sc.scales, "quat": sc.quats, "opacity": sc.opacities[:, None],
            "sh": sc.sh}
    s64 = mv.scene64(sc)
    rows = []
    for k, (cls, i, c) in enumerate(params):
        x0 = np.float32(act0[cls][i, c])
        raw0 = _raw(cls, x0)
        # Activated values of the four fp32 evaluations (as the GPU holds them).
        steps = [+EPS, -EPS, +EPS / 2, -EPS / 2]
        xs = [np.float32(_act(cls, raw0 + s)) for s in steps]
        hs = [_raw(cls, xs[0]) - _raw(cls, xs[1]), _raw(cls, xs[2]) - _raw(cls, xs[3])]
        # The FP64 central difference along the base decisions (gradcheck.rs:118).
        xh = [float(_act(cls, raw0 + H64)), float(_act(cls, raw0 - H64))]
        a = float(grad[cls][i, c]) * {"scales": float(x0), "opacity": float(x0) * (1.0 - float(x0))}.get(cls, 1.0)

        def geom(x):
            p, s, q = sc.positions[i].copy(), sc.scales[i].copy(), sc.quats[i].copy()
            if cls == "pos":
                p[c] = x
            elif cls == "scales":
                s[c] = x
            elif cls == "quat":
                q[c] = x
            return p, s, q

        d32 = np.zeros(4)    # +-L of the GPU images at the four fp32 points
        d64 = np.zeros(4)    # the same, FP64 replays of the GPU's decisions there
        bound = np.zeros(4)  # first-order effect of the images' f32 errors on the loss
        dh = np.zeros(2)     # FP64 along the base decisions at raw +- H64
        a64, a64p = 0.0, [0.0, 0.0]
        flips, kink = 0, False
        for v in views:
            rs = [v.rect(*geom(x)) for x in [x0] + xs]
            rs = [r for r in rs if r is not None]
            if not rs:
                continue
            tx0, ty0 = min(r[0] for r in rs), min(r[1] for r in rs)
            tx1, ty1 = max(r[2] for r in rs), max(r[3] for r in rs)
            wx0, wy0 = tx0 * 16, ty0 * 16
            window = (wx0, wy0, min((tx1 + 1) * 16, v.W) - wx0, min((ty1 + 1) * 16, v.H) - wy0)
            cam = v.cam64.copy()
            cam.crop = window
            # Base evaluation: its decisions and the crop around the window.
            bl, bc, bimg, box = v.forward(n, window)
            ox, oy = window[0] - box[0], window[1] - box[1]
            inner = (slice(oy, oy + window[3]), slice(ox, ox + window[2]))
            outside = np.ones(bimg.shape[:2], bool)
            outside[inner] = False
            base64 = bimg.copy()
            base64[inner] = _sub_replay(s64, cam, bl, bc, cls, i, c, float(x0))
            # dL/d(image) of the view's loss over the crop, at the base point.
            t = v.target[box[1]:box[3], box[0]:box[2]]
            g_img = oracle.l1_ssim(base64, t, LAM, DSSIM_SCALE)[1] * (base64.shape[0] * base64.shape[1]) / (
                v.W * v.H)
            # The FP64 analytic gradient along the base decisions and how far
            # it moves when every input is perturbed by one ulp (2 draws).
            dl_win = np.ascontiguousarray(g_img[inner])
            # An L1 kink: the surfel composites at a pixel whose render is
            # within 1e-5 of the target (a difference straddles it).
            live = np.arange(bl.shape[1])[None, :] < bc[:, None]
            mine = np.any(((bl & 0x3FFFFFFF) == i) & live, axis=1).reshape(window[3], window[2])
            near = np.any(np.abs(base64[inner] - t[inner]) < 1e-5, axis=2)
            kink |= bool(np.any(mine & near))
            a64 += _sub_replay(s64, cam, bl, bc, cls, i, c, float(x0), dl_win)[1]
            for d in range(2):
                a64p[d] += _sub_replay(s64, cam, bl, bc, cls, i, c, float(x0), dl_win,
                                       np.random.default_rng([k, d]))[1]
            errs = []
            buf = {"pos": v.r.positions, "scales": v.r.scales, "quat": v.r.quats, "opacity": v.r.opacities,
                   "sh": v.r.sh}[cls]
            stride = {"pos": 4, "scales": 2, "quat": 4, "opacity": 1, "sh": 48}[cls]
            for e, x in enumerate(xs):
                buf.write(np.array([x], np.float32), stride * i + c)
                lists, counts, img, _ = v.forward(n, window)
                buf.write(np.array([x0], np.float32), stride * i + c)
                # The surfel reaches nothing outside the window.
                assert np.array_equal(img[outside], bimg[outside])
                sgn = 1.0 if e % 2 == 0 else -1.0
                d32[e] += sgn * v.loss(img, box)
                comp = bimg.copy()
                comp[inner] = _sub_replay(s64, cam, lists, counts, cls, i, c, float(x))
                d64[e] += sgn * v.loss(comp, box)
                errs.append(img - comp)
                if e % 2 == 1:
                    # L(+) - L(-) differs between the two by sum dL/dI (err+ - err-)
                    # to first order.
                    bound[e] += float(np.sum(np.abs(g_img) * np.abs(errs[e - 1] - errs[e])))
                flips += 0 if _same_decisions(lists, counts, bl, bc) else 1
            for e, x in enumerate(xh):
                comp = bimg.copy()
                comp[inner] = _sub_replay(s64, cam, bl, bc, cls, i, c, x)
                dh[e] += (1.0 if e == 0 else -1.0) * v.loss(comp, box)
        chain = {"scales": float(x0), "opacity": float(x0) * (1.0 - float(x0))}.get(cls, 1.0)
        rows.append(dict(cls=cls, i=i, c=c, a=a, flips=flips, kink=kink, a64=a64 * chain,
                         sens=max(abs(p - a64) for p in a64p) * chain,
                         fd32=[(d32[0] + d32[1]) / hs[0], (d32[2] + d32[3]) / hs[1]],
                         fd64=[(d64[0] + d64[1]) / hs[0], (d64[2] + d64[3]) / hs[1]],
                         b=[bound[1] / hs[0], bound[3] / hs[1]],
                         fdh=(dh[0] + dh[1]) / (_raw(cls, xh[0]) - _raw(cls, xh[1]))))
        if k % 8 == 7:
            print(f"  {k + 1} of 64 parameters", flush=True)
    a = np.array([r["a"] for r in rows])
    floor = 1e-4 * float(np.sqrt(np.mean(a ** 2)))
    bad1, bad2, bad3, kept, n_ill, n_kink = [], [], [], 0, 0, 0
    for r in rows:
        # 1. fp32 differences against FP64 on the same decisions: within twice
        # the first-order effect of the images' measured f32 errors.
        ok1 = all(abs(r["fd32"][e] - r["fd64"][e]) <= 2.0 * r["b"][e] + 1e-6 * abs(r["fd64"][e]) for e in (0, 1))
        # 2. the FP64 analytic gradient along the base decisions against the
        # FP64 central difference (pins the replay), and the GPU's analytic
        # gradient against it to gradcheck's tolerance with the floor above,
        # unless one-ulp input perturbations move the FP64 value by more
        # than 1e-4 of itself (no f32 evaluation determines it better).
        # Parameters at an L1 kink have no two-sided derivative there.
        ill = r["sens"] > 1e-4 * abs(r["a64"])
        ok2 = r["kink"] or abs(r["a64"] - r["fdh"]) <= 1e-5 * abs(r["fdh"]) + 1e-2 * floor and (
            ill or abs(r["a"] - r["fdh"]) <= 1e-3 * max(abs(r["a"]), abs(r["fdh"])) + floor)
        n_ill += 1 if ill and not r["kink"] else 0
        n_kink += 1 if r["kink"] else 0
        rich = (4.0 * r["fd32"][1] - r["fd32"][0]) / 3.0
        rich_b = (4.0 * r["b"][1] + r["b"][0]) / 3.0
        print(f"  {r['cls']:>7}[{r['i']},{r['c']}]: analytic {r['a']:+.6e} (FP64 {r['a64']:+.6e}, one-ulp "
              f"sensitivity {r['sens']:.1e})  FD64(h=1e-6) {r['fdh']:+.6e}  "
              f"FD32 {r['fd32'][0]:+.6e} vs FP64 same decisions {r['fd64'][0]:+.6e} (f32 bound "
              f"{r['b'][0]:.1e})  decision changes {r['flips']}{'  L1 kink' if r['kink'] else ''}")
        bad1 += [] if ok1 else [r]
        bad2 += [] if ok2 else [r]
        if r["flips"] == 0 and r["cls"] != "pos" and not r["kink"]:
            # 3. Richardson of the fp32 differences against the analytic
            # gradient: the f32 bound, plus 1e-3 for the O(eps^4) remainder
            # (log-scale, rotation, opacity logit and SH steps of 1e-3 are
            # small against every length scale the loss has).
            kept += 1
            if abs(rich - r["a"]) > 2.0 * rich_b + 1e-3 * max(abs(rich), abs(r["a"])) + floor:
                bad3.append(r)
    print(f"  absolute floor {floor:.2e}")
    print(f"  1. GPU fp32 FD vs FP64 FD on the same decisions: {len(rows) - len(bad1)} of {len(rows)}")
    print(f"  2. GPU analytic vs FP64 central difference (base decisions): {len(rows) - len(bad2)} of {len(rows)} "
          f"({n_ill} ill-conditioned: FP64 analytic only; {n_kink} at an L1 kink: not compared)")
    print(f"  3. GPU analytic vs GPU fp32 FD (Richardson, no decision change, non-position): "
          f"{kept - len(bad3)} of {kept}")
    assert not bad1 and not bad2 and not bad3
    assert kept >= 16 and n_kink + n_ill <= 8
    for v in views:
        v.r.close()
