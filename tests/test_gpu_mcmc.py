"""MCMC densification on the device against the restatements pinned in
test_oracle_mcmc.py: the noise kernel (mcmc.wgsl:39-58), the trainer's
relocation (trainer.rs:1558-1690 sampling + relocate.wgsl data movement) and
the noise of captured graph steps."""

import struct

import numpy as np
import pytest

import oracle
from conftest import announce

pytestmark = pytest.mark.gpu

CLASSES = ("pos", "scales", "quat", "opacity", "sh")
WIDTH = {"pos": 4, "scales": 2, "quat": 4, "opacity": 1, "sh": 48}


def _parse_state(blob, n):
    """gs_optim_export_state: header, class descriptors, Adam state, then per
    class raw, m, v."""
    magic, ver, ncls, sbytes = struct.unpack_from("<4I", blob, 0)
    off = 16 + 40 * ncls + sbytes
    raw, m, v = {}, {}, {}
    for k in CLASSES:
        cnt = n * WIDTH[k]
        for d in (raw, m, v):
            a = np.frombuffer(blob, np.float32, cnt, off).copy()
            d[k] = a.reshape(n, WIDTH[k]) if WIDTH[k] > 1 else a
            off += 4 * cnt
    return raw, m, v


def _trainer(gpu, n=3000, seed=5):
    import splattar_amd as sa
    from splattar_amd import scenes
    sc = scenes.make_scene(n, 3, -3.5, 0.5, seed)
    rng = np.random.default_rng(seed)
    op = sc.opacities.copy()
    op[rng.random(n) < 0.1] = 0.002  # dead surfels
    sc.opacities[:] = op
    sc.scales[:5] *= 30.0  # oversized surfels to split
    views = [(sa.Camera(sp.center, sp.quat, sp.focal, 3), np.full((48, 64, 4), 0.3, np.float32))
             for sp in scenes.sphere_views(2, seed, [(64, 48)])]
    return sa.Trainer(gpu, sa.TrainConfig(max_views_per_step=2), sc.as_init(), views), sc


def test_noise_kernel_matches_the_restatement(gpu):
    import splattar_amd as sa
    n = 100_003
    rng = np.random.default_rng(3)
    pos = np.concatenate([rng.normal(0, 1, (n, 3)), np.ones((n, 1))], 1).astype(np.float32)
    sc = np.exp(rng.normal(-4, 1, (n, 2))).astype(np.float32)
    op = rng.uniform(0, 1, n).astype(np.float32)
    bufs = [gpu.alloc(np.float32, s) for s in ((n, 4), (n, 2), (n,))]
    for b, a in zip(bufs, (pos, sc, op)):
        b.write(a)
    announce("mcmc noise kernel")
    sa.mcmc_add_noise(gpu, n, 0x9e3779b9 * 7 & 0xFFFFFFFF, 0.05, 5.0, *bufs)
    got = bufs[0].read()
    for b in bufs:
        b.free()
    want = oracle.mcmc_noise(pos, sc, op, 0x9e3779b9 * 7 & 0xFFFFFFFF, 0.05, 5.0)
    dg, dw = got - pos, want - pos
    assert np.array_equal(got[:, 3], pos[:, 3]) and np.abs(dw).max() > 0
    # Identical PCG draws; exp within an ulp or two of the correctly rounded f32.
    err = np.abs(dg - dw).max()
    print(f"  noise: max |d| {np.abs(dw).max():.3e}, GPU vs restatement {err:.2e}")
    # Per element: one ulp of the result (the final p + amp n rounding) plus
    # 4e-7 of the move (amp through expf).
    assert np.all(np.abs(got - want) <= np.spacing(np.abs(want)) + 4e-7 * np.abs(dw))


def test_trainer_relocation_matches_the_restatement(gpu):
    """Two steps, then one relocation: the pairs the host sampled from the
    activated opacities / scales, the raw parameters and moments after the
    move (every element of unpaired surfels bit-identical, paired ones to
    the ulp of log / exp / sqrt), and the re-activated copies."""
    t, sc = _trainer(gpu)
    n = sc.n
    t.step([0, 1], mode=t.GRAPH)
    t.step([1, 0], mode=t.GRAPH)
    op_act = t.opacities.read().copy()
    sc_act = t.scales.read().copy()
    raw, m, v = _parse_state(t.export_state(), n)
    pairs, rng_w = oracle.mcmc_sample_pairs(op_act, sc_act, 0.005, 8.0, 0xBADC0FFEE)
    announce(f"relocation of {len(pairs)} pairs")
    moved, rng_g = t.mcmc_relocate(0.005, 8.0, 0xBADC0FFEE)
    assert moved == len(pairs) and rng_g == rng_w
    assert sum(1 for p in pairs if p[2] == 1) >= 1 and sum(1 for p in pairs if p[2] == 0) >= 50
    oracle.mcmc_relocate_arrays(pairs, op_act, raw, m, v)
    raw_g, m_g, v_g = _parse_state(t.export_state(), n)
    touched = np.zeros(n, bool)
    for d, a, _, _ in pairs:
        touched[d] = touched[a] = True
    for k in CLASSES:
        assert np.array_equal(raw_g[k][~touched], raw[k][~touched]), k
        assert np.array_equal(m_g[k], m[k]) and np.array_equal(v_g[k], v[k]), k
        np.testing.assert_allclose(raw_g[k][touched], raw[k][touched], rtol=2e-6, atol=2e-7, err_msg=k)
    # Activated copies follow the raw values (exp / sigmoid / identity).
    np.testing.assert_allclose(t.opacities.read(), 1 / (1 + np.exp(-raw_g["opacity"].astype(np.float64))),
                               rtol=1e-5)
    np.testing.assert_allclose(t.positions.read()[:, :3], raw_g["pos"][:, :3], rtol=0, atol=0)
    t.close()


def test_graph_and_plain_steps_apply_the_noise_of_their_step(gpu):
    """Noise on (1e5 x the position rate 1e-4), every other rate 0: each
    step's position change is the restated noise of that step (seed =
    step index x 0x9e3779b9, sigma = noise x rate; activated scales and
    opacities constant) plus Adam's move, which is at most a few x the rate.
    Graph replays and plain steps alike; successive steps draw different
    noise."""
    import splattar_amd as sa
    from splattar_amd import scenes
    n = 3000
    sc = scenes.make_scene(n, 3, -3.5, 0.5, 5)
    views = [(sa.Camera(sp.center, sp.quat, sp.focal, 3), np.full((48, 64, 4), 0.3, np.float32))
             for sp in scenes.sphere_views(2, 5, [(64, 48)])]
    lr = 1e-4
    cfg = sa.TrainConfig(lr_pos=lr, lr_scales=0.0, lr_quat=0.0, lr_opacity=0.0, lr_sh=0.0, max_views_per_step=2)
    for mode in (sa.Trainer.GRAPH, sa.Trainer.PLAIN):
        announce(f"noise in steps, mode {mode}")
        t = sa.Trainer(gpu, cfg, sc.as_init(), views)
        t.set_mcmc_noise(1e5)
        sc_act, op_act = t.scales.read().copy(), t.opacities.read().copy()
        prev = _parse_state(t.export_state(), n)[0]["pos"]
        noises = []
        for it in range(3):
            t.step([it % 2, (it + 1) % 2], mode=mode)
            cur = _parse_state(t.export_state(), n)[0]["pos"]
            want = oracle.mcmc_noise(prev, sc_act, op_act, (it * 0x9e3779b9) & 0xFFFFFFFF, 1e5 * lr) - prev
            dev = np.abs((cur - prev) - want)[:, :3].max()
            print(f"  mode {mode} step {it}: noise max {np.abs(want).max():.3e}, deviation {dev:.2e}")
            assert np.abs(want).max() > 100 * lr and dev <= 4 * lr
            noises.append(want)
            prev = cur
        assert np.array_equal(t.scales.read(), sc_act) and np.array_equal(t.opacities.read(), op_act)
        assert not np.allclose(noises[1], noises[2])
        t.close()
