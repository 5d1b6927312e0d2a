"""GPU tests of the optimizer's device-resident hyperparameters and of exact
resume.

Learning-rate schedule: Trainer::step decays the position rate as
lr_pos * final_factor^(iter / iters) every iteration (first_attempt/crates/
gs-train/src/trainer.rs:1149-1151, schedule_t :407-416). Here the rate is
computed by the step on the device from its own step counter, so plain,
batched and graph-replayed steps follow the same schedule; the checks are
the reference's Adam loop (adam.rs:37-98, |d| < 1e-5) under the schedule and
graph == plain over 10 decaying steps.

Exact resume: the reference snapshots raw parameters only (optim.rs:148-165);
gs_optim_export_state / import_state carry raw, m, v, the step count and the
rates, and a resumed optimizer continues bit for bit.
"""

import numpy as np
import pytest

import oracle
from conftest import announce, assert_steps_close
from test_gpu_train import _params, _render_targets, _small_training_problem

pytestmark = pytest.mark.gpu


def _sched_lr(base, final, iters, i):
    """f32 rate of the step with 0-based index i (trainer.rs:1149-1151)."""
    t = np.float32(i) / np.float32(iters)
    return np.float32(base) * np.float32(np.power(np.float64(final), np.float64(t)))


def test_adam_follows_a_learning_rate_schedule(gpu):
    import splattar_amd as sa
    N, STEPS, BASE, FINAL, ITERS = 1000, 10, 0.01, 0.01, 8
    rng = np.random.default_rng(17)
    init = rng.uniform(-0.8, 0.8, N).astype(np.float32)
    grads = [rng.uniform(-1, 1, N).astype(np.float32) for _ in range(STEPS)]
    gbuf, act = gpu.alloc(np.float32, N), gpu.alloc(np.float32, N)
    opt = sa.Optimizer(gpu)
    opt.add_class("test", N, opt.IDENTITY, 0.5, gbuf, act)
    opt.set_lr_schedule("test", BASE, FINAL, ITERS)
    opt.raw("test").write(init)
    raw, m, v = init.copy(), np.zeros(N, np.float32), np.zeros(N, np.float32)
    announce("adam under an exponential learning-rate schedule")
    for i, g in enumerate(grads):
        gbuf.write(g)
        opt.encode_step()
        want = _sched_lr(BASE, FINAL, ITERS, i)
        got = opt.current_lr("test")
        # powf on the device against numpy's pow: a few ulp.
        assert abs(got - want) <= 4e-7 * want, (i, got, want)
        oracle.adam_step_f32(raw, g, m, v, float(got), 0, i + 1)
    assert np.abs(opt.raw("test").read() - raw).max() < 1e-5
    # set_lr replaces the schedule with a constant rate.
    opt.set_lr("test", 0.25)
    gbuf.write(grads[0])
    opt.encode_step()
    assert opt.current_lr("test") == np.float32(0.25)
    opt.close()
    gbuf.free()
    act.free()


def test_graph_steps_follow_the_position_lr_decay(gpu):
    """10 steps with lr_pos * 0.01^(iter / 10): graph replays (one capture)
    equal plain steps, and the position rate is the schedule's at every step."""
    import splattar_amd as sa
    deg = 2
    sc, tgt, specs = _small_training_problem(deg=deg)
    announce("lr decay: targets")
    targets = _render_targets(gpu, tgt, specs, deg)
    cfg = sa.TrainConfig(lam=0.2, dssim_scale=0.5, lr_pos=4e-3, lr_scales=1e-3, lr_quat=1e-3,
                         lr_opacity=1e-3, lr_sh=1e-3, max_views_per_step=3,
                         pos_lr_final_factor=0.01, iters=10)
    views = [(sa.Camera(sp.center, sp.quat, sp.focal, deg), t) for sp, t in zip(specs, targets)]
    res = {}
    for mode in (sa.Trainer.PLAIN, sa.Trainer.GRAPH):
        t = sa.Trainer(gpu, cfg, sc.as_init(), views)
        o = t.optim()
        lrs = []
        announce(f"lr decay: 10 steps mode {mode}")
        for it in range(10):
            t.step([0, 1, 2], mode=mode)
            lrs.append(o.current_lr("pos"))
        res[mode] = (_params(t), np.array(lrs))
        t.close()
    want = np.array([_sched_lr(4e-3, 0.01, 10, i) for i in range(10)])
    for mode, (_, lrs) in res.items():
        np.testing.assert_allclose(lrs, want, rtol=4e-7, err_msg=f"mode {mode}")
    assert res[0][1][-1] < 0.02 * 4e-3  # decayed to ~0.013 of the base
    p0, pg = res[sa.Trainer.PLAIN][0], res[sa.Trainer.GRAPH][0]
    # Same math; float atomics and the fused geometry pass order the sums differently.
    assert_steps_close(pg, p0, "graph vs plain")


def test_rate_change_between_graph_replays_takes_effect(gpu):
    """A captured step reads its rates from device memory: setting the SH
    rate to 0 between two replays freezes SH in the second replay without a
    recapture, and setting it back moves SH again."""
    import splattar_amd as sa
    deg = 2
    sc, tgt, specs = _small_training_problem(deg=deg)
    targets = _render_targets(gpu, tgt, specs, deg)
    cfg = sa.TrainConfig(lam=0.2, lr_pos=1e-3, lr_scales=1e-3, lr_quat=1e-3, lr_opacity=1e-3,
                         lr_sh=1e-2, max_views_per_step=3)
    views = [(sa.Camera(sp.center, sp.quat, sp.focal, deg), t) for sp, t in zip(specs, targets)]
    t = sa.Trainer(gpu, cfg, sc.as_init(), views)
    o = t.optim()
    announce("rate change between graph replays")
    t.step([0, 1, 2], mode=t.GRAPH)
    sh1, pos1 = t.sh.read().copy(), t.positions.read().copy()
    o.set_lr("sh", 0.0)
    t.step([0, 1, 2], mode=t.GRAPH)
    sh2, pos2 = t.sh.read().copy(), t.positions.read().copy()
    assert np.array_equal(sh2, sh1)
    assert not np.array_equal(pos2, pos1)
    o.set_lr("sh", 1e-2)
    t.step([0, 1, 2], mode=t.GRAPH)
    assert not np.array_equal(t.sh.read(), sh2)
    t.close()


def test_optimizer_state_round_trip_is_exact(gpu):
    """export -> import into a fresh optimizer -> the same steps give the same
    bits as the original (Adam is elementwise: no float-atomic order)."""
    import splattar_amd as sa
    N = 777
    rng = np.random.default_rng(23)
    grads = [rng.normal(size=N).astype(np.float32) for _ in range(6)]
    opts, bufs = [], []
    for _ in range(2):
        g, a = gpu.alloc(np.float32, N), gpu.alloc(np.float32, N)
        g2, a2 = gpu.alloc(np.float32, 2 * N), gpu.alloc(np.float32, 2 * N)
        o = sa.Optimizer(gpu)
        o.add_class("opacity", N, o.SIGMOID, 0.05, g, a)
        o.add_class("scales", 2 * N, o.EXP, 0.005, g2, a2)
        o.set_lr_schedule("scales", 0.005, 0.1, 5)
        o.set_reg("opacity", 0.01)
        opts.append(o)
        bufs.append((g, a, g2, a2))
    A, B = opts
    A.raw("opacity").write(rng.normal(size=N).astype(np.float32))
    A.raw("scales").write(rng.normal(-4, 0.5, 2 * N).astype(np.float32))
    announce("optimizer state round trip")
    for k in range(3):
        bufs[0][0].write(grads[k])
        bufs[0][2].write(np.tile(grads[k], 2))
        A.encode_step()
    blob = A.export_state()
    B.import_state(blob)
    assert B.export_state() == blob
    assert B.step_count() == 3
    assert np.array_equal(bufs[1][1].read(), bufs[0][1].read())  # activated after import
    for k in range(3, 6):
        for o, (g, _, g2, _) in zip(opts, bufs):
            g.write(grads[k])
            g2.write(np.tile(grads[k], 2))
            o.encode_step()
    assert A.export_state() == B.export_state()
    assert np.array_equal(bufs[0][3].read(), bufs[1][3].read())
    assert B.current_lr("scales") == A.current_lr("scales")
    # A blob of another layout or a truncated blob is refused.
    other = sa.Optimizer(gpu)
    g3, a3 = gpu.alloc(np.float32, N), gpu.alloc(np.float32, N)
    other.add_class("opacity", N, other.SIGMOID, 0.05, g3, a3)
    with pytest.raises(sa.GsError):
        other.import_state(blob)
    with pytest.raises(sa.GsError):
        B.import_state(blob[: len(blob) // 2])
    with pytest.raises(sa.GsError):
        B.import_state(b"\0" * len(blob))
    other.close()
    for o in opts:
        o.close()
    for b in (*bufs[0], *bufs[1], g3, a3):
        b.free()


def test_trainer_resume_continues_the_run(gpu):
    """A trainer resumed from an exported state has the exporter's
    parameters bit for bit and continues the same training run: 3 + 3 steps
    resumed equal 6 uninterrupted steps (up to float-atomic order)."""
    import splattar_amd as sa
    deg = 2
    sc, tgt, specs = _small_training_problem(deg=deg)
    targets = _render_targets(gpu, tgt, specs, deg)
    cfg = sa.TrainConfig(lam=0.2, lr_pos=2e-3, lr_scales=1e-3, lr_quat=1e-3, lr_opacity=1e-2,
                         lr_sh=1e-3, max_views_per_step=3, pos_lr_final_factor=0.1, iters=6)
    views = [(sa.Camera(sp.center, sp.quat, sp.focal, deg), t) for sp, t in zip(specs, targets)]
    a = sa.Trainer(gpu, cfg, sc.as_init(), views)
    announce("trainer resume: 3 graph steps + export")
    for it in range(3):
        a.step([it % 3, (it + 1) % 3, (it + 2) % 3], mode=a.GRAPH)
    blob = a.export_state()
    b = sa.Trainer(gpu, cfg, sc.as_init(), views)
    b.import_state(blob)
    pa, pb = _params(a), _params(b)
    for k in pa:
        assert np.array_equal(pa[k], pb[k]), k
    announce("trainer resume: 3 more steps on both")
    for it in range(3, 6):
        for t in (a, b):
            t.step([it % 3, (it + 1) % 3, (it + 2) % 3], mode=t.GRAPH)
    pa, pb = _params(a), _params(b)
    assert_steps_close(pb, pa, "resumed")
    assert b.optim().step_count() == 6
    assert b.optim().current_lr("pos") == a.optim().current_lr("pos")
    a.close()
    b.close()


def test_reads_after_a_step_on_a_caller_stream(gpu):
    """Steps enqueued on a caller's stream: the blocking loss read, the entry
    count and a render wait for them (the trainer records an event on the
    step's stream). The stream comes from the HIP runtime the library itself
    links, default priority, non-blocking."""
    import ctypes

    import splattar_amd as sa
    deg = 2
    sc, tgt, specs = _small_training_problem(deg=deg)
    targets = _render_targets(gpu, tgt, specs, deg)
    cfg = sa.TrainConfig(lam=0.2, lr_pos=1e-3, lr_scales=1e-3, lr_quat=1e-3, lr_opacity=1e-3,
                         lr_sh=1e-3, max_views_per_step=3)
    views = [(sa.Camera(sp.center, sp.quat, sp.focal, deg), t) for sp, t in zip(specs, targets)]
    ref = sa.Trainer(gpu, cfg, sc.as_init(), views)
    t = sa.Trainer(gpu, cfg, sc.as_init(), views)
    hip = ctypes.CDLL("libamdhip64.so.7")
    side = ctypes.c_void_p()
    assert hip.hipStreamCreateWithFlags(ctypes.byref(side), 1) == 0
    announce("steps on a caller stream, then reads")
    for it in range(4):
        ids = [it % 3, (it + 1) % 3, (it + 2) % 3]
        ref.step(ids, mode=ref.GRAPH)
        t.step(ids, mode=t.GRAPH, stream=side.value)
    # No synchronisation of `side` by the caller before the reads.
    l_t, e_t, img_t = t.last_losses(3), t.entry_count(), t.render(1)
    l_r, e_r, img_r = ref.last_losses(3), ref.entry_count(), ref.render(1)
    np.testing.assert_allclose(l_t, l_r, rtol=1e-4)
    assert e_t == e_r
    assert np.abs(img_t - img_r).max() < 1e-4
    assert hip.hipStreamSynchronize(side) == 0
    t.close()
    ref.close()
    assert hip.hipStreamDestroy(side) == 0
