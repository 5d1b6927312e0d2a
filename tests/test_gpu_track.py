"""Photometric tracking on the device (gs_tracker_*, crates/gs-track) against
the restatements pinned in test_oracle_track.py, and the reference's own
behaviour tests (crates/gs-track/tests/recover_pose.rs) on its synthetic
room (crates/gs-map/src/synthetic.rs:22-66)."""

import numpy as np
import pytest

import oracle
from conftest import announce

pytestmark = pytest.mark.gpu

W = H = 320


def _rig(gpu):
    import splattar_amd as sa
    from splattar_amd import scenes
    room = scenes.room(24, 3.0)
    r = sa.Rasterizer(gpu, room.n, W, H, room.n * 64)
    r.upload_scene(room.positions, room.scales, room.quats, room.opacities, room.sh)
    focal = float(np.float32((H * 0.5) / np.tan(np.radians(30.0))))
    return r, room, sa.Tracker(gpu, r, W, H), focal


def _cam(center, quat, focal):
    import splattar_amd as sa
    return sa.Camera(np.asarray(center, np.float32), np.asarray(quat, np.float32), focal, 0)


def _render(r, room, cam):
    r.forward(cam, room.n)
    return r.out_color.read().copy()


def _errors(a, b):
    from splattar_amd.scenes import quat_mul
    pos = float(np.linalg.norm(a.center.astype(np.float64) - b.center.astype(np.float64)))
    qa = a.quat.astype(np.float64) / np.linalg.norm(a.quat)
    qb = b.quat.astype(np.float64) / np.linalg.norm(b.quat)
    rel = quat_mul(np.array([-qb[0], -qb[1], -qb[2], qb[3]]), qa)
    rot = float(np.degrees(2.0 * np.arccos(min(1.0, abs(rel[3])))))
    return pos, rot


def test_photometric_pass_and_one_pose_update_match_the_restatements(gpu):
    import splattar_amd as sa
    from splattar_amd import scenes
    r, room, tr, focal = _rig(gpu)
    truth = _cam([0.4, 0.1, -0.3], scenes.quat_from_rotation_y(0.6), focal)
    frame = _render(r, room, truth)
    tr.set_target(frame)
    start = _cam(truth.center + np.array([0.08, -0.05, 0.06], np.float32),
                 scenes.quat_mul(truth.quat, scenes.quat_from_rotation_x(np.radians(1.5))), focal)
    announce("tracking: photometric pass and one update")
    img = _render(r, room, start)
    res, cov = tr.photometric()
    dl = r.dl_dcolor.read().reshape(-1, 4)
    want_dl, want_res, want_cov = oracle.photometric(img, frame)
    assert np.array_equal(dl, want_dl)
    assert abs(res - want_res) <= 1e-6 * want_res and abs(cov - want_cov) <= 1e-9
    r.backward(room.n, flags=sa._lib.GS_BWD_COLOR_ONLY)
    gcam = r.grad_cam.read()[:12].copy()
    # Two iterations: the first applies the restated update, the second is
    # evaluated at the updated pose.
    cam = _cam(start.center, start.quat, focal)
    cfg = sa.TrackConfig(iterations=2, patience=100)
    rep = tr.track(room.n, cam, 3.0, cfg)
    g6 = oracle.camera_step_grad(gcam, start.quat)
    step = oracle.adam6(np.zeros(6, np.float32), np.zeros(6, np.float32), 1, g6, cfg.lr_rot, cfg.lr_center * 3.0)
    q1, c1 = oracle.apply_camera_step(start.quat, start.center, step)
    print(f"  residual {rep.residuals} vs {want_res:.6f}; step {step}")
    assert rep.iterations == 2 and abs(rep.residuals[0] - want_res) <= 1e-6 * want_res
    np.testing.assert_allclose(rep.last_quat, q1, rtol=0, atol=2e-6)
    np.testing.assert_allclose(rep.last_center, c1, rtol=0, atol=2e-6)
    tr.close()
    r.close()


def test_tracking_recovers_a_displaced_camera(gpu):
    """recover_pose.rs:84-124."""
    import splattar_amd as sa
    from splattar_amd import scenes
    r, room, tr, focal = _rig(gpu)
    truth = _cam([0.4, 0.1, -0.3], scenes.quat_from_rotation_y(0.6), focal)
    start = _cam(truth.center + np.array([0.08, -0.05, 0.06], np.float32),
                 scenes.quat_mul(truth.quat, scenes.quat_from_rotation_x(np.radians(1.5))), focal)
    tr.set_target(_render(r, room, truth))
    before = _errors(start, truth)
    cam = _cam(start.center, start.quat, focal)
    announce("tracking: recover a displaced camera")
    rep = tr.track(room.n, cam, 3.0, sa.TrackConfig())
    after = _errors(cam, truth)
    print(f"  pose error {before[0]:.4f} m / {before[1]:.2f} deg -> {after[0]:.4f} m / {after[1]:.2f} deg in "
          f"{rep.iterations} iters; residual {rep.residuals[0]:.5f} -> {rep.final_residual():.5f}, "
          f"coverage {rep.coverage:.2f}")
    assert rep.improved()
    assert after[0] < before[0] * 0.5 and after[1] < before[1] * 0.5
    tr.close()
    r.close()


def test_tracking_holds_a_correct_pose(gpu):
    """recover_pose.rs:125-139."""
    import splattar_amd as sa
    from splattar_amd import scenes
    r, room, tr, focal = _rig(gpu)
    truth = _cam([-0.2, 0.0, 0.5], scenes.quat_from_rotation_y(-0.3), focal)
    tr.set_target(_render(r, room, truth))
    cam = _cam(truth.center, truth.quat, focal)
    rep = tr.track(room.n, cam, 3.0, sa.TrackConfig())
    pos, rot = _errors(cam, truth)
    print(f"  held pose: drifted {pos:.5f} m / {rot:.3f} deg over {rep.iterations} iters")
    assert pos < 0.01 and rot < 0.2
    tr.close()
    r.close()


def test_the_residual_orders_poses_by_correctness(gpu):
    """recover_pose.rs:141-173: one iteration measures the starting residual."""
    import splattar_amd as sa
    from splattar_amd import scenes
    r, room, tr, focal = _rig(gpu)
    truth = _cam([0.1, 0.0, 0.2], scenes.quat_from_rotation_y(0.2), focal)
    tr.set_target(_render(r, room, truth))

    def probe(off):
        cam = _cam(truth.center + np.array([off, 0.0, 0.0], np.float32), truth.quat, focal)
        return tr.track(room.n, cam, 3.0, sa.TrackConfig(iterations=1)).final_residual()

    exact, near, far = probe(0.0), probe(0.05), probe(0.25)
    print(f"  residual by offset: 0.0 -> {exact:.5f}, 0.05 -> {near:.5f}, 0.25 -> {far:.5f}")
    assert exact < near < far
    tr.close()
    r.close()
