"""The baked 3DGS .ply export (gs_write_3dgs_ply / gs_trainer_write_ply, the
reference's write_3dgs_ply, first_attempt/crates/gs-io/src/ply/writer.rs:
17-79) against the writer restatement (oracle.ply_header / ply_rows, pinned
in test_oracle_ply.py): header bytes identical, every copied field bit-exact,
the three logarithms (opacity logit, two log scales) within 2 ulp of the
correctly rounded f32 value (measured: the device logf is within 2)."""

import os

import numpy as np
import pytest

import oracle
from conftest import announce

pytestmark = pytest.mark.gpu


def _scene(n, seed=9):
    rng = np.random.default_rng(seed)
    pos = rng.normal(0, 1, (n, 3)).astype(np.float32)
    scales = np.exp(rng.normal(-4, 1.5, (n, 2))).astype(np.float32)
    scales[:3] = [[0.0, 1e-12], [1.0, 2.0], [3e-10, 1e-9]]  # the 1e-9 floor
    quats = rng.normal(0, 1, (n, 4)).astype(np.float32)
    op = rng.uniform(0, 1, n).astype(np.float32)
    op[:4] = [0.0, 1.0, 1e-7, 1 - 1e-8]  # the [1e-6, 1 - 1e-6] clamp
    sh = rng.normal(0, 0.3, (n, 16, 3)).astype(np.float32)
    return pos, scales, quats, op, sh


def _check_file(data, count, pos4, scales, quats, op, sh48):
    head = oracle.ply_header(count)
    assert data[:len(head)] == head
    rows = np.frombuffer(data[len(head):], dtype="<f4").reshape(count, 59)
    want = oracle.ply_rows(pos4, scales, quats, op, sh48)
    logs = [51, 52, 53]
    other = [k for k in range(59) if k not in logs]
    assert np.array_equal(rows[:, other].view(np.uint32), want[:, other].view(np.uint32))
    ulp = np.abs(rows[:, logs].view(np.int32).astype(np.int64) - want[:, logs].view(np.int32).astype(np.int64))
    print(f"  {count} rows; log fields within {int(ulp.max())} ulp")
    assert ulp.max() <= 2


@pytest.mark.parametrize("n", [1234, 300_000])
def test_export_matches_the_writer(gpu, tmp_path, n):
    """n = 1234 (a ragged last block) and 300k (two device chunks of 2^18)."""
    import splattar_amd as sa
    pos, scales, quats, op, sh = _scene(n)
    announce(f"ply export n={n}")
    r = sa.Rasterizer(gpu, n, 64, 64, 1024)
    r.upload_scene(pos, scales, quats, op, sh.reshape(n, -1))
    path = os.path.join(tmp_path, "scene.ply")
    sa.write_3dgs_ply(gpu, path, n, r.positions, r.scales, r.quats, r.opacities, r.sh)
    pos4 = r.positions.read()[:n]
    sh48 = r.sh.read()[:n]
    assert np.array_equal(pos4[:, :3], pos) and np.array_equal(sh48, sh.reshape(n, 48))
    with open(path, "rb") as f:
        data = f.read()
    _check_file(data, n, pos4, scales, quats, op, sh48)
    # The device rows alone equal the file's.
    rows = gpu.alloc(np.float32, (n, 59))
    sa.ply_pack_rows(gpu, n, r.positions, r.scales, r.quats, r.opacities, r.sh, rows)
    head = oracle.ply_header(n)
    assert rows.read().astype("<f4").tobytes() == data[len(head):]
    rows.free()
    r.close()


def test_export_of_zero_surfels_and_an_unwritable_path(gpu, tmp_path):
    import splattar_amd as sa
    r = sa.Rasterizer(gpu, 16, 64, 64, 1024)
    path = os.path.join(tmp_path, "empty.ply")
    sa.write_3dgs_ply(gpu, path, 0, r.positions, r.scales, r.quats, r.opacities, r.sh)
    with open(path, "rb") as f:
        assert f.read() == oracle.ply_header(0)
    with pytest.raises(sa.GsError) as e:
        sa.write_3dgs_ply(gpu, os.path.join(tmp_path, "no_such_dir", "x.ply"), 16, r.positions, r.scales,
                          r.quats, r.opacities, r.sh)
    assert "GS_ERR_IO" in str(e.value)
    r.close()


def test_trainer_export_after_steps(gpu, tmp_path):
    """Trainer::read_scene + write_3dgs_ply: the file holds the activated
    parameters after two graph steps (rows against the trainer's buffers)."""
    import splattar_amd as sa
    from splattar_amd import scenes
    sc = scenes.make_scene(700, 2, -3.0, 0.4, 3)
    views = []
    for sp in scenes.sphere_views(2, 5, [(64, 48)]):
        views.append((sa.Camera(sp.center, sp.quat, sp.focal, 2), np.full((48, 64, 4), 0.4, np.float32)))
    t = sa.Trainer(gpu, sa.TrainConfig(max_views_per_step=2), sc.as_init(), views)
    t.step([0, 1], mode=t.GRAPH)
    t.step([1, 0], mode=t.GRAPH)
    path = os.path.join(tmp_path, "trained.ply")
    t.write_ply(path)
    n = sc.n
    pos4, sc2, q, op, sh48 = (t.positions.read(), t.scales.read(), t.quats.read(), t.opacities.read(),
                             t.sh.read())
    t.close()
    with open(path, "rb") as f:
        _check_file(f.read(), n, pos4, sc2, q, op, sh48)
