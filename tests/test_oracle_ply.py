"""The .ply export restatement (oracle.ply_header / ply_rows of
first_attempt/crates/gs-io/src/ply/writer.rs:17-79) pinned by a round trip
through the restated reader (reader.rs:77-147) and by the reader's canonical
layout (layout.rs:115-125, 150-177)."""

import numpy as np

import oracle


def _scene(n=300, seed=4):
    rng = np.random.default_rng(seed)
    pos4 = np.concatenate([rng.normal(0, 1, (n, 3)), np.ones((n, 1))], 1).astype(np.float32)
    scales = np.exp(rng.normal(-4, 1, (n, 2))).astype(np.float32)
    q = rng.normal(0, 1, (n, 4)).astype(np.float32)
    op = rng.uniform(0.01, 0.99, n).astype(np.float32)
    sh = rng.normal(0, 0.3, (n, 48)).astype(np.float32)
    return pos4, scales, q, op, sh


def test_header_is_the_canonical_inria_layout():
    h = oracle.ply_header(7).decode().split("\n")
    assert h[0] == "ply" and h[1] == "format binary_little_endian 1.0" and h[3] == "element vertex 7"
    props = [x.split()[-1] for x in h if x.startswith("property float ")]
    # layout.rs canonical_order: x y z, f_dc_0..2, f_rest_0..N-1, opacity, scale_0..2, rot_0..3.
    want = (["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2"] + [f"f_rest_{i}" for i in range(45)] +
            ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"])
    assert props == want and h[-2] == "end_header" and h[-1] == ""


def test_rows_round_trip_through_the_reader():
    """writer then reader: positions and SH bit-exact (the f_rest
    transposition and its re-interleave are inverse), opacity and scales to
    float32 rounding of logit / exp, the quaternion reordered and normalized."""
    pos4, scales, q, op, sh = _scene()
    data = oracle.ply_header(len(op)) + oracle.ply_rows(pos4, scales, q, op, sh).astype("<f4").tobytes()
    p, shr, o, s, qr = oracle.ply_read_canonical(data)
    assert np.array_equal(p, pos4[:, :3])
    assert np.array_equal(shr.reshape(-1, 48), sh)
    np.testing.assert_allclose(o, op, rtol=2e-6)
    np.testing.assert_allclose(s[:, :2], scales, rtol=2e-6)
    np.testing.assert_allclose(s[:, 2], np.exp(np.float32(-10.0)), rtol=1e-7)
    qn = q / np.linalg.norm(q.astype(np.float64), axis=1, keepdims=True)
    np.testing.assert_allclose(qr, qn, atol=1e-6)


def test_rows_known_values():
    """One hand-checked row: f_rest_0 is coefficient 1 red, f_rest_15
    coefficient 1 green, f_rest_44 coefficient 15 blue; clamps at the
    opacity and scale floors; w first."""
    pos4 = np.array([[1.0, 2.0, 3.0, 1.0]], np.float32)
    sh = np.arange(48, dtype=np.float32).reshape(1, 48)
    r = oracle.ply_rows(pos4, np.array([[0.0, np.e]], np.float32), np.array([[0.1, 0.2, 0.3, 0.9]], np.float32),
                        np.array([1.0], np.float32), sh)[0]
    assert list(r[0:6]) == [1.0, 2.0, 3.0, 0.0, 1.0, 2.0]
    assert r[6] == 3.0 and r[6 + 15] == 4.0 and r[6 + 44] == 47.0
    o = np.float32(1.0 - 1e-6)  # opacity 1 clamped at 1 - 1e-6, then o / (1 - o) in f32
    assert abs(r[51] - np.log(np.float64(o) / (1.0 - np.float64(o)))) < 2e-6 * 13.8
    assert abs(r[52] - np.log(1e-9)) < 1e-5 and abs(r[53] - 1.0) < 1e-6 and r[54] == -10.0
    assert list(r[55:59]) == [np.float32(0.9), np.float32(0.1), np.float32(0.2), np.float32(0.3)]
