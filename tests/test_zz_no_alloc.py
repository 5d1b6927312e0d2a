"""The no-allocation scan entry point (gs_scan_encode_noalloc) against the
handle-based scan (gs_scan_encode, prefix_sum.rs:134-164 semantics), through
the C ABI: two calls in a row, the second smaller, each bit-exact against
the handle's output and against numpy's wrapping cumsum."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[no_alloc] {msg}", flush=True)


def _handle_scan(sa, ctx, vals):
    s = sa.PrefixSum(ctx, len(vals))
    s.data().write(vals)
    s.encode(len(vals))
    out, tot = s.data().read(), int(s.total().read()[0])
    s.close()
    return out, tot


def _expected(vals):
    ex = np.zeros(len(vals), np.uint64)
    ex[1:] = np.cumsum(vals[:-1].astype(np.uint64))
    return (ex & 0xFFFFFFFF).astype(np.uint32), int(vals.astype(np.uint64).sum() & 0xFFFFFFFF)


def test_noalloc_scan_equals_the_handle_scan(gpu):
    import splattar_amd as sa
    L = sa.lib()
    rng = np.random.default_rng(7)
    big = 3_000_017  # 733 partitions, the last one ragged
    buf = gpu.alloc(np.uint32, big)
    total = gpu.alloc(np.uint32, 1)
    for n, hi in ((big, 1 << 31), (5_000, 1 << 20)):
        announce(f"n={n}")
        vals = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
        want, want_total = _handle_scan(sa, gpu, vals)
        ref, ref_total = _expected(vals)
        assert np.array_equal(want, ref) and want_total == ref_total
        buf.write(vals)  # a prefix of the larger buffer on the second call
        sa._lib.check(L.gs_scan_encode_noalloc(gpu.h, None, buf.ptr, n, total.ptr))
        got = buf.read()[:n]
        assert np.array_equal(got, want), f"n={n}: scan differs from gs_scan_encode"
        assert int(total.read()[0]) == want_total
    announce("n=0 writes a zero total")
    sa._lib.check(L.gs_scan_encode_noalloc(gpu.h, None, buf.ptr, 0, total.ptr))
    assert int(total.read()[0]) == 0
    buf.free()
    total.free()


def test_noalloc_scan_rejects_sizes_beyond_its_limit(gpu):
    import splattar_amd as sa
    L = sa.lib()
    total = gpu.alloc(np.uint32, 1)
    # Rejected on the host before any launch: the data pointer is never used.
    st = L.gs_scan_encode_noalloc(gpu.h, None, total.ptr, (1 << 28) + 1, total.ptr)
    assert sa._lib.STATUS_NAMES.get(st) == "GS_ERR_CAPACITY"
    assert b"no-allocation limit" in L.gs_last_error()
    total.free()
