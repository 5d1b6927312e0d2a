"""GPU parity of the integer stages (projection to tile ranges), bit-exact, through the C ABI.

Case lists restate the reference's own GPU tests:
  scan    crates/gs-wgpu/tests/prefix_sum.rs:57-73
  sort    crates/gs-wgpu/tests/sort.rs:27-113 (32-bit keys, as the reference)
          plus 64-bit (tile || depth) keys, both digit widths
  binning crates/gs-kernels/tests/binning.rs:118-158
  project surfel_prep_fwd.wgsl:64-132 rect/key/count vs the fp32 mirror
"""

import numpy as np
import pytest

import oracle
from conftest import announce

pytestmark = pytest.mark.gpu


def test_exclusive_scan_reference_cases(gpu):
    import splattar_amd as sa
    scan = sa.PrefixSum(gpu, 10_000_000)
    cases = []
    for n in [1, 2, 255, 1024, 1025, 4096, 100_003, 1_048_576, 10_000_000]:
        cases.append((f"n={n}", oracle.xorshift32(0x5EED0001 ^ n, n, 16)))
    cases += [("zeros4096", np.zeros(4096, np.uint32)), ("ones2049", np.ones(2049, np.uint32)),
              ("[42]", np.array([42], np.uint32)),
              ("wrap", np.full(1000, 0xFFFFFFF0, np.uint32))]
    for label, vals in cases:
        announce(f"scan {label}")
        scan.data().write(vals)
        scan.encode(len(vals))
        got = scan.data().read(len(vals))
        want, total = oracle.exclusive_scan(vals)
        assert np.array_equal(got, want), label
        assert int(scan.total().read()[0]) == total, label
    scan.close()


@pytest.mark.parametrize("digit_bits", [8, 4])
def test_sort32_reference_cases(gpu, digit_bits):
    import splattar_amd as sa
    s = sa.RadixSorter(gpu, 10_000_000, key_bits=32)
    for n in [0, 1, 255, 256, 1024, 4097, 100_000, 1_000_003, 2_000_000, 4_194_304, 10_000_000]:
        announce(f"sort32 n={n} digit={digit_bits}")
        keys = oracle.xorshift32((0x12345678 ^ ((n * 2654435761) & 0xFFFFFFFF)) & 0xFFFFFFFF, n, 8192)
        pay = np.arange(n, dtype=np.uint32)
        s.keys().write(keys)
        s.payloads().write(pay)
        s.counts().write(np.array([n], np.uint32))
        s.encode(digit_bits=digit_bits)
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(s.keys().read(n), keys[order])
        assert np.array_equal(s.payloads().read(n), pay[order])
    s.close()


def test_sort32_count_above_capacity_sorts_the_capacity(gpu):
    """A device count larger than the sorter's capacity sorts the first
    `capacity` pairs (the kernels clamp the count; nothing past the buffers
    is read or written), with an odd pass count (3 passes of 8 bits) so the
    result is copied back."""
    import splattar_amd as sa
    cap = 10_007
    s = sa.RadixSorter(gpu, cap, key_bits=32)
    keys = oracle.xorshift32(0xC0FFEE, cap, 1 << 24)
    pay = np.arange(cap, dtype=np.uint32)
    announce("sort32 count above capacity")
    s.keys().write(keys)
    s.payloads().write(pay)
    s.counts().write(np.array([cap + 5000], np.uint32))
    s.encode(end_bit=24)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(s.keys().read(cap), keys[order])
    assert np.array_equal(s.payloads().read(cap), pay[order])
    s.close()


def test_sort32_full_range(gpu):
    # sort.rs:82-113: full 32-bit keys at n = 65537.
    import splattar_amd as sa
    n = 65_537
    s = sa.RadixSorter(gpu, n, key_bits=32)
    keys = oracle.xorshift32(0xDEADBEEF, n)
    announce("sort32 full range")
    s.keys().write(keys)
    s.payloads().write(np.arange(n, dtype=np.uint32))
    s.counts().write(np.array([n], np.uint32))
    s.encode()
    assert np.array_equal(s.keys().read(n), np.sort(keys))
    s.close()


@pytest.mark.parametrize("digit_bits,end_bit", [(8, 64), (8, 47), (4, 45)])
def test_sort64_tile_depth_keys(gpu, digit_bits, end_bit):
    import splattar_amd as sa
    n = 3_000_017
    s = sa.RadixSorter(gpu, n, key_bits=64)
    tile = oracle.xorshift32(7, n, 1 << (end_bit - 32))
    depth = oracle.xorshift32(11, n, 4096) | np.uint32(0x3F000000)  # heavy duplicates
    keys = (tile.astype(np.uint64) << np.uint64(32)) | depth.astype(np.uint64)
    pay = np.arange(n, dtype=np.uint32)
    announce(f"sort64 n={n} digit={digit_bits} end_bit={end_bit}")
    s.keys().write(keys)
    s.payloads().write(pay)
    s.counts().write(np.array([n], np.uint32))
    s.encode(begin_bit=0, end_bit=end_bit, digit_bits=digit_bits)
    k, p = oracle.stable_sort(keys, pay)
    assert np.array_equal(s.keys().read(n), k)
    assert np.array_equal(s.payloads().read(n), p)
    s.close()


def _items(num, seed, tiles=64):
    r = np.empty((num, 5), np.uint32)
    import microscene
    r[:] = microscene.xorshift32_stream(seed, num * 5).reshape(num, 5) if num < 20000 else \
        oracle.xorshift32(seed, num * 5).reshape(num, 5)
    x0 = r[:, 0] % tiles
    y0 = r[:, 1] % tiles
    w = r[:, 2] % 4
    h = r[:, 3] % 4
    rects = np.stack([x0, y0, np.minimum(x0 + w, tiles - 1), np.minimum(y0 + h, tiles - 1)], 1)
    return rects.astype(np.uint32), (r[:, 4] % 32).astype(np.uint32)


def _check_binner(gpu, rects, depths, label, digit_bits=8):
    import splattar_amd as sa
    n = len(depths)
    keys, pay, e = oracle.emit_keys(rects, depths, 64)
    b = sa.TileBinner(gpu, max(n, 1), max(e, 1), 64 * 64)
    announce(f"binner {label} n={n} e={e}")
    if n:
        b.rects.write(rects)
        b.depths.write(depths)
    b.encode(n, 64, digit_bits=digit_bits)
    assert int(b.entry_count().read()[0]) == e, label
    k, p = oracle.stable_sort(keys, pay)
    if e:
        assert np.array_equal(b.sorted_entries().read(e), p), label
        assert np.array_equal(b.sorted_keys().read(e), k), label
    assert np.array_equal(b.ranges.read(), oracle.tile_ranges(k, 64 * 64)), label
    assert not b.overflowed()
    b.close()


@pytest.mark.parametrize("num,seed", [(1, 7), (100, 11), (5000, 13), (100_000, 17)])
def test_binning_reference_cases(gpu, num, seed):
    rects, depths = _items(num, seed)
    _check_binner(gpu, rects, depths, f"random seed={seed}")


def test_binning_degenerate_cases(gpu):
    _check_binner(gpu, np.array([[5, 5, 4, 4]], np.uint32), np.array([1], np.uint32), "culled")
    rects = np.tile(np.array([[10, 10, 10, 10]], np.uint32), (500, 1))
    _check_binner(gpu, rects, (np.arange(500) % 3).astype(np.uint32), "stacked", digit_bits=4)


def test_binning_overflow_is_flagged(gpu):
    import splattar_amd as sa
    rects, depths = _items(100, 11)
    keys, pay, e = oracle.emit_keys(rects, depths, 64)
    b = sa.TileBinner(gpu, 100, e // 2, 64 * 64)
    announce("binner overflow")
    b.rects.write(rects)
    b.depths.write(depths)
    b.encode(100, 64)
    assert b.overflowed()
    assert int(b.entry_count().read()[0]) == e // 2
    b.close()


def test_scan_rejects_over_capacity(gpu):
    import splattar_amd as sa
    scan = sa.PrefixSum(gpu, 100)
    with pytest.raises(sa.GsError):
        scan.encode(101)
    scan.close()
