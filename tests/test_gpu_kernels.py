"""Device kernels against the oracle: the reference's kernel test grid
(tests/test_kernels.cpp:147-154), the rotation order, the normalisation."""
import os

import numpy as np
import pytest

import common
import katoann
import oracle as O

pytestmark = pytest.mark.gpu


def draw_bytes(rng, shape):
    c = rng.integers(0, 4, shape)
    u = rng.integers(0, 256, shape)
    return np.where(c == 0, 0, np.where(c == 1, 255, u)).astype(np.uint8)


@pytest.mark.parametrize("dc", [64, 128, 1024, 3072])
def test_sq16_kernel_grid(dc):
    rng = np.random.default_rng(7 + dc)
    ssq = 1.0 / (65535.0 * 65535.0)
    for m in (1, 5, 16):
        for n in (1, 15, 16, 17, 70, 256):
            qh, ql, xh, xl = (draw_bytes(rng, s) for s in ((m, dc), (m, dc), (n, dc), (n, dc)))
            qc = qh.astype(np.int64) * 256 + ql
            xc = xh.astype(np.int64) * 256 + xl
            qn, xn = (qc * qc).sum(1), (xc * xc).sum(1)
            want = O.sq16_l2(qh, ql, xh, xl, qn, xn, ssq)
            got = katoann.sq16_l2(qh, ql, xh, xl, qn, xn, ssq)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (m, n, dc)
            qhn = (qh.astype(np.int64) ** 2).sum(1)
            xhn = (xh.astype(np.int64) ** 2).sum(1)
            want_h = O.sq16_high_l2(qh, xh, qhn, xhn, ssq, dc)
            got_h = katoann.sq16_high_l2(qh, xh, qhn, xhn, ssq)
            assert np.array_equal(got_h.view(np.uint32), want_h.view(np.uint32)), (m, n, dc)
            pd = 64
            while pd < dc:  # leading pd dims of the dc-wide plane, read in place
                qp = np.ascontiguousarray(qh[:, :pd])
                qpn = (qp.astype(np.int64) ** 2).sum(1)
                xpn = (xh[:, :pd].astype(np.int64) ** 2).sum(1)
                want_p = O.sq16_high_l2(qp, xh, qpn, xpn, ssq, pd)
                got_p = katoann.sq16_high_l2(qp, xh, qpn, xpn, ssq)
                assert np.array_equal(got_p.view(np.uint32), want_p.view(np.uint32)), (m, n, dc, pd)
                pd *= 2


@pytest.mark.parametrize("d", [1, 3, 7, 16, 50, 66, 100, 128, 1024, 1030])
def test_rotation_bit_exact(d):
    rng = np.random.default_rng(d)
    for n in (1, 5, 33, 200):
        x = rng.standard_normal((n, d)).astype(np.float32)
        r = rng.standard_normal((d, d)).astype(np.float32) / np.float32(np.sqrt(d))
        want = O.rotate_rows(x, r)
        got = katoann.rotate_rows(x, r)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (n, d)


def test_normalisation_bit_exact():
    rng = np.random.default_rng(5)
    for d in (1, 100, 1024, 3072):
        x = (rng.standard_normal((64, d)) * rng.uniform(0.1, 30, (64, 1))).astype(np.float32)
        x[3] = 0.0
        assert np.array_equal(katoann.normalize_rows(x).view(np.uint32), O.normalize_rows(x).view(np.uint32))


REF_KERNEL_TEST = os.path.join(common.ROOT, "oracle", "_ref", "test_kernels_gfx950")


@pytest.mark.skipif(not os.path.exists(REF_KERNEL_TEST),
                    reason="built by `make -C oracle ref` where the reference sources are")
def test_reference_kernel_test_passes():
    """The reference's own tests/test_kernels.cpp, compiled unchanged against
    include/katoann/kernels.h and linked to libkatoann.so: sq16_l2,
    sq16_high_l2 (also a prefix read in place from a wider plane) and bf16_l2
    on the GPU against its scalar formulas, bit-exact for the 16-bit codes,
    over dims 64..3072 and partial tiles."""
    import subprocess
    r = subprocess.run([REF_KERNEL_TEST], capture_output=True, text=True, timeout=240)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0 and "all kernel checks passed" in r.stdout


REF_INDEX_TEST = os.path.join(common.ROOT, "oracle", "_ref", "test_index_gfx950")


@pytest.mark.skipif(not os.path.exists(REF_INDEX_TEST),
                    reason="built by `make -C oracle ref` where the reference sources are")
def test_reference_index_test_passes():
    """The reference's own tests/test_index.cpp, compiled unchanged against
    include/katoann/index.h and linked to libkatoann.so: build, load, recall of
    the full scan, prefix and high-byte ranking with 4,000 of 40,000
    candidates surviving equal to the full scan, local = threshold selection,
    the per-query coarse search and bf16 precision, on the GPU engine."""
    import subprocess
    print("reference test_index", flush=True)
    r = subprocess.run([REF_INDEX_TEST], capture_output=True, text=True, timeout=600)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert "all index checks passed" in r.stdout
    assert "4,000 of 40,000 candidates surviving equals the full scan" in r.stdout
