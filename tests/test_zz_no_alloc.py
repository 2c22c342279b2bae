"""fit_impedance_noalloc (the allocation-free variant of fit_impedance)
against fit_impedance, through the C ABI: bit-identical outputs on two calls
in a row on one device, the second with a smaller input (so stale scratch
contents of the first call would show), and the size limits rejected."""

import ctypes as C

import numpy as np
import pytest

from wakis_amd import _lib

pytestmark = pytest.mark.gpu


def _bufs(K):
    return [np.empty(K, dtype=np.float64) for _ in range(6)]


def _call(ctx, fn, x, lam, N, K, ds):
    out = _bufs(K)
    rc = fn(ctx.h, _lib.dptr(x), _lib.dptr(lam), x.size, N, K, ds, *[_lib.dptr(b) for b in out])
    return rc, out


def test_noalloc_impedance_matches_allocating_entry_point():
    lib = _lib.load()
    ctx = _lib.Context(4, 4, 4, dtype=np.float64, device=0)
    rng = np.random.default_rng(11)
    # (len, N, K): zero-padded DFT, then a smaller truncated one
    cases = [(20000, 110110, 994), (3001, 2500, 300)]
    try:
        for n, (ln, N, K) in enumerate(cases):
            x = rng.standard_normal(ln)
            lam = np.exp(-np.linspace(-4, 4, ln) ** 2) + 0.1
            ds = 1.0e-4 * (n + 1)
            print(f"noalloc case {n}: len={ln} N={N} K={K}", flush=True)
            rc_a, ref = _call(ctx, lib.fit_impedance, x, lam, N, K, ds)
            assert rc_a == 0, lib.fit_last_error()
            rc_b, got = _call(ctx, lib.fit_impedance_noalloc, x, lam, N, K, ds)
            assert rc_b == 0, lib.fit_last_error()
            for a, b in zip(ref, got):
                np.testing.assert_array_equal(a, b)
            # the DFT itself, bin by bin, against numpy (pins the kernel too)
            X = np.fft.fft(x, n=N)[:K] * ds
            np.testing.assert_allclose(got[2] + 1j * got[3], X, rtol=1e-9,
                                       atol=1e-9 * np.abs(X).max())
        # limits: rejected with FIT_EINVAL and a message, nothing launched
        big = np.zeros(262145)
        rc, _ = _call(ctx, lib.fit_impedance_noalloc, big, big, big.size, 4, 1.0)
        assert rc == -1 and b"scratch limit" in lib.fit_last_error()
        small = np.zeros(16)
        rc, _ = _call(ctx, lib.fit_impedance_noalloc, small, small, 16, 8193, 1.0)
        assert rc == -1
        rc, _ = _call(ctx, lib.fit_impedance_noalloc, small, small, 16 * 1024, 8193, 1.0)
        assert rc == -1 and b"scratch limit" in lib.fit_last_error()
    finally:
        ctx.close()
