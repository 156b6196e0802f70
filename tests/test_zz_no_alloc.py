"""bcmgx_spmm_noalloc (no device allocation; operands staged in the library's
fixed per-device scratch array) against bcmgx_spmm on the same device blocks:
bit for bit (both run the same kernels in the same order), on calls in a row
whose sizes shrink, and the size limit of the scratch array."""
import ctypes as C

import numpy as np
import pytest

import bootcmatchgx_amd as bx
import oracle_py as O
from bootcmatchgx_amd import abi

pytestmark = pytest.mark.gpu


def _spmm(L, name, ctx, A, X, Y, k, ld, alpha, beta, op, kchunk):
    if name == "bcmgx_spmm":
        st = L.bcmgx_spmm(ctx.h, A.h, X.ptr, Y.ptr, k, ld, alpha, beta, op, kchunk, bx.MEM_DEVICE)
    else:
        st = L.bcmgx_spmm_noalloc(ctx.h, A.h, X.ptr, Y.ptr, k, ld, alpha, beta, op, kchunk)
    assert st == abi.OK, L.bcmgx_last_error(ctx.h)
    assert L.bcmgx_context_synchronize(ctx.h) == abi.OK
    return Y.download()


# (matrix, k, ld padding, alpha, beta, op, kchunk): every call smaller than the one before
CALLS = [
    (lambda: O.laplacian3d(27, 40, 40, 40), 32, 3, 1.0, 0.0, 1, 0),
    (lambda: O.laplacian3d(27, 30, 30, 30), 16, 0, -1.0, 1.0, 2, -1),
    (lambda: O.diffusion7p(20, 18, 16, 1, 0.5, 2.0, 4), 5, 7, 0.5, -2.0, 0, 0),
    (lambda: O.laplacian3d(7, 12, 12, 12), 1, 0, 1.5, 0.25, 0, 0),
]


def test_noalloc_spmm_matches_spmm_on_shrinking_calls():
    ctx = bx.Context(0)
    L = bx.lib()
    rng = np.random.default_rng(41)
    for make, k, pad, alpha, beta, op, kchunk in CALLS:
        oa = make()
        n, m, _ = oa.dims()
        r, c, v = oa.arrays()
        A = bx.CSR(ctx, n, m, r, c, v)
        ld = n + pad
        x = rng.normal(size=ld * k)
        y0 = rng.normal(size=ld * k)
        X = bx.DeviceArray(ctx, (ld * k,))
        X.upload(x)
        out = {}
        for name in ("bcmgx_spmm", "bcmgx_spmm_noalloc"):  # bcmgx_spmm first: it tunes k = 1 lanes
            Y = bx.DeviceArray(ctx, (ld * k,))
            Y.upload(y0)
            out[name] = _spmm(L, name, ctx, A, X, Y, k, ld, alpha, beta, op, kchunk)
            Y.close()
        print(f"n={n} k={k} op={op}: noalloc == spmm: {np.array_equal(out['bcmgx_spmm'], out['bcmgx_spmm_noalloc'])}")
        assert np.array_equal(out["bcmgx_spmm"], out["bcmgx_spmm_noalloc"])
        # the padding rows below n in each column are left as they were
        Yn = out["bcmgx_spmm_noalloc"].reshape(k, ld)
        assert np.array_equal(Yn[:, n:], y0.reshape(k, ld)[:, n:])
        # and the product is the oracle's, column by column
        for j in range(min(k, 3)):
            ref = O.spmv(oa, x[j * ld:j * ld + n], y0[j * ld:j * ld + n], alpha, beta, op)
            assert np.allclose(Yn[j, :n], ref, rtol=1e-12, atol=1e-12)
        X.close()
        A.close()
    ctx.close()


def test_noalloc_spmm_rejects_operands_beyond_the_scratch_array():
    ctx = bx.Context(0)
    L = bx.lib()
    n, row, col, val = bx.laplacian3d(27, 42, 42, 42)  # 74,088 rows x 64 columns x 8 B x 2 > 64 MiB
    A = bx.CSR(ctx, n, n, row, col, val)
    X = bx.DeviceArray(ctx, (n * 64,))
    Y = bx.DeviceArray(ctx, (n * 64,))
    st = L.bcmgx_spmm_noalloc(ctx.h, A.h, X.ptr, Y.ptr, 64, n, 1.0, 0.0, 1, 0)
    assert st == abi.ERR_ARG
    assert b"scratch" in L.bcmgx_last_error(ctx.h)
    # 32 columns fit, and the call after the refused one works
    X.upload(np.ones(n * 64))
    st = L.bcmgx_spmm_noalloc(ctx.h, A.h, X.ptr, Y.ptr, 32, n, 1.0, 0.0, 1, 0)
    assert st == abi.OK and L.bcmgx_context_synchronize(ctx.h) == abi.OK
    y = Y.download()[: n * 32].reshape(32, n)
    rs = np.add.reduceat(val, row[:-1])  # row sums = A 1
    assert np.array_equal(y[0], y[31]) and np.allclose(y[0], rs, rtol=0, atol=1e-12)
    X.close()
    Y.close()
    A.close()
    ctx.close()
