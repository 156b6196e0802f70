"""HIP path vs the CPU oracle, through the C ABI (needs an MI355X).

Integer/index results (matching, aggregates, sparsity patterns) must be
bit-identical; setup values (matching weights, prolongator values, Galerkin
coarse operators) are computed in the oracle's operation order without FMA
contraction and are checked bit-exactly too (the stated bar is 1e-12
relative); solve-path values are checked to 1e-12/1e-10 relative and FCG
iteration counts to +-1."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import bootcmatchgx_amd as bx
import oracle_py as O
from bootcmatchgx_amd import abi

pytestmark = pytest.mark.gpu


def dev(ctx, ocsr):
    n, m, _ = ocsr.dims()
    r, c, v = ocsr.arrays()
    return bx.CSR(ctx, n, m, r, c, v)


def rand_spd(n, density, seed, ties=False):
    S = sp.random(n, n, density=density, random_state=seed)
    S = S + S.T
    S.data[:] = -(np.round(S.data * 3) / 3 if ties else S.data)
    S = S + sp.diags(np.asarray(abs(S).sum(1)).ravel() + 1.0)
    S = S.tocsr()
    S.sort_indices()
    return S


CASES = {
    "lap7_12": lambda: O.laplacian3d(7, 12, 12, 12),
    "lap7_9x7x5": lambda: O.laplacian3d(7, 9, 7, 5),
    "lap27_10": lambda: O.laplacian3d(27, 10, 10, 10),
    "lap7_block": lambda: O.laplacian3d(7, 2, 2, 5, 3, 4, 2),
    "diff_exp": lambda: O.diffusion7p(11, 10, 9, 0, 2.0, 0.0, 2),
    "diff_unif": lambda: O.diffusion7p(11, 10, 9, 1, 0.5, 2.0, 4),
    "rand_ties": lambda: O.Csr.from_scipy(rand_spd(500, 0.02, 5, ties=True)),
    "rand": lambda: O.Csr.from_scipy(rand_spd(500, 0.02, 6)),
}


def params(prec, **kw):
    p = bx.Params(solver="FCG", preconditioner=prec, dispnorm=0)
    for k, v in kw.items():
        p.set(k, v)
    return p


def oparams(p):
    q = abi.Params()
    C = __import__("ctypes")
    C.pointer(q)[0] = p.s
    return q


# ------------------------------------------------------------------- SpMV
@pytest.mark.parametrize("name", ["lap7_12", "lap27_10", "rand"])
@pytest.mark.parametrize("lanes", [0, 1, 4, 16, 64, -1, -2, -4, -16, -201, -204, -208, -216,
                                   -1004, -2204, -3002, -1216, -13204, -404, -5404])
def test_spmv_vs_oracle(ctx, name, lanes):
    oa = CASES[name]()
    A = dev(ctx, oa)
    rng = np.random.default_rng(3)
    x, y = rng.normal(size=A.m), rng.normal(size=A.n)
    for op, al, be in [(1, 1.0, 0.0), (2, -1.0, 1.0), (0, 0.7, -1.3)]:
        got = A.spmv(x, y, al, be, op, lanes)
        ref = O.spmv(oa, x, y, al, be, op)
        scale = np.abs(O.spmv(O.Csr.from_scipy(abs(oa.to_scipy())), np.abs(x))) + np.abs(y)
        assert np.all(np.abs(got - ref) <= 1e-14 * scale + 1e-300), (op, lanes)


TILE_CASES = {
    # 7 entries a row: tiles start at odd positions (the paired loads' first entry belongs to the previous tile)
    "lap7_18": lambda: O.laplacian3d(7, 18, 18, 18),
    "lap27_12": lambda: O.laplacian3d(27, 12, 12, 12),
    # 120 entries a row: every tile too long for the index (read by column index)
    "long_rows": lambda: O.Csr.from_scipy(rand_spd(300, 0.4, 7)),
    # about 24 random columns a row over 20000: more distinct columns than a tile indexes
    "scattered": lambda: O.Csr.from_scipy(rand_spd(20000, 0.0006, 9)),
    # indexed and unindexed tiles side by side, empty rows included
    "mixed": lambda: O.Csr.from_scipy(sp.vstack([sp.random(200, 900, density=0.01, random_state=10),
                                                 sp.random(100, 900, density=0.5, random_state=11),
                                                 sp.csr_matrix((30, 900)),
                                                 sp.random(170, 900, density=0.02, random_state=12)]).tocsr()),
}


@pytest.mark.parametrize("name", list(TILE_CASES))
def test_tiled_spmv_equals_stream_bits(ctx, name):
    """The row-tile kernel forms the same products and reduces each row in the same
    order as the paired-load stream kernel with 4 lanes a row: equal bits for every
    op, whether a tile is indexed or read by column index."""
    oa = TILE_CASES[name]()
    A = dev(ctx, oa)
    rng = np.random.default_rng(17)
    x, y = rng.normal(size=A.m), rng.normal(size=A.n)
    for op, al, be in [(1, 1.0, 0.0), (2, -1.0, 1.0), (0, 0.7, -1.3)]:
        ref = A.spmv(x, y, al, be, op, -204)
        for lanes in (-404, -5404, -7404):
            got = A.spmv(x, y, al, be, op, lanes)
            assert np.array_equal(got.view(np.int64), ref.view(np.int64)), (op, lanes)
    A.close()


SPMM_CASES = {
    "lap7_12": CASES["lap7_12"],
    "lap27_10": CASES["lap27_10"],
    "rand": CASES["rand"],
    "long_rows": lambda: O.Csr.from_scipy(rand_spd(300, 0.4, 7)),
    "rect_empty_rows": lambda: O.Csr.from_scipy(sp.random(260, 190, density=0.05, random_state=8, format="csr")),
}


@pytest.mark.parametrize("name", list(SPMM_CASES))
@pytest.mark.parametrize("k,kchunk", [(2, 0), (3, 0), (4, -1), (8, 0), (8, -2), (13, 4), (16, -1), (32, 0), (32, -1),
                                     (32, -2), (32, 8), (64, 0), (64, -1), (64, -2), (64, 16),
                                     (8, -16), (16, -18), (32, -17), (32, -22), (64, -20)])
def test_spmm_columns_bit_exact(ctx, name, k, kchunk):
    """Every column of the multi-vector kernel equals the oracle's sequential
    row sum (same order, no contraction) bit for bit."""
    oa = SPMM_CASES[name]()
    A = dev(ctx, oa)
    rng = np.random.default_rng(k * 7 + kchunk)
    X, Y = rng.normal(size=(A.m, k)), rng.normal(size=(A.n, k))
    for op, al, be in [(1, 1.0, 0.0), (2, -1.0, 1.0), (0, 0.7, -1.3)]:
        got = A.spmm(X, Y, al, be, op, kchunk)
        for c in range(k):
            ref = O.spmv(oa, X[:, c].copy(), Y[:, c].copy(), al, be, op)
            assert np.array_equal(got[:, c], ref), (op, c)


# --------------------------------------------------------------- matching
@pytest.mark.parametrize("name", list(CASES))
def test_matching_bit_exact(ctx, name):
    oa = CASES[name]()
    A = dev(ctx, oa)
    n = A.n
    w = np.random.default_rng(1).uniform(0.5, 2.0, n) if name.startswith("rand") else np.ones(n)
    assert np.array_equal(A.matching(w), O.suitor(oa, w))


def _zero_ah_matrix(n, seed):
    """Symmetric pattern, diagonal values in {0, 2, 4, 6} (explicit zeros kept), and a
    third of the off-diagonal pairs set to a_ij = (a_ii + a_jj) / 2, whose AH is
    exactly 0 for w = 1; pairs of zero diagonals have C_i + C_j = 0."""
    rng = np.random.default_rng(seed)
    m = 4 * n
    i = rng.integers(0, n, m)
    j = rng.integers(0, n, m)
    keep = i != j
    i, j = i[keep], j[keep]
    pairs = np.unique(np.stack([np.minimum(i, j), np.maximum(i, j)], 1), axis=0)
    d = rng.choice([0.0, 2.0, 4.0, 6.0], n)
    v = -rng.uniform(0.1, 3.0, len(pairs))
    z = rng.random(len(pairs)) < 1 / 3
    v[z] = (d[pairs[z, 0]] + d[pairs[z, 1]]) / 2
    r = np.concatenate([pairs[:, 0], pairs[:, 1], np.arange(n)])
    c = np.concatenate([pairs[:, 1], pairs[:, 0], np.arange(n)])
    x = np.concatenate([v, v, d])
    o = np.lexsort((c, r))
    r, c, x = r[o], c[o], x[o]
    row = np.searchsorted(r, np.arange(n + 1)).astype(np.int32)
    return O.Csr.from_arrays(n, n, row, c.astype(np.int32), x)


def test_matching_zero_ah_and_zero_diagonals(ctx):
    """Entries whose AH is exactly 0 weigh log(SUITOR_EPS / (0.999 DBL_EPSILON)), the
    least |AH| being 0 and clamped; pairs of zero diagonals take AH = DBL_EPSILON;
    with w = 1 (exact zeros) and a random w (none), equal to the oracle."""
    for seed in (1, 2):
        oa = _zero_ah_matrix(700, seed)
        A = dev(ctx, oa)
        for w in (np.ones(A.n), np.random.default_rng(seed).uniform(0.5, 2.0, A.n)):
            assert np.array_equal(A.matching(w), O.suitor(oa, w)), seed


@pytest.mark.parametrize("name", ["lap27_40", "lap7_64", "long_rows", "long_rows_ties"])
def test_matching_chains_and_long_rows(ctx, name):
    """Grids large enough for displacement chains to span several rounds, and
    rows longer than the proposal group (the preference list is scanned in
    several slices)."""
    oa = {"lap27_40": lambda: O.laplacian3d(27, 40, 40, 40),
          "lap7_64": lambda: O.laplacian3d(7, 64, 64, 64),
          "long_rows": SPMM_CASES["long_rows"],
          "long_rows_ties": lambda: O.Csr.from_scipy(rand_spd(400, 0.3, 9, ties=True))}[name]()
    A = dev(ctx, oa)
    for seed in (None, 3):
        w = np.ones(A.n) if seed is None else np.random.default_rng(seed).uniform(0.5, 2.0, A.n)
        assert np.array_equal(A.matching(w), O.suitor(oa, w)), seed


@pytest.mark.parametrize("name", ["bcsstk09", "bcsstk13", "bcsstk24"])
def test_matching_harwell_boeing(ctx, golden, name):
    """Structural-engineering SPD matrices of the reference's test data: untied weights."""
    oa = O.read_mtx(os.path.join(golden, name + ".mtx"))
    A = dev(ctx, oa)
    w = np.ones(A.n)
    assert np.array_equal(A.matching(w), O.suitor(oa, w))


@pytest.mark.parametrize("name", ["bcsstk13", "bcsstk24"])
def test_hierarchy_bit_exact_harwell_boeing(ctx, golden, name):
    """The whole BCMG setup on the reference's bcsstk13/24 inputs (their solution files
    hold zeros, so they pin the setup, not a solve): levels, aggregates, prolongator
    values and coarse operators equal the oracle's bit for bit."""
    oa = O.read_mtx(os.path.join(golden, name + ".mtx"))
    A = dev(ctx, oa)
    p = params("BCMG")
    P = bx.prec_setup(ctx, A, p)
    OP = O.Prec(oa, oparams(p))
    assert P.levels() == OP.levels() >= 2
    for l in range(P.levels()):
        r1, r2 = P.level_matrix(l), OP.level_A(l).to_scipy()
        assert np.array_equal(r1.indptr, r2.indptr) and np.array_equal(r1.indices, r2.indices)
        assert np.array_equal(r1.data, r2.data), l
        if l < P.levels() - 1:
            a1, p1 = P.level_prolongator(l)
            a2, p2 = OP.level_P(l)
            assert np.array_equal(a1, a2) and np.array_equal(p1, p2), l


def _dev_from_device(ctx, ocsr):
    """The oracle matrix copied into device buffers, then created from device memory."""
    n, m, _ = ocsr.dims()
    r, c, v = ocsr.arrays()
    bufs = []

    def put(a):
        raw = np.ascontiguousarray(a).tobytes()
        raw += b"\0" * (-len(raw) % 8)
        d = bx.DeviceArray(ctx, (max(len(raw) // 8, 1),))
        if raw:
            d.upload(np.frombuffer(raw, np.float64))
        bufs.append(d)
        return d.ptr

    A = bx.CSR.from_device(ctx, n, m, len(c), put(np.asarray(r, np.int32)), put(np.asarray(c, np.int32)),
                           put(np.asarray(v, np.float64)))
    for d in bufs:
        d.close()
    return A


def test_csr_from_device_memory(ctx):
    """A matrix created from device arrays gets its longest row and its one-entry-per-row
    shape measured on the device, so its Galerkin products take the same one-launch paths
    as a host-created one: the hierarchy and a prolongator-shaped product equal the
    oracle bit for bit."""
    oa = O.laplacian3d(27, 12, 12, 12)
    n = oa.dims()[0]
    A = _dev_from_device(ctx, oa)
    assert np.array_equal(A.to_host()[1], oa.arrays()[1])
    p = params("BCMG")
    P = bx.prec_setup(ctx, A, p)
    OP = O.Prec(oa, oparams(p))
    assert P.levels() == OP.levels() >= 2
    for l in range(P.levels()):
        r1, r2 = P.level_matrix(l), OP.level_A(l).to_scipy()
        assert np.array_equal(r1.indptr, r2.indptr) and np.array_equal(r1.indices, r2.indices)
        assert np.array_equal(r1.data, r2.data), l
    w = np.ones(n)
    agg, pv, nc = O.make_p(O.suitor(oa, w), w)
    oP = O.Csr.from_arrays(n, nc, np.arange(n + 1), agg, pv)
    _assert_same_csr(A.spgemm(_dev_from_device(ctx, oP)), O.spgemm(oa, oP))
    # n entries but not one per row (row 0 empty, row 1 holding two): the general path
    qcol = np.array(agg, np.int32).copy()
    qcol[:2] = np.sort(qcol[:2]) if qcol[0] != qcol[1] else [qcol[0], (qcol[0] + 1) % nc]
    qcol[:2] = np.sort(qcol[:2])
    oQ = O.Csr.from_arrays(n, nc, np.r_[0, 0, np.arange(2, n + 1)], qcol, pv)
    _assert_same_csr(A.spgemm(_dev_from_device(ctx, oQ)), O.spgemm(oa, oQ))


def test_make_prolongator_exact(ctx):
    for name in ["lap7_9x7x5", "rand_ties", "diff_exp"]:
        oa = CASES[name]()
        n = oa.dims()[0]
        w = np.random.default_rng(2).uniform(-1, 2, n)
        w[::17] = 0.0  # bad singletons
        M = O.suitor(oa, np.abs(w) + 0.5)
        a1, p1, c1 = bx.make_prolongator(ctx, M, w)
        a2, p2, c2 = O.make_p(M, w)
        assert c1 == c2 and np.array_equal(a1, a2) and np.array_equal(p1, p2)


# ------------------------------------------------------------------ SpGEMM
def _assert_same_csr(Cg, Co):
    r1, c1, v1 = Cg.to_host()
    r2, c2, v2 = Co.arrays()
    assert np.array_equal(r1, r2) and np.array_equal(c1, c2)
    assert np.array_equal(v1, v2), np.max(np.abs(v1 - v2) / (np.abs(v2) + 1e-300))


@pytest.mark.parametrize("name", ["lap7_12", "lap27_10", "rand", "diff_exp"])
def test_spgemm_galerkin_bit_exact(ctx, name):
    oa = CASES[name]()
    n = oa.dims()[0]
    w = np.ones(n)
    M = O.suitor(oa, w)
    agg, pv, nc = O.make_p(M, w)
    P = O.Csr.from_arrays(n, nc, np.arange(n + 1), agg, pv)
    R = O.transpose(P)
    A, Pd, Rd = dev(ctx, oa), dev(ctx, P), dev(ctx, R)
    AP = A.spgemm(Pd)
    _assert_same_csr(AP, O.spgemm(oa, P))
    _assert_same_csr(Rd.spgemm(AP), O.spgemm(R, O.spgemm(oa, P)))
    _assert_same_csr(Pd.transpose(), R)


@pytest.mark.parametrize("name", ["lap27_10", "long_rows", "rect_empty_rows", "mixed_lengths"])
def test_transpose_equals_scipy(ctx, name):
    """Transposed rows sorted by column: rows of at most 64 entries by the one-wave bitonic
    sort, longer ones (120 and more) by one lane; empty rows and rectangular shapes."""
    S = {"lap27_10": lambda: CASES["lap27_10"]().to_scipy(),
         "long_rows": lambda: rand_spd(300, 0.4, 7),
         "rect_empty_rows": lambda: sp.random(260, 190, density=0.05, random_state=8, format="csr"),
         "mixed_lengths": lambda: sp.hstack([sp.random(400, 30, density=0.5, random_state=13),
                                             sp.random(400, 300, density=0.02, random_state=14)]).tocsr()}[name]()
    S = sp.csr_matrix(S)
    S.sort_indices()
    T = bx.CSR.from_scipy(ctx, S).transpose()
    r, c, v = T.to_host()
    Ts = S.T.tocsr()
    Ts.sort_indices()
    assert np.array_equal(r, Ts.indptr) and np.array_equal(c, Ts.indices)
    assert np.array_equal(v, Ts.data)


def test_spgemm_medium_and_large_rows(ctx):
    # A*A of a 27-point stencil: up to 729 products per row (LDS bin)
    oa = O.laplacian3d(27, 9, 8, 7)
    A = dev(ctx, oa)
    _assert_same_csr(A.spgemm(A), O.spgemm(oa, oa))
    # a few dense rows: > 2048 products per row (global-memory bin)
    rng = np.random.default_rng(4)
    S = sp.random(300, 300, density=0.02, random_state=7, format="lil")
    S[5, :] = rng.normal(size=300)
    S[17, :] = rng.normal(size=300)
    S = S.tocsr()
    S.sort_indices()
    T = sp.random(300, 200, density=0.05, random_state=8, format="lil")
    T[:, 3] = rng.normal(size=(300, 1))
    T = T.tocsr()
    T.sort_indices()
    oS, oT = O.Csr.from_scipy(S), O.Csr.from_scipy(T)
    _assert_same_csr(dev(ctx, oS).spgemm(dev(ctx, oT)), O.spgemm(oS, oT))
    # empty rows and an empty product
    E = O.Csr.from_arrays(4, 4, np.array([0, 0, 1, 1, 1]), np.array([2]), np.array([3.0]))
    _assert_same_csr(dev(ctx, E).spgemm(dev(ctx, E)), O.spgemm(E, E))


def test_spgemm_unit_row_right_operand(ctx):
    """A prolongator-shaped right operand (row k holds entry k) takes the one-product-
    per-entry path: aggregates numbered out of row order (the products then need the
    sort), in row order (they arrive sorted), rows of A of 0 to 120 entries (the
    16-, 32- and 64-lane groups and the LDS bin), equal to the oracle bit for bit.
    Left operands whose longest row has at most 16 or 32 entries take the single
    launch over all rows at A's own entry positions; the product of that launch is
    itself a left operand of the next (its longest row is measured on the device)."""
    rng = np.random.default_rng(21)
    n, nc = 3000, 700
    for top in (30, 16, 32):
        lens = rng.integers(0, top + 1, n)
        if top == 30:
            lens[::97] = 120
        lens[::53] = 0
        lens[1::53] = 1
        lens[2::53] = 2
        arow = np.concatenate([[0], np.cumsum(lens)])
        acol = np.concatenate([np.sort(rng.choice(n, l, replace=False)) for l in lens]).astype(np.int32)
        oA = O.Csr.from_arrays(n, n, arow, acol, rng.normal(size=len(acol)))
        A = dev(ctx, oA)
        for agg in (rng.integers(0, nc, n), np.sort(rng.integers(0, nc, n))):
            P = O.Csr.from_arrays(n, nc, np.arange(n + 1), agg.astype(np.int32), rng.normal(size=n))
            C = A.spgemm(dev(ctx, P))
            oC = O.spgemm(oA, P)
            _assert_same_csr(C, oC)
            # the product (n x nc) times a prolongator of its columns
            Q = O.Csr.from_arrays(nc, 50, np.arange(nc + 1), rng.integers(0, 50, nc).astype(np.int32), rng.normal(size=nc))
            _assert_same_csr(C.spgemm(dev(ctx, Q)), O.spgemm(oC, Q))


def test_spgemm_two_entry_rows_merge(ctx):
    """Left-operand rows with at most two entries (the restriction's rows) are
    merged from two sorted B rows; both orders of k1 < k2 and k1 == k2, rows of
    0, 1 and 2 entries, B rows from empty to 300 entries, overlapping columns."""
    rng = np.random.default_rng(11)
    nb, m = 400, 350
    lens = rng.integers(0, 40, nb)
    lens[::37] = 300
    brow = np.concatenate([[0], np.cumsum(lens)])
    bcol = np.concatenate([np.sort(rng.choice(m, l, replace=False)) for l in lens]).astype(np.int32)
    bval = rng.normal(size=len(bcol))
    oB = O.Csr.from_arrays(nb, m, brow, bcol, bval)
    na = 500
    arow, acol = [0], []
    for i in range(na):
        t = i % 5
        ks = [] if t == 0 else [int(rng.integers(nb))] if t == 1 else sorted(rng.choice(nb, 2, replace=False).tolist())
        if t == 4:
            ks = [int(rng.integers(nb))] * 2  # the same B row twice
        acol += ks
        arow.append(len(acol))
    oA = O.Csr.from_arrays(na, nb, np.array(arow), np.array(acol, np.int32), rng.normal(size=len(acol)))
    _assert_same_csr(dev(ctx, oA).spgemm(dev(ctx, oB)), O.spgemm(oA, oB))


def test_spgemm_two_entry_rows_out_of_order(ctx):
    """The two-run merge falls back to the full sort when a run is out of order:
    B rows with shuffled and repeated columns, A rows listing the larger B row
    first.  Products of one column are summed in the order they are generated
    (A's entry order, then B's), as the oracle does, so the result is bit-exact."""
    rng = np.random.default_rng(12)
    nb, m = 300, 200
    lens = rng.integers(0, 30, nb)
    brow = np.concatenate([[0], np.cumsum(lens)])
    bcol = np.concatenate([rng.integers(0, m, l) for l in lens]).astype(np.int32)  # unsorted, repeats
    oB = O.Csr.from_arrays(nb, m, brow, bcol, rng.normal(size=len(bcol)))
    na = 400
    arow, acol = [0], []
    for i in range(na):
        t = i % 4
        ks = [] if t == 0 else [int(rng.integers(nb))] if t == 1 else sorted(rng.choice(nb, 2, replace=False).tolist())
        if t == 3:
            ks = ks[::-1]  # larger B row first
        acol += ks
        arow.append(len(acol))
    oA = O.Csr.from_arrays(na, nb, np.array(arow), np.array(acol, np.int32), rng.normal(size=len(acol)))
    _assert_same_csr(dev(ctx, oA).spgemm(dev(ctx, oB)), O.spgemm(oA, oB))


def test_spgemm_wide_right_operand(ctx):
    """A right operand with 2^26 or more columns takes the 64-bit (column, position)
    keys.  The expected product comes from the oracle on B with its columns renumbered
    densely in the same order (a monotone map leaves every sum's order unchanged),
    mapped back afterwards."""
    rng = np.random.default_rng(13)
    n, m = 2000, (1 << 26) + 7
    used = np.sort(rng.choice(m, 900, replace=False))
    used[-1] = m - 1
    lens = rng.integers(0, 12, n)
    lens[::41] = 40
    brow = np.concatenate([[0], np.cumsum(lens)])
    bloc = np.concatenate([np.sort(rng.choice(len(used), l, replace=False)) for l in lens]).astype(np.int32)
    bval = rng.normal(size=len(bloc))
    alens = rng.integers(0, 8, n)
    arow = np.concatenate([[0], np.cumsum(alens)])
    acol = np.concatenate([np.sort(rng.choice(n, l, replace=False)) for l in alens]).astype(np.int32)
    oA = O.Csr.from_arrays(n, n, arow, acol, rng.normal(size=len(acol)))
    ref = O.spgemm(oA, O.Csr.from_arrays(n, len(used), brow, bloc, bval))
    rr, rc, rv = ref.arrays()
    B = bx.CSR(ctx, n, m, brow, used[bloc].astype(np.int32), bval)
    C = dev(ctx, oA).spgemm(B)
    cr, cc, cv = C.to_host()
    assert (C.n, C.m) == (n, m)
    np.testing.assert_array_equal(cr, rr)
    np.testing.assert_array_equal(cc, used[rc])
    np.testing.assert_array_equal(cv, rv)


def test_kernel_timers(ctx):
    """The smoother and SpGEMM micro-benchmarks run their kernels and report sane figures."""
    oa = O.laplacian3d(27, 12, 12, 12)
    n = oa.dims()[0]
    A = dev(ctx, oa)
    for prec, k in ((0, 1), (1, 1), (1, 8), (0, 5)):
        assert A.time_smoother(prec, k, 3) > 0
    w = np.ones(n)
    agg, pv, nc = O.make_p(O.suitor(oa, w), w)
    P = dev(ctx, O.Csr.from_arrays(n, nc, np.arange(n + 1), agg, pv))
    ms, nzc = A.time_spgemm(P, 2)
    assert ms > 0 and nzc == A.spgemm(P).nnz
    with pytest.raises(bx.BcmgxError):
        P.time_spgemm(P, 1)  # inner dimensions differ


# --------------------------------------------------------------- hierarchy
HIER = [
    ("lap7_24", lambda: O.laplacian3d(7, 24, 24, 24), {}),
    ("lap27_18", lambda: O.laplacian3d(27, 18, 18, 18), {}),
    ("lap7_odd", lambda: O.laplacian3d(7, 17, 13, 11), {"max_coarse_size": 20}),
    ("diff_exp", lambda: O.diffusion7p(20, 18, 16, 0, 2.0, 0.0, 2), {"max_coarse_size": 50}),
    ("diff_unif", lambda: O.diffusion7p(20, 20, 20, 1, 0.5, 2.0, 4), {}),
    ("one_sweep", lambda: O.laplacian3d(7, 16, 16, 16), {"aggrsweeps": 0, "max_coarse_size": 30}),
]


@pytest.mark.parametrize("name,make,kw", HIER)
def test_hierarchy_bit_exact(ctx, name, make, kw):
    oa = make()
    A = dev(ctx, oa)
    p = params("BCMG", **kw)
    P = bx.prec_setup(ctx, A, p)
    OP = O.Prec(oa, oparams(p))
    assert P.levels() == OP.levels()
    for l in range(P.levels()):
        r1 = P.level_matrix(l)
        r2 = OP.level_A(l).to_scipy()
        assert np.array_equal(r1.indptr, r2.indptr) and np.array_equal(r1.indices, r2.indices)
        assert np.array_equal(r1.data, r2.data), (l, np.max(np.abs(r1.data - r2.data) / np.abs(r2.data)))
        if l < P.levels() - 1:
            a1, p1 = P.level_prolongator(l)
            a2, p2 = OP.level_P(l)
            assert np.array_equal(a1, a2) and np.array_equal(p1, p2)
    info = P.info()
    st = OP.stats()
    assert info["max_coarse_size"] == st["max_coarse_size"]
    assert abs(info["op_cmplx"] - st["op_cmplx"]) < 1e-12
    assert abs(info["avg_cratio"] - st["avg_cratio"]) < 1e-12


# ------------------------------------------------------------- whole solve
@pytest.mark.parametrize("pts", [7, 27])
@pytest.mark.parametrize("prec", ["BCMG", "AFSAI", "L1_JACOBI", "NONE"])
def test_solve_reference_fixture(ctx, golden, pts, prec):
    """launchAndCompare.mk criterion on the reference's 10^3 solutions."""
    oa = O.laplacian3d(pts, 10, 10, 10)
    A = dev(ctx, oa)
    p = params(prec)
    P = bx.prec_setup(ctx, A, p)
    x, out = bx.solve(ctx, A, np.ones(1000), np.zeros(1000), p, P)
    ref = np.loadtxt(os.path.join(golden, f"lap3d_{pts}p_10x10x10_1x1x1_sol.txt"))
    assert out["retv"] == 0
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) <= 1e-5
    _, oi = O.fcg(oa, O.Prec(oa, oparams(p)), oparams(p), np.ones(1000))
    assert abs(out["niter"] - oi["niter"]) <= 1


@pytest.mark.parametrize("prec", ["BCMG", "AFSAI", "L1_JACOBI"])
def test_solve_bcsstk09(ctx, golden, prec):
    oa = O.read_mtx(os.path.join(golden, "bcsstk09.mtx"))
    A = dev(ctx, oa)
    p = params(prec, itnlim=5000)
    x, out = bx.solve(ctx, A, np.ones(A.n), None, p, bx.prec_setup(ctx, A, p))
    ref = np.loadtxt(os.path.join(golden, "bcsstk09_sol.txt"))
    assert out["retv"] == 0 and np.linalg.norm(x - ref) / np.linalg.norm(ref) <= 1e-5


def test_c0_iteration_count_and_residual(ctx):
    """Config C0: 7-point 50^3, b = A x*, rtol 1e-8, 1/1 sweeps."""
    oa = O.laplacian3d(7, 50, 50, 50)
    n = 125000
    xs = np.random.default_rng(0).uniform(-1, 1, n)
    b = O.spmv(oa, xs)
    A = dev(ctx, oa)
    p = params("BCMG", rtol=1e-8, prerelax_sweeps=1, postrelax_sweeps=1)
    P = bx.prec_setup(ctx, A, p)
    x, out = bx.solve(ctx, A, b, None, p, P)
    _, oi = O.fcg(oa, O.Prec(oa, oparams(p)), oparams(p), b)
    assert abs(out["niter"] - oi["niter"]) <= 1
    r = b - O.spmv(oa, x)
    assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(b) * 1.0000001
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-6


def test_prec_apply_matches_oracle(ctx):
    oa = O.laplacian3d(27, 14, 14, 14)
    A = dev(ctx, oa)
    r = np.random.default_rng(5).normal(size=A.n)
    for prec in ["BCMG", "AFSAI", "L1_JACOBI", "NONE"]:
        p = params(prec, prerelax_sweeps=2, postrelax_sweeps=1)
        got = bx.prec_setup(ctx, A, p).apply(r)
        ref = O.Prec(oa, oparams(p)).apply(r)
        assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref), prec


def test_batched_prec_apply_large_level_matches_oracle(ctx):
    """A level above 65536 rows (autotuned persistent SpMM kernels): every
    column of the batched cycle matches the oracle's one-column cycle."""
    oa = O.laplacian3d(27, 48, 48, 48)
    A = dev(ctx, oa)
    R = np.random.default_rng(15).normal(size=(A.n, 8))
    p = params("BCMG", prerelax_sweeps=1, postrelax_sweeps=1)
    P = bx.prec_setup(ctx, A, p)
    got = P.apply(R)
    op = O.Prec(oa, oparams(p))
    for j in range(8):
        ref = op.apply(R[:, j].copy())
        assert np.linalg.norm(got[:, j] - ref) <= 1e-12 * np.linalg.norm(ref), j
    P.finalize()


@pytest.mark.parametrize("cycle", ["W_CYCLE", "H_CYCLE", "VARIABLE_V_CYCLE"])
def test_other_cycles_match_oracle(ctx, cycle):
    oa = O.laplacian3d(7, 20, 20, 20)
    A = dev(ctx, oa)
    p = params("BCMG", cycle_type=cycle, max_coarse_size=40, aggrsweeps=0)
    r = np.random.default_rng(6).normal(size=A.n)
    got = bx.prec_setup(ctx, A, p).apply(r)
    ref = O.Prec(oa, oparams(p)).apply(r)
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)


def test_afsai_factor_vs_oracle(ctx, golden):
    """The stencils' rows take the column-table gradient; the Harwell-Boeing matrix and the
    random one (rows with more than 32 entries left of the diagonal, or more than 192 columns
    in a row's table) go to the merge kernel, partly or wholly."""
    cases = [O.laplacian3d(7, 9, 8, 7), O.laplacian3d(27, 7, 7, 7), O.diffusion7p(8, 8, 8, 1, 0.5, 2.0, 4),
             O.read_mtx(os.path.join(golden, "bcsstk09.mtx")), O.Csr.from_scipy(rand_spd(900, 0.027, 7))]
    for oa in cases:
        A = dev(ctx, oa)
        p = params("AFSAI")
        G = bx.prec_setup(ctx, A, p).afsai_factor()
        Go = O.Prec(oa, oparams(p)).afsai_G().to_scipy()
        assert np.array_equal(G.indptr, Go.indptr) and np.array_equal(G.indices, Go.indices)
        assert np.max(np.abs(G.data - Go.data) / np.abs(Go.data)) <= 1e-12


# ------------------------------------------------------------- extensions
def test_fp32_hierarchy_converges_like_fp64(ctx):
    oa = O.laplacian3d(27, 24, 24, 24)
    n = oa.dims()[0]
    xs = np.random.default_rng(1).uniform(-1, 1, n)
    b = O.spmv(oa, xs)
    A = dev(ctx, oa)
    p64 = params("BCMG", rtol=1e-8, prerelax_sweeps=1, postrelax_sweeps=1)
    p32 = params("BCMG", rtol=1e-8, prerelax_sweeps=1, postrelax_sweeps=1, precision="FP32")
    x64, o64 = bx.solve(ctx, A, b, None, p64, bx.prec_setup(ctx, A, p64))
    x32, o32 = bx.solve(ctx, A, b, None, p32, bx.prec_setup(ctx, A, p32))
    _, oi = O.fcg(oa, O.Prec(oa, oparams(p32)), oparams(p32), b)
    assert o32["retv"] == 0 and abs(o32["niter"] - o64["niter"]) <= 1 and abs(o32["niter"] - oi["niter"]) <= 1
    r = b - O.spmv(oa, x32)
    assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(b) * 1.0000001
    assert np.linalg.norm(x32 - x64) / np.linalg.norm(x64) < 1e-6


@pytest.mark.parametrize("precision", ["FP64", "FP32"])
def test_batched_columns_match_single_solves(ctx, precision):
    oa = O.laplacian3d(7, 30, 30, 30)
    n = oa.dims()[0]
    k = 5
    B = np.random.default_rng(3).normal(size=(n, k))
    B[:, 4] *= 1e-3  # different scales converge independently
    A = dev(ctx, oa)
    p = params("BCMG", rtol=1e-8, prerelax_sweeps=1, postrelax_sweeps=1, precision=precision)
    P = bx.prec_setup(ctx, A, p)
    X, out = bx.solve(ctx, A, B, None, p, P)
    for j in range(k):
        xj, oj = bx.solve(ctx, A, B[:, j], None, p, P)
        assert abs(out["niter_cols"][j] - oj["niter"]) <= 1
        assert np.linalg.norm(X[:, j] - xj) <= 1e-6 * np.linalg.norm(xj)
        r = B[:, j] - O.spmv(oa, X[:, j])
        assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(B[:, j]) * 1.0000001


@pytest.mark.parametrize("precision", ["FP64", "FP32"])
def test_solves_after_wider_batch_reuse_no_stale_graph(ctx, precision):
    """A single-column solve, then a wider batch on the same preconditioner
    (its cycle buffers grow), then the single-column solve again: the replayed
    iteration graph must use the new buffers and give the same bits."""
    oa = O.laplacian3d(27, 20, 20, 20)
    n = oa.dims()[0]
    A = dev(ctx, oa)
    p = params("BCMG", rtol=1e-8, prerelax_sweeps=1, postrelax_sweeps=1, precision=precision)
    P = bx.prec_setup(ctx, A, p)
    b = np.random.default_rng(11).normal(size=n)
    x1, o1 = bx.solve(ctx, A, b, None, p, P)
    B = np.random.default_rng(12).normal(size=(n, 8))
    X8, o8 = bx.solve(ctx, A, B, None, p, P)
    x2, o2 = bx.solve(ctx, A, b, None, p, P)
    assert o1["niter"] == o2["niter"]
    assert np.array_equal(x1, x2)
    for j in range(8):
        r = B[:, j] - O.spmv(oa, X8[:, j])
        assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(B[:, j]) * 1.0000001
    P.finalize()


def test_graph_replay_equals_eager(ctx):
    oa = O.laplacian3d(27, 16, 16, 16)
    b = O.spmv(oa, np.random.default_rng(9).uniform(-1, 1, 4096))
    A = dev(ctx, oa)
    res = []
    for g in (1, 0, 1):
        p = params("BCMG", rtol=1e-8, prerelax_sweeps=1, postrelax_sweeps=1, use_graph=g)
        P = bx.prec_setup(ctx, A, p)
        x1, o1 = bx.solve(ctx, A, b, None, p, P)
        x2, o2 = bx.solve(ctx, A, b, None, p, P)  # the graph is replayed for a second solve
        assert o1["niter"] == o2["niter"] and np.array_equal(x1, x2)
        res.append((x1, o1["niter"]))
    assert res[0][1] == res[1][1] == res[2][1]
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][0], res[2][0])


@pytest.mark.parametrize("precision", ["FP64", "FP32"])
def test_cghs_graph_replay_and_solver_switch(ctx, precision):
    """CG from its replayed iteration graph gives the eager bits; FCG and CG alternate on one
    preconditioner and workspace without disturbing each other."""
    oa = O.laplacian3d(27, 16, 16, 16)
    b = O.spmv(oa, np.random.default_rng(21).uniform(-1, 1, 4096))
    A = dev(ctx, oa)
    res = []
    for g in (1, 0):
        p = params("BCMG", solver="CGHS", rtol=1e-8, prerelax_sweeps=1, postrelax_sweeps=1, precision=precision,
                   use_graph=g)
        P = bx.prec_setup(ctx, A, p)
        x1, o1 = bx.solve(ctx, A, b, None, p, P)
        x2, o2 = bx.solve(ctx, A, b, None, p, P)
        assert o1["retv"] == 0 and o1["niter"] == o2["niter"] and np.array_equal(x1, x2)
        xf, of = bx.solve(ctx, A, b, None, p.copy().set("solver", "FCG"), P)
        x3, o3 = bx.solve(ctx, A, b, None, p, P)
        assert np.array_equal(x1, x3)
        for x in (x1, xf):
            assert np.linalg.norm(b - O.spmv(oa, x)) <= 1e-8 * np.linalg.norm(b) * 1.0000001
        res.append((x1, o1["niter"]))
        P.finalize()
    assert res[0][1] == res[1][1] and np.array_equal(res[0][0], res[1][0])


@pytest.mark.parametrize("itnlim", [2, 3, 4, 5])
@pytest.mark.parametrize("k", [1, 4])
def test_fcg_stop_after_either_iteration_of_a_pair(ctx, itnlim, k):
    """x takes its FCG step every second iteration; stopping after the first iteration of a pair
    (odd itnlim) must still add the waiting step: x equals the oracle's sequential updates."""
    oa = O.laplacian3d(7, 14, 12, 10)
    n = oa.dims()[0]
    A = dev(ctx, oa)
    p = params("BCMG", rtol=1e-30, itnlim=itnlim)
    P = bx.prec_setup(ctx, A, p)
    B = np.random.default_rng(23).normal(size=(n, k))
    X, out = bx.solve(ctx, A, B, None, p, P)
    X = X.reshape(n, k)
    OP = O.Prec(oa, oparams(p))
    for j in range(k):
        xo, io = O.fcg(oa, OP, oparams(p), B[:, j])
        assert out["retv"] == io["retv"] == -1
        assert np.linalg.norm(X[:, j] - xo) / np.linalg.norm(xo) < 1e-12
    P.finalize()


def test_missing_diagonal_is_reported(ctx):
    """The matching needs a diagonal entry in every row (matching.cu:155): setup names the row."""
    S = sp.diags([np.full(40, 4.0), np.full(39, -1.0), np.full(39, -1.0)], [0, -1, 1], format="lil")
    S[7, 7] = 0.0
    S = S.tocsr()
    S.eliminate_zeros()
    S.sort_indices()
    A = dev(ctx, O.Csr.from_scipy(S))
    with pytest.raises(bx.BcmgxError) as e:
        bx.prec_setup(ctx, A, params("BCMG"))
    assert "row 7 has no diagonal" in str(e.value)


@pytest.mark.parametrize("prec,k,guess", [("FP32", 32, "zero"), ("FP32", 32, "rand"), ("FP64", 5, "zero"),
                                           ("FP64", 2, "rand"), ("FP32", 12, "neg0"), ("FP64", 64, "zero"),
                                           ("FP32", 64, "rand"), ("FP32", 3, "zero")])
def test_fcg_device_blocks_match_host_path(ctx, prec, k, guess):
    """FCG on column-major device blocks stages them straight into its buffers
    (tiled layout change, the residual of a zero guess taken as b, the waiting x
    step added on the way out); it gives the host path's bits, leaves padding
    rows alone and counts the same iterations."""
    oa = O.laplacian3d(27, 14, 13, 12)
    A = dev(ctx, oa)
    n = A.n
    ld = n + 3
    rng = np.random.default_rng(k)
    B = rng.normal(size=(n, k))
    X0 = {"zero": np.zeros((n, k)), "rand": rng.normal(size=(n, k)) * 0.1, "neg0": np.full((n, k), -0.0)}[guess]
    p = params("BCMG", precision=prec, rtol=1e-8)
    P = bx.prec_setup(ctx, A, p)
    xh, oh = bx.solve(ctx, A, B, X0, p, P)
    pad = 7.25
    Bd, Xd = bx.DeviceArray(ctx, (ld, k)), bx.DeviceArray(ctx, (ld, k))
    bb = np.full((ld, k), pad)
    bb[:n] = B
    xx = np.full((ld, k), pad)
    xx[:n] = X0
    Bd.upload(bb.T.ravel())
    Xd.upload(xx.T.ravel())
    od = bx.solve_device(ctx, A, Bd.ptr, Xd.ptr, k, p, P, ld=ld)
    xd = Xd.download().reshape(k, ld).T
    assert od["niter"] == oh["niter"]
    assert np.array_equal(xd[:n], np.asarray(xh).reshape(n, k))
    assert np.all(xd[n:] == pad)
    P.finalize()


@pytest.mark.parametrize("k", [4, 8, 32, 12, 64])
def test_fcg_fp32_z_matches_widened_path(ctx, k):
    """With an FP32 hierarchy FCG keeps the cycle's result in FP32 (fcg_z32) and widens it inside
    its kernels: the same operations as the widened-FP64 path, the dot products summed in other
    (fixed) groups; iteration counts within 1, solutions within 1e-10, both meet the tolerance."""
    oa = O.laplacian3d(27, 20, 20, 20)
    n = oa.dims()[0]
    A = dev(ctx, oa)
    B = np.random.default_rng(29).normal(size=(n, k))
    res = {}
    for z in (1, 0):
        p = params("BCMG", rtol=1e-8, prerelax_sweeps=1, postrelax_sweeps=1, precision="FP32", fcg_z32=z)
        P = bx.prec_setup(ctx, A, p)
        X, out = bx.solve(ctx, A, B, None, p, P)
        X2, out2 = bx.solve(ctx, A, B, None, p, P)  # the replayed graph
        assert np.array_equal(X, X2)
        res[z] = (X, out)
        P.finalize()
    (X1, o1), (X0, o0) = res[1], res[0]
    assert o1["retv"] == o0["retv"] == 0
    for j in range(k):
        assert abs(o1["niter_cols"][j] - o0["niter_cols"][j]) <= 1
        assert np.linalg.norm(X1[:, j] - X0[:, j]) <= 1e-10 * np.linalg.norm(X0[:, j])
        r = B[:, j] - O.spmv(oa, X1[:, j])
        assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(B[:, j]) * 1.0000001


def test_itnlim_and_info_file(ctx, tmp_path):
    oa = O.laplacian3d(7, 12, 12, 12)
    A = dev(ctx, oa)
    p = params("BCMG", rtol=1e-30, itnlim=5)
    P = bx.prec_setup(ctx, A, p)
    x, out = bx.solve(ctx, A, np.ones(A.n), None, p, P)
    assert out["retv"] == -1 and out["niter"] == 6
    f = tmp_path / "info.txt"
    bx.dump_info(str(f), p, P, out)
    txt = f.read_text()
    assert "*** Exceeded max iter ***" in txt and "Total iterations" in txt and "Number of levels" in txt


def test_sparse_outer_grad(ctx):
    oa = O.laplacian3d(7, 6, 5, 4)
    A = dev(ctx, oa)
    rng = np.random.default_rng(2)
    lam, x = rng.normal(size=(A.n, 3)), rng.normal(size=(A.n, 3))
    g = A.sparse_outer_grad(lam, x)
    r, c, _ = oa.arrays()
    rows = np.repeat(np.arange(A.n), np.diff(r))
    assert np.allclose(g, -np.sum(lam[rows] * x[c], axis=1), rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("prec", ["NONE", "L1_JACOBI", "BCMG", "AFSAI"])
@pytest.mark.parametrize("name", ["lap7_12", "lap27_10", "rand"])
def test_cghs_matches_oracle(ctx, name, prec):
    """CG (CG_HS.cu): iteration count within 1 of the oracle, solution to 1e-9."""
    oa = CASES[name]()
    n = oa.dims()[0]
    A = dev(ctx, oa)
    p = params(prec, solver="CGHS", rtol=1e-10, itnlim=3000)
    b = np.random.default_rng(11).normal(size=n)
    xo, io = O.cghs(oa, O.Prec(oa, oparams(p)), oparams(p), b)
    P = bx.prec_setup(ctx, A, p)
    x, out = bx.solve(ctx, A, b, None, p, P)
    P.finalize()
    assert out["retv"] == io["retv"] == 0
    assert abs(out["niter"] - io["niter"]) <= 1
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-9


def test_cghs_batched_columns_and_stop_rules(ctx):
    """k columns in one batched CG equal one-at-a-time solves; itnlim and the absolute test."""
    oa = O.laplacian3d(27, 14, 13, 12)
    n = oa.dims()[0]
    A = dev(ctx, oa)
    p = params("BCMG", solver="CGHS", rtol=1e-9)
    P = bx.prec_setup(ctx, A, p)
    B = np.random.default_rng(5).normal(size=(n, 4))
    B[:, 2] *= 1e-3  # scaled: the relative test gives the same count, the absolute one does not
    X, out = bx.solve(ctx, A, B, None, p, P)
    OP = O.Prec(oa, oparams(p))
    for j in range(4):
        xj, oj = bx.solve(ctx, A, B[:, j].copy(), None, p, P)
        # the dot products of a k-column batch sum in another (fixed) order than one column's
        assert np.linalg.norm(X[:, j] - xj) / np.linalg.norm(xj) < 1e-10
        assert abs(out["niter_cols"][j] - oj["niter"]) <= 1
        xo, io = O.cghs(oa, OP, oparams(p), B[:, j])
        assert abs(oj["niter"] - io["niter"]) <= 1
    q = p.copy().set("rtol", 1e-30).set("itnlim", 3)
    _, o3 = bx.solve(ctx, A, B[:, 0].copy(), None, q, P)
    assert o3["retv"] == -1 and o3["niter"] == 3
    q = p.copy().set("rtol", 1e-6).set("stop_criterion", 0)
    x0, o0 = bx.solve(ctx, A, B[:, 1].copy(), None, q, P)
    assert o0["retv"] == 0 and o0["del0"] == 1.0 and o0["exit_res"] < 1e-6
    P.finalize()


@pytest.mark.parametrize("prec", ["NONE", "L1_JACOBI", "BCMG", "AFSAI"])
@pytest.mark.parametrize("s", [1, 2, 3])
def test_sstep_cg_matches_oracle(ctx, prec, s):
    """s-step CG (CGs.cu): outer iteration count within 1 of the oracle, solution to 1e-8."""
    oa = CASES["lap27_10"]() if prec != "NONE" else CASES["lap7_12"]()
    n = oa.dims()[0]
    A = dev(ctx, oa)
    p = params(prec, solver="CGS", sstep=s, rtol=1e-9, itnlim=3000)
    b = np.random.default_rng(13).normal(size=n)
    xo, io = O.cgs(oa, O.Prec(oa, oparams(p)), oparams(p), b)
    P = bx.prec_setup(ctx, A, p)
    x, out = bx.solve(ctx, A, b, None, p, P)
    P.finalize()
    assert out["retv"] == io["retv"] == 0
    assert abs(out["niter"] - io["niter"]) <= 1
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-8


def test_sstep_cg_options_columns_and_breakdown(ctx):
    oa = O.laplacian3d(7, 11, 10, 9)
    n = oa.dims()[0]
    A = dev(ctx, oa)
    B = np.random.default_rng(17).normal(size=(n, 3))
    for kw in (dict(ru_res=0), dict(rec_res_int=3), dict()):
        p = params("L1_JACOBI", solver="CGS", sstep=2, rtol=1e-9, **kw)
        P = bx.prec_setup(ctx, A, p)
        X, out = bx.solve(ctx, A, B, None, p, P)
        for j in range(3):
            xo, io = O.cgs(oa, O.Prec(oa, oparams(p)), oparams(p), B[:, j])
            assert abs(out["niter_cols"][j] - io["niter"]) <= 1
            assert np.linalg.norm(X[:, j] - xo) / np.linalg.norm(xo) < 1e-8
        P.finalize()
    # a negative definite operator: the s x s Cholesky fails and retv > 0 (scalarWorkMO)
    N = O.Csr.from_arrays(50, 50, np.arange(51), np.arange(50), -np.linspace(1, 2, 50))
    Nd = dev(ctx, N)
    p = params("NONE", solver="CGS", sstep=1)
    _, out = bx.solve(ctx, Nd, np.ones(50), None, p, None)
    assert out["retv"] > 0


@pytest.mark.parametrize("solver", ["PIPELINED_CGS", "CGS_CUBLAS"])
@pytest.mark.parametrize("prec", ["NONE", "L1_JACOBI", "BCMG", "AFSAI"])
@pytest.mark.parametrize("s", [1, 3])
def test_pipelined_and_cublas_sstep_match_oracle(ctx, solver, prec, s):
    """pipelinedCGs.cu against the pipelined oracle; CGscublas.cu (the s-step method on BLAS
    calls) against the s-step oracle."""
    oa = CASES["lap27_10"]() if prec != "NONE" else CASES["lap7_12"]()
    n = oa.dims()[0]
    A = dev(ctx, oa)
    p = params(prec, solver=solver, sstep=s, rtol=1e-9, itnlim=3000)
    b = np.random.default_rng(19).normal(size=n)
    ref = O.pipecgs if solver == "PIPELINED_CGS" else O.cgs
    q = oparams(p)
    if solver == "CGS_CUBLAS":
        q.solver = abi.SOLVER["CGS"]
    xo, io = ref(oa, O.Prec(oa, q), q, b)
    P = bx.prec_setup(ctx, A, p)
    x, out = bx.solve(ctx, A, b, None, p, P)
    P.finalize()
    assert out["retv"] == io["retv"] == 0
    assert abs(out["niter"] - io["niter"]) <= 1
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-8


def test_unknown_solver_is_rejected(ctx):
    """An unknown solver name parses to INVALID and solve() refuses it (solve.cu:42-44)."""
    oa = CASES["lap7_12"]()
    A = dev(ctx, oa)
    p = params("BCMG")
    P = bx.prec_setup(ctx, A, p)
    p.set("solver", "GMRES")
    assert p.solver == abi.SOLVER["INVALID"]
    with pytest.raises(bx.BcmgxError) as e:
        bx.solve(ctx, A, np.ones(A.n), None, p, P)
    assert e.value.status == -6 and "SolverType" in str(e.value)
    P.finalize()
