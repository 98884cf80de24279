"""The backend contract (tests/test_backend.cpp of the reference) run through
the C ABI on the GPU, every answer compared with the oracle on the same
inputs.

Bars: elementwise steps and every row sum that fits one LDS tile are
bit-exact (same operation order, no FMA contraction). Rows longer than a tile
and all reductions are summed in a fixed tree, so they match the serial CPU
sums to the tolerances the reference's own cross-check allows
(test_backend.cpp:865-892): SpMV 1e-13, dot 1e-11, two-norm 1e-12, step
terms 1e-10, convergence sums 1e-11 (relative to max(1, |a|, |b|)); maxima
exact.
"""
import numpy as np
import pytest

from anukulan_amd import Matrix

pytestmark = pytest.mark.gpu

INF = np.inf


def rel(a, b):
    return abs(a - b) / max(1.0, abs(a), abs(b))


def csr_random(rows, cols, lens, rng, lo=-3.0, hi=3.0):
    ri, ci = [], []
    for i, L in enumerate(lens):
        for k in range(L):
            ri.append(i)
            ci.append((i * 31 + k * 17) % cols)
    v = rng.uniform(lo, hi, len(ri))
    import _oracle
    s, idx, val = _oracle.csr_from_triplets(rows, cols, ri, ci, v)
    return Matrix(rows, cols, s, idx, val)


def test_allocation_starts_at_zero(ctx):
    v = ctx.allocate(4096)
    assert np.all(v.download() == 0.0)
    v.free()


def test_upload_download_fill_copy(ctx):
    host = 0.5 * np.arange(1000) - 13.0
    v = ctx.vector(host)
    assert np.array_equal(v.download(), host)
    w = ctx.allocate(1000)
    ctx.copy(v, w)
    assert np.array_equal(w.download(), host)
    ctx.fill(v, -2.75)
    assert np.all(v.download() == -2.75)
    ctx.fill(v, 0.0)
    assert np.all(v.download() == 0.0)


def test_multiply_matches_longhand_and_oracle(ctx, orc):
    import _oracle
    s, i, v = _oracle.csr_from_triplets(4, 4, [0, 0, 1, 2, 2, 3], [0, 2, 1, 0, 3, 2],
                                        [1.5, -2.0, 3.25, 4.0, -0.5, 7.0])
    a = Matrix(4, 4, s, i, v)
    at = a.transpose()
    x = np.array([1.0, -2.0, 3.0, 4.0])
    dx, dy = ctx.vector(x), ctx.allocate(4)
    ctx.multiply(a, dx, dy)
    dense = np.zeros((4, 4))
    for r in range(4):
        for k in range(s[r], s[r + 1]):
            dense[r, i[k]] = v[k]
    assert np.array_equal(dy.download(), _oracle.spmv(s, i, v, x))
    assert np.allclose(dy.download(), dense @ x, rtol=0, atol=1e-15)
    dt = ctx.allocate(4)
    ctx.multiply_transpose(at, dx, dt)
    assert np.array_equal(dt.download(), _oracle.spmv(at.start, at.index, at.value, x))
    assert np.allclose(dt.download(), dense.T @ x, rtol=0, atol=1e-15)


@pytest.mark.parametrize("mean", [1, 3, 6, 12, 40])
def test_multiply_across_row_length_regimes(ctx, orc, mean):
    # test_backend.cpp:151-187: 733 ragged rows, every 5th empty, every 100th long,
    # output poisoned first. All rows fit a tile here, so the sums are bit-exact.
    import _oracle
    rng = np.random.default_rng(20260830 + mean)
    rows, cols = 733, 401
    lens = []
    for i in range(rows):
        L = mean
        if i % 5 == 0:
            L = 0
        if i % 100 == 7:
            L = 4 * mean + 37
        lens.append(L)
    a = csr_random(rows, cols, lens, rng)
    x = rng.uniform(-3, 3, cols)
    dx, dy = ctx.vector(x), ctx.allocate(rows)
    ctx.fill(dy, -12345.0)
    ctx.multiply(a, dx, dy)
    assert np.array_equal(dy.download(), _oracle.spmv(a.start, a.index, a.value, x))


@pytest.mark.parametrize("long_len", [1025, 4999, 20000])
def test_rows_longer_than_a_tile(ctx, orc, long_len):
    # Zipf tails: rows past one LDS tile go to a whole workgroup and a tree sum.
    import _oracle
    rng = np.random.default_rng(long_len)
    rows, cols = 300, 30011
    lens = [long_len if i in (3, 150, 299) else (i % 13) for i in range(rows)]
    ri, ci = [], []
    for i, L in enumerate(lens):
        cc = rng.choice(cols, size=L, replace=False)
        ri += [i] * L
        ci += list(cc)
    v = rng.uniform(-2, 2, len(ri))
    s, idx, val = _oracle.csr_from_triplets(rows, cols, ri, ci, v)
    a = Matrix(rows, cols, s, idx, val)
    x = rng.uniform(-1, 1, cols)
    dx, dy = ctx.vector(x), ctx.allocate(rows)
    ctx.fill(dy, -12345.0)
    ctx.multiply(a, dx, dy)
    got, want = dy.download(), _oracle.spmv(s, idx, val, x)
    short = np.array([L <= 1024 for L in lens])
    assert np.array_equal(got[short], want[short])
    for r in np.nonzero(~short)[0]:
        assert rel(got[r], want[r]) <= 1e-13


def test_a_second_matrix_under_a_new_identity(ctx, orc):
    # test_backend.cpp:194-225: a cache keyed by identity, not by address.
    import _oracle
    for rnd in range(2):
        ri, ci, v = [], [], []
        for i in range(40):
            for k in range(rnd + 1):
                ri.append(i)
                ci.append((i + 3 * k) % 12)
                v.append(1.0 + rnd + 0.5 * i)
        s, idx, val = _oracle.csr_from_triplets(40, 12, ri, ci, v)
        a = Matrix(40, 12, s, idx, val)
        x = 1.0 + 0.25 * np.arange(12)
        dx, dy = ctx.vector(x), ctx.allocate(40)
        ctx.fill(dy, -12345.0)
        ctx.multiply(a, dx, dy)
        assert np.array_equal(dy.download(), _oracle.spmv(s, idx, val, x))


def test_prepare_and_release(ctx, orc):
    import _oracle
    s, i, v = _oracle.csr_from_triplets(4, 4, [0, 0, 1, 2, 2, 3], [0, 2, 1, 0, 3, 2],
                                        [1.5, -2.0, 3.25, 4.0, -0.5, 7.0])
    a = Matrix(4, 4, s, i, v)
    o = Matrix(4, 4, [0, 1, 2, 3, 4], [0, 1, 2, 3], [2.0, 2.0, 2.0, 2.0])
    x = np.array([1.0, 2.0, 3.0, 4.0])
    dx, y1, y2 = ctx.vector(x), ctx.allocate(4), ctx.allocate(4)
    ctx.prepare(a)
    ctx.prepare(o)
    ctx.prepare(a)
    ctx.multiply(a, dx, y1)
    ctx.multiply(o, dx, y2)
    assert np.array_equal(y1.download(), _oracle.spmv(s, i, v, x))
    assert np.array_equal(y2.download(), 2.0 * x)
    ctx.release()
    ctx.multiply(a, dx, y1)
    ctx.multiply(o, dx, y2)
    assert np.array_equal(y1.download(), _oracle.spmv(s, i, v, x))
    assert np.array_equal(y2.download(), 2.0 * x)
    ctx.release()


@pytest.mark.parametrize("n", [0, 5, 1000, 200003])
def test_reductions(ctx, orc, n):
    import _oracle
    rng = np.random.default_rng(7)
    a, b = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    if n > 12345:
        a[12345] = -9.5
    if n == 0:
        da = db = ctx.allocate(0)
    else:
        da, db = ctx.vector(a), ctx.vector(b)
    assert rel(ctx.dot(da, db), _oracle.dot(a, b)) <= 1e-11
    assert rel(ctx.two_norm(da), _oracle.two_norm(a)) <= 1e-12
    assert ctx.inf_norm(da) == _oracle.inf_norm(a)


def test_primal_update_is_the_oracle(ctx, orc):
    # test_backend.cpp:276-325, all four shapes the solver runs.
    import _oracle
    x = np.array([0.0, 1.0, -1.0, 5.0, 2.0])
    c = np.array([1.0, -1.0, 0.5, 2.0, 0.0])
    kty = np.array([0.5, 0.5, -0.5, 1.0, 3.0])
    lo = np.array([0.0, -INF, -2.0, 0.0, 1.0])
    hi = np.array([1.0, 1.0, INF, 4.0, 1.5])
    anc = np.array([0.25, -0.5, 0.0, 3.0, 1.25])
    for gamma, h, blend in [(0.0, 1.0, False), (1.0, 1.0, False), (1.0, 0.75, True), (0.5, 0.5, True)]:
        d = {k: ctx.vector(v) for k, v in dict(x=x, c=c, kty=kty, lo=lo, hi=hi, anc=anc).items()}
        xn, dxv, xb = ctx.allocate(5), ctx.allocate(5), ctx.allocate(5)
        ctx.primal_update(5, 0.75, gamma, h, d["x"], d["c"], d["kty"], d["lo"], d["hi"],
                          d["anc"] if blend else None, xn, dxv, xb)
        w = _oracle.primal_update(0.75, gamma, h, x, c, kty, lo, hi, anc if blend else None)
        assert np.array_equal(xn.download(), w[0])
        assert np.array_equal(dxv.download(), w[1])
        assert np.array_equal(xb.download(), w[2])
        # the closed form of the reference test
        v = np.minimum(np.maximum(x - 0.75 * (c - kty), lo), hi)
        want = v + gamma * (v - x)
        if blend:
            want = h * want + (1 - h) * anc
        assert np.allclose(w[0], want, rtol=0, atol=1e-15)


def test_primal_update_allows_aliasing(ctx, orc):
    x = np.array([0.5, -1.0, 2.0, 1.0])
    c = np.array([1.0, 0.0, -1.0, 0.5])
    kty = np.array([0.25, 0.5, 0.0, -1.0])
    lo, hi = np.full(4, -INF), np.full(4, INF)
    dc, dk, dl, dh = ctx.vector(c), ctx.vector(kty), ctx.vector(lo), ctx.vector(hi)
    sep, out, d1, b1 = ctx.vector(x), ctx.allocate(4), ctx.allocate(4), ctx.allocate(4)
    ctx.primal_update(4, 0.5, 1.0, 1.0, sep, dc, dk, dl, dh, None, out, d1, b1)
    inp, d2, b2 = ctx.vector(x), ctx.allocate(4), ctx.allocate(4)
    ctx.primal_update(4, 0.5, 1.0, 1.0, inp, dc, dk, dl, dh, None, inp, d2, b2)
    assert np.array_equal(inp.download(), out.download())
    assert np.array_equal(b2.download(), b1.download())


def test_dual_update_is_the_oracle(ctx, orc):
    # test_backend.cpp:359-406
    import _oracle
    y = np.array([1.0, -1.0, 0.5, -0.5, 2.0])
    q = np.full(5, 2.0)
    kxb = np.full(5, 3.0)
    kx = np.ones(5)
    ay = np.array([0.1, 0.2, 0.3, 0.4, 0.5])
    ak = np.array([-1.0, 0.0, 1.0, 2.0, 3.0])
    for blend in (False, True):
        w = 0.6 if blend else 1.0
        d = {k: ctx.vector(v) for k, v in dict(y=y, q=q, kxb=kxb, kx=kx, ay=ay, ak=ak).items()}
        yn, dyv, ko, kd = (ctx.allocate(5) for _ in range(4))
        ctx.dual_update(5, 2, 0.5, 1.0, w, d["y"], d["q"], d["kxb"], d["kx"],
                        d["ay"] if blend else None, d["ak"] if blend else None, yn, dyv, ko, kd)
        want = _oracle.dual_update(2, 0.5, 1.0, w, y, q, kxb, kx, ay if blend else None,
                                   ak if blend else None)
        for got, exp in zip((yn, dyv, ko, kd), want):
            assert np.array_equal(got.download(), exp)
        v = y + 0.5 * (q - kxb)
        v[2:] = np.maximum(v[2:], 0.0)
        assert np.array_equal(want[1], v - y)


def test_kdx_and_kx_identities_against_a_real_matrix(ctx, orc):
    # test_backend.cpp:408-448: k_dx = K dx and k_out = K x_pdhg, to 1e-13.
    import _oracle
    s, i, v = _oracle.csr_from_triplets(4, 4, [0, 0, 1, 2, 2, 3], [0, 2, 1, 0, 3, 2],
                                        [1.5, -2.0, 3.25, 4.0, -0.5, 7.0])
    a = Matrix(4, 4, s, i, v)
    x = np.array([0.5, -1.0, 2.0, 1.0])
    xp = np.array([1.5, -0.5, 1.0, 2.0])
    xb, diff = 2 * xp - x, xp - x
    kx, kxb, kd_direct, kp_direct = (ctx.allocate(4) for _ in range(4))
    ctx.multiply(a, ctx.vector(x), kx)
    ctx.multiply(a, ctx.vector(xb), kxb)
    ctx.multiply(a, ctx.vector(diff), kd_direct)
    ctx.multiply(a, ctx.vector(xp), kp_direct)
    z = ctx.vector(np.zeros(4))
    yn, dyv, ko, kd = (ctx.allocate(4) for _ in range(4))
    ctx.dual_update(4, 4, 0.0, 0.0, 1.0, z, z, kxb, kx, None, None, yn, dyv, ko, kd)
    assert np.allclose(kd.download(), kd_direct.download(), rtol=0, atol=1e-13)
    assert np.allclose(ko.download(), kp_direct.download(), rtol=0, atol=1e-13)


def test_accumulate_and_average(ctx, orc):
    import _oracle
    a, c = np.array([1.0, 2.0, 3.0, 4.0]), np.full(4, 0.5)
    s = ctx.vector(np.zeros(4))
    ctx.accumulate(4, 2.0, ctx.vector(a), s)
    ctx.accumulate(4, 3.0, ctx.vector(c), s)
    got = s.download()
    assert np.array_equal(got, 2.0 * a + 3.0 * c)
    out = ctx.allocate(4)
    ctx.scale_into(4, 5.0, s, out)
    assert np.array_equal(out.download(), got / 5.0)


@pytest.mark.parametrize("n,m", [(7, 3), (140000, 517), (517, 140000)])
def test_weighted_norm_and_step_terms(ctx, orc, n, m):
    import _oracle
    rng = np.random.default_rng(4242)
    dx, dy, kdx = rng.uniform(-2, 2, n), rng.uniform(-2, 2, m), rng.uniform(-2, 2, m)
    ddx, ddy, dk = ctx.vector(dx), ctx.vector(dy), ctx.vector(kdx)
    for omega in (0.25, 1.0, 4.0, 2.5):
        assert rel(ctx.weighted_norm_squared(ddx, ddy, omega),
                   _oracle.weighted_norm_squared(dx, dy, omega)) <= 1e-12
    inter, mov = ctx.step_size_terms(ddx, ddy, dk, 2.5)
    wi, wm = _oracle.step_size_terms(dx, dy, kdx, 2.5)
    assert inter >= 0.0
    assert rel(inter, wi) <= 1e-10
    assert rel(mov, wm) <= 1e-10


def test_convergence_terms(ctx, orc):
    # test_backend.cpp:533-645: all four bound classes, zero scales.
    import _oracle
    rng = np.random.default_rng(90210)
    n, m, neq = 1301, 907, 401
    x, kty, c = rng.uniform(-2, 2, n), rng.uniform(-2, 2, n), rng.uniform(-2, 2, n)
    cs = 0.25 + 0.5 * (np.arange(n) % 5)
    lo, hi = np.empty(n), np.empty(n)
    cls = np.arange(n) % 4
    lo[cls == 0], hi[cls == 0] = -1.0, 2.0
    lo[cls == 1], hi[cls == 1] = -0.5, INF
    lo[cls == 2], hi[cls == 2] = -INF, 1.5
    lo[cls == 3], hi[cls == 3] = -INF, INF
    cs[7] = 0.0
    y, kx, q = rng.uniform(-2, 2, m), rng.uniform(-2, 2, m), rng.uniform(-2, 2, m)
    rs = 0.5 + 0.25 * (np.arange(m) % 3)
    rs[11] = 0.0
    V = {k: ctx.vector(v) for k, v in dict(x=x, kty=kty, c=c, cs=cs, lo=lo, hi=hi, y=y, kx=kx,
                                               q=q, rs=rs).items()}
    got = ctx.convergence_terms(n, m, V["c"], V["lo"], V["hi"], V["q"], V["rs"], V["cs"], neq,
                                V["x"], V["kx"], V["y"], V["kty"])
    want = _oracle.convergence_terms(c, lo, hi, q, rs, cs, neq, x, kx, y, kty)
    for k in ("primal_sq", "q_dot_y", "dual_sq", "bound_term", "c_dot_x"):
        assert rel(got[k], want[k]) <= 1e-11, k
    assert got["primal_inf"] == want["primal_inf"]
    assert got["dual_inf"] == want["dual_inf"]
    assert np.isfinite(got["bound_term"]) and np.isfinite(got["dual_sq"])


def test_large_random_consistency(ctx, orc):
    import _oracle
    rng = np.random.default_rng(20260825)
    n = 2000
    ri = np.repeat(np.arange(n), 5)
    ci = np.array([(i * 7 + k * 311) % n for i in range(n) for k in range(5)])
    s, idx, val = _oracle.csr_from_triplets(n, n, ri, ci, rng.uniform(-3, 3, 5 * n))
    a = Matrix(n, n, s, idx, val)
    x = rng.uniform(-3, 3, n)
    dy = ctx.allocate(n)
    ctx.multiply(a, ctx.vector(x), dy)
    assert np.array_equal(dy.download(), _oracle.spmv(s, idx, val, x))
    assert np.isfinite(ctx.two_norm(dy))
