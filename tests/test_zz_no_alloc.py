"""The allocation-free convergence terms (anuk_convergence_terms_noalloc)
against anuk_convergence_terms, through the C ABI, on two calls in a row,
the second smaller than the first. Both run the same kernels with the same
fixed reduction tree, so the bar is bit equality of all seven terms; the
allocating entry point itself is pinned against the oracle in
tests/test_gpu_backend.py. The second call is the smaller one: it must
fold only the partials it wrote itself, since the larger first call left
stale partials in the rest of the shared array."""
import ctypes

import numpy as np
import pytest

from anukulan_amd import Context

INF = np.inf


def problem(rng, n, m):
    x, kty, c = rng.uniform(-2, 2, n), rng.uniform(-2, 2, n), rng.uniform(-2, 2, n)
    cs = 0.25 + 0.5 * (np.arange(n) % 5)
    lo, hi = np.empty(n), np.empty(n)
    cls = np.arange(n) % 4
    lo[cls == 0], hi[cls == 0] = -1.0, 2.0
    lo[cls == 1], hi[cls == 1] = -0.5, INF
    lo[cls == 2], hi[cls == 2] = -INF, 1.5
    lo[cls == 3], hi[cls == 3] = -INF, INF
    cs[min(7, n - 1)] = 0.0
    y, kx, q = rng.uniform(-2, 2, m), rng.uniform(-2, 2, m), rng.uniform(-2, 2, m)
    rs = 0.5 + 0.25 * (np.arange(m) % 3)
    rs[min(11, m - 1)] = 0.0
    return dict(x=x, kty=kty, c=c, cs=cs, lo=lo, hi=hi, y=y, kx=kx, q=q, rs=rs)


@pytest.mark.gpu
def test_noalloc_matches_allocating_entry_point():
    ctx = Context(0)
    try:
        rng = np.random.default_rng(31415)
        # A large call (1024 reduction blocks per side: every partial slot
        # written), then a small one (a few blocks).
        for n, m, neq in ((2_000_000, 1_500_000, 700_000), (1301, 907, 401)):
            print(f"noalloc convergence terms n={n} m={m}", flush=True)
            host = problem(rng, n, m)
            V = {k: ctx.vector(v) for k, v in host.items()}
            args = (n, m, V["c"], V["lo"], V["hi"], V["q"], V["rs"], V["cs"], neq, V["x"],
                    V["kx"], V["y"], V["kty"])
            want = ctx.convergence_terms(*args)
            got = ctx.convergence_terms_noalloc(*args)
            for k in want:
                assert got[k] == want[k], (n, k, got[k], want[k])
            again = ctx.convergence_terms_noalloc(*args)
            assert again == got
            for v in V.values():
                v.free()
        # Bad sizes are refused before any launch, as by the allocating entry.
        from anukulan_amd._lib import ConvTerms, lib
        t = ConvTerms()
        rc = lib.anuk_convergence_terms_noalloc(ctx.h, -1, 3, None, None, None, None, None,
                                                ctypes.byref(t))
        assert rc == 1, rc  # ANUK_ERR_INVALID
        assert "negative" in lib.anuk_last_error(ctx.h).decode()
    finally:
        ctx.close()
