"""The adjoint's checker, checked (CPU only).

tests/_sensitivity.py computes exact basis sensitivities from the reference
simplex's final basis. Here it is held to central finite differences through
re-solves with the same simplex, on a nondegenerate LP (the committed
fixture's coordinates and differences), and to closed forms on a 2x2 LP whose
basis inverse is known by hand. A transposed solve, a sign slip on the slack
columns or a wrong bound assignment fails one of them.
"""
import os

import numpy as np
import pytest

import _sensitivity as S
from anukulan_amd import generators as G

FIX = os.path.join(os.path.dirname(__file__), "golden", "adjoint_reference.npz")


def test_closed_form_two_by_two():
    # min -x0 - x1  s.t.  x0 + 2 x1 >= -inf... written as  -x0 - 2x1 >= -4,
    # -3x0 - x1 >= -6, x >= 0: the vertex x = (1.6, 1.2) has both rows binding
    # and both columns basic, so x = B^-1 q with B = [[-1, -2], [-3, -1]].
    e = {"start": np.array([0, 2, 4]), "index": np.array([0, 1, 0, 1]),
         "value": np.array([-1.0, -2.0, -3.0, -1.0]), "c": np.array([-1.0, -1.0]),
         "q": np.array([-4.0, -6.0]), "lower": np.zeros(2), "upper": np.full(2, np.inf)}
    B = np.array([[-1.0, -2.0], [-3.0, -1.0]])
    x = np.linalg.solve(B, e["q"])
    w = np.array([0.7, -1.3])
    gc, gq, gl, gu = S.basis_sensitivity(e, [0, 1], x, w_x=w, w_y=np.array([0.2, 0.5]))
    assert np.allclose(gq, np.linalg.solve(B.T, w), rtol=0, atol=1e-14)
    assert np.allclose(gc, np.linalg.solve(B, [0.2, 0.5]), rtol=0, atol=1e-14)
    assert not gl.any() and not gu.any()


def test_basis_sensitivities_equal_finite_differences(ref):
    if not os.path.exists(FIX):
        pytest.skip("adjoint fixture not generated")
    fx = np.load(FIX)
    rlp = ref.RefLp.from_natural(G.random_lp(m=300, n=700, seed=5))
    e = rlp.export()
    sx = rlp.simplex()
    grads = S.basis_sensitivity(e, sx["basic"], sx["x"], fx["random_w_x"], fx["random_w_y"])
    for f, g in zip(("c", "q", "lower", "upper"), grads):
        assert np.array_equal(g, fx[f"random_grad_{f}"]), f
    got = np.array([grads[f][i] for f, i in fx["random_coords"]])
    fd = fx["random_fd"]
    err = np.abs(got - fd) / np.maximum(1.0, np.abs(fd))
    assert err.max() <= 1e-4, err.max()
    # and a fresh difference the fixture does not hold: the first q row
    loss = lambda x, y: float(fx["random_w_x"] @ x + fx["random_w_y"] @ y)  # noqa: E731
    d = S.central_difference(ref, e, "q", 0, loss)
    assert abs(d - grads[1][0]) <= 1e-4 * max(1.0, abs(d))


def test_basis_directions_at_a_nondegenerate_vertex(ref):
    # At a nondegenerate vertex no basic variable sits at a bound, so the
    # basis stays feasible both ways along every coordinate (the checker the
    # adjoint test uses to recognise coordinates a degenerate basis cannot
    # differentiate).
    import _sensitivity as S
    from anukulan_amd import generators as G
    rlp = ref.RefLp.from_natural(G.random_lp(m=60, n=150, seed=9))
    e = rlp.export()
    sx = rlp.simplex()
    assert sx["status"] == 0
    for i in range(0, len(e["q"]), 7):
        assert S.basis_directions(e, sx["basic"], sx["x"], "q", i) == (True, True)
