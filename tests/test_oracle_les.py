"""Pins of the oracle's Compute_Smagorinsky_Constant_1 (les.c:75-1141).

For a linear velocity u = A x on a uniform Cartesian grid every piece of the
dynamic procedure has a closed form away from the boundary:

* the centred and one-sided differences are exact, so S = sym(A) everywhere;
* the Simpson test filter (weights 1-4-1 per direction, rhs2.c:499-523)
  reproduces linear functions and has second moment h^2/3 per direction, so
  L_ab = filter(u_a u_b) - filter(u_a) filter(u_b) = sum_d A_ad A_bd h_d^2 / 3;
* the test-filter width is (sum of the 27 weights x 1/8-1/4-1/2-1)^(1/3) =
  (8 V)^(1/3) = 2 Delta, so M = (Delta^2 - (2 Delta)^2) |S| S = -3 Delta^2 |S| S;
* in Cartesian form the curvilinear contraction with the covariant tensor
  reduces to L:M and M:M, giving C = 0.5 L:M / (M:M + 1e-4) (les_eps is the
  file-scope 1e-4 of les.c:27).

A wrong filter moment, test-filter width, metric contraction, sign or eps
moves C far beyond round-off.
"""
import numpy as np
import pytest

import pyoracle as po


def linear_box(A, h=(0.11, 0.07, 0.09), n=(9, 9, 9), les=2, max_cs=1.0):
    nx, ny, nz = n
    mx, my, mz = nx + 2, ny + 2, nz + 2
    coor = np.zeros((mz, my, mx, 3))
    K, J, I = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    coor[: nz + 1, : ny + 1, : nx + 1, 0] = I * h[0]
    coor[: nz + 1, : ny + 1, : nx + 1, 1] = J * h[1]
    coor[: nz + 1, : ny + 1, : nx + 1, 2] = K * h[2]
    p = po.make_params(mx, my, mz, per=(0, 0, 0), bctype=(1,) * 6, les=les, max_cs=max_cs)
    o = po.Oracle(p, coor)
    K, J, I = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    X = np.stack([(I - 0.5) * h[0], (J - 0.5) * h[1], (K - 0.5) * h[2]], -1)  # ghosts extrapolated
    o.set("ucat", X @ np.asarray(A).T)
    return p, o, h


def closed_form_c(A, h):
    A = np.asarray(A, dtype=float)
    S = 0.5 * (A + A.T)
    Sabs = np.sqrt(2 * np.sum(S * S))
    L = np.einsum("ad,bd,d->ab", A, A, np.asarray(h) ** 2 / 3)
    delta2 = np.prod(h) ** (2 / 3)
    M = -3 * delta2 * Sabs * S
    return 0.5 * np.sum(L * M) / (np.sum(M * M) + 1e-4)


DEEP = (slice(3, -3),) * 3


@pytest.mark.parametrize("scale", [50.0, 2.0, 0.3])
def test_dynamic_coefficient_linear_field(scale):
    A = scale * np.array([[-1.0, 0.2, 0.1], [0.3, 0.5, -0.2], [0.1, 0.25, 0.45]])
    p, o, h = linear_box(A)
    cs = o.smagorinsky_constant(ti=5)
    want = closed_form_c(A, h)
    assert want > 1e-3
    np.testing.assert_allclose(cs[DEEP], want, rtol=1e-10)
    # boundary ghosts are 0; interior values lie in [0, max_cs]
    assert np.all(cs[0] == 0) and np.all(cs[:, :, -1] == 0)
    assert cs.min() >= 0 and cs.max() <= p.max_cs


def test_dynamic_coefficient_clipping():
    A = 50.0 * np.array([[-1.0, 0.2, 0.1], [0.3, 0.5, -0.2], [0.1, 0.25, 0.45]])
    want = closed_form_c(A, (0.11, 0.07, 0.09))
    p, o, h = linear_box(A, max_cs=0.5 * want)
    cs = o.smagorinsky_constant(ti=5)
    np.testing.assert_array_equal(cs[DEEP], 0.5 * want)
    # a field with L:M < 0 gives a negative coefficient, clipped to 0
    p2, o2, _ = linear_box(-np.asarray(A).T @ np.diag([1.0, -3.0, 1.0]))
    assert closed_form_c(-np.asarray(A).T @ np.diag([1.0, -3.0, 1.0]), h) < 0
    assert np.all(o2.smagorinsky_constant(ti=5)[DEEP] == 0)


def test_smagorinsky_constant_first_steps_and_les1():
    A = 2.0 * np.array([[-1.0, 0.2, 0.1], [0.3, 0.5, -0.2], [0.1, 0.25, 0.45]])
    p, o, h = linear_box(A, les=2)
    assert np.all(o.smagorinsky_constant(ti=0, tistart=0, rstart=0) == 0)   # les.c:77-80
    assert np.all(o.smagorinsky_constant(ti=1, tistart=0, rstart=0) == 0)
    assert np.any(o.smagorinsky_constant(ti=1, tistart=0, rstart=1) != 0)   # a restart skips the guard
    p1, o1, _ = linear_box(A, les=1)
    assert np.all(o1.smagorinsky_constant(ti=3) == 0.01)                    # les.c:82-85
    assert np.all(o1.smagorinsky_constant(ti=1) == 0)


def test_eddy_viscosity_with_dynamic_field():
    """nu_t = Cs (1/aj)^(2/3) |S| with the per-point coefficient."""
    A = 2.0 * np.array([[-1.0, 0.2, 0.1], [0.3, 0.5, -0.2], [0.1, 0.25, 0.45]])
    p, o, h = linear_box(A)
    cs = o.smagorinsky_constant(ti=5)
    o.eddy_viscosity(cs)
    S = 0.5 * (A + A.T)
    want = cs * np.prod(h) ** (2 / 3) * np.sqrt(2 * np.sum(S * S))
    np.testing.assert_allclose(o["nu_t"][1:-1, 1:-1, 1:-1], want[1:-1, 1:-1, 1:-1], rtol=1e-12, atol=1e-300)
