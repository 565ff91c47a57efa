"""GPU parity of the -imp 4 momentum path (FormFunction_SNES residual and the
matrix-free Newton-Krylov solve) against the FP64 oracle.

The residual must match to 1e-12 (normwise).  The solve's iterates are the
solver's own (the reference uses PETSc SNES trust region; parity unpinned for
iterates): the converged velocity is judged by the oracle's residual at it,
which must be below the requested relative tolerance.
"""
import numpy as np
import pytest
import torch

from helpers import channel_pair, relerr, vfs

pytestmark = pytest.mark.gpu


def _state(les=1):
    c, o, ctx = channel_pair(24, 16, 20, les=les)
    rng = np.random.default_rng(11)
    p = 0.1 * rng.standard_normal(o.shape)
    o.set("p", p)
    ctx.upload("P", p)
    if les:
        o.eddy_viscosity()
        ctx.eddy_viscosity()
    o.calc_k_flux()
    ctx.calc_k_flux()
    o.pressure_gradient()
    ctx.pressure_gradient()
    # old state: Ucont_o = Ucont, RHS_o = Formfunction_2(Ucont)
    ctx.store_old_state()
    uo = o["ucont"].copy()
    o.set("rhs", np.zeros(o.shape + (3,)))
    o.formfunction2(1.0)
    rhso = o["rhs"].copy()
    return c, o, ctx, uo, rhso


def _dev(a):
    t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    torch.cuda.synchronize()
    return t


@pytest.mark.parametrize("les", [1, 0], ids=["les", "noles"])
def test_snes_residual(les):
    c, o, ctx, uo, rhso = _state(les)
    assert relerr(ctx.download("UCONT_O"), uo) == 0.0
    assert relerr(ctx.download("RHS_O"), rhso) <= 1e-12
    rng = np.random.default_rng(5)
    u = uo + 1e-3 * rng.standard_normal(uo.shape) * (np.abs(uo) + 1e-3)
    u_d = _dev(u)
    r_d = torch.zeros_like(u_d)
    ctx.formfunction_snes(u_d.data_ptr(), r_d.data_ptr())
    ctx.sync()
    r_gpu = r_d.cpu().numpy()
    r_orc = o.formfunction_snes(u, uo, rhso)
    assert relerr(r_gpu, r_orc) <= 1e-12
    assert relerr(ctx.download("UCAT"), o["ucat"]) <= 1e-12
    ctx.close()


@pytest.mark.parametrize("feul", [0, 1], ids=["no-feul", "feul"])
def test_snes_residual_cell_pass_bitwise(monkeypatch, feul):
    """The residual formed inside Formfunction_2's cell pass (the default on
    the fused path) against the three separate passes (VFS_FUSED_CN=0): the
    same operations in the same order, so the same bits, with and without a
    body force."""
    c, o, ctx, uo, rhso = _state(1)
    rng = np.random.default_rng(9)
    if feul:
        ctx.upload("FEUL", 1e-2 * rng.standard_normal(uo.shape))
    u = uo + 1e-3 * rng.standard_normal(uo.shape) * (np.abs(uo) + 1e-3)
    out = []
    for env in ("1", "0"):
        monkeypatch.setenv("VFS_FUSED_CN", env)
        u_d = _dev(u)
        r_d = torch.zeros_like(u_d)
        ctx.formfunction_snes(u_d.data_ptr(), r_d.data_ptr(), feul)
        ctx.sync()
        out.append((r_d.cpu().numpy(), ctx.download("UCONT"), ctx.download("UCAT")))
    monkeypatch.delenv("VFS_FUSED_CN")
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)
    if not feul:
        assert relerr(out[0][0], o.formfunction_snes(u, uo, rhso)) <= 1e-12
    ctx.close()


def test_implicit_solve_cell_pass_bitwise(monkeypatch):
    """The Newton-Krylov solve with the residual and the Jacobian products'
    difference quotients formed in the cell pass, against the separate passes
    (VFS_FUSED_CN=0): the same iterates, hence the same velocity bits and
    counts."""
    out = []
    for env in ("1", "0"):
        monkeypatch.setenv("VFS_FUSED_CN", env)
        c, o, ctx, uo, rhso = _state(1)
        st = ctx.implicit_matrixfree(vfs.default_newton_opts(rtol=1e-6, restart=30))
        out.append((st.newton_iterations, st.linear_iterations, ctx.download("UCONT"), ctx.download("UCAT")))
        ctx.close()
    monkeypatch.delenv("VFS_FUSED_CN")
    assert out[0][:2] == out[1][:2]
    assert np.array_equal(out[0][2], out[1][2]) and np.array_equal(out[0][3], out[1][3])


def test_implicit_solve_root():
    c, o, ctx, uo, rhso = _state(1)
    tol = 1e-6
    st = ctx.implicit_matrixfree(vfs.default_newton_opts(rtol=tol, restart=30))
    assert st.fnorm <= tol * st.fnorm0 and st.newton_iterations >= 1
    u = ctx.download("UCONT")
    r0 = o.formfunction_snes(uo, uo, rhso)
    r = o.formfunction_snes(u, uo, rhso)
    assert np.linalg.norm(r) <= 2 * tol * np.linalg.norm(r0)
    ctx.close()


def test_cn_fractional_step():
    c, o, ctx = channel_pair(24, 16, 20, les=1)
    st = ctx.fractional_step_cn(vfs.default_newton_opts(rtol=1e-6), vfs.default_solve_opts(tol=1e-10))
    assert st.newton_iterations >= 1 and st.poisson_relres <= 1e-10
    assert st.maxdiv <= 1e-8
    ctx.close()
