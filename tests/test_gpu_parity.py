"""Parity of the HIP path against the FP64 CPU oracle, function by function,
on the seeded channel case (periodic x/z, no-slip walls, curvilinear grid).

Tolerances: stencil kernels 1e-12 normwise relative (max|gpu-cpu|/max|cpu|);
the Poisson solution is judged by its own residual (<= 1e-8 relative, checked
with the oracle's operator) and by closeness to the oracle's converged
solution, which can differ by about cond(A) * 1e-8 since the two Krylov
histories differ.
"""
import numpy as np
import pytest

from helpers import channel_pair, relerr, vfs

pytestmark = pytest.mark.gpu

RHS_TOL = 1e-12


# the last case has sizes off every tile width (odd wall-normal count: a line
# point without a partner in the paired smoother pass; one multigrid level)
@pytest.fixture(scope="module",
                params=[(1, 1, 1, (24, 16, 20)), (1, 0, 0, (24, 16, 20)), (0, 1, 1, (24, 16, 20)), (1, 1, 1, (18, 13, 22))],
                ids=["les-2nd-lap", "les-4th", "nolesi-2nd", "odd-18x13x22"])
def pair(request):
    les, so, lap, (nx, ny, nz) = request.param
    c, o, ctx = channel_pair(nx, ny, nz, les=les, second_order=so, laplacian=lap)
    yield c, o, ctx
    ctx.close()


def test_metrics(pair):
    c, o, ctx = pair
    for name in vfs.METRIC_FIELDS:
        g = ctx.download(name)
        ref = o.m[name.lower()]
        assert relerr(g, ref) <= 1e-13, name


def test_contra2cart(pair):
    c, o, ctx = pair
    assert relerr(ctx.download("UCAT"), o["ucat"]) <= RHS_TOL
    assert relerr(ctx.download("UCONT"), o["ucont"]) <= RHS_TOL


def test_eddy_viscosity(pair):
    c, o, ctx = pair
    o.eddy_viscosity()
    ctx.eddy_viscosity()
    assert relerr(ctx.download("NU_T"), o["nu_t"]) <= RHS_TOL


def test_pressure_gradient(pair):
    c, o, ctx = pair
    rng = np.random.default_rng(7)
    p = rng.standard_normal(o.shape)
    o.set("p", p)
    ctx.upload("P", p)
    o.calc_k_flux()
    ctx.calc_k_flux()
    o.pressure_gradient()
    ctx.pressure_gradient()
    assert relerr(ctx.download("DP"), o["dp"]) <= RHS_TOL
    assert relerr(ctx.download("P"), o["p"]) == 0.0


def test_formfunction2(pair):
    c, o, ctx = pair
    o.eddy_viscosity()
    ctx.upload("NU_T", o["nu_t"])
    rng = np.random.default_rng(11)
    rhs0 = rng.standard_normal(o["rhs"].shape)
    o.set("rhs", rhs0)
    ctx.upload("RHS", rhs0)
    o.formfunction2(0.7)
    ctx.formfunction2(0.7)
    assert relerr(ctx.download("RHS"), o["rhs"]) <= RHS_TOL


def test_blanking_free_instances(pair, monkeypatch):
    """With no blanked point the eddy viscosity, the face and cell passes,
    the pressure gradient, the projection and Contra2Cart run without their
    blanking tests: the same bits as with them (VFS_NOBLANK=0).  Blanked
    cells (nvert = 1 on a small block) reset that choice, and the kernels with
    their tests match the oracle."""
    c, o, ctx = pair
    names = ("RHS", "NU_T", "UCONT", "UCAT", "P", "PHI", "DP")
    saved = {f: ctx.download(f) for f in names}
    saved_o = {f: o[f].copy() for f in o.fields}
    o.eddy_viscosity()
    ctx.upload("NU_T", o["nu_t"])
    rng = np.random.default_rng(21)
    p = rng.standard_normal(o.shape)
    phi = rng.standard_normal(o.shape)
    zero = np.zeros(o["rhs"].shape)

    def gpu():
        for f, v in (("UCONT", saved["UCONT"]), ("UCAT", saved["UCAT"]), ("P", p), ("PHI", phi), ("RHS", zero)):
            ctx.upload(f, v)
        ctx.eddy_viscosity()
        ctx.formfunction2(1.0)
        ctx.calc_k_flux()
        ctx.pressure_gradient()
        ctx.projection()
        return [ctx.download(f) for f in ("NU_T", "RHS", "DP", "UCONT")]

    def cpu():
        for f, v in (("ucont", saved["UCONT"]), ("ucat", saved["UCAT"]), ("p", p), ("phi", phi), ("rhs", zero)):
            o.set(f, v)
        o.eddy_viscosity()
        o.formfunction2(1.0)
        o.calc_k_flux()
        o.pressure_gradient()
        o.projection()
        return [o[f] for f in ("nu_t", "rhs", "dp", "ucont")]

    try:
        no_tests = gpu()
        monkeypatch.setenv("VFS_NOBLANK", "0")
        with_tests = gpu()
        monkeypatch.delenv("VFS_NOBLANK")
        for a_, b_ in zip(no_tests, with_tests):
            assert np.array_equal(a_, b_)
        for a_, b_ in zip(no_tests, cpu()):
            assert relerr(a_, b_) <= RHS_TOL
        mz, my, mx = o.shape
        nv = np.zeros(o.shape)
        nv[mz // 2, my // 2: my // 2 + 2, 3:6] = 1.0
        o.m["nvert"][...] = nv
        ctx.upload("NVERT", nv)
        blanked = gpu()
        for a_, b_ in zip(blanked, cpu()):
            assert relerr(a_, b_) <= RHS_TOL
        assert not np.array_equal(blanked[1], no_tests[1])
    finally:
        o.m["nvert"][...] = 0.0
        ctx.upload("NVERT", np.zeros(o.shape))
        for f, v in saved.items():  # the fixture is shared by the module's tests
            ctx.upload(f, v)
        for f, v in saved_o.items():
            o.set(f, v)


def test_stored_face_metrics_paths(pair):
    """Once a metric array may have been written from outside (here: its
    device pointer was handed out) the face pass reads the stored face-metric
    arrays instead of rebuilding face metrics from the cells; both ways give
    the same bits, and the oracle's result.  Pressure_Gradient and Projection
    (which always read the stored arrays) are run in both states too."""
    c, o, ctx = pair
    saved = {f: ctx.download(f) for f in ("P", "PHI", "RHS", "DP", "UCONT", "UCAT", "NU_T")}
    o.eddy_viscosity()
    ctx.upload("NU_T", o["nu_t"])
    rng = np.random.default_rng(12)
    p = rng.standard_normal(o.shape)
    phi = rng.standard_normal(o.shape)
    ucont0 = ctx.download("UCONT")

    def run():
        ctx.upload("UCONT", ucont0)
        ctx.contra2cart()
        ctx.upload("RHS", np.zeros(o["rhs"].shape))
        ctx.formfunction2(1.0)
        ctx.upload("P", p)
        ctx.pressure_gradient()
        ctx.upload("PHI", phi)
        ctx.projection()
        out = [ctx.download(f) for f in ("RHS", "DP", "UCONT")]
        ctx.upload("UCONT", ucont0)
        ctx.contra2cart()
        return out

    derived = run()
    ctx.device_ptr("ICSI")
    stored = run()
    for d, s_ in zip(derived, stored):
        assert np.array_equal(d, s_)
    o.set("ucont", ucont0)
    o.contra2cart()
    o.set("rhs", np.zeros(o["rhs"].shape))
    o.formfunction2(1.0)
    assert relerr(stored[0], o["rhs"]) <= RHS_TOL
    ctx.form_metrics()  # the same metrics again: back on the cell-metric path
    for f, v in saved.items():  # the fixture is shared by the module's tests
        ctx.upload(f, v)


def test_fused_flux_divergence(pair, monkeypatch):
    """The second-order face pass that forms Fp itself (no flux planes in
    HBM) gives the same bits as the face pass + flux-plane periodic fix +
    cell Fp pass, for Formfunction_2 and for a whole RK4 step."""
    c, o, ctx = pair
    saved = {f: ctx.download(f) for f in ("RHS", "UCONT", "UCAT", "NU_T", "DP")}
    o.eddy_viscosity()
    ctx.upload("NU_T", o["nu_t"])

    def run():
        for f in ("UCONT", "UCAT", "DP"):
            ctx.upload(f, saved[f])
        ctx.upload("RHS", np.zeros(o["rhs"].shape))
        ctx.formfunction2(1.0)
        rhs = ctx.download("RHS")
        ctx.runge_kutta()
        return rhs, ctx.download("UCONT"), ctx.download("UCAT")

    fused = run()
    monkeypatch.setenv("VFS_FUSED_FP", "0")
    planes = run()
    monkeypatch.delenv("VFS_FUSED_FP")
    for a_, b_ in zip(fused, planes):
        assert np.array_equal(a_, b_)
    for f, v in saved.items():  # the fixture is shared by the module's tests
        ctx.upload(f, v)


def test_runge_kutta(pair):
    c, o, ctx = pair
    o.eddy_viscosity(); ctx.upload("NU_T", o["nu_t"])
    o.calc_k_flux(); ctx.calc_k_flux()
    o.pressure_gradient(); ctx.pressure_gradient()
    o.runge_kutta()
    ctx.runge_kutta()
    assert relerr(ctx.download("UCONT"), o["ucont"]) <= RHS_TOL
    assert relerr(ctx.download("UCAT"), o["ucat"]) <= RHS_TOL


def test_poisson_rhs_apply(pair):
    c, o, ctx = pair
    o.poisson_rhs()
    ctx.poisson_rhs()
    assert relerr(ctx.download("B"), o["b"]) <= RHS_TOL
    rng = np.random.default_rng(3)
    x = rng.standard_normal(o.shape)
    ctx.upload("PHI", x)
    ctx.poisson_apply_dev(ctx.device_ptr("PHI"), ctx.device_ptr("RHS"))
    # RHS is a vector field; the operator wrote the first third of it
    y = ctx.download("RHS").reshape(-1)[: x.size].reshape(o.shape)
    ref = o.poisson_apply(x)
    assert relerr(y, ref) <= RHS_TOL


@pytest.mark.parametrize("fp32", [1, 0])
def test_poisson_solve(pair, fp32):
    c, o, ctx = pair
    o.poisson_rhs()
    ctx.poisson_rhs()
    ctx.upload("PHI", np.zeros(o.shape))
    st = ctx.poisson_solve(vfs.default_solve_opts(tol=1e-8, mg_fp32=fp32, use_guess=0))
    assert st.relres <= 1e-8
    phi = ctx.download("PHI")
    b = o["b"]
    r = b - o.poisson_apply(phi)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1.01e-8
    # both solvers converged hard: the solutions must then agree closely
    ctx.upload("PHI", np.zeros(o.shape))
    st = ctx.poisson_solve(vfs.default_solve_opts(tol=1e-11, mg_fp32=fp32, use_guess=0, maxit=400))
    assert st.relres <= 1e-11
    o.set("phi", np.zeros(o.shape))
    it, rel = o.poisson_solve(tol=1e-11)
    assert rel <= 1e-11
    assert relerr(ctx.download("PHI"), o["phi"]) <= 1e-6


@pytest.mark.parametrize("fp32", [1, 0])
def test_poisson_graph_replays(pair, fp32, monkeypatch):
    """The coarse V-cycle levels run as a captured HIP graph: every replay
    (with other eager work and synchronisations in between) gives the same
    bits as launching the same kernels directly."""
    c, o, ctx = pair
    ctx.poisson_rhs()
    opts = vfs.default_solve_opts(tol=1e-10, mg_fp32=fp32, use_guess=0)
    monkeypatch.setenv("VFS_MG_GRAPH", "0")
    ctx.upload("PHI", np.zeros(o.shape))
    ctx.poisson_solve(opts)
    eager = ctx.download("PHI")
    monkeypatch.delenv("VFS_MG_GRAPH")
    for rep in range(3):
        ctx.upload("PHI", np.zeros(o.shape))
        st = ctx.poisson_solve(opts)
        assert st.relres <= 1e-10
        assert np.array_equal(ctx.download("PHI"), eager), rep
        ctx.contra2cart()
        ctx.sync()


def test_k_marching_kernels(pair, monkeypatch):
    """The k-marching eddy viscosity and divergence give the bits of the
    thread-per-point kernels (VFS_EDDY_KM=0, VFS_REDUCE_KM=0); the k-marching
    Poisson right-hand side sums in another order (1e-14)."""
    c, o, ctx = pair
    saved = {f: ctx.download(f) for f in ("NU_T", "B")}
    out = {}
    for km in ("1", "0"):
        monkeypatch.setenv("VFS_EDDY_KM", km)
        monkeypatch.setenv("VFS_REDUCE_KM", km)
        ctx.upload("NU_T", np.full(o.shape, 7.0))
        ctx.upload("B", np.full(o.shape, 7.0))
        ctx.eddy_viscosity()
        d = ctx.divergence()
        ctx.poisson_rhs()
        out[km] = (ctx.download("NU_T"), d, ctx.download("B"))
    monkeypatch.delenv("VFS_EDDY_KM")
    monkeypatch.delenv("VFS_REDUCE_KM")
    assert np.array_equal(out["1"][0], out["0"][0])
    assert out["1"][1] == out["0"][1]
    assert relerr(out["1"][2], out["0"][2]) <= 1e-14
    for f, v in saved.items():  # the fixture is shared by the module's tests
        ctx.upload(f, v)


def test_update_projection_divergence(pair):
    c, o, ctx = pair
    o.poisson_rhs()
    o.set("phi", np.zeros(o.shape))
    o.poisson_solve(tol=1e-10)
    phi = o["phi"].copy()
    rng = np.random.default_rng(5)
    p = rng.standard_normal(o.shape)
    o.set("p", p); ctx.upload("P", p)
    ctx.upload("PHI", phi)
    o.update_pressure(); ctx.update_pressure()
    assert relerr(ctx.download("P"), o["p"]) <= RHS_TOL
    assert relerr(ctx.download("PHI"), o["phi"]) == 0.0
    o.projection(); ctx.projection()
    assert relerr(ctx.download("UCONT"), o["ucont"]) <= RHS_TOL
    assert relerr(ctx.download("UCAT"), o["ucat"]) <= RHS_TOL
    d_cpu = o.divergence()
    d_gpu = ctx.divergence()
    assert abs(d_gpu - d_cpu) <= 1e-12 * max(1.0, abs(d_cpu))


def test_poisson_solve_after_reused_memory():
    """Regression: buffers of a new solver must not inherit garbage from freed
    device memory (a NaN left in a never-written ghost cell used to poison the
    line smoother through a zero coefficient)."""
    c, o, ctx = channel_pair(24, 16, 20, les=0)
    shape = o.shape
    junk = vfs.Context(ctx.params)
    for name in vfs.FIELDS:
        junk.upload(name, np.full(junk.field_shape(name), np.nan))
    junk.close()
    o.runge_kutta()
    ctx.upload("UCONT", o["ucont"])
    ctx.poisson_rhs()
    ctx.upload("PHI", np.zeros(shape))
    st = ctx.poisson_solve(vfs.default_solve_opts(tol=1e-8, mg_fp32=1, use_guess=0))
    assert st.relres <= 1e-8 and np.isfinite(ctx.download("PHI")).all()
    ctx.close()


def _blank(o, ctx):
    """A solid block (nvert 3) wrapped in a layer of boundary cells (nvert 1)
    inside the channel, in the oracle and the context."""
    nv = np.zeros(o.shape)
    mz, my, mx = o.shape
    nv[mz // 2 - 3: mz // 2 + 3, 3: 9, mx // 2 - 4: mx // 2 + 3] = 1.0
    nv[mz // 2 - 2: mz // 2 + 2, 4: 8, mx // 2 - 3: mx // 2 + 2] = 3.0
    o.m["nvert"][...] = nv
    ctx.upload("NVERT", nv)
    return nv


@pytest.mark.parametrize("fp32", [1, 0])
def test_poisson_blanked(fp32):
    """setup_lidx2's unknowns (the row mask), PoissonLHSNew's rows next to
    blanked cells (one-sided cross terms) and the blanked coarse levels, on a
    fresh context whose hierarchy is built with the blanking: right-hand side
    and operator at 1e-12, the solve to 1e-8 with the oracle's operator."""
    c, o, ctx = channel_pair(24, 16, 20, les=0)
    with ctx:
        nv = _blank(o, ctx)
        o.runge_kutta()
        ctx.upload("UCONT", o["ucont"])
        o.poisson_rhs()
        ctx.poisson_rhs()
        assert relerr(ctx.download("B"), o["b"]) <= RHS_TOL
        assert np.all(o["b"][nv > 0.1] == 0)
        rng = np.random.default_rng(4)
        x = rng.standard_normal(o.shape)
        ctx.upload("PHI", x)
        ctx.poisson_apply_dev(ctx.device_ptr("PHI"), ctx.device_ptr("RHS"))
        y = ctx.download("RHS").reshape(-1)[: x.size].reshape(o.shape)
        assert relerr(y, o.poisson_apply(x)) <= RHS_TOL
        ctx.upload("PHI", np.zeros(o.shape))
        st = ctx.poisson_solve(vfs.default_solve_opts(tol=1e-8, mg_fp32=fp32, use_guess=0, maxit=400))
        assert st.relres <= 1e-8
        phi = ctx.download("PHI")
        b = o["b"]
        assert np.linalg.norm(b - o.poisson_apply(phi)) / np.linalg.norm(b) <= 1.01e-8


def test_poisson_rebuilt_after_blanking_changes():
    """A solve, then an upload of a blanked nvert, then a solve again on the
    same context: the operator hierarchy is rebuilt from the new blanking
    (PoissonLHSNew is re-run whenever nvert may change, poisson_hypre.c:534-550)
    and the second solution satisfies the oracle's blanked operator."""
    c, o, ctx = channel_pair(24, 16, 20, les=0)
    with ctx:
        opts = vfs.default_solve_opts(tol=1e-8, mg_fp32=1, use_guess=0, maxit=400)
        ctx.poisson_rhs()
        ctx.upload("PHI", np.zeros(o.shape))
        assert ctx.poisson_solve(opts).relres <= 1e-8
        _blank(o, ctx)
        o.poisson_rhs()
        ctx.poisson_rhs()
        ctx.upload("PHI", np.zeros(o.shape))
        st = ctx.poisson_solve(opts)
        phi = ctx.download("PHI")
        b = o["b"]
        assert st.relres <= 1e-8
        assert np.linalg.norm(b - o.poisson_apply(phi)) / np.linalg.norm(b) <= 1.01e-8


def test_write_through_kept_pointer():
    """Ucont written through a device pointer kept from an earlier call, then
    declared with vfs_field_written: the next Contra2Cart recomputes Ucat."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    c, o, ctx = channel_pair(16, 12, 16, les=0)
    with ctx:
        ptr = ctx.device_ptr("UCONT")
        ctx.contra2cart()  # Ucat fresh for the current Ucont
        u2 = 0.5 * o["ucont"]
        ctx.sync()
        assert hip.hipMemcpy(ptr, u2.ctypes.data, u2.nbytes, 1) == 0  # hipMemcpyHostToDevice
        ctx.field_written("UCONT")
        ctx.contra2cart()
        o.set("ucont", u2)
        o.contra2cart()
        assert relerr(ctx.download("UCAT"), o["ucat"]) <= RHS_TOL


def test_matrix_free_operator_bitwise(monkeypatch):
    """With VFS_POISSON_MF=1 and no blanked point the finest FP64 operator
    is applied matrix-free from nine face products per point; its products
    and a whole FGMRES solve have the bits of the stored-coefficient path
    (the default)."""
    out = {}
    for mf in ("1", "0"):
        monkeypatch.setenv("VFS_POISSON_MF", mf)
        c, o, ctx = channel_pair(24, 16, 20)
        with ctx:
            rng = np.random.default_rng(3)
            x = rng.standard_normal(o.shape)
            ctx.upload("PHI", x)
            ctx.poisson_apply_dev(ctx.device_ptr("PHI"), ctx.device_ptr("RHS"))
            y = ctx.download("RHS").reshape(-1)[: x.size].copy()
            ctx.runge_kutta()
            ctx.poisson_rhs()
            ctx.upload("PHI", np.zeros(o.shape))
            st = ctx.poisson_solve(vfs.default_solve_opts(tol=1e-10, use_guess=0))
            out[mf] = (y, ctx.download("PHI"), st.iterations)
    monkeypatch.delenv("VFS_POISSON_MF")
    assert np.array_equal(out["1"][0], out["0"][0])
    assert np.array_equal(out["1"][1], out["0"][1]) and out["1"][2] == out["0"][2]
    assert relerr(out["1"][0].reshape(o.shape), o.poisson_apply(x)) <= RHS_TOL
