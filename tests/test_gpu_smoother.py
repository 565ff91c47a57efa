"""The two forms of the FP32 line smoother of the Poisson V-cycle.

The full form folds all sixteen off-line couplings of a row into the line
right-hand side; the compact form, which a level takes when its cross
couplings (the twelve off-line neighbours off the i and k axes) are weak,
keeps the four axis neighbours.  The smoother only preconditions the FP64
FGMRES, so either form must give a solution whose residual with the
oracle's operator meets the tolerance.
"""
import numpy as np
import pytest

from helpers import cases, channel_pair, po, relerr, vfs

pytestmark = pytest.mark.gpu


def _solve(monkeypatch, env, shape_cells, tol=1e-8, mg_fp32=1):
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    c, o, ctx = channel_pair(*shape_cells, les=0)
    try:
        o.poisson_rhs()
        ctx.poisson_rhs()
        ctx.upload("PHI", np.zeros(o.shape))
        opts = vfs.default_solve_opts(tol=tol, mg_fp32=mg_fp32, use_guess=0, maxit=400)
        st = ctx.poisson_solve(opts)
        phi = ctx.download("PHI")
        b = o["b"]
        res = np.linalg.norm(b - o.poisson_apply(phi)) / np.linalg.norm(b)
        return st, phi, res
    finally:
        ctx.close()
        for k in env:
            monkeypatch.delenv(k)


@pytest.mark.parametrize("cells", [(24, 16, 20), (18, 13, 22)], ids=["24x16x20", "odd-18x13x22"])
@pytest.mark.parametrize("compact", ["0", "1"], ids=["full", "compact"])
def test_smoother_forms_converge(monkeypatch, cells, compact):
    st, phi, res = _solve(monkeypatch, {"VFS_SMOOTH_COMPACT": compact}, cells)
    assert st.relres <= 1e-8
    assert res <= 1.01e-8
    assert np.isfinite(phi).all()


@pytest.mark.parametrize("fp32", [1, 0], ids=["fp32", "fp64"])
@pytest.mark.parametrize("path", ["short", "long", "solve"])
def test_line_solve_paths(monkeypatch, path, fp32):
    """Every line-solve form (8-point scan, 16-point FP32 / 8-point FP64 scan,
    one thread per line: the forms long lines take at large grids) forced on
    small grids preconditions to the tolerance, checked with the oracle's
    operator, and agrees with the oracle's converged solution."""
    for cells in ((24, 16, 20), (18, 13, 22)):
        monkeypatch.setenv("VFS_LINE_PATH", path)
        c, o, ctx = channel_pair(*cells, les=0)
        try:
            o.poisson_rhs()
            ctx.poisson_rhs()
            ctx.upload("PHI", np.zeros(o.shape))
            st = ctx.poisson_solve(vfs.default_solve_opts(tol=1e-10, mg_fp32=fp32, use_guess=0, maxit=400))
            phi = ctx.download("PHI")
            b = o["b"]
            assert st.relres <= 1e-10
            assert np.linalg.norm(b - o.poisson_apply(phi)) / np.linalg.norm(b) <= 1.01e-10
            o.set("phi", np.zeros(o.shape))
            o.poisson_solve(tol=1e-10)
            assert relerr(phi, o["phi"]) <= 1e-5
        finally:
            ctx.close()
            monkeypatch.delenv("VFS_LINE_PATH")


@pytest.mark.parametrize("cells", [(24, 16, 20), (18, 13, 22), (40, 70, 36)], ids=["24x16x20", "18x13x22", "40x70x36"])
@pytest.mark.parametrize("form", ["compact-short", "compact-long", "compact-natural", "full", "fp64"])
def test_one_pass_forms_bitwise(monkeypatch, form, cells):
    """The V-cycle's one-pass forms against their two-pass forms, on whole
    FGMRES solves: the compact smoother's relaxation that forms each line
    segment's right-hand sides inside the line solve (it reads only the other
    colour, as the two-pass form does; on coarse levels its first half-sweep
    takes the iterate as zero instead of reading a zero-filled one), and the
    residual formed inside the restriction instead of stored (compact, full
    and FP64 levels), and on the compact path the finest level's vectors in
    the colour-split layout (compact-natural: in the natural one).  Same
    bits, same iteration count."""
    env = {"VFS_LINE_PATH": "short" if form == "compact-short" else "long"}
    env["VFS_SMOOTH_COMPACT"] = "0" if form == "full" else "1"
    fp32 = 0 if form == "fp64" else 1
    env1 = {**env, "VFS_MG_SPLIT": "0"} if form == "compact-natural" else env
    st1, phi1, res1 = _solve(monkeypatch, env1, cells, tol=1e-10, mg_fp32=fp32)
    st0, phi0, res0 = _solve(monkeypatch, {**env, "VFS_ZEBRA_FUSED": "0", "VFS_MG_FUSED_RESTRICT": "0"}, cells, tol=1e-10,
                             mg_fp32=fp32)
    assert st0.relres <= 1e-10 and res0 <= 1.01e-10
    assert st1.iterations == st0.iterations
    assert np.array_equal(phi1, phi0)


def test_compact_smoother_odd_periodic(monkeypatch):
    """An odd periodic interior count (19 cells in x) puts lines of one colour
    next to each other across the wrap; the right-hand side pass reads the
    wrap neighbour's value from before the half-sweep, and the compact form
    still preconditions to the tolerance."""
    st, phi, res = _solve(monkeypatch, {"VFS_SMOOTH_COMPACT": "1"}, (19, 12, 20), tol=1e-11)
    assert st.relres <= 1e-11
    assert res <= 1.01e-11
    # the same with the full form: both converge to the same solution
    st2, phi2, res2 = _solve(monkeypatch, {"VFS_SMOOTH_COMPACT": "0"}, (19, 12, 20), tol=1e-11)
    assert res2 <= 1.01e-11
    assert relerr(phi, phi2) <= 1e-6


def _distorted_channel_pair(nx, ny, nz, amp):
    """channel_pair with the nodes displaced by sines of an 8-node wavelength
    (amplitude amp times the local spacing), which tilts the cell faces and
    makes the cross couplings of the Poisson operator strong."""
    c = cases.channel(nx, ny, nz)
    IM, JM, KM = nx + 1, ny + 1, nz + 1
    K, J, I = np.meshgrid(np.arange(KM), np.arange(JM), np.arange(IM), indexing="ij")
    X = c.coor[:KM, :JM, :IM]
    dy = np.gradient(X[0, :, 0, 1])[None, :, None]
    s = lambda q: np.sin(2 * np.pi * q / 8)
    X[..., 1] += amp * dy * (1 - X[..., 1] ** 2) * s(I) * s(K)
    X[..., 0] += amp * (2 * np.pi / nx) * s(J) * s(K)
    X[..., 2] += amp * (np.pi / nz) * s(I) * s(J)
    op = po.make_params(*c.da, per=c.per, bctype=c.bctype, ren=c.ren, dt=1e-3, inlet_flux=c.inlet_flux, les=0)
    o = po.Oracle(op, c.coor)
    o.set("ucont", c.ucont)
    gp = cases.params_for(vfs, c, 1e-3, les=0)
    ctx = vfs.Context(gp)
    ctx.upload("COOR", c.coor)
    ctx.form_metrics()
    ctx.upload("UCONT", c.ucont)
    return o, ctx


def _cross_strength_cpu(o):
    """max over interior rows of sum |a_cross| / |a_P| from the oracle's
    operator, its 19 coefficients recovered by 27 probe vectors (one per
    residue class of (i, j, k) mod 3; the periodic interior counts are
    multiples of 3, so a row's neighbours fall in distinct classes)."""
    mz, my, mx = o.shape
    K, J, I = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    strength = np.zeros(o.shape)
    diag = np.zeros(o.shape)
    interior = (slice(1, mz - 1), slice(1, my - 1), slice(1, mx - 1))
    for ck in range(3):
        for cj in range(3):
            for ci in range(3):
                x = ((I % 3 == ci) & (J % 3 == cj) & (K % 3 == ck)).astype(float)
                y = o.poisson_apply(x)
                # row p picks up the coefficient of its neighbour in this class
                for dk in (-1, 0, 1):
                    for dj in (-1, 0, 1):
                        for di in (-1, 0, 1):
                            nz_off = (dk != 0) + (dj != 0) + (di != 0)
                            if nz_off == 3:
                                continue
                            qi = (I + di - 1) % (mx - 2) + 1
                            qk = (K + dk - 1) % (mz - 2) + 1
                            qj = J + dj
                            m = (qi % 3 == ci) & (qj % 3 == cj) & (qk % 3 == ck)
                            if nz_off == 0:
                                diag[m] = y[m]
                            elif nz_off == 2:
                                strength[m] += np.abs(y[m])
    return float(np.max(strength[interior] / np.abs(diag[interior])))


@pytest.mark.parametrize("amp,form", [(0.0, "compact"), (0.2, "full")], ids=["smooth", "distorted"])
def test_cross_strength_and_choice(monkeypatch, capfd, amp, form):
    """The level's cross-coupling strength measured on the device equals the
    oracle operator's, and picks the compact form only when it is weak."""
    o, ctx = _distorted_channel_pair(24, 16, 24, amp)
    try:
        monkeypatch.setenv("VFS_SOLVER_TRACE", "1")
        capfd.readouterr()
        ctx.upload("PHI", np.zeros(o.shape))
        ctx.poisson_apply_dev(ctx.device_ptr("PHI"), ctx.device_ptr("RHS"))  # builds the hierarchy
        ctx.sync()
        err = capfd.readouterr().err
    finally:
        ctx.close()
    import re
    levels = re.findall(r"level (\d+)x(\d+)x(\d+): cross-coupling strength ([0-9.eE+-]+), (\w+) FP32 smoother", err)
    assert levels, err
    mx, my, mz, xs, got = levels[0]
    assert (int(mx), int(my), int(mz)) == (26, 18, 26)
    ref = _cross_strength_cpu(o)
    assert abs(float(xs) - ref) <= 1e-6 + 1e-5 * ref
    assert got == form
    assert (ref <= 0.01) == (form == "compact")


def test_deep_level_sweeps(monkeypatch):
    """deep_level / deep_pre / deep_post: the levels from deep_level on smooth
    their own sweep counts.  In one context a default solve, a solve with
    1+2 sweeps from level 1 (the captured coarse levels are recaptured for
    it) and the default again: every solve meets the tolerance with the
    oracle's operator, the deep variant gives a different iterate, and the
    third solve repeats the first bit for bit."""
    c, o, ctx = channel_pair(40, 32, 40, les=0)
    try:
        o.poisson_rhs()
        ctx.poisson_rhs()
        b = o["b"]
        out = []
        for deep in (0, 1, 0):
            ctx.upload("PHI", np.zeros(o.shape))
            st = ctx.poisson_solve(vfs.default_solve_opts(tol=1e-10, use_guess=0, maxit=400, deep_level=deep,
                                                          deep_pre=1, deep_post=2))
            phi = ctx.download("PHI")
            assert st.relres <= 1e-10
            assert np.linalg.norm(b - o.poisson_apply(phi)) / np.linalg.norm(b) <= 1.01e-10
            out.append((st.iterations, phi))
        assert out[0][0] == out[2][0] and np.array_equal(out[0][1], out[2][1])
        assert not np.array_equal(out[0][1], out[1][1])
    finally:
        ctx.close()
