"""The algebraic multigrid preconditioner (the role of HYPRE BoomerAMG,
poisson_hypre.c:34-97) inside the FP64 FGMRES on the GPU: the solution is
judged as the reference's own solve is, at the solution level -- relative
residual <= 1e-8 under the oracle's operator, and agreement with the oracle's
converged solution when both solve to 1e-11 -- on the channel, on an odd-sized
channel, and with blanked cells (where geometric coarsening of the blanking is
poorest)."""
import numpy as np
import pytest

from helpers import channel_pair, relerr, vfs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("size,blank", [((24, 16, 20), False), ((18, 13, 22), False), ((24, 16, 20), True)],
                         ids=["channel", "odd", "blanked"])
def test_amg_solve(size, blank):
    c, o, ctx = channel_pair(*size, les=0)
    with ctx:
        if blank:
            nv = np.zeros(o.shape)
            nv[6:12, 4:9, 7:13] = 1.0
            o.m["nvert"][...] = nv
            ctx.upload("NVERT", nv)
        ctx.runge_kutta()
        ctx.poisson_rhs()
        o.set("ucont", ctx.download("UCONT"))
        o.poisson_rhs()
        assert relerr(ctx.download("B"), o["b"]) <= 1e-12
        b = o["b"]
        ctx.upload("PHI", np.zeros(o.shape))
        st = ctx.poisson_solve(vfs.default_solve_opts(tol=1e-8, use_guess=0, amg=1, maxit=300))
        x = ctx.download("PHI")
        res = np.linalg.norm(b - o.poisson_apply(x)) / np.linalg.norm(b)
        print(f"[amg {size} blank={blank}] iterations {st.iterations} relres {st.relres:.3e} oracle {res:.3e}", flush=True)
        assert st.relres <= 1e-8 and res <= 1.01e-8
        ctx.upload("PHI", np.zeros(o.shape))
        st = ctx.poisson_solve(vfs.default_solve_opts(tol=1e-11, use_guess=0, amg=1, maxit=400))
        assert st.relres <= 1e-11
        o.set("phi", np.zeros(o.shape))
        it, rel = o.poisson_solve(tol=1e-11)
        assert rel <= 1e-11
        assert relerr(ctx.download("PHI"), o["phi"]) <= 1e-6


def test_amg_graph_replays_match_direct_launches(monkeypatch):
    """The V-cycle replayed from its HIP graph gives the bits of the same
    kernels launched directly, over a whole solve (one replay per FGMRES
    iteration, with the Krylov kernels launched eagerly between replays)."""
    out = []
    for graph in ("1", "0"):
        monkeypatch.setenv("VFS_AMG_GRAPH", graph)
        c, o, ctx = channel_pair(24, 16, 20, les=0)
        with ctx:
            ctx.runge_kutta()
            ctx.poisson_rhs()
            ctx.upload("PHI", np.zeros(o.shape))
            st = ctx.poisson_solve(vfs.default_solve_opts(tol=1e-10, use_guess=0, amg=1, maxit=300))
            out.append((st.iterations, ctx.download("PHI")))
    monkeypatch.delenv("VFS_AMG_GRAPH")
    assert out[0][0] == out[1][0] and out[0][0] >= 3
    assert np.array_equal(out[0][1], out[1][1])
