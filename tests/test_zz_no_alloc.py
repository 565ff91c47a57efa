"""vfs_divergence_noalloc, the divergence check (main.c:936-1074) for callers
that cannot allocate device memory, against vfs_divergence on the same
fields: two calls in a row, the second on a smaller grid, each equal to the
context entry point's value (a max reduction: exact), enqueued on the null
stream with the result left in device memory."""
import ctypes as C

import numpy as np
import pytest

from helpers import channel_pair, vfs

pytestmark = pytest.mark.gpu


def _lib():
    return vfs.load_library()


def _device_sync():
    hip = C.CDLL("libamdhip64.so")
    assert hip.hipDeviceSynchronize() == 0


def test_divergence_noalloc_matches_context_entry():
    lib = _lib()
    for size, blanked in (((40, 24, 36), True), ((18, 13, 22), False)):
        c, o, ctx = channel_pair(*size)
        with ctx:
            if blanked:
                nv = np.zeros(ctx.field_shape("NVERT"))
                nv[10:14, 5:9, 6:12] = 1.0
                ctx.upload("NVERT", nv)
            # a field that is not divergence-free: the seeded fluxes plus a perturbation
            rng = np.random.default_rng(sum(size))
            u = ctx.download("UCONT") + 1e-3 * rng.standard_normal(ctx.field_shape("UCONT"))
            ctx.upload("UCONT", u)
            want = ctx.divergence()
            assert want > 0
            out = ctx.device_ptr("B")  # the result lands in B[0]
            ctx.sync()
            rc = lib.vfs_divergence_noalloc(C.byref(ctx.params), ctx.device_ptr("UCONT"), ctx.device_ptr("AJ"),
                                            ctx.device_ptr("NVERT"), out, None)
            assert rc == 0, lib.vfs_last_error()
            _device_sync()
            got = ctx.download("B").reshape(-1)[0]
            assert got == want, (size, got, want)


def test_divergence_noalloc_rejects_bad_sizes():
    lib = _lib()
    p = vfs.make_params(2048, 1024, 1024)  # 2^31 points
    dummy = C.c_void_p(8)  # never dereferenced: the size check comes first
    assert lib.vfs_divergence_noalloc(C.byref(p), dummy, dummy, dummy, dummy, None) == vfs.VFS_EINVAL
    assert b"2^31" in lib.vfs_last_error()
    assert lib.vfs_divergence_noalloc(C.byref(p), None, dummy, dummy, dummy, None) == vfs.VFS_EINVAL
