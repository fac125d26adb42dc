"""GPU: pcisph_mlp_gradients_noalloc, whose device scratch is a static per-GPU array,
against pcisph_mlp_gradients, whose scratch the library allocates, called through the
C ABI on the same network and rows: two calls in a row, the second with fewer rows (the
array keeps the first call's contents), and a call beyond the array's maximum.

The gradient kernels sum with float atomics in no fixed order, so the two entry points
agree to a tolerance, not bit for bit: 1e-5 of the largest gradient component (a stale
accumulator or a wrong row count is off by orders of magnitude more)."""
import ctypes as C

import numpy as np
import pytest

from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu

MAX_ROWS = 65536


def _noalloc(lib, f, x, t, n=None):
    g = np.zeros(M.N_PARAMS, np.float32)
    loss = np.zeros(1, np.float32)
    rc = lib.pcisph_mlp_gradients_noalloc(f._h, x.ctypes.data_as(C.c_void_p), t.ctypes.data_as(C.c_void_p),
                                          C.c_size_t(len(t) if n is None else n), g.ctypes.data_as(C.c_void_p),
                                          loss.ctypes.data_as(C.c_void_p))
    return rc, g, float(loss[0])


def test_noalloc_gradients_match_allocating_entry_point():
    import pcisph_amd
    lib = pcisph_amd.load_library()
    f = pcisph_amd.Fluid()
    f.mlp_set_params(M.init_params(11).astype(np.float32))
    rng = np.random.default_rng(12)
    for n in (MAX_ROWS, 3001):
        print(f"case mlp_gradients_noalloc n={n}", flush=True)
        x = rng.normal(size=(n, 8)).astype(np.float32)
        t = rng.normal(size=n).astype(np.float32)
        g_ref, l_ref = f.mlp_gradients(x, t)
        rc, g, loss = _noalloc(lib, f, x, t)
        assert rc == 0, f._error_message()
        scale = float(np.abs(g_ref).max())
        assert scale > 0
        np.testing.assert_allclose(g, g_ref, rtol=0, atol=1e-5 * scale)
        assert loss == pytest.approx(float(l_ref), rel=1e-5)
    print("case mlp_gradients_noalloc beyond the maximum", flush=True)
    x = np.zeros((MAX_ROWS + 1, 8), np.float32)
    t = np.zeros(MAX_ROWS + 1, np.float32)
    rc, _, _ = _noalloc(lib, f, x, t)
    assert rc == 2 and "65536" in f._error_message()  # PCISPH_ERR_SIZE
    f.close()


def test_noalloc_gradients_need_parameters():
    import pcisph_amd
    lib = pcisph_amd.load_library()
    f = pcisph_amd.Fluid()
    x = np.zeros((4, 8), np.float32)
    t = np.zeros(4, np.float32)
    rc, _, _ = _noalloc(lib, f, x, t)
    assert rc == 1 and "pcisph_mlp_set_params" in f._error_message()  # PCISPH_ERR_INVALID_ARGUMENT
    f.close()
