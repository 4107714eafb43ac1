"""mcx_mlp_forward_noalloc (inputs and outputs staged through the library's
fixed device array, no allocation) against mcx_mlp_forward on the same
network: bit-identical outputs on two calls in a row, the second smaller,
and a clear refusal beyond the array's size."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_forward_noalloc_equals_forward():
    from mcmax_amd import _lib, mlp
    net = mlp.Network(0, seed=12)
    lib = net.lib
    try:
        for n in (50000, 777):
            rng = np.random.default_rng(n)
            x = np.ascontiguousarray(rng.normal(size=(n, mlp.NIN)), dtype=np.float32)
            ref = net.forward_raw(x)
            y = np.full(n, np.nan, dtype=np.float32)
            rc = lib.mcx_mlp_forward_noalloc(net.h, n, _lib.ptr(x), _lib.ptr(y))
            assert rc == 0, lib.mcx_mlp_last_error(net.h).decode()
            assert np.array_equal(y, ref), n
        big = _lib.MLP_NOALLOC_MAXROWS + 1
        x = np.zeros((big, mlp.NIN), dtype=np.float32)
        y = np.zeros(big, dtype=np.float32)
        rc = lib.mcx_mlp_forward_noalloc(net.h, big, _lib.ptr(x), _lib.ptr(y))
        assert rc == _lib.MCX_EINVAL
        assert "exceed" in lib.mcx_mlp_last_error(net.h).decode()
    finally:
        net.close()
