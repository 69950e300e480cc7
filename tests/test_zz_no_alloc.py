"""xpbd_mlp_forward_noalloc: the network forward that allocates no device
memory (rows staged through a fixed file-scope device array), called through
the C-ABI and compared with xpbd_mlp_forward, the entry point it is a variant
of: the same kernel, so the same bits, on two calls in a row (the second with
fewer rows, so rows left over in the shared array from the first call must not
leak into it), and the size limit refused with a clear error."""
import ctypes as C

import numpy as np
import pytest

import xpbd_amd as X

pytestmark = pytest.mark.gpu
MAX_ROWS = 65536


def _network(seed):
    rng = np.random.default_rng(seed)
    net = X.Network(rng.normal(0.0, 0.2, X.MLP_PARAMS).astype(np.float32))
    net.set_scaling(rng.normal(size=X.MLP_IN).astype(np.float32), rng.uniform(0.5, 2.0, X.MLP_IN).astype(np.float32),
                    3e-4)
    return net


def _noalloc(net, x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.full((len(x), 3), np.nan, np.float32)
    rc = X.lib.xpbd_mlp_forward_noalloc(net.m, len(x), x.ctypes.data_as(C.POINTER(C.c_float)),
                                        out.ctypes.data_as(C.POINTER(C.c_float)))
    return rc, out


def test_noalloc_forward_equals_forward_on_two_calls():
    net = _network(3)
    rng = np.random.default_rng(4)
    for n in (MAX_ROWS, 1000):
        x = rng.normal(size=(n, X.MLP_IN)).astype(np.float32) * 2.0
        rc, got = _noalloc(net, x)
        assert rc == 0, X.lib.xpbd_mlp_last_error(net.m)
        ref = net.forward(x)
        assert np.array_equal(got, ref), (n, float(np.max(np.abs(got - ref))))


def test_noalloc_forward_refuses_more_rows_than_its_scratch():
    net = _network(5)
    x = np.zeros((MAX_ROWS + 1, X.MLP_IN), np.float32)
    rc, _ = _noalloc(net, x)
    assert rc == -1  # XPBD_E_INVALID_ARG
    assert b"exceed" in X.lib.xpbd_mlp_last_error(net.m)
    # the network still works afterwards, with and without allocation
    rc, got = _noalloc(net, x[:7])
    assert rc == 0 and np.array_equal(got, net.forward(x[:7]))
