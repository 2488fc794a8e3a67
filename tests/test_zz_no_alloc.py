"""The allocation-free cost order (rsim_cost_order_noalloc) against rsim_cost_order with a caller-provided work buffer,
through the C ABI: two calls in a row on the fixed workspace, the second with a smaller batch, in both modes; the size
limit is refused without touching the GPU."""
import sys

import numpy as np
import pytest

import rsim.cost as cm

pytestmark = pytest.mark.gpu


def note(msg):  # names the case before its GPU work, so a failing run says which
    print(msg, flush=True)
    sys.stdout.flush()


def _preds(rng, n):
    pred = np.zeros((n, 2), np.float32)
    pred[:, 0] = (rng.integers(-40, 240, n) / 64.0).astype(np.float32) + rng.random(n).astype(np.float32) * 0.01
    pred[rng.random(n) < 0.01, 0] = np.nan
    return pred


def test_noalloc_order_equals_the_work_buffer_order(rsim, device):
    lib = rsim.lib()
    limit = 262144
    rng = np.random.default_rng(29)
    for n in (limit, 4099):  # the largest batch first, then a smaller one over the counts the first left behind
        note("rsim_cost_order_noalloc: n = %d" % n)
        pred = _preds(rng, n)
        d_pred = device.upload_words(pred.view(np.uint32).ravel())
        work = device.alloc_words(int(lib.rsim_cost_order_work_words(n)))
        want_perm = device.alloc_words(n)
        got_perm = device.alloc_words(n)
        for mode, code in cm.ORDER_MODES.items():
            rsim._check(lib.rsim_cost_order(device.handle, d_pred.handle, n, code, work.handle, want_perm.handle), "order")
            rsim._check(lib.rsim_cost_order_noalloc(device.handle, d_pred.handle, n, code, got_perm.handle), "order (fixed)")
            device.sync()
            want, got = device.read_words(want_perm), device.read_words(got_perm)
            assert np.array_equal(got, want), (n, mode, np.flatnonzero(got != want)[:5])
            assert np.array_equal(np.sort(got), np.arange(n, dtype=np.uint32))  # a permutation
            assert np.array_equal(got, cm.order_by_cost(pred[:, 0], mode))  # and the host statement's order


def test_noalloc_order_refuses_batches_past_the_limit(rsim, device):
    lib = rsim.lib()
    pred = device.alloc_words(2)
    perm = device.alloc_words(1)
    rc = lib.rsim_cost_order_noalloc(device.handle, pred.handle, 262145, 0, perm.handle)
    assert rc == -1 and b"RSIM_COST_ORDER_NOALLOC_MAX" in lib.rsim_last_error()
    assert lib.rsim_cost_order_noalloc(device.handle, pred.handle, 2, 0, perm.handle) == -2  # short buffers
    assert lib.rsim_cost_order_noalloc(device.handle, pred.handle, 1, 7, perm.handle) == -1  # unknown mode
