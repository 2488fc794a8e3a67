"""Rules of the C ABI that protect callers (include/rsim.h): calls that allocate, free, copy to or from the host or
synchronise refuse while the device's stream records a graph (and the capture survives the refusal), and rsim_close
refuses while worlds, buffers, views or graphs of the device are open. Called through ctypes, as a C caller would."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_calls_refused_while_capturing(rsim, world, device, dev_world):
    print("[case] allocation, copies, frees and syncs refused during a capture", flush=True)
    L = rsim.lib()
    n = 256
    states = rsim.sample_arenas(world, rsim.SCEN_CONTACT, 901, n)
    ctl = device.upload_controls(rsim.sample_controls(902, 0, n))
    want_d = device.upload_states(states)
    want_r = device.alloc_words(n)
    device.sim8(want_d, dev_world, ctl, want_r)
    device.sync()
    d = device.upload_states(states)
    r = device.alloc_words(n)
    spare = device.alloc_words(8)
    host = np.zeros(8, np.uint32)
    device.capture_begin()
    h = ctypes.c_void_p()
    refused = {
        "rsim_alloc_words": L.rsim_alloc_words(device.handle, 8, ctypes.byref(h)),
        "rsim_alloc_states": L.rsim_alloc_states(device.handle, 4, ctypes.byref(h)),
        "rsim_read_words": L.rsim_read_words(device.handle, spare.handle, host.ctypes.data_as(ctypes.c_void_p), 8),
        "rsim_write_words": L.rsim_write_words(device.handle, spare.handle, host.ctypes.data_as(ctypes.c_void_p), 8),
        "rsim_words_free": L.rsim_words_free(spare.handle),
        "rsim_sync": L.rsim_sync(device.handle),
        "rsim_close": L.rsim_close(device.handle),
    }
    device.sim8(d, dev_world, ctl, r)  # asynchronous work is still recorded
    graph = device.capture_end()
    assert all(rc == -1 for rc in refused.values()), refused
    graph.launch()
    device.sync()
    assert np.array_equal(device.read_words(r), device.read_words(want_r))
    assert np.array_equal(device.download_states(d)[:, :2567], device.download_states(want_d)[:, :2567])
    assert np.array_equal(device.read_words(spare), np.zeros(8, np.uint32))  # the refused free left the buffer alone
    graph.free()


def test_close_refused_while_objects_are_open(rsim, world):
    print("[case] rsim_close refused while a world, buffers, a view or a graph are open", flush=True)
    L = rsim.lib()
    dev = ctypes.c_void_p()
    assert L.rsim_open(0, ctypes.byref(dev)) == 0
    w = np.ascontiguousarray(world)
    hw, hs, hs2, hb, hv, hg = (ctypes.c_void_p() for _ in range(6))
    assert L.rsim_upload_world(dev, w.ctypes.data_as(ctypes.c_void_p), ctypes.byref(hw)) == 0
    assert L.rsim_alloc_states(dev, 3, ctypes.byref(hs)) == 0
    assert L.rsim_alloc_states(dev, 3, ctypes.byref(hs2)) == 0
    assert L.rsim_alloc_words(dev, 64, ctypes.byref(hb)) == 0
    assert L.rsim_words_view(hb, 8, 8, ctypes.byref(hv)) == 0
    assert L.rsim_capture_begin(dev) == 0
    assert L.rsim_copy_states(dev, hs2, hs) == 0
    assert L.rsim_capture_end(dev, ctypes.byref(hg)) == 0
    for free, h in ((L.rsim_graph_free, hg), (L.rsim_words_free, hv), (L.rsim_words_free, hb), (L.rsim_states_free, hs),
                    (L.rsim_states_free, hs2), (L.rsim_world_free, hw)):
        assert L.rsim_close(dev) == -1, "closed with objects open"
        assert b"still open" in L.rsim_last_error()
        assert free(h) == 0
    assert L.rsim_close(dev) == 0
