"""Host-side checks of librsim.so (no GPU needed): layout constants, the AoSoA-64 index and pack, the exported C ABI, and
argument validation without a device."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("rsim.h", "rsim_v3.h")]


def test_layout_constants(rsim):
    L = rsim.lib()
    assert L.rsim_arena_words() == 2784 and L.rsim_tick_words() == 2567 and L.rsim_static_words() == 317895
    assert rsim.aosoa_words(0) == 0
    assert rsim.aosoa_words(1) == rsim.aosoa_words(64) == 2784 * 64
    assert rsim.aosoa_words(65) == 2 * 2784 * 64
    assert rsim.aosoa_words(65536) * 4 == 729808896


def test_aosoa_index_closed_form(rsim):
    rng = np.random.default_rng(0)
    for a, w in zip(rng.integers(0, 1 << 20, 2000), rng.integers(0, 2784, 2000)):
        assert rsim.aosoa_index(int(a), int(w)) == (a // 64) * 2784 * 64 + w * 64 + a % 64
    # a block's words of one lane are 64 apart; the 64 lanes of one word are contiguous
    assert rsim.aosoa_index(5, 1) - rsim.aosoa_index(5, 0) == 64
    assert rsim.aosoa_index(6, 7) - rsim.aosoa_index(5, 7) == 1


@pytest.mark.parametrize("n", [1, 63, 64, 65, 200])
def test_pack_round_trip(rsim, n):
    rng = np.random.default_rng(n)
    s = rng.integers(0, 2**32, (n, 2784), dtype=np.uint64).astype(np.uint32)
    p = rsim.aosoa_pack(s)
    assert p.size == rsim.aosoa_words(n)
    assert np.array_equal(rsim.aosoa_unpack(p, n), s)
    # brute-force placement and zero padding lanes
    for a in (0, n - 1, n // 2):
        for w in (0, 1, 2566, 2783):
            assert p[(a // 64) * 2784 * 64 + w * 64 + a % 64] == s[a, w]
    pad = p.reshape(-1, 2784, 64)[-1][:, (n % 64 or 64):]
    assert not pad.any()


def declared_functions():
    text = "".join(open(h).read() for h in HEADERS)
    return sorted(set(re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(rsim_\w+)\s*\(", text, re.M)))


def test_header_symbols_exported(rsim):
    names = declared_functions()
    assert len(names) >= 50 and "rsim_v3_parse_arena" in names, names
    L = ctypes.CDLL(rsim.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, "declared in include/ but not exported: %s" % missing


def test_errors_without_device(rsim):
    L = rsim.lib()
    # null handles are rejected, not dereferenced
    assert L.rsim_sync(None) == -1
    assert L.rsim_sim8(None, None, None, None, None, 64, None) == -1
    assert L.rsim_close(None) == 0
    assert b"null" in L.rsim_last_error()
    assert L.rsim_build_world(9, 0, -650.0, 1.0, 91.25, None) == -1
    assert L.rsim_sample_arenas(None, 0, 0, 0, 1, None) == -1


def test_no_cpu_fallback(rsim):
    """Without a GPU, opening a device fails loudly; nothing steps on the CPU behind the API."""
    import subprocess, sys
    code = ("import sys; sys.path.insert(0, %r); import rsim\n"
            "try:\n    rsim.SimDevice(0)\nexcept rsim.RsimError as e:\n    print('ERR', e)\nelse:\n    print('GPU')\n"
            % os.path.join(ROOT, "rocketsim-cuda_amd"))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120).stdout
    assert out.startswith("ERR") or out.startswith("GPU")
    if out.startswith("ERR"):
        assert "no GPU" in out or "no such GPU" in out


def test_world_builder_same_in_both_builds(rsim, oracle):
    """The product's builder (librsim.so) and the oracle's copy of it produce identical bytes."""
    for p in ((0, 0), (6, 2)):
        assert np.array_equal(rsim.build_world(*p), oracle.build_world(*p))
    w = rsim.build_world(0, 0)
    for sc in range(7):
        assert np.array_equal(rsim.sample_arenas(w, sc, 5, 40), oracle.sample_arenas(w, sc, 5, 40))
    assert np.array_equal(rsim.sample_controls(1, 2, 50), oracle.sample_controls(1, 2, 50))
    # subranges are reproducible on their own
    assert np.array_equal(rsim.sample_arenas(w, 5, 9, 10, first=30), rsim.sample_arenas(w, 5, 9, 40)[30:])


def test_named_layout_matches_the_c_layout(rsim, oracle, world):
    """rsim.layout (the field names of the reference's layout.rs) lays the words out as the C structs do: the landmark
    offsets the C header asserts, the offsets the oracle's C build reports, and values read through the names."""
    from rsim import layout as L
    A, S = L.ARENA, L.STATIC
    assert A.itemsize == 2784 * 4 and S.itemsize == 317895 * 4 and L.TICK_WORDS == 2567
    off = lambda dt, name: dt.fields[name][1]
    assert (off(A, "cars"), off(A, "bodies"), off(A, "manifolds"), off(A, "edges"), off(A, "events")) == (
        96, 2888, 3552, 9532, 10264)
    assert (off(S, "tris"), off(S, "nodes"), off(S, "leaves"), off(S, "table_sin")) == (146044, 531004, 981484, 1206044)
    lay = np.zeros(8, np.uint64)
    oracle.L.rsim_cpu_static_layout(lay.ctypes.data_as(ctypes.c_void_p))
    assert [int(x) for x in lay] == [off(S, "meshes"), L.mesh_info.itemsize, off(L.mesh_info, "node_first"),
                                     off(L.mesh_info, "node_count"), off(S, "nodes"), L.wide_node.itemsize, 16, 3754]
    for car in (0, 1):
        assert oracle.controls_word[car] * 4 == off(A, "cars") + car * L.car.itemsize + off(L.car, "controls")
    sv = L.static_view(world)
    assert list(sv["dyn_body"]) == [20, 21, 22] and sv["ball_body"] == 20
    assert sv["table_sin"][4096] == np.float32(1.0) and sv["table_sin"][0] == 0.0  # sin(pi / 2), sin(0)
    st = rsim.sample_arenas(world, rsim.SCEN_KICKOFF, 1, 4)
    a = L.arena_view(st)
    assert a.shape == (4,) and np.all(a["ball"]["pos"]["x"] == 0) and np.all(a["ball"]["pos"]["y"] == 0)
    assert np.allclose(a["cars"]["boost"], 100.0 / 3.0) and np.all(a["n_active"] == 0)
    a["cars"][0, 1]["boost"] = 7.5  # a write through the view lands in the words
    assert st[0].view(np.float32)[(off(A, "cars") + L.car.itemsize + off(L.car, "boost")) // 4] == 7.5
    leaves = L.arena_leaves()
    assert len(leaves) == 2784 and leaves[2566] == ("events", "i") and leaves[24] == ("cars.dodge_deadzone", "f")
    assert leaves[4:7] == [("ball.pos", "f")] * 3 and leaves[-1] == ("records.point.normal_world_on_b", "f")
