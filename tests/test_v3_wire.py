"""The V3 wire format (include/rsim_v3.h, rsim.v3), on the CPU: V3 arena and static dumps to and from ArenaState /
StaticData, as crates/rocketsim-cuda/src/dump.rs reads and writes them.

No V3 build or V3 dump exists offline, so the codec is pinned by:
  * FNV-1a 64's published test vectors, and the reference's format constants (header lengths 68 / 76, versions 2 / 3,
    BODY records of 219 bytes);
  * byte equality with an independent encoder (tests/v3_encoder.py) written from V3's own export code in
    v3/rocketsim-dump.patch, while the library follows the reference's reader: a field order or width that either side
    gets wrong is a byte difference;
  * round trips of stepped and mid-tick (phase-boundary) states from the oracle, and of every world preset;
  * the same physics after a round trip (the oracle steps the parsed state to the same words);
  * the reader's refusals, one hand-made dump per rule of dump.rs.
"""
import struct

import numpy as np
import pytest

import v3_encoder as enc


@pytest.fixture(scope="module")
def v3(rsim):
    from rsim import v3 as mod
    return mod


@pytest.fixture(scope="module")
def tpl(v3):
    """A template as a V3 dump would give it: opaque static body records (companion bytes clear), the handle counters."""
    t = v3.Template()
    t.config_hash = 0x0123456789ABCDEF
    t.num_handles = 24
    t.next_new_handle = 24
    t.manifold_stride = 0
    rng = np.random.default_rng(5)
    for k in range(v3.NUM_BODIES):
        rec = rng.integers(0, 256, v3.BODY_RECORD_LEN, dtype=np.uint8)
        rec[1:6] = 0
        for j in range(v3.BODY_RECORD_LEN):
            t.static_body[k][j] = int(rec[j])
    return t


@pytest.fixture(scope="module")
def stepped(rsim, oracle, world):
    """(label, state) pairs: contact-heavy and kickoff arenas after 1..10 ticks, and at every phase boundary of a tick."""
    out = []
    for scen, seed, n in ((rsim.SCEN_CONTACT, 7, 24), (rsim.SCEN_KICKOFF, 3, 4), (rsim.SCEN_OPEN, 11, 8)):
        st = rsim.sample_arenas(world, scen, seed, n)
        ctl = rsim.sample_controls(seed, 0, n)
        for i in range(n):
            s = st[i].copy()
            out.append(("scen%d/%d/start" % (scen, i), s.copy()))
            for tick in range(10):
                oracle.set_controls(s, ctl[i])
                oracle.step_tick(s, world)
                out.append(("scen%d/%d/tick%d" % (scen, i, tick + 1), s.copy()))
            oracle.set_controls(s, ctl[i])
            for k in range(10):
                oracle.phase(k, s, world)
                out.append(("scen%d/%d/phase%d" % (scen, i, k), s.copy()))
    return out


EVENTS = 2566  # ArenaState.events: the last tick's events, an output V3 does not export (cleared by the next P0)


def test_fnv1a64_published_vectors(v3):
    # the FNV authors' test vectors for FNV-1a 64
    for data, h in ((b"", 0xCBF29CE484222325), (b"a", 0xAF63DC4C8601EC8C), (b"foobar", 0x85944171F73967E8)):
        assert v3.fnv1a64(data) == h
        assert enc.fnv1a64(data) == h


def test_format_constants(rsim, v3):
    L = v3._lib()
    import ctypes
    # 8 + 23 * 219 + 24 bytes, u16 pair at 5070, u32 pair at 5076, 8-byte tail alignment: 5088 in C and in ctypes
    assert L.rsim_v3_template_bytes() == ctypes.sizeof(v3.Template) == 5088
    assert (v3.ARENA_HEADER_LEN, v3.STATIC_HEADER_LEN, v3.BODY_RECORD_LEN) == (68, 76, 219)
    # a BODY record of the independent encoder is the reference's 219 bytes
    from rsim import layout
    e = enc.Enc()
    enc.io_rigid_body(e, np.zeros((), layout.body))
    assert len(e.b) == 219


@pytest.mark.parametrize("presets", [(0, 0), (1, 5), (2, 6), (6, 3)])
def test_static_dump_matches_v3_writer_and_round_trips(rsim, v3, presets):
    w = rsim.build_world(*presets)
    b = v3.write_static(w)
    assert b == enc.static_dump(w)
    h = v3.static_header(b)
    assert h.hash == enc.fnv1a64(b[76:]) and struct.unpack_from("<Q", b, 20)[0] == len(b) - 76
    assert h.counts[:3] == [23, 3, 16] and h.counts[7:11] == [4, 2, 34, 242] and h.counts[3] == h.counts[6]
    assert np.array_equal(v3.parse_static(b), w)


def test_static_mutators_round_trip(rsim, v3):
    w = rsim.build_world(3, 4, gravity_z=-325.0, boost_accel_scale=1.5, ball_radius=120.0)
    assert np.array_equal(v3.parse_static(v3.write_static(w)), w)


def test_arena_dump_matches_v3_writer(v3, world, tpl, stepped):
    for label, s in stepped:
        assert v3.write_arena(s, world, tpl) == enc.arena_dump(s, world, tpl), label


def test_arena_round_trip(v3, world, tpl, stepped):
    kinds = np.zeros(6, int)
    for label, s in stepped:
        b = v3.write_arena(s, world, tpl)
        s2, t2 = v3.parse_arena(b, world)
        d = np.nonzero(s2 != s)[0]
        assert set(d.tolist()) <= {EVENTS} and s2[EVENTS] == 0, (label, d[:8])
        assert v3.write_arena(s2, world, t2) == b, label
        c = v3.arena_header(b).counts
        kinds += np.array([c[1], c[2], c[3], c[6], c[7], c[10]]) > 0
        assert t2.config_hash == tpl.config_hash and (t2.num_handles, t2.next_new_handle) == (24, 24)
        assert all(bytes(t2.static_body[k]) == bytes(tpl.static_body[k]) for k in range(23) if k < 20)
    # the sample covers every list the format carries
    assert kinds.min() > 0, kinds


def test_round_tripped_state_steps_the_same(rsim, v3, oracle, world, tpl, stepped):
    ctl = rsim.sample_controls(99, 0, 1)[0]
    for label, s in stepped[::7]:
        if "phase" in label:
            continue
        a = s.copy()
        b, _ = v3.parse_arena(v3.write_arena(s, world, tpl), world)
        for _ in range(3):
            for x in (a, b):
                oracle.set_controls(x, ctl)
                oracle.step_tick(x, world)
        assert np.array_equal(a, b), label


def test_parse_skips_v3_island_lists(v3, world, tpl, stepped):
    label, s = stepped[5]
    island = ([0, 1, 4294967295], [0, 1, 2], [-1, 7], [3], [1, 2], [])
    b = enc.arena_dump(s, world, tpl, island=island)
    s2, _ = v3.parse_arena(b, world)
    s_ref, _ = v3.parse_arena(v3.write_arena(s, world, tpl), world)
    assert np.array_equal(s2, s_ref)


def test_static_records_keep_the_first_seen_and_clear_companion(v3, world, tpl, stepped):
    from rsim import v3 as mod
    t = mod.Template()
    ctypes_copy = bytes(tpl)
    import ctypes
    ctypes.memmove(ctypes.addressof(t), ctypes_copy, len(ctypes_copy))
    t.static_body[0][1] = 1  # a companion_id set on a static body: cleared on read
    t.static_body[0][2] = 9
    s = stepped[3][1]
    _, t2 = v3.parse_arena(v3.write_arena(s, world, t), world)
    rec = bytes(t2.static_body[0])
    assert rec[1:6] == b"\0" * 5 and rec[6:] == bytes(tpl.static_body[0])[6:]


# ---- the record fold -------------------------------------------------------------------------------------------------------------
def _rec(a, b, normal=(0, 0, 1), swap=0, on_b=(1, 2, 3), on_a=(4, 5, 6)):
    r = np.zeros(12, np.uint32)
    r[0], r[1], r[2] = swap, a, b
    r[3:12] = np.array(list(on_b) + list(on_a) + list(normal), np.float32).view(np.uint32)
    return r


def test_fold_records(rsim, v3, world):
    from rsim import layout
    sd = layout.static_view(world)
    ball, car0, car1 = (int(x) for x in sd["dyn_body"])
    wall = next(i for i in range(23) if sd["bodies"][i]["user_idx"] == 0)
    s = np.zeros(2784, np.uint32)
    recs = np.stack([
        _rec(car0, ball), _rec(car1, wall, normal=(0.5, 0, 0.8)), _rec(ball, wall, normal=(0, 0, 1)),
        _rec(car1, car0, swap=1), _rec(car1, wall, normal=(0, 1, 0)), _rec(ball, wall, normal=(0.9, 0, 0.1)),
        _rec(wall, ball),  # a world-first record: a no-op in P7
    ])
    f = layout.arena_view(v3.fold_records(s, world, recs))
    assert int(f["n_records"]) == 2
    assert np.array_equal(f["records"][:2].view(np.uint32).reshape(2, 12), recs[[0, 3]])
    assert int(f["fold_events"]) == (1 << 11) | (1 << 12) | (1 << 14)  # car 1 hit world, ball hit world, ball wall
    assert tuple(f["fold_car_normal"][1].tolist()) == (0.0, 1.0, 0.0)  # the last car-world normal of car 1 wins
    assert tuple(f["fold_car_normal"][0].tolist()) == (0.0, 0.0, 0.0)
    # a floor-only ball contact leaves the wall bit clear (|n.z| >= 0.8)
    f2 = layout.arena_view(v3.fold_records(s, world, _rec(ball, wall, normal=(0, 0.6, 0.8))[None]))
    assert int(f2["fold_events"]) == 1 << 12
    # unfold then fold is the identity on the fold fields
    u = v3.unfold_records(f.base.reshape(-1) if f.base is not None else f, world)
    again = layout.arena_view(v3.fold_records(s, world, u))
    for k in ("fold_events", "fold_car_normal", "n_records", "records"):
        assert again[k].tobytes() == f[k].tobytes(), k
    # over the cap, and body indices out of range
    with pytest.raises(v3.FormatError, match="over the cap"):
        v3.fold_records(s, world, np.stack([_rec(car0, ball)] * 17))
    with pytest.raises(v3.FormatError, match="out of range"):
        v3.fold_records(s, world, _rec(23, ball)[None])


# ---- the reader's refusals -------------------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def contact_state(stepped):
    """A stepped contact arena holding manifolds, retained pairs and dynamic ray cells."""
    from rsim import layout
    for label, s in stepped:
        a = layout.arena_view(s)
        if "tick" in label and int(a["n_manifolds"]) >= 2 and int(a["n_pairs"]) >= 2 and int(a["n_dyn_cells"]) >= 2:
            return s
    raise AssertionError("no contact state in the sample")


def _refused(v3, data, world, match):
    with pytest.raises(v3.FormatError, match=match):
        v3.parse_arena(data, world)


def test_arena_header_and_framing_refusals(v3, world, tpl, contact_state):
    b = v3.write_arena(contact_state, world, tpl)
    _refused(v3, b"V3ARENA", world, "magic")
    _refused(v3, b"V3STATIC" + b[8:], world, "magic")
    _refused(v3, b[:8] + struct.pack("<I", 1) + b[12:], world, "unsupported arena dump version 1")
    _refused(v3, b[:-1], world, "truncated")
    _refused(v3, b + b"\0", world, "1 trailing bytes")
    _refused(v3, b[:68] + b"ARNX" + b[72:], world, "section tag ARNA expected")
    h = v3.arena_header(b)
    assert h.version == 2 and h.config_hash == tpl.config_hash and h.counts[0] == 23 and h.counts[8:10] == [2, 34]


@pytest.mark.parametrize("name,value,match", [
    ("ball_only", 1, "a ball-only arena"),
    ("ground_stick_applied", 1, "Snowday ground_stick_applied is set"),
    ("ball_world_records", 1, "ball-world contact records present|count exceeds"),
    ("applied_impulse", 2.5, "non-zero applied_impulse"),
    ("pair_key", 7, "pair_key is not min << 32 \\| max"),
    ("cbt", 0.123, "threshold is not the min"),
    ("free_list", [3], "free_list is not empty"),
    ("static_proxy_index", 999, "static broadphase proxy differs"),
])
def test_arena_field_refusals(v3, world, tpl, contact_state, name, value, match):
    b = enc.arena_dump(contact_state, world, tpl, mutate=lambda n, v: value if n == name else v)
    _refused(v3, b, world, match)


def test_arena_table_refusals(v3, world, tpl, contact_state):
    def mut_table(n, v):
        if n == "manifold_table":
            v = v.copy()
            v[np.flatnonzero(v)[0]] += 1
        return v
    _refused(v3, enc.arena_dump(contact_state, world, tpl, mutate=mut_table), world, "manifold table is not the search")

    def mut_pairs(n, v):
        if n == "pair_indices":
            v = np.append(v, [5]).astype(np.uint32)
        return v
    _refused(v3, enc.arena_dump(contact_state, world, tpl, mutate=mut_pairs), world, "pair index table")

    def mut_cells(n, v):
        return list(reversed(v)) if n == "dyn_cells" else v
    _refused(v3, enc.arena_dump(contact_state, world, tpl, mutate=mut_cells), world, "dynamic ray cells not ascending")


def test_arena_caps_refused(v3, world, tpl, contact_state):
    from rsim import layout
    for field, match in (("n_manifolds", "retained manifolds over the cap"), ("n_active", "active manifolds over the cap"),
                         ("n_pairs", "retained pairs over the cap")):
        s = contact_state.copy()
        layout.arena_view(s)[field] = 17
        with pytest.raises(v3.FormatError, match="over its cap"):
            v3.write_arena(s, world, tpl)  # the writer refuses a state over a cap
        # a dump whose list count is 17 (the encoder writes the count the state holds)
        _refused(v3, enc.arena_dump(s, world, tpl), world, match)


def test_static_refusals(rsim, v3):
    w = rsim.build_world(0, 0)
    b = v3.write_static(w)
    with pytest.raises(v3.FormatError, match="magic"):
        v3.parse_static(b"V3ARENA\0" + b[8:])
    with pytest.raises(v3.FormatError, match="unsupported static dump version 2"):
        v3.parse_static(b[:8] + struct.pack("<I", 2) + b[12:])
    with pytest.raises(v3.FormatError, match="payload length mismatch"):
        v3.parse_static(b + b"\0")
    bad = bytearray(b)
    bad[-1] ^= 1
    with pytest.raises(v3.FormatError, match="payload hash mismatch"):
        v3.parse_static(bytes(bad))
    # header counts: exact for the arena's fixed counts, at most the capacity for the mesh arrays
    for i, v, match in ((0, 22, "count 0 is 22"), (9, 33, "count 9 is 33"), (3, 8021, "count 3 is 8021"),
                        (6, 7503, "count 6"), (4, 100, "mesh totals differ")):
        bad = bytearray(b)
        struct.pack_into("<I", bad, 28 + 4 * i, v)
        with pytest.raises(v3.FormatError, match=match):
            v3.parse_static(bytes(bad))


def test_write_reports_size(v3, world, tpl, contact_state):
    L = v3._lib()
    import ctypes
    n = ctypes.c_uint64(0)
    s = np.ascontiguousarray(contact_state)
    rc = L.rsim_v3_write_arena(s.ctypes.data_as(ctypes.c_void_p), world.ctypes.data_as(ctypes.c_void_p), ctypes.byref(tpl),
                               None, 0, ctypes.byref(n))
    assert rc == -2 and n.value == len(v3.write_arena(s, world, tpl))


# ---- the data path end to end on the GPU ----------------------------------------------------------------------------------------
@pytest.mark.gpu
def test_v3_dumps_through_the_device(rsim, v3, oracle, tpl, device):
    """Static and arena dumps in, three decisions on the GPU, arena dumps out: the same bytes as the oracle stepping the
    original states and writing them (the path crates/rocketsim-cuda's README walks: parse_static, parse_arena, sim8,
    write_arena)."""
    print("[case] V3 dumps: parse, 3 x sim8 on the device, write; 256 contact arenas, presets (1, 4)", flush=True)
    w = rsim.build_world(1, 4)
    world = v3.parse_static(v3.write_static(w))
    n = 256
    st = rsim.sample_arenas(world, rsim.SCEN_CONTACT, 21, n)
    parsed = np.stack([v3.parse_arena(v3.write_arena(s, world, tpl), world)[0] for s in st])
    dw = device.upload_static(world)
    d = device.upload_states(parsed)
    r = device.alloc_words(n)
    cs = st.copy()
    for k in range(3):
        ctl = rsim.sample_controls(21, k, n)
        c = device.upload_controls(ctl)  # held until the decision has run
        device.sim8(d, dw, c, r, 64, None)
        device.sync()
        assert np.array_equal(device.read_words(r), oracle.sim8(cs, w, ctl, threads=8)), "decision %d results" % k
    gs = device.download_states(d)
    for i in range(n):
        assert v3.write_arena(gs[i], world, tpl) == v3.write_arena(cs[i], w, tpl), "arena %d" % i


# ---- robustness: damaged dumps end in a clean accept or a FormatError -----------------------------------------------------------
def _mutations(data, rng, n):
    """n damaged copies of `data`: byte flips, overwritten u32 counts, truncations, insertions and deletions."""
    out = []
    for _ in range(n):
        b = bytearray(data)
        kind = rng.integers(0, 5)
        if kind == 0:
            for _ in range(int(rng.integers(1, 8))):
                b[int(rng.integers(0, len(b)))] ^= int(rng.integers(1, 256))
        elif kind == 1:
            at = int(rng.integers(0, len(b) - 4)) & ~3
            struct.pack_into("<I", b, at, int(rng.choice([0, 1, 17, 65, 255, 4096, 0x7FFFFFFF, 0xFFFFFFFF])))
        elif kind == 2:
            b = b[:int(rng.integers(0, len(b)))]
        elif kind == 3:
            at = int(rng.integers(0, len(b)))
            b[at:at] = bytes(rng.integers(0, 256, int(rng.integers(1, 16)), dtype=np.uint8))
        else:
            at = int(rng.integers(0, len(b)))
            del b[at:at + int(rng.integers(1, 16))]
        out.append(bytes(b))
    return out


def test_damaged_arena_dumps_fail_cleanly(v3, world, tpl, contact_state):
    rng = np.random.default_rng(11)
    good = v3.write_arena(contact_state, world, tpl)
    accepted = 0
    for b in _mutations(good, rng, 1500):
        try:
            s, t = v3.parse_arena(b, world)
        except v3.FormatError:
            continue
        accepted += 1
        v3.write_arena(s, world, t)  # whatever parses can be written back
    assert accepted < 1500


def test_damaged_static_dumps_fail_cleanly(rsim, v3):
    rng = np.random.default_rng(12)
    good = bytearray(v3.write_static(rsim.build_world(0, 0)))
    for b in _mutations(bytes(good), rng, 300):
        b = bytearray(b)
        if len(b) >= 76:  # re-seal the payload so the damage reaches the section reader, not only the hash check
            struct.pack_into("<QQ", b, 12, v3.fnv1a64(bytes(b[76:])), len(b) - 76)
        try:
            v3.parse_static(bytes(b))
        except v3.FormatError:
            pass
