"""The launch schedules the autotuner chooses between, each checked word for word against the oracle:
  - rsim_sim8_split: one decision as launches of 1, 2, 3 or 8 ticks (the state back in HBM between launches), including
    the result word's rare outcomes (a goal on each tick 1..8, the overflow and unimplemented flags), whose bits have to
    survive the launch boundaries;
  - rsim_run_queued: persistent work-groups taking AoSoA blocks from a queue, in index order, in a given permutation and
    in the cost model's heaviest-first block order, with few and with as many work-groups as fit, block sizes 64-256,
    a ragged arena count and a per-arena world table.
Results never depend on the schedule (arenas are independent; reference: crates/rocketsim-cuda/tests/step.rs:455-506)."""
import sys

import numpy as np
import pytest

import forced_cases as fc
from conftest import first_diff
from helpers import TICK_WORDS

pytestmark = pytest.mark.gpu


def note(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def cpu_decisions(oracle, world, states, controls):
    s = states.copy()
    return s, [oracle.sim8(s, world, c, threads=8) for c in controls]


def cpu_decisions_in(oracle, worlds, variant, states, controls):
    s = states.copy()
    return s, [oracle.sim8_in(s, worlds, variant, c, threads=8) for c in controls]


def check(got_states, got_results, cs, cr, what):
    for k, (a, b) in enumerate(zip(got_results, cr)):
        assert np.array_equal(a, b), "%s: decision %d, arena %d: gpu %#x cpu %#x" % (
            what, k, int(np.argmax(a != b)), a[np.argmax(a != b)], b[np.argmax(a != b)])
    fd = first_diff(got_states[:, :TICK_WORDS], cs[:, :TICK_WORDS])
    assert fd is None, "%s: arena %d word %d differs" % (what, fd[0], fd[1])


def split_run(device, dw, states, controls, tpl, variant=None, block=64):
    n = states.shape[0]
    d = device.upload_states(states)
    r = device.alloc_words(n)
    out = []
    for c in controls:
        dc = device.upload_controls(c)
        device.sim8_split(d, dw, dc, r, tpl, variant=variant, block=block)
        device.sync()
        out.append(device.read_words(r))
    return device.download_states(d), out


@pytest.mark.parametrize("tpl", [1, 2, 3, 8])
def test_split_decisions_equal_the_oracle(rsim, oracle, world, device, dev_world, tpl):
    note("sim8_split %d ticks per launch, 1,372 contact + goal-frame arenas and 128 scoring on ticks 1-8, 3 decisions" % tpl)
    goals, _ = fc.goal_states(rsim, world, 7)
    states = np.concatenate([rsim.sample_arenas(world, rsim.SCEN_CONTACT, 301, 700),
                             rsim.sample_arenas(world, rsim.SCEN_GOALFRAME, 302, 1500 - 700 - goals.shape[0]), goals])
    n = states.shape[0]
    ctl = [rsim.sample_controls(303, d, n) for d in range(3)]
    gs, gr = split_run(device, dev_world, states, ctl, tpl)
    cs, cr = cpu_decisions(oracle, world, states, ctl)
    assert any(((r & 0xFF) > 0).any() for r in cr), "the batch scores goals"
    check(gs, gr, cs, cr, "split %d" % tpl)


@pytest.mark.parametrize("tpl", [1, 3])
def test_split_keeps_the_rare_result_bits(rsim, oracle, world, device, dev_world, tpl):
    note("sim8_split %d ticks per launch on the goal-per-tick, overflow and unimplemented cases" % tpl)
    st, _ = fc.goal_states(rsim, world, 7)  # a goal on each tick 1..8 at both goals
    gs, gr = split_run(device, dev_world, st, [np.zeros((st.shape[0], 16), np.uint32)], tpl)
    cs, cr = cpu_decisions(oracle, world, st, [np.zeros((st.shape[0], 16), np.uint32)])
    assert np.array_equal(cr[0] & 0xFF, np.repeat(np.arange(1, 9, dtype=np.uint32), 16))
    check(gs, gr, cs, cr, "split %d goals" % tpl)
    st = fc.overflow_states(rsim, world, 512, 17)
    ctl = [rsim.sample_controls(18, d, 512) for d in range(2)]
    gs, gr = split_run(device, dev_world, st, ctl, tpl)
    cs, cr = cpu_decisions(oracle, world, st, ctl)
    assert sum(int(((r & 0x40000000) != 0).sum()) for r in cr) > 50
    check(gs, gr, cs, cr, "split %d overflow" % tpl)
    uw = fc.unimplemented_world(world)
    st = rsim.sample_arenas(world, rsim.SCEN_CONTACT, 21, 512)
    ctl = [rsim.sample_controls(22, 0, 512)]
    gs, gr = split_run(device, device.upload_static(uw), st, ctl, tpl)
    cs, cr = cpu_decisions(oracle, uw, st, ctl)
    assert ((cr[0] & 0x80000000) != 0).sum() > 50
    check(gs, gr, cs, cr, "split %d unimplemented" % tpl)


def test_split_with_a_world_table(rsim, oracle, device):
    note("sim8_split 2 ticks per launch, 900 arenas over 7 pairings, an index past the table")
    worlds = np.stack([rsim.build_world(p, p) for p in range(7)])
    n = 900
    variant = (np.arange(n) % 7).astype(np.uint32)
    states = rsim.sample_arenas_in(worlds, variant, rsim.SCEN_OPEN_CONTACT, 311)
    ctl = [rsim.sample_controls(312, d, n) for d in range(2)]
    variant_dev = variant.copy()
    variant_dev[5] = 7  # past the table: untouched, flagged
    dw, dv = device.upload_worlds(worlds), device.upload_words(variant_dev)
    gs, gr = split_run(device, dw, states, ctl, 2, variant=dv)
    keep = np.arange(n) != 5
    cs, cr = cpu_decisions_in(oracle, worlds, variant[keep], states[keep], [c[keep] for c in ctl])
    assert all(r[5] == 0x20000000 for r in gr) and np.array_equal(gs[5, :TICK_WORDS], states[5, :TICK_WORDS])
    check(gs[keep], [r[keep] for r in gr], cs, cr, "split with worlds")


def queued_run(device, dw, states, controls, block, order=None, workgroups=0, variant=None):
    n, D = states.shape[0], len(controls)
    d = device.upload_states(states)
    c = device.upload_words(np.concatenate(controls))
    r = device.alloc_words(D * n)
    device.run_queued(d, dw, c, r, D, variant=variant, block=block, order=order, workgroups=workgroups)
    device.sync()
    return device.download_states(d), list(device.read_words(r).reshape(D, n))


@pytest.mark.parametrize("block,workgroups", [(64, 0), (64, 3), (128, 0), (256, 2)])
def test_queued_run_equals_the_oracle(rsim, oracle, world, device, dev_world, block, workgroups):
    note("run_queued block %d workgroups %d, 2,049 arenas of the 1:1:1 mix, 3 decisions, three queue orders" % (block, workgroups))
    n, D = 2049, 3  # 33 AoSoA blocks, the last with one arena
    states = rsim.sample_arenas(world, rsim.SCEN_MIXED, 321, n)
    ctl = [rsim.sample_controls(322, d, n) for d in range(D)]
    cs, cr = cpu_decisions(oracle, world, states, ctl)
    nb = (n + 63) // 64
    rng = np.random.default_rng(323)
    for name, perm in (("index order", None), ("reversed", np.arange(nb)[::-1].copy()), ("shuffled", rng.permutation(nb))):
        order = None
        if perm is not None:
            order = device.alloc_words(nb)
            p32 = np.ascontiguousarray(perm, np.uint32)
            rsim._check(rsim.lib().rsim_set_permutation(device.handle, rsim._ptr(p32), nb, order.handle), "set_permutation")
        gs, gr = queued_run(device, dev_world, states, ctl, block, order, workgroups)
        check(gs, gr, cs, cr, "run_queued %s" % name)


def test_queued_run_in_cost_order_with_worlds(rsim, oracle, device):
    note("run_queued in the cost model's block order, 1,300 arenas in 5 worlds, 2 decisions")
    import rsim.cost as cm
    worlds = np.stack([rsim.build_world(p, (p + 2) % 7) for p in range(5)])
    n, D = 1300, 2
    variant = (np.arange(n) * 3 % 5).astype(np.uint32)
    states = rsim.sample_arenas_in(worlds, variant, rsim.SCEN_OPEN_CONTACT, 331)
    ctl = [rsim.sample_controls(332, d, n) for d in range(D)]
    dw, dv = device.upload_worlds(worlds), device.upload_words(variant)
    d_states = device.upload_states(states)
    model = cm.CostModel(device, seed=3)
    pred = model.predict(model.features(d_states), n)
    work = device.alloc_words(int(rsim.lib().rsim_cost_block_order_work_words(n)))
    order = device.alloc_words((n + 63) // 64)
    rsim._check(rsim.lib().rsim_cost_block_order(device.handle, pred.handle, n, work.handle, order.handle), "block order")
    device.sync()
    o = device.read_words(order)
    p = device.read_words(pred, 2 * n).view(np.float32).reshape(n, 2)[:, 0]
    assert np.array_equal(np.sort(o), np.arange((n + 63) // 64, dtype=np.uint32))
    block_max = np.array([np.nanmax(p[b * 64:(b + 1) * 64]) for b in range((n + 63) // 64)], np.float32)
    assert np.array_equal(o, cm.order_by_cost(block_max, "sorted", wave=1 << 30)), "heaviest block first, stable"
    gs, gr = queued_run(device, dw, states, ctl, 64, order, 0, variant=dv)
    cs, cr = cpu_decisions_in(oracle, worlds, variant, states, ctl)
    check(gs, gr, cs, cr, "run_queued cost order")
