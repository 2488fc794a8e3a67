"""Benchmark of the 8-tick decision path on one MI355X: arena-ticks/s over 65,536 contact-heavy arenas.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--block B]

A step is one decision (8 ticks at 120 Hz, sim8) of all 65,536 arenas, with fresh per-car controls each decision, states
resident in HBM (AoSoA-64). The timed region is K back-to-back decisions on the device stream, bracketed by device
synchronisations and HIP events on that stream. Ticks after a goal are not stepped, so `value` counts the ticks actually
stepped (the scored tick ends an arena's decision); the nominal 8-tick count is reported beside it.

Workload: the contact-heavy mix (40% ball pinched against a wall or fillet with a car driving into it, 30% head-on cars,
30% ball at the goal frame), Octane vs Octane on the procedural Soccar arena, seeded synthetic states and controls.

Prints one JSON line. `roofline` prices the sim8 kernel against HBM at 20,544.5 algorithmic bytes per arena-tick (the
2,567 persistent state words read and written once per tick, plus controls and result amortised over the decision).
`cpu_baseline` is the reference's own CPU build of the path (its phase.cpp compiled in place, when built with the tree;
otherwise the test oracle, bit-identical to it) on the host's cores over the arenas of the same run; its outputs are
word-compared with the GPU's (`word_diffs`).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rocketsim-cuda_amd"))

ARENAS = 65536
B_TICK = 2 * 2567 * 4 + (64 + 4) / 8  # algorithmic HBM bytes per arena-tick
HBM_PEAK = 8.0e12
METRIC = "arena-ticks/s @65,536 arenas, plain vs cost-sorted; 0 word diffs vs CPU"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "sim8_contact65536_traffic.json")


def host_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline_leg(world, states0, controls, gpu_states, gpu_results, sample, threads):
    """The CPU path over arenas [0, sample) of the same run, timed, and word-compared with the GPU's outputs. It is the
    reference's own C++ build (crates/rocketsim-cuda/csrc/phase.cpp compiled in place into oracle/_ref, driven over host
    threads) when that library was built with the tree, otherwise the oracle (the same tick built from this project's
    core, bit-identical to the reference build by the tests)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    if os.path.exists(orc.REF_LIB):
        cpu, kind, what = orc.Reference(), "reference", "the reference's own C++ build (phase.cpp, g++ -O2 -ffp-contract=off)"
    else:
        if not os.path.exists(orc.ORACLE_LIB):
            orc.build(ref=False)
        cpu, kind, what = orc.Oracle(), "port", "CPU build of the same tick (g++ -O2 -ffp-contract=off)"
    s = np.ascontiguousarray(states0[:sample])
    ticks = 0
    diffs = 0
    t0 = time.perf_counter()
    res = []
    for c in controls:
        r = cpu.sim8(s, world, np.ascontiguousarray(c[:sample]), threads=threads)
        res.append(r)
    dt = time.perf_counter() - t0
    for k, r in enumerate(res):
        sc = r & 0xFF
        ticks += int(np.where(sc > 0, sc, 8).sum())
        diffs += int((r != gpu_results[k][:sample]).sum())
    diffs += int((s[:, :2567] != gpu_states[:sample, :2567]).sum())
    return {"value": ticks / dt, "unit": "arena-ticks/s", "cores": threads, "kind": kind,
            "sample": "arenas 0..%d of the same run over all %d timed decisions (%d arena-ticks), %d host threads, %s"
                      % (sample - 1, len(controls), ticks, threads, what),
            "seconds": dt}, diffs


def stepped_ticks(results):
    return sum(int(np.where((r & 0xFF) > 0, r & 0xFF, 8).sum()) for r in results)


LABEL_ARENAS, LABEL_DECISIONS = 65536, 21  # per scenario mix: 3 x 65,536 x 21 = 4,128,768 labelled arena-decisions


def train_cost_model(rsim, dev, world_host, dw):
    """The cost model trained on labels the simulator emits: 65,536 arenas of each of three scenario mixes (seeds apart
    from every timed batch) stepped 21 decisions with one arena per wave and each lane's clock recorded, 4.13 M
    arena-decisions in all."""
    import rsim.cost as cm

    t0 = time.perf_counter()
    F, T, C = [], [], []
    for scen, seed in ((rsim.SCEN_CONTACT, 77), (rsim.SCEN_OPEN_CONTACT, 79), (rsim.SCEN_MIXED, 81)):
        f, t, c = cm.label_costs(dev, dw, rsim.sample_arenas(world_host, scen, seed, LABEL_ARENAS),
                                 [rsim.sample_controls(seed + 1, d, LABEL_ARENAS) for d in range(LABEL_DECISIONS)],
                                 decisions=LABEL_DECISIONS)
        F.append(f), T.append(t), C.append(c)
    feats, targets, contacts = np.concatenate(F), np.concatenate(T), np.concatenate(C)
    label_s = time.perf_counter() - t0
    model = cm.CostModel(dev, seed=1)
    curve = model.train(feats, targets, contacts, epochs=15, batch=4096, lr=3e-3)
    # quality on held-out arenas: Pearson correlation of predicted and measured log cost, trained vs untrained
    f, t, _ = cm.label_costs(dev, dw, rsim.sample_arenas(world_host, rsim.SCEN_OPEN_CONTACT, 91, 4096),
                             [rsim.sample_controls(92, d, 4096) for d in range(3)], decisions=3)
    d_f = dev.upload_words(f.view(np.uint32).ravel())

    def corr(m):
        out = m.predict(d_f, f.shape[0])
        dev.sync()
        return float(np.corrcoef(dev.read_words(out, 2 * f.shape[0]).view(np.float32).reshape(-1, 2)[:, 0], t)[0, 1])

    info = {"samples": int(feats.shape[0]), "epochs": 15, "loss_first": curve[0], "loss_last": curve[-1],
            "label_seconds": label_s, "seconds": time.perf_counter() - t0,
            "heldout_pearson_log_cost": {"trained": corr(model), "untrained": corr(cm.CostModel(dev, seed=1))}}
    return model, info


def compare_orders(rsim, dev, dw, model, states_host, controls, block, plain=None):
    """K decisions of one batch in the plain order (per-launch sim8) and regrouped by predicted cost before every decision
    ('sorted': cost-homogeneous waves, most expensive first; 'balanced': each wave dealt one arena of every cost rank),
    and balanced regrouping before every 8th or 32nd decision only ('balanced_every8', 'balanced_every32': the
    regrouping's cost amortised, the order aging in between). Device time over all K decisions (regrouping included), and whether every result word and
    final state equals the plain run's."""
    import rsim.cost as cm

    n, K = states_host.shape[0], len(controls)
    d_init = dev.upload_states(states_host)
    d_work = dev.alloc_states(n)
    d_ctl = [dev.upload_controls(c) for c in controls]
    d_res = [dev.alloc_words(n) for _ in range(K)]
    out = {"arenas": n, "decisions": K}
    if plain is None:
        dev.copy_states(d_work, d_init)
        dev.sync()
        dev.timer_begin()
        for k in range(K):
            dev.sim8(d_work, dw, d_ctl[k], d_res[k], block)
        ms = dev.timer_end()
        results = [dev.read_words(r) for r in d_res]
        final = dev.download_states(d_work)
        stepped = stepped_ticks(results)
        out["plain"] = {"arena_ticks_per_s": stepped / (ms / 1000.0), "ms_per_decision": ms / K}
    else:
        results, final, ms = plain
        stepped = stepped_ticks(results)
        out["plain"] = {"arena_ticks_per_s": stepped / (ms / 1000.0), "ms_per_decision": ms / K, "note": "the timed run"}
    for name, mode, every in (("sorted", "sorted", 1), ("balanced", "balanced", 1), ("balanced_every8", "balanced", 8),
                              ("balanced_every32", "balanced", 32)):
        dev.copy_states(d_work, d_init)
        g = cm.Regrouper(dev, model, dw, d_work, mode, block)
        dev.sync()
        dev.timer_begin()
        for k in range(K):
            if k % every == 0:
                g.regroup()
            g.step(d_ctl[k], d_res[k])
        ms = dev.timer_end()
        same = all(np.array_equal(dev.read_words(d_res[k]), results[k]) for k in range(K))
        fin = dev.download_states(g.states_in_arena_order())
        same = same and bool(np.array_equal(fin[:, :2567], final[:, :2567]))
        dev.sync()
        dev.timer_begin()
        for k in range(0, K, every):
            g.regroup()
        reg_ms = dev.timer_end()
        out[name] = {"arena_ticks_per_s": stepped / (ms / 1000.0), "ms_per_decision": ms / K,
                     "regroup_ms_per_decision": reg_ms / K, "identical_to_plain": same,
                     "vs_plain": out["plain"]["ms_per_decision"] / (ms / K)}
        del g
    return out


def open_play_leg(rsim, dev, decisions, block, threads):
    """16,384 open-play arenas over the seven hitbox-preset pairings (preset i vs preset i, arena a in pairing a mod 7),
    `decisions` plain per-launch decisions with per-arena worlds; every result word and final state compared with the
    CPU build of the same decisions."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    n = 16384
    worlds = np.stack([rsim.build_world(p, p) for p in range(7)])
    variant = (np.arange(n) % 7).astype(np.uint32)
    states = rsim.sample_arenas_in(worlds, variant, rsim.SCEN_OPEN, 2031)
    controls = [rsim.sample_controls(4646, d, n) for d in range(decisions)]
    d_w = dev.upload_worlds(worlds)
    d_v = dev.upload_words(variant)
    d_s = dev.upload_states(states)
    d_c = [dev.upload_controls(c) for c in controls]
    d_r = [dev.alloc_words(n) for _ in range(decisions)]
    dev.sync()
    dev.timer_begin()
    for k in range(decisions):
        dev.sim8_variants(d_s, d_w, d_v, d_c[k], d_r[k], block)
    ms = dev.timer_end()
    res = [dev.read_words(r) for r in d_r]
    fin = dev.download_states(d_s)
    o = orc.Oracle()
    cpu = states.copy()
    diffs = 0
    t0 = time.perf_counter()
    for k in range(decisions):
        diffs += int((o.sim8_in(cpu, worlds, variant, controls[k], threads=threads) != res[k]).sum())
    cpu_s = time.perf_counter() - t0
    diffs += int((cpu[:, :2567] != fin[:, :2567]).sum())
    stepped = stepped_ticks(res)
    return {"arenas": n, "decisions": decisions, "worlds": 7, "arena_ticks_per_s": stepped / (ms / 1000.0),
            "ms_per_decision": ms / decisions, "kernel": "k_sim8 (per-arena worlds)", "word_diffs_vs_cpu": diffs,
            "compared": "every arena, every decision", "cpu_arena_ticks_per_s": stepped / cpu_s, "cpu_threads": threads}


def scale_leg(rsim, dev, decisions, block, threads, chunk=50):
    """131,072 arenas of the 1:1:1 kickoff / open play / contact mix, each in its own world of a table of 7 hitbox presets
    x 16 mutator draws (gravity z U(-800, -500), boost acceleration x U(0.8, 1.2), ball radius U(91.25, 100)); `decisions`
    decisions as persistent launches of `chunk` decisions each (k_run), all captured in one HIP graph together with a
    gather of a 1% sample of the arenas after every chunk; the sample's result words and states are compared with the CPU
    build after the run."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    n = 131072
    rng = np.random.default_rng(2032)
    worlds = np.stack([rsim.build_world(p, p, float(rng.uniform(-800, -500)), float(rng.uniform(0.8, 1.2)),
                                        float(rng.uniform(91.25, 100))) for p in range(7) for _ in range(16)])
    variant = rng.integers(0, worlds.shape[0], n).astype(np.uint32)
    states = rsim.sample_arenas_in(worlds, variant, rsim.SCEN_MIXED, 2033)
    sample = np.sort(rng.choice(n, n // 100, replace=False)).astype(np.uint32)
    d_w = dev.upload_worlds(worlds)
    d_v = dev.upload_words(variant)
    d_s = dev.upload_states(states)
    d_sample = dev.upload_words(sample)
    chunks = decisions // chunk
    d_c, d_r, d_g, ctl_sample = [], [], [], []
    for c in range(chunks):
        ctl = np.concatenate([rsim.sample_controls(4747, c * chunk + k, n) for k in range(chunk)])
        d_c.append(dev.upload_words(ctl))
        ctl_sample.append(ctl.reshape(chunk, n, 16)[:, sample].copy())
        d_r.append(dev.alloc_words(chunk * n))
        d_g.append(dev.alloc_states(sample.size))
    dev.sync()
    dev.capture_begin()
    for c in range(chunks):
        dev.run(d_s, d_w, d_c[c], d_r[c], chunk, variant=d_v, block=block)
        dev.gather_states(d_s, d_sample, d_g[c])
    graph = dev.capture_end()
    dev.timer_begin()
    graph.launch()
    ms = dev.timer_end()
    graph.free()
    stepped = 0
    o = orc.Oracle()
    cpu = np.ascontiguousarray(states[sample])
    vs = variant[sample]
    diffs = 0
    for c in range(chunks):
        r = dev.read_words(d_r[c]).reshape(chunk, n)
        stepped += stepped_ticks(list(r))
        for k in range(chunk):
            diffs += int((o.sim8_in(cpu, worlds, vs, np.ascontiguousarray(ctl_sample[c][k]), threads=threads) != r[k, sample]).sum())
        diffs += int((cpu[:, :2567] != dev.download_states(d_g[c])[:, :2567]).sum())
    return {"arenas": n, "decisions": chunks * chunk, "worlds": int(worlds.shape[0]), "arena_ticks_per_s": stepped / (ms / 1000.0),
            "ms_per_decision": ms / (chunks * chunk), "kernel": "k_run, %d decisions per launch, one HIP graph" % chunk,
            "word_diffs_vs_cpu": diffs, "compared": "a 1%% sample (%d arenas): result words of every decision, states every %d"
            % (sample.size, chunk), "state_gigabytes": n * 11136 / 1e9}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)  # the contact-heavy configuration's 200 decisions
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--block", type=int, default=64)
    ap.add_argument("--arenas", type=int, default=ARENAS)
    ap.add_argument("--cpu-sample", type=int, default=ARENAS)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cost", action="store_true", help="skip the learned cost-model legs")
    ap.add_argument("--no-configs", action="store_true", help="skip the 16,384-arena open-play and 131,072-arena scale legs")
    ap.add_argument("--scale-decisions", type=int, default=1000)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this stepper runs on one GPU; --gpus %d is not supported (arenas are independent, so N GPUs would "
              "each step their own range of AoSoA blocks)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import rsim

    n, K, W = args.arenas, args.steps, args.warmup
    print("[bench] building the world and %d contact-heavy arenas" % n, file=sys.stderr, flush=True)
    world = rsim.build_world(0, 0)
    states0 = rsim.sample_arenas(world, rsim.SCEN_CONTACT, 2025, n)
    controls = [rsim.sample_controls(4242, d, n) for d in range(W + K)]

    dev = rsim.SimDevice(0)
    dw = dev.upload_static(world)
    d_init = dev.upload_states(states0)
    d_work = dev.alloc_states(n)
    d_ctl = [dev.upload_controls(c) for c in controls]
    d_res = [dev.alloc_words(n) for _ in range(K)]
    d_warm = dev.alloc_words(n)

    print("[bench] warmup %d decisions" % W, file=sys.stderr, flush=True)
    dev.copy_states(d_work, d_init)
    for k in range(W):
        dev.sim8(d_work, dw, d_ctl[K + k], d_warm, args.block)
    dev.sync()

    print("[bench] timing %d decisions" % K, file=sys.stderr, flush=True)
    dev.copy_states(d_work, d_init)
    dev.sync()
    t0 = time.perf_counter()
    dev.timer_begin()
    for k in range(K):
        dev.sim8(d_work, dw, d_ctl[k], d_res[k], args.block)
    gpu_ms = dev.timer_end()  # synchronises
    wall = time.perf_counter() - t0

    results = [dev.read_words(r) for r in d_res]
    final = dev.download_states(d_work)

    # the same K decisions in one launch (rsim_run: each arena held on chip across decisions), checked against the plain run
    print("[bench] the same decisions in one launch", file=sys.stderr, flush=True)
    d_ctl_all = dev.upload_words(np.concatenate(controls[:K]))
    d_res_all = dev.alloc_words(K * n)
    dev.copy_states(d_work, d_init)
    dev.sync()
    t1 = time.perf_counter()
    dev.timer_begin()
    dev.run(d_work, dw, d_ctl_all, d_res_all, K, block=args.block)
    run_ms = dev.timer_end()
    run_wall = time.perf_counter() - t1
    run_same = bool(np.array_equal(dev.download_states(d_work)[:, :2567], final[:, :2567]) and
                    np.array_equal(dev.read_words(d_res_all).reshape(K, n), np.stack(results)))
    stepped = sum(int(np.where((r & 0xFF) > 0, r & 0xFF, 8).sum()) for r in results)
    nominal = n * 8 * K
    flags = sum(int(((r & 0xC0000000) != 0).sum()) for r in results)
    ms_step = wall * 1000.0 / K
    value = stepped / wall
    per_launch_s = gpu_ms / 1000.0 / K
    achieved = B_TICK * (stepped / K) / per_launch_s
    b_dec = (2 * 2567 * 4 + 68) / 8.0
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("bytes_per_launch")

    # the same K decisions replayed from a HIP graph, checked against the plain run
    print("[bench] the same decisions from a HIP graph", file=sys.stderr, flush=True)
    dev.copy_states(d_work, d_init)
    dev.sync()
    dev.capture_begin()
    for k in range(K):
        dev.sim8(d_work, dw, d_ctl[k], d_res[k], args.block)
    graph = dev.capture_end()
    dev.copy_states(d_work, d_init)
    dev.sync()
    dev.timer_begin()
    graph.launch()
    graph_ms = dev.timer_end()
    graph_same = bool(np.array_equal(dev.download_states(d_work)[:, :2567], final[:, :2567]) and
                      all(np.array_equal(dev.read_words(d_res[k]), results[k]) for k in range(K)))
    graph.free()

    cost = None
    if not args.no_cost:
        print("[bench] cost model: labels and training", file=sys.stderr, flush=True)
        model, training = train_cost_model(rsim, dev, world, dw)
        cost = {"training": training}
        print("[bench] cost-regrouped decisions, this batch", file=sys.stderr, flush=True)
        cost["contact65536"] = compare_orders(rsim, dev, dw, model, states0, controls[:K], args.block,
                                              plain=(results, final, gpu_ms))
        k2 = min(K, 20)
        print("[bench] cost-regrouped decisions, 65,536 arenas 1:1 open play and contact", file=sys.stderr, flush=True)
        cost["open_contact65536"] = compare_orders(
            rsim, dev, dw, model, rsim.sample_arenas(world, rsim.SCEN_OPEN_CONTACT, 2026, n),
            [rsim.sample_controls(4343, d, n) for d in range(k2)], args.block)
        print("[bench] cost-regrouped decisions, 131,072 arenas 1:1:1 kickoff, open play, contact", file=sys.stderr, flush=True)
        cost["mixed131072"] = compare_orders(
            rsim, dev, dw, model, rsim.sample_arenas(world, rsim.SCEN_MIXED, 2027, 2 * n),
            [rsim.sample_controls(4444, d, 2 * n) for d in range(k2)], args.block)

    configs = None
    if not args.no_configs:
        print("[bench] 16,384 open-play arenas, 7 preset pairings, 100 decisions", file=sys.stderr, flush=True)
        configs = {"open16384": open_play_leg(rsim, dev, 100, args.block, host_threads())}
        print("[bench] 131,072 arenas, %d worlds, %d decisions in one graph" % (112, args.scale_decisions), file=sys.stderr, flush=True)
        configs["scale131072"] = scale_leg(rsim, dev, args.scale_decisions, args.block, host_threads())

    cpu, diffs = None, None
    if not args.no_cpu:
        print("[bench] CPU baseline on %d arenas" % min(args.cpu_sample, n), file=sys.stderr, flush=True)
        cpu, diffs = cpu_baseline_leg(world, states0, controls[:K], final, results, min(args.cpu_sample, n), host_threads())

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "arena-ticks/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded contact-heavy start states and per-decision controls on the procedural Soccar arena",
        "config": {"workload": "contact-heavy 65,536 arenas (40% wall/fillet pinch, 30% head-on, 30% goal frame), "
                               "Octane vs Octane, 8-tick decisions, plain per-launch sim8",
                   "arenas": n, "decisions": K, "block": args.block, "order": "plain (unsorted)"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "bytes_per_arena_tick": B_TICK, "kernel": "k_sim8", "kernel_ms_avg": gpu_ms / K,
                     "decision_boundary_bytes_per_arena_tick": b_dec,
                     "frac_at_decision_boundary_bytes": b_dec * (stepped / K) / per_launch_s / HBM_PEAK},
        "cpu_baseline": cpu,
        "arena_ticks_stepped": stepped,
        "arena_ticks_nominal": nominal,
        "nominal_arena_ticks_per_s": nominal / wall,
        "flagged_arenas": flags,
        "cost_regrouped": cost,
        "configs": configs,
        "multi_decision_launch": {"arena_ticks_per_s": stepped / run_wall, "kernel_ms": run_ms, "kernel": "k_run",
                                  "identical_to_plain": run_same},
        "graph_replay": {"arena_ticks_per_s": stepped / (graph_ms / 1000.0), "ms": graph_ms, "identical_to_plain": graph_same},
        "word_diffs": diffs,
    }
    print(json.dumps(line), flush=True)
    dev.close()


if __name__ == "__main__":
    main()
