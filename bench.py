"""Benchmark of the 8-tick decision path on one MI355X: arena-ticks/s over 65,536 contact-heavy arenas.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--block B] [--no-cpu] [--no-cost] [--no-configs]

A step is one decision (8 ticks at 120 Hz) of all 65,536 arenas, with fresh per-car controls each decision, states
resident in HBM (AoSoA-64). Ticks after a goal are not stepped, so `value` counts the ticks actually stepped (the scored
tick ends an arena's decision); the nominal 8-tick count is reported beside it.

Workload: the contact-heavy mix (40% ball pinched against a wall or fillet with a car driving into it, 30% head-on cars,
30% ball at the goal frame), Octane vs Octane on the procedural Soccar arena, seeded synthetic states and controls. The
configured run is K decisions with the persistent kernel, HIP-graph capture and the autotuned schedule: the autotuner
(rsim.autotune) measures every candidate schedule on the first decisions of this batch, each captured in a graph, and
the K timed decisions replay the chosen one as one graph, bracketed by device synchronisations (wall clock; HIP events
on the device stream give the kernel time). The plain per-launch kernel (one sim8 launch per decision, the reference's
API) is timed beside it, and every result word and final state of every schedule is compared with it.

Prints one JSON line. `roofline` prices the dominant kernel of the configured run against HBM at 20,544.5 algorithmic
bytes per arena-tick (the 2,567 persistent state words read and written once per tick, plus controls and result
amortised over the decision). `cpu_baseline` is the reference's own CPU build of the path (its phase.cpp compiled in
place, when built with the tree; otherwise the test oracle, bit-identical to it) on the host's cores over the arenas of
the same run; its outputs are word-compared with the GPU's (`word_diffs`).
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
TICK_WORDS = 2567
B_TICK = 2 * TICK_WORDS * 4 + (64 + 4) / 8  # algorithmic HBM bytes per arena-tick
B_DEC = (2 * TICK_WORDS * 4 + 68) / 8.0      # the same with the state read and written once per decision
HBM_PEAK = 8.0e12
METRIC = "arena-ticks/s @65,536 arenas, plain vs cost-sorted; 0 word diffs vs CPU"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "run_contact65536_traffic.json")
RECORD_FILE = os.path.join(ROOT, "profiles", "autotune_contact65536.json")
HOST_SHARE = 16  # host threads a one-GPU box allots to one process


def host_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, HOST_SHARE))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def stepped_ticks(results):
    return sum(int(np.where((r & 0xFF) > 0, r & 0xFF, 8).sum()) for r in results)


def cpu_baseline_leg(world, states0, controls, gpu_states, gpu_results, sample, threads):
    """The CPU path over arenas [0, sample) of the same run on `threads` host threads, timed and word-compared with the
    GPU's outputs, plus a single-thread figure on a smaller sample. It is the reference's own C++ build
    (crates/rocketsim-cuda/csrc/phase.cpp compiled in place into oracle/_ref, driven over host threads) when that library
    was built with the tree, otherwise the oracle (the same tick built from this project's core, bit-identical to the
    reference build by the tests)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    if os.path.exists(orc.REF_LIB):
        cpu, kind, what = orc.Reference(), "reference", "the reference's own C++ build (phase.cpp, g++ -O2 -ffp-contract=off)"
    else:
        if not os.path.exists(orc.ORACLE_LIB):
            orc.build(ref=False)
        cpu, kind, what = orc.Oracle(), "port", "CPU build of the same tick (g++ -O2 -ffp-contract=off)"
    s = np.ascontiguousarray(states0[:sample])
    t0 = time.perf_counter()
    res = [cpu.sim8(s, world, np.ascontiguousarray(c[:sample]), threads=threads) for c in controls]
    dt = time.perf_counter() - t0
    ticks = stepped_ticks(res)
    diffs = sum(int((r != gpu_results[k][:sample]).sum()) for k, r in enumerate(res))
    diffs += int((s[:, :TICK_WORDS] != gpu_states[:sample, :TICK_WORDS]).sum())
    # one thread, on the first 2,048 arenas over the first 20 decisions
    n1, k1 = min(2048, sample), min(20, len(controls))
    s1 = np.ascontiguousarray(states0[:n1])
    t1 = time.perf_counter()
    r1 = [cpu.sim8(s1, world, np.ascontiguousarray(c[:n1]), threads=1) for c in controls[:k1]]
    dt1 = time.perf_counter() - t1
    try:
        visible = len(os.sched_getaffinity(0))
    except AttributeError:
        visible = os.cpu_count()
    return {"value": ticks / dt, "unit": "arena-ticks/s", "cores": threads, "kind": kind,
            "sample": "arenas 0..%d of the same run over all %d timed decisions (%d arena-ticks), %d host threads (the "
                      "host's share for one GPU), %s" % (sample - 1, len(controls), ticks, threads, what),
            "seconds": dt, "cpu_model": cpu_model(), "visible_cores": visible,
            "single_core": {"value": stepped_ticks(r1) / dt1, "unit": "arena-ticks/s",
                            "sample": "arenas 0..%d, decisions 0..%d, one thread" % (n1 - 1, k1 - 1), "seconds": dt1}}, diffs


def train_cost_model(rsim, dev, world_host, dw, label_arenas=65536, label_decisions=21):
    """The cost model trained on labels the simulator emits: `label_arenas` arenas of each of three scenario mixes (seeds
    apart from every timed batch) stepped `label_decisions` decisions with one arena per wave and each lane's clock
    recorded (3 x 65,536 x 21 = 4.13 M arena-decisions by default)."""
    import rsim.cost as cm

    t0 = time.perf_counter()
    F, T, C = [], [], []
    for scen, seed in ((rsim.SCEN_CONTACT, 77), (rsim.SCEN_OPEN_CONTACT, 79), (rsim.SCEN_MIXED, 81)):
        f, t, c = cm.label_costs(dev, dw, rsim.sample_arenas(world_host, scen, seed, label_arenas),
                                 [rsim.sample_controls(seed + 1, d, label_arenas) for d in range(label_decisions)],
                                 decisions=label_decisions)
        F.append(f), T.append(t), C.append(c)
    feats, targets, contacts = np.concatenate(F), np.concatenate(T), np.concatenate(C)
    label_s = time.perf_counter() - t0
    model = cm.CostModel(dev, seed=1)
    curve = model.train(feats, targets, contacts, epochs=15, batch=4096, lr=3e-3)
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


def timed_graph(dev, record, reset):
    """Records the work `record()` queues into a graph, then replays it after `reset()` (untimed), bracketed by
    synchronisations: (device ms by HIP events on the stream, wall seconds)."""
    dev.sync()
    dev.capture_begin()
    record()
    g = dev.capture_end()
    reset()
    dev.sync()
    t0 = time.perf_counter()
    dev.timer_begin()
    g.launch()
    ms = dev.timer_end()
    wall = time.perf_counter() - t0
    g.free()
    return ms, wall


def regrouping_leg(rsim, dev, dw, model, states_host, controls, schedule, every, plain=None):
    """K decisions of one batch on the configured persistent schedule (`every` decisions per launch), in the plain order
    and ordered by the learned cost model before every launch:
      sorted    arenas regrouped into cost-homogeneous waves (features -> MLP -> device counting sort -> state gather),
                the block queue then taking the heaviest waves first;
      balanced  every wave dealt one arena of each cost rank (the waves' costs evened out);
      lpt       no arena moved: the AoSoA blocks queued heaviest first by their largest predicted arena cost (longest
                processing time first over the persistent work-groups).
    Device time of all K decisions, ordering included, and whether every result word and final state equals the plain
    run's."""
    import rsim.cost as cm

    n, K = states_host.shape[0], len(controls)
    chunks = [(k0, min(K, k0 + every)) for k0 in range(0, K, every)]
    init = dev.upload_states(states_host)
    work = dev.alloc_states(n)
    ctl_all = dev.upload_words(np.concatenate(controls))
    res_all = dev.alloc_words(K * n)
    cv = [(k1 - k0, ctl_all.view(k0 * n * 16, (k1 - k0) * n * 16), res_all.view(k0 * n, (k1 - k0) * n)) for k0, k1 in chunks]
    block = schedule.get("block", 64)
    out = {"arenas": n, "decisions": K, "decisions_per_launch": every, "block": block}

    def launch(d, c, r, order=None):
        dev.run_queued(work, dw, c, r, d, block=block, order=order)

    if plain is None:
        ms, _ = timed_graph(dev, lambda: [launch(d, c, r) for d, c, r in cv], lambda: dev.copy_states(work, init))
        results = dev.read_words(res_all).reshape(K, n).copy()
        final = dev.download_states(work)[:, :TICK_WORDS]
    else:
        results, final, ms = plain
    stepped = stepped_ticks(list(results))
    out["plain"] = {"arena_ticks_per_s": stepped / (ms / 1000.0), "ms_per_decision": ms / K}

    def verdict(name, ms, res, fin):
        out[name] = {"arena_ticks_per_s": stepped / (ms / 1000.0), "ms_per_decision": ms / K,
                     "identical_to_plain": bool(np.array_equal(res, results) and np.array_equal(fin, final)),
                     "vs_plain": out["plain"]["ms_per_decision"] / (ms / K)}

    for mode in ("sorted", "balanced"):
        g = cm.Regrouper(dev, model, dw, work, mode, block, max_decisions=every)
        for d, c, r in cv:
            g.views(c, r, d)
        if len(cv) > 1 and cv[-1][0] != every:
            g._run_views(cv[-1][0])

        def rec():
            for d, c, r in cv:
                g.regroup()
                g.run(c, r, d, queued=True)

        ms, _ = timed_graph(dev, rec, lambda: g.reset(init))
        res = dev.read_words(res_all).reshape(K, n)
        fin = dev.download_states(g.states_in_arena_order())[:, :TICK_WORDS]
        verdict(mode, ms, res, fin)
        dev.sync()
        dev.timer_begin()
        for _ in cv:
            g.regroup()
        out[mode]["ordering_ms_per_decision"] = dev.timer_end() / K
        g.reset(init)
        dev.sync()
        del g
    bo = cm.BlockOrder(dev, model, work)

    def rec_lpt():
        for d, c, r in cv:
            launch(d, c, r, bo.update())

    ms, _ = timed_graph(dev, rec_lpt, lambda: dev.copy_states(work, init))
    verdict("lpt", ms, dev.read_words(res_all).reshape(K, n), dev.download_states(work)[:, :TICK_WORDS])
    dev.sync()
    dev.timer_begin()
    for _ in cv:
        bo.update()
    out["lpt"]["ordering_ms_per_decision"] = dev.timer_end() / K
    return out


def cpu_checker(reference=True):
    """The CPU build a configuration leg compares with: the reference's own C++ build when it was built with the tree
    (oracle/_ref) and `reference` is set, else the oracle. Returns (name, step) with step(states, worlds, variant,
    controls, threads) -> result words of one decision, the states stepped in place; the reference takes one StaticData
    per call, so its arenas are stepped world by world (one call per world and decision: fine for a few worlds, slow for
    the 112-world table, whose 1% sample is checked with the oracle)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    if reference and os.path.exists(orc.REF_LIB):
        ref = orc.Reference()

        def step(cpu, worlds, variant, controls, threads):
            res = np.zeros(cpu.shape[0], np.uint32)
            for p in np.unique(variant):
                idx = np.nonzero(variant == p)[0]
                sub = np.ascontiguousarray(cpu[idx])
                res[idx] = ref.sim8(sub, worlds[p], np.ascontiguousarray(controls[idx]), threads=threads)
                cpu[idx] = sub
            return res
        return "the reference's own C++ build", step
    if not os.path.exists(orc.ORACLE_LIB):
        orc.build(ref=False)
    o = orc.Oracle()
    return "the oracle", lambda cpu, worlds, variant, controls, threads: o.sim8_in(cpu, worlds, variant, controls, threads=threads)


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
    what, step = cpu_checker()
    cpu = states.copy()
    diffs = 0
    t0 = time.perf_counter()
    for k in range(decisions):
        diffs += int((step(cpu, worlds, variant, controls[k], threads) != res[k]).sum())
    cpu_s = time.perf_counter() - t0
    diffs += int((cpu[:, :TICK_WORDS] != fin[:, :TICK_WORDS]).sum())
    stepped = stepped_ticks(res)
    return {"arenas": n, "decisions": decisions, "worlds": 7, "arena_ticks_per_s": stepped / (ms / 1000.0),
            "ms_per_decision": ms / decisions, "kernel": "k_sim8 (per-arena worlds)", "word_diffs_vs_cpu": diffs,
            "compared": "every arena, every decision, with %s" % what, "cpu_arena_ticks_per_s": stepped / cpu_s,
            "cpu_threads": threads}


def scale_leg(rsim, dev, model, decisions, block, threads, chunk=50):
    """131,072 arenas of the 1:1:1 kickoff / open play / contact mix, each in its own world of a table of 7 hitbox presets
    x 16 mutator draws (gravity z U(-800, -500), boost acceleration x U(0.8, 1.2), ball radius U(91.25, 100));
    `decisions` decisions in one HIP graph: before every `chunk` decisions the arenas are regrouped by the learned cost
    model (sorted into cost-homogeneous waves), the chunk runs as one persistent launch whose block queue takes the
    heaviest waves first, and a 1% sample of the arenas is gathered. The sample's result words (every decision) and
    states (every chunk) are compared with the CPU build after the run. Without a model, the plain order."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    import rsim.cost as cm

    n = 131072
    rng = np.random.default_rng(2032)
    worlds = np.stack([rsim.build_world(p, p, float(rng.uniform(-800, -500)), float(rng.uniform(0.8, 1.2)),
                                        float(rng.uniform(91.25, 100))) for p in range(7) for _ in range(16)])
    variant = rng.integers(0, worlds.shape[0], n).astype(np.uint32)
    states = rsim.sample_arenas_in(worlds, variant, rsim.SCEN_MIXED, 2033)
    sample = np.sort(rng.choice(n, n // 100, replace=False)).astype(np.uint32)
    padded = np.zeros(n, np.uint32)
    padded[: sample.size] = sample
    d_w = dev.upload_worlds(worlds)
    d_v = dev.upload_words(variant)
    d_init = dev.upload_states(states)
    d_s = dev.alloc_states(n)
    d_pad = dev.upload_words(padded)
    d_slots = dev.alloc_words(n)
    chunks = decisions // chunk
    d_c, d_r, d_g, ctl_sample = [], [], [], []
    for c in range(chunks):
        ctl = np.concatenate([rsim.sample_controls(4747, c * chunk + k, n) for k in range(chunk)])
        d_c.append(dev.upload_words(ctl))
        ctl_sample.append(ctl.reshape(chunk, n, 16)[:, sample].copy())
        d_r.append(dev.alloc_words(chunk * n))
        d_g.append(dev.alloc_states(sample.size))
    g = None
    if model is not None:
        g = cm.Regrouper(dev, model, d_w, d_s, "sorted", block, variant=d_v, max_decisions=chunk)
        for c in range(chunks):
            g.views(d_c[c], d_r[c], chunk)
        g.slots_of(d_pad, d_slots)

    def rec():
        for c in range(chunks):
            if g is not None:
                g.regroup()
                g.run(d_c[c], d_r[c], chunk, queued=True)
                g.slots_of(d_pad, d_slots)
                dev.gather_states(d_s, d_slots, d_g[c])
            else:
                dev.run_queued(d_s, d_w, d_c[c], d_r[c], chunk, variant=d_v, block=block)
                dev.gather_states(d_s, d_pad, d_g[c])

    reset = (lambda: g.reset(d_init)) if g is not None else (lambda: dev.copy_states(d_s, d_init))
    ms, wall = timed_graph(dev, rec, reset)
    stepped = 0
    what, step = cpu_checker(reference=False)
    cpu = np.ascontiguousarray(states[sample])
    vs = variant[sample]
    diffs = 0
    for c in range(chunks):
        r = dev.read_words(d_r[c]).reshape(chunk, n)
        stepped += stepped_ticks(list(r))
        for k in range(chunk):
            diffs += int((step(cpu, worlds, vs, np.ascontiguousarray(ctl_sample[c][k]), threads) != r[k, sample]).sum())
        diffs += int((cpu[:, :TICK_WORDS] != dev.download_states(d_g[c])[:, :TICK_WORDS]).sum())
    achieved = B_TICK * stepped / (ms / 1000.0)
    return {"arenas": n, "decisions": chunks * chunk, "worlds": int(worlds.shape[0]), "arena_ticks_per_s": stepped / (ms / 1000.0),
            "ms_per_decision": ms / (chunks * chunk), "wall_s": wall,
            "kernel": "k_run_queue, %d decisions per launch, %s, one HIP graph" % (
                chunk, "arenas regrouped by the cost model (sorted) before each launch" if g is not None else "plain order"),
            "hbm_frac_at_b_tick": achieved / HBM_PEAK,
            "word_diffs_vs_cpu": diffs, "compared": "a 1%% sample (%d arenas): result words of every decision, states every %d, "
            "with %s" % (sample.size, chunk, what), "state_gigabytes": n * 11136 / 1e9}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)  # the contact-heavy configuration's 200 decisions
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--block", type=int, default=64, help="work-group size of the plain per-launch run")
    ap.add_argument("--arenas", type=int, default=ARENAS)
    ap.add_argument("--cpu-sample", type=int, default=ARENAS)
    ap.add_argument("--tune-decisions", type=int, default=25, help="decisions the autotuner measures each candidate on")
    ap.add_argument("--no-tune", action="store_true",
                    help="take the schedule recorded in profiles/autotune_contact65536.json instead of tuning now")
    ap.add_argument("--schedule", default=None, metavar="KIND:BLOCK",
                    help="run the configured decisions on this schedule (run:64, queue:64 ...: all decisions in one "
                         "launch) instead of tuning or the record; profiles use it to name the kernel they measure")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cost", action="store_true", help="skip the learned cost-model legs")
    ap.add_argument("--no-configs", action="store_true", help="skip the 16,384-arena open-play and 131,072-arena scale legs")
    ap.add_argument("--scale-decisions", type=int, default=1000)
    ap.add_argument("--label-arenas", type=int, default=65536)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this stepper runs on one GPU; --gpus %d is not supported (arenas are independent, so N GPUs would "
              "each step their own range of AoSoA blocks)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import rsim
    from rsim import autotune

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

    # the plain per-launch kernel: one sim8 launch per decision (the reference's API), the reference run for every compare
    print("[bench] plain per-launch: %d decisions" % K, file=sys.stderr, flush=True)
    dev.copy_states(d_work, d_init)
    dev.sync()
    t0 = time.perf_counter()
    dev.timer_begin()
    for k in range(K):
        dev.sim8(d_work, dw, d_ctl[k], d_res[k], args.block)
    plain_ms = dev.timer_end()
    plain_wall = time.perf_counter() - t0
    results = [dev.read_words(r) for r in d_res]
    final = dev.download_states(d_work)
    stepped = stepped_ticks(results)
    nominal = n * 8 * K
    flags = sum(int(((r & 0xC0000000) != 0).sum()) for r in results)
    del d_res

    # the autotuner on the first decisions of this batch; its choice is the configured schedule
    if args.schedule:
        kind, _, blk = args.schedule.partition(":")
        if kind not in ("run", "queue") or not blk.isdigit():
            sys.exit("bench.py: --schedule takes run:BLOCK or queue:BLOCK")
        choice = {"kind": kind, "block": int(blk), "decisions_per_launch": K}
        if kind == "queue":
            choice["order"] = "index"
        tune_rec, tuned_on = None, K
    elif args.no_tune:
        with open(RECORD_FILE) as f:
            rec = json.load(f)
        tune_rec, choice, tuned_on = None, rec["choice"], rec["workload"]["decisions"]
    else:
        print("[bench] autotuning on %d decisions" % args.tune_decisions, file=sys.stderr, flush=True)
        tuned_on = min(args.tune_decisions, K)
        tune_rec = autotune.tune("contact65536", tuned_on, dev=dev, arenas=n)
        choice = tune_rec.get("choice") or {"kind": "run", "block": 64, "decisions_per_launch": K}
    if choice["kind"] != "sim8" and choice.get("decisions_per_launch") == tuned_on:
        choice = dict(choice, decisions_per_launch=K)  # "all decisions in one launch" scales with the run

    # the configured run: K decisions on the chosen schedule, captured in one HIP graph
    print("[bench] configured run: %s, %d decisions in one graph" % (autotune.label(choice), K), file=sys.stderr, flush=True)
    b = autotune.Bench(dev, world[None], None, states0, controls[:K])
    g = b.record(choice)
    dev.copy_states(b.work, b.init)
    dev.sync()
    t1 = time.perf_counter()
    dev.timer_begin()
    g.launch()
    run_ms = dev.timer_end()  # synchronises
    wall = time.perf_counter() - t1
    g.free()
    if choice["kind"] == "sim8":
        run_res = np.stack([dev.read_words(r) for r in b.res])
    else:
        run_res = dev.read_words(b.res_all).reshape(K, n)
    run_same = bool(np.array_equal(run_res, np.stack(results)) and
                    np.array_equal(dev.download_states(b.work)[:, :TICK_WORDS], final[:, :TICK_WORDS]))
    if choice["kind"] == "sim8":
        launches = K * ((8 + choice["ticks_per_launch"] - 1) // choice["ticks_per_launch"])
        kernel = "k_sim8" if choice["ticks_per_launch"] == 8 else "k_sim8_ticks"
    else:
        launches = (K + choice["decisions_per_launch"] - 1) // choice["decisions_per_launch"]
        kernel = "k_run" if choice["kind"] == "run" else "k_run_queue"
    del b
    value = stepped / wall
    per_launch_s = run_ms / 1000.0 / launches
    achieved = B_TICK * (stepped / launches) / per_launch_s
    traffic = None  # the profiled HBM bytes of one launch, when they were measured on this kernel, block and launch size
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            profs = json.load(f)
        for prof in profs if isinstance(profs, list) else [profs]:
            if (prof.get("kernel") == kernel and prof.get("block") == choice.get("block") and
                    prof.get("arena_ticks_per_launch") == stepped / launches):
                traffic = prof.get("bytes_per_launch")

    cost = None
    model = None
    if not args.no_cost:
        print("[bench] cost model: labels and training", file=sys.stderr, flush=True)
        model, training = train_cost_model(rsim, dev, world, dw, label_arenas=args.label_arenas)
        cost = {"training": training}
        every = 25 if K >= 25 else K
        sched = {"block": choice.get("block", 64)}
        print("[bench] cost-ordered decisions, this batch (%d decisions, %d per launch)" % (K, every), file=sys.stderr, flush=True)
        cost["contact65536"] = regrouping_leg(rsim, dev, dw, model, states0, controls[:K], sched, every)
        print("[bench] cost-ordered decisions, 65,536 arenas 1:1 open play and contact", file=sys.stderr, flush=True)
        cost["open_contact65536"] = regrouping_leg(
            rsim, dev, dw, model, rsim.sample_arenas(world, rsim.SCEN_OPEN_CONTACT, 2026, n),
            [rsim.sample_controls(4343, d, n) for d in range(K)], sched, every)

    configs = None
    if not args.no_configs:
        print("[bench] 16,384 open-play arenas, 7 preset pairings, 100 decisions", file=sys.stderr, flush=True)
        configs = {"open16384": open_play_leg(rsim, dev, 100, args.block, host_threads())}
        print("[bench] 131,072 arenas, 112 worlds, %d decisions in one graph" % args.scale_decisions, file=sys.stderr, flush=True)
        configs["scale131072"] = scale_leg(rsim, dev, model, args.scale_decisions, choice.get("block", 64), host_threads())

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
        "ms_per_step": wall * 1000.0 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded contact-heavy start states and per-decision controls on the procedural Soccar arena",
        "config": {"workload": "contact-heavy 65,536 arenas (40% wall/fillet pinch, 30% head-on, 30% goal frame), "
                               "Octane vs Octane, 8-tick decisions, plain (unsorted) order; persistent kernel, one HIP "
                               "graph, autotuned schedule",
                   "arenas": n, "decisions": K, "schedule": dict(choice, label=autotune.label(choice)),
                   "order": "plain (unsorted)"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "bytes_per_arena_tick": B_TICK, "kernel": kernel, "launches": launches,
                     "kernel_ms_avg": run_ms / launches, "arena_ticks_per_launch": stepped / launches,
                     "decision_boundary_bytes_per_arena_tick": B_DEC,
                     "frac_at_decision_boundary_bytes": B_DEC * (stepped / launches) / per_launch_s / HBM_PEAK},
        "cpu_baseline": cpu,
        "arena_ticks_stepped": stepped,
        "arena_ticks_nominal": nominal,
        "nominal_arena_ticks_per_s": nominal / wall,
        "flagged_arenas": flags,
        "identical_to_plain": run_same,
        "plain_per_launch": {"arena_ticks_per_s": stepped / plain_wall, "ms_per_decision": plain_wall * 1000.0 / K,
                             "kernel": "k_sim8", "kernel_ms_avg": plain_ms / K, "block": args.block,
                             "roofline_frac": B_TICK * (stepped / K) / (plain_ms / 1000.0 / K) / HBM_PEAK},
        "autotune": {"given": args.schedule, "choice": choice} if args.schedule else
                    {"record": os.path.relpath(RECORD_FILE, ROOT), "choice": choice} if tune_rec is None else {
            "workload": tune_rec["workload"], "all_identical": tune_rec["all_identical"], "choice": tune_rec["choice"],
            "candidates": [{"label": c["label"], "ms_per_decision": c["ms_per_decision"], "identical": c["identical"]}
                           for c in tune_rec["candidates"]]},
        "cost_regrouped": cost,
        "configs": configs,
        "word_diffs": diffs,
    }
    print(json.dumps(line), flush=True)
    dev.close()


if __name__ == "__main__":
    main()
