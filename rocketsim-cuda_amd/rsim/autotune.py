"""Launch-schedule autotuner for the decision stepper.

For one workload (a batch of arenas, its world table and a run of decisions with known controls) it measures every
candidate schedule on the device, each captured once in a HIP graph and replayed, and records the fastest. The knobs:
  - work-group size: 64, 128 or 256 lanes (1, 2 or 4 waves; the pooled stages need whole waves);
  - ticks per launch, i.e. where the state lives between ticks: 1, 2 or 4 ticks per launch (rsim_sim8_split: the state
    back in the HBM buffer between launches), one decision per launch (8 ticks: rsim_sim8_variants), or several
    decisions per launch (rsim_run: the arena held in the lane's private memory across decisions);
  - scheduling of the multi-decision launch: one work-group per AoSoA block (rsim_run), or persistent work-groups taking
    blocks from a queue (rsim_run_queued), in index order or heaviest block first by the learned cost model.
The AoSoA width is not a knob: it is the state buffer's format (64, one wave), which every entry point and caller shares.

Every candidate's result words of every decision and its final states must equal the plain per-launch run's (block 64,
one decision per launch), which the tests check against the oracle; a candidate that differs is recorded and never chosen.

    python -m rsim.autotune [--workload contact65536|open16384|open_contact65536|mixed131072] [--decisions D] [--out FILE]

The record (JSON, one per workload) names the workload, every candidate's time and the choice; `load_choice` reads it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import (SCEN_CONTACT, SCEN_MIXED, SCEN_OPEN, SCEN_OPEN_CONTACT, SIM8_BLOCK, SimDevice, build_world, lib,
               sample_arenas, sample_arenas_in, sample_controls, _check)

TICK_WORDS = 2567


def workload(name, decisions):
    """(description, worlds [k, STATIC_WORDS], variant or None, states, controls list) of a named workload: the batches
    the benchmark times (contact65536: 65,536 contact-heavy arenas; open16384: 16,384 open-play arenas over the seven
    preset pairings; open_contact65536: 1:1 open play and contact; mixed131072: 131,072 arenas of the 1:1:1 mix in 112
    worlds of presets and mutators)."""
    if name == "contact65536":
        w = build_world(0, 0)
        return ("65,536 contact-heavy arenas, Octane vs Octane", w[None], None, sample_arenas(w, SCEN_CONTACT, 2025, 65536),
                [sample_controls(4242, d, 65536) for d in range(decisions)])
    if name == "open_contact65536":
        w = build_world(0, 0)
        return ("65,536 arenas, 1:1 open play and contact", w[None], None, sample_arenas(w, SCEN_OPEN_CONTACT, 2026, 65536),
                [sample_controls(4343, d, 65536) for d in range(decisions)])
    if name == "open16384":
        worlds = np.stack([build_world(p, p) for p in range(7)])
        variant = (np.arange(16384) % 7).astype(np.uint32)
        return ("16,384 open-play arenas over the seven preset pairings", worlds, variant,
                sample_arenas_in(worlds, variant, SCEN_OPEN, 2031), [sample_controls(4646, d, 16384) for d in range(decisions)])
    if name == "mixed131072":
        rng = np.random.default_rng(2032)
        worlds = np.stack([build_world(p, p, float(rng.uniform(-800, -500)), float(rng.uniform(0.8, 1.2)),
                                       float(rng.uniform(91.25, 100))) for p in range(7) for _ in range(16)])
        variant = rng.integers(0, worlds.shape[0], 131072).astype(np.uint32)
        return ("131,072 arenas of the 1:1:1 kickoff / open / contact mix in 112 worlds", worlds, variant,
                sample_arenas_in(worlds, variant, SCEN_MIXED, 2033), [sample_controls(4747, d, 131072) for d in range(decisions)])
    raise ValueError(name)


def candidates(decisions, with_model):
    """Every schedule to measure, as dicts: kind sim8 (ticks_per_launch 1, 2, 4, 8), run (decisions_per_launch) or
    queue (decisions_per_launch, order)."""
    out = [{"kind": "sim8", "block": b, "ticks_per_launch": 8} for b in (64, 128, 256)]
    out += [{"kind": "sim8", "block": 64, "ticks_per_launch": t} for t in (1, 2, 4)]
    dpl = sorted({d for d in (5, 25, decisions) if d <= decisions})
    out += [{"kind": "run", "block": b, "decisions_per_launch": d} for b in (64, 256) for d in dpl]
    orders = ["index"] + (["cost"] if with_model else [])
    out += [{"kind": "queue", "block": 64, "decisions_per_launch": decisions, "order": o} for o in orders]
    return out


def label(c):
    if c["kind"] == "sim8":
        return "block %d, %d tick(s) per launch" % (c["block"], c["ticks_per_launch"])
    if c["kind"] == "run":
        return "block %d, %d decisions per launch" % (c["block"], c["decisions_per_launch"])
    return "block %d, %d decisions per launch, block queue in %s order" % (c["block"], c["decisions_per_launch"], c["order"])


class Bench:
    """Device buffers of one workload and the plain reference run."""

    def __init__(self, dev, worlds, variant, states, controls):
        self.dev, self.n, self.D = dev, states.shape[0], len(controls)
        self.dw = dev.upload_worlds(worlds)
        self.dv = dev.upload_words(variant) if variant is not None else None
        self.init = dev.upload_states(states)
        self.work = dev.alloc_states(self.n)
        self.ctl = [dev.upload_controls(c) for c in controls]
        self.ctl_all = dev.upload_words(np.concatenate(controls))
        self.res = [dev.alloc_words(self.n) for _ in range(self.D)]
        self.res_all = dev.alloc_words(self.D * self.n)
        self.order = None

    def record(self, c):
        """Captures candidate c's launches for all decisions into a graph (results into res / res_all)."""
        dev = self.dev
        chunks = []
        if c["kind"] != "sim8":  # the per-launch slices of the controls and results, made before the capture
            per = c["decisions_per_launch"]
            for k0 in range(0, self.D, per):
                k1 = min(self.D, k0 + per)
                chunks.append((k1 - k0, self.ctl_all.view(k0 * self.n * 16, (k1 - k0) * self.n * 16),
                               self.res_all.view(k0 * self.n, (k1 - k0) * self.n)))
        self._views = chunks
        dev.capture_begin()
        if c["kind"] == "sim8":
            for k in range(self.D):
                if c["ticks_per_launch"] == 8:
                    dev.sim8_variants(self.work, self.dw, self.dv, self.ctl[k], self.res[k], c["block"])
                else:
                    dev.sim8_split(self.work, self.dw, self.ctl[k], self.res[k], c["ticks_per_launch"], variant=self.dv,
                                   block=c["block"])
        else:
            for d, ctl, res in chunks:
                if c["kind"] == "run":
                    dev.run(self.work, self.dw, ctl, res, d, variant=self.dv, block=c["block"])
                else:
                    dev.run_queued(self.work, self.dw, ctl, res, d, variant=self.dv, block=c["block"],
                                   order=self.order if c["order"] == "cost" else None)
        return dev.capture_end()

    def time(self, c, reps=1):
        dev = self.dev
        g = self.record(c)
        best = None
        for _ in range(reps):
            dev.copy_states(self.work, self.init)
            dev.sync()
            dev.timer_begin()
            g.launch()
            ms = dev.timer_end()
            best = ms if best is None else min(best, ms)
        g.free()
        if c["kind"] == "sim8":
            res = np.stack([dev.read_words(r) for r in self.res])
        else:
            res = dev.read_words(self.res_all).reshape(self.D, self.n)
        return best, res, dev.download_states(self.work)[:, :TICK_WORDS]


def cost_block_order(dev, model, states_buf, n):
    """The heaviest-first AoSoA block order of the start states by the cost model (rsim_cost_block_order)."""
    pred = model.predict(model.features(states_buf), n)
    work = dev.alloc_words(int(lib().rsim_cost_block_order_work_words(n)))
    order = dev.alloc_words((n + 63) // 64)
    _check(lib().rsim_cost_block_order(dev.handle, pred.handle, n, work.handle, order.handle), "rsim_cost_block_order")
    dev.sync()
    return order


def tune(name="contact65536", decisions=25, dev=None, model=None, log=sys.stderr, reps=1, arenas=None):
    """Measures every candidate on the named workload (the first `decisions` decisions; the first `arenas` arenas when
    given, for small trial runs) and returns the record."""
    desc, worlds, variant, states, controls = workload(name, decisions)
    if arenas is not None and arenas < states.shape[0]:
        desc = "the first %d arenas of: %s" % (arenas, desc)
        states, controls = states[:arenas], [c[:arenas] for c in controls]
        variant = variant[:arenas] if variant is not None else None
    own = dev is None
    dev = dev or SimDevice(0)
    b = Bench(dev, worlds, variant, states, controls)
    if model is not None:
        b.order = cost_block_order(dev, model, b.init, b.n)
    plain = {"kind": "sim8", "block": SIM8_BLOCK, "ticks_per_launch": 8}
    _, ref_res, ref_states = b.time(plain)
    stepped = int(np.where((ref_res & 0xFF) > 0, ref_res & 0xFF, 8).sum())
    cands = []
    for c in candidates(decisions, model is not None):
        t0 = time.perf_counter()
        ms, res, st = b.time(c, reps)
        same = bool(np.array_equal(res, ref_res) and np.array_equal(st, ref_states))
        cands.append(dict(c, label=label(c), ms=ms, ms_per_decision=ms / decisions,
                          arena_ticks_per_s=stepped / (ms / 1000.0), identical=same, wall_s=time.perf_counter() - t0))
        print("[autotune] %-62s %8.2f ms  identical %s" % (label(c), ms, same), file=log, flush=True)
    if own:
        dev.close()
    ok = [c for c in cands if c["identical"]]
    best = min(ok, key=lambda c: c["ms"]) if ok else None
    return {"workload": {"name": name, "what": desc, "arenas": int(states.shape[0]), "decisions": decisions,
                         "worlds": int(worlds.shape[0])},
            "method": "each candidate captured once in a HIP graph, replayed from the same start states, device time; "
                      "every decision's result words and the final states compared with the plain per-launch run",
            "candidates": cands, "all_identical": len(ok) == len(cands),
            "choice": {k: best[k] for k in ("kind", "block", "ticks_per_launch", "decisions_per_launch", "order") if k in best}
            if best else None}


def load_choice(path):
    """The chosen schedule of a record written by tune(): a dict with kind, block and its knobs."""
    with open(path) as f:
        return json.load(f)["choice"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="contact65536")
    ap.add_argument("--decisions", type=int, default=25)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    rec = tune(a.workload, a.decisions)
    text = json.dumps(rec, indent=1)
    if a.out:
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        with open(a.out, "w") as f:
            f.write(text + "\n")
    print(text)


if __name__ == "__main__":
    main()
