"""Launch-schedule autotuner for the decision stepper.

For a given batch it measures, on the device, every candidate schedule and records the fastest:
  - work-group size (1 to 4 waves: 64, 128, 192, 256 lanes; also 32),
  - decisions per launch (1 = one sim8 launch per decision, or D in one rsim_run launch with each arena held on chip
    across decisions),
and checks that every candidate leaves identical states and result words (results may not depend on the schedule).

    python -m rsim.autotune [--arenas N] [--decisions D] [--scenario S] [--out FILE]

The record (JSON) names the workload, every candidate's time and the choice; `load_choice` reads it back.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import SCEN_CONTACT, SimDevice, build_world, sample_arenas, sample_controls

BLOCKS = (32, 64, 128, 192, 256)
PER_LAUNCH = (1, 0)  # 1: one decision per launch; 0: all decisions in one launch


def tune(n=65536, decisions=8, scenario=SCEN_CONTACT, seed=2025, blocks=BLOCKS, per_launch=PER_LAUNCH, log=sys.stderr):
    world = build_world(0, 0)
    states = sample_arenas(world, scenario, seed, n)
    controls = [sample_controls(seed + 1, d, n) for d in range(decisions)]
    dev = SimDevice(0)
    dw = dev.upload_static(world)
    d_init = dev.upload_states(states)
    d_work = dev.alloc_states(n)
    d_ctl = [dev.upload_controls(c) for c in controls]
    d_ctl_all = dev.upload_words(np.concatenate(controls))
    d_res = dev.alloc_words(decisions * n)
    d_res1 = dev.alloc_words(n)
    cands, ref = [], None
    for block in blocks:
        for pl in per_launch:
            dev.copy_states(d_work, d_init)
            dev.sync()
            t0 = time.perf_counter()
            dev.timer_begin()
            if pl == 1:
                for d in range(decisions):
                    dev.sim8(d_work, dw, d_ctl[d], d_res1, block)
            else:
                dev.run(d_work, dw, d_ctl_all, d_res, decisions, block=block)
            ms = dev.timer_end()
            wall = time.perf_counter() - t0
            out = dev.download_states(d_work)[:, :2567]
            last = dev.read_words(d_res1) if pl == 1 else dev.read_words(d_res).reshape(decisions, n)[-1]
            if ref is None:
                ref = (out, last)
            same = bool(np.array_equal(out, ref[0]) and np.array_equal(last, ref[1]))
            cands.append({"block": block, "decisions_per_launch": 1 if pl == 1 else decisions, "kernel_ms": ms,
                          "wall_s": wall, "identical": same})
            print("[autotune] block %3d, %2d decision(s) per launch: %.2f ms, identical %s"
                  % (block, cands[-1]["decisions_per_launch"], ms, same), file=log, flush=True)
    dev.close()
    ok = [c for c in cands if c["identical"]]
    best = min(ok, key=lambda c: c["kernel_ms"])
    return {"workload": {"arenas": n, "decisions": decisions, "scenario": scenario, "seed": seed},
            "candidates": cands, "choice": {"block": best["block"], "decisions_per_launch": best["decisions_per_launch"]},
            "all_identical": len(ok) == len(cands)}


def load_choice(path):
    """The chosen (block, decisions_per_launch) of a record written by tune()."""
    with open(path) as f:
        c = json.load(f)["choice"]
    return c["block"], c["decisions_per_launch"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--arenas", type=int, default=65536)
    ap.add_argument("--decisions", type=int, default=8)
    ap.add_argument("--scenario", type=int, default=SCEN_CONTACT)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    rec = tune(a.arenas, a.decisions, a.scenario)
    text = json.dumps(rec, indent=1)
    if a.out:
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        with open(a.out, "w") as f:
            f.write(text + "\n")
    print(text)


if __name__ == "__main__":
    main()
