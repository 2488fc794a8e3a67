"""The autotuner's host logic (no GPU): the candidate space covers every knob, labels name them, and the committed
per-workload records are well formed — every candidate identical to the plain run, the choice the fastest of them."""
import glob
import json
import os

import pytest

from rsim import autotune

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_candidates_cover_every_knob():
    c = autotune.candidates(25, with_model=True)
    kinds = {x["kind"] for x in c}
    assert kinds == {"sim8", "run", "queue"}
    assert {x["ticks_per_launch"] for x in c if x["kind"] == "sim8"} == {1, 2, 4, 8}
    assert {x["block"] for x in c if x["kind"] == "sim8" and x["ticks_per_launch"] == 8} == {64, 128, 256}
    assert {x["decisions_per_launch"] for x in c if x["kind"] == "run"} == {5, 25}
    assert {x["order"] for x in c if x["kind"] == "queue"} == {"index", "cost"}
    assert not [x for x in autotune.candidates(25, with_model=False) if x.get("order") == "cost"]
    assert {x["decisions_per_launch"] for x in autotune.candidates(3, False) if x["kind"] == "run"} == {3}
    assert len({autotune.label(x) for x in c}) == len(c)  # every candidate has a label of its own


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(ROOT, "profiles", "autotune_*.json"))))
def test_recorded_choices(path):
    with open(path) as f:
        rec = json.load(f)
    cands = rec["candidates"]
    assert rec["all_identical"] and all(x["identical"] for x in cands)
    best = min(cands, key=lambda x: x["ms"])
    assert rec["choice"] == {k: best[k] for k in rec["choice"]}
    assert autotune.load_choice(path) == rec["choice"]
    assert rec["workload"]["name"] in os.path.basename(path)
