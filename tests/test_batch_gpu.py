"""Batches of gate-count searches that share a gate set (chz_batch_*): every search's level lines and result are
the ones it gets alone and the reference's golden logs, whatever else shares its kernel launches and whenever it
joins.  Seeded workloads as the bench's (splitmix64)."""
import json
import os

import pytest

from conftest import GOLDEN, golden_cases, manifest, read_log, read_spec, strip_seconds

import cheese_amd
from cheese_amd import workloads as W

pytestmark = pytest.mark.gpu


def _levels_text(levels):
    return "\n".join(cheese_amd.level_line(lv) for lv in levels)


def _same(alone, batched):
    assert alone["rc"] == batched["rc"]
    assert strip_seconds(_levels_text(alone["levels"])) == strip_seconds(_levels_text(batched["levels"]))
    if alone["rc"] == 0:
        assert alone["circuit_idx"] == batched["circuit_idx"]
        assert alone["found"] == batched["found"]
        assert batched["verified"]


def test_golden_logs_in_one_batch():
    """every 3-qubit golden fixture of the default gate set that the reference finished within 10 s, searched
    together: each log equals the reference's"""
    names = [n for n in golden_cases(max_seconds=10)
             if manifest()[n]["n"] == 3 and not manifest()[n]["flags"]]
    specs = [read_spec(n) for n in names]
    outs = cheese_amd.search_many(specs)
    assert sum(o["rc"] != 3 for o in outs) >= 8
    for name, o in zip(names, outs):
        if o["rc"] == 3:        # a fixture written for another gate set cannot share the batch
            continue
        assert o["rc"] == manifest()[name]["rc"], name
        got = strip_seconds(_levels_text(o["levels"]))
        ref = [l for l in strip_seconds(read_log(name)) if "prefixes" in l]
        assert got == ref, name


def test_random_circuits_alone_and_batched():
    """the 32 seeded random 3-qubit circuits (10..13 gates, budget = length): identical alone and in one batch.  The
    batch runs its small levels as captured steps and a search alone as direct launches, so the two paths agree"""
    specs = [W.random_spec(seed)[1] for seed in range(32)]
    st = {}
    batched = cheese_amd.search_many(specs, stats=st)
    assert st["kernel_stats"].get("level_graph", {"launches": 0})["launches"] > 0
    for seed in (0, 5, 9, 17, 21, 31):
        alone = cheese_amd.search(specs[seed], stats=True)
        assert alone["kernel_stats"].get("level_graph", {"launches": 0})["launches"] == 0
        _same(alone, batched[seed])
    for o in batched:
        assert o["rc"] == 0 and o["verified"]


def test_searches_joining_between_levels():
    """continuous batching: searches join while others are several levels deep, with depth-0 and goal targets
    among them; every result is the one it gets alone"""
    specs = [W.random_spec(seed, lmin=4, lmax=9)[1] for seed in range(100, 124)]
    b = cheese_amd.Batch(specs[0])
    outs = {i: {"levels": []} for i in range(len(specs) + 2)}
    order = list(range(len(specs)))
    b.add(specs[order.pop(0)], 0)
    b.add(specs[1], 1, depth=0)              # budget 0, target not the goal: proof at once
    order.remove(1)
    n3 = 3
    ident = W.spec_from_skeleton(W.skeleton(n3), W.unitary(n3, []), 5)
    b.add(ident, len(specs))                # the goal itself: found with no gates
    steps = 0
    while True:
        for tag, kind, d in b.events():
            if kind == "level":
                outs[tag]["levels"].append(d)
            else:
                outs[tag].update(d)
        if order and steps % 2 == 0:
            for _ in range(3):
                if order:
                    i = order.pop(0)
                    b.add(specs[i], i)
        if not b.live and not order:
            break
        assert b.step() == 0
        steps += 1
    b.close()
    assert outs[1]["rc"] == 1 and outs[1]["levels"] == []
    assert outs[len(specs)]["rc"] == 0 and outs[len(specs)]["circuit_idx"] == []
    for i in [0] + list(range(2, len(specs))):
        alone = cheese_amd.search(specs[i])
        o = outs[i]
        o["verified"] = o["rc"] != 0 or cheese_amd.verify_spec(specs[i], o["circuit_idx"])
        _same(alone, o)


def test_batch_rejects_other_gate_sets():
    toff = read_spec("toffoli-d8")
    b = cheese_amd.Batch(toff)
    assert b.add(read_spec("cs-d4"), 0) == 3          # two-qubit spec in a three-qubit batch
    assert b.add("cheese-spec 1\nbogus\n", 1) == 2
    b.close()


def test_reused_context_repeats_its_log():
    """a context reused for repeated searches (buffers, tree pool and instance slots persist) gives the golden
    log every time"""
    ref = [l for l in strip_seconds(read_log("toffoli-d9")) if "prefixes" in l]
    s = cheese_amd.Searcher(read_spec("toffoli-d9"))
    try:
        for _ in range(3):
            out = s.run()
            assert out["rc"] == 1
            assert strip_seconds(_levels_text(out["levels"])) == ref
    finally:
        s.close()


def test_autotuned_schedule_keeps_the_log():
    """chz_tune tries every expansion schedule on the context's own search; the chosen table changes timing only"""
    ref = [l for l in strip_seconds(read_log("toffoli-d9")) if "prefixes" in l]
    s = cheese_amd.Searcher(read_spec("toffoli-d9"))
    try:
        table = s.tune(1)
        assert table and all(v["waves_per_eu"] in (4, 5, 6) for v in table.values())
        out = s.run()
        assert strip_seconds(_levels_text(out["levels"])) == ref
    finally:
        s.close()


@pytest.mark.parametrize("budget", [2e6, 4e7])
def test_memory_planning_makes_searches_wait_without_changing_them(budget):
    """a batch whose per-level memory budget is far below what its levels would take: searches wait (their nodes
    carried unexpanded) and every log and circuit is still the reference's"""
    names = [n for n in golden_cases(max_seconds=10)
             if manifest()[n]["n"] == 3 and not manifest()[n]["flags"] and not n.startswith(("rus13", "t-ladder"))]
    specs = [read_spec(n) for n in names]
    st = {}
    outs = cheese_amd.search_many(specs, budget=budget, stats=st)
    print(f"budget {budget:.0e}: {st['waits']} search-levels waited", flush=True)
    assert st["waits"] > 0
    for name, o in zip(names, outs):
        if o["rc"] == 3:
            continue
        m = manifest()[name]
        assert o["rc"] == m["rc"], name
        lines = [cheese_amd.level_line(l) for l in o["levels"]]
        if o["rc"] == 0:
            lines.append(cheese_amd.result_line(o))
        assert strip_seconds("\n".join(lines)) == strip_seconds(read_log(name)), name
