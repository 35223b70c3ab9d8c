"""The committed fixtures in tests/golden/ reproduce: the seeded generators
give the same instances (digests), and the oracle gives the same costs (a
sample of each configuration; the GPU tests compare against every entry)."""
import hashlib
import json
import os

import pytest

import cuda_a_star_amd as ga
import oracle
from cuda_a_star_amd import workloads as wl

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def digest(cells):
    return hashlib.sha256(cells.tobytes()).hexdigest()[:32]


def test_generator_digests():
    g = load("generators.json")
    for p in g["puzzle"]:
        assert ga.gen_puzzle(p["side"], p["depth"], p["seed"], p["no_backtrack"]) == int(p["state"], 16)
    for q in g["grid"]:
        cells, draws = ga.gen_grid(q["width"], q["height"], q["wall_p"], q["seed"])
        assert draws == q["draws"] and digest(cells) == q["cells_sha256"]


def test_config1_costs():
    for x, g in zip([wl.config1(s) for s in range(64)], load("config1.json")["instances"]):
        assert x.start == int(g["start"], 16)
        rc, res, _ = oracle.astar_puzzle(3, x.start, x.goal, want_path=False)
        assert res.cost == g["cost"]


def test_config2_costs():
    for x, g in zip([wl.config2(s) for s in range(16)], load("config2.json")["instances"]):
        assert x.start == int(g["start"], 16)
        if g["oracle_expanded"] > 700_000:
            continue  # the expensive ones are covered on the GPU side
        rc, res, _ = oracle.astar_puzzle(4, x.start, x.goal, want_path=False)
        assert res.cost == g["cost"]


def test_config3_instance_and_cost():
    # the 4096 x 4096 grid: the generator's cells, and Dijkstra's cost on them
    x, g = wl.config3(0), load("config3.json")["instances"][0]
    assert (x.side, x.start, x.goal) == (g["side"], 0, 4096 * 4096 - 1)
    assert digest(x.cells) == g["cells_sha256"]
    assert oracle.dial_grid(x.cells, 4096, 4096, x.start, x.goal, 4) == g["cost"]


def test_config4_instances_and_sample_costs():
    insts = wl.config4(4, 1024, threads=8)
    gold = load("config4.json")["instances"]
    assert len(insts) == len(gold) == 1024
    for x, g in zip(insts, gold):
        assert x.kind == g["kind"] and x.side == g["side"]
        if x.kind == "puzzle":
            assert x.start == int(g["start"], 16)
        else:
            assert digest(x.cells) == g["cells_sha256"]
    for i in range(0, 1024, 61):
        x, g = insts[i], gold[i]
        if x.kind == "puzzle":
            if g["oracle_expanded"] > 300_000:
                continue
            rc, res, _ = oracle.astar_puzzle(4, x.start, x.goal, want_path=False)
            assert res.cost == g["cost"]
        elif x.side <= 1024:
            assert oracle.dial_grid(x.cells, x.side, x.side, x.start, x.goal, 4) == g["cost"]


def test_config5_stream_instances():
    insts = wl.config5(5, 256, 1500.0, threads=8)
    gold = load("config5.json")["instances"]
    for x, g in zip(insts, gold):
        assert x.kind == g["kind"] and x.side == g["side"]
        assert x.arrival_ns == g["arrival_ns_at_1500qps"]
        if x.kind == "puzzle":
            assert x.start == int(g["start"], 16)
        else:
            assert digest(x.cells) == g["cells_sha256"]


def test_puzzle24_pdb_fixture_boards_and_cheap_costs():
    # the deep 24-puzzle boards regenerate from the workload's stream, and the two
    # cheapest fixture costs reproduce with the pattern-database A*
    gold = load("puzzle24_pdb.json")
    cand = wl.puzzle24_deep(gold["master"], gold["tried"])
    for o in gold["instances"]:
        d, seed, b = cand[o["index"]]
        assert (d, seed, list(b)) == (o["depth"], o["seed"], o["board"])
        assert 80 <= d <= 120
    cheap = sorted(gold["instances"], key=lambda o: o["oracle_pdb_expanded"])[:2]
    for o in cheap:
        rc, res, _ = oracle.astar_board_pdb(o["board"], want_path=False)
        assert rc == oracle.OK and res.cost == o["cost"]
