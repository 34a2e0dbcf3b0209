"""The device-code generator (unicorn_amd/csrc/gen_chem.py), on the CPU.

The split forms of the kernels rest on properties of the generated text: each part of the
rate kernel writes its reactions' slots with exactly the statements of the one-lane
function, and the parts of the split sweep divide every species block's terms without
losing or repeating one.  These tests check those properties and that the committed
chem_gen.inc is what the generator writes.
"""
from __future__ import annotations

import importlib.util
import re
from collections import Counter
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
GEN = ROOT / "unicorn_amd" / "csrc" / "gen_chem.py"


@pytest.fixture(scope="module")
def gen():
    spec = importlib.util.spec_from_file_location("gen_chem_under_test", GEN)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="module")
def text(gen, tmp_path_factory):
    out = tmp_path_factory.mktemp("gen") / "chem_gen.inc"
    gen.main(str(out))
    return out.read_text()


def body(text, name):
    i = text.index(f"void {name}(")
    return text[i:text.index("\n}\n", i)]


STORE = re.compile(r"rc\.st\((\d+), [^;]*;")


def test_committed_file_is_generated(text):
    assert (ROOT / "unicorn_amd" / "csrc" / "chem_gen.inc").read_text() == text


@pytest.mark.parametrize("parts", [2, 4, 8])
def test_rate_parts_write_every_slot_once_with_the_one_lane_statement(text, parts):
    full = {m.group(1): m.group(0) for m in STORE.finditer(body(text, "uc_rates"))}
    assert len(full) == 537
    seen = Counter()
    for p in range(parts):
        b = body(text, f"uc_rates_s{parts}_{p}")
        assert "#pragma clang fp contract(off)" in b
        assert ("UC_SLOT_RHOFA" in b) == (p == 0)   # rho*Y_k: part 0 only
        for m in STORE.finditer(b):
            seen[m.group(1)] += 1
            assert m.group(0) == full[m.group(1)], (parts, p, m.group(1))
    assert set(seen) == set(full) and set(seen.values()) == {1}


def test_rate_parts_balance_transcendentals(gen):
    t = gen.load()
    costs = [gen.rate_cost(r) for r in t["reactions"]]
    assert sum(costs) == 673
    for parts in (2, 4, 8):
        load = [0] * parts
        for i in sorted(range(len(costs)), key=lambda i: (-costs[i], i)):
            p = min(range(parts), key=lambda p: (load[p], p))
            load[p] += costs[i]
        assert max(load) - min(load) <= max(costs), load


@pytest.mark.parametrize("lanes", [2, 4])
def test_split_sweep_partitions_every_block(gen, lanes):
    t = gen.load()
    blocks = gen.sweep_terms(t)
    parts = gen.split_terms(blocks, lanes)
    assert len(parts) == len(blocks)
    for (k, A, B, Lo), pb in zip(blocks, parts):
        prod = [x for x in (A or []) + B if x != "0.0"]
        loss = [x for x in Lo if x != "0.0"]
        assert Counter(x for w in range(lanes) for x in pb[w][0]) == Counter(prod), k
        assert Counter(x for w in range(lanes) for x in pb[w][1]) == Counter(loss), k
        # each part keeps the one-lane order of its share
        for w in range(lanes):
            for mine, whole in ((pb[w][0], prod), (pb[w][1], loss)):
                pos = [whole.index(x) for x in mine]
                assert pos == sorted(pos), (k, w)
