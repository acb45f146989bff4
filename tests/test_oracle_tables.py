"""Pins the oracle's tables and per-footprint metadata.

* Every lookup table the reference hard-codes is regenerated by the oracle
  from the ASTC specification formulas and must hash to the digest of the
  reference's own table (tests/golden/ref_table_hashes.json, written by
  tests/golden/make_ref_table_hashes.py from the reference sources).
* Footprint metadata is checked against counts that follow from the ASTC
  specification (valid block modes, decimation modes, selected
  partitionings) and against structural invariants.
"""
import hashlib
import json
import os

import pytest

import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_table_hashes.json")
TABLE_IDS = ["quant_to_unquant", "weight_scramble", "color_u2q", "color_pquant", "quant_mode_table",
             "prev_next", "integer_of_trits", "integer_of_quints"]
LEVELS = [2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256]


def _digest(values):
    return hashlib.sha256(",".join(str(int(v)) for v in values).encode()).hexdigest()


@pytest.mark.parametrize("tid,name", list(enumerate(TABLE_IDS)))
def test_table_matches_reference_digest(tid, name):
    golden = json.load(open(GOLDEN))[name]
    vals = O.table(tid)
    assert len(vals) == golden["count"]
    assert _digest(vals) == golden["sha256"]


def test_ise_bitcount_closed_form():
    # bits*n + ceil(8n/5) for trits, ceil(7n/3) for quints
    for q, n_levels in enumerate(LEVELS):
        trits = n_levels % 3 == 0
        quints = n_levels % 5 == 0
        base = n_levels // 3 if trits else (n_levels // 5 if quints else n_levels)
        bits = base.bit_length() - 1
        for n in range(0, 65):
            exp = bits * n + (-(-8 * n // 5) if trits else 0) + (-(-7 * n // 3) if quints else 0)
            assert O.lib().oracle_ise_bitcount(n, q) == exp
    assert O.lib().oracle_ise_bitcount(10, 21) == 1024


def test_quant_mode_table_is_best_fit():
    qmt = O.table(4)
    for row in range(10):
        for bits in range(128):
            v = qmt[row * 128 + bits]
            if row == 0:
                assert v == -1
                continue
            fits = [q for q in range(21) if O.lib().oracle_ise_bitcount(2 * row, q) <= bits]
            assert v == (max(fits) if fits else -1)


# valid 1-plane trial modes (weight quant <= QUANT_12 grids), all encodable
# block modes, and selected partitionings for 2/3/4 partitions
SPEC_COUNTS = {
    (4, 4): dict(block_modes=145, trial_dm=6, trial_bm=53, selected_p2=437, selected_p3=329, selected_p4=321),
    (6, 6): dict(block_modes=370, trial_dm=22, trial_bm=198, selected_p2=521, selected_p3=333, selected_p4=305),
    (8, 8): dict(block_modes=555, trial_dm=46, trial_bm=339, selected_p2=701, selected_p3=476, selected_p4=464),
}


@pytest.mark.parametrize("fp", sorted(SPEC_COUNTS))
def test_footprint_counts(fp):
    info = O.footprint_info(*fp)
    for k, v in SPEC_COUNTS[fp].items():
        assert info[k] == v, (k, info[k], v)


@pytest.mark.parametrize("fp", [(4, 4), (5, 4), (6, 6), (8, 5), (8, 8), (10, 6), (12, 12)])
def test_decimation_invariants(fp):
    info = O.footprint_info(*fp)
    for m in range(info["trial_dm"]):
        tc, wc, wx, wy, texels = O.decimation_entry(fp[0], fp[1], m)
        assert tc == fp[0] * fp[1] and wc == wx * wy and wx <= fp[0] and wy <= fp[1]
        for t, entries in enumerate(texels):
            assert 1 <= len(entries) <= 4
            # bilinear infill weights of a texel sum to 16/16
            assert sum(c for _, c in entries) == 16
            assert all(0 <= w < wc for w, _ in entries)
        # corner texels sit exactly on corner weights
        assert texels[0] == [(0, 16)]
        assert texels[tc - 1] == [(wc - 1, 16)]
        if tc == wc:
            assert all(e == [(t, 16)] for t, e in enumerate(texels))


@pytest.mark.parametrize("fp", [(4, 4), (6, 6), (8, 8)])
@pytest.mark.parametrize("pc", [2, 3, 4])
def test_selected_partitionings_are_valid_and_distinct(fp, pc):
    info = O.footprint_info(*fp)
    nsel = info["selected_p%d" % pc]
    seen = set()
    seeds = set()
    for idx in range(1024):
        pot, meta = O.partition_entry(fp[0], fp[1], pc, idx)
        seeds.add(meta["seed"])
        if idx < nsel:
            assert meta["partition_count"] == pc
            assert sorted(set(pot)) == list(range(pc))
            assert [pot.count(p) for p in range(4)] == meta["counts"]
            # canonical form: relabel partitions by first appearance
            remap, canon = {}, []
            for p in pot:
                remap.setdefault(p, len(remap))
                canon.append(remap[p])
            key = tuple(canon)
            assert key not in seen
            seen.add(key)
    # the packed table is a permutation of all 1024 seeds
    assert seeds == set(range(1024))


def test_one_partition_entry():
    pot, meta = O.partition_entry(4, 4, 1, 0)
    assert pot == [0] * 16 and meta["counts"][0] == 16
