"""The learned policy's seeded candidates (LearnedParams.seeds) against the
oracle's seeded restatement.  A predictor whose seed heads rank a fixed
list of seeds for every block (zero seed-head weights, descending biases
on the listed seeds) with no fallback, no early out and every partition
count kept searches each count from exactly those lists: the selected
seeds, ranked by the exhaustive search's line fits when more than 4 are
kept, taken as listed otherwise.

* the candidates each block searches (label buffers) equal the oracle's at
  4x4, 6x6 and 8x8 (the float4-staged and the RGBA8-staged kernel), also
  with the first partitionings of the mismatch order fitted before the
  seeds (mismatch_candidates);
* at 4x4, where every trial mode can be kept, the encoded blocks equal the
  oracle's encoding from the same lists, bit for bit;
* hidden width 256 runs the fused inference and the per-count seed top-k
  kernel, width 128 the separate GEMM / top-k kernels; both agree.
"""
import numpy as np
import pytest

import astc_implementation_amd as A
from astc_implementation_amd import synth
from astc_implementation_amd.predictor import trial_mode_count
import oracle_lib as O

pytestmark = pytest.mark.gpu


def _seed_lists(bx, by, k, rng):
    """(3, k) seed lists for 2, 3, 4 partitions: selected seeds in random
    order with one unselected seed (where the count has one) among them."""
    out = []
    for pc in (2, 3, 4):
        nsel = O.footprint_info(bx, by)["selected_p%d" % pc]
        sel = [O.partition_entry(bx, by, pc, int(e))[1]["seed"] for e in rng.permutation(nsel)[:k]]
        if nsel < 1024:
            sel[int(rng.integers(0, k))] = O.partition_entry(bx, by, pc, int(rng.integers(nsel, 1024)))[1]["seed"]
        out.append(sel)
    return np.array(out, np.int32)


def _crafted_predictor(bx, by, hidden, lists, seed):
    rng = np.random.default_rng(seed)
    nin, nm = 3 * bx * by, trial_mode_count(bx, by)
    nout = 4 + 3 * 1024 + nm
    W1 = rng.normal(0, 0.2, (nin, hidden))
    b1 = rng.normal(0, 0.1, hidden)
    W2 = rng.normal(0, 0.1, (hidden, hidden))
    b2 = rng.normal(0, 0.1, hidden)
    W3 = rng.normal(0, 0.1, (hidden, nout))
    b3 = rng.normal(0, 0.1, nout)
    W3[:, 4:4 + 3 * 1024] = 0.0
    b3[4:4 + 3 * 1024] = -20.0
    for q, lst in enumerate(lists):
        for r, sd in enumerate(lst):
            b3[4 + q * 1024 + sd] = 10.0 - 0.25 * r
    return [np.asarray(w, np.float32) for w in (W1, b1, W2, b2, W3, b3)]


_FULL = dict(modes=64, slots=8, fallback_below=0.0, count_below=0.0, early_out=0.0, full_search_above=0.0)


def _oracle_seed_of(bx, by, pc, entry):
    return -1 if entry < 0 else O.partition_entry(bx, by, pc, entry)[1]["seed"]


@pytest.mark.parametrize("fp", [(4, 4), (6, 6), (8, 8)])
@pytest.mark.parametrize("k,hidden,mis", [(12, 256, 0), (3, 256, 0), (12, 128, 0), (12, 256, 16)])
def test_seeded_candidates_match_oracle(fp, k, hidden, mis):
    """mis > 0: the first `mis` partitionings of each block's mismatch order
    are line-fitted before the seeds (mismatch_candidates)."""
    bx, by = fp
    rng = np.random.default_rng(500 + k + hidden + bx)
    lists = _seed_lists(bx, by, k, rng)
    img = synth.alpha_image(16 * bx + 3, 10 * by + 1, 41) if k == 3 else synth.gradient_image(16 * bx + 3, 10 * by + 1, 40)
    h, w = img.shape[:2]
    enc = A.Encoder(0)
    try:
        enc.set_predictor(bx, by, _crafted_predictor(bx, by, hidden, lists, 7), hidden=hidden, seeds=k,
                          mismatch_candidates=mis, **dict(_FULL, modes=16))
        enc.configure(w, h, bx, by, policy=A.POLICY_LEARNED, labels=True)
        enc.encode(img)
        seeds, _ = enc.candidates()
    finally:
        enc.close()
    nbx = (w + bx - 1) // bx
    for b in range(seeds.shape[0]):
        blk = img[(b // nbx) * by:(b // nbx) * by + by, (b % nbx) * bx:(b % nbx) * bx + bx]
        if blk.shape[:2] != (by, bx):
            continue  # edge blocks: the oracle's one-block view would not clamp the same way
        for q, pc in enumerate((2, 3, 4)):
            exp = [_oracle_seed_of(bx, by, pc, e) for e in O.seeded_candidates(blk, bx, by, 0, pc, lists[q], mis)]
            assert seeds[b, q].tolist() == exp, (fp, k, hidden, b, pc)


@pytest.mark.parametrize("k,hidden,mis", [(12, 256, 0), (3, 256, 0), (12, 128, 0), (12, 256, 16)])
@pytest.mark.parametrize("kind", ["gradient", "alpha"])
def test_seeded_learned_encode_equals_oracle(k, hidden, mis, kind):
    bx = by = 4
    assert trial_mode_count(bx, by) <= 64  # every trial mode kept
    rng = np.random.default_rng(600 + k + hidden)
    lists = _seed_lists(bx, by, k, rng)
    img = (synth.gradient_image if kind == "gradient" else synth.alpha_image)(97, 83, 42)
    enc = A.Encoder(0)
    try:
        enc.set_predictor(bx, by, _crafted_predictor(bx, by, hidden, lists, 8), hidden=hidden, seeds=k,
                          mismatch_candidates=mis, **_FULL)
        enc.configure(97, 83, bx, by, policy=A.POLICY_LEARNED)
        got = enc.encode(img)
        assert enc.fallback_rate() == 0.0
    finally:
        enc.close()
    ref, _, _ = O.encode(img, bx, by, 4, threads=8, want_symbolic=False, policy=O.POLICY_SEEDED, seed_lists=lists,
                         mismatch_candidates=mis)
    assert np.array_equal(got, ref)
