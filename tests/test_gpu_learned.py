"""The trained candidate predictor (astc_implementation_amd/predictor_4x4.npz,
trained on the EXHAUSTIVE search's outcomes) in the LEARNED policy: what the
network gains over the same pruning without it, its quality bound against
the plain searches on an image it was not trained on (training images use
seeds 10000..19999), and the effect of each pruning setting."""
import os

import numpy as np
import pytest

import astc_implementation_amd as A
from astc_implementation_amd import synth
from astc_implementation_amd.train_predictor import load_params

pytestmark = pytest.mark.gpu
WEIGHTS = os.path.join(os.path.dirname(A.__file__), "predictor_4x4.npz")


@pytest.fixture(scope="module")
def setup():
    if not os.path.exists(WEIGHTS):
        pytest.skip("no trained predictor")
    params, hidden = load_params(WEIGHTS)
    img = synth.gradient_image(512, 512, 3)
    enc = A.Encoder(0)
    enc.configure(512, 512, 4, 4)
    plain = enc.encode(img)
    yield enc, img, params, hidden, A.psnr(enc.decode(4, 4, 512, 512, plain), img)
    enc.close()


def _learned_psnr(enc, img, params, hidden, **lp):
    enc.set_predictor(4, 4, params, hidden=hidden, **lp)
    enc.configure(512, 512, 4, 4, policy=A.POLICY_LEARNED)
    blocks = enc.encode(img)
    return A.psnr(enc.decode(4, 4, 512, 512, blocks), img), enc.fallback_rate()


@pytest.fixture(scope="module")
def p_exhaustive(setup):
    enc, img, _, _, _ = setup
    enc.configure(512, 512, 4, 4, policy=A.POLICY_EXHAUSTIVE)
    return A.psnr(enc.decode(4, 4, 512, 512, enc.encode(img)), img)


@pytest.mark.parametrize("prune", [dict(modes=16, candidates=1), dict(modes=53, seeds=32, candidates=1)])
def test_trained_network_beats_uninformed_pruning(setup, p_exhaustive, prune):
    """Same pruning against the exhaustive search (every partition count, no
    fallback, no early out, 8 slots): 16 of the 53 trial modes with the
    exhaustive partition candidates, or every mode with 32 seeds per count
    ranked by their line fits.  The trained network against an all-zero one
    (every logit equal: the first 16 trial modes, the 32 lowest seeds) must
    keep less than half of the uninformed pruning's PSNR loss."""
    enc, img, params, hidden, _ = setup
    lp = dict(fallback_below=0.0, count_below=0.0, slots=8, **prune)
    p_net, _ = _learned_psnr(enc, img, params, hidden, **lp)
    zero = [np.zeros_like(w) for w in params]
    p_zero, _ = _learned_psnr(enc, img, zero, hidden, **lp)
    assert p_net > p_zero, (p_exhaustive, p_net, p_zero)
    assert p_exhaustive - p_net < 0.5 * (p_exhaustive - p_zero), (p_exhaustive, p_net, p_zero)


def test_default_learned_policy_quality(setup):
    """The library's default learned parameters (16 modes, 5 slots, the
    COMPAT positions, no seeds) stay within 0.1 dB of the COMPAT search they
    prune."""
    enc, img, params, hidden, p_plain = setup
    p_net, fb = _learned_psnr(enc, img, params, hidden)
    assert p_plain - p_net <= 0.1, (p_plain, p_net)
    assert fb < 0.05


@pytest.mark.parametrize("bx", [4, 6])
def test_shipped_settings_within_bound_of_exhaustive(bx, record_property):
    """The shipped predictors with the pruning saved in their files, against
    the EXHAUSTIVE search on a 512x512 image none of them was trained or
    tuned on: within 0.05 dB (decoded RGBA PSNR)."""
    img = synth.gradient_image(512, 512, 3)
    enc = A.Encoder(0)
    try:
        enc.configure(512, 512, bx, bx, policy=A.POLICY_EXHAUSTIVE)
        p_ex = A.psnr(enc.decode(bx, bx, 512, 512, enc.encode(img)), img)
        enc.load_predictor(bx, bx)
        enc.configure(512, 512, bx, bx, policy=A.POLICY_LEARNED)
        p_learned = A.psnr(enc.decode(bx, bx, 512, 512, enc.encode(img)), img)
        record_property("dpsnr_db", p_ex - p_learned)
        assert p_ex - p_learned <= 0.05, (bx, p_ex, p_learned)
    finally:
        enc.close()


def _bf16(a):
    """float32 -> nearest bfloat16 (ties to even), as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32)


def _seed_logits(params, img, bx, by):
    """Host forward pass of the network to its 3 x 1024 seed-head logits,
    rounding to bfloat16 where the kernels do (inputs, weights, hidden
    activations)."""
    from scipy.special import erf
    h, w = img.shape[:2]
    bxn, byn = (w + bx - 1) // bx, (h + by - 1) // by
    ys = np.minimum(np.arange(byn)[:, None] * by + np.arange(by)[None, :], h - 1)
    xs = np.minimum(np.arange(bxn)[:, None] * bx + np.arange(bx)[None, :], w - 1)
    blk = img[ys[:, None, :, None], xs[None, :, None, :], :3]  # (byn, bxn, by, bx, 3)
    x = _bf16(blk.reshape(bxn * byn, -1).astype(np.float32) * np.float32(1.0 / 255.0))
    W1, b1, W2, b2, W3, b3 = params

    def gelu(z):
        return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))
    h1 = _bf16(gelu(x @ _bf16(W1) + b1))
    h2 = _bf16(gelu(h1 @ _bf16(W2) + b2))
    return h2 @ _bf16(W3[:, 4:4 + 3072]) + b3[4:4 + 3072]


def test_ranked_partitionings_keep_the_seed_heads_best(setup):
    """partitionings=1 keeps, per block and partition count, only the compat
    position whose seed the seed head scores highest.  The plain search's
    best seed of a count is one of the 4 compat positions, so the seed the
    pruned search used (its label) must score at least as high as it.  The
    host forward pass differs from the kernels in summation order, hence the
    0.05 tolerance on the logits."""
    enc, img, params, hidden, _ = setup
    enc.configure(512, 512, 4, 4, labels=True)
    enc.encode(img)
    lab_plain = enc.labels()
    enc.set_predictor(4, 4, params, hidden=hidden, fallback_below=0.0, partitionings=1)
    enc.configure(512, 512, 4, 4, labels=True, policy=A.POLICY_LEARNED)
    blocks = enc.encode(img)
    lab_one = enc.labels()
    logits = _seed_logits(params, img, 4, 4)
    rows = np.arange(len(lab_one))
    same = 0
    for c in range(3):
        s_one, s_plain = lab_one[:, 3 + c], lab_plain[:, 3 + c]
        assert (s_one >= 0).all() and (s_plain >= 0).all()
        l_one = logits[rows, c * 1024 + s_one]
        l_plain = logits[rows, c * 1024 + s_plain]
        bad = np.nonzero(l_one < l_plain - 0.05)[0]
        assert bad.size == 0, (c, bad[:8], l_one[bad[:8]], l_plain[bad[:8]])
        same += int((s_one == s_plain).sum())
    # the seed head ranks the plain search's winner first more often than
    # chance (1 in 4; measured 0.35 on this image with the head trained on
    # the exhaustive search's winners)
    assert same > 0.3 * 3 * len(rows), same
    dec = enc.decode(4, 4, 512, 512, blocks)
    assert A.psnr(dec, img) > 40.0


def test_partitionings_four_is_the_default(setup):
    enc, img, params, hidden, _ = setup
    outs = []
    for k in (0, 4):
        enc.set_predictor(4, 4, params, hidden=hidden, partitionings=k)
        enc.configure(512, 512, 4, 4, policy=A.POLICY_LEARNED)
        outs.append(enc.encode(img))
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("keep", [1, 2])
def test_refine_partitionings_limits_refined_positions(setup, keep):
    """refine_partitionings=k: per block and partition count at most k of
    the 4 positions get refined candidates (the others report no error),
    every block stays a valid encoding, and blocks that fall back keep the
    full search."""
    enc, img, params, hidden, p_plain = setup
    enc.set_predictor(4, 4, params, hidden=hidden, fallback_below=0.0, refine_partitionings=keep)
    enc.configure(512, 512, 4, 4, labels=True, policy=A.POLICY_LEARNED)
    blocks = enc.encode(img)
    _, errors = enc.candidates()
    refined = (errors < 1e29).sum(axis=2)
    assert refined.max() <= keep and refined.min() >= 1
    assert p_plain - A.psnr(enc.decode(4, 4, 512, 512, blocks), img) < 1.0
    enc.set_predictor(4, 4, params, hidden=hidden, fallback_below=2.0, refine_partitionings=keep)
    enc.configure(512, 512, 4, 4, labels=True, policy=A.POLICY_LEARNED)
    enc.encode(img)
    assert enc.fallback_rate() == 1.0
    _, errors = enc.candidates()
    assert ((errors < 1e29).sum(axis=2) == 4).all()


def _reference_error_limit(bx, by):
    """code/compression_pipeline.cpp:276-282 and :330 (float32 steps)."""
    t = np.float32(bx * by)
    db = np.float32(80.0) - np.float32(19.0) * (np.log(t, dtype=np.float32) / np.log(np.float32(10.0)))
    ws = np.float32(0.30 * 2.25) + np.float32(0.59 * 2.25) + np.float32(0.11 * 2.25) + np.float32(1.0)
    return float(np.float32(0.1) ** (db * np.float32(0.1)) * np.float32(65535.0 * 65535.0) * t * ws)


def test_early_out_skips_counts_below_the_error_limit(setup):
    """early_out=1: a block whose best encoding after a partition count is
    below the reference's error limit searches no further count (all 4
    positions of the next count report no error); blocks that stopped after
    one partition keep a single-partition encoding below the limit, and
    blocks above the limit after one partition search every count."""
    enc, img, params, hidden, p_plain = setup
    limit = _reference_error_limit(4, 4)
    enc.set_predictor(4, 4, params, hidden=hidden, early_out=1.0)
    enc.configure(512, 512, 4, 4, labels=True, policy=A.POLICY_LEARNED)
    blocks = enc.encode(img)
    _, errors = enc.candidates()
    lab = enc.labels()
    sym = enc.symbolic()
    err = np.array([s.errorval for s in sym], np.float32)
    skipped2 = (errors[:, 0, :] >= 1e29).all(axis=1)
    assert 0 < skipped2.sum() < len(err)
    assert (lab[skipped2, 0] == 1).all()
    assert (err[skipped2] < limit).all()
    # a block searched at 2 partitions was above the limit after 1 partition,
    # so its final error cannot be below the limit through a 1-partition block
    searched = ~skipped2
    assert not ((lab[searched, 0] == 1) & (err[searched] < limit) & (errors[searched, 0, :] >= 1e29).all(axis=1)).any()
    assert p_plain - A.psnr(enc.decode(4, 4, 512, 512, blocks), img) < 0.5


@pytest.mark.parametrize("bx,by", [(4, 4), (6, 6)])
def test_shipped_predictor_settings(bx, by):
    """The committed predictors with the pruning saved in their files
    (Encoder.load_predictor), on an image outside their training set: the
    decoded PSNR stays within 0.05 dB of the exhaustive search, every block
    is a legal encoding and the early out skips partitioned searches for
    part of the blocks."""
    img = synth.gradient_image(384, 320, 7)
    enc = A.Encoder(0)
    enc.configure(384, 320, bx, by, policy=A.POLICY_EXHAUSTIVE)
    plain = enc.encode(img)
    p_plain = A.psnr(enc.decode(bx, by, 384, 320, plain), img)
    enc.load_predictor(bx, by)
    enc.configure(384, 320, bx, by, policy=A.POLICY_LEARNED, labels=True)
    learned = enc.encode(img)
    dec = enc.decode(bx, by, 384, 320, learned)
    assert not np.all(dec == np.array([255, 0, 255, 255], np.uint8), axis=-1).any()
    assert p_plain - A.psnr(dec, img) <= 0.05, (p_plain, A.psnr(dec, img))
    _, errors = enc.candidates()
    skipped = (errors[:, 0, :] >= 1e29).all(axis=1)
    assert 0 < skipped.mean() < 1
    enc.close()


def test_full_search_above_marks_hard_blocks(setup):
    """full_search_above: blocks whose one-partition error is above the
    multiple of the error limit are searched like the plain search from two
    partitions on (counted as fallback), which lowers the search's own
    objective (the summed weighted block errors); a multiple no block
    reaches changes nothing."""
    enc, img, params, hidden, _ = setup
    limit = _reference_error_limit(4, 4)
    base = dict(fallback_below=0.0, modes=8, slots=4)
    enc.set_predictor(4, 4, params, hidden=hidden, **base)
    enc.configure(512, 512, 4, 4, policy=A.POLICY_LEARNED)
    ref = enc.encode(img)
    enc.set_predictor(4, 4, params, hidden=hidden, full_search_above=1e12, **base)
    enc.configure(512, 512, 4, 4, policy=A.POLICY_LEARNED)
    assert np.array_equal(enc.encode(img), ref)
    assert enc.fallback_rate() == 0.0
    enc.set_predictor(4, 4, params, hidden=hidden, full_search_above=20.0, **base)
    enc.configure(512, 512, 4, 4, policy=A.POLICY_LEARNED)
    out = enc.encode(img)
    rate = enc.fallback_rate()
    assert 0.0 < rate < 1.0
    err_full = sum(b.errorval for b in enc.symbolic())
    enc.set_predictor(4, 4, params, hidden=hidden, **base)
    enc.configure(512, 512, 4, 4, policy=A.POLICY_LEARNED)
    enc.encode(img)
    assert err_full < sum(b.errorval for b in enc.symbolic())
    assert not np.array_equal(out, ref)


def test_rounds_six_is_the_default_and_fewer_rounds_stay_valid(setup):
    enc, img, params, hidden, p_plain = setup
    outs = {}
    for r in (0, 6, 2):
        enc.set_predictor(4, 4, params, hidden=hidden, fallback_below=0.0, rounds=r)
        enc.configure(512, 512, 4, 4, policy=A.POLICY_LEARNED)
        outs[r] = enc.encode(img)
    assert np.array_equal(outs[0], outs[6])
    dec = enc.decode(4, 4, 512, 512, outs[2])
    assert not np.all(dec == np.array([255, 0, 255, 255], np.uint8), axis=-1).any()
    assert p_plain - A.psnr(dec, img) < 0.5


def test_partitioned_slots_default_and_valid(setup):
    enc, img, params, hidden, p_plain = setup
    outs = {}
    for ps in (0, 5, 2):
        enc.set_predictor(4, 4, params, hidden=hidden, fallback_below=0.0, slots=5, partitioned_slots=ps)
        enc.configure(512, 512, 4, 4, policy=A.POLICY_LEARNED)
        outs[ps] = enc.encode(img)
    assert np.array_equal(outs[0], outs[5])
    assert not np.array_equal(outs[2], outs[5])
    dec = enc.decode(4, 4, 512, 512, outs[2])
    assert not np.all(dec == np.array([255, 0, 255, 255], np.uint8), axis=-1).any()
    assert p_plain - A.psnr(dec, img) < 0.5


def test_count_below_prunes_counts_by_the_count_head(setup):
    """count_below: a block searches a partition count only when the count
    head ranks it first or gives it at least count_below.  The counts a block
    still searches see exactly the candidates they see without pruning (the
    k-means kernel and the plan skip the same blocks), and with count_below
    above every probability a block searches at most its top-ranked count.
    Early out, fallback and the hard-block search are off, so nothing else
    decides which counts are searched."""
    enc, img, params, hidden, _ = setup
    base = dict(fallback_below=0.0, early_out=0.0, full_search_above=0.0)

    def run(cb):
        enc.set_predictor(4, 4, params, hidden=hidden, count_below=cb, **base)
        enc.configure(512, 512, 4, 4, labels=True, policy=A.POLICY_LEARNED)
        enc.encode(img)
        return enc.candidates()

    s0, e0 = run(0.0)
    searched0 = (e0 < 1e29).any(axis=2)  # (block, count 2..4)
    assert searched0.all()
    counts = []
    for cb in (0.3, 1.01):
        s, e = run(cb)
        searched = (e < 1e29).any(axis=2)
        counts.append(int(searched.sum()))
        assert 0 < searched.sum() < searched0.sum()
        assert np.array_equal(e[searched], e0[searched])
        assert np.array_equal(s[searched], s0[searched])
    assert (searched.sum(axis=1) <= 1).all()
    assert counts[1] <= counts[0]
