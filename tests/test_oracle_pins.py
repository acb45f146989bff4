"""Pins of the oracle's float search, one pass at a time (CPU only).

The reference's WGSL passes cannot run here (no WebGPU runtime), so the
oracle's restatement of each pass is checked against what the mathematics
fixes: closed forms on special inputs (collinear and constant blocks,
weights on an exact quantisation lattice, identical angles), float64
recomputations of the quantities a pass defines, and brute-force
enumeration where a pass is a search (quant ranges, combinations, top-N,
mismatch permutations, nearest quantisation level).  Each check is chosen
so that a plausible slip (a dropped term, a wrong sign, a transposed table,
an off-by-one index, a wrong tie order) fails it.  Reference passes:
shaders/pass01..pass17 and pass001..pass007 of Jure-Z/ASTC_implementation.
"""
import itertools

import numpy as np
import pytest

import oracle_lib as O
import astc_spec as S

F32 = np.float32
CW = np.array([F32(0.30) * F32(2.25), F32(0.59) * F32(2.25), F32(0.11) * F32(2.25), F32(1.0)], np.float32)
STEPS_Q = [2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32]  # levels of QUANT_2..QUANT_32


def part_map(rng, tc, pc):
    """Random partition map with every partition used."""
    while True:
        tp = rng.integers(0, pc, tc).astype(np.uint8)
        if len(set(tp.tolist())) == pc:
            return tp


def block_px(rng, tc, alpha_const=True):
    px = rng.integers(0, 256, (tc, 4)).astype(np.float32) / F32(255.0) * F32(65536.0)
    if alpha_const:
        px[:, 3] = F32(65536.0)
    return px.astype(np.float32)


def trial_dm_of(bx, by):
    """{trial decimation index: (wx, wy, weight count)} from the trial block modes."""
    out = {}
    k = 0
    while True:
        info, n = O.trial_mode(bx, by, k)
        wx, wy, _ = S.mode_grid_and_quant(info[4])
        out.setdefault(info[0], (wx, wy, info[3]))
        k += 1
        if k >= n:
            return out


def infill_matrix(bx, by, wx, wy):
    """Texel x weight contribution matrix of the spec's bilinear infill."""
    g = S.grid(bx, by, wx, wy)
    M = np.zeros((bx * by, wx * wy))
    for t, lst in enumerate(g):
        for w, c in lst:
            M[t, w] += c / 16.0
    return M


# ---------------------------------------------------------------- pass01
@pytest.mark.parametrize("fp", [(4, 4), (6, 6)])
@pytest.mark.parametrize("pc", [1, 2, 3, 4])
def test_pass01_gray_segments_give_their_extremes_and_weights(fp, pc):
    """Texels on a segment a + t v per partition, v along (1,1,1,x): the
    positive-deviation direction is parallel to v, so the endpoints are the
    segment's extremes and the ideal weights are t (t = 0 and 1 present)."""
    bx, by = fp
    tc = bx * by
    rng = np.random.default_rng(10 * pc + bx)
    for trial in range(5):
        tp = part_map(rng, tc, pc)
        px = np.zeros((tc, 4), np.float64)
        t = rng.uniform(0, 1, tc)
        a, v = {}, {}
        for p in range(pc):
            idx = np.nonzero(tp == p)[0]
            t[idx[0]] = 0.0
            if len(idx) > 1:
                t[idx[1]] = 1.0
            a[p] = np.array([rng.uniform(2000, 20000)] * 3 + [rng.uniform(1000, 60000)])
            s = rng.uniform(3000, 40000)
            v[p] = np.array([s, s, s, s * (trial % 2)])
            px[idx] = a[p] + t[idx, None] * v[p]
        r = O.pass01(bx, by, px.astype(np.float32), tp, pc)
        for p in range(pc):
            idx = np.nonzero(tp == p)[0]
            if len(idx) < 2:
                continue
            assert np.allclose(r["ep0"][p], a[p], atol=0.05), (p, r["ep0"][p], a[p])
            assert np.allclose(r["ep1"][p], a[p] + v[p], atol=0.05)
            assert np.allclose(r["weights"][idx], t[idx], atol=2e-6)
            assert np.allclose(r["avg"][p], px[idx].mean(axis=0), rtol=1e-6)


def test_pass01_direction_of_collinear_data_is_positive_deviation_closed_form():
    """Collinear texels a + t v (any signs): the sum of squared positive
    deviations is v_c^2 S+ for v_c >= 0 and v_c^2 S- for v_c < 0 (S+- the
    squared positive/negative parts of t - mean t); the direction is that
    vector normalised, and for v >= 0 the weights are still exactly t."""
    rng = np.random.default_rng(5)
    for trial in range(40):
        t = rng.uniform(0, 1, 16)
        t[0], t[1] = 0.0, 1.0
        v = rng.uniform(-20000, 20000, 4) if trial % 2 else rng.uniform(0, 20000, 4)
        a = rng.uniform(25000, 35000, 4)
        px = (a + t[:, None] * v).astype(np.float32)
        r = O.pass01(4, 4, px, np.zeros(16, np.uint8), 1)
        d = t - t.mean()
        sp, sn = (np.maximum(d, 0) ** 2).sum(), (np.maximum(-d, 0) ** 2).sum()
        ref = v * v * np.where(v >= 0, sp, sn)
        ref /= np.linalg.norm(ref)
        assert np.allclose(r["dir"][0], ref, atol=2e-5), (r["dir"][0], ref)
        if trial % 2 == 0:
            assert np.allclose(r["weights"], t, atol=5e-6)


def test_pass01_constant_partition_takes_the_zero_span_path():
    """A constant partition: min projection >= max projection, so the span
    becomes [0, 1e-7] (pass01:140-143); its weights are 0, its weight error
    scale is (1e-7)^2 times the mean channel weight, and (normalize of the
    zero vector, which this build defines as 0/0) its direction is NaN."""
    px = np.tile(np.array([1234.5, 2000.25, 60000.0, 65536.0], np.float32), (16, 1))
    r = O.pass01(4, 4, px, np.zeros(16, np.uint8), 1)
    ew = (((CW[0] + CW[1]) + CW[2]) + CW[3]) / F32(4.0)
    assert np.all(r["weights"] == 0.0)
    assert np.all(r["wes"] == F32(1e-7) * F32(1e-7) * ew)
    assert np.all(np.isnan(r["dir"][0]))
    assert np.array_equal(r["avg"][0], px[0])


def test_pass01_min_weight_cutoff_closed_form():
    """min_weight_cutoff = min over channels and partitions of
    (1 - e0) / (e1 - e0 + 1e-9) among values in (0.5, 10), else 10: with
    the search's 0..65536 texel scale only endpoints near zero reach it."""
    tcn = 16
    base = np.zeros((tcn, 4), np.float32)
    base[:, 3] = 65536.0
    # channel 0 from 0 to 1.5: (1 - 0) / 1.5 = 0.667 < 10
    base[:8, 0] = 0.0
    base[8:, 0] = 1.5
    r = O.pass01(4, 4, base, np.zeros(tcn, np.uint8), 1)
    e0, e1 = r["ep0"][0, 0], r["ep1"][0, 0]
    assert abs(r["cutoff"] - (1.0 - e0) / (e1 - e0 + 1e-9)) < 1e-5
    assert 0.6 < r["cutoff"] < 0.7
    # ordinary image blocks: every ratio is negative or above 10
    rng = np.random.default_rng(1)
    r = O.pass01(4, 4, block_px(rng, 16), np.zeros(16, np.uint8), 1)
    assert r["cutoff"] == 10.0


# ---------------------------------------------------------------- pass02
def _pass02_ref(M, w_ideal, s):
    """Weighted-mean start (the minimiser of sum c s (g - ideal)^2 for each
    weight alone), then one clamped step -16 * (dE/dg / 2) / sum c^2 s on
    E(g) = sum_t s_t (M g - ideal)_t^2, the gradient by central differences
    in float64."""
    cs = M * s[:, None]
    g0 = (cs * w_ideal[:, None]).sum(axis=0) / (cs.sum(axis=0) + 1e-10)

    def E(g):
        r = M @ g - w_ideal
        return float((s * r * r).sum())
    h = 1e-6
    grad = np.array([(E(g0 + h * e) - E(g0 - h * e)) / (2 * h) for e in np.eye(len(g0))])
    H = (M * M * s[:, None]).sum(axis=0) + 1e-10
    step = np.clip(-16.0 * (grad / 2.0) / H, -0.25, 0.25)
    return g0, g0 + step


@pytest.mark.parametrize("fp", [(4, 4), (6, 6), (8, 8)])
def test_pass02_identity_constant_and_newton_step(fp):
    bx, by = fp
    tc = bx * by
    rng = np.random.default_rng(bx)
    dms = trial_dm_of(bx, by)
    for m, (wx, wy, wc) in sorted(dms.items()):
        w = rng.uniform(0, 1, tc).astype(np.float32)
        s = rng.uniform(1e6, 1e9, tc).astype(np.float32)
        dw = O.pass02(bx, by, m, w, s, 0)
        if wx * wy == tc:
            assert np.array_equal(dw, w)  # 1:1 grid: copied
            continue
        # a constant field is reproduced (contributions of every texel sum to one)
        c = F32(rng.uniform(0.1, 0.9))
        assert np.allclose(O.pass02(bx, by, m, np.full(tc, c, np.float32), s, 0), c, atol=1e-6)
        M = infill_matrix(bx, by, wx, wy)
        _, ref = _pass02_ref(M, w.astype(np.float64), s.astype(np.float64))
        assert np.allclose(dw, ref, atol=2e-4), (m, wx, wy, np.abs(dw - ref).max())
        # a constant error scale reads entry 0 only
        s1 = s.copy()
        s1[1:] = rng.uniform(1, 2, tc - 1)
        _, ref1 = _pass02_ref(M, w.astype(np.float64), np.full(tc, float(s1[0])))
        assert np.allclose(O.pass02(bx, by, m, w, s1, 1), ref1, atol=2e-4)


# ---------------------------------------------------------------- pass03
def _wrap(x):
    x = np.asarray(x, np.float64) % 1.0
    return np.where(x > 0.5, x - 1.0, x)


def test_pass03_identical_weights_give_their_phase():
    """All weights equal to j/1023: every step's offset is the phase of that
    one angle, frac((s + 1) j / 1023) wrapped to (-1/2, 1/2]."""
    dms = trial_dm_of(6, 6)
    rng = np.random.default_rng(3)
    for m, (wx, wy, wc) in dms.items():
        for _ in range(5):
            j = int(rng.integers(0, 1024))
            ph = _wrap((np.arange(16) + 1) * j / 1023.0)
            if np.any(np.abs(np.abs(ph) - 0.5) < 1e-4):
                continue
            ao = O.pass03(6, 6, m, np.full(wc, j / 1023.0, np.float32))
            assert np.allclose(ao, ph, atol=3e-6), (m, j, ao - ph)


def test_pass03_lattice_and_symmetric_pairs():
    """Weights on a lattice of spacing 1023/(s+1) samples (s + 1 = 3 or 11
    divides 1023) all share one phase at step s; a symmetric pair j0 +- d
    has its circular mean at j0."""
    rng = np.random.default_rng(4)
    m, (wx, wy, wc) = next((k, v) for k, v in trial_dm_of(6, 6).items() if v[2] >= 16)
    for s, spacing in ((2, 341), (10, 93)):
        j0 = int(rng.integers(0, spacing))
        js = j0 + spacing * rng.integers(0, 1023 // spacing, wc)
        js = np.minimum(js, 1023)
        ao = O.pass03(6, 6, m, (js / 1023.0).astype(np.float32))
        assert abs(ao[s] - _wrap((s + 1) * j0 / 1023.0)) < 3e-6
    for _ in range(10):
        j0 = int(rng.integers(100, 900))
        d = int(rng.integers(1, 20))
        js = np.array([j0 - d, j0 + d] * (wc // 2) + [j0] * (wc % 2))
        ao = O.pass03(6, 6, m, (js / 1023.0).astype(np.float32))
        assert np.allclose(ao[:8], _wrap((np.arange(8) + 1) * j0 / 1023.0), atol=3e-6)


# ---------------------------------------------------------------- pass04 / pass05
def _pass04_ref(dw, ao, steps, mqs):
    """float32 sample positions as the pass computes them, then float64:
    error = sum (x - rint x)^2 / (s+1)^2; cutting the lowest (highest)
    level moves those samples one level: the error changes by
    (d - 1)^2 - d^2 = 1 - 2 d  ((d + 1)^2 - d^2 = 1 + 2 d)."""
    recs = []
    dw = dw.astype(np.float32)
    for sp in range(steps):
        r = F32(sp + 1)
        sval = dw * r - F32(ao[sp])
        rr = np.rint(sval)
        d = (sval - rr).astype(np.float64)
        lo = np.rint(F32(dw.min()) * r - F32(ao[sp]))
        hi = np.rint(F32(dw.max()) * r - F32(ao[sp]))
        es = 1.0 / float(r) ** 2
        cut_low = ((d - 1) ** 2 - d ** 2)[rr == lo].sum() * es
        cut_high = ((d + 1) ** 2 - d ** 2)[rr == hi].sum() * es
        span = int(min(max(hi - lo + 1, 2), mqs + 3))
        recs.append((float(lo), span, (d * d).sum() * es, cut_low, cut_high))
    return recs


@pytest.mark.parametrize("fp", [(4, 4), (6, 6), (8, 8)])
def test_pass04_errors_match_float64_and_lattice_has_zero_error(fp):
    bx, by = fp
    rng = np.random.default_rng(bx + 40)
    for m, (wx, wy, wc) in trial_dm_of(bx, by).items():
        dw = rng.uniform(-0.1, 1.1, wc).astype(np.float32)
        ao = rng.uniform(-0.5, 0.5, 16).astype(np.float32)
        hl, mqs, mql = O.pass04(bx, by, m, dw, ao)
        ref = _pass04_ref(dw, ao, len(hl), mqs)
        for sp, (lo, span, err, cl, ch) in enumerate(ref):
            assert hl[sp, 0] == lo and int(hl[sp, 1]) == span, (m, sp)
            assert np.allclose(hl[sp, 2:], [err, cl, ch], rtol=1e-4, atol=1e-6), (m, sp, hl[sp], err, cl, ch)
        # weights on the lattice (n + off) / (s + 1): zero error at step s,
        # lowest index min n, span max n - min n + 1, each cut = count / (s+1)^2
        sp = int(rng.integers(0, len(hl)))
        off = F32(rng.uniform(-0.3, 0.3))
        n = rng.integers(0, sp + 2, wc)
        dwl = ((n + off) / F32(sp + 1)).astype(np.float32)
        ao2 = ao.copy()
        ao2[sp] = off
        hl2, _, _ = O.pass04(bx, by, m, dwl, ao2)
        assert hl2[sp, 2] < 1e-9
        assert hl2[sp, 0] == n.min()
        assert int(hl2[sp, 1]) == min(max(n.max() - n.min() + 1, 2), mqs + 3)
        es = 1.0 / (sp + 1) ** 2
        assert abs(hl2[sp, 3] - (n == n.min()).sum() * es) < 1e-5
        assert abs(hl2[sp, 4] - (n == n.max()).sum() * es) < 1e-5


def _pass05_brute(hl, ao, mql):
    """For each quant level: among all (step, variant) candidates whose span
    after cutting equals the level count -- variants: as is, cut low, cut
    high, cut both, tried in that order per step, strict < keeps the first
    -- the lowest error one; low = (offset + lowest + cut_low) / (s + 1),
    high = low + (levels - 1) / (s + 1).  No candidate: step 0, no cut."""
    cands = []
    for i, (lo, span, e, cl, ch) in enumerate(hl):
        span = int(span)
        e, cl, ch = F32(e), F32(cl), F32(ch)
        cands += [(span, e, i, 0), (span - 1, e + cl, i, 1), (span - 1, e + ch, i, 0),
                  (span - 2, (e + cl) + ch, i, 1)]
    low, high = [], []
    for q in range(mql + 1):
        qs = STEPS_Q[q]
        best = None
        for c in cands:
            if c[0] == qs and c[1] < F32(1e37) and (best is None or c[1] < best[1]):
                best = c
        i, flag = (best[2], best[3]) if best else (0, 0)
        lwi = F32(hl[i][0]) + F32(flag)
        hwi = (lwi + F32(qs)) - F32(1.0)
        st = F32(1.0) / (F32(1.0) + F32(i))
        low.append((F32(ao[i]) + lwi) * st)
        high.append((F32(ao[i]) + hwi) * st)
    return np.array(low, np.float32), np.array(high, np.float32)


@pytest.mark.parametrize("fp", [(4, 4), (6, 6), (8, 8)])
def test_pass05_matches_brute_force_over_steps_and_cuts(fp):
    bx, by = fp
    rng = np.random.default_rng(bx + 50)
    for m, (wx, wy, wc) in trial_dm_of(bx, by).items():
        for _ in range(3):
            dw = rng.uniform(0, 1, wc).astype(np.float32)
            ao = rng.uniform(-0.5, 0.5, 16).astype(np.float32)
            hl, mqs, mql = O.pass04(bx, by, m, dw, ao)
            low, high = O.pass05(bx, by, m, hl, ao)
            bl, bh = _pass05_brute(hl, ao, mql)
            assert np.array_equal(low, bl) and np.array_equal(high, bh), (m, low, bl, high, bh)


def test_passes02_to_05_recover_a_quantisation_lattice():
    """Ideal weights exactly on the QUANT_4 lattice (n + phi) / 4, 1:1 grid:
    the angular search finds phi, step 3 has (near) zero error, and the
    QUANT_4 range is [phi / 4, (phi + 3) / 4] (the 1/1023 sine-table samples
    bound the offset error)."""
    rng = np.random.default_rng(9)
    m = next(k for k, v in trial_dm_of(4, 4).items() if v[2] == 16)
    for _ in range(20):
        phi = rng.uniform(0.02, 0.2)
        n = rng.integers(0, 4, 16)
        n[:4] = [0, 1, 2, 3]
        w = ((n + phi) / 4.0).astype(np.float32)
        dw = O.pass02(4, 4, m, w, np.ones(16, np.float32), 1)
        ao = O.pass03(4, 4, m, dw)
        assert abs(ao[3] - phi) < 2e-3
        hl, _, mql = O.pass04(4, 4, m, dw, ao)
        assert hl[3, 2] < 1e-4 and int(hl[3, 1]) == 4
        low, high = O.pass05(4, 4, m, hl, ao)
        assert abs(low[2] - phi / 4) < 1e-3 and abs(high[2] - (phi + 3) / 4) < 1e-3


@pytest.mark.parametrize("fp", [(4, 4), (6, 6)])
def test_pass05_ties_keep_the_first_candidate(fp):
    """Synthetic step records with many equal errors (errors, cuts and
    spans from small sets): the first candidate in step and variant order
    wins every tie, and a level without candidates falls back to step 0."""
    bx, by = fp
    rng = np.random.default_rng(bx + 60)
    for m, (wx, wy, wc) in trial_dm_of(bx, by).items():
        _, mqs, mql = O.pass04(bx, by, m, np.zeros(wc, np.float32), np.zeros(16, np.float32))
        steps = len(O.pass04(bx, by, m, np.zeros(wc, np.float32), np.zeros(16, np.float32))[0])
        for _ in range(20):
            hl = np.zeros((steps, 5), np.float32)
            hl[:, 0] = rng.integers(-3, 3, steps)
            hl[:, 1] = rng.integers(2, mqs + 4, steps)
            hl[:, 2] = rng.choice([0.0, 0.25, 0.5], steps)
            hl[:, 3] = rng.choice([0.0, 0.25], steps)
            hl[:, 4] = rng.choice([0.0, 0.25], steps)
            ao = rng.uniform(-0.5, 0.5, 16).astype(np.float32)
            low, high = O.pass05(bx, by, m, hl, ao)
            bl, bh = _pass05_brute(hl, ao, mql)
            assert np.array_equal(low, bl) and np.array_equal(high, bh), (m, hl)


# ---------------------------------------------------------------- pass07
FREE_BITS = [111, 111 - 4 - 10, 108 - 4 - 10, 105 - 4 - 10]  # pass07:5 (bits left for weights and colours)


def _levels_of(wq):
    t = O.table(0)
    off = sum(S.LEVELS[:wq])
    return np.array(t[off:off + S.LEVELS[wq]])


@pytest.mark.parametrize("fp", [(4, 4), (6, 6), (8, 8)])
def test_pass07_nearest_level_error_and_bits(fp):
    """Each weight takes the representable level nearest to its position in
    [low, high] (brute force over all levels of the mode's quant, ties to
    the lower level); the error is the scale-weighted squared infill error
    of those levels (float64); the bit budget is the free bits of the
    partition count minus the ISE size of the weights (closed form); only
    whole groups of four weights are stored."""
    bx, by = fp
    tc = bx * by
    rng = np.random.default_rng(bx + 70)
    _, nbm = O.trial_mode(bx, by, 0)
    for k in range(0, nbm, max(1, nbm // 40)):
        (m, wq, wbits, wc, raw), _ = O.trial_mode(bx, by, k)
        wx, wy, q2 = S.mode_grid_and_quant(raw)
        assert q2 == wq and wx * wy == wc and wbits == S.ise_bitcount(wc, wq)
        pc = int(rng.integers(1, 5))
        w = rng.uniform(0, 1, tc).astype(np.float32)
        s = rng.uniform(1e6, 1e9, tc).astype(np.float32)
        dw = rng.uniform(-0.1, 1.1, wc).astype(np.float32)
        low = F32(rng.uniform(0, 0.4))
        high = F32(rng.uniform(0.6, 1.0))
        err, bits, qw = O.pass07(bx, by, k, pc, w, s, 0, 10.0, dw, low, high)
        free = FREE_BITS[pc - 1] - wbits
        if wq > 11 or free <= 0:
            assert err == F32(1e38) and bits == 0
            continue
        assert bits == free
        lev = _levels_of(wq)
        scale = F32(1.0) / (high - low)
        ix = np.clip(dw * scale - low * scale, 0, 1).astype(np.float32)
        pick = np.array([lev[np.argmin(np.abs(lev - 64.0 * x))] for x in ix])
        # exact ties between two levels go to the lower one (argmin keeps the first)
        keep = (wc // 4) * 4
        assert np.array_equal(qw[:keep], pick[:keep]), (k, qw, pick)
        assert np.all(qw[keep:] == 0)
        M = infill_matrix(bx, by, wx, wy)
        qwf = pick * (float(high) - float(low)) / 64.0 + float(low)
        ref = float((s.astype(np.float64) * (M @ qwf - w) ** 2).sum())
        assert abs(err - ref) <= 1e-4 * ref + 1e-3, (k, err, ref)


def test_pass07_range_fixups():
    """high above 1.02 x min_weight_cutoff becomes 1; an empty range becomes
    [0, 1] (pass07:230-237): both give the weights of the [low, 1] / [0, 1]
    range directly."""
    rng = np.random.default_rng(77)
    w = rng.uniform(0, 1, 16).astype(np.float32)
    s = np.ones(16, np.float32)
    dw = rng.uniform(0, 1, 16).astype(np.float32)
    k = next(k for k in range(53) if O.trial_mode(4, 4, k)[0][3] == 16 and O.trial_mode(4, 4, k)[0][1] <= 11)
    a = O.pass07(4, 4, k, 1, w, s, 0, 0.5, dw, 0.2, 0.6)
    b = O.pass07(4, 4, k, 1, w, s, 0, 10.0, dw, 0.2, 1.0)
    assert a[0] == b[0] and np.array_equal(a[2], b[2])
    c = O.pass07(4, 4, k, 1, w, s, 0, 10.0, dw, 0.7, 0.3)
    d = O.pass07(4, 4, k, 1, w, s, 0, 10.0, dw, 0.0, 1.0)
    assert c[0] == d[0] and np.array_equal(c[2], d[2])


# ---------------------------------------------------------------- pass08
def _line_err(px, point, direction):
    """Channel-weighted squared distance of RGB texels to a line (float64)."""
    with np.errstate(invalid="ignore"):
        d = direction / np.linalg.norm(direction)
    rel = px[:, :3] - point
    proj = point + np.outer(rel @ d, d)
    return float((((px[:, :3] - proj) ** 2) * CW[:3]).sum())


def _pass08_ref(px, tp, pc):
    """Per partition: the uncorrelated line through the RGB mean along the
    strongest positive-deviation axis sum, the same-chroma line through the
    origin along the mean, the luma line through the mean along (1,1,1), the
    luminance line through the origin; errors relative to the uncorrelated
    line scaled 0.7 / 1.5 / 3, and the alpha-drop error 3 x sum (a - 65536)^2."""
    out = []
    for p in range(pc):
        q = px[tp == p].astype(np.float64)
        avg = q[:, :3].mean(axis=0)
        dev = q[:, :3] - avg
        sums = [dev[dev[:, a] > 0].sum(axis=0) for a in range(3)]
        best = sums[0]
        for v in sums[1:]:
            if v @ v > best @ best:
                best = v
        e_unc = _line_err(q, avg, best)
        e_samec = _line_err(q, np.zeros(3), avg)
        e_luma = _line_err(q, avg, np.ones(3))
        e_lum = _line_err(q, np.zeros(3), np.ones(3))
        e_alpha = 3.0 * float(((q[:, 3] - 65536.0) ** 2).sum())
        out.append(((e_samec - e_unc) * 0.7, (e_luma - e_unc) * 1.5, (e_lum - e_unc) * 3.0, e_alpha,
                    max(e_unc, e_samec, e_luma, e_lum)))
    return out


@pytest.mark.parametrize("pc", [1, 2, 3, 4])
def test_pass08_line_errors_match_float64(pc):
    rng = np.random.default_rng(80 + pc)
    for _ in range(20):
        tp = part_map(rng, 16, pc)
        px = block_px(rng, 16, alpha_const=False)
        r = O.pass08(4, 4, px, tp, pc)
        for p, ref in enumerate(_pass08_ref(px, tp, pc)):
            if np.isnan(ref[4]):  # one texel: no direction (0/0), in both
                assert np.all(np.isnan(r[p, :3]) == np.isnan(ref[:3])), (p, r[p], ref)
                continue
            tol = 1e-4 * ref[4] + 1.0
            assert np.allclose(r[p, :3], ref[:3], atol=tol * 3), (p, r[p], ref)
            assert abs(r[p, 3] - ref[3]) <= 1e-5 * ref[3] + 1.0


def test_pass08_special_lines_and_flags():
    """Gray texels on a line through the origin: every line fits (errors ~0);
    coloured texels on a line through the origin: the same-chroma line
    fits, the luminance lines do not; texels on a gray-direction line off
    the gray axis: the luma line fits, the luminance and same-chroma lines
    do not.  can_offset
    needs every RGB endpoint span below 0.12 x 65536; blue contraction is
    off only for gray blocks with constant alpha."""
    t = np.linspace(0.1, 1.0, 16)
    px = np.zeros((16, 4), np.float32)
    px[:, 3] = 65536.0
    px[:, :3] = (t[:, None] * 50000.0).astype(np.float32)
    r = O.pass08(4, 4, px, np.zeros(16, np.uint8), 1)[0]
    assert np.all(np.abs(r[:3]) < 50.0) and r[3] == 0.0 and r[4] == 0.0 and r[5] == 0.0
    px[:, :3] = (t[:, None] * np.array([40000.0, 20000.0, 5000.0])).astype(np.float32)
    r = O.pass08(4, 4, px, np.zeros(16, np.uint8), 1)[0]
    assert abs(r[0]) < 50.0 and r[1] > 1e6 and r[2] > 1e6 and r[5] == 1.0
    px[:, :3] = (np.array([20000.0, 10000.0, 5000.0]) + t[:, None] * 2000.0).astype(np.float32)
    r = O.pass08(4, 4, px, np.zeros(16, np.uint8), 1)[0]
    assert abs(r[1]) < 50.0 and r[0] > 1e5 and r[2] > 1e5 and r[4] == 1.0  # span 1800 < 0.12 * 65536
    px[:, 3] = np.linspace(0, 65536, 16).astype(np.float32)
    r = O.pass08(4, 4, px, np.zeros(16, np.uint8), 1)[0]
    assert r[5] == 1.0 and r[3] > 0


# ---------------------------------------------------------------- pass09
def test_pass09_matches_float64_formulas():
    """Per quant level q (QUANT_6..QUANT_256) and integer count: the
    expected quantisation error (65536^2 / 18) / K^2 with K = levels(q) - 1
    times the channel weights and the partition size, scaled for blue
    contraction / offset encoding (not above QUANT_160), plus the range
    error of endpoints outside [0, 65535] and the format's line error; the
    6- and 4-integer columns take the cheaper of their two formats.  Levels
    below QUANT_6 hold the 1e37 default."""
    rng = np.random.default_rng(90)
    for pc in (1, 2, 3, 4):
        for _ in range(10):
            ep0 = rng.uniform(-3000, 68000, (pc, 4)).astype(np.float32)
            ep1 = rng.uniform(-3000, 68000, (pc, 4)).astype(np.float32)
            counts = rng.integers(1, 10, pc)
            ece = np.zeros((pc, 6), np.float32)
            ece[:, :4] = rng.uniform(0, 1e8, (pc, 4))
            ece[:, 4] = rng.integers(0, 2, pc)
            ece[:, 5] = rng.integers(0, 2, pc)
            err, fmt = O.pass09(pc, ep0, ep1, counts, ece)
            for p in range(pc):
                e0, e1 = ep0[p].astype(np.float64), ep1[p].astype(np.float64)
                rng_err = (np.minimum(e0, 0) ** 2 + np.minimum(e1, 0) ** 2 + np.maximum(e0 - 65535, 0) ** 2 +
                           np.maximum(e1 - 65535, 0) ** 2)
                n = float(counts[p])
                rgb_r = float(rng_err[:3] @ CW[:3]) * 0.5 * n
                a_r = float(rng_err[3]) * 0.5 * n
                wrgb, wa = float(CW[:3].sum()) * n, float(CW[3]) * n
                scl, lum, adrop = float(ece[p, 0]), float(ece[p, 2]), float(ece[p, 3])
                bc, oe = ece[p, 5] != 0, ece[p, 4] != 0
                for q in range(4):
                    assert np.all(err[p, q] == F32(1e37))
                for q in range(4, 21):
                    base = (65536.0 ** 2 / 18.0) / (S.LEVELS[q] - 1) ** 2
                    o_rgba = 0.5 if (oe and q < 19) else 1.0
                    o_rgb = 0.25 if (oe and q < 19) else 1.0
                    e8 = (wrgb + wa) * base * (0.625 if bc else 1.0) * o_rgba + rgb_r + a_r
                    rgb6 = wrgb * base * (0.5 if bc else 1.0) * o_rgb + rgb_r + adrop
                    rgbsa6 = (wrgb + wa) * base + scl + rgb_r + a_r
                    rgbs4 = wrgb * base + rgb_r + adrop + scl
                    la4 = (wrgb + wa) * base + rgb_r + a_r + lum
                    l2 = wrgb * base + rgb_r + adrop + lum
                    ref = [l2, min(rgbs4, la4), min(rgb6, rgbsa6), e8]
                    assert np.allclose(err[p, q], ref, rtol=1e-5), (p, q, err[p, q], ref)
                    assert fmt[p, q, 3] == 12 and fmt[p, q, 0] == 0
                    if abs(rgbsa6 - rgb6) > 1e-4 * rgb6:
                        assert fmt[p, q, 2] == (10 if rgbsa6 < rgb6 else 8)
                    if abs(rgbs4 - la4) > 1e-4 * la4:
                        assert fmt[p, q, 1] == (6 if rgbs4 < la4 else 4)


# ---------------------------------------------------------------- pass10 / pass11
def _random_color_tables(rng, pc, ties=False):
    vals = rng.choice([1.0, 2.0, 3.0], (4, 21, 4)) if ties else rng.uniform(0, 1e6, (4, 21, 4))
    err = vals.astype(np.float32)
    err[:, :4, :] = F32(1e37)
    fmt = rng.integers(0, 14, (4, 21, 4)).astype(np.int32)
    err[pc:] = 0
    return err, fmt


@pytest.mark.parametrize("pc", [2, 3, 4])
@pytest.mark.parametrize("ties", [False, True])
def test_pass10_is_the_brute_force_minimum(pc, ties):
    """Per quant level and total integer-count class: the minimum over all
    per-partition count choices whose classes differ by at most one (sum
    capped at 1e10), the first tuple in loop order on ties."""
    rng = np.random.default_rng(100 + pc + 10 * ties)
    for _ in range(3):
        err, fmt = _random_color_tables(rng, pc, ties)
        ce, cf = O.pass10(pc, err, fmt)
        for q in range(21):
            best = {}
            if q >= 4:
                for tup in itertools.product(range(4), repeat=pc):
                    if max(tup) - min(tup) > 1:
                        continue
                    tot = F32(0.0)
                    for p, i in enumerate(tup):
                        tot = F32(tot + err[p, q, i]) if p else err[p, q, i]
                    tot = min(tot, F32(1e10))
                    slot = sum(tup)
                    if slot not in best or tot < best[slot][0]:
                        best[slot] = (tot, [fmt[p, q, i] for p, i in enumerate(tup)])
            for slot in range(13):
                if slot in best:
                    assert ce[q, slot] == best[slot][0], (q, slot)
                    assert list(cf[q, slot][:pc]) == best[slot][1]
                else:
                    assert ce[q, slot] == F32(1e37)


@pytest.mark.parametrize("pc", [1, 2, 3, 4])
def test_pass11_is_the_brute_force_best_combination(pc):
    """For a block mode's remaining bits: over integer counts whose best
    colour quant (largest level whose ISE fits the bits, closed form) is at
    least QUANT_6, the lowest colour error; one partition skips counts that
    do not fit, more partitions stop at the first that does not (the
    reference's break); total = colour + weight error; the modified level
    (formats matched) gets 2 / 5 / 8 more bits."""
    rng = np.random.default_rng(110 + pc)
    for it in range(200):
        err, fmt = _random_color_tables(rng, pc, ties=it % 2 == 1)
        ce, cf = (None, None) if pc == 1 else O.pass10(pc, err, fmt)
        bits = int(rng.integers(0, 120))
        we = F32(rng.uniform(0, 1e5))
        tot, ql, qlm, fm = O.pass11(pc, we, bits, err, fmt, ce, cf)
        best, bidx = F32(1e37), 0
        if pc == 1:
            for ic in range(1, 5):
                q = S.quant_mode_table(ic, bits)
                if q < 4:
                    continue
                if err[0, q, ic - 1] < best:
                    best, bidx = err[0, q, ic - 1], ic
            rq = S.quant_mode_table(bidx, bits)
            assert (tot, ql, qlm) == (best + we, rq, rq), (bits, tot, ql, best)
            if rq >= 4:
                assert fm[0] == fmt[0, rq, bidx - 1]
        else:
            end = 8 if pc == 2 else 9
            for ic in range(pc, end + 1):
                q = S.quant_mode_table(ic, bits)
                if q < 4:
                    break
                if ce[q, ic - pc] < best:
                    best, bidx = ce[q, ic - pc], ic
            rq = S.quant_mode_table(bidx, bits)
            rqm = S.quant_mode_table(bidx, bits + {2: 2, 3: 5, 4: 8}[pc])
            assert (tot, ql, qlm) == (best + we, rq, rqm), (bits, tot, ql, qlm, best, rq, rqm)
            if rq >= 4:
                assert fm[:pc] == list(cf[rq, bidx - pc][:pc])


# ---------------------------------------------------------------- pass12
def test_pass12_is_a_stable_argsort():
    rng = np.random.default_rng(120)
    for n in (3, 8, 53, 198):
        for _ in range(30):
            e = rng.choice([1.0, 2.0, 3.0, 5.0, 1e37, 2e37], n).astype(np.float32)
            idx, top = O.pass12(e)
            order = [i for i in np.argsort(e, kind="stable") if e[i] < F32(1e37)][:8]
            assert list(idx[:len(order)]) == order
            assert np.all(top[len(order):] == F32(1e37))


# ---------------------------------------------------------------- pass13
def _undec(bx, by, raw, qw):
    wx, wy, _ = S.mode_grid_and_quant(raw)
    M = infill_matrix(bx, by, wx, wy)
    return M @ (np.asarray(qw[:wx * wy], np.float64) / 64.0)


def _trial_with_grid(bx, by, full):
    _, nbm = O.trial_mode(bx, by, 0)
    for k in range(nbm):
        info, _ = O.trial_mode(bx, by, k)
        if (info[3] == bx * by) == full and info[1] <= 7:
            return k, info
    raise AssertionError


@pytest.mark.parametrize("full", [True, False])
@pytest.mark.parametrize("pc", [1, 2])
def test_pass13_is_least_squares(full, pc):
    """For fixed (quantised, infilled) weights w_t the endpoints of each
    channel minimise sum_t (e0 (1 - w_t) + e1 w_t - x_t)^2 (numpy lstsq in
    float64); texels exactly on such a line are recovered exactly.  The
    scale endpoints minimise the same fit of the texels' projections on the
    partition's mean direction (RGBS: scale vector and ratio)."""
    bx, by = 4, 4
    k, (m, wq, wbits, wc, raw) = _trial_with_grid(bx, by, full)
    rng = np.random.default_rng(130 + 2 * pc + full)
    lev = _levels_of(wq)
    for trial in range(20):
        tp = part_map(rng, 16, pc)
        qw = rng.choice(lev, wc)
        w = _undec(bx, by, raw, qw)
        exact = trial % 2 == 0
        px = np.zeros((16, 4))
        for p in range(pc):
            a = rng.uniform(5000, 30000, 4)
            b = rng.uniform(35000, 60000, 4)
            idx = tp == p
            px[idx] = a + w[idx, None] * (b - a)
        if not exact:
            px += rng.normal(0, 800, px.shape)
        px = px.astype(np.float32)
        avg = np.array([px[tp == p].mean(axis=0) if p < pc else np.zeros(4) for p in range(4)], np.float32)
        e0, e1, rgbs = O.pass13(bx, by, k, px, tp, pc, avg, avg, avg, qw)
        for p in range(pc):
            idx = tp == p
            wp = w[idx]
            if wp.max() - wp.min() < 0.3 or idx.sum() < 3:
                continue  # nearly constant weights: the pass keeps its degenerate branch
            A = np.stack([1 - wp, wp], axis=1)
            sol = np.linalg.lstsq(A, px[idx].astype(np.float64), rcond=None)[0]
            tol = 0.05 if exact else 2e-3 * 65536
            assert np.allclose(e0[p], sol[0], atol=tol) and np.allclose(e1[p], sol[1], atol=tol), \
                (trial, p, e0[p], sol[0], e1[p], sol[1])
            # RGBS: projections on the mean direction (weighted by the channel
            # weights as the pass does), fitted by the same 2-parameter line
            sdir = (avg[p, :3] * CW[:3]) / CW[:3]
            sdir = sdir / np.linalg.norm(sdir)
            sc = px[idx, :3].astype(np.float64) @ sdir
            s0, s1 = np.linalg.lstsq(A, sc, rcond=None)[0]
            if s0 < s1:
                assert np.allclose(rgbs[p, :3], sdir * s1, rtol=2e-3, atol=5.0)
                assert abs(rgbs[p, 3] - s0 / s1) < 2e-3


def test_pass13_constant_weights_give_the_average():
    """All weights equal (w_min >= 0.999 w_max): both endpoints become the
    channel-weighted mean of the partition's texels and the RGBS ratio 1."""
    k, (m, wq, wbits, wc, raw) = _trial_with_grid(4, 4, True)
    lev = _levels_of(wq)
    rng = np.random.default_rng(135)
    px = block_px(rng, 16)
    qw = np.full(wc, lev[len(lev) // 2])
    avg = np.zeros((4, 4), np.float32)
    avg[0] = px.mean(axis=0)
    e0, e1, rgbs = O.pass13(4, 4, k, px, np.zeros(16, np.uint8), 1, avg, avg, avg, qw)
    assert np.allclose(e0[0], px.astype(np.float64).mean(axis=0), rtol=1e-5)
    assert np.array_equal(e0[0], e1[0]) and rgbs[0, 3] == 1.0


# ---------------------------------------------------------------- pass14 / pass15
def test_pass15_unpack_is_the_spec_decode():
    """pass15's endpoint unpacking equals the specification's LDR colour
    endpoint decode for every format the encoder emits, on random values."""
    rng = np.random.default_rng(150)
    for fmt in (0, 4, 5, 6, 8, 9, 10, 12, 13):
        for _ in range(300):
            v = rng.integers(0, 256, 8)
            e0, e1 = O.unpack_endpoints(fmt, v)
            s0, s1 = S.ldr_endpoints(fmt, [int(x) for x in v])
            assert list(e0) == list(s0) and list(e1) == list(s1), (fmt, v, e0, s0, e1, s1)


@pytest.mark.parametrize("q", [4, 8, 12, 16, 20])
def test_pass14_pack_then_spec_decode_lands_near_the_endpoints(q):
    """pack_color_endpoints then the spec decode: every packed value is a
    legal unquantised level of q, and the decoded endpoints are within a
    few quantisation steps of the requested ones (blue contraction halves
    the red / green precision; delta formats add one step); the returned
    format is one of the request's family."""
    rng = np.random.default_rng(140 + q)
    step = 255.0 / (S.LEVELS[q] - 1)
    legal = set(O.table(2)[(q - 4) * 512:(q - 3) * 512])
    family = {0: {0, 1}, 4: {4, 5}, 6: {6}, 8: {8, 9}, 10: {10}, 12: {12, 13}}
    for fmt in (0, 4, 8, 12):
        for _ in range(200):
            lo = rng.uniform(0, 200, 4)
            e0 = lo * 257.0
            e1 = np.minimum(lo + rng.uniform(0, 55, 4), 255) * 257.0
            if fmt in (0, 4):
                e0[:3] = e0[0]
                e1[:3] = e1[0]
            if fmt in (0, 8):
                e0[3] = e1[3] = 255 * 257.0
            rgbs = np.zeros(4, np.float32)
            f, v = O.pack_endpoints(e0.astype(np.float32), e1.astype(np.float32), rgbs, fmt, q)
            assert f in family[fmt], (fmt, f)
            n = 2 * (f >> 2) + 2
            assert all(int(x) in legal for x in v[:n])
            d0, d1 = S.ldr_endpoints(f, [int(x) for x in v])
            tol = 2.5 * step + 2.0
            got = np.array(d0 + d1, np.float64)
            want = np.concatenate([e0, e1]) / 257.0
            # the decoder may swap the endpoints (blue contraction); compare as sets of endpoints
            alt = np.array(d1 + d0, np.float64)
            ok = np.all(np.abs(got - want) <= tol) or np.all(np.abs(alt - want) <= tol)
            assert ok, (fmt, q, f, want, got)


# ---------------------------------------------------------------- pass16 / pass17
def _decode_error(px, tp, w, ue0, ue1):
    """pass17's error from its definition: weight rint(64 w), ASTC
    interpolation (e0 (64 - w) + e1 w + 32) >> 6 of the x257 endpoints,
    channel-weighted squared difference to the texels (float64)."""
    w1 = np.rint(np.float32(w) * F32(64.0)).astype(np.int64)
    tot = 0.0
    for t in range(len(px)):
        p = tp[t]
        col = (np.asarray(ue0[p], np.int64) * (64 - w1[t]) + np.asarray(ue1[p], np.int64) * w1[t] + 32) >> 6
        tot += float((((px[t].astype(np.float64) - col) ** 2) * CW).sum())
    return tot


@pytest.mark.parametrize("full", [True, False])
def test_pass17_is_the_spec_reconstruction_error_and_pass16_moves_one_level(full):
    rng = np.random.default_rng(160 + full)
    k, (m, wq, wbits, wc, raw) = _trial_with_grid(4, 4, full)
    lev = list(_levels_of(wq))
    slev = sorted(lev)
    raised = 0
    for trial in range(60):
        pc = int(rng.integers(1, 3))
        tp = part_map(rng, 16, pc)
        px = block_px(rng, 16)
        ue0 = rng.integers(0, 256, (pc, 4)) * 257
        ue1 = rng.integers(0, 256, (pc, 4)) * 257
        qw = rng.choice(lev, wc)
        w = _undec(4, 4, raw, qw)
        e = O.pass17(4, 4, k, px, tp, pc, ue0, ue1, qw)
        ref = _decode_error(px, tp, w.astype(np.float32), ue0, ue1)
        assert abs(e - ref) <= 1e-5 * ref + 1.0, (trial, e, ref)
        if full:  # undecimated weights are exact: the float32 sum in texel order is bit-exact
            w1 = np.rint(F32(64.0) * (np.asarray(qw, np.float32) / F32(64.0))).astype(np.int64)
            tot = F32(0.0)
            for t in range(16):
                p = tp[t]
                col = (ue0[p] * (64 - w1[t]) + ue1[p] * w1[t] + 32) >> 6
                dd = [F32(min(abs(F32(px[t, c]) - F32(col[c])), F32(1e15))) ** 2 for c in range(4)]
                et = F32(F32(F32(dd[0] * CW[0] + dd[1] * CW[1]) + dd[2] * CW[2]) + dd[3] * CW[3])
                tot = F32(tot + min(et, F32(1e37)))
            assert e == tot, (e, tot)
        # pass16 on an undecimated grid: each texel's weight goes up one level
        # when that strictly beats staying and going down, else down when that
        # strictly beats staying (float64 model, near ties skipped)
        if full:
            q2 = O.pass16(4, 4, k, px, tp, pc, ue0, ue1, qw)
            for t in range(16):
                i = slev.index(qw[t])
                cand = {}
                for d in (-1, 0, 1):
                    if 0 <= i + d < len(slev):
                        u = slev[i + d]
                        p = tp[t]
                        col = ue0[p] + (ue1[p] - ue0[p]) / 64.0 * u
                        cand[d] = float((((px[t] - col) ** 2) * CW).sum())
                eb, ed, eu = cand[0], cand.get(-1, np.inf), cand.get(1, np.inf)
                if min(abs(eu - eb), abs(ed - eb), abs(eu - ed)) < 1e-3 * eb + 1e-2:
                    continue
                want = 1 if (eu < eb and eu < ed) else (-1 if ed < eb else 0)
                assert slev.index(q2[t]) - i == want, (trial, t, eb, ed, eu, qw[t], q2[t])
        # pass16: each weight moves at most one level, and the float model
        # error it decides on (unrounded weights) never grows
        q2 = O.pass16(4, 4, k, px, tp, pc, ue0, ue1, qw)
        for a, b in zip(qw, q2):
            assert abs(slev.index(a) - slev.index(b)) <= 1, (a, b)

        def model(qq):
            ww = _undec(4, 4, raw, qq)
            tot = 0.0
            for t in range(16):
                p = tp[t]
                col = ue0[p] + (ue1[p] - ue0[p]) / 64.0 * (ww[t] * 64.0)
                tot += float((((px[t] - col) ** 2) * CW).sum())
            return tot
        assert model(q2) <= model(qw) * (1 + 1e-6) + 1.0
        if O.pass17(4, 4, k, px, tp, pc, ue0, ue1, q2) > e:
            raised += 1
    assert raised <= 3  # the integer rounding of pass17 can undo a float-model gain, rarely


@pytest.mark.parametrize("q", [4, 7, 10, 13, 16, 20])
def test_pass14_representable_endpoints_round_trip_exactly(q):
    """Endpoints that are themselves unquantised levels of q (x257, ordered
    so that no blue contraction applies) come back exactly through pack and
    the spec decode whenever the packer keeps the direct RGB / RGBA / L /
    LA format (its delta variants are lossy)."""
    rng = np.random.default_rng(170 + q)
    lv = sorted(set(O.table(2)[(q - 4) * 512:(q - 3) * 512]))
    for fmt in (0, 4, 8, 12):
        for _ in range(200):
            a = np.array(rng.choice(lv, 4))
            b = np.array(rng.choice(lv, 4))
            if fmt in (0, 4):
                a[:3] = a[0]
                b[:3] = b[0]
            if fmt in (0, 8):
                a[3] = b[3] = 255
            if a[:3].sum() > b[:3].sum():
                a, b = b, a
            if a[:3].sum() == b[:3].sum():
                continue
            f, v = O.pack_endpoints((a * 257.0).astype(np.float32), (b * 257.0).astype(np.float32),
                                    np.zeros(4, np.float32), fmt, q)
            d0, d1 = S.ldr_endpoints(f, [int(x) for x in v])
            if f != fmt:  # a delta encoding was taken: lossy by design, checked above
                continue
            assert list(d0) == list(a) and list(d1) == list(b), (fmt, q, f, a, b, d0, d1)


def test_colour_quant_table_and_direct_rgb_rounding():
    """COLOR_UNQUANT_TO_QUANT[q][2v] is the level of q nearest to v - 1/4
    and [2v + 1] the one nearest to v + 1/4 (closed form over all entries);
    the direct RGB(A) packing quantises each endpoint channel x with
    v = round(x) and the +1/4 entry unless x - v < -0.1 (pass14:414-448)."""
    t = O.table(2)
    for qi in range(17):
        row = t[qi * 512:(qi + 1) * 512]
        lv = np.array(sorted(set(row)))
        for v in range(256):
            assert row[2 * v] == lv[np.argmin(np.abs(lv - (v - 0.25)))]
            assert row[2 * v + 1] == lv[np.argmin(np.abs(lv - (v + 0.25)))]
    rng = np.random.default_rng(180)
    checked = 0
    for q in (5, 8, 11, 14):
        lv = np.array(sorted(set(t[(q - 4) * 512:(q - 3) * 512])))
        for _ in range(400):
            x0 = rng.uniform(0, 120, 4).astype(np.float32)
            x1 = (x0 + rng.uniform(60, 130, 4)).astype(np.float32)
            x1 = np.minimum(x1, F32(255.0))
            x0[3] = x1[3] = 255.0
            f, v = O.pack_endpoints(x0 * F32(257.0), x1 * F32(257.0), np.zeros(4, np.float32), 8, q)
            if f != 8 or v[0] + v[2] + v[4] > v[1] + v[3] + v[5]:
                continue  # a delta or blue-contracted encoding
            for c in range(3):
                for x, got in ((x0[c], v[2 * c]), (x1[c], v[2 * c + 1])):
                    xs = F32(min(max(F32(x * F32(257.0)), F32(0.0)), F32(65535.0))) / F32(257.0)
                    vi = np.rint(xs)
                    target = vi + 0.25 if F32(xs - vi) >= F32(-0.1) else vi - 0.25
                    assert got == lv[np.argmin(np.abs(lv - target))], (q, x, xs, got)
            checked += 1
    assert checked > 50


def test_pass16_decimated_realignment_is_sequential_coordinate_descent():
    """On a decimated grid the weights are visited in index order and each
    one moves to the better neighbouring level when that strictly lowers
    the error of the texels it touches, with the moves of earlier weights
    already applied (pass16:287-365); float64 emulation, near ties skipped."""
    rng = np.random.default_rng(190)
    k, (m, wq, wbits, wc, raw) = _trial_with_grid(4, 4, False)
    wx, wy, _ = S.mode_grid_and_quant(raw)
    M = infill_matrix(4, 4, wx, wy)
    slev = sorted(_levels_of(wq))
    compared = 0
    for trial in range(80):
        pc = int(rng.integers(1, 3))
        tp = part_map(rng, 16, pc)
        px = block_px(rng, 16).astype(np.float64)
        ue0 = rng.integers(0, 256, (pc, 4)) * 257
        ue1 = rng.integers(0, 256, (pc, 4)) * 257
        qw = list(rng.choice(slev, wc))
        got = O.pass16(4, 4, k, px.astype(np.float32), tp, pc, ue0, ue1, qw)
        cur = [float(v) for v in qw]
        tie = False
        for w in range(wc):
            texels = np.nonzero(M[:, w])[0]
            i = slev.index(int(cur[w]))

            def err(val):
                g = np.array(cur)
                g[w] = val
                inf = M @ g  # in 0..64 units
                tot = 0.0
                for t in texels:
                    p = tp[t]
                    col = ue0[p] + (ue1[p] - ue0[p]) / 64.0 * inf[t]
                    tot += float((((px[t] - col) ** 2) * CW).sum())
                return tot
            eb = err(cur[w])
            ed = err(slev[i - 1]) if i > 0 else np.inf
            eu = err(slev[i + 1]) if i + 1 < len(slev) else np.inf
            if min(abs(eu - eb), abs(ed - eb)) < 1e-4 * eb + 1e-2:
                tie = True
                break
            if eu < eb and eu < ed:
                cur[w] = slev[i + 1]
            elif ed < eb:
                cur[w] = slev[i - 1]
        if tie:
            continue
        compared += 1
        assert list(got) == [int(v) for v in cur], (trial, qw, list(got), cur)
    assert compared > 40


# ---------------------------------------------------------------- passes 001-007 (partition candidates)
@pytest.mark.parametrize("fp", [(4, 4), (6, 6), (8, 8)])
@pytest.mark.parametrize("pc", [2, 3, 4])
def test_kmeans_recovers_well_separated_clusters(fp, pc):
    """Texels drawn around pc colours far apart (spread << separation):
    farthest-point seeding puts one centre in each cluster and the four
    assign / update rounds return exactly the clusters (up to labels)."""
    bx, by = fp
    tc = bx * by
    rng = np.random.default_rng(200 + 10 * pc + bx)
    for _ in range(20):
        truth = part_map(rng, tc, pc)
        cols = rng.permutation(np.array([[5000, 5000, 5000], [60000, 5000, 5000], [5000, 60000, 5000],
                                         [5000, 5000, 60000], [60000, 60000, 60000]], np.float64))[:pc]
        px = np.zeros((tc, 4), np.float32)
        px[:, :3] = cols[truth] + rng.uniform(-300, 300, (tc, 3))
        px[:, 3] = 65536.0
        a = O.kmeans(bx, by, px, pc)
        # same partition up to relabelling
        pairs = set(zip(truth.tolist(), a.tolist()))
        assert len(pairs) == pc and len({b for _, b in pairs}) == pc, (truth, a)


def _kmeans_texels(tc):
    return [i if i < tc else 0 for i in range(64)]  # footprints up to 64 texels; entries past the block read texel 0


@pytest.mark.parametrize("fp", [(4, 4), (6, 6), (8, 8)])
@pytest.mark.parametrize("pc", [2, 3, 4])
def test_mismatch_is_the_minimum_over_partition_permutations(fp, pc):
    """pass004: half the minimum, over all matchings of k-means clusters to
    the partitioning's partitions, of the number of differing bits between
    the clusters' masks (64 k-means texels) and the partitions' coverage
    masks (the first min(texels, 64) k-means texels)."""
    bx, by = fp
    tc = bx * by
    kt = _kmeans_texels(tc)
    rng = np.random.default_rng(220 + 10 * pc + bx)
    for _ in range(40):
        assign = part_map(rng, tc, pc)
        e = int(rng.integers(0, 300))
        pot, meta = O.partition_entry(bx, by, pc, e)
        if meta["partition_count"] != pc:
            continue
        a = [0] * 4
        for l in range(64):
            a[assign[kt[l]]] |= 1 << l
        b = [0] * 4
        for l in range(min(tc, 64)):
            b[pot[kt[l]]] |= 1 << l
        best = min(sum(bin(a[p] ^ b[perm[p]]).count("1") for p in range(pc))
                   for perm in itertools.permutations(range(pc)))
        assert O.mismatch(bx, by, assign, pc, e) == best // 2


def _bitonic_256(nsel, score0):
    """pass005's network: 1024 (score, index) pairs, bitonic stages k = 2 ..
    1024, but only the 256 invocations l < 256 compare-exchange with l ^ j."""
    sc = np.array([(score0 if i == 0 else 0) if i < nsel else 0xFFFFFFFF for i in range(1024)], np.uint64)
    ix = np.arange(1024)
    k = 2
    while k <= 1024:
        j = k // 2
        while j > 0:
            for l in range(256):
                pt = l ^ j
                if pt > l:
                    asc = (l & k) == 0
                    sw = sc[l] > sc[pt] or (sc[l] == sc[pt] and ix[l] > ix[pt])
                    if sw == asc:
                        sc[l], sc[pt] = sc[pt], sc[l]
                        ix[l], ix[pt] = ix[pt], ix[l]
            j //= 2
        k *= 2
    return ix[:128]


@pytest.mark.parametrize("nsel", [100, 200, 321, 437, 701])
def test_ordering_of_the_partial_bitonic_network(nsel):
    """pass005 with only entry 0 scored (the compat stage's case): with a
    zero score the first 128 positions are 0..127 (what a full sort gives);
    with a non-zero score entry 0 leaves the first 128 when there are more
    than 128 selected partitionings, and the output equals the 256-lane
    network of the reference run here independently."""
    assert list(O.ordering(nsel, 0)) == list(range(128))
    for s0 in (1, 5, 40):
        got = list(O.ordering(nsel, s0))
        assert got == list(_bitonic_256(nsel, s0))
        if nsel > 128:
            assert 0 not in got and len(set(got)) == 128


def _eval_ref(px, part, pcount, pc):
    """pass006 (float64 where the reference's f32 rounding does not decide an
    integer): per-texel private avg/dir arrays -- a texel's thread holds its
    partition's mean and direction only when its index equals the partition
    index -- u32 truncation of each distance, NaN -> 0xFFFFFFFF, wrapping
    u32 sums, line parameters in 1/4096 steps."""
    tc = len(px)
    cnt = np.bincount(part, minlength=4)
    sums = np.zeros((4, 4))
    for t in range(tc):
        sums[part[t]] += np.floor(px[t].astype(np.float64))
    avg = {p: (sums[p] / cnt[p]).astype(np.float32).astype(np.float64) for p in range(pc) if cnt[p] > 0}
    axis = np.zeros((4, 4, 4), np.int64)
    for t in range(tc):
        p = part[t]
        own = p == t and t < pc and p in avg
        datum = px[t].astype(np.float64) - (avg[p] if own else 0.0)
        di = np.trunc(datum).astype(np.int64)
        for ax in range(4):
            if datum[ax] > 0:
                axis[p, ax] += di
    dirs = {}
    for p in range(pc):
        s = axis[p].astype(np.float64)
        best = s[0]
        for ax in range(1, 4):
            if s[ax] @ s[ax] > best @ best:
                best = s[ax]
        with np.errstate(invalid="ignore"):
            dirs[p] = best / np.linalg.norm(best)
    usum = ssum = 0
    lmin = {p: 2 ** 31 - 1 for p in range(pc)}
    lmax = {p: -(2 ** 31 - 1) for p in range(pc)}
    cw = CW.astype(np.float64)

    def u32(x):
        if np.isnan(x) or x >= 4294967040.0:
            return 0xFFFFFFFF
        return 0 if x < 0 else int(x)
    for t in range(tc):
        p = part[t]
        own = p == t and t < pc
        ud = dirs[t] if own else np.zeros(4)
        av = avg[t] if (own and t in avg) else np.zeros(4)
        x = px[t].astype(np.float64)
        param = x @ ud
        proj = av + ud * ((x - av) @ ud)
        udist = ((x - proj) ** 2) @ cw
        with np.errstate(invalid="ignore"):
            sd = av / np.linalg.norm(av)
        sdist = ((x - sd * (x @ sd)) ** 2) @ cw
        usum = (usum + u32(udist)) & 0xFFFFFFFF
        ssum = (ssum + u32(sdist)) & 0xFFFFFFFF
        pi = int(np.clip(np.trunc(param * 4096.0), -2 ** 31, 2 ** 31 - 1)) if not np.isnan(param) else 2 ** 31 - 1
        lmin[p] = min(lmin[p], pi)
        lmax[p] = max(lmax[p], pi)
    tu, ts = float(np.float32(usum)), float(np.float32(ssum))
    wie = 0.03 if tc <= 20 else (0.04 if tc <= 31 else (0.05 if tc <= 41 else 0.055))
    for p in range(pc):
        ln = max(lmax[p] / 4096.0 - lmin[p] / 4096.0, 1e-7)
        ew = pcount[p] * wie * wie
        tu += ew * ln * ln
        if p == 0 and 0 in avg:
            ts += ew * float(avg[0] @ avg[0])
    return tu, ts


@pytest.mark.parametrize("fp", [(4, 4), (6, 6)])
@pytest.mark.parametrize("pc", [2, 3, 4])
def test_partition_line_fit_errors_match_float64(fp, pc):
    bx, by = fp
    tc = bx * by
    rng = np.random.default_rng(240 + 10 * pc + bx)
    done = 0
    for e in rng.integers(0, 300, 60):
        pot, meta = O.partition_entry(bx, by, pc, int(e))
        if meta["partition_count"] != pc:
            continue
        px = block_px(rng, tc, alpha_const=bool(e % 2))
        eu, es = O.eval_candidate(bx, by, px, pc, int(e))
        ru, rs = _eval_ref(px, np.array(pot), meta["counts"], pc)
        assert abs(eu - ru) <= 2e-5 * abs(ru) + 64.0, (e, eu, ru)
        assert abs(es - rs) <= 2e-5 * abs(rs) + 64.0, (e, es, rs)
        done += 1
    assert done > 20


def test_positions_are_interleaved_stable_top4_without_duplicates():
    """pass007: the 4 best by the uncorrelated error and the 4 best by the
    same-chroma error (stable order, errors >= 1e30 never chosen),
    interleaved u0 s0 u1 s1 ..., first 4 distinct kept, -1 padding."""
    rng = np.random.default_rng(260)
    for it in range(300):
        eu = rng.choice([1.0, 2.0, 3.0, 4.0, 5.0, 1e30, 2e30], 128).astype(np.float32)
        es = rng.choice([1.0, 2.0, 3.0, 1e30], 128).astype(np.float32) if it % 3 else eu.copy()
        if it % 7 == 0:
            eu[:] = 1e30
        bu = [i for i in np.argsort(eu, kind="stable") if eu[i] < F32(1e30)][:4]
        bs = [i for i in np.argsort(es, kind="stable") if es[i] < F32(1e30)][:4]
        bu += [0] * (4 - len(bu))  # unfilled slots hold index 0 (BestChoice(0, 1e30))
        bs += [0] * (4 - len(bs))
        inter = [v for pair in zip(bu, bs) for v in pair]
        out = []
        for v in inter:
            if v not in out and len(out) < 4:
                out.append(int(v))
        out += [-1] * (4 - len(out))
        assert O.select_positions(eu, es) == out


# ---------------------------------------------------------------- exhaustive partition candidates
@pytest.mark.parametrize("fp", [(4, 4), (6, 6), (8, 8)])
@pytest.mark.parametrize("pc", [2, 3, 4])
def test_exhaustive_search_finds_the_partitioning_a_block_is_made_of(fp, pc):
    """A block painted from one selected partitioning (a distinct, well
    separated colour per partition): k-means recovers it, its mismatch is 0
    and both of its line fits are exact, so it is the first candidate."""
    bx, by = fp
    rng = np.random.default_rng(300 + pc + bx)
    cols = np.array([[10, 10, 10, 255], [240, 20, 20, 255], [20, 240, 20, 255], [20, 20, 240, 255]], np.uint8)
    nsel = O.footprint_info(bx, by)["selected_p%d" % pc]
    for e in rng.integers(0, nsel, 12):
        pot, meta = O.partition_entry(bx, by, pc, int(e))
        img = cols[np.array(pot)].reshape(by, bx, 4)
        got = O.exhaustive_candidates(img, bx, by, 0, pc)
        assert got[0] == e, (fp, pc, e, got)
        assert len(set(got)) == 4


@pytest.mark.parametrize("alpha", [False, True])
@pytest.mark.parametrize("pc", [2, 3, 4])
def test_exhaustive_search_with_noise_and_alpha(alpha, pc):
    """Blocks painted from one selected partitioning with per-partition
    colours (and alphas, which switches the line fits to four components)
    plus texel noise well below the colour separation: that partitioning
    stays the first candidate.  (Beyond such clear cases the ranking is
    decided by the reference's float32 line fits, whose cancellations make
    a float64 ranking differ.)"""
    bx = by = 6
    rng = np.random.default_rng(320 + pc + 10 * alpha)
    nsel = O.footprint_info(bx, by)["selected_p%d" % pc]
    for e in rng.integers(0, nsel, 10):
        pot, _ = O.partition_entry(bx, by, pc, int(e))
        cols = rng.permutation(np.array([[20, 20, 20], [230, 30, 30], [30, 230, 30], [30, 30, 230]]))[:pc]
        img = np.zeros((by * bx, 4))
        img[:, :3] = cols[np.array(pot)] + rng.normal(0, 2.0, (bx * by, 3))
        img[:, 3] = (np.array([255, 120, 60, 200])[np.array(pot)] if alpha else 255) + \
            (rng.normal(0, 2.0, bx * by) if alpha else 0)
        img = np.clip(np.rint(img), 0, 255).astype(np.uint8).reshape(by, bx, 4)
        got = O.exhaustive_candidates(img, bx, by, 0, pc)
        assert got[0] == e, (pc, e, got)


# ---------------------------------------------------------------- seeded candidates (learned policy)
def _seed_table(bx, by, pc):
    """seed -> packed entry of the selected partitionings, and the seeds of
    the entries that are not selected."""
    nsel = O.footprint_info(bx, by)["selected_p%d" % pc]
    sel = {O.partition_entry(bx, by, pc, e)[1]["seed"]: e for e in range(nsel)}
    unsel = [O.partition_entry(bx, by, pc, e)[1]["seed"] for e in range(nsel, 1024)]
    return sel, unsel


@pytest.mark.parametrize("fp", [(4, 4), (6, 6)])
@pytest.mark.parametrize("pc", [2, 3, 4])
def test_seeded_candidates_up_to_four_are_taken_as_listed(fp, pc):
    """Up to 4 seeds: the selected ones in list order, unselected seeds
    skipped, a negative seed ends the list, -1 padding."""
    bx, by = fp
    sel, unsel = _seed_table(bx, by, pc)
    rng = np.random.default_rng(400 + pc + bx)
    img = rng.integers(0, 256, (by, bx, 4), dtype=np.uint8)
    seeds_sel = [int(v) for v in rng.permutation(sorted(sel))[:4]]
    lists = [[seeds_sel[0], unsel[0], seeds_sel[1], seeds_sel[2]] if unsel else seeds_sel,
             [seeds_sel[3], -1, seeds_sel[0]], [unsel[-1]] if unsel else [-1], seeds_sel[:2]]
    for lst in lists:
        exp = []
        for sd in lst:
            if sd < 0:
                break
            if sd in sel:
                exp.append(sel[sd])
        exp += [-1] * (4 - len(exp))
        assert O.seeded_candidates(img, bx, by, 0, pc, lst) == exp, (lst, exp)


@pytest.mark.parametrize("fp", [(4, 4), (6, 6)])
@pytest.mark.parametrize("pc", [2, 3, 4])
def test_seeded_candidates_rank_the_painted_partitioning_first(fp, pc):
    """More than 4 seeds are ranked by the line fits: a block painted from
    one selected partitioning puts that partitioning first wherever its seed
    sits in a list of 12 (unselected seeds included), and every candidate
    is one of the listed selected partitionings."""
    bx, by = fp
    sel, unsel = _seed_table(bx, by, pc)
    rng = np.random.default_rng(420 + pc + bx)
    cols = np.array([[10, 10, 10, 255], [240, 20, 20, 255], [20, 240, 20, 255], [20, 20, 240, 255]], np.uint8)
    keys = sorted(sel)
    for it in range(8):
        lst = [int(v) for v in rng.permutation(keys)[:12]]
        if unsel:
            lst[int(rng.integers(0, 12))] = unsel[it % len(unsel)]
        target = lst[it % 12] if lst[it % 12] in sel else lst[(it + 1) % 12]
        pot, _ = O.partition_entry(bx, by, pc, sel[target])
        img = cols[np.array(pot)].reshape(by, bx, 4)
        got = O.seeded_candidates(img, bx, by, 0, pc, lst)
        assert got[0] == sel[target], (pc, lst, target, got)
        listed = {sel[v] for v in lst if v in sel}
        assert all(g in listed for g in got if g >= 0) and len({g for g in got if g >= 0}) == 4


@pytest.mark.parametrize("pc", [2, 3, 4])
def test_seeded_candidates_over_every_selected_seed_equal_the_exhaustive_search(pc):
    """With every selected partitioning's seed listed (in packed order), the
    line fits see the same partitionings as the exhaustive search (which
    visits them in mismatch order): on noisy blocks without tied errors both
    keep the same 4 candidates."""
    bx = by = 4
    sel, unsel = _seed_table(bx, by, pc)
    lst = sorted(sel, key=lambda sd: sel[sd]) + unsel[:5]
    rng = np.random.default_rng(440 + pc)
    for it in range(6):
        img = np.clip(rng.normal(128, 50, (by, bx, 4)), 0, 255).astype(np.uint8)
        if it % 2:
            img[..., 3] = 255
        assert O.seeded_candidates(img, bx, by, 0, pc, lst) == O.exhaustive_candidates(img, bx, by, 0, pc)


@pytest.mark.parametrize("pc", [2, 3, 4])
def test_seeded_candidates_after_the_mismatch_order(pc):
    """More than 4 seeds after the first N partitionings of the mismatch
    order: with N covering every selected partitioning the line fits visit
    exactly the exhaustive search's sequence (the seeds are all among them
    and skipped), so the candidates equal the exhaustive search's even where
    errors tie (a constant block); a block painted from one of the seeds'
    partitionings puts it first, fitted once."""
    bx = by = 4
    sel, unsel = _seed_table(bx, by, pc)
    rng = np.random.default_rng(460 + pc)
    nsel = len(sel)
    lst = [int(v) for v in rng.permutation(sorted(sel))[:12]]
    for it in range(5):
        img = np.clip(rng.normal(128, 50, (by, bx, 4)), 0, 255).astype(np.uint8)
        if it == 4:
            img[:] = 77  # every line fit ties
        assert O.seeded_candidates(img, bx, by, 0, pc, lst, nsel) == O.exhaustive_candidates(img, bx, by, 0, pc)
    cols = np.array([[10, 10, 10, 255], [240, 20, 20, 255], [20, 240, 20, 255], [20, 20, 240, 255]], np.uint8)
    pot, _ = O.partition_entry(bx, by, pc, sel[lst[5]])
    img = cols[np.array(pot)].reshape(by, bx, 4)
    got = O.seeded_candidates(img, bx, by, 0, pc, lst, 2)
    assert got[0] == sel[lst[5]], got
    assert len({g for g in got if g >= 0}) == 4
