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
