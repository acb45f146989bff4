"""Launch-shape tuning (astcg_tuning, astcg_autotune): every setting leaves
the encoded blocks byte-identical (texel staging of the refinement, the
refinement's persistent work queue and its refill threshold against the
workgroup-per-tile kernel, float staging of the exhaustive partition search,
the wave- / lane-per-block choice for exhaustive block lists), and the
autotuner's on-disk cache is written, read back by another encoder (also
through ASTCG_TUNE_CACHE), and gives the same blocks."""
import os
import subprocess
import sys

import numpy as np
import pytest

import astc_implementation_amd as A
from astc_implementation_amd import synth

pytestmark = pytest.mark.gpu


def _encode(enc, img, bx, by, policy, **lp):
    h, w = img.shape[:2]
    if policy == A.POLICY_LEARNED:
        enc.load_predictor(bx, by, **lp)
    enc.configure(w, h, bx, by, policy=policy)
    return enc.encode(img)


@pytest.mark.parametrize("bx,by", [(4, 4), (5, 5), (6, 6), (8, 8)])
def test_every_tuning_gives_the_same_blocks(bx, by):
    img = synth.alpha_image(200, 136, 31)
    enc = A.Encoder(0)
    try:
        base = {p: _encode(enc, img, bx, by, p) for p in (A.POLICY_COMPAT, A.POLICY_EXHAUSTIVE)}
        default = enc.tuning(bx, by)
        assert default == {"refine_stage_max_tc": 25, "ex_f4_max_tc": 16,
                           "exw_max_list": 1 << 20 if bx * by >= 36 else 65536, "refine_queue": 0,
                           "refine_classes": int(bx * by <= 25)}
        for t in ({"refine_stage_max_tc": 0}, {"refine_stage_max_tc": 144}, {"ex_f4_max_tc": 0},
                  {"ex_f4_max_tc": 64}, {"refine_queue": 16}, {"refine_queue": 1}, {"refine_queue": 64},
                  {"refine_queue": 16, "refine_stage_max_tc": 0}, {"refine_queue": 7, "refine_stage_max_tc": 0},
                  {"refine_classes": 1 - int(bx * by <= 25)}, {"refine_classes": 0, "refine_queue": 16}):
            enc.set_tuning(bx, by, **default)
            enc.set_tuning(bx, by, **t)
            for p, ref in base.items():
                assert np.array_equal(_encode(enc, img, bx, by, p), ref), (t, p)
        enc.set_tuning(bx, by, **default)
    finally:
        enc.close()


def test_exhaustive_list_kernels_agree():
    """The learned policy's fallback blocks on the exhaustive search through
    a wave per block (list length up to exw_max_list) or a lane per block."""
    img = synth.gradient_image(320, 256, 33)
    enc = A.Encoder(0)
    try:
        lp = dict(fallback_below=2.0, early_out=0.0, count_below=0.0, full_search_above=0.0)
        enc.set_tuning(4, 4, exw_max_list=1 << 20)
        wave = _encode(enc, img, 4, 4, A.POLICY_LEARNED, **lp)
        enc.set_tuning(4, 4, exw_max_list=0)
        lane = _encode(enc, img, 4, 4, A.POLICY_LEARNED, **lp)
        enc.set_tuning(4, 4, exw_max_list=65536)
        assert np.array_equal(wave, lane)
        assert np.array_equal(wave, _encode(enc, img, 4, 4, A.POLICY_EXHAUSTIVE))
    finally:
        enc.close()


def test_seeded_list_kernels_agree():
    """The learned policy's seeded candidates (48 seeds ranked by line fits)
    through a wave per listed block (lists up to exw_max_list) or a lane per
    block give the same blocks."""
    img = synth.gradient_image(320, 256, 36)
    enc = A.Encoder(0)
    try:
        for bx, by in ((4, 4), (6, 6)):
            default = enc.tuning(bx, by)["exw_max_list"]
            enc.set_tuning(bx, by, exw_max_list=1 << 20)
            wave = _encode(enc, img, bx, by, A.POLICY_LEARNED, seeds=48)
            enc.set_tuning(bx, by, exw_max_list=0)
            lane = _encode(enc, img, bx, by, A.POLICY_LEARNED, seeds=48)
            enc.set_tuning(bx, by, exw_max_list=default)
            assert np.array_equal(wave, lane), (bx, by)
            assert np.array_equal(wave, _encode(enc, img, bx, by, A.POLICY_LEARNED, seeds=48)), (bx, by)
    finally:
        enc.close()


def test_autotune_cache_round_trip(tmp_path):
    cache = str(tmp_path / "tune.txt")
    img = synth.gradient_image(160, 128, 35)
    enc = A.Encoder(0)
    try:
        enc.configure(160, 128, 6, 6)
        ref = enc.encode(img)
        tuned = enc.autotune(6, 6, cache)
        assert enc.tuning(6, 6) == tuned
        # the encoder keeps its configuration
        assert np.array_equal(enc.encode(img), ref)
    finally:
        enc.close()
    lines = open(cache).read().splitlines()
    assert len(lines) == 1 and lines[0].startswith("astcg-tune 2 ") and " 6x6 " in lines[0]
    enc2 = A.Encoder(0)
    try:
        assert enc2.autotune(6, 6, cache) == tuned  # read back, not timed again
        assert len(open(cache).read().splitlines()) == 1
        enc2.configure(160, 128, 6, 6)
        assert np.array_equal(enc2.encode(img), ref)
    finally:
        enc2.close()
    # a fresh process picks the entry up through the environment
    code = ("import astc_implementation_amd as A; e = A.Encoder(0); print(sorted(e.tuning(6, 6).items())); e.close()")
    env = dict(os.environ, ASTCG_TUNE_CACHE=cache)
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120,
                         cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip().splitlines()[-1] == str(sorted(tuned.items()))


def test_learned_and_batch_under_both_refinement_kernels():
    """The learned policy (compacted partitioned-block lists) and a batch of
    uneven textures of two footprints give the same blocks through the
    persistent refinement and the workgroup-per-tile kernel, with and without
    the format classes."""
    img = synth.gradient_image(264, 200, 37)
    texs = [synth.gradient_image(w, h, 40 + i) for i, (w, h) in enumerate([(70, 45), (33, 90), (128, 17)])]
    foot = [(4, 4), (6, 6), (4, 4)]
    enc = A.Encoder(0)
    try:
        outs = []
        for q, c in ((0, 1), (16, 1), (3, 0), (0, 0)):
            for bx, by in ((4, 4), (6, 6)):
                enc.set_tuning(bx, by, refine_queue=q, refine_classes=c)
            learned = _encode(enc, img, 4, 4, A.POLICY_LEARNED)
            batch = enc.encode_batch(texs, foot)
            outs.append((learned, batch))
        for learned, batch in outs[1:]:
            assert np.array_equal(learned, outs[0][0])
            assert all(np.array_equal(a, b) for a, b in zip(batch, outs[0][1]))
    finally:
        enc.close()
