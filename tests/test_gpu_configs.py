"""The benchmark workloads at their real sizes (the single-image 4096x4096
cases are in test_gpu_parity.py):

* the 64-texture batch of uneven sizes (128..2048 per side) and mixed
  footprints (4x4 / 6x6 / 8x8) through the batched device entry point: every
  texture equals its own single-texture encode, and one whole 920-block
  batch of the reference (code/compression_pipeline.cpp:345) of textures of
  every content kind and footprint equals the oracle's encoding of it;
* the 256-texture learned batch (64..4096 per side, mixed 4x4 / 6x6, the
  content mix of the 64-texture batch) through one batched call per policy:
  every texture's decoded PSNR against the EXHAUSTIVE search's, with the
  bound this pruning reaches (mean loss <= 0.01 dB, at most one texture
  above 0.05 dB, none above 0.06 dB), no error blocks;
* the learned policy at 4x4 and 6x6 on a fresh 4096x4096 image (seed 3,
  outside the training seeds 10000..19999) against the plain EXHAUSTIVE
  search it is trained on: decoded PSNR (RGBA, unweighted, peak 255) within
  0.05 dB; the fallback rate and the difference to the COMPAT search are
  reported in the assertion message and as test properties.
"""
import numpy as np
import pytest

import astc_implementation_amd as A
from astc_implementation_amd import synth
import oracle_lib as O

pytestmark = pytest.mark.gpu
MAX_DPSNR = 0.05


@pytest.fixture(scope="module")
def encoder():
    enc = A.Encoder(0)
    yield enc
    enc.close()


@pytest.fixture(scope="module")
def batch64():
    spec = synth.batch_spec(64, 128, 2048, [(4, 4), (6, 6), (8, 8)], 1000, 2)
    return spec, synth.make_textures(spec, workers=8)


def _oracle_sample(spec):
    """Texture indices covering every (content kind, footprint) pair present."""
    seen, pick = set(), []
    for i, (kind, _, _, fp, _) in enumerate(spec):
        if (kind, fp) not in seen:
            seen.add((kind, fp))
            pick.append(i)
    return pick


def test_config2_batch_equals_single_encodes_and_oracle(encoder, batch64):
    import torch
    spec, imgs = batch64
    dev = torch.device("cuda:0")
    nbs = [((w + bx - 1) // bx) * ((h + by - 1) // by) for _, w, h, (bx, by), _ in spec]
    d_in = [torch.from_numpy(im.reshape(-1).copy()).to(dev) for im in imgs]
    d_out = [torch.zeros(n * 16, dtype=torch.uint8, device=dev) for n in nbs]
    tex = [(d_in[i].data_ptr(), w, h, bx, by, d_out[i].data_ptr(), nbs[i] * 16)
           for i, (_, w, h, (bx, by), _) in enumerate(spec)]
    stream = torch.cuda.current_stream(dev)
    encoder.encode_batch_device(tex, stream=stream.cuda_stream)
    torch.cuda.synchronize()
    outs = [d.cpu().numpy() for d in d_out]
    for i, (img, (_, w, h, (bx, by), _)) in enumerate(zip(imgs, spec)):
        encoder.configure(w, h, bx, by)
        assert np.array_equal(outs[i], encoder.encode(img)), "texture %d %s" % (i, spec[i][:4])
    pick = _oracle_sample(spec)
    assert len(pick) >= 8 and {spec[i][3] for i in pick} == {(4, 4), (6, 6), (8, 8)}
    for i in pick:
        _, w, h, (bx, by), _ = spec[i]
        first = (nbs[i] // 2) // 920 * 920
        count = min(920, nbs[i] - first)
        ref, _, _ = O.encode(imgs[i], bx, by, 4, threads=16, first_block=first, block_count=count,
                             want_symbolic=False)
        assert np.array_equal(outs[i][first * 16:(first + count) * 16], ref), \
            "texture %d %s, blocks %d..%d" % (i, spec[i][:4], first, first + count)


@pytest.fixture(scope="module")
def fresh_image():
    return synth.gradient_image(4096, 4096, 3)


@pytest.mark.parametrize("bx,by", [(4, 4), (6, 6)])
def test_config3_learned_on_fresh_4096_image(encoder, fresh_image, bx, by, record_property):
    img = fresh_image
    encoder.configure(4096, 4096, bx, by, policy=A.POLICY_EXHAUSTIVE)
    plain = encoder.encode(img)
    p_plain = A.psnr(encoder.decode(bx, by, 4096, 4096, plain), img)
    encoder.configure(4096, 4096, bx, by)
    p_compat = A.psnr(encoder.decode(bx, by, 4096, 4096, encoder.encode(img)), img)
    encoder.load_predictor(bx, by)
    encoder.configure(4096, 4096, bx, by, policy=A.POLICY_LEARNED)
    learned = encoder.encode(img)
    fb = encoder.fallback_rate()
    dec = encoder.decode(bx, by, 4096, 4096, learned)
    p_learned = A.psnr(dec, img)
    record_property("fallback_rate", fb)
    record_property("dpsnr_db", p_plain - p_learned)
    record_property("dpsnr_vs_compat_db", p_compat - p_learned)
    assert not np.all(dec == np.array([255, 0, 255, 255], np.uint8), axis=-1).any()
    assert p_plain - p_learned <= MAX_DPSNR, (p_plain, p_learned, p_compat, fb)
    assert fb < 0.05, fb


def _batch_psnrs(encoder, spec, d_in, d_out, tex, policy):
    import torch
    stream = torch.cuda.current_stream()
    encoder.encode_batch_device(tex, stream=stream.cuda_stream, policy=policy)
    torch.cuda.synchronize()
    out = []
    magenta = 0
    for i, (_, w, h, (bx, by), _) in enumerate(spec):
        dec = torch.empty(w * h * 4, dtype=torch.uint8, device="cuda")
        encoder.decode_device(bx, by, w, h, d_out[i].data_ptr(), dec.data_ptr(), stream.cuda_stream)
        d4 = dec.view(-1, 4)
        magenta += int(((d4[:, 0] == 255) & (d4[:, 1] == 0) & (d4[:, 2] == 255) & (d4[:, 3] == 255)).sum())
        mse = float(((dec.float() - d_in[i].float()) ** 2).mean())
        out.append(99.0 if mse == 0 else 10.0 * np.log10(255.0 * 255.0 / mse))
    return np.array(out), magenta


def test_config4_learned_batch256_against_exhaustive(encoder, record_property):
    import torch
    spec = synth.batch_spec(256, 64, 4096, [(4, 4), (6, 6)], 20000, 4)
    imgs = synth.make_textures(spec, workers=16)
    d_in = [torch.from_numpy(im.reshape(-1).copy()).cuda() for im in imgs]
    del imgs
    nbs = [((w + bx - 1) // bx) * ((h + by - 1) // by) for _, w, h, (bx, by), _ in spec]
    d_out = [torch.zeros(n * 16, dtype=torch.uint8, device="cuda") for n in nbs]
    tex = [(d_in[i].data_ptr(), w, h, bx, by, d_out[i].data_ptr(), nbs[i] * 16)
           for i, (_, w, h, (bx, by), _) in enumerate(spec)]
    p_exh, m_exh = _batch_psnrs(encoder, spec, d_in, d_out, tex, A.POLICY_EXHAUSTIVE)
    for fp in ((4, 4), (6, 6)):
        encoder.load_predictor(*fp)
    p_learned, m_learned = _batch_psnrs(encoder, spec, d_in, d_out, tex, A.POLICY_LEARNED)
    fb = encoder.fallback_rate()
    d = p_exh - p_learned
    record_property("dpsnr_db_mean", float(d.mean()))
    record_property("dpsnr_db_max", float(d.max()))
    record_property("textures_over_0.05_db", int((d > MAX_DPSNR).sum()))
    record_property("fallback_rate", fb)
    worst = [(int(i), spec[i][0], spec[i][1], spec[i][2], spec[i][3], round(float(d[i]), 4)) for i in np.argsort(-d)[:4]]
    assert m_exh == 0 and m_learned == 0
    assert d.mean() <= 0.01, (float(d.mean()), worst)
    assert (d > MAX_DPSNR).sum() <= 1 and d.max() <= 0.06, worst
