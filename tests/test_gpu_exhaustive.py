"""The EXHAUSTIVE partition policy (the CPU partition search of
code/best_partitionings.cpp:498-737, which the reference never calls, run
over every selected partitioning) on the GPU: bit-exact with the oracle's
restatement of it on the same images, also through the batched and the
no-allocation entry points; it finds better partitionings than the live
passes 001-007, so the decoded PSNR is at least the COMPAT search's."""
import ctypes

import numpy as np
import pytest

import astc_implementation_amd as A
from astc_implementation_amd import synth
import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def encoder():
    enc = A.Encoder(0)
    yield enc
    enc.close()


@pytest.mark.parametrize("bx,by", [(4, 4), (5, 4), (6, 6), (8, 8), (10, 10), (12, 12)])
@pytest.mark.parametrize("kind", ["gradient", "alpha"])
def test_exhaustive_bit_exact(encoder, bx, by, kind):
    img = (synth.gradient_image if kind == "gradient" else synth.alpha_image)(96, 80, 13)
    encoder.configure(96, 80, bx, by, policy=A.POLICY_EXHAUSTIVE)
    gpu = encoder.encode(img)
    ref, _, _ = O.encode(img, bx, by, 4, threads=16, want_symbolic=False, policy=O.POLICY_EXHAUSTIVE)
    bad = np.nonzero(np.any(gpu.reshape(-1, 16) != ref.reshape(-1, 16), axis=1))[0]
    assert bad.size == 0, "blocks differ: %s" % bad[:10].tolist()
    encoder.configure(96, 80, bx, by)
    compat = encoder.encode(img)
    assert A.psnr(A.decode_image(bx, by, 96, 80, gpu), img) >= A.psnr(A.decode_image(bx, by, 96, 80, compat), img)


def test_exhaustive_512_config0_image(encoder):
    img = synth.gradient_image(512, 512, 0)
    encoder.configure(512, 512, 4, 4, policy=A.POLICY_EXHAUSTIVE)
    gpu = encoder.encode(img)
    ref, _, _ = O.encode(img, 4, 4, 4, threads=16, want_symbolic=False, policy=O.POLICY_EXHAUSTIVE)
    assert np.array_equal(gpu, ref)


def test_exhaustive_batch_and_noalloc(encoder):
    import torch
    imgs = [synth.gradient_image(70, 50, 3), synth.normal_map_image(45, 66, 4), synth.flat_noise_image(33, 29, 5)]
    fps = [(4, 4), (6, 6), (8, 8)]
    outs = encoder.encode_batch(imgs, fps, policy=A.POLICY_EXHAUSTIVE)
    for img, (bx, by), out in zip(imgs, fps, outs):
        h, w = img.shape[:2]
        ref, _, _ = O.encode(img, bx, by, 4, threads=16, want_symbolic=False, policy=O.POLICY_EXHAUSTIVE)
        assert np.array_equal(out, ref)
        encoder.configure(w, h, bx, by, policy=A.POLICY_EXHAUSTIVE, chunk_blocks=37)
        d_in = torch.from_numpy(img.reshape(-1).copy()).cuda()
        d_out = torch.zeros(ref.size, dtype=torch.uint8, device="cuda")
        st = A.library().astcg_encode_device_noalloc(encoder._h, ctypes.c_void_p(d_in.data_ptr()),
                                                    ctypes.c_void_p(d_out.data_ptr()), None)
        assert st == A.OK
        torch.cuda.synchronize()
        assert np.array_equal(d_out.cpu().numpy(), ref)


@pytest.mark.parametrize("w,h,chunk", [(160, 120, 500), (1088, 1024, 0)])
def test_learned_policy_on_exhaustive_candidates(encoder, w, h, chunk):
    """LEARNED with candidates = 1 takes its partition candidates from the
    exhaustive search: with every block falling back it reproduces the
    EXHAUSTIVE encode bit for bit (the compacted block list of the
    exhaustive kernel covers exactly the blocks the plan searches; 1200
    blocks run a wave per block, 69632 blocks, above the 65536-block cut,
    a lane per block)."""
    img = synth.gradient_image(w, h, 21)
    encoder.configure(w, h, 4, 4, policy=A.POLICY_EXHAUSTIVE)
    ref = encoder.encode(img)
    encoder.load_predictor(4, 4, fallback_below=2.0, candidates=1, early_out=0.0, count_below=0.0,
                           full_search_above=0.0)
    encoder.configure(w, h, 4, 4, policy=A.POLICY_LEARNED, chunk_blocks=chunk)
    assert np.array_equal(encoder.encode(img), ref)
    assert encoder.fallback_rate() == 1.0
    encoder.load_predictor(4, 4, candidates=1)
    encoder.configure(w, h, 4, 4, policy=A.POLICY_LEARNED)
    out = encoder.encode(img)
    dec = A.decode_image(4, 4, w, h, out)
    assert not np.all(dec == np.array([255, 0, 255, 255], np.uint8), axis=-1).any()
    assert A.psnr(A.decode_image(4, 4, w, h, ref), img) - A.psnr(dec, img) < 0.5


def test_learned_exhaustive_search_limit_matches_oracle(encoder):
    """candidates = N >= 2: the exhaustive search line-fits only the first N
    partitionings of the k-means mismatch order (the reference's
    partition_search_limit); with every block falling back and no pruning
    the blocks equal the oracle's EXHAUSTIVE encode with that limit."""
    img = synth.alpha_image(96, 72, 17)
    ref, _, _ = O.encode(img, 4, 4, 4, threads=16, want_symbolic=False, policy=O.POLICY_EXHAUSTIVE,
                         search_limit=100)
    encoder.load_predictor(4, 4, fallback_below=2.0, candidates=100, early_out=0.0, count_below=0.0,
                           full_search_above=0.0)
    encoder.configure(96, 72, 4, 4, policy=A.POLICY_LEARNED)
    assert np.array_equal(encoder.encode(img), ref)
