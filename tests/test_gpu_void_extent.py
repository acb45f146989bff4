"""ASTCG_FLAG_VOID_EXTENT: blocks whose texels (edge-clamped) are all equal
are written as LDR void-extent blocks; every other block is the search's
block, unchanged.  The reference never emits void-extent blocks
(code/physical_compression.cpp:849), so this is checked against the plain
encode and the ASTC specification's block layout, decoded by the host
decoder."""
import numpy as np
import pytest

import astc_implementation_amd as A
from astc_implementation_amd import synth

pytestmark = pytest.mark.gpu


def constant_blocks(img, bx, by):
    """Mask of blocks whose edge-clamped texels are all equal, and the colour
    of each block's first texel."""
    h, w = img.shape[:2]
    bxn, byn = (w + bx - 1) // bx, (h + by - 1) // by
    ys = np.minimum(np.arange(byn)[:, None] * by + np.arange(by)[None, :], h - 1)
    xs = np.minimum(np.arange(bxn)[:, None] * bx + np.arange(bx)[None, :], w - 1)
    blk = img[ys[:, None, :, None], xs[None, :, None, :]].reshape(bxn * byn, bx * by, 4)
    return (blk == blk[:, :1]).all(axis=(1, 2)), blk[:, 0]


def void_extent_block(rgba):
    """LDR void extent of one colour, byte for byte from the specification:
    mode 0x1FC, LDR, reserved bits 1, all-ones extent, UNORM16 colour."""
    out = np.zeros(16, np.uint8)
    out[:8] = [0xFC, 0xFD, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF]
    out[8:] = np.array([int(c) * 257 for c in rgba], "<u2").view(np.uint8)
    return out


@pytest.fixture(scope="module")
def encoder():
    enc = A.Encoder(0)
    yield enc
    enc.close()


@pytest.mark.parametrize("bx,by", [(4, 4), (6, 6), (8, 8)])
def test_void_extent_blocks(encoder, bx, by):
    # the gradient generator's rectangles are solid colour: many blocks fall
    # entirely inside one; ragged sizes exercise edge clamping
    img = synth.alpha_image(203, 157, 12)
    img[100:157, 150:203] = (10, 200, 30, 128)  # a constant corner, clamped edge blocks included
    encoder.configure(203, 157, bx, by)
    plain = encoder.encode(img).reshape(-1, 16)
    encoder.configure(203, 157, bx, by, void_extent=True)
    ve = encoder.encode(img).reshape(-1, 16)
    const, colour = constant_blocks(img, bx, by)
    assert const.sum() >= 20
    assert np.array_equal(ve[~const], plain[~const])
    for i in np.nonzero(const)[0]:
        assert np.array_equal(ve[i], void_extent_block(colour[i])), i
    dec = A.decode_image(bx, by, 203, 157, ve.reshape(-1))
    assert A.psnr(dec, img) >= A.psnr(A.decode_image(bx, by, 203, 157, plain.reshape(-1)), img)
    assert np.array_equal(encoder.decode(bx, by, 203, 157, ve.reshape(-1)), dec)


def test_void_extent_in_batches(encoder):
    imgs = [synth.gradient_image(64, 48, 5), synth.flat_noise_image(40, 40, 6), np.full((30, 30, 4), 77, np.uint8)]
    fps = [(4, 4), (6, 6), (5, 5)]
    outs = encoder.encode_batch(imgs, fps, void_extent=True)
    for img, (bx, by), out in zip(imgs, fps, outs):
        h, w = img.shape[:2]
        encoder.configure(w, h, bx, by, void_extent=True)
        assert np.array_equal(out, encoder.encode(img))
    assert np.array_equal(A.decode_image(5, 5, 30, 30, outs[2]), imgs[2])
