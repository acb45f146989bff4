"""GPU kernels vs the CPU oracle on the same seeded images: the 128-bit
blocks must be identical (the kernels evaluate every reduction in the same
fixed order as the oracle)."""
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


def _compare(enc, img, bx, by, max_partitions):
    h, w = img.shape[:2]
    enc.configure(w, h, bx, by, max_partitions=max_partitions)
    gpu = enc.encode(img)
    ref, sym, _ = O.encode(img, bx, by, max_partitions, threads=8)
    g = gpu.reshape(-1, 16)
    r = ref.reshape(-1, 16)
    same = np.all(g == r, axis=1)
    return same, g, r, sym


@pytest.mark.parametrize("bx,by", [(4, 4), (6, 6), (8, 8)])
def test_one_partition_bit_exact(encoder, bx, by):
    img = synth.gradient_image(96, 80, 7)
    same, g, r, sym = _compare(encoder, img, bx, by, 1)
    bad = np.nonzero(~same)[0]
    assert bad.size == 0, "blocks differ: %s" % bad[:10].tolist()


@pytest.mark.parametrize("bx,by", [(4, 4), (6, 6), (8, 8)])
def test_all_partitions_bit_exact(encoder, bx, by):
    img = synth.gradient_image(96, 80, 11)
    same, g, r, sym = _compare(encoder, img, bx, by, 4)
    bad = np.nonzero(~same)[0]
    assert bad.size == 0, "blocks differ: %s" % bad[:10].tolist()
