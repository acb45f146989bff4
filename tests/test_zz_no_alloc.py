"""astcg_encode_device_noalloc (the entry point that allocates nothing):
through the C ABI, two calls in a row (the second on a smaller image) give
exactly the blocks of astcg_encode_device for the same configuration; the
image is searched in chunks that fit the fixed scratch array, which the
larger image needs several of.  Settings the entry point cannot serve
without allocating are rejected."""
import ctypes

import numpy as np
import pytest

import astc_implementation_amd as A
from astc_implementation_amd import synth

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _encode(enc, img, bx, by, noalloc, **cfg):
    h, w = img.shape[:2]
    enc.configure(w, h, bx, by, **cfg)
    d_in = torch.from_numpy(img.reshape(-1).copy()).cuda()
    d_out = torch.full((enc.num_blocks * 16,), 0xAB, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    fn = A.library().astcg_encode_device_noalloc if noalloc else A.library().astcg_encode_device
    st = fn(enc._h, ctypes.c_void_p(d_in.data_ptr()), ctypes.c_void_p(d_out.data_ptr()),
            ctypes.c_void_p(stream.cuda_stream))
    assert st == A.OK, A.library().astcg_last_error(enc._h)
    torch.cuda.synchronize()
    return d_out.cpu().numpy()


@pytest.mark.parametrize("bx,by", [(4, 4), (6, 6)])
def test_noalloc_equals_encode_device_on_two_calls(bx, by):
    enc = A.Encoder(0)
    # 4x4: 20 000 blocks, several chunks of the 64 MiB scratch; then a small image
    big = synth.gradient_image(640, 500, 31)
    small = synth.alpha_image(70, 45, 32)
    for img in (big, small):
        ref = _encode(enc, img, bx, by, False)
        out = _encode(enc, img, bx, by, True)
        assert np.array_equal(out, ref), (img.shape, bx, by)
    enc.close()


def test_noalloc_explicit_small_chunks_and_void_extent():
    enc = A.Encoder(0)
    img = synth.flat_noise_image(130, 90, 5)
    img[:40, :40] = (10, 200, 30, 255)  # constant blocks for the void-extent path
    ref = _encode(enc, img, 4, 4, False, void_extent=True)
    out = _encode(enc, img, 4, 4, True, void_extent=True, chunk_blocks=97)
    assert np.array_equal(out, ref)
    enc.close()


def test_noalloc_rejects_what_it_cannot_serve():
    lib = A.library()
    enc = A.Encoder(0)
    d = torch.zeros(64 * 64 * 4, dtype=torch.uint8, device="cuda")
    o = torch.zeros(256 * 16, dtype=torch.uint8, device="cuda")
    args = (enc._h, ctypes.c_void_p(d.data_ptr()), ctypes.c_void_p(o.data_ptr()), None)
    assert lib.astcg_encode_device_noalloc(*args) == 4  # not configured
    enc.configure(64, 64, 4, 4, chunk_blocks=1 << 20)
    assert lib.astcg_encode_device_noalloc(*args) == 1  # chunk beyond the scratch
    assert b"static scratch" in lib.astcg_last_error(enc._h)
    enc.configure(64, 64, 4, 4, labels=True)
    assert lib.astcg_encode_device_noalloc(*args) == 8
    enc.load_predictor(4, 4)
    enc.configure(64, 64, 4, 4, policy=A.POLICY_LEARNED)
    assert lib.astcg_encode_device_noalloc(*args) == 8
    enc.close()
