"""ASTCG_FLAG_GRAPH: the launch chain of an encode is captured into a HIP
graph on the first call and replayed on later calls with the same
arguments.  Every replay must equal the eager encode, with the program's own
eager GPU work and synchronisations between replays; new input contents in
the same buffers are encoded by the replay, and new buffers get a new
capture."""
import numpy as np
import pytest

import astc_implementation_amd as A
from astc_implementation_amd import synth

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _eager_work():
    x = torch.randn(1 << 20, device="cuda")
    y = (x * 3.0).sum()
    torch.cuda.synchronize()
    return float(y)


@pytest.fixture(scope="module")
def encoder():
    enc = A.Encoder(0)
    yield enc
    enc.close()


def test_graph_replays_single_image(encoder):
    w, h = 136, 100
    imgs = [synth.gradient_image(w, h, s) for s in (21, 22)]
    refs = []
    for img in imgs:
        encoder.configure(w, h, 4, 4, chunk_blocks=300)
        refs.append(encoder.encode(img))
    encoder.configure(w, h, 4, 4, chunk_blocks=300, graph=True)
    stream = torch.cuda.Stream()
    d_in = torch.from_numpy(imgs[0].reshape(-1).copy()).cuda()
    d_out = torch.zeros(encoder.num_blocks * 16, dtype=torch.uint8, device="cuda")
    for i in range(4):
        d_out.zero_()
        torch.cuda.synchronize()
        encoder.encode_device(d_in.data_ptr(), d_out.data_ptr(), stream.cuda_stream)
        stream.synchronize()
        assert np.array_equal(d_out.cpu().numpy(), refs[0]), "replay %d" % i
        _eager_work()
    # same buffers, new contents: the replay reads them
    d_in.copy_(torch.from_numpy(imgs[1].reshape(-1).copy()))
    torch.cuda.synchronize()
    encoder.encode_device(d_in.data_ptr(), d_out.data_ptr(), stream.cuda_stream)
    stream.synchronize()
    assert np.array_equal(d_out.cpu().numpy(), refs[1])
    # new output buffer: a new capture
    d_out2 = torch.zeros_like(d_out)
    for _ in range(2):
        encoder.encode_device(d_in.data_ptr(), d_out2.data_ptr(), stream.cuda_stream)
        stream.synchronize()
        assert np.array_equal(d_out2.cpu().numpy(), refs[1])
        _eager_work()


def test_graph_replays_learned_policy(encoder):
    w, h = 120, 96
    img = synth.gradient_image(w, h, 23)
    encoder.load_predictor(4, 4)
    encoder.configure(w, h, 4, 4, policy=A.POLICY_LEARNED)
    ref = encoder.encode(img)
    fb = encoder.fallback_rate()
    encoder.configure(w, h, 4, 4, policy=A.POLICY_LEARNED, graph=True)
    for i in range(3):
        assert np.array_equal(encoder.encode(img), ref), "replay %d" % i
        assert encoder.fallback_rate() == fb
        _eager_work()


def test_graph_replays_mixed_batch(encoder):
    imgs = [synth.gradient_image(70, 50, 31), synth.normal_map_image(45, 61, 32), synth.flat_noise_image(33, 29, 33),
            synth.gradient_image(90, 40, 34)]
    fps = [(4, 4), (6, 6), (8, 8), (4, 4)]
    refs = encoder.encode_batch(imgs, fps)
    stream = torch.cuda.Stream()
    d_in = [torch.from_numpy(im.reshape(-1).copy()).cuda() for im in imgs]
    d_out = [torch.zeros(len(r), dtype=torch.uint8, device="cuda") for r in refs]
    tex = [(d_in[i].data_ptr(), im.shape[1], im.shape[0], bx, by, d_out[i].data_ptr(), len(refs[i]))
           for i, (im, (bx, by)) in enumerate(zip(imgs, fps))]
    for i in range(3):
        for o in d_out:
            o.zero_()
        torch.cuda.synchronize()
        encoder.encode_batch_device(tex, stream=stream.cuda_stream, graph=True)
        stream.synchronize()
        for j, (o, r) in enumerate(zip(d_out, refs)):
            assert np.array_equal(o.cpu().numpy(), r), "replay %d texture %d" % (i, j)
        _eager_work()


def test_new_predictor_drops_captured_chains(encoder):
    """Installing another predictor invalidates captured chains: the next
    encode runs (and captures) the new network, matching its eager encode."""
    w, h = 96, 80
    img = synth.gradient_image(w, h, 24)
    params, hidden, prune = A.load_predictor_file(A.predictor_path(4, 4))
    zero = [np.zeros_like(p) for p in params]
    outs = {}
    for graph in (False, True):
        for name, ps in (("trained", params), ("zero", zero)):
            encoder.set_predictor(4, 4, ps, hidden=hidden, **prune)
            encoder.configure(w, h, 4, 4, policy=A.POLICY_LEARNED, graph=graph)
            outs[(graph, name)] = [encoder.encode(img) for _ in range(2)]
    for name in ("trained", "zero"):
        for o in outs[(True, name)]:
            assert np.array_equal(o, outs[(False, name)][0])
    assert not np.array_equal(outs[(False, "trained")][0], outs[(False, "zero")][0])
