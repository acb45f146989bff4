"""Candidate-predictor kernels (include/astcg_predictor.h) against a plain
PyTorch fp32 reference on the CPU that rounds to bfloat16 at the same points
(GEMM operands, hidden activations, backward deltas); tolerance 2e-2
relative to each tensor's scale, which covers fp32 summation order and
single-ulp bf16 flips of intermediate activations."""
import ctypes

import numpy as np
import pytest

import astc_implementation_amd as A
from astc_implementation_amd import synth
from astc_implementation_amd.predictor import Predictor, HEADS

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
TOL = 2e-2


def _close(gpu, ref, tol=TOL):
    gpu = gpu.detach().float().cpu()
    ref = ref.detach().float().cpu()
    scale = max(float(ref.abs().max()), 1e-6)
    err = float((gpu - ref).abs().max())
    assert err <= tol * scale, "max abs err %g vs scale %g" % (err, scale)


def _bf(t):
    return t.to(torch.bfloat16).float()


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("m,n,k", [(64, 64, 32), (37, 101, 70), (130, 3125, 256), (256, 48, 300)])
def test_gemm_matches_reference(ta, tb, m, n, k):
    g = torch.Generator().manual_seed(m * 7 + n * 3 + k)
    a = torch.randn((k, m) if ta else (m, k), generator=g)
    b = torch.randn((n, k) if tb else (k, n), generator=g)
    bias = torch.randn(n, generator=g)
    P = Predictor((4, 4))
    da, db = a.to(torch.bfloat16).cuda(), b.to(torch.bfloat16).cuda()
    c = torch.full((m, n), 3.0, device="cuda")
    P.gemm(m, n, k, da, a.shape[1], ta, db, b.shape[1], tb, c, n, bias.cuda())
    ref = (_bf(a).t() if ta else _bf(a)) @ (_bf(b).t() if tb else _bf(b)) + bias
    torch.cuda.synchronize()
    _close(c, ref, 1e-4 * k ** 0.5)
    c2 = c.clone()
    P.gemm(m, n, k, da, a.shape[1], ta, db, b.shape[1], tb, c2, n, None, accumulate=True)
    torch.cuda.synchronize()
    _close(c2, 2 * ref - bias, 2e-4 * k ** 0.5)


def _gelu(z):
    return 0.5 * z * (1.0 + torch.erf(z / 2 ** 0.5))


def _gelu_grad(z):
    return 0.5 * (1.0 + torch.erf(z / 2 ** 0.5)) + z * torch.exp(-0.5 * z * z) / (2 * np.pi) ** 0.5


def _reference(P, x, labels):
    W1, b1, W2, b2, W3, b3 = [w.detach().cpu() for w in P.params]
    B = x.shape[0]
    z1 = _bf(x) @ _bf(W1) + b1
    h1 = _bf(_gelu(z1))
    z2 = h1 @ _bf(W2) + b2
    h2 = _bf(_gelu(z2))
    logits = h2 @ _bf(W3) + b3
    dlog = torch.zeros_like(logits)
    loss = torch.zeros(B)
    for hi, (off, ln) in enumerate(zip(P.head_off, P.head_len)):
        lab = labels[:, hi]
        sl = logits[:, off:off + ln]
        lse = torch.logsumexp(sl, dim=1)
        p = torch.softmax(sl, dim=1)
        on = lab >= 0
        oh = torch.zeros_like(p)
        oh[on, lab[on].long()] = 1.0
        dlog[:, off:off + ln] = torch.where(on[:, None], (p - oh) / B, torch.zeros_like(p))
        loss += torch.where(on, lse - sl.gather(1, lab.clamp(min=0).long()[:, None])[:, 0], torch.zeros(B))
    d16 = _bf(dlog)
    gW3 = h2.t() @ d16
    gb3 = dlog.sum(0)
    dz2 = (d16 @ _bf(W3).t()) * _gelu_grad(z2)
    gW2 = h1.t() @ _bf(dz2)
    gb2 = dz2.sum(0)
    dz1 = (_bf(dz2) @ _bf(W2).t()) * _gelu_grad(z1)
    gW1 = _bf(x).t() @ _bf(dz1)
    gb1 = dz1.sum(0)
    return logits, loss, [gW1, gb1, gW2, gb2, gW3, gb3]


@pytest.mark.parametrize("block,B", [((4, 4), 37), ((6, 6), 64)])
def test_forward_backward_match_reference(block, B):
    P = Predictor(block, seed=3)
    g = torch.Generator().manual_seed(B)
    x = torch.rand((B, P.nin), generator=g)
    labels = torch.stack([torch.randint(0, ln, (B,), generator=g) for ln in P.head_len], 1).int()
    labels[::3, 1] = -1  # masked heads
    labels[1::4, 3] = -1
    logits = P.forward(x.to(torch.bfloat16).cuda())
    loss, grads = P.backward(logits, labels.cuda())
    torch.cuda.synchronize()
    r_logits, r_loss, r_grads = _reference(P, x, labels)
    _close(logits, r_logits)
    _close(loss, r_loss)
    for g_gpu, g_ref in zip(grads, r_grads):
        _close(g_gpu, g_ref)


class _RoundBf16(torch.autograd.Function):
    """bf16 rounding in the forward pass, identity in the backward pass (the
    kernels differentiate through their bf16 copies as if they were the fp32
    master weights and activations)."""

    @staticmethod
    def forward(ctx, x):
        return x.to(torch.bfloat16).float()

    @staticmethod
    def backward(ctx, g):
        return g


@pytest.mark.parametrize("block,B,with_rank", [((4, 4), 33, False), ((6, 6), 48, False), ((4, 4), 40, True)])
def test_backward_matches_autograd(block, B, with_rank):
    """The hand-written backward kernels against torch.autograd's gradients
    of the training objective: the mean over rows of the summed softmax
    cross-entropies of the labelled heads (plus, with ranking, of the
    softmax over each partition count's candidate seed logits), built in fp32
    on the CPU from the same bf16 operands as the forward kernels, with
    autograd deriving every gradient (GELU's included)."""
    R = _RoundBf16.apply
    P = Predictor(block, seed=11)
    g = torch.Generator().manual_seed(7 * B)
    x = torch.rand((B, P.nin), generator=g)
    labels = torch.stack([torch.randint(0, ln, (B,), generator=g) for ln in P.head_len], 1).int()
    labels[::5, 2] = -1
    labels[2::7, 4] = -1
    rank = None
    if with_rank:
        seeds = torch.randint(0, 1024, (B, 3, 4), generator=g)
        cols = (4 + torch.arange(3)[None, :, None] * 1024 + seeds).reshape(B, 12).int()
        cols[::6, 5] = -1  # an absent candidate
        tgt = torch.randint(0, 4, (B, 3), generator=g).int()
        tgt[1::5, 0] = -1  # a count without a winner
        tgt[::6, 1] = 0
        rank = (cols.cuda(), tgt.cuda())
    logits_gpu = P.forward(x.to(torch.bfloat16).cuda())
    loss_gpu, grads = P.backward(logits_gpu, labels.cuda(), rank)
    torch.cuda.synchronize()
    W = [w.detach().cpu().clone().requires_grad_(True) for w in P.params]
    W1, b1, W2, b2, W3, b3 = W
    gelu = torch.nn.functional.gelu
    h1 = R(gelu(R(x) @ R(W1) + b1))
    h2 = R(gelu(h1 @ R(W2) + b2))
    logits = h2 @ R(W3) + b3
    row = torch.zeros(B)
    for hi, (off, ln) in enumerate(zip(P.head_off, P.head_len)):
        lab = labels[:, hi].long()
        on = lab >= 0
        ce = torch.nn.functional.cross_entropy(logits[:, off:off + ln], lab.clamp(min=0), reduction="none")
        row = row + torch.where(on, ce, torch.zeros_like(ce))
    if with_rank:
        c = cols.reshape(B, 3, 4).long()
        t = tgt.long()
        for r in range(B):
            for q in range(3):
                if t[r, q] < 0:
                    continue
                keep = c[r, q] >= 0
                sel = logits[r, c[r, q][keep]]
                pos = int(keep[:t[r, q]].sum())
                row[r] = row[r] + torch.logsumexp(sel, 0) - sel[pos]
    (row.sum() / B).backward()
    _close(loss_gpu, row.detach())
    for g_gpu, w in zip(grads, W):
        _close(g_gpu, w.grad)


def test_adam_matches_reference():
    P = Predictor((4, 4), seed=5)
    g = torch.Generator().manual_seed(9)
    grads = [torch.randn(w.shape, generator=g) for w in P.params]
    w0 = [w.detach().cpu().clone() for w in P.params]
    for _ in range(3):
        P.adam([gg.cuda() for gg in grads], lr=1e-2, weight_decay=0.01)
    torch.cuda.synchronize()
    for w, w_init, gg in zip(P.params, w0, grads):
        ref = w_init.clone()
        m = torch.zeros_like(ref)
        v = torch.zeros_like(ref)
        for step in range(1, 4):
            m = 0.9 * m + 0.1 * gg
            v = 0.999 * v + 0.001 * gg * gg
            ref = ref - 1e-2 * ((m / (1 - 0.9 ** step)) / ((v / (1 - 0.999 ** step)).sqrt() + 1e-8) + 0.01 * ref)
        _close(w, ref, 1e-5)


def test_topk_and_features():
    P = Predictor((4, 4))
    g = torch.Generator().manual_seed(2)
    logits = torch.randn((50, P.nout), generator=g)
    logits[0, 10:20] = 5.0  # ties: lowest index first
    idx, prob = P.topk(logits.cuda(), 8)
    torch.cuda.synchronize()
    for r in range(50):
        for h, (off, ln) in enumerate(zip(P.head_off, P.head_len)):
            sl = logits[r, off:off + ln].numpy()
            order = np.lexsort((np.arange(ln), -sl))[:8]
            got = idx[r, h].cpu().numpy()
            # heads shorter than k pad with -1
            assert list(got[:len(order)]) == list(order) and all(got[len(order):] == -1)
            p = np.exp(sl - sl.max()) / np.exp(sl - sl.max()).sum()
            assert np.allclose(prob[r, h].cpu().numpy()[:len(order)], p[order], rtol=1e-4, atol=1e-6)
    img = synth.gradient_image(19, 13, 4)
    x = P.features(torch.from_numpy(img.reshape(-1).copy()).cuda(), 19, 13, 2, 7)
    torch.cuda.synchronize()
    for b in range(7):
        gb = 2 + b
        bx0, by0 = (gb % 5) * 4, (gb // 5) * 4
        ref = []
        for t in range(16):
            px = img[min(by0 + t // 4, 12), min(bx0 + t % 4, 18), :3].astype(np.float32) / 255.0
            ref.extend(px)
        _close(x[b].float(), torch.tensor(ref), 4e-3)


def test_search_labels_are_consistent():
    enc = A.Encoder(0)
    img = synth.gradient_image(64, 48, 12)
    enc.configure(64, 48, 4, 4, labels=True)
    enc.encode(img)
    lab = enc.labels()
    sym = enc.symbolic()
    raws = (ctypes.c_uint16 * 64)()
    n = ctypes.c_uint32()
    A.library().astcg_trial_modes(4, 4, ctypes.byref(n), raws, 64)
    for i in range(len(sym)):
        pc = sym[i].partition_count
        assert lab[i, 0] == pc
        assert raws[lab[i, 2]] == sym[i].block_mode_index
        if pc > 1:
            assert lab[i, 1] == sym[i].partition_index
            assert lab[i, 3 + pc - 2] == lab[i, 1]
        assert all(0 <= lab[i, 3 + j] < 1024 for j in range(3))
    # the 4 compat positions of each count: their best is the count's label
    seeds, errors = enc.candidates()
    assert ((seeds >= 0) & (seeds < 1024)).all() and (errors < 1e30).all()
    best = np.take_along_axis(seeds, np.argmin(errors, axis=2)[:, :, None], axis=2)[:, :, 0]
    assert np.array_equal(best, lab[:, 3:6])
    enc.close()


def test_rank_xent_matches_reference():
    """Softmax over candidate columns (absent ones, a repeated column, a row
    without target) added onto existing gradients and losses."""
    g = torch.Generator().manual_seed(5)
    B, NO = 37, 3129
    logits = torch.randn(B, NO, generator=g) * 2.0
    cols = torch.randint(0, NO, (B, 3, 4), generator=g, dtype=torch.int32)
    cols[0, 1, 2] = -1
    cols[1, 0, :] = cols[1, 0, 0]
    target = torch.randint(0, 4, (B, 3), generator=g, dtype=torch.int32)
    target[2, 0] = -1
    target[0, 1] = 3
    base = torch.randn(B, NO, generator=g) * 0.01
    scale = 0.25
    ref, ref_loss = base.clone(), torch.full((B,), 0.5)
    for r in range(B):
        for h in range(3):
            t = int(target[r, h])
            c = cols[r, h].long()
            if t < 0 or c[t] < 0:
                continue
            ok = c >= 0
            z = logits[r, c[ok]]
            p = torch.softmax(z, 0)
            oh = torch.zeros_like(p)
            oh[int(ok[:t].sum())] = 1.0
            for j, col in enumerate(c[ok].tolist()):
                ref[r, col] += (p[j] - oh[j]) * scale
            ref_loss[r] += -torch.log_softmax(z, 0)[int(ok[:t].sum())]
    P = Predictor((4, 4))
    d32 = base.cuda().contiguous()
    d16 = base.to(torch.bfloat16).cuda().contiguous()
    loss = torch.full((B,), 0.5).cuda()
    lg, cl, tg = logits.cuda().contiguous(), cols.reshape(B, 12).cuda().contiguous(), target.cuda().contiguous()
    A._check(P.lib.astcg_mlp_rank_xent(lg.data_ptr(), B, NO, 3, 4, cl.data_ptr(), tg.data_ptr(), scale,
                                       d16.data_ptr(), d32.data_ptr(), loss.data_ptr(), None))
    torch.cuda.synchronize()
    _close(d32, ref, 1e-5)
    _close(d16, ref, 1e-2)
    _close(loss, ref_loss, 1e-5)


def test_learned_policy_with_full_fallback_equals_plain():
    """Every block below the fallback threshold: the learned policy's plan
    (predictor inference, partition lists built in arbitrary order, mode
    masks, slot counts) must reproduce the plain search bit for bit."""
    P = Predictor((4, 4), seed=1)
    enc = A.Encoder(0)
    img = synth.gradient_image(120, 92, 31)
    enc.configure(120, 92, 4, 4)
    plain = enc.encode(img)
    enc.set_predictor(4, 4, [w.cpu().numpy() for w in P.params], fallback_below=2.0)
    enc.configure(120, 92, 4, 4, policy=A.POLICY_LEARNED, chunk_blocks=300)
    learned = enc.encode(img)
    assert enc.fallback_rate() == 1.0
    assert np.array_equal(learned, plain)
    out = enc.encode_batch([img, img[:40, :50]], [(4, 4), (4, 4)], policy=A.POLICY_LEARNED)
    assert np.array_equal(out[0], plain)
    enc.close()


def test_learned_policy_without_fallback_is_valid():
    """No fallback, the predictor's own candidates only: every block is a
    legal ASTC block (decodes without the error colour) of sane quality."""
    P = Predictor((4, 4), seed=2)
    enc = A.Encoder(0)
    img = synth.gradient_image(96, 64, 8)
    enc.set_predictor(4, 4, [w.cpu().numpy() for w in P.params], fallback_below=0.0, seeds=2, modes=16, slots=2)
    enc.configure(96, 64, 4, 4, policy=A.POLICY_LEARNED)
    blocks = enc.encode(img)
    assert enc.fallback_rate() == 0.0
    dec = enc.decode(4, 4, 96, 64, blocks)
    magenta = np.all(dec == np.array([255, 0, 255, 255], np.uint8), axis=-1)
    assert not magenta.any()
    assert A.psnr(dec, img) > 30.0
    enc.close()


@pytest.mark.parametrize("block", [(4, 4), (6, 6)])
def test_fused_inference_matches_unfused(block):
    """astcg_mlp_infer (one kernel: both hidden layers, the count and mode
    heads, top-k) against the separate GEMM / GELU / top-k kernels: same
    matrix instructions in the same K order, so the same hidden activations,
    indices and probabilities."""
    P = Predictor(block, seed=11)
    g = torch.Generator().manual_seed(12)
    for w in P.params:  # non-zero biases, wider logits
        w.add_(torch.randn(w.shape, generator=g).cuda() * 0.05)
    P.refresh_bf16()
    rows, kmax = 1000, 24
    kmode = min(20, P.nmodes)
    x = torch.rand((rows, P.nin), generator=g).to(torch.bfloat16).cuda()
    logits = P.forward(x)
    h2_ref = P.act[4]
    idx_ref, prob_ref = P.topk(logits, kmax)
    W1, b1, W2, b2, W3, b3 = P.params
    kin = (P.nin + 63) // 64 * 64
    n3 = 4 + P.nmodes
    n3p = (n3 + 15) // 16 * 16
    w1t = torch.zeros((256, kin), device="cuda")
    w1t[:, :P.nin] = W1.t()
    cols = list(range(4)) + [4 + 3072 + j for j in range(P.nmodes)]
    w3t = torch.zeros((n3p, 256), device="cuda")
    w3t[:n3] = W3[:, cols].t()
    b3c = torch.zeros(n3p, device="cuda")
    b3c[:n3] = b3[cols]
    w1t, w2t, w3t = (t.to(torch.bfloat16).contiguous() for t in (w1t, W2.t(), w3t))
    out_idx = torch.full((rows, 5, kmax), -7, dtype=torch.int32, device="cuda")
    out_prob = torch.zeros((rows, 5, kmax), device="cuda")
    h2 = torch.zeros((rows, 256), dtype=torch.bfloat16, device="cuda")
    A._check(P.lib.astcg_mlp_infer(rows, x.data_ptr(), P.nin, P.nin, w1t.data_ptr(), b1.data_ptr(), w2t.data_ptr(),
                                   b2.data_ptr(), w3t.data_ptr(), b3c.data_ptr(), n3, P.nmodes, kmax, 4, kmode,
                                   out_idx.data_ptr(), out_prob.data_ptr(), h2.data_ptr(), None))
    torch.cuda.synchronize()
    assert torch.equal(h2.cpu(), h2_ref.cpu())
    assert torch.equal(out_idx[:, 0, :4].cpu(), idx_ref[:, 0, :4].cpu())
    assert torch.equal(out_idx[:, 4, :kmode].cpu(), idx_ref[:, 4, :kmode].cpu())
    _close(out_prob[:, 0, :4], prob_ref[:, 0, :4], 1e-5)
    _close(out_prob[:, 4, :kmode], prob_ref[:, 4, :kmode], 1e-5)
    assert (out_idx[:, 1:4].cpu() == -7).all()  # the seed heads are not written


@pytest.mark.parametrize("k", [1, 8, 33, 48, 64, 100, 200])
def test_seed_topk_matches_unfused(k):
    """astcg_mlp_seed_topk (one partition count's seed-head logits of the
    listed rows from h2, then their top-k) against the GEMM over every
    output column + top-k of the seed head: same matrix instructions in the
    same K order, so the same indices (two identical seed columns: the
    lower first); unlisted rows and other heads are not written.  k up to
    192 takes the histogram-threshold selection, 200 the repeated arg-max."""
    P = Predictor((4, 4), seed=13)
    g = torch.Generator().manual_seed(14 + k)
    with torch.no_grad():
        for w in P.params:
            w.add_(torch.randn(w.shape, generator=g).cuda() * 0.05)
        for pc in (2, 3, 4):
            c0 = 4 + (pc - 2) * 1024
            P.params[4][:, c0 + 9] = P.params[4][:, c0 + 7]
            P.params[5][c0 + 9] = P.params[5][c0 + 7] + 1.0
            P.params[5][c0 + 7] += 1.0
    P.refresh_bf16()
    rows = 300
    x = torch.rand((rows, P.nin), generator=g).to(torch.bfloat16).cuda()
    logits = P.forward(x)
    h2 = P.act[4]
    idx_ref, _ = P.topk(logits, k)
    W3, b3 = P.params[4], P.params[5]
    listed = torch.tensor(sorted(torch.randperm(rows, generator=g)[:117].tolist()), dtype=torch.int32).cuda()
    count = torch.tensor([listed.numel()], dtype=torch.int32).cuda()
    sel = listed.long().cpu()
    mask = torch.ones(rows, dtype=torch.bool)
    mask[sel] = False
    for pc in (2, 3, 4):
        cols = slice(4 + (pc - 2) * 1024, 4 + (pc - 1) * 1024)
        w3t = W3[:, cols].t().contiguous().to(torch.bfloat16)
        b3s = b3[cols].contiguous()
        kmax = k + 3
        out = torch.full((rows, 5, kmax), -7, dtype=torch.int32, device="cuda")
        A._check(P.lib.astcg_mlp_seed_topk(listed.data_ptr(), count.data_ptr(), rows, h2.data_ptr(), w3t.data_ptr(),
                                           b3s.data_ptr(), k, out.data_ptr(), 5 * kmax, (pc - 1) * kmax, None))
        torch.cuda.synchronize()
        got = out.cpu()
        assert torch.equal(got[sel, pc - 1, :k], idx_ref[:, pc - 1, :k].cpu()[sel])
        assert (got[mask] == -7).all()
        others = [h for h in range(5) if h != pc - 1]
        assert (got[sel][:, others] == -7).all() and (got[sel, pc - 1, k:] == -7).all()


def test_fused_inference_argument_limits():
    P = Predictor((4, 4))
    lib = P.lib
    null = None
    # 8x8 has 339 trial modes: more output columns than the fused kernel takes
    assert lib.astcg_mlp_infer(1, 1, 192, 192, 1, 1, 1, 1, 1, 1, 4 + 339, 339, 16, 4, 16, 1, null, null, null) == 1
    assert lib.astcg_mlp_infer(1, 1, 300, 300, 1, 1, 1, 1, 1, 1, 57, 53, 16, 4, 16, 1, null, null, null) == 1
