"""ctypes binding of the CPU oracle (oracle/liboracle_astc.so).

Test infrastructure: used by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg only.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle_astc.so")


class OrSymbolic(ctypes.Structure):
    _fields_ = [
        ("errorval", ctypes.c_float),
        ("block_mode_index", ctypes.c_uint32),
        ("partition_count", ctypes.c_uint32),
        ("partition_index", ctypes.c_uint32),
        ("partition_formats_matched", ctypes.c_uint32),
        ("quant_mode", ctypes.c_uint32),
        ("_padding1", ctypes.c_uint32),
        ("_padding2", ctypes.c_uint32),
        ("partition_formats", ctypes.c_uint32 * 4),
        ("packed_color_values", ctypes.c_uint32 * 32),
        ("quantized_weights", ctypes.c_uint32 * 64),
    ]


class OrParams(ctypes.Structure):
    _fields_ = [("max_partitions", ctypes.c_int), ("batch_size", ctypes.c_int), ("threads", ctypes.c_int),
                ("first_block", ctypes.c_int), ("block_count", ctypes.c_int), ("policy", ctypes.c_int),
                ("search_limit", ctypes.c_int), ("seed_list", ctypes.c_void_p), ("seed_list_len", ctypes.c_int),
                ("mismatch_candidates", ctypes.c_int)]

POLICY_COMPAT = 0
POLICY_EXHAUSTIVE = 2
POLICY_SEEDED = 3


class OrStats(ctypes.Structure):
    _fields_ = [("blocks", ctypes.c_uint64), ("partitioned_blocks", ctypes.c_uint64),
                ("short_candidate_lists", ctypes.c_uint64)]


_lib = None


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            build_oracle()
        L = ctypes.CDLL(ORACLE_LIB)
        vp, i32 = ctypes.c_void_p, ctypes.c_int
        L.oracle_encode.argtypes = [vp, i32, i32, i32, i32, vp, vp, vp, vp]
        L.oracle_encode.restype = i32
        L.oracle_symbolic_to_physical.argtypes = [i32, i32, vp, vp]
        L.oracle_symbolic_to_physical.restype = i32
        L.oracle_encode_ise.argtypes = [i32, i32, vp, vp, i32]
        L.oracle_encode_ise.restype = None
        L.oracle_ise_bitcount.argtypes = [i32, i32]
        L.oracle_ise_bitcount.restype = i32
        L.oracle_table.argtypes = [i32, vp, i32]
        L.oracle_table.restype = i32
        L.oracle_footprint_info.argtypes = [i32, i32, vp, i32]
        L.oracle_footprint_info.restype = i32
        L.oracle_partition_entry.argtypes = [i32, i32, i32, i32, vp, vp]
        L.oracle_partition_entry.restype = i32
        L.oracle_debug_kmeans_candidates.argtypes = [vp, i32, i32, i32, i32, i32, i32, i32, vp]
        L.oracle_debug_kmeans_candidates.restype = i32
        L.oracle_debug_exhaustive_candidates.argtypes = [vp, i32, i32, i32, i32, i32, i32, vp]
        L.oracle_debug_exhaustive_candidates.restype = i32
        L.oracle_debug_seeded_candidates.argtypes = [vp, i32, i32, i32, i32, i32, i32, vp, i32, i32, vp]
        L.oracle_debug_seeded_candidates.restype = i32
        L.oracle_decimation_entry.argtypes = [i32, i32, i32, vp, i32]
        L.oracle_decimation_entry.restype = i32
        _lib = L
    return _lib


def encode(rgba, bx, by, max_partitions=4, threads=1, first_block=0, block_count=0, batch_size=920,
           want_symbolic=True, policy=POLICY_COMPAT, search_limit=0, seed_lists=None, mismatch_candidates=0):
    """Returns (blocks uint8[n*16], symbolic list or None, stats).  seed_lists
    (POLICY_SEEDED): (3, n) partition seeds for 2, 3, 4 partitions, best
    first, the same for every block."""
    rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
    h, w = rgba.shape[:2]
    nb_total = ((w + bx - 1) // bx) * ((h + by - 1) // by)
    n = block_count if block_count else nb_total - first_block
    out = np.zeros(n * 16, np.uint8)
    sym = (OrSymbolic * n)() if want_symbolic else None
    st = OrStats()
    seeds = None if seed_lists is None else np.ascontiguousarray(seed_lists, dtype=np.int32)
    p = OrParams(max_partitions, batch_size, threads, first_block, block_count, policy, search_limit,
                 seeds.ctypes.data if seeds is not None else None, seeds.shape[1] if seeds is not None else 0,
                 mismatch_candidates)
    r = lib().oracle_encode(rgba.ctypes.data, w, h, bx, by, ctypes.byref(p), out.ctypes.data,
                            ctypes.cast(sym, ctypes.c_void_p) if sym is not None else None, ctypes.byref(st))
    if r != 0:
        raise RuntimeError("oracle_encode failed: %d" % r)
    return out, sym, st


def table(tid):
    n = lib().oracle_table(tid, None, 0)
    buf = (ctypes.c_int32 * n)()
    lib().oracle_table(tid, buf, n)
    return list(buf)


def footprint_info(bx, by):
    c = (ctypes.c_int32 * 8)()
    lib().oracle_footprint_info(bx, by, c, 8)
    keys = ["decimation_modes", "block_modes", "trial_dm", "trial_bm", "selected_p2", "selected_p3",
            "selected_p4", "texel_weight_entries"]
    return dict(zip(keys, list(c)))


def partition_entry(bx, by, pc, idx):
    pot = (ctypes.c_uint8 * 144)()
    sc = (ctypes.c_int32 * 6)()
    n = lib().oracle_partition_entry(bx, by, pc, idx, pot, sc)
    return list(pot)[:n], {"seed": sc[0], "counts": list(sc)[1:5], "partition_count": sc[5]}


def symbolic_to_physical(bx, by, sb):
    out = (ctypes.c_uint8 * 16)()
    r = lib().oracle_symbolic_to_physical(bx, by, ctypes.byref(sb), out)
    if r != 0:
        raise RuntimeError("oracle_symbolic_to_physical failed: %d" % r)
    return bytes(out)


def encode_ise(quant, values, bit_offset=0, nbytes=16):
    vals = np.ascontiguousarray(values, dtype=np.uint8)
    out = np.zeros(nbytes + 1, np.uint8)
    lib().oracle_encode_ise(quant, len(vals), vals.ctypes.data, out.ctypes.data, bit_offset)
    return bytes(out[:nbytes])


def compat_positions(rgba, bx, by, block_index, partition_count, batch_size=920):
    rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
    h, w = rgba.shape[:2]
    pos = (ctypes.c_int32 * 4)()
    r = lib().oracle_debug_kmeans_candidates(rgba.ctypes.data, w, h, bx, by, block_index, partition_count,
                                             batch_size, pos)
    if r != 0:
        raise RuntimeError("oracle_debug_kmeans_candidates failed")
    return list(pos)


def exhaustive_candidates(rgba, bx, by, block_index, partition_count):
    """The 4 packed table entries of the exhaustive partition search."""
    rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
    h, w = rgba.shape[:2]
    pos = (ctypes.c_int32 * 4)()
    if lib().oracle_debug_exhaustive_candidates(rgba.ctypes.data, w, h, bx, by, block_index, partition_count, pos):
        raise RuntimeError("oracle_debug_exhaustive_candidates failed")
    return list(pos)


def seeded_candidates(rgba, bx, by, block_index, partition_count, seeds, mismatch_candidates=0):
    """The 4 packed table entries (-1 = none) of the seeded candidates for
    the seed list `seeds` (best first), after `mismatch_candidates`
    partitionings of the mismatch order when more than 4 seeds are given."""
    rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
    h, w = rgba.shape[:2]
    sd = np.ascontiguousarray(seeds, dtype=np.int32)
    pos = np.zeros(4, np.int32)
    if lib().oracle_debug_seeded_candidates(rgba.ctypes.data, w, h, bx, by, block_index, partition_count,
                                            sd.ctypes.data, sd.size, mismatch_candidates, pos.ctypes.data):
        raise RuntimeError("oracle_debug_seeded_candidates failed")
    return [int(v) for v in pos]

def decimation_entry(bx, by, m):
    """(texel_count, weight_count, wx, wy, [(weight index, 16*contribution) per texel])"""
    cap = 4 + 144 * 9
    buf = (ctypes.c_int32 * cap)()
    n = lib().oracle_decimation_entry(bx, by, m, buf, cap)
    if n < 0:
        return None
    v = list(buf)[:n]
    tc, wc, wx, wy = v[:4]
    texels = []
    for t in range(tc):
        base = 4 + t * 9
        c = v[base]
        texels.append([(v[base + 1 + 2 * j], v[base + 2 + 2 * j]) for j in range(c)])
    return tc, wc, wx, wy, texels


# ---- stage-level exports (one pass at a time, for the oracle's own pins) ----
def _f32(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a


_DBG = None


def dbg():
    """The oracle library with the stage-level exports' signatures set."""
    global _DBG
    if _DBG is not None:
        return _DBG
    L = lib()
    vp, i32, f32, u32 = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_uint32
    sig = {
        "oracle_debug_pass01": (i32, [i32, i32, vp, vp, i32, vp]),
        "oracle_debug_pass02": (i32, [i32, i32, i32, vp, vp, i32, vp]),
        "oracle_debug_pass03": (i32, [i32, i32, i32, vp, vp]),
        "oracle_debug_pass04": (i32, [i32, i32, i32, vp, vp, vp, vp]),
        "oracle_debug_pass05": (i32, [i32, i32, i32, vp, vp, vp, vp]),
        "oracle_debug_trial_mode": (i32, [i32, i32, i32, vp]),
        "oracle_debug_pass07": (i32, [i32, i32, i32, i32, vp, vp, i32, f32, vp, f32, f32, vp, vp, vp]),
        "oracle_debug_pass08": (i32, [i32, i32, vp, vp, i32, vp]),
        "oracle_debug_pass09": (i32, [i32, vp, vp, vp, vp, vp, vp]),
        "oracle_debug_pass10": (i32, [i32, vp, vp, vp, vp]),
        "oracle_debug_pass11": (i32, [i32, f32, i32, vp, vp, vp, vp, vp, vp, vp, vp]),
        "oracle_debug_pass12": (i32, [vp, i32, vp, vp]),
        "oracle_debug_pass13": (i32, [i32, i32, i32, vp, vp, i32, vp, vp, vp, vp]),
        "oracle_debug_pack_endpoints": (i32, [vp, vp, vp, i32, i32, vp]),
        "oracle_debug_unpack_endpoints": (i32, [i32, vp, vp, vp]),
        "oracle_debug_pass16": (i32, [i32, i32, i32, vp, vp, i32, vp, vp, vp]),
        "oracle_debug_pass17": (f32, [i32, i32, i32, vp, vp, i32, vp, vp, vp]),
        "oracle_debug_search_pb": (i32, [i32, i32, vp, vp, i32, i32, vp]),
        "oracle_debug_kmeans": (i32, [i32, i32, vp, i32, vp]),
        "oracle_debug_mismatch": (i32, [i32, i32, vp, i32, i32]),
        "oracle_debug_ordering": (i32, [i32, u32, vp]),
        "oracle_debug_eval_candidate": (i32, [i32, i32, vp, i32, i32, vp, vp]),
        "oracle_debug_select_positions": (i32, [vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _DBG = L
    return L


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def pass01(bx, by, px, tp, pc):
    """dict of avg, dir, ep0, ep1 ([4][4]), weights, wes ([tc]), is_const, cutoff."""
    tc = bx * by
    px = np.ascontiguousarray(px, dtype=np.float32).reshape(tc, 4)
    tp = np.ascontiguousarray(tp, dtype=np.uint8)
    out = np.zeros(64 + 2 * tc + 2, np.float32)
    assert dbg().oracle_debug_pass01(bx, by, _p(px), _p(tp), pc, _p(out)) == tc
    return {"avg": out[0:16].reshape(4, 4), "dir": out[16:32].reshape(4, 4), "ep0": out[32:48].reshape(4, 4),
            "ep1": out[48:64].reshape(4, 4), "weights": out[64:64 + tc], "wes": out[64 + tc:64 + 2 * tc],
            "is_const": int(out[64 + 2 * tc]), "cutoff": float(out[65 + 2 * tc])}


def pass02(bx, by, m, weights, wes, is_const):
    w = np.zeros(144, np.float32)
    w[:bx * by] = weights
    s = np.zeros(144, np.float32)
    s[:bx * by] = wes
    dw = np.zeros(64, np.float32)
    n = dbg().oracle_debug_pass02(bx, by, m, _p(w), _p(s), int(is_const), _p(dw))
    assert n > 0
    return dw[:n]


def pass03(bx, by, m, dw):
    d = np.zeros(64, np.float32)
    d[:len(dw)] = dw
    ao = np.zeros(16, np.float32)
    assert dbg().oracle_debug_pass03(bx, by, m, _p(d), _p(ao)) == 16
    return ao


def pass04(bx, by, m, dw, ao):
    """(records [steps][5] = lowest, span, error, cut_low, cut_high; max_quant_steps, max_quant_level)"""
    d = np.zeros(64, np.float32)
    d[:len(dw)] = dw
    ao = np.ascontiguousarray(ao, dtype=np.float32)
    hl = np.zeros(16 * 5, np.float32)
    info = np.zeros(2, np.int32)
    n = dbg().oracle_debug_pass04(bx, by, m, _p(d), _p(ao), _p(hl), _p(info))
    assert n > 0
    return hl[:n * 5].reshape(n, 5), int(info[0]), int(info[1])


def pass05(bx, by, m, hl, ao):
    h = np.zeros(16 * 5, np.float32)
    h[:hl.size] = np.asarray(hl, np.float32).reshape(-1)
    ao = np.ascontiguousarray(ao, dtype=np.float32)
    low = np.zeros(13, np.float32)
    high = np.zeros(13, np.float32)
    n = dbg().oracle_debug_pass05(bx, by, m, _p(h), _p(ao), _p(low), _p(high))
    assert n >= 0
    return low[:n + 1], high[:n + 1]


def trial_mode(bx, by, k):
    """(decimation trial index, weight quant, weight bits, weight count, raw mode), number of trial modes."""
    info = np.zeros(5, np.int32)
    n = dbg().oracle_debug_trial_mode(bx, by, k, _p(info))
    return tuple(int(v) for v in info), n


def pass07(bx, by, k, pc, weights, wes, is_const, cutoff, dw, low, high):
    w = np.zeros(144, np.float32)
    w[:bx * by] = weights
    s = np.zeros(144, np.float32)
    s[:bx * by] = wes
    d = np.zeros(64, np.float32)
    d[:len(dw)] = dw
    err = np.zeros(1, np.float32)
    bits = np.zeros(1, np.int32)
    qw = np.zeros(64, np.uint8)
    n = dbg().oracle_debug_pass07(bx, by, k, pc, _p(w), _p(s), int(is_const), float(cutoff), _p(d), float(low),
                                  float(high), _p(err), _p(bits), _p(qw))
    assert n > 0
    return float(err[0]), int(bits[0]), qw[:n]


def pass08(bx, by, px, tp, pc):
    tc = bx * by
    px = np.ascontiguousarray(px, dtype=np.float32).reshape(tc, 4)
    tp = np.ascontiguousarray(tp, dtype=np.uint8)
    out = np.zeros(24, np.float32)
    assert dbg().oracle_debug_pass08(bx, by, _p(px), _p(tp), pc, _p(out)) == pc
    return out.reshape(4, 6)[:pc]


def pass09(pc, ep0, ep1, counts, ece6):
    e0 = np.zeros((4, 4), np.float32)
    e1 = np.zeros((4, 4), np.float32)
    e0[:pc] = ep0
    e1[:pc] = ep1
    cn = np.zeros(4, np.int32)
    cn[:pc] = counts
    ec = np.zeros((4, 6), np.float32)
    ec[:pc] = ece6
    err = np.zeros((4, 21, 4), np.float32)
    fmt = np.zeros((4, 21, 4), np.int32)
    assert dbg().oracle_debug_pass09(pc, _p(e0), _p(e1), _p(cn), _p(ec), _p(err), _p(fmt)) == pc
    return err, fmt


def pass10(pc, err, fmt):
    err = np.ascontiguousarray(err, np.float32)
    fmt = np.ascontiguousarray(fmt, np.int32)
    ce = np.zeros((21, 13), np.float32)
    cf = np.zeros((21, 13, 4), np.int32)
    assert dbg().oracle_debug_pass10(pc, _p(err), _p(fmt), _p(ce), _p(cf)) == 0
    return ce, cf


def pass11(pc, weight_error, bitcount, err, fmt, comb_err=None, comb_fmt=None):
    err = np.ascontiguousarray(err, np.float32)
    fmt = np.ascontiguousarray(fmt, np.int32)
    ce = None if comb_err is None else np.ascontiguousarray(comb_err, np.float32)
    cf = None if comb_fmt is None else np.ascontiguousarray(comb_fmt, np.int32)
    tot = np.zeros(1, np.float32)
    ql = np.zeros(1, np.int32)
    qm = np.zeros(1, np.int32)
    f = np.zeros(4, np.int32)
    assert dbg().oracle_debug_pass11(pc, float(weight_error), int(bitcount), _p(err), _p(fmt),
                                     _p(ce) if ce is not None else None, _p(cf) if cf is not None else None,
                                     _p(tot), _p(ql), _p(qm), _p(f)) == 0
    return float(tot[0]), int(ql[0]), int(qm[0]), [int(v) for v in f]


def pass12(errors):
    e = np.ascontiguousarray(errors, np.float32)
    idx = np.zeros(8, np.int32)
    top = np.zeros(8, np.float32)
    dbg().oracle_debug_pass12(_p(e), len(e), _p(idx), _p(top))
    return idx, top


def pass13(bx, by, k, px, tp, pc, avg, ep0, ep1, qw):
    tc = bx * by
    px = np.ascontiguousarray(px, dtype=np.float32).reshape(tc, 4)
    tp = np.ascontiguousarray(tp, dtype=np.uint8)
    a = np.zeros((4, 4), np.float32)
    a[:pc] = np.asarray(avg, np.float32)[:pc]
    ep = np.zeros((2, 4, 4), np.float32)
    ep[0, :pc] = np.asarray(ep0, np.float32)[:pc]
    ep[1, :pc] = np.asarray(ep1, np.float32)[:pc]
    q = np.zeros(64, np.uint8)
    q[:len(qw)] = qw
    rgbs = np.zeros((4, 4), np.float32)
    assert dbg().oracle_debug_pass13(bx, by, k, _p(px), _p(tp), pc, _p(a), _p(ep), _p(q), _p(rgbs)) == 0
    return ep[0], ep[1], rgbs


def pack_endpoints(ep0, ep1, rgbs, fmt, q):
    e0 = np.ascontiguousarray(ep0, np.float32)
    e1 = np.ascontiguousarray(ep1, np.float32)
    r = np.ascontiguousarray(rgbs, np.float32)
    v = np.zeros(8, np.uint32)
    f = dbg().oracle_debug_pack_endpoints(_p(e0), _p(e1), _p(r), int(fmt), int(q), _p(v))
    return f, v


def unpack_endpoints(fmt, v):
    vv = np.zeros(8, np.uint32)
    vv[:len(v)] = v
    e0 = np.zeros(4, np.int32)
    e1 = np.zeros(4, np.int32)
    dbg().oracle_debug_unpack_endpoints(int(fmt), _p(vv), _p(e0), _p(e1))
    return e0, e1


def pass16(bx, by, k, px, tp, pc, ue0, ue1, qw):
    tc = bx * by
    px = np.ascontiguousarray(px, dtype=np.float32).reshape(tc, 4)
    tp = np.ascontiguousarray(tp, dtype=np.uint8)
    u0 = np.zeros((4, 4), np.int32)
    u1 = np.zeros((4, 4), np.int32)
    u0[:pc] = np.asarray(ue0)[:pc]
    u1[:pc] = np.asarray(ue1)[:pc]
    q = np.zeros(64, np.uint8)
    q[:len(qw)] = qw
    assert dbg().oracle_debug_pass16(bx, by, k, _p(px), _p(tp), pc, _p(u0), _p(u1), _p(q)) == 0
    return q[:len(qw)]


def pass17(bx, by, k, px, tp, pc, ue0, ue1, qw):
    tc = bx * by
    px = np.ascontiguousarray(px, dtype=np.float32).reshape(tc, 4)
    tp = np.ascontiguousarray(tp, dtype=np.uint8)
    u0 = np.zeros((4, 4), np.int32)
    u1 = np.zeros((4, 4), np.int32)
    u0[:pc] = np.asarray(ue0)[:pc]
    u1[:pc] = np.asarray(ue1)[:pc]
    q = np.zeros(64, np.uint8)
    q[:len(qw)] = qw
    return float(dbg().oracle_debug_pass17(bx, by, k, _p(px), _p(tp), pc, _p(u0), _p(u1), _p(q)))


def search_pb(bx, by, px, tp, pc, pb_in_batch=0):
    tc = bx * by
    px = np.ascontiguousarray(px, dtype=np.float32).reshape(tc, 4)
    tp = np.ascontiguousarray(tp, dtype=np.uint8)
    s = OrSymbolic()
    assert dbg().oracle_debug_search_pb(bx, by, _p(px), _p(tp), pc, pb_in_batch, ctypes.byref(s)) == 0
    return s


def kmeans(bx, by, px, pc):
    tc = bx * by
    px = np.ascontiguousarray(px, dtype=np.float32).reshape(tc, 4)
    a = np.zeros(144, np.uint8)
    assert dbg().oracle_debug_kmeans(bx, by, _p(px), pc, _p(a)) == tc
    return a[:tc]


def mismatch(bx, by, assign, pc, pidx):
    a = np.zeros(144, np.uint8)
    a[:len(assign)] = assign
    return dbg().oracle_debug_mismatch(bx, by, _p(a), pc, pidx)


def ordering(nsel, score0):
    o = np.zeros(128, np.uint32)
    dbg().oracle_debug_ordering(nsel, score0, _p(o))
    return o


def eval_candidate(bx, by, px, pc, entry):
    tc = bx * by
    px = np.ascontiguousarray(px, dtype=np.float32).reshape(tc, 4)
    eu = ctypes.c_float()
    es = ctypes.c_float()
    assert dbg().oracle_debug_eval_candidate(bx, by, _p(px), pc, entry, ctypes.byref(eu), ctypes.byref(es)) == 0
    return eu.value, es.value


def select_positions(eu, es):
    a = np.ascontiguousarray(eu, np.float32)
    b = np.ascontiguousarray(es, np.float32)
    pos = np.zeros(4, np.int32)
    dbg().oracle_debug_select_positions(_p(a), _p(b), _p(pos))
    return [int(v) for v in pos]
