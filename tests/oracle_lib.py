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
                ("first_block", ctypes.c_int), ("block_count", ctypes.c_int)]


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
        L.oracle_decimation_entry.argtypes = [i32, i32, i32, vp, i32]
        L.oracle_decimation_entry.restype = i32
        _lib = L
    return _lib


def encode(rgba, bx, by, max_partitions=4, threads=1, first_block=0, block_count=0, batch_size=920,
           want_symbolic=True):
    """Returns (blocks uint8[n*16], symbolic list or None, stats)."""
    rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
    h, w = rgba.shape[:2]
    nb_total = ((w + bx - 1) // bx) * ((h + by - 1) // by)
    n = block_count if block_count else nb_total - first_block
    out = np.zeros(n * 16, np.uint8)
    sym = (OrSymbolic * n)() if want_symbolic else None
    st = OrStats()
    p = OrParams(max_partitions, batch_size, threads, first_block, block_count)
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
