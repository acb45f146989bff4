"""astc_implementation_amd: ASTC LDR block encoder for AMD Instinct MI355X.

Python view of the C ABI in include/astcg.h (libastcg.so, built in-tree by
``astc_implementation_amd.build``).  The encoder mirrors the reference's
``ASTCEncoder`` (code/astc.h:724-748): construct on a device, ``configure``
(= secondaryInit) for an image size and footprint, then ``encode``.

The library is required: importing succeeds without it, but every entry
point raises ``RuntimeError`` when libastcg.so is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libastcg.so")

OK = 0
ERR_OUT_OF_MEMORY = 6
STATUS = {
    0: "ok", 1: "invalid argument", 2: "invalid footprint", 3: "buffer too small",
    4: "not configured", 5: "device error", 6: "out of memory", 7: "i/o error", 8: "unsupported",
}

FOOTPRINTS = [(4, 4), (5, 4), (5, 5), (6, 5), (6, 6), (8, 5), (8, 6),
              (10, 5), (10, 6), (8, 8), (10, 8), (10, 10), (12, 10), (12, 12)]


class AstcError(RuntimeError):
    def __init__(self, status, msg=""):
        self.status = status
        super().__init__("astcg: %s%s" % (STATUS.get(status, "status %d" % status), (": " + msg) if msg else ""))


class SymbolicBlock(ctypes.Structure):
    """SymbolicBlock (code/astc.h:639-655), 432 bytes."""
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


class Params(ctypes.Structure):
    _fields_ = [
        ("max_partitions", ctypes.c_uint32),
        ("policy", ctypes.c_uint32),
        ("chunk_blocks", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32 * 4),
    ]


FLAG_PROFILE = 1
FLAG_LABELS = 2
FLAG_VOID_EXTENT = 4
FLAG_GRAPH = 8
FLAG_DUAL_PLANE = 16


class Texture(ctypes.Structure):
    """astcg_texture: one texture of a batch."""
    _fields_ = [
        ("rgba", ctypes.c_void_p),
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("block_x", ctypes.c_uint32),
        ("block_y", ctypes.c_uint32),
        ("out", ctypes.c_void_p),
        ("out_len", ctypes.c_size_t),
    ]


class LearnedParams(ctypes.Structure):
    """astcg_learned_params: pruning of the LEARNED candidate policy."""
    _fields_ = [
        ("seeds", ctypes.c_uint32),
        ("modes", ctypes.c_uint32),
        ("slots", ctypes.c_uint32),
        ("fallback_below", ctypes.c_float),
        ("count_below", ctypes.c_float),
        ("partitionings", ctypes.c_uint32),
        ("refine_partitionings", ctypes.c_uint32),
        ("early_out", ctypes.c_float),
        ("full_search_above", ctypes.c_float),
        ("rounds", ctypes.c_uint32),
        ("partitioned_slots", ctypes.c_uint32),
        ("candidates", ctypes.c_uint32),
        ("count_above", ctypes.c_float),
        ("full_search_modes", ctypes.c_uint32),
        ("mismatch_candidates", ctypes.c_uint32),
        ("fallback_above", ctypes.c_float),
    ]


POLICY_COMPAT = 0
POLICY_LEARNED = 1
POLICY_EXHAUSTIVE = 2


class Tuning(ctypes.Structure):
    """astcg_tuning: launch shapes of one footprint's search kernels."""
    _fields_ = [("refine_stage_max_tc", ctypes.c_uint32), ("ex_f4_max_tc", ctypes.c_uint32),
                ("exw_max_list", ctypes.c_uint32), ("refine_queue", ctypes.c_uint32),
                ("refine_classes", ctypes.c_uint32)]


class KernelTime(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("total_ms", ctypes.c_double), ("launches", ctypes.c_uint64)]


_lib = None


def library():
    """Loads libastcg.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libastcg.so not built: run astc_implementation_amd.build.build()")
    # PyTorch-ROCm ships its own HIP runtime; loading it first makes this
    # library bind to the same runtime, so both can share device memory and
    # streams in one process.
    try:
        import torch  # noqa: F401
    except Exception:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    vp, u32, sz, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_size_t, ctypes.c_int
    sig = {
        "astcg_default_params": (None, [vp]),
        "astcg_create": (i32, [i32, ctypes.POINTER(vp)]),
        "astcg_destroy": (i32, [vp]),
        "astcg_configure": (i32, [vp, u32, u32, u32, u32, vp]),
        "astcg_encode": (i32, [vp, vp, sz, vp, sz]),
        "astcg_encode_device": (i32, [vp, vp, vp, vp]),
        "astcg_encode_device_noalloc": (i32, [vp, vp, vp, vp]),
        "astcg_encode_batch": (i32, [vp, vp, u32, vp]),
        "astcg_decode_device": (i32, [vp, u32, u32, u32, u32, vp, vp, vp]),
        "astcg_get_labels": (i32, [vp, vp, sz]),
        "astcg_get_candidates": (i32, [vp, vp, vp, sz]),
        "astcg_default_learned_params": (None, [vp]),
        "astcg_set_predictor": (i32, [vp, u32, u32, u32, vp, sz, vp]),
        "astcg_fallback_rate": (i32, [vp, ctypes.POINTER(ctypes.c_double)]),
        "astcg_trial_modes": (i32, [u32, u32, ctypes.POINTER(u32), vp, sz]),
        "astcg_decode": (i32, [vp, u32, u32, u32, u32, vp, sz, vp, sz]),
        "astcg_encode_batch_device": (i32, [vp, vp, u32, vp, vp]),
        "astcg_get_symbolic": (i32, [vp, vp, sz]),
        "astcg_block_counts": (i32, [vp, vp, vp, vp]),
        "astcg_symbolic_to_physical": (i32, [u32, u32, vp, vp]),
        "astcg_encode_ise": (i32, [u32, u32, vp, vp, u32]),
        "astcg_ise_bitcount": (u32, [u32, u32]),
        "astcg_store_image": (i32, [u32, u32, u32, u32, vp, sz, ctypes.c_char_p]),
        "astcg_astc_file_in_memory": (i32, [u32, u32, u32, u32, vp, sz, vp, sz, ctypes.POINTER(sz)]),
        "astcg_decode_image": (i32, [u32, u32, u32, u32, vp, sz, vp, sz]),
        "astcg_physical_to_symbolic": (i32, [u32, u32, vp, vp]),
        "astcg_kernel_times": (i32, [vp, vp, sz, ctypes.POINTER(sz)]),
        "astcg_reset_kernel_times": (i32, [vp]),
        "astcg_get_tuning": (i32, [vp, u32, u32, vp]),
        "astcg_set_tuning": (i32, [vp, u32, u32, vp]),
        "astcg_autotune": (i32, [vp, u32, u32, ctypes.c_char_p, vp]),
        "astcg_last_error": (ctypes.c_char_p, [vp]),
        "astcg_status_string": (ctypes.c_char_p, [i32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(status, enc=None):
    if status != OK:
        msg = ""
        if enc is not None:
            msg = (library().astcg_last_error(enc) or b"").decode()
        raise AstcError(status, msg)


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class Encoder:
    """ASTCEncoder on one HIP device (code/compression_pipeline.cpp:200-535)."""

    def __init__(self, device=0):
        lib = library()
        h = ctypes.c_void_p()
        _check(lib.astcg_create(device, ctypes.byref(h)))
        self._h = h
        self.width = self.height = self.bx = self.by = 0

    def close(self):
        if getattr(self, "_h", None):
            library().astcg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def configure(self, width, height, bx, by, max_partitions=4, chunk_blocks=0, profile=False, labels=False,
                  policy=POLICY_COMPAT, void_extent=False, graph=False, dual_plane=False):
        """secondaryInit (compression_pipeline.cpp:263-321)."""
        p = Params()
        library().astcg_default_params(ctypes.byref(p))
        p.max_partitions = max_partitions
        p.chunk_blocks = chunk_blocks
        p.policy = policy
        p.flags = ((FLAG_PROFILE if profile else 0) | (FLAG_LABELS if labels else 0) |
                   (FLAG_VOID_EXTENT if void_extent else 0) | (FLAG_GRAPH if graph else 0) |
                   (FLAG_DUAL_PLANE if dual_plane else 0))
        _check(library().astcg_configure(self._h, width, height, bx, by, ctypes.byref(p)), self._h)
        self.width, self.height, self.bx, self.by = width, height, bx, by

    @property
    def num_blocks(self):
        n = ctypes.c_uint32()
        _check(library().astcg_block_counts(self._h, None, None, ctypes.byref(n)), self._h)
        return n.value

    def encode(self, rgba):
        """encode (compression_pipeline.cpp:323-535): host RGBA8 -> blocks."""
        rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
        out = np.zeros(self.num_blocks * 16, np.uint8)
        _check(library().astcg_encode(self._h, _ptr(rgba), rgba.nbytes, _ptr(out), out.nbytes), self._h)
        return out

    def encode_device(self, d_rgba, d_out, stream=None):
        """Enqueues an encode of device buffers (integer addresses) on a HIP stream."""
        _check(library().astcg_encode_device(self._h, ctypes.c_void_p(d_rgba), ctypes.c_void_p(d_out),
                                             ctypes.c_void_p(stream or 0)), self._h)

    def encode_device_noalloc(self, d_rgba, d_out, stream=None):
        """Same encode as encode_device with no device allocation: the working
        memory is the library's fixed per-GPU scratch array."""
        _check(library().astcg_encode_device_noalloc(self._h, ctypes.c_void_p(d_rgba), ctypes.c_void_p(d_out),
                                                     ctypes.c_void_p(stream or 0)), self._h)

    def encode_batch(self, images, footprints, max_partitions=4, chunk_blocks=0, profile=False,
                     policy=POLICY_COMPAT, void_extent=False, dual_plane=False):
        """Encodes several RGBA8 images (H x W x 4 uint8 arrays) with their own
        footprints [(bx, by), ...] in one batched call; returns the block
        arrays in input order."""
        imgs = [np.ascontiguousarray(im, dtype=np.uint8) for im in images]
        outs, tex = [], (Texture * max(len(imgs), 1))()
        for i, (im, (bx, by)) in enumerate(zip(imgs, footprints)):
            h, w = im.shape[:2]
            out = np.zeros(((w + bx - 1) // bx) * ((h + by - 1) // by) * 16, np.uint8)
            outs.append(out)
            tex[i] = Texture(im.ctypes.data, w, h, bx, by, out.ctypes.data, out.nbytes)
        p = Params()
        library().astcg_default_params(ctypes.byref(p))
        p.max_partitions = max_partitions
        p.chunk_blocks = chunk_blocks
        p.flags = ((FLAG_PROFILE if profile else 0) | (FLAG_VOID_EXTENT if void_extent else 0) |
                   (FLAG_DUAL_PLANE if dual_plane else 0))
        p.policy = policy
        _check(library().astcg_encode_batch(self._h, tex, len(imgs), ctypes.byref(p)), self._h)
        return outs

    def encode_batch_device(self, textures, max_partitions=4, stream=None, policy=POLICY_COMPAT, profile=False,
                            graph=False, void_extent=False):
        """textures: [(d_rgba, width, height, bx, by, d_out, out_len)] with
        device addresses; enqueued on `stream`."""
        tex = (Texture * max(len(textures), 1))()
        for i, t in enumerate(textures):
            tex[i] = Texture(*t)
        p = Params()
        library().astcg_default_params(ctypes.byref(p))
        p.max_partitions = max_partitions
        p.flags = ((FLAG_PROFILE if profile else 0) | (FLAG_GRAPH if graph else 0) |
                   (FLAG_VOID_EXTENT if void_extent else 0))
        p.policy = policy
        _check(library().astcg_encode_batch_device(self._h, tex, len(textures), ctypes.byref(p),
                                                   ctypes.c_void_p(stream or 0)), self._h)

    def decode(self, bx, by, width, height, blocks):
        """GPU ASTC LDR decode of host blocks -> (height, width, 4) uint8."""
        blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
        out = np.zeros((height, width, 4), np.uint8)
        _check(library().astcg_decode(self._h, bx, by, width, height, _ptr(blocks), blocks.nbytes, _ptr(out),
                                      out.nbytes), self._h)
        return out

    def decode_device(self, bx, by, width, height, d_blocks, d_rgba, stream=None):
        _check(library().astcg_decode_device(self._h, bx, by, width, height, ctypes.c_void_p(d_blocks),
                                             ctypes.c_void_p(d_rgba), ctypes.c_void_p(stream or 0)), self._h)

    def kernel_times(self):
        """{kernel: {"total_ms", "launches"}} since the last reset (needs profile=True)."""
        arr = (KernelTime * 16)()
        n = ctypes.c_size_t()
        _check(library().astcg_kernel_times(self._h, arr, 16, ctypes.byref(n)), self._h)
        return {arr[i].name.decode(): {"total_ms": arr[i].total_ms, "launches": arr[i].launches}
                for i in range(n.value)}

    def reset_kernel_times(self):
        _check(library().astcg_reset_kernel_times(self._h), self._h)

    def set_predictor(self, bx, by, params, hidden=256, **learned):
        """Installs predictor weights (six fp32 arrays W1, b1, W2, b2, W3, b3)
        for footprint bx x by; keyword arguments override the pruning
        defaults (seeds, modes, slots, fallback_below, count_below,
        partitionings, refine_partitionings, early_out, full_search_above,
        rounds, partitioned_slots, candidates, count_above, full_search_modes,
        mismatch_candidates, fallback_above)."""
        flat = np.concatenate([np.asarray(w, dtype=np.float32).reshape(-1) for w in params])
        lp = LearnedParams()
        library().astcg_default_learned_params(ctypes.byref(lp))
        for k, v in learned.items():
            setattr(lp, k, v)
        _check(library().astcg_set_predictor(self._h, bx, by, hidden, _ptr(flat), flat.size, ctypes.byref(lp)),
               self._h)

    def load_predictor(self, bx, by, path=None, **learned):
        """Installs a saved predictor (default: the package's
        predictor_<bx>x<by>.npz) with the pruning settings saved with it;
        keyword arguments override them."""
        params, hidden, prune = load_predictor_file(path or predictor_path(bx, by))
        prune.update(learned)
        self.set_predictor(bx, by, params, hidden=hidden, **prune)

    def tuning(self, bx, by):
        """The launch shapes of footprint bx x by (a dict of astcg_tuning)."""
        t = Tuning()
        _check(library().astcg_get_tuning(self._h, bx, by, ctypes.byref(t)), self._h)
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    def set_tuning(self, bx, by, **fields):
        """Changes some launch shapes of footprint bx x by (no result changes)."""
        t = Tuning(**self.tuning(bx, by))
        for k, v in fields.items():
            setattr(t, k, v)
        _check(library().astcg_set_tuning(self._h, bx, by, ctypes.byref(t)), self._h)

    def autotune(self, bx, by, cache_path=None):
        """astcg_autotune: the cached or the measured launch shapes of bx x by."""
        t = Tuning()
        path = os.fsencode(cache_path) if cache_path else None
        _check(library().astcg_autotune(self._h, bx, by, path, ctypes.byref(t)), self._h)
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    def fallback_rate(self):
        r = ctypes.c_double()
        _check(library().astcg_fallback_rate(self._h, ctypes.byref(r)), self._h)
        return r.value

    def labels(self):
        """(num_blocks, 6) int32 search outcome of the last encode made with
        labels=True: pc, seed, trial mode, best seed for 2 / 3 / 4 partitions."""
        out = np.zeros((self.num_blocks, 6), np.int32)
        _check(library().astcg_get_labels(self._h, _ptr(out), self.num_blocks), self._h)
        return out

    def candidates(self):
        """(seeds, errors) of the 4 compat positions of every partition count
        in the last encode made with labels=True: (num_blocks, 3, 4) int16
        and float32 (-1 / 1e30 where a position was not searched)."""
        n = self.num_blocks
        seeds = np.zeros((n, 3, 4), np.int16)
        errors = np.zeros((n, 3, 4), np.float32)
        _check(library().astcg_get_candidates(self._h, _ptr(seeds), _ptr(errors), n), self._h)
        return seeds, errors

    def symbolic(self):
        n = self.num_blocks
        arr = (SymbolicBlock * n)()
        _check(library().astcg_get_symbolic(self._h, arr, n), self._h)
        return arr


def predictor_path(bx, by):
    """The package's trained predictor for footprint bx x by."""
    return os.path.join(_HERE, "predictor_%dx%d.npz" % (bx, by))


def load_predictor_file(path):
    """(params, hidden, pruning) of a saved predictor: the six fp32 arrays
    W1, b1, W2, b2, W3, b3, the hidden width, and the LearnedParams fields
    tuned for it (empty when the file carries none)."""
    z = np.load(path, allow_pickle=False)
    params = [z["p%d" % i].astype(np.float32) for i in range(6)]
    prune = {}
    if "prune_i" in z.files:
        seeds, modes, slots = (int(v) for v in z["prune_i"])
        prune.update(seeds=seeds, modes=modes, slots=slots)
    if "prune_f" in z.files:
        fb, cb = (float(v) for v in z["prune_f"])
        prune.update(fallback_below=fb, count_below=cb)
    if "prune_p" in z.files:
        prune.update(partitionings=int(z["prune_p"]))
    if "prune_r" in z.files:
        prune.update(refine_partitionings=int(z["prune_r"]))
    if "prune_e" in z.files:
        prune.update(early_out=float(z["prune_e"]))
    if "prune_h" in z.files:
        prune.update(full_search_above=float(z["prune_h"]))
    if "prune_n" in z.files:
        prune.update(rounds=int(z["prune_n"]))
    if "prune_s" in z.files:
        prune.update(partitioned_slots=int(z["prune_s"]))
    if "prune_c" in z.files:
        prune.update(candidates=int(z["prune_c"]))
    if "prune_a" in z.files:
        prune.update(count_above=float(z["prune_a"]))
    if "prune_m" in z.files:
        prune.update(full_search_modes=int(z["prune_m"]))
    if "prune_k" in z.files:
        prune.update(mismatch_candidates=int(z["prune_k"]))
    if "prune_l" in z.files:
        prune.update(fallback_above=float(z["prune_l"]))
    return params, int(z["hidden"]), prune


def symbolic_to_physical(bx, by, sb):
    out = (ctypes.c_uint8 * 16)()
    _check(library().astcg_symbolic_to_physical(bx, by, ctypes.byref(sb), out))
    return bytes(out)


def physical_to_symbolic(bx, by, block):
    buf = (ctypes.c_uint8 * 16).from_buffer_copy(bytes(block))
    sb = SymbolicBlock()
    _check(library().astcg_physical_to_symbolic(bx, by, buf, ctypes.byref(sb)))
    return sb


def encode_ise(quant, values, bit_offset=0, nbytes=16):
    vals = np.ascontiguousarray(values, dtype=np.uint8)
    out = np.zeros(nbytes + 1, np.uint8)
    _check(library().astcg_encode_ise(quant, len(vals), _ptr(vals), _ptr(out), bit_offset))
    return bytes(out[:nbytes])


def ise_bitcount(count, quant):
    return library().astcg_ise_bitcount(count, quant)


def decode_image(bx, by, width, height, blocks):
    """ASTC LDR decode to RGBA8 (Khronos specification)."""
    blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
    out = np.zeros((height, width, 4), np.uint8)
    st = library().astcg_decode_image(bx, by, width, height, _ptr(blocks), blocks.nbytes, _ptr(out), out.nbytes)
    if st not in (OK, 8):
        _check(st)
    return out


def astc_file_bytes(bx, by, width, height, blocks):
    blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
    out = np.zeros(blocks.nbytes + 16, np.uint8)
    n = ctypes.c_size_t()
    _check(library().astcg_astc_file_in_memory(bx, by, width, height, _ptr(blocks), blocks.nbytes, _ptr(out),
                                               out.nbytes, ctypes.byref(n)))
    return bytes(out[:n.value])


def store_image(bx, by, width, height, blocks, path):
    blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
    _check(library().astcg_store_image(bx, by, width, height, _ptr(blocks), blocks.nbytes, path.encode()))


# channel weights the search minimises (code/astc.h:60-63)
CHANNEL_WEIGHTS = (0.30 * 2.25, 0.59 * 2.25, 0.11 * 2.25, 1.0)


def psnr(a, b, channels=4, weights=None):
    """PSNR (peak 255) over the first `channels` channels; with `weights`,
    of the weighted mean of the per-channel squared errors."""
    a = a[..., :channels].astype(np.float64)
    b = b[..., :channels].astype(np.float64)
    if weights is None:
        mse = np.mean((a - b) ** 2)
    else:
        wv = np.asarray(weights[:channels], np.float64)
        mse = float((np.mean((a - b) ** 2, axis=tuple(range(a.ndim - 1))) * wv).sum() / wv.sum())
    return float("inf") if mse == 0 else 10.0 * np.log10(255.0 ** 2 / mse)
