"""ASTC specification formulas restated for the tests (independent of the
oracle and of the product): block-mode decode, the weight-grid infill
table, LDR colour endpoint decoding, bounded integer sequence (ISE) bit
counts and decoding, and a complete LDR texel decoder built from them.
Khronos Data Format Specification, ASTC chapter (block mode layout, weight
infill, colour endpoint modes 0/4/5/6/8/9/10/12/13, ISE)."""
import numpy as np

LEVELS = [2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256]


def ldr_endpoints(cem, v):
    """ASTC LDR colour endpoint decode (specification section on colour
    endpoint modes 0, 1, 4, 5, 6, 8, 9, 10, 12, 13)."""
    def bts(a, b):
        b = (b >> 1) | (a & 0x80)
        a = (a >> 1) & 0x3F
        if a & 0x20:
            a -= 0x40
        return a, b

    def blue(c):
        return [(c[0] + c[2]) >> 1, (c[1] + c[2]) >> 1, c[2], c[3]]

    cl = lambda x: max(0, min(255, x))
    if cem == 0:
        return [v[0]] * 3 + [255], [v[1]] * 3 + [255]
    if cem == 1:
        l0 = (v[0] >> 2) | (v[1] & 0xC0)
        l1 = min(l0 + (v[1] & 0x3F), 255)
        return [l0] * 3 + [255], [l1] * 3 + [255]
    if cem == 4:
        return [v[0]] * 3 + [v[2]], [v[1]] * 3 + [v[3]]
    if cem == 5:
        b1, a0 = bts(v[1], v[0])
        b3, a2 = bts(v[3], v[2])
        return [a0] * 3 + [a2], [cl(a0 + b1)] * 3 + [cl(a2 + b3)]
    if cem in (6, 10):
        e0 = [(v[c] * v[3]) >> 8 for c in range(3)] + [255 if cem == 6 else v[4]]
        e1 = [v[0], v[1], v[2], 255 if cem == 6 else v[5]]
        return e0, e1
    if cem in (8, 12):
        c0 = [v[0], v[2], v[4], v[6] if cem == 12 else 255]
        c1 = [v[1], v[3], v[5], v[7] if cem == 12 else 255]
        if sum(c1[:3]) < sum(c0[:3]):
            return blue(c1), blue(c0)
        return c0, c1
    if cem in (9, 13):
        a = [v[0], v[2], v[4], v[6] if cem == 13 else 255]
        b = [v[1], v[3], v[5], v[7] if cem == 13 else 0]
        for c in range(4 if cem == 13 else 3):
            b[c], a[c] = bts(b[c], a[c])
        c0 = list(a)
        c1 = [a[c] + b[c] for c in range(4)]
        if cem == 9:
            c0[3] = c1[3] = 255
        if sum(b[:3]) < 0:
            return [cl(x) for x in blue(c1)], [cl(x) for x in blue(c0)]
        return [cl(x) for x in c0], [cl(x) for x in c1]
    raise ValueError(cem)


def mode_grid_and_quant(m):
    """(grid width, grid height, weight quant level) of a 1-plane 2D block
    mode, decoded from its bits (ASTC specification block mode layout;
    quant level = (R - 2) + 6 H in the QUANT_2..QUANT_32 order)."""
    A_ = (m >> 5) & 3
    if m & 3:
        R = ((m >> 4) & 1) | ((m & 3) << 1)
        B = (m >> 7) & 3
        k = (m >> 2) & 3
        wx, wy = [(B + 4, A_ + 2), (B + 8, A_ + 2), (A_ + 2, B + 8),
                  ((B & 1) + 2, A_ + 2) if (m & 0x100) else (A_ + 2, (B & 1) + 6)][k]
        H = (m >> 9) & 1
    else:
        R = ((m >> 4) & 1) | (((m >> 2) & 3) << 1)
        B = (m >> 9) & 3
        k = (m >> 7) & 3
        wx, wy = [(12, A_ + 2), (A_ + 2, 12), (A_ + 6, B + 6),
                  (6, 10) if ((m >> 5) & 3) == 0 else (10, 6)][k]
        H = 0 if k == 2 else (m >> 9) & 1
    return wx, wy, (R - 2) + 6 * H


def grid(bx, by, wx, wy):
    """Per texel list of (weight index, weight/16), in the order the
    reference's decimation tables store them."""
    out = []
    for y in range(by):
        for x in range(bx):
            xw = (((1024 + bx // 2) // (bx - 1)) * x * (wx - 1) + 32) >> 6
            yw = (((1024 + by // 2) // (by - 1)) * y * (wy - 1) + 32) >> 6
            xf, yf, base = xw & 15, yw & 15, (xw >> 4) + (yw >> 4) * wx
            w3 = (xf * yf + 8) >> 4
            ws = [16 - xf - yf + w3, xf - w3, yf - w3, w3]
            idx = [base, base + 1, base + wx, base + wx + 1]
            out.append([(idx[k], ws[k]) for k in range(4) if ws[k] != 0])
    return out


def ise_shape(q):
    """(bits, trits, quints) of quant level q (0 = QUANT_2 .. 20 = QUANT_256)."""
    n = LEVELS[q]
    for t, div in ((1, 3), (0, 5)):
        if n % div == 0 and (n // div) & (n // div - 1) == 0 and n // div >= 1 and n != 1:
            b = (n // div).bit_length() - 1
            return (b, 1, 0) if div == 3 else (b, 0, 1)
    return (n.bit_length() - 1, 0, 0)


def ise_bitcount(count, q):
    """Bits of `count` ISE symbols at quant level q: bits*count plus
    ceil(8*count/5) for trits or ceil(7*count/3) for quints."""
    b, t, qn = ise_shape(q)
    return b * count + (8 * count + 4) // 5 * t + (7 * count + 2) // 3 * qn


def quant_mode_table(pairs, bits):
    """Largest quant level (0..20) that fits 2*pairs ISE values in `bits`
    bits, -1 when none (the reference's QUANT_MODE_TABLE; the search uses
    only levels >= QUANT_6 = 4)."""
    best = -1
    for q in range(0, 21 if pairs > 0 else 0):
        if ise_bitcount(2 * pairs, q) <= bits:
            best = q
    return best


# ---- bounded integer sequence decoding ---------------------------------------
def read_bits(block, start, n):
    v = int.from_bytes(bytes(block), "little")
    return (v >> start) & ((1 << n) - 1)


_read = read_bits


def _trits_of(T):
    """Five trits from the 8-bit packed value T (specification procedure)."""
    bit = lambda v, i: (v >> i) & 1
    if (T >> 2) & 7 == 7:
        C = (((T >> 5) & 7) << 2) | (T & 3)
        t4 = t3 = 2
    else:
        C = T & 0x1F
        if (T >> 5) & 3 == 3:
            t4, t3 = 2, bit(T, 7)
        else:
            t4, t3 = bit(T, 7), (T >> 5) & 3
    if C & 3 == 3:
        t2, t1, t0 = 2, bit(C, 4), (bit(C, 3) << 1) | (bit(C, 2) & (1 - bit(C, 3)))
    elif (C >> 2) & 3 == 3:
        t2, t1, t0 = 2, 2, C & 3
    else:
        t2, t1, t0 = bit(C, 4), (C >> 2) & 3, (bit(C, 1) << 1) | (bit(C, 0) & (1 - bit(C, 1)))
    return [t0, t1, t2, t3, t4]


def _quints_of(Q):
    """Three quints from the 7-bit packed value Q (specification procedure)."""
    bit = lambda v, i: (v >> i) & 1
    if (Q >> 1) & 3 == 3 and (Q >> 5) & 3 == 0:
        q2 = (bit(Q, 0) << 2) | ((bit(Q, 4) & (1 - bit(Q, 0))) << 1) | (bit(Q, 3) & (1 - bit(Q, 0)))
        q1 = q0 = 4
    else:
        if (Q >> 1) & 3 == 3:
            q2 = 4
            C = (((Q >> 3) & 3) << 3) | ((~(Q >> 5) & 3) << 1) | bit(Q, 0)
        else:
            q2 = (Q >> 5) & 3
            C = Q & 0x1F
        if C & 7 == 5:
            q1, q0 = 4, (C >> 3) & 3
        else:
            q1, q0 = (C >> 3) & 3, C & 7
    return [q0, q1, q2]


def ise_decode(q, count, block, start):
    """`count` values of quant level q stored from bit `start` of the
    little-endian bit string `block` (bytes); returns (digit, low bits)
    pairs: digit = trit / quint or None."""
    b, t, qn = ise_shape(q)
    out = []
    pos = start
    end = start + ise_bitcount(count, q)  # bits past the sequence's length read as zero

    def read_bits(blk, p, n):
        n = max(0, min(n, end - p))
        return _read(blk, p, n)
    if t:
        for g in range(0, count, 5):
            m, T = [], 0
            tbits = [2, 2, 1, 2, 1]
            tshift = [0, 2, 4, 5, 7]
            for i in range(5):
                m.append(read_bits(block, pos, b))
                pos += b
                T |= read_bits(block, pos, tbits[i]) << tshift[i]
                pos += tbits[i]
            tr = _trits_of(T)
            out += [(tr[i], m[i]) for i in range(5)]
    elif qn:
        for g in range(0, count, 3):
            m, Q = [], 0
            qbits = [3, 2, 2]
            qshift = [0, 3, 5]
            for i in range(3):
                m.append(read_bits(block, pos, b))
                pos += b
                Q |= read_bits(block, pos, qbits[i]) << qshift[i]
                pos += qbits[i]
            qu = _quints_of(Q)
            out += [(qu[i], m[i]) for i in range(3)]
    else:
        for i in range(count):
            out.append((None, read_bits(block, pos, b)))
            pos += b
    return out[:count]


def _replicate(v, nbits, width):
    out, have = 0, 0
    while have < width:
        out = (out << nbits) | v
        have += nbits
    return out >> (have - width)


def color_unquant(q, digit, m):
    """Colour endpoint unquantisation to 0..255 (specification: bit
    replication, or the A/B/C/D trit and quint expansion)."""
    b, t, qn = ise_shape(q)
    if not t and not qn:
        return _replicate(m, b, 8)
    a = m & 1
    A = 0x1FF if a else 0
    bits = [(m >> i) & 1 for i in range(b)]
    B = 0
    if t:
        C = {1: 204, 2: 93, 3: 44, 4: 22, 5: 11, 6: 5}[b]
        pat = {1: "", 2: "b000b0bb0", 3: "cb000cbcb", 4: "dcb000dcb", 5: "edcb000ed", 6: "fedcb000f"}[b]
    else:
        C = {1: 113, 2: 54, 3: 26, 4: 13, 5: 6}[b]
        pat = {1: "", 2: "b0000bb00", 3: "cb0000cbc", 4: "dcb0000dc", 5: "edcb0000e"}[b]
    for ch in pat:
        B = (B << 1) | (0 if ch == "0" else bits["bcdef".index(ch) + 1])
    T = digit * C + B
    T ^= A
    return (A & 0x80) | (T >> 2)


def weight_unquant(q, digit, m):
    """Weight unquantisation to 0..64 (specification)."""
    b, t, qn = ise_shape(q)
    if not t and not qn:
        T = _replicate(m, b, 6)
    elif b == 0:
        T = ([0, 32, 63] if t else [0, 16, 32, 47, 63])[digit]
    else:
        a = m & 1
        A = 0x7F if a else 0
        bits = [(m >> i) & 1 for i in range(b)]
        if t:
            C = {1: 50, 2: 23, 3: 11}[b]
            pat = {1: "", 2: "b000b0b", 3: "cb000cb"}[b]
        else:
            C = {1: 28, 2: 13}[b]
            pat = {1: "", 2: "b0000b0"}[b]
        B = 0
        for ch in pat:
            B = (B << 1) | (0 if ch == "0" else bits["bc".index(ch) + 1])
        T = digit * C + B
        T ^= A
        T = (A & 0x20) | (T >> 2)
    return T + 1 if T > 32 else T


# ---- partition function ------------------------------------------------------
def _hash52(p):
    M = 0xFFFFFFFF
    p ^= p >> 15
    p = (p - (p << 17)) & M
    p = (p + (p << 7)) & M
    p = (p + (p << 4)) & M
    p ^= p >> 5
    p = (p + (p << 16)) & M
    p ^= p >> 7
    p ^= p >> 3
    p = (p ^ (p << 6)) & M
    p ^= p >> 17
    return p


def select_partition(seed, x, y, pc, small_block):
    """The specification's partition hash (2D, z = 0)."""
    if small_block:
        x, y = x << 1, y << 1
    seed += (pc - 1) * 1024
    r = _hash52(seed)
    s = [(r >> sh) & 0xF for sh in (0, 4, 8, 12, 16, 20, 24, 28, 18, 22, 26)]
    s.append(((r >> 30) | (r << 2)) & 0xF)
    s = [v * v for v in s]
    if seed & 1:
        sh1, sh2 = (4 if seed & 2 else 5), (6 if pc == 3 else 5)
    else:
        sh1, sh2 = (6 if pc == 3 else 5), (4 if seed & 2 else 5)
    sh3 = sh1 if seed & 0x10 else sh2
    for i in range(8):
        s[i] >>= sh1 if i % 2 == 0 else sh2
    for i in range(8, 12):
        s[i] >>= sh3
    a = (s[0] * x + s[1] * y + (r >> 14)) & 0x3F
    b = (s[2] * x + s[3] * y + (r >> 10)) & 0x3F
    c = (s[4] * x + s[5] * y + (r >> 6)) & 0x3F
    d = (s[6] * x + s[7] * y + (r >> 2)) & 0x3F
    if pc <= 3:
        d = 0
    if pc <= 2:
        c = 0
    if a >= b and a >= c and a >= d:
        return 0
    if b >= c and b >= d:
        return 1
    if c >= d:
        return 2
    return 3


# ---- the whole LDR block decoder ----------------------------------------------
ERROR_COLOUR = [255, 0, 255, 255]


def _mode_fields(m):
    """(wx, wy, weight quant, dual plane) or None for reserved modes."""
    if m & 3 == 0 and (m >> 2) & 3 == 0:
        return None
    wx, wy, q = mode_grid_and_quant(m)
    D = (m >> 10) & 1
    if m & 3 == 0 and (m >> 7) & 3 == 2:
        D = 0
    if m & 3 == 0 and (m >> 7) & 3 == 3 and (m >> 5) & 3 > 1:
        return None
    return wx, wy, q, D


def decode_block(bx, by, block):
    """RGBA8 texels (bx*by lists) of one ASTC LDR block; the error colour
    for illegal blocks (reserved modes, grids larger than the block, weight
    or colour data that does not fit, HDR endpoint modes)."""
    tc = bx * by
    err = [list(ERROR_COLOUR) for _ in range(tc)]
    m = read_bits(block, 0, 11)
    if m & 0x1FF == 0x1FC:
        if m & 0x200:
            return err  # HDR void extent
        col = [read_bits(block, 64 + 16 * c, 16) for c in range(4)]
        return [[round(v * 255 / 65535) for v in col] for _ in range(tc)]
    f = _mode_fields(m)
    if f is None:
        return err
    wx, wy, wq, D = f
    if wx > bx or wy > by:
        return err
    pc = read_bits(block, 11, 2) + 1
    if D and pc == 4:
        return err
    nw = wx * wy * (1 + D)
    wbits = ise_bitcount(nw, wq)
    if nw > 64 or wbits < 24 or wbits > 96:
        return err
    below = 128 - wbits
    if pc == 1:
        cems = [read_bits(block, 13, 4)]
        cstart = 17
        seed = 0
    else:
        seed = read_bits(block, 13, 10)
        cstart = 29
        low = read_bits(block, 23, 6)
        if low & 3 == 0:
            cems = [low >> 2] * pc
        else:
            extra = 3 * pc - 4
            below -= extra
            enc = low | (read_bits(block, below, extra) << 6)
            base = (enc & 3) - 1
            cls = [base + ((enc >> (2 + i)) & 1) for i in range(pc)]
            cems = [(cls[i] << 2) | ((enc >> (2 + pc + 2 * i)) & 3) for i in range(pc)]
    ccs = 0
    if D:
        below -= 2
        ccs = read_bits(block, below, 2)
    if any(c in (2, 3, 7, 11, 14, 15) for c in cems):
        return err  # HDR endpoint modes
    nvals = sum(2 * ((c >> 2) + 1) for c in cems)
    if nvals > 18:
        return err
    cq = -1
    for q in range(4, 21):
        if ise_bitcount(nvals, q) <= below - cstart:
            cq = q
    if cq < 0:
        return err
    vals = [color_unquant(cq, d, mm) for d, mm in ise_decode(cq, nvals, block, cstart)]
    rev = bytes(int("{:08b}".format(x)[::-1], 2) for x in bytes(block)[::-1])
    wv = [weight_unquant(wq, d, mm) for d, mm in ise_decode(wq, nw, rev, 0)]
    eps, vi = [], 0
    for c in cems:
        n = 2 * ((c >> 2) + 1)
        eps.append(ldr_endpoints(c, vals[vi:vi + n]))
        vi += n
    g = grid(bx, by, wx, wy)
    planes = 1 + D
    out = []
    for t in range(tc):
        x, y = t % bx, t // bx
        p = select_partition(seed, x, y, pc, tc < 31) if pc > 1 else 0
        w = [(sum(wv[i * planes + pl] * c for i, c in g[t]) + 8) >> 4 for pl in range(planes)]
        e0, e1 = eps[p]
        px = []
        for c in range(4):
            wt = w[1] if (D and c == ccs) else w[0]
            v16 = (e0[c] * 257 * (64 - wt) + e1[c] * 257 * wt + 32) >> 6
            px.append((v16 * 255 + 32767) // 65535)
        out.append(px)
    return out


def decode_image(bx, by, w, h, blocks):
    """Row-major blocks -> H x W x 4 uint8 image through decode_block."""
    bxn, byn = (w + bx - 1) // bx, (h + by - 1) // by
    img = np.zeros((byn * by, bxn * bx, 4), np.uint8)
    blocks = np.asarray(blocks, np.uint8).reshape(-1, 16)
    for j in range(byn):
        for i in range(bxn):
            tex = np.array(decode_block(bx, by, bytes(blocks[j * bxn + i])), np.uint8).reshape(by, bx, 4)
            img[j * by:(j + 1) * by, i * bx:(i + 1) * bx] = tex
    return img[:h, :w]
