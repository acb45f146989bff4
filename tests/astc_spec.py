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
    endpoint modes 0, 4, 5, 6, 8, 9, 10, 12, 13)."""
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
