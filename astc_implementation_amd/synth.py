"""Seeded synthetic RGBA8 images for parity tests and benchmarks.

The generators follow the workload definitions the benchmark is quoted on:

* ``gradient_image`` (512x512 seed 0 / 4096x4096 seed 1): the mean of three
  random-direction linear colour gradients, plus 4-octave value noise of peak
  amplitude 40/255 per channel (base cell 64 px, halving per octave, octave
  weights 1/2, 1/4, 1/8, 1/16, smoothstep interpolation), then 20
  axis-aligned rectangles of uniform random colour with hard edges; A = 255.
* ``alpha_image`` (4096x4096 seed 2): the same colour content with
  A = clamp(255 * (1 - r / r_max) + noise), r the distance to the centre.
* ``normal_map_image``: unit normals of a 4-octave noise height field in RG,
  B = 0, A = 255.
* ``flat_noise_image``: one uniform colour plus N(0, 3) per channel, A = 255.

All random numbers come from splitmix64(seed) so that every image is
reproducible bit for bit on any machine.
"""
import numpy as np

_MASK = (1 << 64) - 1


class SplitMix64:
    """splitmix64 stream; ``uniform`` returns doubles in [0, 1)."""

    def __init__(self, seed):
        self.state = seed & _MASK

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def uniform_array(self, n, lo=0.0, hi=1.0):
        # vectorised splitmix64 over n consecutive states
        base = np.uint64(self.state)
        steps = (np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15))
        with np.errstate(over="ignore"):
            z = base + steps
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        self.state = (self.state + n * 0x9E3779B97F4A7C15) & _MASK
        return lo + (hi - lo) * ((z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53)))

    def normal_array(self, n):
        u1 = self.uniform_array(n)
        u2 = self.uniform_array(n)
        return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def _smoothstep(t):
    return t * t * (3.0 - 2.0 * t)


def value_noise(rng, h, w, octaves=4, base_cell=64):
    """Sum of octaves of lattice value noise in [-1, 1] (weights 1/2 .. 1/16,
    renormalised so the peak amplitude is 1)."""
    ys = np.arange(h, dtype=np.float64) + 0.5
    xs = np.arange(w, dtype=np.float64) + 0.5
    total = np.zeros((h, w))
    wsum = 0.0
    for o in range(octaves):
        cell = max(base_cell >> o, 1)
        gh, gw = h // cell + 2, w // cell + 2
        lat = rng.uniform_array(gh * gw, -1.0, 1.0).reshape(gh, gw)
        fy, fx = ys / cell, xs / cell
        iy, ix = np.floor(fy).astype(np.int64), np.floor(fx).astype(np.int64)
        ty, tx = _smoothstep(fy - iy)[:, None], _smoothstep(fx - ix)[None, :]
        # interpolate every lattice row along x once, then blend rows along y
        # (the same operations, in the same order, as the bilinear formula)
        rows = lat[:, ix] * (1 - tx) + lat[:, ix + 1] * tx
        v = rows[iy] * (1 - ty) + rows[iy + 1] * ty
        wt = 0.5 ** (o + 1)
        total += wt * v
        wsum += wt
    return total / wsum


def _gradient_rgb(rng, h, w):
    ys = (np.arange(h, dtype=np.float64) + 0.5)[:, None]
    xs = (np.arange(w, dtype=np.float64) + 0.5)[None, :]
    rgb = np.zeros((h, w, 3))
    for _ in range(3):
        theta = rng.uniform(0.0, 2.0 * np.pi)
        c0 = np.array([rng.uniform(0, 255) for _ in range(3)])
        c1 = np.array([rng.uniform(0, 255) for _ in range(3)])
        proj = xs * np.cos(theta) + ys * np.sin(theta)
        lo, hi = proj.min(), proj.max()
        t = (proj - lo) / max(hi - lo, 1e-12)
        rgb += c0[None, None, :] * (1.0 - t[..., None]) + c1[None, None, :] * t[..., None]
    rgb /= 3.0
    for c in range(3):
        rgb[..., c] += 40.0 * value_noise(rng, h, w)
    for _ in range(20):
        x0 = int(rng.uniform(0, w))
        y0 = int(rng.uniform(0, h))
        rw = int(rng.uniform(w / 32.0, w / 4.0))
        rh = int(rng.uniform(h / 32.0, h / 4.0))
        col = np.array([rng.uniform(0, 255) for _ in range(3)])
        rgb[y0:y0 + rh, x0:x0 + rw, :] = col
    return rgb


def _to_u8(a):
    return np.clip(np.rint(a), 0, 255).astype(np.uint8)


def gradient_image(width, height, seed):
    rng = SplitMix64(seed)
    img = np.empty((height, width, 4), np.uint8)
    img[..., :3] = _to_u8(_gradient_rgb(rng, height, width))
    img[..., 3] = 255
    return img


def alpha_image(width, height, seed):
    rng = SplitMix64(seed)
    img = np.empty((height, width, 4), np.uint8)
    img[..., :3] = _to_u8(_gradient_rgb(rng, height, width))
    ys = (np.arange(height) + 0.5)[:, None] - height / 2.0
    xs = (np.arange(width) + 0.5)[None, :] - width / 2.0
    r = np.sqrt(xs * xs + ys * ys)
    a = 255.0 * (1.0 - r / r.max()) + 40.0 * value_noise(rng, height, width)
    img[..., 3] = _to_u8(a)
    return img


def normal_map_image(width, height, seed, strength=4.0):
    rng = SplitMix64(seed)
    hf = value_noise(rng, height, width)
    gy, gx = np.gradient(hf * 64.0)
    n = np.stack([-gx * strength, -gy * strength, np.ones_like(hf)], axis=-1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    img = np.zeros((height, width, 4), np.uint8)
    img[..., 0] = _to_u8((n[..., 0] * 0.5 + 0.5) * 255.0)
    img[..., 1] = _to_u8((n[..., 1] * 0.5 + 0.5) * 255.0)
    img[..., 3] = 255
    return img


def flat_noise_image(width, height, seed, sigma=3.0):
    rng = SplitMix64(seed)
    col = np.array([rng.uniform(0, 255) for _ in range(3)])
    noise = rng.normal_array(height * width * 3).reshape(height, width, 3) * sigma
    img = np.empty((height, width, 4), np.uint8)
    img[..., :3] = _to_u8(col[None, None, :] + noise)
    img[..., 3] = 255
    return img


# ---- texture batches ----------------------------------------------------------
def batch_spec(count, lo, hi, footprints, seed0, order_seed):
    """(kind, width, height, (bx, by), seed) of a batch of `count` textures:
    sides uniform in lo..hi (independently per side, not block multiples),
    footprint uniform over `footprints`, content by index: two of every four
    gradient + noise + rectangles, one normal map, one flat colour + noise;
    texture i uses seed seed0 + i."""
    rng = SplitMix64(order_seed)
    spec = []
    for i in range(count):
        w = lo + int(rng.uniform() * (hi - lo + 1))
        h = lo + int(rng.uniform() * (hi - lo + 1))
        fp = footprints[int(rng.uniform() * len(footprints))]
        kind = ("gradient", "gradient", "normal", "flat")[i % 4]
        spec.append((kind, w, h, fp, seed0 + i))
    return spec


def make_texture(entry):
    kind, w, h, _, seed = entry
    if kind == "gradient":
        return gradient_image(w, h, seed)
    if kind == "normal":
        return normal_map_image(w, h, seed)
    return flat_noise_image(w, h, seed)


def make_textures(spec, workers=1):
    """The images of a batch_spec, generated by `workers` processes."""
    if workers <= 1:
        return [make_texture(e) for e in spec]
    import multiprocessing as mp
    with mp.get_context("fork").Pool(workers) as pool:
        # largest first, so the pool stays busy to the end
        order = sorted(range(len(spec)), key=lambda i: -spec[i][1] * spec[i][2])
        imgs = pool.map(make_texture, [spec[i] for i in order], chunksize=1)
    out = [None] * len(spec)
    for i, img in zip(order, imgs):
        out[i] = img
    return out
