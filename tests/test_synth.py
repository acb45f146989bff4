"""The synthetic workload generators are deterministic (splitmix64 seeded);
their outputs are pinned by SHA-256 so that benchmark and parity inputs are
the same images everywhere."""
import hashlib

import pytest

from astc_implementation_amd import synth

GOLDEN = {
    "gradient_512_seed0": "9c4f2cd964f250bc60a86920b8fb11bf978c4a04fbde978e6db963c3155d4cc4",
    "alpha_256_seed2": "d020a07a855aae10e27f40800fbcf335b411aa7c570728228ad62ac1101d9bdf",
    "normal_128_seed5": "7036090953f82c3e1ebeb43ab4afa01a8938a711404293c8ba3ffc219788ab9e",
    "flat_128_seed7": "f5c7ecb3796576418311ed74194a8ec33c9664a2a2ae24a04e3c2e6fe351423b",
}
GEN = {
    "gradient_512_seed0": (synth.gradient_image, (512, 512, 0)),
    "alpha_256_seed2": (synth.alpha_image, (256, 256, 2)),
    "normal_128_seed5": (synth.normal_map_image, (128, 128, 5)),
    "flat_128_seed7": (synth.flat_noise_image, (128, 128, 7)),
}


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_generator_digest(name):
    f, args = GEN[name]
    img = f(*args)
    assert img.shape == (args[1], args[0], 4)
    assert hashlib.sha256(img.tobytes()).hexdigest() == GOLDEN[name]


def test_splitmix64_known_values():
    # splitmix64 reference outputs for seed 0 (Steele, Lea, Flood 2014)
    r = synth.SplitMix64(0)
    assert [r.next_u64() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    r1 = synth.SplitMix64(0)
    r2 = synth.SplitMix64(0)
    a = r1.uniform_array(5)
    b = [r2.uniform() for _ in range(5)]
    assert list(a) == b
