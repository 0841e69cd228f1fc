"""Integer known answers for the RNG of the reference kernels and host.

wang_hash / xor_shift (kernels/legacy/udpt.cl:1131-1147), the per-pixel
seed (udpt.cl:175-180), `seed / (float) UINT_MAX` (udpt.cl:185), and the
frame seed stream std::mt19937 (src/RendererCore.cpp:61-63, 268).
The expected values are computed independently here (numpy's RandomState
is the reference MT19937 with init_genrand seeding; the hash constants are
checked on values worked out by hand).
"""
import numpy as np

from oracle import pyoracle as O


def test_xor_shift_known_answers():
    L = O.lib()
    assert L.or_xor_shift(1) == 270369           # 1 ^ 1<<13 = 8193; ^ >>17: 8193; ^ <<5: 270369
    # three steps from 1, done in numpy uint32 arithmetic
    s = np.uint32(1)
    for _ in range(3):
        s ^= np.uint32(s << np.uint32(13))
        s ^= np.uint32(s >> np.uint32(17))
        s ^= np.uint32(s << np.uint32(5))
    x = 1
    for _ in range(3):
        x = L.or_xor_shift(x)
    assert x == int(s)


def test_wang_hash_known_answers():
    L = O.lib()
    assert L.or_wang_hash(0) == 3232319850
    assert L.or_wang_hash(1) == 663891101


def test_pixel_seed_and_first_draws():
    L = O.lib()
    rand = 3499211612                              # first std::mt19937(5489) output
    assert O.mt19937(5489, 1)[0] == rand
    seed = L.or_pixel_seed(rand, 0, 0, 512)
    assert seed == 2371417505
    s1 = L.or_xor_shift(seed)
    s2 = L.or_xor_shift(s1)
    assert (s1, s2) == (4001306295, 2576870531)
    assert L.or_u2f(s1) == np.float32(0.93162674)
    assert L.or_u2f(s2) == np.float32(0.59997445)


def test_u2f_is_division_by_2_pow_32():
    L = O.lib()
    for s in (1, 2, 12345, 0x7fffffff, 0xffffffff):
        assert L.or_u2f(s) == np.float32(np.float32(s) / np.float32(4294967296.0))
    assert L.or_u2f(0xffffffff) == 1.0               # (float)UINT_MAX rounds up to 2^32


def test_mt19937_matches_standard():
    # std::mt19937 default seed: the 10000th output is mandated by the C++ standard
    assert O.mt19937(5489, 10000)[-1] == 4123659995
    for seed in (1, 2024, 0xdeadbeef):
        ref = np.random.RandomState(seed).randint(0, 2 ** 32, size=2000, dtype=np.uint64).astype(np.uint32)
        assert np.array_equal(O.mt19937(seed, 2000), ref)


M32 = (1 << 32) - 1


def _unxorshift_right(y, k):
    x = y
    for _ in range(32 // k + 1):
        x = y ^ (x >> k)
    return x & M32


def _wang_inverse(h):
    """Inverse of wang_hash, step by step in reverse (each step is a bijection)."""
    x = _unxorshift_right(h, 15)
    x = (x * pow(0x27d4eb2d, -1, 1 << 32)) & M32
    x = _unxorshift_right(x, 4)
    x = (x * pow(9, -1, 1 << 32)) & M32
    x = _unxorshift_right(x, 16)
    return x ^ 61


def test_zero_hash_rehash():
    """A pixel whose hashed seed is 0 is hashed once more (udpt.cl:179-180)."""
    L = O.lib()
    x0 = _wang_inverse(0)
    assert L.or_wang_hash(x0) == 0
    W = 16                                    # pixel (0,0): multiplier W + 1 = 17 (odd)
    rand = (x0 * pow(17, -1, 1 << 32)) & M32
    assert L.or_pixel_seed(rand, 0, 0, W) == L.or_wang_hash(0) == 3232319850
