"""Pins the oracle's collision RNG (the reference uses numba 0.54.1's xoroshiro128+,
abstract_sph_strategy.py:27, base_kernels.py:108-109) with facts independent of its code:

  - SplitMix64 (the seeding step) against its published output sequence for seed 0;
  - xoroshiro128+'s state transition T is linear over GF(2); with the published shift and
    rotation constants its characteristic polynomial is primitive, so T has order exactly
    2^128 - 1.  A wrong constant almost surely breaks this;
  - the jump function equals T^(2^64) (computed here by 64 squarings of the 128x128 matrix);
  - the uniform draw is (x >> 11) * 2^-53 of the generator's sum output.
"""
import numpy as np
import pytest

import oracle

MASK = (1 << 64) - 1
# distinct prime factors of 2^128 - 1 = F0 F1 ... F6 (Fermat numbers)
PRIMES_2_128_MINUS_1 = [3, 5, 17, 257, 65537, 641, 6700417, 274177, 67280421310721]


def test_splitmix64_published_vector():
    x = 0
    out = []
    for _ in range(3):
        v, x = oracle.splitmix64(x)
        out.append(v)
    assert out == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def _state_to_bits(s0, s1):
    v = s0 | (s1 << 64)
    return np.array([(v >> b) & 1 for b in range(128)], dtype=np.int64)


def _transition_matrix():
    """Column b = T applied to basis state e_b (via the oracle's next)."""
    cols = []
    for b in range(128):
        v = 1 << b
        st = np.array([v & MASK, v >> 64], dtype=np.uint64)
        oracle.rng_next(st)
        cols.append(_state_to_bits(int(st[0]), int(st[1])))
    return np.stack(cols, axis=1)


def _matmul(a, b):
    return (a @ b) & 1


def _matpow(m, e):
    result = np.eye(128, dtype=np.int64)
    base = m.copy()
    while e:
        if e & 1:
            result = _matmul(result, base)
        base = _matmul(base, base)
        e >>= 1
    return result


@pytest.fixture(scope="module")
def T():
    return _transition_matrix()


def test_transition_has_full_period(T):
    order = (1 << 128) - 1
    assert np.array_equal(_matpow(T, order), np.eye(128, dtype=np.int64))
    for p in PRIMES_2_128_MINUS_1:
        assert not np.array_equal(_matpow(T, order // p), np.eye(128, dtype=np.int64)), p


def test_jump_is_two_to_the_64_steps(T):
    J = T.copy()
    for _ in range(64):
        J = _matmul(J, J)
    rng = np.random.default_rng(5)
    for _ in range(3):
        s = [int(x) for x in rng.integers(0, 2 ** 63, size=2, dtype=np.int64)]
        st = np.array(s, dtype=np.uint64)
        oracle.rng_jump(st)
        expect = (J @ _state_to_bits(s[0], s[1])) & 1
        assert np.array_equal(_state_to_bits(int(st[0]), int(st[1])), expect)


def test_streams_and_uniform():
    n = 5
    st = oracle.rng_init(16435234, n).reshape(n, 2)
    z, _ = oracle.splitmix64(16435234)
    assert int(st[0, 0]) == z and int(st[0, 1]) == z
    for i in range(1, n):
        prev = st[i - 1].copy()
        oracle.rng_jump(prev)
        assert np.array_equal(prev, st[i])
    s = st[2].copy()
    a, b = int(s[0]), int(s[1])
    u = oracle.rng_uniform(s)
    assert u == (((a + b) & MASK) >> 11) * 2.0 ** -53
    assert 0.0 <= u < 1.0
