"""Pins of the oracle's MWC64X restatement (RNG.cl:35-156).

Checks that would fail on a plausible transcription mistake (swapped x/c,
wrong constant, skip off by one, wrong float scale):
  - two independent implementations in C (double-and-add vs 128-bit %) agree;
  - a third in Python big integers agrees;
  - skipping k steps equals stepping k times (the defining property);
  - one MWC step is multiplication of the packed value x*A + c by A mod M.
"""
import numpy as np
import pytest

A = 4294883355
M = 18446383549859758079
BASEID = 4077358422479273989


def py_seed(linpos, n):
    v = BASEID * pow(A, linpos * n, M) % M
    return v // A, v % A


def py_step(x, c):
    t = A * x + c
    return t & 0xFFFFFFFF, t >> 32


def test_constants_fit():
    assert M == A * 2**32 - 1
    assert BASEID < M


@pytest.mark.parametrize("linpos,n", [(0, 45), (1, 45), (7, 3), (123456, 45), (465749 + 3, 45),
                                      (5725439 + 7, 45), (2**20 + 5, 3)])
def test_seed_three_ways(orc, linpos, n):
    a = orc.rng_init(linpos, n)
    b = orc.rng_init(linpos, n, fast=True)
    assert tuple(a) == tuple(b)
    assert tuple(int(v) for v in a) == py_seed(linpos, n)


def test_mulmod_powmod_against_python(orc):
    rng = np.random.default_rng(0)
    for _ in range(50):
        a = int(rng.integers(0, 2**63)) % M
        b = int(rng.integers(0, 2**63)) * 2 % M
        assert orc.mulmod(a, b) == a * b % M
        e = int(rng.integers(0, 2**40))
        assert orc.powmod(A, e) == pow(A, e, M)


def test_skip_equals_steps(orc):
    st = orc.rng_init(0, 1)
    for k in range(1, 200):
        orc.rng_step(st)
        assert tuple(st) == tuple(orc.rng_init(k, 1)), k


def test_step_is_multiply_by_A(orc):
    x, c = py_seed(99, 45)
    for _ in range(20):
        v = x * A + c
        x2, c2 = py_step(x, c)
        assert (x2 * A + c2) % M == v * A % M
        x, c = x2, c2


def test_step_and_float_match_python(orc):
    st = orc.rng_init(31337, 45)
    x, c = int(st[0]), int(st[1])
    for _ in range(30):
        f = orc.rng_float(st)
        res = x ^ c
        assert f == np.float32(2.0 ** -32) * np.float32(res)
        x, c = py_step(x, c)
        assert (int(st[0]), int(st[1])) == (x, c)


def test_draws_in_unit_interval_and_uniform(orc):
    d = np.concatenate([orc.rng_draws(lp, 45, 21) for lp in range(2000)])
    assert d.min() >= 0.0 and d.max() <= 1.0
    assert abs(d.mean() - 0.5) < 0.01
