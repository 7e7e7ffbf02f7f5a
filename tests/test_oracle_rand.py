"""Pins the oracle's glibc rand() restatement and getstretch
(core/simulator.cu:802-864) against the C library on this machine."""
import ctypes
import ctypes.util

import numpy as np
import pytest

import oracle as orc


def _libc():
    name = ctypes.util.find_library("c")
    if name is None:
        pytest.skip("no C library found")
    libc = ctypes.CDLL(name)
    libc.rand.restype = ctypes.c_int
    libc.srand.argtypes = [ctypes.c_uint]
    return libc


@pytest.mark.parametrize("seed", [1973, 0, 1, 42, 2**31 - 1, 4294967295])
def test_rand_matches_libc(seed):
    libc = _libc()
    libc.srand(seed)
    g = orc.GlibcRand(seed)
    for _ in range(20000):
        assert g.next() == libc.rand()


def test_getstretch_matches_libc_replay():
    # getstretch(s0) = ((sum_{12} rand()/RAND_MAX - 6) * 0.01 + 1) * (double)s0
    libc = _libc()
    s0 = np.float32(0.02)
    libc.srand(1973)
    expect = []
    for _ in range(500):
        acc = 0.0
        for _ in range(12):
            acc += libc.rand() / 2147483647.0
        acc -= 6.0
        expect.append((acc * 0.01 + 1.0) * float(s0))
    got = orc.stretch_init(500, float(s0), 1973)
    assert np.array_equal(got, np.array(expect))


def test_getstretch_distribution():
    # Irwin-Hall: sum of 12 U(0,1) - 6 has mean 0, variance 1, so
    # s = s0 (1 + 0.01 N) has mean s0 and std 0.01 s0.
    s0 = np.float32(1e-3)
    s = orc.stretch_init(200000, float(s0), 1973)
    assert abs(s.mean() / float(s0) - 1.0) < 1e-4
    assert abs(s.std() / float(s0) - 0.01) < 2e-4
    assert s.min() > float(s0) * 0.94 and s.max() < float(s0) * 1.06


def test_s0_is_float_rounded():
    # pd_stretch_limit_s0 is a float (core/parameters.h:123): 0.02 arrives as
    # float(0.02), which differs from the double 0.02.
    g = orc.GlibcRand(7)
    a = g.getstretch(0.02)
    libc = _libc()
    libc.srand(7)
    acc = sum(libc.rand() / 2147483647.0 for _ in range(12)) - 6.0
    assert a == (acc * 0.01 + 1.0) * float(np.float32(0.02))
    assert a != (acc * 0.01 + 1.0) * 0.02
