"""The counter RNG that replaces the reference's mt19937 (src/component.cpp:5-10).

Pinned by the Random123 known-answer vectors for Philox-4x32-10 (Salmon et al.,
SC'11), for the library's implementation and the oracle's independent one.
"""
import numpy as np
import pytest

import oracle_py as O
import raym0nade_amd as rm

KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_kat_library(ctr, key, expect):
    assert rm.philox(ctr, key) == expect


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_kat_oracle(ctr, key, expect):
    assert O.philox(ctr, key) == expect


def test_library_matches_oracle_on_random_blocks():
    rng = np.random.default_rng(7)
    for _ in range(200):
        c = tuple(int(x) for x in rng.integers(0, 2**32, 4))
        k = tuple(int(x) for x in rng.integers(0, 2**32, 2))
        assert rm.philox(c, k) == O.philox(c, k)


def test_generator_range_and_order():
    # draws lie in the reference's open range [1e-6, 1-1e-6] and follow the
    # (block, lane) order: draw n is word n&3 of block n>>2
    lo, hi = 1.0, 0.0
    for n in range(64):
        u = O.lib().oracle_gen(3, 0x1D1EC700, 5, 9, n)
        lo, hi = min(lo, u), max(hi, u)
        w = O.philox((5, 9, n >> 2, 0), (3, 0x1D1EC700))[n & 3]
        expect = np.float32(1e-6) + (np.float32(1.0 - 1e-6) - np.float32(1e-6)) * np.float32((w >> 8) * 2.0**-24)
        assert np.float32(u) == expect
    assert lo >= np.float32(1e-6) and hi <= np.float32(1 - 1e-6)
