"""Known-answer tests that pin the oracle's integer arithmetic.

PCG32 (uniform_sampler.h:17-57): the published pcg32 demo stream for
pcg32_srandom(42, 54) is a15c02b7 7b47f409 ba1d3330 83d2f293 bfa4784b cbed606e.
Lantern's constructor sets state = seed and discards one output, so
UniformSampler(151, 109) reaches the same state as pcg32_srandom(42, 54):
inc = 2*54+1 = 109, state after seeding = (0*M + 109 + 42) * M + 109.

MurmurHash3 (integrator.cpp:45-73): HashMix is one 4-byte body step of
MurmurHash3_x86_32 and HashFinalize is fmix32; with the length xor
(h ^= 4) they give the published 4-byte digests below.

Philox-4x32-10: the Random123 known-answer vectors.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import binding as ob


def test_pcg32_published_stream():
    out = (C.c_uint32 * 6)()
    ob.lib().orc_pcg32(151, 109, 6, out)
    assert list(out) == [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]


def test_pcg32_float_grid():
    out = np.zeros(4096, np.float32)
    ob.lib().orc_pcg32_float(12345, 7, out.size, out.ctypes.data_as(C.POINTER(C.c_float)))
    u = np.zeros(4096, np.uint32)
    ob.lib().orc_pcg32(12345, 7, u.size, u.ctypes.data_as(C.POINTER(C.c_uint32)))
    # NextFloat = asfloat((u >> 9) | 0x3F800000) - 1: exactly (u >> 9) * 2^-23
    assert np.array_equal(out, (u >> 9).astype(np.float64) * 2.0 ** -23)
    assert out.min() >= 0.0 and out.max() < 1.0


@pytest.mark.parametrize("data,seed,expected", [
    (b"\x00\x00\x00\x00", 0, 0x2362F9DE),
    (b"aaaa", 0x9747B28C, 0x5A97808A),
    (b"abcd", 0x9747B28C, 0xF0478627),
    (b"\xff\xff\xff\xff", 0, 0x76293B50),
    (b"\x21\x43\x65\x87", 0, 0xF55B516B),
    (b"\x21\x43\x65\x87", 0x5082EDEE, 0x2362F9DE),
])
def test_murmur3_published_digests(data, seed, expected):
    L = ob.lib()
    h = L.orc_hash_mix(seed, int.from_bytes(data, "little"))
    assert L.orc_hash_finalize(h ^ 4) == expected


def test_tile_seed_is_mix_mix_finalize():
    L = ob.lib()
    for tile, frame in [(0, 0), (1, 0), (4095, 15), (123, 456)]:
        assert L.orc_tile_seed(tile, frame) == L.orc_hash_finalize(L.orc_hash_mix(L.orc_hash_mix(0, tile), frame))


@pytest.mark.parametrize("ctr,key,expected", [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
])
def test_philox_random123_kat(ctr, key, expected):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    ob.lib().orc_philox4x32_10(c, k, o)
    assert tuple(o) == expected
