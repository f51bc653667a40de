"""Pins the oracle's random-number generators and sin/cos (no GPU)."""
import ctypes as C
import math

import numpy as np

import oracle_py as O


def test_tea_published_vector():
    # TEA (Wheeler & Needham 1994), 32 cycles, key 0, plaintext 0 -> 41ea3a0a 94baa940
    v0, v1 = C.c_uint32(0), C.c_uint32(0)
    key = (C.c_uint32 * 4)(0, 0, 0, 0)
    O.lib.orc_tea_keyed(C.byref(v0), C.byref(v1), key, 32)
    assert (v0.value, v1.value) == (0x41EA3A0A, 0x94BAA940)


def test_init_random_is_tea16_with_reference_key():
    # InitRandom (sampling.h:31-46) = 16 TEA rounds with the fixed key
    for a, b in [(0, 0), (1, 2), (65535, 7), (123456, 99)]:
        v0, v1 = C.c_uint32(a), C.c_uint32(b)
        key = (C.c_uint32 * 4)(0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E)
        O.lib.orc_tea_keyed(C.byref(v0), C.byref(v1), key, 16)
        assert O.lib.orc_init_random(a, b, 16) == v0.value


def test_lcg_numerical_recipes_sequence():
    # ranqd1 of Numerical Recipes from idum = 0
    seed = C.c_uint32(0)
    seq = []
    for _ in range(5):
        O.lib.orc_next_random(C.byref(seed))
        seq.append(seed.value)
    assert seq == [0x3C6EF35F, 0x47502932, 0xD1CCF6E9, 0xAAF95334, 0x6252E503]
    seed = C.c_uint32(0x3C6EF35F - 0x3C6EF35F)
    u = O.lib.orc_next_random(C.byref(seed))
    assert u == np.float32((0x3C6EF35F & 0xFFFFFF) / 16777216.0)


def test_pcg32_demo_sequence():
    # pcg32-demo (O'Neill): srandom(42, 54) -> a15c02b7 7b47f409 ba1d3330 83d2f293 bfa4784b cbed606e
    st, inc = C.c_uint64(), C.c_uint64()
    O.lib.orc_pcg32_seed(C.byref(st), C.byref(inc), 42, 54)
    out = [O.lib.orc_pcg32_next(C.byref(st), inc.value) for _ in range(6)]
    assert out == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def test_sin_cos_within_one_ulp():
    xs = np.concatenate([np.linspace(0, 2 * math.pi, 20001), np.linspace(-50, 50, 5001),
                         np.float32([0.0, 1e-20, 3.1415927, 6.2831855, 1.5707964])]).astype(np.float32)
    for x in xs:
        s = np.float32(O.lib.orc_sinf(float(x)))
        c = np.float32(O.lib.orc_cosf(float(x)))
        ts, tc = math.sin(float(x)), math.cos(float(x))
        assert abs(float(s) - ts) <= float(np.spacing(np.float32(abs(ts)))) + 1e-45, x
        assert abs(float(c) - tc) <= float(np.spacing(np.float32(abs(tc)))) + 1e-45, x
