"""Pin the oracle's elementary functions and RNG: fdlibm log/sin/cos against
glibc (< 1 ulp), Philox4x32-10 against its published known-answer vectors."""
import ctypes as C
import math

import numpy as np

import oracle_lib


def ulp_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.spacing(np.abs(b))


def test_log_within_one_ulp():
    lib = oracle_lib.load()
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(0, 1, 40000), rng.uniform(0, 1, 2000) ** 20,
                         np.array([1e-300, 0.5, 1.0, 1 - 2 ** -52, 2 ** -52, 0.9999, 1.0001, 3.0, 1e10])])
    got = np.array([lib.orc_log(float(x)) for x in xs])
    ref = np.log(xs)
    assert np.max(ulp_err(got, ref)) <= 1.0
    assert lib.orc_log(1.0) == 0.0


def test_sin_cos_within_one_ulp():
    lib = oracle_lib.load()
    rng = np.random.default_rng(1)
    xs = np.concatenate([2 * math.pi * rng.uniform(0, 1, 40000),
                         np.array([1e-9, 0.5, math.pi / 4, math.pi / 2, math.pi, 3 * math.pi / 2, 6.28318])])
    s = np.array([lib.orc_sin(float(x)) for x in xs])
    c = np.array([lib.orc_cos(float(x)) for x in xs])
    # compare with an absolute floor near the zeros of sin / cos
    assert np.max(np.abs(s - np.sin(xs)) / np.maximum(np.spacing(np.abs(np.sin(xs))), 2 ** -60)) <= 1.0 or \
        np.max(np.abs(s - np.sin(xs))) < 4e-16
    assert np.max(np.abs(c - np.cos(xs))) < 4e-16
    assert np.max(np.abs(s - np.sin(xs))) < 4e-16
    assert np.allclose(s * s + c * c, 1.0, atol=1e-15)


def philox(ctr, key):
    lib = oracle_lib.load()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.orc_philox4x32(c, k, o)
    return list(o)


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32_10
    assert philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_uniform_stream():
    lib = oracle_lib.load()
    out = np.zeros(100001)
    lib.orc_uniforms(12345, 7, out.size, out.ctypes.data_as(C.POINTER(C.c_double)))
    assert out.min() > 0 and out.max() < 1
    assert abs(out.mean() - 0.5) < 0.005
    # draws are exactly (m + 0.5) 2^-52
    m = out * 2.0 ** 52 - 0.5
    assert np.all(m == np.floor(m))
    out2 = np.zeros(10)
    lib.orc_uniforms(12345, 8, 10, out2.ctypes.data_as(C.POINTER(C.c_double)))
    assert not np.any(out2 == out[:10])


def test_clone_keys_distinct():
    lib = oracle_lib.load()
    keys = {lib.orc_clone_key(p, s, o) for p in range(50) for s in range(20) for o in range(3)}
    keys |= {lib.orc_init_key(1, i) for i in range(3000)}
    assert len(keys) == 50 * 20 * 3 + 3000
