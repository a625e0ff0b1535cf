"""Pins the oracle's scalar primitives (oracle/oracle_math.h) on the CPU.

Philox4x32-10 against the published known-answer vectors of Salmon et al.
(Random123 kat_vectors); expf / exp / sinf / cosf against libm in higher
precision with ulp bounds; the uniform mapping against its definition.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle_py as O


@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(ctr, key, expect):
    assert [int(x) for x in O.philox4x32_10(ctr, key)] == expect


def test_uniform_stream_definition():
    seed = 0x0123456789ABCDEF
    u = O.uniforms(seed, pixel=77, stage=3, n=12)
    # draw k of the stream = word k%4 of Philox(counter=(k//4, pixel, stage, 0), key=seed)
    words = []
    for blk in range(3):
        words += list(O.philox4x32_10([blk, 77, 3, 0], [seed & 0xffffffff, seed >> 32]))
    expect = np.array([((int(w) >> 8) + 1) * 2.0 ** -24 for w in words], np.float32)
    assert np.array_equal(u, expect)
    assert np.all(u > 0) and np.all(u <= 1)
    # independent streams for other pixels / stages
    assert not np.array_equal(u, O.uniforms(seed, 78, 3, 12))
    assert not np.array_equal(u, O.uniforms(seed, 77, 4, 12))


def _ulp32(x):
    return np.spacing(np.abs(np.float32(x))).astype(np.float64)


def test_expf_ulp_bound():
    L = O.lib()
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-87.3, 88.7, 20000), rng.uniform(-1, 1, 5000), rng.uniform(-20, 0, 5000)])
    xs = xs.astype(np.float32)
    got = np.array([L.orc_expf(float(x)) for x in xs], np.float32)
    ref = np.exp(xs.astype(np.longdouble))
    err = np.abs(got.astype(np.longdouble) - ref) / _ulp32(ref.astype(np.float32))
    assert float(err.max()) <= 1.0


def test_expf_special_values():
    L = O.lib()
    assert L.orc_expf(0.0) == 1.0
    assert math.isinf(L.orc_expf(89.0)) and L.orc_expf(89.0) > 0
    assert L.orc_expf(-104.0) == 0.0
    assert math.isnan(L.orc_expf(float("nan")))
    # subnormal results stay within 1 ulp of the true value
    for x in (-88.0, -95.5, -100.0, -103.0):
        r = float(np.exp(np.longdouble(x)))
        assert abs(L.orc_expf(x) - r) <= np.spacing(np.float32(r))


def test_exp_double_bound():
    L = O.lib()
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-700, 700, 5000), rng.uniform(-2, 2, 5000)])
    for x in xs:
        got = L.orc_exp(float(x))
        ref = np.exp(np.longdouble(x))
        assert abs(np.longdouble(got) - ref) <= 2 * np.spacing(np.float64(ref))


def test_sin_cos_bounds():
    L = O.lib()
    rng = np.random.default_rng(2)
    xs = np.concatenate([rng.uniform(-0.2, 0.2, 5000), rng.uniform(-10, 10, 5000)]).astype(np.float32)
    for f, ref_f in ((L.orc_sinf, np.sin), (L.orc_cosf, np.cos)):
        got = np.array([f(float(x)) for x in xs], np.float64)
        ref = ref_f(xs.astype(np.longdouble)).astype(np.float64)
        err = np.abs(got - ref) / np.maximum(_ulp32(ref), np.spacing(np.float32(1e-30)))
        assert err.max() <= 2.0, err.max()
    assert L.orc_sinf(0.0) == 0.0 and L.orc_cosf(0.0) == 1.0
