"""Pins the deterministic primitives shared by the oracle and the kernels:
Philox4x32-10 against the published known-answer vectors, the FP32 math
against libm in FP64, and the uniform draw range."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import binding

# Random123 known-answer vectors for Philox4x32-10 (Salmon et al., SC'11):
# (counter, key) -> output.
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(ctr, key, want):
    lib = binding.load()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.pmo_philox(c, k, o)
    assert tuple(o) == want


def test_uniform_range_and_mean():
    lib = binding.load()
    u = np.array([lib.pmo_uniform(123, 7, p, 5, d) for p in range(400) for d in range(8)], dtype=np.float64)
    assert u.min() > 0.0 and u.max() <= 1.0
    assert abs(u.mean() - 0.5) < 0.03
    # distinct streams for distinct pixels / passes
    assert lib.pmo_uniform(1, 0, 10, 1, 0) != lib.pmo_uniform(1, 0, 11, 1, 0)
    assert lib.pmo_uniform(1, 0, 10, 1, 0) != lib.pmo_uniform(1, 0, 10, 2, 0)


def _ulp(x):
    return np.spacing(np.float32(abs(x)))


def test_exp_accuracy():
    lib = binding.load()
    xs = np.linspace(-86.0, 88.0, 20001, dtype=np.float32)
    worst = max(abs(lib.pmo_expf(float(x)) - math.exp(float(x))) / math.exp(float(x)) for x in xs)
    assert worst < 2.5e-7  # about 2 ulp
    assert lib.pmo_expf(-100.0) == 0.0
    assert lib.pmo_expf(0.0) == 1.0


def test_sin_cos_accuracy():
    lib = binding.load()
    xs = np.linspace(-20.0, 20.0, 20001, dtype=np.float32)
    es = max(abs(lib.pmo_sinf(float(x)) - math.sin(float(x))) for x in xs)
    ec = max(abs(lib.pmo_cosf(float(x)) - math.cos(float(x))) for x in xs)
    assert es < 2e-7 and ec < 2e-7
    # identities on exactly representable arguments
    assert lib.pmo_sinf(0.0) == 0.0 and lib.pmo_cosf(0.0) == 1.0


def test_atan2_acos_accuracy():
    lib = binding.load()
    rng = np.random.default_rng(0)
    for y, x in rng.uniform(-5, 5, size=(5000, 2)).astype(np.float32):
        assert abs(lib.pmo_atan2f(float(y), float(x)) - math.atan2(float(y), float(x))) < 5e-7
    # quadrant special cases
    assert abs(lib.pmo_atan2f(1.0, 0.0) - math.pi / 2) < 3e-7
    assert abs(lib.pmo_atan2f(0.0, -1.0) - math.pi) < 3e-7
    assert lib.pmo_atan2f(0.0, 0.0) == 0.0
    for x in np.linspace(-1, 1, 2001, dtype=np.float32):
        assert abs(lib.pmo_acosf(float(x)) - math.acos(float(x))) < 5e-7
