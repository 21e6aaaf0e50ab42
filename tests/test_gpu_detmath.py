"""The device's elementary functions of the collision path (pow, sin, cos,
acos, atan2, and log for the sources) against the oracle's, which compile the
same source (start_amd/csrc/detmath.h): every result equal to the bit, over
random arguments on the ranges the step uses and the special values."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import pyoracle as O
from start_amd import engine

from ._parity import log

pytestmark = pytest.mark.gpu
N = 1_000_000


def same_bits(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def args(fn, rng):
    if fn == "sincos":
        x = np.concatenate([rng.uniform(-7.0, 7.0, N), rng.uniform(0.0, 2 * np.pi, N),
                            np.pi / 4 * np.arange(-9, 10), [0.0, -0.0, 1e-300, np.pi, 2 * np.pi, 31.0]])
        return x, None
    if fn == "pow":
        x = np.concatenate([rng.uniform(0.0, 1.0, N), 10.0 ** rng.uniform(-12, 12, N),
                            [0.0, 0.0, np.inf, 1.0, 2.0, 1e-300, 3.0]])
        y = np.concatenate([rng.choice([1 / 1.4, 1 / 1.6, 0.5], N), rng.uniform(-0.6, 0.6, N),
                            [0.3, -0.3, 0.3, 0.7, 10.0, 0.7, 0.0]])
        return x, y
    if fn == "acos":
        x = np.concatenate([rng.uniform(-1.0, 1.0, N), 1.0 - 10.0 ** -rng.uniform(1, 16, N // 4),
                            [-1.0, 1.0, 0.0, -0.0, 0.5, -0.5]])
        return x, None
    if fn == "atan2":
        x = np.concatenate([rng.standard_normal(N) * 10.0 ** rng.uniform(-5, 5, N),
                            [1.0, 1.0, -1.0, -1.0, 0.0, 0.0, -0.0, -0.0]])
        y = np.concatenate([rng.standard_normal(N) * 10.0 ** rng.uniform(-5, 5, N),
                            [0.0, -0.0, 0.0, -0.0, 1.0, -1.0, 1.0, -1.0]])
        return x, y
    x = np.concatenate([1.0 - rng.uniform(0.0, 1.0, N), 10.0 ** rng.uniform(-300, 300, N),
                        [1.0, 2.0, 0.5, 5e-324, np.inf, 0.0]])
    return x, None


@pytest.mark.parametrize("fn", ["pow", "sincos", "acos", "atan2", "log"])
def test_device_math_equals_oracle_math(fn):
    rng = np.random.default_rng(100 + ["pow", "sincos", "acos", "atan2", "log"].index(fn))
    x, y = args(fn, rng)
    log(f"math {fn}: {len(x)} arguments")
    dev = engine.math_batch(fn, x, y)
    ora = O.math_eval(fn, x, y, libm=False)
    if fn == "sincos":
        assert same_bits(dev[0], ora[0]) and same_bits(dev[1], ora[1])
    else:
        assert same_bits(dev, ora)


def test_math_batch_rejects_bad_arguments():
    import ctypes as C

    from start_amd import _lib
    L = _lib.lib()
    x = np.zeros(4)
    p = x.ctypes.data_as(C.POINTER(C.c_double))
    err_arg = -1  # START_ERR_ARG (include/start_gpu.h)
    assert L.start_math_batch(6, 4, p, p, p, p) == err_arg
    assert L.start_math_batch(-1, 4, p, p, p, p) == err_arg
    assert L.start_math_batch(0, -1, p, p, p, p) == err_arg
    assert L.start_math_batch(0, 0, None, None, None, None) == 0  # nothing to do
