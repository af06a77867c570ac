"""The device math library's FP64 exp, log, pow and tanh against the host's libm (Python's math
module calls it), in ulps of the host's result, over the arguments the time step gives them:
exp of -tau, -3020 / T, the saturation curves' exponents and the closures' damping terms; log of
temperatures; pow for the closures' integer powers, the precipitation's Rain^(7/9) and Rain^(4/9),
T^(-1/2) and the equilibrium's 10^(14.82 - 4705 / T); tanh of the SST blending arguments.

The parity bars that say "twice the oracle's one-ulp spread" rest on these functions differing from
the host's by at most an ulp or two: measured here, not assumed."""
import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_D = C.POINTER(C.c_double)
MAX_ULP = 2.0


def _device(fn, x, y=None):
    from atjup_amd import lib
    L = lib.load()
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(np.zeros_like(x) if y is None else y, dtype=np.float64)
    out = np.empty_like(x)
    rc = L.atjup_hip_libm_probe(fn, x.ctypes.data_as(_D), y.ctypes.data_as(_D), out.ctypes.data_as(_D), x.size)
    assert rc == 0, L.atjup_hip_last_error_string()
    return out


def _ulps(dev, host):
    return np.array([abs(a - b) / math.ulp(b) if b != 0 else (0.0 if a == 0 else np.inf)
                     for a, b in zip(dev.tolist(), host.tolist())])


def _check(name, dev, host):
    u = _ulps(dev, host)
    print(f"{name}: n={u.size} exact {np.mean(u == 0):.4f}, max {u.max():.2f} ulp, mean {u.mean():.3f}")
    assert u.max() <= MAX_ULP, name


rng = np.random.default_rng(2024)


def test_exp():
    x = np.concatenate([rng.uniform(-60.0, 30.0, 100000), rng.uniform(-1.0, 1.0, 50000),
                        -3020.0 / rng.uniform(100.0, 2000.0, 50000)])
    _check("exp", _device(0, x), np.array([math.exp(v) for v in x]))


def test_log():
    x = np.concatenate([rng.uniform(30.0, 3000.0, 100000), 10.0 ** rng.uniform(-10.0, 10.0, 100000)])
    _check("log", _device(1, x), np.array([math.log(v) for v in x]))


def test_pow():
    base = 10.0 ** rng.uniform(-12.0, 3.0, 150000)
    expo = rng.choice([2.0, 3.0, 4.0, 7.0 / 9.0, 4.0 / 9.0, -0.5], size=base.size)
    t = rng.uniform(100.0, 2000.0, 50000)
    x = np.concatenate([base, np.full(t.size, 10.0)])
    y = np.concatenate([expo, 14.82 - 4705.0 / t])
    _check("pow", _device(2, x, y), np.array([math.pow(a, b) for a, b in zip(x, y)]))


def test_tanh():
    x = np.concatenate([rng.uniform(0.0, 20.0, 100000), rng.uniform(-3.0, 3.0, 100000)])
    _check("tanh", _device(3, x), np.array([math.tanh(v) for v in x]))


def test_probe_refuses_bad_arguments():
    from atjup_amd import lib
    L = lib.load()
    x = np.ones(4)
    out = np.empty(4)
    assert L.atjup_hip_libm_probe(7, x.ctypes.data_as(_D), x.ctypes.data_as(_D), out.ctypes.data_as(_D), 4) != 0
    assert L.atjup_hip_libm_probe(0, x.ctypes.data_as(_D), x.ctypes.data_as(_D), out.ctypes.data_as(_D), 0) != 0
