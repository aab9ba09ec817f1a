"""The erfcx tables compiled into the real-space kernel: evaluated on the
CPU exactly as the kernel evaluates them, erfc(x) = exp(-x^2) erfcx(x)
matches libm's erfc (the reference's ERFC, realpart.c:154)."""
import math
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HDR = os.path.join(ROOT, "p3m-standalone_amd", "csrc", "erfcx_coeffs.hpp")


def load(kind):
    text = open(HDR).read()
    K = int(re.search(r"kErfcx%sPieces = (\d+)" % kind, text).group(1))
    D = int(re.search(r"kErfcx%sDegree = (\d+)" % kind, text).group(1))
    body = text.split("kErfcx%sCoef" % kind)[1].split("};")[0]
    vals = [float(v.rstrip("f")) for v in re.findall(r"-?\d\.\d+e[-+]\d+f?|-?\d+\.\d+(?:e[-+]\d+)?", body)]
    return K, D, np.array(vals).reshape(K, D + 1)


def erfcx_eval(x, K, D, c, dt):
    y = dt(1.0) / (dt(1.0) + dt(0.5) * dt(x))
    k = min(K - 1, int(y * K))
    t = (dt(y * dt(K)) - dt(k)) - dt(0.5)
    acc = dt(c[k, D])
    for d in range(D - 1, -1, -1):
        acc = dt(acc * t + dt(c[k, d]))
    return acc


def test_double_tables_match_libm_erfc():
    K, D, c = load("Double")
    worst = 0.0
    for x in np.concatenate([np.linspace(0.0, 6.0, 3001), np.linspace(6.0, 25.0, 501)]):
        e = math.exp(-x * x) * float(erfcx_eval(x, K, D, c, np.float64))
        ref = math.erfc(x)
        if ref > 1e-300:
            worst = max(worst, abs(e - ref) / ref)
    # exp(-x^2) of the rounded square dominates at large x; the kernel
    # splits x^2 exactly, which this check does not emulate
    assert worst < 1e-13


@pytest.mark.parametrize("x", [0.0, 0.3, 1.0, 2.5, 3.0, 4.5])
def test_float_tables(x):
    K, D, c = load("Float")
    e = math.exp(-x * x) * float(erfcx_eval(x, K, D, c, np.float32))
    assert abs(e - math.erfc(x)) <= 1e-6 * math.erfc(x)


def load_single():
    text = open(HDR).read()
    a = float(re.search(r"kErfcxSingleA = ([-+0-9.e]+)f", text).group(1))
    b = float(re.search(r"kErfcxSingleB = ([-+0-9.e]+)f", text).group(1))
    x1 = float(re.search(r"kErfcxSingleX1 = ([-+0-9.e]+)f", text).group(1))
    body = text.split("kErfcxSingleCoef")[1].split("};")[0].split("{")[1]
    c = [np.float32(float(v.strip().rstrip("f"))) for v in body.split(",")]
    return a, b, x1, c


def test_single_float_polynomial_matches_libm_erfc():
    """The one-piece FP32 erfcx of the mixed-precision pair kernel, evaluated
    in float32 as the kernel does (t = y a + b by FMA, Horner), against libm's
    erfc up to the clamp point."""
    a, b, x1, c = load_single()
    worst = 0.0
    for x in np.linspace(0.0, x1, 6001):
        y = np.float32(1.0) / (np.float32(1.0) + np.float32(0.5) * np.float32(x))
        t = np.float32(y * np.float32(a) + np.float32(b))
        acc = np.float32(0.0)
        for cc in reversed(c):
            acc = np.float32(acc * t + cc)
        xf = float(np.float32(x))
        ref = math.erfc(xf) * math.exp(xf * xf)
        worst = max(worst, abs(float(acc) - ref) / ref)
    assert worst < 1e-6
