"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/ and by bench.py's CPU-baseline
leg as the checker.  The product library never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    subprocess.check_call(["make", "-s", "-B", "-C", _HERE, "liboracle.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        sig = {
            "orc_weight": (ctypes.c_double, [ctypes.c_int, ctypes.c_double, ctypes.c_int]),
            "orc_weight_d": (ctypes.c_double, [ctypes.c_int, ctypes.c_double, ctypes.c_int]),
            "orc_locate": (ctypes.c_int, [ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, _ip, _ip]),
            "orc_quantized_offset": (ctypes.c_double, [ctypes.c_int]),
            "orc_fft1d": (None, [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "orc_fft3d": (None, [_dp, ctypes.c_int, ctypes.c_int]),
            "orc_rfft3d": (None, [_dp, ctypes.c_int]),
            "orc_irfft3d": (None, [_dp, ctypes.c_int]),
            "orc_dn": (None, [ctypes.c_int, _dp]),
            "orc_nshift": (None, [ctypes.c_int, _dp]),
            "orc_influence": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_int, _dp]),
            "orc_self_force_coeffs": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp]),
            "orc_subtract_self_forces": (None, [ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp,
                                                _dp, _dp, _dp]),
            "orc_kspace_force": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp,
                                                _dp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]),
            "orc_assign_mesh": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp,
                                               _dp, ctypes.c_int, ctypes.c_int, _dp]),
            "orc_realteil": (None, [ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, _dp, ctypes.c_double,
                                    ctypes.c_double, _dp, _dp]),
            "orc_realteil_cells": (None, [ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, _dp,
                                          ctypes.c_double, ctypes.c_double, _dp, _dp, ctypes.c_int,
                                          ctypes.c_int]),
            "orc_ewald_influence": (None, [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp]),
            "orc_ewald_kspace": (None, [ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, _dp, ctypes.c_double,
                                        ctypes.c_int, _dp, _dp]),
            "orc_ewald_error_r": (ctypes.c_double, [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                    ctypes.c_double, ctypes.c_double]),
            "orc_ewald_error_k": (ctypes.c_double, [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                    ctypes.c_int, ctypes.c_double]),
            "orc_ewald_optimal_alpha": (ctypes.c_double, [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                          ctypes.c_double, ctypes.c_int]),
            "orc_realspace_error": (ctypes.c_double, [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                      ctypes.c_double, ctypes.c_double]),
            "orc_kspace_error": (ctypes.c_double, [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                                   ctypes.c_double, ctypes.c_int, _dp]),
            "orc_table_lookup": (ctypes.c_double, [_dp, ctypes.c_double]),
            "orc_force_error": (ctypes.c_double, [ctypes.c_int, _dp, _dp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def weight(j, x, cao):
    return lib().orc_weight(j, x, cao)


def weight_d(j, x, cao):
    return lib().orc_weight_d(j, x, cao)


def locate(r, Hi, cao, ii, M):
    base, arg = ctypes.c_int(), ctypes.c_int()
    nmp = lib().orc_locate(r, Hi, cao, ii, M, ctypes.byref(base), ctypes.byref(arg))
    return nmp, base.value, arg.value


def fft3d(c, sign):
    """c: complex128 (M,M,M) -> new array, unnormalised, sign -1 forward."""
    a = np.ascontiguousarray(c, dtype=np.complex128).copy()
    lib().orc_fft3d(a.ctypes.data_as(_dp), a.shape[0], sign)
    return a


def rfft3d(r):
    M = r.shape[0]
    buf = np.zeros((M, M, M + 2))
    buf[:, :, :M] = r
    lib().orc_rfft3d(_p(buf), M)
    return buf.view(np.complex128).reshape(M, M, M // 2 + 1)


def irfft3d(c):
    M = c.shape[0]
    buf = np.ascontiguousarray(c, dtype=np.complex128).copy()
    lib().orc_irfft3d(buf.ctypes.data_as(_dp), M)
    return buf.view(np.float64).reshape(M, M, M + 2)[:, :, :M].copy()


def dn(M):
    out = np.zeros(M)
    lib().orc_dn(M, _p(out))
    return out


def nshift(M):
    out = np.zeros(M)
    lib().orc_nshift(M, _p(out))
    return out


def influence(method, M, cao, L, alpha, mc=1):
    G = np.zeros(M * M * M)
    rc = lib().orc_influence(method, M, cao, L, alpha, mc, _p(G))
    if rc != 0:
        raise ValueError("bad method")
    return G


def self_force_coeffs(M, cao, L, G):
    b = np.zeros(81)
    lib().orc_self_force_coeffs(M, cao, L, _p(_c(G)), _p(b))
    return b


def assign_mesh(method, sysd, M, cao):
    x, y, z, q = (_c(sysd[k]) for k in "xyzq")
    mesh = np.zeros(2 * M * M * M)
    lib().orc_assign_mesh(method, len(x), sysd["L"], _p(x), _p(y), _p(z), _p(q), M, cao, _p(mesh))
    return mesh


def kspace_force(method, sysd, M, cao, alpha, G=None, b=None, fk=None, f=None):
    """One Kspace_force call; returns (fk, f) with fk accumulated and, for
    methods 2/7, the self force subtracted from f."""
    x, y, z, q = (_c(sysd[k]) for k in "xyzq")
    N, L = len(x), float(sysd["L"])
    if G is None:
        G = influence(method, M, cao, L, alpha)
    if b is None and method in (2, 7) and alpha != 0.0:
        b = self_force_coeffs(M, cao, L, G)
    fk = np.zeros(3 * N) if fk is None else _c(fk).copy()
    f = np.zeros(3 * N) if f is None else _c(f).copy()
    bp = _p(_c(b)) if b is not None else None
    rc = lib().orc_kspace_force(method, N, L, _p(x), _p(y), _p(z), _p(q), M, cao, _p(_c(G)), bp,
                                _p(fk), _p(f))
    if rc != 0:
        raise ValueError("orc_kspace_force failed")
    return fk.reshape(3, N), f.reshape(3, N)


def realteil(sysd, rcut, alpha, cells=False, i_range=None):
    x, y, z, q = (_c(sysd[k]) for k in "xyzq")
    N, L = len(x), float(sysd["L"])
    fr = np.zeros(3 * N)
    en = np.zeros(1)
    if cells:
        b, e = (0, N) if i_range is None else i_range
        lib().orc_realteil_cells(N, L, _p(x), _p(y), _p(z), _p(q), rcut, alpha, _p(fr), _p(en), b, e)
    else:
        lib().orc_realteil(N, L, _p(x), _p(y), _p(z), _p(q), rcut, alpha, _p(fr), _p(en))
    return fr.reshape(3, N), en[0]


def ewald_kspace(sysd, alpha, kmax):
    x, y, z, q = (_c(sysd[k]) for k in "xyzq")
    N, L = len(x), float(sysd["L"])
    G = np.zeros((kmax + 1) ** 3)
    lib().orc_ewald_influence(kmax, L, alpha, _p(G))
    fk = np.zeros(3 * N)
    lib().orc_ewald_kspace(N, L, _p(x), _p(y), _p(z), _p(q), alpha, kmax, _p(G), _p(fk))
    return fk.reshape(3, N)


def ewald_reference(sysd, precision=1e-8):
    """Reference forces as common.c:314-348 computes them: rc = 0.49 L,
    bisected alpha, kmax grown by 2 until the estimate <= precision.  Returns
    (f_total, f_kspace, alpha, kmax)."""
    L, N = float(sysd["L"]), len(sysd["x"])
    q2 = float(np.sum(np.asarray(sysd["q"]) ** 2))
    rcut = 0.49 * L
    kmax = 2
    while True:
        alpha = lib().orc_ewald_optimal_alpha(N, L, q2, rcut, kmax)
        err = np.hypot(lib().orc_ewald_error_r(N, L, q2, rcut, alpha),
                       lib().orc_ewald_error_k(N, L, q2, kmax, alpha))
        if err <= precision:
            break
        kmax += 2
    # The reference's influence function fills n^2 <= (kmax-1)^2 only
    # (ewald.c:74-75 with p->mesh = kmax), then sums over the kmax cube.
    fk = ewald_kspace(sysd, alpha, kmax - 1)
    fr, _ = realteil(sysd, rcut, alpha, cells=N > 2000)
    return fr + fk, fk, alpha, kmax


def realspace_error(N, L, q2, rcut, alpha):
    return lib().orc_realspace_error(N, L, q2, rcut, alpha)


def kspace_error(method, N, L, q2, M, cao, alpha, mc=1, table=None):
    tp = _p(_c(table)) if table is not None else None
    return lib().orc_kspace_error(method, N, L, q2, M, cao, alpha, mc, tp)


def rms(f, fref):
    f = np.asarray(f).reshape(3, -1)
    fref = np.asarray(fref).reshape(3, -1)
    return float(np.sqrt(np.sum((f - fref) ** 2) / f.shape[1]))


# Candidate meshes of the reference's tuner (tuning.c:38), ascending.
TUNE_MESHES = (8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 30, 32, 36, 40, 42, 44, 48, 50, 52, 54,
               56, 60, 64, 66, 70, 72, 78, 80, 84, 88, 90, 96, 98, 100, 104, 108, 110, 112, 120,
               126, 128, 130, 132, 140, 144, 150, 154, 156, 160, 162, 168, 176, 180, 182, 192,
               196, 198, 200, 208, 210, 216, 220, 224, 234, 240, 242, 250, 252, 256, 260, 264,
               270, 280, 288, 294, 300)


def tune_alpha(N, L, q2, rcut, precision):
    """alpha of the tuner (tuning.c:293): the real-space estimate
    (realpart.c:343-346) equal to precision / sqrt(2)."""
    V = L ** 3
    return np.sqrt(-np.log(precision * np.sqrt(N * rcut * V) / (2.0 * np.sqrt(2.0) * q2))) / rcut


def total_error(method, N, L, q2, M, cao, alpha, rcut):
    """Error of the method table: sqrt(real^2 + k^2) (e.g. p3m-ik.c:249-261)."""
    r = realspace_error(N, L, q2, rcut, alpha)
    k = kspace_error(method, N, L, q2, M, cao, alpha)
    return float(np.sqrt(r * r + k * k))


def tune_scan(method, N, L, q2, rcut, cao, precision, timer, error=None):
    """Tune (tuning.c:250-370) with an injected timer(mesh, cao) -> seconds
    (the summed minimum stage time the reference compares) and, optionally,
    an injected error(mesh, cao, alpha) (default: the method's estimate).
    Meshes ascend through TUNE_MESHES; orders descend from 7 (or only `cao`
    when nonzero) while the estimate meets the precision, later meshes from
    one below the last order that met it; a mesh whose order-2 time already
    exceeds the best ends the scan.  Returns (mesh, cao, alpha, error, time)
    of the winner, or None."""
    error = error or (lambda M, c, a: total_error(method, N, L, q2, M, c, a, rcut))
    alpha = tune_alpha(N, L, q2, rcut, precision)
    lo, hi = (cao, cao) if cao else (2, 7)
    last, found = hi, False
    best, best_t, best_mc = None, float("inf"), (float("inf"), 8)
    for M in TUNE_MESHES:
        if found:
            if last <= lo:
                break
            first = last - 1
        else:
            first = last = hi
        if timer(M, lo) > best_t:
            break
        for c in range(first, lo - 1, -1):
            err = error(M, c, alpha)
            if not err <= precision:
                break
            last, found = c, True
            if M >= best_mc[0] and c >= best_mc[1]:
                break
            t = timer(M, c)
            if t < best_t:
                best, best_t, best_mc = (M, c, alpha, err, t), t, (M, c)
    return best
