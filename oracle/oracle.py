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
    """Reference forces as common.c:329-363 computes them: rc = 0.49 L,
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
    # (ewald.c:89-90 with p->mesh = kmax), then sums over the kmax cube.
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


# ---- inhomogeneous error estimate (p3m-common.c:356-719) --------------------

def _sinc(d):
    """sinc of p3m-common.c:71-75."""
    d = np.asarray(d, dtype=np.float64)
    pd = np.pi * d
    out = np.ones_like(d)
    nz = d != 0.0
    out[nz] = np.sin(pd[nz]) / pd[nz]
    return out


def _window(nx, ny, nz, M, cao, power):
    """U (power 1) and U^2 (power 2) of p3m-common.c:370-377."""
    return (_sinc(nx / M) * _sinc(ny / M) * _sinc(nz / M)) ** int(power * cao)


def _wave_numbers(M):
    i = np.arange(M)
    s = np.where(i >= M // 2, i - M, i).astype(np.float64)
    return np.meshgrid(s, s, s, indexing="ij")


def inhomo_kernel_terms(M, L, alpha, cao):
    """Per mesh point n (signed wave numbers, x-major): (n, G, U) with G the
    aliasing-free b / a of p3m-common.c:356-396 and U the window; n = 0
    carries G = 0."""
    n = _wave_numbers(M)
    k2 = (2.0 * np.pi / L) ** 2 * (n[0] ** 2 + n[1] ** 2 + n[2] ** 2)
    zero = k2 == 0.0
    k2s = np.where(zero, 1.0, k2)
    gc = (4.0 * np.pi / k2s) * np.exp(-k2s / (4.0 * alpha * alpha))
    u2 = _window(n[0], n[1], n[2], M, cao, 2)
    G = np.where(zero, 0.0, (u2 * gc * k2s) / (u2 * (k2s * u2)))
    return n, G, gc, u2, zero


def _backward(a):
    return np.fft.ifftn(a) * a.size  # unnormalised, sign +1


def inhomo_k2(M, L, alpha, cao, mc, factorised=True):
    """K^2(r) = sum over the homogeneous and aliasing kernels of
    sum_d Re(K_d(r))^2 (p3m-common.c:490-640).  factorised=False evaluates
    the reference's double sum over (n, n2) literally (tiny M only)."""
    n, G, gc, u2, zero = inhomo_kernel_terms(M, L, alpha, cao)
    K3 = np.zeros((M, M, M))
    k2 = np.where(zero, 0.0, u2 * G - gc)
    for d in range(3):
        K3 += _backward(1j * (-2.0 * np.pi * n[d] / L) * k2).real ** 2
    u = _window(n[0], n[1], n[2], M, cao, 1)
    for mx in range(-mc, mc + 1):
        for my in range(-mc, mc + 1):
            for mz in range(-mc, mc + 1):
                if mx == 0 and my == 0 and mz == 0:
                    continue
                um = _window(n[0] + mx * M, n[1] + my * M, n[2] + mz * M, M, cao, 1)
                um2 = _window(-n[0] - mx * M, -n[1] - my * M, -n[2] - mz * M, M, cao, 1)
                k1 = np.where(zero, 0.0, 2.0 * np.pi * u * um * G)
                kk2 = np.where(zero, 0.0, 2.0 * np.pi * u * um2 * G)
                for d in range(3):
                    if factorised:
                        S = np.sum((n[d] / L) * k1)
                        Kd = -(n[d] / L) * kk2 * S
                    else:
                        a = ((n[d] / L) * k1).ravel()
                        b = ((n[d] / L) * kk2).ravel()
                        Kd = (-np.outer(a, b)).sum(axis=0).reshape(M, M, M)
                    K3 += _backward(1j * Kd).real ** 2
    return K3


def _gather(sysd, mesh_real, M, cao):
    """sum_stencil q_i^2 W mesh (collect_rms_nocf, charge-assign.c:1655-1702)
    with the mesh's own spacing and order."""
    L = sysd["L"]
    Hi = M / L
    out = np.zeros(len(sysd["q"]))
    for i in range(len(out)):
        idx, w = [], []
        for k in "xyz":
            _, base, arg = locate(float(sysd[k][i]), Hi, cao, 0, M)
            xq = lib().orc_quantized_offset(arg)
            idx.append((base + np.arange(cao)) % M)
            w.append(np.array([weight(j, xq, cao) for j in range(cao)]))
        W = np.einsum("a,b,c->abc", *w)
        out[i] = sysd["q"][i] ** 2 * np.sum(mesh_real[np.ix_(*idx)] * W)
    return out


def error_inhomo(sysd, alpha, M, cao, mc, uniform=False):
    """Generic_error_estimate_inhomo (p3m-common.c:437-719) with the gather
    on the estimate's own mesh and order: sqrt(sum_i rms_i / N)."""
    q = np.asarray(sysd["q"], dtype=np.float64)
    if uniform:
        Q = np.full((M, M, M), np.sum(q ** 2) / M ** 3, dtype=np.complex128)
    else:
        sq = dict(sysd)
        sq["q"] = q ** 2
        Q = assign_mesh(0, sq, M, cao).view(np.complex128).reshape(M, M, M)
    K3h = np.fft.fftn(inhomo_k2(M, sysd["L"], alpha, cao, mc))
    Qh = np.fft.fftn(Q)
    dens = _backward(Qh.real * K3h.real + 1j * (Qh.imag * K3h.imag))
    rms = _gather(sysd, dens.real, M, cao)
    return float(np.sqrt(np.sum(rms) / len(q)))


def error_map(sysd, fk, fk_ref, M, cao):
    """Error_map (p3m-common.c:402-435): |fk_ref - fk|^2 per particle
    assigned with order cao; real (M, M, M)."""
    d = dict(sysd)
    d["q"] = np.sum((np.asarray(fk_ref) - np.asarray(fk)) ** 2, axis=0)
    return assign_mesh(0, d, M, cao)[0::2].reshape(M, M, M)
