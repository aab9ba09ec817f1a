"""GPU parity: the HIP library against the oracle, through the C ABI.

FP64 bar: global relative difference ||F_gpu - F_oracle|| / ||F_oracle||
<= 1e-12 (the oracle restates the reference's FP64 arithmetic; what remains
is summation order and ~1-ulp function differences).
"""
import ctypes

import numpy as np
import pytest

from conftest import error_table_row, fold, load_fixture, random_system, ref_kspace_error, rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12

CASES = [
    # method, N, L, M, cao, alpha
    (0, 600, 10.0, 16, 5, 1.1),
    (1, 600, 10.0, 16, 5, 1.1),
    (2, 600, 10.0, 16, 7, 1.1),
    (3, 600, 10.0, 16, 7, 1.1),
    (6, 600, 10.0, 16, 4, 1.1),
    (7, 600, 10.0, 16, 7, 1.1),
    (0, 500, 12.0, 24, 3, 0.9),
    (1, 500, 12.0, 20, 2, 0.9),
    (2, 500, 12.0, 18, 6, 0.9),
    (3, 500, 12.0, 22, 1, 0.9),
    (6, 500, 12.0, 32, 6, 0.9),
    (7, 500, 12.0, 28, 3, 0.9),
    (3, 2000, 20.0, 32, 7, 0.8),
    (1, 2000, 20.0, 32, 5, 0.8),
]


@pytest.mark.parametrize("method,n,L,M,cao,alpha", CASES)
def test_kspace_force_fp64(pkg, orc, method, n, L, M, cao, alpha):
    s = random_system(n, L, seed=method * 100 + M)
    system = pkg.HostSystem.from_dict(s)
    solver = pkg.Solver(method, system, pkg.parameters(M, cao, alpha))
    forces = solver.kspace_force()
    fk_ref, f_ref = orc.kspace_force(method, s, M, cao, alpha)
    if np.linalg.norm(fk_ref) == 0.0:
        # ad at cao 1: the order-1 derivative weight is 0 (window-functions.c:101-103)
        assert np.all(forces.kspace() == 0.0)
        return
    print("method %d M %d cao %d: rel %.3e" % (method, M, cao, rel_err(forces.kspace(), fk_ref)))
    assert rel_err(forces.kspace(), fk_ref) < FP64_TOL
    if method in (2, 7):
        # self-force subtraction lands in f->f
        assert rel_err(forces.total(), f_ref) < 1e-11
    else:
        assert np.all(forces.total() == 0.0)
    solver.close()


@pytest.mark.parametrize("method,M,cao", [(3, 32, 7), (0, 32, 5), (7, 24, 6), (1, 28, 7)])
def test_kspace_force_clustered(pkg, orc, method, M, cao):
    """Gaussian-clustered charges: dense mesh tiles, many particles per tile."""
    s = random_system(20000, 20.0, seed=M + cao, layout="gaussian")
    solver = pkg.Solver(method, pkg.HostSystem.from_dict(s), pkg.parameters(M, cao, 1.0))
    forces = solver.kspace_force()
    fk_ref, f_ref = orc.kspace_force(method, s, M, cao, 1.0)
    assert rel_err(forces.kspace(), fk_ref) < FP64_TOL
    if method in (2, 7):
        assert rel_err(forces.total(), f_ref) < 1e-11
    solver.close()


@pytest.mark.parametrize("method", [0, 1, 2, 3, 6, 7])
def test_influence_function(pkg, orc, method):
    s = random_system(50, 9.0, seed=1)
    M, cao, alpha = 16, 5, 1.3
    solver = pkg.Solver(method, pkg.HostSystem.from_dict(s), pkg.parameters(M, cao, alpha))
    G = solver.influence_function()
    Gref = orc.influence(method, M, cao, 9.0, alpha)
    assert rel_err(G, Gref) < 1e-13
    assert np.all((G == 0.0) == (Gref == 0.0))
    if method in (2, 7):
        b = solver.self_force_coefficients()
        bref = orc.self_force_coeffs(M, cao, 9.0, Gref)
        assert rel_err(b, bref) < 1e-10
    solver.close()


def test_alpha_sweep_accumulates(pkg, orc):
    """Influence_function per alpha, Kspace_force accumulating into f_k
    (main.c:375-386, common.c:305)."""
    s = random_system(300, 10.0, seed=3)
    M, cao = 16, 7
    solver = pkg.Solver(7, pkg.HostSystem.from_dict(s), pkg.parameters(M, cao, 0.0))
    forces = pkg.HostForces(300)
    fk_sum = np.zeros((3, 300))
    f_sum = np.zeros((3, 300))
    for alpha in (0.6, 1.2):
        solver.set_alpha(alpha)
        solver.kspace_force(forces)
        fk, f = orc.kspace_force(7, s, M, cao, alpha)
        fk_sum += fk
        f_sum += f
    assert rel_err(forces.kspace(), fk_sum) < FP64_TOL
    assert rel_err(forces.total(), f_sum) < 1e-11
    solver.close()


@pytest.mark.parametrize("n,L,rcut,alpha", [(700, 10.0, 3.0, 1.0), (400, 7.0, 3.5, 0.9),
                                             (300, 10.0, 4.9, 0.7)])
def test_realteil(pkg, orc, n, L, rcut, alpha):
    s = random_system(n, L, seed=n)
    system = pkg.HostSystem.from_dict(s)
    forces = pkg.realteil(system, rcut, alpha)
    fr_ref, e_ref = orc.realteil(s, rcut, alpha)
    assert rel_err(forces.real(), fr_ref) < 1e-13
    assert abs(system.c.energy - e_ref) <= 1e-12 * abs(e_ref)


def test_realteil_clustered(pkg, orc):
    """Dense home cells (several 64-particle chunks, full neighbour queues)."""
    s = random_system(4000, 12.0, seed=9, layout="gaussian")
    system = pkg.HostSystem.from_dict(s)
    forces = pkg.realteil(system, 3.0, 1.1)
    fr_ref, e_ref = orc.realteil(s, 3.0, 1.1)
    assert rel_err(forces.real(), fr_ref) < 1e-13
    assert abs(system.c.energy - e_ref) <= 1e-12 * abs(e_ref)


def test_calculate_forces_tests_wall(pkg, orc):
    """Full Calculate_forces (real + k space) on the reference's 300-particle
    fixture: agrees with the oracle's P3M and with the fixture's Ewald forces
    at the method's accuracy."""
    s, f_ewald = load_fixture("tests_wall.npz")
    s = fold(s)
    M, cao, alpha, rcut = 32, 7, 1.2, 4.0
    system = pkg.HostSystem.from_dict(s)
    solver = pkg.Solver(3, system, pkg.parameters(M, cao, alpha, rcut))
    forces = solver.calculate_forces()
    fk_ref, _ = orc.kspace_force(3, s, M, cao, alpha)
    fr_ref, _ = orc.realteil(s, rcut, alpha)
    assert rel_err(forces.total(), fk_ref + fr_ref) < FP64_TOL
    assert orc.rms(forces.total(), f_ewald) < 1e-5
    solver.close()


@pytest.mark.parametrize("method", [0, 1, 2, 3, 6, 7])
def test_positions_one_period_outside(pkg, orc, method):
    """The reference takes positions up to one period outside the box (its
    mesh-index wrap, charge-assign.c:37-39): particles moved by +-L give the
    oracle's forces for the same raw positions, and those of the folded
    system to rounding; positions beyond [-L, 2L) are rejected."""
    s = random_system(800, 10.0, seed=70 + method)
    L, M, cao, alpha = 10.0, 16, 5, 1.1
    moved = dict(s)
    rng = np.random.default_rng(method)
    for k in "xyz":
        v = np.array(s[k], copy=True)
        pick = rng.random(v.size) < 0.3
        v[pick] = np.where(v[pick] >= 0.5 * L, v[pick] - L, v[pick] + L)
        moved[k] = v
    solver = pkg.Solver(method, pkg.HostSystem.from_dict(moved), pkg.parameters(M, cao, alpha))
    got = solver.kspace_force()
    fk_ref, f_ref = orc.kspace_force(method, moved, M, cao, alpha)
    fk_in, _ = orc.kspace_force(method, s, M, cao, alpha)
    assert rel_err(got.kspace(), fk_ref) < FP64_TOL
    assert rel_err(got.total() + got.kspace(), f_ref + fk_ref) < FP64_TOL
    assert rel_err(got.kspace(), fk_in) < 1e-9
    far = dict(moved)
    far["x"] = np.array(moved["x"], copy=True)
    far["x"][3] = 2.0 * L
    solver2 = pkg.Solver(method, pkg.HostSystem.from_dict(far), pkg.parameters(M, cao, alpha))
    with pytest.raises(pkg.P3MError) as e:
        solver2.kspace_force()
    assert e.value.code == pkg.ERANGE
    solver.close()
    solver2.close()


@pytest.mark.parametrize("method", [1, 3, 7])
def test_mixed_precision(pkg, orc, method):
    """FP32 mesh / FFT / interpolation: close to the FP64 result, far below
    the method's own error."""
    s = random_system(2000, 16.0, seed=11)
    M, cao, alpha = 32, 5, 1.0
    system = pkg.HostSystem.from_dict(s)
    solver = pkg.Solver(method, system, pkg.parameters(M, cao, alpha), pkg.MIXED)
    fk32 = solver.kspace_force().kspace()
    solver.set_precision(pkg.FP64)
    fk64 = solver.kspace_force().kspace()
    solver.close()
    rel = rel_err(fk32, fk64)
    print("mixed vs fp64 rel %.3e" % rel)
    assert rel < 2e-5


@pytest.mark.parametrize("method", [0, 1, 2, 3, 6, 7])
def test_error_estimates(pkg, orc, method):
    s = random_system(1000, 10.0, seed=5)
    M, cao, alpha, rcut = 16, 5, 1.0, 3.0
    system = pkg.HostSystem.from_dict(s)
    total, kspace = pkg.error_estimate(method, system, pkg.parameters(M, cao, alpha, rcut))
    q2 = float(np.sum(s["q"] ** 2))
    kref = ref_kspace_error(orc, method, 1000, 10.0, q2, M, cao, alpha)
    rref = orc.realspace_error(1000, 10.0, q2, rcut, alpha)
    # the summand alias1 - alias2^2/(...) cancels to ~1e-4 of its terms, so
    # summation order shows at ~1e-10 of the sum
    assert abs(kspace - kref) <= 1e-8 * abs(kref)
    assert abs(total - np.hypot(kref, rref)) <= 1e-8 * np.hypot(kref, rref)


@pytest.mark.parametrize("method", [0, 1, 6])
def test_error_table_negative_lookup(pkg, orc, method):
    """A negative table value means "not tabulated" to the reference, which
    then takes the direct sum: ik starting from the negative lookup, ik-i from
    0, then fabs (p3m-ik.c:267-297, p3m-ik-i.c:301-318).  Checked with a
    registered row of -0.5 at alpha L = 10, where the direct sum is well
    conditioned; the built-in row is restored afterwards."""
    import math

    s = random_system(1000, 10.0, seed=5)
    M, cao, alpha = 16, 5, 1.0
    kind = 1 if method == 1 else 0
    row = np.full(101, -0.5)
    system = pkg.HostSystem.from_dict(s)
    q2 = float(np.sum(s["q"] ** 2))
    try:
        pkg.set_error_table(kind, M, cao, row)
        _, k = pkg.error_estimate(method, system, pkg.parameters(M, cao, alpha, 3.0))
    finally:
        pkg.set_error_table(kind, M, cao, error_table_row(method, M, cao))
    kref = orc.kspace_error(method, 1000, 10.0, q2, M, cao, alpha, table=row)
    direct = orc.kspace_error(method, 1000, 10.0, q2, M, cao, alpha)
    print("method %d: %.9e (oracle %.9e, plain direct sum %.9e)" % (method, k, kref, direct))
    assert abs(k - kref) <= 1e-8 * abs(kref)
    if method == 1:
        assert abs(k - direct) <= 1e-8 * direct
    else:
        # 2 q2 sqrt(|-0.5 + S| / N) / L^2 against 2 q2 sqrt(|-1 + S| / N) / L^2
        assert abs(k - direct) > 0.1 * direct
    # the built-in row is back: the tabulated value again
    _, k2 = pkg.error_estimate(method, system, pkg.parameters(M, cao, alpha, 3.0))
    he = error_table_row(method, M, cao)[10]
    assert abs(k2 - 2.0 * q2 * math.sqrt(he / 1000) / 100.0) <= 1e-14 * k2


@pytest.mark.parametrize("method,M,cao,alpha", [(0, 16, 6, 0.15), (0, 32, 7, 0.6), (1, 16, 4, 0.25),
                                                (1, 16, 6, 0.2), (6, 16, 7, 0.2)])
def test_error_builtin_negative_entries(pkg, orc, method, M, cao, alpha):
    """The reference's tables hold rounding-noise negatives at small alpha L;
    there the estimate is the direct sum (from the lookup for ik), itself a
    sum of cancelling terms at the 1e-17 level, so GPU and oracle agree only
    to a few percent -- and for ik both are far below the estimate the
    untabulated path's -1 start would give."""
    s = random_system(1000, 10.0, seed=5)
    row = error_table_row(method, M, cao)
    aL = alpha * 10.0
    l = int(np.floor(aL))
    assert (1.0 - (aL - l)) * row[l] + (aL - l) * row[l + 1] < 0.0
    system = pkg.HostSystem.from_dict(s)
    _, k = pkg.error_estimate(method, system, pkg.parameters(M, cao, alpha, 3.0))
    q2 = float(np.sum(s["q"] ** 2))
    kref = orc.kspace_error(method, 1000, 10.0, q2, M, cao, alpha, table=row)
    kdirect = orc.kspace_error(method, 1000, 10.0, q2, M, cao, alpha)
    print("method %d M %d cao %d alpha L %.2f: %.6e (oracle %.6e, without the lookup %.6e)" %
          (method, M, cao, aL, k, kref, kdirect))
    assert abs(k - kref) <= 0.05 * abs(kref)
    if method != 1:
        assert k < 1e-3 * kdirect


@pytest.mark.parametrize("name,method_id", [("method_p3m_ik", 0), ("method_p3m_ik_i", 1),
                                            ("method_p3m_ad", 2), ("method_p3m_ad_i", 3),
                                            ("method_p3m_ik_r", 6), ("method_p3m_ad_r", 7)])
def test_method_table_drives_like_main(pkg, orc, name, method_id):
    """The exported method_t objects called through their function pointers
    in the order of the reference driver: Init (main.c:361), then per alpha
    Influence_function (main.c:378) and Kspace_force (common.c:305)
    accumulating into f_k, then Error / Error_k (main.c:432-433)."""
    m = pkg.exported_method(name)
    assert m.method_id == method_id
    s = random_system(400, 10.0, seed=40 + method_id)
    system = pkg.HostSystem.from_dict(s)
    M, cao, rcut = 16, 6, 3.0
    params = pkg.parameters(M, cao, 0.9, rcut)
    pptr = ctypes.pointer(params)
    d = m.Init(system.ptr, pptr)
    assert d
    forces = pkg.HostForces(400)
    fk_sum, f_sum = np.zeros((3, 400)), np.zeros((3, 400))
    for alpha in (0.9, 1.3):
        params.alpha = alpha
        m.Influence_function(system.ptr, pptr, d)
        m.Kspace_force(system.ptr, pptr, d, forces.ptr)
        fk, f = orc.kspace_force(method_id, s, M, cao, alpha)
        fk_sum += fk
        f_sum += f
    assert rel_err(forces.kspace(), fk_sum) < FP64_TOL
    if method_id in (2, 7):
        assert rel_err(forces.total(), f_sum) < 1e-11
    else:
        assert np.all(forces.total() == 0.0)
    q2 = float(np.sum(s["q"] ** 2))
    kref = ref_kspace_error(orc, method_id, 400, 10.0, q2, M, cao, params.alpha)
    rref = orc.realspace_error(400, 10.0, q2, rcut, params.alpha)
    assert abs(m.Error_k(system.ptr, pptr) - kref) <= 1e-8 * abs(kref)
    assert abs(m.Error(system.ptr, pptr) - np.hypot(kref, rref)) <= 1e-8 * np.hypot(kref, rref)
    pkg.check(pkg.lib().p3m_hip_free(d))


def _fft_api(pkg):
    """ctypes handles of the library's internal FFT launchers."""
    L = pkg.lib()
    fn = {}
    for name, res, args in [
        ("p3mk_queue_create", ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p)]),
        ("p3mk_queue_destroy", None, [ctypes.c_void_p]),
        ("p3mk_queue_sync", ctypes.c_int, [ctypes.c_void_p]),
        ("p3mk_fft_create", ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p]),
        ("p3mk_fft_destroy", None, [ctypes.c_void_p]),
        ("p3mk_fft3d", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
        ("p3mk_alloc", ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]),
        ("p3mk_release", None, [ctypes.c_void_p]),
        ("p3mk_copy_h2d", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
        ("p3mk_copy_d2h", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    ]:
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
        fn[name] = f

    for name, res, args in [
        ("p3mk_fft_forward_zy", ctypes.c_int, [ctypes.c_void_p] * 4),
        ("p3mk_fft_backward_yz", ctypes.c_int, [ctypes.c_void_p] * 4),
    ]:
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
        fn[name] = f
    return fn


class _Geom(ctypes.Structure):
    _fields_ = [("M", ctypes.c_int), ("cao", ctypes.c_int), ("r2c", ctypes.c_int),
                ("interlaced", ctypes.c_int), ("fp32", ctypes.c_int), ("L", ctypes.c_double)]


class _FftRig:
    """A queue, a plan and device buffers for one transform geometry."""

    def __init__(self, fn, M, r2c, fp32, nbytes):
        self.fn = fn
        self.q, self.plan = ctypes.c_void_p(), ctypes.c_void_p()
        self.bufs = []
        g = _Geom(M, 1, r2c, 0, fp32, 1.0)
        assert fn["p3mk_queue_create"](ctypes.byref(self.q)) == 0
        assert fn["p3mk_fft_create"](ctypes.byref(self.plan), ctypes.byref(g)) == 0
        for nb in nbytes:
            b = ctypes.c_void_p()
            assert fn["p3mk_alloc"](ctypes.byref(b), nb) == 0
            self.bufs.append(b)

    def put(self, k, a):
        assert self.fn["p3mk_copy_h2d"](self.bufs[k], a.ctypes.data, a.nbytes, self.q) == 0

    def get(self, k, out):
        assert self.fn["p3mk_copy_d2h"](out.ctypes.data, self.bufs[k], out.nbytes, self.q) == 0
        assert self.fn["p3mk_queue_sync"](self.q) == 0
        return out

    def close(self):
        for b in self.bufs:
            self.fn["p3mk_release"](b)
        self.fn["p3mk_fft_destroy"](self.plan)
        self.fn["p3mk_queue_destroy"](self.q)


@pytest.mark.parametrize("M", [16, 18, 20, 22, 24, 26, 28, 30, 36, 42, 44, 48, 64, 128])
def test_fft3d_vs_numpy(pkg, M):
    """The Stockham transform (radices 2..13) against numpy's FFT."""
    fn = _fft_api(pkg)
    rng = np.random.default_rng(M)
    a = (rng.standard_normal((M, M, M)) + 1j * rng.standard_normal((M, M, M))).astype(np.complex128)
    rig = _FftRig(fn, M, 0, 0, [a.nbytes])
    try:
        for sign in (-1, 1):
            rig.put(0, a)
            assert fn["p3mk_fft3d"](rig.plan, rig.bufs[0], sign, rig.q) == 0
            out = rig.get(0, np.empty_like(a))
            ref = np.fft.fftn(a) if sign < 0 else np.fft.ifftn(a) * M ** 3
            assert rel_err(out, ref) < 1e-14
    finally:
        rig.close()


@pytest.mark.parametrize("M", [16, 18, 22, 26, 28, 30, 36, 42, 44, 64, 128])
def test_fft3d_fp32_vs_numpy(pkg, M):
    """The FP32 passes of the mixed form (radices 2..13) against numpy's FP64
    FFT of the same FP32 input: relative error within a few FP32 roundings
    per radix stage."""
    fn = _fft_api(pkg)
    rng = np.random.default_rng(100 + M)
    a = (rng.standard_normal((M, M, M)) + 1j * rng.standard_normal((M, M, M))).astype(np.complex64)
    rig = _FftRig(fn, M, 0, 1, [a.nbytes])
    try:
        for sign in (-1, 1):
            rig.put(0, a)
            assert fn["p3mk_fft3d"](rig.plan, rig.bufs[0], sign, rig.q) == 0
            out = rig.get(0, np.empty_like(a))
            a64 = a.astype(np.complex128)
            ref = np.fft.fftn(a64) if sign < 0 else np.fft.ifftn(a64) * M ** 3
            err = rel_err(out.astype(np.complex128), ref)
            print("M %d sign %d: %.2e" % (M, sign, err))
            assert err < 2e-6
    finally:
        rig.close()


@pytest.mark.parametrize("M,fp32", [(16, 0), (22, 0), (30, 0), (64, 0), (128, 0), (16, 1), (26, 1),
                                    (42, 1), (64, 1), (128, 1)])
def test_fft_r2c_zy_vs_numpy(pkg, M, fp32):
    """The r2c z pass with the y pass (methods 6 and 7) and the inverse y pass
    with the c2r z pass, against numpy: spectrum[x, y, k] = fft(rfft(a, z), y)
    for k <= M/2, and back M^2 irfft(ifft(., y), z)."""
    fn = _fft_api(pkg)
    H = M // 2 + 1
    rdt, cdt = (np.float32, np.complex64) if fp32 else (np.float64, np.complex128)
    rng = np.random.default_rng(200 + M + fp32)
    a = rng.standard_normal((M, M, M)).astype(rdt)
    spec = np.empty((M, M, H), dtype=cdt)
    rig = _FftRig(fn, M, 1, fp32, [a.nbytes, spec.nbytes])
    tol = 2e-6 if fp32 else 1e-14
    try:
        rig.put(0, a)
        assert fn["p3mk_fft_forward_zy"](rig.plan, rig.bufs[0], rig.bufs[1], rig.q) == 0
        rig.get(1, spec)
        ref = np.fft.fft(np.fft.rfft(a.astype(np.float64), axis=2), axis=1)
        ef = rel_err(spec.astype(np.complex128), ref)
        assert fn["p3mk_fft_backward_yz"](rig.plan, rig.bufs[1], rig.bufs[0], rig.q) == 0
        back = rig.get(0, np.empty_like(a))
        eb = rel_err(back.astype(np.float64), M * M * a.astype(np.float64))
        print("M %d fp32 %d: forward %.2e, round trip %.2e" % (M, fp32, ef, eb))
        assert ef < tol
        assert eb < tol
    finally:
        rig.close()


@pytest.mark.gpu
@pytest.mark.parametrize("M,fp32", [(64, 1), (128, 1), (128, 0)])
def test_fft_c2c_zy_vs_numpy(pkg, M, fp32):
    """The z and y passes of the c2c forms in place (at M = 128 in FP32 both
    in one launch, a y-z plane per workgroup in LDS) and their inverse,
    against numpy: fft(fft(a, z), y), and back M^2 a."""
    fn = _fft_api(pkg)
    cdt = np.complex64 if fp32 else np.complex128
    rng = np.random.default_rng(300 + M + fp32)
    a = (rng.standard_normal((M, M, M)) + 1j * rng.standard_normal((M, M, M))).astype(cdt)
    rig = _FftRig(fn, M, 0, fp32, [a.nbytes])
    tol = 2e-6 if fp32 else 1e-14
    try:
        rig.put(0, a)
        assert fn["p3mk_fft_forward_zy"](rig.plan, rig.bufs[0], rig.bufs[0], rig.q) == 0
        spec = rig.get(0, np.empty_like(a))
        a64 = a.astype(np.complex128)
        ef = rel_err(spec.astype(np.complex128), np.fft.fft(np.fft.fft(a64, axis=2), axis=1))
        assert fn["p3mk_fft_backward_yz"](rig.plan, rig.bufs[0], rig.bufs[0], rig.q) == 0
        back = rig.get(0, np.empty_like(a))
        eb = rel_err(back.astype(np.complex128), M * M * a64)
        print("M %d fp32 %d: forward %.2e, round trip %.2e" % (M, fp32, ef, eb))
        assert ef < tol
        assert eb < tol
    finally:
        rig.close()
