"""GPU parity at the five benchmark configurations, at their full sizes.

Each test builds the configuration with the generator the benchmark uses
(`bench.make_system`, `bench.batch_systems`) and compares the HIP library,
called through its C ABI, with the oracle on the same inputs:

  C1  method 7 (ad, r2c), 1000 charges +-1 alternating in a box of side 10,
      mesh 16, cao 7, rcut 3: the driver's alpha sweep 0.1 -> 2.5 step 0.05
      (49 values, main.c:375-455) through the exported method_p3m_ad_r,
      Calculate_forces per alpha (common.c:268-327), Error / Error_k.
  C2  method 1 (ik interlaced), 100k uniform charges, box 46.4, mesh 64,
      cao 5, alpha 0.9: k space on every particle, real space on every
      particle (the oracle's cell list), mixed against FP64.
  C3  method 3 (ad interlaced), 1M charges Gaussian about the box centre,
      box 100, mesh 128, cao 7, rcut 4, alpha 0.75: k space on every
      particle; real space on 2000 rows.
  C4  the 256-system batch (method 3): every system's k-space and real-space
      force against the oracle run on that system alone.
  C5  method 0 (ik), 32768 charges N(0, 0.5) mean-subtracted, box 32,
      mesh 32, cao 5, alpha 1, rcut 3: forward force, and the adjoint of
      loss = sum |F - F_target|^2 (F_target from charges of a second seed)
      in FP64 and in the mixed form against central differences of the
      oracle on 30 coordinates and 20 charges.

The FP64 bar is 1e-12 relative (||F_gpu - F_oracle|| / ||F_oracle||).  The
oracle runs on the host in a thread pool (its C functions keep no global
state and ctypes releases the interpreter lock during each call).
"""
import ctypes
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import bench
from conftest import rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
QUANT = 131072  # weight-quantisation bins per mesh spacing (charge-assign.c:94-97)
WORKERS = max(1, min(16, os.cpu_count() or 1))


def pmap(fn, items):
    with ThreadPoolExecutor(WORKERS) as ex:
        return list(ex.map(fn, items))


def oracle_rows(orc, s, rcut, alpha, begin, end, parts=8):
    """Real-space force of rows [begin, end) on the oracle's cell list."""
    bounds = np.linspace(begin, end, parts + 1).astype(int)

    def one(k):
        fr, _ = orc.realteil(s, rcut, alpha, cells=True, i_range=(bounds[k], bounds[k + 1]))
        return fr[:, bounds[k]:bounds[k + 1]]

    return np.concatenate(pmap(one, range(parts)), axis=1)


# ---- C1 ----------------------------------------------------------------------

def c1_alphas():
    """alpha values of the driver's loop (main.c:375): alpha += 0.05 from 0.1
    while alpha <= 2.5, accumulated in FP64."""
    out, a = [], 0.1
    while a <= 2.5:
        out.append(a)
        a += 0.05
    return out


def test_c1_alpha_sweep_through_method_table(pkg, orc):
    w = bench.WORKLOADS["uniform-1k-ad-r-m16"]
    s = bench.make_system(w)
    N, L, M, cao, rcut = w["N"], w["L"], w["M"], w["cao"], w["rcut"]
    alphas = c1_alphas()
    assert len(alphas) == 49
    m = pkg.exported_method("method_p3m_ad_r")
    assert m.method_id == 7
    system = pkg.HostSystem.from_dict(s)
    params = pkg.parameters(M, cao, 0.0, rcut)
    pptr = ctypes.pointer(params)
    d = m.Init(system.ptr, pptr)
    assert d
    q2 = float(np.sum(s["q"] ** 2))

    def reference(alpha):
        fk, fself = orc.kspace_force(7, s, M, cao, alpha)
        fr, e = orc.realteil(s, rcut, alpha)
        return fk, fr, fself + (fk + fr), e

    refs = pmap(reference, alphas)
    f_ewald = orc.ewald_reference(s)[0]
    worst = 0.0
    for alpha, (fk_ref, fr_ref, f_ref, e_ref) in zip(alphas, refs):
        params.alpha = alpha
        m.Influence_function(system.ptr, pptr, d)
        forces = pkg.HostForces(N)
        system.c.energy = 0.0
        pkg.check(pkg.lib().p3m_hip_calculate_forces(ctypes.pointer(m), system.ptr, pptr, d,
                                                     forces.ptr))
        ek = rel_err(forces.kspace(), fk_ref)
        er = rel_err(forces.real(), fr_ref)
        et = rel_err(forces.total(), f_ref)
        worst = max(worst, ek, er, et)
        assert ek < FP64_TOL, (alpha, ek)
        assert er < FP64_TOL, (alpha, er)
        assert et < FP64_TOL, (alpha, et)
        assert abs(system.c.energy - e_ref) <= 1e-12 * abs(e_ref)
        kref = orc.kspace_error(7, N, L, q2, M, cao, alpha)
        rref = orc.realspace_error(N, L, q2, rcut, alpha)
        est = m.Error(system.ptr, pptr)
        # Error_k = 2 q2 sqrt(he_q / N) / L^2 with he_q a sum of terms that
        # cancel to 1e-6 ... 1e-12 of their size as alpha falls (at alpha 0.1
        # he_q is rounding noise), so he_q itself is compared: to 1e-5, or to
        # 1e-16 of the squared total estimate it enters
        tot = np.hypot(kref, rref)
        ek2 = m.Error_k(system.ptr, pptr) ** 2
        assert abs(ek2 - kref ** 2) <= 1e-5 * kref ** 2 + 1e-16 * tot ** 2, (alpha, ek2, kref ** 2)
        assert abs(est - tot) <= 1e-8 * tot
        # the driver's accuracy column (main.c:446): rms error against the
        # Ewald reference next to the a-priori estimate
        dF = orc.rms(forces.total(), f_ewald)
        assert 0.7 < dF / est < 1.6, (alpha, dF, est)
    print("C1: 49 alphas, worst relative difference %.2e" % worst)
    pkg.check(pkg.lib().p3m_hip_free(d))


# ---- C2 ----------------------------------------------------------------------

def test_c2_uniform_100k_ik_i(pkg, orc):
    w = bench.WORKLOADS["uniform-100k-ik-i-m64"]
    s = bench.make_system(w)
    N, M, cao, alpha, rcut = w["N"], w["M"], w["cao"], w["alpha"], w["rcut"]
    system = pkg.HostSystem.from_dict(s)
    solver = pkg.Solver(1, system, pkg.parameters(M, cao, alpha, rcut))
    fk64 = solver.kspace_force().kspace()
    solver.set_precision(pkg.MIXED)
    fk32 = solver.kspace_force().kspace()
    solver.close()
    fk_ref, _ = orc.kspace_force(1, s, M, cao, alpha)
    ek, em = rel_err(fk64, fk_ref), rel_err(fk32, fk_ref)
    fr = pkg.realteil(system, rcut, alpha).real()
    fr_ref = oracle_rows(orc, s, rcut, alpha, 0, N)
    er = rel_err(fr, fr_ref)
    print("C2: k space %.2e, mixed k space vs oracle %.2e, real space %.2e" % (ek, em, er))
    assert ek < FP64_TOL
    assert er < FP64_TOL
    assert em < 2e-5


# ---- C3 ----------------------------------------------------------------------

@pytest.fixture(scope="module")
def c3():
    w = bench.WORKLOADS["gaussian-1M-ad-i-m128"]
    return w, bench.make_system(w)


def test_c3_gaussian_1M_kspace(pkg, orc, c3):
    w, s = c3
    M, cao, alpha = w["M"], w["cao"], w["alpha"]
    solver = pkg.Solver(3, pkg.HostSystem.from_dict(s), pkg.parameters(M, cao, alpha))
    fk64 = solver.kspace_force().kspace()
    solver.set_precision(pkg.MIXED)
    fk32 = solver.kspace_force().kspace()
    solver.close()
    fk_ref, f_ref = orc.kspace_force(3, s, M, cao, alpha)
    assert np.all(f_ref == 0.0)  # no self-force term for method 3
    ek, em = rel_err(fk64, fk_ref), rel_err(fk32, fk_ref)
    print("C3: k space FP64 %.2e, mixed %.2e (against the oracle, all 1M particles)" % (ek, em))
    assert ek < FP64_TOL
    assert em < 2e-5


def test_c3_gaussian_1M_realspace_rows(pkg, orc, c3):
    w, s = c3
    alpha, rcut = w["alpha"], w["rcut"]
    system = pkg.HostSystem.from_dict(s)
    fr64 = pkg.realteil(system, rcut, alpha).real()
    # a contiguous index range is a random sample: the generator draws the
    # particles independently
    b = int(np.random.default_rng(33).integers(0, w["N"] - 2000))
    fr_ref = oracle_rows(orc, s, rcut, alpha, b, b + 2000)
    er = rel_err(fr64[:, b:b + 2000], fr_ref)
    print("C3: real space FP64, rows %d..%d: %.2e" % (b, b + 2000, er))
    assert er < FP64_TOL


# ---- C4 ----------------------------------------------------------------------

@pytest.fixture(scope="module")
def c4(pkg):
    import torch

    systems, data = bench.batch_systems()
    for sy, d in zip(systems, data):
        d["L"] = sy["L"]
    dev = {k: torch.tensor(np.concatenate([d[k] for d in data]), dtype=torch.float64, device="cuda")
           for k in "xyzq"}
    out = {}
    for name, prec in (("fp64", pkg.FP64), ("mixed", pkg.MIXED)):
        batch = pkg.Batch(3, systems, prec)
        n = batch.ntot
        f = torch.zeros(3 * n, dtype=torch.float64, device="cuda")
        fk, fr = torch.zeros_like(f), torch.zeros_like(f)
        torch.cuda.synchronize()
        batch.calculate_forces_device(*[dev[k].data_ptr() for k in "xyzq"], f.data_ptr(),
                                      fk.data_ptr(), fr.data_ptr(), None)
        batch.synchronize()
        out[name] = {"f": f.cpu().numpy().reshape(3, n), "fk": fk.cpu().numpy().reshape(3, n),
                     "fr": fr.cpu().numpy().reshape(3, n)}
        offsets = batch.offsets
        batch.close()
    return systems, data, offsets, out


def test_c4_batch_kspace_every_system(orc, c4):
    systems, data, off, out = c4
    assert len(systems) == 256

    def one(k):
        S, d = systems[k], data[k]
        fk_ref, _ = orc.kspace_force(3, d, S["M"], S["cao"], S["alpha"])
        return rel_err(out["fp64"]["fk"][:, off[k]:off[k + 1]], fk_ref)

    errs = np.array(pmap(one, range(len(systems))))
    k = int(np.argmax(errs))
    print("C4: k space, worst system %d (n %d, M %d, cao %d): %.2e" %
          (k, systems[k]["n"], systems[k]["M"], systems[k]["cao"], errs[k]))
    assert np.all(errs < FP64_TOL)
    em = rel_err(out["mixed"]["fk"], out["fp64"]["fk"])
    print("C4: mixed k space vs FP64 over all systems %.2e" % em)
    assert em < 2e-5


def test_c4_batch_realspace_every_system(orc, c4):
    systems, data, off, out = c4

    def one(k):
        S, d = systems[k], data[k]
        fr_ref, _ = orc.realteil(d, S["rcut"], S["alpha"], cells=True)
        return (rel_err(out["fp64"]["fr"][:, off[k]:off[k + 1]], fr_ref),
                rel_err(out["mixed"]["fr"][:, off[k]:off[k + 1]], fr_ref))

    errs = np.array(pmap(one, range(len(systems))))
    print("C4: real space, worst system FP64 %.2e, mixed %.2e" % tuple(errs.max(axis=0)))
    assert np.all(errs[:, 0] < FP64_TOL)
    assert np.all(errs[:, 1] < 2e-5)
    # Calculate_forces sums the parts (common.c:312-319)
    ft = out["fp64"]["f"]
    assert rel_err(ft, out["fp64"]["fk"] + out["fp64"]["fr"]) < 1e-15


# ---- C5 ----------------------------------------------------------------------

def snap_to_bins(s, M):
    """Positions moved to the centre of their quantisation bin: the quantised
    forward force is then the smooth B-spline force at the bin's lower edge,
    which is what the adjoint differentiates, and moves by whole bins keep
    that true."""
    out = dict(s)
    Hi = M / s["L"]
    for k in "xyz":
        u = np.asarray(s[k]) * Hi
        u = (np.floor(u * QUANT) + 0.5) / QUANT
        out[k] = np.clip(u / Hi, 0.0, np.nextafter(s["L"], 0.0))
    return out


def neighbourhood(s, j, radius):
    """Indices of the particles within `radius` of particle j (minimum image),
    j first."""
    L = s["L"]
    d = np.stack([np.asarray(s[k]) - s[k][j] for k in "xyz"])
    d -= np.round(d / L) * L
    r = np.sqrt((d ** 2).sum(0))
    idx = np.flatnonzero(r <= radius)
    return np.concatenate([[j], idx[idx != j]]), r


def c5_case():
    w = bench.WORKLOADS["uniform-32k-ik-m32"]
    s = snap_to_bins(bench.make_system(w), w["M"])
    rng6 = np.random.default_rng(6)
    qt = rng6.normal(0.0, 0.5, w["N"])
    qt -= qt.mean()
    return w, s, qt


def test_c5_forward(pkg, orc):
    w, s, _ = c5_case()
    M, cao, alpha, rcut = w["M"], w["cao"], w["alpha"], w["rcut"]
    solver = pkg.Solver(0, pkg.HostSystem.from_dict(s), pkg.parameters(M, cao, alpha, rcut))
    forces = solver.calculate_forces()
    solver.close()
    fk_ref, _ = orc.kspace_force(0, s, M, cao, alpha)
    fr_ref = oracle_rows(orc, s, rcut, alpha, 0, w["N"])
    ek, er = rel_err(forces.kspace(), fk_ref), rel_err(forces.real(), fr_ref)
    et = rel_err(forces.total(), fk_ref + fr_ref)
    print("C5 forward: k %.2e r %.2e total %.2e" % (ek, er, et))
    assert ek < FP64_TOL and er < FP64_TOL and et < FP64_TOL


def test_c5_adjoint_against_finite_differences(pkg, orc):
    """d loss / d (positions, charges) for loss = sum |F - F_target|^2 with F
    the total force (k space + real space): the library's k-space and
    real-space adjoints with the cotangent 2 (F - F_target), against central
    differences of the oracle.  k space: moves by 4 quantisation bins,
    differences of the whole force field; real space: only the pairs of the
    moved particle change, so the difference is taken over the particles
    within rcut + 1 of it (every pair of the moved particle is inside)."""
    w, s, qt = c5_case()
    N, L, M, cao, alpha, rcut = w["N"], w["L"], w["M"], w["cao"], w["alpha"], w["rcut"]
    G = orc.influence(0, M, cao, L, alpha)
    total = lambda t: orc.kspace_force(0, t, M, cao, alpha, G=G)[0] + oracle_rows(orc, t, rcut, alpha, 0, N)
    F = total(s)
    F_target = total(dict(s, q=qt))
    g = 2.0 * (F - F_target)

    rng = np.random.default_rng(55)
    dx = 4.0 / QUANT * L / M
    probes, coords = [], []
    # particles whose pair distances all stay clear of rcut under the move,
    # away from the box faces
    for j in rng.permutation(N):
        if min(min(s[k][j], L - s[k][j]) for k in "xyz") < 1e-3:
            continue
        _, r = neighbourhood(s, j, rcut + 1.0)
        if np.min(np.abs(r - rcut)) > 4 * dx:
            probes.append(int(j))
        if len(probes) == 10:
            break
    for j in probes:
        coords += [(j, d) for d in range(3)]
    qprobes = [int(j) for j in rng.choice(N, 20, replace=False)]

    def fd_coord(jd):
        j, d = jd
        k = "xyz"[d]
        sp, sm = dict(s), dict(s)
        sp[k] = np.array(s[k], copy=True)
        sm[k] = np.array(s[k], copy=True)
        sp[k][j] += dx
        sm[k][j] -= dx
        dk = (orc.kspace_force(0, sp, M, cao, alpha, G=G)[0] -
              orc.kspace_force(0, sm, M, cao, alpha, G=G)[0])
        nb, _ = neighbourhood(s, j, rcut + 1.0)
        sub = lambda t: {**{c: np.asarray(t[c])[nb] for c in "xyzq"}, "L": L}
        dr = orc.realteil(sub(sp), rcut, alpha)[0] - orc.realteil(sub(sm), rcut, alpha)[0]
        return (np.sum(g * dk) + np.sum(g[:, nb] * dr)) / (2 * dx)

    def fd_charge(j):
        h = 1e-3
        sp, sm = dict(s), dict(s)
        sp["q"] = np.array(s["q"], copy=True)
        sm["q"] = np.array(s["q"], copy=True)
        sp["q"][j] += h
        sm["q"][j] -= h
        dk = (orc.kspace_force(0, sp, M, cao, alpha, G=G)[0] -
              orc.kspace_force(0, sm, M, cao, alpha, G=G)[0])
        nb, _ = neighbourhood(s, j, rcut + 1.0)
        sub = lambda t: {**{c: np.asarray(t[c])[nb] for c in "xyzq"}, "L": L}
        dr = orc.realteil(sub(sp), rcut, alpha)[0] - orc.realteil(sub(sm), rcut, alpha)[0]
        return (np.sum(g * dk) + np.sum(g[:, nb] * dr)) / (2 * h)

    fd_pos = np.array(pmap(fd_coord, coords))
    fd_q = np.array(pmap(fd_charge, qprobes))
    system = pkg.HostSystem.from_dict(s)
    rpos, rq = pkg.realteil_vjp(system, rcut, alpha, g)
    ci = (np.array([c[1] for c in coords]), np.array([c[0] for c in coords]))
    for name, prec, tol_pos, tol_q in (("FP64", pkg.FP64, 2e-7, 1e-9),
                                       ("mixed", pkg.MIXED, 1e-4, 1e-4)):
        solver = pkg.Solver(0, system, pkg.parameters(M, cao, alpha, rcut), prec)
        kpos, kq = solver.kspace_force_vjp(g)
        solver.close()
        ep = rel_err((kpos + rpos)[ci], fd_pos)
        eq = rel_err((kq + rq)[qprobes], fd_q)
        print("C5 adjoint %s: %d coordinates %.2e, %d charges %.2e" %
              (name, len(coords), ep, len(qprobes), eq))
        assert ep < tol_pos
        assert eq < tol_q
