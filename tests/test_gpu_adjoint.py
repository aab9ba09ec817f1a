"""GPU tests of the adjoint of Kspace_force (p3m_hip_kspace_force_vjp*).

The adjoint is pinned against central finite differences of the oracle's
forward k-space force.  Positions are placed at the centres of the
reference's weight-quantisation bins (1/131072 of a mesh spacing,
charge-assign.c:94-97) and moved by whole bins, so the quantised forward
force equals the smooth B-spline force at the bins' lower edges, which is
where the adjoint differentiates it; the difference quotient then has only
its O(delta^2) truncation error.
"""
import ctypes

import numpy as np
import pytest

from conftest import random_system, rel_err

pytestmark = pytest.mark.gpu

QUANT = 131072  # bins per mesh spacing


def snap_to_bins(s, M):
    """Positions moved to the centre of their quantisation bin."""
    out = dict(s)
    Hi = M / s["L"]
    for k in "xyz":
        u = np.asarray(s[k]) * Hi
        u = (np.floor(u * QUANT) + 0.5) / QUANT
        out[k] = np.clip(u / Hi, 0.0, np.nextafter(s["L"], 0.0))
    return out


def oracle_force(orc, method, s, M, cao, alpha, G, b):
    fk, f = orc.kspace_force(method, s, M, cao, alpha, G=G, b=b)
    return fk + f


def fd_gradient(orc, method, s, M, cao, alpha, gvec, nbins=8, dq_rel=1e-4):
    """Central differences of J = sum(g * F) w.r.t. every coordinate and
    charge (oracle forward)."""
    G = orc.influence(method, M, cao, s["L"], alpha)
    b = orc.self_force_coeffs(M, cao, s["L"], G) if method in (2, 7) else None
    n = len(s["q"])
    dx = nbins / QUANT * s["L"] / M
    dpos = np.zeros((3, n))
    dq = np.zeros(n)
    for d, k in enumerate("xyz"):
        for j in range(n):
            sp, sm = dict(s), dict(s)
            sp[k] = np.array(s[k], copy=True)
            sm[k] = np.array(s[k], copy=True)
            sp[k][j] += dx
            sm[k][j] -= dx
            Jp = np.sum(gvec * oracle_force(orc, method, sp, M, cao, alpha, G, b))
            Jm = np.sum(gvec * oracle_force(orc, method, sm, M, cao, alpha, G, b))
            dpos[d, j] = (Jp - Jm) / (2 * dx)
    for j in range(n):
        h = dq_rel * max(1.0, abs(s["q"][j]))
        sp, sm = dict(s), dict(s)
        sp["q"] = np.array(s["q"], copy=True)
        sm["q"] = np.array(s["q"], copy=True)
        sp["q"][j] += h
        sm["q"][j] -= h
        Jp = np.sum(gvec * oracle_force(orc, method, sp, M, cao, alpha, G, b))
        Jm = np.sum(gvec * oracle_force(orc, method, sm, M, cao, alpha, G, b))
        dq[j] = (Jp - Jm) / (2 * h)
    return dpos, dq


def interior_system(n, L, seed, margin):
    """Uniform positions kept `margin` away from the box faces, so that the
    finite-difference moves never leave [0, L)."""
    s = random_system(n, L, seed=seed)
    rng = np.random.default_rng(seed + 1)
    for k in "xyz":
        s[k] = rng.uniform(margin, L - margin, n)
    return s


@pytest.mark.parametrize("method,M,cao", [(0, 16, 5), (1, 16, 4), (6, 16, 5), (2, 16, 5),
                                          (3, 16, 6), (7, 16, 7), (0, 12, 7), (2, 12, 3)])
def test_vjp_matches_finite_differences(pkg, orc, method, M, cao):
    L, alpha, n = 8.0, 1.3, 16
    s = snap_to_bins(interior_system(n, L, 70 + method + cao, 1e-3), M)
    rng = np.random.default_rng(5 + method)
    gvec = rng.normal(size=(3, n))
    solver = pkg.Solver(method, pkg.HostSystem.from_dict(s), pkg.parameters(M, cao, alpha))
    dpos, dq = solver.kspace_force_vjp(gvec)
    solver.close()
    fd_pos, fd_q = fd_gradient(orc, method, s, M, cao, alpha, gvec)
    ep, eq = rel_err(dpos, fd_pos), rel_err(dq, fd_q)
    print("method %d M %d cao %d: positions %.2e charges %.2e" % (method, M, cao, ep, eq))
    assert ep < 2e-7
    assert eq < 1e-9


@pytest.mark.parametrize("method", [0, 3, 7])
def test_vjp_mixed_precision(pkg, method):
    s = random_system(3000, 16.0, seed=11 + method)
    rng = np.random.default_rng(3)
    gvec = rng.normal(size=(3, 3000))
    out = []
    for prec in (pkg.FP64, pkg.MIXED):
        solver = pkg.Solver(method, pkg.HostSystem.from_dict(s), pkg.parameters(32, 5, 1.0), prec)
        out.append(solver.kspace_force_vjp(gvec))
        solver.close()
    (p64, q64), (p32, q32) = out
    print("mixed vs fp64: positions %.2e charges %.2e" % (rel_err(p32, p64), rel_err(q32, q64)))
    assert rel_err(p32, p64) < 1e-5
    assert rel_err(q32, q64) < 1e-5


def test_vjp_accumulates_and_device_form(pkg):
    """The host form adds into dpos/dq; the device form gives the same."""
    import torch

    s = random_system(500, 10.0, seed=4)
    rng = np.random.default_rng(8)
    gvec = rng.normal(size=(3, 500))
    system = pkg.HostSystem.from_dict(s)
    solver = pkg.Solver(1, system, pkg.parameters(16, 5, 1.1))
    dpos, dq = solver.kspace_force_vjp(gvec)
    dev = {k: torch.tensor(np.asarray(s[k]), dtype=torch.float64, device="cuda") for k in "xyzq"}
    g = torch.tensor(gvec.reshape(-1), dtype=torch.float64, device="cuda")
    dp = torch.ones(1500, dtype=torch.float64, device="cuda")
    dqq = torch.ones(500, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    solver.kspace_force_vjp_device(500, 10.0, *[dev[k].data_ptr() for k in "xyzq"], g.data_ptr(),
                                   dp.data_ptr(), dqq.data_ptr())
    solver.synchronize()
    assert rel_err(dp.cpu().numpy().reshape(3, 500) - 1.0, dpos) < 1e-14
    assert rel_err(dqq.cpu().numpy() - 1.0, dq) < 1e-14
    solver.close()


def test_autograd_function(pkg, orc):
    """loss = sum |F - F_target|^2 through torch.autograd (ik, the
    differentiable configuration's method), against the finite differences
    of the oracle."""
    import importlib

    import torch

    ag = importlib.import_module("p3m_standalone_amd.autograd")
    M, cao, alpha, L, n = 16, 5, 1.2, 8.0, 12
    s = snap_to_bins(interior_system(n, L, 21, 1e-3), M)
    rng = np.random.default_rng(9)
    target = rng.normal(size=(3, n))
    solver = pkg.Solver(0, pkg.HostSystem.from_dict(s), pkg.parameters(M, cao, alpha))
    pos = torch.tensor(np.stack([s["x"], s["y"], s["z"]]), dtype=torch.float64, device="cuda",
                       requires_grad=True)
    q = torch.tensor(np.asarray(s["q"]), dtype=torch.float64, device="cuda", requires_grad=True)
    F = ag.kspace_force(pos, q, solver)
    Fo = oracle_force(orc, 0, s, M, cao, alpha, orc.influence(0, M, cao, L, alpha), None)
    assert rel_err(F.detach().cpu().numpy(), Fo) < 1e-12
    loss = ((F - torch.tensor(target, device="cuda")) ** 2).sum()
    loss.backward()
    # d loss = sum 2 (F - T) dF: the cotangent is 2 (F - T), fixed at F
    gvec = 2.0 * (Fo - target)
    fd_pos, fd_q = fd_gradient(orc, 0, s, M, cao, alpha, gvec)
    assert rel_err(pos.grad.cpu().numpy(), fd_pos) < 2e-7
    assert rel_err(q.grad.cpu().numpy(), fd_q) < 1e-9
    solver.close()


def test_vjp_against_ewald(pkg, orc):
    """At an accurate mesh the P3M adjoint approaches the derivative of the
    Ewald k-space force (finite differences of the oracle's Ewald sum)."""
    L, n, alpha, M, cao = 8.0, 10, 1.0, 32, 7
    s = interior_system(n, L, 33, 1e-2)
    rng = np.random.default_rng(2)
    gvec = rng.normal(size=(3, n))
    solver = pkg.Solver(0, pkg.HostSystem.from_dict(s), pkg.parameters(M, cao, alpha))
    dpos, dq = solver.kspace_force_vjp(gvec)
    solver.close()
    kmax = 16
    h = 1e-5
    ref = np.zeros((3, n))
    for d, k in enumerate("xyz"):
        for j in range(n):
            sp, sm = dict(s), dict(s)
            sp[k] = np.array(s[k], copy=True)
            sm[k] = np.array(s[k], copy=True)
            sp[k][j] += h
            sm[k][j] -= h
            ref[d, j] = (np.sum(gvec * orc.ewald_kspace(sp, alpha, kmax)) -
                         np.sum(gvec * orc.ewald_kspace(sm, alpha, kmax))) / (2 * h)
    err = rel_err(dpos, ref)
    print("P3M adjoint vs Ewald derivative: %.2e" % err)
    assert err < 1e-4
