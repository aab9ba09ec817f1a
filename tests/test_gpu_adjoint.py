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


# ---- real-space adjoint (p3m_hip_realteil_vjp*) ----------------------------

def pair_distances(s):
    P = np.stack([s["x"], s["y"], s["z"]], 1)
    d = P[:, None, :] - P[None, :, :]
    d -= np.round(d / s["L"]) * s["L"]
    r = np.sqrt((d ** 2).sum(-1))
    return r[np.triu_indices(len(P), 1)]


def away_from_cutoff(n, L, rcut, seed, gap):
    """A uniform system none of whose pair distances is within `gap` of rcut
    (the real-space force jumps there, so finite differences must not cross
    it)."""
    for k in range(100):
        s = interior_system(n, L, seed + 17 * k, 1e-3)
        if np.min(np.abs(pair_distances(s) - rcut)) > gap:
            return s
    raise RuntimeError("no system away from the cutoff")


def fd_realspace(orc, s, rcut, alpha, gvec, idx, h=1e-6, hq=1e-3):
    """Central differences of J = sum g . f_r (oracle Realteil, O(N^2))."""
    J = lambda t: np.sum(gvec * orc.realteil(t, rcut, alpha)[0])
    dpos, dq = np.zeros((3, len(idx))), np.zeros(len(idx))
    for c, j in enumerate(idx):
        for d, k in enumerate("xyz"):
            sp, sm = dict(s), dict(s)
            sp[k] = np.array(s[k], copy=True)
            sm[k] = np.array(s[k], copy=True)
            sp[k][j] += h
            sm[k][j] -= h
            dpos[d, c] = (J(sp) - J(sm)) / (2 * h)
        sp, sm = dict(s), dict(s)
        sp["q"] = np.array(s["q"], copy=True)
        sm["q"] = np.array(s["q"], copy=True)
        sp["q"][j] += hq
        sm["q"][j] -= hq
        dq[c] = (J(sp) - J(sm)) / (2 * hq)
    return dpos, dq


@pytest.mark.parametrize("n,L,rcut,alpha", [(300, 8.0, 3.0, 1.0), (200, 5.0, 2.4, 1.4)])
def test_realteil_vjp_matches_finite_differences(pkg, orc, n, L, rcut, alpha):
    """The real-space adjoint against central differences of the oracle's
    Realteil, for a random cotangent, on particles sampled over the box (the
    second system has 2 cells per side: the all-cells path)."""
    s = away_from_cutoff(n, L, rcut, 41, 1e-4)
    rng = np.random.default_rng(5)
    gvec = rng.normal(size=(3, n))
    dpos, dq = pkg.realteil_vjp(pkg.HostSystem.from_dict(s), rcut, alpha, gvec)
    idx = rng.choice(n, 10, replace=False)
    fd_pos, fd_q = fd_realspace(orc, s, rcut, alpha, gvec, idx)
    print("real-space adjoint vs FD: pos %.2e q %.2e" % (rel_err(dpos[:, idx], fd_pos),
                                                        rel_err(dq[idx], fd_q)))
    assert rel_err(dpos[:, idx], fd_pos) < 1e-7
    assert rel_err(dq[idx], fd_q) < 1e-9


def test_realteil_vjp_device_form_accumulates(pkg):
    """The device entry point equals the host one and adds to its outputs."""
    import torch

    n, L, rcut, alpha = 500, 9.0, 3.0, 1.1
    s = random_system(n, L, seed=44, layout="gaussian")
    g = np.random.default_rng(6).normal(size=(3, n))
    want_pos, want_q = pkg.realteil_vjp(pkg.HostSystem.from_dict(s), rcut, alpha, g)
    solver = pkg.Solver(0, pkg.HostSystem.from_dict(s), pkg.parameters(16, 5, alpha))
    dev = {k: torch.tensor(np.asarray(s[k]), dtype=torch.float64, device="cuda") for k in "xyzq"}
    gd = torch.tensor(g.reshape(-1), dtype=torch.float64, device="cuda")
    dpos = torch.ones(3 * n, dtype=torch.float64, device="cuda")
    dq = torch.ones(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    solver.realteil_vjp_device(n, L, *[dev[k].data_ptr() for k in "xyzq"], rcut, alpha,
                               gd.data_ptr(), dpos.data_ptr(), dq.data_ptr())
    solver.synchronize()
    assert rel_err(dpos.cpu().numpy().reshape(3, n) - 1.0, want_pos) < 1e-12
    assert rel_err(dq.cpu().numpy() - 1.0, want_q) < 1e-12
    solver.close()


def test_total_force_autograd(pkg, orc):
    """loss = sum |F - F_target|^2 over total forces (k space + real space)
    through torch.autograd, against central differences of the oracle's total
    force; positions at bin centres (k space) and away from the cutoff (real
    space)."""
    import importlib

    import torch

    ag = importlib.import_module("p3m_standalone_amd.autograd")
    M, cao, alpha, L, n, rcut = 16, 5, 1.2, 8.0, 12, 3.0
    for k in range(200):
        s = snap_to_bins(interior_system(n, L, 60 + k, 1e-3), M)
        if np.min(np.abs(pair_distances(s) - rcut)) > 1e-3:
            break
    rng = np.random.default_rng(9)
    target = rng.normal(size=(3, n))
    solver = pkg.Solver(0, pkg.HostSystem.from_dict(s), pkg.parameters(M, cao, alpha))
    pos = torch.tensor(np.stack([s["x"], s["y"], s["z"]]), dtype=torch.float64, device="cuda",
                       requires_grad=True)
    q = torch.tensor(np.asarray(s["q"]), dtype=torch.float64, device="cuda", requires_grad=True)
    F = ag.total_force(pos, q, solver, rcut, alpha)
    G = orc.influence(0, M, cao, L, alpha)
    total = lambda t: oracle_force(orc, 0, t, M, cao, alpha, G, None) + orc.realteil(t, rcut, alpha)[0]
    Fo = total(s)
    assert rel_err(F.detach().cpu().numpy(), Fo) < 1e-12
    loss = ((F - torch.tensor(target, device="cuda")) ** 2).sum()
    loss.backward()
    gvec = 2.0 * (Fo - target)
    # k-space part by whole quantisation bins, real-space part with a small step
    fdk_pos, fdk_q = fd_gradient(orc, 0, s, M, cao, alpha, gvec)
    fdr_pos, fdr_q = fd_realspace(orc, s, rcut, alpha, gvec, np.arange(n))
    fd_pos, fd_q = fdk_pos + fdr_pos, fdk_q + fdr_q
    print("total-force adjoint vs FD: pos %.2e q %.2e" % (rel_err(pos.grad.cpu().numpy(), fd_pos),
                                                         rel_err(q.grad.cpu().numpy(), fd_q)))
    assert rel_err(pos.grad.cpu().numpy(), fd_pos) < 2e-7
    assert rel_err(q.grad.cpu().numpy(), fd_q) < 1e-9
    solver.close()
