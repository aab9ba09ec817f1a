"""Real space on cell grids (the half-shell kernel: each unordered pair once,
the reaction added to the partner) against the oracle's Realteil
(realpart.c:124-167): dense cells of several 64-particle chunks next to
sparse ones, clusters that wrap across the periodic boundary, both cell
sizes (side rcut / 2, and side rcut for sparse systems), the device entry
point accumulating into the caller's arrays, and the mixed form."""
import math

import numpy as np
import pytest

from conftest import cell_geometry, rel_err

pytestmark = pytest.mark.gpu


def corner_cluster(n_cluster, n_background, L, sigma, seed, centre=(0.0, 0.0, 0.0)):
    """A Gaussian cluster about `centre` (folded into the box, so at a corner
    it spans the periodic boundary on all three axes) plus a uniform
    background; charges +-1, neutral."""
    rng = np.random.default_rng(seed)
    c = rng.normal(0.0, sigma, (n_cluster, 3)) + np.asarray(centre)
    b = rng.uniform(0.0, L, (n_background, 3))
    p = np.mod(np.concatenate([c, b]), L)
    p[p >= L] = 0.0
    n = n_cluster + n_background
    q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    if n % 2:
        q[-1] = 0.0
    return {"x": p[:, 0].copy(), "y": p[:, 1].copy(), "z": p[:, 2].copy(), "q": q, "L": float(L)}


def cell_counts(s, rcut):
    """Particles per cell of the library's grid."""
    L = s["L"]
    nc, reach = cell_geometry(len(s["q"]), L, rcut)
    c = np.minimum((np.stack([s["x"], s["y"], s["z"]], 1) * nc / L).astype(int), nc - 1)
    return nc, reach, np.bincount((c[:, 0] * nc + c[:, 1]) * nc + c[:, 2], minlength=nc ** 3)


@pytest.mark.parametrize("L,n_background,centre,reach", [
    (14.0, 2500, (0.0, 0.0, 0.0), 2), (14.0, 2500, (7.0, 0.3, 13.9), 2),
    (30.0, 1000, (0.0, 0.0, 0.0), 1), (30.0, 1000, (29.5, 15.0, 0.2), 1)])
def test_wrapped_dense_cluster(pkg, orc, L, n_background, centre, reach):
    """A cluster across the box corner / faces: cells of up to several
    hundred particles (several chunks; the candidate search on the matrix
    cores), sparse background cells (the vector search), image shifts inside
    dense cell pairs; with cells of side rcut / 2 and of side rcut."""
    rcut, alpha = 3.0, 1.0
    s = corner_cluster(2500 if reach == 2 else 2000, n_background, L, 0.9, seed=31, centre=centre)
    nc, got_reach, cnt = cell_counts(s, rcut)
    assert got_reach == reach
    assert nc >= 2 * reach + 1 and cnt.max() > 128 and np.sum((cnt > 0) & (cnt <= 4)) > 50
    system = pkg.HostSystem.from_dict(s)
    forces = pkg.realteil(system, rcut, alpha)
    fr_ref, e_ref = orc.realteil(s, rcut, alpha)
    err = rel_err(forces.real(), fr_ref)
    print("reach %d, max cell %d, rel err %.2e, energy %.15e vs %.15e" % (
        reach, cnt.max(), err, system.c.energy, e_ref))
    assert err < 1e-13
    assert abs(system.c.energy - e_ref) <= 1e-12 * abs(e_ref)


def test_single_dense_cell_of_many_chunks(pkg, orc):
    """One cell of ~600 particles (ten chunks: the item's own chunk, later
    chunks of its cell, earlier ones left to their own items) in an
    otherwise sparse box."""
    L, rcut, alpha = 12.0, 3.0, 1.2
    rng = np.random.default_rng(5)
    nc, _ = cell_geometry(900, L, rcut)
    side = L / nc
    cell = rng.uniform(1 * side + 1e-3, 2 * side - 1e-3, (600, 3))
    rest = rng.uniform(0.0, L, (300, 3))
    p = np.concatenate([cell, rest])
    q = np.where(np.arange(900) % 2 == 0, 1.0, -1.0)
    s = {"x": p[:, 0].copy(), "y": p[:, 1].copy(), "z": p[:, 2].copy(), "q": q, "L": L}
    _, _, cnt = cell_counts(s, rcut)
    assert cnt.max() >= 600
    system = pkg.HostSystem.from_dict(s)
    forces = pkg.realteil(system, rcut, alpha)
    fr_ref, e_ref = orc.realteil(s, rcut, alpha)
    assert rel_err(forces.real(), fr_ref) < 1e-13
    assert abs(system.c.energy - e_ref) <= 1e-12 * abs(e_ref)


def test_newton_third_law(pkg):
    """Each unordered pair enters both partners with opposite sign, so the
    real-space forces of a periodic system sum to zero (to rounding)."""
    L, rcut, alpha = 14.0, 3.0, 1.0
    s = corner_cluster(2000, 500, L, 1.2, seed=8, centre=(7.0, 7.0, 7.0))
    forces = pkg.realteil(pkg.HostSystem.from_dict(s), rcut, alpha)
    f = forces.real()
    assert np.abs(f.sum(axis=1)).max() < 1e-10 * np.abs(f).sum()


@pytest.mark.parametrize("prec_name", ["FP64", "MIXED"])
def test_device_form_accumulates_and_mixed_is_close(pkg, orc, prec_name):
    """The device entry point (both precision forms) adds the real-space
    forces to the caller's array and the energy to its scalar; the mixed form
    has the FP64 form's pair set (its energy agrees to FP32 rounding) and
    forces within 1e-6 of the force norm (its pair energies are summed in
    FP32 per lane, hence 1e-5 on the energy)."""
    import torch

    L, rcut, alpha = 14.0, 3.0, 1.0
    s = corner_cluster(2500, 700, L, 0.9, seed=12)
    n = len(s["q"])
    fr_ref, e_ref = orc.realteil(s, rcut, alpha)
    prec = getattr(pkg, prec_name)
    solver = pkg.Solver(3, pkg.HostSystem.from_dict(s), pkg.parameters(16, 5, alpha, rcut), prec)
    dev = {k: torch.tensor(np.asarray(s[k]), dtype=torch.float64, device="cuda") for k in "xyzq"}
    fr = torch.full((3 * n,), 2.0, dtype=torch.float64, device="cuda")
    en = torch.full((1,), 5.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    solver.realteil_device(n, L, *[dev[k].data_ptr() for k in "xyzq"], rcut, alpha, fr.data_ptr(),
                           en.data_ptr())
    solver.synchronize()
    got = fr.cpu().numpy().reshape(3, n) - 2.0
    e = float(en.item()) - 5.0
    solver.close()
    err = rel_err(got, fr_ref)
    print("%s: rel err %.2e, energy rel %.2e" % (prec_name, err, abs(e - e_ref) / abs(e_ref)))
    if prec_name == "FP64":
        assert err < 1e-13
        assert abs(e - e_ref) <= 1e-12 * abs(e_ref)
    else:
        assert err < 1e-6
        assert abs(e - e_ref) <= 1e-5 * abs(e_ref)


@pytest.mark.parametrize("m,reach", [(400, 1), (1100, 2)])
def test_pairs_at_the_cutoff(pkg, orc, m, reach):
    """Pairs placed within 1e-9 of rcut on either side, in all directions and
    across cell and box boundaries: the candidate search (FP16 operands in
    units of the cell side, with a margin) passes all of them and the exact
    FP64 test keeps exactly the reference's ones; both cell sizes."""
    L, rcut, alpha = 12.0, 3.0, 0.8
    assert cell_geometry(2 * m, L, rcut)[1] == reach
    rng = np.random.default_rng(77)
    a = rng.uniform(0.0, L, (m, 3))
    u = rng.normal(size=(m, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    eps = np.where(np.arange(m) % 2 == 0, -1e-9, 1e-9)
    b = a + (rcut * (1.0 + eps))[:, None] * u
    p = np.mod(np.concatenate([a, b]), L)
    p[p >= L] = 0.0
    q = np.concatenate([np.ones(m), -np.ones(m)])
    s = {"x": p[:, 0].copy(), "y": p[:, 1].copy(), "z": p[:, 2].copy(), "q": q, "L": L}
    system = pkg.HostSystem.from_dict(s)
    forces = pkg.realteil(system, rcut, alpha)
    fr_ref, e_ref = orc.realteil(s, rcut, alpha)
    assert rel_err(forces.real(), fr_ref) < 1e-13
    assert abs(system.c.energy - e_ref) <= 1e-12 * abs(e_ref)


@pytest.mark.parametrize("m,reach", [(400, 1), (1100, 2)])
def test_pairs_at_the_cutoff_mixed(pkg, orc, m, reach):
    """The mixed form on the pairs of test_pairs_at_the_cutoff: their FP32
    squared distances fall inside the band around rcut^2, where the FP64 test
    decides on the FP64 positions with the pair's image shift (many of the
    pairs wrap across the box).  One pair decided the other way would move
    its two particles' forces by the pair force at rcut, f_c = 3.4e-4: each
    particle's error is held to 1e-5 of its force plus f_c / 10 (FP32
    rounding stays near 1e-7 of the force), besides the mixed form's 1e-6
    on the norm."""
    import torch

    L, rcut, alpha = 12.0, 3.0, 0.8
    assert cell_geometry(2 * m, L, rcut)[1] == reach
    rng = np.random.default_rng(78)
    a = rng.uniform(0.0, L, (m, 3))
    u = rng.normal(size=(m, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    eps = np.where(np.arange(m) % 2 == 0, -1e-9, 1e-9)
    b = a + (rcut * (1.0 + eps))[:, None] * u
    p = np.mod(np.concatenate([a, b]), L)
    p[p >= L] = 0.0
    q = np.concatenate([np.ones(m), -np.ones(m)])
    s = {"x": p[:, 0].copy(), "y": p[:, 1].copy(), "z": p[:, 2].copy(), "q": q, "L": L}
    n = 2 * m
    fr_ref, e_ref = orc.realteil(s, rcut, alpha)
    solver = pkg.Solver(3, pkg.HostSystem.from_dict(s), pkg.parameters(16, 5, alpha, rcut), pkg.MIXED)
    dev = {k: torch.tensor(np.asarray(s[k]), dtype=torch.float64, device="cuda") for k in "xyzq"}
    fr = torch.zeros(3 * n, dtype=torch.float64, device="cuda")
    en = torch.zeros(1, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    solver.realteil_device(n, L, *[dev[k].data_ptr() for k in "xyzq"], rcut, alpha, fr.data_ptr(),
                           en.data_ptr())
    solver.synchronize()
    got = fr.cpu().numpy().reshape(3, n)
    e = float(en.item())
    solver.close()
    err = rel_err(got, fr_ref)
    fc = (math.erfc(alpha * rcut) / rcut + 2.0 * alpha / math.sqrt(math.pi) * math.exp(-(alpha * rcut) ** 2)) / rcut**2
    dev_i = np.linalg.norm(got - fr_ref, axis=0)
    bound_i = 1e-5 * np.linalg.norm(fr_ref, axis=0) + 0.1 * fc
    print("mixed at the cutoff (m %d, reach %d): rel err %.2e, worst particle %.2e of its bound"
          % (m, reach, err, (dev_i / bound_i).max()))
    assert (dev_i <= bound_i).all()
    assert err < 1e-6
    assert abs(e - e_ref) <= 1e-5 * abs(e_ref)


def test_stage_ring_pair_kernel_times(pkg):
    """The per-step pair-kernel times of the stage ring: positive, within the
    step's real-space stage (which also holds the cell list), and 0 for a
    step without real space."""
    import torch

    L, rcut, alpha = 14.0, 3.0, 1.0
    s = corner_cluster(2500, 700, L, 0.9, seed=14)
    n = len(s["q"])
    solver = pkg.Solver(3, pkg.HostSystem.from_dict(s), pkg.parameters(16, 5, alpha, rcut), pkg.MIXED)
    dev = {k: torch.tensor(np.asarray(s[k]), dtype=torch.float64, device="cuda") for k in "xyzq"}
    f, fk, fr = (torch.zeros(3 * n, dtype=torch.float64, device="cuda") for _ in range(3))
    ptr = [dev[k].data_ptr() for k in "xyzq"]
    torch.cuda.synchronize()
    solver.stage_ring_start(3)
    for rc in (rcut, rcut, 0.0):
        solver.calculate_forces_device(n, L, *ptr, rc, alpha, f.data_ptr(), fk.data_ptr(), fr.data_ptr(),
                                       None)
    stages = solver.stage_ring_read(3)
    pairs = solver.stage_ring_read_pairs(3)
    solver.stage_ring_start(0)
    solver.close()
    print(stages[:, 4], pairs)
    assert len(pairs) == 3
    assert all(0.0 < pairs[k] <= stages[k, 4] for k in (0, 1))
    assert pairs[2] == 0.0
