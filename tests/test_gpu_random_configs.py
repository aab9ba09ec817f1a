"""Seeded random configurations through the whole force path, against the
oracle: method (ids 0-3, 6, 7), particle count, box, even 13-smooth mesh,
order, alpha and rcut (up to L / 2, so both cell sizes and the full-shell
kernel for boxes shorter than 2.5 rcut), uniform or clustered charges
(clusters centred anywhere, so they wrap), equal or unequal charges.

FP64: k space within 1e-12 of the oracle (realpart.c / p3m-*.c restated),
the self-force subtraction of ad within 1e-11, real space within 1e-13 and
its energy within 1e-12.  Mixed: real space within the larger of 1e-6 and
the fixed-point bound of its closest pair (3 (2u + sqrt(3) s / r_min), s the
quantum, 2^-29 of the cell side: cells of side rcut in a sparse box and a
close pair make it the larger), k space within 2e-5 of the FP64 GPU
result."""
import numpy as np
import pytest

from conftest import mixed_realspace_bound, rel_err

pytestmark = pytest.mark.gpu

MESHES = [8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 30, 32, 36, 40]


def make_config(seed):
    rng = np.random.default_rng(1000 + seed)
    method = int(rng.choice([0, 1, 2, 3, 6, 7]))
    n = int(rng.integers(30, 3001))
    L = float(rng.uniform(4.0, 25.0))
    M = int(rng.choice(MESHES))
    cao = int(rng.integers(1, 8))
    alpha = float(rng.uniform(0.4, 2.0)) * 10.0 / L
    rcut = float(rng.uniform(0.15, 0.49)) * L
    layout = rng.choice(["uniform", "cluster", "mixed"])
    if layout == "uniform":
        p = rng.uniform(0.0, L, (n, 3))
    else:
        sigma = float(rng.uniform(0.03, 0.2)) * L
        k = n if layout == "cluster" else n // 2
        c = rng.normal(0.0, sigma, (k, 3)) + rng.uniform(0.0, L, 3)
        p = np.concatenate([c, rng.uniform(0.0, L, (n - k, 3))])
    p = np.mod(p, L)
    p[p >= L] = 0.0
    if rng.integers(2) == 0:
        q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
        if n % 2:
            q[-1] = 0.0
    else:
        q = rng.uniform(-2.0, 2.0, n)
        q -= q.mean()
    s = {"x": p[:, 0].copy(), "y": p[:, 1].copy(), "z": p[:, 2].copy(), "q": q, "L": L}
    return s, dict(method=method, n=n, L=L, M=M, cao=cao, alpha=alpha, rcut=rcut, layout=str(layout))


@pytest.mark.parametrize("seed", range(16))
def test_random_config(pkg, orc, seed):
    import torch

    s, c = make_config(seed)
    print("seed %d: %s" % (seed, c), flush=True)
    method, n, L, M, cao, alpha, rcut = (c[k] for k in ("method", "n", "L", "M", "cao", "alpha", "rcut"))
    system = pkg.HostSystem.from_dict(s)

    # FP64 k space against the oracle
    solver = pkg.Solver(method, system, pkg.parameters(M, cao, alpha, rcut))
    forces = solver.kspace_force()
    fk_ref, f_ref = orc.kspace_force(method, s, M, cao, alpha)
    fk64 = forces.kspace().copy()
    if np.linalg.norm(fk_ref) == 0.0:
        # ad at cao 1: the order-1 derivative weight is 0
        assert np.all(fk64 == 0.0)
    else:
        err = rel_err(fk64, fk_ref)
        print("  k space FP64 rel %.2e" % err)
        assert err < 1e-12
        if method in (2, 7):
            assert rel_err(forces.total(), f_ref) < 1e-11

    # mixed k space against FP64
    solver.set_precision(pkg.MIXED)
    fk32 = solver.kspace_force().kspace()
    if np.linalg.norm(fk64) > 0.0:
        err = rel_err(fk32, fk64)
        print("  k space mixed vs FP64 rel %.2e" % err)
        assert err < 2e-5
    solver.close()

    # real space, FP64 (host entry) and mixed (device entry), against the oracle
    fr_ref, e_ref = orc.realteil(s, rcut, alpha)
    system = pkg.HostSystem.from_dict(s)
    fr64 = pkg.realteil(system, rcut, alpha).real()
    err = rel_err(fr64, fr_ref)
    print("  real space FP64 rel %.2e" % err)
    assert err < 1e-13
    assert abs(system.c.energy - e_ref) <= 1e-12 * abs(e_ref) + 1e-300

    solver = pkg.Solver(method, pkg.HostSystem.from_dict(s), pkg.parameters(M, cao, alpha, rcut), pkg.MIXED)
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
    if np.linalg.norm(fr_ref) > 0.0:
        err = rel_err(got, fr_ref)
        bound = mixed_realspace_bound(s, rcut)
        print("  real space mixed rel %.2e (bound %.2e)" % (err, bound))
        assert err < bound
        # the pair energies are summed in FP32 per lane: their error is
        # relative to the sum of |pair energy|, bounded by the energy of the
        # same system with all charges made positive
        s_abs = dict(s, q=np.abs(s["q"]))
        _, e_abs = orc.realteil(s_abs, rcut, alpha)
        assert abs(e - e_ref) <= 1e-5 * abs(e_ref) + 1e-6 * abs(e_abs)
