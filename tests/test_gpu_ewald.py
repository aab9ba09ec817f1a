"""GPU tests of the sampled FP64 Ewald sum (p3m_hip_ewald_sample_device)."""
import numpy as np
import pytest

from conftest import load_fixture, random_system, rel_err

pytestmark = pytest.mark.gpu


def dev_system(torch, s):
    return {k: torch.tensor(np.asarray(s[k]), dtype=torch.float64, device="cuda") for k in "xyzq"}


def sample(pkg, torch, s, idx, alpha, rcut, kmax, alpha_split=0.0):
    d = dev_system(torch, s)
    ns = len(idx)
    di = torch.tensor(np.asarray(idx, dtype=np.int32), device="cuda")
    ft = torch.zeros(3 * ns, dtype=torch.float64, device="cuda")
    fk = torch.zeros_like(ft) if alpha_split > 0 else None
    torch.cuda.synchronize()
    pkg.ewald_sample_device(len(s["q"]), s["L"], *[d[k].data_ptr() for k in "xyzq"], ns,
                            di.data_ptr(), alpha, rcut, kmax, ft.data_ptr(), alpha_split,
                            fk.data_ptr() if fk is not None else None)
    return ft.cpu().numpy().reshape(3, ns), (fk.cpu().numpy().reshape(3, ns) if fk is not None else None)


def test_ewald_sample_reproduces_tests_wall(pkg):
    """All 300 particles of the reference's tests-wall fixture against its
    Ewald forces (accurate to ~1e-13 relative)."""
    import torch

    s, ref = load_fixture("tests_wall.npz")
    ref = np.asarray(ref).reshape(3, -1)
    n, L = len(s["q"]), float(s["L"])
    q2 = float(np.sum(np.asarray(s["q"]) ** 2))
    rcut = 0.49 * L
    alpha, kmax = pkg.ewald_parameters(n, L, q2, 1e-11, rcut)
    f, _ = sample(pkg, torch, s, np.arange(n), alpha, rcut, kmax)
    err = rel_err(f, ref)
    print("tests-wall: alpha %.4f kmax %d rel %.2e" % (alpha, kmax, err))
    assert err < 1e-8


def test_ewald_kspace_split_matches_oracle(pkg, orc):
    """The k-space split at alpha2 equals the oracle's Ewald k-space sum at
    alpha2 (the reference's ewald.c arithmetic, restated on the CPU)."""
    import torch

    s = random_system(400, 9.0, seed=12)
    n, L = 400, 9.0
    q2 = float(np.sum(s["q"] ** 2))
    rcut = 4.4
    alpha, kmax = pkg.ewald_parameters(n, L, q2, 1e-10, rcut)
    alpha2 = 1.6  # erfc(alpha2 rcut) ~ 1e-23: the split is untruncated
    idx = np.arange(0, n, 7)
    _, fk = sample(pkg, torch, s, idx, alpha, rcut, kmax, alpha2)
    ref = orc.ewald_kspace(s, alpha2, 24)[:, idx]
    err = rel_err(fk, ref)
    print("k-space split vs oracle Ewald: %.2e" % err)
    assert err < 1e-8


def test_ewald_sample_subset_and_alpha_independence(pkg):
    """Forces on a subset equal those rows of the full evaluation, and the
    converged total does not depend on the splitting parameter."""
    import torch

    s = random_system(600, 10.0, seed=4, layout="gaussian")
    n, L = 600, 10.0
    q2 = float(np.sum(s["q"] ** 2))
    a1, k1 = pkg.ewald_parameters(n, L, q2, 1e-10, 4.9)
    full, _ = sample(pkg, torch, s, np.arange(n), a1, 4.9, k1)
    idx = np.array([3, 77, 599, 0, 250])
    sub, _ = sample(pkg, torch, s, idx, a1, 4.9, k1)
    assert rel_err(sub, full[:, idx]) < 1e-14
    a2, k2 = pkg.ewald_parameters(n, L, q2, 1e-10, 3.5)
    other, _ = sample(pkg, torch, s, np.arange(n), a2, 3.5, k2)
    print("alpha %.3f vs %.3f: %.2e" % (a1, a2, rel_err(other, full)))
    assert rel_err(other, full) < 1e-8
