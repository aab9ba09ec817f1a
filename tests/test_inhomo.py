"""Inhomogeneous error estimate and error map (p3m-common.c:402-719): the
oracle pinned by the reference's literal double sum, the vanishing
homogeneous kernel and the uniform-density closed form; the GPU against it."""
import numpy as np
import pytest

from conftest import random_system


def test_oracle_factorised_kernel_equals_the_double_sum(orc):
    """The aliasing kernels factorise over (n, n2): the reference's O(M^6)
    double sum (p3m-common.c:571-612), evaluated literally, gives the same
    K^2(r)."""
    for M, L, alpha, cao in [(6, 5.0, 1.1, 3), (8, 7.0, 0.9, 2)]:
        fast = orc.inhomo_k2(M, L, alpha, cao, 1)
        slow = orc.inhomo_k2(M, L, alpha, cao, 1, factorised=False)
        assert np.max(np.abs(fast - slow)) <= 1e-12 * np.max(np.abs(slow))


def test_oracle_homogeneous_kernel_vanishes(orc):
    """Without aliasing copies the optimal influence function is G_c / U^2,
    so U^2 G - G_c = 0: the homogeneous kernel is rounding only."""
    M, L, alpha, cao = 8, 7.0, 0.9, 3
    homo = orc.inhomo_k2(M, L, alpha, cao, 0)
    full = orc.inhomo_k2(M, L, alpha, cao, 1)
    assert np.max(homo) <= 1e-20 * np.max(full)


def test_oracle_uniform_density_closed_form(orc):
    """A uniform rho^2 = q2 / M^3 has only the k = 0 component q2, so the
    error density is q2 sum_r K^2(r) everywhere, and with sum_stencil W = 1
    the estimate is sqrt(q2^2 sum_r K^2(r) / N)."""
    s = random_system(60, 5.0, seed=8)
    M, cao, alpha = 8, 3, 1.2
    q2 = float(np.sum(s["q"] ** 2))
    k3 = orc.inhomo_k2(M, s["L"], alpha, cao, 1)
    want = np.sqrt(q2 * q2 * np.sum(k3) / len(s["q"]))
    assert orc.error_inhomo(s, alpha, M, cao, 1, uniform=True) == pytest.approx(want, rel=1e-12)


def test_oracle_error_map_conserves_the_squared_error(orc):
    s = random_system(50, 6.0, seed=9)
    rng = np.random.default_rng(2)
    fk, fr = rng.normal(size=(3, 50)), rng.normal(size=(3, 50))
    em = orc.error_map(s, fk, fr, 8, 4)
    assert np.sum(em) == pytest.approx(np.sum((fr - fk) ** 2), rel=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("n,L,M,cao,mc,uniform", [
    (200, 5.0, 8, 3, 1, False), (200, 5.0, 8, 3, 1, True), (400, 8.0, 16, 5, 1, False),
    (150, 6.0, 12, 4, 2, False)])
def test_error_inhomo_matches_oracle(pkg, orc, n, L, M, cao, mc, uniform):
    s = random_system(n, L, seed=31, layout="gaussian")
    alpha = 1.0
    got = pkg.error_inhomo(pkg.HostSystem.from_dict(s), alpha, mesh=M, cao=cao, mc=mc,
                           uniform=uniform)
    want = orc.error_inhomo(s, alpha, M, cao, mc, uniform=uniform)
    print("inhomogeneous estimate M %d cao %d mc %d uniform %d: %.10e vs %.10e" % (
        M, cao, mc, uniform, got, want))
    assert got == pytest.approx(want, rel=1e-10)


@pytest.mark.gpu
def test_error_inhomo_density_file(pkg, orc, tmp_path):
    s = random_system(100, 6.0, seed=32)
    M, cao = 8, 3
    path = str(tmp_path / "inhomo.dat")
    pkg.error_inhomo(pkg.HostSystem.from_dict(s), 1.0, mesh=M, cao=cao, mc=1, out_file=path)
    rows = np.loadtxt(path)
    assert rows.shape == (M ** 3, 9)
    np.testing.assert_array_equal(rows[:8, :3], [[0, 0, k] for k in range(8)])
    sq = dict(s)
    sq["q"] = s["q"] ** 2
    rho2 = orc.assign_mesh(0, sq, M, cao)[0::2]
    np.testing.assert_allclose(rows[:, 3], rho2, rtol=1e-5, atol=1e-300)  # "%e": 7 digits


@pytest.mark.gpu
def test_error_map_matches_oracle(pkg, orc):
    s = random_system(300, 7.0, seed=33)
    rng = np.random.default_rng(4)
    fk, fr = rng.normal(size=(3, 300)), rng.normal(size=(3, 300))
    for M, cao in [(8, 1), (16, 4)]:
        got = pkg.error_map(pkg.HostSystem.from_dict(s), fk, fr, mesh=M, cao=cao)
        want = orc.error_map(s, fk, fr, M, cao)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14 * np.max(want))


def test_error_inhomo_argument_checks(pkg):
    """Bad meshes, orders and copy counts fail before any device work."""
    s = pkg.HostSystem.from_dict(random_system(10, 5.0, seed=1))
    for kw in [dict(mesh=7), dict(cao=8), dict(mc=-1)]:
        with pytest.raises(RuntimeError):
            pkg.error_inhomo(s, 1.0, **{**dict(mesh=8, cao=3, mc=1), **kw})
