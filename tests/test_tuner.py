"""The tuner (Tune, tuning.c:250-370): the oracle's scan pinned by brute force
and by the real-space estimate it inverts, the library's scan against it with
an injected timer, and one scan timed on the GPU."""
import math

import numpy as np
import pytest

from conftest import random_system


def cost_model(N, a=1e-9, b=1e-9):
    """A timer increasing in mesh and order: assignment-like N cao^3 plus
    FFT-like M^3 log M seconds."""
    return lambda M, c: a * N * c ** 3 + b * M ** 3 * math.log2(M)


_errors = {}


def cached_error(orc, method, N, L, q2, rcut):
    """The oracle's Error (direct sums, seconds at M = 40) memoised per
    candidate."""
    def err(M, c, a):
        key = (method, N, L, q2, rcut, M, c, a)
        if key not in _errors:
            _errors[key] = orc.total_error(method, N, L, q2, M, c, a, rcut)
        return _errors[key]
    return err


def brute_force(orc, method, N, L, q2, rcut, cao, precision, timer, mmax=44):
    """Cheapest candidate meeting the precision among meshes <= mmax; with
    the cost model every larger mesh costs more than the winner (checked)."""
    alpha = orc.tune_alpha(N, L, q2, rcut, precision)
    orders = [cao] if cao else range(2, 8)
    err = cached_error(orc, method, N, L, q2, rcut)
    ok = [(timer(M, c), M, c) for M in orc.TUNE_MESHES if M <= mmax for c in orders
          if err(M, c, alpha) <= precision]
    if not ok:
        return None
    best = min(ok)
    nxt = min(M for M in orc.TUNE_MESHES if M > mmax)
    assert best[0] < timer(nxt, 1)
    return best


def test_oracle_alpha_inverts_realspace_estimate(orc):
    """tuning.c:293 makes the real-space estimate (realpart.c:343-346) equal
    to precision / sqrt(2)."""
    for N, L, q2, rc, prec in [(1000, 10.0, 1000.0, 3.0, 1e-4), (10 ** 5, 46.4, 1e5, 3.0, 1e-3),
                               (300, 7.0, 75.0, 2.5, 1e-6)]:
        a = orc.tune_alpha(N, L, q2, rc, prec)
        assert orc.realspace_error(N, L, q2, rc, a) == pytest.approx(prec / math.sqrt(2.0), rel=1e-12)


# The ad-family estimates are direct sums; the ik estimates of meshes without
# a table row start from -1 (p3m-ik.c:267-297) and meet no precision, so a
# scan of them runs through every mesh.
@pytest.mark.parametrize("method", [2, 3, 7])
@pytest.mark.parametrize("precision", [1e-3, 1e-4])
def test_oracle_scan_finds_the_optimum(orc, method, precision):
    """With a timer increasing in mesh and order and an estimate decreasing
    in both, the scan's pruning loses nothing: it returns the cheapest
    candidate that meets the precision (brute force over all of them)."""
    N, L, q2, rc = 1000, 10.0, 1000.0, 3.0
    timer = cost_model(N)
    got = orc.tune_scan(method, N, L, q2, rc, 0, precision, timer,
                        error=cached_error(orc, method, N, L, q2, rc))
    want = brute_force(orc, method, N, L, q2, rc, 0, precision, timer)
    assert want is not None and got is not None
    assert (got[0], got[1]) == (want[1], want[2])
    assert got[3] <= precision


def test_oracle_scan_fixed_order_and_unreachable_precision(orc):
    N, L, q2, rc = 1000, 10.0, 1000.0, 3.0
    timer = cost_model(N)
    got = orc.tune_scan(3, N, L, q2, rc, 5, 1e-4, timer, error=cached_error(orc, 3, N, L, q2, rc))
    want = brute_force(orc, 3, N, L, q2, rc, 5, 1e-4, timer)
    assert got[1] == 5 and (got[0], got[1]) == (want[1], want[2])
    # an estimate that never meets the precision: nothing is returned
    assert orc.tune_scan(3, N, L, q2, rc, 0, 1e-4, timer, error=lambda M, c, a: 1.0) is None


def test_oracle_scan_stops_on_a_slow_cheapest_order(orc):
    """A mesh whose order-2 time exceeds the best ends the scan: with a timer
    that jumps at M = 24, no mesh beyond it is timed."""
    N, L, q2, rc = 1000, 10.0, 1000.0, 3.0
    seen = []

    def timer(M, c):
        seen.append(M)
        return 1.0 if M >= 24 else 1e-3 * c

    orc.tune_scan(3, N, L, q2, rc, 0, 1e-3, timer, error=cached_error(orc, 3, N, L, q2, rc))
    assert max(seen) == 24


def host_system(pkg, n, L, seed):
    s = random_system(n, L, seed=seed)
    return pkg.HostSystem.from_dict(s), s


@pytest.mark.gpu
@pytest.mark.parametrize("method", [2, 3, 7])
@pytest.mark.parametrize("precision", [1e-3, 1e-4])
def test_scan_matches_oracle_with_injected_timer(pkg, orc, method, precision):
    """The library's scan (p3m_hip_tune_scan) with the cost model as its timer
    visits the same candidates and returns the oracle's winner; its error
    estimates run on the GPU."""
    n, L, rc = 1000, 10.0, 3.0
    system, s = host_system(pkg, n, L, 21)
    q2 = float(np.sum(s["q"] ** 2))
    timer = cost_model(n)
    params = pkg.parameters(8, 0, 0.0, rc)
    p, stats = pkg.tune(method, system, params, precision,
                        timer=lambda mid, M, c, a: timer(M, c))
    want = orc.tune_scan(method, n, L, q2, rc, 0, precision, timer)
    print("method %d prec %.0e: library (%d, %d) oracle (%d, %d)" % (method, precision, p.mesh,
                                                                     p.cao, want[0], want[1]))
    assert (p.mesh, p.cao) == (want[0], want[1])
    assert p.alpha == pytest.approx(want[2], rel=1e-14)
    assert p.precision == pytest.approx(want[3], rel=1e-9)
    assert stats["t"]["min"] == pytest.approx(want[4], rel=1e-14)
    assert p.ip == p.cao - 1 and p.cao3 == p.cao ** 3


@pytest.mark.gpu
def test_scan_without_success_leaves_parameters(pkg):
    system, s = host_system(pkg, 1000, 10.0, 22)
    params = pkg.parameters(16, 0, 0.5, 3.0)
    # no candidate up to M = 300 reaches 1e-30: every mesh is tried, none wins
    p, stats = pkg.tune(3, system, params, 1e-30, timer=lambda mid, M, c, a: 1e-3 * M)
    assert stats["t"]["avg"] == -1.0
    assert (p.mesh, p.cao, p.alpha) == (16, 0, 0.5)


@pytest.mark.gpu
@pytest.mark.parametrize("form", [0, 1])
def test_tune_on_the_gpu(pkg, form):
    """A scan timed by 50 Kspace_force calls per candidate (time_full): the
    winner meets the precision, alpha is the real-space one, and the stage
    statistics are consistent."""
    n, L, rc, prec = 1000, 10.0, 3.0, 1e-4
    system, s = host_system(pkg, n, L, 23)
    q2 = float(np.sum(s["q"] ** 2))
    p, st = pkg.tune(3, system, pkg.parameters(8, 0, 0.0, rc), prec, form=form)
    print("tuned (form %d): mesh %d cao %d alpha %.4f err %.2e t %.3e s" % (
        form, p.mesh, p.cao, p.alpha, p.precision, st["t"]["avg"]))
    assert st["t"]["avg"] > 0
    assert p.precision <= prec
    assert p.alpha == pytest.approx(pkg.tune_alpha(n, L, q2, rc, prec), rel=1e-14)
    tot, _ = pkg.error_estimate(3, system, p)
    assert tot == pytest.approx(p.precision, rel=1e-12)
    for k in ("t_c", "t_g", "t_f"):
        t = st[k]
        assert t["n"] == 50 and 0 < t["min"] <= t["avg"] <= t["max"]
    assert st["t"]["avg"] == pytest.approx(st["t_c"]["avg"] + st["t_g"]["avg"] + st["t_f"]["avg"])
    assert st["t"]["min"] == pytest.approx(st["t_c"]["min"] + st["t_g"]["min"] + st["t_f"]["min"])
