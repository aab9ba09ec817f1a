"""GPU tests of batches of independent systems (p3m_hip_batch_*).

Every system of a batch must get exactly the forces the oracle gives for it
alone: the batched kernels only change which blocks and slots serve which
system, not the arithmetic.
"""
import numpy as np
import pytest

from conftest import random_system, rel_err

pytestmark = pytest.mark.gpu

SMOOTH_EVEN = [16, 18, 20, 22, 24, 28, 30, 32, 36, 40, 42, 44, 48]


def make_batch_systems(nsys, seed, nmin=200, nmax=1500, density=0.5, rcut=3.0, eps=1e-4,
                       meshes=SMOOTH_EVEN):
    """Uneven systems in the style of the batched configuration: N uniform,
    L = (N / density)^(1/3), mesh from the even 11-smooth sizes, cao 3..7,
    alpha from the tuner's real-space formula; neutral +-1 charges."""
    rng = np.random.default_rng(seed)
    systems, data = [], []
    for s in range(nsys):
        n = 2 * int(rng.integers(nmin // 2, nmax // 2 + 1))
        L = (n / density) ** (1.0 / 3.0)
        M = int(rng.choice(meshes))
        cao = int(rng.integers(3, 8))
        d = random_system(n, L, seed=seed * 1000 + s)
        q2 = float(np.sum(d["q"] ** 2))
        alpha = float(np.sqrt(-np.log(eps * np.sqrt(n * rcut * L ** 3) / (2 * np.sqrt(2) * q2))) / rcut)
        systems.append(dict(n=n, L=L, M=M, cao=cao, alpha=alpha, rcut=rcut))
        data.append(d)
    return systems, data


def device_arrays(torch, data):
    cat = {k: np.concatenate([np.asarray(d[k]) for d in data]) for k in "xyzq"}
    return {k: torch.tensor(v, dtype=torch.float64, device="cuda") for k, v in cat.items()}


def run_batch(torch, batch, dev, graph=False):
    n = batch.ntot
    f = torch.zeros(3 * n, dtype=torch.float64, device="cuda")
    fk = torch.zeros_like(f)
    fr = torch.zeros_like(f)
    en = torch.zeros(len(batch.systems), dtype=torch.float64, device="cuda")
    ptr = [dev[k].data_ptr() for k in "xyzq"]
    torch.cuda.synchronize()
    if graph:
        batch.capture(*ptr, f.data_ptr(), fk.data_ptr(), fr.data_ptr(), en.data_ptr())
        batch.replay()
    else:
        batch.calculate_forces_device(*ptr, f.data_ptr(), fk.data_ptr(), fr.data_ptr(), en.data_ptr())
    batch.synchronize()
    return (f.cpu().numpy().reshape(3, n), fk.cpu().numpy().reshape(3, n),
            fr.cpu().numpy().reshape(3, n), en.cpu().numpy())


def per_system(arr, offsets, s):
    return arr[:, offsets[s]:offsets[s + 1]]


@pytest.mark.parametrize("method", [0, 1, 2, 3, 6, 7])
def test_batch_matches_oracle_per_system(pkg, orc, method):
    import torch

    systems, data = make_batch_systems(6, seed=10 + method)
    batch = pkg.Batch(method, systems)
    dev = device_arrays(torch, data)
    f, fk, fr, en = run_batch(torch, batch, dev)
    for s, (S, d) in enumerate(zip(systems, data)):
        fk_ref, fself = orc.kspace_force(method, d, S["M"], S["cao"], S["alpha"])
        fr_ref, e_ref = orc.realteil(d, S["rcut"], S["alpha"])
        e_k = rel_err(per_system(fk, batch.offsets, s), fk_ref)
        e_r = rel_err(per_system(fr, batch.offsets, s), fr_ref)
        e_t = rel_err(per_system(f, batch.offsets, s), fk_ref + fr_ref + fself)
        print("method %d system %d (n %d M %d cao %d): k %.2e r %.2e total %.2e" %
              (method, s, S["n"], S["M"], S["cao"], e_k, e_r, e_t))
        assert e_k < 1e-12
        assert e_r < 1e-13
        assert e_t < 1e-12
        assert abs(en[s] - e_ref) <= 1e-12 * abs(e_ref)
    batch.close()


def test_batch_equals_single_system_solver(pkg):
    """A batch of one system gives the single-system Calculate_forces."""
    import torch

    systems, data = make_batch_systems(1, seed=3, nmin=2000, nmax=2000)
    S, d = systems[0], data[0]
    batch = pkg.Batch(3, systems)
    dev = device_arrays(torch, data)
    f, fk, fr, _ = run_batch(torch, batch, dev)
    solver = pkg.Solver(3, pkg.HostSystem.from_dict(d), pkg.parameters(S["M"], S["cao"], S["alpha"], S["rcut"]))
    ref = solver.calculate_forces()
    solver.close()
    assert rel_err(f, ref.total()) < 1e-13
    assert rel_err(fk, ref.kspace()) < 1e-13
    batch.close()


@pytest.mark.parametrize("method", [1, 7])
def test_batch_mixed_precision(pkg, method):
    import torch

    systems, data = make_batch_systems(12, seed=40 + method, nmin=500, nmax=4000)
    dev = device_arrays(torch, data)
    out = []
    for prec in (pkg.FP64, pkg.MIXED):
        batch = pkg.Batch(method, systems, prec)
        out.append(run_batch(torch, batch, dev))
        batch.close()
    (f64, fk64, _, _), (f32, fk32, _, _) = out
    print("mixed vs fp64: total %.2e k-space %.2e" % (rel_err(f32, f64), rel_err(fk32, fk64)))
    assert rel_err(fk32, fk64) < 2e-5
    assert rel_err(f32, f64) < 2e-5


def test_batch_graph_replays(pkg):
    """Captured graph: replays with new positions, with eager work and
    synchronisation between them, each equal to an eager evaluation."""
    import torch

    systems, data = make_batch_systems(8, seed=77, nmin=300, nmax=2000)
    batch = pkg.Batch(1, systems)
    dev = device_arrays(torch, data)
    n = batch.ntot
    f = torch.zeros(3 * n, dtype=torch.float64, device="cuda")
    fk, fr = torch.zeros_like(f), torch.zeros_like(f)
    ptr = [dev[k].data_ptr() for k in "xyzq"]
    torch.cuda.synchronize()
    batch.capture(*ptr, f.data_ptr(), fk.data_ptr(), fr.data_ptr(), None)
    batch.synchronize()
    rng = np.random.default_rng(5)
    Ls = np.concatenate([[S["L"]] * S["n"] for S in systems])
    for rep in range(3):
        # new positions (eager torch work on the same arrays)
        for k in "xyz":
            dev[k].copy_(torch.tensor(rng.uniform(0, 1, n) * Ls, dtype=torch.float64, device="cuda"))
        torch.cuda.synchronize()
        batch.replay()
        batch.synchronize()
        got = f.cpu().numpy().copy()
        f2, fk2, fr2 = torch.zeros_like(f), torch.zeros_like(f), torch.zeros_like(f)
        batch.calculate_forces_device(*ptr, f2.data_ptr(), fk2.data_ptr(), fr2.data_ptr(), None)
        batch.synchronize()
        ref = f2.cpu().numpy()
        print("replay %d: %.2e" % (rep, rel_err(got, ref)))
        assert rel_err(got, ref) < 1e-13
    batch.close()


def test_batch_autotune_keeps_results(pkg):
    import torch

    systems, data = make_batch_systems(10, seed=91, nmin=500, nmax=3000)
    batch = pkg.Batch(3, systems, pkg.MIXED)
    dev = device_arrays(torch, data)
    before = run_batch(torch, batch, dev)[0]
    n = batch.ntot
    scratch = [torch.zeros(3 * n, dtype=torch.float64, device="cuda") for _ in range(3)]
    tile, grid = batch.autotune(*[dev[k].data_ptr() for k in "xyzq"], *[t.data_ptr() for t in scratch])
    print("autotuned tile side %d, real-space grid %d" % (tile, grid))
    assert tile in (4, 8) and grid > 0
    after = run_batch(torch, batch, dev)[0]
    assert rel_err(after, before) < 1e-6
    batch.close()


def test_batch_without_real_space_and_alpha_change(pkg, orc):
    import torch

    systems, data = make_batch_systems(4, seed=5)
    systems[1]["rcut"] = 0.0
    batch = pkg.Batch(7, systems)
    dev = device_arrays(torch, data)
    batch.set_alpha(2, 1.1)
    systems[2]["alpha"] = 1.1
    f, fk, fr, en = run_batch(torch, batch, dev)
    assert np.all(per_system(fr, batch.offsets, 1) == 0.0) and en[1] == 0.0
    for s in range(4):
        S, d = systems[s], data[s]
        fk_ref, _ = orc.kspace_force(7, d, S["M"], S["cao"], S["alpha"])
        assert rel_err(per_system(fk, batch.offsets, s), fk_ref) < 1e-12
    batch.close()


def test_batch_time_pairs(pkg):
    """The pair-kernel timer of a batch returns a positive time, leaves the
    forces as an ordinary evaluation computes them, and a graph captured
    afterwards (no events recorded in it) replays to the same forces (the
    mixed form to the rounding of its summation order, which varies from
    run to run)."""
    import torch

    systems, data = make_batch_systems(8, seed=93, nmin=500, nmax=2500)
    batch = pkg.Batch(3, systems, pkg.MIXED)
    dev = device_arrays(torch, data)
    before = run_batch(torch, batch, dev)[0]
    n = batch.ntot
    out = [torch.zeros(3 * n, dtype=torch.float64, device="cuda") for _ in range(3)]
    ms = batch.time_pairs(*[dev[k].data_ptr() for k in "xyzq"], *[t.data_ptr() for t in out], reps=3)
    print("batched pair kernel: %.4f ms" % ms)
    assert 0.0 < ms < 1e3
    assert rel_err(out[0].cpu().numpy().reshape(3, n), before) < 1e-6
    after = run_batch(torch, batch, dev, graph=True)[0]
    assert rel_err(after, before) < 1e-6
    batch.close()
