#!/usr/bin/env python3
"""Benchmark of the MI355X P3M force path.

One step = one full force evaluation of the workload (Calculate_forces:
real-space erfc part + k-space part: charge assignment, FFTs with the
influence-function multiply, back-interpolation), device-resident, on the
library's stream.  Prints one JSON line.

Default workload: 1,000,000 charges Gaussian about the centre of a box of
side 100 (sigma 12.5), +-1 neutral, P3M ad interlaced (method 3), mesh 128^3,
cao 7, rcut 4, alpha 0.75 -- the configuration the headline target is quoted
on.  `value` is the mixed-precision form; the FP64 form is timed beside it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector / matrix, dense
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector / matrix, dense
FP32_UNPACKED_TFLOPS = 78.6  # FP32 vector without packed (two-wide) instructions

WORKLOADS = {
    "gaussian-1M-ad-i-m128": dict(method=3, N=1_000_000, L=100.0, M=128, cao=7, rcut=4.0, alpha=0.75,
                                  gen="gaussian", seed=3),
    "uniform-100k-ik-i-m64": dict(method=1, N=100_000, L=46.4, M=64, cao=5, rcut=3.0, alpha=0.9,
                                  gen="uniform", seed=2),
    "uniform-1k-ad-r-m16": dict(method=7, N=1000, L=10.0, M=16, cao=7, rcut=3.0, alpha=1.0,
                                gen="alternating", seed=1),
    "uniform-32k-ik-m32": dict(method=0, N=32768, L=32.0, M=32, cao=5, rcut=3.0, alpha=1.0,
                               gen="normal", seed=5),
}
DEFAULT = "gaussian-1M-ad-i-m128"
# many independent systems in one batch (evaluated with one launch per stage)
BATCH_WORKLOAD = "batch-256-ad-i"
# forward k-space force plus its adjoint (vector-Jacobian product)
ADJOINT_WORKLOAD = "adjoint-32k-ik-m32"
SMOOTH_EVEN_16_48 = [16, 18, 20, 22, 24, 28, 30, 32, 36, 40, 42, 44, 48]


def make_system(w):
    rng = np.random.default_rng(w["seed"])
    N, L = w["N"], w["L"]
    if w["gen"] == "gaussian":
        def coord():
            out = np.empty(0)
            while out.size < N:
                v = rng.normal(L / 2, L / 8, 2 * N)
                out = np.concatenate([out, v[(v >= 0.0) & (v < L)]])
            return out[:N]
        x, y, z = coord(), coord(), coord()
    else:
        x, y, z = (rng.uniform(0.0, L, N) for _ in range(3))
    if w["gen"] == "alternating":
        q = np.where(np.arange(N) % 2 == 0, 1.0, -1.0)
    elif w["gen"] == "normal":
        q = rng.normal(0.0, 0.5, N)
        q -= q.mean()
    else:
        q = rng.permutation(np.repeat([1.0, -1.0], N // 2 + 1)[:N])
    return {"x": x, "y": y, "z": z, "q": q, "L": L}


def algorithmic_bytes(w, s):
    """Bytes per stage from the per-unit figures of the k-space/real-space
    pipeline: s = bytes per value (8 FP64, 4 FP32)."""
    N, M, method = w["N"], w["M"], w["method"]
    nm = 2 if method in (1, 3) else 1
    nout = 1 if method in (2, 3, 7) else 3
    r2c = method in (6, 7)
    C = M * M * (M // 2 + 1) if r2c else M ** 3
    return {
        "assign": N * 4 * s + nm * M ** 3 * s,
        "fft_multiply": 2 * C * 2 * s + C * 8 + nout * 2 * C * 2 * s,
        "interp": N * 4 * s + N * 3 * s + nout * nm * M ** 3 * s,
        "realspace": N * 4 * s + N * 3 * s,
    }


def run_form(pkg, torch, w, system, dev, precision, steps, warmup):
    params = pkg.parameters(w["M"], w["cao"], w["alpha"], w["rcut"])
    solver = pkg.Solver(w["method"], system, params, precision)
    N, L = w["N"], float(w["L"])
    f = torch.zeros(3 * N, dtype=torch.float64, device="cuda")
    fk = torch.zeros_like(f)
    fr = torch.zeros_like(f)
    en = torch.zeros(1, dtype=torch.float64, device="cuda")
    ptr = [t.data_ptr() for t in (dev["x"], dev["y"], dev["z"], dev["q"])]

    def step():
        solver.calculate_forces_device(N, L, *ptr, w["rcut"], w["alpha"], f.data_ptr(), fk.data_ptr(),
                                       fr.data_ptr(), en.data_ptr())

    torch.cuda.synchronize()
    for _ in range(warmup):
        step()
    solver.synchronize()
    # every timed step records its stage events (HIP events on the library's
    # stream, around each stage) into a set of its own; read after the loop
    solver.stage_ring_start(steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    solver.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ring = solver.stage_ring_read(steps)
    pairs_ms = solver.stage_ring_read_pairs(steps)
    solver.stage_ring_start(0)
    st = np.median(ring, axis=0)
    forces = f.cpu().numpy().reshape(3, N)
    kforces = fk.cpu().numpy().reshape(3, N)
    solver.close()
    return {
        "ms_per_step": 1e3 * elapsed / steps,
        "stage_ms": {"assign": st[0], "fft_multiply": st[1], "interp": st[2], "kspace": st[3],
                     "realspace": st[4]},
        # the real-space pair kernel alone (HIP events around its launch)
        "pair_kernel_ms": float(np.median(pairs_ms)) if len(pairs_ms) else 0.0,
        "stage_samples": len(ring),
        "forces": forces,
        "kforces": kforces,
    }


def ewald_gates(pkg, torch, w, dev, results, nsample=1000, precision=1e-6, rcut_e=12.0):
    """Accuracy gates against the FP64 Ewald sum on a seeded sample of
    particles (rms over the sample): (i) total force, (ii) k-space force vs
    the Ewald k-space force at the P3M alpha."""
    N, L = w["N"], float(w["L"])
    q = dev["q"].cpu().numpy()
    q2 = float(np.sum(q * q))
    rcut_e = min(rcut_e, 0.5 * L)
    alpha_e, kmax = pkg.ewald_parameters(N, L, q2, precision, rcut_e)
    idx = np.sort(np.random.default_rng(2024).choice(N, min(nsample, N), replace=False))
    ns = len(idx)
    di = torch.tensor(idx.astype(np.int32), device="cuda")
    ft = torch.zeros(3 * ns, dtype=torch.float64, device="cuda")
    fk = torch.zeros_like(ft)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pkg.ewald_sample_device(N, L, *[dev[k].data_ptr() for k in "xyzq"], ns, di.data_ptr(), alpha_e,
                            rcut_e, kmax, ft.data_ptr(), w["alpha"], fk.data_ptr())
    t_ewald = time.perf_counter() - t0
    ft = ft.cpu().numpy().reshape(3, ns)
    fk = fk.cpu().numpy().reshape(3, ns)

    def rms(a, b):
        return float(np.sqrt(np.sum((a - b) ** 2) / ns))

    out = {"sample": ns, "ewald_alpha": round(alpha_e, 6), "ewald_kmax": kmax, "ewald_rcut": rcut_e,
           "ewald_precision_target": precision, "ewald_seconds": round(t_ewald, 2)}
    for name, r in results.items():
        out[name] = {"rms_total_vs_ewald": rms(r["forces"][:, idx], ft),
                     "rms_kspace_vs_ewald": rms(r["kforces"][:, idx], fk)}
    return out


# Algorithmic flops of one real-space pair within rcut (Realteil,
# realpart.c:124-167): displacement and r^2 8, 1/r, r and alpha r 3,
# exp(-x^2) 11 and erfc(x) 10 (each function counted as a 10-flop polynomial),
# the bracket q_i q_j (erfc/r + c exp)/r^2 6, the force update 6.  The
# kernel evaluates each unordered pair once and adds its reaction to the
# partner (3 more flops), so its work is (ordered pairs / 2) x 47 flops.
PAIR_FLOPS = 44
HALF_PAIR_FLOPS = PAIR_FLOPS + 3


def pair_count(w, sysd, samples=20000):
    """Ordered pairs with minimum-image distance <= rcut: neighbour counts of
    a seeded sample of particles (periodic k-d tree), scaled to N; standard
    error about 0.5% at the default workload."""
    from scipy.spatial import cKDTree

    P = np.stack([sysd["x"], sysd["y"], sysd["z"]], 1)
    tree = cKDTree(P, boxsize=w["L"])
    k = min(samples, len(P))
    idx = np.random.default_rng(12345).choice(len(P), k, replace=False)
    cnt = tree.query_ball_point(P[idx], w["rcut"], return_length=True) - 1
    return float(len(P) * cnt.mean())


def stage_flops(w):
    """VALU flops of assignment and back-interpolation per evaluation."""
    N, p, method = w["N"], w["cao"], w["method"]
    nm = 2 if method in (1, 3) else 1
    ad = method in (2, 3, 7)
    return {
        "assign": N * nm * (3 * p * p + p + p * p + 2 * p ** 3),
        "interp": N * nm * (6 * p * p + (8 if ad else 4) * p ** 3),
    }


def roofline(w, res, s, pairs):
    byts = algorithmic_bytes(w, s)
    flops = stage_flops(w)
    vpeak = FP32_PEAK_TFLOPS if s == 4 else FP64_PEAK_TFLOPS
    stages = {k: res["stage_ms"][k] for k in ("assign", "fft_multiply", "interp", "realspace")}
    dom = max(stages, key=stages.get)
    per = {}
    for k, ms in stages.items():
        gbs = byts[k] / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        per[k] = {"ms": round(ms, 4), "GB/s": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4),
                  "bytes": int(byts[k])}
        if k in flops and ms > 0:
            tf = flops[k] / (ms * 1e-3) / 1e12
            per[k].update({"TFLOP/s": round(tf, 3), "frac_valu": round(tf / vpeak, 4),
                           "flops": int(flops[k])})
    # the pair kernel's own time (HIP events around its launch); the stage
    # time also holds the cell-list build
    rs_ms = res.get("pair_kernel_ms") or stages["realspace"]
    if rs_ms > 0:
        fl = 0.5 * pairs * HALF_PAIR_FLOPS
        tf = fl / (rs_ms * 1e-3) / 1e12
        per["realspace"].update({"pairs": int(pairs), "pair_kernel_ms": round(rs_ms, 4),
                                 "TFLOP/s": round(tf, 3), "frac_valu": round(tf / vpeak, 4),
                                 "flops": int(fl)})
    return dom, per


# Kernel symbol of each stage's dominant launch in the mixed form, as the
# profiler names it.
STAGE_KERNEL = {"realspace": "k_realspace_half<true>", "interp": "k_interp_ad_tile<7, float, false, true>",
                "assign": "k_assign_rows<7, float, false, true, 0>", "fft_multiply": "k_fft_p2<128, float, false>"}


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary of
    this workload (rocprofv3 FETCH_SIZE and WRITE_SIZE runs, scaled by the
    same counters over a plain device copy; the table's last column).  The
    counters cannot be read from inside the process, so the figure is that of
    the profiled build; None when no summary covers the kernel."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.md" % workload)
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 4 and cells[0] == "`%s`" % kernel:
            return float(cells[3])
    return None


def roofline_object(dom, per, workload):
    """The dominant stage of the mixed form against its bound.  Real space
    runs its pair arithmetic on the vector ALUs in FP32 (no MFMA: nothing on
    the path is a dense contraction), so its bound is compute, priced at the
    FP32 dense peak; the mesh stages are priced against HBM."""
    st = per[dom]
    kernel = STAGE_KERNEL.get(dom, dom)
    traffic = measured_traffic(workload, kernel)
    if dom == "realspace":
        return {"bound": "valu", "kernel": kernel + " (cell lists + half-shell pair kernel, FP32 pairs)",
                "achieved": st["TFLOP/s"], "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": st["frac_valu"], "traffic": traffic,
                "peak_unpacked_fp32": FP32_UNPACKED_TFLOPS,
                "frac_unpacked_fp32": round(st["TFLOP/s"] / FP32_UNPACKED_TFLOPS, 4),
                "note": "vector-ALU bound (pair arithmetic in FP32, FP64 LDS accumulation; the only "
                        "matrix-core use is the FP16 candidate distance test); achieved = %d ordered "
                        "pairs within rcut / 2 (each unordered pair once) x %d flops / median time "
                        "of the pair kernel (HIP events around its launch on the library's stream); peak = the published FP32 figure, which counts packed FP32 (the build "
                        "has packed FP32 off: 78.6 TFLOP/s unpacked, frac_unpacked_fp32); traffic = "
                        "HBM bytes per launch from profiles/traffic_%s.md"
                        % (st["pairs"], HALF_PAIR_FLOPS, workload)}
    return {"bound": "hbm", "kernel": kernel, "achieved": st["GB/s"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": st["frac_hbm"], "traffic": traffic}


def host_cpu():
    """(logical CPUs usable by this process, CPU model name)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    return ncpu, model


def cpu_baseline(w, sysd):
    """The oracle (plain-C FP64 restatement of the reference: the reference
    itself needs FFTW and GSL, absent here) on this host: one full k-space
    evaluation plus the real-space force of a 2000-particle sample
    (cell-list restatement of Realteil; the reference's O(N^2) loop is
    infeasible at this size), scaled to N.  One thread, as the reference's
    default build runs; beside it the same with the real-space rows split
    over all usable host CPUs (threads; ctypes releases the interpreter
    lock).  The reference's assignment, FFT and back-interpolation carry no
    OpenMP pragmas, so k space stays on one thread in both."""
    from concurrent.futures import ThreadPoolExecutor

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    N, L, M, cao, alpha = w["N"], w["L"], w["M"], w["cao"], w["alpha"]
    ncpu, model = host_cpu()
    G = orc.influence(w["method"], M, cao, L, alpha)
    b = orc.self_force_coeffs(M, cao, L, G) if w["method"] in (2, 7) else None
    t0 = time.perf_counter()
    orc.kspace_force(w["method"], sysd, M, cao, alpha, G=G, b=b)
    tk = time.perf_counter() - t0
    sample = min(N, 2000)
    t0 = time.perf_counter()
    orc.realteil(sysd, w["rcut"], alpha, cells=True, i_range=(0, sample))
    tr = (time.perf_counter() - t0) * N / sample
    t_eval = tk + tr
    # all cores: ncpu times the sample, one slice per thread
    big = min(N, sample * ncpu)
    bounds = np.linspace(0, big, ncpu + 1).astype(int)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(ncpu) as ex:
        list(ex.map(lambda k: orc.realteil(sysd, w["rcut"], alpha, cells=True,
                                           i_range=(bounds[k], bounds[k + 1])), range(ncpu)))
    tr_all = (time.perf_counter() - t0) * N / big
    return {"value": N / t_eval, "unit": "particle-force evals/s", "cores": 1, "kind": "port",
            "host_cpus": ncpu, "cpu_model": model,
            "sample": "oracle (FP64 C, 1 thread): full k-space evaluation (%.1f s) + real space of "
                      "%d of %d particles on a cell list, scaled (%.1f s)" % (tk, sample, N, tr),
            "all_cores": {"value": N / (tk + tr_all), "cores": ncpu,
                          "sample": "k space as above (1 thread: no OpenMP in the reference's "
                                    "assignment / FFT / interpolation) + real space of %d particles "
                                    "over %d threads, scaled (%.1f s)" % (big, ncpu, tr_all)}}


def batch_systems(nsys=256, seed=4, rcut=3.0, eps=1e-4, density=0.5):
    """256 systems: N even, uniform in [500, 20000]; L = (N / 0.5)^(1/3); mesh
    from the even 11-smooth sizes in [16, 48]; cao 3..7; alpha from the
    tuner's real-space formula at 1e-4; uniform positions, +-1 neutral."""
    rng = np.random.default_rng(seed)
    systems, data = [], []
    for _ in range(nsys):
        n = 2 * int(rng.integers(250, 10001))
        L = (n / density) ** (1.0 / 3.0)
        M = int(rng.choice(SMOOTH_EVEN_16_48))
        cao = int(rng.integers(3, 8))
        x, y, z = (rng.uniform(0.0, L, n) for _ in range(3))
        q = rng.permutation(np.repeat([1.0, -1.0], n // 2))
        q2 = float(n)
        alpha = float(np.sqrt(-np.log(eps * np.sqrt(n * rcut * L ** 3) / (2 * np.sqrt(2) * q2))) / rcut)
        systems.append(dict(n=n, L=L, M=M, cao=cao, alpha=alpha, rcut=rcut))
        data.append(dict(x=x, y=y, z=z, q=q))
    return systems, data


def batch_pair_count(systems, data):
    """Ordered pairs within rcut (minimum image) summed over the systems:
    exact, by a periodic k-d tree per system."""
    from scipy.spatial import cKDTree

    total = 0
    for s, d in zip(systems, data):
        P = np.stack([d["x"], d["y"], d["z"]], 1) % s["L"]
        tree = cKDTree(P, boxsize=s["L"])
        total += int(tree.count_neighbors(tree, s["rcut"])) - len(P)
    return total


def bench_batch(pkg, torch, args):
    """One step = Calculate_forces of all 256 systems (real + k space), as
    one replay of the captured graph of batched launches."""
    method = 3
    systems, data = batch_systems()
    ntot = sum(s["n"] for s in systems)
    dev = {k: torch.tensor(np.concatenate([d[k] for d in data]), dtype=torch.float64, device="cuda")
           for k in "xyzq"}
    ptr = [dev[k].data_ptr() for k in "xyzq"]
    res = {}
    for name, prec in (("fp64", pkg.FP64), ("mixed", pkg.MIXED)):
        batch = pkg.Batch(method, systems, prec)
        f = torch.zeros(3 * ntot, dtype=torch.float64, device="cuda")
        fk, fr = torch.zeros_like(f), torch.zeros_like(f)
        torch.cuda.synchronize()
        tile, grid = batch.autotune(*ptr, f.data_ptr(), fk.data_ptr(), fr.data_ptr())
        batch.capture(*ptr, f.data_ptr(), fk.data_ptr(), fr.data_ptr(), None)
        for _ in range(args.warmup):
            batch.replay()
        batch.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            batch.replay()
        batch.synchronize()
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / args.steps
        forces = f.cpu().numpy()
        # the batched half-shell kernel alone (HIP events around its launch,
        # eager evaluations after the timed replays)
        pair_ms = batch.time_pairs(*ptr, f.data_ptr(), fk.data_ptr(), fr.data_ptr(), reps=5)
        res[name] = {"ms_per_step": round(ms, 4), "value": round(ntot / (ms * 1e-3), 1),
                     "tile_side": tile, "rs_grid": grid, "pair_kernel_ms": round(pair_ms, 4),
                     "forces": forces}
        batch.close()
    rel = float(np.linalg.norm(res["mixed"]["forces"] - res["fp64"]["forces"]) /
                np.linalg.norm(res["fp64"]["forces"]))
    for r in res.values():
        r.pop("forces")
    pairs = batch_pair_count(systems, data)
    fl = 0.5 * pairs * HALF_PAIR_FLOPS
    tf = fl / (res["mixed"]["pair_kernel_ms"] * 1e-3) / 1e12
    roof = {"bound": "valu", "kernel": "k_brealspace_half<true> (all systems' half-shell pairs, FP32)",
            "achieved": round(tf, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tf / FP32_PEAK_TFLOPS, 4),
            "traffic": measured_traffic(BATCH_WORKLOAD, "k_brealspace_half<true>"),
            "peak_unpacked_fp32": FP32_UNPACKED_TFLOPS,
            "frac_unpacked_fp32": round(tf / FP32_UNPACKED_TFLOPS, 4),
            "note": "the largest kernel of the batch step; achieved = %d ordered pairs within rcut "
                    "over all systems (exact, periodic k-d trees) / 2 x %d flops / mean time of the "
                    "kernel (HIP events around its launch, %s ms, eager calls); traffic = HBM bytes "
                    "per launch from profiles/traffic_%s.md"
                    % (pairs, HALF_PAIR_FLOPS, res["mixed"]["pair_kernel_ms"], BATCH_WORKLOAD)}
    return {
        "metric": "particle-force evals/sec at rms err <=1e-4 (mixed vs FP64 GPU), HBM GB/s",
        "value": res["mixed"]["value"], "unit": "particle-force evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["mixed"]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 mesh/FFT/interp/pairs, f64 influence+accumulation",
        "data": "synthetic (seeded numpy generator)",
        "config": {"workload": BATCH_WORKLOAD, "method": method, "systems": len(systems),
                   "N_total": ntot, "N_range": [500, 20000], "mesh": "even 11-smooth 16..48",
                   "cao": "3..7", "rcut": 3.0, "alpha": "tuning.c:293 at 1e-4",
                   "parallelism": "single GPU, one launch per stage over all systems, HIP graph"},
        "roofline": roof,
        "fp64": res["fp64"], "mixed": res["mixed"],
        "accuracy": {"rel_mixed_vs_fp64": rel},
    }


def bench_adjoint(pkg, torch, args):
    """One step = total force of 32768 charges (ik, mesh 32, cao 5, real
    space to rcut 3) plus the adjoint of loss = sum |F - F_target|^2 w.r.t.
    positions and charges (k-space and real-space adjoints); the k-space-only
    step is timed beside it."""
    rng = np.random.default_rng(5)
    N, L, M, cao, alpha, rcut = 32768, 32.0, 32, 5, 1.0, 3.0
    x, y, z = (rng.uniform(0.0, L, N) for _ in range(3))
    q = rng.normal(0.0, 0.5, N)
    q -= q.mean()
    sysd = {"x": x, "y": y, "z": z, "q": q, "L": L}
    dev = {k: torch.tensor(sysd[k], dtype=torch.float64, device="cuda") for k in "xyzq"}
    ptr = [dev[k].data_ptr() for k in "xyzq"]
    res = {}
    for name, prec in (("fp64", pkg.FP64), ("mixed", pkg.MIXED)):
        solver = pkg.Solver(0, pkg.HostSystem.from_dict(sysd), pkg.parameters(M, cao, alpha), prec)
        fk = torch.zeros(3 * N, dtype=torch.float64, device="cuda")
        f = torch.zeros_like(fk)
        fr = torch.zeros_like(fk)
        g = torch.tensor(rng.normal(size=3 * N), dtype=torch.float64, device="cuda")
        dpos = torch.zeros(3 * N, dtype=torch.float64, device="cuda")
        dq = torch.zeros(N, dtype=torch.float64, device="cuda")

        def step_total():
            solver.calculate_forces_device(N, L, *ptr, rcut, alpha, f.data_ptr(), fk.data_ptr(),
                                           fr.data_ptr(), None)
            solver.kspace_force_vjp_device(N, L, *ptr, g.data_ptr(), dpos.data_ptr(), dq.data_ptr())
            solver.realteil_vjp_device(N, L, *ptr, rcut, alpha, g.data_ptr(), dpos.data_ptr(),
                                       dq.data_ptr())

        def step_kspace():
            solver.kspace_force_device(N, L, *ptr, fk.data_ptr(), f.data_ptr())
            solver.kspace_force_vjp_device(N, L, *ptr, g.data_ptr(), dpos.data_ptr(), dq.data_ptr())

        res[name] = {}
        for label, step in (("total", step_total), ("kspace_only", step_kspace)):
            torch.cuda.synchronize()
            for _ in range(args.warmup):
                step()
            solver.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                step()
            solver.synchronize()
            ms = 1e3 * (time.perf_counter() - t0) / args.steps
            res[name][label] = {"ms_per_step": round(ms, 4), "value": round(N / (ms * 1e-3), 1)}
        solver.close()
    return {
        "metric": "particle-force evals/sec (forward + adjoint)", "value": res["mixed"]["total"]["value"],
        "unit": "particle-force evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": res["mixed"]["total"]["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 mesh/FFT/interp/pairs, f64 influence+accumulation; f64 real-space adjoint",
        "data": "synthetic (seeded numpy generator)",
        "config": {"workload": ADJOINT_WORKLOAD, "method": 0, "N": N, "box": L, "mesh": M, "cao": cao,
                   "alpha": alpha, "rcut": rcut, "parallelism": "single GPU"},
        "roofline": None, "fp64": res["fp64"], "mixed": res["mixed"],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default=DEFAULT,
                    choices=sorted(WORKLOADS) + [BATCH_WORKLOAD, ADJOINT_WORKLOAD])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; the multi-GPU path is design-only "
              "(--gpus %d not supported)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    import __graft_entry__ as entry

    pkg = entry.load_package()
    if args.workload == BATCH_WORKLOAD:
        print(json.dumps(bench_batch(pkg, torch, args)))
        return
    if args.workload == ADJOINT_WORKLOAD:
        print(json.dumps(bench_adjoint(pkg, torch, args)))
        return
    w = WORKLOADS[args.workload]
    sysd = make_system(w)
    system = pkg.HostSystem.from_dict(sysd)
    dev = {k: torch.tensor(sysd[k], dtype=torch.float64, device="cuda") for k in "xyzq"}
    N = w["N"]

    r64 = run_form(pkg, torch, w, system, dev, pkg.FP64, args.steps, args.warmup)
    rmx = run_form(pkg, torch, w, system, dev, pkg.MIXED, args.steps, args.warmup)

    f64, fmx = r64["forces"], rmx["forces"]
    rms_mixed_vs_fp64 = float(np.sqrt(np.sum((fmx - f64) ** 2) / N))
    rms_force = float(np.sqrt(np.sum(f64 ** 2) / N))
    params = pkg.parameters(w["M"], w["cao"], w["alpha"], w["rcut"])
    est_total, est_k = pkg.error_estimate(w["method"], system, params)
    gates = ewald_gates(pkg, torch, w, dev, {"fp64": r64, "mixed": rmx})
    # the inhomogeneous k-space estimate (p3m-common.c:437-719) at the
    # method's mesh and order with one shell of aliasing copies, beside the
    # same estimate for a uniform density: their ratio is what clustering
    # does to the k-space error, to compare with the measured ratio of the
    # k-space rms against Ewald to the homogeneous a-priori estimate
    t0 = time.perf_counter()
    est_inhomo = pkg.error_inhomo(system, w["alpha"], mesh=w["M"], cao=w["cao"], mc=1)
    est_uniform = pkg.error_inhomo(system, w["alpha"], mesh=w["M"], cao=w["cao"], mc=1, uniform=True)
    t_inhomo = time.perf_counter() - t0

    pairs = pair_count(w, sysd) if w["rcut"] > 0 else 0.0
    dom, per = roofline(w, rmx, 4, pairs)
    _, per64 = roofline(w, r64, 8, pairs)
    value = N / (rmx["ms_per_step"] * 1e-3)
    value64 = N / (r64["ms_per_step"] * 1e-3)
    out = {
        "metric": "particle-force evals/sec at rms err <=1e-4 (mixed vs FP64 GPU), HBM GB/s",
        "value": round(value, 1),
        "unit": "particle-force evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(rmx["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 mesh/FFT/interp/pairs, f64 influence+accumulation",
        "data": "synthetic (seeded numpy generator)",
        "config": {"workload": args.workload, "method": w["method"], "N": N, "box": w["L"],
                   "mesh": w["M"], "cao": w["cao"], "rcut": w["rcut"], "alpha": w["alpha"],
                   "parallelism": "single GPU"},
        "roofline": roofline_object(dom, per, args.workload),
        "stages_mixed": per,
        "fp64": {"value": round(value64, 1), "ms_per_step": round(r64["ms_per_step"], 4),
                 "stages": per64},
        "speedup_mixed_vs_fp64": round(r64["ms_per_step"] / rmx["ms_per_step"], 3),
        "speedup_mixed_vs_fp64_kspace": round(r64["stage_ms"]["kspace"] / rmx["stage_ms"]["kspace"], 3),
        "speedup_mixed_vs_fp64_realspace": round(r64["stage_ms"]["realspace"] /
                                                 rmx["stage_ms"]["realspace"], 3),
        "accuracy": {"rms_mixed_vs_fp64": rms_mixed_vs_fp64, "rms_force": rms_force,
                     "estimate_total": est_total, "estimate_kspace": est_k, "ewald": gates,
                     "estimate_kspace_inhomo": {
                         "mesh": w["M"], "cao": w["cao"], "mc": 1, "inhomogeneous": est_inhomo,
                         "uniform_density": est_uniform,
                         "ratio": (est_inhomo / est_uniform) if est_uniform > 0 else None,
                         "measured_ratio": gates["fp64"]["rms_kspace_vs_ewald"] / est_k if est_k > 0 else None,
                         "seconds": round(t_inhomo, 2)}},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, sysd)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
