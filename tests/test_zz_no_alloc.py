"""GPU test of the allocation-free Ewald sampler
(p3m_hip_ewald_sample_device_noalloc) through the C ABI: two calls in a
row, the second smaller, each against the allocating entry point
(p3m_hip_ewald_sample_device) on the same inputs; requests beyond the
workspace's limits are rejected before any GPU work."""
import ctypes

import numpy as np
import pytest

from conftest import random_system, rel_err

pytestmark = pytest.mark.gpu


def _noalloc(pkg):
    f = pkg.lib().p3m_hip_ewald_sample_device_noalloc
    i, d, vp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    f.restype = i
    f.argtypes = [i, d, vp, vp, vp, vp, i, vp, d, d, i, d, vp, vp, vp]
    return f


def _run_both(pkg, torch, s, n, nsample, kmax, alpha, rcut, alpha_split, stream, seed):
    dev = {k: torch.tensor(np.asarray(s[k][:n]), dtype=torch.float64, device="cuda") for k in "xyzq"}
    idx = np.sort(np.random.default_rng(seed).choice(n, nsample, replace=False)).astype(np.int32)
    di = torch.tensor(idx, device="cuda")
    out = {}
    for name in ("alloc", "noalloc"):
        ft = torch.zeros(3 * nsample, dtype=torch.float64, device="cuda")
        fk = torch.zeros_like(ft)
        args = [n, float(s["L"])] + [dev[k].data_ptr() for k in "xyzq"] + [
            nsample, di.data_ptr(), alpha, rcut, kmax, alpha_split, ft.data_ptr(), fk.data_ptr()]
        torch.cuda.synchronize()
        if name == "alloc":
            pkg.check(pkg.lib().p3m_hip_ewald_sample_device(*args))
        else:
            pkg.check(_noalloc(pkg)(*args, stream.cuda_stream if stream is not None else None))
            if stream is not None:
                stream.synchronize()
            torch.cuda.synchronize()
        out[name] = (ft.cpu().numpy(), fk.cpu().numpy())
    return out


def test_noalloc_matches_allocating_entry_point(pkg):
    import torch

    s = random_system(3000, 14.0, seed=31)
    # first call: the larger input, on a stream of the caller's
    st = torch.cuda.Stream()
    a = _run_both(pkg, torch, s, 3000, 500, 24, 0.8, 7.0, 1.1, st, 1)
    # second call: fewer particles, samples and lattice vectors, null stream;
    # the workspace still holds the first call's contents
    b = _run_both(pkg, torch, s, 1500, 200, 12, 0.9, 5.0, 1.3, None, 2)
    for name, r in (("first", a), ("second", b)):
        et = rel_err(r["noalloc"][0], r["alloc"][0])
        ek = rel_err(r["noalloc"][1], r["alloc"][1])
        print("%s call: total %.2e, k-space split %.2e" % (name, et, ek))
        assert et < 1e-12
        assert ek < 1e-12
        assert np.linalg.norm(r["alloc"][0]) > 0.0


def test_noalloc_rejects_sizes_beyond_its_workspace(pkg):
    import torch

    s = random_system(100, 10.0, seed=3)
    dev = {k: torch.tensor(np.asarray(s[k]), dtype=torch.float64, device="cuda") for k in "xyzq"}
    di = torch.arange(100, dtype=torch.int32, device="cuda")
    ft = torch.zeros(3 * 40000, dtype=torch.float64, device="cuda")
    base = [100, 10.0] + [dev[k].data_ptr() for k in "xyzq"]
    f = _noalloc(pkg)
    # kmax beyond 63
    assert f(*base, 100, di.data_ptr(), 1.0, 4.0, 64, 0.0, ft.data_ptr(), None, None) == pkg.EINVAL
    assert "kmax" in pkg.lib().p3m_hip_last_error().decode()
    # more samples than 32768 (checked before the index array is read)
    assert f(*base, 40000, di.data_ptr(), 1.0, 4.0, 8, 0.0, ft.data_ptr(), None, None) == pkg.EINVAL
    assert "sampled particles" in pkg.lib().p3m_hip_last_error().decode()
    torch.cuda.synchronize()
    assert np.all(ft.cpu().numpy() == 0.0)
