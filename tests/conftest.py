import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import __graft_entry__ as entry  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def orc():
    import oracle

    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        oracle.build()
    return oracle


@pytest.fixture(scope="session")
def pkg():
    return entry.load_package()


def random_system(n, L, seed, charges="alternating", layout="uniform"):
    rng = np.random.default_rng(seed)
    if layout == "gaussian":
        def coord():
            out = np.empty(0)
            while out.size < n:
                v = rng.normal(L / 2, L / 8, 2 * n)
                out = np.concatenate([out, v[(v >= 0.0) & (v < L)]])
            return out[:n]
        x, y, z = coord(), coord(), coord()
    else:
        x, y, z = (rng.uniform(0.0, L, n) for _ in range(3))
    if charges == "alternating":
        q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    else:
        q = rng.normal(0.0, 0.5, n)
        q -= q.mean()
    return {"x": x, "y": y, "z": z, "q": q, "L": float(L)}


def load_fixture(name):
    d = np.load(os.path.join(GOLDEN, name))
    s = {k: np.array(d[k]) for k in ("x", "y", "z", "q")}
    s["L"] = float(d["L"])
    return s, np.array(d["f_ewald"])


def fold(s):
    """Positions wrapped into [0, L) (forces are periodic)."""
    out = dict(s)
    for k in "xyz":
        out[k] = np.mod(s[k], s["L"])
        out[k][out[k] >= s["L"]] = 0.0
    return out


def cell_geometry(n, L, rcut):
    """The library's real-space cell grid: (cells per side, reach).  Cells
    of side >= rcut / 2 (reach 2) unless those would hold fewer than 6
    particles on average, or the box has fewer than 5 of them per side; then
    side >= rcut (reach 1)."""
    nc2 = min(512, int(np.floor(L * 2 / rcut * (1 - 1e-12))))
    nc1 = min(512, int(np.floor(L / rcut * (1 - 1e-12))))
    if nc2 < 5 or (nc1 >= 3 and n < 6.0 * nc2 ** 3):
        return max(nc1, 1), 1
    return nc2, 2


def mixed_realspace_bound(s, rcut):
    """Bound on the mixed form's relative real-space force error: its pair
    differences are exact differences of coordinates rounded to the quantum
    s = side * 2^-29 (cells of side rcut / 2) or side * 2^-30 (side rcut),
    so a pair r apart carries at most 2u + sqrt(3) s / r relative error in d
    (u = 2^-24), about three times that in its force; the closest pair
    bounds the whole.  Never below 1e-6."""
    from scipy.spatial import cKDTree

    L = s["L"]
    p = np.stack([s["x"], s["y"], s["z"]], 1)
    d, _ = cKDTree(p, boxsize=L * (1 + 1e-12)).query(p, k=2)
    rmin = max(float(d[:, 1].min()), 1e-300)
    nc, reach = cell_geometry(len(s["q"]), L, rcut)
    quantum = L / nc * 2.0 ** -(30 if reach == 1 else 29)
    return max(1e-6, 3.0 * (2.0 * 2.0 ** -24 + 3.0 ** 0.5 * quantum / rmin))


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def error_table_row(method, M, cao):
    """The reference's Q_ik / Q_ik_i row (q_ik.c, q_ik_i.c) that Error_k of
    methods 0, 6 / 1 reads at mesh M, or None (ad methods, or a mesh outside
    the committed fixture)."""
    if method not in (0, 1, 6):
        return None
    d = np.load(os.path.join(GOLDEN, "q_tables.npz"))
    hit = np.flatnonzero(d["meshes"] == M)
    if hit.size == 0:
        return None
    return np.array(d["Q_ik_i" if method == 1 else "Q_ik"][hit[0], cao - 1])


def ref_kspace_error(orc, method, N, L, q2, M, cao, alpha):
    """The oracle's Error_k as the reference computes it: through the
    tabulated row where the reference has one (find_error.c:84-107)."""
    return orc.kspace_error(method, N, L, q2, M, cao, alpha, table=error_table_row(method, M, cao))
