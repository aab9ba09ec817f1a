"""The C ABI: library loads, exports every symbol the header declares, struct
layouts match the header, and argument checks that fail before any GPU work
return errors (no GPU needed)."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT, random_system

HEADER = os.path.join(ROOT, "include", "p3m_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(p3m_hip_\w+)\s*\(", text)))


def declared_objects():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"extern const method_t (\w+);", text)))


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg.lib()
    funcs, objs = declared_functions(), declared_objects()
    assert len(funcs) >= 15 and len(objs) == 6
    for name in funcs + objs:
        assert hasattr(lib, name), name


def test_method_table_matches_reference(pkg):
    # ids, short names and flags of p3m-ik.c:46-49 ... p3m-ad-real.c:39-41
    P3M = 8 | 16 | 32 | 64
    expect = {
        "method_p3m_ik": (0, b"p3m-ik", P3M | 1),
        "method_p3m_ik_i": (1, b"p3m-ik-i", P3M | 1 | 4),
        "method_p3m_ad": (2, b"p3m-ad", P3M | 2 | 128),
        "method_p3m_ad_i": (3, b"p3m-ad-i", P3M | 2 | 4),
        "method_p3m_ik_r": (6, b"p3m-ik-r", P3M | 1),
        "method_p3m_ad_r": (7, b"p3m-ad-r", P3M | 2 | 128),
    }
    for sym, (mid, short, flags) in expect.items():
        m = pkg.exported_method(sym)
        assert m.method_id == mid
        assert m.method_name_short == short
        assert ord(m.flags) == flags & 0xFF
        assert m.Init and m.Influence_function and m.Kspace_force and m.Error and m.Error_k
        assert pkg.method(mid).method_id == mid


PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "p3m_hip.h"
#define P(T, F) printf(#T " " #F " %zu\n", offsetof(T, F))
#define S(T) printf(#T " sizeof %zu\n", sizeof(T))
int main(void) {
  S(vector_array_t); P(vector_array_t, x); P(vector_array_t, fields);
  S(forces_t); P(forces_t, f_r);
  S(system_t); P(system_t, nparticles); P(system_t, p); P(system_t, q); P(system_t, q2);
  P(system_t, reference); P(system_t, epsilon); P(system_t, energy);
  S(parameters_t); P(parameters_t, mesh); P(parameters_t, cao3); P(parameters_t, precision);
  P(parameters_t, tuning);
  S(data_t); P(data_t, G_hat); P(data_t, dQ); P(data_t, ca_ind); P(data_t, cf); P(data_t, inter);
  P(data_t, forward_plans); P(data_t, forward_plan); P(data_t, neighbor_list);
  P(data_t, self_force_corrections); P(data_t, method_data); P(data_t, runtime);
  S(method_t); P(method_t, flags); P(method_t, Init); P(method_t, Error_k);
  S(timing_t); P(timing_t, min); P(timing_t, n);
  S(runtime_stat_t); P(runtime_stat_t, t_c); P(runtime_stat_t, t_f);
  return 0;
}
"""


def test_struct_layout_matches_python_mirror(pkg, tmp_path):
    src = tmp_path / "probe.c"
    src.write_text(PROBE)
    exe = tmp_path / "probe"
    subprocess.check_call(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    out = subprocess.check_output([str(exe)]).decode().split("\n")
    table = {"vector_array_t": pkg.VectorArray, "forces_t": pkg.Forces, "system_t": pkg.System,
             "parameters_t": pkg.Parameters, "data_t": pkg.Data, "method_t": pkg.Method,
             "timing_t": pkg.Timing, "runtime_stat_t": pkg.RuntimeStat}
    for line in out:
        if not line:
            continue
        t, f, v = line.split()
        cls = table[t]
        if f == "sizeof":
            assert ctypes.sizeof(cls) == int(v), t
        else:
            assert getattr(cls, f).offset == int(v), (t, f)


def _init(pkg, method, mesh, cao):
    s = random_system(10, 5.0, seed=1)
    system = pkg.HostSystem.from_dict(s)
    p = pkg.parameters(mesh, cao, 1.0)
    d = ctypes.POINTER(pkg.Data)()
    return pkg.lib().p3m_hip_init(method, system.ptr, ctypes.pointer(p), pkg.FP64, ctypes.byref(d))


@pytest.mark.parametrize("method,mesh,cao", [(5, 16, 5), (4, 16, 5), (0, 16, 8), (1, 16, 0),
                                              (2, 15, 3), (3, 4, 7), (7, 0, 1)])
def test_init_rejects_bad_arguments(pkg, method, mesh, cao):
    """Invalid method / cao / mesh are rejected before any device work."""
    rc = _init(pkg, method, mesh, cao)
    assert rc == pkg.EINVAL
    assert pkg.lib().p3m_hip_last_error().decode() != ""


def test_error_rejects_bad_arguments(pkg):
    s = random_system(10, 5.0, seed=1)
    system = pkg.HostSystem.from_dict(s)
    p = pkg.parameters(16, 9, 1.0, 2.0)
    t, k = ctypes.c_double(), ctypes.c_double()
    assert pkg.lib().p3m_hip_error(0, system.ptr, ctypes.pointer(p), ctypes.byref(t), ctypes.byref(k)) == pkg.EINVAL


def test_error_table_argument_checks(pkg):
    import numpy as np

    t = np.zeros(101)
    tp = t.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert pkg.lib().p3m_hip_set_error_table(2, 16, 5, tp) == pkg.EINVAL
    assert pkg.lib().p3m_hip_set_error_table(0, 16, 8, tp) == pkg.EINVAL
    assert pkg.lib().p3m_hip_set_error_table(0, 16, 5, None) == pkg.EINVAL


def _table_lookup(tab, alpha, length):
    """find_error.c:38-58: linear interpolation on alpha L = 0, 1, ..., 100."""
    import math

    aL = alpha * length
    l = int(math.floor(aL))
    d = aL - l
    return tab[100] if l >= 100 else (1.0 - d) * tab[l] + d * tab[l + 1]


@pytest.mark.parametrize("method", [0, 1, 6])
def test_builtin_error_tables(pkg, method):
    """Error_k of ids 0, 1 and 6 at tabulated meshes reads the reference's
    Q_ik / Q_ik_i tables without any registration (find_error.c:84-107),
    through p3m_hip_error and through the method table's Error_k; both equal
    the estimate formed from the rows of q_ik.c / q_ik_i.c in
    tests/golden/q_tables.npz.  A table hit is host logic: no GPU needed."""
    import numpy as np

    d = np.load(os.path.join(ROOT, "tests", "golden", "q_tables.npz"))
    tabs = d["Q_ik_i" if method == 1 else "Q_ik"]
    s = random_system(1000, 10.0, seed=2)
    system = pkg.HostSystem.from_dict(s)
    q2, L = float(np.sum(s["q"] ** 2)), 10.0
    m = pkg.method(method)
    t, k = ctypes.c_double(), ctypes.c_double()
    checked = 0
    for mi, M in enumerate(d["meshes"]):
        for cao in range(1, 8):
            for aL in (0.0, 0.37, 1.0, 7.5, 20.0, 33.3, 41.76, 99.5, 100.0):
                alpha = aL / L
                p = pkg.parameters(int(M), cao, alpha, 3.0)
                he = _table_lookup(tabs[mi, cao - 1], alpha, L)
                if he < 0.0:
                    # the tables hold a few rounding-noise negatives at small
                    # alpha L; there the reference falls back to the direct
                    # sum (p3m-ik.c:270), a GPU computation
                    continue
                # p3m-ik.c:300, p3m-ik-i.c:319
                want = 2.0 * q2 * np.sqrt(he / 1000) / (L * L)
                rc = pkg.lib().p3m_hip_error(method, system.ptr, ctypes.pointer(p), ctypes.byref(t),
                                            ctypes.byref(k))
                assert rc == pkg.OK
                assert abs(k.value - want) <= 1e-14 * abs(want), (M, cao, aL, k.value, want)
                assert m.Error_k(system.ptr, ctypes.pointer(p)) == k.value
                checked += 1
    assert checked > 200
