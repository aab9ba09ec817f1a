#!/usr/bin/env python3
"""Write the test fixtures under tests/golden/ from data files the reference holds.

Run in the build container only (it reads /root/reference); the GPU box and
the test-suite use the .npz files this script writes, never the reference.

  tests_wall.npz   tests-wall.pos (300 charges, L = 10) and the Ewald forces in
                   columns 2-4 of tests-wall.for
  nacl_*.npz       Data/ReferenceDataVincent NaCl configurations (512 ions,
                   L = 22.728) and the EwaldSum forces (columns 1-3 of .exact)
  q_tables.npz     rows of the error-kernel tables Q_ik / Q_ik_i / Q_ad /
                   Q_ad_i (q_*.c) for meshes 16, 32, 64, 128, cao 1..7,
                   101 points at alpha*L = 0..100
"""
import os
import re
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))
MESH_SUBSET = (16, 32, 64, 128)


def tests_wall():
    with open(os.path.join(REF, "tests-wall.pos")) as fh:
        n = int(fh.readline().split(":")[1])
        L = float(fh.readline().split(":")[1])
        pos = np.array([[float(v) for v in fh.readline().split()] for _ in range(n)])
    frc = np.loadtxt(os.path.join(REF, "tests-wall.for"))
    assert frc.shape == (n, 7) and np.all(frc[:, 0] == np.arange(n))
    np.savez_compressed(os.path.join(OUT, "tests_wall.npz"), L=L, x=pos[:, 0], y=pos[:, 1],
                        z=pos[:, 2], q=pos[:, 3], f_ewald=frc[:, 1:4].T.copy())


def nacl(kind, k):
    base = os.path.join(REF, "Data/ReferenceDataVincent", kind)
    with open(os.path.join(base, "konfig_%d" % k)) as fh:
        n = int(fh.readline())
        L = float(fh.readline())
        pos = np.array([[float(v) for v in fh.readline().split()] for _ in range(n)])
    ex = np.loadtxt(os.path.join(base, "Konfig_%02d.exact" % (k + 1)))
    assert ex.shape == (n, 4)
    name = "nacl_%s_%d.npz" % ("crystal" if "crystal" in kind else "melt", k)
    np.savez_compressed(os.path.join(OUT, name), L=L, x=pos[:, 0], y=pos[:, 1], z=pos[:, 2],
                        q=pos[:, 3], f_ewald=ex[:, 1:4].T.copy())


def q_tables():
    with open(os.path.join(REF, "q.c")) as fh:
        text = fh.read()
    meshes = [int(v) for v in re.search(r"meshes\[\]\s*=\s*\{([^}]*)\}", text).group(1).split(",") if v.strip()]
    out = {"meshes": np.array(MESH_SUBSET), "caos": np.arange(1, 8)}
    for name in ("ik", "ik_i", "ad", "ad_i"):
        with open(os.path.join(REF, "q_%s.c" % name)) as fh:
            body = fh.read().split("=", 1)[1]
        vals = np.array([float(v) for v in re.findall(r"[-+]?\d+\.\d*e[-+]\d+", body)])
        tab = vals.reshape(len(meshes), 7, 101)
        out["Q_" + name] = np.stack([tab[meshes.index(m)] for m in MESH_SUBSET])
    np.savez_compressed(os.path.join(OUT, "q_tables.npz"), **out)


if __name__ == "__main__":
    tests_wall()
    for kind in ("NaCl_crystal_300K", "NaCl_melt"):
        nacl(kind, 0)
    q_tables()
    print("fixtures written to", OUT)
