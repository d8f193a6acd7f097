"""Write the golden fixtures in tests/golden/ from independent references.

  boys.npz        Boys F_m(T) from mpmath (40 digits)
  eri_water.npz   contracted spherical ERI blocks of water / cc-pVDZ, one shell
                  quartet per ordered (La,Lb,Lc,Ld) combination, computed by
                  the McMurchie-Davidson code in md_reference.py
  oneel_water.npz S, T, V of water / cc-pVDZ from the same code
  h2_sto3g.json   textbook H2 / STO-3G values at R = 1.4 bohr (Szabo and
                  Ostlund, Modern Quantum Chemistry, section 3.5.2), typed in

Run:  python tests/golden/make_golden.py   (takes a few minutes)
"""
import itertools
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, ROOT)

import md_reference as md  # noqa: E402
from gtfock_amd import geometry  # noqa: E402

GBS = os.path.join(ROOT, "gtfock_amd", "data", "cc-pvdz.gbs")


def boys_fixture():
    Ts = [0.0, 1e-10, 1e-6, 0.01, 0.3, 1.0, 2.5, 5.0, 7.77, 10.0, 15.0, 20.0, 25.0,
          29.9, 33.3, 34.99, 35.01, 36.5, 40.0, 50.0, 80.0, 150.0, 1000.0]
    ms = list(range(0, 17))
    F = np.array([[float(md.boys_mp(m, T)) for m in ms] for T in Ts])
    np.savez(os.path.join(HERE, "boys.npz"), T=np.array(Ts), m=np.array(ms), F=F)


def eri_fixture():
    atoms, shells = md.molecule_shells(GBS, geometry.to_xyz(geometry.water()))
    byL = {0: [2, 7, 10, 6, 9, 0, 1], 1: [4, 8, 11, 3], 2: [5]}
    quartets = []
    for k, Ls in enumerate(itertools.product(range(3), repeat=4)):
        q = [byL[L][(k + pos) % len(byL[L])] for pos, L in enumerate(Ls)]
        quartets.append(q)
    # heavily contracted s shells on one and two centres
    quartets += [[0, 0, 0, 0], [0, 1, 6, 9], [1, 6, 0, 9], [3, 0, 3, 6]]
    blocks, offs = [], [0]
    for q in quartets:
        blk = md.eri_shell_quartet(*(shells[s] for s in q))  # [a][b][c][d]
        flat = blk.transpose(3, 2, 1, 0).reshape(-1)          # iM fastest
        blocks.append(flat)
        offs.append(offs[-1] + flat.size)
        print("eri", q, flat.size, flush=True)
    np.savez(os.path.join(HERE, "eri_water.npz"), quartets=np.array(quartets, dtype=np.int32),
             offsets=np.array(offs, dtype=np.int64), values=np.concatenate(blocks))


def oneel_fixture():
    atoms, shells = md.molecule_shells(GBS, geometry.to_xyz(geometry.water()))
    dims = [2 * s.L + 1 for s in shells]
    off = np.concatenate([[0], np.cumsum(dims)])
    n = off[-1]
    S, T, V = (np.zeros((n, n)) for _ in range(3))
    for i, si in enumerate(shells):
        for j, sj in enumerate(shells):
            s, t, v = md.one_electron(si, sj, atoms)
            S[off[i]:off[i + 1], off[j]:off[j + 1]] = s
            T[off[i]:off[i + 1], off[j]:off[j + 1]] = t
            V[off[i]:off[i + 1], off[j]:off[j + 1]] = v
    np.savez(os.path.join(HERE, "oneel_water.npz"), S=S, T=T, V=V)


def h2_fixture():
    # Szabo & Ostlund, sec. 3.5.2, STO-3G (zeta = 1.24), R = 1.4 bohr; four
    # decimals as printed.  V_ij^(k): attraction to nucleus k.
    vals = {
        "R_bohr": 1.4,
        "S12": 0.6593, "T11": 0.7600, "T12": 0.2365,
        "V11_1": -1.2266, "V12_1": -0.5974, "V22_1": -0.6538,
        "H11": -1.1204, "H12": -0.9584,
        "eri_1111": 0.7746, "eri_1122": 0.5697, "eri_2111": 0.4441, "eri_2121": 0.2970,
        "E_total": -1.1167, "tolerance": 1.0e-4,
    }
    with open(os.path.join(HERE, "h2_sto3g.json"), "w") as fh:
        json.dump(vals, fh, indent=1)


if __name__ == "__main__":
    what = sys.argv[1:] or ["boys", "h2", "oneel", "eri"]
    if "boys" in what:
        boys_fixture()
    if "h2" in what:
        h2_fixture()
    if "oneel" in what:
        oneel_fixture()
    if "eri" in what:
        eri_fixture()
