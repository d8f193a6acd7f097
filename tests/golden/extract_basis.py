"""Write the basis-set data files this package ships.

The Gaussian94-format blocks are public EMSL Basis Set Exchange data
(cc-pVDZ: T.H. Dunning Jr., J. Chem. Phys. 90, 1007 (1989); STO-3G: Hehre,
Stewart and Pople, J. Chem. Phys. 51, 2657 (1969)).  They are taken element by
element from the basis files that accompany GTFock (data/cc-pvdz.gbs and
data/sto-3g.gbs), keeping only the elements the workloads use, so that the GPU
box (which has no copy of GTFock) can load them.

Run once (it writes into gtfock_amd/data/):
    python tests/golden/extract_basis.py /root/reference/data
"""
import os
import sys

ELEMENTS = ("H", "C", "N", "O")


def extract(src, dst, elements):
    with open(src) as fh:
        lines = fh.read().splitlines()
    header = lines[0].strip()          # "spherical" or "cartesian"
    blocks = {}
    i = 0
    while i < len(lines):
        tok = lines[i].split()
        if len(tok) == 2 and tok[1] == "0" and tok[0].isalpha():
            sym = tok[0]
            j = i + 1
            while j < len(lines) and lines[j].strip() != "****":
                j += 1
            blocks[sym] = lines[i:j + 1]
            i = j + 1
        else:
            i += 1
    out = [header, "", "****"]
    for e in elements:
        out.extend(blocks[e])
    with open(dst, "w") as fh:
        fh.write("\n".join(out) + "\n")


if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/data"
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "..", "gtfock_amd", "data")
    extract(os.path.join(ref, "cc-pvdz.gbs"), os.path.join(here, "cc-pvdz.gbs"), ELEMENTS)
    extract(os.path.join(ref, "sto-3g.gbs"), os.path.join(here, "sto-3g.gbs"), ELEMENTS)
