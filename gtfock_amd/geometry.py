"""Molecule generators for the workloads (geometries in Angstrom, .xyz text).

The .xyz layout is the one GTFock's data/*.xyz files use: atom count, a
second line (blank, or an integer total charge), then "Sym x y z" lines.
"""
import math

import numpy as np

TETRA = math.acos(-1.0 / 3.0)  # 109.4712 degrees


def to_xyz(atoms, charge=None):
    lines = ["%d" % len(atoms), "" if charge is None else "%d" % charge]
    for sym, (x, y, z) in atoms:
        lines.append("%-2s %20.12f %20.12f %20.12f" % (sym, x, y, z))
    return "\n".join(lines) + "\n"


def water():
    """O at the origin, O-H 0.9572 A, H-O-H 104.52 degrees."""
    r, th = 0.9572, math.radians(104.52)
    hx, hz = r * math.sin(th / 2), r * math.cos(th / 2)
    return [("O", (0.0, 0.0, 0.0)), ("H", (hx, 0.0, hz)), ("H", (-hx, 0.0, hz))]


def _tetra_pair(c, n1, n2, r):
    """Two substituents completing a tetrahedron at c whose other neighbours
    are n1 and n2 (directions): symmetric about the n1-c-n2 plane."""
    u1 = (n1 - c) / np.linalg.norm(n1 - c)
    u2 = (n2 - c) / np.linalg.norm(n2 - c)
    b = -(u1 + u2)
    b /= np.linalg.norm(b)
    nrm = np.cross(u1, u2)
    nrm /= np.linalg.norm(nrm)
    half = TETRA / 2
    return [c + r * (b * math.cos(half) + s * nrm * math.sin(half)) for s in (1.0, -1.0)]


def alkane(n, rcc=1.54, rch=1.09):
    """All-trans CnH(2n+2): planar zig-zag backbone, tetrahedral angles."""
    if n < 1:
        raise ValueError("alkane needs n >= 1")
    if n == 1:
        d = rch / math.sqrt(3.0)
        return [("C", (0.0, 0.0, 0.0))] + [("H", (d * sx, d * sy, d * sx * sy))
                                          for sx, sy in ((1, 1), (1, -1), (-1, 1), (-1, -1))]
    half = TETRA / 2
    dx, dy = rcc * math.sin(half), rcc * math.cos(half)
    # backbone carbons, with one virtual carbon continuing the zig-zag at each end
    cs = [np.array([i * dx, (dy / 2) * (1 if i % 2 == 0 else -1), 0.0]) for i in range(-1, n + 1)]
    atoms = [("C", tuple(c)) for c in cs[1:-1]]
    hs = []
    for k in range(1, n + 1):
        c, prev, nxt = cs[k], cs[k - 1], cs[k + 1]
        hs.extend(_tetra_pair(c, prev, nxt, rch))
        if k == 1:                       # terminal: H in place of the virtual carbon
            hs.append(c + rch * (prev - c) / np.linalg.norm(prev - c))
        if k == n:
            hs.append(c + rch * (nxt - c) / np.linalg.norm(nxt - c))
    atoms += [("H", tuple(h)) for h in hs]
    return atoms


def graphene_flake(rings, rcc=1.42, rch=1.09):
    """Hexagonal H-capped flake: rings = 1 (benzene), 2 (C24H12 coronene),
    3 (C54H18 circumcoronene), 4 (C96H24).  Edge carbons (two C neighbours)
    carry one H along the outward bisector of their two C-C bonds."""
    a1 = np.array([1.5 * rcc, math.sqrt(3) / 2 * rcc])
    a2 = np.array([1.5 * rcc, -math.sqrt(3) / 2 * rcc])
    corners = [rcc * np.array([math.cos(math.pi / 3 * k), math.sin(math.pi / 3 * k)]) for k in range(6)]
    pts = []
    R = rings - 1
    for i in range(-R, R + 1):
        for j in range(-R, R + 1):
            if abs(i) <= R and abs(j) <= R and abs(i + j) <= R:
                # hexagon centre; a1 and a2 are 60 degrees apart, so the hex
                # distance from the centre hexagon is max(|i|, |j|, |i + j|)
                center = i * a1 + j * a2
                for c in corners:
                    pts.append(center + c)
    uniq = []
    for p in pts:
        if not any(np.linalg.norm(p - q) < 1e-6 for q in uniq):
            uniq.append(p)
    carbons = sorted(uniq, key=lambda p: (round(math.atan2(p[1], p[0]), 9), round(np.linalg.norm(p), 9)))
    atoms = [("C", (p[0], p[1], 0.0)) for p in carbons]
    for p in carbons:
        nb = [q for q in carbons if 0 < np.linalg.norm(p - q) < rcc * 1.1]
        if len(nb) == 2:
            b = 2 * p - nb[0] - nb[1]
            b /= np.linalg.norm(b)
            h = p + rch * b
            atoms.append(("H", (h[0], h[1], 0.0)))
    return atoms


def formula(atoms):
    cnt = {}
    for s, _ in atoms:
        cnt[s] = cnt.get(s, 0) + 1
    return "".join("%s%s" % (s, cnt[s] if cnt[s] > 1 else "") for s in sorted(cnt, key=lambda s: (s != "C", s)))


def random_rotation(rng):
    """Haar-random rotation matrix (QR of a Gaussian matrix, sign-fixed)."""
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q *= np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    return q


def water_cluster(k, rng, min_oo=2.7):
    """k water monomers (C1 geometry) at Haar-random orientations, O atoms
    uniform in a sphere of radius 1.6 * k^(1/3) + 1.4 A, rejection-sampled
    for O-O >= min_oo."""
    radius = 1.6 * k ** (1.0 / 3.0) + 1.4
    mono = water()
    os_ = []
    atoms = []
    tries = 0
    while len(os_) < k:
        tries += 1
        if tries > 100000:
            radius *= 1.1
            tries = 0
        p = rng.uniform(-radius, radius, size=3)
        if np.linalg.norm(p) > radius:
            continue
        if any(np.linalg.norm(p - q) < min_oo for q in os_):
            continue
        os_.append(p)
        R = random_rotation(rng)
        for s, xyz in mono:
            v = R @ np.array(xyz) + p
            atoms.append((s, tuple(v)))
    return atoms


def batch_molecules(nmol, seed=20261202, max_alkane=20, max_cluster=16):
    """A seeded batch of nmol molecules: the first half all-trans alkanes
    CnH2n+2 with n uniform in 1..max_alkane, the second half water clusters
    of k uniform in 1..max_cluster monomers.  Returns [(name, atoms)]."""
    rng = np.random.default_rng(seed)
    out = []
    na = nmol // 2
    for _ in range(na):
        n = int(rng.integers(1, max_alkane + 1))
        out.append(("C%dH%d" % (n, 2 * n + 2), alkane(n)))
    for _ in range(nmol - na):
        k = int(rng.integers(1, max_cluster + 1))
        out.append(("(H2O)%d" % k, water_cluster(k, rng)))
    return out


def serve_molecules(njobs, seed=20261203, max_rings=4, flake_fraction=0.3):
    """A seeded SCF request stream: each job is, with probability
    1 - flake_fraction, a molecule of the batched distribution (an all-trans
    alkane with n uniform in 1..20 or a water cluster of k uniform in 1..16,
    half each), else an H-capped graphene flake with 2..max_rings rings
    (C24H12, C54H18, C96H24).  Returns [(name, atoms)]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(njobs):
        if rng.uniform() < flake_fraction:
            rings = int(rng.integers(2, max_rings + 1))
            atoms = graphene_flake(rings)
            out.append((formula(atoms), atoms))
        elif rng.uniform() < 0.5:
            n = int(rng.integers(1, 21))
            out.append(("C%dH%d" % (n, 2 * n + 2), alkane(n)))
        else:
            k = int(rng.integers(1, 17))
            out.append(("(H2O)%d" % k, water_cluster(k, rng)))
    return out
