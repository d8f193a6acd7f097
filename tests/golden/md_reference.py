"""Independent integral reference used only to write golden fixtures.

McMurchie-Davidson (J. Comput. Phys. 26, 218 (1978)) Hermite-expansion
integrals in plain Python, with the Boys function from mpmath's confluent
hypergeometric function F_m(T) = 1F1(m+1/2; m+3/2; -T)/(2m+1) at 40 digits.
It shares no code and no recursion with the oracle (which uses Obara-Saika
vertical and Head-Gordon-Pople horizontal recursions), so agreement between
the two pins both.

The basis conventions (primitive and contraction normalisation, real solid
harmonics ordered m = -l..l, Angstrom -> bohr with CODATA 2018) are restated
here from their definitions, not imported from the oracle.
"""
import math
from functools import lru_cache

import mpmath
import numpy as np

mpmath.mp.dps = 40
BOHR_PER_ANGSTROM = 1.0 / 0.529177210903
Z_OF = {"H": 1, "C": 6, "N": 7, "O": 8}


def boys_mp(m, T):
    return mpmath.hyp1f1(m + 0.5, m + 1.5, -mpmath.mpf(T)) / (2 * m + 1)


def boys(m, T):
    return float(boys_mp(m, T))


def dfact(n):
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def read_gbs(path):
    """Gaussian94 basis -> {element: [(L, [(alpha, coef), ...]), ...]}."""
    out = {}
    lines = [l.strip() for l in open(path)]
    i, cur = 0, None
    while i < len(lines):
        tok = lines[i].split()
        i += 1
        if not tok or tok[0].startswith("!"):
            continue
        if tok[0] == "****":
            cur = None
            continue
        if cur is None:
            if len(tok) == 2 and tok[1] == "0" and tok[0].isalpha():
                cur = tok[0]
                out[cur] = []
            continue
        kind, nprim = tok[0], int(tok[1])
        prims = []
        for _ in range(nprim):
            t = lines[i].replace("D", "E").split()
            i += 1
            prims.append([float(x) for x in t])
        if kind.upper() == "SP":
            out[cur].append((0, [(p[0], p[1]) for p in prims]))
            out[cur].append((1, [(p[0], p[2]) for p in prims]))
        else:
            out[cur].append(("SPDF".index(kind.upper()), [(p[0], p[1]) for p in prims]))
    return out


class Shell:
    def __init__(self, L, center, prims):
        self.L = L
        self.center = np.array(center, dtype=float)
        self.alpha = [a for a, _ in prims]
        # primitive normalisation for x^L exp(-a r^2), then contraction renorm
        c = [cf * math.sqrt((4 * a) ** L * (2 * a / math.pi) ** 1.5 / dfact(2 * L - 1)) for a, cf in prims]
        s = 0.0
        for i, ai in enumerate(self.alpha):
            for j, aj in enumerate(self.alpha):
                p = ai + aj
                s += c[i] * c[j] * dfact(2 * L - 1) / (2 * p) ** L * (math.pi / p) ** 1.5
        self.coef = [x / math.sqrt(s) for x in c]

    def carts(self):
        L = self.L
        return [(lx, L - lx - lz, lz) for lx in range(L, -1, -1) for lz in range(0, L - lx + 1)]


def sph_rows(L):
    """Real solid harmonics m = -L..L as rows over the cartesian order above."""
    if L == 0:
        return np.array([[1.0]])
    if L == 1:
        return np.eye(3)
    r3 = math.sqrt(3.0)
    # cart order xx xy xz yy yz zz
    return np.array([
        [0, r3, 0, 0, 0, 0],
        [0, 0, 0, 0, r3, 0],
        [-0.5, 0, 0, -0.5, 0, 1.0],
        [0, 0, r3, 0, 0, 0],
        [0.5 * r3, 0, 0, -0.5 * r3, 0, 0],
    ])


def molecule_shells(gbs, xyz_text):
    basis = read_gbs(gbs)
    rows = xyz_text.strip().splitlines()
    n = int(rows[0])
    atoms = []
    for r in rows[2:2 + n]:
        s, x, y, z = r.split()[:4]
        atoms.append((s, np.array([float(x), float(y), float(z)]) * BOHR_PER_ANGSTROM))
    shells = []
    for s, pos in atoms:
        for L, prims in basis[s]:
            shells.append(Shell(L, pos, prims))
    return atoms, shells


def E_coef(i, j, t, Qx, a, b):
    """Hermite expansion coefficient E^{ij}_t (Helgaker, Joergensen, Olsen 9.5)."""
    p = a + b
    q = a * b / p
    if t < 0 or t > i + j:
        return 0.0
    if i == j == t == 0:
        return math.exp(-q * Qx * Qx)
    if j == 0:
        return (E_coef(i - 1, j, t - 1, Qx, a, b) / (2 * p)
                - (q * Qx / a) * E_coef(i - 1, j, t, Qx, a, b)
                + (t + 1) * E_coef(i - 1, j, t + 1, Qx, a, b))
    return (E_coef(i, j - 1, t - 1, Qx, a, b) / (2 * p)
            + (q * Qx / b) * E_coef(i, j - 1, t, Qx, a, b)
            + (t + 1) * E_coef(i, j - 1, t + 1, Qx, a, b))


def R_herm(tmax, p, PC):
    """All Hermite Coulomb integrals R_{tuv} with t+u+v <= tmax."""
    T = p * float(PC @ PC)
    F = [boys(n, T) for n in range(tmax + 1)]

    @lru_cache(maxsize=None)
    def R(t, u, v, n):
        if t < 0 or u < 0 or v < 0:
            return 0.0
        if t == u == v == 0:
            return (-2 * p) ** n * F[n]
        if t > 0:
            return (t - 1) * R(t - 2, u, v, n + 1) + PC[0] * R(t - 1, u, v, n + 1)
        if u > 0:
            return (u - 1) * R(t, u - 2, v, n + 1) + PC[1] * R(t, u - 1, v, n + 1)
        return (v - 1) * R(t, u, v - 2, n + 1) + PC[2] * R(t, u, v - 1, n + 1)

    return R


def _pair_E(sa, sb, a, b):
    """E^{ab}_{tuv} for all cartesian components: dict (ca, cb) -> array."""
    AB = sa.center - sb.center
    out = {}
    for ia, ca in enumerate(sa.carts()):
        for ib, cb in enumerate(sb.carts()):
            Ex = [E_coef(ca[0], cb[0], t, AB[0], a, b) for t in range(ca[0] + cb[0] + 1)]
            Ey = [E_coef(ca[1], cb[1], u, AB[1], a, b) for u in range(ca[1] + cb[1] + 1)]
            Ez = [E_coef(ca[2], cb[2], v, AB[2], a, b) for v in range(ca[2] + cb[2] + 1)]
            out[(ia, ib)] = (Ex, Ey, Ez)
    return out


def eri_shell_quartet(sa, sb, sc, sd):
    """Spherical (ab|cd) block as a [a][b][c][d] numpy array."""
    na, nb, nc, nd = (len(s.carts()) for s in (sa, sb, sc, sd))
    cart = np.zeros((na, nb, nc, nd))
    Lt = sa.L + sb.L + sc.L + sd.L
    for i, a in enumerate(sa.alpha):
        for j, b in enumerate(sb.alpha):
            p = a + b
            P = (a * sa.center + b * sb.center) / p
            Eab = _pair_E(sa, sb, a, b)
            cab = sa.coef[i] * sb.coef[j]
            for k, c in enumerate(sc.alpha):
                for l, d in enumerate(sd.alpha):
                    q = c + d
                    Q = (c * sc.center + d * sd.center) / q
                    Ecd = _pair_E(sc, sd, c, d)
                    alpha = p * q / (p + q)
                    R = R_herm(Lt, alpha, P - Q)
                    pref = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * cab * sc.coef[k] * sd.coef[l]
                    for (ia, ib), (Ex, Ey, Ez) in Eab.items():
                        for (ic, id_), (Fx, Fy, Fz) in Ecd.items():
                            s = 0.0
                            for t, ex in enumerate(Ex):
                                for u, ey in enumerate(Ey):
                                    for v, ez in enumerate(Ez):
                                        e1 = ex * ey * ez
                                        if e1 == 0.0:
                                            continue
                                        for tau, fx in enumerate(Fx):
                                            for nu, fy in enumerate(Fy):
                                                for phi, fz in enumerate(Fz):
                                                    sgn = -1.0 if (tau + nu + phi) % 2 else 1.0
                                                    s += e1 * sgn * fx * fy * fz * R(t + tau, u + nu, v + phi, 0)
                            cart[ia, ib, ic, id_] += pref * s
    Ta, Tb, Tc, Td = (sph_rows(s.L) for s in (sa, sb, sc, sd))
    return np.einsum("ai,bj,ck,dl,ijkl->abcd", Ta, Tb, Tc, Td, cart)


def one_electron(sa, sb, atoms):
    """(S, T, V) spherical blocks by McMurchie-Davidson."""
    na, nb = len(sa.carts()), len(sb.carts())
    S = np.zeros((na, nb))
    Tk = np.zeros((na, nb))
    V = np.zeros((na, nb))
    AB = sa.center - sb.center
    for i, a in enumerate(sa.alpha):
        for j, b in enumerate(sb.alpha):
            p = a + b
            P = (a * sa.center + b * sb.center) / p
            cc = sa.coef[i] * sb.coef[j]
            Rs = [(Z_OF[s], R_herm(sa.L + sb.L, p, P - pos)) for s, pos in atoms]
            for ia, ca in enumerate(sa.carts()):
                for ib, cb in enumerate(sb.carts()):
                    def s1(d, jj):
                        if jj < 0:
                            return 0.0
                        return E_coef(ca[d], jj, 0, AB[d], a, b) * math.sqrt(math.pi / p)

                    def t1(d):
                        jj = cb[d]
                        return (-2 * b * b * s1(d, jj + 2) + b * (2 * jj + 1) * s1(d, jj)
                                - 0.5 * jj * (jj - 1) * s1(d, jj - 2))

                    sx, sy, sz = s1(0, cb[0]), s1(1, cb[1]), s1(2, cb[2])
                    S[ia, ib] += cc * sx * sy * sz
                    Tk[ia, ib] += cc * (t1(0) * sy * sz + sx * t1(1) * sz + sx * sy * t1(2))
                    Ex = [E_coef(ca[0], cb[0], t, AB[0], a, b) for t in range(ca[0] + cb[0] + 1)]
                    Ey = [E_coef(ca[1], cb[1], u, AB[1], a, b) for u in range(ca[1] + cb[1] + 1)]
                    Ez = [E_coef(ca[2], cb[2], v, AB[2], a, b) for v in range(ca[2] + cb[2] + 1)]
                    for Zc, R in Rs:
                        s = 0.0
                        for t, ex in enumerate(Ex):
                            for u, ey in enumerate(Ey):
                                for v, ez in enumerate(Ez):
                                    s += ex * ey * ez * R(t, u, v, 0)
                        V[ia, ib] -= Zc * cc * 2 * math.pi / p * s
    Ta, Tb = sph_rows(sa.L), sph_rows(sb.L)
    return Ta @ S @ Tb.T, Ta @ Tk @ Tb.T, Ta @ V @ Tb.T
