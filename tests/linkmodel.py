"""Independent numpy model of ASL's LBGK boundary semantics (test infrastructure).

ASL fills ghost entries (g, k) after the collision and then pull-streams
(src/num/aslLBGKBC.cxx, src/num/aslLBGK.cxx:237-250).  This model states the
same thing per link instead: the population a computation node x receives
along k is the value of the last executed BC that has an entry (x - c_k, k),
else the collided population of x - c_k.  It collides in the textbook
velocity form, FP64, on interior computation nodes only, and raises if a
link reads a cell it does not compute.  Agreement with the oracle's ghost-fill
restatement pins the BC semantics both the oracle and the GPU path implement.
"""
import numpy as np

from oracle import template

NOSLIP, CONSTANT_PRESSURE, CONSTANT_VELOCITY, CONSTANT_PRESSURE_VELOCITY = 1, 2, 3, 4
NOSLIP_MAP, PRESSURE_VELOCITY_MAP, VELOCITY_MAP = 5, 6, 7
FACE_DIR = [(0, +1), (0, -1), (1, +1), (1, -1), (2, +1), (2, -1)]  # X0 XE Y0 YE Z0 ZE: axis, inward sign


def feq(rho, u, c):
    cu = np.einsum("a,a...->...", c, u)
    return rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * (u * u).sum(0))


class LinkModel:
    def __init__(self, size, q, nu, compressible=True):
        self.size = tuple(size)
        self.S = tuple(s + 2 for s in size)
        self.q = q
        self.c, self.w, self.inv = template(q)
        self.nu = nu
        self.compressible = compressible
        self.bcs = []

    def init(self, rho, u):
        rho = np.broadcast_to(np.asarray(rho, dtype=float), self.S)
        u = np.broadcast_to(np.asarray(u, dtype=float).reshape((3,) + (1,) * 3 if np.ndim(u) == 1 else (3,) + self.S), (3,) + self.S)
        self.f = np.stack([feq(rho, u, self.c[i]) for i in range(self.q)])

    def add(self, kind, faces=0, mapv=None, p=0.0, u=(0.0, 0.0, 0.0), rot=None):
        self.bcs.append(dict(kind=kind, faces=faces, map=None if mapv is None else np.asarray(mapv, np.float32),
                             p=float(p), u=np.asarray(u, float), rot=None if rot is None else np.asarray(rot, float)))
        return len(self.bcs) - 1

    def _value(self, b, x, k, f):
        """BC value of the link into interior cell x (tuple) along k, from post-collision f."""
        c, inv = self.c, self.inv
        kind = b["kind"]
        uc = float(np.dot(b["u"], c[k]))
        if kind in (NOSLIP, NOSLIP_MAP):
            return f[inv[k]][x]
        if kind == CONSTANT_VELOCITY:
            return f[inv[k]][x] + 6 * uc
        if kind == CONSTANT_PRESSURE:
            return b["p"]
        if kind == CONSTANT_PRESSURE_VELOCITY:
            return b["p"] + 3 * uc
        if kind == VELOCITY_MAP:
            # rho at the ghost node g from the populations its computation-node
            # neighbours sent towards it, momentum = u_wall(pos_g) rho
            g = np.array(x) - c[k]
            num = den = 0.0
            for j in range(1, self.q):
                nb = tuple(g + c[j])
                if b["map"][nb] > 0:
                    num += self.w[j] * f[inv[j]][nb]
                    den += self.w[j]
            rho_g = num / den
            pos = g - 1.0  # ghost-inclusive index -> position (dx = 1, block at 0)
            uw = b["u"] if b["rot"] is None else np.cross(b["rot"][:3], pos - b["rot"][3:])
            return rho_g + 3 * rho_g * float(np.dot(uw, c[k]))
        if kind == PRESSURE_VELOCITY_MAP:
            p, uu = b["p"], float(np.dot(b["u"], b["u"]))
            if self.compressible:
                return p * (1 + 3 * uc + 4.5 * uc * uc - 1.5 * uu)
            return p + 3 * uc + 4.5 * uc * uc - 1.5 * uu
        raise ValueError(kind)

    def _claims(self, b, x, k):
        """Does BC b have an entry (x - c_k, k)?"""
        ck = self.c[k]
        if b["map"] is None:
            for f in range(6):
                if not (b["faces"] >> f) & 1:
                    continue
                axis, sgn = FACE_DIR[f]
                layer = 1 if sgn > 0 else self.S[axis] - 2
                if x[axis] == layer and ck[axis] * sgn > 0:
                    return True
            return False
        m = b["map"]
        g = tuple(np.array(x) - ck)
        return bool(m[g] <= 0 and m[g] > np.float32(-0.9999) and m[x] > 0)

    def run(self, steps, order):
        S, q, c = self.S, self.q, self.c
        solid = np.zeros(S, dtype=bool)
        for b in self.bcs:
            if b["map"] is not None:
                solid |= b["map"] <= 0
        comp = np.zeros(S, dtype=bool)
        comp[1:-1, 1:-1, 1:-1] = True
        comp &= ~solid
        cells = list(zip(*np.nonzero(comp)))
        # link table: for each cell and k, the last BC claiming it (or None)
        table = {}
        for x in cells:
            for k in range(1, q):
                who = None
                for bi in order:
                    if self._claims(self.bcs[bi], x, k):
                        who = bi
                table[(x, k)] = who
                if who is None:
                    src = tuple(np.array(x) - c[k])
                    if not comp[src]:
                        raise ValueError("link %s <- %s along %d reads an uncomputed cell" % (x, src, k))
        om = 1.0 / (3.0 * self.nu + 0.5)
        f = self.f
        for _ in range(steps):
            fin = np.zeros_like(f)
            for x in cells:
                fin[0][x] = f[0][x]
                for k in range(1, q):
                    who = table[(x, k)]
                    if who is None:
                        fin[k][x] = f[k][tuple(np.array(x) - c[k])]
                    else:
                        fin[k][x] = self._value(self.bcs[who], x, k, f)
            rho = np.einsum("i,i...->...", self.w, fin)
            m = np.einsum("i,ia,i...->a...", self.w, c, fin)
            fn = f.copy()
            with np.errstate(invalid="ignore", divide="ignore"):
                u = m / rho
            for i in range(q):
                if self.compressible:
                    fe = feq(rho, u, c[i])
                else:
                    cm = np.einsum("a,a...->...", c[i], m)
                    fe = rho + 3 * cm + 4.5 * cm * cm - 1.5 * (m * m).sum(0)
                upd = fin[i] - om * (fin[i] - fe)
                fn[i][comp] = upd[comp]
            f = fn
            self.rho, self.m = rho, m
        self.f = f
        self.comp = comp
        return comp
