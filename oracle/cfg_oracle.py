"""Test infrastructure, not part of the engine: a pure-Python restatement of the reference's
configuration I/O (lmc/cfg/src/Config.cpp) used as the checker of the engine's lmc_config_*.

It follows the reference's own route, independently of the engine's closed-form ids:
- GenerateFCC (Config.cpp:1060-1095): floating-point relative positions in creation order;
- ReassignLatticeVector (Config.cpp:466-552): lattice ids by sorting positions with the
  tolerance comparator of VectorMatrix.hpp:78-88 (kEpsilon = 1e-4);
- ReadConfig (Config.cpp:554-661): the istream token sequence;
- WriteConfig (Config.cpp:663-710): precision(16), general format until the first
  position switches the stream to std::fixed (VectorMatrix.hpp:41-51);
- UpdateNeighbors (Config.cpp:955-1045): neighbour shells by brute-force distances;
- LatticeJump (Config.cpp:431-456): map_shift from the floating-point displacement.
"""
import functools
import gzip

import numpy as np

K_EPSILON = 1e-4
LATTICE_CONSTANT = 4.046
MASS = {"X": 0.0, "Al": 26.98, "Mg": 24.31, "Zn": 65.38, "Cu": 63.55, "Sn": 118.71,
        "pAl": 26.98, "pMg": 24.31, "pZn": 65.38, "pCu": 63.55, "pSn": 118.71}


def _less(a, b):
    """operator<(Vector_d, Vector_d) with tolerance (VectorMatrix.hpp:78-88)."""
    for d in range(2):
        diff = a[d] - b[d]
        if diff < -K_EPSILON:
            return True
        if diff > K_EPSILON:
            return False
    return a[2] < b[2] - K_EPSILON


def _cmp(a, b):
    return -1 if _less(a, b) else (1 if _less(b, a) else 0)


class OracleConfig:
    """positions[lattice] (relative), element[atom], atom_to_lattice[atom], map_shift[atom]."""

    def __init__(self, basis, positions, elements, atom_to_lattice, map_shift):
        self.basis = np.asarray(basis, float)
        self.positions = [tuple(map(float, p)) for p in positions]
        self.elements = list(elements)
        self.atom_to_lattice = list(atom_to_lattice)
        self.map_shift = [list(map(int, s)) for s in map_shift]

    @property
    def lattice_to_atom(self):
        out = [0] * len(self.atom_to_lattice)
        for a, s in enumerate(self.atom_to_lattice):
            out[s] = a
        return out

    def reassign(self):
        """ReassignLatticeVector: new lattice ids in sorted position order."""
        order = sorted(range(len(self.positions)), key=functools.cmp_to_key(
            lambda i, j: _cmp(self.positions[i], self.positions[j])))
        new_of_old = [0] * len(order)
        for new, old in enumerate(order):
            new_of_old[old] = new
        self.positions = [self.positions[old] for old in order]
        self.atom_to_lattice = [new_of_old[s] for s in self.atom_to_lattice]

    def lattice_jump(self, lhs, rhs):
        """LatticeJump(lhs, rhs) on lattice ids: swap the atoms and update map_shift."""
        l2a = self.lattice_to_atom
        a_l, a_r = l2a[lhs], l2a[rhs]
        disp = [self.positions[rhs][d] - self.positions[lhs][d] for d in range(3)]
        ic = [int(np.floor(s + 0.5)) if abs(s) > 0.5 + 1e-8 else 0 for s in disp]
        for d in range(3):
            self.map_shift[a_l][d] -= ic[d]
            self.map_shift[a_r][d] += ic[d]
        self.atom_to_lattice[a_l], self.atom_to_lattice[a_r] = rhs, lhs

    def write_text(self):
        """WriteConfig."""
        out = [f"Number of particles = {len(self.elements)}\n",
               "A = 1.0 Angstrom (basic length-scale)\n"]
        for r in range(3):
            for c in range(3):
                out.append(f"H0({r + 1},{c + 1}) = {'%.16g' % self.basis[r, c]} A\n")
        out.append(".NO_VELOCITY.\nentry_count = 6\n"
                   "auxiliary[0] = ix\nauxiliary[1] = iy\nauxiliary[2] = iz\n")
        fixed = False
        for a, e in enumerate(self.elements):
            out.append(("%.16f" if fixed else "%.16g") % MASS.get(e, 0.0) + "\n" + e + "\n")
            p = self.positions[self.atom_to_lattice[a]]
            s = self.map_shift[a]
            out.append("%.16f %.16f %.16f %d %d %d\n" % (*p, *s))
            fixed = True
        return "".join(out)

    def write_xyz_text(self):
        """WriteExtendedXyz with no auxiliary or global lists (Config.cpp:712-809)."""
        B = self.basis
        lat = " ".join("%.16f %.16f %.16f" % tuple(B[r]) for r in range(3))
        out = [f"{len(self.elements)}\n",
               f'Lattice="{lat}" pbc="T T T" Properties=species:S:1:pos:R:3:map_shift:I:3\n']
        for a, e in enumerate(self.elements):
            p = self.positions[self.atom_to_lattice[a]]
            cart = [p[0] * B[0, i] + p[1] * B[1, i] + p[2] * B[2, i] for i in range(3)]
            out.append(f"{e} " + "%.16f %.16f %.16f" % tuple(cart) + " %d %d %d \n" % tuple(self.map_shift[a]))
        return "".join(out)

    def neighbour_shells(self, s):
        """Ids of the 1st/2nd/3rd neighbour shells of lattice s, ascending (UpdateNeighbors),
        by minimum-image distances in units of a/2."""
        L = np.array([self.basis[d, d] for d in range(3)])
        n = L[0] / LATTICE_CONSTANT
        P = np.array(self.positions)
        dv = P - P[s]
        dv -= np.round(dv)
        r2 = np.rint(((dv * 2 * n) ** 2).sum(1)).astype(int)
        return [list(np.flatnonzero(r2 == k)) for k in (2, 4, 6)]

    def write_lattice_text(self):
        """WriteLattice."""
        out = [f"{len(self.positions)} positions in total\n"]
        for r in range(3):
            out.append("%.16f %.16f %.16f\n" % tuple(self.basis[r]))
        for s, p in enumerate(self.positions):
            ids = [i for shell in self.neighbour_shells(s) for i in shell]
            out.append("%.16f %.16f %.16f # " % p + "".join(f"{i} " for i in ids) + "\n")
        return "".join(out)


def generate_fcc(n, element):
    """GenerateFCC({n, n, n}, element): lattice and atom ids in creation order."""
    basis = np.diag([LATTICE_CONSTANT * n] * 3)
    pos = []
    for k in range(n):
        for j in range(n):
            for i in range(n):
                x, y, z = float(i), float(j), float(k)
                L = float(n)
                pos += [(x / L, y / L, z / L), ((x + 0.5) / L, (y + 0.5) / L, z / L),
                        ((x + 0.5) / L, y / L, (z + 0.5) / L), (x / L, (y + 0.5) / L, (z + 0.5) / L)]
    N = len(pos)
    return OracleConfig(basis, pos, [element] * N, list(range(N)), [[0, 0, 0]] * N)


def read_text(text):
    """ReadConfig: the reference's token sequence (lattice ids in file order)."""
    i = 0

    def past(ch):
        nonlocal i
        k = text.find(ch, i)
        i = len(text) if k < 0 else k + 1

    def token():
        nonlocal i
        while i < len(text) and text[i].isspace():
            i += 1
        j = i
        while j < len(text) and not text[j].isspace():
            j += 1
        t, i = text[i:j], j
        return t

    past("=")
    N = int(token())
    past("=")
    basis = []
    for _ in range(9):
        past("=")
        basis.append(float(token()))
    past("\n")
    no_vel = False
    if i < len(text) and text[i] == ".":
        no_vel = True
        past("\n")
    past("=")
    entry_count = int(token())
    past("\n")
    while i < len(text):
        while i < len(text) and text[i] in " \t\r\v\f":
            i += 1
        if i >= len(text) or text[i].isdigit() or text[i] in ".-":
            break
        past("\n")
    pos, els, shifts = [], [], []
    for _ in range(N):
        token()  # mass
        els.append(token())
        pos.append(tuple(float(token()) for _ in range(3)))
        shifts.append([int(token()) for _ in range(3)] if entry_count >= 6 and no_vel else [0, 0, 0])
        past("\n")
    return OracleConfig(np.array(basis).reshape(3, 3), pos, els, list(range(N)), shifts)


def read_file(path):
    op = gzip.open if str(path).endswith(".gz") else open
    with op(path, "rt") as f:
        return read_text(f.read())
