"""Configuration files (cfg::Config I/O, lmc/cfg/src/Config.cpp:431-1095): the engine's
lmc_config_* against the pure-Python restatement oracle/cfg_oracle.py, whose formatting is pinned
to the C++ iostream it restates and whose LatticeJump image rule is pinned to its closed form."""
import gzip
import subprocess

import numpy as np
import pytest

import cfg_oracle as CO

from latticemontecarlo_amd import capi
from latticemontecarlo_amd import cfg as CF

SPECIES = ["Al", "Mg", "X", "Zn"]


def test_stream_formats_match_iostream(tmp_path):
    """The oracle writes '%.16g' before the first vector and '%.16f' after it: the output of an
    ostream with precision(16) in the default float format and after std::fixed."""
    vals = [26.98, 24.31, 65.38, 0.0, 16.184, 40.46, 0.125, 1 / 6, 5 / 6, 121.38, 1e-20, 123456789.5]
    src = tmp_path / "fmt.cpp"
    src.write_text(
        "#include <iostream>\n#include <cstdlib>\nint main(int c, char **v) {\n"
        "  std::cout.precision(16);\n"
        "  for (int i = 1; i < c; ++i) std::cout << std::strtod(v[i], nullptr) << '\\n';\n"
        "  std::cout << std::fixed;\n"
        "  for (int i = 1; i < c; ++i) std::cout << std::strtod(v[i], nullptr) << '\\n';\n"
        "}\n")
    exe = tmp_path / "fmt"
    subprocess.run(["g++", "-O1", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), *[repr(v) for v in vals]], check=True, capture_output=True,
                         text=True).stdout.split("\n")
    assert out[:len(vals)] == ["%.16g" % v for v in vals]
    assert out[len(vals):2 * len(vals)] == ["%.16f" % v for v in vals]


def test_lattice_jump_image_rule_closed_form():
    """The reference's floating-point image change floor(shift + 0.5) for |shift| > 0.5 + 1e-8
    equals the integer rule the kernels use, (d > n) - (d < -n) with d the a/2-grid difference,
    for every ordered pair of sites of the 4- and 5-cell boxes."""
    for n in (4, 5):
        m = 2 * n
        X = np.arange(m)
        rel = X / float(m)
        d = X[None, :] - X[:, None]
        shift = rel[None, :] - rel[:, None]
        ref = np.where(np.abs(shift) > 0.5 + 1e-8, np.floor(shift + 0.5), 0).astype(int)
        assert np.array_equal(ref, (d > n).astype(int) - (d < -n).astype(int))


def test_generate_matches_reference_route(tmp_path):
    """GenerateFCC + ReassignLatticeVector: the engine's closed-form lattice ids and
    GenerateFCC atom numbering against the oracle's floating-point positions and tolerance
    sort; WriteConfig byte for byte."""
    n = 4
    c = CF.generate_fcc(n, "Al")
    o = CO.generate_fcc(n, "Al")
    o.reassign()
    l2a, shift = c.atoms()
    assert np.array_equal(l2a, o.lattice_to_atom)
    assert not shift.any()
    assert np.allclose(c.basis, o.basis, rtol=0, atol=0)
    # closed-form positions of the x-major ids
    P = np.array(o.positions) * 2 * n
    ids = P[:, 0].round() * 2 * n * n + P[:, 1].round() * n + (P[:, 2].round() // 2)
    assert np.array_equal(ids, np.arange(4 * n ** 3))
    rng = np.random.default_rng(5)
    codes = rng.choice([0, 0, 0, 1, 3], size=c.num_sites).astype(np.uint8)
    codes[17] = SPECIES.index("X")
    c.update(SPECIES, codes)
    for s, a in enumerate(o.lattice_to_atom):
        o.elements[a] = SPECIES[codes[s]]
    c.write(tmp_path / "g.cfg")
    assert (tmp_path / "g.cfg").read_text() == o.write_text()
    assert np.array_equal(c.codes(SPECIES), codes)


def _scrambled(n, seed):
    """An oracle configuration with random elements, a shuffled atom numbering and map_shifts
    from a random walk of LatticeJumps (lattice ids still in creation order)."""
    rng = np.random.default_rng(seed)
    o = CO.generate_fcc(n, "Al")
    N = len(o.elements)
    o.elements = list(rng.choice(["Al", "Al", "Al", "Mg", "Zn"], size=N))
    perm = rng.permutation(N)
    o.atom_to_lattice = [int(perm[a]) for a in range(N)]
    P = np.array(o.positions)
    for _ in range(300):
        lhs = int(rng.integers(N))
        dv = P - P[lhs]
        dv -= np.round(dv)
        nb = np.flatnonzero(np.abs((dv ** 2).sum(1) * (2 * n) ** 2 - 2) < 1e-6)
        o.lattice_jump(lhs, int(rng.choice(nb)))
    for _ in range(50):  # far swaps as canonical MC makes them
        a, b = rng.integers(N, size=2)
        o.lattice_jump(int(a), int(b))
    return o


@pytest.mark.parametrize("gz", [False, True])
def test_read_matches_reference_route(tmp_path, gz):
    """ReadConfig + ReassignLatticeVector of a file with a shuffled atom order and nonzero
    map_shift: lattice ids, atom ids, elements and map_shift against the oracle; writing it
    back reproduces the file byte for byte."""
    n = 5
    o = _scrambled(n, seed=11)
    text = o.write_text()
    f = tmp_path / ("s.cfg.gz" if gz else "s.cfg")
    if gz:
        with gzip.open(f, "wt") as h:
            h.write(text)
    else:
        f.write_text(text)
    assert any(any(s) for s in o.map_shift)
    c = CF.read_config(f)
    r = CO.read_text(text)
    r.reassign()
    l2a, shift = c.atoms()
    assert np.array_equal(l2a, r.lattice_to_atom)
    assert np.array_equal(shift, np.array(r.map_shift))
    els = [r.elements[a] for a in r.lattice_to_atom]
    assert c.elements(SPECIES) == els
    out = tmp_path / ("o.cfg.gz" if gz else "o.cfg")
    c.write(out)
    back = gzip.open(out, "rt").read() if gz else out.read_text()
    assert back == text


def test_header_variants(tmp_path):
    """Extra header lines, no .NO_VELOCITY. line (map_shift then zero), auxiliary columns, and
    an unknown element string (read as X, Element.hpp:25-26)."""
    o = _scrambled(4, seed=3)
    o.elements[5] = "Qq"
    lines = o.write_text().split("\n")
    body = "\n".join(lines[16:])  # after the 11 header lines and the 5 of entry_count
    hdr = "\n".join(lines[:11]) + "\nentry_count = 3\n  a comment line\n\n"
    f = tmp_path / "v.cfg"
    f.write_text(hdr + body)
    c = CF.read_config(f)
    _, shift = c.atoms()
    assert not shift.any()
    r = CO.read_text(hdr + body)
    r.reassign()
    els = ["X" if r.elements[a] == "Qq" else r.elements[a] for a in r.lattice_to_atom]
    assert c.elements(SPECIES) == els
    # auxiliary columns after map_shift are skipped
    aux = o.write_text().replace("entry_count = 6", "entry_count = 7").split("\n")
    aux = [ln + " 0.5" if ln.count(" ") == 5 else ln for ln in aux]
    f2 = tmp_path / "a.cfg"
    f2.write_text("\n".join(aux))
    _, shift2 = CF.read_config(f2).atoms()
    assert np.array_equal(shift2, np.array(o.map_shift))  # atom ids are the file's line order


def test_map_files_round_trip(tmp_path):
    """WriteLattice / WriteElement / WriteMap against the oracle (neighbour shells by brute-force
    distances) and ReadMap back."""
    n = 4
    o = _scrambled(n, seed=8)
    f = tmp_path / "s.cfg"
    f.write_text(o.write_text())
    c = CF.read_config(f)
    c.write_map(tmp_path / "lattice.txt", tmp_path / "element.txt", tmp_path / "map.txt")
    r = CO.read_text(o.write_text())
    r.reassign()
    assert (tmp_path / "lattice.txt").read_text() == r.write_lattice_text()
    assert (tmp_path / "element.txt").read_text() == "".join(e + "\n" for e in r.elements)
    assert (tmp_path / "map.txt").read_text() == "".join(f"{s}\n" for s in r.atom_to_lattice)
    m = CF.read_map(tmp_path / "lattice.txt", tmp_path / "element.txt", tmp_path / "map.txt")
    assert np.array_equal(m.atoms()[0], c.atoms()[0])
    assert m.elements(SPECIES) == c.elements(SPECIES)


def test_errors(tmp_path):
    with pytest.raises(capi.LmcError, match="Cannot open"):
        CF.read_config(tmp_path / "missing.cfg.gz")
    c = CF.generate_fcc(4, "Al")
    c.write_map(tmp_path / "l.txt", tmp_path / "e.txt", tmp_path / "m.txt")
    lines = (tmp_path / "m.txt").read_text().split("\n")
    (tmp_path / "dup.txt").write_text("\n".join([lines[1]] + lines[1:]))
    with pytest.raises(capi.LmcError, match="Duplicate lattice id in map file: "):
        CF.read_map(tmp_path / "l.txt", tmp_path / "e.txt", tmp_path / "dup.txt")
    (tmp_path / "oob.txt").write_text("\n".join(["256"] + lines[1:]))
    with pytest.raises(capi.LmcError, match="Lattice id out of bounds in map file: 256"):
        CF.read_map(tmp_path / "l.txt", tmp_path / "e.txt", tmp_path / "oob.txt")
    text = CO.generate_fcc(4, "Al").write_text()
    (tmp_path / "few.cfg").write_text(text.replace("Number of particles = 256",
                                                   "Number of particles = 255"))
    with pytest.raises(capi.LmcError, match="not an n x n x n-cell FCC box"):
        CF.read_config(tmp_path / "few.cfg")
    (tmp_path / "off.cfg").write_text(text.replace("0.1250000000000000 0.1250000000000000 0.0000000000000000",
                                                   "0.1300000000000000 0.1250000000000000 0.0000000000000000", 1))
    with pytest.raises(capi.LmcError, match="off the FCC lattice"):
        CF.read_config(tmp_path / "off.cfg")
    with pytest.raises(capi.LmcError, match="not in the species list"):
        c.codes(["Mg", "X", "Zn"])


def test_extended_xyz_matches_reference_route(tmp_path):
    """WriteExtendedXyz of a scrambled configuration (plain and gzip) against the oracle."""
    o = _scrambled(4, seed=5)
    f = tmp_path / "s.cfg"
    f.write_text(o.write_text())
    c = CF.read_config(f)
    c.write_xyz(tmp_path / "s.xyz")
    c.write_xyz(tmp_path / "s.xyz.gz")
    r = CO.read_text(o.write_text())
    r.reassign()
    want = r.write_xyz_text()
    assert (tmp_path / "s.xyz").read_text() == want
    assert gzip.open(tmp_path / "s.xyz.gz", "rt").read() == want
