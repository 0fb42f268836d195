"""Atom ids and map_shift tracked on the GPU (Config::LatticeJump, lmc/cfg/src/Config.cpp:431-456)
against the oracle's floating-point LatticeJump replayed over the same jumps, and a resume from
a written .cfg.gz that continues the trajectory bit for bit."""
import numpy as np
import pytest

import cfg_oracle as CO
import log_oracle as LO
from conftest import random_alloc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _start_atoms(n, R, seed):
    rng = np.random.default_rng(seed)
    N = 4 * n ** 3
    l2a = [rng.permutation(N).astype(np.int64) for _ in range(R)]
    sh = [rng.integers(-3, 4, size=(N, 3)).astype(np.int32) for _ in range(R)]
    return l2a, sh


def _replay(n, l2a, shift, pairs):
    """Oracle configuration with x-major lattice ids after the LatticeJumps `pairs`."""
    o = CO.generate_fcc(n, "Al")
    o.reassign()
    o.atom_to_lattice = [0] * len(l2a)
    for s, a in enumerate(l2a):
        o.atom_to_lattice[int(a)] = s
    o.map_shift = [list(map(int, r)) for r in shift]
    for lhs, rhs in pairs:
        o.lattice_jump(int(lhs), int(rhs))
    return np.array(o.lattice_to_atom), np.array(o.map_shift)


@pytest.mark.parametrize("form", [0, 1, 2, 3])
def test_kmc_atom_tracking_matches_lattice_jump(torch, coeffs, form):
    import latticemontecarlo_amd as L
    path, els = coeffs["e0_3"]
    n, R, steps = 4, 3, 1200
    gb = L.Model(path, els, L.E0)
    occs = [random_alloc(n, gb.species, 0.06, seed=40 + r) for r in range(R)]
    l2a, sh = _start_atoms(n, R, seed=form)
    print(f"kmc atom tracking form={form} n={n} R={R} steps={steps}", flush=True)
    runs = []
    for track in (True, False):
        drv = L.KineticMcFirst(gb, None, [n] * R, occs, 700.0, 17, maximum_steps=10 ** 9)
        drv.set_formulation(form)
        drv.set_trace(steps)
        if track:
            drv.set_atoms(l2a, sh)
        drv.run(steps // 2)
        drv.run(steps - steps // 2)
        runs.append((drv.trace(), drv.occupancy(), drv.atoms() if track else None))
    (tr, occ, (ga, gs)), (tr0, occ0, _) = runs
    assert np.array_equal(tr["sel"], tr0["sel"])  # tracking leaves the trajectory alone
    crossed = 0
    for r in range(R):
        vac = int(np.flatnonzero(occs[r] == gb.species.index("X"))[0])
        pairs = []
        for s in tr["sel"][r]:
            pairs.append((vac, int(s)))
            vac = int(s)
        ra, rs = _replay(n, l2a[r], sh[r], pairs)
        assert np.array_equal(ga[r], ra)
        assert np.array_equal(gs[r], rs)
        crossed += int((rs != sh[r]).any(axis=1).sum())
        # every atom carries its species: the occupancy is the start species of its atom
        start_species = np.empty(4 * n ** 3, np.uint8)
        start_species[l2a[r]] = occs[r]
        assert np.array_equal(occ[r], start_species[ga[r]])
    assert crossed > 0


@pytest.mark.parametrize("lds", [True, False])
def test_cmc_atom_tracking_matches_lattice_jump(torch, coeffs, lds):
    import latticemontecarlo_amd as L
    path, els = coeffs["e0_3n"]
    n, R, trials = 4, 2, 800
    gp = L.Model(path, els, L.PAIRSITE)
    occs = [random_alloc(n, gp.species, 0.1, seed=70 + r, vacancy=False) for r in range(R)]
    l2a, sh = _start_atoms(n, R, seed=9)
    print(f"cmc atom tracking lds={lds}", flush=True)
    drv = L.CanonicalMc(gp, None, [n] * R, occs, 1500.0, 1500.0, 5, maximum_steps=trials)
    drv.set_lds_occupancy(lds)
    drv.set_atoms(l2a, sh)
    drv.set_trace(trials)
    drv.run(trials)
    tr, (ga, gs), occ = drv.trace(), drv.atoms(), drv.occupancy()
    for r in range(R):
        acc = tr["acc"][r].astype(bool)
        assert acc.sum() > 20
        ra, rs = _replay(n, l2a[r], sh[r], zip(tr["p"][r][acc], tr["q"][r][acc]))
        assert np.array_equal(ga[r], ra)
        assert np.array_equal(gs[r], rs)
        start_species = np.empty(4 * n ** 3, np.uint8)
        start_species[l2a[r]] = occs[r]
        assert np.array_equal(occ[r], start_species[ga[r]])


def test_kmc_resume_from_cfg_is_bit_exact(torch, coeffs, tmp_path):
    """2k steps in one run against k steps, a .cfg.gz written and read back, and k more steps
    from restart_steps / restart_time / restart_energy: the same jumps, clock, energy,
    occupancy, atom ids and map_shift."""
    import latticemontecarlo_amd as L
    from latticemontecarlo_amd import cfg as CF
    path, els = coeffs["e0_3"]
    n, k = 5, 600
    gb = L.Model(path, els, L.E0)
    ge = L.Model(path, els, L.ENERGY)
    sp = gb.species
    c0 = CF.generate_fcc(n, "Al")
    occ0 = random_alloc(n, sp, 0.05, seed=3)
    c0.update(sp, occ0)
    l2a0, sh0 = c0.atoms()
    print(f"resume n={n} k={k}", flush=True)

    def driver(occ, l2a, sh, **restart):
        d = L.KineticMcFirst(gb, ge, [n], [occ], 650.0, 99, maximum_steps=10 ** 9, **restart)
        d.set_atoms([l2a], [sh])
        d.set_trace(k)
        return d

    full = driver(occ0, l2a0, sh0)
    full.run(k)
    full_tr1 = full.trace()
    full.set_trace(k)
    full.run(k)
    full_tr2, full_st = full.trace(), full.state()

    half = driver(occ0, l2a0, sh0)
    half.run(k)
    st = half.state()
    assert np.array_equal(half.trace()["sel"], full_tr1["sel"])
    (a,), (s,) = half.atoms()
    c0.update(sp, half.occupancy()[0], a, s)
    c0.write(tmp_path / f"{k}.cfg.gz")
    c1 = CF.read_config(tmp_path / f"{k}.cfg.gz")
    l2a1, sh1 = c1.atoms()
    assert np.array_equal(l2a1, a) and np.array_equal(sh1, s)
    rest = driver(c1.codes(sp), l2a1, sh1, restart_steps=int(st["steps"][0]),
                  restart_time=float(st["time"][0]), restart_energy=float(st["energy"][0]))
    rest.run(k)
    rtr, rst = rest.trace(), rest.state()
    assert np.array_equal(rtr["sel"], full_tr2["sel"])
    assert np.array_equal(rtr["time"], full_tr2["time"])  # bit for bit
    assert rst["steps"][0] == full_st["steps"][0]
    assert rst["time"][0] == full_st["time"][0] and rst["energy"][0] == full_st["energy"][0]
    assert np.array_equal(rest.occupancy()[0], full.occupancy()[0])
    (ra,), (rs,) = rest.atoms()
    (fa,), (fs,) = full.atoms()
    assert np.array_equal(ra, fa) and np.array_equal(rs, fs)


def _expected_log(n, occ, l2a, sh, tr, steps0, e0, t0, T, L, sp, first_fixed, skip_first):
    """kmc_log.txt lines of a trajectory from its trace (without the COM columns)."""
    o = CO.generate_fcc(n, "Al")
    o.reassign()
    o.atom_to_lattice = [0] * len(l2a)
    for s, a in enumerate(l2a):
        o.atom_to_lattice[int(a)] = s
    o.map_shift = [list(map(int, r)) for r in sh]
    vac = int(np.flatnonzero(occ == sp.index("X"))[0])
    edge, m = 4.046 * n, float(2 * n)
    fixed = first_fixed
    energy, time, lines = e0, t0, []
    for i, (s, ea, de, tnew) in enumerate(zip(tr["sel"], tr["ea"], tr["de"], tr["time"])):
        step = steps0 + i
        if LO.reference_due(step, L) and not (skip_first and i == 0):
            l2a_now = o.lattice_to_atom
            X = vac // (2 * n * n)
            Y = (vac - X * 2 * n * n) // n
            Z = 2 * (vac - X * 2 * n * n - Y * n) + ((X + Y) & 1)
            sv = o.map_shift[l2a_now[vac]]
            f = "%.16f" if fixed else "%.16g"
            vals = "\t".join(f % v for v in (time, T, energy, ea, de))
            pos = " ".join("%.16f" % ((c / m + sv[d]) * edge) for d, c in enumerate((X, Y, Z)))
            lines.append(f"{step}\t{vals}\t{l2a_now[int(s)]}\t{pos}")
            fixed = True
        o.lattice_jump(vac, int(s))
        vac = int(s)
        energy += de
        time = tnew
    return lines


def test_kmc_log_matches_reference_format(torch, coeffs, tmp_path):
    """kmc_log.txt written from the GPU records against the lines rebuilt from the trace with the
    reference's floating-point cadence and stream formats, over two writes; a restarted run
    appends without a header and without its first dump. The solute COM columns are compared
    to 1e-9 A."""
    import latticemontecarlo_amd as L
    path, els = coeffs["e0_3"]
    n, R, k, Ld, T = 4, 2, 700, 20, 700.0
    gb = L.Model(path, els, L.E0)
    sp = gb.species
    occs = [random_alloc(n, sp, 0.08, seed=60 + r) for r in range(R)]
    l2a, sh = _start_atoms(n, R, seed=21)
    print(f"kmc log n={n} R={R} steps={2 * k}", flush=True)
    drv = L.KineticMcFirst(gb, None, [n] * R, occs, T, 8, maximum_steps=10 ** 9,
                           solute_disp=[False, True])
    drv.set_atoms(l2a, sh)
    drv.set_log(Ld, capacity=400)
    drv.set_trace(2 * k)
    files = [tmp_path / f"kmc_log_{r}.txt" for r in range(R)]
    drv.run(k)
    drv.write_log(files)
    drv.run(k)
    drv.write_log(files)
    tr = drv.trace()
    st = drv.state()
    (ga, gs) = drv.atoms()
    for r in range(R):
        got = files[r].read_text().split("\n")
        hdr = "steps\ttime\ttemperature\tenergy\tEa\tdE\tselected\tvac1\tvac2\tvac3"
        assert got[0] == hdr + ("\tsolute_com1\tsolute_com2\tsolute_com3" if r == 1 else "")
        assert got[-1] == ""
        want = _expected_log(n, occs[r], l2a[r], sh[r], {key: v[r] for key, v in tr.items()},
                             0, 0.0, 0.0, T, Ld, sp, False, False)
        body = got[1:-1]
        assert len(body) == len(want) and len(want) > 100
        if r == 1:
            com = np.array([[float(x) for x in ln.split("\t")[8].split()] for ln in body])
            assert np.abs(com).max() > 0
            body = ["\t".join(ln.split("\t")[:8]) for ln in body]
        assert body == want
    # restart from the final state: appended lines, no header, no first dump
    occ_now = drv.occupancy()
    rest = L.KineticMcFirst(gb, None, [n] * R, occ_now, T, 8, maximum_steps=10 ** 9,
                            restart_steps=[int(x) for x in st["steps"]],
                            restart_time=list(st["time"]), restart_energy=list(st["energy"]))
    rest.set_atoms(ga, gs)
    rest.set_log(Ld, capacity=100)
    rest.set_trace(60)
    rest.run(60)
    before = [f.read_text() for f in files]
    rest.write_log(files)
    rtr = rest.trace()
    for r in range(R):
        added = files[r].read_text()[len(before[r]):].split("\n")[:-1]
        want = _expected_log(n, occ_now[r], ga[r], gs[r], {key: v[r] for key, v in rtr.items()},
                             int(st["steps"][r]), float(st["energy"][r]), float(st["time"][r]), T,
                             Ld, sp, False, True)
        if r == 1:
            added = ["\t".join(ln.split("\t")[:8]) for ln in added]
        assert added == want and len(want) == 2  # steps 1420 and 1440; 1400 is skipped
