"""cmc_log.txt (CanonicalMcAbstract::Dump, lmc/mc/src/CanonicalMcAbstract.cpp:53-84) and the
simulated-annealing log (SimulatedAnnealing::Dump, SimulatedAnnealing.cpp:75-97) written from
the GPU records by lmc_cmc_write_log, against the lines the reference's Dump prints for the same
trials, replayed from the GPU trace by oracle/log_oracle.py: steps, energies and lowest energies
exactly, temperatures to 1e-12 relative, thermodynamic averages to 1e-9 eV."""
import pytest

import log_oracle as LO
from conftest import random_alloc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _check(got, want, avg_tol):
    """Line by line; avg_tol: tolerance of the fourth column (None: exact)."""
    assert len(got) == len(want)
    for g, w in zip(got, want):
        if w.startswith("steps"):
            assert g == w
            continue
        gs, ws = g.split("\t"), w.split("\t")
        assert len(gs) == 5 and gs[0] == ws[0]
        assert gs[2] == ws[2] and gs[4] == ws[4]  # energy and absolute energy
        assert float(gs[1]) == pytest.approx(float(ws[1]), rel=1e-12)
        if avg_tol is None:
            assert gs[3] == ws[3]
        else:
            assert float(gs[3]) == pytest.approx(float(ws[3]), abs=avg_tol)


@pytest.mark.parametrize("M,lds", [(0, True), (37, False), (2000, True)])
def test_cmc_log_matches_reference(torch, coeffs, tmp_path, M, lds):
    """Two writes of one run (constant T and a ramp), then a restarted run that appends without
    a header and without its first dump. M = 2000 runs 140 000 trials: launches of 65 536 trials
    and a ring of 67 536 energies that wraps."""
    import latticemontecarlo_amd as L
    path, els = coeffs["e0_3n"]
    gp = L.Model(path, els, L.PAIRSITE)
    sp = gp.species
    n, R = 4, 2
    occs = [random_alloc(n, sp, 0.1, seed=80 + r, vacancy=False) for r in range(R)]
    k, Ld = (70000, 1000) if M == 2000 else (3000, 50)
    T0, T1 = [900.0, 700.0], [900.0, 300.0]
    maxs = 2 * k + 3 * Ld + 400  # room for the restarted run
    print(f"cmc log M={M} lds={lds} trials={2 * k}", flush=True)
    drv = L.CanonicalMc(gp, None, [n] * R, occs, T0, T1, 5, maximum_steps=maxs)
    drv.set_lds_occupancy(lds)
    drv.set_log(Ld, capacity=5000, thermodynamic_averaging_steps=M)
    drv.set_trace(2 * k)
    files = [tmp_path / f"cmc_log_{r}.txt" for r in range(R)]
    drv.run(k)
    drv.write_log(files)
    drv.run(k)
    drv.write_log(files)
    tr = drv.trace()
    st = drv.state()
    temps = [lambda s, r=r: T0[r] + (T1[r] - T0[r]) * (s / maxs) for r in range(R)]
    for r in range(R):
        want = LO.cmc_log_lines(0, 0.0, 0.0, tr["de"][r], tr["acc"][r], temps[r], Ld, M)
        got = files[r].read_text().split("\n")
        assert got[-1] == "" and len(want) > 200
        _check(got[:-1], want, 1e-9)
        if M == 0:
            assert all(ln.split("\t")[3] == "0" for ln in got[1:-1])
    # restart from the final state: appended lines, no header, no first dump
    steps0 = [int(x) for x in st["steps"]]
    rest = L.CanonicalMc(gp, None, [n] * R, drv.occupancy(), T0, T1, 5, maximum_steps=maxs,
                         restart_steps=steps0, restart_energy=list(st["energy"]))
    rest.set_lds_occupancy(lds)
    kr = 2 * Ld + 100  # past two more due steps
    rest.set_log(Ld, capacity=100, thermodynamic_averaging_steps=M)
    rest.set_trace(kr)
    rest.run(kr)
    before = [f.read_text() for f in files]
    rest.write_log(files)
    rtr = rest.trace()
    for r in range(R):
        added = files[r].read_text()[len(before[r]):].split("\n")[:-1]
        want = LO.cmc_log_lines(steps0[r], float(st["energy"][r]), 0.0, rtr["de"][r],
                                rtr["acc"][r], temps[r], Ld, M, restarted=True)
        assert len(want) >= 2 and not want[0].startswith("steps")
        _check(added, want, 1e-9)


def test_annealing_log_matches_reference(torch, coeffs, tmp_path):
    """Rows every log_dump_steps steps and at every new lowest energy, over two writes, with
    the schedule's temperature."""
    import latticemontecarlo_amd as L
    path, els = coeffs["e0_3n"]
    gp = L.Model(path, els, L.PAIRSITE)
    sp = gp.species
    n, R, N, T0, Ld = 5, 2, 20000, 2000.0, 500
    occs = [random_alloc(n, sp, 0.1, seed=90 + r, vacancy=False) for r in range(R)]
    print(f"annealing log R={R} trials={N + 1}", flush=True)
    sa = L.SimulatedAnnealing(gp, None, [n] * R, occs, T0, 9, maximum_steps=N)
    sa.set_log(Ld, capacity=20000)
    sa.set_trace(N + 1)
    files = [tmp_path / f"sa_log_{r}.txt" for r in range(R)]
    sa.run(N // 2)
    sa.write_log(files)
    sa.run(N // 2 + 1)
    sa.write_log(files)
    tr = sa.trace()
    an = sa.annealing_state()
    for r in range(R):
        want, s = LO.sa_log_lines(0, 0.0, 0.0, tr["de"][r], tr["acc"][r], T0, N, Ld)
        got = files[r].read_text().split("\n")
        assert got[-1] == ""
        assert len(want) > N // Ld + 10  # improvement rows besides the periodic ones
        _check(got[:-1], want, None)
        assert s.T == pytest.approx(float(an["temperature"][r]), rel=1e-12)


def test_cmc_log_argument_errors(torch, coeffs, tmp_path):
    """lmc_cmc_set_log / lmc_cmc_write_log reject a zero capacity, intervals and windows past
    their limits and a write with the log off; records past the capacity are reported after the
    ones that fit are written."""
    import latticemontecarlo_amd as L
    from latticemontecarlo_amd import capi
    path, els = coeffs["e0_3n"]
    gp = L.Model(path, els, L.PAIRSITE)
    occ = random_alloc(4, gp.species, 0.1, seed=5, vacancy=False)
    drv = L.CanonicalMc(gp, None, [4], [occ], 800.0, 800.0, 1, maximum_steps=1000)
    f = tmp_path / "cmc_log.txt"
    with pytest.raises(capi.LmcError, match="logging is off"):
        drv.write_log([f])
    with pytest.raises(capi.LmcError, match="log capacity must be positive"):
        drv.set_log(10, capacity=0)
    with pytest.raises(capi.LmcError, match="log_dump_steps is too large"):
        drv.set_log(10 ** 18)
    with pytest.raises(capi.LmcError, match="thermodynamic_averaging_steps is too large"):
        drv.set_log(10, thermodynamic_averaging_steps=2 ** 31 + 1)
    drv.set_log(10, capacity=5)
    drv.run(200)  # due steps 0..98 and every 10 after: far more than 5
    with pytest.raises(capi.LmcError, match="log records past the capacity were lost"):
        drv.write_log([f])
    lines = f.read_text().split("\n")
    assert lines[0].startswith("steps\t") and [ln.split("\t")[0] for ln in lines[1:6]] == \
        ["0", "1", "2", "3", "4"]
