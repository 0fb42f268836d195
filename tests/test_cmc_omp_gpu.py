"""CanonicalMcOmp on the GPU (lmc_cmc_set_parallel_trials) against the oracle's restatement of
lmc/mc/src/CanonicalMcOmp.cpp:40-92: pairs, dE (1e-9 eV), acceptances and occupancy bit-exact
over runs that cut rounds between launches, with the occupancy in LDS and in HBM, replicas of
several sizes (the smallest so small that rounds end through the 50-draw limit); one trial per
round is the serial trajectory; the 30^3-cell, Al/Mg/Zn 90/5/5 workload under the 800 -> 300 K
ramp; and cmc_log.txt replayed from the trace."""
import numpy as np
import pytest

import log_oracle as LO
from conftest import random_alloc

pytestmark = pytest.mark.gpu
TOL_E = 1e-9


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _run_pair(coeffs, sizes, K, chunks, lds, fracs=None, t0=800.0, t1=300.0, seed=42,
              maxs=None):
    import latticemontecarlo_amd as L
    import oracle as O
    path, els = coeffs["e0_3n"]
    gp = L.Model(path, els, L.PAIRSITE)
    ge = L.Model(path, els, L.ENERGY)
    R = len(sizes)
    fracs = fracs or [0.05 + 0.01 * r for r in range(R)]
    occs = [random_alloc(n, gp.species, fracs[r], seed=700 + r, vacancy=False)
            for r, n in enumerate(sizes)]
    total = sum(chunks)
    maxs = maxs or total
    print(f"cmc omp sizes={sizes} K={K} chunks={chunks} lds={lds}", flush=True)
    drv = L.CanonicalMc(gp, ge, sizes, occs, t0, t1, seed, maximum_steps=maxs)
    drv.set_lds_occupancy(lds)
    drv.set_parallel_trials(K)
    drv.set_trace(total)
    for k in chunks:
        drv.run(k)
    gtr, gst, gocc = drv.trace(), drv.state(), drv.occupancy()
    left = drv.round_left()
    gp.check_errors()
    refs = []
    for r, n in enumerate(sizes):
        lat = O.Lattice(n)
        op, oe = O.Model(lat, path, els, O.PAIRSITE), O.Model(lat, path, els, O.ENERGY)
        c = O.Cmc(op, oe, lat, occs[r], t0, t1, seed, replica=r)
        tr = c.run_omp(total, maxs, K, trace=True)
        refs.append((c, tr))
    return gtr, gst, gocc, left, refs


def _compare(gtr, gst, gocc, left, refs, total):
    for r, (c, tr) in enumerate(refs):
        st = c.state()
        ran = int(gst["steps"][r])
        assert ran == st["steps"], f"replica {r}"
        k = min(ran, total)
        assert np.array_equal(gtr["p"][r][:k], tr["p"][:k].astype(np.int64)), f"replica {r}"
        assert np.array_equal(gtr["q"][r][:k], tr["q"][:k].astype(np.int64))
        np.testing.assert_allclose(gtr["de"][r][:k], tr["de"][:k], rtol=0, atol=TOL_E)
        assert np.array_equal(gtr["acc"][r][:k], tr["acc"][:k])
        assert np.array_equal(gocc[r], c.occupancy())
        assert int(gst["accepted"][r]) == st["accepted"]
        assert abs(gst["energy"][r] - st["energy"]) <= TOL_E * total
        assert abs(gst["abs_energy"][r] - st["abs_energy"]) <= TOL_E * (total + abs(st["abs_energy"]))
        assert int(left[r]) == c.round_left()


@pytest.mark.parametrize("K,lds", [(64, True), (64, False), (16, True), (5, False)])
def test_cmc_omp_matches_oracle(torch, coeffs, K, lds):
    sizes = [6, 4, 9]
    chunks = [100, 333, 7, 1560]  # cuts rounds between launches
    out = _run_pair(coeffs, sizes, K, chunks, lds)
    _compare(*out, sum(chunks))


def test_cmc_omp_one_trial_per_round_is_serial(torch, coeffs):
    import latticemontecarlo_amd as L
    path, els = coeffs["e0_3n"]
    gp = L.Model(path, els, L.PAIRSITE)
    sizes, trials = [5, 7], 1500
    occs = [random_alloc(n, gp.species, 0.08, seed=60 + r, vacancy=False) for r, n in enumerate(sizes)]
    a = L.CanonicalMc(gp, None, sizes, occs, 900.0, 400.0, 9, maximum_steps=trials)
    b = L.CanonicalMc(gp, None, sizes, occs, 900.0, 400.0, 9, maximum_steps=trials)
    b.set_parallel_trials(1)
    for d in (a, b):
        d.set_trace(trials)
        d.run(700)
        d.run(trials - 700)
    ta, tb = a.trace(), b.trace()
    for key in ("p", "q", "acc"):
        assert np.array_equal(ta[key], tb[key]), key
    np.testing.assert_allclose(ta["de"], tb["de"], rtol=0, atol=TOL_E)
    for x, y in zip(a.occupancy(), b.occupancy()):
        assert np.array_equal(x, y)


def test_cmc_omp_stops_after_the_last_round(torch, coeffs):
    """maximum_steps = 1000: the last round started at a step <= 1000 runs to its end."""
    sizes = [6, 8]
    out = _run_pair(coeffs, sizes, 64, [5000], True, maxs=1000)
    gtr, gst, gocc, left, refs = out
    for r in range(2):
        assert 1001 <= int(gst["steps"][r]) < 1001 + 64 and int(left[r]) == 0
    _compare(*out, 5000)


def test_cmc_omp_30_cells_ramp(torch, coeffs):
    """The canonical workload's replicas: 30^3 cells (108000 sites), Al/Mg/Zn 90/5/5, T linear
    800 -> 300 K over the run, 64 trials per round."""
    out = _run_pair(coeffs, [30, 30], 64, [2000, 2000], True, fracs=[0.05, 0.05], maxs=4000)
    _compare(*out, 4000)


def test_cmc_omp_log_matches_reference(torch, coeffs, tmp_path):
    import latticemontecarlo_amd as L
    path, els = coeffs["e0_3n"]
    gp = L.Model(path, els, L.PAIRSITE)
    n, k, Ld, M = 5, 3000, 50, 37
    occs = [random_alloc(n, gp.species, 0.1, seed=81, vacancy=False)]
    maxs = 2 * k
    drv = L.CanonicalMc(gp, None, [n], occs, 900.0, 300.0, 5, maximum_steps=maxs)
    drv.set_parallel_trials(32)
    drv.set_log(Ld, capacity=5000, thermodynamic_averaging_steps=M)
    drv.set_trace(2 * k)
    f = tmp_path / "cmc_log.txt"
    drv.run(k)
    drv.write_log([f])
    drv.run(k)
    drv.write_log([f])
    tr = drv.trace()
    want = LO.cmc_log_lines(0, 0.0, 0.0, tr["de"][0], tr["acc"][0],
                            lambda s: 900.0 + (300.0 - 900.0) * (s / maxs), Ld, M)
    got = f.read_text().split("\n")
    assert got[-1] == "" and len(got) - 1 == len(want)
    for g, w in zip(got[:-1], want):
        if w.startswith("steps"):
            assert g == w
            continue
        gs, ws = g.split("\t"), w.split("\t")
        assert gs[0] == ws[0] and gs[2] == ws[2] and gs[4] == ws[4]
        assert float(gs[1]) == pytest.approx(float(ws[1]), rel=1e-12)
        assert float(gs[3]) == pytest.approx(float(ws[3]), abs=1e-9)
