"""Pins of the oracle's CanonicalMcOmp restatement (no GPU), lmc/mc/src/CanonicalMcOmp.cpp:40-92.

* K = 1 is CanonicalMcSerial: one event per round drawn from the step's own attempts, so the
  trajectory equals the serial one bit for bit (the sequential-equivalence special case).
* The round structure against an independent restatement of BuildEventVector from the Philox
  stream, the lattice's neighbour shells and the `ct == 50` rule, in a box small enough that
  rounds end by exhausting their draws.
* The exclusion set makes the precomputed dE exact: every event's dE equals the swap dE on the
  occupancy at the moment it is tried (after the round's earlier swaps).
"""
import numpy as np

import oracle as O
from conftest import TERNARY, random_alloc


def _models(coeffs, n):
    path, els = coeffs["e0_3n"]
    lat = O.Lattice(n)
    return lat, O.Model(lat, path, els, O.PAIRSITE)


def test_one_event_per_round_is_the_serial_trajectory(coeffs):
    n, trials = 5, 3000
    lat, op = _models(coeffs, n)
    species = sorted(TERNARY + ["X"])
    occ = random_alloc(n, species, 0.07, seed=3, vacancy=False)
    a = O.Cmc(op, None, lat, occ, 900.0, 400.0, 17, replica=2)
    b = O.Cmc(op, None, lat, occ, 900.0, 400.0, 17, replica=2)
    ta = a.run(trials, trials, trace=True)
    tb = b.run_omp(trials, trials, 1, trace=True)
    for key in ("p", "q", "acc"):
        assert np.array_equal(ta[key], tb[key]), key
    assert np.array_equal(ta["de"], tb["de"])
    assert np.array_equal(a.occupancy(), b.occupancy())
    assert a.state() == b.state() or (a.state()["energy"] == b.state()["energy"])
    assert ta["acc"].sum() > 100


def _bits64(r0, r1):
    return int(r0) | (int(r1) << 32)


def _round_restated(lat, occ, seed, replica, s0, K, N):
    """BuildEventVector (CanonicalMcOmp.cpp:40-78) restated from the Philox stream."""
    unavailable = set()
    att = 0
    events = []
    for _ in range(K):
        ct = 0
        while True:
            while True:  # GenerateLatticeIdJumpPair
                r = O.philox([seed & 0xFFFFFFFF, seed >> 32],
                             [s0 & 0xFFFFFFFF, s0 >> 32, replica, 2 * att + 1])
                att += 1
                p = (_bits64(r[0], r[1]) * N) >> 64
                q = (_bits64(r[2], r[3]) * N) >> 64
                if occ[p] != occ[q]:
                    break
            ct += 1
            if ct == 50:
                break
            if p not in unavailable and q not in unavailable:
                break
        if ct == 50:
            break
        for site in (p, q):
            unavailable.add(site)
            for sh in (1, 2, 3):
                unavailable.update(int(x) for x in lat.neighbors(site, sh))
        events.append((p, q))
    return events


def test_rounds_match_a_restatement_of_build_event_vector(coeffs):
    """n = 4 (256 sites): with K = 64 the exclusion set fills the box, so rounds end through the
    50-draw limit; with K = 3 they fill."""
    n, seed, replica = 4, 123, 1
    lat, op = _models(coeffs, n)
    N = 4 * n ** 3
    species = sorted(TERNARY + ["X"])
    occ0 = random_alloc(n, species, 0.2, seed=9, vacancy=False)
    for K in (3, 64):
        c = O.Cmc(op, None, lat, occ0, 700.0, 700.0, seed, replica=replica)
        occ = occ0.copy()
        s = 0
        lengths = []
        for _ in range(6):  # six rounds, each tried in full
            want = _round_restated(lat, occ, seed, replica, s, K, N)
            tr = c.run_omp(len(want), 10 ** 9, K, trace=True)
            assert c.round_left() == 0
            got = list(zip(tr["p"].tolist(), tr["q"].tolist()))
            assert got == want, (K, s)
            lengths.append(len(want))
            occ = c.occupancy()
            s += len(want)
        if K == 3:
            assert lengths == [3] * 6
        else:
            assert max(lengths) < 64 and min(lengths) >= 1


def test_round_de_is_exact_after_earlier_swaps(coeffs):
    n, K = 8, 24
    lat, op = _models(coeffs, n)
    species = sorted(TERNARY + ["X"])
    occ = random_alloc(n, species, 0.08, seed=4, vacancy=False)
    c = O.Cmc(op, None, lat, occ, 2000.0, 2000.0, 5)
    tried = 0
    for _ in range(200):
        before = c.occupancy()
        tr = c.run_omp(1, 10 ** 9, K, trace=True)
        de_now = op.swap_de(before, int(tr["p"][0]), int(tr["q"][0]))
        assert abs(de_now - tr["de"][0]) < 1e-12
        tried += 1
    assert tried == 200
