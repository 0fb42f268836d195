"""The log restatements of oracle/log_oracle.py, pinned on the CPU: ThermodynamicAveraging by
closed forms, the floating-point dump cadence by its integer form and by the step-by-step form
the canonical-MC kernel evaluates, and the annealing replay by the C++ oracle's
SimulatedAnnealing restatement (which the GPU annealing tests pin in turn)."""
import math

import numpy as np
import pytest

import log_oracle as LO
import oracle as O
from conftest import random_alloc


def test_thermodynamic_average_closed_forms():
    a = LO.ThermodynamicAveraging(0)  # size 0: no averaging, the column reads 0
    a.add(1.0)
    assert a.average(10.0) == 0.0
    a = LO.ThermodynamicAveraging(5)
    for _ in range(3):
        a.add(-2.5)
    assert a.average(40.0) == pytest.approx(-2.5, abs=1e-15)
    # two levels 0 (x30) and d (x20): the Boltzmann average b d e^{-d beta} / (a + b e^{-d beta})
    beta, d = 38.7, 0.013
    a = LO.ThermodynamicAveraging(100)
    for v in [0.0] * 30 + [d] * 20:
        a.add(v)
    x = math.exp(-d * beta)
    assert a.average(beta) == pytest.approx(20 * d * x / (30 + 20 * x), rel=1e-12)
    # beta -> 0: the plain mean
    vals = np.random.default_rng(2).normal(-3.0, 0.05, size=40)
    a = LO.ThermodynamicAveraging(40)
    for v in vals:
        a.add(float(v))
    assert a.average(1e-12) == pytest.approx(vals.mean(), abs=1e-13)
    # the window: only the last `size` energies count
    a, b = LO.ThermodynamicAveraging(16), LO.ThermodynamicAveraging(16)
    for v in vals:
        a.add(float(v))
    for v in vals[-16:]:
        b.add(float(v))
    assert len(a.energies) == 16
    assert a.average(30.0) == pytest.approx(b.average(30.0), abs=1e-12)


def _integer_every(s, L):
    """The cadence in integers, as the kernels evaluate it."""
    if s > 10 * L:
        return L
    p, t = 1, s + 1
    while t >= 100:
        t //= 10
        p *= 10
    return min(p, L)


class _Clock:
    """The canonical-MC kernel's step-by-step cadence (one division per change of cadence)."""

    def __init__(self, L, plain, s):
        self.L, self.plain = L, plain
        self.reset(s)

    def reset(self, s):
        self.every, self.change = self.L, 2 ** 64 - 1
        if not self.plain and s <= 10 * self.L:
            p, t = 1, s + 1
            while t >= 100:
                t //= 10
                p *= 10
            self.every = min(p, self.L)
            self.change = min(100 * p - 1, 10 * self.L + 1)
        self.rem = s % self.every

    def due(self, s):
        if s >= self.change:
            self.reset(s)
        d = self.rem == 0
        self.rem += 1
        if self.rem == self.every:
            self.rem = 0
        return d


@pytest.mark.parametrize("L", [1, 3, 7, 9, 10, 25, 100, 1000])
def test_cadence_integer_and_stepwise_forms(L):
    S = 120000
    ref = [LO.reference_due(s, L) for s in range(S)]
    assert ref == [s % _integer_every(s, L) == 0 for s in range(S)]
    for start in sorted({0, 7, 98, 99, 100, 995, 999, 10 * L, 10 * L + 1, 12345}):
        clk = _Clock(L, False, start)
        assert [clk.due(s) for s in range(start, start + 4000)] == ref[start:start + 4000]
    clk = _Clock(L, True, 5)
    assert [clk.due(s) for s in range(5, 3000)] == [s % L == 0 for s in range(5, 3000)]


@pytest.mark.parametrize("N,k,T0,seed", [(10000, 3000, 5000.0, 4), (4000, 4001, 40.0, 6)])
def test_annealing_replay_matches_oracle_schedule(coeffs, N, k, T0, seed):
    """The replay of the temperature schedule from the acceptance trace ends in the C++
    oracle's temperature, lowest energy and reheat count."""
    path, els = coeffs["e0_3"]
    n = 5
    lat = O.Lattice(n)
    pm = O.Model(lat, path, els, O.PAIRSITE)
    species = sorted(set(els) | {"X"})
    occ = random_alloc(n, species, 0.1, seed=seed, vacancy=False)
    c = O.Cmc(pm, None, lat, occ, T0, T0, seed=seed, anneal=True)
    tr = c.run(k, maximum_steps=N, trace=True)
    st = c.anneal_state()
    lines, s = LO.sa_log_lines(0, 0.0, 0.0, tr["de"], tr["acc"], T0, N, 100)
    assert s.T == pytest.approx(st["temperature"], rel=1e-12)
    assert s.lowest == st["lowest"] and s.reheats == st["reheats"]
    assert s.reheats == (3 if T0 == 40.0 else 0)  # the low-temperature case takes the reheat path
    assert lines[0].startswith("steps\t") and len(lines) > k // 100
