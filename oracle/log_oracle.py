"""Test infrastructure, not part of the engine: restatements of the reference's driver logs, used
as the checker of lmc_kmc_write_log's cadence and of lmc_cmc_write_log.

- reference_due: the dump cadence of KineticMcFirstAbstract::Dump (KineticMcAbstract.cpp:83-98)
  and CanonicalMcAbstract::Dump (CanonicalMcAbstract.cpp:69-81), with the floating-point log10;
- ThermodynamicAveraging (ThermodynamicAveraging.cpp:6-40): a deque of the last `size` energies
  with the running sum, and the Boltzmann-weighted average about the mean;
- cmc_log_lines: CanonicalMcSerial::Simulate (CanonicalMcSerial.cpp:40-51) with Dump
  (CanonicalMcAbstract.cpp:53-84), replayed from a trace of (dE, accepted) per trial;
- sa_log_lines: SimulatedAnnealing::Simulate / Dump / UpdateTemperature
  (SimulatedAnnealing.cpp:75-185), replayed from the same kind of trace.
Numbers are printed as an ostream with precision(16) in the default float format ('%.16g').
"""
import collections
import math

K_BOLTZMANN = 8.617333262145e-5  # eV/K (Constants.hpp)


def reference_due(steps, L):
    """steps % every == 0 with every = pow(10, (ull)(log10(steps + 1) - 1)) clamped to [1, L] up
    to 10 L, and L past it; a negative exponent casts to 0 and is raised to 1."""
    if steps > 10 * L:
        every = L
    else:
        q = int(math.log10(steps + 1) - 1)  # truncation toward zero, as the cast
        every = int(10.0 ** q) if q >= 0 else 0
        every = min(max(every, 1), L)
    return steps % every == 0


class ThermodynamicAveraging:
    """ThermodynamicAveraging(size): AddEnergy / GetThermodynamicAverage(beta)."""

    def __init__(self, size):
        self.size = size
        self.energies = collections.deque()
        self.sum = 0.0

    def add(self, value):
        if self.size == 0:
            return
        if len(self.energies) == self.size:
            self.sum -= self.energies.popleft()
        self.energies.append(value)
        self.sum += value

    def average(self, beta):
        if self.size == 0:
            return 0.0
        mean = self.sum / len(self.energies) if self.energies else 0.0
        partition = 0.0
        weighted = 0.0
        for e in self.energies:
            e -= mean
            x = math.exp(-e * beta)
            weighted += e * x
            partition += x
        return weighted / partition + mean


def _fmt(step, *vals):
    return "\t".join([str(step)] + ["%.16g" % v for v in vals])


def cmc_log_lines(steps0, energy0, abs0, de, acc, temperature, L, averaging_steps,
                  restarted=False):
    """cmc_log.txt lines (header included for a run from step 0) of the trials steps0, steps0 + 1,
    ... with energy changes de[i] and acceptances acc[i]; temperature(step) gives T."""
    avg = ThermodynamicAveraging(averaging_steps)
    energy, abse = energy0, abs0
    skip = restarted
    lines = []
    for i, (d, a) in enumerate(zip(de, acc)):
        step = steps0 + i
        avg.add(energy)
        if skip:
            skip = False
        else:
            if step == 0:
                lines.append("steps\ttemperature\tenergy\taverage_energy\tabsolute_energy")
            if reference_due(step, L):
                T = temperature(step)
                lines.append(_fmt(step, T, energy, avg.average(1.0 / K_BOLTZMANN / T), abse))
        if a:
            energy += d
            abse += d
    return lines


class AnnealSchedule:
    """SimulatedAnnealing's temperature state (SimulatedAnnealing.cpp:59-73, 99-133) with
    maximum_steps N."""

    def __init__(self, t_start, energy, steps0, N):
        self.T = t_start
        self.lowest = self.recent_best = energy
        self.last_improvement, self.last_reheat = steps0, 0
        self.window_trials = self.window_accepts = self.reheats = 0
        self.window = max(1, int(N * 0.001))
        self.trigger = max(1, int(N * 0.05))
        self.cooldown = max(1, int(N * 0.10))
        self.N = max(1, N)

    def update(self, accepted, step, energy):
        self.window_trials += 1
        if accepted:
            self.window_accepts += 1
            if energy < self.recent_best - 1e-4:
                self.recent_best = energy
                self.last_improvement = step
        if self.window_trials >= self.window:
            if self.window_accepts / self.window_trials > 0.50:
                self.T *= 0.99
            self.window_trials = self.window_accepts = 0
        acc_est = self.window_accepts / self.window_trials if self.window_trials > 0 else 1.0
        if (self.reheats < 5 and step - self.last_improvement >= self.trigger
                and step - self.last_reheat >= self.cooldown and acc_est < 0.05):
            self.T *= 1.10
            self.last_improvement = self.last_reheat = step
            self.recent_best = energy
            self.reheats += 1
        self.T *= math.exp(-3.0 / self.N)


def sa_log_lines(steps0, energy0, abs0, de, acc, t_start, N, L):
    """The simulated-annealing log lines (header included for a run from step 0) of the trials
    steps0, steps0 + 1, ...; also returns the final schedule state."""
    s = AnnealSchedule(t_start, energy0, steps0, N)
    energy, abse = energy0, abs0
    lines = []
    for i, (d, a) in enumerate(zip(de, acc)):
        step = steps0 + i
        if step == 0:
            lines.append("steps\ttemperature\tenergy\tlowest_energy\tabsolute_energy")
        if energy < s.lowest - 1e-4:
            s.lowest = energy
            lines.append(_fmt(step, s.T, energy, s.lowest, abse))
        if step % L == 0:
            lines.append(_fmt(step, s.T, energy, s.lowest, abse))
        if a:
            energy += d
            abse += d
        s.update(bool(a), step, energy)
    return lines, s
