"""Exit-time analysis of a vacancy configuration, the host mirror of ansys::ExitTime
(lmc/ansys/include/ExitTime.h, lmc/ansys/src/ExitTime.cpp).

Every migration barrier is evaluated on the GPU by lmc_vacancy_scan: the configuration with its
vacancy turned into the solvent and a vacancy placed at each scanned site in turn. What is left
on the host is small: sorting and averaging the per-site barrier lists, the barrier statistics
of a cluster and its first-neighbour shell, and the absorbing Markov chain of a cluster (a dense
solve over the cluster's sites).

Profile and binding energies (the Markov chain's base energies) are batched site energy changes
on the GPU (lmc_site_change_de) with chemical potentials from api.chemical_potential.

Differences from the reference: sites are lattice ids (atom ids live in cfg.Config), and sets
are iterated in ascending id order instead of hash order (sums differ only by rounding).
"""
import numpy as np

from . import api

K_BOLTZMANN = 8.617333262145e-5  # eV/K (Constants.hpp)
K_PREFACTOR = 1e13               # attempt frequency, Hz (Constants.hpp)


class ExitTime:
    """ExitTime(config, solvent, element_set, escape_temperature, current_temperature, ...)
    (ExitTime.cpp:4-21) over one replica: `occ` (uint8 species codes, host array or device
    tensor) of n^3 cells, the barrier model, and the solvent's species code."""

    def __init__(self, barrier_model, occ, n, solvent, escape_temperature,
                 current_temperature=None, pairsite_model=None):
        torch = api._torch()
        self.model = barrier_model
        self.pairsite = pairsite_model
        self.n = int(n)
        self.N = api.fcc_num_sites(self.n)
        self.solvent = int(solvent)
        self.occ = torch.as_tensor(np.asarray(occ.cpu() if hasattr(occ, "cpu") else occ,
                                              np.uint8)).cuda()
        if self.occ.numel() != self.N:
            raise ValueError(f"occupancy has {self.occ.numel()} sites, expected {self.N}")
        x = barrier_model.species.index("X")
        vac = np.flatnonzero(self.occ.cpu().numpy() == x)
        self.vacancy = int(vac[0]) if vac.size else -1
        self.beta_escape = 1.0 / K_BOLTZMANN / escape_temperature
        self.beta_current = (1.0 / K_BOLTZMANN / current_temperature
                             if current_temperature is not None else None)

    def scan(self, sites):
        """(nb, Ea, dE) [K, 12] host arrays of a vacancy at each of `sites` (see
        api.vacancy_scan)."""
        nb, ea, de = api.vacancy_scan(self.model, self.occ, self.n, self.vacancy, self.solvent,
                                      sites)
        self.model.check_errors()
        return nb.cpu().numpy(), ea.cpu().numpy(), de.cpu().numpy()

    def barrier_list_and_exit_time(self, sites=None):
        """GetBarrierListAndExitTime (ExitTime.cpp:216-244): per site, the 12 migration barriers
        of a vacancy placed there (ascending), their mean, and the exit time
        1 / (nu sum_j exp(-Ea_j beta_escape)). Default: every site."""
        s = np.arange(self.N) if sites is None else np.asarray(sites, np.int64)
        _, ea, _ = self.scan(s)
        lists = np.sort(ea, axis=1)
        means = lists.sum(axis=1) / 12.0
        exit_times = 1.0 / np.exp(-lists * self.beta_escape).sum(axis=1) / K_PREFACTOR
        return lists, means, exit_times

    def jump_energetics(self, sites):
        """GetJumpEnergetics (ExitTime.cpp:272-311) of the set `sites` and its first-neighbour
        shell: the map (site, neighbour) -> (Ea, dE), and per-site arrays [N, 12] of neighbour
        ids (default N), barriers and driving forces (default NaN)."""
        plus = first_shell(self.n, sites)
        nb, ea, de = self.scan(plus)
        nbl = np.full((self.N, 12), self.N, np.int64)
        bl = np.full((self.N, 12), np.nan)
        dl = np.full((self.N, 12), np.nan)
        nbl[plus], bl[plus], dl[plus] = nb, ea, de
        pair = {(int(s), int(nb[i, j])): (float(ea[i, j]), float(de[i, j]))
                for i, s in enumerate(plus) for j in range(12)}
        return pair, nbl, bl, dl

    def _site_change(self, occ, sites, frm, to):
        """Energy changes of sites[i] going from species frm[i] to to[i] in the device occupancy
        occ (one batched launch of lmc_site_change_de over a shared occupancy)."""
        if self.pairsite is None:
            raise ValueError("this analysis needs the pair-site model (pairsite_model=)")
        torch = api._torch()
        K = len(sites)
        dev = occ.device
        off = torch.zeros(K, dtype=torch.int64, device=dev)
        nc = torch.full((K,), self.n, dtype=torch.int32, device=dev)
        s = torch.as_tensor(np.asarray(sites, np.int64), device=dev)
        f = torch.as_tensor(np.asarray(frm, np.uint8), device=dev)
        t = torch.as_tensor(np.asarray(to, np.uint8), device=dev)
        de = torch.empty(K, dtype=torch.float64, device=dev)
        api.check(api.lib().lmc_site_change_de(self.pairsite.h, api.dptr(occ), api.dptr(off),
                                               api.dptr(nc), api.dptr(s), api.dptr(f),
                                               api.dptr(t), K, api.dptr(de), None))
        self.pairsite.check_errors()
        return de.cpu().numpy()

    def _without_vacancy(self):
        occ = self.occ.clone()
        if self.vacancy >= 0:
            occ[self.vacancy] = self.solvent
        return occ

    def profile_energy(self, chemical_potential):
        """GetProfileEnergy (ExitTime.cpp:473-512): per site of the configuration whose vacancy
        holds the solvent, the energy of turning it into a vacancy plus mu(its element) - mu(X);
        at the vacancy's own site the largest value over the elements it could hold.
        chemical_potential: {element: mu} (api.chemical_potential)."""
        sp = self.pairsite.species if self.pairsite is not None else self.model.species
        x = sp.index("X")
        occ = self._without_vacancy()
        host = occ.cpu().numpy()
        mu = np.array([chemical_potential[e] for e in sp])
        de = self._site_change(occ, np.arange(self.N), host, np.full(self.N, x))
        prof = de + mu[host] - mu[x]
        if self.vacancy >= 0:
            els = [c for c, e in enumerate(sp) if e != "X"]
            dv = self._site_change(occ, [self.vacancy] * len(els), els, [x] * len(els))
            prof[self.vacancy] = max(dv[i] + mu[c] - mu[x] for i, c in enumerate(els))
        return prof

    def binding_energy(self, chemical_potential):
        """GetBindingEnergy (ExitTime.cpp:401-429): {element: [N]} over the elements of
        chemical_potential other than X; entry s is the energy of turning site s, holding that
        element, into a vacancy minus mu(X) - mu(element), in the configuration whose vacancy
        holds the solvent."""
        sp = self.pairsite.species
        x = sp.index("X")
        occ = self._without_vacancy()
        els = [e for e in chemical_potential if e != "X"]
        codes = np.repeat([sp.index(e) for e in els], self.N)
        sites = np.tile(np.arange(self.N), len(els))
        de = self._site_change(occ, sites, codes, np.full(sites.size, x))
        out = {}
        for i, e in enumerate(els):
            pot = chemical_potential["X"] - chemical_potential[e]
            out[e] = de[i * self.N:(i + 1) * self.N] - pot
        return out

    @staticmethod
    def markov_escape_time(sites, neighbor_lists, barrier_lists, base_energy, beta):
        """BuildMarkovChain (ExitTime.cpp:165-214): mean time for a vacancy that starts in the
        transient set `sites` (weights exp(-base_energy beta)) to leave it. Each site's 12 jumps
        have rates nu exp(-Ea beta); jumps to sites outside the set are absorbing. Returns
        p^T (I - T)^-1 tau, with T the transient jump probabilities and tau the mean residence
        times."""
        sites = [int(x) for x in sites]
        idx = {s: i for i, s in enumerate(sites)}
        m = len(sites)
        T = np.zeros((m, m))
        p = np.zeros(m)
        tau = np.zeros(m)
        for s, i in idx.items():
            p[i] = np.exp(-base_energy[s] * beta)
            rates = K_PREFACTOR * np.exp(-np.asarray(barrier_lists[s], np.float64) * beta)
            total = rates.sum()
            tau[i] = 1.0 / total
            for j in range(12):
                k = idx.get(int(neighbor_lists[s][j]))
                if k is not None:
                    T[i, k] = rates[j] / total
        p /= p.sum()
        return float(p @ np.linalg.solve(np.eye(m) - T, tau))

    def average_barriers(self, sites, pair_energy):
        """GetAverageBarriers (ExitTime.cpp:313-396): mean and standard deviation of the kinetic
        barriers Ea - dE/2 of each jump class of jump_classes(): [mean_in, mean_to, mean_on,
        mean_off, std_in, std_to, std_on, std_off] (NaN for an empty class). Jumps missing from
        `pair_energy` are evaluated."""
        classes = jump_classes(self.n, sites)
        need = sorted({s for v in classes.values() for s, t in v
                       if (s, t) not in pair_energy})
        if need:
            nb, ea, de = self.scan(need)
            pair_energy = dict(pair_energy)
            for i, s in enumerate(need):
                for j in range(12):
                    pair_energy.setdefault((s, int(nb[i, j])), (float(ea[i, j]), float(de[i, j])))

        def stats(pairs):
            if not pairs:
                return np.nan, np.nan
            v = np.array([pair_energy[p][0] - pair_energy[p][1] / 2 for p in pairs])
            mean = v.sum() / v.size
            return mean, float(np.sqrt(((v - mean) ** 2).sum() / v.size))

        ms = [stats(classes[k]) for k in ("in", "to", "on", "off")]
        return [m for m, _ in ms] + [s for _, s in ms]


def first_shell(n, sites):
    """`sites` and their first neighbours on the n^3-cell lattice, ascending."""
    out = set(int(x) for x in sites)
    for x in list(out):
        out.update(int(y) for y in api.fcc_neighbors(n, x, 1))
    return np.array(sorted(out), np.int64)


def jump_classes(n, sites):
    """The jumps (origin, neighbour) from the set `sites` and its first-neighbour shell, by class
    (ExitTime.cpp:334-356): "in" (inside the set), "to" (set -> shell), "on" (shell -> shell) and
    "off" (shell -> outside); jumps from the shell into the set belong to no class."""
    inside = set(int(x) for x in sites)
    plus = first_shell(n, sites)
    shell = set(int(x) for x in plus)
    out = {"in": [], "to": [], "on": [], "off": []}
    for s in plus:
        s = int(s)
        s_in = s in inside
        for t in sorted(int(x) for x in api.fcc_neighbors(n, s, 1)):
            t_in, t_nn = t in inside, t in shell
            if s_in and t_in:
                out["in"].append((s, t))
            elif s_in and not t_in and t_nn:
                out["to"].append((s, t))
            elif not s_in and not t_in and t_nn:
                out["on"].append((s, t))
            elif not s_in and not t_nn:
                out["off"].append((s, t))
    return out
