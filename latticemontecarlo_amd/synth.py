"""Synthetic inputs in the reference's formats.

* Coefficient JSON in the reference schema (lmc/pred/src/EnergyPredictor.cpp:33-37,
  VacancyMigrationPredictorE0.cpp:28-40, VacancyMigrationPredictorQuartic.cpp:44-63).
  Cluster ECIs are drawn as per-occurrence energies eps ~ N(0, 0.02 eV) for pairs at 1st-3rd
  neighbours and first-neighbour triplets, and stored as theta = eps * N_label with the dE
  normalisers N = 1536/768/3072/2048 (VacancyMigrationPredictorQuartic.cpp:13-14), so that
  dE = sum of eps over the clusters that change. The E0 barrier uses mu_e0 = ln(base),
  sigma_e0 = 1, mu_x = 0, sigma_x = 1, one U row w with w_f = |orbit| eps_f / base and
  eps_f ~ N(0, 0.05 eV) on solute-containing types, so E_kra = base exp(sum eps / base)
  ~= base + sum eps. (The exact linear form base + sum eps is not expressible in the E0 schema.)
* Occupancies: species placed by a seeded uniform shuffle, then one uniformly random site
  becomes the vacancy.
"""
import json
import os
import tempfile

import numpy as np

from . import api

# Migration base barriers (eV); Cu is not given by the workload definitions and is set here.
BASE_BARRIER = {"Al": 0.60, "Mg": 0.55, "Zn": 0.50, "Cu": 0.52, "Sn": 0.50}
DE_NORM = {0: 256, 1: 1536, 2: 768, 3: 3072, 4: 2048, 5: 3072, 6: 6144, 7: 6144}


def eci_count(element_set):
    species = sorted(set(element_set) | {"X"})
    x = species.index("X")
    keys = set()
    S = len(species)
    for a in range(S):
        keys.add((1, 0, (a,)))
        for b in range(S):
            if b == x:
                continue
            for lab in (1, 2, 3):
                keys.add((2, lab, tuple(sorted((a, b)))))
            for c in range(S):
                if c == x:
                    continue
                for lab in (4, 5, 6, 7):
                    keys.add((3, lab, tuple(sorted((a, b, c)))))
    return len(keys)


def layout(element_set):
    """ECI keys and MMM/MM2 orbit layout, read from the engine's own tables."""
    els = sorted(set(element_set) - {"X"})
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "zero.json")
        with open(p, "w") as f:
            json.dump({"Base": {"theta": [0.0] * eci_count(els)}}, f)
        m = api.Model(p, els, api.ENERGY)
        return dict(species=m.species, eci=m.eci_keys(), mmm=m.orbits(0), mm2=m.orbits(1),
                    n_eci=m.layout()[0], F_mmm=m.layout()[1], F_mm2=m.layout()[2])


def _onehot_solvent_type(kind, ne, rs):
    if kind == 0:
        return rs
    if kind == 1:
        return rs * ne + rs
    return rs * ne - rs * (rs - 1) // 2  # triangular index of (rs, rs)


def _feature_arrays(orb, ne):
    F = int(orb["base"][-1] + orb["len"][-1])
    size = np.zeros(F)
    kind = np.zeros(F, np.int64)
    tidx = np.zeros(F, np.int64)
    for o in range(len(orb["size"])):
        b, l = int(orb["base"][o]), int(orb["len"][o])
        size[b:b + l] = orb["size"][o]
        kind[b:b + l] = orb["kind"][o]
        tidx[b:b + l] = np.arange(l)
    return F, size, kind, tidx


def make_coefficients(element_set, kind="E0", seed=0, eci_sigma=0.02, barrier_sigma=0.05,
                      solvent="Al", randomize_norm=False, n_rows=1):
    """Coefficient dict in the reference schema. kind: 'E0' or 'Quartic' (both carry Base)."""
    els = sorted(set(element_set) - {"X"})
    lay = layout(els)
    rng = np.random.default_rng(seed)
    theta = np.zeros(lay["n_eci"])
    for i, (label, _codes) in enumerate(lay["eci"]):
        if label in (1, 2, 3, 4):
            theta[i] = rng.normal(0.0, eci_sigma) * DE_NORM[label]
    out = {"Base": {"theta": theta.tolist()}}
    ne = len(els)
    rs = els.index(solvent) if solvent in els else 0
    F1, size1, kind1, t1 = _feature_arrays(lay["mmm"], ne)
    F2, size2, kind2, t2 = _feature_arrays(lay["mm2"], ne)
    solv1 = t1 == np.array([_onehot_solvent_type(k, ne, rs) for k in kind1])
    solv2 = t2 == np.array([_onehot_solvent_type(k, ne, rs) for k in kind2])
    for e in els:
        base = BASE_BARRIER.get(e, 0.55)
        if kind == "E0":
            rows = []
            for r in range(n_rows):
                eps = rng.normal(0.0, barrier_sigma, F1) * (r == 0 or 0.3)
                eps[solv1] = 0.0
                rows.append((size1 * eps / base).tolist())
            mu = np.zeros(F1)
            sg = np.ones(F1)
            if randomize_norm:
                mu = rng.uniform(-0.05, 0.05, F1)
                sg = rng.uniform(0.5, 2.0, F1)
            out[e] = {"mu_x_mmm": mu.tolist(), "sigma_x_mmm": sg.tolist(), "U_mmm": rows,
                      "theta_e0": [1.0] + [0.5] * (n_rows - 1), "mu_e0": float(np.log(base)),
                      "sigma_e0": 1.0}
        elif kind == "Quartic":
            k = max(n_rows, 2)
            U1 = rng.normal(0.0, 0.01, (k, F1)) * size1
            U2 = rng.normal(0.0, 0.005, (k, F2)) * size2
            U1[:, solv1] = 0.0
            U2[:, solv2] = 0.0
            mu1, sg1, mu2, sg2 = np.zeros(F1), np.ones(F1), np.zeros(F2), np.ones(F2)
            if randomize_norm:
                mu1, mu2 = rng.uniform(-0.05, 0.05, F1), rng.uniform(-0.05, 0.05, F2)
                sg1, sg2 = rng.uniform(0.5, 2.0, F1), rng.uniform(0.5, 2.0, F2)
            out[e] = {"mu_x_mmm": mu1.tolist(), "mu_x_mm2": mu2.tolist(),
                      "sigma_x_mmm": sg1.tolist(), "sigma_x_mm2": sg2.tolist(),
                      "U_mmm": U1.tolist(), "U_mm2": U2.tolist(),
                      "theta_D": [1.0] + [0.5] * (k - 1), "theta_Ks": [1.0] + [-0.5] * (k - 1),
                      "mu_D": float(np.log(2.86)), "mu_Ks": float(np.log(4.7 * base / 0.6)),
                      "sigma_D": 1.0, "sigma_Ks": 1.0}
        else:
            raise ValueError(kind)
    return out


def write_coefficients(path, element_set, kind="E0", **kw):
    with open(path, "w") as f:
        json.dump(make_coefficients(element_set, kind, **kw), f)
    return path


def random_occupancy(n, species, counts, seed, vacancy=True):
    """Occupancy codes for an n^3-cell FCC box: `counts` maps element -> number of sites
    (the first element of `species` order absent from counts fills the rest)."""
    N = 4 * n ** 3
    rng = np.random.default_rng(seed)
    codes = []
    filled = 0
    for e, c in counts.items():
        codes += [species.index(e)] * int(c)
        filled += int(c)
    rest = [s for s in species if s not in counts and s != "X"]
    codes += [species.index(rest[0])] * (N - filled)
    occ = np.array(codes, np.uint8)
    rng.shuffle(occ)
    if vacancy:
        occ[rng.integers(N)] = species.index("X")
    return occ


def composition_counts(n, fractions):
    N = 4 * n ** 3
    return {e: int(round(f * N)) for e, f in fractions.items()}
