"""The engine's exact-integer tables against the oracle's floating-point restatement (no GPU).

Neighbour lists, the 473 state clusters of a jump and the MMM/MM2 orbit clusters (orbit order and
the within-cluster order of non-symmetric pairs, which fixes the one-hot index) must agree as
lattice ids, including jumps that wrap around the periodic box.
"""
import json
import os

import numpy as np
import pytest

import latticemontecarlo_amd as L
import oracle as O
from conftest import TERNARY


@pytest.fixture(scope="module")
def zero_json(tmp_path_factory):
    p = str(tmp_path_factory.mktemp("z") / "zero.json")
    with open(p, "w") as f:
        json.dump({"Base": {"theta": [0.0] * 95}}, f)
    return p


@pytest.mark.parametrize("n", [4, 5, 7])
def test_neighbor_lists(n):
    lat = O.Lattice(n)
    for s in range(0, lat.num_sites, 3):
        for sh in (1, 2, 3):
            assert np.array_equal(lat.neighbors(s, sh), L.fcc_neighbors(n, s, sh))


def _state_key(r):
    return (int(r[0]), tuple(sorted(int(x) for x in r[1:] if x >= 0)))


def _orbit_key(r):
    o, sym, s0, s1 = map(int, r)
    return (o, sym, (s0, s1) if not sym else tuple(sorted((s0, s1))))


@pytest.mark.parametrize("n", [4, 5, 9])
def test_jump_cluster_tables(n, zero_json):
    lat = O.Lattice(n)
    m = O.Model(lat, zero_json, TERNARY, O.ENERGY)
    N = lat.num_sites
    rng = np.random.default_rng(n)
    sites = [0, N - 1, 2 * n * n - 1] + list(rng.integers(0, N, 3))
    for v in sites:
        for nb in lat.neighbors(int(v), 1):
            a = sorted(map(_state_key, m.state_clusters(int(v), int(nb))))
            b = sorted(map(_state_key, L.jump_state_clusters(n, int(v), int(nb))))
            assert a == b
            for which in (0, 1):
                a = sorted(map(_orbit_key, m.orbit_clusters(int(v), int(nb), which)))
                b = sorted(map(_orbit_key, L.jump_orbit_clusters(n, int(v), int(nb), which)))
                assert a == b


def test_site_cluster_count():
    with open(os.path.join(os.path.dirname(__file__), "golden", "table_sizes.json")) as f:
        g = json.load(f)
    from latticemontecarlo_amd import synth
    lay = synth.layout(TERNARY)
    assert lay["n_eci"] == g["ternary"]["n_eci"]
    assert lay["F_mmm"] == g["ternary"]["mmm"] and lay["F_mm2"] == g["ternary"]["mm2"]
