"""GPU parity: the HIP kernels against the CPU oracle on the same seeded inputs.

Bars: neighbour ids, event choice, occupancy and every integer state bit-exact; E_kra, dE, swap
and total energies within 1e-9 eV (absolute); residence time within 1e-12 relative.
"""
import numpy as np
import pytest

from conftest import QUATERNARY, TERNARY, random_alloc

pytestmark = pytest.mark.gpu

TOL_E = 1e-9


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _oracle_models(path, els, n, kinds):
    import oracle as O
    lat = O.Lattice(n)
    return lat, [O.Model(lat, path, els, k) for k in kinds]


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("name,kind", [("e0_3", 1), ("e0_3n", 1), ("e0_4n", 1), ("q_3n", 2),
                                       ("q_4", 2)])
@pytest.mark.parametrize("n", [5, 8])
def test_eval_jumps_matches_oracle(torch, coeffs, name, kind, n, form):
    """Both jump formulations (0 dense, 1 solvent baseline) against the oracle, solute
    fractions 8-28 % per element, vacancies on box corners and edges."""
    import latticemontecarlo_amd as L
    path, els = coeffs[name]
    lat, (om,) = _oracle_models(path, els, n, [kind])
    gm = L.Model(path, els, kind)
    gm.set_formulation(form)
    species = gm.species
    N = 4 * n ** 3
    occs, vacs = [], []
    for r in range(6):
        occ = random_alloc(n, species, 0.08 + 0.04 * r, seed=100 * n + r)
        x = species.index("X")
        v = [0, N - 1, 4 * n * n - 1][r] if r < 3 else int(np.flatnonzero(occ == x)[0])
        if r < 3:  # move the vacancy to a boundary site
            occ[occ == x] = occ[v]
            occ[v] = x
        occs.append(occ)
        vacs.append(v)
    batch = L.ReplicaBatch([n] * len(occs), occs)
    nb, ea, de = L.eval_jumps(gm, batch, vacs)
    torch.cuda.synchronize()
    gm.check_errors()
    nb, ea, de = nb.cpu().numpy(), ea.cpu().numpy(), de.cpu().numpy()
    for r, (occ, v) in enumerate(zip(occs, vacs)):
        ref_nb = lat.neighbors(v, 1)
        assert np.array_equal(nb[r], ref_nb)
        for i, b in enumerate(ref_nb):
            e_ref, d_ref = om.barrier_de(occ, v, int(b))
            assert abs(de[r, i] - d_ref) <= TOL_E, (r, i, de[r, i], d_ref)
            assert abs(ea[r, i] - e_ref) <= TOL_E, (r, i, ea[r, i], e_ref)


@pytest.mark.parametrize("name", ["e0_3n", "q_4"])
def test_swap_and_site_de_match_oracle(torch, coeffs, name):
    import latticemontecarlo_amd as L
    path, els = coeffs[name]
    n = 6
    lat, (om,) = _oracle_models(path, els, n, [3])
    gm = L.Model(path, els, L.PAIRSITE)
    species = gm.species
    N = 4 * n ** 3
    rng = np.random.default_rng(7)
    occ = random_alloc(n, species, 0.12, seed=77)
    pairs = []
    for r in range(48):
        p = int(rng.integers(N))
        if r % 3 == 0:
            q = int(rng.integers(N))  # mostly far pairs
        else:  # near pairs: 1st/2nd/3rd shells
            q = int(rng.choice(lat.neighbors(p, 1 + r % 3)))
        pairs.append((p, q))
    batch = L.ReplicaBatch([n] * len(pairs), [occ] * len(pairs))
    de = L.swap_de(gm, batch, pairs).cpu().numpy()
    for r, (p, q) in enumerate(pairs):
        assert abs(de[r] - om.swap_de(occ, p, q)) <= TOL_E
    sites = [int(rng.integers(N)) for _ in range(24)]
    codes = [int(rng.integers(len(species))) for _ in range(24)]
    batch2 = L.ReplicaBatch([n] * 24, [occ] * 24)
    sde = L.site_de(gm, batch2, sites, codes).cpu().numpy()
    for r in range(24):
        assert abs(sde[r] - om.site_de(occ, sites[r], codes[r])) <= TOL_E


@pytest.mark.parametrize("name", ["e0_3", "e0_4n"])
def test_total_energy_matches_oracle(torch, coeffs, name):
    import latticemontecarlo_amd as L
    path, els = coeffs[name]
    gm = L.Model(path, els, L.ENERGY)
    occs, ns = [], [4, 5, 7]
    for i, n in enumerate(ns):
        occs.append(random_alloc(n, gm.species, 0.1, seed=i))
    batch = L.ReplicaBatch(ns, occs)
    e = L.total_energy(gm, batch).cpu().numpy()
    for i, n in enumerate(ns):
        lat, (om,) = _oracle_models(path, els, n, [0])
        assert abs(e[i] - om.total_energy(occs[i])) <= TOL_E * max(1.0, abs(e[i]))


def _run_kmc_pair(coeffs, name, kind, n, R, steps, temps, form=1, maximum_steps=10 ** 9,
                  occs=None, escape_info=False, **opts):
    """The GPU driver and the oracle's KMC on the same replicas. `n` is one box size for all
    replicas or a list of R sizes (an uneven-size batch); `occs` overrides the random alloys."""
    import latticemontecarlo_amd as L
    import oracle as O
    path, els = coeffs[name]
    gb = L.Model(path, els, kind)
    ge = L.Model(path, els, L.ENERGY)
    species = gb.species
    sizes = list(n) if np.ndim(n) else [n] * R
    assert len(sizes) == R
    if occs is None:
        occs = [random_alloc(sizes[r], species, 0.05 + 0.03 * (r % 4), seed=500 + r)
                for r in range(R)]
    seed = 987654321
    tt = opts.pop("tt", None)
    drv = L.KineticMcFirst(gb, ge, sizes, occs, temps, seed, maximum_steps=maximum_steps, **opts)
    drv.set_formulation(form)
    if tt is not None:
        drv.set_schedule(*tt)
    drv.set_trace(steps)
    drv.run(steps)
    gst = drv.state()
    gtr = drv.trace()
    gocc = drv.occupancy()
    esc = drv.escape_info() if escape_info else None
    models = {}
    refs = []
    for r in range(R):
        nr = sizes[r]
        if nr not in models:
            lat = O.Lattice(nr)
            models[nr] = (lat, O.Model(lat, path, els, kind), O.Model(lat, path, els, 0))
        lat, ob, oe = models[nr]
        k = O.Kmc(ob, oe, lat, occs[r], float(np.broadcast_to(temps, (R,))[r]), seed, replica=r,
                  tt=tt, chain=form == 3, **opts)
        tr = k.run(steps, maximum_steps=maximum_steps, trace=True)
        refs.append((k.state(), tr, k.occupancy()))
    if escape_info:
        return gst, gtr, gocc, refs, esc
    return gst, gtr, gocc, refs


def _compare_kmc(gst, gtr, gocc, refs, steps):
    for r, (st, tr, occ) in enumerate(refs):
        assert np.array_equal(gtr["sel"][r], tr["sel"].astype(np.int64)), f"replica {r}"
        np.testing.assert_allclose(gtr["ea"][r], tr["ea"], rtol=0, atol=TOL_E)
        np.testing.assert_allclose(gtr["de"][r], tr["de"], rtol=0, atol=TOL_E)
        np.testing.assert_allclose(gtr["time"][r], tr["time"], rtol=1e-12, atol=0)
        assert np.array_equal(gocc[r], occ)
        assert int(gst["steps"][r]) == st["steps"] == steps
        assert int(gst["vacancy"][r]) == st["vacancy"]
        assert np.array_equal(gst["disp"][r], st["disp"])
        assert abs(gst["energy"][r] - st["energy"]) <= TOL_E * steps
        assert abs(gst["abs_energy"][r] - st["abs_energy"]) <= TOL_E * (steps + abs(st["abs_energy"]))
        np.testing.assert_allclose(gst["time"][r], st["time"], rtol=1e-12)


@pytest.mark.parametrize("form", [0, 1, 2])
@pytest.mark.parametrize("name,kind", [("e0_3n", 1), ("q_3n", 2)])
def test_kmc_trajectory_matches_oracle(torch, coeffs, name, kind, form):
    """Every kernel shape (0 dense, 1 solvent baseline, 2 dense wide) against the oracle."""
    steps = 400
    gst, gtr, gocc, refs = _run_kmc_pair(coeffs, name, kind, n=6, R=4, steps=steps,
                                         temps=np.array([450.0, 500.0, 600.0, 700.0]), form=form)
    _compare_kmc(gst, gtr, gocc, refs, steps)


@pytest.mark.parametrize("name,kind,opts", [("e0_3n", 1, {}),
                                            ("q_3n", 2, {"rate_corrector": True, "solute_disp": True})])
def test_chain_kmc_trajectory_matches_oracle(torch, coeffs, name, kind, opts):
    """Second-order (chain) steps (formulation 3) against the oracle's KineticMcChainOmpi."""
    steps = 150
    print(f"chain KMC {name} {opts}", flush=True)
    gst, gtr, gocc, refs = _run_kmc_pair(coeffs, name, kind, n=6, R=3, steps=steps,
                                         temps=np.array([450.0, 550.0, 700.0]), form=3, **opts)
    _compare_kmc(gst, gtr, gocc, refs, steps)


def test_kmc_autotune_keeps_results(torch, coeffs, tmp_path):
    """The autotuner times every kernel shape without changing the replicas' state, tracked
    atoms or log records."""
    import latticemontecarlo_amd as L
    path, els = coeffs["e0_3"]
    gb = L.Model(path, els, L.E0)
    n, R = 6, 8
    N = 4 * n ** 3
    occs = [random_alloc(n, gb.species, 0.03, seed=70 + r) for r in range(R)]
    rng = np.random.default_rng(4)
    l2a = [rng.permutation(N) for _ in range(R)]
    a = L.KineticMcFirst(gb, None, [n] * R, occs, 500.0, 3, maximum_steps=10 ** 6)
    b = L.KineticMcFirst(gb, None, [n] * R, occs, 500.0, 3, maximum_steps=10 ** 6)
    for d in (a, b):
        d.set_atoms(l2a)
        d.set_log(10, capacity=1000)
    info = a.autotune(steps=50)
    assert info["formulation"] in (0, 1, 2) and info["ms_dense"] > 0 and info["ms_solvent"] > 0
    assert info["ms_wide"] > 0  # 8 replicas: every one has a CU of its own
    a.run(200)
    b.run(200)
    sa, sb = a.state(), b.state()
    assert np.array_equal(sa["vacancy"], sb["vacancy"])
    assert np.array_equal(sa["steps"], sb["steps"])
    np.testing.assert_allclose(sa["time"], sb["time"], rtol=1e-12)
    (aa, ash), (ba, bsh) = a.atoms(), b.atoms()
    assert all(np.array_equal(x, y) for x, y in zip(aa, ba))
    assert all(np.array_equal(x, y) for x, y in zip(ash, bsh))
    fa, fb = tmp_path / "a.txt", tmp_path / "b.txt"
    a.write_log([fa] + [None] * (R - 1))
    b.write_log([fb] + [None] * (R - 1))
    la, lb = fa.read_text().split("\n"), fb.read_text().split("\n")
    assert len(la) == len(lb) == 99 + 10 + 2  # header, 109 records (0-98, 100, 110-190), end
    # same steps, atoms and positions; times agree to rounding (shapes sum in other orders)
    assert [x.split("\t")[0::6] for x in la] == [x.split("\t")[0::6] for x in lb]


def test_kmc_options_match_oracle(torch, coeffs):
    steps = 300
    tt = (np.array([0.0, 1e-9, 1e-6]), np.array([800.0, 600.0, 400.0]))
    gst, gtr, gocc, refs = _run_kmc_pair(coeffs, "e0_4n", 1, n=5, R=3, steps=steps,
                                         temps=550.0, rate_corrector=True, solute_disp=True,
                                         tt=tt)
    _compare_kmc(gst, gtr, gocc, refs, steps)
    for r, (st, _, _) in enumerate(refs):
        np.testing.assert_allclose(gst["com"][r], st["com"], rtol=0, atol=1e-9)
        np.testing.assert_allclose(gst["temperature"][r], st["temperature"], rtol=1e-12)


@pytest.mark.parametrize("lds", [True, False])
def test_cmc_trajectory_matches_oracle(torch, coeffs, lds):
    """Occupancy held in LDS during the run (replicas of different sizes) or in HBM."""
    import latticemontecarlo_amd as L
    import oracle as O
    path, els = coeffs["e0_3n"]
    sizes, trials = [5, 4, 6], 600
    R = len(sizes)
    gp = L.Model(path, els, L.PAIRSITE)
    ge = L.Model(path, els, L.ENERGY)
    occs = [random_alloc(n, gp.species, 0.05, seed=900 + r, vacancy=False)
            for r, n in enumerate(sizes)]
    seed = 42
    print(f"cmc sizes={sizes} lds={lds}", flush=True)
    drv = L.CanonicalMc(gp, ge, sizes, occs, 800.0, 300.0, seed, maximum_steps=trials)
    drv.set_lds_occupancy(lds)
    drv.set_trace(trials)
    drv.run(trials // 3)
    drv.run(trials - trials // 3)
    gst, gtr, gocc = drv.state(), drv.trace(), drv.occupancy()
    for r, n in enumerate(sizes):
        lat = O.Lattice(n)
        op = O.Model(lat, path, els, O.PAIRSITE)
        oe = O.Model(lat, path, els, O.ENERGY)
        c = O.Cmc(op, oe, lat, occs[r], 800.0, 300.0, seed, replica=r)
        tr = c.run(trials, trials, trace=True)
        st = c.state()
        assert np.array_equal(gtr["p"][r], tr["p"]) and np.array_equal(gtr["q"][r], tr["q"])
        np.testing.assert_allclose(gtr["de"][r], tr["de"], rtol=0, atol=TOL_E)
        assert np.array_equal(gtr["acc"][r], tr["acc"])
        assert np.array_equal(gocc[r], c.occupancy())
        assert int(gst["accepted"][r]) == st["accepted"]
        assert abs(gst["energy"][r] - st["energy"]) <= TOL_E * trials


def test_annealing_matches_oracle(torch, coeffs):
    """SimulatedAnnealing's adaptive schedule (SimulatedAnnealing.cpp:99-185): trials, acceptance,
    temperature, lowest energy and reheats against the oracle, over two launches."""
    import latticemontecarlo_amd as L
    import oracle as O
    path, els = coeffs["e0_3n"]
    n, R, N = 5, 3, 4000
    gp = L.Model(path, els, L.PAIRSITE)
    occs = [random_alloc(n, gp.species, 0.08, seed=950 + r, vacancy=False) for r in range(R)]
    seed, T0 = 77, 2500.0
    print(f"annealing n={n} R={R} N={N}", flush=True)
    drv = L.SimulatedAnnealing(gp, None, [n] * R, occs, T0, seed, maximum_steps=N)
    drv.set_trace(N)
    drv.run(N // 2)
    drv.run(N // 2)
    gst, gtr, gocc, ga = drv.state(), drv.trace(), drv.occupancy(), drv.annealing_state()
    lat = O.Lattice(n)
    op = O.Model(lat, path, els, O.PAIRSITE)
    for r in range(R):
        c = O.Cmc(op, None, lat, occs[r], T0, T0, seed, replica=r, anneal=True)
        tr = c.run(N, N, trace=True)
        st, oa = c.state(), c.anneal_state()
        assert np.array_equal(gtr["p"][r], tr["p"]) and np.array_equal(gtr["q"][r], tr["q"])
        assert np.array_equal(gtr["acc"][r], tr["acc"])
        np.testing.assert_allclose(gtr["de"][r], tr["de"], rtol=0, atol=TOL_E)
        assert np.array_equal(gocc[r], c.occupancy())
        assert int(gst["accepted"][r]) == st["accepted"]
        assert abs(gst["energy"][r] - st["energy"]) <= TOL_E * N
        assert ga["temperature"][r] == pytest.approx(oa["temperature"], rel=1e-12)
        assert abs(ga["lowest"][r] - oa["lowest"]) <= TOL_E * N
        assert int(ga["reheats"][r]) == oa["reheats"]
    assert ga["temperature"].max() < T0


@pytest.mark.parametrize("form", [0, 1, 2, 3])
def test_kmc_graph_replay_matches_eager(torch, coeffs, form):
    """Launches replayed from a captured HIP graph do exactly what the same launches do eagerly."""
    import latticemontecarlo_amd as L
    path, els = coeffs["e0_3n"]
    gb = L.Model(path, els, L.E0)
    n, R = 6, 5
    occs = [random_alloc(n, gb.species, 0.06, seed=300 + r) for r in range(R)]
    temps = np.linspace(450.0, 650.0, R)
    a = L.KineticMcFirst(gb, None, [n] * R, occs, temps, 21, maximum_steps=10 ** 6)
    b = L.KineticMcFirst(gb, None, [n] * R, occs, temps, 21, maximum_steps=10 ** 6)
    a.set_formulation(form)
    b.set_formulation(form)
    for _ in range(6):
        a.run(50)
    b.build_graph(50, 3)
    b.run_graph(50, 3)
    b.run_graph(50, 3)
    sa, sb = a.state(), b.state()
    assert np.array_equal(sa["steps"], sb["steps"]) and np.array_equal(sa["vacancy"], sb["vacancy"])
    assert np.array_equal(sa["time"], sb["time"]) and np.array_equal(sa["energy"], sb["energy"])
    for x, y in zip(a.occupancy(), b.occupancy()):
        assert np.array_equal(x, y)


def test_kmc_autotune_is_not_truncated_by_maximum_steps(torch, coeffs):
    """The autotuner's timing runs ignore the driver's maximum_steps: a driver limited to 5 steps
    times the shapes like one without a limit, and its own replicas do not move."""
    import latticemontecarlo_amd as L
    path, els = coeffs["e0_3"]
    gb = L.Model(path, els, L.E0)
    n, R = 8, 64
    occs = [random_alloc(n, gb.species, 0.03, seed=170 + r) for r in range(R)]
    short = L.KineticMcFirst(gb, None, [n] * R, occs, 500.0, 3, maximum_steps=5)
    free = L.KineticMcFirst(gb, None, [n] * R, occs, 500.0, 3, maximum_steps=10 ** 12)
    a = short.autotune(steps=300)
    b = free.autotune(steps=300)
    print(f"autotune limited {a} unlimited {b}", flush=True)
    for key in ("ms_dense", "ms_solvent", "ms_wide"):
        assert 1 / 3 < a[key] / b[key] < 3, key
    st = short.state()
    assert np.array_equal(st["steps"], np.zeros(R, np.uint64))


def test_kmc_graph_follows_a_new_schedule(torch, coeffs):
    """A captured graph is dropped when the temperature schedule changes: build_graph, then
    set_schedule, then run_graph does what eager launches under the new schedule do."""
    import latticemontecarlo_amd as L
    path, els = coeffs["e0_3n"]
    gb = L.Model(path, els, L.E0)
    n, R = 6, 4
    occs = [random_alloc(n, gb.species, 0.06, seed=330 + r) for r in range(R)]
    tt = (np.array([0.0, 1e-9, 1e-6]), np.array([900.0, 600.0, 400.0]))
    a = L.KineticMcFirst(gb, None, [n] * R, occs, 500.0, 21, maximum_steps=10 ** 6)
    b = L.KineticMcFirst(gb, None, [n] * R, occs, 500.0, 21, maximum_steps=10 ** 6)
    a.set_schedule(*tt)
    for _ in range(4):
        a.run(50)
    b.build_graph(50, 2)
    b.set_schedule(*tt)
    b.run_graph(50, 2)
    b.run_graph(50, 2)
    sa, sb = a.state(), b.state()
    assert np.array_equal(sa["vacancy"], sb["vacancy"]) and np.array_equal(sa["time"], sb["time"])
    assert np.array_equal(sa["temperature"], sb["temperature"])
    for x, y in zip(a.occupancy(), b.occupancy()):
        assert np.array_equal(x, y)
