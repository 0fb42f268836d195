"""GPU checks of the surrogate: training samples against the CE oracle, the MFMA forward pass
and gradients against the MLP oracle, and training on exact samples."""
import numpy as np
import pytest

from conftest import TERNARY, random_alloc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _batch(L, species, n, fracs, seed):
    occs, vacs = [], []
    x = species.index("X")
    for r, f in enumerate(fracs):
        occ = random_alloc(n, species, f, seed=seed + r)
        occs.append(occ)
        vacs.append(int(np.flatnonzero(occ == x)[0]))
    return L.ReplicaBatch([n] * len(occs), occs), occs, vacs


@pytest.mark.parametrize("name,kind", [("e0_3n", 1), ("q_4", 2)])
def test_jump_samples_match_oracle(torch, coeffs, name, kind):
    import latticemontecarlo_amd as L
    import oracle as O
    path, els = coeffs[name]
    n = 6
    gm = L.Model(path, els, kind)
    batch, occs, vacs = _batch(L, gm.species, n, [0.05, 0.1, 0.2], seed=40)
    env, ekra, de = L.jump_samples(gm, batch, vacs)
    gm.check_errors()
    env, ekra, de = env.cpu().numpy(), ekra.cpu().numpy(), de.cpu().numpy()
    lat = O.Lattice(n)
    om = O.Model(lat, path, els, kind)
    for r, (occ, v) in enumerate(zip(occs, vacs)):
        for i, nb in enumerate(sorted(lat.neighbors(v, 1))):
            row = env[12 * r + i]
            sites = L.jump_environment(n, v, int(nb))
            assert np.array_equal(row[:60], occ[sites]) and not row[60:].any()
            ea_ref, de_ref = om.barrier_de(occ, v, int(nb))
            assert abs(de[12 * r + i] - de_ref) <= 1e-9
            assert abs(max(0.0, ekra[12 * r + i] + de[12 * r + i] / 2) - ea_ref) <= 1e-9


def _env_rows(rng, n, species, chans):
    """Random environment rows: mostly solvent, solutes on random slots, vacancy in slot 0."""
    env = np.zeros((n, 64), np.uint8)
    env[:, :60] = species.index("Al")
    for c in chans:
        mask = rng.random((n, 60)) < 0.12
        env[:, :60][mask] = c
    env[:, 0] = species.index("X")
    return env


@pytest.mark.parametrize("name", ["e0_3n", "e0_4n"])
def test_forward_and_gradients_match_oracle(torch, coeffs, name):
    import latticemontecarlo_amd as L
    import mlp_oracle as M
    path, els = coeffs[name]
    gm = L.Model(path, els, L.E0)
    sur = L.Surrogate(gm, n_members=4, seed=11)
    rng = np.random.default_rng(8)
    n = 150  # not a multiple of the 64-row tile
    env = _env_rows(rng, n, gm.species, sur.channel_codes)
    # perturb the parameters so that biases are non-zero
    p = sur.get_params()
    p += rng.uniform(-0.05, 0.05, p.shape).astype(np.float32)
    o = M.layout(sur.inputs)
    p[:, o["W1"] + 60 * sur.channels * 128:o["b1"]] = 0.0
    p[:, o["b3"] + 2:] = 0.0
    sur.set_params(p)
    denv = torch.from_numpy(env).cuda()
    y = sur.forward(denv).cpu().numpy()
    x = M.one_hot(env, sur.channel_codes, sur.inputs)
    t = rng.normal(size=(n, 2)).astype(np.float32)
    g, loss = sur.gradients(denv, torch.from_numpy(t).cuda())
    g = g.cpu().numpy()
    for m in range(4):
        y_ref, _ = M.forward(p[m], x)
        np.testing.assert_allclose(y[:, m], y_ref, rtol=0, atol=2e-3 * max(1.0, np.abs(y_ref).max()))
        g_ref = M.gradients(p[m], x, t.astype(np.float64))
        assert abs(loss[m] / (2 * n) - M.loss(p[m], x, t)) <= 1e-3 * M.loss(p[m], x, t)
        for a, b in (("W1", "b1"), ("b1", "W2"), ("W2", "b2"), ("b2", "W3"), ("W3", "b3")):
            gs, gr = g[m, o[a]:o[b]], g_ref[o[a]:o[b]]
            assert np.linalg.norm(gs - gr) <= 1e-2 * np.linalg.norm(gr) + 1e-7, (m, a)
        np.testing.assert_allclose(g[m, o["b3"]:o["b3"] + 2], g_ref[o["b3"]:o["b3"] + 2],
                                   rtol=1e-3, atol=1e-6)


def test_training_learns_the_barriers(torch, coeffs):
    """Four members trained on the GPU on exact trajectory samples; the ensemble mean beats the
    constant predictor by a wide margin on held-out jumps of other replicas."""
    import latticemontecarlo_amd as L
    path, els = coeffs["e0_3"]
    gm = L.Model(path, els, L.E0)
    sur = _trained(L, gm)
    n = 8
    batch, occs, vacs = _batch(L, gm.species, n, list(np.linspace(0.015, 0.045, 60)), seed=5000)
    env, ekra, de = L.jump_samples(gm, batch, vacs)
    pred = sur.predict(env).cpu().numpy().mean(1)
    ek, d = ekra.cpu().numpy(), de.cpu().numpy()
    ea, ea_p = np.maximum(0, ek + d / 2), np.maximum(0, pred[:, 0] + pred[:, 1] / 2)
    mae_ea, mae_d = np.abs(ea_p - ea).mean(), np.abs(pred[:, 1] - d).mean()
    base_ea, base_d = np.abs(ea - np.median(ea)).mean(), np.abs(d - d.mean()).mean()
    print(f"held-out MAE Ea {mae_ea * 1e3:.2f} meV (constant {base_ea * 1e3:.2f}), "
          f"dE {mae_d * 1e3:.2f} meV (constant {base_d * 1e3:.2f})")
    assert mae_ea < 0.3 * base_ea and mae_d < 0.3 * base_d


def _trained(L, gm, n=8, R=300, seed=3000, epochs=30):
    """A surrogate trained on exact samples of KMC trajectories (augmented by the jump
    symmetries), with the samples' trajectory driver."""
    occs = [random_alloc(n, gm.species, 0.015 + 0.03 * (r % 7) / 7, seed=seed + r) for r in range(R)]
    drv = L.KineticMcFirst(gm, None, [n] * R, occs, 500.0, seed, maximum_steps=10 ** 9)
    envs, eks, des = [], [], []
    for _ in range(30):
        drv.run(5)
        e, k, d = drv.samples()
        envs.append(e)
        eks.append(k)
        des.append(d)
    import torch
    env, ek, de = L.augment(torch.cat(envs), torch.cat(eks), torch.cat(des))
    sur = L.Surrogate(gm, n_members=4, seed=9)
    sur.train(env, ek, de, epochs=epochs, batch=2048, lr=3e-3, seed=2)
    return sur


def test_surrogate_eval_matches_ensemble_and_fallback(torch, coeffs):
    import latticemontecarlo_amd as L
    path, els = coeffs["e0_3"]
    gm = L.Model(path, els, L.E0)
    sur = _trained(L, gm)
    n = 8
    batch, occs, vacs = _batch(L, gm.species, n, list(np.linspace(0.02, 0.05, 10)), seed=77)
    temps = np.linspace(450.0, 650.0, 10)
    nb, ea, de, fl = [x.cpu().numpy() for x in sur.eval_jumps(gm, batch, vacs, temps)]
    nb0, ea0, de0 = [x.cpu().numpy() for x in L.eval_jumps(gm, batch, vacs)]
    env, _, _ = L.jump_samples(gm, batch, vacs)
    pred = sur.predict(env).cpu().numpy().reshape(10, 12, 4, 2)
    assert np.array_equal(nb, nb0)
    ek_m, de_m = pred[..., 0].mean(-1), pred[..., 1].mean(-1)
    ea_s = np.maximum(0.0, ek_m + de_m / 2)
    f = fl.astype(bool)
    np.testing.assert_allclose(ea[~f], ea_s[~f], rtol=2e-6, atol=1e-7)
    np.testing.assert_allclose(de[~f], de_m[~f], rtol=2e-6, atol=1e-7)
    np.testing.assert_allclose(ea[f], ea0[f], rtol=0, atol=1e-9)
    np.testing.assert_allclose(de[f], de0[f], rtol=0, atol=1e-9)
    spread = np.maximum(np.ptp(np.maximum(0.0, pred[..., 0] + pred[..., 1] / 2), -1), np.ptp(pred[..., 1], -1))
    # the specified rule (window 0): no served event has members more than 10 meV apart
    assert not np.any(~f & (spread > 0.01 + 1e-6))
    # with a relevance window, disagreeing events above the window may be served
    _, ea_w, _, fl_w = [x.cpu().numpy() for x in sur.eval_jumps(gm, batch, vacs, temps, window=20.0)]
    f_w = fl_w.astype(bool)
    assert np.all(f_w <= f)  # the window only removes fallbacks
    assert not np.any(~f_w & (spread > 0.01 + 1e-6) & (ea_s <= ea_s.min(1, keepdims=True)))
    np.testing.assert_allclose(ea_w[~f_w], ea_s[~f_w], rtol=2e-6, atol=1e-7)
    # every event falls back with a negative threshold
    _, ea2, de2, fl2 = [x.cpu().numpy() for x in sur.eval_jumps(gm, batch, vacs, temps, threshold=-1.0)]
    assert fl2.all()
    np.testing.assert_allclose(ea2, ea0, rtol=0, atol=1e-9)
    np.testing.assert_allclose(de2, de0, rtol=0, atol=1e-9)


def test_surrogate_serves_events_within_the_error_target(torch, coeffs):
    """What the trained network gains over the method without it: on held-out states, the
    specified fallback rule serves a share of the events from the ensemble (their cluster
    expansion is not evaluated) while the barriers of all events, served and exact, stay within
    the 5 meV MAE target of the exact evaluation."""
    import latticemontecarlo_amd as L
    path, els = coeffs["e0_3"]
    gm = L.Model(path, els, L.E0)
    sur = _trained(L, gm, epochs=60)
    n = 8
    batch, occs, vacs = _batch(L, gm.species, n, list(np.linspace(0.015, 0.045, 60)), seed=7000)
    temps = np.linspace(400.0, 700.0, 60)
    _, ea, _, fl = [x.cpu().numpy() for x in sur.eval_jumps(gm, batch, vacs, temps)]
    _, ea0, _ = [x.cpu().numpy() for x in L.eval_jumps(gm, batch, vacs)]
    served = 1.0 - fl.mean()
    mae_all = np.abs(ea - ea0).mean()
    mae_served = np.abs(ea - ea0)[fl == 0].mean()
    print(f"served {served * 100:.1f}% of {fl.size} events; Ea MAE {mae_all * 1e3:.3f} meV over "
          f"all events, {mae_served * 1e3:.3f} meV over served events")
    assert mae_all <= 0.005
    assert served >= 0.15


def test_surrogate_kmc_with_full_fallback_is_the_plain_trajectory(torch, coeffs):
    import latticemontecarlo_amd as L
    path, els = coeffs["e0_3n"]
    gm = L.Model(path, els, L.E0)
    sur = L.Surrogate(gm, n_members=4, seed=1)  # untrained: every event falls back anyway
    n, R, steps = 6, 6, 300
    occs = [random_alloc(n, gm.species, 0.04 + 0.02 * r, seed=600 + r) for r in range(R)]
    temps = np.linspace(450.0, 700.0, R)
    a = L.KineticMcFirst(gm, None, [n] * R, occs, temps, 11, maximum_steps=10 ** 9)
    b = L.KineticMcFirst(gm, None, [n] * R, occs, temps, 11, maximum_steps=10 ** 9)
    b.set_surrogate(sur, threshold=-1.0)
    a.set_trace(steps)
    b.set_trace(steps)
    a.run(steps)
    b.run(steps)
    ta, tb = a.trace(), b.trace()
    assert np.array_equal(ta["sel"], tb["sel"])
    np.testing.assert_allclose(tb["ea"], ta["ea"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(tb["de"], ta["de"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(tb["time"], ta["time"], rtol=1e-12)
    for x, y in zip(a.occupancy(), b.occupancy()):
        assert np.array_equal(x, y)
    st = b.surrogate_stats()
    assert st["events"] == 12 * R * steps and st["fallbacks"] == st["events"]


def test_surrogate_kmc_energy_stays_exact(torch, coeffs):
    import latticemontecarlo_amd as L
    path, els = coeffs["e0_3"]
    gm = L.Model(path, els, L.E0)
    ge = L.Model(path, els, L.ENERGY)
    sur = _trained(L, gm)
    n, R, steps = 8, 8, 400
    occs = [random_alloc(n, gm.species, 0.03, seed=900 + r) for r in range(R)]
    drv = L.KineticMcFirst(gm, ge, [n] * R, occs, 550.0, 5, maximum_steps=10 ** 9)
    drv.set_surrogate(sur)
    drv.run(steps)
    st = drv.state()
    final = drv.occupancy()
    e_final = L.total_energy(ge, L.ReplicaBatch([n] * R, final)).cpu().numpy()
    np.testing.assert_allclose(st["abs_energy"], e_final, rtol=0, atol=1e-9 * steps)
    for o0, o1 in zip(occs, final):
        assert np.array_equal(np.bincount(o0, minlength=4), np.bincount(o1, minlength=4))
    s = drv.surrogate_stats()
    print(f"surrogate KMC: {s['events']} events, {s['fallbacks']} fell back")
    assert s["events"] == 12 * R * steps and s["fallbacks"] < s["events"]
