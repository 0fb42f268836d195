"""Pins of the surrogate MLP oracle (CPU): gradients against central finite differences in
plain FP64, bf16 rounding against hand-computed values, and the one-hot input layout."""
import numpy as np

import mlp_oracle as M


def _params(kp, rng, scale=0.3):
    o = M.layout(kp)
    p = rng.uniform(-scale, scale, o["P"])
    p[o["b3"] + 2:] = 0.0
    return p


def test_bf16_rounding_known_values():
    one = 1.0
    ulp = 2.0 ** -7  # bf16 spacing at 1.0
    x = np.array([one, one + ulp / 2, one + 3 * ulp / 2, one + ulp / 4, -2.5, 3.0e-3])
    got = M.bf16(x)
    # ties go to even: 1 + ulp/2 -> 1, 1 + 3 ulp/2 -> 1 + 2 ulp; below half -> down
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 2 * ulp and got[3] == 1.0
    assert got[4] == -2.5
    assert abs(got[5] - 3.0e-3) <= 3.0e-3 * 2.0 ** -8


def test_one_hot_layout():
    env = np.zeros((2, 64), np.uint8)
    env[0, 5] = 1
    env[0, 59] = 3
    env[1, 0] = 3
    x = M.one_hot(env, [1, 3], 128)
    assert x.shape == (2, 128)
    assert x[0].sum() == 2 and x[0, 10] == 1 and x[0, 119] == 1
    assert x[1].sum() == 1 and x[1, 1] == 1
    assert not x[:, 120:].any()


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(3)
    kp = 128
    n = 7
    env = rng.integers(0, 4, size=(n, 64)).astype(np.uint8)
    x = M.one_hot(env, [1, 3], kp)
    t = rng.normal(size=(n, 2))
    p = _params(kp, rng)
    g = M.gradients(p, x, t, rounding=False)
    o = M.layout(kp)
    # every parameter block, including W1 rows that are active in this batch
    active_k = np.flatnonzero(x.any(0))
    picks = [o["W1"] + active_k[0] * 128 + 5, o["W1"] + active_k[-1] * 128 + 77,
             o["b1"] + 3, o["W2"] + 1000, o["W2"] + 16383, o["b2"] + 64,
             o["W3"] + 17, o["W3"] + 254, o["b3"], o["b3"] + 1]
    h = 1e-6
    for i in picks:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss(pp, x, t, False) - M.loss(pm, x, t, False)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])
    # rows of W1 whose inputs are never active get no gradient
    inactive = np.setdiff1d(np.arange(kp), active_k)
    assert not g[o["W1"]:o["b1"]].reshape(kp, 128)[inactive].any()


def test_rounded_gradients_close_to_exact():
    """The bf16 rounding points change the gradients only at the bf16 level."""
    rng = np.random.default_rng(5)
    kp = 192
    env = rng.integers(0, 5, size=(40, 64)).astype(np.uint8)
    x = M.one_hot(env, [0, 1, 4], kp)
    t = rng.normal(size=(40, 2))
    p = _params(kp, rng)
    ge = M.gradients(p, x, t, rounding=False)
    gr = M.gradients(p, x, t, rounding=True)
    assert np.linalg.norm(gr - ge) <= 0.05 * np.linalg.norm(ge)


def _xyz(site, n):
    X, rem = divmod(int(site), 2 * n * n)
    Y, Zh = divmod(rem, n)
    return np.array([X, Y, 2 * Zh + ((X + Y) & 1)])


def _d2(a, b, n):
    d = (a - b) % (2 * n)
    d = np.where(d > n, d - 2 * n, d)
    return int((d * d).sum())


def test_jump_symmetries_are_exact(coeffs):
    """Each slot permutation is an isometry of the environment fixing both endpoints, and the
    oracle's Ea and dE are unchanged when the environment is transformed by it."""
    import latticemontecarlo_amd as L
    import oracle as O
    from conftest import random_alloc
    n = 6
    perms = L.jump_symmetries()
    path, els = coeffs["q_4"]
    lat = O.Lattice(n)
    om = O.Model(lat, path, els, O.QUARTIC)
    species = sorted(els + ["X"])
    occ = random_alloc(n, species, 0.15, seed=21)
    v = int(np.flatnonzero(occ == species.index("X"))[0])
    for nb in lat.neighbors(v, 1)[:4]:
        sites = L.jump_environment(n, v, int(nb))
        pos = [_xyz(s, n) for s in sites]
        pv, pn = _xyz(v, n), _xyz(nb, n)
        ea, de = om.barrier_de(occ, v, int(nb))
        for g in range(4):
            p = perms[g]
            assert p[list(sites).index(v)] == list(sites).index(v)
            assert p[list(sites).index(nb)] == list(sites).index(nb)
            for i in range(60):
                assert _d2(pos[i], pv, n) == _d2(pos[p[i]], pv, n)
                assert _d2(pos[i], pn, n) == _d2(pos[p[i]], pn, n)
                for j in range(i):
                    assert _d2(pos[i], pos[j], n) == _d2(pos[p[i]], pos[p[j]], n)
            occ2 = occ.copy()
            occ2[sites] = occ[sites[p]]
            ea2, de2 = om.barrier_de(occ2, v, int(nb))
            assert abs(ea2 - ea) <= 1e-9 and abs(de2 - de) <= 1e-9
