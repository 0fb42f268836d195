"""The largest box the engine takes (n = 1000 cells per edge, 4e9 sites, ids above 2^31): a KMC
trajectory and the jump evaluations next to the periodic boundary equal, bit for bit, those of
the same local configuration in a 24^3-cell box (translation invariance of the cluster
expansion)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _site(n, X, Y, Z):
    m = 2 * n
    X, Y, Z = X % m, Y % m, Z % m
    return X * 2 * n * n + Y * n + Z // 2


def _xyz(n, s):
    X = s // (2 * n * n)
    r = s - X * 2 * n * n
    Y = r // n
    Z = 2 * (r - Y * n) + ((X + Y) & 1)
    m = 2 * n
    return tuple(int(((c + n) % m) - n) for c in (X, Y, Z))  # minimum image around the origin


def _pattern(species, seed):
    """Solute positions (a/2 grid offsets within 6 of the origin) and species."""
    rng = np.random.default_rng(seed)
    pts = {}
    while len(pts) < 40:
        x, y, z = rng.integers(-6, 7, 3)
        if (x + y + z) % 2 or (x, y, z) == (0, 0, 0):
            continue
        pts[(int(x), int(y), int(z))] = species.index(rng.choice(["Mg", "Zn"]))
    return pts


def _occupancy(n, species, pts):
    occ = np.full(4 * n ** 3, species.index("Al"), np.uint8)
    for (x, y, z), c in pts.items():
        occ[_site(n, x, y, z)] = c
    occ[_site(n, 0, 0, 0)] = species.index("X")
    return occ


def test_largest_box_matches_small_box(torch, coeffs):
    import latticemontecarlo_amd as L
    path, els = coeffs["e0_3"]
    gb = L.Model(path, els, L.E0)
    sp = gb.species
    pts = _pattern(sp, seed=5)
    steps = 60
    runs = {}
    for n in (1000, 24):
        print(f"largest-box KMC n={n} ({4 * n ** 3} sites) steps={steps}", flush=True)
        occ = _occupancy(n, sp, pts)
        # the 12 jumps of the vacancy at the origin: ids wrap across every face of the box
        nb, ea, de = L.eval_jumps(gb, L.ReplicaBatch([n], [occ]), [0])
        nbx = [_xyz(n, int(s)) for s in nb.cpu().numpy()[0]]
        drv = L.KineticMcFirst(gb, None, [n], [occ], 600.0, 7, maximum_steps=10 ** 9)
        del occ
        drv.set_trace(steps)
        drv.run(steps)
        tr, st = drv.trace(), drv.state()
        runs[n] = dict(nbx=nbx, ea=ea.cpu().numpy()[0], de=de.cpu().numpy()[0],
                       path=[_xyz(n, int(s)) for s in tr["sel"][0]], time=tr["time"][0],
                       vac=_xyz(n, int(st["vacancy"][0])), big=int(nb.cpu().numpy().max()))
        del drv
        torch.cuda.empty_cache()
    a, b = runs[1000], runs[24]
    assert a["big"] > 2 ** 31  # neighbour ids past the int32 range
    # the 12 events in id order differ between the boxes; compare them by position
    ia = sorted(range(12), key=lambda i: a["nbx"][i])
    ib = sorted(range(12), key=lambda i: b["nbx"][i])
    assert [a["nbx"][i] for i in ia] == [b["nbx"][i] for i in ib]
    assert np.array_equal(a["ea"][ia], b["ea"][ib]) and np.array_equal(a["de"][ia], b["de"][ib])
    assert a["path"] == b["path"] and a["vac"] == b["vac"]
    assert max(max(abs(c) for c in p) for p in a["path"]) < 20  # stayed inside the pattern's box
    assert np.array_equal(a["time"], b["time"])


def _philox(seed, counter, replica, tag):
    import ctypes as C
    from latticemontecarlo_amd.capi import lib
    out = (C.c_uint32 * 4)()
    lib().lmc_philox4x32(seed, counter, replica, tag, out)
    return list(out)


def test_largest_box_canonical_trial(torch, coeffs):
    """CanonicalMc on the 4e9-site box: the first trial's pair is the first Philox attempt (counter
    2 att + 1, site = mulhi64(bits, N)) whose sites differ, as restated on the host, and its dE is
    swap_de's on the same occupancy."""
    import latticemontecarlo_amd as L
    path, els = coeffs["e0_3n"]
    gp = L.Model(path, els, L.PAIRSITE)
    sp = gp.species
    n = 1000
    N = 4 * n ** 3
    print(f"largest-box CMC n={n} ({N} sites)", flush=True)
    rng = np.random.default_rng(3)
    lut = np.array([sp.index("Al")] * 8 + [sp.index("Mg"), sp.index("Zn")], np.uint8)
    occ = lut[rng.integers(0, 10, size=N, dtype=np.uint8)]  # Al 80 %, Mg 10 %, Zn 10 %
    seed, rid = 11, 0
    att = 0
    while True:  # the reference's redraw loop, restated on the host
        u = _philox(seed, 0, rid, 2 * att + 1)
        p = ((u[1] << 32 | u[0]) * N) >> 64
        q = ((u[3] << 32 | u[2]) * N) >> 64
        if occ[p] != occ[q]:
            break
        att += 1
    drv = L.CanonicalMc(gp, None, [n], [occ], 900.0, 900.0, seed, maximum_steps=10)
    drv.set_trace(1)
    drv.run(1)
    tr = drv.trace()
    assert (int(tr["p"][0, 0]), int(tr["q"][0, 0])) == (p, q)
    del drv
    de = L.swap_de(gp, L.ReplicaBatch([n], [occ]), [p, q]).cpu().numpy()[0]
    assert tr["de"][0, 0] == de
