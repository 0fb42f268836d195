"""Argument checks of the batch drivers (lmc_kmc_create, lmc_cmc_create): empty and ragged
batches, box sizes out of range, species codes, vacancy counts and parameters. Every one fails
on the host before any device memory is touched, so these run without a GPU."""
import ctypes as C

import numpy as np
import pytest

from conftest import random_alloc

import latticemontecarlo_amd as L
from latticemontecarlo_amd import capi


@pytest.fixture(scope="module")
def models(coeffs):
    path, els = coeffs["e0_3"]
    return L.Model(path, els, L.E0), L.Model(path, els, L.PAIRSITE)


def _kmc(model, ncell, occs, **kw):
    return L.KineticMcFirst(model, None, ncell, occs, kw.pop("T", 500.0), 1, maximum_steps=10,
                            **kw)


def test_kmc_rejects_bad_batches(models):
    gb, _ = models
    sp = gb.species
    good = random_alloc(4, sp, 0.05, seed=1)
    h = C.c_void_p()
    params = (capi.KmcReplica * 1)()
    nc = np.array([4], np.int32)
    with pytest.raises(capi.LmcError, match="null argument"):
        capi.check(capi.lib().lmc_kmc_create(gb.h, None, 1, None, None, params, C.byref(h)))
    with pytest.raises(capi.LmcError, match="n_replicas must be positive"):  # an empty batch
        capi.check(capi.lib().lmc_kmc_create(gb.h, None, 0, capi.ptr(nc, capi._i32p),
                                             capi.ptr(good, capi._u8p), params, C.byref(h)))
    params2 = (capi.KmcReplica * 2)()
    for n in (3, 1001):  # the C ABI checks box sizes before it reads any occupancy
        nc2 = np.array([4, n], np.int32)
        with pytest.raises(capi.LmcError, match=r"lattice size n must be in \[4, 1000\]"):
            capi.check(capi.lib().lmc_kmc_create(gb.h, None, 2, capi.ptr(nc2, capi._i32p),
                                                 capi.ptr(good, capi._u8p), params2,
                                                 C.byref(h)))
        with pytest.raises(ValueError, match="replica 1"):  # the Python mirror checks sizes
            _kmc(gb, [4, n], [good, good])
    bad = good.copy()
    bad[7] = len(sp)
    with pytest.raises(capi.LmcError, match="species code >= S"):
        _kmc(gb, [4], [bad])
    none = np.where(good == sp.index("X"), sp.index("Al"), good).astype(np.uint8)
    with pytest.raises(capi.LmcError, match="exactly one vacancy"):
        _kmc(gb, [4], [none])
    two = good.copy()
    two[np.flatnonzero(good != sp.index("X"))[0]] = sp.index("X")
    with pytest.raises(capi.LmcError, match="exactly one vacancy"):
        _kmc(gb, [4, 4], [good, two])  # the ragged second replica is the bad one
    with pytest.raises(capi.LmcError, match="temperature must be positive"):
        _kmc(gb, [4], [good], T=0.0)


def test_cmc_rejects_bad_batches(models):
    _, gp = models
    sp = gp.species
    good = random_alloc(4, sp, 0.1, seed=2, vacancy=False)
    one = np.full_like(good, sp.index("Al"))
    with pytest.raises(capi.LmcError, match="at least two species"):
        L.CanonicalMc(gp, None, [4, 4], [good, one], 500.0, 500.0, 1, maximum_steps=10)
    with pytest.raises(capi.LmcError, match="must be positive"):
        L.CanonicalMc(gp, None, [4], [good], 500.0, 500.0, 1, maximum_steps=0)
    with pytest.raises(capi.LmcError, match="must be positive"):
        L.CanonicalMc(gp, None, [4], [good], -1.0, 500.0, 1, maximum_steps=10)


def test_kmc_needs_a_barrier_model(models, coeffs):
    path, els = coeffs["e0_3"]
    ge = L.Model(path, els, L.ENERGY)
    occ = random_alloc(4, ge.species, 0.05, seed=3)
    with pytest.raises(capi.LmcError, match="E0 or Quartic"):
        _kmc(ge, [4], [occ])


def test_log_entry_points_reject_null_handles():
    """The log entry points of both drivers check their handles before any HIP call."""
    lib = capi.lib()
    names = (C.c_char_p * 1)(b"log.txt")
    calls = [lambda: lib.lmc_cmc_set_log(None, 10, 10, 0),
             lambda: lib.lmc_cmc_write_log(None, names),
             lambda: lib.lmc_kmc_set_log(None, 10, 10, None),
             lambda: lib.lmc_kmc_write_log(None, names)]
    for call in calls:
        with pytest.raises(capi.LmcError, match="null (driver|argument)"):
            capi.check(call())


def test_device_buffer_entry_points_reject_null_buffers(models):
    """Null device buffers with a positive count, and negative counts, are argument errors
    raised on the host, before any device call (a null pointer would otherwise reach a
    kernel)."""
    gb, ge = models
    lib = capi.lib()
    x = C.c_void_p(1)  # never dereferenced: every call below fails its argument check first
    cases = [
        ("lmc_eval_jumps", lambda: lib.lmc_eval_jumps(gb.h, None, x, x, x, 1, x, x, x, None)),
        ("lmc_eval_jumps", lambda: lib.lmc_eval_jumps(gb.h, x, x, x, x, 1, x, x, None, None)),
        ("lmc_swap_de", lambda: lib.lmc_swap_de(ge.h, x, x, x, None, 1, x, None)),
        ("lmc_site_de", lambda: lib.lmc_site_de(ge.h, x, x, x, x, None, 1, x, None)),
        ("lmc_site_change_de", lambda: lib.lmc_site_change_de(ge.h, x, x, x, x, None, x, 1, x,
                                                              None)),
        ("lmc_total_energy", lambda: lib.lmc_total_energy(ge.h, x, None, x, 2, x, None)),
        ("lmc_total_encode", lambda: lib.lmc_total_encode(ge.h, None, x, x, 1, x, None)),
        ("lmc_jump_samples", lambda: lib.lmc_jump_samples(gb.h, x, x, x, None, 1, x, x, x, None)),
        ("lmc_eval_pair", lambda: lib.lmc_eval_pair(gb.h, None, 6, 0, 6, None, None)),
    ]
    for name, call in cases:
        with pytest.raises(capi.LmcError, match=f"{name}: null buffer"):
            capi.check(call())
    with pytest.raises(capi.LmcError, match="lmc_swap_de: negative count"):
        capi.check(lib.lmc_swap_de(ge.h, x, x, x, x, -1, x, None))
    with pytest.raises(capi.LmcError, match="lmc_eval_jumps: negative count"):
        capi.check(lib.lmc_eval_jumps(gb.h, x, x, x, x, -3, x, x, x, None))
    ea, de = C.c_double(), C.c_double()
    with pytest.raises(capi.LmcError, match="ncell must be >= 4"):
        capi.check(lib.lmc_eval_pair(gb.h, x, 0, 0, 6, C.byref(ea), C.byref(de)))
    with pytest.raises(capi.LmcError, match="vacancy site out of range"):
        capi.check(lib.lmc_eval_pair(gb.h, x, 6, 10 ** 6, 6, C.byref(ea), C.byref(de)))
    # an empty call is not an error, and launches nothing
    capi.check(lib.lmc_swap_de(ge.h, None, None, None, None, 0, None, None))


def test_python_wrappers_check_shapes_before_any_launch(models):
    """The Python mirror refuses per-replica arrays of the wrong length and occupancies of the
    wrong size before anything reaches the C ABI (the kernels read exactly R entries)."""
    import types

    import torch

    gb, ge = models
    good = random_alloc(4, gb.species, 0.05, seed=3)
    with pytest.raises(ValueError, match="2 box sizes for 1 occupancies"):
        L.ReplicaBatch([4, 4], [good])
    with pytest.raises(ValueError, match="replica 1: 256 sites, expected 500"):
        L.ReplicaBatch([4, 5], [good, good])
    fake = types.SimpleNamespace(R=3, occ=torch.zeros(1, dtype=torch.uint8), offset=None,
                                 ncell=None)
    with pytest.raises(ValueError, match="vac: expected 3 entries for 3 replicas, got 2"):
        L.eval_jumps(gb, fake, [1, 2])
    with pytest.raises(ValueError, match="vac: expected 3"):
        L.jump_samples(gb, fake, [1, 2, 3, 4])
    with pytest.raises(ValueError, match="pairs: expected 3 x 2 entries"):
        L.swap_de(ge, fake, [[1, 2], [3, 4]])
    with pytest.raises(ValueError, match="codes: expected 3 entries"):
        L.site_de(ge, fake, [1, 2, 3], [0, 1])
