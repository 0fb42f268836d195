"""lmc_total_energy_noalloc (no reference counterpart; for callers that cannot allocate) through
the C ABI: its results equal lmc_total_energy's bit for bit on two calls in a row, the second
with a smaller batch, and it rejects a model not yet on the device and a batch beyond its fixed
scratch (both before any launch)."""
import ctypes as C

import numpy as np
import pytest

from conftest import random_alloc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def _noalloc(L, torch, model, batch):
    from latticemontecarlo_amd import capi
    out = torch.full((batch.R,), float("nan"), dtype=torch.float64, device=batch.occ.device)
    capi.check(capi.lib().lmc_total_energy_noalloc(
        model.h, C.c_void_p(batch.occ.data_ptr()), C.c_void_p(batch.offset.data_ptr()),
        C.c_void_p(batch.ncell.data_ptr()), batch.R, C.c_void_p(out.data_ptr()), None))
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_total_energy_noalloc_matches_total_energy(torch, coeffs):
    import latticemontecarlo_amd as L
    from latticemontecarlo_amd import capi
    path, els = coeffs["e0_3n"]
    gm = L.Model(path, els, L.ENERGY)
    sizes_a = [4, 6, 5, 8, 7, 4, 9]
    occ_a = [random_alloc(n, gm.species, 0.06 + 0.02 * i, seed=40 + i) for i, n in enumerate(sizes_a)]
    batch_a = L.ReplicaBatch(sizes_a, occ_a)
    print("noalloc: fresh model rejected", flush=True)
    with pytest.raises(capi.LmcError, match="not on this device"):
        _noalloc(L, torch, gm, batch_a)
    ref_a = L.total_energy(gm, batch_a).cpu().numpy()  # uploads the model's tables
    print("noalloc: call 1 (7 replicas)", flush=True)
    got_a = _noalloc(L, torch, gm, batch_a)
    assert np.array_equal(got_a, ref_a)
    sizes_b = [5, 4, 6]
    occ_b = [random_alloc(n, gm.species, 0.15, seed=90 + i) for i, n in enumerate(sizes_b)]
    batch_b = L.ReplicaBatch(sizes_b, occ_b)
    print("noalloc: call 2 (3 replicas)", flush=True)
    got_b = _noalloc(L, torch, gm, batch_b)
    ref_b = L.total_energy(gm, batch_b).cpu().numpy()
    assert np.array_equal(got_b, ref_b)
    assert not np.array_equal(got_b, got_a[:3])  # the scratch was cleared, not reused


def test_total_energy_noalloc_rejects_oversized_batch(torch, coeffs):
    import latticemontecarlo_amd as L
    from latticemontecarlo_amd import capi
    path, els = coeffs["e0_3n"]
    gm = L.Model(path, els, L.ENERGY)
    batch = L.ReplicaBatch([4], [random_alloc(4, gm.species, 0.1, seed=3)])
    L.total_energy(gm, batch)
    torch.cuda.synchronize()
    out = torch.zeros(1, dtype=torch.float64, device=batch.occ.device)
    rc = capi.lib().lmc_total_energy_noalloc(
        gm.h, C.c_void_p(batch.occ.data_ptr()), C.c_void_p(batch.offset.data_ptr()),
        C.c_void_p(batch.ncell.data_ptr()), 5519, C.c_void_p(out.data_ptr()), None)
    assert rc == -1  # LMC_ERR_ARG
    assert "at most 5518 replicas" in capi.lib().lmc_last_error().decode()
