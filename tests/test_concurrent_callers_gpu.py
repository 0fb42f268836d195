"""Host threads sharing one model, as the reference's KineticMcFirstOmp shares its const predictor
among OpenMP threads (KineticMcFirstOmp.cpp:60-64): the first calls upload the model's tables
concurrently, and every result must equal the serial one bit for bit and the oracle's within
1e-9 eV."""
import ctypes as C
import threading

import numpy as np
import pytest

from conftest import TERNARY, random_alloc

pytestmark = pytest.mark.gpu

TOL_E = 1e-9


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


@pytest.mark.parametrize("name,kind", [("e0_3", 1), ("q_3n", 2)])
def test_threads_share_a_model(torch, coeffs, name, kind):
    import latticemontecarlo_amd as L
    import oracle as O
    from latticemontecarlo_amd import capi

    lib = capi.lib()
    path, els = coeffs[name]
    n = 6
    gm = L.Model(path, els, kind)  # not uploaded yet: the threads' first calls race to do it
    occ = random_alloc(n, gm.species, 0.1, seed=77)
    v = int(np.flatnonzero(occ == gm.species.index("X"))[0])
    d_occ = torch.from_numpy(occ).cuda()
    lat = O.Lattice(n)
    nbs = [int(b) for b in lat.neighbors(v, 1)]
    n_threads, rounds = 8, 4
    start = threading.Barrier(n_threads)
    got = [[None] * len(nbs) for _ in range(n_threads)]
    errors = []

    def worker(w):
        try:
            start.wait()
            for _ in range(rounds):
                for i in range(len(nbs)):
                    j = (i + w) % len(nbs)  # threads ask for the jumps in different orders
                    ea, de = C.c_double(), C.c_double()
                    rc = lib.lmc_eval_pair(gm.h, C.c_void_p(d_occ.data_ptr()), n, v, nbs[j],
                                           C.byref(ea), C.byref(de))
                    if rc != 0:
                        raise RuntimeError(lib.lmc_last_error().decode())
                    if got[w][j] is None:
                        got[w][j] = (ea.value, de.value)
                    elif got[w][j] != (ea.value, de.value):
                        raise AssertionError(f"thread {w}: jump {j} changed between rounds")
        except BaseException as e:  # reported by the main thread
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(w,)) for w in range(n_threads)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads), "a caller thread did not finish"
    assert not errors, errors[0]
    serial = []
    for b in nbs:
        ea, de = C.c_double(), C.c_double()
        assert lib.lmc_eval_pair(gm.h, C.c_void_p(d_occ.data_ptr()), n, v, b, C.byref(ea),
                                 C.byref(de)) == 0
        serial.append((ea.value, de.value))
    for w in range(n_threads):
        assert got[w] == serial, f"thread {w} differs from the serial results"
    om = O.Model(lat, path, els, kind)
    for b, (ea, de) in zip(nbs, serial):
        e_ref, d_ref = om.barrier_de(occ, v, b)
        assert abs(ea - e_ref) <= TOL_E and abs(de - d_ref) <= TOL_E, (b, ea, e_ref, de, d_ref)
