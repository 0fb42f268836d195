"""Python mirror of the reference's predictor and driver interface over the C ABI.

Names follow zhucongx/LatticeMonteCarlo (lmc/pred/include/*.h, lmc/mc/include/*.h). Device
buffers are torch CUDA tensors; every compute call goes to a HIP kernel through liblmc.so.
"""
import ctypes as C

import numpy as np

from . import capi
from .capi import check, dptr, lib, ptr

ENERGY, E0, QUARTIC, PAIRSITE = capi.ENERGY, capi.E0, capi.QUARTIC, capi.PAIRSITE


def _torch():
    import torch
    return torch


class Model:
    """Coefficient model of one predictor kind (see lmc_model_load)."""

    def __init__(self, json_path, element_set, kind):
        els = sorted(set(element_set) - {"X"})
        arr = (C.c_char_p * len(els))(*[e.encode() for e in els])
        h = C.c_void_p()
        check(lib().lmc_model_load(str(json_path).encode(), arr, len(els), kind, C.byref(h)))
        self.h = h
        self.kind = kind
        n = lib().lmc_model_num_species(self.h)
        self.species = [lib().lmc_model_species_name(self.h, c).decode() for c in range(n)]
        self.element_set = [e for e in self.species if e != "X"]

    def __del__(self):
        if getattr(self, "h", None):
            try:
                lib().lmc_model_free(self.h)
            except Exception:  # interpreter shutdown
                pass
            self.h = None

    def code(self, element):
        return lib().lmc_model_species_code(self.h, element.encode())

    def layout(self):
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().lmc_model_layout(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def eci_keys(self):
        n_eci = self.layout()[0]
        out = []
        lab = C.c_int32()
        codes = np.zeros(3, np.int32)
        for i in range(n_eci):
            check(lib().lmc_model_eci_key(self.h, i, C.byref(lab), ptr(codes, capi._i32p)))
            out.append((lab.value, tuple(int(c) for c in codes if c >= 0)))
        return out

    def orbits(self, which):
        cap = 512
        arrs = [np.zeros(cap, np.int32) for _ in range(5)]
        n = C.c_int32()
        check(lib().lmc_model_orbits(self.h, which, cap, C.byref(n),
                                     *[ptr(a, capi._i32p) for a in arrs]))
        k = n.value
        return dict(size=arrs[0][:k], kind=arrs[1][:k], shell=arrs[2][:k], base=arrs[3][:k],
                    len=arrs[4][:k])

    def check_errors(self):
        check(lib().lmc_check_errors(self.h))

    def set_formulation(self, formulation):
        """0 = dense, 1 = solvent baseline (default)."""
        check(lib().lmc_model_set_formulation(self.h, formulation))


def fcc_num_sites(n):
    return int(lib().lmc_fcc_num_sites(n))


def fcc_neighbors(n, site, shell):
    out = np.zeros(24, np.int64)
    k = lib().lmc_fcc_neighbors(n, site, shell, ptr(out, capi._i64p))
    if k < 0:
        check(k)
    return out[:k].copy()


def jump_state_clusters(n, v, nb):
    out = np.zeros((1024, 4), np.int64)
    k = lib().lmc_jump_state_clusters(n, v, nb, ptr(out, capi._i64p), 1024)
    if k < 0:
        check(k)
    return out[:k].copy()


def jump_orbit_clusters(n, v, nb, which):
    out = np.zeros((1024, 4), np.int64)
    k = lib().lmc_jump_orbit_clusters(n, v, nb, which, ptr(out, capi._i64p), 1024)
    if k < 0:
        check(k)
    return out[:k].copy()


def jump_environment(n, v, nb):
    """Lattice ids of the 60 environment slots of the jump v -> nb (surrogate input order)."""
    out = np.zeros(60, np.int64)
    k = lib().lmc_jump_environment(n, v, nb, ptr(out, capi._i64p))
    if k < 0:
        check(k)
    return out


def jump_symmetries():
    """Slot permutations [4, 60] of the jump symmetries that fix both endpoints."""
    out = np.zeros((4, 60), np.int32)
    check(lib().lmc_jump_symmetries(ptr(out, capi._i32p)))
    return out


def augment(env, ekra, de):
    """Training rows under the four endpoint-fixing symmetries (exact: E_kra and dE are
    invariant); torch tensors on any device."""
    torch = _torch()
    perm = torch.from_numpy(jump_symmetries().astype(np.int64)).to(env.device)
    rows = []
    for g in range(4):
        e = env.clone()
        e[:, :60] = env[:, :60][:, perm[g]]
        rows.append(e)
    return torch.cat(rows), ekra.repeat(4), de.repeat(4)


def _per_replica(values, R, what, dtype, width=None):
    """values as a flat array of R (or R x width) entries; a ValueError otherwise, since the
    kernels read exactly that many from the device copy."""
    a = np.asarray(values, dtype)
    want = R if width is None else R * width
    if a.size != want:
        shape = f"{R}" if width is None else f"{R} x {width}"
        raise ValueError(f"{what}: expected {shape} entries for {R} replicas, got {a.size}")
    return a.reshape(-1)


def _device_rows(t, what, width, dtype):
    """A contiguous device tensor of rows of `width` entries of torch dtype `dtype`."""
    if t.dim() != 2 or t.shape[1] != width or t.dtype != dtype or not t.is_contiguous():
        raise ValueError(f"{what}: expected a contiguous {dtype} tensor [n, {width}], got "
                         f"{t.dtype} {list(t.shape)}")
    return t


def _concat_occupancies(ncell, occ_list):
    """(ncell int32 [R], the occupancies back to back as contiguous uint8, per-replica sizes);
    a ValueError unless replica r holds 4 ncell[r]^3 sites, since the engine reads exactly
    that many bytes per replica."""
    nc = np.asarray(ncell, np.int32).reshape(-1)
    occ_list = list(occ_list)
    if len(occ_list) != nc.size:
        raise ValueError(f"{nc.size} box sizes for {len(occ_list)} occupancies")
    sizes = [4 * int(n) ** 3 for n in nc]
    arrays = [np.asarray(o, np.uint8).reshape(-1) for o in occ_list]
    for r, (a, N) in enumerate(zip(arrays, sizes)):
        if a.size != N:
            raise ValueError(f"replica {r}: {a.size} sites, expected {N}")
    occ = np.ascontiguousarray(np.concatenate(arrays)) if arrays else np.zeros(0, np.uint8)
    return nc, occ, sizes


class ReplicaBatch:
    """Occupancies of R replicas on the GPU (concatenated uint8 with offsets)."""

    def __init__(self, ncell, occ_list, device="cuda"):
        torch = _torch()
        self.ncell_host, host, sizes = _concat_occupancies(ncell, occ_list)
        self.offset_host = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        self.occ = torch.from_numpy(host).to(device)
        self.offset = torch.from_numpy(self.offset_host).to(device)
        self.ncell = torch.from_numpy(self.ncell_host).to(device)
        self.R = len(sizes)


def eval_jumps(model, batch, vac, stream=None):
    """12 jumps per replica in neighbour-id order: (nb[R,12], Ea[R,12], dE[R,12])."""
    torch = _torch()
    R = batch.R
    vac_t = torch.as_tensor(_per_replica(vac, R, "vac", np.int64)).to(batch.occ.device)
    nb = torch.empty((R, 12), dtype=torch.int64, device=batch.occ.device)
    ea = torch.empty((R, 12), dtype=torch.float64, device=batch.occ.device)
    de = torch.empty((R, 12), dtype=torch.float64, device=batch.occ.device)
    check(lib().lmc_eval_jumps(model.h, dptr(batch.occ), dptr(batch.offset), dptr(batch.ncell),
                               dptr(vac_t), R, dptr(nb), dptr(ea), dptr(de),
                               capi.stream_ptr(stream)))
    return nb, ea, de


def vacancy_scan(model, occ, n, vacancy, solvent, sites, stream=None):
    """The 12 jumps (neighbour-id order) of a vacancy placed at each of `sites` in the
    configuration `occ` (uint8 device tensor, one replica of n^3 cells) with its vacancy site
    `vacancy` (-1: none) holding species `solvent` instead: (nb, Ea, dE) [K, 12]
    (ExitTime::GetJumpEnergetics, lmc/ansys/src/ExitTime.cpp:272-311)."""
    torch = _torch()
    dev = occ.device
    s = torch.as_tensor(np.asarray(sites, np.int64).reshape(-1)).to(dev)
    K = int(s.numel())
    nb = torch.empty((K, 12), dtype=torch.int64, device=dev)
    ea = torch.empty((K, 12), dtype=torch.float64, device=dev)
    de = torch.empty((K, 12), dtype=torch.float64, device=dev)
    check(lib().lmc_vacancy_scan(model.h, dptr(occ), int(n), int(vacancy), int(solvent), dptr(s), K,
                                 dptr(nb), dptr(ea), dptr(de), capi.stream_ptr(stream)))
    return nb, ea, de


def swap_de(model, batch, pairs, stream=None):
    torch = _torch()
    p = torch.as_tensor(_per_replica(pairs, batch.R, "pairs", np.int64, 2)).to(batch.occ.device)
    de = torch.empty(batch.R, dtype=torch.float64, device=batch.occ.device)
    check(lib().lmc_swap_de(model.h, dptr(batch.occ), dptr(batch.offset), dptr(batch.ncell),
                            dptr(p), batch.R, dptr(de), capi.stream_ptr(stream)))
    return de


def site_de(model, batch, sites, codes, stream=None):
    torch = _torch()
    s = torch.as_tensor(_per_replica(sites, batch.R, "sites", np.int64)).to(batch.occ.device)
    c = torch.as_tensor(_per_replica(codes, batch.R, "codes", np.uint8)).to(batch.occ.device)
    de = torch.empty(batch.R, dtype=torch.float64, device=batch.occ.device)
    check(lib().lmc_site_de(model.h, dptr(batch.occ), dptr(batch.offset), dptr(batch.ncell),
                            dptr(s), dptr(c), batch.R, dptr(de), capi.stream_ptr(stream)))
    return de


def total_energy(model, batch, stream=None):
    torch = _torch()
    e = torch.empty(batch.R, dtype=torch.float64, device=batch.occ.device)
    check(lib().lmc_total_energy(model.h, dptr(batch.occ), dptr(batch.offset),
                                 dptr(batch.ncell), batch.R, dptr(e), capi.stream_ptr(stream)))
    return e


def total_encode(model, batch, stream=None):
    """EnergyPredictor::GetEncode of every replica: [R, n_eci] float64 device tensor (ECI key
    order, counts divided by the total-energy normalisers)."""
    torch = _torch()
    n_eci = C.c_int32()
    check(lib().lmc_model_layout(model.h, C.byref(n_eci), None, None))
    out = torch.empty((batch.R, n_eci.value), dtype=torch.float64, device=batch.occ.device)
    check(lib().lmc_total_encode(model.h, dptr(batch.occ), dptr(batch.offset), dptr(batch.ncell),
                                 batch.R, dptr(out), capi.stream_ptr(stream)))
    return out


def cluster_energy(model, occ, n, sites):
    """EnergyPredictor::GetEnergyOfCluster / GetEncodeOfCluster of lattice ids `sites` in one
    n^3-cell occupancy (uint8 device tensor): (energy, encode [n_eci])."""
    torch = _torch()
    occ = torch.as_tensor(occ).to(dtype=torch.uint8, device="cuda").contiguous()
    s = np.ascontiguousarray(np.asarray(sites, np.int64).ravel())
    n_eci = C.c_int32()
    check(lib().lmc_model_layout(model.h, C.byref(n_eci), None, None))
    enc = np.zeros(n_eci.value)
    e = C.c_double()
    check(lib().lmc_cluster_energy(model.h, C.c_void_p(occ.data_ptr()), int(n),
                                   ptr(s, capi._i64p), s.size, ptr(enc, capi._f64p), C.byref(e)))
    return e.value, enc


def chemical_potential(model, solvent):
    """EnergyPredictor::GetChemicalPotential(solvent): {element: mu} for every species of the
    model (X included; 0 for the solvent), from 15^3-cell boxes on the GPU."""
    mu = np.zeros(len(model.species))
    check(lib().lmc_chemical_potential(model.h, solvent.encode(), ptr(mu, capi._f64p)))
    return {e: float(mu[i]) for i, e in enumerate(model.species)}


def _set_atoms(fn, h, ncell, lattice_to_atom, map_shift):
    """Atom tracking of a driver (lmc_kmc_set_atoms / lmc_cmc_set_atoms): per-replica lists of
    the atom on every site (a permutation of that replica's atom ids) and map_shift [N, 3] by
    atom (None: zero); lattice_to_atom None turns tracking off."""
    if lattice_to_atom is None:
        check(fn(h, None, None))
        return
    sizes = 4 * np.asarray(ncell, np.int64) ** 3
    l2a = np.ascontiguousarray(np.concatenate([np.asarray(a, np.int64) for a in lattice_to_atom]))
    if l2a.size != sizes.sum():
        raise ValueError(f"lattice_to_atom has {l2a.size} entries, expected {int(sizes.sum())}")
    sh = None
    if map_shift is not None:
        sh = np.ascontiguousarray(np.concatenate([np.asarray(m, np.int32).reshape(-1, 3)
                                                  for m in map_shift]))
        if sh.shape[0] != l2a.size:
            raise ValueError("map_shift has the wrong size")
    check(fn(h, ptr(l2a, capi._i64p), None if sh is None else ptr(sh, capi._i32p)))


def _atoms(fn, h, ncell):
    sizes = 4 * np.asarray(ncell, np.int64) ** 3
    l2a = np.zeros(int(sizes.sum()), np.int64)
    sh = np.zeros((int(sizes.sum()), 3), np.int32)
    check(fn(h, ptr(l2a, capi._i64p), ptr(sh, capi._i32p)))
    cut = np.cumsum(sizes)[:-1]
    return np.split(l2a, cut), np.split(sh, cut)


class KineticMcFirst:
    """Batched KineticMcFirst driver: R independent replicas advanced on the GPU."""

    def __init__(self, barrier_model, energy_model, ncell, occ_list, temperature, seed,
                 maximum_steps, replica_ids=None, restart_steps=0, restart_energy=0.0,
                 restart_time=0.0, rate_corrector=False, early_stop=False, solute_disp=False):
        R = len(ncell)
        self.R = R
        self.ncell = np.asarray(ncell, np.int32)
        temps = np.broadcast_to(np.asarray(temperature, np.float64), (R,))
        rid = np.arange(R) if replica_ids is None else np.asarray(replica_ids)
        params = (capi.KmcReplica * R)()
        for r in range(R):
            p = params[r]
            p.temperature = float(temps[r])
            p.seed = int(seed)
            p.replica_id = int(rid[r])
            p.restart_steps = int(np.broadcast_to(restart_steps, (R,))[r])
            p.restart_energy = float(np.broadcast_to(restart_energy, (R,))[r])
            p.restart_time = float(np.broadcast_to(restart_time, (R,))[r])
            p.maximum_steps = int(np.broadcast_to(maximum_steps, (R,))[r])
            p.rate_corrector = int(np.broadcast_to(rate_corrector, (R,))[r])
            p.early_stop = int(np.broadcast_to(early_stop, (R,))[r])
            p.solute_disp = int(np.broadcast_to(solute_disp, (R,))[r])
        _, occ, _ = _concat_occupancies(self.ncell, occ_list)
        self.total_sites = occ.size
        h = C.c_void_p()
        check(lib().lmc_kmc_create(barrier_model.h, energy_model.h if energy_model else None, R,
                                   ptr(self.ncell, capi._i32p), ptr(occ, capi._u8p), params,
                                   C.byref(h)))
        self.h = h
        self.trace_cap = 0
        self._models = (barrier_model, energy_model)

    def __del__(self):
        if getattr(self, "h", None):
            try:
                lib().lmc_kmc_free(self.h)
            except Exception:  # interpreter shutdown
                pass
            self.h = None

    def set_schedule(self, times, temps):
        t = np.ascontiguousarray(times, np.float64)
        T = np.ascontiguousarray(temps, np.float64)
        check(lib().lmc_kmc_set_schedule(self.h, len(t), ptr(t, capi._f64p), ptr(T, capi._f64p)))

    def set_trace(self, capacity):
        check(lib().lmc_kmc_set_trace(self.h, capacity))
        self.trace_cap = capacity

    def set_formulation(self, formulation):
        check(lib().lmc_kmc_set_formulation(self.h, formulation))

    def autotune(self, steps=200):
        """Times each kernel shape for `steps` steps on a scratch copy and keeps the fastest:
        formulation 0 dense, 1 solvent baseline, 2 dense wide (tried only with at most one
        replica per CU); ms_* per batch step (NaN when not tried)."""
        ch, ms = C.c_int32(), (C.c_double * 3)()
        check(lib().lmc_kmc_autotune_shapes(self.h, steps, C.byref(ch), ms))
        return dict(formulation=ch.value, ms_dense=ms[0], ms_solvent=ms[1], ms_wide=ms[2])

    def run(self, steps, stream=None):
        s = capi.stream_ptr(stream) if stream is not None else None
        check(lib().lmc_kmc_run(self.h, steps, s))

    def build_graph(self, steps_per_launch, launches):
        check(lib().lmc_kmc_build_graph(self.h, steps_per_launch, launches))

    def run_graph(self, steps_per_launch, launches, stream=None):
        """`launches` x `steps_per_launch` steps replayed from a captured HIP graph."""
        s = capi.stream_ptr(stream) if stream is not None else None
        check(lib().lmc_kmc_run_graph(self.h, steps_per_launch, launches, s))

    def state(self):
        R = self.R
        out = dict(steps=np.zeros(R, np.uint64), time=np.zeros(R), energy=np.zeros(R),
                   abs_energy=np.zeros(R), temperature=np.zeros(R), vacancy=np.zeros(R, np.int64),
                   disp=np.zeros((R, 3), np.int64), com=np.zeros((R, 3)),
                   escaped=np.zeros(R, np.int32))
        check(lib().lmc_kmc_state(self.h, ptr(out["steps"], capi._u64p),
                                  ptr(out["time"], capi._f64p), ptr(out["energy"], capi._f64p),
                                  ptr(out["abs_energy"], capi._f64p),
                                  ptr(out["temperature"], capi._f64p),
                                  ptr(out["vacancy"], capi._i64p), ptr(out["disp"], capi._i64p),
                                  ptr(out["com"], capi._f64p), ptr(out["escaped"], capi._i32p)))
        return out

    def escape_info(self):
        """Per replica: the time and step of the escape check that stopped it, and the vacancy's
        site before that step's jump (-1 where no escape happened)."""
        R = self.R
        out = dict(t_exit=np.zeros(R), step=np.zeros(R, np.uint64),
                   site_before=np.zeros(R, np.int64))
        check(lib().lmc_kmc_escape_info(self.h, ptr(out["t_exit"], capi._f64p),
                                        ptr(out["step"], capi._u64p),
                                        ptr(out["site_before"], capi._i64p)))
        return out

    def set_atoms(self, lattice_to_atom, map_shift=None):
        """Track atom ids and map_shift through the jumps (Config::LatticeJump); per-replica
        lists as cfg.Config.atoms() returns them. None turns tracking off."""
        _set_atoms(lib().lmc_kmc_set_atoms, self.h, self.ncell, lattice_to_atom, map_shift)

    def atoms(self):
        """(per-replica lattice_to_atom, per-replica map_shift [N, 3])."""
        return _atoms(lib().lmc_kmc_atoms, self.h, self.ncell)

    def set_log(self, log_dump_steps, capacity=100000, box_edge=None):
        """Keep kmc_log.txt records at the reference's dump cadence (0 turns it off); box_edge:
        per-replica cubic box edge in A (default 4.046 n)."""
        edge = None if box_edge is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(box_edge, np.float64), (self.R,)))
        check(lib().lmc_kmc_set_log(self.h, int(log_dump_steps), int(capacity),
                                    None if edge is None else ptr(edge, capi._f64p)))
        self._log_edge = edge

    def write_log(self, filenames):
        """Append the records taken since the last call to one kmc_log.txt-format file per
        replica (None skips a replica)."""
        if isinstance(filenames, (str, bytes)) or hasattr(filenames, "__fspath__"):
            filenames = [filenames]
        if len(filenames) != self.R:
            raise ValueError(f"need {self.R} file names")
        arr = (C.c_char_p * self.R)(*[None if f is None else str(f).encode() for f in filenames])
        check(lib().lmc_kmc_write_log(self.h, arr))

    def occupancy(self):
        out = np.zeros(self.total_sites, np.uint8)
        check(lib().lmc_kmc_occupancy(self.h, ptr(out, capi._u8p)))
        sizes = 4 * self.ncell.astype(np.int64) ** 3
        return np.split(out, np.cumsum(sizes)[:-1])

    def set_surrogate(self, surrogate, threshold=0.01, window=0.0):
        """Evaluate jumps with `surrogate` (None: the cluster expansion) from the next run on."""
        check(lib().lmc_kmc_set_surrogate(self.h, surrogate.h if surrogate else None, threshold,
                                          window))
        self._surrogate = surrogate

    def surrogate_stats(self):
        ev, fb = C.c_uint64(), C.c_uint64()
        check(lib().lmc_kmc_surrogate_stats(self.h, C.byref(ev), C.byref(fb)))
        return dict(events=ev.value, fallbacks=fb.value)

    def samples(self, model=None, stream=None):
        """Exact surrogate training samples (env rows, E_kra, dE) of the 12 jumps of every
        replica in its current state, evaluated on the GPU with `model` (default: the driver's
        barrier model)."""
        torch = _torch()
        occ, off, nc, vac = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(lib().lmc_kmc_device_views(self.h, C.byref(occ), C.byref(off), C.byref(nc),
                                         C.byref(vac)))
        m = model if model is not None else self._models[0]
        env = torch.empty((12 * self.R, 64), dtype=torch.uint8, device="cuda")
        ekra = torch.empty(12 * self.R, dtype=torch.float64, device="cuda")
        de = torch.empty(12 * self.R, dtype=torch.float64, device="cuda")
        check(lib().lmc_jump_samples(m.h, occ, off, nc, vac, self.R, dptr(env), dptr(ekra),
                                     dptr(de), capi.stream_ptr(stream)))
        return env, ekra, de

    def trace(self):
        n = self.R * self.trace_cap
        sel, ea, de, t = np.zeros(n, np.int64), np.zeros(n), np.zeros(n), np.zeros(n)
        check(lib().lmc_kmc_trace(self.h, ptr(sel, capi._i64p), ptr(ea, capi._f64p),
                                  ptr(de, capi._f64p), ptr(t, capi._f64p)))
        shape = (self.R, self.trace_cap)
        return dict(sel=sel.reshape(shape), ea=ea.reshape(shape), de=de.reshape(shape),
                    time=t.reshape(shape))


class KineticMcChain(KineticMcFirst):
    """Batched KineticMcChainOmpi: second-order residence-time steps (formulation 3), the 144
    jumps out of the vacancy's first neighbours evaluated on the GPU each step. Same arguments
    as KineticMcFirst."""

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.set_formulation(3)

    def autotune(self, steps=200):
        """Chain steps have one kernel shape; nothing to choose."""
        return dict(formulation=3, ms_dense=float("nan"), ms_solvent=float("nan"),
                    ms_wide=float("nan"))


class CanonicalMc:
    """Batched CanonicalMcSerial driver (optional linear temperature ramp)."""

    def __init__(self, pairsite_model, energy_model, ncell, occ_list, t_start, t_end, seed,
                 maximum_steps, replica_ids=None, restart_steps=0, restart_energy=0.0):
        R = len(ncell)
        self.R = R
        self.ncell = np.asarray(ncell, np.int32)
        rid = np.arange(R) if replica_ids is None else np.asarray(replica_ids)
        params = (capi.CmcReplica * R)()
        for r in range(R):
            p = params[r]
            p.t_start = float(np.broadcast_to(t_start, (R,))[r])
            p.t_end = float(np.broadcast_to(t_end, (R,))[r])
            p.seed = int(seed)
            p.replica_id = int(rid[r])
            p.restart_steps = int(np.broadcast_to(restart_steps, (R,))[r])
            p.restart_energy = float(np.broadcast_to(restart_energy, (R,))[r])
            p.maximum_steps = int(np.broadcast_to(maximum_steps, (R,))[r])
        _, occ, _ = _concat_occupancies(self.ncell, occ_list)
        self.total_sites = occ.size
        h = C.c_void_p()
        check(lib().lmc_cmc_create(pairsite_model.h, energy_model.h if energy_model else None, R,
                                   ptr(self.ncell, capi._i32p), ptr(occ, capi._u8p), params,
                                   C.byref(h)))
        self.h = h
        self.trace_cap = 0
        self._models = (pairsite_model, energy_model)

    def __del__(self):
        if getattr(self, "h", None):
            try:
                lib().lmc_cmc_free(self.h)
            except Exception:  # interpreter shutdown
                pass
            self.h = None

    def set_trace(self, capacity):
        check(lib().lmc_cmc_set_trace(self.h, capacity))
        self.trace_cap = capacity

    def set_lds_occupancy(self, enable):
        """Hold each replica's occupancy in LDS during runs where it fits (default True), or in
        HBM (leaves the CU's LDS to kernels running beside it). Results are identical."""
        check(lib().lmc_cmc_set_lds_occupancy(self.h, 1 if enable else 0))

    def set_parallel_trials(self, trials_per_round):
        """CanonicalMcOmp (CanonicalMcOmp.cpp:40-92): rounds of up to trials_per_round (1..64)
        swap pairs under the 3NN exclusion set, their dE on the round's starting occupancy, then
        the Metropolis steps in order; 0 restores the serial trials."""
        check(lib().lmc_cmc_set_parallel_trials(self.h, int(trials_per_round)))

    def round_left(self):
        """Per replica, the trials of the current CanonicalMcOmp round not yet run."""
        out = np.zeros(self.R, np.int32)
        check(lib().lmc_cmc_round_left(self.h, ptr(out, capi._i32p)))
        return out

    def run(self, trials, stream=None):
        s = capi.stream_ptr(stream) if stream is not None else None
        check(lib().lmc_cmc_run(self.h, trials, s))

    def state(self):
        R = self.R
        out = dict(steps=np.zeros(R, np.uint64), energy=np.zeros(R), abs_energy=np.zeros(R),
                   accepted=np.zeros(R, np.uint64))
        check(lib().lmc_cmc_state(self.h, ptr(out["steps"], capi._u64p),
                                  ptr(out["energy"], capi._f64p),
                                  ptr(out["abs_energy"], capi._f64p),
                                  ptr(out["accepted"], capi._u64p)))
        return out

    def set_atoms(self, lattice_to_atom, map_shift=None):
        """Track atom ids and map_shift through accepted swaps (Config::LatticeJump)."""
        _set_atoms(lib().lmc_cmc_set_atoms, self.h, self.ncell, lattice_to_atom, map_shift)

    def atoms(self):
        """(per-replica lattice_to_atom, per-replica map_shift [N, 3])."""
        return _atoms(lib().lmc_cmc_atoms, self.h, self.ncell)

    def set_log(self, log_dump_steps, capacity=100000, thermodynamic_averaging_steps=0):
        """Keep cmc_log.txt records at the reference's dump cadence (0 turns it off), with the
        thermodynamic average over the last thermodynamic_averaging_steps energies (canonical
        MC); with annealing, the simulated-annealing log (every log_dump_steps steps and at each
        new lowest energy)."""
        check(lib().lmc_cmc_set_log(self.h, int(log_dump_steps), int(capacity),
                                    int(thermodynamic_averaging_steps)))

    def write_log(self, filenames):
        """Append the records taken since the last call to one log file per replica (None skips
        a replica)."""
        if isinstance(filenames, (str, bytes)) or hasattr(filenames, "__fspath__"):
            filenames = [filenames]
        if len(filenames) != self.R:
            raise ValueError(f"need {self.R} file names")
        arr = (C.c_char_p * self.R)(*[None if f is None else str(f).encode() for f in filenames])
        check(lib().lmc_cmc_write_log(self.h, arr))

    def occupancy(self):
        out = np.zeros(self.total_sites, np.uint8)
        check(lib().lmc_cmc_occupancy(self.h, ptr(out, capi._u8p)))
        sizes = 4 * self.ncell.astype(np.int64) ** 3
        return np.split(out, np.cumsum(sizes)[:-1])

    def trace(self):
        n = self.R * self.trace_cap
        p, q = np.zeros(n, np.int64), np.zeros(n, np.int64)
        de, acc = np.zeros(n), np.zeros(n, np.int8)
        check(lib().lmc_cmc_trace(self.h, ptr(p, capi._i64p), ptr(q, capi._i64p),
                                  ptr(de, capi._f64p), ptr(acc, capi._i8p)))
        shape = (self.R, self.trace_cap)
        return dict(p=p.reshape(shape), q=q.reshape(shape), de=de.reshape(shape),
                    acc=acc.reshape(shape))


class SimulatedAnnealing(CanonicalMc):
    """Batched SimulatedAnnealing (SimulatedAnnealing.cpp:99-185): swap trials with the
    reference's adaptive schedule from T = t_start over maximum_steps trials. Energies start at
    restart_energy (default 0); with solvent given (and an energy model), at the reference's
    starting energy (reference_energy). Arguments as CanonicalMc without t_end."""

    def __init__(self, pairsite_model, energy_model, ncell, occ_list, t_start, seed, maximum_steps,
                 solvent=None, **kwargs):
        if solvent is not None:
            if energy_model is None:
                raise ValueError("the reference's starting energy needs an energy model")
            kwargs["restart_energy"] = [
                SimulatedAnnealing.reference_energy(energy_model, n, o, solvent)
                for n, o in zip(np.broadcast_to(ncell, (len(occ_list),)), occ_list)]
        super().__init__(pairsite_model, energy_model, ncell, occ_list, t_start, t_start, seed,
                         maximum_steps, **kwargs)
        check(lib().lmc_cmc_set_annealing(self.h, 1))

    @staticmethod
    def reference_energy(energy_model, n, occ, solvent):
        """The reference's starting energy (SimulatedAnnealing.cpp:59-72): E(config) - E(pure
        solvent box of the same size) - sum over solute atoms of their chemical potentials."""
        sp = energy_model.species
        occ = np.asarray(occ, np.uint8)
        pure = np.full_like(occ, sp.index(solvent))
        e = total_energy(energy_model, ReplicaBatch([int(n), int(n)], [occ, pure])).cpu().numpy()
        mu = chemical_potential(energy_model, solvent)
        counts = np.bincount(occ, minlength=len(sp))
        solution = sum(mu[el] * float(counts[c]) for c, el in enumerate(sp) if el != solvent)
        return float(e[0] - e[1] - solution)

    def annealing_state(self):
        R = self.R
        T, low, rh = np.zeros(R), np.zeros(R), np.zeros(R, np.uint32)
        check(lib().lmc_cmc_annealing_state(self.h, ptr(T, capi._f64p), ptr(low, capi._f64p),
                                            ptr(rh, C.POINTER(C.c_uint32))))
        return dict(temperature=T, lowest=low, reheats=rh)


def jump_samples(model, batch, vac, stream=None):
    """Exact training samples of the 12 jumps per replica (event order): environment rows
    (uint8 [12R, 64]), E_kra and dE (float64 [12R])."""
    torch = _torch()
    dev = batch.occ.device
    R = batch.R
    vac_t = torch.as_tensor(_per_replica(vac, R, "vac", np.int64)).to(dev)
    env = torch.empty((12 * R, 64), dtype=torch.uint8, device=dev)
    ekra = torch.empty(12 * R, dtype=torch.float64, device=dev)
    de = torch.empty(12 * R, dtype=torch.float64, device=dev)
    check(lib().lmc_jump_samples(model.h, dptr(batch.occ), dptr(batch.offset), dptr(batch.ncell),
                                 dptr(vac_t), R, dptr(env), dptr(ekra), dptr(de),
                                 capi.stream_ptr(stream)))
    return env, ekra, de


class Surrogate:
    """Ensemble of BF16 MLPs standing in for a barrier model's jump evaluation (see
    lmc_surrogate_create); the barrier model stays the fallback and the reference its errors
    are measured against."""

    def __init__(self, barrier_model, n_members=4, seed=0):
        h = C.c_void_p()
        check(lib().lmc_surrogate_create(barrier_model.h, n_members, seed, C.byref(h)))
        self.h = h
        self.model = barrier_model
        m, k, p, c, lg = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        codes = np.zeros(4, np.int32)
        check(lib().lmc_surrogate_layout(self.h, C.byref(m), C.byref(k), C.byref(p), C.byref(c),
                                         ptr(codes, capi._i32p), C.byref(lg)))
        self.members, self.inputs, self.n_params, self.channels = m.value, k.value, p.value, c.value
        self.log_ekra = bool(lg.value)
        self.channel_codes = [int(x) for x in codes[:c.value]]

    def __del__(self):
        if getattr(self, "h", None):
            try:
                lib().lmc_surrogate_free(self.h)
            except Exception:  # interpreter shutdown
                pass
            self.h = None

    def get_params(self):
        out = np.zeros((self.members, self.n_params), np.float32)
        check(lib().lmc_surrogate_get_params(self.h, ptr(out, capi._f32p)))
        return out

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        assert p.shape == (self.members, self.n_params)
        check(lib().lmc_surrogate_set_params(self.h, ptr(p, capi._f32p)))

    def normalization(self):
        mu, sg = np.zeros(2), np.zeros(2)
        check(lib().lmc_surrogate_get_normalization(self.h, ptr(mu, capi._f64p),
                                                    ptr(sg, capi._f64p)))
        return mu, sg

    def set_normalization(self, mu, sigma):
        mu = np.ascontiguousarray(mu, np.float64)
        sg = np.ascontiguousarray(sigma, np.float64)
        check(lib().lmc_surrogate_set_normalization(self.h, ptr(mu, capi._f64p),
                                                    ptr(sg, capi._f64p)))

    def forward(self, env, stream=None):
        """Normalised outputs [n, members, 2] for environment rows env (uint8 [n, 64])."""
        torch = _torch()
        _device_rows(env, "env", 64, torch.uint8)
        n = env.shape[0]
        out = torch.empty((n, self.members, 2), dtype=torch.float32, device=env.device)
        check(lib().lmc_surrogate_forward(self.h, dptr(env), n, dptr(out),
                                          capi.stream_ptr(stream)))
        return out

    def gradients(self, env, target):
        """Per-member gradients [members, n_params] of sum (y - t)^2 / (2n) and the sums of
        squared errors [members] for normalised targets (float32 [n, 2])."""
        torch = _torch()
        _device_rows(env, "env", 64, torch.uint8)
        n = env.shape[0]
        g = torch.empty((self.members, self.n_params), dtype=torch.float32, device=env.device)
        loss = np.zeros(self.members)
        t = _device_rows(target.to(torch.float32).contiguous(), "target", 2, torch.float32)
        if t.shape[0] != n:
            raise ValueError(f"target: {t.shape[0]} rows for {n} environments")
        check(lib().lmc_surrogate_gradients(self.h, dptr(env), dptr(t), n, dptr(g),
                                            ptr(loss, capi._f64p)))
        return g, loss

    def train(self, env, ekra, de, epochs, batch=1024, lr=2e-3, seed=0):
        """Adam training on samples; returns the last epoch's mean squared error per member."""
        torch = _torch()
        _device_rows(env, "env", 64, torch.uint8)
        n = env.shape[0]
        for name, v in (("ekra", ekra), ("de", de)):
            if v.dtype != torch.float64 or v.numel() != n or not v.is_contiguous():
                raise ValueError(f"{name}: expected a contiguous float64 tensor of {n} values")
        mse = np.zeros(self.members)
        check(lib().lmc_surrogate_train(self.h, dptr(env), dptr(ekra), dptr(de), env.shape[0],
                                        epochs, batch, lr, seed, ptr(mse, capi._f64p)))
        return mse

    def eval_jumps(self, model, batch, vac, temps, threshold=0.01, window=0.0, stream=None):
        """The 12 jumps per replica as the surrogate KMC step evaluates them (event order):
        nb, Ea, dE [R, 12] and the fallback flags [R, 12] (1 = exact cluster expansion)."""
        torch = _torch()
        dev = batch.occ.device
        R = batch.R
        vac_t = torch.as_tensor(_per_replica(vac, R, "vac", np.int64)).to(dev)
        T = torch.as_tensor(np.broadcast_to(np.asarray(temps, np.float64), (R,)).copy()).to(dev)
        nb = torch.empty((R, 12), dtype=torch.int64, device=dev)
        ea = torch.empty((R, 12), dtype=torch.float64, device=dev)
        de = torch.empty((R, 12), dtype=torch.float64, device=dev)
        fl = torch.empty((R, 12), dtype=torch.int8, device=dev)
        check(lib().lmc_surrogate_eval_jumps(self.h, model.h, dptr(batch.occ), dptr(batch.offset),
                                             dptr(batch.ncell), dptr(vac_t), dptr(T), R,
                                             threshold, window, dptr(nb), dptr(ea), dptr(de),
                                             dptr(fl), capi.stream_ptr(stream)))
        return nb, ea, de, fl

    def predict(self, env, stream=None):
        """Denormalised (E_kra, dE) of every member: float64 [n, members, 2]."""
        y = self.forward(env, stream).double()
        mu, sg = self.normalization()
        torch = _torch()
        out = y * torch.tensor(sg, device=y.device) + torch.tensor(mu, device=y.device)
        if self.log_ekra:
            out[..., 0] = torch.exp(out[..., 0])
        return out
