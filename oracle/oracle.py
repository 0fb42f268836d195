"""ctypes wrapper of the CPU oracle (liblmc_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU-baseline leg of
bench.py, never by the engine. See lmc_oracle.h for the reference functions each call restates.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblmc_oracle.so")
_lib = None

ENERGY, E0, QUARTIC, PAIRSITE = 0, 1, 2, 3

_u8p = C.POINTER(C.c_uint8)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)


class KmcParams(C.Structure):
    _fields_ = [("temperature", C.c_double), ("seed", C.c_uint64), ("replica", C.c_uint32),
                ("restart_steps", C.c_uint64), ("restart_energy", C.c_double),
                ("restart_time", C.c_double), ("rate_corrector", C.c_int),
                ("early_stop", C.c_int), ("solute_disp", C.c_int), ("n_tt", C.c_int),
                ("tt_time", _f64p), ("tt_temp", _f64p), ("lru_capacity", C.c_int),
                ("omp_threads", C.c_int), ("chain", C.c_int)]


class CmcParams(C.Structure):
    _fields_ = [("t_start", C.c_double), ("t_end", C.c_double), ("seed", C.c_uint64),
                ("replica", C.c_uint32), ("restart_steps", C.c_uint64),
                ("restart_energy", C.c_double), ("anneal", C.c_int)]


def build():
    """Compile the oracle with its Makefile (g++ only)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        sig = {
            "orc_lattice_new": (vp, [C.c_int]),
            "orc_lattice_free": (None, [vp]),
            "orc_lattice_num_sites": (C.c_int, [vp]),
            "orc_lattice_neighbors": (C.c_int, [vp, C.c_int, C.c_int, _i32p]),
            "orc_lattice_position": (None, [vp, C.c_int, _f64p]),
            "orc_model_new": (vp, [vp, C.c_char_p, C.c_char_p, C.c_int, C.c_char_p, C.c_int]),
            "orc_model_free": (None, [vp]),
            "orc_model_num_eci": (C.c_int, [vp]),
            "orc_model_num_features": (C.c_int, [vp, C.c_int]),
            "orc_model_num_species": (C.c_int, [vp]),
            "orc_model_code": (C.c_int, [vp, C.c_char_p]),
            "orc_barrier_de": (C.c_int, [vp, vp, _u8p, C.c_int, C.c_int, _f64p, _f64p]),
            "orc_encode": (C.c_int, [vp, vp, _u8p, C.c_int, C.c_int, C.c_int, _f64p]),
            "orc_swap_de": (C.c_int, [vp, vp, _u8p, C.c_int, C.c_int, _f64p]),
            "orc_site_de": (C.c_int, [vp, vp, _u8p, C.c_int, C.c_int, _f64p]),
            "orc_total_energy": (C.c_int, [vp, vp, _u8p, _f64p]),
            "orc_total_encode": (C.c_int, [vp, vp, _u8p, _f64p]),
            "orc_state_clusters": (C.c_int, [vp, vp, C.c_int, C.c_int, _i32p, C.c_int]),
            "orc_orbit_clusters": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_int, _i32p, C.c_int]),
            "orc_philox": (None, [C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32),
                                  C.POINTER(C.c_uint32)]),
            "orc_kmc_new": (vp, [vp, vp, vp, _u8p, C.POINTER(KmcParams), C.c_char_p, C.c_int]),
            "orc_kmc_free": (None, [vp]),
            "orc_kmc_run": (C.c_int, [vp, C.c_uint64, C.c_uint64, _i32p, _f64p, _f64p, _f64p,
                                      _f64p]),
            "orc_kmc_state": (None, [vp, C.POINTER(C.c_uint64), _f64p, _f64p, _f64p, _f64p,
                                     _i32p, _i64p, _f64p, C.POINTER(C.c_int)]),
            "orc_kmc_occupancy": (None, [vp, _u8p]),
            "orc_kmc_events": (C.c_int, [vp, _i32p, _f64p, _f64p, _f64p, _f64p]),
            "orc_kmc_step_time": (C.c_double, [vp]),
            "orc_kmc_lru_hits": (C.c_uint64, [vp]),
            "orc_cmc_new": (vp, [vp, vp, vp, _u8p, C.POINTER(CmcParams), C.c_char_p, C.c_int]),
            "orc_cmc_anneal_state": (None, [vp, _f64p, _f64p, C.POINTER(C.c_uint32)]),
            "orc_cmc_free": (None, [vp]),
            "orc_cmc_run": (C.c_int, [vp, C.c_uint64, C.c_uint64, _i32p, _i32p, _f64p,
                                      C.POINTER(C.c_int8)]),
            "orc_cmc_omp_run": (C.c_int, [vp, C.c_uint64, C.c_uint64, C.c_int, _i32p, _i32p,
                                          _f64p, C.POINTER(C.c_int8)]),
            "orc_cmc_round_left": (C.c_int, [vp]),
            "orc_cmc_state": (None, [vp, C.POINTER(C.c_uint64), _f64p, _f64p,
                                     C.POINTER(C.c_uint64)]),
            "orc_cmc_occupancy": (None, [vp, _u8p]),
            "orc_cmc_baseline": (C.c_double, [vp, vp, _u8p, C.c_double, C.c_uint64, C.c_int,
                                              C.c_uint64, C.c_uint64, C.c_int]),
            "orc_kmc_baseline": (C.c_double, [vp, vp, vp, _u8p, C.c_double, C.c_uint64,
                                              C.c_int, C.c_int, C.c_int, C.c_uint64,
                                              C.POINTER(C.c_uint64), C.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t)


def _opt(a, t):
    return None if a is None else a.ctypes.data_as(t)


class Lattice:
    """FCC n x n x n lattice with the reference's id convention and neighbour lists."""

    def __init__(self, n):
        self.n = n
        self.h = lib().orc_lattice_new(n)
        if not self.h:
            raise ValueError("lattice needs n >= 4")
        self.num_sites = lib().orc_lattice_num_sites(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_lattice_free(self.h)
            self.h = None

    def neighbors(self, site, shell):
        out = np.zeros(24, np.int32)
        k = lib().orc_lattice_neighbors(self.h, site, shell, _ptr(out, _i32p))
        return out[:k].copy()

    def position(self, site):
        out = np.zeros(3)
        lib().orc_lattice_position(self.h, site, _ptr(out, _f64p))
        return out


class Model:
    """A predictor of one kind (ENERGY, E0, QUARTIC, PAIRSITE) built on `lattice`."""

    def __init__(self, lattice, json_path, elements, kind):
        self.lattice = lattice
        err = C.create_string_buffer(512)
        self.h = lib().orc_model_new(lattice.h, json_path.encode(), ",".join(elements).encode(),
                                     kind, err, 512)
        if not self.h:
            raise RuntimeError(err.value.decode())
        self.kind = kind

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_model_free(self.h)
            self.h = None

    @property
    def num_eci(self):
        return lib().orc_model_num_eci(self.h)

    def num_features(self, which):
        return lib().orc_model_num_features(self.h, which)

    def code(self, element):
        return lib().orc_model_code(self.h, element.encode())

    def barrier_de(self, occ, v, nb):
        ea, de = C.c_double(), C.c_double()
        rc = lib().orc_barrier_de(self.h, self.lattice.h, _ptr(occ, _u8p), v, nb, C.byref(ea),
                                  C.byref(de))
        if rc:
            raise RuntimeError("oracle barrier evaluation failed")
        return ea.value, de.value

    def encode(self, occ, v, nb, which):
        out = np.zeros(self.num_features(0 if which == 0 else 1))
        if lib().orc_encode(self.h, self.lattice.h, _ptr(occ, _u8p), v, nb, which,
                            _ptr(out, _f64p)):
            raise RuntimeError("encode failed")
        return out

    def swap_de(self, occ, p, q):
        de = C.c_double()
        if lib().orc_swap_de(self.h, self.lattice.h, _ptr(occ, _u8p), p, q, C.byref(de)):
            raise RuntimeError("swap dE failed")
        return de.value

    def site_de(self, occ, site, code):
        de = C.c_double()
        if lib().orc_site_de(self.h, self.lattice.h, _ptr(occ, _u8p), site, code, C.byref(de)):
            raise RuntimeError("site dE failed")
        return de.value

    def total_energy(self, occ):
        e = C.c_double()
        if lib().orc_total_energy(self.h, self.lattice.h, _ptr(occ, _u8p), C.byref(e)):
            raise RuntimeError("total energy failed")
        return e.value

    def total_encode(self, occ):
        out = np.zeros(self.num_eci)
        if lib().orc_total_encode(self.h, self.lattice.h, _ptr(occ, _u8p), _ptr(out, _f64p)):
            raise RuntimeError("total encode failed")
        return out

    def state_clusters(self, v, nb):
        out = np.zeros((1024, 4), np.int32)
        k = lib().orc_state_clusters(self.h, self.lattice.h, v, nb, _ptr(out, _i32p), 1024)
        return out[:k].copy()

    def orbit_clusters(self, v, nb, which):
        out = np.zeros((1024, 4), np.int32)
        k = lib().orc_orbit_clusters(self.h, self.lattice.h, v, nb, which, _ptr(out, _i32p), 1024)
        return out[:k].copy()


def philox(key, ctr):
    c = (C.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    o = (C.c_uint32 * 4)()
    lib().orc_philox(int(key[0]) & 0xFFFFFFFF, int(key[1]) & 0xFFFFFFFF, c, o)
    return [o[i] for i in range(4)]


class Kmc:
    """KineticMcFirst trajectory of one replica (chain=True: KineticMcChainOmpi's second-order
    steps)."""

    def __init__(self, barrier_model, energy_model, lattice, occ, temperature, seed, replica=0,
                 restart_steps=0, restart_energy=0.0, restart_time=0.0, rate_corrector=False,
                 early_stop=False, solute_disp=False, tt=None, lru_capacity=0, omp_threads=1,
                 chain=False):
        p = KmcParams()
        p.temperature = temperature
        p.seed = seed
        p.replica = replica
        p.restart_steps = restart_steps
        p.restart_energy = restart_energy
        p.restart_time = restart_time
        p.rate_corrector = int(rate_corrector)
        p.early_stop = int(early_stop)
        p.solute_disp = int(solute_disp)
        self._tt = None
        if tt is not None:
            self._tt = (np.ascontiguousarray(tt[0], np.float64),
                        np.ascontiguousarray(tt[1], np.float64))
            p.n_tt = len(self._tt[0])
            p.tt_time = _ptr(self._tt[0], _f64p)
            p.tt_temp = _ptr(self._tt[1], _f64p)
        p.lru_capacity = lru_capacity
        p.omp_threads = omp_threads
        p.chain = int(chain)
        err = C.create_string_buffer(512)
        occ = np.ascontiguousarray(occ, np.uint8)
        self.h = lib().orc_kmc_new(barrier_model.h, energy_model.h if energy_model else None,
                                   lattice.h, _ptr(occ, _u8p), C.byref(p), err, 512)
        if not self.h:
            raise RuntimeError(err.value.decode())
        self.num_sites = lattice.num_sites

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_kmc_free(self.h)
            self.h = None

    def run(self, steps, maximum_steps=2**62, trace=False):
        tr = None
        if trace:
            tr = dict(sel=np.zeros(steps, np.int32), ea=np.zeros(steps), de=np.zeros(steps),
                      time=np.zeros(steps), rate=np.zeros(steps))
        rc = lib().orc_kmc_run(self.h, steps, maximum_steps,
                               _opt(tr and tr["sel"], _i32p), _opt(tr and tr["ea"], _f64p),
                               _opt(tr and tr["de"], _f64p), _opt(tr and tr["time"], _f64p),
                               _opt(tr and tr["rate"], _f64p))
        if rc:
            raise RuntimeError("oracle KMC failed")
        return tr

    def state(self):
        st, t, e, ae, T = C.c_uint64(), C.c_double(), C.c_double(), C.c_double(), C.c_double()
        vac, esc = C.c_int32(), C.c_int()
        disp, com = np.zeros(3, np.int64), np.zeros(3)
        lib().orc_kmc_state(self.h, C.byref(st), C.byref(t), C.byref(e), C.byref(ae), C.byref(T),
                            C.byref(vac), _ptr(disp, _i64p), _ptr(com, _f64p), C.byref(esc))
        return dict(steps=st.value, time=t.value, energy=e.value, abs_energy=ae.value,
                    temperature=T.value, vacancy=vac.value, disp=disp, com=com,
                    escaped=bool(esc.value))

    def occupancy(self):
        out = np.zeros(self.num_sites, np.uint8)
        lib().orc_kmc_occupancy(self.h, _ptr(out, _u8p))
        return out

    def events(self):
        nb = np.zeros(12, np.int32)
        a = [np.zeros(12) for _ in range(4)]
        if lib().orc_kmc_events(self.h, _ptr(nb, _i32p), *[_ptr(x, _f64p) for x in a]):
            raise RuntimeError("events failed")
        return dict(nb=nb, ea=a[0], de=a[1], rate=a[2], cum=a[3])

    def step_time(self):
        """Residence time of the current step (chain: the second-order t_2)."""
        t = lib().orc_kmc_step_time(self.h)
        if t < 0:
            raise RuntimeError("step time failed")
        return t

    def lru_hits(self):
        return lib().orc_kmc_lru_hits(self.h)


class Cmc:
    """CanonicalMcSerial trajectory of one replica (anneal=True: SimulatedAnnealing's adaptive
    schedule from T = t_start)."""

    def __init__(self, pairsite_model, energy_model, lattice, occ, t_start, t_end, seed,
                 replica=0, restart_steps=0, restart_energy=0.0, anneal=False):
        p = CmcParams(t_start, t_end, seed, replica, restart_steps, restart_energy, int(anneal))
        err = C.create_string_buffer(512)
        occ = np.ascontiguousarray(occ, np.uint8)
        self.h = lib().orc_cmc_new(pairsite_model.h, energy_model.h if energy_model else None,
                                   lattice.h, _ptr(occ, _u8p), C.byref(p), err, 512)
        if not self.h:
            raise RuntimeError(err.value.decode())
        self.num_sites = lattice.num_sites

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_cmc_free(self.h)
            self.h = None

    def run(self, trials, maximum_steps, trace=False):
        tr = None
        if trace:
            tr = dict(p=np.zeros(trials, np.int32), q=np.zeros(trials, np.int32),
                      de=np.zeros(trials), acc=np.zeros(trials, np.int8))
        rc = lib().orc_cmc_run(self.h, trials, maximum_steps, _opt(tr and tr["p"], _i32p),
                               _opt(tr and tr["q"], _i32p), _opt(tr and tr["de"], _f64p),
                               _opt(tr and tr["acc"], C.POINTER(C.c_int8)))
        if rc:
            raise RuntimeError("oracle CMC failed")
        return tr

    def run_omp(self, trials, maximum_steps, K, trace=False):
        """CanonicalMcOmp trials: rounds of up to K pairs (see orc_cmc_omp_run)."""
        tr = None
        if trace:
            tr = dict(p=np.zeros(trials, np.int32), q=np.zeros(trials, np.int32),
                      de=np.zeros(trials), acc=np.zeros(trials, np.int8))
        rc = lib().orc_cmc_omp_run(self.h, trials, maximum_steps, K,
                                   _opt(tr and tr["p"], _i32p), _opt(tr and tr["q"], _i32p),
                                   _opt(tr and tr["de"], _f64p),
                                   _opt(tr and tr["acc"], C.POINTER(C.c_int8)))
        if rc:
            raise RuntimeError("oracle CanonicalMcOmp failed")
        return tr

    def round_left(self):
        return lib().orc_cmc_round_left(self.h)

    def anneal_state(self):
        T, low, rh = C.c_double(), C.c_double(), C.c_uint32()
        lib().orc_cmc_anneal_state(self.h, C.byref(T), C.byref(low), C.byref(rh))
        return dict(temperature=T.value, lowest=low.value, reheats=rh.value)

    def state(self):
        st, e, ae, acc = C.c_uint64(), C.c_double(), C.c_double(), C.c_uint64()
        lib().orc_cmc_state(self.h, C.byref(st), C.byref(e), C.byref(ae), C.byref(acc))
        return dict(steps=st.value, energy=e.value, abs_energy=ae.value, accepted=acc.value)

    def occupancy(self):
        out = np.zeros(self.num_sites, np.uint8)
        lib().orc_cmc_occupancy(self.h, _ptr(out, _u8p))
        return out


def cmc_baseline(pairsite_model, lattice, occ, t_start, seed, replicas, trials, maximum_steps,
                 anneal):
    """Seconds for `replicas` concurrent CanonicalMc (anneal: SimulatedAnnealing) replicas of occ,
    one thread each, `trials` trials each."""
    occ = np.ascontiguousarray(occ, np.uint8)
    secs = lib().orc_cmc_baseline(pairsite_model.h, lattice.h, _ptr(occ, _u8p), t_start, seed,
                                  replicas, trials, maximum_steps, int(anneal))
    if secs < 0:
        raise RuntimeError("orc_cmc_baseline: invalid model or occupancy")
    return secs


def kmc_baseline(barrier_model, energy_model, lattice, occ, temperature, seed, replicas,
                 threads_per_replica, lru_capacity, steps, chain=False):
    hits = C.c_uint64()
    occ = np.ascontiguousarray(occ, np.uint8)
    secs = lib().orc_kmc_baseline(barrier_model.h, energy_model.h if energy_model else None,
                                  lattice.h, _ptr(occ, _u8p), temperature, seed, replicas,
                                  threads_per_replica, lru_capacity, steps, C.byref(hits),
                                  int(chain))
    return secs, hits.value
