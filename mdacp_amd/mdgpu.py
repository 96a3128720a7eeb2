"""ctypes binding of include/mdgpu.h (libmdgpu.so, the gfx950 LJ domain).

There is no CPU fallback: if the HIP library is not built, or no GPU is
present, constructing a Domain raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmdgpu.so")

OK = 0
ERR_TOO_FAST, ERR_INVALID_POSITION, ERR_INVALID_INDEX, ERR_CAPACITY, ERR_HIP, ERR_ARGUMENT = 1, 2, 3, 4, 5, 6
FP64, MIXED = 0, 1
NVE, NOSEHOOVER, LANGEVIN = 0, 1, 2

# the symbols include/mdgpu.h declares
EXPORTS = [
    "mdgpu_create", "mdgpu_destroy", "mdgpu_last_error", "mdgpu_set_precision", "mdgpu_add_particles",
    "mdgpu_particle_number", "mdgpu_get_particles", "mdgpu_set_momenta", "mdgpu_get_ghosts",
    "mdgpu_set_thermostat", "mdgpu_set_seed", "mdgpu_get_zeta", "mdgpu_get_time", "mdgpu_make_pair_list",
    "mdgpu_is_pair_list_expired", "mdgpu_update_position_half", "mdgpu_send_border_particles",
    "mdgpu_calculate_force", "mdgpu_langevin", "mdgpu_heatbath_zeta", "mdgpu_heatbath_momenta",
    "mdgpu_calculate", "mdgpu_observe", "mdgpu_change_scale", "mdgpu_get_pairs", "mdgpu_compute_kick",
    "mdgpu_enable_timing", "mdgpu_get_timing", "mdgpu_get_list_stats", "mdgpu_set_option", "mdgpu_get_option",
    "mdgpu_set_positions", "mdgpu_clear_particles", "mdgpu_set_zeta", "mdgpu_get_pair_indices",
    "mdgpu_observe_kinetic", "mdgpu_create_ensemble", "mdgpu_replica_number", "mdgpu_replica_counts",
    "mdgpu_set_replica_thermostat", "mdgpu_observe_replicas", "mdgpu_get_replica_pairs",
    "mdgpu_observe_noalloc", "mdgpu_create_unit", "mdgpu_unit_send_particles", "mdgpu_unit_receive_particles",
    "mdgpu_unit_adjust_periodic_boundary", "mdgpu_unit_find_border_particles", "mdgpu_group_make_pair_list",
    "mdgpu_group_calculate", "mdgpu_replica_zetas",
]
# directions of a unit's faces (mdconfig.h:25)
D_LEFT, D_RIGHT, D_BACK, D_FORWARD, D_DOWN, D_UP = range(6)
OPT_UNROLL, OPT_KAHAN, OPT_WAVES, OPT_LDS_KB, OPT_BRICK, OPT_GRAPH = 1, 2, 3, 4, 5, 6
OPT_SEARCH_WAVES, OPT_SEARCH_STAGE, OPT_PRUNE, OPT_ROW_SORT, OPT_BRICK_KB, OPT_SEARCH_CELLS = 7, 8, 9, 10, 11, 12
OPT_RUN_AHEAD, OPT_PIPELINE, OPT_MIXED_FRAME, OPT_DISPLACEMENT_CHECK = 13, 14, 15, 16


class Params(ctypes.Structure):
    _fields_ = [("L", ctypes.c_double * 3), ("dt", ctypes.c_double), ("rc", ctypes.c_double),
                ("margin", ctypes.c_double), ("check_list_length", ctypes.c_int),
                ("precision", ctypes.c_int), ("capacity", ctypes.c_int), ("device", ctypes.c_int)]


class Replica(ctypes.Structure):
    _fields_ = [("L", ctypes.c_double * 3), ("T", ctypes.c_double), ("gamma", ctypes.c_double),
                ("seed", ctypes.c_uint64)]


class UnitSpec(ctypes.Structure):
    _fields_ = [("grid", ctypes.c_int * 3), ("pos", ctypes.c_int * 3), ("s", ctypes.c_double * 3),
                ("e", ctypes.c_double * 3)]


class MDGPUError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mdgpu error {code}: {msg}")
        self.code = code


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m mdacp_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        PI, PD, PL = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)
        sig = {
            "mdgpu_create": [ctypes.POINTER(Params), ctypes.POINTER(ctypes.c_void_p)],
            "mdgpu_destroy": [P],
            "mdgpu_set_precision": [P, I],
            "mdgpu_add_particles": [P, I, P, P, P],
            "mdgpu_particle_number": [P, PI, PI],
            "mdgpu_get_particles": [P, P, P, P],
            "mdgpu_set_momenta": [P, P],
            "mdgpu_get_ghosts": [P, P, P, P],
            "mdgpu_set_thermostat": [P, I, D, D, D],
            "mdgpu_set_seed": [P, ctypes.c_uint64],
            "mdgpu_get_zeta": [P, PD],
            "mdgpu_get_time": [P, PD, ctypes.POINTER(ctypes.c_uint64)],
            "mdgpu_make_pair_list": [P],
            "mdgpu_is_pair_list_expired": [P, PI],
            "mdgpu_update_position_half": [P],
            "mdgpu_send_border_particles": [P],
            "mdgpu_calculate_force": [P],
            "mdgpu_langevin": [P],
            "mdgpu_heatbath_zeta": [P, D],
            "mdgpu_heatbath_momenta": [P],
            "mdgpu_calculate": [P, I, PI],
            "mdgpu_observe": [P, PD, PD, PD],
            "mdgpu_observe_noalloc": [P, PD, PD, PD],
            "mdgpu_change_scale": [P, D],
            "mdgpu_get_pairs": [P, PL, P],
            "mdgpu_compute_kick": [P, P],
            "mdgpu_enable_timing": [P, I],
            "mdgpu_get_timing": [P, PI, PD, PD, PI, PD],
            "mdgpu_get_list_stats": [P, PL, PL, PI, PI],
            "mdgpu_set_option": [P, I, I],
            "mdgpu_get_option": [P, I, PI],
            "mdgpu_set_positions": [P, P],
            "mdgpu_clear_particles": [P],
            "mdgpu_set_zeta": [P, D],
            "mdgpu_get_pair_indices": [P, PL, P, P],
            "mdgpu_observe_kinetic": [P, PD],
            "mdgpu_create_ensemble": [ctypes.POINTER(Params), I, I, ctypes.POINTER(Replica), P, P, P, P,
                                      ctypes.POINTER(ctypes.c_void_p)],
            "mdgpu_replica_number": [P, PI],
            "mdgpu_replica_counts": [P, P],
            "mdgpu_set_replica_thermostat": [P, I, D, D, ctypes.c_uint64],
            "mdgpu_observe_replicas": [P, P, P, P],
            "mdgpu_get_replica_pairs": [P, I, PL, P],
            "mdgpu_create_unit": [ctypes.POINTER(Params), ctypes.POINTER(UnitSpec), P,
                                  ctypes.POINTER(ctypes.c_void_p)],
            "mdgpu_unit_send_particles": [P, I, PI],
            "mdgpu_unit_receive_particles": [P, P],
            "mdgpu_unit_adjust_periodic_boundary": [P],
            "mdgpu_unit_find_border_particles": [P],
            "mdgpu_group_make_pair_list": [P],
            "mdgpu_group_calculate": [P, I, PI],
            "mdgpu_replica_zetas": [P, P],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.mdgpu_last_error.argtypes = [P]
        L.mdgpu_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Domain:
    """One MDACP domain resident on the GPU (libmdgpu.so)."""

    def __init__(self, L, dt=0.001, rc=2.5, margin=0.5, precision=FP64, capacity=0,
                 check_list_length=500, device=0):
        lib = load()
        Lx, Ly, Lz = (float(L),) * 3 if np.isscalar(L) else (float(v) for v in L)
        prm = Params((ctypes.c_double * 3)(Lx, Ly, Lz), dt, rc, margin, check_list_length, precision,
                     int(capacity), device)
        h = ctypes.c_void_p()
        rc_ = lib.mdgpu_create(ctypes.byref(prm), ctypes.byref(h))
        if rc_ != OK:
            raise MDGPUError(rc_, "mdgpu_create failed")
        self.h = h
        self.L = np.array([Lx, Ly, Lz])
        self.dt = dt
        self.args = dict(dt=dt, rc=rc, margin=margin, precision=precision, check_list_length=check_list_length,
                         device=device)
        self.thermo = (NVE, 1.0, 0.1, 0.1, 0)

    def close(self):
        if getattr(self, "h", None):
            load().mdgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, code):
        if code != OK:
            raise MDGPUError(code, load().mdgpu_last_error(self.h).decode())

    # ---------------------------------------------------------------- state
    def add_particles(self, q, p=None, ids=None):
        q = np.ascontiguousarray(q, dtype=np.float64)
        n = len(q)
        p = None if p is None else np.ascontiguousarray(p, dtype=np.float64)
        ids = np.arange(n, dtype=np.int32) if ids is None else np.ascontiguousarray(ids, dtype=np.int32)
        self._chk(load().mdgpu_add_particles(self.h, n, _ptr(q), _ptr(p), _ptr(ids)))

    def counts(self):
        pn, tn = ctypes.c_int(), ctypes.c_int()
        self._chk(load().mdgpu_particle_number(self.h, ctypes.byref(pn), ctypes.byref(tn)))
        return pn.value, tn.value

    def particles(self):
        pn, _ = self.counts()
        q = np.zeros((pn, 3)); p = np.zeros((pn, 3)); t = np.zeros(pn, dtype=np.int32)
        self._chk(load().mdgpu_get_particles(self.h, _ptr(q), _ptr(p), _ptr(t)))
        return q, p, t

    def particles_by_id(self):
        q, p, t = self.particles()
        o = np.argsort(t, kind="stable")
        return q[o], p[o], t[o]

    def set_momenta(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        self._chk(load().mdgpu_set_momenta(self.h, _ptr(p)))

    def set_positions(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        self._chk(load().mdgpu_set_positions(self.h, _ptr(q)))

    def clear_particles(self):
        self._chk(load().mdgpu_clear_particles(self.h))

    def set_zeta(self, zeta):
        self._chk(load().mdgpu_set_zeta(self.h, zeta))

    def ghosts(self):
        pn, tn = self.counts()
        ng = tn - pn
        src = np.zeros(ng, dtype=np.int32); code = np.zeros(ng, dtype=np.int32); q = np.zeros((ng, 3))
        self._chk(load().mdgpu_get_ghosts(self.h, _ptr(src), _ptr(code), _ptr(q)))
        return src, code, q

    def set_precision(self, precision):
        self._chk(load().mdgpu_set_precision(self.h, precision))

    def set_thermostat(self, kind, T=1.0, gamma=0.1, tau=0.1):
        self._chk(load().mdgpu_set_thermostat(self.h, kind, T, gamma, tau))
        self.thermo = (kind, T, gamma, tau, self.thermo[4])

    def set_seed(self, seed):
        self._chk(load().mdgpu_set_seed(self.h, seed))
        self.thermo = self.thermo[:4] + (seed,)

    def clone(self, **override):
        """A new domain with the same parameters, options and current particles (list built)."""
        a = dict(self.args, **override)
        c = Domain(self.L, dt=a["dt"], rc=a["rc"], margin=a["margin"], precision=a["precision"],
                   capacity=self.counts()[0], check_list_length=a["check_list_length"], device=a["device"])
        q, p, t = self.particles()
        c.add_particles(q, p, t)
        kind, T, gamma, tau, seed = self.thermo
        c.set_thermostat(kind, T, gamma, tau)
        c.set_seed(seed)
        c.copy_options(self)
        c.make_pair_list()
        return c

    def copy_options(self, other):
        for o in (OPT_UNROLL, OPT_KAHAN, OPT_WAVES, OPT_LDS_KB, OPT_GRAPH, OPT_SEARCH_WAVES, OPT_SEARCH_STAGE,
                  OPT_PRUNE, OPT_SEARCH_CELLS, OPT_RUN_AHEAD):
            self.set_option(o, other.get_option(o))

    def zeta(self):
        z = ctypes.c_double()
        self._chk(load().mdgpu_get_zeta(self.h, ctypes.byref(z)))
        return z.value

    def time(self):
        t, s = ctypes.c_double(), ctypes.c_uint64()
        self._chk(load().mdgpu_get_time(self.h, ctypes.byref(t), ctypes.byref(s)))
        return t.value, s.value

    # ----------------------------------------------------------------- step
    def make_pair_list(self):
        self._chk(load().mdgpu_make_pair_list(self.h))

    def is_pair_list_expired(self):
        e = ctypes.c_int()
        self._chk(load().mdgpu_is_pair_list_expired(self.h, ctypes.byref(e)))
        return bool(e.value)

    def update_position_half(self):
        self._chk(load().mdgpu_update_position_half(self.h))

    def send_border_particles(self):
        self._chk(load().mdgpu_send_border_particles(self.h))

    def calculate_force(self):
        self._chk(load().mdgpu_calculate_force(self.h))

    def langevin(self):
        self._chk(load().mdgpu_langevin(self.h))

    def calculate(self, nsteps=1):
        r = ctypes.c_int()
        self._chk(load().mdgpu_calculate(self.h, int(nsteps), ctypes.byref(r)))
        return r.value

    def change_scale(self, alpha):
        self._chk(load().mdgpu_change_scale(self.h, alpha))
        self.L = self.L * alpha

    # ------------------------------------------------------------- observe
    def kinetic_energy(self):
        """KineticEnergyObserver alone (no pair list needed); waits for the domain's stream."""
        ke = ctypes.c_double()
        self._chk(load().mdgpu_observe_kinetic(self.h, ctypes.byref(ke)))
        return ke.value

    def observe_raw(self):
        ke, pe, vir = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._chk(load().mdgpu_observe(self.h, ctypes.byref(ke), ctypes.byref(pe), ctypes.byref(vir)))
        return ke.value, pe.value, vir.value

    def observables(self):
        """T, P, E per MDManager::Temperature/Pressure/TotalEnergy (mdmanager.cc:390-429), in IEEE
        double arithmetic as there: NaN for a domain without particles."""
        ke, pe, vir = self.observe_raw()
        pn = np.float64(self.counts()[0])
        with np.errstate(divide="ignore", invalid="ignore"):
            T = (ke / pn) / 1.5
            V = self.L[0] * self.L[1] * self.L[2]
            return dict(ke=ke, pe=pe, vir=vir, T=float(T), P=float((T - vir / pn) * pn / V),
                        E=float(ke / pn + pe / pn))

    # -------------------------------------------------------------- export
    def pairs(self):
        n = ctypes.c_int64()
        self._chk(load().mdgpu_get_pairs(self.h, ctypes.byref(n), None))
        out = np.zeros((n.value, 3), dtype=np.int32)
        self._chk(load().mdgpu_get_pairs(self.h, ctypes.byref(n), _ptr(out)))
        return out

    def pair_indices(self):
        """(key, partner) index pairs in particles()/ghosts() order (partner >= pn: ghost)."""
        n = ctypes.c_int64()
        self._chk(load().mdgpu_get_pair_indices(self.h, ctypes.byref(n), None, None))
        key = np.zeros(n.value, dtype=np.int32)
        partner = np.zeros(n.value, dtype=np.int32)
        self._chk(load().mdgpu_get_pair_indices(self.h, ctypes.byref(n), _ptr(key), _ptr(partner)))
        return key, partner

    def kick(self):
        """dp = F*dt from zero momenta, in particles() order."""
        pn, _ = self.counts()
        dp = np.zeros((pn, 3))
        self._chk(load().mdgpu_compute_kick(self.h, _ptr(dp)))
        return dp

    def enable_timing(self, on=True):
        self._chk(load().mdgpu_enable_timing(self.h, int(on)))

    def timing(self):
        fl, fm, tm, rb, rm = ctypes.c_int(), ctypes.c_double(), ctypes.c_double(), ctypes.c_int(), ctypes.c_double()
        self._chk(load().mdgpu_get_timing(self.h, ctypes.byref(fl), ctypes.byref(fm), ctypes.byref(tm),
                                          ctypes.byref(rb), ctypes.byref(rm)))
        return dict(force_launches=fl.value, force_ms=fm.value, total_ms=tm.value, rebuilds=rb.value,
                    rebuild_ms=rm.value)

    def set_option(self, option, value):
        self._chk(load().mdgpu_set_option(self.h, option, int(value)))

    def get_option(self, option):
        v = ctypes.c_int()
        self._chk(load().mdgpu_get_option(self.h, option, ctypes.byref(v)))
        return v.value

    def list_stats(self):
        e, s, c, g = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int(), ctypes.c_int()
        self._chk(load().mdgpu_get_list_stats(self.h, ctypes.byref(e), ctypes.byref(s), ctypes.byref(c),
                                              ctypes.byref(g)))
        return dict(entries=e.value, slots=s.value, ncell=c.value, ghosts=g.value)


class Ensemble(Domain):
    """Independent replicas advanced together (mdgpu_create_ensemble).

    boxes: (R, 3) box edges; counts: particles per replica; q, p: (sum counts, 3),
    replica after replica; T, gamma, seed: per-replica Langevin parameters.
    """

    def __init__(self, boxes, counts, q, p=None, ids=None, T=None, gamma=None, seed=None, thermostat=NVE,
                 dt=0.005, rc=2.5, margin=0.5, precision=FP64, check_list_length=500, device=0):
        lib = load()
        boxes = np.asarray(boxes, dtype=np.float64).reshape(-1, 3)
        R = len(boxes)
        counts = np.ascontiguousarray(counts, dtype=np.int32)
        T = np.ones(R) if T is None else np.broadcast_to(np.asarray(T, dtype=np.float64), (R,))
        gamma = np.ones(R) if gamma is None else np.broadcast_to(np.asarray(gamma, dtype=np.float64), (R,))
        seed = np.arange(R) if seed is None else np.broadcast_to(np.asarray(seed), (R,))
        reps = (Replica * R)()
        for r in range(R):
            reps[r].L = (ctypes.c_double * 3)(*boxes[r])
            reps[r].T = float(T[r])
            reps[r].gamma = float(gamma[r])
            reps[r].seed = int(seed[r])
        q = np.ascontiguousarray(q, dtype=np.float64)
        p = None if p is None else np.ascontiguousarray(p, dtype=np.float64)
        ids = None if ids is None else np.ascontiguousarray(ids, dtype=np.int32)
        prm = Params((ctypes.c_double * 3)(*boxes[0]), dt, rc, margin, check_list_length, precision, 0, device)
        h = ctypes.c_void_p()
        rc_ = lib.mdgpu_create_ensemble(ctypes.byref(prm), thermostat, R, reps, _ptr(counts), _ptr(q), _ptr(p),
                                        _ptr(ids), ctypes.byref(h))
        if rc_ != OK:
            raise MDGPUError(rc_, "mdgpu_create_ensemble failed")
        self.h = h
        self.boxes = boxes
        self.L = boxes[0]
        self.dt = dt
        self.R = R
        self.args = dict(dt=dt, rc=rc, margin=margin, precision=precision, check_list_length=check_list_length,
                         device=device)
        self.ens = dict(T=np.array(T, dtype=np.float64), gamma=np.array(gamma, dtype=np.float64),
                        seed=np.array(seed), thermostat=thermostat)
        self.thermo = (thermostat, 1.0, 0.1, 0.1, 0)

    def clone(self, **override):
        """A new ensemble with the same parameters, options and current particles (lists built)."""
        a = dict(self.args, **override)
        q, p, t = self.particles()
        rep_of = np.repeat(np.arange(self.R), self.replica_counts())
        Lp = self.boxes[rep_of]
        q = np.where(q < 0.0, q + Lp, np.where(q >= Lp, q - Lp, q))  # into [0, L), as a construction wraps
        c = Ensemble(self.boxes, self.replica_counts(), q, p, t, T=self.ens["T"], gamma=self.ens["gamma"],
                     seed=self.ens["seed"], thermostat=self.ens["thermostat"], **a)
        c.copy_options(self)
        if c.get_option(OPT_BRICK_KB) != self.get_option(OPT_BRICK_KB):
            c.set_option(OPT_BRICK_KB, self.get_option(OPT_BRICK_KB))
        c.make_pair_list()
        return c

    def replica_counts(self):
        c = np.zeros(self.R, dtype=np.int32)
        self._chk(load().mdgpu_replica_counts(self.h, _ptr(c)))
        return c

    def replica_slices(self):
        c = self.replica_counts()
        e = np.cumsum(c)
        return [slice(int(a - n), int(a)) for a, n in zip(e, c)]

    def set_replica_thermostat(self, r, T, gamma, seed):
        self._chk(load().mdgpu_set_replica_thermostat(self.h, int(r), T, gamma, int(seed)))

    def replica_zetas(self):
        z = np.zeros(self.R)
        self._chk(load().mdgpu_replica_zetas(self.h, _ptr(z)))
        return z

    def change_scale(self, alpha):
        """MDManager::ChangeScale(alpha) of every replica: boxes and positions scaled, lists rebuilt."""
        self._chk(load().mdgpu_change_scale(self.h, alpha))
        self.boxes = self.boxes * alpha
        self.L = self.boxes[0]

    def observe_replicas_raw(self):
        ke, pe, vir = np.zeros(self.R), np.zeros(self.R), np.zeros(self.R)
        self._chk(load().mdgpu_observe_replicas(self.h, _ptr(ke), _ptr(pe), _ptr(vir)))
        return ke, pe, vir

    def observables(self):
        """T, P, E of every replica (arrays), as MDManager computes them for one system."""
        ke, pe, vir = self.observe_replicas_raw()
        pn = self.replica_counts().astype(np.float64)
        T = (ke / pn) / 1.5
        V = self.boxes.prod(axis=1)
        return dict(ke=ke, pe=pe, vir=vir, T=T, P=(T - vir / pn) * pn / V, E=ke / pn + pe / pn)

    def replica_pairs(self, r):
        n = ctypes.c_int64()
        self._chk(load().mdgpu_get_replica_pairs(self.h, int(r), ctypes.byref(n), None))
        out = np.zeros((n.value, 3), dtype=np.int32)
        self._chk(load().mdgpu_get_replica_pairs(self.h, int(r), ctypes.byref(n), _ptr(out)))
        return out


class Unit(Domain):
    """One unit of a process grid on the GPU (mdgpu_create_unit): the box [s, e) of a grid
    of units over the box L, in the group of `join` (None: a new group)."""

    def __init__(self, L, grid, pos, s, e, join=None, dt=0.001, rc=2.5, margin=0.5, precision=FP64, capacity=0,
                 check_list_length=500, device=0):
        lib = load()
        Lx, Ly, Lz = (float(L),) * 3 if np.isscalar(L) else (float(v) for v in L)
        prm = Params((ctypes.c_double * 3)(Lx, Ly, Lz), dt, rc, margin, check_list_length, precision,
                     int(capacity), device)
        spec = UnitSpec((ctypes.c_int * 3)(*grid), (ctypes.c_int * 3)(*pos), (ctypes.c_double * 3)(*s),
                        (ctypes.c_double * 3)(*e))
        h = ctypes.c_void_p()
        rc_ = lib.mdgpu_create_unit(ctypes.byref(prm), ctypes.byref(spec), join.h if join is not None else None,
                                    ctypes.byref(h))
        if rc_ != OK:
            msg = lib.mdgpu_last_error(join.h).decode() if join is not None else "mdgpu_create_unit failed"
            raise MDGPUError(rc_, msg)
        self.h = h
        self.L = np.array([Lx, Ly, Lz])
        self.grid, self.pos = tuple(grid), tuple(pos)
        self.s, self.e = np.array(s, dtype=float), np.array(e, dtype=float)
        self.dt = dt
        self.args = dict(dt=dt, rc=rc, margin=margin, precision=precision, check_list_length=check_list_length,
                         device=device)
        self.thermo = (NVE, 1.0, 0.1, 0.1, 0)

    def send_particles(self, direction):
        n = ctypes.c_int()
        self._chk(load().mdgpu_unit_send_particles(self.h, direction, ctypes.byref(n)))
        return n.value

    def receive_particles(self, src):
        self._chk(load().mdgpu_unit_receive_particles(self.h, src.h))

    def adjust_periodic_boundary(self):
        self._chk(load().mdgpu_unit_adjust_periodic_boundary(self.h))

    def find_border_particles(self):
        self._chk(load().mdgpu_unit_find_border_particles(self.h))

    def group_make_pair_list(self):
        self._chk(load().mdgpu_group_make_pair_list(self.h))

    def group_calculate(self, nsteps=1):
        """nsteps of MDManager::Calculate for every unit of the group; returns the constructions."""
        r = ctypes.c_int()
        self._chk(load().mdgpu_group_calculate(self.h, nsteps, ctypes.byref(r)))
        return r.value


def unit_boxes(L, grid):
    """Boxes of a grid of units over the box L, unit id = x + gx * (y + gy * z), each
    [ul * pos, ul * pos + ul) per axis with ul = L / g (MDUnit::MDUnit, mdunit.cc:13-24)."""
    L = [float(L)] * 3 if np.isscalar(L) else [float(v) for v in L]
    out = []
    for z in range(grid[2]):
        for y in range(grid[1]):
            for x in range(grid[0]):
                p = (x, y, z)
                s, e = [], []
                for a in range(3):
                    ul = L[a] / float(grid[a])
                    s.append(ul * float(p[a]))
                    e.append(s[-1] + ul)
                out.append((p, s, e))
    return out


def unit_grid(L, grid, **kw):
    """One Unit per box of unit_boxes, all in one group."""
    units = []
    for p, s, e in unit_boxes(L, grid):
        units.append(Unit(L, grid, p, s, e, join=units[0] if units else None, **kw))
    return units


class UnitGroup:
    """The units of a process grid as one system, as MDManager sees its units: particles added
    to every unit (each keeps its box), the group's construction and steps
    (mdgpu_group_make_pair_list, mdgpu_group_calculate), observers summed over the units."""

    def __init__(self, L, grid, **kw):
        self.units = unit_grid(L, grid, **kw)
        self.L = self.units[0].L
        self.grid = tuple(grid)

    def add_particles(self, q, p=None, ids=None):
        for u in self.units:
            u.add_particles(q, p, ids)

    def make_pair_list(self):
        self.units[0].group_make_pair_list()

    def calculate(self, nsteps=1):
        return self.units[0].group_calculate(nsteps)

    def counts(self):
        c = [u.counts() for u in self.units]
        return sum(x[0] for x in c), sum(x[1] for x in c)

    def particles(self):
        q, p, t = zip(*(u.particles() for u in self.units))
        return np.concatenate(q), np.concatenate(p), np.concatenate(t)

    def particles_by_id(self):
        q, p, t = self.particles()
        o = np.argsort(t, kind="stable")
        return q[o], p[o], t[o]

    def set_precision(self, precision):
        for u in self.units:
            u.set_precision(precision)

    def set_thermostat(self, kind, T=1.0, gamma=0.1, tau=0.1):
        for u in self.units:
            u.set_thermostat(kind, T, gamma, tau)

    def kinetic_energy(self):
        return float(sum(u.kinetic_energy() for u in self.units))

    def observe_raw(self):
        return tuple(float(v) for v in np.sum([u.observe_raw() for u in self.units], axis=0))

    def observables(self):
        """T, P, E of the whole box (MDManager::Temperature/Pressure/TotalEnergy summed over
        its units, mdmanager.cc:379-429)."""
        ke, pe, vir = self.observe_raw()
        pn = np.float64(self.counts()[0])
        T = (ke / pn) / 1.5
        V = self.L[0] * self.L[1] * self.L[2]
        return dict(ke=ke, pe=pe, vir=vir, T=float(T), P=float((T - vir / pn) * pn / V), E=float(ke / pn + pe / pn))

    def enable_timing(self, on=True):
        self.units[0].enable_timing(on)

    def timing(self):
        """Force launches and their summed time over all units, the steps' wall time."""
        t = [u.timing() for u in self.units]
        return dict(force_launches=sum(x["force_launches"] for x in t), force_ms=sum(x["force_ms"] for x in t),
                    total_ms=t[0]["total_ms"], rebuilds=t[0]["rebuilds"], rebuild_ms=0.0)

    def list_stats(self):
        st = [u.list_stats() for u in self.units]
        return {k: sum(x[k] for x in st) for k in st[0]}

    def close(self):
        for u in self.units:
            u.close()
