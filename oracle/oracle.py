"""ctypes binding of the oracle restatement and runner for the reference driver.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "mdacp_ref")
# the same driver linked against the MI355X facade instead of the reference's
# MDUnit / MeshList / ForceCalculator (make -C oracle facade_driver)
FACADE_DRIVER_BIN = os.path.join(HERE, "_ref", "mdacp_gpu_driver")

NVE, NOSEHOOVER, LANGEVIN = 0, 1, 2
NOISE_MT19937, NOISE_PHILOX = 0, 1

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ctypes.CDLL(LIB_PATH)
        P, D, I, L64 = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_long
        sig = {
            "ora_create": (P, [D, D, D, D, D, D]),
            "ora_destroy": (None, [P]),
            "ora_set_thermostat": (None, [P, I, D, D, D]),
            "ora_set_noise": (None, [P, I, ctypes.c_uint64]),
            "ora_set_check_list_length": (None, [P, I]),
            "ora_add_particles": (I, [P, I, P, P, P]),
            "ora_set_initial_velocity": (None, [P, D, I]),
            "ora_set_seed": (None, [P, I]),
            "ora_make_pair_list": (I, [P]),
            "ora_calculate": (I, [P]),
            "ora_last_step_rebuilt": (I, [P]),
            "ora_send_border_particles": (None, [P]),
            "ora_change_scale": (None, [P, D]),
            "ora_update_position_half": (None, [P]),
            "ora_calculate_force": (None, [P]),
            "ora_kick_only": (None, [P, P]),
            "ora_kinetic_energy_raw": (D, [P]),
            "ora_potential_energy_raw": (D, [P]),
            "ora_virial_raw": (D, [P]),
            "ora_temperature": (D, [P]),
            "ora_pressure": (D, [P]),
            "ora_total_energy": (D, [P]),
            "ora_sim_time": (D, [P]),
            "ora_zeta": (D, [P]),
            "ora_particle_number": (I, [P]),
            "ora_total_particle_number": (I, [P]),
            "ora_pair_number": (L64, [P]),
            "ora_mesh_dims": (I, [P, P]),
            "ora_get_state": (None, [P, P, P, P]),
            "ora_set_momenta": (None, [P, P]),
            "ora_get_ghosts": (None, [P, P, P, P]),
            "ora_get_pairs": (L64, [P, P]),
            "ora_get_cells": (None, [P, P]),
            "ora_buffer_length": (D, [P]),
            "ora_lifetime": (I, [P]),
            "ora_mt_raw": (None, [ctypes.c_uint32, I, P]),
            "ora_mt_uniform": (None, [ctypes.c_uint32, I, P]),
            "ora_mt_normal": (None, [ctypes.c_uint32, D, I, P]),
            "ora_philox4x32": (None, [P, P, P]),
            "ora_philox_normal3": (None, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, P]),
            "ora_fcc": (L64, [D, D, D, D, L64, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def fcc(L, density):
    """FCC lattice of confmaker.cc:8-44 in a box of side(s) L."""
    Lx, Ly, Lz = (L, L, L) if np.isscalar(L) else L
    n = lib().ora_fcc(Lx, Ly, Lz, density, 0, None)
    x = np.zeros((n, 3))
    lib().ora_fcc(Lx, Ly, Lz, density, n, _ptr(x))
    return x


class OracleSystem:
    """One MDACP domain restated on the CPU (test oracle)."""

    def __init__(self, L, dt, rc, margin):
        Lx, Ly, Lz = (L, L, L) if np.isscalar(L) else L
        self.L = np.array([Lx, Ly, Lz], dtype=float)
        self.h = lib().ora_create(Lx, Ly, Lz, dt, rc, margin)

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_destroy(self.h)
            self.h = None

    def add(self, x, v=None, ids=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        n = len(x)
        v = None if v is None else np.ascontiguousarray(v, dtype=np.float64)
        ids = np.arange(n, dtype=np.int32) if ids is None else np.ascontiguousarray(ids, dtype=np.int32)
        return lib().ora_add_particles(self.h, n, _ptr(x), _ptr(v), _ptr(ids))

    def set_thermostat(self, kind, T=1.0, gamma=0.1, tau=0.1):
        lib().ora_set_thermostat(self.h, kind, T, gamma, tau)

    def set_noise(self, kind, seed=0):
        lib().ora_set_noise(self.h, kind, seed)

    def set_initial_velocity(self, v0, seed=0):
        lib().ora_set_initial_velocity(self.h, v0, seed)

    def set_seed(self, seed):
        lib().ora_set_seed(self.h, seed)

    def make_pair_list(self):
        err = lib().ora_make_pair_list(self.h)
        if err:
            raise RuntimeError(f"oracle make_pair_list error {err}")

    def step(self):
        err = lib().ora_calculate(self.h)
        if err:
            raise RuntimeError(f"oracle step error {err}")
        return bool(lib().ora_last_step_rebuilt(self.h))

    def change_scale(self, alpha):
        lib().ora_change_scale(self.h, alpha)
        self.L *= alpha

    @property
    def pn(self):
        return lib().ora_particle_number(self.h)

    @property
    def tn(self):
        return lib().ora_total_particle_number(self.h)

    def state(self):
        n = self.pn
        q = np.zeros((n, 3)); p = np.zeros((n, 3)); t = np.zeros(n, dtype=np.int32)
        lib().ora_get_state(self.h, _ptr(q), _ptr(p), _ptr(t))
        return q, p, t

    def set_momenta(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        lib().ora_set_momenta(self.h, _ptr(p))

    def ghosts(self):
        ng = self.tn - self.pn
        src = np.zeros(ng, dtype=np.int32); code = np.zeros(ng, dtype=np.int32); q = np.zeros((ng, 3))
        lib().ora_get_ghosts(self.h, _ptr(src), _ptr(code), _ptr(q))
        return src, code, q

    def pairs(self):
        n = lib().ora_pair_number(self.h)
        out = np.zeros((n, 3), dtype=np.int32)
        lib().ora_get_pairs(self.h, _ptr(out))
        return out

    def kick(self):
        dp = np.zeros((self.pn, 3))
        lib().ora_kick_only(self.h, _ptr(dp))
        return dp

    def observables(self):
        f = lib()
        return dict(ke=f.ora_kinetic_energy_raw(self.h), pe=f.ora_potential_energy_raw(self.h),
                    vir=f.ora_virial_raw(self.h), T=f.ora_temperature(self.h),
                    P=f.ora_pressure(self.h), E=f.ora_total_energy(self.h),
                    time=f.ora_sim_time(self.h))

    def buffer_length(self):
        return lib().ora_buffer_length(self.h)


def canonical_pairs(pairs3):
    """Sorted canonical pair set as a structured int64 key array."""
    p = np.asarray(pairs3, dtype=np.int64)
    key = (p[:, 0] << 36) | (p[:, 1] << 5) | p[:, 2]
    return np.sort(key)


# ------------------------------------------------------------ reference runs
def ref_available():
    return os.path.exists(REF_BIN)


def run_reference(L, x, v, ids, steps, dump_every=1, dt=0.001, thermostat="nve", T=1.0,
                  gamma=0.1, tau=0.1, seed=1, init_v0=None, change_scale=None, timeout=600,
                  binary=None, kick=True, dump_file_step=None):
    """Run the reference's MDManager (compiled unmodified, rc=3.0, margin=0.3).

    init_v0: MDManager::SetInitialVelocity(init_v0) after loading (the given
    momenta are replaced).  change_scale = (alpha, step): MDManager::ChangeScale(alpha)
    after that step, followed by a record with scaled=True.  binary: another build of
    the driver program (FACADE_DRIVER_BIN); kick=False leaves the kick section zero.
    dump_file_step: MDManager::SaveAsDumpFile after that step (0: before the first);
    returns (records, the dump as an (n, 6) array of x, y, z, px, py, pz rows in the
    program's own order)."""
    binary = REF_BIN if binary is None else binary
    if not os.path.exists(binary):
        raise RuntimeError(f"{binary} is not built (run `make -C oracle ref` / `facade_driver`)")
    Lx, Ly, Lz = (float(L),) * 3 if np.isscalar(L) else (float(v) for v in L)
    dt, T, gamma, tau = float(dt), float(T), float(gamma), float(tau)
    init_v0 = None if init_v0 is None else float(init_v0)
    with tempfile.TemporaryDirectory() as td:
        spec = os.path.join(td, "spec.cfg")
        lines = [f"SystemSizeX={Lx!r}", f"SystemSizeY={Ly!r}", f"SystemSizeZ={Lz!r}",
                 f"TimeStep={dt!r}", f"AimedTemperature={T!r}", f"HeatbathGamma={gamma!r}",
                 f"HeatbathTau={tau!r}", f"Steps={steps}", f"DumpEvery={dump_every}", f"Seed={seed}"]
        if thermostat != "nve":
            lines.append("ControlTemperature=yes")
            lines.append("HeatbathType=" + ("Langevin" if thermostat == "langevin" else "NoseHoover"))
        if init_v0 is not None:
            lines.append(f"InitV0={init_v0!r}")
        if not kick:
            lines.append("Kick=0")
        if change_scale is not None:
            lines.append(f"ChangeScaleAlpha={float(change_scale[0])!r}")
            lines.append(f"ChangeScaleStep={int(change_scale[1])}")
        if dump_file_step is not None:
            lines.append(f"DumpFileStep={int(dump_file_step)}")
        with open(spec, "w") as f:
            f.write("\n".join(lines) + "\n")
        pf = os.path.join(td, "particles.bin")
        n = len(x)
        xv = np.zeros((n, 6))
        xv[:, :3] = x
        if v is not None:
            xv[:, 3:] = v
        with open(pf, "wb") as f:
            np.array([n], dtype=np.int32).tofile(f)
            xv.tofile(f)
            np.asarray(ids, dtype=np.int32).tofile(f)
        out = os.path.join(td, "out.bin")
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([binary, spec, pf, out], check=True, env=env, cwd=td,
                       stdout=subprocess.DEVNULL, timeout=timeout)
        recs = parse_records(open(out, "rb").read())
        if dump_file_step is None:
            return recs
        return recs, np.fromfile(os.path.join(td, "conf.dump"), dtype=np.float64).reshape(-1, 6)


def parse_records(buf):
    recs = []
    off = 0
    while off < len(buf):
        hdr = np.frombuffer(buf, dtype=np.int64, count=8, offset=off); off += 64
        step, pn, tn, npairs, rebuilt, ncons, scaled = (int(v) for v in hdr[:7])
        if pn < 0:
            recs.append(dict(step=step, rebuilt=bool(rebuilt), dump=False))
            continue
        obs = np.frombuffer(buf, dtype=np.float64, count=8, offset=off); off += 64
        ng = tn - pn

        def take(dtype, count):
            nonlocal off
            a = np.frombuffer(buf, dtype=dtype, count=count, offset=off).copy()
            off += a.nbytes
            return a

        q = take(np.float64, 3 * pn).reshape(pn, 3)
        p = take(np.float64, 3 * pn).reshape(pn, 3)
        t = take(np.int32, pn)
        gsrc = take(np.int32, ng)
        gcode = take(np.int32, ng)
        gq = take(np.float64, 3 * ng).reshape(ng, 3)
        pairs = take(np.int32, 3 * npairs).reshape(npairs, 3)
        pairs_csr = take(np.int32, 3 * npairs).reshape(npairs, 3)
        kick = take(np.float64, 3 * pn).reshape(pn, 3)
        recs.append(dict(step=step, rebuilt=bool(rebuilt), dump=True, scaled=bool(scaled), pn=pn, tn=tn, ncons=ncons,
                         ke=obs[0], pe=obs[1], vir=obs[2], T=obs[3], P=obs[4], E=obs[5],
                         time=obs[6], q=q, p=p, type=t, gsrc=gsrc, gcode=gcode, gq=gq,
                         pairs=pairs, pairs_csr=pairs_csr, kick=kick))
    return recs
