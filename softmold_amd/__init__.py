"""softmold_amd -- MI355X (gfx950) solver for the SoftMold MD step.

Python mirror of the C ABI in include/softmold_hip.h (loaded from the in-tree
libsoftmold_hip.so).  The reference's operator interface for this path is the set of
`mpd::*_device(...)` templates driven by MDcuda.cu (include/cuda/mpdCuda.h:1-19) and the
CellOpt/Verlet/Langevin/Blob classes driven by MD.cpp; `Solver` exposes the same
operations (build + pair forces, CHAIN forces, Langevin, Verlet halves, potential, dU of a
box trial, the MC move) on device-resident state, for one system or for a batch of
replica systems stepped together (`Solver.batch`).

There is no CPU fallback: if the HIP library is missing, importing the solver raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

__all__ = ["Solver", "Mpd", "SmError", "lib", "LIB_PATH"]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsoftmold_hip.so")

SM_OK, SM_ERR_INVALID, SM_ERR_HIP, SM_ERR_UNSUPPORTED, SM_ERR_STATE, SM_ERR_IO = range(6)


class SmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[sm error {code}] {msg}")
        self.code = code


class SmSystem(C.Structure):
    _fields_ = [
        ("n_particles", C.c_int32), ("n_types", C.c_int32),
        ("pos", C.POINTER(C.c_double)), ("type", C.POINTER(C.c_int32)),
        ("vel", C.POINTER(C.c_double)), ("acc", C.POINTER(C.c_double)),
        ("n_chain_blocks", C.c_int32), ("chain", C.POINTER(C.c_int32)),
        ("chain_const", C.POINTER(C.c_double)),
        ("fconst", C.POINTER(C.c_double)), ("uconst", C.POINTER(C.c_double)),
        ("box", C.c_double * 3),
        ("cutoff", C.c_double), ("dt", C.c_double), ("gamma", C.c_double),
        ("temperature", C.c_double), ("delta_lxy", C.c_double), ("tension", C.c_double),
        ("seed", C.c_uint64),
        ("n_bonds", C.c_int32), ("bond", C.POINTER(C.c_int32)), ("bond_const", C.POINTER(C.c_double)),
        ("n_bends", C.c_int32), ("bend", C.POINTER(C.c_int32)), ("bend_const", C.POINTER(C.c_double)),
        ("n_balls", C.c_int32), ("ball", C.POINTER(C.c_int32)), ("ball_const", C.POINTER(C.c_double)),
        ("sigma_temperature", C.c_double),
    ]


# every symbol the header declares (tests check the library exports all of them)
EXPORTS = [
    "sm_last_error", "sm_version", "sm_create", "sm_create_batch", "sm_destroy", "sm_replicas",
    "sm_sync", "sm_download_state", "sm_upload_state", "sm_set_temperature", "sm_init_forces",
    "sm_second_half", "sm_step_first", "sm_step_finish", "sm_run", "sm_mc_stats", "sm_mc_hist",
    "sm_measure", "sm_vesicle_radius", "sm_vesicle_radius_noalloc", "sm_bond_stats", "sm_compute_forces", "sm_cell_keys", "sm_pair_count", "sm_force_pair_count", "sm_pair_list",
    "sm_potential_delta", "sm_tune", "sm_tune_get", "sm_autotune", "sm_set_params", "sm_get_params",
    "sm_adjoint", "sm_param_derivative", "sm_profile_enable",
    "sm_profile_read", "sm_mpd_read", "sm_mpd_write", "sm_mpd_free", "sm_mpd_system",
    "sm_mpd_get", "sm_mpd_set", "sm_mpd_set_state", "sm_gen_liposome", "sm_gen_bilayer",
]

_lib = None


def lib():
    """The loaded HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() first "
                              "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    P = C.POINTER
    vp, d, i32, i64 = C.c_void_p, C.c_double, C.c_int32, C.c_int64
    sig = {
        "sm_last_error": (C.c_char_p, []),
        "sm_version": (C.c_char_p, []),
        "sm_create": (C.c_int, [P(SmSystem), P(vp)]),
        "sm_create_batch": (C.c_int, [P(SmSystem), i32, P(vp)]),
        "sm_destroy": (C.c_int, [vp]),
        "sm_replicas": (C.c_int, [vp, P(i32), P(i32), P(i32)]),
        "sm_sync": (C.c_int, [vp]),
        "sm_download_state": (C.c_int, [vp, P(d), P(d), P(d), P(d)]),
        "sm_upload_state": (C.c_int, [vp, P(d), P(d), P(d), P(d)]),
        "sm_set_temperature": (C.c_int, [vp, d]),
        "sm_init_forces": (C.c_int, [vp]),
        "sm_second_half": (C.c_int, [vp]),
        "sm_step_first": (C.c_int, [vp, i64]),
        "sm_step_finish": (C.c_int, [vp, i64, P(d)]),
        "sm_run": (C.c_int, [vp, i64, i32]),
        "sm_mc_stats": (C.c_int, [vp, P(i64), P(i64)]),
        "sm_mc_hist": (C.c_int, [vp, i32, i32, P(d), P(d), P(i32)]),
        "sm_measure": (C.c_int, [vp, P(d)]),
        "sm_vesicle_radius": (C.c_int, [vp, i32, P(d)]),
        "sm_vesicle_radius_noalloc": (C.c_int, [vp, i32, P(d)]),
        "sm_bond_stats": (C.c_int, [vp, P(d)]),
        "sm_compute_forces": (C.c_int, [vp, P(d), P(d)]),
        "sm_cell_keys": (C.c_int, [vp, P(i32)]),
        "sm_pair_count": (C.c_int, [vp, P(i32), P(i64)]),
        "sm_force_pair_count": (C.c_int, [vp, P(i32), P(i64)]),
        "sm_pair_list": (C.c_int, [vp, P(i32), i64, P(i64)]),
        "sm_potential_delta": (C.c_int, [vp, P(d), P(d), P(d)]),
        "sm_tune": (C.c_int, [vp, C.c_char_p, i32]),
        "sm_tune_get": (C.c_int, [vp, C.c_char_p, P(i32)]),
        "sm_autotune": (C.c_int, [vp, i32, P(d)]),
        "sm_set_params": (C.c_int, [vp, i32, i32, P(d)]),
        "sm_get_params": (C.c_int, [vp, i32, i32, P(d)]),
        "sm_adjoint": (C.c_int, [vp, i32, i32, i32, i32, P(d), P(d)]),
        "sm_param_derivative": (C.c_int, [vp, i32, i32, P(d)]),
        "sm_profile_enable": (C.c_int, [vp, i32]),
        "sm_profile_read": (C.c_int, [vp, i32, P(C.c_char_p), P(d), P(i64), P(i32)]),
        "sm_mpd_read": (C.c_int, [C.c_char_p, P(vp)]),
        "sm_mpd_write": (C.c_int, [vp, C.c_char_p]),
        "sm_mpd_free": (C.c_int, [vp]),
        "sm_mpd_system": (C.c_int, [vp, P(SmSystem)]),
        "sm_mpd_get": (C.c_int, [vp, C.c_char_p, P(d)]),
        "sm_mpd_set": (C.c_int, [vp, C.c_char_p, d]),
        "sm_mpd_set_state": (C.c_int, [vp, P(d), P(d), P(d)]),
        "sm_gen_liposome": (C.c_int, [i32, i32, d, P(vp)]),
        "sm_gen_bilayer": (C.c_int, [i32, i32, d, P(vp)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def _check(rc: int):
    if rc != SM_OK:
        raise SmError(rc, lib().sm_last_error().decode(errors="replace"))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32)) if a is not None else None


def _lp(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64)) if a is not None else None


class Mpd:
    """A `.mpd` system (include/system.h Blob + scriptFormat.h Script), host side."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle

    @classmethod
    def read(cls, name: str) -> "Mpd":
        """Read NAME.mpd (the `MD name` argument convention, scriptFormat.h:45-47)."""
        h = C.c_void_p()
        _check(lib().sm_mpd_read(name.encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def liposome(cls, seed: int, n_lipids: int, density: float) -> "Mpd":
        """generate/liposome.cpp: `liposome name seed nLipids arealDensity`."""
        h = C.c_void_p()
        _check(lib().sm_gen_liposome(seed, n_lipids, density, C.byref(h)))
        return cls(h)

    @classmethod
    def bilayer(cls, seed: int, n_lipids: int, density: float) -> "Mpd":
        """generate/bilayer.cpp without anchors: `bilayer name seed nLipids density 0 0 0 0`."""
        h = C.c_void_p()
        _check(lib().sm_gen_bilayer(seed, n_lipids, density, C.byref(h)))
        return cls(h)

    def write(self, name: str):
        _check(lib().sm_mpd_write(self.h, name.encode()))

    def __getitem__(self, key: str) -> float:
        v = C.c_double()
        _check(lib().sm_mpd_get(self.h, key.encode(), C.byref(v)))
        return v.value

    def __setitem__(self, key: str, value: float):
        _check(lib().sm_mpd_set(self.h, key.encode(), float(value)))

    def system(self) -> dict:
        """Copies of the arrays the solver needs."""
        s = SmSystem()
        _check(lib().sm_mpd_system(self.h, C.byref(s)))
        n, nt, nb = s.n_particles, s.n_types, s.n_chain_blocks
        arr = lambda p, k, dt: np.ctypeslib.as_array(p, shape=(k,)).astype(dt, copy=True)
        return {
            "pos": arr(s.pos, 3 * n, np.float64).reshape(n, 3),
            "type": arr(s.type, n, np.int32),
            "vel": arr(s.vel, 3 * n, np.float64).reshape(n, 3),
            "chain": arr(s.chain, 3 * nb, np.int32).reshape(nb, 3) if nb else np.zeros((0, 3), np.int32),
            "chain_const": arr(s.chain_const, 4 * nb, np.float64).reshape(nb, 4) if nb else np.zeros((0, 4)),
            "fconst": arr(s.fconst, 6 * nt * nt, np.float64),
            "uconst": arr(s.uconst, 6 * nt * nt, np.float64),
            "n_types": nt, "box": np.array(list(s.box)), "cutoff": s.cutoff, "dt": s.dt,
            "gamma": s.gamma, "temperature": s.temperature, "delta_lxy": s.delta_lxy,
            "tension": s.tension, "seed": int(s.seed),
            "bond": arr(s.bond, 2 * s.n_bonds, np.int32).reshape(-1, 2) if s.n_bonds else np.zeros((0, 2), np.int32),
            "bond_const": (arr(s.bond_const, 2 * s.n_bonds, np.float64).reshape(-1, 2) if s.n_bonds
                           else np.zeros((0, 2))),
            "bend": arr(s.bend, 3 * s.n_bends, np.int32).reshape(-1, 3) if s.n_bends else np.zeros((0, 3), np.int32),
            "bend_const": (arr(s.bend_const, 2 * s.n_bends, np.float64).reshape(-1, 2) if s.n_bends
                           else np.zeros((0, 2))),
            "ball": arr(s.ball, 2 * s.n_balls, np.int32).reshape(-1, 2) if s.n_balls else np.zeros((0, 2), np.int32),
            "ball_const": (arr(s.ball_const, 2 * s.n_balls, np.float64).reshape(-1, 2) if s.n_balls
                           else np.zeros((0, 2))),
            "sigma_temperature": s.sigma_temperature,
        }

    def set_state(self, pos, vel, box):
        pos = np.ascontiguousarray(pos, np.float64)
        vel = np.ascontiguousarray(vel, np.float64)
        b = np.ascontiguousarray(box, np.float64)
        _check(lib().sm_mpd_set_state(self.h, _dp(pos), _dp(vel), _dp(b)))

    def close(self):
        if getattr(self, "h", None):
            lib().sm_mpd_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_SYS_KEYS = ("cutoff", "dt", "gamma", "temperature", "delta_lxy", "tension", "seed")


def _pack(d: dict, keep: list) -> SmSystem:
    """An SmSystem viewing numpy copies of a system dict (copies kept alive in `keep`)."""
    k = {
        "pos": np.ascontiguousarray(d["pos"], np.float64).reshape(-1, 3),
        "type": np.ascontiguousarray(d["type"], np.int32),
        "vel": np.ascontiguousarray(d["vel"], np.float64).reshape(-1, 3),
        "fconst": np.ascontiguousarray(d["fconst"], np.float64),
        "uconst": np.ascontiguousarray(d["uconst"], np.float64),
        "chain": np.ascontiguousarray(d["chain"], np.int32).reshape(-1, 3),
        "chain_const": np.ascontiguousarray(d["chain_const"], np.float64).reshape(-1, 4),
        # generic BOND / BEND records (optional keys)
        "bond": np.ascontiguousarray(d.get("bond", np.zeros((0, 2))), np.int32).reshape(-1, 2),
        "bond_const": np.ascontiguousarray(d.get("bond_const", np.zeros((0, 2))), np.float64).reshape(-1, 2),
        "bend": np.ascontiguousarray(d.get("bend", np.zeros((0, 3))), np.int32).reshape(-1, 3),
        "bend_const": np.ascontiguousarray(d.get("bend_const", np.zeros((0, 2))), np.float64).reshape(-1, 2),
        "ball": np.ascontiguousarray(d.get("ball", np.zeros((0, 2))), np.int32).reshape(-1, 2),
        "ball_const": np.ascontiguousarray(d.get("ball_const", np.zeros((0, 2))), np.float64).reshape(-1, 2),
    }
    if d.get("acc") is not None:
        k["acc"] = np.ascontiguousarray(d["acc"], np.float64).reshape(-1, 3)
    keep.append(k)
    s = SmSystem()
    s.n_particles = len(k["type"])
    s.n_types = int(d["n_types"])
    s.pos, s.type, s.vel = _dp(k["pos"]), _ip(k["type"]), _dp(k["vel"])
    s.acc = _dp(k["acc"]) if "acc" in k else None
    s.n_chain_blocks = len(k["chain"])
    s.chain, s.chain_const = _ip(k["chain"]), _dp(k["chain_const"])
    s.fconst, s.uconst = _dp(k["fconst"]), _dp(k["uconst"])
    s.box[0], s.box[1], s.box[2] = [float(x) for x in d["box"]]
    s.cutoff, s.dt, s.gamma, s.temperature = d["cutoff"], d["dt"], d["gamma"], d["temperature"]
    s.delta_lxy, s.tension = d["delta_lxy"], d["tension"]
    s.seed = int(d["seed"]) & 0xFFFFFFFFFFFFFFFF
    s.n_bonds, s.bond, s.bond_const = len(k["bond"]), _ip(k["bond"]), _dp(k["bond_const"])
    s.n_bends, s.bend, s.bend_const = len(k["bend"]), _ip(k["bend"]), _dp(k["bend_const"])
    s.n_balls, s.ball, s.ball_const = len(k["ball"]), _ip(k["ball"]), _dp(k["ball_const"])
    s.sigma_temperature = float(d.get("sigma_temperature", 0.0))
    return s


class Solver:
    """Device-resident MD system(s) driven through the C ABI.

    One system (`Solver(...)`, `from_system`, `from_mpd`) or R replica systems stepped
    together (`Solver.batch([...])`).  Per-replica results come back with a leading
    replica axis; for a single system that axis is dropped.
    """

    def __init__(self, pos, type_, vel, box, fconst, uconst, n_types, chain, chain_const,
                 cutoff=2.0, dt=0.02, gamma=1.0, temperature=3.0, delta_lxy=0.0, tension=0.0,
                 seed=0, acc=None):
        d = dict(pos=pos, type=type_, vel=vel, box=box, fconst=fconst, uconst=uconst,
                 n_types=n_types, chain=chain, chain_const=chain_const, cutoff=cutoff, dt=dt,
                 gamma=gamma, temperature=temperature, delta_lxy=delta_lxy, tension=tension,
                 seed=seed, acc=acc)
        self._create([d])

    def _create(self, systems: list):
        self._keep = []
        arr = (SmSystem * len(systems))(*[_pack(d, self._keep) for d in systems])
        self.h = C.c_void_p()
        if len(systems) == 1:
            _check(lib().sm_create(arr, C.byref(self.h)))
        else:
            _check(lib().sm_create_batch(arr, len(systems), C.byref(self.h)))
        self.nrep = len(systems)
        self.first = np.zeros(self.nrep, np.int32)
        self.count = np.zeros(self.nrep, np.int32)
        _check(lib().sm_replicas(self.h, None, _ip(self.first), _ip(self.count)))
        self.n = int(self.count.sum())

    @classmethod
    def from_system(cls, d: dict, **over) -> "Solver":
        d = dict(d)
        d.update(over)
        s = cls.__new__(cls)
        s._create([d])
        return s

    @classmethod
    def from_mpd(cls, m: Mpd, **over) -> "Solver":
        return cls.from_system(m.system(), **over)

    @classmethod
    def batch(cls, systems: list) -> "Solver":
        """R replica systems sharing constants, cutoff and dt, stepped together."""
        s = cls.__new__(cls)
        s._create([dict(d) for d in systems])
        return s

    def _one(self, a):
        return a[0] if self.nrep == 1 else a

    def close(self):
        if getattr(self, "h", None):
            lib().sm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- the step ----
    def init_forces(self):
        _check(lib().sm_init_forces(self.h))

    def second_half(self):
        _check(lib().sm_second_half(self.h))

    def step_first(self, step: int):
        _check(lib().sm_step_first(self.h, step))

    def step_finish(self, step: int, measure: bool = False):
        if measure:
            out = np.zeros((self.nrep, 3))
            _check(lib().sm_step_finish(self.h, step, _dp(out)))
            return self._one(out)
        _check(lib().sm_step_finish(self.h, step, None))
        return None

    def run(self, first_step: int, nsteps: int):
        _check(lib().sm_run(self.h, first_step, nsteps))

    def sync(self):
        _check(lib().sm_sync(self.h))

    def set_temperature(self, T: float):
        _check(lib().sm_set_temperature(self.h, T))

    def mc_stats(self):
        t, a = np.zeros(self.nrep, np.int64), np.zeros(self.nrep, np.int64)
        _check(lib().sm_mc_stats(self.h, _lp(t), _lp(a)))
        if self.nrep == 1:
            return int(t[0]), int(a[0])
        return t, a

    def mc_hist(self, replica: int = 0):
        """(accepted, rejected) trial histograms over the box fluctuation."""
        a, r, nb = np.zeros(256), np.zeros(256), C.c_int32()
        _check(lib().sm_mc_hist(self.h, replica, 256, _dp(a), _dp(r), C.byref(nb)))
        return a[:nb.value].copy(), r[:nb.value].copy()

    # ---- state ----
    def download(self):
        n = self.n
        pos, vel, acc, box = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((self.nrep, 3))
        _check(lib().sm_download_state(self.h, _dp(pos), _dp(vel), _dp(acc), _dp(box)))
        return pos, vel, acc, self._one(box)

    def upload(self, pos=None, vel=None, acc=None, box=None):
        cv = lambda a: np.ascontiguousarray(a, np.float64) if a is not None else None
        pos, vel, acc, box = cv(pos), cv(vel), cv(acc), cv(box)
        _check(lib().sm_upload_state(self.h, _dp(pos), _dp(vel), _dp(acc), _dp(box)))

    # ---- measurement and parity hooks ----
    def measure(self):
        """(U_pair, U_chain, K) per replica, FP64."""
        out = np.zeros((self.nrep, 3))
        _check(lib().sm_measure(self.h, _dp(out)))
        return self._one(out)

    def vesicle_radius(self, head_type: int = 2, noalloc: bool = False):
        """Mean distance of the head beads from their centre of mass, per replica, computed
        on the device (analysis/vesicleRadius.cpp:71-175).  noalloc: through the entry point
        that allocates no device memory (at most 1024 replicas)."""
        out = np.zeros(self.nrep)
        fn = lib().sm_vesicle_radius_noalloc if noalloc else lib().sm_vesicle_radius
        _check(fn(self.h, int(head_type), _dp(out)))
        return self._one(out)

    def box(self):
        """Box lengths per replica (3 each)."""
        box = np.zeros((self.nrep, 3))
        _check(lib().sm_download_state(self.h, None, None, None, _dp(box)))
        return self._one(box)

    def param_derivative(self, pair):
        """d(U_pair + U_chain)/d(Umin, Umax of `pair`, k_bond, k_bend) per replica, FP64, at
        the current state."""
        out = np.zeros((self.nrep, 4))
        _check(lib().sm_param_derivative(self.h, int(pair[0]), int(pair[1]), _dp(out)))
        return self._one(out)

    def bond_stats(self):
        """(mean bond length, mean bend cos theta, mean bend arm lengths a, b) per replica,
        as the reference's lBond_ and bend_ files give them (dataExtraction.h:861-930)."""
        out = np.zeros((self.nrep, 4))
        _check(lib().sm_bond_stats(self.h, _dp(out)))
        return self._one(out)

    def compute_forces(self):
        n = self.n
        fp, fc = np.zeros((n, 3)), np.zeros((n, 3))
        _check(lib().sm_compute_forces(self.h, _dp(fp), _dp(fc)))
        return fp, fc

    def cell_keys(self):
        k = np.zeros(self.n, np.int32)
        _check(lib().sm_cell_keys(self.h, _ip(k)))
        return k

    def pair_count(self):
        c = np.zeros(self.n, np.int32)
        t = C.c_int64()
        _check(lib().sm_pair_count(self.h, _ip(c), C.byref(t)))
        return c, t.value

    def force_pair_count(self):
        """Per particle: the pair force kernel's own accepted partners with a nonzero force
        term (its culling and FP64 test), and the number of such pairs."""
        c = np.zeros(self.n, np.int32)
        t = C.c_int64()
        _check(lib().sm_force_pair_count(self.h, _ip(c), C.byref(t)))
        return c, t.value

    def pair_list(self):
        _, total = self.pair_count()
        cap = total + 16
        buf = np.zeros((cap, 2), np.int32)
        nout = C.c_int64()
        _check(lib().sm_pair_list(self.h, _ip(buf), cap, C.byref(nout)))
        assert nout.value <= cap
        return buf[:nout.value]

    def potential_delta(self, scale):
        """U(old) - U(scaled) (pair, chain); scale is (3,) or (R, 3)."""
        s = np.ascontiguousarray(np.broadcast_to(np.asarray(scale, np.float64), (self.nrep, 3)))
        a, b = np.zeros(self.nrep), np.zeros(self.nrep)
        _check(lib().sm_potential_delta(self.h, _dp(s), _dp(a), _dp(b)))
        if self.nrep == 1:
            return float(a[0]), float(b[0])
        return a, b

    def tune(self, knob: str, value: int):
        _check(lib().sm_tune(self.h, knob.encode(), int(value)))

    def tune_get(self, knob: str) -> int:
        v = C.c_int32()
        _check(lib().sm_tune_get(self.h, knob.encode(), C.byref(v)))
        return v.value

    def autotune(self, steps: int = 64) -> float:
        """Pick the fastest launch shape on this context's own work; ms per step."""
        ms = C.c_double()
        _check(lib().sm_autotune(self.h, steps, C.byref(ms)))
        return ms.value

    # ---- profiling ----
    def profile(self, on: bool):
        _check(lib().sm_profile_enable(self.h, 1 if on else 0))

    def profile_read(self) -> dict:
        cap = 16
        names = (C.c_char_p * cap)()
        avg = np.zeros(cap)
        cnt = np.zeros(cap, np.int64)
        nout = C.c_int32()
        _check(lib().sm_profile_read(self.h, cap, names, _dp(avg), _lp(cnt), C.byref(nout)))
        return {names[i].decode(): (float(avg[i]), int(cnt[i])) for i in range(nout.value)}
