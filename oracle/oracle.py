"""Test infrastructure: ctypes binding of the FP64 C oracle (liboracle.so) and a small
`.mpd` tokenizer for building its inputs.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this module,
and only as the checker; the product never does.  The `.mpd` grammar followed here is the
reference's token state machine (include/system.h:589-1313, include/fileFormats/
scriptFormat.h:79-94): whitespace-separated tokens, a key followed by its values.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

BOND, BEND, CHAIN, BALL = 6, 7, 8, 19  # molecule type ids (include/algorithms/molecules.h:14-16,28)


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


class OrcSys(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("ntypes", C.c_int32),
        ("pos", C.POINTER(C.c_double)), ("type", C.POINTER(C.c_int32)),
        ("vel", C.POINTER(C.c_double)), ("acc", C.POINTER(C.c_double)),
        ("nchain", C.c_int32), ("chain", C.POINTER(C.c_int32)),
        ("chain_const", C.POINTER(C.c_double)),
        ("fconst", C.POINTER(C.c_double)), ("uconst", C.POINTER(C.c_double)),
        ("box", C.c_double * 3),
        ("cutoff", C.c_double), ("dt", C.c_double), ("gamma", C.c_double),
        ("temp", C.c_double), ("delta_lxy", C.c_double), ("tension", C.c_double),
        ("seed", C.c_uint32),
        ("nbond", C.c_int32), ("bond", C.POINTER(C.c_int32)), ("bond_const", C.POINTER(C.c_double)),
        ("nbend", C.c_int32), ("bend", C.POINTER(C.c_int32)), ("bend_const", C.POINTER(C.c_double)),
        ("nball", C.c_int32), ("ball", C.POINTER(C.c_int32)), ("ball_const", C.POINTER(C.c_double)),
    ]


class OrcMT(C.Structure):
    _fields_ = [("s", C.c_uint32 * 624), ("left", C.c_int), ("next", C.c_int)]


def _declare(L):
    P = C.POINTER
    d, i32, u32, u64, i64 = C.c_double, C.c_int32, C.c_uint32, C.c_uint64, C.c_int64
    sig = {
        "orc_set_threads": (None, [i32]),
        "orc_mt_seed": (None, [P(OrcMT), u32]),
        "orc_mt_u32": (u32, [P(OrcMT)]),
        "orc_mt_rand53": (d, [P(OrcMT)]),
        "orc_philox4x32_10": (None, [P(u32), P(u32), P(u32)]),
        "orc_philox_langevin": (None, [u64, u64, u32, P(d)]),
        "orc_philox_mc": (None, [u64, u64, P(d), P(d)]),
        "orc_cell_grid": (None, [P(d), d, P(i32), P(d)]),
        "orc_cell_keys": (None, [P(OrcSys), P(i32)]),
        "orc_pair_force": (None, [P(OrcSys), P(d)]),
        "orc_pair_count": (None, [P(OrcSys), P(i32)]),
        "orc_pair_count_active": (None, [P(OrcSys), P(i32)]),
        "orc_pair_list": (i64, [P(OrcSys), P(i32), i64]),
        "orc_pair_potential": (d, [P(OrcSys)]),
        "orc_pair_dpotential": (d, [P(OrcSys), P(d)]),
        "orc_chain_force": (None, [P(OrcSys), P(d)]),
        "orc_chain_potential": (d, [P(OrcSys)]),
        "orc_chain_dpotential": (d, [P(OrcSys), P(d)]),
        "orc_langevin_mt": (None, [P(OrcSys), P(OrcMT), i32, d]),
        "orc_langevin_philox": (None, [P(OrcSys), u64, d]),
        "orc_verlet_first": (None, [P(OrcSys)]),
        "orc_verlet_second": (None, [P(OrcSys)]),
        "orc_kinetic": (d, [P(OrcSys)]),
        "orc_vesicle_radius": (d, [P(d), P(i32), i32, P(d), i32]),
        "orc_md_create": (C.c_void_p, [P(OrcSys), i32, i32]),
        "orc_md_destroy": (None, [C.c_void_p]),
        "orc_md_init_forces": (None, [C.c_void_p]),
        "orc_md_run": (i32, [C.c_void_p, i32, i32]),
        "orc_md_trials": (i32, [C.c_void_p]),
        "orc_md_decisions": (C.c_uint64, [C.c_void_p]),
        "orc_md_objective": (d, [C.c_void_p, i32]),
        "orc_md_measure": (None, [C.c_void_p, P(d), P(d)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


# ---------------------------------------------------------------------------------------
# .mpd tokenizer (test infrastructure; the product has its own C++ reader)
# ---------------------------------------------------------------------------------------
_NCONST = {6: 2, 7: 2, 8: 4, 19: 2}    # BOND, BEND, CHAIN, BALL constants (system.h:1049-1110)
_NPART = {6: 2, 7: 3, 8: 3, 19: 2}     # indices per record


def read_mpd(path: str) -> dict:
    """Parse the keys a lipid MD system uses.  Molecule records of types other than
    BOND/BEND/CHAIN are rejected."""
    with open(path) as f:
        tok = f.read().split()
    out: dict = {"present": []}
    i = 0

    def nxt():
        nonlocal i
        t = tok[i]
        i += 1
        return t

    scalars_f = {"gamma", "initialTemp", "finalTemp", "cutoff", "initialTime", "finalTime",
                 "deltaT", "storeInterval", "measureInterval", "deltaLXY", "removeSolvent",
                 "tempStepInterval", "tension"}
    scalars_i = {"seed", "nTypes", "nMolecules", "nParticles"}
    while i < len(tok):
        key = nxt()
        out["present"].append(key)
        if key in scalars_f:
            out[key] = float(nxt())
        elif key in scalars_i:
            out[key] = int(nxt())
        elif key == "periodic":
            nxt()
            out[key] = True
        elif key == "size":
            out[key] = [float(nxt()) for _ in range(3)]
        elif key in ("twoBodyFconst", "twoBodyUconst"):
            nt = out["nTypes"]
            out[key] = np.array([float(nxt()) for _ in range(6 * nt * nt)])
        elif key == "positions":
            n = out["nParticles"]
            vals = [nxt() for _ in range(4 * n)]
            out["type"] = np.array([int(v) for v in vals[0::4]], dtype=np.int32)
            out["pos"] = np.array([[float(vals[4 * k + d]) for d in (1, 2, 3)] for k in range(n)])
        elif key == "velocities":
            n = out["nParticles"]
            out["vel"] = np.array([float(nxt()) for _ in range(3 * n)]).reshape(n, 3)
        elif key == "molecule":
            mols = []
            for _ in range(out["nMolecules"]):
                mtype = int(nxt())
                if mtype not in _NCONST:
                    raise ValueError(f"molecule type {mtype} not supported by the oracle")
                nb = int(nxt())
                consts = [float(nxt()) for _ in range(_NCONST[mtype])]
                recs = [[int(nxt()) for _ in range(_NPART[mtype])] for _ in range(nb)]
                mols.append({"type": mtype, "const": consts, "bonds": recs})
            out["molecule"] = mols
        elif key == "banana":
            pass
        elif key == "solventGamma":
            out[key] = [float(nxt()), float(nxt())]
        elif key == "gammaType":
            out[key] = [float(nxt()) for _ in range(out["nTypes"])]
        else:
            raise ValueError(f"unrecognised .mpd key {key!r}")
    return out


class System:
    """Numpy-owned system the C oracle operates on in place."""

    def __init__(self, pos, type_, vel, box, fconst, uconst, ntypes, chains, chain_const,
                 cutoff=2.0, dt=0.02, gamma=1.0, temp=3.0, delta_lxy=0.0, tension=0.0,
                 seed=0, bonds=None, bond_const=None, bends=None, bend_const=None, balls=None, ball_const=None):
        self.pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3).copy()
        self.type = np.ascontiguousarray(type_, dtype=np.int32).copy()
        self.vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3).copy()
        self.acc = np.zeros_like(self.pos)
        self.fconst = np.ascontiguousarray(fconst, dtype=np.float64).copy()
        self.uconst = np.ascontiguousarray(uconst, dtype=np.float64).copy()
        self.chain = np.ascontiguousarray(chains, dtype=np.int32).reshape(-1, 3).copy()
        self.chain_const = np.ascontiguousarray(chain_const, dtype=np.float64).reshape(-1, 4).copy()
        # generic BOND / BEND records, two constants per record
        self.bond = np.ascontiguousarray(bonds if bonds is not None else [], dtype=np.int32).reshape(-1, 2).copy()
        self.bond_const = np.ascontiguousarray(bond_const if bond_const is not None else [],
                                               dtype=np.float64).reshape(-1, 2).copy()
        self.bend = np.ascontiguousarray(bends if bends is not None else [], dtype=np.int32).reshape(-1, 3).copy()
        self.bend_const = np.ascontiguousarray(bend_const if bend_const is not None else [],
                                               dtype=np.float64).reshape(-1, 2).copy()
        self.ball = np.ascontiguousarray(balls if balls is not None else [], dtype=np.int32).reshape(-1, 2).copy()
        self.ball_const = np.ascontiguousarray(ball_const if ball_const is not None else [],
                                               dtype=np.float64).reshape(-1, 2).copy()
        assert len(self.bond_const) == len(self.bond) and len(self.bend_const) == len(self.bend)
        assert len(self.ball_const) == len(self.ball)
        self.s = OrcSys()
        s = self.s
        s.n = len(self.type)
        s.ntypes = ntypes
        s.pos, s.type, s.vel, s.acc = _dp(self.pos), _ip(self.type), _dp(self.vel), _dp(self.acc)
        s.nchain = len(self.chain)
        s.chain, s.chain_const = _ip(self.chain), _dp(self.chain_const)
        s.fconst, s.uconst = _dp(self.fconst), _dp(self.uconst)
        s.box[0], s.box[1], s.box[2] = box
        s.cutoff, s.dt, s.gamma, s.temp = cutoff, dt, gamma, temp
        s.delta_lxy, s.tension, s.seed = delta_lxy, tension, seed & 0xFFFFFFFF
        s.nbond, s.bond, s.bond_const = len(self.bond), _ip(self.bond), _dp(self.bond_const)
        s.nbend, s.bend, s.bend_const = len(self.bend), _ip(self.bend), _dp(self.bend_const)
        s.nball, s.ball, s.ball_const = len(self.ball), _ip(self.ball), _dp(self.ball_const)

    @classmethod
    def from_mpd(cls, path: str) -> "System":
        m = read_mpd(path)
        chains, cc = [], []
        rec = {BOND: ([], []), BEND: ([], []), BALL: ([], [])}
        for mol in m.get("molecule", []):
            if mol["type"] == CHAIN:
                for r in mol["bonds"]:
                    chains.append(r)
                    cc.append(mol["const"])
            elif mol["type"] == BALL:
                # the force takes the centre from the first record (system.h:1944)
                for r in mol["bonds"]:
                    rec[BALL][0].append([mol["bonds"][0][0], r[1]])
                    rec[BALL][1].append(mol["const"])
            elif mol["type"] in rec:
                for r in mol["bonds"]:
                    rec[mol["type"]][0].append(r)
                    rec[mol["type"]][1].append(mol["const"])
            else:
                raise ValueError("oracle System supports CHAIN, BOND and BEND molecules only")
        return cls(m["pos"], m["type"], m["vel"], m["size"], m["twoBodyFconst"],
                   m["twoBodyUconst"], m["nTypes"], np.array(chains).reshape(-1, 3),
                   np.array(cc).reshape(-1, 4), cutoff=m["cutoff"], dt=m["deltaT"],
                   gamma=m.get("gamma", 0.0), temp=m["initialTemp"],
                   delta_lxy=m.get("deltaLXY", 0.0), tension=m.get("tension", 0.0),
                   seed=m["seed"], bonds=rec[BOND][0], bond_const=rec[BOND][1],
                   bends=rec[BEND][0], bend_const=rec[BEND][1], balls=rec[BALL][0], ball_const=rec[BALL][1])

    @property
    def box(self):
        return np.array(list(self.s.box))

    @property
    def n(self):
        return int(self.s.n)

    def ref(self):
        return C.byref(self.s)

    # --- wrapped oracle calls -----------------------------------------------------
    def cell_grid(self):
        nc = np.zeros(3, np.int32)
        cs = np.zeros(3)
        box = np.array(list(self.s.box))
        lib().orc_cell_grid(_dp(box), self.s.cutoff, _ip(nc), _dp(cs))
        return nc, cs

    def cell_keys(self):
        k = np.zeros(self.n, np.int32)
        lib().orc_cell_keys(self.ref(), _ip(k))
        return k

    def pair_force(self):
        a = np.zeros((self.n, 3))
        lib().orc_pair_force(self.ref(), _dp(a))
        return a

    def pair_count(self):
        c = np.zeros(self.n, np.int32)
        lib().orc_pair_count(self.ref(), _ip(c))
        return c

    def pair_count_active(self):
        """Per particle: partners within the cutoff whose force term is not identically zero
        (r^2 < r_m^2, or a pair row with a nonzero outer coefficient)."""
        c = np.zeros(self.n, np.int32)
        lib().orc_pair_count_active(self.ref(), _ip(c))
        return c

    def pair_list(self):
        cap = int(self.pair_count().sum()) // 2 + 16
        buf = np.zeros((cap, 2), np.int32)
        n = lib().orc_pair_list(self.ref(), _ip(buf), cap)
        assert n <= cap
        return buf[:n]

    def pair_potential(self):
        return lib().orc_pair_potential(self.ref())

    def pair_dpotential(self, scale):
        s = np.ascontiguousarray(scale, dtype=np.float64)
        return lib().orc_pair_dpotential(self.ref(), _dp(s))

    def chain_force(self):
        a = np.zeros((self.n, 3))
        lib().orc_chain_force(self.ref(), _dp(a))
        return a

    def chain_potential(self):
        return lib().orc_chain_potential(self.ref())

    def chain_dpotential(self, scale):
        s = np.ascontiguousarray(scale, dtype=np.float64)
        return lib().orc_chain_dpotential(self.ref(), _dp(s))

    def kinetic(self):
        return lib().orc_kinetic(self.ref())


class MD:
    """The MD.cpp loop on a System (rng_mode 0: reference MT19937 streams, 1: counter RNG)."""

    def __init__(self, system: System, rng_mode: int = 1, mt_streams: int = 1):
        self.sys = system
        self.h = lib().orc_md_create(system.ref(), rng_mode, mt_streams)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_md_destroy(self.h)
            self.h = None

    def init_forces(self):
        lib().orc_md_init_forces(self.h)

    def run(self, start: int, end: int) -> int:
        return lib().orc_md_run(self.h, start, end)

    def trials(self) -> int:
        return lib().orc_md_trials(self.h)

    def measure(self):
        U, K = C.c_double(), C.c_double()
        lib().orc_md_measure(self.h, C.byref(U), C.byref(K))
        return U.value, K.value

    def decisions(self) -> int:
        """Accept (1) / reject (0) bits of the MC trials so far, the latest in bit 0."""
        return int(lib().orc_md_decisions(self.h))

    def objective(self, nsteps: int) -> float:
        """Init forces, steps 1..nsteps; mean U_pair + U_chain after each kick."""
        return lib().orc_md_objective(self.h, nsteps)


def vesicle_radius(pos, type_, box, head: int = 2) -> float:
    """One frame of analysis/vesicleRadius.cpp:71-175 (orc_vesicle_radius)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    type_ = np.ascontiguousarray(type_, dtype=np.int32)
    box = np.ascontiguousarray(box, dtype=np.float64)
    return lib().orc_vesicle_radius(_dp(pos), _ip(type_), len(type_), _dp(box), int(head))


def bond_stats(pos, box, bonds, bends):
    """dataExtraction::compute's bond and bend statistics (dataExtraction.h:861-930,
    1009-1027): mean bond length (0/0 without bonds), then the bends' cos theta and arm
    lengths, divided by the bend count only when it exceeds 1.  Differences are imaged
    with >= / <= against half the box."""
    pos = np.asarray(pos, np.float64)
    box = np.asarray(box, np.float64)

    def img(d):
        d = d.copy()
        for k in range(3):
            d[..., k] = np.where(d[..., k] >= box[k] / 2.0, d[..., k] - box[k], d[..., k])
            d[..., k] = np.where(d[..., k] <= -box[k] / 2.0, d[..., k] + box[k], d[..., k])
        return d

    bonds = np.asarray(bonds, np.int64).reshape(-1, 2)
    bends = np.asarray(bends, np.int64).reshape(-1, 3)
    with np.errstate(invalid="ignore", divide="ignore"):
        lb = np.float64(np.linalg.norm(img(pos[bonds[:, 0]] - pos[bonds[:, 1]]), axis=1).sum()) / np.float64(len(bonds))
    da = img(pos[bends[:, 0]] - pos[bends[:, 1]])
    db = img(pos[bends[:, 1]] - pos[bends[:, 2]])
    ra, rb = np.linalg.norm(da, axis=1), np.linalg.norm(db, axis=1)
    out = np.array([lb, ((da * db).sum(1) / (ra * rb)).sum(), ra.sum(), rb.sum()])
    if len(bends) > 1:
        out[1:] /= len(bends)
    return out


def set_threads(n: int):
    lib().orc_set_threads(n)


def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def philox_langevin(seed, step, particle):
    out = (C.c_double * 3)()
    lib().orc_philox_langevin(seed, step, particle, out)
    return list(out)


def philox_mc(seed, step):
    a, b = C.c_double(), C.c_double()
    lib().orc_philox_mc(seed, step, C.byref(a), C.byref(b))
    return a.value, b.value


def mt_stream(seed):
    m = OrcMT()
    lib().orc_mt_seed(C.byref(m), seed & 0xFFFFFFFF)
    return m


def mt_rand53(m):
    return lib().orc_mt_rand53(C.byref(m))


def mt_u32(m):
    return lib().orc_mt_u32(C.byref(m))
