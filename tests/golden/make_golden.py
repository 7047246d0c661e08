"""Write the golden fixtures in tests/golden/ from the compiled reference.

Runs only in the build container, where the reference sources exist: it builds
oracle/_ref/ with `make -C oracle ref` (the reference's own MD.cpp, generators and the
ref_dump harness compiled from the sources where they lie) and records what those
binaries print and write.  The fixtures are data only: generator checksums, step-0 FP64
forces/keys/neighbour counts/energies, and short single-thread MD trajectories.

usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import shutil
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(REPO, "oracle", "_ref")
SCALE = (1.0001, 1.0001, 0.99980004)  # a box trial scale for the dU fixtures

# name -> reference generator argv (after the name)
GEN = {
    "L500": ["liposome", "1234", "500", "3.45"],
    "B400": ["bilayer", "1234", "400", "3.45", "0", "0", "0", "0"],
    "L5000": ["liposome", "1234", "5000", "3.45"],
    "B4000": ["bilayer", "1234", "4000", "3.45", "0", "0", "0", "0"],
    "L80000": ["liposome", "1234", "80000", "3.45"],
    "L2000s7": ["liposome", "7", "2000", "3.45"],
}
STEP0 = ["L500", "B400"]


def run(cmd, cwd, env=None):
    e = dict(os.environ)
    e["OMP_NUM_THREADS"] = "1"
    if env:
        e.update(env)
    return subprocess.run(cmd, cwd=cwd, env=e, check=True, capture_output=True, text=True)


def parse_kv(text):
    out = {}
    for line in text.splitlines():
        if "=" in line:
            k, v = line.split("=", 1)
            vals = v.split()
            out[k] = [float(x) for x in vals] if len(vals) > 1 else float(vals[0])
    return out


def edit_mpd(path, **kv):
    """Replace scalar keys (or append them) in a .mpd written by the reference."""
    tok = open(path).read().split("\n")
    seen = set()
    for i, line in enumerate(tok):
        words = line.split()
        if words and words[0] in kv:
            tok[i] = f"{words[0]} {kv[words[0]]} "
            seen.add(words[0])
    extra = [f"{k} {v} " for k, v in kv.items() if k not in seen]
    open(path, "w").write("\n".join(tok[:1] + extra + tok[1:]))


def read_dat(path):
    return np.loadtxt(path, ndmin=2)


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    work = tempfile.mkdtemp(prefix="smgold_")
    try:
        sha, step0, traj = {}, {}, {}
        for name, argv in GEN.items():
            run([os.path.join(REF, argv[0]), name] + argv[1:], work)
            sha[name] = {"argv": argv,
                         "sha256": hashlib.sha256(open(os.path.join(work, name + ".mpd"), "rb").read()).hexdigest()}
        # step-0 quantities from the reference's own classes
        for name in STEP0 + ["L5000"]:
            r = run([os.path.join(REF, "ref_dump"), name, name] + [repr(x) for x in SCALE], work)
            kv = parse_kv(r.stdout)
            run([os.path.join(REF, "ref_dump_count"), name, name + "c"], work)
            rec = {"U_pair": kv["U_pair"], "U_chain": kv["U_chain"], "dU_pair": kv["dU_pair"],
                   "dU_chain": kv["dU_chain"], "nCells": kv["nCells"], "cellSize": kv["cellSize"]}
            if name in STEP0:
                rec["keys"] = np.fromfile(os.path.join(work, name + "_keys.i32"), np.int32)
                rec["fpair"] = np.fromfile(os.path.join(work, name + "_fpair.f64")).reshape(-1, 3)
                rec["fchain"] = np.fromfile(os.path.join(work, name + "_fchain.f64")).reshape(-1, 3)
                rec["counts"] = np.fromfile(os.path.join(work, name + "c_fpair.f64")).reshape(-1, 3)[:, 0].astype(np.int32)
            step0[name] = rec
        # short single-thread trajectories of the real MD binary.  The store at the last
        # step happens mid-step (after Verlet::first, MD.cpp:356,373-381).
        cases = {
            "L500": dict(steps=20, edits={}),
            "B400t": dict(steps=40, src="B400", edits={"tension": 1.0}),
        }
        for tname, c in cases.items():
            src = c.get("src", tname)
            d = os.path.join(work, "traj_" + tname)
            os.makedirs(d)
            shutil.copy(os.path.join(work, src + ".mpd"), os.path.join(d, tname + ".mpd"))
            T = c["steps"] * 0.02
            edit_mpd(os.path.join(d, tname + ".mpd"), finalTime=repr(T), storeInterval=repr(T),
                     measureInterval=repr(0.2), **c["edits"])
            shutil.copy(os.path.join(d, tname + ".mpd"), os.path.join(d, "input.mpd"))
            run([os.path.join(REF, "MD"), tname], d)
            rec = {"steps": c["steps"],
                   "input_mpd": open(os.path.join(d, "input.mpd")).read(),
                   "final_mpd": open(os.path.join(d, tname + ".mpd")).read()}
            for dat in ("potential", "kinetic", "size"):
                rec[dat] = read_dat(os.path.join(d, f"{dat}_{tname}.dat"))
            hist = os.path.join(d, f"resizeHist_{tname}.dat")
            if os.path.exists(hist):
                rec["resizeHist"] = read_dat(hist)
            traj[tname] = rec

        with open(os.path.join(HERE, "gen_sha256.json"), "w") as f:
            json.dump(sha, f, indent=1, sort_keys=True)
        arrays, scalars = {}, {}
        for name, rec in step0.items():
            scalars[name] = {k: v for k, v in rec.items() if not isinstance(v, np.ndarray)}
            for k, v in rec.items():
                if isinstance(v, np.ndarray):
                    arrays[f"{name}/{k}"] = v
        scalars["scale"] = list(SCALE)
        with open(os.path.join(HERE, "ref_step0.json"), "w") as f:
            json.dump(scalars, f, indent=1, sort_keys=True)
        np.savez_compressed(os.path.join(HERE, "ref_step0.npz"), **arrays)
        tarr, tmeta = {}, {}
        for name, rec in traj.items():
            tmeta[name] = {"steps": rec["steps"]}
            for k, v in rec.items():
                if isinstance(v, np.ndarray):
                    tarr[f"{name}/{k}"] = v
                elif isinstance(v, str):
                    tarr[f"{name}/{k}"] = np.frombuffer(v.encode(), dtype=np.uint8)
        with open(os.path.join(HERE, "ref_traj.json"), "w") as f:
            json.dump(tmeta, f, indent=1, sort_keys=True)
        np.savez_compressed(os.path.join(HERE, "ref_traj.npz"), **tarr)
        print("fixtures written to", HERE)
    finally:
        shutil.rmtree(work, ignore_errors=True)


if __name__ == "__main__":
    main()
