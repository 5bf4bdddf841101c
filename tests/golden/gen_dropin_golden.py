"""Fixtures of the four reference programs' default runs, for the drop-in program tests.

Runs the reference programs exactly as shipped (built unchanged from their single source files by
`make -C oracle refbin` into oracle/_ref/bin/) with no arguments, in a scratch directory, and keeps
from each run: the HDF5 layout (dataset names, shapes, F64LE type, storage layout), the /t and
coordinate arrays, the first and last lines of stdout / stderr, per frame and field the sums
{sum, sum|.|, max|.|}, SHA-256 digests of the first frame of every field, and the values of the
first and last frame at 1024 fixed cells.  The output files themselves (up to 2 GB) are not kept.

    python tests/golden/gen_dropin_golden.py [--reuse DIR] [--tags m2,m3,eh,ns]

--reuse DIR takes existing runs from DIR/<tag>/ (output file plus stdout.txt / stderr.txt)
instead of running the programs.

The M2 default run takes well over an hour on one core, and runs that differ at round-off level
part at t = 0.18 (a threshold decision in one cell) -- some of them then meet a time step of ~1e-9
at t = 0.23 and never finish.  For M2 the script therefore runs the program with line-buffered
stdout for LOG_ONLY_SECONDS, stops it, and keeps the log lines printed by then (one per stored
frame, t and dt to 7 digits, max|divB| to 12) -- the program writes its file only when it exits.
"""
import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from core_collapse_amd import h5io  # noqa: E402

PROGRAMS = {
    "m2": ("mhd2d_resistive_hlld_ct_gravity", "out2d_mhd.h5",
           ["/rho", "/P", "/Mach", "/u", "/v", "/vabs", "/Bx", "/By", "/Bz", "/divB"]),
    "m3": ("mhd3d_resistive_hlld_ct", "out3d_mhd.h5",
           ["/rho", "/P", "/Mach", "/u", "/v", "/w", "/vabs", "/Bx", "/By", "/Bz", "/divB"]),
    "eh": ("euler2d_gravity_hllc", "out2d_hllc.h5", ["/rho", "/P", "/Mach"]),
    "ns": ("navier_stokes_2d_hllc_poisson", "out_ns2d_poisson.h5",
           ["/rho", "/P", "/Mach", "/u", "/v", "/momx", "/momy", "/Phi"]),
}
NSAMPLE = 1024
LOG_ONLY = {"m2"}
LOG_ONLY_SECONDS = 900


def sample_cells(ncells):
    rng = np.random.default_rng(2024)
    return np.sort(rng.choice(ncells, size=min(NSAMPLE, ncells), replace=False))


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def run_reference(tag, workdir):
    prog = os.path.join(ROOT, "oracle", "_ref", "bin", PROGRAMS[tag][0])
    if not os.path.exists(prog):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "refbin"])
    os.makedirs(workdir, exist_ok=True)
    with open(os.path.join(workdir, "stdout.txt"), "w") as so, open(os.path.join(workdir, "stderr.txt"), "w") as se:
        if tag in LOG_ONLY:
            proc = subprocess.Popen(["stdbuf", "-oL", "-eL", prog], cwd=workdir, stdout=so, stderr=se)
            try:
                proc.wait(timeout=LOG_ONLY_SECONDS)
            except subprocess.TimeoutExpired:
                proc.kill()
                proc.wait()
        else:
            subprocess.check_call([prog], cwd=workdir, stdout=so, stderr=se)


def extract_log_only(tag, workdir):
    with open(os.path.join(workdir, "stdout.txt")) as f:
        out = f.read().splitlines()
    complete = bool(out) and out[-1].startswith("Wrote ")
    if not complete and out and not out[-1].endswith(tuple("0123456789")):
        out = out[:-1]  # a line cut off by the stop
    return {}, {"program": PROGRAMS[tag][0], "file": PROGRAMS[tag][1], "fields": PROGRAMS[tag][2],
                "complete": complete, "stdout": out}


def extract(tag, workdir):
    _, fname, fields = PROGRAMS[tag]
    path = os.path.join(workdir, fname)
    lay = h5io.info(path)
    t = h5io.read(path, "/t")
    arrays = {"t": t}
    for c in ("/x", "/y", "/z"):
        if c in lay:
            arrays[c[1:]] = h5io.read(path, c)
    nt = len(t)
    stats = np.zeros((nt, len(fields), 3))
    first, last = {}, {}
    cells = None
    for fi, name in enumerate(fields):
        for k, fr in enumerate(h5io.frames(path, name)):
            a = fr.reshape(-1)
            stats[k, fi] = (a.sum(), np.abs(a).sum(), np.abs(a).max())
            if cells is None:
                cells = sample_cells(a.size)
            if k == 0:
                first[name] = (digest(fr), a[cells].copy())
            if k == nt - 1:
                last[name] = a[cells].copy()
    arrays["stats"] = stats
    arrays["cells"] = cells
    arrays["first_samples"] = np.stack([first[n][1] for n in fields])
    arrays["last_samples"] = np.stack([last[n] for n in fields])
    with open(os.path.join(workdir, "stdout.txt")) as f:
        out = f.read().splitlines()
    with open(os.path.join(workdir, "stderr.txt")) as f:
        err = f.read().splitlines()
    meta = {
        "program": PROGRAMS[tag][0], "file": fname, "fields": fields,
        "layout": {k: {"shape": list(v["shape"]), "f64le": v["f64le"], "layout": v["layout"]} for k, v in lay.items()},
        "first_sha256": {n: first[n][0] for n in fields},
        "stdout_head": out[:3], "stdout_tail": out[-2:], "stdout_lines": len(out),
        "stderr_head": err[:3], "stderr_tail": err[-2:], "stderr_lines": len(err),
    }
    return arrays, meta


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reuse", help="directory holding <tag>/ runs of the reference programs")
    ap.add_argument("--tags", default="m2,m3,eh,ns")
    a = ap.parse_args()
    meta_path = os.path.join(HERE, "dropin.json")
    meta_all = json.load(open(meta_path)) if os.path.exists(meta_path) else {}
    for tag in a.tags.split(","):
        ext = extract_log_only if tag in LOG_ONLY else extract
        if a.reuse:
            arrays, meta = ext(tag, os.path.join(a.reuse, tag))
        else:
            with tempfile.TemporaryDirectory() as td:
                run_reference(tag, td)
                arrays, meta = ext(tag, td)
        if arrays:
            np.savez_compressed(os.path.join(HERE, "dropin_%s.npz" % tag), **arrays)
            print(tag, "frames", len(arrays["t"]))
        else:
            print(tag, "log lines", len(meta["stdout"]))
        meta_all[tag] = meta
    with open(meta_path, "w") as f:
        json.dump(meta_all, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
