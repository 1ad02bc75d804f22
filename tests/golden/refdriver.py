"""Runs the flang-built reference (oracle/_ref/ref_driver) on a mesh.

Used only by make_golden.py, in the build container where the reference
sources exist, to write golden fixtures.  GPU runs never execute it.
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, os.path.join(ROOT, "freecappuccino-dev_amd"))
from fcp.mesh import write_native  # noqa: E402

DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")


def ensure_built() -> None:
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle", "ref")], check=True)


class RefCase:
    """A temporary case directory holding polyMesh/ plus job inputs/outputs."""

    def __init__(self, mesh=None, foam_dir=None):
        ensure_built()
        self.tmp = tempfile.TemporaryDirectory(prefix="fcpref_")
        self.dir = self.tmp.name
        self.fmt = "native"
        if foam_dir is not None:
            os.symlink(os.path.abspath(foam_dir), os.path.join(self.dir, "polyMesh"))
            self.fmt = "foam"
        else:
            write_native(mesh, self.dir)
        self.mesh = mesh

    def put(self, name, arr):
        np.ascontiguousarray(arr, dtype=np.float64).tofile(os.path.join(self.dir, name))

    def get(self, name, dtype=np.float64):
        return np.fromfile(os.path.join(self.dir, name), dtype=dtype)

    def run(self, jobs):
        with open(os.path.join(self.dir, "jobs.txt"), "w") as f:
            for j in jobs:
                f.write(j + "\n")
        r = subprocess.run([DRIVER, self.fmt, "jobs.txt"], cwd=self.dir,
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError("ref_driver failed:\n" + r.stdout[-4000:])
        return r.stdout

    def dump(self):
        self.run(["dump"])
        g = {}
        for k in ("xc", "yc", "zc", "vol", "arx", "ary", "arz", "xf", "yf", "zf", "facint", "Df"):
            g[k] = self.get(f"g_{k}.bin")
        for k in ("dnw", "srdw", "dns", "srds"):
            p = os.path.join(self.dir, f"g_{k}.bin")
            g[k] = np.fromfile(p, dtype=np.float64) if os.path.exists(p) else np.zeros(0)
        c = {k: self.get(f"c_{k}.bin", np.int32) - 1 for k in ("ia", "ja", "diag", "slot_pn", "slot_np")}
        with open(os.path.join(self.dir, "sizes.txt")) as f:
            sizes = [int(t) for t in f.read().split()]
        return g, c, sizes

    def close(self):
        self.tmp.cleanup()
