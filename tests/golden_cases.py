"""The golden cases under tests/golden/ (written by tests/golden/make_golden.py
from the reference solver's own outputs)."""
import os
import shutil

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RUN_CASES = ["shipped_coarse", "square", "square_super", "rb_jump", "circle", "square_cd", "periodic_channel",
             "periodic_open", "periodic_rb"]
MESH_CASES = ["mesh1", "mesh2", "mesh3", "mesh4"]
ALL_CASES = RUN_CASES + MESH_CASES + ["square_dbg"]


def input_dir(name):
    return os.path.join(GOLDEN, name, "input")


def ref_dir(name):
    return os.path.join(GOLDEN, name, "ref")


def fresh_input(name, tmp):
    """Copy of the case input in a scratch directory (solvers rewrite files)."""
    dst = os.path.join(str(tmp), name)
    shutil.copytree(input_dir(name), dst)
    return dst


def ref_field(name, field):
    return np.fromfile(os.path.join(ref_dir(name), field + ".bin"), dtype=np.float64)


def ref_iterations(name):
    rows = open(os.path.join(ref_dir(name), "Iterations_in_loops.txt")).read().split("\n")[1:]
    return [int(r.split()[2]) for r in rows if r.strip()]


def text_mesh(a):
    """A mesh array as WriteMassiveToFile_1D writes it (%.15e per line)."""
    return "".join("%.15e\n" % x for x in a)


def same_bits(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and bool(np.all(a.view(np.uint64) == b.view(np.uint64)))
