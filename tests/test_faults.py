"""The reference's own fatal paths for the absorption inputs of
tests/fault_cases.py (open_files, control, determine_samples,
read_Hamiltonian, read_dipole from its sources, oracle/_ref), each in a
child process: the fatal message and the exit code the GPU drop-in
must reproduce."""
import json
import os
import subprocess
import sys

import pytest

import fault_cases as FC

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import ctypes as C, json, sys
import numpy as np
p = json.loads(sys.argv[1])
L = C.CDLL(sys.argv[2])
f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
L.ref_absorption_checked.argtypes = [C.c_int, C.c_char_p, C.c_char_p, C.c_float, C.c_float, C.c_float,
                                     C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, f32, f32]
re = np.zeros(p["tmax"], np.float32); im = np.zeros(p["tmax"], np.float32)
L.ref_absorption_checked(p["N"], p["hfile"].encode(), p["dfile"].encode(), p["min1"], p["max1"], p["deltat"],
                         p["tmax"], p["sample"], p["begin"], p["end"], p["length"], re, im)
print("RUN COMPLETED")
"""


@pytest.mark.parametrize("name", sorted(FC.CASES))
def test_reference_fatal_paths(orc, name, tmp_path):
    if not orc.have_ref():
        pytest.skip("reference sources not built here")
    params = FC.write_case(str(tmp_path), name)
    r = subprocess.run([sys.executable, "-c", CHILD, json.dumps(params), orc.REF_SO], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=120)
    _, msg, code = FC.CASES[name]
    assert "RUN COMPLETED" not in r.stdout
    assert r.returncode == code, (r.returncode, r.stdout, r.stderr)
    assert msg in FC.lines(r.stdout), r.stdout
