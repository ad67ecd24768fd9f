"""The absorption() drop-in on the fatal inputs of tests/fault_cases.py,
each in a child process: the same fatal message and exit code as the
reference's own paths (pinned on the CPU by tests/test_faults.py)."""
import json
import os
import subprocess
import sys

import pytest

import fault_cases as FC

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import ctypes as C, json, sys
sys.path.insert(0, sys.argv[2])
from nise_2017_amd import abi
p = json.loads(sys.argv[1])
L = C.CDLL(abi.LIB_PATH)
L.absorption.argtypes = [C.POINTER(abi.TNon)]
L.absorption.restype = None
non = abi.make_non(singles=p["N"], tmax1=p["tmax"], sample=p["sample"], begin=p["begin"], end=p["end"], ts=1,
                   propagation=1, deltat=p["deltat"], min1=p["min1"], max1=p["max1"], length=p["length"],
                   energyFName=p["hfile"], dipoleFName=p["dfile"])
L.absorption(C.byref(non))
print("RUN COMPLETED")
"""


@pytest.mark.parametrize("name", sorted(FC.CASES))
def test_dropin_fatal_paths(name, tmp_path):
    params = FC.write_case(str(tmp_path), name)
    print(f"[faults] {name}", flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD, json.dumps(params), ROOT], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=120)
    _, msg, code = FC.CASES[name]
    assert "RUN COMPLETED" not in r.stdout
    assert r.returncode == code, (r.returncode, r.stdout, r.stderr)
    assert msg in FC.lines(r.stdout), r.stdout
