"""The torch-free ctypes binding (openfermion-fqe_amd/fqehip_ctypes.py):
hipMalloc'd buffers, tables built by the library, fqehip_apply.  It runs
in a child process that never imports torch, against the reference's own
golden outputs (FqeData(2,3,6), (2,1,4), (1,1,2) of
tests/unittest_data/fqe_data, tests/fqe_data_test.py:541-568) and against
the oracle on a norb=8 4a4b sector that takes the split path.

Tolerance: normwise relative error 1e-12 (complex128 throughout).
"""

import json
import os
import subprocess
import sys

import numpy
import pytest

from oracle import fqe_oracle as fo

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "fqe_data_golden.npz")

CHILD = r"""
import json, sys
import numpy
sys.path.insert(0, sys.argv[1])
import fqehip_ctypes as fh

def rel(a, b):
    return float(numpy.linalg.norm(a - b) / numpy.linalg.norm(b))

z = numpy.load(sys.argv[2])
x = numpy.load(sys.argv[3])
err = {}
for tag, (na, nb, norb) in (("020306", (2, 3, 6)), ("020104", (2, 1, 4)),
                            ("010102", (1, 1, 2))):
    c, h1, h2 = z["c_" + tag], z["h1_" + tag], z["h2_" + tag]
    err["12_" + tag] = rel(fh.apply_spatial12(c, h1, h2, na, nb, norb),
                           z["out12_" + tag])
    err["2_" + tag] = rel(fh.apply_spatial12(c, 0 * h1, h2, na, nb, norb),
                          z["out2_" + tag])
op = fh.PreparedOperator(x["h1"], x["h2"], 8)
s, keep = fh.sector(4, 4, 8, op)
out = fh.apply_spatial12(x["c"], x["h1"], x["h2"], 4, 4, 8, op=op)
err["split_norb8"] = rel(out, x["ref"])
print(json.dumps({"torch": "torch" in sys.modules, "split": bool(s.hbb_f64),
                  "err": err}))
"""


def test_ctypes_binding_without_torch(tmp_path):
    print("ctypes binding child process", flush=True)
    c, h1, h2 = fo.synthetic_workload(8, 4, 4, 3, 8)
    ref = fo.apply_12(c, h1, h2, fo.OracleGraph(4, 4, 8))
    inputs = tmp_path / "norb8.npz"
    numpy.savez(inputs, c=c, h1=h1, h2=h2, ref=ref)
    env = {k: v for k, v in os.environ.items() if k != "PYTHONPATH"}
    proc = subprocess.run(
        [sys.executable, "-c", CHILD, os.path.join(ROOT, "openfermion-fqe_amd"),
         GOLDEN, str(inputs)], capture_output=True, text=True, timeout=300,
        env=env)
    assert proc.returncode == 0, proc.stderr[-3000:]
    res = json.loads(proc.stdout.strip().splitlines()[-1])
    print(res, flush=True)
    assert res["torch"] is False
    assert res["split"] is True
    for key, e in res["err"].items():
        assert e < 1e-12, (key, e)
