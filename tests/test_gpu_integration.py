"""INTEGRATION.md's C program, compiled against include/ and linked with
libsplatt.so, run on the GPU over the reference's small.tns fixture; its
MTTKRP value is checked against the oracle's long-double COO MTTKRP of the
same factors."""
import subprocess

import numpy as np
import pytest

from conftest import tensor_path
from test_host_cpu import build_integration_program
import oracle_py as orc

pytestmark = pytest.mark.gpu


def test_integration_program_runs(gpu_available, tmp_path):
    exe = build_integration_program(tmp_path)
    r = subprocess.run([exe, tensor_path("small.tns")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    lines = dict(l.split(" ", 1) for l in r.stdout.strip().splitlines() if " " in l)
    got = float(lines["mttkrp[0][0]"])
    coo = orc.Coo.read(tensor_path("small.tns"))
    dims = [int(i.max()) + 1 for i in coo.inds()]
    R = 16
    mats = [(1.0 / (1.0 + (np.arange(d * R) % 7))).reshape(d, R) for d in dims]
    want = orc.mttkrp_coo(coo, mats, 0)[0, 0]
    assert abs(got - want) <= 1e-6 * abs(want)
    assert "cpd done" in r.stdout
