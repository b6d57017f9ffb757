"""svd_main: the reference's file contract (args.txt "M N", input/A_gm.bin ->
output/sigma.bin, U.bin, Vt.bin; main.cpp:20-24,64,169-174) end to end on the GPU."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from svdcheck import measures

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "paco-svd_amd", "svd_main")


@pytest.mark.parametrize("shape", [(128, 96), (33, 33), (700, 300)])
def test_driver_file_contract(tmp_path, shape):
    print("[case] svd_main", shape, flush=True)
    sys.stdout.flush()
    M, N = shape
    if shape == (128, 96):
        A = np.load(os.path.join(ROOT, "tests", "golden", "c1.npz"))["A"]
    else:
        A = np.random.default_rng(M + N).standard_normal(shape).astype(np.float32)
    (tmp_path / "build").mkdir()
    (tmp_path / "input").mkdir()
    (tmp_path / "output").mkdir()
    (tmp_path / "args.txt").write_text("%d %d\n" % (M, N))
    A.astype(np.float32).tofile(tmp_path / "input" / "A_gm.bin")
    # default paths relative to the working directory, as the reference's build dir
    r = subprocess.run([DRIVER], cwd=tmp_path / "build", capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    s = np.fromfile(tmp_path / "output" / "sigma.bin", dtype=np.float32)
    U = np.fromfile(tmp_path / "output" / "U.bin", dtype=np.float32).reshape(M, M)
    Vt = np.fromfile(tmp_path / "output" / "Vt.bin", dtype=np.float32).reshape(N, N)
    m = measures(A, s, U, Vt)
    print("[case] measures", m, flush=True)
    assert m["sigma_err"] <= 1e-5 and m["orth_u"] <= 1e-5 and m["orth_v"] <= 1e-5 and m["resid_ratio"] <= 10
    s_orc, _, _ = oracle.svd(A.astype(np.float64))
    assert np.max(np.abs(s - s_orc)) <= 1e-5 * s_orc[0]


def test_driver_rejects_m_less_than_n(tmp_path):
    (tmp_path / "build").mkdir()
    (tmp_path / "args.txt").write_text("3 5\n")
    r = subprocess.run([DRIVER], cwd=tmp_path / "build", capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "M must be greater than N" in r.stderr
