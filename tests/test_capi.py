"""C-ABI library: loads, exports every declared symbol, host-side functions (no GPU)."""
import ctypes
import os
import re
import struct

import pytest

import oracle
import pacosvd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names += re.findall(r"^\s*(?:int|void|size_t)\s+\**(\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = pacosvd.lib()
    names = declared_functions()
    assert "paco_upper_bidiagonalization" in names and "paco_svd_DC" in names
    for n in names:
        assert hasattr(L, n), n
    assert set(pacosvd.EXPORTS) <= set(names)


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6, 7, 8, 96, 97, 98, 4096])
def test_tiling_blob_matches_reference_partition(N):
    blob = pacosvd.tiling_blob(N)
    assert len(blob) == pacosvd.lib().getTilingSize(N, 1)
    magic, version, n, cnt, off, nb = struct.unpack("<6I", blob[:24])
    assert magic == 0x4F434150 and version == 1 and n == N and off == 64 and nb == 1
    leaves = struct.unpack("<%dI" % (2 * cnt), blob[off:off + 8 * cnt])
    got = [(leaves[2 * i], leaves[2 * i + 1]) for i in range(cnt)]
    assert got == oracle.leaf_partition(N)


def test_workspace_size():
    L = pacosvd.lib()
    assert L.paco_workspace_size(10, 20) == 0
    sizes = [L.paco_workspace_size(n, n) for n in (1, 8, 64, 512, 4096)]
    assert all(s > 0 for s in sizes) and sizes == sorted(sizes)
    # at least the right secular vectors (N^2 floats)
    assert L.paco_workspace_size(4096, 4096) >= 4 * 4096 * 4096


def test_argument_errors_without_gpu():
    L = pacosvd.lib()
    z = ctypes.c_void_p(0)
    # M < N and null pointers are rejected before any device call
    assert L.paco_upper_bidiagonalization(40, z, 3, 5, z, z, z, z, z, z, z, z) == -1
    assert L.paco_svd_DC(1, z, 3, 5, z, z, z, z, z, z, z, z, z, z) == -1
    assert L.paco_svd_DC(1, z, 9000, 8000, z, z, z, z, z, z, z, z, z, z) == -1
    job = (pacosvd.SvdJob * 1)()
    job[0].M, job[0].N = 4, 8
    assert L.paco_svd_batched(job, 1, z) == -1
    assert L.paco_svd_plain(job, 1, z) == -1
    assert L.paco_set_precision(0) in (0, 1, 2)
