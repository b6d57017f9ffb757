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
    old = pacosvd.set_leaf_size(3)  # leaves of 3 columns: the reference's partition
    try:
        assert blob_leaves(N, 1, 3) == oracle.leaf_partition(N)
    finally:
        pacosvd.set_leaf_size(old)


def test_default_leaf_size():
    assert pacosvd.set_leaf_size(0) == 6


def blob_leaves(N, nb, ls):
    blob = pacosvd.tiling_blob(N, nb)
    assert len(blob) == pacosvd.lib().getTilingSize(N, nb)
    magic, version, n, cnt, off, nbl, lsz = struct.unpack("<7I", blob[:28])
    assert magic == 0x4F434150 and version == 2 and n == N and off == 64 and nbl == nb and lsz == ls
    leaves = struct.unpack("<%dI" % (2 * cnt), blob[off:off + 8 * cnt])
    return [(leaves[2 * i], leaves[2 * i + 1]) for i in range(cnt)]


@pytest.mark.parametrize("ls", [4, 5, 8, 16, 31, 32])
def test_tiling_blob_larger_leaves(ls):
    """Leaves of ls columns: they tile [0, N) in order; every leaf but the last has ls
    columns (ls - 1 for the one before a last leaf that starts a column early), so
    N // ls + 1 leaves for N >= ls; the last, square, leaf has 1 .. ls - 1 columns."""
    old = pacosvd.set_leaf_size(ls)
    try:
        assert pacosvd.set_leaf_size(0) == ls
        for N in list(range(1, 3 * ls + 3)) + [96, 97, 511, 512, 4096, 4097]:
            got = blob_leaves(N, 1, ls)
            assert got[0][0] == 0 and got[-1][1] == N, (N, got)
            assert all(a[1] == b[0] for a, b in zip(got, got[1:])), (N, got)
            assert all(0 < b - a for a, b in got)
            if N < ls:
                assert got == [(0, N)]
                continue
            assert len(got) == N // ls + 1
            assert all(b - a == ls for a, b in got[:-2])
            assert got[-2][1] - got[-2][0] == (ls - 1 if N % ls == 0 else ls)
            assert 1 <= got[-1][1] - got[-1][0] <= ls - 1
    finally:
        pacosvd.set_leaf_size(old)
    # out-of-range sizes only query
    assert pacosvd.set_leaf_size(2) == old and pacosvd.set_leaf_size(33) == old


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


def test_autotune_table_loader(tmp_path):
    """The library reads the autotuner's table at first use: leaf size, grid and resident
    thread entries are parsed (a fresh process, table named by PACO_AUTOTUNE_TABLE)."""
    import json
    import subprocess
    import sys
    t = tmp_path / "table.json"
    t.write_text(json.dumps({"512x512": 512, "leaf_size": 12, "grid_tiles:4096x4096": 12,
                             "grid_switch:8192x4096": 256, "res_threads:0": 256, "bogus": 7}))
    code = ("import sys; sys.path.insert(0, %r); import pacosvd; print(pacosvd.lib().paco_leaf_size())"
            % os.path.join(ROOT, "paco-svd_amd"))
    env = dict(os.environ, PACO_AUTOTUNE_TABLE=str(t))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "12"
    # the committed table keeps the built-in default leaf size
    committed = json.load(open(os.path.join(ROOT, "paco-svd_amd", "autotune_table.json")))
    assert committed.get("leaf_size", 6) == pacosvd.set_leaf_size(0)
