"""The five CLI drivers (cg_crs, dcg_crs, iccg_crs, piccg_crs/picg_crs,
abmc_crs) on the GPU path against the reference's own binaries
(oracle/_ref, built from its sources, run with one OpenMP thread): same
argv, same output line; iteration counts within one, the printed residuals
equal to their last printed digit when the counts agree."""
import os
import re
import subprocess

import numpy as np
import pytest

import iccg_amd as I
from oracle import ref as R
from tests import problems as P

pytestmark = pytest.mark.gpu
BIN = os.path.join(P.ROOT, "iccg_amd", "bin")
LINE = re.compile(r"cg_time \[ms\]=(\d+), converged=(\d), iters=(\d+), rel_resid=([^,]+), "
                  r"\|\|Ax-b\|\|/\|\|b\|\| = (\S+)")


def _write_mtx(path, A):
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real symmetric\n")
        low = [(i, A.colind[k], A.val[k]) for i in range(A.n)
               for k in range(A.rowptr[i], A.rowptr[i + 1]) if A.colind[k] <= i]
        f.write(f"{A.n} {A.n} {len(low)}\n")
        for i, j, v in low:
            f.write(f"{i + 1} {j + 1} {float(v)!r}\n")


def _run(exe, args):
    env = dict(os.environ, OMP_NUM_THREADS="1")
    out = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True,
                         env=env, timeout=120)
    assert out.returncode == 0, out.stderr
    m = LINE.search(out.stdout)
    assert m, out.stdout
    return int(m.group(2)), int(m.group(3)), float(m.group(4)), float(m.group(5))


def _compare(mine, ref):
    """Same convergence flag, iterations within one; when they agree, the
    printed recurrence residuals (6 significant digits) agree to the last
    printed digit, at tol 1e-8 and 1e-14 alike: the GPU's residual is the
    double-double oracle's bit for bit, which on these inputs matches the
    reference's 80-bit sums to every printed digit."""
    assert mine[0] == ref[0]
    assert abs(mine[1] - ref[1]) <= 1
    if mine[1] == ref[1]:
        assert abs(mine[2] - ref[2]) <= 2e-5 * abs(ref[2])


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    G = P.grid(40, 30, coef=(-1.0, 1.0))
    _write_mtx(d / "g.mtx", G)
    color, nc = I.greedy_color(G)
    I.write_color_file(d / "g.col", color, nc)
    bo, nb, bc, nc2 = I.abmc_blocking(G, 16)
    I.write_block_file(d / "g.blk", bo, nb)
    I.write_block_color_file(d / "g.bcol", bc, nc2)
    return d


@pytest.mark.parametrize("exe", ["cg_crs", "dcg_crs", "iccg_crs"])
@pytest.mark.parametrize("mat", ["saad", "grid"])
def test_unordered_variants(exe, mat, files):
    path = P.SAAD if mat == "saad" else files / "g.mtx"
    for tol in (1e-8, 1e-14):
        mine = _run(os.path.join(BIN, exe), [path, tol])
        if R.available() and os.path.exists(R.bin_path(exe)):
            _compare(mine, _run(R.bin_path(exe), [path, tol]))
        else:
            assert mine[0] == 1


def test_multicolor_variants(files):
    for exe in ("piccg_crs", "picg_crs"):
        mine = _run(os.path.join(BIN, exe), [files / "g.mtx", files / "g.col", 1e-8])
        if R.available():
            _compare(mine, _run(R.bin_path("piccg_crs"), [files / "g.mtx", files / "g.col", 1e-8]))
        assert mine[0] == 1 and mine[3] < 1e-7


def test_abmc_variant(files):
    args = [files / "g.mtx", files / "g.blk", files / "g.bcol", 1e-8]
    mine = _run(os.path.join(BIN, "abmc_crs"), args)
    if R.available():
        _compare(mine, _run(R.bin_path("abmc_crs"), args))
    assert mine[0] == 1
    mixed = _run(os.path.join(BIN, "abmc_crs"), args + ["--precision=mixed"])
    assert mixed[0] == 1 and mixed[3] < 1e-8


def test_cli_errors(files, tmp_path):
    out = subprocess.run([os.path.join(BIN, "cg_crs")], capture_output=True, text=True)
    assert out.returncode == 1 and "Usage:" in out.stderr
    bad = tmp_path / "bad.mtx"
    bad.write_text("hello\n")
    out = subprocess.run([os.path.join(BIN, "cg_crs"), str(bad)], capture_output=True, text=True)
    assert out.returncode != 0 and "Invalid MatrixMarket header" in out.stderr
