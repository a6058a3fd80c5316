"""CPU tests of the product's host code in libiccg_hip.so (no GPU calls):
readers, CRS construction, orderings, schedules, validators, generators.
Integer outputs must equal the oracle's bit for bit."""
import os

import numpy as np
import pytest

import iccg_amd as I
from oracle import oracle as O
from tests import problems as P


def _eq(a, b):
    return all(np.array_equal(getattr(a, f), getattr(b, f)) for f in ("rowptr", "colind", "val"))


def test_library_exports_every_header_symbol():
    import ctypes
    import re
    from iccg_amd import _capi
    hdr = open(os.path.join(P.ROOT, "include", "iccg_hip.h")).read()
    decl = set(re.findall(r"\b(iccg_[a-z0-9_]+)\s*\(", hdr))
    lib = ctypes.CDLL(_capi.LIB_PATH)
    missing = [s for s in sorted(decl) if not hasattr(lib, s)]
    assert not missing, missing
    assert decl == set(_capi.exported_symbols())


def test_mm_reader_matches_oracle_and_errors(tmp_path):
    A = I.read_matrix_market_crs(P.SAAD)
    assert _eq(A, O.read_matrix_market(P.SAAD))
    # general + pattern + comments inside the data + a duplicate entry
    f = tmp_path / "g.mtx"
    f.write_text("%%MatrixMarket matrix coordinate pattern general\n% c\n3 3 4\n1 1\n%x\n"
                 "2 3\n3 1\n2 3\n")
    A = I.read_matrix_market_crs(f)
    B = O.read_matrix_market(f)
    assert _eq(A, B) and list(A.val) == [1.0, 2.0, 1.0]
    bad = {"Not a MatrixMarket file": "%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 1\n",
           "Only 'matrix coordinate' is supported": "%%MatrixMarket matrix array real general\n",
           "Unsupported field: complex": "%%MatrixMarket matrix coordinate complex general\n",
           "Unsupported symmetry: hermitian": "%%MatrixMarket matrix coordinate real hermitian\n",
           "Unexpected EOF": "%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1\n",
           "Index out of range": "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n",
           "Invalid MatrixMarket header": "%%MatrixMarket matrix\n"}
    for msg, text in bad.items():
        f.write_text(text)
        with pytest.raises(I.IccgError, match=msg.replace("(", r"\(").replace(")", r"\)")):
            I.read_matrix_market_crs(f)
    with pytest.raises(I.IccgError, match="Failed to open"):
        I.read_matrix_market_crs(tmp_path / "missing.mtx")


@pytest.mark.parametrize("grid", [(17, 13, 1, (-1.0, 1.0), 0.0), (9, 8, 7, (-2.0, 2.0), 1e-6),
                                  (30, 30, 1, (0.0, 0.0), 0.0)])
def test_orderings_match_oracle(grid):
    nx, ny, nz, coef, sh = grid
    G = P.grid(nx, ny, nz, coef=coef, shift=sh)
    Go = P.orc(G)
    color, nc = I.greedy_color(G)
    p = I.BuildPermutationByColor(color)
    assert np.array_equal(p, O.perm_by_color(color))
    L = I.Permute_PtAP_LowerCRS_to_LowerCRS(G, p)
    assert _eq(L, O.ptap_lower(Go, p))
    F = I.expand_lower_to_full(L)
    assert _eq(F, O.expand_lower_to_full(O.ptap_lower(Go, p)))
    cp, rc = I.BuildColorPtrRows_FromPermutationAndColor(color, nc, p)
    cp2, rc2 = O.color_ptr_rows(color, nc, p)
    assert np.array_equal(cp, cp2) and np.array_equal(rc, rc2)
    for B in (3, 8, 64):
        bo, nb, bc, nc2 = I.abmc_blocking(G, B)
        q = I.BuildPermutationByBlockColor(bo, bc)
        assert np.array_equal(q, O.perm_by_block_color(bo, bc))
        s1 = I.BuildABMCSchedule(q, bo, bc, nb, nc2)
        s2 = O.abmc_schedule(q, bo, bc, nb, nc2)
        assert all(np.array_equal(a, b) for a, b in zip(s1, s2))


def test_coo_to_crs_matches_oracle():
    rng = np.random.default_rng(5)
    I_ = rng.integers(0, 50, 1000)
    J_ = rng.integers(0, 50, 1000)
    V_ = rng.standard_normal(1000)
    assert _eq(I.coo_to_crs(50, I_, J_, V_), O.coo_to_crs(50, I_, J_, V_))


def test_greedy_coloring_is_first_fit_and_valid():
    G = P.grid(11, 7, 5, coef=(-1.0, 1.0))
    color, nc = I.greedy_color(G)
    S = P.orc(G).to_scipy()
    for i in range(G.n):  # smallest color unused by lower-index neighbours
        nb = [j for j in S.indices[S.indptr[i]:S.indptr[i + 1]] if j < i]
        used = {color[j] for j in nb}
        assert color[i] == min(c for c in range(nc + 1) if c not in used)
    assert nc == 2  # grid graphs are bipartite


def test_abmc_blocking_valid_and_sized():
    G = P.grid(20, 18, 16, coef=(-2.0, 2.0), shift=1e-6)
    for B in (8, 64):
        bo, nb, bc, nc = I.abmc_blocking(G, B)
        sizes = np.bincount(bo, minlength=nb)
        assert sizes.max() <= B and sizes.min() >= 1
        tot, per = I.check_interblock(G, bo, bc, nb, nc)
        assert tot == 0
        tot2, _ = O.check_interblock(P.orc(G), bo, bc, nb, nc)
        assert tot2 == 0
        o = P.abmc_order(G, block=B)
        th, hz, sb, cb = I.check_hazards(o.A, o.new_of_old, bc[bo], bo)
        th2, hz2, sb2, cb2 = O.check_hazards(P.orc(o.A), o.new_of_old, bc[bo], bo)
        assert th == th2 and np.array_equal(hz, hz2) and np.array_equal(cb, cb2)
        assert cb.sum() == 0
    bo, nb, bc, nc = I.abmc_blocking_grid(16, 12, 8, 4, 4, 4)
    assert nb == 4 * 3 * 2 and nc == 2
    assert I.check_interblock(P.grid(16, 12, 8), bo, bc, nb, nc)[0] == 0


def test_file_formats_round_trip(tmp_path):
    G = P.grid(10, 9)
    color, nc = I.greedy_color(G)
    I.write_color_file(tmp_path / "a.col", color, nc)
    c2, nc2 = I.ReadColorFile_1Based(tmp_path / "a.col", G.n)
    assert np.array_equal(c2, color) and nc2 == nc
    bo, nb, bc, nc = I.abmc_blocking(G, 7)
    I.write_block_file(tmp_path / "a.blk", bo, nb)
    I.write_block_color_file(tmp_path / "a.bcol", bc, nc)
    b2, nb2 = I.ReadBlockFile_1Based(tmp_path / "a.blk", G.n)
    bc2, nc3 = I.ReadBlockColorFile_1Based(tmp_path / "a.bcol", nb2)
    assert np.array_equal(b2, bo) and np.array_equal(bc2, bc) and (nb2, nc3) == (nb, nc)
    # the reference's acceptance rules (color.cpp:14-47, block.cpp:15-89)
    (tmp_path / "x.col").write_text("2\n1 1\n2 3\n")
    with pytest.raises(I.IccgError, match="color_id out of range in color file: 3"):
        I.ReadColorFile_1Based(tmp_path / "x.col", 2)
    (tmp_path / "x.col").write_text("2\n1 1\n")
    with pytest.raises(I.IccgError, match="does not assign all nodes: assigned=1 expected=2"):
        I.ReadColorFile_1Based(tmp_path / "x.col", 2)
    (tmp_path / "x.col").write_text("2\n1 1\n2 2\n1 2\n")  # duplicates overwrite
    c, _ = I.ReadColorFile_1Based(tmp_path / "x.col", 2)
    assert list(c) == [1, 1]
    (tmp_path / "x.blk").write_text("0\n")
    with pytest.raises(I.IccgError, match="nb must be positive"):
        I.ReadBlockFile_1Based(tmp_path / "x.blk", 2)


def test_file_readers_match_reference(tmp_path):
    from oracle import ref as R
    if not R.available():
        pytest.skip("oracle/_ref not built")
    G = P.grid(9, 9)
    bo, nb, bc, nc = I.abmc_blocking(G, 5)
    I.write_block_file(tmp_path / "a.blk", bo, nb)
    I.write_block_color_file(tmp_path / "a.bcol", bc, nc)
    assert np.array_equal(R.read_blk(tmp_path / "a.blk", G.n)[0], bo)
    assert np.array_equal(R.read_bcol(tmp_path / "a.bcol", nb)[0], bc)
    color, ncol = I.greedy_color(G)
    I.write_color_file(tmp_path / "a.col", color, ncol)
    assert np.array_equal(R.read_col(tmp_path / "a.col", G.n)[0], color)


def _gen_grid_py(nx, ny, nz, lo, hi, shift, seed):
    """Restatement of the grid generator (face order x, y, z; x fastest)."""
    M = (1 << 64) - 1
    state = [seed & M]

    def u():
        state[0] = (state[0] + 0x9E3779B97F4A7C15) & M
        z = state[0]
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return ((z ^ (z >> 31)) >> 11) * 2.0 ** -53

    def draw(m):
        return [10.0 ** (lo + (hi - lo) * u()) if lo != hi else 10.0 ** lo for _ in range(m)]

    cx = draw((nx + 1) * ny * nz)
    cy = draw(nx * (ny + 1) * nz) if ny > 1 else None
    cz = draw(nx * ny * (nz + 1)) if nz > 1 else None
    rows = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                r = i + nx * (j + ny * k)
                d = cx[i + (nx + 1) * (j + ny * k)] + cx[i + 1 + (nx + 1) * (j + ny * k)]
                ent = []
                if cy is not None:
                    d = (d + cy[i + nx * (j + (ny + 1) * k)]) + cy[i + nx * (j + 1 + (ny + 1) * k)]
                if cz is not None:
                    d = (d + cz[i + nx * (j + ny * k)]) + cz[i + nx * (j + ny * (k + 1))]
                d = d + shift
                if cz is not None and k > 0:
                    ent.append((r - nx * ny, -cz[i + nx * (j + ny * k)]))
                if cy is not None and j > 0:
                    ent.append((r - nx, -cy[i + nx * (j + (ny + 1) * k)]))
                if i > 0:
                    ent.append((r - 1, -cx[i + (nx + 1) * (j + ny * k)]))
                ent.append((r, d))
                if i < nx - 1:
                    ent.append((r + 1, -cx[i + 1 + (nx + 1) * (j + ny * k)]))
                if cy is not None and j < ny - 1:
                    ent.append((r + nx, -cy[i + nx * (j + 1 + (ny + 1) * k)]))
                if cz is not None and k < nz - 1:
                    ent.append((r + nx * ny, -cz[i + nx * (j + ny * (k + 1))]))
                rows.append(ent)
    return rows


def test_generator_matches_restatement():
    for dims, lo, hi, sh in [((5, 4, 3), -2.0, 2.0, 1e-6), ((6, 5, 1), -1.0, 1.0, 0.0),
                             ((4, 4, 4), 0.0, 0.0, 0.0)]:
        G = I.gen_grid(*dims, coef_exp=(lo, hi), diag_shift=sh, seed=77)
        rows = _gen_grid_py(*dims, lo, hi, sh, 77)
        for i, ent in enumerate(rows):
            s, e = G.rowptr[i], G.rowptr[i + 1]
            assert list(G.colind[s:e]) == [c for c, _ in ent]
            assert list(G.val[s:e]) == [v for _, v in ent]
    L = I.gen_grid(100, 100)  # the 100x100 Laplacian: diag 4, off-diagonal -1
    assert L.n == 10000 and L.nnz == 49600
    assert set(np.unique(L.val)) == {-1.0, 4.0}
    G3 = I.gen_grid(12, 12, 12)
    assert G3.nnz == 7 * 12 ** 3 - 6 * 12 ** 2 and set(np.unique(G3.val)) == {-1.0, 6.0}


def test_normal_generator_box_muller():
    import math
    v = I.gen_normal(5, 123)
    M = (1 << 64) - 1
    st = 123
    out = []
    for _ in range(3):
        us = []
        for _ in range(2):
            st = (st + 0x9E3779B97F4A7C15) & M
            z = st
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
            us.append(((z ^ (z >> 31)) >> 11) * 2.0 ** -53)
        r = math.sqrt(-2.0 * math.log(1.0 - us[0]))
        t = 6.283185307179586 * us[1]
        out += [r * math.cos(t), r * math.sin(t)]
    assert list(v) == out[:5]
    big = I.gen_normal(200000, 1)
    assert abs(big.mean()) < 0.01 and abs(big.std() - 1) < 0.01


def test_validators_flag_bad_colorings():
    G = P.grid(8, 8)
    bo = np.zeros(G.n, np.int32)
    bo[G.n // 2:] = 1
    bc = np.array([0, 0], np.int32)  # adjacent blocks share a color
    tot, per = I.check_interblock(G, bo, bc, 2, 1)
    assert tot == 8 and per[0] == 8
    color = np.zeros(G.n, np.int32)
    th, hz, sb, cb = I.check_hazards(G, np.arange(G.n, dtype=np.int32), color)
    assert th == (G.nnz - G.n) // 2


def test_ptap_and_expand_parallel_and_ordered_paths_match_oracle():
    """PtAP and lower->full take a parallel scatter when every row is
    strictly ascending and the ordered one otherwise; both equal the
    oracle's restatement of color.cpp:66-141 / crs_io.cpp:163-227 bit for
    bit, including rows with duplicate columns (summed in input order) and
    unsorted rows."""
    rng = np.random.default_rng(5)
    A = P.grid(23, 17, 3, coef=(-2.0, 2.0))
    p = rng.permutation(A.n).astype(np.int32)
    Ao = P.orc(A)
    got = I.Permute_PtAP_LowerCRS_to_LowerCRS(A, p)
    ref = O.ptap_lower(Ao, p)
    assert np.array_equal(got.rowptr, ref.rowptr) and np.array_equal(got.colind, ref.colind)
    assert np.array_equal(got.val, ref.val)
    F = I.expand_lower_to_full(got)
    Fr = O.expand_lower_to_full(ref)
    assert np.array_equal(F.rowptr, Fr.rowptr) and np.array_equal(F.colind, Fr.colind)
    assert np.array_equal(F.val, Fr.val)
    # rows out of order and with repeated columns: the ordered path
    n = 40
    rows, cols, vals = [], [], []
    for i in range(n):
        for j in rng.choice(i + 1, size=min(i + 1, 4), replace=False)[::-1]:
            rows.append(i); cols.append(int(j)); vals.append(float(rng.normal()))
        rows.append(i); cols.append(i); vals.append(1.5)  # a repeated diagonal entry
    order = np.argsort(np.array(rows), kind="stable")
    rp = np.zeros(n + 1, np.int32)
    np.add.at(rp, np.array(rows) + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    ci = np.array(cols, np.int32)[order]
    v = np.array(vals)[order]
    U = I.CRS(n, rp, ci, v)
    Uo = O.Csr(n, rp, ci, v)
    p = rng.permutation(n).astype(np.int32)
    got = I.Permute_PtAP_LowerCRS_to_LowerCRS(U, p)
    ref = O.ptap_lower(Uo, p)
    assert np.array_equal(got.colind, ref.colind) and np.array_equal(got.val, ref.val)
    F = I.expand_lower_to_full(U)
    Fr = O.expand_lower_to_full(Uo)
    assert np.array_equal(F.colind, Fr.colind) and np.array_equal(F.val, Fr.val)


def test_c_abi_header_is_plain_c_and_links(tmp_path):
    """include/iccg_hip.h compiles as C99 and a C program that references
    the batch and request-loop entry points links against the library (no
    GPU call is made)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("no C compiler")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "use.c"
    src.write_text(
        '#include "iccg_hip.h"\n'
        "int main(void) {\n"
        "  iccg_server_config cfg = {ICCG_SERVE_BATCHED, ICCG_MIXED, 8, 2, 64,\n"
        "                            ICCG_SOLVE_TRUE_RESID, 1LL << 30};\n"
        "  void *f[] = {(void *)iccg_batch_create, (void *)iccg_batch_run,\n"
        "               (void *)iccg_server_create, (void *)iccg_server_submit,\n"
        "               (void *)iccg_cg_solve};\n"
        "  return (cfg.max_slots == 8 && f[0] && f[4]) ? 0 : 1;\n"
        "}\n")
    exe = tmp_path / "use"
    r = subprocess.run([cc, "-std=c99", "-Wall", "-Werror", "-I", os.path.join(root, "include"),
                        str(src), "-o", str(exe), "-L", os.path.join(root, "iccg_amd", "lib"),
                        "-liccg_hip", "-Wl,-rpath," + os.path.join(root, "iccg_amd", "lib")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_cpp_mirror_headers_compile_with_multi_rhs(tmp_path):
    """include/iccg/*.hpp compile as C++17 with conjugate_gradient and
    conjugate_gradient_multi instantiated for IC0 and linked against the
    library (no GPU call is made: the program only takes their addresses)."""
    import shutil
    import subprocess
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no C++ compiler")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "use.cpp"
    src.write_text(
        '#include "iccg/cg.hpp"\n'
        "int main() {\n"
        "  auto f = &conjugate_gradient<IC0>;\n"
        "  auto g = &conjugate_gradient_multi<IC0>;\n"
        "  return (f && g) ? 0 : 1;\n"
        "}\n")
    exe = tmp_path / "use"
    r = subprocess.run([cxx, "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(root, "include"),
                        str(src), "-o", str(exe), "-L", os.path.join(root, "iccg_amd", "lib"),
                        "-liccg_hip", "-Wl,-rpath," + os.path.join(root, "iccg_amd", "lib")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
