"""CPU tests pinning the oracle (oracle/hcore_oracle.c) -- no GPU needed.

Every known-answer vector the reference's own tests hold for this path (see
tests/golden/make_reference_vectors.py for the file:line of each), the
AdvancedGemmTest norm-wise bound, and mathematical checks chosen so that a
plausible slip (alpha/beta on the wrong block, a transposed factor, a wrong
rank rule) fails one of them: the Eckart-Young truncation bound against numpy's
SVD of the dense update, alpha=0 / beta=0 special cases, closed-form rank rules.
"""
import json
import os

import numpy as np
import pytest

from helpers import EPS, rel_fro
from oracle import oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))
K = GOLD["kernels"]


def test_lapack_resolved():
    assert O.lapack_path() != ""


# ---------------------------------------------------------------- kernel known answers
def test_process_vpointer():
    v = K["process_vpointer"]
    out = np.zeros(10)
    cv = np.array(v["cv"], float)
    b = np.array(v["b"], float)
    O.lib().orc_process_vpointer(v["n"], v["crank"], v["n"], v["beta"], O._ptr(cv), v["n"], O._ptr(out), v["arank"],
                                 O._ptr(b))
    assert out.tolist() == v["expected"]


@pytest.mark.parametrize("key", ["new_rank_abs", "new_rank_rel"])
def test_calculate_new_rank(key):
    v = K[key]
    assert O.calculate_new_rank(v["sigma"], v["acc"], v["truncated"]) == v["expected"]


def test_rank_rule_closed_forms():
    # all sigma >= acc -> size_s; sigma_1 < acc -> 1 (loop starts at i = 1); relative rule
    assert O.calculate_new_rank([5, 4, 3], 1.0) == 3
    assert O.calculate_new_rank([5, 0.1, 3], 1.0) == 1
    assert O.calculate_new_rank([0.5, 0.1], 1.0) == 1      # sigma_0 < acc still gives rank >= 1
    assert O.calculate_new_rank([10, 5, 0.9, 0.2], 0.1, truncated_svd=True) == 2  # 0.9 < 0.1 * 10


def test_calculate_uvptr():
    v = K["uvptr"]
    out = np.zeros(v["vm"] * v["rank"])
    vn = np.array(v["vnew"], float)
    O.lib().orc_calculate_uvptr(v["rank"], v["vm"], O._ptr(out), O._ptr(vn))
    assert out.tolist() == v["expected"]


def test_calculate_vtnew():
    v = K["vtnew_ungqr"]
    vt = np.array(v["vt"], float)
    sg = np.array(v["sigma"], float)
    O.lib().orc_calculate_vtnew(v["rank"], v["min_vm_vn"], O._ptr(sg), O._ptr(vt), v["size_s"])
    assert vt.tolist() == v["expected"]


def test_lacpy_upper_and_laset_lower():
    v = K["lacpy_upper"]
    a = np.array(v["a"], float)
    b = np.zeros(16)
    O.lib().orc_lacpy(b"U", 4, 4, O._ptr(a), 4, O._ptr(b), 4)
    assert b.tolist() == v["expected"]
    v = K["laset_lower"]
    a = np.zeros(16)
    O.lib().orc_laset(b"L", 4, 4, v["off"], v["diag"], O._ptr(a), 4)
    assert a.tolist() == v["expected"]


def test_geqrf_known_answer():
    v = K["geqrf"]
    # the reference stores the 2x6 literal row-major and passes it as a 6x2 column-major matrix
    a = np.array(v["a_rowmajor_T"], float).reshape(-1)
    tau = np.zeros(2)
    assert O.lib().orc_geqrf(6, 2, O._ptr(a), 6, O._ptr(tau)) == 0
    np.testing.assert_allclose(a.reshape(2, 6), np.array(v["expected_rowmajor_T"]), rtol=1e-2, atol=1e-5)
    np.testing.assert_allclose(tau, v["tau"], rtol=1e-2)


@pytest.mark.parametrize("svd_type", [O.GESVD, O.GESDD])
def test_svd_reconstruction(svd_type):
    a = np.asfortranarray(np.array(K["svd_3x3"]["a"], float))
    a0 = a.copy()
    s = np.zeros(3)
    u = np.zeros((3, 3), order="F")
    vt = np.zeros((3, 3), order="F")
    O.lib().orc_svd(svd_type, 3, 3, O._ptr(a), 3, O._ptr(s), O._ptr(u), 3, O._ptr(vt), 3)
    np.testing.assert_allclose(u @ np.diag(s) @ vt, a0, atol=1e-12)
    assert s[2] < 1e-12 * s[0]  # rank 2


# ---------------------------------------------------------------- HCore::Gemm known answers
def _tiles(case):
    arr = {k: np.array(case[k], float) for k in ("A", "B", "Au", "Av", "Bu", "Bv") if k in case}
    k = case["kind"]
    A = O.Tile.dense(arr["A"]) if k[0] == "D" else O.Tile.compressed(arr["Au"], arr["Av"])
    B = O.Tile.dense(arr["B"]) if k[1] == "D" else O.Tile.compressed(arr["Bu"], arr["Bv"])
    out = np.array(case["out"], float)
    m, n = out.shape
    if k[2] == "D":
        C = O.Tile.dense(np.zeros((m, n)))
    else:
        cr = case["c_rank"]
        C = O.Tile.compressed(np.zeros((m, cr)), np.zeros((cr, n)))
    return A, B, C, out


@pytest.mark.parametrize("case", GOLD["gemm"], ids=[c["name"] for c in GOLD["gemm"]])
def test_hcore_gemm_known_answers(case):
    A, B, C, out = _tiles(case)
    acc = EPS if case.get("acc") == "eps" else 1e-9
    assert O.hcore_gemm(1.0, A, B, 1.0, C, O.Params(acc)) == 0
    np.testing.assert_allclose(C.to_dense(), out, rtol=1e-2)


def test_compressed_tile_direct_gemm():
    g = GOLD["compressed_direct"]
    A = O.Tile.compressed(np.array(g["A"], float), np.eye(3))
    B = O.Tile.dense(np.array(g["B"], float))
    C = O.Tile.compressed(np.zeros((3, 2)), np.zeros((2, 2)))
    assert O.hcore_gemm(1.0, A, B, 1.0, C, O.Params(EPS)) == 0
    np.testing.assert_allclose(C.to_dense(), np.array(g["out"], float), rtol=1e-2)
    assert C.rank == 2


# ---------------------------------------------------------------- AdvancedGemmTest bound
COMBOS = ["DDD", "DDC", "DCD", "DCC", "CDD", "CDC", "CCD", "CCC"]


@pytest.mark.skipif(not O.lib().orc_has_latms(), reason="dlatms not exported by the LAPACK runtime")
@pytest.mark.parametrize("n", [100, 200, 300])
@pytest.mark.parametrize("combo", COMBOS)
def test_advanced_gemm_bound(combo, n):
    """tests/api/AdvancedGemmTest.cpp:23-280: latms inputs (sigma_i = 10^-i, iseed {0,0,0,1}),
    alpha=3.5, beta=2.5, acc=1e-4, error = ||Cref - C||_inf / (sqrt(k+2)|alpha| |A||B| + 2|beta||C|) < 3 acc."""
    alpha, beta, acc = 3.5, 2.5, 1e-4
    seed = [0, 0, 0, 1]
    Ad = O.latms_dense(n, n, seed)
    Bd = O.latms_dense(n, n, seed)
    Cd = O.latms_dense(n, n, seed)
    inf = lambda x: np.max(np.sum(np.abs(x), axis=1))
    Cref = beta * Cd + alpha * Ad @ Bd
    A = O.compress_dense_matrix(Ad, acc) if combo[0] == "C" else O.Tile.dense(Ad)
    B = O.compress_dense_matrix(Bd, acc) if combo[1] == "C" else O.Tile.dense(Bd)
    C = O.compress_dense_matrix(Cd, acc) if combo[2] == "C" else O.Tile.dense(Cd)
    if combo == "DDC":
        # DDC re-sizes C to rank min(m,n) (ReadjustTile re-allocates); give the buffer that capacity
        C = O.Tile.compressed(C.U.copy(), C.V.copy(), max_rank=n)
    info = O.hcore_gemm(alpha, A, B, beta, C, O.Params(acc))
    assert info == 0
    err = inf(Cref - C.to_dense()) / (np.sqrt(n + 2) * abs(alpha) * inf(Ad) * inf(Bd) + 2 * abs(beta) * inf(Cd))
    assert err < 3 * acc, err


# ---------------------------------------------------------------- mathematical pins
def _lowrank(rng, m, n, r, decay=0.7, scale=1.0):
    from hcorepp_amd.workloads import lowrank_factors
    return lowrank_factors(rng, m, n, r, decay, scale)


@pytest.mark.parametrize("seed", range(4))
def test_eckart_young_against_numpy(seed):
    """The recompressed C equals numpy's best rank-r approximation of the dense update,
    with r given by the absolute threshold rule."""
    rng = np.random.default_rng(seed)
    m, ra, rb, rc, acc = 90, 8 + seed, 7, 10, 1e-6
    au, av = _lowrank(rng, m, m, ra)
    bu, bv = _lowrank(rng, m, m, rb)
    cu, cv = _lowrank(rng, m, m, rc)
    alpha, beta = 1.5, -0.5
    dense = beta * cu @ cv + alpha * au @ av @ bu @ bv
    C = O.Tile.compressed(cu, cv, max_rank=60)
    info, sig = O.hcore_gemm(alpha, O.Tile.compressed(au, av), O.Tile.compressed(bu, bv), beta, C,
                             O.Params(acc), want_sigma=True)
    assert info == 0
    U, S, Vt = np.linalg.svd(dense)
    r = next((i for i in range(1, len(S)) if S[i] < acc), len(S))
    assert C.rank == r
    best = (U[:, :r] * S[:r]) @ Vt[:r]
    assert rel_fro(C.to_dense(), best) < 1e-11
    np.testing.assert_allclose(sig[:r], S[:r], rtol=1e-10, atol=1e-13 * S[0])
    # U orthonormal, V carries Sigma
    assert np.linalg.norm(C.U.T @ C.U - np.eye(r)) < 1e-12
    np.testing.assert_allclose(np.linalg.norm(C.V, axis=1), S[:r], rtol=1e-10)


def test_alpha_zero_and_beta_zero():
    rng = np.random.default_rng(9)
    au, av = _lowrank(rng, 60, 60, 6)
    bu, bv = _lowrank(rng, 60, 60, 5)
    cu, cv = _lowrank(rng, 60, 60, 7)
    # alpha = 0: pure recompression of C
    C = O.Tile.compressed(cu, cv, max_rank=30)
    O.hcore_gemm(0.0, O.Tile.compressed(au, av), O.Tile.compressed(bu, bv), 1.0, C, O.Params(1e-12))
    assert rel_fro(C.to_dense(), cu @ cv) < 1e-13 and C.rank == 7
    # beta = 0: compression of A*B
    C = O.Tile.compressed(cu, cv, max_rank=30)
    O.hcore_gemm(1.0, O.Tile.compressed(au, av), O.Tile.compressed(bu, bv), 0.0, C, O.Params(1e-12))
    assert rel_fro(C.to_dense(), au @ av @ bu @ bv) < 1e-12 and C.rank == 5


def test_alpha_beta_on_the_right_blocks():
    """A restatement that scaled C by alpha (or A*B by beta) would fail: checked with asymmetric values."""
    rng = np.random.default_rng(10)
    au, av = _lowrank(rng, 40, 40, 4)
    bu, bv = _lowrank(rng, 40, 40, 4)
    cu, cv = _lowrank(rng, 40, 40, 4)
    C = O.Tile.compressed(cu, cv, max_rank=20)
    O.hcore_gemm(2.0, O.Tile.compressed(au, av), O.Tile.compressed(bu, bv), 7.0, C, O.Params(1e-13))
    good = 7.0 * cu @ cv + 2.0 * au @ av @ bu @ bv
    swapped = 2.0 * cu @ cv + 7.0 * au @ av @ bu @ bv
    assert rel_fro(C.to_dense(), good) < 1e-12
    assert rel_fro(C.to_dense(), swapped) > 1e-2


def test_max_rank_clamp_and_fixed_rank():
    rng = np.random.default_rng(12)
    au, av = _lowrank(rng, 50, 50, 10, decay=0.9)
    bu, bv = _lowrank(rng, 50, 50, 10, decay=0.9)
    cu, cv = _lowrank(rng, 50, 50, 6, decay=0.9)
    C = O.Tile.compressed(cu, cv, max_rank=8)
    O.hcore_gemm(1.0, O.Tile.compressed(au, av), O.Tile.compressed(bu, bv), 1.0, C, O.Params(1e-14))
    assert C.rank == 8
    C = O.Tile.compressed(cu, cv, max_rank=30)
    O.hcore_gemm(1.0, O.Tile.compressed(au, av), O.Tile.compressed(bu, bv), 1.0, C, O.Params(1e-14, fixed_rank=5))
    assert C.rank == 5


def test_compression_constructor():
    """Compressed.cpp:75-146: max_rank = min(m,n)/3, VT eats Sigma, rank by the absolute rule."""
    rng = np.random.default_rng(13)
    u, v = _lowrank(rng, 90, 90, 20, decay=0.5)
    a = u @ v
    t, s = O.compress(a, O.Params(1e-6))
    assert t.max_rank == 30
    r = next(i for i in range(1, len(s)) if s[i] < 1e-6)
    assert t.rank == r
    assert rel_fro(t.to_dense(), a) < 1e-5


# ---------------------------------------------------------------------------
# Rank-predictor network oracle (oracle/mlp.py), pinned by finite differences and closed forms.
def _mlp_case(seed=3, n=40):
    from oracle import mlp as M
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(M.F * M.H + M.H + M.H * M.H + M.H + M.H + 1) * 0.3
    X = rng.standard_normal((n, M.F))
    y = rng.standard_normal(n)
    return M, p, X, y


def test_mlp_oracle_gradient_matches_finite_differences():
    M, p, X, y = _mlp_case()
    loss, g = M.loss_grad(p, X, y)
    rng = np.random.default_rng(9)
    idx = rng.choice(p.size, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        num = (M.loss_grad(p + e, X, y)[0] - M.loss_grad(p - e, X, y)[0]) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 * max(1.0, abs(num)), (i, num, g[i])


def test_mlp_oracle_closed_forms():
    """All-zero hidden weights: output = b3 for every sample, so the loss is mean((b3 - y)^2) and
    only b3 has a gradient, 2 mean(b3 - y)."""
    M, p, X, y = _mlp_case()
    q = np.zeros_like(p)
    q[-1] = 0.7
    loss, g = M.loss_grad(q, X, y)
    assert np.isclose(loss, np.mean((0.7 - y) ** 2))
    assert np.allclose(g[:-1], 0.0) and np.isclose(g[-1], 2 * np.mean(0.7 - y))
    # one Adam step from zero moments moves every parameter with a nonzero gradient by ~lr
    p1, m, v = M.adam(q, g, np.zeros_like(q), np.zeros_like(q), 1, 1e-3)
    assert np.isclose(abs(p1[-1] - q[-1]), 1e-3, rtol=1e-4)


def test_bf16_rounding_oracle():
    from oracle import mlp as M
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38], dtype=np.float32)
    r = M.bf16(x)
    # 1 + 2^-8 is a tie: rounds to even (1.0); 1 + 3*2^-8 rounds up to 1 + 2^-6
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0 + 2 ** -6) and r[3] == -2.5


# ---------------------------------------------------------------- Cholesky tile operations
def _chol_oracle():
    from oracle import cholesky as OC
    return OC


def test_oracle_potrf_known_answer():
    v = K["potrf_lower"]
    L = _chol_oracle().potrf(np.array(v["a"], dtype=float))
    assert np.array_equal(L, np.array(v["expected"], dtype=float))


def test_oracle_trsm_known_answer():
    """The reference's Trsm vector solves A x = b with A upper (Side Left); the tile operation
    solves X L^T = V with L lower.  With J the index reversal, L = J A J is lower and
    x^T = (J b)^T L^{-T} reversed."""
    v = K["trsm_left_upper"]
    A = np.array(v["a"], dtype=float)
    b = np.array(v["b"], dtype=float)[:, 0]
    L = A[::-1, ::-1].copy()
    x = _chol_oracle().trsm_v(L, b[::-1][None, :])[0][::-1]
    assert np.allclose(x, np.array(v["expected"], dtype=float)[:, 0], rtol=0, atol=1e-14)


def test_oracle_syrk_known_answer():
    """CrossProduct vector: lower(A A^T).  The compressed syrk with U = A, V = I, C = 0 gives
    lower(-A A^T)."""
    v = K["syrk_lower"]
    A = np.array(v["a"], dtype=float)
    out = _chol_oracle().syrk_compressed(np.zeros((4, 4)), A, np.eye(4))
    exp = np.tril(np.array(v["expected"], dtype=float))
    assert np.allclose(-out, exp, rtol=1e-11, atol=0)
    assert np.array_equal(np.triu(out, 1), np.zeros((4, 4)))


def _small_tlr(acc=1e-5, n=512, nb=128):
    from hcorepp_amd.workloads import TLRMatrix
    M = TLRMatrix(n, nb, 2, 0.1, acc, 20261205, nugget=1e-2)
    diag = [M.tile(k, k)[1] for k in range(M.T)]
    off = {(i, k): M.tile(i, k)[1:] for i in range(M.T) for k in range(i)}
    return M, diag, off


def test_oracle_tlr_cholesky_backward_error():
    """A TLR Cholesky at accuracy 1e-5 of an SPD covariance reproduces A to O(acc): the
    relative backward error is below 0.2 acc and L is close to numpy's dense Cholesky."""
    OC = _chol_oracle()
    M, diag, off = _small_tlr()
    F = OC.TLRCholesky(diag, off, M.acc).factor()
    Ld = F.dense_factor()
    A = M.dense_block_range(0, M.n)
    be = np.linalg.norm(A - Ld @ Ld.T) / np.linalg.norm(A)
    assert be < 2e-6, be
    Lx = np.linalg.cholesky(A)
    assert rel_fro(Ld, Lx) < 1e-4
    assert all(0 < r <= 42 for r in F.ranks().values())


def test_oracle_tlr_cholesky_planted_sign_slip_fails():
    """The backward-error check catches a trailing update applied with the wrong sign."""
    OC = _chol_oracle()
    M, diag, off = _small_tlr()
    F = OC.TLRCholesky(diag, off, M.acc)
    orig = OC.syrk_compressed
    try:
        OC.syrk_compressed = lambda c, u, v: np.tril(c + (u @ (v @ v.T)) @ u.T)
        with pytest.raises(np.linalg.LinAlgError):
            F.factor()
            Ld = F.dense_factor()
            A = M.dense_block_range(0, M.n)
            if np.linalg.norm(A - Ld @ Ld.T) / np.linalg.norm(A) > 1e-3:
                raise np.linalg.LinAlgError("backward error")
    finally:
        OC.syrk_compressed = orig


def test_oracle_under_address_and_undefined_sanitizers(tmp_path):
    """The oracle built with -fsanitize=address,undefined runs every HCore::Gemm combination, the
    compression constructor and the tile-GEMM driver without a sanitizer report (host code only;
    LAPACK itself is not instrumented and runs sequentially: MKL's own threads under the address
    sanitizer returned LAPACK failures that its sequential layer does not)."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path / "oracle_sanitize")
    subprocess.run([gcc, "-O1", "-g", "-fopenmp", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
                    "-fno-sanitize-recover=undefined", os.path.join(root, "oracle", "hcore_oracle.c"),
                    os.path.join(root, "tests", "cpp", "oracle_sanitize.c"), "-ldl", "-lm", "-o", exe], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="halt_on_error=1",
               OMP_NUM_THREADS="2", MKL_THREADING_LAYER="SEQUENTIAL")
    res = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert res.returncode == 0, res.stdout + res.stderr[-3000:]
    assert "oracle_sanitize: ok" in res.stdout
