"""Low modes and deflation (dslash.h:43-112, dslash.cpp:144-300, the deflated
restart cycles of invert_with_cg, dslash.h:496-500).

The low modes are exact: the lowest eigenvectors of the oracle's
multiply_hermop_sym2 on the odd sites of a 4x4x2x4 Ls6 lattice, from numpy's
dense Hermitian eigensolver, compressed per 2^4 block by a QR factorisation
(n_basis = n_vec, so the compression is exact up to complex64 rounding).
Pins: deflate equals the three steps of the reference restated in float64 on
the same complex64 data; its deflated part equals sum_i v_i <v_i, in> /
lambda_i of the exact eigenvectors; the deflated FP64 solve reaches the
oracle's solution in fewer iterations, and so do the mixed FP32 and FP16-storage
solvers started from the deflated guess.
"""

import numpy as np
import pytest

import qlattice_amd as q
from oracle import oracle as O
from tests.helpers import block_order, rand_field, rand_gauge, rel

pytestmark = pytest.mark.gpu

SITE = (4, 4, 2, 4)
LS = 6
BS = (2, 2, 2, 2)
NVEC = 24
MASS = 0.02


@pytest.fixture(scope="module")
def system():
    print("\n[gpu] deflation: dense eigensystem of hermop 4x4x2x4 Ls6", flush=True)
    U = rand_gauge(SITE, 0.3, 81)
    orc = O.OracleInverter(SITE, U, mass=MASS, ls=LS, m5=1.8, mobius_scale=2.0)
    nh = int(np.prod(SITE)) // 2
    dim = nh * LS * 12
    A = np.empty((dim, dim), np.complex128)
    e = np.zeros(dim, np.complex128)
    for k in range(dim):
        e[k] = 1.0
        A[:, k] = orc.apply_half("hermop_sym2", e.reshape(nh, LS, 12), 1).reshape(-1)
        e[k] = 0.0
    A = 0.5 * (A + A.conj().T)
    w, V = np.linalg.eigh(A)
    lam, vec = w[:NVEC], V[:, :NVEC]  # (dim, NVEC)
    # compress per block: eigenvector rows of block b in the layout order
    rows = block_order(SITE, BS)
    nblock, bve = rows.shape
    idx = (rows[:, :, None, None] * LS * 12 + np.arange(LS)[None, None, :, None] * 12
           + np.arange(12)[None, None, None, :]).reshape(nblock, -1)
    bases = np.empty((nblock, NVEC, idx.shape[1]), np.complex64)
    coefs = np.empty((nblock, NVEC, NVEC), np.complex64)
    for b in range(nblock):
        Qb, Rb = np.linalg.qr(vec[idx[b]])  # (bsize, NVEC), (NVEC, NVEC)
        bases[b] = Qb.T
        coefs[b] = Rb.T  # coefs[b, i, j] = R[j, i]
    lm = q.LowModes(BS, lam, bases, coefs)
    return U, orc, lm, lam, vec, idx


def make_inv(U):
    geo = q.Geometry(SITE)
    gf = q.GaugeField(geo)
    gf.data = U
    fa = q.FermionAction(mass=MASS, ls=LS, m5=1.8, mobius_scale=2.0)
    return geo, q.InverterDomainWall(gf=gf, fa=fa)


def restated_deflate(x, lm, idx):
    """dslash.cpp:144-249 in float64 on the complex64 data (input rounded to
    complex64; the produce step adds to the blocked input)."""
    v = x.reshape(-1).astype(np.complex64).astype(np.complex128)
    B = lm.bases.astype(np.complex128)
    C = lm.coefs.astype(np.complex128)
    vb = v[idx]  # (nblock, bsize)
    c = np.einsum("bjk,bk->bj", B.conj(), vb)
    p = np.einsum("bij,bj->i", C.conj(), c) / lm.eigen_values
    c2 = np.einsum("bij,i->bj", C, p)
    out = v.copy()
    out[idx] = vb + np.einsum("bjk,bj->bk", B, c2)
    return out.reshape(x.shape)


def test_deflate_matches_reference_steps_and_projector(system):
    U, orc, lm, lam, vec, idx = system
    geo, inv = make_inv(U)
    nh = geo.total_volume() // 2
    x = rand_field((nh, LS, 12), 82)
    f = q.FermionField5d(geo, LS, 1)
    f.data = x
    zero = q.deflate(None, f, inv).data
    assert not zero.any()  # no low modes attached: out = 0 (dslash.cpp:147-152)
    inv.set_low_modes(lm)
    got = q.deflate(None, f, inv).data
    assert rel(got, restated_deflate(x, lm, idx)) < 2e-6
    x32 = x.astype(np.complex64).astype(np.complex128)
    proj = (vec @ ((vec.conj().T @ x.reshape(-1)) / lam)).reshape(x.shape)
    assert rel(got - x32, proj) < 1e-5
    inv.set_low_modes(None)
    assert not q.deflate(None, f, inv).data.any()


def test_deflated_fp64_solve(system):
    U, orc, lm, lam, vec, idx = system
    geo, inv = make_inv(U)
    src = q.FermionField4d(geo)
    src.data = rand_field((geo.total_volume(), 12), 83)
    inv.set_stop_rsd(1e-10)
    inv.set_max_num_iter(5000)
    plain = inv * src
    it_plain = inv.stats()["total_iter"]
    inv.set_low_modes(lm)
    defl = inv * src
    it_defl = inv.stats()["total_iter"]
    orc.set_params(1e-12, 5000, 300)
    ref, _ = orc.invert_4d(src.data)
    print(f"[gpu] deflated solve: {it_defl} iterations against {it_plain}; lambda {lam[0]:.3e}..{lam[-1]:.3e}",
          flush=True)
    assert rel(defl.data, ref) < 1e-8 and rel(plain.data, ref) < 1e-8
    assert it_defl < it_plain


@pytest.mark.parametrize("solver", ["mixed16", "mixed", "mixed-fused"])
def test_deflated_mixed_solve(system, solver, monkeypatch):
    """The mixed solvers start from x = deflate(b) with the FP64 residual
    b - A x, then iterate as without low modes ("mixed-fused": the FP32
    solver on the fused kernels, the path taken for Ls outside [6, 16])."""
    U, orc, lm, lam, vec, idx = system
    if solver == "mixed-fused":
        monkeypatch.setenv("QLAT_DWF_THREE_PHASE", "0")
    geo, inv = make_inv(U)
    inv.set_solver("mixed" if solver == "mixed-fused" else solver)
    src = q.FermionField4d(geo)
    src.data = rand_field((geo.total_volume(), 12), 84)
    inv.set_stop_rsd(1e-9)
    inv.set_max_num_iter(5000)
    plain = inv * src
    st_plain = inv.stats()
    inv.set_low_modes(lm)
    defl = inv * src
    st = inv.stats()
    orc.set_params(1e-12, 5000, 300)
    ref, _ = orc.invert_4d(src.data)
    print(f"[gpu] deflated {solver} solve: {st['total_iter']} iterations against {st_plain['total_iter']}",
          flush=True)
    assert st["true_residual"] <= 1e-9 and st_plain["true_residual"] <= 1e-9
    assert rel(defl.data, ref) < 1e-6 and rel(plain.data, ref) < 1e-6
    assert st["total_iter"] < st_plain["total_iter"]


def test_deflate_with_low_modes_from_files(system, tmp_path):
    """load_low_modes (dslash.cpp:9-23) of the same system written over two
    nodes: bit-identical deflation from FP32 files, and within the 16-bit
    code's resolution when half the basis and the coefficients are 16-bit."""
    U, orc, lm, lam, vec, idx = system
    geo, inv = make_inv(U)
    nh = geo.total_volume() // 2
    f = q.FermionField5d(geo, LS, 1)
    f.data = rand_field((nh, LS, 12), 85)
    inv.set_low_modes(lm)
    ref = q.deflate(None, f, inv).data
    q.save_compressed_eigen_vectors(str(tmp_path / "f32"), lm, SITE, LS, (4, 4, 2, 2))
    inv.set_low_modes(q.load_low_modes(str(tmp_path / "f32")))
    assert np.array_equal(q.deflate(None, f, inv).data, ref)
    q.save_compressed_eigen_vectors(str(tmp_path / "f16"), lm, SITE, LS, (4, 4, 2, 2), nkeep_single=NVEC // 2,
                                    fp16_coef_exp_share_floats=8)
    inv.set_low_modes(q.load_low_modes(str(tmp_path / "f16")))
    got = q.deflate(None, f, inv).data
    x32 = f.data.astype(np.complex64).astype(np.complex128)
    proj = (vec @ ((vec.conj().T @ f.data.reshape(-1)) / lam)).reshape(f.data.shape)
    print(f"[gpu] deflation from 16-bit files: rel {rel(got - x32, proj):.2e}", flush=True)
    assert rel(got - x32, proj) < 1e-3


def test_deflated_fused_propagator(system, monkeypatch):
    """The multi-RHS FP32 solver on the fused kernels (cg_mixed_multi) starts
    every column from x = deflate(b) too: a 12-column propagator reaches the
    oracle's solution in fewer iterations with the low modes attached."""
    U, orc, lm, lam, vec, idx = system
    monkeypatch.setenv("QLAT_DWF_THREE_PHASE", "0")
    geo, inv = make_inv(U)
    inv.set_solver("mixed")
    inv.set_stop_rsd(1e-9)
    inv.set_max_num_iter(5000)
    src = q.Prop(geo).set_rand(86)
    plain = inv * src
    it_plain = inv.stats()["total_iter"]
    inv.set_low_modes(lm)
    defl = inv * src
    st = inv.stats()
    orc.set_params(1e-12, 5000, 300)
    ref, _ = orc.invert_prop(src.data)
    print(f"[gpu] deflated fused propagator: {st['total_iter']} iterations against {it_plain}", flush=True)
    assert st["true_residual"] <= 1e-9
    assert rel(defl.data, ref) < 1e-6 and rel(plain.data, ref) < 1e-6
    assert st["total_iter"] < it_plain
