"""Fermion force of S = phi^+ (Mpc^+ Mpc)^-1 phi: GPU adjoint kernel vs
(1) central finite differences of the GPU FP64 action along random su(3)
directions, (2) plain FP64 PyTorch autograd through an einsum/roll Mpc,
(3) the oracle's action value.  The reference has no fermion force, so these
checks (mathematics plus an independent implementation) are its pins.
"""

import numpy as np
import pytest
import torch

import qlattice_amd as q
from oracle import oracle as O
from qlattice_amd.gauge import gell_mann_generators
from tests.helpers import rand_field, rand_gauge, rel
from tests.torch_dwf import TorchMpc, gauge_to_lattice, odd_half_to_lattice

pytestmark = pytest.mark.gpu


def su3_direction(nlinks, seed):
    """Random T = sum_a t_a (i T_a) per link (anti-hermitian, traceless)."""
    rng = np.random.default_rng(seed)
    t = rng.standard_normal((nlinks, 8))
    return 1j * np.einsum("na,aij->nij", t, gell_mann_generators())


def expm_antiherm(T):
    """exp(T) for anti-hermitian T via the eigen-decomposition of -iT."""
    w, v = np.linalg.eigh(-1j * T)
    return np.einsum("nij,nj,nkj->nik", v, np.exp(1j * w), v.conj())


def make_case(site, ls, mass, sigma, seed, mobius_scale=2.0):
    U = rand_gauge(site, sigma, seed)
    geo = q.Geometry(site)
    gf = q.GaugeField(geo)
    gf.data = U
    fa = q.FermionAction(mass=mass, ls=ls, m5=1.8, mobius_scale=mobius_scale)
    inv = q.InverterDomainWall(gf=gf, fa=fa)
    inv.set_max_num_iter(5000)
    phi = rand_field((geo.total_volume() // 2, ls, 12), seed + 1)
    return geo, fa, inv, U, phi


def action(inv, U, phi, stop):
    """S(U) as an (hi, lo) pair of the compensated device sum."""
    inv.set_gauge(U)
    S, _, _, _ = q.fermion_force(inv, phi, stop, action_pair=True)
    return S


def central_difference(inv, Uf, T, phi, eps, stop=1e-14):
    """(S(exp(eps T) U) - S(exp(-eps T) U)) / (2 eps), with the two actions
    subtracted as (hi, lo) pairs: hi - hi is exact for nearby values."""
    shape = (Uf.shape[0] // 4, 4, 3, 3)
    sp = action(inv, (expm_antiherm(eps * T) @ Uf).reshape(shape), phi, stop)
    sm = action(inv, (expm_antiherm(-eps * T) @ Uf).reshape(shape), phi, stop)
    return ((sp[0] - sm[0]) + (sp[1] - sm[1])) / (2 * eps)


@pytest.fixture(scope="module")
def small():
    print("\n[gpu] force 4^3x8 Ls6", flush=True)
    return make_case((4, 4, 4, 8), 6, 0.1, 0.4, 61)


def test_action_matches_oracle(small):
    geo, fa, inv, U, phi = small
    S, F, G, it = q.fermion_force(inv, phi, 1e-13, with_grad=True)
    orc = O.OracleInverter(geo.total_site(), U, mass=fa.mass(), ls=fa.ls(), m5=fa.m5(),
                           mobius_scale=fa.mobius_scale())
    X, _ = orc.cg_with_herm_sym_2(phi, stop_rsd=1e-13, max_num_iter=5000)
    S_ref = np.vdot(phi, X).real
    assert abs(S - S_ref) < 1e-10 * abs(S_ref)
    assert inv.stats()["true_residual"] < 1e-12


def test_force_is_projected_u_times_g(small):
    """force = T_A(U G) with G = grad^+, and force is in su(3)."""
    geo, fa, inv, U, phi = small
    inv.set_gauge(U)
    S, F, grad, _ = q.fermion_force(inv, phi, 1e-13, with_grad=True)
    G = np.conj(np.swapaxes(grad, -1, -2))
    M = U @ G
    A = 0.5 * (M - np.conj(np.swapaxes(M, -1, -2)))
    tr = np.trace(A, axis1=-2, axis2=-1) / 3
    TA = A - tr[..., None, None] * np.eye(3)
    assert rel(F, TA) < 1e-13
    assert np.allclose(F, -np.conj(np.swapaxes(F, -1, -2)), atol=1e-12)
    assert np.allclose(np.trace(F, axis1=-2, axis2=-1), 0, atol=1e-12)


def test_force_matches_central_finite_differences(small):
    """dS/deps along U -> exp(eps T) U equals sum Tr[force T] (central FD,
    eps = 1e-5, FP64 action at stop 1e-14), for 8 seeded su(3) directions."""
    geo, fa, inv, U, phi = small
    inv.set_gauge(U)
    _, F, _, _ = q.fermion_force(inv, phi, 1e-14)
    Uf = U.reshape(-1, 3, 3)
    for k in range(8):
        T = su3_direction(Uf.shape[0], 100 + k)
        fd = central_difference(inv, Uf, T, phi, 1e-5)
        an = np.einsum("nij,nji->", F.reshape(-1, 3, 3), T).real
        assert abs(fd - an) < 1e-6 * abs(an), (k, fd, an)
    inv.set_gauge(U)


def test_grad_matches_torch_autograd(small):
    """grad (dS/dU, conjugate Wirtinger) equals PyTorch autograd of
    -Re <X, Mpc^+ Mpc(U) X> with X = (Mpc^+ Mpc)^-1 phi held fixed."""
    geo, fa, inv, U, phi = small
    site, ls = geo.total_site(), fa.ls()
    inv.set_gauge(U)
    S, F, grad, _ = q.fermion_force(inv, phi, 1e-13, with_grad=True)
    orc = O.OracleInverter(site, U, mass=fa.mass(), ls=ls, m5=fa.m5(), mobius_scale=fa.mobius_scale())
    X, _ = orc.cg_with_herm_sym_2(phi, stop_rsd=1e-13, max_num_iter=5000)
    tm = TorchMpc(site, ls, fa.mass(), fa.m5(), fa.bs, fa.cs)
    Ut = gauge_to_lattice(U, site).clone().requires_grad_(True)
    Xt = odd_half_to_lattice(X, site, ls)
    L = -(torch.vdot(Xt.reshape(-1), tm.hermop(Ut, Xt).reshape(-1))).real
    (g,) = torch.autograd.grad(L, Ut)
    g = g.detach().numpy().reshape(U.shape)
    assert rel(grad, g) < 1e-9


def test_torch_autograd_function(small):
    """PseudofermionAction: forward value and backward gradient through
    torch.autograd, checked against a directional finite difference."""
    geo, fa, inv, U, phi = small
    links = torch.from_numpy(U.copy()).requires_grad_(True)
    S = q.pseudofermion_action(links, torch.from_numpy(phi), inv, 1e-13)
    S.backward()
    Uf = U.reshape(-1, 3, 3)
    T = su3_direction(Uf.shape[0], 7)
    fd = central_difference(inv, Uf, T, phi, 1e-5)
    dU = (T @ Uf).reshape(U.shape)
    an = np.real(np.vdot(links.grad.numpy(), dU))
    assert abs(fd - an) < 1e-6 * abs(an)
    inv.set_gauge(U)


def test_config_d_force_finite_differences():
    """16^3x32, Ls=8, mass=0.05 (sigma 0.3): force vs central finite
    differences (eps = 1e-5, FP64 solves to 1e-14) along 32 seeded su(3)
    directions, 1e-6 relative per direction.  S ~ 1e7 here, so FP64
    rounding of S alone (ulp ~ 2e-9) would limit a difference quotient to
    ~1e-4 absolute; the actions are therefore subtracted as compensated
    (hi, lo) pairs."""
    print("\n[gpu] force config D 16^3x32 Ls8, 32 FD directions", flush=True)
    geo, fa, inv, U, phi = make_case((16, 16, 16, 32), 8, 0.05, 0.3, 71)
    inv.set_max_num_iter(20000)
    S, F, _, it = q.fermion_force(inv, phi, 1e-14)
    Uf = U.reshape(-1, 3, 3)
    worst = 0.0
    for k in range(32):
        T = su3_direction(Uf.shape[0], 1000 + k)
        fd = central_difference(inv, Uf, T, phi, 1e-5)
        an = np.einsum("nij,nji->", F.reshape(-1, 3, 3), T).real
        worst = max(worst, abs(fd - an) / abs(an))
        if k % 8 == 7:
            print(f"[gpu] config D force: {k + 1} directions, worst so far {worst:.2e}", flush=True)
    print(f"[gpu] config D force: S={S:.6e} cg iters {it}, worst FD rel diff {worst:.2e}", flush=True)
    assert worst < 1e-6
