"""CPU pins of the PyTorch autograd reference used by the force tests."""

import numpy as np
import torch

from oracle import oracle as O
from oracle import dense
from tests.helpers import ZMOBIUS_OMEGA, rand_field, rand_gauge, rel
from tests.torch_dwf import TorchMpc, gauge_to_lattice, odd_half_to_lattice


def test_torch_hermop_matches_oracle():
    site, ls = (4, 4, 4, 4), 6
    U = rand_gauge(site, 0.4, 81)
    for omega in (None, ZMOBIUS_OMEGA[:ls]):
        orc = O.OracleInverter(site, U, mass=0.07, ls=ls, m5=1.8, mobius_scale=2.0, omega=omega)
        tm = TorchMpc(site, ls, 0.07, 1.8, orc.bs, orc.cs)
        x = rand_field((128, ls, 12), 82)
        ref = orc.apply_half("hermop_sym2", x, 1)
        got = tm.hermop(gauge_to_lattice(U, site), odd_half_to_lattice(x, site, ls))
        got = dense.get_half(dense.to_flat(got.numpy()), site, 1)
        assert rel(got, ref) < 1e-13
        ref = orc.apply_half("mpc_sym2", x, 1)
        got = tm.mpc(gauge_to_lattice(U, site), odd_half_to_lattice(x, site, ls))
        assert rel(dense.get_half(dense.to_flat(got.numpy()), site, 1), ref) < 1e-13


def test_torch_quadratic_form_gradient_by_finite_difference():
    site, ls = (2, 2, 2, 4), 4
    U = rand_gauge(site, 0.4, 83)
    tm = TorchMpc(site, ls, 0.1, 1.8, np.full(ls, 1.5 + 0j), np.full(ls, 0.5 + 0j))
    X = odd_half_to_lattice(rand_field((16, ls, 12), 84), site, ls)
    Ut = gauge_to_lattice(U, site).clone().requires_grad_(True)
    L = -(torch.vdot(X.reshape(-1), tm.hermop(Ut, X).reshape(-1))).real
    (g,) = torch.autograd.grad(L, Ut)
    rng = np.random.default_rng(85)
    D = torch.from_numpy(rng.standard_normal(Ut.shape) + 1j * rng.standard_normal(Ut.shape))
    eps = 1e-6

    def Lof(V):
        return -(torch.vdot(X.reshape(-1), tm.hermop(V, X).reshape(-1))).real.item()

    U0 = Ut.detach()
    fd = (Lof(U0 + eps * D) - Lof(U0 - eps * D)) / (2 * eps)
    an = torch.real(torch.vdot(g.reshape(-1), D.reshape(-1))).item()
    assert abs(fd - an) < 1e-7 * abs(an)
