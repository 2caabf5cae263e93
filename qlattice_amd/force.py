"""Pseudofermion action and fermion force on the MI355X.

S(U) = phi^+ (Mpc^+ Mpc)^-1 phi for an odd half field phi, with Mpc the
sym2 Schur complement of the even-odd Moebius operator (dslash.cpp:1102-1149).
The reference has no fermion force; the conventions follow its gauge force
(hmc.cpp:64-75, matrix-hmc.h:202-213):

* ``force[x, mu] = T_A(U_mu(x) G_mu(x))`` with dS = sum Re Tr[G dU] and
  T_A(M) = (M - M^+)/2 - Tr(M - M^+)/6; for U -> exp(eps T) U (T in su(3))
  dS/deps = sum_x,mu Tr[force T];
* ``grad = G^+`` is the derivative with respect to the complex link entries
  in PyTorch's convention (dS = Re sum conj(grad) dU), which is what
  :class:`PseudofermionAction` returns to autograd.

The adjoint is a hand-written HIP kernel (``k_force``); X = (Mpc^+ Mpc)^-1
phi and Y = Mpc X are recomputed on every call, nothing is stored along a
trajectory.  PyTorch only wraps it for autograd.
"""

import numpy as np

from . import _lib
from .fields import FermionField5d, GaugeField


def set_gauge(inv, gf):
    """Give `inv` new links (InverterDomainWall::setup again,
    dslash.h:154-162)."""
    g = np.ascontiguousarray(gf.data if isinstance(gf, GaugeField) else gf, dtype=np.complex128)
    assert g.shape == (inv.geo().total_volume(), 4, 3, 3)
    _lib.load().inverter_set_gauge(inv.cdata, _lib.ptr(g), False)


def fermion_force(inv, phi, stop_rsd=1e-12, with_grad=False, action_pair=False):
    """Returns (S, force, grad_or_None, cg_iterations) for the odd half
    field `phi` (FermionField5d with eo=1, or an array (V/2, Ls, 12)).

    S = 2 Re <phi, X> - |Mpc X|^2 (second order in the solver error),
    summed with compensation; with action_pair=True, S is returned as
    (hi, lo), hi the FP64 rounding of the sum and lo the rest, so that the
    difference of two nearby actions can be formed to beyond FP64."""
    data = phi.data if isinstance(phi, FermionField5d) else phi
    if isinstance(phi, FermionField5d):
        assert phi.eo == 1, "the pseudofermion lives on odd sites"
    p = np.ascontiguousarray(data, dtype=np.complex128)
    V = inv.geo().total_volume()
    assert p.shape == (V // 2, inv.fa.ls(), 12)
    force = np.empty((V, 4, 3, 3), dtype=np.complex128)
    grad = np.empty_like(force) if with_grad else None
    S, S_lo, it = _lib.load().fermion_force(inv.cdata, _lib.ptr(p), float(stop_rsd), _lib.ptr(force),
                                            _lib.ptr(grad) if with_grad else 0)
    return (S, S_lo) if action_pair else S, force, grad, it


def _torch_function():
    import torch

    class PseudofermionAction(torch.autograd.Function):
        """S(U) = phi^+ (Mpc^+ Mpc)^-1 phi as a differentiable function of the
        links U ((V, 4, 3, 3) complex128 tensor, CPU or GPU).  The backward
        pass is the HIP adjoint kernel; the forward solve is redone per call."""

        @staticmethod
        def forward(ctx, links, phi, inv, stop_rsd):
            U = links.detach().to("cpu").contiguous().numpy()
            set_gauge(inv, U)
            ph = phi.detach().to("cpu").contiguous().numpy() if isinstance(phi, torch.Tensor) else phi
            S, _, grad, _ = fermion_force(inv, ph, stop_rsd, with_grad=True)
            ctx.save_for_backward(torch.from_numpy(grad).to(links.device))
            return torch.tensor(S, dtype=torch.float64, device=links.device)

        @staticmethod
        def backward(ctx, gS):
            (grad,) = ctx.saved_tensors
            return gS * grad, None, None, None

    return PseudofermionAction


_PFA = None


def pseudofermion_action(links, phi, inv, stop_rsd=1e-12):
    """Differentiable S(U) (see :class:`PseudofermionAction`)."""
    global _PFA
    if _PFA is None:
        _PFA = _torch_function()
    return _PFA.apply(links, phi, inv, stop_rsd)
