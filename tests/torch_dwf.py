"""Plain FP64 PyTorch statement of the sym2 Schur complement Mpc (einsum /
roll on the full lattice), used as the autograd reference for the fermion
force.  Written from the operator definitions, independently of both the
HIP library and the C++ oracle:

  (H psi)(x) = -1/2 sum_mu [U_mu(x)(1 - g_mu) psi(x+mu) + U_mu(x-mu)^+ (1 + g_mu) psi(x-mu)]
  H^+ swaps (1 - g_mu) and (1 + g_mu)
  K = P M_ee^-1, P = diag(b) + diag(c) S, per chirality (Ls x Ls)
  Mpc = 1 - H K H K,  Mpc^+ = 1 - K^+ H^+ K^+ H^+      (dslash.cpp:1102-1138)

Fields: (T, Z, Y, X, Ls, 4, 3) complex128; gauge (T, Z, Y, X, 4, 3, 3).
"""

import numpy as np
import torch

from oracle import dense, freefield

AXIS = {0: 3, 1: 2, 2: 1, 3: 0}


class TorchMpc:
    def __init__(self, total_site, ls, mass, m5, bs, cs):
        self.site = total_site
        self.ls = ls
        g = freefield.cps_gammas()
        eye = np.eye(4)
        self.pm = [torch.tensor(0.5 * (eye - gm)) for gm in g]
        self.pp = [torch.tensor(0.5 * (eye + gm)) for gm in g]
        Mp, Mm = dense.ls_matrices(ls, mass, m5, bs, cs)
        Sp = np.zeros((ls, ls), complex)
        Sm = np.zeros((ls, ls), complex)
        for s in range(ls):
            if s >= 1:
                Sp[s, s - 1] = 1
            else:
                Sp[0, ls - 1] = -mass
            if s <= ls - 2:
                Sm[s, s + 1] = 1
            else:
                Sm[ls - 1, 0] = -mass
        P_p = np.diag(bs) + np.diag(cs) @ Sp
        P_m = np.diag(bs) + np.diag(cs) @ Sm
        self.K = [torch.tensor(P_p @ np.linalg.inv(Mp)), torch.tensor(P_m @ np.linalg.inv(Mm))]
        self.Kd = [k.conj().T.contiguous() for k in self.K]

    def apply_s(self, mats, psi):
        up = torch.einsum("st,...tac->...sac", mats[0], psi[..., :2, :])
        lo = torch.einsum("st,...tac->...sac", mats[1], psi[..., 2:, :])
        return torch.cat([up, lo], dim=-2)

    def hop(self, U, psi, dag=False):
        out = torch.zeros_like(psi)
        for mu in range(4):
            ax = AXIS[mu]
            Umu = U[..., mu, :, :]
            fwd = torch.roll(psi, -1, dims=ax)
            Pf = self.pp[mu] if dag else self.pm[mu]
            Pb = self.pm[mu] if dag else self.pp[mu]
            t = torch.einsum("ab,...sbc->...sac", Pf, fwd)
            out = out - torch.einsum("...ij,...saj->...sai", Umu, t)
            Ub = torch.roll(Umu, 1, dims=ax)
            bwd = torch.roll(psi, 1, dims=ax)
            t = torch.einsum("ab,...sbc->...sac", Pb, bwd)
            out = out - torch.einsum("...ji,...saj->...sai", Ub.conj(), t)
        return out

    def mpc(self, U, x):
        t = self.apply_s(self.K, x)
        t = self.hop(U, t)
        t = self.apply_s(self.K, t)
        return x - self.hop(U, t)

    def mpcdag(self, U, y):
        t = self.hop(U, y, dag=True)
        t = self.apply_s(self.Kd, t)
        t = self.hop(U, t, dag=True)
        return y - self.apply_s(self.Kd, t)

    def hermop(self, U, x):
        return self.mpcdag(U, self.mpc(U, x))


def odd_half_to_lattice(half, total_site, ls):
    """(V/2, Ls, 12) odd half field -> (T, Z, Y, X, Ls, 4, 3) torch tensor."""
    V = int(np.prod(total_site))
    full = np.zeros((V, ls, 12), dtype=np.complex128)
    dense.set_half(full, half, total_site, 1)
    return torch.from_numpy(dense.to_lattice(full, total_site, ls).copy())


def gauge_to_lattice(U, total_site):
    return torch.from_numpy(np.ascontiguousarray(dense.gauge_lattice(U, total_site)))
