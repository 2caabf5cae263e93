"""FermionAction: Moebius / zMoebius domain-wall parameters.

Mirrors qlat's ``FermionAction`` (qlat/qlat/fermion_action.py:11-45) over the
C++ struct of fermion-action.h:8-53: b_s = 0.5*scale + 0.5,
c_s = 0.5*scale - 0.5 with scale = mobius_scale = b + c (b - c = 1); with
``omega`` (zMoebius) b_s = 0.5*(1/omega_s + 1), c_s = b_s - 1
(cqlat/fermion-action.cpp:17-37).
"""

import numpy as np


class FermionAction:
    def __init__(self, *, mass, ls, m5, mobius_scale=1.0, omega=None,
                 is_multiplying_dminus=True):
        assert isinstance(mass, float)
        assert isinstance(ls, int)
        assert isinstance(m5, float)
        self._mass = mass
        self._ls = ls
        self._m5 = m5
        self.is_multiplying_dminus = bool(is_multiplying_dminus)
        self.cg_diagonal_mee = 2
        if omega is None:
            self.is_using_zmobius = False
            self._mobius_scale = float(mobius_scale)
            self.bs = np.full(ls, 0.5 * self._mobius_scale + 0.5, dtype=np.complex128)
            self.cs = np.full(ls, 0.5 * self._mobius_scale - 0.5, dtype=np.complex128)
            self._omega = None
        else:
            assert isinstance(omega, list)
            assert ls == len(omega)
            self.is_using_zmobius = True
            self._mobius_scale = 0.0
            om = np.asarray(omega, dtype=np.complex128)
            self.bs = 0.5 * (1.0 / om + 1.0)
            self.cs = self.bs - 1.0
            self._omega = [complex(w) for w in om]

    def mass(self):
        return self._mass

    def ls(self):
        return self._ls

    def m5(self):
        return self._m5

    def omega(self):
        """1/(b_s + c_s) for zMoebius, None otherwise
        (cqlat/fermion-action.cpp:83-100)."""
        if not self.is_using_zmobius:
            return None
        return list(1.0 / (self.bs + self.cs))

    def mobius_scale(self):
        return self._mobius_scale
