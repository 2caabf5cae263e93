"""examples-py/dslash-qed.py's recorded data signatures (dslash-qed.log.json)
reproduced by the oracle's restatement of the QED domain-wall operator and
solvers (oracle/qed.py), with the script's fields drawn from the restated
RngState: the links from the phase field, the operator and its adjoint on
non-unitary links, the CG solutions, the 4D solves, the spin propagators
with and without Wick phases, and the unit-gauge propagator against the
free-field restatement."""

import numpy as np
import pytest

from oracle import freefield
from oracle import qed as OQ
from tests import qed_script as S


class OracleBackend:
    @staticmethod
    def mul(f5, gf, mass, m5, ls, dagger, wick=None):
        return OQ.multiply_m_dwf_qed(f5, gf, mass, m5, ls, S.SITE, wick, dagger)

    @staticmethod
    def cg(f5, gf, mass, m5, ls, dagger, max_iter):
        return OQ.cg_with_m_dwf_qed(f5.reshape(S.V, ls, 4), gf, mass, m5, ls, S.SITE, None, dagger,
                                    max_num_iter=max_iter)[0]

    @staticmethod
    def inv4d(f4, gf, mass, m5, ls, dagger, max_iter):
        return OQ.invert_dwf_qed(f4, gf, mass, m5, ls, S.SITE, None, dagger, max_num_iter=max_iter)[0]

    @staticmethod
    def invqed(sp, gf, mass, m5, ls, wick, dagger, max_iter):
        sol = np.zeros_like(sp)
        for i in range(4):  # qed.cpp:60-70: column i as a 4D source
            sol[:, :, i] = OQ.invert_dwf_qed(np.ascontiguousarray(sp[:, :, i]), gf, mass, m5, ls, S.SITE, wick,
                                             dagger, max_num_iter=max_iter)[0]
        return sol

    @staticmethod
    def free(sp, mass, m5):
        prop = np.einsum("vab,cd->vacbd", sp, np.eye(3)).reshape(S.V, 12, 12)
        out = freefield.free_invert(prop, S.SITE, mass, m5).reshape(S.V, 4, 3, 4, 3)
        return np.ascontiguousarray(out[:, :, 0, :, 0])


@pytest.mark.parametrize("blocks", [(1,), (2,), (3, 4), (5,)])
def test_dslash_qed_goldens_oracle(blocks):
    out = S.replay(OracleBackend, blocks)
    bad = S.compare(out, blocks)
    assert out and not bad, bad
