"""QED domain-wall operator and solvers on the GPU (qed.cpp:119-411) against
the numpy restatement (oracle/qed.py, pinned on the CPU by
tests/test_qed_oracle.py) and against the GPU SU(3) Shamir operator with
links u * 1 on one colour."""

import numpy as np
import pytest

import qlattice_amd as q
from oracle import qed as Q
from tests.helpers import rel

pytestmark = pytest.mark.gpu

SITE = (4, 4, 2, 6)
LS = 6
V = int(np.prod(SITE))


@pytest.fixture(scope="module")
def data():
    print("\n[gpu] QED DWF 4x4x2x6 Ls6", flush=True)
    rng = np.random.default_rng(11)
    u = np.exp(1j * rng.uniform(0, 2 * np.pi, (V, 4)))
    g = u * (1 + 0.2 * rng.standard_normal((V, 4)))
    wick = 1.1 * np.exp(1j * rng.uniform(-0.3, 0.3, SITE[3]))
    x = rng.standard_normal((V, LS, 4)) + 1j * rng.standard_normal((V, LS, 4))
    return u, g, wick, x


@pytest.mark.parametrize("dag", [False, True])
@pytest.mark.parametrize("case", ["unitary", "general"])
def test_multiply_matches_restatement(data, dag, case):
    u, g, wick, x = data
    links, w = (u, None) if case == "unitary" else (g, wick)
    got = q.multiply_m_dwf_qed(x, links, 0.1, 1.8, LS, SITE, w, dag)
    ref = Q.multiply_m_dwf_qed(x, links, 0.1, 1.8, LS, SITE, w, dag)
    assert rel(got, ref) < 1e-13


def test_multiply_matches_gpu_su3_shamir(data):
    u, g, wick, x = data
    geo = q.Geometry(SITE)
    gf = q.GaugeField(geo)
    gf.data = u[..., None, None] * np.eye(3)[None, None]
    inv = q.InverterDomainWall(gf=gf, fa=q.FermionAction(mass=0.1, ls=LS, m5=1.8, mobius_scale=1.0))
    f = q.FermionField5d(geo, LS, 0)
    f.data = np.zeros((V, LS, 12), np.complex128)
    f.data[..., 0::3] = x
    su3 = q.multiply_m(None, f, inv).data
    assert rel(q.multiply_m_dwf_qed(x, u, 0.1, 1.8, LS, SITE), su3[..., 0::3]) < 1e-13


def test_cg_and_invert_match_restatement(data):
    u, g, wick, x = data
    sol, it = q.cg_with_m_dwf_qed(x, g, 0.1, 1.0, LS, SITE, wick, False, 1e-10, 3000)
    ref, it_ref = Q.cg_with_m_dwf_qed(x, g, 0.1, 1.0, LS, SITE, wick, False, 1e-10, 3000)
    assert rel(sol, ref) < 1e-8 and abs(it - it_ref) <= 2, (it, it_ref)
    assert q.cg_with_m_dwf_qed(x, g, 0.1, 1.0, LS, SITE, wick, False, 1e-10, 0)[1] == 0
    assert q.cg_with_m_dwf_qed(x, g, 0.1, 1.0, LS, SITE, wick, False, 1e-10, 3)[1] == 4  # not converged
    src4 = x[:, 0, :]
    for dag in (False, True):
        out, it = q.invert_dwf_qed(src4, g, 0.1, 1.0, LS, SITE, wick, dag, 1e-10, 3000)
        ref4, _ = Q.invert_dwf_qed(src4, g, 0.1, 1.0, LS, SITE, wick, dag, 1e-10, 3000)
        assert rel(out, ref4) < 1e-8, dag
    sp = np.zeros((V, 4, 4), np.complex128)
    sp[0] = np.eye(4)
    sol_sp, _ = q.invert_qed(sp, u, 0.1, 1.0, LS, SITE, None, False, 1e-10, 3000)
    col2, _ = q.invert_dwf_qed(sp[:, :, 2], u, 0.1, 1.0, LS, SITE, None, False, 1e-10, 3000)
    assert rel(sol_sp[:, :, 2], col2) < 1e-14


class _GpuBackend:
    """The GPU entry points behind examples-py/dslash-qed.py's calls."""

    @staticmethod
    def mul(f5, gf, mass, m5, ls, dagger):
        from tests.qed_script import SITE

        return q.multiply_m_dwf_qed(f5, gf, mass, m5, ls, SITE, None, dagger)

    @staticmethod
    def cg(f5, gf, mass, m5, ls, dagger, max_iter):
        from tests.qed_script import SITE

        return q.cg_with_m_dwf_qed(f5, gf, mass, m5, ls, SITE, None, dagger, max_num_iter=max_iter)[0]

    @staticmethod
    def inv4d(f4, gf, mass, m5, ls, dagger, max_iter):
        from tests.qed_script import SITE

        return q.invert_dwf_qed(f4, gf, mass, m5, ls, SITE, None, dagger, max_num_iter=max_iter)[0]

    @staticmethod
    def invqed(sp, gf, mass, m5, ls, wick, dagger, max_iter):
        from tests.qed_script import SITE

        return q.invert_qed(sp, gf, mass, m5, ls, SITE, wick, dagger, max_num_iter=max_iter)[0]

    @staticmethod
    def free(sp, mass, m5):
        from tests.qed_script import SITE, V

        prop = q.Prop(q.Geometry(SITE))
        prop.data = np.einsum("vab,cd->vacbd", sp, np.eye(3)).reshape(V, 12, 12)
        out = q.free_invert(prop, mass, m5).data.reshape(V, 4, 3, 4, 3)
        return np.ascontiguousarray(out[:, :, 0, :, 0])


@pytest.mark.parametrize("blocks", [(1,), (2,), (3, 4), (5,)])
def test_dslash_qed_goldens_gpu(blocks):
    """examples-py/dslash-qed.py's recorded data signatures
    (dslash-qed.log.json) from the GPU operator and solvers, with the
    script's fields drawn from the restated RngState."""
    from tests import qed_script as S

    print(f"\n[gpu] dslash-qed goldens, script blocks {blocks}", flush=True)
    out = S.replay(_GpuBackend, blocks)
    bad = S.compare(out, blocks)
    assert out and not bad, bad
