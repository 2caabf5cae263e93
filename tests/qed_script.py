"""examples-py/dslash-qed.py replayed against a backend: the data
signatures it records, in the order it records them, for comparison with
tests/golden/dslash_qed.json.  A backend supplies

  mul(f5, gf, mass, m5, ls, dagger)                   (V, ls, 4)
  cg(f5, gf, mass, m5, ls, dagger, max_iter)          (V, ls, 4)
  inv4d(f4, gf, mass, m5, ls, dagger, max_iter)       (V, 4)
  invqed(sp, gf, mass, m5, ls, wick, dagger, max_iter) (V, 4, 4)
  free(sp, mass, m5)                                  (V, 4, 4)
"""

import cmath
import json
import os

import numpy as np

import qlattice_amd as q

SITE = (4, 4, 4, 8)
V = 512
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "dslash_qed.json")))["entries"]
WICK = [1.0, 1.0, cmath.rect(1.0, 0.5 * cmath.pi / 2), cmath.rect(1.0, 0.8 * cmath.pi / 2),
        cmath.rect(1.0, 0.8 * cmath.pi / 2), cmath.rect(1.0, 0.5 * cmath.pi / 2), 1.0, 1.0]


_BLOCK_HEADERS = {"test 4d propagator with free gauge field": 2,
                  "test 5d multiplication and inversion with complex gauge field": 3,
                  "test 4d propagator with complex gauge field": 4,
                  "test 4d propagator with Minkowski time": 5}


def golden_values(blocks=(1, 2, 3, 4, 5)):
    """(name, value, eps) of the recorded signatures of the given script
    blocks (a block starts at its section line)."""
    vals, block = [], 1
    for e in GOLD:
        if len(e) == 1:
            block = _BLOCK_HEADERS.get(e[0], block)
        elif block in blocks:
            vals.append((e[0], np.array([complex(*z) for z in e[1]]), e[2]))
    return vals


def check_arr(v, ref, eps):
    """check_log_json's metric (qlat_utils utils.py:531-537)."""
    v, ref = np.asarray(v), np.asarray(ref)
    return 2 * np.linalg.norm(v - ref) <= eps * (np.linalg.norm(v) + np.linalg.norm(ref))


def g_rand(shape, rs, sigma=1.0):
    a = np.zeros(shape, np.complex128)
    return q.set_g_rand(a, rs, 0.0, sigma)


def rel2(a, b):
    return np.vdot(a - b, a - b).real / np.vdot(b, b).real


def replay(be, blocks=(1, 2, 3, 4, 5)):
    rs = q.RngState("seed")
    out = []

    def rec(name, x):
        out.append((name, q.get_data_sig_arr(x, rs, 3)))

    def five_d(gf, mass, m5, ls, it):
        f_in = g_rand((V, ls, 4), rs.split("f_in"))
        rec("f_in data sig", f_in)
        f_out = be.mul(f_in, gf, mass, m5, ls, False)
        rec("f_out data sig", f_out)
        f_in2 = be.cg(be.mul(f_out, gf, mass, m5, ls, True), gf, mass, m5, ls, False, it)
        rec("f_in2 data sig", f_in2)
        assert rel2(f_in2, f_in) < 1e-8
        f_out = be.mul(f_in, gf, mass, m5, ls, True)
        rec("f_out dagger data sig", f_out)
        f_in3 = be.cg(be.mul(f_out, gf, mass, m5, ls, False), gf, mass, m5, ls, True, it)
        rec("f_in3 dagger data sig", f_in3)
        assert rel2(f_in3, f_in) < 1e-8

    def prop(gf, mass, m5, ls, it, wick=None, sigma=1.0, dagger=True):
        sp_in = g_rand((V, 4, 4), rs.split("sp_in"), sigma)
        rec("sp_in data sig", sp_in)
        sp_out = be.invqed(sp_in, gf, mass, m5, ls, wick, False, it)
        rec("sp_out data sig", sp_out)
        if dagger:
            rec("sp_out dagger data sig", be.invqed(sp_in, gf, mass, m5, ls, wick, True, it))
        return sp_in, sp_out

    if 1 in blocks:  # dslash-qed.py:23-99
        ph = np.zeros((V, 4))
        q.set_g_rand(ph, rs.split("gf_qed_phase"), 0.0, 1.0)
        gf = np.exp(1j * ph)
        rec("gf1_qed data sig", gf)
        five_d(gf, 0.1, 1.0, 16, 200)
        f_in4d = g_rand((V, 4), rs.split("f_in4d"))
        rec("f_in4d data sig", f_in4d)
        rec("f_out4d data sig", be.inv4d(f_in4d, gf, 0.1, 1.0, 16, False, 200))
        rec("f_out4d dagger data sig", be.inv4d(f_in4d, gf, 0.1, 1.0, 16, True, 200))
        prop(gf, 0.1, 1.0, 16, 200)
    if 2 in blocks:  # :101-123, the unit gauge field against free_invert
        gf = np.ones((V, 4), np.complex128)
        rec("gf1_qed data sig", gf)
        sp_in, sp_out = prop(gf, 0.1, 1.0, 16, 200, sigma=0.5, dagger=False)
        sp_out2 = be.free(sp_in, 0.1, 1.0)
        rec("sp_out2 data sig", sp_out2)
        assert rel2(sp_out, sp_out2) < 1e-8
    if 3 in blocks or 4 in blocks:  # :125-183, complex (non-unitary) links
        gf = g_rand((V, 4), rs.split("gf_qed"), 0.3) + 1.0
        rec("gf1_qed data sig", gf)
        if 3 in blocks:
            five_d(gf, 0.1, 1.0, 16, 400)
        if 4 in blocks:
            prop(gf, 0.1, 1.0, 16, 400)
    if 5 in blocks:  # :185-240, Wick phases per time slice
        gf = g_rand((V, 4), rs.split("gf_qed"), 0.1) + 1.0
        rec("gf1_qed data sig", gf)
        prop(gf, 0.2, 1.0, 8, 50000, wick=WICK)
    return out


def compare(out, blocks=(1, 2, 3, 4, 5)):
    """The entries of out that miss their recorded value: (name, error, eps)."""
    gold = golden_values(blocks)
    if [g[0] for g in gold] != [o[0] for o in out]:
        return [("order", [o[0] for o in out], [g[0] for g in gold])]
    bad = []
    for (name, val), (_, gval, eps) in zip(out, gold):
        if not check_arr(val, gval, eps):
            err = 2 * np.linalg.norm(val - gval) / (np.linalg.norm(val) + np.linalg.norm(gval))
            bad.append((name, err, eps))
    return bad
