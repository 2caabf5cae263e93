"""The allocation-free Mpc^+ Mpc entry point
(qlat_dwf_multiply_hermop_sym2_device_noalloc) against the allocating one
(qlat_dwf_multiply_hermop_sym2_device) through the C ABI: two calls in a
row, the second on a smaller lattice, each bit-identical to the existing
entry point's result (the same kernels on the same inputs); a lattice
beyond the scratch maximum is rejected with a clear error."""

import numpy as np
import pytest
import torch

import qlattice_amd as q
from qlattice_amd import _lib
from tests.helpers import rand_field, rand_gauge

pytestmark = pytest.mark.gpu


def make_inv(site, ls, seed):
    geo = q.Geometry(site)
    gf = q.GaugeField(geo)
    gf.data = rand_gauge(site, 0.3, seed)
    return geo, q.InverterDomainWall(gf=gf, fa=q.FermionAction(mass=0.05, ls=ls, m5=1.8, mobius_scale=2.0))


def test_noalloc_hermop_matches_allocating_entry_point():
    L = _lib.load()
    cases = [((8, 8, 8, 16), 12, 901), ((4, 4, 4, 8), 6, 902)]  # the second call on a smaller input
    invs = [make_inv(site, ls, seed) for site, ls, seed in cases]
    for (geo, inv), (site, ls, seed) in zip(invs, cases):
        print(f"\n[gpu] no-alloc hermop {site} Ls {ls}", flush=True)
        for eo in (1, 2):
            x = torch.from_numpy(rand_field((geo.total_volume() // 2, ls, 12), seed + eo)).cuda()
            ref = torch.empty_like(x)
            L.multiply_hermop_sym2_device(inv.cdata, _lib.dev_ptr(ref), _lib.dev_ptr(x), eo)
            got = torch.full_like(x, float("nan"))
            L.multiply_hermop_sym2_device_noalloc(inv.cdata, _lib.dev_ptr(got), _lib.dev_ptr(x), eo)
            torch.cuda.synchronize()
            assert torch.equal(got, ref), (site, ls, eo)
            assert torch.linalg.vector_norm(ref).item() > 0


def test_noalloc_rejects_lattices_beyond_the_scratch():
    print("\n[gpu] no-alloc hermop size limit", flush=True)
    geo, inv = make_inv((16, 16, 16, 32), 8, 903)  # 5 fields x 100 MB > 64 MiB
    x = torch.zeros((geo.total_volume() // 2, 8, 12), dtype=torch.complex128, device="cuda")
    with pytest.raises(q.QlatDwfError, match="maximum"):
        _lib.load().multiply_hermop_sym2_device_noalloc(inv.cdata, _lib.dev_ptr(x), _lib.dev_ptr(x), 1)
    assert not x.any()  # untouched
