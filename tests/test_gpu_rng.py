"""set_u_rand on the GPU (qlat_dwf_rng_fill_uniform) against the host
restatement, which the reference's goldens pin (tests/test_rng.py): the
numbers must be identical bit for bit."""

import numpy as np
import pytest

import qlattice_amd as q
from qlattice_amd.rng import set_u_rand, set_u_rand_device

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,upper,lower", [((32, 12, 12), 1.0, 0.0), ((32, 3), 3.141592653589793,
                                                                         -3.141592653589793),
                                                 ((32, 1), 1.0, -1.0), ((32, 5), 2.5, 0.5)],
                         ids=["prop", "u1-phases", "one-real", "odd-count"])
def test_device_fill_equals_host_fill(shape, upper, lower):
    print(f"\n[gpu] rng fill {shape} [{lower}, {upper})", flush=True)
    dt = np.complex128 if len(shape) > 2 or shape[1] == 3 else np.float64
    rs = q.RngState("fill").split(str(shape))
    a = set_u_rand(np.zeros(shape, dt), rs, upper, lower)
    b = set_u_rand_device(np.zeros(shape, dt), rs, upper, lower)
    assert np.array_equal(a, b)


def test_device_fill_after_draws_and_golden_source():
    rs = q.RngState(17)
    for _ in range(5):
        rs.rand_gen()  # newtype keeps the index (5), drops the cache
    a = set_u_rand(np.zeros((16, 7)), rs, 1.0, -1.0)
    b = set_u_rand_device(np.zeros((16, 7)), rs, 1.0, -1.0)
    assert np.array_equal(a, b)
    geo = q.Geometry((4, 4, 4, 8))
    rs = q.RngState("seed").split("src_r")
    host = q.Prop(geo).set_rand(rs)
    dev = q.Prop(geo).set_rand(rs, device=True)
    assert np.array_equal(host.data, dev.data)  # the free-invert golden's source


@pytest.mark.parametrize("shape,center,sigma", [((32, 12, 12), 0.0, 1.0), ((32, 4, 16), 0.0, 0.3),
                                                 ((32, 3), 1.0, 0.5)], ids=["prop", "qed-5d", "odd"])
def test_device_gaussian_fill_matches_host(shape, center, sigma):
    """The same per-site streams, pairs and rejections as set_g_rand; the
    device logarithm may differ from the host's in the last bit."""
    from qlattice_amd.rng import set_g_rand, set_g_rand_device

    print(f"\n[gpu] gaussian fill {shape}", flush=True)
    rs = q.RngState("gauss").split(str(shape))
    a = set_g_rand(np.zeros(shape, np.complex128), rs, center, sigma)
    b = set_g_rand_device(np.zeros(shape, np.complex128), rs, center, sigma)
    assert np.max(np.abs(a - b)) <= 1e-14 * np.max(np.abs(a))
    assert np.mean(a == b) > 0.9
