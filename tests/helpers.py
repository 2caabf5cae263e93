"""Shared fixtures-as-functions for the test suite (inputs only)."""

import numpy as np

from qlattice_amd.gauge import rand_gauge_links, twist_boundary_at_boundary


def unit_gauge(V):
    U = np.zeros((V, 4, 3, 3), dtype=np.complex128)
    for i in range(3):
        U[:, :, i, i] = 1.0
    return U


def rand_gauge(site, sigma, seed, antiperiodic=True):
    U = rand_gauge_links(site, sigma, seed)
    if antiperiodic:
        twist_boundary_at_boundary(U, site, -0.5, 3)
    return U


def rand_field(shape, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def point_src(V, site_index=0):
    p = np.zeros((V, 12, 12), dtype=np.complex128)
    p[site_index] = np.eye(12)
    return p


ZMOBIUS_OMEGA = [
    1.0903256131299373,
    0.9570283702230611,
    0.7048886040934104,
    0.48979921782791747,
    0.328608311201356,
    0.21664245377015995,
    0.14121112711957107,
    0.0907785101745156,
    complex(0.05608303440064219, -0.007537158177840385),
    complex(0.05608303440064219, 0.007537158177840385),
    complex(0.0365221637144842, -0.03343945161367745),
    complex(0.0365221637144842, 0.03343945161367745),
]


def qlat_check(v, ref, eps):
    """The reference's golden-value test: 2|v - ref| / (|v| + |ref|) <= eps
    (qlat_utils/utils.py:455-545)."""
    return 2 * abs(v - ref) / (abs(v) + abs(ref)) <= eps


def block_order(site, bs, eo=1):
    """half index (reference order) of element rows: [block][bidx] in the
    compressed layout (convert_half_vector, commpressed-eigen-io.cpp:925-947)."""
    nb = [site[m] // bs[m] for m in range(4)]
    bve = int(np.prod(bs)) // 2
    rows = []
    for b in range(int(np.prod(nb))):
        bc, r = [], b
        for m in range(4):
            bc.append(r % nb[m])
            r //= nb[m]
        blk = []
        for bidx in range(bve):
            r, lc = 2 * bidx, []
            for m in range(4):
                lc.append(r % bs[m])
                r //= bs[m]
            x = [bc[m] * bs[m] + lc[m] for m in range(4)]
            if (sum(x) & 1) != (1 if eo == 1 else 0):
                x[0] += 1
            full = x[0] + site[0] * (x[1] + site[1] * (x[2] + site[2] * x[3]))
            blk.append(full // 2)
        rows.append(blk)
    return np.array(rows)
