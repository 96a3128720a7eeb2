"""Initial configurations on the host (run once per job, not on the step path).

fcc()               ConfigurationMaker::Execute (confmaker.cc:8-44)
initial_velocity()  Variables::SetInitialVelocity (variables.cc:82-108):
                    std::mt19937(seed) + uniform_real_distribution<double>(0,1),
                    |p| = V0 in a uniformly random direction, mean removed.
"""
import math

import numpy as np


def fcc(L, density):
    """FCC lattice with spacing (4/density)^(1/3), offset 1e-7, inside [0, L)."""
    Lx, Ly, Lz = (float(L),) * 3 if np.isscalar(L) else (float(v) for v in L)
    s = 1.0 / math.pow(density * 0.25, 1.0 / 3.0)
    hs = s * 0.5
    sx, sy, sz = int(Lx / s), int(Ly / s), int(Lz / s)
    e = 0.0000001
    iz, iy, ix = np.meshgrid(np.arange(sz, dtype=np.float64), np.arange(sy, dtype=np.float64),
                             np.arange(sx, dtype=np.float64), indexing="ij")
    ix, iy, iz = ix.ravel() * s, iy.ravel() * s, iz.ravel() * s
    basis = [
        (ix + e, iy + e, iz + e),
        (ix + e, (iy + hs) + e, (iz + hs) + e),
        ((ix + hs) + e, iy + e, (iz + hs) + e),
        ((ix + hs) + e, (iy + hs) + e, iz + e),
    ]
    x = np.empty((len(ix), 4, 3))
    for k, (a, b, c) in enumerate(basis):
        x[:, k, 0], x[:, k, 1], x[:, k, 2] = a, b, c
    x = x.reshape(-1, 3)
    inside = np.all((x >= 0.0) & (x < np.array([Lx, Ly, Lz])), axis=1)
    return np.ascontiguousarray(x[inside])


def _mt19937_raw(seed, n):
    """First n outputs of std::mt19937(seed) (32-bit Mersenne Twister)."""
    bg = np.random.MT19937()
    bg._legacy_seeding(int(seed) & 0xFFFFFFFF)
    return bg.random_raw(n).astype(np.uint64)


def _canonical(raw):
    """generate_canonical<double,53> over pairs of 32-bit draws."""
    lo = raw[0::2].astype(np.float64)
    hi = raw[1::2].astype(np.float64)
    r = (lo + hi * 4294967296.0) / 18446744073709551616.0
    return np.where(r >= 1.0, np.nextafter(1.0, 0.0), r)


def initial_velocity(n, v0, seed):
    u = _canonical(_mt19937_raw(seed, 4 * n))
    z = u[0::2] * 2.0 - 1.0
    phi = u[1::2] * 2.0 * math.pi
    p = np.empty((n, 3))
    cos, sin, sqrt = math.cos, math.sin, math.sqrt
    for i in range(n):
        r = v0 * sqrt(1 - z[i] * z[i])
        p[i, 0] = r * cos(phi[i])
        p[i, 1] = r * sin(phi[i])
    p[:, 2] = v0 * z
    # sequential mean, as the reference sums
    av = [0.0, 0.0, 0.0]
    for d in range(3):
        acc = 0.0
        for v in p[:, d].tolist():
            acc += v
        av[d] = acc / float(n)
    p -= np.array(av)
    return p
