"""mdgpu_observe_noalloc (the observers without any device allocation) against
mdgpu_observe, through the C ABI: the same kernels and summation order over a
fixed device scratch array instead of the domain's growable buffer, so the
results must be the same bits.  Two calls in a row, the second on a smaller
system, then the error for a domain beyond the scratch array's brick limit."""
import ctypes
import math

import numpy as np
import pytest

from mdacp_amd import mdgpu
from mdacp_amd.lattice import fcc, initial_velocity

pytestmark = pytest.mark.gpu


def observe(d, fn):
    ke, pe, vir = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    rc = fn(d.h, ctypes.byref(ke), ctypes.byref(pe), ctypes.byref(vir))
    return rc, (ke.value, pe.value, vir.value)


def domain(L, steps, seed):
    x = fcc(L, 0.5)
    d = mdgpu.Domain(L, dt=0.001, rc=2.5, margin=0.5)
    d.add_particles(x, initial_velocity(len(x), math.sqrt(3.0), seed))
    d.make_pair_list()
    d.calculate(steps)
    return d


def test_observe_noalloc_equals_observe_on_two_calls():
    lib = mdgpu.load()
    results = []
    for L, steps in [(40.0, 30), (16.0, 10)]:  # 32000 particles, then 2048
        print(f"no-alloc observe: L={L} steps={steps}", flush=True)
        d = domain(L, steps, 1)
        rc_a, a = observe(d, lib.mdgpu_observe)
        rc_b, b = observe(d, lib.mdgpu_observe_noalloc)
        assert rc_a == 0 and rc_b == 0, lib.mdgpu_last_error(d.h).decode()
        assert a == b, (L, a, b)
        assert all(np.isfinite(a)) and a[0] > 0.0
        results.append(b)
        d.close()
    assert results[0] != results[1]


def test_observe_noalloc_rejects_more_bricks_than_its_scratch():
    lib = mdgpu.load()
    L = 160.0  # 53 cells per axis: 53^3 one-cell bricks, more than the 131072 the scratch holds
    rng = np.random.default_rng(4)
    x = rng.uniform(0.0, L, (3000, 3))
    d = mdgpu.Domain(L, dt=0.001, rc=2.5, margin=0.5)
    d.add_particles(x, np.zeros_like(x))
    d.set_option(mdgpu.OPT_BRICK, 1 | 1 << 8 | 1 << 16)
    d.make_pair_list()
    print("no-alloc observe: bricks", d.list_stats(), flush=True)
    rc, _ = observe(d, lib.mdgpu_observe_noalloc)
    assert rc == mdgpu.ERR_CAPACITY
    assert "scratch limit" in lib.mdgpu_last_error(d.h).decode()
    rc_a, a = observe(d, lib.mdgpu_observe)
    assert rc_a == 0 and np.isfinite(a[1])
    d.close()
