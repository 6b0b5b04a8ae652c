"""Randomised (seeded, deterministic) parity of the multi-step kernels against
the one-step site kernel: random dimension, form, stencil, ensemble of uneven
sides, steps per pass, launch switches and step count (so that passes,
shallower remainder passes and one-step launches mix).  Every case must give
the site kernel's bits."""
import numpy as np
import pytest

import axs
from fields import random_state

pytestmark = pytest.mark.gpu

PF, NT, GRAPH = axs.AXS_MARCH_PREFETCH, axs.AXS_MARCH_NT_STORE, axs.AXS_STEP_GRAPH
WG512, WG1024 = axs.AXS_MARCH_WG512, axs.AXS_MARCH_WG1024


def _depths(nd, stencil, precision, eom):
    """depths the multi-step kernels are built for (launch flags, rows 0)"""
    if nd == 2:
        if stencil == 0:
            return [2, 3, 4, 6, 8]
        return [1, 2, 3, 4] if precision == "fp32" else [1, 2]
    if stencil == 0:
        return [2, 3, 4] if precision == "fp32" else [2, 3]
    if precision == "fp32":
        return [1, 2] if stencil < 3 else [1]
    return [1]


def _case(seed):
    rng = np.random.default_rng(seed)
    nd = int(rng.choice([2, 3]))
    precision = str(rng.choice(["fp32", "fp64"]))
    eom = "hamiltonian" if precision == "fp64" and rng.random() < 0.25 else "field_rescaled"
    stencil = int(rng.choice([0, 0, 1, 2, 3]))
    hi = 300 if nd == 2 else 150
    nmem = int(rng.integers(1, 4))
    Ns = [int(rng.integers(32, hi // 2 + 1)) * 2 for _ in range(nmem)]
    Ns = [max(N, 64) for N in Ns]
    depth = int(rng.choice(_depths(nd, stencil, precision, eom)))
    flags = (PF if rng.random() < 0.7 else 0) | (NT if rng.random() < 0.7 else 0)
    if rng.random() < 0.3:
        flags |= GRAPH
    if nd == 3:
        flags |= WG1024 if precision == "fp32" else WG512
    nsteps = int(rng.integers(1, 14))
    return nd, precision, eom, stencil, Ns, depth, flags, nsteps


@pytest.mark.parametrize("seed", range(24))
def test_random_configuration_matches_site_kernel(seed):
    nd, precision, eom, stencil, Ns, depth, flags, nsteps = _case(seed)
    states = [random_state(nd, N, seed=seed * 10 + r, ker=True) for r, N in enumerate(Ns)]
    lats = []
    for kernel in ("simple", "march"):
        lat = axs.Lattice(nd, members=Ns, precision=precision, eom=eom, stencil_setting=stencil)
        if kernel == "simple":
            lat.set_kernel("simple")
        else:
            lat.set_kernel("march", depth=depth, flags=flags)
            assert lat.kernel_config()["depth"] == depth
        for r, s in enumerate(states):
            lat.set_state(*s, member=r)
        lat.tau = 1.0
        lat.step(nsteps)
        lats.append(lat)
    a, b = lats
    for r in range(len(Ns)):
        ga, gb = a.get_state(member=r), b.get_state(member=r)
        for k in ga:
            np.testing.assert_array_equal(
                ga[k], gb[k],
                err_msg=f"seed {seed}: nd {nd} {precision} {eom} stencil {stencil} sides {Ns} "
                        f"depth {depth} flags {flags} steps {nsteps}: member {r} {k}")
    assert a.tau == b.tau
