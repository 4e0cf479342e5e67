"""mjg_kinetic_energy_static (the allocation-free entry point) against mjg_kinetic_energy, through
the C ABI: two calls in a row, the second on a smaller system, each bit-identical to the existing
entry point's result (same kernels, same sum order) and within 1e-12 of the float64 sum of the
downloaded velocities; a system above the static scratch size is refused with MJG_INVALID_ARG."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ke_numpy(spec, v):
    return 0.5 * float(np.sum(spec.mass[:, None] * v * v))


def test_static_kinetic_energy_matches_the_allocating_entry_point():
    from mjolnir_amd import capi, workloads
    for n_side, seed in ((16, 31), (8, 32)):      # 4096 then 512 particles
        spec = workloads.lj_fluid(n_side=n_side, seed=seed)
        print(f"[gpu] static kinetic energy, LJ N={spec.n}", flush=True)
        e = capi.Engine(spec)
        e.step(20)
        _, v, _ = e.state()
        ke_static = e.kinetic_energy_static()
        ke_alloc = e.kinetic_energy()
        assert ke_static == ke_alloc, (ke_static, ke_alloc)
        assert abs(ke_static - _ke_numpy(spec, v)) <= 1e-12 * abs(ke_static)
        e.close()


def test_static_kinetic_energy_refuses_more_particles_than_its_scratch():
    from mjolnir_amd import capi
    L = capi.lib()
    h = C.c_void_p()
    n = (1 << 20) + 1
    print(f"[gpu] static kinetic energy refusal, N={n}", flush=True)
    lo, hi = np.zeros(3), np.ones(3)
    assert L.mjg_engine_create(capi.FP64, capi.UNLIMITED, lo, hi, n, 0, C.byref(h)) == capi.OK
    try:
        ke = C.c_double(-1.0)
        assert L.mjg_kinetic_energy_static(h, C.byref(ke)) == capi.INVALID_ARG
        assert b"static scratch" in L.mjg_last_error(h)
        assert ke.value == -1.0
    finally:
        L.mjg_engine_destroy(h)
