"""sjf_dft_flux_noalloc, the allocation-free flux read-out, against sjf_dft_flux.

Both entry points run the same two reduction kernels; the allocation-free one keeps the partial
sums and results in the library's static device buffer instead of a per-call allocation.  Two
calls in a row (a 20480-cell plane at 40 frequencies, which fills all 64 partial sums per
frequency, then a 96-cell plane at 3 frequencies) must each equal the allocating call bit for
bit: the second call must not see what the first left in the shared buffer.
"""
import numpy as np
import pytest

from sim_juncs_amd import materials as M
from sim_juncs_amd import scenario
from sim_juncs_amd.engine import Engine

pytestmark = pytest.mark.gpu

TANGENTIAL_Z = ["Ex", "Ey", "Hx", "Hy"]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_flux_equals_allocating_flux_on_two_calls(precision):
    shape, dx = (160, 128, 40), 0.02
    print(f"[case] noalloc flux {shape} {precision}", flush=True)
    rng = np.random.default_rng(11)
    eng = Engine(shape, dx, precision=precision)
    mats = rng.integers(0, 2, size=shape + (3,))
    eng.set_materials(scenario.pack(mats), [M.VACUUM, M.MaterialSpec(2.13)])
    eng.set_pml([4, 4, 4, 4, 6, 6])
    eng.add_source(**scenario.plane_source(shape, dx, 12, scenario.gaussian_pulse(0.8, 1.5, 0.2, start_fs=0.0)))
    freqs_big = np.linspace(0.9, 1.6, 40)
    big = eng.add_dft((0, 0, 25), (shape[0], shape[1], 26), TANGENTIAL_Z, freqs_big)
    small = eng.add_dft((70, 60, 30), (82, 68, 31), TANGENTIAL_Z, [1.1, 1.25, 1.4])
    eng.step(120)
    a_big = eng.dft_flux_noalloc(big, 2)
    a_small = eng.dft_flux_noalloc(small, 2)
    r_big = eng.dft_flux(big, 2)
    r_small = eng.dft_flux(small, 2)
    assert a_big.shape == (40,) and a_small.shape == (3,)
    assert np.all(np.isfinite(a_big)) and np.abs(r_big).max() > 0 and np.abs(r_small).max() > 0
    assert np.array_equal(a_big, r_big)
    assert np.array_equal(a_small, r_small)
    # and a third call, after a step, on the big plane again
    eng.step(7)
    assert np.array_equal(eng.dft_flux_noalloc(big, 2), eng.dft_flux(big, 2))
    eng.close()
