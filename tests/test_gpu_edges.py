"""Edge cases of the GPU path against the FP64 oracle: extreme bin and prism
counts, ray counts that are not a multiple of the wavefront, empty and
invalid inputs."""
import ctypes

import numpy as np
import pytest

import compound_prism_designer as cpd
from compound_prism_designer import _native as N
from oracle import oracle as O

from test_gpu_fitness import INFO_TOL, check_close, to_oracle

pytestmark = pytest.mark.gpu


def design_with(prisms: int, bins: int):
    """A design of the uneven distribution with `prisms` glasses, re-binned to
    `bins` bins over the same array length."""
    P, D, S, _ = cpd.generate_designs(2, 0, 400)
    for p in P:
        if p.n_glass == prisms:
            q = N.Params()
            ctypes.memmove(ctypes.byref(q), ctypes.byref(p), ctypes.sizeof(q))
            q.bin_count = bins
            q.bin_size = q.linear_slope = q.det_length / bins
            d = N.Design()
            assert N.lib.cpd_build_design(ctypes.byref(q), ctypes.byref(d)) == 0
            return d
    raise AssertionError(f"no {prisms}-prism design")


@pytest.mark.parametrize("prisms,bins", [(1, 1), (1, 1024), (6, 1024), (6, 3), (3, 129)])
def test_extreme_bins_and_prisms_match_oracle(prisms, bins):
    # 1024 bins: 2 LDS columns per bin, folded into the u64 row every 960 rays
    d = design_with(prisms, bins)
    got, st = cpd.fitness_batch([d], N.Sampling(6, 6000, N.MODE_FIXED, 0))
    assert st[0] == 0
    ref, _ = O.fitness(to_oracle(d), 6, 6000, 64, adaptive=False, qrng=O.QRNG_FIXED_POINT)
    check_close(got[0], np.array(ref), f"{prisms} prisms, {bins} bins")


@pytest.mark.parametrize("n_rays", [1, 63, 65, 4097])
def test_ray_counts_off_the_wavefront_multiple(n_rays):
    # the last ray (k = n_rays) is the one whose bin divides by n_rays + 1
    d = design_with(3, 32)
    got, st = cpd.fitness_batch([d], N.Sampling(5, n_rays, N.MODE_FIXED, 0))
    assert st[0] == 0
    ref, _ = O.fitness(to_oracle(d), 5, n_rays, 64, adaptive=False, qrng=O.QRNG_FIXED_POINT)
    check_close(got[0], np.array(ref), f"{n_rays} rays")


def test_empty_and_invalid_inputs():
    d = design_with(2, 16)
    out = (N.Fitness * 1)()
    status = (ctypes.c_int * 1)()
    assert N.lib.cpd_fitness_batch(None, None, None, 0, None, None) == N.OK  # empty batch
    bad = [N.Sampling(0, 4096, N.MODE_FIXED, 0), N.Sampling(4, 0, N.MODE_FIXED, 0),
           N.Sampling(4, 4096, 7, 0)]
    for s in bad:
        got, st = cpd.fitness_batch([d], s)
        assert st[0] == N.ERR_INVALID_ARGUMENT and np.isnan(got[0]).all()
    e = N.Design()
    ctypes.memmove(ctypes.byref(e), ctypes.byref(d), ctypes.sizeof(e))
    e.bin_count = 2048  # above CPD_MAX_BINS
    got, st = cpd.fitness_batch([e, d], N.Sampling(4, 2048, N.MODE_FIXED, 0))
    assert st[0] == N.ERR_INVALID_ARGUMENT and st[1] == 0
    ref, _ = O.fitness(to_oracle(d), 4, 2048, 64, adaptive=False, qrng=O.QRNG_FIXED_POINT)
    check_close(got[1], np.array(ref), "valid design beside an invalid one")
    s = N.Sampling(4, 4096, N.MODE_FIXED, 0)
    f = N.Fitness()
    assert N.lib.cpd_fitness_gpu(None, ctypes.byref(e), ctypes.byref(s), ctypes.byref(f)) == N.ERR_INVALID_ARGUMENT
    seeds = np.array([0.5])
    assert N.lib.cpd_gpu_fitness(None, ctypes.byref(d), seeds.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1,
                                 8, 0, 64, ctypes.byref(f)) == N.ERR_INVALID_ARGUMENT  # nwarp = 0


def test_gpu_fitness_without_seeds_is_not_converged():
    # cuda_fitness with an empty seed list returns None
    d = design_with(3, 32)
    got, st = cpd.gpu_fitness_batch([d], np.array([]), max_n=8, nwarp=2, max_eval=256)
    assert st[0] == N.ERR_NOT_CONVERGED
