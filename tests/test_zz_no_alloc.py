"""cpd_fitness_batch_noalloc (no reference counterpart: for callers that cannot
allocate device memory) against cpd_fitness_batch through the C ABI: two
calls in a row, the second smaller, each bit for bit the batch path's results
(same kernels, same beam tables, exact integer bin sums), with sizes beyond
the scratch maximum refused."""
import ctypes

import numpy as np
import pytest

from compound_prism_designer import _native as N

from workloads import transmitting

pytestmark = pytest.mark.gpu


def call(entry, designs, sampling):
    n = len(designs)
    darr = (N.Design * n)(*designs)
    sarr = (N.Sampling * n)(*sampling)
    out = (N.Fitness * n)()
    status = (ctypes.c_int * n)()
    st = entry(None, darr, sarr, n, out, status)
    return st, np.array([[f.size, f.info, f.deviation] for f in out]), np.array(status[:])


def test_noalloc_matches_batch_on_two_calls_in_a_row():
    _, D, _, _ = transmitting(2, 40)
    # first call: 40 uneven designs (16..128 bins), 24 wavelengths x 8192 rays
    # each, plus one adaptive and one over-long design that it refuses per design
    samp = [N.Sampling(24, 8192, N.MODE_FIXED, 0) for _ in D]
    samp[3] = N.Sampling(24, 8192, N.MODE_ADAPTIVE, 0)
    samp[7] = N.Sampling(24, N.NOALLOC_MAX_RAYS + 64, N.MODE_FIXED, 0)
    st_a, got_a, ps_a = call(N.lib.cpd_fitness_batch_noalloc, D, samp)
    assert st_a == N.OK, N.last_error()
    assert ps_a[3] == N.ERR_INVALID_ARGUMENT and ps_a[7] == N.ERR_INVALID_ARGUMENT
    keep = [i for i in range(len(D)) if i not in (3, 7)]
    st_b, want_a, ps_b = call(N.lib.cpd_fitness_batch, [D[i] for i in keep], [samp[i] for i in keep])
    assert st_b == N.OK and (ps_b == 0).all()
    assert (ps_a[keep] == 0).all()
    assert np.array_equal(got_a[keep], want_a)
    assert np.isnan(got_a[[3, 7]]).all()
    # second call, smaller: 5 designs of fewer wavelengths and rays, so the
    # scratch still holds the first call's larger state beyond them
    small = [D[i] for i in (10, 11, 12, 13, 14)]
    samp2 = [N.Sampling(5 + i, 1000 + 64 * i, N.MODE_FIXED, 0) for i in range(5)]
    st_c, got_b, ps_c = call(N.lib.cpd_fitness_batch_noalloc, small, samp2)
    assert st_c == N.OK and (ps_c == 0).all()
    st_d, want_b, _ = call(N.lib.cpd_fitness_batch, small, samp2)
    assert st_d == N.OK
    assert np.array_equal(got_b, want_b)
    print(f"noalloc: {len(keep)} + 5 designs bit-identical, infos {np.round(want_a[:, 1], 3)}", flush=True)


def test_noalloc_refuses_sizes_beyond_the_scratch():
    _, D, _, _ = transmitting(0, 1)
    d = D[0]
    n = N.NOALLOC_MAX_DESIGNS + 1
    st, _, _ = call(N.lib.cpd_fitness_batch_noalloc, [d] * n, [N.Sampling(1, 64, N.MODE_FIXED, 0)] * n)
    assert st == N.ERR_INVALID_ARGUMENT and b"at most" in N.lib.cpd_last_error()
    # 2^18 + 1 wavelength rows over 3 designs
    rows = [N.Sampling(1 << 17, 64, N.MODE_FIXED, 0), N.Sampling(1 << 17, 64, N.MODE_FIXED, 0),
            N.Sampling(1, 64, N.MODE_FIXED, 0)]
    st, out, _ = call(N.lib.cpd_fitness_batch_noalloc, [d] * 3, rows)
    assert st == N.ERR_INVALID_ARGUMENT and b"scratch maximum" in N.lib.cpd_last_error()
