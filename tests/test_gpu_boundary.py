"""The drop-in binding of the reference's `fitness` (src/fitness.rs:130-147),
step for step as the Rust snippet in the integration guide makes the call: a
repr(C) sampling record of {max_n, max_m, CPD_MODE_ADAPTIVE, 0}, the default
context (NULL), the design record the crate's construction produces, and
DesignFitness from the three doubles.  Its results must be the oracle's
fitness() with the SEM early exits, bit for bit."""
import ctypes

import numpy as np
import pytest

import compound_prism_designer as cpd
from compound_prism_designer import _native as N
from oracle import oracle as O
import kat

from test_gpu_fitness import kat_design, to_oracle

pytestmark = pytest.mark.gpu

CPD_MODE_ADAPTIVE = 1


class CpdSampling(ctypes.Structure):
    # #[repr(C)] pub struct CpdSampling { n_wavelengths: u32, n_rays: u32, mode: u32, reserved: u32 }
    _fields_ = [("n_wavelengths", ctypes.c_uint32), ("n_rays", ctypes.c_uint32), ("mode", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class CpdFitness(ctypes.Structure):
    _fields_ = [("size", ctypes.c_double), ("info", ctypes.c_double), ("deviation", ctypes.c_double)]


def rust_fitness(design, max_n, max_m):
    """pub fn fitness(spec, max_n, max_m) -> DesignFitness<f64>"""
    s = CpdSampling(max_n, max_m, CPD_MODE_ADAPTIVE, 0)
    out = CpdFitness()
    st = N.lib.cpd_fitness_gpu(None, ctypes.byref(design), ctypes.cast(ctypes.byref(s), ctypes.POINTER(N.Sampling)),
                               ctypes.cast(ctypes.byref(out), ctypes.POINTER(N.Fitness)))
    assert st == 0, N.last_error()  # the crate's fitness cannot fail: a device error panics
    return out.size, out.info, out.deviation


def test_record_layouts_match_the_rust_binding():
    assert ctypes.sizeof(CpdSampling) == ctypes.sizeof(N.Sampling) == 16
    assert ctypes.sizeof(CpdFitness) == ctypes.sizeof(N.Fitness) == 24
    assert ctypes.sizeof(N.Design) == 672


def test_known_design_through_the_binding():
    # src/fitness.rs:250-257: fitness(&spec, 16_384, 16_384)
    d = kat_design()
    got = rust_fitness(d, 16384, 16384)
    ref, _ = O.fitness(to_oracle(d), 16384, 16384, 64, adaptive=True, qrng=O.QRNG_FIXED_POINT)
    assert got == ref
    assert abs(got[1] - kat.KAT_INFO) <= kat.KAT_INFO_RMAX * kat.KAT_INFO


@pytest.mark.parametrize("config", [5, 6])
def test_shipped_configurations_through_the_binding(config):
    _, D, _, _ = cpd.generate_designs(config, 40, 4)
    for d in D:
        got = rust_fitness(d, 16384, 16384)
        ref, _ = O.fitness(to_oracle(d), 16384, 16384, 64, adaptive=True, qrng=O.QRNG_FIXED_POINT)
        assert got == ref, (config, got, ref)
