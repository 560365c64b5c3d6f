"""Soundness of the FP32 trace's error margins, checked on the host.

The device traces rays in FP32 and re-traces in FP64 only the rays whose FP32
outcome lies within an error margin of a decision (compound-prism-designer_amd/
csrc/trace_core.h).  If a margin were too small, a ray could land in another
bin than the reference's FP64 trace without being re-traced.  tests/guard/
guard_emu.cpp compiles the same trace32 source for the host, with the
device's approximate reciprocal and square root modelled as the correctly
rounded value moved by a random -1, 0 or +1 ulp per call, and compares every
ray with the FP64 trace (trace64, the reference's operation order).

Per generator (the benchmark distributions 0, 2 and 4, and the reference's
shipped detector and coating settings 5, 6 and 7), designs that land >= 10% of
a probe's rays in a bin, >= 1e7 rays: no unflagged bin mismatch, and the
transmittance of every unflagged hit within the reference's own per-ray
tolerance of its GPU trace, 5e-3 relative (src/cuda_fitness.rs:605-607).  The
FP32 transmittance is least accurate near total internal reflection at the
lens, where cr = sqrt(1 - r^2 (1 - ci^2)) cancels (up to ~2e-3 relative);
a hit moves its bin's probability by t / M, so such rays change information
by far less than the 1e-4 bit tolerance.  Two builds: FP32
contraction as the device compiles it (fast, honouring the FP64 code's
contract(off) pragmas) and none at all."""
import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from compound_prism_designer import _native as N
from workloads import transmitting

HERE = Path(__file__).resolve().parent / "guard"
ROOT = HERE.parent.parent
CLANG = "/opt/rocm/llvm/bin/clang++"
T_RTOL = 5e-3

_libs = {}


def guard_lib(contract: str):
    if contract in _libs:
        return _libs[contract]
    out = HERE / "_build" / f"guard_emu_{contract}.so"
    out.parent.mkdir(exist_ok=True)
    src = HERE / "guard_emu.cpp"
    hdr = ROOT / "compound-prism-designer_amd" / "csrc" / "trace_core.h"
    if not out.exists() or out.stat().st_mtime < max(src.stat().st_mtime, hdr.stat().st_mtime):
        cmd = [CLANG, "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", f"-ffp-contract={contract}",
               "-DCPD_EMULATE_DEVICE_ROUNDING", "-I", str(ROOT / "include"),
               "-I", str(ROOT / "compound-prism-designer_amd" / "csrc"), str(src), "-o", str(out) + ".tmp"]
        subprocess.run(cmd, check=True)
        os.replace(str(out) + ".tmp", out)
    L = ctypes.CDLL(str(out))
    P = ctypes.POINTER
    L.guard_check.argtypes = [P(N.Design), ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                              ctypes.c_int, P(ctypes.c_uint64), ctypes.c_double * 2]
    _libs[contract] = L
    return L


def check(designs, n_lambda, n_rays, mode, contract):
    L = guard_lib(contract)
    n = len(designs)
    arr = (N.Design * n)(*designs)
    stats = (ctypes.c_uint64 * 6)()
    terr = (ctypes.c_double * 2)()
    L.guard_check(arr, n, n_lambda, n_rays, mode, min(16, os.cpu_count() or 1), stats, terr)
    return dict(zip(["rays", "flagged", "hits", "unflagged_mismatch", "t_off", "caught"], stats[:])), tuple(terr)


@pytest.mark.parametrize("config", [0, 2, 4, 5, 6, 7])
def test_no_unflagged_bin_mismatch_with_device_rounding(config):
    _, D, _, _ = transmitting(config, 40, first=1000)
    st, (t_abs, t_rel) = check(D, 16, 16384, 1, "fast-honor-pragmas")
    print(f"config {config}: {st}, flagged {st['flagged'] / st['rays']:.4%}, max |t32 - t64| {t_abs:.2e} "
          f"({t_rel:.2e} relative)")
    assert st["rays"] >= 10_000_000 and st["hits"] >= 1_000_000
    assert st["unflagged_mismatch"] == 0, st
    assert t_rel <= T_RTOL, t_rel


@pytest.mark.parametrize("config", [0, 5, 7])
def test_no_unflagged_bin_mismatch_without_contraction(config):
    _, D, _, _ = transmitting(config, 10, first=3000)
    st, (t_abs, t_rel) = check(D, 16, 16384, 1, "off")
    assert st["unflagged_mismatch"] == 0 and t_rel <= T_RTOL, (st, t_rel)


def test_the_margins_catch_real_fp32_errors():
    # rays whose FP32 bin differs from the FP64 one exist, and every one of them
    # is flagged: without the margins the comparison above would fail
    _, D, _, _ = transmitting(7, 40, first=1000)
    st, _ = check(D, 16, 16384, 1, "fast-honor-pragmas")
    print(st)
    assert st["caught"] > 0 and st["unflagged_mismatch"] == 0


def band_stats(designs, n_lambda, n_rays):
    L = guard_lib("fast-honor-pragmas")
    L.band_check.argtypes = [ctypes.POINTER(N.Design), ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint32,
                             ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]
    n = len(designs)
    arr = (N.Design * n)(*designs)
    st = (ctypes.c_uint64 * 7)()
    L.band_check(arr, n, n_lambda, n_rays, min(16, os.cpu_count() or 1), st)
    return dict(zip(["bands", "past_chain", "skipped", "skipped_with_hit", "with_hit", "dead_wavelengths",
                     "dead_with_hit"], st[:]))


@pytest.mark.parametrize("config,transmit,n_rays", [(1, False, 16384), (2, False, 16384), (4, False, 16384),
                                                    (0, True, 16384), (5, True, 16384), (6, True, 16384),
                                                    (7, True, 16384), (1, False, 12288), (0, True, 12288),
                                                    (1, False, 10000), (0, True, 10000), (5, True, 2500)])
def test_skipped_bands_hold_no_hit(config, transmit, n_rays):
    # the trace kernel skips a band of 64 banded rays when band_class proves
    # (interval arithmetic on the lens and the acceptance cosine) that the FP64
    # trace loses every ray of each sub-band its rays fall in (ranks
    # 16 u .. 16 u + 15 of every block of the tile, tested once per tile); no
    # skipped band may hold an FP64 hit (ray counts that are not a multiple of
    # 4096 leave idle positions in the last block, whose heights the
    # sub-band bounds include)
    import compound_prism_designer as cpd
    D = transmitting(config, 48)[1] if transmit else cpd.generate_designs(config, 0, 96)[1]
    st = band_stats(D, 16, n_rays)
    print(f"config {config}: {st}, skipped {st['skipped'] / st['past_chain']:.3f} of the bands past the chain test")
    assert st["skipped_with_hit"] == 0, st
    assert st["skipped"] > 0
    # a wavelength the test over every beam height proves lost is not traced
    assert st["dead_with_hit"] == 0, st


@pytest.mark.parametrize("config", [0, 2, 4, 5, 6, 7])
def test_affine_fp64_trace_decides_as_the_reference(config):
    # the queue drains trace a flagged ray with trace64_affine, the affine
    # FP64 chain; every ray it decides lands in trace64's bin (the
    # reference's operation order) with the transmittance within 1e-10
    # relative (the reference's 1 - rs^2 loses digits near grazing exit, the
    # identity 4 r ci cr / (r ci + cr)^2 does not), and it decides almost
    # every ray
    L = guard_lib("fast-honor-pragmas")
    L.affine_check.argtypes = [ctypes.POINTER(N.Design), ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint32,
                               ctypes.c_int, ctypes.POINTER(ctypes.c_uint64), ctypes.c_double * 1]
    _, D, _, _ = transmitting(config, 24, first=1000)
    arr = (N.Design * len(D))(*D)
    st = (ctypes.c_uint64 * 4)()
    terr = (ctypes.c_double * 1)()
    L.affine_check(arr, len(D), 16, 16384, min(16, os.cpu_count() or 1), st, terr)
    s = dict(zip(["rays", "undecided", "mismatch", "hits"], st[:]))
    print(f"config {config}: {s}, max relative t difference {terr[0]:.2e}")
    assert s["mismatch"] == 0, s
    assert s["hits"] > 100_000 and s["undecided"] <= 1e-4 * s["rays"], s
    assert terr[0] <= 1e-10

