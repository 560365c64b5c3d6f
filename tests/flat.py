"""Hand-built flat spectrometers whose per-ray outcomes follow in closed form.

One glass of constant index between a vertical entrance face and a nearly flat
exit lens (curvature 1e-7) on the same axis; the detector stands at x = 30
facing -x and runs upwards from y = -0.4, so a ray entering at height y meets
it at position pos = y + 0.4 (to ~1e-7).  Every face is crossed at (nearly)
normal incidence, so an anti-reflection coated design transmits
t = 0.99 * 0.99 (src/spectrometer/ray.rs:336-340: ci > 0.5 on both faces) for
every ray, and an uncoated one (1 - ((n-1)/(n+1))^2)^2.

The glass does not disperse, so p(d|lambda) is the same row at every
wavelength.  Rays are the beam quantiles of the R1 points
u_k = frac(1/2 + k (sqrt(5) - 1) / 2): y_k = y_mean - width norminv(u_k), so
y_k > y_mean exactly when u_k < 1/2.
"""
import math
from decimal import Decimal, getcontext

import kat

T_AR = 0.99 * 0.99  # the two faces' 0.99 factors, multiplied in this order


def flat_params(p, n=1.5, ar=True, exit_angle_deg=0.0, curvature=1e-7):
    """Fill a cpd_params struct (either binding) with the flat design; the
    beam is centred on the lens axis (y_mean = h / 2 = 1.25) with width 0.1."""
    kat.fill_params(p)
    p.n_glass = 1
    for j in range(6):
        p.glass[0][j] = 0.0
    p.glass[0][5] = n
    p.angles[0] = 0.0
    p.angles[1] = math.radians(exit_angle_deg)
    p.lengths[0] = 1.0
    p.curvature = curvature
    p.ar_coated = int(ar)
    p.det_angle = 0.0
    p.det_position[0], p.det_position[1] = 30.0, -0.4
    p.flipped = 0
    p.max_incident_angle = math.radians(60)
    p.y_mean = 1.25
    p.beam_width = 0.1
    return p


def single_bin(p):
    """Narrow beam (width 1e-4, so |y - 1.25| < 5e-4): every ray lands in bin 16
    of the default 32 bins of 0.1 (pos = 1.65 +- 5e-4)."""
    p.beam_width = 1e-4
    return p


def split_bins(p):
    """Two bins of width 1 that meet at pos = 1.65, the position of the beam's
    mean height: ray k lands in bin 1 when u_k < 1/2 (y_k > y_mean), else in
    bin 0.  |u_k - 1/2| > 1/(3 k), so |y_k - y_mean| > 1e-6 for k <= 2^16; the
    lens is centred on the beam axis and moves a ray by ~1e-7 of its distance
    from the axis, and positions round at ~1e-15."""
    p.bin_count = 2
    p.bin_size = 1.0
    p.linear_slope = 1.0
    p.linear_intercept = 0.65
    return p


_ALPHA = None


def r1_below_half(k: int) -> bool:
    """u_k < 1/2 for u_k = frac(1/2 + k alpha), in 60-digit decimal arithmetic."""
    global _ALPHA
    getcontext().prec = 60
    if _ALPHA is None:
        _ALPHA = (Decimal(5).sqrt() - 1) / 2
    u = (Decimal("0.5") + k * _ALPHA) % 1
    return u < Decimal("0.5")
