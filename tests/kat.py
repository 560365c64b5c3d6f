"""The reference's known design (src/fitness.rs:153-254; benches/benchmark.rs:53-145;
src/cuda_fitness.rs:437-531): N-PK52A / N-SF57 / N-FK58 polynomial glasses,
angles [-27.2712308, 34.16326141, -42.93207009, 1.06311416] deg, lengths 0,
height 2.5, width 2, lens curvature 0.21, beam width 0.2 at y 0.95,
wavelengths [0.5, 0.82] um, 32 bins of 0.1 over 3.2, max incidence 60 deg,
detector angle 0 positioned with acceptance 1."""
import math

GLASSES = [
    [-0.19660238, 0.85166448, -1.49929414, 1.35438084, -0.64424681, 1.62434799],   # N-PK52A
    [-1.81746234, 7.71730927, -13.2402884, 11.56821078, -5.23836004, 2.82403194],  # N-SF57
    [-0.15938247, 0.69081086, -1.21697038, 1.10021121, -0.52409733, 1.55979703],   # N-FK58
]
ANGLES_DEG = [-27.2712308, 34.16326141, -42.93207009, 1.06311416]
# src/fitness.rs:255-257
KAT_SIZE, KAT_SIZE_RMAX = 41.3, 5e-3
KAT_INFO, KAT_INFO_RMAX = 0.7407, 1e-2
KAT_DEVIATION, KAT_DEVIATION_RMAX = 0.377159, 1e-3


def fill_params(p):
    """Fill a cpd_params ctypes struct (either binding) with the known design,
    detector position left at (0, 0)."""
    p.n_glass = 3
    for i, g in enumerate(GLASSES):
        for j, c in enumerate(g):
            p.glass[i][j] = c
    for i, a in enumerate(ANGLES_DEG):
        p.angles[i] = math.radians(a)
    for i in range(3):
        p.lengths[i] = 0.0
    p.curvature = 0.21
    p.height = 2.5
    p.width = 2.0
    p.ar_coated = 0
    p.bin_count = 32
    p.bin_size = 0.1
    p.linear_slope = 0.1
    p.linear_intercept = 0.0
    p.det_length = 3.2
    p.max_incident_angle = math.radians(60)
    p.det_angle = 0.0
    p.beam_width = 0.2
    p.y_mean = 0.95
    p.wavelength_lo = 0.5
    p.wavelength_hi = 0.82
    return p
