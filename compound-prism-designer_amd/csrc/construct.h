// construct.h -- a 2D compound-prism spectrometer from its parameters, in
// FP64, for the host (cpd_build_design, cpd_position_detector_array) and the
// device (construct_kernel: a generation built from parameter vectors).
//
// Face planes from the angles and separation lengths
// (create_joined_trapezoids, src/spectrometer/geometry.rs:430-489), the exit
// lens from the last plane (CurvedPlane::from_hyperplane, geometry.rs:280-315),
// the detector array (LinearDetectorArray::new, spectrometer.rs:183-219) and
// its positioning from the two extreme wavelengths' exit rays
// (detector_array_positioning, spectrometer.rs:322-364, and
// fit_ray_difference_surface, geometry.rs:495-542).
#pragma once

#include <math.h>
#include <stdint.h>

#include "cpd.h"
#include "trace_core.h"

namespace cpd {

struct Vec2 {
    double x, y;
};

// Unit vector at angle theta, rotated by 180 degrees: (-cos, -sin)
// (src/spectrometer/vector.rs:336-371, angled_xy().rot_180_xy()).
CPD_HD Vec2 face_normal(double theta) { return Vec2{-cos(theta), -sin(theta)}; }

CPD_HD bool sign_positive(double v) { return !signbit(v); }

// Parameter checks shared by host and device; the host maps the codes to
// messages.
enum ParamFault : int {
    PARAM_OK = 0,
    PARAM_N_GLASS,
    PARAM_HEIGHT,
    PARAM_CURVATURE,
    PARAM_WAVELENGTHS,
    PARAM_BIN_COUNT,
    PARAM_SLOPE,
};

// The prism and beam part (what positioning needs).
CPD_HD int check_prism_params(const cpd_params &p) {
    if (p.n_glass < 1 || p.n_glass > CPD_MAX_PRISMS) return PARAM_N_GLASS;
    if (!(p.height > 0.0)) return PARAM_HEIGHT;
    if (!(p.curvature != 0.0) || !isfinite(p.curvature)) return PARAM_CURVATURE;
    if (!(p.wavelength_hi >= p.wavelength_lo)) return PARAM_WAVELENGTHS;
    return PARAM_OK;
}

// Everything cpd_build_design needs.
CPD_HD int check_design_params(const cpd_params &p) {
    const int f = check_prism_params(p);
    if (f != PARAM_OK) return f;
    if (p.bin_count < 1 || p.bin_count > CPD_MAX_BINS) return PARAM_BIN_COUNT;
    if (!(p.linear_slope > 0.0)) return PARAM_SLOPE;
    return PARAM_OK;
}

// Face planes of the joined trapezoids (create_joined_trapezoids,
// geometry.rs:430-489): faces 0..P-1 into pt/nm (mid-height points and
// normals), the last face into exit_pt/exit_nm.
CPD_HD void trapezoid_faces(double height, const double *angles, const double *lengths, uint32_t P,
                            double pt[][2], double nm[][2], Vec2 &exit_pt, Vec2 &exit_nm) {
#pragma clang fp contract(off)
    const double h2 = height * 0.5;
    // entrance face: mid-height point at x = |tan(theta0)| h/2
    Vec2 n = face_normal(angles[0]);
    double x0 = fabs(n.y / n.x) * h2;
    pt[0][0] = x0;
    pt[0][1] = h2;
    nm[0][0] = n.x;
    nm[0][1] = n.y;
    bool prev_up = sign_positive(n.y);
    double prev_half = x0;  // half of the face's x-extent over the height
    double x = x0;
    // faces 1..P: face i sits lengths[i-1] beyond the previous one, measured
    // between the facing edges of the two trapezoids
    for (uint32_t i = 1; i <= P; ++i) {
        Vec2 ni = face_normal(angles[i]);
        bool up = sign_positive(ni.y);
        double half = fabs(ni.y / ni.x) * h2;
        double gap = (up != prev_up) ? prev_half + half : fabs(prev_half - half);
        x += lengths[i - 1] + gap;
        prev_up = up;
        prev_half = half;
        if (i < P) {
            pt[i][0] = x;
            pt[i][1] = h2;
            nm[i][0] = ni.x;
            nm[i][1] = ni.y;
        } else {
            exit_pt = Vec2{x, h2};
            exit_nm = ni;
        }
    }
}

// The compound prism, the exit lens and the beam of a design.
CPD_HD void build_prism(const cpd_params &p, cpd_design &d) {
#pragma clang fp contract(off)
    const uint32_t P = p.n_glass;
    d.n_glass = P;
    d.ar_coated = p.ar_coated ? 1u : 0u;
    for (uint32_t i = 0; i < P; ++i)
        for (int j = 0; j < CPD_GLASS_COEFFS; ++j) d.glass[i][j] = p.glass[i][j];
    d.height = p.height;
    Vec2 exit_point{0, 0}, exit_normal{0, 0};
    trapezoid_faces(p.height, p.angles, p.lengths, P, d.plane_point, d.plane_normal, exit_point, exit_normal);
    // exit lens: chord = |h / n.x|, R = chord / (2 |k|), centre on the normal
    // through the chord midpoint at the apothem, bounds centred on the arc's
    // midpoint with radius^2 = sagitta^2 + chord^2 / 4
    const double k = p.curvature;
    double chord = fabs(p.height / exit_normal.x);
    double radius = chord * 0.5 / fabs(k);
    double apothem = sqrt(radius * radius - chord * chord * 0.25);
    double sagitta = radius - apothem;
    double sa = copysign(apothem, k), ss = -copysign(sagitta, k);
    d.arc_center[0] = fma(exit_normal.x, sa, exit_point.x);
    d.arc_center[1] = fma(exit_normal.y, sa, exit_point.y);
    d.arc_radius = radius;
    d.arc_midpt[0] = fma(exit_normal.x, ss, exit_point.x);
    d.arc_midpt[1] = fma(exit_normal.y, ss, exit_point.y);
    d.arc_max_dist_sq = sagitta * sagitta + chord * chord * 0.25;

    d.beam_width = p.beam_width;
    d.y_mean = p.y_mean;
    d.wavelength_lo = p.wavelength_lo;
    d.wavelength_hi = p.wavelength_hi;
}

// The detector array at (pos, flipped), with min_ci = cos(max_incident_angle)
// (pylib.rs:906-920).
CPD_HD void build_detector(const cpd_params &p, const double pos[2], uint32_t flipped, cpd_design &d) {
    d.bin_count = p.bin_count;
    d.det_bin_size = p.bin_size;
    d.det_slope = p.linear_slope;
    d.det_intercept = p.linear_intercept;
    d.det_min_ci = cos(p.max_incident_angle);
    d.det_length = p.det_length;
    Vec2 n = face_normal(p.det_angle);
    d.det_normal[0] = n.x;
    d.det_normal[1] = n.y;
    // direction: normal rotated +90 degrees when flipped, -90 otherwise
    if (flipped) {
        d.det_direction[0] = -n.y;
        d.det_direction[1] = n.x;
    } else {
        d.det_direction[0] = n.y;
        d.det_direction[1] = -n.x;
    }
    d.det_position[0] = pos[0];
    d.det_position[1] = pos[1];
}

// The array of length L at angle a spanning the exit rays lo and hi of the
// extreme wavelengths (fit_ray_difference_surface, geometry.rs:495-542).
CPD_HD bool fit_detector(const Ray64 &lo, const Ray64 &hi, double length, double angle, double pos[2],
                         bool &flipped) {
#pragma clang fp contract(off)
    // array direction (-sin a, cos a) scaled by its length
    double sx = -sin(angle) * length, sy = cos(angle) * length;
    double a = hi.dx, b = -lo.dx, c = hi.dy, dd = -lo.dy;
    double det = a * dd - b * c;
    if (det == 0.0) return false;
    double m10 = -c / det, m11 = a / det;
    // distance along the lower ray to the array end that the upper ray meets
    double tx = sx - hi.ox + lo.ox, ty = sy - hi.oy + lo.oy;
    double s = m10 * tx + m11 * ty;
    if (s > 0.0) {
        pos[0] = fma(lo.dx, s, lo.ox);
        pos[1] = fma(lo.dy, s, lo.oy);
        flipped = false;
        return true;
    }
    tx = -sx - hi.ox + lo.ox;
    ty = -sy - hi.oy + lo.oy;
    s = m10 * tx + m11 * ty;
    if (s < 0.0) return false;
    pos[0] = fma(lo.dx, s, lo.ox);
    pos[1] = fma(lo.dy, s, lo.oy);
    flipped = true;
    return true;
}

// Detector position for parameters that passed check_prism_params: the two
// wavelengths that cut (1 - acceptance) / 2 off each end of the range, traced
// from the beam's mean height (detector_array_positioning,
// spectrometer.rs:322-364).
CPD_HD cpd_status position_detector(const cpd_params &p, double acceptance, double pos[2], uint32_t &flipped_out) {
#pragma clang fp contract(off)
    cpd_design d{};
    build_prism(p, d);
    double half_out = fmax(1.0 - acceptance, 0.0) * 0.5;
    double w_lo = lambda_of(d, half_out);
    double w_hi = lambda_of(d, 1.0 - half_out);
    Ray64 lo{0.0, d.y_mean, 1.0, 0.0, 1.0, 1.0}, hi = lo;
    int e = prism64(lo, d, w_lo);
    if (e != CPD_OK) return (cpd_status)e;
    e = prism64(hi, d, w_hi);
    if (e != CPD_OK) return (cpd_status)e;
    if ((lo.ts + lo.tp) * 0.5 <= 1e-3 || (hi.ts + hi.tp) * 0.5 <= 1e-3) return CPD_ERR_ANGULAR_RESPONSE_TOO_WEAK;
    bool flipped = false;
    if (!fit_detector(lo, hi, p.det_length, p.det_angle, pos, flipped)) return CPD_ERR_NO_SURFACE_INTERSECTION;
    flipped_out = flipped ? 1u : 0u;
    return CPD_OK;
}

// The detector position given relative to the exit face (designer.py:236-243):
// det_position holds (distance, angle); the array sits that distance from the
// final face's midpoint (the arc's midpoint, ray.rs:231-240) along the face
// normal (cos a, sin a) of the last face angle, rotated by the angle.  The
// rotation is numpy's R @ v and the sums are unfused, as in the designer.
CPD_HD void position_from_exit(const cpd_params &p, const cpd_design &prism, double pos[2]) {
#pragma clang fp contract(off)
    const double a = p.angles[p.n_glass];
    const double ex = cos(a), ey = sin(a);
    const double c = cos(p.det_position[1]), s = sin(p.det_position[1]);
    const double dx = c * ex + -s * ey, dy = s * ex + c * ey;
    const double dv = p.det_position[0];
    pos[0] = prism.arc_midpt[0] + dx * dv;
    pos[1] = prism.arc_midpt[1] + dy * dv;
}

// The whole design from its parameters: checks, then the detector position
// (acceptance > 0: positioned for that acceptance; CPD_POSITION_FROM_EXIT:
// from the exit face, not flipped; 0: the parameters' position and flip as
// given), then the records.
CPD_HD cpd_status construct_design(const cpd_params &p, double acceptance, cpd_design &d) {
    if (check_design_params(p) != PARAM_OK) return CPD_ERR_INVALID_ARGUMENT;
    double pos[2] = {p.det_position[0], p.det_position[1]};
    uint32_t flipped = p.flipped ? 1u : 0u;
    if (acceptance > 0.0) {
        const cpd_status st = position_detector(p, acceptance, pos, flipped);
        if (st != CPD_OK) return st;
    }
    d = cpd_design{};
    build_prism(p, d);
    if (acceptance < 0.0) {
        position_from_exit(p, d, pos);
        flipped = 0;
    }
    build_detector(p, pos, flipped, d);
    return CPD_OK;
}

}  // namespace cpd
