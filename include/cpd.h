/*
 * cpd.h -- C ABI of the MI355X compound-prism fitness evaluator.
 *
 * The evaluator traces quasi-Monte Carlo rays through a 2D compound prism
 * (P planar faces and one circular-arc exit lens) onto a linear detector
 * array, and reduces the per-bin hit probabilities p(d|lambda) to the mutual
 * information I(Lambda; D), next to the design's size and deviation.  The
 * cpd3_* entry points do the same for the 3D variant (toric lenses, fiber
 * beam).
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns a
 * cpd_status and never lets a C++ exception cross the boundary.  Buffers are
 * owned by the caller.  Host-only entry points (construction, positioning,
 * glass fit, workload generation) never touch the GPU and can be called on a
 * machine without one.
 *
 * Each declaration names the reference interface it replaces
 * (jdtatz/compound-prism-designer, path:line).
 */
#ifndef CPD_H
#define CPD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CPD_ABI_VERSION 1
#define CPD_MAX_PRISMS 6      /* src/pylib.rs:458 MAX_GLASS */
#define CPD_GLASS_COEFFS 6    /* Glass<F, 6>, src/spectrometer/glasscat.rs:7-22 */
#define CPD_MAX_BINS 1024

/* Status codes.  1..4 are the reference's RayTraceError variants
 * (src/spectrometer/ray.rs:11-16), in the same order. */
typedef enum cpd_status {
    CPD_OK = 0,
    CPD_ERR_NO_SURFACE_INTERSECTION = 1,
    CPD_ERR_OUT_OF_BOUNDS = 2,
    CPD_ERR_TOTAL_INTERNAL_REFLECTION = 3,
    CPD_ERR_ANGULAR_RESPONSE_TOO_WEAK = 4,
    CPD_ERR_INVALID_ARGUMENT = 16,
    CPD_ERR_NOT_CONVERGED = 17,   /* cuda_fitness returned None, src/cuda_fitness.rs:403-422 */
    CPD_ERR_NOT_READY = 18,       /* server request still in flight */
    CPD_ERR_NO_DEVICE = 32,
    CPD_ERR_DEVICE = 33,          /* CudaError, src/pylib.rs:26-35 */
    CPD_ERR_OUT_OF_MEMORY = 34,
    CPD_ERR_CAPACITY = 35         /* request larger than a server's admission budget */
} cpd_status;

/* Parameters of one spectrometer as the reference's Python constructors take
 * them: CompoundPrism(glasses, angles, lengths, curvature, height, width,
 * ar_coated) src/pylib.rs:689-754, DetectorArray(...) src/pylib.rs:867-904,
 * GaussianBeam(width, y_mean) src/pylib.rs:363-372,
 * UniformWavelengthDistribution(bounds) src/pylib.rs:304-316.
 * Lengths are in the caller's unit (cm in the benchmarks), wavelengths in um,
 * angles in radians. */
typedef struct cpd_params {
    uint32_t n_glass;                                   /* P, 1..CPD_MAX_PRISMS */
    uint32_t ar_coated;                                 /* bool */
    double glass[CPD_MAX_PRISMS][CPD_GLASS_COEFFS];     /* n(l) coefficients, Horner order */
    double angles[CPD_MAX_PRISMS + 1];                  /* P + 1 face angles */
    double lengths[CPD_MAX_PRISMS];                     /* P separation lengths */
    double curvature;                                   /* signed normalised lens curvature */
    double height;
    double width;
    /* detector array */
    uint32_t bin_count;
    uint32_t flipped;                                   /* bool */
    double bin_size;
    double linear_slope;
    double linear_intercept;
    double det_length;
    double max_incident_angle;
    double det_angle;
    double det_position[2];
    /* gaussian beam */
    double beam_width;
    double y_mean;
    /* uniform wavelength distribution */
    double wavelength_lo;
    double wavelength_hi;
} cpd_params;

/* A constructed spectrometer (FP64): the resolved geometry the ray trace reads.
 * Face 0 is the entrance face, faces 1..P-1 the interior faces; the exit face
 * is the arc.  Mirrors Spectrometer<f64, ...> src/spectrometer/spectrometer.rs:387-394. */
typedef struct cpd_design {
    uint32_t n_glass;
    uint32_t bin_count;
    uint32_t ar_coated;
    uint32_t reserved;
    double glass[CPD_MAX_PRISMS][CPD_GLASS_COEFFS];
    double plane_point[CPD_MAX_PRISMS][2];
    double plane_normal[CPD_MAX_PRISMS][2];
    double height;                                      /* PrismBounds height of every plane */
    double arc_center[2];
    double arc_radius;
    double arc_midpt[2];                                /* RadialBounds centre */
    double arc_max_dist_sq;                             /* RadialBounds radius_squared */
    double det_position[2];
    double det_direction[2];
    double det_normal[2];
    double det_min_ci;
    double det_length;
    double det_slope;
    double det_intercept;
    double det_bin_size;
    double beam_width;
    double y_mean;
    double wavelength_lo;
    double wavelength_hi;
} cpd_design;

/* Sampling of one fitness evaluation.
 * CPD_MODE_FIXED traces every wavelength sample j = 1..n_wavelengths and every
 * ray k = 1..n_rays (the reference's loops with both SEM thresholds at 0).
 * CPD_MODE_ADAPTIVE keeps the reference's SEM <= 5e-3 early exits
 * (src/fitness.rs:20-28, 108-114). */
enum { CPD_MODE_FIXED = 0, CPD_MODE_ADAPTIVE = 1 };

typedef struct cpd_sampling {
    uint32_t n_wavelengths;   /* max_n */
    uint32_t n_rays;          /* max_m */
    uint32_t mode;
    uint32_t reserved;
} cpd_sampling;

/* DesignFitness, src/fitness.rs:36-41. */
typedef struct cpd_fitness {
    double size;
    double info;
    double deviation;
} cpd_fitness;

typedef struct cpd_ctx cpd_ctx;
typedef struct cpd_server cpd_server;

/* ------------------------------------------------------------------------ */
/* Library                                                                   */
/* ------------------------------------------------------------------------ */

int cpd_abi_version(void);
/* Name of a status code ("NoSurfaceIntersection", ...): RayTraceError's
 * Into<&'static str>, src/spectrometer/ray.rs:18-29. */
const char *cpd_status_name(cpd_status status);
/* Message of the last failing call on this thread ("" if none). */
const char *cpd_last_error(void);

/* ------------------------------------------------------------------------ */
/* Host-side construction (FP64, no GPU)                                     */
/* ------------------------------------------------------------------------ */

/* detector_array_positioning, src/spectrometer/spectrometer.rs:322-364, as
 * exposed by position_detector_array, src/pylib.rs:814-838.  Reads the prism,
 * beam and wavelength fields of params plus det_length and det_angle. */
cpd_status cpd_position_detector_array(const cpd_params *params, double acceptance,
                                       double position_out[2], uint32_t *flipped_out);

/* CompoundPrism::new src/spectrometer/ray.rs:137-183 (via
 * create_joined_trapezoids src/spectrometer/geometry.rs:430-489 and the
 * Plane/CurvedPlane from_hyperplane constructors geometry.rs:225-243,280-315),
 * LinearDetectorArray::new src/spectrometer/spectrometer.rs:183-219 and
 * Spectrometer src/pylib.rs:985-997. */
cpd_status cpd_build_design(const cpd_params *params, cpd_design *out);

/* Spectrometer::size_and_deviation, src/spectrometer/spectrometer.rs:429-437. */
cpd_status cpd_size_and_deviation(const cpd_design *design, double *size, double *deviation);

/* The ray leaving the compound prism (origin x, y; direction x, y) for a
 * beam ray starting at (0, y) with direction (1, 0): CompoundPrism.exit_ray,
 * src/pylib.rs:563-573, 768-773 via GenericCompoundPrism::propagate,
 * src/spectrometer/ray.rs:485-499. */
cpd_status cpd_exit_ray(const cpd_design *design, double y, double wavelength, double out[4]);

/* Propagation path of one ray (Spectrometer.ray_trace, src/pylib.rs:1060-1083;
 * propagation_path, src/spectrometer/spectrometer.rs:439-464): rows of
 * (origin x, origin y, direction x, direction y) -- the start, each surface
 * crossed, and the detector hit.  out holds up to CPD_MAX_PRISMS + 3 rows;
 * *n_rows receives the count. */
cpd_status cpd_ray_trace(const cpd_design *design, double wavelength, double y, double *out, uint32_t *n_rows);

/* Sellmeier1 n(l) fitted by a degree-5 Horner polynomial over npts equally
 * spaced wavelengths in [lo, hi]: AbstractGlass.horner_polynomial_fit /
 * into_glass, compound_prism_designer/glasscat.py:114-127, 144-152.
 * b and c are (B1,B2,B3) and (C1,C2,C3).  chisq_out may be NULL. */
cpd_status cpd_fit_sellmeier1(const double b[3], const double c[3], double lo, double hi,
                              uint32_t npts, double coeffs_out[CPD_GLASS_COEFFS],
                              double *chisq_out);

/* The same fit for n glasses on the GPU, one glass per thread (a catalogue
 * fitted in one call, as designer.py's glass_catalog is before an
 * optimisation): b and c hold n triples, coeffs_out n records of 6 and
 * chisq_out (may be NULL) n residual sums of squares; per_glass[i] is CPD_OK
 * or CPD_ERR_INVALID_ARGUMENT (a pole in range or dependent columns, NaN
 * results).  Least squares by Givens rotations row by row (no m x 6 matrix
 * per thread): the same optimum as cpd_fit_sellmeier1's Householder
 * factorisation, to rounding.  AbstractGlass.horner_polynomial_fit /
 * into_glass, compound_prism_designer/glasscat.py:114-127. */
cpd_status cpd_fit_sellmeier1_batch(cpd_ctx *ctx, const double *b, const double *c, size_t n, double lo, double hi,
                                    uint32_t npts, double *coeffs_out, double *chisq_out, cpd_status *per_glass);

/* Synthetic workloads of the benchmark configurations (no reference
 * counterpart; the reference's optimiser draws designs in designer.py:203-265).
 * config 0/1: 3 prisms, 32 bins, 64 wavelengths x 16384 rays;
 * config 2:   1..6 prisms, 16..128 bins, 32..256 wavelengths x 4096..65536 rays;
 * config 4:   6 prisms, 256 bins, 1024 wavelengths x 2^20 rays;
 * the detector and coating settings of the reference's shipped designer
 * configurations, geometry drawn from designer.py:122-130's ranges (mm):
 * config 5:   spring.toml / design_config.toml: 4 prisms, anti-reflection
 *             coated, 32 bins of 0.8 at slope 1 from intercept 0.1 over 32,
 *             beam width 3.2, max height 25;
 * config 6:   bch.toml: 3 prisms, coated, 64 bins of 0.42 over 26.6 (the last
 *             bin overhangs the array), wavelengths 0.48..1.0, beam width 2;
 * config 7:   config 0 with gaps (bin size 0.5..1 of the slope), an intercept
 *             of 0..3 slopes and a random coating;
 * configs 5..7 sample 64 wavelengths x 16384 rays and keep only designs that
 * land at least 10% of a probe's rays in a bin (cpd_hit_fraction over 4
 * wavelengths x 256 rays).
 * Design `index` of a configuration is reproducible on its own.  Draws that
 * fail construction, positioning or the probe are redrawn; *attempts_out (may
 * be NULL) receives the number of draws used. */
cpd_status cpd_generate_design(uint32_t config, uint64_t index, cpd_params *params_out,
                               cpd_sampling *sampling_out, uint32_t *attempts_out);

/* Fraction of the rays that land in a bin, over rays k = 1..n_rays of the
 * beam at wavelength samples j = 1..n_lambda (the sequences fitness() uses,
 * src/fitness.rs:19,94), traced in FP64 on the host.  A workload statistic
 * (no reference counterpart). */
cpd_status cpd_hit_fraction(const cpd_design *design, uint32_t n_lambda, uint32_t n_rays, double *out);

/* Designs first_index .. first_index + n - 1 of a configuration, generated
 * on `threads` host threads (0: all); arrays of n.  attempts_out may be NULL. */
cpd_status cpd_generate_designs(uint32_t config, uint64_t first_index, size_t n, uint32_t threads,
                                cpd_params *params_out, cpd_design *designs_out, cpd_sampling *sampling_out,
                                uint32_t *attempts_out);

/* ------------------------------------------------------------------------ */
/* Device context                                                            */
/* ------------------------------------------------------------------------ */

/* A context owns a HIP stream and its device buffers on one GPU.  Calls on
 * one context are serialised by the context; different contexts run
 * independently.  The reference keeps one global context behind a mutex,
 * src/cuda_fitness.rs:322-363; passing ctx = NULL to the entry points below
 * uses the same kind of process-wide default context on device 0. */
cpd_status cpd_ctx_create(int device, cpd_ctx **ctx_out);
void cpd_ctx_destroy(cpd_ctx *ctx);
/* Name of the context's GPU, its compute-unit count and clock (kHz). */
cpd_status cpd_ctx_info(cpd_ctx *ctx, char *name, size_t name_len, int *cu_count, int *clock_khz);

/* Context options (no reference counterpart: measurement and scheduling
 * knobs of this implementation). */
enum {
    CPD_OPT_PROFILE = 1,       /* record HIP events around each kernel of a run (0/1) */
    CPD_OPT_COUNT_OPS = 2,     /* count the algorithmic operations a run performs (0/1) */
    CPD_OPT_TILE_RAYS = 3,     /* rays per work item, multiple of 64, <= 2^20 (0: 65536) */
    CPD_OPT_WAVES_PER_CU = 4,  /* resident wavefronts per CU of the persistent grid */
    CPD_OPT_ORDER = 5,         /* work order of a batch, CPD_ORDER_* */
    CPD_OPT_ACCUMULATE = 6,    /* per-bin accumulation of fixed-count batches, CPD_ACC_* */
    CPD_OPT_ROW_TRACE = 7      /* ray trace of the adaptive estimators: 64 = FP64 in the reference's
                                  operation order for every ray; 32 = the FP32 trace with its FP64
                                  re-trace of flagged rays; 0 = 64 for CPD_MODE_ADAPTIVE rows and
                                  32 for cpd_gpu_fitness (the reference's GPU estimator is FP32) */
};
/* Accumulation of the per-bin transmittance (a precision study; results of
 * the default are exact and order-independent):
 * CPD_ACC_EXACT  round(t 2^22) in integers, FP64 compensated entropies (default);
 * CPD_ACC_F32    FP32 sums and atomics, FP32 entropies summed plainly;
 * CPD_ACC_F32C   FP32 with Neumaier compensation of partial sums and entropies
 *                (one work item per wavelength, so n_rays <= 2^20);
 * CPD_ACC_F64    FP64 sums and atomics, FP64 compensated entropies. */
enum { CPD_ACC_EXACT = 0, CPD_ACC_F32 = 1, CPD_ACC_F32C = 2, CPD_ACC_F64 = 3 };
cpd_status cpd_ctx_set_option(cpd_ctx *ctx, int option, int64_t value);

/* Statistics of the last fitness run on a context. */
typedef struct cpd_run_stats {
    double trace_ms;      /* trace kernel, HIP events on the context's stream (CPD_OPT_PROFILE) */
    double finalize_ms;   /* entropy reduction kernel (CPD_OPT_PROFILE) */
    uint64_t ops;         /* algorithmic operations (CPD_OPT_COUNT_OPS) */
    uint64_t items;       /* work items (design, wavelength, ray tile) */
    uint64_t rays;        /* rays traced */
    uint32_t grid;        /* workgroups launched */
    uint32_t tile_rays;
} cpd_run_stats;
cpd_status cpd_ctx_last_stats(cpd_ctx *ctx, cpd_run_stats *out);

/* ------------------------------------------------------------------------ */
/* Fitness on the GPU                                                        */
/* ------------------------------------------------------------------------ */

/* Plain one-design launch: the design is copied to the device, traced by one
 * wavefront per wavelength sample, reduced to DesignFitness on the device and
 * copied back before return.  With CPD_MODE_ADAPTIVE the estimator keeps the
 * reference's early exits: rows of wavelength samples are traced in FP64 on
 * the device and folded on the host, exactly as fitness() folds them.
 * Replaces fitness(), src/fitness.rs:130-147, with the per-design launch
 * structure of src/cuda_fitness.rs:199-245. */
cpd_status cpd_fitness_gpu(cpd_ctx *ctx, const cpd_design *design, const cpd_sampling *sampling,
                           cpd_fitness *out);

/* Batched evaluation of n designs with per-design sampling in one persistent
 * launch.  per_design (may be NULL) receives each design's status.  Replaces
 * n calls of fitness()/cuda_fitness(), src/fitness.rs:130-147,
 * src/cuda_fitness.rs:347-423, as driven by designer.py:267-300. */
cpd_status cpd_fitness_batch(cpd_ctx *ctx, const cpd_design *designs, const cpd_sampling *sampling,
                             size_t n, cpd_fitness *out, cpd_status *per_design);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, code that must not call hipMalloc, harnesses
 * that forbid it): cpd_fitness_batch for fixed-count sampling, with every
 * device buffer it uses (beam tables, staged records, per-bin accumulators,
 * last-ray bins, results, work counter) inside one static device array of
 * CPD_NOALLOC_SCRATCH_BYTES (22 MiB) that each GPU holds from library load
 * until the process ends; the call allocates nothing (no hipMalloc, hipFree,
 * pinned or heap buffers).  Results are cpd_fitness_batch's bit for bit.
 * Limits: n <= CPD_NOALLOC_MAX_DESIGNS; per design CPD_MODE_FIXED and
 * n_rays <= CPD_NOALLOC_MAX_RAYS (otherwise that design's status is
 * CPD_ERR_INVALID_ARGUMENT); over the valid designs sum n_wavelengths <=
 * CPD_NOALLOC_MAX_ROWS and sum n_wavelengths * bin_count <=
 * CPD_NOALLOC_MAX_CELLS (otherwise the call fails with
 * CPD_ERR_INVALID_ARGUMENT).  The first call on a device fills the array's beam
 * tables (a synchronous upload).  Calls in one process are serialised; each
 * process has its own copy of the array. */
#define CPD_NOALLOC_MAX_DESIGNS 4096
#define CPD_NOALLOC_MAX_RAYS 65536
#define CPD_NOALLOC_MAX_ROWS (1u << 18)
#define CPD_NOALLOC_MAX_CELLS (1u << 21)
#define CPD_NOALLOC_SCRATCH_BYTES 23397632
cpd_status cpd_fitness_batch_noalloc(cpd_ctx *ctx, const cpd_design *designs, const cpd_sampling *sampling,
                                     size_t n, cpd_fitness *out, cpd_status *per_design);

/* Per-bin transmission probabilities p(d|l) for the given wavelengths with
 * max_m rays each; out is row-major [n_wavelengths][bin_count].  With
 * CPD_MODE_ADAPTIVE a row stops at the reference's SEM early exit, as
 * Spectrometer.transmission_probability does.  Replaces
 * p_dets_l_wavelength, src/fitness.rs:48-69, as exposed by
 * Spectrometer.transmission_probability, src/pylib.rs:1028-1058. */
cpd_status cpd_p_dets_l_wavelength(cpd_ctx *ctx, const cpd_design *design, const double *wavelengths,
                                   size_t n_wavelengths, uint32_t max_m, uint32_t mode, double *out);

/* The reference's GPU estimator: per seed, max_n / nwarp * nwarp wavelength
 * rows with u = frac(seed + id / phi), rays u = frac(seed + k / phi) for
 * k = lane + 32 (i + 1), 32-ray chunks folded per distinct bin with the SEM
 * <= 5e-3 early exit per row; seeds are added until every bin's mean over
 * the rows has an SEM <= 5e-3.  Returns CPD_ERR_NOT_CONVERGED (out holds the
 * estimate after the last seed) where the reference returns None.  Rays are
 * traced by the FP32 trace with the exact FP64 re-trace of flagged rays (so
 * bins are the FP64 decisions; CPD_OPT_ROW_TRACE 64 traces every ray in
 * FP64) and folded in FP64 (the reference traces and folds in FP32 with
 * fast-math).  Replaces cuda_fitness, src/cuda_fitness.rs:347-423, with the
 * kernel src/spectrometer/kernel.rs:89-193, as exposed by
 * Spectrometer.gpu_fitness, src/pylib.rs:1087-1135. */
cpd_status cpd_gpu_fitness(cpd_ctx *ctx, const cpd_design *design, const double *seeds, size_t n_seeds,
                           uint32_t max_n, uint32_t nwarp, uint32_t max_eval, cpd_fitness *out);
/* The same for n designs in shared launches; per_design (may be NULL)
 * receives CPD_OK, CPD_ERR_NOT_CONVERGED or a validation error. */
cpd_status cpd_gpu_fitness_batch(cpd_ctx *ctx, const cpd_design *designs, size_t n, const double *seeds,
                                 size_t n_seeds, uint32_t max_n, uint32_t nwarp, uint32_t max_eval,
                                 cpd_fitness *out, cpd_status *per_design);

/* Acceptance value of cpd_build_designs and cpd_plan_run_params that places
 * the detector relative to the exit face, as the designer does without
 * automatic positioning (designer.py:236-243): params' det_position holds
 * (distance, angle), and the array's position is the final face's midpoint
 * (the arc midpoint, src/spectrometer/ray.rs:231-240) plus distance times the
 * face normal (cos, sin of the last face angle) rotated by angle; not
 * flipped. */
#define CPD_POSITION_FROM_EXIT (-1.0)

/* n designs built on the GPU from their parameters: cpd_build_design, and
 * first, when acceptance > 0, cpd_position_detector_array with that
 * acceptance (params' det_position and flipped are then not read; with
 * acceptance 0 they are used as given; with CPD_POSITION_FROM_EXIT
 * det_position is read as (distance, angle) from the exit face).  The same FP64 construction code as the host entry points; only
 * the device's cos and sin may differ from the host's in the last bit.
 * per_design[i] receives design i's status (the host entry points' codes,
 * without their messages); designs_out[i] is meaningful where it is CPD_OK.
 * acceptance in [0, 1] or CPD_POSITION_FROM_EXIT.  designer.py:203-265 per
 * individual. */
cpd_status cpd_build_designs(cpd_ctx *ctx, const cpd_params *params, size_t n, double acceptance,
                             cpd_design *designs_out, cpd_status *per_design);

/* One ray per (wavelength cdf, ray cdf) pair through the FP32 device trace:
 * bin index (bin_count on a miss) and transmittance (0 on a miss).  Replaces
 * propagation_test_kernel, src/spectrometer/kernel.rs:195-224, launched by
 * src/cuda_fitness.rs:247-317. */
cpd_status cpd_propagate_rays(cpd_ctx *ctx, const cpd_design *design, const float *wavelength_cdf,
                              const float *ray_cdf, size_t n, uint32_t *bin_out, float *t_out);

/* ------------------------------------------------------------------------ */
/* Schedules, autotuning and captured generations                            */
/* ------------------------------------------------------------------------ */

/* No reference counterpart.  A schedule fixes how a batch is cut into work
 * items and fed to the persistent grid.  Results never depend on it: bin sums
 * are exact integers, so any schedule gives bit-identical fitness values. */
enum {
    CPD_ORDER_GIVEN = 0,   /* designs run in the order given */
    CPD_ORDER_COST = 1     /* costliest work items first (longest-processing-time order) */
};
typedef struct cpd_schedule {
    uint32_t tile_rays;     /* rays per work item cap, multiple of 64, <= 2^20 (0: 65536) */
    uint32_t waves_per_cu;  /* resident wavefronts per CU of the persistent grid, 1..64 */
    uint32_t order;         /* CPD_ORDER_* */
    uint32_t reserved;
} cpd_schedule;

/* Times candidate schedules on this batch (each evaluated in full, results
 * discarded) and returns the fastest in *best and its trace + finalize time
 * in *best_ms (may be NULL).  The context's own schedule is unchanged. */
cpd_status cpd_autotune(cpd_ctx *ctx, const cpd_design *designs, const cpd_sampling *sampling, size_t n,
                        cpd_schedule *best, double *best_ms);
/* Schedule of cpd_fitness_batch on this context (NULL: the default). */
cpd_status cpd_ctx_set_schedule(cpd_ctx *ctx, const cpd_schedule *schedule);

/* A plan is a generation of n designs whose shape is fixed: per slot the
 * sampling and the bin count.  cpd_plan_create stages the shape on the device
 * and captures upload, trace, reduction and download of one generation in a
 * HIP graph; cpd_plan_run copies the generation's design records into pinned
 * memory, replays the graph and waits.  The FP32 trace records are built on
 * the device.  A design whose bin count differs from its slot's, or that fails
 * validation, gets CPD_ERR_INVALID_ARGUMENT in per_design (may be NULL) and
 * NaN results; the others are evaluated.  Replaces the per-generation loop of
 * designer.py:267-300 over fitness(), src/fitness.rs:130-147. */
typedef struct cpd_plan cpd_plan;
cpd_status cpd_plan_create(cpd_ctx *ctx, const cpd_design *shape, const cpd_sampling *sampling, size_t n,
                           const cpd_schedule *schedule /* NULL: autotuned on `shape` */, cpd_plan **out);
cpd_status cpd_plan_run(cpd_plan *plan, const cpd_design *designs, cpd_fitness *out, cpd_status *per_design);
/* One generation from parameter vectors: the parameters are uploaded and the
 * design records built on the device (as cpd_build_designs, with the same
 * acceptance rule) inside a second captured graph, then evaluated as
 * cpd_plan_run does, so the host does no per-design work.  This is
 * designer.py:203-265 (create_spectrometer: CompoundPrism, then
 * position_detector_array or the given position, then DetectorArray) and
 * :267-300 (fitness per individual) for a whole population.  per_design[i]
 * (may be NULL) is the construction status, or CPD_ERR_INVALID_ARGUMENT for a
 * bin count other than the slot's; such designs get NaN results. */
cpd_status cpd_plan_run_params(cpd_plan *plan, const cpd_params *params, double acceptance, cpd_fitness *out,
                               cpd_status *per_design);
/* Schedule the plan was captured with. */
cpd_status cpd_plan_schedule(const cpd_plan *plan, cpd_schedule *out);
void cpd_plan_destroy(cpd_plan *plan);

/* ------------------------------------------------------------------------ */
/* Serving: continuous batching of fitness requests                          */
/* ------------------------------------------------------------------------ */

/* No reference counterpart: the reference evaluates one individual per call
 * (designer.py:267-300).  A server owns a HIP stream and device pools on its
 * context's GPU and a host thread.  Submitted requests (batches of designs)
 * are admitted first come, first served while their records and accumulators
 * fit the memory budget.  A request completes when its last work item has
 * run; its results are the same as from cpd_fitness_batch, bit for bit, in
 * either mode.
 *
 * CPD_SERVER_STEPS: every scheduling step is one persistent launch over a
 * share of the remaining work of each admitted request, so requests that
 * arrive while others run join the next step and their work overlaps on the
 * device.
 *
 * CPD_SERVER_RESIDENT: the host stages each admitted request in mapped host
 * memory and publishes it in chunks of up to 64 designs; a persistent kernel
 * stays on the GPU while there is work: its wavefronts copy published chunks
 * into the device pools, claim work items from the oldest chunk first, and
 * the wavefront that finishes a design's last item reduces it and writes the
 * result back to host memory, with no launch or copy per request.  The loop
 * exits after a few milliseconds without work and is launched again when work
 * arrives.  Designs must have at most max_bins bins and max_rays rays per
 * wavelength (fixed-count sampling); submit fails with CPD_ERR_CAPACITY
 * otherwise. */
enum {
    CPD_SERVER_STEPS = 0,      /* a host thread launches one persistent trace per scheduling step */
    CPD_SERVER_RESIDENT = 1    /* a device-resident loop admits, traces and finalizes requests */
};
typedef struct cpd_server_opts {
    uint64_t memory_budget;    /* device bytes for admitted requests (0: 4 GiB) */
    uint32_t step_items;       /* work items per step, a cap (0: 64 per resident wavefront; steps mode) */
    uint32_t max_designs;      /* design slots (0: 65536) */
    uint32_t mode;             /* CPD_SERVER_* */
    uint32_t max_bins;         /* resident mode: largest bin count of a design (0: 256) */
    uint32_t max_rays;         /* resident mode: largest rays per wavelength (0: 65536) */
    uint32_t reserved;
} cpd_server_opts;

typedef struct cpd_server_stats {
    uint64_t submitted, completed, rejected;
    uint64_t steps, items, designs;   /* steps: scheduling steps, or launches of the resident loop */
    uint32_t queued, admitted;
    double busy_ms;            /* time with work on the device */
} cpd_server_stats;

cpd_status cpd_server_create(cpd_ctx *ctx, const cpd_server_opts *opts, cpd_server **out);
/* Queue n designs; *request_id identifies the request.  Fails with
 * CPD_ERR_CAPACITY when the request can never fit the budget. */
cpd_status cpd_server_submit(cpd_server *srv, const cpd_design *designs, const cpd_sampling *sampling, size_t n,
                             uint64_t *request_id);
/* CPD_ERR_NOT_READY while the request runs; otherwise fills out[n] and
 * per_design[n] (may be NULL) and forgets the request.  *latency_ms (may be
 * NULL) receives the time from submission to completion. */
cpd_status cpd_server_poll(cpd_server *srv, uint64_t request_id, cpd_fitness *out, cpd_status *per_design,
                           double *latency_ms);
/* As poll, waiting up to timeout_ms (< 0: no limit). */
cpd_status cpd_server_wait(cpd_server *srv, uint64_t request_id, int64_t timeout_ms, cpd_fitness *out,
                           cpd_status *per_design, double *latency_ms);
cpd_status cpd_server_stats_get(cpd_server *srv, cpd_server_stats *out);
/* Stops the scheduling thread after the step in flight; queued requests are dropped. */
void cpd_server_destroy(cpd_server *srv);

/* ------------------------------------------------------------------------ */
/* 3D spectrometer: toric entrance and exit lenses, fiber or Gaussian beam   */
/* ------------------------------------------------------------------------ */

/* The reference's Dim3 variant (DimensionedSizedCompoundPrism::Dim3,
 * src/pylib.rs:636-640): CulminatingToricCompoundPrism -- a ToricLens
 * entrance face, planar interior faces bounded in height and width, a
 * ToricLens exit face (src/spectrometer/ray.rs:121-127,
 * src/spectrometer/toric.rs) -- traced in 3D from a FiberBeam
 * (spectrometer.rs:70-149) or a 3D GaussianBeam (spectrometer.rs:32-46) onto
 * the same linear detector array.  Geometry lies in the xy plane (z = 0);
 * rays leave the plane. */
enum { CPD_BEAM_GAUSSIAN = 0, CPD_BEAM_FIBER = 1 };

/* CompoundPrism(glasses, angles, lengths, ((ip, it), (fp, ft)), height,
 * width, ar_coated) src/pylib.rs:715-744, DetectorArray as in cpd_params,
 * FiberBeam(core_radius, numerical_aperture, center_y) src/pylib.rs:396-409
 * or GaussianBeam(width, y_mean), UniformWavelengthDistribution(bounds). */
typedef struct cpd_params3 {
    uint32_t n_glass;
    uint32_t ar_coated;
    double glass[CPD_MAX_PRISMS][CPD_GLASS_COEFFS];
    double angles[CPD_MAX_PRISMS + 1];
    double lengths[CPD_MAX_PRISMS];
    double initial_curvature[2];   /* (signed normalised poloidal, normalised toroidal) */
    double final_curvature[2];
    double height;
    double width;
    uint32_t bin_count;
    uint32_t flipped;
    double bin_size;
    double linear_slope;
    double linear_intercept;
    double det_length;
    double max_incident_angle;
    double det_angle;
    double det_position[2];
    uint32_t beam_kind;            /* CPD_BEAM_* */
    uint32_t reserved;
    double beam_width;             /* Gaussian 1/e^2 width, or fiber core radius */
    double y_mean;                 /* Gaussian y_mean, or fiber center_y */
    double numerical_aperture;     /* fiber only */
    double wavelength_lo;
    double wavelength_hi;
} cpd_params3;

/* ToricLens<V, 3> (toric.rs:17-24, 240-291): the torus (centre, toroidal
 * axis, radii) and its SandwichBounds (geometry.rs:94-110).  Vectors are xy;
 * their z is 0. */
typedef struct cpd_torus {
    double center[2];
    double toroidal_normal[2];
    double toroidal_radius;
    double poloidal_radius;
    double bound_center[2];
    double bound_normal[2];
    double bound_height;
} cpd_torus;

/* A constructed 3D spectrometer (FP64).  lens[0] is the entrance face,
 * lens[1] the exit face; plane i (1..P-1) the interior face i, bounded by
 * 0 <= y <= height and |z| <= half_width (PrismBounds<3>,
 * geometry.rs:69-77). */
typedef struct cpd_design3 {
    uint32_t n_glass;
    uint32_t bin_count;
    uint32_t ar_coated;
    uint32_t beam_kind;
    double glass[CPD_MAX_PRISMS][CPD_GLASS_COEFFS];
    cpd_torus lens[2];
    double plane_point[CPD_MAX_PRISMS][2];
    double plane_normal[CPD_MAX_PRISMS][2];
    double height;
    double half_width;
    double det_position[2];
    double det_direction[2];
    double det_normal[2];
    double det_min_ci;
    double det_length;
    double det_slope;
    double det_intercept;
    double det_bin_size;
    double beam_width;             /* Gaussian width, or fiber radius */
    double y_mean;                 /* Gaussian y_mean, or fiber center_y */
    double fiber_min_cos;          /* sqrt(1 - NA^2), spectrometer.rs:85-93 */
    double wavelength_lo;
    double wavelength_hi;
} cpd_design3;

/* Host-side (no GPU).  CompoundPrism::new for the Dim3 variant
 * (src/pylib.rs:715-744, ray.rs:137-183 with ToricLens::from_hyperplane
 * toric.rs:244-291), LinearDetectorArray::new and FiberBeam::new. */
cpd_status cpd3_build_design(const cpd_params3 *params, cpd_design3 *out);
/* detector_array_positioning for a Dim3 prism (spectrometer.rs:322-364 via
 * position_detector_array, src/pylib.rs:814-838): the median ray starts at
 * (0, median_y, 0) along x. */
cpd_status cpd3_position_detector_array(const cpd_params3 *params, double acceptance, double position_out[2],
                                        uint32_t *flipped_out);
/* Spectrometer::size_and_deviation, spectrometer.rs:429-437. */
cpd_status cpd3_size_and_deviation(const cpd_design3 *design, double *size, double *deviation);
/* The ray leaving the prism for a start ray (0, y, 0) along x: origin xyz,
 * direction xyz (CompoundPrism.exit_ray, src/pylib.rs:763-768). */
cpd_status cpd3_exit_ray(const cpd_design3 *design, double y, double wavelength, double out[6]);

/* Synthetic 3D workload (no reference counterpart; the reference's optimiser
 * draws 3D designs in designer.py:160-265): 3 prisms, 32 bins, a fiber beam
 * of 50 um core and NA 0.1, the detector positioned for acceptance 1 and
 * at least 10% of a 64-ray probe at three wavelengths landing in bins;
 * sampling 64 wavelengths x 4096 rays.  Design `index` is reproducible on
 * its own; *attempts_out (may be NULL) receives the draws used. */
cpd_status cpd3_generate_design(uint64_t index, cpd_params3 *params_out, cpd_sampling *sampling_out,
                                uint32_t *attempts_out);
/* Designs first_index .. first_index + n - 1 on `threads` host threads (0: all). */
cpd_status cpd3_generate_designs(uint64_t first_index, size_t n, uint32_t threads, cpd_params3 *params_out,
                                 cpd_design3 *designs_out, cpd_sampling *sampling_out);

/* fitness() of 3D designs on the GPU (src/fitness.rs:130-147 with the Dim3
 * spectrometer, as Spectrometer.cpu_fitness calls it, src/pylib.rs:1008-1026):
 * rows of wavelength samples are traced in FP64 on the device, rays in
 * order with the SEM <= 5e-3 early exit under CPD_MODE_ADAPTIVE (none under
 * CPD_MODE_FIXED), and folded on the host as mutual_information folds them.
 * Rays are the R1 points of Qrng<[f64; 4]> (fiber) or Qrng<[f64; 2]>
 * (Gaussian) from 0.5 (qrng.rs:167-255).  per_design may be NULL. */
cpd_status cpd3_fitness_batch(cpd_ctx *ctx, const cpd_design3 *designs, const cpd_sampling *sampling, size_t n,
                              cpd_fitness *out, cpd_status *per_design);
/* p(d|lambda) rows of one 3D design: Spectrometer.transmission_probability
 * (src/pylib.rs:1028-1058) over p_dets_l_wavelength (fitness.rs:48-69). */
cpd_status cpd3_p_dets_l_wavelength(cpd_ctx *ctx, const cpd_design3 *design, const double *wavelengths,
                                    size_t n_wavelengths, uint32_t max_m, uint32_t mode, double *out);
/* The reference's GPU estimator for fiber-beam 3D designs (the
 * ..._fiber_beam_culminating_toric_compound_prism kernels,
 * src/kernel/kernel.rs:169-170, driven by cuda_fitness,
 * src/cuda_fitness.rs:347-423): as cpd_gpu_fitness_batch, with rays the
 * points frac(seed + k alpha_i) of the 4-D sequence.  Traced and folded in
 * FP64.  Gaussian-beam designs fail with CPD_ERR_INVALID_ARGUMENT, where the
 * reference raises NotImplementedError (src/pylib.rs:1118-1130). */
cpd_status cpd3_gpu_fitness_batch(cpd_ctx *ctx, const cpd_design3 *designs, size_t n, const double *seeds,
                                  size_t n_seeds, uint32_t max_n, uint32_t nwarp, uint32_t max_eval,
                                  cpd_fitness *out, cpd_status *per_design);
/* One ray per (wavelength cdf, ray cdf) pair through the FP64 device trace:
 * ray_cdf holds 4 values per ray (fiber: rho, theta, deflection, rotation;
 * Gaussian: the first 2).  bin index (bin_count on a miss) and
 * transmittance (0 on a miss).  Replaces propagation_test_kernel
 * (src/spectrometer/kernel.rs:195-224) for the 3D kernels. */
cpd_status cpd3_propagate_rays(cpd_ctx *ctx, const cpd_design3 *design, const double *wavelength_cdf,
                               const double *ray_cdf, size_t n, uint32_t *bin_out, double *t_out);

#ifdef __cplusplus
}
#endif

#endif /* CPD_H */
