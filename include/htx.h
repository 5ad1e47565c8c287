/*
 * htx.h -- C ABI of the MI355X-native Hermite-TENO finite-volume residual /
 * time-step path (libhtx.so).
 *
 * The reference (indrow/Hermite-TENO) has no FFI: its solver classes are
 * driven from C++ (main.cpp).  This ABI is what those classes bind to: the
 * C++ facade headers in this directory (FV_System2D.h, FVHermite_System2D.h)
 * implement the reference's class interface over these calls, and Python
 * binds them with ctypes.  Each entry point names the reference member it
 * replaces.
 *
 * Conventions
 *   - Return 0 on success, a negative HTX_E* code on failure; the message is
 *     available from htx_last_error() (per host thread).  Nothing exits the
 *     process and no C++ exception crosses this boundary.
 *   - A handle owns its device buffers, its HIP stream and its control block.
 *     One handle per host thread; handles are not internally synchronised.
 *   - Host state arrays are field-major, interior cells only:
 *     q[(f * ny + j) * ld + i], f over the conserved fields (rho, rho u,
 *     rho v, E), then for the Hermite scheme the x- and y-derivative moments
 *     of the same four fields (12 fields in all).  ld >= nx.
 */
#ifndef HTX_H
#define HTX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HTX_ABI_VERSION 1

enum htx_error {
    HTX_OK = 0,
    HTX_EINVAL = -1,   /* bad argument (grid size 0, unknown enum, null pointer) */
    HTX_EDEVICE = -2,  /* HIP runtime error, or no device */
    HTX_ENOMEM = -3,   /* device allocation failed */
    HTX_ESTATE = -4,   /* call out of order (no state set) */
    HTX_ENOTSUP = -5,  /* combination not built yet */
    HTX_EIO = -6,      /* snapshot file missing, unreadable, truncated or corrupt */
    HTX_ENONFINITE = -7 /* the state holds a NaN or infinite value, or the CFL dt came out
                           non-finite or non-positive; stepping stops (steps become no-ops)
                           until a new state is set */
};

/* HTX_SCHEME_TENO5I: Lagrange::teno5i (Schemes_Lagrange.cpp:1235-1263), the
 * reconstruction FV::System1D ships with.  HTX_SCHEME_WENO5JS /
 * HTX_SCHEME_WENO5Z: Lagrange::weno5JS / weno5Z (Schemes_Lagrange.cpp:358-408),
 * the reference's alternatives for the 1D solver's reconstruction.
 * HTX_SCHEME_WENO5N .. HTX_SCHEME_LINEAR: the remaining 5-point routines of
 * Schemes::Lagrange, weno5N (:440-478), teno5N (:553-593), teno5a
 * (:1265-1294), weno5IS (:596-620), weno5IS_Extended (:649-717),
 * weno5ISmod(u, diff) (:622-647), weno5Z(u, diff) (:410-438), teno5(u, diff)
 * (:514-551) and linear (:1219-1233); diff is the cell size normal to the
 * face.  Each of these takes TENO5's place in the characteristic
 * reconstruction (FV_System1D.cpp:200-201 in 1D; FV_System2D.cpp:379, :394
 * in 2D, where a user of the reference swaps the call), everything else
 * unchanged; 2D builds: fp64 and fp64_fast, 1D: fp64.
 * HTX_SCHEME_HTENO: FVHermite::System2D (Hermite::hteno).  HTX_SCHEME_HWENO,
 * HTX_SCHEME_HTENOI, HTX_SCHEME_HLINEAR: the same solver with Hermite::hweno
 * (Schemes_Hermite.cpp:234-293), htenoi (:670-755) or linear (:757-767) in
 * hteno's place (FVHermite_System2D.cpp:459, :486); fp64 and fp64_fast. */
enum htx_scheme {
    HTX_SCHEME_TENO5 = 0,
    HTX_SCHEME_HTENO = 1,
    HTX_SCHEME_TENO5I = 2,
    HTX_SCHEME_WENO5JS = 3,
    HTX_SCHEME_WENO5Z = 4,
    HTX_SCHEME_WENO5N = 5,
    HTX_SCHEME_TENO5N = 6,
    HTX_SCHEME_TENO5A = 7,
    HTX_SCHEME_WENO5IS = 8,
    HTX_SCHEME_WENO5IS_EXT = 9,
    HTX_SCHEME_WENO5ISMOD = 10,
    HTX_SCHEME_WENO5Z_DIFF = 11,
    HTX_SCHEME_TENO5_DIFF = 12,
    HTX_SCHEME_LINEAR = 13,
    HTX_SCHEME_HWENO = 14,
    HTX_SCHEME_HTENOI = 15,
    HTX_SCHEME_HLINEAR = 16
};
/* HTX_FLUX_CU: Euler1D::CUFlux (Equations_Euler1D.cpp:137-154), the 1D
 * solver's commented alternative (FV_System1D.cpp:234); dim = 1 only. */
enum htx_flux { HTX_FLUX_RUSANOV_CHAR = 0, HTX_FLUX_HLLC = 1, HTX_FLUX_CU = 2 };
enum htx_bc { HTX_BC_ZERO_GRAD = 0, HTX_BC_PERIODIC = 1 };
/* HTX_PREC_FP64: the drop-in; every operation is one IEEE operation in the
 *   reference's order (no FMA contraction, divisions kept), so each step is
 *   bit-identical to an IEEE evaluation of the reference formulas.
 * HTX_PREC_FP64_FAST: FP64 with FMA contraction and one reciprocal per common
 *   denominator outside the TENO decision chain; residuals agree to rounding,
 *   but rounding can move a TENO cut-off decision that sits at its threshold.
 * HTX_PREC_FP32_MIXED: FP32 cut-off decisions and nonlinear weights, FP64
 *   values, projections, fluxes and RK accumulation (TENO5: the TENO5 cut-off
 *   and the transverse WENO weights; Hermite-TENO: the HTENO and TENO5
 *   cut-offs and the transverse WENO weights, the derivative limiter FP64).
 * HTX_PREC_LEARNED: FP32-mixed TENO5 where a learned classifier (see below)
 *   sends the faces it finds smooth to the optimal-weight linear scheme
 *   (Lagrange::linear), guarded by an FP64 positivity check, and the rest
 *   to TENO5; the classifier runs once per step, on the state the step
 *   starts from, and its flags serve all three stages; needs
 *   htx_load_classifier.
 * HTX_PREC_FP32 (TENO5, Hermite-TENO): FP32 bulk arithmetic.  The
 *   FP32-mixed cut-off decisions; the characteristic face values stored in
 *   FP32; the transverse quadrature, back-projection and HLLC / Rusanov flux
 *   (Hermite-TENO: also the derivative fluxes) in FP32; the Roe eigensystem
 *   and characteristic projection, the flux divergence, the RK accumulation
 *   and the CFL reduction in FP64 (Hermite-TENO: the Gauss sums and the
 *   derivative limiter too).  A face (Hermite-TENO: a Gauss point) whose FP32
 *   states have a density or pressure that is not positive or too small to
 *   be resolved in FP32 has its back-projection and flux recomputed in
 *   FP64.
 * HTX_PREC_LEARNED_FP32: HTX_PREC_LEARNED on the FP32 bulk arithmetic of
 *   HTX_PREC_FP32 (needs htx_load_classifier); the classifier's features
 *   count a stencil difference within 1e-5 of its scale (the face's density,
 *   pressure or sound speed) as zero, below which FP32-bulk states carry
 *   rounding noise. */
enum htx_precision {
    HTX_PREC_FP64 = 0,
    HTX_PREC_FP32_MIXED = 1,
    HTX_PREC_LEARNED = 2,
    HTX_PREC_FP64_FAST = 3,
    HTX_PREC_FP32 = 4,
    HTX_PREC_LEARNED_FP32 = 5
};

typedef struct htx_params {
    int struct_size;    /* sizeof(htx_params), for ABI checks */
    int dim;            /* 2, or 1 for FV::System1D (nx cells, 3 fields rho, rho u, E;
                           yb, yt, ny ignored; scheme TENO5 or TENO5I; FP64 only) */
    double xl, xr;      /* ctor (xl, xr, nx, ...), FV_System2D.h:12-15 */
    int nx;
    double yb, yt;
    int ny;
    double cfl, tmax, tinit, gamma;  /* cflnumber, tmax, tInit, specificHeatRatio */
    int scheme;         /* htx_scheme: FV::System2D / FVHermite::System2D */
    int flux;           /* htx_flux: FV_System2D.cpp:462-466 (Rusanov) / :465-466 (HLLC) */
    int bc;             /* htx_bc: FV_System2D.cpp:365-366 */
    int precision;      /* htx_precision */
    int ref_quirks;     /* 1: reproduce the reference's HLLC and maxSpeed quirks */
    int tile_x, tile_y; /* TENO5 stage tile: 0 = the default (16 x 15); 16 x 16 in every
                           precision; the FP64_FAST, FP32_MIXED
                           and LEARNED precisions also build 16 x 8 and 32 x 8.  Every shape
                           gives the same results bit for bit, except that the learned
                           precisions decide their linear transverse quadrature per wavefront */
} htx_params;

typedef struct htx_solver htx_solver;

/* Fill p with the defaults of a 2D TENO5/HLLC FP64 solver (quirks on). */
int htx_params_default(htx_params* p);

/* FV::System2D::System2D / FVHermite::System2D::System2D (FV_System2D.cpp:3-47,
 * FVHermite_System2D.cpp:3-50), and FV::System1D::System1D
 * (FV_System1D.cpp:3-33) for dim = 1: validates the parameters (a zero grid
 * size is HTX_EINVAL where the reference calls exit(1),
 * Schemes_Hermite.cpp:16-25), allocates the device state and sets t = tinit.
 * The 1D state is 3 rows of nx (ld >= nx) in htx_set_state / htx_get_state /
 * htx_rhs; its stages are FV::System1D::rhs + RungeKutta3
 * (FV_System1D.cpp:158-262) and its dt is cfl dx / max |u -+ c|
 * (:270-278). */
int htx_create(const htx_params* p, htx_solver** out);

/* Releases every device resource of the handle.  Null is ignored. */
void htx_destroy(htx_solver* h);

/* Replaces the cell(i,j).u assignment of the reference's initial-condition
 * methods: copies host state (layout above) to the device.  Also resets the
 * CFL accumulator from the new state. */
int htx_set_state(htx_solver* h, const double* q, size_t ld);

/* Copies the device state to host memory (the reference's cell(i,j).u). */
int htx_get_state(htx_solver* h, double* q, size_t ld);

/* FV::System2D::isentropicVortex (FV_System2D.cpp:76-99 with
 * Equations_Euler2D.cpp:513-528, 656-666), generated on the device: vortex
 * of strength 5 centred at (cx, cy) plus nmodes density modes, each 6
 * doubles (k, l, amplitude, phase, Lx, Ly) adding
 * amplitude*sin(2 pi (k x / Lx + l y / Ly) + phase) at fixed u, v, p.
 * nmodes <= 16. */
int htx_init_vortex(htx_solver* h, double cx, double cy, const double* modes, int nmodes);

/* FV::System2D::entropyWave / FVHermite::System2D::entropyWave
 * (FV_System2D.cpp:49-74, FVHermite_System2D.cpp:52-86 with
 * Equations_Euler2D.cpp:430-511), generated on the device: density
 * rho + amplitude sin(pi (x + y)) at constant velocity (u, v) and pressure p,
 * 9x9 Gauss cell averages of the conserved state and, for the Hermite scheme,
 * of its x and y derivatives.  The reference's case is (1, 1, 1, 1, 0.5).
 * Requires rho > |amplitude| and p > 0. */
int htx_init_entropy_wave(htx_solver* h, double rho, double u, double v, double p, double amplitude);

/* printResult's numerical error (FV_System2D.cpp:653-666,
 * FVHermite_System2D.cpp:821-835) computed on the device: the mean over the
 * cells of |rho - rho_exact|, where rho_exact is the 9x9 Gauss cell average
 * of the last analytic initial condition (htx_init_vortex or
 * htx_init_entropy_wave, whose exact solutions are their initial fields
 * advected with the background velocity: (1, 1) for the vortex, (u, v) for
 * the wave) shifted by velocity x t_shift.  The reference evaluates the
 * entropy wave and the Hermite vortex at t_shift = tFinal and the FV vortex
 * at t_shift = 0.  HTX_ESTATE when no such condition was set (any other
 * initial condition or htx_set_state clears it).  The cell sums are added
 * in a fixed tree order (the reference adds them in row order; the two agree
 * to rounding). */
int htx_l1_density_error(htx_solver* h, double t_shift, double* err);

/* Four-quadrant Riemann data, generated on the device and point-sampled at
 * cell centres as FV::System2D::Riemann2D does (FV_System2D.cpp:101-111 with
 * the quadrant template of Equations_Euler2D.cpp:530-617), with the split
 * point and the four states as arguments: states holds 4 primitive rows
 * (rho, u, v, p) for quadrants k = (x >= xs) + 2 (y >= ys).  For the Hermite
 * scheme the derivative moments are the face differences of the point data
 * (FVHermite_System2D.cpp:89-107). */
int htx_init_quadrant(htx_solver* h, double xs, double ys, const double* states);

/* Planar shock into a gas carrying Gaussian density blobs (no reference
 * counterpart; same point sampling): where (x - xs, y - ys) . (nx, ny) < 0
 * the primitive state post[4], elsewhere ambient[4] with
 * rho += A exp(-|x - c|^2 / r^2) for each blob row (cx, cy, A, r);
 * nblob <= 64. */
int htx_init_shock_blobs(htx_solver* h, double xs, double ys, double nx, double ny, const double* post,
                         const double* ambient, const double* blobs, int nblob);

/* One residual evaluation L(q) of the current state, FV::System2D::rhs
 * (FV_System2D.cpp:363-573); written to host memory in the state layout. */
int htx_rhs(htx_solver* h, double* rhs, size_t ld);

/* One SSP-RK3 stage with a given dt: the stage loop of
 * FV::System2D::RungeKutta3 (FV_System2D.cpp:575-612;
 * FVHermite_System2D.cpp:733-782), which the reference keeps private, exposed
 * as a per-stage hook.  Stage 1 forms u1 = un + dt L(un) from the current
 * state (for the Hermite scheme it also limits the derivative moments of un
 * in place, as the reference's first rhs() does), stage 2
 * u2 = 3/4 un + 1/4 (u1 + dt L(u1)), stage 3 (un + 2 (u2 + dt L(u2))) / 3,
 * which replaces the current state.  Time and step count do not change
 * (htx_step advances them); stage 3 leaves the CFL maxima of the new state
 * for the next dt.  Stages run in the order 1, 2, 3 (HTX_ESTATE otherwise;
 * setting a state or stepping starts the order again).  Enqueued without
 * blocking. */
int htx_stage(htx_solver* h, int stage, double dt);

/* The intermediate state u1 (stage 1) or u2 (stage 2) of the last stages
 * run, in the state layout (blocks). */
int htx_get_stage_state(htx_solver* h, int stage, double* q, size_t ld);

/* The dt that the next step would take (FV_System2D.cpp:623-645). */
int htx_compute_dt(htx_solver* h, double* dt);

/* One SSP-RK3 step, FV::System2D::RungeKutta3 (:575-612) with the dt of
 * timeStepping (:622-648).  Blocks; returns the dt taken (0 if t >= tmax). */
int htx_step(htx_solver* h, double* dt_taken);

/* Enqueues n steps without blocking the host; steps past tmax are no-ops. */
int htx_steps_async(htx_solver* h, int n);

/* FV::System2D::timeStepping (:614-651): advance until t >= tfinal (tfinal
 * replaces tmax).  *nsteps receives the number of steps taken (may be null). */
int htx_advance(htx_solver* h, double tfinal, int* nsteps);

/* Simulation time and number of steps taken so far (blocks). */
int htx_get_time(htx_solver* h, double* t, int* nsteps);

/* Waits for all work enqueued on the handle's stream. */
int htx_sync(htx_solver* h);

/* Snapshots (no reference counterpart; the reference only writes the final
 * ASCII .dat, FV_System2D.cpp:668-715): the state, simulation time, step
 * count and configuration in one binary file, little-endian:
 *   char magic[8] "HTXSNAP1"; int32 version (1);
 *   int32 nfield, nx, ny, scheme, flux, bc, precision, ref_quirks;
 *   double xl, xr, yb, yt, cfl, tmax, tinit, gamma, t; int64 nsteps;
 *   double q[nfield][ny][nx] (field-major interior cells, as htx_get_state);
 *   uint64 checksum: sum over the words w_i of q (as uint64) of
 *   w_i (2 i + 1), modulo 2^64.
 * htx_load needs a solver of the same grid, scheme, boundary condition,
 * flux, quirk setting, gamma and domain (HTX_EINVAL naming the field that
 * differs); it restores the state, t and the step count (the solver keeps
 * its own precision, CFL number and end time), so that continuing from a
 * snapshot reproduces the uninterrupted run.  HTX_EIO for a missing,
 * truncated or corrupt file. */
int htx_save(htx_solver* h, const char* path);
int htx_load(htx_solver* h, const char* path);

/* Stage-kernel timing with HIP events on the handle's stream.  While
 * enabled, every stage-kernel launch is bracketed by an event pair;
 * htx_timing_read blocks, returns the summed kernel milliseconds and the
 * number of launches since the last read, and clears them. */
int htx_timing_enable(htx_solver* h, int enable);
int htx_timing_read(htx_solver* h, double* kernel_ms, int* nlaunch);

/* ------------------------------------------------------------------------
 * Learned troubled-face classifier (no reference counterpart).
 *
 * A face's label is 1 when plain FP64 TENO5 keeps every candidate stencil on
 * both sides of the face for all four characteristic fields (cutOff,
 * Schemes_Lagrange.cpp:1141-1149), i.e. when TENO5 reduces to the
 * optimal-weight linear scheme there.  The classifier is an MLP
 * 12 -> 64 -> 64 -> 1 with GELU (x times a C1 cubic approximation of the
 * normal CDF, the same in training and inference) whose inputs are normalised
 * density and pressure differences across the face's 6-cell stencil and TENO
 * smoothness contrasts of density, pressure and velocity.  Parameters are
 * HTX_CLS_NPARAM FP32 values: W1[12][64], b1[64], W2[64][64], b2[64],
 * w3[64], b3 (row index = input).  Training runs on the GPU in FP32; the
 * solver's inference runs the weights in bf16 on the matrix cores.
 * ------------------------------------------------------------------------ */
#define HTX_CLS_NPARAM 5057

typedef struct htx_trainer htx_trainer;

/* A training set of up to capacity faces (<= 2^31) on the device. */
int htx_trainer_create(long long capacity, htx_trainer** out);
void htx_trainer_destroy(htx_trainer* t);

/* Appends the features and labels of every x face and every y face of the
 * solver's current state (2 nx ny faces; TENO5 scheme). */
int htx_trainer_add_state(htx_trainer* t, htx_solver* h);

/* Faces in the set and how many of them are labelled smooth. */
int htx_trainer_size(htx_trainer* t, long long* n, long long* n_smooth);

/* Copies faces [start, start + n): features (12 floats each) and labels;
 * either pointer may be null. */
int htx_trainer_read(htx_trainer* t, long long start, long long n, float* feat, unsigned char* label);

/* Random initial parameters (uniform Xavier weights from seed, zero biases)
 * and a fresh Adam state. */
int htx_trainer_init(htx_trainer* t, unsigned long long seed);
int htx_trainer_set_params(htx_trainer* t, const float* params);
int htx_trainer_get_params(htx_trainer* t, float* params);

/* Adam on the class-balanced logistic loss, troubled faces weighted by
 * troubled_weight; each epoch visits the set once in a seeded order, in
 * minibatches of about batch faces (>= 128).  epoch_loss (may be null)
 * receives the mean minibatch loss of each epoch. */
int htx_trainer_train(htx_trainer* t, int epochs, int batch, double lr, double troubled_weight,
                      unsigned long long seed, double* epoch_loss);

/* Gradient of the training loss over the whole set at the current
 * parameters (one pass, no update) and the loss itself. */
int htx_trainer_gradient(htx_trainer* t, double troubled_weight, float* grad, double* loss);

/* FP32 forward of every face of the set: counts[4] = smooth faces predicted
 * smooth, troubled faces predicted smooth, troubled predicted troubled,
 * smooth predicted troubled, a face being predicted smooth when its logit
 * exceeds threshold; logits (may be null) receives every logit. */
int htx_trainer_eval(htx_trainer* t, double threshold, long long* counts, float* logits);

/* The same with the rounding of the solver's inference: the two matrix
 * layers' weights, the features and layer 1's activations rounded to bf16
 * (sums in FP32).  Thresholds meant for htx_load_classifier are chosen on
 * these logits. */
int htx_trainer_eval_inference(htx_trainer* t, double threshold, long long* counts, float* logits);

/* Loads classifier parameters into a solver (packed to bf16); a face whose
 * logit exceeds threshold takes the linear scheme in precision
 * HTX_PREC_LEARNED, and so does a face whose six stencil states are equal
 * (TENO5 cuts no candidate there), whatever the threshold. */
int htx_load_classifier(htx_solver* h, const float* params, double threshold);

/* Logits (and, if feat is not null, the 12 features) of every face of one
 * family of the current state, dir 0 = x faces (cell (i, j) | (i+1, j)),
 * dir 1 = y faces, indexed j * nx + i, computed with the solver's bf16
 * matrix-core inference. */
int htx_classify(htx_solver* h, int dir, float* logits, float* feat);

/* 1D only: the KXRCF troubled-cell flags (1 = discontinuous) that
 * FV::System1D::rhs computes for its stage input (FV_System1D.cpp:166-192;
 * radius circleRadMod = (dx/2)^1.5, :19), here for the current state: the
 * density at the face upwind of the momentum's sign, reconstructed from the
 * 3-cell stencils of the two cells sharing it, jumps by more than radius
 * |rho|.  flags holds nx bytes, one per cell; the first and last cells read
 * the ghost cells of the boundary condition, as the reference's loop
 * (i = s+1 .. nxx-s-2, s = 2) does.  The reference's 1D scheme ignores the flags (both of its
 * branches take teno5i); they are reported as a classical detector's flag
 * rate. */
int htx_kxrcf_flags(htx_solver* h, unsigned char* flags);

/* 1D only: the LMPE troubled-cell flags of FVHermite::System1D::rhs
 * (FVHermite_System1D.cpp:239-257) for the current state, per field:
 * flags[eq nx + c], 3 nx bytes.  The reference sweeps the cells in order
 * (two ghost cells past each end, filled by the boundary condition) and tests
 * a cell not yet flagged: A = |-u_{i-1} + 2 u_i - u_{i+1}| / 2 > dx / 2 and
 * 0 <= xc <= 3 dx with xc = (-2 u_{i-1} + 3 u_i - u_{i+1}) /
 * (-u_{i-1} + 2 u_i - u_{i+1}) dx flag cells i-1, i and i+1.  Here the sweep
 * starts from clear flags (the reference's reset loop is commented out, so
 * its flags accumulate across steps); the sequential skip rule is kept
 * exactly.  Reported as a classical detector's flag rate. */
int htx_lmpe_flags(htx_solver* h, unsigned char* flags);

/* Learned precision: the classifier flags (1 = smooth) of the last step's
 * state (or of the last htx_rhs) that the stage kernels read: x faces
 * (i, j) | (i+1, j) for i = -1 .. nx at fx[j * (nx + 2) + i + 1] (ny rows),
 * y faces (i, j) | (i, j+1) for j = -1 .. ny at fy[(j + 1) * nx + i]; the
 * faces one cell past each boundary see the boundary condition. */
int htx_classifier_flags(htx_solver* h, unsigned char* fx, unsigned char* fy);

/* Learned precision, summed since the last reset (each face once per
 * stage): faces reconstructed, faces the classifier flagged smooth, and
 * flagged faces that took the linear scheme; a flagged face whose linear
 * values back-project to a non-positive density or pressure (FP64 check)
 * falls back to TENO5.  reset != 0 zeroes the counts after reading. */
int htx_classifier_stats(htx_solver* h, long long* n_faces, long long* n_flagged, long long* n_linear, int reset);

/* ------------------------------------------------------------------------
 * Batches: many independent problems of uneven size advanced together (no
 * reference counterpart; the reference runs one System2D at a time).
 *
 * Every instance has its own parameter block (size, domain, end time tmax,
 * CFL, boundary condition, scheme) and its own dt.  The TENO5 instances of
 * a batch share flux and precision: in fp64_fast and fp32_mixed they run in
 * one persistent launch per RK stage over all their tiles (reference HLLC or
 * characteristic Rusanov); in fp64 (the drop-in's strict build), fp32,
 * learned and learned_fp32 each instance runs the single-problem stage
 * kernel of that precision, one launch per instance and stage, the learned
 * ones after a per-step face-flag launch (these give each instance the
 * single problem's results bit for bit).  Hermite instances share flux and
 * precision: fp64_fast or fp32_mixed in batched Hermite launches, fp64 one
 * strict launch sequence per instance.  A step can be captured once in a
 * hipGraph and replayed.  Instances past their tmax are skipped.
 * ------------------------------------------------------------------------ */
typedef struct htx_batch htx_batch;

int htx_batch_create(const htx_params* ps, int n, htx_batch** out);
void htx_batch_destroy(htx_batch* b);

/* Host state of instance i, the layout of htx_set_state / htx_get_state. */
int htx_batch_set_state(htx_batch* b, int i, const double* q, size_t ld);
int htx_batch_get_state(htx_batch* b, int i, double* q, size_t ld);

/* Four-quadrant Riemann data for instance i, as htx_init_quadrant. */
int htx_batch_init_quadrant(htx_batch* b, int i, double xs, double ys, const double* states);

/* Classifier of the learned precisions, as htx_load_classifier: required
 * before a batch with learned TENO5 instances steps (HTX_ESTATE otherwise). */
int htx_batch_load_classifier(htx_batch* b, const float* params, double threshold);

/* Enqueues n steps of every instance (without blocking). */
int htx_batch_steps(htx_batch* b, int n);

/* Steps until every instance reaches its tmax or max_steps steps were
 * enqueued; *nsteps receives the steps enqueued (may be null). */
int htx_batch_advance(htx_batch* b, int max_steps, int* nsteps);

int htx_batch_get_time(htx_batch* b, int i, double* t, int* nsteps);
int htx_batch_sync(htx_batch* b);

/* 1: capture a step into a hipGraph at the next htx_batch_steps and replay
 * it; 0: launch kernel by kernel. */
int htx_batch_use_graph(htx_batch* b, int enable);

/* Tile shape of the persistent TENO5 launch: 16x16, 32x8 or 8x32.  The
 * direct launches always use the single-problem kernels' default tile:
 * 16x16 is accepted for them and changes nothing, the others return
 * HTX_ENOTSUP. */
int htx_batch_set_tile(htx_batch* b, int tile_x, int tile_y);

/* Autotuner: times `steps` steps of the batch with every built tile shape
 * (from the current state, which is restored afterwards), keeps the fastest
 * and returns it; ms (may be null) receives the per-step milliseconds of the
 * three shapes in the order 16x16, 32x8, 8x32 (-1 for a shape not timed:
 * batches with direct TENO5 launches time their one shape, reported as
 * 16x16). */
int htx_batch_tune(htx_batch* b, int steps, int* tile_x, int* tile_y, double* ms);

/* ------------------------------------------------------------------------
 * One step without allocation.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).
 *
 * htx_step_noalloc performs what htx_step does for a 2D TENO5 solver (one
 * SSP-RK3 step with the CFL dt of the state, clamped to p->tmax; the same
 * kernels, the same results) on a state q in caller-owned device memory
 * (4 fields, field-major, pitch nx: q[(f * ny + j) * nx + i]), in place, at
 * time t; it writes the dt taken (0 when t >= p->tmax) to the device double
 * dt_taken (may be null).  Precisions fp64, fp64_fast and fp32_mixed.  The
 * work is enqueued on `stream` (a hipStream_t; null = the default stream)
 * and the call returns without waiting.
 *
 * The two intermediate stage states and the control block live in a fixed
 * device array of the library (one per GPU, 64 MiB, held until the process
 * ends): nx * ny may not exceed HTX_NOALLOC_MAX_CELLS (HTX_EINVAL).  The call
 * allocates and frees nothing.  Every call on a GPU uses that GPU's array:
 * calls that overlap on one GPU (other host threads, or streams not ordered
 * against each other) race and must be serialised by the caller.
 * ------------------------------------------------------------------------ */
#define HTX_NOALLOC_MAX_CELLS (1024LL * 1024LL)
int htx_step_noalloc(const htx_params* p, double* q, double t, double* dt_taken, void* stream);

/* Message of the last failed call on this host thread. */
const char* htx_last_error(void);

/* HTX_ABI_VERSION of the loaded library. */
int htx_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif /* HTX_H */
