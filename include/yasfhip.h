/*
 * yasfhip: C ABI of the MI355X (gfx950) multiple-sphere T-matrix solve path.
 *
 * Drop-in boundary for the reference's coupling-backend seam
 * (yasfpy/coupling/backends.py:36-66 CouplingBackend.multiply, selected by
 * yasfpy/simulation.py:95-106) and for the solve it drives
 * (yasfpy/scattered_field_coefficients.py:95-156, yasfpy/solver.py:108-121).
 *
 * Conventions
 *  - Complex arrays are interleaved (re, im) doubles; complex64 arrays are
 *    interleaved floats.  Vectors are particle-major: j = s * nmax + n with
 *    nmax = 2 L (L + 2) and the reference's mode order
 *    n = (tau-1) L(L+2) + l^2 - 1 + l + m (yasfpy/functions/cpu_numba.py:395-405).
 *  - Every call returns 0 on success, otherwise a status code; the message of
 *    the last failure on the calling thread is yasfhip_last_error().
 *  - A session binds one HIP stream on the current device.  Calls on one
 *    session are synchronous with respect to the host and not thread-safe.
 *  - Pointers are host pointers unless a function says otherwise.
 */
#ifndef YASFHIP_H
#define YASFHIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct yasfhip_session yasfhip_session;

enum {
    YASFHIP_OK = 0,
    YASFHIP_INVALID_ARGUMENT = 1,
    YASFHIP_DEVICE_ERROR = 2,
    YASFHIP_NOT_PREPARED = 3,
    YASFHIP_INTERNAL = 4
};

/* Arithmetic of the pair GEMM (the O(N^2) part of W x):
 *   YASFHIP_FP64    complex128 on the FP64 matrix cores;
 *   YASFHIP_FP32    complex64 on the FP32 matrix cores;
 *   YASFHIP_FP16X2  pair factors and x each carried as two fp16 terms (hi +
 *                   lo, scaled per order and per mode column by powers of
 *                   two) on the fp16 matrix cores with FP32 accumulation:
 *                   the accuracy of complex64 at the fp16 MFMA rate, over a
 *                   resident table of the complex64 table's size;
 *   YASFHIP_FP16    the hi terms only (relative error ~1e-4), half that table.
 * Every precision evaluates the special functions in FP64 and accumulates
 * the ab5 contraction in FP64.  The two fp16 forms need a resident table
 * (matvec mode TABLE, or AUTO where it fits). */
enum { YASFHIP_FP64 = 0, YASFHIP_FP32 = 1, YASFHIP_FP16X2 = 2, YASFHIP_FP16 = 3 };

/* How the coupling matvec obtains the pair factors P_p^{|dm|} h_p e^{i dm phi}:
 * generated inside the matvec kernel (ON_THE_FLY); generated once per session
 * and channel into a resident HBM table (ntiles*(2L+1)^2*ceil(N/4)*64 complex
 * values) that every later matvec streams (TABLE); or, inside every matvec, a
 * slab of that table (a range of 16-particle target tiles) built and streamed
 * at a time (SLAB: same kernels as TABLE, memory of a slab only).  AUTO uses
 * the table when it needs at most three quarters of the free device memory,
 * else slabs of up to a quarter of it. */
enum {
    YASFHIP_MATVEC_AUTO = 0,
    YASFHIP_MATVEC_ON_THE_FLY = 1,
    YASFHIP_MATVEC_TABLE = 2,
    YASFHIP_MATVEC_SLAB = 3
};

/* Message of the last failed call on this thread ("" if none). */
const char *yasfhip_last_error(void);

/* ab5[j1][j2][p] (nmax x nmax x (2L+1) complex) into out.
 * Replaces Numerics.compute_translation_table (yasfpy/numerics.py:187-280). */
int yasfhip_translation_table(int lmax, double *out);

/* Work counts of the coupling plan: complex MACs the pair kernel issues per
 * particle pair, and the number of nonzero ab5 entries. */
int yasfhip_plan_stats(int lmax, long long *pair_macs_per_pair, long long *nnz_ab5);

/* Create a session for particle positions (N x 3), expansion order lmax
 * (1..8) and real medium wave numbers k_medium[channels].  Replaces the
 * Simulation set-up of the coupling backend (yasfpy/simulation.py:95-167);
 * no O(N^2) tables are built. */
int yasfhip_create(const double *positions, int particle_number, int lmax, const double *k_medium,
                   int channels, yasfhip_session **out);
int yasfhip_destroy(yasfhip_session *s);

/* Diagonal T-matrix per channel, t[channels][N*nmax] complex
 * (Simulation.mie_coefficients[single_unique_array_idx, :, w].ravel(),
 * yasfpy/matrix_apply.py:104-107). */
int yasfhip_set_tmatrix(yasfhip_session *s, const double *t);

/* y = W x for one channel (CouplingBackend.multiply(x, idx),
 * yasfpy/coupling/backends.py:99-125).  precision: YASFHIP_FP64 computes in
 * complex128; YASFHIP_FP32, YASFHIP_FP16X2 and YASFHIP_FP16 run the pair
 * GEMM as described above (geometry and special functions are evaluated in
 * FP64 and rounded; the ab5 contraction accumulates in FP64) with
 * complex128 in/out. */
int yasfhip_coupling_matvec(yasfhip_session *s, int channel, const double *x, double *y,
                            int precision);

/* Allocation-free y = W x: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The same operator as
 * yasfhip_coupling_matvec in FP64 with the pair factors generated on the fly
 * (YASFHIP_MATVEC_ON_THE_FLY), bit for bit, without a session: positions
 * (N x 3 doubles), x and y (N * nmax complex128 each) are device pointers on
 * the current device; everything else (coupling plan, padded positions, the
 * pair GEMM's partial sums) lives in a 64 MiB array of device memory that
 * the library holds per GPU until the process ends.  Enqueued on stream
 * (a hipStream_t, NULL: the null stream) without waiting for it; calls no
 * hipMalloc/hipFree.  At most 32768 particles, lmax 1..8.  Calls on one GPU
 * share the array: calls that overlap on one GPU (other streams, threads)
 * race. */
int yasfhip_coupling_matvec_noalloc(const double *positions, int particle_number, int lmax,
                                    double k_medium, const double *x, double *y, void *stream);

/* y = x - t * (W x) (master_matrix_multiply, yasfpy/matrix_apply.py:66-120). */
int yasfhip_master_matvec(yasfhip_session *s, int channel, const double *x, double *y,
                          int precision);

/* Solve (I - T W) x = b for one channel with GMRES on the device, following
 * SciPy's gmres as the reference calls it (solver.py:108-121: rtol=tol,
 * atol=0, restart in [1, 800], maxiter in restart cycles; left
 * preconditioning, LAPACK-style Givens rotations, true-residual outer check).
 * The Arnoldi basis is orthogonalised by classical Gram-Schmidt with one
 * reorthogonalisation (SciPy: modified Gram-Schmidt; the two agree to
 * rounding), and each restart cycle runs on the device with one host
 * synchronisation per cycle.
 * x_inout holds x0 on entry (the reference defaults x0 = b) and the solution
 * on exit.  use_precond != 0 applies the block-diagonal preconditioner built
 * by yasfhip_precond_build.  info: 0 converged, else maxiter (SciPy).
 * history (may be NULL) receives up to history_cap values presid/||b|| per
 * inner iteration (SciPy's 'pr_norm' callback).  iterations = inner
 * iterations; residual = final ||b - A x|| / ||b||. */
int yasfhip_gmres_solve(yasfhip_session *s, int channel, const double *b, double *x_inout,
                        double rtol, int restart, int maxiter, int use_precond, int *info,
                        int *iterations, double *residual, double *history, int history_cap);

/* yasfhip_gmres_solve with b and x_inout in device memory of the session's
 * device (n complex128 values each): the same solve (solver.py:108-121) for
 * callers whose vectors are already in HBM.  b and x stay in device memory
 * (copied within HBM, never over PCIe); the solve's per-cycle scalars
 * (residual norms, the cycle's residual estimates) reach the host as they do
 * in yasfhip_gmres_solve.  Host pointers, and buffers that do not hold n values
 * inside one device allocation, are rejected (YASFHIP_INVALID_ARGUMENT)
 * before any device work; history stays a host array.  Returns when the
 * solve is done, as yasfhip_gmres_solve does. */
int yasfhip_gmres_solve_device(yasfhip_session *s, int channel, const double *b, double *x_inout,
                               double rtol, int restart, int maxiter, int use_precond, int *info,
                               int *iterations, double *residual, double *history, int history_cap);

/* Mixed-precision solve of (I - T W) x = b: FP64 iterative refinement around
 * GMRES whose operator runs the pair GEMM in complex64.  Each refinement step
 * forms the FP64 residual r = b - A x (FP64 matvec), solves A d = r with
 * SciPy-style GMRES (x0 = 0, the given restart and inner_maxiter cycles, the
 * preconditioner if use_precond, FP64 Arnoldi/Givens) to the relative
 * reduction max(inner_rtol, rtol ||b|| / (2 ||r||)), and adds d to x.  It
 * stops when ||b - A x|| <= rtol ||b|| (FP64 operator), after
 * max_refinements steps, or when a step does not reduce the residual.
 * info: 0 converged, 1 not.  refinements = steps taken; iterations = inner
 * iterations summed over the steps; residual = final FP64 ||b - A x|| / ||b||.
 * history (may be NULL) receives that relative residual before the first
 * step and after each step (up to history_cap values).  The reference solves
 * in FP64 only (yasfpy/solver.py:108-121); this entry point has the same
 * convergence criterion on the FP64 operator. */
int yasfhip_gmres_solve_refined(yasfhip_session *s, int channel, const double *b, double *x_inout,
                                double rtol, double inner_rtol, int restart, int inner_maxiter,
                                int max_refinements, int use_precond, int *info, int *refinements,
                                int *iterations, double *residual, double *history,
                                int history_cap);

/* Operator of the inner GMRES of yasfhip_gmres_solve_refined:
 * YASFHIP_FP32 (default), YASFHIP_FP16X2 or YASFHIP_FP16 (see above).  The
 * refinement's residuals stay FP64.  No counterpart in the reference. */
int yasfhip_set_inner_precision(yasfhip_session *s, int precision);

/* Extinction and scattering cross sections of one channel from the incident
 * (a_inc) and scattered (f) field coefficients, particle-major complex128
 * (Optics.compute_cross_sections, yasfpy/optics.py:55-156):
 *   c_ext = -Re sum conj(a) f * pi / k^2,
 *   c_sca = |sum_{j1,j2} conj(f_j1) W_J[j1,j2] f_j2| * pi / k^2,
 * W_J being the coupling operator with spherical Bessel j_p in place of h_p and
 * the coincident pairs s1 = s2 included (the CPU sum of
 * yasfpy/functions/cpu_numba.py:410-457), evaluated in FP64.  By default
 * C_sca comes from the equal far-field integral (yasfhip_set_scattering_method);
 * the pair sum runs over slabs of a regular pair table in the session's slab
 * buffer where its matvecs stream slabs (SLAB mode, or AUTO when the table
 * did not fit), otherwise with the factors generated on the fly. */
int yasfhip_cross_sections(yasfhip_session *s, int channel, const double *a_inc, const double *f,
                           double *c_ext, double *c_sca);

/* How yasfhip_cross_sections (and a server's completions) compute C_sca:
 * YASFHIP_CSCA_FAR_FIELD (default): C_sca k^2 = integral of |E|^2 over the
 * directions, by the product of an n_theta-point Gauss-Legendre rule in
 * cos(theta) and an n_phi-point rule in phi with n_theta = M + 1, n_phi = 2M + 2,
 * M = ceil(kR + lmax + 15^(2/3) (kR + 1)^(1/3)), R the largest distance of a
 * sphere from the origin: exact for the band-limited far field, it equals
 * the pair sum to rounding at O(directions x N x nmax) work instead of
 * O(N^2 nmax^2) (the far field of compute_electric_field_angle_components,
 * cpu_numba.py:535-606; the integral identity is the reference's own check,
 * tests/test_compare_mstm4.py:37-47).  YASFHIP_CSCA_PAIR_SUM: the regular
 * (J) pair sum of cpu_numba.py:410-457 as described above. */
enum { YASFHIP_CSCA_FAR_FIELD = 0, YASFHIP_CSCA_PAIR_SUM = 1 };
int yasfhip_set_scattering_method(yasfhip_session *s, int method);

/* Far-field angle components of the scattered electric field for
 * num_directions directions (polar[a], azimuthal[a]) in radians
 * (Optics.compute_scattered_e_field_angle_components, yasfpy/optics.py:177-266,
 * and compute_electric_field_angle_components,
 * yasfpy/functions/cpu_numba.py:535-606); e_theta and e_phi receive
 * num_directions complex128 values.  The phase function and polarisation
 * components follow from these on the host (optics.py:269-551). */
int yasfhip_far_field(yasfhip_session *s, int channel, const double *f, int num_directions,
                      const double *polar, const double *azimuthal, double *e_theta, double *e_phi);

/* Block-diagonal preconditioner for one channel: particles are grouped into
 * boxes by box_of_particle[N] (values 0..num_boxes-1); each box's block of
 * I - T W restricted to intra-box pairs is assembled and LU-factored on the
 * device.  Intended semantics of yasfpy/preconditioner.py:1-14, 92-219
 * (receiver T, particle-major indices, FP64). */
int yasfhip_precond_build(yasfhip_session *s, int channel, const int *box_of_particle,
                          int num_boxes);

/* Storage of the preconditioner's LU factors for the apply, taking effect at
 * the next yasfhip_precond_build: YASFHIP_FP32 (default) rounds the factors,
 * computed in FP64, to complex64 -- half the bytes each apply streams, the
 * reference's own block precision (preconditioner.py builds complex64
 * blocks) -- while the substitutions run in FP64 with FP64 diagonal
 * inverses, so the apply stays one fixed linear operator; YASFHIP_FP64 keeps
 * them in complex128 (M_box^-1 to FP64 rounding).  No counterpart in the
 * reference. */
int yasfhip_set_precond_factor_precision(yasfhip_session *s, int precision);

/* z = M^-1 r with the prepared preconditioner (BlockDiagonalPreconditioner.apply,
 * yasfpy/preconditioner.py:177-219). */
int yasfhip_precond_apply(yasfhip_session *s, const double *r, double *z);

/* Select the pair-factor strategy (YASFHIP_MATVEC_*); results agree to
 * rounding.  Switching to ON_THE_FLY frees the resident tables.  No
 * counterpart in the reference (its dense backend stores raw per-pair tables,
 * yasfpy/lookups.py:124-171). */
int yasfhip_set_matvec_mode(yasfhip_session *s, int mode);

/* Mie T-matrix diagonal on the device: t[s][n] (N x nmax complex, the
 * reference's mode order) of spheres with radius[s] and relative refractive
 * index relative_index[s] (complex, sphere index / medium index) in a medium
 * of real wave number k_medium: -b_l on tau = 1, -a_l on tau = 2 in Bohren &
 * Huffman's notation, replicated over m.  Replaces mie_coefficients
 * (yasfpy/mie_coefficients.py:29-161) for one wavelength. */
int yasfhip_mie_coefficients(int lmax, int particle_number, const double *radius, const double *relative_index,
                             double k_medium, double *t);

/* Plane-wave incident coefficients a[s][n] (N x nmax complex) on the device:
 * propagation direction (polar, azimuthal), polarization 1 (TE) or 2 (TM),
 * amplitude, focal point (NULL: origin).  Replaces
 * compute_initial_field_coefficients_planewave
 * (yasfpy/initial_field_coefficients.py:79-142) for one wavelength. */
int yasfhip_planewave_coefficients(int lmax, int particle_number, const double *positions, double k_medium,
                                   double polar, double azimuthal, int polarization, double amplitude,
                                   const double *focal_point, double *a);

/* Near-field approximation of W: pairs of spheres farther apart than radius
 * (same length unit as the positions) do not couple (radius 0: W = 0); a
 * negative or infinite radius restores the exact operator (the default).  Applies to the coupling and
 * master matvecs and the GMRES solves; the preconditioner and the cross
 * sections keep every pair.  Replaces NearFarCouplingBackend
 * (yasfpy/coupling/backends.py:259-459: near = |r_s1 - r_s2|^2 <= cutoff^2,
 * self-coupling excluded, far pairs dropped). */
int yasfhip_set_near_field_radius(yasfhip_session *s, double radius);

/* Target tiles per slab (SLAB and AUTO modes); 0 (default): as many as fit in
 * a quarter of the free memory.  Frees the current slab. */
int yasfhip_set_slab_tiles(yasfhip_session *s, int tiles);

/* Tile-shape autotuner of the resident-table (or slab) GEMM for one
 * precision: times `reps` launches of each shape and keeps the fastest for
 * this session.  Shapes 0, 1, 2: 8 waves per block with jobs of at most 2, 4
 * or 8 16-column tiles per dm; 3, 4: 4 waves with 2 or 4 (fewer waves give
 * each more orders).  FP32 offers 0-4, FP64 0-3, where shape 2 runs blocks of
 * 4 waves with one order per wave (8 tiles of two orders exceed a wave's
 * registers): wider jobs stream each table slice fewer times but hold more
 * accumulators.  Returns the
 * index in *grouping; ms[6] receives the mean times (-1 for shapes not
 * offered; all -1 when the session generates the pair factors on the fly, in
 * which case nothing changes).  The table (or first slab) is built first if
 * needed.  Until tuned (or set), a session uses shape 2 in both.  Results agree between shapes to rounding (the orders a block cuts
 * along K depend on the job's order count). */
int yasfhip_tune_matvec(yasfhip_session *s, int channel, int precision, int reps, int *grouping,
                        double *ms);

/* Fixes the table GEMM shape of one precision (an index offered above)
 * without timing.  No counterpart in the reference. */
int yasfhip_set_table_shape(yasfhip_session *s, int precision, int shape);

/* Timing of device-resident matvecs of one channel on the session stream
 * with HIP events around back-to-back launches: ms_pair = mean duration of
 * the pair kernel over `reps` launches in a row, ms_total = mean duration of
 * the whole matvec (pair + contraction) over `reps` in a row.  The
 * input is a deterministic pseudo-random vector (the GEMMs run measurably
 * faster on all-zero operands, which draw less power); the tuner above
 * times the same input. */
int yasfhip_time_matvec(yasfhip_session *s, int channel, int precision, int reps, double *ms_pair,
                        double *ms_total);

/* ------------------------------------------------------------------------
 * Batched serving (no counterpart in the reference, which solves one
 * Simulation at a time, yasfpy/scattered_field_coefficients.py:95-156).
 *
 * A server owns one HIP stream and one device-memory arena of memory_bytes
 * (0: 70% of the free memory), allocated at creation.  Submitted requests are
 * admitted first-in first-out while their resident state fits in the arena
 * (complex64 pair-factor table, Krylov basis of restart+1 vectors, vectors
 * and Hessenberg state); a request that cannot fit even alone completes with
 * info = -1.  Each admitted request is solved by FP64 iterative refinement
 * around restarted GMRES with the complex64 pair GEMM (x0 = b = t a_inc; one
 * refinement per restart cycle; converged when the FP64 ||b - A x|| <=
 * rtol ||b||; each cycle asked for max(inner_rtol, rtol ||b|| / (2 ||r||)));
 * all requests inside a cycle advance together, one Arnoldi iteration per
 * batched launch sequence (a HIP graph when use_graph != 0).
 * Calls on one server are host-synchronous and not thread-safe.
 * ------------------------------------------------------------------------ */
typedef struct yasfhip_server yasfhip_server;

int yasfhip_server_create(size_t memory_bytes, int restart, double inner_rtol, int use_graph,
                          yasfhip_server **out);
int yasfhip_server_destroy(yasfhip_server *s);

/* Queue one solve: positions (N x 3), lmax, real medium wave number,
 * T-matrix diagonal t[N*nmax] and incident coefficients a_inc[N*nmax]
 * (complex128, particle-major), relative tolerance, at most max_cycles
 * refinement cycles, and whether C_ext / C_sca are computed at completion
 * (as yasfhip_cross_sections).  The inputs are copied.
 * operator_key != 0 declares that the operator I - T W (positions, lmax,
 * wave number, T) is the same for every request with this key, e.g. the
 * incidence directions of one cluster at one wavelength: they share one
 * device copy of the geometry, T-matrix and pair-factor tables; only a_inc
 * differs.  A key already in use with another N, lmax or wave number is an
 * error; positions and T of later requests with the key are not read.
 * Pair-factor memory per operator: see yasfhip_server_set_table_policy. */
int yasfhip_server_submit(yasfhip_server *s, const double *positions, int particle_number, int lmax,
                          double k_medium, const double *t, const double *a_inc, double rtol,
                          int max_cycles, int observables, long long operator_key,
                          long long *request_id);

/* As yasfhip_server_submit, with the system set up on the device at
 * admission: the T-matrix diagonal from the spheres' radii and relative
 * refractive indices (complex, sphere / medium; as yasfhip_mie_coefficients)
 * and a_inc of a plane wave (as yasfhip_planewave_coefficients: direction
 * (polar, azimuthal), polarization 1 TE / 2 TM, amplitude, focal point or
 * NULL for the origin).  With an operator_key, radii and indices of later
 * requests with the key are not read.  Replaces the host set-up of
 * yasfpy/mie_coefficients.py:29-161 and initial_field_coefficients.py:79-142
 * per request. */
int yasfhip_server_submit_spheres(yasfhip_server *s, const double *positions, const double *radius,
                                  const double *relative_index, int particle_number, int lmax, double k_medium,
                                  double polar, double azimuthal, int polarization, double amplitude,
                                  const double *focal_point, double rtol, int max_cycles, int observables,
                                  long long operator_key, long long *request_id);

/* Admission, FP64 residuals of requests between cycles, completions, and one
 * batched Arnoldi iteration of every request inside a cycle.  Reports the
 * active and queued counts and the number of requests completed so far. */
int yasfhip_server_step(yasfhip_server *s, int *active, int *queued, long long *completed);

/* Next completed request (request_id = -1 when none): solution into x_out
 * when x_capacity >= N*nmax complex values, info (0 converged, 1 not within
 * max_cycles, -1 never fitted), refinement cycles, inner iterations, final
 * FP64 relative residual, C_ext, C_sca and the seconds from submission to
 * completion. */
int yasfhip_server_poll(yasfhip_server *s, long long *request_id, double *x_out, long long x_capacity,
                        int *info, int *cycles, int *iterations, double *residual, double *c_ext,
                        double *c_sca, double *latency_s);

/* Tile-shape autotuner of the batched complex64 pair GEMM.  A shape is
 * (waves per (job, target tile) block, column tiles per job); the candidates
 * are, in this order, (2, 2), (4, 2), (8, 2), (4, 4), (8, 4), (8, 8), each
 * usable where ceil((2 lmax + 1) / waves) orders per wave stay within 5 (2 or
 * 4 waves) or 3 (8 waves).  Shapes are kept per lmax and size class
 * (requests of fewer than 1000 particles, and of at least 1000).  Times every
 * usable shape on a synthetic cluster of particle_number spheres, keeps the
 * fastest for lmax and that size class and returns it; ms[6] receives the
 * mean times (-1 where not usable).  Needs an idle server. */
int yasfhip_server_autotune(yasfhip_server *s, int lmax, int particle_number, int reps, int *waves,
                            int *column_tiles, double *ms);

/* Pair-factor memory of operators admitted from now on.  policy 1 (default):
 * resident tables where they fit (complex64 and FP64 within a quarter of the
 * arena; the inner table (complex64 or fp16 terms) within 7/8, with an
 * FP64 slab of up to a sixteenth for the FP64 residual and C_sca), else a
 * slab of the table (a range of target
 * tiles, at most an eighth of the arena) rebuilt range by range in every
 * pair GEMM, once for all requests of the operator; 2: slabs only; 0: no
 * tables or slabs, every GEMM generates the pair factors on the fly (FP64
 * residuals use the residual pool when one is reserved, see below).  The
 * complex64 GEMMs of requests sharing an operator run as one launch whose
 * blocks interleave the requests, so each table slice is streamed once for
 * all of them.
 * slab_tiles > 0 fixes the FP64 16-particle target tiles per slab (complex64
 * GEMMs use twice as many), 0 sizes slabs to the arena. */
int yasfhip_server_set_table_policy(yasfhip_server *s, int policy, int slab_tiles);

/* Pair GEMM of the inner GMRES iterations of operators admitted from now
 * on: YASFHIP_FP32 (default: complex64 tables and GEMMs), YASFHIP_FP16X2 or
 * YASFHIP_FP16 (fp16-term tables and the fp16 pair GEMM; see the precision
 * enum).  Applies to operators that keep a resident table or a slab (the
 * slab then holds fp16 target tiles; the requests of a shared operator run
 * one fp16 GEMM, several per block); on-the-fly operators keep the
 * complex64 GEMM.  FP64 residuals decide convergence in every case.  Needs
 * an idle server (nothing queued or active). */
int yasfhip_server_set_inner_precision(yasfhip_server *s, int precision);

/* Reduction each restart cycle is asked for at least (the inner_rtol of
 * yasfhip_server_create): cycles started from now on aim at
 * max(inner_rtol, rtol ||b|| / (2 ||r||)).  With an operator of complex64
 * accuracy (YASFHIP_FP32, YASFHIP_FP16X2) 1e-6 lets most requests converge
 * after one cycle; the one-term YASFHIP_FP16 operator (relative error
 * ~1e-4) gains nothing below 1e-4.  No counterpart in the reference. */
int yasfhip_server_set_inner_rtol(yasfhip_server *s, double inner_rtol);

/* C_sca method of the completions from now on (YASFHIP_CSCA_FAR_FIELD, the
 * default, or YASFHIP_CSCA_PAIR_SUM; see yasfhip_set_scattering_method). */
int yasfhip_server_set_scattering_method(yasfhip_server *s, int method);

/* Fixes the tile shape used for lmax, both size classes (one of the
 * candidates above). */
int yasfhip_server_set_tile_shape(yasfhip_server *s, int lmax, int waves, int column_tiles);

/* Batched Arnoldi iterations per yasfhip_server_step (1..16, default 1; a
 * HIP graph holds them all): the host synchronises, refreshes, completes and
 * admits once per step.  A request whose cycle ends inside a step waits for
 * the step's end (its kernels skip it); results do not depend on the
 * setting.  No counterpart in the reference. */
int yasfhip_server_set_iterations_per_step(yasfhip_server *s, int iterations);

/* Residual pool: one FP64 slab, carved from the arena, that the FP64
 * residuals of operators holding neither an FP64 table nor a slab of their
 * own use in turn (residuals run one after another on the server's stream),
 * under every table policy, where it holds at least min(16, target tiles) of
 * the operator's FP64 tiles; else those residuals generate their pair factors
 * on the fly.  Servers with an arena of 32 GiB or more reserve
 * memory_bytes / 32 at creation.  bytes = 0 releases the pool; a pool that
 * does not fit the arena's free memory is an error.  Needs an idle server.
 * No counterpart in the reference. */
int yasfhip_server_set_residual_pool(yasfhip_server *s, size_t bytes);

/* FP64 residual pair GEMMs run since creation, by form: forms[0] resident
 * FP64 table, [1] the operator's own slab, [2] the residual pool, [3] on the
 * fly. */
int yasfhip_server_residual_forms(yasfhip_server *s, long long forms[4]);

#ifdef __cplusplus
}
#endif

#endif /* YASFHIP_H */
