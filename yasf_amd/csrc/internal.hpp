// Internal declarations shared by the yasfhip translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace yasfhip {

using cplx = std::complex<double>;

// Largest expansion order the device kernels are instantiated for.
constexpr int kMaxLmax = 8;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// Status codes returned through the C ABI (see include/yasfhip.h).
enum Status : int {
    kOk = 0,
    kInvalidArgument = 1,
    kDeviceError = 2,
    kNotPrepared = 3,
    kInternal = 4,
};

#define YH_CHECK(expr)                                                                   \
    do {                                                                                 \
        hipError_t yh_e_ = (expr);                                                       \
        if (yh_e_ != hipSuccess)                                                         \
            throw ::yasfhip::Error(::yasfhip::kDeviceError,                              \
                                   std::string(#expr) + ": " + hipGetErrorString(yh_e_)); \
    } while (0)

#define YH_REQUIRE(cond, msg)                                                            \
    do {                                                                                 \
        if (!(cond))                                                                     \
            throw ::yasfhip::Error(::yasfhip::kInvalidArgument, msg);                    \
    } while (0)

// ---------------------------------------------------------------- host tables

inline int nmax_of(int lmax) { return 2 * lmax * (lmax + 2); }

struct Mode {
    int tau, l, m;
};
std::vector<Mode> modes_of(int lmax);

// ab5[j1][j2][p] (reference yasfpy/numerics.py:224-272), row-major complex.
std::vector<cplx> translation_table(int lmax);
// The same, built once per process and lmax (thread-safe).
const std::vector<cplx> &translation_table_cached(int lmax);

// One workgroup's share of the pair GEMM for one 16-particle target tile:
// a signed dm, all orders p in [|dm|, 2L] (R = 2L+1-|dm| of them) and the
// column tiles [tile0, tile0+ntile) of the dm's compatible source columns.
// Q slots are laid out per dm independently of how its column tiles are
// grouped into jobs: C tile (ip, column tile it) of the dm goes to slot
// qbase + ip * tstride + it, qbase being the dm's first slot and tstride its
// number of column tiles.
// Only the first nord orders (p = |dm| .. |dm|+nord-1) are computed: higher
// orders have no nonzero ab5 term on the job's column tiles.
struct PairJob {
    int dm, tile0, ntile, qbase, cost, tstride, nord, pad;
};

// Column-tile groupings of the resident-table GEMM's job lists: a dm's column
// tiles are split evenly into jobs of at most kTableGrouping[g] tiles.  Wider
// groups stream each table slice from HBM fewer times but hold more
// accumulators per wave.
// Grid of the device GMRES's reductions (krylov.hip): at most this many
// blocks, each leaving one row of 4 doubles of partial sums.
constexpr int kRedBlocks = 512;
constexpr int kRedPartials = kRedBlocks * 4;  // doubles in one launch's partials

constexpr int kTableGroupings = 3;
constexpr int kTableGrouping[kTableGroupings] = {2, 4, 8};
constexpr int kTableMaxColTiles = 8;
// Column tiles per job of the complex128 GEMM that computes one order at a
// time (table shape 2 in FP64): 6 x 3 accumulators of 8 VGPRs and a
// three-deep operand ring fit a wave's registers; every dm of lmax <= 6 has
// at most 6 column tiles, so each table slice streams once.
constexpr int kTableWideTiles64 = 6;

constexpr int kWavesPerJob = 4;
constexpr int kWaveMaxTiles = 8;                                // C tiles per wave
constexpr int kJobMaxTiles = kWavesPerJob * kWaveMaxTiles;      // C tiles per job
constexpr int kJobMaxColTiles = 2;                              // column tiles per job

// Target modes per block of the ab5 contraction, and its mode groups (blocks
// per target tile): the plan packs the modes of each m1 into them.
constexpr int kContractModes = 16;
__host__ __device__ inline int contract_groups(int nmax) { return (nmax + kContractModes - 1) / kContractModes; }

// A block of the fp16 pair GEMM ("bin"): the orders of one or more jobs (a
// dm's group of column tiles and a range of its orders) dealt to the block's
// kHalfBinWaves waves, one order per wave in the two-term form and two in the
// one-term form, the jobs' column tiles side by side in the block's B stage
// (at most kHalfBinCap[c] per request).  Packing the short order ranges of
// large |dm| together keeps the waves busy: a block's time is a fixed part
// per K-step plus its busiest wave's MFMAs.  Bins of every capacity cover
// the same (order, column tile) products, each once.
constexpr int kHalfBinWaves = 8;
constexpr int kHalfBinTiles = 8;
constexpr int kHalfBinCaps = 3;
constexpr int kHalfBinCap[kHalfBinCaps] = {8, 4, 2};
struct HalfBin {
    int ntile;                           // column tiles of the bin
    int tcol[kHalfBinTiles];             // colmap offset of tile k's 16 source columns (-1: none)
    int wd[kHalfBinWaves];               // wave's dm + 2L (-1: no order)
    int wtb[kHalfBinWaves];              // wave's tiles [wtb, wte)
    int wte[kHalfBinWaves];
    int wip[kHalfBinWaves][2];           // wave's orders p - |dm| (-1: none)
    int wq[kHalfBinWaves][2];            // Q slot of bin tile 0 for that order (tile k: wq + k)
};

// Everything the coupling kernels need for one lmax.
struct CouplingPlan {
    int lmax = 0, nmax = 0, ndm = 0, maxcols = 0, nq = 0;
    std::vector<int> colmap;          // [ndm][maxcols] -> source mode n2 or -1
    std::vector<PairJob> jobs;        // per 16-particle target tile, sorted by cost
    std::vector<PairJob> table_jobs[kTableGroupings];  // resident-table GEMM jobs, sorted by cost
    std::vector<PairJob> table_jobs_w64;  // jobs of up to kTableWideTiles64 column tiles, sorted by cost
    int stage_elems = 0;              // complex elements of one pair-kernel LDS stage
    // CSR of the contraction: rows k < R = contract_groups(nmax) x
    // kContractModes, crow[k] .. crow[k + 1] its entries, crow[R + 1 + k]
    // its target mode n1 (-1: padding row); rows grouped by m1
    std::vector<int> crow;
    std::vector<int> cidx;            // Q column index q*16 + col
    std::vector<double> cval;         // ab5 entry (re, im) interleaved
    std::vector<double> leg_alpha;    // [(2L+1)^2] Legendre recurrence alpha(l, m)
    std::vector<double> leg_beta;     // [(2L+1)^2] Legendre recurrence beta(l, m)
    std::vector<double> leg_diag;     // [2L+1] P_m^m = leg_diag[m] * sin^m
    std::vector<int> order_tiles;     // [ndm][2L+1]: column tiles (prefix) of the dm that order p feeds
    // fp16 GEMM bins, [0] two-term (one order per wave), [1] one-term (two),
    // per capacity kHalfBinCap[c]; half_bin_tiles: their largest job widths
    // summed (the busiest waves' column tiles)
    std::vector<HalfBin> half_bins[2][kHalfBinCaps];
    int half_bin_tiles[2][kHalfBinCaps] = {};
    long long pair_macs_per_pair = 0; // complex MACs the on-the-fly pair kernel issues per pair
    long long nnz_ab5 = 0;
};
CouplingPlan build_coupling_plan(int lmax);
// The same, built once per process and lmax (thread-safe).
const CouplingPlan &coupling_plan_cached(int lmax);

// Device-side views of a CouplingPlan (coupling.hip, pairtable.hip).
struct DevicePlan {
    int lmax, nmax, maxcols, nq, njobs, stage_elems;
    PairJob *jobs;
    PairJob *tjobs[kTableGroupings];
    int ntjobs[kTableGroupings];
    PairJob *tjobs_w64;  // CouplingPlan::table_jobs_w64
    int ntjobs_w64;
    int *colmap, *crow, *cidx;
    double2 *cval;
    double *la, *lb, *ld;
    int *dmoff;  // [ndm] first order slot of each dm in a pair-table target tile
    int *order_tiles;  // [ndm][2L+1] column tiles that order p feeds (a prefix of the dm's tiles)
    HalfBin *hbins[2][kHalfBinCaps];  // CouplingPlan::half_bins
    int nhbins[2][kHalfBinCaps];
    int hbin_tiles[2][kHalfBinCaps];
};

// Records the message of the calling thread's last failure (yasfhip_last_error)
// and returns code.
int set_last_error(const std::string &msg, int code);

// Device copies of a plan's tables (asynchronous on s) and their release.
DevicePlan upload_device_plan(const CouplingPlan &plan, hipStream_t s);
void free_device_plan(DevicePlan &dp);

// On-the-fly pair GEMM (pair factors generated in the kernel).  gate (device
// int, may be null): the kernel does nothing unless *gate == 1 at launch.
// Pairs farther apart than sqrt(near2) are left out of W (the near-field
// approximation); kAllPairs keeps every pair.
constexpr double kAllPairs = HUGE_VAL;
// t0, ntl: only the target tiles [t0, t0 + ntl) (ntl < 0: to the last),
// written to Q from its start.
void launch_pair(const DevicePlan &dp, const double4 *pos, int N, double k, const void *x, bool fp32,
                 void *Q, hipStream_t s, const int *gate = nullptr, double near2 = kAllPairs, int t0 = 0,
                 int ntl = -1);
// Q of the regular (J-translation) operator, coincident pairs included, FP64;
// t0, ntl as launch_pair's.
void launch_pair_regular(const DevicePlan &dp, const double4 *pos, int N, double k, const void *x,
                         void *Q, hipStream_t s, int t0 = 0, int ntl = -1);
// ab5 contraction (FP64 accumulation) + optional master-operator epilogue.
// t0, ntl: only the target tiles [t0, t0 + ntl) (ntl < 0: to the last),
// read from Q from its start.  gate: as launch_pair.
void launch_contract(const DevicePlan &dp, int N, const void *Q, const void *x, const double2 *t,
                     int mode, void *y, bool q_fp32, bool io_fp32, hipStream_t s, int t0 = 0, int ntl = -1,
                     const int *gate = nullptr);
// Resident pair-factor table: bytes, build, and the streaming pair GEMM.
// Both work on target tiles [t0, t0 + ntl) (ntl < 0: to the last tile) of a
// table whose pointer is at tile t0, so a slab of a few tiles can stand in
// for a table too large to keep: build the slab, multiply, next slab.
size_t pair_table_bytes(const DevicePlan &dp, int N, bool fp32);
size_t pair_table_tile_bytes(const DevicePlan &dp, int N, bool fp32);  // one target tile
// regular: factors of the regular (J) translation (FP64 only).
void launch_pair_table_build(const DevicePlan &dp, const double4 *pos, int N, double k, bool fp32,
                             void *table, hipStream_t s, int t0 = 0, int ntl = -1,
                             const int *gate = nullptr, bool regular = false, double near2 = kAllPairs);
// Shapes of the resident-table GEMM: shape % kTableGroupings indexes
// kTableGrouping (column tiles per job), shape / kTableGroupings picks 8 or
// 4 waves per block.  Complex128 offers 8 waves with 2 or 4 column tiles, 4
// waves with 2, and as shape 2 8 waves that compute their orders one at a
// time over jobs of up to kTableWideTiles64 tiles (two orders at once over 8
// tiles do not fit a wave's registers); complex64 offers all but 4 waves
// with 8 tiles.
constexpr int kTableShapes = 2 * kTableGroupings;
constexpr bool table_shape_offered(int shape, bool fp32) {
    return shape >= 0 && shape < kTableShapes && (fp32 ? shape != 5 : shape <= 3);
}
// grouping: a table GEMM shape (above).
void launch_pair_table_gemm(const DevicePlan &dp, int N, const void *table, const void *x, bool fp32,
                            int grouping, void *Q, hipStream_t s, int t0 = 0, int ntl = -1,
                            const int *gate = nullptr);

// fp16 pair GEMM (half.hip): table and x-term sizes, the table build (pexp:
// per-order exponents, dmin: device scratch for the smallest pair distance),
// the conversion of x to its fp16 terms (colmax: nmax unsigned of scratch)
// and the GEMM into the complex64 Q.
size_t half_table_term_bytes(const DevicePlan &dp, int N);
size_t half_x_bytes(const DevicePlan &dp, int N);
void launch_half_table_build(const DevicePlan &dp, const double4 *pos, int N, double k, void *table, bool split,
                             int *pexp, unsigned long long *dmin, hipStream_t s, double near2);
// The build in two parts, for slabs: the order exponents once (dmin:
// scratch), then target tiles [t0, t0 + ntl) into tab (lo term term_halves
// past the hi term).
void launch_half_exponents(const DevicePlan &dp, const double4 *pos, int N, double k, unsigned long long *dmin,
                           int *pexp, hipStream_t s);
void launch_half_table_range(const DevicePlan &dp, const double4 *pos, int N, double k, void *tab,
                             size_t term_halves, bool split, int t0, int ntl, const int *pexp, hipStream_t s,
                             double near2, const int *gate = nullptr);
void launch_half_x(const DevicePlan &dp, int N, const double2 *x, void *xh, bool split, unsigned *colmax, int *xexp,
                   hipStream_t s, const int *gate = nullptr);
void launch_half_gemm(const DevicePlan &dp, int N, const void *table, const void *xh, bool split, const int *pexp,
                      const int *xexp, void *Q, hipStream_t s, const int *gate = nullptr, int t0 = 0, int ntl = -1,
                      size_t term_halves = 0);

// Vector kernels of the device GMRES (krylov.hip).  Reductions leave one row
// of 4 doubles per block in partials (kRedBlocks rows); the last block sums
// them in block order and resets counter.
struct Reducer {
    double *partials;
    unsigned *counter;
};
// w -= h_prev v_prev (when v_prev); out = [Re, Im] <v_next, w> and ||w||^2.
void launch_mgs_step(double2 *w, const double2 *v_prev, const double2 *h_prev, const double2 *v_next,
                     long n, Reducer red, double *out, hipStream_t s);
void launch_scale(double2 *y, const double2 *x, double a, long n, hipStream_t s);
void launch_add(double2 *x, const double2 *d, long n, hipStream_t s);
void launch_update(double2 *x, const double2 *V, const double2 *y, int m, long n, long ld, hipStream_t s);
void launch_residual(double2 *r, const double2 *b, const double2 *ax, long n, Reducer red, double *out,
                     hipStream_t s);
void launch_norm2(const double2 *x, long n, Reducer red, double *out, hipStream_t s);
void launch_to_fp32(float *dst, const double *src, long count, hipStream_t s, const int *gate = nullptr);
void launch_to_fp64(double *dst, const float *src, long count, hipStream_t s);
void launch_fill_test(double *dst, long count, hipStream_t s);  // timing input in [-1, 1)

// Device state of one GMRES cycle of the single-solve driver.  The step
// kernels (operator included) run only while active == 1; the Givens kernel
// of the step that ends the cycle clears it.
struct ArnoldiState {
    int active;       // 1 while the cycle runs
    int j;            // Arnoldi steps done in the cycle
    int breakdown;    // the last step found an invariant subspace
    int pad;
    double target;    // stop once the residual estimate is at most this (SciPy's ptol)
    double presid;    // residual estimate after the last step
    double inv_norm;  // 1 / h[j+1, j] of the last step
    double h0sq;      // ||w||^2 of the current step before orthogonalisation
    double wsq;       // ||w||^2 of the current step after the first projection
};

// Krylov basis, Hessenberg and Givens state of the single-solve GMRES.
struct ArnoldiBuffers {
    double2 *V;              // [restart + 1][n] basis
    long n;
    int restart, nblk;       // nblk: blocks of the Gram-Schmidt passes
    double2 *H;              // [restart][restart + 1] rotated Hessenberg columns
    double *gc;              // [restart] Givens cosines
    double2 *gs;             // [restart] Givens sines
    double2 *g;              // [restart + 1] rotated right-hand side
    double2 *h1, *h2;        // [restart + 1] Gram-Schmidt coefficients of the current step
    double2 *partA, *partB;  // [restart + 2][nblk] per-block partial sums (slot-major)
    double2 *y;              // [restart] cycle-end coefficients
    double *hist;            // [restart] residual estimate after each step of the cycle
    ArnoldiState *st;
};
int arnoldi_blocks(long n);         // nblk for vectors of n elements
int arnoldi_max_restart();          // largest restart the step kernels take
// Cycle start: V[0] /= sqrt(*nrm2) (device scalar ||V[0]||^2), g = [sqrt(*nrm2), 0, ...],
// state = {active, j = 0, target}.
void launch_arnoldi_start(const ArnoldiBuffers &a, const double *nrm2, double target, hipStream_t s);
// Orthogonalisation, Givens update and normalisation of step j (w = V[j+1]
// holds the operator applied to V[j]).
void launch_arnoldi_step(const ArnoldiBuffers &a, int j, hipStream_t s);
// Cycle end: y = H^-1 g over the cycle's steps, x += V y.
void launch_arnoldi_end(const ArnoldiBuffers &a, double2 *x, hipStream_t s);

// Block-diagonal preconditioner (precond.hip).
struct BlockPrecond;
BlockPrecond *precond_create(const CouplingPlan &plan, const std::vector<cplx> &ab5,
                             const double *pos_host, int N, const int *box_of_particle, int num_boxes,
                             double k, const double2 *t_dev, bool fp32_factors, hipStream_t s);
void precond_destroy(BlockPrecond *p);
void precond_apply(BlockPrecond *p, const double2 *r, double2 *z, hipStream_t s, const int *gate = nullptr);

// Observables (observables.hip).
void launch_far_field(const double4 *pos, const double2 *f, int N, int lmax, double k,
                      const double4 *dir, const double *polar, const double *azimuthal, int A,
                      double2 *F, double2 *e_theta, double2 *e_phi, hipStream_t s);
void csca_quadrature_grid(double kR, int lmax, int *n_theta, int *n_phi);
size_t csca_quadrature_partials(int n_theta, int n_phi);
void launch_csca_quadrature(const double4 *pos, const double2 *f, int N, int lmax, double k, int n_theta,
                            int n_phi, double *partials, double *out, hipStream_t s);
// The same for nreq <= csca_quadrature_batch_max(lmax) solutions f[i] of one
// geometry in one launch (the phase factors computed once for all), each
// with its own partials[i] and out[i]; the results equal separate launches'.
int csca_quadrature_batch_max(int lmax);
void launch_csca_quadrature_batch(const double4 *pos, const double2 *const *f, double *const *partials,
                                  double *const *out, int nreq, int N, int lmax, double k, int n_theta, int n_phi,
                                  hipStream_t s);

// Device allocation owned for a scope (freed on every exit path).
template <typename T>
class DeviceBuffer {
  public:
    explicit DeviceBuffer(size_t count) { YH_CHECK(hipMalloc(&p_, sizeof(T) * (count ? count : 1))); }
    ~DeviceBuffer() {
        if (p_)
            (void)hipFree(p_);
    }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    T *get() const { return p_; }

  private:
    T *p_ = nullptr;
};

// Per-request set-up on the device.  sph[s] = (radius, Re rel, Im rel, -) with
// rel = sphere index / medium index; t: [N][nmax] Mie T-matrix diagonal.
void launch_mie(const double4 *sph, int N, int lmax, double k, double2 *t, hipStream_t s);
// a[s][n] = c[n] e^{i k (r_s - r_f) . d(polar, azimuthal)}; focal: host double[3].
void launch_planewave(const double4 *pos, int N, int lmax, double k, double polar, double azimuthal,
                      const double *focal, const double2 *c, double2 *a, hipStream_t s);
// c[n] = 4 A e^{-i m azimuthal} B^dagger_n(polar) on the host (nmax entries).
void planewave_angular(int lmax, double polar, double azimuthal, int pol, double amplitude, double2 *c);

}  // namespace yasfhip
