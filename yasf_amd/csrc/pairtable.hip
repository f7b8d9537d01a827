// Resident pair-factor table and the streaming pair GEMM.
//
// The pair factors G_{s1 s2}[p, dm] depend only on the geometry and the wave
// number, not on x.  When HBM allows (a 2,000-sphere lmax-6 cluster needs
// 10.8 GB in complex128), they are generated once per session and channel, in
// FP64, directly in the 16x16x4 MFMA A-operand layout:
//   table[target tile t][dm d][K-step ks][order ip][lane]      (complex T)
//   lane = (source - 4 ks) * 16 + target-in-tile,  ip = p - |dm|
// so every matvec is a pure streaming complex GEMM with no special functions:
// each wave loads its A fragments straight from HBM (1 KB coalesced per
// fragment), gathers its B fragments from x, and issues the MFMAs, with the
// next K-step's operands in flight.  No LDS, no barriers.
#include "table_gemm.hpp"

namespace yasfhip {

namespace {

// One block per (target tile, 16-source chunk) -- per (target tile, dm,
// 16-source chunk) for the regular factors; thread = one pair (the pair
// factors of all its dm: geometry and Hankel functions formed once).
// Builds target tiles t0.. of the table (table points at tile t0: a slab of
// the full table when t0 > 0 or fewer tiles are launched).  gate: run only
// while *gate == 1 (batched callers), null: always.
// Regular: factors of the regular (J) translation, j_p in place of h_p and
// coincident pairs included.
template <typename T2, bool Regular>
__global__ __launch_bounds__(256) void pair_table_build_kernel(
    const double4 *__restrict__ pos, int N, int N4, double k, int lmax, int ndm,
    const int *__restrict__ dmoff, const double *__restrict__ la, const double *__restrict__ lb,
    const double *__restrict__ ld, T2 *__restrict__ table, int t0, const int *__restrict__ gate,
    double near2) {
    if (gate && *gate != 1)
        return;
    __shared__ int dms[4 * kMaxLmax + 1];
    if (threadIdx.x < ndm)
        dms[threadIdx.x] = dmoff[threadIdx.x];
    __syncthreads();
    const int nchunk = (N + 15) / 16;
    const int c = blockIdx.x % nchunk;
    const int d = Regular ? (blockIdx.x / nchunk) % ndm : 0;
    const int t = blockIdx.x / (nchunk * (Regular ? ndm : 1));
    const int P = 2 * lmax + 1;
    const int SR = P * P;
    const int r = threadIdx.x & 15, kk = threadIdx.x >> 4;
    const int s1 = (t0 + t) * 16 + r, s2 = c * 16 + kk;
    const int ks = s2 >> 2;
    if (ks >= N4)
        return;
    const int lane = (s2 & 3) * 16 + r;
    const double4 p1 = pos[s1];  // pos is padded to a multiple of 16
    const double4 p2 = pos[s2 < N ? s2 : 0];
    const double dx = p1.x - p2.x, dy = p1.y - p2.y, dz = p1.z - p2.z;
    // near2: pairs farther apart than the near-field radius do not couple
    const bool valid = (s1 < N) && (s2 < N) && (Regular || (s1 != s2 && dx * dx + dy * dy + dz * dz <= near2));
    T2 *tile = table + static_cast<size_t>(t) * SR * N4 * 64 + lane;
    if constexpr (Regular) {
        const int dm = d - 2 * lmax;
        const int R = P - (dm < 0 ? -dm : dm);
        pair_series_regular<T2>(dx, dy, dz, k, dm, P, valid, la, lb, ld,
                                tile + (static_cast<size_t>(dms[d]) * N4 + static_cast<size_t>(ks) * R) * 64, 64);
    } else {
        pair_series_all(dx, dy, dz, k, lmax, valid, la, lb, ld, [&](int dd, int ip, double gr, double gi) {
            const int R = P - (dd < 2 * lmax ? 2 * lmax - dd : dd - 2 * lmax);
            T2 v;
            v.x = gr;
            v.y = gi;
            tile[(static_cast<size_t>(dms[dd]) * N4 + static_cast<size_t>(ks) * R + ip) * 64] = v;
        });
    }
}

// One block of WAVES waves per (job, target tile), which deals the job's
// orders out over its waves (table_gemm_block); each table fragment a wave
// loads feeds all of the job's (up to NTLMAX) column tiles.  Fewer waves give
// each wave more orders: more MFMAs per K-step to hide the operand loads
// behind, at more accumulator registers.  The grid covers ntl target tiles
// from t0 (table points at tile t0).
template <typename T, int WAVES, int NIPMAX, int NTLMAX, bool SEQ>
__global__ __launch_bounds__(64 * WAVES) void pair_table_gemm_kernel(
    const typename Cplx<T>::type *__restrict__ table, const typename Cplx<T>::type *__restrict__ x, int N,
    int N4, int nmax, const PairJob *__restrict__ jobs, int ntl, int t0, const int *__restrict__ colmap,
    int maxcols, int lmax, const int *__restrict__ dmoff, const int *__restrict__ order_tiles,
    T *__restrict__ Q, int nq, const int *__restrict__ gate) {
    if (gate && *gate != 1)
        return;
    const int jid = blockIdx.x / ntl;
    const int t = blockIdx.x - jid * ntl;  // tile within the slab
    const int dm = uniform(jobs[jid].dm);
    const int tile0 = uniform(jobs[jid].tile0);
    const int ntile = uniform(jobs[jid].ntile);
    const int qbase = uniform(jobs[jid].qbase);
    const int tstride = uniform(jobs[jid].tstride);
    const int P = 2 * lmax + 1;
    const int R = P - (dm < 0 ? -dm : dm);
    const int d = dm + 2 * lmax;
    const int nord = uniform(jobs[jid].nord);  // orders computed (R is the table's order stride)
    const TableGemmOperands<T> op{table + (static_cast<size_t>(t) * P * P + dmoff[d]) * N4 * 64,
                                  x,
                                  N,
                                  N4,
                                  nmax,
                                  R,
                                  tile0,
                                  order_tiles + d * P + (dm < 0 ? -dm : dm),
                                  colmap + d * maxcols + tile0 * 16};
    T *Qt = Q + 2 * (static_cast<size_t>(t0 + t) * nq + qbase) * 256;
    table_gemm_block<T, WAVES, NIPMAX, NTLMAX, SEQ>(op, nord, ntile, Qt, tstride);
}

// jobs: the job list (count njobs) of the shape's column-tile grouping.
// SEQ: each wave computes its whole orders one at a time (complex128 jobs
// wider than 4 column tiles, whose two-orders-at-once accumulators exceed a
// wave's registers).
template <typename T, int WAVES, int NTLMAX, bool SEQ = false>
void gemm_g(const DevicePlan &dp, const PairJob *jobs, int njobs, int N, const void *table, const void *x, void *Q,
            hipStream_t s, int t0, int ntl, const int *gate) {
    typedef typename Cplx<T>::type T2;
    const int N4 = (N + 3) / 4;
    const dim3 grid(njobs * ntl), block(64 * WAVES);
    const T2 *tb = static_cast<const T2 *>(table);
    const T2 *xv = static_cast<const T2 *>(x);
    T *q = static_cast<T *>(Q);
    // whole orders per wave for 2L+1 <= 15 and <= 17 orders
    constexpr int kNipLo = SEQ ? 1 : table_gemm_nipmax<WAVES>(15), kNipHi = SEQ ? 1 : table_gemm_nipmax<WAVES>(17);
    if (2 * dp.lmax + 1 <= 15 || kNipLo == kNipHi)
        hipLaunchKernelGGL((pair_table_gemm_kernel<T, WAVES, kNipLo, NTLMAX, SEQ>), grid, block, 0, s, tb, xv, N, N4,
                           dp.nmax, jobs, ntl, t0, dp.colmap, dp.maxcols, dp.lmax, dp.dmoff, dp.order_tiles, q, dp.nq,
                           gate);
    else
        hipLaunchKernelGGL((pair_table_gemm_kernel<T, WAVES, kNipHi, NTLMAX, SEQ>), grid, block, 0, s, tb, xv, N, N4,
                           dp.nmax, jobs, ntl, t0, dp.colmap, dp.maxcols, dp.lmax, dp.dmoff, dp.order_tiles, q, dp.nq,
                           gate);
    YH_CHECK(hipGetLastError());
}

template <typename T>
void gemm_t(const DevicePlan &dp, int shape, int N, const void *table, const void *x, void *Q, hipStream_t s,
            int t0, int ntl, const int *gate) {
    static_assert(kTableGrouping[0] == 2 && kTableGrouping[1] == 4 && kTableGrouping[2] == 8,
                  "instances follow the column-tile groupings");
    YH_REQUIRE(table_shape_offered(shape, sizeof(T) == sizeof(float)),
               "table GEMM shape not offered in this precision");
    const int g = shape % kTableGroupings;
    const PairJob *jobs = dp.tjobs[g];
    const int nj = dp.ntjobs[g];
    if (shape < kTableGroupings) {  // 8 waves
        if (g == 0)
            gemm_g<T, 8, 2>(dp, jobs, nj, N, table, x, Q, s, t0, ntl, gate);
        else if (g == 1)
            gemm_g<T, 8, 4>(dp, jobs, nj, N, table, x, Q, s, t0, ntl, gate);
        else if constexpr (sizeof(T) == sizeof(float))
            gemm_g<T, 8, 8>(dp, jobs, nj, N, table, x, Q, s, t0, ntl, gate);
        else  // complex128: one order at a time, jobs of up to kTableWideTiles64 column tiles
            gemm_g<T, 8, kTableWideTiles64, true>(dp, dp.tjobs_w64, dp.ntjobs_w64, N, table, x, Q, s, t0, ntl, gate);
    } else {  // 4 waves
        if (g == 0)
            gemm_g<T, 4, 2>(dp, jobs, nj, N, table, x, Q, s, t0, ntl, gate);
        else if constexpr (sizeof(T) == sizeof(float))
            gemm_g<T, 4, 4>(dp, jobs, nj, N, table, x, Q, s, t0, ntl, gate);
    }
}

}  // namespace

size_t pair_table_bytes(const DevicePlan &dp, int N, bool fp32) {
    const size_t P = 2 * dp.lmax + 1;
    const size_t ntiles1 = (N + 15) / 16, N4 = (N + 3) / 4;
    return ntiles1 * P * P * N4 * 64 * (fp32 ? sizeof(float2) : sizeof(double2));
}

size_t pair_table_tile_bytes(const DevicePlan &dp, int N, bool fp32) {
    const size_t P = 2 * dp.lmax + 1;
    const size_t N4 = (N + 3) / 4;
    return P * P * N4 * 64 * (fp32 ? sizeof(float2) : sizeof(double2));
}

void launch_pair_table_build(const DevicePlan &dp, const double4 *pos, int N, double k, bool fp32,
                             void *table, hipStream_t s, int t0, int ntl, const int *gate, bool regular,
                             double near2) {
    const int ntiles1 = (N + 15) / 16, N4 = (N + 3) / 4, nchunk = (N + 15) / 16;
    if (ntl < 0)
        ntl = ntiles1 - t0;
    YH_REQUIRE(t0 >= 0 && ntl >= 1 && t0 + ntl <= ntiles1, "table tile range outside the targets");
    const int ndm = 4 * dp.lmax + 1;
    const dim3 grid(ntl * (regular ? ndm : 1) * nchunk), block(256);
    YH_REQUIRE(!(regular && fp32), "the regular table is FP64 only");
    if (regular)
        hipLaunchKernelGGL((pair_table_build_kernel<double2, true>), grid, block, 0, s, pos, N, N4, k, dp.lmax,
                           ndm, dp.dmoff, dp.la, dp.lb, dp.ld, static_cast<double2 *>(table), t0, gate, near2);
    else if (fp32)
        hipLaunchKernelGGL((pair_table_build_kernel<float2, false>), grid, block, 0, s, pos, N, N4, k, dp.lmax,
                           ndm, dp.dmoff, dp.la, dp.lb, dp.ld, static_cast<float2 *>(table), t0, gate, near2);
    else
        hipLaunchKernelGGL((pair_table_build_kernel<double2, false>), grid, block, 0, s, pos, N, N4, k, dp.lmax,
                           ndm, dp.dmoff, dp.la, dp.lb, dp.ld, static_cast<double2 *>(table), t0, gate, near2);
    YH_CHECK(hipGetLastError());
}

void launch_pair_table_gemm(const DevicePlan &dp, int N, const void *table, const void *x, bool fp32,
                            int grouping, void *Q, hipStream_t s, int t0, int ntl, const int *gate) {
    YH_REQUIRE(grouping >= 0 && grouping < kTableShapes, "unknown table GEMM shape");
    const int ntiles1 = (N + 15) / 16;
    if (ntl < 0)
        ntl = ntiles1 - t0;
    YH_REQUIRE(t0 >= 0 && ntl >= 1 && t0 + ntl <= ntiles1, "table tile range outside the targets");
    YH_REQUIRE(table_gemm_extent_ok(N, dp.nmax, dp.lmax, fp32 ? sizeof(float2) : sizeof(double2)),
               "cluster too large for the table GEMM's 32-bit offsets");
    if (fp32)
        gemm_t<float>(dp, grouping, N, table, x, Q, s, t0, ntl, gate);
    else
        gemm_t<double>(dp, grouping, N, table, x, Q, s, t0, ntl, gate);
}

}  // namespace yasfhip
