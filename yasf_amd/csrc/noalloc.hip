// Allocation-free coupling matvec (yasfhip_coupling_matvec_noalloc).
//
// The same y = W x as a session's FP64 matvec with the pair factors
// generated on the fly (pair_kernel + contract_kernel), for callers that
// cannot allocate device memory.  Everything beyond the caller's inputs and
// outputs -- the coupling plan of lmax, the padded positions and the Q
// partial sums -- lives in one file-scope device array, `scratch`, of
// kScratchBytes per GPU.  Q for all target tiles would not fit, so the
// target tiles are processed in chunks: pair GEMM of a chunk's tiles into the
// Q region, then their ab5 contraction.  Each (job, target tile) block does
// the same arithmetic as in the session's single launch, so the result is
// the session's bit for bit.
//
// The device address of `scratch` is looked up once per device with
// hipGetSymbolAddress (the array's name is a host-side stand-in in host
// code); every later use goes through that address.  Calls on one GPU share
// the array: overlapping calls (other streams or threads) race.
#include <cstring>
#include <memory>
#include <mutex>

#include "../../include/yasfhip.h"
#include "internal.hpp"

namespace yasfhip {

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // per GPU, held until the process ends
constexpr size_t kPlanRegion = size_t(3) << 20;      // the plan of lmax <= 8 needs 1.9 MB
constexpr int kNoallocMaxParticles = 32768;          // padded positions: 1 MiB
constexpr size_t kPosRegion = size_t(kNoallocMaxParticles) * sizeof(double4);
constexpr size_t kQOffset = kPlanRegion + kPosRegion;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's copy of `scratch` (looked up once per device).
unsigned char *scratch_address() {
    int dev = 0;
    YH_CHECK(hipGetDevice(&dev));
    YH_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        YH_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        YH_REQUIRE(ptr != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = ptr;
    }
    return static_cast<unsigned char *>(g_scratch_ptr[dev]);
}

// A plan's tables as one host image with the byte offsets of each array,
// built once per lmax and kept for the process (host memory only).
struct PlanImage {
    std::vector<unsigned char> bytes;
    size_t jobs, colmap, crow, cidx, cval, la, lb, ld;
};

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

template <typename T>
size_t append(std::vector<unsigned char> &img, const T *data, size_t count) {
    const size_t off = align256(img.size());
    img.resize(off + count * sizeof(T));
    if (count)
        std::memcpy(img.data() + off, data, count * sizeof(T));
    return off;
}

const PlanImage &plan_image(int lmax) {
    static std::mutex m;
    static std::unique_ptr<PlanImage> images[kMaxLmax + 1];
    std::lock_guard<std::mutex> lock(m);
    if (!images[lmax]) {
        const CouplingPlan &pl = coupling_plan_cached(lmax);
        auto im = std::make_unique<PlanImage>();
        auto &b = im->bytes;
        im->jobs = append(b, pl.jobs.data(), pl.jobs.size());
        im->colmap = append(b, pl.colmap.data(), pl.colmap.size());
        im->crow = append(b, pl.crow.data(), pl.crow.size());
        im->cidx = append(b, pl.cidx.data(), pl.cidx.size());
        im->cval = append(b, pl.cval.data(), pl.cval.size());  // (re, im) pairs = double2
        im->la = append(b, pl.leg_alpha.data(), pl.leg_alpha.size());
        im->lb = append(b, pl.leg_beta.data(), pl.leg_beta.size());
        im->ld = append(b, pl.leg_diag.data(), pl.leg_diag.size());
        b.resize(align256(b.size()));
        YH_REQUIRE(b.size() <= kPlanRegion, "coupling plan exceeds the scratch plan region");
        images[lmax] = std::move(im);
    }
    return *images[lmax];
}

__global__ void pad_positions_kernel(const double *__restrict__ p3, int N, int Npad, double4 *__restrict__ p4) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < Npad)
        p4[i] = i < N ? make_double4(p3[3 * i], p3[3 * i + 1], p3[3 * i + 2], 0.0) : make_double4(0, 0, 0, 0);
}

void require_device_range(const void *p, size_t bytes, const char *what) {
    hipPointerAttribute_t attr{};
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess)
        (void)hipGetLastError();
    YH_REQUIRE(e == hipSuccess && attr.type == hipMemoryTypeDevice, std::string(what) + " must be device memory");
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    e = hipMemGetAddressRange(&base, &size, const_cast<void *>(p));
    if (e != hipSuccess)
        (void)hipGetLastError();
    const char *lo = static_cast<const char *>(base), *q = static_cast<const char *>(p);
    YH_REQUIRE(e == hipSuccess && q >= lo && static_cast<size_t>(q - lo) + bytes <= size,
               std::string(what) + " is smaller than its size inside its device allocation");
}

}  // namespace

}  // namespace yasfhip

using namespace yasfhip;

extern "C" int yasfhip_coupling_matvec_noalloc(const double *positions, int particle_number, int lmax,
                                               double k_medium, const double *x, double *y, void *stream) {
    try {
        YH_REQUIRE(positions && x && y, "NULL argument");
        YH_REQUIRE(lmax >= 1 && lmax <= kMaxLmax, "lmax must be in [1, 8]");
        YH_REQUIRE(particle_number >= 1 && particle_number <= kNoallocMaxParticles,
                   "particle_number must be in [1, " + std::to_string(kNoallocMaxParticles) +
                       "] (the scratch array's position region)");
        YH_REQUIRE(std::isfinite(k_medium) && k_medium > 0.0, "k_medium must be real, finite and positive");
        const int N = particle_number;
        const int nmax = nmax_of(lmax);
        const size_t n = static_cast<size_t>(N) * nmax;
        require_device_range(positions, 3 * sizeof(double) * N, "positions");
        require_device_range(x, 16 * n, "x");
        require_device_range(y, 16 * n, "y");
        const PlanImage &im = plan_image(lmax);
        const CouplingPlan &pl = coupling_plan_cached(lmax);
        const size_t tile_q = static_cast<size_t>(pl.nq) * 256 * sizeof(double2);
        const int tiles_per_chunk = static_cast<int>((kScratchBytes - kQOffset) / tile_q);
        YH_REQUIRE(tiles_per_chunk >= 1, "one target tile's partial sums exceed the scratch array");
        unsigned char *base = scratch_address();
        hipStream_t s = static_cast<hipStream_t>(stream);

        DevicePlan dp{};
        dp.lmax = pl.lmax;
        dp.nmax = pl.nmax;
        dp.maxcols = pl.maxcols;
        dp.nq = pl.nq;
        dp.njobs = static_cast<int>(pl.jobs.size());
        dp.stage_elems = pl.stage_elems;
        dp.jobs = reinterpret_cast<PairJob *>(base + im.jobs);
        dp.colmap = reinterpret_cast<int *>(base + im.colmap);
        dp.crow = reinterpret_cast<int *>(base + im.crow);
        dp.cidx = reinterpret_cast<int *>(base + im.cidx);
        dp.cval = reinterpret_cast<double2 *>(base + im.cval);
        dp.la = reinterpret_cast<double *>(base + im.la);
        dp.lb = reinterpret_cast<double *>(base + im.lb);
        dp.ld = reinterpret_cast<double *>(base + im.ld);
        double4 *pos = reinterpret_cast<double4 *>(base + kPlanRegion);
        void *Q = base + kQOffset;

        YH_CHECK(hipMemcpyAsync(base, im.bytes.data(), im.bytes.size(), hipMemcpyHostToDevice, s));
        const int ntiles1 = (N + 15) / 16;
        const int Npad = ntiles1 * 16;
        hipLaunchKernelGGL(pad_positions_kernel, dim3((Npad + 255) / 256), dim3(256), 0, s, positions, N, Npad, pos);
        YH_CHECK(hipGetLastError());
        for (int t0 = 0; t0 < ntiles1; t0 += tiles_per_chunk) {
            const int ntl = std::min(tiles_per_chunk, ntiles1 - t0);
            launch_pair(dp, pos, N, k_medium, x, false, Q, s, nullptr, kAllPairs, t0, ntl);
            launch_contract(dp, N, Q, x, nullptr, 0, y, false, false, s, t0, ntl);
        }
        set_last_error("", kOk);
        return kOk;
    } catch (const Error &e) {
        return set_last_error(e.what(), e.code);
    } catch (const std::exception &e) {
        return set_last_error(e.what(), kInternal);
    }
}
