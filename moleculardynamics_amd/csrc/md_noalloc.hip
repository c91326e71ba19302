// md_lj_forces_noalloc: the brute-force and cell-list LennardJones force
// evaluations of md_lj_forces on caller-owned device arrays, with every
// intermediate buffer in one file-scope device array, so that the call makes
// no device allocation. For callers that cannot allocate (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The scratch array holds, for at most kNoAllocAtoms atoms and kNoAllocCells
// cells: the staged positions (double4, w = upload index), the identity slot
// -> upload index map, the cell index arrays and the per-workgroup partial
// sums, plus the reduced scalars. Each device holds its own copy of the array
// (about 60 MiB) from the first call on it until the process ends. Calls on
// one device share it: calls that overlap on one GPU (other threads, other
// streams) race. Everything a call reads back is written earlier in the same
// call (cell counts are cleared by a fill kernel), so no state carries over
// from a previous call.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "md_kernels.hpp"
#include "../../include/moldyn_hip.h"

namespace {

constexpr long kNoAllocAtoms = MD_NOALLOC_MAX_ATOMS;
constexpr long kNoAllocCells = MD_NOALLOC_MAX_CELLS;
constexpr long kPartials = (kNoAllocAtoms + mdk::kBlock - 1) / mdk::kBlock + 8;

constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Byte offsets of the regions inside the scratch array.
constexpr size_t kOffPos = 0;
constexpr size_t kOffOrig = kOffPos + align256(sizeof(double4) * kNoAllocAtoms);
constexpr size_t kOffCellOf = kOffOrig + align256(sizeof(int) * kNoAllocAtoms);
constexpr size_t kOffTmp = kOffCellOf + align256(sizeof(int) * kNoAllocAtoms);
constexpr size_t kOffAtoms = kOffTmp + align256(sizeof(int) * kNoAllocAtoms);
constexpr size_t kOffCount = kOffAtoms + align256(sizeof(int) * kNoAllocAtoms);
constexpr size_t kOffStart = kOffCount + align256(sizeof(int) * kNoAllocCells);
constexpr size_t kOffFill = kOffStart + align256(sizeof(int) * (kNoAllocCells + 1));
constexpr size_t kOffPartD = kOffFill + align256(sizeof(int) * kNoAllocCells);
constexpr size_t kOffPartL = kOffPartD + align256(sizeof(double) * 3 * kPartials);
constexpr size_t kOffScalars = kOffPartL + align256(sizeof(long long) * 3 * kPartials);
constexpr size_t kScratchBytes = kOffScalars + align256(sizeof(mdk::Scalars));

// The host-side name of this array is not a device address: host code uses
// only the per-device pointer from hipGetSymbolAddress.
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
unsigned char* g_scratch[kMaxDevices] = {};

thread_local std::string g_error;

md_status fail(md_status st, const std::string& msg) {
    g_error = msg;
    return st;
}

// The current device's scratch address, looked up once per device.
md_status scratch_of_current_device(unsigned char** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(MD_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail(MD_ERR_RUNTIME, "device index beyond 64");
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_scratch[dev] == nullptr) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            return fail(MD_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                        hipGetErrorString(e));
        g_scratch[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch[dev];
    return MD_OK;
}

__global__ void k_stage_soa(const double* __restrict__ x, const double* __restrict__ y,
                            const double* __restrict__ z, long n, double4* __restrict__ pos,
                            int* __restrict__ orig) {
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pos[i] = make_double4(x[i], y[i], z[i], static_cast<double>(i));
    orig[i] = static_cast<int>(i);
}

}  // namespace

namespace mdk {
const char* noalloc_error() { return g_error.empty() ? nullptr : g_error.c_str(); }
}  // namespace mdk

md_status md_lj_forces_noalloc(int pair_source, const md_lj_params* p, long n, const double* L,
                               const double* x, const double* y, const double* z, double* fx,
                               double* fy, double* fz, double* energy, double* virial,
                               void* hip_stream) {
    g_error.clear();
    if (p == nullptr || L == nullptr || x == nullptr || y == nullptr || z == nullptr)
        return fail(MD_ERR_INVALID_ARGUMENT, "parameters, box and positions are required");
    if ((fx == nullptr) != (fy == nullptr) || (fx == nullptr) != (fz == nullptr))
        return fail(MD_ERR_INVALID_ARGUMENT, "fx, fy, fz must be all given or all NULL");
    if (!(p->cutoff > 0.0) || !(p->sigma > 0.0))
        return fail(MD_ERR_INVALID_ARGUMENT, "sigma and cutoff must be positive");
    for (int d = 0; d < 3; ++d)
        if (!(L[d] > 0.0)) return fail(MD_ERR_INVALID_ARGUMENT, "box lengths must be positive");
    if (n <= 0) return fail(MD_ERR_INVALID_ARGUMENT, "atom count must be positive");
    if (n > kNoAllocAtoms)
        return fail(MD_ERR_CAPACITY, "md_lj_forces_noalloc: " + std::to_string(n) +
                                         " atoms exceed the scratch maximum of " +
                                         std::to_string(kNoAllocAtoms));
    if (pair_source != MD_PAIRS_BRUTE && pair_source != MD_PAIRS_CELLS)
        return fail(MD_ERR_INVALID_ARGUMENT,
                    "md_lj_forces_noalloc takes MD_PAIRS_BRUTE or MD_PAIRS_CELLS");
    mdk::Grid g{};
    long m = 0;
    if (pair_source == MD_PAIRS_CELLS) {
        g = mdk::cell_grid(L, p->cutoff);
        if (!g.usable)
            return fail(MD_ERR_INVALID_ARGUMENT,
                        "CellList is not usable (fewer than 3 cells per side); fall back to "
                        "the brute-force overload");
        m = static_cast<long>(g.nc[0]) * g.nc[1] * g.nc[2];
        if (m > kNoAllocCells)
            return fail(MD_ERR_CAPACITY, "md_lj_forces_noalloc: " + std::to_string(m) +
                                             " cells exceed the scratch maximum of " +
                                             std::to_string(kNoAllocCells));
    }
    unsigned char* base = nullptr;
    md_status st = scratch_of_current_device(&base);
    if (st != MD_OK) return st;
    hipStream_t s = static_cast<hipStream_t>(hip_stream);

    double4* pos = reinterpret_cast<double4*>(base + kOffPos);
    int* orig = reinterpret_cast<int*>(base + kOffOrig);
    mdk::Scalars* sc = reinterpret_cast<mdk::Scalars*>(base + kOffScalars);
    hipLaunchKernelGGL(k_stage_soa, dim3(static_cast<unsigned>((n + mdk::kBlock - 1) / mdk::kBlock)),
                       dim3(mdk::kBlock), 0, s, x, y, z, n, pos, orig);

    mdk::DirectArgs a{};
    a.n = n;
    a.box = mdk::make_box(L);
    a.lj = mdk::lj_constants(p->sigma, p->epsilon, p->cutoff, p->linear_force_shift);
    a.pos = pos;
    a.frc = nullptr;
    a.fx_out = fx;
    a.fy_out = fy;
    a.fz_out = fz;
    a.orig = orig;
    a.part_d = reinterpret_cast<double*>(base + kOffPartD);
    a.part_l = reinterpret_cast<long long*>(base + kOffPartL);
    a.sc = sc;
    if (pair_source == MD_PAIRS_CELLS) {
        int* cell_of = reinterpret_cast<int*>(base + kOffCellOf);
        int* tmp = reinterpret_cast<int*>(base + kOffTmp);
        int* atoms = reinterpret_cast<int*>(base + kOffAtoms);
        int* count = reinterpret_cast<int*>(base + kOffCount);
        int* start = reinterpret_cast<int*>(base + kOffStart);
        int* fill = reinterpret_cast<int*>(base + kOffFill);
        mdk::launch_bin_index(n, pos, g, cell_of, count, start, fill, tmp, atoms, nullptr, s);
        a.grid = g;
        a.cell_of = cell_of;
        a.cell_start = start;
        a.cell_atoms = atoms;
    }
    mdk::launch_lj_direct(a, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(MD_ERR_HIP, std::string("launch: ") + hipGetErrorString(e));
    double ew[2] = {0.0, 0.0};
    e = hipMemcpyAsync(ew, &sc->energy, sizeof(ew), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return fail(MD_ERR_HIP, std::string("reading E and W back: ") + hipGetErrorString(e));
    if (energy) *energy = ew[0];
    if (virial) *virial = ew[1];
    return MD_OK;
}
