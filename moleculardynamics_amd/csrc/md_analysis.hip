// Trajectory analyses on the device: g(r) pair histograms (Rdf) and the
// ring-buffered MSD / VACF time correlations (Msd, Vacf) over a context's
// device-resident system, so a run that samples them never copies a frame to
// the host.
//
// g(r). Every unordered pair is counted once, by device slot order (slot i <
// slot j), with r = sqrt(|minimum image of r_lo - r_hi|^2) where lo/hi order
// the pair by upload index, the orientation of analyze/rdf.cpp:30-40, then
// bin floor(r / width) when r < rMax. Counts are integers, so the histogram
// is exact whatever order the pairs are visited in. Two pair sources:
//   * all pairs, 256 x 256 tiles of the upper triangle, the j tile staged
//     in LDS (the reference's O(N^2) loop);
//   * a cell grid of cells at least rMax wide when 3 or more fit per axis:
//     each atom against the atoms of the 27 cells around it with a higher
//     slot. The pair set within rMax is the same.
// Each workgroup counts into an LDS histogram (uint32, at most 2^31 pair
// visits per workgroup, enforced by the grid size) and adds it into the
// global uint64 histogram once at its end.
//
// MSD / VACF. Per-atom state is indexed by upload index (SoA, x then y then
// z blocks of n doubles): the previous wrapped position and running
// unwrapped position (MSD), and a ring of period = maxLag + 1 frames. One
// thread per atom gathers its current position or velocity from its device
// slot, updates the state with the reference's arithmetic (msd.cpp:25-34),
// and, for every lag up to min(pushes, maxLag), forms |u - u_origin|^2 or
// v . v_origin and reduces it over the workgroup (fixed tree). A second
// kernel adds the workgroup partials of each lag, in block order, into the
// lag's running sum. Per-atom terms equal the reference's bit for bit; the
// sums over atoms differ from its sequential sums by rounding only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <string>
#include <vector>

#include "../../include/moldyn_hip.h"
#include "md_kernels.hpp"

namespace {

using mdk::Box;
using mdk::kBlock;
using mdk::kWave;

constexpr int kTile = kBlock;          // atoms per tile of the all-pairs source
constexpr int kLdsBins = 8192;         // LDS histogram (32 KiB)
constexpr long kMaxVisitsPerGroup = 1L << 31;

// ---- g(r) ------------------------------------------------------------------

struct RdfArgs {
    long n;
    Box box;
    double r_max, r_max2_hi, width;
    int bins;
    const double4* pos;
    unsigned long long* hist;
    // all pairs
    long ntile;
    long tile_pairs;  // ntile * ntile (the lower half exits at once)
    // cells
    mdk::Grid grid;
    const int* cell_of;
    const int* cell_start;
    const int* cell_atoms;
};

__device__ __forceinline__ void rdf_pair(const RdfArgs& a, const double4& pi, const double4& pj,
                                         unsigned int* lh) {
    double dx, dy, dz;
    mdk::pair_delta(a.box, pi, pj, pi.w > pj.w, dx, dy, dz);
    const double r2 = mdk::norm2_ref(dx, dy, dz);
    if (!(r2 < a.r_max2_hi)) return;  // sqrt(r2) >= r_max for every such r2
    const double r = sqrt(r2);
    if (!(r < a.r_max)) return;
    const double q = r / a.width;
    if (!(q < static_cast<double>(a.bins))) return;
    const int k = static_cast<int>(q);
    if (lh != nullptr)
        atomicAdd(&lh[k], 1u);
    else
        atomicAdd(&a.hist[k], 1ull);
}

template <bool LDS>
__device__ __forceinline__ void rdf_begin(unsigned int* lh, int bins) {
    if (LDS)
        for (int k = threadIdx.x; k < bins; k += kBlock) lh[k] = 0u;
    __syncthreads();
}

template <bool LDS>
__device__ __forceinline__ void rdf_flush(const RdfArgs& a, unsigned int* lh) {
    __syncthreads();
    if (LDS)
        for (int k = threadIdx.x; k < a.bins; k += kBlock)
            if (lh[k] != 0u) atomicAdd(&a.hist[k], static_cast<unsigned long long>(lh[k]));
}

template <bool LDS>
__global__ __launch_bounds__(kBlock) void k_rdf_tiles(RdfArgs a) {
    __shared__ double4 tj[kTile];
    __shared__ unsigned int lh_s[LDS ? kLdsBins : 1];
    unsigned int* lh = LDS ? lh_s : nullptr;
    rdf_begin<LDS>(lh_s, a.bins);
    for (long p = blockIdx.x; p < a.tile_pairs; p += gridDim.x) {
        const long bi = p / a.ntile, bj = p % a.ntile;
        if (bj < bi) continue;  // uniform over the workgroup
        const long j0 = bj * kTile;
        const long jn = a.n - j0 < kTile ? a.n - j0 : kTile;
        __syncthreads();  // the previous tile's readers are done
        if (threadIdx.x < jn) tj[threadIdx.x] = a.pos[j0 + threadIdx.x];
        __syncthreads();
        const long i = bi * kTile + threadIdx.x;
        if (i < a.n) {
            const double4 pi = a.pos[i];
            const int jb = bi == bj ? static_cast<int>(threadIdx.x) + 1 : 0;
            for (int jj = jb; jj < static_cast<int>(jn); ++jj) rdf_pair(a, pi, tj[jj], lh);
        }
    }
    rdf_flush<LDS>(a, lh_s);
}

template <bool LDS>
__global__ __launch_bounds__(kBlock) void k_rdf_cells(RdfArgs a) {
    __shared__ unsigned int lh_s[LDS ? kLdsBins : 1];
    unsigned int* lh = LDS ? lh_s : nullptr;
    rdf_begin<LDS>(lh_s, a.bins);
    const long t = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    if (t < a.n) {
        const double4 pi = a.pos[t];
        const mdk::Grid& g = a.grid;
        const int c = a.cell_of[t];
        const int cx = c % g.nc[0];
        const int cy = (c / g.nc[0]) % g.nc[1];
        const int cz = c / (g.nc[0] * g.nc[1]);
        for (int oz = -1; oz <= 1; ++oz)
            for (int oy = -1; oy <= 1; ++oy)
                for (int ox = -1; ox <= 1; ++ox) {
                    const int nb = (cx + ox + g.nc[0]) % g.nc[0] +
                                   g.nc[0] * ((cy + oy + g.nc[1]) % g.nc[1] +
                                              g.nc[1] * ((cz + oz + g.nc[2]) % g.nc[2]));
                    for (int k = a.cell_start[nb]; k < a.cell_start[nb + 1]; ++k) {
                        const int j = a.cell_atoms[k];
                        if (j > t) rdf_pair(a, pi, a.pos[j], lh);
                    }
                }
    }
    rdf_flush<LDS>(a, lh_s);
}

// ---- MSD / VACF --------------------------------------------------------------

struct CorrArgs {
    long n;
    Box box;
    const double4* src;  // positions (MSD) or velocities (VACF), device slot order
    const int* slot;     // slot[upload index]
    int first;           // first push (MSD: unwrapped = wrapped)
    long ring_slot, period, top;
    double* prev;     // 3n (MSD)
    double* cur;      // 3n
    double* ring;     // period * 3n
    double* partial;  // (top + 1) * gridDim.x
};

template <bool MSD>
__global__ __launch_bounds__(kBlock) void k_corr_push(CorrArgs a) {
#pragma clang fp contract(off)
    __shared__ double sm[kBlock / kWave];
    const long id = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    const bool live = id < a.n;
    const long n = a.n;
    double c0 = 0.0, c1 = 0.0, c2 = 0.0;
    if (live) {
        const double4 w = a.src[a.slot[id]];
        if (MSD) {
            if (a.first) {
                c0 = w.x; c1 = w.y; c2 = w.z;
            } else {
                // unwrapped += minimumImage(wrapped - previous wrapped)
                c0 = a.cur[id] + mdk::fold_ref(w.x - a.prev[id], a.box.L[0]);
                c1 = a.cur[n + id] + mdk::fold_ref(w.y - a.prev[n + id], a.box.L[1]);
                c2 = a.cur[2 * n + id] + mdk::fold_ref(w.z - a.prev[2 * n + id], a.box.L[2]);
            }
            a.prev[id] = w.x; a.prev[n + id] = w.y; a.prev[2 * n + id] = w.z;
        } else {
            c0 = w.x; c1 = w.y; c2 = w.z;
        }
        a.cur[id] = c0; a.cur[n + id] = c1; a.cur[2 * n + id] = c2;
        double* r = a.ring + a.ring_slot * 3 * n;
        r[id] = c0; r[n + id] = c1; r[2 * n + id] = c2;
    }
    for (long lag = 0; lag <= a.top; ++lag) {
        double v = 0.0;
        if (live) {
            const long o = (a.ring_slot + a.period - lag) % a.period;
            const double* r = a.ring + o * 3 * n;
            const double o0 = lag == 0 ? c0 : r[id];
            const double o1 = lag == 0 ? c1 : r[n + id];
            const double o2 = lag == 0 ? c2 : r[2 * n + id];
            if (MSD)
                v = mdk::norm2_ref(c0 - o0, c1 - o1, c2 - o2);
            else
                v = mdk::norm2_dot(c0, c1, c2, o0, o1, o2);
        }
        const double b = mdk::block_reduce(v, sm, mdk::wave_sum);
        if (threadIdx.x == 0) a.partial[lag * gridDim.x + blockIdx.x] = b;
    }
}

// sum[lag] += sum over blocks of partial[lag][block], one workgroup per lag.
__global__ __launch_bounds__(kBlock) void k_corr_sum(const double* __restrict__ partial,
                                                     int nblocks, double* __restrict__ sum) {
    __shared__ double sm[kBlock / kWave];
    const long lag = blockIdx.x;
    double v = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += kBlock) v += partial[lag * nblocks + b];
    const double s = mdk::block_reduce(v, sm, mdk::wave_sum);
    if (threadIdx.x == 0) sum[lag] += s;
}

template <class T>
hipError_t grow(T*& p, long& have, long need) {
    if (need <= have) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    have = 0;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), static_cast<size_t>(need) * sizeof(T));
    if (e == hipSuccess) have = need;
    return e;
}

}  // namespace

// ---- handles -------------------------------------------------------------------

struct md_rdf {
    int device = 0;
    std::string err;
    double r_max = 0.0, width = 0.0;
    long bins = 0;
    long frames = 0, natoms = 0;
    double sum_rho = 0.0;
    unsigned long long* hist = nullptr;
    int *cell_of = nullptr, *tmp = nullptr, *atoms = nullptr;
    long atoms_alloc = 0, tmp_alloc = 0, cell_of_alloc = 0;
    int *count = nullptr, *start = nullptr, *fill = nullptr;
    long count_alloc = 0, start_alloc = 0, fill_alloc = 0;
};

struct md_corr {
    bool msd = true;
    int device = 0;
    std::string err;
    long n = 0, max_lag = 0, period = 0, pushed = 0;
    std::vector<long> count;
    double *prev = nullptr, *cur = nullptr, *ring = nullptr, *sum = nullptr, *partial = nullptr;
    long partial_alloc = 0;
};
struct md_msd : md_corr {};
struct md_vacf : md_corr {};

namespace {

md_status fail(std::string& err, md_status st, const std::string& msg) {
    err = msg;
    return st;
}

#define AN_HIP(err, call)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(err, MD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

md_status corr_create(md_corr* c, bool msd, long natoms, long max_lag) {
    c->msd = msd;
    if (natoms <= 0 || max_lag < 0)
        return fail(c->err, MD_ERR_INVALID_ARGUMENT, "natoms must be positive, maxLag >= 0");
    AN_HIP(c->err, hipGetDevice(&c->device));
    c->n = natoms;
    c->max_lag = max_lag;
    c->period = max_lag + 1;
    c->count.assign(static_cast<size_t>(c->period), 0);
    const size_t n3 = 3 * static_cast<size_t>(natoms) * sizeof(double);
    if (msd) AN_HIP(c->err, hipMalloc(reinterpret_cast<void**>(&c->prev), n3));
    AN_HIP(c->err, hipMalloc(reinterpret_cast<void**>(&c->cur), n3));
    AN_HIP(c->err, hipMalloc(reinterpret_cast<void**>(&c->ring), n3 * static_cast<size_t>(c->period)));
    AN_HIP(c->err, hipMalloc(reinterpret_cast<void**>(&c->sum), sizeof(double) * c->period));
    AN_HIP(c->err, hipMemset(c->sum, 0, sizeof(double) * c->period));
    return MD_OK;
}

void corr_release(md_corr* c) {
    for (double* p : {c->prev, c->cur, c->ring, c->sum, c->partial})
        if (p) (void)hipFree(p);
}

md_status corr_push(md_corr* c, md_ctx* ctx) {
    mdk::CtxView v{};
    if (!mdk::ctx_view(ctx, &v)) return fail(c->err, MD_ERR_STATE, "no system uploaded");
    if (v.device != c->device)
        return fail(c->err, MD_ERR_INVALID_ARGUMENT, "the context is on another device");
    if (v.n != c->n)
        return fail(c->err, MD_ERR_INVALID_ARGUMENT,
                    "the system has " + std::to_string(v.n) + " atoms, the correlation " +
                        std::to_string(c->n));
    const int nb = static_cast<int>((c->n + kBlock - 1) / kBlock);
    const long top = std::min(c->pushed, c->max_lag);
    AN_HIP(c->err, grow(c->partial, c->partial_alloc, static_cast<long>(nb) * c->period));
    CorrArgs a{};
    a.n = c->n;
    a.box = v.box;
    a.src = c->msd ? v.pos : v.vel;
    a.slot = v.slot;
    a.first = c->pushed == 0 ? 1 : 0;
    a.ring_slot = c->pushed % c->period;
    a.period = c->period;
    a.top = top;
    a.prev = c->prev;
    a.cur = c->cur;
    a.ring = c->ring;
    a.partial = c->partial;
    if (c->msd)
        hipLaunchKernelGGL(k_corr_push<true>, dim3(nb), dim3(kBlock), 0, v.stream, a);
    else
        hipLaunchKernelGGL(k_corr_push<false>, dim3(nb), dim3(kBlock), 0, v.stream, a);
    hipLaunchKernelGGL(k_corr_sum, dim3(static_cast<unsigned>(top + 1)), dim3(kBlock), 0, v.stream,
                       c->partial, nb, c->sum);
    AN_HIP(c->err, hipGetLastError());
    for (long lag = 0; lag <= top; ++lag) c->count[lag] += c->n;
    ++c->pushed;
    return MD_OK;
}

// sum / count per lag (msd.cpp:54-62, vacf.cpp:42-50).
md_status corr_mean(md_corr* c, std::vector<double>& m) {
    std::vector<double> s(static_cast<size_t>(c->period));
    AN_HIP(c->err, hipDeviceSynchronize());
    AN_HIP(c->err, hipMemcpy(s.data(), c->sum, sizeof(double) * c->period, hipMemcpyDeviceToHost));
    m.assign(static_cast<size_t>(c->period), 0.0);
    for (long k = 0; k < c->period; ++k)
        if (c->count[k] > 0) m[k] = s[k] / static_cast<double>(c->count[k]);
    return MD_OK;
}

}  // namespace

extern "C" {

// ---- Rdf -------------------------------------------------------------------------

md_status md_rdf_create(double r_max, long bins, md_rdf** out) {
    if (out == nullptr) return MD_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!(r_max > 0.0) || bins <= 0 || bins > (1L << 30)) return MD_ERR_INVALID_ARGUMENT;
    md_rdf* r = new (std::nothrow) md_rdf();
    if (r == nullptr) return MD_ERR_RUNTIME;
    r->r_max = r_max;
    r->bins = bins;
    r->width = r_max / static_cast<double>(bins);
    if (hipGetDevice(&r->device) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&r->hist), sizeof(unsigned long long) * bins) !=
            hipSuccess ||
        hipMemset(r->hist, 0, sizeof(unsigned long long) * bins) != hipSuccess) {
        md_rdf_destroy(r);
        return MD_ERR_HIP;
    }
    *out = r;
    return MD_OK;
}

md_status md_rdf_destroy(md_rdf* r) {
    if (r == nullptr) return MD_OK;
    for (void* p : {static_cast<void*>(r->hist), static_cast<void*>(r->cell_of),
                    static_cast<void*>(r->tmp), static_cast<void*>(r->atoms),
                    static_cast<void*>(r->count), static_cast<void*>(r->start),
                    static_cast<void*>(r->fill)})
        if (p) (void)hipFree(p);
    delete r;
    return MD_OK;
}

const char* md_rdf_last_error(const md_rdf* r) { return r ? r->err.c_str() : "null handle"; }

md_status md_rdf_accumulate(md_rdf* r, md_ctx* ctx, int source) {
    if (r == nullptr) return MD_ERR_INVALID_ARGUMENT;
    mdk::CtxView v{};
    if (!mdk::ctx_view(ctx, &v)) return fail(r->err, MD_ERR_STATE, "no system uploaded");
    if (v.device != r->device)
        return fail(r->err, MD_ERR_INVALID_ARGUMENT, "the context is on another device");
    const double smallest = std::min(v.L[0], std::min(v.L[1], v.L[2]));
    if (r->r_max > 0.5 * smallest)
        return fail(r->err, MD_ERR_INVALID_ARGUMENT,
                    "Rdf: rMax exceeds half the smallest box length; minimum-image distances "
                    "beyond that are ambiguous");
    if (source != MD_RDF_AUTO && source != MD_RDF_ALL_PAIRS && source != MD_RDF_CELLS)
        return fail(r->err, MD_ERR_INVALID_ARGUMENT, "unknown pair source");
    // cells a hair wider than rMax, so that rounding in the binning cannot
    // put a pair within rMax two cells apart
    const mdk::Grid g = mdk::cell_grid(v.L, r->r_max * (1.0 + 1e-9));
    if (source == MD_RDF_CELLS && !g.usable)
        return fail(r->err, MD_ERR_INVALID_ARGUMENT,
                    "fewer than 3 cells of width >= rMax fit per side");
    const bool cells = source == MD_RDF_CELLS || (source == MD_RDF_AUTO && g.usable);
    const long n = v.n;
    RdfArgs a{};
    a.n = n;
    a.box = v.box;
    a.r_max = r->r_max;
    a.r_max2_hi = r->r_max * r->r_max * (1.0 + 1e-12);
    a.width = r->width;
    a.bins = static_cast<int>(r->bins);
    a.pos = v.pos;
    a.hist = r->hist;
    // an LDS bin counts at most the workgroup's pair visits: 2^31 for the
    // tile source (by its grid size), 256 n for the cell source
    const bool lds = r->bins <= kLdsBins && (!cells || n * kBlock < kMaxVisitsPerGroup);
    if (cells) {
        const long m = static_cast<long>(g.nc[0]) * g.nc[1] * g.nc[2];
        AN_HIP(r->err, grow(r->cell_of, r->cell_of_alloc, n));
        AN_HIP(r->err, grow(r->tmp, r->tmp_alloc, n));
        AN_HIP(r->err, grow(r->atoms, r->atoms_alloc, n));
        AN_HIP(r->err, grow(r->count, r->count_alloc, m));
        AN_HIP(r->err, grow(r->start, r->start_alloc, m + 1));
        AN_HIP(r->err, grow(r->fill, r->fill_alloc, m));
        mdk::launch_bin_index(n, v.pos, g, r->cell_of, r->count, r->start, r->fill, r->tmp,
                              r->atoms, nullptr, v.stream);
        a.grid = g;
        a.cell_of = r->cell_of;
        a.cell_start = r->start;
        a.cell_atoms = r->atoms;
        const dim3 grid(static_cast<unsigned>((n + kBlock - 1) / kBlock));
        if (lds)
            hipLaunchKernelGGL(k_rdf_cells<true>, grid, dim3(kBlock), 0, v.stream, a);
        else
            hipLaunchKernelGGL(k_rdf_cells<false>, grid, dim3(kBlock), 0, v.stream, a);
    } else {
        a.ntile = (n + kTile - 1) / kTile;
        a.tile_pairs = a.ntile * a.ntile;
        // enough workgroups to fill the chip, and few enough tile pairs per
        // workgroup that an LDS bin (uint32) cannot wrap
        const long per = kMaxVisitsPerGroup / (static_cast<long>(kTile) * kTile);
        long groups = std::min(a.tile_pairs, 16384L);
        groups = std::max(groups, (a.tile_pairs + per - 1) / per);
        if (groups > 0x7fffffffL) return fail(r->err, MD_ERR_CAPACITY, "too many atoms");
        const dim3 grid(static_cast<unsigned>(groups));
        if (lds)
            hipLaunchKernelGGL(k_rdf_tiles<true>, grid, dim3(kBlock), 0, v.stream, a);
        else
            hipLaunchKernelGGL(k_rdf_tiles<false>, grid, dim3(kBlock), 0, v.stream, a);
    }
    AN_HIP(r->err, hipGetLastError());
    r->natoms = n;
    r->sum_rho += static_cast<double>(n) / (v.L[0] * v.L[1] * v.L[2]);
    ++r->frames;
    return MD_OK;
}

md_status md_rdf_result(md_rdf* r, double* centres, double* g, uint64_t* hist, long* frames) {
    if (r == nullptr) return MD_ERR_INVALID_ARGUMENT;
    std::vector<unsigned long long> h(static_cast<size_t>(r->bins));
    AN_HIP(r->err, hipDeviceSynchronize());
    AN_HIP(r->err, hipMemcpy(h.data(), r->hist, sizeof(unsigned long long) * r->bins,
                             hipMemcpyDeviceToHost));
    const double pi = 3.14159265358979323846;
    const double rho = r->frames > 0 ? r->sum_rho / static_cast<double>(r->frames) : 0.0;
    for (long k = 0; k < r->bins; ++k) {
        if (centres) centres[k] = (static_cast<double>(k) + 0.5) * r->width;
        if (hist) hist[k] = h[k];
        if (g) {
            if (r->frames == 0) {
                g[k] = 0.0;
                continue;
            }
            const double lo = static_cast<double>(k) * r->width;
            const double hi = static_cast<double>(k + 1) * r->width;
            const double shell = (4.0 / 3.0) * pi * (hi * hi * hi - lo * lo * lo);
            const double denom = static_cast<double>(r->frames) * 0.5 *
                                 static_cast<double>(r->natoms) * rho * shell;
            g[k] = denom > 0.0 ? static_cast<double>(h[k]) / denom : 0.0;
        }
    }
    if (frames) *frames = r->frames;
    return MD_OK;
}

// ---- Msd / Vacf --------------------------------------------------------------------

md_status md_msd_create(long natoms, long max_lag, md_msd** out) {
    if (out == nullptr) return MD_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    md_msd* c = new (std::nothrow) md_msd();
    if (c == nullptr) return MD_ERR_RUNTIME;
    md_status st = corr_create(c, true, natoms, max_lag);
    if (st != MD_OK) {
        md_msd_destroy(c);
        return st;
    }
    *out = c;
    return MD_OK;
}

md_status md_msd_destroy(md_msd* c) {
    if (c == nullptr) return MD_OK;
    corr_release(c);
    delete c;
    return MD_OK;
}

const char* md_msd_last_error(const md_msd* c) { return c ? c->err.c_str() : "null handle"; }

md_status md_msd_push(md_msd* c, md_ctx* ctx) {
    return c == nullptr ? MD_ERR_INVALID_ARGUMENT : corr_push(c, ctx);
}

md_status md_msd_result(md_msd* c, double* msd) {
    if (c == nullptr || msd == nullptr) return MD_ERR_INVALID_ARGUMENT;
    std::vector<double> m;
    md_status st = corr_mean(c, m);
    if (st != MD_OK) return st;
    std::copy(m.begin(), m.end(), msd);
    return MD_OK;
}

// Msd::diffusionCoefficient (msd.cpp:64-92), host arithmetic on msd().
md_status md_msd_diffusion(md_msd* c, double dt, long fit_from, long fit_to, double* D) {
    if (c == nullptr || D == nullptr || fit_from < 0 || fit_to < 0) return MD_ERR_INVALID_ARGUMENT;
    std::vector<double> m;
    md_status st = corr_mean(c, m);
    if (st != MD_OK) return st;
    const long upper = std::min(fit_to, c->period);
    double st_ = 0.0, sm = 0.0, stt = 0.0, stm = 0.0;
    long cnt = 0;
    for (long k = fit_from; k < upper; ++k) {
        const double t = static_cast<double>(k) * dt;
        st_ += t;
        sm += m[k];
        stt += t * t;
        stm += t * m[k];
        ++cnt;
    }
    *D = 0.0;
    if (cnt < 2) return MD_OK;
    const double nd = static_cast<double>(cnt);
    const double den = nd * stt - st_ * st_;
    if (den == 0.0) return MD_OK;
    *D = (nd * stm - st_ * sm) / den / 6.0;
    return MD_OK;
}

md_status md_vacf_create(long natoms, long max_lag, md_vacf** out) {
    if (out == nullptr) return MD_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    md_vacf* c = new (std::nothrow) md_vacf();
    if (c == nullptr) return MD_ERR_RUNTIME;
    md_status st = corr_create(c, false, natoms, max_lag);
    if (st != MD_OK) {
        md_vacf_destroy(c);
        return st;
    }
    *out = c;
    return MD_OK;
}

md_status md_vacf_destroy(md_vacf* c) {
    if (c == nullptr) return MD_OK;
    corr_release(c);
    delete c;
    return MD_OK;
}

const char* md_vacf_last_error(const md_vacf* c) { return c ? c->err.c_str() : "null handle"; }

md_status md_vacf_push(md_vacf* c, md_ctx* ctx) {
    return c == nullptr ? MD_ERR_INVALID_ARGUMENT : corr_push(c, ctx);
}

// Vacf::vacf (vacf.cpp:52-63): normalised to C(0); unnormalised when
// `normalised` is 0 (vacf.cpp:42-50).
md_status md_vacf_result(md_vacf* c, int normalised, double* vacf) {
    if (c == nullptr || vacf == nullptr) return MD_ERR_INVALID_ARGUMENT;
    std::vector<double> u;
    md_status st = corr_mean(c, u);
    if (st != MD_OK) return st;
    if (!normalised) {
        std::copy(u.begin(), u.end(), vacf);
        return MD_OK;
    }
    const bool zero = u.empty() || u[0] == 0.0;
    for (size_t k = 0; k < u.size(); ++k) vacf[k] = zero ? 0.0 : u[k] / u[0];
    return MD_OK;
}

// Vacf::diffusionCoefficient (vacf.cpp:65-85).
md_status md_vacf_diffusion(md_vacf* c, double dt, double* D) {
    if (c == nullptr || D == nullptr) return MD_ERR_INVALID_ARGUMENT;
    std::vector<double> u;
    md_status st = corr_mean(c, u);
    if (st != MD_OK) return st;
    long valid = 0;
    for (long k = 0; k < c->period; ++k) {
        if (c->count[k] == 0) break;
        valid = k + 1;
    }
    *D = 0.0;
    if (valid < 2) return MD_OK;
    double integral = 0.0;
    for (long k = 0; k + 1 < valid; ++k) integral += 0.5 * (u[k] + u[k + 1]) * dt;
    *D = integral / 3.0;
    return MD_OK;
}

}  // extern "C"
