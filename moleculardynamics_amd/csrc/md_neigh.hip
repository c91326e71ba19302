// Neighbour-list construction on the device: cell binning, a deterministic
// counting sort of the atom order by (cell, upload index), and the full
// Verlet list with the reference's exact pair test.
//
// Semantics reproduced (neighbor/verlet_list.cpp:8-42, cell_list.cpp:8-68):
// the list holds every unordered pair with
//   norm2(minimumImage(r_i - r_j)) < (rc + skin)^2,  i < j by upload index,
// evaluated with the reference's FP64 operation order, so the pair SET is
// bit-identical to the reference's. Candidates come from the 27 cells around
// an atom's cell (cells at least rc + skin wide, at least 3 per side), or from
// every atom when the grid is unusable (the reference's brute-force fallback).
//
// Every kernel here takes the device rebuild flag and returns at once when it
// is 0, so a multi-step run can launch the rebuild unconditionally and keep
// the decision on the device.
#include "md_common.hpp"
#include "md_kernels.hpp"

namespace mdk {

namespace {

__device__ __forceinline__ int cell_coord(double r, double L, double w, int n) {
    // floor((wrap(r) + L/2) / width) clamped to [0, n-1] (cell_list.cpp:31-38);
    // wrap is the reference fold.
#pragma clang fp contract(off)
    const double f = fold_ref(r, L);
    int k = static_cast<int>(floor((f + 0.5 * L) / w));
    return k < 0 ? 0 : (k >= n ? n - 1 : k);
}

__global__ void k_fill_int(int* a, long n, int v, const int* flag) {
    if (flag != nullptr && *flag == 0) return;
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_bin(const double4* __restrict__ pos, long n, Grid g, int* __restrict__ cell_of,
                      int* __restrict__ cell_count, const int* flag) {
    if (*flag == 0) return;
    const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int c = 0;
    if (g.usable) {
        const double4 p = pos[t];
        const int cx = cell_coord(p.x, g.L[0], g.w[0], g.nc[0]);
        const int cy = cell_coord(p.y, g.L[1], g.w[1], g.nc[1]);
        const int cz = cell_coord(p.z, g.L[2], g.w[2], g.nc[2]);
        c = cx + g.nc[0] * (cy + g.nc[1] * cz);
    }
    cell_of[t] = c;
    atomicAdd(&cell_count[c], 1);
}

// Exclusive scan of cell_count into cell_start (and the running fill
// cursor), one workgroup of 1024 threads, each owning a contiguous chunk.
__global__ __launch_bounds__(1024) void k_scan(const int* __restrict__ cell_count, long ncell,
                                               int* __restrict__ cell_start,
                                               int* __restrict__ cell_fill, const int* flag) {
    if (*flag == 0) return;
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const long chunk = (ncell + 1023) / 1024;
    const long b = t * chunk;
    const long e = b + chunk < ncell ? b + chunk : ncell;
    int s = 0;
    for (long c = b; c < e; ++c) s += cell_count[c];
    part[t] = s;
    __syncthreads();
    // Hillis-Steele inclusive scan over 1024 partials.
    for (int off = 1; off < 1024; off <<= 1) {
        const int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = t > 0 ? part[t - 1] : 0;
    for (long c = b; c < e; ++c) {
        cell_start[c] = run;
        cell_fill[c] = run;
        run += cell_count[c];
    }
    if (t == 1023) cell_start[ncell] = part[1023];
}

__global__ void k_place(const int* __restrict__ cell_of, long n, int* __restrict__ cell_fill,
                        int* __restrict__ order, const int* flag) {
    if (*flag == 0) return;
    const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int at = atomicAdd(&cell_fill[cell_of[t]], 1);
    order[at] = static_cast<int>(t);
}

// Within each cell, order the slots by upload index (insertion sort; cells
// hold a few tens of atoms). Makes the new order a pure function of the
// positions, independent of atomic arrival order.
__global__ void k_sort_cells(const int* __restrict__ cell_start, long ncell,
                             const int* __restrict__ orig, int* __restrict__ order,
                             const int* flag) {
    if (*flag == 0) return;
    const long c = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int b = cell_start[c], e = cell_start[c + 1];
    for (int i = b + 1; i < e; ++i) {
        const int v = order[i];
        const int key = orig[v];
        int j = i - 1;
        while (j >= b && orig[order[j]] > key) {
            order[j + 1] = order[j];
            --j;
        }
        order[j + 1] = v;
    }
}

__global__ void k_permute(const int* __restrict__ order, long n, const double4* __restrict__ pos,
                          const double4* __restrict__ vel, const double4* __restrict__ frc,
                          const int* __restrict__ orig, const int* __restrict__ cell_of,
                          double4* __restrict__ pos2, double4* __restrict__ vel2,
                          double4* __restrict__ frc2, int* __restrict__ orig2,
                          int* __restrict__ cell2, const int* flag) {
    if (*flag == 0) return;
    const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int s = order[t];
    pos2[t] = pos[s];
    vel2[t] = vel[s];
    frc2[t] = frc[s];
    orig2[t] = orig[s];
    cell2[t] = cell_of[s];
}

__global__ void k_commit(long n, const double4* __restrict__ pos2, const double4* __restrict__ vel2,
                         const double4* __restrict__ frc2, const int* __restrict__ orig2,
                         const int* __restrict__ cell2, double4* __restrict__ pos,
                         double4* __restrict__ vel, double4* __restrict__ frc,
                         int* __restrict__ orig, int* __restrict__ slot,
                         int* __restrict__ cell_of, double4* __restrict__ ref, const int* flag) {
    if (*flag == 0) return;
    const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double4 p = pos2[t];
    pos[t] = p;
    ref[t] = p;
    vel[t] = vel2[t];
    frc[t] = frc2[t];
    const int o = orig2[t];
    orig[t] = o;
    slot[o] = static_cast<int>(t);
    cell_of[t] = cell2[t];
}

// One thread per slot: scan the 27 surrounding cells (or all atoms when the
// grid is unusable) and keep every partner within the list radius. The test
// is the reference's exact FP64 expression on the pair oriented by upload
// index, so both rows of a pair see the same bits.
__global__ __launch_bounds__(kBlock) void k_build(const double4* __restrict__ pos,
                                                  const int* __restrict__ orig, long n, Grid g,
                                                  Box box, double rl2,
                                                  const int* __restrict__ cell_of,
                                                  const int* __restrict__ cell_start,
                                                  uint32_t* __restrict__ nbr,
                                                  int* __restrict__ ncnt, long npad, int cap,
                                                  Scalars* sc, const int* flag) {
    if (*flag == 0) return;
    const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double4 pi = pos[t];
    const int oi = orig[t];
    int cnt = 0;
    auto visit = [&](int j) {
        if (j == t) return;
        const bool flip = oi > orig[j];
        double dx, dy, dz;
        pair_delta(box, pi, pos[j], flip, dx, dy, dz);
        if (norm2_ref(dx, dy, dz) < rl2) {
            if (cnt < cap) nbr[static_cast<long>(cnt) * npad + t] = static_cast<uint32_t>(j) | (flip ? kFlipBit : 0u);
            ++cnt;
        }
    };
    if (g.usable) {
        const int c = cell_of[t];
        const int cx = c % g.nc[0];
        const int cy = (c / g.nc[0]) % g.nc[1];
        const int cz = c / (g.nc[0] * g.nc[1]);
        for (int oz = -1; oz <= 1; ++oz) {
            const int zz = (cz + oz + g.nc[2]) % g.nc[2];
            for (int oy = -1; oy <= 1; ++oy) {
                const int yy = (cy + oy + g.nc[1]) % g.nc[1];
                for (int ox = -1; ox <= 1; ++ox) {
                    const int xx = (cx + ox + g.nc[0]) % g.nc[0];
                    const int nb = xx + g.nc[0] * (yy + g.nc[1] * zz);
                    const int e = cell_start[nb + 1];
                    for (int j = cell_start[nb]; j < e; ++j) visit(j);
                }
            }
        }
    } else {
        for (int j = 0; j < static_cast<int>(n); ++j) visit(j);
    }
    ncnt[t] = cnt;
    if (cnt > cap) atomicOr(&sc->overflow, 1);
}

__global__ void k_count_build(Scalars* sc, const int* flag) {
    if (*flag != 0 && threadIdx.x == 0) sc->builds += 1;
}

inline unsigned grid_of(long n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

}  // namespace

void launch_rebuild(const NeighArgs& a, hipStream_t s) {
    const long n = a.n;
    const long ncell = a.grid.usable ? static_cast<long>(a.grid.nc[0]) * a.grid.nc[1] * a.grid.nc[2] : 1;
    hipLaunchKernelGGL(k_fill_int, dim3(grid_of(ncell)), dim3(kBlock), 0, s, a.cell_count, ncell, 0, a.flag);
    hipLaunchKernelGGL(k_bin, dim3(grid_of(n)), dim3(kBlock), 0, s, a.pos, n, a.grid, a.cell_of,
                       a.cell_count, a.flag);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, a.cell_count, ncell, a.cell_start,
                       a.cell_fill, a.flag);
    hipLaunchKernelGGL(k_place, dim3(grid_of(n)), dim3(kBlock), 0, s, a.cell_of, n, a.cell_fill,
                       a.order, a.flag);
    hipLaunchKernelGGL(k_sort_cells, dim3(grid_of(ncell)), dim3(kBlock), 0, s, a.cell_start, ncell,
                       a.orig, a.order, a.flag);
    hipLaunchKernelGGL(k_permute, dim3(grid_of(n)), dim3(kBlock), 0, s, a.order, n, a.pos, a.vel,
                       a.frc, a.orig, a.cell_of, a.pos2, a.vel2, a.frc2, a.orig2, a.cell2, a.flag);
    hipLaunchKernelGGL(k_commit, dim3(grid_of(n)), dim3(kBlock), 0, s, n, a.pos2, a.vel2, a.frc2,
                       a.orig2, a.cell2, a.pos, a.vel, a.frc, a.orig, a.slot, a.cell_of, a.ref,
                       a.flag);
    hipLaunchKernelGGL(k_build, dim3(grid_of(n)), dim3(kBlock), 0, s, a.pos, a.orig, n, a.grid,
                       a.box, a.rl2, a.cell_of, a.cell_start, a.nbr, a.ncnt, a.npad, a.cap, a.sc,
                       a.flag);
    hipLaunchKernelGGL(k_count_build, dim3(1), dim3(64), 0, s, a.sc, a.flag);
}

}  // namespace mdk
