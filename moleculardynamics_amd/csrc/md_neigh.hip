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
// Six kernels: bin -> scan -> place -> sort-within-cell -> permute ->
// commit+build. Every kernel takes the device rebuild flag and returns at
// once when it is 0, so a multi-step run launches the rebuild
// unconditionally and keeps the decision on the device; the kernels use
// grid-stride loops over a capped grid so a skipped rebuild costs little.
// Invariant: cell_count is all zero between rebuilds (k_scan clears it).
#include "md_common.hpp"
#include "md_kernels.hpp"

namespace mdk {

namespace {

constexpr unsigned kGridCap = 2048;

__device__ __forceinline__ int cell_coord(double r, double L, double w, int n) {
    // floor((wrap(r) + L/2) / width) clamped to [0, n-1] (cell_list.cpp:31-38);
    // wrap is the reference fold.
#pragma clang fp contract(off)
    const double f = fold_ref(r, L);
    int k = static_cast<int>(floor((f + 0.5 * L) / w));
    return k < 0 ? 0 : (k >= n ? n - 1 : k);
}

#define GRID_STRIDE(i, n)                                                              \
    for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < (n); \
         i += static_cast<long>(gridDim.x) * blockDim.x)

__global__ void k_bin(const double4* __restrict__ pos, long n, Grid g, int* __restrict__ cell_of,
                      int* __restrict__ cell_count, const int* flag) {
    if (*flag == 0) return;
    GRID_STRIDE(t, n) {
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
}

// Exclusive scan of cell_count into cell_start (and the fill cursor), one
// workgroup of 1024 threads each owning a contiguous chunk; clears
// cell_count afterwards and counts the build.
__global__ __launch_bounds__(1024) void k_scan(int* __restrict__ cell_count, long ncell,
                                               int* __restrict__ cell_start,
                                               int* __restrict__ cell_fill, Scalars* sc,
                                               const int* flag) {
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
    for (int off = 1; off < 1024; off <<= 1) {
        const int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = t > 0 ? part[t - 1] : 0;
    for (long c = b; c < e; ++c) {
        const int k = cell_count[c];
        cell_start[c] = run;
        cell_fill[c] = run;
        cell_count[c] = 0;
        run += k;
    }
    if (t == 1023) {
        cell_start[ncell] = part[1023];
        sc->builds += 1;
    }
}

__global__ void k_place(const int* __restrict__ cell_of, long n, int* __restrict__ cell_fill,
                        int* __restrict__ order, const int* flag) {
    if (*flag == 0) return;
    GRID_STRIDE(t, n) {
        const int at = atomicAdd(&cell_fill[cell_of[t]], 1);
        order[at] = static_cast<int>(t);
    }
}

// Within each cell, order the slots by upload index (insertion sort; cells
// hold a few tens of atoms). Makes the new order a pure function of the
// positions, independent of atomic arrival order.
__global__ void k_sort_cells(const int* __restrict__ cell_start, long ncell,
                             const int* __restrict__ orig, int* __restrict__ order,
                             const int* flag) {
    if (*flag == 0) return;
    GRID_STRIDE(c, ncell) {
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
}

__global__ void k_permute(const int* __restrict__ order, long n, const double4* __restrict__ pos,
                          const double4* __restrict__ vel, const double4* __restrict__ frc,
                          const int* __restrict__ orig, const int* __restrict__ cell_of,
                          double4* __restrict__ pos2, double4* __restrict__ vel2,
                          double4* __restrict__ frc2, int* __restrict__ orig2,
                          int* __restrict__ cell2, const int* flag) {
    if (*flag == 0) return;
    GRID_STRIDE(t, n) {
        const int s = order[t];
        pos2[t] = pos[s];
        vel2[t] = vel[s];
        frc2[t] = frc[s];
        orig2[t] = orig[s];
        cell2[t] = cell_of[s];
    }
}

// Commit the new order and build row t of the list from the new positions.
// One thread per slot scans the 27 surrounding cells (or all atoms when the
// grid is unusable) and keeps every partner within the list radius. The test
// is the reference's exact FP64 expression on the pair oriented by upload
// index, so both rows of a pair see the same bits.
__global__ __launch_bounds__(kBlock) void k_commit_build(
    long n, Grid g, Box box, double rl2, const double4* __restrict__ pos2,
    const double4* __restrict__ vel2, const double4* __restrict__ frc2,
    const int* __restrict__ orig2, const int* __restrict__ cell2, const int* __restrict__ cell_start,
    double4* __restrict__ pos, double4* __restrict__ vel, double4* __restrict__ frc,
    double4* __restrict__ ref, int* __restrict__ orig, int* __restrict__ slot,
    int* __restrict__ cell_of, uint32_t* __restrict__ nbr, int* __restrict__ ncnt, int cap,
    Scalars* sc, const int* flag) {
    if (*flag == 0) return;
    GRID_STRIDE(t, n) {
        const double4 pi = pos2[t];
        const int oi = orig2[t];
        const int ci = cell2[t];
        pos[t] = pi;
        ref[t] = pi;
        vel[t] = vel2[t];
        frc[t] = frc2[t];
        orig[t] = oi;
        slot[oi] = static_cast<int>(t);
        cell_of[t] = ci;
        int cnt = 0;
        auto visit = [&](int j) {
            if (j == t) return;
            const bool flip = oi > orig2[j];
            double dx, dy, dz;
            pair_delta(box, pi, pos2[j], flip, dx, dy, dz);
            if (norm2_ref(dx, dy, dz) < rl2) {
                if (cnt < cap)
                    nbr[nbr_index(t, cnt, cap)] = static_cast<uint32_t>(j) | (flip ? kFlipBit : 0u);
                ++cnt;
            }
        };
        if (g.usable) {
            const int cx = ci % g.nc[0];
            const int cy = (ci / g.nc[0]) % g.nc[1];
            const int cz = ci / (g.nc[0] * g.nc[1]);
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
}

inline unsigned grid_capped(long n) {
    const long g = (n + kBlock - 1) / kBlock;
    return static_cast<unsigned>(g < static_cast<long>(kGridCap) ? (g > 0 ? g : 1) : kGridCap);
}

}  // namespace

void launch_rebuild(const NeighArgs& a, hipStream_t s) {
    const long n = a.n;
    const long ncell = a.grid.usable ? static_cast<long>(a.grid.nc[0]) * a.grid.nc[1] * a.grid.nc[2] : 1;
    const dim3 gn(grid_capped(n)), gc(grid_capped(ncell)), b(kBlock);
    hipLaunchKernelGGL(k_bin, gn, b, 0, s, a.pos, n, a.grid, a.cell_of, a.cell_count, a.flag);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, a.cell_count, ncell, a.cell_start,
                       a.cell_fill, a.sc, a.flag);
    hipLaunchKernelGGL(k_place, gn, b, 0, s, a.cell_of, n, a.cell_fill, a.order, a.flag);
    hipLaunchKernelGGL(k_sort_cells, gc, b, 0, s, a.cell_start, ncell, a.orig, a.order, a.flag);
    hipLaunchKernelGGL(k_permute, gn, b, 0, s, a.order, n, a.pos, a.vel, a.frc, a.orig, a.cell_of,
                       a.pos2, a.vel2, a.frc2, a.orig2, a.cell2, a.flag);
    hipLaunchKernelGGL(k_commit_build, gn, b, 0, s, n, a.grid, a.box, a.rl2, a.pos2, a.vel2,
                       a.frc2, a.orig2, a.cell2, a.cell_start, a.pos, a.vel, a.frc, a.ref, a.orig,
                       a.slot, a.cell_of, a.nbr, a.ncnt, a.cap, a.sc, a.flag);
}

}  // namespace mdk
