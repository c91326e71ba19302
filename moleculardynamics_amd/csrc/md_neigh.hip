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

// Within each cell, order the slots by upload index, which makes the new
// order a pure function of the positions (independent of atomic arrival
// order). One wave per cell: with up to 64 entries each lane holds one and
// its rank is the number of smaller keys (upload indices are distinct);
// larger cells fall back to an insertion sort by one lane.
__global__ __launch_bounds__(kBlock) void k_sort_cells(const int* __restrict__ cell_start, long ncell,
                                                       const int* __restrict__ orig,
                                                       int* __restrict__ order, const int* flag) {
    if (*flag == 0) return;
    const int lane = threadIdx.x & (kWave - 1);
    const long waves = static_cast<long>(gridDim.x) * (kBlock / kWave);
    for (long c = static_cast<long>(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave; c < ncell;
         c += waves) {
        const int b = cell_start[c], e = cell_start[c + 1];
        const int cnt = e - b;
        if (cnt <= kWave) {
            int v = 0, key = 0x7fffffff;
            if (lane < cnt) {
                v = order[b + lane];
                key = orig[v];
            }
            int rank = 0;
            for (int j = 0; j < cnt; ++j) rank += (__shfl(key, j, kWave) < key) ? 1 : 0;
            if (lane < cnt) order[b + rank] = v;
        } else if (lane == 0) {
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

// Tiled commit + build for a usable grid. A workgroup owns a run of up to
// kSegMax consecutive cells along x in one (y, z) row; after the sort those
// cells' atoms are one contiguous slot range. The candidates of the whole
// run -- cells x0-1 .. x1 in the 9 rows (y+-1, z+-1) -- are staged once in
// LDS (x, y, z, upload index, slot), so every pair test reads LDS instead of
// gathering from L2. The exact test and the orientation are those of
// k_commit_build. A run whose candidates exceed the LDS capacity reads them
// from global memory instead.
constexpr int kSegMax = 16;
constexpr int kTileLdsCap = 3072;                   // staged candidates per run
constexpr int kTileLdsBytes = kTileLdsCap * 32;     // x, y, z (FP64) + upload index + slot

__global__ __launch_bounds__(kBlock) void k_commit_build_tiles(
    long n, Grid g, Box box, double rl2, int seg, int nseg, int lds_cap,
    const double4* __restrict__ pos2, const double4* __restrict__ vel2,
    const double4* __restrict__ frc2, const int* __restrict__ orig2, const int* __restrict__ cell2,
    const int* __restrict__ cell_start, double4* __restrict__ pos, double4* __restrict__ vel,
    double4* __restrict__ frc, double4* __restrict__ ref, int* __restrict__ orig,
    int* __restrict__ slot, int* __restrict__ cell_of, uint32_t* __restrict__ nbr,
    int* __restrict__ ncnt, int cap, Scalars* sc, const int* flag) {
    if (*flag == 0) return;
    extern __shared__ double lds[];
    double* lx = lds;
    double* ly = lx + lds_cap;
    double* lz = ly + lds_cap;
    int* lo = reinterpret_cast<int*>(lz + lds_cap);
    int* lg = lo + lds_cap;
    __shared__ int coff[9][kSegMax + 2], ccnt[9][kSegMax + 2], cbeg[9][kSegMax + 2];
    __shared__ int total;

    const int sgi = blockIdx.x % nseg;
    const int y = (blockIdx.x / nseg) % g.nc[1];
    const int z = blockIdx.x / (nseg * g.nc[1]);
    const int x0 = sgi * seg;
    const int x1 = min(x0 + seg, g.nc[0]);
    const int ncl = x1 - x0 + 2;  // candidate cells per row
    if (x0 >= g.nc[0]) return;

    // candidate cell table (one thread), then the staging copy
    if (threadIdx.x == 0) {
        int off = 0;
        for (int r = 0; r < 9; ++r) {
            const int yy = (y + (r % 3) - 1 + g.nc[1]) % g.nc[1];
            const int zz = (z + (r / 3) - 1 + g.nc[2]) % g.nc[2];
            for (int k = 0; k < ncl; ++k) {
                const int xx = (x0 - 1 + k + g.nc[0]) % g.nc[0];
                const int c = xx + g.nc[0] * (yy + g.nc[1] * zz);
                const int b = cell_start[c], e = cell_start[c + 1];
                coff[r][k] = off;
                ccnt[r][k] = e - b;
                cbeg[r][k] = b;
                off += e - b;
            }
        }
        total = off;
    }
    __syncthreads();
    const bool staged = total <= lds_cap;
    if (staged) {
        for (int r = 0; r < 9; ++r)
            for (int k = 0; k < ncl; ++k) {
                const int b = cbeg[r][k], c = ccnt[r][k], o = coff[r][k];
                for (int m = threadIdx.x; m < c; m += blockDim.x) {
                    const double4 p = pos2[b + m];
                    lx[o + m] = p.x;
                    ly[o + m] = p.y;
                    lz[o + m] = p.z;
                    lo[o + m] = orig2[b + m];
                    lg[o + m] = b + m;
                }
            }
    }
    __syncthreads();

    const int home_cell0 = x0 + g.nc[0] * (y + g.nc[1] * z);
    const int hb = cell_start[home_cell0];
    const int he = cell_start[home_cell0 + (x1 - x0)];
    for (int t = hb + threadIdx.x; t < he; t += blockDim.x) {
        const double4 pi = pos2[t];
        const int oi = orig2[t];
        const int ci = cell2[t];
        pos[t] = pi;
        ref[t] = pi;
        vel[t] = vel2[t];
        frc[t] = frc2[t];
        orig[t] = oi;
        slot[oi] = t;
        cell_of[t] = ci;
        const int kx = (ci % g.nc[0]) - x0;  // home cell index in the run: 0..seg-1
        int cnt = 0;
        for (int r = 0; r < 9; ++r) {
            for (int kk = 0; kk < 3; ++kk) {
                const int k = kx + kk;
                const int c = ccnt[r][k];
                if (staged) {
                    const int o = coff[r][k];
                    for (int m = o; m < o + c; ++m) {
                        const int j = lg[m];
                        if (j == t) continue;
                        const bool flip = oi > lo[m];
                        double dx, dy, dz;
                        pair_delta(box, pi, make_double4(lx[m], ly[m], lz[m], 0.0), flip, dx, dy, dz);
                        if (norm2_ref(dx, dy, dz) < rl2) {
                            if (cnt < cap) nbr[nbr_index(t, cnt, cap)] = static_cast<uint32_t>(j) | (flip ? kFlipBit : 0u);
                            ++cnt;
                        }
                    }
                } else {
                    const int b = cbeg[r][k];
                    for (int j = b; j < b + c; ++j) {
                        if (j == t) continue;
                        const bool flip = oi > orig2[j];
                        double dx, dy, dz;
                        pair_delta(box, pi, pos2[j], flip, dx, dy, dz);
                        if (norm2_ref(dx, dy, dz) < rl2) {
                            if (cnt < cap) nbr[nbr_index(t, cnt, cap)] = static_cast<uint32_t>(j) | (flip ? kFlipBit : 0u);
                            ++cnt;
                        }
                    }
                }
            }
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
    if (a.grid.usable) {
        const int nseg = (a.grid.nc[0] + 11) / 12;
        const int seg = (a.grid.nc[0] + nseg - 1) / nseg;
        const unsigned nwg = static_cast<unsigned>(nseg) * a.grid.nc[1] * a.grid.nc[2];
        hipLaunchKernelGGL(k_commit_build_tiles, dim3(nwg), b, kTileLdsBytes, s, n, a.grid, a.box,
                           a.rl2, seg, nseg, kTileLdsCap, a.pos2, a.vel2, a.frc2, a.orig2, a.cell2,
                           a.cell_start, a.pos, a.vel, a.frc, a.ref, a.orig, a.slot, a.cell_of,
                           a.nbr, a.ncnt, a.cap, a.sc, a.flag);
    } else {
        hipLaunchKernelGGL(k_commit_build, gn, b, 0, s, n, a.grid, a.box, a.rl2, a.pos2, a.vel2,
                           a.frc2, a.orig2, a.cell2, a.cell_start, a.pos, a.vel, a.frc, a.ref, a.orig,
                           a.slot, a.cell_of, a.nbr, a.ncnt, a.cap, a.sc, a.flag);
    }
}

hipError_t neigh_init() {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(k_commit_build_tiles),
                               hipFuncAttributeMaxDynamicSharedMemorySize, kTileLdsBytes);
}

}  // namespace mdk
