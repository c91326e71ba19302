// Staging of a tile's candidate cells (see Tiles in md_kernels.hpp), shared
// by the tiled list build and the tiled pair kernel so both see the same
// staged order.
#pragma once

#include "md_common.hpp"
#include "md_kernels.hpp"

namespace mdk {

constexpr int kTileCells = 9 * (kTileSegMax + 2);

struct TileStage {
    int x0, x1, y, z;
    int s2;                    // staged cells per row: x1 - x0 + 2
    int nq;                    // staged cells: 9 s2
    int cell[kTileCells];      // staged cell q = r s2 + j (row r = (dz+1) 3 + (dy+1))
    int off[kTileCells + 1];   // first staged index of cell q; off[nq] = total
};

__device__ __forceinline__ int wave_prefix_incl(int x, int lane) {
    for (int off = 1; off < kWave; off <<= 1) {
        const int y = __shfl_up(x, off, kWave);
        if (lane >= off) x += y;
    }
    return x;
}

// Fills ts for tile `tile` (every thread of the workgroup calls it; the
// table is valid after the call).
__device__ inline void tile_stage_table(TileStage& ts, const Grid& g, int nseg, const int* cell_start,
                                 long tile) {
    const int nx = g.nc[0], ny = g.nc[1], nz = g.nc[2];
    if (threadIdx.x == 0) {
        const int k = static_cast<int>(tile % nseg);
        ts.y = static_cast<int>((tile / nseg) % ny);
        ts.z = static_cast<int>(tile / (static_cast<long>(nseg) * ny));
        ts.x0 = k * nx / nseg;
        ts.x1 = (k + 1) * nx / nseg;
        ts.s2 = ts.x1 - ts.x0 + 2;
        ts.nq = 9 * ts.s2;
    }
    __syncthreads();
    if (threadIdx.x < kWave) {
        const int lane = threadIdx.x;
        int cells[3], cnts[3], local = 0;
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            const int q = lane * 3 + u;
            int cell = 0, cnt = 0;
            if (q < ts.nq) {
                const int r = q / ts.s2, j = q - r * ts.s2;
                const int yy = (ts.y + (r % 3) - 1 + ny) % ny;
                const int zz = (ts.z + (r / 3) - 1 + nz) % nz;
                const int xx = (ts.x0 - 1 + j + nx) % nx;
                cell = xx + nx * (yy + ny * zz);
                cnt = cell_start[cell + 1] - cell_start[cell];
            }
            cells[u] = cell;
            cnts[u] = cnt;
            local += cnt;
        }
        const int incl = wave_prefix_incl(local, lane);
        int run = incl - local;
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            const int q = lane * 3 + u;
            if (q < ts.nq) {
                ts.cell[q] = cells[u];
                ts.off[q] = run;
            }
            run += cnts[u];
        }
        if (lane == kWave - 1) ts.off[ts.nq] = incl;
    }
    __syncthreads();
}

// Global slot of staged element m (off[q] <= m < off[q+1] by binary search).
__device__ __forceinline__ int staged_slot(const TileStage& ts, const int* cell_start, int m) {
    int lo = 0, hi = ts.nq;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (ts.off[mid] <= m) lo = mid;
        else hi = mid;
    }
    return cell_start[ts.cell[lo]] + (m - ts.off[lo]);
}

}  // namespace mdk
