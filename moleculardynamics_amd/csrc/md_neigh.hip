// Neighbour-list construction on the device: cell binning, a deterministic
// counting sort of the atom order by (cell, upload index), and the full
// Verlet list with the reference's exact pair test.
//
// Semantics reproduced (neighbor/verlet_list.cpp:8-42, cell_list.cpp:8-68):
// the list holds every unordered pair with
//   norm2(minimumImage(r_i - r_j)) < (rc + skin)^2,  i < j by upload index,
// evaluated with the reference's FP64 operation order, so the pair SET is
// bit-identical to the reference's. Candidates come from the 5 x 5 x 5 cells
// around an atom's cell (cells at least (rc + skin)/2 wide, at least 5 per
// side), or from every atom when the grid is unusable; every pair within
// rc + skin lies in those cells.
//
// Six kernels: bin -> tile count -> scan -> place -> sort-within-cell and
// permute -> commit+build. Every kernel takes the device rebuild flag and
// returns at once when it is 0, so a multi-step run launches the rebuild
// unconditionally and keeps the decision on the device. The O(N) passes use
// grid-stride loops over a capped grid so a skipped rebuild costs little;
// commit+build runs one thread per slot in the XCD-aware block order (eight
// lanes per slot for systems of at most 32,768 atoms).
// Invariant: cell_count is all zero between rebuilds (k_scan clears it).
#include "md_common.hpp"
#include "md_kernels.hpp"

namespace mdk {

namespace {

constexpr unsigned kGridCap = 2048;

// Cell coordinate of the folded position, floor((wrap(r) + L/2) nc / L)
// clamped to [0, n-1] (the rule of cell_list.cpp:31-38 with the division
// replaced by a product). The list grid's cells are at least
// (rc + skin)/2 (1 + 1e-9) wide, a margin ten million times the rounding of
// this product, so every pair within rc + skin still lies within two cells
// per axis and the 5 x 5 x 5 stencil finds it. fold is the exact fast fold.
__device__ __forceinline__ int cell_coord(double r, const Box& b, int d, double inv_w, int n) {
    const double f = fold(r, b.L[d], b.lo[d], b.hi1[d], b.hi2[d]);
    const int k = static_cast<int>(floor(fma(0.5, b.L[d], f) * inv_w));
    return k < 0 ? 0 : (k >= n ? n - 1 : k);
}

// Whole waves stride over the atoms together (t = base + lane, the active
// lanes a prefix of the wave), so lanes can cooperate.
#define WAVE_STRIDE(base, n)                                                                  \
    for (long base = (static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x) & ~(kWave - 1L); \
         base < (n); base += static_cast<long>(gridDim.x) * blockDim.x)

// Runs of equal keys in consecutive lanes. The atoms are in the previous
// build's cell order, so a wave covers a few cells and its lanes form a few
// runs; one atomic per run (by its first lane, with the run length) replaces
// one per lane on the same counter. rank is the lane's place in its run and
// head the run's first lane.
struct Run {
    int len, rank, head;
    bool first;
};

__device__ __forceinline__ Run wave_run(int key, bool active) {
    const int lane = threadIdx.x & (kWave - 1);
    const int prev = __shfl_up(key, 1, kWave);
    const bool first = active && (lane == 0 || prev != key);
    const unsigned long long heads = __ballot(first);
    const int nact = __popcll(__ballot(active));
    const unsigned long long upto = lane == kWave - 1 ? ~0ull : ((2ull << lane) - 1);
    Run r;
    r.first = first;
    r.head = 63 - __clzll(heads & upto);
    r.rank = lane - r.head;
    const unsigned long long after = heads & ~upto;
    r.len = (after ? __ffsll(static_cast<long long>(after)) - 1 : nact) - lane;
    return r;
}

__global__ void k_bin(const double4* __restrict__ pos, long n, Grid g, Box box,
                      int* __restrict__ cell_of, int* __restrict__ cell_count, const int* flag) {
    if (*flag == 0) return;
    WAVE_STRIDE(base, n) {
        const long t = base + (threadIdx.x & (kWave - 1));
        const bool active = t < n;
        int c = 0;
        if (active && g.usable) {
            const double4 p = pos[t];
            const int cx = cell_coord(p.x, box, 0, g.inv_w[0], g.nc[0]);
            const int cy = cell_coord(p.y, box, 1, g.inv_w[1], g.nc[1]);
            const int cz = cell_coord(p.z, box, 2, g.inv_w[2], g.nc[2]);
            c = cx + g.nc[0] * (cy + g.nc[1] * cz);
        }
        const Run r = wave_run(c, active);
        if (active) cell_of[t] = c;
        if (r.first) atomicAdd(&cell_count[c], r.len);
    }
}

// Exclusive scan of cell_count into cell_start (and the fill cursor) in two
// passes over tiles of kScanTile cells: k_tile_count sums each tile, then
// k_scan gives every tile its offset (the sum of the tile sums before it) and
// scans inside the tile with wave prefix sums. k_scan clears cell_count and
// its last tile writes the total and counts the build.
constexpr int kScanTile = 1024;

__device__ __forceinline__ int wave_prefix(int x, int lane) {
    for (int off = 1; off < kWave; off <<= 1) {
        const int y = __shfl_up(x, off, kWave);
        if (lane >= off) x += y;
    }
    return x;  // inclusive
}

__global__ __launch_bounds__(kScanTile) void k_tile_count(const int* __restrict__ cell_count,
                                                          long ncell, int* __restrict__ tile_sum,
                                                          const int* flag) {
    if (*flag == 0) return;
    __shared__ int ws[kScanTile / kWave];
    const long c = static_cast<long>(blockIdx.x) * kScanTile + threadIdx.x;
    int v = c < ncell ? cell_count[c] : 0;
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) ws[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int k = 0; k < kScanTile / kWave; ++k) s += ws[k];
        tile_sum[blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(kScanTile) void k_scan(int* __restrict__ cell_count, long ncell,
                                                    const int* __restrict__ tile_sum,
                                                    int* __restrict__ cell_start,
                                                    int* __restrict__ cell_fill, Scalars* sc,
                                                    const int* flag) {
    if (*flag == 0) return;
    __shared__ int ws[kScanTile / kWave];
    __shared__ int base_s;
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    if (wv == 0) {
        int s = 0;
        for (int k = lane; k < static_cast<int>(blockIdx.x); k += kWave) s += tile_sum[k];
        for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, kWave);
        if (lane == 0) base_s = s;
    }
    const long c = static_cast<long>(blockIdx.x) * kScanTile + threadIdx.x;
    const int k = c < ncell ? cell_count[c] : 0;
    const int x = wave_prefix(k, lane);
    if (lane == kWave - 1) ws[wv] = x;
    __syncthreads();
    int before = base_s + x - k;
    for (int w = 0; w < wv; ++w) before += ws[w];
    if (c < ncell) {
        cell_start[c] = before;
        cell_fill[c] = before;
        cell_count[c] = 0;
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanTile - 1) {
        cell_start[ncell] = before + k;
        sc->builds += 1;
    }
}

// Halo table of brick b (one workgroup): the exclusive prefix of the atom
// counts of its halo cells in z, y, x order, and the total, into
// tab[tab_off[b] ..]. A halo larger than the workgroup's LDS rows sets bit 2
// of the overflow flag (md_run and the list build then grow the rows or
// turn the bricks off and build again).
__device__ void brick_table(const Grid& g, const Bricks& k, const int* __restrict__ cell_start,
                            int b, Scalars* sc) {
    __shared__ int ws[kBlock / kWave];
    const BrickGeo o = brick_geo(k, g, b);
    const int hx = o.ex + 4, hy = o.ey + 4, hz = o.ez + 4;
    const int H = hx * hy * hz;
    const int per = (H + kBlock - 1) / kBlock;
    const int h0 = min(H, static_cast<int>(threadIdx.x) * per), h1 = min(H, h0 + per);
    auto cell_at = [&](int h) {
        const int ix = h % hx, iy = (h / hx) % hy, iz = h / (hx * hy);
        int x = o.x0 - 2 + ix, y = o.y0 - 2 + iy, z = o.z0 - 2 + iz;
        x += x < 0 ? g.nc[0] : (x >= g.nc[0] ? -g.nc[0] : 0);
        y += y < 0 ? g.nc[1] : (y >= g.nc[1] ? -g.nc[1] : 0);
        z += z < 0 ? g.nc[2] : (z >= g.nc[2] ? -g.nc[2] : 0);
        return x + g.nc[0] * (y + g.nc[1] * z);
    };
    int s = 0;
    for (int h = h0; h < h1; ++h) {
        const int c = cell_at(h);
        s += cell_start[c + 1] - cell_start[c];
    }
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int incl = wave_prefix(s, lane);
    if (lane == kWave - 1) ws[wv] = incl;
    __syncthreads();
    int run = incl - s;
    for (int w = 0; w < wv; ++w) run += ws[w];
    int* t = k.tab + k.tab_off[b];
    for (int h = h0; h < h1; ++h) {
        t[h] = run;
        const int c = cell_at(h);
        run += cell_start[c + 1] - cell_start[c];
    }
    if (threadIdx.x == kBlock - 1) {
        t[H] = run;
        if (run > k.halo_cap) atomicOr(&sc->overflow, 2);
    }
}

// Places every atom's upload index in its cell's range of `order` (in the
// arrival order of the atomics; k_sort_permute ranks them). With bricks on,
// workgroups from gridDim.x - nb on build the bricks' halo tables instead
// (they need only the cell starts).
__global__ void k_place(const int* __restrict__ cell_of, const int* __restrict__ orig, long n,
                        int* __restrict__ cell_fill, int* __restrict__ order, const int* flag,
                        Grid grid, Bricks bk, const int* __restrict__ cell_start, Scalars* sc) {
    if (*flag == 0) return;
    const unsigned place_groups = gridDim.x - (bk.on ? bk.nb : 0);
    if (blockIdx.x >= place_groups) {
        brick_table(grid, bk, cell_start, static_cast<int>(blockIdx.x - place_groups), sc);
        return;
    }
    for (long base = (static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x) & ~(kWave - 1L);
         base < n; base += static_cast<long>(place_groups) * blockDim.x) {
        const long t = base + (threadIdx.x & (kWave - 1));
        const bool active = t < n;
        const int c = active ? cell_of[t] : 0;
        const Run r = wave_run(c, active);
        int at = 0;
        if (r.first) at = atomicAdd(&cell_fill[c], r.len);
        at = __shfl(at, r.head, kWave) + r.rank;
        if (active) order[at] = orig[t];
    }
}

// Within each cell, order the slots by upload index, which makes the new
// order a pure function of the positions (independent of atomic arrival
// order), and gather the state into that order. One thread per atom: its
// rank is the number of atoms of its cell with a smaller upload index (the
// cells, at least (rc + skin)/2 wide, hold a few atoms each). With an
// unusable grid every atom is in cell 0 and its rank is its upload index.
__device__ __forceinline__ void gather_slot(const NeighArgs& a, int dst, long v, int key, int c) {
    const double4 p = a.pos[v];
    a.pos2[dst] = p;
    a.vel2[dst] = a.vel[v];
    a.frc2[dst] = a.frc[v];
    a.orig2[dst] = key;
    a.cell2[dst] = c;
    if (a.qpos2 != nullptr) a.qpos2[dst] = quantize(a.quant, p, key);
}

__global__ __launch_bounds__(kBlock) void k_sort_permute(NeighArgs a) {
    if (*a.flag == 0) return;
    for (long t = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x; t < a.n;
         t += static_cast<long>(gridDim.x) * kBlock) {
        const int key = a.orig[t];
        if (!a.grid.usable) {
            gather_slot(a, key, t, key, 0);
            continue;
        }
        const int c = a.cell_of[t];
        const int b = a.cell_start[c], e = a.cell_start[c + 1];
        int rank = 0;
        for (int k = b; k < e; ++k) rank += (a.order[k] < key) ? 1 : 0;
        gather_slot(a, b + rank, t, key, c);
    }
}

// Commit the new order and build row t of the list from the new positions.
// One thread per slot scans the 5 x 5 x 5 surrounding cells (or all atoms
// when the grid is unusable) and keeps every partner within the list radius.
// The test is the reference's exact FP64 expression on the pair oriented by
// upload index, so both rows of a pair see the same bits.
//
// After the sort, five x-adjacent cells of one (y, z) row are one contiguous
// slot range unless the row wraps at x (then two ranges), so the cells are
// scanned as 25 ranges in the order (z, y) and slot order within a range.
// Candidates are read four at a time -- positions and
// upload indices of the four issued together -- so every thread keeps four
// gathers in flight instead of one dependent load chain per candidate.
//
// Q: the candidates are read as fixed-point positions (Quant): one 16-byte
// load carries a candidate's position and upload index, and the fixed-point
// r2 decides every candidate outside a band of relative half-width band_l
// around rl2. A candidate inside the band is only counted; a row with such a
// candidate (rare) is then built again with the exact FP64 test above, from
// the FP64 positions. The pair set is therefore the same. CUBIC: the box has one
// unit on every axis, so r2 is compared in units squared (the integer
// differences converted exactly, no scaling per axis).
//
// BRICK: entries are the partner's index in the halo of the row atom's brick
// (the halo's LDS order, see Bricks). Each candidate range is a run of
// consecutive cells in one halo row, so its halo indices are consecutive
// too: entry = j + (halo index of the range's first atom - j of it).
template <bool Q, bool CUBIC, bool BRICK>
__global__ __launch_bounds__(kBlock) void k_commit_build(NeighArgs a) {
    if (*a.flag == 0) return;
    const long n = a.n;
    const Grid& g = a.grid;
    const Box& box = a.box;
    const double rl2 = a.rl2;
    const double4* __restrict__ pos2 = a.pos2;
    const int* __restrict__ orig2 = a.orig2;
    const int* __restrict__ cell_start = a.cell_start;
    const int cap = a.cap;
    const long tl = xcd_block() * kBlock + threadIdx.x;
    if (tl >= n) return;
    const int t = static_cast<int>(tl);
    const double4 pi = pos2[t];
    const int oi = orig2[t];
    const int ci = a.cell2[t];
    a.pos[t] = pi;
    a.ref[t] = pi;
    a.vel[t] = a.vel2[t];
    a.frc[t] = a.frc2[t];
    a.orig[t] = oi;
    a.slot[oi] = t;
    a.cell_of[t] = ci;
    int4 qi{};
    if (Q) {
        qi = a.qpos2[t];
        a.qpos[t] = qi;
    }
    // Entries go to a 16-slot ring in this thread's LDS column without a
    // branch: every candidate writes the next free slot and only a kept one
    // advances the count. Whole octets leave as 32-byte writes once eight are
    // pending, checked after every group of up to four candidates, so at most
    // 11 are pending and the ring never overwrites one.
    __shared__ uint32_t ring[16][kBlock];
    const int tid = threadIdx.x;
    uint32_t* row = a.nbr + oct_index(t, 0, cap);
    int cnt = 0, done = 0;  // entries kept / entries written out (a multiple of 8)
    auto write_octet = [&]() {
        const int s = done & 15;
        if (done < cap) {
            uint4* dst = reinterpret_cast<uint4*>(row + static_cast<long>(done >> 3) * 512);
            dst[0] = make_uint4(ring[s][tid], ring[s + 1][tid], ring[s + 2][tid], ring[s + 3][tid]);
            dst[1] = make_uint4(ring[s + 4][tid], ring[s + 5][tid], ring[s + 6][tid],
                                ring[s + 7][tid]);
        }
        done += 8;
    };
    auto drain = [&]() {
        if (cnt - done >= 8) write_octet();
    };
    int delta = 0;  // BRICK: halo index minus slot over the current range
    auto put = [&](int j, bool keep) {
        ring[cnt & 15][tid] = BRICK ? static_cast<uint32_t>(j + delta) : entry_of(j);
        cnt += (keep && j != t) ? 1 : 0;
    };
    auto exact = [&](const double4& q) {
        double dx, dy, dz;
        row_delta_w(box, pi, q, dx, dy, dz);
        return norm2_ref(dx, dy, dz) < rl2;
    };
    const double u2 = CUBIC ? a.quant.unit[0] * a.quant.unit[0] : 1.0;
    unsigned int hits = 0;  // fixed-point decisions inside the band
    const double lo = rl2 * (1.0 - a.quant.band_l) / u2, hi = rl2 * (1.0 + a.quant.band_l) / u2;
    auto visit_q = [&](int j, const int4& qj) {
        double r2;
        if (CUBIC) {
            const double ux = static_cast<double>(static_cast<int>(static_cast<uint32_t>(qi.x) - static_cast<uint32_t>(qj.x)));
            const double uy = static_cast<double>(static_cast<int>(static_cast<uint32_t>(qi.y) - static_cast<uint32_t>(qj.y)));
            const double uz = static_cast<double>(static_cast<int>(static_cast<uint32_t>(qi.z) - static_cast<uint32_t>(qj.z)));
            r2 = ux * ux + uy * uy + uz * uz;
        } else {
            double dx, dy, dz;
            qdelta(a.quant, qi, qj, dx, dy, dz);
            r2 = dx * dx + dy * dy + dz * dz;
        }
        // no branch per candidate: a decision inside the band is only counted,
        // and the row is then built again with the exact test
        hits += (r2 >= lo && r2 < hi) ? 1u : 0u;
        put(j, r2 < lo);
    };
    // candidate loads with 32-bit byte offsets from a scalar base (n < 2^28
    // wherever fixed-point positions are kept)
    const char* qb = reinterpret_cast<const char*>(a.qpos2);
    auto qat = [&](int j) {
        return *reinterpret_cast<const int4*>(qb + (static_cast<uint32_t>(j) << 4));
    };
    auto range = [&](int b, int e, bool fixed) {
        int j = b;
        if (Q && fixed) {
            for (; j + 4 <= e; j += 4) {
                const int4 q0 = qat(j), q1 = qat(j + 1), q2 = qat(j + 2), q3 = qat(j + 3);
                visit_q(j, q0);
                visit_q(j + 1, q1);
                visit_q(j + 2, q2);
                visit_q(j + 3, q3);
                drain();
            }
            for (; j < e; ++j) {
                visit_q(j, qat(j));
                drain();
            }
        } else {
            for (; j + 4 <= e; j += 4) {
                const double4 q0 = pos2[j], q1 = pos2[j + 1], q2 = pos2[j + 2], q3 = pos2[j + 3];
                put(j, exact(q0));
                put(j + 1, exact(q1));
                put(j + 2, exact(q2));
                put(j + 3, exact(q3));
                drain();
            }
            for (; j < e; ++j) {
                put(j, exact(pos2[j]));
                drain();
            }
        }
    };
    auto stencil = [&](bool fixed) {
        if (g.usable) {
            const int nx = g.nc[0], ny = g.nc[1], nz = g.nc[2];
            const int cx = ci % nx;
            const int cy = (ci / nx) % ny;
            const int cz = ci / (nx * ny);
            const int xa = cx >= 2 ? cx - 2 : 0, xb = cx + 2 < nx ? cx + 2 : nx - 1;
            // BRICK: the row atom's brick, and the halo x of the pieces' first
            // cells (global cell gx lies at halo x = gx - x0 + 2, wrapped ones
            // at gx -+ nx - x0 + 2)
            const int* tb = nullptr;
            int hxn = 0, hyn = 0, hy0 = 0, hz0 = 0, hxa = 0, hxw = 0;
            if (BRICK) {
                const Bricks& k = a.bk;
                const int b = brick_of(cx, nx, k.kb[0]) +
                              k.kb[0] * (brick_of(cy, ny, k.kb[1]) + k.kb[1] * brick_of(cz, nz, k.kb[2]));
                const BrickGeo o = brick_geo(k, g, b);
                tb = k.tab + k.tab_off[b];
                hxn = o.ex + 4;
                hyn = o.ey + 4;
                hy0 = cy - o.y0 + 2;
                hz0 = cz - o.z0 + 2;
                hxa = xa - o.x0 + 2;
                hxw = cx < 2 ? cx - o.x0 : nx - o.x0 + 2;
            }
            for (int oz = -2; oz <= 2; ++oz) {
                int zz = cz + oz;
                zz += zz < 0 ? nz : (zz >= nz ? -nz : 0);
                for (int oy = -2; oy <= 2; ++oy) {
                    int yy = cy + oy;
                    yy += yy < 0 ? ny : (yy >= ny ? -ny : 0);
                    const int rb = nx * (yy + ny * zz);
                    const int hrow = BRICK ? ((hz0 + oz) * hyn + (hy0 + oy)) * hxn : 0;
                    int s0 = cell_start[rb + xa];
                    if (BRICK) delta = tb[hrow + hxa] - s0;
                    range(s0, cell_start[rb + xb + 1], fixed);
                    if (cx < 2 || cx + 2 >= nx) {
                        s0 = cell_start[cx < 2 ? rb + nx + cx - 2 : rb];
                        if (BRICK) delta = tb[hrow + hxw] - s0;
                        range(s0, cell_start[cx < 2 ? rb + nx : rb + cx + 3 - nx], fixed);
                    }
                }
            }
        } else {
            range(0, static_cast<int>(n), fixed);
        }
    };
    stencil(Q);
    if (Q && hits != 0) {  // rare: a fixed-point r2 inside the band, the row again exactly
        cnt = 0;
        done = 0;
        stencil(false);
    }
    if (cnt > done) write_octet();  // the last, partial octet (slots past the row are not read)
    a.ncnt[t] = cnt;
    if (cnt > cap) atomicOr(&a.sc->overflow, 1);
}

// Commit + build for small systems (at most kBuildLanesMaxAtoms atoms): one
// thread per slot leaves most of the chip idle and walks about 300 candidates
// in a dependent chain, so kBuildLanes consecutive lanes share a slot. The
// group takes a range kBuildLanes candidates at a time, lane s candidate
// base + s, and appends the kept ones in candidate order through a ballot
// over its lanes: the row, and its order, are the ones k_commit_build writes.
// A band hit in any lane makes the group build its row again exactly.
constexpr long kBuildLanesMaxAtoms = 32768;
constexpr int kBuildLanes = 8;

template <bool Q, bool CUBIC>
__global__ __launch_bounds__(kBlock) void k_commit_build_lanes(NeighArgs a) {
    if (*a.flag == 0) return;
    const long n = a.n;
    const Grid& g = a.grid;
    const double rl2 = a.rl2;
    const int cap = a.cap;
    const long gt = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    const long tl = gt / kBuildLanes;
    const int sub = static_cast<int>(gt & (kBuildLanes - 1));
    if (tl >= n) return;  // whole groups: kBuildLanes divides the workgroup
    const int t = static_cast<int>(tl);
    const double4 pi = a.pos2[t];
    const int oi = a.orig2[t];
    const int ci = a.cell2[t];
    int4 qi{};
    if (Q) qi = a.qpos2[t];
    if (sub == 0) {
        a.pos[t] = pi;
        a.ref[t] = pi;
        a.vel[t] = a.vel2[t];
        a.frc[t] = a.frc2[t];
        a.orig[t] = oi;
        a.slot[oi] = t;
        a.cell_of[t] = ci;
        if (Q) a.qpos[t] = qi;
    }
    const int lane = threadIdx.x & (kWave - 1);
    const int head = lane & ~(kBuildLanes - 1);
    const unsigned long long gmask = ((1ull << kBuildLanes) - 1) << head;
    const unsigned long long below = (1ull << lane) - 1;
    const double u2 = CUBIC ? a.quant.unit[0] * a.quant.unit[0] : 1.0;
    const double lo = rl2 * (1.0 - a.quant.band_l) / u2, hi = rl2 * (1.0 + a.quant.band_l) / u2;
    int cnt = 0;
    unsigned int hits = 0;
    auto keep_entry = [&](bool keep, int j) {
        const unsigned long long m = __ballot(keep) & gmask;
        const int at = cnt + __popcll(m & below);
        if (keep && at < cap) a.nbr[oct_index(t, at, cap)] = entry_of(j);
        cnt += __popcll(m);
    };
    auto range = [&](int b, int e, bool fixed) {
        for (int base = b; base < e; base += kBuildLanes) {
            const int j = base + sub;
            const bool ok = j < e;
            const int jj = ok ? j : b;
            bool keep;
            if (Q && fixed) {
                const int4 qj = a.qpos2[jj];
                double r2;
                if (CUBIC) {
                    const double ux = static_cast<double>(static_cast<int>(static_cast<uint32_t>(qi.x) - static_cast<uint32_t>(qj.x)));
                    const double uy = static_cast<double>(static_cast<int>(static_cast<uint32_t>(qi.y) - static_cast<uint32_t>(qj.y)));
                    const double uz = static_cast<double>(static_cast<int>(static_cast<uint32_t>(qi.z) - static_cast<uint32_t>(qj.z)));
                    r2 = ux * ux + uy * uy + uz * uz;
                } else {
                    double dx, dy, dz;
                    qdelta(a.quant, qi, qj, dx, dy, dz);
                    r2 = dx * dx + dy * dy + dz * dz;
                }
                hits += (ok && r2 >= lo && r2 < hi) ? 1u : 0u;
                keep = ok && r2 < lo && j != t;
            } else {
                double dx, dy, dz;
                row_delta_w(a.box, pi, a.pos2[jj], dx, dy, dz);
                keep = ok && norm2_ref(dx, dy, dz) < rl2 && j != t;
            }
            keep_entry(keep, j);
        }
    };
    // the same 25 ranges in the same order as k_commit_build
    auto stencil = [&](bool fixed) {
        if (g.usable) {
            const int nx = g.nc[0], ny = g.nc[1], nz = g.nc[2];
            const int cx = ci % nx;
            const int cy = (ci / nx) % ny;
            const int cz = ci / (nx * ny);
            const int xa = cx >= 2 ? cx - 2 : 0, xb = cx + 2 < nx ? cx + 2 : nx - 1;
            for (int oz = -2; oz <= 2; ++oz) {
                int zz = cz + oz;
                zz += zz < 0 ? nz : (zz >= nz ? -nz : 0);
                for (int oy = -2; oy <= 2; ++oy) {
                    int yy = cy + oy;
                    yy += yy < 0 ? ny : (yy >= ny ? -ny : 0);
                    const int rb = nx * (yy + ny * zz);
                    range(a.cell_start[rb + xa], a.cell_start[rb + xb + 1], fixed);
                    if (cx < 2)
                        range(a.cell_start[rb + nx + cx - 2], a.cell_start[rb + nx], fixed);
                    else if (cx + 2 >= nx)
                        range(a.cell_start[rb], a.cell_start[rb + cx + 3 - nx], fixed);
                }
            }
        } else {
            range(0, static_cast<int>(n), fixed);
        }
    };
    stencil(Q);
    if (Q && (__ballot(hits != 0) & gmask) != 0) {  // a band hit in the group: the row exactly
        cnt = 0;
        stencil(false);
    }
    if (sub == 0) {
        a.ncnt[t] = cnt;
        if (cnt > cap) atomicOr(&a.sc->overflow, 1);
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
    // k_sort_permute takes one wave per cell
    const dim3 gn(grid_capped(n)), b(kBlock);
    const dim3 gt(scan_tiles(ncell)), bt(kScanTile);
    hipLaunchKernelGGL(k_bin, gn, b, 0, s, a.pos, n, a.grid, a.box, a.cell_of, a.cell_count, a.flag);
    hipLaunchKernelGGL(k_tile_count, gt, bt, 0, s, a.cell_count, ncell, a.tile_sum, a.flag);
    hipLaunchKernelGGL(k_scan, gt, bt, 0, s, a.cell_count, ncell, a.tile_sum, a.cell_start,
                       a.cell_fill, a.sc, a.flag);
    const dim3 gp(gn.x + (a.bk.on ? static_cast<unsigned>(a.bk.nb) : 0u));
    hipLaunchKernelGGL(k_place, gp, b, 0, s, a.cell_of, a.orig, n, a.cell_fill, a.order, a.flag,
                       a.grid, a.bk, a.cell_start, a.sc);
    hipLaunchKernelGGL(k_sort_permute, gn, b, 0, s, a);
    if (n <= kBuildLanesMaxAtoms) {
        const dim3 gl(static_cast<unsigned>((n * kBuildLanes + kBlock - 1) / kBlock));
        if (a.qpos2 != nullptr && a.quant.cubic)
            hipLaunchKernelGGL((k_commit_build_lanes<true, true>), gl, b, 0, s, a);
        else if (a.qpos2 != nullptr)
            hipLaunchKernelGGL((k_commit_build_lanes<true, false>), gl, b, 0, s, a);
        else
            hipLaunchKernelGGL((k_commit_build_lanes<false, false>), gl, b, 0, s, a);
    } else if (a.bk.on && a.quant.cubic)
        hipLaunchKernelGGL((k_commit_build<true, true, true>), dim3(xcd_grid(n)), b, 0, s, a);
    else if (a.bk.on)
        hipLaunchKernelGGL((k_commit_build<true, false, true>), dim3(xcd_grid(n)), b, 0, s, a);
    else if (a.qpos2 != nullptr && a.quant.cubic)
        hipLaunchKernelGGL((k_commit_build<true, true, false>), dim3(xcd_grid(n)), b, 0, s, a);
    else if (a.qpos2 != nullptr)
        hipLaunchKernelGGL((k_commit_build<true, false, false>), dim3(xcd_grid(n)), b, 0, s, a);
    else
        hipLaunchKernelGGL((k_commit_build<false, false, false>), dim3(xcd_grid(n)), b, 0, s, a);
}

long scan_tiles(long ncell) { return (ncell + kScanTile - 1) / kScanTile; }

}  // namespace mdk
