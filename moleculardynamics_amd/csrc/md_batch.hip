// Batched independent replicas: three launches advance every replica by one
// BAOAB step (integrate/langevin.hpp:36-88) with its own Verlet list
// (neighbor/verlet_list.cpp:8-69) and LJ forces
// (forcefield/lennard_jones.cpp:50-88).
//
// Each replica is split into pieces, one workgroup each (the piece count per
// replica is a per-size-class choice of the autotuner); the pieces of all
// replicas form one grid, longest first:
//   1. k_batch_pre: B, A, O, A over the piece's atoms and the piece's two
//      largest displacements since the last build;
//   2. k_batch_plan (one workgroup): each replica's rebuild decision and the
//      force kernel's queue, pieces of rebuilding replicas first;
//   3. k_batch_force: every piece stages its whole replica's positions in
//      LDS (FP64, in slot order: the atoms sorted by cell at the last build).
//      When the sum of the replica's two largest displacements exceeds the
//      skin, the piece re-sorts the replica by cell in LDS (a counting sort;
//      the atoms of a cell ordered by index, so the order is a pure function
//      of the positions and every piece computes the same one) and rebuilds
//      the rows of its own slots by the reference's exact pair test: from the
//      5 x 5 x 5 cells of width >= (rc + skin)/2 around each atom when 5 or
//      more such cells fit on every axis, over all atom pairs otherwise --
//      the same pair set as the reference's cell or brute-force build,
//      including the single-nearest-image semantics of boxes shorter than
//      2 (rc + skin). Then `tpa` lanes per slot walk its
//      row, gathering partners from LDS: FP32 pair terms with 64-bit
//      fixed-point accumulation (or FP64), the final B half-kick, and the
//      piece's E, W, KE partials;
//   4. k_batch_sum: one workgroup sums each replica's piece partials in piece
//      order into the thermo ring and advances the device step counter.
// A segment of identical launches can therefore be captured in a HIP graph
// and replayed.
#include "md_common.hpp"
#include "md_batch.hpp"

namespace mdk {

namespace {


// Cell of a coordinate: floor((wrap(r) + L/2) / width) clamped to [0, n-1]
// (neighbor/cell_list.cpp:31-38), wrap the reference fold.
__device__ __forceinline__ int rep_cell_coord(double r, double L, double w, int n) {
#pragma clang fp contract(off)
    const double f = fold_ref(r, L);
    const int k = static_cast<int>(floor((f + 0.5 * L) / w));
    return k < 0 ? 0 : (k >= n ? n - 1 : k);
}

// The workgroup's LDS: the replica's positions and slot -> atom map, then
// the cell tables of the rebuild.
struct Stage {
    double *xs, *ys, *zs;
    int* ol;    // slot -> atom; during a sort, the atoms grouped by cell
    int* cidx;  // atom -> cell, then atom -> slot
    int* cend;  // per cell: count, start, then end offset
};

__device__ __forceinline__ Stage stage_of(const BatchArgs& a) {
    extern __shared__ double lds[];
    Stage st;
    st.xs = lds;
    st.ys = lds + a.nmax;
    st.zs = lds + 2 * a.nmax;
    st.ol = reinterpret_cast<int*>(lds + 3 * a.nmax);
    st.cidx = st.ol + a.nmax;
    st.cend = st.cidx + a.nmax;
    return st;
}

__device__ __forceinline__ int cell_begin(const Stage& st, int c) { return c > 0 ? st.cend[c - 1] : 0; }

// Counting sort of the replica's atoms by cell into slot order: the atoms of
// a cell in index order. Stages positions and the slot map in LDS; for the
// piece's own atoms [s0, s1) writes the new slot and the build-time position.
__device__ void sort_slots(const BatchArgs& a, const BatchRep& R, const Stage& st, int s0, int s1) {
    const double4* P = a.pos + R.off;
    const int n = R.n;
    const int nx = R.nc[0], ny = R.nc[1], nz = R.nc[2];
    const int ncell = nx * ny * nz;
    const int tid = threadIdx.x, bd = blockDim.x;
    for (int c = tid; c < ncell; c += bd) st.cend[c] = 0;
    __syncthreads();
    for (int i = tid; i < n; i += bd) {
        const double4 p = P[i];
        const int cx = rep_cell_coord(p.x, R.box.L[0], R.w[0], nx);
        const int cy = rep_cell_coord(p.y, R.box.L[1], R.w[1], ny);
        const int cz = rep_cell_coord(p.z, R.box.L[2], R.w[2], nz);
        const int c = cx + nx * (cy + ny * cz);
        st.cidx[i] = c;
        atomicAdd(&st.cend[c], 1);
    }
    __syncthreads();
    if (tid < kWave) {  // exclusive scan of the counts in place by the first wave
        const int lane = tid;
        int run = 0;
        for (int c0 = 0; c0 < ncell; c0 += kWave) {
            const int c = c0 + lane;
            const int v = c < ncell ? st.cend[c] : 0;
            int x = v;
            for (int off = 1; off < kWave; off <<= 1) {
                const int y = __shfl_up(x, off, kWave);
                if (lane >= off) x += y;
            }
            if (c < ncell) st.cend[c] = run + x - v;
            run += __shfl(x, kWave - 1, kWave);
        }
    }
    __syncthreads();
    // group by cell (the start offsets advance to the end offsets)
    for (int i = tid; i < n; i += bd) st.ol[atomicAdd(&st.cend[st.cidx[i]], 1)] = i;
    __syncthreads();
    // slot = cell start + rank of the atom index within its cell
    for (int i = tid; i < n; i += bd) {
        const int c = st.cidx[i], b = cell_begin(st, c), e = st.cend[c];
        int r = 0;
        for (int k = b; k < e; ++k) r += st.ol[k] < i ? 1 : 0;
        st.cidx[i] = b + r;
    }
    __syncthreads();
    for (int i = tid; i < n; i += bd) {
        const int s = st.cidx[i];
        const double4 p = P[i];
        st.ol[s] = i;
        st.xs[s] = p.x;
        st.ys[s] = p.y;
        st.zs[s] = p.z;
        if (i >= s0 && i < s1) {
            a.slot_of[R.off + i] = s;
            a.ref[R.off + i] = p;
        }
    }
    __syncthreads();
}

constexpr int kBuildLanes = 8;  // lanes per slot in the row build

// Appends to a row, in slot order, the slots k in [ks, ke) within rc + skin
// of slot s (neighbor/verlet_list.cpp:26-50's test: the lower atom index
// first). The kBuildLanes lanes of a slot test consecutive candidates
// together; a ballot compacts the accepted ones (`gshift` is the group's
// first lane in the wave).
__device__ __forceinline__ int scan_range(const BatchArgs& a, const BatchRep& R, const Stage& st, int s,
                                          int oi, const double4& pi, int ks, int ke, uint32_t* row,
                                          int cnt, int sub, int gshift) {
    for (int k0 = ks; k0 < ke; k0 += kBuildLanes) {
        const int k = k0 + sub;
        bool acc = false;
        uint32_t ent = 0;
        if (k < ke && k != s) {
            const bool flip = oi > st.ol[k];
            double dx, dy, dz;
            row_delta(R.box, pi, make_double4(st.xs[k], st.ys[k], st.zs[k], 0.0), flip, dx, dy, dz);
            acc = norm2_ref(dx, dy, dz) < a.rl2;
            ent = static_cast<uint32_t>(k) | (flip ? kFlipBit : 0u);
        }
        const uint32_t gm = static_cast<uint32_t>(__ballot(acc) >> gshift) & ((1u << kBuildLanes) - 1u);
        if (acc) {
            const int pos = cnt + __popc(gm & ((1u << sub) - 1u));
            if (pos < a.cap) row[pos] = ent;
        }
        cnt += __popc(gm);
    }
    return cnt;
}

// Rows of the piece's slots [s0, s1), kBuildLanes lanes per slot.
// Cell-sorted replicas (cells of at least (rc + skin)/2, 5 or more per axis)
// scan the 5 x 5 runs of five x-adjacent cells around the slot's cell --
// contiguous in slot order, split in two where the run wraps the x axis;
// every pair within rc + skin lies in such a run. Other replicas scan all
// slots.
__device__ void build_rows(const BatchArgs& a, const BatchRep& R, const Stage& st, bool cells, int s0,
                           int s1) {
    const int nx = R.nc[0], ny = R.nc[1], nz = R.nc[2];
    const int tid = threadIdx.x;
    const int groups = blockDim.x / kBuildLanes, sub = tid % kBuildLanes;
    const int gshift = (tid & (kWave - 1)) & ~(kBuildLanes - 1);
    for (int base = s0; base < s1; base += groups) {
        const int s = base + tid / kBuildLanes;
        if (s >= s1) continue;
        const long g = R.off + s;
        const int oi = st.ol[s];
        const double4 pi = make_double4(st.xs[s], st.ys[s], st.zs[s], 0.0);
        uint32_t* row = a.nbr + g * a.cap;
        int cnt = 0;
        if (cells) {
            const int cx = rep_cell_coord(pi.x, R.box.L[0], R.w[0], nx);
            const int cy = rep_cell_coord(pi.y, R.box.L[1], R.w[1], ny);
            const int cz = rep_cell_coord(pi.z, R.box.L[2], R.w[2], nz);
            const int xa = cx >= 2 ? cx - 2 : 0, xb = cx + 2 < nx ? cx + 2 : nx - 1;
            for (int oz = -2; oz <= 2; ++oz) {
                int zz = cz + oz;
                zz += zz < 0 ? nz : (zz >= nz ? -nz : 0);
                for (int oy = -2; oy <= 2; ++oy) {
                    int yy = cy + oy;
                    yy += yy < 0 ? ny : (yy >= ny ? -ny : 0);
                    const int b = nx * (yy + ny * zz);
                    cnt = scan_range(a, R, st, s, oi, pi, cell_begin(st, b + xa), st.cend[b + xb], row, cnt,
                                     sub, gshift);
                    if (cx < 2)
                        cnt = scan_range(a, R, st, s, oi, pi, cell_begin(st, b + nx + cx - 2),
                                         st.cend[b + nx - 1], row, cnt, sub, gshift);
                    else if (cx + 2 >= nx)
                        cnt = scan_range(a, R, st, s, oi, pi, cell_begin(st, b), st.cend[b + cx + 2 - nx],
                                         row, cnt, sub, gshift);
                }
            }
        } else {
            cnt = scan_range(a, R, st, s, oi, pi, 0, R.n, row, cnt, sub, gshift);
        }
        if (sub == 0) {
            a.ncnt[g] = cnt;
            a.orig[g] = oi;
            if (cnt > a.cap) atomicOr(a.overflow, 1);
        }
    }
}

__global__ __launch_bounds__(256) void k_batch_pre(BatchArgs a) {
    __shared__ double red_d[4], red_d2[4];
    const int4 pc = a.pieces[blockIdx.x];
    const BatchRep R = a.reps[pc.x];
    const uint64_t step = *a.step;
    const double h = 0.5 * a.dt;
    const Box box = R.box;
    const double c1 = R.c1, c2 = R.c2, T = R.T;
    const uint32_t k0 = R.seed_lo, k1 = R.seed_hi;
    const long off = R.off;
    double t1 = 0.0, t2 = 0.0;
    for (int i = pc.y + static_cast<int>(threadIdx.x); i < pc.z; i += blockDim.x) {
        const long g = off + i;
        double4 p = a.pos[g];
        double4 v = a.vel[g];
        const double4 f = a.frc[g];
        const double c = h * (1.0 / v.w);
        v.x = fma(c, f.x, v.x);
        v.y = fma(c, f.y, v.y);
        v.z = fma(c, f.z, v.z);
        p = drift(box, p, v, h);
        double g0, g1, g2;
        normals3(k0, k1, step, static_cast<uint32_t>(i), g0, g1, g2);
        const double k = c2 * sqrt(T / v.w);
        v.x = c1 * v.x + g0 * k;
        v.y = c1 * v.y + g1 * k;
        v.z = c1 * v.z + g2 * k;
        p = drift(box, p, v, h);
        a.pos[g] = p;
        a.vel[g] = v;
        const double4 r = a.ref[g];
        double dx = p.x - r.x, dy = p.y - r.y, dz = p.z - r.z;
        min_image(box, dx, dy, dz);
        top2_merge(t1, t2, sqrt(norm2_ref(dx, dy, dz)), 0.0);
    }
    wave_top2(t1, t2);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    if (lane == 0) {
        red_d[wid] = t1;
        red_d2[wid] = t2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double b1 = 0.0, b2 = 0.0;
        for (int w = 0; w < static_cast<int>(blockDim.x / kWave); ++w) top2_merge(b1, b2, red_d[w], red_d2[w]);
        a.ptop[pc.w] = make_double2(b1, b2);
    }
}

// Workgroup sum (any power-of-two workgroup of wave64s); result in thread 0.
template <class T, class WaveFn>
__device__ __forceinline__ T wg_sum(T v, T* smem, WaveFn wf) {
    v = wf(v);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    __syncthreads();
    if (lane == 0) smem[wid] = v;
    __syncthreads();
    T r = T(0);
    for (int w = 0; w < nw; ++w) r += smem[w];
    return r;
}

// The step's plan (one workgroup of 1024 threads): each replica's rebuild
// decision from its pieces' displacement top-2 (always in mode 0), then the
// launch queue of the force kernel -- the pieces of rebuilding replicas first,
// each part in the static longest-first order, so the long rebuild pieces
// start first.
__global__ __launch_bounds__(1024) void k_batch_plan(BatchArgs a) {
    __shared__ int wsum[16];
    const int tid = threadIdx.x, bd = blockDim.x;
    for (int r = tid; r < a.nrep; r += bd) {
        int rb = 1;
        if (a.mode == 1) {
            const int p0 = a.reps[r].piece0, np = a.reps[r].npiece;
            double b1 = 0.0, b2 = 0.0;
            for (int k = 0; k < np; ++k) {
                const double2 t = a.ptop[p0 + k];
                top2_merge(b1, b2, t.x, t.y);
            }
            rb = b1 + b2 > a.skin ? 1 : 0;
        }
        a.rflag[r] = rb;
    }
    __syncthreads();
    // stable partition: thread t owns pieces [t*per, (t+1)*per)
    const int per = (a.npieces + bd - 1) / bd;
    const int q0 = min(tid * per, a.npieces), q1 = min(q0 + per, a.npieces);
    int c1 = 0;
    for (int q = q0; q < q1; ++q) c1 += a.rflag[a.pieces[q].x];
    // exclusive block scan of c1
    const int lane = tid & (kWave - 1), wid = tid / kWave;
    int x = c1;
    for (int o = 1; o < kWave; o <<= 1) {
        const int y = __shfl_up(x, o, kWave);
        if (lane >= o) x += y;
    }
    if (lane == kWave - 1) wsum[wid] = x;
    __syncthreads();
    int before = 0, total = 0;
    for (int w = 0; w < static_cast<int>(bd / kWave); ++w) {
        before += w < wid ? wsum[w] : 0;
        total += wsum[w];
    }
    int p1 = before + x - c1;          // next slot for a rebuilding piece
    int p0 = total + (q0 - (p1));      // next slot for the others
    for (int q = q0; q < q1; ++q) {
        const int4 pc = a.pieces[q];
        if (a.rflag[pc.x]) a.queue[p1++] = pc;
        else a.queue[p0++] = pc;
    }
}

template <bool MIXED>
__global__ __launch_bounds__(1024) void k_batch_force(BatchArgs a) {
    __shared__ double red_d[16];
    __shared__ long long red_l[16];
    __shared__ int s_rebuild;
    const int4 pc = a.queue[blockIdx.x];
    const BatchRep R = a.reps[pc.x];
    const int n = R.n, s0 = pc.y, s1 = pc.z;
    const long off = R.off;
    const int tid = threadIdx.x, bd = blockDim.x;
    const Stage st = stage_of(a);
    const double h = 0.5 * a.dt;
    if (tid == 0) s_rebuild = a.rflag[pc.x];
    __syncthreads();

    // ---- 1. stage the replica in slot order; rebuild the piece's rows ------
    if (!s_rebuild) {
        for (int i = tid; i < n; i += bd) {
            const double4 p = a.pos[off + i];
            const int s = a.slot_of[off + i];
            st.xs[s] = p.x;
            st.ys[s] = p.y;
            st.zs[s] = p.z;
        }
        for (int s = s0 + tid; s < s1; s += bd) st.ol[s] = a.orig[off + s];
        __syncthreads();
    } else {
        const bool cells = R.usable && a.maxcell > 0;
        if (cells) {
            sort_slots(a, R, st, s0, s1);
        } else {
            for (int i = tid; i < n; i += bd) {
                const double4 p = a.pos[off + i];
                st.xs[i] = p.x;
                st.ys[i] = p.y;
                st.zs[i] = p.z;
                st.ol[i] = i;
                if (i >= s0 && i < s1) a.ref[off + i] = p;
            }
            __syncthreads();
        }
        build_rows(a, R, st, cells, s0, s1);
        if (tid == 0 && s0 == 0) a.builds[pc.x] += 1;
        __syncthreads();
    }

    // ---- 2. forces, second half-kick, piece partials ------------------------
    const int tpa = R.tpa;
    const int groups = bd / tpa;
    const int sub = tid % tpa;
    double e64 = 0.0, w64 = 0.0, ke = 0.0;
    long long ef = 0, wf = 0;
    float big = 0.0f;
    for (int base = s0; base < s1; base += groups) {
        const int s = base + tid / tpa;
        const bool valid = s < s1;
        double fx = 0.0, fy = 0.0, fz = 0.0;
        long long ax = 0, ay = 0, az = 0;
        if (valid) {
            const long g = off + s;
            const int cnt = min(a.ncnt[g], a.cap);
            const uint32_t* row = a.nbr + g * a.cap;
            const double4 pi = make_double4(st.xs[s], st.ys[s], st.zs[s], 0.0);
            unsigned long long bx = 0, by = 0, bz = 0, be = 0, bw = 0, nt = 0;
            for (int k = sub; k < cnt; k += tpa) {
                const uint32_t jj = row[k];
                const int j = static_cast<int>(jj & kIndexMask);
                const bool flip = (jj & kFlipBit) != 0;
                double dx, dy, dz;
                row_delta(R.box, pi, make_double4(st.xs[j], st.ys[j], st.zs[j], 0.0), flip, dx, dy, dz);
                if (MIXED) {
                    const float xf = static_cast<float>(dx), yf = static_cast<float>(dy),
                                zf = static_cast<float>(dz);
                    float inv, u, ww;
                    if (lj_terms_f32(norm2_ref(dx, dy, dz), a.lj.rc2, a.lj.sig2f, a.lj.epsf, inv, u, ww)) {
                        const float c = ww * inv;
                        const float tx = c * xf, ty = c * yf, tz = c * zf;
                        be += fix_biased(u);
                        bw += fix_biased(ww);
                        bx += fix_biased(tx);
                        by += fix_biased(ty);
                        bz += fix_biased(tz);
                        nt += 1;
                        big = fmaxf(big, fmaxf(fmaxf(fabsf(u), fabsf(ww)),
                                               fmaxf(fabsf(tx), fmaxf(fabsf(ty), fabsf(tz)))));
                    }
                } else {
                    const double r2 = norm2_ref(dx, dy, dz);
                    if (r2 < a.lj.rc2) {
                        const double inv = 1.0 / r2;
                        const double s2 = a.lj.sig2 * inv;
                        const double s6 = s2 * s2 * s2;
                        e64 += 4.0 * a.lj.eps * (s6 * s6 - s6);
                        const double ww = 24.0 * a.lj.eps * (2.0 * s6 * s6 - s6);
                        w64 += ww;
                        const double c = ww * inv;
                        fx += c * dx;
                        fy += c * dy;
                        fz += c * dz;
                    }
                }
            }
            if (MIXED) {
                ax = fix_value(bx, nt);
                ay = fix_value(by, nt);
                az = fix_value(bz, nt);
                ef += fix_value(be, nt);
                wf += fix_value(bw, nt);
            }
        }
        // combine the tpa lanes of a slot (lanes of a group are adjacent)
        for (int o = 1; o < tpa; o <<= 1) {
            if (MIXED) {
                ax += __shfl_xor(ax, o, kWave);
                ay += __shfl_xor(ay, o, kWave);
                az += __shfl_xor(az, o, kWave);
            } else {
                fx += __shfl_xor(fx, o, kWave);
                fy += __shfl_xor(fy, o, kWave);
                fz += __shfl_xor(fz, o, kWave);
            }
        }
        if (valid && sub == 0) {
            if (MIXED) {
                fx = static_cast<double>(ax) * kFixInv;
                fy = static_cast<double>(ay) * kFixInv;
                fz = static_cast<double>(az) * kFixInv;
            }
            const long ga = off + st.ol[s];
            a.frc[ga] = make_double4(fx, fy, fz, 0.0);
            if (a.mode == 1) {
                double4 v = a.vel[ga];
                const double c = h * (1.0 / v.w);
                v.x = fma(c, fx, v.x);
                v.y = fma(c, fy, v.y);
                v.z = fma(c, fz, v.z);
                a.vel[ga] = v;
                ke += (0.5 * v.w) * norm2_ref(v.x, v.y, v.z);
            }
        }
    }
    if (!(big < kFixLimit)) atomicOr(a.fix_sat, 1);
    if (a.thermo != nullptr) {
        long long* out = a.ppart + 3L * pc.w;
        if (MIXED) {
            const long long se = wg_sum(ef, red_l, wave_sum_ll);
            const long long sw = wg_sum(wf, red_l, wave_sum_ll);
            if (tid == 0) {
                out[0] = se;
                out[1] = sw;
            }
        } else {
            const double se = wg_sum(e64, red_d, wave_sum);
            const double sw = wg_sum(w64, red_d, wave_sum);
            if (tid == 0) {
                out[0] = __double_as_longlong(se);
                out[1] = __double_as_longlong(sw);
            }
        }
        const double sk = wg_sum(ke, red_d, wave_sum);
        if (tid == 0) out[2] = __double_as_longlong(sk);
    }
}

// Per-replica sums of the piece partials (in piece order) into the thermo
// ring, then the step counter advances (one workgroup: every thread has read
// the step before thread 0 writes it).
__global__ __launch_bounds__(256) void k_batch_sum(BatchArgs a) {
    const uint64_t step = *a.step;
    if (a.thermo != nullptr) {
        for (int r = threadIdx.x; r < a.nrep; r += blockDim.x) {
            const int p0 = a.reps[r].piece0, np = a.reps[r].npiece;
            const long long* pp = a.ppart + 3L * p0;
            double E, W, K = 0.0;
            if (a.mixed) {
                long long se = 0, sw = 0;
                for (int k = 0; k < np; ++k) {
                    se += pp[3 * k];
                    sw += pp[3 * k + 1];
                }
                E = 0.5 * (static_cast<double>(se) * kFixInv);
                W = 0.5 * (static_cast<double>(sw) * kFixInv);
            } else {
                double se = 0.0, sw = 0.0;
                for (int k = 0; k < np; ++k) {
                    se += __longlong_as_double(pp[3 * k]);
                    sw += __longlong_as_double(pp[3 * k + 1]);
                }
                E = 0.5 * se;
                W = 0.5 * sw;
            }
            for (int k = 0; k < np; ++k) K += __longlong_as_double(pp[3 * k + 2]);
            double* th = a.thermo + (static_cast<long>(step % a.thermo_steps) * a.nrep + r) * 4;
            th[0] = E;
            th[1] = W;
            th[2] = K;
            th[3] = static_cast<double>(step);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) *a.step = step + 1;
}

}  // namespace

size_t batch_lds_bytes(int nmax, int maxcell) {
    size_t b = static_cast<size_t>(nmax) * (3 * sizeof(double) + sizeof(int));
    if (maxcell > 0) b += (static_cast<size_t>(nmax) + static_cast<size_t>(maxcell)) * sizeof(int);
    return b;
}

hipError_t launch_batch(const BatchArgs& a, int block, hipStream_t s) {
    const size_t shmem = batch_lds_bytes(a.nmax, a.maxcell);
    if (a.mode == 1) hipLaunchKernelGGL(k_batch_pre, dim3(a.npieces), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_batch_plan, dim3(1), dim3(1024), 0, s, a);
    if (a.mixed)
        hipLaunchKernelGGL(k_batch_force<true>, dim3(a.npieces), dim3(block), shmem, s, a);
    else
        hipLaunchKernelGGL(k_batch_force<false>, dim3(a.npieces), dim3(block), shmem, s, a);
    if (a.mode == 1) hipLaunchKernelGGL(k_batch_sum, dim3(1), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t batch_allow_lds(size_t bytes) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_batch_force<true>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       static_cast<int>(bytes));
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute(reinterpret_cast<const void*>(k_batch_force<false>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
}

}  // namespace mdk
