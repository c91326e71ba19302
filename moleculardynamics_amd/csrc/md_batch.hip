// Batched independent replicas: one kernel launch advances every replica by
// one BAOAB step (integrate/langevin.hpp:36-88) with its own Verlet list
// (neighbor/verlet_list.cpp:8-69) and LJ forces
// (forcefield/lennard_jones.cpp:50-88).
//
// One workgroup owns one replica for the whole step, so every phase that
// needs the replica's atoms together (the top-2 displacement check, the
// rebuild, the force rows) synchronises with workgroup barriers only:
//   1. B, A, O, A per atom; the replica's positions are staged in LDS
//      (x, y, z FP64 arrays; up to 160 KB per workgroup on gfx950) and the
//      displacement top-2 is reduced in the workgroup;
//   2. when the sum of the two largest displacements exceeds the skin, the
//      replica's rows are rebuilt by the reference's exact pair test: from
//      the 27 cells around each atom (a counting sort by cell in LDS) when the
//      replica's cell grid is usable, over all atom pairs in LDS otherwise --
//      the same pair set as the reference's cell or brute-force build,
//      including the single-nearest-image semantics of boxes shorter than
//      2 (rc + skin);
//   3. forces: `tpa` lanes per atom walk its row (a per-size-class choice
//      made by the autotuner), gathering partners from LDS; FP32 pair terms
//      with 64-bit fixed-point accumulation (or FP64); the final B half-kick
//      and the replica's E, W, KE.
// The step counter lives on the device and a one-thread kernel after each
// step advances it, so a segment of identical launches can be captured in a
// HIP graph and replayed.
#include "md_common.hpp"
#include "md_batch.hpp"

namespace mdk {

namespace {

struct U4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox10(U4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(M0, c.x), lo0 = M0 * c.x;
        const uint32_t hi1 = __umulhi(M1, c.z), lo1 = M1 * c.z;
        c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += W0;
        k1 += W1;
    }
    return c;
}

__device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
    const uint64_t v = (static_cast<uint64_t>(a) << 21) ^ static_cast<uint64_t>(b >> 11);
    return (static_cast<double>(v & ((1ULL << 53) - 1)) + 0.5) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ double4 drift(const Box& b, double4 p, const double4& v, double h) {
    p.x = fold(fma(h, v.x, p.x), b.L[0], b.lo[0], b.hi1[0], b.hi2[0]);
    p.y = fold(fma(h, v.y, p.y), b.L[1], b.lo[1], b.hi1[1], b.hi2[1]);
    p.z = fold(fma(h, v.z, p.z), b.L[2], b.lo[2], b.hi1[2], b.hi2[2]);
    return p;
}

// Cell of a coordinate: floor((wrap(r) + L/2) / width) clamped to [0, n-1]
// (neighbor/cell_list.cpp:31-38), wrap the reference fold.
__device__ __forceinline__ int rep_cell_coord(double r, double L, double w, int n) {
#pragma clang fp contract(off)
    const double f = fold_ref(r, L);
    const int k = static_cast<int>(floor((f + 0.5 * L) / w));
    return k < 0 ? 0 : (k >= n ? n - 1 : k);
}

// Rebuild of a replica with a usable cell grid, in the workgroup's LDS:
// counting sort of the atoms by cell (LDS atomics, then the atoms of a cell
// ordered by index, so the order is a pure function of the positions), and
// each atom's row from the 27 surrounding cells in (z, y, x) order with the
// reference's exact pair test. The pair set is the all-pairs set: every pair
// within rc + skin lies in neighbouring cells.
__device__ void rebuild_cells(const BatchArgs& a, const BatchRep& R, const double* xs,
                              const double* ys, const double* zs) {
    extern __shared__ double lds[];
    const int n = R.n;
    int* cidx = reinterpret_cast<int*>(lds + 3 * a.nmax);
    int* tmp = cidx + a.nmax;
    int* srt = tmp + a.nmax;
    int* cstart = srt + a.nmax;
    int* cfill = cstart + a.maxcell + 1;
    const int nx = R.nc[0], ny = R.nc[1], nz = R.nc[2];
    const int ncell = nx * ny * nz;
    const int tid = threadIdx.x, bd = blockDim.x;
    for (int c = tid; c < ncell; c += bd) cfill[c] = 0;
    __syncthreads();
    for (int i = tid; i < n; i += bd) {
        const int cx = rep_cell_coord(xs[i], R.box.L[0], R.w[0], nx);
        const int cy = rep_cell_coord(ys[i], R.box.L[1], R.w[1], ny);
        const int cz = rep_cell_coord(zs[i], R.box.L[2], R.w[2], nz);
        const int c = cx + nx * (cy + ny * cz);
        cidx[i] = c;
        atomicAdd(&cfill[c], 1);
    }
    __syncthreads();
    if (tid < kWave) {  // exclusive scan of the counts by the first wave
        const int lane = tid;
        int run = 0;
        for (int c0 = 0; c0 < ncell; c0 += kWave) {
            const int c = c0 + lane;
            const int v = c < ncell ? cfill[c] : 0;
            int x = v;
            for (int off = 1; off < kWave; off <<= 1) {
                const int y = __shfl_up(x, off, kWave);
                if (lane >= off) x += y;
            }
            if (c < ncell) cstart[c] = run + x - v;
            run += __shfl(x, kWave - 1, kWave);
        }
        if (lane == 0) cstart[ncell] = run;
    }
    __syncthreads();
    for (int c = tid; c < ncell; c += bd) cfill[c] = cstart[c];
    __syncthreads();
    for (int i = tid; i < n; i += bd) tmp[atomicAdd(&cfill[cidx[i]], 1)] = i;
    __syncthreads();
    for (int i = tid; i < n; i += bd) {
        const int c = cidx[i], b = cstart[c], e = cstart[c + 1];
        int r = 0;
        for (int k = b; k < e; ++k) r += tmp[k] < i ? 1 : 0;
        srt[b + r] = i;
    }
    __syncthreads();
    for (int i = tid; i < n; i += bd) {
        const long g = R.off + i;
        const double4 pi = make_double4(xs[i], ys[i], zs[i], 0.0);
        const int c = cidx[i];
        const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
        int cnt = 0;
        uint32_t* row = a.nbr + g * a.cap;
        for (int oz = -1; oz <= 1; ++oz)
            for (int oy = -1; oy <= 1; ++oy)
                for (int ox = -1; ox <= 1; ++ox) {
                    const int nb = (cx + ox + nx) % nx + nx * ((cy + oy + ny) % ny + ny * ((cz + oz + nz) % nz));
                    for (int k = cstart[nb]; k < cstart[nb + 1]; ++k) {
                        const int j = srt[k];
                        if (j == i) continue;
                        const bool flip = i > j;
                        double dx, dy, dz;
                        row_delta(R.box, pi, make_double4(xs[j], ys[j], zs[j], 0.0), flip, dx, dy, dz);
                        if (norm2_ref(dx, dy, dz) < a.rl2) {
                            if (cnt < a.cap) row[cnt] = static_cast<uint32_t>(j) | (flip ? kFlipBit : 0u);
                            ++cnt;
                        }
                    }
                }
        a.ncnt[g] = cnt;
        if (cnt > a.cap) atomicOr(a.overflow, 1);
        a.ref[g] = pi;
    }
    if (tid == 0) a.builds[blockIdx.x] += 1;
    __syncthreads();
}

// Workgroup sum (any power-of-two workgroup of wave64s); result in all threads.
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

template <bool MIXED>
__global__ __launch_bounds__(1024) void k_batch(BatchArgs a) {
    extern __shared__ double lds[];
    __shared__ double red_d[16], red_d2[16];
    __shared__ long long red_l[16];
    __shared__ int s_rebuild;
    const BatchRep R = a.reps[blockIdx.x];
    const int n = R.n;
    double* xs = lds;
    double* ys = lds + a.nmax;
    double* zs = lds + 2 * a.nmax;
    const uint64_t step = *a.step;
    const double h = 0.5 * a.dt;

    // ---- 1. pre-force passes and displacement check -----------------------
    double t1 = 0.0, t2 = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const long g = R.off + i;
        double4 p = a.pos[g];
        if (a.mode == 1) {
            double4 v = a.vel[g];
            const double4 f = a.frc[g];
            const double c = h * (1.0 / v.w);
            v.x = fma(c, f.x, v.x);
            v.y = fma(c, f.y, v.y);
            v.z = fma(c, f.z, v.z);
            p = drift(R.box, p, v, h);
            const U4 q0 = philox10(U4{static_cast<uint32_t>(i), static_cast<uint32_t>(step),
                                      static_cast<uint32_t>(step >> 32), 0u}, R.seed_lo, R.seed_hi);
            const U4 q1 = philox10(U4{static_cast<uint32_t>(i), static_cast<uint32_t>(step),
                                      static_cast<uint32_t>(step >> 32), 1u}, R.seed_lo, R.seed_hi);
            const double r1 = sqrt(-2.0 * log(u53(q0.x, q0.y)));
            const double r2 = sqrt(-2.0 * log(u53(q1.x, q1.y)));
            double s1, c1, s2, c2;
            sincospi(2.0 * u53(q0.z, q0.w), &s1, &c1);
            sincospi(2.0 * u53(q1.z, q1.w), &s2, &c2);
            const double k = R.c2 * sqrt(R.T / v.w);
            v.x = R.c1 * v.x + (r1 * c1) * k;
            v.y = R.c1 * v.y + (r1 * s1) * k;
            v.z = R.c1 * v.z + (r2 * c2) * k;
            p = drift(R.box, p, v, h);
            a.pos[g] = p;
            a.vel[g] = v;
            const double4 r = a.ref[g];
            double dx = p.x - r.x, dy = p.y - r.y, dz = p.z - r.z;
            min_image(R.box, dx, dy, dz);
            const double d = sqrt(norm2_ref(dx, dy, dz));
            top2_merge(t1, t2, d, 0.0);
        }
        xs[i] = p.x;
        ys[i] = p.y;
        zs[i] = p.z;
    }
    wave_top2(t1, t2);
    {
        const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
        if (lane == 0) {
            red_d[wid] = t1;
            red_d2[wid] = t2;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double b1 = 0.0, b2 = 0.0;
            for (int w = 0; w < static_cast<int>(blockDim.x / kWave); ++w) top2_merge(b1, b2, red_d[w], red_d2[w]);
            s_rebuild = (a.mode == 0) || (b1 + b2 > a.skin);
        }
        __syncthreads();
    }

    // ---- 2. rebuild the replica's rows ----------------------------------
    if (s_rebuild && R.usable && a.maxcell > 0) {
        rebuild_cells(a, R, xs, ys, zs);
    } else if (s_rebuild) {
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const long g = R.off + i;
            const double4 pi = make_double4(xs[i], ys[i], zs[i], 0.0);
            int cnt = 0;
            uint32_t* row = a.nbr + g * a.cap;
            for (int j = 0; j < n; ++j) {
                if (j == i) continue;
                const bool flip = i > j;
                double dx, dy, dz;
                row_delta(R.box, pi, make_double4(xs[j], ys[j], zs[j], 0.0), flip, dx, dy, dz);
                if (norm2_ref(dx, dy, dz) < a.rl2) {
                    if (cnt < a.cap) row[cnt] = static_cast<uint32_t>(j) | (flip ? kFlipBit : 0u);
                    ++cnt;
                }
            }
            a.ncnt[g] = cnt;
            if (cnt > a.cap) atomicOr(a.overflow, 1);
            a.ref[g] = pi;
        }
        if (threadIdx.x == 0) a.builds[blockIdx.x] += 1;
        __syncthreads();
    }

    // ---- 3. forces, second half-kick, replica sums ------------------------
    const int tpa = R.tpa;
    const int groups = blockDim.x / tpa;
    const int sub = threadIdx.x % tpa;
    double e64 = 0.0, w64 = 0.0, ke = 0.0;
    long long ef = 0, wf = 0;
    float big = 0.0f;
    for (int base = 0; base < n; base += groups) {
        const int i = base + static_cast<int>(threadIdx.x) / tpa;
        const bool valid = i < n;
        double fx = 0.0, fy = 0.0, fz = 0.0;
        long long ax = 0, ay = 0, az = 0;
        if (valid) {
            const long g = R.off + i;
            const int cnt = min(a.ncnt[g], a.cap);
            const uint32_t* row = a.nbr + g * a.cap;
            const double xi = xs[i], yi = ys[i], zi = zs[i];
            unsigned long long bx = 0, by = 0, bz = 0, be = 0, bw = 0, nt = 0;
            for (int k = sub; k < cnt; k += tpa) {
                const uint32_t jj = row[k];
                const int j = static_cast<int>(jj & kIndexMask);
                const bool flip = (jj & kFlipBit) != 0;
                double dx, dy, dz;
                row_delta(R.box, make_double4(xi, yi, zi, 0.0), make_double4(xs[j], ys[j], zs[j], 0.0),
                          flip, dx, dy, dz);
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
        // combine the tpa lanes of an atom (lanes of a group are adjacent)
        for (int off = 1; off < tpa; off <<= 1) {
            if (MIXED) {
                ax += __shfl_xor(ax, off, kWave);
                ay += __shfl_xor(ay, off, kWave);
                az += __shfl_xor(az, off, kWave);
            } else {
                fx += __shfl_xor(fx, off, kWave);
                fy += __shfl_xor(fy, off, kWave);
                fz += __shfl_xor(fz, off, kWave);
            }
        }
        if (valid && sub == 0) {
            if (MIXED) {
                fx = static_cast<double>(ax) * kFixInv;
                fy = static_cast<double>(ay) * kFixInv;
                fz = static_cast<double>(az) * kFixInv;
            }
            const long g = R.off + i;
            a.frc[g] = make_double4(fx, fy, fz, 0.0);
            if (a.mode == 1) {
                double4 v = a.vel[g];
                const double c = h * (1.0 / v.w);
                v.x = fma(c, fx, v.x);
                v.y = fma(c, fy, v.y);
                v.z = fma(c, fz, v.z);
                a.vel[g] = v;
                ke += (0.5 * v.w) * norm2_ref(v.x, v.y, v.z);
            }
        }
    }
    if (!(big < kFixLimit)) atomicOr(a.fix_sat, 1);
    if (a.thermo != nullptr) {
        double E, W;
        if (MIXED) {
            const long long se = wg_sum(ef, red_l, wave_sum_ll);
            const long long sw = wg_sum(wf, red_l, wave_sum_ll);
            E = 0.5 * (static_cast<double>(se) * kFixInv);
            W = 0.5 * (static_cast<double>(sw) * kFixInv);
        } else {
            E = 0.5 * wg_sum(e64, red_d, wave_sum);
            W = 0.5 * wg_sum(w64, red_d, wave_sum);
        }
        const double K = wg_sum(ke, red_d, wave_sum);
        if (threadIdx.x == 0) {
            double* th = a.thermo + (static_cast<long>(step % a.thermo_steps) * a.nrep + blockIdx.x) * 4;
            th[0] = E;
            th[1] = W;
            th[2] = K;
            th[3] = static_cast<double>(step);
        }
    }
}

// Advances the device step counter after every replica finished the step
// (a separate one-thread launch: stream order makes it the barrier).
__global__ void k_batch_advance(unsigned long long* step) {
    if (threadIdx.x == 0) *step += 1;
}

}  // namespace

size_t batch_lds_bytes(int nmax, int maxcell) {
    size_t b = static_cast<size_t>(3) * nmax * sizeof(double);
    if (maxcell > 0) b += (static_cast<size_t>(3) * nmax + 2 * static_cast<size_t>(maxcell) + 1) * sizeof(int);
    return b;
}

hipError_t launch_batch(const BatchArgs& a, int block, hipStream_t s) {
    const size_t shmem = batch_lds_bytes(a.nmax, a.maxcell);
    if (a.mixed)
        hipLaunchKernelGGL(k_batch<true>, dim3(a.nrep), dim3(block), shmem, s, a);
    else
        hipLaunchKernelGGL(k_batch<false>, dim3(a.nrep), dim3(block), shmem, s, a);
    if (a.mode == 1) hipLaunchKernelGGL(k_batch_advance, dim3(1), dim3(64), 0, s, a.step);
    return hipGetLastError();
}

hipError_t batch_allow_lds(size_t bytes) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_batch<true>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       static_cast<int>(bytes));
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute(reinterpret_cast<const void*>(k_batch<false>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
}

}  // namespace mdk
