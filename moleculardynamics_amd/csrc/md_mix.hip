// Per-type-pair LJ on the device: list build in upload order, forces with
// energy/virial and their parameter derivatives, the adjoint pass (force
// vector-Jacobian product plus the gradients of E and W), deterministic
// reductions, steepest descent and the adjoint's per-atom vector updates.
//
// Pair terms (FP64, operation order of lennard_jones.cpp:50-88, per ordered
// type pair k = type_i * nt + type_j, with sigma, epsilon, cutoff of k):
//   s2 = sigma^2 / r2, s6 = s2^3, u = 4 eps (s6^2 - s6), w = 24 eps (2 s6^2 - s6)
//   LFS: u -= u(rc) + (r - rc) u'(rc), w += r u'(rc)
//   c = w / r2 and F_i += c d_ij with d_ij the minimum image of r_i - r_j.
// Derivatives at a fixed cutoff, y = 24 eps (24 s6^2 - 6 s6) = sigma dw/dsigma:
//   dE/deps = u/eps, dW/deps = w/eps, dc/deps = c/eps
//   dE/dsigma = [w_t - (w(rc) - (r - rc) y(rc)/rc)_LFS] / sigma   (w_t: truncated w)
//   dW/dsigma = [y - (r y(rc)/rc)_LFS] / sigma,  dc/dsigma = (dW/dsigma) / r2
// Second derivatives for J = dF/dx (symmetric, J = -Hessian):
//   dc/d(r2) = 24 eps (-14 s6^2 + 4 s6) / r2^2  [- u'(rc) / (2 r^3)]_LFS
//   dw/d(r2) = 24 eps (-12 s6^2 + 3 s6) / r2    [+ u'(rc) / (2 r)]_LFS
//   (J^T v)_i = sum_j [c (v_i - v_j) + 2 c' d_ij (d_ij . (v_i - v_j))]
//   grad_i W = sum_j 2 w' d_ij,  grad_i E = -F_i.
// Each unordered pair's scalar sums (E, W, the tables) are taken once, in
// the row of its lower upload index; forces are summed per row.
#include "md_mix.hpp"

namespace mdk {

namespace {

struct PairTerms {
    double dx, dy, dz;  // d_ij = r_i - r_j (minimum image)
    double r2, s6, u, w, wt, c;
    double r;           // sqrt(r2) (LFS only)
    int k;
    bool lower;         // i is the pair's lower upload index
};

// Returns false when the pair is outside its type pair's cutoff.
template <bool LFS>
__device__ __forceinline__ bool mix_terms(const MixTab& T, const Box& box, const double4& pi, int ti,
                                          const double4& pj, uint32_t e, PairTerms& p) {
#pragma clang fp contract(off)
    const bool flip = (e & kFlipBit) != 0;
    const int tj = static_cast<int>(pj.w);
    const int k = ti * T.nt + tj;
    double dx, dy, dz;
    pair_delta(box, pi, pj, flip, dx, dy, dz);
    const double r2 = norm2_ref(dx, dy, dz);
    if (!(r2 < T.rc2[k])) return false;
    const double s2 = T.sig2[k] / r2;
    const double s6 = s2 * s2 * s2;
    const double eps = T.eps[k];
    double u = 4.0 * eps * (s6 * s6 - s6);
    const double wt = 24.0 * eps * (2.0 * s6 * s6 - s6);
    double w = wt;
    double r = 0.0;
    if (LFS) {
        r = sqrt(r2);
        u = u - T.urc[k] - (r - T.rc[k]) * T.dudr[k];
        w = w + r * T.dudr[k];
    }
    p.dx = flip ? -dx : dx;
    p.dy = flip ? -dy : dy;
    p.dz = flip ? -dz : dz;
    p.r2 = r2;
    p.s6 = s6;
    p.u = u;
    p.w = w;
    p.wt = wt;
    p.c = w / r2;
    p.r = r;
    p.k = k;
    p.lower = !flip;
    return true;
}

template <bool LFS>
__device__ __forceinline__ void sigma_terms(const MixTab& T, const PairTerms& p, double& dEds,
                                            double& dWds) {
#pragma clang fp contract(off)
    const double eps = T.eps[p.k];
    const double y = 24.0 * eps * (24.0 * p.s6 * p.s6 - 6.0 * p.s6);
    double a = p.wt, b = y;
    if (LFS) {
        a = a - (T.wrc[p.k] - (p.r - T.rc[p.k]) * T.yrc_rc[p.k]);
        b = b - p.r * T.yrc_rc[p.k];
    }
    dEds = a * T.inv_sig[p.k];
    dWds = b * T.inv_sig[p.k];
}

// ---- list ----------------------------------------------------------------

constexpr int kListLanes = 8;  // lanes per atom in the list build

// kListLanes lanes per atom test consecutive candidates together and keep
// them in candidate order by a ballot and a prefix count, so the row is the
// row of the serial scan: with a usable grid (cells of at least rc + skin
// over 2, 5 or more per axis) the 5 x 5 runs of five x-adjacent cells around
// the atom's cell in (z, y) order -- contiguous in the cell-sorted index,
// split in two where a run wraps the x axis -- atoms in index order within a
// cell; every atom otherwise. Every pair within rc + skin lies in such a run.
__global__ __launch_bounds__(kBlock) void k_mix_list(MixListArgs a) {
    const int tid = threadIdx.x;
    const int sub = tid % kListLanes;
    const int gshift = (tid & (kWave - 1)) & ~(kListLanes - 1);
    const unsigned below = (1u << sub) - 1u;
    const long groups = static_cast<long>(gridDim.x) * (kBlock / kListLanes);
    for (long il = static_cast<long>(blockIdx.x) * (kBlock / kListLanes) + tid / kListLanes; il < a.n;
         il += groups) {
        const int i = static_cast<int>(il);
        const double4 pi = a.pos[i];
        uint32_t* row = a.nbr + nbr_index(i, 0, a.cap);
        int cnt = 0;
        auto scan = [&](const int* idx, int b, int e) {
            for (int k0 = b; k0 < e; k0 += kListLanes) {
                const int k = k0 + sub;
                bool keep = false;
                int j = 0;
                if (k < e) {
                    j = idx != nullptr ? idx[k] : k;
                    if (j != i) {
                        double dx, dy, dz;
                        pair_delta(a.box, pi, a.pos[j], i > j, dx, dy, dz);
                        keep = norm2_ref(dx, dy, dz) < a.rl2;
                    }
                }
                const unsigned gm =
                    static_cast<unsigned>(__ballot(keep) >> gshift) & ((1u << kListLanes) - 1u);
                if (keep) {
                    const int at = cnt + __popc(gm & below);
                    if (at < a.cap) row[at * 64] = static_cast<uint32_t>(j) | (i > j ? kFlipBit : 0u);
                }
                cnt += __popc(gm);
            }
        };
        const Grid& g = a.grid;
        if (g.usable) {
            const int nx = g.nc[0], ny = g.nc[1], nz = g.nc[2];
            const int ci = a.cell_of[i];
            const int cx = ci % nx, cy = (ci / nx) % ny, cz = ci / (nx * ny);
            const int xa = cx >= 2 ? cx - 2 : 0, xb = cx + 2 < nx ? cx + 2 : nx - 1;
            for (int oz = -2; oz <= 2; ++oz) {
                int zz = cz + oz;
                zz += zz < 0 ? nz : (zz >= nz ? -nz : 0);
                for (int oy = -2; oy <= 2; ++oy) {
                    int yy = cy + oy;
                    yy += yy < 0 ? ny : (yy >= ny ? -ny : 0);
                    const int rb = nx * (yy + ny * zz);
                    scan(a.cell_atoms, a.cell_start[rb + xa], a.cell_start[rb + xb + 1]);
                    if (cx < 2)
                        scan(a.cell_atoms, a.cell_start[rb + nx + cx - 2], a.cell_start[rb + nx]);
                    else if (cx + 2 >= nx)
                        scan(a.cell_atoms, a.cell_start[rb], a.cell_start[rb + cx + 3 - nx]);
                }
            }
        } else {
            scan(nullptr, 0, static_cast<int>(a.n));
        }
        if (sub == 0) {
            if (a.ref != nullptr) a.ref[i] = pi;
            a.ncnt[i] = cnt;
            if (cnt > a.cap) atomicOr(&a.sc->overflow, 1);
        }
    }
}

// kMixTpa lanes per atom in the force and adjoint passes: lane s of an atom
// takes row entries s, s + kMixTpa, ..., and the lanes' sums are combined by
// a fixed shuffle tree (lanes of an atom are adjacent).
constexpr int kMixTpa = 8;

__device__ __forceinline__ double mix_lane_sum(double v) {
    for (int off = kMixTpa / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// ---- forces (+ parameter derivatives) ---------------------------------------

template <bool LFS, bool DERIV>
__global__ __launch_bounds__(kBlock) void k_mix_force(MixArgs a) {
#pragma clang fp contract(off)
    const long gt = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    const long i = gt / kMixTpa;
    const int sub = static_cast<int>(gt % kMixTpa);
    const bool valid = i < a.n;
    const MixTab& T = a.tab;
    const double4 pi = valid ? a.pos[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    const int ti = static_cast<int>(pi.w);
    const int cnt = valid ? min(a.ncnt[i], a.cap) : 0;
    const uint32_t* row = a.nbr + nbr_index(valid ? i : 0, 0, a.cap);
    double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0, w = 0.0;
    double q[kMaxTypes][kMixQ];
#pragma unroll
    for (int b = 0; b < kMaxTypes; ++b)
#pragma unroll
        for (int s = 0; s < kMixQ; ++s) q[b][s] = 0.0;
    for (int m = sub; m < cnt; m += kMixTpa) {
        const uint32_t en = row[m * 64];
        const double4 pj = a.pos[en & kIndexMask];
        PairTerms p;
        if (!mix_terms<LFS>(T, a.box, pi, ti, pj, en, p)) continue;
        fx += p.c * p.dx;
        fy += p.c * p.dy;
        fz += p.c * p.dz;
        if (!p.lower) continue;
        e += p.u;
        w += p.w;
        if (DERIV) {
            double dEds, dWds;
            sigma_terms<LFS>(T, p, dEds, dWds);
            const double ie = T.inv_eps[p.k];
            const int tj = static_cast<int>(pj.w);
#pragma unroll
            for (int b = 0; b < kMaxTypes; ++b) {
                if (b == tj) {
                    q[b][0] += p.u;
                    q[b][1] += p.w;
                    q[b][2] += p.u * ie;
                    q[b][3] += dEds;
                    q[b][4] += p.w * ie;
                    q[b][5] += dWds;
                }
            }
        }
    }
    fx = mix_lane_sum(fx);
    fy = mix_lane_sum(fy);
    fz = mix_lane_sum(fz);
    e = mix_lane_sum(e);
    w = mix_lane_sum(w);
    if (DERIV) {
#pragma unroll
        for (int b = 0; b < kMaxTypes; ++b)
#pragma unroll
            for (int s = 0; s < kMixQ; ++s) q[b][s] = mix_lane_sum(q[b][s]);
    }
    if (!valid || sub != 0) return;
    a.frc[i] = make_double4(fx, fy, fz, 0.0);
    a.part_ew[2 * i] = e;
    a.part_ew[2 * i + 1] = w;
    if (DERIV) {
        double* o = a.part_q + i * (kMaxTypes * kMixQ);
#pragma unroll
        for (int b = 0; b < kMaxTypes; ++b)
#pragma unroll
            for (int s = 0; s < kMixQ; ++s) o[b * kMixQ + s] = q[b][s];
    }
}

// ---- adjoint pass -------------------------------------------------------------

template <bool LFS>
__global__ __launch_bounds__(kBlock) void k_mix_adjoint(MixArgs a) {
#pragma clang fp contract(off)
    const long gt = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    const long i = gt / kMixTpa;
    const int sub = static_cast<int>(gt % kMixTpa);
    const bool valid = i < a.n;
    const MixTab& T = a.tab;
    const double4 pi = valid ? a.pos[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    const double4 vi = valid ? a.vec[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    const int ti = static_cast<int>(pi.w);
    const int cnt = valid ? min(a.ncnt[i], a.cap) : 0;
    const uint32_t* row = a.nbr + nbr_index(valid ? i : 0, 0, a.cap);
    const double wE = a.w_energy, wW = a.w_virial;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    double G[kMaxTypes][kMixG];
#pragma unroll
    for (int b = 0; b < kMaxTypes; ++b) G[b][0] = G[b][1] = 0.0;
    for (int m = sub; m < cnt; m += kMixTpa) {
        const uint32_t en = row[m * 64];
        const uint32_t j = en & kIndexMask;
        const double4 pj = a.pos[j];
        PairTerms p;
        if (!mix_terms<LFS>(T, a.box, pi, ti, pj, en, p)) continue;
        const double4 vj = a.vec[j];
        const double ax = vi.x - vj.x, ay = vi.y - vj.y, az = vi.z - vj.z;
        const double dv = (p.dx * ax + p.dy * ay) + p.dz * az;
        const double eps = T.eps[p.k];
        const double inv = 1.0 / p.r2;
        const double s12 = p.s6 * p.s6;
        double cp = 24.0 * eps * (-14.0 * s12 + 4.0 * p.s6) * (inv * inv);
        double wp = 24.0 * eps * (-12.0 * s12 + 3.0 * p.s6) * inv;
        if (LFS) {
            cp = cp - 0.5 * T.dudr[p.k] * inv / p.r;
            wp = wp + 0.5 * T.dudr[p.k] / p.r;
        }
        // J^T v (row i of the symmetric Jacobian applied to v)
        const double h = 2.0 * cp * dv;
        gx += p.c * ax + h * p.dx;
        gy += p.c * ay + h * p.dy;
        gz += p.c * az + h * p.dz;
        // direct terms: wE grad E = -wE F, wW grad W = wW 2 w' d
        const double dcoef = -wE * p.c + wW * (2.0 * wp);
        gx += dcoef * p.dx;
        gy += dcoef * p.dy;
        gz += dcoef * p.dz;
        if (!p.lower) continue;
        double dEds, dWds;
        sigma_terms<LFS>(T, p, dEds, dWds);
        const double ie = T.inv_eps[p.k];
        const double g_eps = (p.c * ie) * dv + wE * (p.u * ie) + wW * (p.w * ie);
        const double g_sig = (dWds / p.r2) * dv + wE * dEds + wW * dWds;
        const int tj = static_cast<int>(pj.w);
#pragma unroll
        for (int b = 0; b < kMaxTypes; ++b) {
            if (b == tj) {
                G[b][0] += g_eps;
                G[b][1] += g_sig;
            }
        }
    }
    gx = mix_lane_sum(gx);
    gy = mix_lane_sum(gy);
    gz = mix_lane_sum(gz);
#pragma unroll
    for (int b = 0; b < kMaxTypes; ++b) {
        G[b][0] = mix_lane_sum(G[b][0]);
        G[b][1] = mix_lane_sum(G[b][1]);
    }
    if (!valid || sub != 0) return;
    double4 o = a.accumulate ? a.out[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    o.x += gx;
    o.y += gy;
    o.z += gz;
    a.out[i] = o;
    double* pq = a.part_q + i * (kMaxTypes * kMixG);
#pragma unroll
    for (int b = 0; b < kMaxTypes; ++b) {
        pq[b * kMixG] = G[b][0];
        pq[b * kMixG + 1] = G[b][1];
    }
}

// ---- reductions ---------------------------------------------------------------

// The per-atom reductions below run in one workgroup of kSumThreads
// threads. Thread t sums atoms t, t + kSumThreads, ... into kSumWays
// interleaved partial sums (independent loads in flight), which are then
// combined in a fixed order, the threads by wave shuffles and the waves in
// index order: the result does not depend on timing.
constexpr int kSumThreads = 1024;
constexpr int kSumWays = 4;

__device__ double sum_fixed(double v, double* sm) {
    v = wave_sum(v);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) sm[wid] = v;
    __syncthreads();
    double r = 0.0;
    for (int w = 0; w < kSumThreads / kWave; ++w) r += sm[w];
    return r;
}

__global__ __launch_bounds__(kSumThreads) void k_mix_sum_ew(long n, const double* __restrict__ part,
                                                            double* e_out, double* w_out) {
    __shared__ double sm[kSumThreads / kWave];
    double e[kSumWays] = {}, w[kSumWays] = {};
    for (long i0 = threadIdx.x; i0 < n; i0 += static_cast<long>(kSumThreads) * kSumWays) {
#pragma unroll
        for (int k = 0; k < kSumWays; ++k) {
            const long i = i0 + static_cast<long>(k) * kSumThreads;
            if (i < n) {
                e[k] += part[2 * i];
                w[k] += part[2 * i + 1];
            }
        }
    }
    const double be = sum_fixed((e[0] + e[1]) + (e[2] + e[3]), sm);
    const double bw = sum_fixed((w[0] + w[1]) + (w[2] + w[3]), sm);
    if (threadIdx.x == 0) {
        *e_out = be;
        *w_out = bw;
    }
}

// One workgroup per (a, b, k): sum over atoms of type a of part_q[i][b][k].
__global__ __launch_bounds__(kSumThreads) void k_mix_tables(long n, const double4* __restrict__ pos,
                                                            int nt, int q,
                                                            const double* __restrict__ part,
                                                            double* tables, int accumulate) {
    __shared__ double sm[kSumThreads / kWave];
    const int id = blockIdx.x;  // (a * nt + b) * q + k
    const int k = id % q;
    const int b = (id / q) % nt;
    const int ta = id / (q * nt);
    const long stride = static_cast<long>(kMaxTypes) * q;
    double s[kSumWays] = {};
    for (long i0 = threadIdx.x; i0 < n; i0 += static_cast<long>(kSumThreads) * kSumWays) {
#pragma unroll
        for (int m = 0; m < kSumWays; ++m) {
            const long i = i0 + static_cast<long>(m) * kSumThreads;
            if (i < n && static_cast<int>(pos[i].w) == ta) s[m] += part[i * stride + b * q + k];
        }
    }
    const double r = sum_fixed((s[0] + s[1]) + (s[2] + s[3]), sm);
    if (threadIdx.x == 0) tables[id] = accumulate ? tables[id] + r : r;
}

// ---- steepest descent -----------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_mix_fmax(long n, const double4* __restrict__ frc,
                                                     double* part) {
    __shared__ double sm[kBlock / kWave];
    const long i = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    double m = 0.0;
    if (i < n) {
        const double4 f = frc[i];
        m = sqrt(norm2_ref(f.x, f.y, f.z));
    }
    for (int off = kWave / 2; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0) sm[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) b = fmax(b, sm[w]);
        part[blockIdx.x] = b;
    }
}

__global__ __launch_bounds__(kBlock) void k_mix_gamma(const double* __restrict__ part, int nb,
                                                      double step, double max_disp, double* gamma) {
    __shared__ double sm[kBlock];
    double m = 0.0;
    for (int b = threadIdx.x; b < nb; b += kBlock) m = fmax(m, part[b]);
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double fmx = sm[0];
        double g = step;
        if (fmx > 0.0) {
            const double cap = max_disp / fmx;
            if (cap < g) g = cap;
        }
        *gamma = g;
    }
}

__global__ void k_mix_move(long n, Box box, double4* __restrict__ pos, const double4* __restrict__ frc,
                           const double* gamma) {
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double g = *gamma;
    double4 p = pos[i];
    const double4 f = frc[i];
    p.x = fold(fma(g, f.x, p.x), box.L[0], box.lo[0], box.hi1[0], box.hi2[0]);
    p.y = fold(fma(g, f.y, p.y), box.L[1], box.lo[1], box.hi1[1], box.hi2[1]);
    p.z = fold(fma(g, f.z, p.z), box.L[2], box.lo[2], box.hi1[2], box.hi2[2]);
    pos[i] = p;
}

// ---- adjoint vector updates ---------------------------------------------------

__global__ void k_mix_vec(long n, const double4* __restrict__ vel, const double4* __restrict__ lv,
                          const double4* __restrict__ lu, double half_dt, int frame0,
                          double4* __restrict__ vec) {
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a = half_dt * (1.0 / vel[i].w);
    const double4 u = lu[i];
    double4 r;
    if (frame0) {
        r = make_double4(a * u.x, a * u.y, a * u.z, 0.0);
    } else {
        const double4 v = lv[i];
        r = make_double4(a * (v.x + u.x), a * (v.y + u.y), a * (v.z + u.z), 0.0);
    }
    vec[i] = r;
}

__global__ void k_mix_back(long n, double dt, const double4* __restrict__ lx,
                           double4* __restrict__ lv, double4* __restrict__ lu) {
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 x = lx[i];
    const double4 v = lv[i];
    const double4 u = make_double4(v.x + dt * x.x, v.y + dt * x.y, v.z + dt * x.z, 0.0);
    lv[i] = u;
    lu[i] = u;
}

inline unsigned blocks(long n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

}  // namespace

void launch_mix_list(const MixListArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_mix_list, dim3(blocks(a.n * kListLanes)), dim3(kBlock), 0, s, a);
}

void launch_mix_force(const MixArgs& a, int deriv, hipStream_t s) {
    const dim3 g(blocks(a.n * kMixTpa)), b(kBlock);
    if (a.tab.lfs) {
        if (deriv) hipLaunchKernelGGL((k_mix_force<true, true>), g, b, 0, s, a);
        else hipLaunchKernelGGL((k_mix_force<true, false>), g, b, 0, s, a);
    } else {
        if (deriv) hipLaunchKernelGGL((k_mix_force<false, true>), g, b, 0, s, a);
        else hipLaunchKernelGGL((k_mix_force<false, false>), g, b, 0, s, a);
    }
}

void launch_mix_adjoint(const MixArgs& a, hipStream_t s) {
    const dim3 g(blocks(a.n * kMixTpa)), b(kBlock);
    if (a.tab.lfs) hipLaunchKernelGGL(k_mix_adjoint<true>, g, b, 0, s, a);
    else hipLaunchKernelGGL(k_mix_adjoint<false>, g, b, 0, s, a);
}

void launch_mix_sum_ew(long n, const double* part_ew, double* e_out, double* w_out, hipStream_t s) {
    hipLaunchKernelGGL(k_mix_sum_ew, dim3(1), dim3(kSumThreads), 0, s, n, part_ew, e_out, w_out);
}

void launch_mix_tables(long n, const double4* pos, int nt, int q, const double* part_q,
                       double* tables, int accumulate, hipStream_t s) {
    hipLaunchKernelGGL(k_mix_tables, dim3(nt * nt * q), dim3(kSumThreads), 0, s, n, pos, nt, q,
                       part_q, tables, accumulate);
}

void launch_mix_sd(long n, const Box& box, double4* pos, const double4* frc, double step,
                   double max_disp, double* part_d, double* gamma, hipStream_t s) {
    const unsigned nb = blocks(n);
    hipLaunchKernelGGL(k_mix_fmax, dim3(nb), dim3(kBlock), 0, s, n, frc, part_d);
    hipLaunchKernelGGL(k_mix_gamma, dim3(1), dim3(kBlock), 0, s, part_d, static_cast<int>(nb), step,
                       max_disp, gamma);
    hipLaunchKernelGGL(k_mix_move, dim3(nb), dim3(kBlock), 0, s, n, box, pos, frc, gamma);
}

void launch_mix_vec(long n, const double4* vel, const double4* lv, const double4* lu,
                    double half_dt, int frame0, double4* vec, hipStream_t s) {
    hipLaunchKernelGGL(k_mix_vec, dim3(blocks(n)), dim3(kBlock), 0, s, n, vel, lv, lu, half_dt,
                       frame0, vec);
}

void launch_mix_back(long n, double dt, const double4* lx, double4* lv, double4* lu, hipStream_t s) {
    hipLaunchKernelGGL(k_mix_back, dim3(blocks(n)), dim3(kBlock), 0, s, n, dt, lx, lv, lu);
}

}  // namespace mdk
