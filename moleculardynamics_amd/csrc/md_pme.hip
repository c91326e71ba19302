// Smooth particle-mesh Ewald for the SPC/E reciprocal term.
//
// The reference's reciprocal sum (forcefield/ewald.cpp:112-172) runs over
// the k-vectors n with 0 < |n|^2 < n2max, |n_d| <= kmax:
//   E = sum_n C (2 pi/V) e^{-k^2/4a^2} / k^2 |S(n)|^2,  S(n) = sum_i q_i e^{i k.r_i},
// O(N) per k-vector. PME (Essmann et al., J. Chem. Phys. 103, 8577 (1995))
// spreads the charges onto a K0 x K1 x K2 grid with cardinal B-splines of
// order p, so that S(n) ~ +-b(n) F[Q](n) with F[Q](n) = sum_k Q(k)
// e^{2 pi i n.k/K} and |b_d(m)|^2 = 1/|sum_{j<p-1} M_p(j+1) e^{2 pi i m j/K_d}|^2.
// With G(n) = C (2 pi/V) e^{-k^2/4a^2}/k^2 |b(n)|^2 on the same k-set,
//   E = sum_n G(n) |F[Q](n)|^2,
//   phi(k) = dE/dQ(k) = 2 Re sum_n G(n) conj(F[Q](n)) e^{2 pi i n.k/K},
//   F_i = -q_i sum_k phi(k) grad_i prod_d M_p(u_id - k_d),  u = K frac(x/L + 1/2).
// Because only |n_d| <= kmax enters, F[Q] and phi are pruned separable
// transforms, one axis at a time (no FFT, any grid size):
//   spread -> X (K0 -> nb) -> Y (K1 -> nb) -> Z (K2 -> nb, with G, E and
//   P = G conj(F)) -> Z' (nb -> K2) -> Y' (nb -> K1) -> X' (nb -> K0, phi)
//   -> gather (forces), nb = 2 kmax + 1.
// Charges are spread as 64-bit fixed point (2^-48 units), summed with integer
// LDS atomics per grid region, so the grid, and everything after it, does
// not depend on the order of the atomics. Every sum after that has a fixed
// order. All FP64.
#include "md_common.hpp"
#include "md_water.hpp"

namespace mdk {

namespace {

#define GRID_STRIDE(i, n)                                                              \
    for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < (n); \
         i += static_cast<long>(gridDim.x) * blockDim.x)

// spread charges in 2^-48 units: grid sums up to 2^15 in magnitude fit int64
constexpr double kPmeFix = 281474976710656.0;  // 2^48
constexpr double kPmeFixInv = 1.0 / 281474976710656.0;

// round(v) to nearest even as int64 for |v| < 2^51: v + 1.5 2^52 keeps the
// integer in its low mantissa bits, and subtracting the constant's bit
// pattern as an integer leaves it (two's complement for negative v). The
// same result as __double2ll_rn in three instructions.
__device__ __forceinline__ long long fix_round(double v) {
#pragma clang fp contract(off)
    constexpr double kMagic52 = 6755399441055744.0;  // 1.5 2^52
    return __double_as_longlong(v + kMagic52) - __double_as_longlong(kMagic52);
}

inline unsigned blocks(long n) {
    const long g = (n + kBlock - 1) / kBlock;
    return static_cast<unsigned>(g > 0 ? g : 1);
}

__device__ __forceinline__ double2 cmad(double2 acc, double2 a, double2 b) {
    acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
    return acc;
}

// M_p(w + j), j = 0 .. p-1, and dM_p/dt there, M_{p-1}(w + j) -
// M_{p-1}(w + j - 1), for w in [0, 1), by the recursion
// M_n(t) = (t M_{n-1}(t) + (n - t) M_{n-1}(t - 1)) / (n - 1).
__device__ __forceinline__ void bspline(double w, int p, double* m, double* dm) {
    double c[kPmeMaxOrder];
    c[0] = w;
    c[1] = 1.0 - w;
    for (int j = 2; j < p; ++j) c[j] = 0.0;
    for (int n = 3; n <= p; ++n) {
        if (n == p) {
            dm[0] = c[0];
            for (int j = 1; j < p; ++j) dm[j] = c[j] - c[j - 1];
        }
        const double inv = 1.0 / (n - 1);
        // descending j keeps c[j - 1] the order n-1 value when c[j] is replaced
        for (int j = n - 1; j >= 0; --j) {
            const double t = w + j;
            const double a = j < n - 1 ? c[j] : 0.0;
            const double b = j > 0 ? c[j - 1] : 0.0;
            c[j] = (t * a + (n - t) * b) * inv;
        }
    }
    for (int j = 0; j < p; ++j) m[j] = c[j];
}

// Scaled fractional coordinate u = K frac(x / L + 1/2) in [0, K): the
// grid index floor(u) and the spline argument u - floor(u).
__device__ __forceinline__ int frac_index(double x, double L, int K, double& w) {
    double f = x / L + 0.5;
    f -= floor(f);
    double u = f * K;
    if (u >= K) u -= K;
    const double fl = floor(u);
    w = u - fl;
    return static_cast<int>(fl);
}

// Spline weights and their derivatives, once per site (spread and gather
// both read them).
__global__ void k_pme_weights(WaterArgs a) {
    const PmeArgs& P = a.pme;
    const int p = P.p;
    GRID_STRIDE(t, 3 * a.n) {  // t = 3 i + d
        const long i = t / 3;
        const int d = static_cast<int>(t - 3 * i);
        const double4 r = a.pos[i];
        const double x = d == 0 ? r.x : (d == 1 ? r.y : r.z);
        double w, m[kPmeMaxOrder], dm[kPmeMaxOrder];
        P.base[t] = frac_index(x, a.box.L[d], P.K[d], w);
        bspline(w, p, m, dm);
        for (int j = 0; j < p; ++j) {
            P.W[t * p + j] = m[j];
            P.DW[t * p + j] = dm[j];
        }
    }
}

// The charge grid without global atomics. A workgroup owns the region of
// plane z = blockIdx.y, rows y0 .. y0 + hy - 1 (y0 = blockIdx.x hy), all of
// x, and accumulates it in LDS as 64-bit fixed point: it scans the sites in
// chunks of kSpreadChunk, compacts those whose spline support reaches its
// region into an LDS list, then one thread per listed site adds the site's
// points inside the region (one j2 per plane since p <= K2) with integer LDS
// atomics, whose sums do not depend on their order. The region is then
// written out as doubles.
constexpr int kSpreadChunk = 4096;

__global__ __launch_bounds__(kBlock) void k_pme_spread(WaterArgs a, int hy) {
    extern __shared__ long long acc[];  // [hy][K0]
    __shared__ int list[kSpreadChunk];
    __shared__ int nlist;
    const PmeArgs& P = a.pme;
    const int p = P.p, K0 = P.K[0], K1 = P.K[1], K2 = P.K[2];
    const int z = blockIdx.y;
    const int y0 = blockIdx.x * hy;
    const int rows = min(hy, K1 - y0);
    for (int t = threadIdx.x; t < rows * K0; t += kBlock) acc[t] = 0;
    const int lane = threadIdx.x & (kWave - 1);
    for (long c0 = 0; c0 < a.n; c0 += kSpreadChunk) {
        if (threadIdx.x == 0) nlist = 0;
        __syncthreads();
        // compact the chunk's sites that reach the region (whole waves); each
        // thread's site tests are formed first, their loads all in flight
        constexpr int kPer = kSpreadChunk / kBlock;
        bool hits[kPer];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const long i = c0 + u * kBlock + threadIdx.x;
            int by = 0, bz = 0;
            if (i < a.n) {
                by = P.base[3 * i + 1];
                bz = P.base[3 * i + 2];
            }
            int dz = bz - z;
            dz += dz < 0 ? K2 : 0;
            int dy = by - y0;
            dy += dy < 0 ? K1 : 0;
            hits[u] = i < a.n && dz < p && dy < rows + p - 1;
        }
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const long i = c0 + u * kBlock + threadIdx.x;
            const bool hit = hits[u];
            const unsigned long long m = __ballot(hit);
            int at = 0;
            if (lane == 0 && m) at = atomicAdd(&nlist, __popcll(m));
            at = __shfl(at, 0, kWave);
            if (hit) list[at + __popcll(m & ((1ull << lane) - 1))] = static_cast<int>(i);
        }
        __syncthreads();
        const int cnt = nlist;
        for (int e = threadIdx.x; e < cnt; e += kBlock) {
            const long i = list[e];
            const double q = a.pos[i].w;
            const double* W = P.W + 3 * i * p;
            const int b0 = P.base[3 * i], b1 = P.base[3 * i + 1];
            int j2 = P.base[3 * i + 2] - z;
            j2 += j2 < 0 ? K2 : 0;
            const double w2 = W[2 * p + j2];
            for (int j1 = 0; j1 < p; ++j1) {
                int k1 = b1 - j1;
                k1 += k1 < 0 ? K1 : 0;
                const int r = k1 - y0;
                if (r < 0 || r >= rows) continue;
                const double qw = q * w2 * W[p + j1];
                unsigned long long* row = reinterpret_cast<unsigned long long*>(acc + r * K0);
                for (int j0 = 0; j0 < p; ++j0) {
                    int k0 = b0 - j0;
                    k0 += k0 < 0 ? K0 : 0;
                    atomicAdd(row + k0, static_cast<unsigned long long>(fix_round(qw * W[j0] * kPmeFix)));
                }
            }
        }
        __syncthreads();
    }
    double* out = P.q + (static_cast<long>(z) * K1 + y0) * K0;
    for (int t = threadIdx.x; t < rows * K0; t += kBlock) out[t] = static_cast<double>(acc[t]) * kPmeFixInv;
}

// X: A[row][a] = sum_k0 Q[row][k0] tw0[a][k0], row = k2 K1 + k1.
__global__ void k_pme_x(WaterArgs a) {
    const PmeArgs& P = a.pme;
    const int nb = P.nb, K0 = P.K[0];
    const long rows = static_cast<long>(P.K[1]) * P.K[2];
    GRID_STRIDE(t, rows * nb) {
        const long row = t / nb;
        const int c = static_cast<int>(t % nb);
        const double* q = P.q + row * K0;
        const double2* tw = P.twT[0] + c;  // consecutive lanes, consecutive c
        double2 acc = make_double2(0.0, 0.0);
#pragma unroll 8
        for (int k = 0; k < K0; ++k) {
            const double2 t = tw[static_cast<long>(k) * nb];
            acc.x = fma(q[k], t.x, acc.x);
            acc.y = fma(q[k], t.y, acc.y);
        }
        P.A[t] = acc;
    }
}

// Y: B[k2][a][b] = sum_k1 A[k2][k1][a] tw1[b][k1].
__global__ void k_pme_y(WaterArgs a) {
    const PmeArgs& P = a.pme;
    const int nb = P.nb, K1 = P.K[1];
    GRID_STRIDE(t, static_cast<long>(P.K[2]) * nb * nb) {
        const int b = static_cast<int>(t % nb);
        const int c = static_cast<int>((t / nb) % nb);
        const long k2 = t / (static_cast<long>(nb) * nb);
        const double2* src = P.A + (k2 * K1) * nb + c;
        const double2* tw = P.twT[1] + b;
        double2 acc = make_double2(0.0, 0.0);
#pragma unroll 8
        for (int k = 0; k < K1; ++k)
            acc = cmad(acc, src[static_cast<long>(k) * nb], tw[static_cast<long>(k) * nb]);
        P.B[t] = acc;
    }
}

// Z: F[a][b][c] = sum_k2 B[k2][a][b] tw2[c][k2]; E partials (fixed tree per
// workgroup) and P = G conj(F).
__global__ __launch_bounds__(kBlock) void k_pme_z(WaterArgs a) {
    __shared__ double sm[kBlock / kWave];
    const PmeArgs& P = a.pme;
    const int nb = P.nb, K2 = P.K[2];
    const long m = static_cast<long>(nb) * nb * nb;
    const long t = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    double e = 0.0;
    if (t < m) {
        const int c = static_cast<int>(t % nb);
        const long ab = t / nb;
        const double2* src = P.B + ab;
        const double2* tw = P.twT[2] + c;
        double2 f = make_double2(0.0, 0.0);
#pragma unroll 8
        for (int k = 0; k < K2; ++k)
            f = cmad(f, src[static_cast<long>(k) * nb * nb], tw[static_cast<long>(k) * nb]);
        const double g = P.G[t];
        e = g * (f.x * f.x + f.y * f.y);
        P.P[t] = make_double2(g * f.x, -g * f.y);
    }
    const double be = block_reduce(e, sm, wave_sum);
    if (threadIdx.x == 0) a.part2[blockIdx.x] = be;
}

// Z': Cz[k2][a][b] = sum_c P[a][b][c] tw2[c][k2].
__global__ void k_pme_zi(WaterArgs a) {
    const PmeArgs& P = a.pme;
    const int nb = P.nb, K2 = P.K[2];
    GRID_STRIDE(t, static_cast<long>(K2) * nb * nb) {
        const long ab = t % (static_cast<long>(nb) * nb);
        const int k2 = static_cast<int>(t / (static_cast<long>(nb) * nb));
        const double2* src = P.P + ab * nb;
        double2 acc = make_double2(0.0, 0.0);
#pragma unroll 8
        for (int c = 0; c < nb; ++c) acc = cmad(acc, src[c], P.tw[2][static_cast<long>(c) * K2 + k2]);
        P.Cz[t] = acc;
    }
}

// Y': Dy[k2][k1][a] = sum_b Cz[k2][a][b] tw1[b][k1].
__global__ void k_pme_yi(WaterArgs a) {
    const PmeArgs& P = a.pme;
    const int nb = P.nb, K1 = P.K[1];
    GRID_STRIDE(t, static_cast<long>(P.K[2]) * K1 * nb) {
        const int c = static_cast<int>(t % nb);
        const int k1 = static_cast<int>((t / nb) % K1);
        const long k2 = t / (static_cast<long>(nb) * K1);
        const double2* src = P.Cz + (k2 * nb + c) * nb;
        double2 acc = make_double2(0.0, 0.0);
#pragma unroll 8
        for (int b = 0; b < nb; ++b) acc = cmad(acc, src[b], P.tw[1][static_cast<long>(b) * K1 + k1]);
        P.Dy[t] = acc;
    }
}

// X': phi[k2][k1][k0] = 2 Re sum_a Dy[k2][k1][a] tw0[a][k0].
__global__ void k_pme_xi(WaterArgs a) {
    const PmeArgs& P = a.pme;
    const int nb = P.nb, K0 = P.K[0];
    GRID_STRIDE(t, P.ng) {
        const int k0 = static_cast<int>(t % K0);
        const long row = t / K0;
        const double2* src = P.Dy + row * nb;
        double re = 0.0;
#pragma unroll 8
        for (int c = 0; c < nb; ++c) {
            const double2 tw = P.tw[0][static_cast<long>(c) * K0 + k0];
            re = fma(src[c].x, tw.x, fma(-src[c].y, tw.y, re));
        }
        P.phi[t] = 2.0 * re;
    }
}

// Forces: f_i += -q_i sum_k phi(k) grad_i theta_i(k). One wave per site,
// lane l the grid rows (j2, j1) as in the spread; the lanes' sums are
// combined by a fixed shuffle tree.
__global__ __launch_bounds__(kBlock) void k_pme_gather(WaterArgs a) {
    const PmeArgs& P = a.pme;
    const int p = P.p;
    const int lane = threadIdx.x & (kWave - 1);
    const long i = (static_cast<long>(blockIdx.x) * kBlock + threadIdx.x) / kWave;
    if (i >= a.n) return;  // whole waves
    const double* W = P.W + 3 * i * p;
    const double* D = P.DW + 3 * i * p;
    const int b0 = P.base[3 * i], b1 = P.base[3 * i + 1], b2 = P.base[3 * i + 2];
    double gx = 0.0, gy = 0.0, gz = 0.0;
    for (int l = lane; l < p * p; l += kWave) {
        const int j2 = l / p, j1 = l % p;
        int k2 = b2 - j2;
        k2 += k2 < 0 ? P.K[2] : 0;
        int k1 = b1 - j1;
        k1 += k1 < 0 ? P.K[1] : 0;
        const double* row = P.phi + (static_cast<long>(k2) * P.K[1] + k1) * P.K[0];
        double s0 = 0.0, s1 = 0.0;
        for (int j0 = 0; j0 < p; ++j0) {
            int k0 = b0 - j0;
            k0 += k0 < 0 ? P.K[0] : 0;
            const double ph = row[k0];
            s0 = fma(ph, D[j0], s0);
            s1 = fma(ph, W[j0], s1);
        }
        gx = fma(s0, W[p + j1] * W[2 * p + j2], gx);
        gy = fma(s1, D[p + j1] * W[2 * p + j2], gy);
        gz = fma(s1, W[p + j1] * D[2 * p + j2], gz);
    }
    gx = wave_sum(gx);
    gy = wave_sum(gy);
    gz = wave_sum(gz);
    if (lane == 0) {
        const double q = a.pos[i].w;
        double4 f = a.frc[i];
        f.x -= q * gx * (P.K[0] / a.box.L[0]);
        f.y -= q * gy * (P.K[1] / a.box.L[1]);
        f.z -= q * gz * (P.K[2] / a.box.L[2]);
        a.frc[i] = f;
    }
}

}  // namespace

// Rows per spread workgroup: 16, fewer when a row is long (at most 4,096
// fixed-point accumulators, 32 KB of LDS).
int pme_spread_rows(const PmeArgs& p) {
    int hy = 16;
    while (hy > 1 && hy * p.K[0] > 4096) hy >>= 1;
    return hy;
}

long pme_energy_blocks(const PmeArgs& p) {
    return static_cast<long>(blocks(static_cast<long>(p.nb) * p.nb * p.nb));
}

void water_launch_pme(const WaterArgs& a, hipStream_t s) {
    const PmeArgs& p = a.pme;
    const long nbk = static_cast<long>(p.nb);
    hipLaunchKernelGGL(k_pme_weights, dim3(blocks(3 * a.n)), dim3(kBlock), 0, s, a);
    const int hy = pme_spread_rows(p);
    hipLaunchKernelGGL(k_pme_spread, dim3(static_cast<unsigned>((p.K[1] + hy - 1) / hy), p.K[2]),
                       dim3(kBlock), static_cast<size_t>(hy) * p.K[0] * sizeof(long long), s, a, hy);
    hipLaunchKernelGGL(k_pme_x, dim3(blocks(static_cast<long>(p.K[1]) * p.K[2] * nbk)), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_pme_y, dim3(blocks(p.K[2] * nbk * nbk)), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_pme_z, dim3(static_cast<unsigned>(pme_energy_blocks(p))), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_pme_zi, dim3(blocks(p.K[2] * nbk * nbk)), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_pme_yi, dim3(blocks(static_cast<long>(p.K[2]) * p.K[1] * nbk)), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_pme_xi, dim3(blocks(p.ng)), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_pme_gather, dim3(blocks(a.n * kWave)), dim3(kBlock), 0, s, a);
}

}  // namespace mdk
