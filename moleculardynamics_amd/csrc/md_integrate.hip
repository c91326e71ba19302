// Integrator passes and per-atom utility kernels.
//
// k_pre: everything an integrator does to an atom between two force calls
// that does not need the new force, fused into one pass over pos/vel/frc:
//   velocity Verlet (integrate/velocity_verlet.hpp:48-64):
//     v = fma(h/m, f, v); r = wrap(fma(dt, v, r))
//   Nose-Hoover chain (integrate/nose_hoover.hpp:37-42): the pending chain
//     rescale of the previous half-chains is applied first, then as above
//   BAOAB (integrate/langevin.hpp:36-71):
//     B: v = fma(h/m, f, v); A: r = wrap(fma(h, v, r));
//     O: v = c1 v + (c2 sqrt(T/m)) xi; A: r = wrap(fma(h, v, r))
// followed by the Verlet-list displacement check (verlet_list.cpp:44-62):
// the top-2 of |minimumImage(r - r_ref)| reduced by a fixed tree, and the
// rebuild flag set on the device when their sum exceeds the skin. The
// second half-kick runs inside the force kernel.
//
// BAOAB noise: Philox4x32-10 (Salmon, Moraes, Dror & Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC11), key = seed, counter = (upload index,
// step, stream), two 32-bit words per 53-bit uniform in (0,1), and the
// Box-Muller transform. The stream is a function of the atom's upload
// index, so it does not depend on the device sort order.
#include "md_common.hpp"
#include "md_kernels.hpp"

namespace mdk {

namespace {


__device__ __forceinline__ double displacement(const Box& b, const double4& p, const double4& r) {
#pragma clang fp contract(off)
    double dx = p.x - r.x, dy = p.y - r.y, dz = p.z - r.z;
    min_image(b, dx, dy, dz);
    return sqrt(norm2_ref(dx, dy, dz));
}

// The decision from nb per-workgroup top-2 partials, by one workgroup
// (blockDim.x threads: a multiple of the wave size, at most kFinish):
// strided shares, wave top-2 by shuffles, then the waves' pairs merged by
// thread 0.
constexpr int kFinish = 1024;

__device__ void finish_check(const StepArgs& a, int nb) {
    __shared__ double t1s[kFinish / kWave], t2s[kFinish / kWave];
    double t1 = 0.0, t2 = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        top2_merge(t1, t2, a.part_d[2 * b], a.part_d[2 * b + 1]);
    wave_top2(t1, t2);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        t1s[threadIdx.x / kWave] = t1;
        t2s[threadIdx.x / kWave] = t2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double b1 = 0.0, b2 = 0.0;
        for (int w = 0; w < static_cast<int>(blockDim.x / kWave); ++w) top2_merge(b1, b2, t1s[w], t2s[w]);
        a.sc->top1 = b1;
        a.sc->top2 = b2;
        a.sc->rebuild_flag = (b1 + b2) > a.skin ? 1 : 0;
        a.sc->vscale = 1.0;
    }
}

template <int KIND>
__global__ __launch_bounds__(kBlock) void k_pre(StepArgs a) {
    __shared__ double sm1[kBlock / kWave], sm2[kBlock / kWave];
    if (a.sc->overflow) return;  // a truncated list was built: md_run restores and retries
    const long t = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    double d = 0.0;
    if (t < a.n) {
        double4 p = a.pos[t];
        double4 v = a.vel[t];
        const double4 f = a.frc[t];
        if (KIND == 2) {
            const double s = a.sc->vscale;
            v.x = v.x * s;
            v.y = v.y * s;
            v.z = v.z * s;
        }
        const double c = (0.5 * a.dt) * (1.0 / v.w);
        v.x = fma(c, f.x, v.x);
        v.y = fma(c, f.y, v.y);
        v.z = fma(c, f.z, v.z);
        if (KIND == 1) {
            const double h = 0.5 * a.dt;
            p = drift(a.box, p, v, h);
            double g0, g1, g2;
            normals3(a.seed, a.dev_step ? a.sc->step : a.step, static_cast<uint32_t>(a.orig[t]), g0,
                     g1, g2);
            const double k = a.c2 * sqrt(a.temperature / v.w);
            v.x = a.c1 * v.x + g0 * k;
            v.y = a.c1 * v.y + g1 * k;
            v.z = a.c1 * v.z + g2 * k;
            p = drift(a.box, p, v, h);
        } else {
            p = drift(a.box, p, v, a.dt);
        }
        a.pos[t] = p;
        a.vel[t] = v;
        if (a.qpos != nullptr) a.qpos[t] = quantize(a.quant, p, a.orig[t]);
        if (a.check) d = displacement(a.box, p, a.ref[t]);
    }
    double t1 = d, t2 = 0.0;
    wave_top2(t1, t2);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    if (lane == 0) {
        sm1[wid] = t1;
        sm2[wid] = t2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double b1 = 0.0, b2 = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) top2_merge(b1, b2, sm1[w], sm2[w]);
        a.part_d[2 * blockIdx.x] = b1;
        a.part_d[2 * blockIdx.x + 1] = b2;
    }
}

// One workgroup after k_pre: rebuild decision from the partials; the chain
// rescale applied by k_pre is consumed.
__global__ __launch_bounds__(kFinish) void k_pre_finish(StepArgs a, int nb) {
    if (a.sc->overflow) {  // no further builds once a row has overflowed
        if (threadIdx.x == 0) a.sc->rebuild_flag = 0;
        return;
    }
    if (a.check) {
        finish_check(a, nb);
    } else if (threadIdx.x == 0) {
        a.sc->vscale = 1.0;
    }
}

__global__ __launch_bounds__(kBlock) void k_check(StepArgs a) {
    __shared__ double sm1[kBlock / kWave], sm2[kBlock / kWave];
    const long t = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    double d = 0.0;
    if (t < a.n) d = displacement(a.box, a.pos[t], a.ref[t]);
    double t1 = d, t2 = 0.0;
    wave_top2(t1, t2);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    if (lane == 0) {
        sm1[wid] = t1;
        sm2[wid] = t2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double b1 = 0.0, b2 = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) top2_merge(b1, b2, sm1[w], sm2[w]);
        a.part_d[2 * blockIdx.x] = b1;
        a.part_d[2 * blockIdx.x + 1] = b2;
    }
}

__global__ __launch_bounds__(kFinish) void k_check_finish(StepArgs a, int nb) {
    const double keep = a.sc->vscale;
    finish_check(a, nb);
    if (threadIdx.x == 0) a.sc->vscale = keep;
}

__global__ void k_apply_scale(long n, double4* vel, Scalars* sc) {
    const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const double s = sc->vscale;
    if (t < n) {
        double4 v = vel[t];
        v.x = v.x * s;
        v.y = v.y * s;
        v.z = v.z * s;
        vel[t] = v;
    }
}

__global__ void k_reset_scale(Scalars* sc) {
    if (threadIdx.x == 0) sc->vscale = 1.0;
}

__global__ __launch_bounds__(kBlock) void k_kinetic(long n, const double4* vel, double* part,
                                                    Scalars* sc) {
    __shared__ double sm[kBlock / kWave];
    __shared__ double red[kBlock];
    const long t = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    double ke = 0.0;
    if (t < n) {
        const double4 v = vel[t];
        ke = (0.5 * v.w) * norm2_ref(v.x, v.y, v.z);
    }
    const double b = block_reduce(ke, sm, wave_sum);
    if (threadIdx.x == 0) part[blockIdx.x] = b;
    (void)red;
    (void)sc;
}

__global__ __launch_bounds__(kBlock) void k_kinetic_finish(const double* part, int nb, Scalars* sc) {
    __shared__ double red[kBlock];
    double s = 0.0;
    for (int k = threadIdx.x; k < nb; k += kBlock) s += part[k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) sc->kinetic = red[0];
}

// Mass and momentum sums (core/system.cpp:137-164) into out4 = {M, Px, Py, Pz}.
__global__ __launch_bounds__(kBlock) void k_momentum(long n, const double4* vel, double* part,
                                                     double* out4, unsigned int* ticket) {
    __shared__ double sm[kBlock / kWave];
    __shared__ double red[4][kBlock];
    const long t = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    double q[4] = {0, 0, 0, 0};
    if (t < n) {
        const double4 v = vel[t];
        q[0] = v.w;
        q[1] = v.w * v.x;
        q[2] = v.w * v.y;
        q[3] = v.w * v.z;
    }
    for (int k = 0; k < 4; ++k) {
        const double b = block_reduce(q[k], sm, wave_sum);
        if (threadIdx.x == 0) part[4 * blockIdx.x + k] = b;
    }
    (void)red;
    (void)out4;
    (void)ticket;
}

__global__ __launch_bounds__(kBlock) void k_momentum_finish(const double* part, int nb, double* out4) {
    __shared__ double red[4][kBlock];
    for (int k = 0; k < 4; ++k) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nb; b += kBlock) s += part[4 * b + k];
        red[k][threadIdx.x] = s;
    }
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) out4[k] = red[k][0];
}

__global__ void k_sub_com(long n, double4* vel, const double* out4) {
    const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const double M = out4[0];
    if (t >= n || M == 0.0) return;
    const double cx = out4[1] / M, cy = out4[2] / M, cz = out4[3] / M;
    double4 v = vel[t];
    v.x -= cx;
    v.y -= cy;
    v.z -= cz;
    vel[t] = v;
}

__global__ void k_unpermute(long n, const int* __restrict__ orig, const double4* __restrict__ a,
                            double* __restrict__ x, double* __restrict__ y, double* __restrict__ z) {
    const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double4 v = a[t];
    const int o = orig[t];
    x[o] = v.x;
    y[o] = v.y;
    z[o] = v.z;
}

__global__ void k_permute_in(long n, const int* __restrict__ orig, const double* __restrict__ x,
                             const double* __restrict__ y, const double* __restrict__ z,
                             double4* __restrict__ a, int w_mode) {
    const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int o = orig[t];
    double4 v = a[t];
    v.x = x[o];
    v.y = y[o];
    v.z = z[o];
    if (w_mode == kPermuteWZero) v.w = 0.0;
    if (w_mode == kPermuteWIndex) v.w = static_cast<double>(o);
    a[t] = v;
}

// Single integrator passes, for host-driven steps with an arbitrary force
// callback between them (the reference's step(System&, ForceFn&&)).
__global__ void k_half_kick(long n, double4* __restrict__ vel, const double4* __restrict__ frc,
                            double half_dt) {
    const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double4 v = vel[t];
    const double4 f = frc[t];
    const double c = half_dt * (1.0 / v.w);
    v.x = fma(c, f.x, v.x);
    v.y = fma(c, f.y, v.y);
    v.z = fma(c, f.z, v.z);
    vel[t] = v;
}

__global__ void k_drift(long n, Box box, double4* __restrict__ pos, const double4* __restrict__ vel,
                        double h) {
    const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    pos[t] = drift(box, pos[t], vel[t], h);
}

__global__ void k_langevin_o(long n, double4* __restrict__ vel, const int* __restrict__ orig,
                             double c1, double c2, double temperature, uint64_t seed,
                             uint64_t counter) {
    const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double4 v = vel[t];
    double g0, g1, g2;
    normals3(seed, counter, static_cast<uint32_t>(orig[t]), g0, g1, g2);
    const double k = c2 * sqrt(temperature / v.w);
    v.x = c1 * v.x + g0 * k;
    v.y = c1 * v.y + g1 * k;
    v.z = c1 * v.z + g2 * k;
    vel[t] = v;
}

__global__ void k_scale(long n, double4* __restrict__ vel, double s) {
    const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double4 v = vel[t];
    v.x = v.x * s;
    v.y = v.y * s;
    v.z = v.z * s;
    vel[t] = v;
}

__global__ void k_set_flag(int* flag, int v) {
    if (threadIdx.x == 0) *flag = v;
}

__global__ void k_quantize(long n, Quant q, const double4* __restrict__ pos,
                           const int* __restrict__ orig, int4* __restrict__ qpos) {
    const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    qpos[t] = quantize(q, pos[t], orig[t]);
}

inline unsigned grid_of(long n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

}  // namespace

void launch_pre_step(const StepArgs& a, hipStream_t s) {
    const dim3 g(grid_of(a.n));
    if (a.kind == 1)
        hipLaunchKernelGGL(k_pre<1>, g, dim3(kBlock), 0, s, a);
    else if (a.kind == 2)
        hipLaunchKernelGGL(k_pre<2>, g, dim3(kBlock), 0, s, a);
    else
        hipLaunchKernelGGL(k_pre<0>, g, dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_pre_finish, dim3(1), dim3(kFinish), 0, s, a, static_cast<int>(g.x));
}

void launch_quantize(long n, const Quant& q, const double4* pos, const int* orig, int4* qpos,
                     hipStream_t s) {
    hipLaunchKernelGGL(k_quantize, dim3(grid_of(n)), dim3(kBlock), 0, s, n, q, pos, orig, qpos);
}

void launch_check(const StepArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_check, dim3(grid_of(a.n)), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_check_finish, dim3(1), dim3(kFinish), 0, s, a, static_cast<int>(grid_of(a.n)));
}

void launch_apply_scale(long n, double4* vel, Scalars* sc, hipStream_t s) {
    hipLaunchKernelGGL(k_apply_scale, dim3(grid_of(n)), dim3(kBlock), 0, s, n, vel, sc);
    hipLaunchKernelGGL(k_reset_scale, dim3(1), dim3(64), 0, s, sc);
}

void launch_kinetic(long n, const double4* vel, double* part, Scalars* sc, hipStream_t s) {
    hipLaunchKernelGGL(k_kinetic, dim3(grid_of(n)), dim3(kBlock), 0, s, n, vel, part, sc);
    hipLaunchKernelGGL(k_kinetic_finish, dim3(1), dim3(kBlock), 0, s, part, static_cast<int>(grid_of(n)), sc);
}

void launch_remove_com(long n, double4* vel, double* part, double* out4, hipStream_t s) {
    hipLaunchKernelGGL(k_momentum, dim3(grid_of(n)), dim3(kBlock), 0, s, n, vel, part, out4,
                       static_cast<unsigned int*>(nullptr));
    hipLaunchKernelGGL(k_momentum_finish, dim3(1), dim3(kBlock), 0, s, part, static_cast<int>(grid_of(n)), out4);
    hipLaunchKernelGGL(k_sub_com, dim3(grid_of(n)), dim3(kBlock), 0, s, n, vel, out4);
}

void launch_unpermute(long n, const int* orig, const double4* a, double* x, double* y, double* z,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_unpermute, dim3(grid_of(n)), dim3(kBlock), 0, s, n, orig, a, x, y, z);
}

void launch_permute_in(long n, const int* orig, const double* x, const double* y, const double* z,
                       double4* a, int w_mode, hipStream_t s) {
    hipLaunchKernelGGL(k_permute_in, dim3(grid_of(n)), dim3(kBlock), 0, s, n, orig, x, y, z, a, w_mode);
}

void launch_half_kick(long n, double4* vel, const double4* frc, double half_dt, hipStream_t s) {
    hipLaunchKernelGGL(k_half_kick, dim3(grid_of(n)), dim3(kBlock), 0, s, n, vel, frc, half_dt);
}

void launch_drift(long n, const Box& box, double4* pos, const double4* vel, double h, hipStream_t s) {
    hipLaunchKernelGGL(k_drift, dim3(grid_of(n)), dim3(kBlock), 0, s, n, box, pos, vel, h);
}

void launch_langevin_o(long n, double4* vel, const int* orig, double c1, double c2, double T,
                       uint64_t seed, uint64_t counter, hipStream_t s) {
    hipLaunchKernelGGL(k_langevin_o, dim3(grid_of(n)), dim3(kBlock), 0, s, n, vel, orig, c1, c2, T,
                       seed, counter);
}

void launch_scale(long n, double4* vel, double sfac, hipStream_t s) {
    hipLaunchKernelGGL(k_scale, dim3(grid_of(n)), dim3(kBlock), 0, s, n, vel, sfac);
}

void launch_set_flag(int* flag, int v, hipStream_t s) {
    hipLaunchKernelGGL(k_set_flag, dim3(1), dim3(64), 0, s, flag, v);
}

}  // namespace mdk
