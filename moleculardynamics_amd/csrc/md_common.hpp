// Shared device helpers and the device-side state layout of the MI355X
// moldyn path.
//
// Layout in HBM (all arrays in cell-sorted slot order t = 0..n-1):
//   pos  double4  x, y, z, (w unused)            32 B/atom, gathered by the pair kernel
//   vel  double4  vx, vy, vz, mass               32 B/atom
//   frc  double4  fx, fy, fz, (w unused)         32 B/atom
//   ref  double4  positions at the last list build (displacement check)
//   orig int32    upload index of slot t; slot int32 inverse map
//   nbr  uint32   full Verlet list in the octet layout (oct_index); an entry
//                 is its partner's byte offset in the 16-byte fixed-point
//                 position array, slot << 4 (entry_of / entry_slot)
//   ncnt int32    row length of t
// pos.w holds the atom's upload index (as a double): the orientation of a
// pair (the reference's i < j) is read from the two positions themselves.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace mdk {

constexpr int kBlock = 256;     // 4 wave64s per workgroup
constexpr int kWave = 64;
constexpr uint32_t kFlipBit = 0x80000000u;
constexpr uint32_t kIndexMask = 0x7fffffffu;
constexpr double kFixScale = 4294967296.0;          // 2^32: fixed-point resolution 2.3e-10
constexpr double kFixInv = 1.0 / 4294967296.0;
constexpr double kMagic = 6755399441055744.0;        // 1.5 * 2^52
constexpr unsigned kXcd = 8;                         // accelerator dies (each with its own L2)

// XCD-aware workgroup order. The dispatcher hands workgroup b to die b % 8,
// so consecutive workgroups (neighbouring atoms in cell order) would land on
// eight different L2s. Launched with a grid that is a multiple of 8, logical
// block (b % 8) * (grid / 8) + b / 8 gives every die one contiguous run of
// atoms, whose cell neighbourhoods then share that die's L2.
inline unsigned xcd_grid(long n) {
    const long g = (n + kBlock - 1) / kBlock;
    const long r = ((g > 0 ? g : 1) + kXcd - 1) / kXcd * kXcd;
    return static_cast<unsigned>(r);
}
__device__ __forceinline__ long xcd_block() {
    const unsigned per = gridDim.x / kXcd;
    return static_cast<long>(blockIdx.x % kXcd) * per + blockIdx.x / kXcd;
}

// Device scalars shared by the step kernels (one instance per context).
struct Scalars {
    double energy, virial, kinetic;  // last reductions (FP64)
    double top1, top2;               // displacement top-2 of the last check
    double vscale;                   // pending velocity scale (NHC), 1 otherwise
    long long pairs_in_cut;          // unordered pairs inside the cutoff, last force call
    long long builds;                // list builds performed on the device
    int rebuild_flag;                // 1 when the next rebuild kernels must run
    int overflow;                    // neighbour row overflow seen
    int fix_saturated;               // fixed-point conversion saturated
    int pad;
    unsigned long long step;         // step counter of graph-replayed runs (see StepArgs::dev_step)
};

// Nose-Hoover chain state (MTK96), held on the device so a whole run stays
// there. Chains longer than kMaxChain are refused at configuration.
constexpr int kMaxChain = 16;
struct Chain {
    int M;
    double dof, T;
    double Q[kMaxChain], xi[kMaxChain], p[kMaxChain];
    double iQ[kMaxChain];  // 1 / Q
    double sy[3];          // Suzuki-Yoshida weights w1, 1 - 2 w1, w1 (host std::cbrt)
};

// Calls f(std::integral_constant<int, j>) for j = N-1 .. 0 / 0 .. N-1, so
// that every index inside f is a compile-time constant.
template <int N, class F>
__device__ __forceinline__ void unroll_down(F& f) {
    if constexpr (N > 0) {
        f(std::integral_constant<int, N - 1>{});
        unroll_down<N - 1>(f);
    }
}
template <int N, int J = 0, class F>
__device__ __forceinline__ void unroll_up(F& f) {
    if constexpr (J < N) {
        f(std::integral_constant<int, J>{});
        unroll_up<N, J + 1>(f);
    }
}

struct Box {
    double L[3];
    // thresholds for the exact fast fold (see fold())
    double lo[3], hi1[3], hi2[3];
};

inline Box make_box(const double L[3]) {
    Box b;
    for (int d = 0; d < 3; ++d) {
        b.L[d] = L[d];
        b.lo[d] = 0.4999999 * L[d];
        b.hi1[d] = 0.5000001 * L[d];
        b.hi2[d] = 1.4999999 * L[d];
    }
    return b;
}

// v - L*floor(v/L + 0.5) with IEEE division and no contraction: the
// reference's periodic fold (core/box.cpp:12-14) bit for bit.
__device__ __forceinline__ double fold_ref(double v, double len) {
#pragma clang fp contract(off)
    const double k = floor(v / len + 0.5);
    const double s = len * k;
    return v - s;
}

// The same function, exact, without the division on the common paths:
// |v| <= 0.4999999 L gives k = 0 (returns v); 0.5000001 L <= |v| <= 1.4999999 L
// gives k = sign(v) (returns v - sign(v) L, one correctly rounded subtraction,
// as L*k is exact); the narrow bands between fall back to fold_ref.
__device__ __forceinline__ double fold(double v, double len, double lo, double hi1, double hi2) {
#pragma clang fp contract(off)
    const double a = fabs(v);
    if (a <= lo) return v;
    if (a >= hi1 && a <= hi2) return v - copysign(len, v);
    return fold_ref(v, len);
}

__device__ __forceinline__ void min_image(const Box& b, double& dx, double& dy, double& dz) {
    dx = fold(dx, b.L[0], b.lo[0], b.hi1[0], b.hi2[0]);
    dy = fold(dy, b.L[1], b.lo[1], b.hi1[1], b.hi2[1]);
    dz = fold(dz, b.L[2], b.lo[2], b.hi1[2], b.hi2[2]);
}

// (x*x + y*y) + z*z with separate roundings (core/vec3.hpp:59-65).
__device__ __forceinline__ double norm2_ref(double x, double y, double z) {
#pragma clang fp contract(off)
    double s = x * x;
    s = s + y * y;
    s = s + z * z;
    return s;
}

// (ax*bx + ay*by) + az*bz with separate roundings (core/vec3.hpp:59-61).
__device__ __forceinline__ double norm2_dot(double ax, double ay, double az, double bx, double by,
                                            double bz) {
#pragma clang fp contract(off)
    double s = ax * bx;
    s = s + ay * by;
    s = s + az * bz;
    return s;
}

// Oriented minimum-image difference of the unordered pair (a, b): the
// reference evaluates r_i - r_j with i < j by upload index
// (forcefield/lennard_jones.cpp:71, neighbor/verlet_list.cpp:19). `flip`
// says a is the larger index, so the difference is r_b - r_a.
__device__ __forceinline__ void pair_delta(const Box& b, const double4& pa, const double4& pb,
                                           bool flip, double& dx, double& dy, double& dz) {
#pragma clang fp contract(off)
    // b - a is exactly -(a - b) under round-to-nearest, so a sign select
    // replaces the branch with the same value
    dx = pa.x - pb.x;
    dy = pa.y - pb.y;
    dz = pa.z - pb.z;
    dx = flip ? -dx : dx;
    dy = flip ? -dy : dy;
    dz = flip ? -dz : dz;
    min_image(b, dx, dy, dz);
}

// Main-list entries: the partner's slot shifted left by 4, the byte offset of
// its 16-byte fixed-point position, so the mixed pair kernel adds it to a
// scalar base with no arithmetic. Slots stay below 2^28.
constexpr long kMaxListAtoms = 1L << 28;
__host__ __device__ __forceinline__ uint32_t entry_of(int slot) {
    return static_cast<uint32_t>(slot) << 4;
}
__host__ __device__ __forceinline__ int entry_slot(uint32_t e) { return static_cast<int>(e >> 4); }

// Row-oriented minimum image d = r_i - r_j with the bits of pair_delta's
// result (negated when flip): when every component lies inside the fast band
// (|v| <= 0.4999999 L) the fold is the identity, which is odd, so no
// orientation is needed; otherwise the pair is oriented by upload index,
// folded exactly and turned back to the row's orientation. Squares, and so
// r2 and every cutoff decision, are the same from both rows.
__device__ __forceinline__ void row_delta(const Box& b, const double4& pi, const double4& pj,
                                          bool flip, double& dx, double& dy, double& dz) {
#pragma clang fp contract(off)
    dx = pi.x - pj.x;
    dy = pi.y - pj.y;
    dz = pi.z - pj.z;
    if (!(fabs(dx) <= b.lo[0] && fabs(dy) <= b.lo[1] && fabs(dz) <= b.lo[2])) {
        dx = flip ? -dx : dx;
        dy = flip ? -dy : dy;
        dz = flip ? -dz : dz;
        min_image(b, dx, dy, dz);
        dx = flip ? -dx : dx;
        dy = flip ? -dy : dy;
        dz = flip ? -dz : dz;
    }
}

// row_delta with the orientation read from the upload indices the positions
// carry in w (the pair is oriented i < j by upload index).
__device__ __forceinline__ void row_delta_w(const Box& b, const double4& pi, const double4& pj,
                                            double& dx, double& dy, double& dz) {
#pragma clang fp contract(off)
    dx = pi.x - pj.x;
    dy = pi.y - pj.y;
    dz = pi.z - pj.z;
    if (!(fabs(dx) <= b.lo[0] && fabs(dy) <= b.lo[1] && fabs(dz) <= b.lo[2])) {
        const bool flip = pi.w > pj.w;
        dx = flip ? -dx : dx;
        dy = flip ? -dy : dy;
        dz = flip ? -dz : dz;
        min_image(b, dx, dy, dz);
        dx = flip ? -dx : dx;
        dy = flip ? -dy : dy;
        dz = flip ? -dz : dz;
    }
}

// Slot k of atom t in the octet layout of the main Verlet list: every group
// of 64 consecutive atoms owns a [cap/8][64][8] tile, so entries 8g..8g+7 of
// one atom are 32 contiguous bytes (one whole-sector write by the build, two
// 16-byte loads by the pair kernel) and group g of a wave is 2 KB
// contiguous. cap is a multiple of 8.
__host__ __device__ __forceinline__ long oct_index(long t, int k, int cap) {
    return (((t >> 6) * static_cast<long>(cap >> 3) + (k >> 3)) * 64 + (t & 63)) * 8 + (k & 7);
}

// Slot k of atom t in the wave-blocked neighbour layout: every group of 64
// consecutive atoms owns one contiguous [cap][64] tile, so a wave reading
// slot k of its 64 atoms touches one 256-byte line, and all its slots sit in
// one cap*256-byte span.
__host__ __device__ __forceinline__ long nbr_index(long t, int k, int cap) {
    return ((t >> 6) * static_cast<long>(cap) + k) * 64 + (t & 63);
}

// LJ pair terms of the mixed path (Truncated), written once so every kernel
// that uses them evaluates the same operations in the same order. The cutoff
// test is the FP64 one of the plain path (same pairs as the reference); the
// terms are FP32: r2 rounded to FP32, 1/r2 from the hardware reciprocal plus
// one Newton step, s2 = sig^2/r2, s6 = s2^3, u = 4 eps (s6*s6 - s6) and
// w = 24 eps (2 s6*s6 - s6) with the products fused into the subtraction
// (one rounding where the two terms cancel near the potential minimum).
// Returns false outside the cutoff.
// The terms for a pair already known to be inside the cutoff.
__device__ __forceinline__ void lj_terms_f32_in(double r2d, float sig2, float eps, float& inv,
                                                float& u, float& w) {
    const float r2 = static_cast<float>(r2d);
    float y = __builtin_amdgcn_rcpf(r2);
    const float e = __builtin_fmaf(-r2, y, 1.0f);
    y = __builtin_fmaf(y, e, y);
    inv = y;
    const float s2 = sig2 * inv;
    const float s6 = (s2 * s2) * s2;
    u = (4.0f * eps) * __builtin_fmaf(s6, s6, -s6);
    w = (24.0f * eps) * __builtin_fmaf(2.0f * s6, s6, -s6);
}

__device__ __forceinline__ bool lj_terms_f32(double r2d, double rc2d, float sig2, float eps,
                                             float& inv, float& u, float& w) {
    if (!(r2d < rc2d)) return false;
    lj_terms_f32_in(r2d, sig2, eps, inv, u, w);
    return true;
}

// ---- counter-based normal deviates (BAOAB noise) ------------------------------
//
// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC11): key (k0, k1) = the 64-bit seed, counter = (atom
// upload index, step low, step high, stream). Two 32-bit words give one
// 53-bit uniform in (0, 1), and the Box-Muller transform turns two streams'
// four uniforms into the three normals of an atom's O step.
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

__device__ __forceinline__ void normals3(uint32_t k0, uint32_t k1, uint64_t step, uint32_t id,
                                         double& g0, double& g1, double& g2) {
    const U4 a = philox10(U4{id, static_cast<uint32_t>(step), static_cast<uint32_t>(step >> 32), 0u}, k0, k1);
    const U4 b = philox10(U4{id, static_cast<uint32_t>(step), static_cast<uint32_t>(step >> 32), 1u}, k0, k1);
    const double u1 = u53(a.x, a.y), u2 = u53(a.z, a.w);
    const double u3 = u53(b.x, b.y), u4 = u53(b.z, b.w);
    const double r1 = sqrt(-2.0 * log(u1));
    const double r2 = sqrt(-2.0 * log(u3));
    double s1, c1, s2, c2;
    sincospi(2.0 * u2, &s1, &c1);
    sincospi(2.0 * u4, &s2, &c2);
    g0 = r1 * c1;
    g1 = r1 * s1;
    g2 = r2 * c2;
}

__device__ __forceinline__ void normals3(uint64_t seed, uint64_t step, uint32_t id, double& g0,
                                         double& g1, double& g2) {
    normals3(static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32), step, id, g0, g1, g2);
}

// Position update r = wrap(fma(h, v, r)) with the exact fold.
__device__ __forceinline__ double4 drift(const Box& b, double4 p, const double4& v, double h) {
    p.x = fold(fma(h, v.x, p.x), b.L[0], b.lo[0], b.hi1[0], b.hi2[0]);
    p.y = fold(fma(h, v.y, p.y), b.L[1], b.lo[1], b.hi1[1], b.hi2[1]);
    p.z = fold(fma(h, v.z, p.z), b.L[2], b.lo[2], b.hi1[2], b.hi2[2]);
    return p;
}

// ---- wave64 / workgroup reductions (fixed tree: deterministic) -------------

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}
__device__ __forceinline__ long long wave_sum_ll(long long v) {
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Sum over the workgroup of kBlock threads; result valid in thread 0.
template <class T, class WaveFn>
__device__ __forceinline__ T block_reduce(T v, T* smem, WaveFn wf) {
    v = wf(v);
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    if (lane == 0) smem[wid] = v;
    __syncthreads();
    T r = T(0);
    if (threadIdx.x == 0) {
        for (int w = 0; w < kBlock / kWave; ++w) r += smem[w];
    }
    __syncthreads();
    return r;
}

// Top-2 (largest, second largest, with multiplicity) merge.
__device__ __forceinline__ void top2_merge(double& a1, double& a2, double b1, double b2) {
    if (b1 > a1) {
        a2 = fmax(a1, b2);
        a1 = b1;
    } else {
        a2 = fmax(a2, b1);
    }
}

__device__ __forceinline__ void wave_top2(double& t1, double& t2) {
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const double o1 = __shfl_xor(t1, off, kWave);
        const double o2 = __shfl_xor(t2, off, kWave);
        top2_merge(t1, t2, o1, o2);
    }
}

// 64-bit fixed point (scale 2^32) by biased accumulation: every term adds
// the bit pattern of fma(v, 2^32, 1.5 * 2^52), whose value is
// round(v * 2^32) + 1.5 * 2^52 (ties to even) while |v| < 2^19 (v * 2^32 is
// exact in FP64, so the FMA rounds once). The bias is removed once per
// accumulator of n terms at the end, in 64-bit modular arithmetic: one
// conversion, one FMA and one 64-bit add per term, no branches. Callers track
// the largest |term| and check it against kFixLimit.
constexpr float kFixLimit = 524288.0f;  // 2^19

__device__ __forceinline__ unsigned long long fix_biased(float v) {
    return static_cast<unsigned long long>(
        __double_as_longlong(fma(static_cast<double>(v), kFixScale, kMagic)));
}

// The same for a term already multiplied by 2^32 in FP32 (exact: a power of
// two), whose limit is then kFixLimitScaled: one conversion and one FP64 add.
constexpr float kFixLimitScaled = 2251799813685248.0f;  // 2^51

__device__ __forceinline__ unsigned long long fix_biased_scaled(float vs) {
    return static_cast<unsigned long long>(
        __double_as_longlong(static_cast<double>(vs) + kMagic));
}

__device__ __forceinline__ long long fix_value(unsigned long long acc, unsigned long long n) {
    const unsigned long long bias = static_cast<unsigned long long>(__double_as_longlong(kMagic));
    return static_cast<long long>(acc - n * bias);
}

}  // namespace mdk
