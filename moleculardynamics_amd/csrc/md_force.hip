// Lennard-Jones pair kernels.
//
// k_lj_lanes: a few lanes per atom over its row of the full Verlet list (each
// unordered pair is evaluated from both atoms, so every group of lanes owns
// its atom's force outright: no atomics, a fixed per-lane summation order and
// a fixed combination of the lanes, and a deterministic result). Pair
// arithmetic follows
// forcefield/lennard_jones.cpp:50-88:
//   d = minimumImage(r_i - r_j) (i<j by upload index), r2 = |d|^2,
//   skip unless r2 < rc2, s2 = sig^2/r2, s6 = s2^3,
//   u = 4 eps (s6^2 - s6), w = 24 eps (2 s6^2 - s6),
//   LinearForceShift: u -= u(rc) + (r - rc) u'(rc), w += r u'(rc),
//   f_i = +(w/r2) d, f_j = -(w/r2) d, E = sum u, W = sum w.
// FP64: the minimum image and r2 are evaluated exactly as the reference does
// (same cutoff decisions); s2 and w/r2 use one reciprocal (a last-bit
// difference from the reference's two divisions, far inside 1e-12).
// MIXED: the partner's periodic fixed-point position (Quant) gives d as exact
// integer differences and r2 in FP64 from them; the LJ powers and c = w/r2
// are FP32; each lane sums c d, s^6 and s^12 in FP64 (AccMix), and each
// lane's share of a row's force, energy and virial becomes 64-bit fixed point
// (2^-32), so every sum of lanes or rows is an exact integer addition. A row
// with a cutoff decision too close to rc for the fixed-point r2 is evaluated
// again from the FP64 positions with the reference's test (mix_exact), so the
// pairs inside the cutoff are always the FP64 kernel's. Without fixed-point positions (boxes
// longer than 128 rc) the mixed kernel reads the FP64 positions.
//
// E and W come out of a fixed reduction tree: wave shuffles, a 4-wave LDS
// step, per-workgroup partials, and a finish kernel summing the partials in a
// fixed order. The finish kernel also records the step's thermo
// row and advances the Nose-Hoover chain when one is attached.
#include "md_common.hpp"
#include "md_kernels.hpp"

#include "../../include/moldyn_hip.h"

namespace mdk {

namespace {

// Working copy of the chain for the one thread that advances it, in LDS:
// the links are updated by runtime loops over M, so the code is one link
// body, not an unrolled copy per link (a long straight-line sequence runs
// out of the instruction cache). Divisions by Q use the reciprocals iQ
// computed once on the host.
struct ChainWork {
    int M;
    double dof, T;
    double sy[3];
    double p[kMaxChain], xi[kMaxChain], iQ[kMaxChain];

    __device__ void load(const Chain& c) {
        M = c.M;
        dof = c.dof;
        T = c.T;
        sy[0] = c.sy[0];
        sy[1] = c.sy[1];
        sy[2] = c.sy[2];
        for (int k = 0; k < M; ++k) {
            p[k] = c.p[k];
            xi[k] = c.xi[k];
            iQ[k] = c.iQ[k];
        }
    }
    __device__ void store(Chain& c) const {
        for (int k = 0; k < M; ++k) {
            c.p[k] = p[k];
            c.xi[k] = xi[k];
        }
    }
};

// e^x for the chain's damping and rescale factors. Their arguments are
// -dt/4 .. -dt/2 times a thermostat rate, far below 1/8 in magnitude for any
// stable step, where the Taylor series to x^13 (Horner form, 13 dependent
// FMAs) is exact to the last bit (the first omitted term, x^14/14!, is below
// 1e-23); larger arguments take the library exp. The chain runs on one
// thread, so its latency is the step's: this saves most of an exp's.
__device__ __forceinline__ double chain_exp(double x) {
    if (!(fabs(x) <= 0.125)) return exp(x);
    double p = 1.0 / 6227020800.0;  // 1/13!
    p = fma(p, x, 1.0 / 479001600.0);
    p = fma(p, x, 1.0 / 39916800.0);
    p = fma(p, x, 1.0 / 3628800.0);
    p = fma(p, x, 1.0 / 362880.0);
    p = fma(p, x, 1.0 / 40320.0);
    p = fma(p, x, 1.0 / 5040.0);
    p = fma(p, x, 1.0 / 720.0);
    p = fma(p, x, 1.0 / 120.0);
    p = fma(p, x, 1.0 / 24.0);
    p = fma(p, x, 1.0 / 6.0);
    p = fma(p, x, 0.5);
    p = fma(p, x, 1.0);
    return fma(p, x, 1.0);
}

// Link j of one Suzuki-Yoshida sub-step of length d: the thermostat force g
// on p_j (from the kinetic energy for j = 0, else from p_{j-1}) with the
// damping by p_{j+1} on both sides (p_{j+1} is not touched in between, so
// one exp serves both factors).
__device__ __forceinline__ void chain_link(ChainWork& c, int j, double ke, double d2, double d4) {
    const int M = c.M;
    double damp = 1.0;
    if (j + 1 < M) damp = chain_exp(-d4 * c.p[j + 1] * c.iQ[j + 1]);
    const double g = j == 0 ? 2.0 * ke - c.dof * c.T : c.p[j - 1] * c.p[j - 1] * c.iQ[j - 1] - c.T;
    c.p[j] = (c.p[j] * damp + d2 * g) * damp;
}

// One MTK96 chain propagation of length delta with Suzuki-Yoshida weights
// (integrate/nose_hoover.cpp:63-124). The chain rescales every velocity by a
// common factor, so the kinetic energy after a rescale is KE * f^2; it is
// tracked that way instead of re-summing the velocities. Returns the product
// of the three rescale factors and updates ke.
__device__ __noinline__ double chain_half(ChainWork& c, double& ke, double delta) {
    const int M = c.M;
    double scale = 1.0;
#pragma unroll 1
    for (int s = 0; s < 3; ++s) {
        const double d = c.sy[s] * delta;
        const double d2 = 0.5 * d, d4 = 0.25 * d;
#pragma unroll 1
        for (int j = M - 1; j >= 0; --j) chain_link(c, j, ke, d2, d4);
#pragma unroll 1
        for (int k = 0; k < M; ++k) c.xi[k] += d * c.p[k] * c.iQ[k];
        const double f = chain_exp(-d * c.p[0] * c.iQ[0]);
        scale *= f;
        ke *= f * f;
#pragma unroll 1
        for (int j = 0; j < M; ++j) chain_link(c, j, ke, d2, d4);
    }
    return scale;
}

__device__ double chain_conserved(const ChainWork& c) {
    double acc = c.dof * c.T * c.xi[0];
    for (int k = 1; k < c.M; ++k) acc += c.T * c.xi[k];
    for (int k = 0; k < c.M; ++k) acc += 0.5 * c.p[k] * c.p[k] * c.iQ[k];
    return acc;
}

// The chain's part of a step (one thread): one or two half-chains, the
// conserved contribution, the rescale factor.
struct ChainStep {
    double scale, conserved;
};

// Short chains (M <= 4, the usual 3 links) in registers: the same
// operations in the same order as chain_half over ChainWork, with every
// index a compile-time constant, so each link update is a register
// dependence instead of LDS round trips. The chain runs on one thread of the
// last-arriving workgroup, so its latency is on every step's critical path.

template <int M>
__device__ __forceinline__ ChainStep chain_step_reg(Chain* chain, double& ke, double delta, bool twice) {
    double p[M], xi[M], iQ[M];
    const double dof = chain->dof, T = chain->T;
    const double sy0 = chain->sy[0], sy1 = chain->sy[1], sy2 = chain->sy[2];
    auto load = [&](auto J) {
        constexpr int j = decltype(J)::value;
        p[j] = chain->p[j];
        xi[j] = chain->xi[j];
        iQ[j] = chain->iQ[j];
    };
    unroll_up<M>(load);
    // one MTK96 chain propagation of length delta (chain_half)
    auto half = [&](double& k) {
        double scale = 1.0;
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            const double d = (s == 0 ? sy0 : (s == 1 ? sy1 : sy2)) * delta;
            const double d2 = 0.5 * d, d4 = 0.25 * d;
            auto link = [&](auto J) {
                constexpr int j = decltype(J)::value;
                double damp = 1.0;
                if constexpr (j + 1 < M) damp = chain_exp(-d4 * p[j + 1] * iQ[j + 1]);
                double g;
                if constexpr (j == 0)
                    g = 2.0 * k - dof * T;
                else
                    g = p[j - 1] * p[j - 1] * iQ[j - 1] - T;
                p[j] = (p[j] * damp + d2 * g) * damp;
            };
            auto drift = [&](auto J) {
                constexpr int j = decltype(J)::value;
                xi[j] += d * p[j] * iQ[j];
            };
            unroll_down<M>(link);
            unroll_up<M>(drift);
            const double f = chain_exp(-d * p[0] * iQ[0]);
            scale *= f;
            k *= f * f;
            unroll_up<M>(link);
        }
        return scale;
    };
    ChainStep r;
    r.scale = half(ke);
    double acc = dof * T * xi[0];
    auto cons_xi = [&](auto J) {
        constexpr int j = decltype(J)::value;
        if constexpr (j > 0) acc += T * xi[j];
    };
    auto cons_p = [&](auto J) {
        constexpr int j = decltype(J)::value;
        acc += 0.5 * p[j] * p[j] * iQ[j];
    };
    unroll_up<M>(cons_xi);
    unroll_up<M>(cons_p);
    r.conserved = acc;
    if (twice) {
        double ke2 = ke;
        r.scale *= half(ke2);
    }
    auto store = [&](auto J) {
        constexpr int j = decltype(J)::value;
        chain->p[j] = p[j];
        chain->xi[j] = xi[j];
    };
    unroll_up<M>(store);
    return r;
}

__device__ ChainStep chain_step(Chain* chain, ChainWork& c, double& ke, double delta, bool twice) {
    switch (chain->M) {
        case 1: return chain_step_reg<1>(chain, ke, delta, twice);
        case 2: return chain_step_reg<2>(chain, ke, delta, twice);
        case 3: return chain_step_reg<3>(chain, ke, delta, twice);
        case 4: return chain_step_reg<4>(chain, ke, delta, twice);
        default: break;
    }
    c.load(*chain);
    ChainStep r;
    r.scale = chain_half(c, ke, delta);
    r.conserved = chain_conserved(c);
    if (twice) {
        double ke2 = ke;
        r.scale *= chain_half(c, ke2, delta);
    }
    c.store(*chain);
    return r;
}

// Final reduction, launched after the pair kernel: kFinishGroups workgroups
// each sum a fixed contiguous share of the pair kernel's nb per-workgroup
// partials (a strided share per thread, kUnroll loads in flight, wave
// shuffles, the waves in index order) into one partial of their own; the
// workgroup that finishes last (an atomic ticket, after a release fence)
// sums those kFinishGroups partials with one wave's shuffle tree and
// finishes the step on one thread: E, W, the pair count, the Nose-Hoover
// chain and the thermo row. Every sum has a fixed order, so the result does
// not depend on which workgroup arrives last; the partials are read from
// every die's L2 by many CUs at once instead of one.
constexpr int kFinish = 256;
constexpr int kUnroll = 8;
static_assert(kFinishGroups == kWave, "the last workgroup reduces one partial per lane");

__device__ void finish_step(const ForceArgs& a, bool mixed, int nb) {
    __shared__ double red_d[3][kFinish / kWave];
    __shared__ long long red_l[3][kFinish / kWave];
    __shared__ int last;
    __shared__ ChainWork cw;
    const int g = blockIdx.x;
    const int chunk = (nb + kFinishGroups - 1) / kFinishGroups;
    const int b1 = min(nb, (g + 1) * chunk);
    double sd[3] = {0.0, 0.0, 0.0};
    long long sl[3] = {0, 0, 0};
    for (int b0 = g * chunk + threadIdx.x; b0 < b1; b0 += kFinish * kUnroll) {
        double vd[kUnroll][3];
        long long vl[kUnroll][3];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const int b = b0 + u * kFinish;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                vd[u][q] = b < b1 ? a.part_d[3 * b + q] : 0.0;
                vl[u][q] = b < b1 ? a.part_l[3 * b + q] : 0;
            }
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                sd[q] += vd[u][q];
                sl[q] += vl[u][q];
            }
    }
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const double vd = wave_sum(sd[q]);
        const long long vl = wave_sum_ll(sl[q]);
        if (lane == 0) {
            red_d[q][wid] = vd;
            red_l[q][wid] = vl;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < 3; ++q) {
            double td = 0.0;
            long long tl = 0;
            for (int w = 0; w < kFinish / kWave; ++w) {
                td += red_d[q][w];
                tl += red_l[q][w];
            }
            a.fin_d[3 * g + q] = td;
            a.fin_l[3 * g + q] = tl;
        }
        __threadfence();
        last = atomicAdd(a.fin_count, 1u) == kFinishGroups - 1 ? 1 : 0;
    }
    __syncthreads();
    if (!last || threadIdx.x >= kWave) return;
    __threadfence();
    const volatile double* fd = a.fin_d;
    const volatile long long* fl = a.fin_l;
    double tot_d[3];
    long long tot_l[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        tot_d[q] = wave_sum(fd[3 * lane + q]);
        tot_l[q] = wave_sum_ll(fl[3 * lane + q]);
    }
    if (lane != 0) return;
    *a.fin_count = 0u;  // the next launch's tickets
    double E, W;
    if (mixed) {
        E = 0.5 * (static_cast<double>(tot_l[0]) * kFixInv);
        W = 0.5 * (static_cast<double>(tot_l[1]) * kFixInv);
    } else {
        E = 0.5 * tot_d[0];
        W = 0.5 * tot_d[1];
    }
    double ke = tot_d[2];
    Scalars* sc = a.sc;
    sc->energy = E;
    sc->virial = W;
    const long long pairs = tot_l[2] / 2;  // each pair is a term of both its rows
    sc->pairs_in_cut = pairs;
    double conserved = 0.0;
    if (a.kick) {
        if (a.chain != nullptr) {
            const ChainStep r = chain_step(a.chain, cw, ke, a.chain_delta, a.chain_twice != 0);
            conserved = r.conserved;
            sc->vscale = r.scale;
        }
        sc->kinetic = ke;
    }
    md_thermo* th = static_cast<md_thermo*>(a.thermo);
    if (a.dev_step && a.thermo_base != nullptr)
        th = static_cast<md_thermo*>(a.thermo_base) + (sc->step - a.step0);
    if (th != nullptr) {
        th->potential = E;
        th->virial = W;
        th->kinetic = ke;
        th->conserved = conserved;
        th->pairs = static_cast<double>(pairs);
    }
    if (a.dev_step) sc->step += 1;
}

// Every in-cutoff term is counted in the row that holds it (n). The list is
// full and both rows of a pair take the same decision (the same r2 bits), so
// the pairs inside the cutoff are the sum of n over the rows, halved.
struct Acc64 {
    double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0, w = 0.0;
    int n = 0;
};

template <bool LFS>
__device__ __forceinline__ void pair64(const ForceArgs& a, const double4& pi, const double4& pj,
                                       Acc64& q, bool live = true) {
    double dx, dy, dz;
    row_delta_w(a.box, pi, pj, dx, dy, dz);
    const double r2 = norm2_ref(dx, dy, dz);
    if (live && r2 < a.lj.rc2) {
        // hardware reciprocal refined by two Newton steps (within an ulp of 1/r2)
        double inv = __builtin_amdgcn_rcp(r2);
        double e = fma(-r2, inv, 1.0);
        inv = fma(inv, e, inv);
        e = fma(-r2, inv, 1.0);
        inv = fma(inv, e, inv);
        const double s2 = a.lj.sig2 * inv;
        const double s6 = s2 * s2 * s2;
        double u = 4.0 * a.lj.eps * (s6 * s6 - s6);
        double ww = 24.0 * a.lj.eps * (2.0 * s6 * s6 - s6);
        if (LFS) {
            const double r = sqrt(r2);
            u = u - a.lj.u_rc - (r - a.lj.rc) * a.lj.dudr_rc;
            ww = ww + r * a.lj.dudr_rc;
        }
        q.e += u;
        q.w += ww;
        q.n += 1;
        const double c = ww * inv;
        q.fx += c * dx;
        q.fy += c * dy;
        q.fz += c * dz;
    }
}

// ---- mixed precision ---------------------------------------------------------
//
// Per entry: the coordinate differences are exact integers (fixed-point
// positions, one wrapping 32-bit subtraction per axis), converted to FP64
// once; r2 is formed in FP64 from them, so its only error is the fixed-point
// quantisation (one unit per axis) and it decides the cutoff outside a narrow
// band. The LJ powers are FP32: s2 = sigma^2 / r2 (hardware reciprocal and one
// Newton step), s6 = s2^3, t = 2 s6^2 - s6 (one rounding); c = t / r2. The
// force term c d is one FP64 FMA per axis into the row's sum, and the row
// also sums s6 and t in FP64 (E = 2 eps (T - S6), W = 24 eps T per row).
// FP32 values have 24-bit mantissas, so these FP64 sums are exact to about
// 2^-53 relative, and their order (the lane's walk order) is fixed. Each
// lane's share of a row (force, energy, virial) then becomes 64-bit fixed
// point (2^-32 units), and every combination of lanes or rows -- the lanes
// of a row, the workgroup and global E and W -- is an integer sum,
// independent of order. Two cuts measured slower at 1M atoms and not kept:
// summing s6 and t in FP32 over each batch of four entries, one conversion
// per batch (fewer loop instructions, 3 % slower); and an FP64 butterfly over
// the lanes' sums with only the row's E and W converted (74 fewer
// instructions in the kernel, 3 % slower).
struct AccMix {
    double fx = 0.0, fy = 0.0, fz = 0.0;  // sum of c d (scaled lengths)
    double s6 = 0.0, t = 0.0;             // sums of s^6 and t = 2 s^12 - s^6
    double r = 0.0;                       // LinearForceShift: sum of r (scaled)
    int k = 0;                            // pairs inside the cutoff
};

// The FP32 terms of one entry. A pair outside the cutoff (in false) enters
// with s6 = 0 and adds nothing, so a wave's lanes need no divergent update.
template <bool LFS>
__device__ __forceinline__ void mix_term(const MixK& m, double dx, double dy, double dz, double r2,
                                         bool in, AccMix& q) {
    const float r2f = static_cast<float>(r2);
    float y = __builtin_amdgcn_rcpf(r2f);
    y = __builtin_fmaf(y, __builtin_fmaf(-r2f, y, 1.0f), y);
    const float s2 = m.sig2s * y;
    float s6 = (s2 * s2) * s2;
    s6 = in ? s6 : 0.0f;
    // t = 2 s12 - s6 with one rounding: it cancels near the potential minimum
    const float tt = __builtin_fmaf(2.0f * s6, s6, -s6);
    float t = tt;
    if (LFS) {
        const float rs = in ? __builtin_sqrtf(r2f) : 0.0f;
        t = __builtin_fmaf(m.krs, rs, t);
        q.r += static_cast<double>(rs);
    }
    const double c = static_cast<double>(t * y);
    q.fx = fma(c, dx, q.fx);
    q.fy = fma(c, dy, q.fy);
    q.fz = fma(c, dz, q.fz);
    q.s6 += static_cast<double>(s6);
    q.t += static_cast<double>(tt);
    q.k += in ? 1 : 0;
}

// Entry from fixed-point positions (mixed kernels; an entry is the partner's
// byte offset from the array base, so a gather is a scalar base plus one
// vector register, and only x, y, z are read). CUBIC: one unit on every
// axis, so the integer differences are the scaled lengths (s = u); otherwise
// they are scaled to real lengths per axis (s = 1). The two orientations of
// a pair give integer differences of opposite sign, hence the same r2 bits
// and the same decision in both rows.
template <bool CUBIC>
__device__ __forceinline__ void fixed_delta(const ForceArgs& a, const int4& qi, const int4& qj,
                                            double& dx, double& dy, double& dz) {
    dx = static_cast<double>(static_cast<int>(static_cast<uint32_t>(qi.x) - static_cast<uint32_t>(qj.x)));
    dy = static_cast<double>(static_cast<int>(static_cast<uint32_t>(qi.y) - static_cast<uint32_t>(qj.y)));
    dz = static_cast<double>(static_cast<int>(static_cast<uint32_t>(qi.z) - static_cast<uint32_t>(qj.z)));
    if (!CUBIC) {
        dx *= a.quant.unit[0];
        dy *= a.quant.unit[1];
        dz *= a.quant.unit[2];
    }
}

// Band hits are found by counting: q.k counts the entries with r2 < lo (in
// the cutoff) and mix_fixed_kb counts those with r2 < hi in kb, so a row saw
// a decision inside the band, and goes to the exact evaluation, exactly when
// the two differ. Both counts are per-lane carry adds, which keeps the walk
// free of per-entry boolean chains. MASKED: live false enters as a dead
// entry with r2 = 4 hi, beyond both bounds, so it adds to neither count and
// its terms are exact zeros.
template <bool LFS, bool CUBIC, bool MASKED>
__device__ __forceinline__ void mix_fixed_kb(const ForceArgs& a, const int4& qi, const int4& qj,
                                             bool live, AccMix& q, int& kb) {
    double dx, dy, dz;
    fixed_delta<CUBIC>(a, qi, qj, dx, dy, dz);
    const double r2f = fma(dz, dz, fma(dy, dy, dx * dx));
    const double r2 = MASKED ? (live ? r2f : 4.0 * a.mq.hi) : r2f;
    kb += r2 < a.mq.hi ? 1 : 0;
    mix_term<LFS>(a.mq, dx, dy, dz, r2, r2 < a.mq.lo, q);
}

// Entry from the FP64 positions with the reference's exact cutoff test (the
// FP64 kernel's decision), in the units of m.
template <bool LFS>
__device__ __forceinline__ void mix_exact(const ForceArgs& a, const MixK& m, const double4& pi,
                                          const double4& pj, AccMix& q) {
    double dx, dy, dz;
    row_delta_w(a.box, pi, pj, dx, dy, dz);
    const double r2 = norm2_ref(dx, dy, dz);
    const bool in = r2 < a.lj.rc2;
    const double is = m.inv_s;
    mix_term<LFS>(m, dx * is, dy * is, dz * is, r2 * (is * is), in, q);
}

// 64-bit fixed point (2^-32 units) of v, rounded to nearest. |v| < 2^30
// holds for every physical force, energy and virial of a row; larger values
// raise `sat` (the conversion would overflow near 2^31).
__device__ __forceinline__ long long to_fix(double v, bool& sat) {
    sat = sat || !(fabs(v) < 1073741824.0);
    return __double2ll_rn(v * kFixScale);
}

struct FixRow {
    long long fx, fy, fz, e, w;
};

// A row's (or a lane's share of a row's) sums as fixed point.
template <bool LFS>
__device__ __forceinline__ FixRow mix_out(const MixK& m, const AccMix& q, bool& sat) {
    // E = 4 eps (S12 - S6) = 2 eps (T - S6), W = 24 eps T
    double e = 0.5 * m.e4 * (q.t - q.s6);
    double w = m.e24 * q.t;
    if (LFS) {
        e -= static_cast<double>(q.k) * m.ushift + m.urs * q.r;
        w += m.urs * q.r;
    }
    FixRow r;
    r.fx = to_fix(m.fscale * q.fx, sat);
    r.fy = to_fix(m.fscale * q.fy, sat);
    r.fz = to_fix(m.fscale * q.fz, sat);
    r.e = to_fix(e, sat);
    r.w = to_fix(w, sat);
    return r;
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// The workgroup's partials of a pair kernel in one LDS exchange: each wave's
// shuffle sums, then the waves in index order on thread 0 (the order of
// block_reduce, so the same bits), for the values the precision uses (FP64:
// E, W, KE and the term count; mixed: the fixed-point E, W, the term count
// and KE); the unused slots are written as 0.
// STRIDE: only lanes that are multiples of STRIDE hold values (the first lane
// of each group of a multi-lane kernel), so the shuffle steps below STRIDE
// are skipped.
template <int STRIDE, class T>
__device__ __forceinline__ T wave_sum_strided(T v) {
    for (int off = kWave / 2; off >= STRIDE; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

template <bool MIXED, int WAVES = kBlock / kWave, int STRIDE = 1>
__device__ __forceinline__ void write_partials(const ForceArgs& a, long b, double e, double w,
                                               double ke, long long ef, long long wf,
                                               long long nt) {
    __shared__ double xd[3][WAVES];
    __shared__ long long xl[3][WAVES];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const double v0 = MIXED ? 0.0 : wave_sum_strided<STRIDE>(e);
    const double v1 = MIXED ? 0.0 : wave_sum_strided<STRIDE>(w);
    const double v2 = wave_sum_strided<STRIDE>(ke);
    const long long l0 = MIXED ? wave_sum_strided<STRIDE>(ef) : 0;
    const long long l1 = MIXED ? wave_sum_strided<STRIDE>(wf) : 0;
    const long long l2 = wave_sum_strided<STRIDE>(nt);
    if (lane == 0) {
        xd[0][wid] = v0; xd[1][wid] = v1; xd[2][wid] = v2;
        xl[0][wid] = l0; xl[1][wid] = l1; xl[2][wid] = l2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < 3; ++q) {
            double td = 0.0;
            long long tl = 0;
            for (int k = 0; k < static_cast<int>(blockDim.x / kWave); ++k) {
                td += xd[q][k];
                tl += xl[q][k];
            }
            a.part_d[3 * b + q] = td;
            a.part_l[3 * b + q] = tl;
        }
    }
}

// ---- the Verlet-list pair kernel: several lanes per atom ----
//
// LANES consecutive lanes share an atom: lane s takes entries s, s + LANES,
// ... of its row, a few at a time (walk_lane64, walk_lane_fixed), and the
// group's sums are combined by shuffles in a fixed butterfly (FP64:
// deterministic; mixed: each lane's share becomes fixed point first and the
// group adds integers). A group redoes its row exactly when any of its lanes
// saw a band hit, which is the row's own condition.
//
// Up to kLanesMaxAtoms atoms the lane count fills the chip (16 while
// 16 N <= 262,144, halved until it fits). Above it the split is the
// measured best at 1M atoms (per launch; one lane per atom walking whole
// octets of 8 entries with the next octet's indices in flight took 160 us
// mixed, and one entry at a time per lane 264 us FP64):
//   FP64 (32-byte partners):  8 lanes x 2 entries 212 us, 4 x 4 214 us,
//                             8 x 4 229 us, 16 x 2 296 us;
//   mixed (12-byte partners): 4 lanes x 4 entries 153 us, 4 x 8 152 us,
//                             4 x 6 155 us, 4 x 2 171 us, 8 x 4 180 us,
//                             2 x 8 187 us, 2 x 4 197 us, 16 x 2 304 us.
constexpr long kLanesMaxAtoms = 32768;
constexpr int kFp64Lanes = 8;
constexpr int kMixLanes = 4;
// dispatch_lanes instantiates 16, 8 and 4 lanes per atom
static_assert(kFp64Lanes == 4 || kFp64Lanes == 8 || kFp64Lanes == 16, "lanes per atom: 4, 8 or 16");
static_assert(kMixLanes == 4 || kMixLanes == 8 || kMixLanes == 16, "lanes per atom: 4, 8 or 16");

inline int lj_lanes(long n, bool mixed) {
    if (n > kLanesMaxAtoms) return mixed ? kMixLanes : kFp64Lanes;
    int l = 16;
    while (l > 1 && static_cast<long>(l) * n > 262144) l >>= 1;
    return l;
}

// Workgroups (and partial-sum slots) of the pair kernel for n atoms.
inline long lj_groups(long n, bool mixed) {
    const int l = lj_lanes(n, mixed);
    return (static_cast<long>(l) * n + kBlock - 1) / kBlock;
}

// A mixed row shared by LANES lanes, BATCH entries at a time: the row is cut
// into chunks of LANES * BATCH entries (whole octets) and lane s takes entries
// s * BATCH .. s * BATCH + BATCH - 1 of every chunk, in that order. They are
// consecutive in one octet, so one vector load brings a batch's indices.
// Each walk issues a batch's partner gathers together, then the next
// batch's indices, then the batch's terms, so a lane keeps several memory
// accesses in flight; full batches are unmasked, and the lane's last,
// partial batch reads the row atom's own entry (self) for its slots past the
// row end and adds nothing for them. Every kernel that walks a row this way
// (the slot and brick layouts, the exact redo) takes the same entries per
// lane in the same order, so their sums are the same bits.
constexpr int kFp64Batch = 2;
constexpr int kMixBatch = 4;

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// The BATCH indices of the lane's batch that starts at entry k0 (a multiple
// of BATCH). Past the row's last octet the octet is clamped, so the load
// stays inside the row; those indices are never used. The list is read
// exactly once per step, so the loads are non-temporal: the index stream
// does not evict the partner positions that the gathers reuse from L2.
template <int BATCH>
__device__ __forceinline__ void lane_indices(const uint32_t* row, int cap, int k0, uint32_t (&e)[BATCH]) {
    static_assert(BATCH == 1 || BATCH == 2 || BATCH == 4, "a batch lies inside one octet");
    const int oct = min(k0 >> 3, (cap >> 3) - 1);
    const uint32_t* p = row + static_cast<long>(oct) * 512 + (k0 & 7);
    if constexpr (BATCH == 4) {
        const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
        e[0] = v.x;
        e[1] = v.y;
        e[2] = v.z;
        e[3] = v.w;
    } else if constexpr (BATCH == 2) {
        const u32x2 v = __builtin_nontemporal_load(reinterpret_cast<const u32x2*>(p));
        e[0] = v.x;
        e[1] = v.y;
    } else {
        e[0] = __builtin_nontemporal_load(p);
    }
}

// The lane's entries in walk order (for the rare exact redo of a row).
template <int LANES, int BATCH, class Fn>
__device__ __forceinline__ void for_lane_entries(int sub, int cnt, Fn fn) {
    for (int k0 = sub * BATCH; k0 < cnt; k0 += LANES * BATCH)
        for (int j = 0; j < BATCH && k0 + j < cnt; ++j) fn(k0 + j);
}

// FP64 (32-byte partners) takes the strided split instead: lane s takes
// entries s, s + LANES, ..., BATCH at a time, so the lanes of a group gather
// consecutive entries -- neighbouring partners that share cache lines -- in
// each instruction (8 lanes x 2 entries: 213 us per launch at 1M atoms,
// against 234 us with the chunked split). pos_of(entry) gives the partner's
// double4; slots past the row end reread the lane's first entry and add
// nothing.
template <bool LFS, int LANES, int BATCH, class EntryFn, class PosFn>
__device__ __forceinline__ void walk_lane64(const ForceArgs& a, const double4& pi, int sub, int cnt,
                                            EntryFn entry, PosFn pos_of, Acc64& q) {
    constexpr int kStep = BATCH * LANES;
    uint32_t e[BATCH];
    auto indices = [&](int k0) {  // k0 < cnt
#pragma unroll
        for (int b = 0; b < BATCH; ++b) {
            const int k = k0 + b * LANES;
            e[b] = entry(k < cnt ? k : sub);
        }
    };
    if (sub < cnt) indices(sub);
    for (int k = sub; k < cnt; k += kStep) {
        double4 p[BATCH];
#pragma unroll
        for (int b = 0; b < BATCH; ++b) p[b] = pos_of(e[b]);
        indices(k + kStep < cnt ? k + kStep : k);  // the next batch's (this one's after the last)
#pragma unroll
        for (int b = 0; b < BATCH; ++b) pair64<LFS>(a, pi, p[b], q, k + b * LANES < cnt);
    }
}

// Mixed: at(entry) gives the partner's fixed-point position. Returns the
// lane's count of entries below the band's upper bound (mix_fixed_kb); q.k
// counts those inside the cutoff.
template <bool LFS, bool CUBIC, int LANES, int BATCH, class AtFn>
__device__ __forceinline__ int walk_lane_fixed(const ForceArgs& a, const int4& qi, const uint32_t* row,
                                               uint32_t self, int sub, int cnt, AtFn at, AccMix& q) {
    constexpr int kStep = BATCH * LANES;
    int kb = 0;
    uint32_t e[BATCH];
    int k0 = sub * BATCH;
    if (k0 < cnt) lane_indices<BATCH>(row, a.cap, k0, e);
    for (; k0 + BATCH <= cnt; k0 += kStep) {
        int4 p[BATCH];
#pragma unroll
        for (int j = 0; j < BATCH; ++j) p[j] = at(e[j]);
        lane_indices<BATCH>(row, a.cap, k0 + kStep, e);
#pragma unroll
        for (int j = 0; j < BATCH; ++j) mix_fixed_kb<LFS, CUBIC, false>(a, qi, p[j], true, q, kb);
    }
    if (k0 < cnt) {
        int4 p[BATCH];
#pragma unroll
        for (int j = 0; j < BATCH; ++j) p[j] = at(k0 + j < cnt ? e[j] : self);
#pragma unroll
        for (int j = 0; j < BATCH; ++j)
            mix_fixed_kb<LFS, CUBIC, true>(a, qi, p[j], k0 + j < cnt, q, kb);
    }
    return kb;
}

// The sum over a group of LANES consecutive lanes (LANES = 4, 8 or 16, the
// group aligned to LANES), left in every lane of the group. The partner
// values move by DPP lane permutations on the vector ALU instead of LDS
// permutes: swap neighbours, swap pairs (quad_perm), then pair each quad with
// the other quad of its half-row (row_half_mirror) and each half-row with the
// other half of its row (row_mirror). Each step adds two equal-length sums
// that are the same in both partners, so every lane of the group ends with
// the same bits. Every lane of the wave takes part.
template <int CTRL>
__device__ __forceinline__ int dpp_move(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ long long dpp_move(long long v) {
    const int lo = dpp_move<CTRL>(static_cast<int>(static_cast<unsigned long long>(v)));
    const int hi = dpp_move<CTRL>(static_cast<int>(static_cast<unsigned long long>(v) >> 32));
    return static_cast<long long>((static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) |
                                  static_cast<unsigned>(lo));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    return __longlong_as_double(dpp_move<CTRL>(static_cast<long long>(__double_as_longlong(v))));
}

template <int LANES, class T>
__device__ __forceinline__ T group_sum(T v) {
    static_assert(LANES == 4 || LANES == 8 || LANES == 16, "groups of 4, 8 or 16 lanes");
    v += dpp_move<0xB1>(v);                    // quad_perm [1, 0, 3, 2]
    v += dpp_move<0x4E>(v);                    // quad_perm [2, 3, 0, 1]
    if (LANES >= 8) v += dpp_move<0x141>(v);   // row_half_mirror
    if (LANES >= 16) v += dpp_move<0x140>(v);  // row_mirror
    return v;
}

template <bool LFS, bool MIXED, bool QP, bool CUBIC, bool KICK, int LANES>
__global__ __launch_bounds__(kBlock) void k_lj_lanes(ForceArgs a) {
    const long blk = xcd_block();
    if (blk * kBlock >= static_cast<long>(LANES) * a.n) return;  // padding block of the XCD-aware grid
    if (a.sc->overflow) return;  // a truncated list was built: md_run restores and retries
    const long gt = blk * kBlock + threadIdx.x;
    const long t = gt / LANES;
    const int sub = static_cast<int>(gt & (LANES - 1));
    // a group is whole lanes of one wave (LANES divides 64); every lane of the
    // block takes part in the shuffles, a lane past the last atom with an
    // empty row
    const bool live = t < a.n;
    const long tt = live ? t : 0;
    const int cnt = live ? min(a.ncnt[tt], a.cap) : 0;
    const uint32_t* row = a.nbr + oct_index(tt, 0, a.cap);
    auto entry = [&](int k) { return row[(k >> 3) * 512 + (k & 7)]; };
    const char* pb = reinterpret_cast<const char*>(a.pos);
    auto pos_of = [&](uint32_t e) { return *reinterpret_cast<const double4*>(pb + (static_cast<size_t>(e) << 1)); };
    double e = 0.0, w = 0.0, ke = 0.0;
    long long ef = 0, wf = 0, nt = 0;
    double fx, fy, fz;
    if (!MIXED) {
        Acc64 q;
        const double4 pi = a.pos[tt];
        walk_lane64<LFS, LANES, kFp64Batch>(a, pi, sub, cnt, entry, pos_of, q);
        fx = group_sum<LANES>(q.fx);
        fy = group_sum<LANES>(q.fy);
        fz = group_sum<LANES>(q.fz);
        e = group_sum<LANES>(q.e);
        w = group_sum<LANES>(q.w);
        nt = group_sum<LANES>(static_cast<long long>(q.n));
    } else {
        AccMix q;
        const MixK& m = QP ? a.mq : a.mx;
        auto exact = [&](const MixK& mm) {
            const double4 pi = a.pos[tt];
            for_lane_entries<LANES, kMixBatch>(
                sub, cnt, [&](int k) { mix_exact<LFS>(a, mm, pi, pos_of(entry(k)), q); });
        };
        if (QP) {
            const int4 qi = a.qpos[tt];
            const char* qb = reinterpret_cast<const char*>(a.qpos);
            auto at = [&](uint32_t off) { return *reinterpret_cast<const int4*>(qb + off); };
            const int kb = walk_lane_fixed<LFS, CUBIC, LANES, kMixBatch>(
                a, qi, row, entry_of(static_cast<int>(tt)), sub, cnt, at, q);
            if (group_sum<LANES>(kb != q.k ? 1 : 0) != 0) {  // a band hit in the group's row
                q = AccMix{};
                exact(a.mq);
            }
        } else {
            exact(a.mx);
        }
        bool sat = false;
        const FixRow r = mix_out<LFS>(m, q, sat);
        if (live && sat) atomicOr(&a.sc->fix_saturated, 1);
        fx = static_cast<double>(group_sum<LANES>(r.fx)) * kFixInv;
        fy = static_cast<double>(group_sum<LANES>(r.fy)) * kFixInv;
        fz = static_cast<double>(group_sum<LANES>(r.fz)) * kFixInv;
        ef = group_sum<LANES>(r.e);
        wf = group_sum<LANES>(r.w);
        nt = group_sum<LANES>(static_cast<long long>(q.k));
    }
    if (live && sub == 0) {
        a.frc[t] = make_double4(fx, fy, fz, 0.0);
        if (KICK) {
            double4 v = a.vel[t];
            const double ch = a.half_dt * (1.0 / v.w);
            v.x = fma(ch, fx, v.x);
            v.y = fma(ch, fy, v.y);
            v.z = fma(ch, fz, v.z);
            a.vel[t] = v;
            ke = (0.5 * v.w) * norm2_ref(v.x, v.y, v.z);
        }
    } else {
        e = w = 0.0;
        ef = wf = nt = 0;
    }
    write_partials<MIXED, kBlock / kWave, LANES>(a, blk, e, w, ke, ef, wf, nt);
}


// ---- bricks: partners staged in LDS ----
//
// One workgroup per brick of list cells (Bricks). It first copies its halo
// -- every atom within two cells of the brick, in the halo order the list
// build indexed -- from HBM into LDS, then walks the rows of its own atoms,
// whose entries are halo indices: every partner read is an LDS read instead
// of a gather through the vector L1. The halo rows are contiguous slot
// ranges (two where a row wraps in x), copied by one wave each with
// consecutive lanes on consecutive slots. The rows, their order, the lanes
// per atom and the arithmetic are those of k_lj_lanes, so forces, energies
// and virials are bit for bit the same (FP64 E and W: the same up to the
// grouping of the workgroup sums).
constexpr int kBrickMaxWaves = 8;

__device__ __forceinline__ int wave_incl_sum(int x, int lane) {
    for (int off = 1; off < kWave; off <<= 1) {
        const int y = __shfl_up(x, off, kWave);
        if (lane >= off) x += y;
    }
    return x;
}

// Halo copy plan: per halo row, its first halo index, its first slot and
// length before a wrap in x, and its first slot after the wrap.
constexpr int kMaxHaloRows = 160;
struct HaloRows {
    int dst[kMaxHaloRows + 1];
    int a0[kMaxHaloRows];
    int la[kMaxHaloRows];
    int b0[kMaxHaloRows];
};

// Copies the brick's halo into LDS: store(k, load(g)) for halo index k and
// its slot g. The row plans are formed by all threads at once; then every
// thread copies halo indices tid, tid + blockDim, ... (consecutive lanes on
// consecutive slots of a row), four loads in flight before the stores. Ends
// with a barrier.
template <class T, class Load, class Store>
__device__ __forceinline__ void brick_stage(const ForceArgs& a, const BrickGeo& o, const int* tb,
                                            HaloRows& hr, Load load, Store store) {
    const int nx = a.grid.nc[0], ny = a.grid.nc[1], nz = a.grid.nc[2];
    const int* __restrict__ cs = a.cell_start;
    const int hxn = o.ex + 4, hyn = o.ey + 4, rows = hyn * (o.ez + 4);
    const int xs = o.x0 - 2, xe = o.x0 + o.ex + 2;  // unwrapped cells [xs, xe) of each row
    for (int r = threadIdx.x; r < rows; r += blockDim.x) {
        int y = o.y0 - 2 + r % hyn, z = o.z0 - 2 + r / hyn;
        y += y < 0 ? ny : (y >= ny ? -ny : 0);
        z += z < 0 ? nz : (z >= nz ? -nz : 0);
        const int rb = nx * (y + ny * z);
        // the row's first piece, then (when it wraps) the rest from cell 0
        const int a0 = cs[xs < 0 ? rb + nx + xs : rb + xs];
        hr.a0[r] = a0;
        hr.la[r] = cs[xs < 0 ? rb + nx : rb + min(xe, nx)] - a0;
        hr.b0[r] = cs[rb];
        hr.dst[r] = tb[r * hxn];
    }
    if (threadIdx.x == 0) hr.dst[rows] = tb[rows * hxn];
    __syncthreads();
    const int total = hr.dst[rows];
    auto slot_of = [&](int k) {
        int lo = 0, hi = rows - 1;  // the last row r with dst[r] <= k
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (hr.dst[mid] <= k) lo = mid;
            else hi = mid - 1;
        }
        const int off = k - hr.dst[lo];
        return off < hr.la[lo] ? hr.a0[lo] + off : hr.b0[lo] + (off - hr.la[lo]);
    };
    const int step = blockDim.x;
    for (int k0 = threadIdx.x; k0 < total; k0 += 4 * step) {
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = k0 + u * step;
            if (k < total) v[u] = load(slot_of(k));
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = k0 + u * step;
            if (k < total) store(k, v[u]);
        }
    }
    __syncthreads();
}

// The brick's own rows (cells x0 .. x0+ex-1 of each (y, z) row, never
// wrapped): own atom k lies in row r with pre[r] <= k < pre[r + 1], at slot
// g0[r] + k - pre[r] and halo index l0[r] + k - pre[r]. Wave 0 fills it.
struct OwnRows {
    int pre[kWave + 1];
    int g0[kWave];
    int l0[kWave];
};

__device__ __forceinline__ void brick_own(const ForceArgs& a, const BrickGeo& o, const int* tb,
                                          OwnRows& s) {
    if (threadIdx.x >= kWave) return;
    const int lane = threadIdx.x;
    const int nr = o.ey * o.ez;  // at most 64 (brick widths of at most 8 cells)
    int cnt = 0;
    if (lane < nr) {
        const int oy = lane % o.ey, oz = lane / o.ey;
        const int h = ((oz + 2) * (o.ey + 4) + oy + 2) * (o.ex + 4) + 2;
        const int l0 = tb[h];
        cnt = tb[h + o.ex] - l0;
        s.l0[lane] = l0;
        s.g0[lane] = a.cell_start[a.grid.nc[0] * ((o.y0 + oy) + a.grid.nc[1] * (o.z0 + oz)) + o.x0];
    }
    const int incl = wave_incl_sum(cnt, lane);
    if (lane < nr) s.pre[lane] = incl - cnt;
    if (lane == nr - 1) s.pre[nr] = incl;
}

__device__ __forceinline__ int own_row(const OwnRows& s, int nr, int k) {
    int lo = 0, hi = nr - 1;  // the last row r with pre[r] <= k
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s.pre[mid] <= k) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Mixed: the halo holds (X, Y, Z, slot) per atom, 16 bytes; kMixLanes lanes
// per own atom walk its row as k_lj_lanes does (walk_lane_fixed, partners
// from LDS), so the same entries in the same order per lane and the same
// bits.
template <bool LFS, bool CUBIC, bool KICK>
__global__ __launch_bounds__(kBrickMaxWaves * kWave) void k_lj_brick(ForceArgs a) {
    constexpr int LANES = kMixLanes;
    extern __shared__ int4 halo[];
    __shared__ OwnRows own;
    __shared__ HaloRows hr;
    const long b = xcd_block();
    if (b >= a.bk.nb) return;      // padding block of the XCD-aware grid (whole block)
    if (a.sc->overflow) return;    // a truncated list was built: md_run restores and retries
    const BrickGeo o = brick_geo(a.bk, a.grid, static_cast<int>(b));
    const int* tb = a.bk.tab + a.bk.tab_off[b];
    brick_own(a, o, tb, own);
    brick_stage<int4>(
        a, o, tb, hr,
        [&](int g) {
            const int4 q = a.qpos[g];
            return make_int4(q.x, q.y, q.z, g);
        },
        [&](int k, const int4& v) { halo[k] = v; });
    const int nr = o.ey * o.ez;
    const int nown = own.pre[nr];
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave, nw = blockDim.x / kWave;
    const int sub = lane & (LANES - 1);
    constexpr int kPer = kWave / LANES;  // own atoms per wave chunk
    const MixK& m = a.mq;
    auto at = [&](uint32_t e) { return halo[e]; };
    double ke = 0.0;
    long long ef = 0, wf = 0, nt = 0;
    for (int c = wv; c * kPer < nown; c += nw) {
        const int k = c * kPer + lane / LANES;
        const bool live = k < nown;
        const int kk = live ? k : 0;
        const int r = own_row(own, nr, kk);
        const int i = own.g0[r] + (kk - own.pre[r]);
        const int4 qi = halo[own.l0[r] + (kk - own.pre[r])];
        const int cnt = live ? min(a.ncnt[i], a.cap) : 0;
        const uint32_t* row = a.nbr + oct_index(i, 0, a.cap);
        auto entry = [&](int j) { return row[(j >> 3) * 512 + (j & 7)]; };
        const uint32_t self = static_cast<uint32_t>(own.l0[r] + (kk - own.pre[r]));  // halo index
        AccMix q;
        const int kb = walk_lane_fixed<LFS, CUBIC, LANES, kMixBatch>(a, qi, row, self, sub, cnt, at, q);
        if (group_sum<LANES>(kb != q.k ? 1 : 0) != 0) {  // rare: a band hit, the row again exactly
            q = AccMix{};
            const double4 pi = a.pos[i];
            for_lane_entries<LANES, kMixBatch>(
                sub, cnt, [&](int j) { mix_exact<LFS>(a, m, pi, a.pos[halo[entry(j)].w], q); });
        }
        bool sat = false;
        const FixRow fr = mix_out<LFS>(m, q, sat);
        if (live && sat) atomicOr(&a.sc->fix_saturated, 1);
        const double fx = static_cast<double>(group_sum<LANES>(fr.fx)) * kFixInv;
        const double fy = static_cast<double>(group_sum<LANES>(fr.fy)) * kFixInv;
        const double fz = static_cast<double>(group_sum<LANES>(fr.fz)) * kFixInv;
        const long long ge = group_sum<LANES>(fr.e);
        const long long gw = group_sum<LANES>(fr.w);
        const long long gn = group_sum<LANES>(static_cast<long long>(q.k));
        if (live && sub == 0) {
            ef += ge;
            wf += gw;
            nt += gn;
            a.frc[i] = make_double4(fx, fy, fz, 0.0);
            if (KICK) {
                double4 v = a.vel[i];
                const double ch = a.half_dt * (1.0 / v.w);
                v.x = fma(ch, fx, v.x);
                v.y = fma(ch, fy, v.y);
                v.z = fma(ch, fz, v.z);
                a.vel[i] = v;
                ke += (0.5 * v.w) * norm2_ref(v.x, v.y, v.z);
            }
        }
    }
    write_partials<true, kBrickMaxWaves, LANES>(a, b, 0.0, 0.0, ke, ef, wf, nt);
}

// FP64: the halo holds each atom's FP64 position (double4 with the upload
// index in w, 32 bytes); kFp64Lanes lanes per own atom as k_lj_lanes, so the
// forces are k_lj_lanes' bit for bit.
constexpr int kBrick64MaxWaves = 16;

template <bool LFS, bool KICK>
__global__ __launch_bounds__(kBrick64MaxWaves * kWave) void k_lj_brick64(ForceArgs a) {
    constexpr int LANES = kFp64Lanes;
    extern __shared__ double4 hpos[];
    __shared__ OwnRows own;
    __shared__ HaloRows hr;
    const long b = xcd_block();
    if (b >= a.bk.nb) return;
    if (a.sc->overflow) return;
    const BrickGeo o = brick_geo(a.bk, a.grid, static_cast<int>(b));
    const int* tb = a.bk.tab + a.bk.tab_off[b];
    brick_own(a, o, tb, own);
    brick_stage<double4>(
        a, o, tb, hr, [&](int g) { return a.pos[g]; },
        [&](int k, const double4& v) { hpos[k] = v; });
    const int nr = o.ey * o.ez;
    const int nown = own.pre[nr];
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave, nw = blockDim.x / kWave;
    const int sub = lane & (LANES - 1);
    constexpr int kPer = kWave / LANES;  // own atoms per wave chunk
    double e = 0.0, w = 0.0, ke = 0.0;
    long long nt = 0;
    for (int c = wv; c * kPer < nown; c += nw) {
        const int k = c * kPer + lane / LANES;
        const bool live = k < nown;
        const int kk = live ? k : 0;
        const int r = own_row(own, nr, kk);
        const int i = own.g0[r] + (kk - own.pre[r]);
        const int cnt = live ? min(a.ncnt[i], a.cap) : 0;
        const uint32_t* row = a.nbr + oct_index(i, 0, a.cap);
        Acc64 q;
        const double4 pi = hpos[own.l0[r] + (kk - own.pre[r])];
        for (int j = sub; j < cnt; j += LANES)
            pair64<LFS>(a, pi, hpos[row[(j >> 3) * 512 + (j & 7)]], q);
        const double fx = group_sum<LANES>(q.fx);
        const double fy = group_sum<LANES>(q.fy);
        const double fz = group_sum<LANES>(q.fz);
        const double qe = group_sum<LANES>(q.e);
        const double qw = group_sum<LANES>(q.w);
        const long long qn = group_sum<LANES>(static_cast<long long>(q.n));
        if (live && sub == 0) {
            e += qe;
            w += qw;
            nt += qn;
            a.frc[i] = make_double4(fx, fy, fz, 0.0);
            if (KICK) {
                double4 v = a.vel[i];
                const double ch = a.half_dt * (1.0 / v.w);
                v.x = fma(ch, fx, v.x);
                v.y = fma(ch, fy, v.y);
                v.z = fma(ch, fz, v.z);
                a.vel[i] = v;
                ke += (0.5 * v.w) * norm2_ref(v.x, v.y, v.z);
            }
        }
    }
    write_partials<false, kBrick64MaxWaves, LANES>(a, b, e, w, ke, 0, 0, nt);
}

// Translation of a brick list to slot entries (entry_of), same layout, for
// md_neigh_export_pairs.
__global__ __launch_bounds__(kBrickMaxWaves * kWave) void k_brick_export(ForceArgs a, uint32_t* out) {
    extern __shared__ int hslot[];
    __shared__ OwnRows own;
    __shared__ HaloRows hr;
    const long b = xcd_block();
    if (b >= a.bk.nb) return;
    const BrickGeo o = brick_geo(a.bk, a.grid, static_cast<int>(b));
    const int* tb = a.bk.tab + a.bk.tab_off[b];
    brick_own(a, o, tb, own);
    brick_stage<int>(
        a, o, tb, hr, [&](int g) { return g; }, [&](int k, int g) { hslot[k] = g; });
    const int nr = o.ey * o.ez;
    const int nown = own.pre[nr];
    for (int k = threadIdx.x; k < nown; k += blockDim.x) {
        const int r = own_row(own, nr, k);
        const int i = own.g0[r] + (k - own.pre[r]);
        const int cnt = min(a.ncnt[i], a.cap);
        for (int j = 0; j < cnt; ++j) {
            const long at = oct_index(i, j, a.cap);
            out[at] = entry_of(hslot[a.nbr[at]]);
        }
    }
}

__global__ __launch_bounds__(kFinish) void k_lj_finish(ForceArgs a, int nb) {
    finish_step(a, a.mixed != 0, nb);
}

// lj_lanes is 16 or 8 up to kLanesMaxAtoms atoms, kMixLanes or kFp64Lanes above
template <bool LFS, bool MIXED, bool QP, bool CUBIC, bool KICK>
void dispatch_lanes(const ForceArgs& a, hipStream_t s) {
    const dim3 g(xcd_grid(static_cast<long>(lj_lanes(a.n, MIXED)) * a.n)), b(kBlock);
    switch (lj_lanes(a.n, MIXED)) {
        case 16: hipLaunchKernelGGL((k_lj_lanes<LFS, MIXED, QP, CUBIC, KICK, 16>), g, b, 0, s, a); break;
        case 8: hipLaunchKernelGGL((k_lj_lanes<LFS, MIXED, QP, CUBIC, KICK, 8>), g, b, 0, s, a); break;
        default: hipLaunchKernelGGL((k_lj_lanes<LFS, MIXED, QP, CUBIC, KICK, 4>), g, b, 0, s, a); break;
    }
}

template <bool LFS, bool MIXED, bool QP, bool CUBIC>
void dispatch_kick(const ForceArgs& a, hipStream_t s) {
    if (a.kick)
        dispatch_lanes<LFS, MIXED, QP, CUBIC, true>(a, s);
    else
        dispatch_lanes<LFS, MIXED, QP, CUBIC, false>(a, s);
}

// ---- direct (no list) evaluations: brute force and cells at the cutoff ----

__device__ void finish_direct(const DirectArgs& a, int nb) {
    __shared__ double red[2][kBlock];
    __shared__ long long redl[kBlock];
    double s0 = 0, s1 = 0;
    long long s2 = 0;
    for (int b = threadIdx.x; b < nb; b += kBlock) {
        s0 += a.part_d[3 * b];
        s1 += a.part_d[3 * b + 1];
        s2 += a.part_l[3 * b];
    }
    red[0][threadIdx.x] = s0;
    red[1][threadIdx.x] = s1;
    redl[threadIdx.x] = s2;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            red[0][threadIdx.x] += red[0][threadIdx.x + off];
            red[1][threadIdx.x] += red[1][threadIdx.x + off];
            redl[threadIdx.x] += redl[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a.sc->energy = 0.5 * red[0][0];
        a.sc->virial = 0.5 * red[1][0];
        a.sc->pairs_in_cut = redl[0];
    }
}

template <bool LFS>
__device__ __forceinline__ void direct_pair(const DirectArgs& a, const double4& pi, int oi, int j,
                                            double& e, double& w, double& fx, double& fy,
                                            double& fz, long long& np) {
    const bool flip = oi > a.orig[j];
    double dx, dy, dz;
    pair_delta(a.box, pi, a.pos[j], flip, dx, dy, dz);
    const double r2 = norm2_ref(dx, dy, dz);
    if (!(r2 < a.lj.rc2)) return;
    // the reference's two divisions, for the standalone calls
    const double s2 = a.lj.sig2 / r2;
    const double s6 = s2 * s2 * s2;
    double u = 4.0 * a.lj.eps * (s6 * s6 - s6);
    double ww = 24.0 * a.lj.eps * (2.0 * s6 * s6 - s6);
    if (LFS) {
        const double r = sqrt(r2);
        u = u - a.lj.u_rc - (r - a.lj.rc) * a.lj.dudr_rc;
        ww = ww + r * a.lj.dudr_rc;
    }
    e += u;
    w += ww;
    np += flip ? 0 : 1;
    const double c = flip ? -(ww / r2) : ww / r2;
    fx += c * dx;
    fy += c * dy;
    fz += c * dz;
}

template <bool LFS>
__global__ __launch_bounds__(kBlock) void k_lj_direct(DirectArgs a) {
    __shared__ double sm_d[kBlock / kWave];
    __shared__ long long sm_l[kBlock / kWave];
    const long t = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    double e = 0, w = 0, fx = 0, fy = 0, fz = 0;
    long long np = 0;
    if (t < a.n) {
        const double4 pi = a.pos[t];
        const int oi = a.orig[t];
        if (a.cell_atoms == nullptr) {
            for (int j = 0; j < static_cast<int>(a.n); ++j)
                if (j != t) direct_pair<LFS>(a, pi, oi, j, e, w, fx, fy, fz, np);
        } else {
            const Grid& g = a.grid;
            const int c = a.cell_of[t];
            const int cx = c % g.nc[0];
            const int cy = (c / g.nc[0]) % g.nc[1];
            const int cz = c / (g.nc[0] * g.nc[1]);
            for (int oz = -1; oz <= 1; ++oz)
                for (int oy = -1; oy <= 1; ++oy)
                    for (int ox = -1; ox <= 1; ++ox) {
                        const int nb = (cx + ox + g.nc[0]) % g.nc[0] +
                                       g.nc[0] * ((cy + oy + g.nc[1]) % g.nc[1] +
                                                  g.nc[1] * ((cz + oz + g.nc[2]) % g.nc[2]));
                        for (int k = a.cell_start[nb]; k < a.cell_start[nb + 1]; ++k) {
                            const int j = a.cell_atoms[k];
                            if (j != t) direct_pair<LFS>(a, pi, oi, j, e, w, fx, fy, fz, np);
                        }
                    }
        }
        if (a.frc != nullptr) a.frc[t] = make_double4(fx, fy, fz, 0.0);
        if (a.fx_out != nullptr) {
            a.fx_out[oi] = fx;
            a.fy_out[oi] = fy;
            a.fz_out[oi] = fz;
        }
    }
    const double be = block_reduce(e, sm_d, wave_sum);
    const double bw = block_reduce(w, sm_d, wave_sum);
    const long long bn = block_reduce(np, sm_l, wave_sum_ll);
    if (threadIdx.x == 0) {
        a.part_d[3 * blockIdx.x] = be;
        a.part_d[3 * blockIdx.x + 1] = bw;
        a.part_l[3 * blockIdx.x] = bn;
    }
}

__global__ __launch_bounds__(kBlock) void k_direct_finish(DirectArgs a, int nb) {
    finish_direct(a, nb);
}

// The cell-index kernels return at once when flag is given and 0 (a device
// rebuild decision); a null flag always builds.
__global__ void k_bin_index(const double4* __restrict__ pos, long n, Grid g,
                            int* __restrict__ cell_of, int* __restrict__ cell_count,
                            const int* flag) {
    if (flag != nullptr && *flag == 0) return;
    const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double4 p = pos[t];
    int c[3];
    const double r[3] = {p.x, p.y, p.z};
    for (int d = 0; d < 3; ++d) {
        const double f = fold_ref(r[d], g.L[d]);
        int k = static_cast<int>(floor((f + 0.5 * g.L[d]) / g.w[d]));
        c[d] = k < 0 ? 0 : (k >= g.nc[d] ? g.nc[d] - 1 : k);
    }
    const int cell = c[0] + g.nc[0] * (c[1] + g.nc[1] * c[2]);
    cell_of[t] = cell;
    atomicAdd(&cell_count[cell], 1);
}

__global__ void k_zero_int(int* a, long n, const int* flag) {
    if (flag != nullptr && *flag == 0) return;
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

__global__ __launch_bounds__(1024) void k_scan_plain(const int* __restrict__ cnt, long m,
                                                     int* __restrict__ start,
                                                     int* __restrict__ fill, const int* flag) {
    if (flag != nullptr && *flag == 0) return;
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const long chunk = (m + 1023) / 1024;
    const long b = t * chunk;
    const long e = b + chunk < m ? b + chunk : m;
    int s = 0;
    for (long c = b; c < e; ++c) s += cnt[c];
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
        start[c] = run;
        fill[c] = run;
        run += cnt[c];
    }
    if (t == 1023) start[m] = part[1023];
}

__global__ void k_place_index(const int* __restrict__ cell_of, long n, int* __restrict__ fill,
                              int* __restrict__ cell_atoms, const int* flag) {
    if (flag != nullptr && *flag == 0) return;
    const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    cell_atoms[atomicAdd(&fill[cell_of[t]], 1)] = static_cast<int>(t);
}

// Within each cell, the slot indices placed by k_place_index (in atomic
// arrival order, in tmp) go to cell_atoms in ascending order, so the index
// is a pure function of the positions. One wave per cell: a slot's rank is
// its count of smaller indices in the cell, the other indices read 64 at a
// time and broadcast by shuffles.
__global__ __launch_bounds__(kBlock) void k_rank_index(const int* __restrict__ start, long m,
                                                       const int* __restrict__ tmp,
                                                       int* __restrict__ cell_atoms,
                                                       const int* flag) {
    if (flag != nullptr && *flag == 0) return;
    const int lane = threadIdx.x & (kWave - 1);
    const long waves = static_cast<long>(gridDim.x) * (kBlock / kWave);
    for (long c = static_cast<long>(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave; c < m;
         c += waves) {
        const int b = start[c], cnt = start[c + 1] - b;
        for (int k0 = 0; k0 < cnt; k0 += kWave) {
            const int key = k0 + lane < cnt ? tmp[b + k0 + lane] : 0x7fffffff;
            int rank = 0;
            for (int j0 = 0; j0 < cnt; j0 += kWave) {
                const int other = j0 + lane < cnt ? tmp[b + j0 + lane] : 0x7fffffff;
                const int lim = min(kWave, cnt - j0);
                for (int q = 0; q < lim; ++q) rank += __shfl(other, q, kWave) < key ? 1 : 0;
            }
            if (k0 + lane < cnt) cell_atoms[b + rank] = key;
        }
    }
}

__global__ void k_chain_first(Chain* chain, Scalars* sc, double delta) {
    __shared__ ChainWork cw;
    if (threadIdx.x != 0) return;
    double ke = sc->kinetic;
    sc->vscale = chain_step(chain, cw, ke, delta, false).scale;
}

inline unsigned grid_of(long n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

}  // namespace

void launch_chain_first(Chain* chain, Scalars* sc, double delta, hipStream_t s) {
    hipLaunchKernelGGL(k_chain_first, dim3(1), dim3(64), 0, s, chain, sc, delta);
}

long lj_partials(long n) { return std::max(lj_groups(n, false), lj_groups(n, true)); }

inline unsigned brick_grid(const Bricks& k) {
    return static_cast<unsigned>((k.nb + kXcd - 1) / kXcd * kXcd);
}

template <bool LFS>
void launch_brick(const ForceArgs& a, hipStream_t s) {
    const Bricks& k = a.bk;
    const dim3 g(brick_grid(k));
    if (a.mixed) {
        const dim3 b(static_cast<unsigned>(k.waves_mixed * kWave));
        const size_t lds = static_cast<size_t>(k.halo_cap) * sizeof(int4);
        if (a.quant.cubic) {
            if (a.kick) hipLaunchKernelGGL((k_lj_brick<LFS, true, true>), g, b, lds, s, a);
            else hipLaunchKernelGGL((k_lj_brick<LFS, true, false>), g, b, lds, s, a);
        } else {
            if (a.kick) hipLaunchKernelGGL((k_lj_brick<LFS, false, true>), g, b, lds, s, a);
            else hipLaunchKernelGGL((k_lj_brick<LFS, false, false>), g, b, lds, s, a);
        }
    } else {
        const dim3 b(static_cast<unsigned>(k.waves_fp64 * kWave));
        const size_t lds = static_cast<size_t>(k.halo_cap) * sizeof(double4);
        if (a.kick) hipLaunchKernelGGL((k_lj_brick64<LFS, true>), g, b, lds, s, a);
        else hipLaunchKernelGGL((k_lj_brick64<LFS, false>), g, b, lds, s, a);
    }
}

void launch_brick_export(const ForceArgs& a, uint32_t* out, hipStream_t s) {
    const Bricks& k = a.bk;
    hipLaunchKernelGGL(k_brick_export, dim3(brick_grid(k)), dim3(256),
                       static_cast<size_t>(k.halo_cap) * sizeof(int), s, a, out);
}

void launch_lj_force(const ForceArgs& a, hipStream_t s) {
    if (a.bk.on) {
        if (a.lj.lfs) launch_brick<true>(a, s);
        else launch_brick<false>(a, s);
        return;
    }
    const bool qp = a.qpos != nullptr;
    const bool cubic = a.quant.cubic != 0;
    if (a.lj.lfs) {
        if (!a.mixed) dispatch_kick<true, false, false, false>(a, s);
        else if (qp && cubic) dispatch_kick<true, true, true, true>(a, s);
        else if (qp) dispatch_kick<true, true, true, false>(a, s);
        else dispatch_kick<true, true, false, false>(a, s);
    } else {
        if (!a.mixed) dispatch_kick<false, false, false, false>(a, s);
        else if (qp && cubic) dispatch_kick<false, true, true, true>(a, s);
        else if (qp) dispatch_kick<false, true, true, false>(a, s);
        else dispatch_kick<false, true, false, false>(a, s);
    }
}

void launch_lj_finish(const ForceArgs& a, hipStream_t s) {
    // the pair kernel writes partials for its lj_groups(n) blocks (padding blocks of
    // the XCD-aware grid none)
    const long nb = a.bk.on ? a.bk.nb : lj_groups(a.n, a.mixed != 0);
    hipLaunchKernelGGL(k_lj_finish, dim3(kFinishGroups), dim3(kFinish), 0, s, a,
                       static_cast<int>(nb));
}

MixK mix_constants(const LJConst& k, double s, double band) {
    MixK m{};
    m.inv_s = 1.0 / s;
    m.sig2s = static_cast<float>(k.sig2 / (s * s));
    const double rc2s = k.rc2 / (s * s);
    m.lo = rc2s * (1.0 - band);
    m.hi = rc2s * (1.0 + band);
    m.e4 = 4.0 * k.eps;
    m.e24 = 24.0 * k.eps;
    m.fscale = m.e24 / s;
    if (k.lfs) {
        m.ushift = k.u_rc - k.rc * k.dudr_rc;
        m.urs = k.dudr_rc * s;
        m.krs = k.eps != 0.0 ? static_cast<float>(k.dudr_rc * s / m.e24) : 0.0f;
    }
    return m;
}

LJConst lj_constants(double sigma, double epsilon, double cutoff, int linear_force_shift) {
    LJConst k{};
    const double s2 = (sigma * sigma) / (cutoff * cutoff);
    const double s6 = s2 * s2 * s2;
    const double u_rc = 4.0 * epsilon * (s6 * s6 - s6);
    const double w_rc = 24.0 * epsilon * (2.0 * s6 * s6 - s6);
    k.sig2 = sigma * sigma;
    k.eps = epsilon;
    k.rc = cutoff;
    k.rc2 = cutoff * cutoff;
    k.u_rc = u_rc;
    k.dudr_rc = -w_rc / cutoff;
    k.sig2f = static_cast<float>(k.sig2);
    k.epsf = static_cast<float>(k.eps);
    k.rc2f = static_cast<float>(k.rc2);
    k.rcf = static_cast<float>(k.rc);
    k.u_rcf = static_cast<float>(k.u_rc);
    k.dudr_rcf = static_cast<float>(k.dudr_rc);
    k.e4s = (4.0f * k.epsf) * 4294967296.0f;
    k.e24s = (24.0f * k.epsf) * 4294967296.0f;
    k.u_rcs = k.u_rcf * 4294967296.0f;
    k.dudr_rcs = k.dudr_rcf * 4294967296.0f;
    k.lfs = linear_force_shift ? 1 : 0;
    return k;
}

Grid cell_grid(const double L[3], double cutoff) {
    Grid g{};
    for (int d = 0; d < 3; ++d) {
        g.L[d] = L[d];
        const int k = static_cast<int>(floor(L[d] / cutoff));
        g.nc[d] = k > 1 ? k : 1;
        g.w[d] = L[d] / static_cast<double>(g.nc[d]);
        g.inv_w[d] = static_cast<double>(g.nc[d]) / L[d];
    }
    g.usable = (g.nc[0] >= 3 && g.nc[1] >= 3 && g.nc[2] >= 3) ? 1 : 0;
    return g;
}

void launch_lj_direct(const DirectArgs& a, hipStream_t s) {
    const dim3 g(grid_of(a.n));
    if (a.lj.lfs)
        hipLaunchKernelGGL(k_lj_direct<true>, g, dim3(kBlock), 0, s, a);
    else
        hipLaunchKernelGGL(k_lj_direct<false>, g, dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_direct_finish, dim3(1), dim3(kBlock), 0, s, a, static_cast<int>(g.x));
}

void launch_bin_index(long n, const double4* pos, const Grid& g, int* cell_of, int* cell_count,
                      int* cell_start, int* cell_fill, int* cell_tmp, int* cell_atoms,
                      const int* flag, hipStream_t s) {
    const long m = static_cast<long>(g.nc[0]) * g.nc[1] * g.nc[2];
    hipLaunchKernelGGL(k_zero_int, dim3(grid_of(m)), dim3(kBlock), 0, s, cell_count, m, flag);
    hipLaunchKernelGGL(k_bin_index, dim3(grid_of(n)), dim3(kBlock), 0, s, pos, n, g, cell_of,
                       cell_count, flag);
    hipLaunchKernelGGL(k_scan_plain, dim3(1), dim3(1024), 0, s, cell_count, m, cell_start, cell_fill,
                       flag);
    hipLaunchKernelGGL(k_place_index, dim3(grid_of(n)), dim3(kBlock), 0, s, cell_of, n, cell_fill,
                       cell_tmp, flag);
    const long rank_blocks = (m + kBlock / kWave - 1) / (kBlock / kWave);
    hipLaunchKernelGGL(k_rank_index, dim3(static_cast<unsigned>(rank_blocks < 8192 ? rank_blocks : 8192)),
                       dim3(kBlock), 0, s, cell_start, m, cell_tmp, cell_atoms, flag);
}

}  // namespace mdk
