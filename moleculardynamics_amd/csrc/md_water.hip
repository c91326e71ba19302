// Rigid SPC/E water with Ewald electrostatics on the device.
//
// Reference semantics (forcefield/ewald.cpp:33-175, spce_forcefield.cpp:26-41,
// constrain/rigid_water.cpp:61-161, integrate/constrained_velocity_verlet.hpp:27-57):
//  real   C qq erfc(a r)/r over intermolecular pairs with r < rc (minimum
//         image), w = C qq (2a/sqrt(pi) e^{-a^2 r^2} + erfc(a r)/r), f = (w/r^2) d;
//  LJ     oxygen-oxygen, Truncated at rc (plus the analytic tail, host side);
//  self   -C a/sqrt(pi) sum q^2 (host side);
//  intra  -C qq erf(a r)/r over the three pairs of each molecule (minimum image),
//         w = C qq (2a/sqrt(pi) e^{-a^2 r^2} - erf(a r)/r);
//  recip  sum over 0 < |n|^2 < n2max of C (2 pi/V) e^{-k^2/4a^2}/k^2 |S(k)|^2,
//         S(k) = sum_i q_i e^{i k.r_i}, f_i = 2 C (2 pi/V) e^{..}/k^2 q_i (Re S sin - Im S cos) k.
//
// Layout: sites in molecule order (O, H, H) as double4 (x, y, z, charge),
// velocities double4 (vx, vy, vz, mass). Real space and LJ run over a full
// Verlet list of intermolecular pairs (same wave-blocked rows as the LJ
// path). Reciprocal space uses per-site tables e^{i n k0 x} (n = 0..kmax per
// axis; negative n by conjugation) and only one of each +-k pair (S(-k) is
// the conjugate of S(k), so the +k and -k terms are equal):
//  k_sk     one thread per k-vector, sites staged in LDS tiles, partial S(k)
//           per site tile, then k_sk_reduce sums the tiles in order;
//  k_recip  (site tile, column chunk) blocks: a thread walks its site over
//           the chunk's columns of consecutive nz with one complex product
//           per k-vector, and k_recip_reduce adds the chunks in order.
// The real-space kernel takes kWaterTpa lanes per site, the list build one
// wave per site, so a 12,288-site system fills the chip.
// The mixed path evaluates erfc in FP32 with Abramowitz & Stegun 7.1.26
// (|error| <= 1.5e-7); the k-space sums stay FP64.
#include "md_common.hpp"
#include "md_water.hpp"

namespace mdk {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kSiteTile = 32;  // sites per LDS tile of the S(k) kernel
constexpr int kSegLen = 8;     // k-vectors per column segment of the S(k) kernel
constexpr int kSkSub = 8;      // lanes per segment of the S(k) kernel (sites split among them)
constexpr int kWaterTpa = 16;  // lanes per site of the real-space kernel
constexpr int kWaterBatch = 2;  // row entries per lane in flight in the real-space kernel

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// erfc(x) for x >= 0, Abramowitz & Stegun 7.1.26: 1 - (a1 t + ... + a5 t^5) e^{-x^2},
// t = 1/(1 + p x); the e^{-x^2} factor is shared with the force term.
__device__ __forceinline__ float erfc_as(float x, float ex2) {
    const float p = 0.3275911f;
    const float a1 = 0.254829592f, a2 = -0.284496736f, a3 = 1.421413741f, a4 = -1.453152027f,
                a5 = 1.061405429f;
    const float t = 1.0f / (1.0f + p * x);
    const float poly = t * (a1 + t * (a2 + t * (a3 + t * (a4 + t * a5))));
    return poly * ex2;
}

#define GRID_STRIDE(i, n)                                                              \
    for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < (n); \
         i += static_cast<long>(gridDim.x) * blockDim.x)

// Full list of intermolecular partners within rc + skin: from cells of width
// >= (rc + skin)/2 (index array in site order within a cell), the 5 x 5 runs
// of five x-adjacent cells around the site's cell, when at least 5 cells fit
// per side; all sites otherwise. Every pair within rc + skin lies in such a
// run. One wave per site tests 64 consecutive candidates at once and keeps
// them in candidate order by a ballot and a prefix count, so a row is the
// row of the serial scan (runs in order, sites in index order within a cell).
__global__ __launch_bounds__(kBlock) void k_wlist(WaterArgs a, Grid g, const int* __restrict__ cell_of,
                                                  const int* __restrict__ cell_start,
                                                  const int* __restrict__ cell_atoms) {
    if (a.build_flag != nullptr && *a.build_flag == 0) return;
    const int lane = threadIdx.x & (kWave - 1);
    const unsigned long long below = (1ull << lane) - 1;
    const long waves = static_cast<long>(gridDim.x) * (kBlock / kWave);
    for (long i = static_cast<long>(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave; i < a.n;
         i += waves) {
        const double4 pi = a.pos[i];
        const long mi = i / 3;
        int cnt = 0;
        // candidates idx[b..e), or the sites b..e when idx is null
        auto scan = [&](const int* idx, int b, int e) {
            for (int k0 = b; k0 < e; k0 += kWave) {
                const int k = k0 + lane;
                bool keep = false;
                int j = 0;
                if (k < e) {
                    j = idx != nullptr ? idx[k] : k;
                    if (j / 3 != mi) {
                        double dx, dy, dz;
                        pair_delta(a.box, pi, a.pos[j], i > j, dx, dy, dz);
                        keep = norm2_ref(dx, dy, dz) < a.rl2;
                    }
                }
                const unsigned long long m = __ballot(keep);
                if (keep) {
                    const int at = cnt + __popcll(m & below);
                    if (at < a.cap)
                        a.nbr[nbr_index(i, at, a.cap)] = static_cast<uint32_t>(j) | (i > j ? kFlipBit : 0u);
                }
                cnt += __popcll(m);
            }
        };
        if (g.usable) {
            const int nx = g.nc[0], ny = g.nc[1], nz = g.nc[2];
            const int c = cell_of[i];
            const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
            const int xa = cx >= 2 ? cx - 2 : 0, xb = cx + 2 < nx ? cx + 2 : nx - 1;
            for (int oz = -2; oz <= 2; ++oz) {
                int zz = cz + oz;
                zz += zz < 0 ? nz : (zz >= nz ? -nz : 0);
                for (int oy = -2; oy <= 2; ++oy) {
                    int yy = cy + oy;
                    yy += yy < 0 ? ny : (yy >= ny ? -ny : 0);
                    const int b = nx * (yy + ny * zz);
                    // five x-adjacent cells are contiguous in the index array,
                    // split in two where they wrap the x axis
                    scan(cell_atoms, cell_start[b + xa], cell_start[b + xb + 1]);
                    if (cx < 2)
                        scan(cell_atoms, cell_start[b + nx + cx - 2], cell_start[b + nx]);
                    else if (cx + 2 >= nx)
                        scan(cell_atoms, cell_start[b], cell_start[b + cx + 3 - nx]);
                }
            }
        } else {
            scan(nullptr, 0, static_cast<int>(a.n));
        }
        if (lane == 0) {
            a.ncnt[i] = cnt;
            if (cnt > a.cap) atomicOr(a.overflow, 1);
            a.ref[i] = pi;
        }
    }
}

// Real-space Coulomb + oxygen LJ, intramolecular correction; forces are
// written (not accumulated), the reciprocal kernel adds its part.
// kWaterTpa lanes per site: lane s of a site takes row entries s, s + kWaterTpa,
// ... and the intramolecular partner s (s < 2); the lanes' sums are combined
// by a fixed shuffle tree.
template <bool MIXED>
__global__ __launch_bounds__(kBlock) void k_wreal(WaterArgs a) {
    __shared__ double sm[kBlock / kWave];
    const long g = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    const long i = g / kWaterTpa;
    const int sub = static_cast<int>(g % kWaterTpa);
    double ereal = 0.0, elj = 0.0, eintra = 0.0, wsum = 0.0;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (i < a.n) {
        const double4 pi = a.pos[i];
        const bool oi = (i % 3) == 0;
        const int cnt = min(a.ncnt[i], a.cap);
        const uint32_t* row = a.nbr + nbr_index(i, 0, a.cap);
        auto term = [&](uint32_t jj, const double4& pj) {
            const int j = static_cast<int>(jj & kIndexMask);
            const bool flip = (jj & kFlipBit) != 0;
            double dx, dy, dz;
            pair_delta(a.box, pi, pj, flip, dx, dy, dz);
            const double r2 = norm2_ref(dx, dy, dz);
            if (!(r2 < a.rc2)) return;
            double w = 0.0;
            const double qq = pi.w * pj.w;
            if (MIXED) {
                const float rf = sqrtf(static_cast<float>(r2));
                const float ar = static_cast<float>(a.alpha) * rf;
                const float ex2 = __expf(-ar * ar);
                const float ec = erfc_as(ar, ex2);
                const float cq = static_cast<float>(a.coulomb * qq);
                const float inv_r = 1.0f / rf;
                ereal += static_cast<double>(cq * ec * inv_r);
                w += static_cast<double>(cq * (static_cast<float>(a.c2a) * ex2 + ec * inv_r));
            } else {
                const double r = sqrt(r2);
                const double ec = erfc(a.alpha * r);
                ereal += a.coulomb * qq * ec / r;
                w += a.coulomb * qq * (a.c2a * exp(-a.alpha * a.alpha * r2) + ec / r);
            }
            if (oi && (j % 3) == 0) {
                const double s2 = a.sig2 / r2;
                const double s6 = s2 * s2 * s2;
                elj += 4.0 * a.eps * (s6 * s6 - s6);
                w += 24.0 * a.eps * (2.0 * s6 * s6 - s6);
            }
            wsum += w;
            const double c = flip ? -(w / r2) : w / r2;
            fx += c * dx;
            fy += c * dy;
            fz += c * dz;
        };
        // kWaterBatch entries of the lane at a time: their partner gathers
        // together, then the next batch's indices, then the terms in entry
        // order (the order one entry at a time adds them)
        constexpr int kStep = kWaterBatch * kWaterTpa;
        uint32_t e[kWaterBatch];
        auto indices = [&](int k0) {  // k0 < cnt; slots past the row end reread entry sub
#pragma unroll
            for (int b = 0; b < kWaterBatch; ++b) {
                const int k = k0 + b * kWaterTpa;
                e[b] = row[(k < cnt ? k : sub) * 64];
            }
        };
        if (sub < cnt) indices(sub);
        for (int k = sub; k < cnt; k += kStep) {
            double4 p[kWaterBatch];
            uint32_t cur[kWaterBatch];  // this batch's entries (e takes the next batch's)
#pragma unroll
            for (int b = 0; b < kWaterBatch; ++b) {
                cur[b] = e[b];
                p[b] = a.pos[e[b] & kIndexMask];
            }
            indices(k + kStep < cnt ? k + kStep : k);
#pragma unroll
            for (int b = 0; b < kWaterBatch; ++b)
                if (k + b * kWaterTpa < cnt) term(cur[b], p[b]);
        }
        // intramolecular correction: lanes 0 and 1 of the site take its two partners
        if (sub < 2) {
            const long m0 = (i / 3) * 3;
            const long j = m0 + ((m0 + sub) >= i ? sub + 1 : sub);
            const double4 pj = a.pos[j];
            const bool flip = i > j;
            double dx, dy, dz;
            pair_delta(a.box, pi, pj, flip, dx, dy, dz);
            const double r2 = norm2_ref(dx, dy, dz);
            const double r = sqrt(r2);
            const double qq = pi.w * pj.w;
            const double ef = erf(a.alpha * r);
            eintra += -a.coulomb * qq * ef / r;
            const double w = a.coulomb * qq * (a.c2a * exp(-a.alpha * a.alpha * r2) - ef / r);
            const double c = flip ? -(w / r2) : w / r2;
            fx += c * dx;
            fy += c * dy;
            fz += c * dz;
        }
    }
    for (int off = kWaterTpa / 2; off > 0; off >>= 1) {
        fx += __shfl_xor(fx, off, kWave);
        fy += __shfl_xor(fy, off, kWave);
        fz += __shfl_xor(fz, off, kWave);
    }
    if (i < a.n && sub == 0) a.frc[i] = make_double4(fx, fy, fz, 0.0);
    const double b0 = block_reduce(ereal, sm, wave_sum);
    const double b1 = block_reduce(elj, sm, wave_sum);
    const double b2 = block_reduce(eintra, sm, wave_sum);
    const double b3 = block_reduce(wsum, sm, wave_sum);
    if (threadIdx.x == 0) {
        a.part[4 * blockIdx.x + 0] = b0;
        a.part[4 * blockIdx.x + 1] = b1;
        a.part[4 * blockIdx.x + 2] = b2;
        a.part[4 * blockIdx.x + 3] = b3;
    }
}

// Per-site phase tables: tab[(d * (kmax+1) + n) * nsite + i] = e^{i n k0_d r_d}
// by repeated multiplication (n = 0..kmax).
__global__ void k_weik(WaterArgs a) {
    GRID_STRIDE(i, a.n) {
        const double4 p = a.pos[i];
        const double r[3] = {p.x, p.y, p.z};
        for (int d = 0; d < 3; ++d) {
            double s, c;
            sincos(a.k0[d] * r[d], &s, &c);
            const double2 e1 = make_double2(c, s);
            double2 e = make_double2(1.0, 0.0);
            for (int nn = 0; nn <= a.kmax; ++nn) {
                a.eik[(static_cast<long>(d) * (a.kmax + 1) + nn) * a.n + i] = e;
                e = cmul(e, e1);
            }
        }
    }
}

__device__ __forceinline__ double2 phase(const double2* t, int nn) {
    // e^{i n theta} for n of either sign from the table of n >= 0
    const double2 v = t[nn < 0 ? -nn : nn];
    return nn < 0 ? make_double2(v.x, -v.y) : v;
}

// Partial structure factors: block (segment chunk, site group). A segment is
// up to kSegLen k-vectors of one column (same nx, ny; nz rising by one).
// kSkSub adjacent lanes own a segment and take every kSkSub-th site of the
// group; per site one phase product e^{i(k_x x + k_y y + k_z0 z)} from the
// per-axis tables, then e^{i k0z z} steps along the segment. Sites are staged
// in LDS (their q and the three per-axis tables) kSiteTile at a time; the
// lanes of a segment are combined by a fixed shuffle tree at the end.
__global__ __launch_bounds__(kBlock) void k_sk(WaterArgs a, int sites_per_group) {
    __shared__ double2 tab[kSiteTile][3][kMaxK + 1];
    __shared__ double qs[kSiteTile];
    const int gt = blockIdx.x * kBlock + threadIdx.x;
    const int sg = gt / kSkSub, sub = gt % kSkSub;
    const bool valid = sg < a.nseg;
    int4 sv = make_int4(0, 0, 0, 0);
    if (valid) sv = a.kseg[sg];
    const long s0 = static_cast<long>(blockIdx.y) * sites_per_group;
    const long s1 = min(s0 + sites_per_group, a.n);
    double re[kSegLen], im[kSegLen];
#pragma unroll
    for (int m = 0; m < kSegLen; ++m) {
        re[m] = 0.0;
        im[m] = 0.0;
    }
    for (long t0 = s0; t0 < s1; t0 += kSiteTile) {
        const int cnt = static_cast<int>(min(static_cast<long>(kSiteTile), s1 - t0));
        __syncthreads();
        // thread (d, s) copies one table row; consecutive threads read
        // consecutive sites of the [axis][n][site] table
        for (int ds = threadIdx.x; ds < 3 * kSiteTile; ds += kBlock) {
            const int d = ds / kSiteTile, s = ds % kSiteTile;
            if (s < cnt) {
                const double2* src = a.eik + static_cast<long>(d) * (a.kmax + 1) * a.n + t0 + s;
                for (int nn = 0; nn <= a.kmax; ++nn) tab[s][d][nn] = src[static_cast<long>(nn) * a.n];
            }
        }
        for (int s = threadIdx.x; s < cnt; s += kBlock) qs[s] = a.pos[t0 + s].w;
        __syncthreads();
        if (valid) {
            // two sites at a time: two independent recursion chains
            for (int s = sub; s < cnt; s += 2 * kSkSub) {
                const int s2 = s + kSkSub < cnt ? s + kSkSub : s;
                const double w2 = s + kSkSub < cnt ? qs[s2] : 0.0;
                double2 e = cmul(cmul(phase(tab[s][0], sv.x), phase(tab[s][1], sv.y)),
                                 phase(tab[s][2], sv.z));
                double2 f = cmul(cmul(phase(tab[s2][0], sv.x), phase(tab[s2][1], sv.y)),
                                 phase(tab[s2][2], sv.z));
                const double2 step = tab[s][2][1], step2 = tab[s2][2][1];
                const double q = qs[s];
#pragma unroll
                for (int m = 0; m < kSegLen; ++m) {
                    re[m] += q * e.x + w2 * f.x;
                    im[m] += q * e.y + w2 * f.y;
                    e = cmul(e, step);
                    f = cmul(f, step2);
                }
            }
        }
    }
#pragma unroll
    for (int o = 1; o < kSkSub; o <<= 1) {
#pragma unroll
        for (int m = 0; m < kSegLen; ++m) {
            re[m] += __shfl_xor(re[m], o, kWave);
            im[m] += __shfl_xor(im[m], o, kWave);
        }
    }
    if (valid && sub == 0) {
        double2* out = a.skpart + static_cast<long>(blockIdx.y) * a.nk + a.kseg0[sg];
#pragma unroll
        for (int m = 0; m < kSegLen; ++m)
            if (m < sv.w) out[m] = make_double2(re[m], im[m]);
    }
}

// S(k) = sum of the site-group partials in order; E_k term and the force
// coefficient G(k) = 2 factor(k) (factor(k) = C (2 pi/V) e^{-k^2/4a^2}/k^2;
// both signs of k counted once each).
__global__ void k_sk_reduce(WaterArgs a, int groups) {
    __shared__ double sm[kBlock / kWave];
    const int kv = blockIdx.x * kBlock + threadIdx.x;
    double e = 0.0;
    if (kv < a.nk) {
        double re = 0.0, im = 0.0;
        for (int g = 0; g < groups; ++g) {
            const double2 p = a.skpart[static_cast<long>(g) * a.nk + kv];
            re += p.x;
            im += p.y;
        }
        a.sk[kv] = make_double2(re, im);
        a.skf[kv] = make_double2(4.0 * a.kfac[kv] * re, 4.0 * a.kfac[kv] * im);
        e = 2.0 * a.kfac[kv] * (re * re + im * im);
    }
    const double b = block_reduce(e, sm, wave_sum);
    if (threadIdx.x == 0) a.part2[blockIdx.x] = b;
}

__device__ __forceinline__ double2 site_phase(const double2* t, int nn, long n) {
    // e^{i n theta} of one site for n of either sign; t points at the site's n = 0 entry
    const double2 v = t[static_cast<long>(nn < 0 ? -nn : nn) * n];
    return nn < 0 ? make_double2(v.x, -v.y) : v;
}

// Reciprocal forces: f_i += q_i sum_k [4 factor(k) (Re S sin - Im S cos)] k
// over the half space (the -k term equals the +k term; skf = 4 factor S).
// Block (site tile, column chunk): each thread walks the chunk's columns for
// its site, advancing the phase along a column by e^{i k0z z} (one complex
// product per k-vector after the column's first). Along a column kx and ky
// are fixed, so the force needs only sum t and sum t nz per column. The
// chunk's sums go to fpart[chunk]; k_recip_reduce adds the chunks in order.
__global__ __launch_bounds__(kBlock) void k_recip(WaterArgs a) {
    const long i = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= a.n) return;
    const int chunk = blockIdx.y;
    const int c0 = static_cast<int>(static_cast<long>(a.ncol) * chunk / a.nchunk);
    const int c1 = static_cast<int>(static_cast<long>(a.ncol) * (chunk + 1) / a.nchunk);
    const long n = a.n;
    const double2* tx = a.eik + (0L * (a.kmax + 1)) * n + i;
    const double2* ty = a.eik + (1L * (a.kmax + 1)) * n + i;
    const double2* tz = a.eik + (2L * (a.kmax + 1)) * n + i;
    const double2 step = tz[n];  // e^{i k0z z}
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int c = c0; c < c1; ++c) {
        const int4 col = a.kcol[c];
        const double2* s = a.skf + a.kcol0[c];
        double2 e = cmul(cmul(site_phase(tx, col.x, n), site_phase(ty, col.y, n)),
                         site_phase(tz, col.z, n));
        double sum_t = 0.0, sum_tz = 0.0, nz = static_cast<double>(col.z);
        for (int m = 0; m < col.w; ++m) {
            const double2 sf = s[m];
            const double t = sf.x * e.y - sf.y * e.x;
            sum_t += t;
            sum_tz += t * nz;
            nz += 1.0;
            e = cmul(e, step);
        }
        fx += sum_t * (a.k0[0] * col.x);
        fy += sum_t * (a.k0[1] * col.y);
        fz += sum_tz * a.k0[2];
    }
    a.fpart[chunk * n + i] = make_double4(fx, fy, fz, 0.0);
}

__global__ void k_recip_reduce(WaterArgs a) {
    GRID_STRIDE(i, a.n) {
        double fx = 0.0, fy = 0.0, fz = 0.0;
        for (int c = 0; c < a.nchunk; ++c) {
            const double4 p = a.fpart[c * a.n + i];
            fx += p.x;
            fy += p.y;
            fz += p.z;
        }
        const double q = a.pos[i].w;
        double4 f = a.frc[i];
        f.x += q * fx;
        f.y += q * fy;
        f.z += q * fz;
        a.frc[i] = f;
    }
}

// ---- rigid-body constraints: one thread per molecule ------------------------
// Gauss-Seidel SHAKE on (O,H1), (O,H2), (H1,H2) along the reference bonds until
// every relative violation is within tol; a molecule that converged stays
// unchanged, so iterating molecules independently is the reference's global
// sweep (rigid_water.cpp:61-116). iters[m] receives the sweep count or -1
// (no convergence) / -2 (singular).
__global__ void k_shake(WaterArgs a, const double4* __restrict__ refp, int* __restrict__ status) {
    GRID_STRIDE(m, a.n / 3) {
        const long b = 3 * m;
        double4 r[3] = {a.pos[b], a.pos[b + 1], a.pos[b + 2]};
        const double4 rr[3] = {refp[b], refp[b + 1], refp[b + 2]};
        const double mass[3] = {a.vel[b].w, a.vel[b + 1].w, a.vel[b + 2].w};
        const int ci[3] = {0, 0, 1}, cj[3] = {1, 2, 2};
        const double tgt[3] = {a.oh, a.oh, a.hh};
        int it = 0, res = -1;
        for (; it < a.max_iter; ++it) {
            double worst = 0.0;
            for (int c = 0; c < 3; ++c) {
#pragma clang fp contract(off)
                const int p = ci[c], q = cj[c];
                const double d2 = tgt[c] * tgt[c];
                double dx = r[p].x - r[q].x, dy = r[p].y - r[q].y, dz = r[p].z - r[q].z;
                min_image(a.box, dx, dy, dz);
                const double s = norm2_ref(dx, dy, dz) - d2;
                const double viol = fabs(s) / d2;
                worst = fmax(worst, viol);
                if (viol <= a.tol) continue;
                double ex = rr[p].x - rr[q].x, ey = rr[p].y - rr[q].y, ez = rr[p].z - rr[q].z;
                min_image(a.box, ex, ey, ez);
                const double inv = 1.0 / mass[p] + 1.0 / mass[q];
                const double den = 2.0 * inv * norm2_dot(dx, dy, dz, ex, ey, ez);
                if (fabs(den) < 1e-30) {
                    res = -2;
                    it = a.max_iter;
                    break;
                }
                const double gg = s / den;
                const double gp = gg / mass[p], gq = gg / mass[q];
                r[p].x = r[p].x - ex * gp; r[p].y = r[p].y - ey * gp; r[p].z = r[p].z - ez * gp;
                r[q].x = r[q].x + ex * gq; r[q].y = r[q].y + ey * gq; r[q].z = r[q].z + ez * gq;
            }
            if (res == -2) break;
            if (worst <= a.tol) {
                res = it + 1;
                break;
            }
        }
        a.pos[b] = r[0];
        a.pos[b + 1] = r[1];
        a.pos[b + 2] = r[2];
        status[m] = res;
        if (res < 0 && a.fail != nullptr) atomicMin(&a.fail[0], res);
    }
}

// RATTLE: remove the relative velocity along each bond (rigid_water.cpp:118-161).
__global__ void k_rattle(WaterArgs a, int* __restrict__ status) {
    GRID_STRIDE(m, a.n / 3) {
        const long b = 3 * m;
        const double4 r[3] = {a.pos[b], a.pos[b + 1], a.pos[b + 2]};
        double4 v[3] = {a.vel[b], a.vel[b + 1], a.vel[b + 2]};
        const int ci[3] = {0, 0, 1}, cj[3] = {1, 2, 2};
        const double tgt[3] = {a.oh, a.oh, a.hh};
        int res = -1;
        for (int it = 0; it < a.max_iter; ++it) {
            double worst = 0.0;
            for (int c = 0; c < 3; ++c) {
#pragma clang fp contract(off)
                const int p = ci[c], q = cj[c];
                const double d2 = tgt[c] * tgt[c];
                double dx = r[p].x - r[q].x, dy = r[p].y - r[q].y, dz = r[p].z - r[q].z;
                min_image(a.box, dx, dy, dz);
                const double vx = v[p].x - v[q].x, vy = v[p].y - v[q].y, vz = v[p].z - v[q].z;
                const double rv = norm2_dot(dx, dy, dz, vx, vy, vz);
                const double sc = fabs(rv) / d2;
                worst = fmax(worst, sc);
                if (sc <= a.tol) continue;
                const double inv = 1.0 / v[p].w + 1.0 / v[q].w;
                const double kk = rv / (inv * norm2_ref(dx, dy, dz));
                const double kp = kk / v[p].w, kq = kk / v[q].w;
                v[p].x = v[p].x - dx * kp; v[p].y = v[p].y - dy * kp; v[p].z = v[p].z - dz * kp;
                v[q].x = v[q].x + dx * kq; v[q].y = v[q].y + dy * kq; v[q].z = v[q].z + dz * kq;
            }
            if (worst <= a.tol) {
                res = it + 1;
                break;
            }
        }
        a.vel[b] = v[0];
        a.vel[b + 1] = v[1];
        a.vel[b + 2] = v[2];
        status[m] = res;
        if (res < 0 && a.fail != nullptr) atomicMin(&a.fail[1], res);
    }
}

// ---- SETTLE: closed-form rigid constraints, one thread per molecule ---------
// The fixed points of SHAKE and RATTLE above, solved exactly (Miyamoto &
// Kollman, J. Comput. Chem. 13, 952 (1992); the replacement the reference
// plans, constrain/rigid_water.hpp:20-24). Positions: in the frame of the
// reference molecule (ez its normal, ex along H1->H2), the constrained
// molecule is the canonical triangle O (0, ra), H1 (-rc, -rb), H2 (rc, -rb)
// at the unconstrained centre of mass, rotated by R = Rz(theta) M: M is the
// Rodrigues rotation taking ez to R's third row (s1, s2, s3), which keeps
// every site's out-of-plane coordinate (constraint forces lie along the
// reference bonds, so in the reference plane), and theta makes the in-plane
// angular momentum of the displacement zero (bond forces exert no torque):
// alpha sin + beta cos = T. Velocities: the 3 x 3 linear system that RATTLE
// iterates on, A mu = (r_l . v_l), solved by Cramer's rule. Arithmetic in
// the operation order of the CPU restatement, without contraction.
struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 d3(const double4& p) { return D3{p.x, p.y, p.z}; }
__device__ __forceinline__ D3 dsub(D3 a, D3 b) {
#pragma clang fp contract(off)
    return D3{a.x - b.x, a.y - b.y, a.z - b.z};
}
__device__ __forceinline__ D3 dadd(D3 a, D3 b) {
#pragma clang fp contract(off)
    return D3{a.x + b.x, a.y + b.y, a.z + b.z};
}
__device__ __forceinline__ D3 dscale(D3 a, double s) {
#pragma clang fp contract(off)
    return D3{a.x * s, a.y * s, a.z * s};
}
__device__ __forceinline__ D3 dcross(D3 a, D3 b) {
#pragma clang fp contract(off)
    return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double ddot(D3 a, D3 b) { return norm2_dot(a.x, a.y, a.z, b.x, b.y, b.z); }
__device__ __forceinline__ D3 dimage(const Box& b, D3 d) {
    min_image(b, d.x, d.y, d.z);
    return d;
}

// Displacements D[3] that take the unconstrained sites r[3] onto the rigid
// molecule, against the reference sites r0[3]; false when the step has no
// solution (a molecule turned out of its plane by 90 degrees or more, or a
// degenerate reference).
__device__ __forceinline__ bool settle_shift(const WaterArgs& a, const D3 r0[3], const D3 r[3],
                                             double mo, double mh, D3 D[3]) {
#pragma clang fp contract(off)
    const double rc = a.settle_rc, h = a.settle_h;
    const double mt = mo + 2.0 * mh;
    const double ra = 2.0 * mh * h / mt, rb = h - ra;
    const D3 b0 = dimage(a.box, dsub(r0[1], r0[0]));
    const D3 c0 = dimage(a.box, dsub(r0[2], r0[0]));
    const D3 nz = dcross(b0, c0);
    const double nn = ddot(nz, nz);
    const D3 ez = dscale(nz, 1.0 / sqrt(nn));
    const D3 hx = dsub(c0, b0);
    const D3 ex = dscale(hx, 1.0 / sqrt(ddot(hx, hx)));
    const D3 ey = dcross(ez, ex);
    const D3 b1 = dimage(a.box, dsub(r[1], r[0]));
    const D3 c1 = dimage(a.box, dsub(r[2], r[0]));
    const D3 d = dscale(dadd(dscale(b1, mh), dscale(c1, mh)), 1.0 / mt);
    const D3 u[3] = {dscale(d, -1.0), dsub(b1, d), dsub(c1, d)};
    double ux[3], uy[3], uz[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        ux[k] = ddot(u[k], ex);
        uy[k] = ddot(u[k], ey);
        uz[k] = ddot(u[k], ez);
    }
    const double s2 = uz[0] / ra, s1 = (uz[2] - uz[1]) / (2.0 * rc);
    const double q = 1.0 - s1 * s1 - s2 * s2;
    const double s3 = sqrt(q);
    const double k = 1.0 / (1.0 + s3);
    const double mxx = 1.0 - s1 * s1 * k, myy = 1.0 - s2 * s2 * k, mxy = -s1 * s2 * k;
    const double sxx = 2.0 * mh * rc * rc, syy = mo * ra * ra + 2.0 * mh * rb * rb;
    const double T = -mo * ra * ux[0] + mh * rc * (uy[2] - uy[1]) + mh * rb * (ux[1] + ux[2]);
    const double al = sxx * mxx + syy * myy, be = (sxx - syy) * mxy;
    const double rho2 = al * al + be * be;
    const double disc = rho2 - T * T;
    const double Q = sqrt(disc);
    const double st = (T * al - Q * be) / rho2, ct = (Q * al + T * be) / rho2;
    const double r1x = ct * mxx - st * mxy, r1y = ct * mxy - st * myy;
    const double r2x = st * mxx + ct * mxy, r2y = st * mxy + ct * myy;
    const double px[3] = {0.0, -rc, rc}, py[3] = {ra, -rb, -rb};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double nx = r1x * px[i] + r1y * py[i];
        const double ny = r2x * px[i] + r2y * py[i];
        const double nzc = s1 * px[i] + s2 * py[i];
        const double dx = nx - ux[i], dy = ny - uy[i], dz = nzc - uz[i];
        D[i] = dadd(dadd(dscale(ex, dx), dscale(ey, dy)), dscale(ez, dz));
    }
    return nn > 0.0 && q > 0.0 && disc >= 0.0;
}

// Velocities v[3] (in/out) made rigid at the sites r[3]; false when the
// 3 x 3 system is singular.
__device__ __forceinline__ bool settle_velocities(const WaterArgs& a, const D3 r[3], D3 v[3],
                                                  double mo, double mh) {
#pragma clang fp contract(off)
    const D3 r01 = dimage(a.box, dsub(r[0], r[1]));
    const D3 r02 = dimage(a.box, dsub(r[0], r[2]));
    const D3 r12 = dimage(a.box, dsub(r[1], r[2]));
    const double ioh = 1.0 / mo + 1.0 / mh;
    const double a00 = ioh * ddot(r01, r01), a11 = ioh * ddot(r02, r02),
                 a22 = (2.0 / mh) * ddot(r12, r12);
    const double a01 = ddot(r01, r02) / mo, a02 = -ddot(r01, r12) / mh, a12 = ddot(r02, r12) / mh;
    const double b0 = ddot(r01, dsub(v[0], v[1])), b1 = ddot(r02, dsub(v[0], v[2])),
                 b2 = ddot(r12, dsub(v[1], v[2]));
    const double c00 = a11 * a22 - a12 * a12, c01 = a02 * a12 - a01 * a22, c02 = a01 * a12 - a02 * a11;
    const double c11 = a00 * a22 - a02 * a02, c12 = a01 * a02 - a00 * a12, c22 = a00 * a11 - a01 * a01;
    const double det = a00 * c00 + a01 * c01 + a02 * c02;
    const double inv = 1.0 / det;
    const double mu0 = (c00 * b0 + c01 * b1 + c02 * b2) * inv;
    const double mu1 = (c01 * b0 + c11 * b1 + c12 * b2) * inv;
    const double mu2 = (c02 * b0 + c12 * b1 + c22 * b2) * inv;
    const D3 g01 = dscale(r01, mu0), g02 = dscale(r02, mu1), g12 = dscale(r12, mu2);
    v[0] = dsub(v[0], dscale(dadd(g01, g02), 1.0 / mo));
    v[1] = dadd(v[1], dscale(dsub(g01, g12), 1.0 / mh));
    v[2] = dadd(v[2], dscale(dadd(g02, g12), 1.0 / mh));
    return fabs(det) > 0.0;
}

__device__ __forceinline__ void settle_fail(const WaterArgs& a, int* status, long m, bool ok, int slot) {
    status[m] = ok ? 1 : -2;
    if (!ok && a.fail != nullptr) atomicMin(&a.fail[slot], -2);
}

// RigidWater::constrainPositions by SETTLE: positions against refp.
__global__ void k_settle(WaterArgs a, const double4* __restrict__ refp, int* __restrict__ status) {
    GRID_STRIDE(m, a.n / 3) {
        const long b = 3 * m;
        const D3 r0[3] = {d3(refp[b]), d3(refp[b + 1]), d3(refp[b + 2])};
        const double4 p[3] = {a.pos[b], a.pos[b + 1], a.pos[b + 2]};
        const D3 r[3] = {d3(p[0]), d3(p[1]), d3(p[2])};
        D3 D[3];
        const bool ok = settle_shift(a, r0, r, a.vel[b].w, a.vel[b + 1].w, D);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const D3 q = dadd(r[i], D[i]);
            a.pos[b + i] = make_double4(q.x, q.y, q.z, p[i].w);
        }
        settle_fail(a, status, m, ok, 0);
    }
}

// RigidWater::constrainVelocities by SETTLE.
__global__ void k_settle_vel(WaterArgs a, int* __restrict__ status) {
    GRID_STRIDE(m, a.n / 3) {
        const long b = 3 * m;
        const D3 r[3] = {d3(a.pos[b]), d3(a.pos[b + 1]), d3(a.pos[b + 2])};
        const double4 w[3] = {a.vel[b], a.vel[b + 1], a.vel[b + 2]};
        D3 v[3] = {d3(w[0]), d3(w[1]), d3(w[2])};
        const bool ok = settle_velocities(a, r, v, w[0].w, w[1].w);
#pragma unroll
        for (int i = 0; i < 3; ++i) a.vel[b + i] = make_double4(v[i].x, v[i].y, v[i].z, w[i].w);
        settle_fail(a, status, m, ok, 1);
    }
}

// ConstrainedVelocityVerlet::step up to the forces, per molecule with
// SETTLE: kick v + f (dt/2/m), drift without wrap, the closed-form position
// constraint against the start-of-step sites, and v += minimumImage(r -
// r_free) / dt (constrained_velocity_verlet.hpp:33-47), in the arithmetic of
// k_cvv_first, k_shake's replacement k_settle and k_cvv_correct.
__global__ void k_cvv_settle_first(WaterArgs a, double dt, int* __restrict__ status) {
    GRID_STRIDE(m, a.n / 3) {
#pragma clang fp contract(off)
        const long b = 3 * m;
        D3 r0[3], r[3], v[3];
        double w[3], mass[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double4 p = a.pos[b + i], u = a.vel[b + i], f = a.frc[b + i];
            r0[i] = d3(p);
            w[i] = p.w;
            mass[i] = u.w;
            const double c = 0.5 * dt / u.w;
            v[i] = D3{u.x + f.x * c, u.y + f.y * c, u.z + f.z * c};
            r[i] = D3{p.x + v[i].x * dt, p.y + v[i].y * dt, p.z + v[i].z * dt};
        }
        D3 D[3];
        const bool ok = settle_shift(a, r0, r, mass[0], mass[1], D);
        const double inv = 1.0 / dt;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const D3 q = dadd(r[i], D[i]);
            const D3 sh = dimage(a.box, dsub(q, r[i]));
            a.pos[b + i] = make_double4(q.x, q.y, q.z, w[i]);
            a.vel[b + i] = make_double4(v[i].x + sh.x * inv, v[i].y + sh.y * inv, v[i].z + sh.z * inv,
                                        mass[i]);
        }
        settle_fail(a, status, m, ok, 0);
    }
}

// The step's end with SETTLE: kick v + f (dt/2/m), then the velocity
// constraint (constrained_velocity_verlet.hpp:51-55).
__global__ void k_cvv_settle_second(WaterArgs a, double dt, int* __restrict__ status) {
    GRID_STRIDE(m, a.n / 3) {
#pragma clang fp contract(off)
        const long b = 3 * m;
        D3 r[3], v[3];
        double mass[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double4 u = a.vel[b + i], f = a.frc[b + i];
            r[i] = d3(a.pos[b + i]);
            mass[i] = u.w;
            const double c = 0.5 * dt / u.w;
            v[i] = D3{u.x + f.x * c, u.y + f.y * c, u.z + f.z * c};
        }
        const bool ok = settle_velocities(a, r, v, mass[0], mass[1]);
#pragma unroll
        for (int i = 0; i < 3; ++i) a.vel[b + i] = make_double4(v[i].x, v[i].y, v[i].z, mass[i]);
        settle_fail(a, status, m, ok, 1);
    }
}

// Constrained velocity Verlet, first half: save the start positions, kick
// v + f (dt/2/m) (plain multiply and add), drift without wrap, keep the
// unconstrained positions (constrained_velocity_verlet.hpp:33-42).
__global__ void k_cvv_first(WaterArgs a, double dt, double4* __restrict__ start,
                            double4* __restrict__ free_pos) {
    GRID_STRIDE(i, a.n) {
#pragma clang fp contract(off)
        const double4 p = a.pos[i];
        double4 v = a.vel[i];
        const double4 f = a.frc[i];
        start[i] = p;
        const double c = 0.5 * dt / v.w;
        v.x = v.x + f.x * c;
        v.y = v.y + f.y * c;
        v.z = v.z + f.z * c;
        const double4 np = make_double4(p.x + v.x * dt, p.y + v.y * dt, p.z + v.z * dt, p.w);
        a.pos[i] = np;
        a.vel[i] = v;
        free_pos[i] = np;
    }
}

// v += minimumImage(r_constrained - r_free) / dt (constrained_velocity_verlet.hpp:44-47).
__global__ void k_cvv_correct(WaterArgs a, double dt, const double4* __restrict__ free_pos) {
    GRID_STRIDE(i, a.n) {
#pragma clang fp contract(off)
        const double4 p = a.pos[i], u = free_pos[i];
        double dx = p.x - u.x, dy = p.y - u.y, dz = p.z - u.z;
        min_image(a.box, dx, dy, dz);
        const double inv = 1.0 / dt;
        double4 v = a.vel[i];
        v.x = v.x + dx * inv;
        v.y = v.y + dy * inv;
        v.z = v.z + dz * inv;
        a.vel[i] = v;
    }
}

__global__ void k_cvv_kick(WaterArgs a, double dt) {
    GRID_STRIDE(i, a.n) {
#pragma clang fp contract(off)
        double4 v = a.vel[i];
        const double4 f = a.frc[i];
        const double c = 0.5 * dt / v.w;
        v.x = v.x + f.x * c;
        v.y = v.y + f.y * c;
        v.z = v.z + f.z * c;
        a.vel[i] = v;
    }
}

// Top-2 displacement from the list's reference positions (one workgroup).
__global__ __launch_bounds__(1024) void k_wcheck(WaterArgs a, int* flag) {
    __shared__ double s1[1024], s2[1024];
    double t1 = 0.0, t2 = 0.0;
    for (long i = threadIdx.x; i < a.n; i += 1024) {
        const double4 p = a.pos[i], r = a.ref[i];
        double dx = p.x - r.x, dy = p.y - r.y, dz = p.z - r.z;
        min_image(a.box, dx, dy, dz);
        top2_merge(t1, t2, sqrt(norm2_ref(dx, dy, dz)), 0.0);
    }
    s1[threadIdx.x] = t1;
    s2[threadIdx.x] = t2;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            double x1 = s1[threadIdx.x], x2 = s2[threadIdx.x];
            top2_merge(x1, x2, s1[threadIdx.x + off], s2[threadIdx.x + off]);
            s1[threadIdx.x] = x1;
            s2[threadIdx.x] = x2;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *flag = (s1[0] + s2[0]) > a.skin ? 1 : 0;
}

__global__ __launch_bounds__(kBlock) void k_wkinetic(WaterArgs a) {
    __shared__ double sm[kBlock / kWave];
    const long i = static_cast<long>(blockIdx.x) * kBlock + threadIdx.x;
    double ke = 0.0;
    if (i < a.n) {
        const double4 v = a.vel[i];
        ke = (0.5 * v.w) * norm2_ref(v.x, v.y, v.z);
    }
    const double b = block_reduce(ke, sm, wave_sum);
    if (threadIdx.x == 0) a.part2[blockIdx.x] = b;
}

// In-order sums of the energy partials (one workgroup of kBlock threads):
// thread t takes partials t, t + kBlock, ...; then the threads in index order.
__global__ __launch_bounds__(kBlock) void k_wenergy_sum(WaterArgs a, int nbr, int nbk, double* out) {
    __shared__ double sm[kBlock / kWave];
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nbr; b += kBlock)
        for (int q = 0; q < 4; ++q) v[q] += a.part[4 * b + q];
    for (int b = threadIdx.x; b < nbk; b += kBlock) v[4] += a.part2[b];
    for (int q = 0; q < 5; ++q) {
        const double r = block_reduce(v[q], sm, wave_sum);
        if (threadIdx.x == 0) out[q] = r;
    }
}

__global__ __launch_bounds__(kBlock) void k_wkinetic_sum(WaterArgs a, int nb, double* out) {
    __shared__ double sm[kBlock / kWave];
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += kBlock) v += a.part2[b];
    const double r = block_reduce(v, sm, wave_sum);
    if (threadIdx.x == 0) *out = r;
}

inline unsigned gridn(long n) {
    const long g = (n + kBlock - 1) / kBlock;
    return static_cast<unsigned>(g > 0 ? g : 1);
}

}  // namespace

void water_launch_list(const WaterArgs& a, const Grid& g, const int* cell_of, const int* cell_start,
                       const int* cell_atoms, hipStream_t s) {
    // one wave per site
    hipLaunchKernelGGL(k_wlist, dim3(gridn(a.n * kWave)), dim3(kBlock), 0, s, a, g, cell_of,
                       cell_start, cell_atoms);
}

long water_real_blocks(long n) { return static_cast<long>(gridn(n * kWaterTpa)); }

int water_recip_chunks(long n, int ncol) {
    // enough (site tile, column chunk) blocks to fill the chip
    const long tiles = gridn(n);
    long c = (2048 + tiles - 1) / tiles;
    if (c > ncol) c = ncol;
    return static_cast<int>(c < 1 ? 1 : c);
}

void water_launch_forces(const WaterArgs& a, bool mixed, hipStream_t s) {
    const dim3 gr(static_cast<unsigned>(water_real_blocks(a.n)));
    if (mixed)
        hipLaunchKernelGGL(k_wreal<true>, gr, dim3(kBlock), 0, s, a);
    else
        hipLaunchKernelGGL(k_wreal<false>, gr, dim3(kBlock), 0, s, a);
    if (a.pme.on) {
        water_launch_pme(a, s);
        return;
    }
    hipLaunchKernelGGL(k_weik, dim3(gridn(a.n)), dim3(kBlock), 0, s, a);
    const int groups = water_site_groups(a.n);
    const int spg = static_cast<int>((a.n + groups - 1) / groups);
    hipLaunchKernelGGL(k_sk, dim3((a.nseg * kSkSub + kBlock - 1) / kBlock, groups), dim3(kBlock), 0, s, a,
                       spg);
    hipLaunchKernelGGL(k_sk_reduce, dim3((a.nk + kBlock - 1) / kBlock), dim3(kBlock), 0, s, a, groups);
    hipLaunchKernelGGL(k_recip, dim3(gridn(a.n), a.nchunk), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_recip_reduce, dim3(gridn(a.n)), dim3(kBlock), 0, s, a);
}

// One thread per molecule in single-wave workgroups, so a few thousand
// molecules spread over as many compute units as there are waves.
inline unsigned wave_groups(long m) { return static_cast<unsigned>(m > 0 ? (m + kWave - 1) / kWave : 1); }

void water_launch_shake(const WaterArgs& a, const double4* refp, int* status, hipStream_t s) {
    hipLaunchKernelGGL(k_shake, dim3(wave_groups(a.n / 3)), dim3(kWave), 0, s, a, refp, status);
}

void water_launch_rattle(const WaterArgs& a, int* status, hipStream_t s) {
    hipLaunchKernelGGL(k_rattle, dim3(wave_groups(a.n / 3)), dim3(kWave), 0, s, a, status);
}

void water_launch_settle(const WaterArgs& a, const double4* refp, int* status, hipStream_t s) {
    hipLaunchKernelGGL(k_settle, dim3(wave_groups(a.n / 3)), dim3(kWave), 0, s, a, refp, status);
}

void water_launch_settle_vel(const WaterArgs& a, int* status, hipStream_t s) {
    hipLaunchKernelGGL(k_settle_vel, dim3(wave_groups(a.n / 3)), dim3(kWave), 0, s, a, status);
}

void water_launch_cvv_settle_first(const WaterArgs& a, double dt, int* status, hipStream_t s) {
    hipLaunchKernelGGL(k_cvv_settle_first, dim3(wave_groups(a.n / 3)), dim3(kWave), 0, s, a, dt, status);
}

void water_launch_cvv_settle_second(const WaterArgs& a, double dt, int* status, hipStream_t s) {
    hipLaunchKernelGGL(k_cvv_settle_second, dim3(wave_groups(a.n / 3)), dim3(kWave), 0, s, a, dt, status);
}

void water_launch_cvv_first(const WaterArgs& a, double dt, double4* start, double4* free_pos,
                            hipStream_t s) {
    hipLaunchKernelGGL(k_cvv_first, dim3(gridn(a.n)), dim3(kBlock), 0, s, a, dt, start, free_pos);
}

void water_launch_cvv_correct(const WaterArgs& a, double dt, const double4* free_pos, hipStream_t s) {
    hipLaunchKernelGGL(k_cvv_correct, dim3(gridn(a.n)), dim3(kBlock), 0, s, a, dt, free_pos);
}

void water_launch_cvv_kick(const WaterArgs& a, double dt, hipStream_t s) {
    hipLaunchKernelGGL(k_cvv_kick, dim3(gridn(a.n)), dim3(kBlock), 0, s, a, dt);
}

void water_launch_check(const WaterArgs& a, int* flag, hipStream_t s) {
    hipLaunchKernelGGL(k_wcheck, dim3(1), dim3(1024), 0, s, a, flag);
}

void water_launch_energy_sum(const WaterArgs& a, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_wenergy_sum, dim3(1), dim3(kBlock), 0, s, a,
                       static_cast<int>(water_real_blocks(a.n)), a.nkpart, out);
}

void water_launch_kinetic_sum(const WaterArgs& a, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_wkinetic_sum, dim3(1), dim3(kBlock), 0, s, a, static_cast<int>(gridn(a.n)), out);
}

void water_launch_kinetic(const WaterArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_wkinetic, dim3(gridn(a.n)), dim3(kBlock), 0, s, a);
}

int water_site_groups(long n) {
    const long g = (n + 511) / 512;
    return static_cast<int>(g < 1 ? 1 : (g > 64 ? 64 : g));
}

}  // namespace mdk
