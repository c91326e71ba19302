// Host-side launch interfaces of the device kernels (one struct of
// arguments per kernel family).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "md_common.hpp"

struct md_ctx;

namespace mdk {

// Cell grid of the list build: nc = max(1, floor(L/cutoff)) per axis
// (neighbor/cell_list.cpp:8-14), width w = L/nc; usable when nc >= 3 on
// every axis (cell_list.cpp:44-46).
struct Grid {
    int nc[3];
    double w[3];
    double L[3];
    double inv_w[3];  // nc / L
    int usable;
};

// Periodic fixed-point positions. Each coordinate is stored as a uint32 that
// maps the box length onto 2^32 (X = round((x / L + 1/2) 2^32) mod 2^32),
// and the fourth word is the atom's upload index. The difference of two
// coordinates is then one wrapping 32-bit subtraction, read as int32: the
// minimum image falls out of two's-complement arithmetic with no fold, and
// d = dX (L / 2^32) is formed in FP64. The error of d is below one unit
// L / 2^32 per axis (2.5e-8 at L = 107), and one 16-byte load carries a
// partner's position and orientation key.
struct Quant {
    double inv[3];   // 2^32 / L
    double unit[3];  // L / 2^32
    double band_l;   // relative half-width of the decision band around rl^2
    double band_c;   // the same around rc^2
    int cubic;       // one unit on every axis: r2 may be taken in units squared
};

__device__ __forceinline__ uint32_t quant_axis(double x, double inv) {
    // the int64 conversion keeps values outside the box; the uint32 cast then
    // takes them modulo 2^32, which is the periodic wrap
    return static_cast<uint32_t>(static_cast<long long>(rint(fma(x, inv, 2147483648.0))));
}

__device__ __forceinline__ int4 quantize(const Quant& q, const double4& p, int orig) {
    return make_int4(static_cast<int>(quant_axis(p.x, q.inv[0])),
                     static_cast<int>(quant_axis(p.y, q.inv[1])),
                     static_cast<int>(quant_axis(p.z, q.inv[2])), orig);
}

// d = r_a - r_b (minimum image) from fixed-point positions, FP64.
__device__ __forceinline__ void qdelta(const Quant& q, const int4& a, const int4& b, double& dx,
                                       double& dy, double& dz) {
    dx = static_cast<double>(static_cast<int>(static_cast<uint32_t>(a.x) - static_cast<uint32_t>(b.x))) * q.unit[0];
    dy = static_cast<double>(static_cast<int>(static_cast<uint32_t>(a.y) - static_cast<uint32_t>(b.y))) * q.unit[1];
    dz = static_cast<double>(static_cast<int>(static_cast<uint32_t>(a.z) - static_cast<uint32_t>(b.z))) * q.unit[2];
}

// Bricks of list cells. The cell grid (nc per axis) is cut into kb bricks
// per axis, brick i spanning cells [floor(i nc / kb), floor((i+1) nc / kb)).
// A brick's halo is the box of cells within two cells of it (every list
// partner of an atom lies in the 5 x 5 x 5 cells around its own), taken in
// the order z, y, x (x fastest). When bricks are on, one pair-kernel
// workgroup stages its brick's halo into LDS in that order, and a list
// entry is the partner's index in its row's brick halo instead of a slot:
// tab[tab_off[B] + h] is the number of halo atoms before halo cell h of brick
// B (an exclusive prefix, with the halo total at index H_B), rebuilt with
// each list build.
struct Bricks {
    int on;
    int kb[3];
    int nb;             // bricks (kb[0] kb[1] kb[2])
    int halo_cap;       // halo atoms one workgroup stages (LDS rows)
    int waves_mixed;    // waves per workgroup of the mixed brick kernel
    int waves_fp64;     // waves per workgroup of the FP64 brick kernel
    const int* tab_off; // nb + 1 entries
    int* tab;           // tab_off[nb] entries
};

// First cell of brick i along an axis of n cells cut into kb bricks, and the
// brick of cell c (the i with lo(i) <= c < lo(i + 1)).
__host__ __device__ __forceinline__ int brick_lo(int i, int n, int kb) {
    return static_cast<int>(static_cast<long>(i) * n / kb);
}
__host__ __device__ __forceinline__ int brick_of(int c, int n, int kb) {
    return static_cast<int>((static_cast<long>(c + 1) * kb - 1) / n);
}

// Origin (first cell) and extent in cells of brick b.
struct BrickGeo {
    int x0, y0, z0, ex, ey, ez;
};
__host__ __device__ __forceinline__ BrickGeo brick_geo(const Bricks& k, const Grid& g, int b) {
    const int bx = b % k.kb[0], by = (b / k.kb[0]) % k.kb[1], bz = b / (k.kb[0] * k.kb[1]);
    BrickGeo o;
    o.x0 = brick_lo(bx, g.nc[0], k.kb[0]);
    o.y0 = brick_lo(by, g.nc[1], k.kb[1]);
    o.z0 = brick_lo(bz, g.nc[2], k.kb[2]);
    o.ex = brick_lo(bx + 1, g.nc[0], k.kb[0]) - o.x0;
    o.ey = brick_lo(by + 1, g.nc[1], k.kb[1]) - o.y0;
    o.ez = brick_lo(bz + 1, g.nc[2], k.kb[2]) - o.z0;
    return o;
}

struct NeighArgs {
    long n;
    Grid grid;
    Bricks bk;
    Box box;
    double rl2;  // (rc + skin)^2 as the reference computes it
    double4 *pos, *vel, *frc, *ref;
    int *orig, *slot;
    double4 *pos2, *vel2, *frc2;
    int *orig2, *cell2;
    int *cell_of, *cell_count, *cell_start, *cell_fill, *order;
    int* tile_sum;  // scan_tiles(cell count) entries
    // fixed-point positions (or null): the committed order / the permuted
    // order that the candidate scan reads
    int4 *qpos, *qpos2;
    Quant quant;
    uint32_t* nbr;
    int* ncnt;
    long npad;
    int cap;
    Scalars* sc;
    const int* flag;
};
void launch_rebuild(const NeighArgs& a, hipStream_t s);
// Number of scan tiles (entries of NeighArgs::tile_sum) for ncell cells.
long scan_tiles(long ncell);

// LJ pair kernel over the full Verlet list.
struct LJConst {
    double sig2, eps, rc, rc2, u_rc, dudr_rc;
    float sig2f, epsf, rc2f, rcf, u_rcf, dudr_rcf;
    // FP32 coefficients times the fixed-point scale 2^32 (mixed kernel):
    // 4 eps, 24 eps, u(rc), u'(rc)
    float e4s, e24s, u_rcs, dudr_rcs;
    int lfs;
};

// Constants of the mixed pair terms in one length unit s: lengths in the
// terms are lengths / s. On a cubic box with fixed-point positions s is the
// fixed-point unit u, so an integer coordinate difference is the scaled
// difference itself; otherwise s = 1. Host-computed (mix_constants).
struct MixK {
    float sig2s;     // sigma^2 / s^2
    float krs;       // LinearForceShift: u'(rc) s / (24 eps) (0 when eps = 0)
    double lo, hi;   // fixed-point cutoff band: rc^2 (1 -+ band) / s^2
    double inv_s;    // 1 / s
    double fscale;   // 24 eps / s: the row force is fscale * sum c d
    double e4, e24;  // 4 eps, 24 eps
    double ushift;   // LinearForceShift: u(rc) - rc u'(rc)
    double urs;      // LinearForceShift: u'(rc) s
};

struct ForceArgs {
    long n;
    Box box;
    LJConst lj;
    MixK mq;       // mixed terms from the fixed-point positions (s = u on a cubic box, else 1)
    MixK mx;       // mixed terms from the FP64 positions (s = 1)
    const double4* pos;
    double4* vel;  // kicked when kick != 0
    double4* frc;
    const int* orig;
    const uint32_t* nbr;
    const int* ncnt;
    long npad;
    int cap;         // row capacity: rows longer than this overflowed and are clamped
    const int4* qpos;  // fixed-point positions for the mixed kernel (or null)
    Quant quant;
    // bricks (bk.on: list entries are brick-halo indices, the brick kernels run)
    Grid grid;
    Bricks bk;
    const int* cell_start;
    double half_dt;  // kick coefficient h (v += h/m f) when kick != 0
    int kick;
    int mixed;
    // reductions
    double* part_d;     // per-workgroup FP64 partials, 3 per group (E, W, KE)
    long long* part_l;  // per-workgroup int64 partials, 3 per group (E, W, pairs)
    // the finish kernel's own partials (3 per finish workgroup each) and its
    // arrival ticket (0 between launches)
    double* fin_d;
    long long* fin_l;
    unsigned int* fin_count;
    Scalars* sc;
    Chain* chain;       // NHC: chain update on one thread of the finish (null otherwise)
    double chain_delta; // dt/2 of the chain half-steps
    int chain_twice;    // 1: also propagate the next step's first half-chain
    void* thermo;       // md_thermo* record for this step, or null
    // graph-replayed segments (dev_step 1): the step's record is
    // thermo_base[sc->step - step0] (when thermo_base is not null), and the
    // finish advances sc->step
    int dev_step;
    void* thermo_base;
    unsigned long long step0;
};
// Per-workgroup partial slots that launch_lj_force writes for n atoms (small
// systems use several lanes per atom, hence more workgroups).
long lj_partials(long n);
void launch_lj_force(const ForceArgs& a, hipStream_t s);
// Workgroups of the finish kernel (each owns 3 + 3 fin partials).
constexpr int kFinishGroups = 64;
// Fixed-order reduction of the pair kernel's partials (E, W, KE, pairs),
// thermo row and NHC chain; launched right after launch_lj_force.
void launch_lj_finish(const ForceArgs& a, hipStream_t s);
// Rewrites a brick list (a.bk.on) as slot entries (entry_of) into out, in
// the same octet layout (md_neigh_export_pairs).
void launch_brick_export(const ForceArgs& a, uint32_t* out, hipStream_t s);
// LennardJones(sigma, epsilon, cutoff, truncation) constants: u(rc) and
// du/dr(rc) as lennard_jones.cpp:32-48 computes them, and their FP32 and
// fixed-point-scaled forms.
LJConst lj_constants(double sigma, double epsilon, double cutoff, int linear_force_shift);
// The CellList grid at `cutoff` (cell_list.cpp:8-14): floor(L_d / cutoff)
// cells per axis (at least 1), usable with 3 or more on every axis.
Grid cell_grid(const double L[3], double cutoff);
// MixK for length unit s and relative cutoff band `band` (0: exact decisions).
MixK mix_constants(const LJConst& k, double s, double band);
// One chain propagation of length delta from sc->kinetic; the velocity
// rescale is left pending in sc->vscale.
void launch_chain_first(Chain* chain, Scalars* sc, double delta, hipStream_t s);

// Brute-force (all pairs) and cell-list (at the LJ cutoff, through an index
// array) force / energy evaluations for the standalone API calls.
struct DirectArgs {
    long n;
    Box box;
    LJConst lj;
    const double4* pos;
    double4* frc;  // null: energy/virial only
    // or, when fx_out is set, forces written as three arrays in upload order
    double* fx_out;
    double* fy_out;
    double* fz_out;
    const int* orig;
    // cells (null cell_atoms: brute force)
    Grid grid;
    const int* cell_of;
    const int* cell_start;
    const int* cell_atoms;
    double* part_d;
    long long* part_l;
    Scalars* sc;
};
void launch_lj_direct(const DirectArgs& a, hipStream_t s);
// A context's device system, for the analysis entry points (positions and
// velocities in device slot order; slot[upload index] = slot).
struct CtxView {
    long n;
    Box box;
    double L[3];
    const double4* pos;
    const double4* vel;
    const int* orig;
    const int* slot;
    hipStream_t stream;
    int device;
};
// False for a null context or one without a system.
bool ctx_view(md_ctx* c, CtxView* v);
// The calling thread's last md_lj_forces_noalloc failure, or null.
const char* noalloc_error();
// Bins atoms into cells without permuting them: cell_atoms holds slot
// indices grouped by cell (ordered by slot within a cell).
// flag (or null): a device rebuild decision; 0 skips the build.
void launch_bin_index(long n, const double4* pos, const Grid& g, int* cell_of, int* cell_count,
                      int* cell_start, int* cell_fill, int* cell_tmp, int* cell_atoms,
                      const int* flag, hipStream_t s);

// Integrator per-atom passes before the force call, with the displacement
// check folded in.
struct StepArgs {
    long n;
    Box box;
    double4* pos;
    double4* vel;
    const double4* frc;
    const double4* ref;
    const int* orig;
    int kind;         // md_integrator_kind
    double dt;
    double c1, c2, temperature;  // BAOAB
    uint64_t seed;
    uint64_t step;    // global step counter (BAOAB noise counter)
    int dev_step;     // 1: the counter is sc->step instead (graph-replayed segments)
    double skin;
    int check;        // compute the displacement check and set the rebuild flag
    double* part_d;   // 2 per group (top1, top2)
    Scalars* sc;
    int4* qpos;       // refreshed from the new positions when not null
    Quant quant;
};
void launch_pre_step(const StepArgs& a, hipStream_t s);
// Displacement check alone (md_neigh_needs_rebuild).
void launch_check(const StepArgs& a, hipStream_t s);

// Applies sc->vscale to every velocity and resets it to 1.
void launch_apply_scale(long n, double4* vel, Scalars* sc, hipStream_t s);
// KE reduction into sc->kinetic.
void launch_kinetic(long n, const double4* vel, double* part_d, Scalars* sc, hipStream_t s);
// Momentum reduction and centre-of-mass removal (System::removeCenterOfMassMotion).
void launch_remove_com(long n, double4* vel, double* part_d, double* out4, hipStream_t s);
// Upload-order scatter for download: out[orig[t]] = component.
void launch_unpermute(long n, const int* orig, const double4* a, double* x, double* y, double* z,
                      hipStream_t s);
// Upload-order gather for set_positions / set_velocities / set_forces; the
// w component is zeroed, kept (velocities: the mass) or set to the upload
// index (positions).
constexpr int kPermuteWZero = 0, kPermuteWKeep = 1, kPermuteWIndex = 2;
void launch_permute_in(long n, const int* orig, const double* x, const double* y, const double* z,
                       double4* a, int w_mode, hipStream_t s);
void launch_set_flag(int* flag, int v, hipStream_t s);
// Single integrator passes for host-driven steps.
void launch_half_kick(long n, double4* vel, const double4* frc, double half_dt, hipStream_t s);
void launch_drift(long n, const Box& box, double4* pos, const double4* vel, double h, hipStream_t s);
void launch_langevin_o(long n, double4* vel, const int* orig, double c1, double c2, double T,
                       uint64_t seed, uint64_t counter, hipStream_t s);
void launch_scale(long n, double4* vel, double sfac, hipStream_t s);
// qpos[t] = quantize(pos[t], orig[t]) (before a mixed force call whose
// positions were not written by launch_pre_step).
void launch_quantize(long n, const Quant& q, const double4* pos, const int* orig, int4* qpos,
                     hipStream_t s);

}  // namespace mdk
