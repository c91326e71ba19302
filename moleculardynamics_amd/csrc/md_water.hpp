// Launch interface of the SPC/E water kernels (md_water.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "md_common.hpp"
#include "md_kernels.hpp"

namespace mdk {

constexpr int kMaxK = 24;  // largest supported kmax per axis
constexpr int kPmeMaxOrder = 10;  // largest B-spline order of the PME path

// Smooth particle-mesh Ewald of the reciprocal term (md_pme.hip): grid
// K[0] x K[1] x K[2] (x fastest), B-spline order p, the reference's k-set
// restricted to |n_d| <= kmax, nb = 2 kmax + 1 frequencies per axis.
struct PmeArgs {
    int on;
    int K[3];
    int p;
    int kmax, nb;
    long ng;               // K[0] K[1] K[2]
    double* q;             // [ng] the spread charges
    double* W;             // [n][3][p] spline weights M_p(w + j) per site and axis
    double* DW;            // [n][3][p] their derivatives
    int* base;             // [n][3] grid index floor(u) per site and axis
    double2* A;            // [K2][K1][nb] after the x transform
    double2* B;            // [K2][nb][nb] after the y transform
    double2* P;            // [nb][nb][nb] G conj(F[Q])
    double2* Cz;           // [K2][nb][nb]
    double2* Dy;           // [K2][K1][nb]
    double* phi;           // [ng] dE/dQ
    const double2* tw[3];  // [nb][K_d]: e^{2 pi i (c - kmax) k / K_d}
    const double2* twT[3]; // the same, transposed: [K_d][nb]
    const double* G;       // [nb][nb][nb] influence function (0 off the k-set)
};

struct WaterArgs {
    long n;  // sites (3 per molecule, O H H)
    Box box;
    double4* pos;  // x, y, z, charge
    double4* vel;  // vx, vy, vz, mass
    double4* frc;
    double4* ref;  // positions at the last list build
    uint32_t* nbr;
    int* ncnt;
    int cap;
    double rl2, rc2, skin;
    double alpha, c2a, coulomb;  // c2a = 2 alpha / sqrt(pi)
    double sig2, eps;            // oxygen LJ
    double* part;                // 4 per workgroup: real, lj, intra, virial
    double* part2;               // 1 per workgroup: k-space energy / kinetic energy
    int kmax, nk;                // half-space k-vector count
    const int3* kvec;
    const double* kfac;          // C (2 pi / V) e^{-k^2/4a^2} / k^2
    double k0[3];
    double2* eik;                // [3][kmax+1][n]
    double2* skpart;             // [groups][nk]
    double2* sk;                 // [nk]
    // The half-space k-vectors as columns of consecutive nz: column c holds
    // (nx, ny, first nz, count), its k-vectors are kcol0[c] .. kcol0[c] + count - 1.
    const int4* kcol;
    const int* kcol0;
    int ncol, nchunk;            // columns; column chunks of the reciprocal force kernel
    // the columns cut into segments of at most 8 k-vectors: (nx, ny, first nz,
    // count) and the index of the segment's first k-vector
    const int4* kseg;
    const int* kseg0;
    int nseg;
    double2* skf;                // [nk] 4 factor(k) S(k)
    double4* fpart;              // [nchunk][n] reciprocal force partials
    // device-driven runs: SHAKE / RATTLE failure codes (atomic minimum of the
    // per-molecule status, 0 = none) and the rebuild flag that gates the list
    // kernels (null: always build)
    int* fail;
    const int* build_flag;
    int* overflow;
    double oh, hh, tol;
    int max_iter;
    // SETTLE geometry: half the H-H distance and the distance from O to the
    // H-H midpoint, sqrt(oh^2 - settle_rc^2)
    double settle_rc, settle_h;
    // reciprocal term by PME (pme.on) instead of the k-vector sums; nkpart is
    // the number of k-space energy partials in part2
    PmeArgs pme;
    int nkpart;
};

void water_launch_list(const WaterArgs& a, const Grid& g, const int* cell_of, const int* cell_start,
                       const int* cell_atoms, hipStream_t s);
// Column chunks of the reciprocal force kernel for n sites and ncol columns.
int water_recip_chunks(long n, int ncol);
// Workgroups of the real-space kernel (each writes 4 partials to part).
long water_real_blocks(long n);
void water_launch_forces(const WaterArgs& a, bool mixed, hipStream_t s);
void water_launch_shake(const WaterArgs& a, const double4* refp, int* status, hipStream_t s);
void water_launch_rattle(const WaterArgs& a, int* status, hipStream_t s);
// SETTLE in place of SHAKE / RATTLE: status[m] = 1, or -2 when the molecule's
// step has no solution (atomic minimum into fail[0] / fail[1] when set).
void water_launch_settle(const WaterArgs& a, const double4* refp, int* status, hipStream_t s);
void water_launch_settle_vel(const WaterArgs& a, int* status, hipStream_t s);
// The constrained step with SETTLE in two per-molecule kernels: kick, drift,
// position constraint and velocity correction; then kick and velocity
// constraint.
void water_launch_cvv_settle_first(const WaterArgs& a, double dt, int* status, hipStream_t s);
void water_launch_cvv_settle_second(const WaterArgs& a, double dt, int* status, hipStream_t s);
void water_launch_cvv_first(const WaterArgs& a, double dt, double4* start, double4* free_pos,
                            hipStream_t s);
void water_launch_cvv_correct(const WaterArgs& a, double dt, const double4* free_pos, hipStream_t s);
void water_launch_cvv_kick(const WaterArgs& a, double dt, hipStream_t s);
void water_launch_check(const WaterArgs& a, int* flag, hipStream_t s);
void water_launch_kinetic(const WaterArgs& a, hipStream_t s);
// One workgroup: the real-space partials (4 per real-space workgroup) and the
// k-space partials summed in order into out = {real, lj, intra, virial, recip}.
void water_launch_energy_sum(const WaterArgs& a, double* out, hipStream_t s);
// One workgroup: the kinetic partials of water_launch_kinetic summed into *out.
void water_launch_kinetic_sum(const WaterArgs& a, double* out, hipStream_t s);
int water_site_groups(long n);
// The PME reciprocal term: charges spread, transformed, convolved and
// interpolated back; forces added to frc (after the real-space kernel), the
// energy partials (pme_energy_blocks of them) to part2.
void water_launch_pme(const WaterArgs& a, hipStream_t s);
long pme_energy_blocks(const PmeArgs& p);
// Grid rows per workgroup of the charge spread (its LDS region is that many
// rows of one plane).
int pme_spread_rows(const PmeArgs& p);

}  // namespace mdk
