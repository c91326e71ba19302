// Per-type-pair Lennard-Jones with parameter derivatives and the force
// vector-Jacobian product, for the differentiable path (no reference symbol:
// the reference LJ is single-type, forcefield/lennard_jones.hpp:26; the pair
// terms follow its kernel, lennard_jones.cpp:50-88, with sigma, epsilon and
// the cutoff taken per unordered type pair).
//
// Layout in HBM (upload order; nothing is permuted, so every per-atom array
// of the adjoint lines up with the host's):
//   pos double4 x, y, z, type (small integer stored as a double)
//   vel double4 vx, vy, vz, mass;  frc double4
//   nbr full Verlet list in the wave-blocked layout (nbr_index), bit 31 set
//       when the row atom has the larger upload index
//   part_ew   2 doubles per atom: E and W of the pairs whose lower index it is
//   part_q    per atom, kMaxTypes rows (one per partner type) of kMixQ
//             (force pass) or kMixG (adjoint pass) sums over the same pairs
#pragma once

#include "md_common.hpp"
#include "md_kernels.hpp"

namespace mdk {

constexpr int kMaxTypes = 4;
constexpr int kTypePairs = kMaxTypes * kMaxTypes;
constexpr int kMixQ = 6;  // E, W, dE/deps, dE/dsigma, dW/deps, dW/dsigma
constexpr int kMixG = 2;  // dL/deps, dL/dsigma

// Per ordered type pair k = a * nt + b (symmetric tables).
struct MixTab {
    int nt;
    int lfs;
    double sig2[kTypePairs], eps[kTypePairs], rc[kTypePairs], rc2[kTypePairs];
    double inv_eps[kTypePairs], inv_sig[kTypePairs];
    double urc[kTypePairs], dudr[kTypePairs];  // u(rc), u'(rc) = -w(rc)/rc
    double wrc[kTypePairs], yrc_rc[kTypePairs];  // w(rc), y(rc)/rc
};

struct MixArgs {
    long n;
    Box box;
    MixTab tab;
    const double4* pos;
    double4* frc;
    const uint32_t* nbr;
    const int* ncnt;
    int cap;
    double* part_ew;
    double* part_q;
    // adjoint pass
    const double4* vec;  // cotangent on the forces
    double4* out;        // cotangent on the positions
    double w_energy, w_virial;  // weights of the direct E and W terms
    int accumulate;             // out += (1) or out = (0)
};

struct MixListArgs {
    long n;
    Grid grid;
    Box box;
    double rl2;
    const double4* pos;
    double4* ref;  // positions at this build (displacement check), or null
    const int* cell_of;
    const int* cell_start;
    const int* cell_atoms;
    uint32_t* nbr;
    int* ncnt;
    int cap;
    Scalars* sc;
};

void launch_mix_list(const MixListArgs& a, hipStream_t s);
// deriv = 0: forces, part_ew; deriv = 1: also part_q (kMixQ rows).
void launch_mix_force(const MixArgs& a, int deriv, hipStream_t s);
// out (+)= J^T vec + w_energy grad E + w_virial grad W; part_q (kMixG rows).
void launch_mix_adjoint(const MixArgs& a, hipStream_t s);
// e_out[0], w_out[0] = sums of part_ew (fixed order).
void launch_mix_sum_ew(long n, const double* part_ew, double* e_out, double* w_out, hipStream_t s);
// tables[(a*nt + b)*q + k] (+)= sum over atoms i of type a of part_q[i][b][k].
void launch_mix_tables(long n, const double4* pos, int nt, int q, const double* part_q,
                       double* tables, int accumulate, hipStream_t s);
// One steepest-descent move x <- wrap(x + gamma f), gamma = min(step, max_disp / max|f|).
void launch_mix_sd(long n, const Box& box, double4* pos, const double4* frc, double step,
                   double max_disp, double* part_d, double* gamma, hipStream_t s);
// Adjoint helpers: vec = a (lv + lu) (frame0: a lu), a = (dt/2)/m; and the
// backward drift/kick: lu = lv + dt lx, lv = lu.
void launch_mix_vec(long n, const double4* vel, const double4* lv, const double4* lu,
                    double half_dt, int frame0, double4* vec, hipStream_t s);
void launch_mix_back(long n, double dt, const double4* lx, double4* lv, double4* lu, hipStream_t s);

}  // namespace mdk
