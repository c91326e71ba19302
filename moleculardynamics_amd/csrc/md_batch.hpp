// Launch interface of the batched-replica step kernel (md_batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "md_common.hpp"
#include "md_kernels.hpp"

namespace mdk {

// One replica: its slice [off, off + n) of the flat atom arrays, its box,
// Langevin constants, noise key and the lanes-per-atom of its force rows.
struct BatchRep {
    long off;
    int n;
    int tpa;
    uint32_t seed_lo, seed_hi;
    Box box;
    double T, c1, c2;
    // the replica's cell grid at rc + skin (neighbor/cell_list.cpp:8-14);
    // usable when 3 or more cells fit on every axis
    int nc[3];
    double w[3];
    int usable;
};

struct BatchArgs {
    const BatchRep* reps;
    int nrep;
    int nmax;        // largest replica (LDS staging size)
    int maxcell;     // largest usable cell grid (LDS cell tables), 0: every rebuild all-pairs
    int mode;        // 0: build rows + forces only; 1: one full BAOAB step
    int mixed;
    double4 *pos, *vel, *frc, *ref;
    uint32_t* nbr;   // row-major rows of `cap` entries per atom
    int* ncnt;
    int cap;
    LJConst lj;
    double rl2, skin, dt;
    unsigned long long* step;  // device step counter (advanced by the kernel)
    unsigned int* ticket;
    double* thermo;  // [thermo_steps][nrep][4] ring (E, W, KE, step) or null
    long thermo_steps;
    long long* builds;  // per replica
    int* overflow;
    int* fix_sat;
};

hipError_t launch_batch(const BatchArgs& a, int block, hipStream_t s);
// Dynamic LDS of one workgroup: the positions, and with maxcell > 0 the
// cell tables of the in-LDS rebuild.
size_t batch_lds_bytes(int nmax, int maxcell);
hipError_t batch_allow_lds(size_t bytes);

}  // namespace mdk
