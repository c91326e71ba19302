// Launch interface of the batched-replica step kernels (md_batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "md_common.hpp"
#include "md_kernels.hpp"

namespace mdk {

// One replica: its slice [off, off + n) of the flat atom arrays, its box,
// Langevin constants, noise key, the lanes-per-atom of its force rows and
// the pieces (workgroups) it is split into.
struct BatchRep {
    long off;
    int n;
    int tpa;
    uint32_t seed_lo, seed_hi;
    Box box;
    double T, c1, c2;
    // the replica's rebuild grid: cells of at least (rc + skin)/2 searched
    // 5 x 5 x 5 (the cell index rule of neighbor/cell_list.cpp:31-38 at that
    // width); usable when 5 or more cells fit on every axis
    int nc[3];
    double w[3];
    int usable;
    int piece0;  // the replica's pieces are piece0 .. piece0 + npiece - 1
    int npiece;
};

// A piece is one workgroup's share of a replica: atoms [s0, s1) in the
// integration kernel and slots [s0, s1) in the rebuild/force kernel.
// Fields: x = replica, y = s0, z = s1, w = piece index (replica-contiguous
// numbering of the per-piece partial arrays).
struct BatchArgs {
    const BatchRep* reps;
    const int4* pieces;  // static order (longest first)
    int4* queue;         // this step's launch order of the force kernel
    int* rflag;          // per replica: rebuild this step
    int nrep, npieces;
    int nmax;        // largest replica (LDS staging size)
    int maxcell;     // largest usable cell grid (LDS cell tables), 0: every rebuild all-pairs
    int mode;        // 0: build rows + forces only; 1: one full BAOAB step
    int mixed;
    double4 *pos, *vel, *frc, *ref;  // by atom index
    // Slot order: each replica's atoms sorted by cell at the last build;
    // rows, the LDS staging and the force pass run in slot order.
    int* slot_of;    // atom -> slot
    int* orig;       // slot -> atom
    uint32_t* nbr;   // row-major rows of `cap` slot entries per slot
    int* ncnt;       // by slot
    int cap;
    LJConst lj;
    double rl2, skin, dt;
    double2* ptop;       // per piece: the two largest displacements since the build
    long long* ppart;    // per piece: E, W (fixed point or FP64 bits), KE (FP64 bits)
    unsigned long long* step;  // device step counter (advanced by the sum kernel)
    double* thermo;  // [thermo_steps][nrep][4] ring (E, W, KE, step) or null
    long thermo_steps;
    long long* builds;  // per replica
    int* overflow;
    int* fix_sat;
};

hipError_t launch_batch(const BatchArgs& a, int block, hipStream_t s);
// Dynamic LDS of one rebuild/force workgroup: the positions and slot map,
// and with maxcell > 0 the atom -> cell map and cell offsets of the in-LDS
// rebuild.
size_t batch_lds_bytes(int nmax, int maxcell);
hipError_t batch_allow_lds(size_t bytes);

}  // namespace mdk
