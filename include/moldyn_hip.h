/*
 * moldyn_hip.h -- C ABI of the MI355X (gfx950) force / neighbour / integrator
 * path of moldyn (VishnuR23/MolecularDynamics).
 *
 * Plain pointers and sizes only: no C++ or torch types cross this boundary.
 * All host arrays are in the caller's atom order ("original" indices, the
 * order atoms were uploaded in); the device keeps atoms in a cell-sorted
 * order internally and translates at the boundary. Every function returns
 * an md_status; md_last_error(ctx) describes the last failure. No
 * exception crosses the ABI: the C++ wrappers map MD_ERR_INVALID_ARGUMENT to
 * std::invalid_argument and MD_ERR_RUNTIME / MD_ERR_CAPACITY to
 * std::runtime_error, as the reference throws.
 *
 * Reference interfaces each entry point replaces are cited as
 * file:line of VishnuR23/MolecularDynamics (paths under src/moldyn/).
 *
 * Threading: one context per host thread; calls on a context are ordered on
 * its HIP stream and are not thread-safe.
 */
#ifndef MOLDYN_HIP_H
#define MOLDYN_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum md_status {
    MD_OK = 0,
    MD_ERR_INVALID_ARGUMENT = 1, /* maps to std::invalid_argument */
    MD_ERR_RUNTIME = 2,          /* maps to std::runtime_error */
    MD_ERR_HIP = 3,              /* a HIP runtime call failed */
    MD_ERR_CAPACITY = 4,         /* a neighbour row exceeded its capacity */
    MD_ERR_STATE = 5             /* call order violated (e.g. no system uploaded) */
} md_status;

/* Pair sources of LennardJones::computeForces overloads
 * (forcefield/lennard_jones.hpp:36-48). */
typedef enum md_pair_source {
    MD_PAIRS_BRUTE = 0,  /* every i<j pair: computeForces(System&) */
    MD_PAIRS_CELLS = 1,  /* linked cells at the LJ cutoff: computeForces(System&, const CellList&) */
    MD_PAIRS_VERLET = 2  /* the context's Verlet list: computeForces(System&, const VerletList&) */
} md_pair_source;

/* Arithmetic of the pair kernel. MD_FP64 reproduces the reference's FP64
 * results to 1e-12 relative. MD_MIXED evaluates the pair terms in FP32 and
 * accumulates per-atom forces, energy and virial in deterministic 64-bit
 * fixed point; integration stays FP64. */
typedef enum md_precision { MD_FP64 = 0, MD_MIXED = 1 } md_precision;

typedef enum md_integrator_kind {
    MD_VELOCITY_VERLET = 0, /* integrate/velocity_verlet.hpp:39-80 */
    MD_LANGEVIN_BAOAB = 1,  /* integrate/langevin.hpp:29-97 */
    MD_NOSE_HOOVER = 2      /* integrate/nose_hoover.hpp:26-67 */
} md_integrator_kind;

typedef struct md_ctx md_ctx;

/* Lennard-Jones parameters: LennardJones(sigma, epsilon, cutoff, truncation)
 * (forcefield/lennard_jones.hpp:26; truncation enum :11). */
typedef struct md_lj_params {
    double sigma;
    double epsilon;
    double cutoff;
    int linear_force_shift; /* 0: Truncated, 1: LinearForceShift */
} md_lj_params;

/* Integrator parameters. VelocityVerlet(dt) (velocity_verlet.hpp:41);
 * LangevinBAOAB(dt, friction, temperature, seed) (langevin.hpp:31);
 * NoseHooverChain(dt, temperature, tau, chainLength, dof) (nose_hoover.hpp:31). */
typedef struct md_integrator {
    int kind;          /* md_integrator_kind */
    double dt;
    double friction;   /* BAOAB gamma */
    double temperature;
    double tau;        /* NHC period */
    int chain_length;  /* NHC M >= 1 */
    long dof;          /* NHC degrees of freedom */
    uint64_t seed;     /* BAOAB noise seed (counter-based stream, see md_run) */
} md_integrator;

/* Per-step thermodynamic record written by md_run. */
typedef struct md_thermo {
    double potential; /* pair energy of the step's force evaluation (no LRC) */
    double virial;    /* raw sum of w(r), not divided by 3 */
    double kinetic;   /* sum 0.5 m v^2 at the end of the step */
    double conserved; /* NHC conservedQuantityContribution(), 0 otherwise */
    double pairs;     /* unordered pairs inside the cutoff in the step's force call */
} md_thermo;

/* ---- context ------------------------------------------------------------ */

/* Creates a context on the current HIP device. `hip_stream` is a hipStream_t
 * to order work on, or NULL for a context-owned stream (default priority). */
md_status md_ctx_create(void* hip_stream, md_ctx** out);
md_status md_ctx_destroy(md_ctx* ctx);
const char* md_last_error(const md_ctx* ctx);
/* Blocks until all work queued on the context's stream has finished. */
md_status md_ctx_synchronize(md_ctx* ctx);

/* ---- system state (core/system.hpp:11-60) -------------------------------- */

/* Uploads n atoms in a periodic orthorhombic box of lengths L[3]: positions
 * x,y,z, velocities vx,vy,vz (NULL = zero) and masses m (NULL = 1). Forces
 * start at zero. Replaces System(Box) + addAtom (system.hpp:13-21). */
md_status md_system_upload(md_ctx* ctx, long n, const double* L, const double* x,
                           const double* y, const double* z, const double* vx,
                           const double* vy, const double* vz, const double* m);
/* Copies state back in upload order; any pointer may be NULL. Replaces
 * System::position/velocity/force (system.hpp:26-28). */
md_status md_system_download(md_ctx* ctx, double* x, double* y, double* z, double* vx,
                             double* vy, double* vz, double* fx, double* fy, double* fz);
/* System::setPosition / setVelocity for every atom (system.hpp:33-34). */
md_status md_system_set_positions(md_ctx* ctx, const double* x, const double* y,
                                  const double* z);
md_status md_system_set_velocities(md_ctx* ctx, const double* vx, const double* vy,
                                   const double* vz);
/* System::setForce for every atom (system.hpp:35): forces an integrator
 * pass will use when no force call precedes it. */
md_status md_system_set_forces(md_ctx* ctx, const double* fx, const double* fy, const double* fz);
/* System::kineticEnergy (system.cpp:124-131), FP64 deterministic reduction. */
md_status md_system_kinetic_energy(md_ctx* ctx, double* ke);
/* System::removeCenterOfMassMotion (system.cpp:147-164). */
md_status md_system_remove_com_motion(md_ctx* ctx);
md_status md_system_size(const md_ctx* ctx, long* n);

/* ---- Lennard-Jones (forcefield/lennard_jones.{hpp:24-98,cpp:32-151}) ---- */

/* Sets the pair potential; precomputes u(rc) and du/dr(rc)
 * (lennard_jones.cpp:32-48). */
md_status md_lj_set(md_ctx* ctx, const md_lj_params* p);
/* LennardJones::computeForces overloads (lennard_jones.cpp:94-119): zero the
 * forces, evaluate every offered pair with r2 < rc2, write forces, return
 * E (pair energy, no tail) and W (raw virial sum). MD_PAIRS_CELLS fails with
 * MD_ERR_INVALID_ARGUMENT when fewer than 3 cells per side fit
 * (lennard_jones.cpp:99-111). MD_PAIRS_VERLET uses the list as last built. */
md_status md_lj_forces(md_ctx* ctx, int pair_source, int precision, double* energy,
                       double* virial);
/* LennardJones::computeEnergyVirial (lennard_jones.cpp:90-92): all i<j
 * pairs, forces untouched. */
md_status md_lj_energy_virial(md_ctx* ctx, double* energy, double* virial);
/* LennardJones::longRangeCorrection / longRangeCorrectionPressure
 * (lennard_jones.cpp:121-151); host arithmetic, no device needed. */
md_status md_lj_tail(const md_lj_params* p, long n, const double* L, double* u_lrc,
                     double* p_lrc);

/* ---- neighbour lists (neighbor/{cell_list,verlet_list}.{hpp,cpp}) ------- */

/* CellList(box, cutoff).usable() (cell_list.cpp:8-14, :44-46): 1 when
 * floor(L_d/cutoff) >= 3 on every axis. Host arithmetic. */
int md_cells_usable(const double* L, double cutoff);
/* VerletList(box, cutoff, skin) (verlet_list.hpp:26): configures the list. */
md_status md_neigh_config(md_ctx* ctx, double cutoff, double skin);
/* VerletList::build (verlet_list.cpp:8-42): pairs with
 * |minimum image|^2 < (rc+skin)^2 from the cell grid (brute force when it
 * is unusable), reference positions recorded, build counter incremented.
 * Also re-sorts the device atom order by cell. */
md_status md_neigh_build(md_ctx* ctx);
/* VerletList::needsRebuild (verlet_list.cpp:44-62): sum of the two largest
 * minimum-image displacements since the build exceeds the skin. */
md_status md_neigh_needs_rebuild(md_ctx* ctx, int* flag);
/* VerletList::rebuildIfNeeded (verlet_list.cpp:64-68). */
md_status md_neigh_rebuild_if_needed(md_ctx* ctx, int* rebuilt);
/* VerletList::rebuildCount (verlet_list.cpp:70-72). */
md_status md_neigh_rebuild_count(const md_ctx* ctx, long* count);
/* VerletList::forEachPair (verlet_list.hpp:52-57) as data: the unordered
 * pair set (i<j, upload indices) sorted by (i, j), 2*npairs uint32 values.
 * With cap too small, *npairs receives the count and MD_ERR_CAPACITY is
 * returned. */
md_status md_neigh_export_pairs(md_ctx* ctx, uint32_t* ij, long cap, long* npairs);
/* Neighbour row capacity (entries per atom) of the device list; 0 = auto. */
md_status md_neigh_set_capacity(md_ctx* ctx, int per_atom);
/* The automatic row capacity for n atoms in box L with the given cutoff and
 * skin: ceil(2 * (4/3) pi (rc+skin)^3 rho + 32), at most n - 1, rounded up to
 * a multiple of 8 (0 for invalid arguments). Host arithmetic. No reference
 * counterpart: the reference's list is a growing std::vector
 * (neighbor/verlet_list.cpp:8-42). A row that outgrows the capacity during
 * md_run makes md_run restore its starting state, double the capacity and
 * run again. */
int md_neigh_auto_capacity(long n, const double* L, double cutoff, double skin);
/* Bricks of list cells for the LDS-staged pair kernels: cells per brick side
 * (1 .. 8), -1 automatic (4), 0 off (the default). With bricks on, each
 * pair-kernel workgroup stages the positions of one brick's two-cell halo
 * in LDS and the list rows index that halo. They apply to systems above
 * 32,768 atoms with fixed-point positions whose grid has at least the brick
 * width plus 4 cells per axis; the pair set, the forces, the mixed-path
 * energies and virials are the same bits with and without them. No
 * reference counterpart (a device data layout); the list must be built
 * again afterwards. */
md_status md_neigh_set_bricks(md_ctx* ctx, int cells);
/* Number of bricks of the current list layout (0 when bricks are off). */
md_status md_neigh_bricks(const md_ctx* ctx, int* count);

/* ---- integrators ---------------------------------------------------------- */

/* Selects the integrator for md_run; resets thermostat state.
 * NoseHooverChain throws for chainLength < 1 (nose_hoover.cpp:36-38):
 * MD_ERR_INVALID_ARGUMENT. */
md_status md_integrator_set(md_ctx* ctx, const md_integrator* it);
/* Runs nsteps of the selected integrator entirely on the device, with the
 * force callback of the reference driver (apps/moldyn_run.cpp:581-587):
 * rebuild-if-needed then Verlet-list LJ forces at `precision`. Forces must be
 * current on entry (call md_lj_forces once first, as the reference primes
 * them, velocity_verlet.hpp:34-38). When `thermo` is non-NULL it receives
 * nsteps records. BAOAB noise comes from a counter-based generator keyed by
 * (seed, step, upload index), so it is independent of the device atom order;
 * it is not the reference's sequential mt19937_64 stream. */
md_status md_run(md_ctx* ctx, long nsteps, int precision, md_thermo* thermo);
/* Single per-atom integrator passes, for a host-driven step with an
 * arbitrary force callback between them (the reference's
 * `template <class ForceFn> step(System&, ForceFn&&)`,
 * velocity_verlet.hpp:48-74, langevin.hpp:36-88, nose_hoover.hpp:37-42):
 *   md_half_kick:   v = fma((dt/2) * (1/m), f, v)      (velocity_verlet.hpp:55-58)
 *   md_drift:       r = wrap(fma(h, v, r))             (velocity_verlet.hpp:61-64)
 *   md_langevin_o:  v = c1 v + (c2 sqrt(T/m)) xi, xi from the counter-based
 *                   stream (seed, counter, upload index) (langevin.hpp:58-64)
 *   md_scale_velocities: v = v * s                      (nose_hoover.cpp:119-124) */
md_status md_half_kick(md_ctx* ctx, double dt);
md_status md_drift(md_ctx* ctx, double h);
md_status md_langevin_o(md_ctx* ctx, double c1, double c2, double temperature, uint64_t seed,
                        uint64_t counter);
md_status md_scale_velocities(md_ctx* ctx, double s);
/* LennardJones::computeForces(System&, const CellList&) with the CellList's
 * own cutoff (lennard_jones.cpp:99-114, cell_list.cpp:8-14): candidates are
 * the 27 cells around each atom on a grid of floor(L/cell_cutoff) cells per
 * side; MD_ERR_INVALID_ARGUMENT when fewer than 3 fit. */
md_status md_lj_forces_cells(md_ctx* ctx, double cell_cutoff, double* energy, double* virial);

/* ---- trajectory analysis (analyze/rdf.hpp:16-34, msd.hpp:19-42,
 * vacf.hpp:151-174) -----------------------------------------------------------
 * Device-resident accumulators fed from a context's current positions or
 * velocities (its device copy; no host transfer). Each handle lives on the
 * device current at creation and takes contexts on that device; work is
 * queued on the context's stream; the *_result / *_diffusion calls wait for
 * the device. */
typedef struct md_rdf md_rdf;
typedef struct md_msd md_msd;
typedef struct md_vacf md_vacf;

/* Pair sources of md_rdf_accumulate. */
#define MD_RDF_AUTO 0      /* cells when 3 or more of width >= rMax fit per side, else all pairs */
#define MD_RDF_ALL_PAIRS 1 /* every i<j pair (the reference's loop) */
#define MD_RDF_CELLS 2     /* 27 cells of width >= rMax around each atom */

/* Rdf(rMax, bins) (rdf.hpp:18, rdf.cpp:16-17): bin width rMax / bins. */
md_status md_rdf_create(double r_max, long bins, md_rdf** out);
md_status md_rdf_destroy(md_rdf* rdf);
const char* md_rdf_last_error(const md_rdf* rdf);
/* Rdf::accumulate (rdf.cpp:19-46): counts every unordered pair of the
 * context's system with r = |minimum image| < rMax in bin floor(r / width);
 * MD_ERR_INVALID_ARGUMENT when rMax exceeds half the smallest box length
 * (the reference throws std::invalid_argument). Counts are exact integers,
 * the same for every pair source. */
md_status md_rdf_accumulate(md_rdf* rdf, md_ctx* ctx, int source);
/* Rdf::binCentres, g, frames (rdf.cpp:48-75): bins values each (any pointer
 * may be NULL); hist receives the raw pair counts. */
md_status md_rdf_result(md_rdf* rdf, double* centres, double* g, uint64_t* hist, long* frames);

/* Msd(natoms, maxLag) (msd.hpp:21, msd.cpp:12-20). */
md_status md_msd_create(long natoms, long max_lag, md_msd** out);
md_status md_msd_destroy(md_msd* msd);
const char* md_msd_last_error(const md_msd* msd);
/* Msd::push (msd.cpp:22-52): unwraps incrementally with the minimum image of
 * the step from the previous push, by atom upload index, and adds every lag
 * up to min(pushes, maxLag). MD_ERR_INVALID_ARGUMENT when the context's atom
 * count differs from natoms. */
md_status md_msd_push(md_msd* msd, md_ctx* ctx);
/* Msd::msd (msd.cpp:54-62): max_lag + 1 values. */
md_status md_msd_result(md_msd* msd, double* out);
/* Msd::diffusionCoefficient (msd.cpp:64-92). */
md_status md_msd_diffusion(md_msd* msd, double dt, long fit_from, long fit_to, double* D);

/* Vacf(natoms, maxLag) (vacf.hpp:153, vacf.cpp:11-17). */
md_status md_vacf_create(long natoms, long max_lag, md_vacf** out);
md_status md_vacf_destroy(md_vacf* vacf);
const char* md_vacf_last_error(const md_vacf* vacf);
/* Vacf::push (vacf.cpp:19-40). */
md_status md_vacf_push(md_vacf* vacf, md_ctx* ctx);
/* Vacf::vacf (vacf.cpp:52-63), max_lag + 1 values normalised to lag 0; the
 * unnormalised correlation (vacf.cpp:42-50) when `normalised` is 0. */
md_status md_vacf_result(md_vacf* vacf, int normalised, double* out);
/* Vacf::diffusionCoefficient (vacf.cpp:65-85): Green-Kubo, trapezoid rule. */
md_status md_vacf_diffusion(md_vacf* vacf, double dt, double* D);

/* ---- device peaks ------------------------------------------------------------ */

/* No reference counterpart. Measures the current device's attainable HBM copy
 * bandwidth (bytes read + written per second, 1 GiB buffers) and FP64 / FP32
 * vector FMA rates (flops per second, FMA = 2, FP32 without packed forms),
 * the peaks that roofline fractions are quoted against. Any pointer may be
 * NULL to skip that measurement. Allocates 2 GiB for the duration of the call. */
md_status md_device_peaks(double* hbm_bytes_per_s, double* fp64_flops, double* fp32_flops);

/* ---- allocation-free force evaluation ------------------------------------- */

#define MD_NOALLOC_MAX_ATOMS (1L << 20) /* atoms per md_lj_forces_noalloc call */
#define MD_NOALLOC_MAX_CELLS (1L << 20) /* cells of its CellList grid */

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): md_lj_forces with MD_PAIRS_BRUTE or MD_PAIRS_CELLS, at
 * FP64, on caller-owned device arrays and without a context. It calls no
 * hipMalloc or hipFree: every intermediate buffer lives in one fixed device
 * array of the library (about 60 MiB per device, held from the first call on
 * that device until the process ends).
 *   x, y, z:    device arrays of n positions (box of lengths L[3], host array)
 *   fx, fy, fz: device arrays receiving the n forces (all NULL: E and W only)
 *   energy, virial: host pointers (may be NULL) for E and the raw virial W
 *   hip_stream: the hipStream_t to order the work on (NULL: the null stream);
 *               the call returns when that stream has finished it.
 * n > MD_NOALLOC_MAX_ATOMS or a grid of more than MD_NOALLOC_MAX_CELLS cells
 * is refused with MD_ERR_CAPACITY; an unusable CellList with
 * MD_ERR_INVALID_ARGUMENT. md_last_error(NULL) describes the calling
 * thread's last failure. Calls on one device share the fixed array: calls
 * that overlap on one GPU (other threads or streams) race. Results equal
 * md_lj_forces on a context holding the same atoms in the same order. */
md_status md_lj_forces_noalloc(int pair_source, const md_lj_params* p, long n, const double* L,
                               const double* x, const double* y, const double* z, double* fx,
                               double* fy, double* fz, double* energy, double* virial,
                               void* hip_stream);
/* NoseHooverChain::conservedQuantityContribution (nose_hoover.cpp:45-54)
 * and the chain arrays (length chain_length each; NULL to skip). */
md_status md_nhc_state(md_ctx* ctx, double* conserved, double* xi, double* p_xi);

/* ---- seeded inputs (util/lattice.cpp:15-64, util/random.hpp:37-76) ------- */

/* fccLattice(cells, density, temperature, seed): N = 4 cells^3 atoms, cubic
 * box L = cbrt(N/density) written to L[3], positions folded into
 * [-L/2, L/2), Maxwell-Boltzmann velocities from mt19937_64 + Marsaglia's
 * polar method drawn x,y,z per atom in (ix, iy, iz, basis) order, mass 1,
 * centre-of-mass motion removed. Bit-identical to the reference on the same
 * C++ standard library. Host arithmetic; arrays hold 4*cells^3 values. */
md_status md_fcc_lattice(int cells, double density, double temperature, uint64_t seed,
                         double* L, double* x, double* y, double* z, double* vx, double* vy,
                         double* vz);

/* ---- batched independent replicas ------------------------------------------
 * No reference counterpart as one object: the reference runs one System per
 * process (apps/moldyn_run.cpp:570-777). A batch is many independent LJ
 * systems, each with its own box, Verlet list (verlet_list.cpp:8-69) and
 * LangevinBAOAB thermostat (langevin.hpp:36-88), advanced together: one
 * set of three launches per step for all replicas (each replica split into
 * pieces, one workgroup each), multi-step segments captured in a HIP graph.
 * Forces use MD_MIXED or MD_FP64 pair arithmetic; LJ Truncated. */
typedef struct md_batch md_batch;

/* The replica recipe: from Rng(master_seed) (util/random.hpp:37-48), per
 * replica in order: fcc cells uniform in {3..10} (N = 4 cells^3 in
 * {108, ..., 4000}), density uniform in [0.5, 0.9), temperature uniform in
 * [0.7, 1.5); lattice seed master_seed + r. Host arithmetic. */
md_status md_batch_draw(int count, uint64_t master_seed, int* cells, double* rho, double* T,
                        uint64_t* seeds);
/* Flat inputs: replica r owns atoms [sum n[<r], sum n[<=r]) of x..vz (mass 1);
 * L holds 3 lengths per replica; T and seeds one per replica. */
md_status md_batch_create(void* hip_stream, int nrep, const long* n, const double* L,
                          const double* x, const double* y, const double* z, const double* vx,
                          const double* vy, const double* vz, const double* T,
                          const uint64_t* seeds, const md_lj_params* lj, double skin, double dt,
                          double friction, int precision, md_batch** out);
md_status md_batch_destroy(md_batch* b);
const char* md_batch_last_error(const md_batch* b);
/* Builds every replica's rows and primes the forces (the reference primes
 * its force callback once before the first step, velocity_verlet.hpp:34-38). */
md_status md_batch_prime(md_batch* b);
/* Advances every replica nsteps BAOAB steps; with use_graph, `segment`
 * launches are captured once in a HIP graph and replayed. thermo_out (or
 * NULL) receives nsteps x nrep rows of (E, W, KE, step). */
md_status md_batch_run(md_batch* b, long nsteps, long segment, int use_graph, double* thermo_out);
/* Lanes per atom of each replica's force rows (1, 2, 4, 8, 16) and the
 * workgroup size of the rebuild/force kernel (256, 512, 1024). */
md_status md_batch_set_tpa(md_batch* b, const int* tpa, int block);
/* Workgroups (pieces) each replica is split into: 1 .. min(32, n / 32). */
md_status md_batch_set_pieces(md_batch* b, const int* pieces);
/* Times candidate (workgroup size, per-size-class piece counts and
 * lanes-per-atom) choices on the batch itself from a saved state, applies
 * the fastest, restores the state. Any output may be NULL. */
md_status md_batch_autotune(md_batch* b, int trial_steps, int* tpa_out, int* pieces_out,
                            int* block_out, double* best_ms_per_step);
md_status md_batch_download(md_batch* b, double* x, double* y, double* z, double* vx, double* vy,
                            double* vz, double* fx, double* fy, double* fz);
/* Per-replica list build counts (VerletList::rebuildCount). */
md_status md_batch_rebuild_counts(md_batch* b, long* counts);

/* ---- rigid SPC/E water with Ewald -------------------------------------------
 * SpceForceField (forcefield/spce_forcefield.{hpp:23-51,cpp:8-48}) = oxygen LJ
 * (Truncated + tail) + Ewald (forcefield/ewald.{hpp:49-76,cpp:33-188}) under
 * RigidWater SHAKE/RATTLE (constrain/rigid_water.{hpp:28-66,cpp:30-161}) and
 * ConstrainedVelocityVerlet (integrate/constrained_velocity_verlet.hpp:22-68).
 * Units: A, amu, e, energies in K (E/kB), time in units of 1.0966875353 ps
 * (core/units.hpp:33-40). Sites are O, H, H per molecule. */
typedef struct md_water md_water;

typedef struct md_water_params {
    double qH;          /* 0.42380 e; qO = -2 qH (spce.hpp:17) */
    double sigma_o;     /* 3.16555789 A */
    double epsilon_o;   /* 78.19743111 K */
    double r_oh;        /* 1.0 A */
    double angle_hoh;   /* 109.47 deg */
    double mass_o;      /* 15.9994 amu */
    double mass_h;      /* 1.008 amu */
    double cutoff;      /* real-space and LJ cutoff (SpceForceField default 10 A) */
    int kmax;           /* Ewald(cutoff, kmax) (ewald.hpp:51) */
    int n2max;          /* |n|^2 < n2max; <= 0 selects kmax^2 + 2 (ewald.cpp:119) */
    double alpha;       /* <= 0 selects the reference's 5.6 / min(L) (ewald.cpp:40) */
    double skin;        /* Verlet skin of the real-space list */
    double tolerance;   /* RigidWater tolerance (1e-12) */
    int max_iterations; /* RigidWater maxIterations (500) */
} md_water_params;

/* Ewald parameters for a target RMS force accuracy relative to the force
 * between two unit charges 1 A apart, from the Kolafa-Perram estimates (no
 * reference counterpart: the reference fixes alpha = 5.6/L, ewald.cpp:40). */
md_status md_ewald_parameters(long n, const double* L, double sum_q2, double cutoff,
                              double accuracy, double* alpha, int* kmax);
/* Config-3 input builder (no reference counterpart; the reference rescales
 * a NIST configuration, apps/water_rdf.cpp:53-83): m^3 molecules on a cubic
 * lattice, uniform random orientations (Shoemake quaternions) and
 * Maxwell-Boltzmann velocities from a seeded mt19937_64. */
md_status md_water_lattice(int molecules_per_side, double box_length, uint64_t seed,
                           double temperature, const md_water_params* p, double* x, double* y,
                           double* z, double* vx, double* vy, double* vz);
/* buildSpceSystem (forcefield/spce.cpp:8-40): charges and masses by site. */
md_status md_water_create(void* hip_stream, long n, const double* L, const double* x,
                          const double* y, const double* z, const double* vx, const double* vy,
                          const double* vz, const md_water_params* p, md_water** out);
md_status md_water_destroy(md_water* w);
const char* md_water_last_error(const md_water* w);
/* SpceForceField::computeForces: forces zeroed then dispersion + Ewald;
 * terms[6] = {Edisp, ELRC, Ereal, Erecip, Eself, Eintra} (K), *pe their sum.
 * MD_MIXED evaluates real-space erfc in FP32 (k-space stays FP64). */
md_status md_water_forces(md_water* w, int precision, double* terms, double* pe);
md_status md_water_alpha(const md_water* w, double* alpha, int* nk_half);
md_status md_water_build_list(md_water* w);
/* RigidWater::constrainPositions against reference positions (returns the
 * sweep count; MD_ERR_RUNTIME on non-convergence or singularity). */
md_status md_water_shake(md_water* w, const double* rx, const double* ry, const double* rz,
                         int* iterations);
/* RigidWater::constrainVelocities. */
md_status md_water_rattle(md_water* w, int* iterations);
/* Closed-form rigid constraints (SETTLE, Miyamoto & Kollman 1992), the
 * replacement of RigidWater's SHAKE/RATTLE that the reference plans
 * (constrain/rigid_water.hpp:20-24; absent from its sources): the same fixed
 * points solved exactly, one pass per molecule. md_water_settle is
 * RigidWater::constrainPositions against reference positions,
 * md_water_settle_velocities RigidWater::constrainVelocities;
 * MD_ERR_RUNTIME when a molecule's step has no solution (a reference
 * molecule that is degenerate, or turned by 90 degrees or more). The two
 * hydrogen masses of a molecule must be equal. */
md_status md_water_settle(md_water* w, const double* rx, const double* ry, const double* rz);
md_status md_water_settle_velocities(md_water* w);
/* Constraint algorithm of md_water_run: RigidWater's SHAKE/RATTLE (default)
 * or SETTLE, whose step takes two per-molecule kernels in place of five. */
#define MD_RIGID_SHAKE 0
#define MD_RIGID_SETTLE 1
md_status md_water_set_constraint(md_water* w, int method);
/* Reciprocal-space method of md_water_forces / md_water_run: the
 * reference's k-vector sums (MD_KSPACE_EWALD, the default;
 * forcefield/ewald.cpp:112-172) or smooth particle-mesh Ewald (MD_KSPACE_PME,
 * Essmann et al. 1995) over the same k-vector set (0 < |n|^2 < n2max), with a
 * grid of `grid` points per axis (0: 4.2 kmax rounded up to even; at least
 * 2 kmax + 1) and B-spline order `order` (3 .. 10; 0: 8). The reference names
 * PME as future work for the O(K N) sum and has no counterpart. */
#define MD_KSPACE_EWALD 0
#define MD_KSPACE_PME 1
md_status md_water_set_kspace(md_water* w, int method, int grid, int order);
/* nsteps of ConstrainedVelocityVerlet::step with the SPC/E force field;
 * pe/ke (or NULL) receive per-step values in K. */
md_status md_water_run(md_water* w, long nsteps, double dt, int precision, double* pe, double* ke);
md_status md_water_download(md_water* w, double* x, double* y, double* z, double* vx, double* vy,
                            double* vz, double* fx, double* fy, double* fz);
md_status md_water_set_state(md_water* w, const double* x, const double* y, const double* z,
                             const double* vx, const double* vy, const double* vz);

/* ---- per-type-pair LJ, parameter derivatives and the adjoint ---------------
 * No reference counterpart: the reference LJ is single-type
 * (forcefield/lennard_jones.hpp:26). The pair terms follow
 * LennardJones::computePairSum (lennard_jones.cpp:50-88) with sigma, epsilon
 * and the cutoff taken per unordered type pair; the list is the Verlet list
 * with skin (neighbor/verlet_list.cpp:8-73); the trajectory is
 * VelocityVerlet::step (integrate/velocity_verlet.hpp:48-74). Derivatives are
 * taken at fixed cutoffs. Arrays are in upload order; up to 4 types.
 * Tables of ntypes*ntypes entries are row-major and symmetric. Per-type-pair
 * outputs use the canonical order of unordered pairs (a <= b), row-major:
 * (0,0), (0,1), ..., (0,nt-1), (1,1), ... */
typedef struct md_mix md_mix;

md_status md_mix_create(void* hip_stream, md_mix** out);
md_status md_mix_destroy(md_mix* m);
const char* md_mix_last_error(const md_mix* m);
/* type[i] in 0..3; mass (or NULL: 1) and velocities (or NULL: 0). */
md_status md_mix_upload(md_mix* m, long n, const double* L, const double* x, const double* y,
                        const double* z, const double* vx, const double* vy, const double* vz,
                        const double* mass, const int* type);
/* LennardJones(sigma_ab, epsilon_ab, cutoff_ab, Truncated | LinearForceShift)
 * for every type pair (lennard_jones.cpp:32-48). */
md_status md_mix_set_params(md_mix* m, int ntypes, const double* sigma, const double* epsilon,
                            const double* cutoff, int linear_force_shift);
/* Verlet skin of the list (default 0.3); the list radius is max cutoff + skin. */
md_status md_mix_set_skin(md_mix* m, double skin);
/* Forces into the device state; E and W (raw sum of w, as EnergyVirial).
 * tables (or NULL) receives 6 values per type pair:
 * {E_ab, W_ab, dE/deps_ab, dE/dsigma_ab, dW/deps_ab, dW/dsigma_ab}. */
md_status md_mix_forces(md_mix* m, double* energy, double* virial, double* tables);
/* Vector-Jacobian product at the current positions for a cotangent g on the
 * forces (NULL: zero) and weights on E and W:
 *   o = (dF/dx)^T g + w_energy dE/dx + w_virial dW/dx   ((dF/dx) = -Hessian)
 *   dtheta (2 per type pair) = {g . dF/dp + w_energy dE/dp + w_virial dW/dp}
 *   for p = eps_ab, sigma_ab. */
md_status md_mix_force_vjp(md_mix* m, double w_energy, double w_virial, const double* gx,
                           const double* gy, const double* gz, double* ox, double* oy, double* oz,
                           double* dtheta);
/* Steepest descent: nsteps moves x <- wrap(x + gamma F) with
 * gamma = min(step, max_disp / max_i |F_i|), the list rebuilt when the
 * displacement check asks for it. energies (or NULL) receives nsteps + 1
 * values (before the first move and after each). */
md_status md_mix_minimize(md_mix* m, int nsteps, double step, double max_disp, double* energies);
/* nsteps velocity-Verlet steps, checkpoints every checkpoint_every steps
 * (kept for md_mix_adjoint); energies / virials (or NULL) receive the
 * nsteps + 1 frame values E(x_t), W(x_t), t = 0..nsteps. */
md_status md_mix_run(md_mix* m, double dt, int nsteps, int checkpoint_every, double* energies,
                     double* virials);
/* Adjoint of the last md_mix_run for
 *   L = sum_t [w_energy E(x_t) + w_virial W(x_t)] + g . x_T + gv . v_T
 * (g, gv or NULL for zero). dtheta (2 per type pair) = {dL/deps_ab,
 * dL/dsigma_ab}; d* (or NULL) = dL/dx_0, dL/dv_0. Frames are recomputed from
 * the checkpoints; the final state of the run is restored afterwards. */
md_status md_mix_adjoint(md_mix* m, double w_energy, double w_virial, const double* gx,
                         const double* gy, const double* gz, const double* gvx, const double* gvy,
                         const double* gvz, double* dtheta, double* dx, double* dy, double* dz,
                         double* dvx, double* dvy, double* dvz);
md_status md_mix_download(md_mix* m, double* x, double* y, double* z, double* vx, double* vy,
                          double* vz, double* fx, double* fy, double* fz);
md_status md_mix_set_state(md_mix* m, const double* x, const double* y, const double* z,
                           const double* vx, const double* vy, const double* vz);
md_status md_mix_list_builds(const md_mix* m, long* builds);
/* Config-5 input builder: uniform random positions (u - 1/2) L per
 * coordinate (the recipe of tests/test_analysis.cpp:24-28 with
 * Rng::uniform), then Maxwell-Boltzmann velocities normal() sqrt(T/m) from
 * the same stream and the centre-of-mass motion removed
 * (core/system.cpp:147-164). mass may be NULL (unit masses). */
md_status md_mix_random_state(long n, double box_length, uint64_t seed, double temperature,
                              const double* mass, double* x, double* y, double* z, double* vx,
                              double* vy, double* vz);

/* ---- measurement ---------------------------------------------------------- */

/* When enabled, every launch of the timed kernel groups is bracketed by HIP
 * events on the context's stream. md_timing_read returns the summed device
 * time (ms) and launch count since the last reset, for kernel id 0 = LJ pair
 * kernel, 1 = neighbour build, 2 = integrator kernels. md_timing_enable times
 * all three groups (on != 0) or none; md_timing_select times the groups whose
 * bit (1 << kernel id) is set in mask. Each bracket adds two event records
 * to the stream, a few microseconds of device time per step, and any timing
 * keeps md_run from replaying HIP graphs. (Measurement only: the reference
 * has no counterpart.) */
md_status md_timing_enable(md_ctx* ctx, int on);
md_status md_timing_select(md_ctx* ctx, int mask);
md_status md_timing_read(md_ctx* ctx, int kernel_id, double* total_ms, long* launches);
md_status md_timing_reset(md_ctx* ctx);
/* Candidate-pair statistics of the last force evaluation: pairs evaluated
 * inside the cutoff (each unordered pair once). */
md_status md_last_pair_count(md_ctx* ctx, long* pairs_in_cutoff);

#ifdef __cplusplus
}
#endif

#endif /* MOLDYN_HIP_H */
