// C ABI of the batched-replica path (include/moldyn_hip.h, md_batch_*):
// flat device state for many independent systems, one set of launches per
// step for all of them, HIP-graph-captured multi-step segments, and the
// pieces / lanes-per-atom autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "../../include/moldyn_hip.h"
#include "md_batch.hpp"

struct md_batch {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    int nrep = 0;
    long natoms = 0;
    int nmax = 0;
    int maxcell = 0;
    int block = 512;
    int mixed = 1;
    double dt = 0.0, skin = 0.0, friction = 0.0;
    md_lj_params lj{};
    mdk::LJConst ljc{};
    std::vector<mdk::BatchRep> reps;
    std::vector<long> off;
    mdk::BatchRep* d_reps = nullptr;
    double4 *pos = nullptr, *vel = nullptr, *frc = nullptr, *ref = nullptr;
    int *slot_of = nullptr, *orig = nullptr;
    uint32_t* nbr = nullptr;
    int* ncnt = nullptr;
    int cap = 0;
    // pieces: per replica the density (work estimate) and the largest
    // allowed piece count; the launch-order table on the device
    std::vector<double> rho;
    std::vector<int> max_pieces;
    int piece_slots = 0;
    int npieces = 0;
    int4* d_pieces = nullptr;
    int4* queue = nullptr;
    int* rflag = nullptr;
    double2* ptop = nullptr;
    long long* ppart = nullptr;
    unsigned long long* step = nullptr;  // [0] step counter
    long long* builds = nullptr;
    int* flags = nullptr;                // [0] overflow, [1] fixed-point saturation
    double* thermo = nullptr;
    long thermo_alloc = 0;
    bool primed = false;
};

namespace {

// defaults before tuning: lanes per atom of the force rows, atoms per piece
constexpr int kDefaultTpa = 8;
constexpr int kDefaultPer = 384;

md_status bfail(md_batch* b, md_status s, const std::string& m) {
    if (b) b->err = m;
    return s;
}

#define BHIP(b, call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return bfail(b, MD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
hipError_t dalloc(T*& p, long count) {
    if (p) (void)hipFree(p);
    p = nullptr;
    if (count <= 0) return hipSuccess;
    return hipMalloc(reinterpret_cast<void**>(&p), static_cast<size_t>(count) * sizeof(T));
}

mdk::BatchArgs args_of(md_batch* b, int mode, double* thermo, long thermo_steps) {
    mdk::BatchArgs a{};
    a.reps = b->d_reps;
    a.pieces = b->d_pieces;
    a.queue = b->queue;
    a.rflag = b->rflag;
    a.nrep = b->nrep;
    a.npieces = b->npieces;
    a.nmax = b->nmax;
    a.maxcell = b->maxcell;
    a.mode = mode;
    a.mixed = b->mixed;
    a.pos = b->pos; a.vel = b->vel; a.frc = b->frc; a.ref = b->ref;
    a.slot_of = b->slot_of; a.orig = b->orig;
    a.nbr = b->nbr; a.ncnt = b->ncnt; a.cap = b->cap;
    a.ptop = b->ptop;
    a.ppart = b->ppart;
    a.lj = b->ljc;
    const double rl = b->lj.cutoff + b->skin;
    a.rl2 = rl * rl;
    a.skin = b->skin;
    a.dt = b->dt;
    a.step = b->step;
    a.thermo = thermo;
    a.thermo_steps = thermo_steps > 0 ? thermo_steps : 1;
    a.builds = b->builds;
    a.overflow = b->flags;
    a.fix_sat = b->flags + 1;
    return a;
}

md_status check_flags(md_batch* b) {
    int f[2] = {0, 0};
    BHIP(b, hipStreamSynchronize(b->stream));
    BHIP(b, hipMemcpy(f, b->flags, sizeof(f), hipMemcpyDeviceToHost));
    if (f[0]) return bfail(b, MD_ERR_CAPACITY, "a replica's neighbour row exceeded its capacity");
    if (f[1]) return bfail(b, MD_ERR_RUNTIME, "fixed-point force accumulator saturated (|f| > 2^19)");
    return MD_OK;
}

md_status upload_reps(md_batch* b) {
    BHIP(b, hipMemcpy(b->d_reps, b->reps.data(), b->reps.size() * sizeof(mdk::BatchRep),
                      hipMemcpyHostToDevice));
    return MD_OK;
}

// Splits replica r into reps[r].npiece even ranges and orders all pieces by
// estimated work (atoms x (density + 0.05)), largest first, so the longest
// workgroups start first; the partial arrays keep a replica-contiguous
// numbering (piece0 .. piece0 + npiece - 1).
md_status build_pieces(md_batch* b) {
    struct Item {
        double w;
        int4 pc;
    };
    std::vector<Item> items;
    int p = 0;
    for (int r = 0; r < b->nrep; ++r) {
        mdk::BatchRep& R = b->reps[r];
        R.piece0 = p;
        for (int k = 0; k < R.npiece; ++k) {
            const int s0 = static_cast<int>(static_cast<long>(k) * R.n / R.npiece);
            const int s1 = static_cast<int>(static_cast<long>(k + 1) * R.n / R.npiece);
            items.push_back({(s1 - s0) * (b->rho[r] + 0.05), make_int4(r, s0, s1, p)});
            ++p;
        }
    }
    if (p > b->piece_slots) return bfail(b, MD_ERR_INVALID_ARGUMENT, "too many pieces");
    std::stable_sort(items.begin(), items.end(), [](const Item& x, const Item& y) { return x.w > y.w; });
    std::vector<int4> table(items.size());
    for (size_t k = 0; k < items.size(); ++k) table[k] = items[k].pc;
    b->npieces = p;
    BHIP(b, hipMemcpy(b->d_pieces, table.data(), table.size() * sizeof(int4), hipMemcpyHostToDevice));
    return upload_reps(b);
}

}  // namespace

extern "C" {

// The replica recipe: Rng(master_seed) (mt19937_64 with the reference's
// uniform in (0,1], util/random.hpp:37-48) draws, per replica in order,
// the fcc cell count from {3,...,10} (N = 4 cells^3 in {108, ..., 4000}),
// the density in [0.5, 0.9) and the temperature in [0.7, 1.5); the lattice
// seed is master_seed + r.
md_status md_batch_draw(int count, uint64_t master_seed, int* cells, double* rho, double* T,
                        uint64_t* seeds) {
    if (count < 0 || cells == nullptr || rho == nullptr || T == nullptr || seeds == nullptr)
        return MD_ERR_INVALID_ARGUMENT;
    std::mt19937_64 eng(master_seed);
    std::uniform_real_distribution<double> unif(0.0, 1.0);
    auto uniform = [&]() { return 1.0 - unif(eng); };
    for (int r = 0; r < count; ++r) {
        const int k = std::min(7, static_cast<int>(uniform() * 8.0 - 1e-12));
        cells[r] = 3 + std::max(0, k);
        rho[r] = 0.5 + 0.4 * (1.0 - uniform());
        T[r] = 0.7 + 0.8 * (1.0 - uniform());
        seeds[r] = master_seed + static_cast<uint64_t>(r);
    }
    return MD_OK;
}

md_status md_batch_create(void* hip_stream, int nrep, const long* n, const double* L,
                          const double* x, const double* y, const double* z, const double* vx,
                          const double* vy, const double* vz, const double* T,
                          const uint64_t* seeds, const md_lj_params* lj, double skin, double dt,
                          double friction, int precision, md_batch** out) {
    if (out == nullptr) return MD_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (nrep <= 0 || n == nullptr || L == nullptr || x == nullptr || y == nullptr || z == nullptr ||
        T == nullptr || seeds == nullptr || lj == nullptr || !(dt > 0.0) || !(skin >= 0.0))
        return MD_ERR_INVALID_ARGUMENT;
    md_batch* b = new (std::nothrow) md_batch();
    if (b == nullptr) return MD_ERR_RUNTIME;
    *out = b;
    if (hip_stream) {
        b->stream = static_cast<hipStream_t>(hip_stream);
    } else {
        BHIP(b, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        b->own_stream = true;
    }
    b->nrep = nrep;
    b->dt = dt;
    b->skin = skin;
    b->friction = friction;
    b->lj = *lj;
    b->mixed = precision == MD_MIXED ? 1 : 0;
    {
        mdk::LJConst& k = b->ljc;
        const double s2 = (lj->sigma * lj->sigma) / (lj->cutoff * lj->cutoff);
        const double s6 = s2 * s2 * s2;
        k.sig2 = lj->sigma * lj->sigma;
        k.eps = lj->epsilon;
        k.rc = lj->cutoff;
        k.rc2 = lj->cutoff * lj->cutoff;
        k.u_rc = 4.0 * lj->epsilon * (s6 * s6 - s6);
        k.dudr_rc = -(24.0 * lj->epsilon * (2.0 * s6 * s6 - s6)) / lj->cutoff;
        k.sig2f = static_cast<float>(k.sig2);
        k.epsf = static_cast<float>(k.eps);
        k.rc2f = static_cast<float>(k.rc2);
        k.rcf = static_cast<float>(k.rc);
        k.u_rcf = static_cast<float>(k.u_rc);
        k.dudr_rcf = static_cast<float>(k.dudr_rc);
        k.lfs = lj->linear_force_shift;
        if (lj->linear_force_shift) return bfail(b, MD_ERR_INVALID_ARGUMENT, "batched path: Truncated LJ only");
    }
    const double c1 = std::exp(-friction * dt);
    const double c2 = std::sqrt(1.0 - c1 * c1);
    long total = 0;
    double rho_max = 0.0;
    b->reps.resize(static_cast<size_t>(nrep));
    for (int r = 0; r < nrep; ++r) {
        if (n[r] <= 1) return bfail(b, MD_ERR_INVALID_ARGUMENT, "every replica needs at least 2 atoms");
        mdk::BatchRep& R = b->reps[r];
        R.off = total;
        R.n = static_cast<int>(n[r]);
        R.tpa = kDefaultTpa;
        R.seed_lo = static_cast<uint32_t>(seeds[r]);
        R.seed_hi = static_cast<uint32_t>(seeds[r] >> 32);
        const double Lr[3] = {L[3 * r], L[3 * r + 1], L[3 * r + 2]};
        R.box = mdk::make_box(Lr);
        R.T = T[r];
        R.c1 = c1;
        R.c2 = c2;
        total += n[r];
        b->nmax = std::max(b->nmax, R.n);
        b->rho.push_back(n[r] / (Lr[0] * Lr[1] * Lr[2]));
        rho_max = std::max(rho_max, b->rho.back());
    }
    b->natoms = total;
    const double rl = lj->cutoff + skin;
    long cap = static_cast<long>(std::ceil(2.0 * 4.0 / 3.0 * 3.14159265358979323846 * rl * rl * rl * rho_max + 32.0));
    b->cap = static_cast<int>(std::min<long>((cap + 7) / 8 * 8, b->nmax));
    // each replica's rebuild grid: cells of at least (rc + skin)/2 (with a
    // 1e-9 relative margin, so that rounding cannot put a pair within rc + skin
    // three cells apart), searched 5 x 5 x 5; usable with 5 or more per axis
    for (auto& R : b->reps) {
        bool usable = true;
        for (int d = 0; d < 3; ++d) {
            R.nc[d] = std::max(1, static_cast<int>(std::floor(2.0 * R.box.L[d] / (rl * (1.0 + 1e-9)))));
            R.w[d] = R.box.L[d] / R.nc[d];
            usable = usable && R.nc[d] >= 5;
        }
        R.usable = usable ? 1 : 0;
        if (usable) b->maxcell = std::max(b->maxcell, R.nc[0] * R.nc[1] * R.nc[2]);
    }
    int dev = 0, maxsh = 0;
    BHIP(b, hipGetDevice(&dev));
    BHIP(b, hipDeviceGetAttribute(&maxsh, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    // cell tables when they fit next to the positions, else all-pairs rebuilds
    if (mdk::batch_lds_bytes(b->nmax, b->maxcell) + 2048 > static_cast<size_t>(maxsh)) b->maxcell = 0;
    const size_t lds = mdk::batch_lds_bytes(b->nmax, b->maxcell);
    if (lds + 2048 > static_cast<size_t>(maxsh))
        return bfail(b, MD_ERR_INVALID_ARGUMENT,
                     "largest replica does not fit the workgroup's LDS (" + std::to_string(maxsh) + " B)");
    BHIP(b, mdk::batch_allow_lds(lds));
    // pieces: by default kDefaultPer atoms each; a replica has at most 32
    // pieces of at least 32 atoms
    b->max_pieces.resize(static_cast<size_t>(nrep));
    b->piece_slots = 0;
    for (int r = 0; r < nrep; ++r) {
        const int mp = static_cast<int>(std::max<long>(1, std::min<long>(32, n[r] / 32)));
        const int dp = static_cast<int>((n[r] + kDefaultPer - 1) / kDefaultPer);
        b->max_pieces[r] = mp;
        b->reps[r].npiece = std::max(1, std::min(mp, dp));
        b->piece_slots += mp;
    }
    BHIP(b, dalloc(b->d_reps, nrep));
    BHIP(b, dalloc(b->pos, total)); BHIP(b, dalloc(b->vel, total));
    BHIP(b, dalloc(b->frc, total)); BHIP(b, dalloc(b->ref, total));
    BHIP(b, dalloc(b->slot_of, total)); BHIP(b, dalloc(b->orig, total));
    BHIP(b, dalloc(b->nbr, total * b->cap)); BHIP(b, dalloc(b->ncnt, total));
    BHIP(b, dalloc(b->step, 2)); BHIP(b, dalloc(b->builds, nrep)); BHIP(b, dalloc(b->flags, 2));
    BHIP(b, dalloc(b->d_pieces, b->piece_slots)); BHIP(b, dalloc(b->ptop, b->piece_slots));
    BHIP(b, dalloc(b->queue, b->piece_slots)); BHIP(b, dalloc(b->rflag, nrep));
    BHIP(b, dalloc(b->ppart, 3L * b->piece_slots));
    BHIP(b, hipMemset(b->step, 0, 2 * sizeof(unsigned long long)));
    BHIP(b, hipMemset(b->builds, 0, nrep * sizeof(long long)));
    BHIP(b, hipMemset(b->flags, 0, 2 * sizeof(int)));
    std::vector<double4> p(static_cast<size_t>(total)), v(static_cast<size_t>(total)),
        f(static_cast<size_t>(total), make_double4(0, 0, 0, 0));
    for (long i = 0; i < total; ++i) {
        p[i] = make_double4(x[i], y[i], z[i], 0.0);
        v[i] = make_double4(vx ? vx[i] : 0.0, vy ? vy[i] : 0.0, vz ? vz[i] : 0.0, 1.0);
    }
    BHIP(b, hipMemcpy(b->pos, p.data(), total * sizeof(double4), hipMemcpyHostToDevice));
    BHIP(b, hipMemcpy(b->ref, p.data(), total * sizeof(double4), hipMemcpyHostToDevice));
    BHIP(b, hipMemcpy(b->vel, v.data(), total * sizeof(double4), hipMemcpyHostToDevice));
    BHIP(b, hipMemcpy(b->frc, f.data(), total * sizeof(double4), hipMemcpyHostToDevice));
    // identity slot order until a replica's first cell-sorted build
    std::vector<int> ident(static_cast<size_t>(total));
    for (int r = 0; r < nrep; ++r)
        for (int i = 0; i < n[r]; ++i) ident[static_cast<size_t>(b->reps[r].off + i)] = i;
    BHIP(b, hipMemcpy(b->slot_of, ident.data(), total * sizeof(int), hipMemcpyHostToDevice));
    BHIP(b, hipMemcpy(b->orig, ident.data(), total * sizeof(int), hipMemcpyHostToDevice));
    return build_pieces(b);
}

md_status md_batch_destroy(md_batch* b) {
    if (b == nullptr) return MD_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (void* p : {static_cast<void*>(b->d_reps), static_cast<void*>(b->pos),
                    static_cast<void*>(b->vel), static_cast<void*>(b->frc),
                    static_cast<void*>(b->ref), static_cast<void*>(b->nbr),
                    static_cast<void*>(b->ncnt), static_cast<void*>(b->step),
                    static_cast<void*>(b->builds), static_cast<void*>(b->flags),
                    static_cast<void*>(b->thermo), static_cast<void*>(b->slot_of),
                    static_cast<void*>(b->orig), static_cast<void*>(b->d_pieces),
                    static_cast<void*>(b->ptop), static_cast<void*>(b->ppart),
                    static_cast<void*>(b->queue), static_cast<void*>(b->rflag)})
        if (p) (void)hipFree(p);
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return MD_OK;
}

const char* md_batch_last_error(const md_batch* b) { return b ? b->err.c_str() : "null batch"; }

md_status md_batch_prime(md_batch* b) {
    if (b == nullptr) return MD_ERR_INVALID_ARGUMENT;
    BHIP(b, mdk::launch_batch(args_of(b, 0, nullptr, 0), b->block, b->stream));
    md_status st = check_flags(b);
    if (st == MD_OK) b->primed = true;
    return st;
}

md_status md_batch_set_tpa(md_batch* b, const int* tpa, int block) {
    if (b == nullptr || tpa == nullptr) return MD_ERR_INVALID_ARGUMENT;
    if (block != 256 && block != 512 && block != 1024)
        return bfail(b, MD_ERR_INVALID_ARGUMENT, "block must be 256, 512 or 1024");
    for (int r = 0; r < b->nrep; ++r) {
        const int t = tpa[r];
        if (t != 1 && t != 2 && t != 4 && t != 8 && t != 16)
            return bfail(b, MD_ERR_INVALID_ARGUMENT, "lanes per atom must be 1, 2, 4, 8 or 16");
    }
    BHIP(b, hipStreamSynchronize(b->stream));
    for (int r = 0; r < b->nrep; ++r) b->reps[r].tpa = tpa[r];
    b->block = block;
    return upload_reps(b);
}

md_status md_batch_set_pieces(md_batch* b, const int* pieces) {
    if (b == nullptr || pieces == nullptr) return MD_ERR_INVALID_ARGUMENT;
    for (int r = 0; r < b->nrep; ++r)
        if (pieces[r] < 1 || pieces[r] > b->max_pieces[r])
            return bfail(b, MD_ERR_INVALID_ARGUMENT,
                         "replica " + std::to_string(r) + ": pieces must be in [1, " +
                             std::to_string(b->max_pieces[r]) + "]");
    BHIP(b, hipStreamSynchronize(b->stream));
    for (int r = 0; r < b->nrep; ++r) b->reps[r].npiece = pieces[r];
    return build_pieces(b);
}

md_status md_batch_run(md_batch* b, long nsteps, long segment, int use_graph, double* thermo_out) {
    if (b == nullptr) return MD_ERR_INVALID_ARGUMENT;
    if (!b->primed) return bfail(b, MD_ERR_STATE, "md_batch_prime has not been called");
    if (nsteps <= 0) return MD_OK;
    if (segment <= 0) segment = 100;
    double* th = nullptr;
    long th_steps = 0;
    if (thermo_out) {
        if (nsteps * b->nrep * 4 > b->thermo_alloc) {
            BHIP(b, dalloc(b->thermo, nsteps * b->nrep * 4));
            b->thermo_alloc = nsteps * b->nrep * 4;
        }
        th = b->thermo;
        th_steps = nsteps;
    }
    unsigned long long step0 = 0;
    BHIP(b, hipMemcpy(&step0, b->step, sizeof(step0), hipMemcpyDeviceToHost));
    const mdk::BatchArgs a = args_of(b, 1, th, th_steps);
    long done = 0;
    if (use_graph && nsteps >= segment) {
        // the captured graph and its executable are destroyed on every return
        // path, the early error returns of BHIP included (after the stream
        // has drained, so no launch of them is still pending)
        struct GraphGuard {
            hipStream_t stream;
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            ~GraphGuard() {
                if (exec) {
                    (void)hipStreamSynchronize(stream);
                    (void)hipGraphExecDestroy(exec);
                }
                if (graph) (void)hipGraphDestroy(graph);
            }
        } g{b->stream};
        BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        for (long k = 0; k < segment; ++k) {
            hipError_t e = mdk::launch_batch(a, b->block, b->stream);
            if (e != hipSuccess) {
                (void)hipStreamEndCapture(b->stream, &g.graph);
                return bfail(b, MD_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
            }
        }
        BHIP(b, hipStreamEndCapture(b->stream, &g.graph));
        BHIP(b, hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
        while (done + segment <= nsteps) {
            BHIP(b, hipGraphLaunch(g.exec, b->stream));
            done += segment;
        }
        BHIP(b, hipStreamSynchronize(b->stream));
    }
    for (; done < nsteps; ++done) BHIP(b, mdk::launch_batch(a, b->block, b->stream));
    md_status st = check_flags(b);
    if (st != MD_OK) return st;
    if (thermo_out) {
        // ring slot = step % nsteps: rotate so row k is step step0 + k
        std::vector<double> ring(static_cast<size_t>(nsteps * b->nrep * 4));
        BHIP(b, hipMemcpy(ring.data(), b->thermo, ring.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (long k = 0; k < nsteps; ++k) {
            const long slot = static_cast<long>((step0 + k) % static_cast<unsigned long long>(nsteps));
            std::memcpy(thermo_out + k * b->nrep * 4, ring.data() + slot * b->nrep * 4,
                        b->nrep * 4 * sizeof(double));
        }
    }
    return MD_OK;
}

md_status md_batch_download(md_batch* b, double* x, double* y, double* z, double* vx, double* vy,
                            double* vz, double* fx, double* fy, double* fz) {
    if (b == nullptr) return MD_ERR_INVALID_ARGUMENT;
    const long n = b->natoms;
    std::vector<double4> tmp(static_cast<size_t>(n));
    BHIP(b, hipStreamSynchronize(b->stream));
    double* outs[3][3] = {{x, y, z}, {vx, vy, vz}, {fx, fy, fz}};
    double4* srcs[3] = {b->pos, b->vel, b->frc};
    for (int k = 0; k < 3; ++k) {
        if (!outs[k][0] && !outs[k][1] && !outs[k][2]) continue;
        BHIP(b, hipMemcpy(tmp.data(), srcs[k], n * sizeof(double4), hipMemcpyDeviceToHost));
        for (long i = 0; i < n; ++i) {
            if (outs[k][0]) outs[k][0][i] = tmp[i].x;
            if (outs[k][1]) outs[k][1][i] = tmp[i].y;
            if (outs[k][2]) outs[k][2][i] = tmp[i].z;
        }
    }
    return MD_OK;
}

md_status md_batch_rebuild_counts(md_batch* b, long* counts) {
    if (b == nullptr || counts == nullptr) return MD_ERR_INVALID_ARGUMENT;
    std::vector<long long> c(static_cast<size_t>(b->nrep));
    BHIP(b, hipStreamSynchronize(b->stream));
    BHIP(b, hipMemcpy(c.data(), b->builds, b->nrep * sizeof(long long), hipMemcpyDeviceToHost));
    for (int r = 0; r < b->nrep; ++r) counts[r] = static_cast<long>(c[r]);
    return MD_OK;
}

// Autotuner: for each workgroup size in {256, 512, 1024}, coordinate descent
// over each replica size class (largest first): the class's atoms per piece
// in {128, 192, 256, 384, 512, 768, 1024, whole replica}, then its lanes per
// atom in {1, 2, 4, 8, 16}; each candidate is timed over `trial_steps` steps
// from the same saved state with HIP events. The fastest (block, pieces,
// lanes) is applied and the state restored, so tuning leaves the trajectory
// untouched (fixed-point forces and sums do not depend on the split). The
// trials should span list rebuilds: a rebuilding piece costs several force
// pieces, so tune from a state some steps after md_batch_prime, over a few
// tens of steps.
md_status md_batch_autotune(md_batch* b, int trial_steps, int* tpa_out, int* pieces_out,
                            int* block_out, double* best_ms_per_step) {
    if (b == nullptr) return MD_ERR_INVALID_ARGUMENT;
    if (!b->primed) return bfail(b, MD_ERR_STATE, "md_batch_prime has not been called");
    if (trial_steps < 1) trial_steps = 5;
    const long n = b->natoms;
    // save
    double4 *sp = nullptr, *sv = nullptr, *sf = nullptr, *sr = nullptr;
    uint32_t* snb = nullptr;
    int *sc = nullptr, *sslot = nullptr, *sorig = nullptr;
    unsigned long long sstep[2];
    std::vector<long long> sbuilds(static_cast<size_t>(b->nrep));
    auto release = [&]() {
        for (void* q : {static_cast<void*>(sp), static_cast<void*>(sv), static_cast<void*>(sf),
                        static_cast<void*>(sr), static_cast<void*>(snb), static_cast<void*>(sc),
                        static_cast<void*>(sslot), static_cast<void*>(sorig)})
            if (q) (void)hipFree(q);
    };
    struct Guard {
        std::function<void()> f;
        ~Guard() { f(); }
    } guard{release};
    BHIP(b, dalloc(sp, n)); BHIP(b, dalloc(sv, n)); BHIP(b, dalloc(sf, n)); BHIP(b, dalloc(sr, n));
    BHIP(b, dalloc(snb, n * b->cap)); BHIP(b, dalloc(sc, n));
    BHIP(b, dalloc(sslot, n)); BHIP(b, dalloc(sorig, n));
    BHIP(b, hipStreamSynchronize(b->stream));
    const hipMemcpyKind d2d = hipMemcpyDeviceToDevice;
    BHIP(b, hipMemcpy(sp, b->pos, n * sizeof(double4), d2d));
    BHIP(b, hipMemcpy(sv, b->vel, n * sizeof(double4), d2d));
    BHIP(b, hipMemcpy(sf, b->frc, n * sizeof(double4), d2d));
    BHIP(b, hipMemcpy(sr, b->ref, n * sizeof(double4), d2d));
    BHIP(b, hipMemcpy(snb, b->nbr, n * b->cap * sizeof(uint32_t), d2d));
    BHIP(b, hipMemcpy(sc, b->ncnt, n * sizeof(int), d2d));
    BHIP(b, hipMemcpy(sslot, b->slot_of, n * sizeof(int), d2d));
    BHIP(b, hipMemcpy(sorig, b->orig, n * sizeof(int), d2d));
    BHIP(b, hipMemcpy(sstep, b->step, sizeof(sstep), hipMemcpyDeviceToHost));
    BHIP(b, hipMemcpy(sbuilds.data(), b->builds, b->nrep * sizeof(long long), hipMemcpyDeviceToHost));
    auto restore = [&]() -> md_status {
        BHIP(b, hipMemcpy(b->pos, sp, n * sizeof(double4), d2d));
        BHIP(b, hipMemcpy(b->vel, sv, n * sizeof(double4), d2d));
        BHIP(b, hipMemcpy(b->frc, sf, n * sizeof(double4), d2d));
        BHIP(b, hipMemcpy(b->ref, sr, n * sizeof(double4), d2d));
        BHIP(b, hipMemcpy(b->nbr, snb, n * b->cap * sizeof(uint32_t), d2d));
        BHIP(b, hipMemcpy(b->ncnt, sc, n * sizeof(int), d2d));
        BHIP(b, hipMemcpy(b->slot_of, sslot, n * sizeof(int), d2d));
        BHIP(b, hipMemcpy(b->orig, sorig, n * sizeof(int), d2d));
        BHIP(b, hipMemcpy(b->step, sstep, sizeof(sstep), hipMemcpyHostToDevice));
        BHIP(b, hipMemcpy(b->builds, sbuilds.data(), b->nrep * sizeof(long long), hipMemcpyHostToDevice));
        return MD_OK;
    };
    // size classes, largest first
    std::map<int, std::vector<int>, std::greater<int>> classes;
    for (int r = 0; r < b->nrep; ++r) classes[b->reps[r].n].push_back(r);
    hipEvent_t e0, e1;
    BHIP(b, hipEventCreate(&e0));
    BHIP(b, hipEventCreate(&e1));
    struct Events {
        hipEvent_t a, b;
        ~Events() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } events{e0, e1};
    auto trial = [&](double* ms) -> md_status {
        md_status st = restore();
        if (st != MD_OK) return st;
        st = build_pieces(b);
        if (st != MD_OK) return st;
        const mdk::BatchArgs a = args_of(b, 1, nullptr, 0);
        BHIP(b, hipEventRecord(e0, b->stream));
        for (int k = 0; k < trial_steps; ++k) BHIP(b, mdk::launch_batch(a, b->block, b->stream));
        BHIP(b, hipEventRecord(e1, b->stream));
        BHIP(b, hipEventSynchronize(e1));
        float f = 0.0f;
        BHIP(b, hipEventElapsedTime(&f, e0, e1));
        *ms = f / trial_steps;
        return MD_OK;
    };
    // pieces of a replica for a target of `per` atoms per piece (0: whole)
    auto split = [&](int r, int per) {
        if (per <= 0) return 1;
        const int v = (b->reps[r].n + per - 1) / per;
        return std::max(1, std::min(b->max_pieces[r], v));
    };
    double best = 1e30;
    int best_block = b->block;
    std::vector<int> best_tpa(static_cast<size_t>(b->nrep), kDefaultTpa),
        best_np(static_cast<size_t>(b->nrep), 1);
    const int lanes[5] = {1, 2, 4, 8, 16};
    const int per_opts[8] = {kDefaultPer, 128, 192, 256, 512, 768, 1024, 0};
    for (int block : {256, 512, 1024}) {
        b->block = block;
        for (int r = 0; r < b->nrep; ++r) {
            b->reps[r].tpa = kDefaultTpa;
            b->reps[r].npiece = split(r, kDefaultPer);
        }
        double cur = 0.0;
        md_status st = trial(&cur);
        if (st != MD_OK) return st;
        for (auto& kv : classes) {
            int keep_p = kDefaultPer;
            for (int per : per_opts) {
                if (per == kDefaultPer) continue;
                for (int r : kv.second) b->reps[r].npiece = split(r, per);
                double ms = 0.0;
                st = trial(&ms);
                if (st != MD_OK) return st;
                if (ms < cur) {
                    cur = ms;
                    keep_p = per;
                }
            }
            for (int r : kv.second) b->reps[r].npiece = split(r, keep_p);
            int keep_t = kDefaultTpa;
            for (int t : lanes) {
                if (t == kDefaultTpa) continue;
                for (int r : kv.second) b->reps[r].tpa = t;
                double ms = 0.0;
                st = trial(&ms);
                if (st != MD_OK) return st;
                if (ms < cur) {
                    cur = ms;
                    keep_t = t;
                }
            }
            for (int r : kv.second) b->reps[r].tpa = keep_t;
        }
        if (cur < best) {
            best = cur;
            best_block = block;
            for (int r = 0; r < b->nrep; ++r) {
                best_tpa[r] = b->reps[r].tpa;
                best_np[r] = b->reps[r].npiece;
            }
        }
    }
    b->block = best_block;
    for (int r = 0; r < b->nrep; ++r) {
        b->reps[r].tpa = best_tpa[r];
        b->reps[r].npiece = best_np[r];
    }
    md_status st = restore();
    if (st != MD_OK) return st;
    st = build_pieces(b);
    if (st != MD_OK) return st;
    BHIP(b, hipStreamSynchronize(b->stream));
    if (tpa_out)
        for (int r = 0; r < b->nrep; ++r) tpa_out[r] = best_tpa[r];
    if (pieces_out)
        for (int r = 0; r < b->nrep; ++r) pieces_out[r] = best_np[r];
    if (block_out) *block_out = best_block;
    if (best_ms_per_step) *best_ms_per_step = best;
    return MD_OK;
}

}  // extern "C"
