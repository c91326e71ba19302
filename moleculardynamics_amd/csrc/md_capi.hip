// The C ABI (include/moldyn_hip.h): context, device state, and the host
// orchestration of the kernels in md_neigh.hip, md_force.hip and
// md_integrate.hip. No exception crosses this boundary; failures become
// md_status codes with a message in md_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <string>
#include <utility>
#include <vector>

#include "../../include/moldyn_hip.h"
#include "md_common.hpp"
#include "md_kernels.hpp"

struct md_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;

    long n = 0;
    double L[3] = {0, 0, 0};
    mdk::Box box{};

    double4 *pos = nullptr, *vel = nullptr, *frc = nullptr, *ref = nullptr;
    double4 *pos2 = nullptr, *vel2 = nullptr, *frc2 = nullptr;
    int *orig = nullptr, *slot = nullptr, *orig2 = nullptr, *cell2 = nullptr, *cell_of = nullptr;
    double* stage = nullptr;

    // Verlet list
    bool neigh_configured = false;
    bool list_built = false;
    double rc = 0.0, skin = 0.0;
    mdk::Grid grid{};
    long ncell = 0, ncell_alloc = 0;
    int *cell_count = nullptr, *cell_start = nullptr, *cell_fill = nullptr, *order = nullptr;
    int* tile_sum = nullptr;
    // periodic fixed-point positions of the list (see alloc_neigh): the
    // committed order and the permuted order of a rebuild
    int4 *qpos = nullptr, *qpos2 = nullptr;
    mdk::Quant quant{};
    uint32_t* nbr = nullptr;
    int* ncnt = nullptr;
    long npad = 0;
    int cap = 0, cap_user = 0;
    long nbr_alloc = 0;
    // bricks of list cells for the LDS-staged pair kernels (mdk::Bricks):
    // brick_width 0 off (the default), -1 automatic, else cells per brick
    // side; halo_boost grows the LDS rows after a build found a halo above them
    int brick_width = 0;
    double halo_boost = 1.0;
    mdk::Bricks bk{};
    int *brick_off = nullptr, *brick_tab = nullptr;
    long brick_off_alloc = 0, brick_tab_alloc = 0;

    // cells at the LJ cutoff (MD_PAIRS_CELLS)
    int *ci_cell_of = nullptr, *ci_count = nullptr, *ci_start = nullptr, *ci_fill = nullptr,
        *ci_atoms = nullptr, *ci_tmp = nullptr;
    long ci_ncell_alloc = 0;

    // reductions and device scalars (partial slots for part_alloc workgroups)
    double* part_d = nullptr;
    long long* part_l = nullptr;
    long part_alloc = 0;
    mdk::Scalars* sc = nullptr;
    mdk::Chain* chain = nullptr;
    double* com4 = nullptr;
    // the LJ finish kernel's partials (3 doubles + 3 int64 per finish
    // workgroup) and its arrival ticket, one block
    void* fin = nullptr;

    bool lj_set = false;
    md_lj_params lj{};
    mdk::LJConst ljc{};

    bool integ_set = false;
    md_integrator integ{};
    uint64_t step_counter = 0;
    unsigned long long step_upload = 0;  // source of the device step counter's upload
    bool chain_primed = false;

    md_thermo* thermo_dev = nullptr;
    long thermo_alloc = 0;

    // md_run's step-boundary snapshot (restored after a neighbour-row overflow)
    double4 *snap_pos = nullptr, *snap_vel = nullptr, *snap_frc = nullptr;
    int *snap_orig = nullptr, *snap_slot = nullptr;
    mdk::Scalars* snap_sc = nullptr;
    mdk::Chain* snap_chain = nullptr;
    long snap_alloc = 0;
    uint64_t snap_step = 0;

    int timing = 0;  // bit k: launches of kernel group k bracketed by events
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events[3];
    double ms[3] = {0, 0, 0};
    long launches[3] = {0, 0, 0};
};

namespace {

md_status fail(md_ctx* c, md_status s, const std::string& msg) {
    if (c) c->err = msg;
    return s;
}

#define MD_HIP(ctx, call)                                                                    \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(ctx, MD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
hipError_t alloc(T*& p, long count) {
    if (p) {
        (void)hipFree(p);
        p = nullptr;
    }
    if (count <= 0) return hipSuccess;
    return hipMalloc(reinterpret_cast<void**>(&p), static_cast<size_t>(count) * sizeof(T));
}

template <class T>
void release(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

long blocks_of(long n) { return (n + mdk::kBlock - 1) / mdk::kBlock; }

constexpr size_t kFinBytes = 6 * mdk::kFinishGroups * sizeof(double) + 256;

void set_lj_const(md_ctx* c) {
    const md_lj_params& p = c->lj;
    c->ljc = mdk::lj_constants(p.sigma, p.epsilon, p.cutoff, p.linear_force_shift);
}

mdk::Grid make_grid(const double L[3], double cutoff) { return mdk::cell_grid(L, cutoff); }

// The Verlet rebuild's grid: cells of at least rl/2 (a 1e-9 relative margin
// keeps rounding from putting a pair within rl three cells apart), searched
// 5 x 5 x 5, usable with 5 or more per axis (then every list pair also lies
// within 0.4 L, well inside half a box).
mdk::Grid make_list_grid(const double L[3], double rl) {
    mdk::Grid g{};
    for (int d = 0; d < 3; ++d) {
        g.L[d] = L[d];
        g.nc[d] = std::max(1, static_cast<int>(std::floor(2.0 * L[d] / (rl * (1.0 + 1e-9)))));
        g.w[d] = L[d] / static_cast<double>(g.nc[d]);
        g.inv_w[d] = static_cast<double>(g.nc[d]) / L[d];
    }
    g.usable = (g.nc[0] >= 5 && g.nc[1] >= 5 && g.nc[2] >= 5) ? 1 : 0;
    return g;
}

md_status need_system(md_ctx* c) {
    if (c == nullptr) return MD_ERR_INVALID_ARGUMENT;
    if (c->n <= 0) return fail(c, MD_ERR_STATE, "no system uploaded");
    return MD_OK;
}

md_status read_scalars(md_ctx* c, mdk::Scalars* out) {
    MD_HIP(c, hipMemcpyAsync(out, c->sc, sizeof(mdk::Scalars), hipMemcpyDeviceToHost, c->stream));
    MD_HIP(c, hipStreamSynchronize(c->stream));
    return MD_OK;
}

void ev_begin(md_ctx* c, int id, hipEvent_t* a) {
    if (!((c->timing >> id) & 1)) return;
    (void)hipEventCreate(a);
    (void)hipEventRecord(*a, c->stream);
    (void)id;
}
void ev_end(md_ctx* c, int id, hipEvent_t a) {
    if (!((c->timing >> id) & 1)) return;
    hipEvent_t b;
    (void)hipEventCreate(&b);
    (void)hipEventRecord(b, c->stream);
    c->events[id].emplace_back(a, b);
}

md_status resolve_events(md_ctx* c) {
    for (int id = 0; id < 3; ++id) {
        for (auto& e : c->events[id]) {
            MD_HIP(c, hipEventSynchronize(e.second));
            float ms = 0.0f;
            MD_HIP(c, hipEventElapsedTime(&ms, e.first, e.second));
            c->ms[id] += ms;
            c->launches[id] += 1;
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        c->events[id].clear();
    }
    return MD_OK;
}

// Neighbour row capacity: twice the mean full-list row at the system's
// density plus 32, capped at n - 1, rounded up to a multiple of 8 (the octet
// layout): 192 at rho* = 0.8442 and rc + skin = 2.8 (mean row 77.6).
int capacity_for(long n, const double* L, double rc, double skin) {
    const double rl = rc + skin;
    const double rho = static_cast<double>(n) / (L[0] * L[1] * L[2]);
    const double mean = 4.0 / 3.0 * 3.14159265358979323846 * rl * rl * rl * rho;
    const double want = std::ceil(2.0 * mean + 32.0);
    const long most = std::max(1L, n - 1);
    // compared in FP64 first: a huge density would overflow the conversion
    const long cap = want < static_cast<double>(most) ? static_cast<long>(want) : most;
    return static_cast<int>((cap + 7) / 8 * 8);
}

int auto_capacity(const md_ctx* c) { return capacity_for(c->n, c->L, c->rc, c->skin); }

// Bricks of list cells for the LDS-staged pair kernels (mdk::Bricks), when
// requested with md_neigh_set_bricks. Used for systems above 32,768 atoms
// (below, the pair kernels run several lanes per atom over the whole grid)
// with fixed-point positions, when the grid takes bricks of at most the
// chosen width (4 cells for automatic) with their two-cell halo inside the
// box on every axis (width + 4 <= cells per axis). The LDS rows hold 1.15
// times the expected halo (times halo_boost) plus 64 atoms, at most 1,904
// (the FP64 kernel's 32-byte positions and its static arrays within 64 KB of
// LDS); denser grids keep slot entries. Waves per workgroup: the expected
// own atoms (times 4 lanes for FP64) over 64, plus 2 %.
constexpr long kBrickMinAtoms = 32769;
constexpr int kBrickAutoWidth = 4;
constexpr double kBrickMaxHalo = 1904.0;

md_status setup_bricks(md_ctx* c, bool fixed_point) {
    c->bk = mdk::Bricks{};
    if (!fixed_point || c->n < kBrickMinAtoms || c->brick_width == 0) return MD_OK;
    const int want = c->brick_width > 0 ? c->brick_width : kBrickAutoWidth;
    mdk::Bricks k{};
    int wmax[3];
    for (int d = 0; d < 3; ++d) {
        const int nc = c->grid.nc[d];
        k.kb[d] = std::max(1, (nc + want - 1) / want);  // no brick wider than want
        wmax[d] = (nc + k.kb[d] - 1) / k.kb[d];
        if (wmax[d] + 4 > nc) return MD_OK;  // the halo would wrap onto itself
    }
    if (wmax[1] * wmax[2] > 64) return MD_OK;  // own rows of one wave
    if ((wmax[1] + 4) * (wmax[2] + 4) > 160) return MD_OK;  // halo rows of the copy plan
    const double per_cell = static_cast<double>(c->n) / static_cast<double>(c->ncell);
    const double halo = per_cell * (wmax[0] + 4) * (wmax[1] + 4) * (wmax[2] + 4);
    const double rows = std::ceil(1.15 * c->halo_boost * halo) + 64.0;
    if (rows > kBrickMaxHalo) return MD_OK;
    k.halo_cap = static_cast<int>(rows);
    const double own = per_cell * wmax[0] * wmax[1] * wmax[2];
    k.waves_mixed = std::min(8, std::max(1, static_cast<int>(std::ceil(1.02 * own / 64.0))));
    k.waves_fp64 = std::min(16, std::max(1, static_cast<int>(std::ceil(1.02 * 4.0 * own / 64.0))));
    k.nb = k.kb[0] * k.kb[1] * k.kb[2];
    std::vector<int> off(static_cast<size_t>(k.nb) + 1);
    long tot = 0;
    for (int b = 0; b < k.nb; ++b) {
        const mdk::BrickGeo o = mdk::brick_geo(k, c->grid, b);
        off[b] = static_cast<int>(tot);
        tot += static_cast<long>(o.ex + 4) * (o.ey + 4) * (o.ez + 4) + 1;
        if (tot >= (1L << 31)) return MD_OK;
    }
    off[k.nb] = static_cast<int>(tot);
    if (k.nb + 1 > c->brick_off_alloc) {
        MD_HIP(c, alloc(c->brick_off, k.nb + 1));
        c->brick_off_alloc = k.nb + 1;
    }
    if (tot > c->brick_tab_alloc) {
        MD_HIP(c, alloc(c->brick_tab, tot));
        c->brick_tab_alloc = tot;
    }
    MD_HIP(c, hipMemcpy(c->brick_off, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
    if (k.nb > c->part_alloc) {  // one partial slot per brick workgroup
        MD_HIP(c, alloc(c->part_d, 4L * k.nb + 8));
        MD_HIP(c, alloc(c->part_l, 3L * k.nb + 8));
        c->part_alloc = k.nb;
    }
    k.tab_off = c->brick_off;
    k.tab = c->brick_tab;
    k.on = 1;
    c->bk = k;
    return MD_OK;
}

md_status alloc_neigh(md_ctx* c) {
    c->grid = make_list_grid(c->L, c->rc + c->skin);
    c->ncell = c->grid.usable ? static_cast<long>(c->grid.nc[0]) * c->grid.nc[1] * c->grid.nc[2] : 1;
    if (c->ncell > c->ncell_alloc) {
        MD_HIP(c, alloc(c->cell_count, c->ncell));
        // the rebuild kernels keep cell_count all-zero between builds
        MD_HIP(c, hipMemset(c->cell_count, 0, c->ncell * sizeof(int)));
        MD_HIP(c, alloc(c->cell_start, c->ncell + 1));
        MD_HIP(c, alloc(c->cell_fill, c->ncell));
        MD_HIP(c, alloc(c->tile_sum, mdk::scan_tiles(c->ncell)));
        c->ncell_alloc = c->ncell;
    }
    // Periodic fixed-point positions (unit L / 2^32 per axis) are kept when
    // the cell grid is usable (every list pair then lies well inside half a
    // box, where the int32 wrap is the minimum image) and the unit is at most
    // 2^-25 rc, a quarter of FP32's relative precision at the cutoff
    // (L <= 128 rc). The decision bands bound the r2 error: per axis the
    // difference is within one unit u, so |r2' - r2| <= 2 sqrt(3) r u + 3 u^2.
    double umax = 0.0;
    for (int d = 0; d < 3; ++d) umax = std::max(umax, c->L[d] / 4294967296.0);
    // (and n < 2^28: the pair kernel forms 32-bit byte offsets of partners)
    const bool fixed_point = c->grid.usable && c->rc > 0.0 && umax <= c->rc / 33554432.0 &&
                             c->n < (1L << 28);
    if (fixed_point) {
        MD_HIP(c, alloc(c->qpos, c->n));
        MD_HIP(c, alloc(c->qpos2, c->n));
        for (int d = 0; d < 3; ++d) {
            c->quant.inv[d] = 4294967296.0 / c->L[d];
            c->quant.unit[d] = c->L[d] / 4294967296.0;
        }
        const double rl = c->rc + c->skin;
        c->quant.band_l = 4.0 * umax / rl + 4.0 * (umax / rl) * (umax / rl) + 1e-12;
        c->quant.band_c = 4.0 * umax / c->rc + 4.0 * (umax / c->rc) * (umax / c->rc) + 1e-12;
        c->quant.cubic = (c->quant.unit[0] == c->quant.unit[1] && c->quant.unit[0] == c->quant.unit[2]) ? 1 : 0;
    } else {
        release(c->qpos);
        release(c->qpos2);
        c->quant = mdk::Quant{};
    }
    md_status bst = setup_bricks(c, fixed_point);
    if (bst != MD_OK) return bst;
    c->npad = (c->n + 63) / 64 * 64;
    c->cap = c->cap_user > 0 ? (c->cap_user + 7) / 8 * 8 : auto_capacity(c);
    const long need = static_cast<long>(c->cap) * c->npad;
    if (need > c->nbr_alloc) {
        MD_HIP(c, alloc(c->nbr, need));
        c->nbr_alloc = need;
    }
    return MD_OK;
}

mdk::NeighArgs neigh_args(md_ctx* c) {
    mdk::NeighArgs a{};
    a.n = c->n;
    a.grid = c->grid;
    a.bk = c->bk;
    a.box = c->box;
    a.rl2 = (c->rc + c->skin) * (c->rc + c->skin);
    a.pos = c->pos; a.vel = c->vel; a.frc = c->frc; a.ref = c->ref;
    a.orig = c->orig; a.slot = c->slot;
    a.pos2 = c->pos2; a.vel2 = c->vel2; a.frc2 = c->frc2;
    a.orig2 = c->orig2; a.cell2 = c->cell2;
    a.cell_of = c->cell_of; a.cell_count = c->cell_count; a.cell_start = c->cell_start;
    a.cell_fill = c->cell_fill; a.order = c->order; a.tile_sum = c->tile_sum;
    a.qpos = c->qpos; a.qpos2 = c->qpos2; a.quant = c->quant;
    a.nbr = c->nbr; a.ncnt = c->ncnt; a.npad = c->npad; a.cap = c->cap;
    a.sc = c->sc;
    a.flag = &c->sc->rebuild_flag;
    return a;
}

mdk::ForceArgs force_args(md_ctx* c, int precision) {
    mdk::ForceArgs a{};
    a.n = c->n;
    a.box = c->box;
    a.lj = c->ljc;
    a.pos = c->pos; a.vel = c->vel; a.frc = c->frc; a.orig = c->orig;
    a.nbr = c->nbr; a.ncnt = c->ncnt; a.npad = c->npad; a.cap = c->cap;
    a.mixed = precision == MD_MIXED ? 1 : 0;
    // the mixed kernel reads the quantised positions when the list keeps them
    a.qpos = (a.mixed && c->qpos != nullptr) ? c->qpos : nullptr;
    a.quant = c->quant;
    a.grid = c->grid;
    a.bk = c->bk;
    a.cell_start = c->cell_start;
    // mixed-term constants: lengths in fixed-point units on a cubic box with
    // fixed-point positions, real lengths otherwise
    const double s = (c->qpos != nullptr && c->quant.cubic) ? c->quant.unit[0] : 1.0;
    a.mq = mdk::mix_constants(c->ljc, s, c->qpos != nullptr ? c->quant.band_c : 0.0);
    a.mx = mdk::mix_constants(c->ljc, 1.0, 0.0);
    a.part_d = c->part_d; a.part_l = c->part_l;
    a.fin_d = static_cast<double*>(c->fin);
    a.fin_l = reinterpret_cast<long long*>(a.fin_d + 3 * mdk::kFinishGroups);
    a.fin_count = reinterpret_cast<unsigned int*>(a.fin_l + 3 * mdk::kFinishGroups);
    a.sc = c->sc;
    return a;
}

mdk::StepArgs step_args(md_ctx* c) {
    mdk::StepArgs a{};
    a.n = c->n;
    a.box = c->box;
    a.pos = c->pos; a.vel = c->vel; a.frc = c->frc; a.ref = c->ref; a.orig = c->orig;
    a.kind = c->integ.kind;
    a.dt = c->integ.dt;
    a.c1 = std::exp(-c->integ.friction * c->integ.dt);
    a.c2 = std::sqrt(1.0 - a.c1 * a.c1);
    a.temperature = c->integ.temperature;
    a.seed = c->integ.seed;
    a.skin = c->skin;
    a.check = 1;
    a.part_d = c->part_d;
    a.sc = c->sc;
    return a;
}

// Builds the list now (flag forced to 1), growing the row capacity until no
// row overflows.
md_status build_now(md_ctx* c) {
    for (int attempt = 0; attempt < 4; ++attempt) {
        mdk::launch_set_flag(&c->sc->rebuild_flag, 1, c->stream);
        hipEvent_t e0{};
        ev_begin(c, 1, &e0);
        mdk::launch_rebuild(neigh_args(c), c->stream);
        ev_end(c, 1, e0);
        MD_HIP(c, hipGetLastError());
        mdk::Scalars s{};
        md_status st = read_scalars(c, &s);
        if (st != MD_OK) return st;
        if (!s.overflow) {
            c->list_built = true;
            return MD_OK;
        }
        // grow and retry (the build counter is restored so one build counts once):
        // bit 1 a row longer than the capacity, bit 2 a brick halo above the LDS rows
        int zero = 0;
        long long builds = s.builds - 1;
        MD_HIP(c, hipMemcpyAsync(&c->sc->overflow, &zero, sizeof(int), hipMemcpyHostToDevice, c->stream));
        MD_HIP(c, hipMemcpyAsync(&c->sc->builds, &builds, sizeof(long long), hipMemcpyHostToDevice, c->stream));
        if (s.overflow & 1) c->cap_user = std::min<long>(2L * c->cap, std::max(1L, c->n - 1));
        if (s.overflow & 2) c->halo_boost *= 1.5;
        st = alloc_neigh(c);
        if (st != MD_OK) return st;
    }
    return fail(c, MD_ERR_CAPACITY, "neighbour row capacity could not be satisfied");
}

// Clears the device flags a call reports (fixed-point saturation, row
// overflow), stream-ordered, so that each call reports only its own.
void clear_call_flags(md_ctx* c) {
    mdk::launch_set_flag(&c->sc->fix_saturated, 0, c->stream);
    mdk::launch_set_flag(&c->sc->overflow, 0, c->stream);
}

// The state md_run starts from: positions, velocities, forces, the slot
// permutation, the chain and the device scalars (stream-ordered copies).
md_status save_snapshot(md_ctx* c) {
    const long n = c->n;
    if (c->snap_alloc < n) {
        MD_HIP(c, alloc(c->snap_pos, n)); MD_HIP(c, alloc(c->snap_vel, n));
        MD_HIP(c, alloc(c->snap_frc, n)); MD_HIP(c, alloc(c->snap_orig, n));
        MD_HIP(c, alloc(c->snap_slot, n));
        c->snap_alloc = n;
    }
    if (c->snap_sc == nullptr) MD_HIP(c, alloc(c->snap_sc, 1));
    if (c->snap_chain == nullptr) MD_HIP(c, alloc(c->snap_chain, 1));
    const hipMemcpyKind dd = hipMemcpyDeviceToDevice;
    MD_HIP(c, hipMemcpyAsync(c->snap_pos, c->pos, n * sizeof(double4), dd, c->stream));
    MD_HIP(c, hipMemcpyAsync(c->snap_vel, c->vel, n * sizeof(double4), dd, c->stream));
    MD_HIP(c, hipMemcpyAsync(c->snap_frc, c->frc, n * sizeof(double4), dd, c->stream));
    MD_HIP(c, hipMemcpyAsync(c->snap_orig, c->orig, n * sizeof(int), dd, c->stream));
    MD_HIP(c, hipMemcpyAsync(c->snap_slot, c->slot, n * sizeof(int), dd, c->stream));
    MD_HIP(c, hipMemcpyAsync(c->snap_sc, c->sc, sizeof(mdk::Scalars), dd, c->stream));
    MD_HIP(c, hipMemcpyAsync(c->snap_chain, c->chain, sizeof(mdk::Chain), dd, c->stream));
    c->snap_step = c->step_counter;
    return MD_OK;
}

md_status restore_snapshot(md_ctx* c) {
    const long n = c->n;
    const hipMemcpyKind dd = hipMemcpyDeviceToDevice;
    MD_HIP(c, hipMemcpyAsync(c->pos, c->snap_pos, n * sizeof(double4), dd, c->stream));
    MD_HIP(c, hipMemcpyAsync(c->vel, c->snap_vel, n * sizeof(double4), dd, c->stream));
    MD_HIP(c, hipMemcpyAsync(c->frc, c->snap_frc, n * sizeof(double4), dd, c->stream));
    MD_HIP(c, hipMemcpyAsync(c->orig, c->snap_orig, n * sizeof(int), dd, c->stream));
    MD_HIP(c, hipMemcpyAsync(c->slot, c->snap_slot, n * sizeof(int), dd, c->stream));
    MD_HIP(c, hipMemcpyAsync(c->sc, c->snap_sc, sizeof(mdk::Scalars), dd, c->stream));
    MD_HIP(c, hipMemcpyAsync(c->chain, c->snap_chain, sizeof(mdk::Chain), dd, c->stream));
    c->step_counter = c->snap_step;
    return MD_OK;
}

}  // namespace

namespace mdk {
bool ctx_view(md_ctx* c, CtxView* v) {
    if (c == nullptr || c->n <= 0) return false;
    v->n = c->n;
    v->box = c->box;
    for (int d = 0; d < 3; ++d) v->L[d] = c->L[d];
    v->pos = c->pos;
    v->vel = c->vel;
    v->orig = c->orig;
    v->slot = c->slot;
    v->stream = c->stream;
    v->device = c->device;
    return true;
}
}  // namespace mdk

extern "C" {

md_status md_ctx_create(void* hip_stream, md_ctx** out) {
    if (out == nullptr) return MD_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    md_ctx* c = new (std::nothrow) md_ctx();
    if (c == nullptr) return MD_ERR_RUNTIME;
    if (hipGetDevice(&c->device) != hipSuccess) {
        delete c;
        return MD_ERR_HIP;
    }
    if (hip_stream != nullptr) {
        c->stream = static_cast<hipStream_t>(hip_stream);
    } else {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            return MD_ERR_HIP;
        }
        c->own_stream = true;
    }
    if (hipMalloc(reinterpret_cast<void**>(&c->sc), sizeof(mdk::Scalars)) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&c->chain), sizeof(mdk::Chain)) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&c->com4), 8 * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->fin, kFinBytes) != hipSuccess) {
        md_ctx_destroy(c);
        return MD_ERR_HIP;
    }
    mdk::Scalars s{};
    s.vscale = 1.0;
    double zeros[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (hipMemcpy(c->sc, &s, sizeof(s), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->com4, zeros, sizeof(zeros), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(c->fin, 0, kFinBytes) != hipSuccess) {
        md_ctx_destroy(c);
        return MD_ERR_HIP;
    }
    *out = c;
    return MD_OK;
}

md_status md_ctx_destroy(md_ctx* c) {
    if (c == nullptr) return MD_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& v : c->events)
        for (auto& e : v) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    release(c->pos); release(c->vel); release(c->frc); release(c->ref);
    release(c->pos2); release(c->vel2); release(c->frc2);
    release(c->orig); release(c->slot); release(c->orig2); release(c->cell2); release(c->cell_of);
    release(c->stage);
    release(c->cell_count); release(c->cell_start); release(c->cell_fill); release(c->order);
    release(c->tile_sum);
    release(c->qpos); release(c->qpos2);
    release(c->nbr); release(c->ncnt);
    release(c->brick_off); release(c->brick_tab);
    release(c->ci_cell_of); release(c->ci_count); release(c->ci_start); release(c->ci_fill);
    release(c->ci_atoms); release(c->ci_tmp);
    release(c->part_d); release(c->part_l); release(c->sc); release(c->chain); release(c->com4);
    if (c->fin) (void)hipFree(c->fin);
    c->fin = nullptr;
    release(c->thermo_dev);
    release(c->snap_pos); release(c->snap_vel); release(c->snap_frc);
    release(c->snap_orig); release(c->snap_slot); release(c->snap_sc); release(c->snap_chain);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return MD_OK;
}

const char* md_last_error(const md_ctx* c) {
    if (c != nullptr) return c->err.c_str();
    const char* e = mdk::noalloc_error();
    return e != nullptr ? e : "null context";
}

md_status md_ctx_synchronize(md_ctx* c) {
    if (c == nullptr) return MD_ERR_INVALID_ARGUMENT;
    MD_HIP(c, hipStreamSynchronize(c->stream));
    return MD_OK;
}

md_status md_system_upload(md_ctx* c, long n, const double* L, const double* x, const double* y,
                           const double* z, const double* vx, const double* vy, const double* vz,
                           const double* m) {
    if (c == nullptr) return MD_ERR_INVALID_ARGUMENT;
    if (n <= 0 || n >= mdk::kMaxListAtoms)
        return fail(c, MD_ERR_INVALID_ARGUMENT, "atom count out of range (1 .. 2^28 - 1)");
    if (L == nullptr || x == nullptr || y == nullptr || z == nullptr)
        return fail(c, MD_ERR_INVALID_ARGUMENT, "box and positions are required");
    for (int d = 0; d < 3; ++d)
        if (!(L[d] > 0.0)) return fail(c, MD_ERR_INVALID_ARGUMENT, "box lengths must be positive");
    MD_HIP(c, hipStreamSynchronize(c->stream));
    c->n = n;
    for (int d = 0; d < 3; ++d) c->L[d] = L[d];
    c->box = mdk::make_box(L);
    MD_HIP(c, alloc(c->pos, n)); MD_HIP(c, alloc(c->vel, n)); MD_HIP(c, alloc(c->frc, n));
    MD_HIP(c, alloc(c->ref, n)); MD_HIP(c, alloc(c->pos2, n)); MD_HIP(c, alloc(c->vel2, n));
    MD_HIP(c, alloc(c->frc2, n)); MD_HIP(c, alloc(c->orig, n)); MD_HIP(c, alloc(c->slot, n));
    MD_HIP(c, alloc(c->orig2, n)); MD_HIP(c, alloc(c->cell2, n)); MD_HIP(c, alloc(c->cell_of, n));
    MD_HIP(c, alloc(c->order, n)); MD_HIP(c, alloc(c->ncnt, n));
    MD_HIP(c, alloc(c->stage, 9 * n));
    const long nb = std::max(blocks_of(n), mdk::lj_partials(n));
    MD_HIP(c, alloc(c->part_d, 4 * nb + 8));
    MD_HIP(c, alloc(c->part_l, 3 * nb + 8));
    c->part_alloc = nb;

    std::vector<double4> p(static_cast<size_t>(n)), v(static_cast<size_t>(n)), f(static_cast<size_t>(n));
    std::vector<int> id(static_cast<size_t>(n));
    for (long i = 0; i < n; ++i) {
        p[i] = make_double4(x[i], y[i], z[i], static_cast<double>(i));  // w: upload index
        v[i] = make_double4(vx ? vx[i] : 0.0, vy ? vy[i] : 0.0, vz ? vz[i] : 0.0, m ? m[i] : 1.0);
        f[i] = make_double4(0.0, 0.0, 0.0, 0.0);
        id[i] = static_cast<int>(i);
    }
    MD_HIP(c, hipMemcpy(c->pos, p.data(), n * sizeof(double4), hipMemcpyHostToDevice));
    MD_HIP(c, hipMemcpy(c->ref, p.data(), n * sizeof(double4), hipMemcpyHostToDevice));
    MD_HIP(c, hipMemcpy(c->vel, v.data(), n * sizeof(double4), hipMemcpyHostToDevice));
    MD_HIP(c, hipMemcpy(c->frc, f.data(), n * sizeof(double4), hipMemcpyHostToDevice));
    MD_HIP(c, hipMemcpy(c->orig, id.data(), n * sizeof(int), hipMemcpyHostToDevice));
    MD_HIP(c, hipMemcpy(c->slot, id.data(), n * sizeof(int), hipMemcpyHostToDevice));
    mdk::Scalars s{};
    s.vscale = 1.0;
    MD_HIP(c, hipMemcpy(c->sc, &s, sizeof(s), hipMemcpyHostToDevice));
    c->list_built = false;
    c->chain_primed = false;
    if (c->neigh_configured) {
        md_status st = alloc_neigh(c);
        if (st != MD_OK) return st;
    }
    return MD_OK;
}

md_status md_system_download(md_ctx* c, double* x, double* y, double* z, double* vx, double* vy,
                             double* vz, double* fx, double* fy, double* fz) {
    md_status st = need_system(c);
    if (st != MD_OK) return st;
    const long n = c->n;
    double* s = c->stage;
    mdk::launch_unpermute(n, c->orig, c->pos, s, s + n, s + 2 * n, c->stream);
    mdk::launch_unpermute(n, c->orig, c->vel, s + 3 * n, s + 4 * n, s + 5 * n, c->stream);
    mdk::launch_unpermute(n, c->orig, c->frc, s + 6 * n, s + 7 * n, s + 8 * n, c->stream);
    MD_HIP(c, hipGetLastError());
    double* dst[9] = {x, y, z, vx, vy, vz, fx, fy, fz};
    for (int k = 0; k < 9; ++k)
        if (dst[k])
            MD_HIP(c, hipMemcpyAsync(dst[k], s + k * n, n * sizeof(double), hipMemcpyDeviceToHost,
                                     c->stream));
    MD_HIP(c, hipStreamSynchronize(c->stream));
    return MD_OK;
}

md_status md_system_set_positions(md_ctx* c, const double* x, const double* y, const double* z) {
    md_status st = need_system(c);
    if (st != MD_OK) return st;
    const long n = c->n;
    double* s = c->stage;
    MD_HIP(c, hipMemcpyAsync(s, x, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    MD_HIP(c, hipMemcpyAsync(s + n, y, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    MD_HIP(c, hipMemcpyAsync(s + 2 * n, z, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    mdk::launch_permute_in(n, c->orig, s, s + n, s + 2 * n, c->pos, mdk::kPermuteWIndex, c->stream);
    MD_HIP(c, hipGetLastError());
    MD_HIP(c, hipStreamSynchronize(c->stream));
    return MD_OK;
}

md_status md_system_set_velocities(md_ctx* c, const double* vx, const double* vy, const double* vz) {
    md_status st = need_system(c);
    if (st != MD_OK) return st;
    const long n = c->n;
    double* s = c->stage;
    MD_HIP(c, hipMemcpyAsync(s, vx, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    MD_HIP(c, hipMemcpyAsync(s + n, vy, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    MD_HIP(c, hipMemcpyAsync(s + 2 * n, vz, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    mdk::launch_permute_in(n, c->orig, s, s + n, s + 2 * n, c->vel, mdk::kPermuteWKeep, c->stream);
    MD_HIP(c, hipGetLastError());
    MD_HIP(c, hipStreamSynchronize(c->stream));
    return MD_OK;
}

md_status md_system_set_forces(md_ctx* c, const double* fx, const double* fy, const double* fz) {
    md_status st = need_system(c);
    if (st != MD_OK) return st;
    const long n = c->n;
    double* s = c->stage;
    MD_HIP(c, hipMemcpyAsync(s, fx, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    MD_HIP(c, hipMemcpyAsync(s + n, fy, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    MD_HIP(c, hipMemcpyAsync(s + 2 * n, fz, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    mdk::launch_permute_in(n, c->orig, s, s + n, s + 2 * n, c->frc, mdk::kPermuteWZero, c->stream);
    MD_HIP(c, hipGetLastError());
    MD_HIP(c, hipStreamSynchronize(c->stream));
    return MD_OK;
}

md_status md_system_kinetic_energy(md_ctx* c, double* ke) {
    md_status st = need_system(c);
    if (st != MD_OK) return st;
    mdk::launch_kinetic(c->n, c->vel, c->part_d, c->sc, c->stream);
    MD_HIP(c, hipGetLastError());
    mdk::Scalars s{};
    st = read_scalars(c, &s);
    if (st != MD_OK) return st;
    if (ke) *ke = s.kinetic;
    return MD_OK;
}

md_status md_system_remove_com_motion(md_ctx* c) {
    md_status st = need_system(c);
    if (st != MD_OK) return st;
    mdk::launch_remove_com(c->n, c->vel, c->part_d, c->com4, c->stream);
    MD_HIP(c, hipGetLastError());
    MD_HIP(c, hipStreamSynchronize(c->stream));
    return MD_OK;
}

md_status md_system_size(const md_ctx* c, long* n) {
    if (c == nullptr || n == nullptr) return MD_ERR_INVALID_ARGUMENT;
    *n = c->n;
    return MD_OK;
}

md_status md_lj_set(md_ctx* c, const md_lj_params* p) {
    if (c == nullptr || p == nullptr) return MD_ERR_INVALID_ARGUMENT;
    if (!(p->cutoff > 0.0) || !(p->sigma > 0.0))
        return fail(c, MD_ERR_INVALID_ARGUMENT, "sigma and cutoff must be positive");
    c->lj = *p;
    set_lj_const(c);
    c->lj_set = true;
    return MD_OK;
}

md_status md_lj_forces(md_ctx* c, int source, int precision, double* energy, double* virial) {
    md_status st = need_system(c);
    if (st != MD_OK) return st;
    if (!c->lj_set) return fail(c, MD_ERR_STATE, "md_lj_set has not been called");
    clear_call_flags(c);
    if (source == MD_PAIRS_VERLET) {
        if (!c->list_built) return fail(c, MD_ERR_STATE, "the Verlet list has not been built");
        mdk::ForceArgs a = force_args(c, precision);
        if (a.qpos != nullptr)
            mdk::launch_quantize(c->n, c->quant, c->pos, c->orig, c->qpos, c->stream);
        hipEvent_t e0{};
        ev_begin(c, 0, &e0);
        mdk::launch_lj_force(a, c->stream);
        ev_end(c, 0, e0);
        mdk::launch_lj_finish(a, c->stream);
    } else if (source == MD_PAIRS_BRUTE || source == MD_PAIRS_CELLS) {
        mdk::DirectArgs a{};
        a.n = c->n;
        a.box = c->box;
        a.lj = c->ljc;
        a.pos = c->pos;
        a.frc = c->frc;
        a.orig = c->orig;
        a.part_d = c->part_d;
        a.part_l = c->part_l;
        a.sc = c->sc;
        if (source == MD_PAIRS_CELLS) {
            const mdk::Grid g = make_grid(c->L, c->lj.cutoff);
            if (!g.usable)
                return fail(c, MD_ERR_INVALID_ARGUMENT,
                            "CellList is not usable (fewer than 3 cells per side); fall back to "
                            "the brute-force overload");
            const long m = static_cast<long>(g.nc[0]) * g.nc[1] * g.nc[2];
            if (m > c->ci_ncell_alloc) {
                MD_HIP(c, alloc(c->ci_count, m));
                MD_HIP(c, alloc(c->ci_start, m + 1));
                MD_HIP(c, alloc(c->ci_fill, m));
                c->ci_ncell_alloc = m;
            }
            MD_HIP(c, alloc(c->ci_cell_of, c->n));
            MD_HIP(c, alloc(c->ci_atoms, c->n));
            MD_HIP(c, alloc(c->ci_tmp, c->n));
            mdk::launch_bin_index(c->n, c->pos, g, c->ci_cell_of, c->ci_count, c->ci_start,
                                  c->ci_fill, c->ci_tmp, c->ci_atoms, nullptr, c->stream);
            a.grid = g;
            a.cell_of = c->ci_cell_of;
            a.cell_start = c->ci_start;
            a.cell_atoms = c->ci_atoms;
        }
        mdk::launch_lj_direct(a, c->stream);
    } else {
        return fail(c, MD_ERR_INVALID_ARGUMENT, "unknown pair source");
    }
    MD_HIP(c, hipGetLastError());
    mdk::Scalars s{};
    st = read_scalars(c, &s);
    if (st != MD_OK) return st;
    if (s.fix_saturated)
        return fail(c, MD_ERR_RUNTIME, "fixed-point force accumulator saturated (|f| > 2^19)");
    if (energy) *energy = s.energy;
    if (virial) *virial = s.virial;
    return MD_OK;
}

md_status md_lj_energy_virial(md_ctx* c, double* energy, double* virial) {
    md_status st = need_system(c);
    if (st != MD_OK) return st;
    if (!c->lj_set) return fail(c, MD_ERR_STATE, "md_lj_set has not been called");
    mdk::DirectArgs a{};
    a.n = c->n;
    a.box = c->box;
    a.lj = c->ljc;
    a.pos = c->pos;
    a.frc = nullptr;
    a.orig = c->orig;
    a.part_d = c->part_d;
    a.part_l = c->part_l;
    a.sc = c->sc;
    mdk::launch_lj_direct(a, c->stream);
    MD_HIP(c, hipGetLastError());
    mdk::Scalars s{};
    st = read_scalars(c, &s);
    if (st != MD_OK) return st;
    if (energy) *energy = s.energy;
    if (virial) *virial = s.virial;
    return MD_OK;
}

md_status md_lj_tail(const md_lj_params* p, long n, const double* L, double* u_lrc, double* p_lrc) {
    if (p == nullptr || L == nullptr) return MD_ERR_INVALID_ARGUMENT;
    double u = 0.0, pr = 0.0;
    if (!p->linear_force_shift) {
        const double kPi = 3.14159265358979323846;
        const double nn = static_cast<double>(n);
        const double rho = nn / (L[0] * L[1] * L[2]);
        const double t = p->sigma / p->cutoff;
        const double t3 = t * t * t;
        const double t9 = t3 * t3 * t3;
        const double sig3 = p->sigma * p->sigma * p->sigma;
        u = (8.0 / 3.0) * kPi * nn * rho * p->epsilon * sig3 * ((1.0 / 3.0) * t9 - t3);
        pr = (16.0 / 3.0) * kPi * rho * rho * p->epsilon * sig3 * ((2.0 / 3.0) * t9 - t3);
    }
    if (u_lrc) *u_lrc = u;
    if (p_lrc) *p_lrc = pr;
    return MD_OK;
}

int md_cells_usable(const double* L, double cutoff) {
    if (L == nullptr || !(cutoff > 0.0)) return 0;
    return make_grid(L, cutoff).usable;
}

md_status md_neigh_config(md_ctx* c, double cutoff, double skin) {
    if (c == nullptr) return MD_ERR_INVALID_ARGUMENT;
    if (!(cutoff > 0.0) || !(skin >= 0.0))
        return fail(c, MD_ERR_INVALID_ARGUMENT, "cutoff must be positive and skin non-negative");
    c->rc = cutoff;
    c->skin = skin;
    c->neigh_configured = true;
    c->list_built = false;
    if (c->n > 0) return alloc_neigh(c);
    return MD_OK;
}

md_status md_neigh_build(md_ctx* c) {
    md_status st = need_system(c);
    if (st != MD_OK) return st;
    if (!c->neigh_configured) return fail(c, MD_ERR_STATE, "md_neigh_config has not been called");
    return build_now(c);
}

md_status md_neigh_needs_rebuild(md_ctx* c, int* flag) {
    md_status st = need_system(c);
    if (st != MD_OK) return st;
    if (!c->list_built) return fail(c, MD_ERR_STATE, "the Verlet list has not been built");
    mdk::StepArgs a = step_args(c);
    mdk::launch_check(a, c->stream);
    MD_HIP(c, hipGetLastError());
    mdk::Scalars s{};
    st = read_scalars(c, &s);
    if (st != MD_OK) return st;
    if (flag) *flag = s.rebuild_flag;
    return MD_OK;
}

md_status md_neigh_rebuild_if_needed(md_ctx* c, int* rebuilt) {
    int flag = 0;
    md_status st = md_neigh_needs_rebuild(c, &flag);
    if (st != MD_OK) return st;
    if (flag) st = build_now(c);
    if (rebuilt) *rebuilt = flag;
    return st;
}

md_status md_neigh_rebuild_count(const md_ctx* c, long* count) {
    if (c == nullptr || count == nullptr) return MD_ERR_INVALID_ARGUMENT;
    mdk::Scalars s{};
    md_ctx* m = const_cast<md_ctx*>(c);
    md_status st = read_scalars(m, &s);
    if (st != MD_OK) return st;
    *count = static_cast<long>(s.builds);
    return MD_OK;
}

md_status md_neigh_export_pairs(md_ctx* c, uint32_t* ij, long cap, long* npairs) {
    md_status st = need_system(c);
    if (st != MD_OK) return st;
    if (!c->list_built) return fail(c, MD_ERR_STATE, "the Verlet list has not been built");
    const long n = c->n;
    std::vector<int> cnt(static_cast<size_t>(n)), orig(static_cast<size_t>(n));
    std::vector<uint32_t> nb(static_cast<size_t>(c->cap) * c->npad);
    MD_HIP(c, hipStreamSynchronize(c->stream));
    MD_HIP(c, hipMemcpy(cnt.data(), c->ncnt, n * sizeof(int), hipMemcpyDeviceToHost));
    MD_HIP(c, hipMemcpy(orig.data(), c->orig, n * sizeof(int), hipMemcpyDeviceToHost));
    if (c->bk.on) {
        // brick entries are halo indices: translated to slot entries on the device
        uint32_t* tr = nullptr;
        MD_HIP(c, hipMalloc(reinterpret_cast<void**>(&tr), nb.size() * sizeof(uint32_t)));
        mdk::launch_brick_export(force_args(c, MD_FP64), tr, c->stream);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess)
            e = hipMemcpy(nb.data(), tr, nb.size() * sizeof(uint32_t), hipMemcpyDeviceToHost);
        (void)hipFree(tr);
        MD_HIP(c, e);
    } else {
        MD_HIP(c, hipMemcpy(nb.data(), c->nbr, nb.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    }
    std::vector<uint64_t> keys;
    keys.reserve(static_cast<size_t>(n) * 40);
    for (long t = 0; t < n; ++t) {
        const int k = std::min(cnt[t], c->cap);
        for (int q = 0; q < k; ++q) {
            const uint32_t e = nb[static_cast<size_t>(mdk::oct_index(t, q, c->cap))];
            const int j = mdk::entry_slot(e);
            const uint32_t a = static_cast<uint32_t>(orig[t]), b = static_cast<uint32_t>(orig[j]);
            if (a < b) keys.push_back((static_cast<uint64_t>(a) << 32) | b);
        }
    }
    std::sort(keys.begin(), keys.end());
    if (npairs) *npairs = static_cast<long>(keys.size());
    if (static_cast<long>(keys.size()) > cap)
        return fail(c, MD_ERR_CAPACITY, "pair buffer too small");
    for (size_t k = 0; k < keys.size(); ++k) {
        ij[2 * k] = static_cast<uint32_t>(keys[k] >> 32);
        ij[2 * k + 1] = static_cast<uint32_t>(keys[k] & 0xffffffffu);
    }
    return MD_OK;
}

md_status md_neigh_set_bricks(md_ctx* c, int cells) {
    if (c == nullptr || cells < -1 || cells > 8) return MD_ERR_INVALID_ARGUMENT;
    c->brick_width = cells;
    c->halo_boost = 1.0;
    c->list_built = false;
    if (c->n > 0 && c->neigh_configured) return alloc_neigh(c);
    return MD_OK;
}

md_status md_neigh_bricks(const md_ctx* c, int* count) {
    if (c == nullptr || count == nullptr) return MD_ERR_INVALID_ARGUMENT;
    *count = c->bk.on ? c->bk.nb : 0;
    return MD_OK;
}

md_status md_neigh_set_capacity(md_ctx* c, int per_atom) {
    if (c == nullptr || per_atom < 0) return MD_ERR_INVALID_ARGUMENT;
    c->cap_user = per_atom;
    c->list_built = false;
    if (c->n > 0 && c->neigh_configured) return alloc_neigh(c);
    return MD_OK;
}

md_status md_integrator_set(md_ctx* c, const md_integrator* it) {
    if (c == nullptr || it == nullptr) return MD_ERR_INVALID_ARGUMENT;
    if (!(it->dt > 0.0)) return fail(c, MD_ERR_INVALID_ARGUMENT, "dt must be positive");
    if (it->kind == MD_NOSE_HOOVER) {
        if (it->chain_length < 1)
            return fail(c, MD_ERR_INVALID_ARGUMENT, "NoseHooverChain: chainLength must be >= 1");
        if (it->chain_length > mdk::kMaxChain)
            return fail(c, MD_ERR_INVALID_ARGUMENT, "chain length above the supported maximum (16)");
        mdk::Chain ch{};
        ch.M = it->chain_length;
        ch.dof = static_cast<double>(it->dof);
        ch.T = it->temperature;
        // Q_1 = dof T tau^2, Q_k = T tau^2 (nose_hoover.cpp:45-49)
        for (int k = 0; k < ch.M; ++k) ch.Q[k] = it->temperature * it->tau * it->tau;
        ch.Q[0] = static_cast<double>(it->dof) * it->temperature * it->tau * it->tau;
        for (int k = 0; k < ch.M; ++k) ch.iQ[k] = 1.0 / ch.Q[k];
        // w1 = 1 / (2 - 2^(1/3)) with the host's cbrt, as nose_hoover.cpp:15-17 forms it
        const double w1 = 1.0 / (2.0 - std::cbrt(2.0));
        ch.sy[0] = w1;
        ch.sy[1] = 1.0 - 2.0 * w1;
        ch.sy[2] = w1;
        MD_HIP(c, hipMemcpy(c->chain, &ch, sizeof(ch), hipMemcpyHostToDevice));
    } else if (it->kind != MD_VELOCITY_VERLET && it->kind != MD_LANGEVIN_BAOAB) {
        return fail(c, MD_ERR_INVALID_ARGUMENT, "unknown integrator kind");
    }
    c->integ = *it;
    c->integ_set = true;
    c->step_counter = 0;
    return MD_OK;
}

// md_run replays captured segments of this many steps for systems of at
// most kGraphMaxAtoms atoms (where launch latency, not the kernels, sets the
// step time), unless per-kernel timing is on.
constexpr long kGraphSegment = 32;
constexpr long kGraphMaxAtoms = 262144;

// The steps of one md_run attempt; *overflowed reports the overflow bits of
// the attempt's builds (1: a neighbour row longer than the capacity, 2: a
// brick halo above the LDS rows; the state is then not meaningful and md_run
// restores its snapshot).
static md_status run_steps(md_ctx* c, long nsteps, int precision, md_thermo* thermo,
                           int* overflowed) {
    md_status st = MD_OK;
    *overflowed = 0;
    if (thermo && nsteps > c->thermo_alloc) {
        MD_HIP(c, alloc(c->thermo_dev, nsteps));
        c->thermo_alloc = nsteps;
    }
    const bool nhc = c->integ.kind == MD_NOSE_HOOVER;
    mdk::StepArgs sa = step_args(c);
    mdk::NeighArgs na = neigh_args(c);
    mdk::ForceArgs fa = force_args(c, precision);
    if (fa.qpos != nullptr) {
        // every pre-step pass refreshes the quantised positions for the mixed kernel
        sa.qpos = c->qpos;
        sa.quant = c->quant;
    }
    fa.kick = 1;
    fa.half_dt = 0.5 * c->integ.dt;
    fa.chain = nhc ? c->chain : nullptr;
    fa.chain_delta = 0.5 * c->integ.dt;
    if (nhc) {
        // the first half-chain of the run: kinetic energy of the current
        // velocities, one chain propagation of dt/2, rescale left pending in
        // vscale for the first pre-step pass
        mdk::launch_kinetic(c->n, c->vel, c->part_d, c->sc, c->stream);
        mdk::launch_chain_first(c->chain, c->sc, 0.5 * c->integ.dt, c->stream);
    }
    hipGraphExec_t gexec = nullptr;
    if (c->timing == 0 && c->n <= kGraphMaxAtoms && nsteps > kGraphSegment) {
        // Launch-bound runs replay a captured segment of kGraphSegment steps:
        // the step counter (BAOAB noise) and the thermo row come from the
        // device counter sc->step, advanced by each step's finish. Every
        // captured step propagates the next step's first half-chain, so the
        // run's last step is launched directly.
        c->step_upload = c->step_counter;
        MD_HIP(c, hipMemcpyAsync(&c->sc->step, &c->step_upload, sizeof(unsigned long long),
                                 hipMemcpyHostToDevice, c->stream));
        mdk::StepArgs gs = sa;
        gs.dev_step = 1;
        mdk::ForceArgs gf = fa;
        gf.dev_step = 1;
        gf.thermo = nullptr;
        gf.thermo_base = thermo ? static_cast<void*>(c->thermo_dev) : nullptr;
        gf.step0 = c->step_counter;
        gf.chain_twice = nhc ? 1 : 0;
        auto one_step = [&]() {
            mdk::launch_pre_step(gs, c->stream);
            mdk::launch_rebuild(na, c->stream);
            mdk::launch_lj_force(gf, c->stream);
            mdk::launch_lj_finish(gf, c->stream);
        };
        hipGraph_t graph = nullptr;
        MD_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        for (long k = 0; k < kGraphSegment; ++k) one_step();
        const hipError_t le = hipGetLastError();
        const hipError_t ce = hipStreamEndCapture(c->stream, &graph);
        if (le != hipSuccess || ce != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return fail(c, MD_ERR_HIP, std::string("md_run capture: ") +
                                           hipGetErrorString(le != hipSuccess ? le : ce));
        }
        const hipError_t ie = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        MD_HIP(c, ie);
        const long reps = (nsteps - 1) / kGraphSegment;  // the last step always runs directly
        for (long r = 0; r < reps; ++r) {
            const hipError_t e = hipGraphLaunch(gexec, c->stream);
            if (e != hipSuccess) {
                (void)hipStreamSynchronize(c->stream);
                (void)hipGraphExecDestroy(gexec);
                return fail(c, MD_ERR_HIP, std::string("md_run graph launch: ") + hipGetErrorString(e));
            }
        }
        for (long k = reps * kGraphSegment; k < nsteps; ++k) {
            gf.chain_twice = (nhc && k + 1 < nsteps) ? 1 : 0;
            one_step();
        }
        c->step_counter += static_cast<uint64_t>(nsteps);
    } else {
        for (long k = 0; k < nsteps; ++k) {
            sa.step = c->step_counter++;
            hipEvent_t e2{};
            ev_begin(c, 2, &e2);
            mdk::launch_pre_step(sa, c->stream);
            ev_end(c, 2, e2);
            hipEvent_t e1{};
            ev_begin(c, 1, &e1);
            mdk::launch_rebuild(na, c->stream);
            ev_end(c, 1, e1);
            fa.thermo = thermo ? static_cast<void*>(c->thermo_dev + k) : nullptr;
            fa.chain_twice = (nhc && k + 1 < nsteps) ? 1 : 0;
            hipEvent_t e0{};
            ev_begin(c, 0, &e0);
            mdk::launch_lj_force(fa, c->stream);
            ev_end(c, 0, e0);
            mdk::launch_lj_finish(fa, c->stream);
        }
    }
    if (nhc) mdk::launch_apply_scale(c->n, c->vel, c->sc, c->stream);
    MD_HIP(c, hipGetLastError());
    mdk::Scalars s{};
    st = read_scalars(c, &s);  // synchronises the stream
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (st != MD_OK) return st;
    if (s.overflow) {
        *overflowed = s.overflow;
        return MD_OK;
    }
    if (s.fix_saturated)
        return fail(c, MD_ERR_RUNTIME, "fixed-point force accumulator saturated (|f| > 2^19)");
    if (thermo)
        MD_HIP(c, hipMemcpy(thermo, c->thermo_dev, nsteps * sizeof(md_thermo), hipMemcpyDeviceToHost));
    return MD_OK;
}

// A neighbour row longer than the capacity (or a brick halo above the LDS
// rows) stops the work of the remaining steps (the pre-force pass and the
// pair kernel return at once once the device flag is set). md_run then
// restores the state it started from, doubles the capacity (or grows the
// LDS rows, turning the bricks off past 64 KB), builds the list again and
// runs the whole call again (at most three times), so a caller never sees a
// state advanced with a truncated list. The run that completes holds one
// extra list build.
md_status md_run(md_ctx* c, long nsteps, int precision, md_thermo* thermo) {
    md_status st = need_system(c);
    if (st != MD_OK) return st;
    if (!c->lj_set || !c->integ_set || !c->list_built)
        return fail(c, MD_ERR_STATE, "md_run needs md_lj_set, md_integrator_set and a built list");
    if (nsteps <= 0) return MD_OK;
    clear_call_flags(c);
    st = save_snapshot(c);
    if (st != MD_OK) return st;
    for (int attempt = 0;; ++attempt) {
        int over = 0;
        st = run_steps(c, nsteps, precision, thermo, &over);
        if (st != MD_OK || !over) return st;
        if (attempt == 3 || ((over & 1) && c->cap >= c->n - 1))
            return fail(c, MD_ERR_CAPACITY, "neighbour row capacity could not be satisfied in md_run");
        st = restore_snapshot(c);
        if (st != MD_OK) return st;
        if (over & 1) c->cap_user = static_cast<int>(std::min<long>(2L * c->cap, std::max(1L, c->n - 1)));
        if (over & 2) c->halo_boost *= 1.5;
        st = alloc_neigh(c);
        if (st != MD_OK) return st;
        st = build_now(c);
        if (st != MD_OK) return st;
    }
}

int md_neigh_auto_capacity(long n, const double* L, double cutoff, double skin) {
    if (n <= 0 || L == nullptr || !(cutoff > 0.0) || !(skin >= 0.0)) return 0;
    for (int d = 0; d < 3; ++d)
        if (!(L[d] > 0.0) || !std::isfinite(L[d])) return 0;
    return capacity_for(n, L, cutoff, skin);
}

md_status md_half_kick(md_ctx* c, double dt) {
    md_status st = need_system(c);
    if (st != MD_OK) return st;
    mdk::launch_half_kick(c->n, c->vel, c->frc, 0.5 * dt, c->stream);
    MD_HIP(c, hipGetLastError());
    return MD_OK;
}

md_status md_drift(md_ctx* c, double h) {
    md_status st = need_system(c);
    if (st != MD_OK) return st;
    mdk::launch_drift(c->n, c->box, c->pos, c->vel, h, c->stream);
    MD_HIP(c, hipGetLastError());
    return MD_OK;
}

md_status md_langevin_o(md_ctx* c, double c1, double c2, double temperature, uint64_t seed,
                        uint64_t counter) {
    md_status st = need_system(c);
    if (st != MD_OK) return st;
    mdk::launch_langevin_o(c->n, c->vel, c->orig, c1, c2, temperature, seed, counter, c->stream);
    MD_HIP(c, hipGetLastError());
    return MD_OK;
}

md_status md_scale_velocities(md_ctx* c, double s) {
    md_status st = need_system(c);
    if (st != MD_OK) return st;
    mdk::launch_scale(c->n, c->vel, s, c->stream);
    MD_HIP(c, hipGetLastError());
    return MD_OK;
}

md_status md_lj_forces_cells(md_ctx* c, double cell_cutoff, double* energy, double* virial) {
    md_status st = need_system(c);
    if (st != MD_OK) return st;
    if (!c->lj_set) return fail(c, MD_ERR_STATE, "md_lj_set has not been called");
    if (!(cell_cutoff > 0.0)) return fail(c, MD_ERR_INVALID_ARGUMENT, "cell cutoff must be positive");
    const mdk::Grid g = make_grid(c->L, cell_cutoff);
    if (!g.usable)
        return fail(c, MD_ERR_INVALID_ARGUMENT,
                    "CellList is not usable (fewer than 3 cells per side); fall back to the "
                    "brute-force overload");
    const long m = static_cast<long>(g.nc[0]) * g.nc[1] * g.nc[2];
    if (m > c->ci_ncell_alloc) {
        MD_HIP(c, alloc(c->ci_count, m));
        MD_HIP(c, alloc(c->ci_start, m + 1));
        MD_HIP(c, alloc(c->ci_fill, m));
        c->ci_ncell_alloc = m;
    }
    MD_HIP(c, alloc(c->ci_cell_of, c->n));
    MD_HIP(c, alloc(c->ci_atoms, c->n));
    MD_HIP(c, alloc(c->ci_tmp, c->n));
    mdk::launch_bin_index(c->n, c->pos, g, c->ci_cell_of, c->ci_count, c->ci_start, c->ci_fill,
                          c->ci_tmp, c->ci_atoms, nullptr, c->stream);
    mdk::DirectArgs a{};
    a.n = c->n;
    a.box = c->box;
    a.lj = c->ljc;
    a.pos = c->pos;
    a.frc = c->frc;
    a.orig = c->orig;
    a.part_d = c->part_d;
    a.part_l = c->part_l;
    a.sc = c->sc;
    a.grid = g;
    a.cell_of = c->ci_cell_of;
    a.cell_start = c->ci_start;
    a.cell_atoms = c->ci_atoms;
    mdk::launch_lj_direct(a, c->stream);
    MD_HIP(c, hipGetLastError());
    mdk::Scalars s{};
    st = read_scalars(c, &s);
    if (st != MD_OK) return st;
    if (energy) *energy = s.energy;
    if (virial) *virial = s.virial;
    return MD_OK;
}

md_status md_nhc_state(md_ctx* c, double* conserved, double* xi, double* p_xi) {
    if (c == nullptr) return MD_ERR_INVALID_ARGUMENT;
    mdk::Chain ch{};
    MD_HIP(c, hipStreamSynchronize(c->stream));
    MD_HIP(c, hipMemcpy(&ch, c->chain, sizeof(ch), hipMemcpyDeviceToHost));
    if (c->integ.kind != MD_NOSE_HOOVER) {
        if (conserved) *conserved = 0.0;
        return MD_OK;
    }
    double acc = ch.dof * ch.T * ch.xi[0];
    for (int k = 1; k < ch.M; ++k) acc += ch.T * ch.xi[k];
    for (int k = 0; k < ch.M; ++k) acc += 0.5 * ch.p[k] * ch.p[k] / ch.Q[k];
    if (conserved) *conserved = acc;
    for (int k = 0; k < ch.M; ++k) {
        if (xi) xi[k] = ch.xi[k];
        if (p_xi) p_xi[k] = ch.p[k];
    }
    return MD_OK;
}

md_status md_timing_enable(md_ctx* c, int on) {
    if (c == nullptr) return MD_ERR_INVALID_ARGUMENT;
    c->timing = on != 0 ? 7 : 0;
    return MD_OK;
}

md_status md_timing_select(md_ctx* c, int mask) {
    if (c == nullptr || mask < 0 || mask > 7) return MD_ERR_INVALID_ARGUMENT;
    c->timing = mask;
    return MD_OK;
}

md_status md_timing_read(md_ctx* c, int kernel_id, double* total_ms, long* launches) {
    if (c == nullptr || kernel_id < 0 || kernel_id > 2) return MD_ERR_INVALID_ARGUMENT;
    md_status st = resolve_events(c);
    if (st != MD_OK) return st;
    if (total_ms) *total_ms = c->ms[kernel_id];
    if (launches) *launches = c->launches[kernel_id];
    return MD_OK;
}

md_status md_timing_reset(md_ctx* c) {
    if (c == nullptr) return MD_ERR_INVALID_ARGUMENT;
    md_status st = resolve_events(c);
    for (int k = 0; k < 3; ++k) {
        c->ms[k] = 0.0;
        c->launches[k] = 0;
    }
    return st;
}

// Seeded fcc lattice (util/lattice.cpp:15-64) with the reference's random
// source (util/random.hpp:37-76): mt19937_64, uniform u = 1 - U[0,1) in
// (0,1], Marsaglia polar normals with a cached spare.
md_status md_fcc_lattice(int cells, double density, double temperature, uint64_t seed, double* L,
                         double* x, double* y, double* z, double* vx, double* vy, double* vz) {
#pragma clang fp contract(off)
    if (cells < 1 || !(density > 0.0) || L == nullptr || x == nullptr || y == nullptr ||
        z == nullptr || vx == nullptr || vy == nullptr || vz == nullptr)
        return MD_ERR_INVALID_ARGUMENT;
    std::mt19937_64 eng(seed);
    std::uniform_real_distribution<double> unif(0.0, 1.0);
    bool have_spare = false;
    double spare = 0.0;
    auto uniform = [&]() { return 1.0 - unif(eng); };
    auto normal = [&]() {
        if (have_spare) {
            have_spare = false;
            return spare;
        }
        double a, b, s;
        do {
            a = 2.0 * uniform() - 1.0;
            b = 2.0 * uniform() - 1.0;
            s = a * a + b * b;
        } while (s >= 1.0 || s == 0.0);
        const double k = std::sqrt(-2.0 * std::log(s) / s);
        spare = b * k;
        have_spare = true;
        return a * k;
    };
    const long n = 4L * cells * cells * cells;
    const double len = std::cbrt(static_cast<double>(n) / density);
    const double lat = len / static_cast<double>(cells);
    const double sv = std::sqrt(temperature / 1.0);
    const double basis[4][3] = {{0.0, 0.0, 0.0}, {0.0, 0.5, 0.5}, {0.5, 0.0, 0.5}, {0.5, 0.5, 0.0}};
    auto wrap = [len](double v) { return v - len * std::floor(v / len + 0.5); };
    long k = 0;
    for (int ix = 0; ix < cells; ++ix)
        for (int iy = 0; iy < cells; ++iy)
            for (int iz = 0; iz < cells; ++iz)
                for (const auto& b : basis) {
                    x[k] = wrap(lat * (static_cast<double>(ix) + b[0]) - 0.5 * len);
                    y[k] = wrap(lat * (static_cast<double>(iy) + b[1]) - 0.5 * len);
                    z[k] = wrap(lat * (static_cast<double>(iz) + b[2]) - 0.5 * len);
                    const double g0 = normal();
                    const double g1 = normal();
                    const double g2 = normal();
                    vx[k] = sv * g0;
                    vy[k] = sv * g1;
                    vz[k] = sv * g2;
                    ++k;
                }
    // centre-of-mass removal with unit masses (core/system.cpp:147-164)
    double mass = 0.0, px = 0.0, py = 0.0, pz = 0.0;
    for (long i = 0; i < n; ++i) mass += 1.0;
    for (long i = 0; i < n; ++i) {
        px += 1.0 * vx[i];
        py += 1.0 * vy[i];
        pz += 1.0 * vz[i];
    }
    const double cx = px / mass, cy = py / mass, cz = pz / mass;
    for (long i = 0; i < n; ++i) {
        vx[i] -= cx;
        vy[i] -= cy;
        vz[i] -= cz;
    }
    L[0] = L[1] = L[2] = len;
    return MD_OK;
}

md_status md_last_pair_count(md_ctx* c, long* pairs) {
    if (c == nullptr || pairs == nullptr) return MD_ERR_INVALID_ARGUMENT;
    mdk::Scalars s{};
    md_status st = read_scalars(c, &s);
    if (st != MD_OK) return st;
    *pairs = static_cast<long>(s.pairs_in_cut);
    return MD_OK;
}

}  // extern "C"
