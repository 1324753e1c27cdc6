// C ABI (include/fqsb.h) of the MI355X implementation of the
// FrictionQPotSpringBlock dynamics hot path.
//
// A `Group` holds R systems of one description in contiguous device arrays
// ([R, N] per field, [R, N, nchunk] for the landscape). A single system
// (fqsb_system) is a group of one; a batch (fqsb_batch) is a group of R
// replicas. The host code is the control flow of the reference's member
// functions (detail.h:1141-2050): it launches the kernels of fq_resident.hpp
// and fq_service.hpp and keeps a host mirror of the per-system scalars
// (RunState) that it pushes before and pulls after each launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <exception>
#include <map>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/fqsb.h"
#include "fq_launch.hpp"

#include <mutex>

// Device memory of fqsb_batch_quasistatic_noalloc: the per-event records
// (S, u_frame, mean f_frame, inc: 32 bytes per system and event) that
// fqsb_batch_quasistatic stages in buffers it allocates per call. Every GPU
// holds its own copy until the process ends; calls that overlap on one GPU
// share it and race.
constexpr size_t QS_SCRATCH_BYTES = size_t(8) << 20;
static __device__ __align__(256) unsigned char scratch[QS_SCRATCH_BYTES];

__global__ void k_zero_words(uint64_t* __restrict__ p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        p[i] = 0;
    }
}

namespace {

thread_local std::string g_err;

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        throw Error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
    }
}

#define HIPC(x) hip_check((x), #x)

template <class F>
int guard(F&& f)
{
    try {
        f();
        return 0;
    }
    catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// Zeroed device memory. The fill runs on the null stream, which the groups'
// non-blocking streams are not ordered after, so it is completed here before
// any kernel can touch the memory.
template <class T>
T* dalloc(size_t n)
{
    void* p = nullptr;
    HIPC(hipMalloc(&p, sizeof(T) * (n ? n : 1)));
    HIPC(hipMemsetAsync(p, 0, sizeof(T) * (n ? n : 1), nullptr));
    HIPC(hipStreamSynchronize(nullptr));
    return static_cast<T*>(p);
}

template <class T>
void dfree(T*& p)
{
    if (p) {
        (void)hipFree(p);
        p = nullptr;
    }
}

// steps per resident launch of a minimise-type run (bounds one kernel's duration)
constexpr uint64_t MIN_BLOCK = 1u << 18;
// adjoint: device memory for the tape of the windows recorded by one launch
constexpr size_t ADJ_TAPE_BYTES = size_t(16) << 30;
// adjoint record: chains up to this size run whole on one workgroup (4 x 1024
// threads); longer ones over tiles with halos
constexpr int64_t ADJ_RECORD_RING = 4096;
// In ragged sets, systems above this size run on the tiled path (many
// workgroups per system): with a few hundred systems the resident path
// leaves a large system to one workgroup, the tiled path spreads it over the
// chip (measured on the ragged workload: 5.4e10 vs 4.3e10 particle-steps/s
// with the resident limit of a single workgroup).
constexpr int64_t RAGGED_RESIDENT_MAX = 2048;
// steps per resident launch of timeSteps / flowSteps
constexpr uint64_t STEP_BLOCK = 1u << 16;
// long-range chains (without forcing) above this size step over tiles of
// LR_TILE_PARTICLES particles (one per thread), each tile holding every slip
// of the chain in LDS
constexpr int64_t LR_RESIDENT_MAX = 2048;
constexpr int64_t LR_TILE_PARTICLES = 256;
// largest chain of the resident no-passing kernel (16 particles per thread on
// 512 threads); longer chains sweep over tiles
constexpr int64_t NOPASSING_RESIDENT_MAX = 8192;
// sweeps between host checks of a tiled no-passing run, and particles per tile
constexpr uint64_t NP_BLOCK = 256;
constexpr int64_t NP_TILE_PARTICLES = 4096;
// steps between host checks of a batched long-range run (matrix form)
constexpr uint64_t LR_BLOCK = 128;
// steps between host checks of a tiled run
constexpr uint64_t TILE_BLOCK = 256;

struct Group;
std::vector<bool> nopassing_run(Group& g, double tol, uint64_t niter_tol, uint64_t max_iter);

struct Group {
    fqsb_desc d{};
    int64_t R = 0, rows = 0, cols = 0, nchunk = 0;
    int64_t N = 0;            // largest system of the group (all systems for a uniform batch)
    int64_t total = 0;        // sum of the system sizes
    std::vector<int64_t> off; // off[r]: first element of system r in every field; off[R] = total
    bool uniform = true;      // every system of the same shape and interaction parameters
    const uint8_t* mask = nullptr; // if set: systems with mask[r] == 0 sit out the next runs
    int device = 0;
    int64_t refills = 0; // landscape refills so far (the adjoint needs runs without)
    // adjoint (fq_adjoint.hpp): checkpoints of the last adjoint_forward, tape
    // of a group of windows, adjoint state
    int64_t adj_n = 0, adj_K = 0, adj_nw = 0;
    int adj_flow = 0;
    double adj_vf = 0.0;
    bool adj_valid = false;
    uint64_t adj_gen = 0; // number of adjoint forward passes: names the tape a backward pass may use
    double *ck_u = nullptr, *ck_v = nullptr, *ck_a = nullptr;
    int32_t* ck_j = nullptr;
    std::vector<double> ck_uf; // [nw][R]
    double *tp_u = nullptr, *tp_v = nullptr, *tp_a = nullptr, *tp_uf = nullptr;
    int32_t* tp_j = nullptr;
    double* ck_uf_d = nullptr; // device copy of ck_uf
    double *ad_u = nullptr, *ad_v = nullptr, *ad_a = nullptr, *ad_par = nullptr;
    double *ad_u2 = nullptr, *ad_v2 = nullptr, *ad_a2 = nullptr, *ad_gpart = nullptr;
    size_t ad_gpart_cap = 0;
    int32_t* ad_err = nullptr;
    int64_t ck_cap = 0;  // windows the checkpoint buffers hold
    size_t tp_cap = 0;   // entries each tape buffer holds
    hipEvent_t eva[3] = {nullptr, nullptr, nullptr}; // around record and reverse launches
    float adj_record_ms = 0.0f, adj_reverse_ms = 0.0f; // per launch, last backward
    int pot = 0, inter = 0, ext = 0, minim = 0;
    double *u = nullptr, *v = nullptr, *a = nullptr, *f = nullptr, *f_pot = nullptr, *f_int = nullptr,
           *f_frame = nullptr, *f_damp = nullptr, *f_th = nullptr, *pref = nullptr, *y = nullptr;
    int64_t *start = nullptr, *i_ref = nullptr, *next = nullptr, *dinc = nullptr;
    int32_t* jloc = nullptr;
    uint64_t *st_begin = nullptr, *st_end = nullptr;
    fq::RunState* rs = nullptr;
    fq::SysView* views = nullptr;
    std::vector<fq::SysView> hviews;
    std::vector<fq::RunState> hrs;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // ragged sets: time of the group's part of every operation on its stream
    // (events around it) and the particle-steps (no-passing: particle-sweeps)
    // it took, summed until reset
    hipEvent_t evr[2] = {nullptr, nullptr};
    double busy_ms = 0.0, busy_ps = 0.0;
    int ppt = 0, maxt = 0, threads = 0;
    bool streaming = false; // systems too large for a resident workgroup
    // tiled per-step path for streaming groups (fq_tile.hpp); FQSB_NO_TILES=1
    // keeps them on the one-workgroup streaming kernel
    bool tiled = false;
    int ntiles = 0;
    size_t tile_lds = 0;
    fq::TileDesc* tiles = nullptr;
    int32_t* tile_first = nullptr;
    double* tile_part = nullptr;
    int32_t* tile_flag = nullptr;
    // blocks of tile_k steps per launch for chains (fq_tile.hpp k_tile_block;
    // FQSB_TILE_K=0: one launch per step)
    int tile_k = 0;
    int tb_ntiles = 0;                // tiles of the blocked path (chains: 1024 particles; lattices: rows)
    int tb_kmax = fq::TB_KMAX_STEPS;  // longest block the tiles' halos allow (lattices: TL_K_DEFAULT)
    size_t tb_lds = 0;                // dynamic LDS of a lattice block launch
    fq::TileDesc* tb_tiles = nullptr;
    int32_t* tb_first = nullptr;      // [R + 1]
    double* tile_part_k = nullptr;    // [tb_ntiles][tile_k][2]
    int32_t* tile_fail = nullptr;     // [tb_ntiles]
    // captured blocks of tiled steps (hipGraph), keyed by the run arguments and
    // the number of steps; FQSB_NO_GRAPHS=1 launches the kernels one by one
    struct TileGraph {
        fq::RunArgs ra;
        uint64_t steps;
        hipGraphExec_t exec;
    };
    std::vector<TileGraph> tile_graphs;
    bool use_graphs = true;
    // no-passing chains beyond the resident kernel (or FQSB_NOPASSING_TILES=1):
    // Jacobi sweeps over the tiles of `tiles` (fq_tile.hpp k_np_*), the second
    // iterate buffer in u2 and the current wells' yield positions in yl / yr
    bool np_tiled = false;
    std::vector<TileGraph> np_graphs;
    // FQSB_GENERAL_KERNEL=1: timeSteps / flowSteps on k_resident instead of
    // k_flow (the two give the same results; a switch for comparisons)
    bool general_kernel = false;
    double *u2 = nullptr, *v2 = nullptr, *a2 = nullptr, *yl = nullptr, *yr = nullptr, *yl2 = nullptr, *yr2 = nullptr;
    int32_t* jl2 = nullptr;
    size_t lds_limit = 65536;
    int cus = 256; // compute units of the device (4 SIMDs each)
    // LongRange1d batches in matrix form (fq_longrange.hpp)
    int lr_method = FQSB_LR_EXACT;
    int64_t lr_m = 0;
    double lr_c = 0.0;
    int lr_nsplit = 1;
    hipEvent_t evg0 = nullptr, evg1 = nullptr; // around one product launch per block (timed runs)
    float lr_product_ms = 0.0f;                // average product duration of the last timed run
    // last quasistatic run: time around its device work and number of
    // driver launches (0: driven from the host)
    float qs_ms = 0.0f;
    int64_t qs_launches = 0;
    double *lr_un = nullptr, *lr_part = nullptr, *lr_yl = nullptr, *lr_yr = nullptr;

    Group() = default;
    Group(const Group&) = delete;
    Group& operator=(const Group&) = delete;

    ~Group()
    {
        if (stream) {
            (void)hipStreamSynchronize(stream);
        }
        dfree(u);
        dfree(v);
        dfree(a);
        dfree(f);
        dfree(f_pot);
        dfree(f_int);
        dfree(f_frame);
        dfree(f_damp);
        dfree(f_th);
        dfree(pref);
        dfree(y);
        dfree(start);
        dfree(i_ref);
        dfree(next);
        dfree(dinc);
        dfree(jloc);
        dfree(st_begin);
        dfree(st_end);
        dfree(rs);
        dfree(views);
        dfree(u2);
        dfree(v2);
        dfree(a2);
        dfree(yl);
        dfree(yr);
        dfree(yl2);
        dfree(yr2);
        dfree(jl2);
        dfree(lr_un);
        dfree(lr_part);
        free_adjoint();
        for (auto& tg : tile_graphs) {
            (void)hipGraphExecDestroy(tg.exec);
        }
        for (auto& tg : np_graphs) {
            (void)hipGraphExecDestroy(tg.exec);
        }
        dfree(tiles);
        dfree(tile_first);
        dfree(tile_part);
        dfree(tile_flag);
        dfree(tile_part_k);
        dfree(tile_fail);
        dfree(tb_tiles);
        dfree(tb_first);
        dfree(lr_yl);
        dfree(lr_yr);
        for (auto& e : evr) {
            if (e) {
                (void)hipEventDestroy(e);
            }
        }
        if (ev0) {
            (void)hipEventDestroy(ev0);
        }
        if (ev1) {
            (void)hipEventDestroy(ev1);
        }
        if (evg0) {
            (void)hipEventDestroy(evg0);
        }
        if (evg1) {
            (void)hipEventDestroy(evg1);
        }
        if (stream) {
            (void)hipStreamDestroy(stream);
        }
    }

    void push()
    {
        HIPC(hipMemcpyAsync(rs, hrs.data(), sizeof(fq::RunState) * R, hipMemcpyHostToDevice, stream));
    }

    void pull()
    {
        HIPC(hipMemcpyAsync(hrs.data(), rs, sizeof(fq::RunState) * R, hipMemcpyDeviceToHost, stream));
        HIPC(hipStreamSynchronize(stream));
    }

    template <class T>
    void d2h(T* host, const T* dev, size_t n)
    {
        HIPC(hipMemcpyAsync(host, dev, sizeof(T) * n, hipMemcpyDeviceToHost, stream));
        HIPC(hipStreamSynchronize(stream));
    }

    template <class T>
    void h2d(T* dev, const T* host, size_t n)
    {
        HIPC(hipMemcpyAsync(dev, host, sizeof(T) * n, hipMemcpyHostToDevice, stream));
        HIPC(hipStreamSynchronize(stream));
    }

    double* field(int fld) const
    {
        switch (fld) {
            case FQSB_U:
                return u;
            case FQSB_V:
                return v;
            case FQSB_A:
                return a;
            case FQSB_F:
                return f;
            case FQSB_F_POTENTIAL:
                return f_pot;
            case FQSB_F_INTERACTIONS:
                return f_int;
            case FQSB_F_FRAME:
                return f_frame;
            case FQSB_F_DAMPING:
                return f_damp;
            case FQSB_F_THERMAL:
                if (!ext) {
                    throw Error("system has no external forcing");
                }
                return f_th;
            default:
                throw Error("unknown field");
        }
    }

    void forces(int32_t flags)
    {
        push();
        if (!fq::launch_forces(pot, inter, ext, views, (int)R, flags, stream)) {
            throw Error("no force kernel for this potential / interaction / external combination");
        }
        HIPC(hipGetLastError());
        pull();
    }

    void updated_u() // detail.h:1380-1386
    {
        forces(fq::FF_POT | fq::FF_INT | fq::FF_FRAME | fq::FF_TOTAL);
    }

    void updated_v() // detail.h:1391-1395
    {
        forces(fq::FF_DAMP | fq::FF_TOTAL);
    }

    void updated_inc() // detail.h:1369-1375
    {
        forces(ext ? (fq::FF_THERMAL | fq::FF_TOTAL) : fq::FF_TOTAL);
    }

    void refresh() // detail.h:1310-1315
    {
        forces(fq::FF_POT | fq::FF_INT | fq::FF_FRAME | fq::FF_DAMP | fq::FF_TOTAL | (ext ? fq::FF_THERMAL : 0));
    }

    // Pick the resident variant: fewest particles per thread that fits, then
    // the smallest compiled thread bound (most registers per thread).
    void choose_launch(int want_ppt, int want_threads)
    {
        int vp[64], vm[64];
        int nv = std::min(fq::resident_variants(pot, inter, ext, vp, vm, 64), 64);
        if (nv <= 0) {
            throw Error("no resident kernel for this potential / interaction / external combination");
        }
        // Cuspy-Laplace chains (no external forcing) without a requested
        // shape: 4 particles per thread while the systems' wavefronts fit two
        // per SIMD (a workgroup's step is shortest there: a lone 1024-chain's
        // quasistatic step takes 2.37 us, against 2.63 us with 2 x 512 and
        // 3.04 us with 1 x 1024), else 8 (the most particle-steps per unit of
        // time with many systems: 1024 chains of 1024 particles), when a
        // variant of at most 512 threads holds the largest system that way
        // (variants of 1024 threads have 128 registers per thread and spill).
        // The other classes keep the fewest particles per thread: their
        // costlier forces spill at 4 particles per thread.
        if (want_ppt == 0 && want_threads == 0 && pot == FQSB_POT_CUSPY && inter == FQSB_INT_LAPLACE1D &&
            ext == FQSB_EXT_NONE) {
            for (const int p : {4, 8}) {
                const int thr = (int)((((N + p - 1) / p) + 63) / 64 * 64);
                if (thr > 1024 || fq::resident_lds(inter, N, thr) > lds_limit) {
                    continue;
                }
                if (p == 4 && R * (int64_t)(thr / 64) > 2 * 4 * (int64_t)cus) {
                    continue;
                }
                int best = -1;
                for (int i = 0; i < nv; ++i) {
                    if (vp[i] == p && thr <= vm[i] && vm[i] <= 512 && (best < 0 || vm[i] < vm[best])) {
                        best = i;
                    }
                }
                if (best >= 0) {
                    ppt = p;
                    maxt = vm[best];
                    threads = thr;
                    return;
                }
            }
        }
        int best = -1, best_thr = 0;
        for (int i = 0; i < nv; ++i) {
            const int p = vp[i], mt = vm[i];
            if (want_ppt > 0 && p != want_ppt) {
                continue;
            }
            int64_t need = (N + p - 1) / p;
            int thr = want_threads > 0 ? want_threads : (int)(((need + 63) / 64) * 64);
            if (thr % 64 != 0 || thr > mt || (int64_t)thr * p < N || thr > 1024) {
                continue;
            }
            if (fq::resident_lds(inter, N, thr) > lds_limit) {
                continue;
            }
            if (best < 0 || p < vp[best] || (p == vp[best] && mt < vm[best])) {
                best = i;
                best_thr = thr;
            }
        }
        if (best < 0) {
            throw Error("system size " + std::to_string(N) +
                        " is outside the capacity of the resident kernels for this interaction");
        }
        ppt = vp[best];
        maxt = vm[best];
        threads = best_thr;
    }

    // Resident launch shapes (particles per thread, threads) that fit the
    // group: for each kernel variant, the fewest threads that hold the largest
    // system, and twice that if the variant allows it.
    std::vector<std::pair<int, int>> launch_shapes() const
    {
        std::vector<std::pair<int, int>> out;
        if (streaming || minim == FQSB_MIN_OVERDAMPED) {
            return out;
        }
        int vp[64], vm[64];
        const int nv = std::min(fq::resident_variants(pot, inter, ext, vp, vm, 64), 64);
        for (int i = 0; i < nv; ++i) {
            const int p = vp[i], mt = vm[i];
            const int thr = (int)((((N + p - 1) / p) + 63) / 64 * 64);
            for (int t2 : {thr, 2 * thr}) {
                if (t2 > mt || t2 > 1024 || fq::resident_lds(inter, N, t2) > lds_limit) {
                    continue;
                }
                const std::pair<int, int> sh(p, t2);
                if (std::find(out.begin(), out.end(), sh) == out.end()) {
                    out.push_back(sh);
                }
            }
        }
        return out;
    }

    void launch_resident(const fq::RunArgs& ra)
    {
        if (minim == FQSB_MIN_OVERDAMPED) {
            throw Error("dynamics are not available for no-passing systems");
        }
        if (streaming) {
            if (!fq::launch_stream(pot, inter, ext, views, (int)R, ra, stream)) {
                throw Error("no streaming kernel for this potential / interaction / external combination");
            }
            HIPC(hipGetLastError());
            return;
        }
        size_t lds = fq::resident_lds(inter, N, threads);
        // timeSteps / flowSteps of chains: the pipelined k_flow (same results)
        const bool chain = inter == FQSB_INT_NONE || inter == FQSB_INT_LAPLACE1D || inter == FQSB_INT_QUARTIC1D ||
                           inter == FQSB_INT_QUARTICGRADIENT1D;
        bool unit_mass = true, full = true;
        for (const auto& S : hviews) {
            unit_mass &= S.inv_m == 1.0;
            full &= S.N % ppt == 0;
        }
        if (ra.mode == fq::MODE_STEPS && chain && !general_kernel &&
            fq::launch_flow(pot, inter, ext, ppt, maxt, threads, lds, views, (int)R, ra, stream, unit_mass, full)) {
            HIPC(hipGetLastError());
            return;
        }
        if (!fq::launch_resident(pot, inter, ext, ppt, maxt, threads, lds, views, (int)R, ra, stream)) {
            throw Error("no resident kernel variant for this launch shape");
        }
        HIPC(hipGetLastError());
    }

    void reset_run(int mode)
    {
        for (size_t i = 0; i < hrs.size(); ++i) {
            auto& r = hrs[i];
            r.status = (mask && !mask[i]) ? fq::ST_IDLE : fq::ST_RUNNING;
            r.steps = 0;
            r.below = 0;
            r.below2 = 0;
            r.dec = 0;
            r.prev = std::numeric_limits<double>::infinity();
            if (mode == fq::MODE_TRUNCATE || mode == fq::MODE_MINIMISE) {
                r.act_init = 1;
                r.s_n = 0;
            }
        }
    }

    // steps per block of the blocked tiled path: the requested length within
    // what the tiles' halos allow
    int block_steps() const
    {
        return std::min(tile_k, tb_kmax);
    }

    bool any_running() const
    {
        for (const auto& r : hrs) {
            if (r.status == fq::ST_RUNNING || r.status == fq::ST_BUDGET) {
                return true;
            }
        }
        return false;
    }

    // Run up to `total` steps per system (or until every system stopped) in
    // launches of at most `block` steps. A launch that stops before a step
    // needing a landscape refill (ST_REFILL) is followed by k_refill and a
    // relaunch. Returns the time between events around all launches (ms).
    // Systems over many workgroups: per step one k_tile_step and one
    // k_tile_decide for all systems, checked by the host every `block` steps;
    // a pending refill normalises the state (k_tile_finish), refills
    // (k_refill) and refreshes the cached wells (k_tile_prologue).
    float run_tiled(fq::RunArgs ra, uint64_t total, uint64_t block, bool timed)
    {
        reset_run(ra.mode);
        for (auto& r : hrs) {
            r.target = (int64_t)total;
            r.cur = 0;
            r.tb_redo = 0;
            r.tb_fail = 0;
        }
        push();
        if (timed) {
            HIPC(hipEventRecord(ev0, stream));
        }
        HIPC(fq::launch_tile_prologue(views, tiles, ntiles, stream));
        const bool blocks = tile_k > 0 && tile_part_k != nullptr && fq::tile_block_supported(pot, inter);
        auto normalise = [&] {
            if (!fq::launch_tile_finish(pot, inter, views, tiles, ntiles, stream)) {
                throw Error("no tiled kernel for this potential / interaction");
            }
            HIPC(hipGetLastError());
            pull();
            for (auto& r : hrs) {
                r.cur = 0;
            }
            push();
        };
        auto enqueue = [&](uint64_t nb) {
            if (blocks) { // blocks of K steps (decide commits or redoes)
                const int K = block_steps();
                for (uint64_t s = 0; s < (nb + (uint64_t)K - 1) / (uint64_t)K; ++s) {
                    if (!fq::launch_tile_block(pot, inter, views, tb_tiles, tb_ntiles, tile_part_k, tile_fail, ra, K,
                                               tb_lds, stream)) {
                        throw Error("no blocked tile kernel for this potential / interaction");
                    }
                    HIPC(hipGetLastError());
                    HIPC(fq::launch_tile_block_decide(views, (int)R, tb_first, tile_part_k, tile_fail, ra, K, stream));
                }
                return;
            }
            for (uint64_t s = 0; s < nb; ++s) {
                if (!fq::launch_tile_step(pot, inter, views, tiles, ntiles, tile_lds, tile_part, tile_flag, ra,
                                          stream)) {
                    throw Error("no tiled kernel for this potential / interaction");
                }
                HIPC(hipGetLastError());
                HIPC(fq::launch_tile_decide(views, (int)R, tile_first, tile_part, tile_flag, ra, stream));
            }
        };
        // a block of nb steps as one graph launch (captured once per run
        // arguments and length; the kernels read the run state from device
        // memory, so a replay continues where the last one stopped)
        auto launch_block = [&](uint64_t nb) {
            if (!use_graphs) {
                enqueue(nb);
                return;
            }
            for (const auto& tg : tile_graphs) {
                if (tg.steps == nb && std::memcmp(&tg.ra, &ra, sizeof(ra)) == 0) {
                    HIPC(hipGraphLaunch(tg.exec, stream));
                    return;
                }
            }
            hipGraph_t graph = nullptr;
            HIPC(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            try {
                enqueue(nb);
            }
            catch (...) {
                (void)hipStreamEndCapture(stream, &graph);
                if (graph) {
                    (void)hipGraphDestroy(graph);
                }
                throw;
            }
            HIPC(hipStreamEndCapture(stream, &graph));
            hipGraphExec_t exec = nullptr;
            const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HIPC(e);
            if (tile_graphs.size() >= 8) {
                (void)hipGraphExecDestroy(tile_graphs.front().exec);
                tile_graphs.erase(tile_graphs.begin());
            }
            tile_graphs.push_back(TileGraph{ra, nb, exec});
            HIPC(hipGraphLaunch(exec, stream));
        };
        uint64_t todo = total;
        while (true) {
            const uint64_t nb = std::min<uint64_t>(block, std::max<uint64_t>(todo, 1));
            launch_block(nb);
            pull();
            bool refill = false, more = false;
            int64_t lag = 0;
            for (const auto& r : hrs) {
                refill |= r.status == fq::ST_REFILL;
                if ((r.status == fq::ST_RUNNING || r.status == fq::ST_BUDGET || r.status == fq::ST_REFILL) &&
                    r.steps < r.target) {
                    more |= r.status != fq::ST_REFILL;
                    lag = std::max<int64_t>(lag, r.target - r.steps);
                }
            }
            todo = (uint64_t)lag;
            if (refill) {
                normalise();
                HIPC(fq::launch_refill(views, (int)R, stream));
                ++refills;
                for (auto& r : hrs) {
                    if (r.status == fq::ST_REFILL) {
                        r.status = fq::ST_RUNNING;
                    }
                }
                HIPC(fq::launch_tile_prologue(views, tiles, ntiles, stream));
                continue;
            }
            if (!more) {
                break;
            }
        }
        normalise();
        float ms = 0.0f;
        if (timed) {
            HIPC(hipEventRecord(ev1, stream));
            HIPC(hipEventSynchronize(ev1));
            HIPC(hipEventElapsedTime(&ms, ev0, ev1));
        }
        for (auto& r : hrs) {
            if (r.status == fq::ST_RUNNING) {
                r.status = fq::ST_BUDGET;
            }
        }
        return ms;
    }

    float run(fq::RunArgs ra, uint64_t total, uint64_t block, bool timed)
    {
        if (tiled && (ra.mode == fq::MODE_STEPS || (ra.mode == fq::MODE_MINIMISE && !ra.time_activity))) {
            return run_tiled(ra, total, std::min<uint64_t>(block, TILE_BLOCK), timed);
        }
        reset_run(ra.mode);
        for (auto& r : hrs) {
            r.target = (int64_t)total;
        }
        push();
        if (timed) {
            HIPC(hipEventRecord(ev0, stream));
        }
        ra.max_steps = (int64_t)std::min<uint64_t>(block, (uint64_t)std::numeric_limits<int32_t>::max());
        while (total > 0) {
            launch_resident(ra);
            pull();
            bool refill = false, more = false;
            for (const auto& r : hrs) {
                refill |= r.status == fq::ST_REFILL;
                more |= (r.status == fq::ST_RUNNING || r.status == fq::ST_BUDGET) && r.steps < r.target;
            }
            if (refill) {
                HIPC(fq::launch_refill(views, (int)R, stream));
                ++refills;
                for (auto& r : hrs) {
                    if (r.status == fq::ST_REFILL) {
                        r.status = fq::ST_RUNNING;
                    }
                }
                continue;
            }
            if (!more) {
                break;
            }
        }
        float ms = 0.0f;
        if (timed) {
            HIPC(hipEventRecord(ev1, stream));
            HIPC(hipEventSynchronize(ev1));
            HIPC(hipEventElapsedTime(&ms, ev0, ev1));
        }
        pull();
        return ms;
    }


    void free_adjoint()
    {
        for (double** q : {&ck_u, &ck_v, &ck_a, &tp_u, &tp_v, &tp_a, &tp_uf, &ck_uf_d, &ad_u, &ad_v, &ad_a,
                           &ad_par, &ad_u2, &ad_v2, &ad_a2, &ad_gpart}) {
            dfree(*q);
            *q = nullptr;
        }
        for (int32_t** q : {&ck_j, &tp_j, &ad_err}) {
            dfree(*q);
            *q = nullptr;
        }
        ck_cap = 0;
        tp_cap = 0;
        ad_gpart_cap = 0;
        adj_valid = false;
        for (auto& e : eva) {
            if (e) {
                (void)hipEventDestroy(e);
                e = nullptr;
            }
        }
    }

    fq::AdjBuffers adj_buffers(double* gy) const
    {
        fq::AdjBuffers ab{};
        ab.tu = tp_u;
        ab.tv = tp_v;
        ab.ta = tp_a;
        ab.tj = tp_j;
        ab.tuf = tp_uf;
        ab.au = ad_u;
        ab.av = ad_v;
        ab.aa = ad_a;
        ab.gpar = ad_par;
        ab.gy = gy;
        ab.err = ad_err;
        ab.N = N;
        ab.R = R;
        return ab;
    }

    void adjoint_check() const
    {
        if (pot != FQSB_POT_CUSPY || inter != FQSB_INT_LAPLACE1D || ext != FQSB_EXT_NONE ||
            minim == FQSB_MIN_OVERDAMPED || !uniform) {
            throw Error("the adjoint needs replicas of one shape with the Cuspy potential, Laplace1d "
                        "interactions and no external forcing");
        }
        if (N < 3) {
            throw Error("the adjoint needs chains of at least 3 particles");
        }
    }

    // n steps (timeSteps, or flowSteps at v_frame), keeping a checkpoint every K
    void adjoint_forward(uint64_t n, uint64_t K, bool flow, double v_frame)
    {
        adjoint_check();
        if (K < 1 || n < 1) {
            throw Error("the adjoint needs n >= 1 steps and windows of K >= 1 steps");
        }
        // chains beyond one workgroup run both adjoint kernels over tiles with
        // halos of about K particles on each side
        const int64_t rev_cap = (int64_t)fq::ADJ_REV_THREADS * fq::ADJ_REV_MAX_SLOTS;
        if ((N > ADJ_RECORD_RING && (int64_t)K >= ADJ_RECORD_RING / 2) ||
            (N > rev_cap && 2 * ((int64_t)K + 1) >= rev_cap)) {
            throw Error("the adjoint's tiles do not fit: windows of " + std::to_string(K) +
                        " steps are too long for a chain of " + std::to_string(N) + " particles");
        }
        adj_valid = false;
        const int64_t nw = (int64_t)((n + K - 1) / K);
        const size_t RN = (size_t)total;
        if (nw > ck_cap) {
            dfree(ck_u);
            dfree(ck_v);
            dfree(ck_a);
            dfree(ck_j);
            ck_u = dalloc<double>(RN * (size_t)nw);
            ck_v = dalloc<double>(RN * (size_t)nw);
            ck_a = dalloc<double>(RN * (size_t)nw);
            ck_j = dalloc<int32_t>(RN * (size_t)nw);
            ck_cap = nw;
        }
        ck_uf.assign((size_t)(nw * R), 0.0);
        const int64_t refills0 = refills;
        for (int64_t w = 0; w < nw; ++w) {
            const size_t o = (size_t)w * RN;
            HIPC(hipMemcpyAsync(ck_u + o, u, sizeof(double) * RN, hipMemcpyDeviceToDevice, stream));
            HIPC(hipMemcpyAsync(ck_v + o, v, sizeof(double) * RN, hipMemcpyDeviceToDevice, stream));
            HIPC(hipMemcpyAsync(ck_a + o, a, sizeof(double) * RN, hipMemcpyDeviceToDevice, stream));
            HIPC(hipMemcpyAsync(ck_j + o, jloc, sizeof(int32_t) * RN, hipMemcpyDeviceToDevice, stream));
            for (int64_t r = 0; r < R; ++r) {
                ck_uf[(size_t)(w * R + r)] = hrs[(size_t)r].u_frame;
            }
            const uint64_t nsteps = std::min<uint64_t>(K, n - (uint64_t)w * K);
            steps(nsteps, flow, v_frame, nsteps, false, false);
        }
        if (refills != refills0) {
            throw Error("the trajectory left the landscape window (a refill was needed); the adjoint needs "
                        "a larger nchunk");
        }
        adj_n = (int64_t)n;
        adj_K = (int64_t)K;
        adj_nw = nw;
        adj_flow = flow ? 1 : 0;
        adj_vf = v_frame;
        adj_valid = true;
        ++adj_gen;
    }

    // cotangents of (u, v, a) after the last adjoint_forward -> cotangents of
    // (u, v, a) before it, d/d(mu, k, k_frame, eta) [R][4] and d/dy
    // [R][N][nchunk] (or null); all device pointers
    void adjoint_backward(const double* gu, const double* gv, const double* ga, double* gu0, double* gv0,
                          double* ga0, double* gpar, double* gy)
    {
        if (!adj_valid) {
            throw Error("no adjoint forward run to differentiate");
        }
        adjoint_check();
        const size_t RN = (size_t)total;
        // the record kernel's segments: the whole periodic chain up to
        // ADJ_RECORD_RING particles, else tiles with halos of K particles
        // whose segments (own + 2 K) fit 4 x 1024 threads
        int64_t rnt = 1, rtp = N, rseg = N;
        if (N > ADJ_RECORD_RING) {
            const int64_t cap = ADJ_RECORD_RING - 2 * adj_K;
            if (cap < 1) {
                throw Error("the adjoint's record tiles do not fit: window too long for the chain");
            }
            rnt = (N + cap - 1) / cap;
            rtp = (N + rnt - 1) / rnt;
            rnt = (N + rtp - 1) / rtp;
            rseg = rtp + 2 * adj_K;
        }
        // launch shape of the record kernel: fewest particles per thread that
        // fits (4 x 1024 at 4096 holds its registers without spills; 8 x
        // 512 does not)
        int vp[8], vm[8];
        const int nv = fq::adj_variants(vp, vm, 8);
        int ppt = 0, threads = 0;
        for (int i = 0; i < nv; ++i) {
            const int64_t need = (rseg + vp[i] - 1) / vp[i];
            const int thr = (int)(((need + 63) / 64) * 64);
            if (thr <= vm[i]) {
                ppt = vp[i];
                threads = thr;
                break;
            }
        }
        if (!ppt) {
            throw Error("no adjoint kernel for this system size");
        }
        // tape entries per step: particle-major, N per replica
        const int64_t np = N;
        const size_t RNt = (size_t)(R * np);
        // tiles of the reverse pass: enough workgroups to fill the chip (at
        // least 256), each tile's particles at least four times its halo of
        // K + 1 particles per side (one tile: the whole periodic chain, no
        // halo); FQSB_ADJ_TILES=n asks for n tiles per replica
        const int64_t halo = adj_K + 1;
        int64_t nt = 1;
        if (const char* e = std::getenv("FQSB_ADJ_TILES")) {
            nt = std::max<int64_t>(1, std::atoll(e));
        }
        else {
            while (R * nt < 256 && (N + 2 * nt - 1) / (2 * nt) >= 4 * halo) {
                nt *= 2;
            }
        }
        // at least as many tiles as make a tile and its halos fit the kernel
        const int64_t rev_cap = (int64_t)fq::ADJ_REV_THREADS * fq::ADJ_REV_MAX_SLOTS;
        if (N > rev_cap) {
            nt = std::max<int64_t>(nt, 2);
            while (nt < N && (N + nt - 1) / nt + 2 * halo > rev_cap) {
                ++nt;
            }
        }
        int64_t tp = (N + nt - 1) / nt;
        nt = (N + tp - 1) / tp;
        if (nt > 1 && tp < halo) {
            nt = 1;
            tp = N;
        }
        const int slots = fq::adj_reverse_slots(nt == 1 ? N : tp + 2 * halo);
        if (slots == 0) {
            throw Error("the adjoint's reverse tiles do not fit: window too long for the chain");
        }
        // the tape of a group of windows, recorded by one launch (every window
        // starts from its checkpoint): as many windows as fit in ADJ_TAPE_BYTES
        const size_t per_window = RNt * (size_t)adj_K * (3 * sizeof(double) + sizeof(int32_t)) +
                                  (size_t)(R * adj_K) * sizeof(double);
        const int64_t G = std::max<int64_t>(1, std::min<int64_t>(adj_nw, (int64_t)(ADJ_TAPE_BYTES / per_window)));
        if ((size_t)(G * adj_K) * RNt > tp_cap) {
            dfree(tp_u);
            dfree(tp_v);
            dfree(tp_a);
            dfree(tp_j);
            dfree(tp_uf);
            const size_t steps = (size_t)(G * adj_K);
            tp_u = dalloc<double>(RNt * steps);
            tp_v = dalloc<double>(RNt * steps);
            tp_a = dalloc<double>(RNt * steps);
            tp_j = dalloc<int32_t>(RNt * steps);
            tp_uf = dalloc<double>((size_t)R * steps);
            tp_cap = steps * RNt;
        }
        if (!ad_u) {
            ad_u = dalloc<double>(RN);
            ad_v = dalloc<double>(RN);
            ad_a = dalloc<double>(RN);
            ad_u2 = dalloc<double>(RN);
            ad_v2 = dalloc<double>(RN);
            ad_a2 = dalloc<double>(RN);
            ad_par = dalloc<double>((size_t)R * 4);
            ad_err = dalloc<int32_t>(1);
        }
        if ((size_t)(R * nt * 4) > ad_gpart_cap) {
            dfree(ad_gpart);
            ad_gpart = dalloc<double>((size_t)(R * nt * 4));
            ad_gpart_cap = (size_t)(R * nt * 4);
        }
        dfree(ck_uf_d);
        ck_uf_d = dalloc<double>(ck_uf.size());
        h2d(ck_uf_d, ck_uf.data(), ck_uf.size());
        if (gy) {
            HIPC(hipMemsetAsync(gy, 0, sizeof(double) * RN * (size_t)nchunk, stream));
        }
        HIPC(hipMemsetAsync(ad_par, 0, sizeof(double) * (size_t)R * 4, stream));
        HIPC(hipMemsetAsync(ad_err, 0, sizeof(int32_t), stream));
        HIPC(hipMemcpyAsync(ad_u, gu, sizeof(double) * RN, hipMemcpyDeviceToDevice, stream));
        HIPC(hipMemcpyAsync(ad_v, gv, sizeof(double) * RN, hipMemcpyDeviceToDevice, stream));
        HIPC(hipMemcpyAsync(ad_a, ga, sizeof(double) * RN, hipMemcpyDeviceToDevice, stream));
        for (auto& e : eva) {
            if (!e) {
                HIPC(hipEventCreate(&e));
            }
        }
        double rec_ms = 0.0, rev_ms = 0.0;
        for (int64_t w1 = adj_nw - 1; w1 >= 0; w1 -= G) {
            const int64_t w0 = std::max<int64_t>(0, w1 - G + 1);
            fq::AdjBuffers ab = adj_buffers(gy);
            ab.wu = ck_u + (size_t)w0 * RN;
            ab.wv = ck_v + (size_t)w0 * RN;
            ab.wa = ck_a + (size_t)w0 * RN;
            ab.wj = ck_j + (size_t)w0 * RN;
            ab.wuf = ck_uf_d + (size_t)(w0 * R);
            ab.ws = (int64_t)RN;
            ab.ts = adj_K * (int64_t)RNt;
            ab.np = np;
            ab.K = adj_K;
            ab.n = adj_n;
            ab.w0 = w0;
            ab.rnt = (int32_t)rnt;
            ab.rtp = rtp;
            ab.rhalo = rnt == 1 ? 0 : adj_K;
            HIPC(hipEventRecord(eva[0], stream));
            if (!fq::launch_adj_record(ppt, threads, views, ab, (int)(w1 - w0 + 1), adj_flow, adj_vf, stream)) {
                throw Error("adjoint record kernel not available");
            }
            HIPC(hipGetLastError());
            HIPC(hipEventRecord(eva[1], stream));
            for (int64_t w = w1; w >= w0; --w) {
                fq::AdjBuffers abw = ab;
                const size_t to = (size_t)(w - w0) * (size_t)adj_K;
                abw.tu = tp_u + to * RNt;
                abw.tv = tp_v + to * RNt;
                abw.ta = tp_a + to * RNt;
                abw.tj = tp_j + to * RNt;
                abw.tuf = tp_uf + to * (size_t)R;
                // adjoint state: read from one buffer, written to the other
                abw.au = ad_u;
                abw.av = ad_v;
                abw.aa = ad_a;
                abw.bu = ad_u2;
                abw.bv = ad_v2;
                abw.ba = ad_a2;
                abw.gpart = ad_gpart;
                abw.nt = (int32_t)nt;
                abw.tp = tp;
                abw.halo = halo;
                const int nsteps = (int)std::min<int64_t>(adj_K, adj_n - w * adj_K);
                if (!fq::launch_adj_reverse(views, abw, slots, nsteps, stream)) {
                    throw Error("adjoint reverse kernel not available");
                }
                HIPC(hipGetLastError());
                std::swap(ad_u, ad_u2);
                std::swap(ad_v, ad_v2);
                std::swap(ad_a, ad_a2);
            }
            HIPC(hipEventRecord(eva[2], stream));
            HIPC(hipEventSynchronize(eva[2]));
            float e0 = 0.0f, e1 = 0.0f;
            HIPC(hipEventElapsedTime(&e0, eva[0], eva[1]));
            HIPC(hipEventElapsedTime(&e1, eva[1], eva[2]));
            rec_ms += e0;
            rev_ms += e1;
        }
        adj_record_ms = (float)(rec_ms / (double)adj_nw);
        adj_reverse_ms = (float)(rev_ms / (double)adj_nw);
        int32_t err = 0;
        d2h(&err, ad_err, 1);
        if (err) {
            throw Error("a well walk left the landscape window while recomputing the trajectory");
        }
        HIPC(hipMemcpyAsync(gu0, ad_u, sizeof(double) * RN, hipMemcpyDeviceToDevice, stream));
        HIPC(hipMemcpyAsync(gv0, ad_v, sizeof(double) * RN, hipMemcpyDeviceToDevice, stream));
        HIPC(hipMemcpyAsync(ga0, ad_a, sizeof(double) * RN, hipMemcpyDeviceToDevice, stream));
        HIPC(hipMemcpyAsync(gpar, ad_par, sizeof(double) * (size_t)R * 4, hipMemcpyDeviceToDevice, stream));
        HIPC(hipStreamSynchronize(stream));
    }

    // per-replica mu, k_interactions, k_frame, eta (the parameters the adjoint
    // differentiates); forces are recomputed
    void set_parameters(const double* par)
    {
        for (int64_t r = 0; r < R; ++r) {
            fq::SysView& S = hviews[(size_t)r];
            S.mu = par[4 * r + 0];
            S.k = par[4 * r + 1];
            S.k_frame = par[4 * r + 2];
            S.eta = par[4 * r + 3];
        }
        h2d(views, hviews.data(), (size_t)R);
        refresh();
        adj_valid = false;
    }

    void get_parameters(double* par) const
    {
        for (int64_t r = 0; r < R; ++r) {
            const fq::SysView& S = hviews[(size_t)r];
            par[4 * r + 0] = S.mu;
            par[4 * r + 1] = S.k;
            par[4 * r + 2] = S.k_frame;
            par[4 * r + 3] = S.eta;
        }
    }

    void set_lr_method(int method)
    {
        if (inter != FQSB_INT_LONGRANGE1D) {
            throw Error("the systems have no long-range interactions");
        }
        if (method != FQSB_LR_EXACT && method != FQSB_LR_MATRIX) {
            throw Error("unknown long-range method");
        }
        if (method == FQSB_LR_MATRIX && pot != FQSB_POT_CUSPY) {
            throw Error("the matrix form of long-range interactions needs the Cuspy potential");
        }
        if (method == FQSB_LR_MATRIX && !uniform) {
            throw Error("the matrix form of long-range interactions needs replicas of one shape");
        }
        if (method == FQSB_LR_MATRIX && !lr_un) {
            const size_t RN = (size_t)(R * N);
            lr_un = dalloc<double>(RN);
            // split the summation index so that one product launch has about
            // 1024 workgroups (4 per CU, the occupancy that registers and LDS
            // allow), each range at least 8 K-blocks long; FQSB_LR_SPLIT
            // overrides the count (tuning)
            const int64_t kb = fq::lr_split_blocks();
            const int64_t nkb = (N + kb - 1) / kb;
            const int64_t tiles = ((N + 127) / 128) * ((R + 63) / 64);
            const int64_t cap = std::max<int64_t>(1, std::min<int64_t>(8, nkb / 8));
            int64_t ns = std::max<int64_t>(1, std::min<int64_t>(cap, (1024 + tiles - 1) / tiles));
            if (const char* e = std::getenv("FQSB_LR_SPLIT")) {
                ns = std::max<int64_t>(1, std::min<int64_t>(cap, std::atoll(e)));
            }
            lr_nsplit = (int)ns;
            lr_part = dalloc<double>(RN * (size_t)ns);
            lr_yl = dalloc<double>(RN);
            lr_yr = dalloc<double>(RN);
            // c = sum of one row of K: sum over d = 1 .. N-1 of pref[fold(d)]
            std::vector<double> pr((size_t)N);
            d2h(pr.data(), pref, (size_t)N);
            double c = 0.0;
            for (int64_t dd = 1; dd < N; ++dd) {
                c += pr[(size_t)(dd > lr_m ? N - dd : dd)];
            }
            lr_c = c;
        }
        lr_method = method;
    }

    fq::LrBuffers lr_buffers() const
    {
        return fq::LrBuffers{lr_un, lr_part, lr_yl, lr_yr, N, R, lr_c, lr_nsplit};
    }

    // timeSteps / flowSteps with LongRange1d in matrix form: per step one
    // k_lr_gemm (interaction forces of all replicas) and one k_lr_elem (rest
    // of timeStep). The host checks the run states every `block` steps; a
    // replica whose next slips need a landscape refill stops (ST_REFILL)
    // until k_refill and k_lr_prologue ran, then catches up (per-replica
    // step target).
    float lr_run(fq::RunArgs ra, uint64_t total, uint64_t block, bool timed)
    {
        reset_run(ra.mode);
        for (auto& r : hrs) {
            r.target = (int64_t)total;
        }
        push();
        const fq::LrBuffers lb = lr_buffers();
        if (timed && !evg0) {
            HIPC(hipEventCreate(&evg0));
            HIPC(hipEventCreate(&evg1));
        }
        double product_ms = 0.0;
        int64_t product_n = 0;
        if (timed) {
            HIPC(hipEventRecord(ev0, stream));
        }
        HIPC(fq::launch_lr_prologue(views, lb, stream));
        uint64_t todo = total;
        while (todo > 0) {
            const uint64_t nb = std::min<uint64_t>(block, todo);
            for (uint64_t s = 0; s < nb; ++s) {
                // timed runs: the first product of each block between two events
                const bool probe = timed && s == 0;
                if (probe) {
                    HIPC(hipEventRecord(evg0, stream));
                }
                HIPC(fq::launch_lr_gemm(pref, lb, lr_m, stream));
                if (probe) {
                    HIPC(hipEventRecord(evg1, stream));
                }
                HIPC(fq::launch_lr_elem(ext, views, lb, ra, stream));
            }
            pull();
            if (timed) {
                float e = 0.0f;
                HIPC(hipEventElapsedTime(&e, evg0, evg1));
                product_ms += e;
                ++product_n;
            }
            bool refill = false;
            int64_t lag = 0;
            for (const auto& r : hrs) {
                refill |= r.status == fq::ST_REFILL;
                if (r.status == fq::ST_RUNNING || r.status == fq::ST_REFILL) {
                    lag = std::max<int64_t>(lag, r.target - r.steps);
                }
            }
            if (refill) {
                HIPC(fq::launch_refill(views, (int)R, stream));
                HIPC(fq::launch_lr_prologue(views, lb, stream));
                for (auto& r : hrs) {
                    if (r.status == fq::ST_REFILL) {
                        r.status = fq::ST_RUNNING;
                    }
                }
            }
            todo = (uint64_t)lag;
        }
        float ms = 0.0f;
        if (timed) {
            HIPC(hipEventRecord(ev1, stream));
            HIPC(hipEventSynchronize(ev1));
            HIPC(hipEventElapsedTime(&ms, ev0, ev1));
            lr_product_ms = product_n ? (float)(product_ms / (double)product_n) : 0.0f;
        }
        pull();
        return ms;
    }

    // minimise (detail.h:1754-1785) of every system: batches of long-range
    // chains in matrix form run it on the FP64 matrix cores (the forces of all
    // replicas as one product per step), others on the resident / tiled /
    // streaming kernels
    float minimise_run(fq::RunArgs ra, uint64_t max_iter)
    {
        if (lr_method == FQSB_LR_MATRIX && !ra.time_activity) {
            return lr_run(ra, max_iter, LR_BLOCK, false);
        }
        return run(ra, max_iter, MIN_BLOCK, false);
    }

    float steps(uint64_t n, bool flow, double v_frame, uint64_t block, bool check_each, bool timed)
    {
        fq::RunArgs ra{};
        ra.mode = fq::MODE_STEPS;
        ra.flow = flow ? 1 : 0;
        ra.v_frame = v_frame;
        (void)check_each;
        if (minim == FQSB_MIN_OVERDAMPED) {
            throw Error("dynamics are not available for no-passing systems");
        }
        float ms = lr_method == FQSB_LR_MATRIX ? lr_run(ra, n, block ? block : LR_BLOCK, timed)
                                               : run(ra, n, block ? block : STEP_BLOCK, timed);
        for (const auto& r : hrs) {
            if (r.status == fq::ST_NAN) {
                throw Error("NaN entries found");
            }
        }
        return ms;
    }

    void set_iref_current()
    {
        HIPC(fq::launch_set_iref(views, (int)R, N, stream));
    }

    // quasistatic helpers -----------------------------------------------------

    double max_uniform(int64_t r, int direction, std::vector<double>* all = nullptr)
    {
        if (pot == FQSB_POT_SMOOTH) {
            throw Error("Operation not possible.");
        }
        push();
        if (!fq::launch_max_uniform(pot, views, (int)R, direction, stream)) {
            throw Error("maxUniformDisplacement not available");
        }
        HIPC(hipGetLastError());
        pull();
        if (all) {
            all->resize(R);
            for (int64_t i = 0; i < R; ++i) {
                (*all)[i] = hrs[i].red_flag ? 0.0 : hrs[i].red0;
            }
        }
        return hrs[r].red_flag ? 0.0 : hrs[r].red0;
    }

    // advanceUniformly detail.h:2027-2050 (input_is_frame = false), per system
    void advance_uniformly(const std::vector<double>& du, std::vector<double>* du_frame)
    {
        if (du_frame) {
            du_frame->assign(R, 0.0);
        }
        for (int64_t r = 0; r < R; ++r) {
            const double dp = du[r];
            const fq::SysView& S = hviews[(size_t)r];
            const double dfr = dp * (S.k_frame + S.mu) / S.k_frame;
            hrs[r].du_add = dp;
            hrs[r].u_frame += dfr;
            if (du_frame) {
                (*du_frame)[r] = dfr;
            }
        }
        forces(fq::FF_ADD_U | fq::FF_POT | fq::FF_INT | fq::FF_FRAME | fq::FF_TOTAL);
        for (auto& r : hrs) {
            r.du_add = 0.0;
        }
    }

    // eventDrivenStep detail.h:1933-1960 for every system (with mask set:
    // every system with mask[r] != 0; the others are left as they are, ret 0)
    std::vector<double> event_driven_step(double eps, bool kick, int direction)
    {
        std::vector<double> du(R, 0.0), ret(R, 0.0);
        auto on = [&](int64_t r) { return !mask || mask[r]; };
        if (kick) {
            for (int64_t r = 0; r < R; ++r) {
                du[r] = on(r) ? (direction > 0 ? eps : -eps) : 0.0;
            }
            advance_uniformly(du, &ret);
            return ret;
        }
        std::vector<double> m;
        max_uniform(0, direction, &m);
        bool any = false;
        for (int64_t r = 0; r < R; ++r) {
            if (!on(r) || m[r] < 0.5 * eps) {
                du[r] = 0.0;
            }
            else {
                du[r] = direction > 0 ? m[r] - 0.5 * eps : 0.5 * eps - m[r];
                any = true;
            }
        }
        if (any) {
            advance_uniformly(du, &ret);
            for (int64_t r = 0; r < R; ++r) {
                if (!on(r) || m[r] < 0.5 * eps) {
                    ret[r] = 0.0;
                }
            }
        }
        return ret;
    }

    // minimise_truncate detail.h:1833-1893 for every system from the reference
    // indices i_n (all systems' particles, system after system); ret[r] as
    // System::minimise_truncate
    void minimise_truncate(const int64_t* i_n, uint64_t A_truncate, uint64_t S_truncate, double tol,
                           uint64_t niter_tol, uint64_t max_iter, bool max_iter_is_error, uint64_t* ret)
    {
        if (minim != FQSB_MIN_INERTIAL) {
            throw Error("minimise_truncate is not available for this system");
        }
        h2d(i_ref, i_n, (size_t)total);
        fq::RunArgs ra{};
        ra.mode = fq::MODE_TRUNCATE;
        ra.tol = tol;
        ra.tol2 = tol * tol;
        ra.niter_tol = (int64_t)niter_tol;
        ra.A_truncate = (int64_t)A_truncate;
        ra.S_truncate = (int64_t)S_truncate;
        run(ra, max_iter, MIN_BLOCK, false);
        for (int64_t k = 0; k < R; ++k) {
            const auto& r = hrs[(size_t)k];
            if (r.status == fq::ST_NAN) {
                throw Error("NaN entries found");
            }
            if (r.status == fq::ST_CONVERGED) {
                ret[k] = 0;
            }
            else if (r.status == fq::ST_TRUNCATED) {
                ret[k] = (uint64_t)r.steps;
            }
            else if (max_iter_is_error) {
                throw Error("No convergence found");
            }
            else {
                ret[k] = max_iter + 1;
            }
        }
    }

    // advanceToFixedForce detail.h:1988-1995 for every system: target[r] is
    // the wanted mean frame force of system r (mean taken in particle order)
    void advance_to_fixed_force(const double* target, bool allow_plastic)
    {
        const size_t T = (size_t)total;
        std::vector<int64_t> st(T), st2(T);
        std::vector<int32_t> jl(T), jl2(T);
        std::vector<double> ff(T);
        d2h(st.data(), start, T);
        d2h(jl.data(), jloc, T);
        d2h(ff.data(), f_frame, T);
        std::vector<double> du((size_t)R);
        for (int64_t r = 0; r < R; ++r) {
            double mean = 0.0;
            for (int64_t p = off[(size_t)r]; p < off[(size_t)r + 1]; ++p) {
                mean += ff[(size_t)p];
            }
            mean /= (double)(off[(size_t)r + 1] - off[(size_t)r]);
            du[(size_t)r] = (target[r] - mean) / hviews[(size_t)r].mu;
        }
        advance_uniformly(du, nullptr);
        d2h(st2.data(), start, T);
        d2h(jl2.data(), jloc, T);
        bool same = true;
        for (size_t p = 0; p < T; ++p) {
            same &= st[p] + jl[p] == st2[p] + jl2[p];
        }
        if (!allow_plastic && !same) {
            throw Error("assertion failed (allow_plastic || xt::all(xt::equal(m_chunk->index_at_align(), i_n)))");
        }
    }

    // trigger detail.h:1972-1977 of particle p[r] in every system r (forces
    // are left as they are, as in the reference)
    void trigger_each(const uint64_t* p, double eps, int direction)
    {
        std::vector<int64_t> hp((size_t)R);
        for (int64_t r = 0; r < R; ++r) {
            if ((int64_t)p[r] >= off[(size_t)r + 1] - off[(size_t)r]) {
                throw Error("particle index out of range");
            }
            hp[(size_t)r] = (int64_t)p[r];
        }
        int64_t* dp = dalloc<int64_t>((size_t)R);
        h2d(dp, hp.data(), (size_t)R);
        const hipError_t e =
            fq::launch_trigger_each(views, (int)R, dp, eps, direction, pot == FQSB_POT_CUSPY ? 1 : 0, stream);
        const hipError_t e2 = hipStreamSynchronize(stream);
        dfree(dp);
        HIPC(e);
        HIPC(e2);
    }

    // The quasistatic protocol of the reference's examples
    // (examples/Line1d_Cuspy_Laplace.py:39-61) for every system (with mask
    // set: the masked-in ones): events k = first .. first + nevents - 1,
    //   k == 0:      u_frame = 0, minimise
    //   k odd:       eventDrivenStep(eps, false, direction)
    //   k even > 0:  eventDrivenStep(eps, true, direction), minimise
    // or, with kicks set, eventDrivenStep(eps, true, direction) and minimise
    // for every k > 0.
    // recording after event e, at [r * nevents + e] of the host arrays (any
    // may be null): S (sum of well-index changes), u_frame, mean f_frame, inc.
    // Systems that fit a resident workgroup run the whole protocol on the
    // device (k_quasi, one launch per step budget, refill or end); the others
    // are driven event by event from here. Returns the time around the device
    // work (ms).
    // `records`: device memory for the records (4 x R x nevents words, no
    // allocation in the call; the device-side driver only), or null to
    // allocate it here.
    float quasistatic(int64_t first, int64_t nevents, double eps, int direction, int kicks, double tol,
                      uint64_t niter_tol, uint64_t max_iter, int64_t* S, double* uf, double* ff, int64_t* inc_out,
                      void* records = nullptr)
    {
        if (first < 0 || nevents < 0) {
            throw Error("the first event and the number of events must be non-negative");
        }
        if (!(eps > 0.0) || !std::isfinite(eps)) {
            throw Error("eps must be positive and finite");
        }
        if (direction != 1 && direction != -1) {
            throw Error("direction must be +1 or -1");
        }
        if (!(tol < 1.0) || max_iter < 1) {
            throw Error("minimise needs tol < 1 and max_iter >= 1");
        }
        if (minim == FQSB_MIN_NONE) {
            throw Error("Minimisation not implementated");
        }
        // an elastic event needs maxUniformDisplacement, which Smooth raises
        const bool elastic = !kicks && nevents > 0 && (nevents > 1 || (first & 1));
        if (pot == FQSB_POT_SMOOTH && elastic) {
            throw Error("Operation not possible.");
        }
        if (ext) {
            throw Error("eventDrivenStep is not available for this system");
        }
        if (nevents == 0) {
            return 0.0f;
        }
        const size_t n = (size_t)(R * nevents);
        const bool own = records == nullptr;
        int64_t *dS, *dinc;
        double *duf, *dff;
        if (own) {
            dS = dalloc<int64_t>(n);
            duf = dalloc<double>(n);
            dff = dalloc<double>(n);
            dinc = dalloc<int64_t>(n);
        }
        else {
            if (streaming || minim != FQSB_MIN_INERTIAL) {
                throw Error("the quasistatic protocol without allocation needs the device-side driver "
                            "(systems of one resident workgroup, inertial minimisation)");
            }
            dS = (int64_t*)records;
            duf = (double*)(dS + n);
            dff = duf + n;
            dinc = (int64_t*)(dff + n);
            // the records start at zero, as the allocated ones do
            hipLaunchKernelGGL(k_zero_words, dim3((unsigned)std::min<size_t>((4 * n + 255) / 256, 1024)), dim3(256),
                               0, stream, (uint64_t*)records, 4 * n);
            HIPC(hipGetLastError());
        }
        fq::QsArgs qa{};
        qa.first = first;
        qa.nevents = nevents;
        qa.eps = eps;
        qa.direction = direction;
        qa.kicks = kicks ? 1 : 0;
        qa.max_iter = (int64_t)std::min<uint64_t>(max_iter, (uint64_t)std::numeric_limits<int64_t>::max());
        qa.out_S = dS;
        qa.out_uf = duf;
        qa.out_ff = dff;
        qa.out_inc = dinc;
        float ms = 0.0f;
        qs_launches = 0;
        try {
            set_iref_current();
            const bool device = !streaming && minim == FQSB_MIN_INERTIAL &&
                                (!own || std::getenv("FQSB_HOST_QUASISTATIC") == nullptr);
            if (!device || !quasistatic_device(qa, tol, niter_tol, ms)) {
                if (!own) {
                    throw Error("the quasistatic protocol without allocation needs the device-side driver, which "
                                "is not compiled for this system class");
                }
                ms = quasistatic_host(qa, tol, niter_tol, max_iter);
            }
            auto fetch = [&](auto* host, const auto* dev) {
                if (host) {
                    d2h(host, dev, n);
                }
            };
            fetch(S, dS);
            fetch(uf, duf);
            fetch(ff, dff);
            fetch(inc_out, dinc);
        }
        catch (...) {
            (void)hipStreamSynchronize(stream);
            if (own) {
                dfree(dS);
                dfree(duf);
                dfree(dff);
                dfree(dinc);
            }
            throw;
        }
        if (own) {
            dfree(dS);
            dfree(duf);
            dfree(dff);
            dfree(dinc);
        }
        qs_ms = ms;
        return ms;
    }

    // k_quasi for the whole protocol; false if it is not instantiated for this
    // group (the caller then drives the events from the host)
    bool quasistatic_device(const fq::QsArgs& qa, double tol, uint64_t niter_tol, float& ms)
    {
        for (size_t i = 0; i < hrs.size(); ++i) {
            auto& r = hrs[i];
            r.status = (mask && !mask[i]) ? fq::ST_IDLE : fq::ST_RUNNING;
            r.steps = 0;
            r.below = 0;
            r.below2 = 0;
            r.dec = 0;
            r.prev = std::numeric_limits<double>::infinity();
            r.qs_event = 0;
            r.qs_phase = fq::QS_BEGIN;
            r.qs_isum = 0;
        }
        fq::RunArgs ra{};
        ra.mode = fq::MODE_MINIMISE;
        ra.tol = tol;
        ra.tol2 = tol * tol;
        ra.niter_tol = (int64_t)niter_tol;
        // steps per launch (FQSB_QS_BLOCK: another budget, e.g. to exercise relaunches)
        ra.max_steps = (int64_t)MIN_BLOCK;
        if (const char* e = std::getenv("FQSB_QS_BLOCK")) {
            ra.max_steps = std::max<int64_t>(1, std::min<int64_t>(std::atoll(e), (int64_t)MIN_BLOCK));
        }
        const size_t lds = fq::quasi_lds(inter, N, ppt, maxt, threads);
        bool unit_mass = true, full = true;
        for (const auto& S : hviews) {
            unit_mass &= S.inv_m == 1.0;
            full &= S.N % ppt == 0;
        }
        push();
        HIPC(hipEventRecord(ev0, stream));
        for (bool first_launch = true;; first_launch = false) {
            if (!fq::launch_quasi(pot, inter, ppt, maxt, threads, lds, views, (int)R, ra, qa, stream, unit_mass,
                                  full)) {
                if (first_launch) {
                    return false;
                }
                throw Error("quasistatic kernel not available");
            }
            HIPC(hipGetLastError());
            ++qs_launches;
            pull();
            bool refill = false, more = false;
            for (const auto& r : hrs) {
                if (r.status == fq::ST_NAN) {
                    throw Error("NaN entries found");
                }
                if (r.status == fq::ST_MAXITER) {
                    throw Error("No convergence found");
                }
                refill |= r.status == fq::ST_REFILL;
                more |= r.status == fq::ST_RUNNING || r.status == fq::ST_BUDGET;
            }
            if (refill) {
                HIPC(fq::launch_refill(views, (int)R, stream));
                ++refills;
                for (auto& r : hrs) {
                    if (r.status == fq::ST_REFILL) {
                        r.status = fq::ST_RUNNING;
                    }
                }
                continue;
            }
            if (!more) {
                break;
            }
        }
        HIPC(hipEventRecord(ev1, stream));
        HIPC(hipEventSynchronize(ev1));
        HIPC(hipEventElapsedTime(&ms, ev0, ev1));
        return true;
    }

    // the protocol event by event with the group's operations
    float quasistatic_host(const fq::QsArgs& qa, double tol, uint64_t niter_tol, uint64_t max_iter)
    {
        auto on = [&](int64_t r) { return !mask || mask[r]; };
        auto record = [&](int64_t e, int begin) {
            for (int64_t r = 0; r < R; ++r) {
                if (!on(r)) {
                    hrs[(size_t)r].status = fq::ST_IDLE;
                }
                else if (hrs[(size_t)r].status == fq::ST_IDLE) {
                    hrs[(size_t)r].status = fq::ST_CONVERGED;
                }
            }
            push();
            HIPC(fq::launch_qs_record(views, (int)R, qa, e, begin, stream));
            pull();
        };
        HIPC(hipEventRecord(ev0, stream));
        for (int64_t e = 0; e < qa.nevents; ++e) {
            const int64_t k = qa.first + e;
            record(e, 1);
            if (k == 0) {
                for (int64_t r = 0; r < R; ++r) {
                    if (on(r)) {
                        hrs[(size_t)r].u_frame = 0.0;
                    }
                }
                forces(fq::FF_FRAME | fq::FF_TOTAL); // u_frame setter, detail.h:1253-1258
            }
            else {
                event_driven_step(qa.eps, qa.kicks || (k & 1) == 0, qa.direction);
            }
            if (qa.kicks || (k & 1) == 0) {
                if (minim == FQSB_MIN_OVERDAMPED) {
                    auto conv = nopassing_run(*this, tol, niter_tol, max_iter);
                    for (int64_t r = 0; r < R; ++r) {
                        if (on(r) && !conv[(size_t)r]) {
                            throw Error(hrs[(size_t)r].status == fq::ST_NAN ? "NaN entries found"
                                                                            : "No convergence found");
                        }
                    }
                }
                else {
                    fq::RunArgs ra{};
                    ra.mode = fq::MODE_MINIMISE;
                    ra.tol = tol;
                    ra.tol2 = tol * tol;
                    ra.niter_tol = (int64_t)niter_tol;
                    minimise_run(ra, max_iter);
                    for (int64_t r = 0; r < R; ++r) {
                        const auto st = hrs[(size_t)r].status;
                        if (on(r) && st != fq::ST_CONVERGED) {
                            throw Error(st == fq::ST_NAN ? "NaN entries found" : "No convergence found");
                        }
                    }
                }
            }
            record(e, 0);
        }
        float ms = 0.0f;
        HIPC(hipEventRecord(ev1, stream));
        HIPC(hipEventSynchronize(ev1));
        HIPC(hipEventElapsedTime(&ms, ev0, ev1));
        return ms;
    }
};

uint64_t default_stream_state(uint64_t seed, uint64_t& inc)
{
    uint64_t st;
    fq::pcg_seed(seed, fq::PCG_DEFAULT_STREAM, st, inc);
    return st;
}

// Checks one description; returns its number of particles.
int64_t validate_desc(const fqsb_desc& d)
{
    if (d.rank != 1 && d.rank != 2) {
        throw Error("rank must be 1 or 2");
    }
    if (d.distribution < 0 || d.distribution > FQSB_DIST_NORMAL) {
        throw Error("Unknown distribution: " + std::to_string(d.distribution));
    }
    if (d.distribution == FQSB_DIST_GAMMA || d.distribution == FQSB_DIST_NORMAL) {
        throw Error("Distribution not supported for the yield-position landscape");
    }
    if (d.nparameters < 0 || d.nparameters > 4) {
        throw Error("Too many parameters for distribution");
    }
    const int maxpar = (d.distribution == FQSB_DIST_POWER || d.distribution == FQSB_DIST_PARETO ||
                        d.distribution == FQSB_DIST_WEIBULL)
                           ? 3
                           : 2;
    if (d.nparameters > maxpar) {
        throw Error("Too many parameters for distribution");
    }
    if (d.nchunk < 8) {
        throw Error("nchunk must be at least 8");
    }
    const int64_t rows = (int64_t)d.rows, cols = d.rank == 1 ? 1 : (int64_t)d.cols;
    const int64_t N = rows * cols;
    if (N < 1) {
        throw Error("empty system");
    }
    const int inter = d.interactions;
    if ((inter == FQSB_INT_LAPLACE1D || inter == FQSB_INT_QUARTIC1D || inter == FQSB_INT_QUARTICGRADIENT1D) &&
        N < 3) {
        throw Error("a chain needs at least 3 particles");
    }
    if ((inter == FQSB_INT_LAPLACE2D || inter == FQSB_INT_QUARTICGRADIENT2D) && d.rank != 2) {
        throw Error("2d interactions need rank 2");
    }
    if (d.minimisation == FQSB_MIN_OVERDAMPED && (inter != FQSB_INT_LAPLACE1D || d.potential != FQSB_POT_CUSPY)) {
        throw Error("no-passing minimisation needs Cuspy + Laplace1d");
    }
    if (d.external == FQSB_EXT_RANDOMFORCING && (!d.dinc_init || !d.dinc)) {
        throw Error("random forcing needs dinc_init and dinc");
    }
    return N;
}

// A group of systems that share rank, potential, interactions, external
// forcing and minimisation (one kernel instantiation), each with its own size
// and parameters. Fields are stored system after system ([sum N]); the
// landscape of system r is its [N_r, nchunk_r] block.
Group* build_group_multi(const std::vector<fqsb_desc>& ds, int64_t resident_max = 0)
{
    if (ds.empty()) {
        throw Error("at least one replica is needed");
    }
    const int64_t R = (int64_t)ds.size();
    std::vector<int64_t> Ns((size_t)R);
    for (int64_t r = 0; r < R; ++r) {
        const fqsb_desc& d = ds[(size_t)r];
        Ns[(size_t)r] = validate_desc(d);
        const fqsb_desc& d0 = ds[0];
        if (d.rank != d0.rank || d.potential != d0.potential || d.interactions != d0.interactions ||
            d.external != d0.external || d.minimisation != d0.minimisation) {
            throw Error("systems of one group must share rank, potential, interactions, forcing and minimisation");
        }
    }
    const fqsb_desc& d0 = ds[0];

    std::unique_ptr<Group> g(new Group());
    g->d = d0;
    g->d.dinc_init = nullptr;
    g->d.dinc = nullptr;
    g->R = R;
    g->pot = d0.potential;
    g->inter = d0.interactions;
    g->ext = d0.external;
    g->minim = d0.minimisation;
    g->off.assign((size_t)R + 1, 0);
    std::vector<size_t> yoff((size_t)R + 1, 0);
    bool uniform = true;
    for (int64_t r = 0; r < R; ++r) {
        const fqsb_desc& d = ds[(size_t)r];
        g->off[(size_t)r + 1] = g->off[(size_t)r] + Ns[(size_t)r];
        yoff[(size_t)r + 1] = yoff[(size_t)r] + (size_t)Ns[(size_t)r] * (size_t)d.nchunk;
        g->N = std::max<int64_t>(g->N, Ns[(size_t)r]);
        uniform &= d.rows == d0.rows && d.cols == d0.cols && d.nchunk == d0.nchunk &&
                   d.k_interactions == d0.k_interactions && d.alpha == d0.alpha;
    }
    g->uniform = uniform;
    g->total = g->off[(size_t)R];
    g->rows = (int64_t)d0.rows;
    g->cols = d0.rank == 1 ? 1 : (int64_t)d0.cols;
    g->nchunk = (int64_t)d0.nchunk;
    const int inter = g->inter;

    int dev = 0;
    HIPC(hipGetDevice(&dev));
    g->device = dev;
    hipDeviceProp_t prop;
    HIPC(hipGetDeviceProperties(&prop, dev));
    g->lds_limit = prop.sharedMemPerBlock - 2048;
    g->cus = prop.multiProcessorCount;
    if (const char* gk = std::getenv("FQSB_GENERAL_KERNEL")) {
        g->general_kernel = gk[0] == '1';
    }
    if (const char* ng = std::getenv("FQSB_NO_GRAPHS")) {
        g->use_graphs = ng[0] != '1';
    }
    HIPC(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    HIPC(hipEventCreate(&g->ev0));
    HIPC(hipEventCreate(&g->ev1));

    const size_t RN = (size_t)g->total;
    g->u = dalloc<double>(RN);
    g->v = dalloc<double>(RN);
    g->a = dalloc<double>(RN);
    g->f = dalloc<double>(RN);
    g->f_pot = dalloc<double>(RN);
    g->f_int = dalloc<double>(RN);
    g->f_frame = dalloc<double>(RN);
    g->f_damp = dalloc<double>(RN);
    g->y = dalloc<double>(yoff[(size_t)R]);
    g->start = dalloc<int64_t>(RN);
    g->i_ref = dalloc<int64_t>(RN);
    g->jloc = dalloc<int32_t>(RN);
    g->st_begin = dalloc<uint64_t>(RN);
    g->st_end = dalloc<uint64_t>(RN);
    g->rs = dalloc<fq::RunState>((size_t)R);
    g->views = dalloc<fq::SysView>((size_t)R);
    if (g->ext) {
        g->f_th = dalloc<double>(RN);
        g->next = dalloc<int64_t>(RN);
        g->dinc = dalloc<int64_t>(RN);
        for (int64_t r = 0; r < R; ++r) {
            g->h2d(g->next + g->off[(size_t)r], ds[(size_t)r].dinc_init, (size_t)Ns[(size_t)r]);
            g->h2d(g->dinc + g->off[(size_t)r], ds[(size_t)r].dinc, (size_t)Ns[(size_t)r]);
        }
    }
    if (inter == FQSB_INT_LONGRANGE1D) { // LongRange1d ctor detail.h:829-844, per system
        g->pref = dalloc<double>(RN);
        for (int64_t r = 0; r < R; ++r) {
            const fqsb_desc& d = ds[(size_t)r];
            const int64_t N = Ns[(size_t)r];
            std::vector<double> pr((size_t)N, 0.0);
            for (int64_t dd = 1; dd < N; ++dd) {
                pr[(size_t)dd] = d.k_interactions / std::pow((double)dd, d.alpha + 1.0);
            }
            g->h2d(g->pref + g->off[(size_t)r], pr.data(), (size_t)N);
        }
        g->lr_m = (g->N - g->N % 2) / 2;
    }

    // resident workgroup if the largest system fits one, else the streaming
    // path (FQSB_FORCE_STREAMING=1 selects the streaming path for every size;
    // FQSB_RESIDENT_MAX=n for systems above n particles)
    if (g->minim != FQSB_MIN_OVERDAMPED) {
        const char* force = std::getenv("FQSB_FORCE_STREAMING");
        const char* rmax = std::getenv("FQSB_RESIDENT_MAX");
        try {
            if (force && force[0] == '1') {
                throw Error("streaming forced");
            }
            const int64_t limit = rmax ? std::atoll(rmax) : resident_max;
            if (limit > 0 && g->N > limit) {
                throw Error("streaming above the resident size limit");
            }
            // long range without forcing above LR_RESIDENT_MAX: the O(N^2) sum of
            // a step spread over tiles (one particle per thread) instead of one
            // workgroup
            if (inter == FQSB_INT_LONGRANGE1D && g->ext == FQSB_EXT_NONE && g->N > LR_RESIDENT_MAX) {
                throw Error("long range over tiles");
            }
            g->choose_launch(0, 0);
        }
        catch (const Error&) {
            g->streaming = true;
            g->u2 = dalloc<double>(RN);
            g->v2 = dalloc<double>(RN);
            g->a2 = dalloc<double>(RN);
            g->yl = dalloc<double>(RN);
            g->yr = dalloc<double>(RN);
            g->yl2 = dalloc<double>(RN);
            g->yr2 = dalloc<double>(RN);
            g->jl2 = dalloc<int32_t>(RN);
        }
    }

    // tiles of the streaming systems (chains: runs of particles; lattices:
    // whole rows, at most rows - 2 per tile so that the halo rows are distinct)
    if (g->streaming && fq::tile_supported(g->pot, inter, g->ext)) {
        const char* nt = std::getenv("FQSB_NO_TILES");
        bool ok = !(nt && nt[0] == '1');
        std::vector<fq::TileDesc> tl;
        std::vector<int32_t> first((size_t)R + 1, 0);
        size_t lds = 0;
        for (int64_t r = 0; r < R && ok; ++r) {
            const fqsb_desc& d = ds[(size_t)r];
            const int64_t N = Ns[(size_t)r];
            const bool lat = inter == FQSB_INT_LAPLACE2D || inter == FQSB_INT_QUARTICGRADIENT2D;
            const int64_t cols = d.rank == 1 ? 1 : (int64_t)d.cols;
            const int64_t rows = (int64_t)d.rows;
            const bool lr = inter == FQSB_INT_LONGRANGE1D;
            int64_t step;
            if (lat) {
                int64_t rpt = std::max<int64_t>(1, 4096 / cols);
                rpt = std::min<int64_t>(rpt, rows - 2);
                ok = rows >= 3 && rpt >= 1;
                step = rpt * cols;
            }
            else {
                ok = N >= 3;
                step = lr ? LR_TILE_PARTICLES : 4096;
            }
            first[(size_t)r] = (int32_t)tl.size();
            for (int64_t p0 = 0; ok && p0 < N; p0 += step) {
                fq::TileDesc td{};
                td.sys = (int32_t)r;
                td.p0 = p0;
                td.p1 = std::min<int64_t>(N, p0 + step);
                tl.push_back(td);
                // LDS: the tile's slips and halo (long range: every slip of the system)
                const int64_t nl = lr ? N : (td.p1 - td.p0) + 2 * (lat ? cols : 1);
                lds = std::max<size_t>(lds, sizeof(double) * (size_t)nl);
            }
        }
        first[(size_t)R] = (int32_t)tl.size();
        if (ok && lds <= g->lds_limit && !tl.empty()) {
            g->tiled = true;
            g->ntiles = (int)tl.size();
            g->tile_lds = lds;
            g->tiles = dalloc<fq::TileDesc>(tl.size());
            g->h2d(g->tiles, tl.data(), tl.size());
            g->tile_first = dalloc<int32_t>((size_t)R + 1);
            g->h2d(g->tile_first, first.data(), (size_t)R + 1);
            g->tile_part = dalloc<double>(2 * tl.size());
            g->tile_k = 128;
            if (const char* e = std::getenv("FQSB_TILE_K")) {
                g->tile_k = std::max(0, std::min(std::atoi(e), fq::TB_KMAX_STEPS));
            }
            const bool lat = inter == FQSB_INT_LAPLACE2D || inter == FQSB_INT_QUARTICGRADIENT2D;
            if (lat) { // lattices: K halo rows and columns on every side of a tile
                int klat = fq::TL_K_DEFAULT;
                if (const char* e = std::getenv("FQSB_LATTICE_K")) {
                    klat = std::atoi(e);
                }
                g->tb_kmax = std::max(1, std::min(klat, (fq::TL_WIDTH - 1) / 2));
            }
            if (g->tile_k > 0 && fq::tile_block_supported(g->pot, inter)) {
                std::vector<fq::TileDesc> tb;
                std::vector<int32_t> tbf((size_t)R + 1, 0);
                size_t sites = 0;
                const int64_t K = g->tb_kmax;
                for (int64_t r = 0; r < R; ++r) {
                    tbf[(size_t)r] = (int32_t)tb.size();
                    if (lat) {
                        // tiles of H x W sites, (H + 2K) x (W + 2K) <= TL_SITES,
                        // W + 2K <= TL_WIDTH, balanced over the lattice
                        const int64_t cols = (int64_t)ds[(size_t)r].cols, rows = (int64_t)ds[(size_t)r].rows;
                        const int64_t ntc = (cols + (fq::TL_WIDTH - 2 * K) - 1) / (fq::TL_WIDTH - 2 * K);
                        const int64_t W = (cols + ntc - 1) / ntc;
                        const int64_t hmax = fq::TL_SITES / (W + 2 * K) - 2 * K;
                        const int64_t ntr = (rows + hmax - 1) / hmax;
                        const int64_t H = (rows + ntr - 1) / ntr;
                        for (int64_t r0 = 0; r0 < rows; r0 += H) {
                            for (int64_t c0 = 0; c0 < cols; c0 += W) {
                                fq::TileDesc td{};
                                td.sys = (int32_t)r;
                                td.r0 = (int32_t)r0;
                                td.nr = (int32_t)std::min(H, rows - r0);
                                td.c0 = (int32_t)c0;
                                td.nc = (int32_t)std::min(W, cols - c0);
                                td.p0 = r0 * cols + c0;
                                td.p1 = td.p0 + td.nc;
                                tb.push_back(td);
                                sites = std::max<size_t>(sites, (size_t)((td.nr + 2 * K) * (td.nc + 2 * K)));
                            }
                        }
                        continue;
                    }
                    for (int64_t p0 = 0; p0 < Ns[(size_t)r]; p0 += fq::TB_TILE_PARTICLES) {
                        fq::TileDesc td{};
                        td.sys = (int32_t)r;
                        td.p0 = p0;
                        td.p1 = std::min<int64_t>(Ns[(size_t)r], p0 + fq::TB_TILE_PARTICLES);
                        tb.push_back(td);
                    }
                }
                tbf[(size_t)R] = (int32_t)tb.size();
                if (sites > (size_t)fq::TL_SITES) {
                    throw Error("internal: a lattice tile exceeds the block kernel's sites");
                }
                g->tb_lds = 2 * sizeof(double) * sites;
                g->tb_ntiles = (int)tb.size();
                g->tb_tiles = dalloc<fq::TileDesc>(tb.size());
                g->h2d(g->tb_tiles, tb.data(), tb.size());
                g->tb_first = dalloc<int32_t>((size_t)R + 1);
                g->h2d(g->tb_first, tbf.data(), (size_t)R + 1);
                // room for blocks of up to TB_KMAX_STEPS (the block length can be
                // changed later, fqsb_ragged_set_group_block)
                g->tile_part_k = dalloc<double>(2 * (size_t)fq::TB_KMAX_STEPS * tb.size());
                g->tile_fail = dalloc<int32_t>(tb.size());
            }
            g->tile_flag = dalloc<int32_t>((size_t)R);
        }
    }

    // no-passing chains beyond the resident no-passing kernel (8192
    // particles): sweeps over tiles of 4096 particles
    if (g->minim == FQSB_MIN_OVERDAMPED) {
        const char* force = std::getenv("FQSB_NOPASSING_TILES");
        if ((force && force[0] == '1') || g->N > NOPASSING_RESIDENT_MAX) {
            std::vector<fq::TileDesc> tl;
            std::vector<int32_t> first((size_t)R + 1, 0);
            for (int64_t r = 0; r < R; ++r) {
                first[(size_t)r] = (int32_t)tl.size();
                for (int64_t p0 = 0; p0 < Ns[(size_t)r]; p0 += NP_TILE_PARTICLES) {
                    fq::TileDesc td{};
                    td.sys = (int32_t)r;
                    td.p0 = p0;
                    td.p1 = std::min<int64_t>(Ns[(size_t)r], p0 + NP_TILE_PARTICLES);
                    tl.push_back(td);
                }
            }
            first[(size_t)R] = (int32_t)tl.size();
            g->np_tiled = true;
            g->ntiles = (int)tl.size();
            g->tiles = dalloc<fq::TileDesc>(tl.size());
            g->h2d(g->tiles, tl.data(), tl.size());
            g->tile_first = dalloc<int32_t>((size_t)R + 1);
            g->h2d(g->tile_first, first.data(), (size_t)R + 1);
            g->tile_part = dalloc<double>(2 * tl.size());
            g->u2 = dalloc<double>(RN);
            g->yl = dalloc<double>(RN);
            g->yr = dalloc<double>(RN);
        }
    }

    g->hviews.resize((size_t)R);
    g->hrs.assign((size_t)R, fq::RunState{});
    for (int64_t r = 0; r < R; ++r) {
        const fqsb_desc& d = ds[(size_t)r];
        const int64_t N = Ns[(size_t)r];
        // distribution parameters with the reference's defaults (see fq_common.hpp)
        const int maxpar = (d.distribution == FQSB_DIST_POWER || d.distribution == FQSB_DIST_PARETO ||
                            d.distribution == FQSB_DIST_WEIBULL)
                               ? 3
                               : 2;
        double p[3] = {1.0, 0.0, 0.0};
        if (maxpar == 3) {
            p[1] = 1.0;
        }
        for (int i = 0; i < d.nparameters; ++i) {
            p[i] = d.parameters[i];
        }
        // alignment (own rule standing in for prrng::alignment(2, 30, 6, false))
        const int64_t n = (int64_t)d.nchunk;
        int64_t bb = d.align_buffer < 0 ? 0 : d.align_buffer;
        bb = std::min<int64_t>(bb, (n - 2) / 4);
        int64_t mm = std::max<int64_t>(d.align_margin, bb);
        mm = std::min<int64_t>(mm, n - 2 - bb);

        fq::SysView& S = g->hviews[(size_t)r];
        std::memset(&S, 0, sizeof(S));
        S.N = N;
        S.rows = (int64_t)d.rows;
        S.cols = d.rank == 1 ? 1 : (int64_t)d.cols;
        S.pot = g->pot;
        S.inter = inter;
        S.ext = g->ext;
        S.m = d.m;
        S.inv_m = 1.0 / d.m;
        S.eta = d.eta;
        S.mu = d.mu;
        S.kappa = d.kappa;
        S.k_frame = d.k_frame;
        S.dt = d.dt;
        S.k = d.k_interactions;
        S.a1 = d.a1;
        S.a2 = d.a2;
        S.k2 = d.k2;
        S.k4 = d.k4;
        S.k4_4 = 0.25 * d.k4;
        S.alpha = d.alpha;
        S.lr_m = (N - N % 2) / 2;
        const size_t o = (size_t)g->off[(size_t)r];
        S.pref = g->pref ? g->pref + o : nullptr;
        S.u = g->u + o;
        S.v = g->v + o;
        S.a = g->a + o;
        S.f = g->f + o;
        S.f_pot = g->f_pot + o;
        S.f_int = g->f_int + o;
        S.f_frame = g->f_frame + o;
        S.f_damp = g->f_damp + o;
        S.f_th = g->f_th ? g->f_th + o : nullptr;
        S.i_ref = g->i_ref + o;
        S.next = g->next ? g->next + o : nullptr;
        S.dinc = g->dinc ? g->dinc + o : nullptr;
        S.mean = d.mean;
        S.stddev = d.stddev;
        S.seed = d.seed;
        S.offset = d.offset;
        if (g->np_tiled) {
            S.u_b = g->u2 + o;
            S.yl_a = g->yl + o;
            S.yr_a = g->yr + o;
        }
        if (g->streaming) {
            S.u_b = g->u2 + o;
            S.v_b = g->v2 + o;
            S.a_b = g->a2 + o;
            S.yl_a = g->yl + o;
            S.yr_a = g->yr + o;
            S.yl_b = g->yl2 + o;
            S.yr_b = g->yr2 + o;
            S.jl_b = g->jl2 + o;
        }
        S.ls.y = g->y + yoff[(size_t)r];
        S.ls.start = g->start + o;
        S.ls.jloc = g->jloc + o;
        S.ls.st_begin = g->st_begin + o;
        S.ls.st_end = g->st_end + o;
        S.ls.nchunk = n;
        S.ls.buffer = bb;
        S.ls.margin = mm;
        S.ls.inc = 1u; // (initseq 0 << 1) | 1
        S.ls.dist = d.distribution;
        S.ls.p0 = p[0];
        S.ls.p1 = p[1];
        S.ls.p2 = p[2];
        S.rs = g->rs + r;
        fq::RunState& rs = g->hrs[(size_t)r];
        rs.inc = 0;
        rs.u_frame = 0.0;
        rs.prev = std::numeric_limits<double>::infinity();
        rs.status = fq::ST_RUNNING;
        uint64_t rinc = 0;
        rs.rng_state = default_stream_state(d.seed_forcing, rinc);
        rs.rng_inc = rinc;
        rs.min_next = std::numeric_limits<int64_t>::max();
    }
    g->h2d(g->views, g->hviews.data(), (size_t)R);
    g->push();
    HIPC(fq::launch_landscape_init(g->views, (int)R, g->N, g->stream));
    g->refresh();
    return g.release();
}

// R replicas of one description: replica r uses seed + r * seed_stride (and
// seed_forcing + r * seed_stride).
Group* build_group(const fqsb_desc* dsc, uint64_t replicas, uint64_t seed_stride)
{
    if (!dsc) {
        throw Error("null description");
    }
    if (replicas < 1) {
        throw Error("at least one replica is needed");
    }
    std::vector<fqsb_desc> ds((size_t)replicas, *dsc);
    for (uint64_t r = 0; r < replicas; ++r) {
        ds[(size_t)r].seed = dsc->seed + r * seed_stride;
        ds[(size_t)r].seed_forcing = dsc->seed_forcing + r * seed_stride;
    }
    return build_group_multi(ds);
}

int64_t checked_p(const Group& g, uint64_t p)
{
    if ((int64_t)p >= g.N) {
        throw Error("particle index out of range");
    }
    return (int64_t)p;
}

// No-passing minimise (detail.h:1694-1753) on the device (Jacobi sweeps) for
// every system of the group (with g.mask: the masked-in ones). Returns per
// system true if the StopList converged.
// Tiled no-passing run (chains beyond the resident kernel): per sweep one
// k_np_sweep, k_np_resid and k_np_decide for every system of the group,
// captured in blocks of NP_BLOCK sweeps as a HIP graph (a replay continues
// from the run state in device memory; sweeps of a stopped system return at
// once), checked by the host after every block.
static void nopassing_tiled(Group& g, const fq::RunArgs& ra, uint64_t max_iter)
{
    const int64_t target = (int64_t)std::min<uint64_t>(max_iter, (uint64_t)std::numeric_limits<int64_t>::max());
    for (auto& r : g.hrs) {
        r.target = target;
        r.cur = 0;
    }
    g.push();
    HIPC(fq::launch_tile_prologue(g.views, g.tiles, g.ntiles, g.stream));
    auto enqueue = [&](uint64_t nb) {
        for (uint64_t s = 0; s < nb; ++s) {
            HIPC(fq::launch_np_sweep(g.views, g.tiles, g.ntiles, (int)g.R, g.tile_first, g.tile_part, ra, g.stream));
        }
    };
    auto launch_block = [&](uint64_t nb) {
        if (!g.use_graphs) {
            enqueue(nb);
            return;
        }
        for (const auto& tg : g.np_graphs) {
            if (tg.steps == nb && std::memcmp(&tg.ra, &ra, sizeof(ra)) == 0) {
                HIPC(hipGraphLaunch(tg.exec, g.stream));
                return;
            }
        }
        hipGraph_t graph = nullptr;
        HIPC(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal));
        try {
            enqueue(nb);
        }
        catch (...) {
            (void)hipStreamEndCapture(g.stream, &graph);
            if (graph) {
                (void)hipGraphDestroy(graph);
            }
            throw;
        }
        HIPC(hipStreamEndCapture(g.stream, &graph));
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPC(e);
        if (g.np_graphs.size() >= 8) {
            (void)hipGraphExecDestroy(g.np_graphs.front().exec);
            g.np_graphs.erase(g.np_graphs.begin());
        }
        g.np_graphs.push_back(Group::TileGraph{ra, nb, exec});
        HIPC(hipGraphLaunch(exec, g.stream));
    };
    while (true) {
        int64_t lag = 0;
        for (const auto& r : g.hrs) {
            if ((r.status == fq::ST_RUNNING || r.status == fq::ST_BUDGET) && r.steps < r.target) {
                lag = std::max<int64_t>(lag, r.target - r.steps);
            }
        }
        if (lag == 0) {
            break;
        }
        launch_block(std::min<uint64_t>(NP_BLOCK, (uint64_t)lag));
        g.pull();
    }
    HIPC(fq::launch_np_finish(g.views, g.tiles, g.ntiles, g.stream));
    for (auto& r : g.hrs) {
        r.cur = 0;
        if (r.status == fq::ST_RUNNING) {
            r.status = fq::ST_BUDGET;
        }
    }
    g.push();
    HIPC(hipStreamSynchronize(g.stream));
}

// particle-sweeps of the last no-passing run (ragged sets' work measure)
static void count_sweeps(Group& g)
{
    for (int64_t r = 0; r < g.R; ++r) {
        g.busy_ps += (double)(g.off[(size_t)r + 1] - g.off[(size_t)r]) * (double)g.hrs[(size_t)r].steps;
    }
}

std::vector<bool> nopassing_run(Group& g, double tol, uint64_t niter_tol, uint64_t max_iter)
{
    for (auto& r : g.hrs) {
        r.qs_first = r.inc;
        r.qs_last = r.inc;
    }
    g.reset_run(fq::MODE_MINIMISE);
    g.push();
    if (g.np_tiled) {
        fq::RunArgs ra{};
        ra.mode = fq::MODE_MINIMISE;
        ra.tol = tol;
        ra.tol2 = tol * tol;
        ra.niter_tol = (int64_t)niter_tol;
        nopassing_tiled(g, ra, max_iter);
        count_sweeps(g);
        std::vector<bool> conv((size_t)g.R);
        for (int64_t r = 0; r < g.R; ++r) {
            conv[(size_t)r] = g.hrs[(size_t)r].status == fq::ST_CONVERGED;
        }
        return conv;
    }
    // particles per thread: the fewest for which the launch fits the kernel's
    // thread bound (1024 up to 8 particles per thread, 512 at 16)
    int ppt = 1;
    while (ppt <= 16 && (g.N + ppt - 1) / ppt > (ppt <= 8 ? 1024 : 512)) {
        ppt *= 2;
    }
    if (ppt > 16) {
        throw Error("system too large for the no-passing kernel (at most 8192 particles)");
    }
    const int64_t need = (g.N + ppt - 1) / ppt;
    const int threads = (int)(((need + 63) / 64) * 64);
    fq::RunArgs ra{};
    ra.mode = fq::MODE_MINIMISE;
    ra.tol = tol;
    ra.tol2 = tol * tol;
    ra.niter_tol = (int64_t)niter_tol;
    uint64_t done = 0;
    auto running = [&] {
        for (const auto& r : g.hrs) {
            if (r.status == fq::ST_RUNNING || r.status == fq::ST_BUDGET) {
                return true;
            }
        }
        return false;
    };
    while (done < max_iter && running()) {
        const uint64_t nb = std::min<uint64_t>(MIN_BLOCK, max_iter - done);
        ra.max_steps = (int64_t)nb;
        if (!fq::launch_nopassing(ppt, threads, sizeof(double) * 4 * threads, g.views, (int)g.R, ra, g.stream)) {
            throw Error("no-passing kernel not available");
        }
        HIPC(hipGetLastError());
        done += nb;
        g.pull();
    }
    count_sweeps(g);
    std::vector<bool> conv((size_t)g.R);
    for (int64_t r = 0; r < g.R; ++r) {
        conv[(size_t)r] = g.hrs[(size_t)r].status == fq::ST_CONVERGED;
    }
    return conv;
}

// Batch state for checkpoints: where each FQSB_STATE_* item lives.
struct StateItem {
    void* dev;       // device array, or null for a per-system scalar in RunState
    size_t count;    // elements
    int bytes;       // bytes per element
    size_t rs_off;   // offset of the scalar in RunState (dev == null)
};

StateItem state_item(Group& g, int what)
{
    const size_t T = (size_t)g.total, R = (size_t)g.R;
    auto thermal = [&](void* p) {
        if (!g.ext) {
            throw Error("system has no external forcing");
        }
        return p;
    };
    switch (what) {
        case FQSB_STATE_Y:
            return {g.y, T * (size_t)g.nchunk, 8, 0};
        case FQSB_STATE_START:
            return {g.start, T, 8, 0};
        case FQSB_STATE_JLOC:
            return {g.jloc, T, 4, 0};
        case FQSB_STATE_ST_BEGIN:
            return {g.st_begin, T, 8, 0};
        case FQSB_STATE_ST_END:
            return {g.st_end, T, 8, 0};
        case FQSB_STATE_F_THERMAL:
            return {thermal(g.f_th), T, 8, 0};
        case FQSB_STATE_NEXT:
            return {thermal(g.next), T, 8, 0};
        case FQSB_STATE_INC:
            return {nullptr, R, 8, offsetof(fq::RunState, inc)};
        case FQSB_STATE_U_FRAME:
            return {nullptr, R, 8, offsetof(fq::RunState, u_frame)};
        case FQSB_STATE_RNG:
            return {nullptr, R, 8, offsetof(fq::RunState, rng_state)};
        case FQSB_STATE_QS_FIRST:
            return {nullptr, R, 8, offsetof(fq::RunState, qs_first)};
        case FQSB_STATE_QS_LAST:
            return {nullptr, R, 8, offsetof(fq::RunState, qs_last)};
        default:
            throw Error("unknown state item");
    }
}

} // namespace

struct fqsb_system {
    std::unique_ptr<Group> g;
};

struct fqsb_batch {
    std::unique_ptr<Group> g;
};

// A ragged set: systems of any classes and sizes, held in groups that share
// a kernel instantiation and a size class (sizes within a factor 2, so one
// launch shape fits every system of a group). Operations run on all groups
// at once, one host thread and one HIP stream per group.
struct fqsb_ragged {
    std::vector<std::unique_ptr<Group>> groups;
    std::vector<std::pair<size_t, int64_t>> where; // system i -> (group, index in group)
    std::vector<int64_t> sizes;                    // N of system i
    std::vector<int64_t> uoff;                     // first element of system i in concatenated arrays
    bool serial = false;                           // groups one after the other (fqsb_ragged_set_serial)

    // the group's part of an operation between two events on its stream, and
    // the particle-steps it took (increments of inc; no-passing sweeps are
    // counted where they run)
    template <class F>
    static void timed(Group& g, F&& op)
    {
        for (auto& e : g.evr) {
            if (!e) {
                HIPC(hipEventCreate(&e));
            }
        }
        std::vector<int64_t> inc0(g.hrs.size());
        for (size_t k = 0; k < g.hrs.size(); ++k) {
            inc0[k] = g.hrs[k].inc;
        }
        HIPC(hipEventRecord(g.evr[0], g.stream));
        op();
        HIPC(hipEventRecord(g.evr[1], g.stream));
        HIPC(hipEventSynchronize(g.evr[1]));
        float ms = 0.0f;
        HIPC(hipEventElapsedTime(&ms, g.evr[0], g.evr[1]));
        g.busy_ms += ms;
        if (g.minim != FQSB_MIN_OVERDAMPED) {
            for (size_t k = 0; k < g.hrs.size(); ++k) {
                g.busy_ps += (double)(g.off[k + 1] - g.off[k]) * (double)(g.hrs[k].inc - inc0[k]);
            }
        }
    }

    // serial (or FQSB_RAGGED_SERIAL=1): run the groups one after the other
    // from the calling thread (each group's time is then its own; also for
    // tools that do not follow launches from many threads)
    template <class F>
    void each(F&& fn)
    {
        const char* sr = std::getenv("FQSB_RAGGED_SERIAL");
        if (serial || (sr && sr[0] == '1')) {
            for (size_t gi = 0; gi < groups.size(); ++gi) {
                HIPC(hipSetDevice(groups[gi]->device));
                timed(*groups[gi], [&] { fn(*groups[gi], gi); });
            }
            return;
        }
        std::vector<std::exception_ptr> err(groups.size());
        std::vector<std::thread> th;
        th.reserve(groups.size());
        for (size_t gi = 0; gi < groups.size(); ++gi) {
            th.emplace_back([&, gi] {
                try {
                    HIPC(hipSetDevice(groups[gi]->device));
                    timed(*groups[gi], [&] { fn(*groups[gi], gi); });
                }
                catch (...) {
                    err[gi] = std::current_exception();
                }
            });
        }
        for (auto& t : th) {
            t.join();
        }
        for (auto& e : err) {
            if (e) {
                std::rethrow_exception(e);
            }
        }
    }

    // per-group masks from a per-system mask (nullptr: all systems)
    std::vector<std::vector<uint8_t>> masks(const uint8_t* active) const
    {
        std::vector<std::vector<uint8_t>> m(groups.size());
        for (size_t gi = 0; gi < groups.size(); ++gi) {
            m[gi].assign((size_t)groups[gi]->R, active ? 0 : 1);
        }
        if (active) {
            for (size_t i = 0; i < where.size(); ++i) {
                m[where[i].first][(size_t)where[i].second] = active[i] ? 1 : 0;
            }
        }
        return m;
    }
};

extern "C" {

const char* fqsb_last_error(void)
{
    return g_err.c_str();
}

const char* fqsb_version(void)
{
    return "frictionqpotspringblock_amd 0.1.0 (gfx950)";
}

int fqsb_create(const fqsb_desc* desc, fqsb_system** out)
{
    return guard([&] {
        *out = nullptr;
        std::unique_ptr<fqsb_system> s(new fqsb_system());
        s->g.reset(build_group(desc, 1, 0));
        *out = s.release();
    });
}

int fqsb_destroy(fqsb_system* s)
{
    delete s;
    return 0;
}

uint64_t fqsb_size(const fqsb_system* s)
{
    return (uint64_t)s->g->N;
}

int fqsb_get(fqsb_system* s, int field, double* out)
{
    return guard([&] {
        Group& g = *s->g;
        g.d2h(out, g.field(field), (size_t)g.N);
    });
}

int fqsb_set(fqsb_system* s, int field, const double* in)
{
    return guard([&] {
        Group& g = *s->g;
        switch (field) {
            case FQSB_U:
                g.h2d(g.u, in, (size_t)g.N);
                g.updated_u();
                break;
            case FQSB_V:
                g.h2d(g.v, in, (size_t)g.N);
                g.updated_v();
                break;
            case FQSB_A:
                g.h2d(g.a, in, (size_t)g.N);
                break;
            case FQSB_F_THERMAL:
                g.h2d(g.field(FQSB_F_THERMAL), in, (size_t)g.N);
                break;
            default:
                throw Error("field is read-only");
        }
    });
}

int fqsb_get_inc(const fqsb_system* s, int64_t* out)
{
    *out = s->g->hrs[0].inc;
    return 0;
}

int fqsb_set_inc(fqsb_system* s, int64_t inc)
{
    return guard([&] {
        Group& g = *s->g;
        g.hrs[0].inc = inc;
        g.hrs[0].qs_first = inc;
        g.hrs[0].qs_last = inc;
        g.updated_inc();
    });
}

int fqsb_set_t(fqsb_system* s, double t)
{
    Group& g = *s->g;
    g.hrs[0].inc = (int64_t)std::round(t / g.d.dt);
    return 0;
}

int fqsb_get_u_frame(const fqsb_system* s, double* out)
{
    *out = s->g->hrs[0].u_frame;
    return 0;
}

int fqsb_set_u_frame(fqsb_system* s, double u_frame)
{
    return guard([&] {
        Group& g = *s->g;
        g.hrs[0].u_frame = u_frame;
        g.forces(fq::FF_FRAME | fq::FF_TOTAL);
    });
}

int fqsb_quasistatic_activity(const fqsb_system* s, int64_t* first, int64_t* last)
{
    *first = s->g->hrs[0].qs_first;
    *last = s->g->hrs[0].qs_last;
    return 0;
}

static double norm_sq(const std::vector<double>& x)
{
    double r = 0.0;
    for (double v : x) {
        r += v * v;
    }
    return r;
}

int fqsb_residual(fqsb_system* s, double* out)
{
    return guard([&] {
        Group& g = *s->g;
        std::vector<double> f((size_t)g.N), ff((size_t)g.N);
        g.d2h(f.data(), g.f, f.size());
        g.d2h(ff.data(), g.f_frame, ff.size());
        const double r_fres = std::sqrt(norm_sq(f));
        const double r_fext = std::sqrt(norm_sq(ff));
        *out = r_fext != 0.0 ? r_fres / r_fext : r_fres;
    });
}

int fqsb_temperature(fqsb_system* s, double* out)
{
    return guard([&] {
        Group& g = *s->g;
        std::vector<double> v((size_t)g.N);
        g.d2h(v.data(), g.v, v.size());
        *out = 0.5 * g.d.m * norm_sq(v) / (double)g.N;
    });
}

int fqsb_refresh(fqsb_system* s)
{
    return guard([&] { s->g->refresh(); });
}

int fqsb_quench(fqsb_system* s)
{
    return guard([&] { s->g->forces(fq::FF_QUENCH | fq::FF_DAMP | fq::FF_TOTAL); });
}

int fqsb_time_steps(fqsb_system* s, uint64_t n)
{
    return guard([&] {
        if (n > 0) {
            s->g->steps(n, false, 0.0, 0, true, false);
        }
    });
}

int fqsb_time_step(fqsb_system* s)
{
    return fqsb_time_steps(s, 1);
}

int fqsb_flow_steps(fqsb_system* s, uint64_t n, double v_frame)
{
    return guard([&] {
        if (n > 0) {
            s->g->steps(n, true, v_frame, 0, true, false);
        }
    });
}

int fqsb_time_steps_until_event(fqsb_system* s, double tol, uint64_t niter_tol, uint64_t max_iter, uint64_t* ret)
{
    return guard([&] {
        Group& g = *s->g;
        g.set_iref_current();
        fq::RunArgs ra{};
        ra.mode = fq::MODE_UNTIL_EVENT;
        ra.tol = tol;
        ra.tol2 = tol * tol;
        ra.niter_tol = (int64_t)niter_tol;
        g.run(ra, max_iter, MIN_BLOCK, false);
        const auto& r = g.hrs[0];
        if (r.status == fq::ST_NAN) {
            throw Error("NaN entries found");
        }
        if (r.status == fq::ST_EVENT) {
            *ret = (uint64_t)r.steps;
        }
        else if (r.status == fq::ST_CONVERGED) {
            *ret = 0;
        }
        else {
            *ret = max_iter + 1;
        }
    });
}


static void minimise_nopassing(Group& g, double tol, uint64_t niter_tol, uint64_t max_iter, int max_iter_is_error,
                               uint64_t* ret)
{
    if (nopassing_run(g, tol, niter_tol, max_iter)[0]) {
        *ret = 0;
        return;
    }
    if (max_iter_is_error) {
        throw Error("No convergence found");
    }
    *ret = max_iter + 1;
}

int fqsb_minimise(fqsb_system* s, double tol, uint64_t niter_tol, uint64_t max_iter, int time_activity,
                  int max_iter_is_error, uint64_t* ret)
{
    return guard([&] {
        Group& g = *s->g;
        if (g.minim == FQSB_MIN_NONE) {
            throw Error("Minimisation not implementated");
        }
        if (g.minim == FQSB_MIN_OVERDAMPED) {
            minimise_nopassing(g, tol, niter_tol, max_iter, max_iter_is_error, ret);
            return;
        }
        if (time_activity) {
            g.set_iref_current();
        }
        fq::RunArgs ra{};
        ra.mode = fq::MODE_MINIMISE;
        ra.time_activity = time_activity ? 1 : 0;
        ra.tol = tol;
        ra.tol2 = tol * tol;
        ra.niter_tol = (int64_t)niter_tol;
        g.run(ra, max_iter, MIN_BLOCK, false);
        const auto& r = g.hrs[0];
        if (r.status == fq::ST_NAN) {
            throw Error("NaN entries found");
        }
        if (r.status == fq::ST_CONVERGED) {
            *ret = 0;
            return;
        }
        if (max_iter_is_error) {
            throw Error("No convergence found");
        }
        *ret = max_iter + 1;
    });
}

int fqsb_minimise_truncate(fqsb_system* s, const int64_t* i_n, uint64_t A_truncate, uint64_t S_truncate, double tol,
                           uint64_t niter_tol, uint64_t max_iter, int time_activity, int max_iter_is_error,
                           uint64_t* ret)
{
    (void)time_activity;
    return guard([&] {
        s->g->minimise_truncate(i_n, A_truncate, S_truncate, tol, niter_tol, max_iter, max_iter_is_error != 0, ret);
    });
}

int fqsb_max_uniform_displacement(fqsb_system* s, int direction, double* ret)
{
    return guard([&] { *ret = s->g->max_uniform(0, direction); });
}

int fqsb_event_driven_step(fqsb_system* s, double eps, int kick, int direction, double* ret)
{
    return guard([&] {
        Group& g = *s->g;
        if (g.ext) {
            throw Error("eventDrivenStep is not available for this system");
        }
        *ret = g.event_driven_step(eps, kick != 0, direction)[0];
    });
}

int fqsb_quasistatic(fqsb_system* s, int64_t first_event, int64_t nevents, double eps, int direction, int kicks,
                     double tol, uint64_t niter_tol, uint64_t max_iter, int64_t* S, double* u_frame, double* f_frame,
                     int64_t* inc)
{
    return guard([&] {
        s->g->quasistatic(first_event, nevents, eps, direction, kicks, tol, niter_tol, max_iter, S, u_frame, f_frame,
                          inc);
    });
}

int fqsb_trigger(fqsb_system* s, uint64_t p, double eps, int direction)
{
    return guard([&] {
        Group& g = *s->g;
        const int64_t pp = checked_p(g, p);
        HIPC(fq::launch_trigger(g.views, pp, eps, direction, g.pot == FQSB_POT_CUSPY ? 1 : 0, g.stream));
        HIPC(hipStreamSynchronize(g.stream));
    });
}

int fqsb_advance_to_fixed_force(fqsb_system* s, double f_frame, int allow_plastic)
{
    return guard([&] { s->g->advance_to_fixed_force(&f_frame, allow_plastic != 0); });
}

uint64_t fqsb_chunk_size(const fqsb_system* s)
{
    return (uint64_t)s->g->nchunk;
}

int fqsb_chunk_data(fqsb_system* s, double* out)
{
    return guard([&] {
        Group& g = *s->g;
        g.d2h(out, g.y, (size_t)(g.N * g.nchunk));
    });
}

int fqsb_chunk_index_at_align(fqsb_system* s, int64_t* out)
{
    return guard([&] {
        Group& g = *s->g;
        std::vector<int32_t> jl((size_t)g.N);
        g.d2h(out, g.start, (size_t)g.N);
        g.d2h(jl.data(), g.jloc, (size_t)g.N);
        for (int64_t p = 0; p < g.N; ++p) {
            out[p] += jl[(size_t)p];
        }
    });
}

int fqsb_chunk_chunk_index_at_align(fqsb_system* s, int64_t* out)
{
    return guard([&] {
        Group& g = *s->g;
        std::vector<int32_t> jl((size_t)g.N);
        g.d2h(jl.data(), g.jloc, (size_t)g.N);
        for (int64_t p = 0; p < g.N; ++p) {
            out[p] = jl[(size_t)p];
        }
    });
}

int fqsb_chunk_start(fqsb_system* s, int64_t* out)
{
    return guard([&] {
        Group& g = *s->g;
        g.d2h(out, g.start, (size_t)g.N);
    });
}

int fqsb_chunk_state_at(fqsb_system* s, const int64_t* index, uint64_t* out)
{
    return guard([&] {
        Group& g = *s->g;
        std::vector<int64_t> st((size_t)g.N);
        std::vector<uint64_t> sb((size_t)g.N);
        g.d2h(st.data(), g.start, st.size());
        g.d2h(sb.data(), g.st_begin, sb.size());
        for (int64_t p = 0; p < g.N; ++p) {
            out[p] = fq::pcg_advance(sb[(size_t)p], 1u, index[p] - st[(size_t)p]);
        }
    });
}

int fqsb_chunk_restore(fqsb_system* s, const uint64_t* state, const double* value, const int64_t* index)
{
    return guard([&] {
        Group& g = *s->g;
        const size_t N = (size_t)g.N;
        uint64_t* dstate = dalloc<uint64_t>(N);
        double* dvalue = dalloc<double>(N);
        int64_t* dindex = dalloc<int64_t>(N);
        try {
            g.h2d(dstate, state, N);
            g.h2d(dvalue, value, N);
            g.h2d(dindex, index, N);
            HIPC(fq::launch_restore(g.views, g.N, dstate, dvalue, dindex, g.stream));
            HIPC(hipStreamSynchronize(g.stream));
        }
        catch (...) {
            dfree(dstate);
            dfree(dvalue);
            dfree(dindex);
            throw;
        }
        dfree(dstate);
        dfree(dvalue);
        dfree(dindex);
    });
}

int fqsb_thermal_state(const fqsb_system* s, uint64_t* out)
{
    *out = s->g->hrs[0].rng_state;
    return 0;
}

int fqsb_thermal_set_state(fqsb_system* s, uint64_t state)
{
    s->g->hrs[0].rng_state = state;
    return 0;
}

int fqsb_thermal_next(fqsb_system* s, int64_t* out)
{
    return guard([&] {
        Group& g = *s->g;
        if (!g.ext) {
            throw Error("system has no external forcing");
        }
        g.d2h(out, g.next, (size_t)g.N);
    });
}

int fqsb_thermal_set_next(fqsb_system* s, const int64_t* in)
{
    return guard([&] {
        Group& g = *s->g;
        if (!g.ext) {
            throw Error("system has no external forcing");
        }
        g.h2d(g.next, in, (size_t)g.N);
    });
}

// --- batch -------------------------------------------------------------------

int fqsb_batch_create(const fqsb_desc* desc, uint64_t replicas, uint64_t seed_stride, fqsb_batch** out)
{
    return guard([&] {
        *out = nullptr;
        std::unique_ptr<fqsb_batch> b(new fqsb_batch());
        b->g.reset(build_group(desc, replicas, seed_stride));
        if (b->g->inter == FQSB_INT_LONGRANGE1D && b->g->pot == FQSB_POT_CUSPY) {
            b->g->set_lr_method(FQSB_LR_MATRIX);
        }
        *out = b.release();
    });
}

int fqsb_batch_set_longrange_method(fqsb_batch* b, int method)
{
    return guard([&] { b->g->set_lr_method(method); });
}

int fqsb_batch_longrange_product_ms(const fqsb_batch* b, float* ms)
{
    return guard([&] {
        if (b->g->inter != FQSB_INT_LONGRANGE1D) {
            throw Error("the systems have no long-range interactions");
        }
        *ms = b->g->lr_product_ms;
    });
}

int fqsb_batch_get_longrange_method(const fqsb_batch* b, int* method)
{
    return guard([&] {
        if (b->g->inter != FQSB_INT_LONGRANGE1D) {
            throw Error("the systems have no long-range interactions");
        }
        *method = b->g->lr_method;
    });
}

int fqsb_batch_destroy(fqsb_batch* b)
{
    delete b;
    return 0;
}

uint64_t fqsb_batch_replicas(const fqsb_batch* b)
{
    return (uint64_t)b->g->R;
}

int fqsb_batch_set_launch(fqsb_batch* b, int ppt, int threads)
{
    return guard([&] {
        if (b->g->streaming) {
            throw Error("the replicas are too large for a resident workgroup; they use the streaming path");
        }
        b->g->choose_launch(ppt, threads);
    });
}

int fqsb_batch_get_launch(const fqsb_batch* b, int* ppt, int* threads)
{
    *ppt = b->g->ppt;
    *threads = b->g->threads;
    return 0;
}

int fqsb_batch_flow_steps(fqsb_batch* b, uint64_t n, double v_frame, uint64_t block, float* ms)
{
    return guard([&] {
        b->g->adj_valid = false;
        float t = b->g->steps(n, true, v_frame, block, false, ms != nullptr);
        if (ms) {
            *ms = t;
        }
    });
}

int fqsb_batch_time_steps(fqsb_batch* b, uint64_t n, uint64_t block, float* ms)
{
    return guard([&] {
        b->g->adj_valid = false;
        float t = b->g->steps(n, false, 0.0, block, false, ms != nullptr);
        if (ms) {
            *ms = t;
        }
    });
}

int fqsb_batch_minimise(fqsb_batch* b, double tol, uint64_t niter_tol, uint64_t max_iter, int max_iter_is_error,
                        uint64_t* ret)
{
    return guard([&] {
        b->g->adj_valid = false;
        Group& g = *b->g;
        if (g.minim != FQSB_MIN_INERTIAL) {
            throw Error("batched minimise needs inertial systems");
        }
        fq::RunArgs ra{};
        ra.mode = fq::MODE_MINIMISE;
        ra.tol = tol;
        ra.tol2 = tol * tol;
        ra.niter_tol = (int64_t)niter_tol;
        g.minimise_run(ra, max_iter);
        for (int64_t r = 0; r < g.R; ++r) {
            const auto& st = g.hrs[(size_t)r];
            if (st.status == fq::ST_NAN) {
                throw Error("NaN entries found");
            }
            if (st.status == fq::ST_CONVERGED) {
                ret[r] = 0;
            }
            else if (max_iter_is_error) {
                throw Error("No convergence found");
            }
            else {
                ret[r] = max_iter + 1;
            }
        }
    });
}

int fqsb_batch_event_driven_step(fqsb_batch* b, double eps, int kick, int direction, double* ret)
{
    return guard([&] {
        b->g->adj_valid = false;
        Group& g = *b->g;
        if (g.ext) {
            throw Error("eventDrivenStep is not available for this system");
        }
        auto r = g.event_driven_step(eps, kick != 0, direction);
        std::copy(r.begin(), r.end(), ret);
    });
}

int fqsb_batch_quasistatic(fqsb_batch* b, int64_t first_event, int64_t nevents, double eps, int direction, int kicks,
                           double tol, uint64_t niter_tol, uint64_t max_iter, int64_t* S, double* u_frame,
                           double* f_frame, int64_t* inc)
{
    return guard([&] {
        b->g->adj_valid = false;
        b->g->quasistatic(first_event, nevents, eps, direction, kicks, tol, niter_tol, max_iter, S, u_frame, f_frame,
                          inc);
    });
}

// Address of `scratch` on each device, taken once per device with that device
// current (the array's name is a host-side stand-in, not a device address).
static std::once_flag g_scratch_once[64];
static void* g_scratch_ptr[64];
static hipError_t g_scratch_err[64];

static void* scratch_on(int device)
{
    if (device < 0 || device >= 64) {
        throw Error("device index out of range for the quasistatic scratch");
    }
    std::call_once(g_scratch_once[device], [device] {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess && cur != device) {
            e = hipSetDevice(device);
        }
        void* p = nullptr;
        if (e == hipSuccess) {
            e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        }
        if (cur != device) {
            (void)hipSetDevice(cur);
        }
        g_scratch_err[device] = e;
        g_scratch_ptr[device] = e == hipSuccess ? p : nullptr;
    });
    HIPC(g_scratch_err[device]);
    return g_scratch_ptr[device];
}

int fqsb_batch_quasistatic_noalloc(fqsb_batch* b, int64_t first_event, int64_t nevents, double eps, int direction,
                                   int kicks, double tol, uint64_t niter_tol, uint64_t max_iter, int64_t* S,
                                   double* u_frame, double* f_frame, int64_t* inc)
{
    return guard([&] {
        Group& g = *b->g;
        const size_t need = 4 * sizeof(int64_t) * (size_t)g.R * (size_t)std::max<int64_t>(nevents, 0);
        if (need > QS_SCRATCH_BYTES) {
            throw Error("fqsb_batch_quasistatic_noalloc holds at most " +
                        std::to_string(QS_SCRATCH_BYTES / (4 * sizeof(int64_t))) +
                        " records (replicas x events) per call; " + std::to_string(g.R * nevents) + " were asked for");
        }
        void* rec = scratch_on(g.device);
        g.adj_valid = false;
        g.quasistatic(first_event, nevents, eps, direction, kicks, tol, niter_tol, max_iter, S, u_frame, f_frame, inc,
                      rec);
    });
}

int fqsb_batch_quasistatic_stats(const fqsb_batch* b, float* ms, int64_t* launches)
{
    return guard([&] {
        *ms = b->g->qs_ms;
        *launches = b->g->qs_launches;
    });
}

int fqsb_batch_minimise_truncate(fqsb_batch* b, const int64_t* i_n, uint64_t A_truncate, uint64_t S_truncate,
                                 double tol, uint64_t niter_tol, uint64_t max_iter, int max_iter_is_error,
                                 uint64_t* ret)
{
    return guard([&] {
        b->g->adj_valid = false;
        b->g->minimise_truncate(i_n, A_truncate, S_truncate, tol, niter_tol, max_iter, max_iter_is_error != 0, ret);
    });
}

int fqsb_batch_trigger(fqsb_batch* b, const uint64_t* p, double eps, int direction)
{
    return guard([&] {
        b->g->adj_valid = false;
        b->g->trigger_each(p, eps, direction);
    });
}

int fqsb_batch_advance_to_fixed_force(fqsb_batch* b, const double* f_frame, int allow_plastic)
{
    return guard([&] {
        b->g->adj_valid = false;
        b->g->advance_to_fixed_force(f_frame, allow_plastic != 0);
    });
}


int fqsb_batch_state_size(fqsb_batch* b, int what, uint64_t* count, int* bytes)
{
    return guard([&] {
        const StateItem it = state_item(*b->g, what);
        *count = (uint64_t)it.count;
        *bytes = it.bytes;
    });
}

int fqsb_batch_get_state(fqsb_batch* b, int what, void* out)
{
    return guard([&] {
        Group& g = *b->g;
        const StateItem it = state_item(g, what);
        if (it.dev) {
            g.d2h((char*)out, (const char*)it.dev, it.count * (size_t)it.bytes);
            return;
        }
        for (size_t r = 0; r < it.count; ++r) {
            std::memcpy((char*)out + r * 8, (const char*)&g.hrs[r] + it.rs_off, 8);
        }
    });
}

int fqsb_batch_set_state(fqsb_batch* b, int what, const void* in)
{
    return guard([&] {
        Group& g = *b->g;
        const StateItem it = state_item(g, what);
        g.adj_valid = false;
        if (what == FQSB_STATE_JLOC) {
            // every kernel reads y[p][jloc] and y[p][jloc + 1] of the window:
            // a checkpoint of another nchunk (or corrupt data) is refused
            // here, before anything reaches the device
            const int32_t* jl = (const int32_t*)in;
            const int64_t hi = (int64_t)g.nchunk - 2;
            for (size_t i = 0; i < it.count; ++i) {
                if (jl[i] < 0 || (int64_t)jl[i] > hi) {
                    throw Error("checkpoint: local well index " + std::to_string(jl[i]) + " of particle " +
                                std::to_string(i) + " is outside the window [0, " + std::to_string(hi) + "]");
                }
            }
        }
        if (it.dev) {
            g.h2d((char*)it.dev, (const char*)in, it.count * (size_t)it.bytes);
        }
        else {
            for (size_t r = 0; r < it.count; ++r) {
                std::memcpy((char*)&g.hrs[r] + it.rs_off, (const char*)in + r * 8, 8);
            }
        }
        if (what == FQSB_STATE_NEXT) { // the kernels' cached minimum of next
            const int64_t* nx = (const int64_t*)in;
            for (int64_t r = 0; r < g.R; ++r) {
                int64_t mn = std::numeric_limits<int64_t>::max();
                for (int64_t p = g.off[(size_t)r]; p < g.off[(size_t)r + 1]; ++p) {
                    mn = std::min(mn, nx[p]);
                }
                g.hrs[(size_t)r].min_next = mn;
            }
        }
        g.push();
        HIPC(hipStreamSynchronize(g.stream));
    });
}

int fqsb_batch_refresh(fqsb_batch* b)
{
    return guard([&] {
        b->g->adj_valid = false;
        b->g->refresh();
    });
}

int fqsb_batch_get(fqsb_batch* b, int field, double* out)
{
    return guard([&] {
        Group& g = *b->g;
        g.d2h(out, g.field(field), (size_t)g.total);
    });
}

int fqsb_batch_set(fqsb_batch* b, int field, const double* in)
{
    return guard([&] {
        Group& g = *b->g;
        const size_t n = (size_t)g.total;
        g.adj_valid = false;
        switch (field) {
            case FQSB_U:
                g.h2d(g.u, in, n);
                g.updated_u();
                break;
            case FQSB_V:
                g.h2d(g.v, in, n);
                g.updated_v();
                break;
            case FQSB_A:
                g.h2d(g.a, in, n);
                break;
            default:
                throw Error("field is read-only");
        }
    });
}

int fqsb_batch_get_u_frame(fqsb_batch* b, double* out)
{
    for (int64_t r = 0; r < b->g->R; ++r) {
        out[r] = b->g->hrs[(size_t)r].u_frame;
    }
    return 0;
}

int fqsb_batch_get_inc(fqsb_batch* b, int64_t* out)
{
    for (int64_t r = 0; r < b->g->R; ++r) {
        out[r] = b->g->hrs[(size_t)r].inc;
    }
    return 0;
}

int fqsb_batch_chunk_index_at_align(fqsb_batch* b, int64_t* out)
{
    return guard([&] {
        Group& g = *b->g;
        const size_t n = (size_t)g.total;
        std::vector<int32_t> jl(n);
        g.d2h(out, g.start, n);
        g.d2h(jl.data(), g.jloc, n);
        for (size_t p = 0; p < n; ++p) {
            out[p] += jl[p];
        }
    });
}

void* fqsb_batch_stream(fqsb_batch* b)
{
    return (void*)b->g->stream;
}

// --- batch: parameters, device-side access, adjoint ----------------------------

int fqsb_batch_set_parameters(fqsb_batch* b, const double* params)
{
    return guard([&] {
        b->g->adj_valid = false;
        b->g->set_parameters(params);
    });
}

int fqsb_batch_get_parameters(const fqsb_batch* b, double* params)
{
    b->g->get_parameters(params);
    return 0;
}

uint64_t fqsb_batch_nchunk(const fqsb_batch* b)
{
    return (uint64_t)b->g->nchunk;
}

int fqsb_batch_get_device(fqsb_batch* b, int field, double* out)
{
    return guard([&] {
        Group& g = *b->g;
        HIPC(hipMemcpyAsync(out, g.field(field), sizeof(double) * (size_t)g.total, hipMemcpyDeviceToDevice,
                            g.stream));
        HIPC(hipStreamSynchronize(g.stream));
    });
}

int fqsb_batch_set_device(fqsb_batch* b, int field, const double* in)
{
    return guard([&] {
        Group& g = *b->g;
        if (field != FQSB_U && field != FQSB_V && field != FQSB_A) {
            throw Error("field is read-only");
        }
        HIPC(hipMemcpyAsync(g.field(field), in, sizeof(double) * (size_t)g.total, hipMemcpyDeviceToDevice,
                            g.stream));
        HIPC(hipStreamSynchronize(g.stream));
        if (field == FQSB_U) {
            g.updated_u();
        }
        else if (field == FQSB_V) {
            g.updated_v();
        }
        g.adj_valid = false;
    });
}

int fqsb_batch_chunk_data_device(fqsb_batch* b, double* out)
{
    return guard([&] {
        Group& g = *b->g;
        if (!g.uniform) {
            throw Error("chunk data of the batch needs replicas of one shape");
        }
        HIPC(hipMemcpyAsync(out, g.y, sizeof(double) * (size_t)g.total * (size_t)g.nchunk,
                            hipMemcpyDeviceToDevice, g.stream));
        HIPC(hipStreamSynchronize(g.stream));
    });
}

int fqsb_batch_set_chunk_data_device(fqsb_batch* b, const double* in)
{
    return guard([&] {
        Group& g = *b->g;
        if (!g.uniform) {
            throw Error("chunk data of the batch needs replicas of one shape");
        }
        HIPC(hipMemcpyAsync(g.y, in, sizeof(double) * (size_t)g.total * (size_t)g.nchunk, hipMemcpyDeviceToDevice,
                            g.stream));
        HIPC(hipStreamSynchronize(g.stream));
        g.refresh();
        g.adj_valid = false;
    });
}

int fqsb_batch_adjoint_kernel_ms(const fqsb_batch* b, float* record_ms, float* reverse_ms)
{
    *record_ms = b->g->adj_record_ms;
    *reverse_ms = b->g->adj_reverse_ms;
    return 0;
}

int fqsb_batch_adjoint_forward(fqsb_batch* b, uint64_t n, uint64_t window, int flow, double v_frame)
{
    return guard([&] { b->g->adjoint_forward(n, window, flow != 0, v_frame); });
}

int fqsb_batch_adjoint_generation(const fqsb_batch* b, uint64_t* gen)
{
    *gen = b->g->adj_valid ? b->g->adj_gen : 0;
    return 0;
}

int fqsb_batch_adjoint_backward_device(fqsb_batch* b, const double* gu, const double* gv, const double* ga,
                                       double* gu0, double* gv0, double* ga0, double* gparams, double* gy)
{
    return guard([&] { b->g->adjoint_backward(gu, gv, ga, gu0, gv0, ga0, gparams, gy); });
}

// --- ragged set ----------------------------------------------------------------

int fqsb_ragged_create(const fqsb_desc* descs, uint64_t n, fqsb_ragged** out)
{
    return guard([&] {
        *out = nullptr;
        if (!descs || n < 1) {
            throw Error("at least one system is needed");
        }
        std::unique_ptr<fqsb_ragged> rg(new fqsb_ragged());
        // key: kernel class and size class (ceil(log2 N))
        typedef std::tuple<int, int, int, int, int, int> Key;
        std::map<Key, std::vector<uint64_t>> classes;
        rg->sizes.resize((size_t)n);
        rg->uoff.resize((size_t)n + 1, 0);
        for (uint64_t i = 0; i < n; ++i) {
            const fqsb_desc& d = descs[i];
            const int64_t N = validate_desc(d);
            rg->sizes[(size_t)i] = N;
            rg->uoff[(size_t)i + 1] = rg->uoff[(size_t)i] + N;
            int lg = 0;
            while ((int64_t(1) << lg) < N) {
                ++lg;
            }
            classes[Key(d.rank, d.potential, d.interactions, d.external, d.minimisation, lg)].push_back(i);
        }
        rg->where.resize((size_t)n);
        int64_t resident_max = RAGGED_RESIDENT_MAX;
        if (const char* e = std::getenv("FQSB_RAGGED_RESIDENT_MAX")) {
            resident_max = std::max<int64_t>(0, std::atoll(e));
        }
        for (auto& kv : classes) {
            std::vector<fqsb_desc> ds;
            for (uint64_t i : kv.second) {
                rg->where[(size_t)i] = std::make_pair(rg->groups.size(), (int64_t)ds.size());
                ds.push_back(descs[i]);
            }
            rg->groups.emplace_back(build_group_multi(ds, resident_max));
        }
        *out = rg.release();
    });
}

int fqsb_ragged_destroy(fqsb_ragged* r)
{
    delete r;
    return 0;
}

uint64_t fqsb_ragged_systems(const fqsb_ragged* r)
{
    return (uint64_t)r->sizes.size();
}

uint64_t fqsb_ragged_groups(const fqsb_ragged* r)
{
    return (uint64_t)r->groups.size();
}

int fqsb_ragged_sizes(const fqsb_ragged* r, uint64_t* out)
{
    for (size_t i = 0; i < r->sizes.size(); ++i) {
        out[i] = (uint64_t)r->sizes[i];
    }
    return 0;
}

int fqsb_ragged_group_info(const fqsb_ragged* r, uint64_t group, int64_t* info)
{
    return guard([&] {
        if (group >= r->groups.size()) {
            throw Error("group index out of range");
        }
        const Group& g = *r->groups[(size_t)group];
        int64_t path = FQSB_PATH_RESIDENT;
        if (g.np_tiled) {
            path = FQSB_PATH_NOPASSING_TILES;
        }
        else if (g.minim == FQSB_MIN_OVERDAMPED) {
            path = FQSB_PATH_NOPASSING_RESIDENT;
        }
        else if (g.tiled) {
            path = FQSB_PATH_TILES;
        }
        else if (g.streaming) {
            path = FQSB_PATH_STREAMING;
        }
        const bool blocked = g.tiled && g.tile_part_k != nullptr;
        const int64_t v[13] = {g.R,         g.N,     path,    g.ppt,   g.threads,
                               blocked ? g.block_steps() : 0,  (int64_t)g.d.rank,
                               g.pot,       g.inter, g.ext,   g.minim, g.total,
                               blocked ? g.tb_kmax : 0};
        std::copy(v, v + 13, info);
    });
}

int fqsb_ragged_group_time(const fqsb_ragged* r, uint64_t group, double* ms, double* particle_steps)
{
    return guard([&] {
        if (group >= r->groups.size()) {
            throw Error("group index out of range");
        }
        *ms = r->groups[(size_t)group]->busy_ms;
        *particle_steps = r->groups[(size_t)group]->busy_ps;
    });
}

int fqsb_ragged_set_serial(fqsb_ragged* r, int serial)
{
    return guard([&] { r->serial = serial != 0; });
}

int fqsb_ragged_reset_group_time(fqsb_ragged* r)
{
    for (auto& g : r->groups) {
        g->busy_ms = 0.0;
        g->busy_ps = 0.0;
    }
    return 0;
}

int fqsb_ragged_group_systems(const fqsb_ragged* r, uint64_t group, uint64_t* idx)
{
    return guard([&] {
        if (group >= r->groups.size()) {
            throw Error("group index out of range");
        }
        for (size_t i = 0; i < r->where.size(); ++i) {
            if (r->where[i].first == group) {
                idx[r->where[i].second] = (uint64_t)i;
            }
        }
    });
}

int fqsb_ragged_group_shapes(const fqsb_ragged* r, uint64_t group, int* ppt, int* threads, int cap, int* count)
{
    return guard([&] {
        if (group >= r->groups.size()) {
            throw Error("group index out of range");
        }
        const auto sh = r->groups[(size_t)group]->launch_shapes();
        const int n = (int)std::min<size_t>(sh.size(), (size_t)std::max(cap, 0));
        for (int i = 0; i < n; ++i) {
            ppt[i] = sh[(size_t)i].first;
            threads[i] = sh[(size_t)i].second;
        }
        *count = (int)sh.size();
    });
}

int fqsb_ragged_set_group_launch(fqsb_ragged* r, uint64_t group, int ppt, int threads)
{
    return guard([&] {
        if (group >= r->groups.size()) {
            throw Error("group index out of range");
        }
        Group& g = *r->groups[(size_t)group];
        if (g.streaming || g.minim == FQSB_MIN_OVERDAMPED) {
            throw Error("the group does not run on the resident kernels");
        }
        g.choose_launch(ppt, threads);
    });
}

int fqsb_ragged_set_group_block(fqsb_ragged* r, uint64_t group, int steps)
{
    return guard([&] {
        if (group >= r->groups.size()) {
            throw Error("group index out of range");
        }
        Group& g = *r->groups[(size_t)group];
        if (!g.tiled || g.tile_part_k == nullptr) {
            throw Error("the group does not run blocks of steps over tiles");
        }
        if (steps < 1 || steps > fq::TB_KMAX_STEPS) {
            throw Error("the block length must be 1 to " + std::to_string(fq::TB_KMAX_STEPS) + " steps");
        }
        g.tile_k = steps;
    });
}

int fqsb_ragged_time_steps(fqsb_ragged* r, uint64_t n)
{
    return guard([&] {
        if (n > 0) {
            r->each([&](Group& g, size_t) { g.steps(n, false, 0.0, 0, true, false); });
        }
    });
}

int fqsb_ragged_flow_steps(fqsb_ragged* r, uint64_t n, double v_frame)
{
    return guard([&] {
        if (n > 0) {
            r->each([&](Group& g, size_t) { g.steps(n, true, v_frame, 0, true, false); });
        }
    });
}

int fqsb_ragged_minimise(fqsb_ragged* r, double tol, uint64_t niter_tol, uint64_t max_iter, int max_iter_is_error,
                         const uint8_t* active, uint64_t* ret)
{
    return guard([&] {
        auto m = r->masks(active);
        std::vector<std::vector<uint64_t>> gret(r->groups.size());
        for (auto& g : r->groups) {
            if (g->minim == FQSB_MIN_NONE) {
                for (size_t i = 0; i < r->where.size(); ++i) {
                    if (r->groups[r->where[i].first].get() == g.get() && (!active || active[i])) {
                        throw Error("Minimisation not implementated");
                    }
                }
            }
        }
        r->each([&](Group& g, size_t gi) {
            gret[gi].assign((size_t)g.R, 0);
            bool any = false;
            for (uint8_t x : m[gi]) {
                any |= x != 0;
            }
            if (!any) {
                return;
            }
            g.mask = m[gi].data();
            try {
                if (g.minim == FQSB_MIN_OVERDAMPED) {
                    auto conv = nopassing_run(g, tol, niter_tol, max_iter);
                    for (int64_t k = 0; k < g.R; ++k) {
                        gret[gi][(size_t)k] = (!m[gi][(size_t)k] || conv[(size_t)k]) ? 0 : max_iter + 1;
                    }
                }
                else {
                    fq::RunArgs ra{};
                    ra.mode = fq::MODE_MINIMISE;
                    ra.tol = tol;
                    ra.tol2 = tol * tol;
                    ra.niter_tol = (int64_t)niter_tol;
                    g.run(ra, max_iter, MIN_BLOCK, false);
                    for (int64_t k = 0; k < g.R; ++k) {
                        const auto& st = g.hrs[(size_t)k];
                        if (st.status == fq::ST_NAN) {
                            throw Error("NaN entries found");
                        }
                        gret[gi][(size_t)k] = (st.status == fq::ST_IDLE || st.status == fq::ST_CONVERGED)
                                                  ? 0
                                                  : max_iter + 1;
                    }
                }
            }
            catch (...) {
                g.mask = nullptr;
                throw;
            }
            g.mask = nullptr;
        });
        for (size_t i = 0; i < r->where.size(); ++i) {
            ret[i] = gret[r->where[i].first][(size_t)r->where[i].second];
            if (ret[i] != 0 && max_iter_is_error) {
                throw Error("No convergence found");
            }
        }
    });
}

int fqsb_ragged_event_driven_step(fqsb_ragged* r, double eps, int kick, int direction, const uint8_t* active,
                                  double* ret)
{
    return guard([&] {
        auto m = r->masks(active);
        std::vector<std::vector<double>> gret(r->groups.size());
        for (size_t gi = 0; gi < r->groups.size(); ++gi) {
            const Group& g = *r->groups[gi];
            bool any = false;
            for (uint8_t x : m[gi]) {
                any |= x != 0;
            }
            if (any && g.ext) {
                throw Error("eventDrivenStep is not available for this system");
            }
            if (any && !kick && g.pot == FQSB_POT_SMOOTH) {
                throw Error("Operation not possible.");
            }
        }
        r->each([&](Group& g, size_t gi) {
            gret[gi].assign((size_t)g.R, 0.0);
            bool any = false;
            for (uint8_t x : m[gi]) {
                any |= x != 0;
            }
            if (!any) {
                return;
            }
            g.mask = m[gi].data();
            try {
                gret[gi] = g.event_driven_step(eps, kick != 0, direction);
            }
            catch (...) {
                g.mask = nullptr;
                throw;
            }
            g.mask = nullptr;
        });
        for (size_t i = 0; i < r->where.size(); ++i) {
            ret[i] = gret[r->where[i].first][(size_t)r->where[i].second];
        }
    });
}

int fqsb_ragged_quasistatic(fqsb_ragged* r, int64_t first_event, int64_t nevents, double eps, int direction,
                            int kicks, double tol, uint64_t niter_tol, uint64_t max_iter, const uint8_t* active,
                            int64_t* S, double* u_frame, double* f_frame, int64_t* inc)
{
    return guard([&] {
        if (nevents < 0) {
            throw Error("the first event and the number of events must be non-negative");
        }
        auto m = r->masks(active);
        const size_t G = r->groups.size();
        std::vector<std::vector<int64_t>> gS(G), ginc(G);
        std::vector<std::vector<double>> guf(G), gff(G);
        r->each([&](Group& g, size_t gi) {
            const size_t n = (size_t)(g.R * nevents);
            gS[gi].assign(n, 0);
            ginc[gi].assign(n, 0);
            guf[gi].assign(n, 0.0);
            gff[gi].assign(n, 0.0);
            bool any = false;
            for (uint8_t x : m[gi]) {
                any |= x != 0;
            }
            if (!any) {
                return;
            }
            g.mask = m[gi].data();
            try {
                g.quasistatic(first_event, nevents, eps, direction, kicks, tol, niter_tol, max_iter, gS[gi].data(),
                              guf[gi].data(), gff[gi].data(), ginc[gi].data());
            }
            catch (...) {
                g.mask = nullptr;
                throw;
            }
            g.mask = nullptr;
        });
        for (size_t i = 0; i < r->where.size(); ++i) {
            const size_t gi = r->where[i].first;
            const size_t o = (size_t)r->where[i].second * (size_t)nevents;
            for (int64_t e = 0; e < nevents; ++e) {
                const size_t d = i * (size_t)nevents + (size_t)e;
                if (S) {
                    S[d] = gS[gi][o + (size_t)e];
                }
                if (u_frame) {
                    u_frame[d] = guf[gi][o + (size_t)e];
                }
                if (f_frame) {
                    f_frame[d] = gff[gi][o + (size_t)e];
                }
                if (inc) {
                    inc[d] = ginc[gi][o + (size_t)e];
                }
            }
        }
    });
}

int fqsb_ragged_get(fqsb_ragged* r, int field, double* out)
{
    return guard([&] {
        for (size_t gi = 0; gi < r->groups.size(); ++gi) {
            Group& g = *r->groups[gi];
            std::vector<double> h((size_t)g.total);
            g.d2h(h.data(), g.field(field), h.size());
            for (size_t i = 0; i < r->where.size(); ++i) {
                if (r->where[i].first == gi) {
                    const int64_t k = r->where[i].second;
                    std::copy(h.begin() + g.off[(size_t)k], h.begin() + g.off[(size_t)k + 1], out + r->uoff[i]);
                }
            }
        }
    });
}

int fqsb_ragged_set(fqsb_ragged* r, int field, const double* in)
{
    return guard([&] {
        if (field != FQSB_U && field != FQSB_V && field != FQSB_A) {
            throw Error("field is read-only");
        }
        for (size_t gi = 0; gi < r->groups.size(); ++gi) {
            Group& g = *r->groups[gi];
            std::vector<double> h((size_t)g.total);
            for (size_t i = 0; i < r->where.size(); ++i) {
                if (r->where[i].first == gi) {
                    const int64_t k = r->where[i].second;
                    std::copy(in + r->uoff[i], in + r->uoff[i + 1], h.begin() + g.off[(size_t)k]);
                }
            }
            g.h2d(g.field(field), h.data(), h.size());
            if (field == FQSB_U) {
                g.updated_u();
            }
            else if (field == FQSB_V) {
                g.updated_v();
            }
        }
    });
}

int fqsb_ragged_get_u_frame(fqsb_ragged* r, double* out)
{
    for (size_t i = 0; i < r->where.size(); ++i) {
        out[i] = r->groups[r->where[i].first]->hrs[(size_t)r->where[i].second].u_frame;
    }
    return 0;
}

int fqsb_ragged_last_steps(fqsb_ragged* r, uint64_t* out)
{
    for (size_t i = 0; i < r->where.size(); ++i) {
        out[i] = (uint64_t)r->groups[r->where[i].first]->hrs[(size_t)r->where[i].second].steps;
    }
    return 0;
}

int fqsb_ragged_get_inc(fqsb_ragged* r, int64_t* out)
{
    for (size_t i = 0; i < r->where.size(); ++i) {
        out[i] = r->groups[r->where[i].first]->hrs[(size_t)r->where[i].second].inc;
    }
    return 0;
}

int fqsb_ragged_chunk_index_at_align(fqsb_ragged* r, int64_t* out)
{
    return guard([&] {
        for (size_t gi = 0; gi < r->groups.size(); ++gi) {
            Group& g = *r->groups[gi];
            std::vector<int64_t> st((size_t)g.total);
            std::vector<int32_t> jl((size_t)g.total);
            g.d2h(st.data(), g.start, st.size());
            g.d2h(jl.data(), g.jloc, jl.size());
            for (size_t i = 0; i < r->where.size(); ++i) {
                if (r->where[i].first == gi) {
                    const int64_t k = r->where[i].second;
                    for (int64_t p = g.off[(size_t)k]; p < g.off[(size_t)k + 1]; ++p) {
                        out[r->uoff[i] + (p - g.off[(size_t)k])] = st[(size_t)p] + jl[(size_t)p];
                    }
                }
            }
        }
    });
}

} // extern "C"
