// ptd_batch.hip -- batches of independent float64 PT solves (include/ptd.h: ptd_batch_*).
//
// Every member is a full problem of its own (2D or 3D, lin / step / nonlin, its own grid and scalars),
// solved over nt physical steps exactly as ptd_solve would solve it alone: the same tile functions,
// the same checking tiling (so the same residual partials and stopping decisions), the same
// iteration counts and the same bits. What changes is the schedule:
//  - members are grouped by (sweep kind, dcase, arith), the kind being the sweep ptd_solve runs for
//    them (three- and two-level 3D, LDS-resident 2D on grids of at most 1024 cores, two-level 2D on
//    larger ones, single-level 2D); one launch covers the tiles of every member of a group (a table maps
//    each block to its member, tile and z chunk depth);
//  - the PT loop of every member runs on the device: after each check block of nout iterations the
//    group's controller kernel applies the reference loop's rules per member (scripts/diff_3D/
//    diff_3D_nonlin_multixpu.jl:104-118) and members that are done drop out of later launches;
//  - each group's check block (its launches, reduction, controller, Hold .= H and its running-member
//    count) is captured once as a HIP graph on the group's own stream and replayed there; the groups
//    advance independently (no group waits for a slower one), the host polling their counts every few
//    replays; groups of 2D grids wider than the 3D ones run several check blocks per replay
//    (group_sub_blocks), as they need proportionally more PT iterations;
//  - the tiling of the non-checking launches (z chunk depth in 3D, rows per block in 2D) is chosen online
//    per group: the first replays try each candidate graph in turn, three rounds (each replay is real
//    progress, results do not depend on the tiling), and the candidate of least median replay time is kept;
//  - each group's graph brackets its launches with two clock kernels that sum the device time the group
//    took over the replays (ptd_batch_group_times).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "ptd_internal.hpp"

using ptd::BatchMember;
using ptd::BatchTile;

namespace {

struct Candidate {
    int tz3d;     // z chunk depth of the non-checking 3D launches
    int rows2d;   // 0: 16 rows, 1: 8 rows per block in the non-checking 2D launches
};
const Candidate kCandidates[] = {{16, 0}, {8, 0}, {32, 0}, {16, 1}};
constexpr int kNumCandidates = sizeof(kCandidates) / sizeof(kCandidates[0]);

struct Group {
    int kind = 0, dcase = 0, arith = 0;  // kind: the sweep of its members (ptd::batch_kind)
    std::vector<int32_t> members;
    int32_t* d_members = nullptr;
    hipStream_t stream = nullptr;
    BatchTile* d_check = nullptr;
    int n_check = 0;
    BatchTile* d_nc[kNumCandidates] = {};
    int n_nc[kNumCandidates] = {};
    BatchTile* d_t2 = nullptr;  // single-level 2D groups: two-level tiles of their pair launches
    int n_t2 = 0;
    BatchTile* d_t3[kNumCandidates] = {};  // three-level groups: tiles of their three-level passes
    int n_t3[kNumCandidates] = {};
    std::vector<int> plan;      // PT iterations of each launch of a check block (block_plan)
    int sub = 1;                // check blocks per replay
    uint64_t* d_span = nullptr; // (start, sum) of the group's device time per replay, summed over replays
    int32_t* d_active = nullptr; // members of the group still running, after each replay
    int32_t* h_active = nullptr; // its pinned host mirror (a slot of ptd_batch::h_active)
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    hipEvent_t ev_poll[2] = {nullptr, nullptr};  // ends of the two sets of replays in flight
    hipGraph_t graph[kNumCandidates] = {};
    hipGraphExec_t exec[kNumCandidates] = {};
    int chosen = -1;                        // tiling candidate of the non-checking launches
    double cand_ms[kNumCandidates] = {};    // median replay time of each candidate while tuning
    int64_t replays = 0;
};

// Check blocks per replay of each group. Every replay advances each running member by the same number
// of check blocks, and PT iteration counts grow with the grid's extent (the accelerated iteration needs
// O(n) iterations per step in 1D, 2D and 3D alike), so 2D members on grids wider than the 3D ones would
// need several times more replays than the 3D members, and would finish alone, on a GPU they cannot
// fill. Such groups run floor(n_group / n_3D) check blocks per replay (at most 8), each closed by its own
// controller, and finish in about as many replays as the 3D members.
constexpr int kMaxSubBlocks = 8;
int group_sub_blocks(const std::vector<ptd_params>& params, const std::vector<int32_t>& members) {
    int64_t n3 = 0, ng = 0;
    auto extent = [](const ptd_params& p) {
        return std::max<int64_t>(p.nx, std::max<int64_t>(p.ny, p.ndim == 3 ? p.nz : 0));
    };
    for (const auto& p : params)
        if (p.ndim == 3) n3 = std::max(n3, extent(p));
    for (int32_t m : members) ng = std::max(ng, extent(params[(size_t)m]));
    if (n3 == 0) return 1;
    return (int)std::min<int64_t>(kMaxSubBlocks, std::max<int64_t>(1, ng / n3));
}

// PT iterations of each launch of one check block of nout iterations. Only the checking launch must
// use the single-problem tiling (its residual partials decide when a member stops); every other launch
// gives the same bits whatever its depth or tiling. 3D members follow the single-problem rule
// (ptd::next_launch_iters: three-level passes for the float64 FAST linear / step members, two-level
// pairs, a checking two-level pass last); 2D grids of at most 1024 cores run LDS-resident launches of up
// to lds2d_max_iters() iterations as their single solves do (the last one checking); the two-level 2D
// groups run pairs; and the single-level 2D groups (LDS-resident iterations switched off, grids too small
// for the two-level 2D sweep) run their first nout - 2 iterations as two-level pairs, then one
// single-level sweep and the single-problem checking launch.
std::vector<int> block_plan(const ptd_params& p) {
    std::vector<int> plan;
    if (ptd::batch_kind(p) == ptd::kSweepLds2D) {  // run_iterations' LDS-resident rule
        const int64_t lmax = ptd::lds2d_max_iters();
        for (int64_t n = p.nout; n > 0;) {
            const int64_t k = n <= lmax ? n : std::min<int64_t>(n - 1, lmax);
            plan.push_back((int)k);
            n -= k;
        }
        return plan;
    }
    if (ptd::batch_kind(p) == ptd::kSweep2D) {
        for (int i = 0; i < p.nout / 2 - 1; ++i) plan.push_back(2);
        plan.push_back(1);
        plan.push_back(1);
        return plan;
    }
    const bool t3 = ptd::batch_t3(p);
    for (int64_t n = p.nout; n > 0;) {
        const int k = ptd::next_launch_iters(n, true, true, t3);
        plan.push_back(k);
        n -= k;
    }
    return plan;
}

// Estimated work of a member: cells times PT iterations, which grow with the grid's extent (the reference's
// published counts: about 6 n iterations over 5 steps for linear and step D, 2 n for D = H^3).
double member_work(const ptd_params& p) {
    const double n = (double)std::max<int64_t>(p.nx, std::max<int64_t>(p.ny, p.ndim == 3 ? p.nz : 0));
    const double cells = (double)p.nx * (double)p.ny * (double)(p.ndim == 3 ? p.nz : 1);
    return cells * n * (p.dcase == PTD_NONLIN ? 2.0 : 6.0);
}

// Streams of the groups. A device runs a few hardware queues (4 by default) and streams beyond them share
// one, where the kernels of different streams run in submission order; so the groups are packed onto
// kLanes streams, the largest estimated work first onto the least loaded lane (PTD_BATCH_LANES overrides).
constexpr int kLanes = 4;
int batch_lanes(int ngroups) {
    const char* e = getenv("PTD_BATCH_LANES");
    const int want = e && atoi(e) > 0 ? atoi(e) : kLanes;
    return std::max(1, std::min(want, ngroups));
}

}  // namespace

struct ptd_batch {
    int device = 0;
    int n = 0;
    int nout = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_fork = nullptr;
    std::vector<hipStream_t> lanes;  // the groups' streams (several groups may share one)
    std::vector<ptd_params> params;
    std::vector<BatchMember> hm;
    BatchMember* dm = nullptr;
    std::vector<void*> allocs;
    std::vector<Group> groups;
    int64_t* d_iters = nullptr;
    double* d_errs = nullptr;
    int nt_cap = 0;
    int32_t* h_active = nullptr;  // pinned, one slot per group
    uint64_t* d_t0 = nullptr;
    bool graphs = false;
    double wall_khz = 0;
    int64_t bytes = 0;
};

namespace {

#define PTD_HIPB(expr)                                                                             \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return ptd::set_error(PTD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int balloc(ptd_batch* b, void** ptr, size_t bytes) {
    *ptr = nullptr;
    if (bytes == 0) return PTD_OK;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        *ptr = nullptr;
        (void)hipGetLastError();  // clear it: later launch checks read the thread's last error
        return ptd::set_error(PTD_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    b->allocs.push_back(*ptr);
    b->bytes += (int64_t)bytes;
    e = hipMemsetAsync(*ptr, 0, bytes, b->stream);  // at creation only, never inside a captured graph
    if (e != hipSuccess) return ptd::set_error(PTD_ERR_HIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
    return PTD_OK;
}

// frees one of the batch's allocations (null: nothing)
void release(ptd_batch* b, void* ptr, size_t bytes) {
    if (!ptr) return;
    auto it = std::find(b->allocs.begin(), b->allocs.end(), ptr);
    if (it == b->allocs.end()) return;
    b->allocs.erase(it);
    b->bytes -= (int64_t)bytes;
    (void)hipFree(ptr);
}

void destroy_graphs(ptd_batch* b) {
    for (auto& g : b->groups)
        for (int c = 0; c < kNumCandidates; ++c) {
            if (g.exec[c]) (void)hipGraphExecDestroy(g.exec[c]);
            if (g.graph[c]) (void)hipGraphDestroy(g.graph[c]);
            g.exec[c] = nullptr;
            g.graph[c] = nullptr;
        }
    b->graphs = false;
}

void free_batch(ptd_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    destroy_graphs(b);
    for (void* p : b->allocs) (void)hipFree(p);
    if (b->h_active) (void)hipHostFree(b->h_active);
    for (auto l : b->lanes) {
        (void)hipStreamSynchronize(l);
        (void)hipStreamDestroy(l);
    }
    for (auto& g : b->groups) {
        if (g.ev_a) (void)hipEventDestroy(g.ev_a);
        if (g.ev_b) (void)hipEventDestroy(g.ev_b);
        for (auto e : g.ev_poll)
            if (e) (void)hipEventDestroy(e);
    }
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

// tiles of one member's three-level pass with z chunk depth tz
int member_tiles3(const ptd_params& p, int tz) {
    const int r = ptd::fused3_rows();
    const int64_t ntx = (p.nx - 2 + 57) / 58, nty = (p.ny - 2 + r - 7) / (r - 6), nbz = (p.nz - 2 + tz - 1) / tz;
    return (int)(ntx * nty * nbz);
}

// tiles of one member's launch: rows per block (two-level 2D: rows per tile chunk), z chunk depth tz (3D)
int member_tiles(const ptd_params& p, int kind, int rows, int tz) {
    if (kind == ptd::kSweepLds2D) return (int)ptd::lds2d_tiles(p);
    if (kind == ptd::kSweep2D2) return (int)(((p.nx - 2 + 59) / 60) * ((p.ny - 2 + rows - 1) / rows));
    if (p.ndim == 3) {
        const int64_t ntx = (p.nx - 2 + 59) / 60, nty = (p.ny - 2 + rows - 5) / (rows - 4);
        const int64_t nbz = (p.nz - 2 + tz - 1) / tz;
        return (int)(ntx * nty * nbz);
    }
    const int64_t ntx = (p.nx - 2 + 61) / 62, nty = (p.ny - 2 + rows - 3) / (rows - 2);
    return (int)(ntx * nty);
}

// the single-problem sweeps' z chunk depth: the checking launch must use their tiling
int check_tz(const ptd_params& p) { return ptd::check_chunk_tz(p); }

int upload_table(ptd_batch* b, const std::vector<BatchTile>& t, BatchTile** d) {
    int st = balloc(b, (void**)d, t.size() * sizeof(BatchTile));
    if (st != PTD_OK) return st;
    if (!t.empty())
        PTD_HIPB(hipMemcpyAsync(*d, t.data(), t.size() * sizeof(BatchTile), hipMemcpyHostToDevice, b->stream));
    return PTD_OK;
}

// one replay of group gi on its own stream: its check blocks of nout iterations, each closed by the
// group's controller and Hold .= H, then its running-member count to the host
int enqueue_group(ptd_batch* b, size_t gi, int cand) {
    Group& g = b->groups[gi];
    const int nm = (int)g.members.size();
    PTD_HIPB(ptd::launch_batch_span(g.stream, g.d_span, false));
    const int lpb = (int)g.plan.size();
    for (int sb = 0; sb < g.sub; ++sb) {
        for (int li = 0; li < lpb; ++li) {
            const bool check = li == lpb - 1;
            if (g.kind == ptd::kSweepLds2D) {  // LDS-resident iterations (one tiling: the checking one)
                PTD_HIPB(ptd::launch_batch_lds2d(g.stream, g.dcase, g.arith, b->dm, g.d_check, g.n_check, li,
                                                 g.plan[(size_t)li], check));
                continue;
            }
            if (g.plan[(size_t)li] == 3) {  // a three-level pass
                PTD_HIPB(ptd::launch_batch_sweep3(g.stream, g.dcase, g.arith, b->dm, g.d_t3[cand], g.n_t3[cand], li));
                continue;
            }
            if (g.kind == ptd::kSweep2D && g.plan[(size_t)li] == 2) {  // a two-level pair
                PTD_HIPB(ptd::launch_batch_sweep(g.stream, ptd::kSweep2D2, g.dcase, g.arith,
                                                 ptd::batch_rows_nocheck(ptd::kSweep2D2, g.dcase, g.arith, 0),
                                                 false, b->dm, g.d_t2, g.n_t2, li));
                continue;
            }
            const int rows = check ? ptd::batch_rows_check(g.kind, g.dcase, g.arith)
                                   : ptd::batch_rows_nocheck(g.kind, g.dcase, g.arith, kCandidates[cand].rows2d);
            PTD_HIPB(ptd::launch_batch_sweep(g.stream, g.kind, g.dcase, g.arith, rows, check, b->dm,
                                             check ? g.d_check : g.d_nc[cand], check ? g.n_check : g.n_nc[cand], li));
        }
        PTD_HIPB(ptd::launch_batch_reduce(g.stream, b->dm, g.d_members, nm));
        PTD_HIPB(ptd::launch_batch_control_list(g.stream, b->dm, g.d_members, nm, b->d_iters, b->d_errs, g.d_active,
                                                b->nt_cap));
        PTD_HIPB(ptd::launch_batch_assign_list(g.stream, b->dm, g.d_members, nm));
    }
    PTD_HIPB(ptd::launch_batch_span(g.stream, g.d_span, true));
    PTD_HIPB(hipMemcpyAsync(g.h_active, g.d_active, sizeof(int32_t), hipMemcpyDeviceToHost, g.stream));
    return PTD_OK;
}

int build_graphs(ptd_batch* b) {
    destroy_graphs(b);
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
        Group& g = b->groups[gi];
        for (int c = 0; c < kNumCandidates; ++c) {
            PTD_HIPB(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal));
            const int st = enqueue_group(b, gi, c);
            hipGraph_t gr = nullptr;
            const hipError_t e = hipStreamEndCapture(g.stream, &gr);
            if (st != PTD_OK) {
                if (gr) (void)hipGraphDestroy(gr);
                return st;
            }
            PTD_HIPB(e);
            g.graph[c] = gr;
            PTD_HIPB(hipGraphInstantiate(&g.exec[c], gr, nullptr, nullptr, 0));
        }
    }
    b->graphs = true;
    return PTD_OK;
}

int field_target(ptd_batch* b, int32_t m, int32_t field, void** ptr, int64_t* count, bool for_write) {
    const BatchMember& M = b->hm[(size_t)m];
    const ptd_params& p = b->params[(size_t)m];
    const int64_t ny = p.ndim >= 2 ? p.ny : 1, nz = p.ndim >= 3 ? p.nz : 1;
    const int64_t mx = p.nx - 2, my = ny - 2, mz = p.ndim >= 3 ? nz - 2 : 1;
    switch (field) {
        case PTD_F_H: *ptr = M.H[M.hpar]; *count = M.nH; return PTD_OK;
        case PTD_F_HOLD: *ptr = M.Hold; *count = M.nH; return PTD_OK;
        case PTD_F_D:
            if (!M.D) return ptd::set_error(PTD_ERR_STATE, "member has no D field (step case only)");
            *ptr = (void*)M.D; *count = M.nH; return PTD_OK;
        case PTD_F_QX: case PTD_F_QY: case PTD_F_QZ: {
            if (for_write) return ptd::set_error(PTD_ERR_STATE, "batch fluxes are read-only");
            const int a = field - PTD_F_QX;
            if (a >= p.ndim) return ptd::set_error(PTD_ERR_STATE, "flux absent for this member");
            *ptr = M.q[a][M.qpar];
            *count = a == 0 ? (p.nx - 1) * my * mz : a == 1 ? mx * (ny - 1) * mz : mx * my * (nz - 1);
            return PTD_OK;
        }
        default: return ptd::set_error(PTD_ERR_STATE, "field not held by a batch member");
    }
}

}  // namespace

extern "C" {

int32_t ptd_batch_create(const ptd_params* params, int32_t n, int32_t device, ptd_batch** out) {
    if (!out) return ptd::set_error(PTD_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!params || n < 1 || n > ptd::kBatchMaxMembers)
        return ptd::set_error(PTD_ERR_ARG, "need 1..1024 members");
    for (int32_t m = 0; m < n; ++m) {
        const ptd_params& p = params[m];
        int st = ptd::validate_params(&p);
        if (st != PTD_OK) return st;
        const std::string who = "member " + std::to_string(m) + ": ";
        if (p.ndim < 2) return ptd::set_error(PTD_ERR_ARG, who + "batches hold 2D and 3D problems");
        if (p.precision != PTD_FP64) return ptd::set_error(PTD_ERR_ARG, who + "batches are float64");
        if (p.form != PTD_PLAIN) return ptd::set_error(PTD_ERR_ARG, who + "batches run the converging (plain) scripts");
        if (p.nout != params[0].nout || p.nout % 2 != 0)
            return ptd::set_error(PTD_ERR_ARG, who + "nout must be even and the same for every member");
        if (p.itmax % p.nout != 0) return ptd::set_error(PTD_ERR_ARG, who + "itmax must be a multiple of nout");
        if (p.ndim == 3 && ptd::max_chunk_tz(p, 8) < 32)
            return ptd::set_error(PTD_ERR_ARG, who + "x-y plane too large for the batched 3D sweep (37 planes under 4 GiB)");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return ptd::set_error(PTD_ERR_NODEV, "no such HIP device");
    ptd_batch* b = new ptd_batch();
    b->device = device;
    b->n = n;
    b->nout = params[0].nout;
    b->params.assign(params, params + n);
    b->hm.resize((size_t)n);
    auto fail_free = [&](int st) {
        free_batch(b);
        return st;
    };
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&b->stream) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess)
        return fail_free(ptd::set_error(PTD_ERR_HIP, "stream / event creation failed"));
    int rate = 0;
    if (hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, device) != hipSuccess || rate <= 0)
        return fail_free(ptd::set_error(PTD_ERR_HIP, "no device wall clock rate"));
    b->wall_khz = (double)rate;

    // fields
    for (int32_t m = 0; m < n; ++m) {
        const ptd_params& p = params[m];
        BatchMember& M = b->hm[(size_t)m];
        std::memset(&M, 0, sizeof(M));
        M.p = ptd::make_kp<double>(p);
        const int64_t ny = p.ny, nz = p.ndim == 3 ? p.nz : 1;
        const int64_t mx = p.nx - 2, my = ny - 2, mz = p.ndim == 3 ? nz - 2 : 1;
        M.nH = p.nx * ny * nz;
        const int64_t nq[3] = {(p.nx - 1) * my * mz, mx * (ny - 1) * mz, p.ndim == 3 ? mx * my * (nz - 1) : 0};
        int st = PTD_OK;
        for (int k = 0; k < 2 && st == PTD_OK; ++k) st = balloc(b, (void**)&M.H[k], (size_t)M.nH * 8);
        if (st == PTD_OK) st = balloc(b, (void**)&M.Hold, (size_t)M.nH * 8);
        if (st == PTD_OK && p.dcase == PTD_STEP) st = balloc(b, (void**)&M.D, (size_t)M.nH * 8);
        for (int a = 0; a < 3 && st == PTD_OK; ++a)
            for (int k = 0; k < 2 && st == PTD_OK; ++k) st = balloc(b, (void**)&M.q[a][k], (size_t)nq[a] * 8);
        M.ntiles_check = member_tiles(p, ptd::batch_kind(p), ptd::batch_rows_check(ptd::batch_kind(p), p.dcase, p.arith),
                                      check_tz(p));
        if (st == PTD_OK) st = balloc(b, (void**)&M.partial, (size_t)M.ntiles_check * sizeof(double2));
        if (st != PTD_OK) return fail_free(st);
        M.itmax = p.itmax;
        M.ndim = p.ndim;
        M.dcase = p.dcase;
        M.arith = p.arith;
        M.nout = p.nout;
        M.lpb = (int32_t)block_plan(p).size();
        M.stop_norm = p.stop_norm;
        M.tol = p.tol;
        M.norm_len = p.norm_len > 0 ? p.norm_len : (double)(mx * my * mz);
    }
    // groups and their tile tables
    for (int32_t m = 0; m < n; ++m) {
        const ptd_params& p = params[m];
        Group* g = nullptr;
        for (auto& x : b->groups)
            if (x.kind == ptd::batch_kind(p) && x.dcase == p.dcase && x.arith == p.arith) g = &x;
        if (!g) {
            b->groups.emplace_back();
            g = &b->groups.back();
            g->kind = ptd::batch_kind(p);
            g->dcase = p.dcase;
            g->arith = p.arith;
            g->plan = block_plan(p);  // a function of (kind, dcase, arith) and the common nout
        }
        if (block_plan(p) != g->plan)  // one graph serves the group: its members share the launch plan
            return fail_free(ptd::set_error(PTD_ERR_ARG, "member " + std::to_string(m) +
                                                             ": check-block plan differs from its group's"));
        g->members.push_back(m);
    }
    for (auto& g : b->groups) g.sub = group_sub_blocks(b->params, g.members);
    {
        const int nl = batch_lanes((int)b->groups.size());
        for (int l = 0; l < nl; ++l) {
            hipStream_t st = nullptr;
            if (hipStreamCreate(&st) != hipSuccess) return fail_free(ptd::set_error(PTD_ERR_HIP, "stream creation failed"));
            b->lanes.push_back(st);
        }
        std::vector<double> work(b->groups.size(), 0.0), load((size_t)nl, 0.0);
        std::vector<size_t> order(b->groups.size());
        for (size_t gi = 0; gi < b->groups.size(); ++gi) {
            order[gi] = gi;
            for (int32_t m : b->groups[gi].members) work[gi] += member_work(params[m]);
        }
        std::sort(order.begin(), order.end(), [&](size_t a, size_t c) { return work[a] > work[c]; });
        for (size_t gi : order) {
            const size_t l = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
            load[l] += work[gi];
            b->groups[gi].stream = b->lanes[l];
        }
    }
    for (auto& g : b->groups) {
        if (hipEventCreate(&g.ev_a) != hipSuccess || hipEventCreate(&g.ev_b) != hipSuccess ||
            hipEventCreateWithFlags(&g.ev_poll[0], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&g.ev_poll[1], hipEventDisableTiming) != hipSuccess)
            return fail_free(ptd::set_error(PTD_ERR_HIP, "event creation failed"));
        int st = balloc(b, (void**)&g.d_members, g.members.size() * sizeof(int32_t));
        if (st == PTD_OK) st = balloc(b, (void**)&g.d_span, 2 * sizeof(uint64_t));
        if (st == PTD_OK) st = balloc(b, (void**)&g.d_active, sizeof(int32_t));
        if (st != PTD_OK) return fail_free(st);
        if (hipMemcpyAsync(g.d_members, g.members.data(), g.members.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                           b->stream) != hipSuccess)
            return fail_free(ptd::set_error(PTD_ERR_HIP, "member table upload failed"));
        std::vector<BatchTile> t;
        for (int32_t m : g.members) {
            const ptd_params& p = params[m];
            const int rows = ptd::batch_rows_check(g.kind, g.dcase, g.arith);
            const int tz = p.ndim == 3 ? check_tz(p) : (g.kind == ptd::kSweep2D2 ? rows : 0);
            const int nt = member_tiles(p, g.kind, rows, tz);
            for (int k = 0; k < nt; ++k) t.push_back({m, k, tz, 0});
        }
        g.n_check = (int)t.size();
        st = upload_table(b, t, &g.d_check);
        if (st != PTD_OK) return fail_free(st);
        for (int c = 0; c < kNumCandidates; ++c) {
            t.clear();
            const int rows = ptd::batch_rows_nocheck(g.kind, g.dcase, g.arith, kCandidates[c].rows2d);
            for (int32_t m : g.members) {
                const ptd_params& p = params[m];
                const int tz = p.ndim == 3 ? (int)std::min<int64_t>(kCandidates[c].tz3d, p.nz - 2)
                                           : (g.kind == ptd::kSweep2D2 ? rows : 0);
                const int nt = member_tiles(p, g.kind, rows, tz);
                for (int k = 0; k < nt; ++k) t.push_back({m, k, tz, 0});
            }
            g.n_nc[c] = (int)t.size();
            st = upload_table(b, t, &g.d_nc[c]);
            if (st != PTD_OK) return fail_free(st);
            if (std::find(g.plan.begin(), g.plan.end(), 3) == g.plan.end()) continue;
            t.clear();
            for (int32_t m : g.members) {
                const ptd_params& p = params[m];
                // a three-level chunk touches tz + 6 planes from its base (32-bit offsets)
                const int tz = (int)std::min<int64_t>(std::min<int64_t>(kCandidates[c].tz3d, p.nz - 2),
                                                      ptd::max_chunk_tz(p, 8) - 1);
                const int nt = member_tiles3(p, tz);
                for (int k = 0; k < nt; ++k) t.push_back({m, k, tz, 0});
            }
            g.n_t3[c] = (int)t.size();
            st = upload_table(b, t, &g.d_t3[c]);
            if (st != PTD_OK) return fail_free(st);
        }
        if (g.kind == ptd::kSweep2D) {
            t.clear();
            const int ty = ptd::batch_rows_nocheck(ptd::kSweep2D2, g.dcase, g.arith, 0);
            for (int32_t m : g.members) {
                const int nt = member_tiles(params[m], ptd::kSweep2D2, ty, ty);
                for (int k = 0; k < nt; ++k) t.push_back({m, k, ty, 0});
            }
            g.n_t2 = (int)t.size();
            st = upload_table(b, t, &g.d_t2);
            if (st != PTD_OK) return fail_free(st);
        }
    }
    int st = balloc(b, (void**)&b->dm, (size_t)n * sizeof(BatchMember));
    if (st == PTD_OK) st = balloc(b, (void**)&b->d_t0, sizeof(uint64_t));
    if (st != PTD_OK) return fail_free(st);
    if (hipHostMalloc((void**)&b->h_active, b->groups.size() * sizeof(int32_t), hipHostMallocDefault) != hipSuccess)
        return fail_free(ptd::set_error(PTD_ERR_NOMEM, "pinned host allocation failed"));
    for (size_t gi = 0; gi < b->groups.size(); ++gi) b->groups[gi].h_active = b->h_active + gi;
    if (hipStreamSynchronize(b->stream) != hipSuccess)
        return fail_free(ptd::set_error(PTD_ERR_HIP, "batch setup failed"));
    *out = b;
    return PTD_OK;
}

int32_t ptd_batch_destroy(ptd_batch* b) {
    if (!b) return ptd::set_error(PTD_ERR_ARG, "batch is NULL");
    free_batch(b);
    return PTD_OK;
}

int32_t ptd_batch_size(const ptd_batch* b) { return b ? b->n : -1; }

int64_t ptd_batch_device_bytes(const ptd_batch* b) { return b ? b->bytes : -1; }

int32_t ptd_batch_set_field(ptd_batch* b, int32_t member, int32_t field, const double* host, int64_t count) {
    if (!b || !host || member < 0 || member >= b->n) return ptd::set_error(PTD_ERR_ARG, "bad batch / member / host pointer");
    PTD_HIPB(hipSetDevice(b->device));
    void* dst = nullptr;
    int64_t want = 0;
    int st = field_target(b, member, field, &dst, &want, true);
    if (st != PTD_OK) return st;
    if (count != want) return ptd::set_error(PTD_ERR_ARG, "count " + std::to_string(count) + " != " + std::to_string(want));
    PTD_HIPB(hipMemcpyAsync(dst, host, (size_t)count * 8, hipMemcpyHostToDevice, b->stream));
    if (field == PTD_F_H) {  // both buffers: the Dirichlet cells are never written by a sweep
        const BatchMember& M = b->hm[(size_t)member];
        PTD_HIPB(hipMemcpyAsync(M.H[M.hpar ^ 1], host, (size_t)count * 8, hipMemcpyHostToDevice, b->stream));
    }
    PTD_HIPB(hipStreamSynchronize(b->stream));
    return PTD_OK;
}

int32_t ptd_batch_get_field(ptd_batch* b, int32_t member, int32_t field, double* host, int64_t count) {
    if (!b || !host || member < 0 || member >= b->n) return ptd::set_error(PTD_ERR_ARG, "bad batch / member / host pointer");
    PTD_HIPB(hipSetDevice(b->device));
    void* src = nullptr;
    int64_t want = 0;
    int st = field_target(b, member, field, &src, &want, false);
    if (st != PTD_OK) return st;
    if (count != want) return ptd::set_error(PTD_ERR_ARG, "count " + std::to_string(count) + " != " + std::to_string(want));
    PTD_HIPB(hipMemcpyAsync(host, src, (size_t)count * 8, hipMemcpyDeviceToHost, b->stream));
    PTD_HIPB(hipStreamSynchronize(b->stream));
    return PTD_OK;
}

int32_t ptd_batch_solve(ptd_batch* b, int32_t nt, int64_t* iters, double* errs, double* latency_ms,
                        int32_t* status, double* total_ms) {
    if (!b || nt < 1) return ptd::set_error(PTD_ERR_ARG, "bad batch or nt < 1");
    PTD_HIPB(hipSetDevice(b->device));
    if (nt > b->nt_cap) {  // per-step outputs on the device; the graphs capture their addresses
        int64_t* di = nullptr;
        double* de = nullptr;
        int st = balloc(b, (void**)&di, (size_t)b->n * nt * sizeof(int64_t));
        if (st == PTD_OK) st = balloc(b, (void**)&de, (size_t)b->n * nt * sizeof(double));
        if (st != PTD_OK) return st;
        PTD_HIPB(hipStreamSynchronize(b->stream));  // no replay may still use the old buffers
        destroy_graphs(b);                            // they hold the old addresses
        release(b, b->d_iters, (size_t)b->n * b->nt_cap * sizeof(int64_t));
        release(b, b->d_errs, (size_t)b->n * b->nt_cap * sizeof(double));
        b->d_iters = di;
        b->d_errs = de;
        b->nt_cap = nt;
    }
    for (auto& M : b->hm) {
        M.active = 1;
        M.step = 0;
        M.nt = nt;
        M.iter = 0;
        M.assign = 0;
        M.status = PTD_OK;
        M.t_done = 0;
    }
    PTD_HIPB(hipMemcpyAsync(b->dm, b->hm.data(), b->hm.size() * sizeof(BatchMember), hipMemcpyHostToDevice, b->stream));
    for (auto& g : b->groups) PTD_HIPB(hipMemsetAsync(g.d_span, 0, 2 * sizeof(uint64_t), b->stream));  // eager
    PTD_HIPB(hipStreamSynchronize(b->stream));
    if (!b->graphs) {
        int st = build_graphs(b);
        if (st != PTD_OK) return st;
    }
    int64_t max_blocks = 0;
    for (const auto& p : b->params) max_blocks = std::max<int64_t>(max_blocks, p.itmax / p.nout);
    max_blocks = max_blocks * nt + 4 * kNumCandidates + 8;

    const auto t0 = std::chrono::steady_clock::now();
    PTD_HIPB(ptd::launch_batch_clock(b->stream, b->d_t0));
    PTD_HIPB(hipEventRecord(b->ev_fork, b->stream));
    const size_t ng = b->groups.size();
    for (size_t gi = 0; gi < ng; ++gi) {
        Group& g = b->groups[gi];
        PTD_HIPB(hipStreamWaitEvent(g.stream, b->ev_fork, 0));  // replays start after the clock stamp
        *g.h_active = (int32_t)g.members.size();
        g.replays = 0;
    }
    auto replay = [&](Group& g, int cand) -> int {
        PTD_HIPB(hipGraphLaunch(g.exec[cand], g.stream));
        ++g.replays;
        return PTD_OK;
    };
    auto sync_all = [&]() -> int {
        for (auto& g : b->groups) PTD_HIPB(hipStreamSynchronize(g.stream));
        return PTD_OK;
    };
    bool tune = false;
    for (auto& g : b->groups) tune = tune || g.chosen < 0;
    if (tune) {
        // online tiling choice per group: one untimed replay, then kTuneRounds rounds over the candidates in
        // turn (so that members finishing during the tuning weigh on every candidate alike); the least
        // median replay time wins. The groups' replays run concurrently, as they do afterwards.
        constexpr int kTuneRounds = 3;
        for (auto& g : b->groups)
            if (g.chosen < 0 && replay(g, 0) != PTD_OK) return PTD_ERR_HIP;
        if (sync_all() != PTD_OK) return PTD_ERR_HIP;
        std::vector<std::vector<double>> ms_of(ng * kNumCandidates);
        for (int r = 0; r < kTuneRounds; ++r)
            for (int c = 0; c < kNumCandidates; ++c) {
                for (size_t gi = 0; gi < ng; ++gi) {
                    Group& g = b->groups[gi];
                    if (g.chosen >= 0 || *g.h_active == 0) continue;
                    PTD_HIPB(hipEventRecord(g.ev_a, g.stream));
                    if (replay(g, c) != PTD_OK) return PTD_ERR_HIP;
                    PTD_HIPB(hipEventRecord(g.ev_b, g.stream));
                }
                for (size_t gi = 0; gi < ng; ++gi) {
                    Group& g = b->groups[gi];
                    if (g.chosen >= 0 || g.replays < 2 + r * kNumCandidates + c) continue;
                    PTD_HIPB(hipEventSynchronize(g.ev_b));
                    float ms = 0;
                    PTD_HIPB(hipEventElapsedTime(&ms, g.ev_a, g.ev_b));
                    ms_of[gi * kNumCandidates + c].push_back(ms);
                }
            }
        for (size_t gi = 0; gi < ng; ++gi) {
            Group& g = b->groups[gi];
            if (g.chosen >= 0) continue;
            int best = 0;
            for (int c = 0; c < kNumCandidates; ++c) {
                auto& v = ms_of[gi * kNumCandidates + c];
                std::sort(v.begin(), v.end());
                g.cand_ms[c] = v.empty() ? HUGE_VAL : v[v.size() / 2];
                if (g.cand_ms[c] < g.cand_ms[best]) best = c;
            }
            g.chosen = best;
        }
    }
    // every running group replays on its own stream, two sets of kReplaysPerPoll replays in flight: when
    // the older set has ended the host reads the group's running count (written by each replay) and, while
    // members run, queues another set, so no group waits for another one or for the host
    constexpr int kReplaysPerPoll = 4;
    if (sync_all() != PTD_OK) return PTD_ERR_HIP;
    std::vector<int> inflight(ng, 0), oldest(ng, 0);
    auto queue_set = [&](size_t gi) -> int {
        Group& g = b->groups[gi];
        if (g.replays > max_blocks) return ptd::set_error(PTD_ERR_STATE, "batch did not finish within itmax");
        for (int r = 0; r < kReplaysPerPoll; ++r)
            if (replay(g, g.chosen) != PTD_OK) return PTD_ERR_HIP;
        PTD_HIPB(hipEventRecord(g.ev_poll[(oldest[gi] + inflight[gi]) & 1], g.stream));
        ++inflight[gi];
        return PTD_OK;
    };
    for (size_t gi = 0; gi < ng; ++gi)
        for (int k = 0; k < 2 && *b->groups[gi].h_active > 0; ++k) {
            const int st = queue_set(gi);
            if (st != PTD_OK) return st;
        }
    for (;;) {
        bool any = false, progressed = false;
        for (size_t gi = 0; gi < ng; ++gi) {
            if (inflight[gi] == 0) continue;
            any = true;
            Group& g = b->groups[gi];
            const hipError_t q = hipEventQuery(g.ev_poll[oldest[gi]]);
            if (q == hipErrorNotReady) continue;
            PTD_HIPB(q);
            progressed = true;
            oldest[gi] ^= 1;
            --inflight[gi];
            if (*g.h_active > 0 && inflight[gi] < 2) {
                const int st = queue_set(gi);
                if (st != PTD_OK) return st;
            }
        }
        if (!any) break;
        if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    if (sync_all() != PTD_OK) return PTD_ERR_HIP;
    const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();

    std::vector<BatchMember> dmh(b->hm.size());
    PTD_HIPB(hipMemcpyAsync(dmh.data(), b->dm, dmh.size() * sizeof(BatchMember), hipMemcpyDeviceToHost, b->stream));
    uint64_t t0d = 0;
    PTD_HIPB(hipMemcpyAsync(&t0d, b->d_t0, sizeof(uint64_t), hipMemcpyDeviceToHost, b->stream));
    std::vector<int64_t> it((size_t)b->n * b->nt_cap);
    std::vector<double> er((size_t)b->n * b->nt_cap);
    PTD_HIPB(hipMemcpyAsync(it.data(), b->d_iters, it.size() * sizeof(int64_t), hipMemcpyDeviceToHost, b->stream));
    PTD_HIPB(hipMemcpyAsync(er.data(), b->d_errs, er.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    PTD_HIPB(hipStreamSynchronize(b->stream));
    int worst = PTD_OK;
    for (int m = 0; m < b->n; ++m) {
        BatchMember& M = b->hm[(size_t)m];
        M.hpar = dmh[(size_t)m].hpar;  // where the iterate and the fluxes ended up
        M.qpar = dmh[(size_t)m].qpar;
        M.status = dmh[(size_t)m].status;
        M.step = dmh[(size_t)m].step;
        if (status) status[m] = M.status;
        if (M.status != PTD_OK) worst = M.status;
        if (latency_ms) latency_ms[m] = (double)(dmh[(size_t)m].t_done - t0d) / b->wall_khz;
        for (int s = 0; s < nt; ++s) {
            const bool done = s < M.step;
            if (iters) iters[(size_t)m * nt + s] = done ? it[(size_t)m * b->nt_cap + s] : 0;
            if (errs) errs[(size_t)m * nt + s] = done ? er[(size_t)m * b->nt_cap + s] : NAN;
        }
    }
    if (total_ms) *total_ms = wall;
    if (worst == PTD_ERR_NAN) return ptd::set_error(PTD_ERR_NAN, "NaN in at least one member (see status)");
    return PTD_OK;
}

int32_t ptd_batch_group_times(const ptd_batch* b, int32_t* ngroups, int32_t* kind, int32_t* dcase, int32_t* members,
                              double* ms, int32_t cap) {
    if (!b) return ptd::set_error(PTD_ERR_ARG, "batch is NULL");
    const int ng = (int)b->groups.size();
    if (ngroups) *ngroups = ng;
    if (hipSetDevice(b->device) != hipSuccess) return ptd::set_error(PTD_ERR_HIP, "hipSetDevice failed");
    for (int g = 0; g < ng && g < cap; ++g) {
        const Group& G = b->groups[(size_t)g];
        uint64_t span[2] = {0, 0};
        if (hipMemcpy(span, G.d_span, sizeof(span), hipMemcpyDeviceToHost) != hipSuccess)
            return ptd::set_error(PTD_ERR_HIP, "reading the group clocks failed");
        if (kind)
            kind[g] = G.kind == ptd::kSweep3D ? 3 : G.kind == ptd::kSweep2D2 ? 22 : G.kind == ptd::kSweepLds2D ? 20 : 2;
        if (dcase) dcase[g] = G.dcase;
        if (members) members[g] = (int32_t)G.members.size();
        if (ms) ms[g] = b->wall_khz > 0 ? (double)span[1] / b->wall_khz : 0.0;
    }
    return PTD_OK;
}

int32_t ptd_batch_tiling(const ptd_batch* b, int32_t* tz3d, int32_t* rows2d, double* candidate_ms, int32_t ncand) {
    if (!b) return ptd::set_error(PTD_ERR_ARG, "batch is NULL");
    // the choices of the largest 3D group and of the largest single-level 2D group; per candidate the sum
    // over the groups of their median replay times while tuning
    const Group *g3 = nullptr, *g2 = nullptr;
    bool tuned = !b->groups.empty();
    for (const auto& g : b->groups) {
        tuned = tuned && g.chosen >= 0;
        if (g.kind == ptd::kSweep3D && (!g3 || g.members.size() > g3->members.size())) g3 = &g;
        if (g.kind == ptd::kSweep2D && (!g2 || g.members.size() > g2->members.size())) g2 = &g;
    }
    if (tz3d) *tz3d = kCandidates[g3 && g3->chosen >= 0 ? g3->chosen : 0].tz3d;
    if (rows2d) *rows2d = kCandidates[g2 && g2->chosen >= 0 ? g2->chosen : 0].rows2d == 1 ? 8 : 16;
    if (candidate_ms)
        for (int k = 0; k < ncand && k < kNumCandidates; ++k) {
            double sum = 0;
            for (const auto& g : b->groups)
                if (std::isfinite(g.cand_ms[k])) sum += g.cand_ms[k];
            candidate_ms[k] = sum;
        }
    return tuned ? PTD_OK : PTD_ERR_STATE;
}

}  // extern "C"
