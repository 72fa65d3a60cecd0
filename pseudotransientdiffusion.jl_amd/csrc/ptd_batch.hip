// ptd_batch.hip -- batches of independent float64 PT solves (include/ptd.h: ptd_batch_*).
//
// Every member is a full problem of its own (2D or 3D, lin / step / nonlin, its own grid and scalars),
// solved over nt physical steps exactly as ptd_solve would solve it alone: the same tile functions,
// the same checking tiling (so the same residual partials and stopping decisions), the same
// iteration counts and the same bits. What changes is the schedule:
//  - members are grouped by (sweep kind, dcase, arith), the kind being the sweep ptd_solve runs for
//    them (two-level 3D, two-level 2D on large grids, single-level 2D); one launch covers the tiles of
//    every member of a group (a table maps each block to its member, tile and z chunk depth);
//  - the PT loop of every member runs on the device: after each check block of nout iterations a
//    controller kernel applies the reference loop's rules per member (scripts/diff_3D/
//    diff_3D_nonlin_multixpu.jl:104-118) and members that are done drop out of later launches;
//  - one check block (all groups on their own streams, then reduction, controller and Hold .= H) is
//    captured once as a HIP graph and replayed until no member is running; groups of 2D grids wider
//    than the 3D ones run several check blocks per replay (group_sub_blocks), as they need
//    proportionally more PT iterations;
//  - the tiling of the non-checking launches (z chunk depth in 3D, rows per block in 2D) is chosen
//    online: the first replays try each candidate graph (each replay is real progress, results do not
//    depend on the tiling) and the fastest is kept.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
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
    int sub = 1;                // check blocks per replay
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

// Launches per check block of nout iterations: two-level sweeps take two iterations per launch. The
// single-level 2D groups (grids too small for ptd_solve to pick the two-level 2D sweep) still run
// their first nout - 2 iterations as two-level pairs, which give the same bits; only the checking
// launch must use the single-problem tiling (its residual partials decide when a member stops), and it
// follows one single-level launch.
int launches_per_block(int kind, int nout) { return kind == ptd::kSweep2D ? nout / 2 + 1 : nout / 2; }

}  // namespace

struct ptd_batch {
    int device = 0;
    int n = 0;
    int nout = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_a = nullptr, ev_b = nullptr;
    std::vector<hipEvent_t> ev_join;
    std::vector<ptd_params> params;
    std::vector<BatchMember> hm;
    BatchMember* dm = nullptr;
    std::vector<void*> allocs;
    std::vector<Group> groups;
    int64_t* d_iters = nullptr;
    double* d_errs = nullptr;
    int nt_cap = 0;
    int32_t* d_active = nullptr;
    int32_t* h_active = nullptr;
    uint64_t* d_t0 = nullptr;
    hipGraph_t graph[kNumCandidates] = {};
    hipGraphExec_t exec[kNumCandidates] = {};
    int chosen = -1;
    double cand_ms[kNumCandidates] = {};
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
    for (int c = 0; c < kNumCandidates; ++c) {
        if (b->exec[c]) (void)hipGraphExecDestroy(b->exec[c]);
        if (b->graph[c]) (void)hipGraphDestroy(b->graph[c]);
        b->exec[c] = nullptr;
        b->graph[c] = nullptr;
    }
}

void free_batch(ptd_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    destroy_graphs(b);
    for (void* p : b->allocs) (void)hipFree(p);
    if (b->h_active) (void)hipHostFree(b->h_active);
    for (auto& g : b->groups)
        if (g.stream) (void)hipStreamDestroy(g.stream);
    for (auto e : b->ev_join) (void)hipEventDestroy(e);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->ev_a) (void)hipEventDestroy(b->ev_a);
    if (b->ev_b) (void)hipEventDestroy(b->ev_b);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

// tiles of one member's launch: rows per block (two-level 2D: rows per tile chunk), z chunk depth tz (3D)
int member_tiles(const ptd_params& p, int kind, int rows, int tz) {
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

// one replay: every group's check blocks of nout iterations on its own stream (each but the last closed
// by the group's own controller and Hold .= H), then reduction, the controller of every member,
// Hold .= H and the running-member count to the host
int enqueue_block(ptd_batch* b, int cand) {
    PTD_HIPB(hipEventRecord(b->ev_fork, b->stream));
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
        Group& g = b->groups[gi];
        const int nm = (int)g.members.size();
        PTD_HIPB(hipStreamWaitEvent(g.stream, b->ev_fork, 0));
        const int lpb = launches_per_block(g.kind, b->nout);
        for (int sb = 0; sb < g.sub; ++sb) {
            if (sb > 0) {
                PTD_HIPB(ptd::launch_batch_control_list(g.stream, b->dm, g.d_members, nm, b->d_iters, b->d_errs,
                                                        b->nt_cap));
                PTD_HIPB(ptd::launch_batch_assign_list(g.stream, b->dm, g.d_members, nm));
            }
            for (int li = 0; li < lpb; ++li) {
                const bool check = li == lpb - 1;
                if (g.kind == ptd::kSweep2D && li < lpb - 2) {  // a two-level pair
                    PTD_HIPB(ptd::launch_batch_sweep(g.stream, ptd::kSweep2D2, g.dcase, g.arith,
                                                     ptd::batch_rows_nocheck(ptd::kSweep2D2, g.dcase, g.arith, 0),
                                                     false, b->dm, g.d_t2, g.n_t2, li));
                    continue;
                }
                const int rows = check ? ptd::batch_rows_check(g.kind, g.dcase, g.arith)
                                       : ptd::batch_rows_nocheck(g.kind, g.dcase, g.arith, kCandidates[cand].rows2d);
                PTD_HIPB(ptd::launch_batch_sweep(g.stream, g.kind, g.dcase, g.arith, rows, check, b->dm,
                                                 check ? g.d_check : g.d_nc[cand], check ? g.n_check : g.n_nc[cand],
                                                 li));
            }
            PTD_HIPB(ptd::launch_batch_reduce(g.stream, b->dm, g.d_members, nm));
        }
        PTD_HIPB(hipEventRecord(b->ev_join[gi], g.stream));
        PTD_HIPB(hipStreamWaitEvent(b->stream, b->ev_join[gi], 0));
    }
    PTD_HIPB(ptd::launch_batch_control(b->stream, b->dm, b->n, b->d_iters, b->d_errs, b->d_active, b->nt_cap));
    PTD_HIPB(ptd::launch_batch_assign(b->stream, b->dm, b->n));
    PTD_HIPB(hipMemcpyAsync(b->h_active, b->d_active, sizeof(int32_t), hipMemcpyDeviceToHost, b->stream));
    return PTD_OK;
}

int build_graphs(ptd_batch* b) {
    destroy_graphs(b);
    for (int c = 0; c < kNumCandidates; ++c) {
        PTD_HIPB(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const int st = enqueue_block(b, c);
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &gr);
        if (st != PTD_OK) {
            if (gr) (void)hipGraphDestroy(gr);
            return st;
        }
        PTD_HIPB(e);
        b->graph[c] = gr;
        PTD_HIPB(hipGraphInstantiate(&b->exec[c], gr, nullptr, nullptr, 0));
    }
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
        hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&b->ev_a) != hipSuccess || hipEventCreate(&b->ev_b) != hipSuccess)
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
        M.lpb = launches_per_block(ptd::batch_kind(p), p.nout);
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
        }
        g->members.push_back(m);
    }
    for (auto& g : b->groups) g.sub = group_sub_blocks(b->params, g.members);
    for (auto& g : b->groups) {
        if (hipStreamCreate(&g.stream) != hipSuccess) return fail_free(ptd::set_error(PTD_ERR_HIP, "stream creation failed"));
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return fail_free(ptd::set_error(PTD_ERR_HIP, "event creation failed"));
        b->ev_join.push_back(e);
        int st = balloc(b, (void**)&g.d_members, g.members.size() * sizeof(int32_t));
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
    if (st == PTD_OK) st = balloc(b, (void**)&b->d_active, sizeof(int32_t));
    if (st == PTD_OK) st = balloc(b, (void**)&b->d_t0, sizeof(uint64_t));
    if (st != PTD_OK) return fail_free(st);
    if (hipHostMalloc((void**)&b->h_active, sizeof(int32_t), hipHostMallocDefault) != hipSuccess)
        return fail_free(ptd::set_error(PTD_ERR_NOMEM, "pinned host allocation failed"));
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
    PTD_HIPB(hipStreamSynchronize(b->stream));
    if (!b->exec[0]) {
        int st = build_graphs(b);
        if (st != PTD_OK) return st;
    }
    int64_t max_blocks = 0;
    for (const auto& p : b->params) max_blocks = std::max<int64_t>(max_blocks, p.itmax / p.nout);
    max_blocks = max_blocks * nt + 4 * kNumCandidates + 8;

    const auto t0 = std::chrono::steady_clock::now();
    PTD_HIPB(ptd::launch_batch_clock(b->stream, b->d_t0));
    int64_t blocks = 0;
    *b->h_active = b->n;
    if (b->chosen < 0) {
        // online tiling choice: one untimed replay, then one timed replay per candidate
        PTD_HIPB(hipGraphLaunch(b->exec[0], b->stream));
        PTD_HIPB(hipStreamSynchronize(b->stream));
        ++blocks;
        int best = 0;
        for (int c = 0; c < kNumCandidates && *b->h_active > 0; ++c) {
            PTD_HIPB(hipEventRecord(b->ev_a, b->stream));
            PTD_HIPB(hipGraphLaunch(b->exec[c], b->stream));
            PTD_HIPB(hipEventRecord(b->ev_b, b->stream));
            PTD_HIPB(hipEventSynchronize(b->ev_b));
            float ms = 0;
            PTD_HIPB(hipEventElapsedTime(&ms, b->ev_a, b->ev_b));
            b->cand_ms[c] = ms;
            if (ms < b->cand_ms[best]) best = c;
            ++blocks;
        }
        b->chosen = best;
    }
    constexpr int kReplaysPerPoll = 4;
    while (*b->h_active > 0) {
        if (blocks > max_blocks) return ptd::set_error(PTD_ERR_STATE, "batch did not finish within itmax");
        for (int r = 0; r < kReplaysPerPoll; ++r) PTD_HIPB(hipGraphLaunch(b->exec[b->chosen], b->stream));
        PTD_HIPB(hipStreamSynchronize(b->stream));
        blocks += kReplaysPerPoll;
    }
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

int32_t ptd_batch_tiling(const ptd_batch* b, int32_t* tz3d, int32_t* rows2d, double* candidate_ms, int32_t ncand) {
    if (!b) return ptd::set_error(PTD_ERR_ARG, "batch is NULL");
    const int c = b->chosen < 0 ? 0 : b->chosen;
    if (tz3d) *tz3d = kCandidates[c].tz3d;
    if (rows2d) *rows2d = kCandidates[c].rows2d == 1 ? 8 : 16;
    if (candidate_ms)
        for (int k = 0; k < ncand && k < kNumCandidates; ++k) candidate_ms[k] = b->cand_ms[k];
    return b->chosen < 0 ? PTD_ERR_STATE : PTD_OK;
}

}  // extern "C"
