// rbc_ensemble.cpp -- batches of independent runs of uneven size, stepped together through
// one captured HIP graph, with a launch-shape autotuner per size class.
//
// The reference runs one simulation per process (run/script_onset_sdc.sh:8 loops over
// processes).  Here every member is a full context (its own factors, buffers and stream);
// K steps of every member are captured once into a graph and each rbc_ensemble_step replays
// it.  Members of one size class on the batched path (FP32 storage, 64..256-point grids,
// L <= 32) advance together: one launch per stage for the whole class (batch_step), so a
// small member no longer pays a full launch per stage; the other members are independent
// branches of the graph.  The first step of a member with no AB2 history is an Euler step
// with other coefficients, so a call whose members include one without history runs its
// first step eagerly for every member before the graph is used.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "rbc_ctx.hpp"

using namespace rbc;

struct rbc_ensemble {
    std::vector<rbc_ctx*> m;
    int device = 0;
    hipStream_t s0 = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_steps = 0;
    std::string err;
    std::vector<double> tune_report;  // per member: ytx, rows threads, ms per step
    bool batching = true;
    std::vector<StepBatch*> batches;       // one per size class run as a batch
    std::vector<std::vector<int>> bmembers;  // member indices of each batch
    std::vector<int> solo;                 // members stepped on their own
    ~rbc_ensemble() {
        if (s0) hipStreamSynchronize(s0);
        for (StepBatch* b : batches) batch_destroy(b);
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        for (hipEvent_t e : join) hipEventDestroy(e);
        if (fork) hipEventDestroy(fork);
        for (rbc_ctx* c : m) rbc_destroy(c);
        if (s0) hipStreamDestroy(s0);
    }
};

namespace {

int efail(rbc_ensemble* e, int code, const std::string& msg) {
    e->err = msg;
    return code;
}

void drop_graph(rbc_ensemble* e) {
    if (e->exec) hipGraphExecDestroy(e->exec);
    if (e->graph) hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
    e->graph_steps = 0;
    for (StepBatch* b : e->batches) batch_destroy(b);
    e->batches.clear();
    e->bmembers.clear();
    e->solo.clear();
}

// Level-space bytes of one member that a step writes in one stage and reads in the next (the
// hybrid rows, 6 fields x nq levels x N x KXp complex).
size_t stage_bytes(const rbc_ctx* c) { return (size_t)6 * c->dg.nq * c->dg.N * c->dg.KXp * c->csz(); }

// Batching pays where a member's kernels are too small to fill the device: their launches and
// ramp-up dominate, and a batch runs the class's blocks together (measured on the 64-run
// workload: the 64^2 members 2.5x faster).  Members of 128^2 and more already fill it and
// overlap with each other as graph branches; batched they ran 5-7 % slower (their stage data,
// in flight for the whole batch at once, no longer stays in the last-level cache between
// stages).  So only classes with at most kBatchMemberBytes of stage data per member batch, in
// batches of at most kBatchBytes.
constexpr size_t kBatchMemberBytes = (size_t)8 << 20;
constexpr size_t kBatchBytes = (size_t)64 << 20;

// size classes of two or more members with the same launch shapes become batches; the rest,
// and members with a hook after the nonlinear term or penalization, step on their own
void build_batches(rbc_ensemble* e) {
    std::map<std::tuple<int, int, int, int, int, int, int>, std::vector<int>> cls;
    for (size_t i = 0; i < e->m.size(); ++i) {
        rbc_ctx* c = e->m[i];
        if (!e->batching || c->post_nonlinear || c->pen.eta > 0.0 || c->rsz != 4 ||
            stage_bytes(c) > kBatchMemberBytes) {
            e->solo.push_back((int)i);
            continue;
        }
        cls[std::make_tuple(c->dg.M, c->dg.N, c->dg.L, c->dg.nq, c->prec, c->b.tune_ytx, c->b.tune_rows_threads)]
            .push_back((int)i);
    }
    for (auto& kv : cls) {
        const std::vector<int>& all = kv.second;
        const size_t per = std::max<size_t>(1, kBatchBytes / stage_bytes(e->m[all[0]]));
        for (size_t j0 = 0; j0 < all.size(); j0 += per) {
            const std::vector<int> idx(all.begin() + j0, all.begin() + std::min(all.size(), j0 + per));
            StepBatch* sb = nullptr;
            if (idx.size() >= 2) {
                std::vector<BatchMember> bm;
                for (int i : idx) {
                    rbc_ctx* c = e->m[i];
                    bm.push_back({&c->dg, &c->b, c->p.Pr, c->p.dt, c->fac.L0, c->fac.L1, c->fac.stride});
                }
                std::string why;
                sb = batch_create(bm, why);
            }
            if (sb) {
                e->batches.push_back(sb);
                e->bmembers.push_back(idx);
            } else {
                e->solo.insert(e->solo.end(), idx.begin(), idx.end());
            }
        }
    }
}

// The first of a call's steps when some member has no AB2 history: every member advances by
// exactly one step here (an Euler step for members without history, an AB2 step for the others),
// so that the graph replays that follow start from members that all have history.
void first_step_all(rbc_ensemble* e) {
    for (rbc_ctx* c : e->m) run_steps(c, 1);
}

void capture(rbc_ensemble* e, int K) {
    drop_graph(e);
    build_batches(e);  // device tables are allocated before the capture begins
    HIPCHK(hipStreamBeginCapture(e->s0, hipStreamCaptureModeThreadLocal));
    HIPCHK(hipEventRecord(e->fork, e->s0));
    for (int i : e->solo) {
        rbc_ctx* c = e->m[i];
        HIPCHK(hipStreamWaitEvent(c->stream, e->fork, 0));
        run_steps(c, K);
        HIPCHK(hipEventRecord(e->join[i], c->stream));
    }
    // each batch is a branch of its own (on its first member's stream), so that size classes
    // overlap as the members' branches do
    for (size_t k = 0; k < e->batches.size(); ++k) {
        const int i0 = e->bmembers[k][0];
        hipStream_t bs = e->m[i0]->stream;
        HIPCHK(hipStreamWaitEvent(bs, e->fork, 0));
        for (int st = 0; st < K; ++st) batch_step(e->batches[k], bs);
        HIPCHK(hipEventRecord(e->join[i0], bs));
    }
    for (int i : e->solo) HIPCHK(hipStreamWaitEvent(e->s0, e->join[i], 0));
    for (const auto& v : e->bmembers) HIPCHK(hipStreamWaitEvent(e->s0, e->join[v[0]], 0));
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamEndCapture(e->s0, &g));
    e->graph = g;
    HIPCHK(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
    e->graph_steps = K;
}

// device copies of the state and history of one member (autotuning steps it and restores)
struct Snapshot {
    std::vector<void*> bufs;
    bool have_hist = false;
};

std::vector<std::pair<void*, size_t>> state_spans(rbc_ctx* c) {
    const size_t field = (size_t)c->dg.P * c->dg.L * c->csz(), vec = (size_t)c->dg.L * c->csz();
    Buffers& b = c->b;
    return {{b.X[0], field}, {b.X[1], field}, {b.X[2], field}, {b.XF, vec}, {b.XG, vec},
            {b.Rp[0], field}, {b.Rp[1], field}, {b.Rp[2], field}, {b.RpF, vec}, {b.RpG, vec},
            {b.health, 2 * sizeof(int)}};
}

void take(rbc_ctx* c, Snapshot& s) {
    s.have_hist = c->have_hist;
    for (auto& sp : state_spans(c)) {
        void* p = nullptr;
        HIPCHK(hipMalloc(&p, sp.second));
        s.bufs.push_back(p);
        HIPCHK(hipMemcpyAsync(p, sp.first, sp.second, hipMemcpyDeviceToDevice, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
}

void restore(rbc_ctx* c, Snapshot& s) {
    auto spans = state_spans(c);
    for (size_t i = 0; i < spans.size(); ++i)
        HIPCHK(hipMemcpyAsync(spans[i].first, s.bufs[i], spans[i].second, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (void* p : s.bufs) hipFree(p);
    s.bufs.clear();
    c->have_hist = s.have_hist;
}

double time_steps(rbc_ctx* c, int reps) {
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    run_steps(c, 1);  // warm-up (and the Euler start)
    HIPCHK(hipEventRecord(a, c->stream));
    run_steps(c, reps);
    HIPCHK(hipEventRecord(b, c->stream));
    HIPCHK(hipEventSynchronize(b));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, a, b));
    hipEventDestroy(a);
    hipEventDestroy(b);
    return ms / reps;
}

}  // namespace

extern "C" {

int rbc_ensemble_create(int n, const rbc_params* params, rbc_ensemble** out) {
    if (!out || n < 1 || !params) return RBC_ERR_ARG;
    *out = nullptr;
    rbc_ensemble* e = new rbc_ensemble;
    e->device = params[0].device;
    for (int i = 0; i < n; ++i) {
        if (params[i].device != e->device) {
            delete e;
            return fail(nullptr, RBC_ERR_ARG, "all members of an ensemble must use one device");
        }
        rbc_ctx* c = nullptr;
        const int rc = rbc_create(&params[i], &c);
        if (rc != RBC_OK) {
            const std::string msg = std::string("member ") + std::to_string(i) + ": " + rbc_last_error(nullptr);
            delete e;
            return fail(nullptr, rc, msg);
        }
        e->m.push_back(c);
    }
    try {
        hipSetDevice(e->device);
        HIPCHK(hipStreamCreateWithFlags(&e->s0, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&e->fork, hipEventDisableTiming));
        e->join.resize(n);
        for (auto& j : e->join) HIPCHK(hipEventCreateWithFlags(&j, hipEventDisableTiming));
    } catch (const HipError& err) {
        const std::string msg = std::string(err.what) + ": " + hipGetErrorString(err.e);
        delete e;
        return fail(nullptr, RBC_ERR_HIP, msg);
    }
    *out = e;
    return RBC_OK;
}

void rbc_ensemble_destroy(rbc_ensemble* e) { delete e; }

const char* rbc_ensemble_last_error(const rbc_ensemble* e) { return e ? e->err.c_str() : ""; }

int rbc_ensemble_size(const rbc_ensemble* e) { return e ? (int)e->m.size() : 0; }

rbc_ctx* rbc_ensemble_member(rbc_ensemble* e, int i) {
    if (!e || i < 0 || i >= (int)e->m.size()) return nullptr;
    return e->m[i];
}

int rbc_ensemble_tune(rbc_ensemble* e, double* report) {
    if (!e) return RBC_ERR_ARG;
    // one search per size class (M, N, L, precision); the first member of a class is timed
    // with every candidate (its state is restored afterwards) and the winner applies to all
    std::map<std::tuple<int, int, int, int>, std::pair<int, int>> best;
    std::map<std::tuple<int, int, int, int>, double> best_ms;
    try {
        hipSetDevice(e->device);
        drop_graph(e);
        for (rbc_ctx* c : e->m) {
            const auto key = std::make_tuple(c->dg.M, c->dg.N, c->dg.L, c->prec);
            if (best.count(key)) continue;
            Snapshot snap;
            take(c, snap);
            const size_t cs = c->csz();
            const char* gen = getenv("RBC_GENERIC_FFT");
            const bool forced = gen && gen[0] == '1';
            // FP32 grids of 64..256 points run the group transforms: their knobs are the kx
            // columns per y-FFT block (2, 4 or 8 waves) and the threads per row block (2, 4 or
            // 8 waves); the generic LDS kernels take the column tile and 64..256 threads
            const bool gy = c->rsz == 4 && !forced && c->dg.N >= 64 && c->dg.N <= 256;
            const bool gr = c->rsz == 4 && !forced && c->dg.M >= 64 && c->dg.M <= 256;
            std::vector<int> txs, ths;
            if (gy) {
                for (int w : {2, 4, 8}) txs.push_back(w * 512 / c->dg.N);
            } else {
                for (int tx = 1; tx <= 16; tx *= 2)
                    if ((size_t)2 * tx * c->dg.N * cs <= 65536) txs.push_back(tx);
            }
            if (gr) {
                for (int w : {2, 4, 8})
                    if (c->dg.N % (w * 512 / c->dg.M) == 0) ths.push_back(64 * w);
            } else {
                ths = {64, 128, 256};
            }
            // coordinate search: y-FFT tile first (rows at their default), then the row kernel
            int btx = txs.back(), bth = 0;
            double bms = 1e30;
            for (int tx : txs) {
                c->b.tune_ytx = tx;
                c->b.tune_rows_threads = 0;
                const double ms = time_steps(c, 3);
                if (ms < bms) {
                    bms = ms;
                    btx = tx;
                }
            }
            for (int th : ths) {
                c->b.tune_ytx = btx;
                c->b.tune_rows_threads = th;
                const double ms = time_steps(c, 3);
                if (ms < bms) {
                    bms = ms;
                    bth = th;
                }
            }
            restore(c, snap);
            best[key] = {btx, bth};
            best_ms[key] = bms;
        }
        for (size_t i = 0; i < e->m.size(); ++i) {
            rbc_ctx* c = e->m[i];
            const auto key = std::make_tuple(c->dg.M, c->dg.N, c->dg.L, c->prec);
            c->b.tune_ytx = best[key].first;
            c->b.tune_rows_threads = best[key].second;
            if (report) {
                report[3 * i] = best[key].first;
                report[3 * i + 1] = best[key].second;
                report[3 * i + 2] = best_ms[key];
            }
        }
    } catch (const HipError& err) {
        return efail(e, RBC_ERR_HIP, std::string(err.what) + ": " + hipGetErrorString(err.e));
    }
    return RBC_OK;
}

int rbc_ensemble_batching(rbc_ensemble* e, int on) {
    if (!e || (on != 0 && on != 1)) return RBC_ERR_ARG;
    try {
        hipSetDevice(e->device);
        HIPCHK(hipStreamSynchronize(e->s0));
        drop_graph(e);
    } catch (const HipError& err) {
        return efail(e, RBC_ERR_HIP, std::string(err.what) + ": " + hipGetErrorString(err.e));
    }
    e->batching = on == 1;
    return RBC_OK;
}

int rbc_ensemble_batched_members(const rbc_ensemble* e) {
    if (!e) return 0;
    int n = 0;
    for (const auto& v : e->bmembers) n += (int)v.size();
    return n;
}

int rbc_ensemble_step(rbc_ensemble* e, int nsteps, int steps_per_launch) {
    if (!e || nsteps < 0 || steps_per_launch < 0) return RBC_ERR_ARG;
    try {
        hipSetDevice(e->device);
        int left = nsteps;
        if (left > 0) {
            bool any_euler = false;
            for (rbc_ctx* c : e->m) any_euler = any_euler || !c->have_hist;
            if (any_euler) {
                first_step_all(e);
                left -= 1;
            }
        }
        if (steps_per_launch == 0) {  // eager: every member's kernels launched one by one
            for (int s = 0; s < left; ++s)
                for (rbc_ctx* c : e->m) run_steps(c, 1);
        } else {
            if (left >= steps_per_launch) {
                // the members' own streams must be idle before their work moves into the graph
                for (rbc_ctx* c : e->m) HIPCHK(hipStreamSynchronize(c->stream));
                if (!e->exec || e->graph_steps != steps_per_launch) capture(e, steps_per_launch);
                for (; left >= steps_per_launch; left -= steps_per_launch) HIPCHK(hipGraphLaunch(e->exec, e->s0));
                HIPCHK(hipStreamSynchronize(e->s0));
            }
            for (int s = 0; s < left; ++s)
                for (rbc_ctx* c : e->m) run_steps(c, 1);
        }
        for (rbc_ctx* c : e->m) HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipStreamSynchronize(e->s0));
    } catch (const HipError& err) {
        return efail(e, RBC_ERR_HIP, std::string(err.what) + ": " + hipGetErrorString(err.e));
    }
    return RBC_OK;
}

}  // extern "C"
