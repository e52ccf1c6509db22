// Several operators (objects of uneven size) solved together: their batched
// GMRES runs advance in lockstep, and every Arnoldi step launches the
// (preconditioned) matvecs of all objects still iterating as ONE recorded
// graph, one branch per object, forked from and joined back to a stream of
// its own.  Small objects that could not fill the GPU alone run side by side
// from a single launch.  No reference counterpart: the reference rebuilds an
// operator per task and solves the tasks one after another
// (Code/BatchRunner.h:470-480).
//
// Per object the recurrences are those of Engine::solve_block (same phases,
// same kernels, same order of operations, the same narrowing of a block once a
// quarter of its columns has converged), so each object's iterates equal a
// solve of that object alone.  The orthogonalisation of all objects is issued
// together and shares one synchronisation per step.  Graphs are recorded per set of
// (object, graph kind, right-hand sides, precision) taking part in a step, so
// an object that finishes or falls back to a plainer preconditioner leads to
// a new recording; the recordings live for the call.
#include <algorithm>
#include <map>
#include <vector>

#include "engine.hpp"

namespace mlfma {

namespace {
struct Recorded {
    std::map<std::vector<int>, hipGraphExec_t> graphs;
    hipStream_t ms = nullptr;
    hipEvent_t fork = nullptr, done = nullptr;
    std::vector<hipEvent_t> ev, join;   // per object: inputs ready; its branch joined (inside a capture)
    ~Recorded() {
        if (ms) hipStreamSynchronize(ms);
        for (auto& kv : graphs) hipGraphExecDestroy(kv.second);
        for (hipEvent_t e : ev)
            if (e) hipEventDestroy(e);
        for (hipEvent_t e : join)
            if (e) hipEventDestroy(e);
        if (fork) hipEventDestroy(fork);
        if (done) hipEventDestroy(done);
        if (ms) hipStreamDestroy(ms);
    }
};
}  // namespace

void solve_multi(const std::vector<Engine*>& es, const mlfma_solve_params& sp, const std::vector<const double2*>& B,
                 const std::vector<double2*>& X, const std::vector<int>& R,
                 const std::vector<mlfma_solve_stats*>& stats) {
    const int K = (int)es.size();
    if (K < 1 || (int)B.size() != K || (int)X.size() != K || (int)R.size() != K || (int)stats.size() != K)
        throw Error(MLFMA_ERR_ARG, "solve_multi: one right-hand-side block, solution block and stats array per object");
    for (int k = 0; k < K; ++k) {
        if (!es[k]) throw Error(MLFMA_ERR_ARG, "solve_multi: null operator");
        if (es[k]->P.device != es[0]->P.device) throw Error(MLFMA_ERR_ARG, "solve_multi: all operators on one device");
        if (!es[k]->graphs_on()) throw Error(MLFMA_ERR_STATE, "solve_multi: recorded graphs must be on (and profiling off)");
        for (int j = 0; j < k; ++j)
            if (es[j] == es[k]) throw Error(MLFMA_ERR_ARG, "solve_multi: an operator appears twice");
    }
    Recorded rec;
    MLFMA_HIP(hipSetDevice(es[0]->P.device));
    MLFMA_HIP(hipStreamCreateWithFlags(&rec.ms, hipStreamNonBlocking));
    MLFMA_HIP(hipEventCreateWithFlags(&rec.fork, hipEventDisableTiming));
    MLFMA_HIP(hipEventCreateWithFlags(&rec.done, hipEventDisableTiming));
    rec.ev.assign(K, nullptr);
    rec.join.assign(K, nullptr);
    for (int k = 0; k < K; ++k) {
        MLFMA_HIP(hipEventCreateWithFlags(&rec.ev[k], hipEventDisableTiming));
        MLFMA_HIP(hipEventCreateWithFlags(&rec.join[k], hipEventDisableTiming));
    }

    // per object a stack of states: a block narrowed to its unconverged
    // columns continues as a child state on top of its parent
    std::vector<std::vector<BlockState>> stk(K);
    std::vector<std::vector<NarrowRecord>> recs(K);
    for (int k = 0; k < K; ++k) {
        stk[k].emplace_back();
        es[k]->bs_start(stk[k].back(), sp, B[k], X[k], R[k]);
    }
    auto top = [&](int k) -> BlockState& { return stk[k].back(); };

    // one graph over the matvec bodies of the objects in `open`
    auto launch = [&](const std::vector<int>& open) {
        std::vector<int> key;
        for (int k : open) {
            key.push_back(k);
            key.push_back(es[k]->bs_graph_kind(top(k)));
            key.push_back(top(k).R);
            key.push_back(es[k]->apply_prec);
        }
        // the inputs each object staged on its own stream come first
        for (int k : open) {
            MLFMA_HIP(hipEventRecord(rec.ev[k], es[k]->st));
            MLFMA_HIP(hipStreamWaitEvent(rec.ms, rec.ev[k], 0));
        }
        auto it = rec.graphs.find(key);
        if (it == rec.graphs.end()) {
            hipGraph_t g = nullptr;
            MLFMA_HIP(hipStreamBeginCapture(rec.ms, hipStreamCaptureModeThreadLocal));
            int captured = -1;
            try {
                MLFMA_HIP(hipEventRecord(rec.fork, rec.ms));
                for (int k : open) {
                    captured = k;
                    MLFMA_HIP(hipStreamWaitEvent(es[k]->st, rec.fork, 0));
                    es[k]->set_capturing(true);
                    es[k]->bs_body(top(k));
                    es[k]->set_capturing(false);
                    MLFMA_HIP(hipEventRecord(rec.join[k], es[k]->st));
                    MLFMA_HIP(hipStreamWaitEvent(rec.ms, rec.join[k], 0));
                }
            } catch (...) {
                if (captured >= 0) es[captured]->set_capturing(false);
                hipStreamEndCapture(rec.ms, &g);
                if (g) hipGraphDestroy(g);
                throw;
            }
            MLFMA_HIP(hipStreamEndCapture(rec.ms, &g));
            hipGraphExec_t ex = nullptr;
            const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            MLFMA_HIP(e);
            it = rec.graphs.emplace(key, ex).first;
        }
        MLFMA_HIP(hipGraphLaunch(it->second, rec.ms));
        MLFMA_HIP(hipEventRecord(rec.done, rec.ms));
        for (int k : open) MLFMA_HIP(hipStreamWaitEvent(es[k]->st, rec.done, 0));
    };

    for (;;) {
        std::vector<int> active;
        for (int k = 0; k < K; ++k)
            if (es[k]->bs_active(top(k))) active.push_back(k);
        if (active.empty()) break;
        for (int k : active) es[k]->bs_cycle_begin(top(k));
        for (;;) {
            std::vector<int> open;
            for (int k : active)
                if (es[k]->bs_step_open(top(k))) open.push_back(k);
            if (open.empty()) break;
            for (int k : open) es[k]->bs_pre(top(k), true);
            launch(open);
            for (int k : open) es[k]->bs_post(top(k), true);
            for (int k : open) es[k]->bs_orth(top(k), 0);
            for (int k : open) MLFMA_HIP(hipStreamSynchronize(es[k]->st));
            for (int k : open) es[k]->bs_orth(top(k), 1);
        }
        for (int k : active) {
            es[k]->bs_cycle_end(top(k), false);
            // once a quarter of the block has converged, its other columns go on
            // as a narrower block (the graphs of the new width are recorded anew)
            if (!es[k]->bs_active(top(k))) continue;
            BlockState child;
            NarrowRecord nr;
            if (es[k]->bs_narrow(top(k), child, nr)) {
                stk[k].push_back(std::move(child));
                recs[k].push_back(std::move(nr));
            }
        }
    }
    for (int k = 0; k < K; ++k) {
        while (stk[k].size() > 1) {
            BlockState child = std::move(stk[k].back());
            stk[k].pop_back();
            es[k]->bs_unnarrow(stk[k].back(), child, recs[k].back());
            recs[k].pop_back();
        }
        MLFMA_HIP(hipStreamSynchronize(es[k]->st));
        es[k]->bs_stats(stk[k][0], stats[k]);
    }
}

}  // namespace mlfma
