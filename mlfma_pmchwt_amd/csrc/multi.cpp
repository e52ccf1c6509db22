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
// same kernels, same order of operations), so each object's iterates equal a
// solve of that object alone; only the narrowing of a block whose columns
// have partly converged is left out (an object keeps its width until it is
// done).  The orthogonalisation phases of all objects are issued together and
// share one synchronisation per phase.  Graphs are recorded per set of
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

    std::vector<BlockState> s(K);
    for (int k = 0; k < K; ++k) es[k]->bs_start(s[k], sp, B[k], X[k], R[k]);

    // one graph over the matvec bodies of the objects in `open`
    auto launch = [&](const std::vector<int>& open) {
        std::vector<int> key;
        for (int k : open) {
            key.push_back(k);
            key.push_back(es[k]->bs_graph_kind(s[k]));
            key.push_back(s[k].R);
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
                    es[k]->bs_body(s[k]);
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
            if (es[k]->bs_active(s[k])) active.push_back(k);
        if (active.empty()) break;
        for (int k : active) es[k]->bs_cycle_begin(s[k]);
        for (;;) {
            std::vector<int> open;
            for (int k : active)
                if (es[k]->bs_step_open(s[k])) open.push_back(k);
            if (open.empty()) break;
            for (int k : open) es[k]->bs_pre(s[k], true);
            launch(open);
            for (int k : open) es[k]->bs_post(s[k], true);
            for (int phase = 0; phase < 4; ++phase) {
                for (int k : open) es[k]->bs_orth(s[k], phase);
                if (phase < 3)
                    for (int k : open) MLFMA_HIP(hipStreamSynchronize(es[k]->st));
            }
        }
        for (int k : active) es[k]->bs_cycle_end(s[k], false);
    }
    for (int k = 0; k < K; ++k) {
        MLFMA_HIP(hipStreamSynchronize(es[k]->st));
        es[k]->bs_stats(s[k], stats[k]);
    }
}

}  // namespace mlfma
