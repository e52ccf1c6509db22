// Batch runner over the C ABI: the reference's batch configuration format
// (Code/BatchRunner.h, Code/batch_config.txt) with a persistent operator per
// (mesh, frequency, material) instead of one operator per task and
// polarisation (Code/BatchRunner.h:422-484).
//
// Parsing follows BatchConfigParser::parse (Code/BatchRunner.h:103-302):
// "key = value" lines, '#' comments, the same keys and list separators, and
// the same expansion of incident_theta_range / incident_phi_range.  Tasks are
// expanded in the order of BatchConfig::generateTasks (:55-100): mesh,
// frequency, incidence, algorithm, material.  Every task has the two
// polarisations of executeSingleTask, "h" (pol 90) then "v" (pol 0).
//
// Running groups the tasks that share one operator (mesh, frequency,
// material, equation, Gauss rule).  Dual (bistatic) tasks of a group are
// solved together: every (incidence, polarisation) is one right-hand side of
// a lockstep batched GMRES (chunks of at most 256), and each solution's RCS is
// taken on the scattering grid.  Mono tasks run one monostatic sweep per
// polarisation per group (the reference's mono mode ignores the incidence
// angle, so every mono task of a group has the same result).  MoM (0) and FMM
// (1) tasks are not run here (status MLFMA_ERR_ARG); CGS tasks are solved one
// right-hand side at a time.  With a concurrency above 1, that many host
// threads take groups (objects) from a shared counter, each with its own
// operator, stream and recorded graphs, so objects of uneven size share the
// GPU instead of running one after another; results do not depend on the
// concurrency.  Results stay in memory; file export is the caller's.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/mlfma.h"

namespace {

struct Material {
    double er = 1.0, ei = 0.0, mr = 1.0, mi = 0.0;
};

struct Task {
    std::string nas;
    double freq = 0, th = 0, ph = 0;
    int algorithm = 2;
    Material mat;
};

struct RunResult {
    int status = -1;   // -1 not run yet, else an MLFMA_* code
    std::vector<double> rcs;   // 3 per grid point
    mlfma_solve_stats stats{};
};

std::string trim(const std::string& s) {
    const size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

std::vector<std::string> split(const std::string& s, char d) {
    std::vector<std::string> out;
    std::stringstream ss(s);
    std::string item;
    while (std::getline(ss, item, d)) {
        item = trim(item);
        if (!item.empty()) out.push_back(item);
    }
    return out;
}

}  // namespace

struct mlfma_batch {
    std::vector<std::string> nas;
    std::vector<double> freqs;
    std::vector<std::pair<double, double>> inc;
    double sth = 0, sthf = 180, sph = 0, sphf = 360, step = 1;
    std::vector<int> algorithms{2};
    int ie = 1, solver = 0, gauss = 7;
    bool mono = false;
    std::vector<Material> mats{Material{}};
    std::vector<Task> tasks;
    std::vector<double> grid_th, grid_ph;
    std::vector<RunResult> res;   // [task][pol], pol 0 = "h", 1 = "v"
    int concurrency = 1;
    std::string err;

    void parse(std::istream& in) {
        bool mu_set = false;
        double mur = 1, mui = 0;
        std::string line;
        while (std::getline(in, line)) {
            if (line.empty() || line[0] == '#') continue;
            const size_t eq = line.find('=');
            if (eq == std::string::npos) continue;
            const std::string key = trim(line.substr(0, eq)), val = trim(line.substr(eq + 1));
            if (key == "nas_files") {
                nas = split(val, ';');
            } else if (key == "frequencies") {
                freqs.clear();
                for (auto& f : split(val, ',')) freqs.push_back(std::stod(f));
            } else if (key == "incident_angles") {
                inc.clear();
                for (auto& pr : split(val, ',')) {
                    const auto p = split(pr, ':');
                    if (p.size() == 2) inc.emplace_back(std::stod(p[0]), std::stod(p[1]));
                }
            } else if (key == "incident_theta_range") {
                const auto p = split(val, ':');
                if (p.size() == 3) {
                    const double a = std::stod(p[0]), b = std::stod(p[1]), s = std::stod(p[2]);
                    const int cnt = (int)std::round((b - a) / s) + 1;
                    for (int i = 0; i < cnt; ++i) inc.emplace_back(a + i * s, 0.0);
                }
            } else if (key == "incident_phi_range") {
                const auto p = split(val, ':');
                if (p.size() == 3) {
                    const double a = std::stod(p[0]), b = std::stod(p[1]), s = std::stod(p[2]);
                    std::vector<std::pair<double, double>> out;
                    if (inc.empty()) {
                        for (double x = a; x <= b + 1e-9; x += s) out.emplace_back(90.0, x);
                    } else {
                        for (const auto& tp : inc)
                            for (double x = a; x <= b + 1e-9; x += s) out.emplace_back(tp.first, x);
                    }
                    inc = out;
                }
            } else if (key == "sca_theta" || key == "sca_phi") {
                const auto p = split(val, ':');
                if (p.size() == 2) {
                    (key == "sca_theta" ? sth : sph) = std::stod(p[0]);
                    (key == "sca_theta" ? sthf : sphf) = std::stod(p[1]);
                }
            } else if (key == "step") {
                step = std::stod(val);
            } else if (key == "algorithm") {
                algorithms.clear();
                for (auto& a : split(val, ',')) algorithms.push_back(std::stoi(a));
                if (algorithms.empty()) algorithms.push_back(2);
            } else if (key == "integral_equation") {
                ie = std::stoi(val);
            } else if (key == "matrix_solver") {
                solver = std::stoi(val);
            } else if (key == "mono_dual") {
                mono = val == "mono";
            } else if (key == "gauss_points") {
                gauss = std::stoi(val);
            } else if (key == "epsilon_r") {
                Material m;
                const auto p = split(val, ',');
                if (p.size() >= 1) m.er = std::stod(p[0]);
                if (p.size() >= 2) m.ei = std::stod(p[1]);
                if (mu_set) m.mr = mur, m.mi = mui;
                mats = {m};
            } else if (key == "epsilon_r_list") {
                mats.clear();
                for (auto& item : split(val, ';')) {
                    const auto p = split(item, ',');
                    if (p.empty()) continue;
                    Material m;
                    m.er = std::stod(p[0]);
                    if (p.size() >= 2) m.ei = std::stod(p[1]);
                    if (mu_set) m.mr = mur, m.mi = mui;
                    mats.push_back(m);
                }
                if (mats.empty()) mats.push_back(Material{});
            } else if (key == "mu_r") {
                const auto p = split(val, ',');
                if (p.size() >= 1) mur = std::stod(p[0]);
                if (p.size() >= 2) mui = std::stod(p[1]);
                mu_set = true;
                for (auto& m : mats) m.mr = mur, m.mi = mui;
            } else if (key == "material_list") {
                mats.clear();
                for (auto& item : split(val, ';')) {
                    const auto p = split(item, ',');
                    if (p.size() < 2) continue;
                    Material m;
                    m.er = std::stod(p[0]);
                    m.ei = std::stod(p[1]);
                    if (p.size() >= 3) m.mr = std::stod(p[2]);
                    if (p.size() >= 4) m.mi = std::stod(p[3]);
                    if (mu_set && p.size() < 3) m.mr = mur, m.mi = mui;
                    mats.push_back(m);
                }
                if (mats.empty()) mats.push_back(Material{});
            }
            // omp_threads and unknown keys: no effect here
        }
        for (const auto& n : nas)
            for (double f : freqs)
                for (const auto& a : inc)
                    for (int alg : algorithms)
                        for (const auto& m : mats) {
                            Task t;
                            t.nas = n;
                            t.freq = f;
                            t.th = a.first;
                            t.ph = a.second;
                            t.algorithm = alg;
                            t.mat = m;
                            tasks.push_back(t);
                        }
        res.assign(tasks.size() * 2, RunResult{});
        int cnt = 0;
        if (step > 0 && mlfma_scatter_grid(sth, sthf, sph, sphf, step, &cnt, nullptr, nullptr) == MLFMA_OK && cnt > 0) {
            grid_th.resize(cnt);
            grid_ph.resize(cnt);
            mlfma_scatter_grid(sth, sthf, sph, sphf, step, &cnt, grid_th.data(), grid_ph.data());
        }
    }
};

namespace {

int load(std::istream& in, mlfma_batch** out) {
    auto b = std::make_unique<mlfma_batch>();
    try {
        b->parse(in);
    } catch (const std::exception&) {
        return MLFMA_ERR_ARG;   // a malformed number (std::stod / std::stoi)
    }
    *out = b.release();
    return MLFMA_OK;
}

}  // namespace

extern "C" {

int mlfma_batch_load(const char* path, mlfma_batch** out) {
    if (!path || !out) return MLFMA_ERR_ARG;
    *out = nullptr;
    std::ifstream f(path);
    if (!f.is_open()) return MLFMA_ERR_ARG;
    return load(f, out);
}

int mlfma_batch_load_text(const char* text, mlfma_batch** out) {
    if (!text || !out) return MLFMA_ERR_ARG;
    *out = nullptr;
    std::istringstream in{std::string(text)};
    return load(in, out);
}

void mlfma_batch_destroy(mlfma_batch* b) { delete b; }

int mlfma_batch_count(const mlfma_batch* b, int* n_tasks, int* n_points) {
    if (!b) return MLFMA_ERR_ARG;
    if (n_tasks) *n_tasks = (int)b->tasks.size();
    if (n_points) *n_points = (int)b->grid_th.size();
    return MLFMA_OK;
}

int mlfma_batch_task(const mlfma_batch* b, int i, double* d7, int* i5, char* nas, int nas_cap) {
    if (!b || i < 0 || i >= (int)b->tasks.size()) return MLFMA_ERR_ARG;
    const Task& t = b->tasks[i];
    if (d7) {
        const double v[7] = {t.freq, t.th, t.ph, t.mat.er, t.mat.ei, t.mat.mr, t.mat.mi};
        std::copy(v, v + 7, d7);
    }
    if (i5) {
        const int v[5] = {t.algorithm, b->ie, b->solver, b->gauss, b->mono ? 1 : 0};
        std::copy(v, v + 5, i5);
    }
    if (nas && nas_cap > 0) {
        const size_t n = std::min(t.nas.size(), (size_t)nas_cap - 1);
        std::copy(t.nas.begin(), t.nas.begin() + n, nas);
        nas[n] = 0;
    }
    return MLFMA_OK;
}

int mlfma_batch_grid(const mlfma_batch* b, double* theta, double* phi) {
    if (!b || !theta || !phi) return MLFMA_ERR_ARG;
    std::copy(b->grid_th.begin(), b->grid_th.end(), theta);
    std::copy(b->grid_ph.begin(), b->grid_ph.end(), phi);
    return MLFMA_OK;
}

int mlfma_batch_run(mlfma_batch* b, const mlfma_params* base, const mlfma_solve_params* sp_in) {
    if (!b) return MLFMA_ERR_ARG;
    const int npts = (int)b->grid_th.size();
    if (npts < 1) return MLFMA_ERR_ARG;
    // operator groups, in task order
    using Key = std::tuple<std::string, double, double, double, double, double>;
    std::map<Key, std::vector<int>> groups;
    std::vector<Key> order;
    for (size_t i = 0; i < b->tasks.size(); ++i) {
        const Task& t = b->tasks[i];
        if (t.algorithm != 2) {   // MoM / FMM are not provided by this engine
            b->res[2 * i].status = b->res[2 * i + 1].status = MLFMA_ERR_ARG;
            continue;
        }
        const Key k{t.nas, t.freq, t.mat.er, t.mat.ei, t.mat.mr, t.mat.mi};
        if (!groups.count(k)) order.push_back(k);
        groups[k].push_back((int)i);
    }
    // one object (operator group) per call; returns its worst status
    auto run_group = [&](const Key& k) -> int {
        int worst = MLFMA_OK;
        const std::vector<int>& ids = groups.at(k);
        const Task& t0 = b->tasks[ids[0]];
        mlfma_params p;
        if (base) p = *base;
        else mlfma_params_default(&p);
        p.integral_equation = b->ie;
        p.gauss_points = b->gauss;
        p.frequency_hz = t0.freq;
        if (b->ie == MLFMA_PMCHWT) {   // the medium applies to PMCHWT only (Code/BatchRunner.h:411-417)
            p.eps_r_re = t0.mat.er, p.eps_r_im = t0.mat.ei, p.mu_r_re = t0.mat.mr, p.mu_r_im = t0.mat.mi;
        } else {
            p.eps_r_re = 1, p.eps_r_im = 0, p.mu_r_re = 1, p.mu_r_im = 0;
        }
        const int runs = b->mono ? 2 : 2 * (int)ids.size();
        p.max_rhs = std::max(p.max_rhs, std::min(runs, 256));
        mlfma_handle* h = nullptr;
        int rc = mlfma_create_nas(&p, t0.nas.c_str(), &h);
        if (rc != MLFMA_OK) {
            for (int i : ids) b->res[2 * i].status = b->res[2 * i + 1].status = rc;
            return rc;
        }
        long long dims[6];
        mlfma_dims(h, dims);
        const int n = (int)dims[1];
        mlfma_solve_params sp;
        if (sp_in) {
            sp = *sp_in;
        } else {   // the drivers' settings (Code/RCS.h:215-219): GMRES(50), 1e-4, 1e-8, n iterations
            sp.solver = b->solver;
            sp.itr_max = n;
            sp.restart = 50;
            sp.tol_rel = 1e-4;
            sp.tol_abs = 1e-8;
            sp.precond = 0;
        }
        if (sp.itr_max <= 0) sp.itr_max = n;
        const double pols[2] = {90.0, 0.0};   // "h" then "v" (Code/BatchRunner.h:403-408)
        if (b->mono) {
            for (int pl = 0; pl < 2; ++pl) {
                std::vector<double> out(3 * (size_t)npts);
                std::vector<mlfma_solve_stats> st(npts);
                mlfma_solve_params s2 = sp;
                s2.solver = MLFMA_SOLVER_GMRES;
                rc = mlfma_monostatic(h, &s2, pols[pl], npts, b->grid_th.data(), b->grid_ph.data(), out.data(), st.data());
                mlfma_solve_stats agg = st[0];
                for (const auto& s : st) {
                    agg.iterations = std::max(agg.iterations, s.iterations);
                    agg.rel_residual = std::max(agg.rel_residual, s.rel_residual);
                    agg.converged = agg.converged && s.converged;
                }
                for (int i : ids) {
                    RunResult& r = b->res[2 * i + pl];
                    r.status = rc;
                    r.rcs = out;
                    r.stats = agg;
                }
                if (rc != MLFMA_OK) worst = rc;
            }
        } else {
            // every (task, polarisation) of the group is one right-hand side
            std::vector<std::pair<int, int>> cols;   // (task, pol)
            for (int i : ids)
                for (int pl = 0; pl < 2; ++pl) cols.emplace_back(i, pl);
            const int R = (int)cols.size();
            for (int c0 = 0; c0 < R; c0 += 256) {
                const int m = std::min(256, R - c0);
                std::vector<double> pol(m), th(m), ph(m);
                for (int j = 0; j < m; ++j) {
                    const Task& t = b->tasks[cols[c0 + j].first];
                    pol[j] = pols[cols[c0 + j].second];
                    th[j] = t.th;
                    ph[j] = t.ph;
                }
                std::vector<double> B(2 * (size_t)n * m), X(2 * (size_t)n * m, 0.0);
                rc = mlfma_excitation_batch(h, m, pol.data(), th.data(), ph.data(), B.data());
                std::vector<mlfma_solve_stats> st(m);
                if (rc == MLFMA_OK) {
                    if (sp.solver == MLFMA_SOLVER_GMRES) {
                        rc = mlfma_solve_batch(h, &sp, n, m, B.data(), X.data(), st.data());
                    } else {   // CGS: one right-hand side at a time
                        for (int j = 0; j < m; ++j) {
                            const int r1 = mlfma_solve(h, &sp, n, B.data() + 2 * (size_t)n * j,
                                                       X.data() + 2 * (size_t)n * j, &st[j]);
                            if (r1 != MLFMA_OK && rc == MLFMA_OK) rc = r1;
                        }
                    }
                }
                for (int j = 0; j < m; ++j) {
                    RunResult& r = b->res[2 * cols[c0 + j].first + cols[c0 + j].second];
                    r.stats = st[j];
                    r.status = rc == MLFMA_OK || rc == MLFMA_ERR_NOT_CONVERGED
                                   ? (st[j].converged ? MLFMA_OK : MLFMA_ERR_NOT_CONVERGED)
                                   : rc;
                    if (rc == MLFMA_OK || rc == MLFMA_ERR_NOT_CONVERGED) {
                        r.rcs.assign(3 * (size_t)npts, 0.0);
                        const int r2 = mlfma_rcs(h, X.data() + 2 * (size_t)n * j, npts, b->grid_th.data(),
                                                 b->grid_ph.data(), pol[j], r.rcs.data());
                        if (r2 != MLFMA_OK) r.status = r2;
                    }
                    if (r.status != MLFMA_OK) worst = r.status;
                }
            }
        }
        mlfma_destroy(h);
        return worst;
    };
    int worst = MLFMA_OK;
    std::mutex mu;
    std::atomic<int> next{0};
    auto worker = [&] {
        for (int g = next++; g < (int)order.size(); g = next++) {
            const int rc = run_group(order[g]);
            if (rc != MLFMA_OK) {
                std::lock_guard<std::mutex> lock(mu);
                worst = rc;
            }
        }
    };
    const int nthreads = std::min(b->concurrency, (int)order.size());
    if (nthreads <= 1) {
        worker();
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < nthreads; ++t) pool.emplace_back(worker);
        for (auto& th : pool) th.join();
    }
    return worst;
}

int mlfma_batch_set_concurrency(mlfma_batch* b, int n) {
    if (!b || n < 1 || n > 16) return MLFMA_ERR_ARG;
    b->concurrency = n;
    return MLFMA_OK;
}

int mlfma_batch_result(const mlfma_batch* b, int task, int pol, double* rcs3, mlfma_solve_stats* stats) {
    if (!b || task < 0 || task >= (int)b->tasks.size() || pol < 0 || pol > 1) return MLFMA_ERR_ARG;
    const RunResult& r = b->res[2 * task + pol];
    if (r.status < 0) return MLFMA_ERR_STATE;
    if (rcs3 && !r.rcs.empty()) std::copy(r.rcs.begin(), r.rcs.end(), rcs3);
    if (stats) *stats = r.stats;
    return r.status;
}

}  // extern "C"
