// Ensembles of independent cases on one GPU (no counterpart in the reference, which runs
// one case per process): a 2-D grid of 96^2..512^2 cells leaves most of the MI355X idle,
// so many cases advance at once. Every case keeps its own solver, hierarchy and HIP
// stream; host worker threads each own a list of cases and drive them, so that the
// kernels of different cases overlap on the GPU while one case's host waits (the solver's
// read-backs, the AMG set-up) proceed. Cases go largest-first to the least loaded worker
// (longest-processing-time scheduling by cell count x steps). Results of a case do not
// depend on the schedule: each solver is deterministic and shares nothing with the others.
#include <algorithm>
#include <chrono>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cuibm_amd.h"

namespace {

struct Job {
    int idx;
    cuibm_solver *s = nullptr;
    long long cells = 0;
    int steps = 0;
};

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// runs fn(job) for every job of every worker, one host thread per worker; returns the
// first failure (code and message) in case order
template <class F>
int run_workers(std::vector<std::vector<Job *>> &lists, F fn, std::string &msg) {
    std::mutex mu;
    int rc = 0, rc_idx = -1;
    std::vector<std::thread> th;
    for (auto &list : lists)
        th.emplace_back([&, list]() {
            for (Job *j : list) {
                const int r = fn(*j);
                if (r != 0) {
                    std::lock_guard<std::mutex> g(mu);
                    if (rc_idx < 0 || j->idx < rc_idx) {
                        rc = r;
                        rc_idx = j->idx;
                        msg = "case " + std::to_string(j->idx) + ": " + cuibm_last_error();
                    }
                    return;  // a failed case ends its worker's list
                }
            }
        });
    for (auto &t : th) t.join();
    return rc;
}

}  // namespace

extern "C" int cuibm_ensemble_run(int n, const char *const *case_folders, int argc, const char *const *argv,
                                  int steps, int workers, long long *cells, double *cd, double *cl, double *wall_s,
                                  double *steps_wall_s, char *err, int err_len) {
    std::string msg;
    auto fail = [&](int rc, const std::string &m) {
        if (err && err_len > 0) {
            const size_t k = std::min((size_t)err_len - 1, m.size());
            std::copy(m.begin(), m.begin() + k, err);
            err[k] = '\0';
        }
        return rc;
    };
    if (n <= 0 || !case_folders) return fail(CUIBM_ERR_ARGUMENT, "no cases");
    if (workers < 1 || workers > 64) return fail(CUIBM_ERR_ARGUMENT, "workers must be 1..64");
    std::vector<Job> jobs(n);
    const double t0 = now_s();
    // create on this thread (parsing only), then size each case for the schedule
    for (int i = 0; i < n; i++) {
        jobs[i].idx = i;
        int rc = cuibm_create(case_folders[i], argc, argv, &jobs[i].s);
        double sc[8] = {0};
        if (rc == 0) {
            rc = cuibm_case_scalars(jobs[i].s, sc, 8);  // nx, ny, nu, dt, nt, ...; count on success
            if (rc > 0) rc = 0;
        }
        if (rc != 0) {
            const std::string m = "case " + std::to_string(i) + ": " + cuibm_last_error();
            for (auto &j : jobs) cuibm_destroy(j.s);
            return fail(rc, m);
        }
        jobs[i].cells = (long long)sc[0] * (long long)sc[1];
        jobs[i].steps = steps > 0 ? steps : (int)sc[4];
    }
    // longest processing time first onto the least loaded worker
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return jobs[a].cells * jobs[a].steps > jobs[b].cells * jobs[b].steps;
    });
    const int nw = std::min(workers, n);
    std::vector<std::vector<Job *>> lists(nw);
    std::vector<long long> load(nw, 0);
    for (int i : order) {
        const int w = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        lists[w].push_back(&jobs[i]);
        load[w] += jobs[i].cells * jobs[i].steps;
    }
    int rc = run_workers(lists, [](Job &j) { return cuibm_initialise(j.s); }, msg);
    double t1 = now_s();
    if (rc == 0) {
        rc = run_workers(lists, [](Job &j) {
            for (int k = 0; k < j.steps; k++) {
                const int r = cuibm_step_time(j.s);
                if (r != 0) return r;
            }
            return 0;
        }, msg);
    }
    if (rc == 0) rc = cuibm_device_synchronize();
    const double t2 = now_s();
    if (rc == 0)
        for (auto &j : jobs) {
            double info[8];
            if (cuibm_info(j.s, info, 8) != 0) continue;
            if (cells) cells[j.idx] = j.cells;
            if (cd) cd[j.idx] = 2.0 * info[5];
            if (cl) cl[j.idx] = 2.0 * info[6];
        }
    for (auto &j : jobs) cuibm_destroy(j.s);
    if (wall_s) *wall_s = t2 - t0;
    if (steps_wall_s) *steps_wall_s = t2 - t1;
    return rc == 0 ? 0 : fail(rc, msg);
}
