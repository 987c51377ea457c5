// server.hpp -- CCS request server of the trajectory engine (host runtime).
// Included by engine.hip after the batch machinery it drives.
//
// Requests (atoms, run parameters, a target relative standard error) pass
//   submit -> setup worker (molecule setup, boundary marches on the GPU,
//   device upload) -> ready -> admitted (slot and device-memory budget) ->
//   active -> done.
// The scheduler thread runs slices: persistent launches over blocks of
// trajectories of every active request (a block is one CCS integral, Philox
// iteration = block index), by default one per launch shape with each
// workgroup staging the molecule of the request it works on.  Requests that
// arrive while a slice runs join the next slice.  As a slice completes, each
// block's CCS is computed on the device (k_block_ccs) and the stopping rule
// is applied on the device block by block in block order (k_stop); the host
// reads back only the blocks taken and the retirement flags.  A request's
// result depends only on its own blocks: it is the result cosims_run gives
// over the same blocks with the server's launch shape, whatever else shared
// the launches.
#pragma once

#include <condition_variable>
#include <deque>
#include <map>
#include <thread>

namespace {

constexpr int64_t kTrajBytes = 150;  // device bytes per in-flight trajectory (TrajBufs + work list)

// ------------------------------------------------ per-block CCS on the device
// One CCS integral of a served request (a block): its trajectory statuses in
// the slice's buffers, its sums from k_reduce, and where its CCS and standard
// error go (src/monteTraj.cpp:378-386).
struct BlockDesc {
    const int8_t *st;
    const double *sums;   // kSums
    double *out;          // omega, omega_err
    int32_t T, upd;       // trajectories; progress-update period (the denominator's reset)
};

// One workgroup a block (integral_ccs_wg).
__global__ void __launch_bounds__(256) k_block_ccs(const BlockDesc *D)
{
    const BlockDesc d = D[blockIdx.x];
    integral_ccs_wg(d.st, d.sums, d.T, d.upd, d.out, d.out + 1);
}

// Running state of a served request on the device: blocks taken into its
// result, whether it has met the stopping rule, and the inverse-variance
// sums of src/monteTraj.cpp:412-425 over its blocks in block order.
struct ReqDev {
    int32_t used, finished, min_blocks, max_blocks;
    double target, we, wm;
};

// The stopping rule of one request over the blocks of a collected slice
// (`upto`: one past its last block there; earlier blocks were collected
// before): blocks are taken in order from `used`, Sum 1/err^2 and
// Sum omega/err^2 accumulated as final_stats accumulates them, and the
// request stops at the first block from min_blocks on where werr / wmean <=
// target, or at max_blocks.  Each taken block's CCS, error and sums are
// copied to the slice's staging record for the host, after the integration
// steps of all the request's blocks in the slice.
struct StopDesc {
    ReqDev *state;
    const double *blk;    // omega, omega_err per block
    const double *sums;   // kSums per block
    int32_t it0, upto;    // the slice's blocks of this request: [it0, upto)
    double *stage;        // steps of the slice's blocks, then (2 + kSums) doubles per slice block
    int32_t *head;        // used before, used after, finished
};

__global__ void __launch_bounds__(64) k_stop(const StopDesc *D, int n)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const StopDesc d = D[e];
    double steps = 0;
    for (int k = d.it0; k < d.upto; k++) steps += d.sums[kSums * k + 3];
    d.stage[0] = steps;
    ReqDev s = *d.state;
    const int u0 = s.used;
    if (!s.finished && s.used != d.it0) {  // an earlier block of this request was never collected
        d.head[0] = u0;
        d.head[1] = u0;
        d.head[2] = -1;
        return;
    }
    for (int k = s.used; k < d.upto && !s.finished; k++) {
        const double om = d.blk[2 * k], er = d.blk[2 * k + 1];
        s.we += 1.0 / (er * er);
        s.wm += om / (er * er);
        s.used = k + 1;
        double *o = d.stage + 1 + (size_t)(2 + kSums) * (k - d.it0);
        o[0] = om;
        o[1] = er;
        for (int q = 0; q < kSums; q++) o[2 + q] = d.sums[kSums * k + q];
        if (s.used >= s.min_blocks) {
            const double werr = 1.0 / sqrt(s.we), wmean = s.wm / s.we;
            if (werr / wmean <= s.target || s.used >= s.max_blocks) s.finished = 1;
        }
    }
    *d.state = s;
    d.head[0] = u0;
    d.head[1] = s.used;
    d.head[2] = s.finished;
}

struct SJob {
    int64_t ticket = 0;
    cosims_params params{};
    std::vector<double> xyz, mass, sigma, eps, charge;
    bool has_charge = false;
    double sigma_max = 0;
    uint32_t job = 0;
    double target = 0.005;
    cosims_mol *mol = nullptr;          // owned, made by a setup worker
    int state = 0;                      // 0 queued, 1 setup, 2 ready, 3 active, 4 done, 5 failed
    int err = 0;
    std::string errmsg;
    int issued = 0;                     // blocks handed to slices
    int used = 0;                       // blocks taken into the result
    std::vector<double> omega, omega_err, sums;  // per block
    bool finished = false;
    DBuf dsums;                         // per-block sums on the device (max_blocks x kSums)
    DBuf dblk, dstate;                  // per-block CCS and error; the stopping rule's state (ReqDev)
    ReqDev init{};                      // host copy of the initial state (its upload is asynchronous)
    size_t bytes = 0;                   // charged against the budget from admission ...
    int inflight = 0;                   // ... until it has retired and no slice in flight holds it
    bool released = false;
    double t_submit = 0, t_ready = 0, t_admit = 0, t_done = 0, setup_ms = 0;
    double setup_phase_ms[4] = {0, 0, 0, 0};  // boundary marches, PCM clustering, multipoles, upload
    cosims_ccs ccs{};
    ~SJob()
    {
        if (mol) cosims_mol_destroy(mol);
    }
};

}  // namespace

struct cosims_server {
    cosims_server_cfg cfg{};
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv;       // scheduler and setup workers wait here
    std::condition_variable done_cv;  // waiters
    bool stopping = false;
    int64_t next_ticket = 1;
    int setups_running = 0;
    std::deque<std::shared_ptr<SJob>> queued, ready;
    std::vector<std::shared_ptr<SJob>> active;
    std::vector<std::shared_ptr<SJob>> retiring;  // retired, still held by a slice in flight
    std::map<int64_t, std::shared_ptr<SJob>> jobs;
    size_t bytes_in_use = 0, bytes_peak = 0, budget = 0;
    size_t slice_reserve = 0;         // trajectory buffers of pipeline_depth slices, outside the budget
    double units_per_ms = 1.0e9;      // force-loop work per kernel ms (learned from the slices)
    cosims_server_stats st{};
    std::vector<std::thread> workers;
    std::thread sched;
    double now_ms() const
    {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

void setup_worker(cosims_server *S)
{
    for (;;) {
        std::shared_ptr<SJob> J;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv.wait(lk, [&] { return S->stopping || !S->queued.empty(); });
            if (S->queued.empty()) return;
            J = S->queued.front();
            S->queued.pop_front();
            J->state = 1;
            S->setups_running++;
        }
        const double t = S->now_ms();
        cosims_atoms a{};
        a.n = (int32_t)J->mass.size();
        a.xyz = J->xyz.data(); a.mass = J->mass.data(); a.sigma = J->sigma.data(); a.epsilon = J->eps.data();
        a.charge = J->has_charge ? J->charge.data() : nullptr;
        a.sigma_max = J->sigma_max;
        cosims_mol *m = nullptr;
        int rc = cosims_mol_create(&J->params, &a, &m);
        const double t_up = S->now_ms();
        if (rc == 0) rc = ensure_device(m);
        const double upload_ms = S->now_ms() - t_up;
        std::string msg = rc ? g_err : std::string();
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->setups_running--;
            J->setup_ms = S->now_ms() - t;
            J->t_ready = S->now_ms();
            if (rc) {
                if (m) cosims_mol_destroy(m);
                J->state = 5;
                J->err = rc;
                J->errmsg = msg;
                J->t_done = J->t_ready;
                S->st.failed++;
                S->done_cv.notify_all();
            } else {
                J->mol = m;
                J->setup_phase_ms[0] = m->prep.boundary_ms;
                J->setup_phase_ms[1] = m->prep.cluster_ms;
                J->setup_phase_ms[2] = m->prep.moments_ms;
                J->setup_phase_ms[3] = upload_ms;
                // the molecule and its per-block sums; the slices' trajectory
                // buffers are reserved once, outside the budget
                J->bytes = m->dbytes + sizeof(double) * kSums * S->cfg.max_blocks;
                J->state = 2;
                S->ready.push_back(J);
            }
        }
        S->cv.notify_all();
    }
}

// Blocks job J still needs beyond those already issued: min_blocks (or the
// prior's estimate) first, then as many as reach the target if the standard
// error keeps falling as 1/sqrt(blocks).  0 while the issued blocks may
// suffice.
int blocks_wanted(const cosims_server *S, const SJob &J)
{
    const auto &c = S->cfg;
    int want;
    if (J.used < c.min_blocks) {
        int first = c.min_blocks;
        if (c.prior_rel_sd > 0) {
            const double r = c.prior_rel_sd / J.target;
            const double f = std::ceil(r * r / (double)c.block_traj - 1e-9);
            first = (int)std::min<double>(std::max<double>(first, f), c.max_blocks);
        }
        want = std::min(first, c.max_blocks) - J.issued;
    } else {
        cosims_ccs tmp{};
        final_stats(J.omega.data(), J.omega_err.data(), J.sums.data(), J.used, c.block_traj, &tmp);
        const double rel = tmp.werr / tmp.wmean;
        const double need = J.used * (rel / J.target) * (rel / J.target);
        want = std::max(J.issued == J.used ? 1 : 0, (int)std::ceil(need) - J.issued);
    }
    return std::max(0, std::min(want, c.max_blocks - J.issued));
}

// Work of one block of J in force-loop units (integration steps x atoms, or
// x clusters for the clustered types): the mean of its finished blocks, or
// before any, straight chords through the ellipsoid at the mean speed for
// 60 % of the trajectories.
double block_cost(const cosims_server *S, const SJob &J)
{
    const Prepared &p = J.mol->prep;
    const double units = (double)std::max(1, group_units(p));
    if (J.used > 0) {
        double steps = 0;
        for (int k = 0; k < J.used; k++) steps += J.sums[kSums * k + 3];
        return std::max(1.0, steps / J.used) * units;
    }
    const double chord = 2 * p.bmax / (p.vd.mean * J.mol->params.dt);
    return 0.6 * (double)S->cfg.block_traj * chord * units;
}

struct SliceEntry {
    std::shared_ptr<SJob> job;  // keeps the molecule alive while the slice is in flight
    int it0, nb;                // first block and block count of this slice
};

// Slice composition.  Shortest remaining work first: requests in order of
// blocks wanted x block cost take blocks until the slice's work budget
// (slice_ms at the measured throughput) is spent, then the slice is topped
// up to min_slice_traj trajectories in the same order; the first request
// always gets its block.  Equal share: every request gets the same number
// of blocks, up to slice_traj trajectories in all.  Requests already
// holding every block they may need get none.
std::vector<SliceEntry> plan_slice(cosims_server *S, const std::vector<std::shared_ptr<SJob>> &act)
{
    const auto &c = S->cfg;
    std::vector<SliceEntry> out;
    std::vector<int> want(act.size()), nb(act.size(), 0);
    std::vector<double> cost(act.size());
    std::vector<size_t> order;
    for (size_t k = 0; k < act.size(); k++) {
        want[k] = blocks_wanted(S, *act[k]);
        // once the stopping rule applies, issue_fraction of the estimate of
        // what is still needed (rounded up): 1 takes the estimate whole (the
        // fewest round trips per request), smaller values keep the blocks past
        // the stopping point fewer when the estimate is noisy
        if (act[k]->used >= c.min_blocks && want[k] > 0)
            want[k] = std::max(1, (int)std::ceil(want[k] * c.issue_fraction - 1e-9));
        cost[k] = block_cost(S, *act[k]);
        if (want[k] > 0) order.push_back(k);
    }
    if (order.empty()) return out;
    if (c.policy == COSIMS_SERVE_EQUAL_SHARE) {
        const int64_t share = std::max<int64_t>(1, c.slice_traj / ((int64_t)order.size() * c.block_traj));
        for (size_t k : order) nb[k] = (int)std::min<int64_t>(want[k], share);
    } else {
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t a, size_t b) { return want[a] * cost[a] < want[b] * cost[b]; });
        double budget = c.slice_ms * S->units_per_ms;
        int64_t traj_left = std::max<int64_t>(c.slice_traj, c.block_traj);
        bool first = true;
        for (size_t k : order) {
            int64_t b = std::min<int64_t>(want[k], (int64_t)(budget / cost[k]));
            b = std::min<int64_t>(b, traj_left / c.block_traj);
            if (b <= 0 && first) b = 1;
            first = false;
            if (b <= 0) continue;
            nb[k] = (int)b;
            budget -= b * cost[k];
            traj_left -= b * c.block_traj;
        }
        // fill: a slice with too few trajectories leaves its launches tail-bound
        int64_t fill = c.min_slice_traj;
        for (size_t k : order) fill -= (int64_t)nb[k] * c.block_traj;
        for (size_t k : order) {
            if (fill <= 0) break;
            const int64_t add = std::min<int64_t>(want[k] - nb[k], (fill + c.block_traj - 1) / c.block_traj);
            if (add <= 0) continue;
            nb[k] += (int)add;
            fill -= add * c.block_traj;
        }
    }
    for (size_t k : order)
        if (nb[k] > 0) out.push_back({act[k], act[k]->issued, nb[k]});
    return out;
}

// Per pipeline place: the descriptors of a slice's per-block CCS and
// stopping-rule kernels and the staging record the host reads back (into
// pinned memory: the read-back is queued with the slice, and a copy into
// pageable memory could hold the issuing thread until it lands).
struct SliceDev {
    DBuf bd, sd, stage, head;
    size_t bd_cap = 0, sd_cap = 0, stage_cap = 0, head_cap = 0;
    std::vector<BlockDesc> bdh;
    std::vector<StopDesc> sdh;
    double *stage_h = nullptr;
    int32_t *head_h = nullptr;
    size_t stage_h_cap = 0, head_h_cap = 0, nd = 0;
    hipEvent_t stop_ev = nullptr;  // recorded after this place's k_stop
    SliceDev() = default;
    SliceDev(const SliceDev &) = delete;
    SliceDev &operator=(const SliceDev &) = delete;
    ~SliceDev()
    {
        if (stage_h) (void)hipHostFree(stage_h);
        if (head_h) (void)hipHostFree(head_h);
        if (stop_ev) (void)hipEventDestroy(stop_ev);
    }
};

// Pinned host capacity for n elements; grown rarely (at least 4096
// elements, doubling), since freeing pinned memory waits for the device.
template <class T>
int reserve_pinned(T *&p, size_t &cap, size_t n)
{
    if (cap >= n) return 0;
    size_t c = std::max<size_t>(4096, cap);
    while (c < n) c *= 2;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    HIP_CHECK(hipHostMalloc((void **)&p, sizeof(T) * c, hipHostMallocDefault));
    cap = c;
    return 0;
}

// A slice in flight: its plan, device buffers and stream.
struct Slice {
    std::vector<SliceEntry> e;
    BatchPlan *P = nullptr;
    SliceDev *D = nullptr;
    hipStream_t s = nullptr;
    int launches = 0;
    double t_issue = 0;
    bool live() const { return !e.empty(); }
};

template <class T>
int reserve_dev(DBuf &b, size_t &cap, size_t n)
{
    if (cap >= n) return 0;
    HIP_CHECK(b.alloc(sizeof(T) * n));
    cap = n;
    return 0;
}

// Issues one slice (sampling, launches, per-block reduction into each
// request's device sums) without waiting for the launches.
int issue_slice(cosims_server *S, Slice &X, hipEvent_t &last_stop)
{
    const auto &c = S->cfg;
    BatchPlan &P = *X.P;
    P.jobs.clear();
    P.n = 0;
    for (SliceEntry &e : X.e) {
        SJob &J = *e.job;
        if (e.it0 + e.nb > c.max_blocks) return fail(COSIMS_E_INVALID, "server: block count beyond max_blocks");
        if (!J.dsums.p) {
            if (J.dsums.alloc_cached(sizeof(double) * kSums * c.max_blocks) != hipSuccess ||
                J.dblk.alloc_cached(sizeof(double) * 2 * c.max_blocks) != hipSuccess ||
                J.dstate.alloc_cached(sizeof(ReqDev)) != hipSuccess)
                return fail(COSIMS_E_NOMEM, "server: per-request device state");
            J.init = ReqDev{0, 0, c.min_blocks, c.max_blocks, J.target, 0.0, 0.0};
            HIP_CHECK(hipMemcpyAsync(J.dstate.p, &J.init, sizeof(ReqDev), hipMemcpyHostToDevice, X.s));
        }
        BatchPlan::Job b{};
        b.mol = J.mol;
        b.seed = J.params.seed;
        b.job = J.job;
        b.iters = e.nb;
        b.T = c.block_traj;
        b.n = (int64_t)b.iters * b.T;
        b.off = P.n;
        b.ft = J.mol->prep.force_type;
        b.units = group_units(J.mol->prep);
        b.it0 = (uint32_t)e.it0;
        const int k = job_class(J.mol->prep);
        b.L = c.lanes > 0 ? c.lanes : job_lanes(J.mol->prep, c.class_lanes);
        b.W = std::max(c.lanes <= 0 && c.class_wg[k] > 0 ? c.class_wg[k] : c.wg_size, b.L);
        P.n += b.n;
        P.jobs.push_back(b);
        J.issued += e.nb;
        J.inflight++;
    }
    if (P.n > INT32_MAX / 4) return fail(COSIMS_E_INVALID, "server: slice too large");
    P.shared = c.own_launches == 0;
    if (int rc = order_plan(P)) return rc;
    if (int rc = stage_batch(P, X.s)) return rc;
    if (int rc = launch_enqueue(P, c.prec, 0, X.s, X.launches)) return rc;
    for (size_t i = 0; i < X.e.size(); i++) {
        const auto &J = P.jobs[i];
        ReduceArgs RA{};
        RA.status = P.B.status.as<int8_t>() + J.off; RA.chi = P.B.chi.as<double>() + J.off;
        RA.cchi = P.B.exitv2.as<double>() + J.off;
        RA.ergy = P.B.ergy.as<double>() + J.off; RA.totsteps = P.B.tot.as<int32_t>() + J.off;
        RA.per_iter = J.T; RA.ccs_const = J.mol->prep.ccs_const;
        RA.out = X.e[i].job->dsums.as<double>() + kSums * X.e[i].it0;
        hipLaunchKernelGGL(k_reduce, dim3(J.iters), dim3(256), 0, X.s, RA);
        HIP_CHECK(hipGetLastError());
    }
    // each block's CCS and standard error on the device (the host reads no statuses)
    SliceDev &D = *X.D;
    D.bdh.clear();
    for (size_t i = 0; i < X.e.size(); i++) {
        const auto &J = P.jobs[i];
        SJob &R = *X.e[i].job;
        const int64_t pr = J.mol->params.print_rate;
        const int32_t upd = (int32_t)(pr > 0 ? pr : std::max<int64_t>(1, (int64_t)(J.T / 5.0)));
        for (int b = 0; b < J.iters; b++)
            D.bdh.push_back({P.B.status.as<int8_t>() + J.off + (int64_t)b * J.T,
                             R.dsums.as<double>() + kSums * (X.e[i].it0 + b), R.dblk.as<double>() + 2 * (X.e[i].it0 + b),
                             (int32_t)J.T, upd});
    }
    if (int rc = reserve_dev<BlockDesc>(D.bd, D.bd_cap, D.bdh.size())) return rc;
    HIP_CHECK(hipMemcpyAsync(D.bd.p, D.bdh.data(), sizeof(BlockDesc) * D.bdh.size(), hipMemcpyHostToDevice, X.s));
    hipLaunchKernelGGL(k_block_ccs, dim3((unsigned)D.bdh.size()), dim3(256), 0, X.s, D.bd.as<BlockDesc>());
    HIP_CHECK(hipGetLastError());
    // The stopping rule (k_stop: every request of the slice, its blocks in
    // block order; blocks of finished requests are ignored) and the
    // read-back of its staging record follow in the slice's stream, so they
    // start as the slice's workgroups drain; launched when the host collects
    // the slice, they waited ~40 ms for a CU slot behind the next slices'
    // persistent workgroups.  Each slice's k_stop runs after the previous
    // slice's (a request's blocks are taken in block order).
    const size_t ne = X.e.size();
    D.sdh.resize(ne);
    size_t nd = 0;  // staging doubles
    for (size_t i = 0; i < ne; i++) nd += 1 + (size_t)(2 + kSums) * X.e[i].nb;
    nd = std::max<size_t>(1, nd);
    if (int rc = reserve_dev<double>(D.stage, D.stage_cap, nd)) return rc;
    if (int rc = reserve_dev<int32_t>(D.head, D.head_cap, 3 * std::max<size_t>(1, ne))) return rc;
    if (int rc = reserve_dev<StopDesc>(D.sd, D.sd_cap, std::max<size_t>(1, ne))) return rc;
    if (int rc = reserve_pinned<double>(D.stage_h, D.stage_h_cap, nd)) return rc;
    if (int rc = reserve_pinned<int32_t>(D.head_h, D.head_h_cap, 3 * std::max<size_t>(1, ne))) return rc;
    D.nd = nd;
    size_t off = 0;
    for (size_t i = 0; i < ne; i++) {
        SJob &J = *X.e[i].job;
        D.sdh[i] = StopDesc{J.dstate.as<ReqDev>(), J.dblk.as<double>(), J.dsums.as<double>(), X.e[i].it0,
                            X.e[i].it0 + X.e[i].nb, D.stage.as<double>() + off, D.head.as<int32_t>() + 3 * i};
        off += 1 + (size_t)(2 + kSums) * X.e[i].nb;
    }
    HIP_CHECK(hipMemcpyAsync(D.sd.p, D.sdh.data(), sizeof(StopDesc) * ne, hipMemcpyHostToDevice, X.s));
    if (last_stop) HIP_CHECK(hipStreamWaitEvent(X.s, last_stop, 0));
    hipLaunchKernelGGL(k_stop, dim3((unsigned)((ne + 63) / 64)), dim3(64), 0, X.s, D.sd.as<StopDesc>(), (int)ne);
    HIP_CHECK(hipGetLastError());
    if (!D.stop_ev) HIP_CHECK(hipEventCreateWithFlags(&D.stop_ev, hipEventDisableTiming));
    HIP_CHECK(hipEventRecord(D.stop_ev, X.s));
    last_stop = D.stop_ev;
    HIP_CHECK(hipMemcpyAsync(D.stage_h, D.stage.p, sizeof(double) * nd, hipMemcpyDeviceToHost, X.s));
    HIP_CHECK(hipMemcpyAsync(D.head_h, D.head.p, sizeof(int32_t) * 3 * ne, hipMemcpyDeviceToHost, X.s));
    return 0;
}

// Waits for a slice (its launches, the per-block CCS, the stopping rule and
// the read-back of the staging record, all queued at issue) and takes the
// blocks each request took, their CCS, errors and sums, and whether it has
// stopped.
int collect_slice(cosims_server *S, Slice &X)
{
    const auto &c = S->cfg;
    BatchPlan &P = *X.P;
    SliceDev &D = *X.D;
    GroupRun R;
    if (int rc = launch_finish(P, X.s, R)) return rc;  // waits for the stream
    const size_t ne = X.e.size();
    size_t off = 0;
    double units = 0;  // force-loop work of the slice (integration steps x atoms, or x clusters)
    for (size_t i = 0; i < ne; i++) {
        SJob &J = *X.e[i].job;
        const auto &B = P.jobs[i];
        const int32_t *h = D.head_h + 3 * i;
        const double *rec = D.stage_h + off;
        units += rec[0] * B.units;
        if (h[2] < 0 || h[0] != J.used) return fail(COSIMS_E_INVALID, "server: blocks out of order");
        for (int k = h[0]; k < h[1]; k++) {
            const double *o = rec + 1 + (2 + kSums) * (size_t)(k - X.e[i].it0);
            J.omega.push_back(o[0]);
            J.omega_err.push_back(o[1]);
            J.sums.insert(J.sums.end(), o + 2, o + 2 + kSums);
        }
        J.used = h[1];
        if (h[2] == 1 && !J.finished) {
            J.finished = true;
            cosims_ccs tmp{};
            final_stats(J.omega.data(), J.omega_err.data(), J.sums.data(), J.used, c.block_traj, &tmp);
            J.ccs = tmp;
            J.ccs.lanes = B.L;
            J.ccs.wg_size = B.W;
        }
        off += 1 + (size_t)(2 + kSums) * X.e[i].nb;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    if (R.ms > 0 && units > 0) S->units_per_ms = 0.5 * S->units_per_ms + 0.5 * units / R.ms;
    S->st.slices++;
    S->st.launches += X.launches;
    S->st.traj += P.n;
    S->st.kernel_ms += R.ms;
    S->st.pair_evals += R.pairs + R.cl[3] + R.cl[4] + R.cl[5];
    S->st.ref_flops += R.flops;
    S->st.force_evals += R.evals;
    S->st.total_steps += R.steps;
    return 0;
}

// Gives a retired request's device bytes back to the budget once no slice in
// flight holds it any more (its molecule and sums stay alive until then).
// Caller holds S->mu.
void release_retired(cosims_server *S)
{
    std::vector<std::shared_ptr<SJob>> keep;
    for (auto &J : S->retiring) {
        if (J->inflight > 0) {
            keep.push_back(J);
        } else if (!J->released) {
            S->bytes_in_use -= J->bytes;
            J->released = true;
        }
    }
    S->retiring.swap(keep);
}

// The scheduler keeps up to pipeline_depth slices in flight: whenever a place is
// free and some request wants blocks, the next slice (planned from the
// results known so far and the blocks already issued) is sampled and
// launched on its own streams, so one slice's tail overlaps the next ones.
// Slices are collected in issue order (a request's blocks are taken in
// block order), polling the oldest so that new slices can be issued while
// it drains.
constexpr int kMaxDepth = 8;

void scheduler(cosims_server *S)
{
    const int kDepth = S->cfg.pipeline_depth;
    BatchPlan plans[kMaxDepth];
    SliceDev devs[kMaxDepth];
    Stream streams[kMaxDepth];
    bool ok = true;
    for (auto &x : streams) ok &= hipStreamCreateWithFlags(&x.s, hipStreamNonBlocking) == hipSuccess;
    Slice fly[kMaxDepth];
    hipEvent_t last_stop = nullptr;  // after the k_stop of the last slice issued
    int head = 0, inflight = 0;      // fly[head] is the oldest slice
    double last_end = 0;         // busy time: union of the slices' issue-to-collect spans
    for (;;) {
        std::vector<std::shared_ptr<SJob>> act;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            for (;;) {
                // admission: FIFO, while a slot and the device-memory budget allow
                while (!S->ready.empty() && (int)S->active.size() < S->cfg.max_active &&
                       (S->active.empty() || S->bytes_in_use + S->ready.front()->bytes <= S->budget)) {
                    auto J = S->ready.front();
                    S->ready.pop_front();
                    J->state = 3;
                    J->t_admit = S->now_ms();
                    S->bytes_in_use += J->bytes;
                    S->bytes_peak = std::max(S->bytes_peak, S->bytes_in_use);
                    S->active.push_back(J);
                }
                if (!S->active.empty() || inflight > 0) break;
                if (S->stopping && S->queued.empty() && S->ready.empty() && S->setups_running == 0) return;
                S->cv.wait(lk);
            }
            act = S->active;
        }
        int rc = ok ? 0 : fail(COSIMS_E_HIP, "server: no stream");
        // fill the free places of the pipeline
        while (!rc && inflight < kDepth) {
            const int q = (head + inflight) % kDepth;
            Slice &X = fly[q];
            X.e = plan_slice(S, act);
            if (!X.live()) break;
            X.P = &plans[q];
            X.D = &devs[q];
            X.s = streams[q].s;
            X.t_issue = S->now_ms();
            rc = issue_slice(S, X, last_stop);
            inflight++;
        }
        Slice &X = fly[head];
        bool collected = false;
        if (!rc && inflight > 0) {
            const hipError_t q = hipStreamQuery(X.s);
            if (q == hipSuccess || inflight == kDepth) {
                rc = collect_slice(S, X);
                collected = true;
            } else if (q == hipErrorNotReady) {
                // the oldest slice still runs and a place is free: wait a little
                // for it, or for a new request to plan into the free place
                std::unique_lock<std::mutex> lk(S->mu);
                S->cv.wait_for(lk, std::chrono::milliseconds(1));
                continue;
            } else {
                rc = fail(COSIMS_E_HIP, std::string("server: ") + hipGetErrorString(q));
            }
        }
        const std::string msg = rc ? g_err : std::string();
        if (rc) {  // drain what is in flight before failing the requests
            for (auto &x : streams) (void)hipStreamSynchronize(x.s);
        }
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (collected) {
                const double now = S->now_ms();
                S->st.busy_ms += now - std::max(X.t_issue, last_end);
                last_end = now;
                for (auto &e : X.e) e.job->inflight--;
            }
            if (rc)  // every stream is drained: no slice holds a request any more
                for (auto &f : fly)
                    for (auto &e : f.e) e.job->inflight = 0;
            std::vector<std::shared_ptr<SJob>> keep;
            for (auto &J : S->active) {
                if (rc || J->finished) {
                    J->state = rc ? 5 : 4;
                    J->err = rc;
                    J->errmsg = msg;
                    J->t_done = S->now_ms();
                    S->retiring.push_back(J);
                    if (rc) S->st.failed++;
                    else S->st.done++;
                } else {
                    keep.push_back(J);
                }
            }
            S->active.swap(keep);
            release_retired(S);
        }
        S->done_cv.notify_all();
        if (rc) {
            for (auto &f : fly) f = Slice();
            inflight = 0;
            continue;
        }
        if (collected) {
            X = Slice();  // releases its requests (molecules freed once no slice holds them)
            head = (head + 1) % kDepth;
            inflight--;
        }
    }
}

}  // namespace

extern "C" {

int cosims_server_create(const cosims_server_cfg *cfg, cosims_server **out)
{
    if (!out) return fail(COSIMS_E_INVALID, "cosims_server_create: null output");
    cosims_server_cfg c{};
    if (cfg) c = *cfg;
    if (c.prec != COSIMS_PREC_FP64 && c.prec != COSIMS_PREC_MIXED) return fail(COSIMS_E_INVALID, "unknown precision");
    if (c.wg_size == 0) c.wg_size = 512;
    if (c.lanes != 0 && c.lanes != 4 && c.lanes != 8 && c.lanes != 16 && c.lanes != 32 && c.lanes != 64)
        return fail(COSIMS_E_INVALID, "server lanes must be 0 (size rule), 4, 8, 16, 32 or 64");
    if (c.wg_size % 64 || c.wg_size > 1024)
        return fail(COSIMS_E_INVALID, "server wg_size must be a multiple of 64, at most 1024");
    for (int k = 0; k < COSIMS_SIZE_CLASSES; k++) {
        const int L = c.class_lanes[k], W = c.class_wg[k];
        if (L != 0 && L != 4 && L != 8 && L != 16 && L != 32 && L != 64)
            return fail(COSIMS_E_INVALID, "server class_lanes must be 0, 4, 8, 16, 32 or 64");
        if (W < 0 || W % 64 || W > 1024)
            return fail(COSIMS_E_INVALID, "server class_wg must be 0 or a multiple of 64, at most 1024");
    }
    if (c.min_slice_traj <= 0) c.min_slice_traj = 65536;
    if (c.pipeline_depth <= 0) c.pipeline_depth = 3;
    if (c.pipeline_depth > kMaxDepth) return fail(COSIMS_E_INVALID, "pipeline_depth at most 8");
    if (c.max_active <= 0) c.max_active = 64;
    if (c.block_traj <= 0) c.block_traj = 10000;
    if (c.slice_traj <= 0) c.slice_traj = 1 << 20;
    if (c.slice_ms <= 0) c.slice_ms = 150;
    if (c.policy != COSIMS_SERVE_SHORTEST_FIRST && c.policy != COSIMS_SERVE_EQUAL_SHARE)
        return fail(COSIMS_E_INVALID, "unknown server policy");
    if (c.min_blocks <= 0) c.min_blocks = 3;
    if (c.max_blocks <= 0) c.max_blocks = 1000;
    if (c.min_blocks > c.max_blocks) return fail(COSIMS_E_INVALID, "min_blocks > max_blocks");
    if (c.setup_threads <= 0) c.setup_threads = 8;
    if (c.issue_fraction == 0) c.issue_fraction = 1.0;
    if (!(c.prior_rel_sd >= 0 && c.prior_rel_sd < 1e3)) return fail(COSIMS_E_INVALID, "prior_rel_sd must be in [0, 1000)");
    if (!(c.issue_fraction > 0 && c.issue_fraction <= 1)) return fail(COSIMS_E_INVALID, "issue_fraction must be in (0, 1]");
    Device d;
    if (int rc = device_init(d)) return rc;
    std::unique_ptr<cosims_server> S(new cosims_server());
    S->cfg = c;
    S->t0 = std::chrono::steady_clock::now();
    if (c.hbm_budget > 0) {
        S->budget = (size_t)c.hbm_budget;
    } else {
        size_t fr = 0, tot = 0;
        HIP_CHECK(hipMemGetInfo(&fr, &tot));
        S->budget = fr / 5 * 4;
    }
    // the trajectory buffers of pipeline_depth slices (a slice holds at most
    // slice_traj + min_slice_traj trajectories, and one block more) come out
    // of the budget before any request is admitted
    S->slice_reserve = (size_t)c.pipeline_depth * (size_t)(c.slice_traj + c.min_slice_traj + c.block_traj) *
                       (size_t)kTrajBytes;
    if (S->slice_reserve >= S->budget)
        return fail(COSIMS_E_NOMEM, "server: hbm_budget does not cover the slices' trajectory buffers (" +
                                        std::to_string(S->slice_reserve) + " bytes)");
    S->budget -= S->slice_reserve;
    cosims_server *raw = S.get();
    for (int k = 0; k < c.setup_threads; k++) S->workers.emplace_back(setup_worker, raw);
    S->sched = std::thread(scheduler, raw);
    *out = S.release();
    return 0;
}

int cosims_server_submit(cosims_server *S, const cosims_request *r, int64_t *ticket)
{
    if (!S || !r || !ticket) return fail(COSIMS_E_INVALID, "cosims_server_submit: null argument");
    const cosims_atoms &a = r->atoms;
    if (a.n <= 0 || !a.xyz || !a.mass || !a.sigma || !a.epsilon)
        return fail(COSIMS_E_INVALID, "cosims_server_submit: empty molecule");
    auto J = std::make_shared<SJob>();
    J->params = r->params;
    J->xyz.assign(a.xyz, a.xyz + 3 * (size_t)a.n);
    J->mass.assign(a.mass, a.mass + a.n);
    J->sigma.assign(a.sigma, a.sigma + a.n);
    J->eps.assign(a.epsilon, a.epsilon + a.n);
    J->has_charge = a.charge != nullptr;
    if (a.charge) J->charge.assign(a.charge, a.charge + a.n);
    J->sigma_max = a.sigma_max;
    J->job = r->job;
    J->target = r->target_rel_err > 0 ? r->target_rel_err : 0.005;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stopping) return fail(COSIMS_E_INVALID, "cosims_server_submit: server is shutting down");
        J->ticket = S->next_ticket++;
        J->t_submit = S->now_ms();
        S->jobs[J->ticket] = J;
        S->queued.push_back(J);
        S->st.submitted++;
        *ticket = J->ticket;
    }
    S->cv.notify_all();
    return 0;
}

int cosims_server_wait(cosims_server *S, int64_t ticket, double timeout_ms, cosims_result *out, double *omega,
                       double *omega_err, int32_t cap)
{
    if (!S || !out) return fail(COSIMS_E_INVALID, "cosims_server_wait: null argument");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->jobs.find(ticket);
    if (it == S->jobs.end()) return fail(COSIMS_E_INVALID, "cosims_server_wait: unknown ticket");
    std::shared_ptr<SJob> J = it->second;
    auto doneq = [&] { return J->state >= 4; };
    if (timeout_ms < 0) {
        S->done_cv.wait(lk, doneq);
    } else if (!S->done_cv.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), doneq)) {
        std::memset(out, 0, sizeof(*out));
        out->status = 1;
        return fail(COSIMS_E_PENDING, "request not finished");
    }
    std::memset(out, 0, sizeof(*out));
    out->status = J->state == 4 ? 0 : J->err;
    out->blocks = J->used;
    out->ccs = J->ccs;
    out->rel_err = J->used > 0 && J->ccs.wmean != 0 ? J->ccs.werr / J->ccs.wmean : 0;
    out->t_submit_ms = J->t_submit;
    out->t_ready_ms = J->t_ready;
    out->t_admit_ms = J->t_admit;
    out->t_done_ms = J->t_done;
    out->setup_ms = J->setup_ms;
    out->block_traj = S->cfg.block_traj;
    out->setup_boundary_ms = J->setup_phase_ms[0];
    out->setup_cluster_ms = J->setup_phase_ms[1];
    out->setup_moments_ms = J->setup_phase_ms[2];
    out->setup_upload_ms = J->setup_phase_ms[3];
    for (int k = 0; k < std::min<int>(cap, J->used); k++) {
        if (omega) omega[k] = J->omega[k];
        if (omega_err) omega_err[k] = J->omega_err[k];
    }
    if (J->state == 5) return fail(J->err, J->errmsg);
    return 0;
}

int cosims_server_stats_get(cosims_server *S, cosims_server_stats *out)
{
    if (!S || !out) return fail(COSIMS_E_INVALID, "cosims_server_stats_get: null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->st;
    out->queued = (int64_t)(S->queued.size() + S->ready.size()) + S->setups_running;
    out->active = (int64_t)S->active.size();
    out->bytes_in_use = (int64_t)S->bytes_in_use;
    out->bytes_peak = (int64_t)S->bytes_peak;
    out->budget = (int64_t)S->budget;
    out->now_ms = S->now_ms();
    return 0;
}

void cosims_server_destroy(cosims_server *S)
{
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stopping = true;
    }
    S->cv.notify_all();
    for (auto &w : S->workers) w.join();
    S->cv.notify_all();
    S->sched.join();
    delete S;
}

}  // extern "C"
