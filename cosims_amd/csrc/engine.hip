// engine.hip -- MI355X (gfx950) He trajectory engine: kernels + C ABI.
//
// Hot path (one GPU):
//   k_sample  Philox4x32-10 rejection sampler of monteTraj::setUpTraj
//             (src/monteTraj.cpp:174-278); one thread per trajectory; hits
//             are compacted into a work list.
//   k_traj    persistent velocity-Verlet kernel of diffEqSolver::verletVel
//             (src/diffEqSolver.cpp:81-211): one trajectory per lane, lanes
//             refilled from the work list with one atomic per wavefront;
//             every iteration evaluates forceAccel::getReducedVec at each
//             lane's position.  A lane that reaches the ellipsoid writes an
//             exit record and takes new work.
//   k_check   energy-conservation test of an exit (getPotential at the start
//             and exit points, FP64): accept, or queue a retry with dt/(k+1)
//             for the next k_traj round (at most three rounds).
//   k_reduce  per-CCS-integral sums of x = const1 const2 (1 - cos chi)
//             (src/monteTraj.cpp:456-477), FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "cosims_engine.h"
#include "kernels.hpp"
#include "prep.hpp"
#include "traj_group.hpp"

using namespace cosims;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIP_CHECK(expr)                                                                            \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(COSIMS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));          \
    } while (0)

enum : int8_t { ST_MISS = 0, ST_OK = 1, ST_FAIL = 2, ST_PENDING = 3, ST_RETRY = 4 };

constexpr double kPiD = 3.14159265358979323846;
constexpr int32_t kMaxStepsPerAttempt = 200000;
constexpr double kPi2D = 1.57079632679489661923;

// ------------------------------------------------------------- Philox ----
__device__ __forceinline__ void philox4x32(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
}

__device__ __forceinline__ double u01(uint32_t hi, uint32_t lo)
{
    uint64_t x = ((uint64_t)hi << 32) | lo;
    return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

// Work-list order of a persistent launch: longest expected trajectories
// first, so that the launch drains on short ones.  The expected length is
// the straight-line crossing time of the ellipsoid in steps of dt0, which
// ranks the integration steps of a trajectory with Spearman 0.93-0.98 on
// 300- to 20,000-atom ions; hits are bucketed on a log scale (16 buckets an
// octave) and the buckets laid out longest first.  The order of the work
// list changes which group runs a trajectory, never its result.
constexpr int kOrderBins = 256;
constexpr bool kOrderWork = true;

__device__ __forceinline__ uint8_t order_bin(double steps)
{
    if (!(steps > 1.0)) return 0;
    return (uint8_t)min((int)(16.0 * log2(steps)), kOrderBins - 1);
}

struct SampleArgs {
    uint32_t key0, key1, job, iteration0;
    int64_t first, n, per_iter;     // per_iter > 0: iteration = j / per_iter
    double gmin, gmax, max_val, vd_a, vd_norm, bmax, maxdist, ea, eb, ec;
    int projection;
    double *p0, *v0, *g;
    int8_t *status;                 // ST_MISS or ST_PENDING(=hit, to run)
    int32_t *hit;                   // optional
    int32_t *items;                 // optional compacted hit list
    int32_t *n_items;
    int32_t item_base;              // added to the indices written to items
    uint8_t *bin;                   // optional: order bucket of each hit (order_bin)
    double dt0;
};

// setUpTraj: rejection sampling of (g, phi, theta, gamma, b) and projection
// of the start point onto the ellipsoid.
__global__ void __launch_bounds__(256) k_sample(SampleArgs A)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool hit = false;
    if (j < A.n) {
        int64_t tj = A.first + j;
        uint32_t it = A.iteration0, traj = (uint32_t)tj;
        if (A.per_iter > 0) {
            it = A.iteration0 + (uint32_t)(tj / A.per_iter);
            traj = (uint32_t)(tj % A.per_iter);
        }
        double g = 0, phi = 0, theta = 0, gamma = 0, b = 0;
        const double bMax = A.bmax;
        bool acc = false;
        for (uint32_t t = 0; t < 4096 && !acc; t++) {
            uint32_t w[16];
            for (int blk = 0; blk < 4; blk++) {
                uint32_t c[4] = {traj, it, A.job, 4 * t + blk};
                philox4x32(c, A.key0, A.key1);
                w[4 * blk] = c[0]; w[4 * blk + 1] = c[1]; w[4 * blk + 2] = c[2]; w[4 * blk + 3] = c[3];
            }
            double u0 = u01(w[0], w[1]), u1 = u01(w[2], w[3]), u2 = u01(w[4], w[5]), u3 = u01(w[6], w[7]);
            double u4 = u01(w[8], w[9]), u5 = u01(w[10], w[11]), u6 = u01(w[12], w[13]), u7 = u01(w[14], w[15]);
            g = u0 * (A.gmax - A.gmin) + A.gmin;
            double gCheck = u1 * A.max_val;
            phi = u2 * 2.0 * kPiD;
            theta = u3 * kPiD;
            double thetaCheck = 0.5 * u4;
            gamma = u5 * 2.0 * kPiD;
            b = u6 * bMax;
            double bCheck = u7 * (2 / bMax);
            if (thetaCheck <= sin(theta) / 2 && bCheck <= 2 * b / (bMax * bMax) &&
                gCheck <= pow(g, 5.0) * exp(-A.vd_a * g * g) * A.vd_norm)
                acc = true;
        }
        const double md = A.maxdist;
        double P0 = b * cos(gamma) * cos(phi) * cos(theta) - b * sin(gamma) * sin(phi) + md * cos(phi) * sin(theta);
        double P1 = b * cos(gamma) * sin(phi) * cos(theta) + b * sin(gamma) * cos(phi) + md * sin(phi) * sin(theta);
        double P2 = md * cos(theta) - b * cos(gamma) * sin(theta);
        double V0 = -g * cos(phi) * sin(theta), V1 = -g * sin(phi) * sin(theta), V2 = -g * cos(theta);
        hit = acc;
        if (acc && A.projection) {
            double v0 = V0 / A.ea, v1 = V1 / A.eb, v2 = V2 / A.ec;
            double p0 = P0 / A.ea, p1 = P1 / A.eb, p2 = P2 / A.ec;
            double c0 = (p1 * v2) - (p2 * v1), c1 = (p2 * v0) - (p0 * v2), c2 = (p0 * v1) - (p1 * v0);
            double vv = (v0 * v0) + (v1 * v1) + (v2 * v2);
            double disc = -((c0 * c0) + (c1 * c1) + (c2 * c2)) + vv;
            if (disc < 0) {
                hit = false;
            } else {
                double pv = (p0 * v0) + (p1 * v1) + (p2 * v2);
                double pp = -(pv / vv) - sqrt(disc) / vv;
                P0 = P0 + V0 * pp; P1 = P1 + V1 * pp; P2 = P2 + V2 * pp;
            }
        }
        A.p0[3 * j] = P0; A.p0[3 * j + 1] = P1; A.p0[3 * j + 2] = P2;
        A.v0[3 * j] = V0; A.v0[3 * j + 1] = V1; A.v0[3 * j + 2] = V2;
        A.g[j] = g;
        if (A.bin && hit) {
            // crossing time of the straight line through the ellipsoid (scaled coordinates)
            const double q0 = P0 / A.ea, q1 = P1 / A.eb, q2 = P2 / A.ec;
            const double w0 = V0 / A.ea, w1 = V1 / A.eb, w2 = V2 / A.ec;
            const double qw = q0 * w0 + q1 * w1 + q2 * w2, ww = w0 * w0 + w1 * w1 + w2 * w2;
            const double qq = q0 * q0 + q1 * q1 + q2 * q2;
            const double disc = qw * qw - ww * (qq - 1.0);
            A.bin[j] = order_bin(disc > 0 ? 2.0 * sqrt(disc) / ww / A.dt0 : 0.0);
        }
        if (A.status) A.status[j] = hit ? ST_PENDING : ST_MISS;
        if (A.hit) A.hit[j] = hit ? 1 : 0;
    }
    if (A.items) {  // wave-aggregated compaction of hits
        const bool h = (j < A.n) && hit;
        const unsigned long long mask = __ballot(h);
        if (mask) {
            const int lane = threadIdx.x & 63;
            const int leader = __ffsll((unsigned long long)mask) - 1;
            int base = 0;
            if (lane == leader) base = atomicAdd(A.n_items, __popcll(mask));
            base = __shfl(base, leader);
            if (h) A.items[base + __popcll(mask & ((1ull << lane) - 1))] = A.item_base + (int32_t)j;
        }
    }
}

// ------------------------------------------------------ trajectory ---
struct TrajArgs {
    MolDev m;
    const int32_t *items;
    int32_t n_items;
    int32_t *counter;
    const double *p0, *v0;
    int32_t *tries, *maxtries;
    int8_t *status;
    double *chi;
    int32_t *steps, *totsteps;
    double *exit_pos, *exit_v2;
    double dt0;
    int32_t max_steps;          // per-attempt step cap (the reference has none)
    unsigned long long *evals;  // [0] integration steps, [1] force evaluations of this launch
};

template <int PREC, int FT>
__global__ void __launch_bounds__(256) k_traj(TrajArgs A)
{
    const MolDev &m = A.m;
    const int lane = threadIdx.x & 63;
    int t = 0, state = 0, tries = 0, maxtries = 8, orbits = 0, steps = 0, tot = 0;
    double px = 0, py = 0, pz = 0, vx = 0, vy = 0, vz = 0, ax = 0, ay = 0, az = 0, dt = 0, rdot = 0;
    bool exhausted = false;  // wave-uniform
    unsigned n_steps = 0, n_evals = 0;
    for (;;) {
        // refill idle lanes: one atomic per wavefront
        if (!exhausted) {
            const bool need = (state == 0);
            const unsigned long long mask = __ballot(need);
            if (mask) {
                const int cnt = __popcll(mask);
                const int leader = __ffsll((unsigned long long)mask) - 1;
                int base = 0;
                if (lane == leader) base = atomicAdd(A.counter, cnt);
                base = __shfl(base, leader);
                if (base + cnt >= A.n_items) exhausted = true;
                if (need) {
                    const int idx = base + __popcll(mask & ((1ull << lane) - 1));
                    if (idx < A.n_items) {
                        t = A.items[idx];
                        tries = A.tries[t];
                        maxtries = A.maxtries[t];
                        dt = A.dt0 / (double)(tries + 1);
                        px = A.p0[3 * t]; py = A.p0[3 * t + 1]; pz = A.p0[3 * t + 2];
                        tot = 0;
                        state = 1;
                    }
                }
            }
        }
        if (!__any(state != 0)) break;

        double ex, ey, ez, hx = 0, hy = 0, hz = 0;
        if (state == 2) {  // src/diffEqSolver.cpp:131-132
            ex = px + vx * dt + ax * dt * dt * 0.5;
            ey = py + vy * dt + ay * dt * dt * 0.5;
            ez = pz + vz * dt + az * dt * dt * 0.5;
            hx = vx + ax * dt * 0.5;
            hy = vy + ay * dt * 0.5;
            hz = vz + az * dt * 0.5;
        } else if (state == 1) {
            ex = px; ey = py; ez = pz;
        } else {
            ex = ey = ez = 1.0e4;  // idle lane: any finite far point
        }
        double fx, fy, fz;
        accel<PREC, FT>(m, ex, ey, ez, fx, fy, fz);
        n_evals += (state != 0);
        n_steps += (state == 2);

        if (state == 1) {  // start of an attempt, src/diffEqSolver.cpp:112-127
            ax = fx; ay = fy; az = fz;
            vx = A.v0[3 * t]; vy = A.v0[3 * t + 1]; vz = A.v0[3 * t + 2];
            orbits = 0;
            steps = 0;
            rdot = px * vx + py * vy + pz * vz;
            state = 2;
        } else if (state == 2) {  // src/diffEqSolver.cpp:133-205
            px = ex; py = ey; pz = ez;
            ax = fx; ay = fy; az = fz;
            vx = hx + ax * dt * 0.5;
            vy = hy + ay * dt * 0.5;
            vz = hz + az * dt * 0.5;
            steps++;
            tot++;
            const double rold = rdot;
            rdot = px * vx + py * vy + pz * vz;
            const double prod = rdot * rold;
            bool track = true;
            if (steps > 2 && prod < 0 && !(rdot > rold)) {
                orbits++;
                if (orbits > 8) track = false;
            }
            // A non-finite position or a runaway attempt can never reach the
            // exit test; it ends as a failed trajectory (status 2).
            const bool runaway = !(fabs(px) + fabs(py) + fabs(pz) < 1.0e30) || steps > A.max_steps;
            if (runaway) {
                A.status[t] = ST_FAIL;
                A.chi[t] = kPi2D;
                A.steps[t] = steps;
                A.totsteps[t] += tot;
                A.tries[t] = tries;
                state = 0;
            } else if (!track) {
                tries++;
                dt = A.dt0 / (double)(tries + 1);
                if (!(tries < maxtries && maxtries < 10)) {  // loop exit: failed trajectory
                    A.status[t] = ST_FAIL;
                    A.chi[t] = kPi2D;
                    A.steps[t] = steps;
                    A.totsteps[t] += tot;
                    A.tries[t] = tries;
                    state = 0;
                }
            } else {
                if (outside(m, px, py, pz) && steps > 10) {
                    // exit: scattering angle (vector3D::Angle, src/vector3D.h:134-143)
                    const double w0 = A.v0[3 * t], w1 = A.v0[3 * t + 1], w2 = A.v0[3 * t + 2];
                    const double t1 = w0 * vx + w1 * vy + w2 * vz;
                    const double t2 = sqrt(w0 * w0 + w1 * w1 + w2 * w2) * sqrt(vx * vx + vy * vy + vz * vz);
                    const double t3 = t1 / t2;
                    A.chi[t] = (t3 > 1) ? 0.0 : ((t3 < -1) ? kPiD : acos(t1 / t2));
                    A.exit_pos[3 * t] = px; A.exit_pos[3 * t + 1] = py; A.exit_pos[3 * t + 2] = pz;
                    A.exit_v2[t] = vx * vx + vy * vy + vz * vz;
                    A.steps[t] = steps;
                    A.totsteps[t] += tot;
                    A.tries[t] = tries;
                    A.maxtries[t] = maxtries;
                    A.status[t] = ST_PENDING;
                    state = 0;
                }
            }
        }
    }
    // per-launch counters: wave sum, one atomic per wavefront
    unsigned long long s = n_steps, e = n_evals;
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off);
        e += __shfl_xor(e, off);
    }
    if (lane == 0) {
        atomicAdd(A.evals, s);
        atomicAdd(A.evals + 1, e);
    }
}

// ------------------------------------------------------ energy check ---
struct CheckArgs {
    MolDev m;
    const int32_t *items;
    int32_t n_items;
    int first_round;
    const double *p0, *v0, *exit_pos, *exit_v2;
    double *e0, *chi, *ergy;
    int8_t *status;
    int32_t *tries, *maxtries, *steps;
    int32_t *next_items, *next_count;
};

// src/diffEqSolver.cpp:108, 178-203
__global__ void __launch_bounds__(256) k_check(CheckArgs A)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool retry = false;
    int t = 0;
    if (i < A.n_items) {
        t = A.items[i];
        if (A.status[t] == ST_PENDING) {
            const MolDev &m = A.m;
            const double hm = 0.5 * m.mu;
            double e0;
            if (A.first_round) {
                const double *p = A.p0 + 3 * t, *v = A.v0 + 3 * t;
                e0 = potential(m, p[0], p[1], p[2]) + hm * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                A.e0[t] = e0;
            } else {
                e0 = A.e0[t];
            }
            const double *q = A.exit_pos + 3 * t;
            const double e1 = potential(m, q[0], q[1], q[2]) + hm * A.exit_v2[t];
            const double pct = 100 * fabs(e0 - e1) / e0;
            if (pct > 0.5) {
                const int tr = A.tries[t] + 1, mt = A.maxtries[t] + 1;
                A.tries[t] = tr;
                A.maxtries[t] = mt;
                if (tr < mt && mt < 10) {
                    A.status[t] = ST_RETRY;
                    retry = true;
                } else {
                    A.status[t] = ST_FAIL;
                    A.chi[t] = kPi2D;
                    A.steps[t] = 0;
                }
            } else {
                A.status[t] = ST_OK;
                A.ergy[t] = pct;
            }
        }
    }
    const unsigned long long mask = __ballot(retry);
    if (mask) {
        const int lane = threadIdx.x & 63;
        const int leader = __ffsll((unsigned long long)mask) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(A.next_count, __popcll(mask));
        base = __shfl(base, leader);
        if (retry) A.next_items[base + __popcll(mask & ((1ull << lane) - 1))] = t;
    }
}

// ------------------------------------------------------------ reduce ---
struct ReduceArgs {
    const int8_t *status;
    const double *chi, *ergy;
    const double *cchi;  // when set: accepted trajectories' cos(chi) (chi taken here, as k_chi does)
    const int32_t *totsteps;
    int64_t per_iter;
    double ccs_const;
    double *out;  // per iteration: total, total2, max_ergy, steps, hits, fails
};

// runThreads x = const1*const2*(1 - cos|chi|) (src/monteTraj.cpp:471-475);
// one workgroup per CCS integral, fixed-order tree.
__global__ void __launch_bounds__(256) k_reduce(ReduceArgs A)
{
    const int it = blockIdx.x;
    const int64_t base = (int64_t)it * A.per_iter;
    double s = 0, s2 = 0, mx = 0, st = 0, nh = 0, nf = 0;
    for (int64_t j = threadIdx.x; j < A.per_iter; j += blockDim.x) {
        const int8_t stt = A.status[base + j];
        double c = (stt == ST_MISS) ? 0.0 : A.chi[base + j];
        if (A.cchi && stt == ST_OK) {
            const double t3 = A.cchi[base + j];
            c = (t3 > 1) ? 0.0 : ((t3 < -1) ? 3.14159265358979323846 : acos(t3));
        }
        const double x = A.ccs_const * (1 - cos(fabs(c)));
        s += x;
        s2 += x * x;
        if (stt == ST_OK) mx = fmax(mx, A.ergy[base + j]);
        if (stt != ST_MISS) {
            nh += 1;
            st += A.totsteps[base + j];
        }
        if (stt == ST_FAIL) nf += 1;
    }
    __shared__ double red[6][256];
    red[0][threadIdx.x] = s; red[1][threadIdx.x] = s2; red[2][threadIdx.x] = mx;
    red[3][threadIdx.x] = st; red[4][threadIdx.x] = nh; red[5][threadIdx.x] = nf;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int k = 0; k < 6; k++) {
                if (k == 2) red[k][threadIdx.x] = fmax(red[k][threadIdx.x], red[k][threadIdx.x + w]);
                else red[k][threadIdx.x] += red[k][threadIdx.x + w];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) A.out[6 * it + threadIdx.x] = red[threadIdx.x][0];
}

// --------------------------------------------------- batch evaluators ---
template <int PREC, int FT>
__global__ void __launch_bounds__(256) k_accel(MolDev m, int64_t n, const double *pos, double *acc)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double x = 1.0e4, y = 1.0e4, z = 1.0e4;
    if (i < n) { x = pos[3 * i]; y = pos[3 * i + 1]; z = pos[3 * i + 2]; }
    double ox, oy, oz;
    accel<PREC, FT>(m, x, y, z, ox, oy, oz);
    if (i < n) { acc[3 * i] = ox; acc[3 * i + 1] = oy; acc[3 * i + 2] = oz; }
}

__global__ void __launch_bounds__(256) k_potential(MolDev m, int64_t n, const double *pos, double *u)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[i] = potential(m, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
}

// FP32 FMA issue rate of the device as this build compiles it (packed FP32
// off): 16 independent v_fma_f32 chains per lane, every result stored.  The
// multiplier and addend are per-lane values (VGPRs): on gfx950 an FP32 VALU
// instruction whose operands are all VGPRs issues in 2 cycles a wavefront,
// one with a scalar (SGPR) operand in 4, so a probe with uniform a, b would
// measure half the rate the force loops can reach.
__global__ void __launch_bounds__(256) k_fma_rate(float *out, int iters, float a, float b)
{
    float v[16];
#pragma unroll
    for (int i = 0; i < 16; i++) v[i] = threadIdx.x * 1e-3f + i;
    const float av = a + threadIdx.x * 1e-9f, bv = b + threadIdx.x * 1e-9f;
    for (int k = 0; k < iters; k++) {
#pragma unroll
        for (int i = 0; i < 16; i++) v[i] = __builtin_fmaf(v[i], av, bv);
    }
    float s = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) s += v[i];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// ------------------------------------------------- kernel selection ---
using TrajFn = void (*)(TrajArgs);
using AccFn = void (*)(MolDev, int64_t, const double *, double *);

template <int P>
TrajFn traj_fn(int ft)
{
    switch (ft) {
    case 1: return k_traj<P, 1>;
    case 2: return k_traj<P, 2>;
    case 3: return k_traj<P, 3>;
    case 4: return k_traj<P, 4>;
    case 5: return k_traj<P, 5>;
    default: return k_traj<P, 6>;
    }
}
TrajFn traj_kernel(int prec, int ft) { return prec == COSIMS_PREC_FP64 ? traj_fn<0>(ft) : traj_fn<1>(ft); }

template <int P>
AccFn acc_fn(int ft)
{
    switch (ft) {
    case 1: return k_accel<P, 1>;
    case 2: return k_accel<P, 2>;
    case 3: return k_accel<P, 3>;
    case 4: return k_accel<P, 4>;
    case 5: return k_accel<P, 5>;
    default: return k_accel<P, 6>;
    }
}
AccFn accel_kernel(int prec, int ft) { return prec == COSIMS_PREC_FP64 ? acc_fn<0>(ft) : acc_fn<1>(ft); }

// ------------------------------------------------------ device state ---
struct Device {
    int id = -1, cus = 0;
    size_t lds_max = 65536;  // largest dynamic LDS one workgroup may opt into
    char name[256] = {0};
};

int device_init(Device &d)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(COSIMS_E_NODEVICE, "no HIP device");
    HIP_CHECK(hipGetDevice(&d.id));
    hipDeviceProp_t prop;
    HIP_CHECK(hipGetDeviceProperties(&prop, d.id));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(COSIMS_E_NODEVICE, std::string("engine is built for gfx950, device is ") + prop.gcnArchName);
    d.cus = prop.multiProcessorCount;
    int optin = 0;
    if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, d.id) == hipSuccess && optin > 0)
        d.lds_max = (size_t)optin;
    else
        d.lds_max = prop.sharedMemPerBlock;
    std::snprintf(d.name, sizeof(d.name), "%s", prop.name);
    return 0;
}

// Device blocks of the short-lived buffers of molecule setup (boundary
// marches) and of served requests: taken with hipMalloc (sizes rounded up to
// a power of two) and handed back to this cache instead of hipFree, which
// waits for the whole device (a setup or a retiring request beside a
// persistent serving slice would otherwise hold its thread until the slice
// ends).  Blocks are reused only after their owner has seen its last use
// complete (a host synchronisation), so any stream may use them next; they
// stay allocated until the process ends.
struct DevCache {
    std::mutex mu;
    std::multimap<std::pair<int, size_t>, void *> blocks;  // (device, bytes) -> free block
};
DevCache &dev_cache()
{
    static DevCache *c = new DevCache();  // never destroyed: blocks outlive static destructors
    return *c;
}
hipError_t dev_cache_get(size_t bytes, int *dev, size_t *cap, void **p)
{
    size_t sz = 256;
    while (sz < bytes) sz <<= 1;
    if (hipGetDevice(dev) != hipSuccess) *dev = 0;
    DevCache &c = dev_cache();
    {
        std::lock_guard<std::mutex> g(c.mu);
        auto it = c.blocks.find({*dev, sz});
        if (it != c.blocks.end()) {
            *p = it->second;
            c.blocks.erase(it);
            *cap = sz;
            return hipSuccess;
        }
    }
    const hipError_t e = hipMalloc(p, sz);
    if (e == hipSuccess) *cap = sz;
    return e;
}
void dev_cache_put(int dev, size_t cap, void *p)
{
    DevCache &c = dev_cache();
    std::lock_guard<std::mutex> g(c.mu);
    c.blocks.insert({{dev, cap}, p});
}

// RAII device buffer; alloc() replaces any previous allocation.  borrow()
// points it at memory it does not own (a slice of the static scratch).
// alloc_cached() takes a block of the device cache (DevCache) and gives it
// back there; the owner releases it only after its last use has completed.
struct DBuf {
    void *p = nullptr;
    bool owned = false, cached = false;
    int dev = 0;
    size_t cap = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { reset(); }
    void reset()
    {
        if (p && owned) {
            if (cached) dev_cache_put(dev, cap, p);
            else (void)hipFree(p);
        }
        p = nullptr;
        owned = false;
        cached = false;
    }
    hipError_t alloc(size_t bytes)
    {
        reset();
        const hipError_t e = hipMalloc(&p, bytes ? bytes : 16);
        owned = e == hipSuccess;
        if (!owned) p = nullptr;
        return e;
    }
    hipError_t alloc_cached(size_t bytes)
    {
        reset();
        const hipError_t e = dev_cache_get(bytes ? bytes : 16, &dev, &cap, &p);
        owned = cached = e == hipSuccess;
        if (!owned) p = nullptr;
        return e;
    }
    void borrow(void *q)
    {
        reset();
        p = q;
    }
    template <typename T> T *as() const { return (T *)p; }
};

// Points b at the 256-byte aligned slice of `bytes` at or after offset *off
// of [base, base + cap) and advances *off past it; false when it does not fit.
bool carve_one(DBuf &b, uint8_t *base, size_t cap, size_t *off, size_t bytes)
{
    const size_t o = (*off + 255) / 256 * 256;
    *off = o + bytes;
    if (o + bytes > cap) return false;
    b.borrow(base + o);
    return true;
}

struct Stream {
    hipStream_t s = nullptr;
    ~Stream() { if (s) (void)hipStreamDestroy(s); }
};

}  // namespace

// ------------------------------------------------- boundary marches ---
// One wavefront per march (impactParam::setBoundry, src/impactParam.cpp:
// 602-736): 64 consecutive points per pass, lane j rebuilding p_{k+j} by the
// same sequential FP64 additions as the host loop and evaluating getPotential
// there in reference atom order (scalar atom loads shared by the wave); the
// first lane that crosses (or passes the axis) ends the march.
__device__ void march_wave(const MolDev &m, const March &w, double mk, int64_t max_points, MarchHit &out)
{
    const int lane = threadIdx.x;
    const double e0 = potential(m, w.start[0], w.start[1], w.start[2]);
    const double sl = (e0 + mk < 0) ? -1.0 : 1.0;
    const double vx = w.dir[0] * sl, vy = w.dir[1] * sl, vz = w.dir[2] * sl;
    const double sx = w.start[0], sy = w.start[1];
    double bx = w.start[0], by = w.start[1], bz = w.start[2];  // point kb
    double qx = w.prev0[0], qy = w.prev0[1], qz = w.prev0[2];  // point kb - 1
    int64_t kb = 0;
    if (w.first) {
        bx = bx + vx; by = by + vy; bz = bz + vz;
        kb = 1;
    }
    for (; kb < max_points; kb += 64) {
        double px = bx, py = by, pz = bz;
        for (int q = 0; q < lane; q++) {
            px = px + vx; py = py + vy; pz = pz + vz;
        }
        const double en = potential(m, px, py, pz);
        const bool inr = kb + lane < max_points;
        const bool cross = inr && (en + mk) * sl < 0;
        const bool axis = inr && w.axis && ((sx == 0 && sy / py <= 0) || (sy == 0 && sx / px <= 0) ||
                                            (sx / px <= 0 && sy / py <= 0));
        const int src = (lane + 63) & 63;
        const double nx = __shfl(px, src), ny = __shfl(py, src), nz = __shfl(pz, src);
        const unsigned long long mask = __ballot(cross || axis);
        if (mask) {
            const int k = __ffsll(mask) - 1;
            if (lane == k) {
                out.found = cross ? 1 : 0;
                out.p[0] = px; out.p[1] = py; out.p[2] = pz;
                out.prev[0] = k ? nx : qx; out.prev[1] = k ? ny : qy; out.prev[2] = k ? nz : qz;
            }
            return;
        }
        qx = __shfl(px, 63); qy = __shfl(py, 63); qz = __shfl(pz, 63);
        bx = qx + vx; by = qy + vy; bz = qz + vz;
    }
    if (lane == 0) out.found = -1;
}

struct MarchArgs {
    MolDev m;          // atoms (a64, reference order), n, charged, dipole_scalar
    const March *w;
    double mk;
    int64_t max_points;
    MarchHit *out;
};

__global__ void __launch_bounds__(64) k_march(MarchArgs A)
{
    march_wave(A.m, A.w[blockIdx.x], A.mk, A.max_points, A.out[blockIdx.x]);
}

// The whole boundary search as one chain of launches: the z-axis walk, the
// ray geometry it implies, and every ray (blocks past the ray count exit).
constexpr int kChainMaxRays = 1024 * kBoundaryRot;
struct ChainState {
    MarchHit z;
    RayGeometry g;
    int32_t count;     // rays of the geometry; > kChainMaxRays: not run here
};
struct ChainArgs {
    MolDev m;
    BoundaryPlan bp;
    double mk;
    int64_t max_points;    // rays
    int64_t zwalk_points;  // the z-axis walk
    ChainState *st;
    MarchHit *hits;
};

__global__ void __launch_bounds__(64) k_chain_zwalk(ChainArgs A)
{
    // z axis: pos.z -= 0.5 * slope, i.e. the step (0, 0, -0.5) * slope
    March w;
    w.start[0] = 0; w.start[1] = 0; w.start[2] = A.bp.z0;
    w.dir[0] = 0; w.dir[1] = 0; w.dir[2] = -0.5;
    w.prev0[0] = 0; w.prev0[1] = 0; w.prev0[2] = A.bp.z0;
    w.first = 1;
    w.axis = 0;
    march_wave(A.m, w, A.mk, A.zwalk_points, A.st->z);
}

// ray_geometry (host) in the same IEEE operations, from the z-axis walk's
// crossing; count = 0 when the walk found none
__device__ RayGeometry chain_geometry(const ChainArgs &A, const MarchHit &z, int &count)
{
    RayGeometry g{};
    count = 0;
    if (z.found != 1) return g;
    double zMax = (z.p[2] < z.prev[2]) ? z.prev[2] : z.p[2];
    zMax *= 1.3;
    g.zmax = zMax;
    g.posmag = A.bp.posmag0 + fabs(zMax - A.bp.maxz);
    const int cz = (int)ceil(zMax);
    g.levels = (int)ceil((double)(25 < cz ? cz : 25));
    g.zinc = 2 * zMax / (g.levels);
    count = (g.levels + 1) * kBoundaryRot;
    return g;
}

// Every block derives the ray geometry from the walk's crossing itself (a
// few scalar operations; no launch of its own, which under a serving load
// would wait for a CU slot like every launch); block 0 publishes it.
__global__ void __launch_bounds__(64) k_chain_rays(ChainArgs A)
{
    ChainState &S = *A.st;
    const MarchHit z = S.z;
    int count;
    const RayGeometry g = chain_geometry(A, z, count);
    const int r = blockIdx.x;
    if (r == 0 && threadIdx.x == 0) {
        S.g = g;
        S.count = count;
    }
    if (count > kChainMaxRays || r >= count) return;
    const int i = r / kBoundaryRot, n = r % kBoundaryRot;
    // ray_march (host) in the same IEEE operations
    March w;
    w.start[0] = g.posmag * A.bp.cos_t[n];
    w.start[1] = g.posmag * A.bp.sin_t[n];
    w.start[2] = -g.zmax + i * g.zinc;
    w.dir[0] = A.bp.dir_x[n]; w.dir[1] = A.bp.dir_y[n]; w.dir[2] = 0.0;
    w.prev0[0] = 0; w.prev0[1] = 0; w.prev0[2] = 0;
    w.first = 0;
    w.axis = 1;
    march_wave(A.m, w, A.mk, A.max_points, A.hits[r]);
}

// Boundary marches on the GPU for prepare(): the molecule's atoms are
// uploaded once, then the chain of launches runs with one wait at its end.
// Its buffers come from the device cache (DBuf::alloc_cached): a setup
// running beside a serving slice does not wait for the slice at every free.
struct GpuMarcher : BoundaryMarcher {
    // One stream for the whole search: the atoms' upload, the marches and
    // the read-backs are ordered on it, so the march kernels see the copy
    // through the stream's own ordering (buffers come from the device cache
    // and may have held another molecule's data, read on another stream).
    Stream st;
    DBuf atoms_buf;
    void *atoms = nullptr;
    int boundary(const Prepared &p, double mk, const BoundaryPlan &bp, MarchHit &zhit, std::vector<MarchHit> &rays,
                 std::string &err) override
    {
        int rc = boundary_(p, mk, bp, zhit, rays);
        if (rc) err = g_err;
        return rc;
    }
    MolDev mol_dev(const Prepared &p) const
    {
        MolDev m{};
        m.n = p.n;
        m.charged = p.charged ? 1 : 0;
        m.dipole_scalar = p.dipole_scalar;
        m.a64 = (const AtomF64 *)atoms;
        return m;
    }
    int stream()
    {
        if (!st.s) HIP_CHECK(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
        return 0;
    }
    int upload_atoms(const Prepared &p)
    {
        if (atoms) return 0;
        Device d;
        if (int rc = device_init(d)) return rc;
        if (int rc = stream()) return rc;
        std::vector<AtomF64> a(p.n);
        for (int i = 0; i < p.n; i++) {
            a[i].x = p.xyz[3 * i]; a[i].y = p.xyz[3 * i + 1]; a[i].z = p.xyz[3 * i + 2];
            a[i].q = p.q[i]; a[i].s6 = p.sig6[i]; a[i].e24 = p.eps24m[i]; a[i].e4 = p.eps4[i]; a[i].pad = 0;
        }
        HIP_CHECK(atoms_buf.alloc_cached(sizeof(AtomF64) * p.n));
        atoms = atoms_buf.p;
        HIP_CHECK(hipMemcpyAsync(atoms, a.data(), sizeof(AtomF64) * p.n, hipMemcpyHostToDevice, st.s));
        HIP_CHECK(hipStreamSynchronize(st.s));  // a[] is freed on return
        return 0;
    }
    int boundary_(const Prepared &p, double mk, const BoundaryPlan &bp, MarchHit &zhit, std::vector<MarchHit> &rays)
    {
        if (int rc = upload_atoms(p)) return rc;
        // the chain state and the hits in one buffer, read back by one copy
        // (each launch and copy behind a serving load waits for a CU slot)
        constexpr size_t kStateBytes = (sizeof(ChainState) + 255) / 256 * 256;
        DBuf buf;
        HIP_CHECK(buf.alloc_cached(kStateBytes + sizeof(MarchHit) * kChainMaxRays));
        ChainArgs A{};
        A.m = mol_dev(p);
        A.bp = bp;
        A.mk = mk;
        A.max_points = kMaxMarchPoints;
        A.zwalk_points = zwalk_points(bp.z0);
        A.st = buf.as<ChainState>();
        A.hits = (MarchHit *)(buf.as<char>() + kStateBytes);
        hipLaunchKernelGGL(k_chain_zwalk, dim3(1), dim3(64), 0, st.s, A);
        hipLaunchKernelGGL(k_chain_rays, dim3(kChainMaxRays), dim3(64), 0, st.s, A);
        HIP_CHECK(hipGetLastError());
        std::vector<char> host(kStateBytes + sizeof(MarchHit) * kChainMaxRays);
        HIP_CHECK(hipMemcpyAsync(host.data(), buf.p, host.size(), hipMemcpyDeviceToHost, st.s));
        HIP_CHECK(hipStreamSynchronize(st.s));
        ChainState S;
        std::memcpy(&S, host.data(), sizeof(S));
        rays.resize(kChainMaxRays);
        std::memcpy(rays.data(), host.data() + kStateBytes, sizeof(MarchHit) * kChainMaxRays);
        zhit = S.z;
        if (S.z.found != 1) {
            rays.clear();
            return 0;
        }
        if (!(S.g.levels >= 25 && S.g.levels < (1 << 20) && S.count == (S.g.levels + 1) * kBoundaryRot))
            return fail(COSIMS_E_SETUP, "boundary search: inconsistent ray geometry");
        if (S.count <= kChainMaxRays) {
            rays.resize(S.count);
            return 0;
        }
        // more rays than the chain runs: the same marches from a host-built list
        std::vector<March> w((size_t)S.count);
        for (int i = 0; i <= S.g.levels; i++)
            for (int n = 0; n < kBoundaryRot; n++) w[(size_t)i * kBoundaryRot + n] = ray_march(bp, S.g, i, n);
        return run_list(p, mk, w, rays);
    }
    int run_list(const Prepared &p, double mk, const std::vector<March> &w, std::vector<MarchHit> &out)
    {
        DBuf dw, dh;
        HIP_CHECK(dw.alloc_cached(sizeof(March) * w.size()));
        HIP_CHECK(dh.alloc_cached(sizeof(MarchHit) * w.size()));
        HIP_CHECK(hipMemcpyAsync(dw.p, w.data(), sizeof(March) * w.size(), hipMemcpyHostToDevice, st.s));
        MarchArgs A{};
        A.m = mol_dev(p);
        A.w = dw.as<March>();
        A.mk = mk;
        A.max_points = kMaxMarchPoints;
        A.out = dh.as<MarchHit>();
        hipLaunchKernelGGL(k_march, dim3((unsigned)w.size()), dim3(64), 0, st.s, A);
        HIP_CHECK(hipGetLastError());
        out.resize(w.size());
        HIP_CHECK(hipMemcpyAsync(out.data(), dh.p, sizeof(MarchHit) * w.size(), hipMemcpyDeviceToHost, st.s));
        HIP_CHECK(hipStreamSynchronize(st.s));
        return 0;
    }
};

int cosims::set_last_error(int code, const std::string &msg) { return fail(code, msg); }

// -------------------------------------------------------------- handle ---
struct cosims_mol {
    Params params;
    Prepared prep;
    double setup_ms = 0;
    // device side, created on first GPU use (std::call_once)
    mutable std::once_flag up_once;
    mutable int up_rc = 0;
    mutable std::string up_err;
    mutable Device dev;
    mutable MolDev md{};
    mutable void *dmem = nullptr;
    mutable const double *soa = nullptr;  // 7 FP64 arrays of npad: x y z q s6 e24 e4
    mutable int npad = 0;
    mutable size_t dbytes = 0;            // device bytes of the molecule
    mutable size_t dcap = 0;              // ... of its block of the device cache
    mutable int ddev = 0;
    mutable MolRec rec{};                 // table record of this molecule (device pointers)
    mutable const MolRec *rec_dev = nullptr;  // the same record in device memory
    // autotuned launch shapes, keyed by precision and run-size bucket:
    // lanes per trajectory, workgroup size
    mutable std::mutex tune_mu;
    mutable std::map<int, std::pair<int, int>> tuned;
    mutable std::atomic<bool> on_device{false};  // device copy made (ensure_device succeeded)
    // a block of the device cache (filled by a synchronous copy): destroying
    // a molecule does not wait for launches of other molecules, as hipFree
    // would; its launches have completed by then
    ~cosims_mol() { if (dmem) dev_cache_put(ddev, dcap, dmem); }
};

namespace {

Params to_params(const cosims_params *p)
{
    Params q;
    q.temp = p->temp; q.dt = p->dt; q.traj = p->traj; q.iter = p->iter; q.charge = p->charge != 0;
    q.seed = p->seed; q.multipole = p->multipole != 0; q.dispersion_cutoff = p->dispersion_cutoff != 0;
    q.dispersion_radius = p->dispersion_radius; q.multipole_radius = p->multipole_radius;
    q.multipole_order = p->multipole_order; q.max_cluster_size = p->max_cluster_size;
    q.projection = p->projection != 0; q.print_rate = p->print_rate; q.total_charge = p->total_charge;
    q.threads = p->threads;
    return q;
}

template <typename T>
void fill_atom(T &a, double x, double y, double z, double q, double s6, double e24, double e4)
{
    a.x = x; a.y = y; a.z = z; a.q = q; a.s6 = s6; a.e24 = e24; a.e4 = e4; a.pad = 0;
}

template <typename T>
void fill_head(ClusterT<T> &h, const Prepared &p, int c)
{
    h.px = p.cl_pos[3 * c]; h.py = p.cl_pos[3 * c + 1]; h.pz = p.cl_pos[3 * c + 2];
    double mr = p.multipole_radius + p.cl_maxdist[c];
    mr *= mr;
    double dr = p.dispersion_radius + p.cl_maxdist[c];
    dr *= dr;
    h.mpr = mr; h.dpr = dr; h.q = p.cl_q[c];
    h.d0 = p.cl_dip[3 * c]; h.d1 = p.cl_dip[3 * c + 1]; h.d2 = p.cl_dip[3 * c + 2];
    for (int k = 0; k < 9; k++) h.Q[k] = p.cl_quad[9 * c + k];
    // the band branch reads data[n] with n the cluster index (src/forceAccel.cpp:353-355)
    const int qi = c < p.n ? c : 0;
    h.qs6 = p.sig6[qi]; h.qe24 = p.eps24m[qi];
    h.start = p.cl_start[c]; h.end = p.cl_start[c + 1]; h.rep = p.cl_replace[c]; h.pad = 0;
}

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

int upload(const cosims_mol *M)
{
    const Prepared &p = M->prep;
    const int n = p.n, ncl = p.ncl;
    const int npad = (n + 63) / 64 * 64;
    size_t o64 = 0, o32 = align_up(o64 + sizeof(AtomF64) * n), oc64 = align_up(o32 + sizeof(AtomF32) * n),
           oc32 = align_up(oc64 + sizeof(AtomF64) * n), oh64 = align_up(oc32 + sizeof(AtomF32) * n),
           oh32 = align_up(oh64 + sizeof(ClusterF64) * ncl), osoa = align_up(oh32 + sizeof(ClusterF32) * ncl),
           oxq = align_up(osoa + sizeof(double) * 7 * npad), ose = align_up(oxq + sizeof(float4) * npad),
           orec = align_up(ose + sizeof(float2) * npad), og0 = align_up(orec + sizeof(MolRec)),
           og1 = align_up(og0 + sizeof(float4) * ncl), og2 = align_up(og1 + sizeof(float4) * ncl),
           ocse = align_up(og2 + sizeof(float4) * ncl), ocxq = align_up(ocse + sizeof(int2) * ncl),
           ocab = align_up(ocxq + sizeof(float4) * (ncl > 0 ? n : 0)),
           ocdm = align_up(ocab + sizeof(float2) * (ncl > 0 ? n : 0)),
           odg0 = align_up(ocdm + sizeof(unsigned long long) * ((ncl + 63) / 64)),
           odg1 = align_up(odg0 + 4 * sizeof(double) * ncl), odg2 = align_up(odg1 + 4 * sizeof(double) * ncl),
           total = align_up(odg2 + 4 * sizeof(double) * ncl);
    std::vector<char> host(total, 0);
    auto *a64 = (AtomF64 *)(host.data() + o64);
    auto *a32 = (AtomF32 *)(host.data() + o32);
    auto *c64 = (AtomF64 *)(host.data() + oc64);
    auto *c32 = (AtomF32 *)(host.data() + oc32);
    auto *h64 = (ClusterF64 *)(host.data() + oh64);
    auto *h32 = (ClusterF32 *)(host.data() + oh32);
    for (int i = 0; i < n; i++) {
        const double *x = &p.xyz[3 * i];
        fill_atom(a64[i], x[0], x[1], x[2], p.q[i], p.sig6[i], p.eps24m[i], p.eps4[i]);
        fill_atom(a32[i], x[0], x[1], x[2], p.q[i], p.sig6[i], p.eps24m[i], p.eps4[i]);
    }
    for (int k = 0; k < (ncl > 0 ? n : 0); k++) {
        const int i = p.cl_atoms[k];
        const double *x = &p.xyz[3 * i];
        fill_atom(c64[k], x[0], x[1], x[2], p.q[i], p.sig6[i], p.eps24m[i], p.eps4[i]);
        fill_atom(c32[k], x[0], x[1], x[2], p.q[i], p.sig6[i], p.eps24m[i], p.eps4[i]);
    }
    for (int c = 0; c < ncl; c++) {
        fill_head(h64[c], p, c);
        fill_head(h32[c], p, c);
    }
    // FP32 cluster SoA of the mixed-precision clustered kernel
    auto *g0 = (float4 *)(host.data() + og0);
    auto *g1 = (float4 *)(host.data() + og1);
    auto *g2 = (float4 *)(host.data() + og2);
    auto *cse = (int2 *)(host.data() + ocse);
    auto *cxq = (float4 *)(host.data() + ocxq);
    auto *cab = (float2 *)(host.data() + ocab);
    auto *cdm = (unsigned long long *)(host.data() + ocdm);
    auto *dg0 = (double *)(host.data() + odg0);
    auto *dg1 = (double *)(host.data() + odg1);
    auto *dg2 = (double *)(host.data() + odg2);
    float dpr_max = 0.0f;
    double dpr_max64 = 0.0;
    int ndip = 0, ndip64 = 0;
    for (int c = 0; c < ncl; c++) {
        const ClusterF32 &h = h32[c];
        dpr_max = std::max(dpr_max, h.dpr);
        const bool dip = p.multipole_order >= 1 && !h.rep;
        g0[c] = make_float4(h.px, h.py, h.pz, h.mpr);
        g1[c] = make_float4(h.q, dip ? h.d0 : 0.0f, dip ? h.d1 : 0.0f, dip ? h.d2 : 0.0f);
        if (g1[c].y != 0.0f || g1[c].z != 0.0f || g1[c].w != 0.0f) cdm[c >> 6] |= 1ull << (c & 63);
        const ClusterF64 &d = h64[c];
        dpr_max64 = std::max(dpr_max64, d.dpr);
        double *e0 = dg0 + 4 * c, *e1 = dg1 + 4 * c, *e2 = dg2 + 4 * c;
        e0[0] = d.px; e0[1] = d.py; e0[2] = d.pz; e0[3] = d.mpr;
        e1[0] = d.q; e1[1] = dip ? d.d0 : 0.0; e1[2] = dip ? d.d1 : 0.0; e1[3] = dip ? d.d2 : 0.0;
        ndip += (cdm[c >> 6] >> (c & 63)) & 1;
        ndip64 += e1[1] != 0.0 || e1[2] != 0.0 || e1[3] != 0.0;
        e2[0] = d.dpr; e2[1] = d.qs6; e2[2] = d.qe24; e2[3] = 0.0;
        float A, B;
        const int qi = c < n ? c : 0;
        lj_constants(p.sig6[qi], p.eps24m[qi], A, B);
        g2[c] = make_float4(h.dpr, A, B, 0.0f);
        cse[c] = make_int2(h.start, h.end);
    }
    for (int k = 0; k < (ncl > 0 ? n : 0); k++) {
        const AtomF32 &a = c32[k];
        const int i = p.cl_atoms[k];
        cxq[k] = make_float4(a.x, a.y, a.z, a.q);
        float A, B;
        lj_constants(p.sig6[i], p.eps24m[i], A, B);
        cab[k] = make_float2(A, B);
    }
    // FP64 SoA in atom order, padded with inert atoms (no charge, no LJ, far away)
    double *soa = (double *)(host.data() + osoa);
    for (int i = 0; i < npad; i++) {
        const bool real = i < n;
        soa[i] = real ? p.xyz[3 * i] : 1.0e6;
        soa[npad + i] = real ? p.xyz[3 * i + 1] : 1.0e6;
        soa[2 * npad + i] = real ? p.xyz[3 * i + 2] : 1.0e6;
        soa[3 * npad + i] = real ? p.q[i] : 0.0;
        soa[4 * npad + i] = real ? p.sig6[i] : 0.0;
        soa[5 * npad + i] = real ? p.eps24m[i] : 0.0;
        soa[6 * npad + i] = real ? p.eps4[i] : 0.0;
    }
    // FP32 SoA of the group kernels (x, y, z, q | LJ constants A, B)
    auto *xq = (float4 *)(host.data() + oxq);
    auto *se = (float2 *)(host.data() + ose);
    for (int i = 0; i < npad; i++) {
        xq[i] = make_float4((float)soa[i], (float)soa[npad + i], (float)soa[2 * npad + i], (float)soa[3 * npad + i]);
        float A, B;
        lj_constants(soa[4 * npad + i], soa[5 * npad + i], A, B);
        se[i] = make_float2(A, B);
    }
    HIP_CHECK(dev_cache_get(total, &M->ddev, &M->dcap, &M->dmem));
    M->dbytes = total;
    char *d = (char *)M->dmem;
    MolDev &m = M->md;
    m.n = n; m.ncl = ncl; m.force_type = p.force_type; m.charged = p.charged ? 1 : 0;
    m.order = p.multipole_order; m.projection = p.projection ? 1 : 0;
    m.mu = p.mu; m.dipole_scalar = p.dipole_scalar; m.a = p.a; m.b = p.b; m.c = p.c;
    m.maxrad = std::max(p.a, std::max(p.b, p.c));
    m.ia = 1.0 / p.a; m.ib = 1.0 / p.b; m.ic = 1.0 / p.c;
    m.a64 = (const AtomF64 *)(d + o64); m.a32 = (const AtomF32 *)(d + o32);
    m.c64 = (const AtomF64 *)(d + oc64); m.c32 = (const AtomF32 *)(d + oc32);
    m.h64 = (const ClusterF64 *)(d + oh64); m.h32 = (const ClusterF32 *)(d + oh32);
    m.cg0 = (const float4 *)(d + og0); m.cg1 = (const float4 *)(d + og1); m.cg2 = (const float4 *)(d + og2);
    m.cse = (const int2 *)(d + ocse); m.cxq = (const float4 *)(d + ocxq); m.cab = (const float2 *)(d + ocab);
    m.cdm = (const unsigned long long *)(d + ocdm);
    m.dg0 = (const double *)(d + odg0); m.dg1 = (const double *)(d + odg1); m.dg2 = (const double *)(d + odg2);
    m.cl_dpr_max64 = dpr_max64;
    m.cl_dpr_max = dpr_max;
    m.cl_ndip = ndip;
    m.cl_ndip64 = ndip64;
    M->soa = (const double *)(d + osoa);
    M->npad = npad;
    MolRec &r = M->rec;
    r.m = m;
    r.xq = (const float4 *)(d + oxq);
    r.se = (const float2 *)(d + ose);
    r.soa = M->soa;
    r.npad = npad;
    r.dt0 = M->params.dt;
    std::memcpy(host.data() + orec, &r, sizeof(MolRec));
    M->rec_dev = (const MolRec *)(d + orec);
    HIP_CHECK(hipMemcpy(M->dmem, host.data(), total, hipMemcpyHostToDevice));
    return 0;
}

// Device upload on first GPU use; every later call sees the same result.
int ensure_device(const cosims_mol *M)
{
    std::call_once(M->up_once, [M]() {
        M->up_rc = device_init(M->dev);
        if (M->up_rc == 0) M->up_rc = upload(M);
        if (M->up_rc) M->up_err = g_err;
        else M->on_device.store(true);
    });
    if (M->up_rc) g_err = M->up_err;
    return M->up_rc;
}

int finish_create(cosims_mol *M, cosims_mol **out, std::chrono::steady_clock::time_point t0)
{
    M->setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = M;
    return 0;
}

int grid_for(const void *fn, int wg, int blocks_per_cu, int cus, int64_t items)
{
    int occ = 0;
    if (blocks_per_cu <= 0) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, wg, 0) != hipSuccess || occ <= 0) occ = 2;
    } else {
        occ = blocks_per_cu;
    }
    int64_t g = (int64_t)occ * cus;
    int64_t need = (items + wg - 1) / wg;
    return (int)std::max<int64_t>(1, std::min(g, need));
}

}  // namespace

// ================================================================ C ABI ===
extern "C" {

const char *cosims_strerror(int code)
{
    switch (code) {
    case COSIMS_OK: return "ok";
    case COSIMS_E_INVALID: return "invalid argument";
    case COSIMS_E_HIP: return "HIP runtime error";
    case COSIMS_E_NOMEM: return "out of memory";
    case COSIMS_E_SETUP: return "molecule setup failed";
    case COSIMS_E_IO: return "file could not be read";
    case COSIMS_E_UNSUPPORTED: return "input outside the defined behaviour of the reference";
    case COSIMS_E_NODEVICE: return "no usable gfx950 device";
    case COSIMS_E_PENDING: return "request not finished";
    default: return "unknown error";
    }
}

const char *cosims_last_error(void) { return g_err.c_str(); }

int cosims_device_sync(void)
{
    HIP_CHECK(hipDeviceSynchronize());
    return 0;
}

int cosims_device_fma_rate(double *tflops)
{
    if (!tflops) return fail(COSIMS_E_INVALID, "cosims_device_fma_rate: null output");
    Device d;
    if (int rc = device_init(d)) return rc;
    const int blocks = d.cus * 32, threads = 256, iters = 20000;
    DBuf buf;
    HIP_CHECK(buf.alloc(sizeof(float) * blocks * threads));
    Stream st;
    HIP_CHECK(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    hipLaunchKernelGGL(k_fma_rate, dim3(blocks), dim3(threads), 0, st.s, buf.as<float>(), 100, 0.999f, 0.001f);
    float best = 1e30f;
    for (int rep = 0; rep < 3; rep++) {
        HIP_CHECK(hipEventRecord(e0, st.s));
        hipLaunchKernelGGL(k_fma_rate, dim3(blocks), dim3(threads), 0, st.s, buf.as<float>(), iters, 0.999f, 0.001f);
        HIP_CHECK(hipEventRecord(e1, st.s));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        best = std::min(best, ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *tflops = 2.0 * 16 * (double)iters * blocks * threads / (best * 1e-3) / 1e12;
    return 0;
}

int cosims_device_info(char *name, size_t cap, int32_t *cus)
{
    Device d;
    int rc = device_init(d);
    if (rc) return rc;
    if (name && cap) std::snprintf(name, cap, "%s", d.name);
    if (cus) *cus = d.cus;
    return 0;
}

void cosims_params_default(cosims_params *p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->temp = 298; p->dt = 0.01; p->traj = 50000; p->iter = 10; p->charge = 0; p->seed = 0;
    p->multipole = 1; p->dispersion_cutoff = 1; p->dispersion_radius = 0; p->multipole_radius = 0;
    p->multipole_order = 1; p->max_cluster_size = 4; p->projection = 1; p->print_rate = 20000;
    p->total_charge = 0; p->threads = 0;
}

int cosims_mol_create(const cosims_params *p, const cosims_atoms *a, cosims_mol **out)
{
    if (!p || !a || !out || a->n <= 0) return fail(COSIMS_E_INVALID, "cosims_mol_create: null argument or empty molecule");
    auto t0 = std::chrono::steady_clock::now();
    std::unique_ptr<cosims_mol> M(new cosims_mol());
    M->params = to_params(p);
    std::string err;
    int rc;
    if (p->setup_gpu) {
        GpuMarcher gm;
        rc = prepare(M->params, a->n, a->xyz, a->mass, a->sigma, a->epsilon, a->charge, a->sigma_max, M->prep, err,
                     &gm);
    } else {
        rc = prepare(M->params, a->n, a->xyz, a->mass, a->sigma, a->epsilon, a->charge, a->sigma_max, M->prep, err);
    }
    if (rc) return fail(rc, err);
    return finish_create(M.release(), out, t0);
}

int cosims_mol_create_prepared(const cosims_params *p, const cosims_prepared *s, cosims_mol **out)
{
    if (!p || !s || !out || s->n <= 0 || !s->xyz || !s->sig6 || !s->eps24m || !s->eps4)
        return fail(COSIMS_E_INVALID, "cosims_mol_create_prepared: missing arrays");
    auto t0 = std::chrono::steady_clock::now();
    std::unique_ptr<cosims_mol> M(new cosims_mol());
    M->params = to_params(p);
    Prepared &m = M->prep;
    const int n = s->n;
    m.n = n;
    m.xyz.assign(s->xyz, s->xyz + 3 * n);
    m.q.assign(n, 0.0);
    if (s->charge) m.q.assign(s->charge, s->charge + n);
    m.sig6.assign(s->sig6, s->sig6 + n);
    m.eps24m.assign(s->eps24m, s->eps24m + n);
    m.eps4.assign(s->eps4, s->eps4 + n);
    m.mu = s->mu;
    m.a = s->a; m.b = s->b; m.c = s->c;
    m.dispersion_radius = s->dispersion_radius;
    m.multipole_radius = s->multipole_radius;
    m.ncl = s->ncl;
    if (m.ncl > 0) {
        if (!s->cl_start || !s->cl_atoms || !s->cl_center || !s->cl_maxdist)
            return fail(COSIMS_E_INVALID, "cosims_mol_create_prepared: cluster arrays missing");
        m.cl_start.assign(s->cl_start, s->cl_start + m.ncl + 1);
        if (m.cl_start[0] != 0 || m.cl_start[m.ncl] != n)
            return fail(COSIMS_E_INVALID, "cluster offsets must cover all atoms");
        for (int c = 0; c < m.ncl; c++)
            if (m.cl_start[c + 1] < m.cl_start[c]) return fail(COSIMS_E_INVALID, "cluster offsets not monotone");
        m.cl_atoms.assign(s->cl_atoms, s->cl_atoms + n);
        for (int32_t i : m.cl_atoms)
            if (i < 0 || i >= n) return fail(COSIMS_E_INVALID, "cluster member index out of range");
        m.cl_center.assign(s->cl_center, s->cl_center + 3 * m.ncl);
        m.cl_maxdist.assign(s->cl_maxdist, s->cl_maxdist + m.ncl);
    }
    int ft = force_type_of(p->charge != 0, p->multipole != 0, p->dispersion_cutoff != 0);
    if (ft != 4 && ft != 6 && m.ncl == 0) return fail(COSIMS_E_INVALID, "cut-off force types need clusters");
    std::string err;
    int rc = complete_prepared(M->params, m, err);
    if (rc) return fail(rc, err);
    return finish_create(M.release(), out, t0);
}

void cosims_mol_destroy(cosims_mol *m) { delete m; }

int cosims_mol_upload(const cosims_mol *m)
{
    if (!m) return fail(COSIMS_E_INVALID, "null handle");
    return ensure_device(m);
}

int cosims_mol_info_get(const cosims_mol *M, cosims_mol_info *o)
{
    if (!M || !o) return fail(COSIMS_E_INVALID, "null argument");
    const Prepared &p = M->prep;
    std::memset(o, 0, sizeof(*o));
    o->n = p.n; o->ncl = p.ncl; o->force_type = p.force_type; o->charged = p.charged;
    o->projection = p.projection; o->multipole_order = p.multipole_order;
    o->mu = p.mu; o->total_mass = p.total_mass; o->dipole_scalar = p.dipole_scalar;
    o->a = p.a; o->b = p.b; o->c = p.c; o->bmax = p.bmax; o->enlarge = p.enlarge;
    o->gmin = p.vd.gmin; o->gmax = p.vd.gmax; o->gmean = p.vd.mean; o->gmode = p.vd.mode;
    o->gsigma = p.vd.sigma; o->vd_a = p.vd.a; o->vd_norm = p.vd.norm; o->max_val = p.vd.max_val;
    o->cdf_frac = p.vd.cdf_frac; o->ccs_const = p.ccs_const; o->min_kin = p.vd.min_kin;
    o->dispersion_radius = p.dispersion_radius; o->multipole_radius = p.multipole_radius;
    std::memcpy(o->rot, p.rot, sizeof(o->rot));
    std::memcpy(o->max_xyz, p.max_xyz, sizeof(o->max_xyz));
    o->fit_abc[0] = p.fit_a; o->fit_abc[1] = p.fit_b; o->fit_abc[2] = p.fit_c; o->fit_rmsd = p.fit_rmsd;
    o->n_boundary = p.n_boundary;
    o->setup_ms = M->setup_ms;
    o->setup_boundary_ms = p.boundary_ms;
    o->setup_cluster_ms = p.cluster_ms;
    o->setup_moments_ms = p.moments_ms;
    return 0;
}

int cosims_mol_atoms(const cosims_mol *M, double *xyz, double *q, double *s6, double *e24, double *e4)
{
    if (!M) return fail(COSIMS_E_INVALID, "null handle");
    const Prepared &p = M->prep;
    if (xyz) std::memcpy(xyz, p.xyz.data(), sizeof(double) * 3 * p.n);
    if (q) std::memcpy(q, p.q.data(), sizeof(double) * p.n);
    if (s6) std::memcpy(s6, p.sig6.data(), sizeof(double) * p.n);
    if (e24) std::memcpy(e24, p.eps24m.data(), sizeof(double) * p.n);
    if (e4) std::memcpy(e4, p.eps4.data(), sizeof(double) * p.n);
    return 0;
}

int cosims_mol_clusters(const cosims_mol *M, int32_t *start, int32_t *atoms, double *center, double *maxdist,
                        double *pos, double *q, double *dip, double *quad, int32_t *rep)
{
    if (!M) return fail(COSIMS_E_INVALID, "null handle");
    const Prepared &p = M->prep;
    const int c = p.ncl;
    if (c == 0) return 0;
    if (start) std::memcpy(start, p.cl_start.data(), sizeof(int32_t) * (c + 1));
    if (atoms) std::memcpy(atoms, p.cl_atoms.data(), sizeof(int32_t) * p.n);
    if (center) std::memcpy(center, p.cl_center.data(), sizeof(double) * 3 * c);
    if (maxdist) std::memcpy(maxdist, p.cl_maxdist.data(), sizeof(double) * c);
    if (pos) std::memcpy(pos, p.cl_pos.data(), sizeof(double) * 3 * c);
    if (q) std::memcpy(q, p.cl_q.data(), sizeof(double) * c);
    if (dip) std::memcpy(dip, p.cl_dip.data(), sizeof(double) * 3 * c);
    if (quad) std::memcpy(quad, p.cl_quad.data(), sizeof(double) * 9 * c);
    if (rep) std::memcpy(rep, p.cl_replace.data(), sizeof(int32_t) * c);
    return 0;
}

int cosims_accel(const cosims_mol *M, int32_t prec, int64_t n, const double *pos, double *acc)
{
    if (!M || (n > 0 && (!pos || !acc)) || n < 0) return fail(COSIMS_E_INVALID, "cosims_accel: bad arguments");
    if (prec != COSIMS_PREC_FP64 && prec != COSIMS_PREC_MIXED) return fail(COSIMS_E_INVALID, "unknown precision");
    if (n == 0) return 0;
    if (int rc = ensure_device(M)) return rc;
    Stream st;
    HIP_CHECK(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
    DBuf dp, da;
    HIP_CHECK(dp.alloc(sizeof(double) * 3 * n));
    HIP_CHECK(da.alloc(sizeof(double) * 3 * n));
    HIP_CHECK(hipMemcpyAsync(dp.p, pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st.s));
    AccFn fn = accel_kernel(prec, M->prep.force_type);
    hipLaunchKernelGGL(fn, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st.s, M->md, n, dp.as<double>(),
                       da.as<double>());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(acc, da.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, st.s));
    HIP_CHECK(hipStreamSynchronize(st.s));
    return 0;
}

int cosims_potential(const cosims_mol *M, int64_t n, const double *pos, double *u)
{
    if (!M || n < 0 || (n > 0 && (!pos || !u))) return fail(COSIMS_E_INVALID, "cosims_potential: bad arguments");
    if (n == 0) return 0;
    if (int rc = ensure_device(M)) return rc;
    Stream st;
    HIP_CHECK(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
    DBuf dp, du;
    HIP_CHECK(dp.alloc(sizeof(double) * 3 * n));
    HIP_CHECK(du.alloc(sizeof(double) * n));
    HIP_CHECK(hipMemcpyAsync(dp.p, pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st.s));
    hipLaunchKernelGGL(k_potential, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st.s, M->md, n,
                       dp.as<double>(), du.as<double>());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(u, du.p, sizeof(double) * n, hipMemcpyDeviceToHost, st.s));
    HIP_CHECK(hipStreamSynchronize(st.s));
    return 0;
}

}  // extern "C"

namespace {

SampleArgs sample_args(const cosims_mol *M, uint64_t seed, uint32_t job)
{
    const Prepared &p = M->prep;
    SampleArgs A{};
    A.key0 = (uint32_t)seed; A.key1 = (uint32_t)(seed >> 32); A.job = job;
    A.gmin = p.vd.gmin; A.gmax = p.vd.gmax; A.max_val = p.vd.max_val; A.vd_a = p.vd.a; A.vd_norm = p.vd.norm;
    A.bmax = p.bmax;
    A.maxdist = (double)(float)p.bmax;  // monteTraj::maxDist is a float (src/monteTraj.h:67, :87)
    A.ea = p.a; A.eb = p.b; A.ec = p.c; A.projection = p.projection ? 1 : 0;
    return A;
}

// Counters of TrajBufs::evals: the lane kernel's 2 per attempt round, or a
// group launch's kEvalSlots.
constexpr int kEvalBufSlots = 16;
static_assert(kEvalBufSlots >= 2 * COSIMS_MAX_ROUNDS && kEvalBufSlots >= kEvalSlots, "evals slots");

// Work buffers of one batch of trajectories (device).
struct TrajBufs {
    DBuf p0, v0, g, status, chi, steps, tot, tries, maxtries, exitp, exitv2, e0, ergy, items[2], counters, evals,
        bin, order;
    int alloc(int64_t n)
    {
        HIP_CHECK(bin.alloc(n));
        HIP_CHECK(order.alloc(sizeof(int32_t) * 2 * kOrderBins));
        HIP_CHECK(p0.alloc(sizeof(double) * 3 * n));
        HIP_CHECK(v0.alloc(sizeof(double) * 3 * n));
        HIP_CHECK(g.alloc(sizeof(double) * n));
        HIP_CHECK(status.alloc(n));
        HIP_CHECK(chi.alloc(sizeof(double) * n));
        HIP_CHECK(steps.alloc(sizeof(int32_t) * n));
        HIP_CHECK(tot.alloc(sizeof(int32_t) * n));
        HIP_CHECK(tries.alloc(sizeof(int32_t) * n));
        HIP_CHECK(maxtries.alloc(sizeof(int32_t) * n));
        HIP_CHECK(exitp.alloc(sizeof(double) * 3 * n));
        HIP_CHECK(exitv2.alloc(sizeof(double) * n));
        HIP_CHECK(e0.alloc(sizeof(double) * n));
        HIP_CHECK(ergy.alloc(sizeof(double) * n));
        HIP_CHECK(items[0].alloc(sizeof(int32_t) * n));
        HIP_CHECK(items[1].alloc(sizeof(int32_t) * n));
        HIP_CHECK(counters.alloc(sizeof(int32_t) * 64));
        HIP_CHECK(evals.alloc(sizeof(unsigned long long) * kEvalBufSlots));
        return 0;
    }
    // The group-kernel path's buffers for n trajectories as 256-byte aligned
    // slices of [base, base + cap) from offset *off on (advanced past them);
    // the lane kernel's attempt buffers (tries, maxtries, exitp, e0) stay
    // unset.  false when they do not fit.
    bool carve(uint8_t *base, size_t cap, size_t *off, int64_t n)
    {
        bool ok = true;
        auto take = [&](DBuf &b, size_t bytes) { ok = ok && carve_one(b, base, cap, off, bytes); };
        take(bin, n);
        take(order, sizeof(int32_t) * 2 * kOrderBins);
        take(p0, sizeof(double) * 3 * n);
        take(v0, sizeof(double) * 3 * n);
        take(g, sizeof(double) * n);
        take(status, n);
        take(chi, sizeof(double) * n);
        take(steps, sizeof(int32_t) * n);
        take(tot, sizeof(int32_t) * n);
        take(exitv2, sizeof(double) * n);
        take(ergy, sizeof(double) * n);
        take(items[0], sizeof(int32_t) * n);
        take(items[1], sizeof(int32_t) * n);
        take(counters, sizeof(int32_t) * 64);
        take(evals, sizeof(unsigned long long) * kEvalBufSlots);
        return ok;
    }
};

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_copy_i32(const int32_t *src, int32_t *dst)
{
    if (threadIdx.x == 0) *dst = *src;
}

// Counting sort of a work list by order bucket, longest first:
// items[b .. e) -> out[b .. e), b = *first (0 when first is NULL), e = *last
// (n when last is NULL); bin is indexed by item.  hist: kOrderBins counts,
// then kOrderBins cursors, zero on entry.  With ends set, workgroup row y
// sorts work-list position p = p0 + y on its own: b = ends[p - 1] (0 for
// p = 0), e = ends[p], its own 2 kOrderBins of hist.
constexpr int kOrderChunk = 2048;  // items a workgroup of the scatter pass places

__device__ __forceinline__ void order_range(const int32_t *first, const int32_t *last, int64_t n, const int32_t *ends,
                                            int p0, int64_t &b, int64_t &e, int32_t *&hist)
{
    if (ends) {
        const int p = p0 + (int)blockIdx.y;
        b = p ? ends[p - 1] : 0;
        e = ends[p];
        hist += (size_t)2 * kOrderBins * blockIdx.y;
        return;
    }
    b = first ? *first : 0;
    e = last ? *last : n;
}

__global__ void __launch_bounds__(256) k_order_hist(const int32_t *items, const int32_t *first, const int32_t *last,
                                                    int64_t n, const uint8_t *bin, int32_t *hist,
                                                    const int32_t *ends, int p0)
{
    __shared__ int32_t h[kOrderBins];
    int64_t b, e;
    order_range(first, last, n, ends, p0, b, e, hist);
    for (int i = threadIdx.x; i < kOrderBins; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (int64_t i = b + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e; i += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&h[bin[items[i]]], 1);
    __syncthreads();
    for (int i = threadIdx.x; i < kOrderBins; i += blockDim.x)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

__global__ void __launch_bounds__(256) k_order_scatter(const int32_t *items, int32_t *out, const int32_t *first,
                                                       const int32_t *last, int64_t n, const uint8_t *bin,
                                                       int32_t *hist, const int32_t *ends, int p0)
{
    __shared__ int32_t start[kOrderBins], cnt[kOrderBins], base[kOrderBins];
    int64_t b, e;
    order_range(first, last, n, ends, p0, b, e, hist);
    const int64_t c0 = b + (int64_t)blockIdx.x * kOrderChunk;
    if (c0 >= e) return;  // uniform over the workgroup
    const int64_t c1 = min(c0 + (int64_t)kOrderChunk, e);
    for (int i = threadIdx.x; i < kOrderBins; i += blockDim.x) cnt[i] = 0;
    if (threadIdx.x == 0) {  // bucket starts, longest bucket first
        int32_t acc = 0;
        for (int k = kOrderBins - 1; k >= 0; k--) {
            start[k] = acc;
            acc += hist[k];
        }
    }
    __syncthreads();
    constexpr int kPer = kOrderChunk / 256;
    int32_t it[kPer], rank[kPer];
    uint8_t bk[kPer];
#pragma unroll
    for (int r = 0; r < kPer; r++) {
        const int64_t i = c0 + r * 256 + threadIdx.x;
        it[r] = i < c1 ? items[i] : -1;
        bk[r] = it[r] >= 0 ? bin[it[r]] : 0;
        rank[r] = it[r] >= 0 ? atomicAdd(&cnt[bk[r]], 1) : 0;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kOrderBins; k += blockDim.x)
        base[k] = cnt[k] ? atomicAdd(&hist[kOrderBins + k], cnt[k]) : 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kPer; r++)
        if (it[r] >= 0) out[b + start[bk[r]] + base[bk[r]] + rank[r]] = it[r];
}

// Enqueues the reorder of items[*first .. *last) (bounds on the device, or
// [0, n) when last is NULL) into out at the same positions; n bounds the
// range.  hist: 2 kOrderBins int32 of scratch, cleared here unless the
// caller has (clear false).
int order_items(const int32_t *items, int32_t *out, const int32_t *first, const int32_t *last, int64_t n,
                const uint8_t *bin, int32_t *hist, hipStream_t s, bool clear = true)
{
    if (n <= 0) return 0;
    if (clear)
        hipLaunchKernelGGL(k_fill_i32, dim3(2 * kOrderBins / 64), dim3(64), 0, s, hist, (int64_t)2 * kOrderBins, 0);
    const unsigned nh = (unsigned)std::min<int64_t>(1024, (n + 4095) / 4096);
    hipLaunchKernelGGL(k_order_hist, dim3(nh), dim3(256), 0, s, items, first, last, n, bin, hist,
                       (const int32_t *)nullptr, 0);
    const unsigned nc = (unsigned)((n + kOrderChunk - 1) / kOrderChunk);
    hipLaunchKernelGGL(k_order_scatter, dim3(nc), dim3(256), 0, s, items, out, first, last, n, bin, hist,
                       (const int32_t *)nullptr, 0);
    HIP_CHECK(hipGetLastError());
    return 0;
}

// The same for work-list positions p0 .. p0 + npos - 1 at once, each range
// [ends[p - 1], ends[p]) sorted on its own (at most max_n items each); hist:
// 2 kOrderBins int32 per position, zero on entry.
int order_positions(const int32_t *items, int32_t *out, const int32_t *ends, int p0, int npos, int64_t max_n,
                    const uint8_t *bin, int32_t *hist, hipStream_t s)
{
    if (npos <= 0 || max_n <= 0) return 0;
    const unsigned nh = (unsigned)std::min<int64_t>(64, (max_n + 4095) / 4096);
    hipLaunchKernelGGL(k_order_hist, dim3(nh, (unsigned)npos), dim3(256), 0, s, items, (const int32_t *)nullptr,
                       (const int32_t *)nullptr, (int64_t)0, bin, hist, ends, p0);
    const unsigned nc = (unsigned)((max_n + kOrderChunk - 1) / kOrderChunk);
    hipLaunchKernelGGL(k_order_scatter, dim3(nc, (unsigned)npos), dim3(256), 0, s, items, out,
                       (const int32_t *)nullptr, (const int32_t *)nullptr, (int64_t)0, bin, hist, ends, p0);
    HIP_CHECK(hipGetLastError());
    return 0;
}



// Runs the attempt rounds over the hit list in B.items[0] (count in
// counters[0]).  Returns per-round stats.
struct RoundStats {
    int rounds = 0;
    double traj_ms[COSIMS_MAX_ROUNDS] = {0};
    int64_t items[COSIMS_MAX_ROUNDS] = {0};
    int64_t steps[COSIMS_MAX_ROUNDS] = {0}, evals[COSIMS_MAX_ROUNDS] = {0};
    double check_ms = 0;
    int64_t retries = 0;
};

int run_rounds(const cosims_mol *M, int prec, int64_t n, TrajBufs &B, hipStream_t s, int wg, int bpc,
               RoundStats &R)
{
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_fill_i32, dim3(nb), dim3(256), 0, s, B.tries.as<int32_t>(), n, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(nb), dim3(256), 0, s, B.maxtries.as<int32_t>(), n, 8);
    hipLaunchKernelGGL(k_fill_i32, dim3(nb), dim3(256), 0, s, B.tot.as<int32_t>(), n, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, B.evals.as<int32_t>(), 2 * kEvalBufSlots, 0);
    HIP_CHECK(hipGetLastError());
    int32_t *ctr = B.counters.as<int32_t>();
    int32_t count = 0;
    HIP_CHECK(hipMemcpyAsync(&count, ctr, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    TrajFn fn = traj_kernel(prec, M->prep.force_type);
    hipEvent_t e[4];
    for (auto &x : e) HIP_CHECK(hipEventCreate(&x));
    int cur = 0;
    int rc = 0;
    for (int r = 0; r < COSIMS_MAX_ROUNDS && count > 0; r++) {
        // counters: [1] = work counter of this round, [2] = retry count
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, ctr + 1, 2, 0);
        TrajArgs T{};
        T.m = M->md;
        T.items = B.items[cur].as<int32_t>();
        T.n_items = count;
        T.counter = ctr + 1;
        T.p0 = B.p0.as<double>(); T.v0 = B.v0.as<double>();
        T.tries = B.tries.as<int32_t>(); T.maxtries = B.maxtries.as<int32_t>();
        T.status = B.status.as<int8_t>(); T.chi = B.chi.as<double>();
        T.steps = B.steps.as<int32_t>(); T.totsteps = B.tot.as<int32_t>();
        T.exit_pos = B.exitp.as<double>(); T.exit_v2 = B.exitv2.as<double>();
        T.dt0 = M->params.dt;
        T.max_steps = kMaxStepsPerAttempt;
        T.evals = B.evals.as<unsigned long long>() + 2 * r;
        const int grid = grid_for((const void *)fn, wg, bpc, M->dev.cus, count);
        HIP_CHECK(hipEventRecord(e[0], s));
        hipLaunchKernelGGL(fn, dim3(grid), dim3(wg), 0, s, T);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipEventRecord(e[1], s));
        CheckArgs C{};
        C.m = M->md; C.items = T.items; C.n_items = count; C.first_round = (r == 0);
        C.p0 = T.p0; C.v0 = T.v0; C.exit_pos = T.exit_pos; C.exit_v2 = T.exit_v2;
        C.e0 = B.e0.as<double>(); C.chi = T.chi; C.ergy = B.ergy.as<double>(); C.status = T.status;
        C.tries = T.tries; C.maxtries = T.maxtries; C.steps = T.steps;
        C.next_items = B.items[cur ^ 1].as<int32_t>(); C.next_count = ctr + 2;
        hipLaunchKernelGGL(k_check, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, C);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipEventRecord(e[2], s));
        int32_t next = 0;
        HIP_CHECK(hipMemcpyAsync(&next, ctr + 2, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        float ms = 0, ms2 = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, e[0], e[1]));
        HIP_CHECK(hipEventElapsedTime(&ms2, e[1], e[2]));
        R.traj_ms[r] = ms;
        R.items[r] = count;
        R.check_ms += ms2;
        R.rounds = r + 1;
        R.retries += next;
        count = next;
        cur ^= 1;
    }
    for (auto &x : e) (void)hipEventDestroy(x);
    unsigned long long ev[2 * COSIMS_MAX_ROUNDS];
    HIP_CHECK(hipMemcpy(ev, B.evals.p, sizeof(ev), hipMemcpyDeviceToHost));
    for (int r = 0; r < R.rounds; r++) {
        R.steps[r] = (int64_t)ev[2 * r];
        R.evals[r] = (int64_t)ev[2 * r + 1];
    }
    if (count > 0) rc = fail(COSIMS_E_HIP, "attempt rounds exceeded the reference's retry bound");
    return rc;
}

}  // namespace


namespace {

// ------------------------------------------------- group-mode launches ---
using GrpFn = void (*)(GroupArgs);
using MultiFn = void (*)(MultiArgs);

template <int P, int FT>
GrpFn grp_fn_l(int L)
{
    switch (L) {
    case 4: return k_traj_grp<P, FT, 4>;
    case 8: return k_traj_grp<P, FT, 8>;
    case 16: return k_traj_grp<P, FT, 16>;
    case 32: return k_traj_grp<P, FT, 32>;
    case 64: return k_traj_grp<P, FT, 64>;
    default: return nullptr;
    }
}

GrpFn group_kernel(int prec, int ft, int L)
{
    if (ft == 4) return prec == COSIMS_PREC_FP64 ? grp_fn_l<0, 4>(L) : grp_fn_l<1, 4>(L);
    if (ft == 6) return prec == COSIMS_PREC_FP64 ? grp_fn_l<0, 6>(L) : grp_fn_l<1, 6>(L);
    return nullptr;
}

// The LDS-header clustered kernel (mixed precision, 64 lanes).
GrpFn cl_lds_kernel(int ft)
{
    switch (ft) {
    case 1: return k_traj_cl_lds<1>;
    case 2: return k_traj_cl_lds<2>;
    case 3: return k_traj_cl_lds<3>;
    case 5: return k_traj_cl_lds<5>;
    default: return nullptr;
    }
}

using LdsMultiFn = void (*)(LdsMultiArgs);

template <int P, int FT>
LdsMultiFn ldsm_fn_l(int L)
{
    switch (L) {
    case 4: return k_traj_lds_multi<P, FT, 4>;
    case 8: return k_traj_lds_multi<P, FT, 8>;
    case 16: return k_traj_lds_multi<P, FT, 16>;
    case 32: return k_traj_lds_multi<P, FT, 32>;
    case 64: return k_traj_lds_multi<P, FT, 64>;
    default: return nullptr;
    }
}

LdsMultiFn lds_multi_kernel(int prec, int ft, int L)
{
    if (ft == 4) return prec == COSIMS_PREC_FP64 ? ldsm_fn_l<0, 4>(L) : ldsm_fn_l<1, 4>(L);
    if (ft == 6) return prec == COSIMS_PREC_FP64 ? ldsm_fn_l<0, 6>(L) : ldsm_fn_l<1, 6>(L);
    return nullptr;
}

using RateFn = void (*)(GroupArgs, int, double *);

template <int P, int FT>
RateFn rate_fn_l(int L)
{
    switch (L) {
    case 4: return k_force_rate<P, FT, 4>;
    case 8: return k_force_rate<P, FT, 8>;
    case 16: return k_force_rate<P, FT, 16>;
    case 32: return k_force_rate<P, FT, 32>;
    case 64: return k_force_rate<P, FT, 64>;
    default: return nullptr;
    }
}

RateFn rate_kernel(int prec, int ft, int L)
{
    if (ft == 4) return prec == COSIMS_PREC_FP64 ? rate_fn_l<0, 4>(L) : rate_fn_l<1, 4>(L);
    if (ft == 6) return prec == COSIMS_PREC_FP64 ? rate_fn_l<0, 6>(L) : rate_fn_l<1, 6>(L);
    return nullptr;
}

template <int P, int FT>
GrpFn tile_fn_l(int L)
{
    switch (L) {
    case 4: return k_traj_tile<P, FT, 4>;
    case 8: return k_traj_tile<P, FT, 8>;
    case 16: return k_traj_tile<P, FT, 16>;
    case 32: return k_traj_tile<P, FT, 32>;
    case 64: return k_traj_tile<P, FT, 64>;
    default: return nullptr;
    }
}

GrpFn tile_kernel(int prec, int ft, int L)
{
    if (ft == 4) return prec == COSIMS_PREC_FP64 ? tile_fn_l<0, 4>(L) : tile_fn_l<1, 4>(L);
    if (ft == 6) return prec == COSIMS_PREC_FP64 ? tile_fn_l<0, 6>(L) : tile_fn_l<1, 6>(L);
    return nullptr;
}

template <int P, int FT>
MultiFn multi_fn_l(int L)
{
    switch (L) {
    case 4: return k_traj_multi<P, FT, 4>;
    case 8: return k_traj_multi<P, FT, 8>;
    case 16: return k_traj_multi<P, FT, 16>;
    case 32: return k_traj_multi<P, FT, 32>;
    case 64: return k_traj_multi<P, FT, 64>;
    default: return nullptr;
    }
}

template <int P>
MultiFn multi_fn(int ft, int L)
{
    switch (ft) {
    case 1: return multi_fn_l<P, 1>(L);
    case 2: return multi_fn_l<P, 2>(L);
    case 3: return multi_fn_l<P, 3>(L);
    case 4: return multi_fn_l<P, 4>(L);
    case 5: return multi_fn_l<P, 5>(L);
    case 6: return multi_fn_l<P, 6>(L);
    default: return nullptr;
    }
}

MultiFn multi_kernel(int prec, int ft, int L) { return prec == COSIMS_PREC_FP64 ? multi_fn<0>(ft, L) : multi_fn<1>(ft, L); }

size_t multi_lds_bytes(int ft, int L, int wg)
{
    return multi_uses_lists(ft, L) ? (size_t)(wg / 64) * (kListInts + kCountInts) * sizeof(int) : 0;
}

size_t group_lds_bytes(const cosims_mol *M, int prec)
{
    return (size_t)M->npad * (prec == COSIMS_PREC_FP64 ? 6 * sizeof(double) : sizeof(float4) + sizeof(float2));
}

// The LDS-staged kernel serves the exact force types when the molecule fits
// one workgroup's LDS, the tiled kernel the larger ones; the clustered types
// run the table kernel.
bool lds_mode(const cosims_mol *M, int prec)
{
    const int ft = M->prep.force_type;
    return (ft == 4 || ft == 6) && group_lds_bytes(M, prec) <= M->dev.lds_max;
}
bool tile_mode(const cosims_mol *M, int prec)
{
    const int ft = M->prep.force_type;
    return (ft == 4 || ft == 6) && !lds_mode(M, prec);
}

// Atoms per LDS tile of k_traj_tile at workgroup size wg: the share of a
// CU's LDS that leaves room for 1024 / wg workgroups, a multiple of 64.
int tile_atoms(const cosims_mol *M, int prec, int wg)
{
    const size_t per_atom = prec == COSIMS_PREC_FP64 ? 6 * sizeof(double) : sizeof(float4) + sizeof(float2);
    const size_t budget = (size_t)M->dev.lds_max * (size_t)wg / 1024;
    const int t = (int)(budget / per_atom) / 64 * 64;
    return std::max(64, std::min(t, M->npad));
}
size_t tile_lds_bytes(int prec, int tile)
{
    return (size_t)tile * (prec == COSIMS_PREC_FP64 ? 6 * sizeof(double) : sizeof(float4) + sizeof(float2));
}

// A clustered-type molecule runs k_traj_cl_lds (cluster headers in LDS,
// one 1024-thread workgroup a CU) in mixed precision at 64 lanes when its
// multipole order is at most 1 and its headers fit LDS beside the rings of
// 16 wavefronts; otherwise the table kernel (headers through L2).
constexpr int kClLdsWg = 1024;
bool cl_lds_mode(const cosims_mol *M, int prec, int L)
{
    return prec == COSIMS_PREC_MIXED && L == 64 && cl_lds_kernel(M->prep.force_type) != nullptr &&
           M->md.order <= 1 && M->md.ncl > 0 && cl_lds_bytes(M->md.ncl, kClLdsWg) <= M->dev.lds_max;
}

// Launch arguments of the LDS-staged and tiled kernels for one molecule.
GroupArgs group_args(const cosims_mol *M)
{
    GroupArgs G{};
    G.m = M->md;
    G.npad = M->npad;
    G.soa = M->soa;
    G.dt0 = M->params.dt;
    G.xq = M->rec.xq;
    G.se = M->rec.se;
    return G;
}

// Work units one lane of a group walks per force evaluation: atoms, or
// clusters for the clustered types.
int group_units(const Prepared &p)
{
    const int ft = p.force_type;
    return (ft == 4 || ft == 6) ? p.n : p.ncl;
}

// The clustered force types run one group per wavefront (64 lanes: far
// clusters by lane, near and LJ clusters from ballot-built lists), whatever
// their cluster count: with fewer lanes a trajectory each lane steps through
// the branches of its own clusters, 4-14x slower on molecules of 300-5,000
// atoms (38-671 clusters).  Their size rule, tuner and batch classes give
// them 64 lanes.
bool clustered_type(int ft) { return ft == 1 || ft == 2 || ft == 3 || ft == 5; }

struct GroupRun {
    double ms = 0;
    int64_t steps = 0, evals = 0, pots = 0, pairs = 0;
    int64_t cl[6] = {0, 0, 0, 0, 0, 0};  // clustered-type branch counts (ClCount order)
    double flops = 0;                    // reference FLOPs (ref_flops)
    int grid = 0;
};

// Reference FLOPs of force evaluations, integration steps and potential
// atom pairs of one force type (multipole order for the far clusters),
// counted per branch as the reference writes them (the table in
// cosims_ccs.ref_flops' comment).
double ref_flops(int force_type, int order, int64_t evals, int64_t steps, int64_t pot_pairs, int64_t pairs,
                 const int64_t *cl)
{
    const double hdr = (double)cl[0], far = (double)cl[1], fdip = (double)cl[2], exlj = (double)cl[3],
                 exq = (double)cl[4], lj = (double)cl[5];
    double f = 0;
    switch (force_type) {
    case 4: f = 58.0 * (double)pairs + 21.0 * (double)evals; break;
    case 6: f = 22.0 * (double)pairs; break;
    case 1:
    case 2:
        f = 12.0 * (hdr - far) + 49.0 * far + 67.0 * fdip + (order >= 2 ? 89.0 * far : 0.0) +
            62.0 * exlj + 48.0 * exq + 22.0 * lj + 27.0 * (double)evals;
        break;
    case 3: f = 10.0 * hdr + 62.0 * exlj + 48.0 * exq + 21.0 * (double)evals; break;
    default: f = 10.0 * hdr + 22.0 * lj; break;  // type 5
    }
    const bool charged = force_type != 5 && force_type != 6;
    return f + 50.0 * (double)steps + (charged ? 24.0 : 15.0) * (double)pot_pairs;
}

TrajIO traj_io(TrajBufs &B, const int32_t *items, int n_items)
{
    TrajIO io{};
    io.items = items;
    io.n_items = n_items;
    io.counter = B.counters.as<int32_t>() + 1;
    io.p0 = B.p0.as<double>(); io.v0 = B.v0.as<double>();
    io.status = B.status.as<int8_t>(); io.chi = B.chi.as<double>(); io.ergy = B.ergy.as<double>();
    io.cchi = B.exitv2.as<double>();  // the lane kernel's exit buffer holds cos(chi) here
    io.steps = B.steps.as<int32_t>(); io.totsteps = B.tot.as<int32_t>();
    io.max_steps = kMaxStepsPerAttempt;
    io.evals = B.evals.as<unsigned long long>();
    return io;
}

// Grid of a persistent group launch: occupancy-limited, never more
// workgroups than the items can fill.
int64_t group_grid(const void *fn, int wg, size_t lds, int bpc, int cus, int64_t n_items, int L)
{
    int occ = bpc;
    if (occ <= 0 && (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, wg, lds) != hipSuccess || occ <= 0))
        occ = 1;
    const int groups_per_wg = wg / L;
    int64_t grid = (int64_t)occ * cus;
    return std::max<int64_t>(1, std::min<int64_t>(grid, (n_items + groups_per_wg - 1) / groups_per_wg));
}

// One persistent launch of fn(args) on stream s, timed with HIP events
// unless `timed` is off (out.ms = 0: no events are created), then acos of
// the accepted exits and the launch's counters.
template <class Args>
int timed_group_launch(void (*fn)(Args), const Args &args, int64_t grid, int wg, size_t lds, hipStream_t s,
                       TrajBufs &B, int n_items, GroupRun &out, bool timed = true)
{
    int32_t *ctr = B.counters.as<int32_t>();
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, ctr + 1, 2, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, B.evals.as<int32_t>(), 2 * kEvalBufSlots, 0);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (timed) {
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        HIP_CHECK(hipEventRecord(e0, s));
    }
    hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(wg), lds, s, args);
    HIP_CHECK(hipGetLastError());
    if (timed) HIP_CHECK(hipEventRecord(e1, s));
    hipLaunchKernelGGL(k_chi, dim3((unsigned)((n_items + 255) / 256)), dim3(256), 0, s, args.io.items, n_items,
                       args.io.status, args.io.cchi, args.io.chi, (const int32_t *)nullptr, (const int32_t *)nullptr);
    HIP_CHECK(hipGetLastError());
    unsigned long long ev[kEvalSlots];
    HIP_CHECK(hipMemcpyAsync(ev, B.evals.p, sizeof(ev), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    float ms = 0;
    if (timed) {
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    out.ms = ms;
    out.steps = (int64_t)ev[0];
    out.evals = (int64_t)ev[1];
    out.pots = (int64_t)ev[2];
    out.pairs = (int64_t)ev[3];
    for (int k = 0; k < 6; k++) out.cl[k] = (int64_t)ev[4 + k];
    out.grid = (int)grid;
    return 0;
}

// Table-kernel launch over the item list items[0..n_items): mols is a device
// array of MolRec, tmol the molecule of each trajectory (NULL: molecule 0).
int run_multi(const MolRec *mols, const int32_t *tmol, int ft, int prec, int cus, const int32_t *items, int n_items,
              TrajBufs &B, hipStream_t s, int L, int wg, int bpc, GroupRun &out, bool timed = true)
{
    MultiFn fn = multi_kernel(prec, ft, L);
    if (!fn) return fail(COSIMS_E_INVALID, "no table kernel for this force type / lane count");
    MultiArgs G{};
    G.io = traj_io(B, items, n_items);
    G.mols = mols;
    G.tmol = tmol;
    const size_t lds = multi_lds_bytes(ft, L, wg);
    const int64_t grid = group_grid((const void *)fn, wg, lds, bpc, cus, n_items, L);
    return timed_group_launch(fn, G, grid, wg, lds, s, B, n_items, out, timed);
}

// One persistent group launch of one molecule over the item list
// items[0 .. n_items) (B.items[0] when items is NULL).
int run_group(const cosims_mol *M, int prec, int n_items, TrajBufs &B, hipStream_t s, int L, int wg, int bpc,
              GroupRun &out, const int32_t *items = nullptr, bool timed = true)
{
    if (!items) items = B.items[0].as<int32_t>();
    const bool tiled = tile_mode(M, prec);
    if (cl_lds_mode(M, prec, L)) {
        GrpFn fn = cl_lds_kernel(M->prep.force_type);
        GroupArgs G = group_args(M);
        const size_t lds = cl_lds_bytes(M->md.ncl, kClLdsWg);
        HIP_CHECK(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        G.io = traj_io(B, items, n_items);
        const int64_t grid = group_grid((const void *)fn, kClLdsWg, lds, bpc, M->dev.cus, n_items, L);
        return timed_group_launch(fn, G, grid, kClLdsWg, lds, s, B, n_items, out, timed);
    }
    if (!tiled && !lds_mode(M, prec))
        return run_multi(M->rec_dev, nullptr, M->prep.force_type, prec, M->dev.cus, items, n_items, B, s, L, wg, bpc,
                         out, timed);
    GrpFn fn = tiled ? tile_kernel(prec, M->prep.force_type, L) : group_kernel(prec, M->prep.force_type, L);
    if (!fn) return fail(COSIMS_E_INVALID, "no group kernel for this force type / lane count");
    GroupArgs G = group_args(M);
    size_t lds = group_lds_bytes(M, prec);
    if (tiled) {
        G.tile = tile_atoms(M, prec, wg);
        lds = tile_lds_bytes(prec, G.tile);
    }
    HIP_CHECK(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    G.io = traj_io(B, items, n_items);
    const int64_t grid = group_grid((const void *)fn, wg, lds, bpc, M->dev.cus, n_items, L);
    return timed_group_launch(fn, G, grid, wg, lds, s, B, n_items, out, timed);
}

// k_sample into B with hit compaction into B.items[0] (count in counters[0]);
// timed with HIP events when ms is set.
int sample_into(const cosims_mol *M, uint64_t seed, uint32_t job, int64_t n, int64_t per_iter, TrajBufs &B,
                hipStream_t s, float *ms)
{
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, B.counters.as<int32_t>(), 64, 0);
    SampleArgs A = sample_args(M, seed, job);
    A.first = 0; A.n = n; A.per_iter = per_iter;
    A.p0 = B.p0.as<double>(); A.v0 = B.v0.as<double>(); A.g = B.g.as<double>();
    A.status = B.status.as<int8_t>(); A.items = B.items[0].as<int32_t>(); A.n_items = B.counters.as<int32_t>();
    A.bin = B.bin.as<uint8_t>(); A.dt0 = M->params.dt;
    if (!ms) {
        hipLaunchKernelGGL(k_sample, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, A);
        HIP_CHECK(hipGetLastError());
        return 0;
    }
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    HIP_CHECK(hipEventRecord(e0, s));
    hipLaunchKernelGGL(k_sample, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, A);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(e1, s));
    HIP_CHECK(hipEventSynchronize(e1));
    HIP_CHECK(hipEventElapsedTime(ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 0;
}

// Per-integral sums of k_reduce: total, total^2, max ergy, steps, hits, fails.
constexpr int kSums = 6;

// CCS of one integral (src/monteTraj.cpp:378-386) from its sums and the
// trajectory status codes.  Denominator currSuccTraj of one OpenMP thread
// (src/monteTraj.cpp:456-521): failures (a miss repeats the previous success
// flag) counted since the last progress update every print_rate trajectories.
void integral_ccs(const int8_t *st8, const double *sums, int64_t T, int64_t print_rate, double &omega,
                  double &omega_err)
{
    const int64_t upd = print_rate > 0 ? print_rate : std::max<int64_t>(1, (int64_t)(T / 5.0));
    bool succ = false;
    int64_t orbit = 0;
    for (int64_t i = 0; i < T; i++) {
        if (st8[i] != ST_MISS) succ = (st8[i] == ST_OK);
        if (!succ) orbit++;
        if ((i + 1) % upd == 0) orbit = 0;
    }
    const double d = (double)(T - orbit);
    const double tot = sums[0], tot2 = sums[1];
    omega = tot / d;
    omega_err = std::sqrt((tot2 / d - (tot / d) * (tot / d)) / d);
}

// integral_ccs on the device, one workgroup of 256 threads an integral.
// The denominator of one thread's integral (src/monteTraj.cpp:456-521) is T
// minus the trajectories since the last progress update whose success flag
// is off, a miss repeating the flag of the trajectory before it.  The last
// window [w0, T) is scanned in parallel: each thread takes a chunk, finds
// its last non-miss status, and with the last non-miss status before its
// chunk (an exclusive scan over the chunks, or before w0) counts the
// trajectories of its chunk that see the flag off.
__device__ void integral_ccs_wg(const int8_t *st, const double *sums, int T, int upd, double *omega,
                                double *omega_err)
{
    __shared__ int8_t last[256];
    __shared__ int8_t carry;
    __shared__ int orbit;
    const int tid = threadIdx.x;
    const int w0 = (T / upd) * upd;  // orbit counts reset at every multiple of upd
    const int len = T - w0;
    if (tid == 0) {
        int8_t c = -1;  // no non-miss yet: the flag is off
        for (int j = w0 - 1; j >= 0; j--)
            if (st[j] != ST_MISS) {
                c = st[j];
                break;
            }
        carry = c;
        orbit = 0;
    }
    const int chunk = (len + 255) / 256;
    const int a = min(len, tid * chunk), b = min(len, a + chunk);
    int8_t mine = -1;
    for (int i = a; i < b; i++)
        if (st[w0 + i] != ST_MISS) mine = st[w0 + i];
    last[tid] = mine;
    __syncthreads();
    int8_t in = carry;  // last non-miss status before this chunk
    for (int t = tid - 1; t >= 0; t--)
        if (last[t] >= 0) {
            in = last[t];
            break;
        }
    bool succ = in == ST_OK;
    int off = 0;
    for (int i = a; i < b; i++) {
        const int8_t s = st[w0 + i];
        if (s != ST_MISS) succ = s == ST_OK;
        off += succ ? 0 : 1;
    }
    if (off) atomicAdd(&orbit, off);
    __syncthreads();
    if (tid == 0) {
        const double den = (double)(T - orbit);
        const double tot = sums[0], tot2 = sums[1];
        *omega = tot / den;
        *omega_err = sqrt((tot2 / den - (tot / den) * (tot / den)) / den);
    }
}

// The iters integrals of a run of T trajectories each, workgroup = integral.
__global__ void __launch_bounds__(256) k_iter_ccs(const int8_t *status, const double *sums, int T, int upd,
                                                  double *omega, double *omega_err)
{
    const int it = blockIdx.x;
    integral_ccs_wg(status + (int64_t)it * T, sums + kSums * it, T, upd, omega + it, omega_err + it);
}

// Trajectory status codes as cosims_run_cfg.status gives them: 0 miss,
// 1 success, 2 failure.
__global__ void k_status_out(const int8_t *st, int64_t n, int32_t *out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = st[i] == ST_OK ? 1 : (st[i] == ST_MISS ? 0 : 2);
}

// Final statistics of monteTraj::runTraj (src/monteTraj.cpp:412-425) over
// iters integrals, plus the counters of their sums.
void final_stats(const double *omega, const double *omega_err, const double *sums, int iters, int64_t T,
                 cosims_ccs *out)
{
    double mx = 0;
    for (int k = 0; k < iters; k++) {
        mx = std::max(mx, sums[kSums * k + 2]);
        out->total_steps += (int64_t)sums[kSums * k + 3];
        out->n_hit += (int64_t)sums[kSums * k + 4];
        out->n_fail += (int64_t)sums[kSums * k + 5];
    }
    double avg = 0, sd = 0, we = 0, wm = 0;
    for (int k = 0; k < iters; k++) avg += omega[k];
    avg = avg / (double)iters;
    for (int k = 0; k < iters; k++) {
        sd += (omega[k] - avg) * (omega[k] - avg);
        we += 1.0 / (omega_err[k] * omega_err[k]);
        wm += omega[k] / (omega_err[k] * omega_err[k]);
    }
    wm /= we;
    out->mean = avg;
    out->sd = std::sqrt(sd / (double)iters);
    out->wmean = wm;
    out->werr = 1.0 / std::sqrt(we);
    out->max_ergy = mx;
    out->n_traj = (int64_t)iters * T;
}

void finish_ccs(const int8_t *status, const double *sums, int iters, int64_t T, int64_t print_rate, double *omega,
                double *omega_err, cosims_ccs *out)
{
    for (int k = 0; k < iters; k++)
        integral_ccs(status + (int64_t)k * T, sums + kSums * k, T, print_rate, omega[k], omega_err[k]);
    final_stats(omega, omega_err, sums, iters, T, out);
}

// Launch-shape search shared by the single-molecule and batch tuners:
// every lanes-per-trajectory L (the atom tile each lane strides over) x
// workgroup size W (W/L trajectories per workgroup) is timed on a small
// item set; the best W of each of the three fastest L is re-timed twice
// (the faster run counts: launch times vary by a few per cent) on the large
// set, where the end-of-launch tail weighs as in the real run (fewer lanes
// per trajectory mean slower trajectories, hence longer tails, which a small
// set over-weights).  run(set, L, W, ms) times one launch.  only64: the
// clustered types, whose lane count is 64 (clustered_type): W alone.
template <class RunFn>
int tune_shapes(int units, RunFn run, int &L_out, int &W_out, bool only64 = false)
{
    struct Cand {
        int L, W;
        double ms;
    };
    std::vector<Cand> cands;
    const int Ls[] = {4, 8, 16, 32, 64};
    const int Ws[] = {256, 512, 1024};
    // screen: each lane count once on the smallest set; those more than 3x
    // slower than the best are dropped (on large molecules of the clustered
    // types the slow shapes would otherwise cost minutes of tuning)
    std::vector<int> keep;
    {
        std::vector<std::pair<int, double>> sc;
        double best = 0;
        for (int L : Ls) {
            if (only64 ? L != 64 : (L > 4 && units < 8 * L)) continue;  // too few atoms per lane
            double ms = 0;
            if (int rc = run(2, L, 512, ms)) return rc;
            sc.push_back({L, ms});
            best = sc.size() == 1 ? ms : std::min(best, ms);
        }
        for (const auto &x : sc)
            if (x.second <= 3.0 * best) keep.push_back(x.first);
    }
    for (int L : keep) {
        for (int W : Ws) {
            double ms = 0;
            if (int rc = run(0, L, W, ms)) return rc;
            cands.push_back({L, W, ms});
        }
    }
    // finalists: the best workgroup size of each of the three best lane counts
    std::sort(cands.begin(), cands.end(), [](const Cand &a, const Cand &b) { return a.ms < b.ms; });
    std::vector<Cand> fin;
    for (const Cand &c : cands) {
        bool seen = false;
        for (const Cand &f : fin) seen |= f.L == c.L;
        if (!seen && fin.size() < 3) fin.push_back(c);
    }
    cands.swap(fin);
    // each finalist three times on sets of the run's size (set 3: another
    // Philox stream), the mean counts: with few lanes a trajectory the
    // longest trajectories run about as long as the launch, so a launch's
    // end can differ from run to run on the same trajectories (2,000 atoms
    // at 16 lanes without the raised issue priority of long trajectories:
    // 520 or 630 ms; with it 514-533 ms).  Within 2 % of the best, the most
    // lanes win (the steadier drain).
    for (Cand &c : cands)
        if (int rc = run(1, c.L, c.W, c.ms)) return rc;
    for (int set : {3, 1})
        for (Cand &c : cands) {
            double b = 0;
            if (int rc = run(set, c.L, c.W, b)) return rc;
            c.ms += b;
        }
    std::sort(cands.begin(), cands.end(), [](const Cand &a, const Cand &b) { return a.ms < b.ms; });
    Cand pick = cands[0];
    for (const Cand &c : cands)
        if (c.ms <= 1.02 * cands[0].ms && c.L > pick.L) pick = c;
    L_out = pick.L;
    W_out = pick.W;
    return 0;
}

// Single-molecule autotuner, timed on the molecule itself (fixed Philox
// stream).  The final stage runs on as many trajectories as the run it tunes
// for (capped: 4096 per CU mixed, 1024 FP64), because the end-of-launch tail
// weighs differently on small and large launches; results are cached on the
// handle per precision and power-of-two size of that set.
// Run size the tuner times a run of n_hint trajectories at, and the key of
// its tuned shape in cosims_mol::tuned (precision x power-of-two bucket).
int64_t tune_size(const cosims_mol *M, int prec, int64_t n_hint, int *key)
{
    const int pi = prec == COSIMS_PREC_FP64 ? 0 : 1;
    const int64_t cap = (int64_t)M->dev.cus * (pi ? 4096 : 1024);
    const int64_t n2 = std::max<int64_t>(16384, std::min<int64_t>(n_hint, cap));
    int bucket = 0;
    while ((int64_t(1) << (bucket + 1)) <= n2) bucket++;
    *key = 64 * pi + bucket;
    return n2;
}

int autotune(const cosims_mol *M, int prec, int64_t n_hint, int &L_out, int &wg_out)
{
    std::lock_guard<std::mutex> lk(M->tune_mu);
    int key = 0;
    const int64_t n2 = tune_size(M, prec, n_hint, &key);
    auto it = M->tuned.find(key);
    if (it != M->tuned.end()) {
        L_out = it->second.first;
        wg_out = it->second.second;
        return 0;
    }
    if (clustered_type(M->prep.force_type) && cl_lds_mode(M, prec, 64)) {
        // the clustered types run 64 lanes; with the headers in LDS the
        // workgroup is 1024 threads whatever is asked: nothing to time
        L_out = 64;
        wg_out = kClLdsWg;
        M->tuned[key] = {L_out, wg_out};
        return 0;
    }
    Stream st;
    HIP_CHECK(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
    // stage 1 holds several rounds even at 4 lanes per trajectory (16 trajectories per wavefront)
    const int64_t nset[4] = {std::min<int64_t>(n2, std::max<int64_t>(16384, (int64_t)M->dev.cus * 256)), n2,
                             std::min<int64_t>(n2, (int64_t)M->dev.cus * 16), n2};
    TrajBufs B;
    if (int rc = B.alloc(std::max(nset[1], nset[0]))) return rc;
    int cur = -1;
    int32_t hits = 0;
    auto run = [&](int set, int L, int W, double &ms) -> int {
        if (set != cur) {
            float sms = 0;
            if (int rc = sample_into(M, 0x5eed5eedULL, 0xfffffffeu - set, nset[set], nset[set], B, st.s, &sms))
                return rc;
            HIP_CHECK(hipMemcpy(&hits, B.counters.p, sizeof(int32_t), hipMemcpyDeviceToHost));
            if (kOrderWork)  // the order cosims_run launches in
                if (int rc = order_items(B.items[0].as<int32_t>(), B.items[1].as<int32_t>(), nullptr, nullptr, hits,
                                         B.bin.as<uint8_t>(), B.order.as<int32_t>(), st.s))
                    return rc;
            cur = set;
        }
        GroupRun R;
        if (int rc = run_group(M, prec, hits, B, st.s, L, W, 0, R,
                               kOrderWork ? B.items[1].as<int32_t>() : nullptr))
            return rc;
        ms = R.ms;
        return 0;
    };
    if (int rc = tune_shapes(group_units(M->prep), run, L_out, wg_out, clustered_type(M->prep.force_type)))
        return rc;
    M->tuned[key] = {L_out, wg_out};
    return 0;
}

}  // namespace

namespace {

// ------------------------------------------------------------- batches ---
constexpr int kMaxSegs = 64;        // launch shapes of one batch (force type x lanes x workgroup size)
constexpr int kMaxLaunches = 8192;  // kernel launches of one batch (own launches + the table runs between them)
constexpr int kMaxOwnLaunches = 2048;         // jobs with a launch of their own, per batch
constexpr double kOwnLaunchWork = 65536.0;    // trajectories x atoms below which a job shares a table launch
constexpr int kSegStreams = 4;      // streams the launches of a batch are spread over

// Lanes per trajectory from the units one force evaluation sweeps (atoms,
// or clusters): as many as leave each lane at least 48 units per step, from
// 4 to 64.  Size class k = the molecules this gives 4 * 2^k lanes.
int lanes_rule(int units)
{
    int L = 4;
    while (L < 64 && units >= 48 * L) L *= 2;
    return L;
}
int size_class(int units)
{
    int k = 0;
    for (int L = lanes_rule(units); L > 4; L >>= 1) k++;
    return k;
}

// The clustered force types take the 64-lane size class (clustered_type).
int job_class(const Prepared &p)
{
    return clustered_type(p.force_type) ? COSIMS_SIZE_CLASSES - 1 : size_class(group_units(p));
}
// Lanes of a job without a fixed lane count: 64 for the clustered types,
// else its class's tuned lanes, else the size rule.
int job_lanes(const Prepared &p, const int *class_lanes)
{
    if (clustered_type(p.force_type)) return 64;
    const int c = size_class(group_units(p));
    return class_lanes[c] > 0 ? class_lanes[c] : lanes_rule(group_units(p));
}

// A batch laid out on the device: per-trajectory buffers of all jobs back
// to back, the molecule of each trajectory, the MolRec table, and the work
// list ordered by launch (force type, lanes, workgroup size), larger
// molecules first within a launch.
struct BatchPlan {
    struct Job {
        const cosims_mol *mol;
        uint64_t seed;
        uint32_t job;
        int iters;
        int64_t T, n, off;  // trajectories per integral, total, first slot
        int ft, units;
        uint32_t it0;       // Philox iteration of the first integral
        int L, W;           // launch shape of this job's trajectories
    };
    std::vector<Job> jobs;
    std::vector<int> order;              // job indices in work-list order
    struct Seg {
        int ft, L, W;
        int i0, i1;                      // its jobs: work-list positions order[i0 .. i1)
        int64_t traj;                    // trajectories of its jobs (their hits: on the device, job_end)
    };
    std::vector<Seg> segs;
    int64_t n = 0;
    int cus = 0;
    int nlaunch = 0;                     // launches of the last launch_enqueue
    std::vector<int> launch_i0;          // first work-list position of each of them
    bool shared = false;                 // every job of a launch shape in one persistent launch
    TrajBufs B;
    DBuf tmol, table, seg_ctr, seg_evals, job_end, order_scr;
    // shared launches with the molecule in LDS: job of each work-list position, per-position
    // refill counters and sort scratch, the first position of each workgroup (host copy kept
    // for the asynchronous upload)
    DBuf pos_job, pos_ctr, pos_scr, wg_pos;
    int64_t pos_cap = 0, wg_cap = 0;
    std::vector<int32_t> wg_pos_host;
    int64_t cap = 0, table_cap = 0, job_end_cap = 0;  // allocated trajectories / table entries / bounds (reused)
    int order_cap = 0;                   // launches order_scr holds work-list sort scratch for
    float sample_ms = 0;
    std::vector<MolRec> recs;            // host side of the table upload
    std::vector<hipStream_t> streams;    // default priority, made on first use
    hipEvent_t ev0 = nullptr, ev1 = nullptr;  // first launch start / last launch end
    hipEvent_t es0 = nullptr, es1 = nullptr;  // sampling start / end
    std::vector<hipEvent_t> joins;
    BatchPlan() = default;
    BatchPlan(const BatchPlan &) = delete;
    BatchPlan &operator=(const BatchPlan &) = delete;
    ~BatchPlan()
    {
        for (hipEvent_t j : joins) (void)hipEventDestroy(j);
        for (hipEvent_t e : {ev0, ev1, es0, es1})
            if (e) (void)hipEventDestroy(e);
        for (hipStream_t x : streams) (void)hipStreamDestroy(x);
    }
};

// Work-list order: by launch shape, larger molecules first.
int order_plan(BatchPlan &P)
{
    const int nj = (int)P.jobs.size();
    P.cus = P.jobs[0].mol->dev.cus;
    for (const auto &J : P.jobs)
        if (J.mol->dev.id != P.jobs[0].mol->dev.id)
            return fail(COSIMS_E_INVALID, "batch: molecules on different devices");
    P.order.resize(nj);
    for (int k = 0; k < nj; k++) P.order[k] = k;
    std::stable_sort(P.order.begin(), P.order.end(), [&](int a, int b) {
        const auto &x = P.jobs[a], &y = P.jobs[b];
        if (x.ft != y.ft) return x.ft < y.ft;
        if (x.L != y.L) return x.L < y.L;
        if (x.W != y.W) return x.W < y.W;
        return x.units > y.units;
    });
    return 0;
}

// Launch shape of one job: a fixed lane count, the size rule, or the tuned
// table of its size class.
struct ShapeTable {
    int lanes = 0;                                   // > 0: every job
    int wg = 512;
    int class_lanes[COSIMS_SIZE_CLASSES] = {0};      // used when lanes == 0 and non-zero
    int class_wg[COSIMS_SIZE_CLASSES] = {0};
    void apply(BatchPlan::Job &b) const
    {
        const int c = job_class(b.mol->prep);
        b.L = lanes > 0 ? lanes : job_lanes(b.mol->prep, class_lanes);
        b.W = (lanes <= 0 && class_wg[c]) ? class_wg[c] : wg;
        if (b.W % b.L) b.W = std::max(b.W, b.L);
    }
};

// Validates the jobs and computes the layout (host only).
int plan_batch(const cosims_batch_job *jobs, int32_t njobs, int64_t cap_per_job, const ShapeTable &shape,
               BatchPlan &P)
{
    if (!jobs || njobs <= 0) return fail(COSIMS_E_INVALID, "cosims_run_batch: no jobs");
    P.jobs.clear();
    P.n = 0;
    for (int k = 0; k < njobs; k++) {
        const cosims_batch_job &J = jobs[k];
        if (!J.mol) return fail(COSIMS_E_INVALID, "cosims_run_batch: job without molecule");
        if (int rc = ensure_device(J.mol)) return rc;
        BatchPlan::Job b{};
        b.mol = J.mol;
        b.seed = J.seed;
        b.job = J.job;
        b.iters = J.iter > 0 ? J.iter : J.mol->params.iter;
        b.T = J.traj > 0 ? J.traj : J.mol->params.traj;
        if (b.iters <= 0 || b.T <= 0) return fail(COSIMS_E_INVALID, "cosims_run_batch: job without trajectories");
        if (cap_per_job > 0) {  // tuning subsets: one integral of at most cap trajectories
            b.iters = 1;
            b.T = std::min<int64_t>(b.T, cap_per_job);
        }
        b.n = (int64_t)b.iters * b.T;
        b.off = P.n;
        b.ft = J.mol->prep.force_type;
        b.units = group_units(J.mol->prep);
        shape.apply(b);
        P.n += b.n;
        if (P.n > INT32_MAX / 4) return fail(COSIMS_E_INVALID, "cosims_run_batch: too many trajectories for one batch");
        P.jobs.push_back(b);
    }
    return order_plan(P);
}

// Device buffers, molecule table, and the sampled work list of a plan.
int stage_batch(BatchPlan &P, hipStream_t s)
{
    const int nj = (int)P.jobs.size();
    if (P.cap < P.n) {
        if (int rc = P.B.alloc(P.n)) return rc;
        HIP_CHECK(P.tmol.alloc(sizeof(int32_t) * P.n));
        P.cap = P.n;
    }
    if (P.table_cap < nj) {
        HIP_CHECK(P.table.alloc(sizeof(MolRec) * nj));
        P.table_cap = nj;
    }
    if (P.job_end_cap < nj) {
        HIP_CHECK(P.job_end.alloc(sizeof(int32_t) * nj));
        P.job_end_cap = nj;
    }
    // the host copy lives on the plan: the upload is asynchronous
    P.recs.resize(nj);
    for (int k = 0; k < nj; k++) P.recs[k] = P.jobs[k].mol->rec;
    HIP_CHECK(hipMemcpyAsync(P.table.p, P.recs.data(), sizeof(MolRec) * nj, hipMemcpyHostToDevice, s));
    if (P.pos_cap < nj) {
        HIP_CHECK(P.pos_job.alloc(sizeof(int32_t) * nj));
        HIP_CHECK(P.pos_ctr.alloc(sizeof(int32_t) * nj));
        HIP_CHECK(P.pos_scr.alloc(sizeof(int32_t) * 2 * kOrderBins * nj));
        P.pos_cap = nj;
    }
    // job of each work-list position (P.order lives on the plan: the upload is asynchronous)
    HIP_CHECK(hipMemcpyAsync(P.pos_job.p, P.order.data(), sizeof(int32_t) * nj, hipMemcpyHostToDevice, s));
    int32_t *ctr = P.B.counters.as<int32_t>();
    if (!P.es0) {
        HIP_CHECK(hipEventCreate(&P.es0));
        HIP_CHECK(hipEventCreate(&P.es1));
    }
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, ctr, 64, 0);
    HIP_CHECK(hipEventRecord(P.es0, s));
    P.segs.clear();
    int64_t seg_traj = 0;
    int seg_i0 = 0;
    for (size_t i = 0; i < P.order.size(); i++) {
        const int k = P.order[i];
        const BatchPlan::Job &J = P.jobs[k];
        const unsigned nb = (unsigned)((J.n + 255) / 256);
        hipLaunchKernelGGL(k_fill_i32, dim3(nb), dim3(256), 0, s, P.tmol.as<int32_t>() + J.off, J.n, (int32_t)k);
        SampleArgs A = sample_args(J.mol, J.seed, J.job);
        A.first = 0; A.n = J.n; A.per_iter = J.T; A.iteration0 = J.it0;
        A.p0 = P.B.p0.as<double>() + 3 * J.off; A.v0 = P.B.v0.as<double>() + 3 * J.off;
        A.g = P.B.g.as<double>() + J.off; A.status = P.B.status.as<int8_t>() + J.off;
        A.items = P.B.items[0].as<int32_t>(); A.n_items = ctr; A.item_base = (int32_t)J.off;
        A.bin = P.B.bin.as<uint8_t>() + J.off; A.dt0 = J.mol->params.dt;
        hipLaunchKernelGGL(k_sample, dim3(nb), dim3(256), 0, s, A);
        HIP_CHECK(hipGetLastError());
        // this job's hits end where the item counter stands now; the bound
        // stays on the device (no host wait here)
        hipLaunchKernelGGL(k_copy_i32, dim3(1), dim3(1), 0, s, ctr, P.job_end.as<int32_t>() + i);
        seg_traj += J.n;
        const bool last = i + 1 == P.order.size();
        const BatchPlan::Job *nx = last ? nullptr : &P.jobs[P.order[i + 1]];
        if (last || nx->ft != J.ft || nx->L != J.L || nx->W != J.W) {
            if ((int)P.segs.size() == kMaxSegs) return fail(COSIMS_E_INVALID, "batch: too many distinct launch shapes");
            P.segs.push_back({J.ft, J.L, J.W, seg_i0, (int)i + 1, seg_traj});
            seg_traj = 0;
            seg_i0 = (int)i + 1;
        }
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(P.es1, s));
    return 0;
}

// The launches of a staged plan.
//   Own launches (the default, P.shared false): a job of an exact force type
//   runs one launch of its own, the LDS-staged kernel when its molecule fits
//   one workgroup's LDS, the tiled kernel beyond; the clustered jobs of a
//   launch shape, tiny jobs and jobs beyond kMaxOwnLaunches, in consecutive
//   runs of the work list, share table-kernel launches.
//   Shared launches (P.shared): the jobs of a launch shape whose molecule
//   fits LDS share one persistent k_traj_lds_multi launch (each workgroup
//   stages the molecule of the job it works on, and moves to the job with the
//   most trajectories left when its job's range is spent); larger exact-type
//   jobs keep tiled launches of their own; clustered jobs share table-kernel
//   launches.
// All kernels give the same bits for a trajectory.  Launches are spread over
// kSegStreams streams, the most work first, so one launch's tail overlaps the
// next ones.  launch_enqueue only enqueues (the stream s carries the join;
// kernels run on the plan's own streams); launch_finish waits for them and
// adds their time (first launch to last one's end) and counters to tot.
int launch_enqueue(BatchPlan &P, int prec, int bpc, hipStream_t s, int &launches)
{
    launches = 0;
    enum Kind { OWN, TABLE, LDS_MULTI };
    struct Launch {
        int seg, i0, i1;  // work-list positions order[i0 .. i1)
        Kind kind;
        int64_t traj;
        double work;
        int64_t max_n;    // most trajectories of one of its jobs
    };
    std::vector<Launch> plan;
    int own = 0;
    for (int g = 0; g < (int)P.segs.size(); g++) {
        const auto &sg = P.segs[g];
        int run0 = -1;
        Kind run_kind = TABLE;
        int64_t run_traj = 0, run_max = 0;
        double run_work = 0;
        for (int i = sg.i0; i <= sg.i1; i++) {
            const BatchPlan::Job *J = i < sg.i1 ? &P.jobs[P.order[i]] : nullptr;
            Kind k = TABLE;
            bool mine = false;
            if (J) {
                const bool exact = group_kernel(prec, J->ft, J->L) != nullptr;
                if (P.shared) {
                    if (exact && lds_mode(J->mol, prec)) k = LDS_MULTI;
                    else if (exact && tile_mode(J->mol, prec)) mine = true;
                } else {
                    // a launch of its own: the LDS-staged kernel, or the tiled one beyond LDS
                    // (tiny jobs, and jobs past the per-batch count, share table launches)
                    mine = (lds_mode(J->mol, prec) || tile_mode(J->mol, prec)) && exact &&
                           (double)J->n * J->units >= kOwnLaunchWork && own < kMaxOwnLaunches;
                }
                own += mine;
            }
            if (run0 >= 0 && (!J || mine || k != run_kind)) {  // close the run of shared jobs
                plan.push_back({g, run0, i, run_kind, run_traj, run_work, run_max});
                run0 = -1;
                run_traj = 0;
                run_work = 0;
                run_max = 0;
            }
            if (!J) break;
            if (mine) {
                plan.push_back({g, i, i + 1, OWN, J->n, (double)J->n * J->units, J->n});
            } else {
                if (run0 < 0) {
                    run0 = i;
                    run_kind = k;
                }
                run_traj += J->n;
                run_work += (double)J->n * J->units;
                run_max = std::max(run_max, J->n);
            }
        }
    }
    const int nl = (int)plan.size();
    if (nl > kMaxLaunches) return fail(COSIMS_E_INVALID, "batch: too many launches");
    P.nlaunch = nl;
    P.launch_i0.resize(nl);
    for (int li = 0; li < nl; li++) P.launch_i0[li] = plan[li].i0;
    if (!P.seg_ctr.p) {
        HIP_CHECK(P.seg_ctr.alloc(sizeof(int32_t) * kMaxLaunches));
        HIP_CHECK(P.seg_evals.alloc(sizeof(unsigned long long) * kEvalSlots * kMaxLaunches));
    }
    if (kOrderWork && nl > P.order_cap) {
        HIP_CHECK(P.order_scr.alloc(sizeof(int32_t) * 2 * kOrderBins * nl));
        P.order_cap = nl;
    }
    while ((int)P.streams.size() < std::min(nl, kSegStreams)) {
        hipStream_t x;
        HIP_CHECK(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
        P.streams.push_back(x);
    }
    if (!P.ev0) {
        HIP_CHECK(hipEventCreate(&P.ev0));
        HIP_CHECK(hipEventCreate(&P.ev1));
    }
    for (hipEvent_t j : P.joins) (void)hipEventDestroy(j);
    P.joins.clear();
    if (nl > 0) {
        const int nc = std::max(1, (nl + 63) / 64);
        hipLaunchKernelGGL(k_fill_i32, dim3(nc), dim3(64), 0, s, P.seg_ctr.as<int32_t>(), (int64_t)nl, 0);
        hipLaunchKernelGGL(k_fill_i32, dim3(2 * kEvalSlots * nc), dim3(64), 0, s, P.seg_evals.as<int32_t>(),
                           (int64_t)2 * kEvalSlots * nl, 0);
    }
    std::vector<int> order(nl);
    for (int i = 0; i < nl; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return plan[a].work > plan[b].work; });
    const int32_t *ends = P.job_end.as<int32_t>();
    const int nj = (int)P.jobs.size();
    bool any_multi = false;
    for (const Launch &l : plan) any_multi |= l.kind == LDS_MULTI;
    if (any_multi) {
        const int64_t nb = (int64_t)nj + (int64_t)2 * kOrderBins * nj;
        hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((nj + 255) / 256)), dim3(256), 0, s, P.pos_ctr.as<int32_t>(),
                           (int64_t)nj, 0);
        hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((nb - nj + 255) / 256)), dim3(256), 0, s,
                           P.pos_scr.as<int32_t>(), nb - nj, 0);
    }
    // every launch's range of the work list, longest expected trajectories
    // first, into items[1], before the launches start: the small sort
    // kernels would otherwise wait for CU slots behind running launches.
    // Shared LDS launches sort each job's range on its own (a workgroup
    // drains one job at a time).
    if (kOrderWork && nl > 0) {
        const int64_t ns = (int64_t)2 * kOrderBins * nl;
        hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, s, P.order_scr.as<int32_t>(),
                           ns, 0);
        for (int li = 0; li < nl; li++) {
            const Launch &l = plan[li];
            int rc = 0;
            if (l.kind == LDS_MULTI)
                rc = order_positions(P.B.items[0].as<int32_t>(), P.B.items[1].as<int32_t>(), ends, l.i0, l.i1 - l.i0,
                                     l.max_n, P.B.bin.as<uint8_t>(), P.pos_scr.as<int32_t>() + 2 * kOrderBins * l.i0,
                                     s);
            else
                rc = order_items(P.B.items[0].as<int32_t>(), P.B.items[1].as<int32_t>(),
                                 l.i0 ? ends + l.i0 - 1 : nullptr, ends + l.i1 - 1, l.traj, P.B.bin.as<uint8_t>(),
                                 P.order_scr.as<int32_t>() + 2 * kOrderBins * li, s, false);
            if (rc) return rc;
        }
    }
    // shared LDS launches: grids and each workgroup's first job, in proportion to the jobs'
    // work (trajectories x atoms), so that they end about together and few workgroups move
    struct MultiShape {
        LdsMultiFn fn = nullptr;
        size_t lds = 0;
        int64_t grid = 0, wg0 = 0;
    };
    std::vector<MultiShape> ms(nl);
    int64_t wg_total = 0;
    for (int li = 0; li < nl; li++) {
        const Launch &l = plan[li];
        if (l.kind != LDS_MULTI) continue;
        const auto &g = P.segs[l.seg];
        MultiShape &m = ms[li];
        m.fn = lds_multi_kernel(prec, g.ft, g.L);
        if (!m.fn) return fail(COSIMS_E_INVALID, "no shared LDS kernel for this force type / lane count");
        int npad = 0;
        for (int i = l.i0; i < l.i1; i++) npad = std::max(npad, P.jobs[P.order[i]].mol->npad);
        m.lds = kCtlBytes + (size_t)npad * (prec == COSIMS_PREC_FP64 ? 6 * sizeof(double)
                                                                      : sizeof(float4) + sizeof(float2));
        HIP_CHECK(hipFuncSetAttribute((const void *)m.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)m.lds));
        m.grid = group_grid((const void *)m.fn, g.W, m.lds, bpc, P.cus, l.traj, g.L);
        m.wg0 = wg_total;
        wg_total += m.grid;
    }
    if (wg_total > 0) {
        if (P.wg_cap < wg_total) {
            HIP_CHECK(P.wg_pos.alloc(sizeof(int32_t) * wg_total));
            P.wg_cap = wg_total;
        }
        P.wg_pos_host.assign((size_t)wg_total, 0);
        for (int li = 0; li < nl; li++) {
            const Launch &l = plan[li];
            if (l.kind != LDS_MULTI) continue;
            const MultiShape &m = ms[li];
            int p = l.i0;
            double cum = 0, cum_p = (double)P.jobs[P.order[p]].n * P.jobs[P.order[p]].units;
            for (int64_t w = 0; w < m.grid; w++) {
                const double target = ((double)w + 0.5) / (double)m.grid * l.work;
                while (p + 1 < l.i1 && cum_p <= target) {
                    cum = cum_p;
                    p++;
                    cum_p = cum + (double)P.jobs[P.order[p]].n * P.jobs[P.order[p]].units;
                }
                P.wg_pos_host[(size_t)(m.wg0 + w)] = p;
            }
        }
        HIP_CHECK(hipMemcpyAsync(P.wg_pos.p, P.wg_pos_host.data(), sizeof(int32_t) * wg_total, hipMemcpyHostToDevice,
                                 s));
    }
    HIP_CHECK(hipEventRecord(P.ev0, s));
    for (int q = 0; q < nl; q++) {
        const int li = order[q];
        const Launch &l = plan[li];
        const auto &g = P.segs[l.seg];
        const int cnt = (int)l.traj;  // upper bound of the launch's work list (its hits)
        if (cnt <= 0) continue;
        hipStream_t x = P.streams[launches % P.streams.size()];
        HIP_CHECK(hipStreamWaitEvent(x, P.ev0, 0));
        const int32_t *first = l.i0 ? ends + l.i0 - 1 : nullptr;
        const int32_t *last = ends + l.i1 - 1;
        TrajIO io = traj_io(P.B, (kOrderWork ? P.B.items[1] : P.B.items[0]).as<int32_t>(), cnt);
        io.counter = P.seg_ctr.as<int32_t>() + li;
        io.evals = P.seg_evals.as<unsigned long long>() + kEvalSlots * li;
        if (l.kind == OWN) {
            const cosims_mol *M = P.jobs[P.order[l.i0]].mol;
            const bool tiled = tile_mode(M, prec);
            GrpFn fn = tiled ? tile_kernel(prec, g.ft, g.L) : group_kernel(prec, g.ft, g.L);
            GroupArgs G = group_args(M);
            size_t lds = group_lds_bytes(M, prec);
            if (tiled) {
                G.tile = tile_atoms(M, prec, g.W);
                lds = tile_lds_bytes(prec, G.tile);
            }
            HIP_CHECK(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            G.io = io;
            G.first = first;
            G.last = last;
            const int64_t grid = group_grid((const void *)fn, g.W, lds, bpc, P.cus, cnt, g.L);
            hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(g.W), lds, x, G);
        } else if (l.kind == LDS_MULTI) {
            const MultiShape &m = ms[li];
            LdsMultiArgs G{};
            G.io = io;  // items: the whole sorted work list; each job's range from ends
            G.mols = P.table.as<MolRec>();
            G.pos_job = P.pos_job.as<int32_t>();
            G.ends = ends;
            G.pos_ctr = P.pos_ctr.as<int32_t>();
            G.wg_pos = P.wg_pos.as<int32_t>() + m.wg0;
            G.pos0 = l.i0;
            G.npos = l.i1 - l.i0;
            hipLaunchKernelGGL(m.fn, dim3((unsigned)m.grid), dim3(g.W), m.lds, x, G);
        } else {
            MultiFn fn = multi_kernel(prec, g.ft, g.L);
            if (!fn) return fail(COSIMS_E_INVALID, "no table kernel for this force type / lane count");
            MultiArgs G{};
            G.io = io;
            G.mols = P.table.as<MolRec>();
            G.tmol = P.tmol.as<int32_t>();
            G.first = first;
            G.last = last;
            const size_t lds = multi_lds_bytes(g.ft, g.L, g.W);
            const int64_t grid = group_grid((const void *)fn, g.W, lds, bpc, P.cus, cnt, g.L);
            hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(g.W), lds, x, G);
        }
        HIP_CHECK(hipGetLastError());
        // no acos pass here: k_reduce takes chi from the stored cosine (ReduceArgs::cchi)
        hipEvent_t j;
        HIP_CHECK(hipEventCreateWithFlags(&j, hipEventDisableTiming));
        P.joins.push_back(j);
        HIP_CHECK(hipEventRecord(j, x));
        HIP_CHECK(hipStreamWaitEvent(s, j, 0));
        launches++;
    }
    HIP_CHECK(hipEventRecord(P.ev1, s));
    return 0;
}

int launch_finish(BatchPlan &P, hipStream_t s, GroupRun &tot)
{
    const int nseg = P.nlaunch;
    std::vector<unsigned long long> ev(kEvalSlots * (size_t)std::max(1, nseg));
    if (nseg > 0)
        HIP_CHECK(hipMemcpyAsync(ev.data(), P.seg_evals.p, sizeof(unsigned long long) * kEvalSlots * nseg,
                                 hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, P.ev0, P.ev1));
    tot.ms += ms;
    HIP_CHECK(hipEventElapsedTime(&P.sample_ms, P.es0, P.es1));
    for (int i = 0; i < nseg; i++) {
        const unsigned long long *e = ev.data() + (size_t)kEvalSlots * i;
        int64_t cl[6];
        for (int k = 0; k < 6; k++) {
            cl[k] = (int64_t)e[4 + k];
            tot.cl[k] += cl[k];
        }
        tot.steps += (int64_t)e[0];
        tot.evals += (int64_t)e[1];
        tot.pots += (int64_t)e[2];
        tot.pairs += (int64_t)e[3];
        const BatchPlan::Job &J = P.jobs[P.order[P.launch_i0[i]]];
        tot.flops += ref_flops(J.ft, J.mol->prep.multipole_order, (int64_t)e[1], (int64_t)e[0], (int64_t)e[2],
                               (int64_t)e[3], cl);
    }
    return 0;
}

int launch_batch(BatchPlan &P, int prec, int bpc, hipStream_t s, GroupRun &tot, int &launches)
{
    if (int rc = launch_enqueue(P, prec, bpc, s, launches)) {
        (void)hipStreamSynchronize(s);
        return rc;
    }
    return launch_finish(P, s, tot);
}

// Size-class autotuner of a batch: for every size class present, the
// lanes x workgroup search of tune_shapes on that class's jobs alone.
int tune_batch(const cosims_batch_job *jobs, int32_t njobs, int prec, ShapeTable &out, bool shared = true)
{
    if (!jobs || njobs <= 0) return fail(COSIMS_E_INVALID, "cosims_tune_batch: no jobs");
    Stream st;
    HIP_CHECK(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
    std::vector<cosims_batch_job> by_class[COSIMS_SIZE_CLASSES];
    for (int k = 0; k < njobs; k++) {
        if (!jobs[k].mol) return fail(COSIMS_E_INVALID, "cosims_tune_batch: job without molecule");
        by_class[job_class(jobs[k].mol->prep)].push_back(jobs[k]);
    }
    out.lanes = 0;
    for (int c = 0; c < COSIMS_SIZE_CLASSES; c++) {
        if (by_class[c].empty()) continue;
        // the class's exact-type jobs decide its shape (its clustered jobs run 64 lanes
        // whatever the table says); a class of clustered jobs alone tunes W at 64 lanes
        std::vector<cosims_batch_job> exact;
        for (const auto &j : by_class[c])
            if (!clustered_type(j.mol->prep.force_type)) exact.push_back(j);
        const bool only64 = exact.empty();
        const std::vector<cosims_batch_job> &pool = only64 ? by_class[c] : exact;
        int units = 0;  // largest molecule of the class: the lane counts worth trying
        for (const auto &j : pool) units = std::max(units, group_units(j.mol->prep));
        // tuning subsets: at most 8 jobs spread over the class, each with 1/8 (search), 1/2
        // (finalists) and 1/32 (screen) of the fewest trajectories a job of the subset has.
        // Jobs that fit LDS run one launch each, so the sets must be launch-sized like the
        // batch's: on small sets the end-of-launch tail would decide the shape.
        const size_t m = pool.size(), take = std::min<size_t>(8, m);
        std::vector<cosims_batch_job> sub;
        int64_t T = INT64_MAX;
        for (size_t i = 0; i < take; i++) {
            const cosims_batch_job &j = pool[i * m / take];
            sub.push_back(j);
            const int64_t it = j.iter > 0 ? j.iter : j.mol->params.iter;
            const int64_t tr = j.traj > 0 ? j.traj : j.mol->params.traj;
            T = std::min<int64_t>(T, std::max<int64_t>(1, it) * std::max<int64_t>(1, tr));
        }
        auto cap = [&](int64_t d) { return std::min<int64_t>(65536, std::max<int64_t>(256, T / d)); };
        const int64_t caps[4] = {cap(8), cap(2), cap(32), cap(2)};
        BatchPlan set;
        auto run = [&](int k, int L, int W, double &ms) -> int {
            ShapeTable fixed;
            fixed.lanes = L;
            fixed.wg = W;
            if (int rc = plan_batch(sub.data(), (int32_t)sub.size(), caps[k], fixed, set)) return rc;
            set.shared = shared;
            for (auto &J : set.jobs) J.job = 0xfffffff0u - (uint32_t)k;  // a stream of its own
            if (int rc = stage_batch(set, st.s)) return rc;
            GroupRun R;
            int launches = 0;
            if (int rc = launch_batch(set, prec, 0, st.s, R, launches)) return rc;
            ms = R.ms;
            return 0;
        };
        int L = 0, W = 0;
        if (int rc = tune_shapes(units, run, L, W, only64)) return rc;
        out.class_lanes[c] = L;
        out.class_wg[c] = W;
    }
    return 0;
}

}  // namespace

// ------------------------------------------------ allocation-free runs ---
// Device scratch of cosims_run_noalloc: the group-kernel path's per-trajectory
// buffers (102 bytes a trajectory: initial conditions, order buckets, status
// and outcome, work lists, status codes out) plus the integrals' sums and
// CCS, for at most COSIMS_NOALLOC_MAX_TRAJ trajectories in at most
// COSIMS_NOALLOC_MAX_ITER integrals: 27.3 MB on every device the library runs
// on, held until the process ends.  In host code `scratch` names a host-side
// stand-in: kernels and copies take the device address hipGetSymbolAddress
// returns for the current device (scratch_of).
constexpr size_t kNoAllocBytes = (size_t)COSIMS_NOALLOC_MAX_TRAJ * 104 + 64 * 1024;
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

namespace {

constexpr int kMaxDevices = 64;

struct ScratchDev {
    std::once_flag once;
    hipError_t rc = hipSuccess;
    uint8_t *ptr = nullptr;  // device address of this device's copy of scratch
    std::mutex mu;           // calls on one device share its copy: they run one at a time
};
ScratchDev g_scratch[kMaxDevices];

// The scratch of device dev (the current device), its address looked up
// once, on the first call on that device.
int scratch_of(int dev, ScratchDev *&out)
{
    if (dev < 0 || dev >= kMaxDevices) return fail(COSIMS_E_INVALID, "cosims_run_noalloc: device index too large");
    ScratchDev &S = g_scratch[dev];
    std::call_once(S.once, [&S] {
        void *p = nullptr;
        S.rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        S.ptr = S.rc == hipSuccess ? (uint8_t *)p : nullptr;
    });
    if (S.rc != hipSuccess || !S.ptr)
        return fail(COSIMS_E_HIP, std::string("cosims_run_noalloc: hipGetSymbolAddress(scratch): ") +
                                      hipGetErrorString(S.rc));
    out = &S;
    return 0;
}

// Trajectory buffers and stream cosims_run keeps between calls on a device
// (sized for the largest run so far): a call takes them when no other call
// holds them, else allocates its own.  Every buffer a run reads it writes
// first, so nothing of the previous run shows through.
struct RunPool {
    std::mutex mu;
    bool busy = false;
    TrajBufs B;
    int64_t cap = 0;               // trajectories B holds
    DBuf aux;                      // per-integral sums, CCS and error
    int cap_iters = 0;
    Stream st;
};
// never destroyed: a static destructor's HIP calls would run after the
// runtime has shut down at process exit
RunPool *const g_run_pool = new RunPool[kMaxDevices];

struct RunLease {
    RunPool *pool = nullptr;       // the device's pool, or
    TrajBufs own;                  // buffers of this call alone
    DBuf own_aux;
    Stream own_st;
    TrajBufs *B = nullptr;
    double *aux = nullptr;
    hipStream_t s = nullptr;
    RunLease() = default;
    RunLease(const RunLease &) = delete;
    ~RunLease()
    {
        if (pool) {
            std::lock_guard<std::mutex> lk(pool->mu);
            pool->busy = false;
        }
    }
};

// Buffers for n trajectories in iters integrals on device dev (current).
int lease_run(int dev, int64_t n, int iters, RunLease &L)
{
    RunPool *P = (dev >= 0 && dev < kMaxDevices) ? g_run_pool + dev : nullptr;
    if (P) {
        std::lock_guard<std::mutex> lk(P->mu);
        if (!P->busy) {
            P->busy = true;
            L.pool = P;
        }
    }
    const size_t aux_bytes = sizeof(double) * (kSums + 2) * (size_t)iters;
    if (L.pool) {
        if (P->cap < n) {
            P->cap = 0;
            if (int rc = P->B.alloc(n)) return rc;
            P->cap = n;
        }
        if (P->cap_iters < iters) {
            P->cap_iters = 0;
            HIP_CHECK(P->aux.alloc(aux_bytes));
            P->cap_iters = iters;
        }
        if (!P->st.s) HIP_CHECK(hipStreamCreateWithFlags(&P->st.s, hipStreamNonBlocking));
        L.B = &P->B;
        L.aux = P->aux.as<double>();
        L.s = P->st.s;
        return 0;
    }
    if (int rc = L.own.alloc(n)) return rc;
    HIP_CHECK(L.own_aux.alloc(aux_bytes));
    HIP_CHECK(hipStreamCreateWithFlags(&L.own_st.s, hipStreamNonBlocking));
    L.B = &L.own;
    L.aux = L.own_aux.as<double>();
    L.s = L.own_st.s;
    return 0;
}

}  // namespace

extern "C" {

int cosims_sample(const cosims_mol *M, uint64_t seed, uint32_t job, uint32_t iteration, int64_t first, int64_t n,
                  double *p0, double *v0, double *g, int32_t *hit)
{
    if (!M || n < 0 || first < 0 || first + n > (int64_t)UINT32_MAX) return fail(COSIMS_E_INVALID, "cosims_sample: bad range");
    if (n == 0) return 0;
    if (int rc = ensure_device(M)) return rc;
    Stream st;
    HIP_CHECK(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
    DBuf dp, dv, dg, dh;
    HIP_CHECK(dp.alloc(sizeof(double) * 3 * n));
    HIP_CHECK(dv.alloc(sizeof(double) * 3 * n));
    HIP_CHECK(dg.alloc(sizeof(double) * n));
    HIP_CHECK(dh.alloc(sizeof(int32_t) * n));
    SampleArgs A = sample_args(M, seed, job);
    A.iteration0 = iteration; A.first = first; A.n = n; A.per_iter = 0;
    A.p0 = dp.as<double>(); A.v0 = dv.as<double>(); A.g = dg.as<double>(); A.hit = dh.as<int32_t>();
    hipLaunchKernelGGL(k_sample, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st.s, A);
    HIP_CHECK(hipGetLastError());
    if (p0) HIP_CHECK(hipMemcpyAsync(p0, dp.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, st.s));
    if (v0) HIP_CHECK(hipMemcpyAsync(v0, dv.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, st.s));
    if (g) HIP_CHECK(hipMemcpyAsync(g, dg.p, sizeof(double) * n, hipMemcpyDeviceToHost, st.s));
    if (hit) HIP_CHECK(hipMemcpyAsync(hit, dh.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st.s));
    HIP_CHECK(hipStreamSynchronize(st.s));
    return 0;
}

int cosims_sample_reference(const cosims_mol *M, uint64_t seed, int32_t iters, int64_t traj, double *p0, double *v0,
                            double *g, int32_t *hit)
{
    if (!M || iters <= 0 || traj <= 0 || !p0 || !v0 || !g || !hit)
        return fail(COSIMS_E_INVALID, "cosims_sample_reference: bad arguments");
    sample_compat(M->prep, seed, iters, traj, p0, v0, g, hit);
    return 0;
}

int cosims_verlet(const cosims_mol *M, int32_t prec, int64_t n, const double *p0, const double *v0, double *chi,
                  int32_t *ok, int32_t *steps, double *ergy, int64_t *total_steps)
{
    if (!M || n < 0 || (n > 0 && (!p0 || !v0 || !chi || !ok || !steps || !ergy)))
        return fail(COSIMS_E_INVALID, "cosims_verlet: bad arguments");
    if (prec != COSIMS_PREC_FP64 && prec != COSIMS_PREC_MIXED) return fail(COSIMS_E_INVALID, "unknown precision");
    if (n > INT32_MAX / 4) return fail(COSIMS_E_INVALID, "cosims_verlet: batch too large");
    if (n == 0) return 0;
    if (int rc = ensure_device(M)) return rc;
    // FP64 runs the lane-per-trajectory kernel: atom-order sums, bit-identical
    // to the reference; mixed precision runs the tuned group kernel.
    int L = 1, wg = 256;
    if (prec == COSIMS_PREC_MIXED)
        if (int rc = autotune(M, prec, n, L, wg)) return rc;
    Stream st;
    HIP_CHECK(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
    TrajBufs B;
    int rc = B.alloc(n);
    if (rc) return rc;
    HIP_CHECK(hipMemcpyAsync(B.p0.p, p0, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st.s));
    HIP_CHECK(hipMemcpyAsync(B.v0.p, v0, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st.s));
    std::vector<int32_t> idx(n);
    for (int64_t i = 0; i < n; i++) idx[i] = (int32_t)i;
    HIP_CHECK(hipMemcpyAsync(B.items[0].p, idx.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st.s));
    int32_t cnt[3] = {(int32_t)n, 0, 0};
    HIP_CHECK(hipMemcpyAsync(B.counters.p, cnt, sizeof(cnt), hipMemcpyHostToDevice, st.s));
    std::vector<int8_t> s8(n, ST_PENDING);
    HIP_CHECK(hipMemcpyAsync(B.status.p, s8.data(), n, hipMemcpyHostToDevice, st.s));
    if (L == 1) {
        RoundStats R;
        rc = run_rounds(M, prec, n, B, st.s, 256, 0, R);
    } else {
        GroupRun G;
        rc = run_group(M, prec, (int)n, B, st.s, L, wg, 0, G);
    }
    if (rc) return rc;
    std::vector<int32_t> tot(n);
    HIP_CHECK(hipMemcpyAsync(chi, B.chi.p, sizeof(double) * n, hipMemcpyDeviceToHost, st.s));
    HIP_CHECK(hipMemcpyAsync(s8.data(), B.status.p, n, hipMemcpyDeviceToHost, st.s));
    HIP_CHECK(hipMemcpyAsync(steps, B.steps.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st.s));
    HIP_CHECK(hipMemcpyAsync(ergy, B.ergy.p, sizeof(double) * n, hipMemcpyDeviceToHost, st.s));
    HIP_CHECK(hipMemcpyAsync(tot.data(), B.tot.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st.s));
    HIP_CHECK(hipStreamSynchronize(st.s));
    for (int64_t i = 0; i < n; i++) {
        ok[i] = s8[i] == ST_OK ? 1 : 0;
        if (!ok[i]) ergy[i] = 0;
        if (total_steps) total_steps[i] = tot[i];
    }
    return 0;
}

int cosims_run(const cosims_mol *M, const cosims_params *P, const cosims_run_cfg *cfg, double *omega,
               double *omega_err, cosims_ccs *out)
{
    if (!M || !P || !omega || !omega_err || !out) return fail(COSIMS_E_INVALID, "cosims_run: null argument");
    cosims_run_cfg c{};
    if (cfg) c = *cfg;
    const int prec = c.prec;
    if (prec != COSIMS_PREC_FP64 && prec != COSIMS_PREC_MIXED) return fail(COSIMS_E_INVALID, "unknown precision");
    const int iters = c.iter > 0 ? c.iter : P->iter;
    const int64_t T = c.traj > 0 ? c.traj : P->traj;
    if (iters <= 0 || T <= 0) return fail(COSIMS_E_INVALID, "cosims_run: no trajectories");
    const int64_t n = (int64_t)iters * T;
    if (n > INT32_MAX / 4) return fail(COSIMS_E_INVALID, "cosims_run: too many trajectories for one batch");
    if (c.lanes != 0 && c.lanes != 1 && c.lanes != 4 && c.lanes != 8 && c.lanes != 16 && c.lanes != 32 &&
        c.lanes != 64)
        return fail(COSIMS_E_INVALID, "lanes must be 0 (tuned), 1, 4, 8, 16, 32 or 64");
    std::memset(out, 0, sizeof(*out));
    if (int rc = ensure_device(M)) return rc;
    // launch shape: explicit, or tuned
    int L = c.lanes, wg = c.wg_size;
    if (L == 0) {
        int tl = 0, tw = 0;
        if (int rc = autotune(M, prec, n, tl, tw)) return rc;
        L = tl;
        if (wg <= 0) wg = tw;
    }
    if (wg <= 0) wg = 256;
    if (wg % 64 || wg > (L == 1 ? 256 : 1024) || wg % L)
        return fail(COSIMS_E_INVALID, "wg_size must be a multiple of 64 (at most 256 lane-mode, 1024 group-mode)");
    auto t0 = std::chrono::steady_clock::now();
    int dev = -1;
    HIP_CHECK(hipGetDevice(&dev));
    RunLease lease;
    int rc = lease_run(dev, n, iters, lease);
    if (rc) return rc;
    TrajBufs &B = *lease.B;
    const hipStream_t s = lease.s;
    float sms = 0;
    if (c.rng == COSIMS_RNG_REFERENCE) {
        // the reference's stream, drawn on the host and uploaded
        std::vector<double> hp0(3 * n), hv0(3 * n), hg(n);
        std::vector<int32_t> hh(n), items;
        std::vector<int8_t> s8(n);
        sample_compat(M->prep, P->seed, iters, T, hp0.data(), hv0.data(), hg.data(), hh.data());
        items.reserve(n);
        for (int64_t i = 0; i < n; i++) {
            s8[i] = hh[i] ? ST_PENDING : ST_MISS;
            if (hh[i]) items.push_back((int32_t)i);
        }
        int32_t cnt[64] = {0};
        cnt[0] = (int32_t)items.size();
        HIP_CHECK(hipMemcpyAsync(B.p0.p, hp0.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(B.v0.p, hv0.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(B.g.p, hg.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(B.status.p, s8.data(), n, hipMemcpyHostToDevice, s));
        if (!items.empty())
            HIP_CHECK(hipMemcpyAsync(B.items[0].p, items.data(), sizeof(int32_t) * items.size(),
                                     hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(B.counters.p, cnt, sizeof(cnt), hipMemcpyHostToDevice, s));
        HIP_CHECK(hipStreamSynchronize(s));
    } else if (c.rng == COSIMS_RNG_PHILOX) {
        if ((rc = sample_into(M, P->seed, c.job, n, T, B, s, &sms))) return rc;
    } else {
        return fail(COSIMS_E_INVALID, "unknown rng");
    }
    int32_t hits = 0;
    HIP_CHECK(hipMemcpyAsync(&hits, B.counters.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (L == 1) {
        RoundStats R;
        rc = run_rounds(M, prec, n, B, s, wg, c.blocks_per_cu, R);
        if (rc) return rc;
        out->n_retry = R.retries;
        out->n_rounds = R.rounds;
        for (int r = 0; r < R.rounds; r++) {
            out->traj_kernel_ms[r] = R.traj_ms[r];
            out->round_items[r] = R.items[r];
            out->round_steps[r] = R.steps[r];
            out->round_evals[r] = R.evals[r];
        }
        out->check_kernel_ms = R.check_ms;
    } else if (hits > 0) {
        // longest expected trajectories first (the reference stream carries no order buckets)
        const int32_t *work = nullptr;
        if (kOrderWork && c.rng == COSIMS_RNG_PHILOX) {
            work = B.items[1].as<int32_t>();
            if ((rc = order_items(B.items[0].as<int32_t>(), B.items[1].as<int32_t>(), nullptr, nullptr, hits,
                                  B.bin.as<uint8_t>(), B.order.as<int32_t>(), s)))
                return rc;
        }
        GroupRun G;
        rc = run_group(M, prec, hits, B, s, L, wg, c.blocks_per_cu, G, work);
        if (rc) return rc;
        out->n_rounds = 1;
        out->traj_kernel_ms[0] = G.ms;
        out->round_items[0] = hits;
        out->round_steps[0] = G.steps;
        out->round_evals[0] = G.evals;
        out->pot_evals = G.pots;
        for (int k = 0; k < 6; k++) out->cl_counts[k] = G.cl[k];
        out->ref_flops = ref_flops(M->prep.force_type, M->prep.multipole_order, G.evals, G.steps, G.pots, G.pairs,
                                   G.cl);
    }
    out->lanes = L;
    out->wg_size = wg;
    out->sample_kernel_ms = sms;
    // per-integral sums (k_reduce), then each integral's CCS and error with
    // the reference's denominator (k_iter_ccs): only iters values come back
    double *red = lease.aux, *om = red + kSums * iters, *er = om + iters;
    ReduceArgs RA{};
    RA.status = B.status.as<int8_t>(); RA.chi = B.chi.as<double>(); RA.ergy = B.ergy.as<double>();
    RA.totsteps = B.tot.as<int32_t>(); RA.per_iter = T; RA.ccs_const = M->prep.ccs_const; RA.out = red;
    hipLaunchKernelGGL(k_reduce, dim3(iters), dim3(256), 0, s, RA);
    HIP_CHECK(hipGetLastError());
    const int64_t pr = P->print_rate;
    const int upd = (int)(pr > 0 ? std::min<int64_t>(pr, INT32_MAX) : std::max<int64_t>(1, (int64_t)(T / 5.0)));
    hipLaunchKernelGGL(k_iter_ccs, dim3(iters), dim3(256), 0, s, (const int8_t *)B.status.p, (const double *)red,
                       (int)T, upd, om, er);
    HIP_CHECK(hipGetLastError());
    std::vector<double> sums(kSums * iters);
    std::vector<int8_t> status(c.status ? n : 0);
    HIP_CHECK(hipMemcpyAsync(sums.data(), red, sizeof(double) * kSums * iters, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(omega, om, sizeof(double) * iters, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(omega_err, er, sizeof(double) * iters, hipMemcpyDeviceToHost, s));
    if (c.status) HIP_CHECK(hipMemcpyAsync(status.data(), B.status.p, n, hipMemcpyDeviceToHost, s));
    if (c.chi) HIP_CHECK(hipMemcpyAsync(c.chi, B.chi.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    if (c.g) HIP_CHECK(hipMemcpyAsync(c.g, B.g.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (c.status)
        for (int64_t i = 0; i < n; i++)
            c.status[i] = status[i] == ST_OK ? 1 : (status[i] == ST_MISS ? 0 : 2);
    final_stats(omega, omega_err, sums.data(), iters, T, out);
    out->n_traj = n;
    const int ft = M->prep.force_type;
    // clustered types: the member atoms their branches evaluated
    const int64_t npair = (ft == 4 || ft == 6) ? M->prep.n : 0;
    int64_t evals = 0;
    for (int r = 0; r < out->n_rounds; r++) evals += out->round_evals[r];
    out->pair_evals = npair * evals + out->cl_counts[3] + out->cl_counts[4] + out->cl_counts[5];
    if (L == 1 && npair > 0) {  // the lane kernel counts steps and evaluations, not potential pairs
        int64_t steps = 0;
        for (int r = 0; r < out->n_rounds; r++) steps += out->round_steps[r];
        const int64_t no_cl[6] = {0, 0, 0, 0, 0, 0};
        out->ref_flops = ref_flops(ft, M->prep.multipole_order, evals, steps, 0, out->pair_evals, no_cl);
    }
    out->run_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

int cosims_run_noalloc(const cosims_mol *M, const cosims_params *P, const cosims_run_cfg *cfg, double *omega,
                       double *omega_err, cosims_ccs *out)
{
    if (!M || !P || !cfg || !omega || !omega_err || !out)
        return fail(COSIMS_E_INVALID, "cosims_run_noalloc: null argument");
    const cosims_run_cfg c = *cfg;
    const int prec = c.prec;
    if (prec != COSIMS_PREC_FP64 && prec != COSIMS_PREC_MIXED) return fail(COSIMS_E_INVALID, "unknown precision");
    const int iters = c.iter > 0 ? c.iter : P->iter;
    const int64_t T = c.traj > 0 ? c.traj : P->traj;
    if (iters <= 0 || T <= 0) return fail(COSIMS_E_INVALID, "cosims_run_noalloc: no trajectories");
    if (iters > COSIMS_NOALLOC_MAX_ITER || T > COSIMS_NOALLOC_MAX_TRAJ || (int64_t)iters * T > COSIMS_NOALLOC_MAX_TRAJ)
        return fail(COSIMS_E_INVALID, "cosims_run_noalloc: " + std::to_string(iters) + " x " + std::to_string(T) +
                                          " trajectories exceed its static scratch (at most " +
                                          std::to_string(COSIMS_NOALLOC_MAX_TRAJ) + " trajectories in at most " +
                                          std::to_string(COSIMS_NOALLOC_MAX_ITER) + " integrals): use cosims_run");
    const int64_t n = (int64_t)iters * T;
    if (c.rng != COSIMS_RNG_PHILOX)
        return fail(COSIMS_E_INVALID, "cosims_run_noalloc: Philox initial conditions only (the reference's stream "
                                      "is drawn on the host)");
    if (!M->on_device.load())
        return fail(COSIMS_E_INVALID, "cosims_run_noalloc: molecule not on the device (call cosims_mol_upload first)");
    int dev = -1;
    HIP_CHECK(hipGetDevice(&dev));
    if (dev != M->dev.id)
        return fail(COSIMS_E_INVALID, "cosims_run_noalloc: the molecule is on device " + std::to_string(M->dev.id) +
                                          ", device " + std::to_string(dev) + " is current");
    int L = c.lanes, wg = c.wg_size;
    if (L == 0) {  // the shape cosims_run tuned for runs of this size, if it has
        std::lock_guard<std::mutex> lk(M->tune_mu);
        int key = 0;
        (void)tune_size(M, prec, n, &key);
        auto it = M->tuned.find(key);
        if (it == M->tuned.end())
            return fail(COSIMS_E_INVALID, "cosims_run_noalloc: no tuned launch shape for runs of this size: give "
                                          "lanes and wg_size, or tune with cosims_run (lanes 0) first");
        L = it->second.first;
        if (wg <= 0) wg = it->second.second;
    }
    if (L != 4 && L != 8 && L != 16 && L != 32 && L != 64)
        return fail(COSIMS_E_INVALID, "cosims_run_noalloc: lanes must be 0 (tuned before), 4, 8, 16, 32 or 64");
    if (wg <= 0 || wg % 64 || wg > 1024 || wg % L)
        return fail(COSIMS_E_INVALID, "cosims_run_noalloc: wg_size must be a multiple of 64 and of lanes, at most 1024");
    ScratchDev *S = nullptr;
    if (int rc = scratch_of(dev, S)) return rc;
    std::lock_guard<std::mutex> lk(S->mu);
    std::memset(out, 0, sizeof(*out));
    const auto t0 = std::chrono::steady_clock::now();
    TrajBufs B;
    DBuf red, om, er, st32;
    size_t off = 0;
    bool fits = B.carve(S->ptr, kNoAllocBytes, &off, n);
    fits = fits && carve_one(red, S->ptr, kNoAllocBytes, &off, sizeof(double) * kSums * iters);
    fits = fits && carve_one(om, S->ptr, kNoAllocBytes, &off, sizeof(double) * iters);
    fits = fits && carve_one(er, S->ptr, kNoAllocBytes, &off, sizeof(double) * iters);
    fits = fits && carve_one(st32, S->ptr, kNoAllocBytes, &off, sizeof(int32_t) * n);
    if (!fits) return fail(COSIMS_E_NOMEM, "cosims_run_noalloc: scratch layout larger than the static array");
    // every buffer read below is written first in this call (sample_into and
    // the launch clear their counters): nothing of the last call shows through
    const hipStream_t s = hipStreamPerThread;
    if (int rc = sample_into(M, P->seed, c.job, n, T, B, s, nullptr)) return rc;
    int32_t hits = 0;
    HIP_CHECK(hipMemcpyAsync(&hits, B.counters.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (hits > 0) {
        const int32_t *work = nullptr;
        if (kOrderWork) {  // longest expected trajectories first, as cosims_run launches them
            work = B.items[1].as<int32_t>();
            if (int rc = order_items(B.items[0].as<int32_t>(), B.items[1].as<int32_t>(), nullptr, nullptr, hits,
                                     B.bin.as<uint8_t>(), B.order.as<int32_t>(), s))
                return rc;
        }
        GroupRun G;
        if (int rc = run_group(M, prec, hits, B, s, L, wg, c.blocks_per_cu, G, work, false)) return rc;
        out->n_rounds = 1;
        out->round_items[0] = hits;
        out->round_steps[0] = G.steps;
        out->round_evals[0] = G.evals;
        out->pot_evals = G.pots;
        for (int k = 0; k < 6; k++) out->cl_counts[k] = G.cl[k];
        out->ref_flops = ref_flops(M->prep.force_type, M->prep.multipole_order, G.evals, G.steps, G.pots, G.pairs,
                                   G.cl);
    }
    out->lanes = L;
    out->wg_size = wg;
    ReduceArgs RA{};
    RA.status = B.status.as<int8_t>(); RA.chi = B.chi.as<double>(); RA.ergy = B.ergy.as<double>();
    RA.totsteps = B.tot.as<int32_t>(); RA.per_iter = T; RA.ccs_const = M->prep.ccs_const; RA.out = red.as<double>();
    hipLaunchKernelGGL(k_reduce, dim3(iters), dim3(256), 0, s, RA);
    HIP_CHECK(hipGetLastError());
    const int64_t pr = P->print_rate;
    const int upd = (int)(pr > 0 ? std::min<int64_t>(pr, INT32_MAX) : std::max<int64_t>(1, (int64_t)(T / 5.0)));
    hipLaunchKernelGGL(k_iter_ccs, dim3(iters), dim3(256), 0, s, (const int8_t *)B.status.p,
                       (const double *)red.p, (int)T, upd, om.as<double>(), er.as<double>());
    HIP_CHECK(hipGetLastError());
    if (c.status) {
        hipLaunchKernelGGL(k_status_out, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           (const int8_t *)B.status.p, n, st32.as<int32_t>());
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(c.status, st32.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
    }
    double sums[kSums * COSIMS_NOALLOC_MAX_ITER];
    HIP_CHECK(hipMemcpyAsync(sums, red.p, sizeof(double) * kSums * iters, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(omega, om.p, sizeof(double) * iters, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(omega_err, er.p, sizeof(double) * iters, hipMemcpyDeviceToHost, s));
    if (c.chi) HIP_CHECK(hipMemcpyAsync(c.chi, B.chi.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    if (c.g) HIP_CHECK(hipMemcpyAsync(c.g, B.g.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    final_stats(omega, omega_err, sums, iters, T, out);
    out->n_traj = n;
    const int ft = M->prep.force_type;
    const int64_t npair = (ft == 4 || ft == 6) ? M->prep.n : 0;
    out->pair_evals = npair * out->round_evals[0] + out->cl_counts[3] + out->cl_counts[4] + out->cl_counts[5];
    out->run_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

int cosims_force_rate(const cosims_mol *M, int32_t prec, int32_t lanes, int32_t wg_size, double *tflops,
                      double *pairs_per_s, double *kernel_ms)
{
    if (!M || !tflops) return fail(COSIMS_E_INVALID, "cosims_force_rate: null argument");
    if (prec != COSIMS_PREC_FP64 && prec != COSIMS_PREC_MIXED) return fail(COSIMS_E_INVALID, "unknown precision");
    if (lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64)
        return fail(COSIMS_E_INVALID, "cosims_force_rate: lanes must be 4, 8, 16, 32 or 64");
    if (wg_size <= 0 || wg_size % 64 || wg_size > 1024)
        return fail(COSIMS_E_INVALID, "cosims_force_rate: wg_size must be a multiple of 64, at most 1024");
    if (int rc = ensure_device(M)) return rc;
    if (!lds_mode(M, prec))
        return fail(COSIMS_E_UNSUPPORTED, "cosims_force_rate: an exact force type whose molecule fits LDS only");
    const int ft = M->prep.force_type;
    RateFn fn = rate_kernel(prec, ft, lanes);
    const size_t lds = group_lds_bytes(M, prec);
    HIP_CHECK(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void *)fn, wg_size, lds) != hipSuccess || occ <= 0)
        occ = 1;
    const int64_t grid = (int64_t)occ * M->dev.cus;
    const int64_t groups = grid * wg_size / lanes;
    const int64_t n = M->prep.n;
    // about 150 ms a launch at the trajectory kernels' pair rates
    const double target_pairs = (prec == COSIMS_PREC_MIXED ? 1.5e12 : 0.3e12) * 0.15;
    const int evals = (int)std::max<double>(4.0, std::min<double>(1e6, target_pairs / ((double)groups * n)));
    GroupArgs G = group_args(M);
    DBuf sink;
    HIP_CHECK(sink.alloc(sizeof(double)));
    Stream st;
    HIP_CHECK(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(wg_size), lds, st.s, G, std::max(1, evals / 8),
                       sink.as<double>());
    HIP_CHECK(hipGetLastError());
    float best = 1e30f;
    for (int rep = 0; rep < 3; rep++) {
        HIP_CHECK(hipEventRecord(e0, st.s));
        hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(wg_size), lds, st.s, G, evals, sink.as<double>());
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipEventRecord(e1, st.s));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        best = std::min(best, ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // reference FLOPs: type 4 58 a pair + 21 an evaluation (src/forceAccel.cpp:764-802),
    // type 6 22 a pair (:865-880); padding atoms are not counted
    const double ev = (double)groups * evals, pairs = ev * (double)n;
    const double flops = ft == 4 ? 58.0 * pairs + 21.0 * ev : 22.0 * pairs;
    *tflops = flops / (best * 1e-3) / 1e12;
    if (pairs_per_s) *pairs_per_s = pairs / (best * 1e-3);
    if (kernel_ms) *kernel_ms = best;
    return 0;
}

int cosims_tune_batch(const cosims_batch_job *jobs, int32_t njobs, int32_t prec, int32_t *class_lanes,
                      int32_t *class_wg)
{
    if (!class_lanes || !class_wg) return fail(COSIMS_E_INVALID, "cosims_tune_batch: null output");
    if (prec != COSIMS_PREC_FP64 && prec != COSIMS_PREC_MIXED) return fail(COSIMS_E_INVALID, "unknown precision");
    ShapeTable T;
    if (int rc = tune_batch(jobs, njobs, prec, T)) return rc;
    for (int c = 0; c < COSIMS_SIZE_CLASSES; c++) {
        class_lanes[c] = T.class_lanes[c];
        class_wg[c] = T.class_wg[c];
    }
    return 0;
}

int cosims_run_batch(cosims_batch_job *jobs, int32_t njobs, const cosims_batch_cfg *cfg, cosims_batch_stats *stats)
{
    if (!cfg || !stats) return fail(COSIMS_E_INVALID, "cosims_run_batch: null argument");
    const int prec = cfg->prec;
    if (prec != COSIMS_PREC_FP64 && prec != COSIMS_PREC_MIXED) return fail(COSIMS_E_INVALID, "unknown precision");
    for (int k = 0; jobs && k < njobs; k++)
        if (!jobs[k].omega || !jobs[k].omega_err || !jobs[k].out)
            return fail(COSIMS_E_INVALID, "cosims_run_batch: job without outputs");
    auto lanes_ok = [](int L) { return L == 4 || L == 8 || L == 16 || L == 32 || L == 64; };
    auto wg_ok = [](int W) { return W > 0 && W % 64 == 0 && W <= 1024; };
    ShapeTable shape;
    shape.lanes = cfg->lanes > 0 ? cfg->lanes : 0;
    shape.wg = cfg->wg_size > 0 ? cfg->wg_size : 512;
    if ((cfg->lanes > 0 && !lanes_ok(cfg->lanes)) || cfg->lanes < -1)
        return fail(COSIMS_E_INVALID, "lanes must be -1 (tuned per size class), 0 (size rule), 4, 8, 16, 32 or 64");
    if (!wg_ok(shape.wg)) return fail(COSIMS_E_INVALID, "wg_size must be a multiple of 64, at most 1024");
    for (int c = 0; c < COSIMS_SIZE_CLASSES; c++) {
        if (cfg->class_lanes[c] && !lanes_ok(cfg->class_lanes[c])) return fail(COSIMS_E_INVALID, "bad class_lanes");
        if (cfg->class_wg[c] && !wg_ok(cfg->class_wg[c])) return fail(COSIMS_E_INVALID, "bad class_wg");
        shape.class_lanes[c] = cfg->class_lanes[c];
        shape.class_wg[c] = cfg->class_wg[c];
    }
    std::memset(stats, 0, sizeof(*stats));
    auto t0 = std::chrono::steady_clock::now();
    if (cfg->lanes == -1) {
        bool given = false;
        for (int c = 0; c < COSIMS_SIZE_CLASSES; c++) given |= cfg->class_lanes[c] != 0;
        if (!given)
            if (int rc = tune_batch(jobs, njobs, prec, shape, cfg->own_launches == 0)) return rc;
    }
    BatchPlan P;
    if (int rc = plan_batch(jobs, njobs, 0, shape, P)) return rc;
    P.shared = cfg->own_launches == 0;
    Stream st;
    HIP_CHECK(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
    if (int rc = stage_batch(P, st.s)) return rc;
    GroupRun R;
    int launches = 0;
    if (int rc = launch_batch(P, prec, cfg->blocks_per_cu, st.s, R, launches)) return rc;
    // per-integral sums of every job (k_reduce per job), then the host statistics
    int64_t total_iters = 0;
    for (const auto &J : P.jobs) total_iters += J.iters;
    DBuf red;
    HIP_CHECK(red.alloc(sizeof(double) * kSums * total_iters));
    int64_t it0 = 0;
    std::vector<int64_t> it_base(P.jobs.size());
    for (size_t k = 0; k < P.jobs.size(); k++) {
        const auto &J = P.jobs[k];
        it_base[k] = it0;
        ReduceArgs RA{};
        RA.status = P.B.status.as<int8_t>() + J.off; RA.chi = P.B.chi.as<double>() + J.off;
        RA.cchi = P.B.exitv2.as<double>() + J.off;
        RA.ergy = P.B.ergy.as<double>() + J.off; RA.totsteps = P.B.tot.as<int32_t>() + J.off;
        RA.per_iter = J.T; RA.ccs_const = J.mol->prep.ccs_const; RA.out = red.as<double>() + kSums * it0;
        hipLaunchKernelGGL(k_reduce, dim3(J.iters), dim3(256), 0, st.s, RA);
        HIP_CHECK(hipGetLastError());
        it0 += J.iters;
    }
    std::vector<double> sums(kSums * total_iters);
    std::vector<int8_t> status(P.n);
    HIP_CHECK(hipMemcpyAsync(sums.data(), red.p, sizeof(double) * sums.size(), hipMemcpyDeviceToHost, st.s));
    HIP_CHECK(hipMemcpyAsync(status.data(), P.B.status.p, P.n, hipMemcpyDeviceToHost, st.s));
    HIP_CHECK(hipStreamSynchronize(st.s));
    for (size_t k = 0; k < P.jobs.size(); k++) {
        const auto &J = P.jobs[k];
        cosims_ccs *o = jobs[k].out;
        std::memset(o, 0, sizeof(*o));
        finish_ccs(status.data() + J.off, sums.data() + kSums * it_base[k], J.iters, J.T, J.mol->params.print_rate,
                   jobs[k].omega, jobs[k].omega_err, o);
        o->lanes = J.L;
        o->wg_size = J.W;
        stats->n_hit += o->n_hit;
        stats->total_steps += o->total_steps;
    }
    for (int c = 0; c < COSIMS_SIZE_CLASSES; c++) {
        stats->class_lanes[c] = shape.class_lanes[c];
        stats->class_wg[c] = shape.class_wg[c];
    }
    stats->lanes = shape.lanes;
    stats->wg_size = shape.wg;
    stats->launches = launches;
    stats->kernel_ms = R.ms;
    stats->sample_ms = P.sample_ms;
    stats->n_traj = P.n;
    stats->force_evals = R.evals;
    stats->pot_evals = R.pots;
    stats->pair_evals = R.pairs + R.cl[3] + R.cl[4] + R.cl[5];
    stats->ref_flops = R.flops;
    stats->run_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

}  // extern "C"

#include "server.hpp"
