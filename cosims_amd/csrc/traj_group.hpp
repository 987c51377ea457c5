// traj_group.hpp -- group-per-trajectory persistent kernel (gfx950).
//
// L lanes (a power of two, 4..64) integrate one trajectory together: every
// lane of the group holds the same FP64 Verlet state and takes atoms j,
// j+L, j+2L, ... of the force sum; a butterfly of lane shuffles gives every
// lane the full sums.  A wavefront therefore carries 64/L trajectories and
// a step costs N/L pairs per lane instead of N, so a long trajectory no
// longer holds a whole launch hostage.  The energy-conservation test
// (src/diffEqSolver.cpp:108, 178-203) runs inside the kernel as an FP64
// getPotential across the group, and a failed attempt restarts at once
// with dt/(k+1): one launch covers every attempt of every trajectory.
//
// The molecule is staged once per workgroup in LDS (SoA: FP32 x,y,z,q |
// sigma^6, 24 eps/mu as float4 + float2, or the FP64 equivalents), padded to
// a multiple of 64 with inert atoms (charge and LJ parameters zero, far
// away) so every lane runs the same trip count.  The FP64 potential reads
// global SoA arrays (the rare path).
#pragma once

#include "kernels.hpp"

namespace cosims {

struct GroupArgs {
    MolDev m;
    int npad;                              // atoms incl. padding, multiple of 64
    const double *sx, *sy, *sz, *sq, *s6, *se24, *se4;  // global FP64 SoA (npad)
    const int32_t *items;
    int32_t n_items;
    int32_t *counter;
    const double *p0, *v0;
    int8_t *status;
    double *chi, *cchi, *ergy;             // cchi: cos(chi) of accepted trajectories
    int32_t *steps, *totsteps;
    double dt0;
    int32_t max_steps;
    unsigned long long *evals;             // [0] steps, [1] force evaluations, [2] potential evaluations
};

// chi = vector3D::Angle of the accepted exit (src/vector3D.h:134-143) from
// the stored cosine, for the items of one group launch.
__global__ void __launch_bounds__(256) k_chi(const int32_t *items, int n, const int8_t *status, const double *cchi,
                                             double *chi)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int t = items[i];
    if (status[t] != 1) return;
    const double t3 = cchi[t];
    chi[t] = (t3 > 1) ? 0.0 : ((t3 < -1) ? 3.14159265358979323846 : acos(t3));
}

template <int L, typename T>
__device__ __forceinline__ T gsum(T v)
{
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// FP32 pair kernel of the exact-sum types, partial sums of this lane.
template <int FT>
struct Part32 {
    float fx = 0, fy = 0, fz = 0, ex = 0, ey = 0, ez = 0, sxx = 0, syy = 0, szz = 0, sxy = 0, sxz = 0, syz = 0;
};

#pragma clang fp contract(fast)
template <int FT, int L>
__device__ __forceinline__ void grp_force32(const float4 *__restrict__ xq, const float2 *__restrict__ se, int npad,
                                            int j, float px, float py, float pz, double ds, double &ox, double &oy,
                                            double &oz)
{
    Part32<FT> P;
#pragma unroll 2
    for (int k = j; k < npad; k += L) {
        const float4 a = xq[k];
        const float2 b = se[k];
        float dx = px - a.x, dy = py - a.y, dz = pz - a.z;
        float r2 = dx * dx + dy * dy + dz * dz;
        if constexpr (FT == 4) {
            float ir = __builtin_amdgcn_rsqf(r2);
            float ir2 = ir * ir;
            float t2 = b.x * ir2 * ir2 * ir2;
            float lj = b.y * ir2 * t2 * (2.0f * t2 - 1.0f);
            P.fx += lj * dx; P.fy += lj * dy; P.fz += lj * dz;
            float qr3 = a.w * ir2 * ir;
            P.ex += qr3 * dx; P.ey += qr3 * dy; P.ez += qr3 * dz;
            float qr5 = qr3 * ir2;
            float u = qr5 * dx, v = qr5 * dy, w = qr5 * dz;
            P.sxx += u * dx; P.syy += v * dy; P.szz += w * dz;
            P.sxy += u * dy; P.sxz += u * dz; P.syz += v * dz;
        } else {
            float ir2 = __builtin_amdgcn_rcpf(r2);
            float t2 = b.x * ir2 * ir2 * ir2;
            float lj = b.y * ir2 * t2 * (2.0f * t2 - 1.0f);
            P.fx += lj * dx; P.fy += lj * dy; P.fz += lj * dz;
        }
    }
    P.fx = gsum<L>(P.fx); P.fy = gsum<L>(P.fy); P.fz = gsum<L>(P.fz);
    if constexpr (FT == 4) {
        P.ex = gsum<L>(P.ex); P.ey = gsum<L>(P.ey); P.ez = gsum<L>(P.ez);
        P.sxx = gsum<L>(P.sxx); P.syy = gsum<L>(P.syy); P.szz = gsum<L>(P.szz);
        P.sxy = gsum<L>(P.sxy); P.sxz = gsum<L>(P.sxz); P.syz = gsum<L>(P.syz);
        double tr = (double)P.sxx + (double)P.syy + (double)P.szz;
        double Txx = 3.0 * P.sxx - tr, Tyy = 3.0 * P.syy - tr, Tzz = 3.0 * P.szz - tr;
        double Txy = 3.0 * P.sxy, Txz = 3.0 * P.sxz, Tyz = 3.0 * P.syz;
        ox = (double)P.fx - ds * (Txx * P.ex + Txy * P.ey + Txz * P.ez);
        oy = (double)P.fy - ds * (Txy * P.ex + Tyy * P.ey + Tyz * P.ez);
        oz = (double)P.fz - ds * (Txz * P.ex + Tyz * P.ey + Tzz * P.ez);
    } else {
        ox = P.fx; oy = P.fy; oz = P.fz;
    }
}
#pragma clang fp contract(off)

// FP64 reference-form pair terms (src/forceAccel.cpp:764-797, 865-880),
// partial sums of this lane, then the group butterfly and the epilogue
// (:800-802).
template <int FT, int L>
__device__ __forceinline__ void grp_force64(const double *__restrict__ X, int npad, int j, double px, double py,
                                            double pz, double ds, double &ox, double &oy, double &oz)
{
    const double *Y = X + npad, *Z = X + 2 * npad, *Q = X + 3 * npad, *S6 = X + 4 * npad, *E24 = X + 5 * npad;
    double sXX = 0, sYY = 0, sZZ = 0, sXY = 0, sXZ = 0, sYZ = 0, dX = 0, dY = 0, dZ = 0, tx = 0, ty = 0, tz = 0;
#pragma unroll 2
    for (int k = j; k < npad; k += L) {
        double xD = px - X[k], yD = py - Y[k], zD = pz - Z[k];
        double x2 = xD * xD, y2 = yD * yD, z2 = zD * zD;
        double r2 = x2 + y2 + z2;
        double r6 = r2 * r2 * r2;
        double t2 = S6[k] / r6;
        double t1 = 2 * t2 * t2;
        double sc = E24[k] * (t1 - t2) / r2;
        tx += sc * xD; ty += sc * yD; tz += sc * zD;
        if constexpr (FT == 4) {
            double r = __builtin_sqrt(r2);
            double r3 = r2 * r;
            double r_3 = 1.0 / r3;
            double r_5 = r_3 / r2;
            r_3 *= Q[k];
            r_5 *= Q[k];
            dX += xD * r_3; dY += yD * r_3; dZ += zD * r_3;
            sXX += (3.0 * x2 - r2) * r_5;
            sYY += (3.0 * y2 - r2) * r_5;
            sZZ += (3.0 * z2 - r2) * r_5;
            sXY += 3 * xD * yD * r_5;
            sXZ += 3 * xD * zD * r_5;
            sYZ += 3 * yD * zD * r_5;
        }
    }
    tx = gsum<L>(tx); ty = gsum<L>(ty); tz = gsum<L>(tz);
    if constexpr (FT == 4) {
        dX = gsum<L>(dX); dY = gsum<L>(dY); dZ = gsum<L>(dZ);
        sXX = gsum<L>(sXX); sYY = gsum<L>(sYY); sZZ = gsum<L>(sZZ);
        sXY = gsum<L>(sXY); sXZ = gsum<L>(sXZ); sYZ = gsum<L>(sYZ);
        tx -= ds * (sXX * dX + sXY * dY + sXZ * dZ);
        ty -= ds * (sXY * dX + sYY * dY + sYZ * dZ);
        tz -= ds * (sXZ * dX + sYZ * dY + sZZ * dZ);
    }
    ox = tx; oy = ty; oz = tz;
}

// getPotential across the group (src/forceAccel.cpp:943-982), FP64.
template <int L>
__device__ __forceinline__ double grp_potential(const GroupArgs &A, int j, double px, double py, double pz)
{
    double total = 0, sX = 0, sY = 0, sZ = 0;
    const bool charged = A.m.charged;
    for (int k = j; k < A.npad; k += L) {
        double xD = px - A.sx[k], yD = py - A.sy[k], zD = pz - A.sz[k];
        double r2 = xD * xD + yD * yD + zD * zD;
        double r6 = r2 * r2 * r2;
        double t2 = A.s6[k] / r6;
        double t1 = t2 * t2;
        total += A.se4[k] * (t1 - t2);
        if (charged) {
            double r_3 = A.sq[k] / (r2 * __builtin_sqrt(r2));
            sX += xD * r_3; sY += yD * r_3; sZ += zD * r_3;
        }
    }
    total = gsum<L>(total);
    if (charged) {
        sX = gsum<L>(sX); sY = gsum<L>(sY); sZ = gsum<L>(sZ);
        total -= A.m.dipole_scalar * 0.5 * (sX * sX + sY * sY + sZ * sZ);
    }
    return total;
}

enum : int { PH_IDLE = 0, PH_INIT = 1, PH_STEP = 2, PH_CHECK = 3 };

template <int PREC, int FT, int L>
__global__ void __launch_bounds__(1024) k_traj_grp(GroupArgs A)
{
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const MolDev &m = A.m;
    const int npad = A.npad;
    // stage the molecule (once per workgroup; the kernel is persistent)
    if constexpr (PREC == 1) {
        float4 *xq = (float4 *)smem;
        float2 *se = (float2 *)(xq + npad);
        for (int i = threadIdx.x; i < npad; i += blockDim.x) {
            xq[i] = make_float4((float)A.sx[i], (float)A.sy[i], (float)A.sz[i], (float)A.sq[i]);
            se[i] = make_float2((float)A.s6[i], (float)A.se24[i]);
        }
    } else {
        double *X = (double *)smem;
        for (int i = threadIdx.x; i < npad; i += blockDim.x) {
            X[i] = A.sx[i]; X[npad + i] = A.sy[i]; X[2 * npad + i] = A.sz[i];
            X[3 * npad + i] = A.sq[i]; X[4 * npad + i] = A.s6[i]; X[5 * npad + i] = A.se24[i];
        }
    }
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int g0 = lane & ~(L - 1);   // first lane of my group
    const int j = lane & (L - 1);     // my index in the group
    const double hm = 0.5 * m.mu;
    int t = 0, phase = PH_IDLE, tries = 0, maxtries = 8, orbits = 0, steps = 0, tot = 0;
    bool needE0 = false;
    double px = 0, py = 0, pz = 0, vx = 0, vy = 0, vz = 0, ax = 0, ay = 0, az = 0, dt = 0, rdot = 0;
    double e0 = 0, chi = 0;
    bool exhausted = false;
    unsigned n_steps = 0, n_evals = 0, n_pots = 0;

    for (;;) {
        if (!exhausted) {
            const bool need = (phase == PH_IDLE);
            const unsigned long long mask = __ballot(need && j == 0);
            if (mask) {
                const int cnt = __popcll(mask);
                const int leader = __ffsll((unsigned long long)mask) - 1;
                int base = 0;
                if (lane == leader) base = atomicAdd(A.counter, cnt);
                base = __shfl(base, leader);
                if (base + cnt >= A.n_items) exhausted = true;
                if (need) {
                    const int idx = base + __popcll(mask & ((1ull << g0) - 1));
                    if (idx < A.n_items) {
                        t = A.items[idx];
                        tries = 0; maxtries = 8;
                        dt = A.dt0;
                        px = A.p0[3 * t]; py = A.p0[3 * t + 1]; pz = A.p0[3 * t + 2];
                        vx = A.v0[3 * t]; vy = A.v0[3 * t + 1]; vz = A.v0[3 * t + 2];
                        tot = 0;
                        needE0 = true;
                        phase = PH_INIT;
                    }
                }
            }
        }
        if (!__any(phase != PH_IDLE)) break;

        // position of this group's force / potential evaluation
        // (src/diffEqSolver.cpp:131; recomputed after the force call instead
        // of being held live across it -- same operands, same result)
        double ex, ey, ez;
        if (phase == PH_STEP) {
            ex = px + vx * dt + ax * dt * dt * 0.5;
            ey = py + vy * dt + ay * dt * dt * 0.5;
            ez = pz + vz * dt + az * dt * dt * 0.5;
        } else if (phase == PH_IDLE) {
            ex = ey = ez = 1.0e4;
        } else {  // INIT: start point; CHECK: exit point
            ex = px; ey = py; ez = pz;
        }
        const bool needF = (phase == PH_INIT || phase == PH_STEP);
        const bool needU = (phase == PH_INIT && needE0) || phase == PH_CHECK;
        double fx = 0, fy = 0, fz = 0, u = 0;
        if (__any(needF)) {
            if constexpr (PREC == 1) {
                const float4 *xq = (const float4 *)smem;
                grp_force32<FT, L>(xq, (const float2 *)(xq + npad), npad, j, (float)ex, (float)ey, (float)ez,
                                   m.dipole_scalar, fx, fy, fz);
            } else {
                grp_force64<FT, L>((const double *)smem, npad, j, ex, ey, ez, m.dipole_scalar, fx, fy, fz);
            }
        }
        if (__any(needU)) u = grp_potential<L>(A, j, ex, ey, ez);
        if (j == 0) {
            n_evals += needF;
            n_steps += (phase == PH_STEP);
            n_pots += needU;
        }

        if (phase == PH_INIT) {  // src/diffEqSolver.cpp:112-127
            if (needE0) {
                e0 = u + hm * (vx * vx + vy * vy + vz * vz);
                needE0 = false;
            }
            ax = fx; ay = fy; az = fz;
            orbits = 0;
            steps = 0;
            rdot = px * vx + py * vy + pz * vz;
            phase = PH_STEP;
        } else if (phase == PH_STEP) {  // src/diffEqSolver.cpp:132-205
            px = px + vx * dt + ax * dt * dt * 0.5;
            py = py + vy * dt + ay * dt * dt * 0.5;
            pz = pz + vz * dt + az * dt * dt * 0.5;
            vx = vx + ax * dt * 0.5;
            vy = vy + ay * dt * 0.5;
            vz = vz + az * dt * 0.5;
            ax = fx; ay = fy; az = fz;
            vx = vx + ax * dt * 0.5;
            vy = vy + ay * dt * 0.5;
            vz = vz + az * dt * 0.5;
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
            const bool runaway = !(fabs(px) + fabs(py) + fabs(pz) < 1.0e30) || steps > A.max_steps;
            if (runaway) {
                if (j == 0) {
                    A.status[t] = 2; A.chi[t] = 1.57079632679489661923; A.steps[t] = steps; A.totsteps[t] = tot;
                }
                phase = PH_IDLE;
            } else if (!track) {
                tries++;
                dt = A.dt0 / (double)(tries + 1);
                if (!(tries < maxtries && maxtries < 10)) {
                    if (j == 0) {
                        A.status[t] = 2; A.chi[t] = 1.57079632679489661923; A.steps[t] = steps;
                        A.totsteps[t] = tot;
                    }
                    phase = PH_IDLE;
                }
            } else {
                const double tx = px / m.a, ty = py / m.b, tz = pz / m.c;
                const double tm2 = tx * tx + ty * ty + tz * tz;
                if (((tm2 > 1 && m.projection) || (tm2 > m.maxrad && !m.projection)) && steps > 10) {
                    const double w0 = A.v0[3 * t], w1 = A.v0[3 * t + 1], w2 = A.v0[3 * t + 2];
                    const double t1 = w0 * vx + w1 * vy + w2 * vz;
                    const double t2 = sqrt(w0 * w0 + w1 * w1 + w2 * w2) * sqrt(vx * vx + vy * vy + vz * vz);
                    chi = t1 / t2;  // cos(chi); acos is taken after the launch (k_chi)
                    phase = PH_CHECK;
                }
            }
        } else if (phase == PH_CHECK) {  // src/diffEqSolver.cpp:178-203
            const double e1 = u + hm * (vx * vx + vy * vy + vz * vz);
            const double pct = 100 * fabs(e0 - e1) / e0;
            if (pct > 0.5) {
                tries++;
                maxtries++;
                dt = A.dt0 / (double)(tries + 1);
                if (tries < maxtries && maxtries < 10) {
                    px = A.p0[3 * t]; py = A.p0[3 * t + 1]; pz = A.p0[3 * t + 2];
                    vx = A.v0[3 * t]; vy = A.v0[3 * t + 1]; vz = A.v0[3 * t + 2];
                    phase = PH_INIT;  // restart; E0 is kept
                } else {
                    if (j == 0) {
                        A.status[t] = 2; A.chi[t] = 1.57079632679489661923; A.steps[t] = 0; A.totsteps[t] = tot;
                    }
                    phase = PH_IDLE;
                }
            } else {
                if (j == 0) {
                    A.status[t] = 1; A.cchi[t] = chi; A.ergy[t] = pct; A.steps[t] = steps; A.totsteps[t] = tot;
                }
                phase = PH_IDLE;
            }
        }
    }
    unsigned long long s = n_steps, e = n_evals, p = n_pots;
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off);
        e += __shfl_xor(e, off);
        p += __shfl_xor(p, off);
    }
    if (lane == 0) {
        atomicAdd(A.evals, s);
        atomicAdd(A.evals + 1, e);
        atomicAdd(A.evals + 2, p);
    }
}

}  // namespace cosims
