// kernels.hpp -- device data layout and force evaluators of the trajectory
// engine (gfx950).  Included by engine.hip only.
//
// Layout: the prepared molecule lives in one device allocation as
//   atoms (reference atom order)   AtomF64[n] | AtomF32[n]
//   atoms (cluster member order)   AtomF64[n] | AtomF32[n]
//   cluster headers                ClusterF64[ncl] | ClusterF32[ncl]
// and is read through constant-address-space pointers: every lane of a
// wavefront walks the same atom at the same time, so the loads are
// wave-uniform and compile to scalar (SMEM) loads into SGPRs that the VALU
// reads as operands -- no LDS staging, no per-lane replication.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define CAS __attribute__((address_space(4)))

namespace cosims {

struct __attribute__((aligned(64))) AtomF64 {
    double x, y, z, q, s6, e24, e4, pad;
};
struct __attribute__((aligned(32))) AtomF32 {
    float x, y, z, q, s6, e24, e4, pad;
};
// Cluster header: position (centre or centre of charge), squared near/far
// radii (R + rho)^2, moments, and the LJ parameters of atom index = cluster
// index that the band branch of the reference uses (src/forceAccel.cpp:353-355).
template <typename T>
struct ClusterT {
    T px, py, pz, mpr, dpr, q, d0, d1, d2;
    T Q[9];
    T qs6, qe24;
    int32_t start, end, rep, pad;
};
using ClusterF64 = ClusterT<double>;
using ClusterF32 = ClusterT<float>;

struct MolDev {
    int n, ncl, force_type, charged, order, projection;
    double mu, dipole_scalar, a, b, c, maxrad;
    double ia, ib, ic;    // 1/a, 1/b, 1/c (IEEE), for outside()
    const AtomF64 *a64;   // atom order
    const AtomF32 *a32;
    const AtomF64 *c64;   // cluster member order
    const AtomF32 *c32;
    const ClusterF64 *h64;
    const ClusterF32 *h32;
    // FP32 cluster SoA of the mixed-precision clustered kernel (ncl each):
    // cg0 (x, y, z, (R_m + rho)^2), cg1 (charge, dipole; the dipole zero when
    // the centre was replaced or the order is 0), cg2 ((R_d + rho)^2 and the
    // band branch's LJ constants A, B), cse (first, end) of the members in
    // cluster order; cxq / cab: the members' (x, y, z, q) and LJ constants.
    const float4 *cg0, *cg1, *cg2;
    const int2 *cse;
    const float4 *cxq;
    const float2 *cab;
    // bit c % 64 of cdm[c / 64]: cluster c has a nonzero dipole in cg1 (for
    // the branch counts of the far field)
    const unsigned long long *cdm;
    // FP64 cluster SoA of the FP64 one-group-a-wavefront kernel, 4 doubles a
    // cluster each: dg0 (x, y, z, (R_m + rho)^2), dg1 (charge, dipole; zero as
    // in cg1), dg2 ((R_d + rho)^2, the band branch's sigma^6 and 24 eps/mu, 0)
    const double *dg0, *dg1, *dg2;
    double cl_dpr_max64;  // largest (R_d + rho)^2 of dg2
    float cl_dpr_max;     // largest (R_d + rho)^2 of cg2 (FP32)
    int cl_ndip, cl_ndip64;  // clusters with a nonzero dipole in cg1 (the bits of cdm), in dg1
};

// Outside the boundary ellipsoid (src/diffEqSolver.cpp:99, 174): the
// reference's |p / (a, b, c)|^2 > 1 with projection, > max(a, b, c) without
// (its test as written).  The products with the reciprocals carry a few ulp
// of error against the quotients, so they decide wherever they lie more than
// 1e-12 (relative) from the threshold; closer, the quotients decide.  The
// answer is always the divisions' answer (outside_div), without three FP64
// divisions a step.  The FP64 group kernels keep outside_div: at the edge of
// 128 VGPRs, the extra live values of the fast path would spill.
__device__ __forceinline__ bool outside_div(const MolDev &m, double px, double py, double pz)
{
    const double tx = px / m.a, ty = py / m.b, tz = pz / m.c;
    const double tm2 = tx * tx + ty * ty + tz * tz;
    return (tm2 > 1 && m.projection) || (tm2 > m.maxrad && !m.projection);
}
__device__ __forceinline__ bool outside(const MolDev &m, double px, double py, double pz)
{
    const double thr = m.projection ? 1.0 : m.maxrad;
    const double ux = px * m.ia, uy = py * m.ib, uz = pz * m.ic;
    const double q = ux * ux + uy * uy + uz * uz;
    if (fabs(q - thr) > 1e-12 * thr) return q > thr;
    return outside_div(m, px, py, pz);
}

template <typename T>
__device__ __forceinline__ const CAS T *cas(const T *p)
{
    return (const CAS T *)p;
}

// ---------------------------------------------------------------------------
// Arithmetic policy: FP64 is IEEE (correctly rounded div and sqrt, no
// contraction: the translation unit is built with -ffp-contract=off);
// FP32 uses the hardware reciprocal / square root.
template <typename T> struct Ops;
template <> struct Ops<double> {
    static __device__ __forceinline__ double div(double a, double b) { return a / b; }
    static __device__ __forceinline__ double sqrt_(double x) { return __builtin_sqrt(x); }
};
template <> struct Ops<float> {
    static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
};
template <typename T> struct Layout;
template <> struct Layout<double> {
    static __device__ __forceinline__ const CAS AtomF64 *atoms(const MolDev &m) { return cas(m.a64); }
    static __device__ __forceinline__ const CAS AtomF64 *catoms(const MolDev &m) { return cas(m.c64); }
    static __device__ __forceinline__ const CAS ClusterF64 *heads(const MolDev &m) { return cas(m.h64); }
};
template <> struct Layout<float> {
    static __device__ __forceinline__ const CAS AtomF32 *atoms(const MolDev &m) { return cas(m.a32); }
    static __device__ __forceinline__ const CAS AtomF32 *catoms(const MolDev &m) { return cas(m.c32); }
    static __device__ __forceinline__ const CAS ClusterF32 *heads(const MolDev &m) { return cas(m.h32); }
};

// ---------------------------------------------------------------------------
// Reference-form evaluators (operand order of src/forceAccel.cpp).

// chargeDipoleLJ, type 4  (src/forceAccel.cpp:751-805)
template <typename T>
__device__ __forceinline__ void acc_type4(const MolDev &m, T px, T py, T pz, T &ox, T &oy, T &oz)
{
    using O = Ops<T>;
    const auto A = Layout<T>::atoms(m);
    T sXX = 0, sYY = 0, sZZ = 0, sXY = 0, sXZ = 0, sYZ = 0, dX = 0, dY = 0, dZ = 0, tx = 0, ty = 0, tz = 0;
#pragma unroll 2
    for (int n = 0; n < m.n; n++) {
        T xD = px - A[n].x, yD = py - A[n].y, zD = pz - A[n].z;
        T x2 = xD * xD, y2 = yD * yD, z2 = zD * zD;
        T r2 = x2 + y2 + z2;
        T r = O::sqrt_(r2);
        T r3 = r2 * r;
        T r_3 = O::div(T(1), r3);
        T r_5 = O::div(r_3, r2);
        T r6 = r2 * r2 * r2;
        T q = A[n].q;
        r_3 *= q;
        r_5 *= q;
        T t2 = O::div(A[n].s6, r6);
        T t1 = T(2) * t2 * t2;
        T sc = O::div(A[n].e24 * (t1 - t2), r2);
        tx += sc * xD; ty += sc * yD; tz += sc * zD;
        dX += xD * r_3; dY += yD * r_3; dZ += zD * r_3;
        sXX += (T(3) * x2 - r2) * r_5;
        sYY += (T(3) * y2 - r2) * r_5;
        sZZ += (T(3) * z2 - r2) * r_5;
        sXY += T(3) * xD * yD * r_5;
        sXZ += T(3) * xD * zD * r_5;
        sYZ += T(3) * yD * zD * r_5;
    }
    const T ds = (T)m.dipole_scalar;
    tx -= ds * (sXX * dX + sXY * dY + sXZ * dZ);
    ty -= ds * (sXY * dX + sYY * dY + sYZ * dZ);
    tz -= ds * (sXZ * dX + sYZ * dY + sZZ * dZ);
    ox = tx; oy = ty; oz = tz;
}

// dispersionFull, type 6  (src/forceAccel.cpp:856-884)
template <typename T>
__device__ __forceinline__ void acc_type6(const MolDev &m, T px, T py, T pz, T &ox, T &oy, T &oz)
{
    using O = Ops<T>;
    const auto A = Layout<T>::atoms(m);
    T tx = 0, ty = 0, tz = 0;
#pragma unroll 2
    for (int n = 0; n < m.n; n++) {
        T xD = px - A[n].x, yD = py - A[n].y, zD = pz - A[n].z;
        T r2 = xD * xD + yD * yD + zD * zD;
        T r6 = r2 * r2 * r2;
        T t2 = O::div(A[n].s6, r6);
        T t1 = T(2) * t2 * t2;
        T sc = O::div(A[n].e24 * (t1 - t2), r2);
        tx += sc * xD; ty += sc * yD; tz += sc * zD;
    }
    ox = tx; oy = ty; oz = tz;
}

// Branch counts of the clustered types' force evaluations, this lane's
// share: clusters examined, far clusters (multipole terms), far clusters with
// dipole terms, member atoms with charges and LJ, with charges only, with
// LJ only.  They price a launch in the reference's FLOPs per branch.
struct ClCount {
    unsigned hdr = 0, far = 0, fdip = 0, exlj = 0, exq = 0, lj = 0;
};

template <typename T>
struct ClAcc {
    T fX = 0, fY = 0, fZ = 0, sXY = 0, sXZ = 0, sYZ = 0, aXY = 0, aXZ = 0, aYZ = 0, dXX = 0, dYY = 0, dZZ = 0,
      sumX = 0, sumY = 0, sumZ = 0;
};

// 1 / sqrt(r2) and 1 / x in FP64 for the group kernels: v_rsq_f64 / v_rcp_f64
// refined by two Newton steps (within a few ulp of the IEEE quotients).  The
// group kernels sum in lane-split order and are not bit-identical to the
// reference anyway; the lane kernel and cosims_accel keep the IEEE
// divisions and square roots of the reference.
__device__ __forceinline__ double rsqrt64(double r2)
{
    double y = __builtin_amdgcn_rsq(r2);
#pragma unroll
    for (int k = 0; k < 2; k++) {
        // y += y (1 - r2 y^2) / 2
        const double e = __builtin_fma(-(r2 * y), y, 1.0);
        y = __builtin_fma(0.5 * y, e, y);
    }
    return y;
}
__device__ __forceinline__ double rcp64(double x)
{
    double y = __builtin_amdgcn_rcp(x);
#pragma unroll
    for (int k = 0; k < 2; k++) {
        // y += y (1 - x y)
        const double e = __builtin_fma(-x, y, 1.0);
        y = __builtin_fma(y, e, y);
    }
    return y;
}

// exact charges (+ LJ) over a cluster's members, src/forceAccel.cpp:366-403 / :408-436
// FAST (FP64 group kernels): 1 / r from rsqrt64 instead of the IEEE
// quotients, the same terms otherwise.
template <typename T, bool LJ, bool FAST = false>
__device__ __forceinline__ void cl_exact(const MolDev &m, int s, int e, T px, T py, T pz, ClAcc<T> &A)
{
    using O = Ops<T>;
    constexpr bool kF = FAST && sizeof(T) == 8;
    const auto C = Layout<T>::catoms(m);
    for (int k = s; k < e; k++) {
        T charge = C[k].q;
        T d0 = px - C[k].x, d1 = py - C[k].y, d2 = pz - C[k].z;
        T r2 = d0 * d0 + d1 * d1 + d2 * d2;
        T ir2 = 0, r_5;
        if constexpr (kF) {
            const T ir = rsqrt64(r2);
            ir2 = ir * ir;
            r_5 = ir2 * ir2 * ir;
        } else {
            r_5 = O::div(T(1), r2 * r2 * O::sqrt_(r2));
        }
        T r_3 = r_5 * r2;
        r_3 *= charge;
        r_5 *= charge;
        A.fX += d0 * r_3; A.fY += d1 * r_3; A.fZ += d2 * r_3;
        A.dXX += (r2 - T(3) * d0 * d0) * r_5;
        A.dYY += (r2 - T(3) * d1 * d1) * r_5;
        A.dZZ += (r2 - T(3) * d2 * d2) * r_5;
        A.sXY -= T(3) * d0 * d1 * r_5;
        A.sXZ -= T(3) * d0 * d2 * r_5;
        A.sYZ -= T(3) * d1 * d2 * r_5;
        if (LJ) {
            T r6 = r2 * r2 * r2;
            T t2 = kF ? C[k].s6 * (ir2 * ir2 * ir2) : O::div(C[k].s6, r6);
            T t1 = T(2) * t2 * t2;
            T sc = kF ? C[k].e24 * (t1 - t2) * ir2 : O::div(C[k].e24 * (t1 - t2), r2);
            A.sumX += sc * d0; A.sumY += sc * d1; A.sumZ += sc * d2;
        }
    }
}

// LJ over a cluster's members with per-atom or fixed parameters,
// src/forceAccel.cpp:341-360 (fixed = data[cluster index]) and :563-582
template <typename T, bool FIXED, bool FAST = false>
__device__ __forceinline__ void cl_lj(const MolDev &m, int s, int e, T fs6, T fe24, T px, T py, T pz, ClAcc<T> &A)
{
    using O = Ops<T>;
    constexpr bool kF = FAST && sizeof(T) == 8;
    const auto C = Layout<T>::catoms(m);
    for (int k = s; k < e; k++) {
        T d0 = px - C[k].x, d1 = py - C[k].y, d2 = pz - C[k].z;
        T r2 = d0 * d0 + d1 * d1 + d2 * d2;
        T t2, sc;
        if constexpr (kF) {
            const T ir2 = rcp64(r2);
            t2 = (FIXED ? fs6 : (T)C[k].s6) * (ir2 * ir2 * ir2);
            const T t1 = T(2) * t2 * t2;
            sc = (FIXED ? fe24 : (T)C[k].e24) * (t1 - t2) * ir2;
        } else {
            T r6 = r2 * r2 * r2;
            t2 = O::div(FIXED ? fs6 : (T)C[k].s6, r6);
            T t1 = T(2) * t2 * t2;
            sc = O::div((FIXED ? fe24 : (T)C[k].e24) * (t1 - t2), r2);
        }
        A.sumX += sc * d0; A.sumY += sc * d1; A.sumZ += sc * d2;
    }
}

// The same two per-atom loops with the variant chosen at run time (used
// where lanes handle clusters of different kinds side by side).
template <typename T, bool FAST = false>
__device__ __forceinline__ void cl_exact_rt(const MolDev &m, int s, int e, bool lj, T px, T py, T pz, ClAcc<T> &A)
{
    if (lj) cl_exact<T, true, FAST>(m, s, e, px, py, pz, A);
    else cl_exact<T, false, FAST>(m, s, e, px, py, pz, A);
}
template <typename T, bool FAST = false>
__device__ __forceinline__ void cl_lj_rt(const MolDev &m, int s, int e, bool fixed, T fs6, T fe24, T px, T py, T pz,
                                         ClAcc<T> &A)
{
    if (fixed) cl_lj<T, true, FAST>(m, s, e, fs6, fe24, px, py, pz, A);
    else cl_lj<T, false, FAST>(m, s, e, T(0), T(0), px, py, pz, A);
}

// far-field multipole of one cluster, src/forceAccel.cpp:258-337
template <typename T, bool FAST = false>
__device__ __forceinline__ void cl_far(const MolDev &m, const CAS ClusterT<T> &H, T dx0, T dx1, T dx2, T r2,
                                       ClAcc<T> &A)
{
    using O = Ops<T>;
    const int order = m.order;
    const bool dip = order >= 1 && !H.rep;
    T xx = dx0 * dx0, yy = dx1 * dx1, zz = dx2 * dx2;
    T r_2, r_1;
    if constexpr (FAST && sizeof(T) == 8) {
        r_1 = rsqrt64(r2);
        r_2 = r_1 * r_1;
    } else {
        r_2 = O::div(T(1), r2);
        r_1 = O::div(T(1), O::sqrt_(r2));
    }
    T r_3 = r_2 * r_1;
    T r_5 = r_3 * r_2;
    T r_7 = 0, const1 = 0, const2 = 0, tM0 = 0, tM1 = 0, tM2 = 0;
    T xy = dx0 * dx1, xz = dx0 * dx2, yz = dx1 * dx2;
    T cR5 = H.q * r_5;
    T cR3 = cR5 * r2;
    if (dip) {
        r_7 = r_5 * r_2;
        const1 = T(3) * (H.d0 * dx0 + H.d1 * dx1 + H.d2 * dx2);
    }
    if (order >= 2) {
        r_7 = r_5 * r_2;
        tM0 = (H.Q[0] * dx0 + H.Q[3] * dx1 + H.Q[6] * dx2) * r_7;
        tM1 = (H.Q[1] * dx0 + H.Q[4] * dx1 + H.Q[7] * dx2) * r_7;
        tM2 = (H.Q[2] * dx0 + H.Q[5] * dx1 + H.Q[8] * dx2) * r_7;
        const2 = T(5) * (tM0 * dx0 + tM1 * dx1 + tM2 * dx2) * r_2 * T(0.5);
    }
    A.fX += cR3 * dx0; A.fY += cR3 * dx1; A.fZ += cR3 * dx2;
    A.dXX += (r2 - T(3) * xx) * cR5;
    A.dYY += (r2 - T(3) * yy) * cR5;
    A.dZZ += (r2 - T(3) * zz) * cR5;
    A.sXY -= T(3) * xy * cR5;
    A.sXZ -= T(3) * xz * cR5;
    A.sYZ -= T(3) * yz * cR5;
    if (dip) {
        A.fX += (const1 * dx0 - r2 * H.d0) * r_5;
        A.fY += (const1 * dx1 - r2 * H.d1) * r_5;
        A.fZ += (const1 * dx2 - r2 * H.d2) * r_5;
        A.dXX += (r2 - T(5) * xx) * r_7 * const1;
        A.dYY += (r2 - T(5) * yy) * r_7 * const1;
        A.dZZ += (r2 - T(5) * zz) * r_7 * const1;
        A.sXY -= T(5) * xy * r_7 * const1;
        A.sXZ -= T(5) * xz * r_7 * const1;
        A.sYZ -= T(5) * yz * r_7 * const1;
        A.aXY += T(3) * (H.d1 * dx0 - H.d0 * dx1) * r_5;
        A.aXZ += T(3) * (H.d2 * dx0 - H.d0 * dx2) * r_5;
        A.aYZ += T(3) * (H.d2 * dx1 - H.d1 * dx2) * r_5;
    }
    if (order >= 2) {
        A.fX += (const2 * dx0 - tM0) * r2;
        A.fY += (const2 * dx1 - tM1) * r2;
        A.fZ += (const2 * dx2 - tM2) * r2;
        A.dXX += ((r2 - T(7) * xx) * const2 - H.Q[0] * r_5);
        A.dYY += ((r2 - T(7) * yy) * const2 - H.Q[4] * r_5);
        A.dZZ += ((r2 - T(7) * zz) * const2 - H.Q[8] * r_5);
        A.sXY -= (T(7) * xy * const2 + H.Q[1] * r_5);
        A.sXZ -= (T(7) * xz * const2 + H.Q[2] * r_5);
        A.sYZ -= (T(7) * yz * const2 + H.Q[5] * r_5);
        A.aXY += T(5) * (tM1 * dx0 - tM0 * dx1);
        A.aXZ += T(5) * (tM2 * dx0 - tM0 * dx2);
        A.aYZ += T(5) * (tM2 * dx1 - tM1 * dx2);
    }
}

// One cluster of the clustered force types: the loop bodies of
// src/forceAccel.cpp:244-438 (type 1, multipoleCutoff), :470-626 (type 2,
// multipoleOnly), :655-742 (type 3, dispersionOnly), :819-851 (type 5,
// dispersionNoCharge).
// With CNT the branches taken are counted into C.
template <typename T, int FT, bool CNT = false, bool FAST = false>
__device__ __forceinline__ void cl_one(const MolDev &m, const CAS ClusterT<T> &H, T px, T py, T pz, ClAcc<T> &A,
                                       ClCount &C)
{
    T dx0 = px - H.px, dx1 = py - H.py, dx2 = pz - H.pz;
    const unsigned na = (unsigned)(H.end - H.start);
    if (CNT) C.hdr++;
    if constexpr (FT == 1 || FT == 2) {
        T xx = dx0 * dx0, yy = dx1 * dx1, zz = dx2 * dx2;
        T r2 = xx + yy + zz;
        const int s = H.start, e = H.end;
        if (r2 > H.mpr) {
            cl_far<T, FAST>(m, H, dx0, dx1, dx2, r2, A);
            if (CNT) {
                C.far++;
                C.fdip += (m.order >= 1 && !H.rep) ? 1u : 0u;
            }
            if (FT == 1) {
                if (!(r2 > H.dpr)) {
                    cl_lj<T, true, FAST>(m, s, e, H.qs6, H.qe24, px, py, pz, A);
                    if (CNT) C.lj += na;
                }
            } else {
                cl_lj<T, false, FAST>(m, s, e, T(0), T(0), px, py, pz, A);
                if (CNT) C.lj += na;
            }
        } else if (FT == 2 || !(r2 > H.dpr)) {
            cl_exact<T, true, FAST>(m, s, e, px, py, pz, A);
            if (CNT) C.exlj += na;
        } else {
            cl_exact<T, false, FAST>(m, s, e, px, py, pz, A);
            if (CNT) C.exq += na;
        }
    } else {
        T r2 = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
        if constexpr (FT == 3) {
            if (!(r2 > H.dpr)) {
                cl_exact<T, true, FAST>(m, H.start, H.end, px, py, pz, A);
                if (CNT) C.exlj += na;
            } else {
                cl_exact<T, false, FAST>(m, H.start, H.end, px, py, pz, A);
                if (CNT) C.exq += na;
            }
        } else {
            if (!(r2 > H.dpr)) {
                cl_lj<T, false, FAST>(m, H.start, H.end, T(0), T(0), px, py, pz, A);
                if (CNT) C.lj += na;
            }
        }
    }
}

// Polarisation epilogue of the clustered types (src/forceAccel.cpp:440-442
// and :628-630 with the antisymmetric aMat part, :744-746 without it, none
// for type 5).
template <typename T, int FT>
__device__ __forceinline__ void cl_epilogue(const MolDev &m, ClAcc<T> &A, T &ox, T &oy, T &oz)
{
    const T ds = (T)m.dipole_scalar;
    if constexpr (FT == 1 || FT == 2) {
        A.sumX += ds * (A.dXX * A.fX + (A.sXY + A.aXY) * A.fY + (A.sXZ + A.aXZ) * A.fZ);
        A.sumY += ds * ((A.sXY - A.aXY) * A.fX + A.dYY * A.fY + (A.sYZ + A.aYZ) * A.fZ);
        A.sumZ += ds * ((A.sXZ - A.aXZ) * A.fX + (A.sYZ - A.aYZ) * A.fY + A.dZZ * A.fZ);
    } else if constexpr (FT == 3) {
        A.sumX += ds * (A.dXX * A.fX + A.sXY * A.fY + A.sXZ * A.fZ);
        A.sumY += ds * (A.sXY * A.fX + A.dYY * A.fY + A.sYZ * A.fZ);
        A.sumZ += ds * (A.sXZ * A.fX + A.sYZ * A.fY + A.dZZ * A.fZ);
    }
    ox = A.sumX; oy = A.sumY; oz = A.sumZ;
}

// Clustered force types over clusters first, first+stride, ...: stride 1 is
// the reference's cluster order; the group kernel splits clusters over the
// lanes of a group and sums the partial accumulators afterwards.
template <typename T, int FT, bool CNT = false, bool FAST = false>
__device__ __forceinline__ void cl_sum(const MolDev &m, int first, int stride, T px, T py, T pz, ClAcc<T> &A,
                                       ClCount &C)
{
    const auto H = Layout<T>::heads(m);
    for (int n = first; n < m.ncl; n += stride) cl_one<T, FT, CNT, FAST>(m, H[n], px, py, pz, A, C);
}

template <typename T, int FT>
__device__ __forceinline__ void acc_clustered(const MolDev &m, T px, T py, T pz, T &ox, T &oy, T &oz)
{
    ClAcc<T> A;
    ClCount C;
    cl_sum<T, FT>(m, 0, 1, px, py, pz, A, C);
    cl_epilogue<T, FT>(m, A, ox, oy, oz);
}

// ---------------------------------------------------------------------------
// FP32 fast evaluators for the exact-sum types.  Same physics as the
// reference form, rearranged for the VALU: one v_rsq_f32 per pair, the
// field-gradient tensor accumulated as S_ab = sum q d_a d_b / r^5 with
// T_ab = 3 S_ab - delta_ab tr(S) formed once after the loop.
#pragma clang fp contract(fast)
__device__ __forceinline__ void acc4_fast(const MolDev &m, float px, float py, float pz, double &ox, double &oy,
                                          double &oz)
{
    const CAS AtomF32 *A = cas(m.a32);
    float fx = 0, fy = 0, fz = 0, ex = 0, ey = 0, ez = 0;
    float sxx = 0, syy = 0, szz = 0, sxy = 0, sxz = 0, syz = 0;
#pragma unroll 4
    for (int n = 0; n < m.n; n++) {
        float dx = px - A[n].x, dy = py - A[n].y, dz = pz - A[n].z;
        float r2 = dx * dx + dy * dy + dz * dz;
        float ir = __builtin_amdgcn_rsqf(r2);
        float ir2 = ir * ir;
        float ir6 = ir2 * ir2 * ir2;
        float t2 = A[n].s6 * ir6;
        float lj = A[n].e24 * ir2 * t2 * (2.0f * t2 - 1.0f);
        fx += lj * dx; fy += lj * dy; fz += lj * dz;
        float qr3 = A[n].q * ir2 * ir;
        ex += qr3 * dx; ey += qr3 * dy; ez += qr3 * dz;
        float qr5 = qr3 * ir2;
        float a = qr5 * dx, b = qr5 * dy, c = qr5 * dz;
        sxx += a * dx; syy += b * dy; szz += c * dz;
        sxy += a * dy; sxz += a * dz; syz += b * dz;
    }
    double tr = (double)sxx + (double)syy + (double)szz;
    double Txx = 3.0 * sxx - tr, Tyy = 3.0 * syy - tr, Tzz = 3.0 * szz - tr;
    double Txy = 3.0 * sxy, Txz = 3.0 * sxz, Tyz = 3.0 * syz;
    const double ds = m.dipole_scalar;
    ox = (double)fx - ds * (Txx * ex + Txy * ey + Txz * ez);
    oy = (double)fy - ds * (Txy * ex + Tyy * ey + Tyz * ez);
    oz = (double)fz - ds * (Txz * ex + Tyz * ey + Tzz * ez);
}

__device__ __forceinline__ void acc6_fast(const MolDev &m, float px, float py, float pz, double &ox, double &oy,
                                          double &oz)
{
    const CAS AtomF32 *A = cas(m.a32);
    float fx = 0, fy = 0, fz = 0;
#pragma unroll 4
    for (int n = 0; n < m.n; n++) {
        float dx = px - A[n].x, dy = py - A[n].y, dz = pz - A[n].z;
        float r2 = dx * dx + dy * dy + dz * dz;
        float ir2 = __builtin_amdgcn_rcpf(r2);
        float ir6 = ir2 * ir2 * ir2;
        float t2 = A[n].s6 * ir6;
        float lj = A[n].e24 * ir2 * t2 * (2.0f * t2 - 1.0f);
        fx += lj * dx; fy += lj * dy; fz += lj * dz;
    }
    ox = fx; oy = fy; oz = fz;
}
#pragma clang fp contract(off)

// Force dispatch: PREC 0 = FP64 reference form, 1 = FP32 sum.
template <int PREC, int FT>
__device__ __forceinline__ void accel(const MolDev &m, double px, double py, double pz, double &ox, double &oy,
                                      double &oz)
{
    if constexpr (PREC == 0) {
        if constexpr (FT == 4) acc_type4<double>(m, px, py, pz, ox, oy, oz);
        else if constexpr (FT == 6) acc_type6<double>(m, px, py, pz, ox, oy, oz);
        else acc_clustered<double, FT>(m, px, py, pz, ox, oy, oz);
    } else {
        const float fx = (float)px, fy = (float)py, fz = (float)pz;
        if constexpr (FT == 4) acc4_fast(m, fx, fy, fz, ox, oy, oz);
        else if constexpr (FT == 6) acc6_fast(m, fx, fy, fz, ox, oy, oz);
        else {
            float a, b, c;
            acc_clustered<float, FT>(m, fx, fy, fz, a, b, c);
            ox = a; oy = b; oz = c;
        }
    }
}

// getPotential, FP64 reference form (src/forceAccel.cpp:943-982)
__device__ __forceinline__ double potential(const MolDev &m, double px, double py, double pz)
{
    const CAS AtomF64 *A = cas(m.a64);
    double total = 0, sX = 0, sY = 0, sZ = 0;
    if (m.charged) {
#pragma unroll 2
        for (int n = 0; n < m.n; n++) {
            double xD = px - A[n].x, yD = py - A[n].y, zD = pz - A[n].z;
            double r2 = xD * xD + yD * yD + zD * zD;
            double r6 = r2 * r2 * r2;
            double t2 = A[n].s6 / r6;
            double t1 = t2 * t2;
            total += A[n].e4 * (t1 - t2);
            double r_3 = A[n].q / (r2 * __builtin_sqrt(r2));
            sX += xD * r_3; sY += yD * r_3; sZ += zD * r_3;
        }
        total -= m.dipole_scalar * 0.5 * (sX * sX + sY * sY + sZ * sZ);
    } else {
#pragma unroll 2
        for (int n = 0; n < m.n; n++) {
            double xD = px - A[n].x, yD = py - A[n].y, zD = pz - A[n].z;
            double r2 = xD * xD + yD * yD + zD * zD;
            double r6 = r2 * r2 * r2;
            double t2 = A[n].s6 / r6;
            double t1 = t2 * t2;
            total += A[n].e4 * (t1 - t2);
        }
    }
    return total;
}

}  // namespace cosims
