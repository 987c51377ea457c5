// traj_group.hpp -- group-per-trajectory persistent kernels (gfx950).
//
// L lanes (a power of two, 4..64) integrate one trajectory together: every
// lane of the group holds the same FP64 Verlet state and takes atoms (or
// clusters) j, j+L, j+2L, ... of the force sum; a butterfly of lane shuffles
// gives every lane the full sums.  A wavefront therefore carries 64/L
// trajectories and a step costs N/L pairs per lane instead of N, so a long
// trajectory no longer holds a whole launch hostage.  The energy-conservation
// test (src/diffEqSolver.cpp:108, 178-203) runs inside the kernel as an FP64
// getPotential across the group, and a failed attempt restarts at once with
// dt/(k+1): one launch covers every attempt of every trajectory.
//
// Three kernels share the trajectory state machine (group_loop):
//   k_traj_grp    one molecule, staged once per workgroup in LDS (SoA: FP32
//                 x,y,z,q | LJ constants A, B as float4 + float2, or the FP64
//                 x,y,z,q,sigma^6,24 eps/mu), exact force types 4 and 6;
//   k_traj_tile   one molecule too large for LDS, exact types, streamed
//                 through LDS in tiles by the whole workgroup, which
//                 integrates in lock-step;
//   k_traj_multi  any number of molecules of one force type, read from HBM
//                 through L2 (a table of MolRec, one molecule per trajectory):
//                 molecules of uneven size in one launch, and the clustered
//                 force types 1, 2, 3, 5 (lanes split the clusters).
// Atom arrays are padded to a multiple of 64 with inert atoms (charge and LJ
// parameters zero, far away) so every lane of a group runs the same trip
// count.
#pragma once

#include "kernels.hpp"
#include <type_traits>

#define GLB __attribute__((address_space(1)))

namespace cosims {

template <typename T>
__device__ __forceinline__ const GLB T *glb(const T *p)
{
    return (const GLB T *)p;
}

// Native vector types for the FP32 SoA (HIP's float4 / float2 class types
// cannot be copied out of an explicit address space).
typedef float f4n __attribute__((ext_vector_type(4)));
typedef float f2n __attribute__((ext_vector_type(2)));
typedef int i2n __attribute__((ext_vector_type(2)));

// Per-trajectory arrays of one persistent launch.
struct TrajIO {
    const int32_t *items;
    int32_t n_items;
    int32_t *counter;
    const double *p0, *v0;
    int8_t *status;
    double *chi, *cchi, *ergy;             // cchi: cos(chi) of accepted trajectories
    int32_t *steps, *totsteps;
    int32_t max_steps;
    unsigned long long *evals;             // [0] steps, [1] force evaluations, [2] atom pairs of the
                                           // potential (getPotential) evaluations,
                                           // [3] atom pairs of the exact force types' evaluations,
                                           // [4..9] clustered-type branch counts (ClCount order)
};
constexpr int kEvalSlots = 10;             // counters of one launch in TrajIO::evals

// Branch counts of one wavefront's clustered force evaluations (the 64-lane
// clustered kernels), kept in LDS so that they hold no registers: row k
// (ClCount order) has an entry per lane; rows 0-2 (clusters examined, far,
// far with dipole terms) are counted wave-wide into lane 0's entry, rows
// 3-5 (member atoms by branch) by each lane into its own.
constexpr int kCountInts = 6 * 64;  // LDS ints per wavefront
struct WaveCount {
    unsigned *row;
    __device__ void clear() const
    {
        const int lane = threadIdx.x & 63;
#pragma unroll
        for (int k = 0; k < 6; k++) row[k * 64 + lane] = 0;
    }
    __device__ void lane_add(int k, unsigned v) const { row[k * 64 + (threadIdx.x & 63)] += v; }
    __device__ void wave_add(int k, unsigned v) const  // v wave-uniform
    {
        if ((threadIdx.x & 63) == 0) row[k * 64] += v;
    }
};

// Adds this wavefront's six branch counts to evals[4..9] (all lanes call it).
__device__ __forceinline__ void flush_counts(unsigned long long *evals, const WaveCount &W)
{
    unsigned long long v[6];
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] = W.row[k * 64 + (threadIdx.x & 63)];
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < 6; k++) v[k] += __shfl_xor(v[k], off);
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < 6; k++)
            if (v[k]) atomicAdd(evals + 4 + k, v[k]);
    }
}

struct GroupArgs {
    TrajIO io;
    MolDev m;
    int npad;                              // atoms incl. padding, multiple of 64
    const double *soa;                     // global FP64 SoA: x y z q s6 e24 e4, npad each
    double dt0;
    // when set, the work list is io.items[*first .. *last) (first NULL: from
    // 0), bounds written on the device before the launch runs
    const int32_t *first, *last;
    // k_traj_tile: the FP32 SoA in global memory and the atoms per LDS tile
    // (a multiple of 64)
    const float4 *xq;
    const float2 *se;
    int tile;
};

// One molecule of a multi-molecule launch.
struct MolRec {
    MolDev m;
    const float4 *xq;                      // FP32 x, y, z, q (npad)
    const float2 *se;                      // FP32 LJ constants A, B of grp_force32 (npad)
    const double *soa;                     // FP64 SoA as in GroupArgs
    int32_t npad, pad_;
    double dt0;
};

struct MultiArgs {
    TrajIO io;
    const MolRec *mols;
    const int32_t *tmol;                   // molecule of each trajectory (NULL: all molecule 0)
    // when set, the work list is io.items[*first .. *last) (first NULL: from 0),
    // bounds written on the device by the sampler before the launch runs
    const int32_t *first, *last;
};

// Work-list bounds held on the device (see MultiArgs::first / last).
__device__ __forceinline__ void resolve_range(TrajIO &io, const int32_t *first, const int32_t *last)
{
    if (!last) return;
    const int b = first ? *first : 0;
    io.items += b;
    io.n_items = *last - b;
}

// chi = vector3D::Angle of the accepted exit (src/vector3D.h:134-143) from
// the stored cosine, for the items of one group launch (items[0..n), or
// items[*first .. *last) when last is set).
__global__ void __launch_bounds__(256) k_chi(const int32_t *items, int n, const int8_t *status, const double *cchi,
                                             double *chi, const int32_t *first, const int32_t *last)
{
    if (last) {
        const int b = first ? *first : 0;
        items += b;
        n = *last - b;
    }
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int t = items[i];
    if (status[t] != 1) return;
    const double t3 = cchi[t];
    chi[t] = (t3 > 1) ? 0.0 : ((t3 < -1) ? 3.14159265358979323846 : acos(t3));
}

// Sum over the L lanes of a group (aligned, L a power of two <= 64), the
// result in every lane, on the VALU: DPP quad permutations, half-row and row
// mirrors (within 16 lanes), then the gfx950 row and half-wave swaps.  Each
// step adds two operands that every lane of the group sees in the same
// order, so all lanes hold bitwise-identical sums (every lane carries the
// same Verlet state).  Needs all 64 lanes active.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_step(float v)
{
    return v + __builtin_bit_cast(float, dpp_u32<CTRL>(__builtin_bit_cast(uint32_t, v)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_step(double v)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint64_t w = (uint64_t)dpp_u32<CTRL>((uint32_t)u) | ((uint64_t)dpp_u32<CTRL>((uint32_t)(u >> 32)) << 32);
    return v + __builtin_bit_cast(double, w);
}
__device__ __forceinline__ float swap_step(float v, bool half_wave)
{
    const uint32_t u = __builtin_bit_cast(uint32_t, v);
    const auto r = half_wave ? __builtin_amdgcn_permlane32_swap(u, u, false, false)
                             : __builtin_amdgcn_permlane16_swap(u, u, false, false);
    return __builtin_bit_cast(float, (uint32_t)r[0]) + __builtin_bit_cast(float, (uint32_t)r[1]);
}
__device__ __forceinline__ double swap_step(double v, bool half_wave)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)u, hi = (uint32_t)(u >> 32);
    const auto a = half_wave ? __builtin_amdgcn_permlane32_swap(lo, lo, false, false)
                             : __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto b = half_wave ? __builtin_amdgcn_permlane32_swap(hi, hi, false, false)
                             : __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    const double x = __builtin_bit_cast(double, (uint64_t)(uint32_t)a[0] | ((uint64_t)(uint32_t)b[0] << 32));
    const double y = __builtin_bit_cast(double, (uint64_t)(uint32_t)a[1] | ((uint64_t)(uint32_t)b[1] << 32));
    return x + y;
}
template <int L, typename T>
__device__ __forceinline__ T gsum(T v)
{
    static_assert(L >= 4 && L <= 64 && (L & (L - 1)) == 0, "group size");
    v = dpp_step<0xB1>(v);                       // quad_perm [1,0,3,2]
    v = dpp_step<0x4E>(v);                       // quad_perm [2,3,0,1]
    if constexpr (L >= 8) v = dpp_step<0x141>(v);   // row_half_mirror
    if constexpr (L >= 16) v = dpp_step<0x140>(v);  // row_mirror
    if constexpr (L >= 32) v = swap_step(v, false); // rows 2k <-> 2k+1
    if constexpr (L >= 64) v = swap_step(v, true);  // half-waves
    return v;
}

// FP32 pair kernel of the exact-sum types over atoms j, j+L, ... < npad,
// same physics as src/forceAccel.cpp:764-802 rearranged for the VALU (30
// instructions a pair): per atom x, y, z, q and the LJ constants
// A = 2 (24 eps/mu) sigma^12, B = (24 eps/mu) sigma^6, so the LJ factor is
// r^-8 (A r^-6 - B) (src/forceAccel.cpp:785-788: 24 eps/mu (2 t^2 - t) / r^2
// with t = sigma^6 r^-6); one v_rsq_f32; the field gradient accumulated as
// S_ab = sum q d_a d_b r^-5 (xx, yy, xy, xz, yz) and its trace sum q r^-3,
// with T = 3 S - tr(S) formed once after the loop (the reference sums
// (3 d_a d_b - r^2 delta_ab) q r^-5 term by term, :790-797).
#pragma clang fp contract(fast)
// Partial sums of one lane: LJ force 0-2, field 3-5, S xx yy xy xz yz 6-10,
// trace 11.  One pair (atom a = x, y, z, q; c = A, B) added in place; the
// LDS and the global-memory loops both use it, so their sums are bitwise
// identical.
template <int FT>
__device__ __forceinline__ void pair32(f4n a, f2n c, float px, float py, float pz, float (&s)[12])
{
    const float dx = px - a.x, dy = py - a.y, dz = pz - a.z;
    const float r2 = dx * dx + dy * dy + dz * dz;
    float ir2, ir;
    if constexpr (FT == 4) {
        ir = __builtin_amdgcn_rsqf(r2);
        ir2 = ir * ir;
    } else {
        ir2 = __builtin_amdgcn_rcpf(r2);
    }
    const float ir4 = ir2 * ir2;
    const float ir6 = ir4 * ir2;
    const float lj = (ir4 * ir4) * (c.x * ir6 - c.y);
    s[0] += lj * dx; s[1] += lj * dy; s[2] += lj * dz;
    if constexpr (FT == 4) {
        const float qr3 = a.w * ir2 * ir;
        s[3] += qr3 * dx; s[4] += qr3 * dy; s[5] += qr3 * dz;
        s[11] += qr3;
        const float qr5 = qr3 * ir2;
        const float u = qr5 * dx, v = qr5 * dy;
        s[6] += u * dx; s[7] += v * dy;
        s[8] += u * dy; s[9] += u * dz; s[10] += v * dz;
    }
}

// Group sums of the partials and the polarisation epilogue (FP64).  (Summing
// the field first and letting each lane apply it to its own gradient
// partials needs 6 group sums instead of 12, but its two dependent chains
// cost the lock-step tiled kernel 3-9 % and gained nothing measurable in the
// LDS kernel; all exact-type kernels share this one, so they stay bitwise
// identical.)
template <int FT, int L>
__device__ __forceinline__ void force32_epilogue(float (&s)[12], double ds, double &ox, double &oy, double &oz)
{
    const float fx = gsum<L>(s[0]), fy = gsum<L>(s[1]), fz = gsum<L>(s[2]);
    if constexpr (FT == 4) {
        const float ex = gsum<L>(s[3]), ey = gsum<L>(s[4]), ez = gsum<L>(s[5]);
        const float sxx = gsum<L>(s[6]), syy = gsum<L>(s[7]), tr = gsum<L>(s[11]);
        const float sxy = gsum<L>(s[8]), sxz = gsum<L>(s[9]), syz = gsum<L>(s[10]);
        const double t = tr, Sxx = sxx, Syy = syy;
        const double Txx = 3.0 * Sxx - t, Tyy = 3.0 * Syy - t, Tzz = 2.0 * t - 3.0 * Sxx - 3.0 * Syy;
        const double Txy = 3.0 * sxy, Txz = 3.0 * sxz, Tyz = 3.0 * syz;
        ox = (double)fx - ds * (Txx * ex + Txy * ey + Txz * ez);
        oy = (double)fy - ds * (Txy * ex + Tyy * ey + Tyz * ez);
        oz = (double)fz - ds * (Txz * ex + Tyz * ey + Tzz * ez);
    } else {
        ox = fx; oy = fy; oz = fz;
    }
}

// The pairs of lane j with atoms j, j + L, ... < cnt in LDS (cnt a multiple
// of L), the next atom read into registers of its own while the current
// pair is computed: loaded where used, each LDS read waits out its latency
// in full, and the loop runs 10-12 % slower at 4-6 wavefronts a SIMD.
template <int FT, int L, typename P4, typename P2>
__device__ __forceinline__ void lds_pairs32(P4 xq, P2 ab, int cnt, int j, float px, float py, float pz,
                                            float (&s)[12])
{
    if (cnt <= 0) return;
    f4n a = xq[j];
    f2n c = ab[j];
#pragma unroll 4
    for (int kk = L; kk < cnt; kk += L) {
        const f4n an = xq[kk + j];
        const f2n cn = ab[kk + j];
        pair32<FT>(a, c, px, py, pz, s);
        a = an;
        c = cn;
    }
    pair32<FT>(a, c, px, py, pz, s);
}

// Atoms from LDS.
template <int FT, int L, typename P4, typename P2>
__device__ __forceinline__ void grp_force32(P4 xq, P2 ab, int npad, int j, float px, float py, float pz, double ds,
                                            double &ox, double &oy, double &oz)
{
    float s[12];
#pragma unroll
    for (int k = 0; k < 12; k++) s[k] = 0.0f;
    lds_pairs32<FT, L>(xq, ab, npad, j, px, py, pz, s);
    force32_epilogue<FT, L>(s, ds, ox, oy, oz);
}

// Atoms from global memory (L2, hundreds of cycles): kSlots pairs are in
// flight per lane.  Slot i holds pair kk + i L; as soon as its pair is added
// the slot is reloaded with the pair kSlots further on, so each load has the
// arithmetic of kSlots - 1 pairs to land and the slots never move between
// registers.  npad is a multiple of 64 and so of L: a slot past the end is
// skipped (the test is the same for the whole wavefront) and its loads
// re-read the lane's last pair (in bounds, unused).  Same pairs in the same
// order as grp_force32.
constexpr int kSlots = 3;

template <int FT, int L>
__device__ __forceinline__ void grp_force32_glb(const GLB f4n *xq, const GLB f2n *ab, int npad, int j, float px,
                                                float py, float pz, double ds, double &ox, double &oy, double &oz)
{
    float s[12];
#pragma unroll
    for (int k = 0; k < 12; k++) s[k] = 0.0f;
    if (npad > 0) {
        const int last = npad - L + j;
        f4n a[kSlots];
        f2n c[kSlots];
#pragma unroll
        for (int i = 0; i < kSlots; i++) {
            const int k = min(i * L + j, last);
            a[i] = xq[k];
            c[i] = ab[k];
        }
        for (int kk = 0; kk < npad; kk += kSlots * L) {
#pragma unroll
            for (int i = 0; i < kSlots; i++) {
                if (kk + i * L < npad) pair32<FT>(a[i], c[i], px, py, pz, s);
                const int kn = min(kk + (i + kSlots) * L + j, last);
                a[i] = xq[kn];
                c[i] = ab[kn];
            }
        }
    }
    force32_epilogue<FT, L>(s, ds, ox, oy, oz);
}
#pragma clang fp contract(off)

// FP32 LJ constants of one atom for grp_force32 (identical on host and device:
// IEEE double products, no contraction, then one rounding to float).
__host__ __device__ inline void lj_constants(double s6, double e24, float &A, float &B)
{
    A = (float)(2.0 * e24 * s6 * s6);
    B = (float)(e24 * s6);
}

// FP64 pair terms of the group kernels (src/forceAccel.cpp:764-797,
// 865-880) of one atom (x, y, z, q, sigma^6, 24 eps/mu), added to this
// lane's partial sums: LJ 0-2, field 3-5, and the field gradient's
// independent parts as in pair32 (q x^2 / r^5, q y^2 / r^5, q xy / r^5,
// q xz / r^5, q yz / r^5 in 6-10 and q / r^3 in 11: the zz part follows from
// the trace in the epilogue).  The reference's four IEEE divisions and
// square root a pair become products with ir = 1 / sqrt(r^2) (v_rsq_f64
// refined by two Newton steps, within a few ulp of the quotients; 1 / r^2
// from v_rcp_f64 the same way when no odd power is needed), the sums are
// fused multiply-adds, and the LJ term is e24 t (2 t - 1) / r^2 with t =
// sigma^6 / r^6: 39 FP64 operations a pair for force type 4 where the
// reference's form takes 64.  These kernels sum in lane-split order and are
// not bit-identical to the reference anyway; the lane kernel keeps the IEEE
// operations and the reference's order.
#pragma clang fp contract(fast)
template <int FT>
__device__ __forceinline__ void pair64(double ax, double ay, double az, double aq, double as6, double ae24,
                                       double px, double py, double pz, double (&s)[12])
{
    const double xD = px - ax, yD = py - ay, zD = pz - az;
    const double r2 = xD * xD + yD * yD + zD * zD;
    double ir = 0, ir2;
    if constexpr (FT == 4) {
        ir = rsqrt64(r2);
        ir2 = ir * ir;
    } else {
        ir2 = rcp64(r2);
    }
    const double t = as6 * (ir2 * ir2 * ir2);
    const double sc = ae24 * t * (2.0 * t - 1.0) * ir2;
    s[0] += sc * xD; s[1] += sc * yD; s[2] += sc * zD;
    if constexpr (FT == 4) {
        const double qr3 = aq * (ir * ir2);
        s[3] += xD * qr3; s[4] += yD * qr3; s[5] += zD * qr3;
        s[11] += qr3;
        const double qr5 = qr3 * ir2;
        const double u = qr5 * xD, v = qr5 * yD;
        s[6] += u * xD; s[7] += v * yD;
        s[8] += u * yD; s[9] += u * zD; s[10] += v * zD;
    }
}
#pragma clang fp contract(off)

// Group sums of the FP64 partials and the epilogue (:800-802): the gradient
// tensor T_ab = sum q (3 a b - r^2 delta_ab) / r^5 from the partials, its
// zz part from the trace (T_xx + T_yy + T_zz = 0).
template <int FT, int L>
__device__ __forceinline__ void force64_epilogue(double (&s)[12], double ds, double &ox, double &oy, double &oz)
{
    double tx = gsum<L>(s[0]), ty = gsum<L>(s[1]), tz = gsum<L>(s[2]);
    if constexpr (FT == 4) {
        const double dX = gsum<L>(s[3]), dY = gsum<L>(s[4]), dZ = gsum<L>(s[5]);
        const double Sxx = gsum<L>(s[6]), Syy = gsum<L>(s[7]), t = gsum<L>(s[11]);
        const double Sxy = gsum<L>(s[8]), Sxz = gsum<L>(s[9]), Syz = gsum<L>(s[10]);
        const double Txx = 3.0 * Sxx - t, Tyy = 3.0 * Syy - t, Tzz = 2.0 * t - 3.0 * Sxx - 3.0 * Syy;
        const double Txy = 3.0 * Sxy, Txz = 3.0 * Sxz, Tyz = 3.0 * Syz;
        tx -= ds * (Txx * dX + Txy * dY + Txz * dZ);
        ty -= ds * (Txy * dX + Tyy * dY + Tyz * dZ);
        tz -= ds * (Txz * dX + Tyz * dY + Tzz * dZ);
    }
    ox = tx; oy = ty; oz = tz;
}

// FP64 pair terms over atoms j, j + L, ... < npad of the SoA X (x y z q s6
// e24 with stride npad), then the group sums and the epilogue: pair64 and
// force64_epilogue, the same pairs in the same order as the tiled kernel's,
// so the LDS, table and tiled kernels give the same bits.  From LDS (X a
// generic pointer): npad is a multiple of 64, so every lane makes the same
// npad / L passes, and the next atom is read into registers of its own one
// pair ahead, so the reads of the unrolled pairs go out together.  From
// global memory (the table kernel), the plain loop; the tiled kernel, whose
// other live state leaves no room for the prefetch registers, has a loop of
// its own (two pairs a pass at 32 and 64 lanes, no registers held across
// passes).
template <int FT, int L, typename PD>
__device__ __forceinline__ void grp_force64(PD X, int npad, int j, double px, double py, double pz, double ds,
                                            double &ox, double &oy, double &oz)
{
    const PD Y = X + npad, Z = X + 2 * npad, Q = X + 3 * npad, S6 = X + 4 * npad, E24 = X + 5 * npad;
    double s[12];
#pragma unroll
    for (int k = 0; k < 12; k++) s[k] = 0.0;
    const int cnt = npad / L;
    if constexpr (!std::is_same<PD, const double *>::value) {
        for (int k = j; k < npad; k += L) pair64<FT>(X[k], Y[k], Z[k], Q[k], S6[k], E24[k], px, py, pz, s);
    } else if (cnt > 0) {
        double ax = X[j], ay = Y[j], az = Z[j], aq = Q[j], as = S6[j], ae = E24[j];
#pragma unroll 4
        for (int i = 1; i < cnt; i++) {
            const int k = j + i * L;
            const double bx = X[k], by = Y[k], bz = Z[k], bq = Q[k], bs = S6[k], be = E24[k];
            pair64<FT>(ax, ay, az, aq, as, ae, px, py, pz, s);
            ax = bx; ay = by; az = bz; aq = bq; as = bs; ae = be;
        }
        pair64<FT>(ax, ay, az, aq, as, ae, px, py, pz, s);
    }
    force64_epilogue<FT, L>(s, ds, ox, oy, oz);
}

// Clustered force types with the clusters split over the group's lanes:
// the reference's per-cluster terms (cl_one), partial accumulators summed
// across the group, then the polarisation epilogue.
template <typename T, int FT, int L>
__device__ __forceinline__ void grp_clustered(const MolDev &m, int j, T px, T py, T pz, double &ox, double &oy,
                                              double &oz)
{
    ClAcc<T> A;
    ClCount C;  // not kept: branch counts come from the 64-lane kernels only
    cl_sum<T, FT, false, true>(m, j, L, px, py, pz, A, C);
    A.sumX = gsum<L>(A.sumX); A.sumY = gsum<L>(A.sumY); A.sumZ = gsum<L>(A.sumZ);
    if constexpr (FT != 5) {
        A.fX = gsum<L>(A.fX); A.fY = gsum<L>(A.fY); A.fZ = gsum<L>(A.fZ);
        A.dXX = gsum<L>(A.dXX); A.dYY = gsum<L>(A.dYY); A.dZZ = gsum<L>(A.dZZ);
        A.sXY = gsum<L>(A.sXY); A.sXZ = gsum<L>(A.sXZ); A.sYZ = gsum<L>(A.sYZ);
    }
    if constexpr (FT == 1 || FT == 2) {
        A.aXY = gsum<L>(A.aXY); A.aXZ = gsum<L>(A.aXZ); A.aYZ = gsum<L>(A.aYZ);
    }
    T a, b, c;
    cl_epilogue<T, FT>(m, A, a, b, c);
    ox = a; oy = b; oz = c;
}

// Clustered force types with one group per wavefront (L = 64).  Lanes
// classify clusters j, j+64, ... in chunks of 64 x kChunk: a far cluster's
// multipole terms are added at once (most clusters are far, so most lanes
// are busy), while near clusters (exact charges, with or without LJ) and LJ
// clusters (band or per-atom LJ) are appended to two lists in LDS by ballot
// and then taken lane by lane.  Every lane of a list loop runs the same kind
// of per-atom loop, instead of every lane stepping through all three
// branches of whichever clusters its neighbours hold.  Terms and
// accumulators are those of cl_one (reference per-cluster arithmetic).
constexpr int kChunk = 4;
constexpr int kListInts = 2 * 64 * kChunk;  // LDS ints per wavefront

template <typename T, int FT>
__device__ __forceinline__ void wave_clustered(const MolDev &m, int ncl, int lane, T px, T py, T pz, int *lists,
                                               double &ox, double &oy, double &oz, const WaveCount &W)
{
    unsigned n_hdr = 0, n_far = 0, n_fdip = 0;  // wave-uniform
    ClAcc<T> A;
    const auto H = Layout<T>::heads(m);
    int *ex_list = lists, *lj_list = lists + 64 * kChunk;
    const unsigned long long below = (1ull << lane) - 1;
    for (int base = 0; base < ncl; base += 64 * kChunk) {
        int nex = 0, nlj = 0;
#pragma unroll
        for (int q = 0; q < kChunk; q++) {
            const int c = base + q * 64 + lane;
            int kex = 0, klj = 0;  // exact: 1 charges, 2 charges + LJ; LJ: 1 per-atom, 2 fixed parameters
            bool fr = false, dp = false;
            if (c < ncl) {
                const CAS ClusterT<T> &h = H[c];
                const T dx0 = px - h.px, dx1 = py - h.py, dx2 = pz - h.pz;
                const T r2 = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
                if constexpr (FT == 1 || FT == 2) {
                    if (r2 > h.mpr) {
                        cl_far<T, true>(m, h, dx0, dx1, dx2, r2, A);
                        fr = true;
                        dp = m.order >= 1 && !h.rep;
                        if (FT == 2) klj = 1;
                        else if (!(r2 > h.dpr)) klj = 2;
                    } else {
                        kex = (FT == 2 || !(r2 > h.dpr)) ? 2 : 1;
                    }
                } else if constexpr (FT == 3) {
                    kex = !(r2 > h.dpr) ? 2 : 1;
                } else {
                    if (!(r2 > h.dpr)) klj = 1;
                }
            }
            const unsigned long long me = __ballot(kex != 0), ml = __ballot(klj != 0);
            if (kex) ex_list[nex + __popcll(me & below)] = 4 * c + kex;
            if (klj) lj_list[nlj + __popcll(ml & below)] = 4 * c + klj;
            nex += __popcll(me);
            nlj += __popcll(ml);
            n_hdr += __popcll(__ballot(c < ncl));
            n_far += __popcll(__ballot(fr));
            n_fdip += __popcll(__ballot(dp));
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int e = lane; e < nex; e += 64) {
            const int v = ex_list[e];
            const CAS ClusterT<T> &h = H[v >> 2];
            cl_exact_rt<T, true>(m, h.start, h.end, (v & 3) == 2, px, py, pz, A);
            W.lane_add((v & 3) == 2 ? 3 : 4, (unsigned)(h.end - h.start));
        }
        for (int e = lane; e < nlj; e += 64) {
            const int v = lj_list[e];
            const CAS ClusterT<T> &h = H[v >> 2];
            cl_lj_rt<T, true>(m, h.start, h.end, (v & 3) == 2, h.qs6, h.qe24, px, py, pz, A);
            W.lane_add(5, (unsigned)(h.end - h.start));
        }
        // the next chunk overwrites the lists: every lane is past its reads
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    W.wave_add(0, n_hdr);
    W.wave_add(1, n_far);
    W.wave_add(2, n_fdip);
    A.sumX = gsum<64>(A.sumX); A.sumY = gsum<64>(A.sumY); A.sumZ = gsum<64>(A.sumZ);
    if constexpr (FT != 5) {
        A.fX = gsum<64>(A.fX); A.fY = gsum<64>(A.fY); A.fZ = gsum<64>(A.fZ);
        A.dXX = gsum<64>(A.dXX); A.dYY = gsum<64>(A.dYY); A.dZZ = gsum<64>(A.dZZ);
        A.sXY = gsum<64>(A.sXY); A.sXZ = gsum<64>(A.sXZ); A.sYZ = gsum<64>(A.sYZ);
    }
    if constexpr (FT == 1 || FT == 2) {
        A.aXY = gsum<64>(A.aXY); A.aXZ = gsum<64>(A.aXZ); A.aYZ = gsum<64>(A.aYZ);
    }
    T a, b, c;
    cl_epilogue<T, FT>(m, A, a, b, c);
    ox = a; oy = b; oz = c;
}

// Mixed-precision clustered force types 1, 2, 3, 5 with one group per
// wavefront (multipole order 0 or 1), the FP32 counterpart of
// wave_clustered.  Same branches and terms as cl_one / cl_epilogue
// (src/forceAccel.cpp:244-442, 470-630, 655-746, 819-851), regrouped:
//   * lane c classifies cluster c of each 64-cluster slice from the SoA
//     headers (one float4 load for the test, one more for a far cluster) and
//     adds a far cluster's monopole + dipole terms at once;
//   * near clusters (exact charges, with or without LJ) and LJ clusters
//     (per-atom, or the band branch's fixed parameters) go to two 128-entry
//     rings in LDS by ballot, and each ring is emptied 64 entries at a time,
//     lane by lane, so the per-atom loops run with all lanes busy;
//   * the field gradient is kept as M_ab = sum w d_a d_b plus a trace term:
//     a far cluster's diagonal r^2 (cR5 + c1 r^-7) - d_a^2 (3 cR5 + 5 c1 r^-7)
//     and off-diagonal -d_a d_b (3 cR5 + 5 c1 r^-7) (c1 = 3 p.d, cR5 = q r^-5)
//     are w = 3 cR5 + 5 c1 r^-7 and trace cR3 + c1 r^-5; an atom's are
//     w = 3 q r^-5 and trace q r^-3; the gradient is T = M - trace I;
//   * LJ as r^-8 (A r^-6 - B) with the per-atom constants of grp_force32;
//     one v_rsq_f32 per far cluster and per atom, v_rcp_f32 for LJ only.
// The epilogue (dipole_scalar x (T and the antisymmetric part) . E) runs
// in FP64 after the FP32 wavefront sums.
constexpr int kRing = 128;  // ring entries per list; two lists per wavefront

#pragma clang fp contract(fast)
template <bool LJ>
__device__ __forceinline__ void cl32_atom(f4n a, f2n ab, float px, float py, float pz, float (&s)[16])
{
    const float dx = px - a.x, dy = py - a.y, dz = pz - a.z;
    const float r2 = dx * dx + dy * dy + dz * dz;
    const float ir = __builtin_amdgcn_rsqf(r2);
    const float ir2 = ir * ir;
    if (LJ) {
        const float ir4 = ir2 * ir2;
        const float lj = (ir4 * ir4) * (ab.x * (ir4 * ir2) - ab.y);
        s[0] += lj * dx; s[1] += lj * dy; s[2] += lj * dz;
    }
    const float qr3 = a.w * ir2 * ir;
    s[3] += qr3 * dx; s[4] += qr3 * dy; s[5] += qr3 * dz;
    s[12] += qr3;
    const float w = 3.0f * (qr3 * ir2);
    const float u = w * dx, v = w * dy;
    s[6] += u * dx; s[9] += u * dy; s[10] += u * dz;
    s[7] += v * dy; s[11] += v * dz;
    s[8] += (w * dz) * dz;
}

__device__ __forceinline__ void cl32_lj(f4n a, f2n ab, float px, float py, float pz, float (&s)[16])
{
    const float dx = px - a.x, dy = py - a.y, dz = pz - a.z;
    const float r2 = dx * dx + dy * dy + dz * dz;
    const float ir2 = __builtin_amdgcn_rcpf(r2);
    const float ir6 = ir2 * ir2 * ir2;
    const float lj = (ir6 * ir2) * (ab.x * ir6 - ab.y);
    s[0] += lj * dx; s[1] += lj * dy; s[2] += lj * dz;
}

// far-field monopole + dipole of one cluster (src/forceAccel.cpp:258-316)
__device__ __forceinline__ void cl32_far(f4n g, float dx, float dy, float dz, float r2, float (&s)[16])
{
    const float ir = __builtin_amdgcn_rsqf(r2);
    const float ir2 = ir * ir;
    const float r3 = ir2 * ir, r5 = r3 * ir2, r7 = r5 * ir2;
    const float c1 = 3.0f * (g.y * dx + g.z * dy + g.w * dz);
    const float cR3 = g.x * r3, cR5 = g.x * r5;
    const float al = c1 * r5 + cR3;
    s[3] += al * dx - r3 * g.y; s[4] += al * dy - r3 * g.z; s[5] += al * dz - r3 * g.w;
    s[12] += al;
    const float w = 3.0f * cR5 + 5.0f * (c1 * r7);
    const float u = w * dx, v = w * dy;
    s[6] += u * dx; s[9] += u * dy; s[10] += u * dz;
    s[7] += v * dy; s[11] += v * dz;
    s[8] += (w * dz) * dz;
    const float s5 = 3.0f * r5;
    const float sx = s5 * dx, sy = s5 * dy, sz = s5 * dz;
    s[13] += sx * g.z - sy * g.y;
    s[14] += sx * g.w - sz * g.y;
    s[15] += sy * g.w - sz * g.z;
}
#pragma clang fp contract(off)

__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Members [k0, k1) of one cluster in cluster order, kUnroll at a time with
// their loads issued together (indices clamped to the cluster, the extra
// slots contribute nothing: charge and LJ constants zeroed).  MODE 0:
// charges; 1: charges + LJ; 2: LJ with the members' constants; 3: LJ with
// the band branch's fixed constants.
constexpr int kUnroll = 2;

template <int MODE>
__device__ __forceinline__ void cl32_members(const MolDev &m, int k0, int k1, f2n fixed, float px, float py,
                                             float pz, float (&s)[16])
{
    const GLB f4n *X = (const GLB f4n *)glb(m.cxq);
    const GLB f2n *AB = (const GLB f2n *)glb(m.cab);
    for (int k = k0; k < k1; k += kUnroll) {
        f4n a[kUnroll];
        f2n b[kUnroll];
#pragma unroll
        for (int i = 0; i < kUnroll; i++) {
            const int kk = min(k + i, k1 - 1);
            a[i] = X[kk];
            b[i] = (MODE == 1 || MODE == 2) ? AB[kk] : fixed;
        }
#pragma unroll
        for (int i = 0; i < kUnroll; i++) {
            const bool on = k + i < k1;
            if (!on) {
                a[i].w = 0.0f;
                b[i] = f2n{0.0f, 0.0f};
            }
            if constexpr (MODE == 0) cl32_atom<false>(a[i], b[i], px, py, pz, s);
            else if constexpr (MODE == 1) cl32_atom<true>(a[i], b[i], px, py, pz, s);
            else cl32_lj(a[i], b[i], px, py, pz, s);
        }
    }
}

// Empties cnt (<= 64) ring entries from head: lane k takes entry k's
// cluster and its members.  Entries are 4 c + kind.
template <bool EXACT>
__device__ __forceinline__ void cl32_flush(const MolDev &m, const int *ring, int head, int cnt, int lane, float px,
                                           float py, float pz, float (&s)[16], const WaveCount &W)
{
    if (lane >= cnt) return;
    const int v = ring[(head + lane) & (kRing - 1)];
    const int c = v >> 2, kind = v & 3;
    const i2n se = ((const GLB i2n *)glb(m.cse))[c];
    W.lane_add(!EXACT ? 5 : (kind == 2 ? 3 : 4), (unsigned)(se.y - se.x));
    if constexpr (EXACT) {
        if (kind == 2) cl32_members<1>(m, se.x, se.y, f2n{0.0f, 0.0f}, px, py, pz, s);
        else cl32_members<0>(m, se.x, se.y, f2n{0.0f, 0.0f}, px, py, pz, s);
    } else {
        if (kind == 2) {
            const f4n g2 = ((const GLB f4n *)glb(m.cg2))[c];
            cl32_members<3>(m, se.x, se.y, f2n{g2.y, g2.z}, px, py, pz, s);
        } else {
            cl32_members<2>(m, se.x, se.y, f2n{0.0f, 0.0f}, px, py, pz, s);
        }
    }
}

// Where wave_clustered32 reads the cluster headers of the slice loop: the
// molecule's SoA in global memory (through L2), or a copy of cg0, cg1 and
// the (R_d + rho)^2 of cg2 staged in LDS by the workgroup.
struct GlbHeaders {
    const GLB f4n *G0, *G1, *G2;
    const GLB unsigned long long *M;
    __device__ explicit GlbHeaders(const MolDev &m)
        : G0((const GLB f4n *)glb(m.cg0)), G1((const GLB f4n *)glb(m.cg1)), G2((const GLB f4n *)glb(m.cg2)),
          M((const GLB unsigned long long *)glb(m.cdm))
    {
    }
    __device__ f4n g0(int c) const { return G0[c]; }
    __device__ f4n g1(int c) const { return G1[c]; }
    __device__ float dr2(int c) const { return G2[c].x; }
    // dipole bits of 64-cluster slice k (cdm)
    __device__ unsigned long long dip(int k) const { return M[k]; }
    // (R_d + rho)^2 read only where the LJ cut-off decides
    static constexpr bool kPrefetchDr2 = false;
};
struct LdsHeaders {
    const f4n *G0, *G1;
    const float *D;
    const unsigned long long *M;  // the molecule's per-slice dipole bits (cdm)
    __device__ f4n g0(int c) const { return G0[c]; }
    __device__ f4n g1(int c) const { return G1[c]; }
    __device__ float dr2(int c) const { return D[c]; }
    __device__ unsigned long long dip(int k) const { return M[k]; }
    // (R_d + rho)^2 read with the next slice's headers
    static constexpr bool kPrefetchDr2 = true;
};
// LDS bytes of the staged headers: cg0, cg1 (16 B each), (R_d + rho)^2 and
// the dipole bits (8 B a 64-cluster slice, 8-byte aligned after the floats).
__host__ __device__ constexpr size_t cl_header_bytes(int ncl)
{
    return ((size_t)ncl * 36 + 7) / 8 * 8 + (size_t)((ncl + 63) / 64) * 8;
}

template <int FT, class HD>
__device__ __forceinline__ void wave_clustered32(const MolDev &m, const HD &hd, int ncl, int lane, float px, float py,
                                                 float pz, int *lists, double &ox, double &oy, double &oz,
                                                 const WaveCount &W)
{
    // one trajectory a wavefront: the cluster count (0 for an idle group) is
    // the same on every lane; read as a scalar, the slice loop and its ring
    // counters stay scalar
    ncl = __builtin_amdgcn_readfirstlane(ncl);
    // branch counts: every cluster is examined once an evaluation and nearly
    // every one is far, so the slices count the exceptions, the clusters
    // that are not far and those of them with a dipole (wave-uniform)
    unsigned n_near = 0, n_near_dip = 0;
    constexpr bool kFar = FT == 1 || FT == 2;
    // s: LJ force 0-2, field 3-5, M xx yy zz xy xz yz 6-11, trace 12, antisymmetric xy xz yz 13-15
    float s[16];
#pragma unroll
    for (int k = 0; k < 16; k++) s[k] = 0.0f;
    int *ex_ring = lists, *lj_ring = lists + kRing;
    int ex_head = 0, ex_n = 0, lj_head = 0, lj_n = 0;
    const unsigned long long below = (1ull << lane) - 1;
    const float dpr_max = m.cl_dpr_max;  // no cluster's (R_d + rho)^2 exceeds it
    // headers of the next slice are loaded while the current one is worked on
    f4n n0 = f4n{0.0f, 0.0f, 0.0f, 0.0f}, n1 = n0;
    float n2 = 0.0f;
    if (ncl > 0) {
        const int c = min(lane, ncl - 1);
        n0 = hd.g0(c);
        if (kFar) n1 = hd.g1(c);
        if constexpr (HD::kPrefetchDr2) n2 = hd.dr2(c);
    }
    for (int base = 0; base < ncl; base += 64) {
        const int c = base + lane;
        f4n h = n0, g = n1;
        float d2 = n2;
        if constexpr (!HD::kPrefetchDr2) {
            if (base + 64 < ncl) {
                const int cn = min(c + 64, ncl - 1);
                n0 = hd.g0(cn);
                if (kFar) n1 = hd.g1(cn);
            }
        }
        // ring entries 4 c + kind; exact: 1 charges, 2 charges + LJ; LJ: 1 per-atom, 2 fixed parameters
        int kex = 0, klj = 0;
        bool fr = false, near = false, ljc = false, disp = false;
        if constexpr (kFar) {
            // nearly every cluster is far: every lane adds far-field terms, a
            // near cluster's (and a lane past the last cluster's) with zero
            // moments and r^2 = 1, so that each of its terms is exactly +-0
            // and leaves the sums' bits as they were; no branch around them
            const bool valid = c < ncl;
            const float dx = px - h.x, dy = py - h.y, dz = pz - h.z;
            const float r2 = dx * dx + dy * dy + dz * dz;
            disp = HD::kPrefetchDr2 ? !(r2 > d2) : (!(r2 > dpr_max) && !(r2 > hd.dr2(min(c, ncl - 1))));
            fr = valid && r2 > h.w;
            near = valid && !fr;
            ljc = FT == 2 ? fr : (fr && disp);
            const f4n gz = fr ? g : f4n{0.0f, 0.0f, 0.0f, 0.0f};
            cl32_far(gz, dx, dy, dz, fr ? r2 : 1.0f, s);
        } else if (c < ncl) {
            const float dx = px - h.x, dy = py - h.y, dz = pz - h.z;
            const float r2 = dx * dx + dy * dy + dz * dz;
            // within (R_d + rho)^2 of this cluster (its LJ cut-off); dpr_max
            // bounds every cluster's, so it only spares the read of dr2
            disp = HD::kPrefetchDr2 ? !(r2 > d2) : (!(r2 > dpr_max) && !(r2 > hd.dr2(c)));
            if constexpr (FT == 3) {
                kex = disp ? 2 : 1;
            } else {
                if (disp) klj = 1;
            }
            near = kex != 0;
            ljc = klj != 0;
        }
        // the ballots come straight from the conditions; a ring entry's kind
        // is formed only on the (rare) lanes that append
        const unsigned long long me = __ballot(near), ml = __ballot(ljc);
        if (near) {
            if constexpr (kFar) kex = (FT == 2 || disp) ? 2 : 1;
            ex_ring[(ex_head + ex_n + __popcll(me & below)) & (kRing - 1)] = 4 * c + kex;
        }
        if (ljc) {
            if constexpr (kFar) klj = FT == 2 ? 1 : 2;
            lj_ring[(lj_head + lj_n + __popcll(ml & below)) & (kRing - 1)] = 4 * c + klj;
        }
        if constexpr (kFar) {
            // the clusters that are not far (me), and those with a dipole (the
            // dipole is zeroed for a centre-replaced cluster and at order 0)
            if (me) {
                n_near += __popcll(me);
                n_near_dip += __popcll(me & hd.dip(base >> 6));
            }
        }
        ex_n += __popcll(me);
        lj_n += __popcll(ml);
        wave_sync_lds();
        if (ex_n >= 64) {
            cl32_flush<true>(m, ex_ring, ex_head, 64, lane, px, py, pz, s, W);
            ex_head = (ex_head + 64) & (kRing - 1);
            ex_n -= 64;
        }
        if (lj_n >= 64) {
            cl32_flush<false>(m, lj_ring, lj_head, 64, lane, px, py, pz, s, W);
            lj_head = (lj_head + 64) & (kRing - 1);
            lj_n -= 64;
        }
        if constexpr (HD::kPrefetchDr2) {  // LDS: the next slice's headers, read after this one is done
            if (base + 64 < ncl) {
                const int cn = min(c + 64, ncl - 1);
                n0 = hd.g0(cn);
                if (kFar) n1 = hd.g1(cn);
                n2 = hd.dr2(cn);
            }
        }
        // the next slice's appends reuse the flushed slots: every lane is past its reads
        wave_sync_lds();
    }
    if (ex_n) cl32_flush<true>(m, ex_ring, ex_head, ex_n, lane, px, py, pz, s, W);
    if (lj_n) cl32_flush<false>(m, lj_ring, lj_head, lj_n, lane, px, py, pz, s, W);
    W.wave_add(0, ncl);
    W.wave_add(1, kFar ? ncl - n_near : 0);
    W.wave_add(2, kFar && ncl > 0 ? m.cl_ndip - n_near_dip : 0);
    wave_sync_lds();
#pragma unroll
    for (int k = 0; k < 16; k++) {
        if (FT == 5 && k >= 3) break;
        if (FT == 3 && k >= 13) break;
        s[k] = gsum<64>(s[k]);
    }
    double fx = s[0], fy = s[1], fz = s[2];
    if constexpr (FT != 5) {
        const double ds = m.dipole_scalar;
        const double Ex = s[3], Ey = s[4], Ez = s[5], t = s[12];
        const double Txx = (double)s[6] - t, Tyy = (double)s[7] - t, Tzz = (double)s[8] - t;
        const double Txy = s[9], Txz = s[10], Tyz = s[11];
        double Axy = 0, Axz = 0, Ayz = 0;
        if constexpr (kFar) { Axy = s[13]; Axz = s[14]; Ayz = s[15]; }
        fx -= ds * (Txx * Ex + (Txy - Axy) * Ey + (Txz - Axz) * Ez);
        fy -= ds * ((Txy + Axy) * Ex + Tyy * Ey + (Tyz - Ayz) * Ez);
        fz -= ds * ((Txz + Axz) * Ex + (Tyz + Ayz) * Ey + Tzz * Ez);
    }
    ox = fx; oy = fy; oz = fz;
}

// FP64 counterparts of cl32_atom / cl32_lj / cl32_far and the slice loop of
// wave_clustered32, for the FP64 group kernels of the clustered types with
// one group per wavefront (multipole order 0 or 1): the same branches,
// rings and regrouped far field (M_ab = sum w d_a d_b plus a trace term, the
// antisymmetric part as three components), in FP64 with 1/r from rsqrt64 and
// 1/r^2 from rcp64; headers from the FP64 SoA dg0, dg1, dg2, members from
// the FP64 member records.  LJ in the reference's form t2 = sigma^6 r^-6,
// (24 eps/mu)(2 t2^2 - t2) r^-2 (src/forceAccel.cpp:341-360, 366-403).
typedef double d4n __attribute__((ext_vector_type(4)));

template <bool LJ>
__device__ __forceinline__ void cl64_atom(double ax, double ay, double az, double aq, double s6, double e24,
                                          double px, double py, double pz, double (&s)[16])
{
    const double dx = px - ax, dy = py - ay, dz = pz - az;
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double ir = rsqrt64(r2);
    const double ir2 = ir * ir;
    if (LJ) {
        const double t2 = s6 * (ir2 * ir2 * ir2);
        const double t1 = 2.0 * t2 * t2;
        const double sc = e24 * (t1 - t2) * ir2;
        s[0] += sc * dx; s[1] += sc * dy; s[2] += sc * dz;
    }
    const double qr3 = aq * ir2 * ir;
    s[3] += qr3 * dx; s[4] += qr3 * dy; s[5] += qr3 * dz;
    s[12] += qr3;
    const double w = 3.0 * (qr3 * ir2);
    const double u = w * dx, v = w * dy;
    s[6] += u * dx; s[9] += u * dy; s[10] += u * dz;
    s[7] += v * dy; s[11] += v * dz;
    s[8] += (w * dz) * dz;
}

__device__ __forceinline__ void cl64_lj(double ax, double ay, double az, double s6, double e24, double px, double py,
                                        double pz, double (&s)[16])
{
    const double dx = px - ax, dy = py - ay, dz = pz - az;
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double ir2 = rcp64(r2);
    const double t2 = s6 * (ir2 * ir2 * ir2);
    const double t1 = 2.0 * t2 * t2;
    const double sc = e24 * (t1 - t2) * ir2;
    s[0] += sc * dx; s[1] += sc * dy; s[2] += sc * dz;
}

// far-field monopole + dipole of one cluster (src/forceAccel.cpp:258-316)
__device__ __forceinline__ void cl64_far(d4n g, double dx, double dy, double dz, double r2, double (&s)[16])
{
    const double ir = rsqrt64(r2);
    const double ir2 = ir * ir;
    const double r3 = ir2 * ir, r5 = r3 * ir2, r7 = r5 * ir2;
    const double c1 = 3.0 * (g.y * dx + g.z * dy + g.w * dz);
    const double cR3 = g.x * r3, cR5 = g.x * r5;
    const double al = c1 * r5 + cR3;
    s[3] += al * dx - r3 * g.y; s[4] += al * dy - r3 * g.z; s[5] += al * dz - r3 * g.w;
    s[12] += al;
    const double w = 3.0 * cR5 + 5.0 * (c1 * r7);
    const double u = w * dx, v = w * dy;
    s[6] += u * dx; s[9] += u * dy; s[10] += u * dz;
    s[7] += v * dy; s[11] += v * dz;
    s[8] += (w * dz) * dz;
    const double s5 = 3.0 * r5;
    const double sx = s5 * dx, sy = s5 * dy, sz = s5 * dz;
    s[13] += sx * g.z - sy * g.y;
    s[14] += sx * g.w - sz * g.y;
    s[15] += sy * g.w - sz * g.z;
}

// Members [k0, k1) of one cluster in cluster order, two at a time (MODE as
// in cl32_members); the extra slot re-reads the last member and adds zero.
template <int MODE>
__device__ __forceinline__ void cl64_members(const MolDev &m, int k0, int k1, double fs6, double fe24, double px,
                                             double py, double pz, double (&s)[16])
{
    // a member record (AtomF64, 64 B) as two vectors: x y z q | s6 e24 e4 pad
    const GLB d4n *C = (const GLB d4n *)glb(m.c64);
    for (int k = k0; k < k1; k += 2) {
        d4n a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int kk = min(k + i, k1 - 1);
            a[i] = C[2 * kk];
            if constexpr (MODE != 0 && MODE != 3) b[i] = C[2 * kk + 1];
        }
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const bool on = k + i < k1;
            const double q = on ? a[i].w : 0.0;
            const double s6 = on ? (MODE == 3 ? fs6 : (MODE == 0 ? 0.0 : b[i].x)) : 0.0;
            const double e24 = on ? (MODE == 3 ? fe24 : (MODE == 0 ? 0.0 : b[i].y)) : 0.0;
            if constexpr (MODE == 0) cl64_atom<false>(a[i].x, a[i].y, a[i].z, q, s6, e24, px, py, pz, s);
            else if constexpr (MODE == 1) cl64_atom<true>(a[i].x, a[i].y, a[i].z, q, s6, e24, px, py, pz, s);
            else cl64_lj(a[i].x, a[i].y, a[i].z, s6, e24, px, py, pz, s);
        }
    }
}

template <bool EXACT>
__device__ __forceinline__ void cl64_flush(const MolDev &m, const int *ring, int head, int cnt, int lane, double px,
                                           double py, double pz, double (&s)[16], const WaveCount &W)
{
    if (lane >= cnt) return;
    const int v = ring[(head + lane) & (kRing - 1)];
    const int c = v >> 2, kind = v & 3;
    const i2n se = ((const GLB i2n *)glb(m.cse))[c];
    W.lane_add(!EXACT ? 5 : (kind == 2 ? 3 : 4), (unsigned)(se.y - se.x));
    if constexpr (EXACT) {
        if (kind == 2) cl64_members<1>(m, se.x, se.y, 0.0, 0.0, px, py, pz, s);
        else cl64_members<0>(m, se.x, se.y, 0.0, 0.0, px, py, pz, s);
    } else {
        if (kind == 2) {
            const d4n g2 = ((const GLB d4n *)glb(m.dg2))[c];
            cl64_members<3>(m, se.x, se.y, g2.y, g2.z, px, py, pz, s);
        } else {
            cl64_members<2>(m, se.x, se.y, 0.0, 0.0, px, py, pz, s);
        }
    }
}

template <int FT>
__device__ __forceinline__ void wave_clustered64(const MolDev &m, int ncl, int lane, double px, double py, double pz,
                                                 int *lists, double &ox, double &oy, double &oz, const WaveCount &W)
{
    ncl = __builtin_amdgcn_readfirstlane(ncl);  // wave-uniform, as in wave_clustered32
    unsigned n_near = 0, n_near_dip = 0;  // wave-uniform; the exceptions, as in wave_clustered32
    constexpr bool kFar = FT == 1 || FT == 2;
    // s: LJ force 0-2, field 3-5, M xx yy zz xy xz yz 6-11, trace 12, antisymmetric xy xz yz 13-15
    double s[16];
#pragma unroll
    for (int k = 0; k < 16; k++) s[k] = 0.0;
    int *ex_ring = lists, *lj_ring = lists + kRing;
    int ex_head = 0, ex_n = 0, lj_head = 0, lj_n = 0;
    const unsigned long long below = (1ull << lane) - 1;
    const GLB d4n *G0 = (const GLB d4n *)glb(m.dg0);
    const GLB d4n *G1 = (const GLB d4n *)glb(m.dg1);
    const GLB d4n *G2 = (const GLB d4n *)glb(m.dg2);
    const double dpr_max = m.cl_dpr_max64;  // no cluster's (R_d + rho)^2 exceeds it
    // headers of the next slice are loaded while the current one is worked on
    d4n n0 = d4n{0.0, 0.0, 0.0, 0.0}, n1 = n0;
    if (ncl > 0) {
        const int c = min(lane, ncl - 1);
        n0 = G0[c];
        if (kFar) n1 = G1[c];
    }
    for (int base = 0; base < ncl; base += 64) {
        const int c = base + lane;
        const d4n h = n0, g = n1;
        if (base + 64 < ncl) {
            const int cn = min(c + 64, ncl - 1);
            n0 = G0[cn];
            if (kFar) n1 = G1[cn];
        }
        int kex = 0, klj = 0;  // exact: 1 charges, 2 charges + LJ; LJ: 1 per-atom, 2 fixed parameters
        bool fr = false;
        if constexpr (kFar) {
            // as in wave_clustered32: every lane adds far-field terms, a near
            // cluster's (and a lane past the last cluster's) with zero moments
            // and r^2 = 1, each term exactly +-0; no branch around them
            const bool valid = c < ncl;
            const double dx = px - h.x, dy = py - h.y, dz = pz - h.z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            const bool disp = !(r2 > dpr_max) && !(r2 > G2[min(c, ncl - 1)].x);
            fr = valid && r2 > h.w;
            const d4n gz = fr ? g : d4n{0.0, 0.0, 0.0, 0.0};
            cl64_far(gz, dx, dy, dz, fr ? r2 : 1.0, s);
            if (fr) klj = FT == 2 ? 1 : (disp ? 2 : 0);
            else if (valid) kex = (FT == 2 || disp) ? 2 : 1;
        } else if (c < ncl) {
            const double dx = px - h.x, dy = py - h.y, dz = pz - h.z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            // within (R_d + rho)^2 of this cluster (its LJ cut-off)
            const bool disp = !(r2 > dpr_max) && !(r2 > G2[c].x);
            if constexpr (FT == 3) {
                kex = disp ? 2 : 1;
            } else {
                if (disp) klj = 1;
            }
        }
        const unsigned long long me = __ballot(kex != 0), ml = __ballot(klj != 0);
        if (kex) ex_ring[(ex_head + ex_n + __popcll(me & below)) & (kRing - 1)] = 4 * c + kex;
        if (klj) lj_ring[(lj_head + lj_n + __popcll(ml & below)) & (kRing - 1)] = 4 * c + klj;
        if constexpr (kFar) {
            const bool near = c < ncl && !fr;
            const unsigned long long nf = __ballot(near);
            if (nf) {
                n_near += __popcll(nf);
                n_near_dip += __popcll(__ballot(near && (g.y != 0.0 || g.z != 0.0 || g.w != 0.0)));
            }
        }
        ex_n += __popcll(me);
        lj_n += __popcll(ml);
        wave_sync_lds();
        if (ex_n >= 64) {
            cl64_flush<true>(m, ex_ring, ex_head, 64, lane, px, py, pz, s, W);
            ex_head = (ex_head + 64) & (kRing - 1);
            ex_n -= 64;
        }
        if (lj_n >= 64) {
            cl64_flush<false>(m, lj_ring, lj_head, 64, lane, px, py, pz, s, W);
            lj_head = (lj_head + 64) & (kRing - 1);
            lj_n -= 64;
        }
        // the next slice's appends reuse the flushed slots: every lane is past its reads
        wave_sync_lds();
    }
    if (ex_n) cl64_flush<true>(m, ex_ring, ex_head, ex_n, lane, px, py, pz, s, W);
    if (lj_n) cl64_flush<false>(m, lj_ring, lj_head, lj_n, lane, px, py, pz, s, W);
    W.wave_add(0, ncl);
    W.wave_add(1, kFar ? ncl - n_near : 0);
    W.wave_add(2, kFar && ncl > 0 ? m.cl_ndip64 - n_near_dip : 0);
    wave_sync_lds();
#pragma unroll
    for (int k = 0; k < 16; k++) {
        if (FT == 5 && k >= 3) break;
        if (FT == 3 && k >= 13) break;
        s[k] = gsum<64>(s[k]);
    }
    double fx = s[0], fy = s[1], fz = s[2];
    if constexpr (FT != 5) {
        const double ds = m.dipole_scalar;
        const double Ex = s[3], Ey = s[4], Ez = s[5], t = s[12];
        const double Txx = s[6] - t, Tyy = s[7] - t, Tzz = s[8] - t;
        const double Txy = s[9], Txz = s[10], Tyz = s[11];
        double Axy = 0, Axz = 0, Ayz = 0;
        if constexpr (kFar) { Axy = s[13]; Axz = s[14]; Ayz = s[15]; }
        fx -= ds * (Txx * Ex + (Txy - Axy) * Ey + (Txz - Axz) * Ez);
        fy -= ds * ((Txy + Axy) * Ex + Tyy * Ey + (Tyz - Ayz) * Ez);
        fz -= ds * ((Txz + Axz) * Ex + (Tyz + Ayz) * Ey + Tzz * Ez);
    }
    ox = fx; oy = fy; oz = fz;
}

// getPotential across the group (src/forceAccel.cpp:943-982), FP64, over
// the SoA X (x y z q s6 e24 e4, stride npad).
template <int L, typename PD>
__device__ __forceinline__ double grp_potential(PD X, int npad, bool charged, double dipole_scalar, int j,
                                                double px, double py, double pz)
{
    const PD Y = X + npad, Z = X + 2 * npad, Q = X + 3 * npad, S6 = X + 4 * npad, E4 = X + 6 * npad;
    double total = 0, sX = 0, sY = 0, sZ = 0;
    for (int k = j; k < npad; k += L) {
        double xD = px - X[k], yD = py - Y[k], zD = pz - Z[k];
        double r2 = xD * xD + yD * yD + zD * zD;
        double r6 = r2 * r2 * r2;
        double t2 = S6[k] / r6;
        double t1 = t2 * t2;
        total += E4[k] * (t1 - t2);
        if (charged) {
            double r_3 = Q[k] / (r2 * __builtin_sqrt(r2));
            sX += xD * r_3; sY += yD * r_3; sZ += zD * r_3;
        }
    }
    total = gsum<L>(total);
    if (charged) {
        sX = gsum<L>(sX); sY = gsum<L>(sY); sZ = gsum<L>(sZ);
        total -= dipole_scalar * 0.5 * (sX * sX + sY * sY + sZ * sZ);
    }
    return total;
}

// Molecule source of k_traj_grp: one molecule, atoms in LDS, constants in
// kernel arguments (wave-uniform).
template <int PREC, int FT, int L>
struct LdsSource {
    static constexpr bool kCoop = false;
    static constexpr bool kMulti = false;
    const GroupArgs &A;
    const char *smem;
    __device__ void bind(int) {}
    __device__ void flush(const TrajIO &) const {}
    __device__ int pairs() const { return (FT == 4 || FT == 6) ? A.m.n : 0; }
    __device__ int atoms() const { return A.m.n; }
    __device__ double mu() const { return A.m.mu; }
    __device__ double dt0() const { return A.dt0; }
    __device__ bool exits(double px, double py, double pz) const
    {
        return PREC == 1 ? outside(A.m, px, py, pz) : outside_div(A.m, px, py, pz);
    }
    // idle groups evaluate at a far point: the trip count stays wave-uniform
    __device__ void force(bool, int j, double ex, double ey, double ez, double &fx, double &fy, double &fz) const
    {
        const int npad = A.npad;
        if constexpr (PREC == 1) {
            const f4n *xq = (const f4n *)smem;
            grp_force32<FT, L>(xq, (const f2n *)(xq + npad), npad, j, (float)ex, (float)ey, (float)ez,
                               A.m.dipole_scalar, fx, fy, fz);
        } else {
            grp_force64<FT, L>((const double *)smem, npad, j, ex, ey, ez, A.m.dipole_scalar, fx, fy, fz);
        }
    }
    __device__ double potential(bool, int j, double ex, double ey, double ez) const
    {
        return grp_potential<L>(glb(A.soa), A.npad, A.m.charged, A.m.dipole_scalar, j, ex, ey, ez);
    }
};

// Molecule source of k_traj_multi: the molecule of the group's trajectory,
// read from the MolRec table.  Idle groups skip the sums.
template <int PREC, int FT, int L>
struct TableSource {
    static constexpr bool kCoop = false;
    static constexpr bool kMulti = false;
    const MultiArgs &A;
    int *lists;    // this wavefront's cluster lists in LDS (clustered types, L = 64)
    WaveCount W;   // this wavefront's branch counts in LDS (clustered types, L = 64)
    int mi = 0;
    __device__ void bind(int t) { mi = A.tmol ? A.tmol[t] : 0; }
    __device__ void flush(const TrajIO &io) const
    {
        if constexpr (FT != 4 && FT != 6 && L == 64) flush_counts(io.evals, W);
    }
    __device__ const MolRec &rec() const { return A.mols[mi]; }
    __device__ int pairs() const { return (FT == 4 || FT == 6) ? rec().m.n : 0; }
    __device__ int atoms() const { return rec().m.n; }
    __device__ double mu() const { return rec().m.mu; }
    __device__ double dt0() const { return rec().dt0; }
    __device__ bool exits(double px, double py, double pz) const
    {
        return PREC == 1 ? outside(rec().m, px, py, pz) : outside_div(rec().m, px, py, pz);
    }
    __device__ void force(bool active, int j, double ex, double ey, double ez, double &fx, double &fy,
                          double &fz) const
    {
        const MolRec &r = rec();
        if constexpr (FT == 4 || FT == 6) {
            const int npad = active ? r.npad : 0;
            if constexpr (PREC == 1)
                grp_force32_glb<FT, L>(glb((const f4n *)r.xq), glb((const f2n *)r.se), npad, j, (float)ex, (float)ey,
                                       (float)ez, r.m.dipole_scalar, fx, fy, fz);
            else
                grp_force64<FT, L>(glb(r.soa), npad, j, ex, ey, ez, r.m.dipole_scalar, fx, fy, fz);
        } else if constexpr (L == 64) {
            const int ncl = active ? r.m.ncl : 0;
            if constexpr (PREC == 1) {
                if (r.m.order <= 1)
                    wave_clustered32<FT>(r.m, GlbHeaders(r.m), ncl, j, (float)ex, (float)ey, (float)ez, lists, fx,
                                         fy, fz, W);
                else
                    wave_clustered<float, FT>(r.m, ncl, j, (float)ex, (float)ey, (float)ez, lists, fx, fy, fz, W);
            } else if (r.m.order <= 1) {
                wave_clustered64<FT>(r.m, ncl, j, ex, ey, ez, lists, fx, fy, fz, W);
            } else {
                wave_clustered<double, FT>(r.m, ncl, j, ex, ey, ez, lists, fx, fy, fz, W);
            }
        } else {
            MolDev m = r.m;
            if (!active) m.ncl = 0;
            if constexpr (PREC == 1)
                grp_clustered<float, FT, L>(m, j, (float)ex, (float)ey, (float)ez, fx, fy, fz);
            else
                grp_clustered<double, FT, L>(m, j, ex, ey, ez, fx, fy, fz);
        }
    }
    __device__ double potential(bool active, int j, double ex, double ey, double ez) const
    {
        const MolRec &r = rec();
        return grp_potential<L>(glb(r.soa), active ? r.npad : 0, r.m.charged, r.m.dipole_scalar, j, ex, ey, ez);
    }
};

// Molecule source of k_traj_cl_lds: one molecule of a clustered force type
// (multipole order 0 or 1), mixed precision, one group of 64 lanes a
// wavefront; the cluster headers of the slice loop come from the
// workgroup's LDS copy (LdsHeaders), the members of near clusters and the
// potential's atoms from global memory.  Same arithmetic in the same order
// as TableSource, so the two give the same bits.
template <int FT>
struct LdsClusterSource {
    static constexpr bool kCoop = false;
    static constexpr bool kMulti = false;
    const GroupArgs &A;
    int *lists;    // this wavefront's cluster rings
    WaveCount W;   // this wavefront's branch counts
    LdsHeaders H;
    __device__ void bind(int) {}
    __device__ void flush(const TrajIO &io) const { flush_counts(io.evals, W); }
    __device__ int pairs() const { return 0; }
    __device__ int atoms() const { return A.m.n; }
    __device__ double mu() const { return A.m.mu; }
    __device__ double dt0() const { return A.dt0; }
    __device__ bool exits(double px, double py, double pz) const { return outside(A.m, px, py, pz); }
    __device__ void force(bool active, int j, double ex, double ey, double ez, double &fx, double &fy,
                          double &fz) const
    {
        wave_clustered32<FT>(A.m, H, active ? A.m.ncl : 0, j, (float)ex, (float)ey, (float)ez, lists, fx, fy, fz, W);
    }
    __device__ double potential(bool active, int j, double ex, double ey, double ez) const
    {
        return grp_potential<64>(glb(A.soa), active ? A.npad : 0, A.m.charged, A.m.dipole_scalar, j, ex, ey, ez);
    }
};

// Molecule source of k_traj_tile: one molecule too large for LDS, streamed
// through LDS in tiles of A.tile atoms by the whole workgroup.  force() is
// collective: every thread of the workgroup calls it (group_loop's kCoop
// mode decides for the workgroup), stages each tile between two barriers,
// and the wavefronts with a group that needs the force add the tile's
// pairs.  Lane j of a group still sums atoms j, j+L, ... in order (tiles
// start at multiples of 64), with the per-pair arithmetic and group sums of
// the other kernels: the same bits as k_traj_grp and k_traj_multi.
template <int PREC, int FT, int L>
struct TileSource {
    static constexpr bool kCoop = true;
    static constexpr bool kMulti = false;
    const GroupArgs &A;
    char *smem;
    __device__ void bind(int) {}
    __device__ void flush(const TrajIO &) const {}
    __device__ int pairs() const { return A.m.n; }
    __device__ int atoms() const { return A.m.n; }
    __device__ double mu() const { return A.m.mu; }
    __device__ double dt0() const { return A.dt0; }
    __device__ bool exits(double px, double py, double pz) const
    {
        return PREC == 1 ? outside(A.m, px, py, pz) : outside_div(A.m, px, py, pz);
    }
    __device__ void force(bool active, int j, double ex, double ey, double ez, double &fx, double &fy,
                          double &fz) const
    {
        const bool wave = __any(active);  // this wavefront adds pairs (idle groups at a far point)
        const int npad = A.npad, T = A.tile;
        if constexpr (PREC == 1) {
            f4n *xq = (f4n *)smem;
            f2n *ab = (f2n *)(xq + T);
            const GLB f4n *gx = (const GLB f4n *)glb(A.xq);
            const GLB f2n *ga = (const GLB f2n *)glb(A.se);
            const float px = (float)ex, py = (float)ey, pz = (float)ez;
            float s[12];
#pragma unroll
            for (int k = 0; k < 12; k++) s[k] = 0.0f;
            for (int base = 0; base < npad; base += T) {
                const int cnt = min(T, npad - base);
                __syncthreads();  // every reader of the previous tile is done
                for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
                    xq[i] = gx[base + i];
                    ab[i] = ga[base + i];
                }
                __syncthreads();
                if (wave) lds_pairs32<FT, L>(xq, ab, cnt, j, px, py, pz, s);
            }
            if (wave) force32_epilogue<FT, L>(s, A.m.dipole_scalar, fx, fy, fz);
        } else {
            double *X = (double *)smem;  // x y z q s6 e24, T each
            double s[12];
#pragma unroll
            for (int k = 0; k < 12; k++) s[k] = 0.0;
            for (int base = 0; base < npad; base += T) {
                const int cnt = min(T, npad - base);
                __syncthreads();
                for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
#pragma unroll
                    for (int f = 0; f < 6; f++) X[f * T + i] = A.soa[f * npad + base + i];
                }
                __syncthreads();
                if (wave) {
                    // cnt and T are multiples of 64: every lane makes cnt / L
                    // passes; at 32 and 64 lanes two pairs a pass (their atoms
                    // read two at a time), with fewer lanes (more state live)
                    // one, which needs no registers beyond 128
                    const int passes = cnt / L;
                    if constexpr (L >= 32) {
#pragma unroll 2
                        for (int i = 0; i < passes; i++) {
                            const int k = j + i * L;
                            pair64<FT>(X[k], X[T + k], X[2 * T + k], X[3 * T + k], X[4 * T + k], X[5 * T + k], ex,
                                       ey, ez, s);
                        }
                    } else {
                        for (int i = 0; i < passes; i++) {
                            const int k = j + i * L;
                            pair64<FT>(X[k], X[T + k], X[2 * T + k], X[3 * T + k], X[4 * T + k], X[5 * T + k], ex,
                                       ey, ez, s);
                        }
                    }
                }
            }
            if (wave) force64_epilogue<FT, L>(s, A.m.dipole_scalar, fx, fy, fz);
        }
    }
    __device__ double potential(bool, int j, double ex, double ey, double ez) const
    {
        return grp_potential<L>(glb(A.soa), A.npad, A.m.charged, A.m.dipole_scalar, j, ex, ey, ez);
    }
};

// Molecule source of k_traj_lds_multi: the jobs of one launch shape, each
// workgroup working on one job at a time with that job's molecule staged in
// its LDS (the layout of LdsSource).  A job is a position of the batch's work
// list with a range of trajectories of its own, ordered longest expected
// first; the workgroup's groups refill from the range of the staged job
// only.  When that range is exhausted and no wavefront of the workgroup still
// holds one of its trajectories, one wavefront (the stager) picks the job
// with the most trajectories left, stages its molecule and publishes it; the
// others wait (s_sleep) and then refill from the new range.  Lane j of a
// group sums atoms j, j+L, ... with the force call of LdsSource, so every
// trajectory gives the bits it gives in a launch of its own.
//
// Why no wavefront can read the LDS while it is restaged: a wavefront
// evaluates forces only while one of its groups holds a trajectory, a group
// takes a trajectory only from the range of the job its wavefront last saw
// published, and the stager stages only after its own refill found that range
// exhausted (ranges only shrink) and with `active`, the wavefronts counted in
// before their last refill and not yet counted out, at zero.
struct WgCtl {
    int pos;     // position whose molecule is in LDS; -1: nothing left
    int gen;     // publications so far
    int active;  // wavefronts that may hold a trajectory of the staged job
    int stager;  // 1 while a wavefront restages
};
constexpr int kCtlBytes = 64;  // LDS before the molecule

struct LdsMultiArgs {
    TrajIO io;               // io.items: the whole work list; counter / n_items per job (pos_ctr, ends)
    const MolRec *mols;      // molecule table, indexed by job
    const int32_t *pos_job;  // job of each work-list position
    const int32_t *ends;     // end of each position's range of io.items (written on the device)
    int32_t *pos_ctr;        // per-position refill counters, zero at launch
    const int32_t *wg_pos;   // first position of each workgroup
    int pos0, npos;          // this launch's positions
};

template <int PREC, int FT, int L>
struct LdsMultiSource {
    static constexpr bool kCoop = false;
    static constexpr bool kMulti = true;
    const LdsMultiArgs &A;
    char *mol;
    WgCtl *ctl;
    const MolRec *rec = nullptr;
    int gen = 0;
    bool counted = false;
    __device__ void bind(int) {}
    __device__ void flush(const TrajIO &) const {}
    __device__ int pairs() const { return rec->m.n; }
    __device__ int atoms() const { return rec->m.n; }
    __device__ double mu() const { return rec->m.mu; }
    __device__ double dt0() const { return rec->dt0; }
    __device__ bool exits(double px, double py, double pz) const
    {
        return PREC == 1 ? outside(rec->m, px, py, pz) : outside_div(rec->m, px, py, pz);
    }
    __device__ void force(bool, int j, double ex, double ey, double ez, double &fx, double &fy, double &fz) const
    {
        const int npad = rec->npad;
        if constexpr (PREC == 1) {
            const f4n *xq = (const f4n *)mol;
            grp_force32<FT, L>(xq, (const f2n *)(xq + npad), npad, j, (float)ex, (float)ey, (float)ez,
                               rec->m.dipole_scalar, fx, fy, fz);
        } else {
            grp_force64<FT, L>((const double *)mol, npad, j, ex, ey, ez, rec->m.dipole_scalar, fx, fy, fz);
        }
    }
    __device__ double potential(bool, int j, double ex, double ey, double ez) const
    {
        return grp_potential<L>(glb(rec->soa), rec->npad, rec->m.charged, rec->m.dipole_scalar, j, ex, ey, ez);
    }
    // refill from position p's range (p wave-uniform)
    __device__ void use(int p, TrajIO &io)
    {
        p = __builtin_amdgcn_readfirstlane(p);
        rec = A.mols + A.pos_job[p];
        const int b = p ? A.ends[p - 1] : 0;
        io.items = A.io.items + b;
        io.n_items = A.ends[p] - b;
        io.counter = A.pos_ctr + p;
    }
    __device__ void count_in()
    {
        if (!counted) {
            if ((threadIdx.x & 63) == 0) atomicAdd(&ctl->active, 1);
            counted = true;
        }
    }
    __device__ void count_out()
    {
        if (counted) {
            if ((threadIdx.x & 63) == 0) atomicSub(&ctl->active, 1);
            counted = false;
        }
    }
    // the molecule of job rec into LDS by this wavefront alone
    __device__ void stage_wave(const MolRec *r) const
    {
        const int lane = threadIdx.x & 63, npad = r->npad;
        if constexpr (PREC == 1) {
            const GLB f4n *gx = (const GLB f4n *)glb(r->xq);
            const GLB f2n *ga = (const GLB f2n *)glb(r->se);
            f4n *lx = (f4n *)mol;
            f2n *la = (f2n *)(lx + npad);
            for (int i = lane; i < npad; i += 4 * 64) {  // npad is a multiple of 64
                f4n x[4];
                f2n a[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int k = min(i + 64 * u, npad - 64 + lane);
                    x[u] = gx[k];
                    a[u] = ga[k];
                }
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (i + 64 * u < npad) {
                        lx[i + 64 * u] = x[u];
                        la[i + 64 * u] = a[u];
                    }
            }
        } else {
            const GLB double *gs = glb(r->soa);
            double *ls = (double *)mol;
            const int tot = 6 * npad;
            for (int i = lane; i < tot; i += 4 * 64) {
                double x[4];
#pragma unroll
                for (int u = 0; u < 4; u++) x[u] = gs[min(i + 64 * u, tot - 64 + lane)];
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (i + 64 * u < tot) ls[i + 64 * u] = x[u];
            }
        }
    }
    // position with the most trajectories left (-1: none), wave-uniform
    __device__ int pick() const
    {
        const int lane = threadIdx.x & 63;
        int best = -1, left = 0;
        for (int q = lane; q < A.npos; q += 64) {
            const int p = A.pos0 + q;
            const int b = p ? A.ends[p - 1] : 0;
            const int l = A.ends[p] - b - __hip_atomic_load(A.pos_ctr + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (l > left) {
                left = l;
                best = p;
            }
        }
        for (int off = 32; off > 0; off >>= 1) {  // largest left, then lowest position
            const int l2 = __shfl_xor(left, off), b2 = __shfl_xor(best, off);
            if (l2 > left || (l2 == left && l2 > 0 && b2 < best)) {
                left = l2;
                best = b2;
            }
        }
        return __builtin_amdgcn_readfirstlane(left > 0 ? best : -1);
    }
    // This wavefront's groups are idle and its job's range is exhausted:
    // wait for the next publication (restaging it when the workgroup holds no
    // trajectory), then refill from the new range.  false: nothing is left.
    __device__ bool next_job(TrajIO &io)
    {
        const int lane = threadIdx.x & 63;
        for (;;) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            const int g = __builtin_amdgcn_readfirstlane(__hip_atomic_load(&ctl->gen, __ATOMIC_ACQUIRE,
                                                                          __HIP_MEMORY_SCOPE_WORKGROUP));
            if (g != gen) {
                gen = g;
                const int p = __builtin_amdgcn_readfirstlane(
                    __hip_atomic_load(&ctl->pos, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
                if (p < 0) return false;
                use(p, io);
                return true;
            }
            const int act = __builtin_amdgcn_readfirstlane(
                __hip_atomic_load(&ctl->active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
            if (act == 0) {
                int won = 0;
                if (lane == 0) won = atomicCAS(&ctl->stager, 0, 1) == 0;
                won = __builtin_amdgcn_readfirstlane(won);
                if (won) {
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                    const int g2 = __builtin_amdgcn_readfirstlane(
                        __hip_atomic_load(&ctl->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
                    if (g2 == gen) {  // nobody published since this wavefront last looked
                        const int p = pick();
                        if (p >= 0) stage_wave(A.mols + A.pos_job[p]);
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                        if (lane == 0) {
                            __hip_atomic_store(&ctl->pos, p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                            __hip_atomic_fetch_add(&ctl->gen, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                        }
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                    if (lane == 0) __hip_atomic_store(&ctl->stager, 0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                    continue;
                }
            }
            __builtin_amdgcn_s_sleep(8);
        }
    }
};

enum : int { PH_IDLE = 0, PH_INIT = 1, PH_STEP = 2, PH_CHECK = 3 };

// Integration steps past which a trajectory counts as long (the 99th
// percentile of the steps of a 2,000-atom ion is ~1,100).
constexpr int kLongSteps = 1024;

// The trajectory state machine of diffEqSolver::verletVel
// (src/diffEqSolver.cpp:81-211) for the group of lanes this lane belongs
// to, refilled from the work list until it is exhausted.
template <int L, class Src>
__device__ __forceinline__ void group_loop(TrajIO A, Src &S)
{
    const int lane = threadIdx.x & 63;
    const int g0 = lane & ~(L - 1);   // first lane of my group
    const int j = lane & (L - 1);     // my index in the group
    int t = 0, phase = PH_IDLE, tries = 0, maxtries = 8, orbits = 0, steps = 0, tot = 0;
    bool needE0 = false;
    double px = 0, py = 0, pz = 0, vx = 0, vy = 0, vz = 0, ax = 0, ay = 0, az = 0, dt = 0, rdot = 0;
    double e0 = 0, chi = 0;
    bool exhausted = false, prio = false;
    unsigned n_steps = 0, n_evals = 0;
    unsigned long long n_pairs = 0, n_pots = 0;

    for (;;) {
        if (!exhausted) {
            const bool need = (phase == PH_IDLE);
            const unsigned long long mask = __ballot(need && j == 0);
            if (mask) {
                if constexpr (Src::kMulti) S.count_in();  // before taking trajectories of the staged job
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
                        S.bind(t);
                        tries = 0; maxtries = 8;
                        dt = S.dt0();
                        px = A.p0[3 * t]; py = A.p0[3 * t + 1]; pz = A.p0[3 * t + 2];
                        vx = A.v0[3 * t]; vy = A.v0[3 * t + 1]; vz = A.v0[3 * t + 2];
                        tot = 0;
                        needE0 = true;
                        phase = PH_INIT;
                    }
                }
            }
        }
        // a collective source keeps every wavefront of the workgroup in the
        // loop (and in its barriers) until the workgroup has no trajectory left
        if constexpr (Src::kCoop) {
            if (!__syncthreads_or(phase != PH_IDLE)) break;
        } else {
            if (!__any(phase != PH_IDLE)) {
                if constexpr (Src::kMulti) {
                    S.count_out();
                    if (S.next_job(A)) {
                        exhausted = false;
                        continue;
                    }
                }
                break;
            }
            // a wavefront carrying a long trajectory issues first: a launch
            // ends with its longest trajectories, and the SIMD's other
            // wavefronts fill the cycles around them (issue order only)
            const bool hi = __any(phase != PH_IDLE && tot >= kLongSteps);
            if (hi != prio) {
                if (hi) __builtin_amdgcn_s_setprio(2);
                else __builtin_amdgcn_s_setprio(0);
                prio = hi;
            }
        }

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
        if constexpr (Src::kCoop) {
            if (__syncthreads_or(needF)) S.force(needF, j, ex, ey, ez, fx, fy, fz);
        } else {
            if (__any(needF)) S.force(needF, j, ex, ey, ez, fx, fy, fz);
        }
        if (__any(needU)) u = S.potential(needU, j, ex, ey, ez);
        if (j == 0) {
            n_evals += needF;
            n_steps += (phase == PH_STEP);
            if (needU) n_pots += (unsigned)S.atoms();
            if (needF) n_pairs += (unsigned)S.pairs();
        }

        if (phase == PH_INIT) {  // src/diffEqSolver.cpp:112-127
            if (needE0) {
                e0 = u + 0.5 * S.mu() * (vx * vx + vy * vy + vz * vz);
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
            // A non-finite position or a runaway attempt can never reach the
            // exit test; it ends as a failed trajectory (status 2).
            const bool runaway = !(fabs(px) + fabs(py) + fabs(pz) < 1.0e30) || steps > A.max_steps;
            if (runaway) {
                if (j == 0) {
                    A.status[t] = 2; A.chi[t] = 1.57079632679489661923; A.steps[t] = steps; A.totsteps[t] = tot;
                }
                phase = PH_IDLE;
            } else if (!track) {
                tries++;
                dt = S.dt0() / (double)(tries + 1);
                if (!(tries < maxtries && maxtries < 10)) {
                    if (j == 0) {
                        A.status[t] = 2; A.chi[t] = 1.57079632679489661923; A.steps[t] = steps;
                        A.totsteps[t] = tot;
                    }
                    phase = PH_IDLE;
                }
            } else if (S.exits(px, py, pz) && steps > 10) {
                const double w0 = A.v0[3 * t], w1 = A.v0[3 * t + 1], w2 = A.v0[3 * t + 2];
                const double t1 = w0 * vx + w1 * vy + w2 * vz;
                const double t2 = sqrt(w0 * w0 + w1 * w1 + w2 * w2) * sqrt(vx * vx + vy * vy + vz * vz);
                chi = t1 / t2;  // cos(chi); acos is taken after the launch (k_chi)
                phase = PH_CHECK;
            }
        } else if (phase == PH_CHECK) {  // src/diffEqSolver.cpp:178-203
            const double e1 = u + 0.5 * S.mu() * (vx * vx + vy * vy + vz * vz);
            const double pct = 100 * fabs(e0 - e1) / e0;
            if (pct > 0.5) {
                tries++;
                maxtries++;
                dt = S.dt0() / (double)(tries + 1);
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
    unsigned long long s = n_steps, e = n_evals, p = n_pots, q = n_pairs;
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off);
        e += __shfl_xor(e, off);
        p += __shfl_xor(p, off);
        q += __shfl_xor(q, off);
    }
    if (lane == 0) {
        atomicAdd(A.evals, s);
        atomicAdd(A.evals + 1, e);
        atomicAdd(A.evals + 2, p);
        atomicAdd(A.evals + 3, q);
    }
    S.flush(A);
}

// The molecule of A into LDS for LdsSource: FP32 x, y, z, q | A, B, or the
// FP64 x, y, z, q, sigma^6, 24 eps/mu (npad each), by the whole workgroup.
template <int PREC>
__device__ __forceinline__ void stage_molecule(const GroupArgs &A, char *smem)
{
    const int npad = A.npad;
    const GLB double *S = glb(A.soa);
    if constexpr (PREC == 1) {
        float4 *xq = (float4 *)smem;
        float2 *se = (float2 *)(xq + npad);
        for (int i = threadIdx.x; i < npad; i += blockDim.x) {
            xq[i] = make_float4((float)S[i], (float)S[npad + i], (float)S[2 * npad + i], (float)S[3 * npad + i]);
            float a, b;
            lj_constants(S[4 * npad + i], S[5 * npad + i], a, b);
            se[i] = make_float2(a, b);
        }
    } else {
        double *X = (double *)smem;
        for (int i = threadIdx.x; i < 6 * npad; i += blockDim.x) X[i] = S[i];
    }
}

template <int PREC, int FT, int L>
__global__ void __launch_bounds__(1024) k_traj_grp(GroupArgs A)
{
    extern __shared__ __attribute__((aligned(16))) char smem[];
    // stage the molecule (once per workgroup; the kernel is persistent)
    stage_molecule<PREC>(A, smem);
    __syncthreads();
    TrajIO io = A.io;
    resolve_range(io, A.first, A.last);
    LdsSource<PREC, FT, L> src{A, smem};
    group_loop<L>(io, src);
}

// Rate probe of the force sum alone: the molecule staged in LDS as for
// k_traj_grp, then every group of L lanes evaluates the force `evals` times
// with the trajectory kernels' own force call (LdsSource::force: the same
// per-pair arithmetic, group sums and epilogue), at points 20 % beyond the
// boundary ellipsoid's a axis (the exact types' pair loop has no
// data-dependent branch, so the point does not change the work).  No Verlet
// state machine, no potential checks, and every group does the same work, so
// the launch has no drain: its pair rate is the ceiling of the trajectory
// kernel's instruction mix.  The sum of the forces goes to *sink only if it
// equals a value it never takes (so the work is not optimised away).
template <int PREC, int FT, int L>
__global__ void __launch_bounds__(1024) k_force_rate(GroupArgs A, int evals, double *sink)
{
    extern __shared__ __attribute__((aligned(16))) char smem[];
    stage_molecule<PREC>(A, smem);
    __syncthreads();
    LdsSource<PREC, FT, L> src{A, smem};
    const int j = threadIdx.x & (L - 1);
    const int grp = (blockIdx.x * blockDim.x + threadIdx.x) / L;
    const double x0 = 1.2 * A.m.a, y0 = 0.05 * A.m.b * (double)(grp % 7 - 3), z0 = 0.05 * A.m.c * (double)(grp % 5 - 2);
    double acc = 0;
    for (int e = 0; e < evals; e++) {
        double fx, fy, fz;
        src.force(true, j, x0 + 1e-3 * e, y0, z0, fx, fy, fz);
        acc += fx + fy + fz;
    }
    if (acc == -1.0e300) *sink = acc;
}

// The jobs of one launch shape in one persistent launch, each workgroup
// working on one job at a time with that job's molecule in its LDS
// (LdsMultiSource).  Dynamic LDS: the control block, then the largest
// molecule of the launch.
template <int PREC, int FT, int L>
__global__ void __launch_bounds__(1024) k_traj_lds_multi(LdsMultiArgs A)
{
    extern __shared__ __attribute__((aligned(16))) char smem[];
    WgCtl *ctl = (WgCtl *)smem;
    char *mol = smem + kCtlBytes;
    const int p0 = A.wg_pos[blockIdx.x];
    const MolRec *r0 = A.mols + A.pos_job[p0];
    {
        GroupArgs G{};
        G.npad = r0->npad;
        G.soa = r0->soa;
        stage_molecule<PREC>(G, mol);
    }
    if (threadIdx.x == 0) {
        ctl->pos = p0;
        ctl->gen = 0;
        ctl->active = 0;
        ctl->stager = 0;
    }
    __syncthreads();
    LdsMultiSource<PREC, FT, L> src{A, mol, ctl};
    TrajIO io = A.io;
    src.use(p0, io);
    group_loop<L>(io, src);
}

// Molecules beyond one workgroup's LDS, exact force types: the molecule
// streamed through LDS in tiles (TileSource); every wavefront stays in the
// trajectory loop until its workgroup is done.
template <int PREC, int FT, int L>
__global__ void __launch_bounds__(1024) k_traj_tile(GroupArgs A)
{
    extern __shared__ __attribute__((aligned(16))) char smem[];
    TrajIO io = A.io;
    resolve_range(io, A.first, A.last);
    TileSource<PREC, FT, L> src{A, smem};
    group_loop<L>(io, src);
}

// Dynamic LDS of k_traj_multi: the cluster lists of each wavefront and then
// every wavefront's branch counts when clustered types run one group per
// wavefront, else none.
__host__ __device__ constexpr bool multi_uses_lists(int ft, int L) { return ft != 4 && ft != 6 && L == 64; }

template <int PREC, int FT, int L>
__global__ void __launch_bounds__(1024) k_traj_multi(MultiArgs A)
{
    extern __shared__ int lds_lists[];
    constexpr bool kLists = multi_uses_lists(FT, L);
    int *lists = kLists ? lds_lists + (threadIdx.x >> 6) * kListInts : nullptr;
    WaveCount W{kLists ? (unsigned *)(lds_lists + (blockDim.x >> 6) * kListInts) + (threadIdx.x >> 6) * kCountInts
                       : nullptr};
    if (kLists) W.clear();
    TrajIO io = A.io;
    resolve_range(io, A.first, A.last);
    TableSource<PREC, FT, L> src{A, lists, W};
    group_loop<L>(io, src);
}

// A clustered-type molecule whose cluster headers fit LDS, mixed
// precision, 64 lanes a trajectory (LdsClusterSource).  Dynamic LDS: each
// wavefront's two cluster rings (kClRingInts) and branch counts, then the
// headers (cl_header_bytes), staged once per workgroup.
constexpr int kClRingInts = 2 * kRing;
__host__ __device__ constexpr size_t cl_lds_bytes(int ncl, int wg)
{
    return (size_t)(wg / 64) * (kClRingInts + kCountInts) * sizeof(int) + cl_header_bytes(ncl);
}

template <int FT>
__global__ void __launch_bounds__(1024) k_traj_cl_lds(GroupArgs A)
{
    extern __shared__ __attribute__((aligned(16))) int lds_cl[];
    const int nw = blockDim.x >> 6, w = threadIdx.x >> 6;
    int *lists = lds_cl + w * kClRingInts;
    WaveCount W{(unsigned *)(lds_cl + nw * kClRingInts) + w * kCountInts};
    W.clear();
    const int ncl = A.m.ncl;
    f4n *h0 = (f4n *)(lds_cl + nw * (kClRingInts + kCountInts));
    f4n *h1 = h0 + ncl;
    float *d2 = (float *)(h1 + ncl);
    unsigned long long *dm = (unsigned long long *)((char *)h0 + ((size_t)ncl * 36 + 7) / 8 * 8);
    const GLB f4n *G0 = (const GLB f4n *)glb(A.m.cg0);
    const GLB f4n *G1 = (const GLB f4n *)glb(A.m.cg1);
    const GLB f4n *G2 = (const GLB f4n *)glb(A.m.cg2);
    for (int i = threadIdx.x; i < ncl; i += blockDim.x) {
        h0[i] = G0[i];
        h1[i] = G1[i];
        d2[i] = G2[i].x;
    }
    for (int i = threadIdx.x; i < (ncl + 63) / 64; i += blockDim.x) dm[i] = A.m.cdm[i];
    __syncthreads();
    TrajIO io = A.io;
    resolve_range(io, A.first, A.last);
    LdsClusterSource<FT> src{A, lists, W, LdsHeaders{h0, h1, d2, dm}};
    group_loop<64>(io, src);
}

}  // namespace cosims
