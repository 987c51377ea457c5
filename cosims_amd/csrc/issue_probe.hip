// issue_probe.hip -- VALU issue-rate probes for gfx950 (a measurement
// program beside the engine, built by cosims_amd/build.py as
// cosims_amd/issue_probe; not part of the library).
//
// Each probe runs a loop of known wave-instruction counts on every CU and
// reports wave-instructions issued per CU-cycle, with the shader clock read
// from clock64() against the constant-rate wall clock of one wavefront.  The
// probes show the rates the force loops' ceiling rests on:
//   * an FP32 VALU instruction with a scalar (SGPR) operand: 4 cycles a
//     wavefront (one wave-instruction per CU-cycle over the 4 SIMDs);
//   * the same instruction with all operands in VGPRs: 2 cycles (two per
//     CU-cycle), the rate of the published 157.3 TFLOP/s FP32 peak without
//     packed FP32;
//   * v_rsq_f32: 8 cycles;
//   * the exact force types' FP32 pair arithmetic (pair32: 30 VALU
//     instructions and one v_rsq_f32 a pair) with the atoms in registers
//     and from LDS, at 2 to 6 wavefronts a SIMD (probes 7, 8), and the LDS
//     loop with the next atom prefetched or two pairs a step (probe 10);
//   * FP32 VALU issue with the operands' VGPR banks chosen (probe 9).
// Usage: issue_probe [probe number]   (no argument: all)
#include <cstdio>
#include <cstdlib>

#include "traj_group.hpp"

using namespace cosims;

#define CHK(x)                                                                              \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) {                                                             \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);     \
            std::exit(1);                                                                   \
        }                                                                                   \
    } while (0)

// 16 independent chains a lane; OP selects the instruction.  a, b are uniform
// (SGPR operands); av, bv per-lane (VGPR operands).
template <int OP>
__global__ void __launch_bounds__(256) k_issue(float *out, int iters, float a, float b, long long *clk)
{
    float v[16];
#pragma unroll
    for (int i = 0; i < 16; i++) v[i] = threadIdx.x * 1e-3f + i + 1.0f;
    const float av = a + threadIdx.x * 1e-9f, bv = b + threadIdx.x * 1e-9f;
    const long long t0 = clock64(), w0 = wall_clock64();
    for (int k = 0; k < iters; k++) {
#pragma unroll
        for (int i = 0; i < 16; i++) {
            if constexpr (OP == 0) v[i] = __builtin_fmaf(v[i], a, b);          // v_fma_f32, SGPR operands
            else if constexpr (OP == 1) v[i] = v[i] * a;                       // v_mul_f32, SGPR operand
            else if constexpr (OP == 2) v[i] = __builtin_fmaf(v[i], av, bv);   // v_fma_f32, VGPR operands
            else if constexpr (OP == 3) v[i] = v[i] * av;                      // v_mul_f32, VGPR operands
            else if constexpr (OP == 4) v[i] = __builtin_fmaf(av, bv, v[i]);   // v_fmac_f32, VGPR operands
            else if constexpr (OP == 5) v[i] = __builtin_amdgcn_rsqf(v[i]);    // v_rsq_f32
            else if constexpr (OP == 6) v[i] = (i < 2) ? __builtin_amdgcn_rsqf(v[i]) : __builtin_fmaf(av, bv, v[i]);
        }
    }
    const long long t1 = clock64(), w1 = wall_clock64();
    float s = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) s += v[i];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        clk[0] = t1 - t0;
        clk[1] = w1 - w0;
    }
}

// pair32 with 8 atoms held in registers (no memory traffic in the loop)
__global__ void __launch_bounds__(512) k_pair_regs(float *out, int iters, long long *clk)
{
    f4n a[8];
    f2n c[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        a[k] = f4n{1.0f + threadIdx.x * 1e-3f + k, 2.0f - k * 0.1f, 0.5f * k, 0.1f * (k - 4)};
        c[k] = f2n{1e5f + k, 1e2f - k};
    }
    const float px = 30.0f + threadIdx.x * 1e-4f, py = 1.0f, pz = -2.0f;
    float s[12];
#pragma unroll
    for (int k = 0; k < 12; k++) s[k] = 0.0f;
    const long long t0 = clock64(), w0 = wall_clock64();
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int k = 0; k < 8; k++) pair32<4>(a[k], c[k], px + it * 1e-6f, py, pz, s);
    }
    const long long t1 = clock64(), w1 = wall_clock64();
    float t = 0;
#pragma unroll
    for (int k = 0; k < 12; k++) t += s[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = t;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        clk[0] = t1 - t0;
        clk[1] = w1 - w0;
    }
}

// pair32 over npad atoms in LDS, lanes j, j+16, ... (16-lane groups)
__global__ void __launch_bounds__(512) k_pair_lds(float *out, int iters, int npad, long long *clk)
{
    extern __shared__ __attribute__((aligned(16))) char smem[];
    f4n *xq = (f4n *)smem;
    f2n *ab = (f2n *)(xq + npad);
    for (int i = threadIdx.x; i < npad; i += blockDim.x) {
        xq[i] = f4n{1.0f * (i % 37), 0.5f * (i % 11), 0.25f * (i % 13), 0.01f * (i % 7)};
        ab[i] = f2n{1e5f, 1e2f};
    }
    __syncthreads();
    const int j = threadIdx.x & 15;
    const float px = 60.0f + threadIdx.x * 1e-4f, py = 1.0f, pz = -2.0f;
    float s[12];
#pragma unroll
    for (int k = 0; k < 12; k++) s[k] = 0.0f;
    const long long t0 = clock64(), w0 = wall_clock64();
    for (int it = 0; it < iters; it++) {
#pragma unroll 4
        for (int kk = 0; kk < npad; kk += 16) pair32<4>(xq[kk + j], ab[kk + j], px + it * 1e-6f, py, pz, s);
    }
    const long long t1 = clock64(), w1 = wall_clock64();
    float t = 0;
#pragma unroll
    for (int k = 0; k < 12; k++) t += s[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = t;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        clk[0] = t1 - t0;
        clk[1] = w1 - w0;
    }
}

// VALU issue with the operand registers chosen (inline asm, VALU only; the
// VGPR bank of vN is N mod 4): 16 independent accumulators, the sources
// shared.  Variant V selects the instruction and the banks.
template <int V>
__global__ void __launch_bounds__(256) k_bank(float *out, int iters, long long *clk)
{
    asm volatile("v_mov_b32 v1, 1.0\n v_mov_b32 v2, 0.5\n v_mov_b32 v5, 0.5\n" ::: "v1", "v2", "v5");
    asm volatile("v_mov_b32 v32, 0\n""v_mov_b32 v33, 0\n""v_mov_b32 v34, 0\n""v_mov_b32 v35, 0\n""v_mov_b32 v36, 0\n""v_mov_b32 v37, 0\n""v_mov_b32 v38, 0\n""v_mov_b32 v39, 0\n""v_mov_b32 v40, 0\n""v_mov_b32 v41, 0\n""v_mov_b32 v42, 0\n""v_mov_b32 v43, 0\n""v_mov_b32 v44, 0\n""v_mov_b32 v45, 0\n""v_mov_b32 v46, 0\n""v_mov_b32 v47, 0\n""v_mov_b32 v48, 0\n""v_mov_b32 v49, 0\n""v_mov_b32 v50, 0\n""v_mov_b32 v51, 0\n""v_mov_b32 v52, 0\n""v_mov_b32 v53, 0\n""v_mov_b32 v54, 0\n""v_mov_b32 v55, 0\n""v_mov_b32 v56, 0\n""v_mov_b32 v57, 0\n""v_mov_b32 v58, 0\n""v_mov_b32 v59, 0\n""v_mov_b32 v60, 0\n""v_mov_b32 v61, 0\n""v_mov_b32 v62, 0\n""v_mov_b32 v63, 0\n""v_mov_b32 v64, 0\n""v_mov_b32 v65, 0\n""v_mov_b32 v66, 0\n""v_mov_b32 v67, 0\n""v_mov_b32 v68, 0\n""v_mov_b32 v69, 0\n""v_mov_b32 v70, 0\n""v_mov_b32 v71, 0\n""v_mov_b32 v72, 0\n""v_mov_b32 v73, 0\n""v_mov_b32 v74, 0\n""v_mov_b32 v75, 0\n""v_mov_b32 v76, 0\n""v_mov_b32 v77, 0\n""v_mov_b32 v78, 0\n""v_mov_b32 v79, 0\n""v_mov_b32 v80, 0\n""v_mov_b32 v81, 0\n""v_mov_b32 v82, 0\n""v_mov_b32 v83, 0\n""v_mov_b32 v84, 0\n""v_mov_b32 v85, 0\n""v_mov_b32 v86, 0\n""v_mov_b32 v87, 0\n""v_mov_b32 v88, 0\n""v_mov_b32 v89, 0\n""v_mov_b32 v90, 0\n""v_mov_b32 v91, 0\n""v_mov_b32 v92, 0\n""v_mov_b32 v93, 0\n""v_mov_b32 v94, 0\n""v_mov_b32 v95, 0\n" ::: "v32","v33","v34","v35","v36","v37","v38","v39","v40","v41","v42","v43","v44","v45","v46","v47","v48","v49","v50","v51","v52","v53","v54","v55","v56","v57","v58","v59","v60","v61","v62","v63","v64","v65","v66","v67","v68","v69","v70","v71","v72","v73","v74","v75","v76","v77","v78","v79","v80","v81","v82","v83","v84","v85","v86","v87","v88","v89","v90","v91","v92","v93","v94","v95","v1","v2","v5");
    const long long t0 = clock64(), w0 = wall_clock64();
    for (int k = 0; k < iters; k++) {
        if constexpr (V == 0)  // fmac: dst, src0, src1 in three banks
            asm volatile("v_fmac_f32 v32, v1, v2\n""v_fmac_f32 v36, v1, v2\n""v_fmac_f32 v40, v1, v2\n""v_fmac_f32 v44, v1, v2\n""v_fmac_f32 v48, v1, v2\n""v_fmac_f32 v52, v1, v2\n""v_fmac_f32 v56, v1, v2\n""v_fmac_f32 v60, v1, v2\n""v_fmac_f32 v64, v1, v2\n""v_fmac_f32 v68, v1, v2\n""v_fmac_f32 v72, v1, v2\n""v_fmac_f32 v76, v1, v2\n""v_fmac_f32 v80, v1, v2\n""v_fmac_f32 v84, v1, v2\n""v_fmac_f32 v88, v1, v2\n""v_fmac_f32 v92, v1, v2\n" ::: "v32","v33","v34","v35","v36","v37","v38","v39","v40","v41","v42","v43","v44","v45","v46","v47","v48","v49","v50","v51","v52","v53","v54","v55","v56","v57","v58","v59","v60","v61","v62","v63","v64","v65","v66","v67","v68","v69","v70","v71","v72","v73","v74","v75","v76","v77","v78","v79","v80","v81","v82","v83","v84","v85","v86","v87","v88","v89","v90","v91","v92","v93","v94","v95","v1","v2","v5");
        else if constexpr (V == 1)  // fmac: src0 and src1 in one bank
            asm volatile("v_fmac_f32 v32, v1, v5\n""v_fmac_f32 v36, v1, v5\n""v_fmac_f32 v40, v1, v5\n""v_fmac_f32 v44, v1, v5\n""v_fmac_f32 v48, v1, v5\n""v_fmac_f32 v52, v1, v5\n""v_fmac_f32 v56, v1, v5\n""v_fmac_f32 v60, v1, v5\n""v_fmac_f32 v64, v1, v5\n""v_fmac_f32 v68, v1, v5\n""v_fmac_f32 v72, v1, v5\n""v_fmac_f32 v76, v1, v5\n""v_fmac_f32 v80, v1, v5\n""v_fmac_f32 v84, v1, v5\n""v_fmac_f32 v88, v1, v5\n""v_fmac_f32 v92, v1, v5\n" ::: "v32","v33","v34","v35","v36","v37","v38","v39","v40","v41","v42","v43","v44","v45","v46","v47","v48","v49","v50","v51","v52","v53","v54","v55","v56","v57","v58","v59","v60","v61","v62","v63","v64","v65","v66","v67","v68","v69","v70","v71","v72","v73","v74","v75","v76","v77","v78","v79","v80","v81","v82","v83","v84","v85","v86","v87","v88","v89","v90","v91","v92","v93","v94","v95","v1","v2","v5");
        else if constexpr (V == 2)  // fmac: dst (accumulator) in src0's bank
            asm volatile("v_fmac_f32 v33, v1, v2\n""v_fmac_f32 v37, v1, v2\n""v_fmac_f32 v41, v1, v2\n""v_fmac_f32 v45, v1, v2\n""v_fmac_f32 v49, v1, v2\n""v_fmac_f32 v53, v1, v2\n""v_fmac_f32 v57, v1, v2\n""v_fmac_f32 v61, v1, v2\n""v_fmac_f32 v65, v1, v2\n""v_fmac_f32 v69, v1, v2\n""v_fmac_f32 v73, v1, v2\n""v_fmac_f32 v77, v1, v2\n""v_fmac_f32 v81, v1, v2\n""v_fmac_f32 v85, v1, v2\n""v_fmac_f32 v89, v1, v2\n""v_fmac_f32 v93, v1, v2\n" ::: "v32","v33","v34","v35","v36","v37","v38","v39","v40","v41","v42","v43","v44","v45","v46","v47","v48","v49","v50","v51","v52","v53","v54","v55","v56","v57","v58","v59","v60","v61","v62","v63","v64","v65","v66","v67","v68","v69","v70","v71","v72","v73","v74","v75","v76","v77","v78","v79","v80","v81","v82","v83","v84","v85","v86","v87","v88","v89","v90","v91","v92","v93","v94","v95","v1","v2","v5");
        else if constexpr (V == 3)  // fmac: dst, src0, src1 in one bank
            asm volatile("v_fmac_f32 v33, v1, v5\n""v_fmac_f32 v37, v1, v5\n""v_fmac_f32 v41, v1, v5\n""v_fmac_f32 v45, v1, v5\n""v_fmac_f32 v49, v1, v5\n""v_fmac_f32 v53, v1, v5\n""v_fmac_f32 v57, v1, v5\n""v_fmac_f32 v61, v1, v5\n""v_fmac_f32 v65, v1, v5\n""v_fmac_f32 v69, v1, v5\n""v_fmac_f32 v73, v1, v5\n""v_fmac_f32 v77, v1, v5\n""v_fmac_f32 v81, v1, v5\n""v_fmac_f32 v85, v1, v5\n""v_fmac_f32 v89, v1, v5\n""v_fmac_f32 v93, v1, v5\n" ::: "v32","v33","v34","v35","v36","v37","v38","v39","v40","v41","v42","v43","v44","v45","v46","v47","v48","v49","v50","v51","v52","v53","v54","v55","v56","v57","v58","v59","v60","v61","v62","v63","v64","v65","v66","v67","v68","v69","v70","v71","v72","v73","v74","v75","v76","v77","v78","v79","v80","v81","v82","v83","v84","v85","v86","v87","v88","v89","v90","v91","v92","v93","v94","v95","v1","v2","v5");
        else if constexpr (V == 4)  // mul: src0 = src1
            asm volatile("v_mul_f32 v32, v1, v1\n""v_mul_f32 v36, v1, v1\n""v_mul_f32 v40, v1, v1\n""v_mul_f32 v44, v1, v1\n""v_mul_f32 v48, v1, v1\n""v_mul_f32 v52, v1, v1\n""v_mul_f32 v56, v1, v1\n""v_mul_f32 v60, v1, v1\n""v_mul_f32 v64, v1, v1\n""v_mul_f32 v68, v1, v1\n""v_mul_f32 v72, v1, v1\n""v_mul_f32 v76, v1, v1\n""v_mul_f32 v80, v1, v1\n""v_mul_f32 v84, v1, v1\n""v_mul_f32 v88, v1, v1\n""v_mul_f32 v92, v1, v1\n" ::: "v32","v33","v34","v35","v36","v37","v38","v39","v40","v41","v42","v43","v44","v45","v46","v47","v48","v49","v50","v51","v52","v53","v54","v55","v56","v57","v58","v59","v60","v61","v62","v63","v64","v65","v66","v67","v68","v69","v70","v71","v72","v73","v74","v75","v76","v77","v78","v79","v80","v81","v82","v83","v84","v85","v86","v87","v88","v89","v90","v91","v92","v93","v94","v95","v1","v2","v5");
        else if constexpr (V == 5)  // mul: src0 and src1 in one bank
            asm volatile("v_mul_f32 v32, v1, v5\n""v_mul_f32 v36, v1, v5\n""v_mul_f32 v40, v1, v5\n""v_mul_f32 v44, v1, v5\n""v_mul_f32 v48, v1, v5\n""v_mul_f32 v52, v1, v5\n""v_mul_f32 v56, v1, v5\n""v_mul_f32 v60, v1, v5\n""v_mul_f32 v64, v1, v5\n""v_mul_f32 v68, v1, v5\n""v_mul_f32 v72, v1, v5\n""v_mul_f32 v76, v1, v5\n""v_mul_f32 v80, v1, v5\n""v_mul_f32 v84, v1, v5\n""v_mul_f32 v88, v1, v5\n""v_mul_f32 v92, v1, v5\n" ::: "v32","v33","v34","v35","v36","v37","v38","v39","v40","v41","v42","v43","v44","v45","v46","v47","v48","v49","v50","v51","v52","v53","v54","v55","v56","v57","v58","v59","v60","v61","v62","v63","v64","v65","v66","v67","v68","v69","v70","v71","v72","v73","v74","v75","v76","v77","v78","v79","v80","v81","v82","v83","v84","v85","v86","v87","v88","v89","v90","v91","v92","v93","v94","v95","v1","v2","v5");
        else if constexpr (V == 6)  // mul: three banks
            asm volatile("v_mul_f32 v32, v1, v2\n""v_mul_f32 v36, v1, v2\n""v_mul_f32 v40, v1, v2\n""v_mul_f32 v44, v1, v2\n""v_mul_f32 v48, v1, v2\n""v_mul_f32 v52, v1, v2\n""v_mul_f32 v56, v1, v2\n""v_mul_f32 v60, v1, v2\n""v_mul_f32 v64, v1, v2\n""v_mul_f32 v68, v1, v2\n""v_mul_f32 v72, v1, v2\n""v_mul_f32 v76, v1, v2\n""v_mul_f32 v80, v1, v2\n""v_mul_f32 v84, v1, v2\n""v_mul_f32 v88, v1, v2\n""v_mul_f32 v92, v1, v2\n" ::: "v32","v33","v34","v35","v36","v37","v38","v39","v40","v41","v42","v43","v44","v45","v46","v47","v48","v49","v50","v51","v52","v53","v54","v55","v56","v57","v58","v59","v60","v61","v62","v63","v64","v65","v66","v67","v68","v69","v70","v71","v72","v73","v74","v75","v76","v77","v78","v79","v80","v81","v82","v83","v84","v85","v86","v87","v88","v89","v90","v91","v92","v93","v94","v95","v1","v2","v5");
        else if constexpr (V == 7)  // fmac: an inline constant source
            asm volatile("v_fmac_f32 v32, 0.5, v2\n""v_fmac_f32 v36, 0.5, v2\n""v_fmac_f32 v40, 0.5, v2\n""v_fmac_f32 v44, 0.5, v2\n""v_fmac_f32 v48, 0.5, v2\n""v_fmac_f32 v52, 0.5, v2\n""v_fmac_f32 v56, 0.5, v2\n""v_fmac_f32 v60, 0.5, v2\n""v_fmac_f32 v64, 0.5, v2\n""v_fmac_f32 v68, 0.5, v2\n""v_fmac_f32 v72, 0.5, v2\n""v_fmac_f32 v76, 0.5, v2\n""v_fmac_f32 v80, 0.5, v2\n""v_fmac_f32 v84, 0.5, v2\n""v_fmac_f32 v88, 0.5, v2\n""v_fmac_f32 v92, 0.5, v2\n" ::: "v32","v33","v34","v35","v36","v37","v38","v39","v40","v41","v42","v43","v44","v45","v46","v47","v48","v49","v50","v51","v52","v53","v54","v55","v56","v57","v58","v59","v60","v61","v62","v63","v64","v65","v66","v67","v68","v69","v70","v71","v72","v73","v74","v75","v76","v77","v78","v79","v80","v81","v82","v83","v84","v85","v86","v87","v88","v89","v90","v91","v92","v93","v94","v95","v1","v2","v5");
        else if constexpr (V == 8)  // mul: a literal constant source
            asm volatile("v_mul_f32 v32, 0x3dcccccd, v2\n""v_mul_f32 v36, 0x3dcccccd, v2\n""v_mul_f32 v40, 0x3dcccccd, v2\n""v_mul_f32 v44, 0x3dcccccd, v2\n""v_mul_f32 v48, 0x3dcccccd, v2\n""v_mul_f32 v52, 0x3dcccccd, v2\n""v_mul_f32 v56, 0x3dcccccd, v2\n""v_mul_f32 v60, 0x3dcccccd, v2\n""v_mul_f32 v64, 0x3dcccccd, v2\n""v_mul_f32 v68, 0x3dcccccd, v2\n""v_mul_f32 v72, 0x3dcccccd, v2\n""v_mul_f32 v76, 0x3dcccccd, v2\n""v_mul_f32 v80, 0x3dcccccd, v2\n""v_mul_f32 v84, 0x3dcccccd, v2\n""v_mul_f32 v88, 0x3dcccccd, v2\n""v_mul_f32 v92, 0x3dcccccd, v2\n" ::: "v32","v33","v34","v35","v36","v37","v38","v39","v40","v41","v42","v43","v44","v45","v46","v47","v48","v49","v50","v51","v52","v53","v54","v55","v56","v57","v58","v59","v60","v61","v62","v63","v64","v65","v66","v67","v68","v69","v70","v71","v72","v73","v74","v75","v76","v77","v78","v79","v80","v81","v82","v83","v84","v85","v86","v87","v88","v89","v90","v91","v92","v93","v94","v95","v1","v2","v5");
    }
    const long long t1 = clock64(), w1 = wall_clock64();
    float r;
    asm volatile("v_mov_b32 %0, v32" : "=v"(r));
    out[blockIdx.x * blockDim.x + threadIdx.x] = r;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        clk[0] = t1 - t0;
        clk[1] = w1 - w0;
    }
}

// pair32 over npad atoms in LDS, lanes j, j+16, ... (16-lane groups), the
// probe point per lane (all operands in VGPRs, as in the trajectory
// kernels): atoms loaded where used (MODE 0), the next pair's atom
// prefetched into registers of its own (MODE 1) or two pairs a step (MODE 2)
template <int MODE>
__global__ void __launch_bounds__(512) k_pair_lds_v(float *out, int iters, int npad, long long *clk)
{
    extern __shared__ __attribute__((aligned(16))) char smem[];
    f4n *xq = (f4n *)smem;
    f2n *ab = (f2n *)(xq + npad);
    for (int i = threadIdx.x; i < npad; i += blockDim.x) {
        xq[i] = f4n{1.0f * (i % 37), 0.5f * (i % 11), 0.25f * (i % 13), 0.01f * (i % 7)};
        ab[i] = f2n{1e5f, 1e2f};
    }
    __syncthreads();
    const int j = threadIdx.x & 15;
    const float px = 60.0f + threadIdx.x * 1e-4f, py = 1.0f + threadIdx.x * 1e-5f, pz = -2.0f - threadIdx.x * 1e-5f;
    float s[12];
#pragma unroll
    for (int k = 0; k < 12; k++) s[k] = 0.0f;
    const long long t0 = clock64(), w0 = wall_clock64();
    for (int it = 0; it < iters; it++) {
        const float qx = px + it * 1e-6f;
        if constexpr (MODE == 0) {
#pragma unroll 4
            for (int kk = 0; kk < npad; kk += 16) pair32<4>(xq[kk + j], ab[kk + j], qx, py, pz, s);
        } else if constexpr (MODE == 1) {
            f4n a = xq[j];
            f2n c = ab[j];
#pragma unroll 4
            for (int kk = 16; kk < npad; kk += 16) {
                const f4n an = xq[kk + j];
                const f2n cn = ab[kk + j];
                pair32<4>(a, c, qx, py, pz, s);
                a = an;
                c = cn;
            }
            pair32<4>(a, c, qx, py, pz, s);
        } else {
#pragma unroll 2
            for (int kk = 0; kk < npad; kk += 32) {
                const f4n a0 = xq[kk + j], a1 = xq[kk + 16 + j];
                const f2n c0 = ab[kk + j], c1 = ab[kk + 16 + j];
                pair32<4>(a0, c0, qx, py, pz, s);
                pair32<4>(a1, c1, qx, py, pz, s);
            }
        }
    }
    const long long t1 = clock64(), w1 = wall_clock64();
    float t = 0;
#pragma unroll
    for (int k = 0; k < 12; k++) t += s[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = t;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        clk[0] = t1 - t0;
        clk[1] = w1 - w0;
    }
}

namespace {

int g_cus = 0, g_wclk_khz = 0;
float *g_buf = nullptr;
long long *g_clk = nullptr;

// best of three launches: (ms, shader MHz of block 0's first wavefront)
template <class Launch>
void time3(Launch launch, float &best, double &mhz)
{
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    best = 1e30f;
    for (int r = 0; r < 3; r++) {
        CHK(hipEventRecord(e0, 0));
        launch();
        CHK(hipGetLastError());
        CHK(hipEventRecord(e1, 0));
        CHK(hipEventSynchronize(e1));
        float ms = 0;
        CHK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            long long hc[2];
            CHK(hipMemcpy(hc, g_clk, sizeof(hc), hipMemcpyDeviceToHost));
            mhz = (double)hc[0] / ((double)hc[1] / (g_wclk_khz * 1e3)) / 1e6;
        }
    }
    CHK(hipEventDestroy(e0));
    CHK(hipEventDestroy(e1));
}

template <int OP>
void issue(const char *name, int iters)
{
    const int blocks = g_cus * 32, threads = 256;
    float ms;
    double mhz = 0;
    time3([&] { hipLaunchKernelGGL(k_issue<OP>, dim3(blocks), dim3(threads), 0, 0, g_buf, iters, 0.999f, 0.001f,
                                   g_clk); }, ms, mhz);
    const double instr = (double)blocks * threads / 64 * 16.0 * iters;
    std::printf("%-34s %8.3f ms  %5.0f MHz  %.3f wave-instr/CU-cycle\n", name, ms, mhz,
                instr / (ms * 1e-3 * mhz * 1e6 * g_cus));
}

void pair_probe(bool lds, int waves_per_simd)
{
    const int bpc = waves_per_simd / 2;  // 512-thread workgroups: 2 wavefronts a SIMD each
    const int blocks = g_cus * bpc, npad = 2048;
    const size_t bytes = (size_t)npad * 24;
    float ms;
    double mhz = 0;
    double lane_pairs;
    if (lds) {
        const int iters = 60;
        CHK(hipFuncSetAttribute((const void *)k_pair_lds, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
        time3([&] { hipLaunchKernelGGL(k_pair_lds, dim3(blocks), dim3(512), bytes, 0, g_buf, iters, npad, g_clk); },
              ms, mhz);
        lane_pairs = (double)blocks * 512 * iters * (npad / 16);
    } else {
        const int iters = 4000;
        time3([&] { hipLaunchKernelGGL(k_pair_regs, dim3(blocks), dim3(512), 0, 0, g_buf, iters, g_clk); }, ms,
              mhz);
        lane_pairs = (double)blocks * 512 * iters * 8.0;
    }
    std::printf("pair32, atoms in %-9s %d waves/SIMD %8.3f ms  %5.0f MHz  %.3e lane-pairs/s  %.1f TFLOP/s at 58 a pair\n",
                lds ? "LDS," : "registers,", waves_per_simd, ms, mhz, lane_pairs / (ms * 1e-3),
                lane_pairs * 58 / (ms * 1e-3) / 1e12);
}

template <int V>
void bank(const char *name, int iters)
{
    const int blocks = g_cus * 32, threads = 256;
    float ms;
    double mhz = 0;
    time3([&] { hipLaunchKernelGGL(k_bank<V>, dim3(blocks), dim3(threads), 0, 0, g_buf, iters, g_clk); }, ms, mhz);
    const double instr = (double)blocks * threads / 64 * 16.0 * iters;
    std::printf("%-44s %8.3f ms  %5.0f MHz  %.3f wave-instr/CU-cycle\n", name, ms, mhz,
                instr / (ms * 1e-3 * mhz * 1e6 * g_cus));
}

template <int MODE>
void pair_lds_variant(int waves_per_simd, int npad)
{
    const int bpc = waves_per_simd / 2;
    const int blocks = g_cus * bpc, iters = 60 * 2048 / npad;
    const size_t bytes = (size_t)npad * 24;
    float ms;
    double mhz = 0;
    CHK(hipFuncSetAttribute((const void *)k_pair_lds_v<MODE>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)bytes));
    time3([&] { hipLaunchKernelGGL(k_pair_lds_v<MODE>, dim3(blocks), dim3(512), bytes, 0, g_buf, iters, npad,
                                   g_clk); }, ms, mhz);
    const double lane_pairs = (double)blocks * 512 * iters * (npad / 16);
    std::printf("pair32 LDS, %-22s %d waves/SIMD npad %d %8.3f ms  %5.0f MHz  %.3e lane-pairs/s  %.1f TFLOP/s\n",
                MODE == 0 ? "loaded where used," : MODE == 1 ? "next pair prefetched," : "two pairs a step,",
                waves_per_simd, npad, ms, mhz,
                lane_pairs / (ms * 1e-3), lane_pairs * 58 / (ms * 1e-3) / 1e12);
}

}  // namespace

int main(int argc, char **argv)
{
    const int only = argc > 1 ? std::atoi(argv[1]) : -1;
    hipDeviceProp_t p;
    CHK(hipGetDeviceProperties(&p, 0));
    g_cus = p.multiProcessorCount;
    CHK(hipDeviceGetAttribute(&g_wclk_khz, hipDeviceAttributeWallClockRate, 0));
    std::printf("%s, %d CUs\n", p.gcnArchName, g_cus);
    CHK(hipMalloc(&g_buf, sizeof(float) * g_cus * 32 * 512));
    CHK(hipMalloc(&g_clk, 2 * sizeof(long long)));
    const int it = 20000;
    if (only < 0 || only == 0) issue<0>("v_fma_f32, SGPR operands", it);
    if (only < 0 || only == 1) issue<1>("v_mul_f32, SGPR operand", it);
    if (only < 0 || only == 2) issue<2>("v_fma_f32, VGPR operands", it);
    if (only < 0 || only == 3) issue<3>("v_mul_f32, VGPR operands", it);
    if (only < 0 || only == 4) issue<4>("v_fmac_f32, VGPR operands", it);
    if (only < 0 || only == 5) issue<5>("v_rsq_f32", it / 4);
    if (only < 0 || only == 6) issue<6>("14 v_fmac_f32 (VGPR) + 2 v_rsq_f32", it);
    if (only < 0 || only == 7)
        for (int w : {2, 4}) pair_probe(false, w);
    if (only < 0 || only == 8)
        for (int w : {2, 4, 6}) pair_probe(true, w);
    if (only < 0 || only == 9) {
        bank<0>("fmac: dst, src0, src1 in three banks", it);
        bank<1>("fmac: src0 and src1 in one bank", it);
        bank<2>("fmac: dst (accumulator) in src0's bank", it);
        bank<3>("fmac: dst, src0, src1 in one bank", it);
        bank<4>("mul: src0 = src1", it);
        bank<5>("mul: src0 and src1 in one bank", it);
        bank<6>("mul: three banks", it);
        bank<7>("fmac: an inline constant source", it);
        bank<8>("mul: a literal constant source", it);
    }
    if (only < 0 || only == 10)
        for (int w : {2, 4, 6}) {
            pair_lds_variant<0>(w, 2048);
            pair_lds_variant<1>(w, 2048);
            pair_lds_variant<2>(w, 2048);
        }
    return 0;
}
