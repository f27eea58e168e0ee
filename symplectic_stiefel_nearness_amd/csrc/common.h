// Shared device helpers for the gfx950 (CDNA4) kernels of the Dual solver.
//
// Conventions used by every kernel in this library:
//  * m = 2k is the column count of a target; small m x m matrices are stored
//    row-major, padded to mp = 16 * ceil(m / 16) with zero padding, so that the
//    FP64 MFMA 16x16x4 tiles never need edge masks.
//  * wave64 everywhere; lane = threadIdx.x & 63.
//  * FP64 MFMA operand layout (v_mfma_f64_16x16x4f64), lane l = r + 16 q
//    (r = l & 15, q = l >> 4):
//        A (16 x 4):  lane holds A[r][q]
//        B (4 x 16):  lane holds B[q][r]
//        D (16 x 16): lane holds D[q + 4 i][r], i = 0..3 (rows strided by 4)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ssn {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int WAVE = 64;

__device__ __forceinline__ d4 mfma(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ d4 zero4() {
    d4 z; z[0] = 0.0; z[1] = 0.0; z[2] = 0.0; z[3] = 0.0;
    return z;
}

__host__ __device__ __forceinline__ int pad16(int m) { return (m + 15) & ~15; }

// DPP move of a double (two 32-bit halves); lanes without a source read 0.
template <int CTRL, int ROWMASK, int BANKMASK>
__device__ __forceinline__ double dpp(double v) {
    const long long x = __double_as_longlong(v);
    int lo = (int)(x & 0xffffffffLL), hi = (int)(x >> 32);
    lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, ROWMASK, BANKMASK, true);
    hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, ROWMASK, BANKMASK, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(x & 0xffffffffLL), lane);
    const int hi = __builtin_amdgcn_readlane((int)(x >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Deterministic wave-level sum with DPP (fixed order): quad butterflies, 16-lane row
// rotations, then row broadcasts; lane 63 ends with the total, returned wave-uniform.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp<0xB1, 0xF, 0xF>(v);    // quad_perm [1,0,3,2]
    v += dpp<0x4E, 0xF, 0xF>(v);    // quad_perm [2,3,0,1]
    v += dpp<0x124, 0xF, 0xF>(v);   // row_ror:4
    v += dpp<0x128, 0xF, 0xF>(v);   // row_ror:8
    v += dpp<0x142, 0xA, 0xF>(v);   // row_bcast:15 into rows 1, 3
    v += dpp<0x143, 0xC, 0xF>(v);   // row_bcast:31 into rows 2, 3
    return readlane_d(v, 63);
}

// Launcher results that are not HIP errors. Both values lie outside HIP's error enumeration,
// so no HIP call returns them: LAUNCH_UNSUPPORTED says the shape does not fit this device
// (the C ABI reports SSN_ERR_UNSUPPORTED); LAUNCH_DECLINED says a launcher does not handle
// the shape and the caller takes another path. Every other value is a HIP error (SSN_ERR_HIP).
constexpr hipError_t LAUNCH_UNSUPPORTED = static_cast<hipError_t>(2000);
constexpr hipError_t LAUNCH_DECLINED = static_cast<hipError_t>(2001);

// Iteration caps and the stopping tolerance of the Newton / PCG / Armijo iteration; the
// defaults are the reference's constants (dual_recommended_hvp.py:15-16).
struct Caps {
    int max_newton = 100;       // Newton updates; reaching the cap is ST_NEWTON_CAP
    int max_cg = 150;           // CG iterations per Newton step
    int max_backtrack = 40;     // Armijo trials per Newton step; exhausting them is ST_NO_DESCENT
    double feas_tol = 1e-8;     // a target stops when its scaled residual < feas_tol / 2
    // launch form, not a cap: the eigensolver may take every CU for its one-launch form (1), or
    // must leave CUs to other work running beside the solve (0: the two-launch form)
    int eigh_whole_device = 1;
};

// Status codes shared by kernels and the C ABI (include/ssn.h mirrors them).
enum Status : int {
    ST_OK = 0,
    ST_NO_DESCENT = 1,   // Armijo backtracking exhausted
    ST_NEWTON_CAP = 2,   // Newton iteration cap reached
    ST_NONFINITE = 3,    // non-finite Gram / residual
    ST_EIGH = 4,         // Jacobi eigensolver did not converge
};

}  // namespace ssn
