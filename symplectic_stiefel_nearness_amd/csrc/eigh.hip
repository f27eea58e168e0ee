// Eigensolver launches of the Dual solver: the per-target rotation-only Jacobi kernel
// (eigh_kernel, in engine.h) and the eigenvector replay over every CU; the same two
// steps for the uneven-batch plan. Replaces torch.linalg.eigh at
// src/symplectic_projection/dual_recommended_hvp.py:108.
#include "engine.h"

namespace ssn {

__global__ __launch_bounds__(RP_ROWS * WAVE)
void eigh_replay_kernel(EighArgs a)
{
    __shared__ double2 par[2][RP_CHUNK * 64];
    eigh_replay_body(a.ws + (long long)blockIdx.y * target_doubles(a.mp), a.m, a.mp, blockIdx.x, par);
}

static size_t eigh_lds_doubles(int m, bool v_in_lds) {
    return 48 + (size_t)jac_area(m) + (v_in_lds ? (size_t)(m + 1) * (m + 1) : 0);   // V and its dummy row
}

#ifndef SSN_EIGH_REPLAY
#define SSN_EIGH_REPLAY 1
#endif

hipError_t launch_eigh(hipStream_t stream, const double* Cr, long long sC_b, int ldc, int batch,
                              int m, double* ws, size_t lds_limit) {
    if (SSN_EIGH_REPLAY) {
        // rotations only in the per-target kernel, then the eigenvectors replayed row by row
        const size_t lds = eigh_lds_doubles(m, false) * 8;
        if (lds > lds_limit) return LAUNCH_UNSUPPORTED;
        EighArgs e;
        e.Cr = Cr; e.sC_b = sC_b; e.ldc = ldc; e.m = m; e.mp = pad16(m); e.ws = ws; e.v_in_lds = 0;
        const int np = m / 2;
        const int blocks = (np * (np - 1) / 2 + JU - 1) / JU;
        if (blocks > JB_MAX) return LAUNCH_UNSUPPORTED;
        const bool small = blocks <= JB_BATCH;
        const void* fn = small ? (const void*)eigh_kernel<JB_BATCH, false, true> : (const void*)eigh_kernel<JB_MAX, false, true>;
        hipError_t err = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (err != hipSuccess) return err;
        if (small) hipLaunchKernelGGL((eigh_kernel<JB_BATCH, false, true>), dim3(batch), dim3(ET), lds, stream, e);
        else hipLaunchKernelGGL((eigh_kernel<JB_MAX, false, true>), dim3(batch), dim3(ET), lds, stream, e);
        err = hipGetLastError();
        if (err != hipSuccess) return err;
        hipLaunchKernelGGL(eigh_replay_kernel, dim3(e.mp / RP_ROWS, batch), dim3(RP_ROWS * WAVE), 0, stream, e);
        return hipGetLastError();
    }
    const bool v_in_lds = eigh_lds_doubles(m, true) * 8 <= lds_limit;
    const size_t lds = eigh_lds_doubles(m, v_in_lds) * 8;
    if (lds > lds_limit) return LAUNCH_UNSUPPORTED;
    EighArgs e;
    e.Cr = Cr; e.sC_b = sC_b; e.ldc = ldc; e.m = m; e.mp = pad16(m); e.ws = ws;
    e.v_in_lds = v_in_lds ? 1 : 0;
    const int np = m / 2;
    const int blocks = (np * (np - 1) / 2 + JU - 1) / JU;          // off-diagonal blocks per thread
    const int rows = (m + JU / np - 1) / (JU / np);                 // eigenvector rows per thread
    const bool small = blocks <= JB_BATCH && rows <= JV_SMALL;
    if (blocks > JB_MAX || rows > 2 * JV_HALF) return LAUNCH_UNSUPPORTED;
    const void* fn = small ? (v_in_lds ? (const void*)eigh_kernel<JB_BATCH, true> : (const void*)eigh_kernel<JB_BATCH, false>)
                           : (v_in_lds ? (const void*)eigh_kernel<JB_MAX, true> : (const void*)eigh_kernel<JB_MAX, false>);
    hipError_t err = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (err != hipSuccess) return err;
    if (small && v_in_lds) hipLaunchKernelGGL((eigh_kernel<JB_BATCH, true>), dim3(batch), dim3(ET), lds, stream, e);
    else if (small) hipLaunchKernelGGL((eigh_kernel<JB_BATCH, false>), dim3(batch), dim3(ET), lds, stream, e);
    else if (v_in_lds) hipLaunchKernelGGL((eigh_kernel<JB_MAX, true>), dim3(batch), dim3(ET), lds, stream, e);
    else hipLaunchKernelGGL((eigh_kernel<JB_MAX, false>), dim3(batch), dim3(ET), lds, stream, e);
    return hipGetLastError();
}

hipError_t launch_dual_eigh(hipStream_t stream, const double* Cr, long long sC_b, int ldc, int batch, int m,
                            double* ws, size_t lds_limit)
{
    return launch_eigh(stream, Cr, sC_b, ldc, batch, m, ws, lds_limit);
}

// ---------------------------------------------------------------- uneven batches
// One workgroup per target of the plan; each runs the rotation-only body of its own order
// (the eigenvectors follow in eigh_replay_uneven_kernel). The dynamic LDS of a launch is the
// largest of its targets' needs, which covers the rotation-only body.
__global__ __launch_bounds__(ET)
void eigh_uneven_kernel(const UTarget* __restrict__ tg)
{
    const UTarget& t = tg[blockIdx.x];
    EighArgs a;
    a.Cr = t.Cr; a.sC_b = 0; a.ldc = t.m; a.m = t.m; a.mp = t.mp; a.ws = t.dual; a.v_in_lds = 0;
    if (t.eigh_small) eigh_body<JB_BATCH, false, true>(a, 0);
    else eigh_body<JB_MAX, false, true>(a, 0);
}

// Eigenvector replay of every target of the plan: blockIdx.y = target, blockIdx.x = row block
// (grid sized for the largest order; blocks past a target's rows return at once).
__global__ __launch_bounds__(RP_ROWS * WAVE)
void eigh_replay_uneven_kernel(const UTarget* __restrict__ tg)
{
    __shared__ double2 par[2][RP_CHUNK * 64];
    const UTarget& t = tg[blockIdx.y];
    eigh_replay_body(t.dual, t.m, t.mp, blockIdx.x, par);
}

// eigensolver variant of order m on a device with lds_limit bytes of LDS per workgroup
void eigh_variant(int m, size_t lds_limit, int* small, int* vlds, size_t* lds_bytes) {
    const int np = m / 2;
    const int blocks = (np * (np - 1) / 2 + JU - 1) / JU;
    const int rows = (m + JU / np - 1) / (JU / np);
    *small = (blocks <= JB_BATCH && rows <= JV_SMALL) ? 1 : 0;
    *vlds = eigh_lds_doubles(m, true) * 8 <= lds_limit ? 1 : 0;
    *lds_bytes = eigh_lds_doubles(m, *vlds != 0) * 8;
}

hipError_t launch_eigh_uneven(hipStream_t stream, const UTarget* tg, int count, size_t eigh_lds)
{
    hipError_t err = hipFuncSetAttribute((const void*)eigh_uneven_kernel,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)eigh_lds);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(eigh_uneven_kernel, dim3(count), dim3(ET), eigh_lds, stream, tg);
    err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(eigh_replay_uneven_kernel, dim3(MAX_ORDER_PADDED / RP_ROWS, count), dim3(RP_ROWS * WAVE), 0,
                       stream, tg);
    return hipGetLastError();
}

}  // namespace ssn
