// Newton / PCG / Armijo launches of the Dual solver (dual_body, in engine.h): one 8-wave
// workgroup per target. Replaces dual_batch (src/symplectic_projection/dual_recommended_hvp.py:99-191).
#include "engine.h"

namespace ssn {

hipError_t launch_eigh(hipStream_t stream, const double* Cr, long long sC_b, int ldc, int batch,
                       int m, double* ws, size_t lds_limit, bool whole_device);

template <int NT>
__global__ __launch_bounds__(DT)
void dual_kernel(DualArgs a)
{
    dual_body<NT>(a, blockIdx.x);
}

long long workspace_target_doubles(int m) { return target_doubles(pad16(m)); }

// ---------------------------------------------------------------- launchers
size_t dual_workspace_doubles(int batch, int m) {
    return (size_t)batch * (size_t)target_doubles(pad16(m));
}

template <int NT>
static hipError_t launch_nt(hipStream_t stream, DualArgs a, int batch, size_t lds_limit) {
    const size_t lds = (size_t)Cfg<NT>::LDS_END * 8;
    if (lds > lds_limit) return LAUNCH_UNSUPPORTED;
    hipError_t err = hipFuncSetAttribute((const void*)dual_kernel<NT>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(dual_kernel<NT>, dim3(batch), dim3(DT), lds, stream, a);
    return hipGetLastError();
}

hipError_t launch_dual_newton(hipStream_t stream, const double* Cr, const double* Ci, long long sC_b, int ldc,
                              int batch, int m, double* omega, long long sO_b, int ldo,
                              double* Pf, double* Qf, long long* iters, int* status, int* counters,
                              double* ws, size_t lds_limit, const Caps& caps)
{
    const int mp = pad16(m);
    DualArgs a;
    a.caps = caps;
    a.Cr = Cr; a.Ci = Ci; a.sC_b = sC_b; a.ldc = ldc; a.m = m; a.mp = mp;
    a.omega = omega; a.sO_b = sO_b; a.ldo = ldo; a.Pf = Pf; a.Qf = Qf;
    a.iters = iters; a.status = status; a.counters = counters; a.ws = ws; a.v_in_lds = 0;
    switch (mp / 16) {
        case 1: return launch_nt<1>(stream, a, batch, lds_limit);
        case 2: return launch_nt<2>(stream, a, batch, lds_limit);
        case 3: return launch_nt<3>(stream, a, batch, lds_limit);
        case 4: return launch_nt<4>(stream, a, batch, lds_limit);
        case 5: return launch_nt<5>(stream, a, batch, lds_limit);
        case 6: return launch_nt<6>(stream, a, batch, lds_limit);
        case 7: return launch_nt<7>(stream, a, batch, lds_limit);
        case 8: return launch_nt<8>(stream, a, batch, lds_limit);
        default: return LAUNCH_UNSUPPORTED;
    }
}

hipError_t launch_dual(hipStream_t stream, const double* Cr, const double* Ci, long long sC_b, int ldc,
                       int batch, int m, double* omega, long long sO_b, int ldo,
                       double* Pf, double* Qf, long long* iters, int* status, int* counters,
                       double* ws, size_t lds_limit, const Caps& caps)
{
    hipError_t err = launch_eigh(stream, Cr, sC_b, ldc, batch, m, ws, lds_limit, caps.eigh_whole_device != 0);
    if (err != hipSuccess) return err;
    return launch_dual_newton(stream, Cr, Ci, sC_b, ldc, batch, m, omega, sO_b, ldo, Pf, Qf,
                              iters, status, counters, ws, lds_limit, caps);
}

size_t dual_lds_bytes(int m) {
    switch (pad16(m) / 16) {
        case 1: return (size_t)Cfg<1>::LDS_END * 8;
        case 2: return (size_t)Cfg<2>::LDS_END * 8;
        case 3: return (size_t)Cfg<3>::LDS_END * 8;
        case 4: return (size_t)Cfg<4>::LDS_END * 8;
        case 5: return (size_t)Cfg<5>::LDS_END * 8;
        case 6: return (size_t)Cfg<6>::LDS_END * 8;
        case 7: return (size_t)Cfg<7>::LDS_END * 8;
        case 8: return (size_t)Cfg<8>::LDS_END * 8;
        default: return 0;
    }
}

}  // namespace ssn
