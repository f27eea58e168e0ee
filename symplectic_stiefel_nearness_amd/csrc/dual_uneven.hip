// Newton stage of the uneven-batch plan: one workgroup per target, each running the engine
// body of its own order (a plan of one order launches that order's body alone). No reference counterpart (the reference solves one shape per call,
// src/symplectic_projection/dual_recommended_hvp.py:231-242).
#include "engine.h"

namespace ssn {

hipError_t launch_eigh_uneven(hipStream_t stream, const UTarget* tg, int count, size_t eigh_lds);

__device__ __forceinline__ DualArgs uneven_args(const UTarget& t) {
    DualArgs a;
    a.Cr = t.Cr; a.Ci = t.Ci; a.sC_b = 0; a.ldc = t.m; a.m = t.m; a.mp = t.mp;
    a.omega = t.omega; a.sO_b = 0; a.ldo = (int)t.ldo; a.Pf = t.Pf; a.Qf = t.Qf;
    a.iters = t.iters; a.status = t.status; a.counters = t.counters; a.ws = t.dual; a.v_in_lds = 0;
    a.caps = Caps();
    return a;
}

// Plans of mixed orders: every target runs the body of its own order.
__global__ __launch_bounds__(DT)
void dual_uneven_kernel(const UTarget* __restrict__ tg)
{
    const UTarget& t = tg[blockIdx.x];
    const DualArgs a = uneven_args(t);
    switch (t.mp / 16) {
        case 1: dual_body<1>(a, 0); break;
        case 2: dual_body<2>(a, 0); break;
        case 3: dual_body<3>(a, 0); break;
        case 4: dual_body<4>(a, 0); break;
        case 5: dual_body<5>(a, 0); break;
        case 6: dual_body<6>(a, 0); break;
        case 7: dual_body<7>(a, 0); break;
        case 8: dual_body<8>(a, 0); break;
        default: break;
    }
}

// Plans of one order NT: the body of that order alone, so the kernel's register allocation is
// that of the homogeneous launch (the all-orders kernel spills more).
template <int NT>
__global__ __launch_bounds__(DT)
void dual_uneven_order_kernel(const UTarget* __restrict__ tg)
{
    dual_body<NT>(uneven_args(tg[blockIdx.x]), 0);
}

template <int NT>
static hipError_t launch_order(hipStream_t stream, const UTarget* tg, int count, size_t dual_lds) {
    hipError_t err = hipFuncSetAttribute((const void*)dual_uneven_order_kernel<NT>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)dual_lds);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(dual_uneven_order_kernel<NT>, dim3(count), dim3(DT), dual_lds, stream, tg);
    return hipGetLastError();
}

// single_nt: the tile order shared by every target of the plan, or 0 when the orders differ.
hipError_t launch_dual_uneven(hipStream_t stream, const UTarget* tg, int count, size_t eigh_lds, size_t dual_lds,
                              int single_nt)
{
    hipError_t err = launch_eigh_uneven(stream, tg, count, eigh_lds);
    if (err != hipSuccess) return err;
    switch (single_nt) {
        case 1: return launch_order<1>(stream, tg, count, dual_lds);
        case 2: return launch_order<2>(stream, tg, count, dual_lds);
        case 3: return launch_order<3>(stream, tg, count, dual_lds);
        case 4: return launch_order<4>(stream, tg, count, dual_lds);
        case 5: return launch_order<5>(stream, tg, count, dual_lds);
        case 6: return launch_order<6>(stream, tg, count, dual_lds);
        case 7: return launch_order<7>(stream, tg, count, dual_lds);
        case 8: return launch_order<8>(stream, tg, count, dual_lds);
        default: break;
    }
    err = hipFuncSetAttribute((const void*)dual_uneven_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)dual_lds);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(dual_uneven_kernel, dim3(count), dim3(DT), dual_lds, stream, tg);
    return hipGetLastError();
}

}  // namespace ssn
