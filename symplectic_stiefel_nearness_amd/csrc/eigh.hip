// Eigensolver launches of the Dual solver: the per-target rotation-only Jacobi kernel
// (eigh_kernel, in engine.h) and the eigenvector replay over every CU; the same two
// steps for the uneven-batch plan. Replaces torch.linalg.eigh at
// src/symplectic_projection/dual_recommended_hvp.py:108.
#include "engine.h"
#include "eigh_tri.h"
#include <cstdlib>

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

// ---------------------------------------------------------------- eigensolver and replay in one launch
// The eigensolver workgroups (blockIdx.x < batch) publish the rotation record as they go, and
// the replay workgroups (RF_ROWS rows of one target each) follow it, so that V is ready a chunk
// of rounds after the last rotation instead of one replay launch later. Every workgroup takes
// more than half of a CU's LDS, so each has a CU to itself, and the launch is cooperative, so
// all of them are resident at once: every wait of a replay workgroup is on an eigensolver
// workgroup that is running (and that never waits on the replay).
// The replay's staging buffer lives in the dynamic LDS (which the replay workgroups do not
// otherwise use), not in static LDS: static LDS would sit below the eigensolver's area and
// move its base, and the eigensolver measured 3x slower with its area 32 KB higher.
template <int JB>
__global__ __launch_bounds__(ET)
void eigh_coop_kernel(EighArgs a, int batch, int rpt, long long spin_limit)
{
    if ((int)blockIdx.x < batch) {
        eigh_body<JB, false, true, true>(a, blockIdx.x);
    } else {
        extern __shared__ double dyn_lds[];
        double2 (*par)[RP_CHUNK * 64] = reinterpret_cast<double2 (*)[RP_CHUNK * 64]>(dyn_lds);
        unsigned long long* shared_p = reinterpret_cast<unsigned long long*>(dyn_lds + 2 * 2 * RP_CHUNK * 64);
        const int w = (int)blockIdx.x - batch, b = w / rpt, rb = w - b * rpt;
        eigh_replay_follow(a.ws + (long long)b * target_doubles(a.mp), a.m, a.mp, rb, spin_limit, par, shared_p);
    }
}

// each target's progress word and replay-failure flag, cleared before the cooperative launch
__global__ __launch_bounds__(256)
void eigh_progress_clear_kernel(double* ws, int mp, int batch)
{
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= batch) return;
    double* extra = ws + (long long)b * target_doubles(mp) + (long long)S_COUNT * mp * mp + mp;
    reinterpret_cast<unsigned long long*>(extra)[EX_PROGRESS] = 0ull;
    extra[EX_REPLAY_FAILED] = 0.0;
}

// over half of a CU's LDS (one workgroup per CU), and room for the replay's staging buffer
constexpr size_t COOP_MIN_DYN_LDS = 96 * 1024;

// Polls a replay workgroup makes before it gives up on its eigensolver (seconds of waiting;
// SSN_REPLAY_SPIN_LIMIT overrides it, e.g. 0 to exercise the failure path).
static long long replay_spin_limit() {
    const char* e = getenv("SSN_REPLAY_SPIN_LIMIT");
    return e ? atoll(e) : (1LL << 22);
}

// CUs and cooperative-launch support of the current device, looked up once per device
static bool coop_grid_fits(int wgs) {
    constexpr int MAXDEV = 64;
    static int cus[MAXDEV] = {}, coop[MAXDEV] = {}, known[MAXDEV] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAXDEV) return false;
    if (!known[dev]) {
        int c = 0, k = 0;
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) c = 0;
        if (hipDeviceGetAttribute(&k, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess) k = 0;
        cus[dev] = c; coop[dev] = k; known[dev] = 1;
    }
    return coop[dev] && wgs <= cus[dev];
}

// The one-launch form when it fits (one workgroup per CU for every eigensolver and replay
// workgroup); returns hipErrorCooperativeLaunchTooLarge when it does not, so that the caller
// takes the two-launch form. SSN_EIGH_COOP=0 disables it.
template <int JB>
static hipError_t launch_eigh_coop(hipStream_t stream, const EighArgs& e, int batch, size_t lds, size_t lds_limit) {
    const char* env = getenv("SSN_EIGH_COOP");
    if (env && atoi(env) == 0) return hipErrorCooperativeLaunchTooLarge;
    const int rpt = (e.mp + RF_ROWS - 1) / RF_ROWS;
    const int grid = batch * (1 + rpt);
    const size_t dyn = lds > COOP_MIN_DYN_LDS ? lds : COOP_MIN_DYN_LDS;
    static_assert(sizeof(double2) * 2 * RP_CHUNK * 64 + sizeof(unsigned long long) <= COOP_MIN_DYN_LDS, "staging");
    if (dyn > lds_limit || !coop_grid_fits(grid)) return hipErrorCooperativeLaunchTooLarge;
    const void* fn = (const void*)eigh_coop_kernel<JB>;
    hipError_t err = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(eigh_progress_clear_kernel, dim3((batch + 255) / 256), dim3(256), 0, stream, e.ws, e.mp, batch);
    err = hipGetLastError();
    if (err != hipSuccess) return err;
    EighArgs a = e;
    int nb = batch, r = rpt;
    long long spin = replay_spin_limit();
    void* args[] = {&a, &nb, &r, &spin};
    err = hipLaunchCooperativeKernel(fn, dim3(grid), dim3(ET), args, (unsigned)dyn, stream);
    if (err != hipSuccess) (void)hipGetLastError();      // the two-launch form follows
    return err;
}

// ---------------------------------------------------------------- tridiagonal eigensolver
// Every order (padded order <= 128): eigh_tri_kernel, one workgroup per target, with the Jacobi
// eigensolver as its per-target fallback. LAUNCH_DECLINED when the LDS does not fit the device,
// or when SSN_EIGH_TRI=0 selects the Jacobi forms (diagnostics).
static bool tri_enabled() {
    const char* env = getenv("SSN_EIGH_TRI");
    return !(env && atoi(env) == 0);
}

template <int NT>
static hipError_t launch_tri_nt(hipStream_t stream, const EighArgs& e, int batch, size_t lds_limit) {
    const size_t lds = (size_t)tri_lds_doubles<NT>(e.m) * 8;
    if (lds > lds_limit) return LAUNCH_DECLINED;
    hipError_t err = hipFuncSetAttribute((const void*)eigh_tri_kernel<NT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)lds);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(eigh_tri_kernel<NT>, dim3(batch), dim3(ET), lds, stream, e);
    return hipGetLastError();
}

static hipError_t launch_eigh_tri(hipStream_t stream, const double* Cr, long long sC_b, int ldc, int batch, int m,
                                  double* ws, size_t lds_limit) {
    if (!tri_enabled() || m < 2 || batch <= 0) return LAUNCH_DECLINED;
    EighArgs e;
    e.Cr = Cr; e.sC_b = sC_b; e.ldc = ldc; e.m = m; e.mp = pad16(m); e.ws = ws; e.v_in_lds = e.mp <= 96 ? 1 : 0;
    switch (e.mp / 16) {
        case 1: return launch_tri_nt<1>(stream, e, batch, lds_limit);
        case 2: return launch_tri_nt<2>(stream, e, batch, lds_limit);
        case 3: return launch_tri_nt<3>(stream, e, batch, lds_limit);
        case 4: return launch_tri_nt<4>(stream, e, batch, lds_limit);
        case 5: return launch_tri_nt<5>(stream, e, batch, lds_limit);
        case 6: return launch_tri_nt<6>(stream, e, batch, lds_limit);
        case 7: return launch_tri_nt<7>(stream, e, batch, lds_limit);
        case 8: return launch_tri_nt<8>(stream, e, batch, lds_limit);
        default: return LAUNCH_DECLINED;
    }
}

// whole_device: the one-launch form (eigensolver and replay together, every CU for the length
// of the eigensolver) may be used; otherwise, or when it does not fit, the two launches.
hipError_t launch_eigh(hipStream_t stream, const double* Cr, long long sC_b, int ldc, int batch,
                              int m, double* ws, size_t lds_limit, bool whole_device) {
    {
        const hipError_t te = launch_eigh_tri(stream, Cr, sC_b, ldc, batch, m, ws, lds_limit);
        if (te != LAUNCH_DECLINED) return te;
    }
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
        if (batch > 0 && whole_device) {
            const hipError_t ce = small ? launch_eigh_coop<JB_BATCH>(stream, e, batch, lds, lds_limit)
                                        : launch_eigh_coop<JB_MAX>(stream, e, batch, lds, lds_limit);
            if (ce == hipSuccess) return hipSuccess;
            if (ce != hipErrorCooperativeLaunchTooLarge) return ce;
        }
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
    return launch_eigh(stream, Cr, sC_b, ldc, batch, m, ws, lds_limit, true);
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

// Uneven plans with the tridiagonal eigensolver: one workgroup per target, each running the
// eigensolver of its own order (its Jacobi fallback included), V formed in the same launch.
__global__ __launch_bounds__(ET)
void eigh_tri_uneven_kernel(const UTarget* __restrict__ tg)
{
    const UTarget& t = tg[blockIdx.x];
    EighArgs a;
    a.Cr = t.Cr; a.sC_b = 0; a.ldc = t.m; a.m = t.m; a.mp = t.mp; a.ws = t.dual; a.v_in_lds = t.mp <= 96 ? 1 : 0;
    switch (t.mp / 16) {
        case 1: eigh_tri_target<1>(a, 0); break;
        case 2: eigh_tri_target<2>(a, 0); break;
        case 3: eigh_tri_target<3>(a, 0); break;
        case 4: eigh_tri_target<4>(a, 0); break;
        case 5: eigh_tri_target<5>(a, 0); break;
        case 6: eigh_tri_target<6>(a, 0); break;
        case 7: eigh_tri_target<7>(a, 0); break;
        case 8: eigh_tri_target<8>(a, 0); break;
        default: break;
    }
}

// dynamic LDS bytes the tridiagonal eigensolver of order m needs (its fallback included)
size_t eigh_tri_lds_bytes(int m) {
    switch (pad16(m) / 16) {
        case 1: return (size_t)tri_lds_doubles<1>(m) * 8;
        case 2: return (size_t)tri_lds_doubles<2>(m) * 8;
        case 3: return (size_t)tri_lds_doubles<3>(m) * 8;
        case 4: return (size_t)tri_lds_doubles<4>(m) * 8;
        case 5: return (size_t)tri_lds_doubles<5>(m) * 8;
        case 6: return (size_t)tri_lds_doubles<6>(m) * 8;
        case 7: return (size_t)tri_lds_doubles<7>(m) * 8;
        case 8: return (size_t)tri_lds_doubles<8>(m) * 8;
        default: return 0;
    }
}

// eigh_lds: the plan's largest need over both eigensolver forms (eigh_variant, eigh_tri_lds_bytes)
hipError_t launch_eigh_uneven(hipStream_t stream, const UTarget* tg, int count, size_t eigh_lds)
{
    if (tri_enabled()) {
        hipError_t err = hipFuncSetAttribute((const void*)eigh_tri_uneven_kernel,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)eigh_lds);
        if (err != hipSuccess) return err;
        hipLaunchKernelGGL(eigh_tri_uneven_kernel, dim3(count), dim3(ET), eigh_lds, stream, tg);
        return hipGetLastError();
    }
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
