// Newton / PCG / Armijo launches of the Dual solver (dual_body, in engine.h): one 8-wave
// workgroup per target, between the changes of basis done on every CU. Replaces dual_batch
// (src/symplectic_projection/dual_recommended_hvp.py:99-191).
#include "engine.h"

namespace ssn {

hipError_t launch_eigh(hipStream_t stream, const double* Cr, long long sC_b, int ldc, int batch,
                       int m, double* ws, size_t lds_limit, bool whole_device);

template <int NT>
__global__ __launch_bounds__(DT)
void dual_kernel(DualArgs a)
{
    dual_body<NT, true>(a, blockIdx.x);
}

// ---------------------------------------------------------------- change of basis, every CU
// The products before and after the iteration (dual_recommended_hvp.py:110-111 and :58-64) as
// chains Out[:, J] = L (R[:, J]) over 16-column blocks J, one workgroup per (target, block,
// chain) over all targets, so that they take the whole device instead of each target's one CU:
// wave I forms tile I of R[:, J] into LDS, then tile I of Out[:, J] from it. Every tile is the
// MFMA chain of the engine's quadrant products (k ascending, the same operand values up to exact
// sign changes of skew or symmetric operands), so the results are bitwise those of the
// in-engine form.
//   STAGE 0 (chains 0, 1): T4 = V^T ((Ci / scale) V), T1 = V^T J(V) (one product: R = J(V) is formed
//                          directly; J(V) = [V_bot; -V_top] over the leading m rows)
//   STAGE 1 (chains 0-2):  T4 = V (W V^T), Pf = V (P V^T), Qf = V (Q V^T)
// Workgroup x = b + batch u: the workgroups of target b are batch apart, so with a batch that is
// a multiple of 8 they share an XCD (workgroups go to the XCDs round-robin) and its L2.
template <int NT, int STAGE>
__global__ __launch_bounds__(DT)
void dual_basis_kernel(DualArgs a, int batch)
{
    constexpr int MP = 16 * NT, K4 = MP / 4;
    constexpr long long MM = (long long)MP * MP;
    __shared__ double Rl[MP * 17];                    // R[:, J]: MP x 16, row stride 17
    const int b = blockIdx.x % batch, u = blockIdx.x / batch;
    const int ch = u / NT, J = u - ch * NT;           // chain, column block
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
    const int m = a.m, k2 = m / 2;
    double* ws = a.ws + (long long)b * target_doubles(MP);
    auto slot = [&](int sl) { return ws + sl * MM; };
    const double* V = slot(S_V);
    const double scale = ws[S_COUNT * MM + MP];      // extra[0]
    const double* Cib = a.Ci + (long long)b * a.sC_b;
    const int col = 16 * J + r;
    const bool direct = STAGE == 0 && ch == 1;        // T1 = V^T J(V): R = J(V) needs no product
    // first product: tile I = wave of R[:, J]
    if (wave < NT) {
        const int I = wave, row = 16 * I + r;
        d4 acc = zero4();
        if (direct) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int rr = 16 * I + q + 4 * i;
                acc[i] = rr < k2 ? V[(rr + k2) * MP + col] : (rr < m ? -V[(rr - k2) * MP + col] : 0.0);
            }
        } else {
            // A read transposed (coalesced), the sign on B: Ci / scale and W exactly skew, P
            // symmetric, Q through Q^T
            const double* At = STAGE == 0 ? nullptr : slot(ch == 0 ? S_W : (ch == 1 ? S_P : S_QT));
            const double bs = (STAGE == 0 || ch == 0) ? -1.0 : 1.0;
            double xa[K4], yb[K4];
#pragma unroll
            for (int kb = 0; kb < K4; ++kb) {
                const int k = 4 * kb + q;
                if constexpr (STAGE == 0) {
                    xa[kb] = (row < m && k < m) ? Cib[(long long)k * a.ldc + row] / scale : 0.0;
                    yb[kb] = bs * V[k * MP + col];
                } else {
                    xa[kb] = At[k * MP + row];
                    yb[kb] = bs * V[col * MP + k];       // V^T[k][col]
                }
            }
#pragma unroll
            for (int kb = 0; kb < K4; ++kb) acc = mfma(xa[kb], yb[kb], acc);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) Rl[(16 * I + q + 4 * i) * 17 + r] = acc[i];
    }
    __syncthreads();
    // second product: tile I of Out[:, J] = L R[:, J], L = V^T (stage 0) or V (stage 1)
    if (wave < NT) {
        const int I = wave, row = 16 * I + r;
        double xa[K4], yb[K4];
#pragma unroll
        for (int kb = 0; kb < K4; ++kb) {
            const int k = 4 * kb + q;
            xa[kb] = STAGE == 0 ? V[k * MP + row] : V[row * MP + k];
            yb[kb] = Rl[k * 17 + r];
        }
        d4 acc = zero4();
#pragma unroll
        for (int kb = 0; kb < K4; ++kb) acc = mfma(xa[kb], yb[kb], acc);
        double* out;
        if constexpr (STAGE == 0) out = slot(ch == 0 ? S_T4 : S_T1);
        else out = ch == 0 ? slot(S_T4) : (ch == 1 ? a.Pf : a.Qf) + (long long)b * MM;
#pragma unroll
        for (int i = 0; i < 4; ++i) out[(16 * I + q + 4 * i) * MP + col] = acc[i];
    }
}

// Cie = skew(T4), Jse = skew(T1) / scale, and the starting W = 0, P = I, Q = Q^T = 0
// (dual_recommended_hvp.py:110-116); workgroup x = b + batch j covers elements DT j ... of target b.
template <int NT>
__global__ __launch_bounds__(DT)
void dual_setup_finish_kernel(DualArgs a, int batch)
{
    constexpr int MP = 16 * NT;
    constexpr long long MM = (long long)MP * MP;
    const int b = blockIdx.x % batch, e = (blockIdx.x / batch) * DT + threadIdx.x;
    if (e >= MM) return;
    double* ws = a.ws + (long long)b * target_doubles(MP);
    auto slot = [&](int sl) { return ws + sl * MM; };
    const double scale = ws[S_COUNT * MM + MP];
    const int i = e / MP, j = e - (e / MP) * MP, m = a.m;
    const double* T4 = slot(S_T4);
    const double* T1 = slot(S_T1);
    slot(S_CIE)[e] = (T4[e] - T4[j * MP + i]) / 2;
    slot(S_JSE)[e] = ((T1[e] - T1[j * MP + i]) / 2) / scale;
    slot(S_W)[e] = 0.0;
    slot(S_P)[e] = (i == j && i < m) ? 1.0 : 0.0;
    slot(S_Q)[e] = 0.0;
    slot(S_QT)[e] = 0.0;
}

// omega = skew(T4) in the original basis (dual_recommended_hvp.py:58-64); workgroup x = b + batch j.
template <int NT>
__global__ __launch_bounds__(DT)
void dual_omega_kernel(DualArgs a, int batch)
{
    constexpr int MP = 16 * NT;
    constexpr long long MM = (long long)MP * MP;
    const int b = blockIdx.x % batch, e = (blockIdx.x / batch) * DT + threadIdx.x, m = a.m;
    if (e >= m * m) return;
    const double* T4 = a.ws + (long long)b * target_doubles(MP) + S_T4 * MM;
    const int i = e / m, j = e - (e / m) * m;
    a.omega[(long long)b * a.sO_b + (long long)i * a.ldo + j] = (T4[i * MP + j] - T4[j * MP + i]) / 2;
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
    constexpr int MP = 16 * NT;
    const unsigned ew = (unsigned)((MP * MP + DT - 1) / DT);
    // into the eigenbasis, on every CU
    hipLaunchKernelGGL((dual_basis_kernel<NT, 0>), dim3(2 * NT * batch), dim3(DT), 0, stream, a, batch);
    hipLaunchKernelGGL(dual_setup_finish_kernel<NT>, dim3(ew * batch), dim3(DT), 0, stream, a, batch);
    // the iteration, one workgroup per target
    hipLaunchKernelGGL(dual_kernel<NT>, dim3(batch), dim3(DT), lds, stream, a);
    // back to the original basis, on every CU
    hipLaunchKernelGGL((dual_basis_kernel<NT, 1>), dim3(3 * NT * batch), dim3(DT), 0, stream, a, batch);
    const unsigned om = (unsigned)((a.m * a.m + DT - 1) / DT);
    hipLaunchKernelGGL(dual_omega_kernel<NT>, dim3(om * batch), dim3(DT), 0, stream, a, batch);
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
