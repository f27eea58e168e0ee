// Adjoint (backward) kernel of the differentiable projection; no reference counterpart
// (the reference is forward-only).
#include "engine.h"

namespace ssn {

// ---------------------------------------------------------------- adjoint (backward)
// Implicit differentiation of X(T) = B N, N = (I + i omega)^-1, through the dual optimality
// condition (no reference counterpart; the reference is forward-only). Per target, from the
// forward workspace (eigenbasis V, lambda, scale, and the final P, Q, Sr, Si in the eigenbasis)
// and the cross Gram M = B^H G_c (slots S_G = Re, S_D = Im):
//     g      = skew(-Im(N M N))                          (dL/d omega, original basis)
//     lam    = H^-1 g   by PCG in the eigenbasis with the solver's HVP and preconditioner;
//              the solver's H is that of phi / scale, so H_s x = (V^T g V) / scale
//     K      = N lam N  ->  Kr (slot S_R), Ki (slot S_Z0), original basis
// The streaming kernel then forms dL/dB = G_c N + 2i B K.
struct AdjArgs {
    double* ws;                    // solver workspace, one target_doubles(mp) block per target
    const double* Pf; const double* Qf;   // [b][mp][mp] N = Pf + i Qf, original basis
    int m, mp;
    int* counters;                 // [b][4]: CG iterations, status (0 ok, 1 not converged), ticks, 0
    double tol;                    // relative residual of the Hessian solve
};

template <int NT>
__global__ __launch_bounds__(DT)
void adjoint_kernel(AdjArgs a)
{
    using Cf = Cfg<NT>;
    constexpr int MP = Cf::MP;
    constexpr long long MM = (long long)MP * MP;
    extern __shared__ double lds[];
    Ctx c;
    c.tid = threadIdx.x; c.wave = __builtin_amdgcn_readfirstlane(c.tid >> 6); c.lane = c.tid & 63;
    c.r = c.lane & 15; c.q = c.lane >> 4; c.m = a.m; c.red = lds; c.phase = 0;
    const int b = blockIdx.x;
    const unsigned long long t_start = wall_clock64();
    double* ws = a.ws + (long long)b * target_doubles(MP);
    auto slot = [&](int s) { return ws + s * MM; };
    const double* lam = ws + S_COUNT * MM;
    const double scale = lam[MP];
    double* lam_s = lds + Cf::OFF_LAM;
    for (int i = c.tid; i < MP; i += DT) lam_s[i] = lam[i];
    const double* V = slot(S_V);
    const double* Pe = slot(S_P);  const double* Qe = slot(S_Q);
    const double* Sr = slot(S_SR); const double* Si = slot(S_SI);
    const double* Mr = slot(S_G);  const double* Mi = slot(S_D);
    double* X = slot(S_W);   double* R = slot(S_R);   double* Kr = slot(S_R);  double* Ki = slot(S_Z0);
    double* Pp = slot(S_PP); double* Hp = slot(S_HP); double* U = slot(S_U);   double* Vv = slot(S_VV);
    double* T1 = slot(S_T1); double* T2 = slot(S_T2); double* T3 = slot(S_T3);
    double* T4 = slot(S_T4); double* T5 = slot(S_T5);
    const double* Pf = a.Pf + (long long)b * MM;
    const double* Qf = a.Qf + (long long)b * MM;
    sync();

    // Y = M N: Yr = Mr Pf - Mi Qf (T1), Yi = Mr Qf + Mi Pf (T2); Im W = Pf Yi + Qf Yr (T3)
    gemm2<NT>(c, lds, Mr, Pf, Mi, Qf, -1.0, [&](int i, int j, double v) { T1[i * MP + j] = v; });
    gemm2<NT>(c, lds, Mr, Qf, Mi, Pf, 1.0, [&](int i, int j, double v) { T2[i * MP + j] = v; });
    gemm2<NT>(c, lds, Pf, T2, Qf, T1, 1.0, [&](int i, int j, double v) { T3[i * MP + j] = v; });
    each<NT>(c, [&](int i, int j, int e) { T4[e] = (T3[j * MP + i] - T3[e]) / 2; });   // g = skew(-Im W)
    // rhs = skew(V^T g V) / scale
    gemm<NT, false, false>(c, lds, T4, V, [&](int i, int j, double v) { T5[i * MP + j] = v; });
    gemm<NT, true, false>(c, lds, V, T5, [&](int i, int j, double v) { T1[i * MP + j] = v; });
    double rz = 0.0, rr0 = 0.0;
    each<NT>(c, [&](int i, int j, int e) {
        const double rv = ((T1[e] - T1[j * MP + i]) / 2) / scale;
        const double z = rv / (lam_s[i] + lam_s[j]);
        R[e] = rv; X[e] = 0.0; Pp[e] = z;
        rz = fma(rv, z, rz);
        rr0 = fma(rv, rv, rr0);
    });
    wg_sum2(c, rz, rr0);
    const double target = a.tol * sqrt(rr0);
    int it = 0, status = 1;
    if (!(sqrt(rr0) > 0.0)) status = 0;
#pragma unroll 1
    for (; !ChainCfg<NT>::OK && it < MAX_CG && status; ++it) {
        gemm<NT, false, false>(c, lds, Pp, Sr, [&](int i, int j, double v) { U[i * MP + j] = v; });
        gemm<NT, false, false>(c, lds, Pp, Si, [&](int i, int j, double v) { Vv[i * MP + j] = v; });
        gemm2<NT>(c, lds, Pe, U, Qe, Vv, -1.0, [&](int i, int j, double v) { T2[i * MP + j] = v; });
        double curv = 0.0;
        each<NT>(c, [&](int i, int j, int e) {
            const double h = (i == j) ? 0.0 : T2[e] - T2[j * MP + i];
            Hp[e] = h;
            curv = fma(Pp[e], h, curv);
        });
        const double curvature = wg_sum(c, curv);
        if (!(curvature > 0.0)) break;
        const double alpha = rz / curvature;
        double rr = 0.0, rzn = 0.0;
        each<NT>(c, [&](int i, int j, int e) {
            X[e] = X[e] + alpha * Pp[e];
            const double rv = R[e] - alpha * Hp[e];
            R[e] = rv;
            const double z = rv / (lam_s[i] + lam_s[j]);
            T3[e] = z;
            rr = fma(rv, rv, rr);
            rzn = fma(rv, z, rzn);
        });
        wg_sum2(c, rr, rzn);
        if (sqrt(rr) <= target) { status = 0; ++it; break; }
        const double beta = rzn / rz;
        each<NT>(c, [&](int, int, int e) { Pp[e] = T3[e] + beta * Pp[e]; });
        rz = rzn;
        sync();
    }
    if constexpr (ChainCfg<NT>::OK) {
        int how;
        const int nh = cg_chained<NT, true>(c, lds, [&](int i, int j) { return ((T1[i * MP + j] - T1[j * MP + i]) / 2) / scale; },
                                            Sr, Si, Pe, slot(S_QT), lam_s, a.tol, X, nullptr, how,
                                            MAX_CG, R, Pp, Hp, U);
        it = how == 0 ? nh - 1 : nh;
        status = how == 1 ? 0 : 1;
    }
    // lam = skew(V X V^T) (T3), K = N lam N: Zr = lam Pf (T4), Zi = lam Qf (T5),
    // Kr = Pf Zr - Qf Zi, Ki = Pf Zi + Qf Zr
    gemm<NT, false, true>(c, lds, X, V, [&](int i, int j, double v) { T1[i * MP + j] = v; });
    gemm<NT, false, false>(c, lds, V, T1, [&](int i, int j, double v) { T2[i * MP + j] = v; });
    each<NT>(c, [&](int i, int j, int e) { T3[e] = (T2[e] - T2[j * MP + i]) / 2; });
    gemm<NT, false, false>(c, lds, T3, Pf, [&](int i, int j, double v) { T4[i * MP + j] = v; });
    gemm<NT, false, false>(c, lds, T3, Qf, [&](int i, int j, double v) { T5[i * MP + j] = v; });
    gemm2<NT>(c, lds, Pf, T4, Qf, T5, -1.0, [&](int i, int j, double v) { Kr[i * MP + j] = v; });
    gemm2<NT>(c, lds, Pf, T5, Qf, T4, 1.0, [&](int i, int j, double v) { Ki[i * MP + j] = v; });
    if (c.tid == 0 && a.counters) {
        int* cnt = a.counters + 4 * b;
        cnt[0] = it; cnt[1] = status; cnt[2] = (int)(wall_clock64() - t_start); cnt[3] = 0;
    }
}

template <int NT>
static hipError_t launch_adj_nt(hipStream_t stream, AdjArgs a, int batch, size_t lds_limit) {
    const size_t lds = (size_t)Cfg<NT>::LDS_END * 8;
    if (lds > lds_limit) return LAUNCH_UNSUPPORTED;
    hipError_t err = hipFuncSetAttribute((const void*)adjoint_kernel<NT>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(adjoint_kernel<NT>, dim3(batch), dim3(DT), lds, stream, a);
    return hipGetLastError();
}

hipError_t launch_adjoint(hipStream_t stream, double* ws, const double* Pf, const double* Qf, int batch, int m,
                          int* counters, double tol, size_t lds_limit)
{
    const int mp = pad16(m);
    AdjArgs a;
    a.ws = ws; a.Pf = Pf; a.Qf = Qf; a.m = m; a.mp = mp; a.counters = counters; a.tol = tol;
    switch (mp / 16) {
        case 1: return launch_adj_nt<1>(stream, a, batch, lds_limit);
        case 2: return launch_adj_nt<2>(stream, a, batch, lds_limit);
        case 3: return launch_adj_nt<3>(stream, a, batch, lds_limit);
        case 4: return launch_adj_nt<4>(stream, a, batch, lds_limit);
        case 5: return launch_adj_nt<5>(stream, a, batch, lds_limit);
        case 6: return launch_adj_nt<6>(stream, a, batch, lds_limit);
        case 7: return launch_adj_nt<7>(stream, a, batch, lds_limit);
        case 8: return launch_adj_nt<8>(stream, a, batch, lds_limit);
        default: return LAUNCH_UNSUPPORTED;
    }
}

}  // namespace ssn
