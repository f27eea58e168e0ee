// Per-target dual engine: the whole Newton / PCG / Armijo iteration of the Dual
// solver runs on device, one workgroup (8 waves) per target, with no host
// synchronisation. Replaces dual_batch (src/symplectic_projection/dual_recommended_hvp.py:99-191)
// and the fused kernels it launches (hvp_fused.cu:1-420, cholesky_inverse_kernel.cu:1-103),
// plus torch.linalg.eigh / cholesky_ex / cholesky_solve on that path.
//
// The batched reference masks finished targets; per target, its trajectory is exactly
// the single-target trajectory (a masked target only ever receives zero updates), so
// each workgroup runs its own loop and reports a per-target status.
//
// Small m x m matrices live in a per-target global workspace (L2 / Infinity-Cache
// resident), padded to mp = 16 ceil(m/16) with zero padding so every product runs as
// whole v_mfma_f64_16x16x4f64 tiles. LDS holds what is latency-bound: the Jacobi
// eigensolver's matrix (and its eigenvectors when they fit), the packed Cholesky
// factor and its inverse, and reduction scratch. All reductions are deterministic.
#include "common.h"

// element-wise updates mirror the reference's explicitly rounded (non-contracted) ops
#pragma clang fp contract(off)

namespace ssn {

constexpr int DW = 8;                 // waves per target
constexpr int DT = DW * WAVE;         // threads per target
constexpr int MAX_NEWTON = 100, MAX_CG = 150, MAX_BACKTRACK = 40;   // dual_recommended_hvp.py:16
constexpr double FEAS_TOL = 1e-8;                                    // dual_recommended_hvp.py:15
constexpr int MAX_SWEEPS = 40;

// workspace slots, each mp*mp doubles
enum Slot {
    S_V = 0, S_CIE, S_JSE, S_W, S_P, S_Q, S_SR, S_SI, S_G, S_D, S_R, S_Z0, S_PP, S_HP,
    S_U, S_VV, S_T1, S_T2, S_T3, S_T4, S_T5, S_COUNT
};

struct DualArgs {
    const double* Cr; const double* Ci; long long sC_b; int ldc;
    int m, mp;
    double* omega; long long sO_b; int ldo;      // m x m output (original basis)
    double* Pf; double* Qf;                       // [b][mp][mp] final P, Q (original basis)
    long long* iters; int* status; int* counters; // counters [b][4]: newton, cg, backtracks, hvps
    double* ws;                                   // [b][S_COUNT][mp*mp] + [b][mp] values
    int v_in_lds;
};

struct Ctx {
    int tid, wave, lane, r, q, m, mp, nt;
    double* red;       // LDS [DW]
    double* ws;        // this target's slots
    __device__ double* slot(int s) const { return ws + (long long)s * mp * mp; }
};

__device__ __forceinline__ void sync() { __syncthreads(); }

// Deterministic workgroup sum; every thread returns the same value.
__device__ double wg_sum(const Ctx& c, double v) {
    v = wave_sum(v);
    if (c.lane == 0) c.red[c.wave] = v;
    sync();
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < DW; ++w) s += c.red[w];
    sync();
    return s;
}

// ---------------------------------------------------------------- MFMA tile products
// acc += op(A) op(B) for one 16x16 tile (I, J), K = mp. A, B row-major mp x mp.
template <bool TA, bool TB>
__device__ __forceinline__ void tile_mac(const double* __restrict__ A, const double* __restrict__ B,
                                         int mp, int I, int J, int r, int q, d4& acc)
{
    for (int k0 = 0; k0 < mp; k0 += 16) {
        double a[4], b[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = k0 + 4 * u + q;
            a[u] = TA ? A[k * mp + I * 16 + r] : A[(I * 16 + r) * mp + k];
            b[u] = TB ? B[(J * 16 + r) * mp + k] : B[k * mp + J * 16 + r];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) acc = mfma(a[u], b[u], acc);
    }
}

__device__ __forceinline__ void tile_mac_neg(const double* __restrict__ A, const double* __restrict__ B,
                                             int mp, int I, int J, int r, int q, d4& acc)
{
    // acc -= A B (negating the A operand is exact)
    for (int k0 = 0; k0 < mp; k0 += 16) {
        double a[4], b[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = k0 + 4 * u + q;
            a[u] = -A[(I * 16 + r) * mp + k];
            b[u] = B[k * mp + J * 16 + r];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) acc = mfma(a[u], b[u], acc);
    }
}

// C = op(A) op(B), with an epilogue f(row, col, value) per element.
template <bool TA, bool TB, class F>
__device__ void gemm(const Ctx& c, const double* A, const double* B, F f) {
    const int nt = c.nt;
    for (int t = c.wave; t < nt * nt; t += DW) {
        const int I = t / nt, J = t - (t / nt) * nt;
        d4 acc = zero4();
        tile_mac<TA, TB>(A, B, c.mp, I, J, c.r, c.q, acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) f(I * 16 + c.q + 4 * i, J * 16 + c.r, acc[i]);
    }
}

// C = A1 B1 - A2 B2 (all non-transposed), epilogue f.
template <class F>
__device__ void gemm_diff(const Ctx& c, const double* A1, const double* B1,
                          const double* A2, const double* B2, F f) {
    const int nt = c.nt;
    for (int t = c.wave; t < nt * nt; t += DW) {
        const int I = t / nt, J = t - (t / nt) * nt;
        d4 acc = zero4();
        tile_mac<false, false>(A1, B1, c.mp, I, J, c.r, c.q, acc);
        tile_mac_neg(A2, B2, c.mp, I, J, c.r, c.q, acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) f(I * 16 + c.q + 4 * i, J * 16 + c.r, acc[i]);
    }
}

// C = A1 B1^T + s A2 B2^T (s = +1 or -1), epilogue f.
template <class F>
__device__ void gemm_nt2(const Ctx& c, const double* A1, const double* B1,
                         const double* A2, const double* B2, bool negate_second, F f) {
    const int nt = c.nt, mp = c.mp;
    for (int t = c.wave; t < nt * nt; t += DW) {
        const int I = t / nt, J = t - (t / nt) * nt;
        d4 acc = zero4();
        tile_mac<false, true>(A1, B1, mp, I, J, c.r, c.q, acc);
        for (int k0 = 0; k0 < mp; k0 += 16) {
            double a[4], b[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int k = k0 + 4 * u + c.q;
                const double av = A2[(I * 16 + c.r) * mp + k];
                a[u] = negate_second ? -av : av;
                b[u] = B2[(J * 16 + c.r) * mp + k];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) acc = mfma(a[u], b[u], acc);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) f(I * 16 + c.q + 4 * i, J * 16 + c.r, acc[i]);
    }
}

// ---------------------------------------------------------------- element-wise helpers
template <class F>
__device__ __forceinline__ void each(const Ctx& c, F f) {
    const int total = c.mp * c.mp;
    for (int e = c.tid; e < total; e += DT) f(e / c.mp, e - (e / c.mp) * c.mp, e);
}

__device__ double dot(const Ctx& c, const double* X, const double* Y) {
    double s = 0.0;
    each(c, [&](int, int, int e) { s = fma(X[e], Y[e], s); });
    return wg_sum(c, s);
}

// skew(R / (lambda_i + lambda_j)) computed at both positions (hvp_fused.cu:374-387).
__device__ void precondition(const Ctx& c, const double* R, const double* lam, double* out, double sign) {
    each(c, [&](int i, int j, int e) {
        const double a = (sign * R[e]) / (lam[i] + lam[j]);
        const double t = (sign * R[j * c.mp + i]) / (lam[j] + lam[i]);
        out[e] = (a - t) * 0.5;
    });
}

// ---------------------------------------------------------------- Jacobi eigensolver
// Cyclic two-sided Jacobi with the round-robin (circle) ordering: each round applies
// m/2 disjoint rotations in parallel. A (m x m, stride mp) is in LDS; V (stride mp) is
// in LDS or global. Rotations are skipped when |a_pq| <= eps sqrt(|a_pp a_qq|), which
// gives eigenvalues with backward error O(eps ||A||). Returns sweeps used (or -1).
__device__ int jacobi(const Ctx& c, double* A, double* V, int* pairs, double* rot, int* flag) {
    const int m = c.m, mp = c.mp, np = m / 2;
    for (int e = c.tid; e < m * mp; e += DT) {
        const int i = e / mp, j = e - (e / mp) * mp;
        V[e] = (i == j) ? 1.0 : 0.0;
    }
    for (int sweep = 0; sweep < MAX_SWEEPS; ++sweep) {
        if (c.tid == 0) *flag = 0;
        sync();
        for (int round = 0; round < m - 1; ++round) {
            for (int s = c.tid; s < np; s += DT) {
                // positions s and m-1-s; position 0 holds player 0, position j>0 holds
                // player ((j - 1 + round) mod (m - 1)) + 1
                const int pa = s == 0 ? 0 : ((s - 1 + round) % (m - 1)) + 1;
                const int pb = ((m - 1 - s - 1 + round) % (m - 1)) + 1;
                const int p = min(pa, pb), qq = max(pa, pb);
                const double app = A[p * mp + p], aqq = A[qq * mp + qq], apq = A[p * mp + qq];
                double cs = 1.0, sn = 0.0, t = 0.0;
                if (apq != 0.0 && fabs(apq) > 2.220446049250313e-16 * sqrt(fabs(app) * fabs(aqq))) {
                    const double theta = (aqq - app) / (2.0 * apq);
                    t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
                    if (theta < 0.0) t = -t;
                    cs = 1.0 / sqrt(t * t + 1.0);
                    sn = t * cs;
                    *flag = 1;
                }
                pairs[2 * s] = p; pairs[2 * s + 1] = qq;
                rot[3 * s] = cs; rot[3 * s + 1] = sn; rot[3 * s + 2] = t;
            }
            sync();
            // 2x2 blocks (P, P') with P <= P'; the owner writes the block and its mirror
            for (int e = c.tid; e < np * np; e += DT) {
                const int P = e / np, P2 = e - (e / np) * np;
                if (P > P2) continue;
                const int p = pairs[2 * P], qq = pairs[2 * P + 1];
                const double cs = rot[3 * P], sn = rot[3 * P + 1];
                if (P == P2) {
                    const double t = rot[3 * P + 2];
                    if (sn != 0.0) {
                        const double apq = A[p * mp + qq];
                        A[p * mp + p] -= t * apq;
                        A[qq * mp + qq] += t * apq;
                        A[p * mp + qq] = 0.0;
                        A[qq * mp + p] = 0.0;
                    }
                    continue;
                }
                const int p2 = pairs[2 * P2], q2 = pairs[2 * P2 + 1];
                const double c2 = rot[3 * P2], s2 = rot[3 * P2 + 1];
                if (sn == 0.0 && s2 == 0.0) continue;
                const double x00 = A[p * mp + p2], x01 = A[p * mp + q2];
                const double x10 = A[qq * mp + p2], x11 = A[qq * mp + q2];
                // right rotation (columns p2, q2), then left rotation (rows p, q)
                const double y00 = c2 * x00 - s2 * x01, y01 = s2 * x00 + c2 * x01;
                const double y10 = c2 * x10 - s2 * x11, y11 = s2 * x10 + c2 * x11;
                const double z00 = cs * y00 - sn * y10, z10 = sn * y00 + cs * y10;
                const double z01 = cs * y01 - sn * y11, z11 = sn * y01 + cs * y11;
                A[p * mp + p2] = z00; A[p * mp + q2] = z01;
                A[qq * mp + p2] = z10; A[qq * mp + q2] = z11;
                A[p2 * mp + p] = z00; A[q2 * mp + p] = z01;
                A[p2 * mp + qq] = z10; A[q2 * mp + qq] = z11;
            }
            for (int e = c.tid; e < m * np; e += DT) {
                const int i = e / np, P = e - (e / np) * np;
                const double sn = rot[3 * P + 1];
                if (sn == 0.0) continue;
                const double cs = rot[3 * P];
                const int p = pairs[2 * P], qq = pairs[2 * P + 1];
                const double vp = V[i * mp + p], vq = V[i * mp + qq];
                V[i * mp + p] = cs * vp - sn * vq;
                V[i * mp + qq] = sn * vp + cs * vq;
            }
            sync();
        }
        const int any = *flag;
        sync();
        if (!any) return sweep + 1;
    }
    return -1;
}

// ---------------------------------------------------------------- Cholesky in LDS (packed)
__device__ __forceinline__ int pk(int i, int j) { return i * (i + 1) / 2 + j; }   // j <= i

// Lower Cholesky of K (global, stride mp, leading m x m block) into packed LDS L.
// Returns false when a pivot is not positive (torch.linalg.cholesky_ex info != 0).
__device__ bool cholesky(const Ctx& c, const double* K, double* L) {
    const int m = c.m, mp = c.mp;
    for (int e = c.tid; e < m * m; e += DT) {
        const int i = e / m, j = e - (e / m) * m;
        if (j <= i) L[pk(i, j)] = K[i * mp + j];
    }
    sync();
    for (int j = 0; j < m; ++j) {
        const double d = L[pk(j, j)];
        if (!(d > 0.0)) { sync(); return false; }
        const double ljj = sqrt(d);
        const double inv = 1.0 / ljj;
        sync();     // every thread has read the pivot before it is overwritten
        for (int i = j + 1 + c.tid; i < m; i += DT) L[pk(i, j)] *= inv;
        if (c.tid == 0) L[pk(j, j)] = ljj;
        sync();
        const int len = m - j - 1;
        for (int e = c.tid; e < len * len; e += DT) {
            const int i = j + 1 + e / len, k = j + 1 + (e - (e / len) * len);
            if (k <= i) L[pk(i, k)] -= L[pk(i, j)] * L[pk(k, j)];
        }
        sync();
    }
    return true;
}

// R = L^{-1} (packed lower, LDS) by right-looking forward substitution, then written to
// global Rg (stride mp, zero elsewhere in the mp x mp block).
__device__ void tri_inverse(const Ctx& c, const double* L, double* Rp, double* Rg) {
    const int m = c.m, mp = c.mp;
    for (int e = c.tid; e < m * (m + 1) / 2; e += DT) Rp[e] = 0.0;
    sync();
    for (int e = c.tid; e < m; e += DT) Rp[pk(e, e)] = 1.0;
    sync();
    // B := I; for i: X[i][0..i] = B[i][0..i] / L[i][i]; B[k][j] -= L[k][i] X[i][j] (k > i, j <= i)
    for (int i = 0; i < m; ++i) {
        const double inv = 1.0 / L[pk(i, i)];
        for (int j = c.tid; j <= i; j += DT) Rp[pk(i, j)] *= inv;
        sync();
        const int rows = m - i - 1, cols = i + 1;
        for (int e = c.tid; e < rows * cols; e += DT) {
            const int k = i + 1 + e / cols, j = e - (e / cols) * cols;
            Rp[pk(k, j)] -= L[pk(k, i)] * Rp[pk(i, j)];
        }
        sync();
    }
    each(c, [&](int i, int j, int e) { Rg[e] = (i < m && j <= i) ? Rp[pk(i, j)] : 0.0; });
    sync();
}

// ---------------------------------------------------------------- the engine
__global__ __launch_bounds__(DT)
void dual_kernel(DualArgs a)
{
    extern __shared__ double lds[];
    Ctx c;
    c.tid = threadIdx.x; c.wave = c.tid >> 6; c.lane = c.tid & 63;
    c.r = c.lane & 15; c.q = c.lane >> 4;
    c.m = a.m; c.mp = a.mp; c.nt = a.mp / 16;
    const int b = blockIdx.x;
    const int m = c.m, mp = c.mp;
    const long long mm = (long long)mp * mp;
    c.red = lds;
    int* flag = reinterpret_cast<int*>(lds + DW);
    double* big = lds + 16;
    c.ws = a.ws + (long long)b * (S_COUNT * mm + mp);
    double* lam = c.ws + S_COUNT * mm;

    double* V = c.slot(S_V);    double* Cie = c.slot(S_CIE); double* Jse = c.slot(S_JSE);
    double* W = c.slot(S_W);    double* P = c.slot(S_P);     double* Q = c.slot(S_Q);
    double* Sr = c.slot(S_SR);  double* Si = c.slot(S_SI);   double* G = c.slot(S_G);
    double* Dd = c.slot(S_D);   double* R = c.slot(S_R);     double* Z0 = c.slot(S_Z0);
    double* Pp = c.slot(S_PP);  double* Hp = c.slot(S_HP);   double* U = c.slot(S_U);
    double* Vv = c.slot(S_VV);  double* T1 = c.slot(S_T1);   double* T2 = c.slot(S_T2);
    double* T3 = c.slot(S_T3);  double* T4 = c.slot(S_T4);   double* T5 = c.slot(S_T5);

    const double* Crb = a.Cr + (long long)b * a.sC_b;
    const double* Cib = a.Ci + (long long)b * a.sC_b;

    // 1. scale = max(1, tr(Cr)/m); Cr/scale into LDS, Ci/scale into T1 (dual_recommended_hvp.py:106-107)
    double tr = 0.0;
    for (int i = c.tid; i < m; i += DT) tr += Crb[(long long)i * a.ldc + i];
    tr = wg_sum(c, tr);
    const double scale = fmax(tr / m, 1.0);
    double* Aj = big;                                   // Jacobi matrix, stride mp
    int* pairs = reinterpret_cast<int*>(big + m * mp);
    double* rot = big + m * mp + m;                     // 3 * (m/2) doubles after the pair ints
    double* Vj = a.v_in_lds ? (rot + 3 * (m / 2) + 2) : T5;
    each(c, [&](int i, int j, int e) {
        const bool in = i < m && j < m;
        // lower triangle, as eigh(UPLO='L') reads it
        if (i < m) Aj[i * mp + j] = in ? Crb[(long long)max(i, j) * a.ldc + min(i, j)] / scale : 0.0;
        T1[e] = in ? Cib[(long long)i * a.ldc + j] / scale : 0.0;
    });
    sync();

    // 2. eigh(Cr / scale) (dual_recommended_hvp.py:108) by Jacobi, sorted ascending
    const int sweeps = jacobi(c, Aj, Vj, pairs, rot, flag);
    for (int i = c.tid; i < mp; i += DT) {
        if (i >= m) { lam[i] = 1.0; continue; }       // padding: keeps lambda_i + lambda_j > 0
        const double li = Aj[i * mp + i];
        int rank = 0;
        for (int j = 0; j < m; ++j) {
            const double lj = Aj[j * mp + j];
            rank += (lj < li || (lj == li && j < i)) ? 1 : 0;
        }
        pairs[i] = rank;     // pairs[] is free after the sweeps
    }
    sync();
    for (int i = c.tid; i < m; i += DT) lam[pairs[i]] = Aj[i * mp + i];
    each(c, [&](int i, int j, int e) {
        // V[:, rank(j)] = Vj[:, j]; zero padding
        if (i < m && j < m) V[i * mp + pairs[j]] = Vj[i * mp + j];
        else V[e] = 0.0;
    });
    sync();

    // 3. eigenbasis: Ci <- skew(V^T Ci V), J <- skew(V^T J(V)) / scale (:110-111, :53-55)
    gemm<false, false>(c, T1, V, [&](int i, int j, double v) { T2[i * mp + j] = v; });
    each(c, [&](int i, int j, int e) {
        // J(V) = [V_bot; -V_top] over the leading m rows
        const int k = m / 2;
        T3[e] = (i < k) ? V[(i + k) * mp + j] : (i < m ? -V[(i - k) * mp + j] : 0.0);
    });
    sync();
    gemm<true, false>(c, V, T2, [&](int i, int j, double v) { T4[i * mp + j] = v; });
    gemm<true, false>(c, V, T3, [&](int i, int j, double v) { T1[i * mp + j] = v; });
    sync();
    each(c, [&](int i, int j, int e) {
        Cie[e] = (T4[e] - T4[j * mp + i]) / 2;
        Jse[e] = ((T1[e] - T1[j * mp + i]) / 2) / scale;
        W[e] = 0.0;
        P[e] = (i == j && i < m) ? 1.0 : 0.0;
        Q[e] = 0.0;
    });
    sync();

    long long iters = 0;
    int status = sweeps < 0 ? ST_EIGH : ST_OK;
    int n_cg = 0, n_bt = 0, n_hvp = 0;
    const double sqrt_m = sqrt((double)m);
    double* Lp = big;                              // packed Cholesky factor
    double* Rp = big + m * (m + 1) / 2;            // packed inverse factor

    for (int ns = 0; ns < MAX_NEWTON && status == ST_OK; ++ns) {
        // -- sandwich S = N C N^H (:120-125)
        if (ns == 0) {
            each(c, [&](int i, int j, int e) { Sr[e] = (i == j) ? (i < m ? lam[i] : 0.0) : 0.0; Si[e] = Cie[e]; });
        } else {
            gemm<false, false>(c, W, P, [&](int i, int j, double v) { Q[i * mp + j] = -v; });
            sync();
            // Yr = P diag(lam) - Q Ci -> T1 ; Yi = P Ci + Q diag(lam) -> T2 (hvp_fused.cu:39-80)
            gemm<false, false>(c, Q, Cie, [&](int i, int j, double v) {
                const double d = j < m ? lam[j] : 0.0;
                T1[i * mp + j] = P[i * mp + j] * d - v;
            });
            gemm<false, false>(c, P, Cie, [&](int i, int j, double v) {
                const double d = j < m ? lam[j] : 0.0;
                T2[i * mp + j] = v + Q[i * mp + j] * d;
            });
            sync();
            // Ur = Yr P^T + Yi Q^T -> T3 ; Ui = Yi P^T - Yr Q^T -> T4 (recovery_fused.cu on m x m)
            gemm_nt2(c, T1, P, T2, Q, false, [&](int i, int j, double v) { T3[i * mp + j] = v; });
            gemm_nt2(c, T2, P, T1, Q, true, [&](int i, int j, double v) { T4[i * mp + j] = v; });
            sync();
            each(c, [&](int i, int j, int e) {
                Sr[e] = (T3[e] + T3[j * mp + i]) / 2;
                Si[e] = (T4[e] - T4[j * mp + i]) / 2;
            });
        }
        sync();
        // -- gradient and residual (:126-129)
        double gg = 0.0;
        each(c, [&](int, int, int e) { const double g = Jse[e] - Si[e]; G[e] = g; gg = fma(g, g, gg); });
        const double gnorm = sqrt(wg_sum(c, gg));
        const double residual = gnorm * scale / sqrt_m;
        if (!(residual >= FEAS_TOL / 2)) {
            if (residual != residual) status = ST_NONFINITE;
            break;
        }
        ++iters;

        // -- direction (:138-162)
        if (ns == 0) {
            precondition(c, G, lam, Dd, -1.0);
            sync();
            each(c, [&](int, int, int e) { Z0[e] = Dd[e]; });
        } else {
            precondition(c, G, lam, Z0, -1.0);
            each(c, [&](int, int, int e) { R[e] = -G[e]; Dd[e] = 0.0; });
            sync();
            double rz = dot(c, R, Z0);
            each(c, [&](int, int, int e) { Pp[e] = Z0[e]; });
            const double tol = fmin(fmax(sqrt(residual), 1e-8), 0.1) * gnorm;
            sync();
            for (int it = 0; it < MAX_CG; ++it) {
                // Hp = hvp(P, Q, Sr, Si, p): U = p Sr, V = p Si, Z = P U - Q V, H = Z - Z^T
                gemm<false, false>(c, Pp, Sr, [&](int i, int j, double v) { U[i * mp + j] = v; });
                gemm<false, false>(c, Pp, Si, [&](int i, int j, double v) { Vv[i * mp + j] = v; });
                sync();
                gemm_diff(c, P, U, Q, Vv, [&](int i, int j, double v) { T1[i * mp + j] = v; });
                sync();
                each(c, [&](int i, int j, int e) { Hp[e] = (i == j) ? 0.0 : T1[e] - T1[j * mp + i]; });
                sync();
                ++n_hvp; ++n_cg;
                const double curvature = dot(c, Pp, Hp);
                if (!(curvature > 0.0)) break;
                const double alpha = rz / curvature;
                double rr = 0.0;
                each(c, [&](int, int, int e) {
                    Dd[e] = Dd[e] + alpha * Pp[e];
                    const double rv = R[e] - alpha * Hp[e];
                    R[e] = rv;
                    rr = fma(rv, rv, rr);
                });
                const double rnorm = sqrt(wg_sum(c, rr));
                if (!(rnorm > tol)) break;
                precondition(c, R, lam, T2, 1.0);
                sync();
                const double next_rz = dot(c, R, T2);
                const double beta = next_rz / rz;
                each(c, [&](int, int, int e) { Pp[e] = T2[e] + beta * Pp[e]; });
                rz = next_rz;
                sync();
            }
        }
        sync();
        // -- fallback to the preconditioned gradient (:163-164)
        double slope = dot(c, G, Dd);
        if (slope >= 0.0) {
            each(c, [&](int, int, int e) { Dd[e] = Z0[e]; });
            sync();
            slope = dot(c, G, Dd);
        }

        // -- Armijo backtracking with the exact change (:166-188)
        double step = 1.0;
        bool accepted = false;
        for (int bt = 0; bt < MAX_BACKTRACK; ++bt) {
            ++n_bt;
            each(c, [&](int, int, int e) { T3[e] = step * Dd[e]; });                // delta
            sync();
            each(c, [&](int i, int j, int e) {
                T1[e] = ((W[e] + T3[e]) - (W[j * mp + i] + T3[j * mp + i])) / 2;   // trial
            });
            sync();
            gemm<false, false>(c, T1, T1, [&](int i, int j, double v) {
                T2[i * mp + j] = (i == j ? 1.0 : 0.0) + v;                          // K = I + trial^2
            });
            sync();
            const bool valid = cholesky(c, T2, Lp);
            if (valid) {
                tri_inverse(c, Lp, Rp, T4);                                          // R = L^-1
                gemm<true, false>(c, T4, T4, [&](int i, int j, double v) { T2[i * mp + j] = v; });  // P' = R^T R
                gemm<false, false>(c, T3, Sr, [&](int i, int j, double v) { U[i * mp + j] = v; });  // delta Sr
                gemm<false, false>(c, T3, Si, [&](int i, int j, double v) { Vv[i * mp + j] = v; }); // delta Si
                sync();
                gemm<false, false>(c, T2, T3, [&](int i, int j, double v) { T4[i * mp + j] = v; }); // P' delta
                gemm<false, false>(c, T1, Vv, [&](int i, int j, double v) { T5[i * mp + j] = v; }); // trial delta Si
                sync();
                double tail = 0.0, first = 0.0;
                each(c, [&](int, int, int e) {
                    tail = fma(T4[e], U[e] + T5[e], tail);
                    first = fma(G[e], T3[e], first);
                });
                tail = wg_sum(c, tail);
                first = wg_sum(c, first);
                const double change = first + tail;
                if (change <= 1e-4 * step * slope) {
                    each(c, [&](int, int, int e) { W[e] = T1[e]; P[e] = T2[e]; });
                    sync();
                    accepted = true;
                    break;
                }
            }
            step = step / 2;
        }
        if (!accepted) { status = ST_NO_DESCENT; break; }
        if (ns == MAX_NEWTON - 1) status = ST_NEWTON_CAP;
    }
    sync();

    // 4. back to the original basis (:58-64): omega = skew(V W V^T), P, Q rotated
    gemm<false, true>(c, W, V, [&](int i, int j, double v) { T1[i * mp + j] = v; });
    gemm<false, true>(c, P, V, [&](int i, int j, double v) { T2[i * mp + j] = v; });
    gemm<false, true>(c, Q, V, [&](int i, int j, double v) { T3[i * mp + j] = v; });
    sync();
    double* Pf = a.Pf + (long long)b * mm;
    double* Qf = a.Qf + (long long)b * mm;
    gemm<false, false>(c, V, T1, [&](int i, int j, double v) { T4[i * mp + j] = v; });
    gemm<false, false>(c, V, T2, [&](int i, int j, double v) { Pf[i * mp + j] = v; });
    gemm<false, false>(c, V, T3, [&](int i, int j, double v) { Qf[i * mp + j] = v; });
    sync();
    double* Ob = a.omega + (long long)b * a.sO_b;
    for (int e = c.tid; e < m * m; e += DT) {
        const int i = e / m, j = e - (e / m) * m;
        Ob[(long long)i * a.ldo + j] = (T4[i * mp + j] - T4[j * mp + i]) / 2;
    }
    if (c.tid == 0) {
        a.iters[b] = iters;
        a.status[b] = status;
        if (a.counters) {
            a.counters[4 * b + 0] = (int)iters;
            a.counters[4 * b + 1] = n_cg;
            a.counters[4 * b + 2] = n_bt;
            a.counters[4 * b + 3] = n_hvp;
        }
    }
}

// ---------------------------------------------------------------- launcher
size_t dual_workspace_doubles(int batch, int m) {
    const long long mp = pad16(m);
    return (size_t)batch * (S_COUNT * mp * mp + mp);
}

static size_t dual_lds_bytes(int m, bool v_in_lds) {
    const int mp = pad16(m);
    const size_t head = 16 * sizeof(double);
    const size_t jac = (size_t)m * mp * 8 + (size_t)m * 8 + 3 * (m / 2) * 8 + 16 +
                       (v_in_lds ? (size_t)m * mp * 8 : 0);
    const size_t chol = (size_t)m * (m + 1) * 8;
    return head + (jac > chol ? jac : chol);
}

hipError_t launch_dual(hipStream_t stream, const double* Cr, const double* Ci, long long sC_b, int ldc,
                       int batch, int m, double* omega, long long sO_b, int ldo,
                       double* Pf, double* Qf, long long* iters, int* status, int* counters,
                       double* ws, size_t lds_limit)
{
    const int mp = pad16(m);
    bool v_in_lds = dual_lds_bytes(m, true) <= lds_limit;
    const size_t lds = dual_lds_bytes(m, v_in_lds);
    if (lds > lds_limit) return hipErrorInvalidValue;
    DualArgs a;
    a.Cr = Cr; a.Ci = Ci; a.sC_b = sC_b; a.ldc = ldc; a.m = m; a.mp = mp;
    a.omega = omega; a.sO_b = sO_b; a.ldo = ldo; a.Pf = Pf; a.Qf = Qf;
    a.iters = iters; a.status = status; a.counters = counters; a.ws = ws;
    a.v_in_lds = v_in_lds ? 1 : 0;
    hipError_t err = hipFuncSetAttribute((const void*)dual_kernel,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(dual_kernel, dim3(batch), dim3(DT), lds, stream, a);
    return hipGetLastError();
}

}  // namespace ssn
