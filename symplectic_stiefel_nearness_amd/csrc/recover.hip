// Recovery of the primal solution X = [A P^T + D Q^T ; D P^T - A Q^T] for a batch
// of targets T = [A; D], i.e. X_c = B N^H with B = A + iD, N = P + iQ.
// Replaces recover_fused (src/symplectic_projection/recovery_fused.cu:1-41) and the
// chunk loop of recover_matmul (dual_recommended_hvp.py:194-228).
//
// Row r of [X_top | X_bot] is the row [a_r | d_r] times the 2mp x 2mp matrix
//     M = [[P^T, -Q^T], [Q^T, P^T]],
// so the kernel is a streaming GEMM (n x 2mp) * (2mp x 2mp) on FP64 MFMA. M is
// supplied pre-swizzled in MFMA B-fragment order (built once per target by
// `build_recovery_fragments`), so every B-operand load of a wave is one
// contiguous 512-byte read that all workgroups of the target share through L2.
#include "common.h"
#include "uneven.h"

namespace ssn {

constexpr int REC_WAVES = 4;
constexpr int REC_THREADS = REC_WAVES * WAVE;
constexpr int REC_ROWS = REC_WAVES * 16;    // rows of A (and of D) per workgroup

// frag[b][ks][J][lane] = M[4 ks + (lane >> 4)][16 J + (lane & 15)], ks < 2mp/4, J < 2mp/16.
// P, Q are [batch][mp][mp] row-major with zero padding.
// Element e (< 4 mp^2) of one target's fragment-ordered recovery matrix.
__device__ __forceinline__ void recovery_fragment_element(const double* __restrict__ Pb, const double* __restrict__ Qb,
                                                          int mp, double* __restrict__ fb, int e)
{
    const int m2 = 2 * mp;
    const int lane = e & 63;
    const int J = (e >> 6) % (m2 / 16);
    const int ks = (e >> 6) / (m2 / 16);
    const int k = 4 * ks + (lane >> 4), j = 16 * J + (lane & 15);
    double v;
    if (k < mp) v = (j < mp) ? Pb[j * mp + k] : -Qb[(j - mp) * mp + k];           // P^T | -Q^T
    else        v = (j < mp) ? Qb[j * mp + (k - mp)] : Pb[(j - mp) * mp + (k - mp)]; // Q^T |  P^T
    fb[e] = v;
}

__global__ void build_recovery_fragments(const double* __restrict__ P, const double* __restrict__ Q,
                                         int mp, double* __restrict__ frag)
{
    const int b = blockIdx.y;
    const int m2 = 2 * mp;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m2 * m2) return;
    recovery_fragment_element(P + (long long)b * mp * mp, Q + (long long)b * mp * mp, mp,
                              frag + (long long)b * m2 * m2, e);
}

// Streaming product of the rows of up to two 2n x m inputs with a fragment-ordered matrix:
//   [out_top | out_bot](r) = [a_r | d_r (| gt_r | gb_r)] * M,   M of size (2 NIN mp) x 2mp,
// where a/d (gt/gb) are the top/bottom halves of T (G). NIN = 1 is the recovery, NIN = 2 the
// backward product dL/dT = [Re; Im](G_c N + 2i B K).
// Body of one workgroup of REC_THREADS threads on rows [row0, row0 + REC_ROWS) of target b.
template <int NT, int NIN>
__device__ __forceinline__ void stream_gemm_body(const double* __restrict__ T, long long sT_b, long long sT_r,
                                                 const double* __restrict__ G, long long sG_b, long long sG_r,
                                                 const double* __restrict__ frag, int n, int m,
                                                 double* __restrict__ X, long long sX_b, long long sX_r,
                                                 int row0, int b)
{
    constexpr int MP = NT * 16;
    constexpr int NJ = 2 * NT;            // output tiles per row block (top + bottom)
    constexpr int KS = 2 * NIN * MP / 4;  // k-steps
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = lane & 15, q = lane >> 4;
    const int row = row0 + wave * 16 + r;      // A-operand row of this lane
    const bool row_ok = row < n;
    const double* Tb = T + (long long)b * sT_b;
    const double* Arow = Tb + (long long)row * sT_r;
    const double* Drow = Tb + (long long)(row + n) * sT_r;
    const double* Gtrow = nullptr;
    const double* Gbrow = nullptr;
    if (NIN == 2) {
        const double* Gb = G + (long long)b * sG_b;
        Gtrow = Gb + (long long)row * sG_r;
        Gbrow = Gb + (long long)(row + n) * sG_r;
    }
    const double* F = frag + (long long)b * (2 * NIN * MP) * (2 * MP);

    d4 acc[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) acc[j] = zero4();

    // k-step ks reads column 4 ks + q of part (4 ks) / MP of the row (MP is a multiple of 4, so
    // the part is the same for every lane); the row elements are requested DA k-steps ahead
    // (they stream from HBM) and the fragments of M one k-step ahead (L2)
    auto aload = [&](int ks) -> double {
        const int part = (4 * ks) / MP, kk = 4 * ks - part * MP + q;
        const double* src = part == 0 ? Arow : part == 1 ? Drow : part == 2 ? Gtrow : Gbrow;
        return (row_ok && kk < m) ? src[kk] : 0.0;
    };
    constexpr int DA = 8;
    double ar[DA], br[2][NJ];
#pragma unroll
    for (int d = 0; d < DA; ++d) ar[d] = aload(d);
#pragma unroll
    for (int j = 0; j < NJ; ++j) br[0][j] = F[j * 64 + lane];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
        const double a = ar[ks % DA];
        if (ks + DA < KS) ar[ks % DA] = aload(ks + DA);
        if (ks + 1 < KS) {
            const double* Fn = F + (long long)(ks + 1) * NJ * 64 + lane;
#pragma unroll
            for (int j = 0; j < NJ; ++j) br[(ks + 1) & 1][j] = Fn[j * 64];
        }
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[j] = mfma(a, br[ks & 1][j], acc[j]);
    }

    // lane holds D[q + 4i][r] of each tile: rows wave*16 + q + 4i, column 16 J + r
    double* Xb = X + (long long)b * sX_b;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const bool top = j < NT;
        const int col = (top ? j : j - NT) * 16 + r;
        if (col >= m) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int orow = row0 + wave * 16 + q + 4 * i;
            if (orow < n)
                Xb[(long long)(top ? orow : orow + n) * sX_r + col] = acc[j][i];
        }
    }
}

template <int NT, int NIN>
__global__ __launch_bounds__(REC_THREADS)
void stream_gemm_kernel(const double* __restrict__ T, long long sT_b, long long sT_r,
                        const double* __restrict__ G, long long sG_b, long long sG_r,
                        const double* __restrict__ frag, int n, int m,
                        double* __restrict__ X, long long sX_b, long long sX_r)
{
    stream_gemm_body<NT, NIN>(T, sT_b, sT_r, G, sG_b, sG_r, frag, n, m, X, sX_b, sX_r,
                              blockIdx.x * REC_ROWS, blockIdx.y);
}

// Backward fragments: M_bwd (4mp x 2mp) with row blocks
//   a: [-2 Ki, 2 Kr],  d: [-2 Kr, -2 Ki],  gt: [P, Q],  gb: [-Q, P]
// so that [a | d | gt | gb] M_bwd = [Re | Im](G_c N + 2i B K), in MFMA B-fragment order.
// Kr, Ki are read at batch stride sK_b; P, Q are [batch][mp][mp].
__global__ void build_backward_fragments(const double* __restrict__ Kr, const double* __restrict__ Ki,
                                         long long sK_b, const double* __restrict__ P,
                                         const double* __restrict__ Q, int mp, double* __restrict__ frag)
{
    const int b = blockIdx.y;
    const int rows = 4 * mp, cols = 2 * mp;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * cols) return;
    const int lane = e & 63;
    const int J = (e >> 6) % (cols / 16);
    const int ks = (e >> 6) / (cols / 16);
    const int k = 4 * ks + (lane >> 4), j = 16 * J + (lane & 15);
    const int part = k / mp, kk = k - part * mp;
    const bool left = j < mp;
    const int jj = left ? j : j - mp;
    const double* kr = Kr + (long long)b * sK_b;
    const double* ki = Ki + (long long)b * sK_b;
    const double* Pb = P + (long long)b * mp * mp;
    const double* Qb = Q + (long long)b * mp * mp;
    const int o = kk * mp + jj;
    double v;
    switch (part) {
        case 0: v = left ? -2.0 * ki[o] : 2.0 * kr[o]; break;
        case 1: v = left ? -2.0 * kr[o] : -2.0 * ki[o]; break;
        case 2: v = left ? Pb[o] : Qb[o]; break;
        default: v = left ? -Qb[o] : Pb[o]; break;
    }
    frag[(long long)b * rows * cols + e] = v;
}

hipError_t launch_recovery_fragments(hipStream_t stream, const double* P, const double* Q,
                                     int batch, int m, double* frag)
{
    const int mp = pad16(m);
    const int total = 4 * mp * mp;
    dim3 grid((total + 255) / 256, batch), block(256);
    hipLaunchKernelGGL(build_recovery_fragments, grid, block, 0, stream, P, Q, mp, frag);
    return hipGetLastError();
}

template <int NIN>
static hipError_t launch_stream(hipStream_t stream, const double* T, long long sT_b, long long sT_r,
                                const double* G, long long sG_b, long long sG_r, const double* frag,
                                int batch, int n, int m, double* X, long long sX_b, long long sX_r)
{
    const int nt = pad16(m) / 16;
    dim3 grid((n + REC_ROWS - 1) / REC_ROWS, batch), block(REC_THREADS);
#define SSN_REC_CASE(NTV) \
    case NTV: hipLaunchKernelGGL((stream_gemm_kernel<NTV, NIN>), grid, block, 0, stream, \
                                 T, sT_b, sT_r, G, sG_b, sG_r, frag, n, m, X, sX_b, sX_r); break;
    switch (nt) {
        SSN_REC_CASE(1) SSN_REC_CASE(2) SSN_REC_CASE(3) SSN_REC_CASE(4)
        SSN_REC_CASE(5) SSN_REC_CASE(6) SSN_REC_CASE(7) SSN_REC_CASE(8)
        default: return LAUNCH_UNSUPPORTED;
    }
#undef SSN_REC_CASE
    return hipGetLastError();
}

hipError_t launch_recover(hipStream_t stream, const double* T, long long sT_b, long long sT_r,
                          const double* frag, int batch, int n, int m,
                          double* X, long long sX_b, long long sX_r)
{
    return launch_stream<1>(stream, T, sT_b, sT_r, nullptr, 0, 0, frag, batch, n, m, X, sX_b, sX_r);
}

hipError_t launch_backward_stream(hipStream_t stream, const double* T, long long sT_b, long long sT_r,
                                  const double* G, long long sG_b, long long sG_r,
                                  const double* Kr, const double* Ki, long long sK_b,
                                  const double* P, const double* Q, double* frag,
                                  int batch, int n, int m, double* dT, long long sD_b, long long sD_r)
{
    const int mp = pad16(m);
    const int total = 8 * mp * mp;
    hipLaunchKernelGGL(build_backward_fragments, dim3((total + 255) / 256, batch), dim3(256), 0, stream,
                       Kr, Ki, sK_b, P, Q, mp, frag);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_stream<2>(stream, T, sT_b, sT_r, G, sG_b, sG_r, frag, batch, n, m, dT, sD_b, sD_r);
}


// ---------------------------------------------------------------- recovery with P, Q in LDS
// The same product X = [a | d] M, M = [[P^T, -Q^T], [Q^T, P^T]], for mp <= 96: each workgroup
// holds its target's P and Q in LDS (2 mp (mp + 1) doubles, odd row stride) and streams row
// tiles of that target, so the B fragments M[4 ks + q][16 J + r] = +-(P or Q)[16 J' + r][k'] are
// LDS reads instead of L2 reads of a pre-built fragment array (no fragment pass either). The
// products and their order are those of stream_gemm_kernel, so X is bitwise the same.
constexpr int RL_THREADS = 512;
constexpr int RL_WAVES = RL_THREADS / WAVE;

template <int NT>
__host__ __device__ constexpr int recover_lds_doubles() { return 2 * (16 * NT) * (16 * NT + 1); }

template <int NT>
__global__ __launch_bounds__(RL_THREADS)
void recover_lds_kernel(const double* __restrict__ T, long long sT_b, long long sT_r,
                        const double* __restrict__ P, const double* __restrict__ Q, int n, int m,
                        double* __restrict__ X, long long sX_b, long long sX_r, int wpt)
{
    constexpr int MP = 16 * NT, LS = MP + 1, NJ = 2 * NT, KS = 2 * MP / 4;
    extern __shared__ double sh[];
    double* Ps = sh;
    double* Qs = sh + MP * LS;
    const int b = blockIdx.y;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, r = lane & 15, q = lane >> 4;
    const double* Pb = P + (long long)b * MP * MP;
    const double* Qb = Q + (long long)b * MP * MP;
    for (int e = tid; e < MP * MP; e += RL_THREADS) {
        const int i = e / MP, j = e - (e / MP) * MP;
        Ps[i * LS + j] = Pb[e];
        Qs[i * LS + j] = Qb[e];
    }
    __syncthreads();
    const double* Tb = T + (long long)b * sT_b;
    double* Xb = X + (long long)b * sX_b;
    const int ntiles = (n + 15) / 16;
    for (int tile = blockIdx.x * RL_WAVES + wave; tile < ntiles; tile += wpt * RL_WAVES) {
        const int row = tile * 16 + r;
        const bool row_ok = row < n;
        const double* Arow = Tb + (long long)row * sT_r;
        const double* Drow = Tb + (long long)(row + n) * sT_r;
        auto aload = [&](int ks) -> double {
            const int part = (4 * ks) / MP, kk = 4 * ks - part * MP + q;
            return (row_ok && kk < m) ? (part ? Drow : Arow)[kk] : 0.0;
        };
        d4 acc[NJ];
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[j] = zero4();
        constexpr int DA = 8;
        double ar[DA];
#pragma unroll
        for (int d = 0; d < DA; ++d) ar[d] = aload(d);
        const double* Pl = Ps + r * LS + q;
        const double* Ql = Qs + r * LS + q;
        // B fragment J of k-step ks: M[k][j] = [P^T | -Q^T] (k < mp) or [Q^T | P^T] (k >= mp),
        // i.e. +-(P or Q)[16 J' + r][k'], read from LDS one k-step ahead
        auto bload = [&](int ks, double (&dst)[NJ]) {
            const int part = (4 * ks) / MP, kk0 = 4 * ks - part * MP;
#pragma unroll
            for (int J = 0; J < NJ; ++J) {
                const int jc = J / NT, Jm = J - jc * NT;
                const double v = ((part == jc) ? Pl : Ql)[16 * Jm * LS + kk0];
                dst[J] = (part == 0 && jc == 1) ? -v : v;
            }
        };
        double br[2][NJ];
        bload(0, br[0]);
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const double a = ar[ks % DA];
            if (ks + DA < KS) ar[ks % DA] = aload(ks + DA);
            if (ks + 1 < KS) bload(ks + 1, br[(ks + 1) & 1]);
#pragma unroll
            for (int J = 0; J < NJ; ++J) acc[J] = mfma(a, br[ks & 1][J], acc[J]);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const bool top = j < NT;
            const int col = (top ? j : j - NT) * 16 + r;
            if (col >= m) continue;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int orow = tile * 16 + q + 4 * i;
                if (orow < n)
                    Xb[(long long)(top ? orow : orow + n) * sX_r + col] = acc[j][i];
            }
        }
    }
}

// Recovery straight from P and Q (no fragment array) when the LDS holds both: mp <= 96.
// Returns LAUNCH_DECLINED otherwise (the caller builds fragments and uses launch_recover).
hipError_t launch_recover_pq(hipStream_t stream, const double* T, long long sT_b, long long sT_r,
                             const double* P, const double* Q, int batch, int n, int m,
                             double* X, long long sX_b, long long sX_r, size_t lds_limit)
{
    const int nt = pad16(m) / 16;
    if (nt > 6) return LAUNCH_DECLINED;
    const int ntiles = (n + 15) / 16;
    int wpt = (256 + batch - 1) / batch;                      // about one workgroup per CU
    const int most = (ntiles + RL_WAVES - 1) / RL_WAVES;
    if (wpt > most) wpt = most;
    if (wpt < 1) wpt = 1;
    dim3 grid(wpt, batch), block(RL_THREADS);
#define SSN_RL_CASE(NTV) \
    case NTV: { \
        const size_t lds = (size_t)recover_lds_doubles<NTV>() * 8; \
        if (lds > lds_limit) return LAUNCH_DECLINED; \
        hipError_t err = hipFuncSetAttribute((const void*)recover_lds_kernel<NTV>, \
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
        if (err != hipSuccess) return err; \
        hipLaunchKernelGGL(recover_lds_kernel<NTV>, grid, block, lds, stream, T, sT_b, sT_r, P, Q, n, m, \
                           X, sX_b, sX_r, wpt); \
        break; }
    switch (nt) {
        SSN_RL_CASE(1) SSN_RL_CASE(2) SSN_RL_CASE(3) SSN_RL_CASE(4) SSN_RL_CASE(5) SSN_RL_CASE(6)
        default: return LAUNCH_DECLINED;
    }
#undef SSN_RL_CASE
    return hipGetLastError();
}

// ---------------------------------------------------------------- uneven batches
__global__ void recovery_fragments_uneven_kernel(const UTarget* __restrict__ tg)
{
    const UTarget& t = tg[blockIdx.y];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < 4 * t.mp * t.mp) recovery_fragment_element(t.Pf, t.Qf, t.mp, t.frag, e);
}

// Persistent grid over the (target, first row) items of a plan: each item is rec_rows rows
// (a multiple of REC_ROWS) of one target, streamed REC_ROWS at a time.
__global__ __launch_bounds__(REC_THREADS)
void recover_uneven_kernel(const UTarget* __restrict__ tg, const UItem* __restrict__ items, int nitems)
{
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const UItem w = items[it];
        const UTarget& t = tg[w.target];
        for (int row0 = w.index; row0 < w.index + t.rec_rows && row0 < t.n; row0 += REC_ROWS) {
            switch (t.mp / 16) {
#define SSN_RU_CASE(NTV) \
                case NTV: stream_gemm_body<NTV, 1>(t.T, 0, t.sT_r, nullptr, 0, 0, t.frag, t.n, t.m, t.X, 0, t.sX_r, \
                                                   row0, 0); break;
                SSN_RU_CASE(1) SSN_RU_CASE(2) SSN_RU_CASE(3) SSN_RU_CASE(4)
                SSN_RU_CASE(5) SSN_RU_CASE(6) SSN_RU_CASE(7) SSN_RU_CASE(8)
#undef SSN_RU_CASE
                default: break;
            }
        }
    }
}

hipError_t launch_recover_uneven(hipStream_t stream, const UTarget* tg, const UItem* items, int nitems,
                                 int count, int max_mp)
{
    hipLaunchKernelGGL(recovery_fragments_uneven_kernel, dim3((4 * max_mp * max_mp + 255) / 256, count), dim3(256),
                       0, stream, tg);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || nitems == 0) return e;
    const int grid = nitems < 1024 ? nitems : 1024;
    hipLaunchKernelGGL(recover_uneven_kernel, dim3(grid), dim3(REC_THREADS), 0, stream, tg, items, nitems);
    return hipGetLastError();
}

}  // namespace ssn
