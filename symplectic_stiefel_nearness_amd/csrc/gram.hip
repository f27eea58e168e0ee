// Tall-skinny Gram of a batch of targets T = [A; D] (2n x m, FP64):
//   Cr = A^T A + D^T D  (symmetric),  Ci = A^T D - D^T A  (skew).
// Replaces the reference's gram_offdiag / gram_diagonal kernels
// (src/symplectic_projection/gram_triangle_reflect.cu:1-104) and the chunk loop of
// gram_chunks (dual_recommended_hvp.py:79-96).
//
// Design (gfx950): one workgroup of 8 waves per (row chunk, target). Rows are
// streamed from HBM through a double-buffered LDS stage of 16 rows of A and D;
// every wave owns a fixed list of 16x16 output tiles and accumulates them with
// v_mfma_f64_16x16x4f64 (the row index is the MFMA K dimension). Only the
// upper tiles of A^T A + D^T D are formed (one accumulator takes both
// products); all tiles of A^T D are formed and the skew part is taken in the
// deterministic second stage, which sums the chunk partials in a fixed order.
#include "common.h"
#include "uneven.h"

#include <type_traits>

namespace ssn {

constexpr int GRAM_WAVES = 8;
constexpr int GRAM_THREADS = GRAM_WAVES * WAVE;
constexpr int GRAM_ROWS = 16;       // rows of A (and of D) per LDS stage

template <int NT>
struct GramShape {
    static constexpr int MP = NT * 16;
    static constexpr int NCR = NT * (NT + 1) / 2;   // upper tiles of Cr
    static constexpr int NAD = NT * NT;             // all tiles of A^T D
    static constexpr int NTOT = NCR + NAD;
    static constexpr int LDS_STRIDE = MP + 2;       // doubles per staged row
    static constexpr int STAGE = 2 * GRAM_ROWS * LDS_STRIDE;   // A rows then D rows
    static constexpr int PER_THREAD = (2 * GRAM_ROWS * MP + GRAM_THREADS - 1) / GRAM_THREADS;
};

// Tile t of the plan: t < NCR is the upper tile (I <= J) of A^T A + D^T D in column-major
// upper order, the others are the tiles of A^T D in row-major order.
struct GramTile { int I, J; bool cr; };

constexpr GramTile gram_tile(int t, int nt) {
    const int ncr = nt * (nt + 1) / 2;
    if (t < ncr) {
        int col = 0, rem = t;
        while (rem > col) { rem -= col + 1; ++col; }
        return GramTile{rem, col, true};
    }
    return GramTile{(t - ncr) / nt, (t - ncr) % nt, false};
}

// Compile-time assignment of the tiles to the waves of a workgroup: tiles in plan order, each to
// the wave with the fewest MFMAs so far (an A^T A + D^T D tile costs two per k-step, an A^T D
// tile one), so the busiest wave issues ceil(MFMAs / 8) per k-step.
template <int NT>
struct GramPlan {
    static constexpr int MAXW = 24;
    struct Table { int tile[GRAM_WAVES][MAXW]; int count[GRAM_WAVES]; int maxt; };
    static constexpr Table make() {
        Table tb{};
        int load[GRAM_WAVES] = {};
        for (int t = 0; t < GramShape<NT>::NTOT; ++t) {
            int w = 0;
            for (int v = 1; v < GRAM_WAVES; ++v)
                if (load[v] < load[w]) w = v;
            tb.tile[w][tb.count[w]++] = t;
            load[w] += gram_tile(t, NT).cr ? 2 : 1;
        }
        for (int v = 0; v < GRAM_WAVES; ++v)
            if (tb.count[v] > tb.maxt) tb.maxt = tb.count[v];
        return tb;
    }
    static constexpr Table table = make();
    static_assert(table.maxt <= MAXW, "tile plan overflow");
};

template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// partial layout: [batch][chunks][2][MP][MP]; slot 0 = A^T A + D^T D (upper tiles),
// slot 1 = A^T D (all tiles).
// The chunk loop of wave W: every thread stages rows, the wave accumulates the tiles of its
// plan from the 2 NT column fragments of each k-step, read once from LDS.
template <int NT, int W>
__device__ __forceinline__ void gram_partial_wave(const double* __restrict__ Tb, long long sT_r, int n, int m,
                                                  int row_begin, int row_end, double* __restrict__ out,
                                                  double* __restrict__ stage_lds)
{
    using S = GramShape<NT>;
    constexpr int MP = S::MP;
    constexpr int NW = GramPlan<NT>::table.count[W];

    const int tid = threadIdx.x, lane = tid & 63;
    const int r = lane & 15, q = lane >> 4;

    d4 acc[NW > 0 ? NW : 1];
#pragma unroll
    for (int s = 0; s < NW; ++s) acc[s] = zero4();

    // Static per-thread load slots: element e of a stage -> (half, row, col).
    double pre[S::PER_THREAD];
    auto load_stage = [&](int r0) {
#pragma unroll
        for (int s = 0; s < S::PER_THREAD; ++s) {
            const int e = tid + s * GRAM_THREADS;
            double v = 0.0;
            if (e < 2 * GRAM_ROWS * MP) {
                const int half = e / (GRAM_ROWS * MP);
                const int rem = e - half * GRAM_ROWS * MP;
                const int rr = rem / MP, cc = rem - (rem / MP) * MP;
                const int row = r0 + rr;
                if (row < row_end && cc < m)
                    v = Tb[(long long)(row + half * n) * sT_r + cc];
            }
            pre[s] = v;
        }
    };
    auto store_stage = [&](double* dst) {
#pragma unroll
        for (int s = 0; s < S::PER_THREAD; ++s) {
            const int e = tid + s * GRAM_THREADS;
            if (e < 2 * GRAM_ROWS * MP) {
                const int half = e / (GRAM_ROWS * MP);
                const int rem = e - half * GRAM_ROWS * MP;
                const int rr = rem / MP, cc = rem - (rem / MP) * MP;
                dst[(half * GRAM_ROWS + rr) * S::LDS_STRIDE + cc] = pre[s];
            }
        }
    };

    // one barrier per stage: the buffer written at stage st+1 is the other one, and
    // the barrier of stage st+1 orders every read of stage st before the write at st+2
    const int nstages = (row_end - row_begin + GRAM_ROWS - 1) / GRAM_ROWS;
    if (nstages > 0) load_stage(row_begin);
    for (int st = 0; st < nstages; ++st) {
        double* cur = stage_lds + (st & 1) * S::STAGE;
        store_stage(cur);
        __syncthreads();
        if (st + 1 < nstages) load_stage(row_begin + (st + 1) * GRAM_ROWS);
        const double* sA = cur;
        const double* sD = cur + GRAM_ROWS * S::LDS_STRIDE;
#pragma unroll
        for (int kq = 0; kq < GRAM_ROWS / 4; ++kq) {
            const int row = (kq * 4 + q) * S::LDS_STRIDE + r;
            double fa[NT], fd[NT];      // fragments this wave does not use are never read
#pragma unroll
            for (int j = 0; j < NT; ++j) { fa[j] = sA[row + 16 * j]; fd[j] = sD[row + 16 * j]; }
            static_for<0, NW>([&](auto si) {
                constexpr int s = decltype(si)::value;
                constexpr GramTile g = gram_tile(GramPlan<NT>::table.tile[W][s], NT);
                if constexpr (g.cr) {
                    acc[s] = mfma(fa[g.I], fa[g.J], acc[s]);
                    acc[s] = mfma(fd[g.I], fd[g.J], acc[s]);
                } else {
                    acc[s] = mfma(fa[g.I], fd[g.J], acc[s]);
                }
            });
        }
    }

    static_for<0, NW>([&](auto si) {
        constexpr int s = decltype(si)::value;
        constexpr GramTile g = gram_tile(GramPlan<NT>::table.tile[W][s], NT);
        double* o = out + (g.cr ? 0 : MP * MP);
#pragma unroll
        for (int i = 0; i < 4; ++i)
            o[(g.I * 16 + q + 4 * i) * MP + g.J * 16 + r] = acc[s][i];
    });
}

// Body of one (row chunk, target) workgroup; `stage` is 2 * GramShape<NT>::STAGE doubles of LDS.
template <int NT>
__device__ __forceinline__ void gram_partial_body(const double* __restrict__ T, long long sT_b, long long sT_r,
                                                  int n, int m, int chunk_rows, int chunks,
                                                  double* __restrict__ partial, int chunk, int b,
                                                  double* __restrict__ stage_lds)
{
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int row_begin = chunk * chunk_rows;
    const int row_end = min(n, row_begin + chunk_rows);
    const double* Tb = T + (long long)b * sT_b;
    double* out = partial + ((long long)b * chunks + chunk) * 2 * GramShape<NT>::MP * GramShape<NT>::MP;
    switch (wave) {
#define SSN_GW(W) case W: gram_partial_wave<NT, W>(Tb, sT_r, n, m, row_begin, row_end, out, stage_lds); break;
        SSN_GW(0) SSN_GW(1) SSN_GW(2) SSN_GW(3) SSN_GW(4) SSN_GW(5) SSN_GW(6) SSN_GW(7)
#undef SSN_GW
    }
}

template <int NT>
__global__ __launch_bounds__(GRAM_THREADS) __attribute__((amdgpu_waves_per_eu(NT <= 6 ? 4 : 1)))   // 2 WGs per CU
void gram_partial_kernel(const double* __restrict__ T, long long sT_b, long long sT_r,
                         int n, int m, int chunk_rows, int chunks,
                         double* __restrict__ partial)
{
    __shared__ double stage[2 * GramShape<NT>::STAGE];
    gram_partial_body<NT>(T, sT_b, sT_r, n, m, chunk_rows, chunks, partial, blockIdx.x, blockIdx.y, stage);
}

// Deterministic second stage: Cr, Ci [batch][m][m] (leading dimension ldc) from the chunk
// partials, summed in ascending chunk order, one workgroup of 256 threads per (tile pair I <= J,
// target): thread (r, c) sums element (r, c) of the upper tile (I, J) of slot 0 (Cr) and of
// tiles (I, J), (J, I) of slot 1 (A^T D) over the chunks, so every chunk read is a contiguous
// 16-double row; the transposed operands are then exchanged through LDS. Cr is mirrored from its
// upper triangle and Ci = S - S^T from the A^T D sums, so symmetry and skewness are exact.
// l0 .. l2 are LDS tiles.
__device__ __forceinline__ void gram_reduce_tile_body(const double* __restrict__ p, int chunks, int m, int mp,
                                                      double* __restrict__ Crb, double* __restrict__ Cib,
                                                      int ldc, int I, int J, double* l0, double* l1, double* l2)
{
    constexpr int LS = 17;
    const int r = threadIdx.x >> 4, c = threadIdx.x & 15;
    const int oij = (I * 16 + r) * mp + J * 16 + c;     // element (r, c) of tile (I, J)
    const int oji = (J * 16 + r) * mp + I * 16 + c;     // element (r, c) of tile (J, I)
    double cr = 0.0, sij = 0.0, sji = 0.0;
    for (int k = 0; k < chunks; ++k) {
        const double* pc = p + (long long)k * 2 * mp * mp;
        cr += pc[oij];
        sij += pc[mp * mp + oij];
        sji += pc[mp * mp + oji];
    }
    l0[r * LS + c] = cr;
    l1[r * LS + c] = sij;
    l2[r * LS + c] = sji;
    __syncthreads();
    {   // tile (I, J): i = 16I + r, j = 16J + c
        const int i = I * 16 + r, j = J * 16 + c;
        double vr, vi;
        if (I < J) {                              // i < j
            vr = cr;
            vi = sij - l2[c * LS + r];
        } else if (r <= c) {                      // diagonal tile, upper part (lo = i)
            vr = cr;
            vi = (r == c) ? 0.0 : sij - l1[c * LS + r];
        } else {                                  // diagonal tile, lower part (lo = j)
            vr = l0[c * LS + r];
            vi = -(l1[c * LS + r] - sij);
        }
        if (i < m && j < m) {
            Crb[(long long)i * ldc + j] = vr;
            Cib[(long long)i * ldc + j] = vi;
        }
    }
    if (I < J) {   // tile (J, I): i = 16J + r > j = 16I + c, lo = j
        const int i = J * 16 + r, j = I * 16 + c;
        if (i < m && j < m) {
            Crb[(long long)i * ldc + j] = l0[c * LS + r];
            Cib[(long long)i * ldc + j] = -(l1[c * LS + r] - sji);
        }
    }
}

// blockIdx.x -> upper tile pair (I, J), I <= J, row-major over the nt x nt tiles
__device__ __forceinline__ void tile_pair(int idx, int nt, int& I, int& J) {
    int rem = idx;
    I = 0;
    while (rem >= nt - I) { rem -= nt - I; ++I; }
    J = I + rem;
}

__global__ __launch_bounds__(256)
void gram_reduce_tiles_kernel(const double* __restrict__ partial, int chunks, int m, int mp,
                              double* __restrict__ Cr, double* __restrict__ Ci,
                              long long sC_b, int ldc)
{
    __shared__ double l0[16 * 17], l1[16 * 17], l2[16 * 17];
    const int b = blockIdx.y;
    int I, J;
    tile_pair(blockIdx.x, mp / 16, I, J);
    gram_reduce_tile_body(partial + (long long)b * chunks * 2 * mp * mp, chunks, m, mp,
                          Cr + (long long)b * sC_b, Ci + (long long)b * sC_b, ldc, I, J, l0, l1, l2);
}

// ---------------------------------------------------------------- uneven batches
// Persistent grid over the (target, chunk) items of one order of a plan (largest first).
template <int NT>
__global__ __launch_bounds__(GRAM_THREADS) __attribute__((amdgpu_waves_per_eu(NT <= 6 ? 4 : 1)))
void gram_uneven_kernel(const UTarget* __restrict__ tg, const UItem* __restrict__ items, int nitems)
{
    __shared__ double stage[2 * GramShape<NT>::STAGE];
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const UItem w = items[it];
        const UTarget& t = tg[w.target];
        gram_partial_body<NT>(t.T, 0, t.sT_r, t.n, t.m, t.chunk_rows, t.chunks, t.partial, w.index, 0, stage);
        __syncthreads();      // the next item restages the LDS buffers
    }
}

// Cr, Ci (ld m) of every target of a plan from its chunk partials: blockIdx.y = target,
// blockIdx.x = tile pair (grid sized for the largest order; pairs past a target's return at once).
__global__ __launch_bounds__(256)
void gram_reduce_uneven_kernel(const UTarget* __restrict__ tg)
{
    __shared__ double l0[16 * 17], l1[16 * 17], l2[16 * 17];
    const UTarget& t = tg[blockIdx.y];
    const int nt = t.mp / 16;
    if ((int)blockIdx.x >= nt * (nt + 1) / 2) return;
    int I, J;
    tile_pair(blockIdx.x, nt, I, J);
    gram_reduce_tile_body(t.partial, t.chunks, t.m, t.mp, t.Cr, t.Ci, t.m, I, J, l0, l1, l2);
}

// ---------------------------------------------------------------- cross Gram (backward)
// M = B^H G_c for B = A + iD and G_c = G_top + i G_bot:
//   Mr = A^T G_top + D^T G_bot,   Mi = A^T G_bot - D^T G_top   (full mp x mp).
// Same streaming structure as the Gram: a 16-row LDS stage of A, D, G_top, G_bot per
// step, 8 waves each owning a fixed list of output tiles (both parts accumulated).
template <int NT>
struct CrossShape {
    static constexpr int MP = NT * 16;
    static constexpr int NTOT = NT * NT;
    static constexpr int MAXT = (NTOT + GRAM_WAVES - 1) / GRAM_WAVES;
    static constexpr int LDS_STRIDE = MP + 2;
    static constexpr int STAGE = 4 * GRAM_ROWS * LDS_STRIDE;   // A, D, Gt, Gb rows
    static constexpr int PER_THREAD = (4 * GRAM_ROWS * MP + GRAM_THREADS - 1) / GRAM_THREADS;
};

// Wave W of a cross Gram workgroup owns the units [9W, 9W + 9) of the 2 NT^2 units (tile t,
// part p) in row-major tile order (p = 0: Mr, p = 1: Mi; two MFMAs per unit per k-step), so
// every wave issues the same number of MFMAs and reads each column fragment it uses once per
// k-step. Per unit the products are accumulated in the same order as a whole-tile owner would.
template <int NT, int W>
__device__ __forceinline__ void cross_gram_wave(const double* __restrict__ Tb, long long sT_r,
                                                const double* __restrict__ Gb, long long sG_r,
                                                int n, int m, int row_begin, int row_end,
                                                double* __restrict__ out, double* __restrict__ stage_lds)
{
    using S = CrossShape<NT>;
    constexpr int MP = S::MP;
    constexpr int NU = 2 * NT * NT;
    constexpr int U0 = NU * W / GRAM_WAVES, U1 = NU * (W + 1) / GRAM_WAVES;
    constexpr int NW = U1 - U0;
    const int tid = threadIdx.x, lane = tid & 63;
    const int r = lane & 15, q = lane >> 4;

    d4 acc[NW > 0 ? NW : 1];
#pragma unroll
    for (int s = 0; s < NW; ++s) acc[s] = zero4();

    double pre[S::PER_THREAD];
    auto load_stage = [&](int r0) {
#pragma unroll
        for (int s = 0; s < S::PER_THREAD; ++s) {
            const int e = tid + s * GRAM_THREADS;
            double v = 0.0;
            if (e < 4 * GRAM_ROWS * MP) {
                const int part = e / (GRAM_ROWS * MP);          // 0 A, 1 D, 2 Gt, 3 Gb
                const int rem = e - part * GRAM_ROWS * MP;
                const int rr = rem / MP, cc = rem - (rem / MP) * MP;
                const int row = r0 + rr;
                if (row < row_end && cc < m) {
                    const int grow = row + (part & 1) * n;
                    v = (part < 2) ? Tb[(long long)grow * sT_r + cc] : Gb[(long long)grow * sG_r + cc];
                }
            }
            pre[s] = v;
        }
    };
    auto store_stage = [&](double* dst) {
#pragma unroll
        for (int s = 0; s < S::PER_THREAD; ++s) {
            const int e = tid + s * GRAM_THREADS;
            if (e < 4 * GRAM_ROWS * MP) {
                const int part = e / (GRAM_ROWS * MP);
                const int rem = e - part * GRAM_ROWS * MP;
                const int rr = rem / MP, cc = rem - (rem / MP) * MP;
                dst[(part * GRAM_ROWS + rr) * S::LDS_STRIDE + cc] = pre[s];
            }
        }
    };
    const int nstages = (row_end - row_begin + GRAM_ROWS - 1) / GRAM_ROWS;
    if (nstages > 0) load_stage(row_begin);
    for (int st = 0; st < nstages; ++st) {
        double* cur = stage_lds + (st & 1) * S::STAGE;
        store_stage(cur);
        __syncthreads();
        if (st + 1 < nstages) load_stage(row_begin + (st + 1) * GRAM_ROWS);
        const double* sA = cur;
        const double* sD = cur + GRAM_ROWS * S::LDS_STRIDE;
        const double* sGt = cur + 2 * GRAM_ROWS * S::LDS_STRIDE;
        const double* sGb = cur + 3 * GRAM_ROWS * S::LDS_STRIDE;
#pragma unroll
        for (int kq = 0; kq < GRAM_ROWS / 4; ++kq) {
            const int row = (kq * 4 + q) * S::LDS_STRIDE + r;
            double fa[NT], fd[NT], ft[NT], fb[NT];   // fragments this wave does not use are never read
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                fa[j] = sA[row + 16 * j]; fd[j] = sD[row + 16 * j];
                ft[j] = sGt[row + 16 * j]; fb[j] = sGb[row + 16 * j];
            }
            static_for<0, NW>([&](auto si) {
                constexpr int s = decltype(si)::value;
                constexpr int u = U0 + s, I = (u / 2) / NT, J = (u / 2) % NT;
                if constexpr (u % 2 == 0) {
                    acc[s] = mfma(fa[I], ft[J], acc[s]);
                    acc[s] = mfma(fd[I], fb[J], acc[s]);
                } else {
                    acc[s] = mfma(fa[I], fb[J], acc[s]);
                    acc[s] = mfma(-fd[I], ft[J], acc[s]);
                }
            });
        }
    }
    static_for<0, NW>([&](auto si) {
        constexpr int s = decltype(si)::value;
        constexpr int u = U0 + s, I = (u / 2) / NT, J = (u / 2) % NT;
        double* o = out + (u % 2) * MP * MP;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            o[(I * 16 + q + 4 * i) * MP + J * 16 + r] = acc[s][i];
    });
}

template <int NT>
__global__ __launch_bounds__(GRAM_THREADS)
void cross_gram_partial_kernel(const double* __restrict__ T, long long sT_b, long long sT_r,
                               const double* __restrict__ G, long long sG_b, long long sG_r,
                               int n, int m, int chunk_rows, int chunks, double* __restrict__ partial)
{
    using S = CrossShape<NT>;
    constexpr int MP = S::MP;
    __shared__ double stage[2][S::STAGE];
    const int chunk = blockIdx.x, b = blockIdx.y;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int row_begin = chunk * chunk_rows;
    const int row_end = min(n, row_begin + chunk_rows);
    const double* Tb = T + (long long)b * sT_b;
    const double* Gb = G + (long long)b * sG_b;
    double* out = partial + ((long long)b * chunks + chunk) * 2 * MP * MP;
    switch (wave) {
#define SSN_CW(W) case W: cross_gram_wave<NT, W>(Tb, sT_r, Gb, sG_r, n, m, row_begin, row_end, out, &stage[0][0]); break;
        SSN_CW(0) SSN_CW(1) SSN_CW(2) SSN_CW(3) SSN_CW(4) SSN_CW(5) SSN_CW(6) SSN_CW(7)
#undef SSN_CW
    }
}

// Sums the chunk partials in ascending order into Mr, Mi (each [mp][mp] at the given
// per-target strides, typically inside the solver workspace).
__global__ void cross_gram_reduce_kernel(const double* __restrict__ partial, int chunks, int mp,
                                         double* __restrict__ Mr, double* __restrict__ Mi, long long sM_b)
{
    const int b = blockIdx.y;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= mp * mp) return;
    const double* p = partial + (long long)b * chunks * 2 * mp * mp;
    double re = 0.0, im = 0.0;
    for (int c = 0; c < chunks; ++c) {
        re += p[(long long)c * 2 * mp * mp + e];
        im += p[(long long)c * 2 * mp * mp + mp * mp + e];
    }
    Mr[(long long)b * sM_b + e] = re;
    Mi[(long long)b * sM_b + e] = im;
}

}  // namespace ssn

// ------------------------------------------------------------------ launchers
namespace ssn {

#ifndef SSN_GRAM_TARGET_WGS
#define SSN_GRAM_TARGET_WGS 512
#endif

int gram_chunk_count(int batch, int n) {
    // enough workgroups to cover the chip about twice; at least 256 rows per chunk
    const int target_wgs = SSN_GRAM_TARGET_WGS;
    int chunks = (target_wgs + batch - 1) / batch;
    const int max_chunks = (n + 255) / 256;
    if (chunks > max_chunks) chunks = max_chunks;
    if (chunks < 1) chunks = 1;
    return chunks;
}

hipError_t launch_gram_partial(hipStream_t stream, const double* T, long long sT_b, long long sT_r,
                               int batch, int n, int m, int chunks, double* partial)
{
    const int mp = pad16(m);
    const int nt = mp / 16;
    int chunk_rows = (n + chunks - 1) / chunks;
    chunk_rows = (chunk_rows + GRAM_ROWS - 1) / GRAM_ROWS * GRAM_ROWS;
    dim3 grid(chunks, batch), block(GRAM_THREADS);
#define SSN_GRAM_CASE(NTV) \
    case NTV: hipLaunchKernelGGL(gram_partial_kernel<NTV>, grid, block, 0, stream, \
                                 T, sT_b, sT_r, n, m, chunk_rows, chunks, partial); break;
    switch (nt) {
        SSN_GRAM_CASE(1) SSN_GRAM_CASE(2) SSN_GRAM_CASE(3) SSN_GRAM_CASE(4)
        SSN_GRAM_CASE(5) SSN_GRAM_CASE(6) SSN_GRAM_CASE(7) SSN_GRAM_CASE(8)
        default: return LAUNCH_UNSUPPORTED;
    }
#undef SSN_GRAM_CASE
    return hipGetLastError();
}

hipError_t launch_cross_gram(hipStream_t stream, const double* T, long long sT_b, long long sT_r,
                             const double* G, long long sG_b, long long sG_r, int batch, int n, int m,
                             int chunks, double* partial, double* Mr, double* Mi, long long sM_b)
{
    const int mp = pad16(m);
    const int nt = mp / 16;
    int chunk_rows = (n + chunks - 1) / chunks;
    chunk_rows = (chunk_rows + GRAM_ROWS - 1) / GRAM_ROWS * GRAM_ROWS;
    dim3 grid(chunks, batch), block(GRAM_THREADS);
#define SSN_CROSS_CASE(NTV) \
    case NTV: hipLaunchKernelGGL(cross_gram_partial_kernel<NTV>, grid, block, 0, stream, \
                                 T, sT_b, sT_r, G, sG_b, sG_r, n, m, chunk_rows, chunks, partial); break;
    switch (nt) {
        SSN_CROSS_CASE(1) SSN_CROSS_CASE(2) SSN_CROSS_CASE(3) SSN_CROSS_CASE(4)
        SSN_CROSS_CASE(5) SSN_CROSS_CASE(6) SSN_CROSS_CASE(7) SSN_CROSS_CASE(8)
        default: return LAUNCH_UNSUPPORTED;
    }
#undef SSN_CROSS_CASE
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    dim3 g2((mp * mp + 255) / 256, batch);
    hipLaunchKernelGGL(cross_gram_reduce_kernel, g2, dim3(256), 0, stream, partial, chunks, mp, Mr, Mi, sM_b);
    return hipGetLastError();
}

hipError_t launch_gram_reduce(hipStream_t stream, const double* partial, int batch, int chunks, int m,
                              double* Cr, double* Ci, long long sC_b, int ldc)
{
    const int mp = pad16(m), nt = mp / 16;
    dim3 grid(nt * (nt + 1) / 2, batch), block(256);
    hipLaunchKernelGGL(gram_reduce_tiles_kernel, grid, block, 0, stream, partial, chunks, m, mp,
                       Cr, Ci, sC_b, ldc);
    return hipGetLastError();
}

hipError_t launch_gram_uneven(hipStream_t stream, const UTarget* tg, const UItem* items, const int* begin,
                              int count)
{
    int max_pairs = 1;                  // tile pairs of the largest order present
    for (int nt = 1; nt <= 8; ++nt)
        if (begin[nt] > begin[nt - 1]) max_pairs = nt * (nt + 1) / 2;
    for (int nt = 1; nt <= 8; ++nt) {
        const int first = begin[nt - 1], nitems = begin[nt] - first;
        if (nitems <= 0) continue;
        const int grid = nitems < 512 ? nitems : 512;     // two resident workgroups per CU
#define SSN_GU_CASE(NTV) \
        case NTV: hipLaunchKernelGGL(gram_uneven_kernel<NTV>, dim3(grid), dim3(GRAM_THREADS), 0, stream, \
                                     tg, items + first, nitems); break;
        switch (nt) {
            SSN_GU_CASE(1) SSN_GU_CASE(2) SSN_GU_CASE(3) SSN_GU_CASE(4)
            SSN_GU_CASE(5) SSN_GU_CASE(6) SSN_GU_CASE(7) SSN_GU_CASE(8)
        }
#undef SSN_GU_CASE
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(gram_reduce_uneven_kernel, dim3(max_pairs, count), dim3(256), 0, stream, tg);
    return hipGetLastError();
}

}  // namespace ssn
