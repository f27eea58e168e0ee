// Split-precision Gram and recovery: FP64 operands cut into BF16 slices, multiplied on
// the BF16 MFMA (v_mfma_f32_16x16x32_bf16) with FP32 accumulation, reassembled to FP64.
//
// Slicing (Ozaki-style, exact): a value x in a group that shares the power-of-two scale
// 2^E (max |x| < 2^E) is written u = x 2^-E in (-1, 1) and cut as
//     c_1 = rint(128 u),  u_1 = 128 u - c_1,  c_a = rint(256 u_{a-1}),  u_a = 256 u_{a-1} - c_a,
// so u = 2^-7 sum_a s_a + remainder with slices s_a = c_a 2^(-8 (a - 1)), |c_a| <= 128 (at
// most 8 significant bits: exact in BF16) and |remainder| <= 2^(-8 S - 1). A product of two
// values is 2^-14 sum_{a,b} s_a s_b. The slice products of orders a + b = 2, 3 and 4 are
// summed in FP32 accumulators of their own, which stay exact: order d's terms are integer
// multiples of 2^(-8 (d - 2)) of magnitude <= 2^(14 - 8 (d - 2)), so the sum stays within
// 24 bits while (#products per K index) * K * 2^14 <= 2^24. The products of orders
// 5 .. S + 1 share a fourth accumulator whose FP32 rounding is 2^-24 of a sum 2^-24 below
// the total. Orders above S + 1 are dropped. The sums are converted to FP64 and added there.
//
// MFMA 16x16x32 BF16 operand layout (gfx950), lane l = r + 16 g:
//     A (16 x 32): lane holds A[r][8 g + j], j = 0..7
//     B (32 x 16): lane holds B[8 g + j][r]
//     D (16 x 16): lane holds D[4 g + i][r], i = 0..3
#include "common.h"

namespace ssn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int SPLIT_S = 5;          // slices per value (39 bits below the group maximum)
constexpr int SPLIT_THREADS = 512;
constexpr int SPLIT_WAVES = SPLIT_THREADS / WAVE;
constexpr int SG_ROWS = 32;         // rows per Gram stage (one MFMA K step)
// LDS length of a slice column (Gram) or row (recovery) of 32 bf16: 80 bytes, so that the 16-byte
// operand reads of 8 consecutive lanes (8 columns or rows) fall in distinct banks (a 64-byte
// stride put every other lane on the same banks)
constexpr int SL_LEN = 40;
constexpr int NACC = 4;             // accumulators: orders 2, 3, 4, then 5 .. S + 1

__device__ __forceinline__ f4 mfma_bf16(bf16x8 a, bf16x8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f4 zero_f4() { f4 z = {0.f, 0.f, 0.f, 0.f}; return z; }

__host__ __device__ constexpr int acc_of(int a, int b) { return a + b <= 4 ? a + b - 2 : 3; }

// exponent E with |v| < 2^E for every |v| <= vmax (0 for vmax == 0)
__device__ __forceinline__ int scale_exponent(double vmax) {
    int e = 0;
    (void)frexp(vmax, &e);
    return vmax > 0.0 ? e : 0;
}

// Cuts 8 values (times `scale`) into S BF16 slices packed as MFMA operands. The slices
// are exact in BF16, so the BF16 bits are the upper half of the FP32 bits (no rounding
// step); two slices share one 32-bit word (v_perm_b32).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <int S>
__device__ __forceinline__ void cut8(const double (&v)[8], double scale, bf16x8 (&out)[S]) {
    unsigned int bits[S][8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        double x = v[j] * scale * 128.0;
        float w = 1.0f;
#pragma unroll
        for (int a = 0; a < S; ++a) {
            const double q = rint(x);
            bits[a][j] = __float_as_uint((float)q * w);    // c_a 2^(-8 a'), exact
            x = (x - q) * 256.0;
            w *= 0x1p-8f;
        }
    }
#pragma unroll
    for (int a = 0; a < S; ++a) {
        u32x4 wv;
#pragma unroll
        for (int p = 0; p < 4; ++p) wv[p] = __builtin_amdgcn_perm(bits[a][2 * p + 1], bits[a][2 * p], 0x07060302u);
        out[a] = __builtin_bit_cast(bf16x8, wv);
    }
}

// the accumulator sums as one FP64 value, times 2^-14 (smallest first)
__device__ __forceinline__ double reassemble(const f4 (&acc)[NACC], int i) {
    return ((((double)acc[3][i] + (double)acc[2][i]) + (double)acc[1][i]) + (double)acc[0][i]) * 0x1p-14;
}

// multiply-accumulate all slice products of order <= S + 1 into the three accumulators
template <int S>
__device__ __forceinline__ void mfma_products(const bf16x8 (&x)[S], const bf16x8 (&y)[S], f4 (&acc)[NACC]) {
#pragma unroll
    for (int a = 1; a <= S; ++a)
#pragma unroll
        for (int b = 1; b <= S + 1 - a; ++b)
            acc[acc_of(a, b)] = mfma_bf16(x[a - 1], y[b - 1], acc[acc_of(a, b)]);
}

// ---------------------------------------------------------------- split Gram
// One workgroup per (row chunk, tile group, target). Pass 1 takes the maximum of
// |A[:, c]| and |D[:, c]| over the chunk (one shared exponent per column, so that
// A^T A + D^T D and A^T D share their scale 2^(E_i + E_j)). Pass 2 streams the chunk in
// stages of 32 rows: every thread cuts 8 rows of one column into S slices and stores
// them column-major in LDS, [slice][A|D][col][32 rows, padded to SL_LEN], so that an MFMA operand is one
// 16-byte read. The tile group's waves own fixed output tiles (upper tiles of
// A^T A + D^T D, all tiles of A^T D). The order-4 accumulator of an A^T A + D^T D tile
// takes 6 products per row, so every 160 rows the accumulators are folded into FP64.
// The FP64 partial is written in the layout of the FP64 Gram, whose reduction finishes it.
template <int NT, int G>
struct SGShape {
    static constexpr int MP = NT * 16;
    static constexpr int NCR = NT * (NT + 1) / 2;
    static constexpr int NAD = NT * NT;
    static constexpr int NTOT = NCR + NAD;
    static constexpr int MAXT = (NTOT + G * SPLIT_WAVES - 1) / (G * SPLIT_WAVES);
    static constexpr int UNITS = 2 * MP * (SG_ROWS / 8);        // (A|D, col, 8-row group)
    static constexpr int UPT = (UNITS + SPLIT_THREADS - 1) / SPLIT_THREADS;
    static constexpr int SLICE = 2 * MP * SL_LEN;               // bf16 per slice
    static constexpr int FOLD = 160 / SG_ROWS;                  // stages between FP64 folds
};

__device__ __forceinline__ void sg_tile(int t, int nt, int ncr, int& I, int& J, bool& is_cr) {
    if (t < ncr) {
        is_cr = true;
        int col = 0, rem = t;
        while (rem > col) { rem -= col + 1; ++col; }
        I = rem; J = col;
    } else {
        is_cr = false;
        const int u = t - ncr;
        I = u / nt; J = u % nt;
    }
}

template <int NT, int G, int S>
__global__ __launch_bounds__(SPLIT_THREADS)
void split_gram_kernel(const double* __restrict__ T, long long sT_b, long long sT_r,
                       int n, int m, int chunk_rows, int chunks, double* __restrict__ partial)
{
    using Sh = SGShape<NT, G>;
    constexpr int MP = Sh::MP;
    __shared__ __align__(16) __bf16 sl[S * Sh::SLICE];
    __shared__ double colmax[4][MP];
    __shared__ double colscale[MP];

    // the G tile groups of a chunk read the same rows: when the grid's rows of workgroups are whole
    // multiples of 8 G, block x = 8 G u + 8 p + v (v < 8) is group p of chunk 8 u + v, so the groups
    // of a chunk are 8 blocks apart and run on the same XCD (workgroups go to the 8 XCDs round-robin)
    // and the later reads of a row hit that XCD's L2
    int chunk, grp;
    if (G > 1 && gridDim.x % (8 * G) == 0) {
        const int x = blockIdx.x, u = x / (8 * G), rem = x - u * 8 * G;
        grp = rem / 8;
        chunk = 8 * u + (rem - grp * 8);
    } else {
        chunk = blockIdx.x / G; grp = blockIdx.x % G;
    }
    const int b = blockIdx.y;
    const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int r = lane & 15, g = lane >> 4;
    const int row_begin = chunk * chunk_rows;
    const int row_end = min(n, row_begin + chunk_rows);
    const double* Tb = T + (long long)b * sT_b;

    // pass 1: per-column maximum over the chunk, A and D together
    {
        const int c = tid % 128, ro = tid / 128;
        double mx = 0.0;
        if (c < m) {
            // 8 rows per iteration: 16 independent loads in flight, row offsets uniform
            int row = row_begin + ro;
            const double* pa = Tb + (long long)row * sT_r + c;
            const long long dn = (long long)n * sT_r, st4 = 4 * sT_r;
            for (; row + 28 < row_end; row += 32, pa += 8 * st4) {
                double v[16];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    v[2 * u] = pa[u * st4];
                    v[2 * u + 1] = pa[u * st4 + dn];
                }
#pragma unroll
                for (int u = 0; u < 16; ++u) mx = fmax(mx, fabs(v[u]));
            }
            for (; row < row_end; row += 4, pa += st4) mx = fmax(mx, fmax(fabs(pa[0]), fabs(pa[dn])));
        }
        if (c < MP) colmax[ro][c] = mx;
    }
    __syncthreads();
    for (int c = tid; c < MP; c += SPLIT_THREADS) {
        const double mx = fmax(fmax(colmax[0][c], colmax[1][c]), fmax(colmax[2][c], colmax[3][c]));
        colscale[c] = ldexp(1.0, -scale_exponent(mx));
    }
    __syncthreads();

    int tI[Sh::MAXT], tJ[Sh::MAXT];
    bool tcr[Sh::MAXT], tok[Sh::MAXT];
#pragma unroll
    for (int s = 0; s < Sh::MAXT; ++s) {
        const int t = grp + G * (wave + s * SPLIT_WAVES);
        tok[s] = t < Sh::NTOT;
        sg_tile(tok[s] ? t : 0, NT, Sh::NCR, tI[s], tJ[s], tcr[s]);
    }
    f4 acc[Sh::MAXT][NACC];
    d4 acc64[Sh::MAXT];
#pragma unroll
    for (int s = 0; s < Sh::MAXT; ++s) {
        acc64[s] = zero4();
#pragma unroll
        for (int d = 0; d < NACC; ++d) acc[s][d] = zero_f4();
    }
    auto fold = [&]() {
#pragma unroll
        for (int s = 0; s < Sh::MAXT; ++s) {
#pragma unroll
            for (int i = 0; i < 4; ++i) acc64[s][i] += reassemble(acc[s], i);
#pragma unroll
            for (int d = 0; d < NACC; ++d) acc[s][d] = zero_f4();
        }
    };

    // staging unit u -> (8-row group g8, column c, half h), g8 fastest: a wave's slice
    // stores are 1 KB contiguous and its loads read 16 columns x 4 rows per instruction
    double pre[Sh::UPT][8];
    auto load = [&](int r0) {
#pragma unroll
        for (int k = 0; k < Sh::UPT; ++k) {
            const int u = tid + k * SPLIT_THREADS;
            const int g8 = u % 4, c = (u / 4) % MP, h = u / (4 * MP);
            const int first = r0 + 8 * g8;
            const int cnt = (u < Sh::UNITS && c < m) ? min(8, row_end - first) : 0;
            const double* p = Tb + (long long)(first + h * n) * sT_r + c;
#pragma unroll
            for (int j = 0; j < 8; ++j) pre[k][j] = j < cnt ? p[j * sT_r] : 0.0;
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int k = 0; k < Sh::UPT; ++k) {
            const int u = tid + k * SPLIT_THREADS;
            if (u >= Sh::UNITS) continue;
            const int g8 = u % 4, c = (u / 4) % MP, h = u / (4 * MP);
            bf16x8 out[S];
            cut8<S>(pre[k], colscale[c], out);
#pragma unroll
            for (int a = 0; a < S; ++a)
                *reinterpret_cast<bf16x8*>(sl + a * Sh::SLICE + (h * MP + c) * SL_LEN + 8 * g8) = out[a];
        }
    };
    auto frag = [&](int col, bf16x8 (&x)[S]) {
#pragma unroll
        for (int a = 0; a < S; ++a)
            x[a] = *reinterpret_cast<const bf16x8*>(sl + a * Sh::SLICE + (col + r) * SL_LEN + 8 * g);
    };

    const int nstages = (row_end - row_begin + SG_ROWS - 1) / SG_ROWS;
    if (nstages > 0) load(row_begin);
    for (int st = 0; st < nstages; ++st) {
        store();
        __syncthreads();
        if (st + 1 < nstages) load(row_begin + (st + 1) * SG_ROWS);
#pragma unroll
        for (int s = 0; s < Sh::MAXT; ++s) {
            if (!tok[s]) continue;
            const int I = tI[s] * 16, J = tJ[s] * 16;
            bf16x8 x[S], y[S];
            frag(I, x);                               // A columns I
            frag((tcr[s] ? 0 : MP) + J, y);           // A columns J (Cr) or D columns J (A^T D)
            mfma_products<S>(x, y, acc[s]);
            if (tcr[s]) {
                frag(MP + I, x);                      // D columns I and J
                frag(MP + J, y);
                mfma_products<S>(x, y, acc[s]);
            }
        }
        if ((st + 1) % Sh::FOLD == 0 || st + 1 == nstages) fold();
        __syncthreads();
    }

    double* out = partial + ((long long)b * chunks + chunk) * 2 * MP * MP;
#pragma unroll
    for (int s = 0; s < Sh::MAXT; ++s) {
        if (!tok[s]) continue;
        double* o = out + (tcr[s] ? 0 : MP * MP);
        const int j = tJ[s] * 16 + r;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = tI[s] * 16 + 4 * g + i;
            // undo the column scales: 2^(E_row + E_j)
            o[row * MP + j] = acc64[s][i] / (colscale[row] * colscale[j]);
        }
    }
}

// ---------------------------------------------------------------- split recovery
// X = W M with W = [A | D] (n x 2 mp, row r of A next to row r of D) and
// M = [[P^T, -Q^T], [Q^T, P^T]] (2 mp x 2 mp). Each row of W shares one exponent (its row
// maximum), each column of M one exponent (prepared with M's slices by
// split_recovery_fragments). The order-4 sum takes 3 products per K index and
// K = 2 mp <= 256, so one FP32 pass is exact.
//
// Fragments: frag[b] = [S][KS][2 NT][64 lanes] bf16x8 (B operands of every (k step,
// column tile)), then colpow[b] = [2 mp] doubles 2^(F_c).
template <int S>
__host__ __device__ __forceinline__ long long split_frag_doubles(int mp) {
    // slices: S * (2 mp)^2 bf16 = S * mp^2 doubles; then 2 mp column powers
    return (long long)S * mp * mp + 2 * mp;
}

template <int S>
__global__ void split_recovery_fragments(const double* __restrict__ P, const double* __restrict__ Q, int mp,
                                         double* __restrict__ frag)
{
    // one workgroup per (column tile J, target); 256 threads
    const int J = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
    const int K = 2 * mp, KS = K / 32, NJ = K / 16;
    const double* Pb = P + (long long)b * mp * mp;
    const double* Qb = Q + (long long)b * mp * mp;
    __bf16* fb = reinterpret_cast<__bf16*>(frag + (long long)b * split_frag_doubles<S>(mp));
    double* colpow = frag + (long long)b * split_frag_doubles<S>(mp) + (long long)S * mp * mp;
    __shared__ double red[16][17];
    __shared__ double csc[16];
    auto Mv = [&](int k, int c) -> double {
        const bool kd = k >= mp, cb = c >= mp;
        const int kk = kd ? k - mp : k, cc = cb ? c - mp : c;
        // [[P^T, -Q^T], [Q^T, P^T]]
        if (!kd) return cb ? -Qb[cc * mp + kk] : Pb[cc * mp + kk];
        return cb ? Pb[cc * mp + kk] : Qb[cc * mp + kk];
    };
    {
        const int c = tid & 15, kq = tid >> 4;     // 16 columns x 16 k-strides
        double mx = 0.0;
        for (int k = kq; k < K; k += 16) mx = fmax(mx, fabs(Mv(k, 16 * J + c)));
        red[kq][c] = mx;
    }
    __syncthreads();
    if (tid < 16) {
        double mx = 0.0;
        for (int kq = 0; kq < 16; ++kq) mx = fmax(mx, red[kq][tid]);
        const int e = scale_exponent(mx);
        csc[tid] = ldexp(1.0, -e);
        colpow[16 * J + tid] = ldexp(1.0, e);
    }
    __syncthreads();
    for (int u = tid; u < KS * 64; u += blockDim.x) {
        const int ks = u / 64, lane = u % 64, r = lane & 15, g = lane >> 4;
        double v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = Mv(32 * ks + 8 * g + j, 16 * J + r);
        bf16x8 out[S];
        cut8<S>(v, csc[r], out);
#pragma unroll
        for (int a = 0; a < S; ++a)
            *reinterpret_cast<bf16x8*>(fb + (((long long)(a * KS + ks) * NJ + J) * 64 + lane) * 8) = out[a];
    }
}

// Workgroup of 8 waves over ROWS rows: the row exponents are taken cooperatively (TPR
// threads per row), then for every K step the workgroup cuts its ROWS x 32 values into
// slices in LDS once (double-buffered, one barrier per step) and every wave multiplies two
// row blocks of 16 by its TPW column tiles, so each B fragment it loads serves 32 rows.
// Waves: RBP row-block pairs x CG column groups (NT <= 4: 4 x 2, else 2 x 4).
template <int NT>
struct SRShape {
    static constexpr int MP = NT * 16;
    static constexpr int CG = NT <= 4 ? 2 : 4;
    static constexpr int RBP = SPLIT_WAVES / CG;
    static constexpr int ROWS = 32 * RBP;
    static constexpr int TPW = (2 * NT + CG - 1) / CG;   // column tiles per wave
    static constexpr int KS = 2 * MP / 32;               // K steps
    static constexpr int TPR = SPLIT_THREADS / ROWS;     // threads per row (exponents)
    static constexpr int SEGLEN = 2 * MP / TPR;          // values per thread (exponents)
    static constexpr int UNITS = ROWS * 4;               // (row, 8-value group) per K step
};

template <int NT, int S>
__global__ __launch_bounds__(SPLIT_THREADS)
void split_recover_kernel(const double* __restrict__ T, long long sT_b, long long sT_r,
                          const double* __restrict__ frag, int n, int m,
                          double* __restrict__ X, long long sX_b, long long sX_r)
{
    using Sh = SRShape<NT>;
    constexpr int MP = Sh::MP, KS = Sh::KS, NJ = 2 * NT, ROWS = Sh::ROWS;
    __shared__ __align__(16) __bf16 as_lds[2][S][ROWS * SL_LEN];
    __shared__ double rscale[ROWS];
    __shared__ int rexp[ROWS];
    const int b = blockIdx.y, tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int r = lane & 15, g = lane >> 4;
    const int row0 = blockIdx.x * ROWS;
    const double* Tb = T + (long long)b * sT_b;
    const __bf16* fb = reinterpret_cast<const __bf16*>(frag + (long long)b * split_frag_doubles<S>(MP));
    const double* colpow = frag + (long long)b * split_frag_doubles<S>(MP) + (long long)S * MP * MP;

    // the values W[row][k0 .. k0 + len) (one part: k0 and the run stay in A or in D), zero
    // outside the target, through one base pointer
    auto wrun = [&](int row, int k0, int len, double* v) {
        const int part = k0 >= MP ? 1 : 0, c0 = k0 - part * MP;
        const int cnt = row < n ? min(len, m - c0) : 0;
        const double* p = Tb + (long long)(row + part * n) * sT_r + c0;
        for (int j = 0; j < len; ++j) v[j] = j < cnt ? p[j] : 0.0;
    };
    // row exponents: TPR consecutive threads share a row, each takes SEGLEN values
    {
        const int lr = tid / Sh::TPR, seg = tid % Sh::TPR;
        double mx = 0.0;
        double v[Sh::SEGLEN];
        wrun(row0 + lr, seg * Sh::SEGLEN, Sh::SEGLEN, v);
#pragma unroll
        for (int j = 0; j < Sh::SEGLEN; ++j) mx = fmax(mx, fabs(v[j]));
#pragma unroll
        for (int o = 1; o < Sh::TPR; o <<= 1) mx = fmax(mx, __shfl_xor(mx, o));
        if (seg == 0) {
            const int e = scale_exponent(mx);
            rexp[lr] = e;
            rscale[lr] = ldexp(1.0, -e);
        }
    }
    __syncthreads();

    const int rbp = wave / Sh::CG, cg = wave % Sh::CG;
    f4 acc[2][Sh::TPW][NACC];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int t = 0; t < Sh::TPW; ++t)
#pragma unroll
            for (int d = 0; d < NACC; ++d) acc[h][t][d] = zero_f4();

    // cutting unit: (row, 8 consecutive k of the step)
    auto cut_step = [&](int ks, int buf) {
        for (int u = tid; u < Sh::UNITS; u += SPLIT_THREADS) {
            const int lr = u / 4, g8 = u % 4;
            double v[8];
            wrun(row0 + lr, 32 * ks + 8 * g8, 8, v);
            bf16x8 out[S];
            cut8<S>(v, rscale[lr], out);
#pragma unroll
            for (int a = 0; a < S; ++a)
                *reinterpret_cast<bf16x8*>(&as_lds[buf][a][lr * SL_LEN + 8 * g8]) = out[a];
        }
    };

#pragma unroll 1
    for (int ks = 0; ks < KS; ++ks) {
        const int buf = ks & 1;
        cut_step(ks, buf);
        __syncthreads();          // slices of step ks ready; step ks - 1's buffer no longer read
        bf16x8 x0[S], x1[S];
#pragma unroll
        for (int a = 0; a < S; ++a) {
            x0[a] = *reinterpret_cast<const bf16x8*>(&as_lds[buf][a][(rbp * 32 + r) * SL_LEN + 8 * g]);
            x1[a] = *reinterpret_cast<const bf16x8*>(&as_lds[buf][a][(rbp * 32 + 16 + r) * SL_LEN + 8 * g]);
        }
#pragma unroll
        for (int t = 0; t < Sh::TPW; ++t) {
            const int J = cg * Sh::TPW + t;
            if (J >= NJ) continue;
            bf16x8 y[S];
#pragma unroll
            for (int a = 0; a < S; ++a)
                y[a] = *reinterpret_cast<const bf16x8*>(fb + (((long long)(a * KS + ks) * NJ + J) * 64 + lane) * 8);
            mfma_products<S>(x0, y, acc[0][t]);
            mfma_products<S>(x1, y, acc[1][t]);
        }
    }
    double* Xb = X + (long long)b * sX_b;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int lb = rbp * 32 + 16 * h;                   // local first row of the block
#pragma unroll
        for (int t = 0; t < Sh::TPW; ++t) {
            const int J = cg * Sh::TPW + t;
            if (J >= NJ) continue;
            const int c = J * 16 + r;
            const bool bottom = c >= MP;
            const int cc = bottom ? c - MP : c;
            const double cp = colpow[c];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int lr = lb + 4 * g + i, rr = row0 + lr;
                const double v = reassemble(acc[h][t], i);
                if (rr < n && cc < m)
                    Xb[(long long)(rr + (bottom ? n : 0)) * sX_r + cc] = ldexp(v, rexp[lr]) * cp;
            }
        }
    }
}

// ---------------------------------------------------------------- launchers
int split_gram_groups(int nt) { return nt <= 3 ? 1 : (nt <= 7 ? 2 : 4); }

hipError_t launch_split_gram(hipStream_t stream, const double* T, long long sT_b, long long sT_r,
                             int batch, int n, int m, int chunks, double* partial)
{
    const int mp = pad16(m), nt = mp / 16;
    const int G = split_gram_groups(nt);
    int chunk_rows = (n + chunks - 1) / chunks;
    chunk_rows = (chunk_rows + SG_ROWS - 1) / SG_ROWS * SG_ROWS;
    dim3 grid(chunks * G, batch), block(SPLIT_THREADS);
#define SSN_SG_CASE(NTV, GV) \
    case NTV: hipLaunchKernelGGL((split_gram_kernel<NTV, GV, SPLIT_S>), grid, block, 0, stream, \
                                 T, sT_b, sT_r, n, m, chunk_rows, chunks, partial); break;
    switch (nt) {
        SSN_SG_CASE(1, 1) SSN_SG_CASE(2, 1) SSN_SG_CASE(3, 1) SSN_SG_CASE(4, 2)
        SSN_SG_CASE(5, 2) SSN_SG_CASE(6, 2) SSN_SG_CASE(7, 2) SSN_SG_CASE(8, 4)
        default: return LAUNCH_UNSUPPORTED;
    }
#undef SSN_SG_CASE
    return hipGetLastError();
}

long long split_fragment_doubles(int m) { return split_frag_doubles<SPLIT_S>(pad16(m)); }

hipError_t launch_split_recover(hipStream_t stream, const double* T, long long sT_b, long long sT_r,
                                const double* P, const double* Q, double* frag, int batch, int n, int m,
                                double* X, long long sX_b, long long sX_r)
{
    const int mp = pad16(m), nt = mp / 16;
    hipLaunchKernelGGL((split_recovery_fragments<SPLIT_S>), dim3(2 * nt, batch), dim3(256), 0, stream,
                       P, Q, mp, frag);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
#define SSN_SR_CASE(NTV) \
    case NTV: hipLaunchKernelGGL((split_recover_kernel<NTV, SPLIT_S>), \
                                 dim3((n + SRShape<NTV>::ROWS - 1) / SRShape<NTV>::ROWS, batch), \
                                 dim3(SPLIT_THREADS), 0, stream, T, sT_b, sT_r, frag, n, m, X, sX_b, sX_r); break;
    switch (nt) {
        SSN_SR_CASE(1) SSN_SR_CASE(2) SSN_SR_CASE(3) SSN_SR_CASE(4)
        SSN_SR_CASE(5) SSN_SR_CASE(6) SSN_SR_CASE(7) SSN_SR_CASE(8)
        default: return LAUNCH_UNSUPPORTED;
    }
#undef SSN_SR_CASE
    return hipGetLastError();
}

}  // namespace ssn
