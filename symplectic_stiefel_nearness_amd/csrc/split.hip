// Split-precision Gram and recovery: FP64 operands cut into BF16 slices, multiplied on
// the BF16 MFMA (v_mfma_f32_16x16x32_bf16) with FP32 accumulation, reassembled to FP64.
//
// Slicing (Ozaki-style, exact): a value x in a group that shares the power-of-two scale
// 2^E (max |x| < 2^E) is written u = x 2^-E in (-1, 1) and cut as
//     c_1 = rint(128 u),  u_1 = 128 u - c_1,  c_a = rint(256 u_{a-1}),  u_a = 256 u_{a-1} - c_a,
// so u = sum_a c_a 2^(-7 - 8 (a - 1)) + remainder, |remainder| <= 2^(-8 S + 1) / 2, with every
// |c_a| <= 128: each c_a is an integer of at most 8 significant bits, exact in BF16. The
// product of slices a and b is an integer of magnitude <= 2^14 with weight
// 2^(-14 - 8 (a + b - 2)); products of one order d = a + b are summed in their own FP32
// accumulator, which stays an exact integer while (d - 1) * K * 2^14 < 2^24 (K = summed
// length). Products of order d > S + 1 are dropped (below the truncation level). The
// order sums are converted to FP64, weighted and added: the only rounding happens in FP64.
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

__device__ __forceinline__ f4 mfma_bf16(bf16x8 a, bf16x8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f4 zero_f4() { f4 z = {0.f, 0.f, 0.f, 0.f}; return z; }

// weight of the order-d sum (d = a + b >= 2): 2^(-14 - 8 (d - 2))
__device__ __forceinline__ double order_weight(int d) { return ldexp(1.0, -14 - 8 * (d - 2)); }

// exponent E with |v| < 2^E for every |v| <= vmax (0 for vmax == 0)
__device__ __forceinline__ int scale_exponent(double vmax) {
    int e = 0;
    (void)frexp(vmax, &e);
    return vmax > 0.0 ? e : 0;
}

template <int S>
__device__ __forceinline__ void cut(double u, float (&c)[S]) {
    double v = u * 128.0;
#pragma unroll
    for (int a = 0; a < S; ++a) {
        const double q = rint(v);
        c[a] = (float)q;
        v = (v - q) * 256.0;
    }
}

// rows of K summed exactly in one FP32 order accumulator whose largest order takes
// C * S products per K index (C = 1, or 2 for the two chains of A^T A + D^T D)
template <int S, int C>
struct SplitExact {
    static constexpr int K = ((1 << 24) / (C * S * (1 << 14))) / 32 * 32;   // S = 5: 192 / 96
};

// ---------------------------------------------------------------- split Gram
// One workgroup per (row chunk, tile group, target). Pass 1 takes the maximum of
// |A[:, c]| and |D[:, c]| over the chunk (one shared exponent per column, so that
// A^T A + D^T D and A^T D share their scale 2^(E_i + E_j)). Pass 2 streams the chunk in
// stages of 32 rows: every thread cuts 8 rows of one column into S slices and stores
// them column-major in LDS, [slice][A|D][col][32 rows], so that an MFMA operand is one
// 16-byte read. The tile group's waves own fixed output tiles (upper tiles of
// A^T A + D^T D, all tiles of A^T D) with one FP32 accumulator per order; every
// SplitExact<S>::K rows the orders are folded into FP64 accumulators. The FP64 partial is
// written in the layout of the FP64 Gram so the same reduction kernel finishes it.
template <int NT, int G>
struct SGShape {
    static constexpr int MP = NT * 16;
    static constexpr int NCR = NT * (NT + 1) / 2;
    static constexpr int NAD = NT * NT;
    static constexpr int NTOT = NCR + NAD;
    static constexpr int MAXT = (NTOT + G * SPLIT_WAVES - 1) / (G * SPLIT_WAVES);
    static constexpr int UNITS = 2 * MP * (SG_ROWS / 8);        // (A|D, col, 8-row group)
    static constexpr int UPT = (UNITS + SPLIT_THREADS - 1) / SPLIT_THREADS;
    static constexpr int SLICE = 2 * MP * SG_ROWS;              // bf16 per slice
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

    const int chunk = blockIdx.x / G, grp = blockIdx.x % G, b = blockIdx.y;
    const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int r = lane & 15, g = lane >> 4;
    const int row_begin = chunk * chunk_rows;
    const int row_end = min(n, row_begin + chunk_rows);
    const double* Tb = T + (long long)b * sT_b;

    // pass 1: per-column maximum over the chunk, A and D together
    {
        const int c = tid % 128, ro = tid / 128;
        double mx = 0.0;
        if (c < m)
            for (int row = row_begin + ro; row < row_end; row += 4)
                mx = fmax(mx, fmax(fabs(Tb[(long long)row * sT_r + c]), fabs(Tb[(long long)(row + n) * sT_r + c])));
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
    f4 acc[Sh::MAXT][S];
    d4 acc64[Sh::MAXT];
#pragma unroll
    for (int s = 0; s < Sh::MAXT; ++s) {
        acc64[s] = zero4();
#pragma unroll
        for (int d = 0; d < S; ++d) acc[s][d] = zero_f4();
    }
    auto fold = [&]() {
#pragma unroll
        for (int s = 0; s < Sh::MAXT; ++s) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                double v = 0.0;
#pragma unroll
                for (int d = S - 1; d >= 0; --d) v += (double)acc[s][d][i] * order_weight(d + 2);
                acc64[s][i] += v;
            }
#pragma unroll
            for (int d = 0; d < S; ++d) acc[s][d] = zero_f4();
        }
    };

    // staging unit u -> (half, column, 8-row group); lanes run along columns (coalesced)
    double pre[Sh::UPT][8];
    auto load = [&](int r0) {
#pragma unroll
        for (int k = 0; k < Sh::UPT; ++k) {
            const int u = tid + k * SPLIT_THREADS;
            const int c = u % MP, rest = u / MP, g8 = rest % 4, h = rest / 4;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int row = r0 + 8 * g8 + j;
                double v = 0.0;
                if (u < Sh::UNITS && c < m && row < row_end) v = Tb[(long long)(row + h * n) * sT_r + c];
                pre[k][j] = v;
            }
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int k = 0; k < Sh::UPT; ++k) {
            const int u = tid + k * SPLIT_THREADS;
            if (u >= Sh::UNITS) continue;
            const int c = u % MP, rest = u / MP, g8 = rest % 4, h = rest / 4;
            const double sc = colscale[c];
            bf16x8 out[S];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                float cs[S];
                cut<S>(pre[k][j] * sc, cs);
#pragma unroll
                for (int a = 0; a < S; ++a) out[a][j] = (__bf16)cs[a];
            }
#pragma unroll
            for (int a = 0; a < S; ++a)
                *reinterpret_cast<bf16x8*>(sl + a * Sh::SLICE + (h * MP + c) * SG_ROWS + 8 * g8) = out[a];
        }
    };

    const int nstages = (row_end - row_begin + SG_ROWS - 1) / SG_ROWS;
    constexpr int FOLD = SplitExact<S, 2>::K / SG_ROWS;
    if (nstages > 0) load(row_begin);
    for (int st = 0; st < nstages; ++st) {
        store();
        __syncthreads();
        if (st + 1 < nstages) load(row_begin + (st + 1) * SG_ROWS);
#pragma unroll
        for (int s = 0; s < Sh::MAXT; ++s) {
            if (!tok[s]) continue;
            const int I = tI[s] * 16, J = tJ[s] * 16;
            // operand fragments of this tile: A columns I (and J for Cr), D columns J (and I)
            bf16x8 aI[S], xJ[S];
#pragma unroll
            for (int a = 0; a < S; ++a) {
                aI[a] = *reinterpret_cast<const bf16x8*>(sl + a * Sh::SLICE + (I + r) * SG_ROWS + 8 * g);
                xJ[a] = *reinterpret_cast<const bf16x8*>(sl + a * Sh::SLICE + ((tcr[s] ? 0 : MP) + J + r) * SG_ROWS + 8 * g);
            }
#pragma unroll
            for (int d = 2; d <= S + 1; ++d)
#pragma unroll
                for (int a = 1; a < d; ++a) {
                    const int bb = d - a;
                    if (a > S || bb > S) continue;
                    acc[s][d - 2] = mfma_bf16(aI[a - 1], xJ[bb - 1], acc[s][d - 2]);
                }
            if (tcr[s]) {
                bf16x8 dI[S], dJ[S];
#pragma unroll
                for (int a = 0; a < S; ++a) {
                    dI[a] = *reinterpret_cast<const bf16x8*>(sl + a * Sh::SLICE + (MP + I + r) * SG_ROWS + 8 * g);
                    dJ[a] = *reinterpret_cast<const bf16x8*>(sl + a * Sh::SLICE + (MP + J + r) * SG_ROWS + 8 * g);
                }
#pragma unroll
                for (int d = 2; d <= S + 1; ++d)
#pragma unroll
                    for (int a = 1; a < d; ++a) {
                        const int bb = d - a;
                        if (a > S || bb > S) continue;
                        acc[s][d - 2] = mfma_bf16(dI[a - 1], dJ[bb - 1], acc[s][d - 2]);
                    }
            }
        }
        if ((st + 1) % FOLD == 0 || st + 1 == nstages) fold();
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
// M = [[P^T, -Q^T], [Q^T, P^T]] (2 mp x 2 mp). Each row of W shares one exponent (row
// maximum, taken in registers), each column of M one exponent (prepared with M's slices by
// split_recovery_fragments). K = 2 mp is summed in segments of at most SplitExact<S>::K.
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
        bf16x8 out[S];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            float cs[S];
            cut<S>(Mv(32 * ks + 8 * g + j, 16 * J + r) * csc[r], cs);
#pragma unroll
            for (int a = 0; a < S; ++a) out[a][j] = (__bf16)cs[a];
        }
#pragma unroll
        for (int a = 0; a < S; ++a)
            *reinterpret_cast<bf16x8*>(fb + (((long long)(a * KS + ks) * NJ + J) * 64 + lane) * 8) = out[a];
    }
}

// Workgroup: 8 waves = RB row blocks of 16 rows x CG column groups (NT <= 6: 4 x 2,
// else 2 x 4); a wave owns up to ceil(2 NT / CG) column tiles of its 16 rows.
template <int NT>
struct SRShape {
    static constexpr int MP = NT * 16;
    static constexpr int CG = NT <= 6 ? 2 : 4;
    static constexpr int RB = SPLIT_WAVES / CG;
    static constexpr int ROWS = 16 * RB;
    static constexpr int TPW = (2 * NT + CG - 1) / CG;   // column tiles per wave
    static constexpr int KS = 2 * MP / 32;               // K steps
};

template <int NT, int S>
__global__ __launch_bounds__(SPLIT_THREADS)
void split_recover_kernel(const double* __restrict__ T, long long sT_b, long long sT_r,
                          const double* __restrict__ frag, int n, int m,
                          double* __restrict__ X, long long sX_b, long long sX_r)
{
    using Sh = SRShape<NT>;
    constexpr int MP = Sh::MP, KS = Sh::KS, NJ = 2 * NT;
    constexpr int SEG = SplitExact<S, 1>::K / 32;          // K steps per exact segment
    const int b = blockIdx.y;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int r = lane & 15, g = lane >> 4;
    const int rb = wave / Sh::CG, cg = wave % Sh::CG;
    const int row = blockIdx.x * Sh::ROWS + rb * 16 + r;
    const bool row_ok = row < n;
    const double* Tb = T + (long long)b * sT_b;
    const __bf16* fb = reinterpret_cast<const __bf16*>(frag + (long long)b * split_frag_doubles<S>(MP));
    const double* colpow = frag + (long long)b * split_frag_doubles<S>(MP) + (long long)S * MP * MP;

    // the 8 values of W[row][32 ks + 8 g .. + 7] (one part: A if k < MP else D)
    auto wvals = [&](int ks, double (&v)[8]) {
        const int k0 = 32 * ks + 8 * g;
        const int part = k0 / MP, c0 = k0 - part * MP;
        const double* src = Tb + (long long)(row + part * n) * sT_r + c0;
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = (row_ok && c0 + j < m) ? src[j] : 0.0;
    };
    // row maximum: this lane's k values, then the 4 lanes of the row (lanes r + 16 g)
    double mx = 0.0;
#pragma unroll 1
    for (int ks = 0; ks < KS; ++ks) {
        double v[8];
        wvals(ks, v);
#pragma unroll
        for (int j = 0; j < 8; ++j) mx = fmax(mx, fabs(v[j]));
    }
    mx = fmax(mx, __shfl_xor(mx, 16));
    mx = fmax(mx, __shfl_xor(mx, 32));
    const int erow = scale_exponent(mx);
    const double rsc = ldexp(1.0, -erow);

    // K <= one exact segment (mp <= 96): the order sums go straight to the output
    constexpr bool ONESEG = KS <= SEG;
    f4 acc[Sh::TPW][S];
    d4 acc64[ONESEG ? 1 : Sh::TPW];
#pragma unroll
    for (int t = 0; t < (ONESEG ? 1 : Sh::TPW); ++t) acc64[t] = zero4();
#pragma unroll
    for (int t = 0; t < Sh::TPW; ++t) {
#pragma unroll
        for (int d = 0; d < S; ++d) acc[t][d] = zero_f4();
    }
#pragma unroll 1
    for (int ks = 0; ks < KS; ++ks) {
        double v[8];
        wvals(ks, v);
        bf16x8 as[S];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            float cs[S];
            cut<S>(v[j] * rsc, cs);
#pragma unroll
            for (int a = 0; a < S; ++a) as[a][j] = (__bf16)cs[a];
        }
#pragma unroll
        for (int t = 0; t < Sh::TPW; ++t) {
            const int J = cg * Sh::TPW + t;
            if (J >= NJ) continue;
            bf16x8 bs[S];
#pragma unroll
            for (int a = 0; a < S; ++a)
                bs[a] = *reinterpret_cast<const bf16x8*>(fb + (((long long)(a * KS + ks) * NJ + J) * 64 + lane) * 8);
#pragma unroll
            for (int d = 2; d <= S + 1; ++d)
#pragma unroll
                for (int a = 1; a < d; ++a) {
                    const int bb = d - a;
                    if (a > S || bb > S) continue;
                    acc[t][d - 2] = mfma_bf16(as[a - 1], bs[bb - 1], acc[t][d - 2]);
                }
        }
        if (!ONESEG && ((ks + 1) % SEG == 0 || ks + 1 == KS)) {
#pragma unroll
            for (int t = 0; t < (ONESEG ? 1 : Sh::TPW); ++t) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    double s_ = 0.0;
#pragma unroll
                    for (int d = S - 1; d >= 0; --d) s_ += (double)acc[t][d][i] * order_weight(d + 2);
                    acc64[t][i] += s_;
                }
#pragma unroll
                for (int d = 0; d < S; ++d) acc[t][d] = zero_f4();
            }
        }
    }
    // lane holds rows 4 g + i of the block, column 16 J + r: exponents of those rows
    int erows[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) erows[i] = __shfl(erow, 4 * g + i);
    double* Xb = X + (long long)b * sX_b;
    const int row0 = blockIdx.x * Sh::ROWS + rb * 16;
#pragma unroll
    for (int t = 0; t < Sh::TPW; ++t) {
        if (cg * Sh::TPW + t >= NJ) continue;
        const int c = (cg * Sh::TPW + t) * 16 + r;
        const bool bottom = c >= MP;
        const int cc = bottom ? c - MP : c;
        const double cp = colpow[c];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int rr = row0 + 4 * g + i;
            double v;
            if (ONESEG) {
                v = 0.0;
#pragma unroll
                for (int d = S - 1; d >= 0; --d) v += (double)acc[t][d][i] * order_weight(d + 2);
            } else {
                v = acc64[ONESEG ? 0 : t][i];
            }
            if (rr < n && cc < m)
                Xb[(long long)(rr + (bottom ? n : 0)) * sX_r + cc] = ldexp(v, erows[i]) * cp;
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
        default: return hipErrorInvalidValue;
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
        default: return hipErrorInvalidValue;
    }
#undef SSN_SR_CASE
    return hipGetLastError();
}

}  // namespace ssn
