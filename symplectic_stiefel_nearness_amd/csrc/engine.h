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
// Layout. Small m x m matrices live in a per-target global workspace (L2 / Infinity
// Cache resident), padded to mp = 16 NT with zero padding. Every product is an FP64
// MFMA GEMM whose operands are staged through LDS in 16-wide K panels (double
// buffered); the 4 SIMDs of the CU each own one quadrant of the output tiles and the
// two waves of a SIMD split that quadrant's tile rows, so every wave keeps a register
// block of accumulators and each operand fragment feeds several MFMAs. The Cholesky
// factor of I + trial^2 and its inverse stay in LDS as packed 16 x 16 lower blocks
// (blocked right-looking factorisation; MFMA panel, trailing and inverse updates).
// The eigendecomposition of the scaled Cr runs first, in its own kernel
// (cyclic Jacobi with the matrix, and the eigenvectors when they fit, in LDS).
// All reductions are deterministic.
//
// The CG iterates r, z, p, d and the gradient are exactly skew (built from skew()
// outputs and from H = Z - Z^T, whose two halves are negated copies), so the
// reference's preconditioner skew(R / den) equals R / den bit for bit; it is computed
// that way.
#pragma once
#include "common.h"
#include "uneven.h"

// element-wise updates mirror the reference's explicitly rounded (non-contracted) ops
#pragma clang fp contract(off)

// Tuning of the chained CG products: k-blocks of A fragments in flight in the second product,
// and whether each k-block is fenced from the scheduler (keeps the loads from being hoisted).
#ifndef SSN_CG_PREFETCH
#define SSN_CG_PREFETCH 3
#endif
#ifndef SSN_QUAD_SCHED_BARRIER
#define SSN_QUAD_SCHED_BARRIER 1
#endif
#ifndef SSN_QUAD_PREFETCH
#define SSN_QUAD_PREFETCH 2     // k-steps of operand fragments in flight in the quadrant products
#endif
#ifndef SSN_CG_SCHED_BARRIER
#define SSN_CG_SCHED_BARRIER 1
#endif

namespace ssn {

constexpr int DW = 8;                 // waves per target
constexpr int DT = DW * WAVE;         // threads per target
constexpr int MAX_CG = 150;           // dual_recommended_hvp.py:16 (the adjoint's cap; the solver's come in Caps)
constexpr int MAX_SWEEPS = 40;
constexpr int NCOUNT = 12;            // counters per target

// workspace slots, each mp*mp doubles
enum Slot {
    S_V = 0, S_CIE, S_JSE, S_W, S_P, S_Q, S_SR, S_SI, S_G, S_D, S_R, S_Z0, S_PP, S_HP,
    S_U, S_VV, S_T1, S_T2, S_T3, S_T4, S_T5, S_QT, S_COUNT
};
static_assert(S_G == 8 && S_D == 9 && S_R == 10 && S_Z0 == 11, "slot numbers used by capi.hip");

struct DualArgs {
    const double* Cr; const double* Ci; long long sC_b; int ldc;
    int m, mp;
    double* omega; long long sO_b; int ldo;      // m x m output (original basis)
    double* Pf; double* Qf;                       // [b][mp][mp] final P, Q (original basis)
    long long* iters; int* status; int* counters; // counters [b][12] (include/ssn.h)
    double* ws;                                   // [b][S_COUNT][mp*mp] + [b][mp] values
    int v_in_lds;
    Caps caps;                                    // iteration caps and tolerance
};

template <int NT>
struct Cfg {
    static constexpr int MP = 16 * NT;
    static constexpr int KP = (NT % 2 == 0 && NT <= 6) ? 32 : 16;  // K panel width (divides MP)
    static constexpr int AS = KP + 1;             // A panel row stride (doubles)
    static constexpr int BS = MP + 2;             // B panel row stride
    static constexpr int PANEL = MP * AS + KP * BS;
    static constexpr int PER = 2 * KP * MP / DT;  // staged elements per thread and panel
    static constexpr int DEPTH = KP == 32 ? 1 : 2;   // panels loaded ahead (a 32-wide panel
                                                     // computes long enough to cover a load)
    static constexpr int H = (NT + 1) / 2;        // quadrant split
    static constexpr int MAXC = H;                // tile columns per wave
    static constexpr int MAXR = (H + 1) / 2;      // tile rows per wave
    static constexpr int NBLK = NT * (NT + 1) / 2;   // packed lower 16x16 blocks
    // LDS layout (doubles): [0,48) scratch; two K panels; packed blocks; block-row temp; lambda
    static constexpr int OFF_PANEL = 48;
    static constexpr int OFF_BLK = OFF_PANEL + 2 * PANEL;
    static constexpr int OFF_ROW = OFF_BLK + NBLK * 256;
    static constexpr int OFF_LAM = OFF_ROW + (NT > 1 ? NT - 1 : 1) * 256;
    static constexpr int LDS_END = OFF_LAM + MP;
};

// LDS scratch [0, 48): two alternating reduction buffers of 16 doubles, then flags.
constexpr int RED_BUFS = 2, RED_SLOTS = 16, FLAG_JACOBI = 32, FLAG_CHOL = 33;

struct Ctx {
    int tid, wave, lane, r, q;
    int m;
    double* red;                 // LDS scratch
    mutable int phase;           // which reduction buffer the next sum uses
};

__device__ __forceinline__ void sync() { __syncthreads(); }

// An opaque copy of a per-lane index. The engine inlines dozens of GEMM / element-wise
// sites into one loop nest; without this, the compiler hoists every site's per-lane
// address arithmetic out of the solver loops, keeps it all live and spills.
__device__ __forceinline__ int opaque(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

__device__ __forceinline__ double readlane(double v, int lane) { return readlane_d(v, lane); }

// Deterministic workgroup sums with one barrier: the two reduction buffers alternate,
// and a wave cannot reuse a buffer before every wave has passed the barrier of the
// sum in between. Every thread returns the same values.
template <int WAVES>
__device__ void block_sum2(const Ctx& c, double& a, double& b) {
    a = wave_sum(a);
    b = wave_sum(b);
    double* buf = c.red + RED_SLOTS * c.phase;
    c.phase ^= 1;
    if (c.lane == 0) { buf[c.wave] = a; buf[8 + c.wave] = b; }
    sync();
    double s = 0.0, t = 0.0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) { s += buf[w]; t += buf[8 + w]; }
    a = s; b = t;
}

__device__ __forceinline__ void wg_sum2(const Ctx& c, double& a, double& b) { block_sum2<DW>(c, a, b); }

__device__ __forceinline__ double wg_sum(const Ctx& c, double v) {
    double z = 0.0;
    block_sum2<DW>(c, v, z);
    return v;
}

// ---------------------------------------------------------------- tile ownership
// SIMD s = wave & 3 owns quadrant (s >> 1, s & 1) of the NT x NT tile grid; the two
// waves of a SIMD (w and w + 4) split the quadrant's tile rows.
struct WT { int r0, nr, c0, nc; };

template <int NT>
__device__ __forceinline__ WT wave_tiles(int wave) {
    constexpr int H = Cfg<NT>::H;
    const int simd = wave & 3, half = wave >> 2;
    const int qr0 = (simd >> 1) ? H : 0, qr1 = (simd >> 1) ? NT : H;
    const int qc0 = (simd & 1) ? H : 0, qc1 = (simd & 1) ? NT : H;
    const int rows = qr1 - qr0, split = (rows + 1) / 2;
    WT w;
    w.c0 = qc0; w.nc = qc1 - qc0;
    w.r0 = half ? qr0 + split : qr0;
    w.nr = half ? rows - split : split;
    if (w.nr <= 0 || w.nc <= 0) { w.nr = 0; w.nc = 0; }
    return w;
}

template <int NT>
struct Acc {
    d4 v[Cfg<NT>::MAXR][Cfg<NT>::MAXC];
    __device__ __forceinline__ void zero() {
#pragma unroll
        for (int i = 0; i < Cfg<NT>::MAXR; ++i)
#pragma unroll
            for (int j = 0; j < Cfg<NT>::MAXC; ++j) v[i][j] = zero4();
    }
};

// Epilogue over this wave's tiles: f(row, col, value); lane holds rows q + 4e, column r.
template <int NT, class F>
__device__ __forceinline__ void for_tiles(const Ctx& c, const WT& w, const Acc<NT>& acc, F f) {
    const int q = opaque(c.q), r = opaque(c.r);
#pragma unroll
    for (int i = 0; i < Cfg<NT>::MAXR; ++i)
#pragma unroll
        for (int j = 0; j < Cfg<NT>::MAXC; ++j)
            if (i < w.nr && j < w.nc) {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    f(16 * (w.r0 + i) + q + 4 * e, 16 * (w.c0 + j) + r, acc.v[i][j][e]);
            }
}

// ---------------------------------------------------------------- staged GEMM
// acc += sum over terms t of op(A_t) op(B_t), K = MP per term. Element sources:
//   ga(t, i, k) -> op(A_t)[i][k],  gb(t, k, j) -> op(B_t)[k][j]  (global memory).
// TA / TB say which index is contiguous in memory, for coalesced staging.
template <int NT, bool TA, bool TB, class GA, class GB>
__device__ __forceinline__ void gemm_acc(const Ctx& c, double* lds, int nterms, GA ga, GB gb, Acc<NT>& acc, const WT& w) {
    using C = Cfg<NT>;
    constexpr int MP = C::MP, KP = C::KP, PER = C::PER;
    constexpr int PPT = MP / KP;            // panels per term
    // when the A part of a panel is a whole number of thread rows, element s of every
    // thread is an A or a B element independently of the thread
    constexpr bool SPLIT_STATIC = (KP * MP) % DT == 0;
    double* buf0 = lds + C::OFF_PANEL;
    constexpr int DEPTH = C::DEPTH;
    double pre[DEPTH][PER];                 // panels in flight in registers
    const int npanels = nterms * PPT;
    // opaque copies of the lane indices keep the per-thread staging offsets of the many
    // inlined GEMM sites from being hoisted out of the solver loops
    const int tid = opaque(c.tid);
    const int lr = opaque(c.r), lq = opaque(c.q);
    auto load = [&](int p, double (&dst)[PER]) {
        const int t = p / PPT, k0 = (p - t * PPT) * KP;
#pragma unroll
        for (int s = 0; s < PER; ++s) {
            const int e = tid + s * DT;
            double v;
            if (SPLIT_STATIC ? (s * DT < KP * MP) : (e < KP * MP)) {   // A panel element
                int i, kk;
                if (TA) { kk = e / MP; i = e - kk * MP; } else { i = e / KP; kk = e - i * KP; }
                v = ga(t, i, k0 + kk);
            } else {                         // B panel element
                const int f = e - KP * MP;
                int kk, j;
                if (TB) { j = f / KP; kk = f - j * KP; } else { kk = f / MP; j = f - kk * MP; }
                v = gb(t, k0 + kk, j);
            }
            dst[s] = v;
        }
    };
    auto store = [&](const double (&src)[PER], double* dst) {
#pragma unroll
        for (int s = 0; s < PER; ++s) {
            const int e = tid + s * DT;
            if (SPLIT_STATIC ? (s * DT < KP * MP) : (e < KP * MP)) {
                int i, kk;
                if (TA) { kk = e / MP; i = e - kk * MP; } else { i = e / KP; kk = e - i * KP; }
                dst[i * C::AS + kk] = src[s];
            } else {
                const int f = e - KP * MP;
                int kk, j;
                if (TB) { j = f / KP; kk = f - j * KP; } else { kk = f / MP; j = f - kk * MP; }
                dst[MP * C::AS + kk * C::BS + j] = src[s];
            }
        }
    };
    sync();    // prior writes of the operands (and prior panel reads) are complete
#pragma unroll
    for (int d = 0; d < DEPTH; ++d) if (d < npanels) load(d, pre[d]);
#pragma unroll 1
    for (int p = 0; p < npanels; ++p) {
        // buffer (p & 1) was last read by panel p - 2, before every wave passed the
        // barrier of panel p - 1
        double* cur = buf0 + (p & 1) * C::PANEL;
        const int slot = DEPTH == 1 ? 0 : (p & 1);
        if (slot) store(pre[DEPTH - 1], cur); else store(pre[0], cur);
        sync();
        if (p + DEPTH < npanels) { if (slot) load(p + DEPTH, pre[DEPTH - 1]); else load(p + DEPTH, pre[0]); }
        const double* As = cur;
        const double* Bs = cur + MP * C::AS;
#pragma unroll
        for (int k = 0; k < KP; k += 4) {
            double a[C::MAXR], b[C::MAXC];
#pragma unroll
            for (int i = 0; i < C::MAXR; ++i)
                a[i] = (i < w.nr) ? As[(16 * (w.r0 + i) + lr) * C::AS + k + lq] : 0.0;
#pragma unroll
            for (int j = 0; j < C::MAXC; ++j)
                b[j] = (j < w.nc) ? Bs[(k + lq) * C::BS + 16 * (w.c0 + j) + lr] : 0.0;
#pragma unroll
            for (int i = 0; i < C::MAXR; ++i)
#pragma unroll
                for (int j = 0; j < C::MAXC; ++j)
                    if (i < w.nr && j < w.nc) acc.v[i][j] = mfma(a[i], b[j], acc.v[i][j]);
        }
    }
}

// C = op(A) op(B) with global row-major MP x MP operands; epilogue f(row, col, value).
template <int NT, bool TA, bool TB, class F>
__device__ __forceinline__ void gemm(const Ctx& c, double* lds, const double* A, const double* B, F f) {
    constexpr int MP = Cfg<NT>::MP;
    const WT w = wave_tiles<NT>(c.wave);
    Acc<NT> acc; acc.zero();
    gemm_acc<NT, TA, TB>(c, lds, 1,
        [&](int, int i, int k) { return TA ? A[k * MP + i] : A[i * MP + k]; },
        [&](int, int k, int j) { return TB ? B[j * MP + k] : B[k * MP + j]; }, acc, w);
    for_tiles<NT>(c, w, acc, f);
    sync();    // epilogue results are visible to every wave
}

// C = A1 B1 + s A2 B2 (non-transposed operands), epilogue f.
template <int NT, class F>
__device__ __forceinline__ void gemm2(const Ctx& c, double* lds, const double* A1, const double* B1,
                      const double* A2, const double* B2, double s, F f) {
    constexpr int MP = Cfg<NT>::MP;
    const WT w = wave_tiles<NT>(c.wave);
    Acc<NT> acc; acc.zero();
    gemm_acc<NT, false, false>(c, lds, 2,
        [&](int t, int i, int k) { return t ? s * A2[i * MP + k] : A1[i * MP + k]; },
        [&](int t, int k, int j) { return t ? B2[k * MP + j] : B1[k * MP + j]; }, acc, w);
    for_tiles<NT>(c, w, acc, f);
    sync();
}

// C = A1 B1^T + s A2 B2^T, epilogue f.
template <int NT, class F>
__device__ __forceinline__ void gemm2_nt(const Ctx& c, double* lds, const double* A1, const double* B1,
                         const double* A2, const double* B2, double s, F f) {
    constexpr int MP = Cfg<NT>::MP;
    const WT w = wave_tiles<NT>(c.wave);
    Acc<NT> acc; acc.zero();
    gemm_acc<NT, false, true>(c, lds, 2,
        [&](int t, int i, int k) { return t ? s * A2[i * MP + k] : A1[i * MP + k]; },
        [&](int t, int k, int j) { return t ? B2[j * MP + k] : B1[j * MP + k]; }, acc, w);
    for_tiles<NT>(c, w, acc, f);
    sync();
}

// ---------------------------------------------------------------- element-wise helper
template <int NT, class F>
__device__ __forceinline__ void each(const Ctx& c, F f) {
    constexpr int MP = Cfg<NT>::MP;
#pragma unroll 2
    for (int e = opaque(c.tid); e < MP * MP; e += DT) f(e / MP, e % MP, e);
}


// Element pass over this thread's elements e = tid + s DT (the order of `each`) that issues the
// loads of a group of elements before any of the group's stores: ld(e, v) loads into v,
// op(i, j, e, v) computes and stores. The opaque thread index keeps the element addresses
// inside the solver loops.
template <int NT, int NV, class LD, class OP>
__device__ __forceinline__ void each_loaded(const Ctx& c, LD ld, OP op) {
    constexpr int MP = Cfg<NT>::MP, PE = (MP * MP + DT - 1) / DT;
    constexpr int GB = PE < 6 ? PE : 6;              // elements per group of loads
    const int t0 = opaque(c.tid);
#pragma unroll
    for (int g = 0; g < PE; g += GB) {
        double v[GB][NV];
#pragma unroll
        for (int u = 0; u < GB; ++u) {
            const int e = t0 + (g + u) * DT;
            if (g + u < PE && (PE * DT == MP * MP || e < MP * MP)) ld(e, v[u]);
        }
#pragma unroll
        for (int u = 0; u < GB; ++u) {
            const int e = t0 + (g + u) * DT;
            if (g + u < PE && (PE * DT == MP * MP || e < MP * MP)) op(e / MP, e % MP, e, v[u]);
        }
    }
}

// 1 / sqrt(d): hardware estimate refined by two Newton steps (a few ulp).
__device__ __forceinline__ double rsqrt_refined(double d) {
    double y = __builtin_amdgcn_rsq(d);
    y = y * (1.5 - 0.5 * d * y * y);
    y = y * (1.5 - 0.5 * d * y * y);
    return y;
}

// ---------------------------------------------------------------- Jacobi eigensolver
// Cyclic two-sided Jacobi with the round-robin (circle) tournament: each round applies
// m/2 disjoint rotations in parallel, on the fixed slot pairs (s, m-1-s); between rounds
// the player in slot j moves to slot next(j) = j + 1 over 1..m-1 (slot 0 fixed, m-1 -> 1).
//
// A is kept in SLOT order, upper triangle only, as 2x2 pair blocks: block (P, P2), P <= P2,
// holds rows (P, m-1-P) x columns (P2, m-1-P2), its top row pair in one double2 array and
// its bottom row pair in another. Two buffers alternate: a round reads every block of one
// and writes each rotated element straight to its place in the other after the slot move.
// Since the move is the same every round, every read and write address is fixed for the
// whole solve and computed once. Wave 0 owns the diagonal blocks; during round r it also
// recomputes, from round r's inputs and with the same operations, the three elements each
// pair of round r + 1 needs, and writes round r + 1's rotation parameters. One barrier per
// round therefore suffices.
//
// V is kept in PLAYER order and updated in place (slot x holds player 0 if x = 0, else
// 1 + ((x - 1 - r) mod (m - 1)) at round r of a sweep). After m - 1 rounds every player
// is back in its own slot, so at sweep boundaries slot, player and column agree.
//
// A rotation is applied when |a_pq| > eps sqrt(|a_pp a_qq|); the iteration stops after
// the first sweep at whose end no pair exceeds that threshold.
// Eigenvector replay (eigh_replay_body, eigh_replay_follow) and its hand-off from the
// eigensolver: rows per replay workgroup, rounds staged per LDS chunk, and the progress word
// (per target, in extra slot EX_PROGRESS: rounds recorded so far, PROGRESS_FINAL once every
// round and the eigenvalue ranks are in the record).
constexpr int RP_ROWS = 8;          // rows (waves) per workgroup of eigh_replay_kernel
constexpr int RP_CHUNK = 16;        // rounds staged per LDS chunk (and published per step)
constexpr int MAX_ORDER_PADDED = 128; // largest padded order (NT = 8)
constexpr int EX_PROGRESS = 8;      // extra slot: progress word (unsigned long long)
constexpr int EX_REPLAY_FAILED = 9; // extra slot: nonzero when a following replay gave up
constexpr unsigned long long PROGRESS_FINAL = 1ull << 40;

constexpr int JW = 64;        // wave 0: diagonal blocks and rotation parameters
constexpr int JU = 512 - JW;  // threads that update off-diagonal blocks and V
constexpr int JB_MAX = 5;     // off-diagonal blocks per update thread: ceil(64*63/2 / 448)
constexpr int JB_BATCH = 3;   // blocks gathered before their stores
constexpr int JV_SMALL = 11;  // eigenvector rows per update thread, one batch (kernel for m <= 98)
constexpr int JV_HALF = 10;   // rows per batch, two batches: 20 >= ceil(128 / (448 / 64))
constexpr double JAC_EPS = 2.220446049250313e-16;

__host__ __device__ __forceinline__ int jac_blocks(int m) { return (m / 2) * (m / 2 + 1) / 2; }
// LDS doubles of the Jacobi area: two A buffers (4 nb each), two parameter buffers
// ([cos, sin] double2 x np, tan x np, pad np), then m ints of ranks
__host__ __device__ __forceinline__ int jac_area(int m) { return 8 * jac_blocks(m) + 8 * (m / 2) + m / 2 + 2; }

__device__ __forceinline__ int player(int x, int r, int m) {
    if (x == 0) return 0;
    int y = x - 1 - r;
    if (y < 0) y += m - 1;
    return 1 + y;
}

// Rotation (cos, sin, tan) that annihilates a_pq: tan = sign(theta) / (|theta| + sqrt(theta^2 + 1)),
// theta = (a_qq - a_pp) / (2 a_pq), evaluated without dividing by a_pq.
__device__ __forceinline__ void jac_param(double app, double aqq, double apq, double& cs, double& sn, double& t) {
    cs = 1.0; sn = 0.0; t = 0.0;
    if (apq != 0.0 && fabs(apq) > JAC_EPS * sqrt(fabs(app) * fabs(aqq))) {
        const double d = aqq - app, e = 2.0 * apq;
        const double den = fabs(d) + sqrt(fma(d, d, e * e));
        t = (den > 0.0) ? e / den : 1.0;
        if (d < 0.0) t = -t;
        cs = rsqrt_refined(fma(t, t, 1.0));
        sn = t * cs;
    }
}

// G1^T X G2 for the block X = [x00 x01; x10 x11] (rows rotated by r1, columns by r2).
__device__ __forceinline__ void jac_block(double2 r1, double2 r2, double x00, double x01, double x10, double x11,
                                          double& z00, double& z01, double& z10, double& z11) {
    const double y00 = fma(r2.x, x00, -(r2.y * x01)), y01 = fma(r2.y, x00, r2.x * x01);
    const double y10 = fma(r2.x, x10, -(r2.y * x11)), y11 = fma(r2.y, x10, r2.x * x11);
    z00 = fma(r1.x, y00, -(r1.y * y10)); z10 = fma(r1.y, y00, r1.x * y10);
    z01 = fma(r1.x, y01, -(r1.y * y11)); z11 = fma(r1.y, y01, r1.x * y11);
}

// Runs the sweeps on the slot-ordered A in J (buffer 0 filled by the caller) and the
// player-ordered V (m x m, row stride LV, initialised here). Returns the sweep count
// (-1 without convergence) and in `fin` the buffer holding the final A.
// REPLAY: V is not formed here; the update threads of slot pair s (the first group) store the
// round's (cos, sin) of pair s at rot[(sweep (m - 1) + round) np + s] for eigh_replay_kernel.
// PUBLISH (with REPLAY): the record is written with agent-scope stores at the start of each
// round, and every RP_CHUNK rounds the first update thread waits for its wave's stores and then
// stores the number of rounds recorded into *progress, for replay workgroups that run at the
// same time (eigh_replay_follow).
template <int JB, bool V_IN_LDS, bool REPLAY = false, bool PUBLISH = false>
__device__ int jacobi(const Ctx& c, double* J, double* V, int LV, int vdummy, int* flag, const double** fin,
                      double2* rot = nullptr, unsigned long long* progress = nullptr) {
    static_assert(!PUBLISH || REPLAY, "publishing needs the rotation record");
    const int m = c.m, np = m / 2, nb = np * (np + 1) / 2;
    double* A0 = J;
    double* A1 = J + 4 * nb;
    double* R0 = J + 8 * nb;
    double* R1 = R0 + 4 * np;
    auto pr = [&](int x) { return x < np ? x : m - 1 - x; };
    auto nx = [&](int x) { return x == 0 ? 0 : (x == m - 1 ? 1 : x + 1); };
    auto pv = [&](int x) { return x == 0 ? 0 : (x == 1 ? m - 1 : x - 1); };
    auto blk = [&](int P, int P2) { return P * np - (P * (P - 1)) / 2 + (P2 - P); };
    auto eoff = [&](int x, int y) {
        int px = pr(x), hx = x >= np ? 1 : 0, py = pr(y), hy = y >= np ? 1 : 0;
        if (px > py || (px == py && hx > hy)) { int u = px; px = py; py = u; u = hx; hx = hy; hy = u; }
        return (hx ? 2 * nb : 0) + 2 * blk(px, py) + hy;
    };
    // value of slot element (x, y) after the round that reads Ao with parameters Ro
    auto rotated = [&](const double* Ao, const double* Ro, int x, int y) -> double {
        int px = pr(x), hx = x >= np ? 1 : 0, py = pr(y), hy = y >= np ? 1 : 0;
        if (px > py || (px == py && hx > hy)) { int u = px; px = py; py = u; u = hx; hx = hy; hy = u; }
        const double2* cs = reinterpret_cast<const double2*>(Ro);
        const int b = blk(px, py);
        const double2 tp = reinterpret_cast<const double2*>(Ao)[b];
        if (px == py) {
            if (hx != hy) return cs[px].y != 0.0 ? 0.0 : tp.y;
            const double t = Ro[2 * np + px];
            return hx ? fma(t, tp.y, Ao[2 * nb + 2 * b + 1]) : fma(-t, tp.y, tp.x);
        }
        const double2 bt = reinterpret_cast<const double2*>(Ao + 2 * nb)[b];
        double z00, z01, z10, z11;
        jac_block(cs[px], cs[py], tp.x, tp.y, bt.x, bt.y, z00, z01, z10, z11);
        return hx ? (hy ? z11 : z10) : (hy ? z01 : z00);
    };

    if constexpr (!REPLAY) {
        for (int e = c.tid; e < m * m; e += 512) {
            const int i = e / m, j = e - (e / m) * m;
            V[i * LV + j] = (i == j) ? 1.0 : 0.0;
        }
    }
    auto big = [&](double v, double di, double dj) {
        return v != 0.0 && fabs(v) > JAC_EPS * sqrt(fabs(di) * fabs(dj));
    };
    // The two roles run separate copies of the sweep loop, so that each keeps only its own
    // state in registers. Both pass the same barriers in the same order: one per round and
    // one per sweep, after which the sweep's convergence flag (two slots, alternating) is
    // read. Slot s & 1 is cleared by thread 0 at the start of sweep s and set during the
    // check at its end; the previous reads of that slot lie at least one barrier behind.
    int sweeps = -1, cur = 0;
    if (c.wave == 0) {
        // wave 0: diagonal block S and the slots whose elements pair S needs next round
        const bool dg = c.tid < np;
        const int S = dg ? c.tid : 0, Sq = m - 1 - S;
        const int drd = 2 * blk(S, S);
        const int dw0 = eoff(nx(S), nx(S)), dw1 = eoff(nx(S), nx(Sq)), dw2 = eoff(nx(Sq), nx(Sq));
        const int lx = pv(S), ly = pv(Sq);
        // A0 is ready (written by the caller before a barrier): parameters of the first round
        if (dg) {
            const double2 tp = reinterpret_cast<const double2*>(A0)[drd / 2];
            double cs, sn, t;
            jac_param(tp.x, A0[2 * nb + drd + 1], tp.y, cs, sn, t);
            reinterpret_cast<double2*>(R0)[S] = make_double2(cs, sn);
            R0[2 * np + S] = t;
        }
        sync();
        for (int sweep = 0; sweep < MAX_SWEEPS; ++sweep) {
            if (c.tid == 0) flag[sweep & 1] = 0;
#pragma unroll 1
            for (int round = 0; round < m - 1; ++round) {
                const double* Ao = cur ? A1 : A0;
                double* An = cur ? A0 : A1;
                const double* Ro = cur ? R1 : R0;
                double* Rn = cur ? R0 : R1;
                if (dg) {
                    const double2* cso = reinterpret_cast<const double2*>(Ro);
                    const double2 tp = *reinterpret_cast<const double2*>(Ao + drd);
                    const double d2 = Ao[2 * nb + drd + 1];
                    const double t = Ro[2 * np + S], sn = cso[S].y;
                    const double app = rotated(Ao, Ro, lx, lx), aqq = rotated(Ao, Ro, ly, ly);
                    const double apq = rotated(Ao, Ro, lx, ly);
                    double cn, sn2, tn;
                    jac_param(app, aqq, apq, cn, sn2, tn);
                    reinterpret_cast<double2*>(Rn)[S] = make_double2(cn, sn2);
                    Rn[2 * np + S] = tn;
                    An[dw0] = fma(-t, tp.y, tp.x);
                    An[dw1] = sn != 0.0 ? 0.0 : tp.y;
                    An[dw2] = fma(t, tp.y, d2);
                }
                sync();
                cur ^= 1;
            }
            const double* Af = cur ? A1 : A0;
            if (dg && big(Af[drd + 1], Af[drd], Af[2 * nb + drd + 1])) flag[sweep & 1] = 1;
            sync();
            if (!flag[sweep & 1]) { sweeps = sweep + 1; break; }
        }
    } else {
        // update threads: fixed off-diagonal blocks (P < P2) and one slot pair of V
        const int ut = c.tid - JW;
        static_assert(JB <= JB_MAX, "block list");
        int rd[JB], wr[JB][4], bP[JB], bQ[JB];
#pragma unroll
        for (int k = 0; k < JB; ++k) {
            const int e = ut + k * JU;
            bP[k] = -1; bQ[k] = 0; rd[k] = 0;
            wr[k][0] = wr[k][1] = wr[k][2] = wr[k][3] = 0;
            if (e < nb - np) {
                int P = 0, start = 0, len = np - 1;
                while (e >= start + len) { start += len; ++P; --len; }
                const int P2 = P + 1 + (e - start), q = m - 1 - P, q2 = m - 1 - P2;
                bP[k] = P; bQ[k] = P2; rd[k] = 2 * blk(P, P2);
                wr[k][0] = eoff(nx(P), nx(P2)); wr[k][1] = eoff(nx(P), nx(q2));
                wr[k][2] = eoff(nx(q), nx(P2)); wr[k][3] = eoff(nx(q), nx(q2));
            }
        }
        const int vgroups = JU / np;
        const int vS = ut % np, vG = ut / np;
        const bool v_on = vG < vgroups;
        constexpr int JV = JB <= JB_BATCH ? JV_SMALL : JV_HALF, NVB = JB <= JB_BATCH ? 1 : 2;
        sync();
        for (int sweep = 0; sweep < MAX_SWEEPS; ++sweep) {
#pragma unroll 1
            for (int round = 0; round < m - 1; ++round) {
                const double* Ao = cur ? A1 : A0;
                double* An = cur ? A0 : A1;
                const double2* cso = reinterpret_cast<const double2*>(cur ? R1 : R0);
                // V in global memory: the eigenvector rows of this round are requested before
                // the block work, so their latency overlaps it (the columns depend on the round
                // only; the rotation is applied after the blocks)
                double gvv[V_IN_LDS || REPLAY ? 1 : NVB * JV][2];
                int gva[V_IN_LDS || REPLAY ? 1 : NVB * JV];
                int gvp = 0, gvq = 0;
                if constexpr (!V_IN_LDS && !REPLAY) {
                    if (v_on) {
                        const int s_ = opaque(vS), g_ = opaque(vG);
                        gvp = player(s_, round, m); gvq = player(m - 1 - s_, round, m);
#pragma unroll
                        for (int k = 0; k < NVB * JV; ++k) {
                            const int i = g_ + k * vgroups;
                            gva[k] = i < m ? (int)__umul24(i, LV) : vdummy;
                            gvv[k][0] = V[gva[k] + gvp]; gvv[k][1] = V[gva[k] + gvq];
                        }
                    }
                }
                if constexpr (PUBLISH) {
                    // the record first, so that it has long left when the count is published
                    if (ut < np) {
                        const double2 rv = cso[ut];
                        unsigned long long* d = reinterpret_cast<unsigned long long*>(rot + (sweep * (m - 1) + round) * np + ut);
                        __hip_atomic_store(d, (unsigned long long)__double_as_longlong(rv.x), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                        __hip_atomic_store(d + 1, (unsigned long long)__double_as_longlong(rv.y), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
                // Addresses are formed (and made opaque) while gathering; the stores use
                // them as plain registers: an asm statement between stores would force a
                // wait for every outstanding LDS operation.
#pragma unroll
                for (int k0 = 0; k0 < JB; k0 += JB_BATCH) {
                    double2 xt[JB_BATCH], xb[JB_BATCH], r1[JB_BATCH], r2[JB_BATCH];
                    int wa[JB_BATCH][4];
#pragma unroll
                    for (int u = 0; u < JB_BATCH; ++u) {
                        const int k = k0 + u;
                        if (k >= JB || bP[k] < 0) continue;
                        const int o = opaque(rd[k]);
                        xt[u] = *reinterpret_cast<const double2*>(Ao + o);
                        xb[u] = *reinterpret_cast<const double2*>(Ao + 2 * nb + o);
                        r1[u] = cso[opaque(bP[k])];
                        r2[u] = cso[opaque(bQ[k])];
#pragma unroll
                        for (int v = 0; v < 4; ++v) wa[u][v] = opaque(wr[k][v]);
                    }
#pragma unroll
                    for (int u = 0; u < JB_BATCH; ++u) {
                        const int k = k0 + u;
                        if (k >= JB || bP[k] < 0) continue;
                        double z00, z01, z10, z11;
                        jac_block(r1[u], r2[u], xt[u].x, xt[u].y, xb[u].x, xb[u].y, z00, z01, z10, z11);
                        An[wa[u][0]] = z00; An[wa[u][1]] = z01; An[wa[u][2]] = z10; An[wa[u][3]] = z11;
                    }
                }
                if constexpr (PUBLISH) {
                    // every record store of this wave (all of them: np <= 64) completes before
                    // the count that covers it is visible
                    const int done = sweep * (m - 1) + round + 1;
                    if (ut == 0 && done % RP_CHUNK == 0) {
                        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
                        __hip_atomic_store(progress, (unsigned long long)done, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                    }
                } else if constexpr (REPLAY) {
                    if (ut < np) rot[(sweep * (m - 1) + round) * np + ut] = cso[ut];
                } else if (!V_IN_LDS && v_on) {
                    const double2 rv = cso[opaque(vS)];
#pragma unroll
                    for (int k = 0; k < NVB * JV; ++k) {
                        V[gva[k] + gvp] = fma(rv.x, gvv[k][0], -(rv.y * gvv[k][1]));
                        V[gva[k] + gvq] = fma(rv.y, gvv[k][0], rv.x * gvv[k][1]);
                    }
                } else if (v_on) {
                    const int s_ = opaque(vS), g_ = opaque(vG);   // row offsets stay in the loop
                    const int vp = player(s_, round, m), vq = player(m - 1 - s_, round, m);
                    const double2 rv = cso[s_];
                    // batches of rows: all loads of a batch are issued before its stores;
                    // rows past m use the dummy row (no branch)
#pragma unroll
                    for (int h = 0; h < NVB; ++h) {
                        double vv[JV][2];
                        int va[JV];
#pragma unroll
                        for (int k = 0; k < JV; ++k) {
                            const int i = g_ + (h * JV + k) * vgroups;
                            va[k] = i < m ? (int)__umul24(i, LV) : vdummy;
                            vv[k][0] = V[va[k] + vp]; vv[k][1] = V[va[k] + vq];
                        }
#pragma unroll
                        for (int k = 0; k < JV; ++k) {
                            V[va[k] + vp] = fma(rv.x, vv[k][0], -(rv.y * vv[k][1]));
                            V[va[k] + vq] = fma(rv.y, vv[k][0], rv.x * vv[k][1]);
                        }
                    }
                }
                sync();
                cur ^= 1;
            }
            // convergence: does any pair of this thread's blocks exceed the threshold?
            const double* Af = cur ? A1 : A0;
            auto diag = [&](int x) {
                const int bb = blk(pr(x), pr(x));
                return x < np ? Af[2 * bb] : Af[2 * nb + 2 * bb + 1];
            };
            bool any = false;
#pragma unroll
            for (int k = 0; k < JB; ++k) {
                if (bP[k] < 0) continue;
                const int P = bP[k], P2 = bQ[k];
                const double dP = diag(P), dq = diag(m - 1 - P), dP2 = diag(P2), dq2 = diag(m - 1 - P2);
                any = any || big(Af[rd[k]], dP, dP2) || big(Af[rd[k] + 1], dP, dq2) ||
                      big(Af[2 * nb + rd[k]], dq, dP2) || big(Af[2 * nb + rd[k] + 1], dq, dq2);
            }
            if (any) flag[sweep & 1] = 1;
            sync();
            if (!flag[sweep & 1]) { sweeps = sweep + 1; break; }
        }
    }
    *fin = cur ? A1 : A0;
    return sweeps;
}

// ---------------------------------------------------------------- blocked Cholesky / inverse
// Packed lower 16x16 blocks in LDS: block (I, J), J <= I, at index I(I+1)/2 + J, row-major.
__device__ __forceinline__ int bidx(int I, int J) { return I * (I + 1) / 2 + J; }

// One wave: Cholesky of the 16x16 SPD block B (lower part read), replaced in place by
// the inverse X of its factor (lower, zero above). Lane i < 16 holds row i in registers.
// During the factorisation values move across lanes with readlane (to scalar registers: the
// engine has no vector registers to spare for broadcast copies there); for the inverse, L is
// written back to B and read from it (broadcast reads), which takes the readlanes off the
// inverse's dependent chain. Each pivot takes one reciprocal square root
// (hardware estimate plus two Newton steps) instead of a square root and a divide, and the
// inverse is right-looking, so the dependent chains stay short. The updates of a pivot run on
// every lane: lanes i <= j only change entries above the diagonal, which nothing reads, and no
// lane mask has to be formed. Returns false on a non-positive pivot.
__device__ __forceinline__ bool diag_factor_invert(double* B, int lane) {
    double row[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) row[k] = (lane < 16) ? B[lane * 16 + k] : 0.0;
    bool ok = true;
    double rinv[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const double d = readlane(row[j], j);
        if (!(d > 0.0)) ok = false;
        const double y = rsqrt_refined(d);        // 1 / L_jj
        rinv[j] = y;
        if (lane == j) row[j] = d * y;            // L_jj
        else if (lane > j) row[j] = row[j] * y;   // L_ij
#pragma unroll
        for (int k = j + 1; k < 16; ++k) {
            const double lkj = readlane(row[j], k);
            row[k] = row[k] - row[j] * lkj;
        }
    }
    // column `lane` of X = L^-1, right-looking: x_k /= L_kk, then x_i -= L_ik x_k (i > k)
    double x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) x[i] = (lane == i) ? 1.0 : 0.0;
    // L to B (lane i writes row i), then L_ik read from B (one address per read: a broadcast);
    // the wave's LDS operations complete in order, so no barrier is needed
    if (lane < 16) {
#pragma unroll
        for (int k = 0; k < 16; ++k) B[lane * 16 + k] = row[k];
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        x[k] = x[k] * rinv[k];
#pragma unroll
        for (int i = k + 1; i < 16; ++i) x[i] = x[i] - B[i * 16 + k] * x[k];
    }
    if (lane < 16) {
#pragma unroll
        for (int i = 0; i < 16; ++i) B[i * 16 + lane] = x[i];
    }
    return ok;
}

// acc(D layout) -= L_IJ L_KJ^T for one 16x16 block pair, in place on block (I, K).
__device__ __forceinline__ void trailing_block(double* Bk, const double* Li, const double* Lk, int r, int q) {
    d4 acc;
#pragma unroll
    for (int e = 0; e < 4; ++e) acc[e] = Bk[(q + 4 * e) * 16 + r];
    double a[4], b[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) { a[s] = -Li[r * 16 + 4 * s + q]; b[s] = Lk[r * 16 + 4 * s + q]; }
#pragma unroll
    for (int s = 0; s < 4; ++s) acc = mfma(a[s], b[s], acc);
#pragma unroll
    for (int e = 0; e < 4; ++e) Bk[(q + 4 * e) * 16 + r] = acc[e];
}

// Blocked right-looking Cholesky of the packed lower blocks with one-block lookahead
// (wave 0 updates and factors diagonal block J+1 while the other waves finish the
// trailing update of step J), then the blocked inverse R = L^-1 in place (diagonal
// blocks already hold their inverses). Returns validity.
template <int NT>
__device__ bool cholesky_inverse(const Ctx& c, double* lds) {
    using C = Cfg<NT>;
    double* blk = lds + C::OFF_BLK;
    double* tmp = lds + C::OFF_ROW;
    int* okflag = reinterpret_cast<int*>(c.red + FLAG_CHOL);
    const int r = opaque(c.r), q = opaque(c.q);
    bool ok0 = true;                        // wave 0's running validity
    if (c.wave == 0) ok0 = diag_factor_invert(blk, c.lane);
    sync();
#pragma unroll 1
    for (int J = 0; J < NT; ++J) {
        // panel: L_IJ = A_IJ X_JJ^T, I > J
        const double* X = blk + bidx(J, J) * 256;
        for (int I = J + 1 + c.wave; I < NT; I += DW) {
            double* B = blk + bidx(I, J) * 256;
            d4 acc = zero4();
            double a[4], b[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) { a[s] = B[r * 16 + 4 * s + q]; b[s] = X[r * 16 + 4 * s + q]; }
#pragma unroll
            for (int s = 0; s < 4; ++s) acc = mfma(a[s], b[s], acc);
#pragma unroll
            for (int e = 0; e < 4; ++e) B[(q + 4 * e) * 16 + r] = acc[e];
        }
        sync();
        if (J + 1 == NT) break;
        // trailing update A_IK -= L_IJ L_KJ^T, J < K <= I; block (J+1, J+1) goes to wave 0,
        // which then factors it right away; the other waves share the remaining blocks
        if (c.wave == 0) {
            const int I = J + 1;
            trailing_block(blk + bidx(I, I) * 256, blk + bidx(I, J) * 256, blk + bidx(I, J) * 256, r, q);
            ok0 = diag_factor_invert(blk + bidx(I, I) * 256, c.lane) && ok0;
        } else {
            const int rem = NT - 1 - J;
            const int count = rem * (rem + 1) / 2;          // block 0 of the list is (J+1, J+1)
            for (int t = c.wave; t < count; t += DW - 1) {
                int I = J + 1, base = 0;
                while (t >= base + (I - J)) { base += I - J; ++I; }
                const int K = J + 1 + (t - base);
                trailing_block(blk + bidx(I, K) * 256, blk + bidx(I, J) * 256, blk + bidx(K, J) * 256, r, q);
            }
        }
        sync();
    }
    if (c.wave == 0 && c.lane == 0) *okflag = ok0 ? 1 : 0;
    sync();
    const bool valid = *okflag != 0;
    if (!valid) return false;
    // inverse: R_IJ = -X_II sum_{K=J}^{I-1} L_IK R_KJ, block rows top to bottom
#pragma unroll 1
    for (int I = 1; I < NT; ++I) {
        const int J = c.wave;
        if (J < I) {
            d4 S = zero4();
            for (int K = J; K < I; ++K) {
                const double* Lik = blk + bidx(I, K) * 256;
                const double* Rkj = blk + bidx(K, J) * 256;
                double a[4], b[4];
#pragma unroll
                for (int s = 0; s < 4; ++s) { a[s] = Lik[r * 16 + 4 * s + q]; b[s] = Rkj[(4 * s + q) * 16 + r]; }
#pragma unroll
                for (int s = 0; s < 4; ++s) S = mfma(a[s], b[s], S);
            }
            const double* Xi = blk + bidx(I, I) * 256;
            d4 acc = zero4();
#pragma unroll
            for (int s = 0; s < 4; ++s) acc = mfma(-Xi[r * 16 + 4 * s + q], S[s], acc);
#pragma unroll
            for (int e = 0; e < 4; ++e) tmp[J * 256 + (q + 4 * e) * 16 + r] = acc[e];
        }
        sync();
        for (int e = c.tid; e < I * 256; e += DT) blk[bidx(I, 0) * 256 + e] = tmp[e];
        sync();
    }
    return true;
}

// P = R^T R from the packed inverse factor (upper tiles computed, mirrored), written to
// global P (row-major MP x MP) with the padding rows / columns set to zero.
template <int NT>
__device__ void inverse_product(const Ctx& c, const double* lds, double* P) {
    using C = Cfg<NT>;
    constexpr int MP = C::MP;
    const double* blk = lds + C::OFF_BLK;
    const int r = opaque(c.r), q = opaque(c.q), m = c.m;
    for (int t = c.wave; t < C::NBLK; t += DW) {
        int I = 0, J = 0, base = 0;
        while (t >= base + (NT - I)) { base += NT - I; ++I; }
        J = I + (t - base);                       // I <= J
        d4 acc = zero4();
        for (int K = J; K < NT; ++K) {
            const double* Rki = blk + bidx(K, I) * 256;
            const double* Rkj = blk + bidx(K, J) * 256;
            double a[4], b[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) { a[s] = Rki[(4 * s + q) * 16 + r]; b[s] = Rkj[(4 * s + q) * 16 + r]; }
#pragma unroll
            for (int s = 0; s < 4; ++s) acc = mfma(a[s], b[s], acc);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int i = 16 * I + q + 4 * e, j = 16 * J + r;
            const double v = (i < m && j < m) ? acc[e] : 0.0;
            P[i * MP + j] = v;
            if (I != J) P[j * MP + i] = v;
        }
    }
    sync();
}

// ---------------------------------------------------------------- eigensolver kernel
// Per target (one 8-wave workgroup): scale = max(1, tr(Cr)/m) (dual_recommended_hvp.py:106), eigh(Cr / scale)
// (:108) by Jacobi, eigenvalues sorted ascending. Writes the sorted eigenvalues (padding
// 1.0, which keeps lambda_i + lambda_j > 0), the eigenvectors (zero padded, slot S_V),
// scale, the sweep count and the elapsed ticks into the target's workspace.
constexpr int ET = 512;               // threads of the eigensolver kernel
constexpr int EW = ET / WAVE;

struct EighArgs {
    const double* Cr; long long sC_b; int ldc;
    int m, mp;
    double* ws;
    int v_in_lds;
};

// Rotation record of the eigensolver (cos, sin per slot pair and round, MAX_SWEEPS sweeps of
// mp - 1 rounds at most), then the eigenvalue ranks (mp ints), after lambda and the extras.
__host__ __device__ __forceinline__ long long rot_record_doubles(int mp) {
    return (long long)MAX_SWEEPS * (mp - 1) * mp + mp;
}
__host__ __device__ __forceinline__ long long target_doubles(int mp) {
    return (long long)S_COUNT * mp * mp + mp + 16 + rot_record_doubles(mp);
}
__host__ __device__ __forceinline__ long long rot_record_offset(int mp) {
    return (long long)S_COUNT * mp * mp + mp + 16;
}

// Body of target b's eigensolver workgroup (ET threads, dynamic LDS eigh_lds_doubles).
// PUBLISH: the rotation record, the ranks and the count of recorded rounds go out for replay
// workgroups running at the same time (eigh_replay_follow); the progress word in extra slot
// EX_PROGRESS (and EX_REPLAY_FAILED) must have been cleared before the launch.
template <int JB, bool V_IN_LDS, bool REPLAY = false, bool PUBLISH = false>
__device__ __forceinline__ void eigh_body(const EighArgs& a, int b)
{
    extern __shared__ double lds[];
    Ctx c;
    c.tid = threadIdx.x; c.wave = __builtin_amdgcn_readfirstlane(c.tid >> 6); c.lane = c.tid & 63;
    c.r = c.lane & 15; c.q = c.lane >> 4; c.m = a.m; c.red = lds; c.phase = 0;
    const int m = a.m, mp = a.mp;
    const unsigned long long t0 = wall_clock64();
    double* ws = a.ws + (long long)b * target_doubles(mp);
    double* Vout = ws + (long long)S_V * mp * mp;
    double* lam = ws + (long long)S_COUNT * mp * mp;
    double* extra = lam + mp;
    const double* Crb = a.Cr + (long long)b * a.sC_b;

    // LDS: [0, 48) reductions and flags, then the Jacobi area, then V when it fits
    double* J = lds + 48;
    const int np = m / 2, nb = np * (np + 1) / 2;
    int* pairs = reinterpret_cast<int*>(J + 8 * nb + 8 * np);   // m ints: ranks after the sweeps
    int* flag = reinterpret_cast<int*>(c.red + FLAG_JACOBI);
    const int LV = V_IN_LDS ? m + 1 : mp;                      // odd LDS row stride

    if (c.wave == 0) {
        double t = 0.0;
        for (int i = c.lane; i < m; i += WAVE) t += Crb[(long long)i * a.ldc + i];
        t = wave_sum(t);
        if (c.lane == 0) c.red[40] = t;
    }
    sync();
    const double scale = fmax(c.red[40] / m, 1.0);
    for (int e = c.tid; e < m * m; e += ET) {
        const int i = e / m, j = e - (e / m) * m;
        if (i <= j) {
            // slot (= index) pair blocks of the upper triangle; Cr's lower triangle is read,
            // as eigh(UPLO='L') reads it
            const int pi = i < np ? i : m - 1 - i, hi = i >= np ? 1 : 0;
            const int pj = j < np ? j : m - 1 - j, hj = j >= np ? 1 : 0;
            int P = pi, P2 = pj, h1 = hi, h2 = hj;
            if (P > P2 || (P == P2 && h1 > h2)) { P = pj; P2 = pi; h1 = hj; h2 = hi; }
            const int blkidx = P * np - (P * (P - 1)) / 2 + (P2 - P);
            J[(h1 ? 2 * nb : 0) + 2 * blkidx + h2] = Crb[(long long)j * a.ldc + i] / scale;
        }
    }
    sync();
    const double* Af = nullptr;
    // V's dummy row (target of row indices past m): one extra LDS row, or slot T4 below T5
    double* Vj = V_IN_LDS ? J + jac_area(m) : ws + (long long)S_T5 * mp * mp;
    const int vdummy = V_IN_LDS ? m * LV : -(mp * mp);
    double2* rot = reinterpret_cast<double2*>(ws + rot_record_offset(mp));
    int* granks = reinterpret_cast<int*>(ws + rot_record_offset(mp) + (long long)MAX_SWEEPS * (mp - 1) * mp);
    unsigned long long* progress = reinterpret_cast<unsigned long long*>(extra + EX_PROGRESS);
    const int sweeps = jacobi<JB, V_IN_LDS, REPLAY, PUBLISH>(c, J, Vj, LV, vdummy, flag, &Af, rot, progress);
    auto eig = [&](int x) {
        const int px = x < np ? x : m - 1 - x;
        const int bi = 2 * (px * np - (px * (px - 1)) / 2);
        return x < np ? Af[bi] : Af[2 * nb + bi + 1];
    };
    for (int i = c.tid; i < m; i += ET) {
        const double li = eig(i);
        int rank = 0;
        for (int j = 0; j < m; ++j) {
            const double lj = eig(j);
            rank += (lj < li || (lj == li && j < i)) ? 1 : 0;
        }
        pairs[i] = rank;
        if constexpr (REPLAY && !PUBLISH) granks[i] = rank;
    }
    sync();
    if constexpr (PUBLISH) {
        // wave 1 (which wrote the whole record) writes the ranks, waits for all of its stores
        // and publishes the final count
        if (c.wave == 1) {
            for (int i = c.lane; i < m; i += WAVE)
                __hip_atomic_store(granks + i, pairs[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (c.lane == 0) {
                const unsigned long long total = (unsigned long long)(sweeps < 0 ? MAX_SWEEPS : sweeps) * (m - 1);
                __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
                __hip_atomic_store(progress, total | PROGRESS_FINAL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
    for (int i = c.tid; i < mp; i += ET) {
        if (i < m) lam[pairs[i]] = eig(i);
        else lam[i] = 1.0;
    }
    if constexpr (!REPLAY) {
        for (int e = c.tid; e < mp * mp; e += ET) {
            const int i = e / mp, j = e - (e / mp) * mp;
            if (i < m && j < m) Vout[i * mp + pairs[j]] = Vj[i * LV + j];
            else Vout[e] = 0.0;
        }
    }
    if (c.tid == 0) {
        extra[0] = scale;
        extra[1] = (double)sweeps;
        extra[2] = (double)(wall_clock64() - t0);
        extra[3] = V_IN_LDS ? 1.0 : 0.0;
        if constexpr (!PUBLISH) extra[EX_REPLAY_FAILED] = 0.0;   // no following replay to fail
    }
}

template <int JB, bool V_IN_LDS, bool REPLAY = false>
__global__ __launch_bounds__(ET)
void eigh_kernel(EighArgs a)
{
    eigh_body<JB, V_IN_LDS, REPLAY>(a, blockIdx.x);
}

// The eigenvectors from the eigensolver's rotation record. Row i of V only ever changes by the
// column rotations, so every row is an independent replay of the same rounds: one wave per row,
// 8 rows per workgroup, the round parameters staged through LDS for the 8 waves. Lane s holds the
// row's entries at slots s and m - 1 - s (slot order: at round 0 of every sweep slot = column);
// a round applies pair s's rotation with the eigensolver's expressions (the same operations on
// the same values, so V is bitwise the in-kernel one) and then moves the tournament (the entry at
// slot x goes to slot x + 1 over 1 .. m - 1, m - 1 to 1): top entries shift up one lane, bottom
// entries down one lane. At the end the columns are ordered by the eigenvalue ranks.
// Rows [8 blk, 8 blk + 8) of the target whose solver block is ws (order m, padded mp); the whole
// workgroup returns at once when blk is past the target's rows.
__device__ __forceinline__ void eigh_replay_body(const double* ws, int m, int mp, int blk, double2 (*par)[RP_CHUNK * 64])
{
    if (blk * RP_ROWS >= mp) return;
    const int np = m / 2;
    const double2* rot = reinterpret_cast<const double2*>(ws + rot_record_offset(mp));
    const int* granks = reinterpret_cast<const int*>(ws + rot_record_offset(mp) + (long long)MAX_SWEEPS * (mp - 1) * mp);
    const double* extra = ws + (long long)S_COUNT * mp * mp + mp;
    double* Vout = const_cast<double*>(ws) + (long long)S_V * mp * mp;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = blk * RP_ROWS + wave;                         // row of V (uniform per wave)
    const int sw = (int)extra[1];
    const int rounds = (sw < 0 ? MAX_SWEEPS : sw) * (m - 1);
    const int s = lane < np ? lane : np - 1;                    // lanes past np mirror the last pair
    double top = (i == s) ? 1.0 : 0.0, bot = (i == m - 1 - s) ? 1.0 : 0.0;
    const int nchunks = (rounds + RP_CHUNK - 1) / RP_CHUNK;
    auto stage = [&](int ch, int buf) {
        for (int e = threadIdx.x; e < RP_CHUNK * np; e += RP_ROWS * WAVE) {
            const int rr = e / np, ss = e - rr * np;
            const int round = ch * RP_CHUNK + rr;
            par[buf][rr * 64 + ss] = round < rounds ? rot[(long long)round * np + ss] : make_double2(1.0, 0.0);
        }
    };
    // one round: pair rotation, then the tournament move (new top(s) = top(s - 1) for s >= 2,
    // bot(0) for s = 1, top(0) for s = 0; new bot(s) = bot(s + 1) for s < np - 1, top(np - 1) for
    // s = np - 1); lanes 0 and np - 1 are read with readlane, the neighbours with whole-wave DPP
    // shifts (wave_shr:1 / wave_shl:1; the lanes without a source do not use the value)
    auto round_step = [&](double2 rv) {
        const double t = fma(rv.x, top, -(rv.y * bot));
        const double u = fma(rv.y, top, rv.x * bot);
        if (np == 1) { top = t; bot = u; return; }      // m = 2: the two slots stay
        const double up = dpp<0x138, 0xF, 0xF>(t);      // t of lane - 1
        const double dn = dpp<0x130, 0xF, 0xF>(u);      // u of lane + 1
        const double b0 = readlane_d(u, 0);
        const double tl = readlane_d(t, np - 1);
        top = s == 0 ? t : (s == 1 ? b0 : up);
        bot = s == np - 1 ? tl : dn;
    };
    if (nchunks > 0) stage(0, 0);
    for (int ch = 0; ch < nchunks; ++ch) {
        __syncthreads();                                        // chunk ch staged; chunk ch - 1 read
        if (ch + 1 < nchunks) stage(ch + 1, (ch + 1) & 1);
        const double2* pc = par[ch & 1];
        const int nr = min(RP_CHUNK, rounds - ch * RP_CHUNK);
        if (nr == RP_CHUNK) {
            double2 rv[RP_CHUNK];                               // the chunk's parameters in registers
#pragma unroll
            for (int rr = 0; rr < RP_CHUNK; ++rr) rv[rr] = pc[rr * 64 + s];
#pragma unroll
            for (int rr = 0; rr < RP_CHUNK; ++rr) round_step(rv[rr]);
        } else {
            for (int rr = 0; rr < nr; ++rr) round_step(pc[rr * 64 + s]);
        }
    }
    if (i >= mp) return;
    double* row = Vout + (long long)i * mp;
    if (i < m) {
        if (lane < np) {
            row[granks[lane]] = top;
            row[granks[m - 1 - lane]] = bot;
        }
        for (int j = m + lane; j < mp; j += 64) row[j] = 0.0;
    } else {
        for (int j = lane; j < mp; j += 64) row[j] = 0.0;
    }
}


// Rows [32 rb, 32 rb + 32) of target V's replay while the target's eigensolver workgroup is
// still rotating (the workgroups of one cooperative launch, eigh_coop_kernel). Each wave holds
// four rows (the lanes as in eigh_replay_body, with the same expressions per row, so V is
// bitwise the other replay's). Before staging a chunk of RP_CHUNK rounds, thread 0 polls the
// target's progress word (agent-scope loads, with a sleep between polls) until the chunk is
// recorded or the final count is out; the record and the ranks are read with agent-scope
// loads. A wait that exceeds `spin_limit` polls marks the target failed (extra slot
// EX_REPLAY_FAILED, which the dual engine turns into ST_EIGH) and ends the workgroup, so
// every wait ends. Returns false on that failure.
constexpr int RF_ROWS_PER_WAVE = 4;
constexpr int RF_ROWS = RF_ROWS_PER_WAVE * (ET / WAVE);      // rows per following workgroup
__device__ __forceinline__ bool eigh_replay_follow(double* ws, int m, int mp, int rb, long long spin_limit,
                                                   double2 (*par)[RP_CHUNK * 64], unsigned long long* shared_p)
{
    const int np = m / 2;
    const double2* rot = reinterpret_cast<const double2*>(ws + rot_record_offset(mp));
    const int* granks = reinterpret_cast<const int*>(ws + rot_record_offset(mp) + (long long)MAX_SWEEPS * (mp - 1) * mp);
    double* extra = ws + (long long)S_COUNT * mp * mp + mp;
    const unsigned long long* progress = reinterpret_cast<const unsigned long long*>(extra + EX_PROGRESS);
    double* Vout = ws + (long long)S_V * mp * mp;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int s = lane < np ? lane : np - 1;                    // lanes past np mirror the last pair
    int row[RF_ROWS_PER_WAVE];
    double top[RF_ROWS_PER_WAVE], bot[RF_ROWS_PER_WAVE];
#pragma unroll
    for (int q = 0; q < RF_ROWS_PER_WAVE; ++q) {
        row[q] = rb * RF_ROWS + RF_ROWS_PER_WAVE * wave + q;
        top[q] = (row[q] == s) ? 1.0 : 0.0;
        bot[q] = (row[q] == m - 1 - s) ? 1.0 : 0.0;
    }
    auto round_step = [&](double2 rv) {
#pragma unroll
        for (int q = 0; q < RF_ROWS_PER_WAVE; ++q) {
            const double t = fma(rv.x, top[q], -(rv.y * bot[q]));
            const double u = fma(rv.y, top[q], rv.x * bot[q]);
            if (np == 1) { top[q] = t; bot[q] = u; continue; }
            const double up = dpp<0x138, 0xF, 0xF>(t);
            const double dn = dpp<0x130, 0xF, 0xF>(u);
            const double b0 = readlane_d(u, 0);
            const double tl = readlane_d(t, np - 1);
            top[q] = s == 0 ? t : (s == 1 ? b0 : up);
            bot[q] = s == np - 1 ? tl : dn;
        }
    };
    bool failed = false;
    for (int ch = 0;; ++ch) {
        const unsigned long long need = (unsigned long long)(ch + 1) * RP_CHUNK;
        if (threadIdx.x == 0) {
            unsigned long long p = __hip_atomic_load(progress, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            long long spins = 0;
            while (!(p & PROGRESS_FINAL) && p < need) {
                __builtin_amdgcn_s_sleep(2);
                if (++spins > spin_limit) { p = ~0ull; break; }
                p = __hip_atomic_load(progress, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            *shared_p = p;
        }
        __syncthreads();                                        // also: chunk ch - 2's buffer no longer read
        const unsigned long long p = *shared_p;
        if (p == ~0ull) { failed = true; break; }
        const long long avail = (long long)(p & (PROGRESS_FINAL - 1));
        const int nr = (int)min((long long)RP_CHUNK, avail - (long long)ch * RP_CHUNK);
        if (nr <= 0) break;                                     // final count reached
        double2* pc = par[ch & 1];
        for (int e = threadIdx.x; e < nr * np; e += ET) {
            const int rr = e / np, ss = e - rr * np;
            const unsigned long long* src =
                reinterpret_cast<const unsigned long long*>(rot + (long long)(ch * RP_CHUNK + rr) * np + ss);
            const unsigned long long cx = __hip_atomic_load(src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned long long sx = __hip_atomic_load(src + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            pc[rr * 64 + ss] = make_double2(__longlong_as_double((long long)cx), __longlong_as_double((long long)sx));
        }
        __syncthreads();                                        // chunk ch staged
        if (nr == RP_CHUNK) {
            double2 rv[RP_CHUNK];
#pragma unroll
            for (int rr = 0; rr < RP_CHUNK; ++rr) rv[rr] = pc[rr * 64 + s];
#pragma unroll
            for (int rr = 0; rr < RP_CHUNK; ++rr) round_step(rv[rr]);
        } else {
            for (int rr = 0; rr < nr; ++rr) round_step(pc[rr * 64 + s]);
        }
        if (nr < RP_CHUNK) break;                               // a short chunk is the last one
    }
    if (failed) {
        if (threadIdx.x == 0)
            __hip_atomic_store(reinterpret_cast<unsigned long long*>(extra + EX_REPLAY_FAILED),
                               (unsigned long long)__double_as_longlong(1.0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return false;
    }
#pragma unroll
    for (int q = 0; q < RF_ROWS_PER_WAVE; ++q) {
        const int i = row[q];
        if (i >= mp) continue;
        double* out = Vout + (long long)i * mp;
        if (i < m) {
            if (lane < np) {
                out[__hip_atomic_load(granks + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)] = top[q];
                out[__hip_atomic_load(granks + m - 1 - lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)] = bot[q];
            }
            for (int j = m + lane; j < mp; j += 64) out[j] = 0.0;
        } else {
            for (int j = lane; j < mp; j += 64) out[j] = 0.0;
        }
    }
    return true;
}

// ---------------------------------------------------------------- chained-HVP PCG
// The preconditioned CG of the Newton step (dual_recommended_hvp.py:141-162, with the HVP of
// :74-76 = hvp_fused.cu:1-148) with the whole iteration held on chip, for even NT <= 6:
//   * the search direction p (skew) lives in LDS as a full MP x LS matrix;
//   * the HVP H(p) = Z - Z^T, Z = P (p Sr) - Q (p Si), runs as chained MFMA products: a wave
//     owns column block J of one or both terms; its first product U[:, J] = p S_t[:, J] ends
//     in the accumulator layout (lane r + 16 q holds U[16 I + q + 4 e][16 J + r]), which is
//     exactly the B-operand layout of k-block 4 I + e of the second product, so
//     Z[:, J] = M_t U[:, J] (M_0 = P, M_1 = -Q) takes U from registers. The items are
//     (column block, term) pairs, three per SIMD at NT = 6;
//   * Z goes to LDS (over p, after every wave has finished reading p) for the transpose;
//   * all CG iterates are exactly skew, so each thread owns strictly upper elements: the
//     direction in registers, the residual and the solution in LDS arrays indexed by the
//     element number; the inner products are twice the upper sums.
// P is exactly symmetric (mirrored inverse product), so its A fragment P[16 I + r][4 kb + q]
// is read as P[4 kb + q][16 I + r], a coalesced row segment; Q through its transposed copy.
// At NT = 8 only p fits the LDS (128 x 129 doubles): the residual, the solution and the owned
// elements of p itself live in the workspace (RXG), read and written once per element pass.
template <int NT>
struct ChainCfg {
    static constexpr int MP = 16 * NT;
    static constexpr int LS = MP + 1;               // LDS row stride (odd: conflict-free columns)
    static constexpr int UN = MP * (MP - 1) / 2;    // strictly upper elements
    static constexpr int UPT = (UN + DT - 1) / DT;  // owned per thread
    static constexpr int OFF = 48;                  // after the reduction scratch
    static constexpr int OFF_R = OFF + MP * LS;     // residual, then solution, UN doubles each
    static constexpr int OFF_X = OFF_R + UN;
    static constexpr bool RXG = NT == 8;            // residual, solution and p's elements in global memory
    static constexpr bool OK = (NT == 2 || NT == 4 || NT == 6 || NT == 8);
};
static_assert(ChainCfg<6>::OFF_X + ChainCfg<6>::UN <= Cfg<6>::OFF_LAM, "chained CG LDS overlaps lambda");
static_assert(ChainCfg<4>::OFF_X + ChainCfg<4>::UN <= Cfg<4>::OFF_LAM, "chained CG LDS overlaps lambda");
static_assert(ChainCfg<2>::OFF_X + ChainCfg<2>::UN <= Cfg<2>::OFF_LAM, "chained CG LDS overlaps lambda");
static_assert(ChainCfg<8>::OFF + ChainCfg<8>::MP * ChainCfg<8>::LS <= Cfg<8>::OFF_LAM, "chained CG LDS overlaps lambda");
static_assert(2 * ChainCfg<8>::UN <= ChainCfg<8>::MP * ChainCfg<8>::MP, "residual and solution fit one slot");

// Column block J and term mask (bit t: term t) of a wave's chained-HVP items.
template <int NT>
__device__ __forceinline__ void chain_items(int wave, int& J, int& tmask) {
    if (NT == 6) {
        if (wave < 4) { J = wave; tmask = 3; }
        else { J = 4 + (wave & 1); tmask = 1 << ((wave >> 1) & 1); }
    } else if (NT == 8) {
        J = wave; tmask = 3;
    } else if (NT == 4) {
        J = wave & 3; tmask = 1 << (wave >> 2);
    } else {   // NT == 2
        J = wave & 1; tmask = wave < 4 ? 1 << (wave >> 1) : 0;
    }
}

// Strictly upper element u in [0, MP (MP - 1) / 2): rows a and MP - 1 - a fold into one
// row of MP - 1 elements.
template <int MP>
__device__ __forceinline__ void upper_ij(int u, int& i, int& j) {
    const int a = u / (MP - 1), b = u - a * (MP - 1);
    if (b < MP - 1 - a) { i = a; j = a + 1 + b; } else { i = MP - 1 - a; j = b + 1; }
}

// Raw buffer resource over n doubles at p (loads past the end return 0).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const double* p, int n) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, n * 8, 0x00020000);
}
// One double at byte offset vo (per lane) + so (uniform).
__device__ __forceinline__ double buf_ld(__amdgpu_buffer_rsrc_t rs, int vo, int so) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, vo, so, 0));
}

// Z[:, J] of one term (t = 0: P (p Sr); t = 1: -Q (p Si)) accumulated into accZ (QT = Q^T). The
// operands in global memory are read through buffer resources with one per-lane offset and
// uniform k / row-block offsets, so no per-load address registers are kept live.
template <int NT>
__device__ __forceinline__ void chain_term(int t, int J, int r, int q, const double* M,
                                           const double* Sr, const double* Si, const double* P, const double* QT,
                                           d4 (&accZ)[NT]) {
    constexpr int MP = 16 * NT, LS = MP + 1, K4 = 4 * NT;
    const __amdgpu_buffer_rsrc_t rS = buf_rsrc(t == 0 ? Sr : Si, MP * MP);
    const int vS = opaque((q * MP + 16 * J + r) * 8);
    const double* Mp = M + opaque(r * LS + q);
    d4 accU[NT];
#pragma unroll
    for (int I = 0; I < NT; ++I) accU[I] = zero4();
    // U[:, J] = p S_t[:, J]: A = p from LDS, B = rows of S_t (coalesced)
    double bq[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) bq[u] = buf_ld(rS, vS, 4 * u * MP * 8);
#pragma unroll 1
    for (int kb0 = 0; kb0 < K4; kb0 += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int kb = kb0 + u;
            const double bcur = bq[u];
            bq[u] = buf_ld(rS, vS, 4 * (kb + 4) * MP * 8);   // past the end: 0, unused
            double a[NT];
#pragma unroll
            for (int I = 0; I < NT; ++I) a[I] = Mp[16 * I * LS + 4 * kb];
#pragma unroll
            for (int I = 0; I < NT; ++I) accU[I] = mfma(a[I], bcur, accU[I]);
        }
    }
    // Z[:, J] += M_t U[:, J]; U[:, J] in registers is the B operand. The A fragment
    // M[16 I + r][4 kb + q] is read from P[4 kb + q][16 I + r] (P exactly symmetric) or from
    // the transposed copy QT of Q: both are coalesced row segments.
    const __amdgpu_buffer_rsrc_t rA = buf_rsrc(t == 0 ? P : QT, MP * MP);
    const int vA = opaque((q * MP + r) * 8);
    constexpr int ks = 4 * MP * 8;                       // A-fragment step per k-block (bytes)
    constexpr int is = 16 * 8;                           // A-fragment step per row block
    // the sign of -Q is carried by the B operand ((-Q) V = Q (-V) product by product), so the
    // prefetched A fragments are used as loaded
    constexpr int D = SSN_CG_PREFETCH;
    double an[D][NT];
#pragma unroll
    for (int d = 0; d < D; ++d)
#pragma unroll
        for (int I = 0; I < NT; ++I) an[d][I] = buf_ld(rA, vA, d * ks + I * is);
#pragma unroll
    for (int kb = 0; kb < K4; ++kb) {
        const double bu = accU[kb >> 2][kb & 3];
        const double bb = t == 0 ? bu : -bu;
#pragma unroll
        for (int I = 0; I < NT; ++I) accZ[I] = mfma(an[kb % D][I], bb, accZ[I]);
        if (kb + D < K4) {
#pragma unroll
            for (int I = 0; I < NT; ++I) an[kb % D][I] = buf_ld(rA, vA, (kb + D) * ks + I * is);
        }
        if (SSN_CG_SCHED_BARRIER) __builtin_amdgcn_sched_barrier(0);
    }
}

// NT = 8: item (t, J) forms U[:, J] = p S_t[:, J] as chain_term does, then Z_t[:, J] = M_t U[:, J]
// (M_0 = P, M_1 = -Q) in groups of RG row tiles, each group stored to zs (row-major MP x MP), so that
// only one group of accumulators is live next to U.
template <int NT, int RG>
__device__ __forceinline__ void chain_term_staged(int t, int J, int r, int q, const double* M,
                                                  const double* Sr, const double* Si, const double* P,
                                                  const double* QT, double* zs) {
    constexpr int MP = 16 * NT, LS = MP + 1, K4 = 4 * NT;
    const __amdgpu_buffer_rsrc_t rS = buf_rsrc(t == 0 ? Sr : Si, MP * MP);
    const int vS = opaque((q * MP + 16 * J + r) * 8);
    const double* Mp = M + opaque(r * LS + q);
    d4 accU[NT];
#pragma unroll
    for (int I = 0; I < NT; ++I) accU[I] = zero4();
    double bq[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) bq[u] = buf_ld(rS, vS, 4 * u * MP * 8);
#pragma unroll 1
    for (int kb0 = 0; kb0 < K4; kb0 += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int kb = kb0 + u;
            const double bcur = bq[u];
            bq[u] = buf_ld(rS, vS, 4 * (kb + 4) * MP * 8);   // past the end: 0, unused
            double a[NT];
#pragma unroll
            for (int I = 0; I < NT; ++I) a[I] = Mp[16 * I * LS + 4 * kb];
#pragma unroll
            for (int I = 0; I < NT; ++I) accU[I] = mfma(a[I], bcur, accU[I]);
        }
    }
    const __amdgpu_buffer_rsrc_t rA = buf_rsrc(t == 0 ? P : QT, MP * MP);
    const int vA = opaque((q * MP + r) * 8);
    constexpr int ks = 4 * MP * 8, is = 16 * 8, D = 4;     // A fragments in flight (4 k-blocks of 4 MFMAs)
    double* zo = zs + opaque((q * MP + 16 * J + r));
#pragma unroll
    for (int g0 = 0; g0 < NT; g0 += RG) {
        d4 accZ[RG];
#pragma unroll
        for (int i = 0; i < RG; ++i) accZ[i] = zero4();
        double an[D][RG];
#pragma unroll
        for (int d = 0; d < D; ++d)
#pragma unroll
            for (int i = 0; i < RG; ++i) an[d][i] = buf_ld(rA, vA, d * ks + (g0 + i) * is);
#pragma unroll
        for (int kb = 0; kb < K4; ++kb) {
            const double bu = accU[kb >> 2][kb & 3];
            const double bb = t == 0 ? bu : -bu;
#pragma unroll
            for (int i = 0; i < RG; ++i) accZ[i] = mfma(an[kb % D][i], bb, accZ[i]);
            if (kb + D < K4) {
#pragma unroll
                for (int i = 0; i < RG; ++i) an[kb % D][i] = buf_ld(rA, vA, (kb + D) * ks + (g0 + i) * is);
            }
            if (SSN_CG_SCHED_BARRIER) __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int i = 0; i < RG; ++i)
#pragma unroll
            for (int e = 0; e < 4; ++e) zo[(16 * (g0 + i) + 4 * e) * MP] = accZ[i][e];
    }
}

// PCG on the scaled dual Hessian from the initial residual init(i, j) (i < j). Returns the
// number of Hessian-vector products; `how` = 0 (curvature not positive), 1 (tolerance met),
// 2 (iteration cap). ADJ selects the adjoint's rule (stop when |r| <= tol |r0|, none if
// r0 = 0) instead of the forward's (stop unless |r| > tol). Writes the solution (full skew)
// to X and, if Z0 is given, the preconditioned initial residual (full skew) to Z0.
// With K::RXG, rx (2 UN doubles: residual, then solution), pg (UN doubles: p's owned elements)
// and zs0, zs1 (MP^2 doubles each: the two terms of Z, Z = term 0 + term 1) are free workspace
// areas that hold what does not fit the LDS.
template <int NT, bool ADJ, class INIT>
__device__ __forceinline__ int cg_chained(const Ctx& c, double* lds, INIT init,
                                          const double* Sr, const double* Si, const double* P, const double* QT,
                                          const double* lam_s, double tol, double* Xg, double* Z0g, int& how,
                                          int max_cg, double* rx = nullptr, double* pg = nullptr,
                                          double* zs0 = nullptr, double* zs1 = nullptr)
{
    using K = ChainCfg<NT>;
    constexpr int MP = K::MP, LS = K::LS, UPT = K::UPT, UN = K::UN;
    constexpr bool RXG = K::RXG;
    double* M = lds + K::OFF;
    double* Rl = RXG ? rx : lds + K::OFF_R;
    double* Xl = RXG ? rx + UN : lds + K::OFF_X;
    const int tid = c.tid, r = c.r, q = c.q;
    int J, tmask;
    chain_items<NT>(c.wave, J, tmask);
    J = __builtin_amdgcn_readfirstlane(J);
    tmask = __builtin_amdgcn_readfirstlane(tmask);

    // p's owned elements: registers, or the workspace area pg (RXG)
    double Pd[RXG ? 1 : UPT];
    auto pd_get = [&](int s, int u) -> double { if constexpr (RXG) return pg[u]; else return Pd[s]; };
    auto pd_set = [&](int s, int u, double v) { if constexpr (RXG) pg[u] = v; else Pd[s] = v; };
    double rz = 0.0, rr0 = 0.0;
    sync();    // earlier users of this LDS area are done
#pragma unroll
    for (int s = 0; s < UPT; ++s) {
        const int u = tid + s * DT;
        if constexpr (!RXG) Pd[s] = 0.0;
        if (u < UN) {
            int i, j;
            upper_ij<MP>(u, i, j);
            const double rv = init(i, j);
            const double z = rv / (lam_s[i] + lam_s[j]);
            Rl[u] = rv; Xl[u] = 0.0; pd_set(s, u, z);
            rz = fma(rv, z, rz);
            rr0 = fma(rv, rv, rr0);
            M[i * LS + j] = z; M[j * LS + i] = -z;
        }
    }
    for (int t = tid; t < MP; t += DT) M[t * LS + t] = 0.0;
    wg_sum2(c, rz, rr0);    // (its barrier also completes p in LDS)
    // the preconditioned initial residual (full skew) copied row by row from p in LDS, so the
    // global stores are coalesced (their completion overlaps the first products)
    if (Z0g) {
        for (int e = tid; e < MP * MP; e += DT) {
            const int i = e / MP, j = e - (e / MP) * MP;
            Z0g[e] = M[i * LS + j];
        }
    }
    rz *= 2.0; rr0 *= 2.0;
    const double thr = ADJ ? tol * sqrt(rr0) : tol;
    int nh = 0;
    how = 2;
    if (ADJ && !(sqrt(rr0) > 0.0)) how = 1;
#pragma unroll 1
    for (int it = 0; it < max_cg && !(ADJ && how == 1); ++it) {
        ++nh;
        if constexpr (RXG) {
            // wave J: both terms of column block J, staged, then Z[:, J] = term 0 + term 1 into LDS
            chain_term_staged<NT, 4>(0, J, r, q, M, Sr, Si, P, QT, zs0);
            chain_term_staged<NT, 4>(1, J, r, q, M, Sr, Si, P, QT, zs1);
            __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
            sync();    // every wave has finished reading p
            const int lo = opaque(c.lane);
#pragma unroll
            for (int g0 = 0; g0 < 32; g0 += 8) {
                double v0[8], v1[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int e = lo + (g0 + u) * 64, i = e >> 4, j = 16 * J + (e & 15);
                    v0[u] = zs0[i * MP + j]; v1[u] = zs1[i * MP + j];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int e = lo + (g0 + u) * 64, i = e >> 4, j = 16 * J + (e & 15);
                    M[i * LS + j] = v0[u] + v1[u];
                }
            }
        }
        d4 accZ[RXG ? 1 : NT];
        if constexpr (!RXG) {
#pragma unroll
        for (int I = 0; I < NT; ++I) accZ[I] = zero4();
        if (tmask & 1) chain_term<NT>(0, J, r, q, M, Sr, Si, P, QT, accZ);
        if (tmask & 2) chain_term<NT>(1, J, r, q, M, Sr, Si, P, QT, accZ);
        }
        sync();    // every wave has finished reading p (RXG: has written its column block of Z)
        if (!RXG && (tmask == 3 || tmask == 1)) {
#pragma unroll
            for (int I = 0; I < NT; ++I)
#pragma unroll
                for (int e = 0; e < 4; ++e) M[(16 * I + q + 4 * e) * LS + 16 * J + r] = accZ[I][e];
        }
        sync();
        if (!RXG && tmask == 2) {
#pragma unroll
            for (int I = 0; I < NT; ++I)
#pragma unroll
                for (int e = 0; e < 4; ++e) M[(16 * I + q + 4 * e) * LS + 16 * J + r] += accZ[I][e];
        }
        sync();
        // H = Z - Z^T on the owned elements; curvature <p, Hp>. The element indices are
        // recomputed from an opaque copy of the thread index every iteration rather than
        // kept live across the products.
        const int ti = opaque(tid);
        double H[UPT];
        double curv = 0.0;
#pragma unroll
        for (int s = 0; s < UPT; ++s) {
            const int u = ti + s * DT;
            H[s] = 0.0;
            if (u < UN) {
                int i, j;
                upper_ij<MP>(u, i, j);
                const double h = M[i * LS + j] - M[j * LS + i];
                H[s] = h;
                curv = fma(pd_get(s, u), h, curv);
            }
        }
        const double curvature = 2.0 * wg_sum(c, curv);
        if (!(curvature > 0.0)) { how = 0; break; }
        const double alpha = rz / curvature;
        double rr = 0.0, rzn = 0.0;
        if constexpr (RXG) {
            // the workspace arrays are distinct, which the compiler cannot tell: the loads of a
            // group of elements are issued before the group's stores
            constexpr int GB = 8;
#pragma unroll
            for (int g0 = 0; g0 < UPT; g0 += GB) {
                double xv[GB], rv0[GB], pv0[GB];
#pragma unroll
                for (int k = 0; k < GB; ++k) {
                    const int u = ti + (g0 + k) * DT;
                    if (g0 + k < UPT && u < UN) { xv[k] = Xl[u]; rv0[k] = Rl[u]; pv0[k] = pg[u]; }
                }
#pragma unroll
                for (int k = 0; k < GB; ++k) {
                    const int s2 = g0 + k, u = ti + s2 * DT;
                    if (s2 < UPT && u < UN) {
                        int i, j;
                        upper_ij<MP>(u, i, j);
                        Xl[u] = xv[k] + alpha * pv0[k];
                        const double rv = rv0[k] - alpha * H[s2];
                        Rl[u] = rv;
                        const double z = rv / (lam_s[i] + lam_s[j]);
                        H[s2] = z;                  // H now holds the preconditioned residual
                        rr = fma(rv, rv, rr);
                        rzn = fma(rv, z, rzn);
                    }
                }
            }
        } else {
#pragma unroll
        for (int s = 0; s < UPT; ++s) {
            const int u = ti + s * DT;
            if (u < UN) {
                int i, j;
                upper_ij<MP>(u, i, j);
                Xl[u] = Xl[u] + alpha * pd_get(s, u);
                const double rv = Rl[u] - alpha * H[s];
                Rl[u] = rv;
                const double z = rv / (lam_s[i] + lam_s[j]);
                H[s] = z;                       // H now holds the preconditioned residual
                rr = fma(rv, rv, rr);
                rzn = fma(rv, z, rzn);
            }
        }
        }
        wg_sum2(c, rr, rzn);    // also: every thread has read Z
        rr *= 2.0; rzn *= 2.0;
        if (ADJ ? (sqrt(rr) <= thr) : !(sqrt(rr) > thr)) { how = 1; break; }
        const double beta = rzn / rz;
        double pold[RXG ? UPT : 1];
        if constexpr (RXG) {
#pragma unroll
            for (int s = 0; s < UPT; ++s) {
                const int u = ti + s * DT;
                pold[s] = u < UN ? pg[u] : 0.0;
            }
        }
#pragma unroll
        for (int s = 0; s < UPT; ++s) {
            const int u = ti + s * DT;
            if (u < UN) {
                int i, j;
                upper_ij<MP>(u, i, j);
                const double pv = H[s] + beta * (RXG ? pold[RXG ? s : 0] : pd_get(s, u));
                pd_set(s, u, pv);
                M[i * LS + j] = pv; M[j * LS + i] = -pv;
            }
        }
        for (int t = tid; t < MP; t += DT) M[t * LS + t] = 0.0;
        rz = rzn;
        sync();
    }
    sync();
    // the solution (full skew) through p's LDS area, then copied row by row (coalesced stores)
#pragma unroll
    for (int s = 0; s < UPT; ++s) {
        const int u = tid + s * DT;
        if (u < UN) {
            int i, j;
            upper_ij<MP>(u, i, j);
            const double x = Xl[u];
            M[i * LS + j] = x; M[j * LS + i] = -x;
        }
    }
    for (int t = tid; t < MP; t += DT) M[t * LS + t] = 0.0;
    sync();
    for (int e = tid; e < MP * MP; e += DT) {
        const int i = e / MP, j = e - (e / MP) * MP;
        Xg[e] = M[i * LS + j];
    }
    sync();
    return nh;
}

// K = I + trial^2 on the packed lower 16 x 16 blocks, from a copy of trial in LDS (row stride
// MP + 1, in the K-panel area of the staged products). Wave w forms the lower tiles (I >= J)
// t = w, w + 8, ... of trial trial, each over k in ascending order of 4-blocks: the same products
// in the same order as the staged product, so K is bitwise the same; the upper tiles, which
// the Cholesky factorisation does not read, are not formed.
template <int NT>
struct TrialCfg {
    static constexpr int MP = 16 * NT, LS = MP + 1, NB = NT * (NT + 1) / 2;
    static constexpr int MAXT = (NB + 7) / 8;
    static constexpr bool OK = MP * LS <= 2 * Cfg<NT>::PANEL;    // the copy fits the panel area
};

template <int NT>
__device__ __forceinline__ void trial_square_lower(const Ctx& c, double* lds) {
    using TC = TrialCfg<NT>;
    constexpr int LS = TC::LS, K4 = 4 * NT;
    const double* Tl = lds + Cfg<NT>::OFF_PANEL;
    double* blk = lds + Cfg<NT>::OFF_BLK;
    const int r = opaque(c.r), q = opaque(c.q);
    // tiles t = wave + 8 s; a wave with fewer than MAXT tiles repeats the last tile (and does not
    // store it), so the k loop has no branches
    int tI[TC::MAXT], tJ[TC::MAXT];
#pragma unroll
    for (int s = 0; s < TC::MAXT; ++s) {
        const int t = c.wave + 8 * s;
        int I = 0, rem = t < TC::NB ? t : TC::NB - 1;
        while (rem > I) { rem -= I + 1; ++I; }
        tI[s] = I; tJ[s] = rem;
    }
    d4 acc[TC::MAXT];
#pragma unroll
    for (int s = 0; s < TC::MAXT; ++s) acc[s] = zero4();
    // trial is exactly skew, so the A fragment trial[16 I + r][4 kb + q] is read as
    // -trial[4 kb + q][16 I + r]: a row segment like the B fragment's (no bank conflicts between
    // the rows of a fragment), bitwise the same value
#pragma unroll
    for (int kb = 0; kb < K4; ++kb) {
#pragma unroll
        for (int s = 0; s < TC::MAXT; ++s) {
            const double a = -Tl[(4 * kb + q) * LS + 16 * tI[s] + r];
            const double b = Tl[(4 * kb + q) * LS + 16 * tJ[s] + r];
            acc[s] = mfma(a, b, acc[s]);
        }
    }
#pragma unroll
    for (int s = 0; s < TC::MAXT; ++s) {
        if (c.wave + 8 * s < TC::NB) {
            double* o = blk + bidx(tI[s], tJ[s]) * 256;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int i = q + 4 * e;       // row and column within the block
                o[i * 16 + r] = ((tI[s] == tJ[s] && i == r) ? 1.0 : 0.0) + acc[s][e];
            }
        }
    }
}

// ---------------------------------------------------------------- sandwich in quadrants
// S = N C N^H in the eigenbasis (dual_recommended_hvp.py:120-129, hvp_sandwich_front
// hvp_fused.cu:39-80 and the m x m recovery), N = P + iQ, C = diag(lambda) + i Ci, as two
// rounds of quadrant products with the intermediate in LDS. The products are formed
// transposed so that every global operand is read as coalesced row segments (P exactly
// symmetric, Ci read as stored, QT = Q^T):
//   round A:  Yr^T = diag(lambda) P - Ci^T Q^T,   Yi^T = Ci^T P + diag(lambda) Q^T
//   round B:  T3^T = P Yr^T + Q Yi^T,             T4^T = P Yi^T - Q Yr^T
// then Sr = sym(T3), Si = skew(T4), G = J - Si. Each transposed product takes the same
// products in the same order as the untransposed one, so the results are bitwise those of
// Yr = P diag - Q Ci, Yi = P Ci + Q diag, T3 = Yr P^T + Yi Q^T, T4 = Yi P^T - Yr Q^T.
// Wave w computes quadrant (w & 1, (w >> 1) & 1) of output w >> 2 (both outputs of a round
// have the same work, and waves w, w + 4 share a SIMD).
template <int NT>
struct QuadCfg {
    static constexpr bool OK = (NT == 2 || NT == 4 || NT == 6);   // two m x m buffers fit the LDS
    static constexpr int MP = 16 * NT, H = NT / 2, K4 = 4 * NT;
    static constexpr int LY = MP + 1;              // LDS row stride of the two m x m outputs
    static constexpr int OFF = 48;
    static constexpr int PERE = MP * MP / DT;      // elements per thread in the final pass
};
static_assert(QuadCfg<6>::OFF + 2 * QuadCfg<6>::MP * QuadCfg<6>::LY <= Cfg<6>::OFF_LAM, "sandwich LDS");
static_assert(QuadCfg<4>::OFF + 2 * QuadCfg<4>::MP * QuadCfg<4>::LY <= Cfg<4>::OFF_LAM, "sandwich LDS");
static_assert(QuadCfg<2>::OFF + 2 * QuadCfg<2>::MP * QuadCfg<2>::LY <= Cfg<2>::OFF_LAM, "sandwich LDS");

// acc[i][j] += sum over k of A(16 (I0 + i) + ., k) B(k, 16 (J0 + j) + .), K = MP, with the A
// fragment read from global at a[(4 kb + q) MP + 16 (I0 + i) + r] (vA holds the lane and I0
// part) and the B fragment given by bfrag(kb, j), scaled by bsign.
template <int NT, class BF>
__device__ __forceinline__ void quad_mma(__amdgpu_buffer_rsrc_t rA, int vA, double bsign, BF bfrag,
                                         d4 (&acc)[QuadCfg<NT>::H][QuadCfg<NT>::H]) {
    constexpr int MP = QuadCfg<NT>::MP, H = QuadCfg<NT>::H, K4 = QuadCfg<NT>::K4;
    constexpr int D = SSN_QUAD_PREFETCH;
    double ar[D][H];
#pragma unroll
    for (int d = 0; d < D; ++d)
#pragma unroll
        for (int i = 0; i < H; ++i) ar[d][i] = buf_ld(rA, vA, (4 * d * MP + 16 * i) * 8);
#pragma unroll
    for (int kb = 0; kb < K4; ++kb) {
        double b[H];
#pragma unroll
        for (int j = 0; j < H; ++j) b[j] = bsign * bfrag(kb, j);
#pragma unroll
        for (int i = 0; i < H; ++i)
#pragma unroll
            for (int j = 0; j < H; ++j) acc[i][j] = mfma(ar[kb % D][i], b[j], acc[i][j]);
        if (kb + D < K4) {
#pragma unroll
            for (int i = 0; i < H; ++i) ar[kb % D][i] = buf_ld(rA, vA, (4 * (kb + D) * MP + 16 * i) * 8);
        }
        if (SSN_QUAD_SCHED_BARRIER) __builtin_amdgcn_sched_barrier(0);
    }
}

// quad_mma with the B fragment read from global as well, at b[(4 kb + q) MP + 16 (J0 + j) + r]
// (vB holds the lane and J0 part) and scaled by bsign, both operands requested D k-steps ahead.
template <int NT>
__device__ __forceinline__ void quad_mma_gg(__amdgpu_buffer_rsrc_t rA, int vA, __amdgpu_buffer_rsrc_t rB, int vB,
                                            double bsign, d4 (&acc)[QuadCfg<NT>::H][QuadCfg<NT>::H]) {
    constexpr int MP = QuadCfg<NT>::MP, H = QuadCfg<NT>::H, K4 = QuadCfg<NT>::K4;
    constexpr int D = SSN_QUAD_PREFETCH;
    double ar[D][H], br[D][H];
#pragma unroll
    for (int d = 0; d < D; ++d)
#pragma unroll
        for (int i = 0; i < H; ++i) {
            ar[d][i] = buf_ld(rA, vA, (4 * d * MP + 16 * i) * 8);
            br[d][i] = buf_ld(rB, vB, (4 * d * MP + 16 * i) * 8);
        }
#pragma unroll
    for (int kb = 0; kb < K4; ++kb) {
        double b[H];
#pragma unroll
        for (int j = 0; j < H; ++j) b[j] = bsign * br[kb % D][j];
#pragma unroll
        for (int i = 0; i < H; ++i)
#pragma unroll
            for (int j = 0; j < H; ++j) acc[i][j] = mfma(ar[kb % D][i], b[j], acc[i][j]);
        if (kb + D < K4) {
#pragma unroll
            for (int i = 0; i < H; ++i) {
                ar[kb % D][i] = buf_ld(rA, vA, (4 * (kb + D) * MP + 16 * i) * 8);
                br[kb % D][i] = buf_ld(rB, vB, (4 * (kb + D) * MP + 16 * i) * 8);
            }
        }
        if (SSN_QUAD_SCHED_BARRIER) __builtin_amdgcn_sched_barrier(0);
    }
}

// One or two m x m products in quadrants, both operands from global (row-major, MP x MP):
// C0 = s0 a0^T b0 on waves 0-3 and, when a1 is given, C1 = s1 a1^T b1 on waves 4-7, one quadrant
// per wave; store(o, i, j, v) receives element (i, j) of output o. The transposed A operand is read
// as coalesced row segments; a product X B with X symmetric or exactly skew takes a = X (s = 1 or
// -1), which gives the same products in the same order as the untransposed form. Opens and ends
// with a barrier.
template <int NT, class ST>
__device__ __forceinline__ void quad_gemm2(const Ctx& c, const double* a0, const double* b0, double s0,
                                           const double* a1, const double* b1, double s1, ST store) {
    using K = QuadCfg<NT>;
    constexpr int MP = K::MP, H = K::H;
    const int w = c.wave, r = opaque(c.r), q = opaque(c.q);
    const int o = w >> 2, quad = w & 3, I0 = (quad & 1) * H, J0 = ((quad >> 1) & 1) * H;
    sync();
    if (o == 0 || a1 != nullptr) {
        const double* A = o ? a1 : a0;
        const double* B = o ? b1 : b0;
        const int lane_off = (q * MP + r) * 8;
        d4 acc[H][H];
#pragma unroll
        for (int i = 0; i < H; ++i)
#pragma unroll
            for (int j = 0; j < H; ++j) acc[i][j] = zero4();
        quad_mma_gg<NT>(buf_rsrc(A, MP * MP), opaque(lane_off + 16 * I0 * 8), buf_rsrc(B, MP * MP),
                        opaque(lane_off + 16 * J0 * 8), o ? s1 : s0, acc);
#pragma unroll
        for (int i = 0; i < H; ++i)
#pragma unroll
            for (int j = 0; j < H; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e) store(o, 16 * (I0 + i) + q + 4 * e, 16 * (J0 + j) + r, acc[i][j][e]);
    }
    sync();
}

template <int NT>
__device__ __forceinline__ void sandwich_quad(const Ctx& c, double* lds, const double* Cie, const double* P,
                                              const double* QT, const double* lam_s, int m, const double* Jse,
                                              double* Sr, double* Si, double* G, double& gg) {
    using K = QuadCfg<NT>;
    constexpr int MP = K::MP, H = K::H, LY = K::LY;
    double* Y = lds + K::OFF;
    const int w = c.wave, r = opaque(c.r), q = opaque(c.q);
    const int y = w >> 2, I0 = (w & 1) * H, J0 = ((w >> 1) & 1) * H;
    double* Yo = Y + y * MP * LY;
    const int lane_off = (q * MP + r) * 8;
    d4 acc[H][H];
    // round A: A = Ci^T (Ci read as stored), B = Q^T or P
    {
#pragma unroll
        for (int i = 0; i < H; ++i)
#pragma unroll
            for (int j = 0; j < H; ++j) acc[i][j] = zero4();
        const __amdgpu_buffer_rsrc_t rB = buf_rsrc(y == 0 ? QT : P, MP * MP);
        const int vB = opaque(lane_off + 16 * J0 * 8);
        sync();    // the workspace operands and this LDS area are free
        quad_mma_gg<NT>(buf_rsrc(Cie, MP * MP), opaque(lane_off + 16 * I0 * 8), rB, vB, 1.0, acc);
        const double* E = y == 0 ? P : QT;       // P[a][col] (= P[col][a]) or QT[a][col] (= Q[col][a])
#pragma unroll
        for (int i = 0; i < H; ++i)
#pragma unroll
            for (int j = 0; j < H; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int a = 16 * (I0 + i) + q + 4 * e, col = 16 * (J0 + j) + r;
                    const double d = a < m ? lam_s[a] : 0.0;
                    const double ev = E[a * MP + col];
                    const double v = acc[i][j][e];
                    Yo[a * LY + col] = y == 0 ? ev * d - v : v + ev * d;
                }
    }
    sync();
    // round B: T3^T = P Yr^T + Q Yi^T, T4^T = P Yi^T - Q Yr^T (B operands from LDS)
    {
#pragma unroll
        for (int i = 0; i < H; ++i)
#pragma unroll
            for (int j = 0; j < H; ++j) acc[i][j] = zero4();
        const double* Y0 = Y + (y == 0 ? 0 : MP * LY) + q * LY + 16 * J0 + r;    // P term
        const double* Y1 = Y + (y == 0 ? MP * LY : 0) + q * LY + 16 * J0 + r;    // Q term
        const int vA = opaque(lane_off + 16 * I0 * 8);
        quad_mma<NT>(buf_rsrc(P, MP * MP), vA, 1.0, [&](int kb, int j) { return Y0[4 * kb * LY + 16 * j]; }, acc);
        quad_mma<NT>(buf_rsrc(QT, MP * MP), vA, y == 0 ? 1.0 : -1.0,
                     [&](int kb, int j) { return Y1[4 * kb * LY + 16 * j]; }, acc);
    }
    sync();    // every wave has read Yr^T, Yi^T
#pragma unroll
    for (int i = 0; i < H; ++i)
#pragma unroll
        for (int j = 0; j < H; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                Yo[(16 * (I0 + i) + q + 4 * e) * LY + 16 * (J0 + j) + r] = acc[i][j][e];
    sync();
    // Sr = (T3 + T3^T) / 2, Si = (T4 - T4^T) / 2, G = J - Si, with T3[i][j] = T3^T[j][i]
    const double* T3t = Y;
    const double* T4t = Y + MP * LY;
    // (an opaque thread index keeps the element addresses from being hoisted out of the
    // Newton loop and kept live)
    const int t0 = opaque(c.tid);
    double jv[K::PERE];
#pragma unroll
    for (int s = 0; s < K::PERE; ++s) jv[s] = Jse[t0 + s * DT];
#pragma unroll
    for (int s = 0; s < K::PERE; ++s) {
        const int e = t0 + s * DT, i = e / MP, j = e - (e / MP) * MP;
        Sr[e] = (T3t[j * LY + i] + T3t[i * LY + j]) / 2;
        const double si = (T4t[j * LY + i] - T4t[i * LY + j]) / 2;
        Si[e] = si;
        const double g = jv[s] - si;
        G[e] = g;
        gg = fma(g, g, gg);
    }
}

// ---------------------------------------------------------------- sandwich and line-search tail in strips
// For NT = 8, where the LDS holds one m x m matrix (128 x 129 doubles) but not the two of the
// quadrant forms: the same products as sandwich_quad and ls_tail_quad, computed in row strips of
// the outputs (tile row I, all NT tile columns; wave w takes tile row w), with one intermediate in
// the LDS and the other in a free workspace slot.

// acc[j] += sum over k of A(16 I + ., k) B(k, col j), K = MP = 16 NT, j < TJ: A fragment read from
// global at a[(4 kb + q) MP + 16 I + r] (vA), B fragment bfrag(kb, j) times bsign.
template <int NT, int TJ, class BF>
__device__ __forceinline__ void row_strip_mma(__amdgpu_buffer_rsrc_t rA, int vA, double bsign, BF bfrag, d4 (&acc)[TJ]) {
    constexpr int MP = 16 * NT, K4 = 4 * NT, D = 2;
    // A and B fragments are loaded D k-blocks ahead of their MFMAs
    double ar[D], br[D][TJ];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        ar[d] = buf_ld(rA, vA, 4 * d * MP * 8);
#pragma unroll
        for (int j = 0; j < TJ; ++j) br[d][j] = bfrag(d, j);
    }
#pragma unroll
    for (int kb = 0; kb < K4; ++kb) {
        const double a = ar[kb % D];
        double b[TJ];
#pragma unroll
        for (int j = 0; j < TJ; ++j) b[j] = bsign * br[kb % D][j];
        if (kb + D < K4) {
            ar[kb % D] = buf_ld(rA, vA, 4 * (kb + D) * MP * 8);
#pragma unroll
            for (int j = 0; j < TJ; ++j) br[kb % D][j] = bfrag(kb + D, j);
        }
#pragma unroll
        for (int j = 0; j < TJ; ++j) acc[j] = mfma(a, b[j], acc[j]);
    }
}

// Q = -W P and its transpose QT (NT = 8, wave w: tile row w). W is exactly skew, so its A fragment
// W[16 I + r][4 kb + q] is read as -W[4 kb + q][16 I + r], a coalesced row segment: the products
// are those of the staged form (-W[i][k]) P[k][j] = -(W[i][k] P[k][j]), in the same order.
template <int NT>
__device__ __forceinline__ void strip_q(const Ctx& c, const double* W, const double* P, double* Q, double* QT) {
    constexpr int MP = 16 * NT;
    const int I = c.wave, r = opaque(c.r), q = opaque(c.q);
    const int lane_off = (q * MP + r) * 8;
    const __amdgpu_buffer_rsrc_t rB = buf_rsrc(P, MP * MP);
    const int vB = opaque(lane_off);
    d4 acc[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[j] = zero4();
    sync();    // W and P are written
    row_strip_mma<NT, NT>(buf_rsrc(W, MP * MP), opaque(lane_off + 16 * I * 8), 1.0,
                          [&](int kb, int j) { return buf_ld(rB, vB, (4 * kb * MP + 16 * j) * 8); }, acc);
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int a = 16 * I + q + 4 * e, col = 16 * j + r;
            Q[a * MP + col] = acc[j][e];
            QT[col * MP + a] = acc[j][e];
        }
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    sync();
}

// K = I + trial^2 into the packed lower blocks (NT = 8, wave w: tile row w, from the trial in the
// workspace slot T1). trial is exactly skew, so its A fragment trial[16 I + r][4 kb + q] is read as
// -trial[4 kb + q][16 I + r] with the sign carried by the B operand: the products and their order
// are those of the staged product, so K is bitwise the same. The upper tiles are formed and dropped.
template <int NT>
__device__ __forceinline__ void trial_square_strips(const Ctx& c, double* lds, const double* T1) {
    constexpr int MP = 16 * NT;
    double* blk = lds + Cfg<NT>::OFF_BLK;
    const int I = c.wave, r = opaque(c.r), q = opaque(c.q);
    const int lane_off = (q * MP + r) * 8;
    const __amdgpu_buffer_rsrc_t rT = buf_rsrc(T1, MP * MP);
    const int vB = opaque(lane_off);
    d4 acc[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[j] = zero4();
    row_strip_mma<NT, NT>(rT, opaque(lane_off + 16 * I * 8), -1.0,
                          [&](int kb, int j) { return buf_ld(rT, vB, (4 * kb * MP + 16 * j) * 8); }, acc);
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        if (j > I) continue;
        double* o = blk + bidx(I, j) * 256;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int i = q + 4 * e;       // row and column within the block
            o[i * 16 + r] = ((I == j && i == r) ? 1.0 : 0.0) + acc[j][e];
        }
    }
}

// S = N C N^H and the gradient as sandwich_quad, for NT = 8 (wave w: tile row w of each output).
// Yr^T and then T4^T live in the LDS, Yi^T in the workspace slot Yg, T3^T in the slot Tg.
template <int NT>
__device__ __forceinline__ void sandwich_strips(const Ctx& c, double* lds, const double* Cie, const double* P,
                                                const double* QT, const double* lam_s, int m, const double* Jse,
                                                double* Sr, double* Si, double* G, double& gg, double* Yg, double* Tg) {
    constexpr int MP = 16 * NT, LY = MP + 1, PERE = MP * MP / DT;
    static_assert(NT == DW, "one tile row per wave");
    double* Yl = lds + 48;
    const int w = c.wave, r = opaque(c.r), q = opaque(c.q);
    const int I = w;
    const int lane_off = (q * MP + r) * 8;
    const int vA = opaque(lane_off + 16 * I * 8), vB = opaque(lane_off);
    d4 acc[NT];
    sync();    // the workspace operands and this LDS area are free
    // round A: Yr^T = diag(lambda) P - Ci^T Q^T (y = 0, to LDS), Yi^T = Ci^T P + diag(lambda) Q^T (y = 1, to Yg)
#pragma unroll 1
    for (int y = 0; y < 2; ++y) {
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[j] = zero4();
        const __amdgpu_buffer_rsrc_t rB = buf_rsrc(y == 0 ? QT : P, MP * MP);
        row_strip_mma<NT, NT>(buf_rsrc(Cie, MP * MP), vA, 1.0,
                          [&](int kb, int j) { return buf_ld(rB, vB, (4 * kb * MP + 16 * j) * 8); }, acc);
        const double* E = y == 0 ? P : QT;
        // the loads of half a strip before its stores (Yg and E are distinct slots, which the
        // compiler cannot tell)
#pragma unroll
        for (int j0 = 0; j0 < NT; j0 += NT / 2) {
            double ev[NT / 2][4];
#pragma unroll
            for (int j = 0; j < NT / 2; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e) ev[j][e] = E[(16 * I + q + 4 * e) * MP + 16 * (j0 + j) + r];
#pragma unroll
            for (int j = 0; j < NT / 2; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int a = 16 * I + q + 4 * e, col = 16 * (j0 + j) + r;
                    const double d = a < m ? lam_s[a] : 0.0;
                    const double v = acc[j0 + j][e];
                    if (y == 0) Yl[a * LY + col] = ev[j][e] * d - v;
                    else Yg[a * MP + col] = v + ev[j][e] * d;
                }
        }
    }
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    sync();
    // round B: T3^T = P Yr^T + Q Yi^T (to Tg), then T4^T = P Yi^T - Q Yr^T (registers, then LDS)
    const __amdgpu_buffer_rsrc_t rY = buf_rsrc(Yg, MP * MP);
    const double* Ylq = Yl + q * LY + r;
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[j] = zero4();
    row_strip_mma<NT, NT>(buf_rsrc(P, MP * MP), vA, 1.0, [&](int kb, int j) { return Ylq[4 * kb * LY + 16 * j]; }, acc);
    row_strip_mma<NT, NT>(buf_rsrc(QT, MP * MP), vA, 1.0,
                      [&](int kb, int j) { return buf_ld(rY, vB, (4 * kb * MP + 16 * j) * 8); }, acc);
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) Tg[(16 * I + q + 4 * e) * MP + 16 * j + r] = acc[j][e];
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[j] = zero4();
    row_strip_mma<NT, NT>(buf_rsrc(P, MP * MP), vA, 1.0,
                      [&](int kb, int j) { return buf_ld(rY, vB, (4 * kb * MP + 16 * j) * 8); }, acc);
    row_strip_mma<NT, NT>(buf_rsrc(QT, MP * MP), vA, -1.0, [&](int kb, int j) { return Ylq[4 * kb * LY + 16 * j]; }, acc);
    sync();    // every wave has read Yr^T
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) Yl[(16 * I + q + 4 * e) * LY + 16 * j + r] = acc[j][e];
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    sync();
    // Sr = (T3 + T3^T) / 2, Si = (T4 - T4^T) / 2, G = J - Si, with T3[i][j] = T3^T[j][i]
    const int t0 = opaque(c.tid);
    constexpr int GB = 8;                    // elements whose loads precede their stores
#pragma unroll
    for (int s0 = 0; s0 < PERE; s0 += GB) {
        double jv[GB], ta[GB], tb[GB];
#pragma unroll
        for (int k = 0; k < GB; ++k) {
            const int e = t0 + (s0 + k) * DT, i = e / MP, j = e - (e / MP) * MP;
            jv[k] = Jse[e]; ta[k] = Tg[j * MP + i]; tb[k] = Tg[i * MP + j];
        }
#pragma unroll
        for (int k = 0; k < GB; ++k) {
            const int e = t0 + (s0 + k) * DT, i = e / MP, j = e - (e / MP) * MP;
            Sr[e] = (ta[k] + tb[k]) / 2;
            const double si = (Yl[j * LY + i] - Yl[i * LY + j]) / 2;
            Si[e] = si;
            const double g = jv[k] - si;
            G[e] = g;
            gg = fma(g, g, gg);
        }
    }
}

// The exact change's tail <P' delta, delta Sr + trial (delta Si)> as ls_tail_quad, for NT = 8:
// V = delta Si (LDS) and U = delta Sr (slot Ug) in strips, then per strip S = U + trial V in
// registers and T4 = P' delta in two halves of tile columns, summed against S. Returns this
// thread's partial sum.
template <int NT>
__device__ __forceinline__ double ls_tail_strips(const Ctx& c, double* lds, const double* T1, const double* T2,
                                                 const double* T3, const double* Sr, const double* Si, double* Ug) {
    constexpr int MP = 16 * NT, LY = MP + 1, HJ = NT / 2;
    static_assert(NT == DW, "one tile row per wave");
    double* Vl = lds + 48;
    const int w = c.wave, r = opaque(c.r), q = opaque(c.q);
    const int I = w;
    const int lane_off = (q * MP + r) * 8;
    const int vA = opaque(lane_off + 16 * I * 8), vB = opaque(lane_off);
    const __amdgpu_buffer_rsrc_t rD = buf_rsrc(T3, MP * MP);
    d4 acc[NT];
    sync();    // the LDS blocks are consumed, T1..T3 are written
    // round 1: V = delta Si (y = 0, LDS), U = delta Sr (y = 1, Ug); A = delta^T with the sign on B
#pragma unroll 1
    for (int y = 0; y < 2; ++y) {
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[j] = zero4();
        const __amdgpu_buffer_rsrc_t rS = buf_rsrc(y == 0 ? Si : Sr, MP * MP);
        row_strip_mma<NT, NT>(rD, vA, -1.0, [&](int kb, int j) { return buf_ld(rS, vB, (4 * kb * MP + 16 * j) * 8); }, acc);
#pragma unroll
        for (int j = 0; j < NT; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int a = 16 * I + q + 4 * e, col = 16 * j + r;
                if (y == 0) Vl[a * LY + col] = acc[j][e];
                else Ug[a * MP + col] = acc[j][e];
            }
    }
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    sync();
    // round 2: S = U + trial V (registers), T4 = P' delta against it
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[j] = zero4();
    const double* Vlq = Vl + q * LY + r;
    row_strip_mma<NT, NT>(buf_rsrc(T1, MP * MP), vA, -1.0, [&](int kb, int j) { return Vlq[4 * kb * LY + 16 * j]; }, acc);
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) acc[j][e] = Ug[(16 * I + q + 4 * e) * MP + 16 * j + r] + acc[j][e];
    double tail = 0.0;
    const __amdgpu_buffer_rsrc_t rP = buf_rsrc(T2, MP * MP);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        d4 acc2[HJ];
#pragma unroll
        for (int j = 0; j < HJ; ++j) acc2[j] = zero4();
        row_strip_mma<NT, HJ>(rP, vA, 1.0, [&](int kb, int j) { return buf_ld(rD, vB, (4 * kb * MP + 16 * (HJ * h + j)) * 8); },
                          acc2);
#pragma unroll
        for (int j = 0; j < HJ; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) tail = fma(acc2[j][e], acc[HJ * h + j][e], tail);
    }
    return tail;
}

// ---------------------------------------------------------------- line-search tail in quadrants
// The exact change of the Armijo test (hvp_line_partial / hvp_line_finish, hvp_fused.cu:240-367):
// tail = <P' delta, delta Sr + trial (delta Si)>, from delta (T3, exactly skew), trial (T1,
// exactly skew), P' (T2, exactly symmetric) in global memory. Two rounds of quadrant products:
//   round 1: waves 0-3: V = delta Si (quadrant w) -> LDS;  waves 4-7: U = delta Sr (registers)
//   round 2: waves 0-3: T4 = P' delta (registers);  waves 4-7: S = U + trial V -> LDS
// then waves 0-3 sum T4 * S over their quadrant. A fragments of the skew operands are read
// transposed (coalesced) with the sign carried by the B operand, so every product equals the
// untransposed one bit for bit. Returns this thread's partial sum.
template <int NT>
__device__ __forceinline__ double ls_tail_quad(const Ctx& c, double* lds, const double* T1, const double* T2,
                                               const double* T3, const double* Sr, const double* Si) {
    using K = QuadCfg<NT>;
    constexpr int MP = K::MP, H = K::H, LY = K::LY;
    double* Vl = lds + K::OFF;            // V, then S in the second buffer
    double* Sl = Vl + MP * LY;
    const int w = c.wave, r = opaque(c.r), q = opaque(c.q);
    const int quad = w & 3, I0 = (quad & 1) * H, J0 = ((quad >> 1) & 1) * H;
    const int lane_off = (q * MP + r) * 8;
    const int vA = opaque(lane_off + 16 * I0 * 8);
    const int vB = opaque(lane_off + 16 * J0 * 8);
    const __amdgpu_buffer_rsrc_t rD = buf_rsrc(T3, MP * MP);
    d4 acc[H][H], acc2[H][H];
#pragma unroll
    for (int i = 0; i < H; ++i)
#pragma unroll
        for (int j = 0; j < H; ++j) { acc[i][j] = zero4(); acc2[i][j] = zero4(); }
    sync();    // the LDS blocks are consumed, T1..T3 are written
    {
        const __amdgpu_buffer_rsrc_t rS = buf_rsrc(w < 4 ? Si : Sr, MP * MP);
        quad_mma_gg<NT>(rD, vA, rS, vB, -1.0, acc);
    }
    if (w < 4) {
#pragma unroll
        for (int i = 0; i < H; ++i)
#pragma unroll
            for (int j = 0; j < H; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    Vl[(16 * (I0 + i) + q + 4 * e) * LY + 16 * (J0 + j) + r] = acc[i][j][e];
    }
    sync();
    if (w < 4) {
        // T4 = P' delta
#pragma unroll
        for (int i = 0; i < H; ++i)
#pragma unroll
            for (int j = 0; j < H; ++j) acc[i][j] = zero4();
        quad_mma_gg<NT>(buf_rsrc(T2, MP * MP), vA, rD, vB, 1.0, acc);
    } else {
        // S = U + trial V
        const double* Vb = Vl + q * LY + 16 * J0 + r;
        quad_mma<NT>(buf_rsrc(T1, MP * MP), vA, -1.0, [&](int kb, int j) { return Vb[4 * kb * LY + 16 * j]; }, acc2);
#pragma unroll
        for (int i = 0; i < H; ++i)
#pragma unroll
            for (int j = 0; j < H; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    Sl[(16 * (I0 + i) + q + 4 * e) * LY + 16 * (J0 + j) + r] = acc[i][j][e] + acc2[i][j][e];
    }
    sync();
    double tail = 0.0;
    if (w < 4) {
#pragma unroll
        for (int i = 0; i < H; ++i)
#pragma unroll
            for (int j = 0; j < H; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    tail = fma(acc[i][j][e], Sl[(16 * (I0 + i) + q + 4 * e) * LY + 16 * (J0 + j) + r], tail);
    }
    return tail;
}

// ---------------------------------------------------------------- the engine
// Body of target b's workgroup (DT threads, dynamic LDS Cfg<NT>::LDS_END doubles).
// EXT: the change into the eigenbasis before the iteration (Cie, Jse and the starting W, P, Q,
// QT in the workspace) and the change back after it (omega, Pf, Qf from W, P, Q) are done by the
// caller's own launches (dual_tiles_kernel and the element passes in dual.hip), not here.
template <int NT, bool EXT = false>
__device__ __forceinline__ void dual_body(const DualArgs& a, int b)
{
    using Cf = Cfg<NT>;
    constexpr int MP = Cf::MP;
    constexpr long long MM = (long long)MP * MP;
    extern __shared__ double lds[];
    Ctx c;
    c.tid = threadIdx.x; c.wave = __builtin_amdgcn_readfirstlane(c.tid >> 6); c.lane = c.tid & 63;
    c.r = c.lane & 15; c.q = c.lane >> 4;
    c.m = a.m;
    c.red = lds;
    c.phase = 0;
    const int m = c.m;
    const unsigned long long t_start = wall_clock64();
    unsigned long long t_cg = 0, t_ls = 0, t_setup = 0;

    double* ws = a.ws + (long long)b * target_doubles(MP);
    auto slot = [&](int s) { return ws + s * MM; };
    double* lam = ws + S_COUNT * MM;
    const double* extra = lam + MP;
    double* V = slot(S_V);    double* Cie = slot(S_CIE); double* Jse = slot(S_JSE);
    double* W = slot(S_W);    double* P = slot(S_P);     double* Q = slot(S_Q);
    double* Sr = slot(S_SR);  double* Si = slot(S_SI);   double* G = slot(S_G);
    double* Dd = slot(S_D);   double* R = slot(S_R);     double* Z0 = slot(S_Z0);
    double* Pp = slot(S_PP);  double* Hp = slot(S_HP);   double* U = slot(S_U);
    double* Vv = slot(S_VV);  double* T1 = slot(S_T1);   double* T2 = slot(S_T2);
    double* T3 = slot(S_T3);  double* T4 = slot(S_T4);  double* T5 = slot(S_T5);
    double* QT = slot(S_QT);
    double* lam_s = lds + Cf::OFF_LAM;

    const double* Cib = a.Ci + (long long)b * a.sC_b;
    const double scale = extra[0];
    const int sweeps = (int)extra[1];
    const unsigned long long t_eigh = (unsigned long long)extra[2];
    const int v_in_lds = (int)extra[3];
    for (int i = c.tid; i < MP; i += DT) lam_s[i] = lam[i];
    if constexpr (!EXT) {
    // Ci / scale into T1 (dual_recommended_hvp.py:107)
    each<NT>(c, [&](int i, int j, int e) {
        T1[e] = (i < m && j < m) ? Cib[(long long)i * a.ldc + j] / scale : 0.0;
    });
    sync();

    // 3. eigenbasis: Ci <- skew(V^T Ci V), J <- skew(V^T J(V)) / scale (:110-111, :53-55)
    if constexpr (QuadCfg<NT>::OK) {
        // T2 = T1 V (T1 = Ci / scale is exactly skew: a = T1, s = -1); then T4 = V^T T2, T1 = V^T T3
        quad_gemm2<NT>(c, T1, V, -1.0, nullptr, nullptr, 0.0,
                       [&](int, int i, int j, double v) { T2[i * MP + j] = v; });
        each<NT>(c, [&](int i, int j, int e) {
            const int k = m / 2;      // J(V) = [V_bot; -V_top] over the leading m rows
            T3[e] = (i < k) ? V[(i + k) * MP + j] : (i < m ? -V[(i - k) * MP + j] : 0.0);
        });
        quad_gemm2<NT>(c, V, T2, 1.0, V, T3, 1.0,
                       [&](int o, int i, int j, double v) { (o ? T1 : T4)[i * MP + j] = v; });
    } else {
        gemm<NT, false, false>(c, lds, T1, V, [&](int i, int j, double v) { T2[i * MP + j] = v; });
        each<NT>(c, [&](int i, int j, int e) {
            const int k = m / 2;      // J(V) = [V_bot; -V_top] over the leading m rows
            T3[e] = (i < k) ? V[(i + k) * MP + j] : (i < m ? -V[(i - k) * MP + j] : 0.0);
        });
        sync();
        gemm<NT, true, false>(c, lds, V, T2, [&](int i, int j, double v) { T4[i * MP + j] = v; });
        gemm<NT, true, false>(c, lds, V, T3, [&](int i, int j, double v) { T1[i * MP + j] = v; });
    }
    each<NT>(c, [&](int i, int j, int e) {
        Cie[e] = (T4[e] - T4[j * MP + i]) / 2;
        Jse[e] = ((T1[e] - T1[j * MP + i]) / 2) / scale;
        W[e] = 0.0;
        P[e] = (i == j && i < m) ? 1.0 : 0.0;
        Q[e] = 0.0;
        QT[e] = 0.0;
    });
    sync();
    } else {
        sync();
    }
    t_setup = wall_clock64() - t_start;

    long long iters = 0;
    const Caps caps = a.caps;
    // max_newton = 0 leaves the Newton loop at once with the cap reached, as the reference's
    // `for newton_step in range(MAX_NEWTON)` followed by the cap error does (:118, :191)
    int status = (sweeps < 0 || extra[EX_REPLAY_FAILED] != 0.0) ? ST_EIGH
                                                                : (caps.max_newton <= 0 ? ST_NEWTON_CAP : ST_OK);
    int n_cg = 0, n_bt = 0, n_hvp = 0;
    const double sqrt_m = sqrt((double)m);

#pragma unroll 1
    for (int ns = 0; ns < caps.max_newton && status == ST_OK; ++ns) {
        // -- sandwich S = N C N^H and the gradient J - Si (:120-129)
        double gg = 0.0;
        if (ns == 0) {
            each<NT>(c, [&](int i, int j, int e) {
                Sr[e] = (i == j && i < m) ? lam_s[i] : 0.0;
                const double si = Cie[e];
                Si[e] = si;
                const double g = Jse[e] - si;
                G[e] = g;
                gg = fma(g, g, gg);
            });
        } else if constexpr (QuadCfg<NT>::OK) {
            // Q = -W P = W^T P (W exactly skew); Q^T is staged in the sandwich's LDS area (free
            // here; odd row stride) and then stored row by row, so both global stores are coalesced
            double* Ql = lds + QuadCfg<NT>::OFF;
            constexpr int LY = QuadCfg<NT>::LY;
            quad_gemm2<NT>(c, W, P, 1.0, nullptr, nullptr, 0.0, [&](int, int i, int j, double v) {
                Q[i * MP + j] = v; Ql[j * LY + i] = v;
            });
            each<NT>(c, [&](int i, int j, int e) { QT[e] = Ql[i * LY + j]; });
            sandwich_quad<NT>(c, lds, Cie, P, QT, lam_s, m, Jse, Sr, Si, G, gg);
        } else if constexpr (NT == 8) {
            strip_q<NT>(c, W, P, Q, QT);
            sandwich_strips<NT>(c, lds, Cie, P, QT, lam_s, m, Jse, Sr, Si, G, gg, Vv, T5);
        } else {
            gemm<NT, false, false>(c, lds, W, P, [&](int i, int j, double v) {
                Q[i * MP + j] = -v; QT[j * MP + i] = -v;
            });
            // Yr = P diag(lam) - Q Ci -> T1 ; Yi = P Ci + Q diag(lam) -> T2 (hvp_fused.cu:39-80)
            gemm<NT, false, false>(c, lds, Q, Cie, [&](int i, int j, double v) {
                const double d = j < m ? lam_s[j] : 0.0;
                T1[i * MP + j] = P[i * MP + j] * d - v;
            });
            gemm<NT, false, false>(c, lds, P, Cie, [&](int i, int j, double v) {
                const double d = j < m ? lam_s[j] : 0.0;
                T2[i * MP + j] = v + Q[i * MP + j] * d;
            });
            // Ur = Yr P^T + Yi Q^T -> T3 ; Ui = Yi P^T - Yr Q^T -> T4 (recovery_fused.cu on m x m)
            gemm2_nt<NT>(c, lds, T1, P, T2, Q, 1.0, [&](int i, int j, double v) { T3[i * MP + j] = v; });
            gemm2_nt<NT>(c, lds, T2, P, T1, Q, -1.0, [&](int i, int j, double v) { T4[i * MP + j] = v; });
            each<NT>(c, [&](int i, int j, int e) {
                Sr[e] = (T3[e] + T3[j * MP + i]) / 2;
                const double si = (T4[e] - T4[j * MP + i]) / 2;
                Si[e] = si;
                const double g = Jse[e] - si;
                G[e] = g;
                gg = fma(g, g, gg);
            });
        }
        const double gnorm = sqrt(wg_sum(c, gg));
        const double residual = gnorm * scale / sqrt_m;
        if (!(residual >= caps.feas_tol / 2)) {
            if (residual != residual) status = ST_NONFINITE;
            break;
        }
        ++iters;

        // -- direction (:138-162); precondition(R) = R / (lam_i + lam_j) for skew R
        if (ns == 0) {
            each<NT>(c, [&](int i, int j, int e) {
                const double z = (-G[e]) / (lam_s[i] + lam_s[j]);
                Dd[e] = z; Z0[e] = z;
            });
            sync();
        } else if constexpr (ChainCfg<NT>::OK) {
            const double tol = fmin(fmax(sqrt(residual), 1e-8), 0.1) * gnorm;
            const unsigned long long t0 = wall_clock64();
            int how;
            const int nh = cg_chained<NT, false>(c, lds, [&](int i, int j) { return -G[i * MP + j]; },
                                                 Sr, Si, P, QT, lam_s, tol, Dd, Z0, how, caps.max_cg, R, Pp, Hp, U);
            n_hvp += nh; n_cg += nh;
            t_cg += wall_clock64() - t0;
        } else {
            double rz = 0.0, dummy = 0.0;
            each<NT>(c, [&](int i, int j, int e) {
                const double rv = -G[e];
                const double z = rv / (lam_s[i] + lam_s[j]);
                R[e] = rv; Dd[e] = 0.0; Z0[e] = z; Pp[e] = z;
                rz = fma(rv, z, rz);
            });
            wg_sum2(c, rz, dummy);
            const double tol = fmin(fmax(sqrt(residual), 1e-8), 0.1) * gnorm;
            const unsigned long long t0 = wall_clock64();
#pragma unroll 1
            for (int it = 0; it < caps.max_cg; ++it) {
                // Hp = hvp(P, Q, Sr, Si, p): U = p Sr, V = p Si, Z = P U - Q V, H = Z - Z^T
                gemm<NT, false, false>(c, lds, Pp, Sr, [&](int i, int j, double v) { U[i * MP + j] = v; });
                gemm<NT, false, false>(c, lds, Pp, Si, [&](int i, int j, double v) { Vv[i * MP + j] = v; });
                gemm2<NT>(c, lds, P, U, Q, Vv, -1.0, [&](int i, int j, double v) { T1[i * MP + j] = v; });
                ++n_hvp; ++n_cg;
                double curv = 0.0;
                each<NT>(c, [&](int i, int j, int e) {
                    const double h = (i == j) ? 0.0 : T1[e] - T1[j * MP + i];
                    Hp[e] = h;
                    curv = fma(Pp[e], h, curv);
                });
                const double curvature = wg_sum(c, curv);
                if (!(curvature > 0.0)) break;
                const double alpha = rz / curvature;
                double rr = 0.0, rzn = 0.0;
                each<NT>(c, [&](int i, int j, int e) {
                    Dd[e] = Dd[e] + alpha * Pp[e];
                    const double rv = R[e] - alpha * Hp[e];
                    R[e] = rv;
                    const double z = rv / (lam_s[i] + lam_s[j]);
                    T2[e] = z;
                    rr = fma(rv, rv, rr);
                    rzn = fma(rv, z, rzn);
                });
                wg_sum2(c, rr, rzn);
                if (!(sqrt(rr) > tol)) break;
                const double beta = rzn / rz;
                each<NT>(c, [&](int, int, int e) { Pp[e] = T2[e] + beta * Pp[e]; });
                rz = rzn;
                sync();
            }
            sync();
            t_cg += wall_clock64() - t0;
        }
        // -- fallback to the preconditioned gradient (:163-164)
        double slope = 0.0, dummy2 = 0.0;
        each_loaded<NT, 2>(c, [&](int e, double (&v)[2]) { v[0] = G[e]; v[1] = Dd[e]; },
                           [&](int, int, int, const double (&v)[2]) { slope = fma(v[0], v[1], slope); });
        wg_sum2(c, slope, dummy2);
        if (slope >= 0.0) {
            double s2 = 0.0, d3 = 0.0;
            each<NT>(c, [&](int, int, int e) { Dd[e] = Z0[e]; s2 = fma(G[e], Z0[e], s2); });
            wg_sum2(c, s2, d3);
            slope = s2;
        }

        // -- Armijo backtracking with the exact change (:166-188)
        double step = 1.0;
        bool accepted = false;
        const unsigned long long t1 = wall_clock64();
#pragma unroll 1
        for (int bt = 0; bt < caps.max_backtrack; ++bt) {
            ++n_bt;
            double first = 0.0;
            // delta = step d; trial = skew(omega + delta) (:168-170). omega and d are exactly
            // skew, so with a = omega_ij + delta_ij the transposed sum is exactly -a and
            // (a - (-a)) / 2 = a: the skew part needs no transposed reads and is bitwise the same.
            double* Tl = lds + Cf::OFF_PANEL;                          // trial copy (row stride MP + 1)
            each_loaded<NT, 3>(c, [&](int e, double (&v)[3]) { v[0] = Dd[e]; v[1] = W[e]; v[2] = G[e]; },
                               [&](int i, int j, int e, const double (&v)[3]) {
                const double de = step * v[0];
                const double tr = v[1] + de;
                T3[e] = de;                                              // delta
                T1[e] = tr;                                              // trial
                if constexpr (TrialCfg<NT>::OK) Tl[i * (MP + 1) + j] = tr;
                first = fma(v[2], de, first);
            });
            sync();
            // K = I + trial^2 into the packed lower blocks
            if constexpr (TrialCfg<NT>::OK) {
                trial_square_lower<NT>(c, lds);
            } else if constexpr (NT == 8) {
                trial_square_strips<NT>(c, lds, T1);
            } else {
                double* blk = lds + Cf::OFF_BLK;
                gemm<NT, false, false>(c, lds, T1, T1, [&](int i, int j, double v) {
                    const int I = i >> 4, J = j >> 4;
                    if (J <= I) blk[bidx(I, J) * 256 + (i & 15) * 16 + (j & 15)] = (i == j ? 1.0 : 0.0) + v;
                });
            }
            sync();
            const bool valid = cholesky_inverse<NT>(c, lds);
            if (valid) {
                inverse_product<NT>(c, lds, T2);                        // P' = (L L^T)^-1
                double tail = 0.0;
                if constexpr (QuadCfg<NT>::OK) {
                    tail = ls_tail_quad<NT>(c, lds, T1, T2, T3, Sr, Si);
                } else if constexpr (NT == 8) {
                    tail = ls_tail_strips<NT>(c, lds, T1, T2, T3, Sr, Si, Vv);
                } else {
                    gemm<NT, false, false>(c, lds, T3, Sr, [&](int i, int j, double v) { U[i * MP + j] = v; });
                    gemm<NT, false, false>(c, lds, T3, Si, [&](int i, int j, double v) { Vv[i * MP + j] = v; });
                    gemm<NT, false, false>(c, lds, T2, T3, [&](int i, int j, double v) { T4[i * MP + j] = v; });
                    gemm<NT, false, false>(c, lds, T1, Vv, [&](int i, int j, double v) {
                        tail = fma(T4[i * MP + j], U[i * MP + j] + v, tail);   // V (delta Sr + trial delta Si)
                    });
                }
                wg_sum2(c, tail, first);
                const double change = first + tail;
                if (change <= 1e-4 * step * slope) {
                    if constexpr (EXT) {
                        // accept: the trial and P' become W and P by exchanging the slots' roles
                        // (no copy); the slots are put back after the loop
                        double* tw = W; W = T1; T1 = tw;
                        double* tp = P; P = T2; T2 = tp;
                    } else {
                        each_loaded<NT, 2>(c, [&](int e, double (&v)[2]) { v[0] = T1[e]; v[1] = T2[e]; },
                                           [&](int, int, int e, const double (&v)[2]) { W[e] = v[0]; P[e] = v[1]; });
                    }
                    sync();
                    accepted = true;
                    break;
                }
            }
            step = step / 2;
        }
        t_ls += wall_clock64() - t1;
        if (!accepted) { status = ST_NO_DESCENT; break; }
        if (ns == caps.max_newton - 1) status = ST_NEWTON_CAP;
    }
    sync();
    if (EXT && W != slot(S_W)) {
        // W and P in the trial slots after an odd number of exchanges: back to their own slots
        // (the change of basis, the omega pass and the backward read them there)
        double* const Ws = slot(S_W);
        double* const Ps = slot(S_P);
        each_loaded<NT, 2>(c, [&](int e, double (&v)[2]) { v[0] = W[e]; v[1] = P[e]; },
                           [&](int, int, int e, const double (&v)[2]) { Ws[e] = v[0]; Ps[e] = v[1]; });
        T1 = W; T2 = P; W = Ws; P = Ps;
        sync();
    }

    // 4. back to the original basis (:58-64): omega = skew(V W V^T), P, Q rotated
    if constexpr (!EXT) {
    double* Pf = a.Pf + (long long)b * MM;
    double* Qf = a.Qf + (long long)b * MM;
    if constexpr (QuadCfg<NT>::OK) {
        // V^T into T5; T1 = W V^T (a = W, s = -1: W exactly skew), T2 = P V^T (P exactly
        // symmetric), T3 = Q V^T (a = Q^T); then T4 = V T1, Pf = V T2, Qf = V T3 (a = V^T)
        each<NT>(c, [&](int i, int j, int e) { T5[e] = V[j * MP + i]; });
        quad_gemm2<NT>(c, W, T5, -1.0, P, T5, 1.0,
                       [&](int o, int i, int j, double v) { (o ? T2 : T1)[i * MP + j] = v; });
        quad_gemm2<NT>(c, QT, T5, 1.0, T5, T1, 1.0,
                       [&](int o, int i, int j, double v) { (o ? T4 : T3)[i * MP + j] = v; });
        quad_gemm2<NT>(c, T5, T2, 1.0, T5, T3, 1.0,
                       [&](int o, int i, int j, double v) { (o ? Qf : Pf)[i * MP + j] = v; });
    } else {
        gemm<NT, false, true>(c, lds, W, V, [&](int i, int j, double v) { T1[i * MP + j] = v; });
        gemm<NT, false, true>(c, lds, P, V, [&](int i, int j, double v) { T2[i * MP + j] = v; });
        gemm<NT, false, true>(c, lds, Q, V, [&](int i, int j, double v) { T3[i * MP + j] = v; });
        gemm<NT, false, false>(c, lds, V, T1, [&](int i, int j, double v) { T4[i * MP + j] = v; });
        gemm<NT, false, false>(c, lds, V, T2, [&](int i, int j, double v) { Pf[i * MP + j] = v; });
        gemm<NT, false, false>(c, lds, V, T3, [&](int i, int j, double v) { Qf[i * MP + j] = v; });
    }
    double* Ob = a.omega + (long long)b * a.sO_b;
    for (int e = c.tid; e < m * m; e += DT) {
        const int i = e / m, j = e - (e / m) * m;
        Ob[(long long)i * a.ldo + j] = (T4[i * MP + j] - T4[j * MP + i]) / 2;
    }
    }
    if (c.tid == 0) {
        a.iters[b] = iters;
        a.status[b] = status;
        if (a.counters) {
            int* cnt = a.counters + NCOUNT * b;
            cnt[0] = (int)iters; cnt[1] = n_cg; cnt[2] = n_bt; cnt[3] = n_hvp;
            cnt[4] = sweeps; cnt[5] = v_in_lds;
            cnt[6] = (int)t_eigh; cnt[7] = (int)t_setup; cnt[8] = (int)t_cg; cnt[9] = (int)t_ls;
            cnt[10] = (int)(wall_clock64() - t_start); cnt[11] = 0;
        }
    }
}


}  // namespace ssn
