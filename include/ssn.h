/* libssn.so — MI355X (gfx950) Dual solver for the nearest real symplectic Stiefel matrix:
 *   min_X ||X - T||_F^2  s.t.  X^T J_2n X = J_2k,   T of shape 2n x 2k (m = 2k), FP64.
 *
 * C ABI: plain pointers and sizes. Every pointer argument except `stream` is a DEVICE
 * pointer owned by the caller. Every call is asynchronous on `stream` (a hipStream_t,
 * or NULL for the default stream); nothing is allocated inside the library: scratch
 * comes in through `ws` / `ws_bytes`, sized by ssn_workspace_bytes(). Functions return
 * SSN_OK (0) or a negative SSN_ERR_* code and never throw. Per-target solver outcomes
 * are written to the device array `status` (SSN_ST_*).
 *
 * Matrices are row-major with element stride 1: a batch of matrices is addressed as
 * base + b * batch_stride + row * row_stride + col. "mp" below is ssn_padded_order(m)
 * = 16 * ceil(m / 16); [batch][mp][mp] arrays are dense and zero padded.
 *
 * Supported shapes in this build: n >= 1, m even, 2 <= m <= 128.
 */
#ifndef SSN_H
#define SSN_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SSN_OK 0
#define SSN_ERR_ARGS (-1)        /* bad shape / null pointer / leading dimension        */
#define SSN_ERR_HIP (-2)         /* a HIP call failed; see ssn_last_hip_error()         */
#define SSN_ERR_WORKSPACE (-3)   /* ws missing or smaller than ssn_workspace_bytes()    */
#define SSN_ERR_UNSUPPORTED (-4) /* shape not supported on this device (e.g. LDS size)  */
#define SSN_ERR_INIT (-5)        /* no current HIP device                               */

/* per-target status values written by ssn_dual / ssn_solve */
#define SSN_ST_OK 0
#define SSN_ST_NO_DESCENT 1      /* Armijo backtracking exhausted (dual_recommended_hvp.py:189-190) */
#define SSN_ST_NEWTON_CAP 2      /* Newton cap (MAX_NEWTON = 100) reached (dual_recommended_hvp.py:191) */
#define SSN_ST_NONFINITE 3       /* non-finite residual                                             */
#define SSN_ST_EIGH 4            /* Jacobi eigensolver did not converge                             */

/* Replaces initialize_kernels() (src/symplectic_projection/__init__.py:10-19): there is
 * nothing to JIT-compile; it caches the current device's LDS limit and the device address of
 * the static scratch of ssn_solve_static on that device. Idempotent. */
int ssn_init(void);

const char* ssn_version(void);
int ssn_padded_order(int m);
int ssn_last_hip_error(void);
const char* ssn_status_string(int status);

/* Host-to-host copy of `bytes` bytes from src to dst on `threads` host threads (clamped to 1..64):
 * the copy of a CPU target into pinned staging before its host-to-device transfer. Replaces
 * `host.copy_(T[:, start:stop])` into the pinned buffer of gram_chunks
 * (dual_recommended_hvp.py:86-90), whose rate otherwise follows the caller's intra-op thread
 * setting. Host-only (no HIP call). SSN_ERR_ARGS for a null pointer with bytes > 0. */
int ssn_host_copy(void* dst, const void* src, size_t bytes, int threads);

/* Bytes of device scratch needed by ssn_solve (and sufficient for ssn_gram, ssn_dual,
 * ssn_recover) for `batch` targets of shape 2n x m. 0 for unsupported shapes. Non-decreasing in
 * `batch`: a workspace sized for a batch serves every smaller batch of the same shape. */
size_t ssn_workspace_bytes(int batch, int n, int m);

/* Cr = A^T A + D^T D, Ci = A^T D - D^T A for T = [A; D] ([batch][2n][m], strides sT_b, sT_r).
 * Cr, Ci: [batch][m][m] with strides sC_b, ldc.
 * Replaces gram_add_ / gram_offdiag / gram_diagonal and the chunk loop of gram_chunks
 * (gram_triangle_reflect.py:19-36, gram_triangle_reflect.cu:1-104, dual_recommended_hvp.py:79-96). */
int ssn_gram(void* stream, const double* T, long long sT_b, long long sT_r, int batch, int n, int m,
             double* Cr, double* Ci, long long sC_b, int ldc, void* ws, size_t ws_bytes);

/* Dual Newton / PCG / Armijo from the Gram parts, entirely on device, one workgroup per target.
 * Outputs: omega [batch][m][m] (strides sO_b, ldo) in the original basis; P, Q [batch][mp][mp],
 * the inverse parts (I + i omega)^-1 = P + iQ used by recovery; iters [batch] Newton updates;
 * status [batch]; counters [batch][12] (may be NULL) = Newton updates, CG iterations, backtracks,
 * HVPs, eigensolver rounds (Jacobi sweeps, or multisection rounds of the tridiagonal eigensolver),
 * eigensolver form (0 Jacobi with V in global memory, 1 Jacobi with V in LDS, 2 tridiagonal),
 * then device wall-clock ticks (10 ns) spent in the eigensolver, in the whole setup (scaling,
 * eigensolver, basis change), in CG, in the line search, in the kernel, 0.
 * Replaces dual_batch (dual_recommended_hvp.py:99-191) with hvp_inner / hvp_outer /
 * hvp_outer_off16 / hvp_outer_diag16 / hvp_sandwich_front / hvp_line_partial / hvp_line_finish /
 * hvp_precondition / hvp_cg_update / hvp_cg_next (hvp_fused.cu:1-420, hvp_fused.py:22-143),
 * triangular_inverse / inverse_product (cholesky_inverse_kernel.cu:1-103) and the torch.linalg
 * eigh / cholesky_ex / cholesky_solve calls on that path. */
int ssn_dual(void* stream, const double* Cr, const double* Ci, long long sC_b, int ldc, int batch, int m,
             double* omega, long long sO_b, int ldo, double* P, double* Q,
             long long* iters, int* status, int* counters, void* ws, size_t ws_bytes);

/* The two halves of ssn_dual, for callers that time or overlap them separately: ssn_dual_eigh
 * computes scale = max(1, tr(Cr)/m) and the eigendecomposition of Cr / scale
 * (dual_recommended_hvp.py:106-108, torch.linalg.eigh) into `ws`; ssn_dual_newton then runs
 * the Newton / PCG / Armijo iteration from that `ws` (same outputs as ssn_dual). The same ws
 * (at least ssn_workspace_bytes(batch, 1, m)) must be passed to both, in that order. */
int ssn_dual_eigh(void* stream, const double* Cr, long long sC_b, int ldc, int batch, int m,
                  void* ws, size_t ws_bytes);
int ssn_dual_newton(void* stream, const double* Cr, const double* Ci, long long sC_b, int ldc, int batch, int m,
                    double* omega, long long sO_b, int ldo, double* P, double* Q,
                    long long* iters, int* status, int* counters, void* ws, size_t ws_bytes);

/* X = [A P^T + D Q^T ; D P^T - A Q^T] with P, Q [batch][mp][mp] from ssn_dual.
 * X: [batch][2n][m], strides sX_b, sX_r.
 * Replaces recover_block / recover_fused and recover_matmul
 * (recovery_fused.py:19-30, recovery_fused.cu:1-41, dual_recommended_hvp.py:194-228). */
int ssn_recover(void* stream, const double* T, long long sT_b, long long sT_r, const double* P,
                const double* Q, int batch, int n, int m, double* X, long long sX_b, long long sX_r,
                void* ws, size_t ws_bytes);

/* Gram -> dual -> recovery in one call (solve, dual_recommended_hvp.py:231-242), device in and
 * device out. The host-side mirror maps SSN_ST_NO_DESCENT / SSN_ST_NEWTON_CAP to the reference's
 * batch-wide RuntimeError. */
int ssn_solve(void* stream, const double* T, long long sT_b, long long sT_r, int batch, int n, int m,
              double* X, long long sX_b, long long sX_r, double* omega, long long sO_b, int ldo,
              long long* iters, int* status, int* counters, void* ws, size_t ws_bytes);

/* Iteration caps and stopping tolerance of the dual iteration, and the precision of the Gram and
 * recovery, for ssn_solve_opts. The reference fixes the caps and the tolerance as module
 * constants (dual_recommended_hvp.py:15-16: FEAS_TOL = 1e-8, MAX_NEWTON = 100, MAX_CG = 150,
 * MAX_BACKTRACK = 40); ssn_default_options() fills those values, with which ssn_solve_opts is
 * ssn_solve (split = 0) or ssn_solve_split (split = 1). */
typedef struct ssn_options {
    int max_newton;      /* Newton updates, 0 .. 2^20; reaching it sets SSN_ST_NEWTON_CAP (0: at once)      */
    int max_cg;          /* CG iterations per Newton step, 0 .. 2^20 (0: zero direction, so the step
                            falls back to the preconditioned gradient, dual_recommended_hvp.py:163-164) */
    int max_backtrack;   /* Armijo trials per Newton step, 0 .. 1100; exhausting them sets SSN_ST_NO_DESCENT */
    int split;           /* 0: FP64 Gram and recovery; 1: split precision (ssn_solve_split)                */
    double feas_tol;     /* a target stops once its scaled residual < feas_tol / 2; finite, >= 0           */
    int shared_device;   /* 0: the eigensolver may take every CU for a moment (its one-launch form, which
                            overlaps the eigenvector replay with the rotations); 1: other work runs on
                            the device beside this solve (several batches in flight): the two-launch
                            form, which leaves CUs free. Same results bit for bit either way.          */
} ssn_options;

void ssn_default_options(ssn_options* opts);

/* ssn_solve with the caps, tolerance and precision of `opts` (host pointer, read during the call).
 * SSN_ERR_ARGS for a null or out-of-range `opts`. Replaces solve (dual_recommended_hvp.py:231-242)
 * run with other values of the module constants of :15-16. */
int ssn_solve_opts(void* stream, const ssn_options* opts, const double* T, long long sT_b, long long sT_r,
                   int batch, int n, int m, double* X, long long sX_b, long long sX_r, double* omega,
                   long long sO_b, int ldo, long long* iters, int* status, int* counters, void* ws,
                   size_t ws_bytes);

/* Allocation-free form of ssn_solve (same arguments without ws / ws_bytes, same results bit for
 * bit). No counterpart in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * workspace is a static device array of SSN_STATIC_SCRATCH_BYTES inside the library, one per GPU,
 * held from ssn_init (or from the first call on that GPU) until the process ends; the call
 * allocates and frees nothing. Returns SSN_ERR_WORKSPACE when ssn_workspace_bytes(batch, n, m)
 * exceeds SSN_STATIC_SCRATCH_BYTES (e.g. more than 16 targets of 8192 x 96, or more than 2 of
 * 16384 x 128). All calls on one GPU share that GPU's array and it keeps the last call's
 * contents (the solve needs none of it cleared): calls that overlap in time on one GPU (several
 * streams or host threads) race on it, and the caller must serialise them. */
#define SSN_STATIC_SCRATCH_BYTES (128ull << 20)
int ssn_solve_static(void* stream, const double* T, long long sT_b, long long sT_r, int batch, int n, int m,
                     double* X, long long sX_b, long long sX_r, double* omega, long long sO_b, int ldo,
                     long long* iters, int* status, int* counters);

/* Split-precision forms of ssn_gram, ssn_recover and ssn_solve (same arguments and workspace).
 * No reference counterpart (the reference's Gram and recovery are FP64 only,
 * gram_triangle_reflect.cu:1-104, recovery_fused.cu:1-41). Every FP64 operand of the Gram and
 * of the recovery is cut into 5 BF16 slices under a power-of-two scale shared by its group
 * (a column of [A | D] per row chunk for the Gram; a row of [A | D], and a column of the
 * recovery matrix, for the recovery); slice products run on the BF16 MFMA with exact FP32
 * accumulation per order and are reassembled in FP64. The dual phase is FP64 as in ssn_dual.
 * Relative error of Cr, Ci and X is about 2^-39 of the group maxima. */
int ssn_gram_split(void* stream, const double* T, long long sT_b, long long sT_r, int batch, int n, int m,
                   double* Cr, double* Ci, long long sC_b, int ldc, void* ws, size_t ws_bytes);
int ssn_recover_split(void* stream, const double* T, long long sT_b, long long sT_r, const double* P,
                      const double* Q, int batch, int n, int m, double* X, long long sX_b, long long sX_r,
                      void* ws, size_t ws_bytes);
int ssn_solve_split(void* stream, const double* T, long long sT_b, long long sT_r, int batch, int n, int m,
                    double* X, long long sX_b, long long sX_r, double* omega, long long sO_b, int ldo,
                    long long* iters, int* status, int* counters, void* ws, size_t ws_bytes);

/* Backward of the projection (no counterpart in the reference, which is forward-only): for the
 * upstream gradient G = dL/dX ([batch][2n][m], strides sG_b, sG_r) writes dT = dL/dT. `ws` must be
 * the workspace of the ssn_solve call that produced X, unmodified since. Implicit differentiation
 * through the dual optimality condition: cross Gram M = B^H G_c (streaming), per target
 * g = skew(-Im(N M N)), lam = H^-1 g by PCG (the forward's Hessian-vector product and
 * preconditioner, relative residual 1e-13), K = N lam N, then dL/dB = G_c N + 2i B K (streaming).
 * counters [batch][4] (may be NULL): adjoint CG iterations, 0 if the Hessian solve converged,
 * device ticks (10 ns), 0. */
int ssn_backward(void* stream, const double* T, long long sT_b, long long sT_r,
                 const double* G, long long sG_b, long long sG_r, int batch, int n, int m,
                 double* dT, long long sD_b, long long sD_r, int* counters, void* ws, size_t ws_bytes);

/* ---- uneven batches: targets of mixed (n, m) in one launch per stage ---------------------
 * No reference counterpart (the reference solves one shape per call, dual_recommended_hvp.py:
 * 231-242). Use: bytes = ssn_uneven_workspace_bytes(t, count); allocate ws; plan_bytes =
 * ssn_uneven_plan_bytes(t, count); ssn_uneven_plan(t, count, ws, bytes, plan, plan_bytes) fills the
 * HOST buffer `plan` (device pointers in it point into ws); copy `plan` to the first plan_bytes
 * of ws; then ssn_solve_uneven(stream, plan, ws, bytes) launches the stages (Gram, eigensolver,
 * Newton, recovery), each one launch over every target, largest work first. The same plan and
 * ws may be reused (and the launches captured in a graph) while the targets' shapes and
 * pointers stay the same. gram_chunk_rows (multiple of 16, 0 = 512) and rec_rows (multiple of
 * 64, 0 = 64) are the tile shapes an autotuner may choose per target. */
typedef struct ssn_target {
    const double* T;   long long sT_r;     /* target [2n][m], row stride in elements        */
    double* X;         long long sX_r;     /* output [2n][m]                                */
    double* omega;     long long ldo;      /* output [m][m]                                 */
    long long* iters;  int* status;        /* 1 element each                                */
    int* counters;                         /* 12 ints (as ssn_dual) or NULL                 */
    int n, m;
    int gram_chunk_rows, rec_rows;
} ssn_target;

size_t ssn_uneven_workspace_bytes(const ssn_target* targets, int count);
size_t ssn_uneven_plan_bytes(const ssn_target* targets, int count);
int ssn_uneven_plan(const ssn_target* targets, int count, void* ws, size_t ws_bytes, void* plan,
                    size_t plan_bytes);
int ssn_solve_uneven(void* stream, const void* plan, void* ws, size_t ws_bytes);

/* The stages of ssn_solve_uneven selected by `stages` (an OR of the flags below), in stage order,
 * for callers that time or overlap them: Gram (partials and their reduction into Cr, Ci), dual
 * (eigensolver and Newton iteration, from Cr, Ci), recovery (X from P, Q). A stage reads what the
 * earlier stages left in `ws`; SSN_STAGE_ALL is ssn_solve_uneven. */
#define SSN_STAGE_GRAM 1
#define SSN_STAGE_DUAL 2
#define SSN_STAGE_RECOVER 4
#define SSN_STAGE_ALL 7
int ssn_solve_uneven_stages(void* stream, const void* plan, void* ws, size_t ws_bytes, int stages);

#ifdef __cplusplus
}
#endif

#endif /* SSN_H */
