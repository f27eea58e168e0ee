// extern "C" entry points of libssn.so (declared in include/ssn.h).
// Plain pointers and sizes only; every device pointer is caller-owned, every call
// is asynchronous on the caller's stream, and no entry point throws or allocates:
// scratch comes in through the `ws` argument, sized by ssn_workspace_bytes().
#include <hip/hip_runtime.h>
#include <stddef.h>
#include "common.h"

namespace ssn {
int gram_chunk_count(int batch, int n);
hipError_t launch_gram_partial(hipStream_t, const double*, long long, long long, int, int, int, int, double*);
hipError_t launch_gram_reduce(hipStream_t, const double*, int, int, int, double*, double*, long long, int);
hipError_t launch_recovery_fragments(hipStream_t, const double*, const double*, int, int, double*);
hipError_t launch_recover(hipStream_t, const double*, long long, long long, const double*, int, int, int,
                          double*, long long, long long);
size_t dual_workspace_doubles(int batch, int m);
hipError_t launch_dual(hipStream_t, const double*, const double*, long long, int, int, int,
                       double*, long long, int, double*, double*, long long*, int*, int*, double*, size_t);
hipError_t launch_dual_eigh(hipStream_t, const double*, long long, int, int, int, double*, size_t);
hipError_t launch_cross_gram(hipStream_t, const double*, long long, long long, const double*, long long, long long,
                             int, int, int, int, double*, double*, double*, long long);
hipError_t launch_adjoint(hipStream_t, double*, const double*, const double*, int, int, int*, double, size_t);
hipError_t launch_backward_stream(hipStream_t, const double*, long long, long long, const double*, long long,
                                  long long, const double*, const double*, long long, const double*,
                                  const double*, double*, int, int, int, double*, long long, long long);
long long workspace_target_doubles(int m);
hipError_t launch_dual_newton(hipStream_t, const double*, const double*, long long, int, int, int,
                              double*, long long, int, double*, double*, long long*, int*, int*, double*, size_t);
hipError_t launch_split_gram(hipStream_t, const double*, long long, long long, int, int, int, int, double*);
long long split_fragment_doubles(int m);
hipError_t launch_split_recover(hipStream_t, const double*, long long, long long, const double*, const double*,
                                double*, int, int, int, double*, long long, long long);
}  // namespace ssn

using namespace ssn;

#define SSN_OK 0
#define SSN_ERR_ARGS (-1)
#define SSN_ERR_HIP (-2)
#define SSN_ERR_WORKSPACE (-3)
#define SSN_ERR_UNSUPPORTED (-4)
#define SSN_ERR_INIT (-5)

// solver workspace slots used by the backward (dual.hip: S_G, S_D, S_R, S_Z0)
#define SLOT_MR 8
#define SLOT_MI 9
#define SLOT_KR 10
#define SLOT_KI 11

namespace {

constexpr int MAX_DEVICES = 64;
size_t g_lds_limit[MAX_DEVICES];
bool g_ready[MAX_DEVICES];
hipError_t g_last_hip = hipSuccess;

int current_device_ready(size_t* lds) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return SSN_ERR_INIT;
    if (!g_ready[dev]) {
        int v = 0;
        hipError_t e = hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
        if (e != hipSuccess) { g_last_hip = e; return SSN_ERR_HIP; }
        g_lds_limit[dev] = (size_t)v;
        g_ready[dev] = true;
    }
    if (lds) *lds = g_lds_limit[dev];
    return SSN_OK;
}

int hip_status(hipError_t e) {
    if (e == hipSuccess) return SSN_OK;
    g_last_hip = e;
    return e == hipErrorInvalidValue ? SSN_ERR_UNSUPPORTED : SSN_ERR_HIP;
}

bool shape_ok(int batch, int n, int m) {
    return batch >= 0 && n >= 1 && m >= 2 && (m % 2) == 0 && m <= 128;
}

// workspace layout (doubles): [gram partials][Cr][Ci][dual slots][Pf][Qf][fragments]; the
// fragment region holds the recovery's 2mp x 2mp or the backward's 4mp x 2mp matrices
struct Layout {
    size_t partial, cr, ci, dual, pf, qf, frag, total;
    int chunks;
};

Layout layout(int batch, int n, int m) {
    Layout L;
    const size_t mp = (size_t)pad16(m);
    L.chunks = gram_chunk_count(batch > 0 ? batch : 1, n);
    size_t off = 0;
    L.partial = off; off += (size_t)batch * L.chunks * 2 * mp * mp;
    L.cr = off;      off += (size_t)batch * m * m;
    L.ci = off;      off += (size_t)batch * m * m;
    L.dual = off;    off += dual_workspace_doubles(batch, m);
    L.pf = off;      off += (size_t)batch * mp * mp;
    L.qf = off;      off += (size_t)batch * mp * mp;
    L.frag = off;    off += (size_t)batch * 8 * mp * mp;
    L.total = off;
    return L;
}

}  // namespace

extern "C" {

int ssn_init(void) { return current_device_ready(nullptr); }

const char* ssn_version(void) { return "ssn-gfx950 0.1.0"; }

int ssn_padded_order(int m) { return pad16(m); }

int ssn_last_hip_error(void) { return (int)g_last_hip; }

const char* ssn_status_string(int code) {
    switch (code) {
        case ST_OK: return "ok";
        case ST_NO_DESCENT: return "line search found no descent step";
        case ST_NEWTON_CAP: return "Newton iteration cap reached";
        case ST_NONFINITE: return "non-finite residual";
        case ST_EIGH: return "eigensolver did not converge";
        default: return "unknown status";
    }
}

size_t ssn_workspace_bytes(int batch, int n, int m) {
    if (!shape_ok(batch, n, m)) return 0;
    return layout(batch, n, m).total * sizeof(double);
}

static int gram_impl(bool split, void* stream, const double* T, long long sT_b, long long sT_r, int batch,
                     int n, int m, double* Cr, double* Ci, long long sC_b, int ldc, void* ws, size_t ws_bytes)
{
    if (!shape_ok(batch, n, m) || !T || !Cr || !Ci || ldc < m) return SSN_ERR_ARGS;
    if (batch == 0) return SSN_OK;
    int rc = current_device_ready(nullptr);
    if (rc) return rc;
    Layout L = layout(batch, n, m);
    if (!ws || ws_bytes < (L.partial + (size_t)batch * L.chunks * 2 * pad16(m) * pad16(m)) * sizeof(double))
        return SSN_ERR_WORKSPACE;
    hipStream_t s = (hipStream_t)stream;
    double* part = (double*)ws + L.partial;
    rc = hip_status(split ? launch_split_gram(s, T, sT_b, sT_r, batch, n, m, L.chunks, part)
                          : launch_gram_partial(s, T, sT_b, sT_r, batch, n, m, L.chunks, part));
    if (rc) return rc;
    return hip_status(launch_gram_reduce(s, part, batch, L.chunks, m, Cr, Ci, sC_b, ldc));
}

int ssn_gram(void* stream, const double* T, long long sT_b, long long sT_r, int batch, int n, int m,
             double* Cr, double* Ci, long long sC_b, int ldc, void* ws, size_t ws_bytes)
{
    return gram_impl(false, stream, T, sT_b, sT_r, batch, n, m, Cr, Ci, sC_b, ldc, ws, ws_bytes);
}

int ssn_gram_split(void* stream, const double* T, long long sT_b, long long sT_r, int batch, int n, int m,
                   double* Cr, double* Ci, long long sC_b, int ldc, void* ws, size_t ws_bytes)
{
    return gram_impl(true, stream, T, sT_b, sT_r, batch, n, m, Cr, Ci, sC_b, ldc, ws, ws_bytes);
}

int ssn_dual(void* stream, const double* Cr, const double* Ci, long long sC_b, int ldc, int batch, int m,
             double* omega, long long sO_b, int ldo, double* P, double* Q,
             long long* iters, int* status, int* counters, void* ws, size_t ws_bytes)
{
    if (!shape_ok(batch, 1, m) || !Cr || !Ci || !omega || !P || !Q || !iters || !status ||
        ldc < m || ldo < m) return SSN_ERR_ARGS;
    if (batch == 0) return SSN_OK;
    size_t lds = 0;
    int rc = current_device_ready(&lds);
    if (rc) return rc;
    if (!ws || ws_bytes < dual_workspace_doubles(batch, m) * sizeof(double)) return SSN_ERR_WORKSPACE;
    return hip_status(launch_dual((hipStream_t)stream, Cr, Ci, sC_b, ldc, batch, m, omega, sO_b, ldo,
                                  P, Q, iters, status, counters, (double*)ws, lds));
}

int ssn_dual_eigh(void* stream, const double* Cr, long long sC_b, int ldc, int batch, int m,
                  void* ws, size_t ws_bytes)
{
    if (!shape_ok(batch, 1, m) || !Cr || ldc < m) return SSN_ERR_ARGS;
    if (batch == 0) return SSN_OK;
    size_t lds = 0;
    int rc = current_device_ready(&lds);
    if (rc) return rc;
    if (!ws || ws_bytes < dual_workspace_doubles(batch, m) * sizeof(double)) return SSN_ERR_WORKSPACE;
    return hip_status(launch_dual_eigh((hipStream_t)stream, Cr, sC_b, ldc, batch, m, (double*)ws, lds));
}

int ssn_dual_newton(void* stream, const double* Cr, const double* Ci, long long sC_b, int ldc, int batch, int m,
                    double* omega, long long sO_b, int ldo, double* P, double* Q,
                    long long* iters, int* status, int* counters, void* ws, size_t ws_bytes)
{
    if (!shape_ok(batch, 1, m) || !Cr || !Ci || !omega || !P || !Q || !iters || !status ||
        ldc < m || ldo < m) return SSN_ERR_ARGS;
    if (batch == 0) return SSN_OK;
    size_t lds = 0;
    int rc = current_device_ready(&lds);
    if (rc) return rc;
    if (!ws || ws_bytes < dual_workspace_doubles(batch, m) * sizeof(double)) return SSN_ERR_WORKSPACE;
    return hip_status(launch_dual_newton((hipStream_t)stream, Cr, Ci, sC_b, ldc, batch, m, omega, sO_b, ldo,
                                         P, Q, iters, status, counters, (double*)ws, lds));
}

int ssn_recover(void* stream, const double* T, long long sT_b, long long sT_r, const double* P,
                const double* Q, int batch, int n, int m, double* X, long long sX_b, long long sX_r,
                void* ws, size_t ws_bytes)
{
    if (!shape_ok(batch, n, m) || !T || !P || !Q || !X) return SSN_ERR_ARGS;
    if (batch == 0) return SSN_OK;
    int rc = current_device_ready(nullptr);
    if (rc) return rc;
    const size_t need = (size_t)batch * 4 * pad16(m) * pad16(m) * sizeof(double);
    if (!ws || ws_bytes < need) return SSN_ERR_WORKSPACE;
    hipStream_t s = (hipStream_t)stream;
    rc = hip_status(launch_recovery_fragments(s, P, Q, batch, m, (double*)ws));
    if (rc) return rc;
    return hip_status(launch_recover(s, T, sT_b, sT_r, (const double*)ws, batch, n, m, X, sX_b, sX_r));
}

int ssn_recover_split(void* stream, const double* T, long long sT_b, long long sT_r, const double* P,
                      const double* Q, int batch, int n, int m, double* X, long long sX_b, long long sX_r,
                      void* ws, size_t ws_bytes)
{
    if (!shape_ok(batch, n, m) || !T || !P || !Q || !X) return SSN_ERR_ARGS;
    if (batch == 0) return SSN_OK;
    int rc = current_device_ready(nullptr);
    if (rc) return rc;
    const size_t need = (size_t)batch * split_fragment_doubles(m) * sizeof(double);
    if (!ws || ws_bytes < need) return SSN_ERR_WORKSPACE;
    return hip_status(launch_split_recover((hipStream_t)stream, T, sT_b, sT_r, P, Q, (double*)ws, batch, n, m,
                                           X, sX_b, sX_r));
}

static int solve_impl(bool split, void* stream, const double* T, long long sT_b, long long sT_r, int batch,
                      int n, int m, double* X, long long sX_b, long long sX_r, double* omega, long long sO_b,
                      int ldo, long long* iters, int* status, int* counters, void* ws, size_t ws_bytes)
{
    if (!shape_ok(batch, n, m) || !T || !X || !omega || !iters || !status || ldo < m) return SSN_ERR_ARGS;
    if (batch == 0) return SSN_OK;
    size_t lds = 0;
    int rc = current_device_ready(&lds);
    if (rc) return rc;
    Layout L = layout(batch, n, m);
    if (!ws || ws_bytes < L.total * sizeof(double)) return SSN_ERR_WORKSPACE;
    hipStream_t s = (hipStream_t)stream;
    double* w = (double*)ws;
    rc = hip_status(split ? launch_split_gram(s, T, sT_b, sT_r, batch, n, m, L.chunks, w + L.partial)
                          : launch_gram_partial(s, T, sT_b, sT_r, batch, n, m, L.chunks, w + L.partial));
    if (rc) return rc;
    rc = hip_status(launch_gram_reduce(s, w + L.partial, batch, L.chunks, m, w + L.cr, w + L.ci,
                                       (long long)m * m, m));
    if (rc) return rc;
    rc = hip_status(launch_dual(s, w + L.cr, w + L.ci, (long long)m * m, m, batch, m, omega, sO_b, ldo,
                                w + L.pf, w + L.qf, iters, status, counters, w + L.dual, lds));
    if (rc) return rc;
    if (split)
        return hip_status(launch_split_recover(s, T, sT_b, sT_r, w + L.pf, w + L.qf, w + L.frag, batch, n, m,
                                               X, sX_b, sX_r));
    rc = hip_status(launch_recovery_fragments(s, w + L.pf, w + L.qf, batch, m, w + L.frag));
    if (rc) return rc;
    return hip_status(launch_recover(s, T, sT_b, sT_r, w + L.frag, batch, n, m, X, sX_b, sX_r));
}

int ssn_solve(void* stream, const double* T, long long sT_b, long long sT_r, int batch, int n, int m,
              double* X, long long sX_b, long long sX_r, double* omega, long long sO_b, int ldo,
              long long* iters, int* status, int* counters, void* ws, size_t ws_bytes)
{
    return solve_impl(false, stream, T, sT_b, sT_r, batch, n, m, X, sX_b, sX_r, omega, sO_b, ldo, iters, status,
                      counters, ws, ws_bytes);
}

int ssn_solve_split(void* stream, const double* T, long long sT_b, long long sT_r, int batch, int n, int m,
                    double* X, long long sX_b, long long sX_r, double* omega, long long sO_b, int ldo,
                    long long* iters, int* status, int* counters, void* ws, size_t ws_bytes)
{
    return solve_impl(true, stream, T, sT_b, sT_r, batch, n, m, X, sX_b, sX_r, omega, sO_b, ldo, iters, status,
                      counters, ws, ws_bytes);
}

int ssn_backward(void* stream, const double* T, long long sT_b, long long sT_r,
                 const double* G, long long sG_b, long long sG_r, int batch, int n, int m,
                 double* dT, long long sD_b, long long sD_r, int* counters, void* ws, size_t ws_bytes)
{
    if (!shape_ok(batch, n, m) || !T || !G || !dT) return SSN_ERR_ARGS;
    if (batch == 0) return SSN_OK;
    size_t lds = 0;
    int rc = current_device_ready(&lds);
    if (rc) return rc;
    Layout L = layout(batch, n, m);
    if (!ws || ws_bytes < L.total * sizeof(double)) return SSN_ERR_WORKSPACE;
    hipStream_t s = (hipStream_t)stream;
    double* w = (double*)ws;
    double* dual = w + L.dual;
    const long long td = workspace_target_doubles(m);
    const long long mm = (long long)pad16(m) * pad16(m);
    // cross Gram M = B^H G_c into the solver slots of M's real / imaginary parts
    rc = hip_status(launch_cross_gram(s, T, sT_b, sT_r, G, sG_b, sG_r, batch, n, m, L.chunks, w + L.partial,
                                      dual + SLOT_MR * mm, dual + SLOT_MI * mm, td));
    if (rc) return rc;
    rc = hip_status(launch_adjoint(s, dual, w + L.pf, w + L.qf, batch, m, counters, 1e-13, lds));
    if (rc) return rc;
    return hip_status(launch_backward_stream(s, T, sT_b, sT_r, G, sG_b, sG_r, dual + SLOT_KR * mm,
                                             dual + SLOT_KI * mm, td, w + L.pf, w + L.qf, w + L.frag,
                                             batch, n, m, dT, sD_b, sD_r));
}

}  // extern "C"
