// extern "C" entry points of libssn.so (declared in include/ssn.h).
// Plain pointers and sizes only; every device pointer is caller-owned, every call
// is asynchronous on the caller's stream, and no entry point throws or allocates:
// scratch comes in through the `ws` argument, sized by ssn_workspace_bytes(), except in
// ssn_solve_static, whose scratch is the library's static device array.
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>
#include "common.h"
#include "uneven.h"
#include "../../include/ssn.h"

namespace ssn {
int gram_chunk_count(int batch, int n);
hipError_t launch_gram_partial(hipStream_t, const double*, long long, long long, int, int, int, int, double*);
hipError_t launch_gram_reduce(hipStream_t, const double*, int, int, int, double*, double*, long long, int);
hipError_t launch_recovery_fragments(hipStream_t, const double*, const double*, int, int, double*);
hipError_t launch_recover_pq(hipStream_t, const double*, long long, long long, const double*, const double*, int,
                             int, int, double*, long long, long long, size_t);
hipError_t launch_recover(hipStream_t, const double*, long long, long long, const double*, int, int, int,
                          double*, long long, long long);
size_t dual_workspace_doubles(int batch, int m);
hipError_t launch_dual(hipStream_t, const double*, const double*, long long, int, int, int,
                       double*, long long, int, double*, double*, long long*, int*, int*, double*, size_t,
                       const Caps&);
hipError_t launch_dual_eigh(hipStream_t, const double*, long long, int, int, int, double*, size_t);
hipError_t launch_cross_gram(hipStream_t, const double*, long long, long long, const double*, long long, long long,
                             int, int, int, int, double*, double*, double*, long long);
hipError_t launch_adjoint(hipStream_t, double*, const double*, const double*, int, int, int*, double, size_t);
hipError_t launch_backward_stream(hipStream_t, const double*, long long, long long, const double*, long long,
                                  long long, const double*, const double*, long long, const double*,
                                  const double*, double*, int, int, int, double*, long long, long long);
long long workspace_target_doubles(int m);
hipError_t launch_dual_newton(hipStream_t, const double*, const double*, long long, int, int, int,
                              double*, long long, int, double*, double*, long long*, int*, int*, double*, size_t,
                              const Caps&);
hipError_t launch_split_gram(hipStream_t, const double*, long long, long long, int, int, int, int, double*);
long long split_fragment_doubles(int m);
hipError_t launch_split_recover(hipStream_t, const double*, long long, long long, const double*, const double*,
                                double*, int, int, int, double*, long long, long long);
hipError_t launch_gram_uneven(hipStream_t, const UTarget*, const UItem*, const int*, int);
hipError_t launch_dual_uneven(hipStream_t, const UTarget*, int, size_t, size_t, int);
hipError_t launch_recover_uneven(hipStream_t, const UTarget*, const UItem*, int, int, int);
void eigh_variant(int m, size_t lds_limit, int* small, int* vlds, size_t* lds_bytes);
size_t eigh_tri_lds_bytes(int m);
size_t dual_lds_bytes(int m);
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

// Workspace of ssn_solve_static: a static device array, one copy per GPU (each GPU holds its own
// at its own address). In host code `scratch` names a host-side stand-in, not device memory: the
// library only ever uses the device address that hipGetSymbolAddress returns for the current
// device (g_scratch below), never the name itself.
static __device__ __align__(256) double scratch[SSN_STATIC_SCRATCH_BYTES / sizeof(double)];

namespace {

constexpr int MAX_DEVICES = 64;
size_t g_lds_limit[MAX_DEVICES];
bool g_ready[MAX_DEVICES];
void* g_scratch[MAX_DEVICES];          // device address of `scratch` on each device, once looked up
hipError_t g_last_hip = hipSuccess;

// Device address of the static scratch on device `dev` (current), looked up once per device.
int static_scratch(int dev, void** out) {
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) { g_last_hip = e != hipSuccess ? e : hipErrorInvalidSymbol; return SSN_ERR_HIP; }
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return SSN_OK;
}

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
    if (e == LAUNCH_UNSUPPORTED) return SSN_ERR_UNSUPPORTED;
    g_last_hip = e;
    return SSN_ERR_HIP;
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

// Gram partial slabs (batch x chunks) needed by any batch of at most `batch` targets: the
// chunk count of a batch falls as the batch grows, so batch * chunks alone is not monotonic
// (51 targets take 561 slabs, 52 take 520). Sizing for the worst smaller batch makes
// ssn_workspace_bytes() non-decreasing in the batch, so a workspace sized for a batch also
// serves every smaller one. A batch of at least the Gram's target workgroup count (512, at
// most 2048 when tuned) takes one chunk per target, so batches past 2048 need no search.
size_t partial_slabs(int batch, int n) {
    size_t worst = (size_t)batch * gram_chunk_count(batch > 0 ? batch : 1, n);
    const int lim = std::min(batch, 2048);
    for (int b = 1; b <= lim; ++b) worst = std::max(worst, (size_t)b * gram_chunk_count(b, n));
    return worst;
}

Layout layout(int batch, int n, int m) {
    Layout L;
    const size_t mp = (size_t)pad16(m);
    L.chunks = gram_chunk_count(batch > 0 ? batch : 1, n);
    size_t off = 0;
    L.partial = off; off += partial_slabs(batch, n) * 2 * mp * mp;
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

int ssn_init(void) {
    const int rc = current_device_ready(nullptr);
    if (rc) return rc;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return SSN_ERR_INIT;
    void* p = nullptr;
    return static_scratch(dev, &p);
}

const char* ssn_version(void) { return "ssn-gfx950 0.1.0"; }

int ssn_host_copy(void* dst, const void* src, size_t bytes, int threads) {
    if (bytes == 0) return SSN_OK;
    if (!dst || !src) return SSN_ERR_ARGS;
    threads = std::max(1, std::min(threads, 64));
    const size_t page = 4096;
    size_t per = (bytes + threads - 1) / threads;
    per = (per + page - 1) / page * page;                 // page-aligned slices
    char* d = static_cast<char*>(dst);
    const char* sp = static_cast<const char*>(src);
    std::vector<std::thread> pool;
    for (size_t off = per; off < bytes; off += per)
        pool.emplace_back([=] { std::memcpy(d + off, sp + off, std::min(per, bytes - off)); });
    std::memcpy(d, sp, std::min(per, bytes));             // the caller's thread takes the first slice
    for (auto& t : pool) t.join();
    return SSN_OK;
}

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
                                  P, Q, iters, status, counters, (double*)ws, lds, Caps()));
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
                                         P, Q, iters, status, counters, (double*)ws, lds, Caps()));
}

int ssn_recover(void* stream, const double* T, long long sT_b, long long sT_r, const double* P,
                const double* Q, int batch, int n, int m, double* X, long long sX_b, long long sX_r,
                void* ws, size_t ws_bytes)
{
    if (!shape_ok(batch, n, m) || !T || !P || !Q || !X) return SSN_ERR_ARGS;
    if (batch == 0) return SSN_OK;
    size_t lds = 0;
    int rc = current_device_ready(&lds);
    if (rc) return rc;
    const size_t need = (size_t)batch * 4 * pad16(m) * pad16(m) * sizeof(double);
    if (!ws || ws_bytes < need) return SSN_ERR_WORKSPACE;
    hipStream_t s = (hipStream_t)stream;
    const hipError_t e = launch_recover_pq(s, T, sT_b, sT_r, P, Q, batch, n, m, X, sX_b, sX_r, lds);
    if (e != LAUNCH_DECLINED) return hip_status(e);
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

static int solve_impl(bool split, const Caps& caps, void* stream, const double* T, long long sT_b, long long sT_r,
                      int batch, int n, int m, double* X, long long sX_b, long long sX_r, double* omega,
                      long long sO_b, int ldo, long long* iters, int* status, int* counters, void* ws,
                      size_t ws_bytes)
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
                                w + L.pf, w + L.qf, iters, status, counters, w + L.dual, lds, caps));
    if (rc) return rc;
    if (split)
        return hip_status(launch_split_recover(s, T, sT_b, sT_r, w + L.pf, w + L.qf, w + L.frag, batch, n, m,
                                               X, sX_b, sX_r));
    const hipError_t e = launch_recover_pq(s, T, sT_b, sT_r, w + L.pf, w + L.qf, batch, n, m, X, sX_b, sX_r, lds);
    if (e != LAUNCH_DECLINED) return hip_status(e);
    rc = hip_status(launch_recovery_fragments(s, w + L.pf, w + L.qf, batch, m, w + L.frag));
    if (rc) return rc;
    return hip_status(launch_recover(s, T, sT_b, sT_r, w + L.frag, batch, n, m, X, sX_b, sX_r));
}

int ssn_solve(void* stream, const double* T, long long sT_b, long long sT_r, int batch, int n, int m,
              double* X, long long sX_b, long long sX_r, double* omega, long long sO_b, int ldo,
              long long* iters, int* status, int* counters, void* ws, size_t ws_bytes)
{
    return solve_impl(false, Caps(), stream, T, sT_b, sT_r, batch, n, m, X, sX_b, sX_r, omega, sO_b, ldo, iters,
                      status, counters, ws, ws_bytes);
}

int ssn_solve_split(void* stream, const double* T, long long sT_b, long long sT_r, int batch, int n, int m,
                    double* X, long long sX_b, long long sX_r, double* omega, long long sO_b, int ldo,
                    long long* iters, int* status, int* counters, void* ws, size_t ws_bytes)
{
    return solve_impl(true, Caps(), stream, T, sT_b, sT_r, batch, n, m, X, sX_b, sX_r, omega, sO_b, ldo, iters,
                      status, counters, ws, ws_bytes);
}

int ssn_solve_static(void* stream, const double* T, long long sT_b, long long sT_r, int batch, int n, int m,
                     double* X, long long sX_b, long long sX_r, double* omega, long long sO_b, int ldo,
                     long long* iters, int* status, int* counters)
{
    if (!shape_ok(batch, n, m) || !T || !X || !omega || !iters || !status || ldo < m) return SSN_ERR_ARGS;
    if (batch == 0) return SSN_OK;
    if (layout(batch, n, m).total * sizeof(double) > SSN_STATIC_SCRATCH_BYTES) return SSN_ERR_WORKSPACE;
    int rc = current_device_ready(nullptr);
    if (rc) return rc;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return SSN_ERR_INIT;
    void* ws = nullptr;
    rc = static_scratch(dev, &ws);
    if (rc) return rc;
    // the eigensolver's two-launch form: plain kernel launches only (its one-launch form is a
    // cooperative launch, which graph capture and allocation-free callers are not given)
    Caps caps;
    caps.eigh_whole_device = 0;
    return solve_impl(false, caps, stream, T, sT_b, sT_r, batch, n, m, X, sX_b, sX_r, omega, sO_b, ldo, iters,
                      status, counters, ws, SSN_STATIC_SCRATCH_BYTES);
}

void ssn_default_options(ssn_options* opts)
{
    if (!opts) return;
    const Caps d;
    opts->max_newton = d.max_newton;
    opts->max_cg = d.max_cg;
    opts->max_backtrack = d.max_backtrack;
    opts->split = 0;
    opts->feas_tol = d.feas_tol;
    opts->shared_device = 0;
}

int ssn_solve_opts(void* stream, const ssn_options* opts, const double* T, long long sT_b, long long sT_r,
                   int batch, int n, int m, double* X, long long sX_b, long long sX_r, double* omega,
                   long long sO_b, int ldo, long long* iters, int* status, int* counters, void* ws,
                   size_t ws_bytes)
{
    if (!opts) return SSN_ERR_ARGS;
    constexpr int CAP_LIMIT = 1 << 20;
    if (opts->max_newton < 0 || opts->max_newton > CAP_LIMIT || opts->max_cg < 0 || opts->max_cg > CAP_LIMIT ||
        opts->max_backtrack < 0 || opts->max_backtrack > 1100 || (opts->split != 0 && opts->split != 1) ||
        !(opts->feas_tol >= 0.0) || !(opts->feas_tol < 1e300) || (opts->shared_device != 0 && opts->shared_device != 1))
        return SSN_ERR_ARGS;
    Caps caps;
    caps.max_newton = opts->max_newton;
    caps.max_cg = opts->max_cg;
    caps.max_backtrack = opts->max_backtrack;
    caps.feas_tol = opts->feas_tol;
    caps.eigh_whole_device = opts->shared_device ? 0 : 1;
    return solve_impl(opts->split != 0, caps, stream, T, sT_b, sT_r, batch, n, m, X, sX_b, sX_r, omega, sO_b, ldo,
                      iters, status, counters, ws, ws_bytes);
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

// ---------------------------------------------------------------- uneven batches
namespace {

constexpr size_t UNEVEN_ALIGN = 256;
constexpr int UNEVEN_DEFAULT_CHUNK = 512;
constexpr size_t GFX950_LDS_BYTES = 163840;     // planning default when no device is current

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

bool target_ok(const ssn_target& t) {
    return t.T && t.X && t.omega && t.iters && t.status && t.n >= 1 && t.m >= 2 && (t.m % 2) == 0 &&
           t.m <= 128 && t.ldo >= t.m && t.sT_r >= t.m && t.sX_r >= t.m &&
           t.gram_chunk_rows >= 0 && t.rec_rows >= 0;
}

int chunk_rows_of(const ssn_target& t) {
    const int c = t.gram_chunk_rows > 0 ? t.gram_chunk_rows : UNEVEN_DEFAULT_CHUNK;
    return (c + 15) / 16 * 16;
}
int rec_rows_of(const ssn_target& t) {
    const int r = t.rec_rows > 0 ? t.rec_rows : 64;
    return (r + 63) / 64 * 64;
}

// Sizes of one target's workspace regions (doubles), in layout order.
struct UTargetSizes {
    size_t partial, cr, ci, dual, pf, qf, frag;
    size_t total() const {
        return align_up(partial * 8, UNEVEN_ALIGN) + align_up(cr * 8, UNEVEN_ALIGN) +
               align_up(ci * 8, UNEVEN_ALIGN) + align_up(dual * 8, UNEVEN_ALIGN) +
               align_up(pf * 8, UNEVEN_ALIGN) + align_up(qf * 8, UNEVEN_ALIGN) + align_up(frag * 8, UNEVEN_ALIGN);
    }
};

UTargetSizes sizes_of(const ssn_target& t) {
    const size_t mp = (size_t)pad16(t.m);
    const size_t chunks = (size_t)((t.n + chunk_rows_of(t) - 1) / chunk_rows_of(t));
    UTargetSizes z;
    z.partial = chunks * 2 * mp * mp;
    z.cr = z.ci = (size_t)t.m * t.m;
    z.dual = (size_t)workspace_target_doubles(t.m);
    z.pf = z.qf = mp * mp;
    z.frag = 4 * mp * mp;
    return z;
}

struct UCounts { int gram_items, rec_items; };

UCounts counts_of(const ssn_target* t, int count) {
    UCounts c{0, 0};
    for (int i = 0; i < count; ++i) {
        c.gram_items += (t[i].n + chunk_rows_of(t[i]) - 1) / chunk_rows_of(t[i]);
        c.rec_items += (t[i].n + rec_rows_of(t[i]) - 1) / rec_rows_of(t[i]);
    }
    return c;
}

size_t plan_bytes_of(const ssn_target* t, int count) {
    const UCounts c = counts_of(t, count);
    return align_up(sizeof(UPlanHeader), UNEVEN_ALIGN) + align_up((size_t)count * sizeof(UTarget), UNEVEN_ALIGN) +
           align_up((size_t)c.gram_items * sizeof(UItem), UNEVEN_ALIGN) +
           align_up((size_t)c.rec_items * sizeof(UItem), UNEVEN_ALIGN);
}

bool targets_ok(const ssn_target* t, int count) {
    if (!t || count <= 0) return false;
    for (int i = 0; i < count; ++i)
        if (!target_ok(t[i])) return false;
    return true;
}

}  // namespace

size_t ssn_uneven_plan_bytes(const ssn_target* targets, int count) {
    if (!targets_ok(targets, count)) return 0;
    return plan_bytes_of(targets, count);
}

size_t ssn_uneven_workspace_bytes(const ssn_target* targets, int count) {
    if (!targets_ok(targets, count)) return 0;
    size_t total = align_up(plan_bytes_of(targets, count), UNEVEN_ALIGN);
    for (int i = 0; i < count; ++i) total += sizes_of(targets[i]).total();
    return total;
}

int ssn_uneven_plan(const ssn_target* targets, int count, void* ws, size_t ws_bytes, void* plan, size_t plan_bytes)
{
    if (!targets_ok(targets, count) || !ws || !plan) return SSN_ERR_ARGS;
    const size_t need_plan = plan_bytes_of(targets, count);
    const size_t need_ws = ssn_uneven_workspace_bytes(targets, count);
    if (plan_bytes < need_plan) return SSN_ERR_ARGS;
    if (ws_bytes < need_ws) return SSN_ERR_WORKSPACE;
    size_t lds_limit = GFX950_LDS_BYTES;
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
        lds_limit = (size_t)v;

    char* pb = (char*)plan;
    UPlanHeader* h = (UPlanHeader*)pb;
    const UCounts c = counts_of(targets, count);
    h->count = count; h->gram_items = c.gram_items; h->rec_items = c.rec_items;
    h->targets_off = (long long)align_up(sizeof(UPlanHeader), UNEVEN_ALIGN);
    h->gram_off = h->targets_off + (long long)align_up((size_t)count * sizeof(UTarget), UNEVEN_ALIGN);
    h->rec_off = h->gram_off + (long long)align_up((size_t)c.gram_items * sizeof(UItem), UNEVEN_ALIGN);
    h->plan_bytes = (long long)need_plan;
    h->ws_bytes = (long long)need_ws;
    UTarget* tg = (UTarget*)(pb + h->targets_off);
    UItem* gi = (UItem*)(pb + h->gram_off);
    UItem* ri = (UItem*)(pb + h->rec_off);

    char* w = (char*)ws + align_up(need_plan, UNEVEN_ALIGN);
    size_t eigh_lds = 0, dual_lds = 0;
    int max_mp = 0;
    std::vector<std::pair<long long, UItem>> gw, rw;
    for (int i = 0; i < count; ++i) {
        const ssn_target& t = targets[i];
        const UTargetSizes z = sizes_of(t);
        UTarget& u = tg[i];
        u.T = t.T; u.sT_r = t.sT_r; u.X = t.X; u.sX_r = t.sX_r; u.omega = t.omega; u.ldo = t.ldo;
        u.iters = t.iters; u.status = t.status; u.counters = t.counters;
        u.n = t.n; u.m = t.m; u.mp = pad16(t.m);
        u.chunk_rows = chunk_rows_of(t); u.rec_rows = rec_rows_of(t);
        u.chunks = (t.n + u.chunk_rows - 1) / u.chunk_rows;
        auto take = [&](size_t doubles) { double* p = (double*)w; w += align_up(doubles * 8, UNEVEN_ALIGN); return p; };
        u.partial = take(z.partial); u.Cr = take(z.cr); u.Ci = take(z.ci); u.dual = take(z.dual);
        u.Pf = take(z.pf); u.Qf = take(z.qf); u.frag = take(z.frag);
        size_t el = 0;
        eigh_variant(t.m, lds_limit, &u.eigh_small, &u.eigh_vlds, &el);
        el = std::max(el, eigh_tri_lds_bytes(t.m));      // either eigensolver form may run
        const size_t dl = dual_lds_bytes(t.m);
        if (el > lds_limit || dl > lds_limit) return SSN_ERR_UNSUPPORTED;
        eigh_lds = std::max(eigh_lds, el);
        dual_lds = std::max(dual_lds, dl);
        max_mp = std::max(max_mp, u.mp);
        const long long mp2 = (long long)u.mp * u.mp;
        for (int k = 0; k < u.chunks; ++k) {
            const long long rows = std::min(u.chunk_rows, t.n - k * u.chunk_rows);
            gw.push_back({rows * mp2, UItem{i, k}});
        }
        for (int r0 = 0; r0 < t.n; r0 += u.rec_rows) {
            const long long rows = std::min(u.rec_rows, t.n - r0);
            rw.push_back({rows * mp2, UItem{i, r0}});
        }
    }
    // largest work first (ties: target, then index), so the persistent grids end together;
    // the Gram items are grouped by order first (one launch per order)
    auto by_cost = [](const std::pair<long long, UItem>& a, const std::pair<long long, UItem>& b) {
        if (a.first != b.first) return a.first > b.first;
        if (a.second.target != b.second.target) return a.second.target < b.second.target;
        return a.second.index < b.second.index;
    };
    auto by_order_cost = [&](const std::pair<long long, UItem>& a, const std::pair<long long, UItem>& b) {
        const int oa = tg[a.second.target].mp, ob = tg[b.second.target].mp;
        if (oa != ob) return oa < ob;
        return by_cost(a, b);
    };
    std::sort(gw.begin(), gw.end(), by_order_cost);
    std::sort(rw.begin(), rw.end(), by_cost);
    for (int k = 0; k <= 8; ++k) h->gram_begin[k] = 0;
    for (size_t k = 0; k < gw.size(); ++k) h->gram_begin[tg[gw[k].second.target].mp / 16]++;
    for (int k = 1; k <= 8; ++k) h->gram_begin[k] += h->gram_begin[k - 1];
    for (size_t k = 0; k < gw.size(); ++k) gi[k] = gw[k].second;
    for (size_t k = 0; k < rw.size(); ++k) ri[k] = rw[k].second;
    h->eigh_lds = (long long)eigh_lds;
    h->dual_lds = (long long)dual_lds;
    h->max_mp = max_mp;
    return SSN_OK;
}

int ssn_solve_uneven(void* stream, const void* plan, void* ws, size_t ws_bytes)
{
    return ssn_solve_uneven_stages(stream, plan, ws, ws_bytes, SSN_STAGE_ALL);
}

int ssn_solve_uneven_stages(void* stream, const void* plan, void* ws, size_t ws_bytes, int stages)
{
    if (!plan || !ws || stages < 0 || stages > SSN_STAGE_ALL) return SSN_ERR_ARGS;
    const UPlanHeader* h = (const UPlanHeader*)plan;
    if (h->count <= 0 || ws_bytes < (size_t)h->ws_bytes) return SSN_ERR_WORKSPACE;
    int rc = current_device_ready(nullptr);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    const char* dp = (const char*)ws;           // the device copy of the plan
    const UTarget* tg = (const UTarget*)(dp + h->targets_off);
    const UItem* gi = (const UItem*)(dp + h->gram_off);
    const UItem* ri = (const UItem*)(dp + h->rec_off);
    if (stages & SSN_STAGE_GRAM) {
        rc = hip_status(launch_gram_uneven(s, tg, gi, h->gram_begin, h->count));
        if (rc) return rc;
    }
    if (stages & SSN_STAGE_DUAL) {
        int single_nt = 0, orders = 0;       // one order in the plan: its specialised launch
        for (int k = 1; k <= 8; ++k)
            if (h->gram_begin[k] > h->gram_begin[k - 1]) { single_nt = k; ++orders; }
        rc = hip_status(launch_dual_uneven(s, tg, h->count, (size_t)h->eigh_lds, (size_t)h->dual_lds,
                                           orders == 1 ? single_nt : 0));
        if (rc) return rc;
    }
    if (stages & SSN_STAGE_RECOVER) return hip_status(launch_recover_uneven(s, tg, ri, h->rec_items, h->count, h->max_mp));
    return SSN_OK;
}

}  // extern "C"
