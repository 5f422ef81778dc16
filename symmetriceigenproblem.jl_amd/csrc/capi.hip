// capi.hip -- the extern "C" surface of libsep_amd (declared in include/sep.h):
// handle lifetime, error reporting, argument validation and host<->device staging.
#include <cstdarg>
#include <cstdlib>
#include <cstring>

#include "sep_internal.hpp"

namespace sep {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int DevBuf::reserve(size_t want)
{
    if (want <= bytes && p) return SEP_OK;
    if (p) {
        (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    if (want == 0) return SEP_OK;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        set_error("hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
        return SEP_ERR_NOMEM;
    }
    bytes = want;
    return SEP_OK;
}

// Make the handle's device current.  hipSetDevice is skipped when it already is: a caller
// capturing a stream into a graph (global or thread-local capture mode) may not call it.
static int use_device(const sep_handle* h)
{
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && cur == h->device) return SEP_OK;
    SEP_HIP(hipSetDevice(h->device));
    return SEP_OK;
}

int timing_mark(sep_handle* h, hipStream_t st)
{
    if (!h || !h->timing) return -1;
    if (h->tev_used == (int)h->tev.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        h->tev.push_back(e);
    }
    const int i = h->tev_used++;
    if (hipEventRecord(h->tev[i], st) != hipSuccess) return -1;
    return i;
}

void DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

}  // namespace sep

using namespace sep;

namespace {

template <typename T>
int bulk_rmul_dev_checked(sep_handle* h, int64_t m, int64_t n, T* A, int64_t lda, const T* cs,
                          int64_t k)
{
    if (!h) {
        set_error("sep_bulk_rmul: null handle");
        return SEP_ERR_ARG;
    }
    if (m < 0 || n < 0 || k < 0 || lda < (m > 1 ? m : 1)) {
        set_error("sep_bulk_rmul: invalid sizes m=%lld n=%lld k=%lld lda=%lld", (long long)m,
                  (long long)n, (long long)k, (long long)lda);
        return SEP_ERR_ARG;
    }
    if (m == 0 || n <= 1 || k == 0) return SEP_OK;
    if (!A || !cs) {
        set_error("sep_bulk_rmul: null pointer");
        return SEP_ERR_ARG;
    }
    if (int rc = use_device(h)) return rc;
    return bulk_apply_dev<T>(h, h->stream, h->apply_work[0], m, n, A, lda, cs, n - 1, k);
}

template <typename T>
int bulk_rmul_host(sep_handle* h, int64_t m, int64_t n, T* A, int64_t lda, const T* cs, int64_t k)
{
    if (!h) {
        set_error("sep_bulk_rmul: null handle");
        return SEP_ERR_ARG;
    }
    if (m < 0 || n < 0 || k < 0 || lda < (m > 1 ? m : 1)) {
        set_error("sep_bulk_rmul: invalid sizes");
        return SEP_ERR_ARG;
    }
    if (m == 0 || n <= 1 || k == 0) return SEP_OK;
    if (!A || !cs) {
        set_error("sep_bulk_rmul: null pointer");
        return SEP_ERR_ARG;
    }
    if (int rc = use_device(h)) return rc;
    const size_t abytes = sizeof(T) * (size_t)m * (size_t)n;
    const size_t gbytes = sizeof(T) * 2 * (size_t)(n - 1) * (size_t)k;
    if (int rc = h->host_io_A.reserve(abytes)) return rc;
    if (int rc = h->host_io_cs.reserve(gbytes)) return rc;
    T* dA = h->host_io_A.as<T>();
    SEP_HIP(hipMemcpy2DAsync(dA, sizeof(T) * m, A, sizeof(T) * lda, sizeof(T) * m, n,
                             hipMemcpyHostToDevice, h->stream));
    SEP_HIP(hipMemcpyAsync(h->host_io_cs.p, cs, gbytes, hipMemcpyHostToDevice, h->stream));
    if (int rc = bulk_apply_dev<T>(h, h->stream, h->apply_work[0], m, n, dA, m,
                                   h->host_io_cs.as<T>(), n - 1, k))
        return rc;
    SEP_HIP(hipMemcpy2DAsync(A, sizeof(T) * lda, dA, sizeof(T) * m, sizeof(T) * m, n,
                             hipMemcpyDeviceToHost, h->stream));
    SEP_HIP(hipStreamSynchronize(h->stream));
    return SEP_OK;
}

template <typename T>
int qr_checked(sep_handle* h, int64_t n, T* dv, T* ev, T* Q, int64_t ldq, int64_t max_layers,
               T tol, int64_t maxiter, int32_t* converged, int64_t* iter, int64_t* n_rot)
{
    if (!h) {
        set_error("sep_qr_algorithm: null handle");
        return SEP_ERR_ARG;
    }
    if (n < 0 || (n > 0 && (!dv || (n > 1 && !ev)))) {
        set_error("sep_qr_algorithm: invalid n or null dv/ev");
        return SEP_ERR_ARG;
    }
    if (n % 2 != 0 && !h->allow_odd) {
        set_error("sep_qr_algorithm: n=%lld must be even (the reference asserts iseven(n); "
                  "sep_set_odd_order lifts this)",
                  (long long)n);
        return SEP_ERR_ODD_N;
    }
    if (max_layers < 1 || max_layers > 4096) {
        set_error("sep_qr_algorithm: max_layers=%lld outside [1, 4096]", (long long)max_layers);
        return SEP_ERR_BAD_LAYERS;
    }
    if (Q && ldq < (n > 1 ? n : 1)) {
        set_error("sep_qr_algorithm: ldq=%lld < n=%lld", (long long)ldq, (long long)n);
        return SEP_ERR_ARG;
    }
    if (int rc = use_device(h)) return rc;
    return qr_algorithm_dev<T>(h, n, dv, ev, Q, ldq, max_layers, tol, maxiter, converged, iter,
                               n_rot);
}

template <typename T>
int qr_host(sep_handle* h, int64_t n, T* dv, T* ev, T* Q, int64_t ldq, int64_t max_layers, T tol,
            int64_t maxiter, int32_t* converged, int64_t* iter, int64_t* n_rot)
{
    if (!h) {
        set_error("sep_qr_algorithm: null handle");
        return SEP_ERR_ARG;
    }
    if (n < 0 || (n > 0 && (!dv || (n > 1 && !ev)))) {
        set_error("sep_qr_algorithm: invalid n or null dv/ev");
        return SEP_ERR_ARG;
    }
    if (n % 2 != 0 && !h->allow_odd) {
        set_error("sep_qr_algorithm: n=%lld must be even (the reference asserts iseven(n); "
                  "sep_set_odd_order lifts this)",
                  (long long)n);
        return SEP_ERR_ODD_N;
    }
    if (Q && ldq < (n > 1 ? n : 1)) {
        set_error("sep_qr_algorithm: ldq=%lld < n=%lld", (long long)ldq, (long long)n);
        return SEP_ERR_ARG;
    }
    if (n == 0) {
        if (converged) *converged = 1;
        if (iter) *iter = 0;
        if (n_rot) *n_rot = 0;
        return SEP_OK;
    }
    if (int rc = use_device(h)) return rc;
    if (int rc = h->host_io_d.reserve(sizeof(T) * (size_t)n)) return rc;
    if (int rc = h->host_io_e.reserve(sizeof(T) * (size_t)(n > 1 ? n - 1 : 1))) return rc;
    T* dd = h->host_io_d.as<T>();
    T* de = h->host_io_e.as<T>();
    T* dQ = nullptr;
    SEP_HIP(hipMemcpyAsync(dd, dv, sizeof(T) * n, hipMemcpyHostToDevice, h->stream));
    if (n > 1)
        SEP_HIP(hipMemcpyAsync(de, ev, sizeof(T) * (n - 1), hipMemcpyHostToDevice, h->stream));
    if (Q) {
        if (int rc = h->host_io_A.reserve(sizeof(T) * (size_t)n * (size_t)n)) return rc;
        dQ = h->host_io_A.as<T>();
        SEP_HIP(hipMemcpy2DAsync(dQ, sizeof(T) * n, Q, sizeof(T) * ldq, sizeof(T) * n, n,
                                 hipMemcpyHostToDevice, h->stream));
    }
    int rc = qr_checked<T>(h, n, dd, de, dQ, n, max_layers, tol, maxiter, converged, iter, n_rot);
    if (rc < 0) return rc;
    SEP_HIP(hipMemcpyAsync(dv, dd, sizeof(T) * n, hipMemcpyDeviceToHost, h->stream));
    if (n > 1)
        SEP_HIP(hipMemcpyAsync(ev, de, sizeof(T) * (n - 1), hipMemcpyDeviceToHost, h->stream));
    if (Q)
        SEP_HIP(hipMemcpy2DAsync(Q, sizeof(T) * ldq, dQ, sizeof(T) * n, sizeof(T) * n, n,
                                 hipMemcpyDeviceToHost, h->stream));
    SEP_HIP(hipStreamSynchronize(h->stream));
    return rc;
}

}  // namespace

extern "C" {

const char* sep_last_error(void) { return sep::g_err; }

const char* sep_version(void)
{
    return "libsep_amd 0.1 gfx950 (systolic wave-order Givens apply, device Wilkinson chase, "
           "batched solver, FP64-MFMA backward)";
}

int sep_create(sep_handle** out, int device)
{
    if (!out) {
        set_error("sep_create: null output pointer");
        return SEP_ERR_ARG;
    }
    *out = nullptr;
    int count = 0;
    SEP_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) {
        set_error("sep_create: device %d not present (%d devices)", device, count);
        return SEP_ERR_ARG;
    }
    SEP_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    SEP_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error("sep_create: device %d is %s; this library is built for gfx950 only", device,
                  prop.gcnArchName);
        return SEP_ERR_UNSUPPORTED;
    }
    if (int rc = scratch_init(device)) return rc;
    sep_handle* h = new sep_handle();
    h->device = device;
    h->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    int xcc = 0;
    if (hipDeviceGetAttribute(&xcc, hipDeviceAttributeNumberOfXccs, device) == hipSuccess &&
        xcc > 0)
        h->num_xcds = xcc;
    if (const char* ap = std::getenv("SEP_APPLY")) {
        // apply geometry (unset: chosen by the number of rows and layers)
        static const struct { const char* name; int geom; } kGeoms[] = {
            {"wide", 2}, {"small", 5}, {"r16deep", 7}, {"r16s16", 8}, {"r16s24", 10}};
        for (const auto& g : kGeoms)
            if (std::strcmp(ap, g.name) == 0) h->apply_geom = g.geom;
    }
    if (const char* bg = std::getenv("SEP_BATCHED_GRAPH")) h->batch_eager = std::strcmp(bg, "0") == 0;
    // Blocking streams: work on them is ordered after earlier work on the legacy default
    // stream and before later work there, so device arrays a caller produced or consumes
    // on the default stream need no extra synchronisation.
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamDefault);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->chase_stream, hipStreamDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_a, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_b, hipEventDisableTiming);
    if (e != hipSuccess) {
        set_error("sep_create: %s", hipGetErrorString(e));
        sep_destroy(h);
        return SEP_ERR_HIP;
    }
    *out = h;
    return SEP_OK;
}

int sep_destroy(sep_handle* h)
{
    if (!h) return SEP_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->chase_stream) (void)hipStreamSynchronize(h->chase_stream);
    h->host_io_A.release();
    h->host_io_cs.release();
    h->host_io_d.release();
    h->host_io_e.release();
    for (auto& w : h->apply_work) {
        w.table.release();
        w.aux.release();
        w.sync.release();
    }
    for (auto& g : h->grid) g.release();
    h->state.release();
    h->counters.release();
    h->lt.release();
    h->batch_meta.release();
    h->batch_scratch.release();
    h->bwd_meta.release();
    h->bwd_M.release();
    h->bwd_part.release();
    h->trd_vec.release();
    h->trd_part.release();
    h->trd_blk.release();
    h->sort_rank.release();
    release_blas(h);
    for (auto e : h->tev) (void)hipEventDestroy(e);
    if (h->batch_graph) (void)hipGraphExecDestroy(h->batch_graph);
    for (auto bs : h->batch_stream)
        if (bs) {
            (void)hipStreamSynchronize(bs);
            (void)hipStreamDestroy(bs);
        }
    for (auto e : h->batch_ev)
        if (e) (void)hipEventDestroy(e);
    if (h->batch_pinned) (void)hipHostFree(h->batch_pinned);
    if (h->solve_pinned) (void)hipHostFree(h->solve_pinned);
    for (auto e : h->solve_ev)
        if (e) (void)hipEventDestroy(e);
    if (h->ev_a) (void)hipEventDestroy(h->ev_a);
    if (h->ev_b) (void)hipEventDestroy(h->ev_b);
    if (h->chase_stream) (void)hipStreamDestroy(h->chase_stream);
    if (h->stream && h->own_stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return SEP_OK;
}

int sep_set_timing(sep_handle* h, int enable)
{
    if (!h) {
        set_error("sep_set_timing: null handle");
        return SEP_ERR_ARG;
    }
    if (int rc = use_device(h)) return rc;
    SEP_HIP(hipStreamSynchronize(h->stream));
    h->timing = enable != 0;
    h->tapply.clear();
    h->tprep.clear();
    h->tev_used = 0;
    return SEP_OK;
}

int sep_get_timing(sep_handle* h, double* apply_ms, double* prep_ms, int64_t* launches)
{
    if (!h) {
        set_error("sep_get_timing: null handle");
        return SEP_ERR_ARG;
    }
    if (int rc = use_device(h)) return rc;
    SEP_HIP(hipStreamSynchronize(h->stream));
    double a = 0, p = 0;
    for (auto& pr : h->tapply) {
        float ms = 0;
        SEP_HIP(hipEventElapsedTime(&ms, h->tev[pr.first], h->tev[pr.second]));
        a += ms;
    }
    for (auto& pr : h->tprep) {
        float ms = 0;
        SEP_HIP(hipEventElapsedTime(&ms, h->tev[pr.first], h->tev[pr.second]));
        p += ms;
    }
    if (apply_ms) *apply_ms = a;
    if (prep_ms) *prep_ms = p;
    if (launches) *launches = (int64_t)h->tapply.size();
    return SEP_OK;
}

int sep_set_stream(sep_handle* h, void* stream)
{
    if (!h) {
        set_error("sep_set_stream: null handle");
        return SEP_ERR_ARG;
    }
    if (int rc = use_device(h)) return rc;
    if (h->own_stream && h->stream) SEP_HIP(hipStreamSynchronize(h->stream));
    if (stream) {
        if (h->own_stream && h->stream) SEP_HIP(hipStreamDestroy(h->stream));
        h->stream = static_cast<hipStream_t>(stream);
        h->own_stream = false;
    } else if (!h->own_stream) {
        SEP_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamDefault));
        h->own_stream = true;
    }
    return SEP_OK;
}

int sep_synchronize(sep_handle* h)
{
    if (!h) {
        set_error("sep_synchronize: null handle");
        return SEP_ERR_ARG;
    }
    if (int rc = use_device(h)) return rc;
    SEP_HIP(hipStreamSynchronize(h->stream));
    SEP_HIP(hipStreamSynchronize(h->chase_stream));
    return SEP_OK;
}

int sep_bulk_rmul_f64(sep_handle* h, int64_t m, int64_t n, double* A, int64_t lda,
                      const double* cs, int64_t k)
{
    return bulk_rmul_host<double>(h, m, n, A, lda, cs, k);
}
int sep_bulk_rmul_f32(sep_handle* h, int64_t m, int64_t n, float* A, int64_t lda, const float* cs,
                      int64_t k)
{
    return bulk_rmul_host<float>(h, m, n, A, lda, cs, k);
}
int sep_bulk_rmul_f64_dev(sep_handle* h, int64_t m, int64_t n, double* A, int64_t lda,
                          const double* cs, int64_t k)
{
    return bulk_rmul_dev_checked<double>(h, m, n, A, lda, cs, k);
}
int sep_bulk_rmul_f64_dev_noalloc(sep_handle* h, int64_t m, int64_t n, double* A, int64_t lda,
                                  const double* cs, int64_t k)
{
    if (!h) {
        set_error("sep_bulk_rmul_f64_dev_noalloc: null handle");
        return SEP_ERR_ARG;
    }
    if (m < 0 || n < 0 || k < 0 || lda < (m > 1 ? m : 1)) {
        set_error("sep_bulk_rmul_f64_dev_noalloc: invalid sizes m=%lld n=%lld k=%lld lda=%lld",
                  (long long)m, (long long)n, (long long)k, (long long)lda);
        return SEP_ERR_ARG;
    }
    if (m == 0 || n <= 1 || k == 0) return SEP_OK;
    if (!A || !cs) {
        set_error("sep_bulk_rmul_f64_dev_noalloc: null pointer");
        return SEP_ERR_ARG;
    }
    if (int rc = use_device(h)) return rc;
    return bulk_apply_noalloc_f64(h, m, n, A, lda, cs, k);
}
int sep_bulk_rmul_f32_dev(sep_handle* h, int64_t m, int64_t n, float* A, int64_t lda,
                          const float* cs, int64_t k)
{
    return bulk_rmul_dev_checked<float>(h, m, n, A, lda, cs, k);
}

int sep_qr_algorithm_f64(sep_handle* h, int64_t n, double* dv, double* ev, double* Q,
                         int64_t ldq, int64_t max_layers, double tol, int64_t maxiter,
                         int32_t* converged, int64_t* iter, int64_t* n_rotations)
{
    return qr_host<double>(h, n, dv, ev, Q, ldq, max_layers, tol, maxiter, converged, iter,
                           n_rotations);
}
int sep_qr_algorithm_f32(sep_handle* h, int64_t n, float* dv, float* ev, float* Q, int64_t ldq,
                         int64_t max_layers, float tol, int64_t maxiter, int32_t* converged,
                         int64_t* iter, int64_t* n_rotations)
{
    return qr_host<float>(h, n, dv, ev, Q, ldq, max_layers, tol, maxiter, converged, iter,
                          n_rotations);
}
int sep_qr_algorithm_f64_dev(sep_handle* h, int64_t n, double* dv, double* ev, double* Q,
                             int64_t ldq, int64_t max_layers, double tol, int64_t maxiter,
                             int32_t* converged, int64_t* iter, int64_t* n_rotations)
{
    return qr_checked<double>(h, n, dv, ev, Q, ldq, max_layers, tol, maxiter, converged, iter,
                              n_rotations);
}
int sep_qr_algorithm_f32_dev(sep_handle* h, int64_t n, float* dv, float* ev, float* Q,
                             int64_t ldq, int64_t max_layers, float tol, int64_t maxiter,
                             int32_t* converged, int64_t* iter, int64_t* n_rotations)
{
    return qr_checked<float>(h, n, dv, ev, Q, ldq, max_layers, tol, maxiter, converged, iter,
                             n_rotations);
}

int sep_qr_batched_f64_dev(sep_handle* h, int64_t B, const int64_t* n, const int64_t* off_d,
                           const int64_t* off_q, double* dv, double* ev, double* Q,
                           int64_t max_layers, int32_t* converged, int64_t* iters)
{
    if (!h) {
        set_error("sep_qr_batched: null handle");
        return SEP_ERR_ARG;
    }
    if (int rc = use_device(h)) return rc;
    return qr_batched_f64(h, B, n, off_d, off_q, dv, ev, Q, max_layers, converged, iters);
}

int sep_tune_batched_f64(sep_handle* h, int64_t B, const int64_t* n, const int64_t* off_d,
                         const int64_t* off_q, const double* dv, const double* ev,
                         int32_t* choice)
{
    if (!h) {
        set_error("sep_tune_batched: null handle");
        return SEP_ERR_ARG;
    }
    if (int rc = use_device(h)) return rc;
    return tune_batched_f64(h, B, n, off_d, off_q, dv, ev, choice);
}

int sep_set_batched_config(sep_handle* h, int32_t fused_depth, int32_t rows_per_lane)
{
    if (!h) {
        set_error("sep_set_batched_config: null handle");
        return SEP_ERR_ARG;
    }
    return set_batched_config(h, fused_depth, rows_per_lane);
}

int sep_set_chase_rotations(sep_handle* h, int32_t mode)
{
    if (!h) {
        set_error("sep_set_chase_rotations: null handle");
        return SEP_ERR_ARG;
    }
    if (mode != SEP_CHASE_EXACT && mode != SEP_CHASE_FAST) {
        set_error("sep_set_chase_rotations: mode %d is neither SEP_CHASE_EXACT nor "
                  "SEP_CHASE_FAST", (int)mode);
        return SEP_ERR_ARG;
    }
    h->chase_fast = mode == SEP_CHASE_FAST;
    return SEP_OK;
}

int sep_set_odd_order(sep_handle* h, int32_t allow)
{
    if (!h) {
        set_error("sep_set_odd_order: null handle");
        return SEP_ERR_ARG;
    }
    h->allow_odd = allow != 0;
    return SEP_OK;
}

int sep_eig_backward_f64_dev(sep_handle* h, int64_t n, const double* lam, const double* Q,
                             int64_t ldq, const double* lam_bar, const double* Q_bar,
                             double* d_bar, double* e_bar)
{
    if (!h) {
        set_error("sep_eig_backward: null handle");
        return SEP_ERR_ARG;
    }
    if (int rc = use_device(h)) return rc;
    return eig_backward_f64(h, n, lam, Q, ldq, lam_bar, Q_bar, d_bar, e_bar);
}

int sep_eig_backward_batched_f64_dev(sep_handle* h, int64_t B, const int64_t* n,
                                     const int64_t* off_d, const int64_t* off_q,
                                     const double* lam, const double* Q, const double* lam_bar,
                                     const double* Q_bar, double* d_bar, double* e_bar)
{
    if (!h) {
        set_error("sep_eig_backward_batched: null handle");
        return SEP_ERR_ARG;
    }
    if (int rc = use_device(h)) return rc;
    return eig_backward_batched_f64(h, B, n, off_d, off_q, lam, Q, lam_bar, Q_bar, d_bar, e_bar);
}

int sep_sort_eigen_f64_dev(sep_handle* h, int64_t n, const double* lam, const double* Q,
                           int64_t ldq, double* lam_out, double* Q_out, int64_t ldo,
                           int64_t* perm)
{
    if (!h) {
        set_error("sep_sort_eigen: null handle");
        return SEP_ERR_ARG;
    }
    if (int rc = use_device(h)) return rc;
    return sort_eigen_dev<double>(h, n, lam, Q, ldq, lam_out, Q_out, ldo, perm);
}

int sep_sort_eigen_f32_dev(sep_handle* h, int64_t n, const float* lam, const float* Q,
                           int64_t ldq, float* lam_out, float* Q_out, int64_t ldo, int64_t* perm)
{
    if (!h) {
        set_error("sep_sort_eigen: null handle");
        return SEP_ERR_ARG;
    }
    if (int rc = use_device(h)) return rc;
    return sort_eigen_dev<float>(h, n, lam, Q, ldq, lam_out, Q_out, ldo, perm);
}

int sep_sytrd_f64_dev(sep_handle* h, int64_t n, double* A, int64_t lda, double* d, double* e,
                      double* tau)
{
    if (!h) {
        set_error("sep_sytrd: null handle");
        return SEP_ERR_ARG;
    }
    if (int rc = use_device(h)) return rc;
    return sytrd_f64(h, n, A, lda, d, e, tau);
}

int sep_orgtr_f64_dev(sep_handle* h, int64_t n, const double* A, int64_t lda, const double* tau,
                      double* Q, int64_t ldq)
{
    if (!h) {
        set_error("sep_orgtr: null handle");
        return SEP_ERR_ARG;
    }
    if (int rc = use_device(h)) return rc;
    return orgtr_f64(h, n, A, lda, tau, Q, ldq);
}

}  // extern "C"
