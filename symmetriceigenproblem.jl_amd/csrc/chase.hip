// chase.hip -- qr_algorithm! (src/bulge_chasing.jl:146-288) on one GPU.
//
// Host loop over blocks of at most max_layers sweeps.  Block b is chased by a
// one-wavefront kernel on the handle's chase stream into grid buffer b % 2, then
// applied to Q by the systolic apply on the apply stream, while block b+1 is already
// being chased into the other grid buffer.  The host reads the 16-word block state
// (pinned) after each chase to decide whether the solve continues, exactly where the
// reference's outer loop tests did_a_thing / iter > maxiter.
#include "chase.hpp"
#include "sep_internal.hpp"

namespace sep {
namespace {

template <typename T>
__global__ void __launch_bounds__(256) k_fill_identity(T* __restrict__ grid, int64_t pairs)
{
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < pairs;
         q += (int64_t)gridDim.x * blockDim.x) {
        grid[2 * q] = T(1);
        grid[2 * q + 1] = T(0);
    }
}

// Problems up to this order keep dv, ev and layer_thickness in LDS for the whole block
// (20 bytes per index in Float64: 120 KiB beside the chase's 24.6 KiB of chunk staging).
// The chase's own memory traffic is a chain of dependent round trips -- the deflation scan
// from `to` downwards, the thickness maximum, the shift's and the first rotation's
// operands, each chunk's inputs -- which take ~0.1 us each from LDS against ~0.7 us from
// L2; only the (c, s) grid goes to global memory.  Larger problems chase in global memory.
constexpr int kChaseLdsN = 6144;

template <typename T, bool kFast, bool kLds>
__global__ void __launch_bounds__(64) k_chase_block(T* dv, T* ev, int* lt, T* grid, int n,
                                                    int max_layers, T tol, int64_t maxiter,
                                                    int64_t* st)
{
    // the chase is the solve's critical path: its wave issues ahead of the apply of the
    // previous block running beside it
    __builtin_amdgcn_s_setprio(3);
    const int lane = (int)threadIdx.x;
    if constexpr (kLds) {
        // dynamic LDS (chase_lds_bytes(n)): it lies after the chase's static chunk staging,
        // whose per-step accesses then keep small immediate offsets
        extern __shared__ __align__(16) unsigned char chase_lds[];
        T* s_dv = reinterpret_cast<T*>(chase_lds);
        T* s_ev = s_dv + n;
        int* s_lt = reinterpret_cast<int*>(s_ev + n);
        for (int q = lane; q < n; q += 64) s_dv[q] = dv[q];
        for (int q = lane; q < n - 1; q += 64) s_ev[q] = ev[q];
        __syncthreads();
        // layer_thickness lives only within a block (chase_block resets it), so the LDS
        // copy needs neither loading nor writing back
        chase::Problem<T> P{s_dv, s_ev, s_lt, grid, n, max_layers, (int64_t)(n - 1)};
        chase::chase_block<T, true, kFast, true>(P, tol, maxiter, st, lane);
        __syncthreads();
        for (int q = lane; q < n; q += 64) dv[q] = s_dv[q];
        for (int q = lane; q < n - 1; q += 64) ev[q] = s_ev[q];
    } else {
        chase::Problem<T> P{dv, ev, lt, grid, n, max_layers, (int64_t)(n - 1)};
        chase::chase_block<T, true, kFast>(P, tol, maxiter, st, lane);
    }
}

// dv, ev (n entries each) and layer_thickness of the LDS-resident chase
template <typename T>
constexpr size_t chase_lds_bytes(int n)
{
    return (2 * sizeof(T) + sizeof(int)) * (size_t)n;
}

template <typename T, bool kFast>
void launch_chase(hipStream_t s, T* dv, T* ev, int* lt, T* grid, int n, int max_layers, T tol,
                  int64_t maxiter, int64_t* st)
{
    if (n <= kChaseLdsN)
        hipLaunchKernelGGL((k_chase_block<T, kFast, true>), dim3(1), dim3(64),
                           (unsigned)chase_lds_bytes<T>(n), s, dv, ev, lt, grid, n, max_layers,
                           tol, maxiter, st);
    else
        hipLaunchKernelGGL((k_chase_block<T, kFast, false>), dim3(1), dim3(64), 0, s, dv, ev, lt,
                           grid, n, max_layers, tol, maxiter, st);
}

template <typename T>
int fill_identity(hipStream_t s, T* grid, int64_t pairs)
{
    const int blocks = (int)std::min<int64_t>((pairs + 255) / 256, 4096);
    hipLaunchKernelGGL((k_fill_identity<T>), dim3(blocks), dim3(256), 0, s, grid, pairs);
    SEP_CHECK_LAUNCH();
    return SEP_OK;
}

}  // namespace

// The solve's pinned state words (two sets) and its four events, created on the handle's
// first solve and kept until sep_destroy.
static int solve_resources(sep_handle* h)
{
    if (!h->solve_pinned)
        SEP_HIP(hipHostMalloc((void**)&h->solve_pinned, sizeof(int64_t) * 2 * chase::ST_WORDS,
                              hipHostMallocDefault));
    for (auto& e : h->solve_ev)
        if (!e) SEP_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return SEP_OK;
}

template <typename T>
int qr_algorithm_dev(sep_handle* h, int64_t n64, T* dv, T* ev, T* Q, int64_t ldq,
                     int64_t max_layers, T tol, int64_t maxiter, int32_t* converged,
                     int64_t* iter_out, int64_t* nrot_out)
{
    if (n64 > INT32_MAX / 2) {
        set_error("qr_algorithm: n=%lld too large", (long long)n64);
        return SEP_ERR_ARG;
    }
    const int n = (int)n64;
    if (n <= 1) {  // nothing to chase (the reference's loop breaks at once)
        if (converged) *converged = 1;
        if (iter_out) *iter_out = 0;
        if (nrot_out) *nrot_out = 0;
        return SEP_OK;
    }
    hipStream_t cs = h->chase_stream, as = h->stream;
    const int64_t pairs = (int64_t)(n - 1) * max_layers;
    for (int b = 0; b < 2; ++b)
        if (int rc = h->grid[b].reserve(sizeof(T) * 2 * (size_t)pairs)) return rc;
    if (int rc = h->lt.reserve(sizeof(int) * (size_t)n)) return rc;
    if (int rc = h->state.reserve(sizeof(int64_t) * chase::ST_WORDS)) return rc;
    if (int rc = solve_resources(h)) return rc;
    int64_t* const hst = h->solve_pinned;
    int64_t* dst = h->state.as<int64_t>();

    // the chase must see dv/ev as the caller's stream left them
    SEP_HIP(hipEventRecord(h->ev_b, as));
    SEP_HIP(hipStreamWaitEvent(cs, h->ev_b, 0));

    int64_t init[chase::ST_WORDS] = {0};
    init[chase::ST_TO] = n;
    init[chase::ST_LBEA] = n;
    init[chase::ST_ITER] = 0;
    // the previous solve's last read of the pinned words has completed (it synchronised)
    std::memcpy(hst, init, sizeof(init));
    SEP_HIP(hipMemcpyAsync(dst, hst, sizeof(init), hipMemcpyHostToDevice, cs));

    hipEvent_t applied[2] = {h->solve_ev[0], h->solve_ev[1]};
    hipEvent_t chased[2] = {h->solve_ev[2], h->solve_ev[3]};
    bool applied_pending[2] = {false, false};
    // the fast rotation is a Float64 option (Float32 chases always take the exact one)
    const bool fast = std::is_same<T, double>::value && h->chase_fast;

    // Block b chases into grid[b % 2] and copies its state words to hst[b % 2].  It starts
    // from the state words block b - 1 left on the device, so the host enqueues block b + 1
    // before it waits for block b's words: the chase stream runs block after block without
    // a host round trip between them (n = 2000: 19 blocks, whose gaps fell from 148 to
    // 26 us on average, 2.7 to 0.5 ms per solve, in kernel traces).  The block
    // enqueued after the last one changes nothing: after a block that found nothing to
    // chase, `to` is at the start and the next one stops at once; after one that stopped
    // at maxiter, the next one stops before its first scan.
    auto enqueue_block = [&](int slot) -> int {
        T* grid = h->grid[slot].as<T>();
        if (applied_pending[slot])  // the apply of two blocks ago still reads this grid
            SEP_HIP(hipStreamWaitEvent(cs, applied[slot], 0));
        if (int rc = fill_identity<T>(cs, grid, pairs)) return rc;
        if (fast)
            launch_chase<T, true>(cs, dv, ev, h->lt.as<int>(), grid, n, (int)max_layers, tol,
                                  maxiter, dst);
        else
            launch_chase<T, false>(cs, dv, ev, h->lt.as<int>(), grid, n, (int)max_layers, tol,
                                   maxiter, dst);
        SEP_CHECK_LAUNCH();
        SEP_HIP(hipMemcpyAsync(hst + slot * chase::ST_WORDS, dst,
                               sizeof(int64_t) * chase::ST_WORDS, hipMemcpyDeviceToHost, cs));
        SEP_HIP(hipEventRecord(chased[slot], cs));
        return SEP_OK;
    };

    int64_t total_rot = 0, iter = 0;
    int conv = 1;
    int cur = 0;
    int rc = enqueue_block(0);
    while (rc == SEP_OK) {
        if ((rc = enqueue_block(cur ^ 1))) break;
        if (hipEventSynchronize(chased[cur]) != hipSuccess) {
            set_error("chase: state read-back failed: %s", hipGetErrorString(hipGetLastError()));
            rc = SEP_ERR_HIP;
            break;
        }
        const int64_t* st = hst + cur * chase::ST_WORDS;
        iter = st[chase::ST_ITER];
        if (st[chase::ST_EARLY]) {  // :180-182 -- return without applying this block
            conv = 0;
            break;
        }
        if (!st[chase::ST_DID]) break;  // :279
        total_rot += st[chase::ST_NROT];
        if (Q) {
            const int64_t ilo = st[chase::ST_ILO], ihi = st[chase::ST_IHI];
            const int64_t depth = st[chase::ST_MAXLAYER];
            if (hipStreamWaitEvent(as, chased[cur], 0) != hipSuccess) { rc = SEP_ERR_HIP; break; }
            // rotations ilo..ihi (1-based) touch columns ilo-1 .. ihi (0-based)
            rc = bulk_apply_dev<T>(h, as, h->apply_work[cur], n, ihi - ilo + 2,
                                   Q + (ilo - 1) * ldq, ldq, h->grid[cur].as<T>() + 2 * (ilo - 1),
                                   n - 1, depth, /*rotations=*/true);
            if (rc) break;
            if (hipEventRecord(applied[cur], as) != hipSuccess) { rc = SEP_ERR_HIP; break; }
            applied_pending[cur] = true;
        }
        cur ^= 1;
    }
    // the caller's stream (apply stream) must also see the final dv/ev (and the block
    // enqueued after the last one has finished with them)
    (void)hipEventRecord(chased[0], cs);
    (void)hipStreamWaitEvent(as, chased[0], 0);
    hipError_t se = hipStreamSynchronize(as);
    if (se != hipSuccess && rc == SEP_OK) {
        set_error("qr_algorithm: %s", hipGetErrorString(se));
        rc = SEP_ERR_HIP;
    }
    if (rc) return rc;
    if (converged) *converged = conv;
    if (iter_out) *iter_out = iter;
    if (nrot_out) *nrot_out = total_rot;
    return conv ? SEP_OK : SEP_NOT_CONVERGED;
}

template int qr_algorithm_dev<double>(sep_handle*, int64_t, double*, double*, double*, int64_t,
                                      int64_t, double, int64_t, int32_t*, int64_t*, int64_t*);
template int qr_algorithm_dev<float>(sep_handle*, int64_t, float*, float*, float*, int64_t,
                                     int64_t, float, int64_t, int32_t*, int64_t*, int64_t*);

}  // namespace sep
