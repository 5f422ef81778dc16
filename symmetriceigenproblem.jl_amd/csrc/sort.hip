// sort.hip -- eigenvalues in ascending order with the eigenvector columns permuted to
// match (the sorted-output convenience beside eigen_decomp!, src/bulge_chasing.jl:290-299;
// qr_algorithm! itself returns the eigenvalues in deflation order).
//
// The order is a total order on the bit patterns (IEEE order for numbers, -0 before +0,
// NaNs with the sign bit set first and the others last), ties broken by index, so the
// result is a permutation for any input.  k_rank gives each eigenvalue its rank by
// counting the smaller keys (n^2 / 2 comparisons spread over the chip; n = 16384 is
// 1.3e8 compares, tens of microseconds), k_gather moves eigenvalue j and column j of Q to
// position rank[j] of the outputs: one read and one write of Q, coalesced down columns.
#include "sep_internal.hpp"

namespace sep {
namespace {

template <typename T>
__device__ __forceinline__ uint64_t order_key(T x)
{
    uint64_t b;
    if constexpr (sizeof(T) == 8)
        b = __builtin_bit_cast(uint64_t, x);
    else
        b = (uint64_t)__builtin_bit_cast(uint32_t, x) << 32;
    return (b >> 63) ? ~b : (b | (1ull << 63));
}

// rank[j] = #{i : key_i < key_j or (key_i == key_j and i < j)}; one thread per j, the keys
// staged through LDS in tiles of 256.
template <typename T>
__global__ void __launch_bounds__(256) k_rank(const T* __restrict__ lam, int n,
                                              int* __restrict__ rank)
{
    __shared__ uint64_t tile[256];
    const int j = blockIdx.x * 256 + threadIdx.x;
    const uint64_t kj = j < n ? order_key(lam[j]) : 0;
    int r = 0;
    for (int t0 = 0; t0 < n; t0 += 256) {
        const int i = t0 + threadIdx.x;
        tile[threadIdx.x] = i < n ? order_key(lam[i]) : ~0ull;
        __syncthreads();
        const int lim = min(256, n - t0);
        for (int u = 0; u < lim; ++u) {
            const uint64_t ki = tile[u];
            r += (ki < kj) || (ki == kj && t0 + u < j);
        }
        __syncthreads();
    }
    if (j < n) rank[j] = r;
}

// Column j of Q (and lam[j]) to position rank[j]; perm[rank[j]] = j.  Block (x, y): rows
// 256 x .. 256 x + 255 of columns y, y + gridDim.y, ...
template <typename T>
__global__ void __launch_bounds__(256) k_gather(const T* __restrict__ lam, const T* __restrict__ Q,
                                                int64_t ldq, const int* __restrict__ rank,
                                                int n, T* __restrict__ lam_out,
                                                T* __restrict__ Q_out, int64_t ldo,
                                                int64_t* __restrict__ perm)
{
    const int row = blockIdx.x * 256 + threadIdx.x;
    for (int j = blockIdx.y; j < n; j += gridDim.y) {
        const int r = rank[j];
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            lam_out[r] = lam[j];
            if (perm) perm[r] = j;
        }
        if (Q && row < n) Q_out[(int64_t)r * ldo + row] = Q[(int64_t)j * ldq + row];
    }
}

}  // namespace

template <typename T>
int sort_eigen_dev(sep_handle* h, int64_t n, const T* lam, const T* Q, int64_t ldq, T* lam_out,
                   T* Q_out, int64_t ldo, int64_t* perm)
{
    if (n < 0 || n > (INT32_MAX / 2) || (n > 0 && (!lam || !lam_out)) ||
        (Q && (!Q_out || ldq < std::max<int64_t>(n, 1) || ldo < std::max<int64_t>(n, 1)))) {
        set_error("sep_sort_eigen: invalid n=%lld, pointer or leading dimension", (long long)n);
        return SEP_ERR_ARG;
    }
    if (n == 0) return SEP_OK;
    if (Q && ((Q_out < Q + ldq * n && Q < Q_out + ldo * n))) {
        set_error("sep_sort_eigen: Q_out overlaps Q (the permutation is out of place)");
        return SEP_ERR_ARG;
    }
    if (int rc = h->sort_rank.reserve(sizeof(int) * (size_t)n)) return rc;
    int* rank = h->sort_rank.as<int>();
    hipStream_t st = h->stream;
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_rank<T>, dim3(nb), dim3(256), 0, st, lam, (int)n, rank);
    SEP_CHECK_LAUNCH();
    const unsigned cols = (unsigned)std::min<int64_t>(n, 4096);
    hipLaunchKernelGGL(k_gather<T>, dim3(Q ? nb : 1, cols), dim3(256), 0, st, lam, Q, ldq, rank,
                       (int)n, lam_out, Q_out, ldo, perm);
    SEP_CHECK_LAUNCH();
    return SEP_OK;
}

template int sort_eigen_dev<double>(sep_handle*, int64_t, const double*, const double*, int64_t,
                                    double*, double*, int64_t, int64_t*);
template int sort_eigen_dev<float>(sep_handle*, int64_t, const float*, const float*, int64_t,
                                   float*, float*, int64_t, int64_t*);

}  // namespace sep
