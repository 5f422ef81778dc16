/*
 * Oracle body, included twice by sep_oracle.c: once with T=double / SUF=_f64 and
 * once with T=float / SUF=_f32.  TEST INFRASTRUCTURE ONLY (see sep_oracle.c).
 *
 * Indices inside these functions are kept 1-based, exactly as in the Julia source,
 * through the DV/EV/LT/GIV accessors, so each line can be read against
 * src/bulge_chasing.jl line by line.
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SUF)

/* ---- LinearAlgebra.givensAlgorithm (classic LAPACK DLARTG) -----------------
 * Called from src/simple_givens.jl:20-23 (get_rotation), which is used at
 * src/bulge_chasing.jl:24 and :96.  Julia's stdlib is not on disk; this follows
 * the published DLARTG algorithm (LAPACK 3.2, the version Julia's stdlib mirrored
 * in the SIMDPirates era): g==0 -> (1,0,f); f==0 -> (0,1,g); otherwise
 * r = sqrt(f^2+g^2) with power-of-two rescaling for extreme magnitudes,
 * c = f/r, s = g/r, and if |f| > |g| and c < 0 all three are negated. */
void FN(orc_givens)(T f, T g, T *cs, T *sn, T *r)
{
    /* safmn2 = 2^trunc(log2(floatmin/eps)/2): 2^-485 (Float64), 2^-51 (Float32) */
    const T safmn2 = SAFMN2;
    const T safmx2 = (T)1 / safmn2;
    if (g == 0) {
        *cs = 1; *sn = 0; *r = f;
        return;
    }
    if (f == 0) {
        *cs = 0; *sn = 1; *r = g;
        return;
    }
    T f1 = f, g1 = g, rr, c, s;
    T scale = FABS(f1) > FABS(g1) ? FABS(f1) : FABS(g1);
    if (scale >= safmx2) {
        int count = 0;
        do {
            count++;
            f1 *= safmn2;
            g1 *= safmn2;
            scale = FABS(f1) > FABS(g1) ? FABS(f1) : FABS(g1);
        } while (scale >= safmx2 && count < 20);
        rr = SQRT(f1 * f1 + g1 * g1);
        c = f1 / rr;
        s = g1 / rr;
        for (int i = 0; i < count; i++) rr *= safmx2;
    } else if (scale <= safmn2) {
        int count = 0;
        do {
            count++;
            f1 *= safmx2;
            g1 *= safmx2;
            scale = FABS(f1) > FABS(g1) ? FABS(f1) : FABS(g1);
        } while (scale <= safmn2);
        rr = SQRT(f1 * f1 + g1 * g1);
        c = f1 / rr;
        s = g1 / rr;
        for (int i = 0; i < count; i++) rr *= safmn2;
    } else {
        rr = SQRT(f1 * f1 + g1 * g1);
        c = f1 / rr;
        s = g1 / rr;
    }
    if (FABS(f) > FABS(g) && c < 0) {
        c = -c; s = -s; rr = -rr;
    }
    *cs = c; *sn = s; *r = rr;
}

/* 1-based accessors matching the Julia indexing */
#define DV(i) dv[(i) - 1]
#define EV(i) ev[(i) - 1]
#define LT(i) lt[(i) - 1]
#define GIV_C(i, L) givens[2 * (((int64_t)(i) - 1) + (int64_t)(n - 1) * ((L) - 1))]
#define GIV_S(i, L) givens[2 * (((int64_t)(i) - 1) + (int64_t)(n - 1) * ((L) - 1)) + 1]

/* is_offdiagonal_small, src/bulge_chasing.jl:7-8 */
static inline int FN(is_small)(const T *dv, const T *ev, int64_t i, T tol)
{
    return FABS(EV(i)) <= tol * (FABS(DV(i)) + FABS(DV(i + 1)));
}

/* Chronological rotation log (optional): every rotation the chase creates, in
 * creation order, as (index, c, s).  Used by the tests to pin the per-index layer
 * packing of src/bulge_chasing.jl:61-63,137-139 against plain sequential order. */
typedef struct {
    int64_t *idx;
    T *c;
    T *s;
    int64_t len, cap;
} FN(orc_rotlog);

static inline void FN(log_rot)(FN(orc_rotlog) *log, int64_t i, T c, T s)
{
    if (!log || log->len >= log->cap) return;
    log->idx[log->len] = i;
    log->c[log->len] = c;
    log->s[log->len] = s;
    log->len++;
}

/* single_shift!, src/bulge_chasing.jl:11-144.  givens is the (n-1) x max_layers
 * interleaved (c,s) grid, column-major (a Julia Matrix{Tuple{T,T}}). */
void FN(orc_single_shift)(int64_t n, T *dv, T *ev, int64_t from, int64_t to, T mu,
                          T *givens, int64_t *lt, void *logp)
{
    FN(orc_rotlog) *log = (FN(orc_rotlog) *)logp;
    T c, s, nrm;
    /* :17-24 */
    T H11 = DV(from);
    T H21 = EV(from);
    T p1 = H11 - mu;
    T p2 = H21;
    FN(orc_givens)(p1, p2, &c, &s, &nrm);
    /* :27-47 */
    T H22 = DV(from + 1);
    /* x*y + z*w is evaluated as fma(x, y, z*w): single_shift! is @fastmath, which lets
     * the compiler contract; the device chase uses the identical contraction. */
    T H11p = FMA(c, H11, s * H21);
    T H21p = FMA(-s, H11, c * H21);
    T H12p = FMA(c, H21, s * H22);
    T H22p = FMA(-s, H21, c * H22);
    T H11pp = FMA(H11p, c, H12p * s);
    T H12pp = FMA(H11p, -s, H12p * c);
    T H22pp = FMA(H21p, -s, H22p * c);
    DV(from + 0) = H11pp;
    DV(from + 1) = H22pp;
    EV(from + 0) = H12pp;
    /* :50-59 */
    T bulge = 0;
    if (from + 2 <= to) {
        T H23 = EV(from + 1);
        T H13p = s * H23;
        T H23p = c * H23;
        bulge = H13p;
        EV(from + 1) = H23p;
    }
    /* :62-63 */
    int64_t layer = ++LT(from);
    GIV_C(from, layer) = c;
    GIV_S(from, layer) = s;
    FN(log_rot)(log, from, c, s);
    /* :91-140 */
    for (int64_t i = from + 1; i <= to - 1; i++) {
        p1 = EV(i - 1);
        p2 = bulge;
        FN(orc_givens)(p1, p2, &c, &s, &nrm);
        EV(i - 1) = nrm;
        T B11 = DV(i + 0);
        T B22 = DV(i + 1);
        T B21 = EV(i + 0);
        T B11p = FMA(c, B11, s * B21);
        T B21p = FMA(-s, B11, c * B21);
        T B12p = FMA(c, B21, s * B22);
        T B22p = FMA(-s, B21, c * B22);
        T B11pp = FMA(B11p, c, B12p * s);
        T B12pp = FMA(B11p, -s, B12p * c);
        T B22pp = FMA(B21p, -s, B22p * c);
        DV(i + 0) = B11pp;
        DV(i + 1) = B22pp;
        EV(i + 0) = B12pp;
        if (i + 2 <= to) {
            T B23 = EV(i + 1);
            T B13p = s * B23;
            T B23p = c * B23;
            bulge = B13p;
            EV(i + 1) = B23p;
        }
        layer = ++LT(i);
        GIV_C(i, layer) = c;
        GIV_S(i, layer) = s;
        FN(log_rot)(log, i, c, s);
    }
}

/* The Wilkinson shift of src/bulge_chasing.jl:239-265 (scaled trailing 2x2). */
T FN(orc_wilkinson_shift)(T H11, T H12, T H22)
{
    T scale = FABS(H11) + 2 * FABS(H12) + FABS(H22);
    H11 /= scale;
    H12 /= scale;
    H22 /= scale;
    T t = (H11 + H22) / 2;
    T d = (H11 - t) * (H22 - t) - H12 * H12;
    T sqrt_discr = SQRT(FABS(d));
    T l1 = t + sqrt_discr;
    T l2 = t - sqrt_discr;
    T lam = FABS(H22 - l1) < FABS(H22 - l2) ? l1 : l2;
    lam *= scale;
    return lam;
}

/* Rotation2 rmul! (src/simple_givens.jl:53-64) over all m rows:
 * a1' = a1 c + a2 s,  a2' = a1 (-s) + a2 c  on columns (i, i+1), i 1-based. */
static inline void FN(rmul_one)(int64_t m, T *A, int64_t lda, int64_t i, T c, T s)
{
    T *a1 = A + (i - 1) * lda;
    T *a2 = A + i * lda;
    for (int64_t j = 0; j < m; j++) {
        T x = a1[j], y = a2[j];
        a1[j] = x * c + y * s;
        a2[j] = x * -s + y * c;
    }
}

/* The reference's identity skip.  bulk_wave_order_2x2_rmul! (src/wave.jl:62-162) hands
 * the grid to its kernels in fused blocks and skips a block only when every s in it is
 * zero (src/wave.jl:83,98,116,137,152).  A rotation with s == 0 and c != 1 is therefore
 * applied (it scales or negates its columns) unless its whole block is s-zero.  The
 * block of rotation (i, L), 1-based: wave i' = start_col + j + 1 pairs layers j (odd) and
 * j + 1 and takes rows start_col + 1, start_col + 2 of layer j and start_col,
 * start_col + 1 of layer j + 1 (src/wave.jl:78-81), with start_col even, and the edge
 * kernels drop the rows outside 1..n-1 (src/wave.jl:92-95,131-134).  So for L odd the
 * block has start_col = 2 floor((i-1)/2), for L even start_col = 2 floor(i/2).  Every
 * (i, L) lies in exactly one block when the wave preconditions hold (n-1 odd, k even,
 * k <= n-2); the same blocks are used for any other shape. */
static int FN(block_skipped)(const T *cs, int64_t n, int64_t i, int64_t L, int64_t k)
{
    const int64_t j = (L % 2 == 1) ? L : L - 1;
    const int64_t sc = (L % 2 == 1) ? 2 * ((i - 1) / 2) : 2 * (i / 2);
    const int64_t rows[4] = {sc + 1, sc + 2, sc, sc + 1};
    const int64_t lays[4] = {j, j, j + 1, j + 1};
    for (int q = 0; q < 4; q++) {
        if (rows[q] < 1 || rows[q] > n - 1 || lays[q] > k) continue;
        if (cs[2 * ((rows[q] - 1) + (n - 1) * (lays[q] - 1)) + 1] != 0) return 0;
    }
    return 1;
}

/* Layer-major sequential application of an (n-1) x k grid: the product
 * A * prod_{L=1..k} prod_{i=1..n-1} G(i,L)^T that bulk_wave_order_2x2_rmul!
 * (src/wave.jl:62-162) computes, with its fused-block identity skip (block_skipped). */
void FN(orc_bulk_rmul_seq)(int64_t m, int64_t n, T *A, int64_t lda, const T *cs, int64_t k)
{
    for (int64_t L = 1; L <= k; L++)
        for (int64_t i = 1; i <= n - 1; i++) {
            T c = cs[2 * ((i - 1) + (n - 1) * (L - 1))];
            T s = cs[2 * ((i - 1) + (n - 1) * (L - 1)) + 1];
            if (FN(block_skipped)(cs, n, i, L, k)) continue;
            FN(rmul_one)(m, A, lda, i, c, s);
        }
}

/* The same layer-major product with the rotation evaluated as two fused multiply-adds,
 * a1' = fma(a1, c, a2 s), a2' = fma(a1, -s, a2 c): one reading of the reference's
 * @fastmath / SIMD contraction of `s * col3 + c * col2` (src/bulk_givens.jl:44-57), and
 * the form the batched GPU solver computes. */
static inline void FN(rmul_one_fma)(int64_t m, T *A, int64_t lda, int64_t i, T c, T s)
{
    T *a1 = A + (i - 1) * lda;
    T *a2 = A + i * lda;
    for (int64_t j = 0; j < m; j++) {
        T x = a1[j], y = a2[j];
        a1[j] = FMA(x, c, y * s);
        a2[j] = FMA(x, -s, y * c);
    }
}

void FN(orc_bulk_rmul_seq_fma)(int64_t m, int64_t n, T *A, int64_t lda, const T *cs, int64_t k)
{
    for (int64_t L = 1; L <= k; L++)
        for (int64_t i = 1; i <= n - 1; i++) {
            T c = cs[2 * ((i - 1) + (n - 1) * (L - 1))];
            T s = cs[2 * ((i - 1) + (n - 1) * (L - 1)) + 1];
            if (FN(block_skipped)(cs, n, i, L, k)) continue;
            FN(rmul_one_fma)(m, A, lda, i, c, s);
        }
}

/* Apply a chronological rotation log one rotation at a time. */
void FN(orc_apply_log)(int64_t m, T *A, int64_t lda, const int64_t *idx, const T *c,
                       const T *s, int64_t len)
{
    for (int64_t r = 0; r < len; r++) FN(rmul_one)(m, A, lda, idx[r], c[r], s[r]);
}

/* ---- wave order, src/wave.jl:62-162 + src/bulk_givens.jl ------------------ */

/* givens_kernel!(A, ::Fused2x2), src/bulk_givens.jl:4-103: rotations
 * 1=(L, i+1) on cols 2,3; 2=(L, i+2) on cols 3,4; 3=(L+1, i) on cols 1,2;
 * 4=(L+1, i+1) on cols 2,3, where col 1 is column i (1-based start_col+1). */
static void FN(kern_full)(int64_t m, T *A, int64_t lda, int64_t start_col, T c1, T s1, T c2,
                          T s2, T c3, T s3, T c4, T s4)
{
    T *k1 = A + (start_col - 1) * lda;
    T *k2 = k1 + lda, *k3 = k2 + lda, *k4 = k3 + lda;
    for (int64_t j = 0; j < m; j++) {
        T col1 = k1[j], col2 = k2[j], col3 = k3[j], col4 = k4[j];
        T col2p = s1 * col3 + c1 * col2;
        T col3p = c1 * col3 - s1 * col2;
        T col3pp = s2 * col4 + c2 * col3p;
        T col4p = c2 * col4 - s2 * col3p;
        T col1p = s3 * col2p + c3 * col1;
        T col2pp = c3 * col2p - s3 * col1;
        T col2ppp = s4 * col3pp + c4 * col2pp;
        T col3ppp = c4 * col3pp - s4 * col2pp;
        k1[j] = col1p;
        k2[j] = col2ppp;
        k3[j] = col3ppp;
        k4[j] = col4p;
    }
}

/* givens_kernel!(A, ::Fused2x2_remainder1), src/bulk_givens.jl:105-190 (left edge,
 * start_col = 0: columns 1..3, rotations 1, 2, 4). */
static void FN(kern_rem1)(int64_t m, T *A, int64_t lda, T c1, T s1, T c2, T s2, T c4, T s4)
{
    T *k2 = A, *k3 = A + lda, *k4 = A + 2 * lda;
    for (int64_t j = 0; j < m; j++) {
        T col2 = k2[j], col3 = k3[j], col4 = k4[j];
        T col2p = s1 * col3 + c1 * col2;
        T col3p = c1 * col3 - s1 * col2;
        T col3pp = s2 * col4 + c2 * col3p;
        T col4p = c2 * col4 - s2 * col3p;
        T col2ppp = s4 * col3pp + c4 * col2p;
        T col3ppp = c4 * col3pp - s4 * col2p;
        k2[j] = col2ppp;
        k3[j] = col3ppp;
        k4[j] = col4p;
    }
}

/* givens_kernel!(A, ::Fused2x2_remainder2), src/bulk_givens.jl:192-279 (right edge,
 * start_col = n-2 in wave.jl's numbering: columns start_col..start_col+2). */
static void FN(kern_rem2)(int64_t m, T *A, int64_t lda, int64_t start_col, T c1, T s1, T c3,
                          T s3, T c4, T s4)
{
    T *k1 = A + (start_col - 1) * lda;
    T *k2 = k1 + lda, *k3 = k2 + lda;
    for (int64_t j = 0; j < m; j++) {
        T col1 = k1[j], col2 = k2[j], col3 = k3[j];
        T col2p = s1 * col3 + c1 * col2;
        T col3p = c1 * col3 - s1 * col2;
        T col1p = s3 * col2p + c3 * col1;
        T col2pp = c3 * col2p - s3 * col1;
        T col2ppp = s4 * col3p + c4 * col2pp;
        T col3ppp = c4 * col3p - s4 * col2pp;
        k1[j] = col1p;
        k2[j] = col2ppp;
        k3[j] = col3ppp;
    }
}

#define GC(i, L) cs[2 * (((int64_t)(i) - 1) + gm * ((int64_t)(L) - 1))]
#define GS(i, L) cs[2 * (((int64_t)(i) - 1) + gm * ((int64_t)(L) - 1)) + 1]

/* bulk_wave_order_2x2_rmul!, src/wave.jl:62-162, on the m rows of A.  The grid has
 * gm = N-1 rows (must be odd) and k layers (must be even, k <= N-2).  Returns 0 or -1
 * on a violated precondition (the reference asserts / reads out of bounds). */
int FN(orc_bulk_wave)(int64_t m, int64_t N, T *A, int64_t lda, const T *cs, int64_t k)
{
    const int64_t gm = N - 1, gn = k;
    if (gm % 2 != 1 || gn % 2 != 0 || k > N - 2) return -1;
    const int64_t max_m = 2 * (gm / 2);
    /* Fan in, :73-103 */
    for (int64_t i = 2; i <= gn; i += 2) {
        for (int64_t j = 1; j <= i - 2; j += 2) {
            int64_t sc = i - j - 1;
            T c1 = GC(sc + 1, j), s1 = GS(sc + 1, j);
            T c2 = GC(sc + 2, j), s2 = GS(sc + 2, j);
            T c3 = GC(sc + 0, j + 1), s3 = GS(sc + 0, j + 1);
            T c4 = GC(sc + 1, j + 1), s4 = GS(sc + 1, j + 1);
            if (s1 == 0 && s2 == 0 && s3 == 0 && s4 == 0) continue;
            FN(kern_full)(m, A, lda, sc, c1, s1, c2, s2, c3, s3, c4, s4);
        }
        {
            T c1 = GC(1, i - 1), s1 = GS(1, i - 1);
            T c2 = GC(2, i - 1), s2 = GS(2, i - 1);
            T c4 = GC(1, i), s4 = GS(1, i);
            if (!(s1 == 0 && s2 == 0 && s4 == 0)) FN(kern_rem1)(m, A, lda, c1, s1, c2, s2, c4, s4);
        }
    }
    /* Fast zone, :106-122 */
    for (int64_t i = gn + 2; i <= max_m; i += 2) {
        for (int64_t j = 1; j <= gn; j += 2) {
            int64_t sc = i - j - 1;
            T c1 = GC(sc + 1, j), s1 = GS(sc + 1, j);
            T c2 = GC(sc + 2, j), s2 = GS(sc + 2, j);
            T c3 = GC(sc + 0, j + 1), s3 = GS(sc + 0, j + 1);
            T c4 = GC(sc + 1, j + 1), s4 = GS(sc + 1, j + 1);
            if (s1 == 0 && s2 == 0 && s3 == 0 && s4 == 0) continue;
            FN(kern_full)(m, A, lda, sc, c1, s1, c2, s2, c3, s3, c4, s4);
        }
    }
    /* Fan out, :125-158 */
    for (int64_t i = max_m + 2; i <= gm + gn; i += 2) {
        {
            int64_t sc = max_m;
            int64_t layer = i - 1 - max_m;
            T c1 = GC(sc + 1, layer), s1 = GS(sc + 1, layer);
            T c3 = GC(sc + 0, layer + 1), s3 = GS(sc + 0, layer + 1);
            T c4 = GC(sc + 1, layer + 1), s4 = GS(sc + 1, layer + 1);
            if (!(s1 == 0 && s3 == 0 && s4 == 0))
                FN(kern_rem2)(m, A, lda, sc, c1, s1, c3, s3, c4, s4);
        }
        for (int64_t j = (i + 1 - max_m); j <= gn; j += 2) {
            int64_t sc = i - j - 1;
            T c1 = GC(sc + 1, j), s1 = GS(sc + 1, j);
            T c2 = GC(sc + 2, j), s2 = GS(sc + 2, j);
            T c3 = GC(sc + 0, j + 1), s3 = GS(sc + 0, j + 1);
            T c4 = GC(sc + 1, j + 1), s4 = GS(sc + 1, j + 1);
            if (s1 == 0 && s2 == 0 && s3 == 0 && s4 == 0) continue;
            FN(kern_full)(m, A, lda, sc, c1, s1, c2, s2, c3, s3, c4, s4);
        }
    }
    return 0;
}
#undef GC
#undef GS

/* bulk_wave_order_2x2_rmul_parallel!, src/wave.jl:36-60: p row chunks, each a
 * multiple of the vector width W except the last, each running the serial wave. */
typedef struct {
    int64_t m, N, lda, k;
    T *A;
    const T *cs;
} FN(wave_job);

static void *FN(wave_worker)(void *arg)
{
    FN(wave_job) *j = (FN(wave_job) *)arg;
    if (j->m > 0) FN(orc_bulk_wave)(j->m, j->N, j->A, j->lda, j->cs, j->k);
    return NULL;
}

int FN(orc_bulk_wave_parallel)(int64_t m, int64_t N, T *A, int64_t lda, const T *cs, int64_t k,
                               int p)
{
    if ((N - 1) % 2 != 1 || k % 2 != 0 || k > N - 2) return -1;
    if (p <= 1) return FN(orc_bulk_wave)(m, N, A, lda, cs, k);
    const int64_t W = VEC_W;
    const int64_t part = W * ((m / p) / W);
    pthread_t th[256];
    FN(wave_job) jobs[256];
    if (p > 256) p = 256;
    for (int i = 1; i <= p; i++) {
        int64_t from = 1 + part * (i - 1);
        int64_t to = i == p ? m : 1 + part * i - 1;
        jobs[i - 1].m = to - from + 1;
        jobs[i - 1].N = N;
        jobs[i - 1].lda = lda;
        jobs[i - 1].k = k;
        jobs[i - 1].A = A + (from - 1);
        jobs[i - 1].cs = cs;
        pthread_create(&th[i - 1], NULL, FN(wave_worker), &jobs[i - 1]);
    }
    for (int i = 0; i < p; i++) pthread_join(th[i], NULL);
    return 0;
}

/* qr_algorithm!, src/bulge_chasing.jl:146-288.  dv (n), ev (n-1), Q column-major
 * (m rows = n, leading dimension ldq; Q may be NULL for eigenvalues only, which the
 * reference does not offer but which leaves dv/ev identical).  Return value 0 on
 * success, -1 for odd n (the @assert at :155) unless allow_odd (the assert is the
 * reference's "limitation for now": nothing but its wave kernels' preconditions
 * depends on the parity, and apply modes 0, 2 and 3 have none), -2 for max_layers < 1,
 * -3 when apply mode 1 meets a grid outside the wave preconditions, -4 for n < 1 (the
 * reference cannot allocate its (n-1) x max_layers grid there); *converged / *iter as the
 * reference's return tuple; *nrot counts the rotations the chase created.
 * apply_mode: 0 = layer-major sequential grid apply, 1 = wave order (requires the
 * wave preconditions, k even and k <= n-2), 2 = chronological log (no grid),
 * 3 = layer-major with the fused-multiply-add rotation. */
int FN(orc_qr_algorithm)(int64_t n, T *dv, T *ev, T *Q, int64_t ldq, int64_t max_layers, T tol,
                         int64_t maxiter, int32_t *converged, int64_t *iter_out, int64_t *nrot,
                         int apply_mode, int64_t *nblocks, int allow_odd)
{
    if (n % 2 != 0 && !allow_odd) return -1; /* :155 */
    if (max_layers < 1) return -2;
    if (n < 1) return -4; /* :158 allocates an (n-1) x max_layers grid: no such array */
    int64_t start = 1, to = n, iter = 0;
    T *givens = (T *)malloc(sizeof(T) * 2 * (size_t)(n - 1) * (size_t)max_layers);
    int64_t *lt = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    FN(orc_rotlog) log;
    FN(orc_rotlog) *logp = NULL;
    int64_t total = 0, blocks = 0;
    if (apply_mode == 2) {
        log.cap = (n - 1) * max_layers;
        log.idx = (int64_t *)malloc(sizeof(int64_t) * (size_t)log.cap);
        log.c = (T *)malloc(sizeof(T) * (size_t)log.cap);
        log.s = (T *)malloc(sizeof(T) * (size_t)log.cap);
        logp = &log;
    }
    int64_t last_block_ends_at = to;
    int ret = 0;
    int conv = 1;
    for (;;) {
        for (int64_t q = 0; q < (n - 1) * max_layers; q++) {
            givens[2 * q] = 1;
            givens[2 * q + 1] = 0;
        }
        int did = 0;
        for (int64_t q = 0; q < n; q++) lt[q] = 0;
        if (logp) logp->len = 0;
        int early = 0;
        for (;;) {
            if (iter > maxiter) { /* :180-182 */
                conv = 0;
                early = 1;
                break;
            }
            int64_t from = to;
            while (from > start && !FN(is_small)(dv, ev, from - 1, tol)) from -= 1;
            if (to <= start) break;
            if (from == to) {
                EV(from - 1) = 0;
                if (to == last_block_ends_at) last_block_ends_at -= 1;
                to -= 1;
            } else {
                int64_t thickness = 0;
                for (int64_t q = from; q <= to; q++)
                    if (LT(q) > thickness) thickness = LT(q);
                if (thickness == max_layers) {
                    to = from - 1;
                } else {
                    T lam = FN(orc_wilkinson_shift)(DV(to - 1), EV(to - 1), DV(to));
                    did = 1;
                    FN(orc_single_shift)(n, dv, ev, from, to, lam, givens, lt, logp);
                    total += to - from;
                    iter += 1;
                }
            }
            if (to <= start) break;
        }
        if (early) break;
        if (!did) break;
        blocks++;
        if (Q) {
            if (apply_mode == 0) {
                FN(orc_bulk_rmul_seq)(n, n, Q, ldq, givens, max_layers);
            } else if (apply_mode == 3) {
                FN(orc_bulk_rmul_seq_fma)(n, n, Q, ldq, givens, max_layers);
            } else if (apply_mode == 1) {
                if (FN(orc_bulk_wave)(n, n, Q, ldq, givens, max_layers) != 0) {
                    ret = -3;
                    break;
                }
            } else {
                FN(orc_apply_log)(n, Q, ldq, log.idx, log.c, log.s, log.len);
            }
        }
        to = last_block_ends_at;
    }
    if (converged) *converged = conv;
    if (iter_out) *iter_out = iter;
    if (nrot) *nrot = total;
    if (nblocks) *nblocks = blocks;
    free(givens);
    free(lt);
    if (logp) {
        free(log.idx);
        free(log.c);
        free(log.s);
    }
    return ret;
}

/* One chase block of qr_algorithm! without the apply: runs the inner loop of
 * src/bulge_chasing.jl:178-276 from the given state and fills the caller's grid.
 * state[0] = to, state[1] = last_block_ends_at, state[2] = iter, state[3] = did,
 * state[4] = not-converged flag (maxiter hit), state[5] = rotations created. */
void FN(orc_chase_block)(int64_t n, T *dv, T *ev, int64_t max_layers, T tol, int64_t maxiter,
                         T *givens, int64_t *lt, int64_t *state)
{
    int64_t start = 1, to = state[0], last_block_ends_at = state[1], iter = state[2];
    int64_t total = 0;
    for (int64_t q = 0; q < (n - 1) * max_layers; q++) {
        givens[2 * q] = 1;
        givens[2 * q + 1] = 0;
    }
    int did = 0, early = 0;
    for (int64_t q = 0; q < n; q++) lt[q] = 0;
    for (;;) {
        if (iter > maxiter) {
            early = 1;
            break;
        }
        int64_t from = to;
        while (from > start && !FN(is_small)(dv, ev, from - 1, tol)) from -= 1;
        if (to <= start) break;
        if (from == to) {
            EV(from - 1) = 0;
            if (to == last_block_ends_at) last_block_ends_at -= 1;
            to -= 1;
        } else {
            int64_t thickness = 0;
            for (int64_t q = from; q <= to; q++)
                if (LT(q) > thickness) thickness = LT(q);
            if (thickness == max_layers) {
                to = from - 1;
            } else {
                T lam = FN(orc_wilkinson_shift)(DV(to - 1), EV(to - 1), DV(to));
                did = 1;
                FN(orc_single_shift)(n, dv, ev, from, to, lam, givens, lt, NULL);
                total += to - from;
                iter += 1;
            }
        }
        if (to <= start) break;
    }
    state[0] = last_block_ends_at; /* :284, the next block starts here */
    state[1] = last_block_ends_at;
    state[2] = iter;
    state[3] = did;
    state[4] = early;
    state[5] = total;
}

#undef DV
#undef EV
#undef LT
#undef GIV_C
#undef GIV_S
#undef FN
#undef CAT
#undef CAT_
