/*
 * sanitize_check_impl.h -- type-generic body of sanitize_check.c, included once for
 * Float64 and once for Float32 (T, SUF, EPS, BIG, TINY set by the includer).
 *
 * TEST INFRASTRUCTURE.  Each check drives one oracle entry point over small inputs that
 * reach its edge paths (identity rotations, deflation, block limits, odd orders, empty and
 * one-element problems, padded leading dimensions, the iteration limit) and verifies the
 * result against what the mathematics fixes, so that the sanitizer build runs the code on
 * inputs whose outputs are also checked.  Returns the number of failed checks.
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(PFX(name), SUF)
#define LOC(name) CAT(name, SUF)

void FN(orc_givens)(T f, T g, T *cs, T *sn, T *r);
T FN(orc_wilkinson_shift)(T H11, T H12, T H22);
void FN(orc_bulk_rmul_seq)(int64_t m, int64_t n, T *A, int64_t lda, const T *cs, int64_t k);
void FN(orc_bulk_rmul_seq_fma)(int64_t m, int64_t n, T *A, int64_t lda, const T *cs,
                               int64_t k);
int FN(orc_bulk_wave)(int64_t m, int64_t N, T *A, int64_t lda, const T *cs, int64_t k);
int FN(orc_bulk_wave_parallel)(int64_t m, int64_t N, T *A, int64_t lda, const T *cs, int64_t k,
                               int p);
int FN(orc_qr_algorithm)(int64_t n, T *dv, T *ev, T *Q, int64_t ldq, int64_t max_layers, T tol,
                         int64_t maxiter, int32_t *converged, int64_t *iter_out, int64_t *nrot,
                         int apply_mode, int64_t *nblocks, int allow_odd);
void FN(orc_chase_block)(int64_t n, T *dv, T *ev, int64_t max_layers, T tol, int64_t maxiter,
                         T *givens, int64_t *lt, int64_t *state);

#define FAIL(...)                                                                          \
    do {                                                                                   \
        fprintf(stderr, "FAIL %s: ", STR(SUF));                                            \
        fprintf(stderr, __VA_ARGS__);                                                      \
        fputc('\n', stderr);                                                               \
        fails++;                                                                           \
    } while (0)

/* DLARTG's defining identities: c^2 + s^2 = 1 and [c s; -s c] [f; g] = [r; 0]. */
static int LOC(check_givens)(void)
{
    int fails = 0;
    const T cases[][2] = {{0, 0},  {3, 0},     {0, -2},     {1, 1},       {-5, 2},
                          {2, -5}, {BIG, BIG}, {-BIG, TINY}, {TINY, TINY}, {TINY, -BIG},
                          {1, TINY}};
    for (size_t q = 0; q < sizeof(cases) / sizeof(cases[0]); q++) {
        T f = cases[q][0], g = cases[q][1], c, s, r;
        FN(orc_givens)(f, g, &c, &s, &r);
        double sc = fmax(fabs((double)f), fabs((double)g));
        if (!(fabs((double)c * c + (double)s * s - 1) <= 4 * EPS))
            FAIL("givens(%g, %g): c^2 + s^2 - 1 = %g", (double)f, (double)g,
                 (double)c * c + (double)s * s - 1);
        if (!(fabs((double)c * f + (double)s * g - r) <= 4 * EPS * sc))
            FAIL("givens(%g, %g): c f + s g != r", (double)f, (double)g);
        if (!(fabs(-(double)s * f + (double)c * g) <= 4 * EPS * sc))
            FAIL("givens(%g, %g): -s f + c g != 0", (double)f, (double)g);
    }
    return fails;
}

/* The shift is the eigenvalue of the trailing 2x2 nearer to H22. */
static int LOC(check_wilkinson)(void)
{
    int fails = 0;
    const T cases[][3] = {{1, 1, 1}, {2, 0.5, -1}, {-3, TINY, 4}, {BIG, BIG, 0}, {0, 1, 0}};
    for (size_t q = 0; q < sizeof(cases) / sizeof(cases[0]); q++) {
        double a = cases[q][0], b = cases[q][1], d = cases[q][2];
        T lam = FN(orc_wilkinson_shift)(cases[q][0], cases[q][1], cases[q][2]);
        double s = fabs(a) + 2 * fabs(b) + fabs(d);
        double t = (a / s + d / s) / 2, h = hypot((a / s - d / s) / 2, b / s);
        double l1 = (t + h) * s, l2 = (t - h) * s;
        double want = fabs(d - l1) < fabs(d - l2) ? l1 : l2;
        if (!(fabs(lam - want) <= 16 * EPS * s))
            FAIL("wilkinson(%g, %g, %g) = %g, want %g", a, b, d, (double)lam, want);
    }
    return fails;
}

static void LOC(rand_grid)(uint64_t *st, T *cs, int64_t rows, int64_t k)
{
    for (int64_t q = 0; q < rows * k; q++) {
        double th = 3.14159265358979 * urand(st);
        if (urand(st) > 0.6) th = 0; /* identity rotations: the skip paths */
        cs[2 * q] = (T)cos(th);
        cs[2 * q + 1] = (T)sin(th);
    }
}

/* The three grid appliers against one another on an m x N block with a padded leading
 * dimension; the rows beyond m (the padding) must stay untouched. */
static int LOC(check_bulk)(int64_t m, int64_t N, int64_t k, int p)
{
    int fails = 0;
    uint64_t st = 0x9E3779B97F4A7C15ull ^ (uint64_t)(m * 131 + N * 7 + k);
    const int64_t lda = m + 3, rows = N - 1;
    T *cs = malloc(sizeof(T) * 2 * (size_t)(rows > 0 ? rows : 1) * (size_t)(k > 0 ? k : 1));
    T *A0 = malloc(sizeof(T) * (size_t)(lda * N));
    T *A1 = malloc(sizeof(T) * (size_t)(lda * N));
    T *A2 = malloc(sizeof(T) * (size_t)(lda * N));
    T *A3 = malloc(sizeof(T) * (size_t)(lda * N));
    LOC(rand_grid)(&st, cs, rows, k);
    for (int64_t q = 0; q < lda * N; q++) A0[q] = (T)urand(&st);
    memcpy(A1, A0, sizeof(T) * (size_t)(lda * N));
    memcpy(A2, A0, sizeof(T) * (size_t)(lda * N));
    memcpy(A3, A0, sizeof(T) * (size_t)(lda * N));
    FN(orc_bulk_rmul_seq)(m, N, A0, lda, cs, k);
    FN(orc_bulk_rmul_seq_fma)(m, N, A1, lda, cs, k);
    int wave_ok = (rows % 2 == 1) && (k % 2 == 0) && (k <= N - 2);
    int r2 = FN(orc_bulk_wave)(m, N, A2, lda, cs, k);
    int r3 = FN(orc_bulk_wave_parallel)(m, N, A3, lda, cs, k, p);
    if (r2 != (wave_ok ? 0 : -1) || r3 != (wave_ok ? 0 : -1))
        FAIL("bulk m=%lld N=%lld k=%lld: wave returned %d / %d", (long long)m, (long long)N,
             (long long)k, r2, r3);
    const double tol = 8 * EPS * (double)(k + 2);
    for (int64_t j = 0; j < N; j++)
        for (int64_t i = 0; i < lda; i++) {
            int64_t q = i + j * lda;
            if (i >= m) {
                if (A1[q] != A0[q] || A2[q] != A0[q] || A3[q] != A0[q]) {
                    FAIL("bulk m=%lld N=%lld k=%lld: padding row %lld changed", (long long)m,
                         (long long)N, (long long)k, (long long)i);
                    i = lda, j = N;
                }
                continue;
            }
            if (fabs((double)A1[q] - A0[q]) > tol ||
                (wave_ok && (fabs((double)A2[q] - A0[q]) > tol || A3[q] != A2[q]))) {
                FAIL("bulk m=%lld N=%lld k=%lld: entry (%lld, %lld) differs", (long long)m,
                     (long long)N, (long long)k, (long long)i, (long long)j);
                i = lda, j = N;
            }
        }
    free(cs);
    free(A0);
    free(A1);
    free(A2);
    free(A3);
    return fails;
}

/* Residual max_j ||T q_j - lam_j q_j||_inf and orthogonality max |Q^T Q - I|, both in
 * units of n * eps * max(1, ||T||_inf). */
static void LOC(eig_errors)(int64_t n, const T *d0, const T *e0, const T *lam, const T *Q,
                            int64_t ldq, double *res, double *orth)
{
    double tn = 1;
    for (int64_t i = 0; i < n; i++) {
        double row = fabs((double)d0[i]);
        if (i > 0) row += fabs((double)e0[i - 1]);
        if (i + 1 < n) row += fabs((double)e0[i]);
        tn = fmax(tn, row);
    }
    double r = 0, o = 0;
    for (int64_t j = 0; j < n; j++) {
        const T *q = Q + j * ldq;
        for (int64_t i = 0; i < n; i++) {
            double v = ((double)d0[i] - lam[j]) * q[i];
            if (i > 0) v += (double)e0[i - 1] * q[i - 1];
            if (i + 1 < n) v += (double)e0[i] * q[i + 1];
            r = fmax(r, fabs(v));
        }
        for (int64_t l = 0; l <= j; l++) {
            double dot = 0;
            for (int64_t i = 0; i < n; i++) dot += (double)q[i] * Q[i + l * ldq];
            o = fmax(o, fabs(dot - (l == j)));
        }
    }
    *res = r / (n * EPS * tn);
    *orth = o / (n * EPS);
}

static void LOC(make_tridiag)(uint64_t *st, int spec, int64_t n, T *d, T *e)
{
    for (int64_t i = 0; i < n; i++) {
        switch (spec) {
        case SPEC_RANDOM: d[i] = (T)urand(st); break;
        case SPEC_SPLIT: d[i] = (T)urand(st); break;
        case SPEC_GRADED: d[i] = (T)ldexp(1.0, -(int)(i % 40)); break;
        case SPEC_WILKINSON: d[i] = (T)fabs((double)(n - 1) / 2 - (double)i); break;
        default: d[i] = 0;
        }
        if (i + 1 < n) {
            switch (spec) {
            case SPEC_RANDOM: e[i] = (T)urand(st); break;
            case SPEC_SPLIT: e[i] = (i % 5 == 4) ? 0 : (T)urand(st); break;
            case SPEC_GRADED: e[i] = (T)ldexp(1.0, -(int)(i % 40) - 1); break;
            case SPEC_WILKINSON: e[i] = 1; break;
            default: e[i] = 0;
            }
        }
    }
}

/* qr_algorithm! in every apply mode, then the same solve block by block through
 * orc_chase_block + the layer-major apply, which must reproduce it bitwise. */
static int LOC(check_qr)(int64_t n, int64_t k, int spec, int allow_odd)
{
    int fails = 0;
    uint64_t st = 0xD1B54A32D192ED03ull ^ (uint64_t)(n * 977 + k * 13 + spec);
    const int64_t ldq = n + 2, nq = n > 0 ? n : 1;
    T *d0 = malloc(sizeof(T) * (size_t)nq), *e0 = malloc(sizeof(T) * (size_t)nq);
    T *d = malloc(sizeof(T) * (size_t)nq), *e = malloc(sizeof(T) * (size_t)nq);
    T *Q = malloc(sizeof(T) * (size_t)(ldq * nq)), *Qm0 = malloc(sizeof(T) * (size_t)(ldq * nq));
    T *dm0 = malloc(sizeof(T) * (size_t)nq);
    LOC(make_tridiag)(&st, spec, n, d0, e0);
    const int64_t maxiter = 100 * (n > 0 ? n : 1);
    const int odd = n % 2 != 0;
    int64_t iter0 = -1;
    for (int mode = 0; mode < 4; mode++) {
        memcpy(d, d0, sizeof(T) * (size_t)nq);
        memcpy(e, e0, sizeof(T) * (size_t)nq);
        for (int64_t q = 0; q < ldq * nq; q++) Q[q] = (q % ldq == q / ldq) ? 1 : (q % ldq < n ? 0 : 7);
        int32_t conv = -1;
        int64_t it = -1, nrot = -1, nblk = -1;
        int rc = FN(orc_qr_algorithm)(n, d, e, Q, ldq, k, (T)EPS, maxiter, &conv, &it, &nrot, mode,
                                      &nblk, allow_odd);
        if (n < 1) {
            if (rc != (odd && !allow_odd ? -1 : -4)) FAIL("qr n=%lld: rc %d", (long long)n, rc);
            continue;
        }
        if (odd && !allow_odd) {
            if (rc != -1) FAIL("qr odd n=%lld without allow_odd: rc %d", (long long)n, rc);
            continue;
        }
        if (mode == 1 && (odd || k % 2 != 0 || k > n - 2)) {
            /* outside the wave preconditions: refused once a block has rotations */
            if (rc != -3 && !(rc == 0 && nblk == 0))
                FAIL("qr n=%lld k=%lld mode 1: rc %d", (long long)n, (long long)k, rc);
            continue;
        }
        if (rc != 0 || conv != 1) {
            FAIL("qr n=%lld k=%lld spec %d mode %d: rc %d converged %d", (long long)n,
                 (long long)k, spec, mode, rc, conv);
            continue;
        }
        if (mode == 0) {
            iter0 = it;
            memcpy(Qm0, Q, sizeof(T) * (size_t)(ldq * nq));
            memcpy(dm0, d, sizeof(T) * (size_t)nq);
        } else if (it != iter0 || memcmp(d, dm0, sizeof(T) * (size_t)n) != 0) {
            FAIL("qr n=%lld mode %d: the chase differs from mode 0", (long long)n, mode);
        }
        for (int64_t j = 0; j < n; j++)
            for (int64_t i = n; i < ldq; i++)
                if (Q[i + j * ldq] != 7) {
                    FAIL("qr n=%lld mode %d: padding row %lld written", (long long)n, mode,
                         (long long)i);
                    i = ldq, j = n;
                }
        double res, orth;
        LOC(eig_errors)(n, d0, e0, d, Q, ldq, &res, &orth);
        if (!(res <= 8 && orth <= 8))
            FAIL("qr n=%lld k=%lld spec %d mode %d: residual %.2f orthogonality %.2f (n eps)",
                 (long long)n, (long long)k, spec, mode, res, orth);
    }
    if (n >= 1 && !(odd && !allow_odd) && iter0 >= 0) {
        /* block by block */
        memcpy(d, d0, sizeof(T) * (size_t)nq);
        memcpy(e, e0, sizeof(T) * (size_t)nq);
        for (int64_t q = 0; q < ldq * nq; q++) Q[q] = (q % ldq == q / ldq) ? 1 : (q % ldq < n ? 0 : 7);
        const int64_t rows = n - 1;
        T *grid = malloc(sizeof(T) * 2 * (size_t)(rows > 0 ? rows : 1) * (size_t)k);
        int64_t *lt = malloc(sizeof(int64_t) * (size_t)n);
        int64_t state[6] = {n, n, 0, 0, 0, 0};
        for (int64_t b = 0; b <= maxiter + 1; b++) {
            FN(orc_chase_block)(n, d, e, k, (T)EPS, maxiter, grid, lt, state);
            if (state[4] || !state[3]) break;
            FN(orc_bulk_rmul_seq)(n, n, Q, ldq, grid, k);
        }
        if (state[4] || state[2] != iter0 || memcmp(d, dm0, sizeof(T) * (size_t)n) != 0)
            FAIL("chase_block n=%lld: iterations %lld vs %lld", (long long)n,
                 (long long)state[2], (long long)iter0);
        for (int64_t j = 0; j < n; j++)
            if (memcmp(Q + j * ldq, Qm0 + j * ldq, sizeof(T) * (size_t)n) != 0) {
                FAIL("chase_block n=%lld: Q column %lld differs from qr_algorithm", (long long)n,
                     (long long)j);
                break;
            }
        free(grid);
        free(lt);
    }
    /* the iteration limit: not converged, nothing out of bounds */
    if (n >= 4 && !odd) {
        memcpy(d, d0, sizeof(T) * (size_t)nq);
        memcpy(e, e0, sizeof(T) * (size_t)nq);
        int32_t conv = -1;
        int64_t it = -1, nrot = -1, nblk = -1;
        int rc = FN(orc_qr_algorithm)(n, d, e, NULL, ldq, k, (T)EPS, 1, &conv, &it, &nrot, 0,
                                      &nblk, 0);
        if (rc != 0 || (spec != SPEC_ZERO && conv != 0))
            FAIL("qr n=%lld maxiter 1: rc %d converged %d", (long long)n, rc, conv);
    }
    free(d0);
    free(e0);
    free(d);
    free(e);
    free(Q);
    free(Qm0);
    free(dm0);
    return fails;
}

static int LOC(run_all)(void)
{
    int fails = LOC(check_givens)() + LOC(check_wilkinson)();
    const int64_t bulk[][3] = {{1, 2, 0},   {5, 3, 1},   {37, 24, 6},  {64, 40, 2},
                               {70, 33, 8}, {33, 12, 10}, {16, 9, 4}, {0, 10, 2}};
    for (size_t q = 0; q < sizeof(bulk) / sizeof(bulk[0]); q++)
        fails += LOC(check_bulk)(bulk[q][0], bulk[q][1], bulk[q][2], 3);
    fails += LOC(check_bulk)(200, 30, 6, 4);
    const int64_t ns[] = {0, 1, 2, 3, 4, 7, 16, 33, 64, 130};
    const int64_t ks[] = {1, 2, 8};
    for (size_t a = 0; a < sizeof(ns) / sizeof(ns[0]); a++)
        for (size_t b = 0; b < sizeof(ks) / sizeof(ks[0]); b++)
            for (int spec = SPEC_RANDOM; spec <= SPEC_ZERO; spec++)
                for (int odd_ok = 0; odd_ok < 2; odd_ok++)
                    fails += LOC(check_qr)(ns[a], ks[b], spec, odd_ok);
    return fails;
}

#undef FAIL
#undef LOC
#undef FN
#undef CAT
#undef CAT_
