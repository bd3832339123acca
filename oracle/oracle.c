/*
 * oracle.c — CPU FP64 oracle for the multifrontal numeric phase.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP engine is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg load it.  The product path never calls it.
 *
 * It is a plain-C restatement of the reference's serial front loop
 * (paths relative to the CliqueTrees.jl tree, src/Multifrontal.jl/src/):
 *   or_chol_impl   chol_impl!        cholesky.jl:361-392
 *                  chol_loop_snd!    cholesky.jl:434-556
 *                  chol_loop_nod!    cholesky.jl:560-692
 *                  chol_factor!      cholesky.jl:694-737
 *                  chol_send!        cholesky.jl:950-983 (addscattertri!, utils.jl:555-579)
 *   potrf_rec      potrf! recursive  blas/potrf.jl:135-178, potrf2! blas/potrf.jl:3-41
 *   regularize     no_regularization.jl:27-42, dynamic_regularization.jl:75-90,
 *                  GMW81 regularization/gmw81.jl:109-155 (initialize :50-101)
 *   chol_factor_gmw chol_factor! GMW81 specialization cholesky.jl:739-875
 *   trsm_rlt       trsx2_fwd!(:R)    blas/trsx.jl:11-31 (B <- B A^-T, A lower)
 *   or_div_fwd     div_impl! forward divide.jl:417-450, div_fwd_loop_snd!/nod! 493-821,
 *                  div_fwd_update!   divide.jl:1206-1248
 *   or_div_bwd     div_impl! backward, div_bwd_loop_snd!/nod! divide.jl:864-1204,
 *                  div_bwd_update!   divide.jl:1250-1293
 *   or_selinv      selinv_impl!      selinv.jl:72-424 (selected inversion of a
 *                  Cholesky factor, root first)
 *
 * The dense kernels are the reference's own pure-Julia fallbacks restated in
 * C.  For Float64 the reference calls OpenBLAS instead (potrf.jl:87-103,
 * trsx.jl:156-177, syrk.jl:137-159, gemx.jl:3-41, syr.jl:3-11); or_set_blas() installs
 * host BLAS/LAPACK entry points (Fortran ABI, 32-bit integers) so the same
 * loop can run at BLAS speed for the CPU baseline.  All index arrays are the
 * reference's 1-based Int64 arrays.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t i64;

/* ---- optional host BLAS (Fortran ABI) ---------------------------------- */
typedef void (*f_dpotrf)(const char*, const int*, double*, const int*, int*);
typedef void (*f_dtrsm)(const char*, const char*, const char*, const char*, const int*,
                        const int*, const double*, const double*, const int*, double*,
                        const int*);
typedef void (*f_dsyrk)(const char*, const char*, const int*, const int*, const double*,
                        const double*, const int*, const double*, double*, const int*);
typedef void (*f_dgemm)(const char*, const char*, const int*, const int*, const int*,
                        const double*, const double*, const int*, const double*, const int*,
                        const double*, double*, const int*);

static f_dpotrf B_potrf = 0;
static f_dtrsm B_trsm = 0;
static f_dsyrk B_syrk = 0;
typedef void (*f_dsyr)(const char*, const int*, const double*, const double*, const int*,
                       double*, const int*);
static f_dgemm B_gemm = 0;
static f_dsyr B_syr = 0;

/* syr (may be NULL) serves the rank-1 update of nn = 1 fronts, as syr! calls
 * BLAS.syr! for Float64 (blas/syr.jl:3-11). */
void or_set_blas_syr(void* syr) { B_syr = (f_dsyr)syr; }

void or_set_blas(void* potrf, void* trsm, void* syrk, void* gemm) {
    B_potrf = (f_dpotrf)potrf;
    B_trsm = (f_dtrsm)trsm;
    B_syrk = (f_dsyrk)syrk;
    B_gemm = (f_dgemm)gemm;
}

/* ---- regularization ------------------------------------------------------ */
enum { REG_NONE = 0, REG_DYNAMIC = 1, REG_GMW81 = 2 };
typedef struct {
    int kind;
    double delta, eps; /* GMW81: delta = minimum pivot, eps = beta */
} reg_t;

/* Julia's max for floats: NaN if either argument is NaN */
static double jmax(double a, double b) {
    if (a != a || b != b) return a + b;
    return a > b ? a : b;
}

/* GMW81 regularize (gmw81.jl:109-155): s * max(|Djj|, Aimax^2 / beta, delta),
 * Aimax the largest magnitude below the pivot in the pivot's front column */
static double regularize_gmw(const reg_t* R, double s, double Djj, double Aimax) {
    return s * jmax(jmax(fabs(Djj), Aimax * Aimax / R->eps), R->delta);
}

/* gmw81_beta (gmw81.jl:74-101) over the assembled factor storage:
 * max(gamma, xi / sqrt(n^2 - 1), eps) with gamma the largest diagonal and xi
 * the largest off-diagonal magnitude */
double or_gmw81_beta(i64 nfr, const i64* resptr, const i64* sepptr, const i64* Dptr,
                     const i64* Lptr, const double* Dval, const double* Lval) {
    double gamma = 0.0, xi = 0.0;
    i64 n = 0;
    for (i64 j = 0; j < nfr; ++j) {
        i64 nn = resptr[j + 1] - resptr[j], na = sepptr[j + 1] - sepptr[j];
        const double* D = Dval + Dptr[j] - 1;
        const double* L = Lval + Lptr[j] - 1;
        n += nn;
        for (i64 c = 0; c < nn; ++c) {
            gamma = jmax(gamma, fabs(D[c * nn + c]));
            for (i64 r = c + 1; r < nn; ++r) xi = jmax(xi, fabs(D[c * nn + r]));
            for (i64 r = 0; r < na; ++r) xi = jmax(xi, fabs(L[c * na + r]));
        }
    }
    return jmax(jmax(gamma, xi / sqrt((double)n * (double)n - 1.0)), 2.220446049250313e-16);
}

/* regularize(R, S, Djj, j): 0 signals a failed pivot. */
static double regularize(const reg_t* R, double s, double Djj) {
    if (R->kind == REG_DYNAMIC) {
        if (s * Djj < R->eps) return R->delta * s;
        return Djj;
    }
    if (s != 0.0 && !(s * Djj > 0.0)) return 0.0;
    return Djj;
}

/* ---- dense kernels (column-major, leading dimension ld) ----------------- */
#define A_(M, ld, i, j) (M)[(i64)(j) * (ld) + (i)]

/* potrf2!(:L) blas/potrf.jl:3-41: unblocked right-looking; diag=0 Cholesky, 1 LDLt */
static i64 potrf2(double* A, i64 lda, i64 n, double* d, const double* S, const reg_t* R,
                  int diag) {
    for (i64 j = 0; j < n; ++j) {
        double Ajj = regularize(R, S[j], A_(A, lda, j, j));
        if (Ajj == 0.0) return j + 1;
        if (diag == 0) {
            Ajj = sqrt(Ajj);
            A_(A, lda, j, j) = Ajj;
        } else {
            d[j] = Ajj;
        }
        double iD = 1.0 / Ajj;
        for (i64 i = j + 1; i < n; ++i) A_(A, lda, i, j) *= iD;
        for (i64 k = j + 1; k < n; ++k) {
            double Akj = A_(A, lda, k, j);
            double f = diag == 0 ? Akj : Ajj * Akj;
            A_(A, lda, k, k) -= f * Akj;
            for (i64 i = k + 1; i < n; ++i) A_(A, lda, i, k) -= A_(A, lda, i, j) * f;
        }
    }
    return 0;
}

/* B (m x n) <- B A^-T with A n x n lower (unit if diag==1): trsx2_fwd!(:R) */
static void trsm_rlt(const double* A, i64 lda, double* B, i64 ldb, i64 m, i64 n, int unit) {
    if (B_trsm && m > 0 && n > 0) {
        int mi = (int)m, ni = (int)n, la = (int)lda, lb = (int)ldb;
        double one = 1.0;
        B_trsm("R", "L", "T", unit ? "U" : "N", &mi, &ni, &one, A, &la, B, &lb);
        return;
    }
    for (i64 j = 0; j < n; ++j) {
        for (i64 k = 0; k < j; ++k) {
            double Ajk = A_(A, lda, j, k);
            for (i64 i = 0; i < m; ++i) A_(B, ldb, i, j) -= A_(B, ldb, i, k) * Ajk;
        }
        if (!unit) {
            double iA = 1.0 / A_(A, lda, j, j);
            for (i64 i = 0; i < m; ++i) A_(B, ldb, i, j) *= iA;
        }
    }
}

/* C (n x n, lower) -= A diag(d) A^T, A n x k (d == NULL: identity) */
static void syrk_ln(const double* A, i64 lda, const double* d, double* C, i64 ldc, i64 n,
                    i64 k) {
    if (!d && B_syrk && n > 0 && k > 0) {
        int ni = (int)n, ki = (int)k, la = (int)lda, lc = (int)ldc;
        double m1 = -1.0, one = 1.0;
        B_syrk("L", "N", &ni, &ki, &m1, A, &la, &one, C, &lc);
        return;
    }
    if (d && B_gemm && n > 0 && k > 0) {
        /* W = A diag(d); C_lower -= W A^T by column blocks of C: the block
         * below each diagonal block by GEMM, the diagonal block's lower
         * triangle by loops (the strict upper triangle stays untouched, as
         * syr2k!(:L) leaves it, blas/syr2k.jl:4-12) */
        double* W = (double*)malloc(sizeof(double) * (size_t)(n * k));
        for (i64 c = 0; c < k; ++c)
            for (i64 i = 0; i < n; ++i) W[c * n + i] = A_(A, lda, i, c) * d[c];
        const i64 nb = 128;
        for (i64 j0 = 0; j0 < n; j0 += nb) {
            i64 jb = n - j0 < nb ? n - j0 : nb;
            for (i64 c = 0; c < k; ++c)
                for (i64 j = j0; j < j0 + jb; ++j) {
                    double f = A_(A, lda, j, c);
                    if (f == 0.0) continue;
                    for (i64 i = j; i < j0 + jb; ++i) A_(C, ldc, i, j) -= W[c * n + i] * f;
                }
            if (n - j0 - jb <= 0) continue;
            int mi = (int)(n - j0 - jb), ni = (int)jb, ki = (int)k, lw = (int)n, la = (int)lda,
                lc = (int)ldc;
            double m1 = -1.0, one = 1.0;
            B_gemm("N", "T", &mi, &ni, &ki, &m1, W + j0 + jb, &lw, A + j0, &la, &one,
                   C + j0 * ldc + j0 + jb, &lc);
        }
        free(W);
        return;
    }
    for (i64 c = 0; c < k; ++c) {
        for (i64 j = 0; j < n; ++j) {
            double f = A_(A, lda, j, c) * (d ? d[c] : 1.0);
            if (f == 0.0) continue;
            for (i64 i = j; i < n; ++i) A_(C, ldc, i, j) -= A_(A, lda, i, c) * f;
        }
    }
}

/* potrf! recursive, blas/potrf.jl:135-178 (LAPACK for Cholesky without regularization) */
static i64 potrf_rec(double* A, i64 lda, i64 n, double* d, const double* S, const reg_t* R,
                     int diag) {
    if (diag == 0 && R->kind == REG_NONE && B_potrf) {
        /* potrf.jl:101-103: LAPACK.potrf! ignores signs (they are Ones here) */
        int ni = (int)n, la = (int)lda, info = 0;
        if (n == 0) return 0;
        B_potrf("L", &ni, A, &la, &info);
        return info;
    }
    if (n <= 64) return potrf2(A, lda, n, d, S, R, diag);
    i64 n1 = 1;
    while (n1 * 2 <= n) n1 *= 2;
    n1 /= 2;
    i64 info = potrf_rec(A, lda, n1, d, S, R, diag);
    if (info) return info;
    double* B = A + n1; /* rows n1.., cols 0..n1-1 */
    trsm_rlt(A, lda, B, lda, n - n1, n1, diag);
    if (diag) {
        for (i64 c = 0; c < n1; ++c) {
            double idc = 1.0 / d[c];
            for (i64 i = 0; i < n - n1; ++i) A_(B, lda, i, c) *= idc;
        }
    }
    syrk_ln(B, lda, diag ? d : NULL, A + n1 * lda + n1, lda, n - n1, n1);
    info = potrf_rec(A + n1 * lda + n1, lda, n - n1, d + n1, S + n1, R, diag);
    if (info) return n1 + info;
    return 0;
}

/* chol_factor! for GMW81 (cholesky.jl:739-875): the stacked panel [D; L]
 * (D nn x nn lower, L na x nn) factored column by column, each pivot
 * regularized with the largest magnitude below it in the whole front column
 * as it stands after the previous pivots.  The reference groups the columns
 * in blocks of THRESHOLD = 64 and applies the updates of a block to the
 * trailing columns with syrk!/gemm!; the products are the same. */
static void chol_factor_gmw(double* D, i64 nn, double* L, i64 na, double* d, const double* S,
                            const reg_t* R, int diag) {
    for (i64 j = 0; j < nn; ++j) {
        double amax = 0.0;
        for (i64 i = j + 1; i < nn; ++i) amax = jmax(amax, fabs(A_(D, nn, i, j)));
        for (i64 i = 0; i < na; ++i) amax = jmax(amax, fabs(A_(L, na, i, j)));
        double Djj = regularize_gmw(R, S[j], A_(D, nn, j, j), amax);
        if (diag == 0) {
            Djj = sqrt(Djj);
            A_(D, nn, j, j) = Djj;
        } else {
            d[j] = Djj;
        }
        double iD = 1.0 / Djj;
        for (i64 i = j + 1; i < nn; ++i) A_(D, nn, i, j) *= iD;
        for (i64 i = 0; i < na; ++i) A_(L, na, i, j) *= iD;
        for (i64 k = j + 1; k < nn; ++k) {
            double Dkj = A_(D, nn, k, j);
            double f = diag == 0 ? Dkj : Djj * Dkj;
            A_(D, nn, k, k) -= f * Dkj;
            for (i64 i = k + 1; i < nn; ++i) A_(D, nn, i, k) -= A_(D, nn, i, j) * f;
            for (i64 i = 0; i < na; ++i) A_(L, na, i, k) -= A_(L, na, i, j) * f;
        }
    }
}

/* ---- symbolic view ------------------------------------------------------- */
typedef struct {
    i64 n, nfr;
    const i64 *resptr, *sepptr, *reltgt, *chdptr, *chdtgt, *Dptr, *Lptr;
} sym_t;

/* addscattertri!(F, M, inj, :L): F[inj[i], inj[j]] += M[i, j], j <= i */
static void send_tri(double* F, i64 ldf, const double* M, i64 na, const i64* inj) {
    for (i64 j = 0; j < na; ++j) {
        i64 fj = inj[j] - 1;
        for (i64 i = j; i < na; ++i) A_(F, ldf, inj[i] - 1, fj) += A_(M, na, i, j);
    }
}

/*
 * chol_impl! (cholesky.jl:361-392).  diag: 0 = Cholesky (:N), 1 = LDLt (:U).
 * Mval: workspace of nMval doubles; Mptr: nMptr+1 entries; Fval: nFval^2.
 * Returns 0, or the first failing column in permuted order (1-based).
 */
i64 or_chol_range(i64 n, i64 nfr, const i64* resptr, const i64* sepptr, const i64* reltgt,
                  const i64* chdptr, const i64* chdtgt, const i64* Dptr, const i64* Lptr,
                  double* Dval, double* Lval, double* d, const double* signs, int diag,
                  int reg_kind, double delta, double eps, double* Mval, i64* Mptr,
                  double* Fval, i64 j0, i64 j1);

i64 or_chol_impl(i64 n, i64 nfr, const i64* resptr, const i64* sepptr, const i64* reltgt,
                 const i64* chdptr, const i64* chdtgt, const i64* Dptr, const i64* Lptr,
                 double* Dval, double* Lval, double* d, const double* signs, int diag,
                 int reg_kind, double delta, double eps, double* Mval, i64* Mptr,
                 double* Fval) {
    return or_chol_range(n, nfr, resptr, sepptr, reltgt, chdptr, chdtgt, Dptr, Lptr, Dval, Lval,
                         d, signs, diag, reg_kind, delta, eps, Mval, Mptr, Fval, 0, nfr);
}

/*
 * The front loop over fronts [j0, j1) only.  When [j0, j1) is a whole
 * subtree of the postorder (a front and all its descendants) the loop is
 * self-contained; the update matrix of its root is left on the stack.
 */
i64 or_chol_range(i64 n, i64 nfr, const i64* resptr, const i64* sepptr, const i64* reltgt,
                  const i64* chdptr, const i64* chdtgt, const i64* Dptr, const i64* Lptr,
                  double* Dval, double* Lval, double* d, const double* signs, int diag,
                  int reg_kind, double delta, double eps, double* Mval, i64* Mptr,
                  double* Fval, i64 j0, i64 j1) {
    reg_t R = {reg_kind, delta, eps};
    (void)n;
    /* initialize(L, signs, GMW81()) (gmw81.jl:50-63): a negative beta takes
     * gmw81_beta of the assembled matrix */
    if (reg_kind == REG_GMW81 && eps < 0.0)
        R.eps = or_gmw81_beta(nfr, resptr, sepptr, Dptr, Lptr, Dval, Lval);
    i64 ns = 0;
    Mptr[1] = 0; /* 1-based stack index, 0-based offsets */
    for (i64 j = j0; j < j1; ++j) {
        i64 r0 = resptr[j] - 1;
        i64 nn = resptr[j + 1] - resptr[j];
        i64 na = sepptr[j + 1] - sepptr[j];
        i64 nj = nn + na;
        double* F = Fval;
        for (i64 c = 0; c < nj; ++c)
            for (i64 r = c; r < nj; ++r) A_(F, nj, r, c) = 0.0;
        for (i64 q = chdptr[j + 1] - 1; q >= chdptr[j]; --q) {
            i64 i = chdtgt[q - 1] - 1;
            i64 ma = sepptr[i + 1] - sepptr[i];
            send_tri(F, nj, Mval + Mptr[ns], ma, reltgt + sepptr[i] - 1);
            ns -= 1;
        }
        double* D11 = Dval + Dptr[j] - 1;
        double* L21 = Lval + Lptr[j] - 1;
        if (nn == 1) {
            /* chol_loop_nod! */
            double d11 = D11[0] + F[0];
            for (i64 r = 0; r < na; ++r) L21[r] += F[1 + r];
            if (R.kind == REG_GMW81) {
                double amax = 0.0;
                for (i64 r = 0; r < na; ++r) amax = jmax(amax, fabs(L21[r]));
                d11 = regularize_gmw(&R, signs[r0], d11, amax);
            } else {
                d11 = regularize(&R, signs[r0], d11);
            }
            if (d11 == 0.0) return 1 + r0;
            if (diag == 0) {
                d11 = sqrt(d11);
                D11[0] = d11;
            } else {
                d[r0] = d11;
            }
            if (na > 0) {
                ns += 1;
                i64 strt = Mptr[ns];
                Mptr[ns + 1] = strt + na * na;
                double* M = Mval + strt;
                for (i64 c = 0; c < na; ++c)
                    for (i64 r = c; r < na; ++r) A_(M, na, r, c) = A_(F, nj, 1 + r, 1 + c);
                for (i64 r = 0; r < na; ++r) L21[r] /= d11;
                double alpha = diag ? -d11 : -1.0;
                if (B_syr) {
                    int ni = (int)na, one = 1;
                    B_syr("L", &ni, &alpha, L21, &one, M, &ni);
                } else {
                    for (i64 c = 0; c < na; ++c) {
                        double f = alpha * L21[c];
                        for (i64 r = c; r < na; ++r) A_(M, na, r, c) += L21[r] * f;
                    }
                }
            }
            continue;
        }
        /* chol_loop_snd!: L <- L + F */
        for (i64 c = 0; c < nn; ++c) {
            for (i64 r = c; r < nn; ++r) A_(D11, nn, r, c) += A_(F, nj, r, c);
            for (i64 r = 0; r < na; ++r) A_(L21, na, r, c) += A_(F, nj, nn + r, c);
        }
        /* chol_factor! */
        double* dj = diag ? d + r0 : NULL;
        double dummy_d[1];
        if (R.kind == REG_GMW81) {
            chol_factor_gmw(D11, nn, L21, na, dj, signs + r0, &R, diag);
        } else {
            i64 info = potrf_rec(D11, nn, nn, dj ? dj : dummy_d, signs + r0, &R, diag);
            if (info) return info + r0;
        }
        if (na > 0) {
            if (R.kind != REG_GMW81) trsm_rlt(D11, nn, L21, na, na, nn, diag);
            if (diag && R.kind != REG_GMW81) {
                for (i64 c = 0; c < nn; ++c) {
                    double idc = 1.0 / dj[c];
                    for (i64 r = 0; r < na; ++r) A_(L21, na, r, c) *= idc;
                }
            }
            ns += 1;
            i64 strt = Mptr[ns];
            Mptr[ns + 1] = strt + na * na;
            double* M = Mval + strt;
            for (i64 c = 0; c < na; ++c)
                for (i64 r = c; r < na; ++r) A_(M, na, r, c) = A_(F, nj, nn + r, nn + c);
            syrk_ln(L21, na, dj, M, na, na, nn);
        }
    }
    return 0;
}

/* C (m x nrhs) -= A (m x k) X (k x nrhs) ; trans: A^T (A is k x m) */
static void gemm_sub(const double* A, i64 lda, int trans, const double* X, i64 ldx, double* C,
                     i64 ldc, i64 m, i64 k, i64 nrhs) {
    if (m == 0 || k == 0 || nrhs == 0) return;
    if (B_gemm) {
        int mi = (int)m, ni = (int)nrhs, ki = (int)k, la = (int)lda, lx = (int)ldx,
            lc = (int)ldc;
        double m1 = -1.0, one = 1.0;
        B_gemm(trans ? "T" : "N", "N", &mi, &ni, &ki, &m1, A, &la, X, &lx, &one, C, &lc);
        return;
    }
    for (i64 c = 0; c < nrhs; ++c) {
        if (!trans) {
            for (i64 kk = 0; kk < k; ++kk) {
                double x = A_(X, ldx, kk, c);
                for (i64 r = 0; r < m; ++r) A_(C, ldc, r, c) -= A_(A, lda, r, kk) * x;
            }
        } else {
            for (i64 r = 0; r < m; ++r) {
                double s = 0.0;
                for (i64 kk = 0; kk < k; ++kk) s += A_(A, lda, kk, r) * A_(X, ldx, kk, c);
                A_(C, ldc, r, c) -= s;
            }
        }
    }
}

/* X (n x nrhs) <- A^-1 X (fwd, A lower) or A^-T X (bwd) */
static void trsm_left(const double* A, i64 lda, double* X, i64 ldx, i64 n, i64 nrhs, int bwd,
                      int unit) {
    if (B_trsm && n > 0 && nrhs > 0) {
        int ni = (int)n, mi = (int)nrhs, la = (int)lda, lx = (int)ldx;
        double one = 1.0;
        B_trsm("L", "L", bwd ? "T" : "N", unit ? "U" : "N", &ni, &mi, &one, A, &la, X, &lx);
        return;
    }
    for (i64 c = 0; c < nrhs; ++c) {
        double* x = X + c * ldx;
        if (!bwd) {
            for (i64 j = 0; j < n; ++j) {
                for (i64 k = 0; k < j; ++k) x[j] -= A_(A, lda, j, k) * x[k];
                if (!unit) x[j] *= 1.0 / A_(A, lda, j, j);
            }
        } else {
            for (i64 j = n - 1; j >= 0; --j) {
                if (!unit) x[j] *= 1.0 / A_(A, lda, j, j);
                for (i64 k = 0; k < j; ++k) x[k] -= A_(A, lda, j, k) * x[j];
            }
        }
    }
}

/*
 * Forward solve C <- L^-1 C (div_impl!, forward branch), C is n x nrhs
 * (ld ldc) in permuted order.  Mval: nNval*nrhs, Mptr: nMptr+1, Fval: nFval*nrhs.
 */
void or_div_fwd(i64 nfr, const i64* resptr, const i64* sepptr, const i64* reltgt,
                const i64* chdptr, const i64* chdtgt, const i64* Dptr, const i64* Lptr,
                const double* Dval, const double* Lval, int diag, double* C, i64 ldc,
                i64 nrhs, double* Mval, i64* Mptr, double* Fval) {
    i64 ns = 0;
    Mptr[1] = 0;
    for (i64 j = 0; j < nfr; ++j) {
        i64 r0 = resptr[j] - 1;
        i64 nn = resptr[j + 1] - resptr[j];
        i64 na = sepptr[j + 1] - sepptr[j];
        i64 nj = nn + na;
        double* F = Fval;
        memset(F, 0, sizeof(double) * (size_t)(nj * nrhs));
        for (i64 q = chdptr[j + 1] - 1; q >= chdptr[j]; --q) {
            i64 i = chdtgt[q - 1] - 1;
            i64 ma = sepptr[i + 1] - sepptr[i];
            const i64* inj = reltgt + sepptr[i] - 1;
            const double* M = Mval + Mptr[ns];
            for (i64 c = 0; c < nrhs; ++c)
                for (i64 r = 0; r < ma; ++r) A_(F, nj, inj[r] - 1, c) += A_(M, ma, r, c);
            ns -= 1;
        }
        double* C1 = C + r0;
        for (i64 c = 0; c < nrhs; ++c)
            for (i64 r = 0; r < nn; ++r) A_(C1, ldc, r, c) += A_(F, nj, r, c);
        const double* D11 = Dval + Dptr[j] - 1;
        const double* L21 = Lval + Lptr[j] - 1;
        trsm_left(D11, nn, C1, ldc, nn, nrhs, 0, diag);
        if (na > 0) {
            ns += 1;
            i64 strt = Mptr[ns];
            Mptr[ns + 1] = strt + na * nrhs;
            double* M = Mval + strt;
            for (i64 c = 0; c < nrhs; ++c)
                for (i64 r = 0; r < na; ++r) A_(M, na, r, c) = A_(F, nj, nn + r, c);
            gemm_sub(L21, na, 0, C1, ldc, M, na, na, nn, nrhs);
        }
    }
}

/* Backward solve C <- L^-T C (div_impl!, backward branch). */
void or_div_bwd(i64 nfr, const i64* resptr, const i64* sepptr, const i64* reltgt,
                const i64* chdptr, const i64* chdtgt, const i64* Dptr, const i64* Lptr,
                const double* Dval, const double* Lval, int diag, double* C, i64 ldc,
                i64 nrhs, double* Mval, i64* Mptr, double* Fval) {
    i64 ns = 0;
    Mptr[1] = 0;
    for (i64 j = nfr - 1; j >= 0; --j) {
        i64 r0 = resptr[j] - 1;
        i64 nn = resptr[j + 1] - resptr[j];
        i64 na = sepptr[j + 1] - sepptr[j];
        i64 nj = nn + na;
        double* C1 = C + r0;
        const double* D11 = Dval + Dptr[j] - 1;
        const double* L21 = Lval + Lptr[j] - 1;
        const double* M2 = NULL;
        if (na > 0) {
            M2 = Mval + Mptr[ns];
            ns -= 1;
            gemm_sub(L21, na, 1, M2, na, C1, ldc, nn, na, nrhs);
        }
        trsm_left(D11, nn, C1, ldc, nn, nrhs, 1, diag);
        double* F = Fval;
        for (i64 c = 0; c < nrhs; ++c) {
            for (i64 r = 0; r < nn; ++r) A_(F, nj, r, c) = A_(C1, ldc, r, c);
            for (i64 r = 0; r < na; ++r) A_(F, nj, nn + r, c) = A_(M2, na, r, c);
        }
        for (i64 q = chdptr[j]; q < chdptr[j + 1]; ++q) {
            i64 i = chdtgt[q - 1] - 1;
            i64 ma = sepptr[i + 1] - sepptr[i];
            const i64* inj = reltgt + sepptr[i] - 1;
            ns += 1;
            i64 strt = Mptr[ns];
            Mptr[ns + 1] = strt + ma * nrhs;
            double* M = Mval + strt;
            for (i64 c = 0; c < nrhs; ++c)
                for (i64 r = 0; r < ma; ++r) A_(M, ma, r, c) = A_(F, nj, inj[r] - 1, c);
        }
    }
}

/* ---- selected inversion -------------------------------------------------- */

/*
 * or_selinv: the entries of (P A P')^-1 on the pattern of the Cholesky factor,
 * written over the factor in place (Dval: the lower triangle of each diagonal
 * block, strict upper set to 0; Lval: the rows below).  Restates
 *   selinv_impl!       selinv.jl:72-94   (root first: j = nfr, ..., 1)
 *   selinv_loop_snd!   selinv.jl:122-260
 *   selinv_loop_nod!   selinv.jl:262-379 (nn = 1)
 *   selinv_send!       selinv.jl:381-424 (copygathertri!, utils.jl:529-553)
 * with trsm!(:R, :L, :N), symm!(:L, :L), trrk! (= syr2k! with alpha / 2 for
 * Float64, blas/trrk.jl:5-8), trtri! and syrk!(:L, :C) as plain loops.  The
 * stack discipline is the reference's: the parent pushes one update matrix
 * M_i = R_i' F R_i per child in increasing child order, and the next front
 * in the reverse walk (its last child) pops the top.  Mptr is a 1-based
 * stack index with 0-based offsets; Mval must hold the stack's peak.
 * Returns 0.
 */
i64 or_selinv(i64 nfr, const i64* resptr, const i64* sepptr, const i64* reltgt,
              const i64* chdptr, const i64* chdtgt, const i64* Dptr, const i64* Lptr,
              double* Dval, double* Lval, double* Mval, i64* Mptr, double* Fval, double* Xval) {
    i64 ns = 0;
    Mptr[1] = 0;
    for (i64 j = nfr - 1; j >= 0; --j) {
        i64 nn = resptr[j + 1] - resptr[j];
        i64 na = sepptr[j + 1] - sepptr[j];
        i64 nj = nn + na;
        double* F = Fval;
        double* D11 = Dval + Dptr[j] - 1;
        double* L21 = Lval + Lptr[j] - 1;
        const double* M = NULL;
        if (na > 0) {
            M = Mval + Mptr[ns];
            ns -= 1;
            /* F22 <- M22 (copytri!, lower) */
            for (i64 c = 0; c < na; ++c)
                for (i64 r = c; r < na; ++r) A_(F, nj, nn + r, nn + c) = A_(M, na, r, c);
        }
        if (nn == 1) {
            /* selinv_loop_nod! */
            double d11 = D11[0];
            double f11 = 0.0;
            if (na > 0) {
                for (i64 r = 0; r < na; ++r) L21[r] /= d11;
                /* f21 <- -M22 l21 (symv!, lower) */
                for (i64 r = 0; r < na; ++r) {
                    double s = 0.0;
                    for (i64 k = 0; k < na; ++k)
                        s += (r >= k ? A_(M, na, r, k) : A_(M, na, k, r)) * L21[k];
                    A_(F, nj, 1 + r, 0) = -s;
                }
                double dot = 0.0;
                for (i64 r = 0; r < na; ++r) dot += L21[r] * A_(F, nj, 1 + r, 0);
                f11 -= dot;
            }
            d11 = 1.0 / d11;
            f11 += d11 * d11;
            D11[0] = f11;
            F[0] = f11;
            for (i64 r = 0; r < na; ++r) L21[r] = A_(F, nj, 1 + r, 0);
        } else {
            /* selinv_loop_snd! */
            for (i64 c = 0; c < nn; ++c)
                for (i64 r = 0; r < nn; ++r) A_(F, nj, r, c) = 0.0;
            if (na > 0) {
                /* L21 <- L21 D11^-1 (trsm!(:R, :L, :N, :N)): columns last to first */
                for (i64 c = nn - 1; c >= 0; --c) {
                    for (i64 k = c + 1; k < nn; ++k) {
                        double a = A_(D11, nn, k, c);
                        for (i64 r = 0; r < na; ++r) A_(L21, na, r, c) -= A_(L21, na, r, k) * a;
                    }
                    double inv = 1.0 / A_(D11, nn, c, c);
                    for (i64 r = 0; r < na; ++r) A_(L21, na, r, c) *= inv;
                }
                /* F21 <- -M22 L21 (symm!(:L, :L)) */
                for (i64 c = 0; c < nn; ++c)
                    for (i64 r = 0; r < na; ++r) {
                        double s = 0.0;
                        for (i64 k = 0; k < na; ++k)
                            s += (r >= k ? A_(M, na, r, k) : A_(M, na, k, r)) * A_(L21, na, k, c);
                        A_(F, nj, nn + r, c) = -s;
                    }
                /* F11 <- F11 - (L21' F21 + F21' L21) / 2 on the lower triangle (trrk!) */
                for (i64 c = 0; c < nn; ++c)
                    for (i64 r = c; r < nn; ++r) {
                        double s = 0.0;
                        for (i64 k = 0; k < na; ++k)
                            s += A_(L21, na, k, r) * A_(F, nj, nn + k, c) +
                                 A_(F, nj, nn + k, r) * A_(L21, na, k, c);
                        A_(F, nj, r, c) -= 0.5 * s;
                    }
            }
            /* D11 <- D11^-1 (trtri!(:L, :N)) into X, column by column */
            double* X = Xval;
            for (i64 c = 0; c < nn; ++c) {
                for (i64 r = 0; r < c; ++r) A_(X, nn, r, c) = 0.0;
                A_(X, nn, c, c) = 1.0 / A_(D11, nn, c, c);
                for (i64 r = c + 1; r < nn; ++r) {
                    double s = 0.0;
                    for (i64 k = c; k < r; ++k) s += A_(D11, nn, r, k) * A_(X, nn, k, c);
                    A_(X, nn, r, c) = -s / A_(D11, nn, r, r);
                }
            }
            /* F11 <- F11 + D11' D11 (syrk!(:L, :C) with D11 = tril(D11^-1)) */
            for (i64 c = 0; c < nn; ++c)
                for (i64 r = c; r < nn; ++r) {
                    double s = 0.0;
                    for (i64 k = r; k < nn; ++k) s += A_(X, nn, k, r) * A_(X, nn, k, c);
                    A_(F, nj, r, c) += s;
                }
            /* D11 <- F11 (copyrec!: the strict upper triangle of F11 is 0), L21 <- F21 */
            for (i64 c = 0; c < nn; ++c) {
                for (i64 r = 0; r < nn; ++r) A_(D11, nn, r, c) = A_(F, nj, r, c);
                for (i64 r = 0; r < na; ++r) A_(L21, na, r, c) = A_(F, nj, nn + r, c);
            }
        }
        /* selinv_send!: M_i <- R_i' F R_i (lower) for each child, increasing order */
        for (i64 q = chdptr[j]; q < chdptr[j + 1]; ++q) {
            i64 i = chdtgt[q - 1] - 1;
            i64 ma = sepptr[i + 1] - sepptr[i];
            const i64* inj = reltgt + sepptr[i] - 1;
            ns += 1;
            i64 strt = Mptr[ns];
            Mptr[ns + 1] = strt + ma * ma;
            double* Mi = Mval + strt;
            for (i64 c = 0; c < ma; ++c)
                for (i64 r = c; r < ma; ++r) A_(Mi, ma, r, c) = A_(F, nj, inj[r] - 1, inj[c] - 1);
        }
    }
    return 0;
}

/* ---- completion ---------------------------------------------------------- */

/* In-place lower Cholesky of the n x n lower triangle of A (ld lda), the
 * unblocked column algorithm; returns 0 or the 1-based failing column. */
static i64 chol_lower(double* A, i64 lda, i64 n) {
    for (i64 j = 0; j < n; ++j) {
        double d = A_(A, lda, j, j);
        if (!(d > 0.0)) return j + 1;
        d = sqrt(d);
        A_(A, lda, j, j) = d;
        for (i64 i = j + 1; i < n; ++i) A_(A, lda, i, j) /= d;
        for (i64 k = j + 1; k < n; ++k) {
            double l = A_(A, lda, k, j);
            for (i64 i = k; i < n; ++i) A_(A, lda, i, k) -= A_(A, lda, i, j) * l;
        }
    }
    return 0;
}

/*
 * or_complete: complete_impl! (complete.jl:46-72) with complete_loop_snd!
 * (:108-234), complete_loop_nod! (:237-402) and complete_send! (:404-451):
 * Dval/Lval hold a matrix Sigma on the factor's pattern (lower triangles);
 * on return they hold the Cholesky factor of the inverse of Sigma's
 * maximum-determinant positive definite completion.  Per front, root first:
 * F <- [Sigma11; Sigma21] and M22 from the stack (the parent's frontal
 * matrix gathered through rel), C = chol(M22), L21 <- C^-1 L21,
 * D11 <- D11 - L21' L21, L21 <- -C^-T L21, D11 <- R with D11 = R' R
 * (R lower: revtri!, potrf!(:U), revtri!), L21 <- L21 R^-1, D11 <- R^-1
 * (trtri!); children gather their M from F, which keeps the unmodified
 * Sigma.  Returns 0, the failing local column of a potrf!, or 1 for a
 * non-positive nn = 1 pivot (complete_loop_nod!'s convention).
 */
i64 or_complete(i64 nfr, const i64* resptr, const i64* sepptr, const i64* reltgt,
                const i64* chdptr, const i64* chdtgt, const i64* Dptr, const i64* Lptr,
                double* Dval, double* Lval, double* Mval, i64* Mptr, double* Fval) {
    i64 ns = 0;
    Mptr[1] = 0;
    for (i64 j = nfr - 1; j >= 0; --j) {
        i64 nn = resptr[j + 1] - resptr[j];
        i64 na = sepptr[j + 1] - sepptr[j];
        i64 nj = nn + na;
        double* F = Fval;
        double* D11 = Dval + Dptr[j] - 1;
        double* L21 = Lval + Lptr[j] - 1;
        /* F11 <- D11 (lower), F21 <- L21 */
        for (i64 c = 0; c < nn; ++c) {
            for (i64 r = c; r < nn; ++r) A_(F, nj, r, c) = A_(D11, nn, r, c);
            for (i64 r = 0; r < na; ++r) A_(F, nj, nn + r, c) = A_(L21, na, r, c);
        }
        double* M = NULL;
        if (na > 0) {
            M = Mval + Mptr[ns];
            ns -= 1;
            for (i64 c = 0; c < na; ++c)
                for (i64 r = c; r < na; ++r) A_(F, nj, nn + r, nn + c) = A_(M, na, r, c);
            i64 info = chol_lower(M, na, na);
            if (info) return info;
            /* L21 <- C^-1 L21 (forward substitution, column by column) */
            for (i64 c = 0; c < nn; ++c)
                for (i64 r = 0; r < na; ++r) {
                    double s = A_(L21, na, r, c);
                    for (i64 k = 0; k < r; ++k) s -= A_(M, na, r, k) * A_(L21, na, k, c);
                    A_(L21, na, r, c) = s / A_(M, na, r, r);
                }
            if (nn == 1) {
                double dot = 0.0;
                for (i64 r = 0; r < na; ++r) dot += L21[r] * L21[r];
                D11[0] -= dot;
            } else {
                /* D11 <- D11 - L21' L21 (lower) */
                for (i64 c = 0; c < nn; ++c)
                    for (i64 r = c; r < nn; ++r) {
                        double s = 0.0;
                        for (i64 k = 0; k < na; ++k) s += A_(L21, na, k, r) * A_(L21, na, k, c);
                        A_(D11, nn, r, c) -= s;
                    }
            }
            /* L21 <- -C^-T L21 (back substitution) */
            for (i64 c = 0; c < nn; ++c)
                for (i64 r = na - 1; r >= 0; --r) {
                    double s = A_(L21, na, r, c);
                    for (i64 k = r + 1; k < na; ++k) s -= A_(M, na, k, r) * A_(L21, na, k, c);
                    A_(L21, na, r, c) = s / A_(M, na, r, r);
                }
            for (i64 c = 0; c < nn; ++c)
                for (i64 r = 0; r < na; ++r) A_(L21, na, r, c) = -A_(L21, na, r, c);
        }
        if (nn == 1) {
            double d11 = D11[0];
            if (!(d11 > 0.0)) return 1;
            d11 = sqrt(d11);
            for (i64 r = 0; r < na; ++r) L21[r] /= d11;
            D11[0] = 1.0 / d11;
        } else {
            /* D11 = R' R with R lower: the lower Cholesky of the reversed
             * matrix J D11 J is J R' J (revtri!, potrf!(:U), revtri!) */
            double* Rv = (double*)malloc((size_t)(nn * nn) * sizeof(double));
            for (i64 c = 0; c < nn; ++c)
                for (i64 r = c; r < nn; ++r) A_(Rv, nn, nn - 1 - c, nn - 1 - r) = A_(D11, nn, r, c);
            /* Rv now holds J D11 J in its lower triangle (entry (n-1-c, n-1-r) with n-1-c >= n-1-r) */
            i64 info = chol_lower(Rv, nn, nn);
            if (info) {
                free(Rv);
                return info;
            }
            /* R = J Rv' J: R(r, c) = Rv(n-1-c, n-1-r), lower */
            for (i64 c = 0; c < nn; ++c) {
                for (i64 r = 0; r < c; ++r) A_(D11, nn, r, c) = 0.0;
                for (i64 r = c; r < nn; ++r) A_(D11, nn, r, c) = A_(Rv, nn, nn - 1 - c, nn - 1 - r);
            }
            free(Rv);
            /* L21 <- L21 R^-1 (right solve, R lower: columns last to first) */
            for (i64 c = nn - 1; c >= 0; --c) {
                for (i64 k = c + 1; k < nn; ++k) {
                    double a = A_(D11, nn, k, c);
                    for (i64 r = 0; r < na; ++r) A_(L21, na, r, c) -= A_(L21, na, r, k) * a;
                }
                double inv = 1.0 / A_(D11, nn, c, c);
                for (i64 r = 0; r < na; ++r) A_(L21, na, r, c) *= inv;
            }
            /* D11 <- R^-1 (lower) */
            double* X = (double*)malloc((size_t)(nn * nn) * sizeof(double));
            for (i64 c = 0; c < nn; ++c) {
                for (i64 r = 0; r < c; ++r) A_(X, nn, r, c) = 0.0;
                A_(X, nn, c, c) = 1.0 / A_(D11, nn, c, c);
                for (i64 r = c + 1; r < nn; ++r) {
                    double s = 0.0;
                    for (i64 k = c; k < r; ++k) s += A_(D11, nn, r, k) * A_(X, nn, k, c);
                    A_(X, nn, r, c) = -s / A_(D11, nn, r, r);
                }
            }
            memcpy(D11, X, (size_t)(nn * nn) * sizeof(double));
            free(X);
        }
        /* complete_send!: M_i <- R_i' F R_i (lower), increasing child order */
        for (i64 q = chdptr[j]; q < chdptr[j + 1]; ++q) {
            i64 i = chdtgt[q - 1] - 1;
            i64 ma = sepptr[i + 1] - sepptr[i];
            const i64* inj = reltgt + sepptr[i] - 1;
            ns += 1;
            i64 strt = Mptr[ns];
            Mptr[ns + 1] = strt + ma * ma;
            double* Mi = Mval + strt;
            for (i64 c = 0; c < ma; ++c)
                for (i64 r = c; r < ma; ++r) A_(Mi, ma, r, c) = A_(F, nj, inj[r] - 1, inj[c] - 1);
        }
    }
    return 0;
}
