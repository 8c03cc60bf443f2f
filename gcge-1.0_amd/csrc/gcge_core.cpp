// gcge_core.cpp -- the GCG eigensolver driven through GCGE_OPS.
//
// Same algorithm, parameters and control flow as the reference's default
// path (src/gcge_eigsol.c, gcge_xpw.c, gcge_cg.c, gcge_orthogonal.c,
// gcge_rayleighritz.c, gcge_workspace.c, gcge_solver.c).  What changes is
// how the large operations are issued: the reference loops over single
// vectors (one MatDotVec / VecAxpby / VecInnerProd per column); here every
// large-scale step is one block operation on device multivectors:
//   - block CG: one SpMM+dot, one p-update and one fused x/r update per sweep
//     for all active columns (GCGE_BCG, gcge_cg.c:137-302);
//   - projection V2 -= V1 (V1^T B V2): one Gram and one fused linear
//     combination (GCGE_SubOrthonormalization, gcge_orthogonal.c:1880-1929);
//   - residuals: one fused kernel per batch of pairs
//     (GCGE_CheckConvergence, gcge_eigsol.c:559-686).
// The small dense work (subspace matrices, in-subspace orthogonalization,
// dsyevx) stays on the host in FP64 and calls the same LAPACK routines as the
// reference.  Where the reference exchanges vector pointers between V and
// evec/V_tmp (MultiVecSwap), this code copies: the columns that the swap
// would hand back to evec/V_tmp are never read before being overwritten.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gcge.h"

// error latch of the library (hip_backend.cpp): read by GCGE_HIP_GetLastError
namespace gh {
void latch(int code, const char *fmt, ...);
bool failed();
}  // namespace gh

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Sec {
    double *acc_total, *acc_iter, t0;
    Sec(double *tot, double *it) : acc_total(tot), acc_iter(it), t0(now_s()) {}
    ~Sec()
    {
        double d = now_s() - t0;
        *acc_total += d;
        *acc_iter = d;
    }
};
#define GCGE_TIMER(name, para) Sec _sec_##name(&(para)->stat_para->part_time_total->name, \
                                               &(para)->stat_para->part_time_one_iter->name)

inline void set2(GCGE_INT *a, int x, int y)
{
    a[0] = x;
    a[1] = y;
}

// ------------------------------------------------------------------------
// Orthonormalization (src/gcge_orthogonal.c)
// ------------------------------------------------------------------------

// GCGE_Orthonormalization, gcge_orthogonal.c:112-305.  Block version: the
// B-inner products of column `current` with columns 0..current come from
// one Gram call, the projection is one fused combination.
void orth_gs(void **V, int start, int *end, void *B, GCGE_OPS *ops, GCGE_PARA *para, void **V_tmp,
             double *d_tmp)
{
    GCGE_ORTH_PARA *op = para->orth_para;
    GCGE_INT ms[2], me[2];
    double vin = 0.0, vout = 0.0, norm_value = 0.0, ipv = 0.0, ratio = 1.0;
    const double sum_res = para->sum_res, crit = op->criterion_tol;
    for (int current = start; current < *end; ++current) {
        int reorth_count = 0;
        if (sum_res > 0.0 && sum_res < crit) {
            // column version, :139-208 (its reorthogonalization counter is
            // never incremented in the reference; capped here)
            void *vc, *t2 = nullptr;
            ops->GetVecFromMultiVec(V, current, &vc);
            if (B) ops->GetVecFromMultiVec(V_tmp, 1, &t2);
            if (!B) {
                ops->VecInnerProd(vc, vc, &norm_value);
            } else {
                ops->MatDotVec(B, vc, t2);
                ops->VecInnerProd(vc, t2, &norm_value);
            }
            vout = std::sqrt(norm_value);
            int guard = 0;
            do {
                for (int j = 0; j < current; j++) {
                    void *vj;
                    ops->GetVecFromMultiVec(V, j, &vj);
                    if (!B) {
                        ops->VecInnerProd(vc, vj, &ipv);
                    } else {
                        ops->MatDotVec(B, vc, t2);
                        ops->VecInnerProd(vj, t2, &ipv);
                    }
                    ops->VecAxpby(-ipv, vj, 1.0, vc);
                }
                vin = vout;
                if (!B) {
                    ops->VecInnerProd(vc, vc, &norm_value);
                } else {
                    ops->MatDotVec(B, vc, t2);
                    ops->VecInnerProd(vc, t2, &norm_value);
                }
                vout = std::sqrt(norm_value);
                ratio = vout / vin;
                guard++;
            } while (ratio < op->reorth_tol && reorth_count < op->max_reorth_time && vout > op->orth_zero_tol &&
                     guard < 1000);
        } else {
            // block version, :209-273
            do {
                void **T = V_tmp;
                int tc = 0;
                if (!B) {
                    T = V;
                    tc = current;
                } else {
                    set2(ms, current, 0);
                    set2(me, current + 1, 1);
                    ops->MatDotMultiVec(B, V, V_tmp, ms, me, ops);
                }
                set2(ms, 0, tc);
                set2(me, current + 1, tc + 1);
                ops->MultiVecInnerProd(V, T, d_tmp, (char *)"ns", ms, me, current + 1, ops);
                if (current != 0) {
                    set2(ms, 0, current);
                    set2(me, current, current + 1);
                    ops->MultiVecLinearCombAxpby(V, V, ms, me, d_tmp, current, -1.0, 1.0, ops);
                }
                vin = std::sqrt(d_tmp[current]);
                d_tmp[current] = d_tmp[current] - ops->ArrayDotArray(d_tmp, d_tmp, current);
                vout = std::sqrt(d_tmp[current]);
                ratio = vout / vin;
                reorth_count++;
            } while (ratio < op->reorth_tol && reorth_count < op->max_reorth_time && vout > op->orth_zero_tol);
        }
        if (!std::isfinite(vin)) {
            // a non-finite B-norm (NaN or Inf in the vector or in B) would pass
            // as a zero vector and be deflated; a NaN vout from a negative norm
            // estimate is the reference's zero-vector signal and is kept
            gh::latch(191, "Orthonormalization: B-norm of column %d is %g", current, vin);
            *end = current;
            return;
        }
        if (vout > op->orth_zero_tol) {
            ops->MultiVecAxpbyColumn(0.0, V, current, 1 / vout, V, current, ops);
        } else {
            set2(ms, current, *end - 1);
            set2(me, current + 1, *end);
            ops->MultiVecSwap(V, V, ms, me, ops);
            (*end)--;
            current--;
            if (op->print_orth_zero == 1)
                printf("In Orthonormalization, there is a zero vector!, current = %d, start = %d, end = %d\n",
                       current, start, *end);
        }
    }
}

// GCGE_SubOrthonormalization, gcge_orthogonal.c:1880-1929:
// V2 = V2 - V1 (V1^T B V2), as one Gram and one fused combination.  With
// wn2 (not null), the squared B-norms of the V2 columns before the
// projection are written there as well.
void sub_orth(void **V, const GCGE_INT *start, const GCGE_INT *end, void *B, void **V_tmp, double *d, GCGE_OPS *ops,
              double *wn2 = nullptr)
{
    const int len1 = end[0] - start[0], len2 = end[1] - start[1];
    GCGE_INT ms[2], me[2];
    void **T = V;
    int tc = start[1];
    if (B) {
        set2(ms, start[1], 0);
        set2(me, end[1], len2);
        ops->MatDotMultiVec(B, V, V_tmp, ms, me, ops);
        T = V_tmp;
        tc = 0;
    }
    set2(ms, start[0], tc);
    set2(me, end[0], tc + len2);
    ops->MultiVecInnerProd(V, T, d, (char *)"ns", ms, me, len1, ops);
    if (wn2) {
        set2(ms, start[1], tc);
        set2(me, end[1], tc + len2);
        ops->MultiVecInnerProdDiag(V, T, wn2, ms, me, ops);
    }
    set2(ms, start[0], start[1]);
    set2(me, end[0], end[1]);
    ops->MultiVecLinearCombAxpby(V, V, ms, me, d, len1, -1.0, 1.0, ops);
}

// GCGE_SubOrthonormalizationSelfBGS, gcge_orthogonal.c:2035-2115.
void self_bgs(void **V, int start, int *end, void *B, GCGE_PARA *para, GCGE_OPS *ops, GCGE_WORKSPACE *ws)
{
    GCGE_INT ms[2], me[2];
    double *d = ws->subspace_dtmp;
    const double tol = para->orth_para->orth_zero_tol;
    std::vector<double> coef;
    for (int current = start; current < *end; current++) {
        void **T = V;
        int tc = current;
        if (B) {
            set2(ms, current, 0);
            set2(me, current + 1, 1);
            ops->MatDotMultiVec(B, V, ws->V_tmp, ms, me, ops);
            T = ws->V_tmp;
            tc = 0;
        }
        set2(ms, current, tc);
        set2(me, *end, tc + 1);
        const int length = *end - current;
        ops->MultiVecInnerProd(V, T, d, (char *)"ns", ms, me, length, ops);
        const double norm_value = std::sqrt(d[0]);
        if (norm_value > tol) {
            ops->MultiVecAxpbyColumn(0.0, V, current, 1.0 / norm_value, V, current, ops);
            if (current + 1 < *end) {
                // V[j] += (-d[j-current]/norm) V[current], j > current
                coef.assign(*end - current - 1, 0.0);
                for (int j = current + 1; j < *end; j++) coef[j - current - 1] = -d[j - current] / norm_value;
                set2(ms, current, current + 1);
                set2(me, current + 1, *end);
                ops->MultiVecLinearCombAxpby(V, V, ms, me, coef.data(), 1, 1.0, 1.0, ops);
            }
        } else {
            set2(ms, current, (*end) - 1);
            set2(me, current + 1, *end);
            ops->MultiVecSwap(V, V, ms, me, ops);
            (*end)--;
            current--;
            if (para->orth_para->print_orth_zero == 1)
                printf("In GCGE_SubOrthonormalizationSelfBGS, there is a zero vector!, current = %d, start = %d, "
                       "end = %d\n",
                       current, start, *end);
        }
    }
}

// GCGE_MultiOrthonormalization, gcge_orthogonal.c:1945-2033.
void multi_orth(void **V, int start, int *end, void *B, GCGE_OPS *ops, GCGE_PARA *para, GCGE_WORKSPACE *ws)
{
    const int length = *end - start;
    const int length_1 = length / 2, length_2 = length - length_1;
    int mid = start + length_1;
    GCGE_INT ms[2], me[2];
    const double tol = para->orth_para->scbgs_reorth_tol;
    if (length <= para->orth_para->max_direct_orth_length) {
        self_bgs(V, start, end, B, para, ops, ws);
        return;
    }
    multi_orth(V, start, &mid, B, ops, para, ws);
    if (mid + length_2 < *end) {
        const int new_end_2 = mid + length_2, n_zero = *end - new_end_2;
        set2(ms, mid, new_end_2);
        set2(me, mid + n_zero, *end);
        ops->MultiVecSwap(V, V, ms, me, ops);
        *end = new_end_2;
    }
    if (mid - start > 0) {
        double value_inner = 1.0;
        int iter = 0;
        while (value_inner > tol) {
            iter++;
            set2(ms, start, mid);
            set2(me, mid, *end);
            sub_orth(V, ms, me, B, ws->V_tmp, ws->subspace_dtmp, ops);
            // the reference sums length_1*length_2 entries with the
            // pre-deflation length_1 (gcge_orthogonal.c:2014)
            value_inner = 0.0;
            for (int i = 0; i < length_1 * length_2; i++) value_inner += ws->subspace_dtmp[i] * ws->subspace_dtmp[i];
            value_inner = std::sqrt(value_inner);
            if (iter >= para->orth_para->max_reorth_time) break;
        }
        if (value_inner > tol && para->print_eval == 1)
            printf("in GCGE_SubOrthonormalization, V1: %d-%d, V2: %d-%d, reorth_count: %d, value_inner: %e\n", start,
                   mid - 1, mid, *end, para->orth_para->max_reorth_time, value_inner);
    }
    multi_orth(V, mid, end, B, ops, para, ws);
}

// GCGE_StableMultiOrthonormalization, gcge_orthogonal.c:1833-1877.
void stable_multi_orth(void **V, int start, int *end, void *B, GCGE_OPS *ops, GCGE_PARA *para, GCGE_WORKSPACE *ws)
{
    GCGE_INT ms[2], me[2];
    double value_inner = 1.0;
    int iter = 0;
    set2(ms, 0, start);
    set2(me, start, *end);
    while (value_inner > para->orth_para->scbgs_reorth_tol) {
        iter++;
        sub_orth(V, ms, me, B, ws->V_tmp, ws->subspace_dtmp, ops);
        value_inner = 0.0;
        for (int i = 0; i < start * (*end - start); i++) value_inner += ws->subspace_dtmp[i] * ws->subspace_dtmp[i];
        value_inner = std::sqrt(value_inner);
        if (iter >= para->orth_para->max_reorth_time) break;
    }
    multi_orth(V, start, end, B, ops, para, ws);
}

// ---- the other W orthonormalizations of the reference (w_orth_type "bgs",
// "cbgs", "scbgs"), issued with the same per-step structure on block slots.

// Y = B X for columns [xs, xe) of X into columns [ys, ...) of Y; a copy
// when B is NULL (the reference's VecAxpby(1.0, x, 0.0, y)).
void b_times(void *B, void **X, int xs, int xe, void **Y, int ys, GCGE_OPS *ops)
{
    if (xe <= xs) return;
    GCGE_INT ms[2], me[2];
    set2(ms, xs, ys);
    set2(me, xe, ys + (xe - xs));
    if (B) ops->MatDotMultiVec(B, X, Y, ms, me, ops);
    else ops->MultiVecCopy(X, Y, ms, me, ops);
}

// V[j] -= d[j - s] V[cur] for j in [s, e): one rank-1 combination.
void minus_rank1(void **V, int cur, int s, int e, const double *d, double scale, GCGE_OPS *ops,
                 std::vector<double> &coef)
{
    if (e <= s) return;
    GCGE_INT ms[2], me[2];
    coef.resize(e - s);
    for (int j = s; j < e; j++) coef[j - s] = -d[j - s] / scale;
    set2(ms, cur, s);
    set2(me, cur + 1, e);
    ops->MultiVecLinearCombAxpby(V, V, ms, me, coef.data(), 1, 1.0, 1.0, ops);
}

// Modified Gram-Schmidt of V[start:end] against itself in the B inner
// product, column by column, with zero columns swapped to the end and
// dropped: the W-self loop of GCGE_BOrthonormalization
// (gcge_orthogonal.c:459-526) and of GCGE_SCBOrth_Self (:1196-1279).
void w_self_mgs(void **V, int start, int *end, void *B, GCGE_OPS *ops, GCGE_PARA *para, GCGE_WORKSPACE *ws,
                const char *who)
{
    GCGE_ORTH_PARA *op = para->orth_para;
    double *d = ws->subspace_dtmp;
    GCGE_INT ms[2], me[2];
    std::vector<double> coef;
    for (int cur = start; cur < *end; cur++) {
        b_times(B, V, cur, cur + 1, ws->CG_p, 0, ops);
        set2(ms, cur, 0);
        set2(me, *end, 1);
        ops->MultiVecInnerProd(V, ws->CG_p, d, (char *)"ns", ms, me, *end - cur, ops);
        const double norm_value = std::sqrt(d[0]);
        if (norm_value > op->orth_zero_tol) {
            ops->MultiVecAxpbyColumn(0.0, V, cur, 1.0 / norm_value, V, cur, ops);
            minus_rank1(V, cur, cur + 1, *end, d + 1, norm_value, ops, coef);
        } else {
            set2(ms, cur, *end - 1);
            set2(me, cur + 1, *end);
            ops->MultiVecSwap(V, V, ms, me, ops);
            (*end)--;
            cur--;
            if (op->print_orth_zero == 1)
                printf("In %s, there is a zero vector!, current = %d, start = %d, end = %d\n", who, cur, start, *end);
        }
    }
}

// GCGE_BOrthonormalization, gcge_orthogonal.c:352-528: evec = B V[0:nev]
// and V_tmp = B V[nev:dim_xp] once; then twice: W -= v_i (v_i^T B W) for
// each X and P column in turn, and the W-self MGS.
void orth_bgs(void **V, int start, int *end, void *B, GCGE_OPS *ops, GCGE_PARA *para, GCGE_WORKSPACE *ws)
{
    const int nev = para->nev, dim_xp = ws->dim_xp;
    double *d = ws->subspace_dtmp;
    GCGE_INT ms[2], me[2];
    std::vector<double> coef;
    b_times(B, V, 0, nev, ws->evec, 0, ops);
    b_times(B, V, nev, dim_xp, ws->V_tmp, 0, ops);
    for (int pass = 0; pass < 2; pass++) {
        const int wl = *end - start;
        for (int cur = 0; cur < dim_xp; cur++) {
            void **BV = cur < nev ? ws->evec : ws->V_tmp;
            const int bc = cur < nev ? cur : cur - nev;
            set2(ms, start, bc);
            set2(me, *end, bc + 1);
            ops->MultiVecInnerProd(V, BV, d, (char *)"ns", ms, me, wl, ops);
            minus_rank1(V, cur, start, *end, d, 1.0, ops, coef);
        }
        w_self_mgs(V, start, end, B, ops, para, ws, "Orthonormalization");
    }
}

// W -= Vp (Vp^T B W) with Vp = V[p0:p1] and W = V[start:end]: V_tmp = B W,
// d = Vp^T V_tmp (leading dimension lda), V_tmp = Vp d, W -= V_tmp
// (gcge_orthogonal.c:822-855).  Returns sqrt of the sum of the squares of
// d[0:nsum] (the reference's value_inner, read with the same extent).
double project_out(void **V, int p0, int p1, int start, int end, int lda, int nsum, void *B, GCGE_OPS *ops,
                   GCGE_WORKSPACE *ws)
{
    const int wl = end - start;
    double *d = ws->subspace_dtmp;
    GCGE_INT ms[2], me[2];
    if (B) {
        set2(ms, start, 0);
        set2(me, end, wl);
        ops->MatDotMultiVec(B, V, ws->V_tmp, ms, me, ops);
    } else {
        set2(ms, start, 0);
        set2(me, end, wl);
        ops->MultiVecAxpby(1.0, V, 0.0, ws->V_tmp, ms, me, ops);
    }
    set2(ms, p0, 0);
    set2(me, p1, wl);
    ops->MultiVecInnerProd(V, ws->V_tmp, d, (char *)"ns", ms, me, lda, ops);
    double value_inner = 0.0;
    for (int i = 0; i < nsum; i++) value_inner += d[i] * d[i];
    ops->MultiVecLinearComb(V, ws->V_tmp, ms, me, d, lda, nullptr, -1, ops);
    set2(ms, 0, start);
    set2(me, wl, end);
    ops->MultiVecAxpby(-1.0, ws->V_tmp, 1.0, V, ms, me, ops);
    return std::sqrt(value_inner);
}

// GCGE_CBOrthonormalization, gcge_orthogonal.c:755-963: two block
// projections against V[0:start], then W = W Y Theta^{-1/2} from the
// eigenpairs of W^T B W (dsyevx), dropping eigenvalues below orth_zero_tol.
void orth_cbgs(void **V, int start, int *end, void *B, GCGE_OPS *ops, GCGE_PARA *para, GCGE_WORKSPACE *ws)
{
    GCGE_ORTH_PARA *op = para->orth_para;
    double *d = ws->subspace_dtmp;
    GCGE_INT ms[2], me[2];
    int wl = *end - start;
    for (int pass = 0; pass < 2; pass++) {
        wl = *end - start;
        project_out(V, 0, start, start, *end, start, 0, B, ops, ws);
    }
    // M = W^T B W
    b_times(B, V, start, *end, ws->V_tmp, 0, ops);
    set2(ms, start, 0);
    set2(me, *end, wl);
    ops->MultiVecInnerProd(V, ws->V_tmp, d, (char *)"ns", ms, me, wl, ops);
    bool nonzero = false;
    for (int j = 0; j < wl * wl; j++)
        if (d[j] > op->orth_zero_tol) {
            nonzero = true;
            break;
        }
    if (!nonzero) {
        printf("W all zero!\n");
        *end = start;
        return;
    }
    // the reference's buffers: eigenvectors and eigenvalues in subspace_evec,
    // LAPACK work behind the matrix in subspace_dtmp, integer work in subspace_itmp
    double *tmp_evec = ws->subspace_evec, *tmp_eval = tmp_evec + (size_t)wl * wl, *dwork = d + (size_t)wl * wl;
    int *itmp = ws->subspace_itmp, *isuppz = itmp + 10 * wl, *ifail = itmp + 5 * wl;
    double vl = 0.0, vu = 0.0, abstol = 0.0;
    int il = 1, iu = wl, m = wl, lwork = 8 * wl, liwork = 5 * wl, info = 0;
    ops->DenseMatEigenSolver((char *)"V", (char *)"I", (char *)"U", &wl, d, &wl, &vl, &vu, &il, &iu, &abstol, &m,
                             tmp_eval, tmp_evec, &wl, isuppz, dwork, &lwork, itmp, &liwork, ifail, &info);
    int nz_start = 0;
    for (int j = 0; j < wl; j++) {
        if (tmp_eval[j] < op->orth_zero_tol) nz_start = j + 1;
        else ops->ArrayScale(1.0 / std::sqrt(tmp_eval[j]), tmp_evec + (size_t)j * wl, wl);
    }
    if (nz_start) {
        printf("tmp_eval_nonzero_start: %d\n", nz_start);
        for (int j = 0; j < wl; j++) printf("tmp_eval[%d]: %e\n", j, tmp_eval[j]);
    }
    const int wl_new = wl - nz_start;
    // V_tmp = W Y[:, nz_start:].  The reference passes the new column count
    // as the leading dimension (gcge_orthogonal.c:938-945), which reads other
    // coefficients once a column has been dropped (nz_start > 0); here Y is
    // read with its own leading dimension wl.  Without dropped columns the two
    // agree.
    set2(ms, start, 0);
    set2(me, *end, wl_new);
    ops->MultiVecLinearComb(V, ws->V_tmp, ms, me, tmp_evec + (size_t)nz_start * wl, wl, nullptr, -1, ops);
    *end = start + wl_new;
    set2(ms, start, 0);
    set2(me, *end, wl_new);
    ops->MultiVecSwap(V, ws->V_tmp, ms, me, ops);
}

// GCGE_SCBOrthonormalization, gcge_orthogonal.c:967-1113: block projections
// of W against V[0:start] until their size is below scbgs_reorth_tol (at most
// max_reorth_time), then W in blocks of w_orth_block_size columns: each block
// is projected against the W columns before it (GCGE_SCBOrth_Minus,
// :1116-1193) and orthonormalized by MGS scbgs_wself_max_reorth_time times
// (GCGE_SCBOrth_Self, :1196-1279); a block's dropped columns are swapped to
// the end of W.
void orth_scbgs(void **V, int start, int *end, void *B, GCGE_OPS *ops, GCGE_PARA *para, GCGE_WORKSPACE *ws)
{
    GCGE_ORTH_PARA *op = para->orth_para;
    const double tol = op->scbgs_reorth_tol;
    GCGE_INT ms[2], me[2];
    {
        const int wl = *end - start;  // fixed over the passes, as in the reference
        double value_inner = 1.0;
        int iter = 0;
        while (value_inner > tol) {
            iter++;
            value_inner = project_out(V, 0, start, start, start + wl, start, start * wl, B, ops, ws);
            if (iter >= op->max_reorth_time) break;
        }
    }
    auto self = [&](int ws0, int *we) {
        for (int it = 0; it < op->scbgs_wself_max_reorth_time; it++)
            w_self_mgs(V, ws0, we, B, ops, para, ws, "Orthonormalization");
    };
    auto minus = [&](int ws0, int we) {
        // the sum runs over (ws0 - dim_xp) * wl entries of a Gram stored with
        // leading dimension ws0, as the reference reads it (:1149-1155)
        const int wl = we - ws0, dim_xp = ws->dim_xp;
        double value_inner = 1.0;
        int iter = 0;
        while (value_inner > tol) {
            iter++;
            value_inner = project_out(V, dim_xp, ws0, ws0, we, ws0, (ws0 - dim_xp) * wl, B, ops, ws);
            if (iter >= op->max_reorth_time) break;
        }
    };
    const int step = std::min(*end - start, op->w_orth_block_size);
    int w_start = start, w_end = std::min(w_start + step, *end);
    self(w_start, &w_end);
    while (w_end < *end) {
        w_start = w_end;
        w_end = std::min(w_start + step, *end);
        const int old_w_end = w_end;
        minus(w_start, w_end);
        self(w_start, &w_end);
        if (w_end < old_w_end) {
            *end -= old_w_end - w_end;
            set2(ms, w_end, *end);
            set2(me, old_w_end, *end + old_w_end - w_end);
            ops->MultiVecSwap(V, V, ms, me, ops);
        }
    }
}

}  // namespace

int gcge_chol_inverse(int n, double *S, int lds);            // gcge_dense.cpp
int gcge_sym_eig(int n, double *S, int lds, double *w, double *Z);
int gcge_sym_eig_range(int n, const double *S, int lds, int il, int iu, double *w, double *Z, int ldz);

namespace {

// Block B-orthonormalization of W = V[start:end] against V[0:start] and
// itself (w_orth_type "cholqr"; no counterpart in the reference, whose
// "multi" recursion issues O(nw log nw) small products).
//
// Each round: two projection passes W -= V1 (V1^T B W), then two
// factorization passes on the Gram S = W^T B W scaled to unit diagonal
// (D S D):
//   - Cholesky D S D = R^T R in FP64 on the host, W <- W D R^{-1}, when every
//     R_jj is at least sqrt(tau);
//   - otherwise the eigendecomposition D S D = Q diag(theta) Q^T, keeping
//     the directions with theta_i > tau theta_max: W <- W D Q_k theta_k^{-1/2}
//     (the block shrinks, as the reference deflates dependent vectors,
//     gcge_orthogonal.c:1991-2001).
// tau = 1e4 basis_eps bounds the condition of the factored Gram, so the
// second pass restores orthonormality to working accuracy.  When a factor
// pass took the eigendecomposition route or had an R_jj below 1e-2, the
// round ends with a check of V1^T B W and another round runs while its
// largest entry exceeds 1e3 basis_eps (at most three rounds).  Columns with
// a B-norm at or below orth_zero_tol are dropped first.
// One factorization pass; returns the smallest R_jj when the Cholesky route
// was taken, 0 otherwise.
double orth_factor_pass(void **V, int start, int *end, void *B, GCGE_OPS *ops, GCGE_PARA *para,
                        GCGE_WORKSPACE *ws)
{
    GCGE_INT ms[2], me[2];
    const double zero_tol = para->orth_para->orth_zero_tol;
    const double tau = 1e4 * para->basis_eps;
    int nw = *end - start;
    if (nw <= 0) return 0.0;
    // drop vanishing columns (B-norms from the Gram diagonal)
    std::vector<double> S, Z, w, C, d;
    for (;;) {
        nw = *end - start;
        if (nw <= 0) return 0.0;
        void **T = V;
        int tc = start;
        if (B) {
            set2(ms, start, 0);
            set2(me, *end, nw);
            ops->MatDotMultiVec(B, V, ws->V_tmp, ms, me, ops);
            T = ws->V_tmp;
            tc = 0;
        }
        S.assign((size_t)nw * nw, 0.0);
        set2(ms, start, tc);
        set2(me, *end, tc + nw);
        ops->MultiVecInnerProd(V, T, S.data(), (char *)"sym", ms, me, nw, ops);
        int drop = -1;
        for (int i = 0; i < nw && drop < 0; i++)
            if (!(S[(size_t)i * nw + i] > zero_tol * zero_tol)) drop = i;
        if (drop < 0) break;
        set2(ms, start + drop, *end - 1);
        set2(me, start + drop + 1, *end);
        ops->MultiVecSwap(V, V, ms, me, ops);
        (*end)--;
        if (para->orth_para->print_orth_zero == 1)
            printf("In CholQR orthonormalization, column %d vanishes; end = %d\n", start + drop, *end);
    }
    d.resize(nw);
    for (int i = 0; i < nw; i++) d[i] = 1.0 / std::sqrt(S[(size_t)i * nw + i]);
    for (int j = 0; j < nw; j++)
        for (int i = 0; i < nw; i++) S[(size_t)j * nw + i] *= d[i] * d[j];
    std::vector<double> S0(S);
    int k = nw;
    C.assign((size_t)nw * nw, 0.0);
    const int info = gcge_chol_inverse(nw, S.data(), nw);
    bool chol_ok = info == 0;
    if (info < 0) {
        fprintf(stderr, "GCGE: CholQR needs dpotrf/dtrtri from the LAPACK library\n");
        abort();
    }
    double rmin = 1.0;
    if (chol_ok)
        for (int j = 0; j < nw; j++) rmin = std::min(rmin, 1.0 / S[(size_t)j * nw + j]);
    chol_ok = chol_ok && rmin > std::sqrt(tau);
    if (chol_ok) {
        for (int j = 0; j < nw; j++)
            for (int i = 0; i <= j; i++) C[(size_t)j * nw + i] = d[i] * S[(size_t)j * nw + i];
    } else {
        w.resize(nw);
        Z.resize((size_t)nw * nw);
        if (gcge_sym_eig(nw, S0.data(), nw, w.data(), Z.data()) != 0) {
            fprintf(stderr, "GCGE: eigendecomposition of the W Gram failed\n");
            abort();
        }
        const double tmax = w[nw - 1];
        int first = 0;
        while (first < nw && !(w[first] > tau * tmax)) first++;
        k = nw - first;
        // column c of C: D Z(:, first + c) / sqrt(theta)
        for (int c = 0; c < k; c++) {
            const double sc = 1.0 / std::sqrt(w[first + c]);
            for (int i = 0; i < nw; i++) C[(size_t)c * nw + i] = d[i] * Z[(size_t)(first + c) * nw + i] * sc;
        }
        if (k < nw && para->orth_para->print_orth_zero == 1)
            printf("In CholQR orthonormalization, %d dependent directions dropped; end = %d\n", nw - k, start + k);
    }
    if (k > 0) {
        set2(ms, start, 0);
        set2(me, *end, k);
        ops->MultiVecLinearComb(V, ws->V_tmp, ms, me, C.data(), nw, nullptr, 0, ops);
        set2(ms, 0, start);
        set2(me, k, start + k);
        ops->MultiVecCopy(ws->V_tmp, V, ms, me, ops);
    }
    *end = start + k;
    return chol_ok ? rmin : 0.0;
}

void cholqr_orth(void **V, int start, int *end, void *B, GCGE_OPS *ops, GCGE_PARA *para, GCGE_WORKSPACE *ws)
{
    GCGE_INT ms[2], me[2];
    const double check_tol = 1e3 * para->basis_eps;
    std::vector<double> G, wn2;
    // W from the FP32 defect correction is the correction E alone (compute_w),
    // mostly orthogonal to V1 already: when the first projection removes at
    // most half of every column's squared B-norm, the second is skipped
    // ("twice is enough": one pass then loses no more than a factor 2 of
    // orthogonality to cancellation, and the factor pass and the check below
    // still apply)
    const bool correction = para->bcg_precision == 1 && para->if_use_bcg == 1 && !ops->LinearSolver;
    for (int round = 0; round < 3 && *end > start; round++) {
        if (start > 0) {
            bool again = true;
            for (int pass = 0; pass < 2 && *end > start && again; pass++) {
                set2(ms, 0, start);
                set2(me, start, *end);
                const int len2 = *end - start;
                const bool test = correction && pass == 0;
                if (test) wn2.assign(len2, 0.0);
                sub_orth(V, ms, me, B, ws->V_tmp, ws->subspace_dtmp, ops, test ? wn2.data() : nullptr);
                if (test) {
                    again = false;
                    for (int j = 0; j < len2 && !again; j++) {
                        double g2 = 0.0;
                        for (int i = 0; i < start; i++) {
                            const double g = ws->subspace_dtmp[(size_t)j * start + i];
                            g2 += g * g;
                        }
                        if (!(g2 <= 0.5 * wn2[j])) again = true;
                    }
                }
            }
        }
        double rmin = 1.0;
        for (int pass = 0; pass < 2 && *end > start; pass++)
            rmin = std::min(rmin, pass == 0 ? orth_factor_pass(V, start, end, B, ops, para, ws)
                                            : std::max(orth_factor_pass(V, start, end, B, ops, para, ws), 0.0));
        if (start == 0 || *end <= start) return;
        // W was orthogonal to V1 to working accuracy after the projections; a
        // Cholesky factor with R_jj >= 1e-2 amplifies that overlap at most
        // 100 times, so the check is needed only after a worse factorization
        if (rmin >= 1e-2) return;
        // check V1^T B W
        const int nw = *end - start;
        void **T = V;
        int tc = start;
        if (B) {
            set2(ms, start, 0);
            set2(me, *end, nw);
            ops->MatDotMultiVec(B, V, ws->V_tmp, ms, me, ops);
            T = ws->V_tmp;
            tc = 0;
        }
        G.assign((size_t)start * nw, 0.0);
        set2(ms, 0, tc);
        set2(me, start, tc + nw);
        ops->MultiVecInnerProd(V, T, G.data(), (char *)"ns", ms, me, start, ops);
        double gmax = 0.0;
        for (double g : G) gmax = std::max(gmax, std::fabs(g));
        if (gmax <= check_tol) return;
    }
}

// GCGE_MultiOrthonormalizationInSubspace, gcge_orthogonal.c:2130-2204
// (host FP64; the second product updates ldV rows as the reference's does).
void multi_orth_insub(double *Vd, int ldV, int nrows, int start, int *end, GCGE_ORTH_PARA *op, GCGE_WORKSPACE *ws,
                      GCGE_OPS *ops)
{
    const int length = *end - start;
    int length_1 = length / 2;
    const int length_2 = length - length_1;
    int mid = start + length_1;
    if (length == 1) {
        const double v = ops->ArrayNorm(Vd + (size_t)start * ldV, nrows);
        if (v > op->orth_zero_tol) ops->ArrayScale(1.0 / v, Vd + (size_t)start * ldV, nrows);
        else *end = start;
        return;
    }
    multi_orth_insub(Vd, ldV, nrows, start, &mid, op, ws, ops);
    if (mid + length_2 < *end) {
        const int new_end_2 = mid + length_2, n_zero = *end - new_end_2;
        length_1 -= n_zero;
        std::memcpy(Vd + (size_t)mid * ldV, Vd + (size_t)new_end_2 * ldV, (size_t)n_zero * ldV * sizeof(double));
        *end = new_end_2;
    }
    if (length_1 > 0) {
        for (int rc = 0; rc < op->max_reorth_time; rc++) {
            double alpha = 1.0, beta = 0.0;
            int l1 = length_1, l2 = length_2, nr = nrows, ld = ldV;
            ops->DenseMatDotDenseMat((char *)"T", (char *)"N", &l1, &l2, &nr, &alpha, Vd + (size_t)start * ldV, &ld,
                                     Vd + (size_t)mid * ldV, &ld, &beta, ws->subspace_dtmp, &l1);
            alpha = -1.0;
            beta = 1.0;
            ops->DenseMatDotDenseMat((char *)"N", (char *)"N", &ld, &l2, &l1, &alpha, Vd + (size_t)start * ldV, &ld,
                                     ws->subspace_dtmp, &l1, &beta, Vd + (size_t)mid * ldV, &ld);
        }
    }
    multi_orth_insub(Vd, ldV, nrows, mid, end, op, ws, ops);
}

// GCGE_StableMultiOrthonormalizationInSubspace, gcge_orthogonal.c:2215-2247.
void stable_multi_orth_insub(double *Vd, int ldV, int nrows, int start, int *end, GCGE_ORTH_PARA *op,
                             GCGE_WORKSPACE *ws, GCGE_OPS *ops)
{
    if (start != 0) {
        for (int rc = 0; rc < op->max_reorth_time; rc++) {
            int len = *end - start, st = start, nr = nrows, ld = ldV;
            double alpha = 1.0, beta = 0.0;
            ops->DenseMatDotDenseMat((char *)"T", (char *)"N", &st, &len, &nr, &alpha, Vd, &ld,
                                     Vd + (size_t)start * ldV, &ld, &beta, ws->subspace_dtmp, &st);
            alpha = -1.0;
            beta = 1.0;
            ops->DenseMatDotDenseMat((char *)"N", (char *)"N", &ld, &len, &st, &alpha, Vd, &ld, ws->subspace_dtmp, &st,
                                     &beta, Vd + (size_t)start * ldV, &ld);
        }
    }
    multi_orth_insub(Vd, ldV, nrows, start, end, op, ws, ops);
}

// ---- the other in-subspace orthonormalizations of ComputeP's coefficients
// (p_orth_type "gs", "bgs", "scbgs"; gcge_xpw.c:363-385), host FP64 through
// the dense slots, as the reference issues them.

// GCGE_OrthonormalizationInSubspace, gcge_orthogonal.c:581-645: modified
// Gram-Schmidt with reorthogonalization; zero columns replaced by the last.
void gs_insub(double *Vd, int ldV, int nrows, int start, int *end, GCGE_ORTH_PARA *op, GCGE_OPS *ops)
{
    for (int current = start; current < *end; ++current) {
        double *vc = Vd + (size_t)current * ldV;
        double vout = ops->ArrayNorm(vc, nrows);
        if (current != 0) {
            int reorth_count = 0;
            double vin, ratio;
            do {
                vin = vout;
                for (int idx = 0; idx < current; idx++) {
                    const double ip = ops->ArrayDotArray(Vd + (size_t)idx * ldV, vc, nrows);
                    ops->ArrayAXPBY(-ip, Vd + (size_t)idx * ldV, 1.0, vc, nrows);
                }
                vout = ops->ArrayNorm(vc, nrows);
                ratio = vout / vin;
                reorth_count++;
            } while (ratio < op->reorth_tol && reorth_count < op->max_reorth_time && vout > op->orth_zero_tol);
        }
        if (vout > op->orth_zero_tol) {
            ops->ArrayScale(1.0 / vout, vc, nrows);
        } else {
            if (current != *end - 1) ops->ArrayCopy(Vd + (size_t)(*end - 1) * ldV, vc, nrows);
            (*end)--;
            current--;
        }
    }
}

// the self-orthonormalization sweep of the two block variants: each column
// normalized, its components removed from the later ones (skipping
// components below 10 orth_zero_tol), zero columns replaced by the last
// (gcge_orthogonal.c:1538-1575, :1674-1714)
void self_insub(double *Vd, int ldV, int nrows, int start, int *end, GCGE_ORTH_PARA *op, double *d, GCGE_OPS *ops)
{
    const double tol = op->orth_zero_tol;
    for (int current = start; current < *end; ++current) {
        int len = *end - current, one = 1, nr = nrows, ld = ldV;
        double alpha = 1.0, beta = 0.0;
        double *vc = Vd + (size_t)current * ldV;
        ops->DenseMatDotDenseMat((char *)"T", (char *)"N", &len, &one, &nr, &alpha, vc, &ld, vc, &ld, &beta, d, &len);
        const double norm = std::sqrt(d[0]);
        if (norm > tol) {
            ops->ArrayScale(1.0 / norm, vc, nrows);
            for (int j = current + 1; j < *end; j++)
                if (std::fabs(d[j - current] / norm) > 10.0 * tol)
                    ops->ArrayAXPBY(-d[j - current] / norm, vc, 1.0, Vd + (size_t)j * ldV, nrows);
        } else {
            if (current != *end - 1) ops->ArrayCopy(Vd + (size_t)(*end - 1) * ldV, vc, nrows);
            (*end)--;
            current--;
        }
    }
}

// GCGE_BOrthonormalizationInSubspace, gcge_orthogonal.c:1501-1578: two
// passes of column-by-column projection against V[0:start] (components
// below 10 orth_zero_tol skipped) followed by the self sweep.
void bgs_insub(double *Vd, int ldV, int nrows, int start, int *end, GCGE_ORTH_PARA *op, double *d, GCGE_OPS *ops)
{
    const double tol = op->orth_zero_tol;
    for (int pass = 0; pass < 2; pass++) {
        int len = *end - start, one = 1, nr = nrows, ld = ldV;
        for (int current = 0; current < start; ++current) {
            double alpha = 1.0, beta = 0.0;
            ops->DenseMatDotDenseMat((char *)"T", (char *)"N", &len, &one, &nr, &alpha, Vd + (size_t)start * ldV, &ld,
                                     Vd + (size_t)current * ldV, &ld, &beta, d, &len);
            for (int j = start; j < *end; j++)
                if (std::fabs(d[j - start]) > 10.0 * tol)
                    ops->ArrayAXPBY(-d[j - start], Vd + (size_t)current * ldV, 1.0, Vd + (size_t)j * ldV, nrows);
        }
        self_insub(Vd, ldV, nrows, start, end, op, d, ops);
    }
}

// GCGE_SCBOrthonormalizationInSubspace, gcge_orthogonal.c:1619-1718:
// max_reorth_time block projections V2 -= V1 (V1^T V2) (the update covers ldV
// rows, as in the reference), then one self sweep.
void scbgs_insub(double *Vd, int ldV, int nrows, int start, int *end, GCGE_ORTH_PARA *op, double *d, GCGE_OPS *ops)
{
    if (start != 0) {
        for (int rc = 0; rc < op->max_reorth_time; rc++) {
            int len = *end - start, st = start, nr = nrows, ld = ldV;
            double alpha = 1.0, beta = 0.0;
            ops->DenseMatDotDenseMat((char *)"T", (char *)"N", &st, &len, &nr, &alpha, Vd, &ld,
                                     Vd + (size_t)start * ldV, &ld, &beta, d, &st);
            alpha = -1.0;
            beta = 1.0;
            ops->DenseMatDotDenseMat((char *)"N", (char *)"N", &ld, &len, &st, &alpha, Vd, &ld, d, &st, &beta,
                                     Vd + (size_t)start * ldV, &ld);
        }
    }
    self_insub(Vd, ldV, nrows, start, end, op, d, ops);
}

// ------------------------------------------------------------------------
// CG solvers (src/gcge_cg.c)
// ------------------------------------------------------------------------

// GCGE_CG, gcge_cg.c:26-113 (only with if_use_bcg != 1), per-vector ops.
void cg_single(void *A, void *b, void *x, GCGE_OPS *ops, GCGE_PARA *para, void **V_tmp1, void **V_tmp2)
{
    void *r, *p, *w;
    int niter, max_it = para->cg_max_it, type = para->cg_type;
    double rate = para->cg_rate, alpha, beta, rho1 = 0.0, rho2, error, last_error, pTw, wTw;
    ops->GetVecFromMultiVec(V_tmp1, 1, &r);
    ops->GetVecFromMultiVec(V_tmp2, 0, &p);
    ops->GetVecFromMultiVec(V_tmp2, 1, &w);
    ops->MatDotVec(A, x, r);
    ops->VecAxpby(1.0, b, -1.0, r);
    ops->VecInnerProd(r, r, &rho2);
    error = std::sqrt(rho2);
    niter = 1;
    last_error = error;
    while ((last_error / error >= rate) && (niter < max_it)) {
        if (niter == 1) {
            ops->VecAxpby(1.0, r, 0.0, p);
        } else {
            beta = rho2 / rho1;
            ops->VecAxpby(1.0, r, beta, p);
        }
        ops->MatDotVec(A, p, w);
        if (type == 1) {
            ops->VecInnerProd(p, w, &pTw);
            alpha = rho2 / pTw;
            ops->VecAxpby(alpha, p, 1.0, x);
            ops->VecAxpby(-alpha, w, 1.0, r);
            rho1 = rho2;
            ops->VecInnerProd(r, r, &rho2);
        } else {
            ops->VecInnerProd(p, w, &pTw);
            ops->VecInnerProd(w, w, &wTw);
            alpha = rho2 / pTw;
            ops->VecAxpby(alpha, p, 1.0, x);
            ops->VecAxpby(-alpha, w, 1.0, r);
            rho1 = rho2;
            rho2 = rho1 - 2.0 * alpha * pTw + alpha * alpha * wTw;
        }
        last_error = std::sqrt(rho2);
        niter++;
    }
}

// GCGE_BCG, gcge_cg.c:137-302, one block operation per step for all active
// columns.  The reference rebuilds its active list with unlock[idx] = idx
// (:289-293) starting from the identity, so after each sweep the active
// columns are 0..num_unlock-1 -- the first num_unlock columns, not the
// unconverged ones.  That rule is kept.
// x_is_zero: the initial guess is exactly zero, so the first residual
// r = b - A*0 is b itself and its matrix product is skipped.
// With para->bcg_precond the CG is preconditioned by the diagonal D of A:
// rho = r.(D^-1 r) drives alpha and beta and p starts from D^-1 r (formed by
// the search step, never stored), while the stopping rule keeps the
// reference's ||r_k|| / ||r_0|| (e2 = r.r).
void bcg(void *A, void **RHS, void **V, int x_start, int x_length, GCGE_OPS *ops, GCGE_PARA *para, void **CG_P,
         void **CG_W, double *dtmp, int *itmp, bool x_is_zero = false)
{
    if (x_length <= 0) return;
    const bool pre = para->bcg_precond != 0;
    std::vector<double> e2(pre ? x_length : 0);
    double *rho1 = dtmp, *rho2 = rho1 + x_length, *error = rho2 + x_length;
    double *last_error = error + x_length, *ptw = last_error + x_length;
    int *unlock = itmp;
    const int max_it = para->cg_max_it;
    const double rate = para->cg_rate;
    GCGE_INT ms[2], me[2];
    std::vector<double> b(x_length), alpha(x_length), rr(x_length);

    if (!x_is_zero) {
        set2(ms, x_start, 0);
        set2(me, x_start + x_length, x_length);
        ops->MatDotMultiVec(A, V, CG_W, ms, me, ops);
        set2(ms, 0, 0);
        set2(me, x_length, x_length);
        ops->MultiVecAxpby(-1.0, CG_W, 1.0, RHS, ms, me, ops);
    }
    set2(ms, 0, 0);
    set2(me, x_length, x_length);
    if (pre) {
        // alpha = 0 leaves r as it is (the a == 0 branch of y = a x + b y)
        std::vector<double> zero(x_length, 0.0);
        ops->BlockPCGResidualUpdate(A, CG_W, RHS, x_length, zero.data(), rho2, e2.data(), ops);
    } else {
        ops->MultiVecInnerProdDiag(RHS, RHS, rho2, ms, me, ops);
    }
    for (int idx = 0; idx < x_length; idx++) unlock[idx] = idx;
    int num_unlock = x_length;
    for (int idx = 0; idx < x_length; idx++) error[idx] = std::sqrt(pre ? e2[idx] : rho2[idx]);
    // The solution update x += alpha p of a step is deferred into the search
    // direction update of the next step (p is read there before it is
    // overwritten); every element sees the reference's roundings in the
    // reference's order (gcge_cg.c:197-286).
    int niter = 1, na_prev = 0;
    while (num_unlock > 0 && niter < max_it) {
        const int na = num_unlock;  // active columns are unlock[0..na) = 0..na-1
        set2(ms, 0, 0);
        set2(me, na, na);
        if (niter == 1) {
            if (pre) {
                // p = D^-1 r (beta 0: the copy branch), nothing to add to x yet
                std::fill(b.begin(), b.end(), 0.0);
                ops->BlockPCGSearchUpdate(A, V, x_start, CG_P, RHS, 0, na, alpha.data(), b.data(), ops);
            } else {
                ops->MultiVecAxpby(1.0, RHS, 0.0, CG_P, ms, me, ops);
            }
        } else {
            for (int id = 0; id < na; id++) b[id] = rho2[id] / rho1[id];
            if (pre) ops->BlockPCGSearchUpdate(A, V, x_start, CG_P, RHS, na_prev, na, alpha.data(), b.data(), ops);
            else ops->BlockCGSearchUpdate(V, x_start, CG_P, RHS, na_prev, na, alpha.data(), b.data(), ops);
        }
        ops->MatDotMultiVecInnerProdDiag(A, CG_P, CG_W, ms, me, ptw, ops);
        for (int id = 0; id < na; id++) {
            alpha[id] = rho2[id] / ptw[id];
            rho1[id] = rho2[id];
        }
        if (pre) {
            std::vector<double> rz(na);
            ops->BlockPCGResidualUpdate(A, CG_W, RHS, na, alpha.data(), rz.data(), rr.data(), ops);
            for (int id = 0; id < na; id++) {
                rho2[id] = rz[id];
                e2[id] = rr[id];
            }
        } else {
            ops->BlockCGResidualUpdate(CG_W, RHS, na, alpha.data(), rr.data(), ops);
            for (int id = 0; id < na; id++) rho2[id] = rr[id];
        }
        na_prev = na;
        for (int idx = 0; idx < x_length; idx++)
            last_error[idx] = std::sqrt(std::fabs(pre ? e2[idx] : rho2[idx]));
        num_unlock = 0;
        for (int idx = 0; idx < x_length; idx++) {
            if (last_error[idx] / error[idx] >= rate) {
                unlock[idx] = idx;
                num_unlock++;
            }
        }
        niter++;
    }
    if (na_prev > 0) ops->BlockCGSearchUpdate(V, x_start, CG_P, RHS, na_prev, 0, alpha.data(), b.data(), ops);
    para->bcg_calls++;
    para->bcg_iters_total += niter - 1;
}

// ------------------------------------------------------------------------
// X, W, P (src/gcge_xpw.c)
// ------------------------------------------------------------------------

// GCGE_ComputeX, gcge_xpw.c:47-76.
void compute_x(void **V, double *coef, void **X, int size_X, int size_V, GCGE_OPS *ops, GCGE_WORKSPACE *ws)
{
    GCGE_INT s[2], e[2];
    const int sx = ws->num_soft_locked_in_last_iter;
    set2(s, sx, sx);
    set2(e, size_V, size_X);
    ops->MultiVecLinearComb(V, X, s, e, coef + (size_t)sx * size_V + sx, size_V, nullptr, 0, ops);
}

// FP32 work columns of the inner solves (bcg_precision == 1), created once
// per workspace from A through the backend's precision-aware constructor.
bool bcg32_ready(void *A, int cols, GCGE_OPS *ops, GCGE_WORKSPACE *ws)
{
    if (ws->bcg32[0] && ws->bcg32_cols >= cols) return true;
    if (!ops->MultiVecCreateByMatPrecision) {
        fprintf(stderr, "GCGE: bcg_precision 1 needs MultiVecCreateByMatPrecision\n");
        return false;
    }
    for (int i = 0; i < 4; i++)
        if (ws->bcg32[i]) ops->MultiVecDestroy(&ws->bcg32[i], ws->bcg32_cols, ops);
    const int n = ws->CG_p_size > cols ? ws->CG_p_size : cols;
    for (int i = 0; i < 4; i++) ops->MultiVecCreateByMatPrecision(&ws->bcg32[i], n, A, 1, ops);
    ws->bcg32_cols = n;
    return ws->bcg32[0] && ws->bcg32[1] && ws->bcg32[2] && ws->bcg32[3];
}

// the preconditioned block CG needs its two slots
bool pcg_ready(GCGE_OPS *ops, GCGE_PARA *para)
{
    if (!para->bcg_precond || (ops->BlockPCGResidualUpdate && ops->BlockPCGSearchUpdate)) return true;
    fprintf(stderr, "GCGE: bcg_precond 1 needs the BlockPCGResidualUpdate and BlockPCGSearchUpdate slots\n");
    return false;
}

// GCGE_ComputeW, gcge_xpw.c:118-217: W_j starts from x_j (|lambda| <= 1) or
// x_j / lambda; the right-hand side is lambda B x_j or B x_j.
void compute_w(void *A, void *B, void **V, double *eval, GCGE_OPS *ops, GCGE_PARA *para, GCGE_WORKSPACE *ws)
{
    const int w_start = ws->dim_xp, w_length = ws->unconv_bs;
    if (w_length <= 0) return;
    GCGE_INT ms[2], me[2];
    std::vector<double> ra(w_length, 0.0), rb(w_length), wa(w_length, 0.0), wb(w_length);
    // W_j <- x_{unlock[j]}
    ops->MultiVecGather(V, ws->unlock, w_length, V, w_start, ops);
    set2(ms, w_start, 0);
    set2(me, w_start + w_length, w_length);
    if (!B) ops->MultiVecAxpby(1.0, V, 0.0, ws->V_tmp, ms, me, ops);
    else ops->MatDotMultiVec(B, V, ws->V_tmp, ms, me, ops);
    bool scale_rhs = false, scale_w = false;
    for (int j = 0; j < w_length; j++) {
        const double lam = eval[ws->unlock[j]];
        if (std::fabs(lam) <= 1.0) {
            rb[j] = lam;  // rhs = lambda * rhs
            wb[j] = 1.0;
            scale_rhs = true;
        } else {
            rb[j] = 1.0;
            wb[j] = 1.0 / lam;  // w = x / lambda
            scale_w = true;
        }
    }
    set2(ms, 0, 0);
    set2(me, w_length, w_length);
    if (scale_rhs) ops->MultiVecAxpbyDiag(ra.data(), ws->V_tmp, rb.data(), ws->V_tmp, ms, me, ops);
    set2(ms, w_start, w_start);
    set2(me, w_start + w_length, w_start + w_length);
    if (scale_w) ops->MultiVecAxpbyDiag(wa.data(), V, wb.data(), V, ms, me, ops);

    if (ops->LinearSolver) {
        for (int idx = 0; idx < w_length; idx++) {
            void *rhs, *wc;
            ops->GetVecFromMultiVec(ws->V_tmp, idx, &rhs);
            ops->GetVecFromMultiVec(V, w_start + idx, &wc);
            ops->LinearSolver(A, rhs, wc, ops);
            ops->RestoreVecForMultiVec(V, w_start + idx, &wc);
            ops->RestoreVecForMultiVec(ws->V_tmp, idx, &rhs);
        }
    } else if (para->if_use_bcg == 1 && para->bcg_precision == 1) {
        // Defect correction: r0 = rhs - A W0 in FP64 (the first step of the
        // reference's BCG), CG for the correction A E = r0 from E = 0 on FP32
        // copies (the same CG iterates in exact arithmetic, and the same
        // ||r_k|| / ||r_0|| stopping rule); W0 + E and E span the same space
        // with X.  FP32 rounding is relative to the correction, which shrinks
        // as X converges.
        if (!bcg32_ready(A, w_length, ops, ws) || !pcg_ready(ops, para)) return;
        set2(ms, w_start, 0);
        set2(me, w_start + w_length, w_length);
        ops->MatDotMultiVec(A, V, ws->evec, ms, me, ops);
        set2(ms, 0, 0);
        set2(me, w_length, w_length);
        ops->MultiVecAxpby(-1.0, ws->evec, 1.0, ws->V_tmp, ms, me, ops);
        ops->MultiVecCopy(ws->V_tmp, ws->bcg32[0], ms, me, ops);
        ops->MultiVecAxpby(0.0, ws->bcg32[1], 0.0, ws->bcg32[1], ms, me, ops);
        bcg(A, ws->bcg32[0], ws->bcg32[1], 0, w_length, ops, para, ws->bcg32[2], ws->bcg32[3], ws->subspace_dtmp,
            ws->subspace_itmp, true);
        // W = W0 + E spans the same space with X as E alone (W0 is a column of X
        // scaled), so E, widened to FP64, is the W block: it is mostly
        // orthogonal to X already, which the W orthonormalization uses
        set2(ms, 0, w_start);
        set2(me, w_length, w_start + w_length);
        ops->MultiVecCopy(ws->bcg32[1], V, ms, me, ops);
    } else if (para->if_use_bcg == 1) {
        if (!pcg_ready(ops, para)) return;
        bcg(A, ws->V_tmp, V, w_start, w_length, ops, para, ws->CG_p, ws->evec, ws->subspace_dtmp, ws->subspace_itmp);
    } else {
        for (int idx = 0; idx < w_length; idx++) {
            void *rhs, *wc;
            ops->GetVecFromMultiVec(ws->V_tmp, idx, &rhs);
            ops->GetVecFromMultiVec(V, w_start + idx, &wc);
            cg_single(A, rhs, wc, ops, para, ws->V_tmp, ws->evec);
        }
    }
}

// GCGE_ComputeP, gcge_xpw.c:286-425 ("multi" subspace orthogonalization).
void compute_p(double *coef, void **V, GCGE_OPS *ops, GCGE_PARA *para, GCGE_WORKSPACE *ws)
{
    GCGE_TIMER(p_orth_time, para);
    const int ldc = ws->dim_xpw;
    int p_start = ws->dim_x, p_ncols = ws->unconv_bs;
    const int xx_nrows = ws->last_dim_x, nev = para->nev;
    GCGE_INT ms[2], me[2];
    double *px = coef + (size_t)p_start * ldc, *pp = px + xx_nrows, *xp = coef + xx_nrows;
    std::memset(px, 0, (size_t)p_ncols * ldc * sizeof(double));
    for (int cp = 0; cp < p_ncols; cp++) {
        const int cx = ws->unlock[cp];
        std::memcpy(pp + (size_t)cp * ldc, xp + (size_t)cx * ldc, (size_t)(ldc - xx_nrows) * sizeof(double));
    }
    const int orth_start = ws->num_soft_locked_in_last_iter;
    const int orth_length = ldc - orth_start;
    p_start -= orth_start;
    int p_end = p_start + p_ncols;
    double *sub = coef + (size_t)orth_start * ldc + orth_start;
    const char *pt = para->p_orth_type;
    if (std::strcmp(pt, "scbgs") == 0)
        scbgs_insub(sub, ldc, orth_length, p_start, &p_end, para->orth_para, ws->subspace_dtmp, ops);
    else if (std::strcmp(pt, "bgs") == 0)
        bgs_insub(sub, ldc, orth_length, p_start, &p_end, para->orth_para, ws->subspace_dtmp, ops);
    else if (std::strcmp(pt, "multi") == 0)
        stable_multi_orth_insub(sub, ldc, orth_length, p_start, &p_end, para->orth_para, ws, ops);
    else  // "gs", and any other name, as in the reference (gcge_xpw.c:381-385)
        gs_insub(sub, ldc, orth_length, p_start, &p_end, para->orth_para, ops);
    p_end += orth_start;
    const int p_end_in_V_tmp = p_end - nev;
    {
        GCGE_TIMER(p_axpy_time, para);
        if (p_end_in_V_tmp > 0) {
            set2(ms, orth_start, 0);
            set2(me, ldc, p_end_in_V_tmp);
            ops->MultiVecLinearComb(V, ws->V_tmp, ms, me, coef + (size_t)nev * ldc + orth_start, ldc, nullptr, -1,
                                    ops);
            set2(ms, 0, nev);
            set2(me, p_end_in_V_tmp, p_end);
            ops->MultiVecCopy(ws->V_tmp, V, ms, me, ops);
        }
    }
    ws->dim_xp = p_end;
}

// ------------------------------------------------------------------------
// Rayleigh-Ritz (src/gcge_rayleighritz.c)
// ------------------------------------------------------------------------

// GCGE_ComputeSubspaceMatrixVTAW, gcge_rayleighritz.c:35-81.
void rr_vtaw(void **V, void *A, int start_W, int end_V, double *mat, GCGE_OPS *ops, void **work)
{
    const int w_length = end_V - start_W, xp_length = end_V - w_length;
    GCGE_INT ms[2], me[2];
    set2(ms, start_W, 0);
    set2(me, end_V, w_length);
    ops->MatDotMultiVec(A, V, work, ms, me, ops);
    set2(ms, 0, 0);
    set2(me, xp_length, w_length);
    ops->MultiVecInnerProd(V, work, mat, (char *)"nonsym", ms, me, end_V, ops);
    set2(ms, start_W, 0);
    set2(me, end_V, w_length);
    ops->MultiVecInnerProd(V, work, mat + xp_length, (char *)"sym", ms, me, end_V, ops);
}

// GCGE_ComputeSubspaceMatrix, gcge_rayleighritz.c:113-200 (live branch).
void rr_matrix(void *A, void **V, GCGE_OPS *ops, GCGE_PARA *para, GCGE_WORKSPACE *ws)
{
    GCGE_TIMER(rr_mat_time, para);
    int dim_xp = ws->dim_xp, ldm = ws->dim_xpw;
    if (dim_xp != 0) {
        int lde = ws->last_dim_xpw, dim_x = ws->dim_x, dim_p = dim_xp - dim_x;
        double alpha = 1.0, beta = 0.0;
        double *pp = ws->subspace_evec + (size_t)dim_x * lde;
        double *AApp = ws->subspace_dtmp;
        double *PTAP = ws->subspace_matrix + (size_t)dim_x * ldm + dim_x;
        ops->DenseSymMatDotDenseMat((char *)"L", (char *)"U", &lde, &dim_p, &alpha, ws->subspace_matrix, &lde, pp,
                                    &lde, &beta, AApp, &lde);
        std::memset(ws->subspace_matrix, 0, (size_t)ldm * ldm * sizeof(double));
        ops->DenseMatDotDenseMat((char *)"T", (char *)"N", &dim_p, &dim_p, &lde, &alpha, pp, &lde,
                                 ws->subspace_dtmp, &lde, &beta, PTAP, &ldm);
        for (int i = 0; i < dim_x; i++) ws->subspace_matrix[(size_t)i * ldm + i] = ws->eval[i];
    }
    rr_vtaw(V, A, dim_xp, ldm, ws->subspace_matrix + (size_t)dim_xp * ldm, ops, ws->evec);
}

// GCGE_ComputeSubspaceEigenpairs, gcge_rayleighritz.c:364-590 (one process).
void rr_eigenpairs(double *subspace_matrix, double *eval, double *subspace_evec, GCGE_OPS *ops, GCGE_PARA *para,
                   GCGE_WORKSPACE *ws)
{
    GCGE_TIMER(rr_eigen_time, para);
    int ldm = ws->dim_xpw, info = 0;
    double *temp = ws->subspace_dtmp, *dwork = temp + (size_t)ldm * ldm;
    double vl = 0.0, vu = 0.0, abstol = 0.0;
    int rr_eigen_start = ws->unlock[0];
    if (para->opt_rr_eig_partly == 0) rr_eigen_start = 0;
    int il = rr_eigen_start;
    for (int i = rr_eigen_start - 1; i > -1; i--)
        if (std::fabs(eval[i] - eval[rr_eigen_start]) / eval[rr_eigen_start] < para->multi_tol_for_lock) il -= 1;
    rr_eigen_start = il;
    il += 1;
    ws->num_soft_locked_in_last_iter = rr_eigen_start;
    int iu = (ws->max_dim_x < ldm) ? ws->max_dim_x : ldm;
    int m = iu - il + 1, lwork = 8 * ldm, liwork = 5 * ldm;
    int *isuppz = ws->subspace_itmp + 10 * ldm, *ifail = ws->subspace_itmp + 5 * ldm;
    for (int i = 0; i < ws->dim_xpw - ws->dim_xp; i++)
        std::memset(subspace_matrix + (size_t)(ws->dim_xp + i) * ldm, 0, (size_t)rr_eigen_start * sizeof(double));
    std::memcpy(temp, subspace_matrix, (size_t)ldm * ldm * sizeof(double));
    for (size_t t = 0; t < (size_t)ldm * ldm; t++)
        if (!std::isfinite(temp[t])) {
            // NaN or Inf from the large-scale products: no LAPACK call on it
            gh::latch(192, "Rayleigh-Ritz: subspace matrix entry (%d, %d) is %g", (int)(t % ldm), (int)(t / ldm),
                      temp[t]);
            for (int i = rr_eigen_start; i < ldm; i++) eval[i] = temp[t];
            return;
        }
    int nrows = ldm - rr_eigen_start, ncols = ldm;
    il -= rr_eigen_start;
    iu -= rr_eigen_start;
    if (para->rr_eig_full)
        info = gcge_sym_eig_range(nrows, temp + (size_t)rr_eigen_start * ncols + rr_eigen_start, ncols, il, iu,
                                  eval + rr_eigen_start, subspace_evec + (size_t)rr_eigen_start * ldm + rr_eigen_start,
                                  ldm);
    else
        ops->DenseMatEigenSolver((char *)"V", (char *)"I", (char *)"U", &nrows,
                                 temp + (size_t)rr_eigen_start * ncols + rr_eigen_start, &ncols, &vl, &vu, &il, &iu,
                                 &abstol, &m, eval + rr_eigen_start,
                                 subspace_evec + (size_t)rr_eigen_start * ldm + rr_eigen_start, &ldm, isuppz, dwork,
                                 &lwork, ws->subspace_itmp, &liwork, ifail, &info);
    if (info != 0) fprintf(stderr, "GCGE: dense eigensolver returned info = %d\n", info);
}

}  // namespace

// ------------------------------------------------------------------------
// Convergence and multiplicity (src/gcge_eigsol.c)
// ------------------------------------------------------------------------

// GCGE_CheckConvergence, gcge_eigsol.c:559-686.  The residuals of the pairs
// the reference visits are computed in batches: pairs up to position
// max_conv_idx + 9 of unlock[] are always visited (the early exit needs
// j > max_conv_idx + 8), so each batch extends exactly to that horizon and no
// residual is computed that the reference would not compute.  A residual
// that is NaN or infinite leaves its pair unconverged and latches error 190
// (the reference counts a NaN as converged); GCGE_Solve then stops.
extern "C" void GCGE_CheckConvergence(void *A, void *B, GCGE_DOUBLE *eval, void **evec, GCGE_OPS *ops,
                                      GCGE_PARA *para, GCGE_WORKSPACE *ws)
{
    int *unlock = ws->unlock, unconv_bs = 0, max_conv_idx = 0, max_ind = 0, min_ind = 0;
    const int num_unlock = para->num_unlock;
    double max_res = 0.0, min_res = 0.0, sum_res = 0.0;
    const char *ct = para->conv_type;
    const bool ns = std::strcmp(ct, "NS") == 0;
    std::vector<int> idx;
    std::vector<double> lam, rn, xn, bn;
    int j = 0;
    bool done = false;
    while (j < num_unlock && !done) {
        const int horizon = std::min(num_unlock, max_conv_idx + 10);
        const int k = horizon - j;
        idx.assign(unlock + j, unlock + horizon);
        lam.resize(k);
        rn.resize(k);
        xn.resize(k);
        bn.resize(k);
        for (int t = 0; t < k; t++) lam[t] = eval[idx[t]];
        ops->MultiVecResidualNorms(A, B, evec, idx.data(), k, lam.data(), rn.data(), xn.data(), ns ? bn.data() : nullptr,
                                   ops);
        for (int t = 0; t < k && !done; t++, j++) {
            const int i = idx[t];
            const double res_norm = rn[t], evec_norm = xn[t];
            double residual;
            if (std::strcmp(ct, "R") == 0)
                residual = res_norm / evec_norm / (1.0 > std::fabs(eval[i]) ? 1.0 : std::fabs(eval[i]));
            else if (std::strcmp(ct, "O") == 0)
                residual = res_norm / evec_norm / (para->conv_omega_norm + std::fabs(eval[i]));
            else if (ns)
                residual = res_norm / (std::fabs(eval[i]) * bn[t]);
            else
                residual = res_norm / evec_norm;
            para->res[i] = residual;
            sum_res += residual;
            const bool finite = std::isfinite(residual);
            if (!finite)
                gh::latch(190, "GCGE_CheckConvergence: residual of eigenpair %d is %g (eigenvalue %g)", i, residual,
                          eval[i]);
            if (j == 0) {
                max_res = residual;
                min_res = residual;
            } else {
                if (!(residual <= max_res) && !std::isnan(max_res)) {  // a NaN residual becomes the maximum
                    max_res = residual;
                    max_ind = i;
                }
                if (residual < min_res) {
                    min_res = residual;
                    min_ind = i;
                }
            }
            // the reference tests residual > ev_tol (gcge_eigsol.c:630), which
            // passes a NaN residual as converged; here it stays unconverged
            if (!(residual <= para->ev_tol)) {
                unlock[unconv_bs] = i;
                unconv_bs += 1;
                if (j > max_conv_idx + 8) {
                    for (int kk = j + 1; kk < num_unlock; kk++) {
                        unlock[unconv_bs] = unlock[kk];
                        unconv_bs += 1;
                    }
                    done = true;
                }
            } else {
                max_conv_idx = j;
            }
        }
    }
    if (!para->lock_converged && unconv_bs > 0) {
        // every pair stays in the W and P blocks until all have converged
        for (int i = 0; i < para->nev; i++) unlock[i] = i;
        unconv_bs = para->nev;
    }
    para->num_unlock = unconv_bs;
    ws->unconv_bs = (unconv_bs < para->block_size) ? unconv_bs : para->block_size;
    para->max_res = max_res;
    para->max_ind = max_ind;
    para->min_res = min_res;
    para->min_ind = min_ind;
    para->sum_res = sum_res;
    para->num_iter += 1;
}

// GCGE_CheckEvalMultiplicity, gcge_eigsol.c:696-718 (fixed 0.2 threshold).
extern "C" void GCGE_CheckEvalMultiplicity(GCGE_INT start, GCGE_INT nev, GCGE_INT end, GCGE_INT *dim_x,
                                           GCGE_DOUBLE *eval)
{
    int tmp = start;
    if (start < nev) {
        fprintf(stderr, "GCGE_CheckEvalMultiplicity: start < nev\n");
        return;
    }
    for (int i = start; i < end; i++) {
        if (std::fabs(std::fabs(eval[tmp] / eval[nev]) - 1) < 0.2) tmp += 1;
        else break;
    }
    *dim_x = tmp;
}

// ------------------------------------------------------------------------
// Driver: GCGE_Solve, gcge_eigsol.c:65-542
// ------------------------------------------------------------------------
extern "C" void GCGE_Solve(void *A, void *B, GCGE_DOUBLE *eval, void **evec, GCGE_PARA *para, GCGE_OPS *ops,
                           GCGE_WORKSPACE *ws)
{
    const double t_start = now_s();
    const int nev = para->nev, ev_max_it = para->ev_max_it;
    void **V = ws->V;
    GCGE_INT ms[2], me[2];
    ws->evec = evec;
    para->bcg_calls = 0;
    para->bcg_iters_total = 0;
    void *Orth_mat, *RR_mat;
    if (std::strcmp(para->orth_type, "B")) {
        Orth_mat = A;
        RR_mat = B;
    } else {
        Orth_mat = B;
        RR_mat = A;
    }
    if (!RR_mat) {
        fprintf(stderr, "GCGE_Solve: orth_type \"A\" needs B\n");
        return;
    }

    // With w_orth_type "cholqr" the initial block is B-orthonormalized the same
    // way (one block factorization instead of the column-by-column
    // Gram-Schmidt of the reference, about 0.09 s against a few ms for 100
    // columns of 4.1M rows), when V_tmp can hold the block.
    const bool chol_init = std::strcmp(para->w_orth_type, "cholqr") == 0 && ws->V_tmp_size >= nev;
    auto orth_init = [&](int start, int *end) {
        if (chol_init)
            cholqr_orth(evec, start, end, Orth_mat, ops, para, ws);
        else
            orth_gs(evec, start, end, Orth_mat, ops, para, ws->V_tmp, ws->subspace_dtmp);
    };
    {
        GCGE_TIMER(x_orth_time, para);
        if (para->given_init_evec == 0) {
            int num_init_evec = 0;
            do {
                ops->MultiVecSetRandomValue(evec, num_init_evec, nev - num_init_evec, ops);
                if (std::strcmp(para->conv_type, "O") == 0 && num_init_evec == 0) {
                    // omega-norm estimate of A (:135-173)
                    double omega_F = 0.0, AO_F = 0.0, t;
                    std::vector<double> nrm(nev), a0(nev, 0.0), sc(nev);
                    set2(ms, 0, 0);
                    set2(me, nev, nev);
                    ops->MultiVecInnerProdDiag(evec, evec, nrm.data(), ms, me, ops);
                    for (int i = 0; i < nev; i++) sc[i] = 1.0 / std::sqrt(nrm[i]);
                    ops->MultiVecAxpbyDiag(a0.data(), evec, sc.data(), evec, ms, me, ops);
                    ops->MultiVecInnerProdDiag(evec, evec, nrm.data(), ms, me, ops);
                    for (int i = 0; i < nev; i++) omega_F += nrm[i];
                    ops->MatDotMultiVec(A, evec, V, ms, me, ops);
                    ops->MultiVecInnerProdDiag(V, V, nrm.data(), ms, me, ops);
                    for (int i = 0; i < nev; i++) {
                        t = nrm[i];
                        AO_F += t;
                    }
                    para->conv_omega_norm = std::sqrt(AO_F / omega_F);
                }
                if (para->dirichlet_boundary == 1 && ops->SetDirichletBoundary)
                    ops->SetDirichletBoundary(evec, nev, A, B);
                ws->dim_x = nev;
                orth_init(num_init_evec, &ws->dim_x);
                num_init_evec = ws->dim_x;
            } while (num_init_evec < nev && !gh::failed());
        } else {
            orth_init(0, &ws->dim_x);
        }
    }
    GCGE_PrintParaInfo(para);
    if (gh::failed()) {
        // a latched error (e.g. a non-finite norm of the initial block): no iteration
        para->solve_seconds = now_s() - t_start;
        return;
    }

    set2(ms, 0, 0);
    set2(me, nev, nev);
    ops->MultiVecCopy(evec, V, ms, me, ops);
    ws->dim_xpw = ws->dim_x;
    ws->last_dim_x = ws->dim_x;
    ws->dim_xp = 0;
    rr_matrix(RR_mat, V, ops, para, ws);
    rr_eigenpairs(ws->subspace_matrix, ws->eval, ws->subspace_evec, ops, para, ws);
    {
        GCGE_TIMER(x_axpy_time, para);
        compute_x(V, ws->subspace_evec, evec, nev, ws->dim_xpw, ops, ws);
    }
    {
        GCGE_TIMER(conv_time, para);
        GCGE_CheckConvergence(A, B, ws->eval, evec, ops, para, ws);
    }
    if (para->print_result || para->print_eval) GCGE_PrintIterationInfo(ws->eval, para);

    set2(ms, ws->num_soft_locked_in_last_iter, ws->num_soft_locked_in_last_iter);
    set2(me, nev, nev);
    ops->MultiVecCopy(evec, V, ms, me, ops);
    ws->dim_xp = ws->dim_x;
    {
        GCGE_TIMER(w_line_time, para);
        compute_w(A, B, V, ws->eval, ops, para, ws);
    }
    ws->last_dim_xpw = ws->dim_xpw;
    ws->dim_xpw = ws->dim_xp + ws->unconv_bs;
    {
        GCGE_TIMER(w_orth_time, para);
        if (std::strcmp(para->w_orth_type, "cholqr") == 0)
            cholqr_orth(V, ws->dim_x, &ws->dim_xpw, Orth_mat, ops, para, ws);
        else
            orth_gs(V, ws->dim_x, &ws->dim_xpw, Orth_mat, ops, para, ws->V_tmp, ws->subspace_dtmp);
    }
    ws->dim_xp = nev;
    rr_matrix(RR_mat, V, ops, para, ws);
    rr_eigenpairs(ws->subspace_matrix, ws->eval, ws->subspace_evec, ops, para, ws);
    {
        GCGE_TIMER(x_axpy_time, para);
        compute_x(V, ws->subspace_evec, evec, nev, ws->dim_xpw, ops, ws);
    }
    {
        GCGE_TIMER(conv_time, para);
        GCGE_CheckConvergence(A, B, ws->eval, evec, ops, para, ws);
    }
    if (para->print_result || para->print_eval) GCGE_PrintIterationInfo(ws->eval, para);

    double best_res = para->max_res, peak_res = para->max_res;
    int best_iter = para->num_iter;
    bool watching = false;
    para->stopped_on_stagnation = 0;
    // lowest-max_res iterate while watched: eigenvectors (locked ones merged
    // in from V) and eigenvalues
    void **snap = nullptr;
    std::vector<double> snap_eval, snap_pair_res;
    double snap_res = 0.0;
    auto take_snapshot = [&]() {
        if (!snap) ops->MultiVecCreateByMat(&snap, nev, A, ops);
        GCGE_INT s0[2] = {0, 0}, e0[2] = {nev, nev};
        ops->MultiVecCopy(evec, snap, s0, e0, ops);
        const int nsl = ws->num_soft_locked_in_last_iter;
        if (nsl > 0) {
            GCGE_INT e1[2] = {nsl, nsl};
            ops->MultiVecCopy(V, snap, s0, e1, ops);
        }
        snap_eval.assign(ws->eval, ws->eval + nev);
        snap_pair_res.assign(para->res, para->res + nev);
        snap_res = para->max_res;
    };
    while (para->num_unlock > 0 && para->num_iter < ev_max_it && !gh::failed()) {
        peak_res = std::max(peak_res, para->max_res);
        if (para->stagnation_window > 0 && para->stagnation_start > 0.0 &&
            !(para->max_res <= para->stagnation_start * peak_res) && !watching) {
            // not watched yet: residuals of a random start first grow
            best_res = para->max_res;
            best_iter = para->num_iter;
        } else if (para->stagnation_window > 0) {
            watching = true;
            if (!snap || para->max_res < snap_res) take_snapshot();
            if (para->max_res > 10.0 * snap_res) {
                para->stopped_on_stagnation = 2;
                break;
            }
            if (para->max_res < para->stagnation_ratio * best_res) {
                best_res = para->max_res;
                best_iter = para->num_iter;
            } else if (para->num_iter - best_iter >= para->stagnation_window) {
                para->stopped_on_stagnation = 1;
                break;
            }
        }
        compute_p(ws->subspace_evec, V, ops, para, ws);
        {
            GCGE_TIMER(x_axpy_time, para);
            set2(ms, ws->num_soft_locked_in_last_iter, ws->num_soft_locked_in_last_iter);
            set2(me, nev, nev);
            ops->MultiVecCopy(evec, V, ms, me, ops);
        }
        {
            GCGE_TIMER(w_line_time, para);
            compute_w(A, B, V, ws->eval, ops, para, ws);
        }
        ws->last_dim_xpw = ws->dim_xpw;
        ws->dim_xpw = ws->dim_xp + ws->unconv_bs;
        {
            GCGE_TIMER(w_orth_time, para);
            const char *wt = para->w_orth_type;
            if (std::strcmp(wt, "multi") == 0)
                stable_multi_orth(V, ws->dim_xp, &ws->dim_xpw, Orth_mat, ops, para, ws);
            else if (std::strcmp(wt, "cholqr") == 0)
                cholqr_orth(V, ws->dim_xp, &ws->dim_xpw, Orth_mat, ops, para, ws);
            else if (std::strcmp(wt, "bgs") == 0)
                orth_bgs(V, ws->dim_xp, &ws->dim_xpw, Orth_mat, ops, para, ws);
            else if (std::strcmp(wt, "cbgs") == 0)
                orth_cbgs(V, ws->dim_xp, &ws->dim_xpw, Orth_mat, ops, para, ws);
            else if (std::strcmp(wt, "scbgs") == 0)
                orth_scbgs(V, ws->dim_xp, &ws->dim_xpw, Orth_mat, ops, para, ws);
            else  // "gs", and any other name, as in the reference (gcge_eigsol.c:468-472)
                orth_gs(V, ws->dim_xp, &ws->dim_xpw, Orth_mat, ops, para, ws->V_tmp, ws->subspace_dtmp);
        }
        rr_matrix(RR_mat, V, ops, para, ws);
        rr_eigenpairs(ws->subspace_matrix, ws->eval, ws->subspace_evec, ops, para, ws);
        ws->last_dim_x = ws->dim_x;
        GCGE_CheckEvalMultiplicity(nev, nev, ws->max_dim_x, &ws->dim_x, ws->eval);
        {
            GCGE_TIMER(x_axpy_time, para);
            compute_x(V, ws->subspace_evec, evec, nev, ws->dim_xpw, ops, ws);
        }
        {
            GCGE_TIMER(conv_time, para);
            GCGE_CheckConvergence(A, B, ws->eval, evec, ops, para, ws);
        }
        if (para->print_result || para->print_eval) GCGE_PrintIterationInfo(ws->eval, para);
    }

    if (snap && para->num_unlock > 0 && snap_res < para->max_res) {
        // stopped without converging: hand back the best iterate seen
        set2(ms, 0, 0);
        set2(me, nev, nev);
        ops->MultiVecCopy(snap, evec, ms, me, ops);
        std::memcpy(ws->eval, snap_eval.data(), (size_t)nev * sizeof(double));
        std::memcpy(para->res, snap_pair_res.data(), (size_t)nev * sizeof(double));
        para->max_res = snap_res;
        ws->num_soft_locked_in_last_iter = 0;  // the copy already holds the locked vectors
    }
    if (snap) ops->MultiVecDestroy(&snap, nev, ops);
    std::memcpy(eval, ws->eval, (size_t)nev * sizeof(double));
    set2(ms, 0, 0);
    set2(me, ws->num_soft_locked_in_last_iter, ws->num_soft_locked_in_last_iter);
    if (me[0] > 0) ops->MultiVecCopy(V, evec, ms, me, ops);
    para->solve_seconds = now_s() - t_start;
    GCGE_PrintFinalInfo(eval, para);
}

// ------------------------------------------------------------------------
// Workspace, ops table and solver object
// ------------------------------------------------------------------------

extern "C" void GCGE_WORKSPACE_Create(GCGE_WORKSPACE **workspace)
{
    *workspace = (GCGE_WORKSPACE *)calloc(1, sizeof(GCGE_WORKSPACE));
}

// GCGE_WORKSPACE_Setup, gcge_workspace.c:61-119 (same sizes).
extern "C" void GCGE_WORKSPACE_Setup(GCGE_WORKSPACE *ws, GCGE_PARA *para, GCGE_OPS *ops, void *A)
{
    const int nev = para->nev, bs = para->block_size;
    const int max_dim_x_tmp = (nev * 1.25 < nev + 8) ? (int)(nev * 1.25) : (nev + 8);
    const int max_dim_x = (max_dim_x_tmp > nev + 3) ? max_dim_x_tmp : (nev + 3);
    ws->V_size = max_dim_x + 2 * bs;
    ws->V_tmp_size = max_dim_x - nev + bs;
    ws->CG_p_size = bs;
    ops->MultiVecCreateByMat(&ws->V, ws->V_size, A, ops);
    ops->MultiVecCreateByMat(&ws->V_tmp, ws->V_tmp_size, A, ops);
    ops->MultiVecCreateByMat(&ws->CG_p, ws->CG_p_size, A, ops);
    ws->max_dim_x = max_dim_x;
    ws->dim_x = nev;
    ws->dim_xp = 0;
    ws->dim_xpw = nev;
    ws->conv_bs = 0;
    ws->num_soft_locked_in_last_iter = 0;
    ws->unconv_bs = (nev < bs) ? nev : bs;
    const int max_dim_xpw = max_dim_x + 2 * bs;
    ws->eval = (double *)calloc(max_dim_xpw, sizeof(double));
    ws->subspace_dtmp = (double *)calloc((size_t)max_dim_xpw * max_dim_xpw + 40 * (size_t)max_dim_xpw, sizeof(double));
    ws->subspace_itmp = (int *)calloc(100 * (size_t)max_dim_xpw, sizeof(int));
    ws->subspace_evec = (double *)calloc((size_t)max_dim_xpw * max_dim_xpw, sizeof(double));
    ws->subspace_matrix = (double *)calloc((size_t)max_dim_xpw * max_dim_xpw, sizeof(double));
    ws->unlock = (int *)calloc(max_dim_x, sizeof(int));
    for (int i = 0; i < max_dim_x; i++) ws->unlock[i] = i;
}

extern "C" void GCGE_WORKSPACE_Free(GCGE_WORKSPACE **workspace, GCGE_PARA *, GCGE_OPS *ops)
{
    if (!workspace || !*workspace) return;
    GCGE_WORKSPACE *ws = *workspace;
    free(ws->unlock);
    free(ws->subspace_matrix);
    free(ws->subspace_evec);
    free(ws->eval);
    free(ws->subspace_dtmp);
    free(ws->subspace_itmp);
    if (ws->V) {
        ops->MultiVecDestroy(&ws->V, ws->V_size, ops);
        ops->MultiVecDestroy(&ws->V_tmp, ws->V_tmp_size, ops);
        ops->MultiVecDestroy(&ws->CG_p, ws->CG_p_size, ops);
    }
    for (int i = 0; i < 4; i++)
        if (ws->bcg32[i]) ops->MultiVecDestroy(&ws->bcg32[i], ws->bcg32_cols, ops);
    if (ws->mixed32) {
        GCGE_SOLVER *s32 = (GCGE_SOLVER *)ws->mixed32;
        ws->mixed32 = nullptr;
        GCGE_SOLVER_Free_All(&s32);
    }
    free(ws);
    *workspace = nullptr;
}

extern "C" void GCGE_OPS_Create(GCGE_OPS **ops) { *ops = (GCGE_OPS *)calloc(1, sizeof(GCGE_OPS)); }

extern "C" void GCGE_SetDenseOps(GCGE_OPS *ops);

// GCGE_OPS_Setup, gcge_ops.c:470-605: required slots are checked; missing
// dense slots get the LAPACK defaults.  This build has no per-column
// multivector defaults: the block-kernel slots must be provided by the
// backend (GCGE_HIP_SetOps fills them all).
extern "C" GCGE_INT GCGE_OPS_Setup(GCGE_OPS *ops)
{
    struct {
        const void *p;
        const char *name;
    } req[] = {
        {(const void *)ops->MatDotVec, "MatDotVec"},
        {(const void *)ops->VecAxpby, "VecAxpby"},
        {(const void *)ops->VecInnerProd, "VecInnerProd"},
        {(const void *)ops->VecLocalInnerProd, "VecLocalInnerProd"},
        {(const void *)ops->VecCreateByVec, "VecCreateByVec"},
        {(const void *)ops->VecCreateByMat, "VecCreateByMat"},
        {(const void *)ops->VecDestroy, "VecDestroy"},
        {(const void *)ops->VecSetRandomValue, "VecSetRandomValue"},
        {(const void *)ops->MultiVecCreateByMat, "MultiVecCreateByMat"},
        {(const void *)ops->MultiVecDestroy, "MultiVecDestroy"},
        {(const void *)ops->MultiVecSetRandomValue, "MultiVecSetRandomValue"},
        {(const void *)ops->MatDotMultiVec, "MatDotMultiVec"},
        {(const void *)ops->MultiVecAxpby, "MultiVecAxpby"},
        {(const void *)ops->MultiVecAxpbyColumn, "MultiVecAxpbyColumn"},
        {(const void *)ops->MultiVecLinearComb, "MultiVecLinearComb"},
        {(const void *)ops->MultiVecInnerProd, "MultiVecInnerProd"},
        {(const void *)ops->MultiVecSwap, "MultiVecSwap"},
        {(const void *)ops->GetVecFromMultiVec, "GetVecFromMultiVec"},
        {(const void *)ops->RestoreVecForMultiVec, "RestoreVecForMultiVec"},
        {(const void *)ops->MultiVecAxpbyDiag, "MultiVecAxpbyDiag"},
        {(const void *)ops->MultiVecInnerProdDiag, "MultiVecInnerProdDiag"},
        {(const void *)ops->MultiVecLinearCombAxpby, "MultiVecLinearCombAxpby"},
        {(const void *)ops->MatDotMultiVecInnerProdDiag, "MatDotMultiVecInnerProdDiag"},
        {(const void *)ops->BlockCGUpdate, "BlockCGUpdate"},
        {(const void *)ops->MultiVecResidualNorms, "MultiVecResidualNorms"},
        {(const void *)ops->MultiVecGather, "MultiVecGather"},
        {(const void *)ops->MultiVecCopy, "MultiVecCopy"},
        {(const void *)ops->BlockCGResidualUpdate, "BlockCGResidualUpdate"},
        {(const void *)ops->BlockCGSearchUpdate, "BlockCGSearchUpdate"},
    };
    for (auto &r : req) {
        if (!r.p) {
            printf("ERROR: Please provide the %s operation!\n", r.name);
            return 1;
        }
    }
    if (!ops->DenseMatEigenSolver || !ops->DenseMatDotDenseMat || !ops->DenseSymMatDotDenseMat ||
        !ops->ArrayDotArray || !ops->ArrayNorm || !ops->ArrayAXPBY || !ops->ArrayCopy || !ops->ArrayScale) {
        GCGE_OPS tmp;
        std::memset(&tmp, 0, sizeof(tmp));
        GCGE_SetDenseOps(&tmp);
        if (!ops->DenseMatEigenSolver) ops->DenseMatEigenSolver = tmp.DenseMatEigenSolver;
        if (!ops->DenseMatDotDenseMat) ops->DenseMatDotDenseMat = tmp.DenseMatDotDenseMat;
        if (!ops->DenseSymMatDotDenseMat) ops->DenseSymMatDotDenseMat = tmp.DenseSymMatDotDenseMat;
        if (!ops->ArrayDotArray) ops->ArrayDotArray = tmp.ArrayDotArray;
        if (!ops->ArrayNorm) ops->ArrayNorm = tmp.ArrayNorm;
        if (!ops->ArrayAXPBY) ops->ArrayAXPBY = tmp.ArrayAXPBY;
        if (!ops->ArrayCopy) ops->ArrayCopy = tmp.ArrayCopy;
        if (!ops->ArrayScale) ops->ArrayScale = tmp.ArrayScale;
    }
    return 0;
}

extern "C" void GCGE_OPS_Free(GCGE_OPS **ops)
{
    free(*ops);
    *ops = nullptr;
}

extern "C" void GCGE_OPS_SetLinearSolverWorkspace(GCGE_OPS *ops, void *w) { ops->linear_solver_workspace = w; }

extern "C" void GCGE_SOLVER_Create(GCGE_SOLVER **solver)
{
    GCGE_SOLVER *s = (GCGE_SOLVER *)calloc(1, sizeof(GCGE_SOLVER));
    GCGE_PARA_Create(&s->para);
    GCGE_OPS_Create(&s->ops);
    GCGE_WORKSPACE_Create(&s->workspace);
    *solver = s;
}

extern "C" void GCGE_SOLVER_Free(GCGE_SOLVER **solver)
{
    if (!solver || !*solver) return;
    GCGE_WORKSPACE_Free(&(*solver)->workspace, (*solver)->para, (*solver)->ops);
    GCGE_PARA_Free(&(*solver)->para);
    GCGE_OPS_Free(&(*solver)->ops);
    free(*solver);
    *solver = nullptr;
}

extern "C" void GCGE_SOLVER_Free_All(GCGE_SOLVER **solver)
{
    if (!solver || !*solver) return;
    GCGE_SOLVER *s = *solver;
    if (s->evec && s->ops->MultiVecDestroy) s->ops->MultiVecDestroy(&s->evec, s->para->nev, s->ops);
    free(s->eval);
    s->eval = nullptr;
    GCGE_SOLVER_Free(solver);
}

// GCGE_SOLVER_Setup, gcge_solver.c:56-61.  A failed ops setup leaves the
// workspace empty, and GCGE_SOLVER_Solve then refuses to run.
extern "C" void GCGE_SOLVER_Setup(GCGE_SOLVER *solver)
{
    GCGE_PARA_Setup(solver->para);
    if (GCGE_OPS_Setup(solver->ops) != 0) return;
    GCGE_WORKSPACE_Setup(solver->workspace, solver->para, solver->ops, solver->A);
}

extern "C" void GCGE_SOLVER_SetMatA(GCGE_SOLVER *s, void *A) { s->A = A; }
extern "C" void GCGE_SOLVER_SetMatB(GCGE_SOLVER *s, void *B) { s->B = B; }
extern "C" void GCGE_SOLVER_SetEigenvalues(GCGE_SOLVER *s, GCGE_DOUBLE *eval) { s->eval = eval; }
extern "C" void GCGE_SOLVER_SetEigenvectors(GCGE_SOLVER *s, void **evec) { s->evec = evec; }
extern "C" void GCGE_SOLVER_SetNumEigen(GCGE_SOLVER *s, GCGE_INT nev) { GCGE_PARA_SetNumEigen(s->para, nev); }
extern "C" void GCGE_SOLVER_SetOpsLinearSolverWorkspace(GCGE_SOLVER *s, void *w)
{
    GCGE_OPS_SetLinearSolverWorkspace(s->ops, w);
}
extern "C" void GCGE_SOLVER_GetEigenvalues(GCGE_SOLVER *s, GCGE_DOUBLE **eval) { *eval = s->eval; }
extern "C" void GCGE_SOLVER_GetEigenvectors(GCGE_SOLVER *s, void ***evec) { *evec = s->evec; }

// The per-solve state starts as GCGE_PARA_Setup and GCGE_WORKSPACE_Setup
// leave it, so a set-up solver can solve again (with new initial vectors or
// new matrix values) without reallocating its workspace; for the first solve
// this changes nothing.
extern "C" void GCGE_SOLVER_Solve(GCGE_SOLVER *s)
{
    if (!s->workspace || !s->workspace->V) {
        fprintf(stderr, "GCGE_SOLVER_Solve: solver is not set up (missing operations?)\n");
        return;
    }
    GCGE_PARA *P = s->para;
    GCGE_WORKSPACE *ws = s->workspace;
    const int nev = P->nev;
    P->num_iter = 0;
    P->num_unlock = nev;
    P->sum_res = 0.0;
    for (int i = 0; i < nev; i++) P->res[i] = 0.0;
    // part times describe this solve (a reused solver starts them at zero)
    if (P->stat_para && P->stat_para->part_time_total) *P->stat_para->part_time_total = GCGE_PART_TIME{};
    ws->dim_x = nev;
    ws->dim_xp = 0;
    ws->dim_xpw = nev;
    ws->conv_bs = 0;
    ws->num_soft_locked_in_last_iter = 0;
    ws->unconv_bs = (nev < P->block_size) ? nev : P->block_size;
    for (int i = 0; i < ws->max_dim_x; i++) ws->unlock[i] = i;
    GCGE_Solve(s->A, s->B, s->eval, s->evec, s->para, s->ops, s->workspace);
}
