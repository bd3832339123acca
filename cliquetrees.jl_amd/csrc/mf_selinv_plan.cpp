// mf_selinv_plan.cpp — schedule of the selected inversion of a Cholesky
// factor (selinv!, src/Multifrontal.jl/src/selinv.jl:38-424) over the
// internal fronts of a plan, and a host executor of that schedule.
//
// Per internal front J (nn columns, na update rows, root first), with the
// factor panel [L11; L21] and M = Sigma(sep, sep) gathered from the parent's
// frontal matrix (selinv_send!, selinv.jl:381-424):
//   X        = L11^-1                          (trtri!, selinv.jl:236-239)
//   Lt       = L21 X                           (trsm!,  selinv.jl:210-217)
//   Sigma21  = -M Lt                           (symm!,  selinv.jl:218-225)
//   Sigma11  = X' X - Lt' Sigma21              (trrk! and syrk!, selinv.jl:226-246)
// and the children then gather their M from [Sigma11, Sigma21'; Sigma21, M].
//
// Big fronts (nJ > SMALL_NJ) are tiled by NB = 64.  X is built from the
// factor's inverted diagonal tiles (the W pool) by the 2 x 2 block recursion
//   [A 0; B C]^-1 = [A^-1 0; -C^-1 B A^-1, C^-1]
// over column-tile ranges (split [a, b) at m = (a + b) / 2): one phase of
// T = B A^-1 tiles, one of -C^-1 T tiles per recursion depth, deepest first,
// so the chain of dependent steps is 2 log2(nn / 64) long instead of nn / 64.
// Every other product is a phase of independent 64 x 64 tiles.  A small
// front (nJ <= SMALL_NJ) is one task that does all of it in LDS.
//
// Tickets: the independent phases (X, Lt) of the fronts at depth d + 1 come
// before the dependent ones (gather, Sigma21, Sigma11) of depth d, so while a
// level waits for its parents the next level's inversions fill the device.
// Each task waits for the earlier phases of its own front (counter J >= the
// number of tasks in those phases) and, from the gather on, for its parent's
// completion (counter P >= total[P]); both are advanced only by tasks with
// smaller tickets, so the order is deadlock-free on any number of resident
// workgroups.  mf_selinv_check runs the schedule on the host in ticket order
// and checks every wait and every buffer bound on the way.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mf_plan.h"

namespace mf {

namespace {

int tile_len(int64_t total, int i) { return (int)std::min<int64_t>(NB, total - (int64_t)NB * i); }
int ntiles(int64_t total) { return (int)((total + NB - 1) / NB); }

SelTask gemm_task(int J, int8_t arrC, int64_t c_off, int ldc, int m, int n, bool lower) {
    SelTask t;
    std::memset(&t, 0, sizeof t);
    t.kind = ST_GEMM;
    t.front = J;
    t.arrC = arrC;
    t.c_off = c_off;
    t.ldc = ldc;
    t.m = m;
    t.n = n;
    t.lower = lower ? 1 : 0;
    t.wait[0] = t.wait[1] = -1;
    return t;
}

void add_part(SelTask& t, int8_t arrA, int64_t a_off, int lda, int ta, int8_t arrB, int64_t b_off,
              int ldb, int tb, int k, int sgn) {
    SelPart& p = t.p[t.np++];
    p.arrA = arrA;
    p.a_off = a_off;
    p.lda = lda;
    p.ta = (int8_t)ta;
    p.arrB = arrB;
    p.b_off = b_off;
    p.ldb = ldb;
    p.tb = (int8_t)tb;
    p.k = k;
    p.sgn = sgn;
}

SelTask plain_task(int kind, int J, int a) {
    SelTask t;
    std::memset(&t, 0, sizeof t);
    t.kind = kind;
    t.front = J;
    t.a = a;
    t.wait[0] = t.wait[1] = -1;
    return t;
}

}  // namespace

std::string build_selinv(const HostPlan& H, SelPlan& S) {
    const int32_t NJ = (int32_t)H.fronts.size();
    S = SelPlan();
    S.total.assign(NJ, 0);
    S.xoff.assign(NJ, -1);
    S.ltoff.assign(NJ, -1);
    S.moff.assign(NJ, -1);
    std::vector<int32_t> depth(NJ, 0);
    for (int32_t J = NJ - 1; J >= 0; --J) {
        const int32_t p = H.fronts[J].parent;
        depth[J] = p < 0 ? 0 : depth[p] + 1;
        S.max_depth = std::max(S.max_depth, depth[J]);
    }
    for (int32_t J = 0; J < NJ; ++J) {
        const Front& F = H.fronts[J];
        if (!H.is_small[J]) {
            if (F.woff < 0) return "big front without inverted diagonal tiles";
            S.xoff[J] = S.x_elems;
            S.x_elems += (int64_t)F.nn * F.nn;
            if (F.na > 0) {
                S.ltoff[J] = S.lt_elems;
                S.lt_elems += (int64_t)F.na * F.nn;
            }
        }
        if (F.na > 0) {
            S.moff[J] = S.m_elems;
            S.m_elems += (int64_t)F.na * F.na;
        }
    }
    // per front: its independent tasks (A) and its dependent tasks (B), each
    // tagged with the number of the front's tasks in earlier phases
    std::vector<std::vector<SelTask>> A(NJ), B(NJ);
    for (int32_t J = 0; J < NJ; ++J) {
        const Front& F = H.fronts[J];
        const int nn = F.nn, na = F.na, nJ = F.nn + F.na;
        int32_t done = 0;  // tasks of J in finished phases
        auto close_phase = [&](std::vector<SelTask>& v, size_t first) {
            for (size_t i = first; i < v.size(); ++i) {
                v[i].wait[0] = done > 0 ? J : -1;
                v[i].target[0] = done;
            }
            done += (int32_t)(v.size() - first);
        };
        if (H.is_small[J]) {
            SelTask t = plain_task(ST_SMALL, J, 0);
            B[J].push_back(t);
            close_phase(B[J], 0);
        } else {
            const int nk = ntiles(nn), nr = ntiles(na);
            const int64_t po = F.poff, xo = S.xoff[J], lo = S.ltoff[J], mo = S.moff[J];
            // X diagonal tiles from W
            size_t f0 = A[J].size();
            for (int K = 0; K < nk; ++K) A[J].push_back(plain_task(ST_XDIAG, J, K));
            close_phase(A[J], f0);
            // recursion nodes by depth
            std::vector<std::vector<std::pair<int, int>>> lev;
            std::vector<std::pair<std::pair<int, int>, int>> stack{{{0, nk}, 0}};
            while (!stack.empty()) {
                auto [ab, l] = stack.back();
                stack.pop_back();
                if (ab.second - ab.first < 2) continue;
                if ((int)lev.size() <= l) lev.resize(l + 1);
                lev[l].push_back(ab);
                const int m = (ab.first + ab.second) / 2;
                stack.push_back({{ab.first, m}, l + 1});
                stack.push_back({{m, ab.second}, l + 1});
            }
            for (int l = (int)lev.size() - 1; l >= 0; --l) {
                // T(i, j) = sum_{k in [j, m)} L(i, k) X(k, j), i in [m, b), j in [a, m)
                f0 = A[J].size();
                for (auto [a, b] : lev[l]) {
                    const int m = (a + b) / 2;
                    for (int i = m; i < b; ++i)
                        for (int j = a; j < m; ++j) {
                            SelTask t = gemm_task(J, SP_T, xo + (int64_t)NB * i + (int64_t)NB * j * nn, nn,
                                                  tile_len(nn, i), NB, false);
                            add_part(t, SP_PANEL, po + (int64_t)NB * i + (int64_t)NB * j * nJ, nJ, 0, SP_X,
                                     xo + (int64_t)NB * j + (int64_t)NB * j * nn, nn, 1, NB * (m - j), 1);
                            A[J].push_back(t);
                        }
                }
                close_phase(A[J], f0);
                // X(i, j) = -sum_{k in [m, i]} X(i, k) T(k, j)
                f0 = A[J].size();
                for (auto [a, b] : lev[l]) {
                    const int m = (a + b) / 2;
                    for (int i = m; i < b; ++i)
                        for (int j = a; j < m; ++j) {
                            SelTask t = gemm_task(J, SP_X, xo + (int64_t)NB * i + (int64_t)NB * j * nn, nn,
                                                  tile_len(nn, i), NB, false);
                            add_part(t, SP_X, xo + (int64_t)NB * i + (int64_t)NB * m * nn, nn, 0, SP_T,
                                     xo + (int64_t)NB * m + (int64_t)NB * j * nn, nn, 1,
                                     NB * (i - m) + tile_len(nn, i), -1);
                            A[J].push_back(t);
                        }
                }
                close_phase(A[J], f0);
            }
            if (na > 0) {
                // Lt(i, j) = sum_{k >= 64 j} L21(i, k) X(k, j)
                f0 = A[J].size();
                for (int i = 0; i < nr; ++i)
                    for (int j = 0; j < nk; ++j) {
                        SelTask t = gemm_task(J, SP_LT, lo + (int64_t)NB * i + (int64_t)NB * j * na, na,
                                              tile_len(na, i), tile_len(nn, j), false);
                        add_part(t, SP_PANEL, po + nn + (int64_t)NB * i + (int64_t)NB * j * nJ, nJ, 0, SP_X,
                                 xo + (int64_t)NB * j + (int64_t)NB * j * nn, nn, 1, nn - NB * j, 1);
                        A[J].push_back(t);
                    }
                close_phase(A[J], f0);
                // gather M from the parent, one column tile per task
                f0 = B[J].size();
                for (int jb = 0; jb < nr; ++jb) B[J].push_back(plain_task(ST_GATHER, J, jb));
                // the B list continues the phase count of the A list
                close_phase(B[J], f0);
                // Sigma21(i, j) = -sum_k M(i, k) Lt(k, j)
                f0 = B[J].size();
                for (int i = 0; i < nr; ++i)
                    for (int j = 0; j < nk; ++j) {
                        SelTask t = gemm_task(J, SP_SIG, po + nn + (int64_t)NB * i + (int64_t)NB * j * nJ, nJ,
                                              tile_len(na, i), tile_len(nn, j), false);
                        add_part(t, SP_M, mo + (int64_t)NB * i, na, 0, SP_LT, lo + (int64_t)NB * j * na, na, 1,
                                 na, -1);
                        B[J].push_back(t);
                    }
                close_phase(B[J], f0);
            }
            // Sigma11(i, j) = sum_{k >= 64 i} X(k, i) X(k, j) - sum_k Lt(k, i) Sigma21(k, j), i >= j
            f0 = B[J].size();
            for (int i = 0; i < nk; ++i)
                for (int j = 0; j <= i; ++j) {
                    SelTask t = gemm_task(J, SP_SIG, po + (int64_t)NB * i + (int64_t)NB * j * nJ, nJ,
                                          tile_len(nn, i), tile_len(nn, j), i == j);
                    add_part(t, SP_X, xo + (int64_t)NB * i + (int64_t)NB * i * nn, nn, 1, SP_X,
                             xo + (int64_t)NB * i + (int64_t)NB * j * nn, nn, 1, nn - NB * i, 1);
                    if (na > 0)
                        add_part(t, SP_LT, lo + (int64_t)NB * i * na, na, 1, SP_SIG,
                                 po + nn + (int64_t)NB * j * nJ, nJ, 1, na, -1);
                    B[J].push_back(t);
                }
            close_phase(B[J], f0);
        }
        // the parent's completion gates the front's dependent tasks (the first
        // of them reads the parent's frontal matrix)
        if (F.parent >= 0)
            for (SelTask& t : B[J]) t.wait[1] = F.parent;
        S.total[J] = done;
    }
    for (int32_t J = 0; J < NJ; ++J)
        for (SelTask& t : B[J])
            if (t.wait[1] >= 0) t.target[1] = S.total[t.wait[1]];
    // ticket order: A(0), then per depth d: A(d + 1), B(d); fronts of a depth
    // by decreasing size (the longest chains first)
    std::vector<std::vector<int32_t>> bydepth(S.max_depth + 1);
    for (int32_t J = NJ - 1; J >= 0; --J) bydepth[depth[J]].push_back(J);
    for (auto& v : bydepth)
        std::stable_sort(v.begin(), v.end(), [&](int32_t a, int32_t b) {
            const Front &fa = H.fronts[a], &fb = H.fronts[b];
            return (int64_t)fa.nn * (fa.nn + fa.na) > (int64_t)fb.nn * (fb.nn + fb.na);
        });
    auto emit = [&](std::vector<std::vector<SelTask>>& L, int d) {
        for (int32_t J : bydepth[d])
            for (const SelTask& t : L[J]) S.task.push_back(t);
    };
    emit(A, 0);
    for (int d = 0; d <= S.max_depth; ++d) {
        if (d + 1 <= S.max_depth) emit(A, d + 1);
        emit(B, d);
    }
    for (const SelTask& t : S.task) {
        if (t.kind == ST_GEMM) {
            ++S.n_gemm;
            for (int q = 0; q < t.np; ++q) S.flops += 2.0 * t.m * t.n * t.p[q].k;
        } else if (t.kind == ST_SMALL) {
            ++S.n_small;
            const Front& F = H.fronts[t.front];
            const double nn = F.nn, na = F.na;
            S.flops += 2.0 * na * na * nn + 2.0 * na * nn * nn + 2.0 / 3.0 * nn * nn * nn;
        } else if (t.kind == ST_GATHER) {
            ++S.n_gather;
        }
    }
    // bounds of every product against its buffers
    const int64_t size[SP_N] = {H.panel_elems, H.w_elems, S.x_elems, S.x_elems, S.lt_elems, S.m_elems,
                                H.panel_elems};
    auto span = [](int64_t off, int rows, int cols, int ld, int trans) {
        // last element of a rows x cols operand (cols: the k extent)
        return trans ? off + (cols - 1) + (int64_t)(rows - 1) * ld : off + (rows - 1) + (int64_t)(cols - 1) * ld;
    };
    for (const SelTask& t : S.task) {
        if (t.kind != ST_GEMM) continue;
        if (t.m <= 0 || t.n <= 0 || t.m > NB || t.n > NB) return "selinv tile shape out of range";
        if (t.c_off < 0 || span(t.c_off, t.m, t.n, t.ldc, 0) >= size[t.arrC]) return "selinv output out of bounds";
        for (int q = 0; q < t.np; ++q) {
            const SelPart& p = t.p[q];
            if (p.k <= 0) return "selinv empty product";
            if (p.a_off < 0 || span(p.a_off, t.m, p.k, p.lda, p.ta) >= size[p.arrA])
                return "selinv operand A out of bounds";
            if (p.b_off < 0 || span(p.b_off, t.n, p.k, p.ldb, p.tb) >= size[p.arrB])
                return "selinv operand B out of bounds";
        }
    }
    return std::string();
}

// ---- the completion's backward solves ---------------------------------------
//
// Per big root J of a completion forest (nn_f = nJ' columns, factor
// L = chol(K') in the panel, inverted 64 x 64 diagonal tiles in W), with
// the separator part padded to na' = 64 s rows and nn = nnr[J] residual
// columns after it: Y = L^-T [0; I_nn] (nJ' x nn, ld nJ', in SP_X at
// xoff[J]), whose rows are the last nn rows of L^-1.  RHS column tile c
// has its identity block in row tile s + c (= top), so its chain starts
// there and runs up, I = top - 1 down to 0:
//   U(I, c) = -sum_{K = I+1}^{min(I+H-1, top)} L(K, I)' Y(K, c)   (SP_T)
//   R(I, c) = -sum_{K = I+H}^{top} L(K, I)' Y(K, c)                (SP_M)
//   Y(I, c) = W_I' U(I, c) + W_I' R(I, c),  Y(top, c) = W_top' I   (SP_LT: identity)
// The chain's own tasks (U and Y, at most H - 1 tiles of reduction) are
// phases of the front's counter; R(I, c), the long rest of the reduction,
// can start once Y(I + H, c) is done and runs beside the chain's next
// H - 1 steps, on a counter of its own that Y(I, c) waits for.  The fronts'
// phases are interleaved round-robin (the longest chains first), each R
// task right after the phase that completes its inputs.
std::string build_cpl_solve(const HostPlan& H, const std::vector<int32_t>& pad_na,
                            const std::vector<int32_t>& nnr, SelPlan& S) {
    constexpr int LOOK = 8;  // H: tiles of the reduction on the chain
    const int32_t NJ = (int32_t)H.fronts.size();
    S = SelPlan();
    S.total.assign(NJ, 0);
    S.xoff.assign(NJ, -1);
    S.ltoff.assign(NJ, -1);
    S.moff.assign(NJ, -1);
    struct Phase {
        std::vector<SelTask> main, extra;  // extra: R tasks (own counters), after main
    };
    std::vector<std::vector<Phase>> ph(NJ);
    int32_t nctr = NJ;
    for (int32_t J = 0; J < NJ; ++J) {
        if (H.is_small[J]) continue;
        const Front& F = H.fronts[J];
        const int nJ = F.nn, s = pad_na[J] / NB, nn = nnr[J], nrt = ntiles(nn);
        if (F.na != 0 || F.woff < 0 || pad_na[J] % NB || pad_na[J] + nn != nJ || nn <= 0)
            return "completion solve: front " + std::to_string(J) + " is not a padded forest root";
        const int64_t po = F.poff, wo = F.woff, yo = S.x_elems;
        S.xoff[J] = yo;
        S.x_elems += (int64_t)nJ * NB * nrt;
        const int top_max = s + nrt - 1;
        // phase index of Y(I): the phases run U(I), Y(I) for I = top_max .. 0
        // (no U phase at I = top_max), so Y(I) is phase 2 (top_max - I)
        std::vector<Phase> P(2 * (top_max + 1) - 1);
        auto yph = [&](int I) { return 2 * (top_max - I); };
        auto rows_to = [&](int K) { return std::min<int64_t>((int64_t)NB * (K + 1), nJ); };  // end of tile K
        for (int c = 0; c < nrt; ++c) {
            const int top = s + c, mc = tile_len(nn, c);
            for (int I = top; I >= 0; --I) {
                const int m = tile_len(nJ, I);
                const int64_t cy = (int64_t)NB * I + (int64_t)NB * c * nJ;  // tile (I, c)
                SelTask y = gemm_task(J, SP_X, yo + cy, nJ, m, mc, false);
                if (I == top) {
                    add_part(y, SP_W, wo + (int64_t)NB * I, nJ, 1, SP_LT, 0, NB, 1, m, 1);
                    P[yph(I)].main.push_back(y);
                    continue;
                }
                const int knear = std::min(I + LOOK - 1, top);
                SelTask u = gemm_task(J, SP_T, yo + cy, nJ, m, mc, false);
                add_part(u, SP_PANEL, po + (int64_t)NB * (I + 1) + (int64_t)NB * I * nJ, nJ, 1, SP_X,
                         yo + (int64_t)NB * (I + 1) + (int64_t)NB * c * nJ, nJ, 1,
                         (int)(rows_to(knear) - (int64_t)NB * (I + 1)), -1);
                P[yph(I) - 1].main.push_back(u);
                add_part(y, SP_W, wo + (int64_t)NB * I, nJ, 1, SP_T, yo + cy, nJ, 1, m, 1);
                if (I + LOOK <= top) {
                    SelTask r = gemm_task(nctr, SP_M, yo + cy, nJ, m, mc, false);
                    add_part(r, SP_PANEL, po + (int64_t)NB * (I + LOOK) + (int64_t)NB * I * nJ, nJ, 1, SP_X,
                             yo + (int64_t)NB * (I + LOOK) + (int64_t)NB * c * nJ, nJ, 1,
                             (int)(rows_to(top) - (int64_t)NB * (I + LOOK)), -1);
                    r.b = J;  // the front whose chain it waits for
                    P[yph(I + LOOK)].extra.push_back(r);
                    add_part(y, SP_W, wo + (int64_t)NB * I, nJ, 1, SP_M, yo + cy, nJ, 1, m, 1);
                    y.wait[1] = nctr;
                    y.target[1] = 1;
                    ++nctr;
                }
                P[yph(I)].main.push_back(y);
            }
        }
        for (Phase& q : P)
            if (!q.main.empty() || !q.extra.empty()) ph[J].push_back(std::move(q));
    }
    S.ncounters = nctr;
    // ticket order: phase p of every front, for p = 0, 1, ...; fronts with
    // more phases first; a chain task waits for its front's earlier phases,
    // an R task for its front's phases through its own
    std::vector<int32_t> order;
    for (int32_t J = 0; J < NJ; ++J)
        if (!ph[J].empty()) order.push_back(J);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return ph[a].size() > ph[b].size(); });
    std::vector<int32_t> done(NJ, 0);
    for (size_t p = 0; !order.empty() && p < ph[order[0]].size(); ++p)
        for (int32_t J : order) {
            if (p >= ph[J].size()) break;  // sorted: the rest have fewer phases
            for (SelTask t : ph[J][p].main) {
                t.wait[0] = done[J] > 0 ? J : -1;
                t.target[0] = done[J];
                S.task.push_back(t);
            }
            done[J] += (int32_t)ph[J][p].main.size();
            for (SelTask t : ph[J][p].extra) {
                t.wait[0] = J;
                t.target[0] = done[J];
                S.task.push_back(t);
            }
        }
    for (int32_t J = 0; J < NJ; ++J) S.total[J] = done[J];
    for (const SelTask& t : S.task) {
        ++S.n_gemm;
        for (int q = 0; q < t.np; ++q) S.flops += 2.0 * t.m * t.n * t.p[q].k;
    }
    // bounds of every product against its buffers, and every wait against
    // the tasks of smaller tickets
    const int64_t size[SP_N] = {H.panel_elems, H.w_elems, S.x_elems, S.x_elems, (int64_t)NB * NB, S.x_elems, 0};
    auto last = [](int64_t off, int rows, int cols, int ld, int trans) {
        return trans ? off + (cols - 1) + (int64_t)(rows - 1) * ld : off + (rows - 1) + (int64_t)(cols - 1) * ld;
    };
    std::vector<int32_t> seen(nctr, 0);
    for (const SelTask& t : S.task) {
        if (t.m <= 0 || t.n <= 0 || t.m > NB || t.n > NB || t.np < 1) return "completion solve: bad tile";
        if (t.c_off < 0 || last(t.c_off, t.m, t.n, t.ldc, 0) >= size[t.arrC]) return "completion solve: output out of bounds";
        for (int q = 0; q < t.np; ++q) {
            const SelPart& pp = t.p[q];
            if (pp.k <= 0) return "completion solve: empty product";
            if (pp.a_off < 0 || last(pp.a_off, t.m, pp.k, pp.lda, pp.ta) >= size[pp.arrA])
                return "completion solve: operand A out of bounds";
            if (pp.b_off < 0 || last(pp.b_off, t.n, pp.k, pp.ldb, pp.tb) >= size[pp.arrB])
                return "completion solve: operand B out of bounds";
        }
        for (int w = 0; w < 2; ++w)
            if (t.wait[w] >= 0 && (t.wait[w] >= nctr || seen[t.wait[w]] < t.target[w]))
                return "completion solve: wait on a later ticket";
        if (t.front < 0 || t.front >= nctr) return "completion solve: counter out of range";
        ++seen[t.front];
    }
    return std::string();
}

// ---- host executor of the schedule (plan validation) -----------------------

namespace {

struct HostSel {
    const HostPlan& H;
    const SelPlan& S;
    std::vector<double> pool[SP_N];
    std::vector<int32_t> cnt;
    HostSel(const HostPlan& h, const SelPlan& s) : H(h), S(s) {}

    double fp(int32_t P, int p, int q) const {  // lower entry (p >= q) of P's frontal matrix
        const Front& F = H.fronts[P];
        const int nJ = F.nn + F.na;
        if (q < F.nn) return pool[SP_SIG][F.poff + p + (int64_t)q * nJ];
        return pool[SP_M][S.moff[P] + (p - F.nn) + (int64_t)(q - F.nn) * F.na];
    }

    void gather(int32_t J, int b0, int b1) {
        const Front& F = H.fronts[J];
        const int32_t* rel = H.rel.data() + F.reloff;
        double* M = pool[SP_M].data() + S.moff[J];
        for (int b = b0; b < b1; ++b)
            for (int a = b; a < F.na; ++a) {
                const double v = fp(F.parent, rel[a], rel[b]);
                M[a + (int64_t)b * F.na] = v;
                M[b + (int64_t)a * F.na] = v;
            }
    }

    void gemm(const SelTask& t) {
        std::vector<double> acc((size_t)t.m * t.n, 0.0);
        for (int q = 0; q < t.np; ++q) {
            const SelPart& p = t.p[q];
            const double* Ab = pool[p.arrA].data() + p.a_off;
            const double* Bb = pool[p.arrB].data() + p.b_off;
            for (int c = 0; c < t.n; ++c)
                for (int r = 0; r < t.m; ++r) {
                    double s = 0.0;
                    for (int k = 0; k < p.k; ++k) {
                        const double a = p.ta ? Ab[k + (int64_t)r * p.lda] : Ab[r + (int64_t)k * p.lda];
                        const double b = p.tb ? Bb[k + (int64_t)c * p.ldb] : Bb[c + (int64_t)k * p.ldb];
                        s += a * b;
                    }
                    acc[r + (size_t)c * t.m] += p.sgn * s;
                }
        }
        double* C = pool[t.arrC].data() + t.c_off;
        for (int c = 0; c < t.n; ++c)
            for (int r = 0; r < t.m; ++r)
                if (!t.lower || r >= c) C[r + (int64_t)c * t.ldc] = acc[r + (size_t)c * t.m];
    }

    // The small-front task: the algorithm of the LDS kernel, in plain loops.
    void small(int32_t J) {
        const Front& F = H.fronts[J];
        const int nn = F.nn, na = F.na, nJ = nn + na;
        const double* Lp = pool[SP_PANEL].data() + F.poff;
        std::vector<double> L(Lp, Lp + (size_t)nJ * nn), M((size_t)na * na), S21((size_t)na * nn);
        if (na > 0) {
            gather(J, 0, na);
            std::copy(pool[SP_M].begin() + S.moff[J], pool[SP_M].begin() + S.moff[J] + (int64_t)na * na,
                      M.begin());
        }
        auto l = [&](int r, int c) -> double& { return L[r + (size_t)c * nJ]; };
        // Lt = L21 L11^-1, last column first
        for (int c = nn - 1; c >= 0; --c) {
            const double inv = 1.0 / l(c, c);
            for (int r = nn; r < nJ; ++r) l(r, c) *= inv;
            for (int j = 0; j < c; ++j)
                for (int r = nn; r < nJ; ++r) l(r, j) -= l(r, c) * l(c, j);
        }
        // X = L11^-1 in place, last column first
        for (int j = nn - 1; j >= 0; --j) {
            const double xjj = 1.0 / l(j, j);
            std::vector<double> col(nn, 0.0);
            for (int i = j + 1; i < nn; ++i) {
                double s = 0.0;
                for (int k = j + 1; k <= i; ++k) s += l(i, k) * l(k, j);
                col[i] = -xjj * s;
            }
            for (int i = j + 1; i < nn; ++i) l(i, j) = col[i];
            l(j, j) = xjj;
        }
        // Sigma21 = -M Lt
        for (int c = 0; c < nn; ++c)
            for (int r = 0; r < na; ++r) {
                double s = 0.0;
                for (int k = 0; k < na; ++k) s += M[r + (size_t)k * na] * l(nn + k, c);
                S21[r + (size_t)c * na] = -s;
            }
        double* Sg = pool[SP_SIG].data() + F.poff;
        for (int c = 0; c < nn; ++c) {
            for (int r = 0; r < nn; ++r) {
                double v = 0.0;
                if (r >= c) {
                    for (int k = r; k < nn; ++k) v += l(k, r) * l(k, c);
                    double s = 0.0;
                    for (int k = 0; k < na; ++k) s += l(nn + k, r) * S21[k + (size_t)c * na];
                    v -= s;
                }
                Sg[r + (int64_t)c * nJ] = v;
            }
            for (int r = 0; r < na; ++r) Sg[nn + r + (int64_t)c * nJ] = S21[r + (size_t)c * na];
        }
    }

    std::string run() {
        cnt.assign(H.fronts.size(), 0);
        for (size_t ti = 0; ti < S.task.size(); ++ti) {
            const SelTask& t = S.task[ti];
            for (int w = 0; w < 2; ++w)
                if (t.wait[w] >= 0 && cnt[t.wait[w]] < t.target[w])
                    return "selinv wait not satisfied in ticket order (ticket " + std::to_string(ti) + ")";
            const Front& F = H.fronts[t.front];
            switch (t.kind) {
                case ST_XDIAG: {
                    const int kb = tile_len(F.nn, t.a);
                    for (int c = 0; c < kb; ++c)
                        for (int r = 0; r < kb; ++r)
                            pool[SP_X][S.xoff[t.front] + (int64_t)NB * t.a + r + (int64_t)(NB * t.a + c) * F.nn] =
                                pool[SP_W][F.woff + (int64_t)NB * t.a + r + (int64_t)c * F.nn];
                    break;
                }
                case ST_GEMM: gemm(t); break;
                case ST_GATHER: gather(t.front, NB * t.a, std::min(F.na, NB * t.a + NB)); break;
                case ST_SMALL: small(t.front); break;
                default: return "selinv task kind";
            }
            ++cnt[t.front];
        }
        for (size_t J = 0; J < H.fronts.size(); ++J)
            if (cnt[J] != S.total[J]) return "selinv front incomplete";
        return std::string();
    }
};

}  // namespace

}  // namespace mf

// Host-only entry (built into libmfsym.so): builds the plan and the
// selected-inversion schedule, then runs the schedule on the host in ticket
// order from a Cholesky factor in the reference layout (uplo = :L): panels
// and inverted 64 x 64 diagonal tiles are formed from Dval/Lval, every task
// runs as the device kernel would (checking its waits and bounds), and the
// selected inverse is written to SDval/SLval in the same layout.  stats: tasks,
// GEMM tiles, small-front tasks, gather tasks, flops, X/T, Lt and M elements,
// maximum depth.
extern "C" int mf_selinv_check(int64_t n, int64_t nfr, const int64_t* resptr, const int64_t* sepptr,
                               const int64_t* septgt, const int64_t* reltgt, const int64_t* chdptr,
                               const int64_t* chdtgt, const int64_t* Dptr, const int64_t* Lptr,
                               const double* Dval, const double* Lval, double* SDval, double* SLval,
                               int64_t* stats, int64_t nstats, char* msg, int64_t msglen) {
    using namespace mf;
    HostPlan H;
    SelPlan S;
    std::string err = build_plan(H, n, nfr, resptr, sepptr, septgt, reltgt, chdptr, chdtgt, Dptr, Lptr,
                                 PlanTuning());
    if (err.empty()) err = build_selinv(H, S);
    if (err.empty() && Dval && Lval && SDval && SLval) {
        HostSel X(H, S);
        X.pool[SP_PANEL].assign(H.panel_elems, 0.0);
        X.pool[SP_W].assign(H.w_elems, 0.0);
        X.pool[SP_X].assign(S.x_elems, 0.0);
        X.pool[SP_T].assign(S.x_elems, 0.0);
        X.pool[SP_LT].assign(S.lt_elems, 0.0);
        X.pool[SP_M].assign(S.m_elems, 0.0);
        X.pool[SP_SIG].assign(H.panel_elems, 0.0);
        // panels from the reference layout (the inverse of the copy-out)
        for (int64_t j = 0; j < nfr; ++j) {
            const OutFront& o = H.out[j];
            const Front& F = H.fronts[o.J];
            const int nJ = F.nn + F.na, cl = o.r0 - F.c0;
            for (int c = 0; c < o.nn; ++c) {
                double* col = X.pool[SP_PANEL].data() + F.poff + (int64_t)(cl + c) * nJ;
                for (int r = c; r < o.nn; ++r) col[cl + r] = Dval[o.Dp + r + (int64_t)c * o.nn];
                for (int i = 0; i < o.na; ++i) col[H.septpos[o.sepoff + i]] = Lval[o.Lp + i + (int64_t)c * o.na];
            }
        }
        // inverted diagonal tiles of the big fronts (lower, zero above)
        for (size_t J = 0; J < H.fronts.size(); ++J) {
            const Front& F = H.fronts[J];
            if (F.woff < 0) continue;
            const int nJ = F.nn + F.na;
            for (int K = 0; K * NB < F.nn; ++K) {
                const int kb = tile_len(F.nn, K);
                const double* Lk = X.pool[SP_PANEL].data() + F.poff + (int64_t)NB * K + (int64_t)NB * K * nJ;
                double* W = X.pool[SP_W].data() + F.woff + (int64_t)NB * K;
                for (int c = 0; c < kb; ++c) {
                    for (int r = 0; r < kb; ++r) W[r + (int64_t)c * F.nn] = 0.0;
                    W[c + (int64_t)c * F.nn] = 1.0 / Lk[c + (int64_t)c * nJ];
                    for (int r = c + 1; r < kb; ++r) {
                        double s = 0.0;
                        for (int k = c; k < r; ++k) s += Lk[r + (int64_t)k * nJ] * W[k + (int64_t)c * F.nn];
                        W[r + (int64_t)c * F.nn] = -s / Lk[r + (int64_t)r * nJ];
                    }
                }
            }
        }
        err = X.run();
        if (err.empty()) {
            for (int64_t j = 0; j < nfr; ++j) {
                const OutFront& o = H.out[j];
                const Front& F = H.fronts[o.J];
                const int nJ = F.nn + F.na, cl = o.r0 - F.c0;
                for (int c = 0; c < o.nn; ++c) {
                    const double* col = X.pool[SP_SIG].data() + F.poff + (int64_t)(cl + c) * nJ;
                    for (int r = 0; r < o.nn; ++r) SDval[o.Dp + r + (int64_t)c * o.nn] = r >= c ? col[cl + r] : 0.0;
                    for (int i = 0; i < o.na; ++i) SLval[o.Lp + i + (int64_t)c * o.na] = col[H.septpos[o.sepoff + i]];
                }
            }
        }
    }
    if (msg && msglen > 0) {
        size_t k = std::min<size_t>(err.size(), (size_t)msglen - 1);
        std::copy(err.begin(), err.begin() + k, msg);
        msg[k] = 0;
    }
    const int64_t v[] = {(int64_t)S.task.size(), S.n_gemm, S.n_small, S.n_gather, (int64_t)S.flops,
                         S.x_elems, S.lt_elems, S.m_elems, S.max_depth};
    for (int64_t i = 0; stats && i < nstats && i < (int64_t)(sizeof(v) / sizeof(v[0])); ++i) stats[i] = v[i];
    return err.empty() ? 0 : -1;
}

// Host-only entry (built into libmfsym.so): the completion's backward-solve
// schedule on a forest of nfr roots (separator na[u], residual nn[u]; a
// big root padded as the library pads it), run on the host in ticket order
// from the Cholesky factor of a random positive definite K per root (seeded)
// with its inverted 64 x 64 diagonal tiles, every wait checked and every
// value read checked to be written; *worst = the largest relative
// difference of the rows the completion reads against L^-T e by plain
// backward substitution.  stats: tasks, flops, Y elements.
extern "C" int mf_cpl_solve_check(int64_t nfr, const int64_t* na, const int64_t* nn, uint64_t seed, double* worst,
                                  int64_t* stats, int64_t nstats, char* msg, int64_t msglen) {
    using namespace mf;
    auto fail_msg = [&](const std::string& m) {
        if (msg && msglen > 0) {
            std::strncpy(msg, m.c_str(), (size_t)msglen - 1);
            msg[msglen - 1] = 0;
        }
        return -1;
    };
    if (nfr <= 0 || !na || !nn || !worst) return fail_msg("bad argument");
    auto padded = [&](int64_t u) { return na[u] + nn[u] > SMALL_NJ ? (na[u] + 63) / 64 * 64 : na[u]; };
    std::vector<int64_t> resptr(nfr + 1, 1), ones(nfr + 2, 1), Dptr(nfr + 1, 1), dummy(1, 1);
    for (int64_t u = 0; u < nfr; ++u) {
        if (na[u] < 0 || nn[u] <= 0) return fail_msg("bad front size");
        const int64_t nj = nn[u] + padded(u);
        resptr[u + 1] = resptr[u] + nj;
        Dptr[u + 1] = Dptr[u] + nj * nj;
    }
    HostPlan H;
    PlanTuning tune;
    tune.factor_only = true;  // as the library plans the completion's forests
    std::string err = build_plan(H, resptr[nfr] - 1, nfr, resptr.data(), ones.data(), dummy.data(), dummy.data(),
                                 ones.data(), dummy.data(), Dptr.data(), ones.data(), tune);
    if (!err.empty()) return fail_msg(err);
    const int32_t NJ = (int32_t)H.fronts.size();
    std::vector<int32_t> pad(NJ, 0), nnr(NJ, 0);
    for (int64_t u = 0; u < nfr; ++u) {
        const int32_t J = H.owner[u];
        pad[J] = (int32_t)padded(u);
        nnr[J] = (int32_t)nn[u];
    }
    SelPlan S;
    err = build_cpl_solve(H, pad, nnr, S);
    if (!err.empty()) return fail_msg(err);
    const double unset = std::nan("");
    std::vector<double> panel(H.panel_elems, 0.0), W(H.w_elems, 0.0), Y(S.x_elems, unset), U(S.x_elems, unset),
        R(S.x_elems, unset), eye(NB * NB, 0.0);
    for (int i = 0; i < NB; ++i) eye[i + NB * i] = 1.0;
    uint64_t st = seed * 6364136223846793005ull + 1442695040888963407ull;
    auto rnd = [&]() {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return (double)(st >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    };
    std::vector<std::vector<double>> Ls(NJ);
    for (int32_t J = 0; J < NJ; ++J) {
        const Front& F = H.fronts[J];
        const int n = F.nn;
        std::vector<double> G((size_t)n * n), L((size_t)n * n, 0.0);
        for (double& g : G) g = rnd();
        for (int j = 0; j < n; ++j)
            for (int i = j; i < n; ++i) {
                double s = i == j ? (double)n : 0.0;
                for (int k = 0; k < n; ++k) s += G[i + (size_t)k * n] * G[j + (size_t)k * n];
                L[i + (size_t)j * n] = s;  // K, lower triangle
            }
        for (int j = 0; j < n; ++j) {  // Cholesky in place
            double d = L[j + (size_t)j * n];
            for (int k = 0; k < j; ++k) d -= L[j + (size_t)k * n] * L[j + (size_t)k * n];
            L[j + (size_t)j * n] = std::sqrt(d);
            for (int i = j + 1; i < n; ++i) {
                double s = L[i + (size_t)j * n];
                for (int k = 0; k < j; ++k) s -= L[i + (size_t)k * n] * L[j + (size_t)k * n];
                L[i + (size_t)j * n] = s / L[j + (size_t)j * n];
            }
        }
        for (int j = 0; j < n; ++j)
            for (int i = j; i < n; ++i) panel[F.poff + i + (size_t)j * n] = L[i + (size_t)j * n];
        if (!H.is_small[J])
            for (int k0 = 0; k0 < n; k0 += NB) {  // W: (row k0 + r, column c) at woff + k0 + r + c n
                const int kb = std::min(NB, n - k0);
                for (int c = 0; c < kb; ++c) {
                    std::vector<double> x(kb, 0.0);
                    x[c] = 1.0;
                    for (int r = 0; r < kb; ++r) {
                        double s = x[r];
                        for (int k = 0; k < r; ++k) s -= L[k0 + r + (size_t)(k0 + k) * n] * x[k];
                        x[r] = s / L[k0 + r + (size_t)(k0 + r) * n];
                    }
                    for (int r = 0; r < kb; ++r) W[F.woff + k0 + r + (size_t)c * n] = x[r];
                }
            }
        Ls[J] = std::move(L);
    }
    double* base[SP_N] = {panel.data(), W.data(), Y.data(), U.data(), eye.data(), R.data(), nullptr};
    std::vector<int32_t> cnt(std::max<int64_t>(S.ncounters, NJ), 0);
    std::vector<double> acc((size_t)NB * NB);
    for (const SelTask& t : S.task) {
        for (int w = 0; w < 2; ++w)
            if (t.wait[w] >= 0 && cnt[t.wait[w]] < t.target[w]) return fail_msg("completion solve: wait not met");
        std::fill(acc.begin(), acc.end(), 0.0);
        for (int q = 0; q < t.np; ++q) {
            const SelPart& p = t.p[q];
            const double* A = base[p.arrA] + p.a_off;
            const double* B = base[p.arrB] + p.b_off;
            for (int c = 0; c < t.n; ++c)
                for (int r = 0; r < t.m; ++r) {
                    double s = 0.0;
                    for (int k = 0; k < p.k; ++k) {
                        const double a = p.ta ? A[k + (size_t)r * p.lda] : A[r + (size_t)k * p.lda];
                        const double b = p.tb ? B[k + (size_t)c * p.ldb] : B[c + (size_t)k * p.ldb];
                        if (std::isnan(a) || std::isnan(b)) return fail_msg("completion solve: reads an unwritten value");
                        s += a * b;
                    }
                    acc[r + (size_t)NB * c] += p.sgn * s;
                }
        }
        double* C = base[t.arrC] + t.c_off;
        for (int c = 0; c < t.n; ++c)
            for (int r = 0; r < t.m; ++r) C[r + (size_t)c * t.ldc] = acc[r + (size_t)NB * c];
        ++cnt[t.front];
    }
    double w = 0.0;
    for (int32_t J = 0; J < NJ; ++J) {
        if (H.is_small[J]) continue;
        const int n = H.fronts[J].nn, p = pad[J];
        const std::vector<double>& L = Ls[J];
        for (int q = 0; q < nnr[J]; ++q) {
            std::vector<double> x(n, 0.0);
            x[p + q] = 1.0;
            for (int i = n - 1; i >= 0; --i) {
                double s = x[i];
                for (int k = i + 1; k < n; ++k) s -= L[k + (size_t)i * n] * x[k];
                x[i] = s / L[i + (size_t)i * n];
            }
            double xm = 0.0;
            for (int i = 0; i <= p + q; ++i) xm = std::max(xm, std::fabs(x[i]));
            for (int i = 0; i <= p + q; ++i)
                w = std::max(w, std::fabs(Y[S.xoff[J] + i + (size_t)q * n] - x[i]) / xm);
        }
    }
    *worst = w;
    const int64_t v[] = {(int64_t)S.task.size(), (int64_t)std::llround(S.flops), S.x_elems};
    for (int64_t i = 0; stats && i < nstats && i < 3; ++i) stats[i] = v[i];
    return 0;
}
