// mf_plan.cpp — host planner of libmfhip.so.
//
// Input: the reference's ChordalSymbolic arrays (1-based Int64,
// src/Multifrontal.jl/src/chordal_symbolic.jl:15-28).  Output: HostPlan.
//   1. internal fronts: chains j -> pnt(j) with pnt(j) = j + 1 (j is the last
//      child, so columns are contiguous) merged while the explicit zeros stay
//      under `zero_frac` of the merged panel (or the panel is tiny);
//   2. small region (whole subtree with nJ <= SMALL_NJ): one persistent
//      launch in ticket order; big region: levels by height;
//   3. memory: panel pool (all fronts), U pool (small region: private slots;
//      big region: slots reused after the parent's extend-add), W pool;
//   4. static task lists of every launch of the factorization and the solves,
//      the dataflow form of the tile-DAG, its tickets by list scheduling, and
//      checks of every task's bounds and of both ticket orders.
// build_plan runs these as the stages of class Planner, one method each.
#include "mf_plan.h"

#include <algorithm>
#include <cmath>
#include <deque>
#include <functional>
#include <map>
#include <numeric>
#include <queue>

namespace mf {

namespace {

// First-fit allocator over a linear address space with coalescing.
struct Allocator {
    std::map<int64_t, int64_t> free_;  // offset -> size
    int64_t top = 0;
    int64_t peak = 0;
    int64_t alloc(int64_t size) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= size) {
                int64_t off = it->first;
                int64_t rest = it->second - size;
                free_.erase(it);
                if (rest > 0) free_[off + size] = rest;
                return off;
            }
        }
        int64_t off = top;
        top += size;
        peak = std::max(peak, top);
        return off;
    }
    void release(int64_t off, int64_t size) {
        auto it = free_.emplace(off, size).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
                it = pv;
            }
        }
        // give back a free block that ends at the top
        auto last = std::prev(free_.end());
        if (last->first + last->second == top) {
            top = last->first;
            free_.erase(last);
        }
    }
};

inline int64_t rup(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

// The planner's stages.  Each stage reads what the earlier ones left in the
// HostPlan (and the few shared values below) and returns an empty string or
// an error message; build_plan runs them in order.
class Planner {
  public:
    Planner(HostPlan& plan, const PlanTuning& tune)
        : P(plan), zero_frac(tune.zero_frac), UPD_GROUP(tune.upd_group), EA_ROWS(tune.ea_rows),
          WIDE_K(tune.wide_k), SIM_WORKERS(tune.sim_workers), FUSE(tune.fuse != 0) {
        P.tune = tune;
    }
    // the reference's arrays, 0-based, and their consistency
    std::string read_input(int64_t n, int64_t nfr, const int64_t* resptr, const int64_t* sepptr,
                           const int64_t* septgt, const int64_t* reltgt, const int64_t* chdptr,
                           const int64_t* chdtgt, const int64_t* Dptr, const int64_t* Lptr);
    std::string amalgamate();     // 1. internal fronts, children, rel maps, copy-out map
    std::string classify();       // 2. small region, levels of the big region
    std::string layout_memory();  // 3. panel, U (slot reuse), W pools, solve workspaces
    std::string factor_tasks();   // 4. extend-add, diagonal and tile-DAG tasks
    std::string solve_tasks();    // gather tables and persistent solve task lists
    std::string check_bounds();   // every task's accesses inside its pools
    std::string dataflow();       // predecessor counts and successor lists of the DAG
    std::string schedule();       // tickets by list scheduling (and the GMW81 order)
    std::string check_order();    // every wait of both orders targets an earlier ticket
    std::string statistics();

  private:
    HostPlan& P;
    const double zero_frac;
    const int32_t UPD_GROUP, EA_ROWS, WIDE_K, SIM_WORKERS;
    const bool FUSE;
    int64_t n = 0, nfr = 0, nsep = 0;  // matrix order, reference fronts, separator entries
    int32_t NJ = 0, nlev = 0;          // internal fronts, levels of the big region
    std::vector<uint8_t> in_small_region;  // per internal front: whole subtree small
    // GMW81: (diagonal-tile task, last update of a tile below it in its block
    // column) edges, in creation-order task numbers
    std::vector<std::pair<int32_t, int32_t>> gmw_edges;
    int64_t nnref(int64_t j) const { return P.resptr[j + 1] - P.resptr[j]; }
    int64_t naref(int64_t j) const { return P.sepptr[j + 1] - P.sepptr[j]; }
};

std::string Planner::read_input(int64_t n, int64_t nfr, const int64_t* resptr, const int64_t* sepptr,
                                const int64_t* septgt, const int64_t* reltgt, const int64_t* chdptr,
                                const int64_t* chdtgt, const int64_t* Dptr, const int64_t* Lptr) {
    (void)reltgt;
    if (n < 0 || nfr < 0) return "negative size";
    if (n > (int64_t)INT32_MAX / 2) return "n too large";
    P.n = n;
    P.nfr = nfr;
    this->n = n;
    this->nfr = nfr;
    P.resptr.resize(nfr + 1);
    P.sepptr.resize(nfr + 1);
    P.Dptr.resize(nfr + 1);
    P.Lptr.resize(nfr + 1);
    for (int64_t j = 0; j <= nfr; ++j) {
        P.resptr[j] = resptr[j] - 1;
        P.sepptr[j] = sepptr[j] - 1;
        P.Dptr[j] = Dptr[j] - 1;
        P.Lptr[j] = Lptr[j] - 1;
    }
    if (P.resptr[0] != 0 || P.resptr[nfr] != n) return "resptr does not partition 1..n";
    nsep = P.sepptr[nfr];
    P.septgt.resize(nsep);
    for (int64_t p = 0; p < nsep; ++p) {
        P.septgt[p] = septgt[p] - 1;
        if (P.septgt[p] < 0 || P.septgt[p] >= n) return "septgt out of range";
    }
    P.ndz = P.Dptr[nfr];
    P.nlz = P.Lptr[nfr];
    // reference parents from the children lists (chd has nfr+2 pointers)
    P.pnt.assign(nfr, -1);
    for (int64_t j = 0; j < nfr; ++j) {
        for (int64_t q = chdptr[j] - 1; q < chdptr[j + 1] - 1; ++q) {
            int64_t c = chdtgt[q] - 1;
            if (c < 0 || c >= nfr || c >= j) return "children must precede parents";
            P.pnt[c] = (int32_t)j;
        }
    }
    for (int64_t j = 0; j < nfr; ++j) {
        if (nnref(j) <= 0) return "empty front";
        if (Dptr[j + 1] - Dptr[j] != nnref(j) * nnref(j)) return "Dptr inconsistent";
        if (Lptr[j + 1] - Lptr[j] != nnref(j) * naref(j)) return "Lptr inconsistent";
        for (int64_t p = P.sepptr[j]; p < P.sepptr[j + 1]; ++p) {
            if (P.septgt[p] < P.resptr[j + 1]) return "separator row inside its front";
            if (p > P.sepptr[j] && P.septgt[p] <= P.septgt[p - 1]) return "separator not sorted";
        }
    }
    return std::string();
}

std::string Planner::amalgamate() {
    // ---- 1. amalgamation --------------------------------------------------
    std::vector<int32_t> gowner(nfr);
    std::vector<int64_t> gcols(nfr), gzeros(nfr), gc0(nfr);
    std::vector<int32_t> gtop(nfr);
    for (int64_t j = 0; j < nfr; ++j) {
        gowner[j] = (int32_t)j;
        gcols[j] = nnref(j);
        gzeros[j] = 0;
        gc0[j] = P.resptr[j];
        gtop[j] = (int32_t)j;
    }
    for (int64_t j = 0; j < nfr; ++j) {
        int64_t p = P.pnt[j];
        if (p != j + 1) continue;
        int32_t G = gowner[j];
        int64_t c1 = gcols[G], c2 = nnref(p), naj = naref(j), nap = naref(p);
        int64_t extra = c1 * (c2 + nap - naj);
        int64_t C = c1 + c2;
        int64_t total = C * (C + 1) / 2 + C * nap;
        int64_t z = gzeros[G] + extra;
        if ((double)z <= zero_frac * (double)total || C <= 8) {
            gowner[p] = G;
            gcols[G] = C;
            gzeros[G] = z;
            gtop[G] = (int32_t)p;
        }
    }
    // internal fronts, ordered by their top member (a postorder)
    std::vector<int32_t> groups;
    for (int64_t j = 0; j < nfr; ++j)
        if (gowner[j] == j) groups.push_back((int32_t)j);
    std::sort(groups.begin(), groups.end(), [&](int32_t a, int32_t b) { return gtop[a] < gtop[b]; });
    NJ = (int32_t)groups.size();
    std::vector<int32_t> gid(nfr, -1);
    for (int32_t J = 0; J < NJ; ++J) gid[groups[J]] = J;
    P.owner.resize(nfr);
    for (int64_t j = 0; j < nfr; ++j) P.owner[j] = gid[gowner[j]];
    P.fronts.resize(NJ);
    P.rows.clear();
    for (int32_t J = 0; J < NJ; ++J) {
        int32_t G = groups[J], top = gtop[G];
        Front& F = P.fronts[J];
        F.c0 = (int32_t)gc0[G];
        F.nn = (int32_t)gcols[G];
        F.na = (int32_t)naref(top);
        F.rowoff = (int32_t)P.rows.size();
        for (int32_t c = 0; c < F.nn; ++c) P.rows.push_back(F.c0 + c);
        for (int64_t p = P.sepptr[top]; p < P.sepptr[top + 1]; ++p)
            P.rows.push_back((int32_t)P.septgt[p]);
        F.parent = P.pnt[top] < 0 ? -1 : P.owner[P.pnt[top]];
        if (F.parent >= 0 && F.parent <= J) return "internal tree not topological";
        P.max_nj = std::max<int64_t>(P.max_nj, F.nn + F.na);
    }
    // children (accumulation order: decreasing index, like the reference's
    // reverse traversal of the children list) and rel maps
    {
        std::vector<std::vector<int32_t>> kids(NJ);
        for (int32_t J = 0; J < NJ; ++J)
            if (P.fronts[J].parent >= 0) kids[P.fronts[J].parent].push_back(J);
        P.chd.clear();
        for (int32_t J = 0; J < NJ; ++J) {
            P.fronts[J].chd0 = (int32_t)P.chd.size();
            for (auto it = kids[J].rbegin(); it != kids[J].rend(); ++it) P.chd.push_back(*it);
            P.fronts[J].chd1 = (int32_t)P.chd.size();
        }
        P.rel.clear();
        for (int32_t K = 0; K < NJ; ++K) {
            Front& F = P.fronts[K];
            F.reloff = (int32_t)P.rel.size();
            if (F.parent < 0) {
                if (F.na != 0) return "root front with update rows";
                continue;
            }
            const Front& Pf = P.fronts[F.parent];
            const int32_t* prow = P.rows.data() + Pf.rowoff;
            int32_t pn = Pf.nn + Pf.na;
            for (int32_t t = 0; t < F.na; ++t) {
                int32_t v = P.rows[F.rowoff + F.nn + t];
                const int32_t* it = std::lower_bound(prow, prow + pn, v);
                if (it == prow + pn || *it != v) return "separator row missing in parent";
                P.rel.push_back((int32_t)(it - prow));
            }
        }
    }
    // copy-out map
    P.out.resize(nfr);
    P.septpos.resize(nsep);
    for (int64_t j = 0; j < nfr; ++j) {
        OutFront& o = P.out[j];
        o.Dp = P.Dptr[j];
        o.Lp = P.Lptr[j];
        o.J = P.owner[j];
        o.r0 = (int32_t)P.resptr[j];
        o.nn = (int32_t)nnref(j);
        o.na = (int32_t)naref(j);
        o.sepoff = (int32_t)P.sepptr[j];
        const Front& F = P.fronts[o.J];
        const int32_t* frow = P.rows.data() + F.rowoff;
        int32_t fn = F.nn + F.na;
        for (int64_t p = P.sepptr[j]; p < P.sepptr[j + 1]; ++p) {
            int32_t v = (int32_t)P.septgt[p];
            const int32_t* it = std::lower_bound(frow, frow + fn, v);
            if (it == frow + fn || *it != v) return "reference separator row missing";
            P.septpos[p] = (int32_t)(it - frow);
        }
    }
    return std::string();
}

std::string Planner::classify() {
    // ---- 2. classification and levels --------------------------------------
    P.is_small.assign(NJ, 0);
    in_small_region.assign(NJ, 0);
    std::vector<int32_t> slev(NJ, 0), blev(NJ, 0);
    for (int32_t J = 0; J < NJ; ++J) {
        const Front& F = P.fronts[J];
        P.is_small[J] = (F.nn + F.na) <= SMALL_NJ;
        bool reg = P.is_small[J];
        int32_t sl = 0, bl = 0;
        for (int32_t q = F.chd0; q < F.chd1; ++q) {
            int32_t K = P.chd[q];
            if (!in_small_region[K]) reg = false;
            sl = std::max(sl, slev[K] + 1);
            if (!in_small_region[K]) bl = std::max(bl, blev[K] + 1);
        }
        in_small_region[J] = reg;
        slev[J] = sl;
        blev[J] = bl;
    }
    P.small_order.clear();
    nlev = 0;
    for (int32_t J = 0; J < NJ; ++J) {
        if (in_small_region[J]) {
            P.small_order.push_back(J);
            P.small_levels = std::max<int64_t>(P.small_levels, slev[J] + 1);
        } else {
            nlev = std::max(nlev, blev[J] + 1);
        }
    }
    std::stable_sort(P.small_order.begin(), P.small_order.end(),
                     [&](int32_t a, int32_t b) { return slev[a] < slev[b]; });
    P.n_small_region = (int64_t)P.small_order.size();
    // tickets of the small region: up to four consecutive fronts of nJ <= TINY_NJ
    // share a ticket (one wavefront each); larger small fronts take one
    P.small_units.clear();
    for (int32_t i = 0; i < (int32_t)P.small_order.size();) {
        const Front& F = P.fronts[P.small_order[i]];
        if (F.nn + F.na > TINY_NJ) {
            P.small_units.push_back(i << 3);
            ++i;
            continue;
        }
        int32_t c = 1;
        while (c < 4 && i + c < (int32_t)P.small_order.size()) {
            const Front& G = P.fronts[P.small_order[i + c]];
            if (G.nn + G.na > TINY_NJ) break;
            ++c;
        }
        P.small_units.push_back((i << 3) | c);
        i += c;
    }
    P.levels.assign(nlev, Level());
    for (int32_t J = 0; J < NJ; ++J) {
        if (in_small_region[J]) continue;
        if (P.is_small[J]) P.levels[blev[J]].small_list.push_back(J);
        else {
            P.levels[blev[J]].big_list.push_back(J);
            P.n_big++;
        }
    }
    return std::string();
}

std::string Planner::layout_memory() {
    // ---- 3. memory -----------------------------------------------------------
    int64_t poff = 0;
    for (int32_t J = 0; J < NJ; ++J) {
        Front& F = P.fronts[J];
        F.poff = poff;
        poff += rup((int64_t)(F.nn + F.na) * F.nn, 32);
        F.uoff = -1;
        F.woff = -1;
    }
    P.panel_elems = poff;
    int64_t uoff = 0;
    for (int32_t J : P.small_order) {
        Front& F = P.fronts[J];
        if (F.na > 0) {
            F.uoff = uoff;
            uoff += rup((int64_t)F.na * F.na, 32);
        }
    }
    int64_t small_u = uoff;
    Allocator A;
    // Last user of every address range of the big U region: a front that
    // reuses the range must wait until the previous user's parent has read it
    // (its extend-add), since the tile-DAG runs levels concurrently.
    std::map<int64_t, std::pair<int64_t, int32_t>> paint;  // begin -> (end, front)
    std::vector<std::vector<int32_t>> uw(NJ);
    auto claim = [&](int32_t J, int64_t b, int64_t e) {
        auto it = paint.upper_bound(b);
        if (it != paint.begin()) --it;
        std::vector<std::pair<int64_t, std::pair<int64_t, int32_t>>> keep;
        while (it != paint.end() && it->first < e) {
            int64_t b0 = it->first, e0 = it->second.first;
            int32_t K = it->second.second;
            if (e0 <= b) { ++it; continue; }
            int32_t c = P.fronts[K].parent;
            if (c >= 0 && std::find(uw[J].begin(), uw[J].end(), c) == uw[J].end()) uw[J].push_back(c);
            if (b0 < b) keep.push_back({b0, {b, K}});
            if (e0 > e) keep.push_back({e, {e0, K}});
            it = paint.erase(it);
        }
        for (auto& kv : keep) paint[kv.first] = kv.second;
        paint[b] = {e, J};
    };
    for (int32_t l = 0; l < nlev; ++l) {
        Level& L = P.levels[l];
        std::vector<int32_t> here(L.small_list);
        here.insert(here.end(), L.big_list.begin(), L.big_list.end());
        for (int32_t J : here) {
            Front& F = P.fronts[J];
            if (F.na > 0) {
                int64_t sz = rup((int64_t)F.na * F.na, 32);
                int64_t o = A.alloc(sz);
                F.uoff = small_u + o;
                claim(J, o, o + sz);
            }
        }
        for (int32_t J : here) {
            const Front& F = P.fronts[J];
            for (int32_t q = F.chd0; q < F.chd1; ++q) {
                int32_t K = P.chd[q];
                const Front& Kf = P.fronts[K];
                if (!in_small_region[K] && Kf.na > 0)
                    A.release(Kf.uoff - small_u, rup((int64_t)Kf.na * Kf.na, 32));
            }
        }
    }
    P.u_elems = small_u + A.peak + 32;
    P.uwait_ptr.assign(NJ + 1, 0);
    P.uwait.clear();
    for (int32_t J = 0; J < NJ; ++J) {
        P.uwait.insert(P.uwait.end(), uw[J].begin(), uw[J].end());
        P.uwait_ptr[J + 1] = (int32_t)P.uwait.size();
    }
    int64_t woff = 0;
    for (int32_t J = 0; J < NJ; ++J) {
        Front& F = P.fronts[J];
        if (!P.is_small[J] && !in_small_region[J]) {
            F.woff = woff;
            woff += rup((int64_t)F.nn * NB, 32);
        }
    }
    P.w_elems = woff + 32;
    P.mvoff1.resize(NJ);
    int64_t mv = 0;
    for (int32_t J = 0; J < NJ; ++J) {
        P.mvoff1[J] = mv;
        mv += P.fronts[J].na;
    }
    // backward separator partials of big fronts with long separators: one
    // 64-row slot per (diagonal block, chunk of SEP_TILES separator tiles)
    P.sboff1.assign(NJ, -1);
    for (int32_t J = 0; J < NJ; ++J) {
        const Front& F = P.fronts[J];
        const int32_t nsep = (F.na + NB - 1) / NB;
        if (P.is_small[J] || nsep <= SEP_TILES) continue;
        const int32_t pk = (F.nn + NB - 1) / NB, nch = (nsep + SEP_TILES - 1) / SEP_TILES;
        P.sboff1[J] = mv;
        mv += (int64_t)pk * nch * NB;
    }
    P.mv_elems1 = mv + 1;
    return std::string();
}

std::string Planner::factor_tasks() {
    // ---- 4. task lists ---------------------------------------------------------
    P.level_small.clear();
    P.level_big.clear();
    P.ea.clear();
    P.ea_src.clear();
    P.wstart.clear();
    P.wsoff.assign(NJ, -1);
    P.gemm.clear();
    P.diag.clear();

    auto add_gemm = [&](int8_t mode, int8_t aA, int64_t ao, int32_t lda, int8_t aB, int64_t bo,
                        int32_t ldb, int8_t aC, int64_t co, int32_t ldc, int32_t m, int32_t nn_,
                        int32_t k, int32_t scol, int8_t lower, int8_t wide = 0) {
        GemmTask t{};
        t.wide = wide;
        t.a_off = ao;
        t.b_off = bo;
        t.c_off = co;
        t.lda = lda;
        t.ldb = ldb;
        t.ldc = ldc;
        t.m = m;
        t.n = nn_;
        t.k = k;
        t.s_col = scol;
        t.arrA = aA;
        t.arrB = aB;
        t.arrC = aC;
        t.mode = mode;
        t.lower = lower;
        P.gemm.push_back(t);
        P.gemm_flops += lower ? (double)m * (m + 1) * k : 2.0 * m * nn_ * k;
    };
    for (int32_t l = 0; l < nlev; ++l) {
        Level& L = P.levels[l];
        L.smallL0 = (int32_t)P.level_small.size();
        P.level_small.insert(P.level_small.end(), L.small_list.begin(), L.small_list.end());
        L.smallL1 = (int32_t)P.level_small.size();
        L.bigL0 = (int32_t)P.level_big.size();
        P.level_big.insert(P.level_big.end(), L.big_list.begin(), L.big_list.end());
        L.bigL1 = (int32_t)P.level_big.size();
        // extend-add tasks
        L.ea0 = (int32_t)P.ea.size();
        for (int32_t J : L.big_list) {
            const Front& F = P.fronts[J];
            int32_t nj = F.nn + F.na;
            std::vector<std::vector<int32_t>> src(nj);
            for (int32_t q = F.chd0; q < F.chd1; ++q) {
                int32_t K = P.chd[q];
                const Front& Kf = P.fronts[K];
                for (int32_t t = 0; t < Kf.na; ++t) {
                    int32_t pc = P.rel[Kf.reloff + t];
                    src[pc].push_back(K);
                    src[pc].push_back(t);
                }
            }
            for (int32_t pc = 0; pc < nj; ++pc) {
                if (pc < F.nn && src[pc].empty()) continue;
                EaTask e{};
                e.front = J;
                e.col = pc;
                e.row0 = pc;
                e.row1 = nj;
                e.src0 = (int32_t)(P.ea_src.size() / 2);
                P.ea_src.insert(P.ea_src.end(), src[pc].begin(), src[pc].end());
                e.src1 = (int32_t)(P.ea_src.size() / 2);
                P.ea.push_back(e);
            }
        }
        L.ea1 = (int32_t)P.ea.size();
        // window tables of the children read by this level's extend-add:
        // wstart[wsoff[K] + w] = first row i of child K with rel_K[i] >= EA_WIN*w
        for (int32_t J : L.big_list) {
            const Front& F = P.fronts[J];
            int32_t nw = (F.nn + F.na + EA_WIN - 1) / EA_WIN;
            for (int32_t q = F.chd0; q < F.chd1; ++q) {
                int32_t K = P.chd[q];
                const Front& Kf = P.fronts[K];
                P.wsoff[K] = (int32_t)P.wstart.size();
                int32_t i = 0;
                for (int32_t w = 0; w <= nw; ++w) {
                    while (i < Kf.na && P.rel[Kf.reloff + i] < w * EA_WIN) ++i;
                    P.wstart.push_back(i);
                }
            }
        }
        // diagonal tiles (the inverse-only pass of solves on given factors)
        for (int32_t J : L.big_list) {
            const Front& F = P.fronts[J];
            for (int32_t k0 = 0; k0 < F.nn; k0 += NB) {
                DiagTask d{};
                d.front = J;
                d.k0 = k0;
                d.kb = std::min(NB, F.nn - k0);
                P.diag.push_back(d);
            }
        }
    }

    // ---- tile-DAG factorization of the big region -----------------------------
    // Row blocks of a big front: panel blocks [64K, 64K+kb_K) for K < p, then
    // update rows [nn + 64m, ...).  Tile (I, K) of the panel carries a counter
    // = number of panel columns applied; K+1 once it is final (factored or
    // solved).  Order per level: small fronts, extend-add, then per panel step
    // K: POTRF(K), TRSM(., K), the look-ahead column updates (., K+1, K), the
    // other updates; finally the update-matrix tiles (SYRK, K = nn).
    {
        P.dag.clear();
        P.dagf.assign(NJ, DagFront{0, 0, 0, 0, 0, 0});
        int64_t toff = 0;
        for (int32_t J : P.small_order) P.dagf[J].n_done = 1;
        auto row0 = [&](const Front& F, int32_t p, int32_t I) { return I < p ? I * NB : F.nn + (I - p) * NB; };
        auto rlen = [&](const Front& F, int32_t p, int32_t I) {
            return I < p ? std::min(NB, F.nn - I * NB) : std::min(NB, F.na - (I - p) * NB);
        };
        for (int32_t l = 0; l < nlev; ++l) {
            const Level& L = P.levels[l];
            for (int32_t J : L.small_list) {
                P.dagf[J].n_done = 1;
                P.dagf[J].n_ea = 1;
                P.dag.push_back(DagTask{DT_SMALL, J, 0, 0, 0, 0});
            }
            for (int32_t J : L.big_list) {
                const Front& F = P.fronts[J];
                DagFront& D = P.dagf[J];
                D.p = (F.nn + NB - 1) / NB;
                D.R = D.p + (F.na + NB - 1) / NB;
                D.toff = (int32_t)toff;
                toff += (int64_t)D.R * D.p;
                D.n_done = 0;  // counted with the update-matrix tiles below
            }
            // extend-add groups: consecutive columns of one front, about
            // EA_ROWS parent rows (each wave of the task takes every fourth
            // column).  A group holds panel columns or update-matrix columns,
            // never both: the panel's first diagonal tile waits only for the
            // panel groups, the update-matrix tiles for all of them.
            for (int32_t e = L.ea0; e < L.ea1;) {
                int32_t J = P.ea[e].front, e1 = e + 1;
                const bool pan = P.ea[e].col < P.fronts[J].nn;
                int64_t rows = P.ea[e].row1 - P.ea[e].row0;
                while (e1 < L.ea1 && e1 - e < 64 && P.ea[e1].front == J && rows < EA_ROWS &&
                       (P.ea[e1].col < P.fronts[J].nn) == pan) {
                    rows += P.ea[e1].row1 - P.ea[e1].row0;
                    ++e1;
                }
                P.dag.push_back(DagTask{DT_EA, J, e, e1, pan ? 1 : 0, 0});
                P.dagf[J].n_ea++;
                if (pan) P.dagf[J].n_eap++;
                e = e1;
            }
            int32_t maxp = 0;
            for (int32_t J : L.big_list) maxp = std::max(maxp, P.dagf[J].p);
            // With fusion (tune.fuse), the diagonal-tile task of step K also runs
            // the panel solve of tile (K+1, K) and the look-ahead update of the
            // next diagonal tile (K+1, K+1): the critical chain
            // POTRF -> TRSM -> update -> POTRF then hands over between workgroups
            // once per step instead of three times.  b / c of the task hold the
            // two products' GemmTask indices (-1: none).
            std::vector<int64_t> potrf_at(NJ, -1);
            for (int32_t K = 0; K < maxp; ++K) {
                for (int32_t J : L.big_list)
                    if (K < P.dagf[J].p) {
                        potrf_at[J] = (int64_t)P.dag.size();
                        P.dag.push_back(DagTask{DT_POTRF, J, K, -1, -1, 0});
                    }
                for (int32_t J : L.big_list) {
                    const Front& F = P.fronts[J];
                    const DagFront& D = P.dagf[J];
                    if (K >= D.p) continue;
                    int32_t nj = F.nn + F.na, k0 = K * NB, kb = rlen(F, D.p, K);
                    for (int32_t I = K + 1; I < D.R; ++I) {
                        int64_t ao = F.poff + row0(F, D.p, I) + (int64_t)k0 * nj;
                        if (FUSE && I == K + 1 && K + 1 < D.p)
                            P.dag[potrf_at[J]].b = (int32_t)P.gemm.size();
                        else
                            P.dag.push_back(DagTask{DT_TRSM, J, (int32_t)P.gemm.size(), I, K, 0});
                        add_gemm(G_TRSM, ARR_PANEL, ao, nj, ARR_W, F.woff + k0, F.nn, ARR_PANEL, ao, nj,
                                 rlen(F, D.p, I), kb, kb, F.c0 + k0, 0);
                    }
                }
                // trailing updates grouped by UPD_GROUP panel steps: tile
                // (I, Jc) receives the columns [K0, K1) in one product once
                // block column K1 - 1 is final.  The last step before Jc (the
                // look-ahead column, which carries the next diagonal tile) is
                // a product of its own, so the critical chain POTRF -> TRSM ->
                // update -> POTRF runs a single-step update; the steps before
                // it go in a partial group one step earlier.  Per step K:
                //   pass 0: Jc = K + 1, columns [K, K + 1)
                //   pass 1: Jc = K + 2, columns [K0, K + 1) (partial group)
                //   pass 2: Jc >= K + 2, columns [K0, K + 1) (complete group)
                const int32_t K0g = (K / UPD_GROUP) * UPD_GROUP;
                const bool full = (K + 1) % UPD_GROUP == 0;
                for (int pass = 0; pass < 3; ++pass)
                    for (int32_t J : L.big_list) {
                        const Front& F = P.fronts[J];
                        const DagFront& D = P.dagf[J];
                        if (K >= D.p) continue;
                        if (pass == 1 && full) continue;   // a complete group: pass 2
                        if (pass == 2 && !full) continue;  // group not complete yet
                        const int32_t K1 = K + 1, K0 = pass == 0 ? K : K0g;
                        const int32_t nj = F.nn + F.na, k0 = K0 * NB, kk = (K1 - K0) * NB;
                        const int32_t jc0 = pass == 0 ? K + 1 : K + 2;
                        const int32_t jc1 = pass == 2 ? D.p : std::min(jc0 + 1, D.p);
                        for (int32_t Jc = jc0; Jc < jc1; ++Jc)
                            for (int32_t I = Jc; I < D.R; ++I) {
                                if (FUSE && pass == 0 && I == Jc && Jc == K + 1 && Jc < D.p)
                                    P.dag[potrf_at[J]].c = (int32_t)P.gemm.size();
                                else
                                    P.dag.push_back(DagTask{DT_UPD, J, (int32_t)P.gemm.size(), I, Jc, K0 | (K1 << 16)});
                                add_gemm(G_SUB, ARR_PANEL, F.poff + row0(F, D.p, I) + (int64_t)k0 * nj, nj,
                                         ARR_PANEL, F.poff + row0(F, D.p, Jc) + (int64_t)k0 * nj, nj,
                                         ARR_PANEL, F.poff + row0(F, D.p, I) + (int64_t)(Jc * NB) * nj, nj,
                                         rlen(F, D.p, I), rlen(F, D.p, Jc), kk, F.c0 + k0, I == Jc ? 1 : 0);
                            }
                    }
            }
            // update-matrix tiles: 128 x 128 (two row blocks by two column
            // blocks) when the k-extent nn is long enough to amortize them
            for (int32_t J : L.big_list) {
                const Front& F = P.fronts[J];
                DagFront& D = P.dagf[J];
                int32_t nj = F.nn + F.na;
                const int32_t step = (F.nn >= WIDE_K && F.na > NB) ? 2 : 1;
                auto urow = [&](int32_t I) { return I >= D.R ? F.na : row0(F, D.p, I) - F.nn; };
                D.n_done = 0;
                for (int32_t Jc = D.p; Jc < D.R; Jc += step)
                    for (int32_t I = Jc; I < D.R; I += step) {
                        const int32_t i2 = std::min(I + step, D.R), j2 = std::min(Jc + step, D.R);
                        const int32_t r = urow(I), c = urow(Jc), m = urow(i2) - r, nc = urow(j2) - c;
                        const int32_t fl = (i2 - I > 1 ? 1 : 0) | (j2 - Jc > 1 ? 2 : 0);
                        P.dag.push_back(DagTask{DT_SYRK, J, (int32_t)P.gemm.size(), I, Jc, fl});
                        add_gemm(G_SUB, ARR_PANEL, F.poff + F.nn + r, nj, ARR_PANEL, F.poff + F.nn + c, nj,
                                 ARR_U, F.uoff + r + (int64_t)c * F.na, F.na, m, nc, F.nn, F.c0,
                                 I == Jc ? 1 : 0, step == 2 ? 1 : 0);
                        D.n_done++;
                    }
            }
        }
        P.dag_counters = toff + 1;
    }
    return std::string();
}

std::string Planner::solve_tasks() {
    // ---- solve gather tables ----------------------------------------------------
    // A solve task of a big front J covers one row block b of J (diagonal
    // blocks [64b, 64b + 64) for b < p, then update rows [nn + 64(b-p), ...)).
    // For each child K of J, sgstart[sgoff[K] + b] is the first row i of K
    // with rel_K[i] >= the first row of block b (rel is increasing), b = 0..R,
    // so the rows of K landing in block b are [tab[b], tab[b+1]).
    P.sgoff.assign(NJ, -1);
    P.sgstart.clear();
    for (int32_t J = 0; J < NJ; ++J) {
        if (P.is_small[J]) continue;
        const Front& F = P.fronts[J];
        const int32_t p = (F.nn + NB - 1) / NB, R = p + (F.na + NB - 1) / NB;
        auto brow = [&](int32_t b) { return b < p ? std::min(b * NB, F.nn) : std::min(F.nn + (b - p) * NB, F.nn + F.na); };
        for (int32_t q = F.chd0; q < F.chd1; ++q) {
            const int32_t K = P.chd[q];
            const Front& Kf = P.fronts[K];
            P.sgoff[K] = (int32_t)P.sgstart.size();
            int32_t i = 0;
            for (int32_t b = 0; b <= R; ++b) {
                const int32_t r0 = b < R ? brow(b) : F.nn + F.na;
                while (i < Kf.na && P.rel[Kf.reloff + i] < r0) ++i;
                P.sgstart.push_back(i);
            }
        }
    }

    for (int32_t J = 0; J < NJ; ++J) {
        if (P.is_small[J]) continue;
        const Front& F = P.fronts[J];
        const int32_t R = (F.nn + NB - 1) / NB + (F.na + NB - 1) / NB;
        for (int32_t q = F.chd0; q < F.chd1; ++q) {
            const int32_t K = P.chd[q], o = P.sgoff[K];
            if (o < 0 || (size_t)o + R + 1 > P.sgstart.size()) return "solve gather table";
            for (int32_t b = 0; b <= R; ++b)
                if (P.sgstart[o + b] < 0 || P.sgstart[o + b] > P.fronts[K].na ||
                    (b > 0 && P.sgstart[o + b] < P.sgstart[o + b - 1]))
                    return "solve gather table range";
        }
    }

    // ---- persistent solve task lists ------------------------------------------
    // Each direction is one ticket list.  Forward tickets are ordered by a
    // host list schedule (SIM_WORKERS workgroups, longest remaining path
    // first), so that a parent front's tasks are handed out as soon as its
    // children finish instead of after every task of the children's level:
    // small front = 1 task; big front = diagonal blocks 0..p-1 (a chain), then
    // update-row tasks (they need every block); every task of a front follows
    // every task of its children.  Backward: small front = 1 task; big front
    // = diagonal blocks p-1..0 (a chain) after every task of its parent.
    P.ftask.clear();
    P.btask.clear();
    P.ntf.assign(NJ, 0);
    P.ntb.assign(NJ, 0);
    {
        auto nblk = [&](int32_t J) { return (P.fronts[J].nn + NB - 1) / NB; };
        auto nrow = [&](int32_t J) { return (P.fronts[J].na + NB - 1) / NB; };
        // estimated task times (us) of the cost model
        const double c_small = 7.5, c_diag = 12.0;
        auto c_rows = [&](int32_t J) { return 3.0 + 1.5 * nblk(J); };
        std::vector<std::vector<int32_t>> kids(NJ);
        for (int32_t J = 0; J < NJ; ++J)
            if (P.fronts[J].parent >= 0) kids[P.fronts[J].parent].push_back(J);
        // list schedule of tasks given as (front, step); step -1 = small front
        struct SimTask {
            int32_t J, step;
            double dur, bl;
        };
        auto schedule = [&](std::vector<SimTask>& tasks, std::vector<int32_t>& indeg,
                            const std::function<void(int32_t, std::vector<int32_t>&)>& succ) {
            std::vector<int32_t> order;
            order.reserve(tasks.size());
            typedef std::pair<double, int32_t> Item;
            std::priority_queue<Item> ready;                                        // by bl
            std::priority_queue<Item, std::vector<Item>, std::greater<Item>> run;  // by end
            for (int32_t t = 0; t < (int32_t)tasks.size(); ++t)
                if (indeg[t] == 0) ready.push({tasks[t].bl, t});
            double now = 0;
            int busy = 0;
            std::vector<int32_t> out;
            while (!ready.empty() || !run.empty()) {
                while (busy < SIM_WORKERS && !ready.empty()) {
                    const int32_t t = ready.top().second;
                    ready.pop();
                    order.push_back(t);
                    run.push({now + tasks[t].dur, t});
                    ++busy;
                }
                if (run.empty()) break;
                now = run.top().first;
                const int32_t t = run.top().second;
                run.pop();
                --busy;
                out.clear();
                succ(t, out);
                for (int32_t u : out)
                    if (--indeg[u] == 0) ready.push({tasks[u].bl, u});
            }
            return order;
        };
        // forward
        {
            std::vector<int32_t> first(NJ), count(NJ);
            std::vector<SimTask> tasks;
            for (int32_t J = 0; J < NJ; ++J) {
                first[J] = (int32_t)tasks.size();
                if (P.is_small[J]) {
                    tasks.push_back({J, -1, c_small, 0});
                } else {
                    for (int32_t k = 0; k < nblk(J); ++k) tasks.push_back({J, k, c_diag, 0});
                    for (int32_t r = 0; r < nrow(J); ++r) tasks.push_back({J, nblk(J) + r, c_rows(J), 0});
                }
                count[J] = (int32_t)tasks.size() - first[J];
            }
            // entry = first task (small front or block 0); bottom levels root first
            std::vector<double> entry_bl(NJ, 0);
            for (int32_t J = NJ - 1; J >= 0; --J) {
                const int32_t par = P.fronts[J].parent;
                const double pb = par >= 0 ? entry_bl[par] : 0.0;
                if (P.is_small[J]) {
                    tasks[first[J]].bl = c_small + pb;
                } else {
                    const int32_t pk = nblk(J), nr = nrow(J);
                    double rows_bl = 0;
                    for (int32_t r = 0; r < nr; ++r) {
                        SimTask& x = tasks[first[J] + pk + r];
                        x.bl = x.dur + pb;
                        rows_bl = std::max(rows_bl, x.bl);
                    }
                    double nxt = std::max(pb, rows_bl);
                    for (int32_t k = pk - 1; k >= 0; --k) {
                        SimTask& x = tasks[first[J] + k];
                        x.bl = x.dur + nxt;
                        nxt = std::max(pb, x.bl);
                    }
                }
                entry_bl[J] = tasks[first[J]].bl;
            }
            std::vector<int32_t> indeg(tasks.size(), 0);
            std::vector<int32_t> left(NJ, 0);  // unfinished tasks of the children
            for (int32_t J = 0; J < NJ; ++J) {
                for (int32_t K : kids[J]) left[J] += count[K];
                if (P.is_small[J]) {
                    indeg[first[J]] = left[J] > 0 ? 1 : 0;
                } else {
                    const int32_t pk = nblk(J);
                    indeg[first[J]] = left[J] > 0 ? 1 : 0;
                    for (int32_t k = 1; k < pk; ++k) indeg[first[J] + k] = 1;
                    for (int32_t r = 0; r < nrow(J); ++r) indeg[first[J] + pk + r] = 1;
                }
            }
            auto succ = [&](int32_t t, std::vector<int32_t>& out) {
                const SimTask& x = tasks[t];
                const int32_t J = x.J;
                if (!P.is_small[J]) {
                    const int32_t pk = nblk(J);
                    if (x.step < pk - 1) out.push_back(first[J] + x.step + 1);
                    if (x.step == pk - 1)
                        for (int32_t r = 0; r < nrow(J); ++r) out.push_back(first[J] + pk + r);
                }
                const int32_t par = P.fronts[J].parent;
                if (par >= 0 && --left[par] == 0) out.push_back(first[par]);
            };
            const std::vector<int32_t> order = schedule(tasks, indeg, succ);
            if (order.size() != tasks.size()) return "solve schedule (forward) incomplete";
            for (int32_t t : order) {
                const SimTask& x = tasks[t];
                const Front& F = P.fronts[x.J];
                const int32_t pk = P.is_small[x.J] ? 0 : nblk(x.J);
                if (x.step < 0) P.ftask.push_back(PTask{PT_SMALL, x.J, 0, 0});
                else if (x.step < pk)
                    P.ftask.push_back(PTask{PT_DIAG, x.J, x.step * NB, std::min(NB, F.nn - x.step * NB)});
                else {
                    const int32_t r0 = F.nn + (x.step - pk) * NB;
                    P.ftask.push_back(PTask{PT_ROWS, x.J, r0, std::min(F.nn + F.na, r0 + NB)});
                }
                P.ntf[x.J]++;
            }
        }
        // backward: the big region by depth (distance from the root): a
        // front's backward tasks all become ready when its parent finishes,
        // so the fronts of one depth are ready together.  Per depth: the
        // small fronts living in the big region, then the split separator
        // products of the depth's big fronts (their diagonal tasks wait for
        // them), then the diagonal chains interleaved by block, last block
        // first.  (By height instead, a shallow front with a short subtree
        // waits behind deeper fronts of larger height; a list schedule like
        // the forward one was measured slower: 7.4 ms against 4.5.)
        std::vector<int32_t> depth(NJ, 0);
        for (int32_t J = NJ - 1; J >= 0; --J)
            if (P.fronts[J].parent >= 0) depth[J] = depth[P.fronts[J].parent] + 1;
        {
            int32_t maxd = 0;
            for (int32_t J = 0; J < NJ; ++J)
                if (!in_small_region[J]) maxd = std::max(maxd, depth[J]);
            std::vector<std::vector<int32_t>> small_at(maxd + 1), big_at(maxd + 1);
            for (int32_t J = NJ - 1; J >= 0; --J) {
                if (in_small_region[J]) continue;
                (P.is_small[J] ? small_at : big_at)[depth[J]].push_back(J);
            }
            for (int32_t d = 0; d <= maxd; ++d) {
                for (int32_t J : small_at[d]) {
                    P.btask.push_back(PTask{PT_SMALL, J, 0, 0});
                    P.ntb[J] = 1;
                }
                for (int32_t J : big_at[d]) {
                    if (P.sboff1[J] < 0) continue;
                    const Front& F = P.fronts[J];
                    const int32_t nsep = (F.na + NB - 1) / NB;
                    for (int32_t k = nblk(J) - 1; k >= 0; --k)
                        for (int32_t t0 = 0; t0 < nsep; t0 += SEP_TILES) {
                            P.btask.push_back(PTask{PT_SEPB, J, k * NB, t0, std::min(nsep, t0 + SEP_TILES)});
                            P.ntb[J]++;
                        }
                }
                int32_t maxp = 0;
                for (int32_t J : big_at[d]) maxp = std::max(maxp, nblk(J));
                for (int32_t t = 0; t < maxp; ++t)
                    for (int32_t J : big_at[d]) {
                        int32_t pk = nblk(J), k = pk - 1 - t;
                        if (k >= 0) {
                            const Front& F = P.fronts[J];
                            P.btask.push_back(PTask{PT_DIAG, J, k * NB, std::min(NB, F.nn - k * NB)});
                            P.ntb[J]++;
                        }
                    }
            }
        }
        // the small region by depth (distance from the root): a front's
        // children follow right after its own depth, however low their
        // subtrees are (height order would put a short subtree's fronts at
        // the end of the list and make them wait for unrelated tickets)
        std::vector<int32_t> border(P.small_order.rbegin(), P.small_order.rend());
        std::stable_sort(border.begin(), border.end(), [&](int32_t a, int32_t b) { return depth[a] < depth[b]; });
        for (int32_t J : border) {
            P.btask.push_back(PTask{PT_SMALL, J, 0, 0});
            P.ntb[J] = 1;
        }
    }

    // consecutive tickets of tiny fronts share one ticket, one wavefront each
    // (a front may wait for another front of the same ticket: that one runs
    // on another wavefront of the workgroup, so it makes progress)
    for (std::vector<PTask>* tl : {&P.ftask, &P.btask}) {
        std::vector<PTask> out;
        out.reserve(tl->size());
        auto tiny = [&](const PTask& t) {
            const Front& F = P.fronts[t.front];
            return t.kind == PT_SMALL && F.nn + F.na <= TINY_NJ;
        };
        for (size_t i = 0; i < tl->size();) {
            if (!tiny((*tl)[i])) {
                out.push_back((*tl)[i]);
                if (out.back().kind != PT_SEPB) out.back().c = 0;
                ++i;
                continue;
            }
            int32_t fr[4] = {-1, -1, -1, -1};
            int n = 0;
            while (n < 4 && i < tl->size() && tiny((*tl)[i])) fr[n++] = (*tl)[i++].front;
            out.push_back(PTask{PT_TINY, fr[0], fr[1], fr[2], fr[3]});
        }
        tl->swap(out);
    }
    return std::string();
}

std::string Planner::check_bounds() {
    // ---- bounds validation of every task against the pools ----------------------
    {
        auto pool = [&](int8_t a) {
            return a == ARR_PANEL ? P.panel_elems : (a == ARR_U ? P.u_elems : P.w_elems);
        };
        auto inside = [&](int8_t a, int64_t off, int32_t rows, int32_t cols, int32_t ld) {
            if (rows <= 0 || cols <= 0) return true;
            if (off < 0 || ld < rows) return false;
            return off + (rows - 1) + (int64_t)(cols - 1) * ld < pool(a);
        };
        for (const GemmTask& t : P.gemm) {
            const int32_t edge = t.wide ? 2 * NB : NB;
            if (t.m <= 0 || t.n <= 0 || t.k <= 0 || t.m > edge || t.n > edge) return "gemm tile shape";
            if (t.wide && (t.mode != G_SUB || t.arrC != ARR_U)) return "wide tile kind";
            if (!inside(t.arrA, t.a_off, t.m, t.k, t.lda) || !inside(t.arrB, t.b_off, t.n, t.k, t.ldb) ||
                !inside(t.arrC, t.c_off, t.m, t.n, t.ldc))
                return "gemm task out of bounds";
            if (t.s_col < 0 || t.s_col + t.k > n || (t.mode == G_TRSM && t.s_col + t.n > n))
                return "gemm scaling out of range";
        }
        for (const PTask& t : P.btask) {
            if (t.kind != PT_SEPB) continue;
            const Front& F = P.fronts[t.front];
            const int32_t nsep = (F.na + NB - 1) / NB, pk = (F.nn + NB - 1) / NB;
            const int32_t nch = (nsep + SEP_TILES - 1) / SEP_TILES;
            if (P.sboff1[t.front] < 0 || t.a < 0 || t.a % NB != 0 || t.a >= F.nn || t.b < 0 ||
                t.b % SEP_TILES != 0 || t.c <= t.b || t.c > nsep || t.c - t.b > SEP_TILES)
                return "backward separator task";
            if (P.sboff1[t.front] + (int64_t)pk * nch * NB > P.mv_elems1) return "separator partials out of bounds";
        }
        for (const DiagTask& t : P.diag) {
            const Front& F = P.fronts[t.front];
            if (t.kb <= 0 || t.kb > NB || t.k0 + t.kb > F.nn || F.woff < 0) return "diag task";
            if (!inside(ARR_W, F.woff + t.k0, t.kb, t.kb, F.nn)) return "diag W out of bounds";
        }
        for (const EaTask& e : P.ea) {
            const Front& F = P.fronts[e.front];
            int32_t nj = F.nn + F.na;
            if (e.col < 0 || e.col >= nj || e.row0 != e.col || e.row1 != nj) return "ea task";
            if (e.col >= F.nn && (F.uoff < 0 || F.uoff + (int64_t)F.na * F.na > P.u_elems))
                return "ea U out of bounds";
            for (int32_t s = e.src0; s < e.src1; ++s) {
                int32_t K = P.ea_src[2 * s], tc = P.ea_src[2 * s + 1];
                const Front& Kf = P.fronts[K];
                if (Kf.parent != e.front || tc < 0 || tc >= Kf.na) return "ea source";
                if (Kf.uoff < 0 || Kf.uoff + (int64_t)Kf.na * Kf.na > P.u_elems) return "ea child U";
                int32_t nw = (nj + EA_WIN - 1) / EA_WIN;
                if (P.wsoff[K] < 0 || (int64_t)P.wsoff[K] + nw >= (int64_t)P.wstart.size())
                    return "ea window table";
                for (int32_t w = 0; w < nw; ++w) {
                    int32_t i0 = P.wstart[P.wsoff[K] + w], i1 = P.wstart[P.wsoff[K] + w + 1];
                    if (i0 < 0 || i1 < i0 || i1 > Kf.na) return "ea window range";
                    for (int32_t i = i0; i < i1; ++i) {
                        int32_t pr = P.rel[Kf.reloff + i];
                        if (pr < w * EA_WIN || pr >= (w + 1) * EA_WIN || pr >= nj) return "ea window row";
                    }
                }
            }
        }
        for (size_t J = 0; J < P.fronts.size(); ++J) {
            const Front& F = P.fronts[J];
            int32_t nj = F.nn + F.na;
            if (F.poff + (int64_t)nj * F.nn > P.panel_elems) return "panel out of bounds";
            if (F.na > 0 && (F.uoff < 0 || F.uoff + (int64_t)F.na * F.na > P.u_elems)) return "U slot";
            if (P.is_small[J] && nj > SMALL_NJ) return "small front too big";
            for (int32_t t = 0; t < F.na && F.parent >= 0; ++t) {
                int32_t r = P.rel[F.reloff + t];
                const Front& Pf = P.fronts[F.parent];
                if (r < 0 || r >= Pf.nn + Pf.na) return "rel out of range";
            }
        }
        for (const std::vector<PTask>* tl : {&P.ftask, &P.btask}) {
            if (P.tune.factor_only) break;  // no solve task lists
            int64_t nsmall = (int64_t)P.small_order.size();
            for (const Level& L : P.levels) nsmall += (int64_t)L.small_list.size();
            int64_t seen = 0;
            for (const PTask& t : *tl) {
                if (t.kind == PT_SMALL) ++seen;
                if (t.kind == PT_TINY)
                    for (int32_t J : {t.front, t.a, t.b, t.c}) seen += J >= 0;
            }
            if (seen != nsmall) return "solve task count";
            for (const PTask& t : *tl) {
                const Front& F = P.fronts[t.front];
                int32_t nj = F.nn + F.na;
                if (t.kind == PT_TINY) {
                    for (int32_t J : {t.front, t.a, t.b, t.c})
                        if (J >= NJ || (J >= 0 && P.fronts[J].nn + P.fronts[J].na > TINY_NJ))
                            return "tiny solve task";
                    if (t.front < 0) return "tiny solve task";
                } else if (t.kind == PT_SMALL) {
                    if (nj > SMALL_NJ) return "small solve task on a big front";
                } else if (t.kind == PT_DIAG) {
                    if (F.woff < 0 || t.a < 0 || t.b <= 0 || t.b > NB || t.a + t.b > F.nn || t.a % NB)
                        return "diag solve task";
                } else if (t.kind == PT_ROWS) {
                    if (t.a < F.nn || t.b > nj || t.b <= t.a || t.b - t.a > NB || (t.a - F.nn) % NB ||
                        (t.b != nj && t.b - t.a != NB))
                        return "rows solve task";
                } else if (t.kind == PT_SEPB) {
                    if (tl != &P.btask) return "separator task outside the backward solve";
                } else {
                    return "solve task kind";
                }
            }
        }
        // every dependency points to an earlier ticket
        if (!P.tune.factor_only) {
            // positions 4 i + slot (slot = wavefront inside a tiny ticket)
            std::vector<int64_t> first(NJ, -1), last(NJ, -1);
            auto mark = [&](const std::vector<PTask>& tl) {
                std::fill(first.begin(), first.end(), -1);
                for (int64_t i = 0; i < (int64_t)tl.size(); ++i) {
                    const PTask& t = tl[i];
                    const int32_t fr[4] = {t.front, t.kind == PT_TINY ? t.a : -1,
                                           t.kind == PT_TINY ? t.b : -1, t.kind == PT_TINY ? t.c : -1};
                    for (int slot = 0; slot < 4; ++slot) {
                        const int32_t J = fr[slot];
                        if (J < 0) continue;
                        if (first[J] < 0) first[J] = 4 * i + slot;
                        last[J] = 4 * i + slot;
                    }
                }
            };
            mark(P.ftask);
            for (int32_t J = 0; J < NJ; ++J)
                for (int32_t q = P.fronts[J].chd0; q < P.fronts[J].chd1; ++q)
                    if (last[P.chd[q]] >= first[J]) return "forward solve order";
            mark(P.btask);
            for (int32_t J = 0; J < NJ; ++J)
                if (P.fronts[J].parent >= 0 && last[P.fronts[J].parent] >= first[J])
                    return "backward solve order";
            // inside a front: the diagonal chain in its direction, update rows
            // after the last diagonal block
            std::vector<int32_t> nextk(NJ, 0), seen_rows(NJ, 0);
            for (const PTask& t : P.ftask) {
                if (t.kind == PT_DIAG) {
                    if (t.a / NB != nextk[t.front] || seen_rows[t.front]) return "forward chain order";
                    nextk[t.front]++;
                } else if (t.kind == PT_ROWS) {
                    if (nextk[t.front] * NB < P.fronts[t.front].nn) return "forward rows before the chain";
                    seen_rows[t.front] = 1;
                }
            }
            for (int32_t J = 0; J < NJ; ++J) nextk[J] = (P.fronts[J].nn + NB - 1) / NB - 1;
            for (const PTask& t : P.btask)
                if (t.kind == PT_DIAG) {
                    if (t.a / NB != nextk[t.front]) return "backward chain order";
                    nextk[t.front]--;
                }
            // a front's separator tasks come before its diagonal tasks (which
            // wait for all of them)
            std::vector<uint8_t> diag_seen(NJ, 0);
            std::vector<int32_t> nsepb(NJ, 0);
            for (const PTask& t : P.btask) {
                if (t.kind == PT_DIAG) diag_seen[t.front] = 1;
                if (t.kind == PT_SEPB) {
                    if (diag_seen[t.front]) return "backward separator task after a diagonal task";
                    nsepb[t.front]++;
                }
            }
            for (int32_t J = 0; J < NJ; ++J) {
                const Front& F = P.fronts[J];
                const int32_t nsep = (F.na + NB - 1) / NB;
                const int32_t want = P.sboff1[J] < 0 ? 0 : ((F.nn + NB - 1) / NB) * ((nsep + SEP_TILES - 1) / SEP_TILES);
                if (nsepb[J] != want) return "backward separator task count";
            }
        }
    }
    return std::string();
}

std::string Planner::dataflow() {
    // ---- dataflow form of the DAG ------------------------------------------------
    // Predecessor edges per tile: POTRF(K) <- last update of (K,K) (or every
    // extend-add task of the front for K = 0); TRSM(I,K) <- POTRF(K), last
    // update of (I,K); UPD(I,Jc,K) <- the final tiles (I,K), (Jc,K) and the
    // previous update of (I,Jc); SYRK <- the final last-column tiles of its row
    // blocks.  A host run of the dataflow checks that every task executes.
    {
        const int32_t NT = (int32_t)P.dag.size();
        std::vector<int32_t> fin(P.dag_counters + 1, -1), lastw(P.dag_counters + 1, -1);
        auto tile = [&](int32_t J, int32_t I, int32_t K) {
            return (int64_t)P.dagf[J].toff + (int64_t)K * P.dagf[J].R + I;
        };
        P.gtask0.assign(NJ, 0);
        P.gtask1.assign(NJ, 0);
        std::vector<uint8_t> in_dag(NJ, 0);
        for (int32_t t = 0; t < NT; ++t) {
            const DagTask& d = P.dag[t];
            in_dag[d.front] = 1;
            if (d.kind == DT_EA || d.kind == DT_SMALL) {
                if (P.gtask1[d.front] == 0) P.gtask0[d.front] = t;
                else if (P.gtask1[d.front] != t) return "dataflow: gate tasks of a front not contiguous";
                P.gtask1[d.front] = t + 1;
            }
        }
        std::vector<std::pair<int32_t, int32_t>> edges;  // (successor, predecessor)
        edges.reserve(4 * (size_t)NT);
        auto edge = [&](int32_t pred, int32_t t) -> bool {
            if (pred < 0 || pred >= t) return false;
            edges.push_back({t, pred});
            return true;
        };
        for (int32_t t = 0; t < NT; ++t) {
            const DagTask& d = P.dag[t];
            const int32_t J = d.front;
            switch (d.kind) {
                case DT_POTRF:
                    if (d.a == 0) {
                        for (int32_t e = P.gtask0[J]; e < P.gtask1[J]; ++e)
                            if (P.dag[e].kind != DT_EA || P.dag[e].c)
                                if (!edge(e, t)) return "dataflow: potrf edge";
                    } else if (!edge(lastw[tile(J, d.a, d.a)], t)) {
                        return "dataflow: potrf without its update";
                    }
                    for (int32_t I = d.a + 1; d.a > 0 && I < P.dagf[J].R; ++I) {
                        const int32_t pr = lastw[tile(J, I, d.a)];
                        if (pr < 0 || pr >= t) return "dataflow: block column update after its diagonal tile";
                        gmw_edges.push_back({t, pr});
                    }
                    fin[tile(J, d.a, d.a)] = t;
                    if (d.b >= 0) {  // fused panel solve of (a+1, a)
                        if (d.a > 0 && !edge(lastw[tile(J, d.a + 1, d.a)], t))
                            return "dataflow: fused trsm without its update";
                        fin[tile(J, d.a + 1, d.a)] = t;
                    }
                    if (d.c >= 0) {  // fused look-ahead update of (a+1, a+1)
                        if (d.b < 0) return "dataflow: fused update without the fused trsm";
                        if (d.a > 0 && !edge(lastw[tile(J, d.a + 1, d.a + 1)], t))
                            return "dataflow: fused update order";
                        lastw[tile(J, d.a + 1, d.a + 1)] = t;
                    }
                    break;
                case DT_TRSM:
                    if (!edge(fin[tile(J, d.c, d.c)], t)) return "dataflow: trsm without potrf";
                    if (d.c > 0 && !edge(lastw[tile(J, d.b, d.c)], t)) return "dataflow: trsm without its update";
                    fin[tile(J, d.b, d.c)] = t;
                    break;
                case DT_UPD: {
                    const int32_t K0 = d.d & 0xffff, K1 = d.d >> 16;
                    if (!edge(fin[tile(J, d.b, K1 - 1)], t) || !edge(fin[tile(J, d.c, K1 - 1)], t))
                        return "dataflow: update without its operands";
                    if (K0 > 0 && !edge(lastw[tile(J, d.b, d.c)], t)) return "dataflow: update order";
                    lastw[tile(J, d.b, d.c)] = t;
                    break;
                }
                case DT_SYRK: {
                    const int32_t p = P.dagf[J].p;
                    const int32_t rows[4] = {d.b, d.c, (d.d & 1) ? d.b + 1 : d.b, (d.d & 2) ? d.c + 1 : d.c};
                    for (int32_t I : rows)
                        if (!edge(fin[tile(J, I, p - 1)], t)) return "dataflow: syrk without its rows";
                    // the front's update-matrix extend-add groups: one event per
                    // front (all groups done) instead of an edge per (group, tile)
                    // pair, which would be quadratic in the front size
                    for (int32_t e = P.gtask0[J]; e < P.gtask1[J]; ++e)
                        if (P.dag[e].kind == DT_EA && !P.dag[e].c && e >= t)
                            return "dataflow: syrk before its extend-add";
                    break;
                }
                default: break;
            }
        }
        std::sort(edges.begin(), edges.end());
        edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
        P.dep0.assign(NT, 0);
        std::vector<int32_t> nsucc(NT + 1, 0);
        for (auto& e : edges) {
            P.dep0[e.first]++;
            nsucc[e.second + 1]++;
        }
        // the update-matrix extend-add event of each front and its waiters
        P.n_eau.assign(NJ, 0);
        P.eau_ptr.assign(NJ + 1, 0);
        for (int32_t t = 0; t < NT; ++t) {
            const DagTask& d = P.dag[t];
            if (d.kind == DT_EA && !d.c) P.n_eau[d.front]++;
        }
        for (int32_t t = 0; t < NT; ++t)
            if (P.dag[t].kind == DT_SYRK && P.n_eau[P.dag[t].front] > 0) P.eau_ptr[P.dag[t].front + 1]++;
        for (int32_t J = 0; J < NJ; ++J) P.eau_ptr[J + 1] += P.eau_ptr[J];
        P.eau_wait.assign(P.eau_ptr[NJ], 0);
        {
            std::vector<int32_t> pos(P.eau_ptr.begin(), P.eau_ptr.end() - 1);
            for (int32_t t = 0; t < NT; ++t)
                if (P.dag[t].kind == DT_SYRK && P.n_eau[P.dag[t].front] > 0) {
                    P.eau_wait[pos[P.dag[t].front]++] = t;
                    P.dep0[t]++;
                }
        }
        if (edges.size() > (size_t)INT32_MAX) return "dataflow: too many edges";
        P.succ_ptr.assign(NT + 1, 0);
        for (int32_t t = 0; t < NT; ++t) P.succ_ptr[t + 1] = P.succ_ptr[t] + nsucc[t + 1];
        P.succ.assign(edges.size(), 0);
        {
            std::vector<int32_t> pos(P.succ_ptr.begin(), P.succ_ptr.end() - 1);
            for (auto& e : edges) P.succ[pos[e.second]++] = e.first;
        }
        // gates
        P.gate0.assign(NJ, 0);
        for (int32_t J = 0; J < NJ; ++J) {
            if (!in_dag[J] || P.gtask1[J] == 0) continue;
            const Front& F = P.fronts[J];
            for (int32_t q = F.chd0; q < F.chd1; ++q)
                if (in_dag[P.chd[q]]) P.gate0[J]++;
            P.gate0[J] += P.uwait_ptr[J + 1] - P.uwait_ptr[J];
        }
        for (int32_t J = 0; J < NJ; ++J)
            for (int32_t q = P.uwait_ptr[J]; q < P.uwait_ptr[J + 1]; ++q)
                if (!in_dag[P.uwait[q]]) return "dataflow: U reuse waits on a front outside the DAG";
        P.uwaiters_ptr.assign(NJ + 1, 0);
        for (int32_t J = 0; J < NJ; ++J)
            for (int32_t q = P.uwait_ptr[J]; q < P.uwait_ptr[J + 1]; ++q) P.uwaiters_ptr[P.uwait[q] + 1]++;
        for (int32_t J = 0; J < NJ; ++J) P.uwaiters_ptr[J + 1] += P.uwaiters_ptr[J];
        P.uwaiters.assign(P.uwait.size(), 0);
        {
            std::vector<int32_t> pos(P.uwaiters_ptr.begin(), P.uwaiters_ptr.end() - 1);
            for (int32_t J = 0; J < NJ; ++J)
                for (int32_t q = P.uwait_ptr[J]; q < P.uwait_ptr[J + 1]; ++q) P.uwaiters[pos[P.uwait[q]]++] = J;
        }
        // priority: the panel's critical path (extend-add, diagonal tiles,
        // panel solves, the look-ahead column) before trailing and update-matrix tiles
        P.hi.assign(NT, 0);
        P.n_hi = P.n_lo = 0;
        for (int32_t t = 0; t < NT; ++t) {
            const DagTask& d = P.dag[t];
            const bool h = d.kind == DT_EA || d.kind == DT_SMALL || d.kind == DT_POTRF ||
                           d.kind == DT_TRSM || (d.kind == DT_UPD && d.c == (d.d >> 16));
            P.hi[t] = h ? 1 : 0;
            (h ? P.n_hi : P.n_lo)++;
        }
        auto gated = [&](int32_t t) { return P.dag[t].kind == DT_EA || P.dag[t].kind == DT_SMALL; };
        P.init_hi.clear();
        P.init_lo.clear();
        for (int32_t t = 0; t < NT; ++t) {
            const bool ready = gated(t) ? P.gate0[P.dag[t].front] == 0 : P.dep0[t] == 0;
            if (ready) (P.hi[t] ? P.init_hi : P.init_lo).push_back(t);
        }
        // host run of the dataflow (the kernel's completion rules)
        std::vector<int32_t> dep(P.dep0), gate(P.gate0), eacnt(NJ, 0), done(NJ, 0), seen(NT, 0),
            eaucnt(NJ, 0);
        std::deque<int32_t> qh(P.init_hi.begin(), P.init_hi.end()), ql(P.init_lo.begin(), P.init_lo.end());
        int64_t executed = 0;
        auto push = [&](int32_t s) { (P.hi[s] ? qh : ql).push_back(s); };
        std::function<void(int32_t)> gate_dec = [&](int32_t G) {
            if (--gate[G] == 0)
                for (int32_t e = P.gtask0[G]; e < P.gtask1[G]; ++e) push(e);
        };
        auto consumed = [&](int32_t c) {
            for (int32_t q = P.uwaiters_ptr[c]; q < P.uwaiters_ptr[c + 1]; ++q) gate_dec(P.uwaiters[q]);
        };
        auto front_done = [&](int32_t J) {
            const int32_t par = P.fronts[J].parent;
            if (par >= 0) {
                if (!in_dag[par] || P.gtask1[par] == 0) return false;
                gate_dec(par);
            }
            return true;
        };
        while (!qh.empty() || !ql.empty()) {
            std::deque<int32_t>& q = qh.empty() ? ql : qh;
            const int32_t t = q.front();
            q.pop_front();
            if (seen[t]++) return "dataflow: task released twice";
            ++executed;
            for (int32_t k = P.succ_ptr[t]; k < P.succ_ptr[t + 1]; ++k)
                if (--dep[P.succ[k]] == 0) push(P.succ[k]);
            const DagTask& d = P.dag[t];
            if (d.kind == DT_EA) {
                if (++eacnt[d.front] == P.dagf[d.front].n_ea) consumed(d.front);
                if (!d.c && ++eaucnt[d.front] == P.n_eau[d.front])
                    for (int32_t q = P.eau_ptr[d.front]; q < P.eau_ptr[d.front + 1]; ++q)
                        if (--dep[P.eau_wait[q]] == 0) push(P.eau_wait[q]);
            } else if (d.kind == DT_SMALL) {
                consumed(d.front);
                if (!front_done(d.front)) return "dataflow: parent of a small front has no gate";
            } else if (d.kind == DT_SYRK) {
                if (++done[d.front] == P.dagf[d.front].n_done && !front_done(d.front))
                    return "dataflow: parent of a front has no gate";
            }
        }
        if (executed != NT) return "dataflow: not every task runs (cycle or missing release)";
        if ((int64_t)P.init_hi.size() > P.n_hi || (int64_t)P.init_lo.size() > P.n_lo) return "dataflow: queue sizes";
    }
    return std::string();
}

std::string Planner::schedule() {
    // ---- ticket order by list scheduling ------------------------------------------
    // The persistent kernel hands out tasks in ticket order and a task waits
    // for its predecessors, so the order decides how much the workgroups wait.
    // Simulate the launch on the host: SIM_WORKERS workgroups, an estimated
    // duration per task, ready tasks taken by longest remaining path to the end
    // (critical path first).  The simulated start order is a topological order
    // (a task starts after its predecessors finish) and becomes the tickets.
    {
        const int32_t NT = (int32_t)P.dag.size();
        std::vector<double> cost(NT), bl(NT, 0.0), bgate(NJ, 0.0), bsyrk(NJ, 0.0);
        for (int32_t t = 0; t < NT; ++t) {
            const DagTask& d = P.dag[t];
            // microseconds per task on one workgroup (measured on MI355X with
            // two workgroups per CU): dispatch + extend-add rows / panel tile /
            // MFMA work at the rate each tile shape reaches
            double c = 5.0;
            switch (d.kind) {
                case DT_EA: {
                    double rows = 0;
                    for (int32_t e = d.a; e < d.b; ++e) rows += P.ea[e].row1 - P.ea[e].row0;
                    c += rows * 0.0053;
                    break;
                }
                case DT_SMALL: c += 20.0; break;
                case DT_POTRF:
                    c += 26.0;
                    for (int32_t gi : {d.b, d.c})
                        if (gi >= 0) {
                            const GemmTask& g = P.gemm[gi];
                            c += 2.0 * g.m * g.n * (double)g.k / 7.5e4;
                        }
                    break;
                default: {
                    const GemmTask& g = P.gemm[d.a];
                    c += 2.0 * g.m * g.n * (double)g.k / (g.wide ? 1.3e5 : 7.5e4);
                    break;
                }
            }
            cost[t] = c;
        }
        // list-schedules the tasks under the successor lists (sp, sc) and
        // predecessor counts d0; the start order goes to order
        auto simulate = [&](const std::vector<int32_t>& sp, const std::vector<int32_t>& sc,
                            const std::vector<int32_t>& d0, std::vector<int32_t>& order,
                            double& makespan) -> const char* {
            std::fill(bl.begin(), bl.end(), 0.0);
            std::fill(bgate.begin(), bgate.end(), 0.0);
            std::fill(bsyrk.begin(), bsyrk.end(), 0.0);
            // longest path to the end, in reverse topological (static) order
            for (int32_t t = NT - 1; t >= 0; --t) {
                const DagTask& d = P.dag[t];
                double m = 0;
                for (int32_t k = sp[t]; k < sp[t + 1]; ++k) m = std::max(m, bl[sc[k]]);
                const int32_t J = d.front;
                if (d.kind == DT_EA && !d.c) m = std::max(m, bsyrk[J]);  // the front's update-matrix tiles
                if (d.kind == DT_SYRK || d.kind == DT_SMALL) {
                    const int32_t par = P.fronts[J].parent;
                    if (par >= 0) m = std::max(m, bgate[par]);
                }
                if (d.kind == DT_EA || d.kind == DT_SMALL)
                    for (int32_t q = P.uwaiters_ptr[J]; q < P.uwaiters_ptr[J + 1]; ++q)
                        m = std::max(m, bgate[P.uwaiters[q]]);
                bl[t] = cost[t] + m;
                if (d.kind == DT_EA || d.kind == DT_SMALL) bgate[J] = std::max(bgate[J], bl[t]);
                if (d.kind == DT_SYRK) bsyrk[J] = std::max(bsyrk[J], bl[t]);
            }
            typedef std::pair<double, int32_t> Item;
            std::priority_queue<Item> ready;                                     // (bl, -t)
            std::priority_queue<Item, std::vector<Item>, std::greater<Item>> run;  // (finish, t)
            std::vector<int32_t> dep(d0), gate(P.gate0), eacnt(NJ, 0), done(NJ, 0), eaucnt(NJ, 0);
            auto push = [&](int32_t s) { ready.push({bl[s], -s}); };
            for (int32_t t : P.init_hi) push(t);
            for (int32_t t : P.init_lo) push(t);
            auto gate_dec = [&](int32_t G) {
                if (--gate[G] == 0)
                    for (int32_t e = P.gtask0[G]; e < P.gtask1[G]; ++e) push(e);
            };
            order.clear();
            order.reserve(NT);
            double now = 0;
            int32_t busy = 0;
            while ((int32_t)order.size() < NT) {
                while (busy < SIM_WORKERS && !ready.empty()) {
                    const int32_t t = -ready.top().second;
                    ready.pop();
                    order.push_back(t);
                    run.push({now + cost[t], t});
                    ++busy;
                }
                if (run.empty()) return "list schedule: no runnable task";
                const Item fin = run.top();
                run.pop();
                now = fin.first;
                --busy;
                const int32_t t = fin.second;
                for (int32_t k = sp[t]; k < sp[t + 1]; ++k)
                    if (--dep[sc[k]] == 0) push(sc[k]);
                const DagTask& d = P.dag[t];
                const int32_t J = d.front;
                bool consumed = false, fdone = false;
                if (d.kind == DT_EA && !d.c && ++eaucnt[J] == P.n_eau[J])
                    for (int32_t q = P.eau_ptr[J]; q < P.eau_ptr[J + 1]; ++q)
                        if (--dep[P.eau_wait[q]] == 0) push(P.eau_wait[q]);
                if (d.kind == DT_EA) consumed = ++eacnt[J] == P.dagf[J].n_ea;
                else if (d.kind == DT_SMALL) consumed = fdone = true;
                else if (d.kind == DT_SYRK) fdone = ++done[J] == P.dagf[J].n_done;
                if (consumed)
                    for (int32_t q = P.uwaiters_ptr[J]; q < P.uwaiters_ptr[J + 1]; ++q) gate_dec(P.uwaiters[q]);
                if (fdone && P.fronts[J].parent >= 0) gate_dec(P.fronts[J].parent);
            }
            makespan = now;
            return nullptr;
        };
        std::vector<int32_t> order, order_gmw;
        if (const char* e = simulate(P.succ_ptr, P.succ, P.dep0, order, P.sim_makespan_us)) return e;
        // GMW81 launches: each diagonal-tile task also waits for the last
        // update of every tile below it in its block column, and takes about
        // 200 us (it reads the block column once per pivot)
        if (!P.tune.factor_only) {
            std::vector<std::pair<int32_t, int32_t>> ge(gmw_edges);  // (successor, predecessor)
            for (int32_t t = 0; t < NT; ++t)
                for (int32_t k = P.succ_ptr[t]; k < P.succ_ptr[t + 1]; ++k) ge.push_back({P.succ[k], t});
            std::sort(ge.begin(), ge.end(), [](const std::pair<int32_t, int32_t>& x,
                                               const std::pair<int32_t, int32_t>& y) {
                return x.second != y.second ? x.second < y.second : x.first < y.first;
            });
            ge.erase(std::unique(ge.begin(), ge.end()), ge.end());
            // predecessor counts: the block-column edges that are not already
            // edges (the fused panel solve's update is one), plus the events
            // that P.dep0 counts besides its edges
            std::vector<int32_t> sp2(NT + 1, 0), sc2(ge.size()), d2(P.dep0);
            for (int32_t t = 0; t < NT; ++t)
                for (int32_t k = P.succ_ptr[t]; k < P.succ_ptr[t + 1]; ++k) d2[P.succ[k]]--;
            for (auto& e : ge) {
                sp2[e.second + 1]++;
                d2[e.first]++;
            }
            for (int32_t t = 0; t < NT; ++t) sp2[t + 1] += sp2[t];
            for (size_t k = 0; k < ge.size(); ++k) sc2[k] = ge[k].first;
            const std::vector<double> cost0(cost);
            for (int32_t t = 0; t < NT; ++t)
                if (P.dag[t].kind == DT_POTRF) cost[t] += 175.0;
            double ms = 0;
            if (const char* e = simulate(sp2, sc2, d2, order_gmw, ms)) return std::string("gmw81: ") + e;
            cost = cost0;
        }
        std::vector<DagTask> dag2(NT), dag3(order_gmw.size());
        for (int32_t i = 0; i < NT; ++i) dag2[i] = P.dag[order[i]];
        for (size_t i = 0; i < order_gmw.size(); ++i) dag3[i] = P.dag[order_gmw[i]];
        P.dag.swap(dag2);
        P.dag_gmw.swap(dag3);
    }
    return std::string();
}

std::string Planner::check_order() {
    // ---- DAG order validation: every wait targets an earlier ticket ------------
    // Run on the list-scheduled tickets and on the creation order that GMW81
    // factorizations use (gmw: a diagonal-tile task also waits until every
    // tile of its block column is updated through the previous column).
    auto validate_dag = [&](const std::vector<DagTask>& dg, bool gmw) -> const char* {
        const int32_t NT = (int32_t)dg.size();
        std::vector<int32_t> last_ea(NJ, -1), last_pea(NJ, -1), first_panel(NJ, INT32_MAX), last_done(NJ, -1);
        std::vector<int32_t> n_ea(NJ, 0), n_eap(NJ, 0);
        std::vector<int32_t> setter(P.dag_counters + 1, -1);  // ticket that finalizes (I, K) at value K+1
        std::map<int64_t, int32_t> upd_at;                      // (tile, K) -> ticket of UPD
        auto cnt_id = [&](int32_t J, int32_t I, int32_t K) {
            return (int64_t)P.dagf[J].toff + (int64_t)K * P.dagf[J].R + I;
        };
        for (int32_t t = 0; t < NT; ++t) {
            const DagTask& d = dg[t];
            const DagFront& D = P.dagf[d.front];
            switch (d.kind) {
                case DT_EA:
                    last_ea[d.front] = t;
                    n_ea[d.front]++;
                    if (d.c) {
                        last_pea[d.front] = t;
                        n_eap[d.front]++;
                    }
                    break;
                case DT_SMALL: last_done[d.front] = t; break;
                case DT_POTRF:
                    setter[cnt_id(d.front, d.a, d.a)] = t;
                    first_panel[d.front] = std::min(first_panel[d.front], t);
                    if (d.b >= 0) setter[cnt_id(d.front, d.a + 1, d.a)] = t;
                    if (d.c >= 0) upd_at[cnt_id(d.front, d.a + 1, d.a + 1) * 4096 + d.a] = t;
                    break;
                case DT_TRSM: setter[cnt_id(d.front, d.b, d.c)] = t; break;
                case DT_UPD: upd_at[cnt_id(d.front, d.b, d.c) * 4096 + (d.d >> 16) - 1] = t; break;
                case DT_SYRK: last_done[d.front] = t; break;
                default: return "dag task kind";
            }
            if (d.kind >= DT_POTRF && (d.front < 0 || D.p <= 0)) return "dag task on a front without panel";
        }
        for (int32_t J = 0; J < NJ; ++J)
            if (n_ea[J] > 0 && (n_ea[J] != P.dagf[J].n_ea || n_eap[J] != P.dagf[J].n_eap))
                return "dag: extend-add counts";
        std::vector<int32_t> last_consume(NJ, -1);
        for (int32_t t = 0; t < NT; ++t)
            if (dg[t].kind == DT_EA || dg[t].kind == DT_SMALL) last_consume[dg[t].front] = t;
        for (int32_t t = 0; t < NT; ++t) {
            const DagTask& d = dg[t];
            const Front& F = P.fronts[d.front];
            if (d.kind == DT_EA || d.kind == DT_SMALL) {
                for (int32_t q = P.uwait_ptr[d.front]; q < P.uwait_ptr[d.front + 1]; ++q) {
                    int32_t c = P.uwait[q];
                    if (c == d.front || last_consume[c] >= t) return "dag: U reuse before the previous user is consumed";
                    if (P.dagf[c].n_ea < 1) return "dag: consumer without extend-add";
                }
                for (int32_t q = F.chd0; q < F.chd1; ++q) {
                    int32_t K = P.chd[q];
                    if (P.dagf[K].n_done > 0 && last_done[K] >= t && P.dagf[K].n_done != 1) return "dag: child after parent";
                    if (last_done[K] >= t) return "dag: child completes after parent extend-add";
                }
                continue;
            }
            if (last_pea[d.front] >= t) return "dag: panel task before extend-add";
            if (d.kind == DT_SYRK && last_ea[d.front] >= t) return "dag: update-matrix tile before extend-add";
            auto need = [&](int32_t I, int32_t K) { return setter[cnt_id(d.front, I, K)]; };
            if (d.kind == DT_POTRF && d.a > 0) {
                auto it = upd_at.find(cnt_id(d.front, d.a, d.a) * 4096 + d.a - 1);
                if (it == upd_at.end() || it->second >= t) return "dag: potrf before its update";
                for (int32_t I = d.a + 1; gmw && I < P.dagf[d.front].R; ++I) {
                    auto ig = upd_at.find(cnt_id(d.front, I, d.a) * 4096 + d.a - 1);
                    if (ig == upd_at.end() || ig->second >= t) return "dag: potrf before its block column";
                }
                // the fused products: their tiles carry the updates through a - 1
                if (d.b >= 0) {
                    auto it2 = upd_at.find(cnt_id(d.front, d.a + 1, d.a) * 4096 + d.a - 1);
                    if (it2 == upd_at.end() || it2->second >= t) return "dag: fused trsm before its update";
                }
                if (d.c >= 0) {
                    auto it3 = upd_at.find(cnt_id(d.front, d.a + 1, d.a + 1) * 4096 + d.a - 1);
                    if (it3 == upd_at.end() || it3->second >= t) return "dag: fused update order";
                }
            }
            if (d.kind == DT_TRSM) {
                if (need(d.c, d.c) < 0 || need(d.c, d.c) >= t) return "dag: trsm before potrf";
                if (d.c > 0) {
                    auto it = upd_at.find(cnt_id(d.front, d.b, d.c) * 4096 + d.c - 1);
                    if (it == upd_at.end() || it->second >= t) return "dag: trsm before its update";
                }
            }
            if (d.kind == DT_UPD) {
                const int32_t K0 = d.d & 0xffff, K1 = d.d >> 16;
                if (K1 <= K0 || K1 > d.c || K1 - K0 > UPD_GROUP) return "dag: update group";
                if (need(d.b, K1 - 1) < 0 || need(d.b, K1 - 1) >= t || need(d.c, K1 - 1) < 0 ||
                    need(d.c, K1 - 1) >= t)
                    return "dag: update before its operands";
                if (K0 > 0) {
                    auto it = upd_at.find(cnt_id(d.front, d.b, d.c) * 4096 + K0 - 1);
                    if (it == upd_at.end() || it->second >= t) return "dag: update order";
                }
            }
            if (d.kind == DT_SYRK) {
                int32_t p = P.dagf[d.front].p, R = P.dagf[d.front].R;
                int32_t rows[4] = {d.b, d.c, (d.d & 1) ? d.b + 1 : d.b, (d.d & 2) ? d.c + 1 : d.c};
                for (int32_t I : rows)
                    if (I >= R || need(I, p - 1) < 0 || need(I, p - 1) >= t) return "dag: syrk before its rows";
                const GemmTask& gt = P.gemm[d.a];
                if ((gt.m > NB) != ((d.d & 1) != 0) || (gt.n > NB) != ((d.d & 2) != 0))
                    return "dag: syrk tile and its row blocks disagree";
            }
        }
        return nullptr;
    };
    if (const char* e = validate_dag(P.dag, false)) return e;
    if (!P.tune.factor_only)
        if (const char* e = validate_dag(P.dag_gmw, true)) return std::string("gmw81 order: ") + e;
    return std::string();
}

std::string Planner::statistics() {
    // ---- statistics ------------------------------------------------------------
    for (int64_t j = 0; j < nfr; ++j) {
        double a = (double)nnref(j), b = (double)naref(j);
        P.ref_flops += a * a * a / 3.0 + b * a * a + b * b * a;
    }
    for (int32_t J = 0; J < NJ; ++J) {
        double a = P.fronts[J].nn, b = P.fronts[J].na;
        P.engine_flops += a * a * a / 3.0 + b * a * a + b * b * a;
    }
    return std::string();
}

}  // namespace

std::string build_plan(HostPlan& P, int64_t n, int64_t nfr, const int64_t* resptr,
                       const int64_t* sepptr, const int64_t* septgt, const int64_t* reltgt,
                       const int64_t* chdptr, const int64_t* chdtgt, const int64_t* Dptr,
                       const int64_t* Lptr, const PlanTuning& tune) {
    Planner pl(P, tune);
    std::string e = pl.read_input(n, nfr, resptr, sepptr, septgt, reltgt, chdptr, chdtgt, Dptr, Lptr);
    if (e.empty()) e = pl.amalgamate();
    if (e.empty()) e = pl.classify();
    if (e.empty()) e = pl.layout_memory();
    if (e.empty()) e = pl.factor_tasks();
    if (e.empty() && !tune.factor_only) e = pl.solve_tasks();
    if (e.empty()) e = pl.check_bounds();
    if (e.empty()) e = pl.dataflow();
    if (e.empty()) e = pl.schedule();
    if (e.empty()) e = pl.check_order();
    if (e.empty()) e = pl.statistics();
    return e;
}

std::string tuning_from(const int64_t* v, int64_t nv, PlanTuning* t) {
    auto get = [&](int64_t i, int64_t def) { return (v && i < nv && v[i] > 0) ? v[i] : def; };
    const int64_t ug = get(0, t->upd_group), er = get(1, t->ea_rows), wk = get(2, t->wide_k),
                  sw = get(3, t->sim_workers), zf = get(4, (int64_t)llround(t->zero_frac * 1e6));
    const int64_t pr = (v && nv > 5 && v[5] >= 0) ? v[5] : t->prio;
    const int64_t fu = (v && nv > 6 && v[6] >= 0) ? v[6] : t->fuse;
    const int64_t wc = (v && nv > 7 && v[7] >= 0) ? v[7] : t->wg_per_cu;
    if (ug < 1 || ug > 64) return "upd_group out of range (1..64)";
    if (er < EA_WIN || er > (1 << 24)) return "ea_rows out of range";
    if (wk < 1 || wk > (1 << 20)) return "wide_k out of range";
    if (sw < 1 || sw > 65536) return "sim_workers out of range";
    if (zf > 1000000) return "zero fraction above 1";
    if (pr < 0 || pr >= (1 << (DT_KINDS + 1))) return "priority mask out of range";
    if (fu != 0 && fu != 1) return "fuse must be 0 or 1";
    if (wc < 0 || wc > 16) return "workgroups per CU out of range (0..16)";
    t->upd_group = (int32_t)ug;
    t->ea_rows = (int32_t)er;
    t->wide_k = (int32_t)wk;
    t->sim_workers = (int32_t)sw;
    t->zero_frac = (double)zf * 1e-6;
    t->prio = (int32_t)pr;
    t->fuse = (int32_t)fu;
    t->wg_per_cu = (int32_t)wc;
    return std::string();
}

}  // namespace mf

// Host-only entry (built into libmfsym.so): runs the planner with its bounds
// validation, without any device, and reports a few sizes.
extern "C" int mf_plan_check(int64_t n, int64_t nfr, const int64_t* resptr, const int64_t* sepptr,
                             const int64_t* septgt, const int64_t* reltgt, const int64_t* chdptr,
                             const int64_t* chdtgt, const int64_t* Dptr, const int64_t* Lptr,
                             int64_t* stats, int64_t nstats, char* msg, int64_t msglen) {
    mf::HostPlan P;
    std::string err = mf::build_plan(P, n, nfr, resptr, sepptr, septgt, reltgt, chdptr, chdtgt,
                                     Dptr, Lptr, mf::PlanTuning());
    if (msg && msglen > 0) {
        size_t k = std::min<size_t>(err.size(), (size_t)msglen - 1);
        std::copy(err.begin(), err.begin() + k, msg);
        msg[k] = 0;
    }
    int64_t ea_u = 0, ea_p = 0, ea_src_rows = 0;
    for (const mf::EaTask& e : P.ea) {
        const mf::Front& F = P.fronts[e.front];
        (e.col >= F.nn ? ea_u : ea_p) += e.row1 - e.row0;
        for (int32_t q = e.src0; q < e.src1; ++q)
            ea_src_rows += P.fronts[P.ea_src[2 * q]].na - P.ea_src[2 * q + 1];
    }
    int64_t v[] = {(int64_t)P.fronts.size(), P.n_small_region, P.n_big, (int64_t)P.levels.size(),
                   P.panel_elems, P.u_elems, P.w_elems, (int64_t)P.gemm.size(),
                   (int64_t)P.ea.size(), (int64_t)(P.ftask.size() + P.btask.size()), (int64_t)P.ref_flops,
                   (int64_t)P.engine_flops, P.max_nj, (int64_t)P.dag.size(), P.dag_counters,
                   (int64_t)P.sim_makespan_us, ea_u, ea_p, ea_src_rows};
    for (int64_t i = 0; stats && i < nstats && i < (int64_t)(sizeof(v) / sizeof(v[0])); ++i)
        stats[i] = v[i];
    return err.empty() ? 0 : -1;
}
