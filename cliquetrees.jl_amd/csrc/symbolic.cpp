// symbolic.cpp — host symbolic analysis (libmfsym.so), C++ restatement of the
// CliqueTrees.jl symbolic layer that feeds the multifrontal numeric phase.
//
// Pipeline (reference paths relative to the CliqueTrees.jl tree):
//   1. permuted "upper" graph      sympermute!_impl!  src/bipartite_graphs.jl:275-330
//   2. elimination tree            etree_impl!        src/parent.jl:54-90 (Liu, path compression)
//   3. column counts of L          supcnt_impl!       src/parent.jl:201-297 (the counts are
//                                  unique; computed here by row-subtree traversal)
//   4. supernode partition         stree_impl!        src/supernode_types.jl:41-159
//                                  (Maximal / Fundamental / Nodal, same child scan order)
//   5. postorder of supernode tree postorder!_impl!   src/parent.jl:420-458,527-537,575-598
//   6. vertex relabelling          supernodetree_impl! src/supernode_trees.jl:147-174
//   7. separators (sorted)         cliquetree_impl!   src/clique_trees.jl:146-234
//   8. rel maps, Dptr/Lptr, maxima ChordalSymbolic(res, sep, tree)
//                                  src/Multifrontal.jl/src/chordal_symbolic.jl:302-378
// Internally everything is 0-based; mfs_export converts to the reference's
// 1-based Int64 arrays.
#include "../../include/mfsym.h"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

using i64 = int64_t;

struct mfs_result {
    i64 n = 0, nfr = 0;
    std::vector<i64> perm;                    // final position -> original vertex
    std::vector<i64> invp;                    // original vertex -> final position
    std::vector<i64> resptr, sepptr, septgt, reltgt;
    std::vector<i64> chdptr, chdtgt, pnt, idx, Dptr, Lptr;
    i64 nMptr = 1, nMval = 0, nNval = 0, nFval = 0;
};

namespace {

// Full symmetric adjacency without the diagonal, deduplicated (0-based).
bool build_adjacency(i64 n, const i64* colptr, const i64* rowval,
                     std::vector<i64>& ptr, std::vector<i64>& tgt) {
    if (n < 0 || colptr[0] != 1) return false;
    i64 nnz = colptr[n] - 1;
    std::vector<i64> deg(n + 1, 0);
    for (i64 j = 0; j < n; ++j) {
        if (colptr[j + 1] < colptr[j]) return false;
        for (i64 p = colptr[j] - 1; p < colptr[j + 1] - 1; ++p) {
            i64 i = rowval[p] - 1;
            if (i < 0 || i >= n) return false;
            if (i == j) continue;
            deg[i]++;
            deg[j]++;
        }
    }
    (void)nnz;
    ptr.assign(n + 1, 0);
    for (i64 v = 0; v < n; ++v) ptr[v + 1] = ptr[v] + deg[v];
    tgt.assign(ptr[n], 0);
    std::vector<i64> pos(ptr.begin(), ptr.end() - 1);
    for (i64 j = 0; j < n; ++j) {
        for (i64 p = colptr[j] - 1; p < colptr[j + 1] - 1; ++p) {
            i64 i = rowval[p] - 1;
            if (i == j) continue;
            tgt[pos[i]++] = j;
            tgt[pos[j]++] = i;
        }
    }
    // sort + dedup each list
    std::vector<i64> nptr(n + 1, 0);
    i64 w = 0;
    for (i64 v = 0; v < n; ++v) {
        i64 a = ptr[v], b = ptr[v + 1];
        std::sort(tgt.begin() + a, tgt.begin() + b);
        nptr[v] = w;
        i64 last = -1;
        for (i64 p = a; p < b; ++p) {
            if (tgt[p] != last) { tgt[w++] = tgt[p]; last = tgt[p]; }
        }
    }
    nptr[n] = w;
    tgt.resize(w);
    ptr.swap(nptr);
    return true;
}

}  // namespace

extern "C" int mfs_symbolic(int64_t n, const int64_t* colptr, const int64_t* rowval,
                            const int64_t* order, int snd_kind, mfs_result** out) {
    if (!out || n < 0 || (n > 0 && (!colptr || !order))) return -1;
    if (snd_kind < 0 || snd_kind > 2) return -2;
    *out = nullptr;
    mfs_result* R = new (std::nothrow) mfs_result();
    if (!R) return -3;
    R->n = n;

    std::vector<i64> aptr, atgt;
    if (!build_adjacency(n, colptr, rowval, aptr, atgt)) { delete R; return -4; }

    // order / index (0-based): order0[k] = vertex eliminated k-th
    std::vector<i64> order0(n), index(n, -1);
    for (i64 k = 0; k < n; ++k) {
        i64 v = order[k] - 1;
        if (v < 0 || v >= n || index[v] != -1) { delete R; return -5; }
        order0[k] = v;
        index[v] = k;
    }

    // 1. permuted upper graph: for permuted column i, rows r < i.
    std::vector<i64> uptr(n + 1, 0), utgt;
    for (i64 i = 0; i < n; ++i) {
        i64 v = order0[i];
        for (i64 p = aptr[v]; p < aptr[v + 1]; ++p)
            if (index[atgt[p]] < i) uptr[i + 1]++;
    }
    for (i64 i = 0; i < n; ++i) uptr[i + 1] += uptr[i];
    utgt.resize(uptr[n]);
    {
        std::vector<i64> pos(uptr.begin(), uptr.end() - 1);
        for (i64 i = 0; i < n; ++i) {
            i64 v = order0[i];
            for (i64 p = aptr[v]; p < aptr[v + 1]; ++p) {
                i64 r = index[atgt[p]];
                if (r < i) utgt[pos[i]++] = r;
            }
        }
    }

    // 2. elimination tree (Liu, path compression): parent[i] = -1 for roots.
    std::vector<i64> parent(n, -1), ancestor(n, -1);
    for (i64 i = 0; i < n; ++i) {
        for (i64 p = uptr[i]; p < uptr[i + 1]; ++p) {
            i64 r = utgt[p];
            while (ancestor[r] != -1 && ancestor[r] != i) {
                i64 t = ancestor[r];
                ancestor[r] = i;
                r = t;
            }
            if (ancestor[r] == -1) {
                ancestor[r] = i;
                parent[r] = i;
            }
        }
    }

    // 3. column counts (diagonal included) by row-subtree traversal:
    //    row i of L has a nonzero in column k iff k lies on the etree path
    //    from some r in upper(i) towards i.
    std::vector<i64> colcount(n, 1), mark(n, -1);
    for (i64 i = 0; i < n; ++i) {
        mark[i] = i;
        for (i64 p = uptr[i]; p < uptr[i + 1]; ++p) {
            i64 k = utgt[p];
            while (k != -1 && mark[k] != i) {
                colcount[k]++;
                mark[k] = i;
                k = parent[k];
            }
        }
    }

    // children of the etree in increasing order (reverse!_impl!, parent.jl:151-193)
    std::vector<i64> cptr(n + 1, 0), ctgt(n > 0 ? n : 0);
    for (i64 i = 0; i < n; ++i)
        if (parent[i] != -1) cptr[parent[i] + 1]++;
    for (i64 i = 0; i < n; ++i) cptr[i + 1] += cptr[i];
    {
        std::vector<i64> pos(cptr.begin(), cptr.end() - 1);
        for (i64 i = 0; i < n; ++i)
            if (parent[i] != -1) ctgt[pos[parent[i]]++] = i;
    }

    // 4. supernode partition (stree_impl!): nic = supernode of a vertex,
    //    newv[s] = first vertex of s, sprnt/sanc = supernode parent and the
    //    vertex of the parent supernode that s attaches to (-1: none).
    std::vector<i64> nic(n, -1), newv(n, -1), sprnt(n, -1), sanc(n, -1);
    i64 nsnd = 0;
    if (snd_kind == 1) {  // Nodal
        for (i64 i = 0; i < n; ++i) {
            newv[i] = i;
            nic[i] = i;
            sprnt[i] = parent[i];
            sanc[i] = parent[i];
        }
        nsnd = n;
    } else {
        for (i64 v = 0; v < n; ++v) {
            i64 u = -1;
            if (snd_kind == 0) {  // Maximal: first child with colcount = colcount(v)+1
                for (i64 p = cptr[v]; p < cptr[v + 1]; ++p) {
                    i64 s = ctgt[p];
                    if (colcount[s] == colcount[v] + 1) { u = s; break; }
                }
            } else {  // Fundamental: a single child with colcount = colcount(v)+1
                if (cptr[v + 1] - cptr[v] == 1) {
                    i64 s = ctgt[cptr[v]];
                    if (colcount[s] == colcount[v] + 1) u = s;
                }
            }
            if (u != -1) {
                nic[v] = nic[u];
                if (snd_kind == 0) {
                    for (i64 p = cptr[v]; p < cptr[v + 1]; ++p) {
                        i64 s = ctgt[p];
                        if (s != u) { sprnt[nic[s]] = nic[v]; sanc[nic[s]] = v; }
                    }
                }
            } else {
                nic[v] = nsnd;
                newv[nsnd] = v;
                sprnt[nsnd] = -1;
                sanc[nsnd] = -1;
                nsnd++;
                for (i64 p = cptr[v]; p < cptr[v + 1]; ++p) {
                    i64 s = ctgt[p];
                    sprnt[nic[s]] = nic[v];
                    sanc[nic[s]] = v;
                }
            }
        }
    }

    // 5. postorder of the supernode forest (postorder_impl!, parent.jl:420-458)
    std::vector<i64> child(nsnd, -1), brother(nsnd, -1), post(nsnd, -1), stack;
    stack.reserve(nsnd);
    i64 root = -1;
    for (i64 i = nsnd - 1; i >= 0; --i) {
        i64 j = sprnt[i];
        if (j == -1) { brother[i] = root; root = i; }
        else { brother[i] = child[j]; child[j] = i; }
    }
    for (i64 r = root; r != -1; r = brother[r]) stack.push_back(r);
    for (i64 i = 0; i < nsnd; ++i) {
        i64 j = stack.back();
        stack.pop_back();
        while (child[j] != -1) {
            stack.push_back(j);
            i64 c = child[j];
            child[j] = brother[c];
            j = c;
        }
        post[j] = i;
    }
    std::vector<i64> fpnt(nsnd, -1), elmorder(nsnd);  // parent in new labels
    for (i64 s = 0; s < nsnd; ++s) {
        fpnt[post[s]] = (sprnt[s] == -1) ? -1 : post[sprnt[s]];
        elmorder[post[s]] = s;
    }

    // 6. relabel vertices supernode by supernode (supernode_trees.jl:154-166)
    std::vector<i64> sndptr(nsnd + 1, 0), lab(n, -1), nnv(nsnd), nav(nsnd);
    for (i64 i = 0; i < nsnd; ++i) {
        i64 j = elmorder[i];
        i64 u = newv[j];
        i64 p = sndptr[i];
        lab[u] = p;
        for (i64 v = parent[u]; v != -1; v = parent[v]) {
            if (v == sanc[j]) break;
            lab[v] = ++p;
        }
        nnv[i] = p - sndptr[i] + 1;
        nav[i] = colcount[u] - nnv[i];
        sndptr[i + 1] = p + 1;
    }
    if (sndptr[nsnd] != n) { delete R; return -6; }

    R->nfr = nsnd;
    R->perm.assign(n, 0);
    R->invp.assign(n, 0);
    for (i64 v = 0; v < n; ++v) {
        R->perm[lab[v]] = order0[v];
        R->invp[order0[v]] = lab[v];
    }

    // final supernode children lists (increasing labels)
    std::vector<i64> fcptr(nsnd + 2, 0), fctgt(nsnd);
    for (i64 i = 0; i < nsnd; ++i) {
        i64 pj = fpnt[i] == -1 ? nsnd : fpnt[i];
        fcptr[pj + 1]++;
    }
    for (i64 i = 0; i <= nsnd; ++i) fcptr[i + 1] += fcptr[i];
    {
        std::vector<i64> pos(fcptr.begin(), fcptr.end() - 1);
        for (i64 i = 0; i < nsnd; ++i) {
            i64 pj = fpnt[i] == -1 ? nsnd : fpnt[i];
            fctgt[pos[pj]++] = i;
        }
    }

    // 7. separators: sep(j) = sorted({A-neighbours of first(j)} U sep(children)) \ {<= last(j)}
    std::vector<i64> sepp(nsnd + 1, 0);
    for (i64 i = 0; i < nsnd; ++i) sepp[i + 1] = sepp[i] + nav[i];
    std::vector<i64> sept(sepp[nsnd]);
    {
        std::vector<i64> buf, tmp;
        for (i64 j = 0; j < nsnd; ++j) {
            i64 last = sndptr[j + 1] - 1;
            buf.clear();
            i64 f = R->perm[sndptr[j]];
            for (i64 p = aptr[f]; p < aptr[f + 1]; ++p) {
                i64 w = R->invp[atgt[p]];
                if (w > last) buf.push_back(w);
            }
            std::sort(buf.begin(), buf.end());
            for (i64 q = fcptr[j]; q < fcptr[j + 1]; ++q) {
                i64 c = fctgt[q];
                tmp.clear();
                i64 a = sepp[c], b = sepp[c + 1];
                while (a < b && sept[a] <= last) ++a;
                std::merge(buf.begin(), buf.end(), sept.begin() + a, sept.begin() + b,
                           std::back_inserter(tmp));
                tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
                buf.swap(tmp);
            }
            if ((i64)buf.size() != nav[j]) { delete R; return -7; }
            std::copy(buf.begin(), buf.end(), sept.begin() + sepp[j]);
        }
    }

    // 8. ChordalSymbolic(res, sep, tree): rel maps, pointers, workspace maxima
    R->resptr.resize(nsnd + 1);
    R->sepptr.resize(nsnd + 1);
    for (i64 j = 0; j <= nsnd; ++j) {
        R->resptr[j] = sndptr[j] + 1;
        R->sepptr[j] = sepp[j] + 1;
    }
    R->septgt.resize(sept.size());
    for (size_t p = 0; p < sept.size(); ++p) R->septgt[p] = sept[p] + 1;
    R->reltgt.assign(sept.size(), 0);
    R->chdptr.resize(nsnd + 2);
    for (i64 j = 0; j < nsnd + 2; ++j) R->chdptr[j] = fcptr[j] + 1;
    R->chdtgt.resize(nsnd);
    for (i64 j = 0; j < nsnd; ++j) R->chdtgt[j] = fctgt[j] + 1;
    R->pnt.resize(nsnd);
    for (i64 j = 0; j < nsnd; ++j) R->pnt[j] = fpnt[j] + 1;  // 0 = root
    R->idx.assign(n, 0);
    R->Dptr.resize(nsnd + 1);
    R->Lptr.resize(nsnd + 1);

    i64 nMptr = 1, jMptr = 1, nMval = 0, jMval = 0, nNval = 0, jNval = 0, nFval = 0;
    i64 Dp = 0, Lp = 0;
    for (i64 j = 0; j < nsnd; ++j) {
        R->Dptr[j] = Dp + 1;
        R->Lptr[j] = Lp + 1;
        i64 nn = nnv[j], na = nav[j], nj = nn + na;
        i64 r0 = sndptr[j];
        for (i64 w = r0; w < r0 + nn; ++w) R->idx[w] = j + 1;
        for (i64 q = fcptr[j]; q < fcptr[j + 1]; ++q) {
            i64 i = fctgt[q];
            // rel: position (1-based) of each v in sep(i) inside [res(j); sep(j)]
            i64 s = sepp[j];
            for (i64 p = sepp[i]; p < sepp[i + 1]; ++p) {
                i64 v = sept[p];
                if (v < r0 + nn) {
                    R->reltgt[p] = v - r0 + 1;
                } else {
                    while (s < sepp[j + 1] && sept[s] < v) ++s;
                    if (s >= sepp[j + 1] || sept[s] != v) { delete R; return -8; }
                    R->reltgt[p] = nn + (s - sepp[j]) + 1;
                }
            }
            i64 ma = nav[i];
            jMptr -= 1;
            jMval -= ma * ma;
            jNval -= ma;
        }
        if (na > 0) {
            jMptr += 1;
            jMval += na * na;
            jNval += na;
            nMptr = std::max(nMptr, jMptr);
            nMval = std::max(nMval, jMval);
            nNval = std::max(nNval, jNval);
        }
        nFval = std::max(nFval, nj);
        Dp += nn * nn;
        Lp += nn * na;
    }
    R->Dptr[nsnd] = Dp + 1;
    R->Lptr[nsnd] = Lp + 1;
    R->nMptr = nMptr;
    R->nMval = nMval;
    R->nNval = nNval;
    R->nFval = nFval;
    *out = R;
    return 0;
}

extern "C" int mfs_sizes(const mfs_result* r, int64_t* sizes) {
    if (!r || !sizes) return -1;
    sizes[MFS_N] = r->n;
    sizes[MFS_NFR] = r->nfr;
    sizes[MFS_NSEP] = (i64)r->septgt.size();
    sizes[MFS_NDZ] = r->nfr ? r->Dptr[r->nfr] - 1 : 0;
    sizes[MFS_NLZ] = r->nfr ? r->Lptr[r->nfr] - 1 : 0;
    sizes[MFS_NMPTR] = r->nMptr;
    sizes[MFS_NMVAL] = r->nMval;
    sizes[MFS_NNVAL] = r->nNval;
    sizes[MFS_NFVAL] = r->nFval;
    return 0;
}

extern "C" int mfs_export(const mfs_result* r, int64_t* perm, int64_t* resptr,
                          int64_t* sepptr, int64_t* septgt, int64_t* reltgt,
                          int64_t* chdptr, int64_t* chdtgt, int64_t* pnt, int64_t* idx,
                          int64_t* Dptr, int64_t* Lptr) {
    if (!r) return -1;
    auto put = [](int64_t* dst, const std::vector<i64>& src, i64 add) {
        if (!dst) return;
        for (size_t k = 0; k < src.size(); ++k) dst[k] = src[k] + add;
    };
    put(perm, r->perm, 1);
    put(resptr, r->resptr, 0);
    put(sepptr, r->sepptr, 0);
    put(septgt, r->septgt, 0);
    put(reltgt, r->reltgt, 0);
    put(chdptr, r->chdptr, 0);
    put(chdtgt, r->chdtgt, 0);
    put(pnt, r->pnt, 0);
    put(idx, r->idx, 0);
    put(Dptr, r->Dptr, 0);
    put(Lptr, r->Lptr, 0);
    return 0;
}

extern "C" void mfs_free(mfs_result* r) { delete r; }

// flatindex(S, v, w, Val(:L)) (chordal_symbolic.jl:535-563) for every stored
// entry of A, in final labels v = invp[row], w = invp[col]; entries of the
// strict upper triangle of PAP^T map to 0.
extern "C" int mfs_flatindices(const mfs_result* r, int64_t n, const int64_t* colptr,
                               const int64_t* rowval, int64_t* flat) {
    if (!r || n != r->n || !flat) return -1;
    i64 ndz = r->nfr ? r->Dptr[r->nfr] - 1 : 0;
    for (i64 col = 0; col < n; ++col) {
        for (i64 p = colptr[col] - 1; p < colptr[col + 1] - 1; ++p) {
            i64 row = rowval[p] - 1;
            if (row < 0 || row >= n) return -2;
            i64 v = r->invp[row], w = r->invp[col];
            flat[p] = 0;
            if (v < w) continue;
            i64 j = r->idx[w] - 1;
            i64 r0 = r->resptr[j] - 1, r1 = r->resptr[j + 1] - 1;
            i64 nn = r1 - r0;
            i64 s0 = r->sepptr[j] - 1, s1 = r->sepptr[j + 1] - 1;
            i64 na = s1 - s0;
            i64 wl = w - r0;
            if (v < r1) {
                flat[p] = r->Dptr[j] + (v - r0) + wl * nn;
            } else {
                auto b = r->septgt.begin() + s0, e = r->septgt.begin() + s1;
                auto it = std::lower_bound(b, e, v + 1);
                if (it != e && *it == v + 1) {
                    i64 i = it - b;
                    flat[p] = r->Lptr[j] + ndz + i + wl * na;
                }
            }
        }
    }
    return 0;
}

// Geometric nested dissection of a box grid: split the longest side at its
// middle plane, order both halves recursively, then the plane.
namespace {
struct NDState {
    i64 nx, ny, nz, leaf;
    i64* order;
    i64 k;
};

void nd_box(NDState& s, i64 x0, i64 x1, i64 y0, i64 y1, i64 z0, i64 z1) {
    i64 lx = x1 - x0, ly = y1 - y0, lz = z1 - z0;
    if (lx <= 0 || ly <= 0 || lz <= 0) return;
    if (lx * ly * lz <= s.leaf) {
        for (i64 z = z0; z < z1; ++z)
            for (i64 y = y0; y < y1; ++y)
                for (i64 x = x0; x < x1; ++x)
                    s.order[s.k++] = 1 + x + s.nx * (y + s.ny * z);
        return;
    }
    if (lx >= ly && lx >= lz) {
        i64 m = x0 + lx / 2;
        nd_box(s, x0, m, y0, y1, z0, z1);
        nd_box(s, m + 1, x1, y0, y1, z0, z1);
        nd_box(s, m, m + 1, y0, y1, z0, z1);
    } else if (ly >= lz) {
        i64 m = y0 + ly / 2;
        nd_box(s, x0, x1, y0, m, z0, z1);
        nd_box(s, x0, x1, m + 1, y1, z0, z1);
        nd_box(s, x0, x1, m, m + 1, z0, z1);
    } else {
        i64 m = z0 + lz / 2;
        nd_box(s, x0, x1, y0, y1, z0, m);
        nd_box(s, x0, x1, y0, y1, m + 1, z1);
        nd_box(s, x0, x1, y0, y1, m, m + 1);
    }
}
}  // namespace

extern "C" int mfs_grid_nd(int64_t nx, int64_t ny, int64_t nz, int64_t leaf, int64_t* order) {
    if (nx < 1 || ny < 1 || nz < 1 || !order) return -1;
    NDState s{nx, ny, nz, leaf < 1 ? 1 : leaf, order, 0};
    nd_box(s, 0, nx, 0, ny, 0, nz);
    return s.k == nx * ny * nz ? 0 : -2;
}

namespace mfs {
// shared with the MMD ordering (mmd.cpp)
bool build_adjacency(int64_t n, const int64_t* colptr, const int64_t* rowval,
                     std::vector<int64_t>& ptr, std::vector<int64_t>& tgt) {
    return ::build_adjacency(n, colptr, rowval, ptr, tgt);
}
}  // namespace mfs

// Multiple minimum degree: implemented in mmd.cpp when present in the build.
#ifndef MFS_HAVE_MMD
extern "C" int mfs_mmd(int64_t, const int64_t*, const int64_t*, int64_t, int64_t*, int64_t*) {
    return -100;
}
#endif
