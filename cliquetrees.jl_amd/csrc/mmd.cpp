// mmd.cpp — multiple minimum degree ordering (Liu, "Modification of the
// minimum degree algorithm by multiple elimination", ACM TOMS 11 (1985);
// George & Liu, SIAM Rev. 31 (1989)), restated for the reference's
// `mmd(neqns, xadj, adjncy; delta)` (src/MMDLib.jl/src/MMDLib.jl:19-60,
// called by `permutation(weights, graph, ::MMD)` at
// src/elimination_algorithms.jl:830-833 on the loop-free graph of
// src/bipartite_graphs.jl:393-413).
//
// The ordering must be identical to the reference's, so the order of every
// list operation follows the published algorithm as the reference runs it:
// degree buckets are LIFO stacks, elements are scanned in quotient-graph
// storage order, supernodes absorb in the order met, and the final numbering
// gives each absorbed node the next number after its representative's.
//
// Quotient graph storage (1-based, 0 terminates a list, -e links to the
// storage of element e):
//   adj[xadj[v] .. xadj[v+1]-1]  neighbours of v (uneliminated nodes and
//                                elements), reused for the reach of an
//                                eliminated node
// Per node:
//   head[d]  first node of degree bucket d;  next[v] / prev[v] bucket links
//            (prev[v] = -d when v heads bucket d, 0 when v is in no bucket);
//            after elimination next[v] = -(order number), after absorption
//            next[v] = -(representative)
//   size/wgt supernode size / weight (0: absorbed)
//   link     chains the eliminated nodes of a pass, the elements met while
//            forming a reach, and the two update queues
//   mark     visit stamps;  need  > 0 when the degree must be recomputed
//            (2: the node is adjacent to exactly two elements)
#include <algorithm>
#include <cstdint>
#include <limits>
#include <vector>

namespace {

using i64 = int64_t;

class MultipleMinimumDegree {
public:
    MultipleMinimumDegree(i64 n, std::vector<i64> xadj, std::vector<i64> adj, i64 delta)
        : n_(n), delta_(delta), xadj_(std::move(xadj)), adj_(std::move(adj)) {}

    // Returns invp: invp[v] = 1-based position of vertex v (1-based vector).
    std::vector<i64> run() {
        const i64 total = n_;  // unit vertex weights
        wgt_.assign(n_ + 1, 1);
        next_.assign(n_ + 1, 0);
        prev_.assign(n_ + 1, 0);
        size_.assign(n_ + 1, 1);
        swgt_.assign(n_ + 1, 1);
        link_.assign(n_ + 1, 0);
        need_.assign(n_ + 1, 0);
        mark_.assign(n_ + 1, 0);
        head_.assign(total + delta_ + 2, 0);
        i64 mindeg = total, num = 1;
        tag_ = 1;
        for (i64 v = 1; v <= n_; ++v) {
            if (xadj_[v] < xadj_[v + 1]) {
                i64 d = 1;
                for (i64 i = xadj_[v]; i < xadj_[v + 1]; ++i) d += wgt_[adj_[i]];
                push(v, d);
                if (d < mindeg) mindeg = d;
            } else {  // isolated vertex: numbered at once
                mark_[v] = kBig;
                next_[v] = -num;
                ++num;
            }
        }
        while (num <= n_) {
            while (head_[mindeg] <= 0) ++mindeg;
            const i64 limit = std::max(mindeg, mindeg + delta_);
            i64 elim = 0;  // nodes eliminated in this pass
            bool last = false;
            for (;;) {
                i64 v = head_[mindeg];
                while (v <= 0) {
                    if (++mindeg > limit) break;
                    v = head_[mindeg];
                }
                if (v <= 0) break;
                const i64 nx = next_[v];
                head_[mindeg] = nx;
                if (nx > 0) prev_[nx] = -mindeg;
                next_[v] = -num;
                if (num + size_[v] > n_) {
                    last = true;
                    break;
                }
                bump_tag(1);
                eliminate(v);
                num += size_[v];
                link_[v] = elim;
                elim = v;
            }
            if (last || num > n_) break;
            mindeg = update(elim, mindeg);
        }
        return number();
    }

private:
    static constexpr i64 kBig = std::numeric_limits<i64>::max() - 100000;

    void push(i64 v, i64 d) {
        const i64 f = head_[d];
        head_[d] = v;
        next_[v] = f;
        if (f > 0) prev_[f] = v;
        prev_[v] = -d;
    }

    void unlink(i64 v) {
        const i64 p = prev_[v];
        if (p == 0) return;
        const i64 nx = next_[v];
        if (nx > 0) prev_[nx] = p;
        if (p > 0) next_[p] = nx;
        else head_[-p] = nx;
    }

    // tag += step, restarting the stamps when they would overflow
    void bump_tag(i64 step) {
        tag_ += step;
        if (tag_ >= kBig) {
            tag_ = 1;
            for (i64 v = 1; v <= n_; ++v)
                if (mark_[v] < kBig) mark_[v] = 0;
        }
    }

    // Eliminates v: its reach (uneliminated neighbours plus the nodes of its
    // adjacent elements) replaces its adjacency, then every reached node is
    // taken out of its bucket, its list is purged of covered entries and v is
    // added as an element neighbour; a node left with no other neighbour is
    // absorbed into v.
    void eliminate(i64 v) {
        mark_[v] = tag_;
        i64 elm = 0;
        i64 w = xadj_[v], wend = xadj_[v + 1] - 1;  // write cursor / last slot
        for (i64 i = xadj_[v]; i <= xadj_[v + 1] - 1; ++i) {
            const i64 u = adj_[i];
            if (u == 0) break;
            if (mark_[u] < tag_) {
                mark_[u] = tag_;
                if (next_[u] >= 0) adj_[w++] = u;
                else {
                    link_[u] = elm;
                    elm = u;
                }
            }
        }
        while (elm > 0) {
            adj_[wend] = -elm;
            i64 j = xadj_[elm], jend = xadj_[elm + 1];
            i64 u = adj_[j];
            while (u != 0) {
                if (u < 0) {
                    j = xadj_[-u];
                    jend = xadj_[-u + 1];
                } else {
                    if (mark_[u] < tag_ && next_[u] >= 0) {
                        mark_[u] = tag_;
                        while (w >= wend) {  // continue in the storage of a merged element
                            const i64 e = -adj_[wend];
                            w = xadj_[e];
                            wend = xadj_[e + 1] - 1;
                        }
                        adj_[w++] = u;
                    }
                    ++j;
                }
                if (j >= jend) break;
                u = adj_[j];
            }
            elm = link_[elm];
        }
        if (w <= wend) adj_[w] = 0;

        i64 i = xadj_[v], iend = xadj_[v + 1];
        i64 r = adj_[i];
        while (r != 0) {
            if (r < 0) {
                i = xadj_[-r];
                iend = xadj_[-r + 1];
            } else {
                unlink(r);
                i64 q = xadj_[r];
                for (i64 j = xadj_[r]; j <= xadj_[r + 1] - 1; ++j) {
                    const i64 u = adj_[j];
                    if (u == 0) break;
                    if (mark_[u] < tag_) adj_[q++] = u;
                }
                const i64 left = q - xadj_[r];
                if (left <= 0) {  // indistinguishable from v: absorb
                    size_[v] += size_[r];
                    size_[r] = 0;
                    swgt_[v] += swgt_[r];
                    swgt_[r] = 0;
                    next_[r] = -v;
                    mark_[r] = kBig;
                } else {
                    need_[r] = left + 1;
                    adj_[q++] = v;
                    if (q < xadj_[r + 1]) adj_[q] = 0;
                }
                prev_[r] = 0;
                ++i;
            }
            if (i >= iend) break;
            r = adj_[i];
        }
    }

    void place(i64 u, i64 d, i64& mindeg) {
        d -= swgt_[u];
        push(u, d);
        need_[u] = 0;
        if (d < mindeg) mindeg = d;
    }

    // Recomputes the external degrees of the nodes adjacent to the elements
    // formed in this pass (list through link_ from `elim`).
    i64 update(i64 elim, i64 mindeg) {
        const i64 limit = mindeg + delta_;
        for (i64 e = elim; e > 0; e = link_[e]) {
            i64 mtag = tag_ + limit;
            if (mtag >= kBig) {
                tag_ = 1;
                mtag = tag_ + limit;
                for (i64 v = 1; v <= n_; ++v)
                    if (mark_[v] < kBig) mark_[v] = 0;
            }
            i64 q2 = 0, qx = 0, ew = 0;
            i64 i = xadj_[e], iend = xadj_[e + 1];
            i64 u = adj_[i];
            while (u != 0) {
                if (u < 0) {
                    i = xadj_[-u];
                    iend = xadj_[-u + 1];
                } else {
                    if (size_[u] != 0) {
                        ew += swgt_[u];
                        mark_[u] = mtag;
                        if (need_[u] > 0) {
                            if (need_[u] != 2) {
                                link_[u] = qx;
                                qx = u;
                            } else {
                                link_[u] = q2;
                                q2 = u;
                            }
                        }
                    }
                    ++i;
                }
                if (i >= iend) break;
                u = adj_[i];
            }
            // nodes adjacent to exactly two elements
            for (u = q2; u > 0; u = link_[u]) {
                if (need_[u] <= 0) continue;
                ++tag_;
                i64 deg = ew + 1;
                const i64 s = xadj_[u];
                i64 other = adj_[s];
                if (other == e) other = adj_[s + 1];
                if (next_[other] >= 0) {
                    deg += swgt_[other];
                } else {
                    i64 j = xadj_[other], jend = xadj_[other + 1];
                    i64 x = adj_[j];
                    while (x != 0) {
                        if (x < 0) {
                            j = xadj_[-x];
                            jend = xadj_[-x + 1];
                        } else {
                            if (x != u && size_[x] != 0) {
                                if (mark_[x] < tag_) {
                                    mark_[x] = tag_;
                                    deg += swgt_[x];
                                } else if (need_[x] > 0) {
                                    if (need_[x] == 2) {  // indistinguishable from u
                                        size_[u] += size_[x];
                                        size_[x] = 0;
                                        swgt_[u] += swgt_[x];
                                        swgt_[x] = 0;
                                        mark_[x] = kBig;
                                        next_[x] = -u;
                                    }
                                    need_[x] = 0;
                                    prev_[x] = 0;
                                }
                            }
                            ++j;
                        }
                        if (j >= jend) break;
                        x = adj_[j];
                    }
                }
                place(u, deg, mindeg);
            }
            // the other nodes
            for (u = qx; u > 0; u = link_[u]) {
                if (need_[u] <= 0) continue;
                ++tag_;
                i64 deg = ew + 1;
                for (i64 k = xadj_[u]; k <= xadj_[u + 1] - 1; ++k) {
                    const i64 y = adj_[k];
                    if (y == 0) break;
                    if (mark_[y] >= tag_) continue;
                    mark_[y] = tag_;
                    if (next_[y] >= 0) {
                        deg += swgt_[y];
                        continue;
                    }
                    i64 j = xadj_[y], jend = xadj_[y + 1];
                    i64 x = adj_[j];
                    while (x != 0) {
                        if (x < 0) {
                            j = xadj_[-x];
                            jend = xadj_[-x + 1];
                        } else {
                            if (mark_[x] < tag_) {
                                mark_[x] = tag_;
                                deg += swgt_[x];
                            }
                            ++j;
                        }
                        if (j >= jend) break;
                        x = adj_[j];
                    }
                }
                place(u, deg, mindeg);
            }
            tag_ = mtag;
        }
        return mindeg;
    }

    // Representatives hold -(first number of their block); absorbed nodes
    // point (through chains) to their representative and take the next
    // numbers of its block in vertex order.
    std::vector<i64> number() {
        for (i64 v = 1; v <= n_; ++v) {
            if (size_[v] > 0) {
                next_[v] = -next_[v];
                size_[v] = next_[v];  // last number used in the block
            } else {
                i64 root = -next_[v];
                while (size_[root] == 0) root = -next_[root];
                for (i64 x = v; x != root;) {
                    const i64 up = -next_[x];
                    next_[x] = -root;
                    x = up;
                }
            }
        }
        for (i64 v = 1; v <= n_; ++v) {
            if (size_[v] == 0) {
                const i64 root = -next_[v];
                next_[v] = ++size_[root];
            }
        }
        return next_;
    }

    i64 n_, delta_, tag_ = 1;
    std::vector<i64> xadj_, adj_;
    std::vector<i64> wgt_, next_, prev_, size_, swgt_, link_, need_, mark_, head_;
};

}  // namespace

namespace mfs {
bool build_adjacency(int64_t n, const int64_t* colptr, const int64_t* rowval,
                     std::vector<int64_t>& ptr, std::vector<int64_t>& tgt);
}

// invp[v] (1-based) = position of vertex v; perm (optional) its inverse.
extern "C" int mfs_mmd(int64_t n, const int64_t* colptr, const int64_t* rowval, int64_t delta,
                       int64_t* invp, int64_t* perm) {
    if (n < 0 || !colptr || (n > 0 && (!rowval || !invp)) || delta < 0) return -1;
    if (n == 0) return 0;
    std::vector<int64_t> ptr, tgt;
    if (!mfs::build_adjacency(n, colptr, rowval, ptr, tgt)) return -4;
    // 1-based quotient-graph storage
    std::vector<i64> xadj(n + 2), adj(tgt.size() + 2, 0);
    for (i64 v = 0; v <= n; ++v) xadj[v + 1] = ptr[v] + 1;
    for (size_t k = 0; k < tgt.size(); ++k) adj[k + 1] = tgt[k] + 1;
    MultipleMinimumDegree m(n, std::move(xadj), std::move(adj), delta);
    std::vector<i64> pos = m.run();
    std::vector<char> seen(n + 1, 0);
    for (i64 v = 1; v <= n; ++v) {
        const i64 p = pos[v];
        if (p < 1 || p > n || seen[p]) return -2;
        seen[p] = 1;
        invp[v - 1] = p;
        if (perm) perm[p - 1] = v;
    }
    return 0;
}
