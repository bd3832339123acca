/*
 * mfsym.h — host-side symbolic analysis for the multifrontal engine.
 *
 * This is the C-ABI of libmfsym.so, a C++ restatement of the symbolic layer
 * of CliqueTrees.jl that produces the `ChordalSymbolic` arrays consumed by the
 * numeric engine (libmfhip.so, include/mfhip.h).  Every array that crosses
 * this boundary uses the reference's own conventions: Int64, 1-based indices,
 * CSR pointer arrays of length count+1 whose first entry is 1.
 *
 * Reference interfaces replaced (paths relative to the CliqueTrees.jl tree):
 *   mfs_symbolic   <- `symbolic(A; alg=perm, snd=Maximal())`
 *                     src/Multifrontal.jl/src/chordal_symbolic.jl:72-91, which
 *                     runs `cliquetree` src/clique_trees.jl:106-144
 *                     (`supernodetree_impl!` src/supernode_trees.jl:107-180,
 *                     `cliquetree_impl!` src/clique_trees.jl:146-234) and
 *                     `ChordalSymbolic(res, sep, tree)`
 *                     src/Multifrontal.jl/src/chordal_symbolic.jl:302-378.
 *   mfs_flatindices <- `flatindices(S, A, Val(:L))`
 *                     src/Multifrontal.jl/src/chordal_symbolic.jl:595-752.
 *   mfs_mmd        <- `mmd(n, colptr, rowval; delta)` src/MMDLib.jl/src/MMDLib.jl:19-60
 *                     (the multiple-minimum-degree ordering, `MMD()`).
 *   mfs_grid_nd    <- no counterpart in the reference: a geometric nested
 *                     dissection of a box grid, passed to `symbolic` as a
 *                     user permutation (src/elimination_algorithms.jl:776-779).
 *
 * Return codes: 0 success, negative on invalid input (never aborts).
 */
#ifndef MFSYM_H
#define MFSYM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mfs_result mfs_result;

/* Sizes reported by mfs_sizes (indices into the int64 array it fills). */
enum {
    MFS_N = 0,      /* number of columns                                   */
    MFS_NFR,        /* number of fronts (supernodes)                       */
    MFS_NSEP,       /* total separator entries = length of septgt/reltgt   */
    MFS_NDZ,        /* length of Dval (sum nn^2)                           */
    MFS_NLZ,        /* length of Lval (sum nn*na)                          */
    MFS_NMPTR,      /* workspace maxima, chordal_symbolic.jl:309-367        */
    MFS_NMVAL,
    MFS_NNVAL,
    MFS_NFVAL,
    MFS_NSIZES
};

/*
 * Symbolic factorization of the symmetric pattern (colptr, rowval) (CSC,
 * 1-based; either triangle or both; the diagonal is ignored) under the
 * elimination order `order` (order[k] = vertex eliminated k-th, 1-based).
 * snd_kind: 0 = Maximal supernodes (the reference default), 1 = Nodal
 * (one column per front), 2 = Fundamental.
 */
int mfs_symbolic(int64_t n, const int64_t* colptr, const int64_t* rowval,
                 const int64_t* order, int snd_kind, mfs_result** out);

int mfs_sizes(const mfs_result* r, int64_t* sizes /* MFS_NSIZES */);

/*
 * Copy the ChordalSymbolic arrays out.  Lengths: perm n; resptr, sepptr,
 * Dptr, Lptr nfr+1; septgt, reltgt nsep; chdptr nfr+2; chdtgt nfr; pnt nfr;
 * idx n.  Any pointer may be NULL to skip that array.
 */
int mfs_export(const mfs_result* r, int64_t* perm, int64_t* resptr,
               int64_t* sepptr, int64_t* septgt, int64_t* reltgt,
               int64_t* chdptr, int64_t* chdtgt, int64_t* pnt, int64_t* idx,
               int64_t* Dptr, int64_t* Lptr);

void mfs_free(mfs_result* r);

/*
 * Flat index of every stored entry of the lower triangle of P A P^T:
 * for each nonzero k of A (CSC, 1-based, full symmetric pattern or lower
 * triangle), flat[k] is its 1-based slot in [Dval; Lval] when the entry
 * (perm-mapped) lies in the lower triangle and is stored, else 0.
 */
int mfs_flatindices(const mfs_result* r, int64_t n, const int64_t* colptr,
                    const int64_t* rowval, int64_t* flat);

/* Multiple minimum degree ordering; writes invp (1-based), perm optional. */
int mfs_mmd(int64_t n, const int64_t* colptr, const int64_t* rowval,
            int64_t delta, int64_t* invp, int64_t* perm);

/*
 * Geometric nested dissection of an nx*ny*nz grid (vertex (x,y,z) has label
 * 1 + x + nx*(y + ny*z)).  Writes the elimination order (order[k] = vertex
 * eliminated k-th, 1-based).  Boxes with at most `leaf` vertices are ordered
 * lexicographically.
 */
int mfs_grid_nd(int64_t nx, int64_t ny, int64_t nz, int64_t leaf, int64_t* order);

/*
 * Runs the numeric engine's planner (internal fronts, level schedule, memory
 * layout, task lists and their bounds validation) on the host only, for the
 * ChordalSymbolic arrays of mfs_export.  No counterpart in the reference: it
 * lets the planner be checked without a GPU.  stats (optional): internal
 * fronts, small-region fronts, big fronts, levels, panel/U/W elements, GEMM,
 * extend-add and solve tasks, reference and engine flops, max front size.
 * Returns 0, or -1 with the reason in msg.
 */
int mf_plan_check(int64_t n, int64_t nfr, const int64_t* resptr, const int64_t* sepptr,
                  const int64_t* septgt, const int64_t* reltgt, const int64_t* chdptr,
                  const int64_t* chdtgt, const int64_t* Dptr, const int64_t* Lptr,
                  int64_t* stats, int64_t nstats, char* msg, int64_t msglen);

/* Host-only check of the selected-inversion schedule (libmfhip.so's
 * mf_selinv): builds the plan and the schedule, then runs every task on the
 * host in ticket order, checking each wait and each buffer bound, from a
 * Cholesky factor in the reference layout (Dval, Lval, uplo = :L); writes the
 * selected inverse to SDval/SLval (same layout) when all four arrays are
 * given.  stats: tasks, GEMM tiles, small-front tasks, gather tasks, flops,
 * X elements, Lt elements, M elements, maximum depth.  Returns 0 or -1 (msg). */
int mf_selinv_check(int64_t n, int64_t nfr, const int64_t* resptr, const int64_t* sepptr,
                    const int64_t* septgt, const int64_t* reltgt, const int64_t* chdptr,
                    const int64_t* chdtgt, const int64_t* Dptr, const int64_t* Lptr,
                    const double* Dval, const double* Lval, double* SDval, double* SLval,
                    int64_t* stats, int64_t nstats, char* msg, int64_t msglen);

/* Host-only check of the completion's backward-solve schedule
 * (libmfhip.so's mf_complete_impl; no counterpart in the reference): a
 * forest of nfr roots with separator na[u] and residual nn[u] (big roots
 * padded as the library pads them), the schedule run on the host in ticket
 * order from the Cholesky factor of a random positive definite matrix per
 * root (seed), each wait and each value read checked; *worst = the largest
 * difference of the rows the completion reads from backward substitution,
 * relative to each column's largest entry.  stats: tasks, flops, Y
 * elements.  Returns 0 or -1 (msg). */
int mf_cpl_solve_check(int64_t nfr, const int64_t* na, const int64_t* nn, uint64_t seed, double* worst,
                       int64_t* stats, int64_t nstats, char* msg, int64_t msglen);

#ifdef __cplusplus
}
#endif

#endif /* MFSYM_H */
