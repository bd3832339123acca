// mf_internal.h — data structures shared by the planner (host) and the
// kernels of libmfhip.so.
//
// Layout in HBM (one factor object):
//   panel pool  For every internal front J, the stacked panel [D_JJ; L_J]
//               (nJ x nnJ, column-major, ld = nJ) — the F11/F21 part of the
//               reference's frontal matrix (cholesky.jl:467-500).
//   U pool      Update matrix U_J = F22 - L_J L_J' (naJ x naJ, ld = naJ),
//               lower triangle meaningful (cholesky.jl:531-553).
//   W pool      Inverses of the 64x64 diagonal blocks of big fronts
//               (nnJ x 64, ld = nnJ), reused by the solves.
//   Dval/Lval/d The reference's own factor arrays, written by the copy-out
//               kernel from the panels.
// An internal front is a chain of reference fronts j_1 -> j_2 -> ... (each
// the last child of the next, hence contiguous in postorder and in columns)
// merged when the explicit zeros it introduces stay below a bound.  Entries
// that are structurally zero in the reference factor stay exactly zero.
#pragma once

#include <cstdint>

namespace mf {

constexpr int NB = 64;          // panel width and GEMM tile edge of big fronts
constexpr int SMALL_NJ = 64;    // fronts with nJ <= SMALL_NJ use the fused LDS kernel
constexpr int TINY_NJ = 32;     // small fronts that one wavefront factors (four per ticket)
constexpr int EA_WIN = 1024;    // parent rows per extend-add window (one wave)
constexpr int EA_ROWS = 16384;  // parent rows per extend-add task (about)
constexpr int WIDE_K = 128;     // update-matrix tiles are 128 x 128 when nn >= WIDE_K
constexpr int UPD_GROUP = 8;    // panel steps applied by one trailing-update tile
constexpr int SIM_WORKERS = 512;  // resident workgroups assumed when ordering the DAG tickets
constexpr int SEP_TILES = 8;    // separator tiles per backward partial-product task (fronts whose
                                // separator has more are split over several tasks)

struct Front {
    int64_t poff;    // panel offset (elements) in the panel pool
    int64_t uoff;    // U offset (elements) in the U pool, -1 if na == 0
    int64_t woff;    // W offset (elements), -1 for small fronts
    int32_t c0;      // first global (permuted) column, 0-based
    int32_t nn;      // eliminated columns
    int32_t na;      // update rows
    int32_t rowoff;  // offset of the row list (nn + na entries) in rows[]
    int32_t chd0;    // children of J in chd[] (in accumulation order)
    int32_t chd1;
    int32_t reloff;  // offset of J's rel map (na entries) in rel[]: positions of
                     // J's update rows inside its parent's row list
    int32_t parent;  // internal parent, -1 for roots
};

// Extend-add task: parent front, parent column, parent row range.
struct EaTask {
    int32_t front;
    int32_t col;
    int32_t row0;
    int32_t row1;
    int32_t src0;   // sources [src0, src1) in ea_src (child, child column)
    int32_t src1;
};

// Device form of an extend-add column: destination (panel or U element of
// parent row 0, so dst[p] is parent row p) and, per source, the child's U
// column, rel map and window table, so the kernel reads them in one step.
struct EaDst {
    int64_t off;
    int32_t to_u;
    int32_t pad;
};
struct EaSrc {
    int64_t u_off;    // U element (0, tcol) of the child
    int32_t rel_off;  // child's rel map in rel[]
    int32_t ws_off;   // child's window table in wstart[]
    int32_t tcol;     // child column (its first row)
    int32_t i0, i1;   // child rows landing in the column's first window
    int32_t pad;
};

enum : int8_t { ARR_PANEL = 0, ARR_U = 1, ARR_W = 2 };
enum : int8_t {
    G_SUB = 0,      // C -= A diag(s) B'            (trailing / SYRK updates)
    G_TRSM = 1,     // C  = A W' [diag(1/s)]        (panel solve, A aliases C)
};

// One 64x64 output tile of a batched GEMM.
struct GemmTask {
    int64_t a_off, b_off, c_off;
    int32_t lda, ldb, ldc;
    int32_t m, n, k;
    int32_t s_col;  // global column of s[0] (LDLt pivots), -1: none
    int8_t arrA, arrB, arrC, mode;
    int8_t lower;   // store only r >= c (diagonal tile)
    int8_t wide;    // 128 x 128 tile (m, n <= 2 NB), else 64 x 64
    int8_t pad[2];
};

// Factor one diagonal block of a big front and invert it.
struct DiagTask {
    int32_t front;
    int32_t k0;     // first local column of the block
    int32_t kb;     // block width (<= NB)
    int32_t pad;
};

// Task of the persistent tile-DAG factorization of the big region.
enum : int32_t { DT_EA = 0, DT_SMALL = 1, DT_POTRF = 2, DT_TRSM = 3, DT_UPD = 4, DT_SYRK = 5,
                 DT_KINDS = 6 };
struct DagTask {
    int32_t kind;
    int32_t front;
    int32_t a;   // DT_EA: first EaTask; DT_POTRF: panel block K; else: GemmTask index
    int32_t b;   // DT_EA: end EaTask; DT_TRSM/UPD/SYRK: row block I
    int32_t c;   // DT_EA: 1 for panel columns, 0 for update-matrix columns;
                 // DT_TRSM: panel block K; DT_UPD/SYRK: column block Jc
    int32_t d;   // DT_UPD: panel blocks [K0, K1) as K0 | K1 << 16; DT_SYRK: bit 0 /
                 // bit 1: the tile also covers row block I+1 / column block Jc+1
};

// Per-front constants of the DAG.
struct DagFront {
    int32_t toff;    // offset of the front's tile counters (K * R + I)
    int32_t R;       // row blocks (panel blocks p, then update-row blocks)
    int32_t p;       // panel (column) blocks
    int32_t n_ea;    // extend-add tasks (all; 1 for a small front)
    int32_t n_done;  // tasks whose completion makes U final (SYRK tiles, or 1 for small)
    int32_t n_eap;   // extend-add tasks into panel columns: the panel waits only for these
};

// Copy-out of one reference front from its internal panel.
struct OutFront {
    int64_t Dp, Lp;      // 0-based offsets in Dval / Lval
    int32_t J;           // internal front
    int32_t r0;          // first global column of the reference front
    int32_t nn, na;
    int32_t sepoff;      // offset in septpos[] of this front's separator rows
    int32_t pad;
};

// ---- selected inversion (k_selinv, mf_selinv.hip) ----
// Buffers a selected-inversion task reads or writes (index into DevSel::base):
// the factor's panel and W pools, per big front X = L11^-1 (nn x nn, ld nn),
// the recursion's scratch T (same layout), Lt = L21 X (na x nn, ld na), the
// gathered update matrix M (na x na, ld na, both triangles) and the output
// panel pool Sig (the panel layout, holding [Sigma11; Sigma21]).
enum : int8_t { SP_PANEL = 0, SP_W = 1, SP_X = 2, SP_T = 3, SP_LT = 4, SP_M = 5, SP_SIG = 6,
                SP_N = 7 };
enum : int32_t { ST_XDIAG = 0, ST_GEMM = 1, ST_GATHER = 2, ST_SMALL = 3, ST_KINDS = 4 };
// One product of a tile task: C += sgn * A B' over k, where A(r, k) sits at
// base[arrA] + a_off + r + k * lda (ta = 0) or + k + r * lda (ta = 1) and
// B(c, k) at base[arrB] + b_off + c + k * ldb (tb = 0) or + k + c * ldb.
struct SelPart {
    int64_t a_off, b_off;
    int32_t lda, ldb;
    int32_t k;
    int8_t arrA, arrB, ta, tb;
    int32_t sgn;  // +1 or -1
    int32_t pad;
};
struct SelTask {
    SelPart p[2];     // ST_GEMM: np products summed into one m x n tile
    int64_t c_off;    // ST_GEMM: C(r, c) at base[arrC] + c_off + r + c * ldc
    int32_t ldc, m, n;
    int8_t arrC, np, lower, pad;  // lower: store r >= c only (diagonal tile)
    int32_t kind;
    int32_t front;
    int32_t a;        // ST_XDIAG: diagonal tile; ST_GATHER: column tile of M
    int32_t b;
    int32_t wait[2];  // counters that must reach target (-1: none): the front's
    int32_t target[2];//   own earlier phases, the parent's completion
};

}  // namespace mf
