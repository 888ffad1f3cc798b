// lsg_plan.hpp -- host-side flattening of a LoAdSG sparse grid into the
// depth-blocked layout the gfx950 kernels work on, and the per-stage work-item
// tables of the matrix-free prewavelet matvec.
//
// Data layout in HBM (one copy per batch instance, instance-major):
//   vector      : one block per depth T (depth list order), n_T = prod 2^(T_i-1)
//                 entries, lexicographic in the local odd positions, dim 0 fastest.
//   nodal       : one dense box per depth T, |G_T| = prod (2^T_i - 1) entries: the
//                 level-T full grid without boundary, dim 0 fastest (the reference's
//                 MultiLevelVector restricted to G_T, multilevelvector.h:72-144).
//   case buffer : per case c with prolongation dims, one box per depth of the case.
//   W buffer    : per (apply item, restriction step) whose value a later item reads.
//   part buffer : per (case, target depth) block of case-partial outputs.
#pragma once
#include <array>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace lsg {

constexpr int MAXD = 8;      // maximal dimension
constexpr int SP_CELL_MAXD = 4;   // sparse apply: cell-wise element products up to this dimension
constexpr int MAXLEV = 15;   // maximal 1D depth (MaximumDepth in library/source/abbrevi.h:12)

using Depth = std::array<int, MAXD>;

// ---- work items shared with the device (plain structs) ----------------------
// calcNodal: depth block -> nodal box.
struct NodalItem {
    int32_t T[MAXD];
    int64_t in_off;    // vector block offset
    int64_t out_off;   // nodal box offset
};

// Prolongation prefix sweep along one dimension for a family of depths that
// differ only in that dimension (T[dd] = t0, t0+1, ...).
struct ProlItem {
    int32_t T[MAXD];   // depth of the first box
    int32_t dd;        // sweep dimension
    int32_t nt;        // number of boxes
    int32_t from_nodal;  // bit q: box q's input is in the nodal buffer (else the case buffer)
    int32_t pad;
    int64_t in_off[MAXLEV];
    int64_t out_off[MAXLEV];
    int32_t tidx[MAXLEV];   // depth index of each box
};

// Stencil + restriction chain + prewavelet conversion for one (case, depth).
struct ApplyItem {
    int32_t T[MAXD];
    int32_t k;            // number of restriction dims
    int32_t r[MAXD];      // restriction dims, ascending
    int32_t in_nodal;     // 1: input box from nodal buffer, 0: from case buffer
    int32_t tidx;         // depth index of T
    int64_t in_off;
    int64_t rd_off[MAXD + 1];   // gM values added after step j (0 = after the stencil); -1 none
    int64_t w_off[MAXD + 1];    // W_j written after restriction j (1..k); -1 if never read
    int64_t out_off;            // part buffer block of the target depth
    int64_t z_off;              // scratch of the sparse apply stage (sparse grids only)
    int32_t ctidx[MAXD + 1];    // depth index after j restrictions (j = 0..k)
    int32_t shp;                // first of the k + 1 box shapes (Plan::shapes) of this item
};

// Shape of a dense box (dimension 0 fastest) with the reciprocals of its extents and
// strides for the exact float division of point indices (correctly rounded 1/x).
struct BoxShape {
    int32_t d;
    int32_t total;
    int32_t n[MAXD];      // extent per dim
    int32_t S[MAXD];      // stride per dim
    float rn[MAXD];       // 1/n
    float rS[MAXD];       // 1/S
};

// A family of apply items along the last prolongation dimension fd of a case
// (depths T, T+e_fd, ...): one workgroup prolongs along fd box by box and applies
// each box.  fd = -1: a single item (case without prolongation dims).
struct FamilyItem {
    int32_t fd;
    int32_t nt;
    int32_t first;      // index of the first ApplyItem
    int32_t pad;
};

// Sum of the case partials of one depth block (fixed case order).
struct SumItem {
    int32_t T[MAXD];
    int32_t ncase;
    int32_t list_start;
    int64_t blk_off;
    int64_t count;      // points in the block
};

struct Plan {
    std::vector<NodalItem> nodal;
    std::vector<std::vector<ProlItem>> prol;        // per sweep dimension
    std::vector<ApplyItem> applyItems;              // all apply items, family by family
    std::vector<std::vector<FamilyItem>> apply;     // per wavefront level
    std::vector<SumItem> sum;
    std::vector<int64_t> sumList;                   // part offsets for SumItem
    std::vector<BoxShape> shapes;                   // per apply item: the box after 0..k restrictions
    int64_t caseSize = 0;
    int64_t wSize = 0;
    int64_t partSize = 0;
    int64_t zSize = 0;          // scratch of the sparse apply stage
    int64_t maxBox = 0;         // largest |G_T|
    int64_t maxBlock = 0;       // largest n_T
    int64_t maxFamily = 0;      // largest box in a prolongation family
    int64_t nApply = 0;
    double streamValues = 0;    // streaming-model values per matvec
    double streamApplyValues = 0;   // of which the stages the apply kernel performs (last prolongation
                                    // dim, stencil, restrictions with gM, prewavelet conversion)
    double movedValues = 0;     // values this implementation reads+writes per matvec
    double applyValues = 0;     // of which the apply stage (input box, gM reads, W writes, output block)
};

// Gather tables of the sparse path (grids whose depth blocks are not full, or with
// hanging nodes): G_T is the set of stored points of depth <= T (the reference's
// MultiLevelAdaptiveSparseGrid, multilevelSparseGrid.h:13-143), kept as a list in slot
// order; every stage becomes a gather over such lists.
struct SparseTables {
    // calcNodal: CSR over multilevel rows (boxOff[t] + i): sources (device index, coefficient)
    std::vector<int64_t> nodPtr;
    std::vector<int32_t> nodCol;
    std::vector<double> nodVal;
    // prolongation into depth t along dd from t - e_dd: per target row (a, b) local
    // indices in G_{t-e_dd}; b == -2: copy a; else 0.5 (src[a] + src[b]) (-1 = absent)
    std::vector<int64_t> prolOff;          // [t * d + dd], -1 if t - e_dd is not a depth
    std::vector<int32_t> prolAB;           // pairs
    // cells of depth T+1 in the grid (cells/CellStructure.h:23-61), per depth: corner local
    // indices (2^d per cell, -1 = boundary corner) and the cell's lower position per dim
    std::vector<int64_t> cellOff;          // [t], first cell of depth t
    std::vector<int32_t> cellCorner;
    std::vector<int32_t> cellLo;
    // cells of each depth by colour (parity of the lower position, bit i = dimension i):
    // cellCol[cellColOff[t * (2^d + 1) + col] ..] are the local cell indices of colour col
    std::vector<int64_t> cellColOff;
    std::vector<int32_t> cellCol;
    // cells around each multilevel row: CSR (cell id, corner id of the row's point)
    std::vector<int64_t> rcPtr;
    std::vector<int32_t> rcCell;
    std::vector<int32_t> rcCorner;
    std::vector<int32_t> rcY;              // cell * 2^d + corner: the row's entry of the cell results
    // restriction from depth t along dd into t - e_dd: per target row (self, left, right)
    // local indices in G_t (-1 = absent)
    std::vector<int64_t> restOff;          // [t * d + dd]
    std::vector<int32_t> restSLR;          // triples
    // ConvertToPrewavelet2: CSR over device rows (points of depth t in block order):
    // (local index in G_t, coefficient); hanging nodes have empty rows
    std::vector<int64_t> convPtr;
    std::vector<int32_t> convCol;
    std::vector<double> convVal;
    // level-T_own position of each point (device order), for the diagonal
    std::vector<int32_t> pos;
    std::vector<int32_t> tOfDev;     // depth index of each point (device order)
    std::vector<int32_t> depthT;     // depths, nd x d
};

struct Grid {
    bool sparse = false;   // sparse path (see SparseTables)
    SparseTables sp;
    int d = 0;
    int64_t n = 0;
    std::vector<int32_t> idx;      // n x d, reference slot order
    std::vector<uint8_t> active;
    std::vector<Depth> depths;     // DepthList order (first appearance)
    std::map<Depth, int> depthIdx;
    std::vector<int64_t> blkOff, blkSize;   // vector blocks per depth
    std::vector<int64_t> boxOff, boxSize;   // nodal boxes per depth
    int64_t nml = 0;
    std::vector<int64_t> devOfSlot;
    std::vector<int64_t> slotOfDev;
    Depth maxDepth{};
    int maxLevel = 0;              // largest 1D depth
    Plan plan;
    // dense grids: the plan with every prolongation dimension materialised and single-item
    // apply families (buildSinglePlan; for small batches)
    std::shared_ptr<Plan> planSingle;
};
// Builds g.planSingle (dense grids only).
int buildSinglePlan(Grid& g, std::string& err);

// Basis transformations: one item per depth, its level-T box G_T (every position of the
// level-T full grid; on adaptive grids some positions are not stored).
struct TrItem {
    int32_t T[MAXD];
    int32_t n[MAXD];      // box extents 2^T_i - 1
    int32_t S[MAXD];      // box strides (dimension 0 fastest)
    int32_t d;
    int32_t total;        // |G_T|
    int64_t box_off;      // first entry of the box in the gather table and the scratch
};

// Adaptive grids: one connected component of the points of one depth T
// (ZusammenhangsKomponente, sgrid/komponente.cpp) with its box and the factors of the
// prewavelet line matrix along every dimension.
struct TrComp {
    int32_t mn[MAXD], mx[MAXD];   // box corners, level-T positions (1-based)
    int32_t size[MAXD];           // line length along each dimension (from the lower corner)
    int64_t fac[MAXD];            // first entry of that line's factors in TrTables::afac
};

// box entry status on adaptive grids
enum : uint8_t { TR_MISSING = 0, TR_HANGING = 1, TR_ACTIVE = 2 };

struct TrTables {
    bool adaptive = false;                     // grid with hanging nodes or incomplete boxes
    std::vector<TrItem> items;                 // per depth, depth list order
    std::vector<int32_t> gidx;                 // per box entry: device index of its point (-1: not stored)
    std::vector<int32_t> levelList;            // depth ids, |T|_1 descending
    std::vector<std::array<int32_t, 2>> levels;   // (first in levelList, count) per |T|_1, descending
    std::vector<int64_t> levelMaxBox;          // largest |G_T| per level
    // hierarchical parents (poles) of every point along every dimension, device indices:
    // poles[(i * n + p) * 2 + {0, 1}] = left / right pole of point p along dim i, -1 on
    // the boundary
    std::vector<int32_t> poles;
    // points by their depth along each dimension: byLevel[i * n + k] for
    // k in [lvlOff[i * (MAXLEV + 2) + t], lvlOff[i * (MAXLEV + 2) + t + 1]) have depth t along dim i
    std::vector<int32_t> byLevel;
    std::vector<int64_t> lvlOff;
    // Scatter-add of box results onto points, as a gather per point: CSR whose rows are
    // device indices and whose entries are scratch positions.  Row groups: one per
    // |T|_1 (same order as levels), then one over all depths.
    std::vector<int32_t> rowDev;
    std::vector<int64_t> rowPtr;               // rows + 1
    std::vector<int32_t> col;
    std::vector<std::array<int64_t, 2>> groups;   // (first row, number of rows)
    // LU factors (no pivoting) of the line matrices of levels 1..maxLevel:
    // fac[a * facTot + facBase(t) + k], a = 0..4 for l1, l2, u0, u1, u2
    std::vector<double> fac;
    int64_t facTot = 0;
    int64_t maxBox = 0;
    int64_t boxTot = 0;                        // sum of the box sizes (scratch per instance)
    // adaptive grids only
    std::vector<uint8_t> bstat;                // per box entry: TR_MISSING / TR_HANGING / TR_ACTIVE
    std::vector<int32_t> seq;                  // depth ids in the reference's transformation order
    std::vector<int32_t> compOff;              // per depth t: components compOff[t] .. compOff[t + 1]
    std::vector<TrComp> comps;
    std::vector<double> afac;                  // l1, l2, u0, u1, u2 arrays of afacTot entries each
    int64_t afacTot = 0;
};
inline int64_t facBase(int t) { return (int64_t(1) << t) - t - 1; }
// Tables of the basis transformations (dense boxes on regular grids, component-wise line
// solves on adaptive ones).
int buildTransform(const Grid& g, TrTables& tr, std::string& err);

// Build the grid from a reference slot-order index table.  Returns 0 or a
// negative status with a message in err.  forceSparse: build the gather tables of the
// sparse path even for a grid with complete depth blocks (cell-wise coefficients); such a
// grid keeps the dense device order.
int buildGrid(int d, int64_t n, const int32_t* idx, const uint8_t* active, Grid& g, std::string& err,
              bool forceSparse = false);
// 1D prewavelet of odd position k at level t as (position, coefficient) pairs of level-t
// nodal hats (IndexDimension::nextFive, indices/index.cc:180-327); returns the count.
int prewaveletStencil1D(int32_t k, int t, int32_t pos[5], double c[5]);
// Reference slot order of AddRecursiveSonsOfPoint(center, level).
void regularIndexTable(int d, int level, std::vector<int32_t>& idx);

// Depth-block preconditioner of the inner (mixed-precision) solves: for every depth block
// T the Galerkin block of the constant-coefficient Laplacian on the prewavelets of depth T,
// B_T = sum_i K_(T_i) (x) prod_(j != i) M_(T_j) with the 1D prewavelet stiffness / mass
// matrices of level t (no counterpart in the reference, whose CG uses the diagonal).  On
// adaptive grids the block is taken over the active points of depth T (hanging points: the
// identity, as the operator's diagonal is 1 there).  Blocks of at most maxDense points keep
// their inverse.  Larger ones whose active points form a product set S_1 x ... x S_d of 1D
// positions (every depth block of a dense grid; the blocks of grids refined towards a corner)
// use the fast diagonalisation of the Kronecker sum: with V_i^T M_i V_i = I and
// V_i^T K_i V_i = diag(lam_i) for the restrictions K_i = K_(T_i)[S_i, S_i],
// M_i = M_(T_i)[S_i, S_i], B^-1 = (x)_i V_i diag(1 / sum_i lam_i) (x)_i V_i^T; up to
// FD_MAXM positions per dimension and FD_MAXN points.  Other blocks keep the diagonal.
constexpr int FD_MAXM = 256;      // 1D eigenproblems up to 256 x 256
constexpr int FD_MAXN = 2048;     // largest product set on the fast-diagonalisation path
enum PreMode : int32_t { PRE_DIAG = 0, PRE_DENSE = 1, PRE_FD = 2 };
struct PreBlock {
    int64_t off;      // first point of the block (device order)
    int32_t n;        // points of the block
    int32_t mode;     // PreMode
    int32_t inv;      // PRE_DENSE: first entry of the inverse in BlockPre::inv (n x n, column k
                      // contiguous); PRE_FD: first entry of the block's point order in
                      // BlockPre::fdPerm, or -1 when it is the identity
    int32_t np;       // PRE_FD: points of the product set (the first np of the point order;
                      // the other n - np are hanging points)
    int32_t m[MAXD];  // PRE_FD: |S_i| (the product set is ordered dimension 0 fastest)
    int32_t fv[MAXD]; // PRE_FD: V_i at fdV[fv[i]] (row j = position j of S_i, column k =
                      // eigenvector k)
    int32_t fl[MAXD]; // PRE_FD: lam_i at fdLam[fl[i]]
};
struct BlockPre {
    std::vector<PreBlock> blocks;   // depth-list order
    std::vector<double> inv;        // block inverses
    std::vector<double> d0;         // diagonal of the constant-coefficient blocks, device order
                                    // (1 at hanging points)
    int maxN = 0;                   // largest block with an inverse
    std::vector<double> fdV, fdLam; // 1D factors of the fast diagonalisation
    std::vector<int32_t> fdList;    // blocks on the fast-diagonalisation path
    std::vector<int32_t> fdPerm;    // point orders (offsets inside the block)
    int maxFd = 0;                  // largest product set
};
int buildBlockPre(const Grid& g, int maxDense, BlockPre& out, std::string& err);

inline int64_t boxCount(const int32_t* T, int d) {
    int64_t s = 1;
    for (int i = 0; i < d; ++i) s *= (int64_t(1) << T[i]) - 1;
    return s;
}
inline int64_t blockCount(const int32_t* T, int d) {
    int64_t s = 1;
    for (int i = 0; i < d; ++i) s *= int64_t(1) << (T[i] - 1);
    return s;
}

}  // namespace lsg
