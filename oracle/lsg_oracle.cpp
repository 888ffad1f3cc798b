// lsg_oracle.cpp -- CPU restatement of LoAdSG's matrix-free prewavelet matvec,
// diagonal preconditioner and CG on sparse grids with zero Dirichlet boundaries.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke()
// entry point and bench.py's cpu_baseline leg may load it.  It is never on the
// product path.  It follows the reference algorithm step by step, in plain C++
// with hash maps instead of the reference's prime/second tables, in the
// reference's own slot order (the order of VectorSparseG's data table).
//
// Each function cites the reference file:line it restates (paths relative to
// library/source/).  Parity status of each part (how it is pinned) is listed at
// the bottom of this header:
//   * index algebra (Idx1 helpers)       : exhaustive closed-form tests + golden
//   * regular grid slot order            : golden index tables from the compiled reference
//   * calcNodal / caseFunction / Convert : golden matvecs (Poisson, HelmHoltz, variable
//                                          coefficient) from the compiled reference, plus
//                                          dense brute force Psi^T K Psi for constant
//                                          coefficients (exact by prewavelet semi-orthogonality)
//   * precond3 diagonal                  : golden diagonal + brute force psi^T A_T psi
//   * solveHomogen                       : golden iteration counts and solutions
//   * basis transformations (Trafo)      : golden CalcHierarchicalBasis / calcPrewByNodal /
//                                          calcNodalByPrew / H1 and L2 norms on regular and
//                                          adaptive grids, brute-force basis evaluation on
//                                          regular grids, round trips on adaptive grids
//
// Build: see oracle/Makefile (g++ -O2 -fopenmp -shared).
#include <cstdint>
#include <cstring>
#include <cmath>
#include <vector>
#include <array>
#include <unordered_map>
#include <algorithm>
#include <list>
#include <memory>
#include <stdexcept>

namespace orc {

constexpr int MAXD = 8;

// ---------------------------------------------------------------------------
// 1D index algebra (indices/index.h:657-876).  Index 0 = left boundary (x=0),
// 1 = right boundary (x=1), 2 = 0.5, sons 2i / 2i+1.
// ---------------------------------------------------------------------------
static inline int depth1(int32_t i) {            // index.h:812-817
    if (i == 0) return 0;
    return 31 - __builtin_clz((uint32_t)i);
}
static inline int internDepth1(int32_t i) {      // index.h:825-834
    if (i == 0) return -1;
    return 31 - __builtin_clz((uint32_t)i);
}
static inline double coord1(int32_t i) {         // index.h:718-730
    if (i == 0) return 0.0;
    int t = depth1(i);
    double h = 1.0 / double(1u << t);
    int32_t n = i & ~(1 << t);
    return double(2 * n + 1) * h;
}
static inline int32_t poleLeft(int32_t i) {      // index.h:749-755
    return i >> (__builtin_ctz((uint32_t)i) + 1);
}
static inline int32_t poleRight(int32_t i) {     // index.h:782-789
    return i >> (__builtin_ctz(~(uint32_t)i) + 1);
}
static inline int32_t leftAtLevel(int32_t i, int level) {   // index.h:839-851
    int my = depth1(i);
    int l = level - my;
    if (l == 0) return poleLeft(i);
    i = i << 1;
    for (int k = 1; k < l; ++k) i = (i << 1) + 1;
    return i;
}
static inline int32_t rightAtLevel(int32_t i, int level) {  // index.h:859-871
    int my = internDepth1(i);
    int l = level - my;
    if (l == 0) return poleRight(i);
    i = (i << 1) + 1;
    for (int k = 1; k < l; ++k) i = i << 1;
    return i;
}
static inline bool isBoundary1(int32_t i) { return i == 0 || i == 1; }

// 1D prewavelet of a point at depth t (its own level), as nodal values at
// level t: nextFiveP (indices/index.cc:252-327) restricted to one dimension.
// The multi-dimensional 5^d stencil of nextFiveP is the tensor product of
// these lists (the compass skip logic only drops entries whose step would hit
// the boundary, see index.cc:264-310).
struct Entry1 { int32_t idx; double c; };
static int prewavelet1D(int32_t i, int t, Entry1 out[5]) {
    int n = 0;
    // Mitte: 0.9 if depth > 1 and next to the boundary (index.cc:312-316)
    double cm = 1.0;
    if (depth1(i) > 1) {
        if (poleLeft(i) == 0 || poleRight(i) == 1) cm = 0.9;
    }
    out[n++] = {i, cm};
    int32_t l = poleLeft(i);
    if (!isBoundary1(l)) {                       // Links (index.cc:264-276)
        out[n++] = {l, -0.6};
        out[n++] = {leftAtLevel(l, t), 0.1};     // LinksLinks (index.cc:278-288)
    }
    int32_t r = poleRight(i);
    if (!isBoundary1(r)) {                       // Rechts / RechtsRechts (index.cc:289-310)
        out[n++] = {r, -0.6};
        out[n++] = {rightAtLevel(r, t), 0.1};
    }
    return n;
}

// ---------------------------------------------------------------------------
// Point keys and depths
// ---------------------------------------------------------------------------
struct Key {
    std::array<int32_t, MAXD> v{};
    bool operator==(const Key& o) const { return v == o.v; }
};
struct KeyHash {
    size_t operator()(const Key& k) const {
        uint64_t h = 1469598103934665603ull;
        for (int i = 0; i < MAXD; ++i) { h ^= (uint32_t)k.v[i]; h *= 1099511628211ull; }
        return (size_t)h;
    }
};
using Depth = std::array<int, MAXD>;
struct DepthHash {
    size_t operator()(const Depth& k) const {
        uint64_t h = 1469598103934665603ull;
        for (int i = 0; i < MAXD; ++i) { h ^= (uint32_t)k[i]; h *= 1099511628211ull; }
        return (size_t)h;
    }
};

// Coefficient families shared with the product library (same formulas, evaluated
// here directly, not through the product's separable tables).
enum CoefKind { COEF_CONST = 0, COEF_CFG2 = 1, COEF_CFG3 = 2, COEF_EXPSIN = 3, COEF_GEN = 10 };
// COEF_GEN: a non-separable pair, a = 1 + x1 x2 + sin(pi (x1 + x2)), c = 1 + sin(pi x1 x2) / 2
// (x2 = x1 when d = 1); the product library receives it sampled per cell.
struct Coef {
    int kind = COEF_CONST;
    double a0 = 1.0, c0 = 0.0;   // COEF_CONST: a = a0, c = c0
    std::vector<double> xi;      // COEF_EXPSIN: a = exp(sum_k xi_k sin(k pi x_{k mod d}) / k)
    double a(const double* x, int d) const {
        switch (kind) {
            case COEF_CONST: return a0;
            case COEF_CFG2: { double p = 1.0; for (int i = 0; i < d; ++i) p *= std::cos(M_PI * x[i]); return 1.0 + 0.5 * p; }
            case COEF_CFG3: { double s = 1.0; for (int i = 0; i < d; ++i) s += x[i] * x[i]; return s; }
            case COEF_EXPSIN: {
                double s = 0.0;
                for (size_t k = 1; k <= xi.size(); ++k)
                    s += xi[k - 1] * std::sin(double(k) * M_PI * x[k % d]) / double(k);
                return std::exp(s);
            }
            case COEF_GEN: return 1.0 + x[0] * x[1 % d] + std::sin(M_PI * (x[0] + x[1 % d]));
        }
        return 0.0;
    }
    double c(const double* x, int d) const {
        switch (kind) {
            case COEF_CONST: return c0;
            case COEF_CFG2: return 1.0 + x[0];
            case COEF_GEN: return 1.0 + 0.5 * std::sin(M_PI * x[0] * x[1 % d]);
            default: return 0.0;
        }
    }
};

// ---------------------------------------------------------------------------
// Grid (sgrid/sparseGrid.h:86-387): points in slot order + active flags.
// ---------------------------------------------------------------------------
struct Grid {
    int d = 0;
    std::vector<Key> pts;             // slot -> point
    std::vector<uint8_t> active;      // slot -> active (false = hanging node)
    std::unordered_map<Key, int64_t, KeyHash> slot;
    std::vector<Depth> depthOf;       // slot -> depth
    // DepthList (iterator/depthIterator.h:16-48): depths in first-appearance order
    std::vector<Depth> depths;
    std::unordered_map<Depth, int, DepthHash> depthIdx;
    std::vector<std::vector<int64_t>> slotsOfDepth;   // MultiDepthHashGrid::_mapPosToGridPos
    Depth maxDepth{};

    bool occupied(const Key& k, int64_t& s) const {
        auto it = slot.find(k);
        if (it == slot.end()) return false;
        s = it->second;
        return true;
    }
    void build() {
        slot.clear(); depthOf.resize(pts.size());
        for (int64_t s = 0; s < (int64_t)pts.size(); ++s) {
            slot[pts[s]] = s;
            Depth T{};
            for (int i = 0; i < d; ++i) T[i] = depth1(pts[s].v[i]);
            depthOf[s] = T;
            auto it = depthIdx.find(T);
            if (it == depthIdx.end()) {
                depthIdx[T] = (int)depths.size();
                depths.push_back(T);
                slotsOfDepth.emplace_back();
            }
            slotsOfDepth[depthIdx[T]].push_back(s);
        }
        maxDepth = Depth{};
        for (auto& T : depths) for (int i = 0; i < d; ++i) maxDepth[i] = std::max(maxDepth[i], T[i]);
    }
    bool included(const Depth& T) const { return depthIdx.count(T) != 0; }
};

// AddRecursiveSonsOfPoint(center, level) (sgrid/sparseGrid.cc:557-566): DFS,
// slot = first-visit order.  Every path to a point has the same length, so a
// revisit explores an identical subtree and can be skipped without changing the
// first-visit order.
static void addRecursiveSons(Grid& g, const Key& I, int level,
                             std::unordered_map<Key, char, KeyHash>& seen) {
    if (seen.count(I)) return;
    seen[I] = 1;
    g.pts.push_back(I);
    g.active.push_back(1);
    if (level > 0) {
        for (int dd = 0; dd < g.d; ++dd) {
            Key L = I; L.v[dd] = I.v[dd] << 1;
            addRecursiveSons(g, L, level - 1, seen);
            Key R = I; R.v[dd] = (I.v[dd] << 1) + 1;
            addRecursiveSons(g, R, level - 1, seen);
        }
    }
}

// ---------------------------------------------------------------------------
// Multilevel index (sgrid/multilevelSparseGrid.cpp:9-34): for each depth T of
// the list, G_T = all grid points p with depth(p) <= T componentwise.
// ---------------------------------------------------------------------------
struct MultiLevel {
    std::vector<std::vector<int64_t>> slotsG;                       // per depth: slots of G_T
    std::vector<std::unordered_map<int64_t, int64_t>> posInG;       // per depth: slot -> position
    void build(const Grid& g) {
        size_t nd = g.depths.size();
        slotsG.assign(nd, {}); posInG.assign(nd, {});
        for (size_t t = 0; t < nd; ++t) {
            const Depth& T = g.depths[t];
            // GetNextSmallerDepthsIterator order does not matter for values.
            for (int64_t s = 0; s < (int64_t)g.pts.size(); ++s) {
                bool le = true;
                for (int i = 0; i < g.d; ++i) if (g.depthOf[s][i] > T[i]) { le = false; break; }
                if (le) { posInG[t][s] = (int64_t)slotsG[t].size(); slotsG[t].push_back(s); }
            }
        }
    }
};
using MLVec = std::vector<std::vector<double>>;   // per depth: values on G_T

// ---------------------------------------------------------------------------
// Operator: Q1 element matrices on a cell (stencils/Stencil.h)
// ---------------------------------------------------------------------------
struct Op {
    Coef coef;
    // integration(cell, dirP, dirQ): Poisson (Stencil.h:1707-1752) times a(center)
    // plus HelmHoltz (Stencil.h:900-919) times c(center).
    double element(int d, const double* h, const double* center, unsigned p, unsigned q) const {
        double a = coef.a(center, d), c = coef.c(center, d);
        double stiff = 0.0;
        if (a != 0.0) {
            for (int i = 0; i < d; ++i) {
                double prod = (((p >> i) & 1u) == ((q >> i) & 1u)) ? 1.0 / h[i] : -1.0 / h[i];
                for (int j = 0; j < d; ++j) {
                    if (j == i) continue;
                    prod *= (((p >> j) & 1u) == ((q >> j) & 1u)) ? h[j] / 3.0 : h[j] / 6.0;
                }
                stiff += prod;
            }
        }
        double mass = 0.0;
        if (c != 0.0) {
            mass = 1.0;
            for (int j = 0; j < d; ++j)
                mass *= (((p >> j) & 1u) == ((q >> j) & 1u)) ? h[j] / 3.0 : h[j] / 6.0;
        }
        return a * stiff + c * mass;
    }
};

struct Cell { Key lo; };   // lower-left corner index at level T (may be boundary index 0)

// Cells of depth T+1 that are in the grid: all non-boundary corners occupied
// (cells/CellStructure.h:23-61).  Enumerated per level-T full grid.
static void cellsOfDepth(const Grid& g, const Depth& T, std::vector<std::array<int32_t, MAXD>>& cornersLoHi,
                         std::vector<std::array<double, MAXD>>& centers) {
    int d = g.d;
    // 1D: level-T grid nodes at positions 0..2^t (0 and 2^t = boundary)
    std::vector<std::vector<int32_t>> nodes(d);
    for (int i = 0; i < d; ++i) {
        int t = T[i];
        int n = 1 << t;
        nodes[i].resize(n + 1);
        nodes[i][0] = 0; nodes[i][n] = 1;
        for (int k = 1; k < n; ++k) {
            // position k at level t -> index: depth s = t - ctz(k), n' = k >> (ctz(k)+1)
            int tz = __builtin_ctz((unsigned)k);
            int s = t - tz;
            int32_t local = k >> (tz + 1);
            nodes[i][k] = (1 << s) + local;
        }
    }
    std::vector<int> cnt(d, 0);
    while (true) {
        // cell with lower corner position cnt[i]
        bool ok = true;
        for (unsigned c = 0; c < (1u << d) && ok; ++c) {
            Key K{};
            bool interior = true;
            for (int i = 0; i < d; ++i) {
                int pos = cnt[i] + ((c >> i) & 1u);
                K.v[i] = nodes[i][pos];
                if (isBoundary1(K.v[i])) interior = false;
            }
            int64_t s;
            if (interior && !g.occupied(K, s)) ok = false;
        }
        if (ok) {
            std::array<int32_t, MAXD> lo{};
            std::array<double, MAXD> ctr{};
            for (int i = 0; i < d; ++i) {
                lo[i] = cnt[i];
                ctr[i] = (double(cnt[i]) + 0.5) / double(1 << T[i]);
            }
            cornersLoHi.push_back(lo);
            centers.push_back(ctr);
        }
        int i = 0;
        for (; i < d; ++i) {
            if (++cnt[i] < (1 << T[i])) break;
            cnt[i] = 0;
        }
        if (i == d) break;
    }
}

struct Oracle {
    Grid g;
    MultiLevel ml;
    Op op;
    // cached cells per depth index (cells of depth T+1)
    std::vector<std::vector<std::array<int32_t, MAXD>>> cellLo;
    std::vector<std::vector<std::array<double, MAXD>>> cellCtr;
    std::vector<std::vector<std::vector<int32_t>>> levelNodes;   // per depth, per dim, position -> index
    // Stored element matrices (LocalStiffnessMatrices, LocalStiffnessMatrices.cpp:453-531),
    // when set: per depth, per cell of cellLo, the 2^(d-1) (2^d + 1) entries of corners p <= q
    // in the order of the nested CellIndexIterator loops (cells/iterator_neu.cpp:34-49)
    std::vector<std::vector<double>> mats;
    std::vector<std::unordered_map<Key, int64_t, KeyHash>> cellOf;   // per depth: lower position -> cell
    int64_t nEnt() const { return (int64_t(1) << (g.d - 1)) * ((int64_t(1) << g.d) + 1); }
    static int entryOf(unsigned p, unsigned q, unsigned nc) {
        if (p > q) std::swap(p, q);
        return (int)(p * nc - (p * (p - 1)) / 2 + (q - p));
    }

    void init() {
        g.build();
        ml.build(g);
        size_t nd = g.depths.size();
        cellLo.assign(nd, {}); cellCtr.assign(nd, {}); levelNodes.assign(nd, {});
        for (size_t t = 0; t < nd; ++t) {
            cellsOfDepth(g, g.depths[t], cellLo[t], cellCtr[t]);
            cellOf.emplace_back();
            for (size_t c = 0; c < cellLo[t].size(); ++c) {
                Key K{};
                for (int i = 0; i < g.d; ++i) K.v[i] = cellLo[t][c][i];
                cellOf[t][K] = (int64_t)c;
            }
            levelNodes[t].resize(g.d);
            for (int i = 0; i < g.d; ++i) {
                int L = g.depths[t][i];
                int n = 1 << L;
                auto& nodes = levelNodes[t][i];
                nodes.resize(n + 1);
                nodes[0] = 0; nodes[n] = 1;
                for (int k = 1; k < n; ++k) {
                    int tz = __builtin_ctz((unsigned)k);
                    nodes[k] = (1 << (L - tz)) + (k >> (tz + 1));
                }
            }
        }
    }
    int64_t N() const { return (int64_t)g.pts.size(); }

    // tensor-product prewavelet stencil of point s at its own depth
    void stencil5(int64_t s, std::vector<std::pair<Key, double>>& out) const {
        int d = g.d;
        Entry1 e[MAXD][5];
        int ne[MAXD];
        for (int i = 0; i < d; ++i) ne[i] = prewavelet1D(g.pts[s].v[i], g.depthOf[s][i], e[i]);
        out.clear();
        std::vector<int> c(d, 0);
        while (true) {
            Key J{};
            double v = 1.0;
            for (int i = 0; i < d; ++i) { J.v[i] = e[i][c[i]].idx; v *= e[i][c[i]].c; }
            out.push_back({J, v});
            int i = 0;
            for (; i < d; ++i) { if (++c[i] < ne[i]) break; c[i] = 0; }
            if (i == d) break;
        }
    }

    // calcNodal(nodal_u, prew, T) (MatrixVectorMultiplication/MatrixVectorHomogen.cpp:43-74)
    // followed by nodal.setMultiLevelValuesOMP (extemp/multilevelvector.h:200-223).
    void calcNodal(const std::vector<double>& x, MLVec& nodal) const {
        std::vector<std::pair<Key, double>> st;
        for (size_t t = 0; t < g.depths.size(); ++t) {
            std::vector<double> un(N(), 0.0);
            for (int64_t s : g.slotsOfDepth[t]) {
                if (!g.active[s]) continue;
                double coeff = x[s];
                stencil5(s, st);
                for (auto& [J, c] : st) {
                    int64_t k;
                    if (g.occupied(J, k)) un[k] += coeff * c;
                }
            }
            nodal[t].assign(ml.slotsG[t].size(), 0.0);
            for (size_t p = 0; p < ml.slotsG[t].size(); ++p) nodal[t][p] = un[ml.slotsG[t][p]];
        }
    }

    // DepthList::sortDepths / recursiveDepth (iterator/depthIterator.cpp:7-41,88-93)
    void recursiveDepth(const bool* R, bool checkProl, int dd, Depth& T, std::vector<int>& out) const {
        int D = g.d;
        if (checkProl && dd > 0) {
            if (!R[dd]) {
                for (int t = 1; t <= g.maxDepth[dd]; ++t) { T[dd] = t; recursiveDepth(R, checkProl, dd - 1, T, out); }
            } else recursiveDepth(R, checkProl, dd - 1, T, out);
        }
        if (checkProl && dd == 0) {
            if (!R[dd]) {
                for (int t = 1; t <= g.maxDepth[dd]; ++t) { T[dd] = t; recursiveDepth(R, false, D - 1, T, out); }
            } else recursiveDepth(R, false, D - 1, T, out);
        }
        if (!checkProl && dd >= 0) {
            if (R[dd]) {
                for (int t = g.maxDepth[dd]; t > 1; --t) { T[dd] = t; recursiveDepth(R, checkProl, dd - 1, T, out); }
            } else recursiveDepth(R, checkProl, dd - 1, T, out);
        }
        if (dd < 0) {
            auto it = g.depthIdx.find(T);
            if (it != g.depthIdx.end()) out.push_back(it->second);
        }
    }
    std::vector<int> sortDepths(const bool* R) const {
        std::vector<int> out;
        Depth T{};
        recursiveDepth(R, true, g.d - 1, T, out);
        return out;
    }

    // slot-vector <-> multilevel helpers (extemp/vector.cc:600-675, multilevelvector.h:146-298)
    void setFromML(std::vector<double>& v, const MLVec& a, int t) const {     // setMultiLevelValues2
        for (size_t p = 0; p < ml.slotsG[t].size(); ++p) v[ml.slotsG[t][p]] = a[t][p];
    }
    void addFromML(std::vector<double>& v, const MLVec& a, int t) const {     // addMultiLevelValues2
        for (size_t p = 0; p < ml.slotsG[t].size(); ++p) v[ml.slotsG[t][p]] += a[t][p];
    }
    void mlAdd(MLVec& a, const std::vector<double>& v, int t) const {         // MultiLevelVector::addMultiLevelValues
        for (size_t p = 0; p < ml.slotsG[t].size(); ++p) a[t][p] += v[ml.slotsG[t][p]];
    }
    void mlSet(MLVec& a, const std::vector<double>& v, int t) const {         // MultiLevelVector::setMultiLevelValues2
        for (size_t p = 0; p < ml.slotsG[t].size(); ++p) a[t][p] = v[ml.slotsG[t][p]];
    }
    double getVal(const std::vector<double>& v, const Key& K) const {          // VectorSparseG::getValue(I) (vector.h:331-336)
        int64_t k;
        if (g.occupied(K, k)) return v[k];
        return 0.0;
    }

    // prolongation1D_inplace (MatrixVectorHomogen.h:444-474)
    void prolongation1D(std::vector<double>& v, int t, int dd) const {
        for (size_t ti = 0; ti < g.depths.size(); ++ti) {
            if (g.depths[ti][dd] != t) continue;
            for (int64_t s : g.slotsOfDepth[ti]) {
                const Key& I = g.pts[s];
                double value = v[s];
                if (!isBoundary1(I.v[dd])) {
                    Key R = I; R.v[dd] = poleRight(I.v[dd]);
                    Key L = I; L.v[dd] = poleLeft(I.v[dd]);
                    value = value + 0.5 * (getVal(v, R) + getVal(v, L));
                }
                v[s] = value;
            }
        }
    }

    // restriction1D_inverted_inplace (MatrixVectorHomogen.h:485-530)
    void restriction1D(std::vector<double>& v, int t, int dd) const {
        for (size_t ti = 0; ti < g.depths.size(); ++ti) {
            const Depth& Tl = g.depths[ti];
            if (Tl[dd] != t + 1) continue;
            bool allPos = true;
            for (int i = 0; i < g.d; ++i) if (Tl[i] <= 0) allPos = false;
            if (!allPos) continue;
            for (int64_t s : g.slotsOfDepth[ti]) {
                const Key& I = g.pts[s];
                double value = 0.5 * v[s];
                if (t >= 0) {
                    Key R = I; R.v[dd] = poleRight(I.v[dd]);
                    if (R.v[dd] != 1) { int64_t k; if (g.occupied(R, k)) v[k] += value; }
                    Key L = I; L.v[dd] = poleLeft(I.v[dd]);
                    if (L.v[dd] != 1) { int64_t k; if (g.occupied(L, k)) v[k] += value; }
                }
            }
        }
    }

    // applyStencilGlobal_symmetry + StencilTemplate::applyStencilOnCell
    // (MatrixVectorHomogen.h:1157-1189, stencils/Stencil.h:187-246)
    void applyStencil(const std::vector<double>& u, std::vector<double>& z, int ti) const {
        const Depth& T = g.depths[ti];
        int d = g.d;
        double h[MAXD];
        for (int i = 0; i < d; ++i) h[i] = 1.0 / double(1 << T[i]);
        unsigned nc = 1u << d;
        std::vector<int64_t> ks(nc);
        std::vector<uint8_t> occ(nc);
        for (size_t c = 0; c < cellLo[ti].size(); ++c) {
            const auto& lo = cellLo[ti][c];
            for (unsigned p = 0; p < nc; ++p) {
                Key K{};
                bool interior = true;
                for (int i = 0; i < d; ++i) {
                    K.v[i] = levelNodes[ti][i][lo[i] + ((p >> i) & 1u)];
                    if (isBoundary1(K.v[i])) interior = false;
                }
                int64_t k = -1;
                occ[p] = interior && g.occupied(K, k);
                ks[p] = k;
            }
            for (unsigned p = 0; p < nc; ++p) {
                if (!occ[p]) continue;
                for (unsigned q = p; q < nc; ++q) {
                    if (!occ[q]) continue;
                    double val = mats.empty() ? op.element(d, h, cellCtr[ti][c].data(), p, q)
                                              : mats[ti][c * nEnt() + entryOf(p, q, nc)];
                    if (p == q) z[ks[p]] += val * u[ks[p]];
                    else { z[ks[q]] += val * u[ks[p]]; z[ks[p]] += val * u[ks[q]]; }
                }
            }
        }
    }

    // ConvertToPrewavelet2 (MatrixVectorHomogen.h:540-562)
    void convertToPrewavelet(const std::vector<double>& z, std::vector<double>& Ax, int ti) const {
        std::vector<std::pair<Key, double>> st;
        for (int64_t s : g.slotsOfDepth[ti]) {
            if (!g.active[s]) continue;
            stencil5(s, st);
            double coeff = 0.0;
            for (auto& [J, c] : st) coeff = coeff + getVal(z, J) * c;
            Ax[s] = coeff;
        }
    }

    // caseFunction (MatrixVectorHomogen.h:965-1090)
    void caseFunction(const MLVec& nodal, const bool* R, std::vector<double>& AxNeu) const {
        int d = g.d;
        size_t nd = g.depths.size();
        MLVec P = nodal;                         // P += nodal assigns (multilevelvector.h:112-116)
        MLVec gM(nd);
        for (size_t t = 0; t < nd; ++t) gM[t].assign(ml.slotsG[t].size(), 0.0);
        std::vector<int> sorted = sortDepths(R);
        if (sorted.empty()) return;
        std::vector<double> unew(N()), u(N()), z(N()), gv(N());
        for (int dd = 0; dd < d; ++dd) {
            if (R[dd]) continue;
            for (int ti : sorted) {
                Depth Tf = g.depths[ti];
                int t = Tf[dd] + 1;
                Tf[dd] = t;
                auto it = g.depthIdx.find(Tf);
                if (it == g.depthIdx.end()) continue;
                std::fill(unew.begin(), unew.end(), 0.0);
                setFromML(unew, P, ti);
                prolongation1D(unew, t, dd);
                mlAdd(P, unew, it->second);
            }
        }
        for (int ti : sorted) {
            std::fill(u.begin(), u.end(), 0.0);
            setFromML(u, P, ti);
            std::fill(z.begin(), z.end(), 0.0);
            applyStencil(u, z, ti);
            addFromML(z, gM, ti);
            Depth Tc = g.depths[ti];
            for (int dd = 0; dd < d; ++dd) {
                if (!R[dd]) continue;
                Tc[dd] = g.depths[ti][dd] - 1;
                auto it = g.depthIdx.find(Tc);
                if (it != g.depthIdx.end()) {
                    restriction1D(z, Tc[dd], dd);
                    gv = z;
                    addFromML(z, gM, it->second);
                    mlSet(gM, gv, it->second);
                }
            }
            auto it = g.depthIdx.find(Tc);
            if (it != g.depthIdx.end()) convertToPrewavelet(z, AxNeu, it->second);
        }
    }

    // multiplication (MatrixVectorHomogen.h:660-759), serial non-MPI branch; cases by
    // CasesIterator (tests/old_versions/MatrixVectorMultiplicationPrewavelets.h:23-50)
    void matvec(const std::vector<double>& x, std::vector<double>& Ax) const {
        size_t nd = g.depths.size();
        MLVec nodal(nd);
        calcNodal(x, nodal);
        Ax.assign(N(), 0.0);
        std::vector<double> AxNeu(N());
        bool R[MAXD];
        for (int c = 0; c < (1 << g.d); ++c) {
            for (int i = 0; i < g.d; ++i) R[i] = (c >> i) & 1;
            std::fill(AxNeu.begin(), AxNeu.end(), 0.0);
            caseFunction(nodal, R, AxNeu);
            for (int64_t s = 0; s < N(); ++s) Ax[s] = Ax[s] + AxNeu[s];
        }
    }

    // (A_T v)_J for a point J of G_T: sum over in-grid cells of depth T+1 containing J
    // (Preconditioning::applyStencilLocal, applications/cg_method.h:691-709, and
    // StencilTemplate::returnValue / cellInGrid, stencils/Stencil.h:667-719)
    double applyLocal(int ti, const Key& J, const std::unordered_map<Key, double, KeyHash>& v) const {
        const Depth& T = g.depths[ti];
        int d = g.d;
        double h[MAXD];
        int pos[MAXD];
        for (int i = 0; i < d; ++i) {
            h[i] = 1.0 / double(1 << T[i]);
            // position of J at level T
            int32_t idx = J.v[i];
            int s = depth1(idx);
            int32_t n = idx & ~(1 << s);
            pos[i] = (2 * n + 1) << (T[i] - s);
        }
        double val = 0.0;
        unsigned nc = 1u << d;
        for (unsigned cs = 0; cs < nc; ++cs) {   // cells around J: lower corner pos - 1 + bit
            int lo[MAXD];
            for (int i = 0; i < d; ++i) lo[i] = pos[i] - 1 + ((cs >> i) & 1u);
            // J's corner id in this cell: bit i = 1 if J is the upper node in dim i
            unsigned pj = 0;
            for (int i = 0; i < d; ++i) if (((cs >> i) & 1u) == 0) pj |= (1u << i);
            // in grid?
            bool ok = true;
            Key Ks[1 << 8];
            bool interior[1 << 8];
            for (unsigned q = 0; q < nc && ok; ++q) {
                Key K{};
                bool in = true;
                for (int i = 0; i < d; ++i) {
                    K.v[i] = levelNodes[ti][i][lo[i] + ((q >> i) & 1u)];
                    if (isBoundary1(K.v[i])) in = false;
                }
                Ks[q] = K; interior[q] = in;
                int64_t k;
                if (in && !g.occupied(K, k)) ok = false;
            }
            if (!ok) continue;
            double ctr[MAXD];
            for (int i = 0; i < d; ++i) ctr[i] = (double(lo[i]) + 0.5) * h[i];
            int64_t cell = -1;
            if (!mats.empty()) {
                Key L{};
                for (int i = 0; i < d; ++i) L.v[i] = lo[i];
                cell = cellOf[ti].at(L);
            }
            for (unsigned q = 0; q < nc; ++q) {
                if (!interior[q]) continue;
                auto it = v.find(Ks[q]);
                if (it == v.end()) continue;
                const double e = cell < 0 ? op.element(d, h, ctr, pj, q) : mats[ti][cell * nEnt() + entryOf(pj, q, nc)];
                val += e * it->second;
            }
        }
        return val;
    }

    // Preconditioning::precond3 (applications/cg_method.h:322-451): z_p = psi_p^T A_T psi_p
    void precond(std::vector<double>& z) const {
        z.assign(N(), 0.0);
        std::vector<std::pair<Key, double>> st;
        for (size_t ti = 0; ti < g.depths.size(); ++ti) {
            for (int64_t s : g.slotsOfDepth[ti]) {
                if (!g.active[s]) continue;
                stencil5(s, st);
                std::unordered_map<Key, double, KeyHash> v;
                for (auto& [J, c] : st) { int64_t k; if (g.occupied(J, k)) v[J] = c; }
                double coeff = 0.0;
                for (auto& [J, c] : st) {
                    int64_t k;
                    if (!g.occupied(J, k)) continue;
                    coeff += c * applyLocal((int)ti, J, v);
                }
                z[s] = coeff;
            }
        }
    }

    // CG::solveHomogen (applications/cg_method.h:759-886); absolute stop on r^T D^-1 r.
    int cg(double eps, std::vector<double>& x, const std::vector<double>& f, int maxIt,
           std::vector<double>* deltaHist) const {
        int64_t n = N();
        std::vector<double> Dg;
        precond(Dg);
        std::vector<double> r(n), h(n, 0.0), dvec(n), gvec(n);
        matvec(x, r);
        for (int64_t i = 0; i < n; ++i) r[i] = f[i] - r[i];
        for (int64_t i = 0; i < n; ++i) if (g.active[i]) h[i] = r[i];
        for (int64_t i = 0; i < n; ++i) if (g.active[i]) h[i] = h[i] / Dg[i];
        for (int64_t i = 0; i < n; ++i) dvec[i] = -1.0 * h[i];
        double delta = 0.0;
        for (int64_t i = 0; i < n; ++i) delta = delta + r[i] * h[i];
        if (deltaHist) deltaHist->push_back(delta);
        int k = 0;
        for (int it = 1; it <= maxIt && delta > eps; ++it) {
            k = it;
            matvec(dvec, gvec);
            double dg = 0.0;
            for (int64_t i = 0; i < n; ++i) dg = dg + dvec[i] * gvec[i];
            double tau = delta / dg;
            for (int64_t i = 0; i < n; ++i) r[i] = r[i] + tau * gvec[i];
            for (int64_t i = 0; i < n; ++i) x[i] = x[i] - tau * dvec[i];
            for (int64_t i = 0; i < n; ++i) h[i] = 1.0 * r[i];
            for (int64_t i = 0; i < n; ++i) if (g.active[i]) h[i] = h[i] / Dg[i];
            double dp = 0.0;
            for (int64_t i = 0; i < n; ++i) dp = dp + r[i] * h[i];
            if (deltaHist) deltaHist->push_back(dp);
            if (dp < eps * eps) break;
            double beta = dp / delta;
            delta = dp;
            for (int64_t i = 0; i < n; ++i) dvec[i] = -1.0 * h[i] + beta * dvec[i];
        }
        return k;
    }
};


// ---------------------------------------------------------------------------
// Basis transformations (BasisTransformations/BasisTransformations.cpp) on grids
// whose multilevel boxes are complete: for every depth T of the grid, every point of
// the level-T full grid G_T is stored (regular grids).  Vectors are in slot order.
// ---------------------------------------------------------------------------
struct Trafo {
    const Grid& g;
    const MultiLevel& ml;
    // depths in the order ListOfDepthOrderedSubgrids visits them from gotoEnd() through
    // previous() (sgrid/ListOfDepthOrderedGrids.cc:77-137, 219-244): one list per |T|_1 in
    // first-appearance order over the slots, walked from the largest |T|_1 down and each
    // list from its back, so depths of equal |T|_1 come in reverse first-appearance order
    std::vector<int> order;
    Trafo(const Grid& g_, const MultiLevel& ml_) : g(g_), ml(ml_) {
        order.resize(g.depths.size());
        for (size_t t = 0; t < order.size(); ++t) order[t] = (int)(order.size() - 1 - t);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sum(g.depths[a]) > sum(g.depths[b]); });
    }
    int sum(const Depth& T) const { int s = 0; for (int i = 0; i < g.d; ++i) s += T[i]; return s; }
    bool leq(const Depth& A, const Depth& B) const {
        for (int i = 0; i < g.d; ++i) if (A[i] > B[i]) return false;
        return true;
    }
    // ShiftOperator Links/Rechts (sgoperation/shift.*): the value at the neighbour of the
    // point's own depth in dimension i, i.e. at its pole; 0 on the boundary.
    double pole(const std::vector<double>& u, int64_t s, int i, bool left) const {
        Key k = g.pts[s];
        int32_t p = left ? poleLeft(k.v[i]) : poleRight(k.v[i]);
        if (isBoundary1(p)) return 0.0;
        k.v[i] = p;
        int64_t q;
        return g.occupied(k, q) ? u[q] : 0.0;
    }
    // CalcHierarchicalBasis (BasisTransformations.cpp:13-37, 390-409): per dimension,
    // finest subgrids first, u -= (Links(u) + Rechts(u)) / 2; `upTo` restricts it to the
    // subgrids of depth <= upTo.
    void hierarchize(std::vector<double>& u, const Depth* upTo = nullptr) const {
        for (int i = 0; i < g.d; ++i)
            for (int t : order) {
                if (upTo && !leq(g.depths[t], *upTo)) continue;
                for (int64_t s : g.slotsOfDepth[t])
                    if (g.active[s]) u[s] = u[s] - 0.5 * (pole(u, s, i, true) + pole(u, s, i, false));
            }
    }
    // CalcNodalBasis (BasisTransformations.cpp:174-198): coarsest subgrids first.
    void dehierarchize(std::vector<double>& u) const {
        for (int i = 0; i < g.d; ++i)
            for (auto it = order.rbegin(); it != order.rend(); ++it)
                for (int64_t s : g.slotsOfDepth[*it])
                    if (g.active[s]) u[s] = u[s] + 0.5 * (pole(u, s, i, true) + pole(u, s, i, false));
    }
    // PrewaveletMatrixHomogen::create (sgoperation/matrix.cpp:74-139) for a full line of
    // level t (n = 2^t - 1 positions, 1-based): column j is the basis function at position
    // j written in level-t nodal values -- the prewavelet (0.1, -0.6, 1, -0.6, 0.1; 0.9 next
    // to the boundary) at the positions of depth t, the level-(t-1) hat (0.5, 1, 0.5) at
    // the coarser ones.  Row-major n x n.
    static std::vector<double> lineMatrix(int t) {
        const int n = (1 << t) - 1;
        std::vector<double> M((size_t)n * n, 0.0);
        auto set = [&](int r, int c, double v) { if (r >= 1 && r <= n) M[(size_t)(r - 1) * n + (c - 1)] = v; };
        for (int j = 1; j <= n; ++j) {
            if (j & 1) {
                if (t == 1) { set(j, j, 1.0); continue; }
                if (j == 1) { set(j, j, 0.9); set(j + 1, j, -0.6); set(j + 2, j, 0.1); }
                else if (j == n) { set(j, j, 0.9); set(j - 1, j, -0.6); set(j - 2, j, 0.1); }
                else { set(j, j, 1.0); set(j - 1, j, -0.6); set(j - 2, j, 0.1); set(j + 1, j, -0.6); set(j + 2, j, 0.1); }
            } else {
                set(j, j, 1.0); set(j - 1, j, 0.5); set(j + 1, j, 0.5);
            }
        }
        return M;
    }
    // Solve M x = b by Gaussian elimination with partial pivoting (the reference uses a
    // banded LR without pivoting, PrewaveletMatrixHomogen::solve, matrix.cpp:141-155;
    // the solution is the same up to rounding).
    static void solveDense(std::vector<double> M, std::vector<double>& b) {
        const int n = (int)b.size();
        for (int c = 0; c < n; ++c) {
            int p = c;
            for (int r = c + 1; r < n; ++r) if (std::fabs(M[(size_t)r * n + c]) > std::fabs(M[(size_t)p * n + c])) p = r;
            if (p != c) {
                for (int k = 0; k < n; ++k) std::swap(M[(size_t)c * n + k], M[(size_t)p * n + k]);
                std::swap(b[c], b[p]);
            }
            for (int r = c + 1; r < n; ++r) {
                double f = M[(size_t)r * n + c] / M[(size_t)c * n + c];
                if (f == 0.0) continue;
                for (int k = c; k < n; ++k) M[(size_t)r * n + k] -= f * M[(size_t)c * n + k];
                b[r] -= f * b[c];
            }
        }
        for (int r = n - 1; r >= 0; --r) {
            double s = b[r];
            for (int k = r + 1; k < n; ++k) s -= M[(size_t)r * n + k] * b[k];
            b[r] = s / M[(size_t)r * n + r];
        }
    }
    // Points by their level-T positions (1 .. 2^T_i - 1 interior; 0 and 2^T_i boundary).
    static int32_t idxAtPos(int32_t k, int t) {
        if (k <= 0) return 0;
        if (k >= (1 << t)) return 1;
        const int tz = __builtin_ctz((unsigned)k);
        return (1 << (t - tz)) + (k >> (tz + 1));
    }
    static int32_t posOfIdx(int32_t v, int t) {
        if (v == 0) return 0;
        if (v == 1) return 1 << t;
        const int sd = depth1(v);
        return (2 * (v - (1 << sd)) + 1) << (t - sd);
    }
    // slot at level-T positions k[] (-1: not stored, or a boundary position)
    int64_t slotAtPos(const Depth& T, const int32_t* k) const {
        Key key{};
        for (int i = 0; i < g.d; ++i) {
            if (k[i] <= 0 || k[i] >= (1 << T[i])) return -1;
            key.v[i] = idxAtPos(k[i], T[i]);
        }
        int64_t q = -1;
        return g.occupied(key, q) ? q : -1;
    }
    // stored and worked on (workonindex: active; hanging nodes are not worked on,
    // sgrid/sparseGrid.h:189-193)
    bool work(const Depth& T, const int32_t* k, int64_t* out = nullptr) const {
        const int64_t q = slotAtPos(T, k);
        if (out) *out = q;
        return q >= 0 && g.active[q];
    }
    // ZusammenhangsKomponente::recursiveMarkArbeite (sgrid/komponente.cpp:43-73): depth-T
    // points connected through the level-T 5^d stencil (IndexDimension::nextFive,
    // indices/index.cc:124-176: a step along dimension i stops when its first
    // level-T step reaches the boundary); the box [mn, mx] spans every stored, worked-on
    // stencil point met on the way.
    void markComponent(const Depth& T, int64_t s, std::vector<char>& marked, int32_t* mn, int32_t* mx) const {
        const int d = g.d;
        marked[s] = 1;
        int32_t k0[MAXD];
        for (int i = 0; i < d; ++i) k0[i] = posOfIdx(g.pts[s].v[i], T[i]);
        int o[MAXD];
        for (int i = 0; i < d; ++i) o[i] = -2;
        while (true) {
            bool ok = true;
            int32_t k[MAXD];
            for (int i = 0; i < d && ok; ++i) {
                if (o[i] < 0 && k0[i] - 1 < 1) ok = false;
                if (o[i] > 0 && k0[i] + 1 > (1 << T[i]) - 1) ok = false;
                k[i] = k0[i] + o[i];
            }
            if (ok) {
                int64_t q;
                if (work(T, k, &q) && !marked[q]) {
                    for (int i = 0; i < d; ++i) { mn[i] = std::min(mn[i], k[i]); mx[i] = std::max(mx[i], k[i]); }
                    if (g.depthOf[q] == T) markComponent(T, q, marked, mn, mx);
                }
            }
            int i = 0;
            while (i < d && ++o[i] > 2) o[i++] = -2;
            if (i == d) break;
        }
        for (int i = 0; i < d; ++i) { mn[i] = std::min(mn[i], k0[i]); mx[i] = std::max(mx[i], k0[i]); }
    }
    // PrewaveletMatrixHomogen::create (sgoperation/matrix.cpp:74-139) for the line along dim
    // that starts at level-T position k[]: column j is the basis function at line position
    // j in level-T nodal values -- the prewavelet (0.1, -0.6, 1, -0.6, 0.1; 0.9 next to the
    // boundary) at positions of depth T_dim, the coarser hat (0.5, 1, 0.5) elsewhere; an
    // entry exists only where that neighbour is stored and worked on.  Entries that would
    // fall outside the line are dropped.  Row-major size x size.
    std::vector<double> lineMatrixAt(const Depth& T, const int32_t* kStart, int dim, int size) const {
        std::vector<double> M((size_t)size * size, 0.0);
        auto set = [&](int r, int c, double v) { if (r >= 0 && r < size) M[(size_t)r * size + c] = v; };
        const int t = T[dim];
        int32_t k[MAXD];
        for (int i = 0; i < g.d; ++i) k[i] = kStart[i];
        auto workAt = [&](int off) {
            int32_t kk[MAXD];
            for (int i = 0; i < g.d; ++i) kk[i] = k[i];
            kk[dim] += off;
            return work(T, kk);
        };
        for (int c = 0; c < size && work(T, k); ++c, ++k[dim]) {
            const int dpos = t - __builtin_ctz((unsigned)k[dim]);       // depth of the position along dim
            if (dpos == t) {
                if (k[dim] - 1 == 0) {
                    set(c, c, 0.9);
                    if (t == 1) { set(c, c, 1.0); break; }
                    if (workAt(1)) set(c + 1, c, -0.6);
                    if (workAt(2)) set(c + 2, c, 0.1);
                } else if (k[dim] + 1 == (1 << t)) {
                    set(c, c, 0.9);
                    if (workAt(-1)) set(c - 1, c, -0.6);
                    if (workAt(-2)) set(c - 2, c, 0.1);
                } else {
                    set(c, c, 1.0);
                    if (workAt(-1)) set(c - 1, c, -0.6);
                    if (workAt(-2)) set(c - 2, c, 0.1);
                    if (workAt(1)) set(c + 1, c, -0.6);
                    if (workAt(2)) set(c + 2, c, 0.1);
                }
            } else {
                set(c, c, 1.0);
                if (workAt(-1)) set(c - 1, c, 0.5);
                if (workAt(1)) set(c + 1, c, 0.5);
            }
        }
        return M;
    }
    // SolvePrewavelet / GetPrewaveletCoefficients (BasisTransformations.cpp:100-172,
    // 468-484): for each connected component of depth-T points (found in the order of
    // their first point among the depth-T points, ZusammenhangsKomponente::findAnicht,
    // komponente.cpp:9-41), along every dimension, every line of the component's box is
    // replaced by the solution of M x = line (M from the line through the box's lower
    // corner, of the length of its stored run); the values at worked-on points are
    // written back, and those at depth-T points are their prewavelet coefficients.
    void solvePrewavelet(const Depth& T, std::vector<double>& u, std::vector<double>& prew) const {
        const int d = g.d;
        const int t = g.depthIdx.at(T);
        std::vector<char> marked(g.pts.size(), 0);
        for (int64_t s0 : g.slotsOfDepth[t]) {
            if (marked[s0] || !g.active[s0]) continue;
            int32_t mn[MAXD], mx[MAXD];
            for (int i = 0; i < d; ++i) mn[i] = mx[i] = posOfIdx(g.pts[s0].v[i], T[i]);
            markComponent(T, s0, marked, mn, mx);
            for (int dim = 0; dim < d; ++dim) {
                int size = 1;
                {
                    int32_t k[MAXD];
                    for (int i = 0; i < d; ++i) k[i] = mn[i];
                    for (;;) {
                        ++k[dim];
                        if (k[dim] >= (1 << T[dim]) || !work(T, k)) break;
                        ++size;
                    }
                }
                const std::vector<double> M = lineMatrixAt(T, mn, dim, size);
                // every line start: the box's positions with dim fixed at mn[dim]
                int32_t k[MAXD];
                for (int i = 0; i < d; ++i) k[i] = mn[i];
                std::vector<double> line(size);
                std::vector<int64_t> sl(size);
                while (true) {
                    int32_t kk[MAXD];
                    for (int i = 0; i < d; ++i) kk[i] = k[i];
                    for (int q = 0; q < size; ++q, ++kk[dim]) {
                        sl[q] = slotAtPos(T, kk);
                        line[q] = sl[q] >= 0 ? u[sl[q]] : 0.0;
                    }
                    solveDense(M, line);
                    for (int q = 0; q < size; ++q) {
                        const int64_t c = sl[q];
                        if (c < 0 || !g.active[c]) continue;
                        if (g.depthOf[c] == T) prew[c] = line[q];
                        u[c] = line[q];
                    }
                    int i = 0;
                    for (; i < d; ++i) {
                        if (i == dim) continue;
                        if (++k[i] <= mx[i]) break;
                        k[i] = mn[i];
                    }
                    if (i == d) break;
                }
            }
        }
    }
    // CalcHierbyPrew (BasisTransformations.cpp:321-353): the depth-T prewavelets of prew
    // as nodal values (the 5^d stencil of nextFive) hierarchized on the subgrids <= T.
    void hierByPrew(const Depth& T, int t, const std::vector<double>& prew, std::vector<double>& hier) const {
        hier.assign(g.pts.size(), 0.0);
        const int d = g.d;
        for (int64_t s : g.slotsOfDepth[t]) {
            Entry1 e[MAXD][5];
            int ne[MAXD];
            for (int i = 0; i < d; ++i) ne[i] = prewavelet1D(g.pts[s].v[i], T[i], e[i]);
            int c[MAXD] = {};
            while (true) {
                Key k{};
                double w = prew[s];
                for (int i = 0; i < d; ++i) { k.v[i] = e[i][c[i]].idx; w *= e[i][c[i]].c; }
                int64_t q;
                if (g.occupied(k, q) && g.active[q]) hier[q] += w;
                int i = 0;
                while (i < d && ++c[i] == ne[i]) c[i++] = 0;
                if (i == d) break;
            }
        }
        hierarchize(hier, &T);
    }
    // calcPrewByNodal (BasisTransformations.cpp:486-545)
    void prewByNodal(const std::vector<double>& u0, std::vector<double>& prew) const {
        std::vector<double> hier = u0, u = u0, hnew;
        hierarchize(hier);
        prew.assign(g.pts.size(), 0.0);
        for (int t : order) {
            const Depth& T = g.depths[t];
            solvePrewavelet(T, u, prew);
            hierByPrew(T, t, prew, hnew);
            for (size_t s = 0; s < hier.size(); ++s)
                if (leq(g.depthOf[s], T) && g.active[s]) hier[s] = hier[s] - hnew[s];
            u = hier;
            dehierarchize(u);
        }
    }
    // calcNodalByPrew (BasisTransformations.cpp:670-707)
    void nodalByPrew(const std::vector<double>& prew, std::vector<double>& u) const {
        std::vector<double> hier(g.pts.size(), 0.0), hnew;
        for (int t : order) {
            const Depth& T = g.depths[t];
            hierByPrew(T, t, prew, hnew);
            for (size_t s = 0; s < hier.size(); ++s)
                if (leq(g.depthOf[s], T) && g.active[s]) hier[s] = hier[s] + hnew[s];
        }
        u = hier;
        dehierarchize(u);
    }
};

}  // namespace orc

// ---------------------------------------------------------------------------
// C ABI for ctypes (tests only)
// ---------------------------------------------------------------------------
extern "C" {

int64_t orc_regular_grid(int d, int level, int32_t* idx_out, uint8_t* active_out, int64_t cap) {
    using namespace orc;
    if (d < 1 || d > MAXD) return -1;
    Grid g; g.d = d;
    std::unordered_map<Key, char, KeyHash> seen;
    Key c{};
    for (int i = 0; i < d; ++i) c.v[i] = 2;
    addRecursiveSons(g, c, level, seen);
    int64_t n = (int64_t)g.pts.size();
    if (idx_out && cap >= n) {
        for (int64_t s = 0; s < n; ++s) {
            for (int i = 0; i < d; ++i) idx_out[s * d + i] = g.pts[s].v[i];
            if (active_out) active_out[s] = g.active[s];
        }
    }
    return n;
}

void* orc_create(int d, int64_t n, const int32_t* idx, const uint8_t* active) {
    using namespace orc;
    if (d < 1 || d > MAXD) return nullptr;
    auto* o = new Oracle();
    o->g.d = d;
    o->g.pts.resize(n);
    o->g.active.resize(n);
    for (int64_t s = 0; s < n; ++s) {
        for (int i = 0; i < d; ++i) o->g.pts[s].v[i] = idx[s * d + i];
        o->g.active[s] = active ? active[s] : 1;
    }
    o->init();
    return o;
}

void orc_destroy(void* h) { delete (orc::Oracle*)h; }

int orc_set_coef(void* h, int kind, double a0, double c0, const double* xi, int nxi) {
    auto* o = (orc::Oracle*)h;
    o->op.coef.kind = kind;
    o->op.coef.a0 = a0;
    o->op.coef.c0 = c0;
    o->op.coef.xi.assign(xi, xi + (xi ? nxi : 0));
    return 0;
}

int orc_num_depths(void* h) { return (int)((orc::Oracle*)h)->g.depths.size(); }

// Cells of depth T+1 of every depth (depth-list order, cellsOfDepth order within a depth):
// level-T boxes prod_i [pos_i 2^-T_i, (pos_i + 1) 2^-T_i].  Returns the count.
int64_t orc_cell_boxes(void* h, int32_t* level, int32_t* pos) {
    auto* o = (orc::Oracle*)h;
    int64_t k = 0;
    for (size_t t = 0; t < o->g.depths.size(); ++t)
        for (const auto& lo : o->cellLo[t]) {
            for (int i = 0; i < o->g.d; ++i) {
                if (level) level[k * o->g.d + i] = o->g.depths[t][i];
                if (pos) pos[k * o->g.d + i] = lo[i];
            }
            ++k;
        }
    return k;
}

// Stored element matrices for the cells of orc_cell_boxes (ncells x 2^(d-1) (2^d + 1));
// NULL returns to the coefficient's element matrices.
int orc_set_cell_matrices(void* h, const double* entries) {
    auto* o = (orc::Oracle*)h;
    o->mats.clear();
    if (!entries) return 0;
    const int64_t ne = o->nEnt();
    int64_t k = 0;
    for (size_t t = 0; t < o->g.depths.size(); ++t) {
        const int64_t nc = (int64_t)o->cellLo[t].size();
        o->mats.emplace_back(entries + k * ne, entries + (k + nc) * ne);
        k += nc;
    }
    return 0;
}

void orc_depths(void* h, int32_t* out) {
    auto* o = (orc::Oracle*)h;
    for (size_t t = 0; t < o->g.depths.size(); ++t)
        for (int i = 0; i < o->g.d; ++i) out[t * o->g.d + i] = o->g.depths[t][i];
}

int64_t orc_multilevel_size(void* h) {
    auto* o = (orc::Oracle*)h;
    int64_t s = 0;
    for (auto& v : o->ml.slotsG) s += (int64_t)v.size();
    return s;
}

int orc_sorted_depths(void* h, int caseno, int32_t* out) {
    auto* o = (orc::Oracle*)h;
    bool R[orc::MAXD];
    for (int i = 0; i < o->g.d; ++i) R[i] = (caseno >> i) & 1;
    auto s = o->sortDepths(R);
    if (out) for (size_t k = 0; k < s.size(); ++k) out[k] = s[k];
    return (int)s.size();
}

int orc_matvec(void* h, const double* x, double* y) {
    auto* o = (orc::Oracle*)h;
    std::vector<double> xv(x, x + o->N()), yv;
    o->matvec(xv, yv);
    std::memcpy(y, yv.data(), sizeof(double) * o->N());
    return 0;
}

int orc_precond(void* h, double* z) {
    auto* o = (orc::Oracle*)h;
    std::vector<double> zv;
    o->precond(zv);
    std::memcpy(z, zv.data(), sizeof(double) * o->N());
    return 0;
}

int orc_cg(void* h, double eps, const double* f, double* x, int maxit, double* delta_hist, int hist_cap) {
    auto* o = (orc::Oracle*)h;
    std::vector<double> fv(f, f + o->N()), xv(x, x + o->N()), hist;
    int it = o->cg(eps, xv, fv, maxit, &hist);
    std::memcpy(x, xv.data(), sizeof(double) * o->N());
    if (delta_hist) for (int k = 0; k < hist_cap && k < (int)hist.size(); ++k) delta_hist[k] = hist[k];
    return it;
}

// Basis transformations (complete grids only; returns -1 otherwise).
// what: 0 hierarchize, 1 dehierarchize, 2 prewavelets from nodal values,
// 3 nodal values from prewavelets.
int orc_transform(void* h, int what, const double* in, double* out) {
    auto* o = (orc::Oracle*)h;
    orc::Trafo tr(o->g, o->ml);
    std::vector<double> a(in, in + o->N()), b;
    switch (what) {
        case 0: tr.hierarchize(a); b = a; break;
        case 1: tr.dehierarchize(a); b = a; break;
        case 2: tr.prewByNodal(a, b); break;
        case 3: tr.nodalByPrew(a, b); break;
        default: return -2;
    }
    std::memcpy(out, b.data(), sizeof(double) * o->N());
    return 0;
}

}  // extern "C"
