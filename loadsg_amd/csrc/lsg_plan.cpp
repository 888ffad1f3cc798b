// lsg_plan.cpp -- grid flattening and matvec schedule (host side).
//
// The schedule restates the data flow of MatrixVectorHomogen::caseFunction
// (library/source/MatrixVectorMultiplication/MatrixVectorHomogen.h:965-1090):
//   * cases = restriction subsets R of the dimensions (CasesIterator,
//     tests/old_versions/MatrixVectorMultiplicationPrewavelets.h:23-50);
//   * per case, the depths S_c of DepthList::sortDepths
//     (iterator/depthIterator.cpp:7-41,88-93) in the reference order;
//   * prolongation: P[T+e_d] += I_d P[T] for every prolongation dim d, in
//     ascending d and sorted-depth order (MatrixVectorHomogen.h:997-1019);
//   * for every T of S_c: z = A_T P[T] + gM[T], then for each restriction dim
//     (ascending) restrict, g = z, z += gM[Tc], gM[Tc] = g, and finally
//     ConvertToPrewavelet2(z, Ax, Tc) (MatrixVectorHomogen.h:1023-1085).
// The gM exchanges are replayed symbolically in the reference order, so every
// read of gM is bound to the exact earlier write it sees.  That turns the
// reference's sequential loop into a dependency DAG whose items at equal depth in
// the DAG (wavefront levels) run concurrently on the GPU with identical data flow.
#include "lsg_plan.hpp"
#include "../../include/loadsg_amd.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <set>
#include <unordered_set>

namespace lsg {

namespace {

inline int depth1(int32_t i) { return i <= 1 ? 0 : 31 - __builtin_clz((uint32_t)i); }

struct Key2 {
    uint64_t a = 0, b = 0;
    bool operator==(const Key2& o) const { return a == o.a && b == o.b; }
};
struct Key2Hash {
    size_t operator()(const Key2& k) const { return (size_t)(k.a * 0x9E3779B97F4A7C15ull ^ (k.b + 0x632BE59BD9B4E019ull)); }
};
inline Key2 packKey(const int32_t* v, int d) {
    Key2 k;
    for (int i = 0; i < d; ++i) {
        uint64_t x = (uint64_t)(uint32_t)v[i] & 0xFFFFull;
        if (i < 4) k.a |= x << (16 * i);
        else k.b |= x << (16 * (i - 4));
    }
    return k;
}

void dfsSons(int d, std::array<int32_t, MAXD>& I, int level, std::unordered_set<Key2, Key2Hash>& seen,
             std::vector<int32_t>& out) {
    Key2 k = packKey(I.data(), d);
    if (seen.count(k)) return;     // every path to a point has the same length
    seen.insert(k);
    for (int i = 0; i < d; ++i) out.push_back(I[i]);
    if (level <= 0) return;
    for (int dd = 0; dd < d; ++dd) {
        int32_t save = I[dd];
        I[dd] = save << 1;
        dfsSons(d, I, level - 1, seen, out);
        I[dd] = (save << 1) + 1;
        dfsSons(d, I, level - 1, seen, out);
        I[dd] = save;
    }
}

// DepthList::recursiveDepth (iterator/depthIterator.cpp:7-41)
void recursiveDepth(const Grid& g, const bool* R, bool checkProl, int dd, Depth& T, std::vector<int>& out) {
    int D = g.d;
    if (checkProl && dd > 0) {
        if (!R[dd]) {
            for (int t = 1; t <= g.maxDepth[dd]; ++t) { T[dd] = t; recursiveDepth(g, R, checkProl, dd - 1, T, out); }
        } else recursiveDepth(g, R, checkProl, dd - 1, T, out);
    }
    if (checkProl && dd == 0) {
        if (!R[dd]) {
            for (int t = 1; t <= g.maxDepth[dd]; ++t) { T[dd] = t; recursiveDepth(g, R, false, D - 1, T, out); }
        } else recursiveDepth(g, R, false, D - 1, T, out);
    }
    if (!checkProl && dd >= 0) {
        if (R[dd]) {
            for (int t = g.maxDepth[dd]; t > 1; --t) { T[dd] = t; recursiveDepth(g, R, checkProl, dd - 1, T, out); }
        } else recursiveDepth(g, R, checkProl, dd - 1, T, out);
    }
    if (dd < 0) {
        auto it = g.depthIdx.find(T);
        if (it != g.depthIdx.end()) out.push_back(it->second);
    }
}

int64_t boxOf(const Depth& T, int d) {
    int64_t s = 1;
    for (int i = 0; i < d; ++i) s *= (int64_t(1) << T[i]) - 1;
    return s;
}
int64_t blockOf(const Depth& T, int d) {
    int64_t s = 1;
    for (int i = 0; i < d; ++i) s *= int64_t(1) << (T[i] - 1);
    return s;
}

// ---- sparse path ------------------------------------------------------------
// 1D index <-> position on the level-t grid (positions 1 .. 2^t - 1 are interior;
// indices/index.h:657-876: index 2^s + n sits at position (2n+1) 2^(t-s)).
inline int32_t posAt(int32_t v, int t) {
    const int s = depth1(v);
    return (2 * (v - (1 << s)) + 1) << (t - s);
}
inline int32_t idxAt(int32_t k, int t) {
    const int tz = __builtin_ctz((uint32_t)k);
    return (1 << (t - tz)) + (k >> (tz + 1));
}
// 1D prewavelet of the point at odd position k, level t: (position, coefficient)
// pairs (IndexDimension::nextFive, indices/index.cc:180-327): centre 1 (0.9 next to
// the boundary for t >= 2), neighbours -0.6, second neighbours 0.1, cut at the boundary.
int prewavelet1D(int32_t k, int t, int32_t pos[5], double c[5]) {
    const int32_t n = (1 << t);
    int m = 0;
    pos[m] = k;
    c[m++] = (t >= 2 && (k == 1 || k == n - 1)) ? 0.9 : 1.0;
    if (k - 1 >= 1) { pos[m] = k - 1; c[m++] = -0.6; pos[m] = k - 2; c[m++] = 0.1; }
    if (k + 1 <= n - 1) { pos[m] = k + 1; c[m++] = -0.6; pos[m] = k + 2; c[m++] = 0.1; }
    return m;
}

int buildSparse(Grid& g, const std::vector<int>& tOf, std::string& err) {
    const int d = g.d;
    const int nd = (int)g.depths.size();
    const int64_t n = g.n;
    SparseTables& S = g.sp;
    S = SparseTables();
    std::unordered_map<Key2, int64_t, Key2Hash> slotOf;
    slotOf.reserve((size_t)n * 2);
    for (int64_t s = 0; s < n; ++s) slotOf[packKey(&g.idx[s * d], d)] = s;
    auto find = [&](const int32_t* v) -> int64_t {
        auto it = slotOf.find(packKey(v, d));
        return it == slotOf.end() ? -1 : it->second;
    };
    // G_t in slot order and the local index of every slot in it
    std::vector<std::vector<int64_t>> G(nd);
    std::vector<std::unordered_map<int64_t, int32_t>> loc(nd);
    int64_t off = 0;
    for (int t = 0; t < nd; ++t) {
        const Depth& T = g.depths[t];
        for (int64_t s = 0; s < n; ++s) {
            const Depth& Ts = g.depths[tOf[s]];
            bool le = true;
            for (int i = 0; i < d; ++i) if (Ts[i] > T[i]) { le = false; break; }
            if (le) { loc[t][s] = (int32_t)G[t].size(); G[t].push_back(s); }
        }
        if (G[t].size() >= (size_t)INT32_MAX) { err = "multilevel set too large"; return LSG_ERR_CAPACITY; }
        g.boxSize[t] = (int64_t)G[t].size();
        g.boxOff[t] = off;
        off += g.boxSize[t];
    }
    g.nml = off;
    auto locOf = [&](int t, const int32_t* v) -> int32_t {
        const int64_t s = find(v);
        if (s < 0) return -1;
        auto it = loc[t].find(s);
        return it == loc[t].end() ? -1 : it->second;
    };

    // calcNodal rows: every active point of depth T scatters its prewavelet to the stored
    // points of its stencil (MatrixVectorHomogen.cpp:43-74), gathered here by target row
    {
        std::vector<std::vector<std::pair<int64_t, double>>> rows(g.nml);
        for (int64_t s = 0; s < n; ++s) {
            if (!g.active[s]) continue;
            const int t = tOf[s];
            const Depth& T = g.depths[t];
            int32_t p1[MAXD][5];
            double c1[MAXD][5];
            int m1[MAXD];
            for (int i = 0; i < d; ++i) m1[i] = prewavelet1D(posAt(g.idx[s * d + i], T[i]), T[i], p1[i], c1[i]);
            int cnt[MAXD] = {};
            while (true) {
                int32_t J[MAXD];
                double c = 1.0;
                for (int i = 0; i < d; ++i) { J[i] = idxAt(p1[i][cnt[i]], T[i]); c *= c1[i][cnt[i]]; }
                const int32_t l = locOf(t, J);
                if (l >= 0) rows[g.boxOff[t] + l].push_back({g.devOfSlot[s], c});
                int i = 0;
                for (; i < d; ++i) { if (++cnt[i] < m1[i]) break; cnt[i] = 0; }
                if (i == d) break;
            }
        }
        S.nodPtr.assign(g.nml + 1, 0);
        for (int64_t r = 0; r < g.nml; ++r) S.nodPtr[r + 1] = S.nodPtr[r] + (int64_t)rows[r].size();
        for (auto& row : rows)
            for (auto& e : row) { S.nodCol.push_back((int32_t)e.first); S.nodVal.push_back(e.second); }
    }
    // prolongation gathers (prolongation1D_inplace, MatrixVectorHomogen.h:444-474)
    S.prolOff.assign((size_t)nd * d, -1);
    for (int t = 0; t < nd; ++t)
        for (int dd = 0; dd < d; ++dd) {
            Depth Tc = g.depths[t];
            if (Tc[dd] < 2) continue;
            Tc[dd] -= 1;
            auto it = g.depthIdx.find(Tc);
            if (it == g.depthIdx.end()) continue;
            const int tc = it->second;
            S.prolOff[(size_t)t * d + dd] = (int64_t)S.prolAB.size() / 2;
            for (int64_t s : G[t]) {
                const int32_t* v = &g.idx[s * d];
                auto lc = loc[tc].find(s);
                if (lc != loc[tc].end()) { S.prolAB.push_back(lc->second); S.prolAB.push_back(-2); continue; }
                // depth_dd(s) = T_dd: the poles along dd (positions k -+ 1 at level T_dd)
                const int tdd = g.depths[t][dd];
                const int32_t k = posAt(v[dd], tdd);
                int32_t w[MAXD];
                for (int i = 0; i < d; ++i) w[i] = v[i];
                int32_t a = -1, b = -1;
                if (k - 1 >= 1) { w[dd] = idxAt(k - 1, tdd); a = locOf(tc, w); }
                if (k + 1 <= (1 << tdd) - 1) { w[dd] = idxAt(k + 1, tdd); b = locOf(tc, w); }
                S.prolAB.push_back(a);
                S.prolAB.push_back(b);
            }
        }
    // cells of depth T+1 in the grid: cells of the level-T full grid whose interior corners
    // are all stored (cells/CellStructure.h:23-61), and for every row the cells around it
    S.cellOff.assign(nd + 1, 0);
    S.rcPtr.assign(g.nml + 1, 0);
    {
        const unsigned nc = 1u << d;
        std::vector<std::vector<std::pair<int32_t, int32_t>>> rc(g.nml);
        for (int t = 0; t < nd; ++t) {
            const Depth& T = g.depths[t];
            S.cellOff[t] = (int64_t)S.cellLo.size() / d;
            std::unordered_map<Key2, int32_t, Key2Hash> seenCell;
            for (int64_t s : G[t]) {
                int32_t k[MAXD];
                for (int i = 0; i < d; ++i) k[i] = posAt(g.idx[s * d + i], T[i]);
                for (unsigned cs = 0; cs < nc; ++cs) {
                    int32_t lo[MAXD];
                    bool inside = true;
                    for (int i = 0; i < d; ++i) {
                        lo[i] = k[i] - 1 + (int32_t)((cs >> i) & 1u);
                        if (lo[i] < 0 || lo[i] >= (1 << T[i])) inside = false;
                    }
                    if (!inside) continue;
                    const Key2 ck = packKey(lo, d);
                    int32_t cid;
                    auto sc = seenCell.find(ck);
                    if (sc != seenCell.end()) {
                        cid = sc->second;
                    } else {
                        int32_t corner[1 << MAXD];
                        bool ok = true;
                        for (unsigned q = 0; q < nc && ok; ++q) {
                            int32_t J[MAXD];
                            bool interior = true;
                            for (int i = 0; i < d; ++i) {
                                const int32_t p = lo[i] + (int32_t)((q >> i) & 1u);
                                if (p == 0 || p == (1 << T[i])) { interior = false; break; }
                                J[i] = idxAt(p, T[i]);
                            }
                            if (!interior) { corner[q] = -1; continue; }
                            corner[q] = locOf(t, J);
                            if (corner[q] < 0) ok = false;
                        }
                        if (!ok) cid = -1;
                        else {
                            cid = (int32_t)((int64_t)S.cellLo.size() / d - S.cellOff[t]);
                            for (unsigned q = 0; q < nc; ++q) S.cellCorner.push_back(corner[q]);
                            for (int i = 0; i < d; ++i) S.cellLo.push_back(lo[i]);
                        }
                        seenCell[ck] = cid;
                    }
                    if (cid < 0) continue;
                    // corner of s in this cell: bit i set if s is the upper node in dim i
                    int32_t cj = 0;
                    for (int i = 0; i < d; ++i) if (((cs >> i) & 1u) == 0) cj |= 1 << i;
                    rc[g.boxOff[t] + loc[t][s]].push_back({cid, cj});
                }
            }
        }
        S.cellOff[nd] = (int64_t)S.cellLo.size() / d;
        for (int t = 0; t < nd; ++t)
            if ((S.cellOff[t + 1] - S.cellOff[t]) * (int64_t)nc >= (int64_t)INT32_MAX) {
                err = "too many cells in one depth";
                return LSG_ERR_CAPACITY;
            }
        // cells of each depth grouped by the parity of their position (2^d colours): two
        // cells of one colour share no corner
        S.cellColOff.assign((size_t)nd * (nc + 1), 0);
        S.cellCol.reserve((size_t)S.cellOff[nd]);
        for (int t = 0; t < nd; ++t) {
            const int64_t c0 = S.cellOff[t], n0 = S.cellOff[t + 1] - c0;
            std::vector<std::vector<int32_t>> byCol(nc);
            for (int64_t c = 0; c < n0; ++c) {
                unsigned col = 0;
                for (int i = 0; i < d; ++i) col |= (unsigned)(S.cellLo[(size_t)(c0 + c) * d + i] & 1) << i;
                byCol[col].push_back((int32_t)c);
            }
            for (unsigned col = 0; col < nc; ++col) {
                S.cellColOff[(size_t)t * (nc + 1) + col] = (int64_t)S.cellCol.size();
                S.cellCol.insert(S.cellCol.end(), byCol[col].begin(), byCol[col].end());
            }
            S.cellColOff[(size_t)t * (nc + 1) + nc] = (int64_t)S.cellCol.size();
        }
        for (int64_t r = 0; r < g.nml; ++r) S.rcPtr[r + 1] = S.rcPtr[r] + (int64_t)rc[r].size();
        for (auto& row : rc)
            for (auto& e : row) {
                S.rcCell.push_back(e.first);
                S.rcCorner.push_back(e.second);
                S.rcY.push_back((int32_t)((int64_t)e.first * nc + e.second));
            }
    }
    // restriction gathers (restriction1D_inverted_inplace, MatrixVectorHomogen.h:485-530):
    // z'(i) = z(i) + 0.5 (z(left child) + z(right child)) for i in G_{T - e_dd}
    S.restOff.assign((size_t)nd * d, -1);
    for (int t = 0; t < nd; ++t)
        for (int dd = 0; dd < d; ++dd) {
            Depth Tc = g.depths[t];
            if (Tc[dd] < 2) continue;
            Tc[dd] -= 1;
            auto it = g.depthIdx.find(Tc);
            if (it == g.depthIdx.end()) continue;
            const int tc = it->second;
            S.restOff[(size_t)t * d + dd] = (int64_t)S.restSLR.size() / 3;
            const int tdd = g.depths[t][dd];
            for (int64_t s : G[tc]) {
                const int32_t* v = &g.idx[s * d];
                int32_t w[MAXD];
                for (int i = 0; i < d; ++i) w[i] = v[i];
                const int32_t k = posAt(v[dd], tdd);     // even position at level T_dd
                S.restSLR.push_back(loc[t].at(s));
                int32_t a = -1, b = -1;
                if (k - 1 >= 1) { w[dd] = idxAt(k - 1, tdd); a = locOf(t, w); }
                if (k + 1 <= (1 << tdd) - 1) { w[dd] = idxAt(k + 1, tdd); b = locOf(t, w); }
                S.restSLR.push_back(a);
                S.restSLR.push_back(b);
            }
        }
    // prewavelet conversion rows (ConvertToPrewavelet2, MatrixVectorHomogen.h:540-562)
    S.convPtr.assign(n + 1, 0);
    S.pos.assign((size_t)n * d, 0);
    S.tOfDev.assign(n, 0);
    S.depthT.assign((size_t)nd * d, 0);
    for (int t = 0; t < nd; ++t)
        for (int i = 0; i < d; ++i) S.depthT[(size_t)t * d + i] = g.depths[t][i];
    for (int64_t s = 0; s < n; ++s) S.tOfDev[g.devOfSlot[s]] = tOf[s];
    {
        std::vector<std::vector<std::pair<int32_t, double>>> rows(n);
        for (int64_t s = 0; s < n; ++s) {
            const int t = tOf[s];
            const Depth& T = g.depths[t];
            const int64_t dev = g.devOfSlot[s];
            for (int i = 0; i < d; ++i) S.pos[dev * d + i] = posAt(g.idx[s * d + i], T[i]);
            if (!g.active[s]) continue;
            int32_t p1[MAXD][5];
            double c1[MAXD][5];
            int m1[MAXD];
            for (int i = 0; i < d; ++i) m1[i] = prewavelet1D(posAt(g.idx[s * d + i], T[i]), T[i], p1[i], c1[i]);
            int cnt[MAXD] = {};
            while (true) {
                int32_t J[MAXD];
                double c = 1.0;
                for (int i = 0; i < d; ++i) { J[i] = idxAt(p1[i][cnt[i]], T[i]); c *= c1[i][cnt[i]]; }
                const int32_t l = locOf(t, J);
                if (l >= 0) rows[dev].push_back({l, c});
                int i = 0;
                for (; i < d; ++i) { if (++cnt[i] < m1[i]) break; cnt[i] = 0; }
                if (i == d) break;
            }
        }
        for (int64_t r = 0; r < n; ++r) S.convPtr[r + 1] = S.convPtr[r] + (int64_t)rows[r].size();
        for (auto& row : rows)
            for (auto& e : row) { S.convCol.push_back(e.first); S.convVal.push_back(e.second); }
    }
    return 0;
}

// allProl: materialise every prolongation dimension of a case in the case buffer, so that
// every apply item is a single depth (sparse grids; dense grids for small batches, where
// splitting the families along the last prolongation dimension shortens the critical path
// of each wavefront level).  Otherwise the last prolongation dimension is fused into the
// apply stage (families).
int buildPlan(Grid& g, Plan& P, bool allProl, std::string& err) {
    const int d = g.d;
    const int nd = (int)g.depths.size();
    P = Plan();
    // multilevel set / block sizes by depth (boxes and blocks of the dense layout, point
    // lists of the sparse one)
    auto boxSz = [&](const Depth& T) -> int64_t { return g.boxSize[g.depthIdx.at(T)]; };
    auto toDepth = [&](const int32_t* T) { Depth D{}; for (int i = 0; i < d; ++i) D[i] = T[i]; return D; };

    // --- calcNodal items
    for (int t = 0; t < nd; ++t) {
        NodalItem it{};
        for (int i = 0; i < d; ++i) it.T[i] = g.depths[t][i];
        it.in_off = g.blkOff[t];
        it.out_off = g.boxOff[t];
        P.nodal.push_back(it);
        P.maxBox = std::max(P.maxBox, g.boxSize[t]);
        P.maxBlock = std::max(P.maxBlock, g.blkSize[t]);
    }

    struct Read { int item = -1; int step = -1; };
    struct Tmp {
        int c, t;
        int k;
        int r[MAXD];
        Read reads[MAXD + 1];
        int outDepth;
        int level = 0;
        int64_t caseOff = -1;
        int fam = -1;
        bool wNeeded[MAXD + 1] = {};
        int64_t wOff[MAXD + 1];
    };
    std::vector<Tmp> items;
    struct Fam { int c; int fd; std::vector<int> ts; };
    std::vector<Fam> famList;
    P.prol.assign(d, {});
    std::vector<std::vector<int64_t>> partOf(nd, std::vector<int64_t>(size_t(1) << d, -1));
    double stream = double(g.n) + double(g.nml) + double(g.n);   // calcNodal in/out + final write
    double streamApply = 0.0;   // of which the stages the apply kernel performs

    // chains along dd over a depth list: maximal runs of depths T, T + e_dd, ... (the
    // reference links T -> T + e_dd only when the finer depth exists,
    // MatrixVectorHomogen.h:1006; where T + e_dd is missing it adds nothing)
    auto chainsOf = [&](int dd, const std::vector<int>& list, std::map<Depth, std::vector<int>>& fam) -> int {
        std::map<Depth, std::vector<int>> lines;
        for (int t : list) {
            Depth K = g.depths[t];
            K[dd] = 0;
            lines[K].push_back(t);
        }
        fam.clear();
        for (auto& kv : lines) {
            auto ts = kv.second;
            std::sort(ts.begin(), ts.end(), [&](int a, int b) { return g.depths[a][dd] < g.depths[b][dd]; });
            size_t q0 = 0;
            for (size_t q = 1; q <= ts.size(); ++q)
                if (q == ts.size() || g.depths[ts[q]][dd] != g.depths[ts[q - 1]][dd] + 1) {
                    Depth K = kv.first;
                    K[dd] = g.depths[ts[q0]][dd];      // keyed by the first depth of the chain
                    fam[K].assign(ts.begin() + q0, ts.begin() + q);
                    if (q - q0 > (size_t)MAXLEV) { err = "family too long"; return LSG_ERR_UNSUPPORTED; }
                    q0 = q;
                }
        }
        return 0;
    };

    // Prolongation stages shared between cases.  A case prolongates along its
    // prolongation dims p1 < p2 < ... in ascending order, over the lines of its depth
    // list S_c = {T : T_i >= 2 for the restriction dims i} (DepthList::sortDepths).  A
    // line along a prolongation dim is either a whole line of the grid or absent from
    // S_c, so the result after the first j prolongation dims depends only on the prefix
    // Q = {p1..pj} and is the same arithmetic for every case with that prefix: each
    // prefix is computed once, over the union of the depth lists of its cases.  Dense
    // grids leave the last prolongation dim to the apply stage; sparse grids
    // materialise every prolongation dim.
    const int ncase = 1 << d;
    std::vector<std::vector<int>> sortedOf(ncase);
    std::map<unsigned, std::set<int>> prefixDepths;       // prefix mask -> union of depth lists
    for (int c = 0; c < ncase; ++c) {
        bool R[MAXD] = {};
        for (int i = 0; i < d; ++i) R[i] = (c >> i) & 1;
        Depth T0{};
        recursiveDepth(g, R, true, d - 1, T0, sortedOf[c]);
        if (sortedOf[c].empty()) continue;
        std::vector<int> Pd;
        for (int i = 0; i < d; ++i) if (!R[i]) Pd.push_back(i);
        const int nStages = allProl ? (int)Pd.size() : (int)Pd.size() - 1;
        unsigned Q = 0;
        for (int j = 0; j < nStages; ++j) {
            Q |= 1u << Pd[j];
            prefixDepths[Q].insert(sortedOf[c].begin(), sortedOf[c].end());
        }
    }
    // Where P_Q[t] lives: the first box of a chain has nothing coarser along dd, so
    // P_Q equals P_parent there and is not copied: it aliases the parent's values (the
    // nodal box when the parent is empty).  Chains of one box need no work at all.
    // (std::map visits a prefix after its parent, whose mask is smaller.)
    struct Src { bool nodal; int64_t off; };
    std::map<unsigned, std::unordered_map<int, Src>> prefixSrc;
    for (auto& [Q, ts] : prefixDepths) {
        const int dd = 31 - __builtin_clz(Q);
        const unsigned parent = Q & ~(1u << dd);
        auto srcOf = [&](unsigned Qp, int t) -> Src {
            return Qp == 0 ? Src{true, g.boxOff[t]} : prefixSrc.at(Qp).at(t);
        };
        std::vector<int> list(ts.begin(), ts.end());
        std::map<Depth, std::vector<int>> fam;
        if (int rc = chainsOf(dd, list, fam)) return rc;
        auto& mine = prefixSrc[Q];
        for (auto& [K, tq] : fam) {
            mine[tq[0]] = srcOf(parent, tq[0]);
            for (size_t q = 1; q < tq.size(); ++q) {
                mine[tq[q]] = Src{false, P.caseSize};
                P.caseSize += g.boxSize[tq[q]];
            }
            if (tq.size() < 2) continue;
            ProlItem pi{};
            for (int i = 0; i < d; ++i) pi.T[i] = g.depths[tq[0]][i];
            pi.dd = dd;
            pi.nt = (int)tq.size();
            pi.from_nodal = 0;
            for (size_t q = 0; q < tq.size(); ++q) {
                const Src in = srcOf(parent, tq[q]);
                if (in.nodal) pi.from_nodal |= 1 << q;
                pi.in_off[q] = in.off;
                pi.out_off[q] = q == 0 ? -1 : mine[tq[q]].off;
                pi.tidx[q] = tq[q];
                P.maxFamily = std::max(P.maxFamily, g.boxSize[tq[q]]);
            }
            P.prol[dd].push_back(pi);
        }
    }

    for (int c = 0; c < ncase; ++c) {
        bool R[MAXD] = {};
        for (int i = 0; i < d; ++i) R[i] = (c >> i) & 1;
        const std::vector<int>& sorted = sortedOf[c];
        if (sorted.empty()) continue;
        int lastP = -1, nP = 0;
        for (int i = 0; i < d; ++i) if (!R[i]) { lastP = i; ++nP; }
        auto families = [&](int dd, std::map<Depth, std::vector<int>>& fam) -> int {
            return chainsOf(dd, sorted, fam);
        };
        // the apply stage reads the boxes prolongated along all prolongation dims but
        // the last one (sparse grids: all of them)
        // (-1: the nodal box itself, whose offset is boxOff[t])
        std::unordered_map<int, int64_t> caseOff;
        if (nP >= (allProl ? 1 : 2)) {
            unsigned Q = 0;
            for (int i = 0; i < d; ++i) if (!R[i] && (allProl || i != lastP)) Q |= 1u << i;
            for (int t : sorted) {
                const Src& sr = prefixSrc.at(Q).at(t);
                caseOff[t] = sr.nodal ? -1 : sr.off;
            }
        }
        // streaming-model prolongation traffic (every prolongation dim, reference order)
        for (int dd = 0; dd < d; ++dd) {
            if (R[dd]) continue;
            std::map<Depth, std::vector<int>> fam;
            if (int rc = families(dd, fam)) return rc;
            for (auto& [K, ts] : fam)
                for (size_t q = 1; q < ts.size(); ++q) {
                    const double v = double(g.boxSize[ts[q - 1]]) + 2.0 * double(g.boxSize[ts[q]]);
                    stream += v;
                    if (dd == lastP && !allProl) streamApply += v;   // fused into the apply stage
                }
        }
        // apply families along the last prolongation dim (single items if none)
        std::vector<std::vector<int>> applyFam;
        if (lastP >= 0 && !allProl) {
            std::map<Depth, std::vector<int>> fam;
            if (int rc = families(lastP, fam)) return rc;
            for (auto& kv : fam) applyFam.push_back(kv.second);
        } else {
            for (int t : sorted) applyFam.push_back({t});
        }
        std::unordered_map<int, int> famOf;
        for (size_t f = 0; f < applyFam.size(); ++f)
            for (int t : applyFam[f]) famOf[t] = (int)f;
        const int famBase = (int)famList.size();
        for (auto& ts : applyFam) famList.push_back({c, allProl ? -1 : lastP, ts});

        // symbolic replay of the gM exchanges in the reference order
        std::unordered_map<int, Read> state;   // depth -> latest write
        for (int t : sorted) {
            Tmp X{};
            X.c = c; X.t = t; X.k = 0;
            for (int i = 0; i < d; ++i) if (R[i]) X.r[X.k++] = i;
            for (int j = 0; j <= MAXD; ++j) { X.reads[j] = Read(); X.wOff[j] = -1; }
            X.caseOff = nP >= (allProl ? 1 : 2) ? caseOff[t] : -1;
            X.fam = famBase + famOf[t];
            int me = (int)items.size();
            auto st = state.find(t);
            if (st != state.end()) X.reads[0] = st->second;
            Depth Tc = g.depths[t];
            stream += 3.0 * double(g.boxSize[t]);
            streamApply += 3.0 * double(g.boxSize[t]);
            for (int j = 1; j <= X.k; ++j) {
                int rr = X.r[j - 1];
                Depth prev = Tc;
                Tc[rr] -= 1;
                auto it = g.depthIdx.find(Tc);
                if (it == g.depthIdx.end()) {
                    err = "coarse depth of a restriction missing from the depth list";
                    return LSG_ERR_UNSUPPORTED;
                }
                int tc = it->second;
                stream += double(boxSz(prev)) + 3.0 * double(g.boxSize[tc]);
                streamApply += double(boxSz(prev)) + 3.0 * double(g.boxSize[tc]);
                auto s2 = state.find(tc);
                if (s2 != state.end()) X.reads[j] = s2->second;
                state[tc] = Read{me, j};
            }
            auto it = g.depthIdx.find(Tc);
            if (it == g.depthIdx.end()) { err = "target depth missing"; return LSG_ERR_UNSUPPORTED; }
            X.outDepth = it->second;
            stream += double(g.boxSize[X.outDepth]) + 2.0 * double(g.blkSize[X.outDepth]);
            streamApply += double(g.boxSize[X.outDepth]) + 2.0 * double(g.blkSize[X.outDepth]);
            if (partOf[X.outDepth][c] >= 0) { err = "two depths of one case convert to the same depth"; return LSG_ERR_UNSUPPORTED; }
            partOf[X.outDepth][c] = P.partSize;
            P.partSize += g.blkSize[X.outDepth];
            // dependencies and levels
            for (int j = 0; j <= X.k; ++j) {
                if (X.reads[j].item >= 0) {
                    Tmp& Y = items[X.reads[j].item];
                    Y.wNeeded[X.reads[j].step] = true;
                    X.level = std::max(X.level, Y.level + 1);
                }
            }
            items.push_back(X);
        }
    }

    // W buffer allocation
    for (auto& X : items) {
        Depth Tc = g.depths[X.t];
        for (int j = 1; j <= X.k; ++j) {
            Tc[X.r[j - 1]] -= 1;
            if (X.wNeeded[j]) { X.wOff[j] = P.wSize; P.wSize += boxSz(Tc); }
        }
    }
    // wavefront level of a family: one more than the deepest family any of its items
    // reads a gM value from (fixpoint over the family DAG; a family never reads from
    // itself because every read refers to a depth finer in a restriction dim)
    std::vector<int> famLevel(famList.size(), 0);
    for (bool changed = true; changed;) {
        changed = false;
        for (auto& X : items)
            for (int j = 0; j <= X.k; ++j) {
                if (X.reads[j].item < 0) continue;
                int fy = items[X.reads[j].item].fam;
                if (fy == X.fam) { err = "family reads its own gM values"; return LSG_ERR_UNSUPPORTED; }
                if (famLevel[X.fam] < famLevel[fy] + 1) {
                    famLevel[X.fam] = famLevel[fy] + 1;
                    changed = true;
                    if (famLevel[X.fam] > (int)famList.size()) { err = "cyclic family dependencies"; return LSG_ERR_UNSUPPORTED; }
                }
            }
    }
    int nlev = 0;
    for (int v : famLevel) nlev = std::max(nlev, v + 1);
    P.apply.assign(nlev, {});
    std::map<std::pair<int, int>, int> itemOf;   // (case, depth) -> item
    for (size_t q = 0; q < items.size(); ++q) itemOf[{items[q].c, items[q].t}] = (int)q;
    auto emit = [&](const Tmp& X) {
        ApplyItem a{};
        for (int i = 0; i < d; ++i) a.T[i] = g.depths[X.t][i];
        a.k = X.k;
        for (int j = 0; j < X.k; ++j) a.r[j] = X.r[j];
        a.in_nodal = X.caseOff < 0 ? 1 : 0;
        a.tidx = X.t;
        a.in_off = X.caseOff < 0 ? g.boxOff[X.t] : X.caseOff;
        {
            Depth Tc = g.depths[X.t];
            a.ctidx[0] = X.t;
            for (int j = 1; j <= X.k; ++j) {
                Tc[X.r[j - 1]] -= 1;
                a.ctidx[j] = g.depthIdx.at(Tc);
            }
        }
        // box shapes after 0..k restrictions (the dense apply kernels read them as they are)
        a.shp = (int32_t)P.shapes.size();
        {
            BoxShape sh{};
            sh.d = d;
            for (int i = 0; i < d; ++i) sh.n[i] = (1 << a.T[i]) - 1;
            for (int j = 0; j <= X.k; ++j) {
                if (j > 0) sh.n[X.r[j - 1]] = (sh.n[X.r[j - 1]] - 1) >> 1;
                int st = 1;
                for (int i = 0; i < d; ++i) {
                    sh.S[i] = st;
                    sh.rS[i] = 1.0f / float(st);
                    sh.rn[i] = 1.0f / float(sh.n[i]);
                    st *= sh.n[i];
                }
                sh.total = st;
                P.shapes.push_back(sh);
            }
        }
        a.z_off = -1;
        if (g.sparse) {      // stencil result and restriction chain, two rows of |G_T|
            a.z_off = P.zSize;    // (+ the element products of the cells of depth T+1)
            P.zSize += 2 * g.boxSize[X.t];
            if (g.d <= SP_CELL_MAXD) P.zSize += (g.sp.cellOff[X.t + 1] - g.sp.cellOff[X.t]) << g.d;
        }
        for (int j = 0; j <= MAXD; ++j) { a.rd_off[j] = -1; a.w_off[j] = -1; }
        for (int j = 0; j <= X.k; ++j)
            if (X.reads[j].item >= 0) a.rd_off[j] = items[X.reads[j].item].wOff[X.reads[j].step];
        for (int j = 1; j <= X.k; ++j) a.w_off[j] = X.wOff[j];
        a.out_off = partOf[X.outDepth][X.c];
        P.applyItems.push_back(a);
    };
    for (size_t f = 0; f < famList.size(); ++f) {
        const Fam& F = famList[f];
        FamilyItem fi{};
        fi.fd = F.fd;
        fi.nt = (int)F.ts.size();
        fi.first = (int)P.applyItems.size();
        for (int t : F.ts) emit(items[itemOf[{F.c, t}]]);
        P.apply[famLevel[f]].push_back(fi);
    }
    P.nApply = (int64_t)items.size();

    // sum items
    for (int t = 0; t < nd; ++t) {
        SumItem s{};
        for (int i = 0; i < d; ++i) s.T[i] = g.depths[t][i];
        s.blk_off = g.blkOff[t];
        s.count = g.blkSize[t];
        s.list_start = (int)P.sumList.size();
        for (int c = 0; c < ncase; ++c)
            if (partOf[t][c] >= 0) P.sumList.push_back(partOf[t][c]);
        s.ncase = (int)P.sumList.size() - s.list_start;
        P.sum.push_back(s);
    }
    P.streamValues = stream;
    P.streamApplyValues = streamApply;

    // values this implementation moves per matvec (each kernel reads its inputs and
    // writes its outputs once)
    double mv = 0.0;
    for (int t = 0; t < nd; ++t) mv += double(g.blkSize[t]) + double(g.boxSize[t]);
    for (auto& v : P.prol)
        for (auto& it : v)
            for (int q = 0; q < it.nt; ++q) mv += 2.0 * double(g.boxSize[it.tidx[q]]);
    const double beforeApply = mv;
    for (auto& a : P.applyItems) {
        Depth T = toDepth(a.T);
        mv += double(boxSz(T));
        if (a.rd_off[0] >= 0) mv += double(boxSz(T));
        for (int j = 1; j <= a.k; ++j) {
            T[a.r[j - 1]] -= 1;
            if (a.w_off[j] >= 0) mv += double(boxSz(T));
            if (a.rd_off[j] >= 0) mv += double(boxSz(T));
        }
        mv += double(g.blkSize[g.depthIdx.at(T)]);
    }
    P.applyValues = mv - beforeApply;
    for (auto& s : P.sum) mv += double(s.count) * double(s.ncase + 1);
    P.movedValues = mv;
    return 0;
}

}  // namespace

void regularIndexTable(int d, int level, std::vector<int32_t>& idx) {
    idx.clear();
    std::unordered_set<Key2, Key2Hash> seen;
    std::array<int32_t, MAXD> I{};
    for (int i = 0; i < d; ++i) I[i] = 2;
    dfsSons(d, I, level, seen, idx);
}

int prewaveletStencil1D(int32_t k, int t, int32_t pos[5], double c[5]) { return prewavelet1D(k, t, pos, c); }

int buildGrid(int d, int64_t n, const int32_t* idx, const uint8_t* active, Grid& g, std::string& err,
              bool forceSparse) {
    if (d < 1 || d > MAXD || n < 1 || !idx) { err = "bad arguments"; return LSG_ERR_ARG; }
    g = Grid();
    g.d = d;
    g.n = n;
    g.idx.assign(idx, idx + n * d);
    g.active.assign(n, 1);
    if (active) g.active.assign(active, active + n);
    std::unordered_set<Key2, Key2Hash> seen;
    std::vector<int> tOf(n);
    for (int64_t s = 0; s < n; ++s) {
        Depth T{};
        for (int i = 0; i < d; ++i) {
            int32_t v = idx[s * d + i];
            if (v < 2) { err = "boundary points are not part of a zero-Dirichlet grid"; return LSG_ERR_ARG; }
            int t = depth1(v);
            if (t > MAXLEV) { err = "1D depth above 15"; return LSG_ERR_UNSUPPORTED; }
            if (v - (1 << t) >= (1 << (t - 1))) { err = "invalid 1D index (outside the unit interval)"; return LSG_ERR_ARG; }
            T[i] = t;
        }
        if (!seen.insert(packKey(&idx[s * d], d)).second) { err = "duplicate point"; return LSG_ERR_ARG; }
        auto it = g.depthIdx.find(T);
        if (it == g.depthIdx.end()) {
            int id = (int)g.depths.size();
            g.depthIdx[T] = id;
            g.depths.push_back(T);
            tOf[s] = id;
        } else tOf[s] = it->second;
    }
    const int nd = (int)g.depths.size();
    g.maxDepth = Depth{};
    for (auto& T : g.depths)
        for (int i = 0; i < d; ++i) {
            g.maxDepth[i] = std::max(g.maxDepth[i], T[i]);
            g.maxLevel = std::max(g.maxLevel, T[i]);
        }
    // block layout; grids with hanging nodes or partly filled depth blocks take the
    // sparse path (multilevel point lists instead of dense boxes)
    std::vector<int64_t> cnt(nd, 0);
    for (int64_t s = 0; s < n; ++s) cnt[tOf[s]]++;
    g.blkOff.resize(nd); g.blkSize.resize(nd); g.boxOff.resize(nd); g.boxSize.resize(nd);
    g.sparse = false;
    for (int64_t s = 0; s < n; ++s) if (!g.active[s]) g.sparse = true;
    for (int t = 0; t < nd; ++t) if (cnt[t] != blockOf(g.depths[t], d)) g.sparse = true;
    // a grid with complete depth blocks keeps the dense device order even when it is
    // forced onto the sparse path (cell-wise coefficients), so vectors of both paths match
    const bool denseOrder = !g.sparse;
    int64_t off = 0, boff = 0;
    for (int t = 0; t < nd; ++t) {
        g.blkSize[t] = cnt[t];
        g.blkOff[t] = off; off += g.blkSize[t];
        g.boxSize[t] = boxOf(g.depths[t], d);
        g.boxOff[t] = boff; boff += g.boxSize[t];
        // downward closed (every restriction target exists)
        for (int i = 0; i < d; ++i) {
            if (g.depths[t][i] >= 2) {
                Depth Q = g.depths[t];
                Q[i] -= 1;
                if (!g.depthIdx.count(Q)) { err = "depth list not downward closed"; return LSG_ERR_UNSUPPORTED; }
            }
        }
    }
    g.nml = boff;
    g.devOfSlot.resize(n);
    g.slotOfDev.resize(n);
    if (!denseOrder) {
        // device order: depth blocks, slot order inside a block
        std::vector<int64_t> fill(nd, 0);
        for (int64_t s = 0; s < n; ++s) {
            const int64_t dev = g.blkOff[tOf[s]] + fill[tOf[s]]++;
            g.devOfSlot[s] = dev;
            g.slotOfDev[dev] = s;
        }
        if (int rc = buildSparse(g, tOf, err)) return rc;
        return buildPlan(g, g.plan, true, err);
    }
    for (int64_t s = 0; s < n; ++s) {
        const Depth& T = g.depths[tOf[s]];
        int64_t loc = 0, mul = 1;
        for (int i = 0; i < d; ++i) {
            int32_t v = idx[s * d + i];
            int64_t j = v - (1 << T[i]);
            loc += j * mul;
            mul *= int64_t(1) << (T[i] - 1);
        }
        int64_t dev = g.blkOff[tOf[s]] + loc;
        g.devOfSlot[s] = dev;
        g.slotOfDev[dev] = s;
    }
    if (forceSparse) {
        g.sparse = true;
        if (int rc = buildSparse(g, tOf, err)) return rc;
    }
    return buildPlan(g, g.plan, g.sparse, err);
}

// 1D Galerkin entries of the level-t prewavelets j and l (odd positions 2j + 1 and 2l + 1 of
// the level-t full grid, zero Dirichlet ends): K = Psi^T K_t Psi, M = Psi^T M_t Psi with the
// nodal K_t = 2^t tridiag(-1, 2, -1) and M_t = 2^-t / 6 tridiag(1, 4, 1).  A prewavelet spans
// five nodal positions, so entries with |j - l| > 2 vanish.
static void prewavelet1DPair(int t, int j, int l, double& k, double& m) {
    k = m = 0.0;
    if (std::abs(j - l) > 2) return;
    int32_t pa[5], pb[5];
    double ca[5], cb[5];
    const int na = prewavelet1D(2 * j + 1, t, pa, ca), nb = prewavelet1D(2 * l + 1, t, pb, cb);
    const double h = std::ldexp(1.0, -t);
    for (int e = 0; e < na; ++e)
        for (int f = 0; f < nb; ++f) {
            const int dq = std::abs(pa[e] - pb[f]);
            const double c = ca[e] * cb[f];
            if (dq == 0) { k += 2.0 * c / h; m += 4.0 * c * h / 6.0; }
            else if (dq == 1) { k -= c / h; m += c * h / 6.0; }
        }
}

// The restrictions K[S, S], M[S, S] of the level-t matrices to the positions S (row-major).
static void prewavelet1DMatrices(int t, const std::vector<int32_t>& S, std::vector<double>& K,
                                 std::vector<double>& M) {
    const int m = (int)S.size();
    K.assign((size_t)m * m, 0.0);
    M.assign((size_t)m * m, 0.0);
    for (int a = 0; a < m; ++a)
        for (int b = 0; b < m; ++b)
            prewavelet1DPair(t, S[a], S[b], K[(size_t)a * m + b], M[(size_t)a * m + b]);
}

// Generalised symmetric eigenproblem K v = lam M v (M positive definite, m x m row-major):
// M = L L^T, C = L^-1 K L^-T, C = W diag(lam) W^T by cyclic Jacobi rotations, V = L^-T W.
static int generalisedEigen(int m, const std::vector<double>& K, const std::vector<double>& M, std::vector<double>& V,
                            std::vector<double>& lam) {
    std::vector<double> L((size_t)m * m, 0.0);
    for (int j = 0; j < m; ++j) {
        double s = M[(size_t)j * m + j];
        for (int q = 0; q < j; ++q) s -= L[(size_t)j * m + q] * L[(size_t)j * m + q];
        if (!(s > 0.0)) return -1;
        L[(size_t)j * m + j] = std::sqrt(s);
        for (int i = j + 1; i < m; ++i) {
            double v = M[(size_t)i * m + j];
            for (int q = 0; q < j; ++q) v -= L[(size_t)i * m + q] * L[(size_t)j * m + q];
            L[(size_t)i * m + j] = v / L[(size_t)j * m + j];
        }
    }
    // Y = L^-1 K (forward substitution per column), C = Y L^-T = (L^-1 Y^T)^T
    std::vector<double> Y((size_t)m * m), C((size_t)m * m);
    for (int c = 0; c < m; ++c)
        for (int i = 0; i < m; ++i) {
            double v = K[(size_t)i * m + c];
            for (int q = 0; q < i; ++q) v -= L[(size_t)i * m + q] * Y[(size_t)q * m + c];
            Y[(size_t)i * m + c] = v / L[(size_t)i * m + i];
        }
    for (int r = 0; r < m; ++r)            // row r of C = L^-1 applied to row r of Y, as a column
        for (int i = 0; i < m; ++i) {
            double v = Y[(size_t)r * m + i];
            for (int q = 0; q < i; ++q) v -= L[(size_t)i * m + q] * C[(size_t)r * m + q];
            C[(size_t)r * m + i] = v / L[(size_t)i * m + i];
        }
    for (int i = 0; i < m; ++i)
        for (int j = i + 1; j < m; ++j) {
            const double a = 0.5 * (C[(size_t)i * m + j] + C[(size_t)j * m + i]);
            C[(size_t)i * m + j] = C[(size_t)j * m + i] = a;
        }
    std::vector<double> W((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) W[(size_t)i * m + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) {
                const double c2 = C[(size_t)i * m + j] * C[(size_t)i * m + j];
                tot += c2;
                if (i != j) off += c2;
            }
        if (off <= 1e-30 * tot) break;
        for (int p = 0; p < m - 1; ++p)
            for (int q = p + 1; q < m; ++q) {
                const double apq = C[(size_t)p * m + q];
                if (std::fabs(apq) < 1e-300) continue;
                const double app = C[(size_t)p * m + p], aqq = C[(size_t)q * m + q];
                const double th = 0.5 * (aqq - app) / apq;
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < m; ++k) {       // columns p, q of C
                    const double akp = C[(size_t)k * m + p], akq = C[(size_t)k * m + q];
                    C[(size_t)k * m + p] = c * akp - s * akq;
                    C[(size_t)k * m + q] = s * akp + c * akq;
                }
                for (int k = 0; k < m; ++k) {       // rows p, q of C
                    const double apk = C[(size_t)p * m + k], aqk = C[(size_t)q * m + k];
                    C[(size_t)p * m + k] = c * apk - s * aqk;
                    C[(size_t)q * m + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < m; ++k) {       // W <- W J
                    const double wkp = W[(size_t)k * m + p], wkq = W[(size_t)k * m + q];
                    W[(size_t)k * m + p] = c * wkp - s * wkq;
                    W[(size_t)k * m + q] = s * wkp + c * wkq;
                }
            }
    }
    lam.resize(m);
    for (int i = 0; i < m; ++i) lam[i] = C[(size_t)i * m + i];
    V.assign((size_t)m * m, 0.0);                 // V = L^-T W (back substitution per column)
    for (int c = 0; c < m; ++c)
        for (int i = m - 1; i >= 0; --i) {
            double v = W[(size_t)i * m + c];
            for (int q = i + 1; q < m; ++q) v -= L[(size_t)q * m + i] * V[(size_t)q * m + c];
            V[(size_t)i * m + c] = v / L[(size_t)i * m + i];
        }
    return 0;
}

int buildBlockPre(const Grid& g, int maxDense, BlockPre& out, std::string& err) {
    const int d = g.d;
    const int nd = (int)g.depths.size();
    out = BlockPre();
    out.d0.assign((size_t)g.n, 1.0);
    // 1D factors of the fast diagonalisation, shared by the blocks with the same (t, S)
    std::map<std::pair<int, std::vector<int32_t>>, std::pair<int32_t, int32_t>> factors;
    std::vector<int32_t> loc;       // local 1D positions of the block's points, nb x d
    std::vector<int32_t> act;       // block offsets of its active points
    for (int t = 0; t < nd; ++t) {
        const Depth& T = g.depths[t];
        const int64_t nb = g.blkSize[t];
        PreBlock pb{};
        pb.off = g.blkOff[t];
        pb.n = (int32_t)nb;
        pb.mode = PRE_DIAG;
        pb.inv = -1;
        loc.assign((size_t)nb * d, 0);
        act.clear();
        for (int64_t p = 0; p < nb; ++p) {
            const int64_t s = g.slotOfDev[(size_t)(pb.off + p)];
            for (int i = 0; i < d; ++i) loc[(size_t)p * d + i] = g.idx[(size_t)s * d + i] - (1 << T[i]);
            if (g.active[(size_t)s]) act.push_back((int32_t)p);
        }
        // diagonal of B_T at the active points
        for (int32_t p : act) {
            double kd[MAXD], md[MAXD], sum = 0.0;
            for (int j = 0; j < d; ++j)
                prewavelet1DPair(T[j], loc[(size_t)p * d + j], loc[(size_t)p * d + j], kd[j], md[j]);
            for (int i = 0; i < d; ++i) {
                double prod = kd[i];
                for (int j = 0; j < d; ++j)
                    if (j != i) prod *= md[j];
                sum += prod;
            }
            out.d0[(size_t)(pb.off + p)] = sum;
        }
        // the 1D position sets of the active points and whether they form their product
        std::vector<std::vector<int32_t>> S(d);
        int64_t prod = act.empty() ? 0 : 1;
        for (int i = 0; i < d; ++i) {
            for (int32_t p : act) S[i].push_back(loc[(size_t)p * d + i]);
            std::sort(S[i].begin(), S[i].end());
            S[i].erase(std::unique(S[i].begin(), S[i].end()), S[i].end());
            prod *= (int64_t)S[i].size();
        }
        bool fd = nb > maxDense && !act.empty() && prod == (int64_t)act.size() && prod <= FD_MAXN;
        for (int i = 0; i < d && fd; ++i) fd = (int)S[i].size() <= FD_MAXM;
        if (fd) {
            for (int i = 0; i < d; ++i) {
                auto key = std::make_pair((int)T[i], S[i]);
                auto it = factors.find(key);
                if (it == factors.end()) {
                    std::vector<double> K, M, V, lam;
                    prewavelet1DMatrices(T[i], S[i], K, M);
                    if (generalisedEigen((int)S[i].size(), K, M, V, lam)) {
                        err = "1D prewavelet mass matrix not positive definite";
                        return LSG_ERR_ARG;
                    }
                    it = factors.emplace(key, std::make_pair((int32_t)out.fdV.size(), (int32_t)out.fdLam.size())).first;
                    out.fdV.insert(out.fdV.end(), V.begin(), V.end());
                    out.fdLam.insert(out.fdLam.end(), lam.begin(), lam.end());
                }
                pb.m[i] = (int32_t)S[i].size();
                pb.fv[i] = it->second.first;
                pb.fl[i] = it->second.second;
            }
            // point order: the product set dimension 0 fastest, then the hanging points
            std::vector<int32_t> order((size_t)nb, -1);
            std::vector<char> used((size_t)nb, 0);
            for (int32_t p : act) {
                int64_t r = 0, mul = 1;
                for (int i = 0; i < d; ++i) {
                    const int32_t k = (int32_t)(std::lower_bound(S[i].begin(), S[i].end(), loc[(size_t)p * d + i]) -
                                                S[i].begin());
                    r += k * mul;
                    mul *= S[i].size();
                }
                order[(size_t)r] = p;
                used[(size_t)p] = 1;
            }
            int64_t fill = prod;
            for (int64_t p = 0; p < nb; ++p)
                if (!used[(size_t)p]) order[(size_t)fill++] = (int32_t)p;
            bool identity = true;
            for (int64_t r = 0; r < nb && identity; ++r) identity = order[(size_t)r] == r;
            if (!identity) {
                pb.inv = (int32_t)out.fdPerm.size();
                out.fdPerm.insert(out.fdPerm.end(), order.begin(), order.end());
            }
            pb.np = (int32_t)prod;
            pb.mode = PRE_FD;
            out.maxFd = std::max(out.maxFd, (int)prod);
            out.fdList.push_back((int32_t)out.blocks.size());
        } else if (nb <= maxDense) {
            // Cholesky factor of the block (the identity at hanging points), then its inverse
            // column by column
            const int n = (int)nb;
            // 1D matrices over the positions the block's points take along every dimension
            std::vector<std::vector<int32_t>> P(d);
            std::vector<std::vector<double>> K1(d), M1(d);
            std::vector<int32_t> rk((size_t)n * d);
            for (int i = 0; i < d; ++i) {
                for (int p = 0; p < n; ++p) P[i].push_back(loc[(size_t)p * d + i]);
                std::sort(P[i].begin(), P[i].end());
                P[i].erase(std::unique(P[i].begin(), P[i].end()), P[i].end());
                for (int p = 0; p < n; ++p)
                    rk[(size_t)p * d + i] = (int32_t)(std::lower_bound(P[i].begin(), P[i].end(),
                                                                       loc[(size_t)p * d + i]) - P[i].begin());
                prewavelet1DMatrices(T[i], P[i], K1[i], M1[i]);
            }
            std::vector<double> A((size_t)n * n, 0.0), L((size_t)n * n, 0.0);
            for (int a = 0; a < n; ++a) A[(size_t)a * n + a] = 1.0;
            for (int32_t a : act)
                for (int32_t b : act) {
                    double sum = 0.0;
                    for (int i = 0; i < d; ++i) {
                        double prod = 1.0;
                        for (int j = 0; j < d && prod != 0.0; ++j) {
                            const size_t e = (size_t)rk[(size_t)a * d + j] * P[j].size() + rk[(size_t)b * d + j];
                            prod *= j == i ? K1[j][e] : M1[j][e];
                        }
                        sum += prod;
                    }
                    A[(size_t)a * n + b] = sum;
                }
            for (int j = 0; j < n; ++j) {
                double v0 = A[(size_t)j * n + j];
                for (int q = 0; q < j; ++q) v0 -= L[(size_t)j * n + q] * L[(size_t)j * n + q];
                if (!(v0 > 0.0)) { err = "depth block not positive definite"; return LSG_ERR_ARG; }
                L[(size_t)j * n + j] = std::sqrt(v0);
                for (int i = j + 1; i < n; ++i) {
                    double v = A[(size_t)i * n + j];
                    for (int q = 0; q < j; ++q) v -= L[(size_t)i * n + q] * L[(size_t)j * n + q];
                    L[(size_t)i * n + j] = v / L[(size_t)j * n + j];
                }
            }
            pb.mode = PRE_DENSE;
            pb.inv = (int32_t)out.inv.size();
            out.inv.resize(out.inv.size() + (size_t)n * n);
            double* inv = &out.inv[(size_t)pb.inv];
            std::vector<double> y(n);
            for (int k = 0; k < n; ++k) {   // column k of the inverse: L L^T x = e_k
                for (int i = 0; i < n; ++i) {
                    double v = (i == k) ? 1.0 : 0.0;
                    for (int q = 0; q < i; ++q) v -= L[(size_t)i * n + q] * y[q];
                    y[i] = v / L[(size_t)i * n + i];
                }
                for (int i = n - 1; i >= 0; --i) {
                    double v = y[i];
                    for (int q = i + 1; q < n; ++q) v -= L[(size_t)q * n + i] * inv[(size_t)k * n + q];
                    inv[(size_t)k * n + i] = v / L[(size_t)i * n + i];
                }
            }
            out.maxN = std::max(out.maxN, n);
        }
        out.blocks.push_back(pb);
    }
    if ((int64_t)out.inv.size() > INT32_MAX || (int64_t)out.fdV.size() > INT32_MAX ||
        (int64_t)out.fdPerm.size() > INT32_MAX) {
        err = "depth-block preconditioner tables too large";
        return LSG_ERR_UNSUPPORTED;
    }
    return LSG_OK;
}

int buildSinglePlan(Grid& g, std::string& err) {
    if (g.planSingle) return 0;
    if (g.sparse) { err = "sparse grids have single-item plans already"; return LSG_ERR_ARG; }
    auto P = std::make_shared<Plan>();
    if (int rc = buildPlan(g, *P, true, err)) return rc;
    g.planSingle = P;
    return 0;
}


// ---------------------------------------------------------------------------
// Basis transformations (BasisTransformations/BasisTransformations.cpp)
// ---------------------------------------------------------------------------
namespace {

// LU factors without pivoting of the level-t line matrix of PrewaveletMatrixHomogen::create
// (sgoperation/matrix.cpp:74-139): column j (1-based) is the basis function at line
// position j in level-t nodal values, the prewavelet (0.1, -0.6, 1, -0.6, 0.1; 0.9 next
// to the boundary, 1 at level 1) at the odd positions and the level-(t-1) hat
// (0.5, 1, 0.5) at the even ones.  The matrix is pentadiagonal and its pivots lie in
// [0.9, 4/3], as in the reference's own unpivoted LR (matrix.cpp:141-182).
// LU factors without pivoting of a pentadiagonal n x n matrix given by its band
// B[r][c - r + 2] (1-based r, c): l1[k] = L[k][k-1], l2[k] = L[k][k-2], u0/u1/u2[k] =
// U[k][k..k+2] (0-based k).  False on a zero pivot.
bool bandFactor(int n, std::vector<std::array<double, 5>>& B, double* l1, double* l2, double* u0, double* u1,
                double* u2) {
    auto at = [&](int r, int c) -> double& { return B[r][c - r + 2]; };
    for (int c = 1; c <= n; ++c) {
        l1[c - 1] = 0.0;
        l2[c - 1] = 0.0;
    }
    for (int c = 1; c <= n; ++c) {
        const double p = at(c, c);
        if (p == 0.0) return false;
        for (int r = c + 1; r <= std::min(n, c + 2); ++r) {
            const double f = at(r, c) / p;
            (r == c + 1 ? l1 : l2)[r - 1] = f;
            for (int k = c; k <= std::min(n, c + 2); ++k) at(r, k) -= f * at(c, k);
        }
        u0[c - 1] = p;
        u1[c - 1] = c + 1 <= n ? at(c, c + 1) : 0.0;
        u2[c - 1] = c + 2 <= n ? at(c, c + 2) : 0.0;
    }
    return true;
}

// LU factors without pivoting of the level-t line matrix of PrewaveletMatrixHomogen::create
// (sgoperation/matrix.cpp:74-139): column j (1-based) is the basis function at line
// position j in level-t nodal values, the prewavelet (0.1, -0.6, 1, -0.6, 0.1; 0.9 next
// to the boundary, 1 at level 1) at the odd positions and the level-(t-1) hat
// (0.5, 1, 0.5) at the even ones.  The matrix is pentadiagonal and its pivots lie in
// [0.9, 4/3], as in the reference's own unpivoted LR (matrix.cpp:141-182).
void lineFactors(int t, double* l1, double* l2, double* u0, double* u1, double* u2) {
    const int n = (1 << t) - 1;
    std::vector<std::array<double, 5>> B(n + 2);
    for (auto& r : B) r.fill(0.0);
    auto at = [&](int r, int c) -> double& { return B[r][c - r + 2]; };
    for (int j = 1; j <= n; ++j) {
        auto set = [&](int r, double v) { if (r >= 1 && r <= n) at(r, j) = v; };
        if (j & 1) {
            if (t == 1) { set(j, 1.0); continue; }
            set(j, (j == 1 || j == n) ? 0.9 : 1.0);
            set(j - 1, -0.6); set(j + 1, -0.6); set(j - 2, 0.1); set(j + 2, 0.1);
        } else {
            set(j, 1.0); set(j - 1, 0.5); set(j + 1, 0.5);
        }
    }
    bandFactor(n, B, l1, l2, u0, u1, u2);
}

// ---- adaptive grids

struct AdaptiveBox {
    const TrItem& it;
    const int32_t* gidx;     // this box's entries
    const uint8_t* st;
    int d;
    // box entry of level-T positions k (-1 outside the interior)
    int64_t entry(const int32_t* k) const {
        int64_t p = 0;
        for (int i = 0; i < d; ++i) {
            if (k[i] < 1 || k[i] > it.n[i]) return -1;
            p += int64_t(k[i] - 1) * it.S[i];
        }
        return p;
    }
    bool work(const int32_t* k) const {   // stored and worked on (active)
        const int64_t p = entry(k);
        return p >= 0 && st[p] == TR_ACTIVE;
    }
    bool ownDepth(const int32_t* k) const {
        for (int i = 0; i < d; ++i) if (!(k[i] & 1)) return false;
        return true;
    }
};

// Components of the depth-T points (ZusammenhangsKomponente::findAnicht and
// recursiveMarkArbeite, sgrid/komponente.cpp:9-73): in the order of their first point
// among the depth-T points in slot order, each the closure of its points under the
// level-T 5^d stencil of IndexDimension::nextFive (indices/index.cc:124-176: a step along
// dimension i is taken only when the first level-T step stays inside); the box spans
// every stored, worked-on stencil point met.  Line matrices per dimension follow
// GetPrewaveletCoefficients and PrewaveletMatrixHomogen::create
// (BasisTransformations.cpp:100-172, sgoperation/matrix.cpp:74-139).
int adaptiveComponents(const AdaptiveBox& B, const std::vector<std::array<int32_t, MAXD>>& ownPts, TrTables& tr,
                       std::string& err) {
    const int d = B.d;
    const TrItem& it = B.it;
    std::vector<char> marked((size_t)it.total, 0);
    for (const auto& k0 : ownPts) {
        const int64_t p0 = B.entry(k0.data());
        if (marked[p0] || B.st[p0] != TR_ACTIVE) continue;
        TrComp c{};
        for (int i = 0; i < d; ++i) c.mn[i] = c.mx[i] = k0[i];
        std::vector<std::array<int32_t, MAXD>> stack{k0};
        marked[p0] = 1;
        while (!stack.empty()) {
            const auto k = stack.back();
            stack.pop_back();
            int o[MAXD];
            for (int i = 0; i < d; ++i) o[i] = -2;
            while (true) {
                bool ok = true;
                int32_t q[MAXD];
                for (int i = 0; i < d && ok; ++i) {
                    if (o[i] < 0 && k[i] - 1 < 1) ok = false;
                    if (o[i] > 0 && k[i] + 1 > it.n[i]) ok = false;
                    q[i] = k[i] + o[i];
                }
                if (ok && B.work(q)) {
                    const int64_t pq = B.entry(q);
                    if (!marked[pq]) {
                        for (int i = 0; i < d; ++i) { c.mn[i] = std::min(c.mn[i], q[i]); c.mx[i] = std::max(c.mx[i], q[i]); }
                        if (B.ownDepth(q)) {
                            marked[pq] = 1;
                            std::array<int32_t, MAXD> nq{};
                            for (int i = 0; i < d; ++i) nq[i] = q[i];
                            stack.push_back(nq);
                        }
                    }
                }
                int i = 0;
                while (i < d && ++o[i] > 2) o[i++] = -2;
                if (i == d) break;
            }
        }
        // line matrices through the lower corner, one per dimension
        for (int dim = 0; dim < d; ++dim) {
            int32_t k[MAXD];
            for (int i = 0; i < d; ++i) k[i] = c.mn[i];
            int size = 1;
            for (;;) {
                ++k[dim];
                if (k[dim] > it.n[dim] || !B.work(k)) break;
                ++size;
            }
            c.size[dim] = size;
            std::vector<std::array<double, 5>> M(size + 2);
            for (auto& r : M) r.fill(0.0);
            auto set = [&](int r, int col, double v) {   // 0-based row/column, dropped outside the line
                if (r >= 0 && r < size) M[r + 1][col - r + 2] = v;
            };
            const int t = it.T[dim];
            for (int i = 0; i < d; ++i) k[i] = c.mn[i];
            auto workAt = [&](int off) {
                int32_t kk[MAXD];
                for (int i = 0; i < d; ++i) kk[i] = k[i];
                kk[dim] += off;
                return B.work(kk);
            };
            for (int col = 0; col < size && B.work(k); ++col, ++k[dim]) {
                if (k[dim] & 1) {   // depth T_dim: prewavelet column
                    if (k[dim] == 1) {
                        set(col, col, 0.9);
                        if (t == 1) { set(col, col, 1.0); break; }
                        if (workAt(1)) set(col + 1, col, -0.6);
                        if (workAt(2)) set(col + 2, col, 0.1);
                    } else if (k[dim] == it.n[dim]) {
                        set(col, col, 0.9);
                        if (workAt(-1)) set(col - 1, col, -0.6);
                        if (workAt(-2)) set(col - 2, col, 0.1);
                    } else {
                        set(col, col, 1.0);
                        if (workAt(-1)) set(col - 1, col, -0.6);
                        if (workAt(-2)) set(col - 2, col, 0.1);
                        if (workAt(1)) set(col + 1, col, -0.6);
                        if (workAt(2)) set(col + 2, col, 0.1);
                    }
                } else {            // coarser: hat column
                    set(col, col, 1.0);
                    if (workAt(-1)) set(col - 1, col, 0.5);
                    if (workAt(1)) set(col + 1, col, 0.5);
                }
            }
            c.fac[dim] = tr.afacTot;
            tr.afacTot += size;
            const size_t o = tr.afac.size();
            tr.afac.resize(o + 5 * (size_t)size);
            double* f = tr.afac.data() + o;
            if (!bandFactor(size, M, f, f + size, f + 2 * size, f + 3 * size, f + 4 * size)) {
                err = "singular prewavelet line matrix (a component's lower box corner is not an active point)";
                return LSG_ERR_UNSUPPORTED;
            }
        }
        tr.comps.push_back(c);
    }
    return LSG_OK;
}

int buildTransformAdaptive(const Grid& g, TrTables& tr, std::string& err) {
    const int d = g.d;
    const int nd = (int)g.depths.size();
    tr.adaptive = true;
    tr.items.assign(nd, TrItem{});
    std::unordered_map<Key2, int32_t, Key2Hash> devOf;
    devOf.reserve((size_t)g.n * 2);
    for (int64_t s = 0; s < g.n; ++s) devOf[packKey(&g.idx[s * d], d)] = (int32_t)g.devOfSlot[s];
    auto find = [&](const int32_t* v) -> int32_t {
        auto f = devOf.find(packKey(v, d));
        return f == devOf.end() ? -1 : f->second;
    };
    std::vector<uint8_t> actDev(g.n);
    for (int64_t s = 0; s < g.n; ++s) actDev[g.devOfSlot[s]] = g.active[s];
    int64_t off = 0;
    tr.maxBox = 0;
    for (int t = 0; t < nd; ++t) {
        TrItem& it = tr.items[t];
        it.d = d;
        int64_t S = 1;
        for (int i = 0; i < d; ++i) {
            it.T[i] = g.depths[t][i];
            it.n[i] = (1 << it.T[i]) - 1;
            it.S[i] = (int32_t)S;
            S *= it.n[i];
        }
        if (S >= (int64_t)INT32_MAX) { err = "depth box too large"; return LSG_ERR_CAPACITY; }
        it.total = (int32_t)S;
        it.box_off = off;
        off += S;
        tr.maxBox = std::max<int64_t>(tr.maxBox, S);
    }
    tr.boxTot = off;
    if (off >= (int64_t)INT32_MAX) { err = "depth boxes too large"; return LSG_ERR_CAPACITY; }
    tr.gidx.assign((size_t)off, -1);
    tr.bstat.assign((size_t)off, TR_MISSING);
    for (int t = 0; t < nd; ++t) {
        const TrItem& it = tr.items[t];
        for (int32_t p = 0; p < it.total; ++p) {
            int32_t v[MAXD];
            int32_t r = p;
            for (int i = 0; i < d; ++i) {
                v[i] = idxAt(r % it.n[i] + 1, it.T[i]);
                r /= it.n[i];
            }
            const int32_t q = find(v);
            tr.gidx[it.box_off + p] = q;
            if (q >= 0) tr.bstat[it.box_off + p] = actDev[q] ? TR_ACTIVE : TR_HANGING;
        }
    }
    // poles along every dimension (missing or boundary: -1); a hanging point is never
    // changed by a (de)hierarchization, so its poles are -1 as well
    tr.poles.assign((size_t)d * g.n * 2, -1);
    for (int64_t s = 0; s < g.n; ++s) {
        if (!g.active[s]) continue;
        const int64_t dev = g.devOfSlot[s];
        int32_t v[MAXD];
        for (int i = 0; i < d; ++i) v[i] = g.idx[s * d + i];
        for (int i = 0; i < d; ++i) {
            const int t = depth1(v[i]);
            const int32_t k = posAt(v[i], t), keep = v[i];
            if (k - 1 >= 1) { v[i] = idxAt(k - 1, t); tr.poles[((size_t)i * g.n + dev) * 2] = find(v); }
            if (k + 1 <= (1 << t) - 1) { v[i] = idxAt(k + 1, t); tr.poles[((size_t)i * g.n + dev) * 2 + 1] = find(v); }
            v[i] = keep;
        }
    }
    // the reference's order: by |T|_1 descending, equal |T|_1 in reverse first-appearance
    // order (ListOfDepthOrderedSubgrids from gotoEnd() through previous(),
    // sgrid/ListOfDepthOrderedGrids.cc:77-137, 219-244)
    tr.seq.resize(nd);
    for (int t = 0; t < nd; ++t) tr.seq[t] = nd - 1 - t;
    auto l1 = [&](int t) { int s = 0; for (int i = 0; i < d; ++i) s += g.depths[t][i]; return s; };
    std::stable_sort(tr.seq.begin(), tr.seq.end(), [&](int a, int b) { return l1(a) > l1(b); });
    // components per depth, from the depth's points in slot order
    std::vector<std::vector<std::array<int32_t, MAXD>>> own(nd);
    for (int64_t s = 0; s < g.n; ++s) {
        std::array<int32_t, MAXD> k{};
        Depth T{};
        for (int i = 0; i < d; ++i) T[i] = depth1(g.idx[s * d + i]);
        const int t = g.depthIdx.at(T);
        for (int i = 0; i < d; ++i) k[i] = posAt(g.idx[s * d + i], T[i]);
        own[t].push_back(k);
    }
    tr.compOff.assign(nd + 1, 0);
    tr.comps.clear();
    tr.afac.clear();
    tr.afacTot = 0;
    for (int t = 0; t < nd; ++t) {
        tr.compOff[t] = (int32_t)tr.comps.size();
        AdaptiveBox B{tr.items[t], tr.gidx.data() + tr.items[t].box_off, tr.bstat.data() + tr.items[t].box_off, d};
        if (int rc = adaptiveComponents(B, own[t], tr, err)) return rc;
    }
    tr.compOff[nd] = (int32_t)tr.comps.size();
    // the factor arrays were appended per line as (l1, l2, u0, u1, u2) blocks; regroup
    // them into five arrays of afacTot entries
    {
        std::vector<double> f(5 * (size_t)tr.afacTot);
        size_t o = 0;
        for (const TrComp& c : tr.comps)
            for (int i = 0; i < d; ++i) {
                for (int a = 0; a < 5; ++a)
                    for (int q = 0; q < c.size[i]; ++q) f[(size_t)a * tr.afacTot + c.fac[i] + q] = tr.afac[o + (size_t)a * c.size[i] + q];
                o += 5 * (size_t)c.size[i];
            }
        tr.afac.swap(f);
    }
    // points by their depth along each dimension, for the dehierarchization
    tr.byLevel.assign((size_t)d * g.n, 0);
    tr.lvlOff.assign((size_t)d * (MAXLEV + 2), 0);
    for (int i = 0; i < d; ++i) {
        int64_t* lo = &tr.lvlOff[(size_t)i * (MAXLEV + 2)];
        std::vector<int64_t> cnt(MAXLEV + 2, 0);
        for (int t = 0; t < nd; ++t) cnt[g.depths[t][i]] += g.blkSize[t];
        for (int t = 1; t <= MAXLEV + 1; ++t) lo[t] = lo[t - 1] + cnt[t - 1];
        std::vector<int64_t> fill(lo, lo + MAXLEV + 2);
        for (int t = 0; t < nd; ++t)
            for (int64_t q = 0; q < g.blkSize[t]; ++q)
                tr.byLevel[(size_t)i * g.n + fill[g.depths[t][i]]++] = (int32_t)(g.blkOff[t] + q);
    }
    // calcNodalByPrew: the box results summed per stored active point (one group over
    // all depths)
    {
        std::vector<std::pair<int32_t, int32_t>> pairs;
        for (int t = 0; t < nd; ++t) {
            const TrItem& it = tr.items[t];
            for (int32_t p = 0; p < it.total; ++p)
                if (tr.bstat[it.box_off + p] == TR_ACTIVE) pairs.push_back({tr.gidx[it.box_off + p], (int32_t)(it.box_off + p)});
        }
        std::stable_sort(pairs.begin(), pairs.end(),
                         [](const std::pair<int32_t, int32_t>& a, const std::pair<int32_t, int32_t>& b) { return a.first < b.first; });
        tr.rowDev.clear();
        tr.col.clear();
        tr.rowPtr.assign(1, 0);
        for (size_t k = 0; k < pairs.size(); ++k) {
            if (k == 0 || pairs[k].first != pairs[k - 1].first) {
                if (k) tr.rowPtr.push_back((int64_t)tr.col.size());
                tr.rowDev.push_back(pairs[k].first);
            }
            tr.col.push_back(pairs[k].second);
        }
        if (!pairs.empty()) tr.rowPtr.push_back((int64_t)tr.col.size());
        tr.groups.assign(1, {0, (int64_t)tr.rowDev.size()});
    }
    tr.levelList.clear();
    tr.levels.clear();
    tr.levelMaxBox.clear();
    tr.fac.clear();
    tr.facTot = 0;
    return LSG_OK;
}

}  // namespace

int buildTransform(const Grid& g, TrTables& tr, std::string& err) {
    tr = TrTables();
    if (g.sparse) return buildTransformAdaptive(g, tr, err);
    const int d = g.d;
    const int nd = (int)g.depths.size();
    tr.items.assign(nd, TrItem{});
    tr.gidx.assign((size_t)g.nml, 0);
    tr.boxTot = g.nml;
    tr.maxBox = 0;
    for (int t = 0; t < nd; ++t) {
        TrItem& it = tr.items[t];
        it.d = d;
        int32_t S = 1;
        for (int i = 0; i < d; ++i) {
            it.T[i] = g.depths[t][i];
            it.n[i] = (1 << it.T[i]) - 1;
            it.S[i] = S;
            S *= it.n[i];
        }
        it.total = S;
        it.box_off = g.boxOff[t];
        tr.maxBox = std::max<int64_t>(tr.maxBox, S);
        for (int32_t p = 0; p < S; ++p) {
            Depth Q{};
            int64_t loc = 0, mul = 1;
            int32_t r = p;
            for (int i = 0; i < d; ++i) {
                const int k = r % it.n[i] + 1;
                r /= it.n[i];
                const int tz = __builtin_ctz((unsigned)k);
                Q[i] = it.T[i] - tz;
                loc += int64_t(k >> (tz + 1)) * mul;
                mul *= int64_t(1) << (Q[i] - 1);
            }
            auto f = g.depthIdx.find(Q);
            if (f == g.depthIdx.end()) { err = "depth box not complete"; return LSG_ERR_UNSUPPORTED; }
            tr.gidx[it.box_off + p] = (int32_t)(g.blkOff[f->second] + loc);
        }
    }
    // poles: the neighbours of a point of depth T inside its own box G_T
    tr.poles.assign((size_t)d * g.n * 2, -1);
    for (int t = 0; t < nd; ++t) {
        const TrItem& it = tr.items[t];
        for (int32_t p = 0; p < it.total; ++p) {
            int32_t r = p;
            bool own = true;
            int k[MAXD];
            for (int i = 0; i < d; ++i) {
                k[i] = r % it.n[i];
                r /= it.n[i];
                if (k[i] & 1) own = false;   // 0-based even = depth T_i
            }
            if (!own) continue;
            const int32_t dev = tr.gidx[it.box_off + p];
            for (int i = 0; i < d; ++i) {
                if (k[i] >= 1) tr.poles[((size_t)i * g.n + dev) * 2] = tr.gidx[it.box_off + p - it.S[i]];
                if (k[i] + 1 < it.n[i]) tr.poles[((size_t)i * g.n + dev) * 2 + 1] = tr.gidx[it.box_off + p + it.S[i]];
            }
        }
    }
    // points grouped by their depth along each dimension (coarse to fine), for the
    // level-by-level dehierarchization
    tr.byLevel.assign((size_t)d * g.n, 0);
    tr.lvlOff.assign((size_t)d * (MAXLEV + 2), 0);
    for (int i = 0; i < d; ++i) {
        int64_t* off = &tr.lvlOff[(size_t)i * (MAXLEV + 2)];
        std::vector<int64_t> cnt(MAXLEV + 2, 0);
        for (int t = 0; t < nd; ++t) cnt[g.depths[t][i]] += g.blkSize[t];
        for (int t = 1; t <= MAXLEV + 1; ++t) off[t] = off[t - 1] + cnt[t - 1];
        std::vector<int64_t> fill(off, off + MAXLEV + 2);
        for (int t = 0; t < nd; ++t)
            for (int64_t q = 0; q < g.blkSize[t]; ++q)
                tr.byLevel[(size_t)i * g.n + fill[g.depths[t][i]]++] = (int32_t)(g.blkOff[t] + q);
    }
    // depths grouped by |T|_1, finest first
    std::map<int, std::vector<int32_t>, std::greater<int>> bySum;
    for (int t = 0; t < nd; ++t) {
        int s = 0;
        for (int i = 0; i < d; ++i) s += g.depths[t][i];
        bySum[s].push_back(t);
    }
    tr.levelList.clear();
    tr.levels.clear();
    tr.levelMaxBox.clear();
    tr.rowDev.clear();
    tr.col.clear();
    tr.rowPtr.assign(1, 0);
    tr.groups.clear();
    auto addGroup = [&](const std::vector<int32_t>& ts) {
        std::vector<std::pair<int32_t, int32_t>> pairs;   // (device index, scratch position)
        for (int32_t t : ts) {
            const TrItem& it = tr.items[t];
            for (int32_t p = 0; p < it.total; ++p)
                pairs.push_back({tr.gidx[it.box_off + p], (int32_t)(it.box_off + p)});
        }
        std::stable_sort(pairs.begin(), pairs.end(),
                         [](const std::pair<int32_t, int32_t>& a, const std::pair<int32_t, int32_t>& b) { return a.first < b.first; });
        const int64_t first = (int64_t)tr.rowDev.size();
        for (size_t k = 0; k < pairs.size(); ++k) {
            if (k == 0 || pairs[k].first != pairs[k - 1].first) {
                if (k) tr.rowPtr.push_back((int64_t)tr.col.size());
                tr.rowDev.push_back(pairs[k].first);
            }
            tr.col.push_back(pairs[k].second);
        }
        if (!pairs.empty()) tr.rowPtr.push_back((int64_t)tr.col.size());
        tr.groups.push_back({first, (int64_t)tr.rowDev.size() - first});
    };
    std::vector<int32_t> all;
    for (auto& kv : bySum) {
        tr.levels.push_back({(int32_t)tr.levelList.size(), (int32_t)kv.second.size()});
        int64_t mb = 0;
        for (int32_t t : kv.second) mb = std::max<int64_t>(mb, tr.items[t].total);
        tr.levelMaxBox.push_back(mb);
        tr.levelList.insert(tr.levelList.end(), kv.second.begin(), kv.second.end());
        addGroup(kv.second);
        all.insert(all.end(), kv.second.begin(), kv.second.end());
    }
    std::sort(all.begin(), all.end());
    addGroup(all);
    // line-matrix factors
    tr.facTot = facBase(g.maxLevel + 1);
    tr.fac.assign((size_t)5 * tr.facTot, 0.0);
    for (int t = 1; t <= g.maxLevel; ++t) {
        double* b = tr.fac.data() + facBase(t);
        lineFactors(t, b, b + tr.facTot, b + 2 * tr.facTot, b + 3 * tr.facTot, b + 4 * tr.facTot);
    }
    return LSG_OK;
}

}  // namespace lsg
