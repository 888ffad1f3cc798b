// lsg_gridgen.cpp -- host producer of locally adaptive sparse grids with zero Dirichlet
// boundaries: the point set and slot order that the reference's AdaptiveSparseGrid builds
// from AddPoint / AddRecursiveSonsOfPoint followed by completeDirichletGrid
// (sgrid/sparseGrid.cc:482-566, 587-621, 740-787; sgrid/localTensorProduct.cc:29-180,
// 189-213; sgrid/GridGeneration.cpp:12-38, 124-160, 333-378), and the refinement towards
// a corner used for the locally adaptive benchmark grids.
//
// Slots are the reference's data-table order: the reference hands out the lowest free
// entry of its second table and never removes a point, so slot = insertion order, and
// every step below adds points in the order the reference's loops reach them.
//
// Index algebra (indices/index.h:657-876): 1D index 0 / 1 are the boundaries, interior
// index v of depth t = floor(log2 v) sits at level-t position 2 (v - 2^t) + 1; a point is
// stored with all indices >= 2 (interior).  Positions below are level-t positions.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/loadsg_amd.h"

namespace lsg {
namespace {

constexpr int GMAXD = 8;

inline int depthOf(int32_t v) { return v <= 1 ? 0 : 31 - __builtin_clz((uint32_t)v); }
inline int32_t posOf(int32_t v, int t) {   // level-t position of index v (depth <= t)
    if (v == 0) return 0;
    if (v == 1) return 1 << t;
    const int s = depthOf(v);
    return (2 * (v - (1 << s)) + 1) << (t - s);
}
inline int32_t idxOf(int32_t k, int t) {   // index of level-t position k (0 and 2^t: boundary)
    if (k <= 0) return 0;
    if (k >= (1 << t)) return 1;
    const int tz = __builtin_ctz((uint32_t)k);
    return (1 << (t - tz)) + (k >> (tz + 1));
}
// own-level neighbours (IndexOneD::nextLeft / nextRight without a level)
inline int32_t ownLeft(int32_t v) { return v >> (__builtin_ctz((uint32_t)v) + 1); }
inline int32_t ownRight(int32_t v) { return v >> (__builtin_ctz(~(uint32_t)v) + 1); }

struct Key {
    uint64_t a = 0, b = 0;
    bool operator==(const Key& o) const { return a == o.a && b == o.b; }
};
struct KeyHash {
    size_t operator()(const Key& k) const {
        uint64_t h = k.a * 0x9E3779B97F4A7C15ull ^ (k.b + 0x632BE59BD9B4E019ull + (k.a << 6) + (k.a >> 2));
        return (size_t)(h ^ (h >> 29));
    }
};

}  // namespace

// The growing grid.
struct GridGen {
    int d = 0;
    std::vector<int32_t> idx;       // slot-major, d per slot
    std::vector<uint8_t> active;
    std::unordered_map<Key, int64_t, KeyHash> slotOf;
    // points of each depth in insertion order (MultiDepthHashGrid, _mapPosToGridPos)
    std::unordered_map<Key, std::vector<int64_t>, KeyHash> byDepth;

    int64_t n() const { return (int64_t)active.size(); }
    const int32_t* pt(int64_t s) const { return &idx[(size_t)s * d]; }
    Key key(const int32_t* v) const {
        Key k;
        for (int i = 0; i < d; ++i) {
            const uint64_t x = (uint64_t)(uint16_t)v[i];
            if (i < 4) k.a |= x << (16 * i);
            else k.b |= x << (16 * (i - 4));
        }
        return k;
    }
    Key depthKey(const int32_t* v) const {
        int32_t T[GMAXD];
        for (int i = 0; i < d; ++i) T[i] = depthOf(v[i]);
        return key(T);
    }
    int64_t find(const int32_t* v) const {
        auto f = slotOf.find(key(v));
        return f == slotOf.end() ? -1 : f->second;
    }
    static bool interior(const int32_t* v, int d) {
        for (int i = 0; i < d; ++i)
            if (v[i] < 2) return false;
        return true;
    }
    int64_t append(const int32_t* v, bool act) {
        const int64_t s = n();
        idx.insert(idx.end(), v, v + d);
        active.push_back(act ? 1 : 0);
        slotOf.emplace(key(v), s);
        byDepth[depthKey(v)].push_back(s);
        return s;
    }
    // AddPoint(I) (sparseGrid.cc:482-517): a new point is active, an existing one is made
    // active; true if the point is new
    bool addPoint(const int32_t* v) {
        const int64_t s = find(v);
        if (s >= 0) {
            active[s] = 1;
            return false;
        }
        append(v, true);
        return true;
    }
    // AddPoint(I, flag) (sparseGrid.cc:519-555): a new point is active iff flag; an
    // existing one is left as it is
    bool addPoint(const int32_t* v, bool flag) {
        if (find(v) >= 0) return false;
        append(v, flag);
        return true;
    }
};

namespace {

// AddRecursiveSonsOfPoint (sparseGrid.cc:557-566): the point, then depth-first the sons
// in dimension order, left before right
void addRecursiveSons(GridGen& g, std::array<int32_t, GMAXD> I, int level) {
    g.addPoint(I.data());
    if (level <= 0) return;
    for (int i = 0; i < g.d; ++i) {
        std::array<int32_t, GMAXD> J = I;
        J[i] = 2 * I[i];
        addRecursiveSons(g, J, level - 1);
        J[i] = 2 * I[i] + 1;
        addRecursiveSons(g, J, level - 1);
    }
}

// The level-T 5^d stencil around a point of depth T, visited as a MultiDimFiveCompass loop
// over IndexDimension::nextFive (indices/index.h:436-480, index.cc:124-176): directions
// (0, -1, +1, -2, +2) per dimension, dimension 0 fastest; a direction whose first level-T
// step reaches the boundary is skipped.
template <typename F>
void fiveStencil(const GridGen& g, const int32_t* I, const int* T, F&& visit) {
    const int d = g.d;
    int32_t k0[GMAXD];
    for (int i = 0; i < d; ++i) k0[i] = posOf(I[i], T[i]);
    static const int off[5] = {0, -1, 1, -2, 2};
    int dir[GMAXD] = {};
    while (true) {
        bool ok = true;
        int32_t J[GMAXD];
        for (int i = 0; i < d && ok; ++i) {
            const int o = off[dir[i]];
            if (o < 0 && k0[i] - 1 < 1) ok = false;
            if (o > 0 && k0[i] + 1 > (1 << T[i]) - 1) ok = false;
            J[i] = idxOf(k0[i] + o, T[i]);
        }
        if (ok) visit(J);
        int i = 0;
        while (i < d && ++dir[i] == 5) dir[i++] = 0;
        if (i == d) break;
    }
}

// completeGridWithoutBoundary (sparseGrid.cc:740-769): the own-level neighbours of every
// point, repeated until nothing is added
void completeWithoutBoundary(GridGen& g) {
    for (bool changed = true; changed;) {
        changed = false;
        const int64_t end = g.n();
        for (int64_t s = 0; s < end; ++s) {
            for (int i = 0; i < g.d; ++i) {
                int32_t J[GMAXD];
                std::memcpy(J, g.pt(s), sizeof(int32_t) * g.d);
                J[i] = ownLeft(g.pt(s)[i]);
                if (GridGen::interior(J, g.d) && g.find(J) < 0 && g.addPoint(J)) changed = true;
                std::memcpy(J, g.pt(s), sizeof(int32_t) * g.d);
                J[i] = ownRight(g.pt(s)[i]);
                if (GridGen::interior(J, g.d) && g.find(J) < 0 && g.addPoint(J)) changed = true;
            }
        }
    }
}

// Depths in the order ListOfDepthOrderedSubgrids visits them from gotoBegin() through
// next() (sgrid/ListOfDepthOrderedGrids.cc:77-137, 246-270): by |T|_1 ascending, each
// |T|_1 in first-appearance order over the slots.
std::vector<std::array<int, GMAXD>> depthsAscending(const GridGen& g) {
    std::vector<std::array<int, GMAXD>> list;
    std::unordered_map<Key, int, KeyHash> seen;
    for (int64_t s = 0; s < g.n(); ++s) {
        const Key k = g.depthKey(g.pt(s));
        if (seen.count(k)) continue;
        seen[k] = 1;
        std::array<int, GMAXD> T{};
        for (int i = 0; i < g.d; ++i) T[i] = depthOf(g.pt(s)[i]);
        list.push_back(T);
    }
    auto l1 = [&](const std::array<int, GMAXD>& T) { int s = 0; for (int i = 0; i < g.d; ++i) s += T[i]; return s; };
    std::stable_sort(list.begin(), list.end(), [&](const auto& a, const auto& b) { return l1(a) < l1(b); });
    return list;
}

// CompleteToLocalTensorProductGrid (localTensorProduct.cc:189-213, LocalForTensor
// :29-180): per depth T, every active point of depth T not yet visited (scanning the slots
// upwards, new points included) spans the box of its active level-T stencil neighbours;
// every level-T position of the box is added (active) or made active.
void completeLocalTensor(GridGen& g) {
    const int d = g.d;
    std::vector<char> marked;
    for (const auto& Tarr : depthsAscending(g)) {
        const int* T = Tarr.data();
        auto isT = [&](int64_t s) {
            for (int i = 0; i < d; ++i)
                if (depthOf(g.pt(s)[i]) != T[i]) return false;
            return true;
        };
        for (int64_t s = 0; s < g.n(); ++s) {
            if ((int64_t)marked.size() < g.n()) marked.resize(g.n(), 0);
            if (marked[s] || !g.active[s] || !isT(s)) continue;
            marked[s] = 1;
            int32_t mn[GMAXD], mx[GMAXD];
            for (int i = 0; i < d; ++i) mn[i] = mx[i] = posOf(g.pt(s)[i], T[i]);
            fiveStencil(g, g.pt(s), T, [&](const int32_t* J) {
                const int64_t q = g.find(J);
                if (q < 0 || !g.active[q]) return;
                for (int i = 0; i < d; ++i) {
                    const int32_t k = posOf(J[i], T[i]);
                    mn[i] = std::min(mn[i], k);
                    mx[i] = std::max(mx[i], k);
                }
            });
            // checkRechteck: RectangularIterator over the box, dimension 0 fastest
            int32_t k[GMAXD];
            for (int i = 0; i < d; ++i) k[i] = mn[i];
            while (true) {
                int32_t J[GMAXD];
                for (int i = 0; i < d; ++i) J[i] = idxOf(k[i], T[i]);
                const int64_t q = g.find(J);
                if (q < 0) g.addPoint(J, true);
                else g.active[q] = 1;
                int i = 0;
                for (; i < d; ++i) {
                    if (k[i] < mx[i]) { ++k[i]; break; }
                    k[i] = mn[i];
                }
                if (i == d) break;
            }
        }
    }
}

// addHangingNodes (sparseGrid.cc:587-621): the missing level-T stencil points of every
// active point (T its depth) become hanging nodes
void addHangingNodes(GridGen& g) {
    const int64_t end = g.n();
    for (int64_t s = 0; s < end; ++s) {
        if (!g.active[s]) continue;
        int T[GMAXD];
        for (int i = 0; i < g.d; ++i) T[i] = depthOf(g.pt(s)[i]);
        int32_t I[GMAXD];
        std::memcpy(I, g.pt(s), sizeof(int32_t) * g.d);
        fiveStencil(g, I, T, [&](const int32_t* J) {
            if (g.find(J) < 0) g.append(J, false);
        });
    }
}

// support (GridGeneration.cpp:12-38) along one dimension: level-t positions k +- 3 of a
// point at position k of depth t, cut at the boundary, or the whole interval for t <= 2;
// returned as level-L positions (L >= every depth) with L = 16
constexpr int SL = 16;
inline void support1(int32_t v, int32_t& lo, int32_t& hi) {
    const int t = depthOf(v);
    if (t <= 2) { lo = 0; hi = 1 << SL; return; }
    const int32_t k = posOf(v, t);
    lo = k == 1 ? 0 : (k - 3) << (SL - t);
    hi = k == (1 << t) - 1 ? (1 << SL) : (k + 3) << (SL - t);
}

// Stored points of one depth T counted over boxes: a d-dimensional prefix sum over the
// depth's odd level-T positions (2^(T_i - 1) per dimension), rebuilt after points of the
// depth are added.  Depths whose table would exceed 2^24 entries count by scanning.
struct DepthCount {
    int d = 0;
    int32_t ext[GMAXD] = {};
    int64_t stride[GMAXD] = {};
    std::vector<int32_t> ps;
    int64_t builtFor = -1;    // number of the depth's points the table counts
    bool scan = false;

    void build(const GridGen& g, const int* T, const std::vector<int64_t>& pts) {
        d = g.d;
        int64_t tot = 1;
        for (int i = 0; i < d; ++i) {
            ext[i] = 1 << (T[i] - 1);
            stride[i] = tot;
            tot *= ext[i] + 1;
            if (tot > (int64_t(1) << 24)) { scan = true; builtFor = (int64_t)pts.size(); return; }
        }
        ps.assign((size_t)tot, 0);
        for (int64_t s : pts) {
            int64_t o = 0;
            for (int i = 0; i < d; ++i) o += (int64_t)((posOf(g.pt(s)[i], T[i]) - 1) / 2 + 1) * stride[i];
            ps[(size_t)o] += 1;
        }
        for (int i = 0; i < d; ++i)   // inclusive prefix sums along each dimension
            for (int64_t o = 0; o < tot; ++o)
                if ((o / stride[i]) % (ext[i] + 1) > 0) ps[(size_t)o] += ps[(size_t)(o - stride[i])];
        builtFor = (int64_t)pts.size();
    }
    // points with odd positions in [a_i, b_i] (odd bounds, a_i <= b_i)
    int64_t count(const GridGen& g, const int* T, const std::vector<int64_t>& pts, const int32_t* a,
                  const int32_t* b) const {
        if (scan) {
            int64_t c = 0;
            for (int64_t s : pts) {
                bool in = true;
                for (int i = 0; i < d && in; ++i) {
                    const int32_t k = posOf(g.pt(s)[i], T[i]);
                    in = k >= a[i] && k <= b[i];
                }
                c += in;
            }
            return c;
        }
        int64_t c = 0;
        for (unsigned m = 0; m < (1u << d); ++m) {
            int64_t o = 0;
            int sign = 1;
            bool zero = false;
            for (int i = 0; i < d; ++i) {
                int32_t j;
                if ((m >> i) & 1u) { j = (a[i] - 1) / 2; sign = -sign; }   // below the box
                else j = (b[i] - 1) / 2 + 1;
                if (j == 0) { zero = true; break; }
                o += (int64_t)j * stride[i];
            }
            if (!zero) c += sign * (int64_t)ps[(size_t)o];
        }
        return c;
    }
};

// refill (GridGeneration.cpp:157-195) for the active points p (depth T1) and q (depth
// T2), T = max(T1, T2): if their supports overlap and a point of depth T is stored in the
// overlap box, every missing depth-T position of the box becomes a hanging node.
bool refill(GridGen& g, const int32_t* lo1, const int32_t* hi1, const int32_t* lo2, const int32_t* hi2, const int* T,
            const std::vector<int64_t>& ofT, DepthCount& dc) {
    const int d = g.d;
    int32_t a[GMAXD], b[GMAXD];   // first and last odd level-T position of the box
    int64_t odd = 1;
    for (int i = 0; i < d; ++i) {
        if (hi2[i] < lo1[i] || hi1[i] < lo2[i]) return false;
        const int32_t lo = std::max(lo1[i], lo2[i]) >> (SL - T[i]);
        const int32_t hi = std::min(hi1[i], hi2[i]) >> (SL - T[i]);
        a[i] = lo | 1;
        b[i] = (hi & 1) ? hi : hi - 1;
        if (b[i] < a[i]) return false;   // no position of depth T_i in the box
        odd *= (b[i] - a[i]) / 2 + 1;
    }
    // stored depth-T points inside the box: none, or all, and nothing is added
    if (dc.builtFor != (int64_t)ofT.size()) dc.build(g, T, ofT);
    const int64_t inside = dc.count(g, T, ofT, a, b);
    if (inside == 0 || inside == odd) return false;
    bool added = false;
    int32_t k[GMAXD];
    for (int i = 0; i < d; ++i) k[i] = a[i];
    while (true) {   // RectangularIteratorExact order: dimension 0 fastest
        int32_t J[GMAXD];
        for (int i = 0; i < d; ++i) J[i] = idxOf(k[i], T[i]);
        if (g.find(J) < 0) {
            g.append(J, false);
            added = true;
        }
        int i = 0;
        for (; i < d; ++i) {
            if (k[i] + 2 <= b[i]) { k[i] += 2; break; }
            k[i] = a[i];
        }
        if (i == d) break;
    }
    return added;
}

// addPoints2 (GridGeneration.cpp:333-378): for every pair of depths (first-appearance
// order, DepthList, iterator/depthIterator.h:16-48) that differ in at least two
// dimensions, every pair of active points, repeated until nothing is added
void addPoints2(GridGen& g) {
    const int d = g.d;
    std::vector<std::array<int, GMAXD>> list;
    {
        std::unordered_map<Key, int, KeyHash> seen;
        for (int64_t s = 0; s < g.n(); ++s) {
            const Key k = g.depthKey(g.pt(s));
            if (seen.count(k)) continue;
            seen[k] = 1;
            std::array<int, GMAXD> T{};
            for (int i = 0; i < d; ++i) T[i] = depthOf(g.pt(s)[i]);
            list.push_back(T);
        }
    }
    auto keyOfDepth = [&](const int* T) {
        int32_t v[GMAXD];
        for (int i = 0; i < d; ++i) v[i] = T[i];
        return g.key(v);
    };
    // supports of the points, level-SL positions
    std::vector<int32_t> slo, shi;
    std::unordered_map<Key, DepthCount, KeyHash> counts;
    auto supports = [&]() {
        const int64_t n0 = (int64_t)slo.size() / d;
        slo.resize((size_t)g.n() * d);
        shi.resize((size_t)g.n() * d);
        for (int64_t s = n0; s < g.n(); ++s)
            for (int i = 0; i < d; ++i) support1(g.pt(s)[i], slo[(size_t)s * d + i], shi[(size_t)s * d + i]);
    };
    for (bool changed = true; changed;) {
        changed = false;
        for (size_t a = 0; a < list.size(); ++a) {
            for (size_t b = a + 1; b < list.size(); ++b) {
                int differ = 0;
                int T[GMAXD];
                for (int i = 0; i < d; ++i) {
                    differ += list[a][i] != list[b][i];
                    T[i] = std::max(list[a][i], list[b][i]);
                }
                if (differ < 2) continue;
                // a box is filled only when it holds a stored depth-T point, and only refill
                // adds depth-T points: without one now, none appears for this pair
                auto f = g.byDepth.find(keyOfDepth(T));
                if (f == g.byDepth.end()) continue;
                const std::vector<int64_t>& ofT = f->second;   // grows as refill adds
                DepthCount& dc = counts[f->first];
                supports();
                // the point lists grow only by hanging nodes, which are never paired
                std::vector<int64_t> act1, act2;
                for (int64_t p : g.byDepth[keyOfDepth(list[a].data())])
                    if (g.active[p]) act1.push_back(p);
                for (int64_t q : g.byDepth[keyOfDepth(list[b].data())])
                    if (g.active[q]) act2.push_back(q);
                for (int64_t p : act1) {
                    for (int64_t q : act2) {
                        const int32_t *lo1 = &slo[(size_t)p * d], *hi1 = &shi[(size_t)p * d];
                        const int32_t *lo2 = &slo[(size_t)q * d], *hi2 = &shi[(size_t)q * d];
                        bool ov = true;
                        for (int i = 0; i < d && ov; ++i) ov = !(hi2[i] < lo1[i] || hi1[i] < lo2[i]);
                        if (!ov) continue;
                        if (refill(g, lo1, hi1, lo2, hi2, T, ofT, dc)) {
                            changed = true;
                            supports();
                        }
                    }
                }
            }
        }
    }
}

}  // namespace

}  // namespace lsg

// ------------------------------------------------------------------------------ C ABI
using lsg::GridGen;

struct lsg_gridgen {
    GridGen g;
};

extern "C" {

lsg_gridgen* lsg_gridgen_create(int d) {
    if (d < 1 || d > lsg::GMAXD) return nullptr;
    auto* h = new lsg_gridgen;
    h->g.d = d;
    return h;
}

void lsg_gridgen_destroy(lsg_gridgen* h) { delete h; }

static int checkPoint(const lsg_gridgen* h, const int32_t* v) {
    for (int i = 0; i < h->g.d; ++i) {
        if (v[i] < 2) return LSG_ERR_ARG;                      // boundary points are not stored
        if (lsg::depthOf(v[i]) > 15) return LSG_ERR_UNSUPPORTED;
    }
    return LSG_OK;
}

int lsg_gridgen_add_point(lsg_gridgen* h, const int32_t* idx) {
    if (!h || !idx) return LSG_ERR_ARG;
    if (int rc = checkPoint(h, idx)) return rc;
    h->g.addPoint(idx);
    return LSG_OK;
}

int lsg_gridgen_add_recursive_sons(lsg_gridgen* h, const int32_t* idx, int level) {
    if (!h || !idx || level < 0) return LSG_ERR_ARG;
    if (int rc = checkPoint(h, idx)) return rc;
    for (int i = 0; i < h->g.d; ++i)
        if (lsg::depthOf(idx[i]) + level > 15) return LSG_ERR_UNSUPPORTED;
    std::array<int32_t, lsg::GMAXD> I{};
    for (int i = 0; i < h->g.d; ++i) I[i] = idx[i];
    lsg::addRecursiveSons(h->g, I, level);
    return LSG_OK;
}

int lsg_gridgen_refine_towards(lsg_gridgen* h, double radius, int max_level, int corners) {
    if (!h || radius < 0 || max_level < 0 || corners < 1 || corners > 2) return LSG_ERR_ARG;
    GridGen& g = h->g;
    const int d = g.d;
    for (bool changed = true; changed;) {
        changed = false;
        const int64_t end = g.n();
        for (int64_t s = 0; s < end; ++s) {
            if (!g.active[s]) continue;
            // distance from the corner(s) to the support box [x - 2^-t, x + 2^-t]
            double r2 = 0.0, q2 = 0.0;
            for (int i = 0; i < d; ++i) {
                const int32_t v = g.pt(s)[i];
                const int t = lsg::depthOf(v);
                const double x = double(2 * (v - (1 << t)) + 1) / double(1 << t);
                const double lo = x - 1.0 / double(1 << t), hi = 1.0 - (x + 1.0 / double(1 << t));
                if (lo > 0.0) r2 += lo * lo;
                if (hi > 0.0) q2 += hi * hi;
            }
            if (corners == 2) r2 = std::min(r2, q2);
            if (std::sqrt(r2) >= radius) continue;
            int32_t I[lsg::GMAXD];
            std::memcpy(I, g.pt(s), sizeof(int32_t) * d);
            for (int k = 0; k < d; ++k) {
                for (int son = 0; son < 2; ++son) {
                    int32_t J[lsg::GMAXD];
                    std::memcpy(J, I, sizeof(int32_t) * d);
                    J[k] = 2 * I[k] + son;
                    int lsum = 0;
                    for (int q = 0; q < d; ++q) lsum += lsg::depthOf(J[q]) - 1;
                    if (lsum > max_level || lsg::depthOf(J[k]) > 15) continue;
                    if (g.find(J) < 0) {
                        g.addPoint(J);
                        changed = true;
                    }
                }
            }
        }
    }
    return LSG_OK;
}

int lsg_gridgen_complete_dirichlet(lsg_gridgen* h) {
    if (!h) return LSG_ERR_ARG;
    lsg::completeWithoutBoundary(h->g);
    lsg::completeLocalTensor(h->g);
    lsg::completeWithoutBoundary(h->g);
    lsg::addHangingNodes(h->g);
    lsg::addPoints2(h->g);
    return LSG_OK;
}

int64_t lsg_gridgen_size(const lsg_gridgen* h) { return h ? h->g.n() : -1; }

int lsg_gridgen_points(const lsg_gridgen* h, int32_t* idx, uint8_t* active) {
    if (!h || !idx || !active) return LSG_ERR_ARG;
    std::memcpy(idx, h->g.idx.data(), h->g.idx.size() * sizeof(int32_t));
    std::memcpy(active, h->g.active.data(), h->g.active.size());
    return LSG_OK;
}

}  // extern "C"
