// ref_harness.cpp -- drives the compiled LoAdSG reference (built from the
// unmodified sources under /root/reference by oracle/Makefile) to write golden
// vectors as .npy files.  TEST INFRASTRUCTURE ONLY: its outputs are fixtures
// for tests/ and a CPU timing sample for bench.py's cpu_baseline leg.
//
// It uses the reference's own public API exactly as its example drivers do
// (examples_doxygen/poisson_regular.cpp:39-120):
//   AdaptiveSparseGrid::AddRecursiveSonsOfPoint, completeDirichletGrid,
//   MultiLevelAdaptiveSparseGrid, VectorSparseG, calcPrewByNodal,
//   MatrixVectorHomogen(grid, mgrid, true, false)::multiplication,
//   Preconditioning<S>, CG::solveHomogen.
// The variable-coefficient operator is supplied through the reference's plug-in
// hook StencilTemplate::integration / localValue (stencils/Stencil.h:735-748):
// element matrix = a(cell centre) * Poisson + c(cell centre) * HelmHoltz.
//
// Usage:
//   ref_harness regular  <level> <coef> <outdir> [what...]
//   ref_harness adaptive <startlevel> <radius> <maxlevel> <coef> <outdir> [what...]
//   ref_harness adaptive2 <startlevel> <radius> <maxlevel> <coef> <outdir> [what...]
//   ref_harness points <pointfile> <outdir> [what...]
//   ref_harness timemv   <level> <coef> <reps>            (timing sample, prints JSON)
//   ref_harness timetr   <level>                          (calcPrewByNodal / calcNodalByPrew timing, JSON)
//   ref_harness timesolve <level> <coef> [eps]           (the driver's whole PCG solve timed, JSON)
// coef: const:<a0>:<c0> | cfg2 | cfg3 | expsin:<xi1>,<xi2>,... | gen
//       (gen: a = 1 + x1 x2 + sin(pi (x1 + x2)), c = 1 + sin(pi x1 x2) / 2, not separable)
// what: matvec precond cg transform stored   (default: all)
#include "source/sparseEXPDE.h"
#include "ref_plugin.hpp"

#include <fstream>
#include <string>
#include <vector>
#include <cstdio>
#include <cstring>
#include <cmath>
#include <chrono>
#include <sstream>

namespace {

// --- minimal .npy writer ----------------------------------------------------
void writeNpy(const std::string& path, const char* descr, const void* data, size_t elsize,
              const std::vector<size_t>& shape) {
    std::string shp = "(";
    for (size_t i = 0; i < shape.size(); ++i) {
        shp += std::to_string(shape[i]);
        shp += (shape.size() == 1 || i + 1 < shape.size()) ? "," : "";
        if (i + 1 < shape.size()) shp += " ";
    }
    shp += ")";
    std::string hdr = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': " + shp + ", }";
    size_t total = 10 + hdr.size() + 1;
    size_t pad = (64 - total % 64) % 64;
    hdr += std::string(pad, ' ');
    hdr += "\n";
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) { std::perror(path.c_str()); std::exit(3); }
    const char magic[] = "\x93NUMPY";
    std::fwrite(magic, 1, 6, f);
    unsigned char ver[2] = {1, 0};
    std::fwrite(ver, 1, 2, f);
    uint16_t hl = (uint16_t)hdr.size();
    std::fwrite(&hl, 2, 1, f);
    std::fwrite(hdr.data(), 1, hdr.size(), f);
    size_t n = 1;
    for (auto s : shape) n *= s;
    std::fwrite(data, elsize, n, f);
    std::fclose(f);
}

void writeVec(const std::string& dir, const std::string& name, VectorSparseG& v, size_t n) {
    std::vector<double> buf(n);
    for (size_t i = 0; i < n; ++i) buf[i] = v.getValue((unsigned long)i);
    writeNpy(dir + "/" + name + ".npy", "<f8", buf.data(), 8, {n});
}

double sinFunc(IndexDimension I) {       // poisson_regular.cpp:5-11
    double v = 1.0;
    for (int d = 0; d < DimensionSparseGrid; d++) v *= std::sin(M_PI * I.coordinate(d));
    return v;
}

bool want(const std::vector<std::string>& w, const char* s) {
    if (w.empty()) return true;
    for (auto& x : w) if (x == s) return true;
    return false;
}

void runAll(AdaptiveSparseGrid& grid, const CoefSpec& cs, const std::string& out,
            const std::vector<std::string>& what) {
    MultiLevelAdaptiveSparseGrid mgrid(&grid);
    size_t n = grid.getMaximalOccupiedSecondTable();
    const int d = DimensionSparseGrid;
    std::vector<int32_t> idx(n * d);
    std::vector<uint8_t> act(n);
    for (size_t i = 0; i < n; ++i) {
        IndexDimension I = grid.getIndexOfTable(i);
        for (int k = 0; k < d; ++k) idx[i * d + k] = I.getIndex(k);
        act[i] = grid.getActiveTable()[i] ? 1 : 0;
    }
    writeNpy(out + "/idx.npy", "<i4", idx.data(), 4, {n, (size_t)d});
    writeNpy(out + "/active.npy", "|u1", act.data(), 1, {n});

    MatrixVectorHomogen m(grid, mgrid, true, false);
    Poisson poisson(grid);
    HelmHoltz helm(grid);
    VarCoeffStencil var(grid, cs);

    VectorSparseG x(grid), y(grid);
    for (size_t i = 0; i < n; ++i) x.setValue((unsigned long)i, act[i] ? 1.0 / (1.0 + double(i % 7)) : 0.0);
    writeVec(out, "x", x, n);

    if (want(what, "matvec")) {
        y = 0.0; m.multiplication(x, y, poisson); writeVec(out, "Ax_poisson", y, n);
        y = 0.0; m.multiplication(x, y, helm);    writeVec(out, "Ax_helm", y, n);
        y = 0.0; m.multiplication(x, y, var);     writeVec(out, "Ax_var", y, n);
    }
    if (want(what, "precond")) {
        VectorSparseG z1(grid), z2(grid);
        { Preconditioning<Poisson> P1(z1, poisson); }
        { Preconditioning<VarCoeffStencil> P2(z2, var); }
        writeVec(out, "diag_poisson", z1, n);
        writeVec(out, "diag_var", z2, n);
    }
    if (want(what, "stored")) {
        // Stored element matrices (LocalStiffnessMatrices, built by the plug-in's
        // integration() once per cell) applied by MatrixVectorHomogen::multiplication's
        // StoreLocalStiffnessMatrix branch (MatrixVectorHomogen.h:753-756,170-245), and the
        // same plug-in applied on the fly (applyStencilOnCell, Stencil.h:187-246).
        QuadStencil qs(grid, cs);
        LocalStiffnessMatricesMemoryDistribution lsm(grid, qs, 1);
        y = 0.0; m.multiplication(x, y, lsm); writeVec(out, "Ax_stored", y, n);
        y = 0.0; m.multiplication(x, y, qs);  writeVec(out, "Ax_stored_otf", y, n);
        // Preconditioning::precondLocalStiffnessMatrix (cg_method.h:544-641) for this
        // process: D_p = psi_p^T A_T psi_p, A_T applied by the stored matrices of depth
        // T = depth(p) on the prewavelet stencil of p.  (The reference's copy compares an
        // uninitialised world_rank with the owning node when MPI is off; here the single
        // process owns every depth.  The work vectors are cleared at hanging nodes too, so
        // every D_p is the definition, without values left over from earlier points.)
        VectorSparseG z(grid), v(grid), vout(grid);
        z = 0.0;
        grid.WorkOnHangingNodes = true;
        v = 0.0; vout = 0.0;
        for (size_t i = 0; i < n; ++i) {
            if (!act[i]) continue;
            IndexDimension I = grid.getIndexOfTable(i);
            Depth T(I);
            std::vector<IndexDimension> Js;
            std::vector<double> cs5;
            for (MultiDimFiveCompass mc; mc.goon(); ++mc) {
                double bc = 1.0;
                bool next = true;
                IndexDimension J = I.nextFive(&mc, T, &bc, &next);
                if (next) { v.setValue(J, bc); Js.push_back(J); cs5.push_back(bc); }
            }
            lsm.applyLocalStiffnessMatricesOnIndices_onNode(v, vout, T, Js);
            double coeff = 0.0;
            for (size_t j = 0; j < Js.size(); ++j) coeff += vout.getValue(Js[j]) * cs5[j];
            v = 0.0;
            vout = 0.0;
            z.setValue((unsigned long)i, coeff);
        }
        grid.WorkOnHangingNodes = false;
        writeVec(out, "diag_stored", z, n);
    }
    if (want(what, "cg")) {
        // RHS in the reference driver's convention f_h = M * prew(f_I)
        // (poisson_regular.cpp:57-113) for f = d*pi^2*prod sin(pi x_i).
        VectorSparseG fvec(grid), ftest(grid), prew(grid), u(grid);
        fvec = 0.0;
        for (size_t i = 0; i < n; ++i) {
            if (!act[i]) continue;
            IndexDimension I = grid.getIndexOfTable(i);
            fvec.setValue((unsigned long)i, double(d) * M_PI * M_PI * sinFunc(I));
        }
        prew = 0.0;
        calcPrewByNodal(prew, fvec);
        writeVec(out, "f_prew", prew, n);
        ftest = 0.0;
        m.multiplication(prew, ftest, helm);
        writeVec(out, "f_rhs", ftest, n);
        int it = 0;
        double tp = 0.0;
        VectorSparseG sol(grid);
        sol = 0.0;
        CG::solveHomogen(1e-10, sol, ftest, &it, u, m, poisson, &tp);
        writeVec(out, "cg_poisson_x", sol, n);
        int32_t its = it;
        writeNpy(out + "/cg_poisson_iters.npy", "<i4", &its, 4, {1});
        VectorSparseG sol2(grid);
        sol2 = 0.0;
        it = 0;
        CG::solveHomogen(1e-10, sol2, ftest, &it, u, m, var, &tp);
        writeVec(out, "cg_var_x", sol2, n);
        its = it;
        writeNpy(out + "/cg_var_iters.npy", "<i4", &its, 4, {1});
    }
    if (want(what, "transform")) {
        // Basis transformations (BasisTransformations.cpp) and the norms built on them
        // (applications/H1_Norm.cpp:8-78) for nodal values u = prod sin(pi x_i) + x_1
        // at the active points.
        VectorSparseG u(grid), pr(grid), nod(grid), hier(grid), xin(grid);
        u = 0.0;
        for (size_t i = 0; i < n; ++i) {
            if (!act[i]) continue;
            IndexDimension I = grid.getIndexOfTable(i);
            u.setValue((unsigned long)i, sinFunc(I) + I.coordinate(0));
        }
        writeVec(out, "tr_u", u, n);
        hier = u;
        CalcHierarchicalBasis(hier);
        writeVec(out, "tr_hier", hier, n);
        pr = 0.0;
        calcPrewByNodal(pr, u);
        writeVec(out, "tr_prew", pr, n);
        xin = x;
        nod = 0.0;
        calcNodalByPrew(xin, nod);
        writeVec(out, "tr_nodal_of_x", nod, n);
        H1 h1(grid, mgrid);
        L2 l2(grid, mgrid);
        double nrm[2] = {h1.getValue(u), l2.getValue(u)};
        writeNpy(out + "/tr_norms.npy", "<f8", nrm, 8, {2});
    }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) { std::fprintf(stderr, "usage: see header\n"); return 2; }
    std::string mode = argv[1];
    if (mode == "regular" && argc >= 5) {
        int level = std::atoi(argv[2]);
        CoefSpec cs = parseCoef(argv[3]);
        std::string out = argv[4];
        std::vector<std::string> what(argv + 5, argv + argc);
        AdaptiveSparseGrid grid;
        IndexDimension center;
        grid.AddRecursiveSonsOfPoint(center, level);
        runAll(grid, cs, out, what);
        return 0;
    }
    if ((mode == "adaptive" || mode == "adaptive2") && argc >= 7) {
        // Config-3 style locally adaptive grid: a regular start grid, then repeated
        // refinement (both sons in every dimension) of active points whose support
        // box is closer than `radius` to the origin (adaptive2: to the origin or to the
        // opposite corner (1, ..., 1), two refined regions that do not touch), while the
        // son's sum of (depth_i - 1) <= maxlevel, then the reference's
        // completeDirichletGrid().
        const bool twoCorners = mode == "adaptive2";
        int start = std::atoi(argv[2]);
        double radius = std::atof(argv[3]);
        int maxlevel = std::atoi(argv[4]);
        CoefSpec cs = parseCoef(argv[5]);
        std::string out = argv[6];
        std::vector<std::string> what(argv + 7, argv + argc);
        AdaptiveSparseGrid grid;
        IndexDimension center;
        grid.AddRecursiveSonsOfPoint(center, start);
        for (bool changed = true; changed;) {
            changed = false;
            unsigned long end = grid.getMaximalOccupiedSecondTable();
            for (unsigned long i = 0; i < end; ++i) {
                if (!grid.getActiveTable()[i]) continue;
                IndexDimension I = grid.getIndexOfTable(i);
                // distance from the origin to the point's support box
                // [x_k - 2^-depth_k, x_k + 2^-depth_k]
                double r2 = 0.0, q2 = 0.0;
                for (int k = 0; k < DimensionSparseGrid; ++k) {
                    double lo = I.coordinate(k) - 1.0 / double(1 << I.getDepth(k));
                    if (lo > 0.0) r2 += lo * lo;
                    double hi = 1.0 - (I.coordinate(k) + 1.0 / double(1 << I.getDepth(k)));
                    if (hi > 0.0) q2 += hi * hi;
                }
                if (twoCorners) r2 = std::min(r2, q2);
                if (std::sqrt(r2) >= radius) continue;
                for (int k = 0; k < DimensionSparseGrid; ++k) {
                    IndexDimension S[2] = {I.leftSon(k), I.rightSon(k)};
                    for (auto& J : S) {
                        int lsum = 0;
                        for (int q = 0; q < DimensionSparseGrid; ++q) lsum += J.getDepth(q) - 1;
                        if (lsum > maxlevel) continue;
                        unsigned long kk;
                        if (!grid.occupied(kk, J)) { grid.AddPoint(J); changed = true; }
                    }
                }
            }
        }
        grid.completeDirichletGrid();
        runAll(grid, cs, out, what);
        return 0;
    }
    if (mode == "points" && argc >= 4) {
        // AddPoint for every point of a text file (d 1D indices per line, in order), then
        // completeDirichletGrid(); writes the grid (and the numerics named after outdir).
        std::ifstream in(argv[2]);
        std::string out = argv[3];
        std::vector<std::string> what(argv + 4, argv + argc);
        AdaptiveSparseGrid grid;
        for (;;) {
            IndexDimension I;
            bool ok = true;
            for (int k = 0; k < DimensionSparseGrid && ok; ++k) {
                long v;
                if (!(in >> v)) ok = false;
                else I.replace(k, (indexInteger)v);
            }
            if (!ok) break;
            grid.AddPoint(I);
        }
        grid.completeDirichletGrid();
        CoefSpec cs = parseCoef("const:1:0");
        runAll(grid, cs, out, what);
        return 0;
    }
    if (mode == "samplemv" && argc >= 5) {
        // CPU-baseline sample of one variable-coefficient matvec: the reference's own
        // multiplication() steps (MatrixVectorHomogen.h:660-749) driven member by
        // member (this file is compiled with -fno-access-control): calcNodal for all
        // depths, then cases in a fixed permuted order until the time budget is used.
        // The full matvec time is estimated by scaling the sampled case time with the
        // per-case cell counts sum_{T in sorted(c)} prod 2^T_i.
        int level = std::atoi(argv[2]);
        CoefSpec cs = parseCoef(argv[3]);
        double budget = std::atof(argv[4]);
        const int d = DimensionSparseGrid;
        auto t0 = std::chrono::steady_clock::now();
        AdaptiveSparseGrid grid;
        IndexDimension center;
        grid.AddRecursiveSonsOfPoint(center, level);
        MultiLevelAdaptiveSparseGrid mgrid(&grid);
        MatrixVectorHomogen m(grid, mgrid, true, false);
        VarCoeffStencil var(grid, cs);
        auto t1 = std::chrono::steady_clock::now();
        VectorSparseG prew(grid), Ax(grid);
        size_t n = grid.getMaximalOccupiedSecondTable();
        for (size_t i = 0; i < n; ++i) prew.setValue((unsigned long)i, 1.0 / (1.0 + double(i % 7)));
        auto secs = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
            return std::chrono::duration<double>(b - a).count();
        };
        auto a0 = std::chrono::steady_clock::now();
        Ax = 0.0;
        m.nodal = 0.0;
        grid.WorkOnHangingNodes = true;
#pragma omp parallel
        {
            MultiDepthHashGrid mdg = *grid.getMultiDepthHashGrid();
            VectorSparseG u_nodal(m.u.getSparseGrid());
            int i = 0;
            for (auto it = m.depthList.begin_all(); it != m.depthList.end_all(); ++it) {
                if (i % omp_get_num_threads() == omp_get_thread_num()) {
                    Depth T = *it;
                    u_nodal = 0.0;
                    m.calcNodal(u_nodal, prew, T);
                    m.nodal.setMultiLevelValuesOMP(u_nodal, T, mdg);
                }
                i++;
            }
        }
        grid.WorkOnHangingNodes = false;
        double tNodal = secs(a0, std::chrono::steady_clock::now());
        const int ncase = 1 << d;
        std::vector<double> w(ncase, 0.0);
        for (int c = 0; c < ncase; ++c) {
            bool R[DimensionSparseGrid];
            for (int i = 0; i < d; ++i) R[i] = (c >> i) & 1;
            m.depthList.sortDepths(R);
            for (Depth T : *m.depthList.getSortierteTiefen()) {
                double cells = 1.0;
                for (int i = 0; i < d; ++i) cells *= double(1 << T.at(i));
                w[c] += cells;
            }
        }
        double tCases = 0.0, wSampled = 0.0, wAll = 0.0;
        for (double v : w) wAll += v;
        std::string done;
        for (int k = 0; k < ncase && tCases < budget; ++k) {
            int c = (k * 11 + 0) % ncase;
            bool R[DimensionSparseGrid];
            for (int i = 0; i < d; ++i) R[i] = (c >> i) & 1;
            auto b0 = std::chrono::steady_clock::now();
            grid.WorkOnHangingNodes = true;
            m.Ax_neu = 0.0;
            m.gM = 0.0;
            m.caseFunction(prew, R, var);
            grid.WorkOnHangingNodes = false;
            Ax = Ax + m.Ax_neu;
            tCases += secs(b0, std::chrono::steady_clock::now());
            wSampled += w[c];
            done += (done.empty() ? "" : ",") + std::to_string(c);
        }
        double est = tNodal + tCases * (wAll / wSampled);
        std::printf("{\"n\": %zu, \"setup_s\": %.3f, \"nodal_s\": %.4f, \"timed_s\": %.3f, \"cases\": \"%s\", "
                    "\"case_weight_fraction\": %.4f, \"matvec_s_estimate\": %.4f}\n",
                    n, secs(t0, t1), tNodal, tNodal + tCases, done.c_str(), wSampled / wAll, est);
        return 0;
    }
    if (mode == "timetr" && argc >= 3) {
        // calcPrewByNodal and calcNodalByPrew (BasisTransformations.cpp:486-545, 670-707)
        // on the regular grid of the given level, timed once each; prints JSON.
        int level = std::atoi(argv[2]);
        AdaptiveSparseGrid grid;
        IndexDimension center;
        grid.AddRecursiveSonsOfPoint(center, level);
        size_t n = grid.getMaximalOccupiedSecondTable();
        VectorSparseG u(grid), pr(grid), nod(grid);
        u = 0.0;
        for (size_t i = 0; i < n; ++i) {
            IndexDimension I = grid.getIndexOfTable(i);
            u.setValue((unsigned long)i, sinFunc(I) + I.coordinate(0));
        }
        pr = 0.0;
        auto t0 = std::chrono::steady_clock::now();
        calcPrewByNodal(pr, u);
        auto t1 = std::chrono::steady_clock::now();
        nod = 0.0;
        calcNodalByPrew(pr, nod);
        auto t2 = std::chrono::steady_clock::now();
        double e = 0.0;
        for (size_t i = 0; i < n; ++i) e = std::max(e, std::fabs(nod.getValue((unsigned long)i) - u.getValue((unsigned long)i)));
        std::printf("{\"n\": %zu, \"prew_by_nodal_s\": %.6f, \"nodal_by_prew_s\": %.6f, \"roundtrip_err\": %.3e}\n", n,
                    std::chrono::duration<double>(t1 - t0).count(), std::chrono::duration<double>(t2 - t1).count(), e);
        return 0;
    }
    if (mode == "timemv" && argc >= 5) {
        int level = std::atoi(argv[2]);
        CoefSpec cs = parseCoef(argv[3]);
        int reps = std::atoi(argv[4]);
        AdaptiveSparseGrid grid;
        IndexDimension center;
        auto t0 = std::chrono::steady_clock::now();
        grid.AddRecursiveSonsOfPoint(center, level);
        MultiLevelAdaptiveSparseGrid mgrid(&grid);
        MatrixVectorHomogen m(grid, mgrid, true, false);
        VarCoeffStencil var(grid, cs);
        auto t1 = std::chrono::steady_clock::now();
        VectorSparseG x(grid), y(grid);
        size_t n = grid.getMaximalOccupiedSecondTable();
        for (size_t i = 0; i < n; ++i) x.setValue((unsigned long)i, 1.0 / (1.0 + double(i % 7)));
        std::vector<double> ts;
        double chk = 0.0;
        for (int r = 0; r < reps; ++r) {
            auto a = std::chrono::steady_clock::now();
            y = 0.0;
            m.multiplication(x, y, var);
            auto b = std::chrono::steady_clock::now();
            ts.push_back(std::chrono::duration<double>(b - a).count());
            chk += y.getValue((unsigned long)0);
        }
        std::sort(ts.begin(), ts.end());
        std::printf("{\"n\": %zu, \"setup_s\": %.6f, \"matvec_s_median\": %.6f, \"reps\": %d, \"check\": %.17g}\n",
                    n, std::chrono::duration<double>(t1 - t0).count(), ts[ts.size() / 2], reps, chk);
        return 0;
    }
    if (mode == "timesolve" && argc >= 4) {
        // CPU timing of the reference driver's whole solve on a regular grid
        // (examples_doxygen/poisson_regular.cpp:39-120): grid and operator set-up, the
        // right-hand side f_h = M prew(f_I) of f = d pi^2 prod sin(pi x_i), one matvec, and
        // CG::solveHomogen to eps (its diagonal preconditioner built inside, cg_method.h:805).
        int level = std::atoi(argv[2]);
        CoefSpec cs = parseCoef(argv[3]);
        double eps = argc >= 5 ? std::atof(argv[4]) : 1e-10;
        const int d = DimensionSparseGrid;
        auto now = [] { return std::chrono::steady_clock::now(); };
        auto secs = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
            return std::chrono::duration<double>(b - a).count();
        };
        auto t0 = now();
        AdaptiveSparseGrid grid;
        IndexDimension center;
        grid.AddRecursiveSonsOfPoint(center, level);
        MultiLevelAdaptiveSparseGrid mgrid(&grid);
        MatrixVectorHomogen m(grid, mgrid, true, false);
        VarCoeffStencil var(grid, cs);
        HelmHoltz helm(grid);
        size_t n = grid.getMaximalOccupiedSecondTable();
        auto t1 = now();
        VectorSparseG fvec(grid), prew(grid), f(grid), x(grid), u(grid);
        fvec = 0.0;
        for (size_t i = 0; i < n; ++i) {
            if (!grid.getActiveTable()[i]) continue;
            IndexDimension I = grid.getIndexOfTable(i);
            fvec.setValue((unsigned long)i, double(d) * M_PI * M_PI * sinFunc(I));
        }
        prew = 0.0;
        calcPrewByNodal(prew, fvec);
        f = 0.0;
        m.multiplication(prew, f, helm);
        auto t2 = now();
        VectorSparseG y(grid);
        y = 0.0;
        m.multiplication(prew, y, var);
        auto t3 = now();
        int it = 0;
        double tp = 0.0;
        x = 0.0;
        CG::solveHomogen(eps, x, f, &it, u, m, var, &tp);
        auto t4 = now();
        std::printf("{\"n\": %zu, \"setup_s\": %.6f, \"rhs_s\": %.6f, \"matvec_s\": %.6f, \"cg_s\": %.6f, "
                    "\"iterations\": %d, \"x0\": %.17g}\n",
                    n, secs(t0, t1), secs(t1, t2), secs(t2, t3), secs(t3, t4), it, x.getValue((unsigned long)0));
        return 0;
    }
    std::fprintf(stderr, "bad arguments\n");
    return 2;
}
