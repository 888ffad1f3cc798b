// shim_check.cpp -- checks the C++ drop-in (include/loadsg_amd_shim.hpp) against the
// compiled LoAdSG reference inside one program.  TEST INFRASTRUCTURE ONLY.
//
// Both sides see the same reference objects: an AdaptiveSparseGrid built with
// AddRecursiveSonsOfPoint(center, level) (sgrid/sparseGrid.cc:557-566) and
// VectorSparseG vectors in slot order.  The reference side runs
// ::MatrixVectorHomogen::multiplication (MatrixVectorHomogen.h:660-759) and
// ::CG::solveHomogen (applications/cg_method.h:759-886) on the CPU; the drop-in
// side runs lsg_shim::MatrixVectorHomogen / lsg_shim::CG with the same calls on the GPU.
// The same is done for calcPrewByNodal / calcNodalByPrew (BasisTransformations.cpp:486-707)
// and H1 / L2 ::getValue (applications/H1_Norm.cpp:8-78), and for a general coefficient:
// the reference with the plug-in of ref_plugin.hpp (a(centre) Poisson + c(centre)
// HelmHoltz per cell, non-separable a and c), the drop-in with
// lsg_shim::Coefficient::functions of the same a and c.  Stored element matrices:
// LocalStiffnessMatricesMemoryDistribution of the quadrature plug-in (QuadStencil) on the
// reference side, lsg_shim::LocalStiffnessMatrices of the same plug-in on the drop-in side
// (matvec; addStencil(Poisson) against the reference's two products added; CG checked by
// the reference's stored-matrix residual of the drop-in's solution).
//
// Usage: shim_check <level>      prints one JSON line; exit status 0 when
//   matvecs agree to 1e-12 (relative, max norm), CG iteration counts are equal and
//   the CG solutions agree to 1e-10, transforms and norms to 1e-12.
#include "source/sparseEXPDE.h"
#include "loadsg_amd_shim.hpp"
#include "ref_plugin.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace {

double relDiff(VectorSparseG& a, VectorSparseG& b, size_t n) {
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < n; ++i) {
        num = std::fmax(num, std::fabs(a.getDatatableVector()[i] - b.getDatatableVector()[i]));
        den = std::fmax(den, std::fabs(a.getDatatableVector()[i]));
    }
    return den > 0.0 ? num / den : num;
}

}  // namespace

int main(int argc, char** argv) {
    const int level = argc > 1 ? std::atoi(argv[1]) : 5;
    if (!lsg_device_available()) {
        std::printf("{\"error\": \"no GPU\"}\n");
        return 2;
    }
    AdaptiveSparseGrid grid;
    IndexDimension center;
    grid.AddRecursiveSonsOfPoint(center, level);
    MultiLevelAdaptiveSparseGrid mgrid(&grid);
    const size_t n = grid.getMaximalOccupiedSecondTable();

    ::MatrixVectorHomogen ref(grid, mgrid, true, false);
    lsg_shim::MatrixVectorHomogen dev(grid, mgrid, true, false);
    Poisson poisson(grid);
    HelmHoltz helm(grid);

    VectorSparseG x(grid), yRef(grid), yDev(grid);
    for (size_t i = 0; i < n; ++i) x.setValue((unsigned long)i, 1.0 / (1.0 + double(i % 7)) - 0.3);

    ref.multiplication(x, yRef, poisson);
    dev.multiplication(x, yDev, poisson);
    const double mvPoisson = relDiff(yRef, yDev, n);
    ref.multiplication(x, yRef, helm);
    dev.multiplication(x, yDev, helm);
    const double mvHelm = relDiff(yRef, yDev, n);

    // CG on f = M x (HelmHoltz product), reference stopping rule, eps 1e-8
    VectorSparseG f(grid), solRef(grid), solDev(grid), usol(grid);
    ref.multiplication(x, f, helm);
    solRef = 0.0;
    solDev = 0.0;
    int itRef = 0, itDev = 0;
    double tpRef = 0.0, tpDev = 0.0;
    ::CG::solveHomogen(1e-8, solRef, f, &itRef, usol, ref, poisson, &tpRef);
    lsg_shim::CG::solveHomogen(1e-8, solDev, f, &itDev, usol, dev, poisson, &tpDev);
    const double cgRel = relDiff(solRef, solDev, n);

    // basis transformations and norms (BasisTransformations.cpp, H1_Norm.cpp)
    VectorSparseG u(grid), pRef(grid), pDev(grid), nRef(grid), nDev(grid);
    for (size_t i = 0; i < n; ++i) u.setValue((unsigned long)i, std::sin(0.37 * double(i)) + 0.1);
    pRef = 0.0;
    calcPrewByNodal(pRef, u);
    lsg_shim::calcPrewByNodal(dev, pDev, u);
    const double trPrew = relDiff(pRef, pDev, n);
    nRef = 0.0;
    calcNodalByPrew(x, nRef);
    lsg_shim::calcNodalByPrew(dev, x, nDev);
    const double trNodal = relDiff(nRef, nDev, n);
    const double h1Ref = ::H1(grid, mgrid).getValue(u), h1Dev = lsg_shim::H1(grid, mgrid).getValue(u);
    const double l2Ref = ::L2(grid, mgrid).getValue(u), l2Dev = lsg_shim::L2(grid, mgrid).getValue(u);
    const double normRel = std::fmax(std::fabs(h1Ref - h1Dev) / h1Ref, std::fabs(l2Ref - l2Dev) / l2Ref);

    // general (non-separable) coefficient: reference plug-in vs host functions
    CoefSpec gen;
    gen.kind = 10;
    VarCoeffStencil var(grid, gen);
    const lsg_shim::Coefficient fun = lsg_shim::Coefficient::functions(
        "gen", [&](const double* p) { return gen.a(p); }, [&](const double* p) { return gen.c(p); });
    ref.multiplication(x, yRef, var);
    dev.multiplication(x, yDev, fun);
    const double mvGen = relDiff(yRef, yDev, n);
    VectorSparseG solRefG(grid), solDevG(grid);
    solRefG = 0.0;
    solDevG = 0.0;
    int itRefG = 0, itDevG = 0;
    ::CG::solveHomogen(1e-8, solRefG, f, &itRefG, usol, ref, var, &tpRef);
    lsg_shim::CG::solveHomogen(1e-8, solDevG, f, &itDevG, usol, dev, fun, &tpDev);
    const double cgRelG = relDiff(solRefG, solDevG, n);

    // stored element matrices of a plug-in whose coefficient varies inside each cell
    QuadStencil quad(grid, gen);
    LocalStiffnessMatricesMemoryDistribution lsmRef(grid, quad, 1);
    lsg_shim::LocalStiffnessMatrices lsmDev(grid, quad, 1);
    ref.multiplication(x, yRef, lsmRef);
    dev.multiplication(x, yDev, lsmDev);
    const double mvStored = relDiff(yRef, yDev, n);
    VectorSparseG yP(grid);
    ref.multiplication(x, yP, poisson);
    for (size_t i = 0; i < n; ++i) yRef.setValue((unsigned long)i, yRef.getValue((unsigned long)i) + yP.getValue((unsigned long)i));
    lsmDev.addStencil(poisson);
    dev.multiplication(x, yDev, lsmDev);
    const double mvStoredAdd = relDiff(yRef, yDev, n);
    // CG with eps = 1e-24 ||f||^2: the reference's loop runs while r^T D^-1 r > eps
    // (cg_method.h:852), so ||r|| / ||f|| ends near 1e-12 sqrt(max D), and the residual of the
    // solution under the reference's stored-matrix operator stays of that order unless the
    // two operators differ
    VectorSparseG solS(grid), rS(grid);
    solS = 0.0;
    int itS = 0;
    double fn = 0.0;
    for (size_t i = 0; i < n; ++i) fn += f.getValue((unsigned long)i) * f.getValue((unsigned long)i);
    lsg_shim::LocalStiffnessMatrices lsmSolve(grid, quad, 1);
    lsg_shim::CG::solveHomogen(1e-24 * fn, solS, f, &itS, usol, dev, lsmSolve, &tpDev);
    ref.multiplication(solS, rS, lsmRef);
    double rn = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const double r = f.getValue((unsigned long)i) - rS.getValue((unsigned long)i);
        rn += r * r;
    }
    const double cgStoredRes = std::sqrt(rn / fn);

    const bool ok = mvPoisson < 1e-12 && mvHelm < 1e-12 && itRef == itDev && cgRel < 1e-10 && trPrew < 1e-12 &&
                    trNodal < 1e-12 && normRel < 1e-12 && mvGen < 1e-12 && itRefG == itDevG && cgRelG < 1e-10 &&
                    mvStored < 1e-12 && mvStoredAdd < 1e-12 && cgStoredRes < 1e-9;
    std::printf("{\"d\": %d, \"level\": %d, \"n\": %zu, \"matvec_poisson_rel\": %.3e, \"matvec_helm_rel\": %.3e, "
                "\"cg_iters_ref\": %d, \"cg_iters_dev\": %d, \"cg_x_rel\": %.3e, \"prew_by_nodal_rel\": %.3e, "
                "\"nodal_by_prew_rel\": %.3e, \"norms_rel\": %.3e, \"matvec_gen_rel\": %.3e, "
                "\"cg_gen_iters_ref\": %d, \"cg_gen_iters_dev\": %d, \"cg_gen_x_rel\": %.3e, \"matvec_stored_rel\": %.3e, "
                "\"matvec_stored_add_rel\": %.3e, \"cg_stored_iters\": %d, \"cg_stored_ref_residual\": %.3e, \"ok\": %s}\n",
                DimensionSparseGrid, level, n, mvPoisson, mvHelm, itRef, itDev, cgRel, trPrew, trNodal, normRel, mvGen,
                itRefG, itDevG, cgRelG, mvStored, mvStoredAdd, itS, cgStoredRes, ok ? "true" : "false");
    return ok ? 0 : 1;
}
