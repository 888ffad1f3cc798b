// ref_plugin.hpp -- the variable-coefficient StencilTemplate plug-in that the reference
// harness (ref_harness.cpp) and the drop-in check (shim_check.cpp) hand to the compiled
// LoAdSG reference.  TEST INFRASTRUCTURE ONLY.  Include after source/sparseEXPDE.h,
// inside no namespace.
//
// The reference has no variable-diffusion plug-in; this one goes through its hook
// StencilTemplate::integration / localValue (stencils/Stencil.h:735-748): the element
// matrix of a cell is a(cell centre) * Poisson + c(cell centre) * HelmHoltz, with the
// reference's own Poisson / HelmHoltz integrals.
#ifndef LSG_REF_PLUGIN_HPP
#define LSG_REF_PLUGIN_HPP

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <string>
#include <vector>

struct CoefSpec {
    int kind = 0;  // 0 const, 1 cfg2, 2 cfg3, 3 expsin, 10 gen (non-separable a and c)
    double a0 = 1.0, c0 = 0.0;
    std::vector<double> xi;
    double a(const double* x) const {
        const int d = DimensionSparseGrid;
        switch (kind) {
            case 0: return a0;
            case 1: { double p = 1.0; for (int i = 0; i < d; ++i) p *= std::cos(M_PI * x[i]); return 1.0 + 0.5 * p; }
            case 2: { double s = 1.0; for (int i = 0; i < d; ++i) s += x[i] * x[i]; return s; }
            case 3: {
                double s = 0.0;
                for (size_t k = 1; k <= xi.size(); ++k)
                    s += xi[k - 1] * std::sin(double(k) * M_PI * x[k % d]) / double(k);
                return std::exp(s);
            }
            case 10: return 1.0 + x[0] * x[1 % d] + std::sin(M_PI * (x[0] + x[1 % d]));
        }
        return 0.0;
    }
    double c(const double* x) const {
        const int d = DimensionSparseGrid;
        switch (kind) {
            case 0: return c0;
            case 1: return 1.0 + x[0];
            case 10: return 1.0 + 0.5 * std::sin(M_PI * x[0] * x[1 % d]);
            default: return 0.0;
        }
    }
};

CoefSpec parseCoef(const std::string& s) {
    CoefSpec c;
    if (s.rfind("const:", 0) == 0) {
        c.kind = 0;
        std::sscanf(s.c_str() + 6, "%lf:%lf", &c.a0, &c.c0);
    } else if (s == "gen") c.kind = 10;
    else if (s == "cfg2") c.kind = 1;
    else if (s == "cfg3") c.kind = 2;
    else if (s.rfind("expsin:", 0) == 0) {
        c.kind = 3;
        std::stringstream ss(s.substr(7));
        std::string tok;
        while (std::getline(ss, tok, ',')) c.xi.push_back(std::stod(tok));
    } else {
        std::fprintf(stderr, "bad coef %s\n", s.c_str());
        std::exit(2);
    }
    return c;
}

// Variable-coefficient plug-in: midpoint value of a and c on each cell times
// the reference's own Poisson / HelmHoltz element integrals.
class VarCoeffStencil : public StencilTemplate {
public:
    VarCoeffStencil(AdaptiveSparseGrid& g, const CoefSpec& cs) : StencilTemplate(g), P(g), H(g), spec(cs) {}
    TypeMatrixVectorMultiplication getTypeMatrixVectorMultiplication() { return StencilOnTheFly; }
    inline void initialize(Depth& T_) {
        StencilTemplate::initialize(T_);
        P.initialize(T_);
        H.initialize(T_);
    }
    inline double integration(CellDimension& cell, CellIndexDirection& dirP, CellIndexDirection& dirQ) {
        double x[DimensionSparseGrid];
        for (int i = 0; i < DimensionSparseGrid; ++i) x[i] = cell.coordinate(i);
        double a = spec.a(x), c = spec.c(x);
        double v = 0.0;
        if (a != 0.0) v += a * P.integration(cell, dirP, dirQ);
        if (c != 0.0) v += c * H.integration(cell, dirP, dirQ);
        return v;
    }
    inline double localValue(const IndexDimension& center, const IndexDimension& cell, const MultiDimCompass& mc) const {
        double x[DimensionSparseGrid];
        for (int i = 0; i < DimensionSparseGrid; ++i) x[i] = cell.coordinate(i) + 0.5 * meshwidth[i];
        double a = spec.a(x), c = spec.c(x);
        double v = 0.0;
        if (a != 0.0) v += a * P.localValue(center, cell, mc);
        if (c != 0.0) v += c * H.localValue(center, cell, mc);
        return v;
    }
    Poisson P;
    HelmHoltz H;
    CoefSpec spec;
};

// Element matrices by quadrature, for the stored-matrix path: integration(cell, dirP, dirQ)
// = int_cell a(x) grad phi_p . grad phi_q + c(x) phi_p phi_q with the tensor 2-point
// Gauss-Legendre rule (nodes m_i +- h_i / (2 sqrt 3), weights h_i / 2).  a and c vary inside
// the cell, so the matrices are not a(centre) Poisson + c(centre) HelmHoltz.
class QuadStencil : public StencilTemplate {
public:
    QuadStencil(AdaptiveSparseGrid& g, const CoefSpec& cs) : StencilTemplate(g), spec(cs) {}
    TypeMatrixVectorMultiplication getTypeMatrixVectorMultiplication() { return StencilOnTheFly; }
    inline double integration(CellDimension& cell, CellIndexDirection& dirP, CellIndexDirection& dirQ) {
        const int d = DimensionSparseGrid;
        double lo[DimensionSparseGrid], h[DimensionSparseGrid];
        bool rp[DimensionSparseGrid], rq[DimensionSparseGrid];
        for (int i = 0; i < d; ++i) {
            lo[i] = cell.getLeftBoundary(i);
            h[i] = cell.getRightBoundary(i) - lo[i];
            rp[i] = dirP.getDirection(i) == Right;
            rq[i] = dirQ.getDirection(i) == Right;
        }
        const double g = 0.5 / std::sqrt(3.0);
        double sum = 0.0;
        for (int m = 0; m < (1 << d); ++m) {
            double x[DimensionSparseGrid], fp[DimensionSparseGrid], fq[DimensionSparseGrid];
            double w = 1.0;
            for (int i = 0; i < d; ++i) {
                const double t = ((m >> i) & 1) ? 0.5 + g : 0.5 - g;   // local coordinate in [0, 1]
                x[i] = lo[i] + t * h[i];
                fp[i] = rp[i] ? t : 1.0 - t;
                fq[i] = rq[i] ? t : 1.0 - t;
                w *= 0.5 * h[i];
            }
            double mass = 1.0;
            for (int i = 0; i < d; ++i) mass *= fp[i] * fq[i];
            double stiff = 0.0;
            for (int i = 0; i < d; ++i) {
                double prod = (rp[i] == rq[i] ? 1.0 : -1.0) / (h[i] * h[i]);
                for (int j = 0; j < d; ++j)
                    if (j != i) prod *= fp[j] * fq[j];
                stiff += prod;
            }
            sum += w * (spec.a(x) * stiff + spec.c(x) * mass);
        }
        return sum;
    }
    CoefSpec spec;
};

#endif
