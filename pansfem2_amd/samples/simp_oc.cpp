// SIMP compliance minimisation of a cantilever with the density filter and
// OC, written against the PANSFEM2 facade (include/pansfem2/pansfem2.hpp):
// the same calls a driver written for the reference makes for this loop
// (mesh, boundary conditions, element matrices, Assembling, CSR, ScalingCG,
// reactions, sensitivities, DensityFilter, OC with a volume closure), with
// the assembly, solve, filter and OC candidates running on the GPU.
//
//   simp_oc nx ny iters [out_prefix] [H]
//
// Grid nx x ny of unit quad4 elements, left edge clamped, unit downward load
// at the bottom-right node, E1 = 1, E0 = 1e-9, nu = 0.3, p = 3, volume
// fraction 0.5, filter radius 1.5, ScalingCG(100000, 1e-10), OC(0.5, 0,
// 1e9, 1e-3, 0.2).  Prints "k f g lambda" per iteration; with out_prefix,
// writes the final design variables and their filtered densities as text,
// and the last iteration's mesh, u, r and densities as
// <out_prefix>_result<k>.vtk in the reference sample's layout.
// With "H": HeavisideFilter with the density samples' beta continuation
// (beta 0.5, doubled at every 40th iteration from k = 0) instead of
// DensityFilter.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <numeric>
#include <string>
#include <vector>

#include "pansfem2/pansfem2.hpp"

using namespace PANSFEM2;

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s nx ny iters [out_prefix]\n", argv[0]);
        return 2;
    }
    const int nx = std::atoi(argv[1]), ny = std::atoi(argv[2]), iters = std::atoi(argv[3]);
    try {
        SquareMesh<double> mesh((double)nx, (double)ny, nx, ny);
        std::vector<Vector<double> > x = mesh.GenerateNodes();
        std::vector<std::vector<int> > elements = mesh.GenerateElements();
        const int ne = (int)elements.size();
        auto ufixed = mesh.GenerateFixedlist({0, 1}, [](Vector<double> p) { return std::fabs(p(0)) < 1.0e-5; });
        auto qfixed = mesh.GenerateFixedlist({1}, [&](Vector<double> p) {
            return std::fabs(p(0) - nx) < 1.0e-5 && std::fabs(p(1)) < 1.0e-5;
        });
        for (auto& q : qfixed) q.second = -1.0;

        // filter neighbourhoods: centroids within R, weight (R - d)/R, in
        // ascending element order (windowed search over the grid)
        const double R = 1.5;
        std::vector<Vector<double> > cg(ne);
        for (int e = 0; e < ne; e++) cg[e] = CenterOfGravity(x, elements[e]);
        std::vector<std::vector<int> > neighbors(ne);
        std::vector<std::vector<double> > w(ne);
        const int win = (int)std::ceil(R) + 1;
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++) {
                const int e = ny * i + j;
                for (int a = std::max(0, i - win); a <= std::min(nx - 1, i + win); a++)
                    for (int b = std::max(0, j - win); b <= std::min(ny - 1, j + win); b++) {
                        const int f = ny * a + b;
                        const double d = (cg[e] - cg[f]).Norm();
                        if (d <= R) {
                            neighbors[e].push_back(f);
                            w[e].push_back((R - d) / R);
                        }
                    }
            }
        const bool heaviside = argc > 5 && std::string(argv[5]) == "H";
        DensityFilter<double> dfilter(ne, neighbors, w);
        HeavisideFilter<double> hfilter(ne, neighbors, w);
        double beta = 0.5;
        auto vars = [&](const std::vector<double>& x) {
            return heaviside ? hfilter.GetFilteredVariables(x) : dfilter.GetFilteredVariables(x);
        };
        auto sens = [&](const std::vector<double>& x, const std::vector<double>& d) {
            return heaviside ? hfilter.GetFilteredSensitivitis(x, d) : dfilter.GetFilteredSensitivitis(x, d);
        };

        std::vector<double> s(ne, 0.5);
        const double E0 = 1.0e-9, E1 = 1.0, nu = 0.3, p = 3.0, vf = 0.5, scale0 = 1.0, scale1 = 1.0;
        OC<double> optimizer(ne, 0.5, 0.0, 1.0e9, 1.0e-3, 0.2, std::vector<double>(ne, 0.0),
                             std::vector<double>(ne, 1.0));
        std::vector<double> rho;
        for (int k = 0; k < iters; k++) {
            if (heaviside && k % 40 == 0) {
                beta *= 2.0;
                hfilter.UpdateBeta(beta);
            }
            rho = vars(s);
            double g = 0.0;
            std::vector<double> dgdrho(ne);
            for (int i = 0; i < ne; i++) {
                g += scale1 * rho[i] / (vf * ne);
                dgdrho[i] = scale1 / (vf * ne);
            }
            g -= 1.0 * scale1;

            // displacement
            std::vector<Vector<double> > u(x.size(), Vector<double>(2));
            std::vector<std::vector<int> > n2g(x.size(), std::vector<int>(2, 0));
            SetDirichlet(u, n2g, ufixed);
            int ndof = Renumbering(n2g);
            LILCSR<double> K(ndof, ndof);
            std::vector<double> F(ndof, 0.0);
            for (int e = 0; e < ne; e++) {
                const double E = E1 * std::pow(rho[e], p) + E0 * (1.0 - std::pow(rho[e], p));
                std::vector<std::vector<std::pair<int, int> > > n2e;
                Matrix<double> Ke;
                PlaneStrainStiffness<double, ShapeFunction4Square, Gauss4Square>(Ke, n2e, elements[e], {0, 1}, x, E, nu,
                                                                                 1.0);
                Assembling(K, F, u, Ke, n2g, n2e, elements[e]);
            }
            Assembling(F, qfixed, n2g);
            CSR<double> Kmod(K);
            std::vector<double> result = ScalingCG(Kmod, F, 100000, 1.0e-10);
            Disassembling(u, result, n2g);

            // reactions with every DOF free, compliance f = u . r
            RemoveBoundaryConditions(n2g);
            ndof = Renumbering(n2g);
            std::vector<double> RF(ndof, 0.0);
            std::vector<Vector<double> > r(x.size(), Vector<double>(2));
            for (int e = 0; e < ne; e++) {
                const double E = E1 * std::pow(rho[e], p) + E0 * (1.0 - std::pow(rho[e], p));
                std::vector<std::vector<std::pair<int, int> > > n2e;
                Matrix<double> Ke;
                PlaneStrainStiffness<double, ShapeFunction4Square, Gauss4Square>(Ke, n2e, elements[e], {0, 1}, x, E, nu,
                                                                                 1.0);
                Vector<double> fe = Ke * ElementVector(u, n2e, elements[e]);
                Assembling(RF, fe, n2g, n2e, elements[e]);
            }
            Disassembling(r, RF, n2g);
            const double f = scale0 * std::inner_product(u.begin(), u.end(), r.begin(), 0.0);

            // sensitivities with the unit-modulus element matrix
            std::vector<double> dfdrho(ne);
            for (int e = 0; e < ne; e++) {
                std::vector<std::vector<std::pair<int, int> > > n2e;
                Matrix<double> Ke;
                PlaneStrainStiffness<double, ShapeFunction4Square, Gauss4Square>(Ke, n2e, elements[e], {0, 1}, x, 1.0,
                                                                                 nu, 1.0);
                Vector<double> ue = ElementVector(u, n2e, elements[e]);
                dfdrho[e] = -scale0 * p * (-E0 + E1) * std::pow(rho[e], p - 1.0) * (ue * (Ke * ue));
            }
            std::vector<double> dfds = sens(s, dfdrho);
            std::vector<double> dgds = sens(s, dgdrho);

            // post process of the last iteration, as the reference sample
            // writes every iteration (sample_optimize_density_oc.cpp:175-182)
            if (argc > 4 && k == iters - 1) {
                std::ofstream fout(std::string(argv[4]) + "_result" + std::to_string(k) + ".vtk");
                MakeHeadderToVTK(fout);
                AddPointsToVTK(x, fout);
                AddElementToVTK(elements, fout);
                AddElementTypes(std::vector<int>(elements.size(), 9), fout);
                AddPointVectors(u, "u", fout, true);
                AddPointVectors(r, "r", fout, false);
                AddElementScalers(rho, "s", fout, true);
            }

            if (optimizer.IsConvergence(f)) break;
            optimizer.UpdateVariables(s, f, dfds, g, dgds, [&](std::vector<double> xk1) {
                double gk = 0.0;
                std::vector<double> rk = vars(xk1);
                for (int i = 0; i < ne; i++) gk += scale1 * rk[i] / (vf * ne);
                return gk - 1.0 * scale1;
            });
            std::printf("%d %.17g %.17g %.17g\n", k, f, g, optimizer.LastLambda());
            std::fflush(stdout);
        }
        rho = vars(s);  // densities of the final design
        if (argc > 4) {
            std::ofstream fs(std::string(argv[4]) + "_s.txt"), fr(std::string(argv[4]) + "_rho.txt");
            fs << std::setprecision(17);
            fr << std::setprecision(17);
            for (double v : s) fs << v << "\n";
            for (double v : rho) fr << v << "\n";
        }
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "error: %s\n", ex.what());
        return 1;
    }
    return 0;
}
