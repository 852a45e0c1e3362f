// Host-only check of the C++ facade's element routines (no device calls):
// prints Ke of PlaneStrainStiffness and SolidLinearIsotropicElastic for the
// element coordinates given on stdin, one value per line in %a (exact) form.
//   input: "q4 E nu t x0 y0 ... x3 y3" or "h8 E nu x0 y0 z0 ... x7 y7 z7"
#include <cstdio>
#include <iostream>
#include <string>

#include "pansfem2/pansfem2.hpp"

using namespace PANSFEM2;

int main() {
    std::string kind;
    while (std::cin >> kind) {
        const int nen = kind == "q4" ? 4 : 8, dim = kind == "q4" ? 2 : 3;
        double E, nu, t = 1.0;
        std::cin >> E >> nu;
        if (kind == "q4") std::cin >> t;
        std::vector<Vector<double> > x(nen, Vector<double>(dim));
        std::vector<int> element(nen);
        for (int a = 0; a < nen; a++) {
            element[a] = a;
            for (int d = 0; d < dim; d++) std::cin >> x[a](d);
        }
        Matrix<double> Ke;
        std::vector<std::vector<std::pair<int, int> > > n2e;
        if (kind == "q4")
            PlaneStrainStiffness<double, ShapeFunction4Square, Gauss4Square>(Ke, n2e, element, {0, 1}, x, E, nu, t);
        else
            SolidLinearIsotropicElastic<double, ShapeFunction8Cubic, Gauss8Cubic>(Ke, n2e, element, {0, 1, 2}, x, E, nu);
        for (int i = 0; i < Ke.ROW(); i++)
            for (int j = 0; j < Ke.COL(); j++) std::printf("%a\n", Ke(i, j));
    }
    return 0;
}
