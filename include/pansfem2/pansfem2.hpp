// C++ drop-in facade of PANSFEM2's topology-optimisation path over the
// MI355X library (libpansfem2_amd.so, C-ABI in pansfem2_amd.h).
//
// A driver written against the reference's headers for this path compiles
// against this one header instead.  The names, template parameters,
// signatures and semantics follow the reference:
//   Vector<T>, Matrix<T>            src/LinearAlgebra/Models/{Vector,Matrix}.h
//   ShapeFunction4Square/8Cubic     src/FEM/Controller/ShapeFunction.h:159-329
//   Gauss4Square, Gauss8Cubic       src/FEM/Controller/GaussIntegration.h:133-253
//   PlaneStrainStiffness            src/FEM/Equation/PlaneStrain.h:19-58
//   SolidLinearIsotropicElastic     src/FEM/Equation/Solid.h:19-64
//   CenterOfGravity, ElementVector  src/FEM/Equation/General.h:70-96
//   SetDirichlet, RemoveBoundaryConditions  src/FEM/Controller/BoundaryCondition.h:19-72
//   Assembling, Disassembling, Renumbering  src/FEM/Controller/Assembling.h:46-186
//   SquareMesh<T>                   src/PrePost/Mesher/SquareMesh.h:19-207
//   LILCSR<T>, CSR<T>, ScalingCG    src/LinearAlgebra/Models/{LILCSR,CSR}.h, Solvers/CG.h:419-453
//   DensityFilter<T>                src/Optimize/Filter/DensityFilter.h:15-71
//   HeavisideFilter<T>              src/Optimize/Filter/HeavisideFilter.h:15-99
//   SensitivityFilter<T>, SensitivityFilter2<T>  src/Optimize/Filter/SensitivityFilter.h:14-101
//   OC<T>                           src/Optimize/Solver/OC.h:15-107
//   MMA<T>                          src/Optimize/Solver/MMA.h:15-509
//   CONLIN<T>                       src/Optimize/Solver/CONLIN.h:18-455
//   MakeHeadderToVTK, AddPointsToVTK, AddElementToVTK, AddElementTypes,
//   AddPointScalers, AddPointVectors, AddElementScalers
//                                   src/PrePost/Export/ExportToVTK.h:18-137
//
// Where the work goes:
//   * element matrices, meshes, DOF maps, boundary conditions and the
//     element-vector helpers are host code (as in the reference), evaluated in
//     the reference's operation order so that every value is bit-identical;
//   * CSR<T>(LILCSR<T>&) builds the matrix on the GPU (closed-form structure
//     on the registered grid, values from the recorded element matrices);
//     ScalingCG runs Jacobi-PCG on the GPU; DensityFilter, the OC bisection
//     candidates and MMA run on the GPU.
// The device path covers structured grids: the mesh must come from
// SquareMesh<T> (quad4) or CubicMesh<T> (hex8, no counterpart in the
// reference), and elements must be assembled once each, in any order.  Other
// meshes raise PANSFEM2::DeviceError instead of running on the CPU.
// LILCSR<T>/CSR<T>/ScalingCG are global (not in namespace PANSFEM2), as in
// the reference.
//
// Errors: library failures throw PANSFEM2::DeviceError (std::runtime_error)
// with the library's message.  ScalingCG that reaches itrmax prints
// "Convergence:faild" and returns the last iterate, as CG.h:451-452 does.
#pragma once

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <initializer_list>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../pansfem2_amd.h"

namespace PANSFEM2 {

// ============================================================ device ======
struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// The process-wide library context (one GPU, one grid at a time).
class Device {
public:
    static Device& Get() {
        static Device dev;
        return dev;
    }
    // Choose the GPU before the first device call (default 0).
    static void Select(int device) { Get().device_ = device; }

    pf_ctx* ctx() {
        if (!ctx_) check(pf_create(&ctx_, device_), "pf_create");
        return ctx_;
    }
    void check(int rc, const char* what) {
        if (rc < 0) {
            std::string msg = std::string(what) + ": ";
            msg += ctx_ ? pf_last_error(ctx_) : "library error";
            throw DeviceError(msg);
        }
    }

    // grid registration (SquareMesh / CubicMesh constructors)
    void grid2d(int nx, int ny, double lx, double ly) {
        if (dim_ == 2 && nx_ == nx && ny_ == ny && lx_ == lx && ly_ == ly) return;
        check(pf_grid2d(ctx(), nx, ny, lx, ly), "pf_grid2d");
        dim_ = 2, nx_ = nx, ny_ = ny, nz_ = 0, lx_ = lx, ly_ = ly, lz_ = 0.0;
        ++generation_;
        fixed_.clear();
        have_fixed_ = false;  // a new grid has no DOF map on the device yet
        filter_owner_ = nullptr;
    }
    void grid3d(int nx, int ny, int nz, double lx, double ly, double lz) {
        if (dim_ == 3 && nx_ == nx && ny_ == ny && nz_ == nz && lx_ == lx && ly_ == ly && lz_ == lz) return;
        check(pf_grid3d(ctx(), nx, ny, nz, lx, ly, lz), "pf_grid3d");
        dim_ = 3, nx_ = nx, ny_ = ny, nz_ = nz, lx_ = lx, ly_ = ly, lz_ = lz;
        ++generation_;
        fixed_.clear();
        have_fixed_ = false;
        filter_owner_ = nullptr;
    }
    int dim() const { return dim_; }
    int nodes() const { return dim_ == 2 ? (nx_ + 1) * (ny_ + 1) : (nx_ + 1) * (ny_ + 1) * (nz_ + 1); }
    int elements() const { return dim_ == 2 ? nx_ * ny_ : nx_ * ny_ * nz_; }
    int nodes_per_element() const { return dim_ == 2 ? 4 : 8; }
    int dofs_per_node() const { return dim_; }
    long generation() const { return generation_; }

    // Grid element id of a node list in the grid's element node order, or -1.
    int element_id(const std::vector<int>& el) const {
        if (dim_ == 2) {
            if (el.size() != 4 || el[0] < 0) return -1;
            const int i = el[0] / (ny_ + 1), j = el[0] % (ny_ + 1);
            if (i >= nx_ || j >= ny_) return -1;
            const int n0 = (ny_ + 1) * i + j, n1 = (ny_ + 1) * (i + 1) + j;
            if (el[1] != n1 || el[2] != n1 + 1 || el[3] != n0 + 1) return -1;
            return ny_ * i + j;
        }
        if (dim_ == 3) {
            if (el.size() != 8 || el[0] < 0) return -1;
            const int P = (ny_ + 1) * (nz_ + 1);
            const int i = el[0] / P, j = (el[0] % P) / (nz_ + 1), k = el[0] % (nz_ + 1);
            if (i >= nx_ || j >= ny_ || k >= nz_) return -1;
            auto n = [&](int a, int b, int c) { return P * (i + a) + (nz_ + 1) * (j + b) + (k + c); };
            // ShapeFunction8Cubic order: bottom face (z = k) CCW, then top face
            const int want[8] = {n(0, 0, 0), n(1, 0, 0), n(1, 1, 0), n(0, 1, 0),
                                 n(0, 0, 1), n(1, 0, 1), n(1, 1, 1), n(0, 1, 1)};
            for (int q = 0; q < 8; q++)
                if (el[q] != want[q]) return -1;
            return ny_ * nz_ * i + nz_ * j + k;
        }
        return -1;
    }

    // Send the fixed pattern of a nodetoglobal table (-1 entries) to the
    // device unless it is the one already there; returns the free-DOF count.
    int dirichlet(const std::vector<std::vector<int> >& n2g) {
        std::vector<int32_t> node, dof;
        for (size_t n = 0; n < n2g.size(); n++)
            for (size_t d = 0; d < n2g[n].size(); d++)
                if (n2g[n][d] == -1) node.push_back((int32_t)n), dof.push_back((int32_t)d);
        std::vector<int32_t> key(node.size() * 2);
        for (size_t q = 0; q < node.size(); q++) key[2 * q] = node[q], key[2 * q + 1] = dof[q];
        if (have_fixed_ && key == fixed_) return nfree_;
        int32_t nfree = 0;
        check(pf_set_dirichlet(ctx(), node.data(), dof.data(), (int)node.size(), &nfree), "pf_set_dirichlet");
        fixed_ = key;
        have_fixed_ = true;
        nfree_ = nfree;
        ++generation_;
        return nfree_;
    }

    // which filter's neighbour lists are on the device
    const void* filter_owner_ = nullptr;
    long filter_generation_ = -1;

    ~Device() {
        if (ctx_) pf_destroy(ctx_);
    }

private:
    Device() = default;
    pf_ctx* ctx_ = nullptr;
    int device_ = 0;
    int dim_ = 0, nx_ = 0, ny_ = 0, nz_ = 0;
    double lx_ = 0, ly_ = 0, lz_ = 0;
    long generation_ = 0;
    std::vector<int32_t> fixed_;
    bool have_fixed_ = false;
    int nfree_ = 0;
};

// ====================================================== Vector / Matrix ==
// Dense small vectors and matrices with the reference's interface and the
// reference's association order in every product (sums from zero in index
// order), so host-side element computations round as the reference's do.
template <class T>
class Matrix;

template <class T>
class Vector {
public:
    Vector() = default;
    explicit Vector(int size) : v_(size, T()) {}
    Vector(const std::initializer_list<T>& vals) : v_(vals) {}
    Vector(const std::vector<T>& vals) : v_(vals) {}
    Vector(const Matrix<T>& m);  // column (or row) matrix to vector

    int SIZE() const { return (int)v_.size(); }
    T& operator()(int i) {
        assert(0 <= i && i < SIZE());
        return v_[i];
    }
    const T& operator()(int i) const { return v_[i]; }

    Vector& operator+=(const Vector& o) {
        assert(SIZE() == o.SIZE());
        for (int i = 0; i < SIZE(); i++) v_[i] += o.v_[i];
        return *this;
    }
    Vector& operator-=(const Vector& o) {
        assert(SIZE() == o.SIZE());
        for (int i = 0; i < SIZE(); i++) v_[i] -= o.v_[i];
        return *this;
    }
    Vector& operator*=(T a) {
        for (auto& x : v_) x *= a;
        return *this;
    }
    Vector& operator/=(T a) {
        for (auto& x : v_) x /= a;
        return *this;
    }
    Vector operator+(const Vector& o) const { return Vector(*this) += o; }
    Vector operator-(const Vector& o) const { return Vector(*this) -= o; }
    Vector operator-() const {
        Vector r(*this);
        for (auto& x : r.v_) x = -x;
        return r;
    }
    T operator*(const Vector& o) const {  // inner product
        assert(SIZE() == o.SIZE());
        T s = T();
        for (int i = 0; i < SIZE(); i++) s += v_[i] * o.v_[i];
        return s;
    }
    Matrix<T> operator*(const Matrix<T>& m) const;  // column vector times a 1-row matrix
    Vector operator*(T a) const { return Vector(*this) *= a; }
    Vector operator/(T a) const { return Vector(*this) /= a; }
    friend Vector operator*(T a, const Vector& x) {
        Vector r(x);
        for (auto& y : r.v_) y = a * y;
        return r;
    }

    T Norm() const {
        T s = T();
        for (const auto& x : v_) s += x * x;
        return std::sqrt(s);
    }
    Matrix<T> Transpose() const;
    Vector Vstack(const Vector& o) const {
        Vector r(*this);
        r.v_.insert(r.v_.end(), o.v_.begin(), o.v_.end());
        return r;
    }
    Vector Segment(int head, int tail) const {
        assert(0 <= head && head <= tail && tail <= SIZE());
        return Vector(std::vector<T>(v_.begin() + head, v_.begin() + tail));
    }
    Vector Normal() const { return *this / Norm(); }
    friend std::ostream& operator<<(std::ostream& out, const Vector& x) {  // one value per line (Vector.h:288-293)
        for (int i = 0; i < x.SIZE(); i++) out << x.v_[i] << std::endl;
        return out;
    }
    const std::vector<T>& data() const { return v_; }

private:
    std::vector<T> v_;
    friend class Matrix<T>;
};

template <class T>
Vector<T> VectorProduct(const Vector<T>& a, const Vector<T>& b) {
    assert(a.SIZE() == 3 && b.SIZE() == 3);
    return Vector<T>({a(1) * b(2) - a(2) * b(1), a(2) * b(0) - a(0) * b(2), a(0) * b(1) - a(1) * b(0)});
}

template <class T>
class Matrix {
public:
    Matrix() = default;
    Matrix(int row, int col) : r_(row), c_(col), v_((size_t)row * col, T()) {}
    Matrix(const Vector<T>& x) : r_(x.SIZE()), c_(1), v_(x.v_) {}

    int ROW() const { return r_; }
    int COL() const { return c_; }
    T& operator()(int i, int j) {
        assert(0 <= i && i < r_ && 0 <= j && j < c_);
        return v_[(size_t)i * c_ + j];
    }
    const T& operator()(int i, int j) const { return v_[(size_t)i * c_ + j]; }

    Matrix& operator+=(const Matrix& o) {
        assert(r_ == o.r_ && c_ == o.c_);
        for (size_t k = 0; k < v_.size(); k++) v_[k] += o.v_[k];
        return *this;
    }
    Matrix& operator-=(const Matrix& o) {
        assert(r_ == o.r_ && c_ == o.c_);
        for (size_t k = 0; k < v_.size(); k++) v_[k] -= o.v_[k];
        return *this;
    }
    Matrix& operator*=(T a) {
        for (auto& x : v_) x *= a;
        return *this;
    }
    Matrix& operator/=(T a) {
        for (auto& x : v_) x /= a;
        return *this;
    }
    Matrix operator+(const Matrix& o) const { return Matrix(*this) += o; }
    Matrix operator-(const Matrix& o) const { return Matrix(*this) -= o; }
    Matrix operator-() const {
        Matrix m(*this);
        for (auto& x : m.v_) x = -x;
        return m;
    }
    Matrix operator*(const Matrix& o) const {
        assert(c_ == o.r_);
        Matrix m(r_, o.c_);
        for (int i = 0; i < r_; i++)
            for (int j = 0; j < o.c_; j++) {
                T s = T();
                for (int k = 0; k < c_; k++) s += (*this)(i, k) * o(k, j);
                m(i, j) = s;
            }
        return m;
    }
    Vector<T> operator*(const Vector<T>& x) const {
        assert(c_ == x.SIZE());
        Vector<T> y(r_);
        for (int i = 0; i < r_; i++) {
            T s = T();
            for (int k = 0; k < c_; k++) s += (*this)(i, k) * x(k);
            y(i) = s;
        }
        return y;
    }
    Matrix operator*(T a) const { return Matrix(*this) *= a; }
    Matrix operator/(T a) const { return Matrix(*this) /= a; }
    friend Matrix operator*(T a, const Matrix& m) {
        Matrix r(m);
        for (auto& x : r.v_) x = a * x;
        return r;
    }

    Matrix Transpose() const {
        Matrix t(c_, r_);
        for (int i = 0; i < r_; i++)
            for (int j = 0; j < c_; j++) t(j, i) = (*this)(i, j);
        return t;
    }
    // 2x2 and 3x3 closed forms in the reference's term order (Matrix.h:326-334);
    // larger matrices by cofactor expansion along row 0
    T Determinant() const {
        assert(r_ == c_);
        const auto& v = v_;
        if (r_ == 1) return v[0];
        if (r_ == 2) return v[0] * v[3] - v[1] * v[2];
        if (r_ == 3)
            return -v[8] * v[1] * v[3] - v[7] * v[5] * v[0] - v[2] * v[4] * v[6] + v[6] * v[1] * v[5] +
                   v[7] * v[3] * v[2] + v[0] * v[4] * v[8];
        T s = T();
        for (int j = 0; j < c_; j++) {
            T c = (*this)(0, j) * Cofactor(0, j).Determinant();
            s += (j % 2 == 0) ? c : -c;
        }
        return s;
    }
    // the matrix without row i and column j
    Matrix Cofactor(int i, int j) const {
        Matrix m(r_ - 1, c_ - 1);
        for (int a = 0, p = 0; a < r_; a++) {
            if (a == i) continue;
            for (int b = 0, q = 0; b < c_; b++) {
                if (b == j) continue;
                m(p, q++) = (*this)(a, b);
            }
            p++;
        }
        return m;
    }
    // adjugate over the determinant (Matrix.h:346-359): entry (i, j) is
    // (-1)^(i+j) det(Cofactor(j, i)), then every entry divided by det
    Matrix Inverse() const {
        assert(r_ == c_);
        const T det = Determinant();
        Matrix inv(r_, c_);
        if (r_ == 2) {
            inv(0, 0) = v_[3];
            inv(0, 1) = -v_[1];
            inv(1, 0) = -v_[2];
            inv(1, 1) = v_[0];
        } else {
            for (int i = 0; i < r_; i++)
                for (int j = 0; j < c_; j++) {
                    T c = Cofactor(j, i).Determinant();
                    inv(i, j) = ((i + j) % 2 == 0) ? c : -c;
                }
        }
        inv /= det;
        return inv;
    }
    Matrix Vstack(const Matrix& o) const {
        assert(c_ == o.c_ || r_ == 0);
        Matrix m(r_ + o.r_, o.c_);
        std::copy(v_.begin(), v_.end(), m.v_.begin());
        std::copy(o.v_.begin(), o.v_.end(), m.v_.begin() + v_.size());
        return m;
    }
    Matrix Hstack(const Matrix& o) const {
        assert(r_ == o.r_ || c_ == 0);
        Matrix m(o.r_, c_ + o.c_);
        for (int i = 0; i < o.r_; i++) {
            for (int j = 0; j < c_; j++) m(i, j) = (*this)(i, j);
            for (int j = 0; j < o.c_; j++) m(i, c_ + j) = o(i, j);
        }
        return m;
    }
    Matrix Block(int row, int col, int h, int w) const {
        Matrix m(h, w);
        for (int i = 0; i < h; i++)
            for (int j = 0; j < w; j++) m(i, j) = (*this)(row + i, col + j);
        return m;
    }
    friend std::ostream& operator<<(std::ostream& out, const Matrix& m) {  // rows, tab after each value
        for (int i = 0; i < m.r_; i++) {
            for (int j = 0; j < m.c_; j++) out << m(i, j) << "\t";
            out << std::endl;
        }
        return out;
    }
    const std::vector<T>& data() const { return v_; }

private:
    int r_ = 0, c_ = 0;
    std::vector<T> v_;
    friend class Vector<T>;
};

template <class T>
Vector<T>::Vector(const Matrix<T>& m) : v_(m.v_) {
    assert(m.ROW() == 1 || m.COL() == 1);
}
template <class T>
Matrix<T> Vector<T>::Transpose() const {
    Matrix<T> m(1, SIZE());
    for (int i = 0; i < SIZE(); i++) m(0, i) = v_[i];
    return m;
}
template <class T>
Matrix<T> Vector<T>::operator*(const Matrix<T>& m) const {
    return Matrix<T>(*this) * m;
}
template <class T>
Matrix<T> Identity(int n) {
    Matrix<T> m(n, n);
    for (int i = 0; i < n; i++) m(i, i) = T(1);
    return m;
}
template <class T>
Matrix<T> Diagonal(const Vector<T>& x) {
    Matrix<T> m(x.SIZE(), x.SIZE());
    for (int i = 0; i < x.SIZE(); i++) m(i, i) = x(i);
    return m;
}

// ================================================ shape functions ========
// Bilinear quad (nodes CCW from (-1,-1)) and trilinear hex (bottom face CCW,
// then top face), with their 2x2 and 2x2x2 Gauss rules (weights 1); the
// point order of the rules is the reference's.
template <class T>
struct ShapeFunction4Square {
    static const int d = 2;
    static const int n = 4;
    static const std::vector<Vector<T> > Points;
    static Vector<T> N(Vector<T> r) {
        return Vector<T>({0.25 * (1.0 - r(0)) * (1.0 - r(1)), 0.25 * (1.0 + r(0)) * (1.0 - r(1)),
                          0.25 * (1.0 + r(0)) * (1.0 + r(1)), 0.25 * (1.0 - r(0)) * (1.0 + r(1))});
    }
    static Matrix<T> dNdr(Vector<T> r) {
        Matrix<T> m(d, n);
        m(0, 0) = -0.25 * (1.0 - r(1)); m(0, 1) = 0.25 * (1.0 - r(1));
        m(0, 2) = 0.25 * (1.0 + r(1));  m(0, 3) = -0.25 * (1.0 + r(1));
        m(1, 0) = -0.25 * (1.0 - r(0)); m(1, 1) = -0.25 * (1.0 + r(0));
        m(1, 2) = 0.25 * (1.0 + r(0));  m(1, 3) = 0.25 * (1.0 - r(0));
        return m;
    }
};
template <class T>
const std::vector<Vector<T> > ShapeFunction4Square<T>::Points = {{-1.0, -1.0}, {1.0, -1.0}, {1.0, 1.0}, {-1.0, 1.0}};

template <class T>
struct ShapeFunction8Cubic {
    static const int d = 3;
    static const int n = 8;
    static const std::vector<Vector<T> > Points;
    static Vector<T> N(Vector<T> r) {
        Vector<T> v(n);
        for (int a = 0; a < n; a++) {
            const Vector<T>& P = Points[a];
            v(a) = 0.125 * (1.0 + P(0) * r(0)) * (1.0 + P(1) * r(1)) * (1.0 + P(2) * r(2));
        }
        return v;
    }
    static Matrix<T> dNdr(Vector<T> r) {
        Matrix<T> m(d, n);
        const T xm = 1.0 - r(0), xp = 1.0 + r(0), ym = 1.0 - r(1), yp = 1.0 + r(1), zm = 1.0 - r(2),
                zp = 1.0 + r(2);
        // d/dr0 (factor order: the y term, then the z term)
        m(0, 0) = -0.125 * ym * zm; m(0, 1) = 0.125 * ym * zm; m(0, 2) = 0.125 * yp * zm; m(0, 3) = -0.125 * yp * zm;
        m(0, 4) = -0.125 * ym * zp; m(0, 5) = 0.125 * ym * zp; m(0, 6) = 0.125 * yp * zp; m(0, 7) = -0.125 * yp * zp;
        // d/dr1 (the z term, then the x term)
        m(1, 0) = -0.125 * zm * xm; m(1, 1) = -0.125 * zm * xp; m(1, 2) = 0.125 * zm * xp; m(1, 3) = 0.125 * zm * xm;
        m(1, 4) = -0.125 * zp * xm; m(1, 5) = -0.125 * zp * xp; m(1, 6) = 0.125 * zp * xp; m(1, 7) = 0.125 * zp * xm;
        // d/dr2 (the x term, then the y term)
        m(2, 0) = -0.125 * xm * ym; m(2, 1) = -0.125 * xp * ym; m(2, 2) = -0.125 * xp * yp; m(2, 3) = -0.125 * xm * yp;
        m(2, 4) = 0.125 * xm * ym;  m(2, 5) = 0.125 * xp * ym;  m(2, 6) = 0.125 * xp * yp;  m(2, 7) = 0.125 * xm * yp;
        return m;
    }
};
template <class T>
const std::vector<Vector<T> > ShapeFunction8Cubic<T>::Points = {{-1.0, -1.0, -1.0}, {1.0, -1.0, -1.0}, {1.0, 1.0, -1.0},
                                                                {-1.0, 1.0, -1.0},  {-1.0, -1.0, 1.0}, {1.0, -1.0, 1.0},
                                                                {1.0, 1.0, 1.0},    {-1.0, 1.0, 1.0}};

template <class T>
struct Gauss4Square {
    static const int N = 4;
    static const std::vector<Vector<T> > Points;
    static const std::vector<std::vector<T> > Weights;
};
template <class T>
const std::vector<Vector<T> > Gauss4Square<T>::Points = {{-1.0 / std::sqrt(3.0), -1.0 / std::sqrt(3.0)},
                                                         {1.0 / std::sqrt(3.0), -1.0 / std::sqrt(3.0)},
                                                         {-1.0 / std::sqrt(3.0), 1.0 / std::sqrt(3.0)},
                                                         {1.0 / std::sqrt(3.0), 1.0 / std::sqrt(3.0)}};
template <class T>
const std::vector<std::vector<T> > Gauss4Square<T>::Weights = {{1.0, 1.0}, {1.0, 1.0}, {1.0, 1.0}, {1.0, 1.0}};

template <class T>
struct Gauss8Cubic {
    static const int N = 8;
    static const std::vector<Vector<T> > Points;
    static const std::vector<std::vector<T> > Weights;
};
template <class T>
const std::vector<Vector<T> > Gauss8Cubic<T>::Points = []() {
    const T a = 1.0 / std::sqrt(3.0), m = -1.0 / std::sqrt(3.0);
    return std::vector<Vector<T> >{{m, m, m}, {a, m, m}, {a, a, m}, {m, a, m},
                                   {m, m, a}, {a, m, a}, {a, a, a}, {m, a, a}};
}();
template <class T>
const std::vector<std::vector<T> > Gauss8Cubic<T>::Weights(8, std::vector<T>{1.0, 1.0, 1.0});

// ================================================ element matrices =======
// Ke = sum_g B^T D B J t w0 w1 with D the plane-strain elasticity matrix;
// nodetoelement[i] = {(doulist[0], 2i), (doulist[1], 2i+1)}.
template <class T, template <class> class SF, template <class> class IC>
void PlaneStrainStiffness(Matrix<T>& Ke, std::vector<std::vector<std::pair<int, int> > >& nodetoelement,
                          const std::vector<int>& element, const std::vector<int>& doulist, std::vector<Vector<T> >& x,
                          T E, T V, T t) {
    assert(doulist.size() == 2);
    const int ne = (int)element.size();
    Ke = Matrix<T>(2 * ne, 2 * ne);
    nodetoelement.assign(ne, std::vector<std::pair<int, int> >(2));
    for (int i = 0; i < ne; i++) {
        nodetoelement[i][0] = {doulist[0], 2 * i};
        nodetoelement[i][1] = {doulist[1], 2 * i + 1};
    }
    Matrix<T> X(ne, 2);
    for (int i = 0; i < ne; i++) X(i, 0) = x[element[i]](0), X(i, 1) = x[element[i]](1);
    Matrix<T> D(3, 3);
    D(0, 0) = 1.0 - V; D(0, 1) = V;       D(0, 2) = T();
    D(1, 0) = V;       D(1, 1) = 1.0 - V; D(1, 2) = T();
    D(2, 0) = T();     D(2, 1) = T();     D(2, 2) = 0.5 * (1.0 - 2.0 * V);
    D *= E / ((1.0 - 2.0 * V) * (1.0 + V));
    for (int g = 0; g < IC<T>::N; g++) {
        Matrix<T> dNdr = SF<T>::dNdr(IC<T>::Points[g]);
        Matrix<T> dXdr = dNdr * X;
        T J = dXdr.Determinant();
        Matrix<T> dNdX = dXdr.Inverse() * dNdr;
        Matrix<T> B(3, 2 * ne);
        for (int n = 0; n < ne; n++) {
            B(0, 2 * n) = dNdX(0, n);
            B(1, 2 * n + 1) = dNdX(1, n);
            B(2, 2 * n) = dNdX(1, n);
            B(2, 2 * n + 1) = dNdX(0, n);
        }
        Ke += B.Transpose() * D * B * J * t * IC<T>::Weights[g][0] * IC<T>::Weights[g][1];
    }
}

// Ke = sum_g B^T C B J w0 w1 w2 with C the isotropic elasticity matrix
// (strain order xx, yy, zz, xy, yz, zx).
template <class T, template <class> class SF, template <class> class IC>
void SolidLinearIsotropicElastic(Matrix<T>& Ke, std::vector<std::vector<std::pair<int, int> > >& nodetoelement,
                                 const std::vector<int>& element, const std::vector<int>& doulist,
                                 std::vector<Vector<T> >& x, T E, T V) {
    assert(doulist.size() == 3);
    const int ne = (int)element.size();
    Ke = Matrix<T>(3 * ne, 3 * ne);
    nodetoelement.assign(ne, std::vector<std::pair<int, int> >(3));
    for (int i = 0; i < ne; i++)
        for (int d = 0; d < 3; d++) nodetoelement[i][d] = {doulist[d], 3 * i + d};
    Matrix<T> X(ne, 3);
    for (int i = 0; i < ne; i++)
        for (int d = 0; d < 3; d++) X(i, d) = x[element[i]](d);
    Matrix<T> C(6, 6);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) C(a, b) = (a == b) ? 1.0 - V : V;
    for (int a = 3; a < 6; a++) C(a, a) = 0.5 * (1.0 - 2.0 * V);
    C *= E / ((1.0 + V) * (1.0 - 2.0 * V));
    for (int g = 0; g < IC<T>::N; g++) {
        Matrix<T> dNdr = SF<T>::dNdr(IC<T>::Points[g]);
        Matrix<T> dXdr = dNdr * X;
        T J = dXdr.Determinant();
        Matrix<T> dNdX = dXdr.Inverse() * dNdr;
        Matrix<T> B(6, 3 * ne);
        for (int n = 0; n < ne; n++) {
            const T dx = dNdX(0, n), dy = dNdX(1, n), dz = dNdX(2, n);
            B(0, 3 * n) = dx;
            B(1, 3 * n + 1) = dy;
            B(2, 3 * n + 2) = dz;
            B(3, 3 * n) = dy, B(3, 3 * n + 1) = dx;
            B(4, 3 * n + 1) = dz, B(4, 3 * n + 2) = dy;
            B(5, 3 * n) = dz, B(5, 3 * n + 2) = dx;
        }
        Ke += B.Transpose() * C * B * J * IC<T>::Weights[g][0] * IC<T>::Weights[g][1] * IC<T>::Weights[g][2];
    }
}

template <class T>
Vector<T> CenterOfGravity(std::vector<Vector<T> >& x, std::vector<int>& element) {
    Vector<T> c(x[0].SIZE());
    for (int i : element) c += x[i];
    return c / (T)element.size();
}

template <class T>
Vector<T> ElementVector(std::vector<Vector<T> >& u, const std::vector<std::vector<std::pair<int, int> > >& n2e,
                        const std::vector<int>& element) {
    int size = 0;
    for (const auto& l : n2e) size += (int)l.size();
    Vector<T> v(size);
    for (size_t i = 0; i < n2e.size(); i++)
        for (const auto& dou : n2e[i]) v(dou.second) = u[element[i]](dou.first);
    return v;
}

// ============================================================ meshes =====
// SquareMesh<T>(lx, ly, nx, ny): node (ny+1) i + j at (lx i/nx, ly j/ny),
// element ny i + j with nodes BL, BR, TR, TL.  Constructing one registers
// the grid with the device.
template <class T>
class SquareMesh {
public:
    SquareMesh(T lx, T ly, int nx, int ny) : x_(lx), y_(ly), nx_(nx), ny_(ny) {
        Device::Get().grid2d(nx, ny, (double)lx, (double)ly);
    }
    std::vector<Vector<T> > GenerateNodes() const {
        std::vector<Vector<T> > nodes;
        nodes.reserve((size_t)(nx_ + 1) * (ny_ + 1));
        for (int i = 0; i < nx_ + 1; i++)
            for (int j = 0; j < ny_ + 1; j++) nodes.push_back(Vector<T>({x_ * (i / (T)nx_), y_ * (j / (T)ny_)}));
        return nodes;
    }
    std::vector<std::vector<int> > GenerateElements() const {
        std::vector<std::vector<int> > el;
        el.reserve((size_t)nx_ * ny_);
        for (int i = 0; i < nx_; i++)
            for (int j = 0; j < ny_; j++) {
                const int n0 = (ny_ + 1) * i + j, n1 = (ny_ + 1) * (i + 1) + j;
                el.push_back({n0, n1, n1 + 1, n0 + 1});
            }
        return el;
    }
    template <class F>
    std::vector<std::pair<std::pair<int, int>, T> > GenerateFixedlist(std::vector<int> ulist, F iscorrespond) const {
        assert(!ulist.empty() && 0 <= *std::min_element(ulist.begin(), ulist.end()));
        std::vector<std::pair<std::pair<int, int>, T> > fixed;
        for (int i = 0; i < nx_ + 1; i++)
            for (int j = 0; j < ny_ + 1; j++)
                if (iscorrespond(Vector<T>({x_ * (i / (T)nx_), y_ * (j / (T)ny_)})))
                    for (int d : ulist) fixed.push_back({{(ny_ + 1) * i + j, d}, T()});
        return fixed;
    }

private:
    T x_, y_;
    int nx_, ny_;
};

// CubicMesh<T>(lx, ly, lz, nx, ny, nz): structured hex8 grid (no
// counterpart in the reference, which has no 3D mesher).  Node
// ((ny+1)(nz+1)) i + (nz+1) j + k at (lx i/nx, ly j/ny, lz k/nz); element
// ny nz i + nz j + k with ShapeFunction8Cubic's node order.
template <class T>
class CubicMesh {
public:
    CubicMesh(T lx, T ly, T lz, int nx, int ny, int nz) : x_(lx), y_(ly), z_(lz), nx_(nx), ny_(ny), nz_(nz) {
        Device::Get().grid3d(nx, ny, nz, (double)lx, (double)ly, (double)lz);
    }
    Vector<T> Node(int i, int j, int k) const {
        return Vector<T>({x_ * (i / (T)nx_), y_ * (j / (T)ny_), z_ * (k / (T)nz_)});
    }
    std::vector<Vector<T> > GenerateNodes() const {
        std::vector<Vector<T> > nodes;
        nodes.reserve((size_t)(nx_ + 1) * (ny_ + 1) * (nz_ + 1));
        for (int i = 0; i < nx_ + 1; i++)
            for (int j = 0; j < ny_ + 1; j++)
                for (int k = 0; k < nz_ + 1; k++) nodes.push_back(Node(i, j, k));
        return nodes;
    }
    std::vector<std::vector<int> > GenerateElements() const {
        std::vector<std::vector<int> > el;
        el.reserve((size_t)nx_ * ny_ * nz_);
        const int P = (ny_ + 1) * (nz_ + 1);
        for (int i = 0; i < nx_; i++)
            for (int j = 0; j < ny_; j++)
                for (int k = 0; k < nz_; k++) {
                    auto n = [&](int a, int b, int c) { return P * (i + a) + (nz_ + 1) * (j + b) + (k + c); };
                    el.push_back({n(0, 0, 0), n(1, 0, 0), n(1, 1, 0), n(0, 1, 0), n(0, 0, 1), n(1, 0, 1), n(1, 1, 1),
                                  n(0, 1, 1)});
                }
        return el;
    }
    template <class F>
    std::vector<std::pair<std::pair<int, int>, T> > GenerateFixedlist(std::vector<int> ulist, F iscorrespond) const {
        std::vector<std::pair<std::pair<int, int>, T> > fixed;
        const int P = (ny_ + 1) * (nz_ + 1);
        for (int i = 0; i < nx_ + 1; i++)
            for (int j = 0; j < ny_ + 1; j++)
                for (int k = 0; k < nz_ + 1; k++)
                    if (iscorrespond(Node(i, j, k)))
                        for (int d : ulist) fixed.push_back({{P * i + (nz_ + 1) * j + k, d}, T()});
        return fixed;
    }

private:
    T x_, y_, z_;
    int nx_, ny_, nz_;
};

// ============================================= boundary conditions ======
template <class T>
void SetDirichlet(std::vector<Vector<T> >& u, std::vector<std::vector<int> >& n2g,
                  const std::vector<std::pair<std::pair<int, int>, T> >& ufixed) {
    for (const auto& f : ufixed) {
        u[f.first.first](f.first.second) = f.second;
        n2g[f.first.first][f.first.second] = -1;
    }
}
template <class T>
void SetDirichlet(std::vector<std::vector<int> >& n2g, const std::vector<std::pair<std::pair<int, int>, T> >& ufixed) {
    for (const auto& f : ufixed) n2g[f.first.first][f.first.second] = -1;
}
inline void RemoveBoundaryConditions(std::vector<std::vector<int> >& n2g) {
    for (auto& node : n2g)
        for (auto& d : node) d = 0;
}
inline int Renumbering(std::vector<std::vector<int> >& n2g) {
    int count = 0;
    for (auto& node : n2g)
        for (auto& d : node)
            if (d != -1) d = count++;
    return count;
}

}  // namespace PANSFEM2

// ================================================== sparse matrices ======
// LILCSR<T> records element matrices per grid element (Assembling); CSR<T>
// (LILCSR<T>&) assembles them on the GPU and is a handle to the device
// matrix: operator* and ScalingCG run there.
template <class T>
class CSR;

template <class T>
class LILCSR {
public:
    LILCSR(int rows, int cols) : ROWS(rows), COLS(cols) {}
    const int ROWS;
    const int COLS;

    // element matrix of grid element e (called by Assembling)
    void record(int e, const PANSFEM2::Matrix<T>& Ke, const std::vector<std::vector<int> >& n2g) {
        auto& dev = PANSFEM2::Device::Get();
        const int ne = dev.elements();
        const int nd = dev.nodes_per_element() * dev.dofs_per_node();
        if (ke_.empty()) {
            ke_.assign((size_t)ne * nd * nd, T());
            seen_.assign(ne, 0);
            n2g_ = n2g;
            generation_ = dev.generation();
        }
        if (Ke.ROW() != nd || Ke.COL() != nd)
            throw PANSFEM2::DeviceError("Assembling: element matrix size does not match the grid's elements");
        if (seen_[e])
            throw PANSFEM2::DeviceError("Assembling: element assembled twice (one call per element is supported)");
        seen_[e] = 1;
        std::copy(Ke.data().begin(), Ke.data().end(), ke_.begin() + (size_t)e * nd * nd);
    }

private:
    std::vector<T> ke_;
    std::vector<char> seen_;
    std::vector<std::vector<int> > n2g_;
    long generation_ = -1;
    friend class CSR<T>;
};

template <class T>
class CSR {
public:
    explicit CSR(LILCSR<T>& K) : ROWS(K.ROWS), COLS(K.COLS) {
        static_assert(std::is_same<T, double>::value, "the device path computes in double");
        auto& dev = PANSFEM2::Device::Get();
        if (K.ke_.empty()) throw PANSFEM2::DeviceError("CSR(LILCSR&): nothing was assembled");
        for (char c : K.seen_)
            if (!c) throw PANSFEM2::DeviceError("CSR(LILCSR&): every grid element must be assembled once");
        const int nfree = dev.dirichlet(K.n2g_);
        if (nfree != ROWS || ROWS != COLS)
            throw PANSFEM2::DeviceError("CSR(LILCSR&): matrix size does not match the DOF map of the assembly");
        int64_t nnz = 0;
        dev.check(pf_assemble_csr_elements(dev.ctx(), K.ke_.data(), &nnz), "pf_assemble_csr_elements");
        nnz_ = nnz;
        generation_ = dev.generation();
        current() = this;
    }
    ~CSR() {
        if (current() == this) current() = nullptr;
    }
    CSR(const CSR&) = delete;
    CSR& operator=(const CSR&) = delete;

    const int ROWS;
    const int COLS;

    const std::vector<T> operator*(const std::vector<T>& x) {
        require();
        std::vector<T> y(ROWS);
        auto& dev = PANSFEM2::Device::Get();
        dev.check(pf_apply_K(dev.ctx(), PF_OP_CSR64, x.data(), y.data()), "pf_apply_K");
        return y;
    }
    T get(int row, int col) const {
        const_cast<CSR*>(this)->download();
        for (int k = indptr_[row]; k < indptr_[row + 1]; k++)
            if (indices_[k] == col) return data_[k];
        return T();
    }
    int64_t nnz() const { return nnz_; }
    // device matrix this handle refers to is still the context's current one
    void require() const {
        if (current() != this || PANSFEM2::Device::Get().generation() != generation_)
            throw PANSFEM2::DeviceError("CSR: the device holds a newer matrix; rebuild this one");
    }

private:
    static CSR*& current() {
        static CSR* c = nullptr;
        return c;
    }
    void download() {
        if (!indptr_.empty()) return;
        require();
        auto& dev = PANSFEM2::Device::Get();
        indptr_.resize(ROWS + 1);
        indices_.resize(std::max<int64_t>(nnz_, 1));
        data_.resize(std::max<int64_t>(nnz_, 1));
        dev.check(pf_get_csr(dev.ctx(), indptr_.data(), indices_.data(), data_.data()), "pf_get_csr");
    }
    int64_t nnz_ = 0;
    long generation_ = -1;
    std::vector<int32_t> indptr_, indices_;
    std::vector<T> data_;
};

// Jacobi-preconditioned CG from x = 0 to ||r|| < eps ||b|| on the device.
template <class T>
std::vector<T> ScalingCG(CSR<T>& A, const std::vector<T>& b, int itrmax, T eps) {
    A.require();
    auto& dev = PANSFEM2::Device::Get();
    std::vector<T> x(A.ROWS, T());
    pf_solve_stats st{};
    const int rc = pf_pcg(dev.ctx(), PF_OP_CSR64, b.data(), x.data(), eps, itrmax, &st);
    dev.check(rc, "pf_pcg");
    if (rc == PF_NOT_CONVERGED) std::cout << "Convergence:faild" << std::endl;
    return x;
}

namespace PANSFEM2 {

// ======================================================== assembling =====
// K(r, c) += Ke(a, b) for free-free pairs (recorded for the device
// assembly); F[r] -= Ke(a, b) u_c for fixed columns, on the host.
template <class T>
void Assembling(LILCSR<T>& K, std::vector<T>& F, std::vector<Vector<T> >& u, Matrix<T>& Ke,
                const std::vector<std::vector<int> >& n2g, const std::vector<std::vector<std::pair<int, int> > >& n2e,
                const std::vector<int>& element) {
    auto& dev = Device::Get();
    const int e = dev.element_id(element);
    if (e < 0) throw DeviceError("Assembling: element is not an element of the registered grid");
    const int ndpn = dev.dofs_per_node();
    for (size_t i = 0; i < n2e.size(); i++) {
        if ((int)n2e[i].size() != ndpn) throw DeviceError("Assembling: unexpected nodetoelement layout");
        for (int d = 0; d < ndpn; d++)
            if (n2e[i][d].first != d || n2e[i][d].second != ndpn * (int)i + d)
                throw DeviceError("Assembling: unexpected nodetoelement layout");
    }
    K.record(e, Ke, n2g);
    for (size_t i = 0; i < element.size(); i++)
        for (const auto& di : n2e[i]) {
            const int r = n2g[element[i]][di.first];
            if (r == -1) continue;
            for (size_t j = 0; j < element.size(); j++)
                for (const auto& dj : n2e[j])
                    if (n2g[element[j]][dj.first] == -1) F[r] -= Ke(di.second, dj.second) * u[element[j]](dj.first);
        }
}

template <class T>
void Assembling(std::vector<T>& F, Vector<T>& Fe, const std::vector<std::vector<int> >& n2g,
                const std::vector<std::vector<std::pair<int, int> > >& n2e, const std::vector<int>& element) {
    for (size_t i = 0; i < element.size(); i++)
        for (const auto& di : n2e[i]) {
            const int r = n2g[element[i]][di.first];
            if (r != -1) F[r] += Fe(di.second);
        }
}

template <class T>
void Assembling(std::vector<T>& F, const std::vector<std::pair<std::pair<int, int>, T> >& f,
                const std::vector<std::vector<int> >& n2g) {
    for (const auto& q : f) {
        const int r = n2g[q.first.first][q.first.second];
        if (r != -1) F[r] += q.second;
    }
}

template <class T>
void Disassembling(std::vector<Vector<T> >& u, const std::vector<T>& result, const std::vector<std::vector<int> >& n2g) {
    for (size_t n = 0; n < n2g.size(); n++)
        for (size_t d = 0; d < n2g[n].size(); d++)
            if (n2g[n][d] != -1) u[n](d) = result[n2g[n][d]];
}

// =========================================================== filter ======
namespace detail {
// Neighbour lists of a filter, uploaded to the device when a filter call
// finds another filter's lists (or another grid) there.
template <class T>
class FilterLists {
public:
    FilterLists() = default;
    FilterLists(int n, const std::vector<std::vector<int> >& neighbors, const std::vector<std::vector<T> >& w) : n_(n) {
        static_assert(std::is_same<T, double>::value, "the device path computes in double");
        if ((int)neighbors.size() != n || (int)w.size() != n) throw DeviceError("filter: list sizes differ from n");
        ptr_.assign(1, 0);
        for (int i = 0; i < n; i++) {
            if (neighbors[i].size() != w[i].size()) throw DeviceError("filter: neighbour and weight lists differ");
            idx_.insert(idx_.end(), neighbors[i].begin(), neighbors[i].end());
            w_.insert(w_.end(), w[i].begin(), w[i].end());
            ptr_.push_back((int32_t)idx_.size());
        }
    }
    void ensure() {
        auto& dev = Device::Get();
        if (n_ != dev.elements()) throw DeviceError("filter: n differs from the registered grid's elements");
        if (dev.filter_owner_ != this || dev.filter_generation_ != dev.generation()) {
            dev.check(pf_filter_set_lists(dev.ctx(), ptr_.data(), idx_.data(), w_.data()), "pf_filter_set_lists");
            dev.filter_owner_ = this;
            dev.filter_generation_ = dev.generation();
        }
    }
    int n_ = 0;

private:
    std::vector<int32_t> ptr_, idx_;
    std::vector<T> w_;
};
}  // namespace detail

// rho_i = sum_j w_ij s_j / sum_j w_ij over the given neighbour lists, on the
// device; GetFilteredSensitivitis is the same row-normalised sum.
template <class T>
class DensityFilter {
public:
    DensityFilter(int n, std::vector<std::vector<int> > neighbors, std::vector<std::vector<T> > w)
        : lists_(n, neighbors, w) {}
    std::vector<T> GetFilteredVariables(std::vector<T> s) { return run(PF_FILTER_VARIABLES, s); }
    std::vector<T> GetFilteredSensitivitis(std::vector<T> /*s*/, std::vector<T> dfdrho) {
        return run(PF_FILTER_SENSITIVITIES, dfdrho);
    }

private:
    std::vector<T> run(int mode, const std::vector<T>& in) {
        lists_.ensure();
        auto& dev = Device::Get();
        std::vector<T> out(lists_.n_);
        dev.check(pf_filter(dev.ctx(), mode, in.data(), out.data()), "pf_filter");
        return out;
    }
    detail::FilterLists<T> lists_;
};

// Sensitivity filters on the device over the given lists:
// SensitivityFilter (Sigmund's scheme) divides the weighted sum of s_j df_j
// by s_i times the weight sum, SensitivityFilter2 by the s-weighted sum.
template <class T, int Scheme>
class SensitivityFilterOf {
public:
    SensitivityFilterOf(int n, std::vector<std::vector<int> > neighbors, std::vector<std::vector<T> > w)
        : lists_(n, neighbors, w) {}
    std::vector<T> GetFilteredSensitivitis(std::vector<T> s, std::vector<T> dfdrho) {
        lists_.ensure();
        auto& dev = Device::Get();
        std::vector<T> out(lists_.n_);
        dev.check(pf_sensitivity_filter(dev.ctx(), Scheme, s.data(), dfdrho.data(), out.data()),
                  "pf_sensitivity_filter");
        return out;
    }

private:
    detail::FilterLists<T> lists_;
};
template <class T>
using SensitivityFilter = SensitivityFilterOf<T, 1>;
template <class T>
using SensitivityFilter2 = SensitivityFilterOf<T, 2>;

// The density filter followed by the tanh projection
// rho = (tanh(beta/2) + tanh(beta (s~ - 1/2))) / (2 tanh(beta/2)), with the
// chain-rule sensitivities; beta is set by UpdateBeta (1 initially).
template <class T>
class HeavisideFilter {
public:
    HeavisideFilter() = default;
    HeavisideFilter(int n, std::vector<std::vector<int> > neighbors, std::vector<std::vector<T> > w)
        : lists_(n, neighbors, w) {}
    void UpdateBeta(T beta) { beta_ = beta; }
    std::vector<T> GetFilteredVariables(std::vector<T> s) { return run(0, s, nullptr); }
    std::vector<T> GetFilteredSensitivitis(std::vector<T> s, std::vector<T> dfdrho) { return run(1, s, &dfdrho); }

private:
    std::vector<T> run(int mode, const std::vector<T>& s, const std::vector<T>* dfdrho) {
        lists_.ensure();
        auto& dev = Device::Get();
        std::vector<T> out(lists_.n_);
        dev.check(pf_heaviside(dev.ctx(), beta_, mode, s.data(), dfdrho ? dfdrho->data() : nullptr, out.data()),
                  "pf_heaviside");
        return out;
    }
    detail::FilterLists<T> lists_;
    T beta_ = 1.0;
};

// ============================================================== OC =======
// Bisection on lambda; the candidates are computed on the device and the
// caller's constraint closure gkp1 decides each step (OC.h:76-107).
template <class T>
class OC {
public:
    OC(int n, T iota, T lambdamin, T lambdamax, T lambdaeps, T movelimit, const std::vector<T>& xmin,
       const std::vector<T>& xmax)
        : n_(n), iota_(iota), lmin_(lambdamin), lmax_(lambdamax), leps_(lambdaeps), move_(movelimit), xmin_(xmin),
          xmax_(xmax) {}
    bool IsConvergence(T f) { return std::fabs(f - prev_) / (f + prev_) < eps_; }
    template <class F>
    void UpdateVariables(std::vector<T>& xk, T f, std::vector<T> dfdx, T /*g*/, std::vector<T> dgdx, F gkp1) {
        auto& dev = Device::Get();
        if (n_ != dev.elements()) throw DeviceError("OC: n differs from the registered grid's elements");
        T l0 = lmin_, l1 = lmax_, lambda = T();
        std::vector<T> xkp1(n_);
        while ((l1 - l0) / (l1 + l0) > leps_) {
            lambda = 0.5 * (l1 + l0);
            dev.check(pf_oc_candidate(dev.ctx(), xk.data(), dfdx.data(), dgdx.data(), lambda, iota_, move_,
                                      xkp1.data()),
                      "pf_oc_candidate");
            if (gkp1(xkp1) > T()) l0 = lambda;
            else l1 = lambda;
        }
        last_lambda_ = lambda;
        prev_ = f;
        k_++;
        xk = xkp1;
    }
    T LastLambda() const { return last_lambda_; }

private:
    int n_;
    T iota_, lmin_, lmax_, leps_, move_;
    std::vector<T> xmin_, xmax_;  // kept for the interface; the reference's update does not use them
    T prev_ = T(), eps_ = 1.0e-5, last_lambda_ = T();
    int k_ = 0;
};

// ============================================================== MMA ======
// Svanberg's MMA with one constraint on the device (state kept there).
template <class T>
class MMA {
public:
    MMA(int n, int m, T a0, std::vector<T> a, std::vector<T> c, std::vector<T> d, const std::vector<T>& xmin,
        const std::vector<T>& xmax)
        : n_(n), m_(m) {
        if (m != 1) throw DeviceError("MMA: the device solver handles one constraint (m = 1)");
        if ((int)xmin.size() != n || (int)xmax.size() != n) throw DeviceError("MMA: bound sizes differ from n");
        for (int i = 1; i < n; i++)
            if (xmin[i] != xmin[0] || xmax[i] != xmax[0]) throw DeviceError("MMA: the device solver takes uniform bounds");
        prm_.m = 1;
        prm_.a0 = a0;
        prm_.a = a.at(0);
        prm_.c = c.at(0);
        prm_.d = d.at(0);
        prm_.xmin = xmin.at(0);
        prm_.xmax = xmax.at(0);
        // SetParameters defaults of MMA.h:63-88
        prm_.raa0 = 1.0e-5, prm_.albefa = 0.1, prm_.move = 0.5, prm_.asyinit = 0.5, prm_.asydecr = 0.7,
        prm_.asyincr = 1.2, prm_.epsvalue = 1.0e-5;
        prm_.max_newton = 1000;
    }
    void SetParameters(T raa0, T albefa, T move, T asyinit, T asydecr, T asyincr, T epsvalue) {
        prm_.raa0 = raa0, prm_.albefa = albefa, prm_.move = move, prm_.asyinit = asyinit, prm_.asydecr = asydecr,
        prm_.asyincr = asyincr, prm_.epsvalue = epsvalue;
        eps_ = epsvalue;
        initialised_ = false;
    }
    bool IsConvergence(T f) { return std::fabs(f - prev_) / (f + prev_) < eps_; }
    void UpdateVariables(std::vector<T>& xk, T f, std::vector<T> dfdx, std::vector<T> g,
                         std::vector<std::vector<T> > dgdx) {
        auto& dev = Device::Get();
        if (n_ != dev.elements()) throw DeviceError("MMA: n differs from the registered grid's elements");
        if (g.size() != 1 || dgdx.size() != 1) throw DeviceError("MMA: one constraint expected");
        if (!initialised_ || dev.generation() != generation_) {
            dev.check(pf_mma_init(dev.ctx(), &prm_), "pf_mma_init");
            initialised_ = true;
            generation_ = dev.generation();
        }
        int steps = 0;
        dev.check(pf_mma_update(dev.ctx(), xk.data(), f, dfdx.data(), g[0], dgdx[0].data(), &steps), "pf_mma_update");
        prev_ = f;
    }

private:
    int n_, m_;
    pf_mma_params prm_{};
    bool initialised_ = false;
    long generation_ = -1;
    T prev_ = T(), eps_ = 1.0e-5;
};

// =========================================================== CONLIN ======
// Fleury's convex linearisation with one constraint on the device (the
// interior-point subproblem solver is MMA's).
template <class T>
class CONLIN {
public:
    CONLIN(int n, int m, T a0, std::vector<T> a, std::vector<T> c, std::vector<T> d, const std::vector<T>& xmin,
           const std::vector<T>& xmax)
        : n_(n) {
        if (m != 1) throw DeviceError("CONLIN: the device solver handles one constraint (m = 1)");
        if ((int)xmin.size() != n || (int)xmax.size() != n) throw DeviceError("CONLIN: bound sizes differ from n");
        for (int i = 1; i < n; i++)
            if (xmin[i] != xmin[0] || xmax[i] != xmax[0])
                throw DeviceError("CONLIN: the device solver takes uniform bounds");
        prm_.m = 1;
        prm_.a0 = a0;
        prm_.a = a.at(0);
        prm_.c = c.at(0);
        prm_.d = d.at(0);
        prm_.xmin = xmin.at(0);
        prm_.xmax = xmax.at(0);
        prm_.move = 0.5;  // CONLIN.h:66
        prm_.epsvalue = 1.0e-5;
        prm_.max_newton = 1000;
    }
    void SetParameters(T move, T epsvalue) {
        prm_.move = move;
        prm_.epsvalue = epsvalue;
        eps_ = epsvalue;
        initialised_ = false;
    }
    bool IsConvergence(T f) { return std::fabs(f - prev_) / (f + prev_) < eps_; }
    void UpdateVariables(std::vector<T>& xk, T f, std::vector<T> dfdx, std::vector<T> g,
                         std::vector<std::vector<T> > dgdx) {
        auto& dev = Device::Get();
        if (n_ != dev.elements()) throw DeviceError("CONLIN: n differs from the registered grid's elements");
        if (g.size() != 1 || dgdx.size() != 1) throw DeviceError("CONLIN: one constraint expected");
        if (!initialised_ || dev.generation() != generation_) {
            dev.check(pf_conlin_init(dev.ctx(), &prm_), "pf_conlin_init");
            initialised_ = true;
            generation_ = dev.generation();
        }
        int steps = 0;
        dev.check(pf_conlin_update(dev.ctx(), xk.data(), f, dfdx.data(), g[0], dgdx[0].data(), &steps),
                  "pf_conlin_update");
        prev_ = f;
    }

private:
    int n_;
    pf_mma_params prm_{};
    bool initialised_ = false;
    long generation_ = -1;
    T prev_ = T(), eps_ = 1.0e-5;
};

// ========================================================= VTK export ====
// Legacy ASCII VTK, formatted as the reference's exporter writes it
// (src/PrePost/Export/ExportToVTK.h:18-137): default stream formatting of
// every value, tabs and newlines in the same places.
inline void MakeHeadderToVTK(std::ofstream& out) {
    out << "# vtk DataFile Version 4.1\n" << "vtk output\n" << "ASCII\n" << "DATASET UNSTRUCTURED_GRID\n";
}

template <class T>
void AddPointsToVTK(const std::vector<Vector<T> >& nodes, std::ofstream& out) {
    out << "\nPOINTS\t" << nodes.size() << "\tfloat\n";
    for (const auto& node : nodes) {
        int i = 0;
        for (; i < node.SIZE(); i++) out << node(i) << "\t";
        for (; i < 3; i++) out << T() << "\t";
        out << std::endl;
    }
}

inline void AddElementToVTK(const std::vector<std::vector<int> >& elements, std::ofstream& out) {
    int datanum = 0;
    for (const auto& e : elements) datanum += (int)e.size() + 1;
    out << "\nCELLS " << elements.size() << "\t" << datanum << "\n";
    for (const auto& e : elements) {
        out << e.size() << "\t";
        for (int nd : e) out << nd << "\t";
        out << std::endl;
    }
}

inline void AddElementTypes(const std::vector<int>& types, std::ofstream& out) {
    out << "\nCELL_TYPES\t" << types.size() << "\n";
    for (int t : types) out << t << "\n";
}

template <class T>
void AddPointScalers(const std::vector<T>& values, std::string symbol, std::ofstream& out, bool isheader) {
    if (isheader) out << "\nPOINT_DATA\t" << values.size() << "\n";
    out << "SCALARS " << symbol << " float\n" << "LOOKUP_TABLE default\n";
    for (const auto& v : values) out << v << std::endl;
}

template <class T>
void AddPointScalers(const std::vector<Vector<T> >& values, std::string symbol, std::ofstream& out, bool isheader) {
    if (isheader) out << "\nPOINT_DATA\t" << values.size() << "\n";
    out << "SCALARS " << symbol << " float\n" << "LOOKUP_TABLE default\n";
    for (const auto& v : values) out << v;
}

template <class T>
void AddPointVectors(const std::vector<Vector<T> >& values, std::string symbol, std::ofstream& out, bool isheader) {
    if (isheader) out << "\nPOINT_DATA\t" << values.size() << "\n";
    out << "VECTORS " << symbol << " float\n";
    for (const auto& v : values) {
        int i = 0;
        for (; i < v.SIZE(); i++) out << v(i) << "\t";
        for (; i < 3; i++) out << T() << "\t";
        out << std::endl;
    }
}

template <class T>
void AddElementScalers(const std::vector<T>& values, std::string symbol, std::ofstream& out, bool isheader) {
    if (isheader) out << "\nCELL_DATA\t" << values.size() << "\n";
    out << "SCALARS " << symbol << " float\n" << "LOOKUP_TABLE default\n";
    for (const auto& v : values) out << v << std::endl;
}

}  // namespace PANSFEM2
