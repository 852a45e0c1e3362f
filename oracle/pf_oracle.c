/* CPU oracle for the PANSFEM2 SIMP path -- TEST INFRASTRUCTURE ONLY.
 *
 * A restatement of the reference's algorithm, routine by routine, in plain
 * C.  It is the checker for the HIP library: only tests/, the smoke entry and
 * the CPU-baseline leg of bench.py load it.  Built with -ffp-contract=off so
 * that every a*b+c rounds twice, as the reference's x86-64 build does.
 * Each routine cites the reference file:line it follows.
 */
#include "pf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

void orc_free(void* p) { free(p); }

/* pow(v, 2.0) as the reference's g++ -O3 build evaluates it: the constant
 * exponent 2.0 is folded to v*v (other exponents stay libm pow calls). */
static double pow2(double v) { return v * v; }
static double pow_folded(double v, double e) { return e == 2.0 ? v * v : pow(v, e); }

/* ---------------------------------------------------------------- mesh -- */
void orc_grid2d_nodes(int nx, int ny, double lx, double ly, double* xy) {
    /* SquareMesh.h:60-70 */
    for (int i = 0; i < nx + 1; i++)
        for (int j = 0; j < ny + 1; j++) {
            long n = (long)(ny + 1) * i + j;
            xy[2 * n + 0] = lx * (i / (double)nx);
            xy[2 * n + 1] = ly * (j / (double)ny);
        }
}

void orc_grid2d_elements(int nx, int ny, int* conn) {
    /* SquareMesh.h:97-106: BL, BR, TR, TL */
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            long e = (long)ny * i + j;
            conn[4 * e + 0] = (ny + 1) * i + j;
            conn[4 * e + 1] = (ny + 1) * (i + 1) + j;
            conn[4 * e + 2] = (ny + 1) * (i + 1) + (j + 1);
            conn[4 * e + 3] = (ny + 1) * i + (j + 1);
        }
}

int orc_dofmap(int nn, int ndpn, const int* fixed_node, const int* fixed_dof, int nfixed, int* n2g) {
    /* BoundaryCondition.h:19-25 then Assembling.h:175-186 */
    for (long i = 0; i < (long)nn * ndpn; i++) n2g[i] = 0;
    for (int k = 0; k < nfixed; k++) n2g[(long)fixed_node[k] * ndpn + fixed_dof[k]] = -1;
    int N = 0;
    for (long i = 0; i < (long)nn * ndpn; i++)
        if (n2g[i] != -1) n2g[i] = N++;
    return N;
}

/* ------------------------------------------------- dense element algebra -- */
/* Matrix<T>::operator*(Matrix) (src/LinearAlgebra/Models/Matrix.h:243-255):
 * each entry starts at 0 and accumulates over k in order. */
static void mm(const double* A, const double* B, double* C, int m, int k, int n) {
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++) {
            double v = 0.0;
            for (int q = 0; q < k; q++) v += A[i * k + q] * B[q * n + j];
            C[i * n + j] = v;
        }
}

static void transpose(const double* A, double* T, int m, int n) {
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m; j++) T[i * m + j] = A[j * n + i];
}

/* Matrix<T>::Determinant, 2x2 and 3x3 closed forms (Matrix.h:326-334). */
static double det2(const double* v) { return v[0] * v[3] - v[1] * v[2]; }
static double det3(const double* v) {
    return -v[8] * v[1] * v[3] - v[7] * v[5] * v[0] - v[2] * v[4] * v[6] + v[6] * v[1] * v[5] + v[7] * v[3] * v[2] +
           v[0] * v[4] * v[8];
}

/* Matrix<T>::Inverse via cofactors (Matrix.h:346-385): entry (i,j) is
 * (-1)^(i+j) * det(Cofactor(j,i)), then every entry is divided by det. */
static void inv2(const double* a, double* out) {
    double det = det2(a);
    out[0] = a[3] / det;
    out[1] = (-a[1]) / det;
    out[2] = (-a[2]) / det;
    out[3] = a[0] / det;
}

static void inv3(const double* a, double* out) {
    double det = det3(a);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            /* Cofactor(j, i): drop row j and column i */
            double c[4];
            int q = 0;
            for (int r = 0; r < 3; r++) {
                if (r == j) continue;
                for (int s = 0; s < 3; s++) {
                    if (s == i) continue;
                    c[q++] = a[r * 3 + s];
                }
            }
            double v = det2(c);
            out[i * 3 + j] = ((i + j) % 2 == 0) ? v : -v;
        }
    for (int k = 0; k < 9; k++) out[k] = out[k] / det;
}

/* ShapeFunction4Square<T>::dNdr (src/FEM/Controller/ShapeFunction.h:185-191). */
static void dndr_q4(double r0, double r1, double* d) {
    d[0] = -0.25 * (1.0 - r1); d[1] = 0.25 * (1.0 - r1); d[2] = 0.25 * (1.0 + r1); d[3] = -0.25 * (1.0 + r1);
    d[4] = -0.25 * (1.0 - r0); d[5] = -0.25 * (1.0 + r0); d[6] = 0.25 * (1.0 + r0); d[7] = 0.25 * (1.0 - r0);
}

/* ShapeFunction8Cubic<T>::dNdr (ShapeFunction.h:317-329). */
static void dndr_h8(double r0, double r1, double r2, double* d) {
    d[0] = -0.125 * (1.0 - r1) * (1.0 - r2); d[1] = 0.125 * (1.0 - r1) * (1.0 - r2);
    d[2] = 0.125 * (1.0 + r1) * (1.0 - r2);  d[3] = -0.125 * (1.0 + r1) * (1.0 - r2);
    d[4] = -0.125 * (1.0 - r1) * (1.0 + r2); d[5] = 0.125 * (1.0 - r1) * (1.0 + r2);
    d[6] = 0.125 * (1.0 + r1) * (1.0 + r2);  d[7] = -0.125 * (1.0 + r1) * (1.0 + r2);
    d[8] = -0.125 * (1.0 - r2) * (1.0 - r0);  d[9] = -0.125 * (1.0 - r2) * (1.0 + r0);
    d[10] = 0.125 * (1.0 - r2) * (1.0 + r0);  d[11] = 0.125 * (1.0 - r2) * (1.0 - r0);
    d[12] = -0.125 * (1.0 + r2) * (1.0 - r0); d[13] = -0.125 * (1.0 + r2) * (1.0 + r0);
    d[14] = 0.125 * (1.0 + r2) * (1.0 + r0);  d[15] = 0.125 * (1.0 + r2) * (1.0 - r0);
    d[16] = -0.125 * (1.0 - r0) * (1.0 - r1); d[17] = -0.125 * (1.0 + r0) * (1.0 - r1);
    d[18] = -0.125 * (1.0 + r0) * (1.0 + r1); d[19] = -0.125 * (1.0 - r0) * (1.0 + r1);
    d[20] = 0.125 * (1.0 - r0) * (1.0 - r1);  d[21] = 0.125 * (1.0 + r0) * (1.0 - r1);
    d[22] = 0.125 * (1.0 + r0) * (1.0 + r1);  d[23] = 0.125 * (1.0 - r0) * (1.0 + r1);
}

void orc_ke_planestrain_q4(const double* X, double E, double nu, double t, double* Ke) {
    /* PlaneStrain.h:19-58 with Gauss4Square (GaussIntegration.h:133-157):
     * points (-,-), (+,-), (-,+), (+,+) at 1/sqrt(3), weights 1. */
    const double a = 1.0 / sqrt(3.0);
    const double gp[4][2] = {{-1.0 / sqrt(3.0), -1.0 / sqrt(3.0)}, {a, -1.0 / sqrt(3.0)}, {-1.0 / sqrt(3.0), a}, {a, a}};
    double D[9] = {1.0 - nu, nu, 0.0, nu, 1.0 - nu, 0.0, 0.0, 0.0, 0.5 * (1.0 - 2.0 * nu)};
    double c = E / ((1.0 - 2.0 * nu) * (1.0 + nu));
    for (int k = 0; k < 9; k++) D[k] *= c;
    for (int k = 0; k < 64; k++) Ke[k] = 0.0;
    for (int g = 0; g < 4; g++) {
        double dN[8], dXdr[4], inv[4], dNdX[8], B[24], Bt[24], BtD[24], BtDB[64];
        dndr_q4(gp[g][0], gp[g][1], dN);
        mm(dN, X, dXdr, 2, 4, 2);
        double J = det2(dXdr);
        inv2(dXdr, inv);
        mm(inv, dN, dNdX, 2, 2, 4);
        for (int n = 0; n < 4; n++) {
            B[0 * 8 + 2 * n] = dNdX[0 * 4 + n]; B[0 * 8 + 2 * n + 1] = 0.0;
            B[1 * 8 + 2 * n] = 0.0;             B[1 * 8 + 2 * n + 1] = dNdX[1 * 4 + n];
            B[2 * 8 + 2 * n] = dNdX[1 * 4 + n]; B[2 * 8 + 2 * n + 1] = dNdX[0 * 4 + n];
        }
        transpose(B, Bt, 3, 8);
        mm(Bt, D, BtD, 8, 3, 3);
        mm(BtD, B, BtDB, 8, 3, 8);
        for (int k = 0; k < 64; k++) Ke[k] += BtDB[k] * J * t * 1.0 * 1.0;
    }
}

void orc_ke_solid_h8(const double* X, double E, double nu, double* Ke) {
    /* Solid.h:19-64 with Gauss8Cubic (GaussIntegration.h:221-253): points
     * counter-clockwise per z layer, weights 1. */
    const double a = 1.0 / sqrt(3.0), m = -1.0 / sqrt(3.0);
    const double gp[8][3] = {{m, m, m}, {a, m, m}, {a, a, m}, {m, a, m}, {m, m, a}, {a, m, a}, {a, a, a}, {m, a, a}};
    double C[36];
    for (int k = 0; k < 36; k++) C[k] = 0.0;
    C[0] = 1.0 - nu; C[1] = nu; C[2] = nu;
    C[6] = nu; C[7] = 1.0 - nu; C[8] = nu;
    C[12] = nu; C[13] = nu; C[14] = 1.0 - nu;
    C[21] = 0.5 * (1.0 - 2.0 * nu); C[28] = 0.5 * (1.0 - 2.0 * nu); C[35] = 0.5 * (1.0 - 2.0 * nu);
    double c = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
    for (int k = 0; k < 36; k++) C[k] *= c;
    for (int k = 0; k < 576; k++) Ke[k] = 0.0;
    for (int g = 0; g < 8; g++) {
        double dN[24], dXdr[9], inv[9], dNdX[24], B[144], Bt[144], BtC[144], BtCB[576];
        dndr_h8(gp[g][0], gp[g][1], gp[g][2], dN);
        mm(dN, X, dXdr, 3, 8, 3);
        double J = det3(dXdr);
        inv3(dXdr, inv);
        mm(inv, dN, dNdX, 3, 3, 8);
        for (int k = 0; k < 144; k++) B[k] = 0.0;
        for (int n = 0; n < 8; n++) {
            double dx = dNdX[0 * 8 + n], dy = dNdX[1 * 8 + n], dz = dNdX[2 * 8 + n];
            B[0 * 24 + 3 * n] = dx;
            B[1 * 24 + 3 * n + 1] = dy;
            B[2 * 24 + 3 * n + 2] = dz;
            B[3 * 24 + 3 * n] = dy; B[3 * 24 + 3 * n + 1] = dx;
            B[4 * 24 + 3 * n + 1] = dz; B[4 * 24 + 3 * n + 2] = dy;
            B[5 * 24 + 3 * n] = dz; B[5 * 24 + 3 * n + 2] = dx;
        }
        transpose(B, Bt, 6, 24);
        mm(Bt, C, BtC, 24, 6, 6);
        mm(BtC, B, BtCB, 24, 6, 24);
        for (int k = 0; k < 576; k++) Ke[k] += BtCB[k] * J * 1.0 * 1.0 * 1.0;
    }
}

/* ------------------------------------------------------------ assembly -- */
typedef struct { int col; double val; } lil_ent;
typedef struct { lil_ent* e; int n, cap; } lil_row;

static void lil_add(lil_row* r, int col, double v) {
    /* LILCSR::get then LILCSR::set (LILCSR.h:92-114): linear scan, append if absent */
    for (int k = 0; k < r->n; k++)
        if (r->e[k].col == col) { r->e[k].val = r->e[k].val + v; return; }
    if (r->n == r->cap) {
        r->cap = r->cap ? 2 * r->cap : 8;
        r->e = (lil_ent*)realloc(r->e, sizeof(lil_ent) * r->cap);
    }
    r->e[r->n].col = col;
    r->e[r->n].val = 0.0 + v;
    r->n++;
}

static int cmp_ent(const void* a, const void* b) {
    const lil_ent* x = (const lil_ent*)a;
    const lil_ent* y = (const lil_ent*)b;
    if (x->col != y->col) return x->col < y->col ? -1 : 1;
    return (x->val < y->val) ? -1 : (x->val > y->val);
}

long orc_assemble_csr(int N, int ne, int nen, int ndpn, const int* conn, const int* n2g, const double* ke_all,
                      const double* ufix, double* F, int** indptr, int** indices, double** data) {
    const int nd = nen * ndpn;
    lil_row* rows = (lil_row*)calloc((size_t)N, sizeof(lil_row));
    for (int e = 0; e < ne; e++) {
        const double* Ke = ke_all + (long)e * nd * nd;
        const int* el = conn + (long)e * nen;
        /* Assembling.h:46-66: i (node), doui, j (node), douj */
        for (int i = 0; i < nen; i++)
            for (int di = 0; di < ndpn; di++) {
                int r = n2g[(long)el[i] * ndpn + di];
                if (r == -1) continue;
                for (int j = 0; j < nen; j++)
                    for (int dj = 0; dj < ndpn; dj++) {
                        int c = n2g[(long)el[j] * ndpn + dj];
                        double kv = Ke[(i * ndpn + di) * nd + (j * ndpn + dj)];
                        if (c != -1) lil_add(&rows[r], c, kv);
                        else F[r] -= kv * ufix[(long)el[j] * ndpn + dj];
                    }
            }
    }
    long nnz = 0;
    for (int r = 0; r < N; r++) nnz += rows[r].n;
    *indptr = (int*)malloc(sizeof(int) * (size_t)(N + 1));
    *indices = (int*)malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
    *data = (double*)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    long k = 0;
    (*indptr)[0] = 0;
    for (int r = 0; r < N; r++) {
        /* CSR.h:92-105: std::sort of (col, value) pairs per row */
        qsort(rows[r].e, (size_t)rows[r].n, sizeof(lil_ent), cmp_ent);
        for (int q = 0; q < rows[r].n; q++) {
            (*indices)[k] = rows[r].e[q].col;
            (*data)[k] = rows[r].e[q].val;
            k++;
        }
        (*indptr)[r + 1] = (int)k;
        free(rows[r].e);
    }
    free(rows);
    return nnz;
}

/* ------------------------------------------------------ sparse / solver -- */
void orc_csr_spmv(int N, const int* indptr, const int* indices, const double* data, const double* x, double* y) {
    /* CSR.h:108-122: each row summed serially from 0 in column order */
    for (int i = 0; i < N; i++) {
        double v = 0.0;
        for (int j = indptr[i]; j < indptr[i + 1]; j++) v += data[j] * x[indices[j]];
        y[i] = v;
    }
}

void orc_csr_diag(int N, const int* indptr, const int* indices, const double* data, double* D) {
    /* CSR::get by lower_bound (CSR.h:154-167) */
    for (int i = 0; i < N; i++) {
        D[i] = 0.0;
        int lo = indptr[i], hi = indptr[i + 1];
        while (lo < hi) {
            int mid = lo + (hi - lo) / 2;
            if (indices[mid] < i) lo = mid + 1; else hi = mid;
        }
        if (lo < indptr[i + 1] && indices[lo] == i) D[i] = data[lo];
    }
}

static double dot_serial(int n, const double* a, const double* b) {
    /* std::inner_product with init T() */
    double s = 0.0;
    for (int i = 0; i < n; i++) s = s + a[i] * b[i];
    return s;
}

int orc_scaling_cg(int N, const int* indptr, const int* indices, const double* data, const double* b, int itrmax,
                   double eps, double* x) {
    /* CG.h:419-453 */
    double* D = (double*)malloc(sizeof(double) * (size_t)N);
    double* r = (double*)malloc(sizeof(double) * (size_t)N);
    double* p = (double*)malloc(sizeof(double) * (size_t)N);
    double* z = (double*)malloc(sizeof(double) * (size_t)N);
    double* Ap = (double*)malloc(sizeof(double) * (size_t)N);
    orc_csr_diag(N, indptr, indices, data, D);
    for (int i = 0; i < N; i++) x[i] = 0.0;
    orc_csr_spmv(N, indptr, indices, data, x, Ap);
    for (int i = 0; i < N; i++) r[i] = b[i] - Ap[i];          /* subtract (CG.h:17-27) */
    for (int i = 0; i < N; i++) p[i] = r[i] / D[i];           /* Scaling (CG.h:408-415) */
    for (int i = 0; i < N; i++) z[i] = p[i];
    double bnorm = sqrt(dot_serial(N, b, b));
    double zr = dot_serial(N, z, r);
    int ret = -1;
    for (int k = 0; k < itrmax; ++k) {
        orc_csr_spmv(N, indptr, indices, data, p, Ap);
        double alpha = zr / dot_serial(N, p, Ap);
        for (int i = 0; i < N; i++) x[i] = x[i] + alpha * p[i];       /* xexpay (CG.h:31-38) */
        for (int i = 0; i < N; i++) r[i] = r[i] + (-alpha) * Ap[i];
        for (int i = 0; i < N; i++) z[i] = r[i] / D[i];
        double zr1 = dot_serial(N, z, r);
        double beta = zr1 / zr;
        for (int i = 0; i < N; i++) p[i] = beta * p[i] + z[i];       /* xeaxpy (CG.h:42-49) */
        zr = zr1;
        double rnorm = sqrt(dot_serial(N, r, r));
        if (rnorm < eps * bnorm) { ret = k + 1; break; }
    }
    free(D); free(r); free(p); free(z); free(Ap);
    return ret;
}

/* -------------------------------------------------------------- filter -- */
void orc_centroids(int ne, int nen, int dim, const int* conn, const double* nodes, double* cg) {
    /* General.h:70-77: sum of node vectors, then divided by the node count */
    for (int e = 0; e < ne; e++) {
        double c[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < nen; k++)
            for (int d = 0; d < dim; d++) c[d] += nodes[(long)conn[(long)e * nen + k] * dim + d];
        for (int d = 0; d < dim; d++) cg[(long)e * dim + d] = c[d] / (double)nen;
    }
}

static double dist2d(const double* a, const double* b) {
    /* (cg_i - cg_j).Norm(): sum of pow(v, 2.0), then sqrt (Vector.h:297-303) */
    double v0 = a[0] - b[0], v1 = a[1] - b[1];
    double s = 0.0;
    s += pow2(v0);
    s += pow2(v1);
    return sqrt(s);
}

long orc_filter_build(int nx, int ny, const double* cg, double R, int brute, int* ptr, int** idx, double** w) {
    int ne = nx * ny;
    int win = (int)ceil(R) + 1;
    long cap = (long)ne * (2 * win + 1) * (2 * win + 1);
    if (brute) cap = (long)ne * 32;
    int* I = (int*)malloc(sizeof(int) * (size_t)cap);
    double* W = (double*)malloc(sizeof(double) * (size_t)cap);
    long k = 0;
    ptr[0] = 0;
    for (int i = 0; i < ne; i++) {
        if (brute) {
            for (int j = 0; j < ne; j++) {
                double d = dist2d(cg + 2L * i, cg + 2L * j);
                if (d <= R) {
                    if (k == cap) { cap *= 2; I = (int*)realloc(I, sizeof(int) * cap); W = (double*)realloc(W, sizeof(double) * cap); }
                    I[k] = j; W[k] = (R - dist2d(cg + 2L * i, cg + 2L * j)) / R; k++;
                }
            }
        } else {
            int ix = i / ny, iy = i % ny;
            int x0 = ix - win < 0 ? 0 : ix - win, x1 = ix + win > nx - 1 ? nx - 1 : ix + win;
            int y0 = iy - win < 0 ? 0 : iy - win, y1 = iy + win > ny - 1 ? ny - 1 : iy + win;
            for (int jx = x0; jx <= x1; jx++)
                for (int jy = y0; jy <= y1; jy++) {
                    int j = ny * jx + jy;
                    double d = dist2d(cg + 2L * i, cg + 2L * j);
                    if (d <= R) { I[k] = j; W[k] = (R - dist2d(cg + 2L * i, cg + 2L * j)) / R; k++; }
                }
        }
        ptr[i + 1] = (int)k;
    }
    *idx = I;
    *w = W;
    return k;
}

long orc_filter_build_brute(int ne, int dim, const double* cg, double R, int* ptr, int** idx, double** w) {
    /* the sample driver's all-pairs search (sample_optimize_density_oc.cpp:53-60)
     * for any dimension: Norm() sums pow(v, 2.0) per component (Vector.h:297-303) */
    long cap = (long)ne * 32, k = 0;
    int* I = (int*)malloc(sizeof(int) * (size_t)cap);
    double* W = (double*)malloc(sizeof(double) * (size_t)cap);
    ptr[0] = 0;
    for (int i = 0; i < ne; i++) {
        for (int j = 0; j < ne; j++) {
            double s = 0.0;
            for (int d = 0; d < dim; d++) s += pow2(cg[(long)i * dim + d] - cg[(long)j * dim + d]);
            double dist = sqrt(s);
            if (dist <= R) {
                if (k == cap) { cap *= 2; I = (int*)realloc(I, sizeof(int) * cap); W = (double*)realloc(W, sizeof(double) * cap); }
                I[k] = j; W[k] = (R - dist) / R; k++;
            }
        }
        ptr[i + 1] = (int)k;
    }
    *idx = I;
    *w = W;
    return k;
}

void orc_filter_apply(int ne, const int* ptr, const int* idx, const double* w, const double* s, double* rho) {
    /* DensityFilter.h:45-56 */
    for (int i = 0; i < ne; i++) {
        double r = 0.0, wsum = 0.0;
        for (int j = ptr[i]; j < ptr[i + 1]; j++) {
            r += w[j] * s[idx[j]];
            wsum += w[j];
        }
        rho[i] = r / wsum;
    }
}

void orc_filter_sens(int ne, const int* ptr, const int* idx, const double* w, const double* dfdrho, double* dfds) {
    /* DensityFilter.h:60-71: normalised by element i's own weight sum */
    for (int i = 0; i < ne; i++) {
        double r = 0.0, wsum = 0.0;
        for (int j = ptr[i]; j < ptr[i + 1]; j++) {
            r += dfdrho[idx[j]] * w[j];
            wsum += w[j];
        }
        dfds[i] = r / wsum;
    }
}

/* ------------------------------------------------------------------ OC -- */
static double std_max(double a, double b) { return (a < b) ? b : a; } /* std::max */
static double std_min(double a, double b) { return (b < a) ? b : a; } /* std::min */

/* SensitivityFilter<T> (scheme 1, src/Optimize/Filter/SensitivityFilter.h:44-55):
 * out_i = sum_j (w_ij s_j) dfds_j / (sum_j w_ij * s_i);
 * SensitivityFilter2<T> (scheme 2, SensitivityFilter.h:88-99):
 * out_i = sum_j (w_ij s_j) dfds_j / sum_j (w_ij s_j). */
void orc_sensitivity_filter(int ne, const int* ptr, const int* idx, const double* w, const double* s,
                            const double* dfds, int scheme, double* out) {
    for (int i = 0; i < ne; i++) {
        double acc = 0.0, den = 0.0;
        for (int k = ptr[i]; k < ptr[i + 1]; k++) {
            acc += w[k] * s[idx[k]] * dfds[idx[k]];
            den += scheme == 2 ? w[k] * s[idx[k]] : w[k];
        }
        out[i] = scheme == 2 ? acc / den : acc / (den * s[i]);
    }
}

/* HeavisideFilter<T>::GetFilteredVariables (src/Optimize/Filter/HeavisideFilter.h:62-75):
 * rho_i = 0.5 (tanh(beta/2) + tanh(beta (s~_i - 1/2))) / tanh(beta/2) with
 * s~_i = (sum_j w_ij s_j) / (sum_j w_ij). */
void orc_heaviside_vars(int ne, const int* ptr, const int* idx, const double* w, double beta, const double* s,
                        double* rho) {
    for (int i = 0; i < ne; i++) {
        double wssum = 0.0, wsum = 0.0;
        for (int k = ptr[i]; k < ptr[i + 1]; k++) {
            wssum += w[k] * s[idx[k]];
            wsum += w[k];
        }
        rho[i] = 0.5 * (tanh(0.5 * beta) + tanh(beta * (wssum / wsum - 0.5))) / tanh(0.5 * beta);
    }
}

/* HeavisideFilter<T>::GetFilteredSensitivitis (HeavisideFilter.h:78-99): the
 * projection's derivative at every element, then the row-normalised sum of
 * dfdrho_j drho_j/ds~_j w_ij; pow(t, 2.0) is t*t as the reference's build
 * folds it. */
void orc_heaviside_sens(int ne, const int* ptr, const int* idx, const double* w, double beta, const double* s,
                        const double* dfdrho, double* dfds) {
    double* dr = (double*)malloc(sizeof(double) * (size_t)(ne > 0 ? ne : 1));
    for (int i = 0; i < ne; i++) {
        double wssum = 0.0, wsum = 0.0;
        for (int k = ptr[i]; k < ptr[i + 1]; k++) {
            wssum += w[k] * s[idx[k]];
            wsum += w[k];
        }
        double t = tanh(beta * (wssum / wsum - 0.5));
        dr[i] = 0.5 * beta * (1.0 - t * t) / tanh(0.5 * beta);
    }
    for (int i = 0; i < ne; i++) {
        double r = 0.0, wsum = 0.0;
        for (int k = ptr[i]; k < ptr[i + 1]; k++) {
            r += dfdrho[idx[k]] * dr[idx[k]] * w[k];
            wsum += w[k];
        }
        dfds[i] = r / wsum;
    }
    free(dr);
}

int orc_oc_update(int n, double* x, const double* dfdx, const double* dgdx, double iota, double lmin, double lmax,
                  double leps, double move, const int* fptr, const int* fidx, const double* fw, double vf,
                  double scale1, double* lambda_out, int* decisions, int max_dec) {
    return orc_oc_update_beta(n, x, dfdx, dgdx, iota, lmin, lmax, leps, move, fptr, fidx, fw, vf, scale1, 0.0,
                              lambda_out, decisions, max_dec);
}

int orc_oc_update_beta(int n, double* x, const double* dfdx, const double* dgdx, double iota, double lmin,
                       double lmax, double leps, double move, const int* fptr, const int* fidx, const double* fw,
                       double vf, double scale1, double beta, double* lambda_out, int* decisions, int max_dec) {
    /* OC.h:76-107; the closure is sample_optimize_density_oc.cpp:199-206,
     * filtering with HeavisideFilter at beta when beta > 0 */
    double l0 = lmin, l1 = lmax, lambda = 0.0;
    double* xk1 = (double*)calloc((size_t)n, sizeof(double));
    double* rr = (double*)malloc(sizeof(double) * (size_t)n);
    int steps = 0;
    while ((l1 - l0) / (l1 + l0) > leps) {
        lambda = 0.5 * (l1 + l0);
        for (int i = 0; i < n; i++) {
            xk1[i] = pow(-dfdx[i] / (dgdx[i] * lambda), iota) * x[i];
            if (xk1[i] < std_max(0.0, (1.0 - move) * x[i])) xk1[i] = std_max(0.0, (1.0 - move) * x[i]);
            else if (xk1[i] > std_min(1.0, (1.0 + move) * x[i])) xk1[i] = std_min(1.0, (1.0 + move) * x[i]);
        }
        if (beta > 0.0) orc_heaviside_vars(n, fptr, fidx, fw, beta, xk1, rr);
        else orc_filter_apply(n, fptr, fidx, fw, xk1, rr);
        double g = 0.0;
        for (int i = 0; i < n; i++) g += scale1 * rr[i] / (vf * n);
        g = g - 1.0 * scale1;
        int dec = g > 0.0;
        if (decisions && steps < max_dec) decisions[steps] = dec;
        if (dec) l0 = lambda; else l1 = lambda;
        steps++;
    }
    memcpy(x, xk1, sizeof(double) * (size_t)n);
    if (lambda_out) *lambda_out = lambda;
    free(xk1);
    free(rr);
    return steps;
}

/* ----------------------------------------------------------------- MMA -- */
static double* dalloc(long n) { return (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

orc_mma* orc_mma_create(int n, int m, double a0, const double* a, const double* c, const double* d, const double* xmin,
                        const double* xmax) {
    /* MMA.h:63-88 */
    orc_mma* s = (orc_mma*)calloc(1, sizeof(orc_mma));
    s->n = n; s->m = m; s->k = 0;
    s->previousvalue = 0.0; s->epsvalue = 1.0e-5;
    s->a0 = a0;
    s->a = dalloc(m); s->c = dalloc(m); s->d = dalloc(m);
    memcpy(s->a, a, sizeof(double) * m); memcpy(s->c, c, sizeof(double) * m); memcpy(s->d, d, sizeof(double) * m);
    s->xmin = dalloc(n); s->xmax = dalloc(n);
    memcpy(s->xmin, xmin, sizeof(double) * n); memcpy(s->xmax, xmax, sizeof(double) * n);
    s->xkm1 = dalloc(n); s->xkm2 = dalloc(n); s->L = dalloc(n); s->U = dalloc(n);
    s->raa0 = 1.0e-5; s->albefa = 0.1; s->move = 0.5; s->asyinit = 0.5; s->asydecr = 0.7; s->asyincr = 1.2;
    return s;
}

void orc_mma_set_parameters(orc_mma* s, double raa0, double albefa, double move, double asyinit, double asydecr,
                            double asyincr, double epsvalue) {
    /* MMA.h:91-100 */
    s->raa0 = raa0; s->albefa = albefa; s->move = move; s->asyinit = asyinit;
    s->asydecr = asydecr; s->asyincr = asyincr; s->epsvalue = epsvalue;
}

int orc_mma_is_convergence(orc_mma* s, double f) {
    /* MMA.h:107-113 */
    return fabs(f - s->previousvalue) / (f + s->previousvalue) < s->epsvalue;
}

void orc_mma_destroy(orc_mma* s) {
    if (!s) return;
    free(s->a); free(s->c); free(s->d); free(s->xmin); free(s->xmax);
    free(s->xkm1); free(s->xkm2); free(s->L); free(s->U);
    free(s);
}

static double max3(double a, double b, double c) { double r = a; if (r < b) r = b; if (r < c) r = c; return r; }
static double min3(double a, double b, double c) { double r = a; if (b < r) r = b; if (c < r) r = c; return r; }
static double max4(double a, double b, double c, double d) { double r = max3(a, b, c); if (r < d) r = d; return r; }
static double max5(double a, double b, double c, double d, double e) { double r = max4(a, b, c, d); if (r < e) r = e; return r; }

typedef struct {
    const double *p, *q, *p0, *q0, *alpha, *beta, *b; /* p, q: m x n */
} mma_sub;

/* MMA<T>::KKTNorm (MMA.h:422-461). */
static double mma_kktnorm(const orc_mma* s, const mma_sub* S, const double* x, const double* y, double z, const double* lambda,
                          const double* gsi, const double* ita, const double* mu, double zeta, const double* sv, double eps) {
    const int n = s->n, m = s->m;
    double norm = 0.0;
    double* g = dalloc(m);
    double* pl = dalloc(n);
    double* ql = dalloc(n);
    for (int j = 0; j < n; j++) {
        pl[j] = S->p0[j];
        ql[j] = S->q0[j];
        for (int i = 0; i < m; i++) {
            pl[j] += lambda[i] * S->p[(long)i * n + j];
            ql[j] += lambda[i] * S->q[(long)i * n + j];
            g[i] += S->p[(long)i * n + j] / (s->U[j] - x[j]) + S->q[(long)i * n + j] / (x[j] - s->L[j]);
        }
    }
    for (int j = 0; j < n; j++) {
        norm += pow2(pl[j] / pow2(s->U[j] - x[j]) - ql[j] / pow2(x[j] - s->L[j]) - gsi[j] + ita[j]);
        norm += pow2(gsi[j] * (x[j] - S->alpha[j]) - eps);
        norm += pow2(ita[j] * (S->beta[j] - x[j]) - eps);
    }
    for (int i = 0; i < m; i++) {
        norm += pow2(s->c[i] + s->d[i] * y[i] - lambda[i] - mu[i]);
        norm += pow2(g[i] - s->a[i] * z - y[i] + sv[i] - S->b[i]);
        norm += pow2(mu[i] * y[i] - eps);
        norm += pow2(lambda[i] * sv[i] - eps);
    }
    double la = 0.0;
    for (int i = 0; i < m; i++) la = la + lambda[i] * s->a[i];
    norm += pow2(s->a0 - zeta - la);
    norm += pow2(zeta * z - eps);
    free(g); free(pl); free(ql);
    return sqrt(norm);
}

/* MMA<T>::solvels: Gaussian elimination with partial pivoting (MMA.h:464-509). */
static void mma_solvels(int nb, double* A, double* b, double* x) {
    for (int i = 0; i < nb - 1; i++) {
        double pivot = fabs(A[i * nb + i]);
        int pi = i;
        for (int j = i + 1; j < nb; j++)
            if (pivot < fabs(A[j * nb + i])) { pivot = fabs(A[j * nb + i]); pi = j; }
        if (pi != i) {
            double tmp = b[i]; b[i] = b[pi]; b[pi] = tmp;
            for (int j = i; j < nb; j++) { tmp = A[i * nb + j]; A[i * nb + j] = A[pi * nb + j]; A[pi * nb + j] = tmp; }
        }
        for (int j = i + 1; j < nb; j++) {
            for (int k = i + 1; k < nb; k++) A[j * nb + k] -= A[i * nb + k] * A[j * nb + i] / A[i * nb + i];
            b[j] -= b[i] * A[j * nb + i] / A[i * nb + i];
        }
    }
    for (int i = nb - 1; i >= 0; i--) {
        x[i] = b[i];
        for (int j = nb - 1; j > i; j--) x[i] -= x[j] * A[i * nb + j];
        x[i] /= A[i * nb + i];
    }
}

void orc_mma_update(orc_mma* s, double* xk, double f, const double* dfdx, const double* gv, const double* dgdx) {
    const int n = s->n, m = s->m;
    double *L = s->L, *U = s->U, *xmin = s->xmin, *xmax = s->xmax;
    /* asymptotes (MMA.h:118-142) */
    if (s->k < 2) {
        for (int j = 0; j < n; j++) {
            L[j] = xk[j] - s->asyinit * (xmax[j] - xmin[j]);
            U[j] = xk[j] + s->asyinit * (xmax[j] - xmin[j]);
            L[j] = std_min(std_max(xk[j] - 10.0 * (xmax[j] - xmin[j]), L[j]), xk[j] - 0.01 * (xmax[j] - xmin[j]));
            U[j] = std_min(std_max(xk[j] + 0.01 * (xmax[j] - xmin[j]), U[j]), xk[j] + 10.0 * (xmax[j] - xmin[j]));
        }
    } else {
        for (int j = 0; j < n; j++) {
            double tmp = (xk[j] - s->xkm1[j]) * (s->xkm1[j] - s->xkm2[j]);
            if (tmp < 0.0) {
                L[j] = xk[j] - s->asydecr * (s->xkm1[j] - L[j]);
                U[j] = xk[j] + s->asydecr * (U[j] - s->xkm1[j]);
            } else if (tmp > 0.0) {
                L[j] = xk[j] - s->asyincr * (s->xkm1[j] - L[j]);
                U[j] = xk[j] + s->asyincr * (U[j] - s->xkm1[j]);
            } else {
                L[j] = xk[j] - (s->xkm1[j] - L[j]);
                U[j] = xk[j] + (U[j] - s->xkm1[j]);
            }
            L[j] = std_min(std_max(xk[j] - 10.0 * (xmax[j] - xmin[j]), L[j]), xk[j] - 0.01 * (xmax[j] - xmin[j]));
            U[j] = std_min(std_max(xk[j] + 0.01 * (xmax[j] - xmin[j]), U[j]), xk[j] + 10.0 * (xmax[j] - xmin[j]));
        }
    }
    /* move limits (MMA.h:145-150) */
    double* alpha = dalloc(n);
    double* beta = dalloc(n);
    for (int j = 0; j < n; j++) {
        alpha[j] = max3(xmin[j], L[j] + s->albefa * (xk[j] - L[j]), xk[j] - s->move * (xmax[j] - xmin[j]));
        beta[j] = min3(xmax[j], U[j] - s->albefa * (U[j] - xk[j]), xk[j] + s->move * (xmax[j] - xmin[j]));
    }
    /* p0, q0, p, q, b (MMA.h:153-175) */
    double* p0 = dalloc(n);
    double* q0 = dalloc(n);
    for (int j = 0; j < n; j++) {
        double dp = std_max(dfdx[j], 0.0), dm = std_max(-dfdx[j], 0.0);
        p0[j] = pow2(U[j] - xk[j]) * (1.001 * dp + 0.001 * dm + s->raa0 / (xmax[j] - xmin[j]));
        q0[j] = pow2(xk[j] - L[j]) * (0.001 * dp + 1.001 * dm + s->raa0 / (xmax[j] - xmin[j]));
    }
    double* p = dalloc((long)m * n);
    double* q = dalloc((long)m * n);
    double* b = dalloc(m);
    for (int i = 0; i < m; i++) {
        b[i] = -gv[i];
        for (int j = 0; j < n; j++) {
            double dp = std_max(dgdx[(long)i * n + j], 0.0), dm = std_max(-dgdx[(long)i * n + j], 0.0);
            p[(long)i * n + j] = pow2(U[j] - xk[j]) * (1.001 * dp + 0.001 * dm + s->raa0 / (xmax[j] - xmin[j]));
            q[(long)i * n + j] = pow2(xk[j] - L[j]) * (0.001 * dp + 1.001 * dm + s->raa0 / (xmax[j] - xmin[j]));
            b[i] += p[(long)i * n + j] / (U[j] - xk[j]) + q[(long)i * n + j] / (xk[j] - L[j]);
        }
    }
    mma_sub S = {p, q, p0, q0, alpha, beta, b};
    /* primal-dual interior point (MMA.h:178-411) */
    double eps = 1.0;
    double* x = dalloc(n);
    double* y = dalloc(m);
    double z = 1.0, zeta = 1.0;
    double* lambda = dalloc(m);
    double* sv = dalloc(m);
    double* gsi = dalloc(n);
    double* ita = dalloc(n);
    double* mu = dalloc(m);
    for (int i = 0; i < m; i++) { y[i] = 1.0; lambda[i] = 1.0; sv[i] = 1.0; mu[i] = std_max(1.0, 0.5 * s->c[i]); }
    for (int j = 0; j < n; j++) {
        x[j] = 0.5 * (alpha[j] + beta[j]);
        gsi[j] = std_max(1.0, 1.0 / (x[j] - alpha[j]));
        ita[j] = std_max(1.0, 1.0 / (beta[j] - x[j]));
    }
    double *pl = dalloc(n), *ql = dalloc(n), *G = dalloc((long)m * n), *Dx = dalloc(n), *dlx = dalloc(n);
    double *Dy = dalloc(m), *Dl = dalloc(m), *dly = dalloc(m), *dll = dalloc(m), *Dlly = dalloc(m), *dlly = dalloc(m);
    double *dx = dalloc(n), *dy = dalloc(m), *dlam = dalloc(m), *dgsi = dalloc(n), *dita = dalloc(n), *dmu = dalloc(m), *ds = dalloc(m);
    double *xp = dalloc(n), *yp = dalloc(m), *lp = dalloc(m), *gp = dalloc(n), *ip = dalloc(n), *mp = dalloc(m), *sp = dalloc(m);
    int steps = 0;
    for (int l = 0; eps > 1.0e-7; l++) {
        steps++;
        for (int j = 0; j < n; j++) {
            pl[j] = p0[j];
            ql[j] = q0[j];
            for (int i = 0; i < m; i++) { pl[j] += lambda[i] * p[(long)i * n + j]; ql[j] += lambda[i] * q[(long)i * n + j]; }
        }
        for (int i = 0; i < m; i++)
            for (int j = 0; j < n; j++)
                G[(long)i * n + j] = p[(long)i * n + j] / pow2(U[j] - x[j]) - q[(long)i * n + j] / pow2(x[j] - L[j]);
        for (int j = 0; j < n; j++) {
            Dx[j] = 2.0 * pl[j] / pow(U[j] - x[j], 3.0) + 2.0 * ql[j] / pow(x[j] - L[j], 3.0) + gsi[j] / (x[j] - alpha[j]) +
                    ita[j] / (beta[j] - x[j]);
            dlx[j] = pl[j] / pow2(U[j] - x[j]) - ql[j] / pow2(x[j] - L[j]) - eps / (x[j] - alpha[j]) + eps / (beta[j] - x[j]);
        }
        for (int i = 0; i < m; i++) {
            Dy[i] = s->d[i] + mu[i] / y[i];
            Dl[i] = sv[i] / lambda[i];
            dly[i] = s->c[i] + s->d[i] * y[i] - lambda[i] - eps / y[i];
        }
        double la = 0.0;
        for (int i = 0; i < m; i++) la = la + lambda[i] * s->a[i];
        double dlz = s->a0 - eps / z - la;
        for (int i = 0; i < m; i++) {
            dll[i] = -s->a[i] * z - y[i] - b[i] + eps / lambda[i];
            for (int j = 0; j < n; j++) dll[i] += p[(long)i * n + j] / (U[j] - x[j]) + q[(long)i * n + j] / (x[j] - L[j]);
            Dlly[i] = Dl[i] + 1.0 / Dy[i];
            dlly[i] = dll[i] + dly[i] / Dy[i];
        }
        double dz = 0.0, dzeta = 0.0;
        for (int i = 0; i < m; i++) { dy[i] = 0.0; dlam[i] = 0.0; dmu[i] = 0.0; ds[i] = 0.0; }
        for (int j = 0; j < n; j++) { dx[j] = 0.0; dgsi[j] = 0.0; dita[j] = 0.0; }
        if (n > m) {
            int nb = m + 1;
            double* A = dalloc((long)nb * nb);
            double* B = dalloc(nb);
            double* sol = dalloc(nb);
            for (int ii = 0; ii < m; ii++) {
                for (int jj = 0; jj < m; jj++)
                    for (int kk = 0; kk < n; kk++) A[ii * nb + jj] += G[(long)ii * n + kk] * G[(long)jj * n + kk] / Dx[kk];
                A[ii * nb + ii] += Dlly[ii];
                A[ii * nb + m] = s->a[ii];
                A[m * nb + ii] = s->a[ii];
            }
            A[m * nb + m] = -zeta / z;
            for (int ii = 0; ii < m; ii++) {
                B[ii] = dlly[ii];
                for (int jj = 0; jj < n; jj++) B[ii] -= G[(long)ii * n + jj] * dlx[jj] / Dx[jj];
            }
            B[m] = dlz;
            mma_solvels(nb, A, B, sol);
            for (int i = 0; i < m; i++) dlam[i] = sol[i];
            dz = sol[m];
            for (int j = 0; j < n; j++) {
                dx[j] = -dlx[j] / Dx[j];
                for (int i = 0; i < m; i++) dx[j] -= G[(long)i * n + j] * dlam[i] / Dx[j];
            }
            free(A); free(B); free(sol);
        } else {
            int nb = n + 1;
            double* A = dalloc((long)nb * nb);
            double* B = dalloc(nb);
            double* sol = dalloc(nb);
            for (int ii = 0; ii < n; ii++) {
                for (int jj = 0; jj < n; jj++)
                    for (int kk = 0; kk < m; kk++) A[ii * nb + jj] += G[(long)kk * n + ii] * G[(long)kk * n + jj] / Dlly[kk];
                A[ii * nb + ii] += Dx[ii];
                for (int jj = 0; jj < m; jj++) {
                    A[ii * nb + n] -= G[(long)jj * n + ii] * s->a[jj] / Dlly[jj];
                    A[n * nb + ii] -= G[(long)jj * n + ii] * s->a[jj] / Dlly[jj];
                    A[n * nb + n] += s->a[jj] * s->a[jj] / Dlly[jj];
                }
            }
            A[n * nb + n] += zeta / z;
            for (int ii = 0; ii < n; ii++) {
                B[ii] = -dlx[ii];
                for (int jj = 0; jj < m; jj++) B[ii] -= G[(long)jj * n + ii] * dlly[jj] / Dlly[jj];
            }
            B[n] = -dlz;
            for (int jj = 0; jj < m; jj++) B[n] += s->a[jj] * dlly[jj] / Dlly[jj];
            mma_solvels(nb, A, B, sol);
            for (int j = 0; j < n; j++) dx[j] = sol[j];
            dz = sol[n];
            for (int i = 0; i < m; i++) {
                dlam[i] = -s->a[i] * dz / Dlly[i] + dlly[i] / Dlly[i];
                for (int j = 0; j < n; j++) dlam[i] += G[(long)i * n + j] * dx[j] / Dlly[i];
            }
            free(A); free(B); free(sol);
        }
        for (int i = 0; i < m; i++) {
            dy[i] = dlam[i] / Dy[i] - dly[i] / Dy[i];
            dmu[i] = -mu[i] * dy[i] / y[i] - mu[i] + eps / y[i];
            ds[i] = -sv[i] * dlam[i] / lambda[i] - sv[i] + eps / lambda[i];
        }
        for (int j = 0; j < n; j++) {
            dgsi[j] = -gsi[j] * dx[j] / (x[j] - alpha[j]) - gsi[j] + eps / (x[j] - alpha[j]);
            dita[j] = ita[j] * dx[j] / (beta[j] - x[j]) - ita[j] + eps / (beta[j] - x[j]);
        }
        dzeta = -zeta * dz / z - zeta + eps / z;
        /* step length (MMA.h:346-360) */
        double txmax = 0.0;
        for (int j = 0; j < n; j++) {
            double t = max4(-1.01 * dx[j] / (x[j] - alpha[j]), 1.01 * dx[j] / (beta[j] - x[j]), -1.01 * dgsi[j] / gsi[j],
                            -1.01 * dita[j] / ita[j]);
            if (txmax < t) txmax = t;
        }
        double tymax = 0.0;
        for (int i = 0; i < m; i++) {
            double t = max4(-1.01 * dy[i] / y[i], -1.01 * dlam[i] / lambda[i], -1.01 * dmu[i] / mu[i], -1.01 * ds[i] / sv[i]);
            if (tymax < t) tymax = t;
        }
        double tau = 1.0 / max5(1.0, txmax, tymax, -1.01 * dz / z, -1.01 * dzeta / zeta);
        double dw = mma_kktnorm(s, &S, x, y, z, lambda, gsi, ita, mu, zeta, sv, eps);
        double zp = 0.0, zetap = 0.0;
        for (int ll = 0; ll < 50; ll++) {
            for (int j = 0; j < n; j++) { xp[j] = x[j] + tau * dx[j]; gp[j] = gsi[j] + tau * dgsi[j]; ip[j] = ita[j] + tau * dita[j]; }
            for (int i = 0; i < m; i++) {
                yp[i] = y[i] + tau * dy[i]; lp[i] = lambda[i] + tau * dlam[i];
                mp[i] = mu[i] + tau * dmu[i]; sp[i] = sv[i] + tau * ds[i];
            }
            zp = z + tau * dz;
            zetap = zeta + tau * dzeta;
            double dw1 = mma_kktnorm(s, &S, xp, yp, zp, lp, gp, ip, mp, zetap, sp, eps);
            if (dw1 < dw) break;
            tau *= 0.5;
        }
        memcpy(x, xp, sizeof(double) * n); memcpy(y, yp, sizeof(double) * m); z = zp;
        memcpy(lambda, lp, sizeof(double) * m); memcpy(gsi, gp, sizeof(double) * n); memcpy(ita, ip, sizeof(double) * n);
        memcpy(mu, mp, sizeof(double) * m); zeta = zetap; memcpy(sv, sp, sizeof(double) * m);
        double dw1 = mma_kktnorm(s, &S, x, y, z, lambda, gsi, ita, mu, zeta, sv, eps);
        if (dw1 < 0.9 * eps) eps *= 0.1; else eps *= 1.0;
    }
    s->newton_steps = steps;
    s->previousvalue = f;
    s->k++;
    memcpy(s->xkm2, s->xkm1, sizeof(double) * n);
    memcpy(s->xkm1, xk, sizeof(double) * n);
    memcpy(xk, x, sizeof(double) * n);
    free(alpha); free(beta); free(p0); free(q0); free(p); free(q); free(b);
    free(x); free(y); free(lambda); free(sv); free(gsi); free(ita); free(mu);
    free(pl); free(ql); free(G); free(Dx); free(dlx); free(Dy); free(Dl); free(dly); free(dll); free(Dlly); free(dlly);
    free(dx); free(dy); free(dlam); free(dgsi); free(dita); free(dmu); free(ds);
    free(xp); free(yp); free(lp); free(gp); free(ip); free(mp); free(sp);
}

/* ----------------------------------------------------------- SIMP loop -- */
/* -------------------------------------------------------------- CONLIN -- */
/* CONLIN<T> (src/Optimize/Solver/CONLIN.h:18-455): the objective and each
 * constraint are linearised in x where their derivative is positive and in
 * 1/x where it is negative; the separable subproblem is solved by the same
 * primal-dual interior-point iteration as MMA's. */
orc_conlin* orc_conlin_create(int n, int m, double a0, const double* a, const double* c, const double* d,
                              const double* xmin, const double* xmax) {
    orc_conlin* s = (orc_conlin*)calloc(1, sizeof(orc_conlin));
    s->n = n; s->m = m; s->k = 0;
    s->previousvalue = 0.0; s->epsvalue = 1.0e-5; s->move = 0.5;
    s->a0 = a0;
    s->a = dalloc(m); s->c = dalloc(m); s->d = dalloc(m);
    memcpy(s->a, a, sizeof(double) * m); memcpy(s->c, c, sizeof(double) * m); memcpy(s->d, d, sizeof(double) * m);
    s->xmin = dalloc(n); s->xmax = dalloc(n);
    memcpy(s->xmin, xmin, sizeof(double) * n); memcpy(s->xmax, xmax, sizeof(double) * n);
    return s;
}

void orc_conlin_set_parameters(orc_conlin* s, double move, double epsvalue) { s->move = move; s->epsvalue = epsvalue; }

int orc_conlin_is_convergence(orc_conlin* s, double f) {
    return fabs(f - s->previousvalue) / (f + s->previousvalue) < s->epsvalue;
}

void orc_conlin_destroy(orc_conlin* s) {
    if (!s) return;
    free(s->a); free(s->c); free(s->d); free(s->xmin); free(s->xmax);
    free(s);
}

/* CONLIN<T>::KKTNorm (CONLIN.h:371-413) */
static double conlin_kktnorm(const orc_conlin* s, const double* p, const double* q, const double* p0,
                             const double* q0, const double* alpha, const double* beta, const double* b,
                             const double* x, const double* y, double z, const double* lambda, const double* gsi,
                             const double* ita, const double* mu, double zeta, const double* sv, double eps) {
    const int n = s->n, m = s->m;
    double norm = 0.0;
    double* g = dalloc(m);
    double* pl = dalloc(n);
    double* ql = dalloc(n);
    for (int j = 0; j < n; j++) {
        pl[j] = p0[j];
        ql[j] = q0[j];
        for (int i = 0; i < m; i++) {
            pl[j] += lambda[i] * p[(long)i * n + j];
            ql[j] += lambda[i] * q[(long)i * n + j];
            g[i] += p[(long)i * n + j] * x[j] + q[(long)i * n + j] / x[j];
        }
    }
    for (int j = 0; j < n; j++) {
        norm += pow2(pl[j] - ql[j] / pow2(x[j]) - gsi[j] + ita[j]);
        norm += pow2(gsi[j] * (x[j] - alpha[j]) - eps);
        norm += pow2(ita[j] * (beta[j] - x[j]) - eps);
    }
    for (int i = 0; i < m; i++) {
        norm += pow2(s->c[i] + s->d[i] * y[i] - lambda[i] - mu[i]);
        norm += pow2(g[i] - s->a[i] * z - y[i] + sv[i] - b[i]);
        norm += pow2(mu[i] * y[i] - eps);
        norm += pow2(lambda[i] * sv[i] - eps);
    }
    double la = 0.0;
    for (int i = 0; i < m; i++) la = la + lambda[i] * s->a[i];
    norm += pow2(s->a0 - zeta - la);
    norm += pow2(zeta * z - eps);
    free(g); free(pl); free(ql);
    return sqrt(norm);
}

/* CONLIN<T>::UpdateVariables (CONLIN.h:90-365) */
void orc_conlin_update(orc_conlin* s, double* xk, double f, const double* dfdx, const double* gv, const double* dgdx) {
    const int n = s->n, m = s->m;
    const double *xmin = s->xmin, *xmax = s->xmax;
    double* alpha = dalloc(n);
    double* beta = dalloc(n);
    for (int j = 0; j < n; j++) {
        alpha[j] = std_max(xmin[j], xk[j] - s->move * (xmax[j] - xmin[j]));
        beta[j] = std_min(xmax[j], xk[j] + s->move * (xmax[j] - xmin[j]));
    }
    double* p0 = dalloc(n);
    double* q0 = dalloc(n);
    for (int j = 0; j < n; j++) {
        p0[j] = std_max(dfdx[j], 0.0);
        q0[j] = std_max(-dfdx[j], 0.0) * pow2(xk[j]);
    }
    double* p = dalloc((long)m * n);
    double* q = dalloc((long)m * n);
    double* b = dalloc(m);
    for (int i = 0; i < m; i++) {
        b[i] = -gv[i];
        for (int j = 0; j < n; j++) {
            p[(long)i * n + j] = std_max(dgdx[(long)i * n + j], 0.0);
            q[(long)i * n + j] = std_max(-dgdx[(long)i * n + j], 0.0) * pow2(xk[j]);
            b[i] += p[(long)i * n + j] * xk[j] + q[(long)i * n + j] / xk[j];
        }
    }
    double eps = 1.0;
    double* x = dalloc(n);
    double* y = dalloc(m);
    double z = 1.0, zeta = 1.0;
    double* lambda = dalloc(m);
    double* sv = dalloc(m);
    double* gsi = dalloc(n);
    double* ita = dalloc(n);
    double* mu = dalloc(m);
    for (int i = 0; i < m; i++) { y[i] = 1.0; lambda[i] = 1.0; sv[i] = 1.0; mu[i] = std_max(1.0, 0.5 * s->c[i]); }
    for (int j = 0; j < n; j++) {
        x[j] = 0.5 * (alpha[j] + beta[j]);
        gsi[j] = std_max(1.0, 1.0 / (x[j] - alpha[j]));
        ita[j] = std_max(1.0, 1.0 / (beta[j] - x[j]));
    }
    double *pl = dalloc(n), *ql = dalloc(n), *G = dalloc((long)m * n), *Dx = dalloc(n), *dlx = dalloc(n);
    double *Dy = dalloc(m), *Dl = dalloc(m), *dly = dalloc(m), *dll = dalloc(m), *Dlly = dalloc(m), *dlly = dalloc(m);
    double *dx = dalloc(n), *dy = dalloc(m), *dlam = dalloc(m), *dgsi = dalloc(n), *dita = dalloc(n), *dmu = dalloc(m), *ds = dalloc(m);
    double *xp = dalloc(n), *yp = dalloc(m), *lp = dalloc(m), *gp = dalloc(n), *ip = dalloc(n), *mp = dalloc(m), *sp = dalloc(m);
    int steps = 0;
    for (int l = 0; eps > 1.0e-7; l++) {
        steps++;
        for (int j = 0; j < n; j++) {
            pl[j] = p0[j];
            ql[j] = q0[j];
            for (int i = 0; i < m; i++) { pl[j] += lambda[i] * p[(long)i * n + j]; ql[j] += lambda[i] * q[(long)i * n + j]; }
        }
        for (int i = 0; i < m; i++)
            for (int j = 0; j < n; j++) G[(long)i * n + j] = p[(long)i * n + j] - q[(long)i * n + j] / pow2(x[j]);
        for (int j = 0; j < n; j++) {
            Dx[j] = 2.0 * ql[j] / pow(x[j], 3.0) + gsi[j] / (x[j] - alpha[j]) + ita[j] / (beta[j] - x[j]);
            dlx[j] = pl[j] - ql[j] / pow2(x[j]) - eps / (x[j] - alpha[j]) + eps / (beta[j] - x[j]);
        }
        for (int i = 0; i < m; i++) {
            Dy[i] = s->d[i] + mu[i] / y[i];
            Dl[i] = sv[i] / lambda[i];
            dly[i] = s->c[i] + s->d[i] * y[i] - lambda[i] - eps / y[i];
        }
        double la = 0.0;
        for (int i = 0; i < m; i++) la = la + lambda[i] * s->a[i];
        double dlz = s->a0 - eps / z - la;
        for (int i = 0; i < m; i++) {
            dll[i] = -s->a[i] * z - y[i] - b[i] + eps / lambda[i];
            for (int j = 0; j < n; j++) dll[i] += p[(long)i * n + j] * x[j] + q[(long)i * n + j] / x[j];
            Dlly[i] = Dl[i] + 1.0 / Dy[i];
            dlly[i] = dll[i] + dly[i] / Dy[i];
        }
        double dz = 0.0, dzeta = 0.0;
        for (int i = 0; i < m; i++) { dy[i] = 0.0; dlam[i] = 0.0; dmu[i] = 0.0; ds[i] = 0.0; }
        for (int j = 0; j < n; j++) { dx[j] = 0.0; dgsi[j] = 0.0; dita[j] = 0.0; }
        if (n > m) {
            int nb = m + 1;
            double* A = dalloc((long)nb * nb);
            double* B = dalloc(nb);
            double* sol = dalloc(nb);
            for (int ii = 0; ii < m; ii++) {
                for (int jj = 0; jj < m; jj++)
                    for (int kk = 0; kk < n; kk++) A[ii * nb + jj] += G[(long)ii * n + kk] * G[(long)jj * n + kk] / Dx[kk];
                A[ii * nb + ii] += Dlly[ii];
                A[ii * nb + m] = s->a[ii];
                A[m * nb + ii] = s->a[ii];
            }
            A[m * nb + m] = -zeta / z;
            for (int ii = 0; ii < m; ii++) {
                B[ii] = dlly[ii];
                for (int jj = 0; jj < n; jj++) B[ii] -= G[(long)ii * n + jj] * dlx[jj] / Dx[jj];
            }
            B[m] = dlz;
            mma_solvels(nb, A, B, sol);
            for (int i = 0; i < m; i++) dlam[i] = sol[i];
            dz = sol[m];
            for (int j = 0; j < n; j++) {
                dx[j] = -dlx[j] / Dx[j];
                for (int i = 0; i < m; i++) dx[j] -= G[(long)i * n + j] * dlam[i] / Dx[j];
            }
            free(A); free(B); free(sol);
        } else {
            int nb = n + 1;
            double* A = dalloc((long)nb * nb);
            double* B = dalloc(nb);
            double* sol = dalloc(nb);
            for (int ii = 0; ii < n; ii++) {
                for (int jj = 0; jj < n; jj++)
                    for (int kk = 0; kk < m; kk++) A[ii * nb + jj] += G[(long)kk * n + ii] * G[(long)kk * n + jj] / Dlly[kk];
                A[ii * nb + ii] += Dx[ii];
                for (int jj = 0; jj < m; jj++) {
                    A[ii * nb + n] -= G[(long)jj * n + ii] * s->a[jj] / Dlly[jj];
                    A[n * nb + ii] -= G[(long)jj * n + ii] * s->a[jj] / Dlly[jj];
                    A[n * nb + n] += s->a[jj] * s->a[jj] / Dlly[jj];
                }
            }
            A[n * nb + n] += zeta / z;
            for (int ii = 0; ii < n; ii++) {
                B[ii] = -dlx[ii];
                for (int jj = 0; jj < m; jj++) B[ii] -= G[(long)jj * n + ii] * dlly[jj] / Dlly[jj];
            }
            B[n] = -dlz;
            for (int jj = 0; jj < m; jj++) B[n] += s->a[jj] * dlly[jj] / Dlly[jj];
            mma_solvels(nb, A, B, sol);
            for (int j = 0; j < n; j++) dx[j] = sol[j];
            dz = sol[n];
            for (int i = 0; i < m; i++) {
                dlam[i] = -s->a[i] * dz / Dlly[i] + dlly[i] / Dlly[i];
                for (int j = 0; j < n; j++) dlam[i] += G[(long)i * n + j] * dx[j] / Dlly[i];
            }
            free(A); free(B); free(sol);
        }
        for (int i = 0; i < m; i++) {
            dy[i] = dlam[i] / Dy[i] - dly[i] / Dy[i];
            dmu[i] = -mu[i] * dy[i] / y[i] - mu[i] + eps / y[i];
            ds[i] = -sv[i] * dlam[i] / lambda[i] - sv[i] + eps / lambda[i];
        }
        for (int j = 0; j < n; j++) {
            dgsi[j] = -gsi[j] * dx[j] / (x[j] - alpha[j]) - gsi[j] + eps / (x[j] - alpha[j]);
            dita[j] = ita[j] * dx[j] / (beta[j] - x[j]) - ita[j] + eps / (beta[j] - x[j]);
        }
        dzeta = -zeta * dz / z - zeta + eps / z;
        double txmax = 0.0;
        for (int j = 0; j < n; j++) {
            double t = max4(-1.01 * dx[j] / (x[j] - alpha[j]), 1.01 * dx[j] / (beta[j] - x[j]), -1.01 * dgsi[j] / gsi[j],
                            -1.01 * dita[j] / ita[j]);
            if (txmax < t) txmax = t;
        }
        double tymax = 0.0;
        for (int i = 0; i < m; i++) {
            double t = max4(-1.01 * dy[i] / y[i], -1.01 * dlam[i] / lambda[i], -1.01 * dmu[i] / mu[i], -1.01 * ds[i] / sv[i]);
            if (tymax < t) tymax = t;
        }
        double tau = 1.0 / max5(1.0, txmax, tymax, -1.01 * dz / z, -1.01 * dzeta / zeta);
        double dw = conlin_kktnorm(s, p, q, p0, q0, alpha, beta, b, x, y, z, lambda, gsi, ita, mu, zeta, sv, eps);
        double zp = 0.0, zetap = 0.0;
        for (int ll = 0; ll < 50; ll++) {
            for (int j = 0; j < n; j++) { xp[j] = x[j] + tau * dx[j]; gp[j] = gsi[j] + tau * dgsi[j]; ip[j] = ita[j] + tau * dita[j]; }
            for (int i = 0; i < m; i++) {
                yp[i] = y[i] + tau * dy[i]; lp[i] = lambda[i] + tau * dlam[i];
                mp[i] = mu[i] + tau * dmu[i]; sp[i] = sv[i] + tau * ds[i];
            }
            zp = z + tau * dz;
            zetap = zeta + tau * dzeta;
            double dw1 = conlin_kktnorm(s, p, q, p0, q0, alpha, beta, b, xp, yp, zp, lp, gp, ip, mp, zetap, sp, eps);
            if (dw1 < dw) break;
            tau *= 0.5;
        }
        memcpy(x, xp, sizeof(double) * n); memcpy(y, yp, sizeof(double) * m); z = zp;
        memcpy(lambda, lp, sizeof(double) * m); memcpy(gsi, gp, sizeof(double) * n); memcpy(ita, ip, sizeof(double) * n);
        memcpy(mu, mp, sizeof(double) * m); zeta = zetap; memcpy(sv, sp, sizeof(double) * m);
        double dw2 = conlin_kktnorm(s, p, q, p0, q0, alpha, beta, b, x, y, z, lambda, gsi, ita, mu, zeta, sv, eps);
        if (dw2 < 0.9 * eps) eps *= 0.1;
        else eps *= 1.0;
    }
    s->previousvalue = f;
    s->k++;
    s->newton_steps = steps;
    memcpy(xk, x, sizeof(double) * n);
    free(alpha); free(beta); free(p0); free(q0); free(p); free(q); free(b);
    free(x); free(y); free(lambda); free(sv); free(gsi); free(ita); free(mu);
    free(pl); free(ql); free(G); free(Dx); free(dlx); free(Dy); free(Dl); free(dly); free(dll); free(Dlly); free(dlly);
    free(dx); free(dy); free(dlam); free(dgsi); free(dita); free(dmu); free(ds);
    free(xp); free(yp); free(lp); free(gp); free(ip); free(mp); free(sp);
}

void orc_simp_default_params(orc_simp_params* p) {
    p->E1 = 1.0; p->E0 = 1.0e-9; p->nu = 0.3; p->t = 1.0; p->p = 3.0; p->vf = 0.5; p->R = 1.5;
    p->scale0 = 1.0; p->scale1 = 1.0; p->cg_eps = 1.0e-10; p->cg_itrmax = 100000; p->use_mma = 0;
    p->oc_iota = 0.5; p->oc_lmin = 0.0; p->oc_lmax = 1.0e9; p->oc_leps = 1.0e-3; p->oc_move = 0.2;
    p->mma_a0 = 1.0; p->mma_c = 1.0e4; p->mma_xmin = 0.01; p->mma_xmax = 1.0; p->mma_raa0 = 1.0e-5; p->mma_albefa = 0.1;
    p->mma_move = 0.2; p->mma_asyinit = 0.5; p->mma_asydecr = 0.7; p->mma_asyincr = 1.2; p->mma_eps = 1.0e-6;
    p->heaviside = 0; p->beta0 = 0.5; p->beta_factor = 2.0; p->beta_every = 40;
}

int orc_simp2d(int nx, int ny, int iters, const orc_simp_params* P, double* s, double* rho_out, double* f_hist,
               double* g_hist, double* lambda_hist, int* cg_hist, double* u0_free) {
    /* sample_optimize_density_oc.cpp:22-208 (OC) / sample_optimize_density_mma.cpp:80-204 (MMA) */
    const int nn = (nx + 1) * (ny + 1), ne = nx * ny;
    double* xy = dalloc(2L * nn);
    int* conn = (int*)malloc(sizeof(int) * 4 * (size_t)ne);
    orc_grid2d_nodes(nx, ny, (double)nx, (double)ny, xy);
    orc_grid2d_elements(nx, ny, conn);
    /* clamp |x| < 1e-5 on dofs {0,1}; unit downward load at (lx, 0) */
    int nfix = 0;
    int* fn = (int*)malloc(sizeof(int) * 2 * (size_t)nn);
    int* fd = (int*)malloc(sizeof(int) * 2 * (size_t)nn);
    int load_node = -1;
    for (int i = 0; i < nx + 1; i++)
        for (int j = 0; j < ny + 1; j++) {
            int n = (ny + 1) * i + j;
            double X0 = (double)nx * (i / (double)nx), X1 = (double)ny * (j / (double)ny);
            if (fabs(X0) < 1.0e-5) { fn[nfix] = n; fd[nfix++] = 0; fn[nfix] = n; fd[nfix++] = 1; }
            if (fabs(X0 - (double)nx) < 1.0e-5 && fabs(X1) < 1.0e-5) load_node = n;
        }
    int* n2g = (int*)malloc(sizeof(int) * 2 * (size_t)nn);
    int N = orc_dofmap(nn, 2, fn, fd, nfix, n2g);
    double* cg = dalloc(2L * ne);
    orc_centroids(ne, 4, 2, conn, xy, cg);
    int* fptr = (int*)malloc(sizeof(int) * (size_t)(ne + 1));
    int* fidx;
    double* fw;
    orc_filter_build(nx, ny, cg, P->R, 0, fptr, &fidx, &fw);

    double *rho = dalloc(ne), *dgdrho = dalloc(ne), *dfdrho = dalloc(ne), *dfds = dalloc(ne), *dgds = dalloc(ne);
    double *ke_all = dalloc(64L * ne), *F = dalloc(N), *ufix = dalloc(2L * nn), *uf = dalloc(N), *u = dalloc(2L * nn);
    double *rf = dalloc(2L * nn), *ke1 = dalloc(64L * ne);
    /* E = 1 element matrices do not depend on rho */
    for (int e = 0; e < ne; e++) {
        double X[8];
        for (int k = 0; k < 4; k++) { X[2 * k] = xy[2L * conn[4L * e + k]]; X[2 * k + 1] = xy[2L * conn[4L * e + k] + 1]; }
        orc_ke_planestrain_q4(X, 1.0, P->nu, P->t, ke1 + 64L * e);
    }
    orc_mma* mma = NULL;
    orc_conlin* conlin = NULL;
    if (P->use_mma == 2) {
        /* CONLIN with the CONLIN sample's constants (sample_optimize_density_CONLIN.cpp:80-85) */
        double a = 0.0, c = P->mma_c, d = 0.0;
        double *xmin = dalloc(ne), *xmax = dalloc(ne);
        for (int i = 0; i < ne; i++) { xmin[i] = P->mma_xmin; xmax[i] = P->mma_xmax; }
        conlin = orc_conlin_create(ne, 1, P->mma_a0, &a, &c, &d, xmin, xmax);
        orc_conlin_set_parameters(conlin, P->mma_move, P->mma_eps);
        free(xmin); free(xmax);
    } else if (P->use_mma) {
        double a = 0.0, c = P->mma_c, d = 0.0;
        double *xmin = dalloc(ne), *xmax = dalloc(ne);
        for (int i = 0; i < ne; i++) { xmin[i] = P->mma_xmin; xmax[i] = P->mma_xmax; }
        mma = orc_mma_create(ne, 1, P->mma_a0, &a, &c, &d, xmin, xmax);
        orc_mma_set_parameters(mma, P->mma_raa0, P->mma_albefa, P->mma_move, P->mma_asyinit, P->mma_asydecr, P->mma_asyincr, P->mma_eps);
        free(xmin); free(xmax);
    }
    double beta = P->beta0;
    for (int k = 0; k < iters; k++) {
        /* beta continuation of the density samples (sample_optimize_density_oc.cpp:89-93) */
        if (P->heaviside && k % P->beta_every == 0) beta *= P->beta_factor;
        if (P->heaviside) orc_heaviside_vars(ne, fptr, fidx, fw, beta, s, rho);
        else orc_filter_apply(ne, fptr, fidx, fw, s, rho);
        double g = 0.0;
        for (int i = 0; i < ne; i++) {
            g += P->scale1 * rho[i] / (P->vf * ne);
            dgdrho[i] = P->scale1 / (P->vf * ne);
        }
        g -= 1.0 * P->scale1;
        for (int e = 0; e < ne; e++) {
            double E = P->E1 * pow_folded(rho[e], P->p) + P->E0 * (1.0 - pow_folded(rho[e], P->p));
            double X[8];
            for (int q = 0; q < 4; q++) { X[2 * q] = xy[2L * conn[4L * e + q]]; X[2 * q + 1] = xy[2L * conn[4L * e + q] + 1]; }
            orc_ke_planestrain_q4(X, E, P->nu, P->t, ke_all + 64L * e);
        }
        for (int i = 0; i < N; i++) F[i] = 0.0;
        int *ip, *ix;
        double* dv;
        orc_assemble_csr(N, ne, 4, 2, conn, n2g, ke_all, ufix, F, &ip, &ix, &dv);
        if (load_node >= 0 && n2g[2L * load_node + 1] != -1) F[n2g[2L * load_node + 1]] += -1.0;
        int it = orc_scaling_cg(N, ip, ix, dv, F, P->cg_itrmax, P->cg_eps, uf);
        if (cg_hist) cg_hist[k] = it;
        if (k == 0 && u0_free) memcpy(u0_free, uf, sizeof(double) * (size_t)N);
        for (long i = 0; i < 2L * nn; i++) u[i] = (n2g[i] != -1) ? uf[n2g[i]] : ufix[i];
        free(ip); free(ix); free(dv);
        /* reactions with all DOFs renumbered (node-major), then f = u . r */
        for (long i = 0; i < 2L * nn; i++) rf[i] = 0.0;
        for (int e = 0; e < ne; e++) {
            const double* Ke = ke_all + 64L * e;
            double ue[8], ke_ue[8];
            for (int q = 0; q < 4; q++) { ue[2 * q] = u[2L * conn[4L * e + q]]; ue[2 * q + 1] = u[2L * conn[4L * e + q] + 1]; }
            for (int a = 0; a < 8; a++) { double v = 0.0; for (int b2 = 0; b2 < 8; b2++) v += Ke[a * 8 + b2] * ue[b2]; ke_ue[a] = v; }
            for (int q = 0; q < 4; q++) { rf[2L * conn[4L * e + q]] += ke_ue[2 * q]; rf[2L * conn[4L * e + q] + 1] += ke_ue[2 * q + 1]; }
        }
        double f = 0.0;
        for (int nidx = 0; nidx < nn; nidx++) {
            double dn = 0.0;
            dn += u[2L * nidx] * rf[2L * nidx];
            dn += u[2L * nidx + 1] * rf[2L * nidx + 1];
            f = f + dn;
        }
        f = P->scale0 * f;
        for (int e = 0; e < ne; e++) {
            const double* Ke = ke1 + 64L * e;
            double ue[8], ke_ue[8];
            for (int q = 0; q < 4; q++) { ue[2 * q] = u[2L * conn[4L * e + q]]; ue[2 * q + 1] = u[2L * conn[4L * e + q] + 1]; }
            for (int a = 0; a < 8; a++) { double v = 0.0; for (int b2 = 0; b2 < 8; b2++) v += Ke[a * 8 + b2] * ue[b2]; ke_ue[a] = v; }
            double uku = 0.0;
            for (int a = 0; a < 8; a++) uku += ue[a] * ke_ue[a];
            dfdrho[e] = -P->scale0 * P->p * (-P->E0 + P->E1) * pow_folded(rho[e], P->p - 1.0) * uku;
        }
        if (P->heaviside) {
            orc_heaviside_sens(ne, fptr, fidx, fw, beta, s, dfdrho, dfds);
            orc_heaviside_sens(ne, fptr, fidx, fw, beta, s, dgdrho, dgds);
        } else {
            orc_filter_sens(ne, fptr, fidx, fw, dfdrho, dfds);
            orc_filter_sens(ne, fptr, fidx, fw, dgdrho, dgds);
        }
        if (f_hist) f_hist[k] = f;
        if (g_hist) g_hist[k] = g;
        double lam = 0.0;
        if (!P->use_mma) {
            orc_oc_update_beta(ne, s, dfds, dgds, P->oc_iota, P->oc_lmin, P->oc_lmax, P->oc_leps, P->oc_move, fptr, fidx,
                               fw, P->vf, P->scale1, P->heaviside ? beta : 0.0, &lam, NULL, 0);
        } else if (conlin) {
            orc_conlin_update(conlin, s, f, dfds, &g, dgds);
        } else {
            orc_mma_update(mma, s, f, dfds, &g, dgds);
        }
        if (lambda_hist) lambda_hist[k] = lam;
    }
    if (rho_out) {
        if (P->heaviside) orc_heaviside_vars(ne, fptr, fidx, fw, beta, s, rho_out);
        else orc_filter_apply(ne, fptr, fidx, fw, s, rho_out);
    }
    orc_mma_destroy(mma);
    orc_conlin_destroy(conlin);
    free(xy); free(conn); free(fn); free(fd); free(n2g); free(cg); free(fptr); free(fidx); free(fw);
    free(rho); free(dgdrho); free(dfdrho); free(dfds); free(dgds); free(ke_all); free(F); free(ufix); free(uf); free(u);
    free(rf); free(ke1);
    return N;
}
