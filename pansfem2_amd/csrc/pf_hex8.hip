// Hex8 (3D solid) kernels of the SIMP loop on a structured grid.
//
// Node (i,j,k) -> ((ny+1)(nz+1)) i + (nz+1) j + k; element (i,j,k) ->
// ny nz i + nz j + k; element nodes in ShapeFunction8Cubic order
// (src/FEM/Controller/ShapeFunction.h:299): the bottom face z = k
// counter-clockwise (i,j), (i+1,j), (i+1,j+1), (i,j+1), then the top face.
// Included by pf_capi.hip (same translation unit, no FP contraction).
#pragma once

namespace pf {

struct G3 {
    int nx, ny, nz;
    __device__ long node(int i, int j, int k) const { return ((long)(ny + 1) * (nz + 1)) * i + (long)(nz + 1) * j + k; }
    __device__ long elem(int i, int j, int k) const { return ((long)ny * nz) * i + (long)nz * j + k; }
    __device__ long nnode() const { return (long)(nx + 1) * (ny + 1) * (nz + 1); }
    __device__ long nelem() const { return (long)nx * ny * nz; }
};

// local index of element corner (di, dj, dk), each 0/1
__device__ inline int loc_h8(int di, int dj, int dk) { return 4 * dk + (dj == 0 ? di : 3 - di); }

// y(node) = sum over the node's elements (ascending id) of E_e times rows
// 3 loc .. 3 loc + 2 of K0 applied to the element vector.  v[a][b][c] holds
// the 3 values of node (i+a-1, j+b-1, k+c-1).
template <class T>
__device__ inline void apply_h8_node(const G3& g, int i, int j, int k, const T (&v)[3][3][3][3],
                                     const T* __restrict__ E, const T* __restrict__ K0, T (&y)[3]) {
#pragma clang fp contract(fast)  // not a check-path kernel: FMA allowed
    y[0] = y[1] = y[2] = 0;
    for (int ea = 0; ea < 2; ea++) {
        int ei = i - 1 + ea;
        if (ei < 0 || ei >= g.nx) continue;
        for (int eb = 0; eb < 2; eb++) {
            int ej = j - 1 + eb;
            if (ej < 0 || ej >= g.ny) continue;
            for (int ec = 0; ec < 2; ec++) {
                int ek = k - 1 + ec;
                if (ek < 0 || ek >= g.nz) continue;
                const T Ee = E[g.elem(ei, ej, ek)];
                const int me = loc_h8(i - ei, j - ej, k - ek);
                T acc[3] = {0, 0, 0};
#pragma unroll
                for (int l = 0; l < 8; l++) {
                    const int dk = l >> 2, dj = ((l & 3) >= 2) ? 1 : 0;
                    const int di = ((l & 3) == 1 || (l & 3) == 2) ? 1 : 0;
                    const T* u = v[ea + di][eb + dj][ec + dk];
#pragma unroll
                    for (int r = 0; r < 3; r++) {
                        const T* row = K0 + (3 * me + r) * 24 + 3 * l;
                        acc[r] += row[0] * u[0];
                        acc[r] += row[1] * u[1];
                        acc[r] += row[2] * u[2];
                    }
                }
                y[0] += Ee * acc[0];
                y[1] += Ee * acc[1];
                y[2] += Ee * acc[2];
            }
        }
    }
}

template <class T>
__device__ inline void gather_h8(const G3& g, int i, int j, int k, const T* __restrict__ x, T (&v)[3][3][3][3]) {
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                int ii = i + a - 1, jj = j + b - 1, kk = k + c - 1;
                if (ii >= 0 && ii <= g.nx && jj >= 0 && jj <= g.ny && kk >= 0 && kk <= g.nz) {
                    long m = g.node(ii, jj, kk);
                    v[a][b][c][0] = x[3 * m];
                    v[a][b][c][1] = x[3 * m + 1];
                    v[a][b][c][2] = x[3 * m + 2];
                } else {
                    v[a][b][c][0] = v[a][b][c][1] = v[a][b][c][2] = 0;
                }
            }
}

__device__ inline void node_ijk(const G3& g, long n, int& i, int& j, int& k) {
    long pl = (long)(g.ny + 1) * (g.nz + 1);
    i = (int)(n / pl);
    long rem = n - (long)i * pl;
    j = (int)(rem / (g.nz + 1));
    k = (int)(rem - (long)j * (g.nz + 1));
}

__device__ inline void elem_ijk(const G3& g, long e, int& i, int& j, int& k) {
    long pl = (long)g.ny * g.nz;
    i = (int)(e / pl);
    long rem = e - (long)i * pl;
    j = (int)(rem / g.nz);
    k = (int)(rem - (long)j * g.nz);
}

// y = K x (nodal), optionally b - K x with the r.r partials.
template <class T>
__global__ void __launch_bounds__(256) k_apply_h8(G3 g, const T* __restrict__ x, T* __restrict__ y,
                                                  const T* __restrict__ E, const T* __restrict__ K0,
                                                  const uint8_t* __restrict__ fix, int mask_fixed,
                                                  const T* __restrict__ b, double* __restrict__ part) {
#pragma clang fp contract(fast)  // not a check-path kernel: FMA allowed
    __shared__ double lds[32];
    double acc[1] = {0.0};
    const long nn = g.nnode();
    for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < nn; n += (long)gridDim.x * blockDim.x) {
        int i, j, k;
        node_ijk(g, n, i, j, k);
        T v[3][3][3][3];
        gather_h8(g, i, j, k, x, v);
        T yv[3];
        apply_h8_node(g, i, j, k, v, E, K0, yv);
        uint8_t f = fix[n];
        for (int d = 0; d < 3; d++) {
            T o = yv[d];
            if (b) o = b[3 * n + d] - o;
            if ((mask_fixed || b) && (f & (1u << d))) o = 0;
            y[3 * n + d] = o;
            acc[0] += (double)o * (double)o;
        }
    }
    if (part) store_partials<1>(acc, lds, part);
}

template <class T>
__global__ void k_diag_h8(G3 g, const T* __restrict__ E, const T* __restrict__ K0, const uint8_t* __restrict__ fix,
                          T* __restrict__ D) {
    const long nn = g.nnode();
    for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < nn; n += (long)gridDim.x * blockDim.x) {
        int i, j, k;
        node_ijk(g, n, i, j, k);
        T d[3] = {0, 0, 0};
        for (int ea = 0; ea < 2; ea++)
            for (int eb = 0; eb < 2; eb++)
                for (int ec = 0; ec < 2; ec++) {
                    int ei = i - 1 + ea, ej = j - 1 + eb, ek = k - 1 + ec;
                    if (ei < 0 || ei >= g.nx || ej < 0 || ej >= g.ny || ek < 0 || ek >= g.nz) continue;
                    T Ee = E[g.elem(ei, ej, ek)];
                    int me = loc_h8(i - ei, j - ej, k - ek);
                    for (int r = 0; r < 3; r++) d[r] += Ee * K0[(3 * me + r) * 24 + 3 * me + r];
                }
        uint8_t f = fix[n];
        for (int r = 0; r < 3; r++) D[3 * n + r] = (f & (1u << r)) ? (T)1 : d[r];
    }
}

// Hex8 PCG is three launches per iteration: K3 (beta from K2's partials,
// convergence test, p = beta p + z), K1 (q = K p, partial p.q), K2 (shared
// with quad4: x, r, z update and the z.r, r.r partials).
template <class T>
__global__ void __launch_bounds__(256) k_pcg_k3(long n, int k, T* __restrict__ p, const T* __restrict__ z,
                                                const double* __restrict__ partB, int nbB, PcgDev* __restrict__ S,
                                                double eps) {
#pragma clang fp contract(fast)  // not a check-path kernel: FMA allowed
    __shared__ double lds[32];
    if (S->done) return;
    double beta;
    bool stop;
    pcg_head(k, partB, nbB, S, eps, lds, beta, stop);
    if (stop) return;
    const T tb = (T)beta;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        p[i] = tb * p[i] + z[i];  // xeaxpy (CG.h:42-49)
}

template <class T>
__global__ void __launch_bounds__(256) k_pcg_k1_h8(G3 g, const T* __restrict__ p, T* __restrict__ q,
                                                   const T* __restrict__ E, const T* __restrict__ K0,
                                                   const uint8_t* __restrict__ fix, double* __restrict__ partA,
                                                   const PcgDev* __restrict__ S) {
#pragma clang fp contract(fast)  // not a check-path kernel: FMA allowed
    __shared__ double lds[32];
    if (S->done) return;
    double acc[1] = {0.0};
    const long nn = g.nnode();
    for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < nn; n += (long)gridDim.x * blockDim.x) {
        int i, j, k;
        node_ijk(g, n, i, j, k);
        T v[3][3][3][3];
        gather_h8(g, i, j, k, p, v);
        T yv[3];
        apply_h8_node(g, i, j, k, v, E, K0, yv);
        uint8_t f = fix[n];
        for (int d = 0; d < 3; d++) {
            T o = (f & (1u << d)) ? (T)0 : yv[d];
            q[3 * n + d] = o;
            acc[0] += (double)v[1][1][1][d] * (double)o;
        }
    }
    store_partials<1>(acc, lds, partA);
}

// ---------------------------------------------------- filter / compliance --
__global__ void k_centroids_h8(G3 g, double lx, double ly, double lz, double* __restrict__ cg) {
    const long ne = g.nelem();
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < ne; e += (long)gridDim.x * blockDim.x) {
        int i, j, k;
        elem_ijk(g, e, i, j, k);
        double X[2] = {lx * (i / (double)g.nx), lx * ((i + 1) / (double)g.nx)};
        double Y[2] = {ly * (j / (double)g.ny), ly * ((j + 1) / (double)g.ny)};
        double Z[2] = {lz * (k / (double)g.nz), lz * ((k + 1) / (double)g.nz)};
        double c[3] = {0.0, 0.0, 0.0};
        for (int l = 0; l < 8; l++) {
            const int dk = l >> 2, dj = ((l & 3) >= 2) ? 1 : 0, di = ((l & 3) == 1 || (l & 3) == 2) ? 1 : 0;
            c[0] += X[di];
            c[1] += Y[dj];
            c[2] += Z[dk];
        }
        cg[3 * e] = c[0] / 8.0;
        cg[3 * e + 1] = c[1] / 8.0;
        cg[3 * e + 2] = c[2] / 8.0;
    }
}

__device__ inline double filter_one_h8(const G3& g, int win, double R, const double* __restrict__ cg,
                                       const double* __restrict__ in, long e) {
    int i, j, k;
    elem_ijk(g, e, i, j, k);
    double acc = 0.0, wsum = 0.0;
    for (int ii = max(i - win, 0); ii <= min(i + win, g.nx - 1); ii++)
        for (int jj = max(j - win, 0); jj <= min(j + win, g.ny - 1); jj++)
            for (int kk = max(k - win, 0); kk <= min(k + win, g.nz - 1); kk++) {
                long m = g.elem(ii, jj, kk);
                double v0 = cg[3 * e] - cg[3 * m], v1 = cg[3 * e + 1] - cg[3 * m + 1], v2 = cg[3 * e + 2] - cg[3 * m + 2];
                double s = 0.0;
                s += v0 * v0;
                s += v1 * v1;
                s += v2 * v2;
                double d = sqrt(s);
                if (d <= R) {
                    double w = (R - d) / R;
                    acc += w * in[m];
                    wsum += w;
                }
            }
    return acc / wsum;
}

__global__ void __launch_bounds__(256) k_filter_h8(G3 g, int win, double R, const double* __restrict__ cg,
                                                   const double* __restrict__ in, double* __restrict__ out) {
    const long ne = g.nelem();
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < ne; e += (long)gridDim.x * blockDim.x)
        out[e] = filter_one_h8(g, win, R, cg, in, e);
}

__global__ void __launch_bounds__(256) k_oc_volume_h8(G3 g, int win, double R, const double* __restrict__ cg,
                                                      const double* __restrict__ xplus, double vf, double scale1,
                                                      double beta, double* __restrict__ part,
                                                      const OcDev* __restrict__ S) {
    __shared__ double lds[32];
    if (S->done) return;
    const long ne = g.nelem();
    const double den = vf * (double)ne;
    double acc[1] = {0.0};
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < ne; e += (long)gridDim.x * blockDim.x) {
        double v = filter_one_h8(g, win, R, cg, xplus, e);
        if (beta > 0.0) v = heaviside_rho(v, beta);
        acc[0] += scale1 * v / den;
    }
    store_partials<1>(acc, lds, part);
}

__global__ void __launch_bounds__(256) k_comp_sens_h8(G3 g, const double* __restrict__ u, const double* __restrict__ E,
                                                      const double* __restrict__ rho, const double* __restrict__ K0,
                                                      double coef, double pm1, double* __restrict__ dfdrho,
                                                      double* __restrict__ part) {
    __shared__ double lds[32];
    const long ne = g.nelem();
    double acc[1] = {0.0};
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < ne; e += (long)gridDim.x * blockDim.x) {
        int i, j, k;
        elem_ijk(g, e, i, j, k);
        double ue[24];
        for (int l = 0; l < 8; l++) {
            const int dk = l >> 2, dj = ((l & 3) >= 2) ? 1 : 0, di = ((l & 3) == 1 || (l & 3) == 2) ? 1 : 0;
            long m = g.node(i + di, j + dj, k + dk);
            ue[3 * l] = u[3 * m];
            ue[3 * l + 1] = u[3 * m + 1];
            ue[3 * l + 2] = u[3 * m + 2];
        }
        double c = 0.0;
        for (int a = 0; a < 24; a++) {
            double ka = 0.0;
            for (int b = 0; b < 24; b++) ka += K0[a * 24 + b] * ue[b];
            c += ue[a] * ka;
        }
        dfdrho[e] = coef * simp_pow(rho[e], pm1) * c;
        acc[0] += E[e] * c;
    }
    store_partials<1>(acc, lds, part);
}

// ------------------------------------------------------- element matrix ---
// SolidLinearIsotropicElastic<T, ShapeFunction8Cubic, Gauss8Cubic>
// (src/FEM/Equation/Solid.h:19-64) in the reference's operation order.
__device__ inline double det3_ref(const double* v) {
    return -v[8] * v[1] * v[3] - v[7] * v[5] * v[0] - v[2] * v[4] * v[6] + v[6] * v[1] * v[5] + v[7] * v[3] * v[2] +
           v[0] * v[4] * v[8];
}

__device__ void ke_h8_ref(const double* X, double E, double nu, double* Ke) {
    const double a = 1.0 / sqrt(3.0), m = -1.0 / sqrt(3.0);
    const double gp[8][3] = {{m, m, m}, {a, m, m}, {a, a, m}, {m, a, m}, {m, m, a}, {a, m, a}, {a, a, a}, {m, a, a}};
    double C[36];
    for (int q = 0; q < 36; q++) C[q] = 0.0;
    C[0] = 1.0 - nu; C[1] = nu; C[2] = nu;
    C[6] = nu; C[7] = 1.0 - nu; C[8] = nu;
    C[12] = nu; C[13] = nu; C[14] = 1.0 - nu;
    C[21] = 0.5 * (1.0 - 2.0 * nu); C[28] = 0.5 * (1.0 - 2.0 * nu); C[35] = 0.5 * (1.0 - 2.0 * nu);
    double c = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
    for (int q = 0; q < 36; q++) C[q] *= c;
    for (int q = 0; q < 576; q++) Ke[q] = 0.0;
    for (int g = 0; g < 8; g++) {
        double r0 = gp[g][0], r1 = gp[g][1], r2 = gp[g][2];
        double dN[24] = {-0.125 * (1.0 - r1) * (1.0 - r2), 0.125 * (1.0 - r1) * (1.0 - r2),
                         0.125 * (1.0 + r1) * (1.0 - r2), -0.125 * (1.0 + r1) * (1.0 - r2),
                         -0.125 * (1.0 - r1) * (1.0 + r2), 0.125 * (1.0 - r1) * (1.0 + r2),
                         0.125 * (1.0 + r1) * (1.0 + r2), -0.125 * (1.0 + r1) * (1.0 + r2),
                         -0.125 * (1.0 - r2) * (1.0 - r0), -0.125 * (1.0 - r2) * (1.0 + r0),
                         0.125 * (1.0 - r2) * (1.0 + r0), 0.125 * (1.0 - r2) * (1.0 - r0),
                         -0.125 * (1.0 + r2) * (1.0 - r0), -0.125 * (1.0 + r2) * (1.0 + r0),
                         0.125 * (1.0 + r2) * (1.0 + r0), 0.125 * (1.0 + r2) * (1.0 - r0),
                         -0.125 * (1.0 - r0) * (1.0 - r1), -0.125 * (1.0 + r0) * (1.0 - r1),
                         -0.125 * (1.0 + r0) * (1.0 + r1), -0.125 * (1.0 - r0) * (1.0 + r1),
                         0.125 * (1.0 - r0) * (1.0 - r1), 0.125 * (1.0 + r0) * (1.0 - r1),
                         0.125 * (1.0 + r0) * (1.0 + r1), 0.125 * (1.0 - r0) * (1.0 + r1)};
        double J3[9], inv[9], dNdX[24];
        mm_dev(dN, X, J3, 3, 8, 3);
        double det = det3_ref(J3);
        for (int ii = 0; ii < 3; ii++)
            for (int jj = 0; jj < 3; jj++) {
                double cf[4];
                int qq = 0;
                for (int r = 0; r < 3; r++) {
                    if (r == jj) continue;
                    for (int s = 0; s < 3; s++) {
                        if (s == ii) continue;
                        cf[qq++] = J3[r * 3 + s];
                    }
                }
                double v = cf[0] * cf[3] - cf[1] * cf[2];
                inv[ii * 3 + jj] = ((ii + jj) % 2 == 0) ? v : -v;
            }
        for (int q = 0; q < 9; q++) inv[q] = inv[q] / det;
        mm_dev(inv, dN, dNdX, 3, 3, 8);
        // B^T C B accumulated column block by column block: BtC (24x6) then
        // (BtC) B (24x24), each entry summed from 0 over the inner index
        double B[144];
        for (int q = 0; q < 144; q++) B[q] = 0.0;
        for (int nd = 0; nd < 8; nd++) {
            double dx = dNdX[nd], dy = dNdX[8 + nd], dz = dNdX[16 + nd];
            B[0 * 24 + 3 * nd] = dx;
            B[1 * 24 + 3 * nd + 1] = dy;
            B[2 * 24 + 3 * nd + 2] = dz;
            B[3 * 24 + 3 * nd] = dy; B[3 * 24 + 3 * nd + 1] = dx;
            B[4 * 24 + 3 * nd + 1] = dz; B[4 * 24 + 3 * nd + 2] = dy;
            B[5 * 24 + 3 * nd] = dz; B[5 * 24 + 3 * nd + 2] = dx;
        }
        double BtC[144];
        for (int r = 0; r < 24; r++)
            for (int s = 0; s < 6; s++) {
                double v = 0.0;
                for (int q = 0; q < 6; q++) v += B[q * 24 + r] * C[q * 6 + s];
                BtC[r * 6 + s] = v;
            }
        for (int r = 0; r < 24; r++)
            for (int s = 0; s < 24; s++) {
                double v = 0.0;
                for (int q = 0; q < 6; q++) v += BtC[r * 6 + q] * B[q * 24 + s];
                Ke[r * 24 + s] += v * det * 1.0 * 1.0 * 1.0;
            }
    }
}

__device__ inline void h8_coords(const G3& g, double lx, double ly, double lz, int i, int j, int k, double* X) {
    double xs[2] = {lx * (i / (double)g.nx), lx * ((i + 1) / (double)g.nx)};
    double ys[2] = {ly * (j / (double)g.ny), ly * ((j + 1) / (double)g.ny)};
    double zs[2] = {lz * (k / (double)g.nz), lz * ((k + 1) / (double)g.nz)};
    for (int l = 0; l < 8; l++) {
        const int dk = l >> 2, dj = ((l & 3) >= 2) ? 1 : 0, di = ((l & 3) == 1 || (l & 3) == 2) ? 1 : 0;
        X[3 * l] = xs[di];
        X[3 * l + 1] = ys[dj];
        X[3 * l + 2] = zs[dk];
    }
}

__global__ void k_ke_one_h8(G3 g, double lx, double ly, double lz, int e, double E, double nu,
                            double* __restrict__ Ke) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int i, j, k;
    elem_ijk(g, e, i, j, k);
    double X[24];
    h8_coords(g, lx, ly, lz, i, j, k, X);
    ke_h8_ref(X, E, nu, Ke);
}

__global__ void k_ke_all_h8(G3 g, double lx, double ly, double lz, const double* __restrict__ E, double nu,
                            double* __restrict__ ke_all) {
    const long ne = g.nelem();
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < ne; e += (long)gridDim.x * blockDim.x) {
        int i, j, k;
        elem_ijk(g, e, i, j, k);
        double X[24];
        h8_coords(g, lx, ly, lz, i, j, k, X);
        ke_h8_ref(X, E[e], nu, ke_all + e * 576);
    }
}

// ----------------------------------------------------------------- CSR ----
__global__ void k_csr_count_h8(G3 g, const int32_t* __restrict__ n2g, const uint8_t* __restrict__ fix,
                               int32_t* __restrict__ rowlen) {
    const long nn = g.nnode();
    for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < nn; n += (long)gridDim.x * blockDim.x) {
        int i, j, k;
        node_ijk(g, n, i, j, k);
        int cnt = 0;
        for (int ii = i - 1; ii <= i + 1; ii++)
            for (int jj = j - 1; jj <= j + 1; jj++)
                for (int kk = k - 1; kk <= k + 1; kk++) {
                    if (ii < 0 || ii > g.nx || jj < 0 || jj > g.ny || kk < 0 || kk > g.nz) continue;
                    uint8_t f = fix[g.node(ii, jj, kk)];
                    cnt += !(f & 1) + !(f & 2) + !(f & 4);
                }
        for (int d = 0; d < 3; d++) {
            int r = n2g[3 * n + d];
            if (r >= 0) rowlen[r] = cnt;
        }
    }
}

__global__ void k_csr_fill_h8(G3 g, const int32_t* __restrict__ n2g, const int32_t* __restrict__ indptr,
                              const double* __restrict__ ke_all, int32_t* __restrict__ indices,
                              double* __restrict__ data) {
    const long nn = g.nnode();
    for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < nn; n += (long)gridDim.x * blockDim.x) {
        int i, j, k;
        node_ijk(g, n, i, j, k);
        for (int d = 0; d < 3; d++) {
            int row = n2g[3 * n + d];
            if (row < 0) continue;
            int pos = indptr[row];
            for (int ii = i - 1; ii <= i + 1; ii++)
                for (int jj = j - 1; jj <= j + 1; jj++)
                    for (int kk = k - 1; kk <= k + 1; kk++) {
                        if (ii < 0 || ii > g.nx || jj < 0 || jj > g.ny || kk < 0 || kk > g.nz) continue;
                        long m = g.node(ii, jj, kk);
                        for (int d2 = 0; d2 < 3; d2++) {
                            int col = n2g[3 * m + d2];
                            if (col < 0) continue;
                            double v = 0.0;
                            int ei0 = max(max(i, ii) - 1, 0), ei1 = min(min(i, ii), g.nx - 1);
                            int ej0 = max(max(j, jj) - 1, 0), ej1 = min(min(j, jj), g.ny - 1);
                            int ek0 = max(max(k, kk) - 1, 0), ek1 = min(min(k, kk), g.nz - 1);
                            for (int ei = ei0; ei <= ei1; ei++)
                                for (int ej = ej0; ej <= ej1; ej++)
                                    for (int ek = ek0; ek <= ek1; ek++) {
                                        const double* Ke = ke_all + g.elem(ei, ej, ek) * 576;
                                        int a = 3 * loc_h8(i - ei, j - ej, k - ek) + d;
                                        int b = 3 * loc_h8(ii - ei, jj - ej, kk - ek) + d2;
                                        v = v + Ke[a * 24 + b];
                                    }
                            indices[pos] = col;
                            data[pos] = v;
                            pos++;
                        }
                    }
        }
    }
}

}  // namespace pf
