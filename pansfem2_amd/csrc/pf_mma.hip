// MMA (Svanberg) and CONLIN for one constraint on the device: the O(n) passes of
// MMA<T>::UpdateVariables (src/Optimize/Solver/MMA.h:116-419) as kernels
// with fixed-order workgroup partials; the O(1) scalar algebra of the
// primal-dual interior-point iteration (the (m+1)^2 = 2x2 Newton system
// solved by solvels, MMA.h:464-509) runs on the host between launches.
// Included by pf_capi.hip (same translation unit, no FP contraction).
#pragma once

namespace pf {

// Asymptotes (MMA.h:118-142), move limits (145-150), p0, q0, p, q and the
// b-sum terms (153-175).  xmin/xmax are uniform.
__global__ void __launch_bounds__(256) k_mma_prep(long n, int early, const double* __restrict__ xk,
                                                  const double* __restrict__ xkm1, const double* __restrict__ xkm2,
                                                  double* __restrict__ L, double* __restrict__ U, double xmin,
                                                  double xmax, double asyinit, double asydecr, double asyincr,
                                                  double albefa, double move, double raa0,
                                                  const double* __restrict__ dfdx, const double* __restrict__ dgdx,
                                                  double* __restrict__ alpha, double* __restrict__ beta,
                                                  double* __restrict__ p0, double* __restrict__ q0,
                                                  double* __restrict__ p, double* __restrict__ q,
                                                  double* __restrict__ part) {
    __shared__ double lds[32];
    double acc[1] = {0.0};
    const double dx = xmax - xmin;
    for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
        double x = xk[j], l, u;
        if (early) {
            l = x - asyinit * dx;
            u = x + asyinit * dx;
        } else {
            double tmp = (x - xkm1[j]) * (xkm1[j] - xkm2[j]);
            double lo = L[j], up = U[j], x1 = xkm1[j];
            if (tmp < 0.0) {
                l = x - asydecr * (x1 - lo);
                u = x + asydecr * (up - x1);
            } else if (tmp > 0.0) {
                l = x - asyincr * (x1 - lo);
                u = x + asyincr * (up - x1);
            } else {
                l = x - (x1 - lo);
                u = x + (up - x1);
            }
        }
        // std::min(std::max(a, b), c)
        double a1 = x - 10.0 * dx, c1 = x - 0.01 * dx;
        double m1 = (a1 < l) ? l : a1;
        l = (c1 < m1) ? c1 : m1;
        double a2 = x + 0.01 * dx, c2 = x + 10.0 * dx;
        double m2 = (a2 < u) ? u : a2;
        u = (c2 < m2) ? c2 : m2;
        L[j] = l;
        U[j] = u;
        // std::max({xmin, L + albefa (x - L), x - move dx}) / std::min({...})
        double al = xmin, t1 = l + albefa * (x - l), t2 = x - move * dx;
        if (al < t1) al = t1;
        if (al < t2) al = t2;
        double be = xmax, s1 = u - albefa * (u - x), s2 = x + move * dx;
        if (s1 < be) be = s1;
        if (s2 < be) be = s2;
        alpha[j] = al;
        beta[j] = be;
        double ux = u - x, xl = x - l;
        double df = dfdx[j];
        double dp = (df < 0.0) ? 0.0 : df, dm = (-df < 0.0) ? 0.0 : -df;  // std::max(v, T())
        p0[j] = (ux * ux) * (1.001 * dp + 0.001 * dm + raa0 / dx);
        q0[j] = (xl * xl) * (0.001 * dp + 1.001 * dm + raa0 / dx);
        double dg = dgdx[j];
        double gp = (dg < 0.0) ? 0.0 : dg, gm = (-dg < 0.0) ? 0.0 : -dg;
        double pj = (ux * ux) * (1.001 * gp + 0.001 * gm + raa0 / dx);
        double qj = (xl * xl) * (0.001 * gp + 1.001 * gm + raa0 / dx);
        p[j] = pj;
        q[j] = qj;
        acc[0] += pj / ux + qj / xl;
    }
    store_partials<1>(acc, lds, part);
}

// Interior-point start (MMA.h:192-196).
__global__ void k_mma_start(long n, const double* __restrict__ alpha, const double* __restrict__ beta,
                            double* __restrict__ x, double* __restrict__ gsi, double* __restrict__ ita) {
    for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
        double xv = 0.5 * (alpha[j] + beta[j]);
        double g = 1.0 / (xv - alpha[j]), t = 1.0 / (beta[j] - xv);
        x[j] = xv;
        gsi[j] = (1.0 < g) ? g : 1.0;
        ita[j] = (1.0 < t) ? t : 1.0;
    }
}

// KKT residual terms of equations (5.9a,e,f) at one point (MMA.h:439-443).
__device__ inline double mma_kkt_terms(double x, double gs, double it, double lam, double eps, double al, double be,
                                       double l, double u, double p0, double q0, double p, double q, double& gsum) {
    double pl = p0 + lam * p, ql = q0 + lam * q;
    double ux = u - x, xl = x - l;
    gsum = p / ux + q / xl;
    double a = pl / (ux * ux) - ql / (xl * xl) - gs + it;
    double e = gs * (x - al) - eps;
    double f = it * (be - x) - eps;
    double s = 0.0;
    s += a * a;
    s += e * e;
    s += f * f;
    return s;
}

// Newton coefficients (MMA.h:201-222) and the n-sums of the 2x2 system
// (MMA.h:260-278): partials of sum(p/(U-x) + q/(x-L)), sum(G^2/Dx),
// sum(G dlx/Dx) and the KKT norm terms at the current point.
__global__ void __launch_bounds__(256) k_mma_newton_a(long n, double lam, double eps, const double* __restrict__ x,
                                                      const double* __restrict__ gsi, const double* __restrict__ ita,
                                                      const double* __restrict__ alpha, const double* __restrict__ beta,
                                                      const double* __restrict__ L, const double* __restrict__ U,
                                                      const double* __restrict__ p0, const double* __restrict__ q0,
                                                      const double* __restrict__ p, const double* __restrict__ q,
                                                      double* __restrict__ G, double* __restrict__ Dx,
                                                      double* __restrict__ dlx, double* __restrict__ part) {
    __shared__ double lds[64];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
        double xv = x[j], l = L[j], u = U[j], al = alpha[j], be = beta[j];
        double pl = p0[j] + lam * p[j], ql = q0[j] + lam * q[j];
        double ux = u - xv, xl = xv - l;
        double ux2 = ux * ux, xl2 = xl * xl;
        double g = p[j] / ux2 - q[j] / xl2;
        double d = 2.0 * pl / cube_rn(ux) + 2.0 * ql / cube_rn(xl) + gsi[j] / (xv - al) + ita[j] / (be - xv);
        double dl = pl / ux2 - ql / xl2 - eps / (xv - al) + eps / (be - xv);
        G[j] = g;
        Dx[j] = d;
        dlx[j] = dl;
        double gs;
        double kt = mma_kkt_terms(xv, gsi[j], ita[j], lam, eps, al, be, l, u, p0[j], q0[j], p[j], q[j], gs);
        acc[0] += gs;
        acc[1] += g * g / d;
        acc[2] += g * dl / d;
        acc[3] += kt;
    }
    store_partials<4>(acc, lds, part);
}

// Newton direction of the primal variables (MMA.h:283-290, 303-306) and the
// step bound over j (MMA.h:347-352): partial maxima.
__global__ void __launch_bounds__(256) k_mma_newton_b(long n, double dlam, double eps, const double* __restrict__ x,
                                                      const double* __restrict__ gsi, const double* __restrict__ ita,
                                                      const double* __restrict__ alpha, const double* __restrict__ beta,
                                                      const double* __restrict__ G, const double* __restrict__ Dx,
                                                      const double* __restrict__ dlx, double* __restrict__ dx,
                                                      double* __restrict__ dgsi, double* __restrict__ dita,
                                                      double* __restrict__ part) {
    __shared__ double lds[16];
    double mx = 0.0;  // txmax starts at T()
    for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
        double xv = x[j], al = alpha[j], be = beta[j], d = Dx[j];
        double v = -dlx[j] / d;
        v -= G[j] * dlam / d;
        double gs = gsi[j], it = ita[j];
        double dg = -gs * v / (xv - al) - gs + eps / (xv - al);
        double di = it * v / (be - xv) - it + eps / (be - xv);
        dx[j] = v;
        dgsi[j] = dg;
        dita[j] = di;
        double t = -1.01 * v / (xv - al);
        double t2 = 1.01 * v / (be - xv);
        double t3 = -1.01 * dg / gs;
        double t4 = -1.01 * di / it;
        if (t < t2) t = t2;
        if (t < t3) t = t3;
        if (t < t4) t = t4;
        if (mx < t) mx = t;
    }
    // block max (order-free)
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        double o = __shfl_xor(mx, off, 64);
        mx = (mx < o) ? o : mx;
    }
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = lds[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = (m < lds[w]) ? lds[w] : m;
        part[blockIdx.x] = m;
    }
}

// Trial point of the line search (MMA.h:372-385): writes x + tau dx etc. and
// the partial KKT sums at that point.
__global__ void __launch_bounds__(256) k_mma_trial(long n, double tau, double lam, double eps,
                                                   const double* __restrict__ x, const double* __restrict__ gsi,
                                                   const double* __restrict__ ita, const double* __restrict__ dx,
                                                   const double* __restrict__ dgsi, const double* __restrict__ dita,
                                                   const double* __restrict__ alpha, const double* __restrict__ beta,
                                                   const double* __restrict__ L, const double* __restrict__ U,
                                                   const double* __restrict__ p0, const double* __restrict__ q0,
                                                   const double* __restrict__ p, const double* __restrict__ q,
                                                   double* __restrict__ xp, double* __restrict__ gp,
                                                   double* __restrict__ ip, double* __restrict__ part) {
    __shared__ double lds[32];
    double acc[2] = {0.0, 0.0};
    for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
        double xv = x[j] + tau * dx[j];
        double gv = gsi[j] + tau * dgsi[j];
        double iv = ita[j] + tau * dita[j];
        xp[j] = xv;
        gp[j] = gv;
        ip[j] = iv;
        double gs;
        acc[0] += mma_kkt_terms(xv, gv, iv, lam, eps, alpha[j], beta[j], L[j], U[j], p0[j], q0[j], p[j], q[j], gs);
        acc[1] += gs;
    }
    store_partials<2>(acc, lds, part);
}

// ------------------------------------------------------------- CONLIN ----
// CONLIN<T> (src/Optimize/Solver/CONLIN.h:90-365) shares the interior-point
// driver; only the approximation differs: p x + q / x with p, q the positive
// and negative parts of the derivative (q scaled by x_k^2), move limits
// without asymptotes.

// Move limits (CONLIN.h:92-96), p0, q0 (99-105), p, q and the b-sum terms
// (108-120).
__global__ void __launch_bounds__(256) k_conlin_prep(long n, const double* __restrict__ xk, double xmin, double xmax,
                                                     double move, const double* __restrict__ dfdx,
                                                     const double* __restrict__ dgdx, double* __restrict__ alpha,
                                                     double* __restrict__ beta, double* __restrict__ p0,
                                                     double* __restrict__ q0, double* __restrict__ p,
                                                     double* __restrict__ q, double* __restrict__ part) {
    __shared__ double lds[32];
    double acc[1] = {0.0};
    const double dx = xmax - xmin;
    for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
        const double x = xk[j];
        // std::max({xmin, x - move dx}), std::min({xmax, x + move dx})
        double al = xmin, t1 = x - move * dx;
        if (al < t1) al = t1;
        double be = xmax, t2 = x + move * dx;
        if (t2 < be) be = t2;
        alpha[j] = al;
        beta[j] = be;
        const double x2 = x * x;  // pow(x, 2.0)
        const double df = dfdx[j], dg = dgdx[j];
        p0[j] = (df < 0.0) ? 0.0 : df;
        q0[j] = ((-df < 0.0) ? 0.0 : -df) * x2;
        const double pj = (dg < 0.0) ? 0.0 : dg;
        const double qj = ((-dg < 0.0) ? 0.0 : -dg) * x2;
        p[j] = pj;
        q[j] = qj;
        acc[0] += pj * x + qj / x;
    }
    store_partials<1>(acc, lds, part);
}

// KKT residual terms of equations (5.9a,e,f) (CONLIN.h:392-396).
__device__ inline double conlin_kkt_terms(double x, double gs, double it, double lam, double eps, double al,
                                          double be, double p0, double q0, double p, double q, double& gsum) {
    const double pl = p0 + lam * p, ql = q0 + lam * q;
    gsum = p * x + q / x;
    const double a = pl - ql / (x * x) - gs + it;
    const double e = gs * (x - al) - eps;
    const double f = it * (be - x) - eps;
    double s = 0.0;
    s += a * a;
    s += e * e;
    s += f * f;
    return s;
}

// Newton coefficients (CONLIN.h:143-170) and the n-sums of the 2x2 system,
// as k_mma_newton_a.
__global__ void __launch_bounds__(256) k_conlin_newton_a(long n, double lam, double eps, const double* __restrict__ x,
                                                         const double* __restrict__ gsi, const double* __restrict__ ita,
                                                         const double* __restrict__ alpha,
                                                         const double* __restrict__ beta, const double* __restrict__ p0,
                                                         const double* __restrict__ q0, const double* __restrict__ p,
                                                         const double* __restrict__ q, double* __restrict__ G,
                                                         double* __restrict__ Dx, double* __restrict__ dlx,
                                                         double* __restrict__ part) {
    __shared__ double lds[64];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
        const double xv = x[j], al = alpha[j], be = beta[j];
        const double pl = p0[j] + lam * p[j], ql = q0[j] + lam * q[j];
        const double x2 = xv * xv;
        const double g = p[j] - q[j] / x2;
        const double d = 2.0 * ql / cube_rn(xv) + gsi[j] / (xv - al) + ita[j] / (be - xv);
        const double dl = pl - ql / x2 - eps / (xv - al) + eps / (be - xv);
        G[j] = g;
        Dx[j] = d;
        dlx[j] = dl;
        double gs;
        const double kt = conlin_kkt_terms(xv, gsi[j], ita[j], lam, eps, al, be, p0[j], q0[j], p[j], q[j], gs);
        acc[0] += gs;
        acc[1] += g * g / d;
        acc[2] += g * dl / d;
        acc[3] += kt;
    }
    store_partials<4>(acc, lds, part);
}

// Trial point of the line search (CONLIN.h:318-340), as k_mma_trial.
__global__ void __launch_bounds__(256) k_conlin_trial(long n, double tau, double lam, double eps,
                                                      const double* __restrict__ x, const double* __restrict__ gsi,
                                                      const double* __restrict__ ita, const double* __restrict__ dx,
                                                      const double* __restrict__ dgsi, const double* __restrict__ dita,
                                                      const double* __restrict__ alpha, const double* __restrict__ beta,
                                                      const double* __restrict__ p0, const double* __restrict__ q0,
                                                      const double* __restrict__ p, const double* __restrict__ q,
                                                      double* __restrict__ xp, double* __restrict__ gp,
                                                      double* __restrict__ ip, double* __restrict__ part) {
    __shared__ double lds[32];
    double acc[2] = {0.0, 0.0};
    for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
        const double xv = x[j] + tau * dx[j];
        const double gv = gsi[j] + tau * dgsi[j];
        const double iv = ita[j] + tau * dita[j];
        xp[j] = xv;
        gp[j] = gv;
        ip[j] = iv;
        double gs;
        acc[0] += conlin_kkt_terms(xv, gv, iv, lam, eps, alpha[j], beta[j], p0[j], q0[j], p[j], q[j], gs);
        acc[1] += gs;
    }
    store_partials<2>(acc, lds, part);
}

}  // namespace pf
