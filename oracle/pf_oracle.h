/* CPU oracle for the PANSFEM2 SIMP topology-optimisation path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker that tests/, the smoke entry
 * and the CPU-baseline leg of bench.py compare the HIP path against.  Nothing
 * in the product (libpansfem2_amd.so, include/pansfem2/) links or calls it.
 *
 * Every routine restates one reference routine operation for operation (same
 * loop order, same association of floating-point operations, no FMA
 * contraction: the library is built with -ffp-contract=off), so that its
 * results are bit-identical to the reference compiled with g++ -O3 -fopenmp
 * on x86-64.  The reference file:line each routine follows is given beside it;
 * tests/test_oracle.py checks each against the golden vectors written by the
 * reference itself (oracle/_ref) and against closed forms.
 */
#ifndef PF_ORACLE_H
#define PF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* SquareMesh<T>::GenerateNodes (src/PrePost/Mesher/SquareMesh.h:60-70):
 * node (ny+1)*i+j at {lx*(i/(double)nx), ly*(j/(double)ny)}; xy is nn x 2. */
void orc_grid2d_nodes(int nx, int ny, double lx, double ly, double* xy);
/* SquareMesh<T>::GenerateElements (SquareMesh.h:97-106): element ny*i+j. */
void orc_grid2d_elements(int nx, int ny, int* conn);

/* SetDirichlet + Renumbering (src/FEM/Controller/BoundaryCondition.h:19-25,
 * src/FEM/Controller/Assembling.h:175-186).  n2g is nn x ndpn; fixed entries
 * become -1, free ones are numbered in (node, dof) order.  Returns N. */
int orc_dofmap(int nn, int ndpn, const int* fixed_node, const int* fixed_dof, int nfixed, int* n2g);

/* PlaneStrainStiffness<T, ShapeFunction4Square, Gauss4Square>
 * (src/FEM/Equation/PlaneStrain.h:19-58).  X is 4 x 2 node coordinates. */
void orc_ke_planestrain_q4(const double* X, double E, double nu, double t, double* Ke);
/* SolidLinearIsotropicElastic<T, ShapeFunction8Cubic, Gauss8Cubic>
 * (src/FEM/Equation/Solid.h:19-64).  X is 8 x 3 node coordinates. */
void orc_ke_solid_h8(const double* X, double E, double nu, double* Ke);

/* Assembling(LILCSR&, F, u, Ke, ...) over all elements in order, then
 * CSR(LILCSR&) (Assembling.h:46-66, LILCSR.h:92-114, CSR.h:92-105).
 * conn: ne x nen; n2g: nn x ndpn; ke_all: ne x nd x nd (nd = nen*ndpn);
 * ufix: nn x ndpn prescribed values (Dirichlet columns move to F).
 * Arrays are malloc'd; release them with orc_free. Returns nnz. */
long orc_assemble_csr(int N, int ne, int nen, int ndpn, const int* conn, const int* n2g,
                      const double* ke_all, const double* ufix, double* F,
                      int** indptr, int** indices, double** data);
void orc_free(void* p);

/* CSR<T>::operator* (src/LinearAlgebra/Models/CSR.h:108-122). */
void orc_csr_spmv(int N, const int* indptr, const int* indices, const double* data, const double* x, double* y);
/* GetDiagonal via CSR::get (src/LinearAlgebra/Solvers/CG.h:397-404, CSR.h:154-167). */
void orc_csr_diag(int N, const int* indptr, const int* indices, const double* data, double* D);
/* ScalingCG (CG.h:419-453).  Returns the number of loop iterations run
 * (the index k+1 of the converging iteration) or -1 when itrmax is reached
 * (the last iterate is still returned in x, as CG.h:451-452 does). */
int orc_scaling_cg(int N, const int* indptr, const int* indices, const double* data,
                   const double* b, int itrmax, double eps, double* x);

/* CenterOfGravity (src/FEM/Equation/General.h:70-77). */
void orc_centroids(int ne, int nen, int dim, const int* conn, const double* nodes, double* cg);
/* Neighbour lists and weights of the density filter as the sample driver
 * builds them (sample/optimize/sample_optimize_density_oc.cpp:50-60):
 * ascending j with |cg_i - cg_j| <= R, w = (R - |.|)/R.  brute = 1 tests all
 * pairs; brute = 0 tests a window of +-(ceil(R)+1) elements on an nx x ny grid
 * (same candidates that can pass, same order).  ptr has ne+1 entries; idx/w
 * are malloc'd (orc_free). Returns the total entry count. */
long orc_filter_build(int nx, int ny, const double* cg, double R, int brute, int* ptr, int** idx, double** w);
/* All-pairs neighbour search for any dimension (dim = 2 or 3), cg is ne x dim. */
long orc_filter_build_brute(int ne, int dim, const double* cg, double R, int* ptr, int** idx, double** w);
/* DensityFilter<T>::GetFilteredVariables (src/Optimize/Filter/DensityFilter.h:45-56). */
void orc_filter_apply(int ne, const int* ptr, const int* idx, const double* w, const double* s, double* rho);
/* DensityFilter<T>::GetFilteredSensitivitis (DensityFilter.h:60-71): row-normalised. */
void orc_filter_sens(int ne, const int* ptr, const int* idx, const double* w, const double* dfdrho, double* dfds);

/* OC<T>::UpdateVariables (src/Optimize/Solver/OC.h:76-107) with the
 * sample's volume closure (sample_optimize_density_oc.cpp:198-207): g(x+) =
 * sum(scale1*filter(x+)_i/(vf*ne)) - scale1.  x is updated in place.
 * Returns the number of bisection steps; *lambda_out is the last lambda
 * tested; decisions (may be NULL, capacity max_dec) receive 1 for g>0. */
int orc_oc_update(int n, double* x, const double* dfdx, const double* dgdx, double iota, double lmin,
                  double lmax, double leps, double move, const int* fptr, const int* fidx, const double* fw,
                  double vf, double scale1, double* lambda_out, int* decisions, int max_dec);

/* SensitivityFilter<T> (scheme 1) and SensitivityFilter2<T> (scheme 2)
 * (src/Optimize/Filter/SensitivityFilter.h:44-55,88-99) over the lists. */
void orc_sensitivity_filter(int ne, const int* ptr, const int* idx, const double* w, const double* s,
                            const double* dfds, int scheme, double* out);
/* HeavisideFilter<T> (src/Optimize/Filter/HeavisideFilter.h:62-99) over the
 * same neighbour lists: projected filtered variables and filtered
 * sensitivities at a given beta. */
void orc_heaviside_vars(int ne, const int* ptr, const int* idx, const double* w, double beta, const double* s,
                        double* rho);
void orc_heaviside_sens(int ne, const int* ptr, const int* idx, const double* w, double beta, const double* s,
                        const double* dfdrho, double* dfds);
/* orc_oc_update with the closure filtering through HeavisideFilter at beta
 * (beta <= 0: the density filter). */
int orc_oc_update_beta(int n, double* x, const double* dfdx, const double* dgdx, double iota, double lmin,
                       double lmax, double leps, double move, const int* fptr, const int* fidx, const double* fw,
                       double vf, double scale1, double beta, double* lambda_out, int* decisions, int max_dec);

/* MMA<T> state (src/Optimize/Solver/MMA.h:15-100). */
typedef struct orc_mma {
    int n, m, k;
    double previousvalue, epsvalue;
    double a0, raa0, albefa, move, asyinit, asydecr, asyincr;
    double *a, *c, *d;               /* m */
    double *xmin, *xmax;             /* n */
    double *xkm1, *xkm2, *L, *U;     /* n */
    int newton_steps;                /* inner iterations of the last update */
} orc_mma;
orc_mma* orc_mma_create(int n, int m, double a0, const double* a, const double* c, const double* d,
                        const double* xmin, const double* xmax);
void orc_mma_set_parameters(orc_mma* s, double raa0, double albefa, double move, double asyinit,
                            double asydecr, double asyincr, double epsvalue);
int orc_mma_is_convergence(orc_mma* s, double f);
/* MMA<T>::UpdateVariables (MMA.h:116-419): dgdx is m x n row-major. */
void orc_mma_update(orc_mma* s, double* xk, double f, const double* dfdx, const double* g, const double* dgdx);
void orc_mma_destroy(orc_mma* s);

/* CONLIN<T> state (src/Optimize/Solver/CONLIN.h:18-68) and its calls
 * (SetParameters :71-74, IsConvergence :82-87, UpdateVariables :90-365). */
typedef struct orc_conlin {
    int n, m, k;
    double previousvalue, epsvalue, move, a0;
    double *a, *c, *d, *xmin, *xmax;
    int newton_steps;
} orc_conlin;
orc_conlin* orc_conlin_create(int n, int m, double a0, const double* a, const double* c, const double* d,
                              const double* xmin, const double* xmax);
void orc_conlin_set_parameters(orc_conlin* s, double move, double epsvalue);
int orc_conlin_is_convergence(orc_conlin* s, double f);
void orc_conlin_update(orc_conlin* s, double* xk, double f, const double* dfdx, const double* g, const double* dgdx);
void orc_conlin_destroy(orc_conlin* s);

/* The SIMP loop of the sample driver (sample_optimize_density_oc.cpp:83-208,
 * sample_optimize_density_mma.cpp:88-204) on an nx x ny grid of unit quad4
 * elements with the density filter, left edge clamped and a unit downward
 * load at the bottom-right node.  Per-iteration histories (length iters) are
 * written to f_hist/g_hist/lambda_hist/cg_hist when not NULL; s (ne) holds the
 * design variables on entry (normally 0.5) and on exit. */
typedef struct orc_simp_params {
    double E1, E0, nu, t, p, vf, R, scale0, scale1;
    double cg_eps; int cg_itrmax;
    int use_mma;  /* 0: OC, 1: MMA, 2: CONLIN (with the mma_* constants it shares: a0, c, bounds, move, eps) */
    double oc_iota, oc_lmin, oc_lmax, oc_leps, oc_move;
    double mma_a0, mma_c, mma_xmin, mma_xmax, mma_raa0, mma_albefa, mma_move, mma_asyinit, mma_asydecr, mma_asyincr, mma_eps;
    /* HeavisideFilter instead of DensityFilter, with beta multiplied by
     * beta_factor at every beta_every-th iteration from k = 0 */
    int heaviside; double beta0, beta_factor; int beta_every;
} orc_simp_params;
void orc_simp_default_params(orc_simp_params* p);
int orc_simp2d(int nx, int ny, int iters, const orc_simp_params* prm, double* s, double* rho_out,
               double* f_hist, double* g_hist, double* lambda_hist, int* cg_hist, double* u0_free);

#ifdef __cplusplus
}
#endif
#endif
