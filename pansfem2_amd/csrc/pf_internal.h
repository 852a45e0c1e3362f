// Internal declarations shared by the HIP translation units of libpansfem2_amd.
//
// Data layout in HBM (structured grids only):
//   * nodes are numbered as SquareMesh does (src/PrePost/Mesher/SquareMesh.h:60-70):
//     2D node (i,j) -> (ny+1)*i + j, element (i,j) -> ny*i + j;
//     3D node (i,j,k) -> ((ny+1)*(nz+1))*i + (nz+1)*j + k, element -> ny*nz*i + nz*j + k.
//   * every solver vector is "nodal": ndpn consecutive values per node, fixed DOFs
//     held at zero, so the EbE operator and all vector kernels index without
//     indirection.  Free-DOF (reference-numbered) vectors exist only at the
//     boundary and are converted with g2n / n2g.
//   * per-node fixity is one byte, bit d set when DOF d is fixed.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pansfem2_amd.h"

namespace pf {

constexpr int kBlock = 256;        // threads per workgroup for the 1D kernels
constexpr int kMaxPartials = 8192;  // upper bound on workgroups of a reducing launch

// Scalars of one Jacobi-PCG solve, resident on the device.  The kernels of
// iteration k read what the previous kernel's last workgroup wrote.
struct PcgScalars {
    double zr;      // z.r of the current iterate (CG.h:428,440)
    double pq;      // p.Ap
    double alpha, beta;
    double rr;      // r.r
    double bnorm, eps;
    int k, itrmax, done, status;  // status: 0 running, 1 converged, 2 itrmax reached, 3 breakdown
};

// OC bisection state (OC.h:80-99), resident on the device.
struct OcScalars {
    double l0, l1, lambda, leps;
    double g;        // constraint value at the last tested lambda
    int steps, done, max_steps, pad;
};

// Reduction scratch: per-workgroup partials and the arrival counter of the
// "last workgroup reduces" pattern.  One set per context; launches that reduce
// are serialised on the context's stream.
struct ReduceScratch {
    double* partials;    // kMaxPartials * 4 doubles
    unsigned* counter;   // one counter, returned to zero by the last workgroup
    double* out;         // 16 doubles of results
};

}  // namespace pf

struct pf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // grid
    int dim = 0;
    int nx = 0, ny = 0, nz = 0;
    long nn = 0, ne = 0;
    int ndpn = 0, nd = 0;
    double lx = 0, ly = 0, lz = 0;

    // DOF map
    int N = 0;
    std::vector<int32_t> h_n2g;
    int32_t* d_n2g = nullptr;   // nn*ndpn, -1 for fixed
    int32_t* d_g2n = nullptr;   // N, nodal index of each free DOF
    uint8_t* d_fix = nullptr;   // nn, fixity bits

    // shared unit element matrix (E = 1)
    std::vector<double> h_K0;
    double* d_K0 = nullptr;
    float* d_K0f = nullptr;

    // material / design
    double E1 = 1.0, E0 = 1e-9, p = 3.0, nu = 0.3, t = 1.0;
    double* d_rho = nullptr;
    double* d_E = nullptr;
    float* d_Ef = nullptr;
    bool have_E = false;

    // nodal work vectors (nv = nn*ndpn)
    long nv = 0;
    double *d_x = nullptr, *d_r = nullptr, *d_z = nullptr, *d_q = nullptr, *d_p0 = nullptr, *d_p1 = nullptr;
    double *d_D = nullptr, *d_b = nullptr, *d_t0 = nullptr, *d_t1 = nullptr;
    float *f_x = nullptr, *f_r = nullptr, *f_z = nullptr, *f_q = nullptr, *f_p0 = nullptr, *f_p1 = nullptr;
    int pcg_form = 0;  // PF_PCG_SCALINGCG or PF_PCG_SINGLE_REDUCTION (quad4 matrix-free solves)
    // second buffers of the single-reduction solver's s = K p
    double *d_s0 = nullptr, *d_s1 = nullptr;
    float *f_s0 = nullptr, *f_s1 = nullptr;
    float *f_D = nullptr, *f_b = nullptr;

    // element work vectors (ne)
    double *d_e0 = nullptr, *d_e1 = nullptr, *d_e2 = nullptr, *d_e3 = nullptr, *d_e4 = nullptr, *d_e5 = nullptr;
    double* d_e6 = nullptr;     // projection derivative of the Heaviside filter
    double* d_cg = nullptr;     // element centroids (ne*dim)

    // point loads (nodal)
    double* d_load = nullptr;

    // assembled CSR (free-DOF numbering)
    long nnz = -1;
    int32_t *d_indptr = nullptr, *d_indices = nullptr;
    double* d_data = nullptr;
    double* d_ke_all = nullptr;  // ne*nd*nd element matrices in reference arithmetic

    // filter: the grid rule (pf_filter_setup) or caller-given neighbour
    // lists (pf_filter_set_lists, CSR-like ptr/idx/w over elements)
    double R = 0.0;
    int win = 0;
    bool have_filter = false;
    bool have_lists = false;
    int32_t *d_fptr = nullptr, *d_fidx = nullptr;
    double* d_fw = nullptr;

    // device scalars
    pf::PcgScalars* d_pcg = nullptr;
    pf::OcScalars* d_oc = nullptr;
    pf::ReduceScratch red{};
    pf::PcgScalars* h_pcg = nullptr;  // pinned mirror
    pf::OcScalars* h_oc = nullptr;    // pinned mirror
    double* h_red = nullptr;          // pinned mirror of red.out

    // MMA state
    struct MmaState* mma = nullptr;
    // learned preconditioner
    struct LearnedNet* net = nullptr;

    // multigrid base preconditioner of the flexible CG (pf_mg_setup)
    struct MgState* mg = nullptr;
};

// ---------------------------------------------------------------- errors ----
#define PF_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);                     \
            return PF_ERR_HIP;                                                                \
        }                                                                                     \
    } while (0)

#define PF_CHECK(cond, code, msg)                                                             \
    do {                                                                                      \
        if (!(cond)) {                                                                        \
            ctx->err = (msg);                                                                 \
            return (code);                                                                    \
        }                                                                                     \
    } while (0)

#define PF_TRY(call)                                                                          \
    do {                                                                                      \
        int rc_ = (call);                                                                     \
        if (rc_ < 0) return rc_;                                                              \
    } while (0)

// ------------------------------------------------- launchers (per .hip) ----
namespace pf {
inline unsigned nblocks(long n, int bs = kBlock) { return (unsigned)((n + bs - 1) / bs); }

// pf_kernels.hip
int launch_E_from_rho(pf_ctx* ctx);
int launch_diag(pf_ctx* ctx, bool f32);
int launch_apply(pf_ctx* ctx, bool f32, const void* x, void* y, bool mask_fixed);
int launch_pcg_init(pf_ctx* ctx, bool f32, const void* b, double eps, int itrmax, bool csr);
int launch_pcg_iter(pf_ctx* ctx, bool f32, int parity, bool csr);
int launch_free_to_nodal(pf_ctx* ctx, const double* xf, double* xn);
int launch_nodal_to_free(pf_ctx* ctx, const double* xn, double* xf);
int launch_filter(pf_ctx* ctx, int mode, const double* in, double* out);
int launch_volume(pf_ctx* ctx, const double* rho, double vf, double scale1, double* g_dev);
int launch_comp_sens(pf_ctx* ctx, const double* u, double scale0, double* dfdrho, double* f_dev);
int launch_oc_init(pf_ctx* ctx, const pf_oc_params* prm, double* xplus);
int launch_oc_step(pf_ctx* ctx, const pf_oc_params* prm, const double* x, const double* dfdx, const double* dgdx,
                   double* xplus);
int launch_fill(pf_ctx* ctx, double* p, long n, double v);
int launch_ke_all(pf_ctx* ctx, const double* E);
int launch_csr_count(pf_ctx* ctx, int32_t* rowlen);
int launch_csr_fill(pf_ctx* ctx);
int launch_csr_spmv(pf_ctx* ctx, const double* xf, double* yf);
int launch_csr_diag(pf_ctx* ctx, double* Df);
int launch_scan(pf_ctx* ctx, const int32_t* rowlen, int32_t* indptr, int n);
int launch_residual(pf_ctx* ctx, const double* b, const double* x, double* r, double* rnorm2_dev);
int launch_dot(pf_ctx* ctx, const double* a, const double* b, long n, double* out_dev);
}  // namespace pf
