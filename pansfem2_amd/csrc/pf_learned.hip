// Learned preconditioner: a 4-level convolutional network on the nodal grid
// of a quad4 problem, BF16 weights and activations with FP32 accumulation on
// the MFMA units (v_mfma_f32_16x16x16_bf16), trained on the device with its
// own forward / data-gradient / weight-gradient kernels and Adam.  No
// counterpart in the reference.  Included by pf_capi.hip.
//
// Tensors are H x W x C, channels fastest, C a multiple of 16; H = nx + 1 and
// W = ny + 1 at the finest level (node (i, j) is pixel (i, j)).
//
// MFMA operand layout (16x16x16, one wavefront; checked on gfx950):
//   A (16 x 16, M x K): lane l holds A[l % 16][4 (l / 16) + t], t = 0..3
//   B (16 x 16, K x N): lane l holds B[4 (l / 16) + t][l % 16]
//   D (16 x 16, M x N): lane l holds D[4 (l / 16) + t][l % 16]
// A convolution is O^T (Cout x pixels) = W (Cout x 9 Cin) . X_col (9 Cin x pixels):
// W is stored [tap][Cout][Cin] so A fragments are 8-byte loads, X fragments
// are 8-byte loads of 4 channels of one pixel, D fragments 8-byte stores.
#pragma once

namespace pf {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t bf16;

union F32Bits {
    float f;
    uint32_t u;
};

__device__ __host__ inline bf16 f2bf(float f) {
    // round to nearest even (no NaN inputs occur here)
    F32Bits b;
    b.f = f;
    uint32_t u = b.u;
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf16)(u >> 16);
}

__device__ __host__ inline float bf2f(bf16 h) {
    F32Bits b;
    b.u = (uint32_t)h << 16;
    return b.f;
}

__device__ inline bf16x4 ld4(const bf16* p) { return *reinterpret_cast<const bf16x4*>(p); }

// ------------------------------------------------------------- forward ---
// One workgroup = one image row, 4 wavefronts x 16 pixels.  mode: 0 linear,
// 1 ReLU.  If Yf is non-null the FP32 result is written there instead.
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_conv_fwd(int H, int W, const bf16* __restrict__ X, const bf16* __restrict__ Wt,
                                                  bf16* __restrict__ Y, float* __restrict__ Yf, int relu) {
    const int i = blockIdx.x;
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int j0 = blockIdx.y * 64 + wave * 16;
    if (j0 >= W) return;
    const int px = j0 + (l & 15);
    const int g = l >> 4;
    f32x4 acc[COUT / 16];
#pragma unroll
    for (int c = 0; c < COUT / 16; c++) acc[c] = f32x4{0.f, 0.f, 0.f, 0.f};
    const bf16x4 zero = {0, 0, 0, 0};
#pragma unroll
    for (int tap = 0; tap < 9; tap++) {
        const int ii = i + tap / 3 - 1, jj = px + tap % 3 - 1;
        const bool in = ii >= 0 && ii < H && jj >= 0 && jj < W;
        const bf16* xp = X + ((long)ii * W + jj) * CIN;
#pragma unroll
        for (int kc = 0; kc < CIN / 16; kc++) {
            const int ci = kc * 16 + 4 * g;
            bf16x4 b = in ? ld4(xp + ci) : zero;
#pragma unroll
            for (int c = 0; c < COUT / 16; c++) {
                const int co = c * 16 + (l & 15);
                bf16x4 a = ld4(Wt + ((long)(tap * COUT + co) * CIN + ci));
                acc[c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[c], 0, 0, 0);
            }
        }
    }
    if (px >= W) return;
    const long o = ((long)i * W + px) * COUT;
#pragma unroll
    for (int c = 0; c < COUT / 16; c++) {
        float v[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            v[t] = acc[c][t];
            if (relu && v[t] < 0.f) v[t] = 0.f;
        }
        const int co = c * 16 + 4 * g;
        if (Yf) {
#pragma unroll
            for (int t = 0; t < 4; t++) Yf[o + co + t] = v[t];
        } else {
            bf16x4 pk = {(short)f2bf(v[0]), (short)f2bf(v[1]), (short)f2bf(v[2]), (short)f2bf(v[3])};
            *reinterpret_cast<bf16x4*>(Y + o + co) = pk;
        }
    }
}

// -------------------------------------------------------- weight gradient --
// dW[tap][co][ci] = sum_p dZ[p][co] X[p + off(tap)][ci]: per workgroup a
// range of `ppb` pixels, FP32 partials per workgroup (reduced afterwards in
// a fixed order).
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_conv_wgrad(int H, int W, const bf16* __restrict__ X,
                                                    const bf16* __restrict__ dZ, float* __restrict__ part, int ppb) {
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 4;
    const long npx = (long)H * W;
    const long p_begin = (long)blockIdx.x * ppb;
    const long p_end = min(npx, p_begin + ppb);
    constexpr int CT = COUT / 16, KT = CIN / 16, ITEMS = 9 * CT * KT;
    float* out = part + (long)blockIdx.x * 9 * COUT * CIN;
    for (int item = wave; item < ITEMS; item += 4) {
        const int tap = item / (CT * KT), rem = item - tap * CT * KT, ct = rem / KT, kt = rem - ct * KT;
        const int di = tap / 3 - 1, dj = tap % 3 - 1;
        const int co = ct * 16 + (l & 15), ci = kt * 16 + (l & 15);
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (long p0 = p_begin; p0 < p_end; p0 += 16) {
            bf16x4 a, b;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const long p = p0 + 4 * g + t;
                short av = 0, bv = 0;
                if (p < p_end) {
                    av = (short)dZ[p * COUT + co];
                    const int pi = (int)(p / W), pj = (int)(p - (long)pi * W);
                    const int ii = pi + di, jj = pj + dj;
                    if (ii >= 0 && ii < H && jj >= 0 && jj < W) bv = (short)X[((long)ii * W + jj) * CIN + ci];
                }
                a[t] = av;
                b[t] = bv;
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int cor = ct * 16 + 4 * g + t;
            out[((long)tap * COUT + cor) * CIN + ci] = acc[t];
        }
    }
}

__global__ void k_wgrad_reduce(const float* __restrict__ part, int nparts, int n, float* __restrict__ dW) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int b = 0; b < nparts; b++) s += part[(long)b * n + q];
        dW[q] = s;
    }
}

// ------------------------------------------------- elementwise / resample --
__global__ void k_relu_bwd(long n, const bf16* __restrict__ Y, bf16* __restrict__ dY) {
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
        if (!(bf2f(Y[q]) > 0.f)) dY[q] = 0;
}

// average pooling 2x2 (ceil), coarse (Hc, Wc) from fine (H, W)
__global__ void k_pool2(int H, int W, int C, const bf16* __restrict__ in, bf16* __restrict__ out) {
    const int Hc = (H + 1) / 2, Wc = (W + 1) / 2;
    const long n = (long)Hc * Wc * C;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int c = (int)(q % C);
        const long pc = q / C;
        const int I = (int)(pc / Wc), J = (int)(pc - (long)I * Wc);
        float s = 0.f;
        int cnt = 0;
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                int i = 2 * I + a, j = 2 * J + b;
                if (i < H && j < W) {
                    s += bf2f(in[((long)i * W + j) * C + c]);
                    cnt++;
                }
            }
        out[q] = f2bf(s / (float)cnt);
    }
}

// d_fine += d_coarse / count (backward of k_pool2, accumulated)
__global__ void k_pool2_bwd_add(int H, int W, int C, const bf16* __restrict__ dout, bf16* __restrict__ din) {
    const int Wc = (W + 1) / 2;
    const long n = (long)H * W * C;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int c = (int)(q % C);
        const long p = q / C;
        const int i = (int)(p / W), j = (int)(p - (long)i * W);
        const int I = i / 2, J = j / 2;
        const int cnt = ((2 * I + 1 < H) ? 2 : 1) * ((2 * J + 1 < W) ? 2 : 1);
        din[q] = f2bf(bf2f(din[q]) + bf2f(dout[((long)I * Wc + J) * C + c]) / (float)cnt);
    }
}

// out (fine) = nearest-upsampled coarse + skip
__global__ void k_up2_add(int H, int W, int C, const bf16* __restrict__ coarse, const bf16* __restrict__ skip,
                          bf16* __restrict__ out) {
    const int Wc = (W + 1) / 2;
    const long n = (long)H * W * C;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int c = (int)(q % C);
        const long p = q / C;
        const int i = (int)(p / W), j = (int)(p - (long)i * W);
        out[q] = f2bf(bf2f(coarse[((long)(i / 2) * Wc + j / 2) * C + c]) + bf2f(skip[q]));
    }
}

// d_coarse = sum of the fine gradient over each 2x2 block
__global__ void k_up2_bwd(int H, int W, int C, const bf16* __restrict__ dfine, bf16* __restrict__ dcoarse) {
    const int Hc = (H + 1) / 2, Wc = (W + 1) / 2;
    const long n = (long)Hc * Wc * C;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int c = (int)(q % C);
        const long pc = q / C;
        const int I = (int)(pc / Wc), J = (int)(pc - (long)I * Wc);
        float s = 0.f;
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                int i = 2 * I + a, j = 2 * J + b;
                if (i < H && j < W) s += bf2f(dfine[((long)i * W + j) * C + c]);
            }
        dcoarse[q] = f2bf(s);
    }
}

__global__ void k_copy_bf16(long n, const bf16* __restrict__ a, bf16* __restrict__ b) {
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) b[q] = a[q];
}

// ------------------------------------------------------- network input ----
// x0[p] = (2 rho_node - 1, r_x / s, r_y / s, 0 ...), 16 channels, where
// rho_node is the mean density of the node's elements and s the RMS of r.
__global__ void k_net_input(int nx, int ny, const double* __restrict__ rho, const double* __restrict__ r, double rs,
                            bf16* __restrict__ x0) {
    const long nn = (long)(nx + 1) * (ny + 1);
    for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < nn; n += (long)gridDim.x * blockDim.x) {
        int i = (int)(n / (ny + 1)), j = (int)(n - (long)i * (ny + 1));
        double s = 0.0;
        int cnt = 0;
        for (int a = -1; a <= 0; a++)
            for (int b = -1; b <= 0; b++) {
                int ei = i + a, ej = j + b;
                if (ei >= 0 && ei < nx && ej >= 0 && ej < ny) {
                    s += rho[(long)ny * ei + ej];
                    cnt++;
                }
            }
        float f[16] = {0};
        f[0] = (float)(2.0 * (s / cnt) - 1.0);
        f[1] = (float)(r[2 * n] / rs);
        f[2] = (float)(r[2 * n + 1] / rs);
        for (int c = 0; c < 16; c++) x0[n * 16 + c] = f2bf(f[c]);
    }
}

// Up to 32 dot products against one vector in a single pass:
// part[block * nb + j] = sum_i a_i * b_j[i], fixed-order block sums.
struct PtrPack {
    const double* p[32];
};
struct CoefPack {
    double c[32];
};

// NB: the accumulator count (>= nb), so that they stay in registers.
template <int NB>
__global__ void __launch_bounds__(256) k_multidot(long n, int nb, const double* __restrict__ a, PtrPack B,
                                                  double* __restrict__ part) {
    __shared__ double lds[16 * NB];
    double acc[NB];
#pragma unroll
    for (int j = 0; j < NB; j++) acc[j] = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double ai = a[i];
#pragma unroll
        for (int j = 0; j < NB; j++)
            if (j < nb) acc[j] += ai * B.p[j][i];
    }
    block_sum<NB>(acc, lds);
    if (threadIdx.x == 0)
        for (int j = 0; j < nb; j++) part[blockIdx.x * nb + j] = acc[j];
}

// The direction and its image without an operator application: with
// z = a t + c y and K z = a Kt + c Ky,
//   p = z - sum_j c_j P_j,  q = K p = K z - sum_j c_j Q_j.
__global__ void __launch_bounds__(256) k_fcg_direction2(long n, int m, double a, double c, const double* __restrict__ t,
                                                        const double* __restrict__ y, const double* __restrict__ Kt,
                                                        const double* __restrict__ Ky, PtrPack P, PtrPack Q,
                                                        CoefPack cf, double* __restrict__ p, double* __restrict__ q) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double v = a * t[i] + c * y[i];
        double w = a * Kt[i] + c * Ky[i];
        for (int j = 0; j < m; j++) {
            v -= cf.c[j] * P.p[j][i];
            w -= cf.c[j] * Q.p[j][i];
        }
        p[i] = v;
        q[i] = w;
    }
}

// x += alpha p, r -= alpha q
__global__ void __launch_bounds__(256) k_step2(long n, double alpha, const double* __restrict__ p,
                                               const double* __restrict__ q, double* __restrict__ x,
                                               double* __restrict__ r) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        x[i] += alpha * p[i];
        r[i] -= alpha * q[i];
    }
}

// y channels 0, 1 -> nodal FP64 vector (0 on fixed DOFs)
__global__ void k_net_y_nodal(long nn, const float* __restrict__ y, const uint8_t* __restrict__ fix,
                              double* __restrict__ v) {
    for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < nn; n += (long)gridDim.x * blockDim.x) {
        uint8_t f = fix[n];
        for (int c = 0; c < 2; c++) v[2 * n + c] = (f & (1u << c)) ? 0.0 : (double)y[n * 16 + c];
    }
}

// nodal FP64 gradient -> 16-channel BF16 output gradient (channels 0, 1)
__global__ void k_net_dy_from_nodal(long nn, const double* __restrict__ g, const uint8_t* __restrict__ fix,
                                    bf16* __restrict__ dy) {
    for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < nn; n += (long)gridDim.x * blockDim.x) {
        uint8_t f = fix[n];
        float v[16] = {0};
        for (int c = 0; c < 2; c++) v[c] = (f & (1u << c)) ? 0.f : (float)g[2 * n + c];
        for (int c = 0; c < 16; c++) dy[n * 16 + c] = f2bf(v[c]);
    }
}

// Smooth random displacement field for synthetic training pairs (K v, v):
// a sum over scales s = 2^l of bilinearly interpolated lattice noise of
// amplitude s (hash-based, seeded), zero on fixed DOFs.
__device__ inline double lattice_noise(uint64_t seed, int l, int I, int J, int c) {
    uint64_t z = seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(l + 1)) ^ ((uint64_t)(uint32_t)I << 21) ^
                 ((uint64_t)(uint32_t)J << 42) ^ (uint64_t)c;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

__global__ void k_random_field(int nx, int ny, uint64_t seed, int levels, const uint8_t* __restrict__ fix,
                               double* __restrict__ v) {
    const long nn = (long)(nx + 1) * (ny + 1);
    for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < nn; n += (long)gridDim.x * blockDim.x) {
        const int i = (int)(n / (ny + 1)), j = (int)(n - (long)i * (ny + 1));
        const uint8_t f = fix[n];
        for (int c = 0; c < 2; c++) {
            double acc = 0.0;
            for (int l = 0; l < levels; l++) {
                const int sp = 1 << l;
                const int I = i / sp, J = j / sp;
                const double fi = (double)(i - I * sp) / sp, fj = (double)(j - J * sp) / sp;
                double a = lattice_noise(seed, l, I, J, c), b = lattice_noise(seed, l, I + 1, J, c);
                double cc = lattice_noise(seed, l, I, J + 1, c), d = lattice_noise(seed, l, I + 1, J + 1, c);
                acc += (double)sp * ((1 - fi) * ((1 - fj) * a + fj * cc) + fi * ((1 - fj) * b + fj * d));
            }
            v[2 * n + c] = (f & (1u << c)) ? 0.0 : acc;
        }
    }
}

// d = (r + s y) ./ D on free DOFs (0 on fixed), y = network output channels 0, 1
__global__ void k_net_output(long nn, const float* __restrict__ y, const double* __restrict__ r,
                             const double* __restrict__ D, const uint8_t* __restrict__ fix, double rs,
                             double* __restrict__ d) {
    for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < nn; n += (long)gridDim.x * blockDim.x) {
        uint8_t f = fix[n];
        for (int c = 0; c < 2; c++) {
            double v = (r[2 * n + c] + rs * (double)y[n * 16 + c]) / D[2 * n + c];
            d[2 * n + c] = (f & (1u << c)) ? 0.0 : v;
        }
    }
}

// loss gradient wrt y: dy = s g ./ D on free DOFs, g = dL/dd
__global__ void k_net_output_bwd(long nn, const double* __restrict__ g, const double* __restrict__ D,
                                 const uint8_t* __restrict__ fix, double rs, bf16* __restrict__ dy) {
    for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < nn; n += (long)gridDim.x * blockDim.x) {
        uint8_t f = fix[n];
        float v[16] = {0};
        for (int c = 0; c < 2; c++) v[c] = (f & (1u << c)) ? 0.f : (float)(rs * g[2 * n + c] / D[2 * n + c]);
        for (int c = 0; c < 16; c++) dy[n * 16 + c] = f2bf(v[c]);
    }
}

// ------------------------------------------------------------------ Adam --
// FP32 master weights; the BF16 copy and its transposed, tap-flipped copy
// (for the data gradient) are refreshed after the step.
__global__ void k_adam(int n, float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ grad, float lr, float b1, float b2, float eps, float c1, float c2,
                       float scale) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        float gq = grad[q] * scale;
        float mq = b1 * m[q] + (1.f - b1) * gq;
        float vq = b2 * v[q] + (1.f - b2) * gq * gq;
        m[q] = mq;
        v[q] = vq;
        w[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
    }
}

// Wb[tap][co][ci] = bf16(W); Wtb[tap'][ci][co] = bf16(W[8 - tap'][co][ci])
__global__ void k_weights_bf16(int cout, int cin, const float* __restrict__ W, bf16* __restrict__ Wb,
                               bf16* __restrict__ Wtb) {
    const int n = 9 * cout * cin;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const int tap = q / (cout * cin), rem = q - tap * cout * cin, co = rem / cin, ci = rem - co * cin;
        bf16 h = f2bf(W[q]);
        Wb[q] = h;
        Wtb[((long)(8 - tap) * cin + ci) * cout + co] = h;
    }
}


// training pair (r, e) -> FP32 dataset slots
__global__ void k_pair_store(long n, const double* __restrict__ r, const double* __restrict__ e,
                             float* __restrict__ rout, float* __restrict__ eout) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        rout[i] = (float)r[i];
        eout[i] = (float)e[i];
    }
}

__global__ void k_d2f(long n, const double* __restrict__ a, float* __restrict__ b) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = (float)a[i];
}

__global__ void k_f2d(long n, const float* __restrict__ a, double* __restrict__ b) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = (double)a[i];
}

__global__ void k_sub_scale(long n, const double* __restrict__ a, const double* __restrict__ b, double s,
                            double* __restrict__ out) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = s * (a[i] - b[i]);
}

}  // namespace pf
