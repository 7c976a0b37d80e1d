// The learned FLD preconditioner on the device: a 3-level U-Net (16/32/64 channels, 3x3 convolutions, BF16
// activations and weights, FP32 accumulation), its backward pass and Adam.
//
// The reference has no learned part (its preconditioners are Jacobi and the multigrid of
// crates/sr-engine/shaders/heat.wgsl:1160-1526); the network's definition is this library's, stated layer by layer in
// the f64 numpy checker oracle/unet.py, whose names the kernels follow.
//
//   * activations are channels-last BF16 (pixel-major, channels contiguous), so a lane's eight consecutive K values
//     of the implicit GEMM are one 16-byte load;
//   * k_conv is the implicit-GEMM 3x3 convolution on the matrix cores: M = pixels (16 per MFMA tile, a 32 x 4
//     pixel tile per block with its input staged in LDS), N = output channels, K = 9 taps x input channels in chunks
//     of 32 (v_mfma_f32_16x16x32_bf16). The same kernel computes the backward-data pass with mirrored, transposed
//     weights and the ReLU mask applied while staging; upsample + concat of the decoder is folded into the staging;
//   * the weight gradient (k_dw) is a reduction over pixels: FP32 on the vector units from LDS tiles, per-block
//     partials summed in a fixed order by k_dw_reduce (deterministic, no atomics);
//   * the output 1x1 layer, the features and the energy-norm loss are elementwise / stencil kernels.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "net.hpp"

namespace sr {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define NTH 256
#define NB_FEAT 1024
#define DW_SEG 32
#define DW_BLOCKS 256
#define LOG_CLAMP 20.0f

__device__ __forceinline__ bf16x8 ld8(const bf16 *p) { return *(const bf16x8 *)p; }
__device__ __forceinline__ void st8(bf16 *p, bf16x8 v) { *(bf16x8 *)p = v; }

// the sum of a cell's four face couplings (rim faces included, as the features read them)
__device__ __forceinline__ float coupling_sum(int w, const float *kx, const float *ky, int i, int y, int x) {
    return kx[(size_t)y * (w + 1) + x] + kx[(size_t)y * (w + 1) + x + 1] + ky[i] + ky[i + w];
}

// ------------------------------------------------------------------------------------------------------- features
// partials of sum (r d^-1/2)^2 and sum log(mean coupling)
__global__ void __launch_bounds__(NTH) k_feat_sums(int w, int h, const float *__restrict__ own,
                                                   const float *__restrict__ kx, const float *__restrict__ ky,
                                                   const float *__restrict__ r, double *part) {
    double v[2] = {0.0, 0.0};
    const int n = w * h;
    for (int i = blockIdx.x * NTH + threadIdx.x; i < n; i += gridDim.x * NTH) {
        const int y = i / w, x = i - y * w;
        const float ks = coupling_sum(w, kx, ky, i, y, x);
        const float rt = r[i] * rsqrtf(own[i] + ks);
        v[0] += (double)rt * rt;
        v[1] += (double)logf(fmaxf(0.25f * ks, 1e-30f));
    }
    block_partials<2>(v, part, gridDim.x);
}

// the three input channels (padded to 8): r d^-1/2 / s, log mean coupling - its mean (clamped), log(own / d);
// d^-1/2 kept for the output's scaling; scal = {s, mean}
__global__ void __launch_bounds__(NTH) k_feat(int w, int h, const float *__restrict__ own, const float *__restrict__ kx,
                                              const float *__restrict__ ky, const float *__restrict__ r,
                                              const double *part, int nb, bf16 *__restrict__ x,
                                              float *__restrict__ dis, float *scal) {
    const double s2 = sum_partials(part, nb), sl = sum_partials(part + nb, nb);
    const int n = w * h;
    const float s = s2 > 0.0 ? (float)sqrt(s2 / n) : 1.0f, mean = (float)(sl / n);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[0] = s;
        scal[1] = mean;
    }
    const float is = 1.0f / s;
    for (int i = blockIdx.x * NTH + threadIdx.x; i < n; i += gridDim.x * NTH) {
        const int y = i / w, xx = i - y * w;
        const float ks = coupling_sum(w, kx, ky, i, y, xx);
        const float d = own[i] + ks;
        const float di = rsqrtf(d);
        dis[i] = di;
        bf16x8 f = {};
        f[0] = (bf16)(r[i] * di * is);
        f[1] = (bf16)fminf(fmaxf(logf(fmaxf(0.25f * ks, 1e-30f)) - mean, -LOG_CLAMP), LOG_CLAMP);
        f[2] = (bf16)logf(own[i] / d);
        st8(x + (size_t)8 * i, f);
    }
}

// ------------------------------------------------------------------------------------------- implicit-GEMM conv 3x3
struct ConvArgs {
    const bf16 *in;    // (H, W, CIN - cup): channels [cup, CIN)
    const bf16 *up;    // (H/2, W/2, cup): channels [0, cup), nearest-upsampled; null when cup == 0
    const bf16 *mask;  // (H, W, CIN): entries whose mask is not > 0 read as zero; or null (only with cup == 0)
    const bf16 *wt;    // (COUT, KP), k = tap * CIN + ci, zero past 9 CIN
    const float *bias; // (COUT) or null
    bf16 *out;         // (H, W, COUT)
    int H, W, cup, relu;
};

// A block computes a tile of TW x TH pixels (a row of TW per wave, two 16-pixel MFMA tiles) x all COUT channels. Its
// input -- rows y0 - 1 .. y0 + TH, columns x0 - 1 .. x0 + TW, every input channel, upsampled and concatenated or
// masked as the layer reads it -- is staged once in LDS, so the nine taps read LDS instead of global memory. Lane l
// of an MFMA holds A[m = l % 16][k = 8 (l / 16) + j], B[k = 8 (l / 16) + j][n = l % 16] and D[m = 4 (l / 16) + j]
// [n = l % 16]; K runs over 9 taps x CIN channels in chunks of 32, zero past 9 CIN.
#define TW 32
#define TH 4
template <int CIN, int COUT>
__global__ void __launch_bounds__(NTH) k_conv(ConvArgs a) {
    constexpr int KP = (9 * CIN + 31) / 32 * 32;
    constexpr int NT = COUT / 16;
    constexpr int CP = CIN + 8;                    // pixel stride in LDS (bf16), staggering the banks
    constexpr int RW = TW + 2, RH = TH + 2, OCT = CIN / 8;
    static_assert(CIN % 8 == 0 && COUT % 16 == 0, "channel counts");
    __shared__ __attribute__((aligned(16))) bf16 tile[RH * RW * CP];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, c16 = lane & 15;
    const int H = a.H, W = a.W, cup = a.cup, cs = CIN - cup;
    const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH;
    // stage the input tile (zero outside the image)
    for (int t = threadIdx.x; t < RH * RW * OCT; t += NTH) {
        const int o = t % OCT, px = t / OCT, lx = px % RW, ly = px / RW;
        const int xx = x0 + lx - 1, yy = y0 + ly - 1, ci = 8 * o;
        bf16x8 v = {};
        if (xx >= 0 && xx < W && yy >= 0 && yy < H) {
            if (ci < cup)
                v = ld8(a.up + ((size_t)(yy >> 1) * (W >> 1) + (xx >> 1)) * cup + ci);
            else
                v = ld8(a.in + ((size_t)yy * W + xx) * cs + (ci - cup));
            if (a.mask) {
                const bf16x8 mv = ld8(a.mask + ((size_t)yy * W + xx) * CIN + ci);
#pragma unroll
                for (int j = 0; j < 8; j++)
                    if (!((float)mv[j] > 0.0f)) v[j] = (bf16)0.0f;
            }
        }
        st8(tile + (size_t)px * CP + ci, v);
    }
    __syncthreads();
    f32x4 acc[2][NT];
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int t = 0; t < NT; t++) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int kc = 0; kc < KP; kc += 32) {
        const int k0 = kc + 8 * g;
        const int tap = k0 / CIN, ci = k0 - tap * CIN;
        const int ty = tap / 3, tx = tap - 3 * ty;
        bf16x8 bv[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) bv[t] = ld8(a.wt + (size_t)(t * 16 + c16) * KP + k0);
#pragma unroll
        for (int m = 0; m < 2; m++) {
            bf16x8 av = {};
            if (tap < 9) av = ld8(tile + (size_t)((wave + ty) * RW + 16 * m + c16 + tx) * CP + ci);
#pragma unroll
            for (int t = 0; t < NT; t++)
                acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv[t], acc[m][t], 0, 0, 0);
        }
    }
    const int y = y0 + wave;
    if (y >= H) return;
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int co = t * 16 + c16;
            const float b = a.bias ? a.bias[co] : 0.0f;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int x = x0 + 16 * m + 4 * g + j;
                if (x < W) {
                    float v = acc[m][t][j] + b;
                    if (a.relu) v = fmaxf(v, 0.0f);
                    a.out[((size_t)y * W + x) * COUT + co] = (bf16)v;
                }
            }
        }
}

template <int CIN, int COUT> static void conv(hipStream_t s, const ConvArgs &a) {
    const dim3 grid((a.W + TW - 1) / TW, (a.H + TH - 1) / TH);
    hipLaunchKernelGGL((k_conv<CIN, COUT>), grid, dim3(NTH), 0, s, a);
}

// 2x2 mean, one thread per (output pixel, channel octet)
__global__ void __launch_bounds__(NTH) k_pool(const bf16 *__restrict__ in, int H, int W, int C, bf16 *__restrict__ out) {
    const int Ho = H / 2, Wo = W / 2, oct = C / 8;
    const int idx = blockIdx.x * NTH + threadIdx.x;
    if (idx >= Ho * Wo * oct) return;
    const int o = idx % oct, p = idx / oct, Y = p / Wo, X = p - Y * Wo;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int dy = 0; dy < 2; dy++)
#pragma unroll
        for (int dx = 0; dx < 2; dx++) {
            const bf16x8 v = ld8(in + ((size_t)(2 * Y + dy) * W + 2 * X + dx) * C + 8 * o);
#pragma unroll
            for (int j = 0; j < 8; j++) s[j] += (float)v[j];
        }
    bf16x8 r;
#pragma unroll
    for (int j = 0; j < 8; j++) r[j] = (bf16)(0.25f * s[j]);
    st8(out + (size_t)p * C + 8 * o, r);
}

// the 1x1 output layer and the scaling back to the system's units: zn = (b + w . d1) s d^-1/2;
// z = zadd + zn and/or err = zn - e
__global__ void __launch_bounds__(NTH) k_out(int n, const bf16 *__restrict__ d1, const float *__restrict__ wout,
                                             const float *scal, const float *__restrict__ dis,
                                             const float *__restrict__ zadd, float *__restrict__ z,
                                             const float *__restrict__ e, float *__restrict__ err) {
    const float s = scal[0];
    float wv[17];
#pragma unroll
    for (int c = 0; c < 17; c++) wv[c] = wout[c];
    for (int i = blockIdx.x * NTH + threadIdx.x; i < n; i += gridDim.x * NTH) {
        const bf16x8 a = ld8(d1 + (size_t)16 * i), b = ld8(d1 + (size_t)16 * i + 8);
        float o = wv[16];
#pragma unroll
        for (int j = 0; j < 8; j++) o += wv[j] * (float)a[j] + wv[8 + j] * (float)b[j];
        const float zn = o * s * dis[i];
        if (err) err[i] = zn - e[i];
        if (z) z[i] = (zadd ? zadd[i] : 0.0f) + zn;
    }
}

// ------------------------------------------------------------------------------------------------------- training
// A v in face form (own v + sum over interior faces k (v_i - v_j)), as the solver's level 0 applies it
__device__ __forceinline__ float apply_a(int w, int h, const float *own, const float *kx, const float *ky,
                                         const float *v, int i, int y, int x) {
    const float vi = v[i];
    float s = own[i] * vi;
    if (x > 0) s += kx[(size_t)y * (w + 1) + x] * (vi - v[i - 1]);
    if (x < w - 1) s += kx[(size_t)y * (w + 1) + x + 1] * (vi - v[i + 1]);
    if (y > 0) s += ky[i] * (vi - v[i - w]);
    if (y < h - 1) s += ky[i + w] * (vi - v[i + w]);
    return s;
}

// aerr = A err; partials of err . A err and e . A e
__global__ void __launch_bounds__(NTH) k_loss(int w, int h, const float *__restrict__ own, const float *__restrict__ kx,
                                              const float *__restrict__ ky, const float *__restrict__ err,
                                              const float *__restrict__ e, float *__restrict__ aerr, double *part) {
    double v[2] = {0.0, 0.0};
    const int n = w * h;
    for (int i = blockIdx.x * NTH + threadIdx.x; i < n; i += gridDim.x * NTH) {
        const int y = i / w, x = i - y * w;
        const float ae = apply_a(w, h, own, kx, ky, err, i, y, x);
        aerr[i] = ae;
        v[0] += (double)err[i] * ae;
        v[1] += (double)e[i] * apply_a(w, h, own, kx, ky, e, i, y, x);
    }
    block_partials<2>(v, part, gridDim.x);
}

// dL/d(out) = 2 A err / (e . A e) * s d^-1/2; then dL/d(d1) = gout w_out and partials of dL/dw_out, dL/db_out
__global__ void __launch_bounds__(NTH) k_out_back(int n, const float *__restrict__ aerr, const double *lpart, int nbl,
                                                  const float *scal, const float *__restrict__ dis,
                                                  const bf16 *__restrict__ d1, const float *__restrict__ wout,
                                                  float *__restrict__ gout, bf16 *__restrict__ g_d1, double *part) {
    const double norm = sum_partials(lpart + nbl, nbl);
    const float f = norm > 0.0 ? (float)(2.0 / norm) * scal[0] : 0.0f;
    double v[17];
#pragma unroll
    for (int c = 0; c < 17; c++) v[c] = 0.0;
    float wv[16];
#pragma unroll
    for (int c = 0; c < 16; c++) wv[c] = wout[c];
    for (int i = blockIdx.x * NTH + threadIdx.x; i < n; i += gridDim.x * NTH) {
        const float go = f * aerr[i] * dis[i];
        gout[i] = go;
        const bf16x8 a = ld8(d1 + (size_t)16 * i), b = ld8(d1 + (size_t)16 * i + 8);
        bf16x8 ga, gb;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            v[j] += (double)((float)a[j] * go);
            v[8 + j] += (double)((float)b[j] * go);
            ga[j] = (bf16)(go * wv[j]);
            gb[j] = (bf16)(go * wv[8 + j]);
        }
        v[16] += go;
        st8(g_d1 + (size_t)16 * i, ga);
        st8(g_d1 + (size_t)16 * i + 8, gb);
    }
    block_partials<17>(v, part, gridDim.x);
}

// out[k] = sum of part[k nb .. k nb + nb) for k < K (fixed order)
__global__ void k_reduce_rows(const double *part, int nb, int K, float *out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    double s = 0.0;
    for (int i = 0; i < nb; i++) s += part[(size_t)k * nb + i];
    out[k] = (float)s;
}

// Weight and bias gradients of one layer: dW[tap][ci][co] = sum_p x[p + tap][ci] ga[p][co], ga = gy (y > 0).
// A block takes 32-pixel row segments in turn; thread (gi, co) accumulates every tap of input channels gi, gi + NG,
// ... for output channel co, reading the staged tiles from LDS (x broadcast across the wave, ga conflict-free).
template <int CIN, int COUT>
__global__ void __launch_bounds__(NTH) k_dw(const bf16 *__restrict__ in, const bf16 *__restrict__ up, int cup,
                                            const bf16 *__restrict__ gy, const bf16 *__restrict__ y, int H, int W,
                                            float *__restrict__ part) {
    constexpr int NG = NTH / COUT;
    constexpr int CPT = CIN >= NG ? CIN / NG : 1;  // input channels per thread
    static_assert(CIN < NG || CIN % NG == 0, "channel split");
    __shared__ float xs[3][DW_SEG + 2][CIN];
    __shared__ float gs[DW_SEG][COUT];
    const int co = threadIdx.x % COUT, gi = threadIdx.x / COUT;
    const bool active = gi < CIN;
    float acc[9][CPT];
#pragma unroll
    for (int t = 0; t < 9; t++)
#pragma unroll
        for (int c = 0; c < CPT; c++) acc[t][c] = 0.0f;
    float bacc = 0.0f;
    const int cs = CIN - cup, spr = (W + DW_SEG - 1) / DW_SEG, nseg = H * spr;
    for (int sg = blockIdx.x; sg < nseg; sg += gridDim.x) {
        const int yr = sg / spr, x0 = (sg - yr * spr) * DW_SEG;
        __syncthreads();
        for (int t = threadIdx.x; t < 3 * (DW_SEG + 2) * CIN; t += NTH) {
            const int c = t % CIN, rest = t / CIN, px = rest % (DW_SEG + 2), ry = rest / (DW_SEG + 2);
            const int yy = yr + ry - 1, xx = x0 + px - 1;
            float val = 0.0f;
            if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                val = c < cup ? (float)up[((size_t)(yy >> 1) * (W >> 1) + (xx >> 1)) * cup + c]
                              : (float)in[((size_t)yy * W + xx) * cs + (c - cup)];
            xs[ry][px][c] = val;
        }
        for (int t = threadIdx.x; t < DW_SEG * COUT; t += NTH) {
            const int c = t % COUT, px = t / COUT, xx = x0 + px;
            float gv = 0.0f;
            if (xx < W) {
                const size_t i = ((size_t)yr * W + xx) * COUT + c;
                gv = (float)y[i] > 0.0f ? (float)gy[i] : 0.0f;
            }
            gs[px][c] = gv;
        }
        __syncthreads();
        if (active) {
            for (int p = 0; p < DW_SEG; p++) {
                const float gv = gs[p][co];
                bacc += gv;
#pragma unroll
                for (int t = 0; t < 9; t++)
#pragma unroll
                    for (int c = 0; c < CPT; c++) acc[t][c] += xs[t / 3][p + t % 3][gi + c * NG] * gv;
            }
        }
    }
    float *pb = part + (size_t)blockIdx.x * (9 * CIN * COUT + COUT);
    if (active) {
#pragma unroll
        for (int t = 0; t < 9; t++)
#pragma unroll
            for (int c = 0; c < CPT; c++) pb[((size_t)t * CIN + gi + c * NG) * COUT + co] = acc[t][c];
        if (gi == 0) pb[9 * CIN * COUT + co] = bacc;
    }
}

// sums the k_dw partials of nb blocks into a layer's gradients in the (3, 3, cin, cout) + (cout) layout
__global__ void k_dw_reduce(const float *__restrict__ part, int nb, int cinp, int cin, int cout, float *gw, float *gb) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    const int nw = 9 * cin * cout;
    if (o >= nw + cout) return;
    const size_t stride = (size_t)9 * cinp * cout + cout;
    size_t src;
    if (o < nw) {
        const int co = o % cout, rest = o / cout, ci = rest % cin, tap = rest / cin;
        src = ((size_t)tap * cinp + ci) * cout + co;
    } else {
        src = (size_t)9 * cinp * cout + (o - nw);
    }
    double s = 0.0;
    for (int b = 0; b < nb; b++) s += part[(size_t)b * stride + src];
    if (o < nw)
        gw[o] = (float)s;
    else
        gb[o - nw] = (float)s;
}

// split the gradient of a concat(up(a), skip) input: g_up[Y][X][c] = sum over the 2x2 children (c < cup);
// g_skip[p][c] = g[p][cup + c]
__global__ void k_split_cat(const bf16 *__restrict__ g, int H, int W, int C, int cup, bf16 *__restrict__ gup,
                            bf16 *__restrict__ gskip) {
    const int idx = blockIdx.x * NTH + threadIdx.x;
    const int cs = C - cup, Wo = W / 2, Ho = H / 2;
    const int nup = Ho * Wo * (cup / 8), nsk = H * W * (cs / 8);
    if (idx < nup) {
        const int oc = cup / 8, o = idx % oc, p = idx / oc, Y = p / Wo, X = p - Y * Wo;
        float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int dy = 0; dy < 2; dy++)
#pragma unroll
            for (int dx = 0; dx < 2; dx++) {
                const bf16x8 v = ld8(g + ((size_t)(2 * Y + dy) * W + 2 * X + dx) * C + 8 * o);
#pragma unroll
                for (int j = 0; j < 8; j++) s[j] += (float)v[j];
            }
        bf16x8 r;
#pragma unroll
        for (int j = 0; j < 8; j++) r[j] = (bf16)s[j];
        st8(gup + (size_t)p * cup + 8 * o, r);
    } else if (idx < nup + nsk) {
        const int k = idx - nup, oc = cs / 8, o = k % oc, p = k / oc;
        st8(gskip + (size_t)p * cs + 8 * o, ld8(g + (size_t)p * C + cup + 8 * o));
    }
}

// g_skip += pool_backward(g_pool): each fine pixel gets a quarter of its parent's gradient
__global__ void k_add_pool_back(bf16 *__restrict__ gskip, const bf16 *__restrict__ gp, int H, int W, int C) {
    const int idx = blockIdx.x * NTH + threadIdx.x;
    const int oc = C / 8;
    if (idx >= H * W * oc) return;
    const int o = idx % oc, p = idx / oc, y = p / W, x = p - y * W;
    const bf16x8 a = ld8(gskip + (size_t)p * C + 8 * o);
    const bf16x8 b = ld8(gp + ((size_t)(y >> 1) * (W >> 1) + (x >> 1)) * C + 8 * o);
    bf16x8 r;
#pragma unroll
    for (int j = 0; j < 8; j++) r[j] = (bf16)((float)a[j] + 0.25f * (float)b[j]);
    st8(gskip + (size_t)p * C + 8 * o, r);
}

// Adam (Kingma and Ba 2015, Algorithm 1) on the FP32 master parameters
__global__ void k_adam(int n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float bc1, float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi, vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// the BF16 weight layouts of one layer from its FP32 (3, 3, cin, cout) master: forward (cout, kp) with
// k = tap * cinp + ci, backward-data (cin, kpb) with k = tap' * cout + co and tap' = 8 - tap
__global__ void k_pack(const float *__restrict__ w, int cin, int cinp, int cout, int kp, int kpb, bf16 *__restrict__ wt,
                       bf16 *__restrict__ wb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nf = cout * kp, nbk = wb ? cin * kpb : 0;
    if (i < nf) {
        const int co = i / kp, k = i - co * kp, tap = k / cinp, ci = k - tap * cinp;
        const float v = (tap < 9 && ci < cin) ? w[((size_t)tap * cin + ci) * cout + co] : 0.0f;
        wt[i] = (bf16)v;
    } else if (i < nf + nbk) {
        const int j = i - nf, ci = j / kpb, k = j - ci * kpb, tb = k / cout, co = k - tb * cout;
        const float v = tb < 9 ? w[((size_t)(8 - tb) * cin + ci) * cout + co] : 0.0f;
        wb[j] = (bf16)v;
    }
}

// ------------------------------------------------------------------------------------------------------ host side
static const int LAYER_SHAPE[10][2] = {{3, 16}, {16, 16}, {16, 32}, {32, 32}, {32, 64},
                                       {64, 64}, {96, 32}, {32, 32}, {48, 16}, {16, 16}};
enum { E1A, E1, E2A, E2, E3A, E3, D2A, D2, D1A, D1 };

Net::~Net() {
    for (auto &l : L) {
        if (l.wt) (void)hipFree(l.wt);
        if (l.wb) (void)hipFree(l.wb);
    }
    void *ps[] = {params, grads, m, v, x, e1a, e1, p1, e2a, e2, p2, e3a, e3, d2a, d2, d1a, d1, g_d1, g_d1a, g_cat1,
                  g_d2, g_d2a, g_cat2, g_e3, g_e3a, g_p2, g_e2, g_e2a, g_p1, g_e1, g_e1a, dis, err, aerr, gout, wpart,
                  part, lpart, opart, scal};
    for (void *q : ps)
        if (q) (void)hipFree(q);
    if (host) (void)hipHostFree(host);
}

template <typename T> static int dalloc(T **p, size_t count) {
    SR_HIP(hipMalloc((void **)p, sizeof(T) * std::max<size_t>(count, 1)));
    SR_HIP(hipMemset(*p, 0, sizeof(T) * std::max<size_t>(count, 1)));
    return SR_OK;
}

int Net::init(int mw, int mh) {
    if (mw <= 0 || mh <= 0 || mw % 4 || mh % 4)
        return fail(SR_ERR_ARGUMENT, "the network's grid sides must be positive multiples of 4");
    max_w = mw;
    max_h = mh;
    size_t off = 0;
    for (int i = 0; i < 10; i++) {
        NetLayer l;
        l.cin = LAYER_SHAPE[i][0];
        l.cout = LAYER_SHAPE[i][1];
        l.cinp = i == E1A ? 8 : l.cin;
        l.kp = (9 * l.cinp + 31) / 32 * 32;
        l.kpb = (9 * l.cout + 31) / 32 * 32;
        l.w_off = off;
        off += (size_t)9 * l.cin * l.cout;
        l.b_off = off;
        off += l.cout;
        SR_TRY(dalloc(&l.wt, (size_t)l.cout * l.kp));
        if (i != E1A) SR_TRY(dalloc(&l.wb, (size_t)l.cin * l.kpb));
        L.push_back(l);
    }
    out_off = off;
    nparams = off + 17;
    SR_TRY(dalloc(&params, nparams));
    SR_TRY(dalloc(&grads, nparams));
    SR_TRY(dalloc(&m, nparams));
    SR_TRY(dalloc(&v, nparams));
    const size_t n = (size_t)mw * mh, n2 = n / 4, n4 = n / 16;
    SR_TRY(dalloc(&x, 8 * n));
    SR_TRY(dalloc(&e1a, 16 * n));
    SR_TRY(dalloc(&e1, 16 * n));
    SR_TRY(dalloc(&p1, 16 * n2));
    SR_TRY(dalloc(&e2a, 32 * n2));
    SR_TRY(dalloc(&e2, 32 * n2));
    SR_TRY(dalloc(&p2, 32 * n4));
    SR_TRY(dalloc(&e3a, 64 * n4));
    SR_TRY(dalloc(&e3, 64 * n4));
    SR_TRY(dalloc(&d2a, 32 * n2));
    SR_TRY(dalloc(&d2, 32 * n2));
    SR_TRY(dalloc(&d1a, 16 * n));
    SR_TRY(dalloc(&d1, 16 * n));
    SR_TRY(dalloc(&dis, n));
    SR_TRY(dalloc(&part, 2 * NB_FEAT));
    SR_TRY(dalloc(&scal, 4));
    SR_HIP(hipHostMalloc((void **)&host, sizeof(double) * 4 * NB_FEAT));
    return SR_OK;
}

int Net::set_params(const float *h, size_t n) {
    if (n != nparams) return fail(SR_ERR_LENGTH, "network parameters: expected " + std::to_string(nparams));
    SR_HIP(hipMemcpy(params, h, sizeof(float) * n, hipMemcpyHostToDevice));
    SR_HIP(hipMemset(m, 0, sizeof(float) * n));
    SR_HIP(hipMemset(v, 0, sizeof(float) * n));
    adam_t = 0;
    packed = false;
    return SR_OK;
}

int Net::get_params(float *h, size_t n) {
    if (n != nparams) return fail(SR_ERR_LENGTH, "network parameters: expected " + std::to_string(nparams));
    SR_HIP(hipMemcpy(h, params, sizeof(float) * n, hipMemcpyDeviceToHost));
    return SR_OK;
}

int Net::get_grads(float *h, size_t n) {
    if (n != nparams) return fail(SR_ERR_LENGTH, "network gradients: expected " + std::to_string(nparams));
    SR_HIP(hipMemcpy(h, grads, sizeof(float) * n, hipMemcpyDeviceToHost));
    return SR_OK;
}

int Net::pack(hipStream_t s) {
    for (auto &l : L) {
        const int total = l.cout * l.kp + (l.wb ? l.cin * l.kpb : 0);
        hipLaunchKernelGGL(k_pack, dim3((total + 255) / 256), dim3(256), 0, s, params + l.w_off, l.cin, l.cinp, l.cout,
                           l.kp, l.kpb, l.wt, l.wb);
    }
    SR_HIP(hipGetLastError());
    packed = true;
    return SR_OK;
}

static ConvArgs cargs(const bf16 *in, const bf16 *up, int cup, const bf16 *mask, const bf16 *wt, const float *bias,
                      bf16 *out, int H, int W, int relu) {
    ConvArgs a;
    a.in = in;
    a.up = up;
    a.mask = mask;
    a.wt = wt;
    a.bias = bias;
    a.out = out;
    a.H = H;
    a.W = W;
    a.cup = cup;
    a.relu = relu;
    return a;
}

static int grid_of(size_t n) { return (int)std::min<size_t>(NB_FEAT, (n + NTH - 1) / NTH); }

int Net::forward(int w, int h, const float *own, const float *kx, const float *ky, const float *r, hipStream_t s) {
    if (w <= 0 || h <= 0 || w % 4 || h % 4 || w > max_w || h > max_h || (size_t)w * h > (size_t)max_w * max_h)
        return fail(SR_ERR_ARGUMENT, "network input size outside the multiples of 4 it was created for");
    if (!own || !kx || !ky || !r) return fail(SR_ERR_ARGUMENT, "network: null system array");
    if (!packed) SR_TRY(pack(s));
    const size_t n = (size_t)w * h;
    const int nb = grid_of(n);
    nb_feat = nb;
    last_w = w;
    last_h = h;
    hipLaunchKernelGGL(k_feat_sums, dim3(nb), dim3(NTH), 0, s, w, h, own, kx, ky, r, part);
    hipLaunchKernelGGL(k_feat, dim3(nb), dim3(NTH), 0, s, w, h, own, kx, ky, r, (const double *)part, nb, x, dis, scal);
    const int w2 = w / 2, h2 = h / 2, w4 = w / 4, h4 = h / 4;
    auto B = [&](int i) { return (const float *)(params + L[i].b_off); };
    conv<8, 16>(s, cargs(x, nullptr, 0, nullptr, L[E1A].wt, B(E1A), e1a, h, w, 1));
    conv<16, 16>(s, cargs(e1a, nullptr, 0, nullptr, L[E1].wt, B(E1), e1, h, w, 1));
    hipLaunchKernelGGL(k_pool, dim3((n / 4 * 2 + NTH - 1) / NTH), dim3(NTH), 0, s, e1, h, w, 16, p1);
    conv<16, 32>(s, cargs(p1, nullptr, 0, nullptr, L[E2A].wt, B(E2A), e2a, h2, w2, 1));
    conv<32, 32>(s, cargs(e2a, nullptr, 0, nullptr, L[E2].wt, B(E2), e2, h2, w2, 1));
    hipLaunchKernelGGL(k_pool, dim3(((size_t)w4 * h4 * 4 + NTH - 1) / NTH), dim3(NTH), 0, s, e2, h2, w2, 32, p2);
    conv<32, 64>(s, cargs(p2, nullptr, 0, nullptr, L[E3A].wt, B(E3A), e3a, h4, w4, 1));
    conv<64, 64>(s, cargs(e3a, nullptr, 0, nullptr, L[E3].wt, B(E3), e3, h4, w4, 1));
    conv<96, 32>(s, cargs(e2, e3, 64, nullptr, L[D2A].wt, B(D2A), d2a, h2, w2, 1));
    conv<32, 32>(s, cargs(d2a, nullptr, 0, nullptr, L[D2].wt, B(D2), d2, h2, w2, 1));
    conv<48, 16>(s, cargs(e1, d2, 32, nullptr, L[D1A].wt, B(D1A), d1a, h, w, 1));
    conv<16, 16>(s, cargs(d1a, nullptr, 0, nullptr, L[D1].wt, B(D1), d1, h, w, 1));
    SR_HIP(hipGetLastError());
    return SR_OK;
}

int Net::apply(int w, int h, const float *own, const float *kx, const float *ky, const float *r, const float *zadd,
               float *z, hipStream_t s) {
    if (!z) return fail(SR_ERR_ARGUMENT, "network: null output");
    SR_TRY(forward(w, h, own, kx, ky, r, s));
    const size_t n = (size_t)w * h;
    hipLaunchKernelGGL(k_out, dim3(grid_of(n)), dim3(NTH), 0, s, (int)n, (const bf16 *)d1,
                       (const float *)(params + out_off), (const float *)scal, (const float *)dis, zadd, z,
                       (const float *)nullptr, (float *)nullptr);
    SR_HIP(hipGetLastError());
    return SR_OK;
}

template <int CIN, int COUT>
static void dw(hipStream_t s, const bf16 *in, const bf16 *up, int cup, const bf16 *gy, const bf16 *y, int H, int W,
               float *wpart, const NetLayer &l, float *grads) {
    const int nseg = H * ((W + DW_SEG - 1) / DW_SEG);
    const int nb = std::min(DW_BLOCKS, nseg);
    hipLaunchKernelGGL((k_dw<CIN, COUT>), dim3(nb), dim3(NTH), 0, s, in, up, cup, gy, y, H, W, wpart);
    const int nout = 9 * l.cin * l.cout + l.cout;
    hipLaunchKernelGGL(k_dw_reduce, dim3((nout + 255) / 256), dim3(256), 0, s, (const float *)wpart, nb, l.cinp,
                       l.cin, l.cout, grads + l.w_off, grads + l.b_off);
}

int Net::loss_grad(int w, int h, const float *own, const float *kx, const float *ky, const float *r, const float *e,
                   double *loss, hipStream_t s) {
    if (!e) return fail(SR_ERR_ARGUMENT, "network: null target");
    if (!g_d1) {  // training buffers, allocated at the first training call
        const size_t n = (size_t)max_w * max_h, n2 = n / 4, n4 = n / 16;
        SR_TRY(dalloc(&g_d1, 16 * n));
        SR_TRY(dalloc(&g_d1a, 16 * n));
        SR_TRY(dalloc(&g_cat1, 48 * n));
        SR_TRY(dalloc(&g_d2, 32 * n2));
        SR_TRY(dalloc(&g_d2a, 32 * n2));
        SR_TRY(dalloc(&g_cat2, 96 * n2));
        SR_TRY(dalloc(&g_e3, 64 * n4));
        SR_TRY(dalloc(&g_e3a, 64 * n4));
        SR_TRY(dalloc(&g_p2, 32 * n4));
        SR_TRY(dalloc(&g_e2, 32 * n2));
        SR_TRY(dalloc(&g_e2a, 32 * n2));
        SR_TRY(dalloc(&g_p1, 16 * n2));
        SR_TRY(dalloc(&g_e1, 16 * n));
        SR_TRY(dalloc(&g_e1a, 16 * n));
        SR_TRY(dalloc(&err, n));
        SR_TRY(dalloc(&aerr, n));
        SR_TRY(dalloc(&gout, n));
        SR_TRY(dalloc(&lpart, 2 * NB_FEAT));
        SR_TRY(dalloc(&opart, 17 * NB_FEAT));
        SR_TRY(dalloc(&wpart, (size_t)DW_BLOCKS * (9 * 64 * 64 + 64)));
    }
    SR_TRY(forward(w, h, own, kx, ky, r, s));
    const size_t n = (size_t)w * h;
    const int nb = grid_of(n);
    const int w2 = w / 2, h2 = h / 2, w4 = w / 4, h4 = h / 4;
    hipLaunchKernelGGL(k_out, dim3(nb), dim3(NTH), 0, s, (int)n, (const bf16 *)d1, (const float *)(params + out_off),
                       (const float *)scal, (const float *)dis, (const float *)nullptr, (float *)nullptr, e, err);
    hipLaunchKernelGGL(k_loss, dim3(nb), dim3(NTH), 0, s, w, h, own, kx, ky, (const float *)err, e, aerr, lpart);
    hipLaunchKernelGGL(k_out_back, dim3(nb), dim3(NTH), 0, s, (int)n, (const float *)aerr, (const double *)lpart, nb,
                       (const float *)scal, (const float *)dis, (const bf16 *)d1, (const float *)(params + out_off),
                       gout, g_d1, opart);
    hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(64), 0, s, (const double *)opart, nb, 17, grads + out_off);
    // decoder, level 1
    dw<16, 16>(s, d1a, nullptr, 0, g_d1, d1, h, w, wpart, L[D1], grads);
    conv<16, 16>(s, cargs(g_d1, nullptr, 0, d1, L[D1].wb, nullptr, g_d1a, h, w, 0));
    dw<48, 16>(s, e1, d2, 32, g_d1a, d1a, h, w, wpart, L[D1A], grads);
    conv<16, 48>(s, cargs(g_d1a, nullptr, 0, d1a, L[D1A].wb, nullptr, g_cat1, h, w, 0));
    {
        const int tot = h2 * w2 * 4 + h * w * 2;
        hipLaunchKernelGGL(k_split_cat, dim3((tot + NTH - 1) / NTH), dim3(NTH), 0, s, (const bf16 *)g_cat1, h, w, 48,
                           32, g_d2, g_e1);
    }
    // decoder, level 2
    dw<32, 32>(s, d2a, nullptr, 0, g_d2, d2, h2, w2, wpart, L[D2], grads);
    conv<32, 32>(s, cargs(g_d2, nullptr, 0, d2, L[D2].wb, nullptr, g_d2a, h2, w2, 0));
    dw<96, 32>(s, e2, e3, 64, g_d2a, d2a, h2, w2, wpart, L[D2A], grads);
    conv<32, 96>(s, cargs(g_d2a, nullptr, 0, d2a, L[D2A].wb, nullptr, g_cat2, h2, w2, 0));
    {
        const int tot = h4 * w4 * 8 + h2 * w2 * 4;
        hipLaunchKernelGGL(k_split_cat, dim3((tot + NTH - 1) / NTH), dim3(NTH), 0, s, (const bf16 *)g_cat2, h2, w2, 96,
                           64, g_e3, g_e2);
    }
    // encoder, level 3
    dw<64, 64>(s, e3a, nullptr, 0, g_e3, e3, h4, w4, wpart, L[E3], grads);
    conv<64, 64>(s, cargs(g_e3, nullptr, 0, e3, L[E3].wb, nullptr, g_e3a, h4, w4, 0));
    dw<32, 64>(s, p2, nullptr, 0, g_e3a, e3a, h4, w4, wpart, L[E3A], grads);
    conv<64, 32>(s, cargs(g_e3a, nullptr, 0, e3a, L[E3A].wb, nullptr, g_p2, h4, w4, 0));
    hipLaunchKernelGGL(k_add_pool_back, dim3((h2 * w2 * 4 + NTH - 1) / NTH), dim3(NTH), 0, s, g_e2, (const bf16 *)g_p2,
                       h2, w2, 32);
    // encoder, level 2
    dw<32, 32>(s, e2a, nullptr, 0, g_e2, e2, h2, w2, wpart, L[E2], grads);
    conv<32, 32>(s, cargs(g_e2, nullptr, 0, e2, L[E2].wb, nullptr, g_e2a, h2, w2, 0));
    dw<16, 32>(s, p1, nullptr, 0, g_e2a, e2a, h2, w2, wpart, L[E2A], grads);
    conv<32, 16>(s, cargs(g_e2a, nullptr, 0, e2a, L[E2A].wb, nullptr, g_p1, h2, w2, 0));
    hipLaunchKernelGGL(k_add_pool_back, dim3((h * w * 2 + NTH - 1) / NTH), dim3(NTH), 0, s, g_e1, (const bf16 *)g_p1, h,
                       w, 16);
    // encoder, level 1
    dw<16, 16>(s, e1a, nullptr, 0, g_e1, e1, h, w, wpart, L[E1], grads);
    conv<16, 16>(s, cargs(g_e1, nullptr, 0, e1, L[E1].wb, nullptr, g_e1a, h, w, 0));
    dw<8, 16>(s, x, nullptr, 0, g_e1a, e1a, h, w, wpart, L[E1A], grads);
    SR_HIP(hipGetLastError());
    if (loss) {
        SR_HIP(hipMemcpyAsync(host, lpart, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, s));
        SR_HIP(hipStreamSynchronize(s));
        double a = 0.0, b = 0.0;
        for (int i = 0; i < nb; i++) {
            a += host[i];
            b += host[nb + i];
        }
        *loss = b > 0.0 ? a / b : 0.0;
    }
    return SR_OK;
}

int Net::adam(double lr, hipStream_t s) {
    adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.0f - powf(b1, (float)adam_t), bc2 = 1.0f - powf(b2, (float)adam_t);
    hipLaunchKernelGGL(k_adam, dim3((nparams + 255) / 256), dim3(256), 0, s, (int)nparams, params,
                       (const float *)grads, m, v, (float)lr, b1, b2, 1e-8f, bc1, bc2);
    SR_HIP(hipGetLastError());
    packed = false;
    return SR_OK;
}

int Net::activation(const char *name, float *out, size_t count) {
    const size_t n = (size_t)last_w * last_h, n2 = n / 4, n4 = n / 16;
    struct A {
        const char *nm;
        const bf16 *p;
        size_t c;
    } t[] = {{"x", x, 8 * n},        {"e1a", e1a, 16 * n}, {"e1", e1, 16 * n},   {"p1", p1, 16 * n2},
             {"e2a", e2a, 32 * n2},  {"e2", e2, 32 * n2},  {"p2", p2, 32 * n4},  {"e3a", e3a, 64 * n4},
             {"e3", e3, 64 * n4},    {"d2a", d2a, 32 * n2}, {"d2", d2, 32 * n2}, {"d1a", d1a, 16 * n},
             {"d1", d1, 16 * n}};
    for (const A &a : t) {
        if (std::strcmp(a.nm, name) != 0) continue;
        if (count != a.c) return fail(SR_ERR_LENGTH, std::string("activation ") + name + ": expected " + std::to_string(a.c));
        std::vector<uint16_t> tmp(a.c);
        SR_HIP(hipMemcpy(tmp.data(), a.p, sizeof(uint16_t) * a.c, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < a.c; i++) {
            const uint32_t bits = (uint32_t)tmp[i] << 16;
            std::memcpy(out + i, &bits, 4);
        }
        return SR_OK;
    }
    return fail(SR_ERR_ARGUMENT, std::string("no activation named ") + name);
}

// HBM bytes of one application by its algorithm: every activation written once and read by its consumers (3x3
// reads counted once), the features' inputs, the output
double Net::bytes_per_apply(int w, int h) const {
    const double n = (double)w * h;
    const double act = 2.0 * (8 + 16 + 16 + 16 + 16) * n                        // x, e1a, e1, d1a, d1
                       + 2.0 * (16 + 32 + 32 + 32 + 32) * n / 4                 // p1, e2a, e2, d2a, d2
                       + 2.0 * (32 + 64 + 64) * n / 16;                         // p2, e3a, e3
    return 2.0 * act + (16.0 + 4.0 + 8.0) * n;
}

double Net::flops_per_apply(int w, int h) const {
    double f = 0.0;
    const double n = (double)w * h;
    const double scale[10] = {1, 1, 0.25, 0.25, 1.0 / 16, 1.0 / 16, 0.25, 0.25, 1, 1};
    for (int i = 0; i < 10; i++) f += 2.0 * 9 * L[i].cin * L[i].cout * n * scale[i];
    return f + 2.0 * 16 * n;
}

} // namespace sr

// ------------------------------------------------------------------------------------------------------------ C ABI
using namespace sr;

struct sr_net {
    Net n;
};

extern "C" int sr_net_create(int max_w, int max_h, sr_net **out) {
    if (!out) return fail(SR_ERR_ARGUMENT, "sr_net_create: null out");
    auto *p = new sr_net();
    const int c = p->n.init(max_w, max_h);
    if (c != SR_OK) {
        delete p;
        return c;
    }
    *out = p;
    return SR_OK;
}

extern "C" int sr_net_param_count(sr_net *p, size_t *n) {
    if (!p || !n) return fail(SR_ERR_ARGUMENT, "sr_net_param_count: null pointer");
    *n = p->n.param_count();
    return SR_OK;
}

extern "C" int sr_net_set_params(sr_net *p, const float *params, size_t n) {
    if (!p || !params) return fail(SR_ERR_ARGUMENT, "sr_net_set_params: null pointer");
    return p->n.set_params(params, n);
}

extern "C" int sr_net_get_params(sr_net *p, float *params, size_t n) {
    if (!p || !params) return fail(SR_ERR_ARGUMENT, "sr_net_get_params: null pointer");
    return p->n.get_params(params, n);
}

extern "C" int sr_net_get_grads(sr_net *p, float *grads, size_t n) {
    if (!p || !grads) return fail(SR_ERR_ARGUMENT, "sr_net_get_grads: null pointer");
    return p->n.get_grads(grads, n);
}

extern "C" int sr_net_apply(sr_net *p, int w, int h, const float *own, const float *kx, const float *ky,
                            const float *r, float *z, void *stream) {
    if (!p) return fail(SR_ERR_ARGUMENT, "sr_net_apply: null network");
    return p->n.apply(w, h, own, kx, ky, r, nullptr, z, (hipStream_t)stream);
}

extern "C" int sr_net_loss_grad(sr_net *p, int w, int h, const float *own, const float *kx, const float *ky,
                                const float *r, const float *e, double *loss, void *stream) {
    if (!p) return fail(SR_ERR_ARGUMENT, "sr_net_loss_grad: null network");
    return p->n.loss_grad(w, h, own, kx, ky, r, e, loss, (hipStream_t)stream);
}

extern "C" int sr_net_adam(sr_net *p, double lr, void *stream) {
    if (!p) return fail(SR_ERR_ARGUMENT, "sr_net_adam: null network");
    if (!(lr > 0.0)) return fail(SR_ERR_ARGUMENT, "sr_net_adam: the learning rate must be positive");
    return p->n.adam(lr, (hipStream_t)stream);
}

extern "C" int sr_net_activation(sr_net *p, const char *name, float *out, size_t count) {
    if (!p || !name || !out) return fail(SR_ERR_ARGUMENT, "sr_net_activation: null pointer");
    return p->n.activation(name, out, count);
}

extern "C" void sr_net_destroy(sr_net *p) { delete p; }

namespace sr {
Net *net_of(sr_net *p) { return p ? &p->n : nullptr; }
}
