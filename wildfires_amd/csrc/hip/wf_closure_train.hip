/*
 * wf_closure_train.hip -- training of the learned subgrid closure on the GPU
 * with the network's own BF16 MFMA forward and backward kernels (FP32
 * accumulation), Adam on FP32 master weights.  No counterpart in the
 * reference.
 *
 * Loss over a batch of rows r with inputs f_r and target viscosity nu_r:
 *   L = w sum_r (softplus(z_r) - nu_r / out_scale)^2,   w = 1 / batch.
 * Rounding points (restated by the oracle's oracle_closure_gradient): the
 * forward of wf_closure.h (bf16 inputs, weights and activations); then
 *   dz  = 2 w (yhat - t) sigmoid(z)                          FP32
 *   dh2 = dz bf16(w3) gelu'(h2);  dh1 = (W2^T bf16(dh2)) gelu'(h1)
 *   dW2 = sum bf16(dh2) a1^T,  dW1 = sum bf16(dh1) x^T       MFMA, samples as K
 *   db2, db1, dw3, db3 from the FP32 deltas                  VALU
 * The weight gradients are straight-through: the bf16 rounding of the
 * weights counts as the identity.
 *
 * Gradient kernel: one wavefront per workgroup; each pass takes 32 rows, in
 * two 16-row tiles for the forward products and as the K dimension of the
 * weight-gradient products.  Per-workgroup partial sums are reduced in a
 * fixed order, so a gradient is deterministic.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <vector>
#include "../../../include/wildfires.h"
#include "wf_closure.h"
#include "wf_kernels.h"

using namespace WF_NS; /* compiled once, in the FP64 namespace */

#define WF_CLT_NP (WF_CL_NPARAMS + 1) /* gradient entries + the loss */
#define WF_CLT_MAXWG 1024

/* per-wavefront LDS of the trainer: activations and deltas of 32 rows in row
 * ([row][unit]) and transposed ([unit][row]) order, bf16 */
struct __attribute__((aligned(16))) WfClTrainLds {
    __bf16 a1s[32 * 64], a1t[64 * 32], a2s[32 * 64];
    __bf16 dh2s[32 * 64], dh2t[64 * 32], dh1t[64 * 32];
    __bf16 xt[16 * 32];
    float dz[32];
};

__device__ __forceinline__ float wf_cl_gelu_d(float h) {
    const float s = 1.0f / (1.0f + expf(-1.702f * h));
    return s + 1.702f * h * s * (1.0f - s);
}

/* trainable entries of the parameter block */
__host__ __device__ __forceinline__ bool wf_cl_trainable(int q) { return q >= WF_CL_W1 && q <= WF_CL_B3; }

/* x: rows of 16 bf16 (12 normalised inputs, 4 zeros); t: targets nu / out_scale.
 * Rows [off, off + cnt); part: per workgroup WF_CLT_NP partial sums. */
__global__ __launch_bounds__(64) void k_cl_grad(WfClosure cl, const wf_bf16x8 *w2bt, const __bf16 *x, const float *t,
                                                long off, long cnt, float w, float *part) {
    __shared__ WfClTrainLds L;
    const int l = threadIdx.x, lr = l & 15, lg = l >> 4;
    const float *prm = cl.prm;
    const wf_bf16x8 *W1 = cl.frag + WF_CL_FRAG_W1, *W2 = cl.frag + WF_CL_FRAG_W2, *W3 = cl.frag + WF_CL_FRAG_W3;
    float w3b[4];
#pragma unroll
    for (int nt = 0; nt < 4; nt++) w3b[nt] = (float)(__bf16)prm[WF_CL_W3 + nt * 16 + lr];
    wf_f32x4 gW2[4][4], gW1[4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
        gW1[a] = wf_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int b = 0; b < 4; b++) gW2[a][b] = wf_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    }
    float gw3[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gb1[4] = {0, 0, 0, 0}, gb3 = 0.0f, loss = 0.0f;
    const long ngroups = (cnt + 31) / 32;
    for (long grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
        const long base = off + grp * 32;
        const long last = off + cnt - 1;
        /* forward */
        wf_f32x4 h1[2][4], h2[2][4];
#pragma unroll
        for (int mt = 0; mt < 2; mt++) {
            long row = base + mt * 16 + lr;
            row = (row <= last) ? row : last;
            wf_bf16x8 a = {};
            if (lg < 2) a = *(const wf_bf16x8 *)&x[row * 16 + 8 * lg];
#pragma unroll
            for (int j = 0; j < 8; j++)
                if (lg < 2) L.xt[(8 * lg + j) * 32 + mt * 16 + lr] = a[j];
#pragma unroll
            for (int nt = 0; nt < 4; nt++) {
                wf_f32x4 d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, W1[nt * 64 + l], wf_f32x4{0, 0, 0, 0}, 0, 0, 0);
                const int n = nt * 16 + lr;
                const float b = prm[WF_CL_B1 + n];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    h1[mt][nt][r] = d[r] + b;
                    const __bf16 v = (__bf16)wf_cl_gelu(h1[mt][nt][r]);
                    const int s = mt * 16 + 4 * lg + r;
                    L.a1s[s * 64 + n] = v;
                    L.a1t[n * 32 + s] = v;
                }
            }
        }
        wf_cl_wave_sync();
#pragma unroll
        for (int mt = 0; mt < 2; mt++) {
            const wf_bf16x8 aa = *(const wf_bf16x8 *)&L.a1s[(mt * 16 + lr) * 64 + 8 * lg];
            const wf_bf16x8 ab = *(const wf_bf16x8 *)&L.a1s[(mt * 16 + lr) * 64 + 32 + 8 * lg];
#pragma unroll
            for (int nt = 0; nt < 4; nt++) {
                wf_f32x4 d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aa, W2[(nt * 2) * 64 + l], wf_f32x4{0, 0, 0, 0}, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ab, W2[(nt * 2 + 1) * 64 + l], d, 0, 0, 0);
                const int n = nt * 16 + lr;
                const float b = prm[WF_CL_B2 + n];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    h2[mt][nt][r] = d[r] + b;
                    L.a2s[(mt * 16 + 4 * lg + r) * 64 + n] = (__bf16)wf_cl_gelu(h2[mt][nt][r]);
                }
            }
        }
        wf_cl_wave_sync();
#pragma unroll
        for (int mt = 0; mt < 2; mt++) {
            const wf_bf16x8 aa = *(const wf_bf16x8 *)&L.a2s[(mt * 16 + lr) * 64 + 8 * lg];
            const wf_bf16x8 ab = *(const wf_bf16x8 *)&L.a2s[(mt * 16 + lr) * 64 + 32 + 8 * lg];
            wf_f32x4 z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aa, W3[l], wf_f32x4{0, 0, 0, 0}, 0, 0, 0);
            z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ab, W3[64 + l], z, 0, 0, 0);
            if (lr == 0) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int s = mt * 16 + 4 * lg + r;
                    const long row = base + s;
                    const float zv = z[r] + prm[WF_CL_B3];
                    const float e = (row <= last) ? wf_cl_softplus(zv) - t[row] : 0.0f;
                    loss += w * e * e;
                    const float dz = 2.0f * w * e * (1.0f / (1.0f + expf(-zv)));
                    gb3 += dz;
                    L.dz[s] = dz;
                }
            }
        }
        wf_cl_wave_sync();
        /* backward: output layer and second hidden layer */
#pragma unroll
        for (int mt = 0; mt < 2; mt++) {
#pragma unroll
            for (int nt = 0; nt < 4; nt++) {
                const int n = nt * 16 + lr;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int s = mt * 16 + 4 * lg + r;
                    const float dz = L.dz[s];
                    gw3[nt] += dz * (float)L.a2s[s * 64 + n];
                    const float d = dz * w3b[nt] * wf_cl_gelu_d(h2[mt][nt][r]);
                    gb2[nt] += d;
                    const __bf16 db = (__bf16)d;
                    L.dh2s[s * 64 + n] = db;
                    L.dh2t[n * 32 + s] = db;
                }
            }
        }
        wf_cl_wave_sync();
        /* dW2 += bf16(dh2)^T a1 over the 32 rows */
#pragma unroll
        for (int mo = 0; mo < 4; mo++) {
            const wf_bf16x8 A = *(const wf_bf16x8 *)&L.dh2t[(mo * 16 + lr) * 32 + 8 * lg];
#pragma unroll
            for (int ki = 0; ki < 4; ki++) {
                const wf_bf16x8 B = *(const wf_bf16x8 *)&L.a1t[(ki * 16 + lr) * 32 + 8 * lg];
                gW2[mo][ki] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, gW2[mo][ki], 0, 0, 0);
            }
        }
        /* da1 = bf16(dh2) W2, then dh1 */
#pragma unroll
        for (int mt = 0; mt < 2; mt++) {
            const wf_bf16x8 Aa = *(const wf_bf16x8 *)&L.dh2s[(mt * 16 + lr) * 64 + 8 * lg];
            const wf_bf16x8 Ab = *(const wf_bf16x8 *)&L.dh2s[(mt * 16 + lr) * 64 + 32 + 8 * lg];
#pragma unroll
            for (int ki = 0; ki < 4; ki++) {
                wf_f32x4 d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Aa, w2bt[(ki * 2) * 64 + l], wf_f32x4{0, 0, 0, 0}, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ab, w2bt[(ki * 2 + 1) * 64 + l], d, 0, 0, 0);
                const int k = ki * 16 + lr;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int s = mt * 16 + 4 * lg + r;
                    const float dh = d[r] * wf_cl_gelu_d(h1[mt][ki][r]);
                    gb1[ki] += dh;
                    L.dh1t[k * 32 + s] = (__bf16)dh;
                }
            }
        }
        wf_cl_wave_sync();
        /* dW1 += bf16(dh1)^T x over the 32 rows (input columns 12..15 are zero) */
        const wf_bf16x8 B = *(const wf_bf16x8 *)&L.xt[lr * 32 + 8 * lg];
#pragma unroll
        for (int mo = 0; mo < 4; mo++) {
            const wf_bf16x8 A = *(const wf_bf16x8 *)&L.dh1t[(mo * 16 + lr) * 32 + 8 * lg];
            gW1[mo] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, gW1[mo], 0, 0, 0);
        }
        wf_cl_wave_sync(); /* this pass's LDS reads are done before the next pass writes */
    }
    /* partial sums of this workgroup */
    float *P = part + (long)blockIdx.x * WF_CLT_NP;
#pragma unroll
    for (int mo = 0; mo < 4; mo++) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int n = mo * 16 + 4 * lg + r;
#pragma unroll
            for (int ki = 0; ki < 4; ki++) P[WF_CL_W2 + n * 64 + ki * 16 + lr] = gW2[mo][ki][r];
            if (lr < WF_CL_NIN) P[WF_CL_W1 + n * WF_CL_NIN + lr] = gW1[mo][r];
        }
    }
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
        float a = gw3[nt], b = gb2[nt], c = gb1[nt];
        a += __shfl_xor(a, 16, 64);
        a += __shfl_xor(a, 32, 64);
        b += __shfl_xor(b, 16, 64);
        b += __shfl_xor(b, 32, 64);
        c += __shfl_xor(c, 16, 64);
        c += __shfl_xor(c, 32, 64);
        if (lg == 0) {
            P[WF_CL_W3 + nt * 16 + lr] = a;
            P[WF_CL_B2 + nt * 16 + lr] = b;
            P[WF_CL_B1 + nt * 16 + lr] = c;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        gb3 += __shfl_xor(gb3, o, 64);
        loss += __shfl_xor(loss, o, 64);
    }
    if (l == 0) {
        P[WF_CL_B3] = gb3;
        P[WF_CL_NPARAMS] = loss;
    }
}

/* grad[q] = sum over the nwg partials (trainable entries and the loss; 0 elsewhere) */
__global__ void k_cl_reduce(const float *part, int nwg, float *grad) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= WF_CLT_NP) return;
    float s = 0.0f;
    if (wf_cl_trainable(q) || q == WF_CL_NPARAMS)
        for (int b = 0; b < nwg; b++) s += part[(long)b * WF_CLT_NP + q];
    grad[q] = s;
}

/* Adam on the trainable entries (bias-corrected, step >= 1) */
__global__ void k_cl_adam(float *prm, float *m, float *v, const float *grad, float lr, float b1, float b2, float eps,
                          int step) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= WF_CL_NPARAMS || !wf_cl_trainable(q)) return;
    const float g = grad[q];
    const float mq = b1 * m[q] + (1.0f - b1) * g;
    const float vq = b2 * v[q] + (1.0f - b2) * g * g;
    m[q] = mq;
    v[q] = vq;
    const float mh = mq / (1.0f - powf(b1, (float)step)), vh = vq / (1.0f - powf(b2, (float)step));
    prm[q] -= lr * mh / (sqrtf(vh) + eps);
}

/* The MFMA B fragments of prm (the forward set of wf_closure.h, and W2 for
 * the da1 product: B[k = n][col = k_in] = W2[n][k_in], 4 column tiles x 2 K-steps). */
__global__ void k_cl_frags(const float *prm, wf_bf16x8 *frag, wf_bf16x8 *w2bt) {
    const int l = threadIdx.x & 63, f = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lr = l & 15, lg = l >> 4;
    wf_bf16x8 o;
    if (f < 4) { /* W1, output tile f */
        const int n = f * 16 + lr;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int k = 8 * lg + j;
            o[j] = (__bf16)((k < WF_CL_NIN) ? prm[WF_CL_W1 + n * WF_CL_NIN + k] : 0.0f);
        }
        frag[WF_CL_FRAG_W1 + f * 64 + l] = o;
    } else if (f < 12) { /* W2, output tile nt, K-step ks */
        const int nt = (f - 4) >> 1, ks = (f - 4) & 1, n = nt * 16 + lr;
#pragma unroll
        for (int j = 0; j < 8; j++) o[j] = (__bf16)prm[WF_CL_W2 + n * WF_CL_NH + ks * 32 + 8 * lg + j];
        frag[WF_CL_FRAG_W2 + (f - 4) * 64 + l] = o;
    } else if (f < 14) { /* w3 as column 0 */
        const int ks = f - 12;
#pragma unroll
        for (int j = 0; j < 8; j++) o[j] = (__bf16)((lr == 0) ? prm[WF_CL_W3 + ks * 32 + 8 * lg + j] : 0.0f);
        frag[WF_CL_FRAG_W3 + ks * 64 + l] = o;
    } else if (f < 22) { /* W2 for da1: column tile ki, K-step ks */
        const int ki = (f - 14) >> 1, ks = (f - 14) & 1, k = ki * 16 + lr;
#pragma unroll
        for (int j = 0; j < 8; j++) o[j] = (__bf16)prm[WF_CL_W2 + (ks * 32 + 8 * lg + j) * WF_CL_NH + k];
        w2bt[(f - 14) * 64 + l] = o;
    }
}

/* ---------------------------------------------------------------- C-ABI ---- */

struct wf_closure_trainer {
    long n;
    float *prm, *m, *v, *part, *grad;
    wf_bf16x8 *frag, *w2bt;
    __bf16 *x;
    float *t;
    int step;
};

#define WF_CLT_HIP(call)                                                                         \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) return wf_set_error(WF_E_DEVICE, hipGetErrorString(e_));           \
    } while (0)

static int wf_clt_frags(wf_closure_trainer *T) {
    hipLaunchKernelGGL(k_cl_frags, dim3(22), dim3(64), 0, 0, T->prm, T->frag, T->w2bt);
    WF_CLT_HIP(hipGetLastError());
    return WF_OK;
}

/* one gradient over rows [off, off + cnt) with weight w into T->grad */
static int wf_clt_grad(wf_closure_trainer *T, long off, long cnt, float w) {
    const long groups = (cnt + 31) / 32;
    const int nwg = (int)((groups < WF_CLT_MAXWG) ? groups : WF_CLT_MAXWG);
    hipLaunchKernelGGL(k_cl_grad, dim3(nwg), dim3(64), 0, 0, WfClosure{T->prm, T->frag}, T->w2bt, T->x, T->t, off,
                       cnt, w, T->part);
    WF_CLT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_cl_reduce, dim3((WF_CLT_NP + 255) / 256), dim3(256), 0, 0, T->part, nwg, T->grad);
    WF_CLT_HIP(hipGetLastError());
    return WF_OK;
}

extern "C" int wf_closure_trainer_destroy(wf_closure_trainer *T) {
    if (T == NULL) return WF_OK;
    (void)hipDeviceSynchronize();
    void *b[] = {T->prm, T->m, T->v, T->part, T->grad, T->frag, T->w2bt, T->x, T->t};
    for (void *p : b)
        if (p) (void)hipFree(p);
    free(T);
    return WF_OK;
}

extern "C" int wf_closure_trainer_create(const float *params, long nparams, const double *features,
                                         const double *nu_target, long n, wf_closure_trainer **out) {
    if (out == NULL || params == NULL || features == NULL || nu_target == NULL || n <= 0)
        return wf_set_error(WF_E_ARG, "closure trainer: null argument or empty data");
    *out = NULL;
    if (nparams != WF_CL_NPARAMS) return wf_set_error(WF_E_ARG, "closure trainer: wrong parameter count");
    for (long q = 0; q < nparams; q++)
        if (!isfinite(params[q])) return wf_set_error(WF_E_ARG, "closure trainer: parameter not finite");
    if (!(params[WF_CL_SCALE] > 0.0f)) return wf_set_error(WF_E_ARG, "closure trainer: out_scale must be positive");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return wf_set_error(WF_E_DEVICE, "no HIP device");
    WF_CLT_HIP(hipSetDevice(0));
    /* normalised bf16 inputs (the rounding of wf_cl_inputs) and targets */
    std::vector<uint16_t> xb((size_t)n * 16, 0);
    std::vector<float> tb(n);
    for (long r = 0; r < n; r++) {
        for (int i = 0; i < WF_CL_NIN; i++) {
            const double f = features[r * WF_CL_NIN + i];
            xb[(size_t)r * 16 + i] =
                wf_bf16_bits((float)((f - (double)params[WF_CL_MEAN + i]) * (double)params[WF_CL_ISTD + i]));
        }
        tb[r] = (float)(nu_target[r] / (double)params[WF_CL_SCALE]);
    }
    wf_closure_trainer *T = (wf_closure_trainer *)calloc(1, sizeof(wf_closure_trainer));
    if (T == NULL) return wf_set_error(WF_E_ALLOC, "closure trainer: host allocation");
    T->n = n;
    hipError_t e = hipMalloc((void **)&T->prm, sizeof(float) * WF_CL_NPARAMS);
    if (e == hipSuccess) e = hipMalloc((void **)&T->m, sizeof(float) * WF_CL_NPARAMS);
    if (e == hipSuccess) e = hipMalloc((void **)&T->v, sizeof(float) * WF_CL_NPARAMS);
    if (e == hipSuccess) e = hipMalloc((void **)&T->part, sizeof(float) * WF_CLT_NP * WF_CLT_MAXWG);
    if (e == hipSuccess) e = hipMalloc((void **)&T->grad, sizeof(float) * WF_CLT_NP);
    if (e == hipSuccess) e = hipMalloc((void **)&T->frag, sizeof(wf_bf16x8) * WF_CL_NFRAG);
    if (e == hipSuccess) e = hipMalloc((void **)&T->w2bt, sizeof(wf_bf16x8) * 8 * 64);
    if (e == hipSuccess) e = hipMalloc((void **)&T->x, sizeof(uint16_t) * 16 * (size_t)n);
    if (e == hipSuccess) e = hipMalloc((void **)&T->t, sizeof(float) * n);
    if (e == hipSuccess) e = hipMemcpy(T->prm, params, sizeof(float) * WF_CL_NPARAMS, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(T->m, 0, sizeof(float) * WF_CL_NPARAMS);
    if (e == hipSuccess) e = hipMemset(T->v, 0, sizeof(float) * WF_CL_NPARAMS);
    if (e == hipSuccess) e = hipMemcpy(T->x, xb.data(), sizeof(uint16_t) * 16 * (size_t)n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(T->t, tb.data(), sizeof(float) * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        wf_closure_trainer_destroy(T);
        return wf_set_error(WF_E_DEVICE, hipGetErrorString(e));
    }
    int rc = wf_clt_frags(T);
    if (rc) {
        wf_closure_trainer_destroy(T);
        return rc;
    }
    *out = T;
    return WF_OK;
}

extern "C" int wf_closure_trainer_gradient(wf_closure_trainer *T, long offset, long count, double weight,
                                           float *grad, double *loss) {
    if (T == NULL || grad == NULL || loss == NULL || offset < 0 || count <= 0 || offset + count > T->n)
        return wf_set_error(WF_E_ARG, "closure trainer: bad row range");
    int rc = wf_clt_grad(T, offset, count, (float)weight);
    if (rc) return rc;
    std::vector<float> g(WF_CLT_NP);
    WF_CLT_HIP(hipMemcpy(g.data(), T->grad, sizeof(float) * WF_CLT_NP, hipMemcpyDeviceToHost));
    memcpy(grad, g.data(), sizeof(float) * WF_CL_NPARAMS);
    *loss = g[WF_CL_NPARAMS];
    return WF_OK;
}

extern "C" int wf_closure_trainer_train(wf_closure_trainer *T, int steps, long batch, double lr, double *last_loss) {
    if (T == NULL || steps < 0 || batch <= 0 || batch > T->n || !(lr > 0.0))
        return wf_set_error(WF_E_ARG, "closure trainer: bad training arguments");
    long off = ((long)T->step * batch) % T->n;
    for (int s = 0; s < steps; s++) {
        if (off + batch > T->n) off = 0;
        int rc = wf_clt_grad(T, off, batch, 1.0f / (float)batch);
        if (rc) return rc;
        T->step++;
        hipLaunchKernelGGL(k_cl_adam, dim3((WF_CL_NPARAMS + 255) / 256), dim3(256), 0, 0, T->prm, T->m, T->v, T->grad,
                           (float)lr, 0.9f, 0.999f, 1e-8f, T->step);
        WF_CLT_HIP(hipGetLastError());
        rc = wf_clt_frags(T);
        if (rc) return rc;
        off += batch;
    }
    if (last_loss) {
        float L = 0.0f;
        WF_CLT_HIP(hipMemcpy(&L, T->grad + WF_CL_NPARAMS, sizeof(float), hipMemcpyDeviceToHost));
        *last_loss = L;
    } else {
        WF_CLT_HIP(hipDeviceSynchronize());
    }
    return WF_OK;
}

extern "C" int wf_closure_trainer_params(wf_closure_trainer *T, float *params) {
    if (T == NULL || params == NULL) return wf_set_error(WF_E_ARG, "closure trainer: null argument");
    WF_CLT_HIP(hipMemcpy(params, T->prm, sizeof(float) * WF_CL_NPARAMS, hipMemcpyDeviceToHost));
    return WF_OK;
}
