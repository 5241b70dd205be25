/*
 * wf_reports.hip -- checkpoint reports as one reduction over the device state
 * (src/cu/functions.cu:82-185): max |u|, |v|, |w| for the CFL number, the
 * minimum and maximum of T and of Y over the nodes inside the initial fuel
 * box, and the first non-finite value in the reference's loop order; plus two
 * diagnostics of the flow itself, the resolved kinetic energy and the
 * fraction of fuel columns whose mean fuel fraction fell below one half, and
 * the fraction of interior nodes whose last derived pass used the learned
 * closure.
 *
 * Every workgroup reduces its nodes (grid-stride) and writes WF_REP_N partial
 * values; the host reduces the partials (a few thousand) in FP64.  Minima and
 * maxima do not depend on the order of reduction, so CFL and the T, Y ranges
 * equal the host computation bit for bit.
 */
#include "wf_common.h"
#include "wf_kernels.h"

namespace WF_NS {

#define WF_REP_NT 256

/* workgroup reduction of one value (op: 0 max, 1 min, 2 sum) into *out */
template <int OP>
__device__ __forceinline__ double wf_rep_reduce(double v, double *red) {
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = (OP == 0) ? ((o > v) ? o : v) : (OP == 1) ? ((o < v) ? o : v) : v + o;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = red[0];
    for (int w = 1; w < WF_REP_NT / 64; w++) {
        const double o = red[w];
        r = (OP == 0) ? ((o > r) ? o : r) : (OP == 1) ? ((o < r) ? o : r) : r + o;
    }
    __syncthreads();
    return r;
}

/* in_box: per column (col = i + Nx j), the reference's test of (x_i, y_j)
 * against the initial fuel box; part: WF_REP_N partials per workgroup;
 * nan_first: the smallest ((i Ny + j) Nz + k) 5 + f of a NaN (field f of
 * u, v, w, T, Y), initialised to ~0 by the caller. */
__global__ __launch_bounds__(WF_REP_NT) void k_reports(WfConst c, WfState X, const wf_real *mu_sgs,
                                                       const int *in_box, double *part,
                                                       unsigned long long *nan_first) {
    __shared__ double red[WF_REP_NT / 64];
    double mu = 0.0, mv = 0.0, mw = 0.0;
    double t_lo = 1e9, t_hi = -1e9, y_lo = 1e9, y_hi = -1e9;
    double ke = 0.0, burned = 0.0, fuel_cols = 0.0, learned = 0.0;
    const long N = c.N;
    for (long q = (long)blockIdx.x * WF_REP_NT + threadIdx.x; q < N; q += (long)gridDim.x * WF_REP_NT) {
        const int k = (int)(q / c.plane);
        const int col = (int)(q - (long)k * c.plane);
        const int j = col / c.Nx, i = col - j * c.Nx;
        const double u = (double)X.u[q], v = (double)X.v[q], w = (double)X.w[q], T = (double)X.T[q];
        const double Y = (k < c.Nzm) ? X.Y[q] : 0.0;
        const double vals[5] = {u, v, w, T, Y};
#pragma unroll
        for (int f = 0; f < 5; f++)
            if (isnan(vals[f]))
                atomicMin(nan_first, ((((unsigned long long)i * c.Ny + j) * c.Nz + k) * 5ull + (unsigned long long)f));
        const double au = fabs(u), av = fabs(v), aw = fabs(w);
        mu = (au > mu) ? au : mu;
        mv = (av > mv) ? av : mv;
        mw = (aw > mw) ? aw : mw;
        if (in_box[col]) {
            t_lo = (T < t_lo) ? T : t_lo;
            t_hi = (T > t_hi) ? T : t_hi;
            if (k < c.Nzm) {
                y_lo = (Y < y_lo) ? Y : y_lo;
                y_hi = (Y > y_hi) ? Y : y_hi;
            }
        }
        /* resolved kinetic energy over the unique nodes of the interior levels */
        if (i < c.Nx - 1 && j < c.Ny - 1 && k >= 1 && k <= c.Nz - 2) {
            ke += 0.5 * (u * u + v * v + w * w);
            /* the last derived pass stored a learned viscosity negative */
            learned += (mu_sgs[q] < WR(0.0)) ? 1.0 : 0.0;
        }
        /* fuel columns: mean Y over the column's fuel levels (k = 0 thread) */
        if (k == 0 && i < c.Nx - 1 && j < c.Ny - 1) {
            const int nzy = c.nzy[col];
            if (nzy > 0) {
                double s = 0.0;
                for (int kk = 0; kk < nzy; kk++) s += X.Y[(long)col + c.plane * kk];
                fuel_cols += 1.0;
                burned += (s < 0.5 * nzy) ? 1.0 : 0.0;
            }
        }
    }
    const double r[WF_REP_N] = {wf_rep_reduce<0>(mu, red),      wf_rep_reduce<0>(mv, red),
                                wf_rep_reduce<0>(mw, red),      wf_rep_reduce<1>(t_lo, red),
                                wf_rep_reduce<0>(t_hi, red),    wf_rep_reduce<1>(y_lo, red),
                                wf_rep_reduce<0>(y_hi, red),    wf_rep_reduce<2>(ke, red),
                                wf_rep_reduce<2>(burned, red),  wf_rep_reduce<2>(fuel_cols, red),
                                wf_rep_reduce<2>(learned, red)};
    if (threadIdx.x == 0) {
#pragma unroll
        for (int f = 0; f < WF_REP_N; f++) part[(long)blockIdx.x * WF_REP_N + f] = r[f];
    }
}

hipError_t wf_launch_reports(const WfConst &c, const WfState &X, const wf_real *mu_sgs, const int *in_box,
                             double *part, int nblocks, unsigned long long *nan_first, hipStream_t s) {
    hipLaunchKernelGGL(k_reports, dim3(nblocks), dim3(WF_REP_NT), 0, s, c, X, mu_sgs, in_box, part, nan_first);
    return hipGetLastError();
}

}  // namespace WF_NS
