/*
 * wf_pressure.hip -- pressure Poisson solve: FFT in x,y, tridiagonal in z.
 *
 * The reference (src/cu/pressure.cu:506-606) forms f on the (Nx-1)(Ny-1)(Nz-1)
 * unique nodes, runs a batched 2-D complex FFT per z-slab, solves one complex
 * tridiagonal system in z per (kx, ky), transforms back and keeps the real part
 * divided by (Nx-1)(Ny-1).  Here f is real, so the x-transform is a real
 * transform that keeps kx bins 0..nx/2 (two rows per complex FFT), the
 * y-transform runs on that half spectrum, and the Thomas solve runs on the
 * nx/2+1 retained kx bins only; the discarded half is the complex conjugate.
 *
 *   P1 k_rhs_fx   : f (or the top pressure plane) -> x half-spectrum   [per row pair]
 *   P2 k_fy       : y-transform of the half spectrum (forward/inverse) [per slab, kx block]
 *   P3 k_thomas   : tridiagonal solve per (kx, ky), pivots precomputed [per system]
 *   P5 k_inv_fx   : inverse x-transform, p, periodic duplicates, |dp|max [per row pair]
 *
 * The variable-density fixed-point loop is controlled on the device
 * (WfIterCtl): every kernel of a solve returns at once after convergence, so
 * the host queues solves without waiting for each norm.
 */
#include "wf_common.h"
#include "wf_fft.h"
#include "wf_kernels.h"

/* One full FFT of `nseq` sequences of run-time length n held in LDS
 * (a -> b -> a ...), for lengths without a compile-time plan.  Returns the
 * buffer that holds the result. */
__device__ wfc *wf_lds_fft(wfc *a, wfc *b, int n, int nseq, const int *rad, const wfc *tw) {
    int Ns = 1;
    for (int p = 0; p < WF_MAX_RADICES; p++) {
        int R = rad[p];
        if (R == 0) break;
        for (int t = threadIdx.x; t < n * nseq; t += blockDim.x) {
            int s = t / n, o = t - s * n;
            b[s * n + o] = wf_stockham_output(a + s * n, n, R, Ns, tw, o);
        }
        __syncthreads();
        wfc *tmp = a;
        a = b;
        b = tmp;
        Ns *= R;
    }
    return a;
}

/* p-independent part of f at node (i,j,k), 1 <= k <= Nz-2
 * (src/cu/pressure.cu:14-101 for constant density, 103-327 otherwise).
 * [Q1] h_c per node with |T_gas - T_solid|. */
__device__ double wf_f_fixed(const WfConst &c, const WfState &Ys, const WfState &Yn, int i, int j, int k) {
    const int Nx = c.Nx, Ny = c.Ny;
    const long P = c.plane;
    const double *z = c.z;
    int im1 = wf_wrap_m(i, 1, Nx), ip1 = wf_wrap_p(i, 1, Nx);
    int jm1 = wf_wrap_m(j, 1, Ny), jp1 = wf_wrap_p(j, 1, Ny);
    long q = wf_q(c, i, j, k);
    long qxm = wf_q(c, im1, j, k), qxp = wf_q(c, ip1, j, k), qym = wf_q(c, i, jm1, k), qyp = wf_q(c, i, jp1, k);
    double dz_km1 = z[k] - z[k - 1], dz_k = z[k + 1] - z[k];
    double u = Ys.u[q], v = Ys.v[q];
    double wz = (Ys.w[q + P] - Ys.w[q - P]) / (dz_k + dz_km1);
    double u_iph = 0.5 * (Ys.u[qxp] + u), u_imh = 0.5 * (u + Ys.u[qxm]);
    double v_jph = 0.5 * (Ys.v[qyp] + v), v_jmh = 0.5 * (v + Ys.v[qym]);
    double ux = (u_iph - u_imh) / c.dx, vy = (v_jph - v_jmh) / c.dy;
    if (c.variable_density == 0) return c.rho_inf * (ux + vy + wz) / c.dt;
    double T = Ys.T[q];
    double Tip = Ys.T[qxp], Tim = Ys.T[qxm], Tjp = Ys.T[qyp], Tjm = Ys.T[qym];
    double Tkp = Ys.T[q + P], Tkm = Ys.T[q - P];
    double rho = c.T_inf * c.rho_inf / Yn.T[q];
    double Tx = (Tip - Tim) / (2 * c.dx), Ty = (Tjp - Tjm) / (2 * c.dy);
    double Txx = (Tip - 2 * T + Tim) / (c.dx * c.dx), Tyy = (Tjp - 2 * T + Tjm) / (c.dy * c.dy);
    double Tz = (Tkp - Tkm) / (dz_k + dz_km1);
    double Tzz = wf_d2m(Tkm, T, Tkp, dz_km1, dz_k);
    double lap_T = Txx + Tyy + Tzz;
    double kT = c.kT, dk = 0.0;
    double Y = 0.0;
    double K_T = wf_rate(c, T);
    double T_gas = T, T_solid = T;
    int col = i + Nx * j;
    int nzy = c.nzy[col];
    if (k <= nzy) {
        if (k < nzy) Y = Ys.Y[q];
        if (Y > 0.0)
            T_gas = c.T_pc + (T - c.T_pc);
        else
            T_gas = T;
    }
    if (c.suth_k == 1) wf_sutherland_k(T, &kT, &dk);
    double qs = c.H_C * Y * K_T / c.c_p;
    if (c.exchange_on) {
        double h_c = c.h_c * dz_k * pow(fabs(T_gas - T_solid) / dz_k, 1.0 / 4.0);
        qs = qs - h_c * c.alpha_s * c.sigma_s * (T_gas - T_solid) / (c.c_p * c.T_inf * c.rho_inf / T_gas);
    }
    double T2 = T * T, T3 = T * T * T;
    double div_U_temp = (dk + 12 * WF_SIGMA_D * c.delta * T2 * (Tx * Tx + Ty * Ty + Tz * Tz) +
                         (kT + 4 * WF_SIGMA_D * c.delta * T3) * lap_T) / (c.T_inf * c.rho_inf * c.c_p) + qs / T;
    return rho * (ux + vy + wz - div_U_temp) / c.dt;
}

/* (grad rho . grad p) / rho at node (i,j,k), 1 <= k <= Nz-2, with rho = rho^n
 * the density of y_n (src/cu/pressure.cu:163-197, 269-282, 319). */
__device__ double wf_f_pgrad(const WfConst &c, const double *rho_n, const double *p, int i, int j, int k) {
    const int Nx = c.Nx, Ny = c.Ny;
    const long P = c.plane;
    const double *z = c.z;
    int im1 = wf_wrap_m(i, 1, Nx), ip1 = wf_wrap_p(i, 1, Nx);
    int jm1 = wf_wrap_m(j, 1, Ny), jp1 = wf_wrap_p(j, 1, Ny);
    long q = wf_q(c, i, j, k);
    long qxm = wf_q(c, im1, j, k), qxp = wf_q(c, ip1, j, k), qym = wf_q(c, i, jm1, k), qyp = wf_q(c, i, jp1, k);
    double dz_km1 = z[k] - z[k - 1], dz_k = z[k + 1] - z[k];
    double rho = rho_n[q];
    double rhox = (rho_n[qxp] - rho_n[qxm]) / (2 * c.dx), rhoy = (rho_n[qyp] - rho_n[qym]) / (2 * c.dy);
    double px = (p[qxp] - p[qxm]) / (2 * c.dx), py = (p[qyp] - p[qym]) / (2 * c.dy);
    double rhoz = (rho_n[q + P] - rho_n[q - P]) / (dz_k + dz_km1);
    double pz = (p[q + P] - p[q - P]) / (dz_k + dz_km1);
    return (rhox * px + rhoy * py + rhoz * pz) / rho;
}

/* rho^n = T_inf rho_inf / T^n on every node, once per step. */
__global__ void k_density(WfConst c, const double *T, double *rho) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < c.N) rho[q] = c.T_inf * c.rho_inf / T[q];
}

__global__ void k_iter_reset(WfIterCtl *ctl) {
    ctl->norm_bits = 0ull;
    ctl->done = 0;
    ctl->iters = 0;
    ctl->solves = 0;
    ctl->err = 0.0;
}

/* After solve m: err = max|p_new - p_old|, stop once err <= tol or after
 * max_iter solves (src/cu/pressure.cu:564-574); iters is the count the
 * reference reports (the index of the converging solve, else max_iter). */
__global__ void k_iter_check(WfIterCtl *ctl, double tol, int m, int max_iter) {
    if (ctl->done) return;
    double err = __longlong_as_double((long long)ctl->norm_bits);
    ctl->err = err;
    ctl->solves = m + 1;
    ctl->norm_bits = 0ull;
    if (err <= tol) {
        ctl->done = 1;
        ctl->iters = m;
    } else {
        ctl->iters = m + 1;
        if (m + 1 >= max_iter) ctl->done = 1;
    }
}

/* Wavefront max of a non-negative double, folded into *bits (IEEE order of
 * non-negative doubles equals the unsigned order of their bit patterns). */
__device__ void wf_atomic_max_nonneg(unsigned long long *bits, double v) {
    unsigned long long b = __double_as_longlong(v);
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(b, off, 64);
        b = (o > b) ? o : b;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(bits, b);
}

/* P1: f on slab k (1..nz-1) for sp.fx_pairs row pairs, forward x half-spectrum.
 * source 0: Poisson RHS (first_iter: also stores its p-independent part in sp.A);
 * source 1: the top pressure plane into sp.Ptop.  NX > 0 selects the
 * compile-time plan of that length. */
template <int NX>
__global__ __launch_bounds__(256) void k_rhs_fx(WfConst c, WfSpectral sp, WfState Ys, WfState Yn, const double *p,
                                                int first_iter, int source) {
    if (source == 0 && sp.ctl->done) return;
    extern __shared__ wfc smem[];
    __shared__ int rad[WF_MAX_RADICES];
    const int nx = (NX > 0) ? NX : sp.nx;
    const int ny = sp.ny, hx = sp.hx, pairs = sp.fx_pairs;
    wfc *tw = smem, *bufa = smem + nx, *bufb = bufa + pairs * nx;
    int k = (source == 1) ? c.Nz - 1 : 1 + blockIdx.y;
    int jpair0 = blockIdx.x * pairs;
    for (int t = threadIdx.x; t < nx; t += blockDim.x) tw[t] = *(const wfc *)&sp.twx[t];
    if (NX == 0 && threadIdx.x < WF_MAX_RADICES) rad[threadIdx.x] = sp.radx[threadIdx.x];
    for (int t = threadIdx.x; t < pairs * nx; t += blockDim.x) {
        int s = t / nx, i = t - s * nx;
        int ja = 2 * (jpair0 + s), jb = ja + 1;
        double fa = 0.0, fb = 0.0;
        if (ja < ny) {
            if (source == 1) {
                fa = p[wf_q(c, i, ja, k)];
                if (jb < ny) fb = p[wf_q(c, i, jb, k)];
            } else {
                long qa = wf_q(c, i, ja, k);
                if (first_iter || c.variable_density == 0) {
                    fa = wf_f_fixed(c, Ys, Yn, i, ja, k);
                    if (c.variable_density) sp.A[qa] = fa;
                } else {
                    fa = sp.A[qa];
                }
                if (c.variable_density) fa = fa + wf_f_pgrad(c, sp.rho, p, i, ja, k);
                if (jb < ny) {
                    long qb = wf_q(c, i, jb, k);
                    if (first_iter || c.variable_density == 0) {
                        fb = wf_f_fixed(c, Ys, Yn, i, jb, k);
                        if (c.variable_density) sp.A[qb] = fb;
                    } else {
                        fb = sp.A[qb];
                    }
                    if (c.variable_density) fb = fb + wf_f_pgrad(c, sp.rho, p, i, jb, k);
                }
            }
        }
        bufa[s * nx + i].re = fa;
        bufa[s * nx + i].im = fb;
    }
    __syncthreads();
    wfc *res;
    if constexpr (NX > 0)
        res = wf_fft_ct<NX>(bufa, bufb, pairs, tw);
    else
        res = wf_lds_fft(bufa, bufb, nx, pairs, rad, tw);
    double2 *dst = (source == 1) ? sp.Ptop : sp.F + (long)k * ny * hx;
    for (int t = threadIdx.x; t < pairs * hx; t += blockDim.x) {
        int s = t / hx, r = t - s * hx;
        int ja = 2 * (jpair0 + s), jb = ja + 1;
        if (ja >= ny) continue;
        int rm = (r == 0) ? 0 : nx - r;
        wfc A, B;
        wf_unpack_pair(res[s * nx + r], res[s * nx + rm], &A, &B);
        dst[(long)ja * hx + r] = make_double2(A.re, A.im);
        if (jb < ny) dst[(long)jb * hx + r] = make_double2(B.re, B.im);
    }
}

/* P2/P4: y-transform of sp.fy_cols kx columns over all ky, one slab. */
template <int NY>
__global__ __launch_bounds__(256) void k_fy(WfSpectral sp, double2 *buf, int slab0, int sign, int guarded) {
    if (guarded && sp.ctl->done) return;
    extern __shared__ wfc smem[];
    __shared__ int rad[WF_MAX_RADICES];
    const int ny = (NY > 0) ? NY : sp.ny;
    const int hx = sp.hx, cols = sp.fy_cols;
    wfc *tw = smem, *bufa = smem + ny, *bufb = bufa + cols * ny;
    int r0 = blockIdx.x * cols;
    long slab = (long)(slab0 + blockIdx.y) * ny * hx;
    for (int t = threadIdx.x; t < ny; t += blockDim.x) {
        double2 w = sp.twy[t];
        tw[t].re = w.x;
        tw[t].im = (sign > 0) ? -w.y : w.y;
    }
    if (NY == 0 && threadIdx.x < WF_MAX_RADICES) rad[threadIdx.x] = sp.rady[threadIdx.x];
    for (int t = threadIdx.x; t < cols * ny; t += blockDim.x) {
        int cc = t % cols, jj = t / cols;
        int r = r0 + cc;
        wfc v = {0.0, 0.0};
        if (r < hx) {
            double2 x = buf[slab + (long)jj * hx + r];
            v.re = x.x;
            v.im = x.y;
        }
        bufa[cc * ny + jj] = v;
    }
    __syncthreads();
    wfc *res;
    if constexpr (NY > 0)
        res = wf_fft_ct<NY>(bufa, bufb, cols, tw);
    else
        res = wf_lds_fft(bufa, bufb, ny, cols, rad, tw);
    for (int t = threadIdx.x; t < cols * ny; t += blockDim.x) {
        int cc = t % cols, jj = t / cols;
        int r = r0 + cc;
        if (r < hx) buf[slab + (long)jj * hx + r] = make_double2(res[cc * ny + jj].re, res[cc * ny + jj].im);
    }
}

/* Thomas pivots u_k of every (kx, ky) system (src/cu/pressure.cu:429-433):
 * u_0 = b_0, u_k = b_k - (a_{k-1}/u_{k-1}) c_{k-1}.  They do not change between
 * steps, so they are formed once. */
__global__ void k_pivots(WfSpectral sp, double *U) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    int s = blockIdx.y;
    if (r >= sp.hx) return;
    const int nz = sp.nz;
    const long slab = (long)sp.ny * sp.hx;
    long o = (long)s * sp.hx + r;
    double K2 = sp.kx2[r] + sp.ky2[s];
    double c0 = sp.c0a - 0.5 * sp.dz0 * K2 - sp.c0b;
    double u = sp.b0;
    U[o] = u;
    for (int k = 1; k < nz; k++) {
        double ck1 = (k == 1) ? c0 : sp.tc[k - 1];
        double bk = (k == nz - 1) ? -(K2 + sp.beta_top) : -(K2 + sp.beta[k]);
        double l = sp.ta[k - 1] / u;
        u = bk - l * ck1;
        U[o + k * slab] = u;
    }
}

/* P3: forward elimination and back substitution per (kx, ky), in place in sp.F
 * (src/cu/pressure.cu:418-443 with the right-hand side of 473-504). */
__global__ void k_thomas(WfSpectral sp) {
    if (sp.ctl->done) return;
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    int s = blockIdx.y;
    if (r >= sp.hx) return;
    const int nz = sp.nz;
    const long slab = (long)sp.ny * sp.hx;
    long o = (long)s * sp.hx + r;
    double2 *F = sp.F;
    double K2 = sp.kx2[r] + sp.ky2[s];
    double c0 = sp.c0a - 0.5 * sp.dz0 * K2 - sp.c0b;
    /* d_0 = (dz_0/2) f^_1 */
    double2 f1 = F[o + slab];
    double yr = 0.5 * sp.dz0 * f1.x, yi = 0.5 * sp.dz0 * f1.y;
    F[o] = make_double2(yr, yi);
    double u_prev = sp.U[o];
    for (int k = 1; k < nz; k++) {
        double2 d = F[o + k * slab];
        if (k == nz - 1) {
            double2 pt = sp.Ptop[o];
            d.x = d.x - pt.x / sp.top_div;
            d.y = d.y - pt.y / sp.top_div;
        }
        double l = sp.ta[k - 1] / u_prev;
        yr = d.x - l * yr;
        yi = d.y - l * yi;
        F[o + k * slab] = make_double2(yr, yi);
        u_prev = sp.U[o + k * slab];
    }
    double xr = yr / u_prev, xi = yi / u_prev;
    F[o + (nz - 1) * slab] = make_double2(xr, xi);
    for (int k = nz - 2; k >= 0; k--) {
        double2 y = F[o + k * slab];
        double ck = (k == 0) ? c0 : sp.tc[k];
        double uk = sp.U[o + k * slab];
        xr = (y.x - ck * xr) / uk;
        xi = (y.y - ck * xi) / uk;
        F[o + k * slab] = make_double2(xr, xi);
    }
}

/* P5: inverse x-transform of sp.fx_pairs row pairs on slab k (0..nz-1):
 * p = Re(.)/(nx ny) on unique nodes, duplicate planes i = Nx-1 and j = Ny-1 set
 * from their images [Q6], and the running max |p_new - p_old| into the
 * iteration control.  Workgroup row blockIdx.y = nz refreshes the periodic
 * copies on the top plane, which the solve never writes on unique nodes. */
template <int NX>
__global__ __launch_bounds__(256) void k_inv_fx(WfConst c, WfSpectral sp, double *p) {
    if (sp.ctl->done) return;
    extern __shared__ wfc smem[];
    __shared__ int rad[WF_MAX_RADICES];
    const int nx = (NX > 0) ? NX : sp.nx;
    const int ny = sp.ny, hx = sp.hx, Nx = c.Nx, pairs = sp.fx_pairs;
    int k = blockIdx.y;
    double dmax = 0.0;
    if (k == sp.nz) {
        if (blockIdx.x == 0) {
            int kt = c.Nz - 1;
            for (int t = threadIdx.x; t < c.Ny + c.Nx - 1; t += blockDim.x) {
                long qd, qs;
                if (t < c.Ny) { /* i = Nx-1 from i = 0 */
                    qd = wf_q(c, c.Nx - 1, t, kt);
                    qs = wf_q(c, 0, (t == c.Ny - 1) ? 0 : t, kt);
                } else { /* j = Ny-1 from j = 0 */
                    qd = wf_q(c, t - c.Ny, c.Ny - 1, kt);
                    qs = wf_q(c, t - c.Ny, 0, kt);
                }
                double v = p[qs];
                double d = fabs(v - p[qd]);
                dmax = (d > dmax) ? d : dmax;
                p[qd] = v;
            }
        }
        wf_atomic_max_nonneg(&sp.ctl->norm_bits, dmax);
        return;
    }
    wfc *tw = smem, *bufa = smem + nx, *bufb = bufa + pairs * nx;
    int jpair0 = blockIdx.x * pairs;
    const double2 *src = sp.F + (long)k * ny * hx;
    for (int t = threadIdx.x; t < nx; t += blockDim.x) {
        double2 w = sp.twx[t];
        tw[t].re = w.x;
        tw[t].im = -w.y;
    }
    if (NX == 0 && threadIdx.x < WF_MAX_RADICES) rad[threadIdx.x] = sp.radx[threadIdx.x];
    for (int t = threadIdx.x; t < pairs * nx; t += blockDim.x) {
        int s = t / nx, r = t - s * nx;
        int ja = 2 * (jpair0 + s), jb = ja + 1;
        wfc A = {0.0, 0.0}, B = {0.0, 0.0};
        int conj = r >= hx;
        int rr = conj ? nx - r : r;
        if (ja < ny) {
            double2 x = src[(long)ja * hx + rr];
            A.re = x.x;
            A.im = x.y;
            if (jb < ny) {
                double2 y = src[(long)jb * hx + rr];
                B.re = y.x;
                B.im = y.y;
            }
        }
        bufa[s * nx + r] = wf_pack_pair(A, B, conj);
    }
    __syncthreads();
    wfc *res;
    if constexpr (NX > 0)
        res = wf_fft_ct<NX>(bufa, bufb, pairs, tw);
    else
        res = wf_lds_fft(bufa, bufb, nx, pairs, rad, tw);
    double scale = (double)nx * (double)ny;
    for (int t = threadIdx.x; t < pairs * Nx; t += blockDim.x) {
        int s = t / Nx, i = t - s * Nx;
        int ja = 2 * (jpair0 + s), jb = ja + 1;
        if (ja >= ny) continue;
        int ii = (i == Nx - 1) ? 0 : i;
        double pa = res[s * nx + ii].re / scale;
        long qa = wf_q(c, i, ja, k);
        double da = fabs(pa - p[qa]);
        dmax = (da > dmax) ? da : dmax;
        p[qa] = pa;
        if (ja == 0) { /* duplicate row j = Ny-1 */
            long qd = wf_q(c, i, c.Ny - 1, k);
            double dd = fabs(pa - p[qd]);
            dmax = (dd > dmax) ? dd : dmax;
            p[qd] = pa;
        }
        if (jb < ny) {
            double pb = res[s * nx + ii].im / scale;
            long qb = wf_q(c, i, jb, k);
            double db = fabs(pb - p[qb]);
            dmax = (db > dmax) ? db : dmax;
            p[qb] = pb;
        }
    }
    wf_atomic_max_nonneg(&sp.ctl->norm_bits, dmax);
}

/* ------------------------------------------------------------- launchers ---- */

#define WF_DISPATCH_N(n, KERNEL, grd, blk, lds, s, ...)                                        \
    do {                                                                                      \
        switch (n) {                                                                          \
            case 63: hipLaunchKernelGGL(KERNEL<63>, grd, blk, lds, s, __VA_ARGS__); break;    \
            case 127: hipLaunchKernelGGL(KERNEL<127>, grd, blk, lds, s, __VA_ARGS__); break;  \
            case 255: hipLaunchKernelGGL(KERNEL<255>, grd, blk, lds, s, __VA_ARGS__); break;  \
            case 511: hipLaunchKernelGGL(KERNEL<511>, grd, blk, lds, s, __VA_ARGS__); break;  \
            default: hipLaunchKernelGGL(KERNEL<0>, grd, blk, lds, s, __VA_ARGS__); break;     \
        }                                                                                     \
    } while (0)

hipError_t wf_launch_density(const WfConst &c, const double *T, double *rho, hipStream_t s) {
    hipLaunchKernelGGL(k_density, dim3((unsigned)((c.N + 255) / 256)), dim3(256), 0, s, c, T, rho);
    return hipGetLastError();
}

hipError_t wf_launch_iter_reset(WfIterCtl *ctl, hipStream_t s) {
    hipLaunchKernelGGL(k_iter_reset, dim3(1), dim3(1), 0, s, ctl);
    return hipGetLastError();
}

hipError_t wf_launch_iter_check(WfIterCtl *ctl, double tol, int m, int max_iter, hipStream_t s) {
    hipLaunchKernelGGL(k_iter_check, dim3(1), dim3(1), 0, s, ctl, tol, m, max_iter);
    return hipGetLastError();
}

hipError_t wf_launch_pressure_rhs_fx(const WfConst &c, const WfSpectral &sp, const WfState &Ys, const WfState &Yn,
                                     const double *p, int first_iter, hipStream_t s) {
    int npairs = (sp.ny + 1) / 2;
    dim3 grd((npairs + sp.fx_pairs - 1) / sp.fx_pairs, sp.nz - 1, 1);
    size_t lds = sizeof(wfc) * (size_t)sp.nx * (1 + 2 * sp.fx_pairs);
    WF_DISPATCH_N(sp.nx, k_rhs_fx, grd, dim3(256), lds, s, c, sp, Ys, Yn, p, first_iter, 0);
    return hipGetLastError();
}

hipError_t wf_launch_ptop_fx(const WfConst &c, const WfSpectral &sp, const double *p, hipStream_t s) {
    int npairs = (sp.ny + 1) / 2;
    dim3 grd((npairs + sp.fx_pairs - 1) / sp.fx_pairs, 1, 1);
    WfState none = {0, 0, 0, 0, 0};
    size_t lds = sizeof(wfc) * (size_t)sp.nx * (1 + 2 * sp.fx_pairs);
    WF_DISPATCH_N(sp.nx, k_rhs_fx, grd, dim3(256), lds, s, c, sp, none, none, p, 0, 1);
    return hipGetLastError();
}

hipError_t wf_launch_fy(const WfSpectral &sp, double2 *buf, int slab0, int nslabs, int sign, int guarded,
                        hipStream_t s) {
    if (nslabs <= 0) return hipSuccess;
    dim3 grd((sp.hx + sp.fy_cols - 1) / sp.fy_cols, nslabs, 1);
    size_t lds = sizeof(wfc) * (size_t)sp.ny * (1 + 2 * sp.fy_cols);
    WF_DISPATCH_N(sp.ny, k_fy, grd, dim3(256), lds, s, sp, buf, slab0, sign, guarded);
    return hipGetLastError();
}

hipError_t wf_launch_pivots(const WfSpectral &sp, double *U, hipStream_t s) {
    dim3 grd((sp.hx + 63) / 64, sp.ny, 1);
    hipLaunchKernelGGL(k_pivots, grd, dim3(64), 0, s, sp, U);
    return hipGetLastError();
}

hipError_t wf_launch_thomas(const WfSpectral &sp, hipStream_t s) {
    dim3 grd((sp.hx + 63) / 64, sp.ny, 1);
    hipLaunchKernelGGL(k_thomas, grd, dim3(64), 0, s, sp);
    return hipGetLastError();
}

hipError_t wf_launch_inverse_fx(const WfConst &c, const WfSpectral &sp, double *p, hipStream_t s) {
    int npairs = (sp.ny + 1) / 2;
    dim3 grd((npairs + sp.fx_pairs - 1) / sp.fx_pairs, sp.nz + 1, 1);
    size_t lds = sizeof(wfc) * (size_t)sp.nx * (1 + 2 * sp.fx_pairs);
    WF_DISPATCH_N(sp.nx, k_inv_fx, grd, dim3(256), lds, s, c, sp, p);
    return hipGetLastError();
}
