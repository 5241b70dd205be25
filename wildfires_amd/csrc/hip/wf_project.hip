/*
 * wf_project.hip -- end of a time step: velocity correction, column boundary
 * conditions, bounds and temperature source in one pass
 * (src/cu/solver.cu:204-223; src/cu/pde.cu:14-137, 789-908;
 * src/cu/functions.cu:325-403).
 *
 * One thread per node writes y_{n+1} into a buffer distinct from y* and y_n,
 * so the nodes a cut or top node extrapolates from are read unmodified.
 */
#include "wf_common.h"
#include "wf_kernels.h"

namespace WF_NS {

/* dp/dz at node q for the velocity correction's difference scheme fd_z
 * (src/cu/pde.cu:831-898). */
__device__ wf_real wf_pz(const WfConst &c, const wf_real *p, long q, int k) {
    const wf_real *z = c.z;
    const long P = c.plane;
    const int Nz = c.Nz;
    int fwd, bwd;
    if (c.fd_z == 0) {
        fwd = (k == 0);
        bwd = (k == Nz - 1);
    } else if (c.fd_z == 1) {
        fwd = (k < Nz - 2);
        bwd = !fwd;
    } else if (c.fd_z == -1) {
        fwd = (k < 2);
        bwd = !fwd;
    } else {
        return WR(0.0);
    }
    if (fwd) return wf_d1f(p[q], p[q + P], p[q + 2 * P], z[k + 1] - z[k], z[k + 2] - z[k + 1]);
    if (bwd) return wf_d1b(p[q - 2 * P], p[q - P], p[q], z[k - 1] - z[k - 2], z[k] - z[k - 1]);
    return (p[q + P] - p[q - P]) / ((z[k] - z[k - 1]) + (z[k + 1] - z[k]));
}

/* u, v, w - dt grad(p) / rho(T^n) at node (i,j,k) (src/cu/pde.cu:816-906). */
__device__ void wf_corrected_velocity(const WfConst &c, const WfState &Ys, const WfState &Yn, const wf_real *p, int i,
                                      int j, int k, wf_real out[3]) {
    int im1 = wf_wrap_m(i, 1, c.Nx), ip1 = wf_wrap_p(i, 1, c.Nx);
    int jm1 = wf_wrap_m(j, 1, c.Ny), jp1 = wf_wrap_p(j, 1, c.Ny);
    long q = wf_q(c, i, j, k);
    wf_real rho = c.T_inf * c.rho_inf / Yn.T[q];
    wf_real px = (p[wf_q(c, ip1, j, k)] - p[wf_q(c, im1, j, k)]) / (2 * c.dx);
    wf_real py = (p[wf_q(c, i, jp1, k)] - p[wf_q(c, i, jm1, k)]) / (2 * c.dy);
    wf_real pz = wf_pz(c, p, q, k);
    if (c.variable_density == 0) rho = c.rho_inf;
    out[0] = Ys.u[q] - c.dt * px / rho;
    out[1] = Ys.v[q] - c.dt * py / rho;
    out[2] = Ys.w[q] - c.dt * pz / rho;
}

__global__ __launch_bounds__(256) void k_project(WfConst c, WfState Ys, WfState Yn, const wf_real *p, WfState out,
                                                 WfSource src) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (i >= c.Nx || j >= c.Ny) return;
    const wf_real *z = c.z;
    const long P = c.plane;
    const int Nz = c.Nz;
    int col = i + c.Nx * j;
    int cut = c.cut[col], nzy = c.nzy[col];
    long q = wf_q(c, i, j, k);
    wf_real u, v, w, T;
    double Y = (k < c.Nzm) ? Ys.Y[q] : 0.0; /* FP64 in both precisions */
    if (k < cut) {
        u = c.u_dead;
        v = c.v_dead;
        w = c.w_dead;
        T = c.T_dead;
        if (k < nzy) Y = c.Y_dead;
    } else if (k == cut) {
        wf_real h0 = z[k + 1] - z[k], h1 = z[k + 2] - z[k + 1];
        wf_real T1 = Ys.T[q + P], T2 = Ys.T[q + 2 * P];
        u = WR(0.0);
        v = WR(0.0);
        w = WR(0.0);
        T = ((h0 + h1) * (h0 + h1) * T1 - h0 * h0 * T2) / (h1 * (2 * h0 + h1));
        if (k < nzy) {
            const double H0 = (double)h0, H1 = (double)h1;
            double Y1 = (k + 1 < nzy) ? Ys.Y[q + P] : 0.0;
            double Y2 = (k + 2 < nzy) ? Ys.Y[q + 2 * P] : 0.0;
            Y = ((H0 + H1) * (H0 + H1) * Y1 - H0 * H0 * Y2) / (H1 * (2 * H0 + H1));
        }
    } else if (k < Nz - 1) {
        wf_real uvw[3];
        wf_corrected_velocity(c, Ys, Yn, p, i, j, k, uvw);
        u = uvw[0];
        v = uvw[1];
        w = uvw[2];
        T = Ys.T[q];
    } else {
        wf_real hm1 = z[k] - z[k - 1], hm2 = z[k - 1] - z[k - 2];
        wf_real a1[3], a2[3];
        wf_corrected_velocity(c, Ys, Yn, p, i, j, k - 1, a1);
        wf_corrected_velocity(c, Ys, Yn, p, i, j, k - 2, a2);
        wf_real den = hm2 * (hm2 + 2 * hm1);
        wf_real s1 = (hm2 + hm1) * (hm2 + hm1), s2 = hm1 * hm1;
        u = (s1 * a1[0] - s2 * a2[0]) / den;
        v = (s1 * a1[1] - s2 * a2[1]) / den;
        w = (s1 * a1[2] - s2 * a2[2]) / den;
        T = (s1 * Ys.T[q - P] - s2 * Ys.T[q - 2 * P]) / den;
    }
    /* bounds (src/cu/pde.cu:101-137) */
    if (T < c.T_min) T = c.T_min;
    if (T > c.T_max) T = c.T_max;
    if (k < c.Nzm) {
        if (Y < c.Y_min) Y = c.Y_min;
        if (Y > c.Y_max) Y = c.Y_max;
    }
    /* temperature source (src/cu/functions.cu:325-403) */
    if (src.mode == 1) {
        wf_real Ts = src.T_src[q];
        if (Ts > T) T = Ts;
    } else if (src.mode == 2) {
        wf_real shape = WR(0.0);
        wf_real xi = src.x[i], yj = src.y[j], zk = z[k];
        if (src.shape == 1)
            shape = exp(-pow((xi - src.xc) / src.sx, WR(2.0)) - pow((yj - src.yc) / src.sy, WR(2.0)) -
                        pow((zk - src.zc) / src.sz, WR(2.0)));
        else if (src.shape == 2)
            shape = (xi >= src.x0 && xi <= src.x1 && yj >= src.y0 && yj <= src.y1 && zk >= src.z0 && zk <= src.z1)
                        ? WR(1.0)
                        : WR(0.0);
        wf_real temperature = src.scale * (src.T_inf + (src.T_source - src.T_inf) * shape);
        if (temperature > T) T = temperature;
    }
    out.u[q] = u;
    out.v[q] = v;
    out.w[q] = w;
    out.T[q] = T;
    if (k < c.Nzm) out.Y[q] = Y;
}

hipError_t wf_launch_project(const WfConst &c, const WfState &Ys, const WfState &Yn, const wf_real *p,
                             const WfState &out, const WfSource &src, hipStream_t s) {
    dim3 blk(64, 4, 1), grd((c.Nx + 63) / 64, (c.Ny + 3) / 4, c.Nz);
    hipLaunchKernelGGL(k_project, grd, blk, 0, s, c, Ys, Yn, p, out, src);
    return hipGetLastError();
}

}  // namespace WF_NS
