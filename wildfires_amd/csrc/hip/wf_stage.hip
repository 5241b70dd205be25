/*
 * wf_stage.hip -- one right-hand-side evaluation per RK stage, fused with the
 * SGS divergence, the immersed-boundary/top conditions and the RK combination.
 *
 * Pass 1 (k_derived) evaluates, per node, the quantities the SGS divergence
 * reads at a neighbour: mu_sgs, the six distinct entries of the modified strain
 * tensor, div_U, and u_tau on the k = 0 plane (src/cu/pde.cu:632-772).
 * Pass 2 (k_stage) evaluates the full RHS of u, v, w, T, Y at a node
 * (src/cu/pde.cu:139-787 plus src/cu/turbulence.cu:255-454), applies the
 * column conditions of src/cu/pde.cu:14-99 by evaluating the raw RHS of the
 * two nodes above a cut node / below the top node in the same thread, and
 * writes the RK combination (src/cu/utils.cu:14-21, src/cu/solver.cu:14-39)
 * instead of storing the stage derivative.
 */
#include "wf_common.h"
#include "wf_kernels.h"

/* ---------------------------------------------------------------- pass 1 ---- */

__global__ __launch_bounds__(256) void k_derived(WfConst c, WfState X, WfDerived D) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (i >= c.Nx || j >= c.Ny) return;
    const int Nx = c.Nx, Ny = c.Ny, Nz = c.Nz;
    const double *z = c.z;
    int im1 = wf_wrap_m(i, 1, Nx), ip1 = wf_wrap_p(i, 1, Nx);
    int jm1 = wf_wrap_m(j, 1, Ny), jp1 = wf_wrap_p(j, 1, Ny);
    long q = wf_q(c, i, j, k);
    long qxm = wf_q(c, im1, j, k), qxp = wf_q(c, ip1, j, k);
    long qym = wf_q(c, i, jm1, k), qyp = wf_q(c, i, jp1, k);
    const long P = c.plane;
    double u = X.u[q], v = X.v[q], w = X.w[q], T = X.T[q];
    double ux = (X.u[qxp] - X.u[qxm]) / (2 * c.dx), uy = (X.u[qyp] - X.u[qym]) / (2 * c.dy);
    double vx = (X.v[qxp] - X.v[qxm]) / (2 * c.dx), vy = (X.v[qyp] - X.v[qym]) / (2 * c.dy);
    double wx = (X.w[qxp] - X.w[qxm]) / (2 * c.dx), wy = (X.w[qyp] - X.w[qym]) / (2 * c.dy);
    double Txm = X.T[qxm], Txp = X.T[qxp], Tym = X.T[qym], Typ = X.T[qyp];
    double Tx = (Txp - Txm) / (2.0 * c.dx), Ty = (Typ - Tym) / (2.0 * c.dy);
    double Txx = (Txp - 2.0 * T + Txm) / (c.dx * c.dx);
    double Tyy = (Typ - 2.0 * T + Tym) / (c.dy * c.dy);
    double uz, vz, wz, Tz, Tzz, T_kp1 = 0.0;
    if (k == 0) {
        double h0 = z[1] - z[0], h1 = z[2] - z[1], h2 = z[3] - z[2];
        uz = wf_d1f(u, X.u[q + P], X.u[q + 2 * P], h0, h1);
        vz = wf_d1f(v, X.v[q + P], X.v[q + 2 * P], h0, h1);
        wz = wf_d1f(w, X.w[q + P], X.w[q + 2 * P], h0, h1);
        T_kp1 = X.T[q + P];
        double T2 = X.T[q + 2 * P], T3 = X.T[q + 3 * P];
        Tz = wf_d1f(T, T_kp1, T2, h0, h1);
        Tzz = wf_d2f(T, T_kp1, T2, T3, h0, h1, h2);
    } else if (k == Nz - 1) {
        double hm1 = z[k] - z[k - 1], hm2 = z[k - 1] - z[k - 2], hm3 = z[k - 2] - z[k - 3];
        uz = wf_d1b(X.u[q - 2 * P], X.u[q - P], u, hm2, hm1);
        vz = wf_d1b(X.v[q - 2 * P], X.v[q - P], v, hm2, hm1);
        wz = wf_d1b(X.w[q - 2 * P], X.w[q - P], w, hm2, hm1);
        double Tm1 = X.T[q - P], Tm2 = X.T[q - 2 * P], Tm3 = X.T[q - 3 * P];
        Tz = wf_d1b(Tm2, Tm1, T, hm2, hm1);
        Tzz = wf_d2b(Tm3, Tm2, Tm1, T, hm3, hm2, hm1);
    } else {
        double hm1 = z[k] - z[k - 1], h0 = z[k + 1] - z[k];
        uz = (X.u[q + P] - X.u[q - P]) / (hm1 + h0);
        vz = (X.v[q + P] - X.v[q - P]) / (hm1 + h0);
        wz = (X.w[q + P] - X.w[q - P]) / (hm1 + h0);
        T_kp1 = X.T[q + P];
        double Tm1 = X.T[q - P];
        Tz = (T_kp1 - Tm1) / (hm1 + h0);
        Tzz = wf_d2m(Tm1, T, T_kp1, hm1, h0);
    }
    double lap_T = Txx + Tyy + Tzz;
    double rho = c.T_inf * c.rho_inf / T;
    double s11 = 2 * ux / 3 - (vy + wz) / 3;
    double s22 = 2 * vy / 3 - (ux + wz) / 3;
    double s33 = 2 * wz / 3 - (ux + vy) / 3;
    double s12 = 0.5 * (uy + vx);
    double s13 = 0.5 * (uz + wx);
    double s23 = 0.5 * (wy + vz);
    double mod_S = sqrt(2.0 * (ux * ux + vy * vy + wz * wz) + (uz + wx) * (uz + wx) + (vx + uy) * (vx + uy) +
                        (wy + vz) * (wy + vz));
    double mu_sgs = rho * c.cd2[k] * mod_S;
    /* T_RHS_tmp of src/cu/pde.cu:706 (needed for div_U) */
    int col = i + Nx * j;
    double mu = c.mu, kT = c.kT, dk = 0.0;
    double q_src = wf_heat_source(c, col, i, j, k, T, T_kp1, X.Y);
    if (c.suth_mu == 1) mu = 1.716e-5 * pow(T / 273.15, 1.5) * (273.15 + 110.4) / (T + 110.4);
    if (c.suth_k == 1) wf_sutherland_k(T, &kT, &dk);
    double nu = mu / rho;
    double T2 = T * T, T3 = T * T * T;
    double T_RHS_tmp = (dk + 12 * WF_SIGMA_D * c.delta * T2 * (Tx * Tx + Ty * Ty + Tz * Tz) +
                        (kT + 4 * WF_SIGMA_D * c.delta * T3) * lap_T) / (rho * c.c_p) + q_src;
    D.mu[q] = mu_sgs;
    D.s11[q] = s11;
    D.s12[q] = s12;
    D.s13[q] = s13;
    D.s22[q] = s22;
    D.s23[q] = s23;
    D.s33[q] = s33;
    D.divU[q] = nu * T_RHS_tmp / T;
    if (k == 0) {
        double a1 = mu * (uz + wx * 0.0), a2 = mu * (vz + wy * 0.0);
        double tau_w = sqrt(a1 * a1 + a2 * a2);
        D.utau[col] = sqrt(tau_w / rho);
    }
}

/* ---------------------------------------------------------------- pass 2 ---- */

/* Raw RHS of u, v, w, T (RHS + SGS, before any boundary condition) at an
 * interior plane 1 <= k <= Nz-2. */
__device__ void wf_rhs_interior(const WfConst &c, const WfState &X, const WfDerived &D, int i, int j, int k,
                                double out[4]) {
    const int Nx = c.Nx, Ny = c.Ny, Nz = c.Nz;
    const double *z = c.z;
    const long P = c.plane;
    int im1 = wf_wrap_m(i, 1, Nx), im2 = wf_wrap_m(i, 2, Nx), ip1 = wf_wrap_p(i, 1, Nx), ip2 = wf_wrap_p(i, 2, Nx);
    int jm1 = wf_wrap_m(j, 1, Ny), jm2 = wf_wrap_m(j, 2, Ny), jp1 = wf_wrap_p(j, 1, Ny), jp2 = wf_wrap_p(j, 2, Ny);
    long q = wf_q(c, i, j, k);
    long qxm1 = wf_q(c, im1, j, k), qxm2 = wf_q(c, im2, j, k), qxp1 = wf_q(c, ip1, j, k), qxp2 = wf_q(c, ip2, j, k);
    long qym1 = wf_q(c, i, jm1, k), qym2 = wf_q(c, i, jm2, k), qyp1 = wf_q(c, i, jp1, k), qyp2 = wf_q(c, i, jp2, k);
    const double *F[4] = {X.u, X.v, X.w, X.T};
    double cc[4], xm1[4], xp1[4], ym1[4], yp1[4], zm1[4], zp1[4];
    double d_im[4], d_ip[4], d_jm[4], d_jp[4], d_km[4], d_kp[4], dxx[4], dyy[4], dz1[4], dzz[4];
    double hm1 = z[k] - z[k - 1], h0 = z[k + 1] - z[k];
#pragma unroll
    for (int f = 0; f < 4; f++) {
        const double *A = F[f];
        double a0 = A[q];
        double am1 = A[qxm1], am2 = A[qxm2], ap1 = A[qxp1], ap2 = A[qxp2];
        double bm1 = A[qym1], bm2 = A[qym2], bp1 = A[qyp1], bp2 = A[qyp2];
        double cm1 = A[q - P], cp1 = A[q + P];
        cc[f] = a0;
        xm1[f] = am1;
        xp1[f] = ap1;
        ym1[f] = bm1;
        yp1[f] = bp1;
        zm1[f] = cm1;
        zp1[f] = cp1;
        d_im[f] = (3 * a0 - 4 * am1 + am2) / (2 * c.dx);
        d_ip[f] = (-3 * a0 + 4 * ap1 - ap2) / (2 * c.dx);
        d_jm[f] = (3 * a0 - 4 * bm1 + bm2) / (2 * c.dy);
        d_jp[f] = (-3 * a0 + 4 * bp1 - bp2) / (2 * c.dy);
        if (k <= 1) {
            double g1 = z[k + 2] - z[k + 1];
            d_km[f] = wf_d1f(a0, cp1, A[q + 2 * P], h0, g1);
        } else {
            double g2 = z[k - 1] - z[k - 2];
            d_km[f] = wf_d1b(A[q - 2 * P], cm1, a0, g2, hm1);
        }
        if (k >= Nz - 2) {
            double g2 = z[k - 1] - z[k - 2];
            d_kp[f] = wf_d1b(A[q - 2 * P], cm1, a0, g2, hm1);
        } else {
            double g1 = z[k + 2] - z[k + 1];
            d_kp[f] = wf_d1f(a0, cp1, A[q + 2 * P], h0, g1);
        }
        dxx[f] = (ap1 - 2.0 * a0 + am1) / (c.dx * c.dx);
        dyy[f] = (bp1 - 2.0 * a0 + bm1) / (c.dy * c.dy);
        dz1[f] = (cp1 - cm1) / (hm1 + h0);
        dzz[f] = wf_d2m(cm1, a0, cp1, hm1, h0);
    }
    double u = cc[0], v = cc[1], w = cc[2], T = cc[3];
    double rho = c.T_inf * c.rho_inf / T;
    double u_plu = (u > 0.0) ? u : 0.0, u_min = (u < 0.0) ? u : 0.0;
    double v_plu = (v > 0.0) ? v : 0.0, v_min = (v < 0.0) ? v : 0.0;
    double w_plu = (w > 0.0) ? w : 0.0, w_min = (w < 0.0) ? w : 0.0;
    double uux = u_plu * d_im[0] + u_min * d_ip[0];
    double vuy = v_plu * d_jm[0] + v_min * d_jp[0];
    double wuz = w_plu * d_km[0] + w_min * d_kp[0];
    double uvx = u_plu * d_im[1] + u_min * d_ip[1];
    double vvy = v_plu * d_jm[1] + v_min * d_jp[1];
    double wvz = w_plu * d_km[1] + w_min * d_kp[1];
    double uwx = u_plu * d_im[2] + u_min * d_ip[2];
    double vwy = v_plu * d_jm[2] + v_min * d_jp[2];
    double wwz = w_plu * d_km[2] + w_min * d_kp[2];
    double uTx = u_plu * d_im[3] + u_min * d_ip[3];
    double vTy = v_plu * d_jm[3] + v_min * d_jp[3];
    double wTz = w_plu * d_km[3] + w_min * d_kp[3];
    double Tx = (xp1[3] - xm1[3]) / (2.0 * c.dx), Ty = (yp1[3] - ym1[3]) / (2.0 * c.dy), Tz = dz1[3];
    double Txx = dxx[3], Tyy = dyy[3], Tzz = dzz[3];
    double lap_u = dxx[0] + dyy[0] + dzz[0];
    double lap_v = dxx[1] + dyy[1] + dzz[1];
    double lap_w = dxx[2] + dyy[2] + dzz[2];
    double lap_T = dxx[3] + dyy[3] + dzz[3];
    int col = i + Nx * j;
    double mu = c.mu, kT = c.kT, dk = 0.0;
    double Y = 0.0;
    double q_src = wf_heat_source_y(c, col, i, j, k, T, zp1[3], X.Y, &Y);
    if (c.suth_mu == 1) mu = 1.716e-5 * pow(T / 273.15, 1.5) * (273.15 + 110.4) / (T + 110.4);
    if (c.suth_k == 1) wf_sutherland_k(T, &kT, &dk);
    double nu = mu / rho;
    double T2 = T * T, T3 = T * T * T;
    double T_RHS_tmp = (dk + 12 * WF_SIGMA_D * c.delta * T2 * (Tx * Tx + Ty * Ty + Tz * Tz) +
                        (kT + 4 * WF_SIGMA_D * c.delta * T3) * lap_T) / (rho * c.c_p) + q_src;
    double T_RHS;
    if (c.temperature_convection == 0)
        T_RHS = T_RHS_tmp - (u * Tx + v * Ty + w * Tz);
    else
        T_RHS = T_RHS_tmp - (uTx + vTy + wTz);
    double mod_U = sqrt(u * u + v * v + w * w);
    double F_x = -0.5 * c.C_d * c.alpha_s * c.sigma_s * Y * mod_U * u;
    double F_y = -0.5 * c.C_d * c.alpha_s * c.sigma_s * Y * mod_U * v;
    double F_z = c.g * (rho - c.rho_inf) / rho - 0.5 * c.C_d * c.alpha_s * c.sigma_s * Y * mod_U * w;
    double u_RHS = nu * lap_u - (uux + vuy + wuz) + F_x;
    double v_RHS = nu * lap_v - (uvx + vvy + wvz) + F_y;
    double w_RHS = nu * lap_w - (uwx + vwy + wwz) + F_z;

    /* SGS divergence (src/cu/turbulence.cu:291-452) */
    double nu_t = c.mu / rho;
    double ut = D.utau[col];
    double ut_ip = D.utau[ip1 + Nx * j], ut_im = D.utau[im1 + Nx * j];
    double ut_jp = D.utau[i + Nx * jp1], ut_jm = D.utau[i + Nx * jm1];
    double fw = wf_fw(wf_zibm(c, col, k), ut, nu_t);
    double fw_ip = wf_fw(wf_zibm(c, ip1 + Nx * j, k), ut_ip, nu_t);
    double fw_im = wf_fw(wf_zibm(c, im1 + Nx * j, k), ut_im, nu_t);
    double fw_jp = wf_fw(wf_zibm(c, i + Nx * jp1, k), ut_jp, nu_t);
    double fw_jm = wf_fw(wf_zibm(c, i + Nx * jm1, k), ut_jm, nu_t);
    double fw_km = wf_fw(wf_zibm(c, col, k - 1), ut, nu_t);
    double fw_kp = wf_fw(wf_zibm(c, col, k + 1), ut, nu_t);
    double mu_c = D.mu[q];
    double mu_ip = D.mu[qxp1], mu_im = D.mu[qxm1], mu_jp = D.mu[qyp1], mu_jm = D.mu[qym1];
    double mu_km = D.mu[q - P], mu_kp = D.mu[q + P];
    mu_c *= fw * fw;
    mu_im *= fw_im * fw_im;
    mu_ip *= fw_ip * fw_ip;
    mu_jm *= fw_jm * fw_jm;
    mu_jp *= fw_jp * fw_jp;
    mu_km *= fw_km * fw_km;
    mu_kp *= fw_kp * fw_kp;
    double S11 = D.s11[q], S12 = D.s12[q], S13 = D.s13[q], S22 = D.s22[q], S23 = D.s23[q], S33 = D.s33[q];
    double mu_x = (mu_ip - mu_im) / (2 * c.dx);
    double mu_y = (mu_jp - mu_jm) / (2 * c.dy);
    double S11_x = (D.s11[qxp1] - D.s11[qxm1]) / (2.0 * c.dx);
    double S12_x = (D.s12[qxp1] - D.s12[qxm1]) / (2.0 * c.dx);
    double S13_x = (D.s13[qxp1] - D.s13[qxm1]) / (2.0 * c.dx);
    double S21_y = (D.s12[qyp1] - D.s12[qym1]) / (2.0 * c.dy);
    double S22_y = (D.s22[qyp1] - D.s22[qym1]) / (2.0 * c.dy);
    double S23_y = (D.s23[qyp1] - D.s23[qym1]) / (2.0 * c.dy);
    double dU_x = (D.divU[qxp1] - D.divU[qxm1]) / (2.0 * c.dx);
    double dU_y = (D.divU[qyp1] - D.divU[qym1]) / (2.0 * c.dy);
    double mu_z = (mu_kp - mu_km) / (h0 + hm1);
    double S31_z = (D.s13[q + P] - D.s13[q - P]) / (h0 + hm1);
    double S32_z = (D.s23[q + P] - D.s23[q - P]) / (h0 + hm1);
    double S33_z = (D.s33[q + P] - D.s33[q - P]) / (h0 + hm1);
    double dU_z = (D.divU[q + P] - D.divU[q - P]) / (h0 + hm1);
    double div_S_x = S11_x + S21_y + S31_z;
    double div_S_y = S12_x + S22_y + S32_z;
    double div_S_z = S13_x + S23_y + S33_z;
    double sgs_x = -2 * (mu_x * S11 + mu_y * S12 + mu_z * S13 + mu_c * div_S_x) / rho;
    double sgs_y = -2 * (mu_x * S12 + mu_y * S22 + mu_z * S23 + mu_c * div_S_y) / rho;
    double sgs_z = -2 * (mu_x * S13 + mu_y * S23 + mu_z * S33 + mu_c * div_S_z) / rho;
    double sgs_q = -(mu_x * Tx + mu_y * Ty + mu_z * Tz + mu_c * (Txx + Tyy + Tzz)) / (rho * c.Pr);
    out[0] = u_RHS + (-sgs_x + dU_x / 3.0);
    out[1] = v_RHS + (-sgs_y + dU_y / 3.0);
    out[2] = w_RHS + (-sgs_z + dU_z / 3.0);
    out[3] = T_RHS - sgs_q;
}

/* Raw fuel RHS -Y_f K(T) Y where the column stores fuel (src/cu/pde.cu:672-675);
 * zero elsewhere. */
__device__ __forceinline__ double wf_rhs_fuel(const WfConst &c, const WfState &X, int col, long q, int k, int nzy) {
    if (k >= nzy) return 0.0;
    double T = X.T[q];
    double Y = X.Y[q];
    return -c.Y_f * wf_rate(c, T) * Y;
}

/* Extrapolation giving a zero one-sided z-derivative at the bottom of the fluid
 * (src/cu/pde.cu:62) and at the top (src/cu/pde.cu:84-87). */
__device__ __forceinline__ double wf_extrap_bottom(double f1, double f2, double h0, double h1) {
    return ((h0 + h1) * (h0 + h1) * f1 - h0 * h0 * f2) / (h1 * (2 * h0 + h1));
}
__device__ __forceinline__ double wf_extrap_top(double fm1, double fm2, double hm2, double hm1) {
    return ((hm2 + hm1) * (hm2 + hm1) * fm1 - hm1 * hm1 * fm2) / (hm2 * (hm2 + 2 * hm1));
}

/* One stage.  mode: WF_MODE_PHI writes the stage derivative to `out`;
 * otherwise `out` receives the next stage input (or y_{n+1}) and `acc` the
 * running combination. */
__global__ __launch_bounds__(256) void k_stage(WfConst c, WfState X, WfDerived D, WfState Yn, WfState acc,
                                               WfState out, int mode, double a) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (i >= c.Nx || j >= c.Ny) return;
    const int Nz = c.Nz;
    const long P = c.plane;
    int col = i + c.Nx * j;
    int cut = c.cut[col], nzy = c.nzy[col];
    long q = wf_q(c, i, j, k);
    double r[5];
    if (k < cut) {
        r[0] = c.u_dead;
        r[1] = c.v_dead;
        r[2] = c.w_dead;
        r[3] = c.T_dead;
        r[4] = (k < nzy) ? c.Y_dead : 0.0;
    } else if (k == cut) {
        const double *z = c.z;
        double h0 = z[k + 1] - z[k], h1 = z[k + 2] - z[k + 1];
        double t1[4], t2[4];
        wf_rhs_interior(c, X, D, i, j, k + 1, t1);
        wf_rhs_interior(c, X, D, i, j, k + 2, t2);
        r[0] = 0.0;
        r[1] = 0.0;
        r[2] = 0.0;
        r[3] = wf_extrap_bottom(t1[3], t2[3], h0, h1);
        if (k < nzy) {
            double Y1 = wf_rhs_fuel(c, X, col, q + P, k + 1, nzy);
            double Y2 = wf_rhs_fuel(c, X, col, q + 2 * P, k + 2, nzy);
            r[4] = wf_extrap_bottom(Y1, Y2, h0, h1);
        } else {
            r[4] = 0.0;
        }
    } else if (k < Nz - 1) {
        wf_rhs_interior(c, X, D, i, j, k, r);
        r[4] = wf_rhs_fuel(c, X, col, q, k, nzy);
    } else {
        const double *z = c.z;
        double hm1 = z[k] - z[k - 1], hm2 = z[k - 1] - z[k - 2];
        double t1[4], t2[4];
        wf_rhs_interior(c, X, D, i, j, k - 1, t1);
        wf_rhs_interior(c, X, D, i, j, k - 2, t2);
#pragma unroll
        for (int f = 0; f < 4; f++) r[f] = wf_extrap_top(t1[f], t2[f], hm2, hm1);
        r[4] = wf_rhs_fuel(c, X, col, q, k, nzy);
    }
    double *o[5] = {out.u, out.v, out.w, out.T, out.Y};
    const double *yn[5] = {Yn.u, Yn.v, Yn.w, Yn.T, Yn.Y};
    double *ac[5] = {acc.u, acc.v, acc.w, acc.T, acc.Y};
    int nf = (k < c.Nzm) ? 5 : 4;
#pragma unroll
    for (int f = 0; f < 5; f++) {
        if (f >= nf) break;
        double kf = r[f];
        double res;
        switch (mode) {
            case WF_MODE_PHI:
                res = kf;
                break;
            case WF_MODE_FIRST: /* acc = k1; next = y_n + a k1 */
                ac[f][q] = kf;
                res = yn[f][q] + a * kf;
                break;
            case WF_MODE_MID: /* acc += 2 k; next = y_n + a k */
                ac[f][q] = ac[f][q] + 2 * kf;
                res = yn[f][q] + a * kf;
                break;
            case WF_MODE_LAST: /* y_{n+1} = y_n + a (acc + k) */
                res = yn[f][q] + a * (ac[f][q] + kf);
                break;
            default: /* WF_MODE_EULER: y_{n+1} = y_n + a k */
                res = yn[f][q] + a * kf;
                break;
        }
        o[f][q] = res;
    }
}

/* --------------------------------------------------------------- layout ---- */

/* Reference layout (k fastest) -> internal layout (i fastest) for `nfields`
 * consecutive scalar fields of nz planes each (nz = Nz, or Nz_Y_max for Y). */
__global__ void k_to_internal(const double *src, double *dst, int Nx, int Ny, int nz) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y;
    int k = blockIdx.z;
    if (i >= Nx) return;
    dst[(long)i + (long)Nx * ((long)j + (long)Ny * k)] = src[(long)k + (long)nz * ((long)j + (long)Ny * i)];
}

__global__ void k_to_reference(const double *src, double *dst, int Nx, int Ny, int nz) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y;
    int i = blockIdx.z;
    if (k >= nz) return;
    dst[(long)k + (long)nz * ((long)j + (long)Ny * i)] = src[(long)i + (long)Nx * ((long)j + (long)Ny * k)];
}

/* ------------------------------------------------------------- launchers ---- */

hipError_t wf_launch_derived(const WfConst &c, const WfState &X, const WfDerived &D, hipStream_t s) {
    dim3 blk(64, 4, 1), grd((c.Nx + 63) / 64, (c.Ny + 3) / 4, c.Nz);
    hipLaunchKernelGGL(k_derived, grd, blk, 0, s, c, X, D);
    return hipGetLastError();
}

hipError_t wf_launch_stage(const WfConst &c, const WfState &X, const WfDerived &D, const WfState &Yn,
                           const WfState &acc, const WfState &out, int mode, double a, hipStream_t s) {
    dim3 blk(64, 4, 1), grd((c.Nx + 63) / 64, (c.Ny + 3) / 4, c.Nz);
    hipLaunchKernelGGL(k_stage, grd, blk, 0, s, c, X, D, Yn, acc, out, mode, a);
    return hipGetLastError();
}

hipError_t wf_launch_to_internal(const double *src, double *dst, int Nx, int Ny, int nz, hipStream_t s) {
    if (nz <= 0) return hipSuccess;
    dim3 blk(64, 1, 1), grd((Nx + 63) / 64, Ny, nz);
    hipLaunchKernelGGL(k_to_internal, grd, blk, 0, s, src, dst, Nx, Ny, nz);
    return hipGetLastError();
}

hipError_t wf_launch_to_reference(const double *src, double *dst, int Nx, int Ny, int nz, hipStream_t s) {
    if (nz <= 0) return hipSuccess;
    dim3 blk(64, 1, 1), grd((nz + 63) / 64, Ny, Nx);
    hipLaunchKernelGGL(k_to_reference, grd, blk, 0, s, src, dst, Nx, Ny, nz);
    return hipGetLastError();
}
