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
    int i, j, k;
    wf_tile_ijk(c, &i, &j, &k);
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

/* 32-bit element offsets from a uniform base (global loads with a scalar base
 * and a 32-bit lane offset; the solver limits Nx*Ny*Nz to 2^28). */
__device__ __forceinline__ double wf_ld(const double *base, unsigned q) {
    return *(const double *)((const char *)base + (q << 3));
}

/* LDS halo tiles of one plane of a WF_TILE_X x WF_TILE_Y node tile. */
#define WF_SH 2 /* state halo (upwind stencils reach i +- 2) */
#define WF_DH 1 /* halo of the derived fields, u_tau and topography */
#define WF_SW (WF_TILE_X + 2 * WF_SH)
#define WF_SHH (WF_TILE_Y + 2 * WF_SH)
#define WF_DW (WF_TILE_X + 2 * WF_DH)
#define WF_DHH (WF_TILE_Y + 2 * WF_DH)
#define WF_NT_STAGE (WF_TILE_X * WF_TILE_Y)

struct WfStageTiles {
    double st[4][WF_SHH * WF_SW]; /* u, v, w, T */
    double dv[7][WF_DHH * WF_DW]; /* mu, s11, s12, s13, s22, s23, divU */
    double ut[WF_DHH * WF_DW];    /* u_tau (k = 0 plane) */
    double tp[WF_DHH * WF_DW];    /* topography */
};

/* Periodic neighbour of tile position g (= i0 - halo + offset) with the
 * reference's wrap (i +- d) mod (n-1): every neighbour a node 0 <= i <= n-1
 * reads (|d| <= 2) lands on its image in [0, n-2].  Positions beyond the last
 * node's neighbours are never read and map to 0. */
__device__ __forceinline__ int wf_tile_wrap(int g, int m) {
    g = (g < 0) ? g + m : g;
    g = (g >= m) ? g - m : g;
    return (g >= m) ? 0 : g;
}

/* Loads of one halo tile (width W, height HH, halo H) of plane k of `src`
 * into registers; all lanes load (out-of-tile entries read node 0) so the
 * loads issue without branches. */
template <int W, int HH, int H>
__device__ __forceinline__ void wf_tile_load(double *v, const double *src, const WfConst &c, int i0, int j0, int k,
                                             int tid) {
    constexpr int NE = W * HH;
    constexpr int PER = (NE + WF_NT_STAGE - 1) / WF_NT_STAGE;
#pragma unroll
    for (int p = 0; p < PER; p++) {
        const int e = tid + p * WF_NT_STAGE;
        const int r = e / W, cx = e - r * W;
        const int gx = wf_tile_wrap(i0 - H + cx, c.Nx - 1);
        const int gy = wf_tile_wrap(j0 - H + r, c.Ny - 1);
        const unsigned q = (e < NE) ? (unsigned)(gx + c.Nx * (gy + c.Ny * k)) : 0u;
        v[p] = wf_ld(src, q);
    }
}

template <int W, int HH>
__device__ __forceinline__ void wf_tile_store(double *dst, const double *v, int tid) {
    constexpr int NE = W * HH;
    constexpr int PER = (NE + WF_NT_STAGE - 1) / WF_NT_STAGE;
#pragma unroll
    for (int p = 0; p < PER; p++) {
        const int e = tid + p * WF_NT_STAGE;
        if (e < NE) dst[e] = v[p];
    }
}

/* Raw RHS of u, v, w, T, Y (RHS + SGS, before any boundary condition) at an
 * interior node 1 <= k <= Nz-2 of the tile: neighbours in x,y from the LDS
 * tiles, the node itself and its z-neighbours from registers.  Every
 * expression is the reference's (src/cu/pde.cu:139-787,
 * src/cu/turbulence.cu:255-454), evaluated in its order. */
struct WfColumnVals {
    double s[4][5];   /* u, v, w, T at k-2, k-1, k, k+1, k+2 (index 2 = k) */
    double mu[3], s13[3], s23[3], s33[3], dU[3]; /* k-1, k, k+1 */
    double s11, s12, s22;
    double Y, utau, topo;
};

__device__ __forceinline__ void wf_rhs_tile(const WfConst &c, const WfStageTiles &L, const WfColumnVals &V, int tx,
                                            int ty, int col, int k, int nzy, double r[5]) {
    const int Nz = c.Nz;
    const double *z = c.z;
    const double hm1 = z[k] - z[k - 1], h0 = z[k + 1] - z[k];
    const double g1 = (k <= Nz - 3) ? z[k + 2] - z[k + 1] : 0.0;
    const double g2 = (k >= 2) ? z[k - 1] - z[k - 2] : 0.0;
    const double u = V.s[0][2], v = V.s[1][2], w = V.s[2][2], T = V.s[3][2];
    const double u_plu = (u > 0.0) ? u : 0.0, u_min = (u < 0.0) ? u : 0.0;
    const double v_plu = (v > 0.0) ? v : 0.0, v_min = (v < 0.0) ? v : 0.0;
    const double w_plu = (w > 0.0) ? w : 0.0, w_min = (w < 0.0) ? w : 0.0;
    const int sc = (ty + WF_SH) * WF_SW + (tx + WF_SH);
    double adv[4], lap[4];
    double Tx = 0.0, Ty = 0.0, Tz = 0.0, Txx = 0.0, Tyy = 0.0, Tzz = 0.0, T_kp1 = 0.0;
#pragma unroll
    for (int f = 0; f < 4; f++) {
        const double *A = L.st[f];
        const double a0 = V.s[f][2];
        const double am1 = A[sc - 1], am2 = A[sc - 2], ap1 = A[sc + 1], ap2 = A[sc + 2];
        const double bm1 = A[sc - WF_SW], bm2 = A[sc - 2 * WF_SW], bp1 = A[sc + WF_SW], bp2 = A[sc + 2 * WF_SW];
        const double cm2 = V.s[f][0], cm1 = V.s[f][1], cp1 = V.s[f][3], cp2 = V.s[f][4];
        const double d_im = (3 * a0 - 4 * am1 + am2) / (2 * c.dx);
        const double d_ip = (-3 * a0 + 4 * ap1 - ap2) / (2 * c.dx);
        const double d_jm = (3 * a0 - 4 * bm1 + bm2) / (2 * c.dy);
        const double d_jp = (-3 * a0 + 4 * bp1 - bp2) / (2 * c.dy);
        const double d_km = (k <= 1) ? wf_d1f(a0, cp1, cp2, h0, g1) : wf_d1b(cm2, cm1, a0, g2, hm1);
        const double d_kp = (k >= Nz - 2) ? wf_d1b(cm2, cm1, a0, g2, hm1) : wf_d1f(a0, cp1, cp2, h0, g1);
        const double dxx = (ap1 - 2.0 * a0 + am1) / (c.dx * c.dx);
        const double dyy = (bp1 - 2.0 * a0 + bm1) / (c.dy * c.dy);
        const double dzz = wf_d2m(cm1, a0, cp1, hm1, h0);
        const double ax = u_plu * d_im + u_min * d_ip;
        const double ay = v_plu * d_jm + v_min * d_jp;
        const double az = w_plu * d_km + w_min * d_kp;
        adv[f] = ax + ay + az;
        lap[f] = dxx + dyy + dzz;
        if (f == 3) {
            Tx = (ap1 - am1) / (2.0 * c.dx);
            Ty = (bp1 - bm1) / (2.0 * c.dy);
            Tz = (cp1 - cm1) / (hm1 + h0);
            Txx = dxx;
            Tyy = dyy;
            Tzz = dzz;
            T_kp1 = cp1;
        }
    }
    const double rho = c.T_inf * c.rho_inf / T;
    /* heat source and fuel zone (src/cu/pde.cu:663-699) */
    double K_T = wf_rate(c, T);
    double T_gas = c.T_inf, T_solid = c.T_inf, Y = 0.0;
    if (k <= nzy) {
        if (k < nzy)
            Y = V.Y;
        else
            T_gas = T_kp1;
        if (Y > 0.0)
            T_solid = T;
        else
            T_gas = T;
    }
    double q_src = c.H_C * Y * K_T / c.c_p;
    if (c.exchange_on) {
        double dz_k = h0;
        double h_c = c.h_c * dz_k * pow(fabs(T_gas - T_solid) / dz_k, 1.0 / 4.0);
        q_src = q_src - h_c * c.alpha_s * c.sigma_s * (T_gas - T_solid) / (c.c_p * c.T_inf * c.rho_inf / T_gas);
    }
    double mu = c.mu, kT = c.kT, dk = 0.0;
    if (c.suth_mu == 1) mu = 1.716e-5 * pow(T / 273.15, 1.5) * (273.15 + 110.4) / (T + 110.4);
    if (c.suth_k == 1) wf_sutherland_k(T, &kT, &dk);
    double nu = mu / rho;
    double T2 = T * T, T3 = T * T * T;
    double T_RHS_tmp = (dk + 12 * WF_SIGMA_D * c.delta * T2 * (Tx * Tx + Ty * Ty + Tz * Tz) +
                        (kT + 4 * WF_SIGMA_D * c.delta * T3) * lap[3]) / (rho * c.c_p) + q_src;
    double T_RHS;
    if (c.temperature_convection == 0)
        T_RHS = T_RHS_tmp - (u * Tx + v * Ty + w * Tz);
    else
        T_RHS = T_RHS_tmp - adv[3];
    double mod_U = sqrt(u * u + v * v + w * w);
    double F_x = -0.5 * c.C_d * c.alpha_s * c.sigma_s * Y * mod_U * u;
    double F_y = -0.5 * c.C_d * c.alpha_s * c.sigma_s * Y * mod_U * v;
    double F_z = c.g * (rho - c.rho_inf) / rho - 0.5 * c.C_d * c.alpha_s * c.sigma_s * Y * mod_U * w;
    double u_RHS = nu * lap[0] - adv[0] + F_x;
    double v_RHS = nu * lap[1] - adv[1] + F_y;
    double w_RHS = nu * lap[2] - adv[2] + F_z;

    /* SGS divergence (src/cu/turbulence.cu:291-452) */
    const int dc = (ty + WF_DH) * WF_DW + (tx + WF_DH);
    double nu_t = c.mu / rho;
    double ut = V.utau;
    double fw = wf_fw(wf_zibm_t(c, V.topo, k), ut, nu_t);
    double fw_ip = wf_fw(wf_zibm_t(c, L.tp[dc + 1], k), L.ut[dc + 1], nu_t);
    double fw_im = wf_fw(wf_zibm_t(c, L.tp[dc - 1], k), L.ut[dc - 1], nu_t);
    double fw_jp = wf_fw(wf_zibm_t(c, L.tp[dc + WF_DW], k), L.ut[dc + WF_DW], nu_t);
    double fw_jm = wf_fw(wf_zibm_t(c, L.tp[dc - WF_DW], k), L.ut[dc - WF_DW], nu_t);
    double fw_km = wf_fw(wf_zibm_t(c, V.topo, k - 1), ut, nu_t);
    double fw_kp = wf_fw(wf_zibm_t(c, V.topo, k + 1), ut, nu_t);
    const double *Mu = L.dv[0];
    double mu_c = V.mu[1];
    double mu_ip = Mu[dc + 1], mu_im = Mu[dc - 1], mu_jp = Mu[dc + WF_DW], mu_jm = Mu[dc - WF_DW];
    double mu_km = V.mu[0], mu_kp = V.mu[2];
    mu_c *= fw * fw;
    mu_im *= fw_im * fw_im;
    mu_ip *= fw_ip * fw_ip;
    mu_jm *= fw_jm * fw_jm;
    mu_jp *= fw_jp * fw_jp;
    mu_km *= fw_km * fw_km;
    mu_kp *= fw_kp * fw_kp;
    double S11 = V.s11, S12 = V.s12, S13 = V.s13[1], S22 = V.s22, S23 = V.s23[1], S33 = V.s33[1];
    const double *s11 = L.dv[1], *s12 = L.dv[2], *s13 = L.dv[3], *s22 = L.dv[4], *s23 = L.dv[5], *dU = L.dv[6];
    double mu_x = (mu_ip - mu_im) / (2 * c.dx);
    double mu_y = (mu_jp - mu_jm) / (2 * c.dy);
    double S11_x = (s11[dc + 1] - s11[dc - 1]) / (2.0 * c.dx);
    double S12_x = (s12[dc + 1] - s12[dc - 1]) / (2.0 * c.dx);
    double S13_x = (s13[dc + 1] - s13[dc - 1]) / (2.0 * c.dx);
    double S21_y = (s12[dc + WF_DW] - s12[dc - WF_DW]) / (2.0 * c.dy);
    double S22_y = (s22[dc + WF_DW] - s22[dc - WF_DW]) / (2.0 * c.dy);
    double S23_y = (s23[dc + WF_DW] - s23[dc - WF_DW]) / (2.0 * c.dy);
    double dU_x = (dU[dc + 1] - dU[dc - 1]) / (2.0 * c.dx);
    double dU_y = (dU[dc + WF_DW] - dU[dc - WF_DW]) / (2.0 * c.dy);
    double mu_z = (mu_kp - mu_km) / (h0 + hm1);
    double S31_z = (V.s13[2] - V.s13[0]) / (h0 + hm1);
    double S32_z = (V.s23[2] - V.s23[0]) / (h0 + hm1);
    double S33_z = (V.s33[2] - V.s33[0]) / (h0 + hm1);
    double dU_z = (V.dU[2] - V.dU[0]) / (h0 + hm1);
    double div_S_x = S11_x + S21_y + S31_z;
    double div_S_y = S12_x + S22_y + S32_z;
    double div_S_z = S13_x + S23_y + S33_z;
    double sgs_x = -2 * (mu_x * S11 + mu_y * S12 + mu_z * S13 + mu_c * div_S_x) / rho;
    double sgs_y = -2 * (mu_x * S12 + mu_y * S22 + mu_z * S23 + mu_c * div_S_y) / rho;
    double sgs_z = -2 * (mu_x * S13 + mu_y * S23 + mu_z * S33 + mu_c * div_S_z) / rho;
    double sgs_q = -(mu_x * Tx + mu_y * Ty + mu_z * Tz + mu_c * (Txx + Tyy + Tzz)) / (rho * c.Pr);
    r[0] = u_RHS + (-sgs_x + dU_x / 3.0);
    r[1] = v_RHS + (-sgs_y + dU_y / 3.0);
    r[2] = w_RHS + (-sgs_z + dU_z / 3.0);
    r[3] = T_RHS - sgs_q;
    /* fuel (src/cu/pde.cu:672-675): -Y_f K(T) Y where the column stores fuel */
    r[4] = (k < nzy) ? -c.Y_f * K_T * V.Y : 0.0;
}

/* Raw RHS of u, v, w, T (RHS + SGS, before any boundary condition) at an
 * interior plane 1 <= k <= Nz-2, every neighbour read from global memory
 * (used on the duplicate planes i = Nx-1, j = Ny-1).  The stencil terms of each field are folded
 * into its advection sum and Laplacian as soon as they are formed, so few
 * intermediate values are live at once; every expression is the reference's. */
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
    const double hm1 = z[k] - z[k - 1], h0 = z[k + 1] - z[k];
    const double g1 = (k <= Nz - 3) ? z[k + 2] - z[k + 1] : 0.0;
    const double g2 = (k >= 2) ? z[k - 1] - z[k - 2] : 0.0;
    const double u = X.u[q], v = X.v[q], w = X.w[q], T = X.T[q];
    const double u_plu = (u > 0.0) ? u : 0.0, u_min = (u < 0.0) ? u : 0.0;
    const double v_plu = (v > 0.0) ? v : 0.0, v_min = (v < 0.0) ? v : 0.0;
    const double w_plu = (w > 0.0) ? w : 0.0, w_min = (w < 0.0) ? w : 0.0;
    double adv[4], lap[4];
    double Tx = 0.0, Ty = 0.0, Tz = 0.0, Txx = 0.0, Tyy = 0.0, Tzz = 0.0, T_kp1 = 0.0;
#pragma unroll
    for (int f = 0; f < 4; f++) {
        const double *A = (f == 0) ? X.u : (f == 1) ? X.v : (f == 2) ? X.w : X.T;
        const double a0 = (f == 0) ? u : (f == 1) ? v : (f == 2) ? w : T;
        const double am1 = A[qxm1], am2 = A[qxm2], ap1 = A[qxp1], ap2 = A[qxp2];
        const double bm1 = A[qym1], bm2 = A[qym2], bp1 = A[qyp1], bp2 = A[qyp2];
        const double cm1 = A[q - P], cp1 = A[q + P];
        const double d_im = (3 * a0 - 4 * am1 + am2) / (2 * c.dx);
        const double d_ip = (-3 * a0 + 4 * ap1 - ap2) / (2 * c.dx);
        const double d_jm = (3 * a0 - 4 * bm1 + bm2) / (2 * c.dy);
        const double d_jp = (-3 * a0 + 4 * bp1 - bp2) / (2 * c.dy);
        double d_km, d_kp;
        if (k <= 1)
            d_km = wf_d1f(a0, cp1, A[q + 2 * P], h0, g1);
        else
            d_km = wf_d1b(A[q - 2 * P], cm1, a0, g2, hm1);
        if (k >= Nz - 2)
            d_kp = wf_d1b(A[q - 2 * P], cm1, a0, g2, hm1);
        else
            d_kp = wf_d1f(a0, cp1, A[q + 2 * P], h0, g1);
        const double dxx = (ap1 - 2.0 * a0 + am1) / (c.dx * c.dx);
        const double dyy = (bp1 - 2.0 * a0 + bm1) / (c.dy * c.dy);
        const double dzz = wf_d2m(cm1, a0, cp1, hm1, h0);
        const double ax = u_plu * d_im + u_min * d_ip;
        const double ay = v_plu * d_jm + v_min * d_jp;
        const double az = w_plu * d_km + w_min * d_kp;
        adv[f] = ax + ay + az;
        lap[f] = dxx + dyy + dzz;
        if (f == 3) {
            Tx = (ap1 - am1) / (2.0 * c.dx);
            Ty = (bp1 - bm1) / (2.0 * c.dy);
            Tz = (cp1 - cm1) / (hm1 + h0);
            Txx = dxx;
            Tyy = dyy;
            Tzz = dzz;
            T_kp1 = cp1;
        }
    }
    const double rho = c.T_inf * c.rho_inf / T;
    const double lap_u = lap[0], lap_v = lap[1], lap_w = lap[2], lap_T = lap[3];
    int col = i + Nx * j;
    double mu = c.mu, kT = c.kT, dk = 0.0;
    double Y = 0.0;
    double q_src = wf_heat_source_y(c, col, i, j, k, T, T_kp1, X.Y, &Y);
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
        T_RHS = T_RHS_tmp - adv[3];
    double mod_U = sqrt(u * u + v * v + w * w);
    double F_x = -0.5 * c.C_d * c.alpha_s * c.sigma_s * Y * mod_U * u;
    double F_y = -0.5 * c.C_d * c.alpha_s * c.sigma_s * Y * mod_U * v;
    double F_z = c.g * (rho - c.rho_inf) / rho - 0.5 * c.C_d * c.alpha_s * c.sigma_s * Y * mod_U * w;
    double u_RHS = nu * lap_u - adv[0] + F_x;
    double v_RHS = nu * lap_v - adv[1] + F_y;
    double w_RHS = nu * lap_w - adv[2] + F_z;

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

/* RK combination of a node (src/cu/utils.cu:14-21, src/cu/solver.cu:14-39).
 * mode WF_MODE_PHI writes the stage derivative r; otherwise `out` receives
 * the next stage input (or y_{n+1}) and `acc` the running combination. */
__device__ __forceinline__ void wf_rk_write(int mode, double a, unsigned q, const double r[5], int nf,
                                            const WfState &Yn, const WfState &acc, const WfState &out) {
    double *o[5] = {out.u, out.v, out.w, out.T, out.Y};
    const double *yn[5] = {Yn.u, Yn.v, Yn.w, Yn.T, Yn.Y};
    double *ac[5] = {acc.u, acc.v, acc.w, acc.T, acc.Y};
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

/* Raw RHS of the nodes cut+1, cut+2, Nz-2 and Nz-3, kept per column for the
 * cut and top nodes (k_stage_columns). */
__device__ __forceinline__ void wf_keep_raw(const WfConst &c, const WfDerived &D, int col, int k, int cut,
                                            const double r[5]) {
    const unsigned pl = (unsigned)c.plane;
    double *keep = (k == cut + 1) ? D.col : (k == cut + 2) ? D.col + 5 * pl : nullptr;
    if (keep) {
#pragma unroll
        for (int f = 0; f < 5; f++) keep[f * pl + col] = r[f];
    }
    keep = (k == c.Nz - 2) ? D.col + 10 * pl : (k == c.Nz - 3) ? D.col + 15 * pl : nullptr;
    if (keep) {
#pragma unroll
        for (int f = 0; f < 5; f++) keep[f * pl + col] = r[f];
    }
}

/* One stage on the nodes below the top plane: dead nodes (k < cut) take the
 * dead-node values, fluid nodes cut < k <= Nz-2 their RHS; the raw RHS of the
 * nodes cut+1, cut+2, Nz-2 and Nz-3 is also kept per column for the cut and
 * top nodes, which k_stage_columns completes (src/cu/pde.cu:14-99). */
__global__ __launch_bounds__(WF_NT_STAGE) void k_stage(WfConst c, WfState X, WfDerived D, WfState Yn, WfState acc,
                                                       WfState out, int mode, double a) {
    __shared__ WfStageTiles L;
    const int t = wf_xcd_tile(blockIdx.x, gridDim.x);
    const int nty = (c.Ny + WF_TILE_Y - 1) / WF_TILE_Y;
    const int k = t % (c.Nz - 1); /* the top plane is completed per column */
    const int colt = t / (c.Nz - 1);
    const int j0 = (colt % nty) * WF_TILE_Y, i0 = (colt / nty) * WF_TILE_X;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + WF_TILE_X * ty;
    const int i = i0 + tx, j = j0 + ty;
    const bool inside = (i < c.Nx) && (j < c.Ny);
    const unsigned P = (unsigned)c.plane;
    const int col = inside ? i + c.Nx * j : 0;
    const unsigned q = (unsigned)col + P * (unsigned)k;
    /* halo tiles of plane k */
    {
        constexpr int PS = (WF_SW * WF_SHH + WF_NT_STAGE - 1) / WF_NT_STAGE;
        constexpr int PD = (WF_DW * WF_DHH + WF_NT_STAGE - 1) / WF_NT_STAGE;
        double vs[4][PS], vd[7][PD], vu[PD], vt[PD];
        wf_tile_load<WF_SW, WF_SHH, WF_SH>(vs[0], X.u, c, i0, j0, k, tid);
        wf_tile_load<WF_SW, WF_SHH, WF_SH>(vs[1], X.v, c, i0, j0, k, tid);
        wf_tile_load<WF_SW, WF_SHH, WF_SH>(vs[2], X.w, c, i0, j0, k, tid);
        wf_tile_load<WF_SW, WF_SHH, WF_SH>(vs[3], X.T, c, i0, j0, k, tid);
        const double *dsrc[7] = {D.mu, D.s11, D.s12, D.s13, D.s22, D.s23, D.divU};
#pragma unroll
        for (int f = 0; f < 7; f++) wf_tile_load<WF_DW, WF_DHH, WF_DH>(vd[f], dsrc[f], c, i0, j0, k, tid);
        wf_tile_load<WF_DW, WF_DHH, WF_DH>(vu, D.utau, c, i0, j0, 0, tid);
        wf_tile_load<WF_DW, WF_DHH, WF_DH>(vt, c.topo, c, i0, j0, 0, tid);
#pragma unroll
        for (int f = 0; f < 4; f++) wf_tile_store<WF_SW, WF_SHH>(L.st[f], vs[f], tid);
#pragma unroll
        for (int f = 0; f < 7; f++) wf_tile_store<WF_DW, WF_DHH>(L.dv[f], vd[f], tid);
        wf_tile_store<WF_DW, WF_DHH>(L.ut, vu, tid);
        wf_tile_store<WF_DW, WF_DHH>(L.tp, vt, tid);
    }
    /* the node's own column (its own values, also on the duplicate planes) */
    WfColumnVals V;
    {
        const unsigned km2 = (k >= 2) ? q - 2 * P : q, kp2 = (k <= c.Nz - 3) ? q + 2 * P : q;
        const unsigned km1 = (k >= 1) ? q - P : q, kp1 = q + P;
        const double *sf[4] = {X.u, X.v, X.w, X.T};
#pragma unroll
        for (int f = 0; f < 4; f++) {
            V.s[f][0] = wf_ld(sf[f], km2);
            V.s[f][1] = wf_ld(sf[f], km1);
            V.s[f][2] = wf_ld(sf[f], q);
            V.s[f][3] = wf_ld(sf[f], kp1);
            V.s[f][4] = wf_ld(sf[f], kp2);
        }
        V.mu[0] = wf_ld(D.mu, km1);
        V.mu[1] = wf_ld(D.mu, q);
        V.mu[2] = wf_ld(D.mu, kp1);
        V.s13[0] = wf_ld(D.s13, km1);
        V.s13[1] = wf_ld(D.s13, q);
        V.s13[2] = wf_ld(D.s13, kp1);
        V.s23[0] = wf_ld(D.s23, km1);
        V.s23[1] = wf_ld(D.s23, q);
        V.s23[2] = wf_ld(D.s23, kp1);
        V.s33[0] = wf_ld(D.s33, km1);
        V.s33[1] = wf_ld(D.s33, q);
        V.s33[2] = wf_ld(D.s33, kp1);
        V.dU[0] = wf_ld(D.divU, km1);
        V.dU[1] = wf_ld(D.divU, q);
        V.dU[2] = wf_ld(D.divU, kp1);
        V.s11 = wf_ld(D.s11, q);
        V.s12 = wf_ld(D.s12, q);
        V.s22 = wf_ld(D.s22, q);
        /* Y stores Nz_Y_max planes (possibly none): other planes read a T node */
        V.Y = wf_ld((c.Nzm > 0) ? X.Y : X.T, (k < c.Nzm) ? q : (unsigned)col);
        V.utau = wf_ld(D.utau, (unsigned)col);
        V.topo = wf_ld(c.topo, (unsigned)col);
    }
    const int cut = c.cut[col], nzy = c.nzy[col];
    __syncthreads();
    /* the duplicate planes i = Nx-1, j = Ny-1 read their x- (y-) neighbours
     * in their own row (column), which the tiles hold as row (column) 0:
     * k_stage_dup evaluates them */
    if (!inside || k == cut || i == c.Nx - 1 || j == c.Ny - 1) return;
    double r[5];
    if (k < cut) {
        r[0] = c.u_dead;
        r[1] = c.v_dead;
        r[2] = c.w_dead;
        r[3] = c.T_dead;
        r[4] = (k < nzy) ? c.Y_dead : 0.0;
    } else {
        wf_rhs_tile(c, L, V, tx, ty, col, k, nzy, r);
        wf_keep_raw(c, D, col, k, cut, r);
    }
    wf_rk_write(mode, a, q, r, (k < c.Nzm) ? 5 : 4, Yn, acc, out);
}

/* The nodes of k_stage on the duplicate planes i = Nx-1 and j = Ny-1, one
 * thread per node, neighbours from global memory. */
__global__ __launch_bounds__(256) void k_stage_dup(WfConst c, WfState X, WfDerived D, WfState Yn, WfState acc,
                                                   WfState out, int mode, double a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (t >= c.Nx + c.Ny - 1) return;
    const int i = (t < c.Ny) ? c.Nx - 1 : t - c.Ny;
    const int j = (t < c.Ny) ? t : c.Ny - 1;
    const int col = i + c.Nx * j;
    const int cut = c.cut[col], nzy = c.nzy[col];
    if (k == cut) return;
    const long q = wf_q(c, i, j, k);
    double r[5];
    if (k < cut) {
        r[0] = c.u_dead;
        r[1] = c.v_dead;
        r[2] = c.w_dead;
        r[3] = c.T_dead;
        r[4] = (k < nzy) ? c.Y_dead : 0.0;
    } else {
        wf_rhs_interior(c, X, D, i, j, k, r);
        r[4] = (k < nzy) ? -c.Y_f * wf_rate(c, X.T[q]) * X.Y[q] : 0.0;
        wf_keep_raw(c, D, col, k, cut, r);
    }
    wf_rk_write(mode, a, (unsigned)q, r, (k < c.Nzm) ? 5 : 4, Yn, acc, out);
}

/* Extrapolation giving a zero one-sided z-derivative at the bottom of the fluid
 * (src/cu/pde.cu:62) and at the top (src/cu/pde.cu:84-87). */
__device__ __forceinline__ double wf_extrap_bottom(double f1, double f2, double h0, double h1) {
    return ((h0 + h1) * (h0 + h1) * f1 - h0 * h0 * f2) / (h1 * (2 * h0 + h1));
}
__device__ __forceinline__ double wf_extrap_top(double fm1, double fm2, double hm2, double hm1) {
    return ((hm2 + hm1) * (hm2 + hm1) * fm1 - hm1 * hm1 * fm2) / (hm2 * (hm2 + 2 * hm1));
}

/* Cut and top node of every column from the raw RHS the stage pass kept
 * (src/cu/pde.cu:45-87): u = v = w = 0 and zero-gradient T, Y at the cut node;
 * zero-gradient u, v, w, T at the top; Y RHS of the top node itself. */
__global__ void k_stage_columns(WfConst c, WfState X, WfDerived D, WfState Yn, WfState acc, WfState out, int mode,
                                double a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= c.Nx) return;
    const unsigned pl = (unsigned)c.plane;
    const int col = i + c.Nx * j;
    const int cut = c.cut[col], nzy = c.nzy[col];
    const double *z = c.z;
    const double *b1 = D.col, *b2 = D.col + 5 * pl, *t1 = D.col + 10 * pl, *t2 = D.col + 15 * pl;
    double r[5];
    {
        const int k = cut;
        const double h0 = z[k + 1] - z[k], h1 = z[k + 2] - z[k + 1];
        r[0] = 0.0;
        r[1] = 0.0;
        r[2] = 0.0;
        r[3] = wf_extrap_bottom(b1[3 * pl + col], b2[3 * pl + col], h0, h1);
        r[4] = (k < nzy) ? wf_extrap_bottom(b1[4 * pl + col], b2[4 * pl + col], h0, h1) : 0.0;
        wf_rk_write(mode, a, (unsigned)col + pl * (unsigned)k, r, (k < c.Nzm) ? 5 : 4, Yn, acc, out);
    }
    {
        const int k = c.Nz - 1;
        const unsigned q = (unsigned)col + pl * (unsigned)k;
        const double hm1 = z[k] - z[k - 1], hm2 = z[k - 1] - z[k - 2];
#pragma unroll
        for (int f = 0; f < 4; f++) r[f] = wf_extrap_top(t1[f * pl + col], t2[f * pl + col], hm2, hm1);
        r[4] = (k < nzy) ? -c.Y_f * wf_rate(c, X.T[q]) * X.Y[q] : 0.0;
        wf_rk_write(mode, a, q, r, (k < c.Nzm) ? 5 : 4, Yn, acc, out);
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
    dim3 blk(WF_TILE_X, WF_TILE_Y, 1), grd(wf_tile_count(c));
    hipLaunchKernelGGL(k_derived, grd, blk, 0, s, c, X, D);
    return hipGetLastError();
}

hipError_t wf_launch_stage(const WfConst &c, const WfState &X, const WfDerived &D, const WfState &Yn,
                           const WfState &acc, const WfState &out, int mode, double a, hipStream_t s) {
    const int tiles = ((c.Nx + WF_TILE_X - 1) / WF_TILE_X) * ((c.Ny + WF_TILE_Y - 1) / WF_TILE_Y) * (c.Nz - 1);
    hipLaunchKernelGGL(k_stage, dim3(tiles), dim3(WF_TILE_X, WF_TILE_Y, 1), 0, s, c, X, D, Yn, acc, out, mode, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_stage_dup, dim3((c.Nx + c.Ny - 1 + 255) / 256, c.Nz - 1), dim3(256), 0, s, c, X, D, Yn, acc,
                       out, mode, a);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_stage_columns, dim3((c.Nx + 63) / 64, c.Ny), dim3(64), 0, s, c, X, D, Yn, acc, out, mode, a);
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
