/*
 * init.c -- initial conditions, the raw-binary initial-condition loader and the
 * state-vector packer (host C).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "wf_host.h"

#define WF_FILENAME 192

/* Analytic initial state (src/cu/functions.cu:187-323): power-law wind along x,
 * a Gaussian or box temperature anomaly, p = p_top on the top plane and 0
 * elsewhere, fuel Y = 1 in the fuel box (with ramps of widths Y0_xa, Y0_ya
 * around it) below topography + Y_h, and the dead-node values under the terrain.
 * Y entries the reference leaves unassigned (its buffer comes from malloc) are
 * set to 0 here. */
void initial_conditions(double *u, double *v, double *w, double *T, double *Y, double *p, Parameters P) {
    int Nx = P.Nx, Ny = P.Ny, Nz = P.Nz, Nzm = P.Nz_Y_max;
    const double *x = P.x, *y = P.y, *z = P.z;
    double xa = P.Y0_xa, ya = P.Y0_ya, Y_h = P.Y_h;
    int gaussian_shape = strcmp(P.T0_shape, "gaussian") == 0;
    int box_shape = strcmp(P.T0_shape, "parallelepiped") == 0;
    memset(Y, 0, (size_t)Nx * Ny * Nzm * sizeof(double));
    for (int i = 0; i < Nx; i++) {
        for (int j = 0; j < Ny; j++) {
            long col = WF_IDX(i, j, 0, Nx, Ny, 1);
            int nzy = P.Nz_Y[col];
            int cut = P.cut_nodes[col];
            double topo = P.topography[col];
            for (int k = 0; k < Nz; k++) {
                long q = WF_IDX(i, j, k, Nx, Ny, Nz);
                u[q] = power_law(z[k], P.u_r, P.z_r, P.alpha_u);
                v[q] = 0.0;
                w[q] = 0.0;
                if (gaussian_shape)
                    T[q] = P.T_inf + (P.T_0 - P.T_inf) * gaussian(x[i], y[j], z[k], P.T0_x_center, P.T0_y_center,
                                                                 P.T0_z_center, P.T0_length, P.T0_width, P.T0_height);
                else if (box_shape)
                    T[q] = P.T_inf + (P.T_0 - P.T_inf) * parallelepiped(x[i], y[j], z[k], P.T0_x_start, P.T0_x_end,
                                                                       P.T0_y_start, P.T0_y_end, P.T0_z_start,
                                                                       P.T0_z_end);
                p[q] = (k == Nz - 1) ? P.p_top : 0.0;
                if (k < nzy) {
                    double z_bottom = Y_h * (y[j] - P.Y0_y_start + ya) / ya;
                    double z_top = Y_h * (P.Y0_y_end + ya - y[j]) / ya;
                    double z_left = Y_h * (x[i] - P.Y0_x_start + xa) / xa;
                    double z_right = Y_h * (P.Y0_x_end + xa - x[i]) / xa;
                    long qy = WF_IDX(i, j, k, Nx, Ny, Nzm);
                    int in_x = P.Y0_x_start <= x[i] && x[i] <= P.Y0_x_end;
                    int in_y = P.Y0_y_start <= y[j] && y[j] <= P.Y0_y_end;
                    int ramp_x = P.Y0_x_start - xa <= x[i] && x[i] <= P.Y0_x_end + xa;
                    int ramp_y = P.Y0_y_start - ya <= y[j] && y[j] <= P.Y0_y_end + ya;
                    if (in_x && in_y && (0 <= z[k] && z[k] <= (topo + Y_h)))
                        Y[qy] = 1.0;
                    else if ((P.Y0_y_start - ya <= y[j] && y[j] < P.Y0_y_start) && ramp_x && (0 <= z[k] && z[k] <= z_bottom))
                        Y[qy] = 1.0;
                    else if ((P.Y0_y_end < y[j] && y[j] <= P.Y0_y_end + ya) && ramp_x && (0 <= z[k] && z[k] <= z_top))
                        Y[qy] = 1.0;
                    else if ((P.Y0_x_start - xa <= x[i] && x[i] < P.Y0_x_start) && ramp_y && (0 <= z[k] && z[k] <= z_left))
                        Y[qy] = 1.0;
                    else if ((P.Y0_x_end < x[i] && x[i] <= P.Y0_x_end + xa) && ramp_y && (0 <= z[k] && z[k] <= z_right))
                        Y[qy] = 1.0;
                }
                if (k < cut) {
                    u[q] = P.u_dead_nodes;
                    v[q] = P.v_dead_nodes;
                    w[q] = P.w_dead_nodes;
                    T[q] = P.T_dead_nodes;
                    if (k < nzy) Y[WF_IDX(i, j, k, Nx, Ny, Nzm)] = P.Y_dead_nodes;
                }
            }
        }
    }
}

/* Initial state from <input_path>{u,v,w,T,Y,p}.bin, raw float64 in the output
 * layout (src/cu/input.cu:3-58).  Exits with the reference's messages when a
 * file is missing or short. */
void initial_conditions_input(double *u, double *v, double *w, double *T, double *Y, double *p, Parameters P) {
    const char *names[6] = {"u", "v", "w", "T", "Y", "p"};
    double *dst[6] = {u, v, w, T, Y, p};
    size_t want[6];
    size_t n3 = (size_t)P.Nx * P.Ny * P.Nz;
    FILE *files[6];
    int missing = 0;
    for (int q = 0; q < 6; q++) {
        char fn[WF_FILENAME];
        snprintf(fn, sizeof(fn), "%s%s.bin", P.input_path, names[q]);
        files[q] = fopen(fn, "rb");
        if (files[q] == NULL) missing = 1;
        want[q] = (q == 4) ? (size_t)P.Nx * P.Ny * P.Nz_Y_max : n3;
    }
    if (missing) {
        log_message(P, "Error opening initial conditions files.\n");
        exit(EXIT_FAILURE);
    }
    int short_read = 0;
    for (int q = 0; q < 6; q++) {
        size_t got = fread(dst[q], sizeof(double), want[q], files[q]);
        if (got != want[q]) short_read = 1;
    }
    if (short_read) {
        log_message(P, "Error reading initial conditions: file size mismatch.\n");
        exit(EXIT_FAILURE);
    }
    for (int q = 0; q < 6; q++) fclose(files[q]);
}

/* y_0 = [u | v | w | T | Y] (src/cu/solver.cu:41-57). */
void create_y_0(double *u, double *v, double *w, double *T, double *Y, double *y_0, Parameters P) {
    size_t n3 = (size_t)P.Nx * P.Ny * P.Nz;
    size_t ny = (size_t)P.Nx * P.Ny * P.Nz_Y_max;
    memcpy(y_0 + P.field_indexes.u, u, n3 * sizeof(double));
    memcpy(y_0 + P.field_indexes.v, v, n3 * sizeof(double));
    memcpy(y_0 + P.field_indexes.w, w, n3 * sizeof(double));
    memcpy(y_0 + P.field_indexes.T, T, n3 * sizeof(double));
    memcpy(y_0 + P.field_indexes.Y, Y, ny * sizeof(double));
}
