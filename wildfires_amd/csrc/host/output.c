/*
 * output.c -- snapshots, domain/terrain files, logs and checkpoint reports (host C).
 *
 * File names, formats and message texts follow src/cu/output.cu:3-98,
 * src/cu/logs.cu:3-154 and src/cu/functions.cu:82-185 so that existing
 * post-processing (src/py/data_conversion.py:29-62) reads the output unchanged.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "wf_host.h"

#define WF_FILENAME 192
#define WF_MIN(a, b) (((a) < (b)) ? (a) : (b))
#define WF_MAX(a, b) (((a) > (b)) ? (a) : (b))

static void wf_write_file(const char *path, const void *data, size_t elem, size_t count) {
    FILE *f = fopen(path, "wb");
    if (f == NULL) {
        fprintf(stderr, "Error opening output file %s\n", path);
        return;
    }
    fwrite(data, elem, count, f);
    fclose(f);
}

/* u,v,w,T,p as Nx*Ny*Nz float64 and Y as Nx*Ny*Nz_Y_max float64, suffix .n. */
void save_data(double *data, double *p, int n, Parameters P) {
    size_t n3 = (size_t)P.Nx * P.Ny * P.Nz;
    size_t ny = (size_t)P.Nx * P.Ny * P.Nz_Y_max;
    char fn[WF_FILENAME];
    snprintf(fn, sizeof(fn), "%su.bin.%d", P.save_path, n);
    wf_write_file(fn, data + P.field_indexes.u, sizeof(double), n3);
    snprintf(fn, sizeof(fn), "%sv.bin.%d", P.save_path, n);
    wf_write_file(fn, data + P.field_indexes.v, sizeof(double), n3);
    snprintf(fn, sizeof(fn), "%sw.bin.%d", P.save_path, n);
    wf_write_file(fn, data + P.field_indexes.w, sizeof(double), n3);
    snprintf(fn, sizeof(fn), "%sT.bin.%d", P.save_path, n);
    wf_write_file(fn, data + P.field_indexes.T, sizeof(double), n3);
    snprintf(fn, sizeof(fn), "%sp.bin.%d", P.save_path, n);
    wf_write_file(fn, p, sizeof(double), n3);
    snprintf(fn, sizeof(fn), "%sY.bin.%d", P.save_path, n);
    wf_write_file(fn, data + P.field_indexes.Y, sizeof(double), ny);
}

/* x.bin, y.bin, z.bin and t.bin with the checkpoint times only. */
void save_domain(Parameters P) {
    char fn[WF_FILENAME];
    snprintf(fn, sizeof(fn), "%sx.bin", P.save_path);
    wf_write_file(fn, P.x, sizeof(double), (size_t)P.Nx);
    snprintf(fn, sizeof(fn), "%sy.bin", P.save_path);
    wf_write_file(fn, P.y, sizeof(double), (size_t)P.Ny);
    snprintf(fn, sizeof(fn), "%sz.bin", P.save_path);
    wf_write_file(fn, P.z, sizeof(double), (size_t)P.Nz);
    snprintf(fn, sizeof(fn), "%st.bin", P.save_path);
    FILE *f = fopen(fn, "wb");
    if (f == NULL) return;
    for (int n = 0; n <= P.Nt; n++)
        if (n % P.NT == 0 || n == P.Nt) fwrite(&P.t[n], sizeof(double), 1, f);
    fclose(f);
}

/* topography.bin (float64) and Nz_Y.bin (int32). */
void save_topography(Parameters P) {
    char fn[WF_FILENAME];
    snprintf(fn, sizeof(fn), "%stopography.bin", P.save_path);
    wf_write_file(fn, P.topography, sizeof(double), (size_t)P.Nx * P.Ny);
    snprintf(fn, sizeof(fn), "%sNz_Y.bin", P.save_path);
    wf_write_file(fn, P.Nz_Y, sizeof(int), (size_t)P.Nx * P.Ny);
}

void log_parameters_report(Parameters P, int to_file) {
    FILE *out = to_file ? P.log_files.parameters : stdout;
    double dz_min = INFINITY, dz_max = -INFINITY;
    setbuf(stdout, NULL);
    if (out == NULL) return;
    fprintf(out, "Simulation name: %s\n", P.sim_id);
    fprintf(out, "Domain:\n");
    fprintf(out, "  [%lf, %lf] x [%lf, %lf] x [%lf, %lf] x [%lf, %lf]\n", P.x_min, P.x_max, P.y_min, P.y_max,
            P.z_min, P.z_max, P.t_min, P.t_max);
    fprintf(out, "Grid size:\n");
    fprintf(out, "  Nx: %d, Ny: %d, Nz: %d, Nt: %d\n", P.Nx, P.Ny, P.Nz, P.Nt);
    if (P.z_domain == 0) {
        fprintf(out, "  dx: %lf, dy: %lf, dz: %lf, dt: %lf\n", P.dx, P.dy, P.dz, P.dt);
    } else {
        for (int k = 0; k < P.Nz - 1; k++) {
            dz_min = WF_MIN(dz_min, P.z[k + 1] - P.z[k]);
            dz_max = WF_MAX(dz_max, P.z[k + 1] - P.z[k]);
        }
        fprintf(out, "  dx: %lf, dy: %lf, dz: [%lf, %lf], dt: %lf\n", P.dx, P.dy, dz_min, dz_max, P.dt);
        if (P.z_domain == 1) fprintf(out, "  Stretched grid in z-direction, k: %lf\n", P.k_nu_grid);
    }
    fprintf(out, "  Time samples: %d\n", P.NT);
    fprintf(out, "Time integration: %s\n", P.method);
    fprintf(out, "Fluid parameters:\n");
    fprintf(out, "  mu: %e, k: %e, c_p: %lf\n", P.mu, P.k, P.c_p);
    fprintf(out, "  rho_inf: %lf, T_inf: %lf,  g: %lf\n", P.rho_inf, P.T_inf, P.g);
    fprintf(out, "  C_s: %lf, Pr: %lf\n", P.C_s, P.Pr);
    if (P.sutherland_viscosity == 1) fprintf(out, "  Sutherland's law for viscosity activated.\n");
    if (P.sutherland_conductivity == 1) fprintf(out, "  Sutherland's law for conductivity activated.\n");
    fprintf(out, "Fuel parameters:\n");
    fprintf(out, "  h_c: %lf, alpha_s: %lf, sigma_s: %lf, delta: %lf\n", P.h_c, P.alpha_s, P.sigma_s, P.delta);
    fprintf(out, "  H_C: %e, A: %e, T_pc: %lf, T_act: %lf\n", P.H_C, P.A, P.T_pc, P.T_act);
    fprintf(out, "  C_d: %lf, Y_f: %lf\n", P.C_d, P.Y_f);
    if (P.input_path[0] == '\0') {
        fprintf(out, "Wind initial condition:\n");
        fprintf(out, "  Type: %s\n", P.U0_type);
        if (strcmp(P.U0_type, "constant") == 0)
            fprintf(out, "  u_r: %lf\n", P.u_r);
        else if (strcmp(P.U0_type, "log") == 0)
            fprintf(out, "  u_z0: %lf, kappa: %lf, d: %lf, u_ast: %lf\n", P.u_z0, 0.41, P.d, P.u_ast);
        else if (strcmp(P.U0_type, "power_law") == 0)
            fprintf(out, "  u_r: %lf, z_r: %lf, alpha_u: %lf\n", P.u_r, P.z_r, P.alpha_u);
        fprintf(out, "Temperature initial condition:\n");
        fprintf(out, "  Shape: %s\n", P.T0_shape);
        fprintf(out, "  T_0: %lf\n", P.T_0);
        fprintf(out, "  x: [%lf, %lf], y: [%lf, %lf], z: [%lf, %lf]\n", P.T0_x_start, P.T0_x_end, P.T0_y_start,
                P.T0_y_end, P.T0_z_start, P.T0_z_end);
        fprintf(out, "  Center: [%lf, %lf, %lf], Length: %lf, Width: %lf, Height: %lf\n", P.T0_x_center,
                P.T0_y_center, P.T0_z_center, P.T0_length, P.T0_width, P.T0_height);
        fprintf(out, "Fuel initial condition:\n");
        fprintf(out, "  Fuel height: %lf\n", P.Y_h);
        fprintf(out, "  x: [%lf, %lf], y: [%lf, %lf]\n", P.Y0_x_start, P.Y0_x_end, P.Y0_y_start, P.Y0_y_end);
        fprintf(out, "  Fuel smooth x: %lf, y: %lf\n", P.Y0_xa, P.Y0_ya);
    }
    if (P.t_source_end > 0) {
        fprintf(out, "Temperature source:\n");
        fprintf(out, "  Shape: %s\n", P.T_source_shape);
        fprintf(out, "  T_source: %lf\n", P.T_source);
        fprintf(out, "  Source time: [%lf, %lf]\n", P.t_source_start, P.t_source_end);
        fprintf(out, "  x: [%lf, %lf], y: [%lf, %lf], z: [%lf, %lf]\n", P.TS_x_start, P.TS_x_end, P.TS_y_start,
                P.TS_y_end, P.TS_z_start, P.TS_z_end);
        fprintf(out, "  Center: [%lf, %lf, %lf], Length: %lf, Width: %lf, Height: %lf\n", P.TS_x_center,
                P.TS_y_center, P.TS_z_center, P.TS_length, P.TS_width, P.TS_height);
    }
    fprintf(out, "Topography:\n");
    fprintf(out, "  Shape: %s\n", P.topo_shape);
    if (strcmp(P.topo_shape, "hill") == 0) {
        fprintf(out, "  Hill center: [%lf, %lf], Length: %lf, Width: %lf, Height: %lf\n", P.hill_center_x,
                P.hill_center_y, P.hill_length, P.hill_width, P.hill_height);
        fprintf(out, "  Dead nodes:  u = %lf, v = %lf, w = %lf, T = %lf, Y = %lf\n", P.u_dead_nodes,
                P.v_dead_nodes, P.w_dead_nodes, P.T_dead_nodes, P.Y_dead_nodes);
    }
    fprintf(out, "Bounds:\n");
    fprintf(out, "  Temperature: [%lf, %lf]\n", P.T_min, P.T_max);
    fprintf(out, "  Fuel: [%lf, %lf]\n", P.Y_min, P.Y_max);
    if (P.variable_density == 1) {
        fprintf(out, "Pressure solver for variable density:\n");
        fprintf(out, "  Tolerance: %e, Max iterations: %d\n", P.pressure_solver_tol, P.pressure_solver_iter);
    }
    fprintf(out, "Velocity correction in z-direction: ");
    if (P.velocity_correction_fd == 0)
        fprintf(out, "Central difference.\n");
    else if (P.velocity_correction_fd == 1)
        fprintf(out, "Forward difference.\n");
    else if (P.velocity_correction_fd == -1)
        fprintf(out, "Backward difference.\n");
    fprintf(out, "Temperature convection: ");
    if (P.temperature_convection == 0)
        fprintf(out, "Finite difference.\n");
    else if (P.temperature_convection == 1)
        fprintf(out, "Upwind.\n");
    if (to_file) fclose(out);
}

void log_parameters(Parameters P) {
    log_parameters_report(P, 0);
    log_parameters_report(P, 1);
}

void log_message(Parameters P, const char *message) {
    setbuf(stdout, NULL);
    if (P.log_files.log != NULL) {
        fprintf(P.log_files.log, "%s\n", message);
        fflush(P.log_files.log);
    }
    printf("%s\n", message);
}

void log_timestep(Parameters P, int n, double t_n, double step_time, double CFL, double T_min, double T_max,
                  double Y_min, double Y_max) {
    FILE *outs[2] = {stdout, P.log_files.log};
    setbuf(stdout, NULL);
    for (int q = 0; q < 2; q++) {
        FILE *o = outs[q];
        if (o == NULL) continue;
        fprintf(o, "\nSimulation checkpoint\n");
        fprintf(o, "Time step: %6d, Simulation time: %lf s\n", n, t_n);
        fprintf(o, "CFL: %lf\n", CFL);
        fprintf(o, "Temperature: Min = %.2f, Max = %.2f\n", T_min, T_max);
        fprintf(o, "Fuel: Min = %.2f, Max = %.2f\n", Y_min, Y_max);
        fprintf(o, "Step time: %lf s\n", step_time);
    }
    if (P.log_files.log != NULL) fflush(P.log_files.log);
}

/* Checkpoint report (src/cu/functions.cu:82-185): CFL over all nodes with the
 * uniform dz, T and Y extrema inside the fuel box, exit on the first NaN in
 * u, v, w, T or Y (scanned in i, j, k order). */
void timestep_reports(double *y_n, double *CFL, double *Y_min, double *Y_max, double *T_min, double *T_max, int n,
                      Parameters P) {
    int Nx = P.Nx, Ny = P.Ny, Nz = P.Nz, Nzm = P.Nz_Y_max;
    double max_u = 0.0, max_v = 0.0, max_w = 0.0;
    double y_lo = 0.0, y_hi = -1e9, t_lo = 1e9, t_hi = -1e9;
    const char *names[5] = {"u", "v", "w", "T", "Y"};
    char message[256];
    for (int i = 0; i < Nx; i++) {
        for (int j = 0; j < Ny; j++) {
            int in_box = P.Y0_x_start <= P.x[i] && P.x[i] <= P.Y0_x_end && P.Y0_y_start <= P.y[j] &&
                         P.y[j] <= P.Y0_y_end;
            for (int k = 0; k < Nz; k++) {
                long q = WF_IDX(i, j, k, Nx, Ny, Nz);
                double vals[5];
                vals[0] = y_n[P.field_indexes.u + q];
                vals[1] = y_n[P.field_indexes.v + q];
                vals[2] = y_n[P.field_indexes.w + q];
                vals[3] = y_n[P.field_indexes.T + q];
                vals[4] = 0;
                if (k < Nzm) {
                    vals[4] = y_n[P.field_indexes.Y + WF_IDX(i, j, k, Nx, Ny, Nzm)];
                    if (in_box) {
                        y_lo = WF_MIN(y_lo, vals[4]);
                        y_hi = WF_MAX(y_hi, vals[4]);
                    }
                }
                for (int f = 0; f < 5; f++) {
                    if (isnan(vals[f])) {
                        sprintf(message,
                                "NaN value found in %s(%f, %f, %f, %f) in node (%d, %d, %d) at timestep %d (%d). Exiting...",
                                names[f], P.x[i], P.y[j], P.z[k], P.t[n], i, j, k, n, n / P.NT);
                        log_message(P, message);
                        exit(1);
                    }
                }
                max_u = WF_MAX(max_u, fabs(vals[0]));
                max_v = WF_MAX(max_v, fabs(vals[1]));
                max_w = WF_MAX(max_w, fabs(vals[2]));
                if (in_box) {
                    t_lo = WF_MIN(t_lo, vals[3]);
                    t_hi = WF_MAX(t_hi, vals[3]);
                }
            }
        }
    }
    *CFL = P.dt * (max_u / P.dx + max_v / P.dy + max_w / P.dz);
    *Y_min = y_lo;
    *Y_max = y_hi;
    *T_min = t_lo;
    *T_max = t_hi;
}
