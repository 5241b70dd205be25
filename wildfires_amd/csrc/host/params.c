/*
 * params.c -- parameter-file front end of the drop-in (host C).
 *
 * Follows src/cu/parameters.cu:13-571: `key = value` lines matched by prefix
 * (strncmp over the key text), every matching key applied in table order so a
 * later duplicate line wins, defaults for the keys the reference defaults and
 * zero for the rest, then grids, wavenumbers, topography, the IBM mask and the
 * state-vector index tables.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

#include "wf_host.h"

#define WF_MAX_LINE 512

/* Value kinds of a parameter key. */
enum { K_INT, K_DBL, K_STR };

/* One recognised key: the text compared with strncmp, the number of characters
 * compared, the offset at which sscanf starts, and where the value goes. */
typedef struct {
    const char *text;
    int cmp_len;
    int value_at;
    int kind;
    size_t field;
    int cap; /* capacity of a string field */
} wf_key;

#define KI(t, n, o, f) {t, n, o, K_INT, offsetof(Parameters, f), 0}
#define KD(t, n, o, f) {t, n, o, K_DBL, offsetof(Parameters, f), 0}
#define KS(t, n, o, f) {t, n, o, K_STR, offsetof(Parameters, f), (int)sizeof(((Parameters *)0)->f)}

/* Keys in the order src/cu/parameters.cu:79-414 tests them.  The compare
 * lengths and value offsets are the reference's, including its short compares
 * ("T_min =" over 6 characters) and the prefix collisions ("k =", "A =", "g ="). */
static const wf_key wf_keys[] = {
    KI("Nx =", 4, 5, Nx),
    KI("Ny =", 4, 5, Ny),
    KI("Nz =", 4, 5, Nz),
    KI("Nt =", 4, 5, Nt),
    KI("NT =", 4, 5, NT),
    KD("x_min =", 7, 8, x_min),
    KD("x_max =", 7, 8, x_max),
    KD("y_min =", 7, 8, y_min),
    KD("y_max =", 7, 8, y_max),
    KD("z_min =", 7, 8, z_min),
    KD("z_max =", 7, 8, z_max),
    KD("t_min =", 7, 8, t_min),
    KD("t_max =", 7, 8, t_max),
    KD("nu =", 4, 5, nu),
    KD("mu =", 4, 5, mu),
    KD("u_r =", 5, 6, u_r),
    KD("z_r =", 5, 6, z_r),
    KD("alpha_u =", 9, 10, alpha_u),
    KD("alpha =", 7, 8, alpha),
    KD("T_inf =", 7, 8, T_inf),
    KD("T_0 =", 5, 6, T_0),
    KD("T_source =", 10, 11, T_source),
    KS("T0_shape =", 10, 11, T0_shape),
    KD("T0_x_start =", 12, 13, T0_x_start),
    KD("T0_x_end =", 10, 11, T0_x_end),
    KD("T0_y_start =", 12, 13, T0_y_start),
    KD("T0_y_end =", 10, 11, T0_y_end),
    KD("T0_z_start =", 12, 13, T0_z_start),
    KD("T0_z_end =", 10, 11, T0_z_end),
    KD("Y0_x_start =", 12, 13, Y0_x_start),
    KD("Y0_x_end =", 10, 11, Y0_x_end),
    KD("Y0_y_start =", 12, 13, Y0_y_start),
    KD("Y0_y_end =", 10, 11, Y0_y_end),
    KD("Y0_xa =", 7, 8, Y0_xa),
    KD("Y0_ya =", 7, 8, Y0_ya),
    KD("Y_h =", 5, 6, Y_h),
    KD("Y_f =", 5, 6, Y_f),
    KD("C_d =", 5, 6, C_d),
    KD("alpha_s =", 9, 10, alpha_s),
    KD("sigma_s =", 9, 10, sigma_s),
    KD("H_C =", 5, 6, H_C),
    KD("A =", 3, 4, A),
    KD("T_act =", 7, 8, T_act),
    KD("T_pc =", 6, 7, T_pc),
    KD("h_c =", 5, 6, h_c),
    KD("c_p =", 5, 6, c_p),
    KD("rho_inf =", 9, 10, rho_inf),
    KS("U0_type =", 9, 10, U0_type),
    KS("T0_shape =", 10, 11, T0_shape),
    KD("g =", 3, 4, g),
    KD("C_s =", 5, 6, C_s),
    KD("Pr =", 4, 5, Pr),
    KD("p_top =", 7, 8, p_top),
    KS("method =", 8, 9, method),
    KI("threads =", 9, 10, n_threads),
    KD("T_min =", 6, 7, T_min),
    KD("T_max =", 6, 7, T_max),
    KD("Y_min =", 6, 7, Y_min),
    KD("Y_max =", 6, 7, Y_max),
    KS("topo_shape =", 11, 12, topo_shape),
    KD("u_dead_nodes =", 13, 14, u_dead_nodes),
    KD("v_dead_nodes =", 13, 14, v_dead_nodes),
    KD("w_dead_nodes =", 13, 14, w_dead_nodes),
    KD("T_dead_nodes =", 13, 14, T_dead_nodes),
    KD("Y_dead_nodes =", 13, 14, Y_dead_nodes),
    KD("hill_center_x =", 14, 15, hill_center_x),
    KD("hill_center_y =", 14, 15, hill_center_y),
    KD("hill_length =", 12, 13, hill_length),
    KD("hill_width =", 11, 12, hill_width),
    KD("hill_height =", 12, 13, hill_height),
    KD("t_source_start =", 16, 17, t_source_start),
    KD("t_source_end =", 14, 15, t_source_end),
    KS("T_source_shape =", 16, 17, T_source_shape),
    KD("delta =", 7, 8, delta),
    KD("k =", 3, 4, k),
    KI("variable_density =", 17, 18, variable_density),
    KD("pressure_solver_tol =", 21, 22, pressure_solver_tol),
    KI("pressure_solver_iter =", 22, 23, pressure_solver_iter),
    KI("pressure_solver_log =", 21, 22, pressure_solver_log),
    KI("z_domain =", 10, 11, z_domain),
    KD("k_nu_grid =", 11, 12, k_nu_grid),
    KI("k_transition =", 14, 15, k_transition),
    KD("z_transition =", 14, 15, z_transition),
    KS("input_path =", 12, 13, input_path),
    KI("velocity_correction_fd =", 24, 25, velocity_correction_fd),
    KI("time_step_log =", 15, 16, time_step_log),
    KI("temperature_convection =", 24, 25, temperature_convection),
    KD("TS_x_start =", 12, 13, TS_x_start),
    KD("TS_x_end =", 10, 11, TS_x_end),
    KD("TS_y_start =", 12, 13, TS_y_start),
    KD("TS_y_end =", 10, 11, TS_y_end),
    KD("TS_z_start =", 12, 13, TS_z_start),
    KD("TS_z_end =", 10, 11, TS_z_end),
    KI("sutherland_viscosity =", 22, 23, sutherland_viscosity),
    KI("sutherland_conductivity =", 25, 26, sutherland_conductivity),
};

/* Defaults of src/cu/parameters.cu:39-76. */
static void wf_set_defaults(Parameters *p) {
    p->mu = 1.802e-5;
    p->nu = 1.47e-5;
    p->z_r = 2.0;
    p->alpha_u = 1.0 / 7.0;
    p->alpha = 2.009e-5;
    p->T_inf = 288.15;
    p->C_d = 0.15;
    p->H_C = 21e6;
    p->A = 1e9;
    p->T_act = 18040.8533;
    p->T_pc = 523.0;
    p->T_0 = 1000.0;
    p->h_c = 1.42;
    p->c_p = 1007.0;
    p->k = 0.02476;
    p->delta = 0.1;
    p->rho_inf = 1.225;
    p->g = -9.807;
    p->C_s = 0.2;
    p->Pr = 0.7329;
    p->variable_density = 1;
    p->pressure_solver_tol = 1e-10;
    p->pressure_solver_iter = 50;
    p->pressure_solver_log = 0;
    p->n_threads = 1;
    p->T_min = 288.15;
    p->T_max = 2500.0;
    p->Y_min = 0.0;
    p->Y_max = 1.0;
    p->u_dead_nodes = 0.0;
    p->v_dead_nodes = 0.0;
    p->w_dead_nodes = 0.0;
    p->T_dead_nodes = 288.15;
    p->Y_dead_nodes = 0.0;
    p->input_path[0] = '\0';
    p->velocity_correction_fd = 0;
    p->time_step_log = 100;
    p->temperature_convection = 1;
}

/* Apply one line of the parameter file: every key whose text is a prefix of
 * the line (over the key's compare length) is read, in table order. */
void wf_parse_parameter_line(Parameters *p, const char *line) {
    size_t nkeys = sizeof(wf_keys) / sizeof(wf_keys[0]);
    size_t len = strlen(line);
    for (size_t q = 0; q < nkeys; q++) {
        const wf_key *key = &wf_keys[q];
        if (strncmp(line, key->text, (size_t)key->cmp_len) != 0) continue;
        const char *val = (len >= (size_t)key->value_at) ? line + key->value_at : line + len;
        char *dst = (char *)p + key->field;
        if (key->kind == K_INT) {
            sscanf(val, "%d", (int *)dst);
        } else if (key->kind == K_DBL) {
            sscanf(val, "%lf", (double *)dst);
        } else {
            /* Bounded %s: the reference's unbounded read would overflow. */
            char fmt[16];
            snprintf(fmt, sizeof(fmt), "%%%ds", key->cap - 1);
            sscanf(val, fmt, dst);
        }
    }
}

/* YYYYMMDDHHMMSS of the local time (src/cu/utils.cu:53-60). */
void generate_current_datetime_string(char *datetime_str, size_t max_size) {
    time_t t = time(NULL);
    struct tm *tm = localtime(&t);
    if (tm != NULL) strftime(datetime_str, max_size, "%Y%m%d%H%M%S", tm);
}

/* hh:mm:ss of a duration (src/cu/utils.cu:62-67). */
void format_seconds(double seconds, char *formatted_time) {
    int hours = (int)(seconds / 3600);
    int minutes = ((int)seconds % 3600) / 60;
    int secs = ((int)seconds % 3600) % 60;
    sprintf(formatted_time, "%02d:%02d:%02d", hours, minutes, secs);
}

/* Fourier sample frequencies in numpy.fft.fftfreq order (src/cu/utils.cu:32-51). */
void fft_freq(double *f, int N, double d) {
    double scale = N * d;
    int half = (N % 2 == 0) ? N / 2 : (N - 1) / 2 + 1;
    for (int i = 0; i < half; i++) f[i] = i / scale;
    for (int i = half; i < N; i++) f[i] = (i - N) / scale;
}

/* z grids (src/cu/functions.cu:58-80). */
void equispaced_domain(double *z, int Nz, double z_min, double dz) {
    for (int k = 0; k < Nz; k++) z[k] = z_min + k * dz;
}

void non_equispaced_domain(double *z, int Nz, double z_min, double z_max, double k) {
    double dz = (z_max - z_min) / (Nz - 1);
    for (int q = 0; q < Nz; q++)
        z[q] = (z_max - z_min) * (exp(k * q * dz / (z_max - z_min)) - 1) / (exp(k) - 1) + z_min;
}

void transition_domain(double *z, int Nz, double z_min, double z_max, double z_t, int k_t) {
    double dz_1 = (z_t - z_min) / (k_t - 1);
    double dz_2 = (z_max - z_t) / (Nz - k_t - 1);
    for (int k = 0; k < Nz; k++) z[k] = (k <= k_t) ? z_min + k * dz_1 : z_t + (k - k_t) * dz_2;
}

/* mkdir -p of a directory path ending in '/'. */
static void wf_make_dirs(const char *path) {
    char buf[512];
    size_t n = strlen(path);
    if (n >= sizeof(buf)) n = sizeof(buf) - 1;
    memcpy(buf, path, n);
    buf[n] = '\0';
    for (size_t q = 1; q < n; q++) {
        if (buf[q] == '/') {
            buf[q] = '\0';
            mkdir(buf, 0777);
            buf[q] = '/';
        }
    }
    mkdir(buf, 0777);
}

/* Grids, wavenumbers, terrain, IBM and index tables from the parsed values
 * (src/cu/parameters.cu:416-523). */
void wf_build_derived(Parameters *p) {
    int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz, Nt = p->Nt;
    p->x = (double *)malloc((size_t)Nx * sizeof(double));
    p->y = (double *)malloc((size_t)Ny * sizeof(double));
    p->z = (double *)malloc((size_t)Nz * sizeof(double));
    p->t = (double *)malloc((size_t)(Nt + 1) * sizeof(double));
    p->r = (double *)malloc((size_t)(Nx - 1) * sizeof(double));
    p->s = (double *)malloc((size_t)(Ny - 1) * sizeof(double));
    p->kx = (double *)malloc((size_t)(Nx - 1) * sizeof(double));
    p->ky = (double *)malloc((size_t)(Ny - 1) * sizeof(double));
    p->cut_nodes = (int *)calloc((size_t)Nx * Ny, sizeof(int));
    p->z_ibm = (double *)calloc((size_t)Nx * Ny * Nz, sizeof(double));
    /* zero-filled: an unknown topo_shape leaves a flat terrain */
    p->topography = (double *)calloc((size_t)Nx * Ny, sizeof(double));
    p->Nz_Y = (int *)calloc((size_t)Nx * Ny, sizeof(int));
    p->dx = (p->x_max - p->x_min) / (Nx - 1);
    p->dy = (p->y_max - p->y_min) / (Ny - 1);
    p->dz = (p->z_max - p->z_min) / (Nz - 1);
    p->dt = (p->t_max - p->t_min) / Nt;
    fft_freq(p->r, Nx - 1, 1.0 / (Nx - 1));
    fft_freq(p->s, Ny - 1, 1.0 / (Ny - 1));
    for (int n = 0; n <= Nt; n++) p->t[n] = p->t_min + n * p->dt;
    for (int i = 0; i < Nx; i++) {
        p->x[i] = p->x_min + i * p->dx;
        if (i < Nx - 1) p->kx[i] = 2 * M_PI * p->r[i] / (p->x_max - p->x_min);
    }
    for (int j = 0; j < Ny; j++) {
        p->y[j] = p->y_min + j * p->dy;
        if (j < Ny - 1) p->ky[j] = 2 * M_PI * p->s[j] / (p->y_max - p->y_min);
    }
    if (p->z_domain == 0)
        equispaced_domain(p->z, Nz, p->z_min, p->dz);
    else if (p->z_domain == 1)
        non_equispaced_domain(p->z, Nz, p->z_min, p->z_max, p->k_nu_grid);
    else if (p->z_domain == 2)
        transition_domain(p->z, Nz, p->z_min, p->z_max, p->z_transition, p->k_transition);
    p->T0_x_center = (p->T0_x_start + p->T0_x_end) / 2;
    p->T0_y_center = (p->T0_y_start + p->T0_y_end) / 2;
    p->T0_z_center = 0.0;
    p->T0_length = p->T0_x_end - p->T0_x_start;
    p->T0_width = p->T0_y_end - p->T0_y_start;
    p->T0_height = p->T0_z_end - p->T0_z_start;
    p->TS_x_center = (p->TS_x_start + p->TS_x_end) / 2.0;
    p->TS_y_center = (p->TS_y_start + p->TS_y_end) / 2.0;
    p->TS_z_center = 0.0;
    p->TS_length = p->TS_x_end - p->TS_x_start;
    p->TS_width = p->TS_y_end - p->TS_y_start;
    p->TS_height = p->TS_z_end - p->TS_z_start;
    if (strcmp(p->topo_shape, "hill") == 0)
        simple_hill(p);
    else if (strcmp(p->topo_shape, "flat") == 0)
        flat_terrain(p);
    ibm_parameters(p);
    int size = Nx * Ny * Nz;
    p->field_indexes.u = 0;
    p->field_indexes.v = size;
    p->field_indexes.w = 2 * size;
    p->field_indexes.T = 3 * size;
    p->field_indexes.Y = 4 * size;
    /* R_turbulence plane offsets (src/cu/parameters.cu:505-523) */
    int *ti = &p->turbulence_indexes.rho;
    for (int q = 0; q < 19; q++) ti[q] = q * size;
}

Parameters read_parameters_file(const char *file_path) {
    Parameters parameters;
    char line[WF_MAX_LINE];
    char path[512];
    memset(&parameters, 0, sizeof(Parameters));
    FILE *f = fopen(file_path, "r");
    if (f == NULL) {
        fprintf(stderr, "Error opening parameters file %s: %s\n", file_path, strerror(errno));
        exit(1);
    }
    generate_current_datetime_string(parameters.sim_id, 32);
    const char *root = getenv("WILDFIRES_OUTPUT_ROOT");
    int need;
    if (root != NULL && root[0] != '\0')
        need = snprintf(parameters.save_path, sizeof(parameters.save_path), "%s/%s/", root, parameters.sim_id);
    else
        need = snprintf(parameters.save_path, sizeof(parameters.save_path), "data/output/%s/", parameters.sim_id);
    if (need >= (int)sizeof(parameters.save_path)) {
        /* save_path is a 64-byte member of Parameters (include/c/structures.h:238) */
        fprintf(stderr, "Output directory path too long (%d characters, at most %d)\n", need,
                (int)sizeof(parameters.save_path) - 1);
        exit(1);
    }
    wf_make_dirs(parameters.save_path);
    snprintf(path, sizeof(path), "%sparameters.txt", parameters.save_path);
    parameters.log_files.parameters = fopen(path, "w");
    snprintf(path, sizeof(path), "%slog.txt", parameters.save_path);
    parameters.log_files.log = fopen(path, "w");
    wf_set_defaults(&parameters);
    while (fgets(line, WF_MAX_LINE, f) != NULL) wf_parse_parameter_line(&parameters, line);
    fclose(f);
    wf_build_derived(&parameters);
    return parameters;
}

void free_parameters(Parameters parameters) {
    free(parameters.x);
    free(parameters.y);
    free(parameters.z);
    free(parameters.t);
    free(parameters.r);
    free(parameters.s);
    free(parameters.kx);
    free(parameters.ky);
    free(parameters.cut_nodes);
    free(parameters.z_ibm);
    free(parameters.topography);
    free(parameters.Nz_Y);
    if (parameters.log_files.log != NULL) fclose(parameters.log_files.log);
}

Parameters *wf_parameters_read(const char *file_path) {
    Parameters *p = (Parameters *)malloc(sizeof(Parameters));
    if (p == NULL) return NULL;
    *p = read_parameters_file(file_path);
    return p;
}

void wf_parameters_free(Parameters *parameters) {
    if (parameters == NULL) return;
    free_parameters(*parameters);
    free(parameters);
}

size_t wf_parameters_sizeof(void) { return sizeof(Parameters); }

long wf_state_size(const Parameters *p) {
    return 4L * p->Nx * p->Ny * p->Nz + (long)p->Nx * p->Ny * p->Nz_Y_max;
}
