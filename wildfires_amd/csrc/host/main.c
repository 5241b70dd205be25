/*
 * main.c -- command-line driver `wildfires_amd <parameters_file>`, the same
 * sequence as src/cu/main.cu:3-94 with host-resident initial state.
 */
#include <stdio.h>
#include <stdlib.h>

#include "../../../include/wildfires.h"

int main(int argc, char *argv[]) {
    if (argc != 2) {
        printf("Usage: %s <parameters_file_path>\n", argv[0]);
        return 1;
    }
    Parameters parameters = read_parameters_file(argv[1]);
    log_parameters(parameters);
    size_t size = (size_t)parameters.Nx * parameters.Ny * parameters.Nz;
    size_t size_Y = (size_t)parameters.Nx * parameters.Ny * parameters.Nz_Y_max;
    double *u = (double *)malloc(size * sizeof(double));
    double *v = (double *)malloc(size * sizeof(double));
    double *w = (double *)malloc(size * sizeof(double));
    double *T = (double *)malloc(size * sizeof(double));
    double *p = (double *)malloc(size * sizeof(double));
    double *Y = (double *)calloc(size_Y > 0 ? size_Y : 1, sizeof(double));
    double *y_0 = (double *)malloc((4 * size + size_Y) * sizeof(double));
    if (parameters.input_path[0] == '\0') {
        log_message(parameters, "Initial conditions...");
        initial_conditions(u, v, w, T, Y, p, parameters);
    } else {
        initial_conditions_input(u, v, w, T, Y, p, parameters);
        log_message(parameters, "Initial conditions loaded from files.");
    }
    log_message(parameters, "Creating y_0...");
    create_y_0(u, v, w, T, Y, y_0, parameters);
    log_message(parameters, "Saving initial data...");
    save_data(y_0, p, 0, parameters);
    save_domain(parameters);
    save_topography(parameters);
    log_message(parameters, "Solving PDE...");
    solve_PDE(y_0, p, parameters);
    log_message(parameters, "Freeing memory...");
    free(u);
    free(v);
    free(w);
    free(T);
    free(p);
    free(Y);
    free(y_0);
    free_parameters(parameters);
    printf("Memory freed!\n");
    printf("Simulation name: %s\n", parameters.sim_id);
    return 0;
}
