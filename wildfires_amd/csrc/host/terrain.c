/*
 * terrain.c -- topography and the immersed-boundary terrain mask (host C).
 *
 * The mask is integer output of FP64 comparisons, so every comparison and
 * every operand is evaluated exactly as src/cu/ibm.cu:3-95 and
 * src/cu/functions.cu:18-32 evaluate them.
 */
#include <math.h>

#include "wf_host.h"

/* u_r (z/z_r)^alpha_u (src/cu/functions.cu:18-20). */
double power_law(double z, double u_r, double z_r, double alpha_u) {
    return u_r * pow(z / z_r, alpha_u);
}

/* exp(-((x-x0)/sx)^2 - ((y-y0)/sy)^2 - ((z-z0)/sz)^2) with the squares taken by
 * pow(.,2.0) as in src/cu/functions.cu:22-24 (no factor 2 in the exponent). */
double gaussian(double x, double y, double z, double x_0, double y_0, double z_0, double sx, double sy, double sz) {
    double ex = pow((x - x_0) / sx, 2.0);
    double ey = pow((y - y_0) / sy, 2.0);
    double ez = pow((z - z_0) / sz, 2.0);
    return exp(-ex - ey - ez);
}

/* Indicator of the closed box (src/cu/functions.cu:26-32). */
double parallelepiped(double x, double y, double z, double x_min, double x_max, double y_min,
                      double y_max, double z_min, double z_max) {
    int inside = x >= x_min && x <= x_max && y >= y_min && y <= y_max && z >= z_min && z <= z_max;
    return inside ? 1.0 : 0.0;
}

/* src/cu/ibm.cu:3-11 */
void flat_terrain(Parameters *parameters) {
    long n = (long)parameters->Nx * parameters->Ny;
    for (long q = 0; q < n; q++) parameters->topography[q] = 0.0;
}

/* src/cu/ibm.cu:13-28 */
void simple_hill(Parameters *parameters) {
    int Nx = parameters->Nx, Ny = parameters->Ny;
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny; j++)
            parameters->topography[WF_IDX(i, j, 0, Nx, Ny, 1)] =
                parameters->hill_height * gaussian(parameters->x[i], parameters->y[j], 0,
                                                   parameters->hill_center_x, parameters->hill_center_y, 0,
                                                   parameters->hill_length, parameters->hill_width, 1);
}

/* Cut node, fuel depth and wall distance per column (src/cu/ibm.cu:57-95):
 *   cut_nodes = last k with topo - dz_k <= z_k <= topo (0 if none),
 *   Nz_Y      = 1 + last k with z_k <= topo + Y_h (0 if none),
 *   z_ibm     = max(z_k - topo, 0),
 * with dz_k the spacing above node k (below it at the top node). */
void ibm_parameters(Parameters *parameters) {
    int Nx = parameters->Nx, Ny = parameters->Ny, Nz = parameters->Nz;
    const double *z = parameters->z;
    double Y_h = parameters->Y_h;
    int nz_y_max = 0;
    for (int i = 0; i < Nx; i++) {
        for (int j = 0; j < Ny; j++) {
            long col = WF_IDX(i, j, 0, Nx, Ny, 1);
            double topo = parameters->topography[col];
            int cut = 0, nzy = 0;
            for (int k = 0; k < Nz; k++) {
                double dz = (k < Nz - 1) ? z[k + 1] - z[k] : z[k] - z[k - 1];
                if ((topo - dz) <= z[k] && z[k] <= topo) cut = k;
                double dist = z[k] - topo;
                parameters->z_ibm[WF_IDX(i, j, k, Nx, Ny, Nz)] = (dist > 0.0) ? dist : 0.0;
                if (z[k] <= topo + Y_h) nzy = k + 1;
            }
            parameters->cut_nodes[col] = cut;
            parameters->Nz_Y[col] = nzy;
            if (nzy > nz_y_max) nz_y_max = nzy;
        }
    }
    parameters->Nz_Y_max = nz_y_max;
}
