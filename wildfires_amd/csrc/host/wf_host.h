/*
 * wf_host.h -- internal declarations shared by the host C front end.
 */
#ifndef WF_HOST_H
#define WF_HOST_H

#include "../../../include/wildfires.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* Scalar-field index of node (i,j,k) in the reference layout (z fastest). */
#define WF_IDX(i, j, k, Nx, Ny, Nz) ((k) + (long)(Nz) * ((j) + (long)(Ny) * (i)))

void wf_parse_parameter_line(Parameters *p, const char *line);
void wf_build_derived(Parameters *p);
void generate_current_datetime_string(char *datetime_str, size_t max_size);
void format_seconds(double seconds, char *formatted_time);

#endif
