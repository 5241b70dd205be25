/* CPU run of the Stockham plan used by the HIP pressure kernels
 * (wildfires_amd/csrc/hip/wf_fft.h), for comparison with numpy.fft.
 * Built by tests/test_fft_plan.py with gcc; test code only. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../../wildfires_amd/csrc/hip/wf_fft.h"

static int factor(int n, int *rad) {
    int c = 0;
    while (n % 4 == 0) { rad[c++] = 4; n /= 4; }
    while (n % 2 == 0) { rad[c++] = 2; n /= 2; }
    for (int p = 3; n > 1; p += 2)
        while (n % p == 0) { rad[c++] = p; n /= p; }
    return c;
}

/* In-place DFT of `count` contiguous sequences of length n (interleaved re,im). */
int plan_fft(double *data, int n, int count, int sign) {
    int rad[WF_MAX_RADICES];
    int nr = factor(n, rad);
    wfc *tw = (wfc *)malloc(sizeof(wfc) * n);
    wfc *a = (wfc *)malloc(sizeof(wfc) * n), *b = (wfc *)malloc(sizeof(wfc) * n);
    for (int e = 0; e < n; e++) {
        double ang = 2.0 * 3.14159265358979323846 * (double)e / (double)n;
        tw[e].re = cos(ang);
        tw[e].im = sign * sin(ang);
    }
    for (int s = 0; s < count; s++) {
        memcpy(a, data + 2L * s * n, sizeof(wfc) * n);
        int Ns = 1;
        for (int p = 0; p < nr; p++) {
            for (int o = 0; o < n; o++) b[o] = wf_stockham_output(a, n, rad[p], Ns, tw, o);
            wfc *t = a; a = b; b = t;
            Ns *= rad[p];
        }
        memcpy(data + 2L * s * n, a, sizeof(wfc) * n);
    }
    free(tw); free(a); free(b);
    return nr;
}

/* Real-pair forward transform: rows x0, x1 (length n) -> half spectra of
 * length n/2+1 each, through one complex FFT and wf_unpack_pair. */
void plan_rfft_pair(const double *x0, const double *x1, int n, double *A, double *B) {
    double *z = (double *)malloc(sizeof(double) * 2 * n);
    for (int i = 0; i < n; i++) { z[2 * i] = x0[i]; z[2 * i + 1] = x1[i]; }
    plan_fft(z, n, 1, -1);
    int h = n / 2 + 1;
    for (int r = 0; r < h; r++) {
        wfc zr = {z[2 * r], z[2 * r + 1]};
        int m = (n - r) % n;
        wfc zm = {z[2 * m], z[2 * m + 1]};
        wfc a, bb;
        wf_unpack_pair(zr, zm, &a, &bb);
        A[2 * r] = a.re; A[2 * r + 1] = a.im;
        B[2 * r] = bb.re; B[2 * r + 1] = bb.im;
    }
    free(z);
}

/* Inverse of plan_rfft_pair (unnormalised): half spectra -> rows x0, x1. */
void plan_irfft_pair(const double *A, const double *B, int n, double *x0, double *x1) {
    double *z = (double *)malloc(sizeof(double) * 2 * n);
    int h = n / 2 + 1;
    for (int r = 0; r < n; r++) {
        int src = (r < h) ? r : n - r;
        wfc a = {A[2 * src], A[2 * src + 1]}, bb = {B[2 * src], B[2 * src + 1]};
        wfc v = wf_pack_pair(a, bb, r >= h);
        z[2 * r] = v.re; z[2 * r + 1] = v.im;
    }
    plan_fft(z, n, 1, +1);
    for (int i = 0; i < n; i++) { x0[i] = z[2 * i]; x1[i] = z[2 * i + 1]; }
    free(z);
}
