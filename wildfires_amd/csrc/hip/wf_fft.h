/*
 * wf_fft.h -- mixed-radix Stockham FFT building blocks for lengths with any
 * prime factors (the periodic x,y node counts Nx-1, Ny-1: 63, 127, 255, 510,
 * 511 ...).  Host/device-neutral so the same indexing runs in the LDS kernels
 * and in the CPU check of the plan.
 *
 * Stage with radix R after radices of product Ns (natural-order Stockham,
 * decimation in time): butterfly j in [0, n/R) reads in[j + r n/R], r < R, and
 * writes out[(j/Ns) Ns R + j%Ns + q Ns], q < R, with
 *   out = sum_r W_{Ns R}^{r (j%Ns + q Ns)} in[j + r n/R],  W_m = exp(sign 2 pi i / m).
 * Each output element is produced independently ("thread per output"), so a
 * pass is a fully parallel loop over n outputs.
 */
#ifndef WF_FFT_H
#define WF_FFT_H

#if defined(__HIPCC__)
#define WF_HD __host__ __device__ __forceinline__
#else
#define WF_HD static inline
#endif

#define WF_MAX_RADICES 16

typedef struct {
    double re, im;
} wfc;

/* Output element o of one Stockham stage over a sequence of length n stored at
 * `in`; tw[e] = exp(sign 2 pi i e / n). */
WF_HD wfc wf_stockham_output(const wfc *in, int n, int R, int Ns, const wfc *tw, int o) {
    int kq = o % Ns;
    int rest = o / Ns;
    int q = rest % R;
    int jb = (rest / R) * Ns + kq;
    int m = n / R;
    int scale = n / (Ns * R);
    int step = ((kq + q * Ns) * scale) % n;
    double ar = 0.0, ai = 0.0;
    int e = 0;
    for (int r = 0; r < R; r++) {
        wfc x = in[jb + r * m];
        wfc w = tw[e];
        ar += w.re * x.re - w.im * x.im;
        ai += w.re * x.im + w.im * x.re;
        e += step;
        if (e >= n) e -= n;
    }
    wfc y;
    y.re = ar;
    y.im = ai;
    return y;
}

/* Spectra of two real sequences a, b packed as z = a + i b, at bin r:
 * A[r] = (Z[r] + conj Z[n-r]) / 2, B[r] = (Z[r] - conj Z[n-r]) / (2i). */
WF_HD void wf_unpack_pair(wfc zr, wfc zm, wfc *A, wfc *B) {
    A->re = 0.5 * (zr.re + zm.re);
    A->im = 0.5 * (zr.im - zm.im);
    B->re = 0.5 * (zr.im + zm.im);
    B->im = 0.5 * (zm.re - zr.re);
}

/* Bin r of Z = A + i B for the inverse of two half spectra (r < n/2+1 read
 * directly, r above it from the conjugate-symmetric partner). */
WF_HD wfc wf_pack_pair(wfc A, wfc B, int conj) {
    wfc z;
    if (!conj) {
        z.re = A.re - B.im;
        z.im = A.im + B.re;
    } else {
        z.re = A.re + B.im;
        z.im = B.re - A.im;
    }
    return z;
}

#endif
