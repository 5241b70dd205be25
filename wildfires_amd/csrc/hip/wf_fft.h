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

#if defined(__HIPCC__)
/* ------------------------------------------------------------------------
 * Compile-time plans for the transform lengths of the supported grids.
 * The radix sequence of N is the one the host factoriser produces (4s, then
 * 2, then odd primes ascending), so every stage's sizes are constants and the
 * index arithmetic folds into multiplications and shifts.
 * ---------------------------------------------------------------------- */

constexpr int wf_radix_of(int n, int idx) {
    int cnt = 0;
    while (n % 4 == 0) {
        if (cnt == idx) return 4;
        cnt++;
        n /= 4;
    }
    while (n % 2 == 0) {
        if (cnt == idx) return 2;
        cnt++;
        n /= 2;
    }
    for (int p = 3; n > 1; p += 2) {
        while (n % p == 0) {
            if (cnt == idx) return p;
            cnt++;
            n /= p;
        }
    }
    return 0;
}

__device__ __forceinline__ wfc wf_cmul(wfc a, wfc b) {
    wfc r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.re * b.im + a.im * b.re;
    return r;
}

/* One Stockham stage of radix R after radices of product NS over `nseq`
 * sequences of length N (same mathematics as wf_stockham_output).  Radices up
 * to 8 run one butterfly per thread with its R values in registers; larger
 * prime radices run one output per thread to keep the workgroup busy. */
template <int N, int R, int NS>
__device__ __forceinline__ void wf_stage_ct(const wfc *in, wfc *out, int nseq, const wfc *tw) {
    constexpr int M = N / R;
    constexpr int SC = N / (NS * R);
    constexpr int WR = N / R; /* W_R^m = tw[m * WR] */
    if constexpr (R <= 8) {
        for (int t = threadIdx.x; t < nseq * M; t += blockDim.x) {
            int s = t / M, j = t - s * M;
            int kq = j % NS;
            const wfc *src = in + s * N + j;
            wfc v[R];
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = src[r * M];
#pragma unroll
            for (int r = 1; r < R; r++) v[r] = wf_cmul(v[r], tw[kq * r * SC]);
            wfc *dst = out + s * N + (j / NS) * NS * R + kq;
#pragma unroll
            for (int q = 0; q < R; q++) {
                wfc acc = v[0];
#pragma unroll
                for (int r = 1; r < R; r++) {
                    wfc p = wf_cmul(v[r], tw[((r * q) % R) * WR]);
                    acc.re += p.re;
                    acc.im += p.im;
                }
                dst[q * NS] = acc;
            }
        }
    } else {
        for (int t = threadIdx.x; t < nseq * N; t += blockDim.x) {
            int s = t / N, o = t - s * N;
            int kq = o % NS;
            int rest = o / NS;
            int q = rest % R;
            int jb = (rest / R) * NS + kq;
            int step = (kq + q * NS) * SC; /* < N */
            const wfc *src = in + s * N + jb;
            double ar = 0.0, ai = 0.0;
            int e = 0;
#pragma unroll 4
            for (int r = 0; r < R; r++) {
                wfc x = src[r * M];
                wfc w = tw[e];
                ar += w.re * x.re - w.im * x.im;
                ai += w.re * x.im + w.im * x.re;
                e += step;
                if (e >= N) e -= N;
            }
            out[s * N + o].re = ar;
            out[s * N + o].im = ai;
        }
    }
}

template <int N, int IDX, int NS>
struct WfStages {
    static __device__ __forceinline__ wfc *run(wfc *a, wfc *b, int nseq, const wfc *tw) {
        constexpr int R = wf_radix_of(N, IDX);
        if constexpr (R == 0) {
            return a;
        } else {
            wf_stage_ct<N, R, NS>(a, b, nseq, tw);
            __syncthreads();
            return WfStages<N, IDX + 1, NS * R>::run(b, a, nseq, tw);
        }
    }
};

/* Full transform of `nseq` sequences of compile-time length N held in LDS
 * buffer a (b is scratch); returns the buffer holding the result. */
template <int N>
__device__ __forceinline__ wfc *wf_fft_ct(wfc *a, wfc *b, int nseq, const wfc *tw) {
    return WfStages<N, 0, 1>::run(a, b, nseq, tw);
}
#endif

#endif /* WF_FFT_H */
