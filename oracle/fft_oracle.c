/*
 * fft_oracle.c -- CPU FP64 oracle for the SiP batched FFT path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this file; only tests/, the smoke check and the CPU-baseline leg of
 * the benchmark use it, and only as the checker.
 *
 * What it restates (all citations relative to the reference tree):
 *   - the transform semantics of include/fft_api.h as pinned by the reference
 *     unit-test goldens:
 *       C2C fwd  = np.fft.fft                tests/ut/unittest/fft/1d/c2c/c2c_data/gen_data.py:42-45
 *       C2C inv  = np.fft.ifft(norm=forward)  (unnormalised e^{+})            gen_data.py:46
 *       R2C fwd  = np.fft.rfft               tests/ut/unittest/fft/1d/r2c/r2c_data/gen_data.py:42-44
 *       R2C inv  = np.fft.ihfft(norm=forward) = conj(rfft)                    gen_data.py:47-48
 *       C2R fwd  = np.fft.hfft(n=N)           tests/ut/unittest/fft/1d/c2r/c2r_data/gen_data.py:42-44
 *       C2R inv  = np.fft.irfft(n=N, norm=forward)                            gen_data.py:47-48
 *       2D/3D C2C = fftn over the trailing axes (inverse unnormalised)
 *         tests/ut/unittest/fft/2d/c2c/c2c_2d_data/gen_data_2d.py:64-67,
 *         tests/ut/unittest/fft/3d/c2c/c2c_3d_data/gen_data_3d.py:37-48
 *       2D/3D R2C = R2C on the last axis, then C2C on the leading FFT axes;
 *       2D/3D C2R = C2C on the leading FFT axes, then C2R on the last axis
 *         (step order of init2DSteps / init3DSteps, core/fft/fft_api.cpp:796-897, :938-992)
 *     numpy's golden comes from pocketfft (a third-party dependency, unpinned
 *     in requirements.txt:9); this file restates the DFT definition itself.
 *   - the unit-test input generator (Numerical-Recipes LCG, seed 2026,
 *     U[-300,300)): tests/ut/framework/test_util/float_util.h:23-36,
 *     tests/ut/framework/test_util/util.cpp:324-338.
 *   - the accuracy metric of the standalone harness (relative L2 vs an FP64
 *     golden): tests/standalone/fft1d_optimized/run.py:503-511.
 *
 * Two independent FP64 evaluators are provided so that each pins the other:
 *   oracle_dft_direct  O(N^2) direct sum in long double with exactly reduced
 *                      root indices ((n*k) mod N);
 *   oracle_fft         recursive mixed-radix decimation-in-time FFT in double
 *                      (any N: each prime factor p is combined with a p-point
 *                      direct DFT), roots computed in long double.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    double re, im;
} cd;

static const long double PI_L = 3.141592653589793238462643383279502884L;

/* e^{sign * 2*pi*i * m / n} in long double, m reduced mod n first. */
static void root_ld(long n, long m, int sign, long double* c, long double* s)
{
    m %= n;
    if (m < 0) m += n;
    long double a = 2.0L * PI_L * (long double)m / (long double)n;
    *c = cosl(a);
    *s = (long double)sign * sinl(a);
}

/* ------------------------------------------------------------------------- */
/* Direct DFT, long double accumulation.                                      */
/* ------------------------------------------------------------------------- */
void oracle_dft_direct(const double* in, double* out, long n, int sign)
{
    long double* wc = (long double*)malloc(sizeof(long double) * (size_t)n);
    long double* ws = (long double*)malloc(sizeof(long double) * (size_t)n);
    for (long m = 0; m < n; m++) root_ld(n, m, sign, &wc[m], &ws[m]);
    for (long k = 0; k < n; k++) {
        long double ar = 0.0L, ai = 0.0L;
        long idx = 0;
        for (long j = 0; j < n; j++) {
            long double xr = in[2 * j], xi = in[2 * j + 1];
            ar += xr * wc[idx] - xi * ws[idx];
            ai += xr * ws[idx] + xi * wc[idx];
            idx += k;
            if (idx >= n) idx -= n;
        }
        out[2 * k] = (double)ar;
        out[2 * k + 1] = (double)ai;
    }
    free(wc);
    free(ws);
}

/* ------------------------------------------------------------------------- */
/* Mixed-radix recursive DIT FFT in double.                                   */
/* ------------------------------------------------------------------------- */
static long smallest_factor(long n)
{
    if (n % 2 == 0) return 2;
    for (long p = 3; p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

/* tw[m] = e^{sign 2 pi i m / N}, m in [0, N) */
static void fft_rec(const cd* in, long istride, cd* out, long n, long tstride, const cd* tw, long N, cd* tmp)
{
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    long p = smallest_factor(n);
    long m = n / p;
    for (long r = 0; r < p; r++) fft_rec(in + r * istride, istride * p, out + r * m, m, tstride * p, tw, N, tmp);
    /* combine: X[k + q m] = sum_r W_n^{r (k + q m)} Y_r[k] */
    cd* t = tmp;           /* p entries */
    cd* res = tmp + p;     /* n entries */
    for (long k = 0; k < m; k++) {
        for (long r = 0; r < p; r++) {
            cd y = out[r * m + k];
            long e = ((r * k) % n) * tstride % N;
            cd w = tw[e];
            t[r].re = y.re * w.re - y.im * w.im;
            t[r].im = y.re * w.im + y.im * w.re;
        }
        for (long q = 0; q < p; q++) {
            double ar = 0.0, ai = 0.0;
            long step = (q * (N / p)) % N; /* W_p^{q} = tw[q N / p] */
            long e = 0;
            for (long r = 0; r < p; r++) {
                cd w = tw[e];
                ar += t[r].re * w.re - t[r].im * w.im;
                ai += t[r].re * w.im + t[r].im * w.re;
                e += step;
                if (e >= N) e -= N;
            }
            res[k + q * m].re = ar;
            res[k + q * m].im = ai;
        }
    }
    memcpy(out, res, sizeof(cd) * (size_t)n);
}

/* in/out: n interleaved complex doubles; may not alias. */
void oracle_fft(const double* in, double* out, long n, int sign)
{
    cd* tw = (cd*)malloc(sizeof(cd) * (size_t)n);
    for (long m = 0; m < n; m++) {
        long double c, s;
        root_ld(n, m, sign, &c, &s);
        tw[m].re = (double)c;
        tw[m].im = (double)s;
    }
    /* scratch: the deepest recursion level needs p + n entries; the largest p is n itself */
    cd* tmp = (cd*)malloc(sizeof(cd) * (size_t)(2 * n + 2));
    fft_rec((const cd*)in, 1, (cd*)out, n, 1, tw, n, tmp);
    free(tmp);
    free(tw);
}

/* ------------------------------------------------------------------------- */
/* Strided 1D complex transform along one axis of a row-major array.          */
/*   data: complex doubles, logical shape [outer][n][inner]; in place.        */
/* ------------------------------------------------------------------------- */
static void c2c_axis(double* data, long outer, long n, long inner, int sign)
{
    double* a = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    double* b = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    for (long o = 0; o < outer; o++) {
        for (long c = 0; c < inner; c++) {
            double* base = data + 2 * (o * n * inner + c);
            for (long i = 0; i < n; i++) {
                a[2 * i] = base[2 * i * inner];
                a[2 * i + 1] = base[2 * i * inner + 1];
            }
            oracle_fft(a, b, n, sign);
            for (long i = 0; i < n; i++) {
                base[2 * i * inner] = b[2 * i];
                base[2 * i * inner + 1] = b[2 * i + 1];
            }
        }
    }
    free(a);
    free(b);
}

static int sign_of(int inverse) { return inverse ? +1 : -1; }

/* ------------------------------------------------------------------------- */
/* Public semantics (float32 / complex64 in, FP64 out).                       */
/*   dims[0..ndim-1]: FFT sizes (X[,Y[,Z]]), the last one is the contiguous   */
/*   axis; batch leading.  Complex arrays are interleaved (re, im).          */
/* ------------------------------------------------------------------------- */

/* C2C: in complex64 [batch, dims...] -> out complex128 [batch, dims...] */
void oracle_c2c_nd(const float* in, double* out, long batch, const long* dims, int ndim, int inverse)
{
    long total = batch;
    for (int d = 0; d < ndim; d++) total *= dims[d];
    for (long i = 0; i < 2 * total; i++) out[i] = (double)in[i];
    /* apply along every FFT axis, last axis first (order does not change the exact result) */
    for (int d = ndim - 1; d >= 0; d--) {
        long inner = 1, outer = batch;
        for (int e = d + 1; e < ndim; e++) inner *= dims[e];
        for (int e = 0; e < d; e++) outer *= dims[e];
        c2c_axis(out, outer, dims[d], inner, sign_of(inverse));
    }
}

/* 1D real->half-complex on rows of length n: in real (double) [rows][n] -> out [rows][n/2+1] */
static void r2c_rows(const double* in, double* out, long rows, long n, int inverse)
{
    long h = n / 2 + 1;
    double* a = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    double* b = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    for (long r = 0; r < rows; r++) {
        for (long i = 0; i < n; i++) {
            a[2 * i] = in[r * n + i];
            a[2 * i + 1] = 0.0;
        }
        oracle_fft(a, b, n, -1);
        for (long k = 0; k < h; k++) {
            out[2 * (r * h + k)] = b[2 * k];
            /* R2C INVERSE = sum x e^{+} = conj(rfft(x)) for real x */
            out[2 * (r * h + k) + 1] = inverse ? -b[2 * k + 1] : b[2 * k + 1];
        }
    }
    free(a);
    free(b);
}

/* 1D half-complex->real on rows: in complex (double) [rows][n/2+1] -> out real [rows][n].
 * The Hermitian extension drops the imaginary part of bin 0 and, for even n,
 * of bin n/2 (numpy irfft behaviour).  FORWARD (hfft) conjugates the input
 * first; both then sum with e^{+}, without normalisation. */
static void c2r_rows(const double* in, double* out, long rows, long n, int inverse)
{
    long h = n / 2 + 1;
    double* a = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    double* b = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    double cs = inverse ? 1.0 : -1.0;
    for (long r = 0; r < rows; r++) {
        const double* x = in + 2 * r * h;
        a[0] = x[0];
        a[1] = 0.0;
        for (long k = 1; k < h; k++) {
            if (2 * k == n) {
                a[2 * k] = x[2 * k];
                a[2 * k + 1] = 0.0;
            } else {
                a[2 * k] = x[2 * k];
                a[2 * k + 1] = cs * x[2 * k + 1];
                a[2 * (n - k)] = x[2 * k];
                a[2 * (n - k) + 1] = -cs * x[2 * k + 1];
            }
        }
        oracle_fft(a, b, n, +1);
        for (long i = 0; i < n; i++) out[r * n + i] = b[2 * i];
    }
    free(a);
    free(b);
}

/* R2C: in float32 [batch, dims...] -> out complex128 [batch, dims[:-1]..., last/2+1] */
void oracle_r2c_nd(const float* in, double* out, long batch, const long* dims, int ndim, int inverse)
{
    long n = dims[ndim - 1], h = n / 2 + 1;
    long rows = batch;
    for (int d = 0; d < ndim - 1; d++) rows *= dims[d];
    double* tmp = (double*)malloc(sizeof(double) * (size_t)(rows * n));
    for (long i = 0; i < rows * n; i++) tmp[i] = (double)in[i];
    r2c_rows(tmp, out, rows, n, inverse);
    free(tmp);
    for (int d = ndim - 2; d >= 0; d--) {
        long inner = h, outer = batch;
        for (int e = d + 1; e < ndim - 1; e++) inner *= dims[e];
        for (int e = 0; e < d; e++) outer *= dims[e];
        c2c_axis(out, outer, dims[d], inner, sign_of(inverse));
    }
}

/* C2R: in complex64 [batch, dims[:-1]..., last/2+1] -> out float64 [batch, dims...] */
void oracle_c2r_nd(const float* in, double* out, long batch, const long* dims, int ndim, int inverse)
{
    long n = dims[ndim - 1], h = n / 2 + 1;
    long rows = batch;
    for (int d = 0; d < ndim - 1; d++) rows *= dims[d];
    double* tmp = (double*)malloc(sizeof(double) * 2 * (size_t)(rows * h));
    for (long i = 0; i < 2 * rows * h; i++) tmp[i] = (double)in[i];
    for (int d = ndim - 2; d >= 0; d--) {
        long inner = h, outer = batch;
        for (int e = d + 1; e < ndim - 1; e++) inner *= dims[e];
        for (int e = 0; e < d; e++) outer *= dims[e];
        c2c_axis(tmp, outer, dims[d], inner, sign_of(inverse));
    }
    c2r_rows(tmp, out, rows, n, inverse);
    free(tmp);
}

/* ------------------------------------------------------------------------- */
/* Unit-test input generator (tests/ut/framework/test_util/float_util.h:23-36) */
/* ------------------------------------------------------------------------- */
void oracle_lcg_uniform(uint32_t* state, float lower, float upper, float* out, long n)
{
    const float inv_u32 = 1.0f / 4294967296.0f;
    uint32_t s = *state;
    for (long i = 0; i < n; i++) {
        s = s * 1664525U + 1013904223U;
        out[i] = lower + ((float)s * inv_u32) * (upper - lower);
    }
    *state = s;
}

/* ------------------------------------------------------------------------- */
/* Relative L2 error of a float result against an FP64 golden                 */
/* (tests/standalone/fft1d_optimized/run.py:503-506: ||out-g|| / ||g||,       */
/*  or ||out-g|| when ||g|| == 0).  count = number of doubles.               */
/* ------------------------------------------------------------------------- */
double oracle_rel_l2(const float* got, const double* golden, long count)
{
    long double num = 0.0L, den = 0.0L;
    for (long i = 0; i < count; i++) {
        long double d = (long double)got[i] - (long double)golden[i];
        num += d * d;
        den += (long double)golden[i] * (long double)golden[i];
    }
    if (den == 0.0L) return (double)sqrtl(num);
    return (double)sqrtl(num / den);
}

/* ------------------------------------------------------------------------- */
/* Batched 2D circular shift (include/fft_api.h:97-122 contract of          */
/* asdFftShift / asdIfftShift; goldens in                                   */
/* tests/ut/unittest/fft/{fftshift,ifftshift}/<case>_data/gen_data.py):           */
/*   out[s][(r + sr) % row][(c + sc) % col] = scale * in[s*dist + r*col + c] */
/* with (sr, sc) = (ceil(row/2), ceil(col/2)) for the forward shift and     */
/* (floor(row/2), floor(col/2)) for the inverse; the scale is applied in    */
/* float32 (each of re, im multiplied by (float)scale) only when apply != 0.*/
/* in/out hold interleaved (re, im) float pairs.                            */
/* ------------------------------------------------------------------------- */
void oracle_shift2d(const float* in, float* out, long row, long col, long batch, long dist, int inverse,
                    double scale, int apply)
{
    const long sr = inverse ? row / 2 : (row + 1) / 2;
    const long sc = inverse ? col / 2 : (col + 1) / 2;
    const float f = (float)scale;
    for (long s = 0; s < batch; s++) {
        for (long r = 0; r < row; r++) {
            for (long c = 0; c < col; c++) {
                const float* src = in + 2 * (s * dist + r * col + c);
                float* dst = out + 2 * (s * row * col + ((r + sr) % row) * col + (c + sc) % col);
                dst[0] = apply ? src[0] * f : src[0];
                dst[1] = apply ? src[1] * f : src[1];
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* ISTFT (asdFftIstftMakePlan / asdFftExecIstft; the only supported form:    */
/* center = true, normalized = false, onesided = false, length = 0,         */
/* returnComplex = true), following core/fft/istft.cpp:288-433 and          */
/* core/fft/fftcore/fft_core_istft_any.cpp (Run: 1/nfft, inverse DFT of each */
/* frame, multiply by the window, overlap-add at stride hop, divide by the  */
/* overlap-added squared window, keep [nfft/2, nfft/2 + out_len)), the      */
/* golden of tests/ut/unittest/fft/istft/istft_data/gen_data_istft.py:37-64.*/
/*   spec: complex64 [channel][nfft][nframes]   window: float32 [nfft]      */
/*   out:  complex128 [channel][out_len], out_len = nfft + hop*(nframes-1)  */
/*         - 2*(nfft/2)                                                      */
/* ------------------------------------------------------------------------- */
void oracle_istft(const float* spec, const float* window, double* out, long channel, long nfft, long nframes,
                  long hop)
{
    const long full = nfft + hop * (nframes - 1);
    const long start = nfft / 2;
    const long out_len = full - 2 * (nfft / 2);
    double* col = (double*)malloc(sizeof(double) * 2 * nfft);
    double* frame = (double*)malloc(sizeof(double) * 2 * nfft);
    double* y = (double*)malloc(sizeof(double) * 2 * full);
    double* env = (double*)malloc(sizeof(double) * full);
    for (long m = 0; m < full; m++) env[m] = 0.0;
    for (long t = 0; t < nframes; t++) {
        for (long n = 0; n < nfft; n++) env[t * hop + n] += (double)window[n] * (double)window[n];
    }
    for (long b = 0; b < channel; b++) {
        for (long m = 0; m < 2 * full; m++) y[m] = 0.0;
        for (long t = 0; t < nframes; t++) {
            for (long k = 0; k < nfft; k++) {  /* frequency column t of channel b */
                const float* src = spec + 2 * ((b * nfft + k) * nframes + t);
                col[2 * k] = src[0];
                col[2 * k + 1] = src[1];
            }
            oracle_fft(col, frame, nfft, +1);
            for (long n = 0; n < nfft; n++) {
                const double w = (double)window[n] / (double)nfft;
                y[2 * (t * hop + n)] += frame[2 * n] * w;
                y[2 * (t * hop + n) + 1] += frame[2 * n + 1] * w;
            }
        }
        for (long i = 0; i < out_len; i++) {
            const double e = env[start + i];
            out[2 * (b * out_len + i)] = y[2 * (start + i)] / e;
            out[2 * (b * out_len + i) + 1] = y[2 * (start + i) + 1] / e;
        }
    }
    free(col);
    free(frame);
    free(y);
    free(env);
}
