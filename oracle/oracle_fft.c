/* FP64 CPU oracle FFT -- TEST INFRASTRUCTURE ONLY (see oracle_fft.h). */
#include "oracle_fft.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static int smallest_prime_factor(int n)
{
    for (int p = 2; (long)p * p <= n; ++p)
        if (n % p == 0) return p;
    return n;
}

/* out[0..n-1] = DFT_n of in[0], in[is], ...; tw[t] = exp(sign 2 pi i t / N),
 * and the n-th roots are tw[t * (N / n)]. */
static void dft_rec(const double complex *in, long is, double complex *out, int n,
                    const double complex *tw, int N, double complex *y)
{
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    int p = smallest_prime_factor(n);
    int m = n / p;
    for (int r = 0; r < p; ++r) dft_rec(in + r * is, is * p, out + (long)r * m, m, tw, N, y);
    long step = N / n;
    for (int k = 0; k < m; ++k) {
        for (int r = 0; r < p; ++r) y[r] = out[(long)r * m + k];
        for (int q = 0; q < p; ++q) {
            double complex acc = 0;
            long kq = k + (long)q * m;
            for (int r = 0; r < p; ++r) acc += y[r] * tw[((r * kq) % n) * step];
            out[kq] = acc;
        }
    }
}

static void twiddles(double complex *tw, int N, int sign)
{
    const long double twopi = 6.283185307179586476925286766559005768L;
    for (int t = 0; t < N; ++t) {
        long double a = twopi * (long double)t / (long double)N;
        tw[t] = (double)cosl(a) + I * (double)(sign * sinl(a));
    }
}

void offt_strided(double complex *x, int n, long s, int sign)
{
    if (n <= 1) return;
    double complex *buf = malloc(sizeof(double complex) * (3 * (size_t)n));
    double complex *in = buf, *out = buf + n, *tw = buf + 2 * (size_t)n;
    double complex *y = malloc(sizeof(double complex) * (size_t)n);
    twiddles(tw, n, sign);
    for (int t = 0; t < n; ++t) in[t] = x[t * s];
    dft_rec(in, 1, out, n, tw, n, y);
    for (int t = 0; t < n; ++t) x[t * s] = out[t];
    free(y);
    free(buf);
}

void offt_r2c(const double *x, double complex *X, int n)
{
    double complex *a = malloc(sizeof(double complex) * (size_t)n);
    for (int t = 0; t < n; ++t) a[t] = x[t];
    offt_strided(a, n, 1, -1);
    for (int k = 0; k <= n / 2; ++k) X[k] = a[k];
    free(a);
}

void offt_c2r(const double complex *X, double *x, int n)
{
    double complex *a = malloc(sizeof(double complex) * (size_t)n);
    int h = n / 2;
    for (int k = 0; k <= h; ++k) a[k] = X[k];
    a[0] = creal(a[0]);
    if (n % 2 == 0) a[h] = creal(a[h]);
    for (int k = h + 1; k < n; ++k) a[k] = conj(X[n - k]);
    offt_strided(a, n, 1, +1);
    for (int t = 0; t < n; ++t) x[t] = creal(a[t]);
    free(a);
}
