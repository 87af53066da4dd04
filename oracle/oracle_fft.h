/*
 * FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * This directory is the checker for the HIP library in
 * distributednavierstokes_amd/: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Nothing in the product path
 * links, imports or calls it.
 *
 * Complex FFT used by the oracle in place of FFTW (the reference's CPU FFT,
 * reached through AbstractFFTs plan_rfft/plan_fft!/plan_brfft at
 * src/poisson.jl:191-194 and src/spectral/setup.jl:97-98,148-151).
 * Algorithm: recursive decimation-in-time over the smallest prime factor,
 * with a direct DFT for each prime factor; any length n >= 1.
 * Pinned against numpy.fft (pocketfft) in tests/test_oracle_fft.py.
 */
#ifndef DNS_ORACLE_FFT_H
#define DNS_ORACLE_FFT_H
#include <complex.h>

/* Unnormalised DFT of the n values x[0], x[s], ..., x[(n-1)s], in place.
 * sign = -1: forward exp(-2 pi i jk/n); sign = +1: backward. */
void offt_strided(double complex *x, int n, long s, int sign);

/* Real-to-complex forward transform of a real line (n values, stride 1)
 * into n/2+1 complex values (unnormalised, like FFTW r2c). */
void offt_r2c(const double *x, double complex *X, int n);

/* Complex-to-real backward transform (unnormalised, like FFTW c2r / brfft):
 * X holds n/2+1 values of a Hermitian spectrum; imaginary parts of X[0] and
 * (even n) X[n/2] are ignored. */
void offt_c2r(const double complex *X, double *x, int n);

#endif
