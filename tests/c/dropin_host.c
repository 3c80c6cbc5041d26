/* dropin_host.c -- a C caller of the fftset API that passes host arrays, as
 * the reference's callers do, and checks every result against FP64 direct
 * sums: FREQOFFSETREAL and COMPLEX forward transforms, inverse(forward) =
 * L * identity, and the convolutions (negacyclic for FREQOFFSETREAL, cyclic
 * for COMPLEX). Exit status 0 when every check passes. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fftset.h"

static const double PI = 3.14159265358979323846;

static double rel_rms(const float *y, const double *ref, unsigned n)
{
	double e = 0.0, r = 0.0;
	for (unsigned i = 0; i < n; i++) {
		e += (y[i] - ref[i]) * (y[i] - ref[i]);
		r += ref[i] * ref[i];
	}
	return sqrt(e / (r > 0.0 ? r : 1.0));
}

static float urand(unsigned *s)
{
	*s = *s * 1664525u + 1013904223u;
	return (float)((*s >> 8) & 0xffffff) / 8388608.0f - 1.0f;
}

/* odd-frequency DFT of N = 2L reals, natural order (the library's bin
 * order): X[q] = sum_n x[n] e^{-i pi (2q+1) n / N}, q < L */
static void fo_dft(const float *x, double *out, unsigned L)
{
	const unsigned N = 2 * L;
	for (unsigned q = 0; q < L; q++) {
		double re = 0.0, im = 0.0;
		for (unsigned n = 0; n < N; n++) {
			const double a = -PI * (double)((2ull * q + 1) * n % (2ull * N)) / N;
			re += x[n] * cos(a);
			im += x[n] * sin(a);
		}
		out[2 * q] = re;
		out[2 * q + 1] = im;
	}
}

static void cx_dft(const float *x, double *out, unsigned L)
{
	for (unsigned k = 0; k < L; k++) {
		double re = 0.0, im = 0.0;
		for (unsigned n = 0; n < L; n++) {
			const double a = -2.0 * PI * (double)((unsigned long long)k * n % L) / L;
			re += x[2 * n] * cos(a) - x[2 * n + 1] * sin(a);
			im += x[2 * n] * sin(a) + x[2 * n + 1] * cos(a);
		}
		out[2 * k] = re;
		out[2 * k + 1] = im;
	}
}

int main(void)
{
	static const unsigned sizes[] = {2, 3, 5, 12, 45, 64, 120, 256, 360, 512};
	struct fftset fc;
	int failures = 0;
	if (fftset_init(&fc)) {
		printf("fftset_init failed\n");
		return 2;
	}
	for (unsigned si = 0; si < sizeof(sizes) / sizeof(sizes[0]); si++) {
		const unsigned L = sizes[si], N = 2 * L;
		const double tol = 1e-6 * (log2((double)L) + 1.0);
		float *x = malloc(N * sizeof(float)), *h = malloc(N * sizeof(float));
		float *y = malloc(N * sizeof(float)), *z = malloc(N * sizeof(float));
		float *k = malloc(N * sizeof(float)), *w = malloc(N * sizeof(float));
		double *ref = malloc(N * sizeof(double));
		unsigned seed = 12345u + L;
		for (unsigned i = 0; i < N; i++) {
			x[i] = urand(&seed);
			h[i] = urand(&seed) * (float)exp(-(double)i / (L / 2.0 + 1.0));
		}
		for (int m = 0; m < 2; m++) {
			const struct fftset_modulation *mod = m ? FFTSET_MODULATION_COMPLEX : FFTSET_MODULATION_FREQ_OFFSET_REAL;
			const struct fftset_fft *f = fftset_create_fft(&fc, mod, L);
			if (!f || f != fftset_create_fft(&fc, mod, L)) {
				printf("L=%u %s: create failed\n", L, m ? "COMPLEX" : "FREQOFFSETREAL");
				failures++;
				continue;
			}
			double e1, e2, e3;
			/* forward */
			fftset_fft_forward(f, y, x, w);
			if (m)
				cx_dft(x, ref, L);
			else
				fo_dft(x, ref, L);
			e1 = rel_rms(y, ref, N);
			/* inverse(forward(x)) = L x, in place */
			memcpy(z, y, N * sizeof(float));
			fftset_fft_inverse(f, z, z, w);
			for (unsigned i = 0; i < N; i++) ref[i] = (double)L * x[i];
			e2 = rel_rms(z, ref, N);
			/* convolution: L * negacyclic_N (FREQOFFSETREAL) or L * cyclic_L (COMPLEX) */
			fftset_fft_conv_get_kernel(f, k, h);
			fftset_fft_conv(f, y, x, k, w);
			if (m) {
				for (unsigned n = 0; n < L; n++) {
					double re = 0.0, im = 0.0;
					for (unsigned j = 0; j < L; j++) {
						const unsigned q = (n + L - j) % L;
						re += x[2 * j] * h[2 * q] - x[2 * j + 1] * h[2 * q + 1];
						im += x[2 * j] * h[2 * q + 1] + x[2 * j + 1] * h[2 * q];
					}
					ref[2 * n] = L * re;
					ref[2 * n + 1] = L * im;
				}
			} else {
				for (unsigned n = 0; n < N; n++) {
					double acc = 0.0;
					for (unsigned j = 0; j < N; j++)
						acc += (j <= n) ? x[j] * h[n - j] : -x[j] * h[N + n - j];
					ref[n] = L * acc;
				}
			}
			e3 = rel_rms(y, ref, N);
			const int ok = e1 <= tol && e2 <= tol && e3 <= tol && fftset_last_error() == FFTSET_OK;
			printf("L=%-4u %-15s forward %.2e  inverse %.2e  conv %.2e  %s\n", L, m ? "COMPLEX" : "FREQOFFSETREAL",
			       e1, e2, e3, ok ? "ok" : "FAIL");
			failures += !ok;
		}
		free(x); free(h); free(y); free(z); free(k); free(w); free(ref);
	}
	if (fftset_recommend_conv_length(1024, 0) != 8192) {
		printf("recommend_conv_length mismatch\n");
		failures++;
	}
	fftset_destroy(&fc);
	printf("%s (%d failures)\n", failures ? "FAILED" : "all ok", failures);
	return failures ? 1 : 0;
}
