/* fftset_oracle.c -- CPU restatement of fftset's scalar (v1f) transform and
 * fast-convolution path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker the GPU library is
 * compared against; it is never linked into, loaded by or called from the
 * product library (fftset_amd/). Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it.
 *
 * What it restates (reference paths are relative to the fftset source tree):
 *   - the min-cost radix planner      src/fftset_vec.c:944-1062
 *   - per-pass FP32 twiddle tables     src/fftset_vec.c:1064-1120
 *   - Stockham (self-sorting) passes   src/fftset_vec.c:54-63, 90-106, 1222-1249
 *   - in-place DIF / mulconj / DIT     src/fftset_vec.c:46-53, 66-89, 847-867, 1161-1220
 *   - FREQOFFSETREAL scalar modulator  src/fftset_mod_freqoffsetreal.c:884-1010
 *   - COMPLEX scalar modulator         src/fftset_mod_cplx.c:242-318
 *   - conv-length recommendation       src/fftset.c:131-167
 * The butterflies are this file's own small DFTs (the same transforms with
 * their own operation order), so results agree with the reference to FP32
 * rounding, not bit for bit.
 *
 * Known reference quirk NOT reproduced: the scalar FREQOFFSETREAL inverse with
 * an odd bin count reads the middle bin from the output buffer
 * (src/fftset_mod_freqoffsetreal.c:995-998); this restatement reads it from the
 * input buffer, which is what the transform means.
 *
 * Also exported: a pthread batch runner (one transform per thread) that is
 * the CPU baseline timed next to the GPU path. */

#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define OR_MOD_FREQ_OFFSET_REAL 0
#define OR_MOD_COMPLEX          1

#define OR_OP_FORWARD 0
#define OR_OP_INVERSE 1
#define OR_OP_GETKERN 2
#define OR_OP_CONV    3

#define OR_MAX_PASSES 40

/* Mutation hooks for the oracle's own tests: building with -DOR_PLANT=n
 * plants one deliberate mistake so a test can show its checks catch it.
 * The default build plants nothing. */
#ifndef OR_PLANT
#define OR_PLANT 0
#endif
#define OR_PLANT_IS(n) (OR_PLANT == (n))

struct or_pass {
	unsigned radix;
	unsigned m;      /* sub-transform length after this pass (lfft_div_radix) */
	float   *tw;     /* m*(radix-1) complex twiddles [j][k-1], NULL for the final pass */
};

struct or_plan {
	unsigned       L;
	int            mod;
	unsigned       np;
	struct or_pass p[OR_MAX_PASSES];
};

/* ------------------------------------------------------------------------ */
/* Planner: min-cost chain of radices for the scalar (vector width 1) passes.
 * Follows build_float_graph (src/fftset_vec.c:1004-1062) over the scalar rows
 * of FFTSET_FLOAT_PASSES (src/fftset_vec.c:944-971): radices 2, 3, 4 may be
 * used at any position, 5, 6, 8, 16 only as the final pass; each pass costs
 * 100 + ((r+2)(vw+3)*10^4)/(r*vw) with vw = 1 (src/fftset_vec.c:1023); a
 * candidate replaces the best so far only when strictly cheaper, scanning
 * radices in table order. The result for a length depends only on that
 * length, so it is memoised. */

static const unsigned OR_RADICES[7] = {2, 3, 4, 5, 6, 8, 16};

struct or_chain {
	unsigned length;
	unsigned cost;
	unsigned n;
	unsigned r[OR_MAX_PASSES];
};

static struct or_chain *g_memo = NULL;
static unsigned g_memo_n = 0, g_memo_cap = 0;
static pthread_mutex_t g_memo_lock = PTHREAD_MUTEX_INITIALIZER;

static unsigned or_radix_cost(unsigned r)
{
	const unsigned vw = 1;
	return 100u + ((r + 2u) * (vw + 3u) * 10000u) / (r * vw);
}

static int or_memo_find(unsigned length, struct or_chain *out)
{
	unsigned i;
	for (i = 0; i < g_memo_n; i++)
		if (g_memo[i].length == length) {
			*out = g_memo[i];
			return 1;
		}
	return 0;
}

static void or_memo_add(const struct or_chain *c)
{
	if (g_memo_n == g_memo_cap) {
		unsigned nc = g_memo_cap ? 2 * g_memo_cap : 64;
		struct or_chain *nm = realloc(g_memo, nc * sizeof(*nm));
		if (!nm)
			return;
		g_memo = nm;
		g_memo_cap = nc;
	}
	g_memo[g_memo_n++] = *c;
}

/* Returns 1 and fills *best when `length` factors into a valid chain. */
static int or_search(unsigned length, struct or_chain *best)
{
	unsigned i;
	if (or_memo_find(length, best))
		return best->cost != 0;
	best->length = length;
	best->cost = 0;
	best->n = 0;
	for (i = 0; i < 7; i++) {
		unsigned r = OR_RADICES[i];
		unsigned c = or_radix_cost(r);
		struct or_chain sub;
		if (length == r) {
			if (best->cost == 0 || best->cost > c) {
				best->cost = c;
				best->n = 1;
				best->r[0] = r;
			}
			continue;
		}
		if (r > 4 || length % r)
			continue;
		if (!or_search(length / r, &sub))
			continue;
		if (sub.n + 1 > OR_MAX_PASSES)
			continue;
		if (best->cost == 0 || best->cost > sub.cost + c) {
			best->cost = sub.cost + c;
			best->n = sub.n + 1;
			best->r[0] = r;
			memcpy(best->r + 1, sub.r, sub.n * sizeof(unsigned));
		}
	}
	or_memo_add(best);
	return best->cost != 0;
}

/* Writes the planned radix chain for `length` into radices[]; returns the
 * number of passes, or 0 when the length is unsupported. */
unsigned oracle_plan_chain(unsigned length, unsigned *radices, unsigned max)
{
	struct or_chain c;
	int ok;
	if (length < 2)
		return 0;
	pthread_mutex_lock(&g_memo_lock);
	ok = or_search(length, &c);
	pthread_mutex_unlock(&g_memo_lock);
	if (!ok || c.n > max)
		return 0;
	memcpy(radices, c.r, c.n * sizeof(unsigned));
	return c.n;
}

/* ------------------------------------------------------------------------ */
/* Small forward DFTs on interleaved complex floats (x[2q], x[2q+1]). */

static void or_dft2(float *x)
{
	float ar = x[0], ai = x[1], br = x[2], bi = x[3];
	x[0] = ar + br; x[1] = ai + bi;
	x[2] = ar - br; x[3] = ai - bi;
}

static void or_dft3(float *x)
{
	const float s3 = 0.86602540378443865f;
	float tr = x[2] + x[4], ti = x[3] + x[5];
	float dr = x[2] - x[4], di = x[3] - x[5];
	float mr = x[0] - 0.5f * tr, mi = x[1] - 0.5f * ti;
	float sr = s3 * dr, si = s3 * di;
	x[0] = x[0] + tr; x[1] = x[1] + ti;
	/* y1 = m - j*s, y2 = m + j*s */
	x[2] = mr + si; x[3] = mi - sr;
	x[4] = mr - si; x[5] = mi + sr;
}

static void or_dft4(float *x)
{
	float ar = x[0] + x[4], ai = x[1] + x[5];
	float br = x[0] - x[4], bi = x[1] - x[5];
	float cr = x[2] + x[6], ci = x[3] + x[7];
	float dr = x[2] - x[6], di = x[3] - x[7];
	x[0] = ar + cr; x[1] = ai + ci;
	x[4] = ar - cr; x[5] = ai - ci;
	/* y1 = b - j*d, y3 = b + j*d */
	x[2] = br + di; x[3] = bi - dr;
	x[6] = br - di; x[7] = bi + dr;
}

static void or_dft5(float *x)
{
	const float c1 = 0.30901699437494742f;   /* cos(2pi/5) */
	const float c2 = -0.80901699437494742f;  /* cos(4pi/5) */
	const float s1 = 0.95105651629515357f;   /* sin(2pi/5) */
	const float s2 = 0.58778525229247313f;   /* sin(4pi/5) */
	float t1r = x[2] + x[8], t1i = x[3] + x[9];
	float t2r = x[4] + x[6], t2i = x[5] + x[7];
	float t3r = x[2] - x[8], t3i = x[3] - x[9];
	float t4r = x[4] - x[6], t4i = x[5] - x[7];
	float a1r = x[0] + c1 * t1r + (OR_PLANT_IS(6) ? 0.0f : c2 * t2r), a1i = x[1] + c1 * t1i + c2 * t2i;
	float a2r = x[0] + c2 * t1r + c1 * t2r, a2i = x[1] + c2 * t1i + c1 * t2i;
	float b1r = s1 * t3r + s2 * t4r, b1i = s1 * t3i + s2 * t4i;
	float b2r = s2 * t3r - s1 * t4r, b2i = s2 * t3i - s1 * t4i;
	x[0] = x[0] + t1r + t2r; x[1] = x[1] + t1i + t2i;
	/* y1 = a1 - j*b1, y4 = a1 + j*b1, y2 = a2 - j*b2, y3 = a2 + j*b2 */
	x[2] = a1r + b1i; x[3] = a1i - b1r;
	x[8] = a1r - b1i; x[9] = a1i + b1r;
	x[4] = a2r + b2i; x[5] = a2i - b2r;
	x[6] = a2r - b2i; x[7] = a2i + b2r;
}

static void or_dft_small(unsigned r, float *x);

/* Composite size r = a*b by one Cooley-Tukey step held in a local array:
 * X[k1 + a*k2] = sum_n2 W_b^{n2 k2} W_r^{n2 k1} sum_n1 x[b n1 + n2] W_a^{n1 k1}. */
static void or_dft_ct(unsigned a, unsigned b, float *x)
{
	float col[2 * 16], tmp[2 * 16], row[2 * 16];
	unsigned r = a * b, n1, n2, k1, k2;
	for (n2 = 0; n2 < b; n2++) {
		for (n1 = 0; n1 < a; n1++) {
			col[2 * n1]     = x[2 * (b * n1 + n2)];
			col[2 * n1 + 1] = x[2 * (b * n1 + n2) + 1];
		}
		or_dft_small(a, col);
		for (k1 = 0; k1 < a; k1++) {
			double ang = -2.0 * M_PI * (double)((n2 * k1) % r) / (double)r;
			float wr = (float)cos(ang), wi = (float)sin(ang);
			float vr = col[2 * k1], vi = col[2 * k1 + 1];
			tmp[2 * (k1 * b + n2)]     = vr * wr - vi * wi;
			tmp[2 * (k1 * b + n2) + 1] = vr * wi + vi * wr;
		}
	}
	for (k1 = 0; k1 < a; k1++) {
		memcpy(row, tmp + 2 * k1 * b, 2 * b * sizeof(float));
		or_dft_small(b, row);
		for (k2 = 0; k2 < b; k2++) {
			x[2 * (k1 + a * k2)]     = row[2 * k2];
			x[2 * (k1 + a * k2) + 1] = row[2 * k2 + 1];
		}
	}
}

static void or_dft_small(unsigned r, float *x)
{
	switch (r) {
	case 1: break;
	case 2: or_dft2(x); break;
	case 3: or_dft3(x); break;
	case 4: or_dft4(x); break;
	case 5: or_dft5(x); break;
	case 6: or_dft_ct(2, 3, x); break;
	case 8: or_dft_ct(2, 4, x); break;
	case 16: or_dft_ct(4, 4, x); break;
	default: abort();
	}
}

static inline void or_cmul(float *v, const float *w)
{
	float r = v[0] * w[0] - v[1] * w[1];
	float i = v[0] * w[1] + v[1] * w[0];
	v[0] = r;
	v[1] = i;
}

/* ------------------------------------------------------------------------ */
/* Passes. Twiddles for a pass of sub-length len = r*m are
 * W_len^{jk} = cosf/sinf(j*k*(-2pi)/len) with the argument formed in float,
 * as src/fftset_vec.c:1103-1109 forms it. */

static float *or_make_twiddles(unsigned r, unsigned m)
{
	unsigned len = r * m, j, k;
	float *tw = malloc(sizeof(float) * 2 * (r - 1) * m);
	float *p = tw;
	if (!tw)
		return NULL;
	for (j = 0; j < m; j++)
		for (k = 1; k < r; k++) {
			float arg = (float)(j * k) * (-(float)M_PI * 2) / (float)len;
			*p++ = cosf(arg);
			*p++ = OR_PLANT_IS(1) ? -sinf(arg) : sinf(arg);
		}
	return tw;
}

/* One Stockham DIF pass over s sub-transforms of length r*m
 * (src/fftset_vec.c:90-106; the final pass with m == 1 is :54-63):
 * out[b*m + j + k*m*s] = W^{jk} * DFT_r(in[b*r*m + j + q*m])[k]. */
static void or_stock_pass(const struct or_pass *p, unsigned s, const float *in, float *out)
{
	unsigned r = p->radix, m = p->m, b, j, q, k;
	float v[2 * 16];
	for (b = 0; b < s; b++)
		for (j = 0; j < m; j++) {
			for (q = 0; q < r; q++) {
				v[2 * q]     = in[2 * (b * r * m + j + q * m)];
				v[2 * q + 1] = in[2 * (b * r * m + j + q * m) + 1];
			}
			or_dft_small(r, v);
			for (k = 0; k < r; k++) {
				if (k && p->tw)
					or_cmul(v + 2 * k, p->tw + 2 * (j * (r - 1) + k - 1));
				unsigned o = OR_PLANT_IS(5) ? (b * m + j + k * m * s + (s > 1 ? 1 : 0)) % (m * s * r)
				                            : b * m + j + k * m * s;
				out[2 * o]     = v[2 * k];
				out[2 * o + 1] = v[2 * k + 1];
			}
		}
}

/* In-place DIF pass over nfft blocks (src/fftset_vec.c:66-77, final pass
 * :46-53): positions b*r*m + j + q*m, twiddle applied after the DFT. */
static void or_dif_pass(const struct or_pass *p, unsigned nfft, float *w)
{
	unsigned r = p->radix, m = p->m, b, j, q;
	float v[2 * 16];
	for (b = 0; b < nfft; b++)
		for (j = 0; j < m; j++) {
			float *base = w + 2 * (b * r * m + j);
			for (q = 0; q < r; q++) {
				v[2 * q]     = base[2 * q * m];
				v[2 * q + 1] = base[2 * q * m + 1];
			}
			or_dft_small(r, v);
			for (q = 0; q < r; q++) {
				if (q && p->tw)
					or_cmul(v + 2 * q, p->tw + 2 * (j * (r - 1) + q - 1));
				base[2 * q * m]     = v[2 * q];
				base[2 * q * m + 1] = v[2 * q + 1];
			}
		}
}

/* In-place DIT pass (src/fftset_vec.c:78-89): the same forward twiddles are
 * applied before a forward DFT, so a DIF chain followed by the reversed DIT
 * chain on conjugated data computes a conjugated inverse. */
static void or_dit_pass(const struct or_pass *p, unsigned nfft, float *w)
{
	unsigned r = p->radix, m = p->m, b, j, q;
	float v[2 * 16];
	for (b = 0; b < nfft; b++)
		for (j = 0; j < m; j++) {
			float *base = w + 2 * (b * r * m + j);
			for (q = 0; q < r; q++) {
				v[2 * q]     = base[2 * q * m];
				v[2 * q + 1] = base[2 * q * m + 1];
				if (q && p->tw)
					or_cmul(v + 2 * q, p->tw + 2 * (j * (r - 1) + q - 1));
			}
			or_dft_small(r, v);
			for (q = 0; q < r; q++) {
				base[2 * q * m]     = v[2 * q];
				base[2 * q * m + 1] = v[2 * q + 1];
			}
		}
}

/* work <- conj(work * kern), elementwise (src/fftset_vec.c:847-867). */
static void or_mulconj(float *w, const float *k, unsigned n)
{
	unsigned i;
	for (i = 0; i < n; i++) {
		float dr = w[2 * i], di = w[2 * i + 1];
		float cr = k[2 * i], ci = k[2 * i + 1];
		w[2 * i]     = dr * cr - di * ci;
		w[2 * i + 1] = OR_PLANT_IS(4) ? (dr * ci + di * cr) : -(dr * ci + di * cr);
	}
}

/* Runs the Stockham chain ping-ponging between in and tmp; returns the buffer
 * holding the result (src/fftset_vec.c:1222-1249). */
static float *or_stockham(const struct or_plan *pl, float *in, float *tmp)
{
	unsigned i, s = 1;
	for (i = 0; i < pl->np; i++) {
		float *t;
		or_stock_pass(&pl->p[i], s, in, tmp);
		s *= pl->p[i].radix;
		t = in;
		in = tmp;
		tmp = t;
	}
	return in;
}

/* DIF chain in place; with a kernel, mulconj then the reversed DIT chain
 * (src/fftset_vec.c:1161-1220). */
static void or_dif_passes(const struct or_pass *p, unsigned np, float *w, const float *kern, unsigned L)
{
	unsigned i, nfft = 1;
	unsigned nf[OR_MAX_PASSES];
	for (i = 0; i < np; i++) {
		nf[i] = nfft;
		or_dif_pass(&p[i], nfft, w);
		nfft *= p[i].radix;
	}
	if (!kern)
		return;
	or_mulconj(w, kern, L);
	for (i = np; i-- > 0;)
		or_dit_pass(&p[i], nf[i], w);
}

/* ------------------------------------------------------------------------ */
/* FREQOFFSETREAL modulation factor e^{-j*pi*i/(2L)} with the argument formed
 * in float as src/fftset_mod_freqoffsetreal.c:897-898 forms it. */
static void or_fo_twiddle(unsigned i, unsigned L, float *twr, float *twi)
{
	float arg = -2.0f * (float)M_PI * (float)i / ((float)L * 4.0f);
	*twr = cosf(arg);
	*twi = OR_PLANT_IS(2) ? -sinf(arg) : sinf(arg);
}

/* src/fftset_mod_freqoffsetreal.c:917-924 and :947-954. */
static void or_fo_first(unsigned L, float *w, const float *x)
{
	unsigned i;
	for (i = 0; i < L; i++) {
		float re = x[i], im = x[L + i], twr, twi;
		or_fo_twiddle(i, L, &twr, &twi);
		w[2 * i]     = re * twr + im * twi;
		w[2 * i + 1] = re * twi - im * twr;
	}
}

/* src/fftset_mod_freqoffsetreal.c:926-933 and :1002-1009. */
static void or_fo_final(unsigned L, float *y, const float *w)
{
	unsigned i;
	for (i = 0; i < L; i++) {
		float re = w[2 * i], im = w[2 * i + 1], twr, twi;
		or_fo_twiddle(i, L, &twr, &twi);
		y[i]     = re * twr - im * twi;
		y[L + i] = re * twi + im * twr;
	}
}

static void or_fo_forward(const struct or_plan *pl, float *out, const float *in, float *work)
{
	unsigned L = pl->L, i;
	float *res;
	or_fo_first(L, work, in);
	res = or_stockham(pl, work, out);
	if (res == out)
		memcpy(work, out, sizeof(float) * 2 * L);
	/* bin 2p <- Z[p], bin 2p+1 <- conj Z[L-1-p] (src/fftset_mod_freqoffsetreal.c:958-971) */
	for (i = 0; i < L / 2; i++) {
		out[4 * i + 0] = work[2 * i];
		out[4 * i + 1] = work[2 * i + 1];
		out[4 * i + 2] = work[2 * L - 2 - 2 * i];
		out[4 * i + 3] = OR_PLANT_IS(3) ? work[2 * L - 1 - 2 * i] : -work[2 * L - 1 - 2 * i];
	}
	if (L & 1) {
		out[4 * i + 0] = work[2 * i];
		out[4 * i + 1] = work[2 * i + 1];
	}
}

static void or_fo_inverse(const struct or_plan *pl, float *out, const float *in, float *work)
{
	unsigned L = pl->L, i;
	float *res;
	/* exact inverse of the bin map, conjugated (src/fftset_mod_freqoffsetreal.c:985-998) */
	for (i = 0; i < L / 2; i++) {
		work[2 * i]             = in[4 * i + 0];
		work[2 * i + 1]         = -in[4 * i + 1];
		work[2 * L - 2 - 2 * i] = in[4 * i + 2];
		work[2 * L - 1 - 2 * i] = in[4 * i + 3];
	}
	if (L & 1) {
		work[2 * i]     = in[4 * i + 0];
		work[2 * i + 1] = -in[4 * i + 1];
	}
	res = or_stockham(pl, work, out);
	if (res == out)
		memcpy(work, out, sizeof(float) * 2 * L);
	or_fo_final(L, out, work);
}

/* src/fftset_mod_freqoffsetreal.c:884-901. */
static void or_fo_get_kernel(const struct or_plan *pl, float *out, const float *in)
{
	unsigned L = pl->L, i;
	for (i = 0; i < L; i++) {
		float re = in[i], im = -in[L + i], twr, twi;
		or_fo_twiddle(i, L, &twr, &twi);
		out[2 * i]     = re * twr - im * twi;
		out[2 * i + 1] = re * twi + im * twr;
	}
	or_dif_passes(pl->p, pl->np, out, NULL, L);
}

/* src/fftset_mod_freqoffsetreal.c:903-934. */
static void or_fo_conv(const struct or_plan *pl, float *out, const float *in, const float *kern, float *work)
{
	or_fo_first(pl->L, work, in);
	or_dif_passes(pl->p, pl->np, work, kern, pl->L);
	or_fo_final(pl->L, out, work);
}

/* COMPLEX scalar modulator (src/fftset_mod_cplx.c:242-318). */
static void or_cx_forward(const struct or_plan *pl, float *out, const float *in, float *work)
{
	float *res;
	memcpy(work, in, sizeof(float) * 2 * pl->L);
	res = or_stockham(pl, work, out);
	if (res != out)
		memcpy(out, work, sizeof(float) * 2 * pl->L);
}

static void or_cx_inverse(const struct or_plan *pl, float *out, const float *in, float *work)
{
	unsigned L = pl->L, i;
	float *res;
	for (i = 0; i < L; i++) {
		work[2 * i]     = in[2 * i];
		work[2 * i + 1] = -in[2 * i + 1];
	}
	res = or_stockham(pl, work, out);
	for (i = 0; i < L; i++) {
		out[2 * i]     = res[2 * i];
		out[2 * i + 1] = -res[2 * i + 1];
	}
}

static void or_cx_get_kernel(const struct or_plan *pl, float *out, const float *in)
{
	memcpy(out, in, sizeof(float) * 2 * pl->L);
	or_dif_passes(pl->p, pl->np, out, NULL, pl->L);
}

static void or_cx_conv(const struct or_plan *pl, float *out, const float *in, const float *kern, float *work)
{
	unsigned L = pl->L, i;
	memcpy(work, in, sizeof(float) * 2 * L);
	or_dif_passes(pl->p, pl->np, work, kern, L);
	for (i = 0; i < L; i++) {
		out[2 * i]     = work[2 * i];
		out[2 * i + 1] = -work[2 * i + 1];
	}
}

/* ------------------------------------------------------------------------ */
/* Exported API (ctypes). */

void oracle_destroy(struct or_plan *pl)
{
	unsigned i;
	if (!pl)
		return;
	for (i = 0; i < pl->np; i++)
		free(pl->p[i].tw);
	free(pl);
}

/* Plans a transform of `complex_bins` bins; NULL when the size is
 * unsupported (the reference leaves that case undefined). */
struct or_plan *oracle_create(int mod, unsigned complex_bins)
{
	unsigned radices[OR_MAX_PASSES], n, i, len;
	struct or_plan *pl;
	if (mod != OR_MOD_FREQ_OFFSET_REAL && mod != OR_MOD_COMPLEX)
		return NULL;
	n = oracle_plan_chain(complex_bins, radices, OR_MAX_PASSES);
	if (!n)
		return NULL;
	pl = calloc(1, sizeof(*pl));
	if (!pl)
		return NULL;
	pl->L = complex_bins;
	pl->mod = mod;
	pl->np = n;
	len = complex_bins;
	for (i = 0; i < n; i++) {
		pl->p[i].radix = radices[i];
		pl->p[i].m = len / radices[i];
		pl->p[i].tw = NULL;
		if (pl->p[i].m > 1) {
			pl->p[i].tw = or_make_twiddles(radices[i], pl->p[i].m);
			if (!pl->p[i].tw) {
				oracle_destroy(pl);
				return NULL;
			}
		}
		len /= radices[i];
	}
	return pl;
}

unsigned oracle_plan_passes(const struct or_plan *pl, unsigned *radices, unsigned max)
{
	unsigned i;
	if (pl->np > max)
		return 0;
	for (i = 0; i < pl->np; i++)
		radices[i] = pl->p[i].radix;
	return pl->np;
}

void oracle_forward(const struct or_plan *pl, float *out, const float *in, float *work)
{
	if (pl->mod == OR_MOD_FREQ_OFFSET_REAL)
		or_fo_forward(pl, out, in, work);
	else
		or_cx_forward(pl, out, in, work);
}

void oracle_inverse(const struct or_plan *pl, float *out, const float *in, float *work)
{
	if (pl->mod == OR_MOD_FREQ_OFFSET_REAL)
		or_fo_inverse(pl, out, in, work);
	else
		or_cx_inverse(pl, out, in, work);
}

void oracle_get_kernel(const struct or_plan *pl, float *out, const float *in)
{
	if (pl->mod == OR_MOD_FREQ_OFFSET_REAL)
		or_fo_get_kernel(pl, out, in);
	else
		or_cx_get_kernel(pl, out, in);
}

void oracle_conv(const struct or_plan *pl, float *out, const float *in, const float *kern, float *work)
{
	if (pl->mod == OR_MOD_FREQ_OFFSET_REAL)
		or_fo_conv(pl, out, in, kern, work);
	else
		or_cx_conv(pl, out, in, kern, work);
}

/* src/fftset.c:131-167. */
static unsigned or_round_up_factorisation(unsigned min)
{
	unsigned length = 1;
	while (length < min) {
		if (length * 2 >= min)
			length *= 2;
		else if (length * 3 >= min)
			length *= 3;
		else
			length *= 4;
	}
	return length;
}

unsigned oracle_recommend_conv_length(unsigned kernel_length, unsigned max_block_size)
{
	const unsigned multiple = 32;
	unsigned target = 8 * kernel_length;
	unsigned block = max_block_size > target ? max_block_size : target;
	unsigned min_real = kernel_length + block - 1;
	return (multiple / 2) * or_round_up_factorisation((min_real + multiple - 1) / multiple);
}

/* ------------------------------------------------------------------------ */
/* CPU baseline: `n` transforms with buffers of 2L floats at stride 2L,
 * one transform per thread at a time, `nthreads` threads. */

struct or_batch_arg {
	const struct or_plan *pl;
	int op;
	unsigned n, first, step;
	float *out;
	const float *in;
	const float *kern;
};

static void *or_batch_worker(void *varg)
{
	struct or_batch_arg *a = varg;
	size_t len2 = 2 * (size_t)a->pl->L;
	float *work = malloc(sizeof(float) * len2);
	unsigned t;
	if (!work)
		return (void *)1;
	for (t = a->first; t < a->n; t += a->step) {
		float *o = a->out + t * len2;
		const float *i = a->in + t * len2;
		switch (a->op) {
		case OR_OP_FORWARD: oracle_forward(a->pl, o, i, work); break;
		case OR_OP_INVERSE: oracle_inverse(a->pl, o, i, work); break;
		case OR_OP_GETKERN: oracle_get_kernel(a->pl, o, i); break;
		case OR_OP_CONV:    oracle_conv(a->pl, o, i, a->kern + t * len2, work); break;
		}
	}
	free(work);
	return NULL;
}

/* Returns 0 on success. */
int oracle_run_batch(const struct or_plan *pl, int op, unsigned n, float *out, const float *in,
                     const float *kern, unsigned nthreads)
{
	pthread_t th[256];
	struct or_batch_arg args[256];
	unsigned i, started = 0;
	int rc = 0;
	if (nthreads < 1)
		nthreads = 1;
	if (nthreads > 256)
		nthreads = 256;
	for (i = 0; i < nthreads; i++) {
		args[i].pl = pl;
		args[i].op = op;
		args[i].n = n;
		args[i].first = i;
		args[i].step = nthreads;
		args[i].out = out;
		args[i].in = in;
		args[i].kern = kern;
	}
	if (nthreads == 1)
		return or_batch_worker(&args[0]) ? -1 : 0;
	for (i = 0; i < nthreads; i++) {
		if (pthread_create(&th[i], NULL, or_batch_worker, &args[i]))
			break;
		started++;
	}
	for (i = 0; i < started; i++) {
		void *ret = NULL;
		pthread_join(th[i], &ret);
		if (ret)
			rc = -1;
	}
	return started == nthreads ? rc : -1;
}
