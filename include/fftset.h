/* fftset.h -- MI355X (gfx950) implementation of the fftset C API.
 *
 * Drop-in for the reference interface in fftset.h of templeblock/fftset:
 * same names, argument meaning and buffer rules, except that every buffer
 * argument is a DEVICE pointer (hipMalloc'd memory, or any memory the GPU can
 * address) aligned to at least 8 bytes (one complex value; hipMalloc gives
 * 256). The synchronous drop-in calls also accept host arrays and stage them
 * through device memory. The drop-in execution calls run on a
 * per-thread HIP stream and return when their results are complete, as the
 * reference's CPU calls do. Extensions (stream-ordered variants, uniform and
 * ragged batches, convolution adjoints) follow the drop-in section and have
 * no counterpart in the reference.
 *
 * Sizes: complex_bins = 2^a 3^b 5^{0,1}, complex_bins >= 2 (the reference's
 * supported set; it leaves other sizes undefined, this library returns NULL).
 * Every buffer holds 2*complex_bins floats. */
#ifndef FFTSET_H
#define FFTSET_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- types */

/* Holds the memory for every transform created from it (plans and the
 * shared FP32 twiddle tables in HBM). A complete type so it can live on the
 * stack (reference fftset.h:29-31, 159-167); the members are private. */
struct fftset {
	void *impl;
	void *reserved[7];
};

/* Opaque (reference fftset.h:33-39). */
struct fftset_modulation;
struct fftset_fft;

/* ---------------------------------------------- initialisation / cleanup */

/* Replaces fftset_init (reference fftset.h:44-47, src/fftset.c:80-85).
 * Binds the set to the current HIP device. Returns 0 on success, non-zero
 * when memory is exhausted or no device is usable. */
int fftset_init(struct fftset *fc);

/* Replaces fftset_destroy (reference fftset.h:49-52, src/fftset.c:87-95).
 * Frees every plan and twiddle table; every fftset_fft from this set becomes
 * invalid. */
void fftset_destroy(struct fftset *fc);

/* ------------------------------------------------------------ modulations */

/* Replaces FFTSET_MODULATION_FREQ_OFFSET_REAL (reference fftset.h:62-73):
 *   X[k] = sum_{n<N} x[n] e^{-2 pi i n (k + 1/2) / N},  N = 2*complex_bins,
 * x real (N floats in), X[k] for k < complex_bins interleaved re/im out.
 * Its convolution is negacyclic of length N, scaled by complex_bins. */
extern const struct fftset_modulation *FFTSET_MODULATION_FREQ_OFFSET_REAL;

/* Replaces FFTSET_MODULATION_COMPLEX (reference fftset.h:75-76): a plain
 * complex DFT of complex_bins points, interleaved re/im, natural order;
 * inverse unnormalised; convolution cyclic, scaled by complex_bins. */
extern const struct fftset_modulation *FFTSET_MODULATION_COMPLEX;

/* Replaces fftset_create_fft (reference fftset.h:78, src/fftset.c:97-129).
 * Returns the existing plan for (modulation, complex_bins) or creates it;
 * NULL for an unsupported size or when memory is exhausted. Not thread-safe
 * (as in the reference). */
const struct fftset_fft *fftset_create_fft(struct fftset *fc, const struct fftset_modulation *modulation,
                                           unsigned complex_bins);

/* -------------------------------------------------------------- execution
 * Buffer rules as in the reference (fftset.h:80-108): input and output may
 * alias; work must alias neither; kernel_buf must alias nothing. Buffers are
 * device pointers. Errors are reported through fftset_last_error(). */

/* Replaces fftset_fft_forward (reference fftset.h:109-115, src/fftset.c:58-67). */
void fftset_fft_forward(const struct fftset_fft *first_pass, float *output_buf, const float *input_buf,
                        float *work_buf);

/* Replaces fftset_fft_inverse (reference fftset.h:117-123, src/fftset.c:69-78). */
void fftset_fft_inverse(const struct fftset_fft *first_pass, float *output_buf, const float *input_buf,
                        float *work_buf);

/* Replaces fftset_fft_conv_get_kernel (reference fftset.h:125-130,
 * src/fftset.c:36-44). The output is an opaque, unnormalised kernel spectrum
 * of 2*complex_bins floats in this library's own order, valid only for
 * the same fftset_fft. input_buf and output_buf must not alias. */
void fftset_fft_conv_get_kernel(const struct fftset_fft *first_pass, float *output_buf, const float *input_buf);

/* Replaces fftset_fft_conv (reference fftset.h:132-139, src/fftset.c:46-56):
 * FREQOFFSETREAL: out = complex_bins * negacyclic_N(in, h);
 * COMPLEX:        out = complex_bins * cyclic(in, h),
 * where kernel_buf = fftset_fft_conv_get_kernel(h). */
void fftset_fft_conv(const struct fftset_fft *first_pass, float *output_buf, const float *input_buf,
                     const float *kernel_buf, float *work_buf);

/* Replaces fftset_recommend_conv_length (reference fftset.h:144-151,
 * src/fftset.c:131-167), bit-exact. The result is in complex bins. */
unsigned fftset_recommend_conv_length(unsigned kernel_length, unsigned max_block_size);

/* ============================================================ extensions
 * None of the following has a counterpart in the reference. */

/* Error codes. */
#define FFTSET_OK               0
#define FFTSET_E_INVALID       (-1)  /* bad argument (NULL plan, bad op, ...) */
#define FFTSET_E_NOMEM         (-2)  /* host or device allocation failed      */
#define FFTSET_E_DEVICE        (-3)  /* a HIP call failed                     */
#define FFTSET_E_UNSUPPORTED   (-4)  /* size or operation not supported       */
#define FFTSET_E_TOO_LARGE     (-5)  /* work space beyond the no-allocation limit
                                        of fftset_fft_execute_noalloc          */

/* Returns the code of the most recent failure on the calling thread and
 * clears it (FFTSET_OK when nothing failed since the last call). */
int fftset_last_error(void);

/* Operations for the batched and stream-ordered entry points. */
#define FFTSET_OP_FORWARD          0
#define FFTSET_OP_INVERSE          1
#define FFTSET_OP_GET_KERNEL       2
#define FFTSET_OP_CONV             3
#define FFTSET_OP_CONV_GRAD_INPUT  4
#define FFTSET_OP_CONV_GRAD_KERNEL 5
#define FFTSET_OP_CONV_SAVE_SPECTRUM 6

/* Uniform batch, stream-ordered (returns before completion): `count`
 * transforms of one plan; transform t reads in + t*in_stride, writes
 * out + t*out_stride (strides in floats, >= 2*complex_bins unless 0 for a
 * buffer shared by all transforms, which is allowed for kern only).
 *   FORWARD / INVERSE / GET_KERNEL: in -> out
 *   CONV:             out = conv(in, kern)
 *   CONV_GRAD_INPUT:  out = adjoint of conv w.r.t. the signal applied to
 *                     in (the output gradient), with kern = get_kernel(h)
 *   CONV_GRAD_KERNEL: out = adjoint w.r.t. the kernel taps applied to kern
 *                     (the output gradient) with in = the signal x
 *   CONV_SAVE_SPECTRUM: out = conv(in, kern), and work (required, stride
 *                     work_stride) receives the signal's spectrum in the
 *                     kernel-spectrum layout, i.e. get_kernel(in). The kernel
 *                     gradient of a later backward pass is then one
 *                     CONV_GRAD_INPUT of the output gradient with that
 *                     spectrum as kern, instead of CONV_GRAD_KERNEL's
 *                     get_kernel of the signal plus that pass.
 * work (may be NULL) provides count*work_stride floats of scratch for the
 * operations that need it (work_stride >= 2*complex_bins; GET_KERNEL needs
 * only 16384 floats per transform, for 61440-bin plans); when NULL the
 * library uses its own, one buffer
 * per stream (so calls on different streams never share it). A buffer that
 * must grow is replaced, and the old one is freed only by fftset_destroy,
 * so work=NULL calls on one stream from several host threads are safe;
 * growing is refused (FFTSET_E_INVALID) while the stream is being captured
 * into a graph (pass work, or use fftset_fft_execute_noalloc, there).
 * stream is a hipStream_t (NULL = the default stream). Returns FFTSET_OK or
 * an error. */
int fftset_fft_execute(const struct fftset_fft *fft, int op, size_t count, float *out, size_t out_stride,
                       const float *in, size_t in_stride, const float *kern, size_t kern_stride, float *work,
                       size_t work_stride, void *stream);

/* fftset_fft_execute with work = NULL, for callers that cannot allocate
 * device memory (code under graph capture, serving loops, harnesses that
 * forbid hipMalloc). No counterpart in the reference. It calls no device or
 * host allocator: the work space that fftset_fft_execute would allocate
 * comes from a static device array of FFTSET_NOALLOC_WORK_FLOATS floats
 * (64 MiB, held by each GPU the library runs on until the process ends).
 * Needs (floats): four-step plans (complex_bins > 16384), every operation
 * but GET_KERNEL: count * 2 * complex_bins, CONV_GRAD_KERNEL twice that;
 * GET_KERNEL of a 61440-bin plan: count * 16384; LDS plans:
 * CONV_GRAD_KERNEL count * 2 * complex_bins, others none.
 * Larger calls fail with FFTSET_E_TOO_LARGE before any work is queued.
 * Every call on one GPU shares that GPU's array: calls that overlap in time
 * on one GPU (different streams or host threads without ordering) race and
 * give wrong results. Other arguments as fftset_fft_execute. */
#define FFTSET_NOALLOC_WORK_FLOATS (16u << 20)
int fftset_fft_execute_noalloc(const struct fftset_fft *fft, int op, size_t count, float *out, size_t out_stride,
                               const float *in, size_t in_stride, const float *kern, size_t kern_stride,
                               void *stream);

/* Adjoints of fftset_fft_conv (synchronous, like the drop-in calls).
 * For FREQOFFSETREAL, with y = conv(x, h) = L * negacyclic(x, h) and an
 * output gradient g = dLoss/dy:
 *   grad_input  = dLoss/dx = L * negacyclic(g, h_flat),  h_flat[0] = h[0],
 *                 h_flat[k] = -h[N-k]  (spectrum: conj of the kernel's)
 *   grad_kernel = dLoss/dh = L * negacyclic(g, x_flat)
 * For COMPLEX the same calls apply the conjugate-transpose maps.
 * kernel_buf = fftset_fft_conv_get_kernel(h); work_buf holds 2L floats and
 * aliases nothing. */
void fftset_fft_conv_grad_input(const struct fftset_fft *first_pass, float *grad_input, const float *grad_output,
                                const float *kernel_buf, float *work_buf);
void fftset_fft_conv_grad_kernel(const struct fftset_fft *first_pass, float *grad_kernel, const float *grad_output,
                                 const float *input_buf, float *work_buf);

/* ---------------------------------------------------------- ragged batches
 * Thousands of transforms of different sizes, modulations and operations
 * run as one graph launch. Buffers are recorded at add time and read at
 * every launch; they must stay valid until the batch is destroyed. */
struct fftset_batch;

#define FFTSET_BATCH_NO_GRAPH      1u  /* launch the kernels directly instead of a captured hipGraph */
#define FFTSET_BATCH_THREADS_512   2u  /* at least 512-lane workgroups (default: each plan's packing) */
#define FFTSET_BATCH_SCHED_SERIAL  4u  /* run the tile groups one after another                      */
#define FFTSET_BATCH_SCHED_PHASED  8u  /* LDS groups beside the first four-step step only            */
#define FFTSET_BATCH_PERSISTENT   16u  /* resident-size grids taking tiles from a counter, instead of
                                            one workgroup per tile                                     */
#define FFTSET_BATCH_FOUR_STEP_TILES 32u /* four-step steps as ragged tiles of one launch per step and
                                            workgroup size, instead of the four-step kernels once per
                                            (class, step) reading per-transform pointer arrays        */

#define FFTSET_BASELINE_PER_GROUP      1  /* one launch per (plan, operation) group and step */
#define FFTSET_BASELINE_PER_TRANSFORM  2  /* one launch per transform and step               */
#define FFTSET_BASELINE_PER_PASS       3  /* one launch per radix pass of this library's chain
                                             for each group, HBM round trip between passes
                                             (four-step groups: one launch per step)           */
#define FFTSET_BASELINE_REFERENCE_PASSES 4 /* the reference's pass structure for each group: its
                                             modulating / splitting first pass, one sweep per
                                             pass of its planner's chain at its vector width,
                                             the separate product sweep of a convolution and
                                             its reorder or demodulating last pass, one launch
                                             per sweep (see fftset_baseline_chain)            */

/* Creates an empty batch on the set; NULL on failure. */
struct fftset_batch *fftset_batch_create(struct fftset *fc);

/* Appends one transform: op is FFTSET_OP_FORWARD, _INVERSE, _GET_KERNEL,
 * _CONV, _CONV_GRAD_INPUT or _CONV_GRAD_KERNEL, with the buffers of
 * fftset_fft_execute for that op (kern for the convolutions; for
 * _CONV_GRAD_KERNEL in = the signal, kern = the output gradient). A kernel
 * gradient runs as a get_kernel of its signal into a spectrum the batch
 * owns, then, after every other transform of the launch, the signal
 * adjoint of its output gradient against that spectrum. Returns the
 * transform's index or an error. */
int fftset_batch_add(struct fftset_batch *b, const struct fftset_fft *fft, int op, float *out, const float *in,
                     const float *kern);

/* Plans the tiles, uploads the descriptors and (unless
 * FFTSET_BATCH_NO_GRAPH) captures the launches into a hipGraph. */
int fftset_batch_commit(struct fftset_batch *b, unsigned flags);

/* Runs the whole batch on a hipStream_t (stream-ordered, asynchronous).
 * Launches of one batch share its four-step work buffers: run them one at
 * a time (on one stream, or synchronised). */
int fftset_batch_launch(struct fftset_batch *b, void *stream);

/* Runs the same work as FFTSET_BASELINE_* separate launches (for
 * comparison). */
int fftset_batch_launch_baseline(struct fftset_batch *b, int mode, void *stream);

/* The pass structure FFTSET_BASELINE_REFERENCE_PASSES runs for
 * (modulation, complex_bins), as the reference plans it
 * (src/fftset_mod_freqoffsetreal.c:1012-1131, src/fftset_mod_cplx.c:320-347,
 * src/fftset_vec.c:944-971, 1004-1062): the radices of the first pass's
 * split into the vector width's sub-transforms ([4] at width 4, [4, 2] at
 * width 8, none at width 1), a 0, then the planner's chain for length
 * complex_bins / width. Returns the number of entries written (at most
 * `max`), 0 for an unsupported size. Host only. */
unsigned fftset_baseline_chain(const struct fftset_modulation *modulation, unsigned complex_bins,
                               unsigned *radices, unsigned max);

/* Algorithmic bytes of one launch: 16 L per forward / inverse / get_kernel,
 * 24 L per convolution, signal adjoint or kernel gradient. */
long long fftset_batch_bytes(const struct fftset_batch *b);

/* Tile count, class count, LDS bytes per workgroup and the number of
 * classes run as four-step transforms. */
int fftset_batch_info(const struct fftset_batch *b, int *ntiles, int *nclasses, int *lds_bytes,
                      int *four_step_classes);

void fftset_batch_destroy(struct fftset_batch *b);

/* --------------------------------------------------------------- tuning
 * The radix chain, transforms per workgroup and (for sizes beyond LDS) the
 * four-step split of a size are picked by timing candidates on the device.
 * A result applies to plans created afterwards on the same set (a plan's
 * kernel spectrum layout depends on its chain, so plans never change):
 * call fftset_autotune before fftset_create_fft for that size. Tables
 * persist as text; FFTSET_AMD_TUNE_FILE names one loaded by fftset_init. */

/* Times the candidates for (modulation, complex_bins) on the operations in
 * `ops` (a mask of 1 << FFTSET_OP_*; 0 = forward + convolution) and records
 * the fastest; *best_ms (may be NULL) receives its time per 64 MiB of
 * transforms. Fails with FFTSET_E_INVALID if the plan exists. */
int fftset_autotune(struct fftset *fc, const struct fftset_modulation *modulation, unsigned complex_bins,
                    unsigned ops, float *best_ms);
int fftset_tune_save(const struct fftset *fc, const char *path);
int fftset_tune_load(struct fftset *fc, const char *path);

/* ----------------------------------------------------- block convolution
 * No counterpart in the reference: the overlap-save driver that the
 * reference's fftset_recommend_conv_length (fftset.h:144-151) sizes
 * transforms for. Causal FIR filtering of long real signals,
 *   y[n] = sum_{j<k} h[j] x[n-j],  0 <= n < nsamples,
 * with x[n] = 0 for n < 0 (or, with FFTSET_BLOCKCONV_HISTORY, read from the
 * k-1 samples the caller keeps in front of every channel). The transform
 * has L = fftset_recommend_conv_length(k, max_block_size) complex bins;
 * every block of 2L - k + 1 samples of every channel is one FREQOFFSETREAL
 * convolution, all of them in one launch. Sizes with L > 16384 are not
 * supported (FFTSET_E_UNSUPPORTED). */
struct fftset_blockconv;

#define FFTSET_BLOCKCONV_HISTORY 1u  /* in[c][-(k-1) .. -1] hold the samples before the block */

/* taps: k floats in device memory (read once). The set must outlive the
 * object. NULL on failure (see fftset_last_error). */
struct fftset_blockconv *fftset_blockconv_create(struct fftset *fc, const float *taps, unsigned kernel_length,
                                                 unsigned max_block_size);

/* Filters `channels` signals of `nsamples` samples (channel c at
 * in + c * in_stride, output at out + c * out_stride; strides in floats).
 * Stream-ordered on a hipStream_t; out must not overlap in. */
int fftset_blockconv_run(struct fftset_blockconv *bc, float *out, size_t out_stride, const float *in,
                         size_t in_stride, size_t nsamples, size_t channels, unsigned flags, void *stream);

/* New samples per block (2L - k + 1) and the transform's complex bins. */
unsigned fftset_blockconv_block_size(const struct fftset_blockconv *bc);
unsigned fftset_blockconv_complex_bins(const struct fftset_blockconv *bc);

void fftset_blockconv_destroy(struct fftset_blockconv *bc);

/* Kernel selection, process-wide (for tests and comparisons). Chains with a
 * compile-time specialisation run it by default; FFTSET_SELECT_GENERIC runs
 * the runtime-chain kernels instead, FFTSET_SELECT_PERSISTENT resident-size
 * grids with register prefetch. 61440-bin transforms (the 240 x 256 split)
 * run by default with the whole transform resident on one CU (one launch,
 * intermediate in LDS and VGPRs), except FREQOFFSETREAL forward and inverse;
 * FFTSET_SELECT_FOUR_STEP_SPLIT runs them with separate column and row
 * launches (intermediate in the work buffer), and
 * FFTSET_SELECT_FOUR_STEP_FUSED with one workgroup per transform and its
 * intermediate in a per-workgroup scratch slot (measured slower on MI355X;
 * kept for comparison). Kernel spectra are interchangeable between the
 * three. Initial value from the environment (FFTSET_AMD_GENERIC=1,
 * FFTSET_AMD_PERSIST=1), read once at load. Returns the previous flags. */
#define FFTSET_SELECT_GENERIC          1u
#define FFTSET_SELECT_PERSISTENT       2u
#define FFTSET_SELECT_FOUR_STEP_FUSED  4u
#define FFTSET_SELECT_FOUR_STEP_SPLIT  8u
unsigned fftset_kernel_select(unsigned flags);
/* Launches of specialised and of runtime-chain kernels issued by the
 * drop-in and uniform-batch calls since the previous call (which resets
 * both counts). */
void fftset_kernel_counts(unsigned long long *specialised, unsigned long long *generic);

/* Plan introspection (for tests, benchmarks and the autotuner). */
unsigned fftset_fft_complex_bins(const struct fftset_fft *fft);
/* HBM bytes held by the set's shared twiddle and modulation tables (no
 * counterpart in the reference, whose tables live in host memory). Tables
 * made only for fftset_autotune's candidates are freed before it returns. */
unsigned long long fftset_table_bytes(const struct fftset *fc);
/* Writes at most `max` radices of the plan's LDS chain; returns the count
 * (for four-step plans: the chains of the column and row transforms, with
 * a 0 separating them). */
unsigned fftset_fft_radices(const struct fftset_fft *fft, unsigned *radices, unsigned max);

#ifdef __cplusplus
}
#endif

#endif /* FFTSET_H */
