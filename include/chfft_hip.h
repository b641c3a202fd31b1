/*
 * chfft_hip.h — C ABI of the MI355X (gfx950) transform engine.
 *
 * Drop-in boundary for chfft 0.3.4's five public transform types
 * (lib.rs:21-25: CFft1D, CFft2D, Dct1D/DctType, Mdct1D, RFft1D). The reference
 * has no FFI of its own; each entry point below replaces the Rust method it
 * cites, and a language binding (a Rust shim over this header, the Python
 * mirror in chfft_amd/) keeps the reference's method names on top of it.
 *
 * Conventions
 *  - Plans are created for one length (or shape) and a batch count, and own
 *    their device tables and scratch. A plan is used by one host thread at a
 *    time (the reference takes &mut self on every method); distinct plans may
 *    run concurrently.
 *  - chfft_exec takes DEVICE pointers and is asynchronous on `stream`
 *    (a hipStream_t; NULL = default stream). chfft_exec_host takes HOST
 *    pointers, copies through the plan's staging buffers and blocks.
 *  - Complex data is interleaved {re, im} (num-complex Complex<T> layout),
 *    batch-major: transform b starts at element b * (its length).
 *  - precision: CHFFT_F32 (FP32 butterflies, FP64-generated tables rounded
 *    once) or CHFFT_F64 (the plain method).
 *  - in == out is allowed for every complex-to-complex plan (the reference's
 *    in-place *i methods, cfft1d.rs:403-455).
 */
#ifndef CHFFT_HIP_H
#define CHFFT_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct chfft_plan_s* chfft_plan;

typedef enum {
  CHFFT_OK = 0,
  /* construct/setup/exec length errors: the reference panics with
     "invalid length" (rfft1d.rs:74-76,116-122; dct1d.rs:54-56,237-243;
     mdct1d.rs:102-104,198-204) */
  CHFFT_ERR_INVALID_LENGTH = 1,
  CHFFT_ERR_INVALID_ARGUMENT = 2, /* null pointer, unknown enum value */
  CHFFT_ERR_UNSUPPORTED = 3,      /* method the transform type does not have */
  CHFFT_ERR_HIP = 4,              /* a HIP runtime call failed */
  CHFFT_ERR_OUT_OF_MEMORY = 5,
  /* a multi-length launch could not complete: one of its grid-wide barriers
     timed out (its workgroups were not all resident), so the output is
     incomplete; reported by chfft_plan_check and chfft_exec_host */
  CHFFT_ERR_INCOMPLETE = 6,
  /* chfft_exec_noalloc: the plan needs more scratch than
     CHFFT_NOALLOC_SCRATCH_BYTES */
  CHFFT_ERR_SCRATCH_LIMIT = 7
} chfft_status;

/* Size of the device array behind chfft_exec_noalloc (96 MiB per GPU). */
#define CHFFT_NOALLOC_SCRATCH_BYTES (96u << 20)

typedef enum { CHFFT_F32 = 0, CHFFT_F64 = 1 } chfft_precision;

/* Methods, named and scaled as the reference's (cfft1d.rs:270-389,
   rfft1d.rs:199-312, cfft2d.rs:150-324, dct1d.rs:98-144, mdct1d.rs:149-195).
   The in-place *i variants are these methods with in == out. */
typedef enum {
  CHFFT_FORWARD = 0,   /* scale 1                (MDCT: sqrt(N/2)) */
  CHFFT_FORWARD0 = 1,  /* scale 1 */
  CHFFT_FORWARDU = 2,  /* 1/sqrt(N)              (DCT: sqrt(2/N); MDCT: 1) */
  CHFFT_FORWARDN = 3,  /* 1/N                    (DCT: 2/N) */
  CHFFT_BACKWARD = 4,  /* 1/N                    (MDCT: 1/sqrt(N/2)) */
  CHFFT_BACKWARD0 = 5, /* 1 */
  CHFFT_BACKWARDU = 6, /* 1/sqrt(N)              (MDCT: 1) */
  CHFFT_BACKWARDN = 7  /* 1/N */
} chfft_method;

typedef enum { CHFFT_DCT_II = 2, CHFFT_DCT_III = 3 } chfft_dct_type;

/* CFft1D::with_len / setup (cfft1d.rs:83-92, 103-206): `batch` transforms of
   `len` complex points. len >= 1; 7-smooth lengths use mixed radix, others
   Bluestein (MAX_PRIME = 7, cfft1d.rs:104). */
int chfft_plan_c2c_1d(chfft_plan* plan, size_t len, size_t batch, int precision);

/* RFft1D::new / setup (rfft1d.rs:73-113): len even (else
   CHFFT_ERR_INVALID_LENGTH). Forward: len reals -> len/2+1 complex per
   transform; backward: the reverse. Any even len: a half length above one
   workgroup's LDS (FP32 16384, FP64 8192 points; 8192 / 4096 when it is not
   7-smooth) runs its complex transform as a four-step over a plan-owned work
   buffer, with the split as a pass of its own (also Dct1D and Mdct1D). */
int chfft_plan_r2c_1d(chfft_plan* plan, size_t len, size_t batch, int precision);

/* CFft2D::with_len / setup (cfft2d.rs:93-124): len_m rows of len_n complex
   points, row-major, `batch` matrices. Columns longer than one workgroup's
   LDS run as rows of the transposed matrices in a plan-owned work buffer. */
int chfft_plan_c2c_2d(chfft_plan* plan, size_t len_m, size_t len_n, size_t batch, int precision);

/* Dct1D::new / setup (dct1d.rs:53-86): dct_type CHFFT_DCT_II or
   CHFFT_DCT_III, len even, is_ortho as the reference's flag. len reals in,
   len reals out. */
int chfft_plan_dct_1d(chfft_plan* plan, int dct_type, size_t len, int is_ortho, size_t batch, int precision);

/* Mdct1D::new / with_sine / with_vorbis / setup (mdct1d.rs:55-137): len % 4
   == 0. `window` holds the len window values w(len, p) (the reference
   evaluates its window closure into the same table, mdct1d.rs:91-93).
   Forward reads `batch` frames of len samples, frame b starting at sample
   b * frame_stride (frame_stride = len for independent frames, len/2 for a
   50 % overlapped signal; 0 means len), and writes len/2 coefficients per
   frame. Backward reads len/2 coefficients and writes len windowed samples
   per frame. A plan whose frame_stride differs from len owns a buffer of
   batch * len samples for chfft_exec_mdct_ola. */
int chfft_plan_mdct_1d(chfft_plan* plan, size_t len, const double* window, size_t batch, size_t frame_stride,
                       int precision);

/* Stream synthesis (no counterpart in the reference, whose Mdct1D::backward,
   mdct1d.rs:244-291, returns windowed frames and leaves their overlap-add to
   the caller): `method` CHFFT_BACKWARD or CHFFT_BACKWARDU of the plan's
   `batch` coefficient frames, then the overlap-add of the windowed frames
   at frame_stride, in frame order, into `out`, which receives
   (batch - 1) * frame_stride + len samples. With a Princen-Bradley window,
   hop len/2 and CHFFT_BACKWARD after CHFFT_FORWARD (or the U pair), the
   samples from len/2 to (batch - 1) * frame_stride + len/2 reproduce the
   analysed signal (TDAC). Device pointers, asynchronous on `stream`. */
int chfft_exec_mdct_ola(chfft_plan plan, const void* in, void* out, int method, void* stream);

/* Uneven batch (no counterpart type in the reference; each transform is the
   reference's CFft1D::forward/backward of its own length, cfft1d.rs:208-230):
   `count` complex transforms of lengths lens[0..count), packed back to back
   in one buffer (transform i starts at element sum(lens[0..i))). One exec is
   ONE persistent kernel launch: each workgroup takes whole transforms, in LDS
   when they fit, else as a four-step (Bluestein for non-7-smooth lengths)
   through its own scratch slot. Tables are generated on the GPU in FP64 and
   rounded once; Bluestein spectra by an FP64 FFT. Methods scale each
   transform by its own length, as CFft1D does. */
int chfft_plan_c2c_multi(chfft_plan* plan, const size_t* lens, size_t count, int precision);

/* Workgroup count of a multi-length plan's persistent launch (0 = default). */
int chfft_plan_set_grid(chfft_plan plan, int grid);

/* Waits for `stream` and reports whether the multi-length launches of this
   plan since the last check completed: CHFFT_ERR_INCOMPLETE when one of
   their grid-wide barriers timed out (the launch needs all its workgroups
   resident at once; it is launched cooperatively where the stream is not
   being captured, which guarantees that, and otherwise relies on the grid
   being sized to the occupancy), else CHFFT_OK. Clears the record. Other
   plan kinds: CHFFT_OK after the wait. */
int chfft_plan_check(chfft_plan plan, void* stream);

/* Autotuner (no counterpart in the reference; it takes the role of the
   reference's fixed factor order, prime_factorization.rs:66-113): times
   the schedules this plan can run on the given device buffers (their
   contents are overwritten), rebuilds the plan with the fastest and returns
   its index (>= 0; 0 is the planner's own choice) or minus a status code.
   `reps` timed executions per candidate, best of two runs. Candidates, as
   chfft_plan_tune specs: a fused length tries every radix order of the
   runtime-radix kernel and workgroup shapes around the planner's (half /
   double the transforms per workgroup, the fewest threads per transform,
   half / double the threads per transform) against the length-specialised
   kernel; a four-step length the persistent schedule, the runtime-radix
   kernels, chunk sizes, other splits N = n1 x n2 and radix orders (a
   Bluestein length fused into one workgroup: the four-step schedules); a
   multi-length plan workgroup counts, work-unit sizes, generic tile widths
   and the runtime-radix tiles. CFft1D and multi-length plans; other kinds
   return 0. */
int chfft_plan_autotune(chfft_plan plan, void* in, void* out, int reps, void* stream);

/* Planner overrides (no counterpart in the reference): rebuilds a CFft1D or
   multi-length plan under `spec`, "key=value" pairs separated by spaces or
   commas ("default" or "" = the planner's own choice):
     generic=1   runtime-radix kernels only (no length-specialised kernel)
     order=k     radix schedule k of each runtime-radix sequence (0 = the
                 planner's; fused lengths: an order the length has)
     tpb=, tpt=  fused length: transforms per workgroup, threads per
                 transform (threads per transform >= length / 16);
                 multi-length (tpb only): the most transforms of a generic
                 four-step tile, a power of two up to 128 (default 64)
     n1=         four-step split N = n1 x n2 (7-smooth factors one
                 workgroup holds)
     chunk=      four-step: transforms per chunk (-1 = the whole batch)
     sched=1     four-step: one persistent launch
     sched=2     four-step: split launches, also for a Bluestein length the
                 planner would fuse into one workgroup
     grid=, unit= multi-length: workgroups of the persistent launch,
                 points per work unit of a four-step phase (>= 1024)
   CHFFT_ERR_INVALID_ARGUMENT for a spec the plan cannot take (the plan is
   left as it was); CHFFT_ERR_UNSUPPORTED for overrides on other plan kinds.
   The device must be idle of this plan's execs (the call waits for the
   device before replacing the plan's buffers). */
int chfft_plan_tune(chfft_plan plan, const char* spec);

/* The spec the plan was built with, a newline, and the last autotune's
   report (every candidate's time), NUL-terminated, truncated to buf_len. */
int chfft_plan_tuning(chfft_plan plan, char* buf, size_t buf_len);

int chfft_plan_destroy(chfft_plan plan);

/* Run `method` on device buffers, asynchronously on `stream`. */
int chfft_exec(chfft_plan plan, const void* in, void* out, int method, void* stream);

/* No counterpart in the reference: for callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc). Same as chfft_exec, but the plan's four-step scratch
   and work buffer are taken from a library-owned device array of
   CHFFT_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until
   the process ends; the call itself allocates and frees nothing. A plan
   whose chfft_plan_scratch_bytes plus work buffer exceed the array gets
   CHFFT_ERR_SCRATCH_LIMIT (e.g. the 43200 x 512 four-step fits in both
   precisions; Bluestein 10007 x 512 in FP64 does not). All calls on one GPU
   share the array: calls that overlap in time on one GPU race. */
int chfft_exec_noalloc(chfft_plan plan, const void* in, void* out, int method, void* stream);

/* Profiling hook (no counterpart in the reference): run only kernel launches
   [first, first + count) of chfft_exec's schedule for `method`, so that a
   caller can bracket each kernel with its own events. Running every range in
   order equals chfft_exec. */
int chfft_exec_range(chfft_plan plan, const void* in, void* out, int method, int first, int count, void* stream);

/* Number of kernel launches chfft_exec makes for `method`. */
int chfft_plan_launch_count(chfft_plan plan, int method);

/* Run `method` on host buffers; blocks until the result is in `out`. */
int chfft_exec_host(chfft_plan plan, const void* in, void* out, int method);

/* Sizes in bytes of one exec's input and output for `method`, and of the
   plan's device scratch. */
int chfft_plan_io_bytes(chfft_plan plan, int method, size_t* in_bytes, size_t* out_bytes);
size_t chfft_plan_scratch_bytes(chfft_plan plan);

/* Human-readable kernel schedule of the plan (launch count, per launch the
   operators, FFT length, radix passes, transforms per workgroup, workgroup
   size). Returns the number of kernel launches one exec makes. */
int chfft_plan_describe(chfft_plan plan, char* buf, size_t buf_len);

const char* chfft_status_string(int status);

#ifdef __cplusplus
}
#endif

#endif /* CHFFT_HIP_H */
