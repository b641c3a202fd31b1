// Launch descriptors shared by the host planner (chfft_plan.hip) and the
// device kernels (chfft_kernels.hip). Plain data only: a KParams is passed by
// value as the kernel argument.
#pragma once
#include <stddef.h>
#include <stdint.h>

namespace chfft {

constexpr int kMaxPass = 24;
// Register-staged LDS passes: each thread holds at most kEmax complex values of
// its butterflies between the LDS read and the LDS write of one pass.
constexpr int kEmax = 16;
// Workgroup size limits of the fused kernel (FP32 / FP64).
constexpr int kMaxThreadsF32 = 1024;
constexpr int kMaxThreadsF64 = 512;
// ... and of the persistent four-step kernel (more live state per thread)
constexpr int kMaxThreadsFsF32 = 512;
constexpr int kMaxThreadsFsF64 = 256;
// ... and of the multi-length persistent kernel (FP32: the 10,000-length
// uneven batch measured 13.5 ms at 512 threads, 12.5 ms at 1024 and 17.3 ms
// at 256 on MI355X; with 1024, lengths up to 16384 run fused in one
// workgroup)
constexpr int kMultiThreadsF32 = 1024;
constexpr int kMultiThreadsF64 = 256;
// LDS per workgroup on gfx950.
constexpr size_t kLdsBytes = 160 * 1024;

// Compact twiddle tables of the length-specialised kernels. Every pass table
// [j-1][k] = W_{pR}^(j k) (span p, radix R) that would exceed kTw2Min entries
// is kept in a two-level form, A[j-1][k mod 16] = W_{pR}^(j (k mod 16)) and
// B[j-1][k / 16] = W_{pR}^(16 j (k / 16)), whose product is the twiddle; the
// other passes keep their full table. The compact tables of all passes follow
// the full tables in the plan's twiddle buffer, pass after pass.
constexpr int kTw2Min = 1024;
constexpr bool tw_two_level(int R, int p) { return p % 16 == 0 && (R - 1) * p > kTw2Min; }
constexpr int tw_compact_entries(int R, int p) {
  return p <= 1 ? 0 : tw_two_level(R, p) ? (R - 1) * (16 + p / 16) : (R - 1) * p;
}

// One radix-R Stockham pass over a length-L sequence resident in LDS.
// p = product of the radices already applied; tws = L / (p * R) is the step of
// the twiddle index j*k*tws into the W_L table.
struct PassDesc {
  int radix;
  int p;
  int tws;
  unsigned pmagic;  // ceil(2^32 / p): i / p == umulhi(i, pmagic) for i < 2^16
  int twoff;        // offset of this pass's table [j-1][k] = W_{pR}^(j*k) in tw
};

struct SeqDesc {
  int L;
  int npass;
  PassDesc pass[kMaxPass];
  const void* tw;  // per-pass twiddle tables, in the kernel's precision
};

// Element idx of transform g lives at
//   (g / ncols) * bdist + (g % ncols) * cstride + idx * estride
// in units of the element type (complex for complex data, scalar for real).
struct IoDesc {
  long long bdist;
  long long cstride;
  long long estride;
  int ncols;
};

// Load-stage operators (what the first LDS image is made from).
enum LoadOp : int {
  LD_PLAIN = 0,     // complex element (also real pairs viewed as complex)
  LD_FS_CHIRP = 1,  // Bluestein column pass: chirp multiply + zero padding
  LD_R2C_BACK = 2,  // RFft1D backward pre-pass (Hermitian un-split)
  LD_DCT2 = 3,      // DCT-II even/odd packing
  LD_DCT3 = 4,      // DCT-III pre-twiddle + split
  LD_MDCT = 5,      // MDCT window + fold + pre-twiddle
  LD_IMDCT = 6,     // IMDCT pre-twiddle
};

// Store-stage operators (what is written from the final LDS image).
enum StoreOp : int {
  ST_PLAIN = 0,
  ST_TW = 1,           // four-step inter-pass twiddle W_N^(c*k)
  ST_FS_DECHIRP = 2,   // Bluestein last column pass: de-chirp, crop to N
  ST_R2C = 3,          // RFft1D forward split
  ST_DCT2 = 4,
  ST_DCT3 = 5,
  ST_MDCT = 6,
  ST_IMDCT = 7,
  ST_TW2 = 8,          // four-step twiddle from two-level tables (twsh), length-specialised tiles
};

struct KParams {
  const void* src;
  void* dst;
  int G;       // transforms in this launch
  int tpb;     // transforms per workgroup
  int Lfft;    // FFT length held in LDS (M for a fused Bluestein)
  int Lp;      // LDS stride between transforms (>= Lfft)
  int L;       // logical inner length (== Lfft unless fused Bluestein)
  int colmajor_ld;  // 1: consecutive threads take consecutive transforms on load
  int colmajor_st;  // same for the store stage
  int conj_in;      // conjugate after the load operator (inverse via conj trick)
  int conj_out;     // conjugate before the store operator
  int has_mid;      // Bluestein spectrum product + second pass sequence
  int has_chirp;    // fused Bluestein chirp (load) and de-chirp (before store)
  SeqDesc seq;
  IoDesc in, out;
  double scale;     // user scaling, applied by the store operator
  double ortho;     // DCT element-0 scaler
  long long N;      // logical length the operator works on
  long long twN;    // four-step twiddle: size of the W^(c*k) table
  int twk, twc;     // four-step twiddle index = k * twk + c * twc
  int twsh;         // > 0: two-level twiddle W^m = tab_a[m >> twsh] * tab_b[m & (2^twsh - 1)], m = c * k
  long long kc, ke; // spectrum index = c * kc + idx * ke
  const void* chirp;
  const void* kspec;
  const void* tab_a;
  const void* tab_b;
  // four-step twiddles by recurrence (length-specialised ST_TW stores): FP64
  // table [r][c] of W^(c r) for r < K0 (the kernel's base rows, K0 = its
  // last-pass butterflies per thread x threads per transform) and, in row K0,
  // the step W^(c p) (p: the last pass's span); null: the full table tab_a.
  // For ST_TW2 stores (the multi-length kernel's tiles): FP64 two-level tables
  // [W^i, i < 2^twsh][W^(i 2^twsh)] from which the tile forms each base and
  // step; null: the FP32 two-level tables tab_a, tab_b, twice per element
  const void* tw_rec;
  int spec;         // multi-length kernel: register-resident tile path of this descriptor (0: generic)
  int nt_out;       // 1: the launch writes the exec's final output (no later launch reads it): a registered
                    // kernel with a non-temporal-output build runs that build
};

// One item of a multi-length batch (one transform). kind 0: fused (descs[desc]);
// 1: four-step (descs[desc], descs[desc+1]); 2: four-step Bluestein (three
// descriptors); 3: length 1 (copy).
struct MultiItem {
  long long off;   // element offset of the transform in the in/out buffers
  long long soff;  // element offset of its four-step intermediate in the scratch buffer
  int n;
  int kind;
  int desc;
  int pad;
};

// A unit of work of the multi-length launch: a whole fused item, or tiles
// [g0, g1) of phase ph of a four-step item. Units are grouped by the launch's
// phases; a grid-wide barrier separates the phases.
struct MultiUnit {
  int item;
  int ph;
  int g0, g1;
};
constexpr int kMultiPhases = 3;
// Bound of each grid-wide barrier wait of the multi-length launch, in ticks
// of the 100 MHz real-time counter (4 s).
constexpr unsigned long long kBarrierTicks = 400000000ull;

// Device table generation job (chfft_tables.hip).
struct RootJob {
  long long off;
  long long count;
  long long den;
  long long p;
  int kind;
  int sh;
};

}  // namespace chfft
