// Length-specialised fused transform kernels for gfx950: the kernels built
// from the tiles of chfft_sfft.h and the registry the planner picks them from.
#include <algorithm>

#include "chfft_sfft.h"

namespace chfft {

template <typename T, int L, int TPT, int TPB, bool COL, int LD, int ST, int MODE, bool EPI, class RS, int AL,
          bool NT = false>
__global__ void __launch_bounds__(TPT * TPB) sfft_kernel(KParams P) {
  extern __shared__ __align__(16) unsigned char smem[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem);
  const int tile = (COL || P.colmajor_st) ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  sfft_tile<T, L, TPT, TPB, COL, LD, ST, MODE, EPI, RS, AL, NT>(P, P.src, P.dst, tile * TPB, P.scale, lds);
}

// Persistent four-step transform with length-specialised phases: one
// transform per workgroup iteration, columns (A) then rows (B) [then inverse
// columns (C) for Bluestein] through the workgroup's own scratch slot, which
// is written and re-read within the same workgroup (L2/MALL resident).
// Phase descriptors live in device memory (Q); src/dst advance by
// src_dist/dst_dist elements per transform.
template <typename T, class A, class B, class C>
__global__ void __launch_bounds__(A::THREADS, 4) sfs_kernel(const KParams* __restrict__ Q, const void* in, void* out,
                                                        double scale, int batch, long long src_dist,
                                                        long long dst_dist, void* scratch, long long slot_elems) {
  static_assert(A::THREADS == B::THREADS && A::THREADS == C::THREADS, "phases share the workgroup");
  extern __shared__ __align__(16) unsigned char smem[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem);
  cx<T>* slot = reinterpret_cast<cx<T>*>(scratch) + (long long)blockIdx.x * slot_elems;
  for (int tr = blockIdx.x; tr < batch; tr += gridDim.x) {
    const cx<T>* src = reinterpret_cast<const cx<T>*>(in) + (long long)tr * src_dist;
    cx<T>* dst = reinterpret_cast<cx<T>*>(out) + (long long)tr * dst_dist;
    const KParams& QA = Q[0];
    const KParams& QB = Q[1];
    constexpr bool three = C::PHASE;
    for (int g0 = 0; g0 < QA.G; g0 += A::TPB_) {
      A::template run<T>(QA, src, slot, g0, 1.0, lds);
      __syncthreads();
    }
    for (int g0 = 0; g0 < QB.G; g0 += B::TPB_) {
      B::template run<T>(QB, slot, three ? (void*)slot : (void*)dst, g0, three ? 1.0 : scale, lds);
      __syncthreads();
    }
    if constexpr (three) {
      const KParams& QC = Q[2];
      for (int g0 = 0; g0 < QC.G; g0 += C::TPB_) {
        C::template run<T>(QC, slot, dst, g0, scale, lds);
        __syncthreads();
      }
    }
  }
}

// A phase of sfs_kernel: one sfft_tile configuration.
template <int L, int TPT, int TPB, bool COL, int LD, int ST, int MODE, bool EPI, class RS>
struct Phase {
  static constexpr int THREADS = TPT * TPB;
  static constexpr int TPB_ = TPB;
  static constexpr bool PHASE = true;
  template <typename T>
  __device__ __forceinline__ static void run(const KParams& P, const void* src, void* dst, int g0, double scale,
                                             cx<T>* lds) {
    sfft_tile<T, L, TPT, TPB, COL, LD, ST, MODE, EPI, RS>(P, src, dst, g0, scale, lds);
  }
};
template <int THREADS_>
struct NoPhase {
  static constexpr int THREADS = THREADS_;
  static constexpr int TPB_ = 1;
  static constexpr bool PHASE = false;
  template <typename T>
  __device__ __forceinline__ static void run(const KParams&, const void*, void*, int, double, cx<T>*) {}
};

// ---------------------------------------------------------------- registry
// Each entry: a kernel instantiation and the schedule the planner must build
// tables for (the radices, threads per transform, transforms per workgroup).
struct SfftEntry {
  int f64, L, tpt, tpb, col, ld, st, mode, epi;
  int nrad;
  int rad[8];
  void (*fn)(KParams);   // the kernel (paired-real operators: aligned buffers)
  void (*fn0)(KParams);  // paired-real operators on unaligned buffers (else fn)
  void (*fnt)(KParams);  // fn with non-temporal output stores (null: none), for launches that write the
                         // exec's final output
};

// Operators with paired-real fast paths get two kernels: one that assumes the
// alignment those paths need and one with per-element accesses only; the
// launch picks one from the buffers (see pair_aligned).
constexpr bool sf_paired(int ld, int st) {
  return ld == LD_DCT2 || ld == LD_MDCT || st == ST_DCT3 || st == ST_MDCT;
}
constexpr int sf_al(int ld, int st, bool aligned) { return sf_paired(ld, st) ? (aligned ? 1 : 0) : 2; }

#define SF_KERNEL(T, L, TPT, TPB, COL, LD, ST, MODE, EPI, AL, ...) \
  &sfft_kernel<T, L, TPT, TPB, COL, LD, ST, MODE, EPI, Rad<__VA_ARGS__>, AL>
#define SF_KERNEL_NT(T, L, TPT, TPB, COL, LD, ST, MODE, EPI, AL, ...) \
  &sfft_kernel<T, L, TPT, TPB, COL, LD, ST, MODE, EPI, Rad<__VA_ARGS__>, AL, true>

#define SF_ENTRY_(NTK, F64, T, L, TPT, TPB, COL, LD, ST, MODE, EPI, NR, ...)                                \
  {F64, L, TPT, TPB, COL, LD, ST, MODE, EPI, NR, {__VA_ARGS__},                                             \
   reinterpret_cast<void (*)(KParams)>(SF_KERNEL(T, L, TPT, TPB, COL, LD, ST, MODE, EPI, sf_al(LD, ST, true), \
                                                 __VA_ARGS__)),                                             \
   reinterpret_cast<void (*)(KParams)>(SF_KERNEL(T, L, TPT, TPB, COL, LD, ST, MODE, EPI, sf_al(LD, ST, false), \
                                                 __VA_ARGS__)),                                             \
   NTK(T, L, TPT, TPB, COL, LD, ST, MODE, EPI, sf_al(LD, ST, true), __VA_ARGS__)}
#define SF_NO_NT(...) nullptr
#define SF_WITH_NT(...) reinterpret_cast<void (*)(KParams)>(SF_KERNEL_NT(__VA_ARGS__))
#define SF_ENTRY(...) SF_ENTRY_(SF_NO_NT, __VA_ARGS__)
#define SF_ENTRY_NT(...) SF_ENTRY_(SF_WITH_NT, __VA_ARGS__)

#define SF_BOTH(L, TPT, TPB, TPB64, COL, LD, ST, MODE, EPI, NR, ...)                    \
  SF_ENTRY(0, float, L, TPT, TPB, COL, LD, ST, MODE, EPI, NR, __VA_ARGS__),              \
      SF_ENTRY(1, double, L, TPT, TPB64, COL, LD, ST, MODE, EPI, NR, __VA_ARGS__)
// the same with a non-temporal-output build of each kernel
#define SF_BOTH_NT(L, TPT, TPB, TPB64, COL, LD, ST, MODE, EPI, NR, ...)                 \
  SF_ENTRY_NT(0, float, L, TPT, TPB, COL, LD, ST, MODE, EPI, NR, __VA_ARGS__),           \
      SF_ENTRY_NT(1, double, L, TPT, TPB64, COL, LD, ST, MODE, EPI, NR, __VA_ARGS__)

static const SfftEntry kSfft[] = {
    // 4096: CFft1D rows, RFft1D N=8192 forward / backward
    SF_BOTH_NT(4096, 256, 1, 1, false, LD_PLAIN, ST_PLAIN, 0, false, 3, 16, 16, 16),
    SF_BOTH_NT(4096, 256, 1, 1, false, LD_PLAIN, ST_R2C, 0, true, 3, 16, 16, 16),
    SF_BOTH_NT(4096, 256, 1, 1, false, LD_R2C_BACK, ST_PLAIN, 0, false, 3, 16, 16, 16),
    // 512: Dct1D / Mdct1D N=1024 / 2048 inner transforms, and CFft1D
    SF_BOTH(512, 32, 8, 4, false, LD_PLAIN, ST_PLAIN, 0, false, 3, 16, 8, 4),
    SF_BOTH(512, 32, 8, 4, false, LD_DCT2, ST_DCT2, 0, true, 3, 16, 8, 4),
    SF_BOTH(512, 32, 8, 4, false, LD_DCT3, ST_DCT3, 0, true, 3, 16, 8, 4),
    SF_BOTH_NT(512, 32, 8, 4, false, LD_MDCT, ST_MDCT, 0, false, 3, 16, 8, 4),
    SF_BOTH(512, 32, 8, 4, false, LD_IMDCT, ST_IMDCT, 0, false, 3, 16, 8, 4),
    // 2048: CFft2D rows, CFft1D
    SF_BOTH(2048, 128, 2, 1, false, LD_PLAIN, ST_PLAIN, 0, false, 3, 16, 16, 8),
    // 2048 columns (CFft2D 2048 x 2048 second axis): 64-byte row pieces of
    // 8 (FP32) / 4 (FP64) columns, one 128 KB tile per CU
    SF_BOTH(2048, 128, 8, 4, true, LD_PLAIN, ST_PLAIN, 0, false, 3, 16, 16, 8),
    // 8192: fused Bluestein for N <= 4096 (N = 4093 and smaller primes), CFft1D
    SF_BOTH_NT(8192, 512, 1, 1, false, LD_PLAIN, ST_PLAIN, 1, false, 4, 16, 16, 8, 4),
    SF_BOTH(8192, 512, 1, 1, false, LD_PLAIN, ST_PLAIN, 0, false, 4, 16, 16, 8, 4),
    // 20160 = 2^6 3^2 5 7 (FP32): fused Bluestein for 8192 < N <= 10080 (N =
    // 10007), the whole padded sequence (20475 slots) in one CU's LDS; 768
    // threads (168 registers each; 960 threads spilled at 128), a multiple
    // of the spans 16 and 192 so that those passes' twiddles repeat across a
    // thread's butterflies (radix 15 first, whose stride-15 first-pass writes
    // spread over the LDS banks, measured 94 us against 87)
    SF_ENTRY_NT(0, float, 20160, 768, 1, false, LD_PLAIN, ST_PLAIN, 1, false, 4, 16, 12, 15, 7),
    // 43200 = 180 x 240 (rows of the column pass 128-byte aligned), two
    // passes each: 180 = 12 x 15, 240 = 16 x 15 (prime-factor radices)
    SF_BOTH(180, 16, 16, 8, true, LD_PLAIN, ST_TW, 0, false, 2, 12, 15),
    SF_BOTH(240, 16, 16, 8, false, LD_PLAIN, ST_PLAIN, 0, true, 2, 16, 15),
    // four-step Bluestein: chirp + columns, rows x spectrum + inverse rows,
    // inverse columns + de-chirp; M = 20480 = 128 x 160 (N = 10007),
    // M = 32768 = 128 x 256, M = 138240 = 360 x 384 (columns as radix
    // 8 x 15 x 3), M = 147456 = 384 x 384
    SF_BOTH(128, 16, 16, 16, true, LD_FS_CHIRP, ST_TW, 0, false, 2, 16, 8),
    SF_BOTH(160, 16, 8, 4, false, LD_PLAIN, ST_TW, 2, false, 2, 16, 10),
    SF_BOTH(256, 16, 4, 4, false, LD_PLAIN, ST_TW, 2, false, 2, 16, 16),
    SF_BOTH(128, 16, 16, 16, true, LD_PLAIN, ST_FS_DECHIRP, 0, false, 2, 16, 8),
    SF_BOTH(384, 32, 16, 8, true, LD_FS_CHIRP, ST_TW, 0, false, 3, 16, 8, 3),
    SF_BOTH(384, 32, 4, 4, false, LD_PLAIN, ST_TW, 2, false, 3, 16, 8, 3),
    SF_BOTH(384, 32, 16, 8, true, LD_PLAIN, ST_FS_DECHIRP, 0, false, 3, 16, 8, 3),
    // M = 131072 = 256 x 512 = 2N - 2 (N = 65537): columns of 256 as
    // radix 16 x 16, rows of 512 as 16 x 8 x 4
    SF_BOTH(256, 16, 16, 8, true, LD_FS_CHIRP, ST_TW, 0, false, 2, 16, 16),
    SF_BOTH(512, 32, 4, 4, false, LD_PLAIN, ST_TW, 2, false, 3, 16, 8, 4),
    SF_BOTH(256, 16, 16, 8, true, LD_PLAIN, ST_FS_DECHIRP, 0, false, 2, 16, 16),
    // (FP32 only: in FP64 M = 138240 measured 3 % slower than 147456)
    SF_ENTRY(0, float, 360, 32, 16, true, LD_FS_CHIRP, ST_TW, 0, false, 3, 8, 15, 3),
    SF_ENTRY(0, float, 360, 32, 16, true, LD_PLAIN, ST_FS_DECHIRP, 0, false, 3, 8, 15, 3),
};

// Persistent four-step instantiations.
struct SfsEntry {
  int f64, n1, n2, nph;
  int threads;
  int tpb[3];
  int Lp[3];
  size_t lds;
  const void* fn;
};

template <int L, int TPB, bool F64, class RS>
constexpr size_t tile_bytes() {
  return lds_need(L, TPB, F64 ? 16 : 8, rad_compact_off(RS{}, rad_count(RS{})));
}
constexpr size_t mx(size_t a, size_t b) { return a > b ? a : b; }

#define SFS2(F64, T, N1, N2, TPT, TPB, RA, RB)                                                                  \
  {F64, N1, N2, 2, TPT * TPB, {TPB, TPB, 1}, {lds_stride(N1, TPB), lds_stride(N2, TPB), 0},                    \
   mx(tile_bytes<N1, TPB, F64, RA>(), tile_bytes<N2, TPB, F64, RB>()),                                          \
   reinterpret_cast<const void*>(                                                                               \
       &sfs_kernel<T, Phase<N1, TPT, TPB, true, LD_PLAIN, ST_TW, 0, false, RA>,                                \
                   Phase<N2, TPT, TPB, false, LD_PLAIN, ST_PLAIN, 0, true, RB>, NoPhase<TPT * TPB>>)}

#define SFS3(F64, T, N1, N2, TPT, TPB, RA, RB)                                                                  \
  {F64, N1, N2, 3, TPT * TPB, {TPB, TPB, TPB}, {lds_stride(N1, TPB), lds_stride(N2, TPB), lds_stride(N1, TPB)}, \
   mx(tile_bytes<N1, TPB, F64, RA>(), tile_bytes<N2, TPB, F64, RB>()),                                          \
   reinterpret_cast<const void*>(                                                                               \
       &sfs_kernel<T, Phase<N1, TPT, TPB, true, LD_FS_CHIRP, ST_TW, 0, false, RA>,                             \
                   Phase<N2, TPT, TPB, false, LD_PLAIN, ST_TW, 2, false, RB>,                                  \
                   Phase<N1, TPT, TPB, true, LD_PLAIN, ST_FS_DECHIRP, 0, false, RA>>)}

static const SfsEntry kSfs[] = {
    SFS2(0, float, 180, 240, 16, 16, Rad<4 COMMA 5 COMMA 3 COMMA 3>, Rad<16 COMMA 5 COMMA 3>),
    SFS2(1, double, 180, 240, 16, 8, Rad<4 COMMA 5 COMMA 3 COMMA 3>, Rad<16 COMMA 5 COMMA 3>),
    SFS3(0, float, 128, 256, 16, 16, Rad<16 COMMA 8>, Rad<16 COMMA 16>),
    SFS3(1, double, 128, 256, 16, 8, Rad<16 COMMA 8>, Rad<16 COMMA 16>),
};

const SfsEntry* sfs_find(bool f64, int n1, int n2, int nph) {
  for (const SfsEntry& e : kSfs)
    if (e.f64 == (int)f64 && e.n1 == n1 && e.n2 == n2 && e.nph == nph) return &e;
  return nullptr;
}

void sfs_info(const SfsEntry* e, int* threads, int* tpb, int* Lp, size_t* lds) {
  *threads = e->threads;
  for (int i = 0; i < 3; i++) {
    tpb[i] = e->tpb[i];
    Lp[i] = e->Lp[i];
  }
  *lds = e->lds;
}

hipError_t sfs_launch(const SfsEntry* e, const KParams* Q, const void* in, void* out, double scale, int batch,
                      long long src_dist, long long dst_dist, void* scratch, long long slot_elems, dim3 grid,
                      hipStream_t st) {
  void* args[] = {&Q, &in, &out, &scale, &batch, &src_dist, &dst_dist, &scratch, &slot_elems};
  return hipLaunchKernel(e->fn, grid, dim3(e->threads), args, e->lds, st);
}

int sfs_occupancy(const SfsEntry* e) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, e->fn, e->threads, e->lds) != hipSuccess) return 1;
  return n > 0 ? n : 1;
}

const SfftEntry* sfft_find(bool f64, int L, int ld, int st, int mode, bool col) {
  for (const SfftEntry& e : kSfft)
    if (e.f64 == (int)f64 && e.L == L && e.ld == ld && e.st == st && e.mode == mode && e.col == (int)col)
      return &e;
  return nullptr;
}

// Smallest length-specialised fused Bluestein (plain load and store, row
// tiles) whose padded length M holds an N-point transform (M >= 2N - 1 and
// N <= M / 2, the half its first and last passes skip); 0 if none.
int sfft_fused_bluestein_m(bool f64, long long N) {
  int best = 0;
  for (const SfftEntry& e : kSfft)
    if (e.f64 == (int)f64 && e.mode == 1 && e.ld == LD_PLAIN && e.st == ST_PLAIN && !e.col && e.L >= 2 * N - 1 &&
        2 * N <= e.L && (best == 0 || e.L < best))
      best = e.L;
  return best;
}

// Whether some registered kernel of this length runs exactly this radix
// schedule.
// Radix schedule of the registered kernel for this length and operator pair
// (rad receives up to 8 radices); 0 if none is registered.
int sfft_schedule_of(bool f64, int L, int ld, int st, int mode, bool col, int* rad) {
  const SfftEntry* e = sfft_find(f64, L, ld, st, mode, col);
  if (!e) return 0;
  for (int q = 0; q < e->nrad; q++) rad[q] = e->rad[q];
  return e->nrad;
}

bool sfft_has_schedule(bool f64, int L, const int* rad, int nrad) {
  for (const SfftEntry& e : kSfft) {
    if (e.f64 != (int)f64 || e.L != L || e.nrad != nrad) continue;
    bool same = true;
    for (int q = 0; q < nrad; q++) same = same && e.rad[q] == rad[q];
    if (same) return true;
  }
  return false;
}

hipError_t sfft_launch(const SfftEntry* e, const KParams& P, dim3 grid, size_t lds, hipStream_t st) {
  void (*fn)(KParams) = e->fn;
  bool al = true;
  if (e->fn0 != e->fn) {
    const size_t esz = e->f64 ? 16 : 8;
    al = (!(e->ld == LD_DCT2 || e->ld == LD_MDCT) || pair_aligned<2>(P.src, P.in.bdist, esz)) &&
         (!(e->st == ST_DCT3 || e->st == ST_MDCT) || pair_aligned<2>(P.dst, P.out.bdist, esz));
    if (!al) fn = e->fn0;
  }
  // the non-temporal build runs the aligned fast paths only
  if (P.nt_out && e->fnt && al) fn = e->fnt;
  hipLaunchKernelGGL(fn, grid, dim3(e->tpt * e->tpb), lds, st, P);
  return hipGetLastError();
}

hipError_t sfft_set_lds_limits() {
  hipError_t r = hipSuccess;
  for (const SfsEntry& e : kSfs) {
    hipError_t x = hipFuncSetAttribute(e.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBytes);
    if (x != hipSuccess) r = x;
  }
  for (const SfftEntry& e : kSfft)
    for (void (*fn)(KParams) : {e.fn, e.fn0, e.fnt}) {
      if (!fn) continue;
      hipError_t x = hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)kLdsBytes);
      if (x != hipSuccess) r = x;
    }
  return r;
}

// LDS bytes an entry needs (two buffers when they fit, else one).
size_t sfft_lds_bytes(const SfftEntry* e) {
  int ntw = 0, p = 1;  // compact table entries
  for (int q = 0; q < e->nrad; q++) {
    ntw += tw_compact_entries(e->rad[q], p);
    p *= e->rad[q];
  }
  return lds_need(e->L, e->tpb, e->f64 ? 16 : 8, ntw, mdct_tab_elems(e->L, e->ld, e->f64 ? 16 : 8));
}

int sfft_lds_stride(const SfftEntry* e) { return lds_stride(e->L, e->tpb); }

}  // namespace chfft
