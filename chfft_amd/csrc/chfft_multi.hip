// The uneven multi-length batch in one persistent launch (the batching form
// of the engine; each transform is CFft1D of its own length, cfft1d.rs:208-230).
// Work units run through the generic runtime-radix tile code of
// chfft_generic.h, or, for the tile shapes that carry most of a typical
// batch's points, through the register-resident length-specialised tiles of
// chfft_sfft.h (the descriptor's `spec`).
#include <hip/hip_runtime.h>

#include "chfft_generic.h"
#include "chfft_sfft.h"

namespace chfft {

// Register-resident tile paths of the FP32 multi-length kernel (1024-thread
// workgroups: TPT * TPB = 1024), X(id, L, TPT, TPB, col, LD, ST, MODE, radices):
//   rows of four-step Bluestein items: spectrum product and inverse rows,
//   two-level twiddle on the store (MODE 2, ST_TW2);
//   columns of four-step Bluestein items: chirp on the load, two-level
//   twiddle on the store; inverse columns with the de-chirp on the store;
//   columns of smooth four-step items (plain load, two-level twiddle) and
//   their rows (plain passes, transposed store through the LDS epilogue:
//   the only plain row store of this kernel, so ST_PLAIN rows take it).
// Lengths: the row and column lengths 2^a {1, 3, 5, 15} that the planner's
// splits of M in (16384, 2^21] produce most often (the 10,000-length uneven
// workload: 12.56 -> 9.95 ms on MI355X with the first 22 of these), and the
// 16 column and 16 row lengths that carry most points of its smooth items
// (66 % and 73 % of them); Bluestein columns of 960 and 1024 and rows of
// 1280 and 1536 (M up to 2^21).
#define CHFFT_MULTI_SPECS(X)                                                         \
  X(1, 1024, 64, 16, false, LD_PLAIN, ST_TW2, 2, 16, 16, 4)                          \
  X(2, 768, 64, 16, false, LD_PLAIN, ST_TW2, 2, 16, 16, 3)                           \
  X(3, 640, 64, 16, false, LD_PLAIN, ST_TW2, 2, 16, 8, 5)                            \
  X(4, 512, 32, 32, false, LD_PLAIN, ST_TW2, 2, 16, 8, 4)                            \
  X(5, 480, 32, 32, false, LD_PLAIN, ST_TW2, 2, 8, 4, 5, 3)                          \
  X(6, 384, 32, 32, false, LD_PLAIN, ST_TW2, 2, 16, 8, 3)                            \
  X(7, 320, 32, 32, false, LD_PLAIN, ST_TW2, 2, 16, 4, 5)                            \
  X(8, 256, 16, 64, false, LD_PLAIN, ST_TW2, 2, 16, 16)                              \
  X(9, 192, 16, 64, false, LD_PLAIN, ST_TW2, 2, 16, 4, 3)                            \
  X(10, 160, 16, 64, false, LD_PLAIN, ST_TW2, 2, 8, 4, 5)                            \
  X(11, 768, 64, 16, true, LD_FS_CHIRP, ST_TW2, 0, 16, 16, 3)                        \
  X(12, 640, 64, 16, true, LD_FS_CHIRP, ST_TW2, 0, 16, 8, 5)                         \
  X(13, 512, 32, 32, true, LD_FS_CHIRP, ST_TW2, 0, 16, 8, 4)                         \
  X(14, 480, 32, 32, true, LD_FS_CHIRP, ST_TW2, 0, 8, 4, 5, 3)                       \
  X(15, 384, 32, 32, true, LD_FS_CHIRP, ST_TW2, 0, 16, 8, 3)                         \
  X(16, 320, 32, 32, true, LD_FS_CHIRP, ST_TW2, 0, 16, 4, 5)                         \
  X(17, 256, 16, 64, true, LD_FS_CHIRP, ST_TW2, 0, 16, 16)                           \
  X(18, 240, 16, 64, true, LD_FS_CHIRP, ST_TW2, 0, 16, 5, 3)                         \
  X(19, 192, 16, 64, true, LD_FS_CHIRP, ST_TW2, 0, 16, 4, 3)                         \
  X(30, 160, 16, 64, true, LD_FS_CHIRP, ST_TW2, 0, 8, 4, 5)                          \
  X(31, 128, 16, 64, true, LD_FS_CHIRP, ST_TW2, 0, 16, 8)                            \
  X(21, 768, 64, 16, true, LD_PLAIN, ST_FS_DECHIRP, 0, 16, 16, 3)                    \
  X(22, 640, 64, 16, true, LD_PLAIN, ST_FS_DECHIRP, 0, 16, 8, 5)                     \
  X(23, 512, 32, 32, true, LD_PLAIN, ST_FS_DECHIRP, 0, 16, 8, 4)                     \
  X(24, 480, 32, 32, true, LD_PLAIN, ST_FS_DECHIRP, 0, 8, 4, 5, 3)                   \
  X(25, 384, 32, 32, true, LD_PLAIN, ST_FS_DECHIRP, 0, 16, 8, 3)                     \
  X(26, 320, 32, 32, true, LD_PLAIN, ST_FS_DECHIRP, 0, 16, 4, 5)                     \
  X(27, 256, 16, 64, true, LD_PLAIN, ST_FS_DECHIRP, 0, 16, 16)                       \
  X(28, 240, 16, 64, true, LD_PLAIN, ST_FS_DECHIRP, 0, 16, 5, 3)                     \
  X(29, 192, 16, 64, true, LD_PLAIN, ST_FS_DECHIRP, 0, 16, 4, 3)                     \
  X(32, 160, 16, 64, true, LD_PLAIN, ST_FS_DECHIRP, 0, 8, 4, 5)                      \
  X(33, 128, 16, 64, true, LD_PLAIN, ST_FS_DECHIRP, 0, 16, 8)                        \
  X(40, 125, 16, 64, true, LD_PLAIN, ST_TW2, 0, 5, 5, 5)                             \
  X(41, 150, 16, 64, true, LD_PLAIN, ST_TW2, 0, 10, 15)                              \
  X(42, 225, 16, 64, true, LD_PLAIN, ST_TW2, 0, 15, 15)                              \
  X(43, 243, 16, 64, true, LD_PLAIN, ST_TW2, 0, 3, 3, 3, 3, 3)                       \
  X(44, 250, 16, 64, true, LD_PLAIN, ST_TW2, 0, 10, 5, 5)                            \
  X(45, 270, 32, 32, true, LD_PLAIN, ST_TW2, 0, 2, 15, 3, 3)                         \
  X(46, 300, 32, 32, true, LD_PLAIN, ST_TW2, 0, 4, 15, 5)                            \
  X(47, 324, 32, 32, true, LD_PLAIN, ST_TW2, 0, 12, 3, 3, 3)                         \
  X(48, 360, 32, 32, true, LD_PLAIN, ST_TW2, 0, 8, 15, 3)                            \
  X(49, 375, 32, 32, true, LD_PLAIN, ST_TW2, 0, 15, 5, 5)                            \
  X(50, 400, 32, 32, true, LD_PLAIN, ST_TW2, 0, 16, 5, 5)                            \
  X(51, 405, 32, 32, true, LD_PLAIN, ST_TW2, 0, 15, 3, 3, 3)                         \
  X(52, 450, 32, 32, true, LD_PLAIN, ST_TW2, 0, 2, 15, 15)                           \
  X(53, 480, 32, 32, true, LD_PLAIN, ST_TW2, 0, 8, 4, 15)                            \
  X(54, 486, 32, 32, true, LD_PLAIN, ST_TW2, 0, 2, 3, 3, 3, 3, 3)                    \
  X(55, 500, 32, 32, true, LD_PLAIN, ST_TW2, 0, 4, 5, 5, 5)                          \
  X(60, 240, 16, 64, false, LD_PLAIN, ST_PLAIN, 0, 16, 15)                           \
  X(61, 256, 16, 64, false, LD_PLAIN, ST_PLAIN, 0, 16, 16)                           \
  X(62, 288, 32, 32, false, LD_PLAIN, ST_PLAIN, 0, 8, 4, 3, 3)                       \
  X(63, 320, 32, 32, false, LD_PLAIN, ST_PLAIN, 0, 16, 4, 5)                         \
  X(64, 375, 32, 32, false, LD_PLAIN, ST_PLAIN, 0, 15, 5, 5)                         \
  X(65, 384, 32, 32, false, LD_PLAIN, ST_PLAIN, 0, 16, 8, 3)                         \
  X(66, 400, 32, 32, false, LD_PLAIN, ST_PLAIN, 0, 16, 5, 5)                         \
  X(67, 432, 32, 32, false, LD_PLAIN, ST_PLAIN, 0, 16, 3, 3, 3)                      \
  X(68, 480, 32, 32, false, LD_PLAIN, ST_PLAIN, 0, 8, 4, 15)                         \
  X(69, 512, 32, 32, false, LD_PLAIN, ST_PLAIN, 0, 16, 8, 4)                         \
  X(70, 576, 64, 16, false, LD_PLAIN, ST_PLAIN, 0, 16, 12, 3)                        \
  X(71, 625, 64, 16, false, LD_PLAIN, ST_PLAIN, 0, 5, 5, 5, 5)                       \
  X(72, 640, 64, 16, false, LD_PLAIN, ST_PLAIN, 0, 16, 8, 5)                         \
  X(73, 720, 64, 16, false, LD_PLAIN, ST_PLAIN, 0, 16, 15, 3)                        \
  X(74, 768, 64, 16, false, LD_PLAIN, ST_PLAIN, 0, 16, 16, 3)                        \
  X(75, 800, 64, 16, false, LD_PLAIN, ST_PLAIN, 0, 16, 10, 5)                        \
  X(80, 1024, 64, 16, true, LD_FS_CHIRP, ST_TW2, 0, 16, 16, 4)                       \
  X(81, 960, 64, 16, true, LD_FS_CHIRP, ST_TW2, 0, 16, 4, 15)                        \
  X(82, 1024, 64, 16, true, LD_PLAIN, ST_FS_DECHIRP, 0, 16, 16, 4)                   \
  X(83, 960, 64, 16, true, LD_PLAIN, ST_FS_DECHIRP, 0, 16, 4, 15)                    \
  X(84, 1280, 128, 8, false, LD_PLAIN, ST_TW2, 2, 16, 16, 5)                         \
  X(85, 1536, 128, 8, false, LD_PLAIN, ST_TW2, 2, 16, 8, 4, 3)

// Each tile path is a function of its own (not inlined), so that the
// register allocation of one shape does not add to the others' or to the
// persistent loop's. Its arguments arrive in vector registers as generic
// pointers; the tile re-derives them (uniform_global, uniform_desc) so that
// the descriptor is read with scalar loads and the data and tables with
// global instructions rather than flat ones, which also count against the
// LDS wait counter and would make every LDS-only wait before one of the
// tile's barriers drain its global loads in flight.
__device__ __forceinline__ uintptr_t uniform_bits(uintptr_t v) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return ((uintptr_t)hi << 32) | lo;
}
// A workgroup-uniform pointer to global memory, as a global-address-space
// pointer.
template <typename T>
__device__ __forceinline__ T* uniform_global(T* p) {
  return as_global((T*)uniform_bits((uintptr_t)p));
}
// The descriptor at the workgroup-uniform address Qd (read-only during the
// launch), read through the constant address space, with its table
// pointers marked global.
__device__ __forceinline__ KParams uniform_desc(const KParams* Qd) {
  static_assert(sizeof(KParams) % 4 == 0, "the descriptor is copied in 32-bit words");
  typedef __attribute__((address_space(4))) const unsigned CW;
  const CW* w = (const CW*)uniform_bits((uintptr_t)Qd);
  KParams Q;
  unsigned* q = reinterpret_cast<unsigned*>(&Q);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(KParams) / 4); i++) q[i] = w[i];
  Q.seq.tw = uniform_global(Q.seq.tw);
  Q.chirp = uniform_global(Q.chirp);
  Q.kspec = uniform_global(Q.kspec);
  Q.tab_a = uniform_global(Q.tab_a);
  Q.tab_b = uniform_global(Q.tab_b);
  Q.tw_rec = uniform_global(Q.tw_rec);
  return Q;
}

#define CHFFT_SPEC_FN(ID, L, TPT, TPB, COL, LD, ST, MODE, ...)                                                 \
  __device__ __attribute__((noinline)) void spec_tile_##ID(const KParams* Qd, const void* src, void* dst,      \
                                                           int g0, double scale, cx<float>* lds) {             \
    const KParams Q = uniform_desc(Qd);                                                                      \
    sfft_tile<float, L, TPT, TPB, COL, LD, ST, MODE, (ST == ST_PLAIN && !COL), Rad<__VA_ARGS__>>(            \
        Q, uniform_global(src), uniform_global(dst), __builtin_amdgcn_readfirstlane(g0), scale, lds);           \
  }
CHFFT_MULTI_SPECS(CHFFT_SPEC_FN)
#undef CHFFT_SPEC_FN

template <typename T>
__device__ __forceinline__ bool spec_tile(int spec, const KParams* Qd, const void* src, void* dst, int g0,
                                          double scale, cx<T>* lds) {
  if constexpr (sizeof(T) == 4) {
    switch (spec) {
#define CHFFT_SPEC_CASE(ID, ...)                    \
  case ID:                                         \
    spec_tile_##ID(Qd, src, dst, g0, scale, lds);  \
    return true;
      CHFFT_MULTI_SPECS(CHFFT_SPEC_CASE)
#undef CHFFT_SPEC_CASE
      default:
        break;
    }
  }
  return false;
}

// Grid-wide barrier of the multi-length launch. The launch is cooperative
// where the stream allows it (every workgroup resident, guaranteed by the
// runtime) and otherwise sized to the occupancy, so a workgroup that arrives
// waits only for the others to finish their share of the phase. The wait is
// bounded: after `ticks` of the 100 MHz real-time counter (4 s by default)
// the workgroup records the failure in sync[1], which stays set until the
// host reads it (chfft_plan_check), and leaves the kernel, so no wave can
// wait forever. `extra` raises the arrival target (0 in use; the host's test
// mode sets it to make every wait time out). The agent-scope fences publish
// each workgroup's scratch writes to every XCD before it arrives, and make
// the others' writes visible after.
__device__ __forceinline__ bool grid_barrier(unsigned* sync, unsigned target, unsigned long long ticks) {
  __shared__ int ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    __hip_atomic_fetch_add(&sync[0], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    int good = 1;
    const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
    // relaxed polls: an acquire load would invalidate this XCD's L2 on every
    // poll, under the workgroups still working on the phase; the fence after
    // the wait orders what follows
    while (__hip_atomic_load(&sync[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(8);
      if (__builtin_amdgcn_s_memrealtime() - t0 > ticks) {
        good = 0;
        __hip_atomic_fetch_or(&sync[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
    }
    ok = good;
  }
  __syncthreads();
  __threadfence();
  return ok != 0;
}

// Clears the barrier counter before each multi-length launch (a kernel rather
// than a memset, so that the pair can be captured in a graph and replayed).
// The failure flag sync[1] is left alone: the host clears it when it reads it.
__global__ void multi_sync_reset(unsigned* sync) {
  if (threadIdx.x == 0) sync[0] = 0u;
}

// Uneven multi-length batch in one persistent launch, in up to three phases
// separated by grid-wide barriers: phase 0 runs the column passes of the
// four-step items (chirp + columns for Bluestein), phase 1 their row passes
// (spectrum product and inverse rows for Bluestein), phase 2 the inverse
// columns + de-chirp of the Bluestein items; whole fused items (in LDS) fill
// whichever phase the planner balanced them into. Within a phase, workgroup w
// takes units w, w + grid, ... (the planner sorts each phase by decreasing
// cost), so a long transform is spread over many workgroups instead of
// bounding the launch. Every loop bound is a unit count and every barrier
// wait is bounded, so every wave exits.
template <typename T>
__global__ void __launch_bounds__(sizeof(T) == 8 ? kMultiThreadsF64 : kMultiThreadsF32)
    __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? 2 : 4)))
    multi_kernel(const MultiItem* __restrict__ items, const MultiUnit* __restrict__ units, int end0, int end1,
                 int end2, const KParams* __restrict__ descs, const void* in, void* out, int scale_kind,
                 void* scratch, unsigned* sync, unsigned long long ticks, unsigned extra) {
  extern __shared__ __align__(16) unsigned char smem[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem);
  const int ends[kMultiPhases] = {end0, end1, end2};
  const int nph = end2 > end1 ? 3 : end1 > end0 ? 2 : 1;
  int u0 = 0;
  for (int ph = 0; ph < nph; ph++) {
    for (int u = u0 + blockIdx.x; u < ends[ph]; u += gridDim.x) {
      const MultiUnit U = units[u];
      const MultiItem m = items[U.item];
      const cx<T>* src = reinterpret_cast<const cx<T>*>(in) + m.off;
      cx<T>* dst = reinterpret_cast<cx<T>*>(out) + m.off;
      cx<T>* slot = reinterpret_cast<cx<T>*>(scratch) + m.soff;
      const double n = (double)m.n;
      const double scale = scale_kind == 0 ? 1.0 : scale_kind == 1 ? 1.0 / sqrt(n) : 1.0 / n;
      if (m.kind == 0) {
        block_transform<T, LD_PLAIN, ST_PLAIN, true>(descs[m.desc], src, dst, lds, 0, 1, scale);
        __syncthreads();
      } else if (m.kind == 3) {
        // length 1: returned unchanged (cfft1d.rs:213-214)
        if (threadIdx.x == 0) dst[0] = src[0];
      } else {
        const KParams& Q = descs[m.desc + U.ph];
        // phase buffers: columns read the input, the last phase writes the
        // output, the rest stay in the item's scratch slot
        const bool last = (m.kind == 1 && U.ph == 1) || U.ph == 2;
        const void* psrc = U.ph == 0 ? (const void*)src : (const void*)slot;
        void* pdst = last ? (void*)dst : (void*)slot;
        const double psc = last ? scale : 1.0;
        for (int g = U.g0; g < U.g1; g += Q.tpb) {
          const int nt = min(Q.tpb, U.g1 - g);
          if (Q.spec && spec_tile<T>(Q.spec, &Q, psrc, pdst, g, psc, lds)) {
            // register-resident tile of a registered shape
          } else if (m.kind == 1) {
            if (U.ph == 0)
              block_transform<T, LD_PLAIN, ST_TW, true>(Q, src, slot, lds, g, nt, 1.0);
            else
              block_transform<T, LD_PLAIN, ST_PLAIN, true>(Q, slot, dst, lds, g, nt, scale);
          } else {
            if (U.ph == 0)
              block_transform<T, LD_FS_CHIRP, ST_TW, true>(Q, src, slot, lds, g, nt, 1.0);
            else if (U.ph == 1)
              block_transform<T, LD_PLAIN, ST_TW, true>(Q, slot, slot, lds, g, nt, 1.0);
            else
              block_transform<T, LD_PLAIN, ST_FS_DECHIRP, true>(Q, slot, dst, lds, g, nt, scale);
          }
          __syncthreads();  // LDS reuse by the next tile
        }
      }
    }
    u0 = ends[ph];
    if (ph + 1 < nph && !grid_barrier(sync, (unsigned)(ph + 1) * gridDim.x + extra, ticks)) return;
  }
}

// coop: launch cooperatively (the runtime then refuses a grid whose
// workgroups cannot all be resident at once); a stream under graph capture
// takes the ordinary launch.
hipError_t launch_multi(bool f64, const MultiItem* items, const MultiUnit* units, const int* ends,
                        const KParams* descs, const void* in, void* out, int scale_kind, void* scratch,
                        unsigned* sync, unsigned long long ticks, unsigned extra, bool coop, dim3 grid, size_t lds,
                        hipStream_t st) {
  hipLaunchKernelGGL(multi_sync_reset, dim3(1), dim3(64), 0, st, sync);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const dim3 block(f64 ? kMultiThreadsF64 : kMultiThreadsF32);
  const void* fn = f64 ? reinterpret_cast<const void*>(&multi_kernel<double>)
                       : reinterpret_cast<const void*>(&multi_kernel<float>);
  int e0 = ends[0], e1 = ends[1], e2 = ends[2];
  void* args[] = {&items, &units, &e0, &e1, &e2, &descs, &in, &out, &scale_kind, &scratch, &sync, &ticks, &extra};
  if (coop) return hipLaunchCooperativeKernel(fn, grid, block, args, (unsigned)lds, st);
  return hipLaunchKernel(fn, grid, block, args, lds, st);
}

int multi_occupancy(bool f64, size_t lds) {
  int n = 0;
  hipError_t e = f64 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, multi_kernel<double>, kMultiThreadsF64, lds)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, multi_kernel<float>, kMultiThreadsF32, lds);
  return (e == hipSuccess && n > 0) ? n : 1;
}

hipError_t multi_set_lds_limits() {
  hipError_t e = hipSuccess;
  const void* ms[2] = {reinterpret_cast<const void*>(&multi_kernel<float>),
                       reinterpret_cast<const void*>(&multi_kernel<double>)};
  for (const void* f : ms) {  // these keep a few bytes of static LDS (the barrier flag)
    hipError_t r = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBytes - 64);
    if (r != hipSuccess) e = r;
  }
  return e;
}

// Host view of the register-resident tile paths, for the planner.
struct MultiSpec {
  int id, L, tpt, tpb, col, ld, st, mode;
  int nrad;
  int rad[6];
  size_t lds;
};

template <int... RS>
constexpr int spec_ntw(Rad<RS...> r) {
  return rad_compact_off(r, rad_count(r));
}

#define CHFFT_SPEC_ROW(ID, L, TPT, TPB, COL, LD, ST, MODE, ...)                                        \
  {ID, L, TPT, TPB, COL, LD, ST, MODE, rad_count(Rad<__VA_ARGS__>{}), {__VA_ARGS__},                  \
   lds_need(L, TPB, 8, spec_ntw(Rad<__VA_ARGS__>{}))},
static const MultiSpec kMultiSpecs[] = {CHFFT_MULTI_SPECS(CHFFT_SPEC_ROW)};
#undef CHFFT_SPEC_ROW

// The FP32 tile path for a phase of length L with these operators, or null.
const MultiSpec* multi_spec_find(int L, int ld, int st, int mode, bool col) {
  for (const MultiSpec& s : kMultiSpecs)
    if (s.L == L && s.ld == ld && s.st == st && s.mode == mode && s.col == (int)col) return &s;
  return nullptr;
}

void multi_spec_info(const MultiSpec* s, int* id, int* tpb, int* nrad, const int** rad, size_t* lds) {
  *id = s->id;
  *tpb = s->tpb;
  *nrad = s->nrad;
  *rad = s->rad;
  *lds = s->lds;
}

}  // namespace chfft
