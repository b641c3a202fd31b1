// CDNA4 (gfx950) transform kernels.
//
// One kernel template covers every fused transform: a workgroup owns `tpb`
// transforms whose working sequence (length Lfft) is resident in LDS.
//   1. load stage   global -> LDS, through a load operator (plain, chirp,
//                   RFft/DCT/MDCT pre-pass), optional conjugation
//   2. passes       radix-2/3/4/5/7/8/16 Stockham autosort passes LDS -> LDS,
//                   inputs staged in registers between two barriers
//   3. (Bluestein)  spectrum product, conjugate, second pass sequence,
//                   de-chirp
//   4. store stage  LDS -> global, through a store operator (plain, four-step
//                   twiddle, de-chirp, RFft/DCT/MDCT post-pass)
// Inverse transforms run the same forward passes between two conjugations.
//
// Reference correspondence (chfft 0.3.4): the passes replace the digit-reversal
// gather and DIT stages of mixed_radix.rs:24-368 and chirpz.rs:27-85; the load
// and store operators replace the pre/post loops of chirpz.rs:97-145,
// rfft1d.rs:128-186, dct1d.rs:245-287 / 367-407 and mdct1d.rs:212-288.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "chfft_generic.h"

namespace chfft {

// FP64 workgroups are capped at 512 threads so that a thread may use up to 256
// VGPRs (16 staged double-complex values without spilling).
template <typename T, int LD, int ST>
__global__ void __launch_bounds__(sizeof(T) == 8 ? 512 : 1024) fused_fft_kernel(KParams P) {
  extern __shared__ __align__(16) unsigned char smem[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem);
  const int g0 = blockIdx.x * P.tpb;
  const int nt = min(P.tpb, P.G - g0);
  block_transform<T, LD, ST>(P, P.src, P.dst, lds, g0, nt, P.scale);
}

// Four-step transform of a whole length, one transform per workgroup
// iteration: the phases (columns + twiddle, rows [+ spectrum product +
// inverse rows + twiddle, inverse columns + de-chirp]) run back to back in the
// same workgroup through its own scratch slot, so the intermediate is written
// and re-read within a few microseconds (L2/MALL resident) instead of
// round-tripping through HBM between two kernel launches. Phase KParams
// describe a single transform; src/dst advance by src_dist/dst_dist elements
// per transform.
template <typename T, int LD, int ST>
__device__ void run_phase(const KParams& P, const void* src, void* dst, cx<T>* lds, double scale) {
  for (int g0 = 0; g0 < P.G; g0 += P.tpb) {
    const int nt = min(P.tpb, P.G - g0);
    block_transform<T, LD, ST>(P, src, dst, lds, g0, nt, scale);
    __syncthreads();  // LDS reuse by the next tile; scratch visibility for the next phase
  }
}

// Q points to the phase descriptors in device memory (uploaded once per plan);
// scale applies to the last phase.
template <typename T, int NPH>
__global__ void __launch_bounds__(sizeof(T) == 8 ? kMaxThreadsFsF64 : kMaxThreadsFsF32)
    fourstep_kernel(const KParams* __restrict__ Q, const void* in, void* out, double scale, int batch,
                    long long src_dist, long long dst_dist, void* scratch, long long slot_elems) {
  extern __shared__ __align__(16) unsigned char smem[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem);
  cx<T>* slot = reinterpret_cast<cx<T>*>(scratch) + (long long)blockIdx.x * slot_elems;
  for (int tr = blockIdx.x; tr < batch; tr += gridDim.x) {
    const cx<T>* src = reinterpret_cast<const cx<T>*>(in) + (long long)tr * src_dist;
    cx<T>* dst = reinterpret_cast<cx<T>*>(out) + (long long)tr * dst_dist;
    if constexpr (NPH == 2) {
      run_phase<T, LD_PLAIN, ST_TW>(Q[0], src, slot, lds, 1.0);
      run_phase<T, LD_PLAIN, ST_PLAIN>(Q[1], slot, dst, lds, scale);
    } else {
      run_phase<T, LD_FS_CHIRP, ST_TW>(Q[0], src, slot, lds, 1.0);
      run_phase<T, LD_PLAIN, ST_TW>(Q[1], slot, slot, lds, 1.0);
      run_phase<T, LD_PLAIN, ST_FS_DECHIRP>(Q[2], slot, dst, lds, scale);
    }
  }
}

// Plain copy used for length-1 transforms (cfft1d.rs:213-214 returns the input
// unchanged, whatever the scaling).
template <typename T>
__global__ void copy_kernel(const T* __restrict__ src, T* __restrict__ dst, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}

// ------------------------------------------------------------------ long-length operator passes
// RFft1D / Dct1D / Mdct1D whose half (quarter) length exceeds one
// workgroup's LDS run their inner complex transform through the four-step
// launches of the complex engine; the operators then become passes of their
// own over a work buffer of P.L complex elements per transform:
//   pre   work[g][n] = load operator element n of transform g (LD_*),
//   post  the store operator over work[g] read from global memory (ST_*).
// The grid's y dimension walks the transforms (grid-stride), x the elements.
template <typename T, int LD>
__global__ void __launch_bounds__(256) pre_kernel(KParams P) {
  cx<T>* work = reinterpret_cast<cx<T>*>(P.dst);
  for (int g = blockIdx.y; g < P.G; g += gridDim.y) {
    const Loader<T, LD> ld = make_loader<T, LD>(P, P.src, g);
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < P.L; n += gridDim.x * blockDim.x) {
      cx<T> v = load_elem<T, LD>(P, ld, n);
      if (P.conj_in) v = cj(v);
      work[(long long)g * P.L + n] = v;
    }
  }
}

template <typename T, int ST>
__global__ void __launch_bounds__(256) post_kernel(KParams P) {
  const cx<T>* work = reinterpret_cast<const cx<T>*>(P.src);
  const bool cjo = P.conj_out != 0;
  const T s = (T)P.scale;
  for (int g = blockIdx.y; g < P.G; g += gridDim.y) {
    const cx<T>* row = work + (long long)g * P.L;
    post_ops<T, ST>(P, [&](int k) {
      const cx<T> z = row[k];
      return cjo ? cj(z) : z;
    }, P.dst, (long long)g * P.out.bdist, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x, s);
  }
}

// Batched out-of-place transpose of complex R x C matrices (row-major,
// matrix b at element b * R * C) through 32 x 32 LDS tiles: the column
// pass of a 2-D transform whose column length exceeds one workgroup's LDS
// runs as rows of the transposed matrix.
template <typename T>
__global__ void __launch_bounds__(256) transpose_kernel(const cx<T>* __restrict__ src, cx<T>* __restrict__ dst,
                                                        int R, int C, int batch) {
  __shared__ cx<T> tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
  const int tilesC = (C + 31) / 32, tilesR = (R + 31) / 32;
  const long long per = (long long)tilesC * tilesR;
  for (long long t = blockIdx.x; t < per * batch; t += gridDim.x) {
    const int b = (int)(t / per);
    const int rem = (int)(t - (long long)b * per);
    const int r0 = (rem / tilesC) * 32, c0 = (rem % tilesC) * 32;
    const cx<T>* s = src + (long long)b * R * C;
    cx<T>* d = dst + (long long)b * R * C;
    for (int i = ty; i < 32; i += 8)
      if (r0 + i < R && c0 + tx < C) tile[i][tx] = s[(long long)(r0 + i) * C + c0 + tx];
    __syncthreads();
    for (int i = ty; i < 32; i += 8)
      if (c0 + i < C && r0 + tx < R) d[(long long)(c0 + i) * R + r0 + tx] = tile[tx][i];
    __syncthreads();
  }
}

// Overlap-add of `frames` windowed IMDCT frames of n samples placed `hop`
// apart (the caller-side step after Mdct1D::backward, mdct1d.rs:244-291):
// out[t] = sum over frames f with 0 <= t - f * hop < n of frames[f][t - f hop],
// summed in increasing f, for t < (frames - 1) * hop + n.
template <typename T>
__global__ void __launch_bounds__(256) ola_kernel(const T* __restrict__ fr, T* __restrict__ out, long long n,
                                                  long long hop, long long frames) {
  const long long total = (frames - 1) * hop + n;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    long long f = t >= n ? (t - n) / hop + 1 : 0;  // first frame reaching t
    const long long f1 = min(t / hop, frames - 1);
    T acc = T(0);
    for (; f <= f1; f++) acc += fr[f * n + (t - f * hop)];
    out[t] = acc;
  }
}

template <typename T>
static hipError_t launch_pre_t(int ld, const KParams& P, hipStream_t st) {
  const dim3 grid((unsigned)std::min((P.L + 255) / 256, 1024), (unsigned)std::min(P.G, 65535));
  switch (ld) {
    case LD_PLAIN: hipLaunchKernelGGL((pre_kernel<T, LD_PLAIN>), grid, dim3(256), 0, st, P); break;
    case LD_R2C_BACK: hipLaunchKernelGGL((pre_kernel<T, LD_R2C_BACK>), grid, dim3(256), 0, st, P); break;
    case LD_DCT2: hipLaunchKernelGGL((pre_kernel<T, LD_DCT2>), grid, dim3(256), 0, st, P); break;
    case LD_DCT3: hipLaunchKernelGGL((pre_kernel<T, LD_DCT3>), grid, dim3(256), 0, st, P); break;
    case LD_MDCT: hipLaunchKernelGGL((pre_kernel<T, LD_MDCT>), grid, dim3(256), 0, st, P); break;
    case LD_IMDCT: hipLaunchKernelGGL((pre_kernel<T, LD_IMDCT>), grid, dim3(256), 0, st, P); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_post_t(int st_op, const KParams& P, hipStream_t st) {
  // every operator loops over at most L / 2 + 1 bins
  const dim3 grid((unsigned)std::min((P.L / 2 + 256) / 256, 1024), (unsigned)std::min(P.G, 65535));
  switch (st_op) {
    case ST_R2C: hipLaunchKernelGGL((post_kernel<T, ST_R2C>), grid, dim3(256), 0, st, P); break;
    case ST_DCT2: hipLaunchKernelGGL((post_kernel<T, ST_DCT2>), grid, dim3(256), 0, st, P); break;
    case ST_DCT3: hipLaunchKernelGGL((post_kernel<T, ST_DCT3>), grid, dim3(256), 0, st, P); break;
    case ST_MDCT: hipLaunchKernelGGL((post_kernel<T, ST_MDCT>), grid, dim3(256), 0, st, P); break;
    case ST_IMDCT: hipLaunchKernelGGL((post_kernel<T, ST_IMDCT>), grid, dim3(256), 0, st, P); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_pre(bool f64, int ld, const KParams& P, hipStream_t st) {
  return f64 ? launch_pre_t<double>(ld, P, st) : launch_pre_t<float>(ld, P, st);
}

hipError_t launch_post(bool f64, int st_op, const KParams& P, hipStream_t st) {
  return f64 ? launch_post_t<double>(st_op, P, st) : launch_post_t<float>(st_op, P, st);
}

hipError_t launch_transpose(bool f64, const void* src, void* dst, int R, int C, int batch, hipStream_t st) {
  const long long tiles = (long long)((R + 31) / 32) * ((C + 31) / 32) * batch;
  const dim3 grid((unsigned)std::min<long long>(tiles, 65536));
  if (f64)
    hipLaunchKernelGGL(transpose_kernel<double>, grid, dim3(256), 0, st, (const cx<double>*)src, (cx<double>*)dst, R,
                       C, batch);
  else
    hipLaunchKernelGGL(transpose_kernel<float>, grid, dim3(256), 0, st, (const cx<float>*)src, (cx<float>*)dst, R, C,
                       batch);
  return hipGetLastError();
}

hipError_t launch_ola(bool f64, const void* frames, void* out, long long n, long long hop, long long count,
                      hipStream_t st) {
  const long long total = (count - 1) * hop + n;
  const dim3 grid((unsigned)std::min<long long>((total + 255) / 256, 65536));
  if (f64)
    hipLaunchKernelGGL(ola_kernel<double>, grid, dim3(256), 0, st, (const double*)frames, (double*)out, n, hop, count);
  else
    hipLaunchKernelGGL(ola_kernel<float>, grid, dim3(256), 0, st, (const float*)frames, (float*)out, n, hop, count);
  return hipGetLastError();
}

// ------------------------------------------------------------------ host-side launch table
using KernelFn = void (*)(KParams);

template <typename T, int LD, int ST>
static hipError_t launch_one(const KParams& P, dim3 grid, dim3 block, size_t lds, hipStream_t st) {
  hipLaunchKernelGGL((fused_fft_kernel<T, LD, ST>), grid, block, lds, st, P);
  return hipGetLastError();
}

#define CHFFT_COMBOS(X)         \
  X(LD_PLAIN, ST_PLAIN)         \
  X(LD_PLAIN, ST_TW)            \
  X(LD_FS_CHIRP, ST_TW)         \
  X(LD_PLAIN, ST_FS_DECHIRP)    \
  X(LD_PLAIN, ST_R2C)           \
  X(LD_R2C_BACK, ST_PLAIN)      \
  X(LD_DCT2, ST_DCT2)           \
  X(LD_DCT3, ST_DCT3)           \
  X(LD_MDCT, ST_MDCT)           \
  X(LD_IMDCT, ST_IMDCT)

template <typename T>
static hipError_t launch_fused_t(int ld, int st_op, const KParams& P, dim3 grid, dim3 block, size_t lds,
                                 hipStream_t st) {
#define CHFFT_CASE(A, B) \
  if (ld == A && st_op == B) return launch_one<T, A, B>(P, grid, block, lds, st);
  CHFFT_COMBOS(CHFFT_CASE)
#undef CHFFT_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_fused(bool f64, int ld, int st_op, const KParams& P, dim3 grid, dim3 block, size_t lds,
                        hipStream_t st) {
  return f64 ? launch_fused_t<double>(ld, st_op, P, grid, block, lds, st)
             : launch_fused_t<float>(ld, st_op, P, grid, block, lds, st);
}

hipError_t set_fused_lds_limits() {
  hipError_t e = hipSuccess;
#define CHFFT_ATTR(A, B)                                                                                       \
  {                                                                                                            \
    hipError_t r1 = hipFuncSetAttribute(reinterpret_cast<const void*>(&fused_fft_kernel<float, A, B>),         \
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBytes);           \
    hipError_t r2 = hipFuncSetAttribute(reinterpret_cast<const void*>(&fused_fft_kernel<double, A, B>),        \
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBytes);           \
    if (r1 != hipSuccess) e = r1;                                                                              \
    if (r2 != hipSuccess) e = r2;                                                                              \
  }
  CHFFT_COMBOS(CHFFT_ATTR)
#undef CHFFT_ATTR
  const void* fs[4] = {reinterpret_cast<const void*>(&fourstep_kernel<float, 2>),
                       reinterpret_cast<const void*>(&fourstep_kernel<float, 3>),
                       reinterpret_cast<const void*>(&fourstep_kernel<double, 2>),
                       reinterpret_cast<const void*>(&fourstep_kernel<double, 3>)};
  for (const void* f : fs) {
    hipError_t r = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBytes);
    if (r != hipSuccess) e = r;
  }
  return e;
}

hipError_t launch_fourstep(bool f64, int nph, const KParams* Q, const void* in, void* out, double scale, int batch,
                           long long src_dist, long long dst_dist, void* scratch, long long slot_elems, dim3 grid,
                           dim3 block, size_t lds, hipStream_t st) {
  if (f64) {
    if (nph == 2)
      hipLaunchKernelGGL((fourstep_kernel<double, 2>), grid, block, lds, st, Q, in, out, scale, batch, src_dist,
                         dst_dist, scratch, slot_elems);
    else
      hipLaunchKernelGGL((fourstep_kernel<double, 3>), grid, block, lds, st, Q, in, out, scale, batch, src_dist,
                         dst_dist, scratch, slot_elems);
  } else {
    if (nph == 2)
      hipLaunchKernelGGL((fourstep_kernel<float, 2>), grid, block, lds, st, Q, in, out, scale, batch, src_dist,
                         dst_dist, scratch, slot_elems);
    else
      hipLaunchKernelGGL((fourstep_kernel<float, 3>), grid, block, lds, st, Q, in, out, scale, batch, src_dist,
                         dst_dist, scratch, slot_elems);
  }
  return hipGetLastError();
}

// Workgroups of the four-step kernel that fit on one CU at once.
int fourstep_occupancy(bool f64, int nph, int block, size_t lds) {
  int n = 0;
  hipError_t e;
  if (f64)
    e = nph == 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fourstep_kernel<double, 2>, block, lds)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fourstep_kernel<double, 3>, block, lds);
  else
    e = nph == 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fourstep_kernel<float, 2>, block, lds)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fourstep_kernel<float, 3>, block, lds);
  return e == hipSuccess ? n : 0;
}

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t st) {
  const long long n = (long long)(bytes / 4);
  if (n == 0) return hipSuccess;
  const int block = 256;
  long long grid = (n + block - 1) / block;
  if (grid > 65536) grid = 65536;
  hipLaunchKernelGGL(copy_kernel<float>, dim3((unsigned)grid), dim3(block), 0, st, (const float*)src,
                     (float*)dst, n);
  return hipGetLastError();
}

}  // namespace chfft
