// One HBM pass for 43200-point FP32 complex transforms (CFft1D::forward and
// friends at N = 43200 = 2^6 3^3 5^2, the reference's mixed-radix path,
// cfft1d.rs:208-230 / mixed_radix.rs:104-166).
//
// A 43200-point FP32 transform (345.6 KB) does not fit one CU's LDS, so the
// split four-step moves it through HBM twice. Here one 960-thread workgroup
// holds a whole transform in registers (45 points per thread, 90 VGPRs) and
// exchanges it through LDS in three rounds:
//
//   N = 45 x 960, n = t + 960 j (t < 960, j < 45), k = k1 + 45 k2
//   X[k1 + 45 k2] = sum_t W_960^(t k2) W_N^(t k1) sum_j x[t + 960 j] W_45^(j k1)
//
//   1. thread t loads x[t + 960 j] for every j (coalesced: lanes on
//      consecutive t) and runs the 45-point DFT over j in registers
//      (prime-factor 5 x 9, no inner twiddles);
//   2. three rounds of 15 k1 values each: multiply by W_N^(t k1), write the
//      15 sequences (length 960, element t) to LDS, run the 960-point DFTs
//      as Stockham passes of radix 4, 16 and 15 over LDS (the radix-15 pass
//      is wave-local and in place: sequence q belongs to wave q), and read
//      back the outputs this thread stores: X[t + 960 j] has
//      k1 = (t + 15 j) mod 45, so round r holds it for the 15 j with
//      j = r - c (mod 3), c = (t mod 45) / 15, all in sequence t mod 15 at
//      elements K0 + 64 i;
//   3. thread t stores X[t + 960 j] for every j (coalesced).
//
// LDS: 15 sequences of 1025 slots (960 elements, one padding slot per 16,
// so that the radix-4 and radix-16 writes and the output reads spread over
// the banks), 123,000 bytes. Methods as the other kernels: the inverse is
// the forward transform between two conjugations, then the scale.
#include <hip/hip_runtime.h>

#include "chfft_device.h"
#include "chfft_internal.h"

namespace chfft {
namespace {

constexpr int kN = 43200, kT = 960, kLS = 1025;

__device__ __forceinline__ int px(int e) { return e + (e >> 4); }

__device__ __forceinline__ void lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// 9-point DFT as 3 x 3 Cooley-Tukey: n = 3a + b, DFT3 over a for each b,
// twiddle W_9^(b k1), DFT3 over b; X[k1 + 3 k2].
template <typename T>
__device__ __forceinline__ void dft9(cx<T>* x) {
  const cx<T> w1 = mk(T(0.7660444431189780352023927), T(-0.6427876096865393263226434));
  const cx<T> w2 = mk(T(0.1736481776669303488517166), T(-0.9848077530122080593667430));
  const cx<T> w4 = mk(T(-0.9396926207859083840541093), T(-0.3420201433256687330440996));
  cx<T> y[3][3];
#pragma unroll
  for (int b = 0; b < 3; b++) {
    y[b][0] = x[b];
    y[b][1] = x[b + 3];
    y[b][2] = x[b + 6];
    dft3(y[b]);
  }
  y[1][1] = y[1][1] * w1;
  y[1][2] = y[1][2] * w2;
  y[2][1] = y[2][1] * w2;
  y[2][2] = y[2][2] * w4;
#pragma unroll
  for (int k1 = 0; k1 < 3; k1++) {
    cx<T> c[3] = {y[0][k1], y[1][k1], y[2][k1]};
    dft3(c);
    x[k1] = c[0];
    x[k1 + 3] = c[1];
    x[k1 + 6] = c[2];
  }
}
struct Dft9 {
  template <typename T>
  __device__ __forceinline__ static void run(cx<T>* x) { dft9(x); }
};
// 45 = 5 x 9 (coprime): input (9 n1 + 5 n2) mod 45, output k = 36 k1 + 10 k2 (mod 45)
template <typename T>
__device__ __forceinline__ void dft45(cx<T>* x) { dft_pfa<T, 5, 9, 36, 10, Dft5, Dft9>(x); }

template <typename V>
__device__ __forceinline__ V ld_at(const void* base, unsigned off) {
  return *reinterpret_cast<const V*>(reinterpret_cast<const char*>(base) + off);
}
template <typename V>
__device__ __forceinline__ void st_at(void* base, unsigned off, V v) {
  *reinterpret_cast<V*>(reinterpret_cast<char*>(base) + off) = v;
}

// tabA[k1 * 960 + t] = W_N^(t k1); tw64[(m - 1) * 4 + k] = W_64^(m k);
// tw960[(m - 1) * 64 + l] = W_960^(m l). Workgroup b takes transforms
// b, b + grid, ...
__global__ void __launch_bounds__(960) onepass43200_kernel(const cx<float>* __restrict__ src,
                                                           cx<float>* __restrict__ dst, int batch, long long din,
                                                           long long dout, const cx<float>* __restrict__ tabA,
                                                           const cx<float>* __restrict__ tw64,
                                                           const cx<float>* __restrict__ tw960, int conj,
                                                           float scale) {
  extern __shared__ cx<float> lds[];
  const int t = threadIdx.x;
  const int w = t >> 6, l = t & 63;
  const int d = t % 15, c = (t % 45) / 15;
  const int ia = t % 240, qa = t / 240;                    // radix-4 pass: butterflies (qa + 4 s, ia)
  const int qb = t / 60, ib = t - 60 * (t / 60), kb = ib & 3;  // radix-16 pass (t < 900)
  for (int b = blockIdx.x; b < batch; b += gridDim.x) {
    const cx<float>* x = src + (long long)b * din;
    cx<float> v[45];
#pragma unroll
    for (int j = 0; j < 45; j++) {
      v[j] = ld_at<cx<float>>(x, (unsigned)(t + kT * j) * 8u);
      if (conj) v[j].im = -v[j].im;
    }
    dft45(v);
    cx<float> o[45];
#pragma unroll
    for (int r = 0; r < 3; r++) {
      // this round's 15 sequences: element t of sequence q is W_N^(t k1) v[k1], k1 = 15 r + q
#pragma unroll
      for (int q = 0; q < 15; q++) {
        const int k1 = 15 * r + q;
        cx<float> y = v[k1];
        if (k1) y = y * tabA[k1 * kT + t];
        lds[q * kLS + px(t)] = y;
      }
      lds_sync();
      // radix 4, span 1: butterfly i of sequence q reads elements i + 240 m, writes 4 i + m
      {
        cx<float> a[4][4];
#pragma unroll
        for (int s = 0; s < 4; s++)
          if (qa + 4 * s < 15) {
#pragma unroll
            for (int m = 0; m < 4; m++) a[s][m] = lds[(qa + 4 * s) * kLS + px(ia) + 255 * m];
          }
        lds_sync();
#pragma unroll
        for (int s = 0; s < 4; s++)
          if (qa + 4 * s < 15) {
            dft4(a[s]);
#pragma unroll
            for (int m = 0; m < 4; m++) lds[(qa + 4 * s) * kLS + px(4 * ia + m)] = a[s][m];
          }
        lds_sync();
      }
      // radix 16, span 4: butterfly i (k = i mod 4) reads i + 60 m, twiddle W_64^(m k),
      // writes 16 (i - k) + k + 4 m
      {
        cx<float> a[16];
        if (t < 900) {
#pragma unroll
          for (int m = 0; m < 16; m++) a[m] = lds[qb * kLS + px(ib + 60 * m)];
#pragma unroll
          for (int m = 1; m < 16; m++) a[m] = a[m] * tw64[(m - 1) * 4 + kb];
        }
        lds_sync();
        if (t < 900) {
          dft16(a);
#pragma unroll
          for (int m = 0; m < 16; m++) lds[qb * kLS + px(16 * (ib - kb) + kb + 4 * m)] = a[m];
        }
        lds_sync();
      }
      // radix 15, span 64: sequence w, butterfly l reads and writes l + 64 m (in place,
      // only this wave touches sequence w), twiddle W_960^(m l)
      {
        cx<float> a[15];
#pragma unroll
        for (int m = 0; m < 15; m++) a[m] = lds[w * kLS + px(l) + 68 * m];
#pragma unroll
        for (int m = 1; m < 15; m++) a[m] = a[m] * tw960[(m - 1) * 64 + l];
        dft15(a);
#pragma unroll
        for (int m = 0; m < 15; m++) lds[w * kLS + px(l) + 68 * m] = a[m];
      }
      lds_sync();
      // outputs X[t + 960 j] for j = 3 i + j0: sequence t mod 15, elements K0 + 64 i
      {
        const int j0 = (r - c + 3) % 3;
        const int K0 = (t + kT * j0) / 45;
#pragma unroll
        for (int i = 0; i < 15; i++) o[15 * r + i] = lds[d * kLS + px(K0) + 68 * i];
      }
      lds_sync();
    }
    cx<float>* y = dst + (long long)b * dout;
#pragma unroll
    for (int j = 0; j < 45; j++) {
      const int i = j / 3, m = j % 3;
      const int rr = (m + c) % 3;  // the round that holds X[t + 960 j]
      // the three candidates pass through registers first: a select of array
      // elements by a run-time index would otherwise become an indexed
      // access and move the whole array to scratch memory
      cx<float> a0 = o[i], a1 = o[15 + i], a2 = o[30 + i];
      asm volatile("" : "+v"(a0.re), "+v"(a0.im), "+v"(a1.re), "+v"(a1.im), "+v"(a2.re), "+v"(a2.im));
      cx<float> X = rr == 0 ? a0 : rr == 1 ? a1 : a2;
      X = scl(X, scale);
      if (conj) X.im = -X.im;
      st_at<cx<float>>(y, (unsigned)(t + kT * j) * 8u, X);
    }
  }
}

}  // namespace

bool onepass_supported(bool f64, long long n) { return !f64 && n == kN; }

size_t onepass_lds_bytes() { return (size_t)15 * kLS * sizeof(cx<float>); }

hipError_t onepass_set_lds_limits() {
  return hipFuncSetAttribute(reinterpret_cast<const void*>(&onepass43200_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)onepass_lds_bytes());
}

hipError_t onepass_launch(const void* src, void* dst, int batch, long long din, long long dout, const void* tabA,
                          const void* tw64, const void* tw960, int conj, float scale, int grid, hipStream_t st) {
  hipLaunchKernelGGL(onepass43200_kernel, dim3((unsigned)grid), dim3(kT), onepass_lds_bytes(), st,
                     reinterpret_cast<const cx<float>*>(src), reinterpret_cast<cx<float>*>(dst), batch, din, dout,
                     reinterpret_cast<const cx<float>*>(tabA), reinterpret_cast<const cx<float>*>(tw64),
                     reinterpret_cast<const cx<float>*>(tw960), conj, scale);
  return hipGetLastError();
}

}  // namespace chfft
