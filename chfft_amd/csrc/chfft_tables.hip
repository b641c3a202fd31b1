// Device-side table generation for plans that hold many lengths at once (the
// uneven multi-length batch). Every table value is computed in FP64 on the GPU
// (sincospi of an exactly reduced rational angle) and rounded once to the
// plan's precision; Bluestein kernel spectra come from an FP64 FFT of the
// chirp, scaled by 1/M and rounded once (the reference computes the same
// spectrum with its own M-point FFT, cfft1d.rs:120-179).
#include "chfft_device.h"

namespace chfft {

// Table jobs (RootJob, chfft_internal.h): out[off + i] = exp(-2 pi i num(i) / den), i < count.
//   kind 0: pass table [j-1][k]: j = i / p + 1, k = i % p, num = j * k mod den
//   kind 1: Bluestein chirp:     num = i * i mod den (den = 2N)
//   kind 2: two-level high part: num = (i << sh) mod den
//   kind 3: two-level low part:  num = i
template <typename T>
__global__ void gen_roots_kernel(const RootJob* __restrict__ jobs, int njobs, cx<T>* __restrict__ arena) {
  for (int jb = blockIdx.x; jb < njobs; jb += gridDim.x) {
    const RootJob J = jobs[jb];
    for (long long i = threadIdx.x; i < J.count; i += blockDim.x) {
      unsigned long long num;
      if (J.kind == 0) {
        const long long j = i / J.p + 1, k = i % J.p;
        num = (unsigned long long)((j * k) % J.den);
      } else if (J.kind == 1) {
        num = (unsigned long long)(((unsigned __int128)i * (unsigned __int128)i) % (unsigned long long)J.den);
      } else if (J.kind == 2) {
        num = (unsigned long long)((i << J.sh) % J.den);
      } else {
        num = (unsigned long long)i;
      }
      double s, c;
      sincospi(-2.0 * (double)num / (double)J.den, &s, &c);
      arena[J.off + i] = mk((T)c, (T)s);
    }
  }
}

// Bluestein kernel sequence rot_i = conj(c_|i|) folded around M/2, in FP64,
// for `count` lengths sharing M (cfft1d.rs:120-140).
__global__ void gen_rot_kernel(const long long* __restrict__ ns, int count, long long M, cx<double>* out) {
  const long long total = (long long)count * M;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / M, i = e - b * M;
    const long long N = ns[b];
    const long long j = i <= M / 2 ? i : M - i;
    cx<double> v = mk(0.0, 0.0);
    if (j < N) {
      const unsigned long long num = (unsigned long long)(((unsigned __int128)j * j) % (unsigned long long)(2 * N));
      double s, c;
      sincospi(-(double)num / (double)N, &s, &c);
      v = mk(c, -s);
    }
    out[e] = v;
  }
}

// dst[i] = (T)(src[i] / M) for the spectra of `count` lengths; spectrum b goes
// to arena offset offs[b].
template <typename T>
__global__ void put_spectra_kernel(const cx<double>* __restrict__ src, int count, long long M, long long n1,
                                   const long long* __restrict__ offs, cx<T>* __restrict__ arena) {
  const long long total = (long long)count * M;
  const double inv = 1.0 / (double)M;
  const long long n2 = n1 > 0 ? M / n1 : 0;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / M, i = e - b * M;
    // four-step items store bin k1 + n1 * k2 at k1 * n2 + k2 (row-major for
    // the row pass, which reads row k1's bins with consecutive lanes)
    const long long k = n1 > 0 ? (i / n2) + n1 * (i % n2) : i;
    const cx<double> v = src[b * M + k];
    arena[offs[b] + i] = mk((T)(v.re * inv), (T)(v.im * inv));
  }
}

hipError_t gen_roots(bool f64, const RootJob* jobs_dev, int njobs, void* arena, hipStream_t st) {
  if (njobs <= 0) return hipSuccess;
  const int grid = njobs < 65535 ? njobs : 65535;
  if (f64)
    hipLaunchKernelGGL(gen_roots_kernel<double>, dim3(grid), dim3(256), 0, st, jobs_dev, njobs,
                       reinterpret_cast<cx<double>*>(arena));
  else
    hipLaunchKernelGGL(gen_roots_kernel<float>, dim3(grid), dim3(256), 0, st, jobs_dev, njobs,
                       reinterpret_cast<cx<float>*>(arena));
  return hipGetLastError();
}

hipError_t gen_rot(const long long* ns_dev, int count, long long M, void* out, hipStream_t st) {
  hipLaunchKernelGGL(gen_rot_kernel, dim3(1024), dim3(256), 0, st, ns_dev, count, M,
                     reinterpret_cast<cx<double>*>(out));
  return hipGetLastError();
}

hipError_t put_spectra(bool f64, const void* src, int count, long long M, long long n1, const long long* offs_dev,
                       void* arena, hipStream_t st) {
  if (f64)
    hipLaunchKernelGGL(put_spectra_kernel<double>, dim3(1024), dim3(256), 0, st,
                       reinterpret_cast<const cx<double>*>(src), count, M, n1, offs_dev,
                       reinterpret_cast<cx<double>*>(arena));
  else
    hipLaunchKernelGGL(put_spectra_kernel<float>, dim3(1024), dim3(256), 0, st,
                       reinterpret_cast<const cx<double>*>(src), count, M, n1, offs_dev,
                       reinterpret_cast<cx<float>*>(arena));
  return hipGetLastError();
}

}  // namespace chfft
