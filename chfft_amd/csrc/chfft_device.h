// Device-side building blocks shared by the generic kernels
// (chfft_kernels.hip) and the length-specialised kernels (chfft_sfft.hip):
// complex arithmetic, radix-2/3/4/5/7/8/16 butterflies, LDS index padding,
// and the load operators that build the first image of a transform from
// global memory (plain, Bluestein chirp, RFft1D/DCT/MDCT pre-passes).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "chfft_internal.h"

namespace chfft {

template <typename T>
struct alignas(2 * sizeof(T)) cx {
  T re, im;
};

// A pointer the caller knows to address global memory, as a global-address-
// space pointer, so that accesses through it compile to global instructions
// rather than flat ones. Pointers read from descriptors in device memory
// (the multi-length kernel's) are otherwise flat, and a flat access also
// counts against the LDS wait counter: the LDS-only wait before each barrier
// would drain the global loads in flight. The integer round trip keeps the
// compiler from folding the address-space casts away.
template <typename T>
__device__ __forceinline__ T* as_global(T* p) {
  typedef __attribute__((address_space(1))) T G;
  return (T*)(G*)(uintptr_t)p;
}

// Store of one output value to global memory; NT: non-temporal (the L2
// streams the line out instead of keeping it), for the exec's final output
// in the kernels compiled for it.
template <bool NT, typename T>
__device__ __forceinline__ void gput(T* p, T v) {
  if constexpr (NT)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}
template <bool NT, typename T>
__device__ __forceinline__ void gput(cx<T>* p, cx<T> v) {
  if constexpr (NT) {
    typedef T V2 __attribute__((ext_vector_type(2)));
    V2 w;
    w.x = v.re;
    w.y = v.im;
    __builtin_nontemporal_store(w, reinterpret_cast<V2*>(p));
  } else {
    *p = v;
  }
}

template <typename T>
__device__ __forceinline__ cx<T> mk(T r, T i) {
  cx<T> c;
  c.re = r;
  c.im = i;
  return c;
}
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) {
  return mk(a.re + b.re, a.im + b.im);
}
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) {
  return mk(a.re - b.re, a.im - b.im);
}
template <typename T>
__device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T>
__device__ __forceinline__ cx<T> scl(cx<T> a, T s) {
  return mk(a.re * s, a.im * s);
}
template <typename T>
__device__ __forceinline__ cx<T> cj(cx<T> a) {
  return mk(a.re, -a.im);
}
// multiply by -i (forward quarter turn of the DFT kernel)
template <typename T>
__device__ __forceinline__ cx<T> mul_mi(cx<T> a) {
  return mk(a.im, -a.re);
}

// ------------------------------------------------------------------ butterflies
// Forward DFT of R points in registers, natural order in and out:
//   X_m = sum_j x_j exp(-2 pi i j m / R)

template <typename T>
__device__ __forceinline__ void bf2(cx<T>& a, cx<T>& b) {
  cx<T> t = a - b;
  a = a + b;
  b = t;
}

template <typename T>
__device__ __forceinline__ void dft2(cx<T>* x) {
  bf2(x[0], x[1]);
}

template <typename T>
__device__ __forceinline__ void dft3(cx<T>* x) {
  const T s3 = T(0.86602540378443864676372317075294);  // sin(2 pi / 3)
  cx<T> s = x[1] + x[2];
  cx<T> d = x[1] - x[2];
  cx<T> a = x[0] - scl(s, T(0.5));
  cx<T> b = scl(mul_mi(d), s3);  // -i sin(2pi/3) (x1 - x2)
  x[0] = x[0] + s;
  x[1] = a + b;
  x[2] = a - b;
}

template <typename T>
__device__ __forceinline__ void dft4(cx<T>* x) {
  cx<T> s02 = x[0] + x[2], d02 = x[0] - x[2];
  cx<T> s13 = x[1] + x[3], d13 = mul_mi(x[1] - x[3]);
  x[0] = s02 + s13;
  x[2] = s02 - s13;
  x[1] = d02 + d13;
  x[3] = d02 - d13;
}

template <typename T>
__device__ __forceinline__ void dft5(cx<T>* x) {
  const T c1 = T(0.30901699437494742410229341718282);   // cos(2 pi / 5)
  const T c2 = T(-0.80901699437494742410229341718282);  // cos(4 pi / 5)
  const T s1 = T(0.95105651629515357211643933337938);   // sin(2 pi / 5)
  const T s2 = T(0.58778525229247312916870595463907);   // sin(4 pi / 5)
  cx<T> s14 = x[1] + x[4], d14 = x[1] - x[4];
  cx<T> s23 = x[2] + x[3], d23 = x[2] - x[3];
  cx<T> a1 = x[0] + scl(s14, c1) + scl(s23, c2);
  cx<T> a2 = x[0] + scl(s14, c2) + scl(s23, c1);
  cx<T> b1 = mul_mi(scl(d14, s1) + scl(d23, s2));
  cx<T> b2 = mul_mi(scl(d14, s2) - scl(d23, s1));
  x[0] = x[0] + s14 + s23;
  x[1] = a1 + b1;
  x[4] = a1 - b1;
  x[2] = a2 + b2;
  x[3] = a2 - b2;
}

template <typename T>
__device__ __forceinline__ void dft7(cx<T>* x) {
  const T c1 = T(0.62348980185873353052500488400424);
  const T c2 = T(-0.22252093395631440428890256449679);
  const T c3 = T(-0.90096886790241912623610231950745);
  const T s1 = T(0.78183148246802980870844452667406);
  const T s2 = T(0.97492791218182360701813168299393);
  const T s3 = T(0.43388373911755812047576833284835);
  cx<T> p1 = x[1] + x[6], m1 = x[1] - x[6];
  cx<T> p2 = x[2] + x[5], m2 = x[2] - x[5];
  cx<T> p3 = x[3] + x[4], m3 = x[3] - x[4];
  cx<T> x0 = x[0];
  cx<T> a1 = x0 + scl(p1, c1) + scl(p2, c2) + scl(p3, c3);
  cx<T> a2 = x0 + scl(p1, c2) + scl(p2, c3) + scl(p3, c1);
  cx<T> a3 = x0 + scl(p1, c3) + scl(p2, c1) + scl(p3, c2);
  cx<T> b1 = mul_mi(scl(m1, s1) + scl(m2, s2) + scl(m3, s3));
  cx<T> b2 = mul_mi(scl(m1, s2) - scl(m2, s3) - scl(m3, s1));
  cx<T> b3 = mul_mi(scl(m1, s3) - scl(m2, s1) + scl(m3, s2));
  x[0] = x0 + p1 + p2 + p3;
  x[1] = a1 + b1;
  x[6] = a1 - b1;
  x[2] = a2 + b2;
  x[5] = a2 - b2;
  x[3] = a3 + b3;
  x[4] = a3 - b3;
}

// Radix-8: decimation in frequency, two radix-4 halves.
template <typename T>
__device__ __forceinline__ void dft8(cx<T>* x) {
  const T r = T(0.70710678118654752440084436210485);
  cx<T> a[4], b[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    a[j] = x[j] + x[j + 4];
    b[j] = x[j] - x[j + 4];
  }
  // b_j *= W8^j
  b[1] = mk((b[1].re + b[1].im) * r, (b[1].im - b[1].re) * r);
  b[2] = mul_mi(b[2]);
  b[3] = mk((b[3].im - b[3].re) * r, -(b[3].re + b[3].im) * r);
  dft4(a);
  dft4(b);
#pragma unroll
  for (int m = 0; m < 4; m++) {
    x[2 * m] = a[m];
    x[2 * m + 1] = b[m];
  }
}

// Radix-16: decimation in frequency, two radix-8 halves.
template <typename T>
__device__ __forceinline__ void dft16(cx<T>* x) {
  const T c1 = T(0.92387953251128675612818318939679);  // cos(pi/8)
  const T s1 = T(0.38268343236508977172845998403040);  // sin(pi/8)
  const T r = T(0.70710678118654752440084436210485);
  cx<T> a[8], b[8];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    a[j] = x[j] + x[j + 8];
    b[j] = x[j] - x[j + 8];
  }
  // b_j *= W16^j = cos(pi j / 8) - i sin(pi j / 8)
  b[1] = b[1] * mk(c1, -s1);
  b[2] = mk((b[2].re + b[2].im) * r, (b[2].im - b[2].re) * r);
  b[3] = b[3] * mk(s1, -c1);
  b[4] = mul_mi(b[4]);
  b[5] = b[5] * mk(-s1, -c1);
  b[6] = mk((b[6].im - b[6].re) * r, -(b[6].re + b[6].im) * r);
  b[7] = b[7] * mk(-c1, -s1);
  dft8(a);
  dft8(b);
#pragma unroll
  for (int m = 0; m < 8; m++) {
    x[2 * m] = a[m];
    x[2 * m + 1] = b[m];
  }
}

// Prime-factor (Good-Thomas) DFTs for R = R1 * R2 with coprime factors: no
// internal twiddles. Input n = (R2 n1 + R1 n2) mod R goes to row n1, column
// n2; R1-point DFTs along each column, R2-point DFTs along each row, and
// output (k1, k2) is X[k] with k = k1 mod R1, k = k2 mod R2 (CRT):
// k = (A k1 + B k2) mod R.
template <typename T, int R1, int R2, int A, int B, class D1, class D2>
__device__ __forceinline__ void dft_pfa(cx<T>* x) {
  constexpr int R = R1 * R2;
  cx<T> m[R2][R1];
#pragma unroll
  for (int n2 = 0; n2 < R2; n2++)
#pragma unroll
    for (int n1 = 0; n1 < R1; n1++) m[n2][n1] = x[(R2 * n1 + R1 * n2) % R];
#pragma unroll
  for (int n2 = 0; n2 < R2; n2++) D1::template run<T>(m[n2]);
#pragma unroll
  for (int k1 = 0; k1 < R1; k1++) {
    cx<T> c[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; n2++) c[n2] = m[n2][k1];
    D2::template run<T>(c);
#pragma unroll
    for (int k2 = 0; k2 < R2; k2++) x[(A * k1 + B * k2) % R] = c[k2];
  }
}
struct Dft2 {
  template <typename T>
  __device__ __forceinline__ static void run(cx<T>* x) { dft2(x); }
};
struct Dft3 {
  template <typename T>
  __device__ __forceinline__ static void run(cx<T>* x) { dft3(x); }
};
struct Dft4 {
  template <typename T>
  __device__ __forceinline__ static void run(cx<T>* x) { dft4(x); }
};
struct Dft5 {
  template <typename T>
  __device__ __forceinline__ static void run(cx<T>* x) { dft5(x); }
};
// 10 = 2 x 5: k = 5 k1 + 6 k2 (mod 10); 12 = 4 x 3: k = 9 k1 + 4 k2 (mod 12);
// 15 = 3 x 5: k = 10 k1 + 6 k2 (mod 15)
template <typename T>
__device__ __forceinline__ void dft10(cx<T>* x) { dft_pfa<T, 2, 5, 5, 6, Dft2, Dft5>(x); }
template <typename T>
__device__ __forceinline__ void dft12(cx<T>* x) { dft_pfa<T, 4, 3, 9, 4, Dft4, Dft3>(x); }
template <typename T>
__device__ __forceinline__ void dft15(cx<T>* x) { dft_pfa<T, 3, 5, 10, 6, Dft3, Dft5>(x); }

template <typename T, int R>
__device__ __forceinline__ void dft(cx<T>* x) {
  if constexpr (R == 2) dft2(x);
  else if constexpr (R == 3) dft3(x);
  else if constexpr (R == 4) dft4(x);
  else if constexpr (R == 5) dft5(x);
  else if constexpr (R == 7) dft7(x);
  else if constexpr (R == 8) dft8(x);
  else if constexpr (R == 10) dft10(x);
  else if constexpr (R == 12) dft12(x);
  else if constexpr (R == 15) dft15(x);
  else if constexpr (R == 16) dft16(x);
}

// ------------------------------------------------------------------ thread map
// A workgroup of tpb * tpt threads owns tpb transforms. For LDS passes and
// row-major global access, thread (t, u) = (tid / tpt, tid % tpt) works on
// transform t; for column-major global access (consecutive threads on
// consecutive transforms, so that strided columns are read in full cache
// lines) it is (tid % tpb, tid / tpb). Both are computed once per thread.
struct TMap {
  int tpt;
  int tr, ur;  // row-major
  int tc, uc;  // column-major
};

__device__ __forceinline__ TMap make_map(int tpb) {
  TMap m;
  m.tpt = blockDim.x / tpb;
  m.tr = threadIdx.x / m.tpt;
  m.ur = threadIdx.x - m.tr * m.tpt;
  m.tc = threadIdx.x % tpb;
  m.uc = threadIdx.x / tpb;
  return m;
}

// LDS element index with one padding slot per 16 elements: Stockham writes of
// the early passes (stride R*p < 16) and column-major tile loads then spread
// over the banks instead of piling onto one.
__device__ __forceinline__ int lx(int n) { return n + (n >> 4); }

// floor(i / p) for i < 2^16, p <= 2^14, with pm = ceil(2^32 / p).
__device__ __forceinline__ int qdiv(int i, unsigned pm) { return (int)__umulhi((unsigned)i, pm); }

// ------------------------------------------------------------------ operators
// Offset of transform g (its element 0) and its column index under IoDesc d.
__device__ __forceinline__ long long io_base(const IoDesc& d, int g, int* col) {
  const int b = g / d.ncols;
  const int c = g - b * d.ncols;
  *col = c;
  return (long long)b * d.bdist + (long long)c * d.cstride;
}

// Per-transform state of the load operator (computed once per thread).
template <typename T, int LD>
struct Loader {
  const cx<T>* cs;  // complex row / column base
  const T* rs;      // real row base
  long long seq0;   // FS_CHIRP: sequence index of element 0 (= column * cstride)
  long long brow;   // FS_CHIRP: batch base
};

template <typename T, int LD>
__device__ __forceinline__ Loader<T, LD> make_loader(const KParams& P, const void* src, int g) {
  Loader<T, LD> L;
  int c;
  const long long base = io_base(P.in, g, &c);
  L.cs = reinterpret_cast<const cx<T>*>(as_global(src)) + base;
  L.rs = reinterpret_cast<const T*>(as_global(src)) + base;
  L.seq0 = (long long)c * P.in.cstride;
  L.brow = base - L.seq0;
  return L;
}

// Element n of the load image of one transform, before conjugation.
template <typename T, int LD>
__device__ __forceinline__ cx<T> load_elem(const KParams& P, const Loader<T, LD>& ld, int n) {
  const cx<T> zero = mk(T(0), T(0));
  if constexpr (LD == LD_PLAIN) {
    return ld.cs[(long long)n * P.in.estride];
  } else if constexpr (LD == LD_FS_CHIRP) {
    const long long s = ld.seq0 + (long long)n * P.in.estride;
    if (s >= P.N) return zero;
    cx<T> v = reinterpret_cast<const cx<T>*>(ld.cs - ld.seq0)[s];
    if (P.conj_in) v = cj(v);  // inverse: conjugate the input, not the chirp
    return v * reinterpret_cast<const cx<T>*>(as_global(P.chirp))[s];
  } else if constexpr (LD == LD_R2C_BACK) {
    // rfft1d.rs:161-177 (complex input, N/2+1 bins per row)
    const cx<T>* X = ld.cs;
    const cx<T>* bc = reinterpret_cast<const cx<T>*>(as_global(P.tab_a));
    const int N = (int)P.N, h = N >> 1, q = (N + 3) >> 2;
    if (n == 0) {
      const T a = X[0].re, z = X[h].re;
      return mk((a + z) * T(0.5), (a - z) * T(0.5));
    }
    if ((N & 3) == 0 && n == q) return cj(X[q]);
    const int i = (n < q) ? n : h - n;
    const cx<T> xi = X[i], xh = X[h - i];
    const cx<T> x = bc[i] * (xi - cj(xh));
    return (n < q) ? xi - x : xh + cj(x);
  } else if constexpr (LD == LD_DCT2) {
    // dct1d.rs:245-263 (real input)
    const T* x = ld.rs;
    const int N = (int)P.N, h = N >> 1, q = h >> 1;
    const T o = (T)P.ortho;
    if (q == 0) return mk(x[0] * o, x[1]);
    if (n == 0) return mk(x[0] * o, x[2]);
    if (n < q) return mk(x[4 * n], x[4 * n + 2]);
    if ((h & 1) && n == q) return mk(x[N - 2], x[N - 1]);
    const int i = h - 1 - n;
    return mk(x[4 * i + 3], x[4 * i + 1]);
  } else if constexpr (LD == LD_DCT3) {
    // dct1d.rs:367-391 (real input)
    const T* x = ld.rs;
    const cx<T>* co = reinterpret_cast<const cx<T>*>(as_global(P.tab_a));
    const cx<T>* om = reinterpret_cast<const cx<T>*>(as_global(P.tab_b));
    const int N = (int)P.N, h = N >> 1, q = (N + 3) >> 2;
    if (n == 0) {
      const T z = x[0] * (T)P.ortho;
      const T hv = x[h] * om[h].re;
      return mk(z + hv, z - hv);
    }
    if ((N & 3) == 0 && n == q) return mk(x[q], x[N - q]) * om[q];
    const int i = (n < q) ? n : h - n;
    const cx<T> j = cj(mk(x[i], x[N - i]) * om[i]);
    const cx<T> k = mk(x[h - i], x[h + i]) * om[h - i];
    const cx<T> xx = co[i] * (j - k);
    return (n < q) ? j - xx : cj(k) + cj(xx);
  } else if constexpr (LD == LD_MDCT) {
    // mdct1d.rs:206-230 (frames `bdist` samples apart)
    const T* x = ld.rs;
    const cx<T>* tw = reinterpret_cast<const cx<T>*>(as_global(P.tab_a));
    const T* w = reinterpret_cast<const T*>(as_global(P.tab_b));
    const int N = (int)P.N, q = N >> 2, h = N >> 1, q3 = h + q, q5 = h + q3, lo = q - (q >> 1);
    const int m = n << 1;
    cx<T> v;
    if (n < lo)
      v = mk(x[q3 - 1 - m] * w[q3 - 1 - m] + x[q3 + m] * w[q3 + m], x[q + m] * w[q + m] - x[q - 1 - m] * w[q - 1 - m]);
    else
      v = mk(x[q3 - 1 - m] * w[q3 - 1 - m] - x[m - q] * w[m - q], x[q + m] * w[q + m] + x[q5 - 1 - m] * w[q5 - 1 - m]);
    return v * tw[n];
  } else {  // LD_IMDCT, mdct1d.rs:261-265
    const T* c = ld.rs;
    const cx<T>* tw = reinterpret_cast<const cx<T>*>(as_global(P.tab_a));
    const int h = (int)(P.N >> 1);
    return mk(c[2 * n], c[h - 1 - 2 * n]) * scl(tw[n], T(-2));
  }
}

// Post-processing store operators (RFft1D split, DCT-II/III and MDCT/IMDCT
// outputs) for one transform whose final complex sequence is read through
// Z(k) (conjugation already applied): bins i = i0, i0 + step, ... of the
// operator's own loop, written to the transform's output at element obase.
// Shared by the fused kernels (Z reads the LDS image) and the stand-alone
// post-pass kernel of the long-length schedules (Z reads global memory).
template <typename T, int ST, bool NT = false, class ZF>
__device__ __forceinline__ void post_ops(const KParams& P, ZF Z, void* dst, long long obase, int i0, int step, T s) {
  if constexpr (ST == ST_R2C) {
    // rfft1d.rs:134-148: N/2+1 bins per row. Bins i and h - i (0 < i < q)
    // come from the same pair Z(i), Z(h - i) and twiddle co[i]: one thread
    // computes both; i = 0 gives bins 0 and h, i = q (N % 4 == 0) bin q.
    const int N = (int)P.N, h = N >> 1, q = (N + 3) >> 2;
    const cx<T>* co = reinterpret_cast<const cx<T>*>(as_global(P.tab_a));
    cx<T>* out = reinterpret_cast<cx<T>*>(as_global(dst)) + obase;
    for (int i = i0; i <= q; i += step) {
      if (i == 0) {
        const cx<T> z = Z(0);
        gput<NT>(&out[0], scl(mk(z.re + z.im, T(0)), s));
        gput<NT>(&out[h], scl(mk(z.re - z.im, T(0)), s));
      } else if (i < q) {
        const cx<T> zi = Z(i), zh = Z(h - i);
        const cx<T> x = co[i] * (zi - cj(zh));
        gput<NT>(&out[i], scl(zi - x, s));
        gput<NT>(&out[h - i], scl(zh + cj(x), s));
      } else if ((N & 3) == 0) {
        gput<NT>(&out[q], scl(cj(Z(q)), s));
      }
    }
  } else if constexpr (ST == ST_DCT2) {
    // dct1d.rs:266-287
    const int N = (int)P.N, h = N >> 1, q2 = (N + 3) >> 2;
    const cx<T>* co = reinterpret_cast<const cx<T>*>(as_global(P.tab_a));
    const cx<T>* om = reinterpret_cast<const cx<T>*>(as_global(P.tab_b));
    T* ret = reinterpret_cast<T*>(as_global(dst)) + obase;
    for (int i = i0; i <= q2; i += step) {
      if (i == 0) {
        const cx<T> z = Z(0);
        gput<NT>(&ret[0], (z.re + z.im) * s);
        gput<NT>(&ret[h], (z.re - z.im) * om[h].re * s);
      } else if (i < q2) {
        const cx<T> zi = Z(i);
        const cx<T> wc = cj(Z(h - i));
        const cx<T> x = co[i] * (zi - wc);
        const cx<T> j = cj(zi - x) * om[i];
        gput<NT>(&ret[i], j.re * s);
        gput<NT>(&ret[N - i], j.im * s);
        const cx<T> k = (wc + x) * om[h - i];
        gput<NT>(&ret[h - i], k.re * s);
        gput<NT>(&ret[h + i], k.im * s);
      } else if ((N & 3) == 0) {
        const cx<T> x = Z(q2) * om[q2];
        gput<NT>(&ret[q2], x.re * s);
        gput<NT>(&ret[N - q2], x.im * s);
      }
    }
  } else if constexpr (ST == ST_DCT3) {
    // dct1d.rs:395-407
    const int N = (int)P.N, h = N >> 1, q2 = h >> 1;
    const int per = q2 + (h & 1);
    T* ret = reinterpret_cast<T*>(as_global(dst)) + obase;
    for (int i = i0; i < per; i += step) {
      const cx<T> a = Z(i);
      if (i < q2) {
        const cx<T> b = Z(h - 1 - i);
        gput<NT>(&ret[4 * i], a.re * s);
        gput<NT>(&ret[4 * i + 1], b.im * s);
        gput<NT>(&ret[4 * i + 2], a.im * s);
        gput<NT>(&ret[4 * i + 3], b.re * s);
      } else {
        gput<NT>(&ret[4 * i], a.re * s);
        gput<NT>(&ret[4 * i + 1], a.im * s);
      }
    }
    (void)N;
  } else if constexpr (ST == ST_MDCT) {
    // mdct1d.rs:234-240
    const int N = (int)P.N, q = N >> 2, h = N >> 1;
    const cx<T>* tw = reinterpret_cast<const cx<T>*>(as_global(P.tab_a));
    T* ret = reinterpret_cast<T*>(as_global(dst)) + obase;
    for (int i = i0; i < q; i += step) {
      const cx<T> v = Z(i) * scl(tw[i], s);
      gput<NT>(&ret[2 * i], -v.re);
      gput<NT>(&ret[h - 1 - 2 * i], v.im);
    }
  } else {  // ST_IMDCT, mdct1d.rs:268-288
    const int N = (int)P.N, q = N >> 2, h = N >> 1, q3 = h + q, q5 = h + q3, lo = q - (q >> 1);
    const cx<T>* tw = reinterpret_cast<const cx<T>*>(as_global(P.tab_a));
    const T* w = reinterpret_cast<const T*>(as_global(P.tab_b));
    T* ret = reinterpret_cast<T*>(as_global(dst)) + obase;
    for (int i = i0; i < q; i += step) {
      const cx<T> v = Z(i) * scl(tw[i], s);
      const int mm = 2 * i;
      gput<NT>(&ret[q3 - 1 - mm], v.re * w[q3 - 1 - mm]);
      gput<NT>(&ret[q + mm], -v.im * w[q + mm]);
      if (i < lo) {
        gput<NT>(&ret[q3 + mm], v.re * w[q3 + mm]);
        gput<NT>(&ret[q - 1 - mm], v.im * w[q - 1 - mm]);
      } else {
        gput<NT>(&ret[mm - q], -v.re * w[mm - q]);
        gput<NT>(&ret[q5 - 1 - mm], -v.im * w[q5 - 1 - mm]);
      }
    }
  }
}


// Store stage from the final LDS image (natural order, padded index), one
// loop per store operator.
template <typename T, int ST, bool NT = false>
__device__ void store_stage(const KParams& P, const TMap& m, void* dst, const cx<T>* lds, int g0, int nt,
                            double scale) {
  const T s = (T)scale;
  if constexpr (ST == ST_PLAIN || ST == ST_TW || ST == ST_FS_DECHIRP) {
    const int t = P.colmajor_st ? m.tc : m.tr;
    const int u = P.colmajor_st ? m.uc : m.ur;
    if (t >= nt) return;
    const cx<T>* row = lds + t * P.Lp;
    int c;
    const long long base = io_base(P.out, g0 + t, &c);
    cx<T>* out = reinterpret_cast<cx<T>*>(as_global(dst)) + base;
    const long long es = P.out.estride;
    // two-level four-step twiddles by FP64 recurrence where the plan has FP64
    // two-level tables (tw_rec: [W^i, i < 2^twsh][W^(i 2^twsh)]): W^(c u) and
    // the step W^(c tpt) from two lookups each (the step only where a second
    // output exists, so that c tpt < c L <= the table's length), each twiddle
    // rounded once
    const cx<double>* t2 = (ST == ST_TW && P.twsh) ? reinterpret_cast<const cx<double>*>(as_global(P.tw_rec)) : nullptr;
    cx<double> wd = mk(1.0, 0.0), wst = mk(1.0, 0.0);
    if (ST == ST_TW && t2) {
      const unsigned sh = (unsigned)P.twsh, lo = (1u << sh) - 1u;
      const unsigned m0 = (unsigned)c * (unsigned)u;
      wd = t2[(1u << sh) + (m0 >> sh)] * t2[m0 & lo];
      if (u + m.tpt < P.L) {
        const unsigned ms = (unsigned)c * (unsigned)m.tpt;
        wst = t2[(1u << sh) + (ms >> sh)] * t2[ms & lo];
      }
    }
    for (int k = u; k < P.L; k += m.tpt) {
      cx<T> v = row[lx(k)];
      if constexpr (ST == ST_PLAIN) {
        if (P.conj_out) v = cj(v);
        gput<NT>(&out[k * es], scl(v, s));
      } else if constexpr (ST == ST_TW) {
        if (t2) {
          v = v * mk((T)wd.re, (T)wd.im);
          wd = wd * wst;
        } else if (P.twsh) {
          // two-level table: W^m = W^(m_hi << sh) * W^(m_lo), both rounded once
          const int mm = c * k;
          v = v * (reinterpret_cast<const cx<T>*>(as_global(P.tab_a))[mm >> P.twsh] *
                   reinterpret_cast<const cx<T>*>(as_global(P.tab_b))[mm & ((1 << P.twsh) - 1)]);
        } else {
          // table of W^(c*k) laid out so that consecutive lanes read
          // consecutive entries: index k * twk + c * twc
          v = v * reinterpret_cast<const cx<T>*>(as_global(P.tab_a))[k * P.twk + c * P.twc];
        }
        if (P.conj_out) v = cj(v);
        gput<NT>(&out[k * es], scl(v, s));
      } else {
        const long long j = (long long)c * P.out.cstride + (long long)k * es;
        if (j < P.N) {
          v = cj(v) * reinterpret_cast<const cx<T>*>(as_global(P.chirp))[j];
          if (P.conj_out) v = cj(v);
          gput<NT>(&reinterpret_cast<cx<T>*>(as_global(dst))[base - (long long)c * P.out.cstride + j], scl(v, s));
        }
      }
    }
    return;
  }
  const int t = m.tr;
  if (t >= nt) return;
  const cx<T>* row = lds + t * P.Lp;
  const bool cjo = P.conj_out != 0;
  post_ops<T, ST, NT>(P, [&](int k) {
    const cx<T> z = row[lx(k)];
    return cjo ? cj(z) : z;
  }, dst, (long long)(g0 + t) * P.out.bdist, m.ur, m.tpt, s);
}

}  // namespace chfft
