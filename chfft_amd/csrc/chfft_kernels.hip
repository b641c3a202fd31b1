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

#include "chfft_internal.h"

namespace chfft {

template <typename T>
struct alignas(2 * sizeof(T)) cx {
  T re, im;
};

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

template <typename T, int R>
__device__ __forceinline__ void dft(cx<T>* x) {
  if constexpr (R == 2) dft2(x);
  else if constexpr (R == 3) dft3(x);
  else if constexpr (R == 4) dft4(x);
  else if constexpr (R == 5) dft5(x);
  else if constexpr (R == 7) dft7(x);
  else if constexpr (R == 8) dft8(x);
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

// floor(i / p) for i < 2^16, p <= 2^14, with pm = ceil(2^32 / p).
__device__ __forceinline__ int qdiv(int i, unsigned pm) { return (int)__umulhi((unsigned)i, pm); }

// ------------------------------------------------------------------ LDS passes
// One Stockham pass (span p -> p*R) over all tpb sequences of the workgroup.
// Butterfly i of a sequence reads x[i + j*L/R], multiplies input j by
// W_{pR}^{j*k} (k = i mod p), and writes X_j to ((i-k)*R + k + j*p).
// Inputs of all of a thread's butterflies are staged in registers between the
// read barrier and the write, so one LDS buffer suffices.
template <typename T, int R>
__device__ __forceinline__ void lds_pass(cx<T>* lds, const TMap& m, int L, int Lp, const PassDesc& pd,
                                         const cx<T>* __restrict__ tw) {
  constexpr int BPT = (kEmax + R - 1) / R;
  const int nbt = L / R;
  const int p = pd.p;
  cx<T> x[BPT][R];
  int obase[BPT];
  cx<T>* base = lds + m.tr * Lp;
#pragma unroll
  for (int q = 0; q < BPT; q++) {
    const int i = m.ur + q * m.tpt;
    obase[q] = -1;
    if (i < nbt) {
      const int k = (p == 1) ? 0 : i - qdiv(i, pd.pmagic) * p;
#pragma unroll
      for (int j = 0; j < R; j++) x[q][j] = base[i + j * nbt];
      if (p > 1) {
        const int step = k * pd.tws;
#pragma unroll
        for (int j = 1; j < R; j++) x[q][j] = x[q][j] * tw[j * step];
      }
      obase[q] = (i - k) * R + k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < BPT; q++) {
    if (obase[q] >= 0) {
      dft<T, R>(x[q]);
      cx<T>* d = base + obase[q];
#pragma unroll
      for (int j = 0; j < R; j++) d[j * p] = x[q][j];
    }
  }
  __syncthreads();
}

template <typename T>
__device__ void run_passes(cx<T>* lds, const TMap& m, const SeqDesc& s, int Lp) {
  const cx<T>* tw = reinterpret_cast<const cx<T>*>(s.tw);
  for (int q = 0; q < s.npass; q++) {
    const PassDesc& pd = s.pass[q];
    switch (pd.radix) {
      case 2: lds_pass<T, 2>(lds, m, s.L, Lp, pd, tw); break;
      case 3: lds_pass<T, 3>(lds, m, s.L, Lp, pd, tw); break;
      case 4: lds_pass<T, 4>(lds, m, s.L, Lp, pd, tw); break;
      case 5: lds_pass<T, 5>(lds, m, s.L, Lp, pd, tw); break;
      case 7: lds_pass<T, 7>(lds, m, s.L, Lp, pd, tw); break;
      case 8: lds_pass<T, 8>(lds, m, s.L, Lp, pd, tw); break;
      case 16: lds_pass<T, 16>(lds, m, s.L, Lp, pd, tw); break;
      default: break;
    }
  }
}

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
  L.cs = reinterpret_cast<const cx<T>*>(src) + base;
  L.rs = reinterpret_cast<const T*>(src) + base;
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
    return v * reinterpret_cast<const cx<T>*>(P.chirp)[s];
  } else if constexpr (LD == LD_R2C_BACK) {
    // rfft1d.rs:161-177 (complex input, N/2+1 bins per row)
    const cx<T>* X = ld.cs;
    const cx<T>* bc = reinterpret_cast<const cx<T>*>(P.tab_a);
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
    const cx<T>* co = reinterpret_cast<const cx<T>*>(P.tab_a);
    const cx<T>* om = reinterpret_cast<const cx<T>*>(P.tab_b);
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
    const cx<T>* tw = reinterpret_cast<const cx<T>*>(P.tab_a);
    const T* w = reinterpret_cast<const T*>(P.tab_b);
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
    const cx<T>* tw = reinterpret_cast<const cx<T>*>(P.tab_a);
    const int h = (int)(P.N >> 1);
    return mk(c[2 * n], c[h - 1 - 2 * n]) * scl(tw[n], T(-2));
  }
}

template <typename T, int LD>
__device__ void load_stage(const KParams& P, const TMap& m, const void* src, cx<T>* lds, int g0, int nt) {
  const int t = P.colmajor_ld ? m.tc : m.tr;
  const int u = P.colmajor_ld ? m.uc : m.ur;
  const cx<T>* chirp = reinterpret_cast<const cx<T>*>(P.chirp);
  cx<T>* row = lds + t * P.Lp;
  if (t < nt) {
    const Loader<T, LD> ld = make_loader<T, LD>(P, src, g0 + t);
    for (int n = u; n < P.Lfft; n += m.tpt) {
      cx<T> v = mk(T(0), T(0));
      if (n < P.L) {  // n >= L only under a fused Bluestein: zero padding
        v = load_elem<T, LD>(P, ld, n);
        if constexpr (LD != LD_FS_CHIRP) {
          if (P.conj_in) v = cj(v);
        }
        if (P.has_chirp) v = v * chirp[n];
      }
      row[n] = v;
    }
  } else {
    for (int n = u; n < P.Lfft; n += m.tpt) row[n] = mk(T(0), T(0));
  }
  __syncthreads();
}

// Bluestein spectrum product (chirpz.rs:115-117) followed by a conjugation so
// that the inverse transform reuses the forward passes.
template <typename T>
__device__ void mid_stage(const KParams& P, const TMap& m, cx<T>* lds, int g0) {
  const cx<T>* K = reinterpret_cast<const cx<T>*>(P.kspec);
  const int c = (g0 + m.tr) % P.in.ncols;
  const long long kb = (long long)c * P.kc;
  cx<T>* row = lds + m.tr * P.Lp;
  for (int n = m.ur; n < P.Lfft; n += m.tpt) row[n] = cj(row[n] * K[kb + (long long)n * P.ke]);
  __syncthreads();
}

// Fused Bluestein de-chirp (chirpz.rs:127-145 with the conjugation undone).
template <typename T>
__device__ void dechirp_stage(const KParams& P, const TMap& m, cx<T>* lds) {
  const cx<T>* chirp = reinterpret_cast<const cx<T>*>(P.chirp);
  cx<T>* row = lds + m.tr * P.Lp;
  for (int n = m.ur; n < P.L; n += m.tpt) row[n] = cj(row[n]) * chirp[n];
  __syncthreads();
}

template <typename T, int ST>
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
    cx<T>* out = reinterpret_cast<cx<T>*>(dst) + base;
    const long long es = P.out.estride;
    for (int k = u; k < P.L; k += m.tpt) {
      cx<T> v = row[k];
      if constexpr (ST == ST_PLAIN) {
        if (P.conj_out) v = cj(v);
        out[k * es] = scl(v, s);
      } else if constexpr (ST == ST_TW) {
        // c < columns and k < rows, so c * k < twN: no reduction needed
        v = v * reinterpret_cast<const cx<T>*>(P.tab_a)[c * k];
        if (P.conj_out) v = cj(v);
        out[k * es] = scl(v, s);
      } else {
        const long long j = (long long)c * P.out.cstride + (long long)k * es;
        if (j < P.N) {
          v = cj(v) * reinterpret_cast<const cx<T>*>(P.chirp)[j];
          if (P.conj_out) v = cj(v);
          reinterpret_cast<cx<T>*>(dst)[base - (long long)c * P.out.cstride + j] = scl(v, s);
        }
      }
    }
    return;
  }
  const int t = m.tr;
  if (t >= nt) return;
  const cx<T>* row = lds + t * P.Lp;
  auto Z = [&](int k) {
    cx<T> z = row[k];
    return P.conj_out ? cj(z) : z;
  };
  const long long obase = (long long)(g0 + t) * P.out.bdist;
  if constexpr (ST == ST_R2C) {
    // rfft1d.rs:134-148: N/2+1 bins per row
    const int N = (int)P.N, h = N >> 1, q = (N + 3) >> 2;
    const cx<T>* co = reinterpret_cast<const cx<T>*>(P.tab_a);
    cx<T>* out = reinterpret_cast<cx<T>*>(dst) + obase;
    for (int r = m.ur; r <= h; r += m.tpt) {
      cx<T> v;
      if (r == 0 || r == h) {
        const cx<T> z = Z(0);
        v = mk(r == 0 ? z.re + z.im : z.re - z.im, T(0));
      } else if ((N & 3) == 0 && r == q) {
        v = cj(Z(q));
      } else {
        const int i = (r < q) ? r : h - r;
        const cx<T> zi = Z(i), zh = Z(h - i);
        const cx<T> x = co[i] * (zi - cj(zh));
        v = (r < q) ? zi - x : zh + cj(x);
      }
      out[r] = scl(v, s);
    }
  } else if constexpr (ST == ST_DCT2) {
    // dct1d.rs:266-287
    const int N = (int)P.N, h = N >> 1, q2 = (N + 3) >> 2;
    const cx<T>* co = reinterpret_cast<const cx<T>*>(P.tab_a);
    const cx<T>* om = reinterpret_cast<const cx<T>*>(P.tab_b);
    T* ret = reinterpret_cast<T*>(dst) + obase;
    for (int i = m.ur; i <= q2; i += m.tpt) {
      if (i == 0) {
        const cx<T> z = Z(0);
        ret[0] = (z.re + z.im) * s;
        ret[h] = (z.re - z.im) * om[h].re * s;
      } else if (i < q2) {
        const cx<T> zi = Z(i);
        const cx<T> wc = cj(Z(h - i));
        const cx<T> x = co[i] * (zi - wc);
        const cx<T> j = cj(zi - x) * om[i];
        ret[i] = j.re * s;
        ret[N - i] = j.im * s;
        const cx<T> k = (wc + x) * om[h - i];
        ret[h - i] = k.re * s;
        ret[h + i] = k.im * s;
      } else if ((N & 3) == 0) {
        const cx<T> x = Z(q2) * om[q2];
        ret[q2] = x.re * s;
        ret[N - q2] = x.im * s;
      }
    }
  } else if constexpr (ST == ST_DCT3) {
    // dct1d.rs:395-407
    const int N = (int)P.N, h = N >> 1, q2 = h >> 1;
    const int per = q2 + (h & 1);
    T* ret = reinterpret_cast<T*>(dst) + obase;
    for (int i = m.ur; i < per; i += m.tpt) {
      const cx<T> a = Z(i);
      if (i < q2) {
        const cx<T> b = Z(h - 1 - i);
        ret[4 * i] = a.re * s;
        ret[4 * i + 1] = b.im * s;
        ret[4 * i + 2] = a.im * s;
        ret[4 * i + 3] = b.re * s;
      } else {
        ret[4 * i] = a.re * s;
        ret[4 * i + 1] = a.im * s;
      }
    }
    (void)N;
  } else if constexpr (ST == ST_MDCT) {
    // mdct1d.rs:234-240
    const int N = (int)P.N, q = N >> 2, h = N >> 1;
    const cx<T>* tw = reinterpret_cast<const cx<T>*>(P.tab_a);
    T* ret = reinterpret_cast<T*>(dst) + obase;
    for (int i = m.ur; i < q; i += m.tpt) {
      const cx<T> v = Z(i) * scl(tw[i], s);
      ret[2 * i] = -v.re;
      ret[h - 1 - 2 * i] = v.im;
    }
  } else {  // ST_IMDCT, mdct1d.rs:268-288
    const int N = (int)P.N, q = N >> 2, h = N >> 1, q3 = h + q, q5 = h + q3, lo = q - (q >> 1);
    const cx<T>* tw = reinterpret_cast<const cx<T>*>(P.tab_a);
    const T* w = reinterpret_cast<const T*>(P.tab_b);
    T* ret = reinterpret_cast<T*>(dst) + obase;
    for (int i = m.ur; i < q; i += m.tpt) {
      const cx<T> v = Z(i) * scl(tw[i], s);
      const int mm = 2 * i;
      ret[q3 - 1 - mm] = v.re * w[q3 - 1 - mm];
      ret[q + mm] = -v.im * w[q + mm];
      if (i < lo) {
        ret[q3 + mm] = v.re * w[q3 + mm];
        ret[q - 1 - mm] = v.im * w[q - 1 - mm];
      } else {
        ret[mm - q] = -v.re * w[mm - q];
        ret[q5 - 1 - mm] = -v.im * w[q5 - 1 - mm];
      }
    }
  }
}

// One LDS-resident batch of tpb transforms: load, passes, [Bluestein middle],
// store. The store stage is not followed by a barrier.
template <typename T, int LD, int ST>
__device__ __forceinline__ void block_transform(const KParams& P, const void* src, void* dst, cx<T>* lds, int g0,
                                                int nt, double scale) {
  const TMap m = make_map(P.tpb);
  load_stage<T, LD>(P, m, src, lds, g0, nt);
  run_passes<T>(lds, m, P.seq, P.Lp);
  if (P.has_mid) {
    mid_stage<T>(P, m, lds, g0);
    run_passes<T>(lds, m, P.seq, P.Lp);
    if (P.has_chirp) dechirp_stage<T>(P, m, lds);
  }
  store_stage<T, ST>(P, m, dst, lds, g0, nt, scale);
}

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
