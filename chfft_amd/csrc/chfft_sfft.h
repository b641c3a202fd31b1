// Length-specialised transform tiles for gfx950 (device templates; the
// kernels and their registry are in chfft_sfft.hip, the multi-length
// kernel in chfft_multi.hip also runs these tiles).
//
// Same job as fused_fft_kernel (chfft_kernels.hip) — one workgroup owns TPB
// transforms of length L — but with the length, the radix schedule and the
// thread layout fixed at compile time:
//   * each thread keeps the inputs of its butterflies in registers; the first
//     pass gathers them straight from global memory through the load operator
//     (coalesced: butterfly i of thread u reads element i + j*L/R0 with
//     i = u + b*TPT), the last pass stores straight to global memory through
//     an element-wise store operator (position i + j*L/R_last), or through the
//     LDS epilogue for operators that pair bins k and N-k;
//   * passes exchange through LDS once each, ping-ponging between two padded
//     buffers when they fit (one barrier per exchange), else one buffer with
//     two barriers;
//   * all index arithmetic (i mod p, twiddle offsets, strides) is constant-
//     folded; twiddles come from per-pass tables laid out [j-1][k] so that
//     consecutive lanes read consecutive entries.
// Fused Bluestein (BLUE) multiplies by the kernel spectrum in registers at the
// last-pass positions, conjugates, re-enters the first-pass layout through one
// exchange, runs the passes again and de-chirps in registers.
//
// Reference correspondence: as for the generic kernels — the passes replace
// mixed_radix.rs:104-368 / chirpz.rs:27-85, the operators the pre/post loops
// of chirpz.rs, rfft1d.rs, dct1d.rs and mdct1d.rs.
#pragma once
#include <type_traits>

#include "chfft_device.h"

namespace chfft {

#define COMMA ,

template <int... RS>
struct Rad {};

template <int... RS>
constexpr int rad_count(Rad<RS...>) {
  return (int)sizeof...(RS);
}
template <int... RS>
constexpr int rad_at(Rad<RS...>, int q) {
  const int a[] = {RS...};
  return a[q];
}
template <int... RS>
constexpr int rad_span(Rad<RS...> r, int q) {
  int p = 1;
  for (int i = 0; i < q; i++) p *= rad_at(r, i);
  return p;
}
// offset of pass q's twiddle table: passes with span p > 1 store (R-1)*p entries
template <int... RS>
constexpr int rad_twoff(Rad<RS...> r, int q) {
  int off = 0;
  for (int i = 0; i < q; i++)
    if (rad_span(r, i) > 1) off += (rad_at(r, i) - 1) * rad_span(r, i);
  return off;
}

// offset of pass q's compact table (tw_compact_entries) from the start of
// the compact region, which follows the full tables
template <int... RS>
constexpr int rad_compact_off(Rad<RS...> r, int q) {
  int off = 0;
  for (int i = 0; i < q; i++) off += tw_compact_entries(rad_at(r, i), rad_span(r, i));
  return off;
}

constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }

// Register budgets (bytes per thread) for twiddles fetched early: a pass's
// table entries one pass ahead, and the four-step output twiddles of a
// direct store together with the data.
#ifndef CHFFT_TW_AHEAD_BYTES
#define CHFFT_TW_AHEAD_BYTES 128
#endif
#ifndef CHFFT_TW_EARLY_BYTES
#define CHFFT_TW_EARLY_BYTES 128
#endif
// Per-pass twiddle tables up to this size (per precision) are copied into
// LDS once per tile and read from there (shorter latency than the L2, and
// off the vector memory path).
#ifndef CHFFT_TW_LDS_BYTES_F32
#define CHFFT_TW_LDS_BYTES_F32 8192
#endif
#ifndef CHFFT_TW_LDS_BYTES_F64
#define CHFFT_TW_LDS_BYTES_F64 2048
#endif
constexpr bool tw_in_lds(int entries, int esz) {
  return entries > 0 && entries * esz <= (esz == 8 ? CHFFT_TW_LDS_BYTES_F32 : CHFFT_TW_LDS_BYTES_F64);
}
// LDS padding: one slot per 2^pad_shift elements. A single transform longer
// than 16384 points (a whole CU's LDS for one sequence) takes one slot per 64
// so that its padded image still fits 160 KB; every other tile one per 16.
constexpr int pad_shift(int L, int tpb) { return (tpb == 1 && L > 16384) ? 6 : 4; }
constexpr int lds_stride(int L, int tpb) {
  const int ps = pad_shift(L, tpb);
  const int lp = L + ((L + (1 << ps) - 1) >> ps);
  return (tpb > 1 && (lp & 1) == 0) ? lp + 1 : lp;
}
// LDS (complex elements) of the MDCT fold's shared tables: the window (4L
// reals, as 2L pairs) and the L fold twiddles, read by every transform of a
// tile (mdct1d.rs:206-240). FP32 only: measured on MI355X (2048-point MDCT
// over 2^24 samples) FP32 50.5 -> 41.3 us, FP64 70 -> 75 us (its tables
// are twice the LDS), so FP64 keeps reading them from the caches.
constexpr int mdct_tab_elems(int L, int ld, int esz) { return ld == LD_MDCT && esz == 8 ? 3 * L : 0; }

// LDS bytes of one tile: one padded buffer, or two (ping-pong) when both
// fit in 16 KB, plus the per-pass twiddle tables when they are kept in LDS,
// plus `extra` complex elements of operator tables.
constexpr size_t lds_need(int L, int tpb, int esz, int ntw, int extra = 0) {
  const size_t one = (size_t)tpb * lds_stride(L, tpb) * esz;
  return (2 * one <= 16384 ? 2 * one : one) + (tw_in_lds(ntw, esz) ? (size_t)ntw * esz : 0) + (size_t)extra * esz;
}

// Workgroup barrier for LDS traffic only: waits for this wave's LDS accesses
// (lgkmcnt) but leaves its global loads in flight, so that twiddles fetched
// for a later pass keep arriving across the exchange. The "memory" clobber
// keeps the compiler from moving memory accesses across it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Global access at a 32-bit byte offset from a workgroup-uniform base: the
// backend then uses the scalar-base + vector-offset form of the global
// instructions, with no 64-bit address arithmetic per element. The planner
// checks that every tile's offsets fit (tile_span_bytes).
template <typename V>
__device__ __forceinline__ V ld_off(const void* base, unsigned off) {
  return *reinterpret_cast<const V*>(reinterpret_cast<const char*>(base) + off);
}
template <typename V, bool NT = false>
__device__ __forceinline__ void st_off(void* base, unsigned off, V v) {
  gput<NT>(reinterpret_cast<V*>(reinterpret_cast<char*>(base) + off), v);
}

// Padded LDS index of element base + j * S: linear in j when S is a multiple
// of 16, so that the offsets fold into the instructions' immediates.
template <int S>
__device__ __forceinline__ int lxs(int base, int j) {
  if constexpr (S % 16 == 0)
    return lx(base) + j * (S + S / 16);
  else
    return lx(base + j * S);
}

// LDS layouts of a tile. Rows: transform t's element n at t * LP + lx(n)
// (padded rows; lanes of one instruction walk one transform). Interleaved:
// element n of transform t at n * TPB + t, for tiles whose consecutive lanes
// take consecutive transforms: a lane group then covers whole bank rows
// without padding, and every offset j * S of a butterfly folds into the
// instruction's immediate.
struct LdsRows {
  template <int S>
  __device__ __forceinline__ static int at(int base, int j) {
    return lxs<S>(base, j);
  }
};
// Rows padded one slot per 64 elements (pad_shift 6): element n at n + n / 64.
struct LdsRows64 {
  template <int S>
  __device__ __forceinline__ static int at(int base, int j) {
    if constexpr (S % 64 == 0)
      return base + (base >> 6) + j * (S + S / 64);
    else
      return (base + j * S) + ((base + j * S) >> 6);
  }
};
template <int TPB>
struct LdsInterleaved {
  template <int S>
  __device__ __forceinline__ static int at(int base, int j) {
    return (base + j * S) * TPB;
  }
};

template <class RS, int L, int TPT>
struct Plan {
  static constexpr int LEN = L;
  // long single-transform tiles hold several butterflies per thread in a
  // full register budget: their butterflies are computed one after the
  // other (a scheduling fence between them) instead of interleaved
  static constexpr bool SERIAL = L > 16384;
  static constexpr int NP = rad_count(RS{});
  static constexpr int R(int q) { return rad_at(RS{}, q); }
  static constexpr int P(int q) { return rad_span(RS{}, q); }
  static constexpr int NB(int q) { return L / R(q); }
  static constexpr int B(int q) { return cdiv(NB(q), TPT); }
  static constexpr bool FULL(int q) { return NB(q) % TPT == 0; }
  static constexpr int EMAX() {
    int e = 0;
    for (int q = 0; q < NP; q++) e = (B(q) * R(q) > e) ? B(q) * R(q) : e;
    return e;
  }
};

// ---------------------------------------------------------------- pass pieces
// Twiddles of pass q in registers, R-1 per butterfly group. When TPT is a
// multiple of the span p, every butterfly of a thread has the same
// k = i mod p and one group serves them all. k < p for any i, so the loads
// need no guard and are issued all together, ahead of the butterflies.
template <typename T, class PL, class RS, int q, int TPT>
struct TwReg {
  static constexpr int R = PL::R(q), p = PL::P(q), B = PL::B(q);
  static constexpr int NG0 = p <= 1 ? 0 : (TPT % p == 0 ? 1 : B);
  // LAZY (long single-transform tiles): a pass whose butterflies need
  // different twiddles reads them per butterfly inside pass_compute instead
  // of holding the whole set in registers
  static constexpr bool LAZY = PL::SERIAL && NG0 > 1;
  static constexpr int NG = LAZY ? 0 : NG0;
  static constexpr int N = NG * (R - 1);
  // small sets are fetched one pass early, before the previous pass computes
  static constexpr bool AHEAD = N * (int)sizeof(cx<T>) <= CHFFT_TW_AHEAD_BYTES;
  cx<T> w[N > 0 ? N : 1];
  // LDT: tw is the LDS copy of the compact tables (two-level passes multiply
  // their two factors); else the full tables in global memory
  template <bool LDT>
  __device__ __forceinline__ void load(int u, const cx<T>* __restrict__ tw) {
    constexpr int off = rad_twoff(RS{}, q);
    constexpr int coff = rad_compact_off(RS{}, q);
    constexpr bool TWO = tw_two_level(R, p);
#pragma unroll
    for (int g = 0; g < NG; g++) {
      const unsigned k = (unsigned)(u + g * TPT) % (unsigned)p;
#pragma unroll
      for (int j = 1; j < R; j++) {
        if constexpr (LDT && TWO)
          w[g * (R - 1) + j - 1] = tw[coff + (j - 1) * 16 + (k & 15)] *
                                   tw[coff + (R - 1) * 16 + (j - 1) * (p / 16) + (k >> 4)];
        else if constexpr (LDT)
          w[g * (R - 1) + j - 1] = tw[coff + (j - 1) * p + k];
        else
          w[g * (R - 1) + j - 1] = ld_off<cx<T>>(tw + off + (j - 1) * p, k * (unsigned)sizeof(cx<T>));
      }
    }
  }
};

template <typename T, class PL, class RS, int q, int TPT>
__device__ __forceinline__ void pass_compute(cx<T>* x, int u, const TwReg<T, PL, RS, q, TPT>& W,
                                             const cx<T>* __restrict__ tw) {
  using TR = TwReg<T, PL, RS, q, TPT>;
  constexpr int R = PL::R(q), NB = PL::NB(q), B = PL::B(q);
#pragma unroll
  for (int b = 0; b < B; b++) {
    const int i = u + b * TPT;
    if (PL::FULL(q) || i < NB) {
      if constexpr (TR::LAZY) {
        // full table [j-1][k] of this pass in global memory, k = i mod p
        constexpr int off = rad_twoff(RS{}, q);
        const unsigned k = (unsigned)i % (unsigned)TR::p;
#pragma unroll
        for (int j = 1; j < R; j++)
          x[b * R + j] = x[b * R + j] * ld_off<cx<T>>(tw + off + (j - 1) * TR::p, k * (unsigned)sizeof(cx<T>));
      } else if constexpr (TR::NG > 0) {
        constexpr int gs = TR::NG == 1 ? 0 : 1;
#pragma unroll
        for (int j = 1; j < R; j++) x[b * R + j] = x[b * R + j] * W.w[gs * b * (R - 1) + j - 1];
      }
      dft<T, R>(x + b * R);
    }
    if constexpr (PL::SERIAL && B > 1) __builtin_amdgcn_sched_barrier(0);
  }
}

// Write the outputs of pass q at their Stockham positions.
template <typename T, class PL, int q, int TPT, class LY>
__device__ __forceinline__ void pass_write(const cx<T>* x, int u, cx<T>* row) {
  constexpr int R = PL::R(q), p = PL::P(q), NB = PL::NB(q), B = PL::B(q);
#pragma unroll
  for (int b = 0; b < B; b++) {
    const int i = u + b * TPT;
    if (PL::FULL(q) || i < NB) {
      const int k = i % p;
      const int o = (i - k) * R + k;
#pragma unroll
      for (int j = 0; j < R; j++) row[LY::template at<p>(o, j)] = x[b * R + j];
    }
    if constexpr (PL::SERIAL) __builtin_amdgcn_sched_barrier(0);
  }
}

// Read the inputs of pass q (element i + j*NB of butterfly i).
template <typename T, class PL, int q, int TPT, class LY>
__device__ __forceinline__ void pass_read(cx<T>* x, int u, const cx<T>* row) {
  constexpr int R = PL::R(q), NB = PL::NB(q), B = PL::B(q);
#pragma unroll
  for (int b = 0; b < B; b++) {
    const int i = u + b * TPT;
    if (PL::FULL(q) || i < NB) {
#pragma unroll
      for (int j = 0; j < R; j++) x[b * R + j] = row[LY::template at<NB>(i, j)];
    }
    if constexpr (PL::SERIAL) __builtin_amdgcn_sched_barrier(0);
  }
}

// Exchange from pass qa's outputs to pass qb's inputs through LDS.
template <typename T, class PL, int qa, int qb, int TPT, bool PP, class LY>
__device__ __forceinline__ void exchange(cx<T>* x, int u, cx<T>* row0, cx<T>* row1, int& cur) {
  cx<T>* row = (PP && cur) ? row1 : row0;
  pass_write<T, PL, qa, TPT, LY>(x, u, row);
  lds_barrier();
  pass_read<T, PL, qb, TPT, LY>(x, u, row);
  if constexpr (PP) {
    cur ^= 1;
  } else {
    lds_barrier();
  }
}

// Passes q.. of the sequence; W holds pass q's twiddles. From global
// memory, pass q+1's twiddles are fetched before pass q computes (small
// sets) or right after, so that their latency overlaps the exchange; from
// the LDS copy (LDT) they are read after the exchange's barrier, which also
// publishes the copy.
template <typename T, class PL, class RS, int q, int TPT, bool PP, class LY, bool LDT>
__device__ __forceinline__ void run_seq(cx<T>* x, int u, const cx<T>* __restrict__ tw, cx<T>* row0, cx<T>* row1,
                                        int& cur, const TwReg<T, PL, RS, q, TPT>& W) {
  if constexpr (q + 1 < PL::NP) {
    using TN = TwReg<T, PL, RS, q + 1, TPT>;
    TN Wn;
    if constexpr (!LDT && TN::AHEAD) Wn.template load<false>(u, tw);
    pass_compute<T, PL, RS, q, TPT>(x, u, W, tw);
    if constexpr (!LDT && !TN::AHEAD) Wn.template load<false>(u, tw);
    exchange<T, PL, q, q + 1, TPT, PP, LY>(x, u, row0, row1, cur);
    if constexpr (LDT) Wn.template load<true>(u, tw);
    run_seq<T, PL, RS, q + 1, TPT, PP, LY, LDT>(x, u, tw, row0, row1, cur, Wn);
  } else {
    pass_compute<T, PL, RS, q, TPT>(x, u, W, tw);
  }
}

template <typename T, class PL, class RS, int TPT, bool PP, class LY, bool LDT>
__device__ __forceinline__ void run_all(cx<T>* x, int u, const cx<T>* __restrict__ tw, cx<T>* row0, cx<T>* row1,
                                        int& cur) {
  const TwReg<T, PL, RS, 0, TPT> W0{};  // the first pass has span 1: no twiddles
  run_seq<T, PL, RS, 0, TPT, PP, LY, LDT>(x, u, tw, row0, row1, cur, W0);
}

// ---------------------------------------------------------------- store operators (element-wise)
template <typename T, int ST, bool NT = false>
__device__ __forceinline__ void store_elem(const KParams& P, void* dst, long long obase, int c, int k, cx<T> v, T s) {
  if constexpr (ST == ST_PLAIN) {
    if (P.conj_out) v = cj(v);
    gput<NT>(&reinterpret_cast<cx<T>*>(dst)[obase + (long long)k * P.out.estride], scl(v, s));
  } else if constexpr (ST == ST_TW) {
    v = v * reinterpret_cast<const cx<T>*>(P.tab_a)[k * P.twk + c * P.twc];
    if (P.conj_out) v = cj(v);
    gput<NT>(&reinterpret_cast<cx<T>*>(dst)[obase + (long long)k * P.out.estride], scl(v, s));
  } else if constexpr (ST == ST_FS_DECHIRP) {
    // four-step Bluestein last pass: element k of column c is sequence index
    // c + k * estride; de-chirp (chirpz.rs:127-145) and crop to N
    const long long j = (long long)c * P.out.cstride + (long long)k * P.out.estride;
    if (j < P.N) {
      v = cj(v) * reinterpret_cast<const cx<T>*>(P.chirp)[j];
      if (P.conj_out) v = cj(v);
      gput<NT>(&reinterpret_cast<cx<T>*>(dst)[obase - (long long)c * P.out.cstride + j], scl(v, s));
    }
  } else if constexpr (ST == ST_MDCT) {
    // mdct1d.rs:234-240
    if (P.conj_out) v = cj(v);
    const int h = (int)(P.N >> 1);
    const cx<T> w = v * scl(reinterpret_cast<const cx<T>*>(P.tab_a)[k], s);
    T* ret = reinterpret_cast<T*>(dst) + obase;
    gput<NT>(&ret[2 * k], -w.re);
    gput<NT>(&ret[h - 1 - 2 * k], w.im);
  } else if constexpr (ST == ST_IMDCT) {
    // mdct1d.rs:268-288
    if (P.conj_out) v = cj(v);
    const int N = (int)P.N, q = N >> 2, h = N >> 1, q3 = h + q, q5 = h + q3, lo = q - (q >> 1);
    const T* w = reinterpret_cast<const T*>(P.tab_b);
    const cx<T> z = v * scl(reinterpret_cast<const cx<T>*>(P.tab_a)[k], s);
    T* ret = reinterpret_cast<T*>(dst) + obase;
    const int mm = 2 * k;
    gput<NT>(&ret[q3 - 1 - mm], z.re * w[q3 - 1 - mm]);
    gput<NT>(&ret[q + mm], -z.im * w[q + mm]);
    if (k < lo) {
      gput<NT>(&ret[q3 + mm], z.re * w[q3 + mm]);
      gput<NT>(&ret[q - 1 - mm], z.im * w[q - 1 - mm]);
    } else {
      gput<NT>(&ret[mm - q], -z.re * w[mm - q]);
      gput<NT>(&ret[q5 - 1 - mm], -z.im * w[q5 - 1 - mm]);
    }
  }
}

// ---------------------------------------------------------------- the kernel
// MODE: 0 plain passes; 1 fused Bluestein (chirp on load, spectrum product,
// second pass sequence, de-chirp); 2 spectrum product + second sequence only
// (the row pass of a four-step Bluestein).
// Natural-order store of a tile from its final LDS image (epilogue path) for
// the element-wise operators PLAIN and TW; CM: consecutive threads on
// consecutive transforms (transposed stores write whole segments).
template <typename T, int L, int TPT, int TPB, int ST, bool CM, bool NT = false>
__device__ __forceinline__ void epi_store(const KParams& P, void* dst, const cx<T>* lds, int LP, int g0, int nt,
                                          T s) {
  const int tid = threadIdx.x;
  const int t = CM ? tid % TPB : tid / TPT;
  const int u = CM ? tid / TPB : tid % TPT;
  if (t >= nt) return;
  const cx<T>* row = lds + t * LP;
  constexpr unsigned ESZ = sizeof(cx<T>);
  int c0, c;
  const long long o0 = io_base(P.out, g0, &c0);
  const long long og = io_base(P.out, g0 + t, &c);
  char* ob = reinterpret_cast<char*>(dst) + o0 * (long long)ESZ;
  const unsigned es = (unsigned)P.out.estride * ESZ;
  const unsigned voff = (unsigned)(og - o0) * ESZ + (unsigned)u * es;
  const unsigned tk = (unsigned)P.twk * ESZ;
  const unsigned toff = ((unsigned)u * (unsigned)P.twk + (unsigned)c * (unsigned)P.twc) * ESZ;
  const bool cjo = P.conj_out != 0;
  constexpr int S = cdiv(L, TPT);
#pragma unroll
  for (int r = 0; r < S; r++) {
    if (L % TPT == 0 || u + r * TPT < L) {
      cx<T> v = row[lxs<TPT>(u, r)];
      if constexpr (ST == ST_TW) v = v * ld_off<cx<T>>(reinterpret_cast<const char*>(P.tab_a) + (unsigned)(r * TPT) * tk, toff);
      if (cjo) v = cj(v);
      st_off<cx<T>, NT>(ob + (unsigned)(r * TPT) * es, voff, scl(v, s));
    }
  }
}

// DCT-II post-processing (dct1d.rs:266-287) from the final LDS image of a
// tile of half length L (N = 2L): thread u of transform t takes the pairs
// i = u + r * TPT, i <= L/2, and writes bins i, N-i, L-i and L+i; tables and
// the row are addressed by 32-bit offsets.
template <typename T, int L, int TPT, int TPB, bool NT = false>
__device__ __forceinline__ void epi_dct2(const KParams& P, void* dst, const cx<T>* lds, int LP, int g0, int nt, T s) {
  constexpr int N = 2 * L, h = L, q2 = L / 2;  // N % 4 == 0 when L is even
  const int tid = threadIdx.x;
  const int t = tid / TPT, u = tid % TPT;
  if (t >= nt) return;
  const cx<T>* row = lds + t * LP;
  const bool cjo = P.conj_out != 0;
  auto Z = [&](int k) {
    const cx<T> z = row[lx(k)];
    return cjo ? cj(z) : z;
  };
  const cx<T>* co = reinterpret_cast<const cx<T>*>(P.tab_a);
  const cx<T>* om = reinterpret_cast<const cx<T>*>(P.tab_b);
  T* ret = reinterpret_cast<T*>(dst) + (long long)(g0 + t) * P.out.bdist;
  constexpr unsigned CE = sizeof(cx<T>), RE = sizeof(T);
  constexpr int NR = cdiv(q2 + 1, TPT);
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int i = u + r * TPT;
    if (r * TPT + TPT - 1 <= q2 || i <= q2) {
      if (i == 0) {
        const cx<T> z = Z(0);
        ret[0] = (z.re + z.im) * s;
        ret[h] = (z.re - z.im) * om[h].re * s;
      } else if (i < q2) {
        const cx<T> zi = Z(i);
        const cx<T> wc = cj(Z(h - i));
        const cx<T> x = ld_off<cx<T>>(co, (unsigned)i * CE) * (zi - wc);
        const cx<T> jj = cj(zi - x) * ld_off<cx<T>>(om, (unsigned)i * CE);
        const cx<T> kk = (wc + x) * ld_off<cx<T>>(om, (unsigned)(h - i) * CE);
        st_off<T, NT>(ret, (unsigned)i * RE, jj.re * s);
        st_off<T, NT>(ret, (unsigned)(N - i) * RE, jj.im * s);
        st_off<T, NT>(ret, (unsigned)(h - i) * RE, kk.re * s);
        st_off<T, NT>(ret, (unsigned)(h + i) * RE, kk.im * s);
      } else {  // i == q2
        const cx<T> x = Z(q2) * om[q2];
        ret[q2] = x.re * s;
        ret[N - q2] = x.im * s;
      }
    }
  }
}

// DCT-III output (dct1d.rs:395-407) from the final LDS image, half length L
// even: pair i < L/2 writes ret[4i .. 4i+3] = (a.re, b.im, a.im, b.re) with
// a = Z(i), b = Z(L-1-i), as two aligned two-real stores.
template <typename T, int L, int TPT, int TPB, bool NT = false>
__device__ __forceinline__ void epi_dct3(const KParams& P, void* dst, const cx<T>* lds, int LP, int g0, int nt, T s) {
  constexpr int h = L, q2 = L / 2;
  const int tid = threadIdx.x;
  const int t = tid / TPT, u = tid % TPT;
  if (t >= nt) return;
  const cx<T>* row = lds + t * LP;
  const bool cjo = P.conj_out != 0;
  T* ret = reinterpret_cast<T*>(dst) + (long long)(g0 + t) * P.out.bdist;
  constexpr unsigned RE = sizeof(T);
  constexpr int NR = cdiv(q2, TPT);
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int i = u + r * TPT;
    if (q2 % TPT == 0 || i < q2) {
      cx<T> a = row[lx(i)], b = row[lx(h - 1 - i)];
      if (cjo) {
        a = cj(a);
        b = cj(b);
      }
      st_off<cx<T>, NT>(ret, (unsigned)(4 * i) * RE, mk(a.re * s, b.im * s));
      st_off<cx<T>, NT>(ret, (unsigned)(4 * i + 2) * RE, mk(a.im * s, b.re * s));
    }
  }
}

// First-pass gather through the generic load operator (any LD, per element).
template <typename T, int L, int TPT, int TPB, bool COL, int LD, int MODE, class RS>
__device__ __forceinline__ void gather_generic(const KParams& P, const void* src, int g0,
                                               cx<T> (&x)[Plan<RS, L, TPT>::EMAX()]) {
  using PL = Plan<RS, L, TPT>;
  constexpr bool BLUE = MODE == 1;
  const int tid = threadIdx.x;
  const int t = COL ? tid % TPB : tid / TPT;
  const int u = COL ? tid / TPB : tid % TPT;
  const int g = g0 + t;
  const bool valid = g < P.G;
  const cx<T>* chirp = reinterpret_cast<const cx<T>*>(P.chirp);
  constexpr int R0 = PL::R(0), NB0 = PL::NB(0), B0 = PL::B(0);
  const Loader<T, LD> ld = make_loader<T, LD>(P, src, valid ? g : 0);
#pragma unroll
  for (int b = 0; b < B0; b++) {
    const int i = u + b * TPT;
    if (PL::FULL(0) || i < NB0) {
#pragma unroll
      for (int j = 0; j < R0; j++) {
        const int n = i + j * NB0;
        cx<T> v = mk(T(0), T(0));
        if (valid && (!BLUE || n < P.L)) {
          v = load_elem<T, LD>(P, ld, n);
          if constexpr (LD != LD_FS_CHIRP) {
            if (P.conj_in) v = cj(v);
          }
          if constexpr (BLUE) v = v * chirp[n];
        }
        x[b * R0 + j] = v;
      }
    }
  }
}

// Fused Bluestein pads N points to M >= 2N, so the upper half of the first
// sequence is zero and only outputs below M / 2 are kept. In a pass of even
// radix R over L points, slot j of a butterfly (element i + j L / R, i < L / R)
// lies in the lower half exactly when 2 j < R: a compile-time test, so the
// compiler drops the zero inputs' arithmetic in the first pass and the unused
// outputs' in the last.
template <int R>
__device__ __forceinline__ constexpr bool lower_half(int j) {
  return R % 2 != 0 || 2 * j < R;
}

// First-pass gather of one tile: element i + j*L/R0 of butterfly i = u + b*TPT
// of transform g0 + t.
// Paired-real fast paths need the real buffer 8-byte (FP32) / 16-byte (FP64)
// aligned with an even distance between transforms. AL selects how a kernel
// decides: 2 tests at run time, 1 takes the fast paths (the kernel has
// tested every such condition once, up front), 0 takes the per-element
// paths. Separate AL = 1 and AL = 0 bodies keep each one's register arrays
// apart (one body holding both paths kept some of them in scratch memory).
template <int AL>
__host__ __device__ __forceinline__ bool pair_aligned(const void* p, long long bdist, size_t esz) {
  if constexpr (AL == 2) return (reinterpret_cast<uintptr_t>(p) & (esz - 1)) == 0 && (bdist & 1) == 0;
  return AL == 1;
}

template <typename T, int L, int TPT, int TPB, bool COL, int LD, int MODE, class RS, int AL>
__device__ __forceinline__ void sfft_gather(const KParams& P, const void* src, int g0,
                                            cx<T> (&x)[Plan<RS, L, TPT>::EMAX()], cx<T>* mtab) {
  using PL = Plan<RS, L, TPT>;
  constexpr bool BLUE = MODE == 1;
  constexpr unsigned ESZ = sizeof(cx<T>);
  const int tid = threadIdx.x;
  const int t = COL ? tid % TPB : tid / TPT;
  const int u = COL ? tid / TPB : tid % TPT;
  const int g = g0 + t;
  const bool valid = g < P.G;
  const cx<T>* chirp = reinterpret_cast<const cx<T>*>(P.chirp);
  constexpr int R0 = PL::R(0), NB0 = PL::NB(0), B0 = PL::B(0);
  if constexpr (LD == LD_PLAIN) {
    // threads past the last transform re-read the tile's first one (their
    // results are never stored), so every load stays in bounds unguarded
    int c0, c;
    const long long b0 = io_base(P.in, g0, &c0);
    const long long bg = io_base(P.in, valid ? g : g0, &c);
    const char* tb = reinterpret_cast<const char*>(src) + b0 * (long long)ESZ;
    const unsigned es = (unsigned)P.in.estride * ESZ;
    const unsigned voff = (unsigned)(bg - b0) * ESZ + (unsigned)u * es;
    const bool cji = P.conj_in != 0;
#pragma unroll
    for (int b = 0; b < B0; b++) {
      const int i = u + b * TPT;
      if (PL::FULL(0) || i < NB0) {
#pragma unroll
        for (int j = 0; j < R0; j++) {
          const int rel = b * TPT + j * NB0;  // element n = u + rel
          cx<T> v = mk(T(0), T(0));
          if (!BLUE || (lower_half<R0>(j) && u + rel < P.L)) {
            v = ld_off<cx<T>>(tb + (unsigned)rel * es, voff);
            if (cji) v = cj(v);
            if constexpr (BLUE) v = v * ld_off<cx<T>>(chirp + rel, (unsigned)u * ESZ);
          }
          x[b * R0 + j] = v;
        }
      }
    }
  } else if constexpr (LD == LD_FS_CHIRP) {
    // four-step Bluestein columns: element n of column c is sequence index
    // s = c * cstride + n * estride of its transform; x[s] * chirp[s] for
    // s < N, zero padding beyond (chirpz.rs:97-113, the inverse conjugating
    // the input first). 32-bit sequence offsets; the chirp is read from a
    // uniform base.
    int c;
    const long long base = io_base(P.in, valid ? g : g0, &c);
    const long long seq0 = (long long)c * P.in.cstride;
    const cx<T>* row = reinterpret_cast<const cx<T>*>(src) + (base - seq0);
    const unsigned es = (unsigned)P.in.estride;
    const unsigned s0 = (unsigned)seq0 + (unsigned)u * es;
    const unsigned N = (unsigned)P.N;
    const bool cji = P.conj_in != 0;
#pragma unroll
    for (int b = 0; b < B0; b++) {
      const int i = u + b * TPT;
      if (PL::FULL(0) || i < NB0) {
#pragma unroll
        for (int j = 0; j < R0; j++) {
          const unsigned sq = s0 + (unsigned)(b * TPT + j * NB0) * es;
          cx<T> v = mk(T(0), T(0));
          if (sq < N) {
            v = ld_off<cx<T>>(row, sq * ESZ);
            if (cji) v = cj(v);
            v = v * ld_off<cx<T>>(chirp, sq * ESZ);
          }
          x[b * R0 + j] = v;
        }
      }
    }
  } else if constexpr (LD == LD_DCT2 && !COL && !BLUE && B0 == 1 && NB0 == TPT && R0 % 2 == 0 &&
                       (TPT & (TPT - 1)) == 0 && TPT <= 64) {
    // DCT-II pack (dct1d.rs:245-263) for an even half length L = N/2 whose
    // first pass gives thread u the elements n = u + j * TPT: element n < N/4
    // is (x[4n], x[4n+2]) and element L-1-n is (x[4n+3], x[4n+1]), both from
    // the same four reals. Lane u reads the four reals of its slots j < R0/2
    // (two contiguous complex-sized loads per slot), keeps element n and hands
    // element L-1-n -- slot R0-1-j of lane TPT-1-u, the mirror lane of the
    // same transform -- across with one lane shuffle per real.
    if (pair_aligned<AL>(src, P.in.bdist, sizeof(cx<T>))) {  // uniform
      const T* xr = reinterpret_cast<const T*>(src) + (long long)(valid ? g : g0) * P.in.bdist;
      const T o = (T)P.ortho;
      const bool cji = P.conj_in != 0;
#pragma unroll
      for (int j = 0; j < R0 / 2; j++) {
        const unsigned n4 = 4u * (unsigned)(u + j * TPT);
        const cx<T> lo = ld_off<cx<T>>(xr, n4 * (unsigned)sizeof(T));
        const cx<T> hi = ld_off<cx<T>>(xr, (n4 + 2) * (unsigned)sizeof(T));
        cx<T> mine = mk(lo.re, hi.re);
        if (j == 0 && u == 0) mine.re = lo.re * o;  // element 0 carries the ortho factor
        const cx<T> other = mk(__shfl_xor(hi.im, TPT - 1), __shfl_xor(lo.im, TPT - 1));
        x[j] = cji ? cj(mine) : mine;
        x[R0 - 1 - j] = cji ? cj(other) : other;
      }
    } else {
      gather_generic<T, L, TPT, TPB, COL, LD, MODE, RS>(P, src, g0, x);
    }
  } else if constexpr (LD == LD_DCT3 && !COL && !BLUE && B0 == 1 && NB0 == TPT && R0 % 2 == 0 &&
                       (TPT & (TPT - 1)) == 0 && TPT <= 64 && L % 2 == 0) {
    // DCT-III pre-processing (dct1d.rs:367-391), half length L = N/2 even,
    // q = L/2: elements n and L-n (0 < n < q) come from the same i = n, so
    // lane u computes both for its slots j < R0/2 (n = u + j * TPT < q) and
    // element L-n goes to the lane that holds it: lane TPT-u at slot
    // R0-1-j, or lane 0 itself at slot R0-j when u = 0. Element 0 and
    // element q have their own formulas.
    const T* xr = reinterpret_cast<const T*>(src) + (long long)(valid ? g : g0) * P.in.bdist;
    const cx<T>* co = reinterpret_cast<const cx<T>*>(P.tab_a);
    const cx<T>* om = reinterpret_cast<const cx<T>*>(P.tab_b);
    constexpr int N = 2 * L, h = L, q = L / 2;
    constexpr unsigned RE = sizeof(T), CE = sizeof(cx<T>);
    const bool cji = P.conj_in != 0;
    const int src_lane = (TPT - u) & (TPT - 1);
    cx<T> mine[R0 / 2], other[R0 / 2];
#pragma unroll
    for (int j = 0; j < R0 / 2; j++) {
      const int i = u + j * TPT;
      const int ni = i == 0 ? 0 : N - i;  // element 0 has its own formula; keep its loads in the row
      const cx<T> a = mk(ld_off<T>(xr, (unsigned)i * RE), ld_off<T>(xr, (unsigned)ni * RE));
      const cx<T> bb = mk(ld_off<T>(xr, (unsigned)(h - i) * RE), ld_off<T>(xr, (unsigned)(h + i) * RE));
      const cx<T> jv = cj(a * ld_off<cx<T>>(om, (unsigned)i * CE));
      const cx<T> kv = bb * ld_off<cx<T>>(om, (unsigned)(h - i) * CE);
      const cx<T> xx = ld_off<cx<T>>(co, (unsigned)i * CE) * (jv - kv);
      mine[j] = jv - xx;
      other[j] = cj(kv) + cj(xx);
    }
    if (u == 0) {  // element 0
      const T z = xr[0] * (T)P.ortho;
      const T hv = xr[h] * om[h].re;
      mine[0] = mk(z + hv, z - hv);
    }
#pragma unroll
    for (int j = 0; j < R0 / 2; j++) {
      x[j] = cji ? cj(mine[j]) : mine[j];
      const int J = R0 - 1 - j;  // slots R0/2 .. R0-1 receive element L - n'
      const cx<T> got = mk(__shfl(other[j].re, src_lane, TPT), __shfl(other[j].im, src_lane, TPT));
      cx<T> v = got;
      if (u == 0) {
        if (J == R0 / 2)  // element q
          v = mk(xr[q], xr[N - q]) * om[q];
        else
          v = other[R0 - J];
      }
      x[J] = cji ? cj(v) : v;
    }
  } else if constexpr (LD == LD_MDCT && !COL && !BLUE && B0 == 1 && NB0 == TPT && R0 % 2 == 0 &&
                       (TPT & (TPT - 1)) == 0 && TPT <= 64) {
    // MDCT fold (mdct1d.rs:206-230), frame length N = 4L: element n < L/2
    // reads the reals at q3-1-m, q3+m, q+m, q-1-m (m = 2n, q = L, q3 = 3L)
    // and element L-1-n their neighbours q+1+m, q-2-m, q3-2-m, q3+1+m, so
    // four aligned pairs of reals (and of window values) serve both. Lane u
    // folds both elements of its slots j < R0/2 and hands element L-1-n to
    // the mirror lane TPT-1-u (slot R0-1-j) with one shuffle per real.
    // The window and the fold twiddles are the same for every transform of
    // the tile: in FP32 the workgroup copies them into LDS (mtab) once, after
    // its signal loads are in flight, and every lane reads them from there.
    if (pair_aligned<AL>(src, P.in.bdist, sizeof(cx<T>))) {  // uniform
      const T* xr = reinterpret_cast<const T*>(src) + (long long)(valid ? g : g0) * P.in.bdist;
      const cx<T>* wr = reinterpret_cast<const cx<T>*>(P.tab_b);  // window pairs
      const cx<T>* tw = reinterpret_cast<const cx<T>*>(P.tab_a);
      constexpr int q = L, q3 = 3 * L;
      constexpr unsigned RE = sizeof(T), CE = sizeof(cx<T>);
      const bool cji = P.conj_in != 0;
      if constexpr (mdct_tab_elems(L, LD, (int)sizeof(cx<T>)) > 0) {
        cx<T> pr[R0 / 2][4];
#pragma unroll
        for (int j = 0; j < R0 / 2; j++) {
          const int m = 2 * (u + j * TPT);
          pr[j][0] = ld_off<cx<T>>(xr, (unsigned)(q + m) * RE);       // (A3, B1)
          pr[j][1] = ld_off<cx<T>>(xr, (unsigned)(q - 2 - m) * RE);   // (B2, A4)
          pr[j][2] = ld_off<cx<T>>(xr, (unsigned)(q3 - 2 - m) * RE);  // (B3, A1)
          pr[j][3] = ld_off<cx<T>>(xr, (unsigned)(q3 + m) * RE);      // (A2, B4)
        }
        constexpr int NT = 3 * L, TH = TPT * TPB;
        const int tid = threadIdx.x;
#pragma unroll
        for (int e0 = 0; e0 < NT; e0 += TH)
          if (e0 + tid < NT)
            mtab[e0 + tid] = e0 + tid < 2 * L ? ld_off<cx<T>>(wr, (unsigned)(e0 + tid) * CE)
                                              : ld_off<cx<T>>(tw, (unsigned)(e0 + tid - 2 * L) * CE);
        lds_barrier();
        const cx<T>* wl = mtab;          // window pair i holds reals 2i, 2i + 1
        const cx<T>* tl = mtab + 2 * L;  // fold twiddles
#pragma unroll
        for (int j = 0; j < R0 / 2; j++) {
          const int n = u + j * TPT, m = 2 * n;
          const cx<T> p1 = pr[j][0], p2 = pr[j][1], p3 = pr[j][2], p4 = pr[j][3];
          const cx<T> w1 = wl[(q + m) / 2];
          const cx<T> w2 = wl[(q - 2 - m) / 2];
          const cx<T> w3 = wl[(q3 - 2 - m) / 2];
          const cx<T> w4 = wl[(q3 + m) / 2];
          cx<T> va = mk(p3.im * w3.im + p4.re * w4.re, p1.re * w1.re - p2.im * w2.im);
          cx<T> vb = mk(p1.im * w1.im - p2.re * w2.re, p3.re * w3.re + p4.im * w4.im);
          va = va * tl[n];
          vb = vb * tl[L - 1 - n];
          const cx<T> other = mk(__shfl_xor(vb.re, TPT - 1), __shfl_xor(vb.im, TPT - 1));
          x[j] = cji ? cj(va) : va;
          x[R0 - 1 - j] = cji ? cj(other) : other;
        }
      } else {
        (void)mtab;
#pragma unroll
        for (int j = 0; j < R0 / 2; j++) {
          const int n = u + j * TPT, m = 2 * n;
          const cx<T> p1 = ld_off<cx<T>>(xr, (unsigned)(q + m) * RE);       // (A3, B1)
          const cx<T> p2 = ld_off<cx<T>>(xr, (unsigned)(q - 2 - m) * RE);   // (B2, A4)
          const cx<T> p3 = ld_off<cx<T>>(xr, (unsigned)(q3 - 2 - m) * RE);  // (B3, A1)
          const cx<T> p4 = ld_off<cx<T>>(xr, (unsigned)(q3 + m) * RE);      // (A2, B4)
          const cx<T> w1 = ld_off<cx<T>>(wr, (unsigned)(q + m) * RE);
          const cx<T> w2 = ld_off<cx<T>>(wr, (unsigned)(q - 2 - m) * RE);
          const cx<T> w3 = ld_off<cx<T>>(wr, (unsigned)(q3 - 2 - m) * RE);
          const cx<T> w4 = ld_off<cx<T>>(wr, (unsigned)(q3 + m) * RE);
          cx<T> va = mk(p3.im * w3.im + p4.re * w4.re, p1.re * w1.re - p2.im * w2.im);
          cx<T> vb = mk(p1.im * w1.im - p2.re * w2.re, p3.re * w3.re + p4.im * w4.im);
          va = va * ld_off<cx<T>>(tw, (unsigned)n * CE);
          vb = vb * ld_off<cx<T>>(tw, (unsigned)(L - 1 - n) * CE);
          const cx<T> other = mk(__shfl_xor(vb.re, TPT - 1), __shfl_xor(vb.im, TPT - 1));
          x[j] = cji ? cj(va) : va;
          x[R0 - 1 - j] = cji ? cj(other) : other;
        }
      }
    } else {
      gather_generic<T, L, TPT, TPB, COL, LD, MODE, RS>(P, src, g0, x);
    }
  } else {
    gather_generic<T, L, TPT, TPB, COL, LD, MODE, RS>(P, src, g0, x);
  }
}

// Four-step twiddles of a direct store, W^(k*twk + c*twc) for the outputs
// k = u + rel of a thread whose transform is column c: fetched with the data
// when they fit a small register budget (their latency then hides behind
// the passes), else read at the store.
template <typename T, class PL, int TPT, int ST, bool EPI>
struct StoreTw {
  static constexpr int NP = PL::NP, RL = PL::R(NP - 1), PLAST = PL::P(NP - 1), BL = PL::B(NP - 1);
  static constexpr bool EARLY = ST == ST_TW && !EPI && BL * RL * sizeof(cx<T>) <= CHFFT_TW_EARLY_BYTES;
  cx<T> w[EARLY ? BL * RL : 1];
  __device__ __forceinline__ void load(const KParams& P, int u, int c) {
    if constexpr (EARLY) {
      if (P.tw_rec) {
        // W^(c k) for k = u + b TPT + j PLAST: base W^(c (u + b TPT)) times
        // j steps W^(c PLAST), in FP64, each result rounded once
        constexpr int K0 = BL * TPT;
        const cx<double>* tr = reinterpret_cast<const cx<double>*>(P.tw_rec);
        const int nc = P.out.ncols;
        const cx<double> st = tr[K0 * nc + c];
#pragma unroll
        for (int b = 0; b < BL; b++) {
          if (PL::FULL(NP - 1) || u + b * TPT < PL::NB(NP - 1)) {
            cx<double> wd = tr[(u + b * TPT) * nc + c];
#pragma unroll
            for (int j = 0; j < RL; j++) {
              w[b * RL + j] = mk((T)wd.re, (T)wd.im);
              if (j + 1 < RL) wd = wd * st;
            }
          }
        }
        return;
      }
      constexpr unsigned ESZ = sizeof(cx<T>);
      const unsigned tk = (unsigned)P.twk * ESZ;
      const unsigned toff = ((unsigned)u * (unsigned)P.twk + (unsigned)c * (unsigned)P.twc) * ESZ;
#pragma unroll
      for (int b = 0; b < BL; b++) {
        if (PL::FULL(NP - 1) || u + b * TPT < PL::NB(NP - 1)) {  // outputs k < L only
#pragma unroll
          for (int j = 0; j < RL; j++)
            w[b * RL + j] =
                ld_off<cx<T>>(reinterpret_cast<const char*>(P.tab_a) + (unsigned)(b * TPT + j * PLAST) * tk, toff);
        }
      }
    }
  }
};

// MODE: 0 plain passes; 1 fused Bluestein (chirp on load, spectrum product,
// second pass sequence, de-chirp); 2 spectrum product + second sequence only
// (the row pass of a four-step Bluestein).
template <typename T, int L, int TPT, int TPB, bool COL, int LD, int ST, int MODE, bool EPI, class RS, int AL, bool NT>
__device__ __forceinline__ void sfft_body(const KParams& P, const void* src, void* dst, int g0, double scale,
                                          cx<T>* lds, const StoreTw<T, Plan<RS, L, TPT>, TPT, ST, EPI>& stw) {
  constexpr bool BLUE = MODE == 1;
  using PL = Plan<RS, L, TPT>;
  constexpr int NP = PL::NP;
  constexpr int E = PL::EMAX();
  constexpr int LP = lds_stride(L, TPB);
  constexpr unsigned ESZ = sizeof(cx<T>);
  // a second LDS buffer saves one barrier per exchange but halves the
  // workgroups that fit on a CU; only tiny tiles take it
  constexpr bool PP = 2ull * TPB * LP * sizeof(cx<T>) <= 16384;
  constexpr int NTW = rad_twoff(RS{}, NP);          // full tables
  constexpr int NTWC = rad_compact_off(RS{}, NP);   // compact tables, after the full ones
  constexpr bool LDT = tw_in_lds(NTWC, (int)sizeof(cx<T>));
  constexpr int MTAB = mdct_tab_elems(L, LD, (int)sizeof(cx<T>));
  static_assert(((PP ? 2 : 1) * TPB * LP + (LDT ? NTWC : 0) + MTAB) * sizeof(cx<T>) <= kLdsBytes,
                "transform tile does not fit in LDS");
  const int tid = threadIdx.x;
  const int t = COL ? tid % TPB : tid / TPT;
  const int u = COL ? tid / TPB : tid % TPT;
  const int g = g0 + t;
  const bool valid = g < P.G;
  // the epilogue reads the final image as padded rows
  constexpr bool IL = COL && !EPI;
  static_assert(!PL::SERIAL || !LDT, "per-butterfly twiddles read the full tables in global memory");
  static_assert(pad_shift(L, TPB) == 4 || (!EPI && !IL && MTAB == 0),
                "the epilogues and operator tables assume one padding slot per 16 elements");
  using LY = std::conditional_t<IL, LdsInterleaved<TPB>, std::conditional_t<pad_shift(L, TPB) == 6, LdsRows64, LdsRows>>;
  cx<T>* row0 = IL ? lds + t : lds + t * LP;
  cx<T>* row1 = PP ? (IL ? lds + TPB * LP + t : lds + (TPB + t) * LP) : row0;
  const cx<T>* tw = reinterpret_cast<const cx<T>*>(P.seq.tw);
  const cx<T>* chirp = reinterpret_cast<const cx<T>*>(P.chirp);
  int cur = 0;

  constexpr int RL = PL::R(NP - 1), PLAST = PL::P(NP - 1), BL = PL::B(NP - 1);
  cx<T> x[E];
  // the MDCT fold's tables, behind the tile buffers and the pass tables
  cx<T>* mtab = lds + (PP ? 2 : 1) * TPB * LP + (LDT ? NTWC : 0);
  sfft_gather<T, L, TPT, TPB, COL, LD, MODE, RS, AL>(P, src, g0, x, mtab);
  // per-pass twiddle tables copied into LDS behind the tile buffers
  const cx<T>* twp = tw;
  if constexpr (LDT) {
    cx<T>* twl = lds + (PP ? 2 : 1) * TPB * LP;
    constexpr int TH = TPT * TPB;
#pragma unroll
    for (int e0 = 0; e0 < NTWC; e0 += TH)
      if (e0 + tid < NTWC) twl[e0 + tid] = ld_off<cx<T>>(tw + NTW + e0, (unsigned)tid * ESZ);
    twp = twl;
  }
  constexpr bool TW_EARLY = StoreTw<T, PL, TPT, ST, EPI>::EARLY;
  int c0o = 0, co = 0;
  long long o0 = 0, og = 0;
  static_assert(!(EPI && ST == ST_TW2), "two-level four-step twiddles are applied by the direct store only");
  if constexpr (ST == ST_PLAIN || ST == ST_TW || ST == ST_TW2) {
    o0 = io_base(P.out, g0, &c0o);
    og = io_base(P.out, valid ? g : g0, &co);
  }
  const unsigned tk = (unsigned)P.twk * ESZ;
  const unsigned toff = ((unsigned)u * (unsigned)P.twk + (unsigned)co * (unsigned)P.twc) * ESZ;
  run_all<T, PL, RS, TPT, PP, LY, LDT>(x, u, twp, row0, row1, cur);

  if constexpr (MODE != 0) {
    // spectrum product at the last-pass positions, conjugate, second sequence
    const cx<T>* K = reinterpret_cast<const cx<T>*>(P.kspec) + (long long)((valid ? g : g0) % P.in.ncols) * P.kc;
    const unsigned ke = (unsigned)P.ke * ESZ;
#pragma unroll
    for (int b = 0; b < BL; b++) {
      const int i = u + b * TPT;
      if (PL::FULL(NP - 1) || i < PL::NB(NP - 1)) {
#pragma unroll
        for (int j = 0; j < RL; j++)
          x[b * RL + j] = cj(x[b * RL + j] * ld_off<cx<T>>(K, (unsigned)(i + j * PLAST) * ke));
      }
      if constexpr (PL::SERIAL) __builtin_amdgcn_sched_barrier(0);
    }
    // long tiles: the second sequence recomputes its LDS and table
    // addresses instead of keeping the first sequence's live in registers
    int u2 = u;
    if constexpr (PL::SERIAL) asm volatile("" : "+v"(u2));
    exchange<T, PL, NP - 1, 0, TPT, PP, LY>(x, u2, row0, row1, cur);
    run_all<T, PL, RS, TPT, PP, LY, LDT>(x, u2, twp, row0, row1, cur);
  }
  if constexpr (BLUE) {
#pragma unroll
    for (int b = 0; b < BL; b++) {
      const int i = u + b * TPT;
      if (PL::FULL(NP - 1) || i < PL::NB(NP - 1)) {
#pragma unroll
        for (int j = 0; j < RL; j++) {
          const int rel = b * TPT + j * PLAST;  // k = u + rel
          if (lower_half<RL>(j) && u + rel < P.L) x[b * RL + j] = cj(x[b * RL + j]) * ld_off<cx<T>>(chirp + rel, (unsigned)u * ESZ);
        }
      }
      if constexpr (PL::SERIAL) __builtin_amdgcn_sched_barrier(0);
    }
  }

  if constexpr (EPI) {
    // through LDS in natural order
    cx<T>* row = (PP && cur) ? row1 : row0;
#pragma unroll
    for (int b = 0; b < BL; b++) {
      const int i = u + b * TPT;
      if (PL::FULL(NP - 1) || i < PL::NB(NP - 1)) {
#pragma unroll
        for (int j = 0; j < RL; j++) row[LY::template at<PLAST>(i, j)] = x[b * RL + j];
      }
    }
    lds_barrier();
    const cx<T>* img = (PP && cur) ? lds + TPB * LP : lds;
    const int nt = min(TPB, P.G - g0);
    if constexpr (ST == ST_PLAIN || ST == ST_TW) {
      if (P.colmajor_st)
        epi_store<T, L, TPT, TPB, ST, true, NT>(P, dst, img, LP, g0, nt, (T)scale);
      else
        epi_store<T, L, TPT, TPB, ST, false, NT>(P, dst, img, LP, g0, nt, (T)scale);
    } else if constexpr (ST == ST_DCT2 && L % 2 == 0) {
      epi_dct2<T, L, TPT, TPB, NT>(P, dst, img, LP, g0, nt, (T)scale);
    } else if constexpr (ST == ST_DCT3 && L % 2 == 0) {
      if (pair_aligned<AL>(dst, P.out.bdist, sizeof(cx<T>))) {
        epi_dct3<T, L, TPT, TPB, NT>(P, dst, img, LP, g0, nt, (T)scale);
      } else {
        const TMap m = make_map(TPB);
        store_stage<T, ST, NT>(P, m, dst, img, g0, nt, scale);
      }
    } else {
      // the planner sets P.Lp == LP for specialised launches
      const TMap m = make_map(TPB);
      store_stage<T, ST, NT>(P, m, dst, img, g0, nt, scale);
    }
  } else {
    if (!valid) return;
    const T s = (T)scale;
    if constexpr (ST == ST_PLAIN || ST == ST_TW || ST == ST_TW2) {
      char* ob = reinterpret_cast<char*>(dst) + o0 * (long long)ESZ;
      const unsigned es = (unsigned)P.out.estride * ESZ;
      const unsigned voff = (unsigned)(og - o0) * ESZ + (unsigned)u * es;
      const bool cjo = P.conj_out != 0;
      // two-level twiddle W^m = tab_a[m >> twsh] * tab_b[m mod 2^twsh],
      // m = (column) * (element index k = u + rel)
      const cx<T>* ta = reinterpret_cast<const cx<T>*>(P.tab_a);
      const cx<T>* tb = reinterpret_cast<const cx<T>*>(P.tab_b);
      const unsigned sh = (unsigned)P.twsh, lo = (1u << sh) - 1u;
      const unsigned mu = (unsigned)co * (unsigned)u;
      // twiddles by recurrence (see StoreTw::load) where the plan has the table
      constexpr bool REC = ST == ST_TW && !TW_EARLY;
      const cx<double>* trc = REC ? reinterpret_cast<const cx<double>*>(P.tw_rec) : nullptr;
      const int trn = P.out.ncols;
      cx<double> wst = mk(1.0, 0.0);
      if (REC && trc) wst = trc[BL * TPT * trn + co];
      // ST_TW2 by recurrence where the plan has FP64 two-level tables (tw_rec:
      // [W^i, i < 2^twsh][W^(i 2^twsh)]): W^(c k) for each butterfly's first
      // output k = u + b TPT and the step W^(c PLAST) from two lookups each, the
      // rest by FP64 products, each twiddle rounded once
      constexpr bool REC2 = ST == ST_TW2;
      const cx<double>* t2 = REC2 ? reinterpret_cast<const cx<double>*>(P.tw_rec) : nullptr;
      if (REC2 && t2) {
        const unsigned ms = (unsigned)co * (unsigned)PLAST;
        wst = t2[(1u << sh) + (ms >> sh)] * t2[ms & lo];
      }
      const bool rec = (REC && trc) || (REC2 && t2);
#pragma unroll
      for (int b = 0; b < BL; b++) {
        const int i = u + b * TPT;
        if (PL::FULL(NP - 1) || i < PL::NB(NP - 1)) {
          cx<double> wd = mk(1.0, 0.0);
          if (REC && trc) wd = trc[i * trn + co];
          if (REC2 && t2) {
            const unsigned m0 = (unsigned)co * (unsigned)i;
            wd = t2[(1u << sh) + (m0 >> sh)] * t2[m0 & lo];
          }
#pragma unroll
          for (int j = 0; j < RL; j++) {
            const int rel = b * TPT + j * PLAST;  // k = u + rel
            if (!BLUE || (lower_half<RL>(j) && u + rel < P.L)) {
              cx<T> v = x[b * RL + j];
              if constexpr (TW_EARLY)
                v = v * stw.w[b * RL + j];
              else if constexpr (ST == ST_TW) {
                if (trc)
                  v = v * mk((T)wd.re, (T)wd.im);
                else
                  v = v * ld_off<cx<T>>(reinterpret_cast<const char*>(P.tab_a) + (unsigned)rel * tk, toff);
              } else if constexpr (ST == ST_TW2) {
                if (t2) {
                  v = v * mk((T)wd.re, (T)wd.im);
                } else {
                  const unsigned m = mu + (unsigned)co * (unsigned)rel;
                  v = v * (ta[m >> sh] * tb[m & lo]);
                }
              }
              if (cjo) v = cj(v);
              st_off<cx<T>, NT>(ob + (unsigned)rel * es, voff, scl(v, s));
            }
            if (rec && j + 1 < RL) wd = wd * wst;
          }
        }
      }
    } else if constexpr (ST == ST_MDCT && !COL && !BLUE && PL::FULL(NP - 1) && (TPT & (TPT - 1)) == 0 &&
                         TPT <= 64) {
      // MDCT output (mdct1d.rs:234-240): bin k gives ret[2k] = -w_k.re and
      // ret[N/2-1-2k] = w_k.im; bins k and L-1-k fill the aligned pair
      // ret[2k], ret[2k+1] = (-w_k.re, w_(L-1-k).im). Bin L-1-k sits in the
      // mirror lane TPT-1-u at the mirrored slot; one shuffle brings its
      // imaginary part over, and each lane stores one pair per bin.
      const cx<T>* tw = reinterpret_cast<const cx<T>*>(P.tab_a);
      T* ret = reinterpret_cast<T*>(dst) + (long long)g * P.out.bdist;
      constexpr unsigned RE = sizeof(T), CE = sizeof(cx<T>);
      const bool cjo = P.conj_out != 0;
      if (pair_aligned<AL>(dst, P.out.bdist, sizeof(cx<T>))) {  // uniform
        T wim[BL * RL], wre[BL * RL];
#pragma unroll
        for (int b = 0; b < BL; b++)
#pragma unroll
          for (int j = 0; j < RL; j++) {
            const int k = u + b * TPT + j * PLAST;
            cx<T> v = x[b * RL + j];
            if (cjo) v = cj(v);
            // the aligned fold (AL == 1 with LD_MDCT, FP32) left the twiddles in LDS
            const cx<T> t = (AL == 1 && mdct_tab_elems(L, LD, (int)sizeof(cx<T>)) > 0)
                                ? mtab[2 * L + k]
                                : ld_off<cx<T>>(tw, (unsigned)k * CE);
            const cx<T> w = v * scl(t, s);
            wre[b * RL + j] = w.re;
            wim[b * RL + j] = w.im;
          }
#pragma unroll
        for (int b = 0; b < BL; b++)
#pragma unroll
          for (int j = 0; j < RL; j++) {
            const int k = u + b * TPT + j * PLAST;
            const T pim = __shfl_xor(wim[(BL - 1 - b) * RL + (RL - 1 - j)], TPT - 1);
            st_off<cx<T>, NT>(ret, (unsigned)(2 * k) * RE, mk(-wre[b * RL + j], pim));
          }
      } else {
        const int h = (int)(P.N >> 1);
#pragma unroll
        for (int b = 0; b < BL; b++)
#pragma unroll
          for (int j = 0; j < RL; j++) {
            const int k = u + b * TPT + j * PLAST;
            cx<T> v = x[b * RL + j];
            if (cjo) v = cj(v);
            const cx<T> w = v * scl(tw[k], s);
            ret[2 * k] = -w.re;
            ret[h - 1 - 2 * k] = w.im;
          }
      }
    } else {
      int c = 0;
      long long obase;
      if constexpr (ST == ST_FS_DECHIRP)
        obase = io_base(P.out, g, &c);
      else
        obase = (long long)g * P.out.bdist;
#pragma unroll
      for (int b = 0; b < BL; b++) {
        const int i = u + b * TPT;
        if (PL::FULL(NP - 1) || i < PL::NB(NP - 1)) {
#pragma unroll
          for (int j = 0; j < RL; j++) {
            const int k = i + j * PLAST;
            if (!BLUE || (lower_half<RL>(j) && k < P.L)) store_elem<T, ST, NT>(P, dst, obase, c, k, x[b * RL + j], s);
          }
        }
      }
    }
  }
}

template <typename T, int L, int TPT, int TPB, bool COL, int LD, int ST, int MODE, bool EPI, class RS, int AL = 2,
          bool NT = false>
__device__ __forceinline__ void sfft_tile(const KParams& P, const void* src, void* dst, int g0, double scale,
                                          cx<T>* lds) {
  StoreTw<T, Plan<RS, L, TPT>, TPT, ST, EPI> stw;
  if constexpr (decltype(stw)::EARLY) {
    const int tid = threadIdx.x;
    const int t = COL ? tid % TPB : tid / TPT;
    const int u = COL ? tid / TPB : tid % TPT;
    int c;
    io_base(P.out, min(g0 + t, P.G - 1), &c);
    stw.load(P, u, c);
  }
  sfft_body<T, L, TPT, TPB, COL, LD, ST, MODE, EPI, RS, AL, NT>(P, src, dst, g0, scale, lds, stw);
}

// Tile index of this workgroup, arranged so that the workgroups one XCD runs
// (hardware assigns workgroup b to XCD b mod 8) take consecutive tiles: tiles
// that share cache lines at their edges then meet in the same L2. Used by
// kernels with column-strided loads or stores (128-byte pieces of lines
// that neighbouring tiles complete); kernels that read and write whole rows
// keep the dispatch order, which measured 5-11 % faster on MI355X for them
// (RFft1D 8192 x 4096, DCT-II 1024 x 16384, Bluestein 65537 x 64) while
// the 43200 row pass with its transposed store ran 18 % slower without the
// arrangement.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  constexpr int X = 8;
  const int x = b % X, q = nb / X, r = nb % X;
  const int first = x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q;
  return first + b / X;
}

}  // namespace chfft
