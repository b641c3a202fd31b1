// Generic (runtime-radix) building blocks of the fused transform: a
// workgroup owns `tpb` transforms whose working sequence (length Lfft) is
// resident in LDS.
//   1. load stage   global -> LDS, through a load operator (plain, chirp,
//                   RFft/DCT/MDCT pre-pass), optional conjugation
//   2. passes       radix-2/3/4/5/7/8/16 (and 10/12/15 where the kernel asks)
//                   Stockham autosort passes LDS -> LDS,
//                   inputs staged in registers between two barriers
//   3. (Bluestein)  spectrum product, conjugate, second pass sequence,
//                   de-chirp
//   4. store stage  LDS -> global, through a store operator (plain, four-step
//                   twiddle, de-chirp, RFft/DCT/MDCT post-pass)
// Used by the fused kernels (chfft_kernels.hip) and the multi-length
// persistent kernel (chfft_multi.hip).
#pragma once
#include "chfft_device.h"

namespace chfft {

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
      for (int j = 0; j < R; j++) x[q][j] = base[lx(i + j * nbt)];
      if (p > 1) {
        // per-pass table laid out [j-1][k]: consecutive lanes read consecutive k
        const cx<T>* twq = tw + pd.twoff + k;
#pragma unroll
        for (int j = 1; j < R; j++) x[q][j] = x[q][j] * twq[(j - 1) * p];
      }
      obase[q] = (i - k) * R + k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < BPT; q++) {
    if (obase[q] >= 0) {
      dft<T, R>(x[q]);
#pragma unroll
      for (int j = 0; j < R; j++) base[lx(obase[q] + j * p)] = x[q][j];
    }
  }
  __syncthreads();
}

// PFA: the schedule may also hold the prime-factor radices 10, 12 and 15
// (fewer passes for lengths with factors 3 and 5); only the kernels that
// take such schedules compile those cases, which cost the others registers.
template <typename T, bool PFA = false>
__device__ void run_passes(cx<T>* lds, const TMap& m, const SeqDesc& s, int Lp) {
  const cx<T>* tw = reinterpret_cast<const cx<T>*>(as_global(s.tw));
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
      default:
        if constexpr (PFA) {
          switch (pd.radix) {
            case 10: lds_pass<T, 10>(lds, m, s.L, Lp, pd, tw); break;
            case 12: lds_pass<T, 12>(lds, m, s.L, Lp, pd, tw); break;
            case 15: lds_pass<T, 15>(lds, m, s.L, Lp, pd, tw); break;
            default: break;
          }
        }
        break;
    }
  }
}

template <typename T, int LD>
__device__ void load_stage(const KParams& P, const TMap& m, const void* src, cx<T>* lds, int g0, int nt) {
  const int t = P.colmajor_ld ? m.tc : m.tr;
  const int u = P.colmajor_ld ? m.uc : m.ur;
  const cx<T>* chirp = reinterpret_cast<const cx<T>*>(as_global(P.chirp));
  cx<T>* row = lds + t * P.Lp;
  if (t < nt) {
    const Loader<T, LD> ld = make_loader<T, LD>(P, src, g0 + t);
    // batches of 8 elements per thread: all loads of a batch are issued
    // before any LDS write, so they are in flight together
    constexpr int NB = 8;
    for (int n0 = u; n0 < P.Lfft; n0 += NB * m.tpt) {
      cx<T> v[NB];
#pragma unroll
      for (int e = 0; e < NB; e++) {
        const int n = n0 + e * m.tpt;
        v[e] = mk(T(0), T(0));
        if (n < P.L) v[e] = load_elem<T, LD>(P, ld, n);  // n >= L only under a fused Bluestein: zero padding
      }
#pragma unroll
      for (int e = 0; e < NB; e++) {
        const int n = n0 + e * m.tpt;
        if (n < P.Lfft) {
          cx<T> x = v[e];
          if (n < P.L) {
            if constexpr (LD != LD_FS_CHIRP) {
              if (P.conj_in) x = cj(x);
            }
            if (P.has_chirp) x = x * chirp[n];
          }
          row[lx(n)] = x;
        }
      }
    }
  } else {
    for (int n = u; n < P.Lfft; n += m.tpt) row[lx(n)] = mk(T(0), T(0));
  }
  __syncthreads();
}

// Bluestein spectrum product (chirpz.rs:115-117) followed by a conjugation so
// that the inverse transform reuses the forward passes.
template <typename T>
__device__ void mid_stage(const KParams& P, const TMap& m, cx<T>* lds, int g0) {
  const cx<T>* K = reinterpret_cast<const cx<T>*>(as_global(P.kspec));
  const int c = (g0 + m.tr) % P.in.ncols;
  const long long kb = (long long)c * P.kc;
  cx<T>* row = lds + m.tr * P.Lp;
  for (int n = m.ur; n < P.Lfft; n += m.tpt) row[lx(n)] = cj(row[lx(n)] * K[kb + (long long)n * P.ke]);
  __syncthreads();
}

// Fused Bluestein de-chirp (chirpz.rs:127-145 with the conjugation undone).
template <typename T>
__device__ void dechirp_stage(const KParams& P, const TMap& m, cx<T>* lds) {
  const cx<T>* chirp = reinterpret_cast<const cx<T>*>(as_global(P.chirp));
  cx<T>* row = lds + m.tr * P.Lp;
  for (int n = m.ur; n < P.L; n += m.tpt) row[lx(n)] = cj(row[lx(n)]) * chirp[n];
  __syncthreads();
}

// One LDS-resident batch of tpb transforms: load, passes, [Bluestein middle],
// store. The store stage is not followed by a barrier.
template <typename T, int LD, int ST, bool PFA = false>
__device__ __forceinline__ void block_transform(const KParams& P, const void* src, void* dst, cx<T>* lds, int g0,
                                                int nt, double scale) {
  const TMap m = make_map(P.tpb);
  load_stage<T, LD>(P, m, src, lds, g0, nt);
  run_passes<T, PFA>(lds, m, P.seq, P.Lp);
  if (P.has_mid) {
    mid_stage<T>(P, m, lds, g0);
    run_passes<T, PFA>(lds, m, P.seq, P.Lp);
    if (P.has_chirp) dechirp_stage<T>(P, m, lds);
  }
  store_stage<T, ST>(P, m, dst, lds, g0, nt, scale);
}

}  // namespace chfft
