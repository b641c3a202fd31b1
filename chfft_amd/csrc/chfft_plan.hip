// Host side of the transform engine: the GPU plan builder and the C ABI
// declared in include/chfft_hip.h.
//
// The plan builder replaces chfft's CFft1D::setup (cfft1d.rs:103-206) and the
// constructors of the derived types. A plan is a short list of kernel launches
// (chfft_kernels.hip) plus device tables generated on the host in long double
// and rounded once to the plan's precision:
//   * length fits one workgroup's LDS  -> one fused launch (Stockham passes,
//     or a fused Bluestein when the length is not 7-smooth);
//   * longer 7-smooth lengths          -> four-step: column pass with the
//     inter-pass twiddle, then row pass with a transposed store (2 launches);
//   * longer non-7-smooth lengths      -> Bluestein over a four-step M-point
//     FFT (3 launches: chirp + columns, rows + spectrum product + inverse rows,
//     inverse columns + de-chirp).
// The MAX_PRIME = 7 rule follows cfft1d.rs:104. The fused Bluestein pads to
// M = 2 * next_pow2(N) as cfft1d.rs:114 does; the four-step Bluestein takes
// the cheapest smooth M >= 2N - 1 (bluestein_length).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/chfft_hip.h"
#include "chfft_internal.h"

namespace chfft {

hipError_t launch_fused(bool f64, int ld, int st, const KParams& P, dim3 grid, dim3 block, size_t lds,
                        hipStream_t stream);
hipError_t set_fused_lds_limits();
hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t stream);
hipError_t launch_fourstep(bool f64, int nph, const KParams* Q, const void* in, void* out, double scale, int batch,
                           long long src_dist, long long dst_dist, void* scratch, long long slot_elems, dim3 grid,
                           dim3 block, size_t lds, hipStream_t stream);
int fourstep_occupancy(bool f64, int nph, int block, size_t lds);
hipError_t launch_multi(bool f64, const MultiItem* items, const MultiUnit* units, const int* ends,
                        const KParams* descs, const void* in, void* out, int scale_kind, void* scratch,
                        unsigned* sync, unsigned long long ticks, unsigned extra, bool coop, dim3 grid, size_t lds,
                        hipStream_t st);
hipError_t launch_pre(bool f64, int ld, const KParams& P, hipStream_t st);
hipError_t launch_post(bool f64, int st_op, const KParams& P, hipStream_t st);
hipError_t launch_transpose(bool f64, const void* src, void* dst, int R, int C, int batch, hipStream_t st);
hipError_t launch_ola(bool f64, const void* frames, void* out, long long n, long long hop, long long count,
                      hipStream_t st);
int multi_occupancy(bool f64, size_t lds);
hipError_t multi_set_lds_limits();
struct MultiSpec;
const MultiSpec* multi_spec_find(int L, int ld, int st, int mode, bool col);
void multi_spec_info(const MultiSpec* s, int* id, int* tpb, int* nrad, const int** rad, size_t* lds);
hipError_t gen_roots(bool f64, const RootJob* jobs_dev, int njobs, void* arena, hipStream_t st);
hipError_t gen_rot(const long long* ns_dev, int count, long long M, void* out, hipStream_t st);
hipError_t put_spectra(bool f64, const void* src, int count, long long M, long long n1, const long long* offs_dev,
                       void* arena, hipStream_t st);
struct SfftEntry {
  int f64, L, tpt, tpb, col, ld, st, mode, epi;
  int nrad;
  int rad[8];
  void (*fn)(KParams);
  void (*fn0)(KParams);
  void (*fnt)(KParams);
};
const SfftEntry* sfft_find(bool f64, int L, int ld, int st, int mode, bool col);
bool sfft_has_schedule(bool f64, int L, const int* rad, int nrad);
int sfft_fused_bluestein_m(bool f64, long long N);
int sfft_schedule_of(bool f64, int L, int ld, int st, int mode, bool col, int* rad);
hipError_t sfft_launch(const SfftEntry* e, const KParams& P, dim3 grid, size_t lds, hipStream_t st);
hipError_t sfft_set_lds_limits();
size_t sfft_lds_bytes(const SfftEntry* e);
int sfft_lds_stride(const SfftEntry* e);
struct SfsEntry;
const SfsEntry* sfs_find(bool f64, int n1, int n2, int nph);
void sfs_info(const SfsEntry* e, int* threads, int* tpb, int* Lp, size_t* lds);
hipError_t sfs_launch(const SfsEntry* e, const KParams* Q, const void* in, void* out, double scale, int batch,
                      long long src_dist, long long dst_dist, void* scratch, long long slot_elems, dim3 grid,
                      hipStream_t st);
int sfs_occupancy(const SfsEntry* e);

namespace {

struct Status {
  int code;
};

// ------------------------------------------------------------------ device memory
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  // moves keep the device address (launch descriptors hold it)
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      if (p) (void)hipFree(p);
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

int dev_alloc(DevBuf& b, size_t bytes) {
  if (b.p && b.bytes >= bytes) return CHFFT_OK;
  if (b.p) {
    (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
  }
  if (bytes == 0) return CHFFT_OK;
  if (hipMalloc(&b.p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    b.p = nullptr;
    return CHFFT_ERR_OUT_OF_MEMORY;
  }
  b.bytes = bytes;
  return CHFFT_OK;
}

// ------------------------------------------------------------------ host tables
using ld_t = long double;
const ld_t kPi = 3.141592653589793238462643383279502884L;

// exp(-2 pi i num / den) with the fraction reduced exactly in integers first.
void unit_root(unsigned long long num, unsigned long long den, ld_t* re, ld_t* im) {
  num %= den;
  ld_t th = -2.0L * kPi * (ld_t)num / (ld_t)den;
  *re = cosl(th);
  *im = sinl(th);
}

struct HostTable {
  std::vector<ld_t> v;  // interleaved re, im (or plain reals)
};

// A host table rounded once to the plan's precision and uploaded.
int upload(const std::vector<ld_t>& v, bool f64, DevBuf& dst) {
  size_t n = v.size();
  size_t bytes = n * (f64 ? 8 : 4);
  int rc = dev_alloc(dst, bytes);
  if (rc) return rc;
  if (f64) {
    std::vector<double> h(n);
    for (size_t i = 0; i < n; i++) h[i] = (double)v[i];
    if (hipMemcpy(dst.p, h.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) return CHFFT_ERR_HIP;
  } else {
    std::vector<float> h(n);
    for (size_t i = 0; i < n; i++) h[i] = (float)v[i];
    if (hipMemcpy(dst.p, h.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) return CHFFT_ERR_HIP;
  }
  return CHFFT_OK;
}

std::vector<ld_t> roots_table(size_t L) {  // W_L^m, m in [0, L)
  std::vector<ld_t> v(2 * L);
  for (size_t m = 0; m < L; m++) unit_root(m, L, &v[2 * m], &v[2 * m + 1]);
  return v;
}

// Bluestein chirp c_n = exp(-pi i (n^2 mod 2N) / N), n in [0, N); the index is
// formed in exact 64-bit integers (cfft1d.rs:123-128, chirpz.rs:97-101).
std::vector<ld_t> chirp_table(size_t N) {
  std::vector<ld_t> v(2 * N);
  const unsigned long long d = 2ull * N;
  for (size_t n = 0; n < N; n++) {
    unsigned long long sq = (unsigned long long)(((unsigned __int128)n * n) % d);
    unit_root(sq, d, &v[2 * n], &v[2 * n + 1]);
  }
  return v;
}

// Forward FFT of a 7-smooth length in double precision with long-double
// twiddles (used once per Bluestein plan for the kernel spectrum): Stockham
// passes of radix 4, 2, 3, 5, 7, each butterfly a direct DFT.
void host_fft(std::vector<double>& re, std::vector<double>& im) {
  const size_t n = re.size();
  if (n <= 1) return;
  std::vector<double> wr(n), wi(n);  // W_n^t, t < n
  for (size_t t = 0; t < n; t++) {
    ld_t a, b;
    unit_root(t, n, &a, &b);
    wr[t] = (double)a;
    wi[t] = (double)b;
  }
  std::vector<int> rad;
  size_t m = n;
  while (m % 4 == 0) {
    rad.push_back(4);
    m /= 4;
  }
  for (int q : {2, 3, 5, 7})
    while (m % q == 0) {
      rad.push_back(q);
      m /= q;
    }
  std::vector<double> yr(n), yi(n);
  size_t p = 1;
  for (int r : rad) {
    const size_t nb = n / r;
    for (size_t i = 0; i < nb; i++) {
      const size_t k = i % p;
      double ar[8], ai[8];
      for (int j = 0; j < r; j++) {  // inputs times W_{p r}^{j k} = W_n^{j k n / (p r)}
        const size_t t = ((size_t)j * k * (n / (p * r))) % n;
        const double xr = re[i + j * nb], xi = im[i + j * nb];
        ar[j] = xr * wr[t] - xi * wi[t];
        ai[j] = xr * wi[t] + xi * wr[t];
      }
      for (int j = 0; j < r; j++) {  // DFT_r: W_r^{j m} = W_n^{j m n / r}
        double sr = 0, si = 0;
        for (int q = 0; q < r; q++) {
          const size_t t = ((size_t)j * q % r) * (n / r);
          sr += ar[q] * wr[t] - ai[q] * wi[t];
          si += ar[q] * wi[t] + ai[q] * wr[t];
        }
        const size_t o = (i - k) * r + k + j * p;
        yr[o] = sr;
        yi[o] = si;
      }
    }
    re.swap(yr);
    im.swap(yi);
    p *= r;
  }
}

// Kernel spectrum K'[k] = FFT_M(rot)[k] / M with rot[i] = conj(c_i) for i < N,
// 0 for N <= i <= M/2, rot[M - i] mirrored above M/2 (cfft1d.rs:120-140, 179;
// the 1/M of chirpz.rs:42-46 is folded in).
std::vector<ld_t> bluestein_spectrum(size_t N, size_t M) {
  std::vector<ld_t> c = chirp_table(N);
  std::vector<double> re(M, 0.0), im(M, 0.0);
  for (size_t i = 0; i < N; i++) {
    re[i] = (double)c[2 * i];
    im[i] = -(double)c[2 * i + 1];
  }
  for (size_t i = M / 2 + 1; i < M; i++) {
    re[i] = re[M - i];
    im[i] = im[M - i];
  }
  host_fft(re, im);
  std::vector<ld_t> v(2 * M);
  for (size_t k = 0; k < M; k++) {
    v[2 * k] = (ld_t)re[k] / (ld_t)M;
    v[2 * k + 1] = (ld_t)im[k] / (ld_t)M;
  }
  return v;
}

// ------------------------------------------------------------------ factorization
bool smooth7(size_t n) {
  if (n == 0) return false;
  for (size_t p : {2, 3, 5, 7})
    while (n % p == 0) n /= p;
  return n == 1;
}

bool smooth5(size_t n) {
  if (n == 0) return false;
  for (size_t p : {2, 3, 5})
    while (n % p == 0) n /= p;
  return n == 1;
}

size_t next_pow2(size_t n) {
  size_t p = 1;
  while (p < n) p <<= 1;
  return p;
}

// Radix schedule for a 7-smooth L: as few passes as the radix set allows
// (16, 8, 4, 2 for the power of two; 3, 5, 7 for the odd part). With `pfa`
// prime-factor radices join coprime factors into one pass: each 3 x 5 pair
// becomes radix 15, a lone 4 with a 3 radix 12, and an odd power of two
// (2^(4k+1)) gives its single 2 to a 5 as radix 10 (180 = 12 x 15,
// 240 = 16 x 15, 160 = 16 x 10): a schedule only the length-specialised
// kernels run (the runtime-radix kernels stop at 16).
std::vector<int> radices_for(size_t L, bool pfa = false) {
  int a = 0;
  while (L % 2 == 0) {
    L /= 2;
    a++;
  }
  int n3 = 0, n5 = 0, n7 = 0;
  while (L % 5 == 0) {
    L /= 5;
    n5++;
  }
  while (L % 3 == 0) {
    L /= 3;
    n3++;
  }
  while (L % 7 == 0) {
    L /= 7;
    n7++;
  }
  const int n15 = pfa ? std::min(n3, n5) : 0;
  n3 -= n15;
  n5 -= n15;
  std::vector<int> r;
  int k16 = a / 4;
  const int rem = a % 4;
  const bool ten = pfa && rem == 1 && n5 > 0;  // 2 x 5
  if (ten) {
    n5--;
  } else if (rem == 1 && k16 > 0) {
    k16--;
    r.push_back(8);
    r.push_back(4);
  } else if (rem == 1) {
    r.push_back(2);
  } else if (rem == 2) {
    r.push_back(4);
  } else if (rem == 3) {
    r.push_back(8);
  }
  for (int i = 0; i < k16; i++) r.insert(r.begin(), 16);
  if (ten) r.push_back(10);
  if (pfa && n3 > 0 && !r.empty() && r.back() == 4 && (r.size() == 1 || r[r.size() - 2] != 8)) {
    r.back() = 12;  // the lone 4 (not the 8 x 4 tail) joins a 3
    n3--;
  }
  for (int i = 0; i < n15; i++) r.push_back(15);
  for (int i = 0; i < n5; i++) r.push_back(5);
  for (int i = 0; i < n3; i++) r.push_back(3);
  for (int i = 0; i < n7; i++) r.push_back(7);
  return r;
}

// Alternative radix schedules of a 7-smooth L for the runtime-radix kernels
// (the autotuner's radix-order knob): [0] the planner's schedule, then its
// reverse, the power-of-two part as radix-8 or radix-4 passes, and the odd
// radices first. Distinct schedules only, each within kMaxPass passes.
std::vector<std::vector<int>> radix_variants(size_t L) {
  std::vector<std::vector<int>> out;
  auto add = [&](const std::vector<int>& r) {
    if ((int)r.size() > kMaxPass) return;
    for (const auto& o : out)
      if (o == r) return;
    out.push_back(r);
  };
  const std::vector<int> base = radices_for(L);
  add(base);
  add(std::vector<int>(base.rbegin(), base.rend()));
  int a = 0;
  for (size_t m = L; m > 1 && m % 2 == 0; m /= 2) a++;
  std::vector<int> odd;
  for (int r : base)
    if (r & 1) odd.push_back(r);
  for (int lb : {3, 2}) {  // radix 8, radix 4
    std::vector<int> r(a / lb, 1 << lb);
    if (a % lb) r.push_back(1 << (a % lb));
    r.insert(r.end(), odd.begin(), odd.end());
    add(r);
  }
  std::vector<int> odd_first = odd;
  for (int r : base)
    if (!(r & 1)) odd_first.push_back(r);
  add(odd_first);
  return out;
}

// Planner choices the autotuner varies; all zero is the planner's own rules.
// Text form: "key=value" pairs separated by spaces or commas (tune_str).
struct Tune {
  int generic = 0;      // 1: runtime-radix kernels only (no length-specialised kernel)
  int order = 0;        // radix_variants() index of each runtime-radix sequence (where it has one)
  int tpb = 0, tpt = 0; // fused launches: transforms per workgroup, threads per transform; multi-length
                        // plans (tpb only): the most transforms of a generic four-step tile (default 64)
  long long n1 = 0;     // four-step split N = n1 x n2
  long long chunk = 0;  // transforms per chunk of the split four-step (-1: the whole batch)
  int sched = 0;        // 1: one persistent four-step launch; 2: split four-step launches, also where
                        // the planner would fuse a padded Bluestein sequence into one workgroup
  int grid = 0;         // multi-length plans: workgroups of the persistent launch
  int unit = 0;         // multi-length plans: points per work unit of a four-step phase
};

std::string tune_str(const Tune& t) {
  std::string s;
  auto put = [&](const char* k, long long v) {
    if (v == 0) return;
    if (!s.empty()) s += ' ';
    s += std::string(k) + "=" + std::to_string(v);
  };
  put("generic", t.generic);
  put("order", t.order);
  put("tpb", t.tpb);
  put("tpt", t.tpt);
  put("n1", t.n1);
  put("chunk", t.chunk);
  put("sched", t.sched);
  put("grid", t.grid);
  put("unit", t.unit);
  return s.empty() ? "default" : s;
}

int tune_parse(const char* s, Tune& t) {
  t = Tune{};
  if (!s) return CHFFT_ERR_INVALID_ARGUMENT;
  std::string str(s);
  for (char& c : str)
    if (c == ',') c = ' ';
  size_t i = 0;
  while (i < str.size()) {
    while (i < str.size() && str[i] == ' ') i++;
    if (i >= str.size()) break;
    size_t j = str.find(' ', i);
    if (j == std::string::npos) j = str.size();
    const std::string tok = str.substr(i, j - i);
    i = j;
    if (tok == "default") continue;
    const size_t eq = tok.find('=');
    if (eq == std::string::npos || eq + 1 >= tok.size()) return CHFFT_ERR_INVALID_ARGUMENT;
    const std::string k = tok.substr(0, eq), v = tok.substr(eq + 1);
    char* end = nullptr;
    const long long x = strtoll(v.c_str(), &end, 10);
    if (!end || *end != '\0') return CHFFT_ERR_INVALID_ARGUMENT;
    const bool small = x >= -1 && x <= (1LL << 30);
    if (!small) return CHFFT_ERR_INVALID_ARGUMENT;
    if (k == "generic" && (x == 0 || x == 1)) t.generic = (int)x;
    else if (k == "order" && x >= 0) t.order = (int)x;
    else if (k == "tpb" && x >= 0) t.tpb = (int)x;
    else if (k == "tpt" && x >= 0) t.tpt = (int)x;
    else if (k == "n1" && x >= 0) t.n1 = x;
    else if (k == "chunk") t.chunk = x;
    else if (k == "sched" && x >= 0 && x <= 2) t.sched = (int)x;
    else if (k == "grid" && x >= 0) t.grid = (int)x;
    else if (k == "unit" && x >= 0) t.unit = (int)x;
    else return CHFFT_ERR_INVALID_ARGUMENT;
  }
  return CHFFT_OK;
}

size_t esz_of(bool f64) { return f64 ? 16 : 8; }

int max_threads(bool f64) { return f64 ? kMaxThreadsF64 : kMaxThreadsF32; }

// Largest sequence one workgroup holds: registers (kEmax per thread) and LDS.
size_t lds_cap(bool f64) {
  return std::min<size_t>((size_t)kEmax * max_threads(f64), kLdsBytes / esz_of(f64));
}

// ------------------------------------------------------------------ launch geometry
struct Geometry {
  int tpb, tpt, Lp, block;
  size_t lds;
};

// Workgroup shape for tpb transforms of length Lfft: tpt threads per
// transform (>= Lfft / kEmax so one pass stages at most kEmax values per
// thread; about 8 elements per thread otherwise), tpb a power of two
// (16 for column-major access: 16 adjacent complex values per row segment),
// block = tpb * tpt a whole number of wavefronts where possible.
Geometry choose_geometry(int Lfft, long long G, bool colmajor, bool f64, int max_t = 0) {
  const int maxT = max_t ? max_t : max_threads(f64);
  const size_t esz = esz_of(f64);
  const int tpt_min = (Lfft + kEmax - 1) / kEmax;
  int tpt_pref = std::max(tpt_min, std::min((Lfft + 7) / 8, 256));
  auto lp_of = [&](int tpb) {
    int Lp = Lfft + (Lfft + 15) / 16;  // one padding slot per 16 elements
    if (tpb > 1 && (Lp & 1) == 0) Lp++;
    return Lp;
  };
  auto fits = [&](int tpb, int tpt) {
    int Lp = lp_of(tpb);
    return (long long)tpb * tpt <= maxT && (size_t)tpb * Lp * esz <= kLdsBytes;
  };
  int tpb;
  if (colmajor) {
    tpb = 16;
  } else {
    tpb = 1;
    while (tpb * 2 * Lfft <= 2048) tpb *= 2;
    while (tpb > 1 && (G + tpb - 1) / tpb < 1024) tpb >>= 1;
  }
  while (tpb > 1 && !fits(tpb, tpt_min)) tpb >>= 1;
  int tpt = tpt_pref;
  const int gran = 64 / std::min(64, tpb);  // tpb is a power of two
  tpt = (tpt + gran - 1) / gran * gran;
  while (tpt > tpt_min && !fits(tpb, tpt)) tpt--;
  if (!fits(tpb, tpt)) tpt = tpt_min;
  Geometry g;
  g.tpb = tpb;
  g.tpt = tpt;
  g.Lp = lp_of(tpb);
  g.block = tpb * tpt;
  g.lds = (size_t)tpb * g.Lp * esz;
  return g;
}

// Workgroup shape of tpb transforms of tpt threads each (the autotuner's
// tile knobs); false where it does not fit the kernel: fewer threads per
// transform than Lfft / kEmax, more threads than the kernel's limit, or more
// LDS than one workgroup gets.
bool geometry_with(int Lfft, int tpb, int tpt, bool f64, Geometry& g) {
  if (tpb < 1 || tpt < 1) return false;
  if (tpt < (Lfft + kEmax - 1) / kEmax || (long long)tpb * tpt > max_threads(f64)) return false;
  int Lp = Lfft + (Lfft + 15) / 16;
  if (tpb > 1 && (Lp & 1) == 0) Lp++;
  const size_t lds = (size_t)tpb * Lp * esz_of(f64);
  if (lds > kLdsBytes) return false;
  g.tpb = tpb;
  g.tpt = tpt;
  g.Lp = Lp;
  g.block = tpb * tpt;
  g.lds = lds;
  return true;
}

// ------------------------------------------------------------------ plan objects
enum Kind { K_C2C = 0, K_R2C = 1, K_C2C2D = 2, K_DCT = 3, K_MDCT = 4, K_MULTI = 5 };
// Buffers a launch reads and writes: the exec's input and output, the
// plan's four-step scratch, and the plan's work buffer (the operand of the
// long-length operator passes and of the transposed 2-D column pass).
enum Slot { S_IN = 0, S_OUT = 1, S_SCRATCH = 2, S_WORK = 3 };
constexpr int kSlots = 4;

struct Launch {
  int kind = 0;  // 0 fused kernel, 1 copy, 2 persistent four-step kernel, 3 specialised fused kernel,
                 // 4 specialised persistent four-step kernel, 5 operator pre-pass (P.L elements per
                 // transform into the work layout), 6 operator post-pass, 7 transpose (P.G matrices
                 // of P.Lfft x P.L complex elements)
  const SfftEntry* sf = nullptr;
  const SfsEntry* sfs = nullptr;
  int ld, st;
  KParams P;
  // kind 2: per-phase descriptors of one transform
  KParams Q[3];
  int nph;
  long long src_dist, dst_dist, slot_elems;
  int batch;
  void* qdev;  // device copy of Q[0..nph), owned by the plan
  dim3 grid, block;
  size_t lds;
  int src, dst;      // Slot
  long long src_off = 0, dst_off = 0;  // bytes into the in / out slots (chunked launches)
  int scaled;        // receives the method's scale
  double scale_mul;  // extra factor (RFft backward: 2)
  std::string what;
};

// A fused sequence (one workgroup per tpb transforms) for a logical inner
// length L: Stockham passes over L, or a fused Bluestein over M.
struct Inner {
  size_t L = 0, Lfft = 0;
  bool bluestein = false;
  SeqDesc seq{};
  DevBuf tw, chirp, kspec;
};

// `order` > 0 picks radix_variants(L)[order] for a runtime-radix sequence
// (ignored where L has fewer variants or with `pfa`).
int build_seq(size_t L, bool f64, SeqDesc& seq, DevBuf& tw, bool pfa = false, int order = 0,
              const std::vector<int>* given = nullptr) {
  std::vector<int> rad = given ? *given : radices_for(L, pfa);
  if (!given && !pfa && order > 0) {
    const std::vector<std::vector<int>> v = radix_variants(L);
    if ((size_t)order < v.size()) rad = v[order];
  }
  if ((int)rad.size() > kMaxPass) return CHFFT_ERR_UNSUPPORTED;
  seq.L = (int)L;
  seq.npass = (int)rad.size();
  std::vector<ld_t> tab;
  int p = 1;
  for (size_t q = 0; q < rad.size(); q++) {
    const int R = rad[q];
    PassDesc& pd = seq.pass[q];
    pd.radix = R;
    pd.p = p;
    pd.tws = (int)(L / ((size_t)p * R));
    pd.pmagic = (unsigned)((((unsigned long long)1 << 32) + (unsigned long long)p - 1) / (unsigned long long)p);
    pd.twoff = (int)(tab.size() / 2);
    if (p > 1) {
      // [j-1][k] = W_{pR}^(j*k), k < p, j = 1..R-1
      for (int j = 1; j < R; j++)
        for (int k = 0; k < p; k++) {
          ld_t re, im;
          unit_root((unsigned long long)j * k, (unsigned long long)p * R, &re, &im);
          tab.push_back(re);
          tab.push_back(im);
        }
    }
    p *= R;
  }
  // compact tables (the length-specialised kernels' form), pass after pass
  p = 1;
  for (size_t q = 0; q < rad.size(); q++) {
    const int R = rad[q];
    auto put = [&](unsigned long long num, unsigned long long den) {
      ld_t re, im;
      unit_root(num, den, &re, &im);
      tab.push_back(re);
      tab.push_back(im);
    };
    if (p > 1 && tw_two_level(R, p)) {
      for (int j = 1; j < R; j++)
        for (int k = 0; k < 16; k++) put((unsigned long long)j * k, (unsigned long long)p * R);
      for (int j = 1; j < R; j++)
        for (int k = 0; k < p / 16; k++) put((unsigned long long)j * k * 16, (unsigned long long)p * R);
    } else if (p > 1) {
      for (int j = 1; j < R; j++)
        for (int k = 0; k < p; k++) put((unsigned long long)j * k, (unsigned long long)p * R);
    }
    p *= R;
  }
  if (tab.empty()) {
    tab.push_back(1.0L);
    tab.push_back(0.0L);
  }
  int rc = upload(tab, f64, tw);
  seq.tw = tw.p;
  return rc;
}

// Four-step inter-pass twiddle table T[k1 * n2 + c2] = W_total^(k1 * c2),
// k1 < n1, c2 < n2 (column pass: k = k1, c = column; Bluestein row pass:
// c = k1, k = element).
std::vector<ld_t> fourstep_table(size_t n1, size_t n2) {
  std::vector<ld_t> v(2 * n1 * n2);
  for (size_t k1 = 0; k1 < n1; k1++)
    for (size_t c = 0; c < n2; c++)
      unit_root((unsigned long long)k1 * c, (unsigned long long)n1 * n2, &v[2 * (k1 * n2 + c)],
                &v[2 * (k1 * n2 + c) + 1]);
  return v;
}

// Can a length-L inner transform be fused into one workgroup?
// Four-step split total = n1 * n2 (n1 <= n2, both 7-smooth): rows of the
// column pass 128-byte aligned (n2 a multiple of 16) if possible, then the
// most balanced; 0 if there is none.
size_t fourstep_split(size_t total, bool f64) {
  (void)f64;
  size_t best1 = 0;
  bool best_aligned = false;
  for (size_t d = 1; d * d <= total; d++) {
    if (total % d != 0 || !smooth7(d) || !smooth7(total / d)) continue;
    const bool aligned = (total / d) % 16 == 0;
    if (aligned || !best_aligned) {
      best1 = d;
      best_aligned = aligned;
    }
  }
  return best1;
}

// Padded length of a four-step Bluestein. The reference pads to
// M = 2 * next_pow2(N) (cfft1d.rs:114); every M >= 2N - 2 holds the same
// linear convolution of the chirped input a (indices 0..N-1) with the chirp
// filter h (indices -(N-1)..N-1), so the results agree to rounding: the
// cyclic convolution over M reads h at (k - j) mod M for k, j < N, and the
// only two lags that share a slot when M = 2N - 2 are N-1 and -(N-1), where
// h is symmetric (h[m] = h[-m] = W^(-m^2/2)) and the slot holds the right
// value for both. Taken here: the 5-smooth M in [2N - 2, 2 * next_pow2(N)]
// with an aligned split (n2 % 16 == 0, n2 fusable) that needs the least work,
// M * (passes of n1 + passes of n2), a factor counting its prime-factor
// schedule where a specialised kernel runs it -- 20480 = 128 x 160 for
// N = 10007 and, for N = 65537, 131072 = 256 x 512 = 2N - 2, instead of
// 32768 and 262144.
size_t bluestein_length(size_t N, bool f64) {
  const size_t hi = 2 * next_pow2(N);
  size_t best = hi;
  double best_cost = 1e300;
  auto passes = [&](size_t L) {
    const std::vector<int> rp = radices_for(L, true);
    const size_t plain = radices_for(L).size();
    return sfft_has_schedule(f64, (int)L, rp.data(), (int)rp.size()) ? std::min(plain, rp.size()) : plain;
  };
  for (size_t m = 2 * N - 2; m <= hi; m++) {
    if (!smooth5(m)) continue;
    const size_t d = fourstep_split(m, f64);
    if (d == 0 || (m / d) % 16 != 0 || (m / d) * 4 > lds_cap(f64)) continue;
    const double cost = (double)m * (double)(passes(d) + passes(m / d));
    if (cost < best_cost) {
      best_cost = cost;
      best = m;
    }
  }
  return best;
}

bool fusable(size_t L, bool f64) {
  const size_t cap = lds_cap(f64);
  if (smooth7(L)) return L <= cap;
  return 2 * next_pow2(L) <= cap;
}

// M_special > 0: a fused Bluestein over that padded length, which only a
// length-specialised kernel runs (its tables follow that kernel's
// prime-factor schedule); else M = 2 * next_pow2(L) as cfft1d.rs:114.
int build_inner(size_t L, bool f64, Inner& in, size_t M_special = 0, int order = 0) {
  in.L = L;
  if (smooth7(L)) {
    in.Lfft = L;
    in.bluestein = false;
    return build_seq(L, f64, in.seq, in.tw, false, order);
  }
  size_t M = M_special ? M_special : 2 * next_pow2(L);
  in.Lfft = M;
  in.bluestein = true;
  // a padded length only a specialised kernel runs: that kernel's schedule
  int rs[8];
  const int nrs = M_special ? sfft_schedule_of(f64, (int)M, LD_PLAIN, ST_PLAIN, 1, false, rs) : 0;
  const std::vector<int> given(rs, rs + nrs);
  int rc = build_seq(M, f64, in.seq, in.tw, M_special != 0, order, nrs ? &given : nullptr);
  if (rc) return rc;
  if ((rc = upload(chirp_table(L), f64, in.chirp))) return rc;
  return upload(bluestein_spectrum(L, M), f64, in.kspec);
}

void fill_inner(KParams& P, const Inner& in) {
  P.Lfft = (int)in.Lfft;
  P.L = (int)in.L;
  P.seq = in.seq;
  P.has_mid = in.bluestein ? 1 : 0;
  P.has_chirp = in.bluestein ? 1 : 0;
  P.chirp = in.chirp.p;
  P.kspec = in.kspec.p;
  P.kc = 0;
  P.ke = 1;
}

int device_cus() {
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) return 256;
  return n;
}

// Largest byte offset of one tile's accesses (tpb consecutive transforms of
// len elements under d) from the tile's first transform; -1 when an offset
// could be negative (transform bases not increasing with g).
long long tile_span_bytes(const IoDesc& d, long long G, int tpb, long long len, size_t esz) {
  if (d.bdist < 0 || d.cstride < 0 || d.estride < 0 || d.ncols < 1) return -1;
  if (d.ncols > 1 && d.bdist < (long long)(d.ncols - 1) * d.cstride) return -1;
  auto base = [&](long long g) { return (g / d.ncols) * d.bdist + (g % d.ncols) * d.cstride; };
  long long worst = 0;
  for (long long g0 = 0; g0 < G; g0 += tpb) worst = std::max(worst, base(std::min(G, g0 + tpb) - 1) - base(g0));
  return (worst + (len - 1) * d.estride + 1) * (long long)esz;
}

std::string seq_str(const SeqDesc& s) {
  std::string r = "L=" + std::to_string(s.L) + " radix[";
  for (int i = 0; i < s.npass; i++) r += (i ? "," : "") + std::to_string(s.pass[i].radix);
  return r + "]";
}

bool same_schedule(const SfftEntry* e, const SeqDesc& s) {
  if (e->nrad != s.npass) return false;
  for (int q = 0; q < e->nrad; q++)
    if (e->rad[q] != s.pass[q].radix) return false;
  return true;
}

// Use a length-specialised kernel for a fused launch when one is registered
// for this length, operator pair and radix schedule: the launch's own
// schedule, or `alt` (the same length under prime-factor radices, with its
// own twiddle tables), which then replaces it.
void try_specialise(Launch& l, bool f64, const SeqDesc* alt = nullptr) {
  if (l.kind != 0) return;
  const int mode = l.P.has_mid ? (l.P.has_chirp ? 1 : 2) : 0;
  const bool col = l.P.colmajor_ld != 0;
  const SfftEntry* e = sfft_find(f64, l.P.Lfft, l.ld, l.st, mode, col);
  if (!e) return;
  const bool use_alt = !same_schedule(e, l.P.seq);
  if (use_alt && !(alt && alt->L == l.P.Lfft && same_schedule(e, *alt))) return;
  // a direct (non-epilogue) store uses the load's thread layout
  if (!e->epi && l.P.colmajor_st != l.P.colmajor_ld) return;
  // plain loads and plain / twiddled stores address each tile by 32-bit
  // byte offsets from its first transform
  const size_t esz = f64 ? 16 : 8;
  const long long lim = 1LL << 31;
  if (l.ld == LD_PLAIN) {
    const long long sp = tile_span_bytes(l.P.in, l.P.G, e->tpb, l.P.Lfft, esz);
    if (sp < 0 || sp >= lim) return;
  }
  if (l.ld == LD_FS_CHIRP) {  // sequence offsets s < N, and every s the tile forms, in 32 bits
    const long long smax = (long long)(l.P.in.ncols - 1) * l.P.in.cstride + (long long)(l.P.Lfft - 1) * l.P.in.estride;
    if (l.P.N * (long long)esz >= lim || smax >= lim || l.P.in.cstride < 0 || l.P.in.estride < 0) return;
  }
  if (l.st == ST_PLAIN || l.st == ST_TW) {
    const long long sp = tile_span_bytes(l.P.out, l.P.G, e->tpb, l.P.Lfft, esz);
    if (sp < 0 || sp >= lim) return;
  }
  if (mode != 0 && (l.P.ke < 0 || (long long)l.P.Lfft * l.P.ke * (long long)esz >= lim)) return;
  if (l.st == ST_TW) {
    if (l.P.twsh != 0) return;
    const long long top = ((long long)(l.P.Lfft - 1) * l.P.twk + (long long)(l.P.out.ncols - 1) * l.P.twc + 1) * esz;
    if (l.P.twk < 0 || l.P.twc < 0 || top >= lim) return;
  }
  l.kind = 3;
  l.sf = e;
  l.P.tpb = e->tpb;
  l.P.Lp = sfft_lds_stride(e);
  l.grid = dim3((unsigned)((l.P.G + e->tpb - 1) / e->tpb));
  l.block = dim3((unsigned)(e->tpt * e->tpb));
  l.lds = sfft_lds_bytes(e);
  if (use_alt) {
    const std::string a = seq_str(*alt), o = seq_str(l.P.seq);
    const size_t at = l.what.find(o);
    if (at != std::string::npos) l.what.replace(at, o.size(), a);
    l.P.seq = *alt;
  }
  l.what = "specialised " + l.what;
}

IoDesc io(long long bdist, long long cstride, long long estride, int ncols) {
  IoDesc d;
  d.bdist = bdist;
  d.cstride = cstride;
  d.estride = estride;
  d.ncols = ncols;
  return d;
}

}  // namespace

// Twiddle recurrence table of a registered four-step store kernel (ST_TW):
// the kernel's thread u of a transform holds outputs k = u + b TPT + j p
// (b < BL, j < RL; RL the last radix, p = L / RL its span, BL TPT >= the
// last pass's butterflies) and forms W_Ntw^(c k) from the base row u + b TPT
// and the step row K0 = BL TPT of this table, in FP64 (each twiddle rounded
// once to the transform's precision): one table read per row of outputs
// instead of one per output. Layout [r][c], c < ncols.
int upload_tw_recurrence(const SfftEntry* e, size_t Ntw, size_t ncols, DevBuf& dst) {
  const int rl = e->rad[e->nrad - 1];
  const int span = e->L / rl;  // also the last pass's butterfly count
  const int K0 = (span + e->tpt - 1) / e->tpt * e->tpt;
  std::vector<ld_t> v(2 * (size_t)(K0 + 1) * ncols);
  for (int r = 0; r <= K0; r++) {
    const unsigned long long m = r < K0 ? (unsigned long long)r : (unsigned long long)span;
    for (size_t c = 0; c < ncols; c++)
      unit_root(m * c, Ntw, &v[2 * ((size_t)r * ncols + c)], &v[2 * ((size_t)r * ncols + c) + 1]);
  }
  return upload(v, true, dst);
}

// Complex-to-complex engine for one length (used by CFft1D, both CFft2D axes).
struct C2CEngine {
  size_t n = 0;
  bool f64 = false;
  int mode = 0;  // 0 copy (n == 1), 1 fused, 2 four-step 7-smooth, 3 four-step Bluestein
  Inner inner;   // mode 1
  size_t n1 = 0, n2 = 0;
  SeqDesc s1{}, s2{};
  SeqDesc x1{}, x2{};  // prime-factor schedules of n1, n2 (L = 0: none registered)
  DevBuf tw1, tw2, twx1, twx2, twN, chirp, kspec;
  DevBuf twr_a, twr_b;  // twiddle recurrence tables of the column / Bluestein row store kernels
  size_t M = 0;
  Tune tune;  // planner overrides (the autotuner's candidates)

  int init(size_t len, bool use_f64) {
    n = len;
    f64 = use_f64;
    // CHFFT_FOURSTEP=split|persistent overrides the batch-size rule below
    // (used by the tests to cover both schedules)
    if (const char* e = getenv("CHFFT_FOURSTEP")) {
      force_split = strcmp(e, "split") == 0;
      force_persistent = strcmp(e, "persistent") == 0;
    }
    if (tune.sched) {
      force_split = tune.sched == 2;
      force_persistent = tune.sched == 1;
    }
    if (len == 1) {
      mode = 0;
      return CHFFT_OK;
    }
    if (fusable(len, f64)) {
      mode = 1;
      return build_inner(len, f64, inner, 0, tune.order);
    }
    // a non-7-smooth length whose padded Bluestein sequence a specialised
    // kernel holds in one CU's LDS: one fused launch instead of three
    // (N = 10007 over M = 20160 in FP32)
    if (!smooth7(len) && !force_split && !force_persistent && !tune.generic)
      if (const int Ms = sfft_fused_bluestein_m(f64, (long long)len)) {
        mode = 1;
        big_fused = true;
        return build_inner(len, f64, inner, (size_t)Ms);
      }
    mode = smooth7(len) ? 2 : 3;
    const size_t total = mode == 2 ? len : bluestein_length(len, f64);
    M = total;
    size_t best1 = fourstep_split(total, f64);
    if (tune.n1) {  // any 7-smooth split whose factors one workgroup each holds
      const size_t t1 = (size_t)tune.n1;
      if (total % t1 || !smooth7(t1) || !smooth7(total / t1) || t1 > lds_cap(f64)) return CHFFT_ERR_INVALID_ARGUMENT;
      best1 = t1;
    }
    if (best1 == 0) return CHFFT_ERR_UNSUPPORTED;
    n1 = best1;
    n2 = total / best1;
    if (n2 * 4 > lds_cap(f64)) return tune.n1 ? CHFFT_ERR_INVALID_ARGUMENT : CHFFT_ERR_UNSUPPORTED;
    int rc;
    if ((rc = build_seq(n1, f64, s1, tw1, false, tune.order))) return rc;
    if ((rc = build_seq(n2, f64, s2, tw2, false, tune.order))) return rc;
    if ((rc = build_pfa(n1, x1, twx1))) return rc;
    if ((rc = build_pfa(n2, x2, twx2))) return rc;
    if ((rc = upload(fourstep_table(n1, n2), f64, twN))) return rc;
    if (!tune.generic) {
      const SfftEntry* ea = sfft_find(f64, (int)n1, mode == 2 ? LD_PLAIN : LD_FS_CHIRP, ST_TW, 0, true);
      if (ea && (rc = upload_tw_recurrence(ea, total, n2, twr_a))) return rc;
      const SfftEntry* eb = mode == 3 ? sfft_find(f64, (int)n2, LD_PLAIN, ST_TW, 2, false) : nullptr;
      if (eb && (rc = upload_tw_recurrence(eb, total, n1, twr_b))) return rc;
    }
    if (mode == 3) {
      if ((rc = upload(chirp_table(len), f64, chirp))) return rc;
      // the row pass reads the spectrum at bins k1 + n1 * k2 (row k1, element
      // k2): stored row-major, [k1][k2], so that consecutive lanes read
      // consecutive entries
      const std::vector<ld_t> K = bluestein_spectrum(len, total);
      std::vector<ld_t> Kt(K.size());
      for (size_t k1 = 0; k1 < n1; k1++)
        for (size_t k2 = 0; k2 < n2; k2++) {
          Kt[2 * (k1 * n2 + k2)] = K[2 * (k1 + n1 * k2)];
          Kt[2 * (k1 * n2 + k2) + 1] = K[2 * (k1 + n1 * k2) + 1];
        }
      if ((rc = upload(Kt, f64, kspec))) return rc;
    }
    return CHFFT_OK;
  }

  // The prime-factor schedule of L when it differs from the plain one and a
  // specialised kernel runs it.
  int build_pfa(size_t L, SeqDesc& seq, DevBuf& tw) const {
    seq = SeqDesc{};
    const std::vector<int> rp = radices_for(L, true);
    if (rp == radices_for(L) || !sfft_has_schedule(f64, (int)L, rp.data(), (int)rp.size())) return CHFFT_OK;
    return build_seq(L, f64, seq, tw, true);
  }

  // scratch of the split (multi-launch) four-step
  size_t scratch_elems(size_t batch) const {
    if (persistent_ok(batch)) return 0;
    return (mode == 2 ? n : mode == 3 ? M : 0) * chunk_of(batch);
  }

  // Appends the launches for `batch` transforms; `inverse` selects the
  // conjugation pair; src/dst slots as given.
  void emit(std::vector<Launch>& out, size_t batch, bool inverse, int src, int dst, long long dist_in,
            long long dist_out) const {
    const int ci = inverse ? 1 : 0;
    if (mode == 0) {
      Launch l{};
      l.kind = 1;  // copy
      l.src = src;
      l.dst = dst;
      l.P.G = (int)batch;
      l.what = "copy (length 1)";
      out.push_back(l);
      return;
    }
    if (mode == 1) {
      Launch l{};
      l.ld = LD_PLAIN;
      l.st = ST_PLAIN;
      fill_inner(l.P, inner);
      l.P.G = (int)batch;
      l.P.in = io(dist_in, 0, 1, 1);
      l.P.out = io(dist_out, 0, 1, 1);
      l.P.conj_in = ci;
      l.P.conj_out = ci;
      Geometry g = choose_geometry((int)inner.Lfft, (long long)batch, false, f64);
      if (tune.tpb || tune.tpt) {  // checked when the plan is built (geometry_ok)
        Geometry t;
        if (geometry_with((int)inner.Lfft, tune.tpb ? tune.tpb : g.tpb, tune.tpt ? tune.tpt : g.tpt, f64, t)) g = t;
      }
      finish(l, g, src, dst, 1);
      l.what = std::string(inner.bluestein ? "fused bluestein N=" + std::to_string(n) + " M " : "fused ") +
               seq_str(inner.seq);
      if (!tune.generic) try_specialise(l, f64);
      // the generic fused kernel cannot hold such a length; the planner only
      // chooses it where the specialised kernel applies (plain row tiles)
      if (big_fused && l.kind != 3) l.kind = -1;
      out.push_back(l);
      return;
    }
    if (persistent_ok(batch)) {
      emit_persistent(out, batch, inverse, src, dst, dist_in, dist_out);
      return;
    }
    // split launches, chunk by chunk through one scratch of chunk transforms
    const size_t cb = chunk_of(batch);
    const long long esz = f64 ? 16 : 8;
    for (size_t c0 = 0; c0 < batch; c0 += cb) {
      const size_t first = out.size();
      emit_split(out, std::min(cb, batch - c0), inverse, src, dst, dist_in, dist_out);
      for (size_t i = first; i < out.size(); i++) {
        Launch& l = out[i];
        if (l.src != S_SCRATCH) l.src_off = (long long)c0 * dist_in * esz;
        if (l.dst != S_SCRATCH) l.dst_off = (long long)c0 * dist_out * esz;
        if (cb < batch) l.what += " (transforms " + std::to_string(c0) + "+)";
      }
    }
  }

  // Transforms per chunk of the split four-step. A chunk's scratch (its
  // intermediate spectra) is written by the column launch and read back by
  // the row launch; a 7-smooth four-step whose scratch would exceed
  // kChunkScratchBytes runs in equal chunks below it, so that the scratch
  // stays in the 256 MB memory-side cache. Measured on MI355X, 43200 x 512
  // with rotated 1 GB buffers: FP32 155 -> 142 us (2 chunks, 88 MB each;
  // 4 chunks: 166 us), FP64 301 -> 286 us (4 chunks; 2 chunks of 177 MB:
  // 325 us). The three-launch Bluestein four-step (N = 10007 x 512,
  // 65537 x 64) ran 10-25 % slower in chunks and keeps the whole batch.
  // CHFFT_FS_CHUNK overrides (transforms per chunk; -1: whole batch; 0:
  // the rule above).
  static constexpr size_t kChunkScratchBytes = (size_t)96 << 20;
  size_t chunk_of(size_t batch) const {
    size_t c = batch;
    if (mode == 2) {
      const size_t per = n * (f64 ? 16 : 8);
      const size_t cmax = std::max<size_t>(1, kChunkScratchBytes / per);
      const size_t k = (batch + cmax - 1) / cmax;
      c = (batch + k - 1) / k;
    }
    if (const char* e = getenv("CHFFT_FS_CHUNK")) {
      const long long v = atoll(e);
      if (v > 0) c = (size_t)v;
      if (v == -1) c = batch;
    }
    if (tune.chunk > 0) c = (size_t)tune.chunk;
    if (tune.chunk == -1) c = batch;
    return std::max<size_t>(1, std::min(c, batch));
  }

  void emit_split(std::vector<Launch>& out, size_t batch, bool inverse, int src, int dst, long long dist_in,
                  long long dist_out) const {
    const int ci = inverse ? 1 : 0;
    if (mode == 2) {
      // K1: columns (length n1, stride n2) + twiddle W_n^(n2*k1) -> scratch
      Launch a{};
      a.ld = LD_PLAIN;
      a.st = ST_TW;
      plain_seq(a.P, s1, n1);
      a.P.G = (int)(batch * n2);
      a.P.in = io(dist_in, 1, (long long)n2, (int)n2);
      a.P.out = io((long long)n, 1, (long long)n2, (int)n2);
      a.P.colmajor_ld = a.P.colmajor_st = 1;
      a.P.conj_in = ci;
      a.P.twN = (long long)n;
      a.P.tab_a = twN.p;
      a.P.twk = (int)n2;
      a.P.twc = 1;
      finish(a, choose_geometry((int)n1, a.P.G, true, f64), src, S_SCRATCH, 0);
      a.what = "four-step columns " + seq_str(s1) + " + twiddle";
      if (!tune.generic) try_specialise(a, f64, &x1);
      if (a.kind == 3 && twr_a.p) a.P.tw_rec = twr_a.p;
    out.push_back(a);
      // K2: rows (length n2) -> out[k1 + n1*k2]
      Launch b{};
      b.ld = LD_PLAIN;
      b.st = ST_PLAIN;
      plain_seq(b.P, s2, n2);
      b.P.G = (int)(batch * n1);
      b.P.in = io((long long)n, (long long)n2, 1, (int)n1);
      b.P.out = io(dist_out, 1, (long long)n1, (int)n1);
      b.P.colmajor_st = 1;
      b.P.conj_out = ci;
      finish(b, choose_geometry((int)n2, b.P.G, true, f64), S_SCRATCH, dst, 1);
      b.what = "four-step rows " + seq_str(s2) + " + transpose";
      if (!tune.generic) try_specialise(b, f64, &x2);
    out.push_back(b);
      return;
    }
    // mode 3: Bluestein over a four-step M = n1 * n2
    Launch a{};
    a.ld = LD_FS_CHIRP;
    a.st = ST_TW;
    plain_seq(a.P, s1, n1);
    a.P.G = (int)(batch * n2);
    a.P.N = (long long)n;
    a.P.chirp = chirp.p;
    a.P.in = io(dist_in, 1, (long long)n2, (int)n2);
    a.P.out = io((long long)M, 1, (long long)n2, (int)n2);
    a.P.colmajor_ld = a.P.colmajor_st = 1;
    a.P.conj_in = ci;
    a.P.twN = (long long)M;
    a.P.tab_a = twN.p;
    a.P.twk = (int)n2;
    a.P.twc = 1;
    finish(a, choose_geometry((int)n1, a.P.G, true, f64), src, S_SCRATCH, 0);
    a.what = "bluestein N=" + std::to_string(n) + " chirp + columns " + seq_str(s1) + " + twiddle";
    if (!tune.generic) try_specialise(a, f64, &x1);
    if (a.kind == 3 && twr_a.p) a.P.tw_rec = twr_a.p;
    out.push_back(a);
    Launch b{};
    b.ld = LD_PLAIN;
    b.st = ST_TW;
    plain_seq(b.P, s2, n2);
    b.P.G = (int)(batch * n1);
    b.P.in = io((long long)M, (long long)n2, 1, (int)n1);
    b.P.out = io((long long)M, (long long)n2, 1, (int)n1);
    b.P.has_mid = 1;
    b.P.kspec = kspec.p;
    b.P.kc = (long long)n2;
    b.P.ke = 1;
    b.P.twN = (long long)M;
    b.P.tab_a = twN.p;
    b.P.twk = 1;
    b.P.twc = (int)n2;
    finish(b, choose_geometry((int)n2, b.P.G, false, f64), S_SCRATCH, S_SCRATCH, 0);
    b.what = "bluestein rows " + seq_str(s2) + " x spectrum, inverse rows + twiddle";
    if (!tune.generic) try_specialise(b, f64, &x2);
    if (b.kind == 3 && twr_b.p) b.P.tw_rec = twr_b.p;
    out.push_back(b);
    Launch c{};
    c.ld = LD_PLAIN;
    c.st = ST_FS_DECHIRP;
    plain_seq(c.P, s1, n1);
    c.P.G = (int)(batch * n2);
    c.P.N = (long long)n;
    c.P.chirp = chirp.p;
    c.P.in = io((long long)M, 1, (long long)n2, (int)n2);
    c.P.out = io(dist_out, 1, (long long)n2, (int)n2);
    c.P.colmajor_ld = c.P.colmajor_st = 1;
    c.P.conj_out = ci;
    finish(c, choose_geometry((int)n1, c.P.G, true, f64), S_SCRATCH, dst, 1);
    c.what = "bluestein inverse columns " + seq_str(s1) + " + de-chirp";
    if (!tune.generic) try_specialise(c, f64, &x1);
    out.push_back(c);
  }

  // One persistent launch pays off once the batch keeps every CU busy with
  // whole transforms; smaller batches split the phases over launches so that
  // all CUs share each transform.
  bool persistent_ok(size_t batch) const {
    if (force_split) return false;
    if (force_persistent) return true;
    // the persistent launch needs a registered specialised instantiation
    // (the generic persistent kernel is kept for forced use only)
    // Measured on MI355X (43200 x 512 FP32): the persistent launch, one
    // transform per workgroup, ran at 449 us against 258 us for the split
    // launches (its tiles run back to back with little memory-level
    // parallelism), so it is only used when requested.
    (void)batch;
    return false;
  }

  static int num_cus() { return device_cus(); }

  int fourstep_grid(size_t batch) const {
    // phases share one workgroup size and LDS size (set in emit_persistent)
    int occ = fourstep_occupancy(f64, mode == 2 ? 2 : 3, p_block, p_lds);
    if (occ < 1) occ = 1;
    long long g = (long long)occ * num_cus();
    if (g > (long long)batch) g = (long long)batch;
    return (int)std::max(1LL, g);
  }

  mutable int p_block = 0;
  mutable size_t p_lds = 0;
  bool force_split = false, force_persistent = false;
  bool big_fused = false;  // mode 1 over a length only a specialised kernel runs

  void emit_persistent(std::vector<Launch>& out, size_t batch, bool inverse, int src, int dst, long long dist_in,
                       long long dist_out) const {
    const int ci = inverse ? 1 : 0;
    Launch l{};
    l.kind = 2;
    l.nph = mode == 2 ? 2 : 3;
    KParams& A = l.Q[0];
    KParams& B = l.Q[1];
    KParams& C = l.Q[2];
    plain_seq(A, s1, n1);
    A.G = (int)n2;
    A.in = io(0, 1, (long long)n2, (int)n2);
    A.out = A.in;
    A.colmajor_ld = A.colmajor_st = 1;
    A.conj_in = ci;
    A.twN = (long long)(mode == 2 ? n : M);
    A.tab_a = twN.p;
    A.twk = (int)n2;
    A.twc = 1;
    plain_seq(B, s2, n2);
    B.G = (int)n1;
    B.in = io(0, (long long)n2, 1, (int)n1);
    if (mode == 2) {
      B.out = io(0, 1, (long long)n1, (int)n1);
      B.colmajor_st = 1;
      B.conj_out = ci;
    } else {
      A.N = (long long)n;
      A.chirp = chirp.p;
      B.out = B.in;
      B.has_mid = 1;
      B.kspec = kspec.p;
      B.kc = (long long)n2;
      B.ke = 1;
      B.twN = (long long)M;
      B.tab_a = twN.p;
      B.twk = 1;
      B.twc = (int)n2;
      plain_seq(C, s1, n1);
      C.G = (int)n2;
      C.in = io(0, 1, (long long)n2, (int)n2);
      C.out = C.in;
      C.colmajor_ld = C.colmajor_st = 1;
      C.N = (long long)n;
      C.chirp = chirp.p;
      C.conj_out = ci;
    }
    if (const SfsEntry* e = tune.generic ? nullptr : sfs_find(f64, (int)n1, (int)n2, l.nph)) {
      int threads, tpb[3], Lp[3];
      size_t lds;
      sfs_info(e, &threads, tpb, Lp, &lds);
      A.tpb = tpb[0];
      A.Lp = Lp[0];
      B.tpb = tpb[1];
      B.Lp = Lp[1];
      C.tpb = tpb[2];
      C.Lp = Lp[2];
      l.kind = 4;
      l.sfs = e;
      l.block = dim3((unsigned)threads);
      l.lds = lds;
      l.batch = (int)batch;
      l.src_dist = dist_in;
      l.dst_dist = dist_out;
      l.slot_elems = (long long)(mode == 2 ? n : M);
      long long gsz = (long long)sfs_occupancy(e) * num_cus();
      l.grid = dim3((unsigned)std::max(1LL, std::min(gsz, (long long)batch)));
      l.src = src;
      l.dst = dst;
      l.scaled = 1;
      l.scale_mul = 1.0;
      l.what = std::string("specialised persistent four-step ") + (mode == 3 ? "bluestein " : "") + "columns " +
               seq_str(s1) + " x rows " + seq_str(s2);
      out.push_back(l);
      return;
    }
    const int fs_max = f64 ? kMaxThreadsFsF64 : kMaxThreadsFsF32;
    Geometry ga = choose_geometry((int)n1, 1 << 20, true, f64, fs_max);
    Geometry gb = choose_geometry((int)n2, 1 << 20, true, f64, fs_max);
    A.tpb = ga.tpb;
    A.Lp = ga.Lp;
    B.tpb = gb.tpb;
    B.Lp = gb.Lp;
    C.tpb = ga.tpb;
    C.Lp = ga.Lp;
    p_block = std::max(ga.block, gb.block);
    p_lds = std::max(ga.lds, gb.lds);
    l.block = dim3((unsigned)p_block);
    l.lds = p_lds;
    l.batch = (int)batch;
    l.src_dist = dist_in;
    l.dst_dist = dist_out;
    l.slot_elems = (long long)(mode == 2 ? n : M);
    l.grid = dim3((unsigned)fourstep_grid(batch));
    l.src = src;
    l.dst = dst;
    l.scaled = 1;
    l.scale_mul = 1.0;
    l.what = std::string(mode == 2 ? "persistent four-step " : "persistent bluestein N=" + std::to_string(n) + " ") +
             "columns " + seq_str(s1) + " x rows " + seq_str(s2) + " (one transform per workgroup iteration)";
    out.push_back(l);
  }

  size_t persistent_scratch_elems(const std::vector<Launch>& L) const {
    size_t need = 0;
    for (const Launch& l : L)
      if (l.kind == 2 || l.kind == 4) need = std::max(need, (size_t)l.grid.x * (size_t)l.slot_elems);
    return need;
  }

  static void plain_seq(KParams& P, const SeqDesc& s, size_t L) {
    P.seq = s;
    P.Lfft = (int)L;
    P.L = (int)L;
  }

  static void finish(Launch& l, const Geometry& g, int src, int dst, int scaled) {
    l.P.tpb = g.tpb;
    l.P.Lp = g.Lp;
    l.grid = dim3((unsigned)((l.P.G + g.tpb - 1) / g.tpb));
    l.block = dim3((unsigned)g.block);
    l.lds = g.lds;
    l.src = src;
    l.dst = dst;
    l.scaled = scaled;
    l.scale_mul = 1.0;
  }
};

}  // namespace chfft

using namespace chfft;

// Scratch of chfft_exec_noalloc: a file-scope device array that every GPU the
// library runs on holds until the process ends. Its size is the entry
// point's limit on a plan's four-step scratch plus work buffer.
static __device__ __align__(256) unsigned char noalloc_scratch[CHFFT_NOALLOC_SCRATCH_BYTES];

struct chfft_plan_s {
  Kind kind = K_C2C;
  bool f64 = false;
  size_t n = 0, m = 0, batch = 0, frame_stride = 0;
  int dct_type = 0;
  bool ortho = false;
  C2CEngine rows, cols;  // c2c / 2-D
  Inner inner;           // r2c / dct / mdct
  DevBuf tab_a, tab_b, tab_c;
  DevBuf scratch;
  DevBuf work;            // S_WORK: operator passes / transposed 2-D columns
  DevBuf ola;             // MDCT: windowed frames before the overlap-add
  bool long_inner = false;  // r2c / dct / mdct inner transform on the four-step engine (rows)
  DevBuf stage_in, stage_out;
  std::vector<std::unique_ptr<DevBuf>> qbufs;  // four-step phase descriptors
  std::vector<Launch> fwd, bwd;
  // multi-length batch (K_MULTI)
  std::vector<size_t> lens;
  size_t total = 0;  // complex elements over all transforms
  int nitems = 0;
  DevBuf arena, items_dev, units_dev, descs_fwd, descs_bwd, sync_dev;
  DevBuf arena64;  // FP64 two-level twiddle tables of the register-resident tiles' four-step stores
  int unit_ends[kMultiPhases] = {0, 0, 0};  // units of phases 0..2 end here
  size_t multi_lds = 0;
  int multi_grid = 0, multi_occ = 1;
  unsigned long long barrier_ticks = kBarrierTicks;
  unsigned barrier_extra = 0;
  std::string multi_desc;
  Tune tune;                // the planner overrides the plan was built with
  std::string tune_report;  // the last autotune's candidates and times
};

namespace chfft {
namespace {

std::once_flag g_lds_once;
hipError_t g_lds_err = hipSuccess;

// Device address of noalloc_scratch on each device, looked up once per
// device with that device current. In host code the array's name is only a
// host-side handle: kernels receive these addresses, never the name.
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_addr[kMaxDevices] = {};

int noalloc_base(void** base) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return CHFFT_ERR_HIP;
  if (dev < 0 || dev >= kMaxDevices) return CHFFT_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_addr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return CHFFT_ERR_HIP;
    }
    g_noalloc_addr[dev] = p;
  }
  *base = g_noalloc_addr[dev];
  return CHFFT_OK;
}

// Library set-up, run by every plan constructor: LDS limits of the kernels
// (once per process) and the no-allocation scratch address of the current
// device (once per device).
int ensure_runtime() {
  std::call_once(g_lds_once, [] {
    g_lds_err = set_fused_lds_limits();
    hipError_t e2 = sfft_set_lds_limits();
    hipError_t e3 = multi_set_lds_limits();
    if (g_lds_err == hipSuccess) g_lds_err = e2;
    if (g_lds_err == hipSuccess) g_lds_err = e3;
  });
  if (g_lds_err != hipSuccess) return CHFFT_ERR_HIP;
  void* base = nullptr;
  return noalloc_base(&base);
}

int check_prec(int precision) {
  return (precision == CHFFT_F32 || precision == CHFFT_F64) ? CHFFT_OK : CHFFT_ERR_INVALID_ARGUMENT;
}

// Fused single-launch recipe for r2c / dct / mdct around plan->inner.
Launch fused_op(const chfft_plan_s* p, int ld, int st, bool inner_inverse, long long in_dist, long long out_dist,
                int G) {
  Launch l{};
  l.ld = ld;
  l.st = st;
  fill_inner(l.P, p->inner);
  l.P.G = G;
  l.P.in = io(in_dist, 0, 1, 1);
  l.P.out = io(out_dist, 0, 1, 1);
  l.P.conj_in = l.P.conj_out = inner_inverse ? 1 : 0;
  Geometry g = choose_geometry((int)p->inner.Lfft, G, false, p->f64);
  C2CEngine::finish(l, g, S_IN, S_OUT, 1);
  l.what = std::string(p->inner.bluestein ? "fused bluestein " : "fused ") + seq_str(p->inner.seq);
  try_specialise(l, p->f64);
  return l;
}

// A stand-alone operator pass of a long-length schedule (kind 5: load
// operator `op` into the work layout; kind 6: store operator `op` from it):
// G transforms of L complex work elements, the operator side at in_dist /
// out_dist elements per transform.
Launch op_pass(int kind, int op, size_t L, long long in_dist, long long out_dist, int src, int dst, size_t G) {
  Launch l{};
  l.kind = kind;
  l.ld = kind == 5 ? op : LD_PLAIN;
  l.st = kind == 6 ? op : ST_PLAIN;
  l.P.G = (int)G;
  l.P.L = l.P.Lfft = (int)L;
  l.P.in = io(in_dist, 0, 1, 1);
  l.P.out = io(out_dist, 0, 1, 1);
  l.src = src;
  l.dst = dst;
  l.scaled = kind == 6 ? 1 : 0;
  l.scale_mul = 1.0;
  static const char* const lds[] = {"plain", "chirp", "RFft un-split", "DCT-II pack", "DCT-III pre-twiddle",
                                    "MDCT fold", "IMDCT pre-twiddle"};
  static const char* const sts[] = {"plain", "twiddle", "de-chirp", "RFft split", "DCT-II post-twiddle",
                                    "DCT-III unpack", "MDCT post-twiddle", "IMDCT unfold"};
  l.what = std::string(kind == 5 ? "operator pre-pass " : "operator post-pass ") + (kind == 5 ? lds[op] : sts[op]) +
           " L=" + std::to_string(L);
  return l;
}

// The inner complex transform of a long-length plan: `batch` transforms of
// h points on the complex engine (p->rows) between the given slots; its
// launches apply no scale (the operator post-pass or the caller does).
int long_inner_init(chfft_plan_s* p, size_t h, size_t batch) {
  int rc = p->rows.init(h, p->f64);
  if (rc) return rc;
  p->rows.force_split = true;
  p->rows.force_persistent = false;
  p->long_inner = true;
  return dev_alloc(p->work, batch * h * esz_of(p->f64));
}

void emit_unscaled(const chfft_plan_s* p, std::vector<Launch>& L, size_t batch, bool inverse, int src, int dst,
                   long long dist) {
  const size_t first = L.size();
  p->rows.emit(L, batch, inverse, src, dst, dist, dist);
  for (size_t i = first; i < L.size(); i++) L[i].scaled = 0;
}

Launch transpose_launch(int src, int dst, size_t R, size_t C, size_t batch) {
  Launch l{};
  l.kind = 7;
  l.P.G = (int)batch;
  l.P.Lfft = (int)R;
  l.P.L = (int)C;
  l.src = src;
  l.dst = dst;
  l.scaled = 0;
  l.scale_mul = 1.0;
  l.what = "transpose " + std::to_string(R) + " x " + std::to_string(C);
  return l;
}

// Launches of a registered kernel that has a non-temporal-output build
// (sfft_kernel's NT variants) use it when they write the exec's output and
// no later launch of the schedule reads it back (an in-place CFft2D column
// pass, an operator post-pass): the L2 then streams those lines out instead
// of keeping them. Measured on MI355X with rotated 1 GB buffers: RFft1D FP64
// 8192 x 4096 120 -> 100 us, MDCT FP32 2048 over 2^24 samples 42 -> 38 us,
// CFft1D FP64 4096 x 256 11.1 -> 10.2 us, Bluestein 4093 x 2048 89.6 ->
// 86.0 us (FP32) and 10007 x 512 87.5 -> 84.1 us (FP32). Kernels without
// such a build measured no gain or a loss: transposed four-step row stores
// (43200 x 512: 145 -> 159 us), DCT pairs and scattered reals, the four-step
// Bluestein de-chirp, and every intermediate that a later launch reads.
void mark_nt_outputs(std::vector<Launch>& L) {
  for (size_t i = 0; i < L.size(); i++) {
    Launch& l = L[i];
    l.P.nt_out = 0;
    if (l.dst != S_OUT || l.kind != 3) continue;
    bool reread = false;
    for (size_t j = i + 1; j < L.size(); j++) reread = reread || L[j].src == S_OUT;
    if (!reread) l.P.nt_out = 1;
  }
}

int finish_plan(chfft_plan_s* p, chfft_plan* out) {
  mark_nt_outputs(p->fwd);
  mark_nt_outputs(p->bwd);
  for (const std::vector<Launch>* L : {&p->fwd, &p->bwd})
    for (const Launch& l : *L)
      if (l.kind < 0) {  // a schedule no kernel runs (see C2CEngine::emit)
        delete p;
        return CHFFT_ERR_UNSUPPORTED;
      }
  size_t need = 0;
  if (p->kind == K_C2C) need = p->rows.scratch_elems(p->batch);
  if (p->kind == K_C2C2D) {
    need = p->rows.scratch_elems(p->batch * p->m);
    if (p->cols.mode >= 2) need = std::max(need, p->cols.scratch_elems(p->batch * p->n));
  }
  if (p->long_inner) need = p->rows.scratch_elems(p->batch);
  need = std::max(need, p->rows.persistent_scratch_elems(p->fwd));
  need = std::max(need, p->rows.persistent_scratch_elems(p->bwd));
  for (std::vector<Launch>* L : {&p->fwd, &p->bwd})
    for (Launch& l : *L)
      if (l.kind == 2 || l.kind == 4) {
        p->qbufs.emplace_back(new DevBuf());
        DevBuf& q = *p->qbufs.back();
        int rc2 = dev_alloc(q, sizeof(KParams) * l.nph);
        if (!rc2 && hipMemcpy(q.p, l.Q, sizeof(KParams) * l.nph, hipMemcpyHostToDevice) != hipSuccess)
          rc2 = CHFFT_ERR_HIP;
        if (rc2) {
          delete p;
          return rc2;
        }
        l.qdev = q.p;
      }
  int rc = dev_alloc(p->scratch, need * esz_of(p->f64));
  if (rc) {
    delete p;
    return rc;
  }
  *out = p;
  return CHFFT_OK;
}

// A CFft1D plan of `batch` transforms of `len` points under the planner
// overrides `t` (all zero: the planner's own choices). Overrides a plan
// cannot take are CHFFT_ERR_INVALID_ARGUMENT: tile knobs on a four-step
// length or a shape the fused kernel cannot run, a radix-order index the
// fused sequence does not have, a split that does not divide the length,
// a workgroup count.
int c2c_create(size_t len, size_t batch, bool f64, const Tune& t, chfft_plan* out) {
  if (t.grid) return CHFFT_ERR_INVALID_ARGUMENT;
  chfft_plan_s* p = new chfft_plan_s();
  p->kind = K_C2C;
  p->f64 = f64;
  p->n = len;
  p->batch = batch;
  p->tune = t;
  p->rows.tune = t;
  int rc = p->rows.init(len, f64);
  if (!rc) {
    const C2CEngine& e = p->rows;
    const bool fs = e.mode == 2 || e.mode == 3;
    if ((t.n1 || t.chunk || t.sched) && !fs) rc = CHFFT_ERR_INVALID_ARGUMENT;
    if ((t.tpb || t.tpt) && e.mode != 1) rc = CHFFT_ERR_INVALID_ARGUMENT;
    if (t.order && e.mode == 0) rc = CHFFT_ERR_INVALID_ARGUMENT;
    if (e.mode == 1) {
      Geometry g = choose_geometry((int)e.inner.Lfft, (long long)batch, false, f64), gt;
      if ((t.tpb || t.tpt) && !geometry_with((int)e.inner.Lfft, t.tpb ? t.tpb : g.tpb, t.tpt ? t.tpt : g.tpt, f64, gt))
        rc = CHFFT_ERR_INVALID_ARGUMENT;
      if (t.order && (e.big_fused || (size_t)t.order >= radix_variants(e.inner.Lfft).size()))
        rc = CHFFT_ERR_INVALID_ARGUMENT;
    }
    if (fs && t.order &&
        (size_t)t.order >= std::max(radix_variants(e.n1).size(), radix_variants(e.n2).size()))
      rc = CHFFT_ERR_INVALID_ARGUMENT;
  }
  if (rc) {
    delete p;
    return rc;
  }
  p->rows.emit(p->fwd, batch, false, S_IN, S_OUT, (long long)len, (long long)len);
  p->rows.emit(p->bwd, batch, true, S_IN, S_OUT, (long long)len, (long long)len);
  return finish_plan(p, out);
}

}  // namespace
}  // namespace chfft

extern "C" {

const char* chfft_status_string(int status) {
  switch (status) {
    case CHFFT_OK: return "ok";
    case CHFFT_ERR_INVALID_LENGTH: return "invalid length";
    case CHFFT_ERR_INVALID_ARGUMENT: return "invalid argument";
    case CHFFT_ERR_UNSUPPORTED: return "unsupported";
    case CHFFT_ERR_HIP: return "HIP runtime error";
    case CHFFT_ERR_OUT_OF_MEMORY: return "out of device memory";
    case CHFFT_ERR_INCOMPLETE: return "incomplete: a grid-wide barrier of the multi-length launch timed out";
    case CHFFT_ERR_SCRATCH_LIMIT: return "the plan's scratch exceeds the no-allocation limit";
  }
  return "unknown status";
}

int chfft_plan_c2c_1d(chfft_plan* plan, size_t len, size_t batch, int precision) {
  if (!plan) return CHFFT_ERR_INVALID_ARGUMENT;
  *plan = nullptr;
  if (check_prec(precision)) return CHFFT_ERR_INVALID_ARGUMENT;
  if (len == 0) return CHFFT_ERR_INVALID_LENGTH;
  if (batch == 0 || batch * len > (size_t)1 << 31) return CHFFT_ERR_INVALID_ARGUMENT;
  int rc = ensure_runtime();
  if (rc) return rc;
  return c2c_create(len, batch, precision == CHFFT_F64, Tune{}, plan);
}

int chfft_plan_c2c_2d(chfft_plan* plan, size_t len_m, size_t len_n, size_t batch, int precision) {
  if (!plan) return CHFFT_ERR_INVALID_ARGUMENT;
  *plan = nullptr;
  if (check_prec(precision)) return CHFFT_ERR_INVALID_ARGUMENT;
  if (len_m == 0 || len_n == 0) return CHFFT_ERR_INVALID_LENGTH;
  if (batch == 0 || batch * len_m * len_n > (size_t)1 << 31) return CHFFT_ERR_INVALID_ARGUMENT;
  int rc = ensure_runtime();
  if (rc) return rc;
  chfft_plan_s* p = new chfft_plan_s();
  p->kind = K_C2C2D;
  p->f64 = precision == CHFFT_F64;
  p->m = len_m;
  p->n = len_n;
  p->batch = batch;
  if ((rc = p->rows.init(len_n, p->f64))) {
    delete p;
    return rc;
  }
  p->rows.force_split = true;  // the column pass below continues from split launches
  p->rows.force_persistent = false;
  // columns: one fused launch in place on `out` when the column length fits
  // one workgroup; otherwise the matrices are transposed into the work
  // buffer, the columns run as rows on the complex engine, and the result
  // is transposed back
  const bool col_fused = len_m > 1 && fusable(len_m, p->f64);
  const bool col_long = len_m > 1 && !col_fused;
  if (col_fused && (rc = build_inner(len_m, p->f64, p->inner))) {
    delete p;
    return rc;
  }
  if (col_long) {
    if ((rc = p->cols.init(len_m, p->f64)) ||
        (rc = dev_alloc(p->work, batch * len_m * len_n * esz_of(p->f64)))) {
      delete p;
      return rc;
    }
    p->cols.force_split = true;
    p->cols.force_persistent = false;
  }
  for (int dir = 0; dir < 2; dir++) {
    std::vector<Launch>& L = dir ? p->bwd : p->fwd;
    // rows: batch * m transforms of length n (cfft2d.rs:340-348); with a
    // column pass to follow they leave the scaling to it and, in an
    // inverse, conjugate on load only (the column pass conjugates on store)
    p->rows.emit(L, batch * len_m, dir == 1, S_IN, S_OUT, (long long)len_n, (long long)len_n);
    if (len_m > 1)
      for (auto& l : L) {
        l.scaled = 0;
        if (dir) l.P.conj_out = 0;
      }
    // an inverse row pass leaves its output conjugated; a length-1 row (plain
    // copy) does not, so the column pass then conjugates on load
    const int cin = p->rows.mode == 0 ? dir : 0;
    if (col_fused) {
      // columns: batch * n transforms of length m, stride n, in place on out
      Launch c{};
      c.ld = LD_PLAIN;
      c.st = ST_PLAIN;
      fill_inner(c.P, p->inner);
      c.P.G = (int)(batch * len_n);
      c.P.in = io((long long)(len_m * len_n), 1, (long long)len_n, (int)len_n);
      c.P.out = c.P.in;
      c.P.colmajor_ld = c.P.colmajor_st = 1;
      c.P.conj_out = dir;
      c.P.conj_in = cin;
      Geometry g = choose_geometry((int)p->inner.Lfft, c.P.G, true, p->f64);
      C2CEngine::finish(c, g, S_OUT, S_OUT, 1);
      c.what = "2-D columns " + seq_str(p->inner.seq);
      try_specialise(c, p->f64);
      L.push_back(c);
    } else if (col_long) {
      L.push_back(transpose_launch(S_OUT, S_WORK, len_m, len_n, batch));
      const size_t first = L.size();
      p->cols.emit(L, batch * len_n, dir == 1, S_WORK, S_WORK, (long long)len_m, (long long)len_m);
      if (dir && !cin)
        for (size_t i = first; i < L.size(); i++)
          if (L[i].src == S_WORK) L[i].P.conj_in = 0;
      for (size_t i = first; i < L.size(); i++) L[i].what = "2-D columns (transposed) " + L[i].what;
      L.push_back(transpose_launch(S_WORK, S_OUT, len_n, len_m, batch));
    }
  }
  return finish_plan(p, plan);
}

int chfft_plan_r2c_1d(chfft_plan* plan, size_t len, size_t batch, int precision) {
  if (!plan) return CHFFT_ERR_INVALID_ARGUMENT;
  *plan = nullptr;
  if (check_prec(precision)) return CHFFT_ERR_INVALID_ARGUMENT;
  if (len == 0 || (len & 1)) return CHFFT_ERR_INVALID_LENGTH;  // rfft1d.rs:74-76
  if (batch == 0 || batch * len > (size_t)1 << 31) return CHFFT_ERR_INVALID_ARGUMENT;
  int rc = ensure_runtime();
  if (rc) return rc;
  const bool f64 = precision == CHFFT_F64;
  const size_t h = len / 2;
  const bool longp = !fusable(h, f64);
  chfft_plan_s* p = new chfft_plan_s();
  p->kind = K_R2C;
  p->f64 = f64;
  p->n = len;
  p->batch = batch;
  if ((rc = longp ? long_inner_init(p, h, batch) : build_inner(h, f64, p->inner))) {
    delete p;
    return rc;
  }
  // coef_i = (1 + i W_N^i) / 2, bcoef_i = (1 - i W_N^-i) / 2 (rfft1d.rs:44-65)
  std::vector<ld_t> co(2 * h), bc(2 * h);
  for (size_t i = 0; i < h; i++) {
    ld_t c, s;
    unit_root(i, len, &c, &s);  // W_N^i = c + i s
    co[2 * i] = (1.0L - s) / 2.0L;
    co[2 * i + 1] = c / 2.0L;
    // W_N^-i = c - i s ; i W^-i = s + i c
    bc[2 * i] = (1.0L - s) / 2.0L;
    bc[2 * i + 1] = -c / 2.0L;
  }
  if ((rc = upload(co, f64, p->tab_a)) || (rc = upload(bc, f64, p->tab_b))) {
    delete p;
    return rc;
  }
  if (longp) {
    // forward: the inner transform reads the len reals as h complex values
    // into the work buffer, the split runs as a post-pass
    emit_unscaled(p, p->fwd, batch, false, S_IN, S_WORK, (long long)h);
    Launch f = op_pass(6, ST_R2C, h, (long long)h, (long long)(h + 1), S_WORK, S_OUT, batch);
    f.P.N = (long long)len;
    f.P.tab_a = p->tab_a.p;
    p->fwd.push_back(f);
    // backward: the un-split as a pre-pass, the inverse inner transform
    // writes the reals and applies the scale (times 2, rfft1d.rs:161)
    Launch b = op_pass(5, LD_R2C_BACK, h, (long long)(h + 1), (long long)h, S_IN, S_WORK, batch);
    b.P.N = (long long)len;
    b.P.tab_a = p->tab_b.p;
    p->bwd.push_back(b);
    const size_t first = p->bwd.size();
    p->rows.emit(p->bwd, batch, true, S_WORK, S_OUT, (long long)h, (long long)h);
    for (size_t i = first; i < p->bwd.size(); i++) p->bwd[i].scale_mul = 2.0;
    return finish_plan(p, plan);
  }
  Launch f = fused_op(p, LD_PLAIN, ST_R2C, false, (long long)h, (long long)(h + 1), (int)batch);
  f.P.N = (long long)len;
  f.P.tab_a = p->tab_a.p;
  p->fwd.push_back(f);
  Launch b = fused_op(p, LD_R2C_BACK, ST_PLAIN, true, (long long)(h + 1), (long long)h, (int)batch);
  b.P.N = (long long)len;
  b.P.tab_a = p->tab_b.p;
  b.scale_mul = 2.0;  // rfft1d.rs:161
  p->bwd.push_back(b);
  return finish_plan(p, plan);
}

int chfft_plan_dct_1d(chfft_plan* plan, int dct_type, size_t len, int is_ortho, size_t batch, int precision) {
  if (!plan) return CHFFT_ERR_INVALID_ARGUMENT;
  *plan = nullptr;
  if (check_prec(precision)) return CHFFT_ERR_INVALID_ARGUMENT;
  if (dct_type != CHFFT_DCT_II && dct_type != CHFFT_DCT_III) return CHFFT_ERR_INVALID_ARGUMENT;
  if (len == 0 || (len & 1)) return CHFFT_ERR_INVALID_LENGTH;  // dct1d.rs:54-56
  if (batch == 0 || batch * len > (size_t)1 << 31) return CHFFT_ERR_INVALID_ARGUMENT;
  int rc = ensure_runtime();
  if (rc) return rc;
  const bool f64 = precision == CHFFT_F64;
  const size_t h = len / 2;
  const bool longp = !fusable(h, f64);
  chfft_plan_s* p = new chfft_plan_s();
  p->kind = K_DCT;
  p->f64 = f64;
  p->n = len;
  p->batch = batch;
  p->dct_type = dct_type;
  p->ortho = is_ortho != 0;
  if ((rc = longp ? long_inner_init(p, h, batch) : build_inner(h, f64, p->inner))) {
    delete p;
    return rc;
  }
  std::vector<ld_t> co(2 * h), om(2 * (h + 1));
  for (size_t i = 0; i < h; i++) {
    ld_t c, s;
    unit_root(i, len, &c, &s);
    if (dct_type == CHFFT_DCT_II) {  // (1 + i W_N^i) / 2, dct1d.rs:208-215
      co[2 * i] = (1.0L - s) / 2.0L;
      co[2 * i + 1] = c / 2.0L;
    } else {  // (1 - i W_N^-i) / 2, dct1d.rs:330-338
      co[2 * i] = (1.0L - s) / 2.0L;
      co[2 * i + 1] = -c / 2.0L;
    }
  }
  for (size_t k = 0; k <= h; k++) {
    ld_t c, s;
    unit_root(k, 4 * len, &c, &s);  // exp(-i pi k / 2N)
    om[2 * k] = c;
    om[2 * k + 1] = dct_type == CHFFT_DCT_II ? -s : s;  // DCT-II uses exp(+i pi k / 2N)
  }
  if ((rc = upload(co, f64, p->tab_a)) || (rc = upload(om, f64, p->tab_b))) {
    delete p;
    return rc;
  }
  const bool two = dct_type == CHFFT_DCT_II;
  // dct1d.rs:51-55 and 174-178
  const double ortho = two ? (p->ortho ? std::sqrt(0.5) : 1.0) : (p->ortho ? std::sqrt(0.5) : 0.5);
  if (longp) {
    // pack / pre-twiddle pass, inner transform on the work buffer (inverse
    // for DCT-III), post-twiddle / unpack pass
    Launch a = op_pass(5, two ? LD_DCT2 : LD_DCT3, h, (long long)len, (long long)h, S_IN, S_WORK, batch);
    Launch c = op_pass(6, two ? ST_DCT2 : ST_DCT3, h, (long long)h, (long long)len, S_WORK, S_OUT, batch);
    for (Launch* l : {&a, &c}) {
      l->P.N = (long long)len;
      l->P.tab_a = p->tab_a.p;
      l->P.tab_b = p->tab_b.p;
      l->P.ortho = ortho;
    }
    p->fwd.push_back(a);
    emit_unscaled(p, p->fwd, batch, !two, S_WORK, S_WORK, (long long)h);
    p->fwd.push_back(c);
    return finish_plan(p, plan);
  }
  Launch f = fused_op(p, two ? LD_DCT2 : LD_DCT3, two ? ST_DCT2 : ST_DCT3, !two, (long long)len, (long long)len,
                      (int)batch);
  f.P.N = (long long)len;
  f.P.tab_a = p->tab_a.p;
  f.P.tab_b = p->tab_b.p;
  f.P.ortho = ortho;
  p->fwd.push_back(f);
  return finish_plan(p, plan);
}

int chfft_plan_mdct_1d(chfft_plan* plan, size_t len, const double* window, size_t batch, size_t frame_stride,
                       int precision) {
  if (!plan) return CHFFT_ERR_INVALID_ARGUMENT;
  *plan = nullptr;
  if (check_prec(precision)) return CHFFT_ERR_INVALID_ARGUMENT;
  if (len == 0 || (len & 3)) return CHFFT_ERR_INVALID_LENGTH;  // mdct1d.rs:102-104
  if (!window) return CHFFT_ERR_INVALID_ARGUMENT;
  if (batch == 0 || batch * len > (size_t)1 << 31) return CHFFT_ERR_INVALID_ARGUMENT;
  int rc = ensure_runtime();
  if (rc) return rc;
  const bool f64 = precision == CHFFT_F64;
  const size_t q = len / 4;
  const bool longp = !fusable(q, f64);
  chfft_plan_s* p = new chfft_plan_s();
  p->kind = K_MDCT;
  p->f64 = f64;
  p->n = len;
  p->batch = batch;
  p->frame_stride = frame_stride ? frame_stride : len;
  if ((rc = longp ? long_inner_init(p, q, batch) : build_inner(q, f64, p->inner))) {
    delete p;
    return rc;
  }
  // frames of an overlapped (or spaced) stream are overlap-added by
  // chfft_exec_mdct_ola from this buffer
  if (p->frame_stride != len && (rc = dev_alloc(p->ola, batch * len * (f64 ? 8 : 4)))) {
    delete p;
    return rc;
  }
  // tw_i = exp(-2 pi i (8i + 1) / 8N) * (2/N)^(1/4) (mdct1d.rs:78-89)
  std::vector<ld_t> tw(2 * q), w(len);
  const ld_t sc = sqrtl(sqrtl(2.0L / (ld_t)len));
  for (size_t i = 0; i < q; i++) {
    ld_t c, s;
    unit_root(8 * i + 1, 8 * len, &c, &s);
    tw[2 * i] = c * sc;
    tw[2 * i + 1] = s * sc;
  }
  for (size_t i = 0; i < len; i++) w[i] = (ld_t)window[i];
  if ((rc = upload(tw, f64, p->tab_a)) || (rc = upload(w, f64, p->tab_b))) {
    delete p;
    return rc;
  }
  if (longp) {
    // fold + pre-twiddle, inner forward transform, post-twiddle (both
    // directions: the IMDCT also runs the forward transform, mdct1d.rs:266)
    Launch a = op_pass(5, LD_MDCT, q, (long long)p->frame_stride, (long long)q, S_IN, S_WORK, batch);
    Launch c = op_pass(6, ST_MDCT, q, (long long)q, (long long)(len / 2), S_WORK, S_OUT, batch);
    Launch d = op_pass(5, LD_IMDCT, q, (long long)(len / 2), (long long)q, S_IN, S_WORK, batch);
    Launch e = op_pass(6, ST_IMDCT, q, (long long)q, (long long)len, S_WORK, S_OUT, batch);
    for (Launch* l : {&a, &c, &d, &e}) {
      l->P.N = (long long)len;
      l->P.tab_a = p->tab_a.p;
      l->P.tab_b = p->tab_b.p;
    }
    p->fwd.push_back(a);
    emit_unscaled(p, p->fwd, batch, false, S_WORK, S_WORK, (long long)q);
    p->fwd.push_back(c);
    p->bwd.push_back(d);
    emit_unscaled(p, p->bwd, batch, false, S_WORK, S_WORK, (long long)q);
    p->bwd.push_back(e);
    return finish_plan(p, plan);
  }
  Launch f = fused_op(p, LD_MDCT, ST_MDCT, false, (long long)p->frame_stride, (long long)(len / 2), (int)batch);
  f.P.N = (long long)len;
  f.P.tab_a = p->tab_a.p;
  f.P.tab_b = p->tab_b.p;
  p->fwd.push_back(f);
  Launch b = fused_op(p, LD_IMDCT, ST_IMDCT, false, (long long)(len / 2), (long long)len, (int)batch);
  b.P.N = (long long)len;
  b.P.tab_a = p->tab_a.p;
  b.P.tab_b = p->tab_b.p;
  p->bwd.push_back(b);
  return finish_plan(p, plan);
}

int chfft_plan_destroy(chfft_plan plan) {
  delete plan;
  return CHFFT_OK;
}

}  // extern "C"

namespace chfft {
namespace {

// Method -> (inverse?, scale), per transform type (see include/chfft_hip.h).
int method_scale(const chfft_plan_s* p, int method, bool* inverse, double* scale) {
  const double N = p->kind == K_C2C2D ? (double)p->m * (double)p->n : (double)p->n;
  switch (p->kind) {
    case K_MULTI:  // scale kind (per transform length): 0 -> 1, 1 -> 1/sqrt(n), 2 -> 1/n
      switch (method) {
        case CHFFT_FORWARD: case CHFFT_FORWARD0: *inverse = false; *scale = 0; return CHFFT_OK;
        case CHFFT_FORWARDU: *inverse = false; *scale = 1; return CHFFT_OK;
        case CHFFT_FORWARDN: *inverse = false; *scale = 2; return CHFFT_OK;
        case CHFFT_BACKWARD: case CHFFT_BACKWARDN: *inverse = true; *scale = 2; return CHFFT_OK;
        case CHFFT_BACKWARD0: *inverse = true; *scale = 0; return CHFFT_OK;
        case CHFFT_BACKWARDU: *inverse = true; *scale = 1; return CHFFT_OK;
      }
      return CHFFT_ERR_INVALID_ARGUMENT;
    case K_C2C:
    case K_R2C:
    case K_C2C2D:
      switch (method) {
        case CHFFT_FORWARD: case CHFFT_FORWARD0: *inverse = false; *scale = 1.0; return CHFFT_OK;
        case CHFFT_FORWARDU: *inverse = false; *scale = 1.0 / std::sqrt(N); return CHFFT_OK;
        case CHFFT_FORWARDN: *inverse = false; *scale = 1.0 / N; return CHFFT_OK;
        case CHFFT_BACKWARD: *inverse = true; *scale = 1.0 / N; return CHFFT_OK;
        case CHFFT_BACKWARD0: *inverse = true; *scale = 1.0; return CHFFT_OK;
        case CHFFT_BACKWARDU: *inverse = true; *scale = 1.0 / std::sqrt(N); return CHFFT_OK;
        case CHFFT_BACKWARDN:
          if (p->kind == K_C2C2D) return CHFFT_ERR_UNSUPPORTED;  // CFft2D has no backwardn
          *inverse = true;
          *scale = 1.0 / N;
          return CHFFT_OK;
      }
      return CHFFT_ERR_INVALID_ARGUMENT;
    case K_DCT: {
      const double u = std::sqrt(2.0 / N);  // dct1d.rs:218
      *inverse = false;
      switch (method) {
        case CHFFT_FORWARD: case CHFFT_FORWARD0: *scale = 1.0; return CHFFT_OK;
        case CHFFT_FORWARDU: *scale = u; return CHFFT_OK;
        case CHFFT_FORWARDN: *scale = u * u; return CHFFT_OK;
      }
      return method >= 0 && method <= 7 ? CHFFT_ERR_UNSUPPORTED : CHFFT_ERR_INVALID_ARGUMENT;
    }
    case K_MDCT: {
      const double h = (double)(p->n >> 1);
      switch (method) {
        case CHFFT_FORWARD: *inverse = false; *scale = std::sqrt(h); return CHFFT_OK;
        case CHFFT_FORWARDU: *inverse = false; *scale = 1.0; return CHFFT_OK;
        case CHFFT_BACKWARD: *inverse = true; *scale = 1.0 / std::sqrt(h); return CHFFT_OK;
        case CHFFT_BACKWARDU: *inverse = true; *scale = 1.0; return CHFFT_OK;
      }
      return method >= 0 && method <= 7 ? CHFFT_ERR_UNSUPPORTED : CHFFT_ERR_INVALID_ARGUMENT;
    }
  }
  return CHFFT_ERR_INVALID_ARGUMENT;
}

void io_bytes(const chfft_plan_s* p, bool inverse, size_t* ib, size_t* ob) {
  const size_t e = p->f64 ? 8 : 4;  // bytes per real
  const size_t B = p->batch, n = p->n;
  switch (p->kind) {
    case K_MULTI: *ib = *ob = 2 * e * p->total; return;
    case K_C2C: *ib = *ob = 2 * e * n * B; return;
    case K_C2C2D: *ib = *ob = 2 * e * n * p->m * B; return;
    case K_R2C:
      if (!inverse) {
        *ib = e * n * B;
        *ob = 2 * e * (n / 2 + 1) * B;
      } else {
        *ib = 2 * e * (n / 2 + 1) * B;
        *ob = e * n * B;
      }
      return;
    case K_DCT: *ib = *ob = e * n * B; return;
    case K_MDCT:
      if (!inverse) {
        *ib = e * ((B - 1) * p->frame_stride + n);
        *ob = e * (n / 2) * B;
      } else {
        *ib = e * (n / 2) * B;
        *ob = e * n * B;
      }
      return;
  }
}

}  // namespace
}  // namespace chfft

extern "C" {

}  // extern "C"

namespace chfft {
namespace {

// Runs launches [first, first + count) of `method`'s schedule with the given
// scratch and work buffers (the plan's own, or a caller's).
int run_launches(chfft_plan plan, const void* in, void* out, void* scratch, void* work, int method, int first,
                 int count, hipStream_t st) {
  if (!plan || !in || !out) return CHFFT_ERR_INVALID_ARGUMENT;
  bool inverse;
  double scale;
  int rc = method_scale(plan, method, &inverse, &scale);
  if (rc) return rc;
  if (plan->kind == K_MULTI) {
    if (first != 0 || count != 1) return count == 0 ? CHFFT_OK : CHFFT_ERR_INVALID_ARGUMENT;
    if (plan->nitems == 0) return CHFFT_OK;
    const KParams* descs = reinterpret_cast<const KParams*>(inverse ? plan->descs_bwd.p : plan->descs_fwd.p);
    // a cooperative launch guarantees that every workgroup is resident for
    // the grid-wide barriers; a stream under capture takes the ordinary
    // launch (the barrier waits stay bounded and chfft_plan_check reports a
    // timed-out one)
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    const bool coop = hipStreamIsCapturing(st, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone;
    (void)hipGetLastError();
    auto go = [&](bool c) {
      return launch_multi(plan->f64, reinterpret_cast<const MultiItem*>(plan->items_dev.p),
                          reinterpret_cast<const MultiUnit*>(plan->units_dev.p), plan->unit_ends, descs, in, out,
                          (int)scale, scratch, reinterpret_cast<unsigned*>(plan->sync_dev.p), plan->barrier_ticks,
                          plan->barrier_extra, c, dim3((unsigned)plan->multi_grid), plan->multi_lds, st);
    };
    hipError_t e = go(coop);
    if (e != hipSuccess && coop) {
      (void)hipGetLastError();
      e = go(false);
    }
    return e == hipSuccess ? CHFFT_OK : CHFFT_ERR_HIP;
  }
  const std::vector<Launch>& L = inverse ? plan->bwd : plan->fwd;
  if (first < 0 || count < 0 || (size_t)first + (size_t)count > L.size()) return CHFFT_ERR_INVALID_ARGUMENT;
  for (int li = first; li < first + count; li++) {
    const Launch& l0 = L[li];
    const void* slots_src[kSlots] = {in, out, scratch, work};
    void* slots_dst[kSlots] = {const_cast<void*>(in), out, scratch, work};
    if (l0.src_off) slots_src[l0.src] = static_cast<const char*>(slots_src[l0.src]) + l0.src_off;
    if (l0.dst_off) slots_dst[l0.dst] = static_cast<char*>(slots_dst[l0.dst]) + l0.dst_off;
    if (l0.kind == 2) {
      if (launch_fourstep(plan->f64, l0.nph, reinterpret_cast<const KParams*>(l0.qdev), slots_src[l0.src],
                          slots_dst[l0.dst], scale * l0.scale_mul, l0.batch, l0.src_dist, l0.dst_dist, scratch,
                          l0.slot_elems, l0.grid, l0.block, l0.lds, st) != hipSuccess)
        return CHFFT_ERR_HIP;
      continue;
    }
    if (l0.kind == 4) {
      if (sfs_launch(l0.sfs, reinterpret_cast<const KParams*>(l0.qdev), slots_src[l0.src], slots_dst[l0.dst],
                     scale * l0.scale_mul, l0.batch, l0.src_dist, l0.dst_dist, scratch, l0.slot_elems, l0.grid,
                     st) != hipSuccess)
        return CHFFT_ERR_HIP;
      continue;
    }
    if (l0.kind == 1) {  // length-1 copy
      size_t ib, ob;
      io_bytes(plan, inverse, &ib, &ob);
      if (slots_src[l0.src] != slots_dst[l0.dst] &&
          launch_copy(slots_src[l0.src], slots_dst[l0.dst], ib, st) != hipSuccess)
        return CHFFT_ERR_HIP;
      continue;
    }
    if (l0.kind == 7) {
      if (launch_transpose(plan->f64, slots_src[l0.src], slots_dst[l0.dst], l0.P.Lfft, l0.P.L, l0.P.G, st) !=
          hipSuccess)
        return CHFFT_ERR_HIP;
      continue;
    }
    KParams P = l0.P;
    P.src = slots_src[l0.src];
    P.dst = slots_dst[l0.dst];
    P.scale = l0.scaled ? scale * l0.scale_mul : 1.0;
    hipError_t e;
    if (l0.kind == 3)
      e = sfft_launch(l0.sf, P, l0.grid, l0.lds, st);
    else if (l0.kind == 5)
      e = launch_pre(plan->f64, l0.ld, P, st);
    else if (l0.kind == 6)
      e = launch_post(plan->f64, l0.st, P, st);
    else
      e = launch_fused(plan->f64, l0.ld, l0.st, P, l0.grid, l0.block, l0.lds, st);
    if (e != hipSuccess) return CHFFT_ERR_HIP;
  }
  return CHFFT_OK;
}

// Reads and clears the multi-length launch's sticky barrier-failure flag
// (stream-ordered after `st`, blocking).
int read_barrier_flag(chfft_plan plan, hipStream_t st) {
  if (hipStreamSynchronize(st) != hipSuccess) return CHFFT_ERR_HIP;
  if (plan->kind != K_MULTI || !plan->sync_dev.p) return CHFFT_OK;
  unsigned flag = 0;
  unsigned* f = reinterpret_cast<unsigned*>(plan->sync_dev.p) + 1;
  if (hipMemcpy(&flag, f, sizeof flag, hipMemcpyDeviceToHost) != hipSuccess) return CHFFT_ERR_HIP;
  if (!flag) return CHFFT_OK;
  if (hipMemset(f, 0, sizeof flag) != hipSuccess) return CHFFT_ERR_HIP;
  return CHFFT_ERR_INCOMPLETE;
}

}  // namespace
}  // namespace chfft

extern "C" {

int chfft_exec_range(chfft_plan plan, const void* in, void* out, int method, int first, int count, void* stream) {
  if (!plan) return CHFFT_ERR_INVALID_ARGUMENT;
  return run_launches(plan, in, out, plan->scratch.p, plan->work.p, method, first, count,
                      reinterpret_cast<hipStream_t>(stream));
}

int chfft_plan_check(chfft_plan plan, void* stream) {
  if (!plan) return CHFFT_ERR_INVALID_ARGUMENT;
  return read_barrier_flag(plan, reinterpret_cast<hipStream_t>(stream));
}

int chfft_exec_mdct_ola(chfft_plan plan, const void* in, void* out, int method, void* stream) {
  if (!plan || !in || !out || plan->kind != K_MDCT) return CHFFT_ERR_INVALID_ARGUMENT;
  bool inverse;
  double scale;
  int rc = method_scale(plan, method, &inverse, &scale);
  if (rc) return rc;
  if (!inverse) return CHFFT_ERR_INVALID_ARGUMENT;  // only the backward methods give frames to overlap-add
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const int n = (int)plan->bwd.size();
  if (plan->frame_stride == plan->n)  // frames back to back: the frames are the signal
    return run_launches(plan, in, out, plan->scratch.p, plan->work.p, method, 0, n, st);
  if ((rc = run_launches(plan, in, plan->ola.p, plan->scratch.p, plan->work.p, method, 0, n, st))) return rc;
  return launch_ola(plan->f64, plan->ola.p, out, (long long)plan->n, (long long)plan->frame_stride,
                    (long long)plan->batch, st) == hipSuccess
             ? CHFFT_OK
             : CHFFT_ERR_HIP;
}

int chfft_exec(chfft_plan plan, const void* in, void* out, int method, void* stream) {
  int n = chfft_plan_launch_count(plan, method);
  if (n < 0) return -n;
  return chfft_exec_range(plan, in, out, method, 0, n, stream);
}

// chfft_exec with the plan's scratch and work buffers replaced by
// noalloc_scratch: the scratch at its start, the work buffer at the next
// 256-byte boundary. Every kernel writes each scratch and work element
// before reading it within one call, so nothing needs clearing; calls that
// overlap on one GPU share the array and race.
int chfft_exec_noalloc(chfft_plan plan, const void* in, void* out, int method, void* stream) {
  if (!plan || !in || !out) return CHFFT_ERR_INVALID_ARGUMENT;
  const size_t sb = (plan->scratch.bytes + 255) / 256 * 256;
  if (sb + plan->work.bytes > (size_t)CHFFT_NOALLOC_SCRATCH_BYTES) return CHFFT_ERR_SCRATCH_LIMIT;
  int n = chfft_plan_launch_count(plan, method);
  if (n < 0) return -n;
  void* base = nullptr;
  int rc = noalloc_base(&base);
  if (rc) return rc;
  return run_launches(plan, in, out, base, static_cast<char*>(base) + sb, method, 0, n,
                      reinterpret_cast<hipStream_t>(stream));
}

int chfft_plan_launch_count(chfft_plan plan, int method) {
  if (!plan) return -CHFFT_ERR_INVALID_ARGUMENT;
  bool inverse;
  double scale;
  int rc = method_scale(plan, method, &inverse, &scale);
  if (rc) return -rc;
  if (plan->kind == K_MULTI) return 1;
  return (int)(inverse ? plan->bwd.size() : plan->fwd.size());
}

int chfft_plan_io_bytes(chfft_plan plan, int method, size_t* in_bytes, size_t* out_bytes) {
  if (!plan || !in_bytes || !out_bytes) return CHFFT_ERR_INVALID_ARGUMENT;
  bool inverse;
  double scale;
  int rc = method_scale(plan, method, &inverse, &scale);
  if (rc) return rc;
  io_bytes(plan, inverse, in_bytes, out_bytes);
  return CHFFT_OK;
}

size_t chfft_plan_scratch_bytes(chfft_plan plan) { return plan ? plan->scratch.bytes : 0; }

int chfft_exec_host(chfft_plan plan, const void* in, void* out, int method) {
  if (!plan || !in || !out) return CHFFT_ERR_INVALID_ARGUMENT;
  size_t ib, ob;
  int rc = chfft_plan_io_bytes(plan, method, &ib, &ob);
  if (rc) return rc;
  if ((rc = dev_alloc(plan->stage_in, ib)) || (rc = dev_alloc(plan->stage_out, ob))) return rc;
  if (hipMemcpy(plan->stage_in.p, in, ib, hipMemcpyHostToDevice) != hipSuccess) return CHFFT_ERR_HIP;
  if ((rc = chfft_exec(plan, plan->stage_in.p, plan->stage_out.p, method, nullptr))) return rc;
  if (hipMemcpy(out, plan->stage_out.p, ob, hipMemcpyDeviceToHost) != hipSuccess) return CHFFT_ERR_HIP;
  // a grid-wide barrier of a multi-length launch that timed out leaves its flag behind
  return read_barrier_flag(plan, nullptr);
}

int chfft_plan_describe(chfft_plan plan, char* buf, size_t buf_len) {
  if (!plan) return -1;
  std::string s;
  if (plan->kind == K_MULTI) s = plan->multi_desc;
  for (int dir = 0; dir < 2; dir++) {
    const std::vector<Launch>& L = dir ? plan->bwd : plan->fwd;
    if (L.empty()) continue;
    s += dir ? "inverse:\n" : "forward:\n";
    for (const Launch& l : L) {
      char line[512];
      // ops = load / store operator of the fused kernels (their template
      // arguments, which tie a profiler's per-dispatch rows to the launch)
      snprintf(line, sizeof line, "  %s | G=%d tpb=%d block=%u grid=%u lds=%zu ops=%d/%d\n", l.what.c_str(), l.P.G,
               l.P.tpb, l.block.x, l.grid.x, l.lds, l.ld, l.st);
      s += line;
    }
  }
  // the planner overrides the plan runs with, and the last autotune's report
  const std::string spec = tune_str(plan->tune);
  if (spec != "default" || !plan->tune_report.empty()) {
    s += "tuning: " + spec + "\n";
    if (!plan->tune_report.empty()) s += plan->tune_report + "\n";
  }
  if (buf && buf_len) {
    size_t k = std::min(buf_len - 1, s.size());
    memcpy(buf, s.data(), k);
    buf[k] = 0;
  }
  return plan->kind == K_MULTI ? 1 : (int)plan->fwd.size();
}

}  // extern "C"

// ---------------------------------------------------------------- multi-length batch
namespace chfft {
namespace {

struct SeqTab {  // a pass sequence whose table lives in the arena
  SeqDesc seq{};
  long long off = 0;  // arena offset of its table
  size_t count = 0;
  bool compact = false;  // also the compact tables of the length-specialised tiles
};

struct LenInfo {
  int kind = 0;  // 0 fused, 1 four-step, 2 four-step Bluestein, 3 copy
  size_t n = 0, Lfft = 0, M = 0, n1 = 0, n2 = 0;
  bool blue = false;
  int seq = -1, s1 = -1, s2 = -1;  // indices into the SeqTab list
  long long chirp = -1, kspec = -1, tw_hi = -1, tw_lo = -1;
  int twsh = 0;
  // register-resident tile paths of the phases (0: generic), their tile
  // widths, SeqTab indices of their schedules and LDS bytes
  int spec[3] = {0, 0, 0}, spec_tpb[3] = {0, 0, 0}, spec_seq[3] = {-1, -1, -1};
  size_t spec_lds[3] = {0, 0, 0};
  long long rec64 = -1;  // FP64 two-level tables in arena64 (register-resident four-step stores)
  int desc = 0;  // first descriptor of this length (per direction)
  int ndesc = 0;
};

int pow2_ceil_log(size_t n) {
  int b = 0;
  while (((size_t)1 << b) < n) b++;
  return b;
}

// tpb for a phase of length L under a fixed workgroup of `block` threads:
// the most transforms (up to maxtpb, 64 unless tuned) for which tpt >= L/16
// and the tile fits in
// LDS. Column phases want at least 16 (128-byte rows); beyond that, a wider
// tile spreads each pass's barrier and twiddle-load latency over more points
// (the multi-length kernel runs one workgroup per CU, with nothing else to
// hide that latency: 16 -> 64 took the 10,000-length batch's smooth
// four-step items from 2.90 to 2.11 ms on MI355X).
int phase_tpb(size_t L, int block, bool colmajor, bool f64, int maxtpb) {
  (void)colmajor;
  int tpb = maxtpb;
  const size_t esz = f64 ? 16 : 8;
  while (tpb > 1) {
    const int tpt = block / tpb;
    size_t Lp = L + (L + 15) / 16;
    if (tpb > 1 && (Lp & 1) == 0) Lp++;
    // (the multi-length kernel also keeps a few bytes of static LDS)
    if ((size_t)tpt * kEmax >= L && (size_t)tpb * Lp * esz + 64 <= kLdsBytes) break;
    tpb >>= 1;
  }
  return tpb;
}

int lp_for(size_t L, int tpb) {
  size_t Lp = L + (L + 15) / 16;
  if (tpb > 1 && (Lp & 1) == 0) Lp++;
  return (int)Lp;
}

}  // namespace
}  // namespace chfft

// The multi-length plan under the planner overrides `t`: `generic` keeps
// every four-step phase on the runtime-radix tiles, `grid` sets the
// persistent launch's workgroup count (at most every workgroup that fits at
// once), `unit` the points per work unit of a four-step phase, `tpb` the
// most transforms of a generic four-step tile (a power of two up to 128).
static int multi_create(const size_t* lens, size_t count, int precision, const Tune& t, chfft_plan* plan) {
  if (!plan) return CHFFT_ERR_INVALID_ARGUMENT;
  *plan = nullptr;
  if ((t.tpb & (t.tpb - 1)) || t.tpb < 0 || t.tpb > 128 || t.tpt || t.order || t.n1 || t.chunk || t.sched ||
      (t.unit && t.unit < 1024))
    return CHFFT_ERR_INVALID_ARGUMENT;
  const int maxtpb = t.tpb ? t.tpb : 64;
  if (check_prec(precision)) return CHFFT_ERR_INVALID_ARGUMENT;
  if (count > 0 && !lens) return CHFFT_ERR_INVALID_ARGUMENT;
  size_t total = 0;
  for (size_t i = 0; i < count; i++) {
    if (lens[i] == 0) return CHFFT_ERR_INVALID_LENGTH;
    total += lens[i];
  }
  if (total > ((size_t)1 << 31)) return CHFFT_ERR_INVALID_ARGUMENT;
  int rc = ensure_runtime();
  if (rc) return rc;
  const bool f64 = precision == CHFFT_F64;
  const size_t esz = esz_of(f64);
  const int block = f64 ? kMultiThreadsF64 : kMultiThreadsF32;
  const size_t capM = (size_t)kEmax * block;  // one workgroup, tpb = 1
  std::unique_ptr<chfft_plan_s> p(new chfft_plan_s());
  p->kind = K_MULTI;
  p->f64 = f64;
  p->lens.assign(lens, lens + count);
  p->total = total;
  p->batch = 1;
  p->tune = t;

  // ---- distinct lengths and their schedules
  std::vector<SeqTab> seqs;
  std::map<std::vector<int>, int> seq_of_sched;  // key: {L, compact?, radices...}
  long long arena = 0;
  // pass sequence of length L with radix schedule `rad` (default: the
  // runtime-radix schedule with the prime-factor radices 10, 12 and 15, which
  // the multi-length kernel's generic passes take: fewer passes for lengths
  // with factors 3 and 5); `compact`: the length-specialised tiles' table
  // layout, the compact tables (two-level where large) after the full ones
  auto seq_for = [&](size_t L, std::vector<int> rad = {}, bool compact = false) {
    if (rad.empty()) rad = radices_for(L, true);
    std::vector<int> key = {(int)L, compact ? 1 : 0};
    key.insert(key.end(), rad.begin(), rad.end());
    auto it = seq_of_sched.find(key);
    if (it != seq_of_sched.end()) return it->second;
    SeqTab t;
    t.seq.L = (int)L;
    t.seq.npass = (int)rad.size();
    t.compact = compact;
    int pp = 1;
    size_t cnt = 0;
    for (size_t q = 0; q < rad.size(); q++) {
      PassDesc& pd = t.seq.pass[q];
      pd.radix = rad[q];
      pd.p = pp;
      pd.tws = (int)(L / ((size_t)pp * rad[q]));
      pd.pmagic = (unsigned)((((unsigned long long)1 << 32) + (unsigned long long)pp - 1) / (unsigned long long)pp);
      pd.twoff = (int)cnt;
      if (pp > 1) cnt += (size_t)(rad[q] - 1) * pp;
      pp *= rad[q];
    }
    if (compact) {
      pp = 1;
      for (size_t q = 0; q < rad.size(); q++) {
        cnt += (size_t)tw_compact_entries(rad[q], pp);
        pp *= rad[q];
      }
    }
    t.count = cnt ? cnt : 1;
    t.off = arena;
    arena += (long long)t.count;
    seqs.push_back(t);
    seq_of_sched[key] = (int)seqs.size() - 1;
    return (int)seqs.size() - 1;
  };
  std::map<size_t, LenInfo> info;
  std::vector<RootJob> jobs;
  for (size_t i = 0; i < count; i++) {
    const size_t n = lens[i];
    if (info.count(n)) continue;
    LenInfo L;
    L.n = n;
    if (n == 1) {
      L.kind = 3;
    } else {
      const bool sm = smooth7(n);
      // Bluestein padding: the smallest M >= 2N - 1 of the form 2^a times
      // 1, 3, 5 or 15 (at most 25 % above the bound, and few distinct M per
      // batch, each of which needs its own spectrum plan) that one workgroup
      // holds or that splits into two such factors; every such M gives the
      // reference's linear convolution (cfft1d.rs:114 pads to 2 * next_pow2(N))
      size_t M = sm ? n : 2 * next_pow2(n);
      if (!sm)
        for (size_t m = 2 * n - 1; m < M; m++) {
          size_t odd = m;
          while (odd % 2 == 0) odd /= 2;
          if (odd != 1 && odd != 3 && odd != 5 && odd != 15) continue;
          const size_t d = m <= capM ? 1 : fourstep_split(m, f64);
          if (d && m / d <= capM) {
            M = m;
            break;
          }
        }
      L.blue = !sm;
      L.M = M;
      if (M <= capM) {
        L.kind = 0;
        L.Lfft = M;
        L.seq = seq_for(M);
      } else {
        L.kind = sm ? 1 : 2;
        size_t best = 0;
        bool best_aligned = false;
        for (size_t d = 1; d * d <= M; d++) {
          if (M % d || !smooth7(d) || !smooth7(M / d)) continue;
          const bool al = (M / d) % 16 == 0;
          if (al || !best_aligned) {
            best = d;
            best_aligned = al;
          }
        }
        if (!best || M / best > capM) return CHFFT_ERR_UNSUPPORTED;
        if (sm && !f64 && !t.generic && getenv("CHFFT_MULTI_GENERIC") == nullptr) {
          // FP32 smooth items: a split whose column and row lengths both have
          // register-resident tiles where one exists (a row length that is a
          // multiple of 16 first, then the most balanced): the 10,000-length
          // uneven workload's smooth points on such tiles go from 48 % to 63 %
          size_t pick = 0;
          bool pick_al = false;
          double pick_gap = 0.0;
          for (size_t d = 2; d <= capM && d < M; d++) {
            if (M % d || M / d > capM) continue;
            if (!multi_spec_find((int)d, LD_PLAIN, ST_TW2, 0, true) ||
                !multi_spec_find((int)(M / d), LD_PLAIN, ST_PLAIN, 0, false))
              continue;
            const bool al = (M / d) % 16 == 0;
            const double gap = std::fabs(std::log((double)d) - std::log((double)(M / d)));
            if (!pick || (al && !pick_al) || (al == pick_al && gap < pick_gap)) {
              pick = d;
              pick_al = al;
              pick_gap = gap;
            }
          }
          if (pick) best = pick;
        }
        L.n1 = best;
        L.n2 = M / best;
        L.s1 = seq_for(L.n1);
        L.s2 = seq_for(L.n2);
        if (!f64 && !t.generic && getenv("CHFFT_MULTI_GENERIC") == nullptr) {
          // phases with a register-resident tile path: columns (+ chirp),
          // rows (Bluestein: spectrum product, inverse rows; smooth:
          // transposed store), inverse columns
          const bool bl = !sm;
          const int nph = bl ? 3 : 2;
          for (int ph = 0; ph < nph; ph++) {
            const MultiSpec* sp =
                ph == 0   ? multi_spec_find((int)L.n1, bl ? LD_FS_CHIRP : LD_PLAIN, ST_TW2, 0, true)
                : ph == 1 ? (bl ? multi_spec_find((int)L.n2, LD_PLAIN, ST_TW2, 2, false)
                                : multi_spec_find((int)L.n2, LD_PLAIN, ST_PLAIN, 0, false))
                          : multi_spec_find((int)L.n1, LD_PLAIN, ST_FS_DECHIRP, 0, true);
            if (!sp) continue;
            int id, tpb, nrad;
            const int* rad;
            size_t slds;
            multi_spec_info(sp, &id, &tpb, &nrad, &rad, &slds);
            L.spec[ph] = id;
            L.spec_tpb[ph] = tpb;
            L.spec_lds[ph] = slds;
            L.spec_seq[ph] = seq_for(ph == 1 ? L.n2 : L.n1, std::vector<int>(rad, rad + nrad), true);
          }
        }
        L.twsh = (pow2_ceil_log(M) + 1) / 2;
        const size_t nhi = (M + ((size_t)1 << L.twsh) - 1) >> L.twsh, nlo = (size_t)1 << L.twsh;
        L.tw_hi = arena;
        jobs.push_back({arena, (long long)nhi, (long long)M, 1, 2, L.twsh});
        arena += (long long)nhi;
        L.tw_lo = arena;
        jobs.push_back({arena, (long long)nlo, (long long)M, 1, 3, 0});
        arena += (long long)nlo;
      }
      if (L.blue) {
        L.chirp = arena;
        jobs.push_back({arena, (long long)n, (long long)(2 * n), 1, 1, 0});
        arena += (long long)n;
        L.kspec = arena;
        arena += (long long)M;
      }
    }
    info[n] = L;
  }
  auto table_jobs = [&](const SeqTab& t) {
    int pp = 1;
    long long c = t.off;
    for (int q = 0; q < t.seq.npass; q++) {
      const PassDesc& pd = t.seq.pass[q];
      if (pp > 1) {
        jobs.push_back({t.off + pd.twoff, (long long)(pd.radix - 1) * pp, (long long)pp * pd.radix, (long long)pp, 0,
                        0});
        c = t.off + pd.twoff + (long long)(pd.radix - 1) * pp;
      }
      pp *= pd.radix;
    }
    if (!t.compact) return;
    // compact tables: [j-1][k mod 16] = W_{pR}^(j k) and [j-1][k / 16] =
    // W_{pR/16}^(j k) where two-level, else the full [j-1][k]
    pp = 1;
    for (int q = 0; q < t.seq.npass; q++) {
      const int R = t.seq.pass[q].radix;
      if (pp > 1 && tw_two_level(R, pp)) {
        jobs.push_back({c, (long long)(R - 1) * 16, (long long)pp * R, 16, 0, 0});
        c += (long long)(R - 1) * 16;
        jobs.push_back({c, (long long)(R - 1) * (pp / 16), (long long)pp * R / 16, (long long)(pp / 16), 0, 0});
        c += (long long)(R - 1) * (pp / 16);
      } else if (pp > 1) {
        jobs.push_back({c, (long long)(R - 1) * pp, (long long)pp * R, (long long)pp, 0, 0});
        c += (long long)(R - 1) * pp;
      }
      pp *= R;
    }
  };
  for (const SeqTab& t : seqs) table_jobs(t);

  // ---- tables on the device (FP64 math, rounded once)
  if ((rc = dev_alloc(p->arena, (size_t)std::max<long long>(arena, 1) * esz))) return rc;
  {
    DevBuf jd;
    if ((rc = dev_alloc(jd, std::max<size_t>(jobs.size(), 1) * sizeof(RootJob)))) return rc;
    if (!jobs.empty() && hipMemcpy(jd.p, jobs.data(), jobs.size() * sizeof(RootJob), hipMemcpyHostToDevice) != hipSuccess)
      return CHFFT_ERR_HIP;
    if (gen_roots(f64, reinterpret_cast<const RootJob*>(jd.p), (int)jobs.size(), p->arena.p, nullptr) != hipSuccess)
      return CHFFT_ERR_HIP;
    if (hipDeviceSynchronize() != hipSuccess) return CHFFT_ERR_HIP;
  }
  // Bluestein spectra: FP64 FFT of the folded chirp, grouped by M, in chunks
  {
    std::map<size_t, std::vector<size_t>> by_m;
    for (auto& kv : info)
      if (kv.second.blue && kv.second.kind != 3) by_m[kv.second.M].push_back(kv.first);
    const size_t chunk_bytes = (size_t)1 << 30;
    for (auto& kv : by_m) {
      const size_t M = kv.first;
      const std::vector<size_t>& ns = kv.second;
      const size_t per = std::max<size_t>(1, chunk_bytes / (M * 16));
      for (size_t b0 = 0; b0 < ns.size(); b0 += per) {
        const size_t cnt = std::min(per, ns.size() - b0);
        std::vector<long long> nv(cnt), offs(cnt);
        for (size_t k = 0; k < cnt; k++) {
          nv[k] = (long long)ns[b0 + k];
          offs[k] = info[ns[b0 + k]].kspec;
        }
        DevBuf nd, od, buf;
        if ((rc = dev_alloc(nd, cnt * 8)) || (rc = dev_alloc(od, cnt * 8)) || (rc = dev_alloc(buf, cnt * M * 16)))
          return rc;
        if (hipMemcpy(nd.p, nv.data(), cnt * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(od.p, offs.data(), cnt * 8, hipMemcpyHostToDevice) != hipSuccess)
          return CHFFT_ERR_HIP;
        if (gen_rot(reinterpret_cast<const long long*>(nd.p), (int)cnt, (long long)M, buf.p, nullptr) != hipSuccess)
          return CHFFT_ERR_HIP;
        chfft_plan sub = nullptr;
        if ((rc = chfft_plan_c2c_1d(&sub, M, cnt, CHFFT_F64))) return rc;
        rc = chfft_exec(sub, buf.p, buf.p, CHFFT_FORWARD, nullptr);
        const LenInfo& L0 = info[ns[b0]];  // every length of this M has the same split
        if (!rc && put_spectra(f64, buf.p, (int)cnt, (long long)M, L0.kind == 2 ? (long long)L0.n1 : 0,
                               reinterpret_cast<const long long*>(od.p),
                               p->arena.p, nullptr) != hipSuccess)
          rc = CHFFT_ERR_HIP;
        if (!rc && hipDeviceSynchronize() != hipSuccess) rc = CHFFT_ERR_HIP;
        chfft_plan_destroy(sub);
        if (rc) return rc;
      }
    }
  }

  // ---- FP64 two-level tables for the four-step twiddle stores, one pair
  // per distinct M: [W_M^i, i < 2^twsh] then [W_M^(i 2^twsh), i <
  // ceil(M / 2^twsh)]; the stores take W^(c k) for a thread's first output
  // and the step between its outputs from them and run the rest by an FP64
  // recurrence, each twiddle rounded once
  {
    std::map<size_t, long long> rec_of_m;
    std::vector<RootJob> jobs64;
    long long a64 = 0;
    for (auto& kv : info) {
      LenInfo& L = kv.second;
      if (L.kind != 1 && L.kind != 2) continue;
      auto it = rec_of_m.find(L.M);
      if (it != rec_of_m.end()) {
        L.rec64 = it->second;
        continue;
      }
      const long long nlo = 1ll << L.twsh, nhi = (long long)((L.M + nlo - 1) >> L.twsh);
      jobs64.push_back({a64, nlo, (long long)L.M, 1, 3, 0});
      jobs64.push_back({a64 + nlo, nhi, (long long)L.M, 1, 2, L.twsh});
      L.rec64 = rec_of_m[L.M] = a64;
      a64 += nlo + nhi;
    }
    if (!jobs64.empty()) {
      if ((rc = dev_alloc(p->arena64, (size_t)a64 * 16))) return rc;
      DevBuf jd;
      if ((rc = dev_alloc(jd, jobs64.size() * sizeof(RootJob)))) return rc;
      if (hipMemcpy(jd.p, jobs64.data(), jobs64.size() * sizeof(RootJob), hipMemcpyHostToDevice) != hipSuccess)
        return CHFFT_ERR_HIP;
      if (gen_roots(true, reinterpret_cast<const RootJob*>(jd.p), (int)jobs64.size(), p->arena64.p, nullptr) !=
          hipSuccess)
        return CHFFT_ERR_HIP;
      if (hipDeviceSynchronize() != hipSuccess) return CHFFT_ERR_HIP;
    }
  }

  // ---- descriptors (per distinct length, both directions) and items
  const char* base = reinterpret_cast<const char*>(p->arena.p);
  auto at = [&](long long off) { return (const void*)(base + (size_t)off * esz); };
  std::vector<KParams> dfw, dbw;
  size_t lds = 0;
  for (auto& kv : info) {
    LenInfo& L = kv.second;
    if (L.kind == 3) continue;
    L.desc = (int)dfw.size();
    for (int dir = 0; dir < 2; dir++) {
      std::vector<KParams>& D = dir ? dbw : dfw;
      if (L.kind == 0) {
        KParams P{};
        P.seq = seqs[L.seq].seq;
        P.seq.tw = at(seqs[L.seq].off);
        P.Lfft = (int)L.Lfft;
        P.L = (int)L.n;
        P.G = 1;
        P.tpb = 1;
        P.Lp = lp_for(L.Lfft, 1);
        P.in = io(0, 0, 1, 1);
        P.out = P.in;
        P.conj_in = P.conj_out = dir;
        P.has_mid = P.has_chirp = L.blue ? 1 : 0;
        if (L.blue) {
          P.chirp = at(L.chirp);
          P.kspec = at(L.kspec);
          P.kc = 0;
          P.ke = 1;
        }
        lds = std::max(lds, (size_t)P.Lp * esz);
        D.push_back(P);
      } else {
        const int ta = phase_tpb(L.n1, block, true, f64, maxtpb), tb = phase_tpb(L.n2, block, L.kind == 1, f64, maxtpb);
        KParams A{}, B{}, C{};
        A.seq = seqs[L.s1].seq;
        A.seq.tw = at(seqs[L.s1].off);
        A.Lfft = A.L = (int)L.n1;
        A.G = (int)L.n2;
        A.tpb = ta;
        A.Lp = lp_for(L.n1, ta);
        A.in = io(0, 1, (long long)L.n2, (int)L.n2);
        A.out = A.in;
        A.colmajor_ld = A.colmajor_st = 1;
        A.conj_in = dir;
        A.twsh = L.twsh;
        A.tab_a = at(L.tw_hi);
        A.tab_b = at(L.tw_lo);
        B.seq = seqs[L.s2].seq;
        B.seq.tw = at(seqs[L.s2].off);
        B.Lfft = B.L = (int)L.n2;
        B.G = (int)L.n1;
        B.tpb = tb;
        B.Lp = lp_for(L.n2, tb);
        B.in = io(0, (long long)L.n2, 1, (int)L.n1);
        if (L.kind == 1) {
          B.out = io(0, 1, (long long)L.n1, (int)L.n1);
          B.colmajor_st = 1;
          B.conj_out = dir;
        } else {
          A.N = (long long)L.n;
          A.chirp = at(L.chirp);
          B.out = B.in;
          B.has_mid = 1;
          B.kspec = at(L.kspec);  // row-major [k1][k2]
          B.kc = (long long)L.n2;
          B.ke = 1;
          B.twsh = L.twsh;
          B.tab_a = at(L.tw_hi);
          B.tab_b = at(L.tw_lo);
          C = A;
          C.conj_in = 0;
          C.twsh = 0;
          C.tab_a = C.tab_b = nullptr;
          C.out = io(0, 1, (long long)L.n2, (int)L.n2);
          C.N = (long long)L.n;
          C.chirp = at(L.chirp);
          C.conj_out = dir;
        }
        lds = std::max(lds, (size_t)ta * A.Lp * esz);
        lds = std::max(lds, (size_t)tb * B.Lp * esz);
        // the four-step twiddle stores (columns; Bluestein rows) run by FP64
        // recurrence from the FP64 two-level tables, in the generic stores
        // and in the register-resident tiles alike
        if (L.rec64 >= 0) {
          const void* t64 = reinterpret_cast<const char*>(p->arena64.p) + (size_t)L.rec64 * 16;
          A.tw_rec = t64;
          if (L.kind == 2) B.tw_rec = t64;
        }
        KParams* Qs[3] = {&A, &B, &C};
        for (int ph = 0; ph < 3; ph++)
          if (L.spec[ph]) {
            KParams& Q = *Qs[ph];
            Q.spec = L.spec[ph];
            Q.tpb = L.spec_tpb[ph];
            Q.seq = seqs[L.spec_seq[ph]].seq;
            Q.seq.tw = at(seqs[L.spec_seq[ph]].off);
            lds = std::max(lds, L.spec_lds[ph]);
          }
        D.push_back(A);
        D.push_back(B);
        if (L.kind == 2) D.push_back(C);
      }
    }
    L.ndesc = (int)dfw.size() - L.desc;
  }
  // items: offsets in the packed buffers and, for four-step items, in the
  // scratch buffer (each keeps its intermediate from one phase to the next)
  std::vector<MultiItem> items(count);
  size_t acc = 0;
  long long sacc = 0;
  for (size_t i = 0; i < count; i++) {
    const LenInfo& L = info[lens[i]];
    items[i].off = (long long)acc;
    acc += lens[i];
    items[i].n = (int)lens[i];
    items[i].kind = L.kind;
    items[i].desc = L.desc;
    items[i].soff = 0;
    if (L.kind == 1 || L.kind == 2) {
      items[i].soff = sacc;
      sacc += (long long)L.M;
    }
  }
  // units: tiles of each four-step phase in chunks of about kChunk points,
  // whole fused items; fused items go to the phase with the least work so
  // far; each phase sorted by decreasing cost (workgroup w takes units w,
  // w + grid, ...)
  const double kChunk = t.unit ? (double)t.unit : 32768.0;
  std::vector<std::pair<double, MultiUnit>> ph_units[kMultiPhases];
  double ph_cost[kMultiPhases] = {0, 0, 0};
  int nph = 1;
  for (size_t i = 0; i < count; i++) {
    const LenInfo& L = info[lens[i]];
    if (L.kind == 1 || L.kind == 2) {
      const int np = L.kind == 1 ? 2 : 3;
      nph = std::max(nph, np);
      for (int ph = 0; ph < np; ph++) {
        const KParams& Q = dfw[L.desc + ph];
        const double tile_pts = (double)Q.tpb * Q.Lfft;
        const int tiles = (Q.G + Q.tpb - 1) / Q.tpb;
        const int per = std::max(1, (int)(kChunk / tile_pts));
        for (int t0 = 0; t0 < tiles; t0 += per) {
          MultiUnit U{(int)i, ph, t0 * Q.tpb, std::min(Q.G, (t0 + per) * Q.tpb)};
          const double c = (double)(U.g1 - U.g0) * Q.Lfft * std::log2((double)Q.Lfft + 1.0) * (ph == 1 && L.kind == 2 ? 2.0 : 1.0);
          ph_units[ph].push_back({c, U});
          ph_cost[ph] += c;
        }
      }
    }
  }
  for (size_t i = 0; i < count; i++) {
    const LenInfo& L = info[lens[i]];
    if (L.kind == 1 || L.kind == 2) continue;
    const double m = (double)std::max<size_t>(L.M, 1);
    const double c = L.kind == 3 ? 1.0 : m * std::log2(m + 1.0) * (L.blue ? 2.0 : 1.0);
    int best = 0;
    for (int ph = 1; ph < nph; ph++)
      if (ph_cost[ph] < ph_cost[best]) best = ph;
    ph_units[best].push_back({c, MultiUnit{(int)i, 0, 0, 1}});
    ph_cost[best] += c;
  }
  std::vector<MultiUnit> units;
  for (int ph = 0; ph < kMultiPhases; ph++) {
    std::stable_sort(ph_units[ph].begin(), ph_units[ph].end(),
                     [](const std::pair<double, MultiUnit>& a, const std::pair<double, MultiUnit>& b) {
                       return a.first > b.first;
                     });
    for (auto& cu : ph_units[ph]) units.push_back(cu.second);
    p->unit_ends[ph] = (int)units.size();
  }
  if ((rc = dev_alloc(p->items_dev, std::max<size_t>(count, 1) * sizeof(MultiItem)))) return rc;
  if ((rc = dev_alloc(p->units_dev, std::max<size_t>(units.size(), 1) * sizeof(MultiUnit)))) return rc;
  if ((rc = dev_alloc(p->descs_fwd, std::max<size_t>(dfw.size(), 1) * sizeof(KParams)))) return rc;
  if ((rc = dev_alloc(p->descs_bwd, std::max<size_t>(dbw.size(), 1) * sizeof(KParams)))) return rc;
  if ((rc = dev_alloc(p->sync_dev, 2 * sizeof(unsigned)))) return rc;
  if ((count && hipMemcpy(p->items_dev.p, items.data(), count * sizeof(MultiItem), hipMemcpyHostToDevice)) ||
      (!units.empty() &&
       hipMemcpy(p->units_dev.p, units.data(), units.size() * sizeof(MultiUnit), hipMemcpyHostToDevice)) ||
      (!dfw.empty() && hipMemcpy(p->descs_fwd.p, dfw.data(), dfw.size() * sizeof(KParams), hipMemcpyHostToDevice)) ||
      (!dbw.empty() && hipMemcpy(p->descs_bwd.p, dbw.data(), dbw.size() * sizeof(KParams), hipMemcpyHostToDevice)) ||
      hipMemset(p->sync_dev.p, 0, 2 * sizeof(unsigned)) != hipSuccess)
    return CHFFT_ERR_HIP;
  p->nitems = (int)count;
  // CHFFT_MULTI_BARRIER_TEST=1 (tests only): every grid-wide barrier waits
  // for one more arrival than the grid makes, for 10 us, so that each wait
  // gives up at once and the failure path runs with every wave exiting
  if (const char* e = getenv("CHFFT_MULTI_BARRIER_TEST"))
    if (strcmp(e, "1") == 0) {
      p->barrier_ticks = 1000;
      p->barrier_extra = 1;
    }
  p->multi_lds = std::max<size_t>(lds, 16);
  p->multi_occ = multi_occupancy(f64, p->multi_lds);
  // every workgroup resident at once (the phases meet at grid-wide barriers)
  p->multi_grid = std::max(1, p->multi_occ * C2CEngine::num_cus());
  if (t.grid > p->multi_grid) return CHFFT_ERR_INVALID_ARGUMENT;
  if (t.grid) p->multi_grid = t.grid;
  if ((rc = dev_alloc(p->scratch, (size_t)std::max<long long>(sacc, 1) * esz))) return rc;
  size_t nf = 0, nf4 = 0, nb4 = 0, nspec = 0;
  for (auto& kv : info) {
    nf += kv.second.kind == 0;
    nf4 += kv.second.kind == 1;
    nb4 += kv.second.kind == 2;
    for (int ph = 0; ph < 3; ph++) nspec += kv.second.spec[ph] != 0;
  }
  p->multi_desc = "multi-length batch: " + std::to_string(count) + " transforms, " + std::to_string(total) +
                  " points, " + std::to_string(info.size()) + " distinct lengths (" + std::to_string(nf) +
                  " fused in LDS, " + std::to_string(nf4) + " four-step, " + std::to_string(nb4) +
                  " four-step Bluestein; " + std::to_string(nspec) +
                  " four-step phases on register-resident tiles) | one persistent launch, " + std::to_string(nph) +
                  " phase(s), " +
                  std::to_string(units.size()) + " work units: grid=" + std::to_string(p->multi_grid) +
                  " block=" + std::to_string(block) + " lds=" + std::to_string(p->multi_lds) +
                  " scratch=" + std::to_string((size_t)std::max<long long>(sacc, 1) * esz) + "\n";
  *plan = p.release();
  return CHFFT_OK;
}

extern "C" int chfft_plan_c2c_multi(chfft_plan* plan, const size_t* lens, size_t count, int precision) {
  return multi_create(lens, count, precision, Tune{}, plan);
}

// Sets the number of workgroups of a multi-length plan's persistent launch
// (the autotuner's knob); 0 restores the default, every workgroup that fits
// at once. Larger grids are refused: the launch's phases meet at grid-wide
// barriers, which need every workgroup resident.
extern "C" int chfft_plan_set_grid(chfft_plan plan, int grid) {
  if (!plan || plan->kind != K_MULTI || grid < 0) return CHFFT_ERR_INVALID_ARGUMENT;
  const int cap = std::max(1, plan->multi_occ * C2CEngine::num_cus());
  if (grid > cap) return CHFFT_ERR_INVALID_ARGUMENT;
  plan->multi_grid = grid ? grid : cap;
  plan->tune.grid = grid == cap ? 0 : grid;
  return CHFFT_OK;
}

// ---------------------------------------------------------------- autotuner
namespace chfft {
namespace {

// Rebuilds plan `p` under the overrides `t` (CFft1D and multi-length plans;
// the other kinds take only the planner's own choices). The plan keeps its
// handle; its buffers and launches are replaced once the device is idle.
int apply_tune(chfft_plan_s* p, const Tune& t) {
  chfft_plan fresh = nullptr;
  int rc;
  if (p->kind == K_C2C)
    rc = c2c_create(p->n, p->batch, p->f64, t, &fresh);
  else if (p->kind == K_MULTI)
    rc = multi_create(p->lens.data(), p->lens.size(), p->f64 ? CHFFT_F64 : CHFFT_F32, t, &fresh);
  else
    return tune_str(t) == "default" ? CHFFT_OK : CHFFT_ERR_UNSUPPORTED;
  if (rc) return rc;
  if (hipDeviceSynchronize() != hipSuccess) {
    delete fresh;
    return CHFFT_ERR_HIP;
  }
  std::string report = std::move(p->tune_report);
  *p = std::move(*fresh);
  p->tune_report = std::move(report);
  delete fresh;
  return CHFFT_OK;
}

// Aligned four-step splits of `total` other than n1 (n2 a multiple of 16,
// both factors held by one workgroup), nearest to the square root first.
std::vector<size_t> alt_splits(size_t total, size_t n1, bool f64, size_t maxn) {
  std::vector<size_t> v;
  for (size_t d = 2; d <= lds_cap(f64) && d < total; d++) {
    if (total % d || d == n1 || !smooth7(d) || !smooth7(total / d)) continue;
    if ((total / d) % 16 || (total / d) * 4 > lds_cap(f64)) continue;
    v.push_back(d);
  }
  const double mid = 0.5 * std::log((double)total);
  std::sort(v.begin(), v.end(), [&](size_t a, size_t b) {
    return std::fabs(std::log((double)a) - mid) < std::fabs(std::log((double)b) - mid);
  });
  if (v.size() > maxn) v.resize(maxn);
  return v;
}

// The autotuner's candidates for plan `p`, the planner's own choice first.
// Fused lengths: every radix schedule of radix_variants() and workgroup
// shapes around the planner's (half / double the transforms per workgroup,
// the fewest threads per transform, half / double the threads) on the
// runtime-radix kernel. Four-step lengths: the persistent schedule, the
// runtime-radix kernels, chunk sizes, other splits and radix orders.
// Multi-length plans: workgroup counts, work-unit sizes, generic tile widths,
// runtime-radix tiles.
std::vector<Tune> tune_candidates(const chfft_plan_s* p) {
  std::vector<Tune> c(1);
  if (p->kind == K_MULTI) {
    const int def = std::max(1, p->multi_occ * C2CEngine::num_cus());
    for (int g : {def / 2, def})
      for (int u : {16384, 0, 65536}) {
        Tune t;
        t.grid = g == def ? 0 : std::max(1, g);
        t.unit = u;
        if (t.grid || t.unit) c.push_back(t);
      }
    for (int w : {16, 32}) {  // narrower generic four-step tiles than the default 64
      Tune t;
      t.tpb = w;
      c.push_back(t);
    }
    if (!p->f64) {
      Tune t;
      t.generic = 1;
      c.push_back(t);
    }
    return c;
  }
  if (p->kind != K_C2C) return c;
  const C2CEngine& e = p->rows;
  if (e.mode == 1 && e.big_fused) {  // the four-step Bluestein schedules instead
    for (int k = 0; k < 3; k++) {
      Tune t;
      t.sched = k == 0 ? 2 : k == 1 ? 1 : 0;
      t.generic = k == 2;
      c.push_back(t);
    }
  } else if (e.mode == 1) {
    const int Lf = (int)e.inner.Lfft;
    const size_t nv = radix_variants((size_t)Lf).size();
    for (size_t o = 0; o < nv; o++) {
      Tune t;
      t.generic = 1;
      t.order = (int)o;
      c.push_back(t);
    }
    const Geometry g = choose_geometry(Lf, (long long)p->batch, false, p->f64);
    const int tmin = (Lf + kEmax - 1) / kEmax;
    const int shapes[5][2] = {{g.tpb / 2, g.tpt}, {g.tpb * 2, g.tpt}, {g.tpb, tmin}, {g.tpb, g.tpt / 2}, {g.tpb, g.tpt * 2}};
    for (const auto& sh : shapes) {
      Geometry gt;
      if (sh[0] == g.tpb && sh[1] == g.tpt) continue;
      if (!geometry_with(Lf, sh[0], sh[1], p->f64, gt)) continue;
      bool dup = false;
      for (const Tune& o : c) dup |= o.tpb == sh[0] && o.tpt == sh[1];
      if (dup) continue;
      Tune t;
      t.generic = 1;
      t.tpb = sh[0];
      t.tpt = sh[1];
      c.push_back(t);
    }
  } else if (e.mode == 2 || e.mode == 3) {
    Tune t;
    t.sched = 1;
    c.push_back(t);
    t = Tune{};
    t.generic = 1;
    c.push_back(t);
    const size_t cb = e.chunk_of(p->batch);
    if (cb < p->batch) {
      t = Tune{};
      t.chunk = -1;
      c.push_back(t);
    }
    if (cb > 1) {
      t = Tune{};
      t.chunk = (long long)(cb / 2);
      c.push_back(t);
    }
    for (size_t d : alt_splits(e.mode == 2 ? e.n : e.M, e.n1, p->f64, 3)) {
      t = Tune{};
      t.n1 = (long long)d;
      c.push_back(t);
    }
    const size_t nv = std::min(radix_variants(e.n1).size(), radix_variants(e.n2).size());
    for (size_t o = 1; o < std::min<size_t>(nv, 3); o++) {
      t = Tune{};
      t.generic = 1;
      t.order = (int)o;
      c.push_back(t);
    }
  }
  return c;
}

// Average time of one forward exec over `reps` back-to-back execs (after
// one untimed exec), best of two such runs.
double time_exec(chfft_plan_s* p, void* in, void* out, int reps, hipStream_t st) {
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess) return -1;
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return -1;
  }
  double best = -1;
  bool ok = chfft_exec(p, in, out, CHFFT_FORWARD, st) == CHFFT_OK;
  for (int run = 0; run < 2 && ok; run++) {
    ok = hipEventRecord(a, st) == hipSuccess;
    for (int r = 0; r < reps && ok; r++) ok = chfft_exec(p, in, out, CHFFT_FORWARD, st) == CHFFT_OK;
    float f = 0;
    ok = ok && hipEventRecord(b, st) == hipSuccess && hipEventSynchronize(b) == hipSuccess &&
         hipEventElapsedTime(&f, a, b) == hipSuccess;
    if (ok && (best < 0 || f / reps < best)) best = f / reps;
  }
  if (ok && chfft_plan_check(p, st) != CHFFT_OK) best = -1;  // a multi-length launch that did not complete
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ok ? best : -1;
}

}  // namespace
}  // namespace chfft

extern "C" int chfft_plan_tune(chfft_plan plan, const char* spec) {
  if (!plan || !spec) return CHFFT_ERR_INVALID_ARGUMENT;
  Tune t;
  if (tune_parse(spec, t)) return CHFFT_ERR_INVALID_ARGUMENT;
  return apply_tune(plan, t);
}

extern "C" int chfft_plan_tuning(chfft_plan plan, char* buf, size_t buf_len) {
  if (!plan || !buf || buf_len == 0) return CHFFT_ERR_INVALID_ARGUMENT;
  const std::string s = tune_str(plan->tune) + "\n" + plan->tune_report;
  const size_t n = std::min(s.size(), buf_len - 1);
  memcpy(buf, s.data(), n);
  buf[n] = '\0';
  return CHFFT_OK;
}

extern "C" int chfft_plan_autotune(chfft_plan plan, void* in, void* out, int reps, void* stream) {
  if (!plan || !in || !out || reps <= 0) return -CHFFT_ERR_INVALID_ARGUMENT;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const std::vector<Tune> cands = tune_candidates(plan);
  std::string report = "autotune (" + std::to_string(reps) + " execs per candidate, ms per exec):";
  int best = -1;
  double best_ms = 1e30;
  for (size_t i = 0; i < cands.size(); i++) {
    chfft_plan fresh = nullptr;
    const int rc = plan->kind == K_MULTI
                       ? multi_create(plan->lens.data(), plan->lens.size(), plan->f64 ? CHFFT_F64 : CHFFT_F32,
                                      cands[i], &fresh)
                       : plan->kind == K_C2C ? c2c_create(plan->n, plan->batch, plan->f64, cands[i], &fresh)
                                             : CHFFT_ERR_UNSUPPORTED;
    if (rc) {
      if (i == 0 && plan->kind != K_C2C && plan->kind != K_MULTI) return 0;  // a single schedule
      continue;
    }
    const double ms = time_exec(fresh, in, out, reps, st);
    (void)hipStreamSynchronize(st);
    delete fresh;
    char line[160];
    snprintf(line, sizeof line, "\n  [%zu] %s: %.4f", i, tune_str(cands[i]).c_str(), ms);
    report += line;
    if (ms > 0 && ms < best_ms) {
      best_ms = ms;
      best = (int)i;
    }
  }
  if (best < 0) return -CHFFT_ERR_HIP;
  const int rc = apply_tune(plan, cands[best]);
  if (rc) return -rc;
  plan->tune_report = report + "\n  -> [" + std::to_string(best) + "] " + tune_str(cands[best]);
  return best;
}
