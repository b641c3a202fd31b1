// CPU oracle: a restatement of chfft 0.3.4's transform engine (Rust crate, MPL-2.0).
//
// TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load this library, and only as the checker / the timed CPU
// baseline. The product (libchfft_hip.so) never links or calls it.
//
// The reference cannot be built here (Rust; no cargo/rustc, no vendored crates),
// so this file restates its algorithms in the same arithmetic order, following
// each reference function cited below (paths relative to the chfft source tree):
//   prime_factorization.rs:14-113   factor order (4s, one 2, odd primes; 4-group deferred)
//   precompute_utils.rs:15-104      twiddle table with quarter/half-period symmetry,
//                                   digit reversal, swap-sequence conversion
//   mixed_radix.rs:24-368           out-of-place / in-place drivers, radix-2/3/4/5/generic DIT
//   chirpz.rs:27-203                Bluestein: chirp, radix-2/4 M-point FFTs, de-chirp
//   cfft1d.rs:103-256               plan setup (MAX_PRIME = 7), dispatch, scalings
//   rfft1d.rs:44-187                real FFT via N/2 complex FFT + split
//   cfft2d.rs:326-363               row/column 2-D FFT
//   dct1d.rs:199-410                DCT-II / DCT-III via N/2 complex FFT
//   mdct1d.rs:42-291                MDCT / IMDCT via N/4 complex FFT
// Complex arithmetic follows num-complex 0.4 (plain (ac-bd, ad+bc) products, no
// fused multiply-add: build with -ffp-contract=off). sin/cos come from the same
// platform libm the Rust f32/f64 methods call.
//
// Parity pinning (what checks each function, beyond retyping): see the tests in
// tests/test_oracle.py — brute-force long-double DFT/DCT/MDCT definitions
// (the reference's own test oracles), numpy/pocketfft golden fixtures, the
// README known-answer vector, round-trip identities, and MDCT TDAC.
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <chrono>
#include <thread>
#include <vector>

namespace oracle {

template <class T>
struct Cx {
  T re, im;
  Cx() : re(0), im(0) {}
  Cx(T r, T i) : re(r), im(i) {}
};
template <class T> inline Cx<T> operator+(Cx<T> a, Cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <class T> inline Cx<T> operator-(Cx<T> a, Cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <class T> inline Cx<T> operator*(Cx<T> a, Cx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class T> inline Cx<T> scale(Cx<T> a, T s) { return {a.re * s, a.im * s}; }
template <class T> inline Cx<T> conj(Cx<T> a) { return {a.re, -a.im}; }
// lib.rs:27-39
template <class T> inline Cx<T> quarter_turn(Cx<T> a) { return {-a.im, a.re}; }
template <class T> inline Cx<T> three_quarter_turn(Cx<T> a) { return {a.im, -a.re}; }
// Complex::i() * w (num-complex Mul): (0*w.re - 1*w.im, 0*w.im + 1*w.re)
template <class T> inline Cx<T> times_i(Cx<T> w) { return Cx<T>(T(0), T(1)) * w; }

template <class T> struct Pi;
template <> struct Pi<float> { static float v() { return 3.14159265358979323846f; } };
template <> struct Pi<double> { static double v() { return 3.14159265358979323846; } };

// ---------------------------------------------------------------- prime_factorization.rs
struct Factor {
  size_t value, count;
};

// FactorIterator::next, prime_factorization.rs:14-60
struct FactorIterator {
  size_t value, prime;
  bool next(size_t* out) {
    for (;;) {
      if (prime == 4) {
        if (value != 0 && (value & 3) == 0) {
          value >>= 2;
          *out = 4;
          return true;
        }
        prime = 2;
      } else if (prime == 2) {
        if ((value & 1) == 0) {
          value >>= 1;
          *out = 2;
          return true;
        }
        prime = 3;
      } else {
        if (value >= prime * prime) {
          if (value % prime == 0) {
            value /= prime;
            *out = prime;
            return true;
          }
          prime += 2;
        } else if (value > 1) {
          *out = value;
          value = 0;
          return true;
        } else {
          return false;
        }
      }
    }
  }
};

// prime_factorization.rs:66-113 (value 0 is rejected by the callers; the
// reference loops forever on it)
std::vector<Factor> prime_factorization(size_t value, size_t max) {
  std::vector<Factor> factors;
  size_t count = 0, prime = 0, fac4count = 0;
  FactorIterator it{value, 4};
  size_t p;
  while (it.next(&p)) {
    if (p > max) return {};
    if (prime != p) {
      if (count > 0) {
        if (prime == 4) {
          fac4count = count;
        } else {
          if (prime > 4 && fac4count != 0) {
            factors.push_back({4, fac4count});
            fac4count = 0;
          }
          factors.push_back({prime, count});
        }
      }
      count = 1;
      prime = p;
    } else {
      count += 1;
    }
  }
  if (count > 0) factors.push_back({prime, count});
  if (fac4count != 0) factors.push_back({4, fac4count});
  return factors;
}

// ---------------------------------------------------------------- precompute_utils.rs
// precompute_utils.rs:15-21: from_polar(1, (-2 * PI / len) * pos), all in T.
template <class T>
Cx<T> calc_omega_item(size_t len, size_t pos) {
  T theta = T(-2.0) * Pi<T>::v() / T(len) * T(pos);
  return {std::cos(theta), std::sin(theta)};
}

// precompute_utils.rs:24-68
template <class T>
std::vector<Cx<T>> calc_omega(size_t len) {
  std::vector<Cx<T>> omega;
  omega.reserve(len + 1);
  omega.push_back(Cx<T>(1, 0));
  if (len != 0 && (len & 3) == 0) {
    size_t q = len >> 2, h = len >> 1;
    for (size_t i = 1; i < q; i++) omega.push_back(calc_omega_item<T>(len, i));
    for (size_t i = q; i < h; i++) {
      Cx<T> t = omega[i - q];
      omega.push_back(Cx<T>(t.im, -t.re));
    }
    for (size_t i = h; i < len; i++) {
      Cx<T> t = omega[i - h];
      omega.push_back(Cx<T>(-t.re, -t.im));
    }
  } else if ((len & 1) == 0) {
    size_t h = std::max<size_t>(len >> 1, 1);
    for (size_t i = 1; i < h; i++) omega.push_back(calc_omega_item<T>(len, i));
    for (size_t i = h; i < len; i++) {
      Cx<T> t = omega[i - h];
      omega.push_back(Cx<T>(-t.re, -t.im));
    }
  } else {
    for (size_t i = 1; i < len; i++) omega.push_back(calc_omega_item<T>(len, i));
  }
  omega.push_back(Cx<T>(1, 0));
  return omega;
}

// precompute_utils.rs:70-91
std::vector<size_t> calc_bitreverse(size_t len, const std::vector<Factor>& factors) {
  std::vector<size_t> ids;
  ids.reserve(len);
  size_t llen = 1;
  ids.push_back(0);
  for (const Factor& f : factors) {
    for (size_t c = 0; c < f.count; c++) {
      for (size_t j = 0; j < llen; j++) ids[j] *= f.value;
      for (size_t i = 1; i < f.value; i++)
        for (size_t j = 0; j < llen; j++) ids.push_back(ids[j] + i);
      llen *= f.value;
    }
  }
  return ids;
}

// precompute_utils.rs:93-104. The reference searches linearly (O(N^2)); this
// keeps an inverse map so the same swap sequence comes out in O(N).
std::vector<size_t> calc_bitreverse2inplace(const std::vector<size_t>& source) {
  size_t n = source.size();
  std::vector<size_t> nums(n), where(n), out(n);
  for (size_t i = 0; i < n; i++) nums[i] = where[i] = i;
  for (size_t i = 0; i < n; i++) {
    size_t r = where[source[i]];  // first j >= i with nums[j] == source[i]
    std::swap(nums[i], nums[r]);
    where[nums[i]] = i;
    where[nums[r]] = r;
    out[i] = r;
  }
  return out;
}

// ---------------------------------------------------------------- mixed_radix.rs
// mixed_radix.rs:168-192
template <class T>
void radix2(Cx<T>* ret, size_t count, size_t& po2, size_t& rad, const Cx<T>* omega, size_t len) {
  for (size_t c = 0; c < count; c++) {
    size_t po2m = po2;
    po2 <<= 1;
    rad >>= 1;
    for (size_t j0 = 0; j0 < po2m; j0++) {
      Cx<T> w1 = omega[rad * j0];
      for (size_t j = j0; j < len; j += po2) {
        size_t pos1 = j + po2m;
        Cx<T> z1 = ret[pos1] * w1;
        ret[pos1] = ret[j] - z1;
        ret[j] = ret[j] + z1;
      }
    }
  }
}

// mixed_radix.rs:194-228
template <class T, class F>
void radix3(Cx<T>* ret, size_t count, size_t& po2, size_t& rad, const Cx<T>* omega, size_t len, F turn) {
  T t3scaler = std::sin(T(-2.0) * Pi<T>::v() / T(3.0));
  for (size_t c = 0; c < count; c++) {
    size_t po2m = po2;
    po2 *= 3;
    rad /= 3;
    for (size_t j0 = 0; j0 < po2m; j0++) {
      size_t wpos = rad * j0;
      Cx<T> w1 = omega[wpos], w2 = omega[wpos << 1];
      for (size_t j = j0; j < len; j += po2) {
        size_t pos1 = j + po2m, pos2 = pos1 + po2m;
        Cx<T> z1 = ret[pos1] * w1;
        Cx<T> z2 = ret[pos2] * w2;
        Cx<T> t1 = z1 + z2;
        Cx<T> t2 = ret[j] - scale(t1, T(0.5));
        Cx<T> t3 = scale(turn(z1 - z2), t3scaler);
        ret[j] = ret[j] + t1;
        ret[pos1] = t2 + t3;
        ret[pos2] = t2 - t3;
      }
    }
  }
}

// mixed_radix.rs:230-270
template <class T, class F>
void radix4(Cx<T>* ret, size_t count, size_t& po2, size_t& rad, const Cx<T>* omega, size_t len, F turn) {
  for (size_t c = 0; c < count; c++) {
    size_t po2m = po2;
    po2 <<= 2;
    rad >>= 2;
    for (size_t j0 = 0; j0 < po2m; j0++) {
      size_t wpos = rad * j0;
      Cx<T> w1 = omega[wpos], w2 = omega[wpos << 1], w3 = omega[wpos * 3];
      for (size_t j = j0; j < len; j += po2) {
        size_t pos1 = j + po2m, pos2 = pos1 + po2m, pos3 = pos2 + po2m;
        Cx<T> wfa = ret[j];
        Cx<T> wfc = ret[pos1] * w1;
        Cx<T> wfb = ret[pos2] * w2;
        Cx<T> wfd = ret[pos3] * w3;
        Cx<T> wfab = wfa + wfb;
        Cx<T> wfamb = wfa - wfb;
        Cx<T> wfcd = wfc + wfd;
        Cx<T> wfcimdi = turn(wfc - wfd);
        ret[j] = wfab + wfcd;
        ret[pos1] = wfamb - wfcimdi;
        ret[pos2] = wfab - wfcd;
        ret[pos3] = wfamb + wfcimdi;
      }
    }
  }
}

// mixed_radix.rs:272-334
template <class T, class F>
void radix5(Cx<T>* ret, size_t count, size_t& po2, size_t& rad, const Cx<T>* omega, size_t len, F turn) {
  const T pi = Pi<T>::v();
  for (size_t c = 0; c < count; c++) {
    size_t po2m = po2;
    po2 *= 5;
    rad /= 5;
    for (size_t j0 = 0; j0 < po2m; j0++) {
      size_t wpos = rad * j0;
      Cx<T> w1 = omega[wpos], w2 = omega[wpos << 1], w3 = omega[wpos * 3], w4 = omega[wpos << 2];
      for (size_t j = j0; j < len; j += po2) {
        size_t pos2 = j + po2m, pos3 = pos2 + po2m, pos4 = pos3 + po2m, pos5 = pos4 + po2m;
        Cx<T> z0 = ret[j];
        Cx<T> z1 = ret[pos2] * w1;
        Cx<T> z2 = ret[pos3] * w2;
        Cx<T> z3 = ret[pos4] * w3;
        Cx<T> z4 = ret[pos5] * w4;
        Cx<T> t1 = z1 + z4;
        Cx<T> t2 = z2 + z3;
        Cx<T> t3 = z1 - z4;
        Cx<T> t4 = z2 - z3;
        Cx<T> t5 = t1 + t2;
        Cx<T> t6 = scale(t1 - t2, T(0.25) * std::sqrt(T(5.0)));
        Cx<T> t7 = z0 - scale(t5, T(0.25));
        Cx<T> t8 = t6 + t7;
        Cx<T> t9 = t7 - t6;
        Cx<T> t10 = turn(scale(t3, std::sin(T(-0.4) * pi)) + scale(t4, std::sin(T(-0.2) * pi)));
        Cx<T> t11 = turn(scale(t3, std::sin(T(-0.2) * pi)) - scale(t4, std::sin(T(-0.4) * pi)));
        ret[j] = z0 + t5;
        ret[pos2] = t8 + t10;
        ret[pos3] = t9 + t11;
        ret[pos4] = t9 - t11;
        ret[pos5] = t8 - t10;
      }
    }
  }
}

// mixed_radix.rs:336-368 (generic O(r^2) butterfly; radix 7 under MAX_PRIME = 7)
template <class T>
void radix_other(Cx<T>* ret, size_t value, size_t count, size_t& po2, size_t& rad, const Cx<T>* omega,
                 size_t len) {
  size_t rot_width = len / value;
  std::vector<Cx<T>> rot(value), z(value);
  std::vector<size_t> pos(value);
  for (size_t i = 0; i < value; i++) rot[i] = omega[rot_width * i];
  for (size_t c = 0; c < count; c++) {
    size_t po2m = po2;
    po2 *= value;
    rad /= value;
    for (size_t j0 = 0; j0 < po2m; j0++) {
      size_t wpos = rad * j0;
      for (size_t j = j0; j < len; j += po2) {
        for (size_t i = 0; i < value; i++) pos[i] = j + po2m * i;
        for (size_t i = 0; i < value; i++) z[i] = ret[pos[i]] * omega[wpos * i];
        for (size_t i = 0; i < value; i++) {
          Cx<T> x = z[0];
          for (size_t l = 1; l < value; l++) x = x + z[l] * rot[(i * l) % value];
          ret[pos[i]] = x;
        }
      }
    }
  }
}

// mixed_radix.rs:104-166
template <class T, class F>
void fft_kernel(Cx<T>* src, size_t len, const Cx<T>* omega, const std::vector<Factor>& factors, F turn) {
  size_t po2 = 1, rad = len;
  for (const Factor& f : factors) {
    switch (f.value) {
      case 2: radix2(src, f.count, po2, rad, omega, len); break;
      case 3: radix3(src, f.count, po2, rad, omega, len, turn); break;
      case 4: radix4(src, f.count, po2, rad, omega, len, turn); break;
      case 5: radix5(src, f.count, po2, rad, omega, len, turn); break;
      default: radix_other(src, f.value, f.count, po2, rad, omega, len); break;
    }
  }
}

template <class T>
struct MixedRadixData {
  std::vector<size_t> ids;
  std::vector<Cx<T>> omega, omega_back;
  std::vector<Factor> factors;
  std::vector<size_t> ids_inplace;  // built lazily (cfft1d.rs:244-247)
};

template <class T>
struct ChirpzData {
  size_t level = 0;
  std::vector<size_t> ids;
  std::vector<Cx<T>> omega, omega_back, src_omega, rot_conj, rot_ft;
  T pow2len_inv = 0;
};

// ---------------------------------------------------------------- chirpz.rs
// chirpz.rs:27-85
template <class T>
void convert_rad2_inplace(Cx<T>* source, size_t len, size_t level, const std::vector<size_t>& ids,
                          const Cx<T>* omega, bool is_back, T pow2len_inv) {
  for (size_t i = 0; i < ids.size(); i++)
    if (i != ids[i]) std::swap(source[i], source[ids[i]]);
  if (is_back)
    for (size_t i = 0; i < len; i++) source[i] = scale(source[i], pow2len_inv);
  size_t po2 = 1, rad = len;
  if ((level & 1) == 1) {
    rad >>= 1;
    po2 = 2;
    for (size_t j = 0; j < rad; j++) {
      size_t a = j << 1, b = a + 1;
      Cx<T> wfb = source[b];
      source[b] = source[a] - wfb;
      source[a] = source[a] + wfb;
    }
  }
  if (is_back)
    radix4(source, level >> 1, po2, rad, omega, len, three_quarter_turn<T>);
  else
    radix4(source, level >> 1, po2, rad, omega, len, quarter_turn<T>);
}

// chirpz.rs:87-146 (out of place) and 148-203 (in place): identical arithmetic.
template <class T>
void convert_chirpz(const Cx<T>* source, Cx<T>* out, size_t srclen, bool is_back, T scaler,
                    const ChirpzData<T>& d) {
  size_t len = size_t(1) << d.level;
  size_t dlen = srclen << 1;
  std::vector<Cx<T>> a(len);
  for (size_t i = 0; i < srclen; i++) {
    Cx<T> w = d.src_omega[(uint64_t(i) * uint64_t(i)) % dlen];
    a[i] = source[i] * w;
  }
  convert_rad2_inplace(a.data(), len, d.level, d.ids, d.omega.data(), false, d.pow2len_inv);
  for (size_t i = 0; i < len; i++) a[i] = a[i] * d.rot_ft[i];
  convert_rad2_inplace(a.data(), len, d.level, d.ids, d.omega_back.data(), true, d.pow2len_inv);
  for (size_t k = 0; k < srclen; k++) {
    size_t i = (k == 0) ? 0 : (is_back ? srclen - k : k);
    out[k] = (scaler != T(1)) ? a[i] * scale(d.rot_conj[i], scaler) : a[i] * d.rot_conj[i];
  }
}

// ---------------------------------------------------------------- cfft1d.rs
template <class T>
struct CFft1D {
  size_t len = 0;
  T scaler_n = 0, scaler_u = 0;
  int kind = 0;  // 0 none, 1 mixed radix, 2 chirp-z
  MixedRadixData<T> mr;
  ChirpzData<T> cz;

  // cfft1d.rs:103-206
  void setup(size_t n) {
    const size_t MAX_PRIME = 7;
    len = n;
    scaler_n = T(1) / T(n);
    scaler_u = std::sqrt(scaler_n);
    std::vector<Factor> factors = prime_factorization(n, MAX_PRIME);
    if (factors.empty()) {
      size_t np2 = 1;
      while (np2 < n) np2 <<= 1;
      size_t pow2len = np2 << 1;
      size_t lv = 0;
      while ((size_t(1) << lv) < pow2len) lv++;
      size_t dlen = n << 1;
      std::vector<Cx<T>> src_omega = calc_omega<T>(dlen);
      std::vector<Cx<T>> rot, rot_conj;
      rot.reserve(pow2len);
      rot_conj.reserve(pow2len);
      for (size_t i = 0; i < n; i++) {
        size_t sq = (uint64_t(i) * uint64_t(i)) % dlen;
        rot_conj.push_back(src_omega[sq]);
        rot.push_back(src_omega[dlen - sq]);
      }
      size_t hlen = (pow2len >> 1) + 1;
      for (size_t i = n; i < hlen; i++) {
        rot_conj.push_back(Cx<T>());
        rot.push_back(Cx<T>());
      }
      for (size_t i = hlen; i < pow2len; i++) {
        rot_conj.push_back(rot_conj[pow2len - i]);
        rot.push_back(rot[pow2len - i]);
      }
      if (kind == 2 && cz.level == lv) {  // level reuse, cfft1d.rs:142-160
        cz.src_omega = std::move(src_omega);
        cz.rot_conj = std::move(rot_conj);
        convert_rad2_inplace(rot.data(), pow2len, lv, cz.ids, cz.omega.data(), false, cz.pow2len_inv);
        cz.rot_ft = std::move(rot);
        return;
      }
      std::vector<Factor> f2 = {{2, lv & 1}, {4, lv >> 1}};
      cz.ids = calc_bitreverse2inplace(calc_bitreverse(n, f2));
      cz.omega = calc_omega<T>(pow2len);
      cz.omega_back.assign(cz.omega.rbegin(), cz.omega.rend());
      cz.pow2len_inv = T(1) / T(pow2len);
      convert_rad2_inplace(rot.data(), pow2len, lv, cz.ids, cz.omega.data(), false, cz.pow2len_inv);
      cz.level = lv;
      cz.src_omega = std::move(src_omega);
      cz.rot_conj = std::move(rot_conj);
      cz.rot_ft = std::move(rot);
      kind = 2;
    } else {
      mr.omega = calc_omega<T>(n);
      mr.omega_back.assign(mr.omega.rbegin(), mr.omega.rend());
      mr.ids = calc_bitreverse(n, factors);
      mr.factors = factors;
      mr.ids_inplace.clear();
      kind = 1;
    }
  }

  // cfft1d.rs:208-230 + mixed_radix.rs:24-63
  void convert(const Cx<T>* src, Cx<T>* out, size_t n, bool is_back, T scaler) {
    if (n <= 1) {
      for (size_t i = 0; i < n; i++) out[i] = src[i];
      return;
    }
    if (n != len) setup(n);
    if (kind == 1) {
      std::vector<Cx<T>> ret(n);
      if (scaler != T(1))
        for (size_t i = 0; i < n; i++) ret[i] = scale(src[mr.ids[i]], scaler);
      else
        for (size_t i = 0; i < n; i++) ret[i] = src[mr.ids[i]];
      if (is_back)
        fft_kernel(ret.data(), n, mr.omega_back.data(), mr.factors, three_quarter_turn<T>);
      else
        fft_kernel(ret.data(), n, mr.omega.data(), mr.factors, quarter_turn<T>);
      std::copy(ret.begin(), ret.end(), out);
    } else {
      std::vector<Cx<T>> tmp(n);
      convert_chirpz(src, tmp.data(), n, is_back, scaler, cz);
      std::copy(tmp.begin(), tmp.end(), out);
    }
  }

  // cfft1d.rs:232-256 + mixed_radix.rs:65-102
  void convert_inplace(Cx<T>* src, size_t n, bool is_back, T scaler) {
    if (n <= 1) return;
    if (n != len) setup(n);
    if (kind == 1) {
      if (mr.ids_inplace.empty()) mr.ids_inplace = calc_bitreverse2inplace(mr.ids);
      for (size_t i = 0; i < n; i++)
        if (i != mr.ids_inplace[i]) std::swap(src[i], src[mr.ids_inplace[i]]);
      if (is_back)
        fft_kernel(src, n, mr.omega_back.data(), mr.factors, three_quarter_turn<T>);
      else
        fft_kernel(src, n, mr.omega.data(), mr.factors, quarter_turn<T>);
      if (scaler != T(1))
        for (size_t i = 0; i < n; i++) src[i] = scale(src[i], scaler);
    } else {
      std::vector<Cx<T>> tmp(n);
      convert_chirpz(src, tmp.data(), n, is_back, scaler, cz);
      std::copy(tmp.begin(), tmp.end(), src);
    }
  }
};

// Method codes shared with the product's C-ABI (include/chfft_hip.h).
enum Method {
  FORWARD = 0, FORWARD0 = 1, FORWARDU = 2, FORWARDN = 3,
  BACKWARD = 4, BACKWARD0 = 5, BACKWARDU = 6, BACKWARDN = 7,
};

// cfft1d.rs:270-455: scaling per method
template <class T>
bool c2c_method(const CFft1D<T>& f, int m, bool* back, T* s) {
  switch (m) {
    case FORWARD: case FORWARD0: *back = false; *s = T(1); return true;
    case FORWARDU: *back = false; *s = f.scaler_u; return true;
    case FORWARDN: *back = false; *s = f.scaler_n; return true;
    case BACKWARD: case BACKWARDN: *back = true; *s = f.scaler_n; return true;
    case BACKWARD0: *back = true; *s = T(1); return true;
    case BACKWARDU: *back = true; *s = f.scaler_u; return true;
  }
  return false;
}

// ---------------------------------------------------------------- rfft1d.rs
template <class T>
struct RFft1D {
  CFft1D<T> fft;
  size_t len = 0;
  std::vector<Cx<T>> coef, bcoef, work;
  T scaler_n = 0, scaler_u = 0;

  // rfft1d.rs:44-65, 73-113
  void setup(size_t n) {
    len = n;
    fft = CFft1D<T>();
    fft.setup(n >> 1);
    scaler_n = T(1) / T(n);
    scaler_u = std::sqrt(scaler_n);
    std::vector<Cx<T>> omega = calc_omega<T>(n);
    coef.clear();
    bcoef.clear();
    for (size_t i = 0; i < (n >> 1); i++)
      coef.push_back(scale(Cx<T>(1, 0) + times_i(omega[i]), T(0.5)));
    for (size_t i = 0; i < (n >> 1); i++)
      bcoef.push_back(scale(Cx<T>(1, 0) - times_i(omega[n - i]), T(0.5)));
    work.assign(n >> 1, Cx<T>());
  }

  // rfft1d.rs:115-151
  void convert(const T* source, Cx<T>* ret, T scaler) {
    size_t t = len >> 1;
    for (size_t i = 0; i < t; i++) work[i] = scale(Cx<T>(source[i << 1], source[(i << 1) + 1]), scaler);
    fft.convert_inplace(work.data(), t, false, T(1));
    size_t hlen = (len + 1) >> 1, qlen = (len + 3) >> 2;
    for (size_t i = 0; i <= t; i++) ret[i] = Cx<T>();
    for (size_t i = 1; i < qlen; i++) {
      Cx<T> x = coef[i] * (work[i] - conj(work[hlen - i]));
      ret[i] = work[i] - x;
      ret[hlen - i] = work[hlen - i] + conj(x);
    }
    ret[0] = Cx<T>(work[0].re + work[0].im, T(0));
    if ((len & 3) == 0) ret[qlen] = conj(work[qlen]);
    ret[hlen] = Cx<T>(work[0].re - work[0].im, T(0));
  }

  // rfft1d.rs:153-187
  void convert_back(const Cx<T>* source, T* ret, T scaler) {
    scaler = scaler * T(2.0);
    size_t hlen = (len + 1) >> 1, qlen = (len + 3) >> 2;
    work[0] = scale(Cx<T>(source[0].re + source[hlen].re, source[0].re - source[hlen].re), T(0.5));
    if ((len & 3) == 0) work[qlen] = conj(source[qlen]);
    for (size_t i = 1; i < qlen; i++) {
      Cx<T> x = bcoef[i] * (source[i] - conj(source[hlen - i]));
      work[i] = source[i] - x;
      work[hlen - i] = source[hlen - i] + conj(x);
    }
    fft.convert_inplace(work.data(), len >> 1, true, T(1));
    for (size_t i = 0; i < hlen; i++) {
      ret[2 * i] = work[i].re * scaler;
      ret[2 * i + 1] = work[i].im * scaler;
    }
  }
};

// ---------------------------------------------------------------- cfft2d.rs
// cfft2d.rs:326-363. Rows (length n) first, transposed scatter, columns
// (length m), scaler applied on the final scatter.
template <class T>
void cfft2d_convert(const Cx<T>* src, Cx<T>* out, size_t m, size_t n, bool is_back, T scaler) {
  CFft1D<T> fft_m, fft_n;
  std::vector<Cx<T>> work(n * m), row(std::max(m, n));
  for (size_t i = 0; i < m; i++) {
    fft_m.convert(src + i * n, row.data(), n, is_back, T(1));
    for (size_t j = 0; j < n; j++) work[j * m + i] = row[j];
  }
  for (size_t i = 0; i < n; i++) {
    fft_n.convert(work.data() + i * m, row.data(), m, is_back, T(1));
    for (size_t j = 0; j < m; j++) out[j * n + i] = scale(row[j], scaler);
  }
}

// ---------------------------------------------------------------- dct1d.rs
template <class T>
struct Dct2 {
  CFft1D<T> cfft;
  std::vector<Cx<T>> omega, coef, work;
  size_t len = 0;
  T ortho_scaler = 1, unitary_scaler = 1;

  // dct1d.rs:200-226
  void setup(size_t n, bool is_ortho) {
    if (n != len) {
      cfft = CFft1D<T>();
      cfft.setup(n >> 1);
      std::vector<Cx<T>> om = calc_omega<T>(n << 2);
      omega.clear();
      for (size_t i = 0; i < (n >> 1) + 1; i++) omega.push_back(om[om.size() - 1 - i]);
      std::vector<Cx<T>> om2 = calc_omega<T>(n);
      coef.clear();
      for (size_t i = 0; i < (n >> 1); i++) coef.push_back(scale(Cx<T>(1, 0) + times_i(om2[i]), T(0.5)));
      work.assign(n >> 1, Cx<T>());
      len = n;
      unitary_scaler = std::sqrt(T(2.0) / T(n));
    }
    ortho_scaler = is_ortho ? std::sqrt(T(0.5)) : T(1);
  }

  // dct1d.rs:236-290
  void convert(const T* source, T* ret, T scaler) {
    size_t hlen = len >> 1, qlen = hlen >> 1;
    if (qlen == 0) {
      work[0] = scale(Cx<T>(source[0] * ortho_scaler, source[1]), scaler);
    } else {
      work[0] = scale(Cx<T>(source[0] * ortho_scaler, source[2]), scaler);
      for (size_t i = 1; i < qlen; i++) work[i] = scale(Cx<T>(source[i << 2], source[(i << 2) + 2]), scaler);
      for (size_t i = 0; i < qlen; i++)
        work[hlen - i - 1] = scale(Cx<T>(source[(i << 2) + 3], source[(i << 2) + 1]), scaler);
      if (hlen & 1) work[qlen] = scale(Cx<T>(source[len - 2], source[len - 1]), scaler);
    }
    cfft.convert_inplace(work.data(), hlen, false, T(1));
    for (size_t i = 0; i < len; i++) ret[i] = T(0);
    size_t q2 = (len + 3) >> 2;
    for (size_t i = 1; i < q2; i++) {
      Cx<T> wconj = conj(work[hlen - i]);
      Cx<T> x = coef[i] * (work[i] - wconj);
      Cx<T> j = conj(work[i] - x) * omega[i];
      ret[i] = j.re;
      ret[len - i] = j.im;
      Cx<T> k = (wconj + x) * omega[hlen - i];
      ret[hlen - i] = k.re;
      ret[hlen + i] = k.im;
    }
    if ((len & 3) == 0) {
      Cx<T> x = work[q2] * omega[q2];
      ret[q2] = x.re;
      ret[len - q2] = x.im;
    }
    ret[0] = work[0].re + work[0].im;
    ret[hlen] = (work[0].re - work[0].im) * omega[hlen].re;
  }
};

template <class T>
struct Dct3 {
  CFft1D<T> cfft;
  std::vector<Cx<T>> omega, coef, work;
  size_t len = 0;
  T ortho_scaler = 0.5, unitary_scaler = 1;

  // dct1d.rs:323-349
  void setup(size_t n, bool is_ortho) {
    if (n != len) {
      cfft = CFft1D<T>();
      cfft.setup(n >> 1);
      std::vector<Cx<T>> om = calc_omega<T>(n << 2);
      omega.assign(om.begin(), om.begin() + (n >> 1) + 1);
      std::vector<Cx<T>> om2 = calc_omega<T>(n);
      coef.clear();
      for (size_t i = 0; i < (n >> 1); i++)
        coef.push_back(scale(Cx<T>(1, 0) - times_i(om2[om2.size() - 1 - i]), T(0.5)));
      work.assign(n >> 1, Cx<T>());
      len = n;
      unitary_scaler = std::sqrt(T(2.0) / T(n));
    }
    ortho_scaler = is_ortho ? std::sqrt(T(0.5)) : T(0.5);
  }

  // dct1d.rs:359-409
  void convert(const T* source, T* ret, T scaler) {
    size_t hlen = len >> 1;
    T zval = source[0] * ortho_scaler;
    T hval = source[hlen] * omega[hlen].re;
    work[0] = scale(Cx<T>(zval + hval, zval - hval), scaler);
    size_t qlen = (len + 3) >> 2;
    if ((len & 3) == 0) work[qlen] = Cx<T>(source[qlen], source[len - qlen]) * scale(omega[qlen], scaler);
    for (size_t i = 1; i < qlen; i++) {
      Cx<T> j = conj(scale(Cx<T>(source[i], source[len - i]) * omega[i], scaler));
      Cx<T> k = scale(Cx<T>(source[hlen - i], source[hlen + i]) * omega[hlen - i], scaler);
      Cx<T> x = coef[i] * (j - k);
      work[i] = j - x;
      work[hlen - i] = conj(k) + conj(x);
    }
    cfft.convert_inplace(work.data(), hlen, true, T(1));
    size_t q2 = hlen >> 1, o = 0;
    for (size_t i = 0; i < q2; i++) {
      ret[o++] = work[i].re;
      ret[o++] = work[hlen - i - 1].im;
      ret[o++] = work[i].im;
      ret[o++] = work[hlen - i - 1].re;
    }
    if (hlen & 1) {
      ret[o++] = work[q2].re;
      ret[o++] = work[q2].im;
    }
  }
};

// ---------------------------------------------------------------- mdct1d.rs
template <class T>
struct Mdct {
  CFft1D<T> fft;
  size_t len = 0;
  std::vector<Cx<T>> twiddle, work;
  T scaler_u = 0, scaler_ui = 0;
  std::vector<T> window;

  // mdct1d.rs:78-89, 101-137. window: n entries (evaluated by the caller,
  // mdct1d.rs:91-93).
  void setup(size_t n, const T* win) {
    len = n;
    fft = CFft1D<T>();
    fft.setup(n >> 2);
    scaler_u = T(1) / std::sqrt(T(n >> 1));
    scaler_ui = std::sqrt(T(n >> 1));
    T sc = std::sqrt(std::sqrt(T(2.0) / T(n)));
    twiddle.clear();
    for (size_t i = 0; i < (n >> 2); i++) twiddle.push_back(scale(calc_omega_item<T>(n << 3, (i << 3) + 1), sc));
    work.assign(n >> 2, Cx<T>());
    window.assign(win, win + n);
  }

  // mdct1d.rs:197-242
  void convert(const T* source, T* ret, T scaler) {
    const T* w = window.data();
    size_t lenh = len >> 1, lenq = lenh >> 1, leno = lenq - (lenq >> 1), len3q = lenh + lenq,
           len5q = lenh + len3q;
    for (size_t i = 0; i < leno; i++) {
      size_t n = i << 1;
      work[i] = Cx<T>(source[len3q - 1 - n] * w[len3q - 1 - n] + source[len3q + n] * w[len3q + n],
                      source[lenq + n] * w[lenq + n] - source[lenq - 1 - n] * w[lenq - 1 - n]) *
                twiddle[i];
    }
    for (size_t i = leno; i < lenq; i++) {
      size_t n = i << 1;
      work[i] = Cx<T>(source[len3q - 1 - n] * w[len3q - 1 - n] - source[n - lenq] * w[n - lenq],
                      source[lenq + n] * w[lenq + n] + source[len5q - 1 - n] * w[len5q - 1 - n]) *
                twiddle[i];
    }
    fft.convert_inplace(work.data(), lenq, false, T(1));
    for (size_t i = 0; i < lenq; i++) {
      size_t n = i << 1;
      Cx<T> t = work[i] * scale(twiddle[i], scaler);
      ret[n] = -t.re;
      ret[lenh - 1 - n] = t.im;
    }
  }

  // mdct1d.rs:244-291
  void convert_back(const T* source, T* ret, T scaler) {
    const T* w = window.data();
    size_t lenh = len >> 1, lenq = lenh >> 1, leno = lenq - (lenq >> 1), len3q = lenh + lenq,
           len5q = lenh + len3q;
    for (size_t i = 0; i < lenq; i++) {
      size_t n = i << 1;
      work[i] = Cx<T>(source[n], source[lenh - 1 - n]) * scale(twiddle[i], T(-2.0));
    }
    fft.convert_inplace(work.data(), lenq, false, T(1));
    for (size_t i = 0; i < len; i++) ret[i] = T(0);
    for (size_t i = 0; i < leno; i++) {
      size_t n = i << 1;
      Cx<T> t = work[i] * scale(twiddle[i], scaler);
      ret[len3q - 1 - n] = t.re * w[len3q - 1 - n];
      ret[len3q + n] = t.re * w[len3q + n];
      ret[lenq + n] = -t.im * w[lenq + n];
      ret[lenq - 1 - n] = t.im * w[lenq - 1 - n];
    }
    for (size_t i = leno; i < lenq; i++) {
      size_t n = i << 1;
      Cx<T> t = work[i] * scale(twiddle[i], scaler);
      ret[len3q - 1 - n] = t.re * w[len3q - 1 - n];
      ret[n - lenq] = -t.re * w[n - lenq];
      ret[lenq + n] = -t.im * w[lenq + n];
      ret[len5q - 1 - n] = -t.im * w[len5q - 1 - n];
    }
  }
};

// mdct1d.rs:42-53
template <class T> T sine_window(size_t l, size_t p) {
  return std::sin(Pi<T>::v() * (T(p) + T(0.5)) / T(l));
}
template <class T> T vorbis_window(size_t l, size_t p) {
  T s = std::sin(Pi<T>::v() * (T(p) + T(0.5)) / T(l));
  return std::sin(s * s * Pi<T>::v() * T(0.5));
}

// Split [0, count) over `threads` workers (the CPU baseline: one transform per
// worker at a time, single-threaded per transform like chfft).
template <class F>
void parallel_for(size_t count, int threads, F f) {
  if (threads <= 1 || count <= 1) {
    f(size_t(0), count);
    return;
  }
  std::vector<std::thread> ts;
  size_t per = (count + threads - 1) / threads;
  for (int t = 0; t < threads; t++) {
    size_t b = t * per, e = std::min(count, b + per);
    if (b >= e) break;
    ts.emplace_back([=] { f(b, e); });
  }
  for (auto& th : ts) th.join();
}

}  // namespace oracle

using namespace oracle;

// ================================================================ C entry points
// All buffers are host memory, batch-major, complex interleaved {re, im}.
// Return 0 on success, -1 on invalid arguments (mirrors the reference panics).
extern "C" {

#define ORACLE_C2C(T, SUF)                                                                         \
  int oracle_c2c_##SUF(const T* in, T* out, size_t n, size_t batch, int method, int inplace,      \
                       int threads) {                                                              \
    if (n == 0) return -1;                                                                         \
    CFft1D<T> probe;                                                                               \
    probe.setup(n);                                                                                \
    bool back;                                                                                     \
    T s;                                                                                           \
    if (!c2c_method(probe, method, &back, &s)) return -1;                                          \
    parallel_for(batch, threads, [&](size_t b0, size_t b1) {                                       \
      CFft1D<T> f;                                                                                 \
      f.setup(n);                                                                                  \
      for (size_t b = b0; b < b1; b++) {                                                           \
        const Cx<T>* src = reinterpret_cast<const Cx<T>*>(in) + b * n;                             \
        Cx<T>* dst = reinterpret_cast<Cx<T>*>(out) + b * n;                                        \
        if (inplace) {                                                                             \
          std::copy(src, src + n, dst);                                                            \
          f.convert_inplace(dst, n, back, s);                                                      \
        } else {                                                                                   \
          f.convert(src, dst, n, back, s);                                                         \
        }                                                                                          \
      }                                                                                            \
    });                                                                                            \
    return 0;                                                                                      \
  }
ORACLE_C2C(double, f64)
ORACLE_C2C(float, f32)

// Plan reuse across lengths (CFft1D::new + a sequence of lengths, replanning on
// each change: cfft1d.rs:216-218, including the Bluestein level reuse 142-160).
int oracle_c2c_sequence_f64(const double* in, double* out, const size_t* lens, size_t count, int method) {
  CFft1D<double> f;
  size_t off = 0;
  for (size_t c = 0; c < count; c++) {
    size_t n = lens[c];
    if (n == 0) return -1;
    bool back;
    double s;
    if (n > 1 && n != f.len) f.setup(n);
    CFft1D<double> probe;
    probe.setup(n);
    if (!c2c_method(probe, method, &back, &s)) return -1;
    f.convert(reinterpret_cast<const Cx<double>*>(in) + off, reinterpret_cast<Cx<double>*>(out) + off, n, back,
              s);
    off += n;
  }
  return 0;
}

// RFft1D (rfft1d.rs:199-312). Forward: in n reals, out n/2+1 complex. Backward: reverse.
#define ORACLE_R2C(T, SUF)                                                                         \
  int oracle_rfft_##SUF(const T* in, T* out, size_t n, size_t batch, int method, int threads) {    \
    if (n == 0 || (n & 1)) return -1;                                                              \
    RFft1D<T> probe;                                                                               \
    probe.setup(n);                                                                                \
    T s;                                                                                           \
    bool back;                                                                                     \
    switch (method) {                                                                              \
      case FORWARD: case FORWARD0: s = T(1); back = false; break;                                  \
      case FORWARDU: s = probe.scaler_u; back = false; break;                                      \
      case FORWARDN: s = probe.scaler_n; back = false; break;                                      \
      case BACKWARD: case BACKWARDN: s = probe.scaler_n; back = true; break;                       \
      case BACKWARD0: s = T(1); back = true; break;                                                \
      case BACKWARDU: s = probe.scaler_u; back = true; break;                                      \
      default: return -1;                                                                          \
    }                                                                                              \
    size_t h = n / 2 + 1;                                                                          \
    parallel_for(batch, threads, [&](size_t b0, size_t b1) {                                       \
      RFft1D<T> f;                                                                                 \
      f.setup(n);                                                                                  \
      for (size_t b = b0; b < b1; b++) {                                                           \
        if (!back)                                                                                 \
          f.convert(in + b * n, reinterpret_cast<Cx<T>*>(out) + b * h, s);                         \
        else                                                                                       \
          f.convert_back(reinterpret_cast<const Cx<T>*>(in) + b * h, out + b * n, s);              \
      }                                                                                            \
    });                                                                                            \
    return 0;                                                                                      \
  }
ORACLE_R2C(double, f64)
ORACLE_R2C(float, f32)

// CFft2D (cfft2d.rs:150-324): scalings 1/(mn), 1/sqrt(mn).
#define ORACLE_2D(T, SUF)                                                                          \
  int oracle_c2c2d_##SUF(const T* in, T* out, size_t m, size_t n, size_t batch, int method,       \
                         int threads) {                                                            \
    if (m == 0 || n == 0) return 0;                                                                \
    T sn = T(1) / T(m * n);                                                                        \
    T su = T(1) / std::sqrt(T(m * n));                                                             \
    T s;                                                                                           \
    bool back;                                                                                     \
    switch (method) {                                                                              \
      case FORWARD: case FORWARD0: s = T(1); back = false; break;                                  \
      case FORWARDU: s = su; back = false; break;                                                  \
      case FORWARDN: s = sn; back = false; break;                                                  \
      case BACKWARD: s = sn; back = true; break;                                                   \
      case BACKWARD0: s = T(1); back = true; break;                                                \
      case BACKWARDU: s = su; back = true; break;                                                  \
      default: return -1;                                                                          \
    }                                                                                              \
    parallel_for(batch, threads, [&](size_t b0, size_t b1) {                                       \
      for (size_t b = b0; b < b1; b++)                                                             \
        cfft2d_convert(reinterpret_cast<const Cx<T>*>(in) + b * m * n,                             \
                       reinterpret_cast<Cx<T>*>(out) + b * m * n, m, n, back, s);                  \
    });                                                                                            \
    return 0;                                                                                      \
  }
ORACLE_2D(double, f64)
ORACLE_2D(float, f32)

// Dct1D (dct1d.rs:98-144). dct_type 2 or 3. Methods: FORWARD/FORWARD0 (1),
// FORWARDU (sqrt(2/n)), FORWARDN (2/n).
#define ORACLE_DCT(T, SUF)                                                                         \
  int oracle_dct_##SUF(const T* in, T* out, size_t n, size_t batch, int dct_type, int is_ortho,    \
                       int method, int threads) {                                                  \
    if (n == 0 || (n & 1)) return -1;                                                              \
    if (dct_type != 2 && dct_type != 3) return -1;                                                 \
    T us = std::sqrt(T(2.0) / T(n));                                                               \
    T s;                                                                                           \
    switch (method) {                                                                              \
      case FORWARD: case FORWARD0: s = T(1); break;                                                \
      case FORWARDU: s = us; break;                                                                \
      case FORWARDN: s = us * us; break;                                                           \
      default: return -1;                                                                          \
    }                                                                                              \
    parallel_for(batch, threads, [&](size_t b0, size_t b1) {                                       \
      if (dct_type == 2) {                                                                         \
        Dct2<T> d;                                                                                 \
        d.setup(n, is_ortho != 0);                                                                 \
        for (size_t b = b0; b < b1; b++) d.convert(in + b * n, out + b * n, s);                    \
      } else {                                                                                     \
        Dct3<T> d;                                                                                 \
        d.setup(n, is_ortho != 0);                                                                 \
        for (size_t b = b0; b < b1; b++) d.convert(in + b * n, out + b * n, s);                    \
      }                                                                                            \
    });                                                                                            \
    return 0;                                                                                      \
  }
ORACLE_DCT(double, f64)
ORACLE_DCT(float, f32)

// Mdct1D (mdct1d.rs:149-195). window: n entries. Forward: frames of n samples
// read at `in_stride` apart (n for independent frames, n/2 for a 50 % overlap
// signal), n/2 coefficients out. Backward: n/2 in, n windowed samples out.
// Methods: FORWARD (sqrt(n/2)), FORWARDU (1), BACKWARD (1/sqrt(n/2)), BACKWARDU (1).
#define ORACLE_MDCT(T, SUF)                                                                        \
  int oracle_mdct_##SUF(const T* in, T* out, size_t n, size_t batch, const T* window,              \
                        size_t in_stride, int method, int threads) {                               \
    if (n == 0 || (n & 3)) return -1;                                                              \
    T s;                                                                                           \
    bool back;                                                                                     \
    switch (method) {                                                                              \
      case FORWARD: s = std::sqrt(T(n >> 1)); back = false; break;                                 \
      case FORWARDU: s = T(1); back = false; break;                                                \
      case BACKWARD: s = T(1) / std::sqrt(T(n >> 1)); back = true; break;                          \
      case BACKWARDU: s = T(1); back = true; break;                                                \
      default: return -1;                                                                          \
    }                                                                                              \
    parallel_for(batch, threads, [&](size_t b0, size_t b1) {                                       \
      Mdct<T> d;                                                                                   \
      d.setup(n, window);                                                                          \
      for (size_t b = b0; b < b1; b++) {                                                           \
        if (!back)                                                                                 \
          d.convert(in + b * in_stride, out + b * (n / 2), s);                                     \
        else                                                                                       \
          d.convert_back(in + b * in_stride, out + b * n, s);                                      \
      }                                                                                            \
    });                                                                                            \
    return 0;                                                                                      \
  }
ORACLE_MDCT(double, f64)
ORACLE_MDCT(float, f32)

// Window tables exactly as mdct1d.rs:42-53 computes them (kind 0 sine, 1 vorbis).
int oracle_window_f64(size_t n, int kind, double* out) {
  for (size_t i = 0; i < n; i++) out[i] = kind == 0 ? sine_window<double>(n, i) : vorbis_window<double>(n, i);
  return 0;
}
int oracle_window_f32(size_t n, int kind, float* out) {
  for (size_t i = 0; i < n; i++) out[i] = kind == 0 ? sine_window<float>(n, i) : vorbis_window<float>(n, i);
  return 0;
}

// Plan internals exposed for pinning tests.
int oracle_factors(size_t n, size_t max, size_t* values, size_t* counts, size_t cap) {
  std::vector<Factor> f = prime_factorization(n, max);
  if (f.size() > cap) return -1;
  for (size_t i = 0; i < f.size(); i++) {
    values[i] = f[i].value;
    counts[i] = f[i].count;
  }
  return int(f.size());
}
int oracle_omega_f64(size_t len, double* out) {
  std::vector<Cx<double>> w = calc_omega<double>(len);
  std::memcpy(out, w.data(), w.size() * sizeof(Cx<double>));
  return 0;
}
int oracle_bitreverse(size_t len, const size_t* values, const size_t* counts, size_t nf, size_t* out,
                      int inplace) {
  std::vector<Factor> f;
  for (size_t i = 0; i < nf; i++) f.push_back({values[i], counts[i]});
  std::vector<size_t> ids = calc_bitreverse(len, f);
  if (ids.size() != len) return -1;
  if (inplace) ids = calc_bitreverse2inplace(ids);
  std::copy(ids.begin(), ids.end(), out);
  return 0;
}

// Long-double brute-force definitions (the reference tests' oracles:
// cfft1d.rs:481-494, dct1d.rs:423-434, mdct1d.rs:305-339) used to pin the
// restatement. sign = -1 forward, +1 backward.
int oracle_dft_brute(const double* in, double* out, size_t n, int sign) {
  const long double pi = 3.141592653589793238462643383279502884L;
  for (size_t k = 0; k < n; k++) {
    long double sr = 0, si = 0;
    for (size_t j = 0; j < n; j++) {
      unsigned long long m = (unsigned long long)((unsigned __int128)j * k % n);
      long double th = sign * 2.0L * pi * (long double)m / (long double)n;
      long double c = cosl(th), s = sinl(th);
      sr += in[2 * j] * c - in[2 * j + 1] * s;
      si += in[2 * j] * s + in[2 * j + 1] * c;
    }
    out[2 * k] = (double)sr;
    out[2 * k + 1] = (double)si;
  }
  return 0;
}
int oracle_dct2_brute(const double* in, double* out, size_t n) {
  const long double pi = 3.141592653589793238462643383279502884L;
  for (size_t k = 0; k < n; k++) {
    long double s = 0;
    for (size_t j = 0; j < n; j++) {
      unsigned long long m = (unsigned long long)(((2 * j + 1) * k) % (4 * n));
      s += in[j] * cosl(pi * (long double)m / (long double)(2 * n));
    }
    out[k] = (double)s;
  }
  return 0;
}
int oracle_mdct_brute(const double* in, double* out, size_t n, const double* window) {
  const long double pi = 3.141592653589793238462643383279502884L;
  for (size_t m = 0; m < n / 2; m++) {
    long double s = 0;
    for (size_t k = 0; k < n; k++) {
      unsigned long long a = (unsigned long long)(((1 + (n >> 1) + (k << 1)) * (1 + (m << 1))) % (4 * n));
      s += window[k] * in[k] * cosl(pi / (long double)(2 * n) * (long double)a);
    }
    out[m] = (double)s;
  }
  return 0;
}
int oracle_imdct_brute(const double* in, double* out, size_t n, const double* window) {
  const long double pi = 3.141592653589793238462643383279502884L;
  for (size_t k = 0; k < n; k++) {
    long double s = 0;
    for (size_t m = 0; m < n / 2; m++) {
      unsigned long long a = (unsigned long long)(((1 + (n >> 1) + (k << 1)) * (1 + (m << 1))) % (4 * n));
      s += in[m] * cosl(pi / (long double)(2 * n) * (long double)a);
    }
    out[k] = (double)(4.0L * window[k] / (long double)n * s);
  }
  return 0;
}

}  // extern "C"

// CPU baseline timer: `reps` passes of a c2c method over the batch on
// `threads` workers. Plans are built before the clock starts (chfft plans once
// with with_len); each pass writes the full output. Returns the best seconds
// per pass, or -1 on bad arguments.
template <class T>
static double time_c2c(const T* in, T* out, size_t n, size_t batch, int method, int threads, int reps) {
  using clk = std::chrono::steady_clock;
  CFft1D<T> probe;
  probe.setup(n);
  bool back;
  T s;
  if (!c2c_method(probe, method, &back, &s)) return -1.0;
  if (threads < 1) threads = 1;
  std::vector<CFft1D<T>> plans(threads);
  for (auto& p : plans) p.setup(n);
  double best = 1e30;
  for (int r = 0; r < reps; r++) {
    auto t0 = clk::now();
    std::vector<std::thread> ts;
    size_t per = (batch + threads - 1) / threads;
    for (int t = 0; t < threads; t++) {
      size_t b0 = t * per, b1 = std::min(batch, b0 + per);
      if (b0 >= b1) break;
      ts.emplace_back([&, t, b0, b1] {
        for (size_t b = b0; b < b1; b++)
          plans[t].convert(reinterpret_cast<const Cx<T>*>(in) + b * n, reinterpret_cast<Cx<T>*>(out) + b * n, n,
                           back, s);
      });
    }
    for (auto& th : ts) th.join();
    auto t1 = clk::now();
    best = std::min(best, std::chrono::duration<double>(t1 - t0).count());
  }
  return best;
}

// CPU baseline timer for the other transform types: `reps` passes of
// run(plan, b) over b in [0, batch) on `threads` workers, each worker with its
// own plan made by make() before the clock starts. Best seconds per pass.
template <class P, class Make, class Run>
static double time_batch(size_t batch, int threads, int reps, Make make, Run run) {
  using clk = std::chrono::steady_clock;
  if (threads < 1) threads = 1;
  std::vector<P> plans;
  plans.reserve(threads);
  for (int t = 0; t < threads; t++) plans.push_back(make());
  double best = 1e30;
  for (int r = 0; r < reps; r++) {
    auto t0 = clk::now();
    std::vector<std::thread> ts;
    size_t per = (batch + threads - 1) / threads;
    for (int t = 0; t < threads; t++) {
      size_t b0 = t * per, b1 = std::min(batch, b0 + per);
      if (b0 >= b1) break;
      ts.emplace_back([&, t, b0, b1] {
        for (size_t b = b0; b < b1; b++) run(plans[t], b);
      });
    }
    for (auto& th : ts) th.join();
    auto t1 = clk::now();
    best = std::min(best, std::chrono::duration<double>(t1 - t0).count());
  }
  return best;
}

template <class T>
static double time_op(int kind, const T* in, T* out, size_t n, size_t m, size_t batch, size_t stride,
                      const T* window, int threads, int reps) {
  switch (kind) {
    case 1:  // RFft1D::forward, n reals -> n/2 + 1 complex
      return time_batch<RFft1D<T>>(batch, threads, reps, [&] { RFft1D<T> f; f.setup(n); return f; },
                                   [&](RFft1D<T>& f, size_t b) {
                                     f.convert(in + b * n, reinterpret_cast<Cx<T>*>(out) + b * (n / 2 + 1), T(1));
                                   });
    case 2:  // Dct1D type II forward
      return time_batch<Dct2<T>>(batch, threads, reps, [&] { Dct2<T> d; d.setup(n, false); return d; },
                                 [&](Dct2<T>& d, size_t b) { d.convert(in + b * n, out + b * n, T(1)); });
    case 3:  // Dct1D type III forward
      return time_batch<Dct3<T>>(batch, threads, reps, [&] { Dct3<T> d; d.setup(n, false); return d; },
                                 [&](Dct3<T>& d, size_t b) { d.convert(in + b * n, out + b * n, T(1)); });
    case 4:  // Mdct1D::forward of frames `stride` samples apart
      if (!window) return -1.0;
      return time_batch<Mdct<T>>(batch, threads, reps, [&] { Mdct<T> d; d.setup(n, window); return d; },
                                 [&](Mdct<T>& d, size_t b) {
                                   d.convert(in + b * stride, out + b * (n / 2), std::sqrt(T(n >> 1)));
                                 });
    case 5:  // CFft2D::forward of m x n matrices (the 2-D convert builds its plans per call)
      return time_batch<int>(batch, threads, reps, [] { return 0; }, [&](int&, size_t b) {
        cfft2d_convert(reinterpret_cast<const Cx<T>*>(in) + b * m * n, reinterpret_cast<Cx<T>*>(out) + b * m * n, m,
                       n, false, T(1));
      });
  }
  return -1.0;
}

// CPU baseline timer of a batch of uneven lengths (CFft1D::forward of each
// transform, packed back to back): one plan per transform, built on the
// workers before the clock starts; transforms are spread over the workers
// longest first (greedy by n log n). Best seconds per pass.
template <class T>
static double time_lens(const T* in, T* out, const size_t* lens, size_t count, int threads, int reps) {
  using clk = std::chrono::steady_clock;
  if (threads < 1) threads = 1;
  std::vector<size_t> off(count + 1, 0), order(count);
  for (size_t i = 0; i < count; i++) {
    off[i + 1] = off[i] + lens[i];
    order[i] = i;
  }
  auto cost = [&](size_t i) { return (double)lens[i] * std::log2((double)lens[i] + 1.0); };
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost(a) > cost(b); });
  std::vector<std::vector<size_t>> part(threads);
  std::vector<double> load(threads, 0.0);
  for (size_t i : order) {
    int t = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    part[t].push_back(i);
    load[t] += cost(i);
  }
  std::vector<CFft1D<T>> plans(count);
  parallel_for((size_t)threads, threads, [&](size_t t0, size_t t1) {
    for (size_t t = t0; t < t1; t++)
      for (size_t i : part[t]) plans[i].setup(lens[i]);
  });
  double best = 1e30;
  for (int r = 0; r < reps; r++) {
    auto a = clk::now();
    parallel_for((size_t)threads, threads, [&](size_t t0, size_t t1) {
      for (size_t t = t0; t < t1; t++)
        for (size_t i : part[t])
          plans[i].convert(reinterpret_cast<const Cx<T>*>(in) + off[i], reinterpret_cast<Cx<T>*>(out) + off[i],
                           lens[i], false, T(1));
    });
    auto b = clk::now();
    best = std::min(best, std::chrono::duration<double>(b - a).count());
  }
  return best;
}

extern "C" {
double oracle_time_c2c(int f64, const void* in, void* out, size_t n, size_t batch, int method, int threads,
                       int reps) {
  if (n < 2) return -1.0;
  return f64 ? time_c2c((const double*)in, (double*)out, n, batch, method, threads, reps)
             : time_c2c((const float*)in, (float*)out, n, batch, method, threads, reps);
}

// kind: 1 RFft1D forward, 2 DCT-II, 3 DCT-III, 4 MDCT forward (frames
// `stride` samples apart; window of n entries in the same precision),
// 5 CFft2D forward (m x n).
double oracle_time_op(int kind, int f64, const void* in, void* out, size_t n, size_t m, size_t batch, size_t stride,
                      const void* window, int threads, int reps) {
  if (n < 2) return -1.0;
  return f64 ? time_op(kind, (const double*)in, (double*)out, n, m, batch, stride, (const double*)window, threads,
                       reps)
             : time_op(kind, (const float*)in, (float*)out, n, m, batch, stride, (const float*)window, threads, reps);
}

double oracle_time_lens(int f64, const void* in, void* out, const size_t* lens, size_t count, int threads,
                        int reps) {
  for (size_t i = 0; i < count; i++)
    if (lens[i] == 0) return -1.0;
  return f64 ? time_lens((const double*)in, (double*)out, lens, count, threads, reps)
             : time_lens((const float*)in, (float*)out, lens, count, threads, reps);
}

}  // extern "C"
