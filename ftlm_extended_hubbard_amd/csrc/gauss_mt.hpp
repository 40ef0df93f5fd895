// gauss_mt.hpp -- the reference's Gaussian start vectors drawn exactly on many host threads.
//
// The reference draws a start vector as std::mt19937 rng(seed) and
// std::normal_distribution<double>(0, 1), two draws (re, im) per entry in index order
// (src/lanczos.cpp:140-168).  That stream is one long sequence; this header splits it over T
// threads without changing a single value:
//
// 1. MT19937 jump-ahead (Haramoto, Matsumoto, Nishimura, Panneton, L'Ecuyer, "Efficient jump ahead
//    for F2-linear random number generators", INFORMS J. Computing 20 (2008)).  The generator is
//    linear over GF(2); its state transition F has characteristic polynomial t^31 phi(t) with phi
//    primitive of degree 19937 (the t^31 part is the 31 low bits of the oldest word, which the
//    recurrence never reads).  phi is found once per process by Berlekamp-Massey on the low bits
//    of 2 * 19937 outputs; F^J applied to a state is then g(F) state with g = t^J mod phi, evaluated
//    Horner-style by stepping a copy of the generator deg(g) times and adding (XOR) the states at
//    the coefficients that are 1.  Thread k of a round of T segments of kSegWords words starts at
//    word k * kSegWords of the round.
// 2. libstdc++'s normal_distribution is Marsaglia's polar method: an attempt reads 4 words (two
//    generate_canonical<double, 53> uniforms of two 32-bit words each), is accepted when
//    0 < x^2 + y^2 <= 1, and gives the pair (y m, x m), m = sqrt(-2 ln r2 / r2), the first returned
//    and the second saved.  Entry i of the reference vector is therefore exactly the i-th accepted
//    attempt of the word stream (attempts sit at word offsets 4a, since every attempt reads 4
//    words).  Each thread evaluates every attempt of its segment (polar_attempt, the library's
//    arithmetic in the library's order) and keeps the accepted pairs as (re, im).
// 3. The accepted entries of the T segments, in segment order, are the next entries of the vector.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

namespace ftlm_gauss {

constexpr int kN = 624, kM = 397, kDeg = 19937;
constexpr uint32_t kUpper = 0x80000000u, kLower = 0x7fffffffu, kMatrixA = 0x9908b0dfu;

// MT19937: the std::mt19937 output sequence.  s[(h + j) % 624] holds x_{n+j}, j = 0..623, where
// n is the number of words produced; the next word is x_{n+624} = x_{n+397} ^ twist(x_n, x_{n+1}),
// and the output is its tempered value.
class Mt19937 {
 public:
  explicit Mt19937(uint32_t seed = 5489u) {
    s_[0] = seed;
    for (int i = 1; i < kN; ++i) s_[i] = 1812433253u * (s_[i - 1] ^ (s_[i - 1] >> 30)) + static_cast<uint32_t>(i);
    h_ = 0;
  }

  static uint32_t temper(uint32_t y) {
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
  }

  // Advance one word; returns the new (untempered) word.
  uint32_t step() {
    const int h1 = h_ + 1 == kN ? 0 : h_ + 1;
    const int hm = h_ + kM >= kN ? h_ + kM - kN : h_ + kM;
    const uint32_t y = (s_[h_] & kUpper) | (s_[h1] & kLower);
    const uint32_t z = s_[hm] ^ (y >> 1) ^ ((y & 1u) ? kMatrixA : 0u);
    s_[h_] = z;
    h_ = h1;
    return z;
  }

  uint32_t operator()() { return temper(step()); }

  // n output words; whole blocks of 624 take the in-place block update when the head is at 0.
  void fill(uint32_t* out, size_t n) {
    size_t k = 0;
    while (k < n && h_ != 0) out[k++] = temper(step());
    while (n - k >= static_cast<size_t>(kN)) {
      int i = 0;
      for (; i < kN - kM; ++i) {
        const uint32_t y = (s_[i] & kUpper) | (s_[i + 1] & kLower);
        s_[i] = s_[i + kM] ^ (y >> 1) ^ ((y & 1u) ? kMatrixA : 0u);
      }
      for (; i < kN - 1; ++i) {
        const uint32_t y = (s_[i] & kUpper) | (s_[i + 1] & kLower);
        s_[i] = s_[i + kM - kN] ^ (y >> 1) ^ ((y & 1u) ? kMatrixA : 0u);
      }
      const uint32_t y = (s_[kN - 1] & kUpper) | (s_[0] & kLower);
      s_[kN - 1] = s_[kM - 1] ^ (y >> 1) ^ ((y & 1u) ? kMatrixA : 0u);
      for (int j = 0; j < kN; ++j) out[k + j] = temper(s_[j]);
      k += kN;
    }
    while (k < n) out[k++] = temper(step());
  }

  // this = g(F) this for a polynomial g over GF(2) (bit i of g = coefficient of t^i).
  void apply_poly(const std::vector<uint64_t>& g) {
    int deg = -1;
    for (int w = static_cast<int>(g.size()) - 1; w >= 0 && deg < 0; --w)
      if (g[w]) deg = 64 * w + 63 - __builtin_clzll(g[w]);
    uint32_t acc[kN] = {};
    Mt19937 cur = *this;
    for (int i = 0; i <= deg; ++i) {
      if ((g[i >> 6] >> (i & 63)) & 1u) {
        const int a = kN - cur.h_;
        for (int j = 0; j < a; ++j) acc[j] ^= cur.s_[cur.h_ + j];
        for (int j = a; j < kN; ++j) acc[j] ^= cur.s_[j - a];
      }
      cur.step();
    }
    std::memcpy(s_, acc, sizeof acc);
    h_ = 0;
  }

 private:
  uint32_t s_[kN];
  int h_ = 0;
};

// ---------------------------------------------------------------- GF(2) polynomials
using Poly = std::vector<uint64_t>;  // bit i = coefficient of t^i

inline int poly_deg(const Poly& p) {
  for (int w = static_cast<int>(p.size()) - 1; w >= 0; --w)
    if (p[w]) return 64 * w + 63 - __builtin_clzll(p[w]);
  return -1;
}

// p ^= q << sh (p must hold the shifted bits).
inline void xor_shifted(Poly& p, const Poly& q, int sh) {
  const int ws = sh >> 6, bs = sh & 63;
  const int nq = static_cast<int>(q.size());
  for (int w = 0; w < nq; ++w) {
    if (!q[w]) continue;
    p[w + ws] ^= q[w] << bs;
    if (bs) p[w + ws + 1] ^= q[w] >> (64 - bs);
  }
}

// The characteristic polynomial phi of the MT19937 transition (degree 19937, leading bit set),
// from Berlekamp-Massey over the low output bits 64 .. 64 + 2 * 19937 of std::mt19937(5489).
inline Poly char_poly() {
  const int N = 2 * kDeg + 64;
  const int W = (N + 64) / 64 + 2;
  std::mt19937 e(5489u);
  e.discard(64);
  std::vector<uint8_t> s(static_cast<size_t>(N));
  for (int n = 0; n < N; ++n) s[n] = static_cast<uint8_t>(e() & 1u);
  // rev bit j = s[N - 1 - j]: the window s[n - 1], s[n - 2], ... is rev bits N - n, N - n + 1, ...
  Poly rev(static_cast<size_t>(W), 0);
  for (int j = 0; j < N; ++j)
    if (s[N - 1 - j]) rev[j >> 6] |= uint64_t{1} << (j & 63);
  auto rev_bits = [&](int o) -> uint64_t {  // 64 bits of rev from bit offset o
    const int w = o >> 6, b = o & 63;
    uint64_t v = rev[w] >> b;
    if (b && w + 1 < W) v |= rev[w + 1] << (64 - b);
    return v;
  };
  Poly C(static_cast<size_t>(W), 0), B(static_cast<size_t>(W), 0), T;
  C[0] = B[0] = 1;
  int L = 0, m = 1;
  for (int n = 0; n < N; ++n) {
    // d = s[n] + sum_{i=1..L} C_i s[n - i]; C_i pairs with rev bit (N - n - 1) + i
    uint64_t acc = 0;
    const int nw = L / 64 + 1;
    for (int w = 0; w < nw; ++w) {
      uint64_t c = C[w];
      if (w == 0) c &= ~uint64_t{1};
      if (!c) continue;
      acc ^= c & rev_bits(N - n - 1 + 64 * w);
    }
    const int d = s[n] ^ (__builtin_popcountll(acc) & 1);
    if (!d) {
      ++m;
    } else if (2 * L <= n) {
      T = C;
      xor_shifted(C, B, m);
      L = n + 1 - L;
      B = T;
      m = 1;
    } else {
      xor_shifted(C, B, m);
      ++m;
    }
  }
  // phi(t) = t^L C(1/t)
  Poly phi(static_cast<size_t>((L + 64) / 64 + 1), 0);
  for (int i = 0; i <= L; ++i)
    if ((C[i >> 6] >> (i & 63)) & 1u) phi[(L - i) >> 6] |= uint64_t{1} << ((L - i) & 63);
  return phi;
}

// p mod phi, p of degree < 2 * deg(phi).
inline void reduce(Poly& p, const Poly& phi, int dphi) {
  for (int k = poly_deg(p); k >= dphi; --k)
    if ((p[k >> 6] >> (k & 63)) & 1u) xor_shifted(p, phi, k - dphi);
}

inline Poly mulmod(const Poly& a, const Poly& b, const Poly& phi, int dphi) {
  Poly p(static_cast<size_t>(2 * (dphi / 64 + 2)), 0);
  const int da = poly_deg(a);
  for (int i = 0; i <= da; ++i)
    if ((a[i >> 6] >> (i & 63)) & 1u) xor_shifted(p, b, i);
  reduce(p, phi, dphi);
  p.resize(static_cast<size_t>(dphi / 64 + 2));
  return p;
}

inline Poly sqrmod(const Poly& a, const Poly& phi, int dphi) {
  Poly p(static_cast<size_t>(2 * (dphi / 64 + 2)), 0);
  for (size_t w = 0; w < a.size(); ++w)
    for (int b = 0; b < 64; ++b)
      if ((a[w] >> b) & 1u) {
        const size_t i = 2 * (64 * w + static_cast<size_t>(b));
        p[i >> 6] |= uint64_t{1} << (i & 63);
      }
  reduce(p, phi, dphi);
  p.resize(static_cast<size_t>(dphi / 64 + 2));
  return p;
}

// t^J mod phi.
inline Poly pow_t(uint64_t J, const Poly& phi, int dphi) {
  Poly r(static_cast<size_t>(dphi / 64 + 2), 0);
  r[0] = 1;
  for (int b = 63; b >= 0; --b) {
    r = sqrmod(r, phi, dphi);
    if ((J >> b) & 1u) {  // r *= t
      Poly p(r.size() + 1, 0);
      xor_shifted(p, r, 1);
      if ((p[dphi >> 6] >> (dphi & 63)) & 1u) xor_shifted(p, phi, 0);
      p.resize(r.size());
      r.swap(p);
    }
  }
  return r;
}

// ---------------------------------------------------------------- the exact parallel draw
constexpr int kSegWords = 1 << 22;  // words per thread and round (a multiple of 4)

// Jump polynomials t^(k * seg) mod phi, k = 1 .. T - 1, computed once per (thread count, segment).
struct JumpTable {
  int T = 0, seg = 0;
  std::vector<Poly> to;  // to[k]
};

inline const JumpTable& jump_table(int T, int seg) {
  static std::mutex mu;
  static std::vector<std::unique_ptr<JumpTable>> cache;
  std::lock_guard<std::mutex> lk(mu);
  for (const auto& t : cache)
    if (t->T == T && t->seg == seg) return *t;
  static Poly phi;
  static int dphi = -1;
  if (dphi < 0) {
    phi = char_poly();
    dphi = poly_deg(phi);
  }
  auto jt = std::make_unique<JumpTable>();
  jt->T = T;
  jt->seg = seg;
  jt->to.resize(static_cast<size_t>(T));
  if (T > 1) {
    jt->to[1] = pow_t(static_cast<uint64_t>(seg), phi, dphi);
    for (int k = 2; k < T; ++k) jt->to[k] = mulmod(jt->to[k - 1], jt->to[1], phi, dphi);
  }
  cache.push_back(std::move(jt));
  return *cache.back();
}

// One polar attempt on 4 words, as std::normal_distribution<double>(0, 1) computes it from a
// 32-bit engine: each uniform is generate_canonical<double, 53> of two words, u = fl(w0 + w1 2^32)
// / 2^64 (nextafter(1, 0) if that rounds to 1); x = 2 u1 - 1, y = 2 u2 - 1, r2 = x^2 + y^2; the
// attempt is accepted unless r2 > 1 or r2 == 0, and then gives first = y m, second = x m with
// m = sqrt(-2 ln r2 / r2) (each times stddev 1 plus mean 0).  Pinned bit for bit against the
// library's own distribution on the serial stream (tests/sanitize/host_checks.cpp).
inline bool polar_attempt(const uint32_t* w, double* first, double* second) {
  constexpr double k32 = 4294967296.0, k64 = 18446744073709551616.0;
  double u1 = (static_cast<double>(w[0]) + static_cast<double>(w[1]) * k32) / k64;
  double u2 = (static_cast<double>(w[2]) + static_cast<double>(w[3]) * k32) / k64;
  if (u1 >= 1.0) u1 = std::nextafter(1.0, 0.0);
  if (u2 >= 1.0) u2 = std::nextafter(1.0, 0.0);
  const double x = 2.0 * u1 - 1.0;
  const double y = 2.0 * u2 - 1.0;
  const double r2 = x * x + y * y;
  if (r2 > 1.0 || r2 == 0.0) return false;
  const double mult = std::sqrt(-2 * std::log(r2) / r2);
  *first = y * mult * 1.0 + 0.0;
  *second = x * mult * 1.0 + 0.0;
  return true;
}

// The reference start of one sample, entries [0, dim) in index order, handed to sink(first, n,
// values) on the calling thread in consecutive chunks.  T threads (T = 1: the serial loop), each
// drawing seg words (a multiple of 4) per round.
inline void draw_sample(unsigned seed, int64_t dim, int T,
                        const std::function<void(int64_t, int64_t, const std::complex<double>*)>& sink,
                        int seg = kSegWords) {
  if (dim <= 0) return;
  if (T <= 1) {
    std::mt19937 rng(seed);
    std::normal_distribution<double> g(0.0, 1.0);
    constexpr int64_t kChunk = 1 << 16;
    std::vector<std::complex<double>> buf(static_cast<size_t>(std::min<int64_t>(kChunk, dim)));
    for (int64_t i0 = 0; i0 < dim; i0 += kChunk) {
      const int64_t n = std::min<int64_t>(kChunk, dim - i0);
      for (int64_t i = 0; i < n; ++i) {
        const double re = g(rng);
        const double im = g(rng);
        buf[static_cast<size_t>(i)] = std::complex<double>(re, im);
      }
      sink(i0, n, buf.data());
    }
    return;
  }
  seg = std::max(4, seg - seg % 4);
  const JumpTable& jt = jump_table(T, seg);
  std::vector<Mt19937> gen(static_cast<size_t>(T), Mt19937(seed));
  std::vector<std::vector<uint32_t>> words(static_cast<size_t>(T));
  // two sets of per-thread results: the calling thread hands one round to the sink while the T
  // worker threads draw the next
  std::vector<std::vector<std::complex<double>>> vals[2];
  vals[0].resize(static_cast<size_t>(T));
  vals[1].resize(static_cast<size_t>(T));
  auto body = [&](int round, int k) {
    Mt19937& e = gen[static_cast<size_t>(k)];
    if (round == 0) {
      if (k > 0) e.apply_poly(jt.to[static_cast<size_t>(k)]);
    } else {
      e.apply_poly(jt.to[static_cast<size_t>(T - 1)]);  // from the end of its last segment
    }
    std::vector<uint32_t>& w = words[static_cast<size_t>(k)];
    w.resize(static_cast<size_t>(seg));
    e.fill(w.data(), static_cast<size_t>(seg));
    std::vector<std::complex<double>>& v = vals[round & 1][static_cast<size_t>(k)];
    v.clear();
    for (int a = 0; a < seg; a += 4) {
      double re, im;
      if (polar_attempt(w.data() + a, &re, &im)) v.emplace_back(re, im);
    }
  };
  auto launch = [&](int round) {
    std::vector<std::thread> th;
    for (int k = 0; k < T; ++k) th.emplace_back(body, round, k);
    return th;
  };
  // a round yields about T * seg / 5.09 entries (an attempt is accepted with probability pi/4)
  int64_t done = 0, drawn = 0;
  std::vector<std::thread> cur = launch(0);
  for (int round = 0; done < dim; ++round) {
    for (auto& x : cur) x.join();
    const std::vector<std::vector<std::complex<double>>>& got = vals[round & 1];
    for (const auto& v : got) drawn += static_cast<int64_t>(v.size());
    cur.clear();
    if (drawn < dim) cur = launch(round + 1);  // next round runs while this one is handed over
    for (int k = 0; k < T && done < dim; ++k) {
      const std::vector<std::complex<double>>& v = got[static_cast<size_t>(k)];
      const int64_t n = std::min<int64_t>(static_cast<int64_t>(v.size()), dim - done);
      sink(done, n, v.data());
      done += n;
    }
  }
  for (auto& x : cur) x.join();
}

}  // namespace ftlm_gauss
