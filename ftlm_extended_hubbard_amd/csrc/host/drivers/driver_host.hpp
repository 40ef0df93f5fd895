// driver_host.hpp -- the host-only pieces of the n(mu) and small bench drivers: --key=value flag
// lookup (the reference drivers' parser, benchmarks/bench_ftlm_nmu_rect.cpp:40-70), log-sum-exp,
// and the grand-canonical n(mu) lines of benchmarks/bench_ftlm_nmu_rect.cpp:186-219.  Kept apart
// from the device calls so that a CPU-only program can run them under the sanitizers.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <ostream>
#include <string>
#include <vector>

namespace ftlm_driver {

// The value of --key=value, or nullptr (the first match wins).
inline const char* find_arg(int argc, char** argv, const char* key) {
  const std::string prefix = std::string("--") + key + "=";
  for (int i = 1; i < argc; ++i)
    if (argv[i] != nullptr && std::strncmp(argv[i], prefix.c_str(), prefix.size()) == 0) return argv[i] + prefix.size();
  return nullptr;
}
inline double arg_double(int argc, char** argv, const char* key, double dflt) {
  const char* v = find_arg(argc, argv, key);
  return v ? std::strtod(v, nullptr) : dflt;
}
inline int arg_int(int argc, char** argv, const char* key, int dflt) {
  const char* v = find_arg(argc, argv, key);
  return v ? std::atoi(v) : dflt;
}
inline bool has_arg(int argc, char** argv, const char* key) { return find_arg(argc, argv, key) != nullptr; }

inline double logsumexp2(double a, double b) {
  if (!std::isfinite(a)) return b;
  if (!std::isfinite(b)) return a;
  const double m = std::max(a, b);
  return m + std::log(std::exp(a - m) + std::exp(b - m));
}

// n(mu) lines of bench_ftlm_nmu_rect.cpp:186-219: for n_mu points of [mu_min, mu_max], the
// grand-canonical <N> / n_sites over the sectors with finite logZ (weights logZ + beta mu N,
// shifted by their maximum).
inline void write_nmu(std::ostream& out, const std::vector<double>& logZ, const std::vector<int>& Nelec, int n_sites,
                      double beta, double mu_min, double mu_max, int n_mu) {
  const int nsec = static_cast<int>(std::min(logZ.size(), Nelec.size()));
  for (int k = 0; k < n_mu; ++k) {
    const double tmu = static_cast<double>(k) / static_cast<double>(n_mu - 1);
    const double mu = mu_min + tmu * (mu_max - mu_min);
    double mx = -std::numeric_limits<double>::infinity();
    for (int i = 0; i < nsec; ++i) {
      if (!std::isfinite(logZ[i])) continue;
      mx = std::max(mx, logZ[i] + beta * mu * static_cast<double>(Nelec[i]));
    }
    if (!std::isfinite(mx)) {
      out << mu << "\t0\n";
      continue;
    }
    double sum_w = 0.0, sum_Nw = 0.0;
    for (int i = 0; i < nsec; ++i) {
      if (!std::isfinite(logZ[i])) continue;
      const double w = std::exp(logZ[i] + beta * mu * static_cast<double>(Nelec[i]) - mx);
      sum_w += w;
      sum_Nw += static_cast<double>(Nelec[i]) * w;
    }
    const double n_avg = (sum_w > 0.0) ? (sum_Nw / sum_w) / static_cast<double>(n_sites) : 0.0;
    out << mu << "\t" << n_avg << "\n";
  }
}

}  // namespace ftlm_driver
