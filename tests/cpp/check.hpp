// Minimal checking helpers for the drop-in C++ API tests (exit code != 0 on failure).
#pragma once

#include <cmath>
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>

#define CHECK(cond)                                                                      \
  do {                                                                                   \
    if (!(cond)) {                                                                       \
      std::cerr << __FILE__ << ":" << __LINE__ << ": check failed: " #cond << std::endl; \
      std::exit(1);                                                                      \
    }                                                                                    \
  } while (0)

#define CHECK_NEAR(a, b, tol)                                                                              \
  do {                                                                                                     \
    const double _a = (a), _b = (b);                                                                       \
    if (!(std::abs(_a - _b) <= (tol))) {                                                                   \
      std::cerr << __FILE__ << ":" << __LINE__ << ": " #a " = " << _a << " vs " #b " = " << _b << std::endl; \
      std::exit(1);                                                                                        \
    }                                                                                                      \
  } while (0)

template <class E, class F>
bool throws(F&& f) {
  try {
    f();
  } catch (const E&) {
    return true;
  } catch (...) {
    return false;
  }
  return false;
}
