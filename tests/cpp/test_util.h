// Minimal check macro for the C++ tests: counts failures, prints each one.
#pragma once
#include <cstdio>

namespace test {
inline int& failures() {
  static int n = 0;
  return n;
}
}  // namespace test

#define EXPECT(cond, ...)                                           \
  do {                                                              \
    if (!(cond)) {                                                  \
      std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);     \
      std::fprintf(stderr, __VA_ARGS__);                            \
      std::fprintf(stderr, "\n");                                   \
      ++test::failures();                                           \
    }                                                               \
  } while (0)
