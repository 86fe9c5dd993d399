// Operation-counting stand-in for double, used to compile the oracle's C
// sources as C++ and count the floating-point operations of the algorithm
// (test infrastructure; the counts give the algorithmic flops per unit of
// work that the benchmark divides by measured time).
#pragma once
#include <math.h>
#include <float.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>

struct FlopCount {
  uint64_t add, mul, div, sqrt, trans;
};
extern FlopCount g_flops;

struct cd {
  double v;
  cd() = default;
  template <class T>
  cd(T x) : v((double)x) {}
  explicit operator double() const { return v; }
  explicit operator int() const { return (int)v; }
  cd operator-() const { return cd(-v); }
  cd &operator+=(cd o) { g_flops.add++; v += o.v; return *this; }
  cd &operator-=(cd o) { g_flops.add++; v -= o.v; return *this; }
  cd &operator*=(cd o) { g_flops.mul++; v *= o.v; return *this; }
  cd &operator/=(cd o) { g_flops.div++; v /= o.v; return *this; }
};
inline cd operator+(cd a, cd b) { g_flops.add++; return cd(a.v + b.v); }
inline cd operator-(cd a, cd b) { g_flops.add++; return cd(a.v - b.v); }
inline cd operator*(cd a, cd b) { g_flops.mul++; return cd(a.v * b.v); }
inline cd operator/(cd a, cd b) { g_flops.div++; return cd(a.v / b.v); }
inline bool operator<(cd a, cd b) { return a.v < b.v; }
inline bool operator>(cd a, cd b) { return a.v > b.v; }
inline bool operator<=(cd a, cd b) { return a.v <= b.v; }
inline bool operator>=(cd a, cd b) { return a.v >= b.v; }
inline bool operator==(cd a, cd b) { return a.v == b.v; }
inline bool operator!=(cd a, cd b) { return a.v != b.v; }
inline cd sqrt(cd a) { g_flops.sqrt++; return cd(::sqrt(a.v)); }
inline cd fabs(cd a) { return cd(::fabs(a.v)); }
inline cd pow(cd a, cd b) { g_flops.trans++; return cd(::pow(a.v, b.v)); }
inline cd log(cd a) { g_flops.trans++; return cd(::log(a.v)); }
inline cd exp(cd a) { g_flops.trans++; return cd(::exp(a.v)); }
inline bool isfinite(cd a) { return std::isfinite(a.v); }
inline bool isnan(cd a) { return std::isnan(a.v); }
