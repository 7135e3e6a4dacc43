// Counts the floating-point operations of one PNAcc pair at a given PN
// order: the post-Newtonian terms of tupan_amd/csrc/pn_terms.inc evaluated
// with a number type that counts each +, -, *, / (negation and the rational
// constants Q(n, d) are free; additions of a term that is structurally zero
// -- an unused PN slot -- are not counted), plus the pair set-up and
// accumulation of pnacc_kernel_common.h:30-40, 66-68 as that file counts
// them (40 operations with smoothed_inv_r1r2r3 as 4).
// Build: g++ -std=c++17 -I tupan_amd/csrc tests/pnacc_flops.cpp
#include <cstdio>
#include <cstdlib>

static long g_ops = 0;

struct Num {
  double v;
  bool zero;  // structurally zero (never assigned a computed value)
  Num() : v(0), zero(true) {}
  Num(int x) : v(x), zero(x == 0) {}
  Num(double x) : v(x), zero(false) {}
  static Num konst(double x) { Num n(x); return n; }
};
static Num op(double v) { ++g_ops; return Num(v); }
static Num operator+(Num a, Num b) {
  if (a.zero) return b;
  if (b.zero) return a;
  return op(a.v + b.v);
}
static Num operator-(Num a, Num b) {
  if (b.zero) return a;
  return op(a.v - b.v);
}
static Num operator*(Num a, Num b) { return op(a.v * b.v); }
static Num operator/(Num a, Num b) { return op(a.v / b.v); }
static Num operator-(Num a) { Num r(-a.v); r.zero = a.zero; return r; }
static Num operator+(Num a) { return a; }
static Num &operator+=(Num &a, Num b) { a = a + b; return a; }
static Num &operator-=(Num &a, Num b) { a = a - b; return a; }
static Num &operator*=(Num &a, Num b) { a = a * b; return a; }
static bool operator>(Num a, Num b) { return a.v > b.v; }
static bool operator<(Num a, Num b) { return a.v < b.v; }
static Num sqrt(Num a) { return op(__builtin_sqrt(a.v)); }

#define __device__
#define __forceinline__ inline
#define double Num
#define Q(n, d) (Num::konst(((n) * 1.0) / (d)))
static const Num kPi2 = Num::konst(9.869604401089358);
struct PnParams {
  unsigned order;
  Num inv1, inv2, inv3, inv4, inv5, inv6, inv7;
};
struct PnAB {
  Num a, b;
};
#include "pn_terms.inc"
#undef double

int main(int argc, char **argv) {
  const unsigned order = argc > 1 ? (unsigned)atoi(argv[1]) : 5;
  PnParams c{order, Num(1e-3), Num(1e-6), Num(1e-9), Num(1e-12), Num(1e-15),
             Num(1e-18), Num(1e-21)};
  Num im(0.3), jm(0.2), rx(0.1), ry(-0.2), rz(0.05), vx(0.01), vy(0.02),
      vz(-0.03), v2(0.0014), ivx(0.011), ivy(0.019), ivz(-0.029), jvx(0.001),
      jvy(-0.001), jvz(0.001), inv_r(4.0), inv_r2(16.0), inv_r3(64.0);
  g_ops = 0;
  PnAB pn = pn_pair(im, jm, rx, ry, rz, vx, vy, vz, v2, ivx, ivy, ivz, jvx,
                    jvy, jvz, inv_r, inv_r2, inv_r3, c);
  (void)pn;
  const long setup = 33;  // pnacc_kernel_common.h:30-40, 66-68
  printf("%ld %ld\n", g_ops, g_ops + setup);
  return 0;
}
