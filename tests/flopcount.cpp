// Floating-point operation count of one hex element of each element routine
// of the CPU oracle (test infrastructure).
//
// The oracle's C source is compiled here as C++ with `double` replaced by a
// value type that counts every addition/subtraction, multiplication,
// division and square root it performs (sign changes, comparisons, fabs,
// fmin and fmax are not counted).  The element is a distorted unit hex with
// seeded nodal data; the count is that of the restatement's expression trees,
// which follow the reference's (assembleNavSto.f90, e3LHS.f90, e3RHS.f90,
// e3STAB.f90, solveQuenching.f90).
//
// Output, one line per case: name add mul div sqrt total.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace fc {
long n_add = 0, n_mul = 0, n_div = 0, n_sqrt = 0;
}

struct fd {
  double v;
  fd() = default;
  fd(double x) : v(x) {}  // NOLINT: implicit, as the literals of the C source
  explicit operator int() const { return (int)v; }
  explicit operator long() const { return (long)v; }
  explicit operator double() const { return v; }
  fd &operator+=(fd o) { ++fc::n_add; v += o.v; return *this; }
  fd &operator-=(fd o) { ++fc::n_add; v -= o.v; return *this; }
  fd &operator*=(fd o) { ++fc::n_mul; v *= o.v; return *this; }
  fd &operator/=(fd o) { ++fc::n_div; v /= o.v; return *this; }
  fd operator-() const { return fd(-v); }
  fd operator+() const { return *this; }
};
inline fd operator+(fd a, fd b) { ++fc::n_add; return fd(a.v + b.v); }
inline fd operator-(fd a, fd b) { ++fc::n_add; return fd(a.v - b.v); }
inline fd operator*(fd a, fd b) { ++fc::n_mul; return fd(a.v * b.v); }
inline fd operator/(fd a, fd b) { ++fc::n_div; return fd(a.v / b.v); }
inline bool operator<(fd a, fd b) { return a.v < b.v; }
inline bool operator>(fd a, fd b) { return a.v > b.v; }
inline bool operator<=(fd a, fd b) { return a.v <= b.v; }
inline bool operator>=(fd a, fd b) { return a.v >= b.v; }
inline bool operator==(fd a, fd b) { return a.v == b.v; }
inline bool operator!=(fd a, fd b) { return a.v != b.v; }
inline fd sqrt(fd a) { ++fc::n_sqrt; return fd(std::sqrt(a.v)); }
inline fd fabs(fd a) { return fd(std::fabs(a.v)); }
inline fd fmax(fd a, fd b) { return fd(std::fmax(a.v, b.v)); }
inline fd fmin(fd a, fd b) { return fd(std::fmin(a.v, b.v)); }

#define double fd
#include "../oracle/oracle.c"
#undef double

namespace {

unsigned long long rng_state = 12345;
double urand() {  // xorshift64*, uniform in [0, 1)
  rng_state ^= rng_state >> 12;
  rng_state ^= rng_state << 25;
  rng_state ^= rng_state >> 27;
  return (double)((rng_state * 2685821657736338717ULL) >> 11) * (1.0 / 9007199254740992.0);
}

void reset() { fc::n_add = fc::n_mul = fc::n_div = fc::n_sqrt = 0; }

void report(const char *name) {
  std::printf("%s %ld %ld %ld %ld %ld\n", name, fc::n_add, fc::n_mul, fc::n_div, fc::n_sqrt,
              fc::n_add + fc::n_mul + fc::n_div + fc::n_sqrt);
}

// linhexshl node order: (xi, eta, zeta) signs of nodes 1..8
const int kSgn[8][3] = {{-1, 1, -1}, {1, 1, -1}, {1, 1, 1},  {-1, 1, 1},
                        {-1, -1, -1}, {1, -1, -1}, {1, -1, 1}, {-1, -1, 1}};

}  // namespace

int main() {
  const int n = 8;
  fd xl[24], dl[24], ul[24], uml[24], acl[24], acml[24], pl[8], phil[8], rphil[8], Tl[8],
      rTl[8];
  for (int a = 0; a < n; ++a) {
    for (int c = 0; c < 3; ++c) {
      const double h = c == 0 ? 0.021 : (c == 1 ? 0.019 : 0.02);
      xl[c * n + a] = 0.5 * h * (kSgn[a][c] + 1) + 0.002 * (urand() - 0.5);
      dl[c * n + a] = 1e-5 * (urand() - 0.5);
      ul[c * n + a] = 0.6 * (urand() - 0.5);
      uml[c * n + a] = 0.1 * (urand() - 0.5);
      acl[c * n + a] = 4.0 * (urand() - 0.5);
      acml[c * n + a] = 0.2 * (urand() - 0.5);
    }
    pl[a] = 200.0 * (urand() - 0.5);
    phil[a] = 0.3 + 0.4 * urand();  // interface cell: both phases
    rphil[a] = 10.0 * (urand() - 0.5);
    Tl[a] = 360.0 + 30.0 * urand();  // around T_sat = 373.15: both mass-transfer branches
    rTl[a] = 100.0 * (urand() - 0.5);
  }
  // NS parameters: almi alfi gami beti delt rho mu kdc gx gy gz
  const fd prm[11] = {0.8333333333333334, 0.6666666666666666, 0.6666666666666667, 0.0, 1e-3,
                      1.0, 1e-3, 0.0, 0.0, 0.0, -9.81};
  // quench material (19 values and the fluid id)
  const fd qp[20] = {1.0, 1000.0, 1e-5, 1e-3, 2000.0, 4200.0, 0.025, 0.6, 7800.0, 1e-3,
                     500.0, 30.0, 373.15, 2.26e6, 0.1, 0.1, 1.0, 0.0, 425.0, 102.0};
  const fd ex[4] = {0.0, 0.0, -9.81, 0.0};
  static fd ke[9 * 64], ge[3 * 64], de[3 * 64], me[64], lse[64], lsue[3 * 64], ulse[3 * 64],
      plse[64], ru[24], rp[8], rls[8], te[64], rt[8];
  fd cfl = 0.0;
  const int flags[2] = {6, 2};
  const char *fname[2] = {"vecmat", "vec"};
  for (int f = 0; f < 2; ++f) {
    char name[64];
    reset();
    orc_element_ns(8, 8, xl, dl, ul, uml, acl, acml, pl, nullptr, prm, 0, flags[f], ke, ge, de,
                   me, ru, rp, &cfl);
    std::snprintf(name, sizeof name, "ns_%s", fname[f]);
    report(name);
    reset();
    orc_element_nsvof(8, 8, xl, dl, ul, uml, acl, pl, phil, rphil, Tl, prm, qp, ex, 1, flags[f],
                      ke, ge, de, me, lse, lsue, ulse, plse, ru, rp, rls, &cfl);
    std::snprintf(name, sizeof name, "nsvof_%s", fname[f]);
    report(name);
    reset();
    orc_element_tem(8, 8, xl, dl, ul, uml, phil, Tl, rTl, nullptr, prm, qp, 1, flags[f], te,
                    rt);
    std::snprintf(name, sizeof name, "tem_%s", fname[f]);
    report(name);
  }
  return 0;
}
