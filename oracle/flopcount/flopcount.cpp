// Flop counts of the CPU oracle (test infrastructure only): the oracle's C sources compiled as
// C++ with `double` replaced by a counting scalar, so that every floating-point +, -, *, /,
// sqrt, log, exp and pow the restated algorithm performs is counted once (fabs, fmin, fmax and
// comparisons are not counted; the source has no explicit fused multiply-add, so a*b+c counts
// 2).  The counting scalar has the layout of a double, so the oracle's problem structs and the
// caller's arrays are used unchanged.  Built by oracle/flopcount/Makefile into
// oracle/flopcount/libflopcount.so and driven by oracle/flopcount/flopcount.py, which writes
// profiles/flops.json (algorithmic flops per DOF-update for the benchmark configurations).
#include <cfloat>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static long long g_flops = 0;

struct fc {
  double v;
  fc() = default;
  constexpr fc(double x) : v(x) {}
  explicit operator double() const { return v; }
  explicit operator int() const { return (int)v; }
  explicit operator float() const { return (float)v; }
  explicit operator long() const { return (long)v; }
  fc &operator+=(fc b) { ++g_flops; v += b.v; return *this; }
  fc &operator-=(fc b) { ++g_flops; v -= b.v; return *this; }
  fc &operator*=(fc b) { ++g_flops; v *= b.v; return *this; }
  fc &operator/=(fc b) { ++g_flops; v /= b.v; return *this; }
  fc operator-() const { return fc(-v); }
  fc operator+() const { return *this; }
};
static_assert(sizeof(fc) == sizeof(double), "layout of a double");
inline fc operator+(fc a, fc b) { ++g_flops; return fc(a.v + b.v); }
inline fc operator-(fc a, fc b) { ++g_flops; return fc(a.v - b.v); }
inline fc operator*(fc a, fc b) { ++g_flops; return fc(a.v * b.v); }
inline fc operator/(fc a, fc b) { ++g_flops; return fc(a.v / b.v); }
inline bool operator<(fc a, fc b) { return a.v < b.v; }
inline bool operator>(fc a, fc b) { return a.v > b.v; }
inline bool operator<=(fc a, fc b) { return a.v <= b.v; }
inline bool operator>=(fc a, fc b) { return a.v >= b.v; }
inline bool operator==(fc a, fc b) { return a.v == b.v; }
inline bool operator!=(fc a, fc b) { return a.v != b.v; }
inline fc sqrt(fc a) { ++g_flops; return fc(std::sqrt(a.v)); }
inline fc log(fc a) { ++g_flops; return fc(std::log(a.v)); }
inline fc exp(fc a) { ++g_flops; return fc(std::exp(a.v)); }
inline fc pow(fc a, fc b) { ++g_flops; return fc(std::pow(a.v, b.v)); }
inline fc sin(fc a) { ++g_flops; return fc(std::sin(a.v)); }
inline fc cos(fc a) { ++g_flops; return fc(std::cos(a.v)); }
inline fc acos(fc a) { ++g_flops; return fc(std::acos(a.v)); }
inline fc fabs(fc a) { return fc(std::fabs(a.v)); }
inline fc fmax(fc a, fc b) { return fc(std::fmax(a.v, b.v)); }
inline fc fmin(fc a, fc b) { return fc(std::fmin(a.v, b.v)); }
inline fc copysign(fc a, fc b) { return fc(std::copysign(a.v, b.v)); }
inline bool isnan(fc a) { return std::isnan(a.v); }

#define double fc
#include "../ora_basis.c"
#include "../ora_mesh.c"
#include "../ora_ns.c"
#include "../ora_nfvse.c"
#include "../ora_mhd.c"
#undef double

extern "C" {
long long flopcount_reset() { long long f = g_flops; g_flops = 0; return f; }
long long flopcount_value() { return g_flops; }

// one TimeStepByLSERKW2 step with CALCTIMESTEP before it (Euler / Navier-Stokes)
long long flopcount_ns_step(const void *P, void *U, void *Ut_tmp, double cfl, double dfl)
{
  g_flops = 0;
  const ora_problem *p = static_cast<const ora_problem *>(P);
  int err = 0;
  fc dt = ora_calc_timestep(p, static_cast<const fc *>(U), fc(cfl), fc(dfl), &err);
  ora_lserk_step(p, static_cast<fc *>(U), static_cast<fc *>(Ut_tmp), fc(0.0), dt);
  return g_flops;
}

// one GLM-MHD step with CALCTIMESTEP before it
long long flopcount_mhd_step(const void *M, void *U, void *Ut_tmp, double cfl, double dfl)
{
  g_flops = 0;
  const ora_mhd *m = static_cast<const ora_mhd *>(M);
  fc dtch1 = ora_mhd_glm_dtch1(m, fc(cfl));
  fc ch = fc(0.5);
  int err = 0;
  fc dt = ora_mhd_calc_timestep(m, static_cast<const fc *>(U), fc(cfl), fc(dfl), dtch1, &ch, &err);
  ora_mhd_lserk_step(m, static_cast<fc *>(U), static_cast<fc *>(Ut_tmp), ch, dt);
  return g_flops;
}

// single flux evaluations
long long flopcount_volflux(int which, double kappa, const double *UL, const double *UR, const double *mL,
                            const double *mR)
{
  fc auxL[6], auxR[6], F[5];
  ora_eval_uaux(fc(kappa), reinterpret_cast<const fc *>(UL), auxL);
  ora_eval_uaux(fc(kappa), reinterpret_cast<const fc *>(UR), auxR);
  g_flops = 0;
  ora_volflux(which, fc(kappa), reinterpret_cast<const fc *>(UL), reinterpret_cast<const fc *>(UR), auxL, auxR,
              reinterpret_cast<const fc *>(mL), reinterpret_cast<const fc *>(mR), F);
  return g_flops;
}

long long flopcount_riemann(int which, double kappa, const double *UL, const double *UR, const double *nv,
                            const double *t1, const double *t2)
{
  fc F[5];
  g_flops = 0;
  ora_riemann(which, fc(kappa), reinterpret_cast<const fc *>(UL), reinterpret_cast<const fc *>(UR),
              reinterpret_cast<const fc *>(nv), reinterpret_cast<const fc *>(t1), reinterpret_cast<const fc *>(t2), F);
  return g_flops;
}

long long flopcount_derigs(const void *M, const double *UL, const double *UR, const double *mL, const double *mR)
{
  fc F[9];
  g_flops = 0;
  ora_mhd_derigs(static_cast<const ora_mhd *>(M), fc(0.5), reinterpret_cast<const fc *>(UL),
                 reinterpret_cast<const fc *>(UR), reinterpret_cast<const fc *>(mL), reinterpret_cast<const fc *>(mR),
                 F);
  return g_flops;
}
}
