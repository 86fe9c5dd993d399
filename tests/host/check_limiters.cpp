// Host check of the device limiters (koral_amd/csrc/dev_grid.hpp compiled as
// plain C++): DMM, DM4 and mp5 must return bit-identical doubles (signed
// zeros, inf and NaN included) to the sign-arithmetic forms of
// finite.c:55-63 and finite.c:228-289, restated below.
#define KD_HOST_CHECK 1
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include "dev_grid.hpp"

namespace ref {
double sgn(double x) { return x > 0. ? 1. : (x < 0. ? -1. : 1.); }
double mn(double a, double b) { return a < b ? a : b; }
double mx(double a, double b) { return a > b ? a : b; }
double DMM(double X, double Y) { return .5 * (sgn(X) + sgn(Y)) * mn(fabs(X), fabs(Y)); }
double DM4(double W, double X, double Y, double Z)
{
  return 0.125 * (sgn(W) + sgn(X)) * fabs((sgn(W) + sgn(Y)) * (sgn(W) + sgn(Z))) *
         mn(mn(fabs(W), fabs(X)), mn(fabs(Y), fabs(Z)));
}
double mp5(double um2, double um1, double u0, double up1, double up2)
{
  const double alpha = 4., epsm = 1.e-10, BB2 = 1.3333333;
  double VOR = (3. * um2 - 20. * um1 + 90. * u0 + 60. * up1 - 5. * up2) / 128.;
  double VMP = u0 + DMM(up1 - u0, alpha * (u0 - um1));
  if ((VOR - u0) * (VOR - VMP) < epsm) return VOR;
  double DJM1 = um2 - 2. * um1 + u0, DJ = um1 - 2. * u0 + up1, DJP1 = u0 - 2. * up1 + up2;
  double DM4JPH = DM4(4. * DJ - DJP1, 4. * DJP1 - DJ, DJ, DJP1);
  double DM4JMH = DM4(4. * DJ - DJM1, 4. * DJM1 - DJ, DJ, DJM1);
  double VUL = u0 + alpha * (u0 - um1), VAV = 0.5 * (u0 + up1);
  double VMD = VAV - 0.5 * DM4JPH, VLC = u0 + 0.5 * (u0 - um1) + BB2 * DM4JMH;
  double VMIN = mx(mn(u0, mn(up1, VMD)), mn(u0, mn(VUL, VLC)));
  double VMAX = mn(mx(u0, mx(up1, VMD)), mx(u0, mx(VUL, VLC)));
  return VOR + DMM(VMIN - VOR, VMAX - VOR);
}
}  // namespace ref

static bool same(double a, double b) { return std::memcmp(&a, &b, 8) == 0 || (std::isnan(a) && std::isnan(b)); }

int main()
{
  const double v[] = {0., -0., 1., -1., 2.5, -3.5, 1e-300, -1e-300, INFINITY, -INFINITY, NAN, 7., -7.};
  const int n = sizeof(v) / sizeof(v[0]);
  long bad = 0, cnt = 0;
  for (int a = 0; a < n; a++)
    for (int b = 0; b < n; b++) {
      cnt++;
      bad += !same(ref::DMM(v[a], v[b]), kd::DMM(v[a], v[b]));
      for (int c = 0; c < n; c++)
        for (int d = 0; d < n; d++) {
          cnt++;
          bad += !same(ref::DM4(v[a], v[b], v[c], v[d]), kd::DM4(v[a], v[b], v[c], v[d]));
        }
    }
  std::mt19937_64 g(7);
  std::normal_distribution<double> nd;
  std::uniform_int_distribution<int> pick(0, n - 1);
  for (long i = 0; i < 1000000; i++) {
    double x[5];
    for (int k = 0; k < 5; k++) x[k] = (i % 7 == 0) ? v[pick(g)] : nd(g) * ((i % 3) ? 1. : 1e-6);
    if (i % 5 == 0) x[2] = x[1] + 1e-12 * nd(g);  // near-flat data around the switch
    cnt += 3;
    bad += !same(ref::DMM(x[0], x[1]), kd::DMM(x[0], x[1]));
    bad += !same(ref::DM4(x[0], x[1], x[2], x[3]), kd::DM4(x[0], x[1], x[2], x[3]));
    bad += !same(ref::mp5(x[0], x[1], x[2], x[3], x[4]), kd::mp5(x[0], x[1], x[2], x[3], x[4]));
  }
  std::printf("checked %ld, mismatches %ld\n", cnt, bad);
  return bad != 0;
}
