// skinny_numeric_tests: the cancellation known-answer test of
// numeric_tests.cu:23-73, for the double (DR) and double-double (HR)
// accumulations of a 1 x 1 TSMTTSM.
//
// a = b = 0.1 over N = 111113 rows, except a[0] = a[1234] = a[N-1] = 2^62,
// a[41234] = -2^62, b[0] = 2^62, b[N-1] = -2^62.  The 2^124 terms and the
// 2^62 * 0.1 terms cancel exactly, leaving (N - 4) fl(0.1)^2, whose nearest
// double is 0x1.15c5c28f5c290p+10 = 1111.0900000000001 (the reference prints
// (N - 2) * 0.01 as "should be", which counts two rows too many).  The FP64
// sum loses the small terms to the 2^124 ones; the dd sum must give the
// exact value to a few ulp (the small terms sit in one double next to 2^124
// while a thread accumulates, so the last bits may differ).  Exit status 0
// iff the HR result is within 4 ulp of the exact value.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "common.hpp"

static void show(const char* what, double v) {
  uint64_t bits;
  std::memcpy(&bits, &v, sizeof(bits));
  std::printf("%-12s %.17g  (%a, bits %016llx)\n", what, v, v,
              (unsigned long long)bits);
}

int main() {
  const int N = 111113;
  const double big = std::ldexp(1.0, 62);
  std::vector<double> a(N, 0.1), b(N, 0.1);
  a[0] = a[1234] = a[N - 1] = big;
  a[41234] = -big;
  b[0] = big;
  b[N - 1] = -big;
  drv::DevBuf<double> da(N), db(N), dc(1), dq(1);
  HIP_CHECK(hipMemcpy(da.p, a.data(), sizeof(double) * N, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(db.p, b.data(), sizeof(double) * N, hipMemcpyHostToDevice));

  skinny::set_abort_on_error(true);
  const bool ok1 = skinny::tsmttsm<double, double>(1000, 1, 1, N, da.p, 1, 1.0, db.p,
                                                   1, 0.0, dc.p, 1);
  const bool ok2 = skinny::tsmttsm<double, PseudoQuad>(1000, 1, 1, N, da.p, 1, 1.0,
                                                       db.p, 1, 0.0, dq.p, 1);
  double c = 0, q = 0;
  HIP_CHECK(hipMemcpy(&c, dc.p, sizeof(double), hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(&q, dq.p, sizeof(double), hipMemcpyDeviceToHost));
  const double exact = 0x1.15c5c28f5c290p+10;  // nearest double to (N - 4) fl(0.1)^2
  show("Double:", c);
  show("PseudoQuad:", q);
  show("Exact:", exact);
  if (!ok1 || !ok2) {
    std::printf("a variant returned false\n");
    return 1;
  }
  const double ulp = std::nextafter(exact, 2 * exact) - exact;
  std::printf("errors in ulp of the exact value: double %.6g, PseudoQuad %.6g\n",
              std::fabs(c - exact) / ulp, std::fabs(q - exact) / ulp);
  const bool good = std::fabs(q - exact) <= 4 * ulp;
  std::printf(good ? "PseudoQuad result is within 4 ulp of the exact value\n"
                   : "PseudoQuad result is off by more than 4 ulp\n");
  return good ? 0 : 1;
}
