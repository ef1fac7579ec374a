// Host build of the device shape-function kernels: checks that the
// branch-free quadratic B-spline evaluation, and its split form with the
// interior-node path, equal the piecewise definition
// (Bsplines.jl:18-83) bit for bit, for every boundary type, at every piece
// boundary, at its neighbouring doubles and on a dense grid.  Prints the
// number of points checked and exits 1 on the first mismatch.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../jump_amd/csrc/mpm_basis.hpp"

static bool same(double a, double b) {
  uint64_t x, y;
  memcpy(&x, &a, 8);
  memcpy(&y, &b, 8);
  return x == y || (isnan(a) && isnan(b));
}

int main() {
  const double knots[] = {-2.0, -1.5, -1.0, -0.5, 0.0, 0.5, 1.0, 1.5, 2.0};
  long checked = 0;
  for (int type = 1; type <= 4; type++) {
    auto check = [&](double r) {
      double n0, d0, n1, d1, n2, d2;
      jh::qbs_eval_branchy(type, r, n0, d0);
      jh::qbs_eval(type, r, n1, d1);
      jh::qbs_eval_split(type, r, n2, d2);  // the interior path for type 3
      checked++;
      if (!same(n0, n1) || !same(d0, d1) || !same(n0, n2) || !same(d0, d2)) {
        printf("mismatch type %d r=%.17g: %.17g/%.17g vs %.17g/%.17g vs %.17g/%.17g\n", type, r, n0, d0, n1, d1, n2,
               d2);
        return false;
      }
      return true;
    };
    for (double k : knots) {
      double lo = k, hi = k;
      for (int s = 0; s < 64; s++) {
        if (!check(lo) || !check(hi)) return 1;
        lo = nextafter(lo, -INFINITY);
        hi = nextafter(hi, INFINITY);
      }
    }
    for (int i = -300000; i <= 300000; i++)
      if (!check(i * 1e-5)) return 1;
    // random doubles in [-2, 2] (a fixed linear congruential sequence)
    uint64_t z = 88172645463325252ull;
    for (int i = 0; i < 1000000; i++) {
      z = z * 6364136223846793005ull + 1442695040888963407ull;
      if (!check(-2.0 + 4.0 * double(z >> 11) * (1.0 / 9007199254740992.0))) return 1;
    }
    if (!check(NAN) || !check(INFINITY) || !check(-INFINITY) || !check(-0.0)) return 1;
  }
  printf("%ld\n", checked);
  return 0;
}
