// Equivalence of the transport's ray-sphere / ray-cone crossings (which skip
// the square root and divisions when the coefficient signs rule out a
// positive root) with the plain quadratics of MCRadiation.f:1274-1355, bit for
// bit, on random and special values (signed zeros, denormals, huge values).
#include "mcx_device.h"
#include <stdio.h>
#include <string.h>
#include <stdlib.h>
#include <random>
#include <math.h>
// the quadratics as the reference states them (MCRadiation.f:1274-1355), no shortcut
static bool ref_sphere(double Rad2, double r, double b, double &v) {
    bool hit = false; v = 1e200;
    double cc = r - Rad2, discr = b * b - 4.0 * cc;
    if (discr >= 0.0) {
        discr = sqrt(discr);
        double q = (b > 0.0) ? -0.5 * (b + discr) : -0.5 * (b - discr);
        double vr1 = q, vr2 = cc / q;
        if (vr1 > 0.0) { v = vr1; hit = true; }
        if (vr2 > 0.0 && vr2 < v) { v = vr2; hit = true; }
    }
    return hit;
}
static bool ref_cone(double Thet, double r, double b, double z, double vz, double &v) {
    bool hit = false; v = 1e200;
    double at = Thet - vz * vz, bt = Thet * b - (2.0 * z) * vz, ct = Thet * r - z * z;
    double discr = bt * bt - (4.0 * at) * ct;
    if (discr >= 0.0) {
        discr = sqrt(discr);
        double q = (bt > 0.0) ? -0.5 * (bt + discr) : -0.5 * (bt - discr);
        double vt1 = q / at, vt2 = ct / q;
        if (vt1 > 0.0) { v = vt1; hit = true; }
        if (vt2 > 0.0 && vt2 < v) { v = vt2; hit = true; }
    }
    return hit;
}
int main(int argc, char **argv) {
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    const double specials[] = {0.0, -0.0, 1e-300, -1e-300, 4.9e-324, -4.9e-324, 1.0, -1.0, 1e300, -1e300, 0.25, 2.0};
    long n = 0, bad = 0;
    auto pick = [&](double scale) { unsigned k = rng() % 16; return k < 12 ? specials[k] : U(rng) * scale; };
    const long N = argc > 1 ? atol(argv[1]) : 1000000;
    for (long it = 0; it < N; it++) {
        double R2 = fabs(pick(10.0)), r = fabs(pick(10.0)), b = pick(10.0), Th = fabs(pick(1.0)), z = pick(3.0), vz = pick(1.0);
        double v1, v2; bool h1 = hit_sphere(R2, r, b, v1), h2 = ref_sphere(R2, r, b, v2);
        n++; if (h1 != h2 || memcmp(&v1, &v2, 8)) { if (bad < 5) printf("sphere %a %a %a\n", R2, r, b); bad++; }
        h1 = hit_cone(Th, r, b, z, vz, v1); h2 = ref_cone(Th, r, b, z, vz, v2);
        n++; if (h1 != h2 || memcmp(&v1, &v2, 8)) { if (bad < 10) printf("cone %a %a %a %a %a\n", Th, r, b, z, vz); bad++; }
    }
    printf("checked %ld, mismatches %ld\n", n, bad);
    return bad != 0;
}
