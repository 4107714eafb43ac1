// mcx_device.h -- device-side building blocks of the MI355X transport:
// Philox4x32-10 packet streams and the cell geometry of the (r, theta) grid.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MCX_DTMAX 10000
#define MCX_DNANG 30

// Modules.f:5-11.  AU is a single-precision literal in the reference.
#define MCX_AU ((double)1.49598e13f)
#define MCX_PI 3.14159265358979323846264338328

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).  Stream of packet p under key `seed`:
// counter = (block, p); each block yields two doubles, taken in order
// (low 64 bits first).  Each double keeps the top 53 bits, centred: (0, 1).
__device__ __forceinline__ void philox_block(uint32_t k0, uint32_t k1, uint64_t blk, uint64_t pid, double &d0,
                                             double &d1) {
    uint32_t c0 = (uint32_t)blk, c1 = (uint32_t)(blk >> 32);
    uint32_t c2 = (uint32_t)pid, c3 = (uint32_t)(pid >> 32);
    uint32_t a = k0, b = k1;
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        uint32_t n0 = hi1 ^ c1 ^ a;
        uint32_t n2 = hi0 ^ c3 ^ b;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        a += 0x9E3779B9u;
        b += 0xBB67AE85u;
    }
    uint64_t v0 = ((uint64_t)c1 << 32) | c0;
    uint64_t v1 = ((uint64_t)c3 << 32) | c2;
    d0 = ((double)(v0 >> 11) + 0.5) * 0x1p-53;
    d1 = ((double)(v1 >> 11) + 0.5) * 0x1p-53;
}

struct PhiloxStream {
    uint32_t k0, k1;
    uint64_t pid;
    uint64_t blk;
    double cache;
    bool have;

    __device__ void init(uint64_t seed, uint64_t p) {
        k0 = (uint32_t)seed;
        k1 = (uint32_t)(seed >> 32);
        pid = p;
        blk = 0;
        have = false;
        cache = 0.0;
    }
    __device__ double next() {
        if (have) {
            have = false;
            return cache;
        }
        double d0;
        philox_block(k0, k1, blk, pid, d0, cache);
        blk++;
        have = true;
        return d0;
    }
};

// Radiation.f:1007-1024
template <typename Gen>
__device__ __forceinline__ void random_direction(Gen &g, double &x, double &y, double &z) {
    double r;
    for (;;) {
        x = 2.0 * g.next() - 1.0;
        y = 2.0 * g.next() - 1.0;
        z = 2.0 * g.next() - 1.0;
        r = x * x + y * y + z * z;
        if (!(r > 1.0)) break;
    }
    r = sqrt(r);
    x = x / r;
    y = y / r;
    z = z / r;
}

// ---------------------------------------------------------------------------
// Ray-sphere and ray-cone crossings, in the numerically stable q-form of the
// reference (MCRadiation.f:1274-1355).  Thet is cos^2 of the cone edge.
__host__ __device__ __forceinline__ bool hit_sphere(double Rad2, double r, double b, double &v) {
    bool hit = false;
    v = 1e200;
    double cc = r - Rad2;
    double discr = b * b - 4.0 * cc;
    // q = -(b + sign(b) sqrt(discr))/2 has the sign bit of -b (b > 0) or of
    // +0 (b <= 0); the roots are q and cc/q.  When neither can be positive
    // the square root and the division are skipped: the result is the same.
    const bool qneg = (b > 0.0);
    if (discr >= 0.0 && (!qneg || (qneg == (bool)__builtin_signbit(cc)))) {
        discr = sqrt(discr);
        double q = (b > 0.0) ? -0.5 * (b + discr) : -0.5 * (b - discr);
        double vr1 = q, vr2 = cc / q;
        if (vr1 > 0.0) { v = vr1; hit = true; }
        if (vr2 > 0.0 && vr2 < v) { v = vr2; hit = true; }
    }
    return hit;
}

__host__ __device__ __forceinline__ bool hit_cone(double Thet, double r, double b, double z, double vz, double &v) {
    bool hit = false;
    v = 1e200;
    double at = Thet - vz * vz;
    double bt = Thet * b - (2.0 * z) * vz;
    double ct = Thet * r - z * z;
    double discr = bt * bt - (4.0 * at) * ct;
    // the roots q/at and ct/q, q with the sign bit of -bt (bt > 0) or of +0
    // (bt <= 0): a root can be positive only when the signs agree
    const bool qneg = (bt > 0.0);
    if (discr >= 0.0 && ((qneg == (bool)__builtin_signbit(at)) || (qneg == (bool)__builtin_signbit(ct)))) {
        discr = sqrt(discr);
        double q = (bt > 0.0) ? -0.5 * (bt + discr) : -0.5 * (bt - discr);
        double vt1 = q / at, vt2 = ct / q;
        if (vt1 > 0.0) { v = vt1; hit = true; }
        if (vt2 > 0.0 && vt2 < v) { v = vt2; hit = true; }
    }
    return hit;
}

__host__ __device__ __forceinline__ bool hit_cone_same(double Thet, double b, double z, double vz, double &v) {
    double bt = Thet * b - (2.0 * z) * vz;
    double at = Thet - vz * vz;
    v = -bt / at;
    return !(v <= 0.0);
}

// Distance to the exit face of cell (i, j) (MCRadiation.f:1150-1272).
// R2/T2 are squared radial / cos-theta edges; ntheta_edges = nt + 1.
// Returns false on "nothing to hit".  edgeNr is updated; jnext is 0-based.
__host__ __device__ __forceinline__ bool exit_face(const double *__restrict__ R2, const double *__restrict__ T2,
                                          int nt, double x, double y, double z, double vx, double vy,
                                          double vz, int i, int jj, bool onEdge, int &edgeNr, double &v,
                                          int &inext, int &jnext) {
    const int NT = nt + 1;
    double r = x * x + y * y + z * z;
    double R1s = R2[i], R2s = R2[i + 1];
    double T1s = T2[jj], T2s = T2[jj + 1];
    double b = 2.0 * (x * vx + y * vy + z * vz);
    int j = jj + 1;
    bool h1, h2, h3, h4;
    double vR1, vR2, vT1, vT2;
    if (!onEdge) {
        h1 = hit_sphere(R1s, r, b, vR1);
        h2 = hit_sphere(R2s, r, b, vR2);
        h3 = hit_cone(T1s, r, b, z, vz, vT1);
        h4 = hit_cone(T2s, r, b, z, vz, vT2);
    } else if (edgeNr == 1) {
        h1 = false; vR1 = 1e200;
        h2 = hit_sphere(R2s, r, b, vR2);
        h3 = hit_cone(T1s, r, b, z, vz, vT1);
        h4 = hit_cone(T2s, r, b, z, vz, vT2);
    } else if (edgeNr == 2) {
        h1 = hit_sphere(R1s, r, b, vR1);
        h2 = true; vR2 = -b;
        h3 = hit_cone(T1s, r, b, z, vz, vT1);
        h4 = hit_cone(T2s, r, b, z, vz, vT2);
    } else if (edgeNr == 3) {
        h1 = hit_sphere(R1s, r, b, vR1);
        h2 = hit_sphere(R2s, r, b, vR2);
        h3 = false; vT1 = 1e200;
        if (j != NT - 1) {
            h4 = hit_cone(T2s, r, b, z, vz, vT2);
        } else {
            vT2 = -z / vz;
            h4 = (vT2 > 0.0);
        }
    } else {
        h1 = hit_sphere(R1s, r, b, vR1);
        h2 = hit_sphere(R2s, r, b, vR2);
        if (j != 1) h3 = hit_cone(T1s, r, b, z, vz, vT1);
        else { h3 = false; vT1 = 1e200; }
        if (j != NT - 1) h4 = hit_cone_same(T2s, b, z, vz, vT2);
        else { h4 = false; vT2 = 1e200; }
    }
    if (!h1 && !h2 && !h3 && !h4) return false;
    v = 1e200;
    int in = i, jn = j;
    if (h1 && vR1 < v) { v = vR1; in = i - 1; jn = j; edgeNr = 2; }
    if (h2 && vR2 < v) { v = vR2; in = i + 1; jn = j; edgeNr = 1; }
    if (h3 && vT1 < v) { v = vT1; in = i; jn = j - 1; edgeNr = 4; }
    if (h4 && vT2 < v) { v = vT2; in = i; jn = j + 1; edgeNr = 3; }
    if (jn == NT) { jn = NT - 1; edgeNr = 4; }
    if (jn == 0) { jn = 1; edgeNr = 3; }
    inext = in;
    jnext = jn - 1;
    return true;
}

// Subroutines.f:1648-1672: rotate (x,y,z) about the unit axis (u,v,w).
__device__ __forceinline__ void rotate_about(double &x, double &y, double &z, double u, double v, double w,
                                             double theta) {
    double sint, cost;
    sincos(theta, &sint, &cost);
    double u2 = u * u, v2 = v * v, w2 = w * w;
    double X = x, Y = y, Z = z;
    double inp = X * u + Y * v + Z * w;
    double y1 = u * inp + (X * (v2 + w2) - u * (v * Y + w * Z)) * cost + (v * Z - w * Y) * sint;
    double y2 = v * inp + (Y * (u2 + w2) - v * (u * X + w * Z)) * cost + (w * X - u * Z) * sint;
    double y3 = w * inp + (Z * (u2 + v2) - w * (u * X + v * Y)) * cost + (u * Y - v * X) * sint;
    x = y1;
    y = y2;
    z = y3;
}

// ---------------------------------------------------------------------------
// Single-precision counterparts of the direction arithmetic, for transport in
// FP32 (precision mode MCX_FP32): hardware sin/cos, log and reciprocal square
// root instead of the FP64 library sequences.  Positions, crossings, optical
// depths, tallies and temperatures stay FP64.
// The crossing quadratics take |v| = 1 (leading coefficient 1, and the
// second root -b of a packet on its own surface), so the FP32 directions
// become unit vectors in double before they reach the geometry: one FP64
// reciprocal square root instead of the FP64 trigonometry.
__device__ __forceinline__ void to_unit(double &x, double &y, double &z, double a, double b, double c) {
    const double q = 1.0 / sqrt(a * a + b * b + c * c);
    x = a * q;
    y = b * q;
    z = c * q;
}

template <typename Gen>
__device__ __forceinline__ void random_direction_f(Gen &g, double &x, double &y, double &z) {
    float a, b, c, r;
    for (;;) {
        a = 2.0f * (float)g.next() - 1.0f;
        b = 2.0f * (float)g.next() - 1.0f;
        c = 2.0f * (float)g.next() - 1.0f;
        r = a * a + b * b + c * c;
        if (!(r > 1.0f) && r > 0.0f) break;
    }
    to_unit(x, y, z, (double)a, (double)b, (double)c);
}

__device__ __forceinline__ void rotate_about_f(double &x, double &y, double &z, double u_, double v_, double w_,
                                               double theta) {
    float sint, cost;
    __sincosf((float)theta, &sint, &cost);
    const float u = (float)u_, v = (float)v_, w = (float)w_;
    const float X = (float)x, Y = (float)y, Z = (float)z;
    const float u2 = u * u, v2 = v * v, w2 = w * w;
    const float inp = X * u + Y * v + Z * w;
    x = (double)(u * inp + (X * (v2 + w2) - u * (v * Y + w * Z)) * cost + (v * Z - w * Y) * sint);
    y = (double)(v * inp + (Y * (u2 + w2) - v * (u * X + w * Z)) * cost + (w * X - u * Z) * sint);
    z = (double)(w * inp + (Z * (u2 + v2) - w * (u * X + v * Y)) * cost + (u * Y - v * X) * sint);
}

// unit vector (in double, see to_unit)
__device__ __forceinline__ void normalize_f(double &x, double &y, double &z) { to_unit(x, y, z, x, y, z); }

// the optical depth to the next interaction, -ln(u), in FP32 (u rounded to 1
// in float: the first-order 1 - u)
__device__ __forceinline__ double tau_draw_f(double u) {
    const float uf = (float)u;
    return (uf < 1.0f) ? (double)(-__logf(uf)) : 1.0 - u;
}

// (real(i))**4 in single precision, as the reference evaluates it in the
// T^4 interpolation (Radiation.f:551, 698).
__device__ __forceinline__ float fpow4(int i) {
    float x = (float)i;
    float x2 = x * x;
    return x2 * x2;
}
