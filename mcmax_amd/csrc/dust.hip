// dust.hip -- dust opacities from optical constants on the device: MCMax's
// ComputePart (ComputePart.f:1-552) for the DHS particle, standard = FILE.
//
// Work and kernels:
//   layered-sphere call, one per (species, wavelength, material, size, vacuum
//   fraction): DMiLay (ComputePart.f:1256-1715, Toon & Ackerman 1981)
//     k_dhs_coef    one thread per call: the down-recurrences of the
//                   logarithmic derivatives into scratch, then the series
//                   coefficients a_n, b_n up to convergence, each written over
//                   the scratch slot it has just consumed;
//     k_dhs_angles  one wavefront per call: the amplitude sums at the 90
//                   cosines and their supplements, as M1, M2, S21, D21 (REAL);
//   homogeneous-sphere item, one per (species, wavelength, material, size)
//   that needs one: MeerhoffMie (ComputePart.f:1883-1930; de Rooij & van der
//   Stap 1984)
//     k_hom_coef    one thread per item: fichid's recurrences, anbn, the cross
//                   sections and the number of terms of the angular sums;
//     k_hom_angles  one wavefront per item: the 90 symmetric angle pairs;
//   k_part_sum      one block per (species, wavelength): ComputePart's loop
//                   over materials, sizes and fractions, each term's
//                   forward-peak normalisation, and its single-precision sums.
//
// The host computes every transcendental starting value the way the reference
// does (sin, cos and exp of the size parameters, the angle cosines, the sine
// weights), so the kernels run only IEEE +, -, *, / without contraction.
#include "mcx_internal.h"

namespace {

constexpr double DPI = 3.14159265358979323846264338328; // Modules.f:7
constexpr int NA = 180, NH = 90;

struct cxd {
    double re, im;
};
__device__ __forceinline__ cxd mk(double r, double i) {
    cxd z;
    z.re = r;
    z.im = i;
    return z;
}
__device__ __forceinline__ cxd cadd(cxd a, cxd b) { return mk(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ cxd csub(cxd a, cxd b) { return mk(a.re - b.re, a.im - b.im); }
__device__ __forceinline__ cxd cmul(cxd a, cxd b) { return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
__device__ __forceinline__ cxd cscl(double s, cxd a) { return mk(s * a.re, s * a.im); }
// Smith's complex division
__device__ __forceinline__ cxd cdiv(cxd a, cxd b) {
    if (fabs(b.re) >= fabs(b.im)) {
        const double r = b.im / b.re, den = b.re + b.im * r;
        return mk((a.re + a.im * r) / den, (a.im - a.re * r) / den);
    }
    const double r = b.re / b.im, den = b.im + b.re * r;
    return mk((a.re * r + a.im) / den, (a.im * r - a.re) / den);
}
__device__ __forceinline__ cxd rdiv(double a, cxd b) { return cdiv(mk(a, 0.0), b); }

// One layered sphere: vacuum or material core of radius rcore inside a shell
// of radius rshell; wave number wvno.  Host-computed starting values.
struct DhsCall {
    double xshell, xcore;              // REAL products rshell*wvno, rcore*wvno
    double cosx, sinx;                 // of xshell (homogeneous wave function)
    double sinx1, cosx1, sinx4, cosx4; // of X1 = Re(m) xshell, X4 = Re(m) xcore
    double ey1, ey4, ey1my4;           // exp(Y1), exp(Y4), exp(Y1 - Y4)
    long long off;                     // first scratch slot (64 bytes each)
    float wvno, sh_re, sh_im, co_re, co_im;
    int nmx1, nmx2;                    // recurrence start; dimension of the sum (0: input error)
    int pad;
};
struct DhsRes {
    int nterms, err;
    float qext, qsca;
};

// One homogeneous sphere (idis = 0, one radius).
struct HomItem {
    double x, perx, sinx, cosx; // size parameter and the values at x
    double m_re, m_im;          // m = e1 - i e2
    double fakt;                // lam^2 / (2 pi)
    long long off_psi;          // doubles
    long long off_D, off_ab;    // complex slots
    int nmax, nfi, nD, pad;
};
struct HomRes {
    double csca, cext;
    int nstop, pad;
};

// One term of ComputePart's sums (ComputePart.f:306-397).
struct Term {
    int kind;        // 0 layered sphere, 1 homogeneous sphere evaluated, 2 homogeneous values of the previous term
    int idx;         // call (kind 0) or item (kind 1)
    float wn, wnrho; // wf(i)*nr(l,k), and that times rho(l) (REAL)
    float r3;        // r1**3 (REAL)
    float rad2;      // rad**2 of the layered sphere (REAL)
    float w2;        // wvno**2 (REAL)
    int pad;
};

__global__ __launch_bounds__(64) void k_dhs_coef(const DhsCall *__restrict__ calls, int first, int ncall,
                                                 long long base, cxd *__restrict__ scr, DhsRes *__restrict__ res) {
    const int c = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= first + ncall) return;
    const DhsCall q = calls[c];
    DhsRes out;
    out.nterms = 0;
    out.err = 1;
    out.qext = 0.0f;
    out.qsca = 0.0f;
    if (q.nmx1 <= 0) {
        res[c] = out;
        return;
    }
    cxd *S = scr + 4 * (q.off - base); // slot n: S[4n] = A_n, S[4n+1..3] = W_n at Z2, Z3, Z4
    const double xshell = q.xshell, xcore = q.xcore;
    const cxd K1 = mk((double)(q.co_re * q.wvno), (double)(q.co_im * q.wvno));
    const cxd K2 = mk((double)(q.sh_re * q.wvno), (double)(q.sh_im * q.wvno));
    const cxd K3 = mk((double)q.wvno, 0.0);
    cxd Z[4];
    Z[0] = mk((double)q.sh_re * xshell, (double)q.sh_im * xshell);
    Z[1] = mk(xshell, 0.0);
    Z[2] = mk((double)q.co_re * xcore, (double)q.co_im * xcore);
    Z[3] = mk((double)q.sh_re * xcore, (double)q.sh_im * xcore);
    const double RX = 1.0 / xshell;
    {
        const cxd rrfx = rdiv(1.0, Z[0]);
        cxd a1 = mk(0.0, 0.0), w1[3] = {mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0)};
        for (int nn = q.nmx1; nn >= 1; nn--) {
            const cxd a = cscl((double)(nn + 1), rrfx);
            a1 = csub(a, rdiv(1.0, cadd(a, a1)));
            S[4 * (size_t)nn] = a1;
#pragma unroll
            for (int m = 0; m < 3; m++) {
                const cxd qq = rdiv((double)(nn + 1), Z[m + 1]);
                w1[m] = csub(qq, rdiv(1.0, cadd(qq, w1[m])));
                S[4 * (size_t)nn + 1 + m] = w1[m];
            }
        }
    }
    cxd wm1 = mk(q.cosx, -q.sinx), wfn0 = mk(q.sinx, q.cosx);
    cxd wfn1 = csub(cscl(RX, wfn0), wm1);
    double ta3 = wfn1.re;
    int N = 1;
    const double sinx1 = q.sinx1, sinx4 = q.sinx4, cosx1 = q.cosx1, cosx4 = q.cosx4;
    const double ey1 = q.ey1, e2y1 = ey1 * ey1, ey4 = q.ey4, ey1my4 = q.ey1my4, ey1py4 = ey1 * ey4;
    const double aa = sinx4 * (ey1py4 + ey1my4), bb = cosx4 * (ey1py4 - ey1my4);
    const double cc = sinx1 * (e2y1 + 1.0), dd = cosx1 * (e2y1 - 1.0);
    const double denom = 1.0 + e2y1 * (4.0 * (sinx1 * sinx1) - 2.0 + e2y1);
    cxd acap = S[4], W0 = S[5], W1 = S[6], W2 = S[7];
    cxd dummy = mk((aa * cc + bb * dd) / denom, (bb * cc - aa * dd) / denom);
    dummy = cdiv(cmul(dummy, cadd(acap, rdiv(N, Z[0]))), cadd(W2, rdiv(N, Z[3])));
    cxd dumsq = cmul(dummy, dummy);
    cxd p24h24 = cadd(mk(0.5, 0.0), cscl(ey4 * ey4, mk(sinx4 * sinx4 - 0.5, cosx4 * sinx4)));
    cxd p24h21 = cadd(cscl(ey1py4, cscl(0.5, mk(sinx1 * sinx4 - cosx1 * cosx4, sinx1 * cosx4 + cosx1 * sinx4))),
                      cscl(ey1my4, cscl(0.5, mk(sinx1 * sinx4 + cosx1 * cosx4, -sinx1 * cosx4 + cosx1 * sinx4))));
    cxd dh1, dh2, dh4;
    {
        cxd z = Z[0];
        dh1 = csub(cdiv(z, cadd(mk(1.0, 0.0), mk(-z.im, z.re))), rdiv(1.0, z));
        z = Z[1];
        dh2 = csub(cdiv(z, cadd(mk(1.0, 0.0), mk(-z.im, z.re))), rdiv(1.0, z));
        z = Z[3];
        dh4 = csub(cdiv(z, cadd(mk(1.0, 0.0), mk(-z.im, z.re))), rdiv(1.0, z));
    }
    p24h24 = cdiv(p24h24, cmul(cadd(dh4, rdiv(N, Z[3])), cadd(W2, rdiv(N, Z[3]))));
    p24h21 = cdiv(p24h21, cmul(cadd(dh1, rdiv(N, Z[0])), cadd(W2, rdiv(N, Z[3]))));

    cxd acoe, bcoe;
    auto coefs = [&]() {
        const cxd U1 = csub(cmul(K3, acap), cmul(K2, W0));
        const cxd U2 = csub(cmul(K3, acap), cmul(K2, dh2));
        const cxd U3 = csub(cmul(K2, acap), cmul(K3, W0));
        const cxd U4 = csub(cmul(K2, acap), cmul(K3, dh2));
        const cxd U5 = csub(cmul(K1, W2), cmul(K2, W1));
        const cxd U6 = csub(cmul(K2, W2), cmul(K1, W1));
        const cxd t7 = csub(cmul(dummy, p24h21), p24h24);
        const cxd U7 = mk(t7.im, -t7.re);
        const cxd U8 = rdiv(ta3, wfn1);
        const cxd dk3 = cmul(dumsq, K3), dk2 = cmul(dumsq, K2);
        const cxd na_ = csub(cadd(cmul(cmul(U1, U5), U7), cmul(K1, U1)), cmul(dk3, U5));
        const cxd da_ = csub(cadd(cmul(cmul(U2, U5), U7), cmul(K1, U2)), cmul(dk3, U5));
        acoe = cdiv(cmul(U8, na_), da_);
        const cxd nb_ = csub(cadd(cmul(cmul(U3, U6), U7), cmul(K2, U3)), cmul(dk2, U6));
        const cxd db_ = csub(cadd(cmul(cmul(U4, U6), U7), cmul(K2, U4)), cmul(dk2, U6));
        bcoe = cdiv(cmul(U8, nb_), db_);
    };
    coefs();
    S[4] = acoe;
    S[5] = bcoe;
    cxd acoem1 = acoe, bcoem1 = bcoe;
    double are = acoe.re, aim = acoe.im, bre = bcoe.re, bim = bcoe.im;
    double dqext = 3.0 * (are + bre);
    double dqsca = 3.0 * (are * are + aim * aim + bre * bre + bim * bim);
    double dgqsc = 0.0;
    int err = 0;
    N = 2;
    for (;;) {
        const double T1 = 2 * N - 1;
        acap = S[4 * (size_t)N];
        W0 = S[4 * (size_t)N + 1];
        W1 = S[4 * (size_t)N + 2];
        W2 = S[4 * (size_t)N + 3];
        wm1 = wfn0;
        wfn0 = wfn1;
        wfn1 = csub(cscl(T1 * RX, wfn0), wm1);
        ta3 = wfn1.re;
        {
            cxd nz = rdiv(N, Z[0]);
            dh1 = cadd(mk(-nz.re, -nz.im), rdiv(1.0, csub(nz, dh1)));
            nz = rdiv(N, Z[1]);
            dh2 = cadd(mk(-nz.re, -nz.im), rdiv(1.0, csub(nz, dh2)));
            nz = rdiv(N, Z[3]);
            dh4 = cadd(mk(-nz.re, -nz.im), rdiv(1.0, csub(nz, dh4)));
        }
        p24h24 = cdiv(p24h24, cmul(cadd(dh4, rdiv(N, Z[3])), cadd(W2, rdiv(N, Z[3]))));
        p24h21 = cdiv(p24h21, cmul(cadd(dh1, rdiv(N, Z[0])), cadd(W2, rdiv(N, Z[3]))));
        dummy = cdiv(cmul(dummy, cadd(acap, rdiv(N, Z[0]))), cadd(W2, rdiv(N, Z[3])));
        dumsq = cmul(dummy, dummy);
        coefs();
        S[4 * (size_t)N] = acoe;
        S[4 * (size_t)N + 1] = bcoe;
        are = acoe.re;
        aim = acoe.im;
        bre = bcoe.re;
        bim = bcoe.im;
        const double am1re = acoem1.re, am1im = acoem1.im, bm1re = bcoem1.re, bm1im = bcoem1.im;
        double t4 = (2 * N - 1.0) / (N * (N - 1.0));
        const double t2 = (N - 1.0) * (N + 1.0) / N;
        dgqsc = dgqsc + t2 * (am1re * are + am1im * aim + bm1re * bre + bm1im * bim) +
                t4 * (am1re * bm1re + am1im * bm1im);
        const double t3 = 2 * N + 1;
        dqext = dqext + t3 * (are + bre);
        t4 = are * are + aim * aim + bre * bre + bim * bim;
        dqsca = dqsca + t3 * t4;
        if (t4 >= 1e-14) {
            N = N + 1;
            if (N > q.nmx2) {
                err = 1;
                break;
            }
            acoem1 = acoe;
            bcoem1 = bcoe;
            continue;
        }
        break;
    }
    if (!err) {
        const double tq = 2.0 * (RX * RX);
        out.qext = (float)(tq * dqext);
        out.qsca = (float)(tq * dqsca);
        out.err = 0;
        out.nterms = N;
    }
    res[c] = out;
}

// M1, M2, S21, D21 of one call at the 90 cosines (k = 0) and their
// supplements (k = 1); out[c][4][90][2].  Lane l serves angles l and l + 64.
__global__ __launch_bounds__(64) void k_dhs_angles(const DhsCall *__restrict__ calls,
                                                   const DhsRes *__restrict__ res, int first, long long base,
                                                   const cxd *__restrict__ scr, const float *__restrict__ mu,
                                                   float *__restrict__ out) {
    const int c = first + blockIdx.x;
    const DhsRes r = res[c];
    if (r.err) return;
    const cxd *S = scr + 4 * (calls[c].off - base);
    const int N = r.nterms;
    const int lane = threadIdx.x;
    const int nj = lane + 64 < NH ? 2 : 1;
    double mj[2], si2[2], pi0[2], pi1[2], ta0[2], ta1[2];
    cxd s1[2][2], s2[2][2];
    const cxd ac1 = cscl(1.5, S[4]), bc1 = cscl(1.5, S[5]);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int j = lane + 64 * h;
        const float m = mu[j < NH ? j : 0];
        mj[h] = (double)m;
        si2[h] = 1.0 - (double)(m * m);
        pi0[h] = 0.0;
        pi1[h] = 1.0;
        ta0[h] = 0.0;
        ta1[h] = mj[h];
        s1[h][0] = cadd(cscl(pi1[h], ac1), cscl(ta1[h], bc1));
        s1[h][1] = csub(cscl(pi1[h], ac1), cscl(ta1[h], bc1));
        s2[h][0] = cadd(cscl(pi1[h], bc1), cscl(ta1[h], ac1));
        s2[h][1] = csub(cscl(pi1[h], bc1), cscl(ta1[h], ac1));
    }
    for (int n = 2; n <= N; n++) {
        const double T1 = 2 * n - 1, T2 = n - 1;
        const double tf = (double)(2 * n + 1) / (double)(n * (n + 1));
        const cxd ac = cscl(tf, S[4 * (size_t)n]), bc = cscl(tf, S[4 * (size_t)n + 1]);
        const bool even = (n % 2) == 0;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const double p3 = (T1 * pi1[h] * mj[h] - n * pi0[h]) / T2;
            const double t3 = mj[h] * (p3 - pi0[h]) - T1 * si2[h] * pi1[h] + ta0[h];
            s1[h][0] = cadd(cadd(s1[h][0], cscl(p3, ac)), cscl(t3, bc));
            s2[h][0] = cadd(cadd(s2[h][0], cscl(p3, bc)), cscl(t3, ac));
            if (even) {
                s1[h][1] = cadd(csub(s1[h][1], cscl(p3, ac)), cscl(t3, bc));
                s2[h][1] = cadd(csub(s2[h][1], cscl(p3, bc)), cscl(t3, ac));
            } else {
                s1[h][1] = csub(cadd(s1[h][1], cscl(p3, ac)), cscl(t3, bc));
                s2[h][1] = csub(cadd(s2[h][1], cscl(p3, bc)), cscl(t3, ac));
            }
            pi0[h] = pi1[h];
            pi1[h] = p3;
            ta0[h] = ta1[h];
            ta1[h] = t3;
        }
    }
    float *o = out + (size_t)c * 4 * NA;
#pragma unroll
    for (int h = 0; h < 2; h++) {
        if (h >= nj) break;
        const int j = lane + 64 * h;
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const cxd a = s1[h][k], b = s2[h][k];
            o[0 * NA + 2 * j + k] = (float)(a.re * a.re + a.im * a.im);
            o[1 * NA + 2 * j + k] = (float)(b.re * b.re + b.im * b.im);
            o[2 * NA + 2 * j + k] = (float)(a.re * b.re + a.im * b.im);
            o[3 * NA + 2 * j + k] = (float)(a.im * b.re - b.im * a.re);
        }
    }
}

// fichid + anbn + the cross-section sums of scatmat (ComputePart.f:2506-2580,
// 2211-2239, 2973-2988) for one radius; a_n, b_n to scratch up to nstop.
__global__ __launch_bounds__(64) void k_hom_coef(const HomItem *__restrict__ items, int first, int nitem,
                                                 double *__restrict__ scrd, cxd *__restrict__ scrc,
                                                 HomRes *__restrict__ res) {
    const int h = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= first + nitem) return;
    const HomItem it = items[h];
    double *psi = scrd + it.off_psi;
    cxd *D = scrc + it.off_D;
    cxd *ab = scrc + it.off_ab;
    const double x = it.x, perx = it.perx;
    const cxd m = mk(it.m_re, it.m_im);
    const cxd z = cscl(x, m), perz = rdiv(1.0, z);
    double pn = 0.0;
    psi[it.nfi] = 0.0;
    for (int n = it.nfi - 1; n >= 0; n--) {
        pn = 1.0 / ((double)(2 * n + 1) / x - pn);
        psi[n] = pn;
    }
    cxd dn = mk(0.0, 0.0);
    D[it.nD] = dn;
    for (int n = it.nD - 1; n >= 1; n--) {
        const cxd zn1 = cscl((double)(n + 1), perz);
        dn = csub(zn1, rdiv(1.0, cadd(dn, zn1)));
        D[n] = dn;
    }
    psi[0] = it.sinx;
    const double psi1 = psi[0] * perx - it.cosx;
    if (fabs(psi1) > 1e-4) {
        psi[1] = psi1;
        for (int n = 2; n <= it.nmax; n++) psi[n] = psi[n] * psi[n - 1];
    } else {
        for (int n = 1; n <= it.nmax; n++) psi[n] = psi[n] * psi[n - 1];
    }
    const cxd perm = rdiv(1.0, m);
    double chim1 = it.cosx, chin = chim1 * perx + it.sinx; // chi(0), chi(1)
    double csca = 0.0, cext = 0.0;
    int nstop = it.nmax;
    for (int n = 1; n <= it.nmax; n++) {
        const cxd zn = mk(psi[n], chin), znm1 = mk(psi[n - 1], chim1);
        const double xn = (double)n * perx;
        const cxd Dn = D[n];
        cxd sv = cadd(cmul(Dn, perm), mk(xn, 0.0));
        const cxd an = cdiv(csub(cscl(psi[n], sv), mk(psi[n - 1], 0.0)), csub(cmul(sv, zn), znm1));
        sv = cadd(cmul(m, Dn), mk(xn, 0.0));
        const cxd bn = cdiv(csub(cscl(psi[n], sv), mk(psi[n - 1], 0.0)), csub(cmul(sv, zn), znm1));
        ab[2 * (size_t)n] = an;
        ab[2 * (size_t)n + 1] = bn;
        const double aux =
            (2.0 * (double)n + 1.0) * fabs((an.re * an.re + an.im * an.im) + (bn.re * bn.re + bn.im * bn.im));
        csca = csca + aux;
        cext = cext + (2.0 * n + 1.0) * (an.re + bn.re);
        if (aux < 1e-8) {
            nstop = n;
            break;
        }
        if (n < it.nmax) {
            const double chn1 = (double)(2 * n + 1) * chin * perx - chim1;
            chim1 = chin;
            chin = chn1;
        }
    }
    HomRes o;
    o.csca = csca;
    o.cext = cext;
    o.nstop = nstop;
    o.pad = 0;
    res[h] = o;
}

// scatmat's symmetric angle loop (ComputePart.f:3094-3135) for one radius:
// out[h][4][180] = F11, F12, F33, F34 from the forward direction, as
// MeerhoffMie returns them; res gets Csca, Cext times lam^2/(2 pi).
__global__ __launch_bounds__(64) void k_hom_angles(const HomItem *__restrict__ items, int first,
                                                   const cxd *__restrict__ scrc, const double *__restrict__ u,
                                                   HomRes *__restrict__ res, double *__restrict__ out) {
    const int h = first + blockIdx.x;
    const HomItem it = items[h];
    const HomRes r = res[h];
    const cxd *ab = scrc + it.off_ab;
    const int lane = threadIdx.x;
    const int nj = lane + 64 < NH ? 2 : 1;
    double uu[2], pin[2], taun[2], pin1[2], taun1[2], dl[2];
    cxd spf[2], smf[2], spb[2], smb[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const int j = lane + 64 * q;
        uu[q] = u[j < NH ? j : 0];
        pin[q] = 1.0;
        pin1[q] = 3.0 * uu[q];
        dl[q] = 3.0 * uu[q] * uu[q] - 1.0;
        taun[q] = uu[q];
        taun1[q] = 2.0 * dl[q] - 1.0;
        spf[q] = smf[q] = spb[q] = smb[q] = mk(0.0, 0.0);
    }
    for (int n = 1; n <= r.nstop; n++) {
        const double facf = (double)(2 * n + 1) / (double)(n * (n + 1));
        const double facb = (n % 2 == 1) ? -facf : facf;
        const cxd an = ab[2 * (size_t)n], bn = ab[2 * (size_t)n + 1];
        const cxd apb = cadd(an, bn), amb = csub(an, bn);
        const cxd fpa = cscl(facf, apb), fma_ = cscl(facf, amb), bpa = cscl(facb, apb), bma = cscl(facb, amb);
#pragma unroll
        for (int q = 0; q < 2; q++) {
            spf[q] = cadd(spf[q], cscl(pin[q] + taun[q], fpa));
            smf[q] = cadd(smf[q], cscl(pin[q] - taun[q], fma_));
            spb[q] = cadd(spb[q], cscl(pin[q] - taun[q], bpa));
            smb[q] = cadd(smb[q], cscl(pin[q] + taun[q], bma));
            if (n == 1) {
                pin[q] = pin1[q];
                taun[q] = taun1[q];
            } else {
                const double pnext = (double)(n + 1) / (double)n * dl[q] + uu[q] * pin[q];
                dl[q] = uu[q] * pnext - pin[q];
                const double tnext = (double)(n + 1) * dl[q] - pin[q];
                pin[q] = pnext;
                taun[q] = tnext;
            }
        }
    }
    const double den = 2.0 * r.csca;
    double *o = out + (size_t)h * 4 * NA;
#pragma unroll
    for (int q = 0; q < 2; q++) {
        if (q >= nj) break;
        const int j = lane + 64 * q;
        const cxd cpf = mk(spf[q].re, -spf[q].im), cmf = mk(smf[q].re, -smf[q].im);
        const cxd cpb = mk(spb[q].re, -spb[q].im), cmb = mk(smb[q].re, -smb[q].im);
        // F(.,j) lies at u(j) (backward of 90 degrees): MeerhoffMie's index 179-j
        const double f1 = 0.0 + (cmul(spf[q], cpf).re + cmul(smf[q], cmf).re);
        const double f2 = 0.0 - (cmul(smf[q], cpf).re + cmul(spf[q], cmf).re);
        const double f3 = 0.0 + (cmul(spf[q], cpf).re - cmul(smf[q], cmf).re);
        const double f4 = 0.0 + -(cmul(smf[q], cpf).im - cmul(spf[q], cmf).im);
        const double b1 = 0.0 + (cmul(spb[q], cpb).re + cmul(smb[q], cmb).re);
        const double b2 = 0.0 - (cmul(smb[q], cpb).re + cmul(spb[q], cmb).re);
        const double b3 = 0.0 + (cmul(spb[q], cpb).re - cmul(smb[q], cmb).re);
        const double b4 = 0.0 + -(cmul(smb[q], cpb).im - cmul(spb[q], cmb).im);
        o[0 * NA + (NA - 1 - j)] = f1 / den;
        o[1 * NA + (NA - 1 - j)] = f2 / den;
        o[2 * NA + (NA - 1 - j)] = f3 / den;
        o[3 * NA + (NA - 1 - j)] = f4 / den;
        o[0 * NA + j] = b1 / den;
        o[1 * NA + j] = b2 / den;
        o[2 * NA + j] = b3 / den;
        o[3 * NA + j] = b4 / den;
    }
    __syncthreads();
    if (lane == 0) {
        HomRes w = r;
        w.csca = r.csca * it.fakt;
        w.cext = r.cext * it.fakt;
        res[h] = w;
    }
}

struct SumArgs {
    const Term *terms;
    const int *toff;           // [nblock+1]
    const DhsRes *dres;
    const float *dout;         // [call][4][180]
    const HomRes *hres;
    const double *hout;        // [item][4][180]
    const double *sinj;        // [180] sin(pi (j - 0.5)/180)
    const double *mscale;      // [nspec]
    double sin0;
    float tot2;
    int nlam;
    double *Kabs, *Ksca, *Kext, *F, *rho; // [nspec*nlam], F [nspec*nlam][6][180], rho [nspec]
};

// ComputePart's sums for one (species, wavelength): thread j owns angle j;
// thread 0 also carries the cross sections, the mass and the volume.
__global__ __launch_bounds__(192) void k_part_sum(SumArgs A) {
    __shared__ double sM[NA], sSin[NA];
    __shared__ float sCsca;
    const int b = blockIdx.x;
    const int j = threadIdx.x;
    const bool act = j < NA;
    if (act) sSin[j] = A.sinj[j];
    double mf[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double csmie = 0.0, cemie = 0.0;
    float acc[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    float csca = 0.0f, cabs = 0.0f, cext = 0.0f, mass = 0.0f, vol = 0.0f;
    __syncthreads();
    for (int t = A.toff[b]; t < A.toff[b + 1]; t++) {
        const Term T = A.terms[t];
        if (T.kind == 0) {
            const DhsRes r = A.dres[T.idx];
            cemie = (double)r.qext * DPI * (double)T.rad2;
            csmie = (double)r.qsca * DPI * (double)T.rad2;
            if (act) {
                const double w2 = (double)T.w2;
                const int jj = j < NH ? j : NA - 1 - j, k = j < NH ? 0 : 1;
                const float *o = A.dout + (size_t)T.idx * 4 * NA + 2 * jj + k;
                const float m1 = o[0], m2 = o[NA], s21 = o[2 * NA], d21 = o[3 * NA];
                const double sp = (double)(m2 + m1), sm = (double)(m2 - m1);
                mf[0] = sp / csmie / w2 * 2.0 * DPI;
                mf[1] = sm / csmie / w2 * 2.0 * DPI;
                mf[2] = sp / csmie / w2 * 2.0 * DPI;
                mf[3] = (double)s21 / csmie / w2 * 2.0 * DPI;
                mf[4] = (double)(-d21) / csmie / w2 * 2.0 * DPI;
                mf[5] = (double)s21 / csmie / w2 * 2.0 * DPI;
            }
        } else {
            if (T.kind == 1) {
                const HomRes r = A.hres[T.idx];
                csmie = r.csca;
                cemie = r.cext;
                if (act) {
                    const double *o = A.hout + (size_t)T.idx * 4 * NA;
                    mf[0] = o[j];
                    mf[1] = o[NA + j];
                    mf[3] = o[2 * NA + j];
                    mf[4] = o[3 * NA + j];
                }
            }
            mf[2] = mf[0];
            mf[5] = mf[3];
        }
        if (act) sM[j] = mf[0];
        __syncthreads();
        if (j == 0) {
            // the forward-peak correction that normalises F11 (REAL sums)
            float tot = 0.0f;
            for (int a = 0; a < NA; a++) tot = (float)((double)tot + sM[a] * sSin[a]);
            double adj = sM[0] + (double)(A.tot2 - tot) / A.sin0;
            if (adj < 0.0) adj = 0.0;
            mf[0] = adj;
        }
        if (act) {
            const double wn = (double)T.wn;
#pragma unroll
            for (int q = 0; q < 6; q++) acc[q] = (float)((double)acc[q] + wn * mf[q] * csmie);
        }
        if (j == 0) {
            const double wn = (double)T.wn;
            cext = (float)((double)cext + wn * cemie);
            csca = (float)((double)csca + wn * csmie);
            cabs = (float)((double)cabs + wn * (cemie - csmie));
            mass = (float)((double)mass + (double)T.wnrho * 4.0 * DPI * (double)T.r3 / 3.0);
            vol = (float)((double)vol + wn * 4.0 * DPI * (double)T.r3 / 3.0);
        }
        __syncthreads();
    }
    const int s = b / A.nlam, il = b % A.nlam;
    if (j == 0) {
        sCsca = csca;
        const double ms = A.mscale[s];
        A.Kabs[b] = 1e4 * (double)cabs / ((double)mass / ms);
        A.Ksca[b] = 1e4 * (double)csca / ((double)mass / ms);
        A.Kext[b] = 1e4 * (double)cext / ((double)mass / ms);
        if (il == A.nlam - 1) A.rho[s] = (double)(mass / vol);
    }
    __syncthreads();
    if (act) {
#pragma unroll
        for (int q = 0; q < 6; q++) A.F[((size_t)b * 6 + q) * NA + j] = (double)(acc[q] / sCsca);
    }
}

struct TmpList {
    std::vector<void *> v;
    ~TmpList() { free_list(v); }
};

// gauleg2 (ComputePart.f:1218-1250)
void gauleg2(float x1, float x2, float *x, float *w, int n) {
    const int m = (n + 1) / 2;
    const double xm = 0.5 * ((double)x2 + (double)x1), xl = 0.5 * ((double)x2 - (double)x1);
    for (int i = 1; i <= m; i++) {
        double z = cos(3.141592654 * (i - 0.25) / (n + 0.5)), z1, pp;
        do {
            double p1 = 1.0, p2 = 0.0, p3;
            for (int j = 1; j <= n; j++) {
                p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            z1 = z;
            z = z1 - p1 / pp;
        } while (fabs(z - z1) > 3e-14);
        x[i - 1] = (float)(xm - xl * z);
        x[n - i] = (float)(xm + xl * z);
        w[i - 1] = (float)(2.0 * xl / ((1.0 - z * z) * pp * pp));
        w[n - i] = w[i - 1];
    }
}

// one (species, wavelength, material, size) of ComputePart's loops
struct Group {
    int s, il, l, k;
    int first, ncall; // its layered-sphere calls (one per fraction) unless solid or too large
    bool hom_only;    // fmax = 0 or x > 10^4: homogeneous spheres throughout
    int item;         // homogeneous item, -1 if none
    float r1, wvno;
};

} // namespace

extern "C" int mcx_compute_parts(mcx_ctx *ctx, int32_t nlam, const double *lam, int32_t nspec,
                                 const mcx_part_spec *specs, mcx_part_out *out, mcx_part_stats *stats) {
    if (!ctx || !lam || !specs || !out || nlam < 1 || nspec < 1) return MCX_EINVAL;
    for (int il = 0; il < nlam; il++)
        if (!(lam[il] > 0.0)) return fail(ctx, MCX_EINVAL, "wavelengths must be positive");
    for (int s = 0; s < nspec; s++) {
        const mcx_part_spec &p = specs[s];
        if (p.nm < 1 || p.ns < 1 || p.nf < 1 || !p.e1 || !p.e2 || !p.frac || !p.rho)
            return fail(ctx, MCX_EINVAL, "species " + std::to_string(s) + ": nm, ns, nf >= 1 and all arrays needed");
        if (!(p.amin > 0.0 && p.amax >= p.amin && p.fmax >= 0.0 && p.fmax < 1.0 && p.rscale > 0.0 && p.mscale > 0.0))
            return fail(ctx, MCX_EINVAL, "species " + std::to_string(s) + ": need 0 < amin <= amax, 0 <= fmax < 1");
        if (p.nf > 1 && (float)p.fmax > 0.0f && (float)p.fmax <= 0.01f)
            return fail(ctx, MCX_EINVAL, "species " + std::to_string(s) + ": nf > 1 needs fmax > 0.01 (ComputePart.f:289-297)");
        for (size_t q = 0; q < (size_t)p.nm * nlam; q++)
            if (!(p.e1[q] > 0.0f && p.e2[q] >= 0.0f))
                return fail(ctx, MCX_EINVAL, "species " + std::to_string(s) + ": refractive index needs n > 0, k >= 0");
        if (!out[s].Kabs || !out[s].Ksca || !out[s].Kext || !out[s].F || !out[s].rho) return MCX_EINVAL;
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    HIPCHK(ctx, hipEventRecord(ctx->ev[0], st));
    TmpList tmp;
    int rc;

    // angles and weights as ComputePart sets them (ComputePart.f:311-314, 380-386)
    float mu[NH];
    for (int i = 1; i <= NH; i++) {
        const double theta = (double)(((float)i - 0.5f) / (float)NH * 3.1415926536f) / 2.0;
        mu[i - 1] = (float)cos(theta);
    }
    double sinj[NA];
    float tot2 = 0.0f;
    for (int j = 1; j <= NA; j++) {
        sinj[j - 1] = sin(DPI * (double)((float)j - 0.5f) / (double)(float)NA);
        tot2 = (float)((double)tot2 + sinj[j - 1]);
    }
    const double sin0 = sin(DPI * (double)0.5f / (double)(float)NA);
    // MeerhoffMie's cosines (ComputePart.f:1895-1899, 3070-3075)
    const double pie = acos(-1.0), radfac = pie / 180.0;
    const double thmin = 180.0 * (double)(1.0f - 0.5f) / (double)(float)NA;
    const double thmax = 180.0 * (double)((float)NA - 0.5f) / (double)(float)NA;
    const double step = (thmax - thmin) / (double)(float)(NA - 1);
    double u[NA];
    for (int iang = 1; iang <= NA; iang++) u[NA - iang] = cos(radfac * (thmin + (double)(iang - 1) * step));

    // ---- enumerate ComputePart's loops ------------------------------------
    std::vector<DhsCall> calls;
    std::vector<Group> groups;
    struct Species {
        std::vector<float> f, wf, r, nr;
        int nf;
    };
    std::vector<Species> sp(nspec);
    long long slots = 0;
    for (int s = 0; s < nspec; s++) {
        const mcx_part_spec &p = specs[s];
        Species &S = sp[s];
        const float minlog = (float)log10(p.amin), maxlog = (float)log10(p.amax), pw = (float)(-p.apow);
        const float maxf = (float)p.fmax;
        int nf = maxf == 0.0f ? 1 : p.nf;
        S.nf = nf;
        S.f.assign(nf, 0.0f);
        S.wf.assign(nf, 0.0f);
        if (nf > 1 && maxf > 0.01f) {
            gauleg2(0.01f, maxf, S.f.data(), S.wf.data(), nf);
        } else if (maxf == 0.0f) {
            S.f[0] = 0.0f;
            S.wf[0] = (float)(1.0 / (double)(float)nf);
        } else {
            S.f[0] = maxf;
            S.wf[0] = 1.0f;
        }
        const int ns = p.ns;
        S.r.assign(ns, 0.0f);
        S.nr.assign((size_t)p.nm * ns, 0.0f);
        for (int l = 0; l < p.nm; l++) {
            if (ns == 1) {
                S.r[0] = (float)pow(10.0, (double)(minlog + maxlog) / 2.0);
                S.nr[(size_t)l * ns] = p.frac[l];
            } else {
                float tot = 0.0f;
                for (int k = 0; k < ns; k++) {
                    const float e = minlog + (maxlog - minlog) * (float)k / (float)(ns - 1);
                    S.r[k] = (float)pow(10.0, (double)e);
                    S.nr[(size_t)l * ns + k] = (float)pow((double)S.r[k], (double)pw + 1.0);
                    tot = tot + S.nr[(size_t)l * ns + k] * (S.r[k] * S.r[k] * S.r[k]);
                }
                for (int k = 0; k < ns; k++) S.nr[(size_t)l * ns + k] = p.frac[l] * S.nr[(size_t)l * ns + k] / tot;
            }
        }
        for (int il = 0; il < nlam; il++) {
            const float wvno = (float)(2.0 * (double)3.1415926536f / lam[il]);
            for (int l = 0; l < p.nm; l++) {
                if (p.frac[l] == 0.0f) continue;
                const float m_re = p.e1[(size_t)l * nlam + il], m_im = -p.e2[(size_t)l * nlam + il];
                for (int k = 0; k < ns; k++) {
                    Group g{};
                    g.s = s;
                    g.il = il;
                    g.l = l;
                    g.k = k;
                    g.item = -1;
                    g.r1 = (float)((double)S.r[k] * p.rscale);
                    g.wvno = wvno;
                    g.first = (int)calls.size();
                    g.hom_only = S.f[0] == 0.0f || (double)(g.r1 * wvno) > 10000.0;
                    if (!g.hom_only) {
                        for (int i = 0; i < nf; i++) {
                            const float rad = (float)((double)g.r1 / pow(1.0 - (double)S.f[i], 1.0 / 3.0));
                            const float rcore = (float)((double)rad * pow((double)S.f[i], 1.0 / 3.0));
                            DhsCall c{};
                            c.wvno = wvno;
                            c.sh_re = m_re;
                            c.sh_im = m_im;
                            c.co_re = 1.0f;
                            c.co_im = 0.0f;
                            c.xshell = (double)(rad * wvno);
                            c.xcore = (double)(rcore * wvno);
                            const double t1 = c.xshell * (double)hypotf(m_re, m_im);
                            int nmx1 = (int)(1.1 * t1), nmx2 = (int)t1;
                            if (nmx1 <= 150) {
                                nmx1 = 150;
                                nmx2 = 135;
                            }
                            const bool bad = wvno <= 0.0f || rad <= 0.0f || rcore <= 0.0f || rcore > rad ||
                                             m_re <= 0.0f || m_im > 0.0f || nmx1 + 1 > 200000;
                            c.nmx1 = bad ? 0 : nmx1;
                            c.nmx2 = nmx2;
                            const double X1 = (double)m_re * c.xshell, Y1 = (double)m_im * c.xshell;
                            const double X4 = (double)m_re * c.xcore, Y4 = (double)m_im * c.xcore;
                            c.cosx = cos(c.xshell);
                            c.sinx = sin(c.xshell);
                            c.sinx1 = sin(X1);
                            c.cosx1 = cos(X1);
                            c.sinx4 = sin(X4);
                            c.cosx4 = cos(X4);
                            c.ey1 = exp(Y1);
                            c.ey4 = exp(Y4);
                            c.ey1my4 = exp(Y1 - Y4);
                            c.off = slots;
                            if (!bad) slots += nmx1 + 2;
                            calls.push_back(c);
                        }
                    }
                    g.ncall = (int)calls.size() - g.first;
                    groups.push_back(g);
                }
            }
        }
    }
    const int ncall = (int)calls.size();

    // ---- homogeneous spheres known before any layered result --------------
    // The too-large (x > 10^4) and solid grains take MeerhoffMie throughout
    // (ComputePart.f:316-330): a few hundred long serial chains on a few
    // CUs, run on a second stream of default priority while the layered
    // spheres use the rest of the device.
    auto make_item = [&](const Group &g, long long &offd, long long &offc, HomItem &it) -> int {
        const mcx_part_spec &p = specs[g.s];
        const double rmie = (double)g.r1, lmie0 = lam[g.il];
        const double lmie = (rmie / lmie0 < 5000.0) ? lmie0 : rmie / 5000.0;
        it.m_re = (double)p.e1[(size_t)g.l * nlam + g.il];
        it.m_im = -(double)p.e2[(size_t)g.l * nlam + g.il];
        const double rtox = 2.0 * pie / lmie;
        it.x = rtox * rmie;
        it.fakt = lmie * lmie / (2.0 * pie);
        it.nmax = (int)(it.x + 4.05 * pow(it.x, 1.0 / 3.0) + 20.0);
        it.nfi = it.nmax + 60;
        const double zabs = it.x * hypot(it.m_re, it.m_im);
        it.nD = (int)(zabs + 4.05 * pow(zabs, 1.0 / 3.0) + 70.0);
        if (it.nD < it.nmax)
            return fail(ctx, MCX_EINVAL, "species " + std::to_string(g.s) +
                                             ": |m| x too small for the homogeneous-sphere series (ComputePart.f:2946-2950)");
        it.perx = 1.0 / it.x;
        it.sinx = sin(it.x);
        it.cosx = cos(it.x);
        it.off_psi = offd;
        offd += it.nfi + 1;
        it.off_D = offc;
        offc += it.nD + 1;
        it.off_ab = offc;
        offc += 2 * ((long long)it.nmax + 1);
        return 0;
    };
    struct Side {
        hipStream_t st = nullptr;
        hipEvent_t ev = nullptr;
        ~Side() {
            if (st) (void)hipStreamSynchronize(st);
            if (ev) (void)hipEventDestroy(ev);
            if (st) (void)hipStreamDestroy(st);
        }
    } side;
    std::vector<HomItem> items;
    long long offd_a = 0, offc_a = 0;
    for (Group &g : groups) {
        if (!g.hom_only) continue;
        HomItem it{};
        if ((rc = make_item(g, offd_a, offc_a, it))) return rc;
        g.item = (int)items.size();
        items.push_back(it);
    }
    // capacity for every group that could need a homogeneous sphere
    const size_t item_cap = std::max<size_t>(groups.size(), 1);
    HomItem *ditems = nullptr;
    HomRes *dhres = nullptr;
    double *dhout = nullptr, *du = nullptr;
    if ((rc = dalloc(ctx, tmp.v, &ditems, item_cap))) return rc;
    if ((rc = dalloc(ctx, tmp.v, &dhres, item_cap))) return rc;
    if ((rc = dalloc(ctx, tmp.v, &dhout, item_cap * 4 * NA))) return rc;
    if ((rc = dupload(ctx, tmp.v, &du, u, NA))) return rc;
    size_t max_scratch = 0;
    if (!items.empty()) {
        const int n = (int)items.size();
        HIPCHK(ctx, hipMemcpyAsync(ditems, items.data(), sizeof(HomItem) * n, hipMemcpyHostToDevice, st));
        double *scrd = nullptr;
        cxd *scrc = nullptr;
        if ((rc = dalloc(ctx, tmp.v, &scrd, (size_t)offd_a))) return rc;
        if ((rc = dalloc(ctx, tmp.v, &scrc, (size_t)offc_a))) return rc;
        max_scratch = (size_t)offd_a * 8 + (size_t)offc_a * 16;
        HIPCHK(ctx, hipStreamCreate(&side.st));
        HIPCHK(ctx, hipEventCreateWithFlags(&side.ev, hipEventDisableTiming));
        // the side stream starts after the uploads on st
        HIPCHK(ctx, hipEventRecord(side.ev, st));
        HIPCHK(ctx, hipStreamWaitEvent(side.st, side.ev, 0));
        k_hom_coef<<<(n + 63) / 64, 64, 0, side.st>>>(ditems, 0, n, scrd, scrc, dhres);
        HIPCHK(ctx, hipGetLastError());
        k_hom_angles<<<n, 64, 0, side.st>>>(ditems, 0, scrc, du, dhres, dhout);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipEventRecord(side.ev, side.st));
    }

    // ---- layered spheres, in batches that fit the scratch budget ----------
    size_t freeb = 0, totb = 0;
    HIPCHK(ctx, hipMemGetInfo(&freeb, &totb));
    const size_t budget = std::min<size_t>(freeb / 2, (size_t)24 << 30);
    DhsCall *dcalls = nullptr;
    DhsRes *dres = nullptr;
    float *dout = nullptr;
    float *dmu = nullptr;
    if ((rc = dupload(ctx, tmp.v, &dcalls, calls.data(), calls.size()))) return rc;
    if ((rc = dalloc(ctx, tmp.v, &dres, calls.size()))) return rc;
    if ((rc = dalloc(ctx, tmp.v, &dout, calls.size() * 4 * NA))) return rc;
    if ((rc = dupload(ctx, tmp.v, &dmu, mu, NH))) return rc;
    cxd *scr = nullptr;
    size_t scr_bytes = 0;
    for (int c0 = 0; c0 < ncall;) {
        int c1 = c0;
        const long long base = calls[c0].off;
        while (c1 < ncall) {
            const long long need = calls[c1].off + (calls[c1].nmx1 > 0 ? calls[c1].nmx1 + 2 : 0) - base;
            if (c1 > c0 && (size_t)need * 64 > budget) break;
            c1++;
        }
        const long long end = calls[c1 - 1].off + (calls[c1 - 1].nmx1 > 0 ? calls[c1 - 1].nmx1 + 2 : 0);
        const size_t bytes = (size_t)std::max<long long>(end - base, 1) * 64;
        if (bytes > scr_bytes) {
            HIPCHK(ctx, hipStreamSynchronize(st));
            if (scr) {
                tmp.v.erase(std::find(tmp.v.begin(), tmp.v.end(), (void *)scr));
                (void)hipFree(scr);
            }
            scr = nullptr;
            scr_bytes = 0;
            if ((rc = dalloc(ctx, tmp.v, (char **)&scr, bytes))) return rc;
            scr_bytes = bytes;
        }
        max_scratch = std::max(max_scratch, bytes);
        const int n = c1 - c0;
        k_dhs_coef<<<(n + 63) / 64, 64, 0, st>>>(dcalls, c0, n, base, scr, dres);
        HIPCHK(ctx, hipGetLastError());
        k_dhs_angles<<<n, 64, 0, st>>>(dcalls, dres, c0, base, scr, dmu, dout);
        HIPCHK(ctx, hipGetLastError());
        c0 = c1;
    }
    std::vector<DhsRes> hres_d(calls.size());
    if (ncall) HIPCHK(ctx, hipMemcpyAsync(hres_d.data(), dres, sizeof(DhsRes) * ncall, hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    if (scr) {
        tmp.v.erase(std::find(tmp.v.begin(), tmp.v.end(), (void *)scr));
        (void)hipFree(scr);
    }
    scr = nullptr;

    // ---- homogeneous spheres of the layered-sphere failures ---------------
    // (the ones of the too-large and solid grains already run on st2)
    long long offd_b = 0, offc_b = 0;
    const int nA = (int)items.size();
    for (Group &g : groups) {
        if (g.hom_only) continue;
        bool need = false;
        for (int i = 0; i < g.ncall && !need; i++) need = hres_d[g.first + i].err != 0;
        if (!need) continue;
        HomItem it{};
        if ((rc = make_item(g, offd_b, offc_b, it))) return rc;
        g.item = (int)items.size();
        items.push_back(it);
    }
    const int nitem = (int)items.size();
    if (nitem > nA) {
        HIPCHK(ctx, hipMemcpyAsync(ditems + nA, items.data() + nA, sizeof(HomItem) * (nitem - nA),
                                   hipMemcpyHostToDevice, st));
        double *scrd = nullptr;
        cxd *scrc = nullptr;
        if ((rc = dalloc(ctx, tmp.v, &scrd, (size_t)offd_b))) return rc;
        if ((rc = dalloc(ctx, tmp.v, &scrc, (size_t)offc_b))) return rc;
        max_scratch = std::max(max_scratch, (size_t)offd_b * 8 + (size_t)offc_b * 16);
        k_hom_coef<<<(nitem - nA + 63) / 64, 64, 0, st>>>(ditems, nA, nitem - nA, scrd, scrc, dhres);
        HIPCHK(ctx, hipGetLastError());
        k_hom_angles<<<nitem - nA, 64, 0, st>>>(ditems, nA, scrc, du, dhres, dhout);
        HIPCHK(ctx, hipGetLastError());
    }
    if (nA) HIPCHK(ctx, hipStreamWaitEvent(st, side.ev, 0));

    // ---- the terms of ComputePart's sums, in its order ------------------
    const int nblock = nspec * nlam;
    std::vector<Term> terms;
    std::vector<int> toff(nblock + 1, 0);
    long long ndhs = 0, nhom = 0, nerr = 0, dterms = 0;
    for (int c = 0; c < ncall; c++)
        if (!hres_d[c].err) dterms += hres_d[c].nterms;
    {
        size_t gi = 0;
        for (int b = 0; b < nblock; b++) {
            toff[b] = (int)terms.size();
            while (gi < groups.size() && groups[gi].s * nlam + groups[gi].il == b) {
                const Group &g = groups[gi++];
                const mcx_part_spec &p = specs[g.s];
                const Species &S = sp[g.s];
                bool err = false;
                for (int i = 0; i < S.nf; i++) {
                    Term T{};
                    T.wn = S.wf[i] * S.nr[(size_t)g.l * p.ns + g.k];
                    T.wnrho = T.wn * p.rho[g.l];
                    T.r3 = g.r1 * g.r1 * g.r1;
                    T.w2 = g.wvno * g.wvno;
                    if (g.hom_only) {
                        T.kind = i == 0 ? 1 : 2;
                        T.idx = g.item;
                    } else {
                        const DhsRes &r = hres_d[g.first + i];
                        if (r.err) err = true;
                        if (err) {
                            T.kind = 1;
                            T.idx = g.item;
                            nerr++;
                        } else {
                            T.kind = 0;
                            T.idx = g.first + i;
                            const float rad = (float)((double)g.r1 / pow(1.0 - (double)S.f[i], 1.0 / 3.0));
                            T.rad2 = rad * rad;
                        }
                    }
                    if (T.kind == 0) ndhs++;
                    if (T.kind == 1) nhom++;
                    terms.push_back(T);
                }
            }
        }
        toff[nblock] = (int)terms.size();
    }
    Term *dterms_ = nullptr;
    int *dtoff = nullptr;
    double *dsin = nullptr, *dms = nullptr, *dK = nullptr, *dF = nullptr, *drho = nullptr;
    std::vector<double> msc(nspec);
    for (int s = 0; s < nspec; s++) msc[s] = specs[s].mscale;
    if ((rc = dupload(ctx, tmp.v, &dterms_, terms.data(), terms.size()))) return rc;
    if ((rc = dupload(ctx, tmp.v, &dtoff, toff.data(), toff.size()))) return rc;
    if ((rc = dupload(ctx, tmp.v, &dsin, sinj, NA))) return rc;
    if ((rc = dupload(ctx, tmp.v, &dms, msc.data(), msc.size()))) return rc;
    if ((rc = dalloc(ctx, tmp.v, &dK, (size_t)3 * nblock))) return rc;
    if ((rc = dalloc(ctx, tmp.v, &dF, (size_t)nblock * 6 * NA))) return rc;
    if ((rc = dalloc(ctx, tmp.v, &drho, (size_t)nspec))) return rc;
    SumArgs A{};
    A.terms = dterms_;
    A.toff = dtoff;
    A.dres = dres;
    A.dout = dout;
    A.hres = dhres;
    A.hout = dhout;
    A.sinj = dsin;
    A.mscale = dms;
    A.sin0 = sin0;
    A.tot2 = tot2;
    A.nlam = nlam;
    A.Kabs = dK;
    A.Ksca = dK + nblock;
    A.Kext = dK + 2 * nblock;
    A.F = dF;
    A.rho = drho;
    k_part_sum<<<nblock, 192, 0, st>>>(A);
    HIPCHK(ctx, hipGetLastError());
    std::vector<double> K((size_t)3 * nblock), Fh((size_t)nblock * 6 * NA), rho(nspec);
    HIPCHK(ctx, hipMemcpyAsync(K.data(), dK, sizeof(double) * K.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipMemcpyAsync(Fh.data(), dF, sizeof(double) * Fh.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipMemcpyAsync(rho.data(), drho, sizeof(double) * nspec, hipMemcpyDeviceToHost, st));
    long long hterms = 0;
    if (nitem) {
        std::vector<HomRes> hr(nitem);
        HIPCHK(ctx, hipMemcpyAsync(hr.data(), dhres, sizeof(HomRes) * nitem, hipMemcpyDeviceToHost, st));
        HIPCHK(ctx, hipStreamSynchronize(st));
        for (const HomRes &r : hr) hterms += r.nstop;
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev[1], st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    for (int s = 0; s < nspec; s++) {
        for (int il = 0; il < nlam; il++) {
            const int b = s * nlam + il;
            out[s].Kabs[il] = K[b];
            out[s].Ksca[il] = K[nblock + b];
            out[s].Kext[il] = K[2 * nblock + b];
        }
        memcpy(out[s].F, Fh.data() + (size_t)s * nlam * 6 * NA, sizeof(double) * nlam * 6 * NA);
        out[s].rho[0] = rho[s];
    }
    if (stats) {
        float ms = 0.0f;
        HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
        stats->ndhs = ndhs;
        stats->nhom = nhom;
        stats->nerr = nerr;
        stats->dhs_terms = dterms;
        stats->hom_terms = hterms;
        stats->scratch_bytes = (int64_t)max_scratch;
        stats->ms_total = ms;
    }
    return MCX_OK;
}
