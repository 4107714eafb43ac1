// sed.hip -- the ray-traced SED of MCMax's Observe/SPECTRUM branch on the
// MI355X (Telescope.f:138-164 -> TracePath, TraceMono.f:2032-2293 -> TraceFlux,
// TraceMono.f:8-578), behind mcx_sed() of include/mcx.h.
//
// Per observation the image rays are traced once through the grid's (sub)cells
// (k_paths, one thread per ray, segments stored flat in HBM); the rays do not
// depend on wavelength.  Per wavelength:
//  * k_mono: monochromatic opacities, albedo and thermal source of every cell,
//    and the EmissionDistribution with its escape optical depths (one thread
//    per cell, tracing the cell's nexits rays);
//  * k_mix: the Mueller matrix and its phi integrals per species mix;
//  * k_scan: the emission cumulative, in the reference's cell order;
//  * k_mc: the scattering Monte Carlo (MakeStarScatter and TraceMono), one
//    wavefront per packet: the lanes follow the packet together and split the
//    180 observer directions of every scattered-field deposit between them;
//  * k_rays: the formal solution along every image ray (one thread per ray).
// The flux quadrature over the image rings runs on the host.
#include <chrono>

#include "mcx_internal.h"

#define NPHIS 180 // observer directions of the scattered field (TraceMono.f:1309)

// Per-call temporaries of the SED and image calls come from the stream-ordered
// pool (hipMallocAsync / hipFreeAsync on the context's stream): a call makes
// about forty of them, and hipFree would synchronise the device for each.
// Every host read of their contents follows a stream synchronisation.
template <typename T>
static int palloc(mcx_ctx *ctx, std::vector<void *> &list, T **p, size_t n) {
    void *q = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMallocAsync(&q, n * sizeof(T), ctx->stream);
    if (e != hipSuccess)
        return fail(ctx, MCX_ENOMEM, "hipMallocAsync of " + std::to_string(n * sizeof(T)) + " bytes failed");
    list.push_back(q);
    *p = (T *)q;
    return 0;
}

template <typename T>
static int pupload(mcx_ctx *ctx, std::vector<void *> &list, T **p, const T *h, size_t n) {
    int r = palloc(ctx, list, p, n);
    if (r) return r;
    HIPCHK(ctx, hipMemcpyAsync(*p, h, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return 0;
}

static void pfree_list(mcx_ctx *ctx, std::vector<void *> &list) {
    for (void *q : list) (void)hipFreeAsync(q, ctx->stream);
    list.clear();
}
#define NSTAR_RAYS 30

struct SedGeom {
    int nr, nt, NT, NRG, MAXRG;
    const int *nrg, *rgoff;   // [nt]
    const double *thetrg;     // [nt][MAXRG+1] cos^2 of sub-cell edges
    const double *thet;       // [nt+1] cone angles
    const double *R2, *T2, *R, *Theta;
};

struct SedMono {
    int l1, l2; // 0-based
    double wl1, wl2;
    const double *Kabs, *Kext, *Ksca, *Alb; // [ncell]
    const int *cls;                         // [ncell] species-mix class
    const double *F;                        // [nclass][6][180]
    const double *MIF11, *MIF12;            // [nclass]
};

struct SedObsDev {
    const double *xsf, *ysf, *zsf, *xin, *yin, *zin; // [NPHIS]
    const double *pisin;                             // [180] pi sin(theta_a)
    int usepol, scat_how;
    double pfstop;
};

// TraceMono.f:2487-2653, the cell exit with theta sub-cells
__device__ bool exit_rg(const SedGeom &G, double x, double y, double z, double vx, double vy, double vz, int i,
                        int jj, int irg, bool onEdge, int &edgeNr, double &v, int &inext, int &jnext,
                        int &irgnext) {
    const int NT = G.NT;
    double r = x * x + y * y + z * z;
    double R1s = G.R2[i], R2s = G.R2[i + 1];
    int j = jj + 1;
    int n = G.nrg[jj];
    double T1s, T2s;
    if (n <= 1 || j == 0 || j == NT) {
        T1s = G.T2[jj];
        T2s = G.T2[jj + 1];
    } else {
        T1s = G.thetrg[(size_t)jj * (G.MAXRG + 1) + irg - 1];
        T2s = G.thetrg[(size_t)jj * (G.MAXRG + 1) + irg];
    }
    double b = 2.0 * (x * vx + y * vy + z * vz);
    bool h1, h2, h3, h4;
    double vR1, vR2, vT1, vT2;
    if (!onEdge) {
        h1 = hit_sphere(R1s, r, b, vR1);
        h2 = hit_sphere(R2s, r, b, vR2);
        h3 = hit_cone(T1s, r, b, z, vz, vT1);
        h4 = hit_cone(T2s, r, b, z, vz, vT2);
    } else if (edgeNr == 1) {
        h1 = false;
        vR1 = 1e200;
        h2 = hit_sphere(R2s, r, b, vR2);
        h3 = hit_cone(T1s, r, b, z, vz, vT1);
        h4 = hit_cone(T2s, r, b, z, vz, vT2);
    } else if (edgeNr == 2) {
        h1 = hit_sphere(R1s, r, b, vR1);
        h2 = true;
        vR2 = -b;
        h3 = hit_cone(T1s, r, b, z, vz, vT1);
        h4 = hit_cone(T2s, r, b, z, vz, vT2);
    } else if (edgeNr == 3) {
        h1 = hit_sphere(R1s, r, b, vR1);
        h2 = hit_sphere(R2s, r, b, vR2);
        h3 = false;
        vT1 = 1e200;
        if (j != NT - 1 || irg != n) {
            h4 = hit_cone(T2s, r, b, z, vz, vT2);
        } else {
            vT2 = -z / vz;
            h4 = (vT2 > 0.0);
        }
    } else {
        h1 = hit_sphere(R1s, r, b, vR1);
        h2 = hit_sphere(R2s, r, b, vR2);
        if (j != 1 || irg != 1) {
            h3 = hit_cone(T1s, r, b, z, vz, vT1);
        } else {
            h3 = false;
            vT1 = 1e200;
        }
        if (j != NT - 1 || irg != n) {
            h4 = hit_cone_same(T2s, b, z, vz, vT2);
        } else {
            h4 = false;
            vT2 = 1e200;
        }
    }
    if (!h1 && !h2 && !h3 && !h4) return false;
    v = 1e200;
    int in = i, jn = j, rn = irg;
    if (h1 && vR1 < v) { v = vR1; in = i - 1; jn = j; rn = irg; edgeNr = 2; }
    if (h2 && vR2 < v) { v = vR2; in = i + 1; jn = j; rn = irg; edgeNr = 1; }
    if (h3 && vT1 < v) {
        v = vT1;
        in = i;
        if (n > 1) {
            rn = irg - 1;
            jn = j;
            if (rn < 1) {
                jn = j - 1;
                rn = (jn >= 1) ? G.nrg[jn - 1] : 1;
            }
        } else {
            jn = j - 1;
            if (jn > 0) rn = G.nrg[jn - 1];
        }
        edgeNr = 4;
    }
    if (h4 && vT2 < v) {
        v = vT2;
        in = i;
        if (n > 1) {
            rn = irg + 1;
            jn = j;
            if (rn > n) {
                jn = j + 1;
                rn = 1;
                if (jn == NT) rn = G.nrg[NT - 2];
            }
        } else {
            jn = j + 1;
            rn = 1;
        }
        edgeNr = 3;
    }
    if (jn == NT) { jn = NT - 1; rn = G.nrg[NT - 2]; edgeNr = 4; }
    if (jn == 0) { jn = 1; rn = 1; edgeNr = 3; }
    inext = in;
    jnext = jn - 1;
    irgnext = rn;
    return true;
}

__host__ __device__ inline int irg_of(const int *nrg, const double *thet, int jj, double theta) {
    int n = nrg[jj];
    if (n <= 1) return 1;
    int irg = (int)((double)(float)n * (theta - thet[jj]) / (thet[jj + 1] - thet[jj])) + 1;
    if (irg > n) irg = n;
    if (irg < 1) irg = 1;
    return irg;
}

// TraceMono.f:2034-2050: azimuth bin of a point for the scattered field
__device__ inline int jphi_of(double x, double y) {
    double phi = (180.0 * acos(x / sqrt(x * x + y * y))) / MCX_PI;
    if (y < 0.0) phi = -phi;
    int jp = (int)(((double)(float)NPHIS * phi) / 360.0);
    jp = jp + 1;
    if (jp < 1) jp = 1 - jp;
    if (jp > NPHIS / 2) jp = NPHIS + 1 - jp;
    return jp;
}

__host__ __device__ inline size_t sf_idx(int NRG, const int *rgoff, int i, int jj, int irg, int a, int side) {
    return ((((size_t)i * NRG + rgoff[jj] + (irg - 1)) * 91 + a) * 2 + (side - 1));
}

// ---------------------------------------------------------------------------
// image rays (TracePath start TraceMono.f:2241-2270, trace2dpath 2350-2430).
// Pass 0 counts segments, pass 1 stores them at off[ray].
__device__ void rot_y(double &x, double &y, double &z, double ang) {
    rotate_about(x, y, z, 0.0, 1.0, 0.0, ang);
}

__global__ void k_paths(SedGeom G, const double *__restrict__ Rimg, int nimg, int nphi, double angle, int pass,
                        int *count, const long long *__restrict__ off, int *seg_c, unsigned char *seg_irg,
                        unsigned char *seg_k, double *seg_v, unsigned char *seg_j1, unsigned char *seg_j2,
                        int *hitstar) {
    const int ray = blockIdx.x * blockDim.x + threadIdx.x;
    if (ray >= nimg * nphi) return;
    const int ir = ray / nphi, k = ray % nphi + 1;
    const double Rad = Rimg[ir];
    const double phi = 1e-5 + MCX_PI * ((double)((float)k - 0.5f)) / (double)nphi;
    const double Rout = G.R[G.nr];
    double x = Rad * cos(phi), y = Rad * sin(phi), z = sqrt(Rout * Rout - Rad * Rad);
    int edgeNr = 2;
    bool onEdge = true;
    rot_y(x, y, z, angle);
    double ct = fabs(z) / Rout;
    int i = G.nr - 1, jj = 0;
    for (int j = 1; j <= G.NT - 1; j++)
        if (ct < G.Theta[j - 1] && ct >= G.Theta[j]) jj = j - 1;
    int irg = irg_of(G.nrg, G.thet, jj, acos(ct));
    double vx = 0.0, vy = 0.0, vz = -1.0;
    rot_y(vx, vy, vz, angle);
    int n = 0, hs = 0;
    long long base = (pass == 1) ? off[ray] : 0;
    for (int guard = 0; guard < 1000000; guard++) {
        double v;
        int in, jn, rn;
        if (!exit_rg(G, x, y, z, vx, vy, vz, i, jj, irg, onEdge, edgeNr, v, in, jn, rn)) {
            hs = -1;
            break;
        }
        onEdge = true;
        const int j1 = jphi_of(x, y);
        const double x2 = x + vx * v, y2 = y + vy * v, z2 = z + vz * v;
        const int kk = (z2 > 0.0) ? 1 : 2;
        const int j2 = jphi_of(x2, y2);
        if (pass == 1) {
            const long long s = base + n;
            seg_c[s] = i * G.nt + jj;
            seg_irg[s] = (unsigned char)irg;
            seg_k[s] = (unsigned char)kk;
            seg_v[s] = v;
            seg_j1[s] = (unsigned char)j1;
            seg_j2[s] = (unsigned char)j2;
        }
        n++;
        x = x + vx * v;
        y = y + vy * v;
        z = z + vz * v;
        if (in >= G.nr) break;
        if (in < 0) {
            hs = 1;
            break;
        }
        i = in;
        jj = jn;
        irg = rn;
    }
    if (pass == 0) count[ray] = n;
    else hitstar[ray] = hs;
}

// ---------------------------------------------------------------------------
// monochromatic cells (TraceMono.f:649-754), thermal source (229-305) and
// EmissionDistribution (956-1119) with trace2exit (2659-2692)
__global__ void k_mono(int nr, int nt, int nlam, int ng, SedMono M, const double *__restrict__ Kabs_s,
                       const double *__restrict__ Kext_s, const double *__restrict__ Ksca_s,
                       const double *__restrict__ w, double *Kabs, double *Kext, double *Ksca, double *Alb) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nr * nt) return;
    double ka = 0.0, ke = 0.0, ks = 0.0;
    for (int s = 0; s < ng; s++) {
        const size_t o1 = (size_t)s * nlam + M.l1, o2 = (size_t)s * nlam + M.l2;
        const double ws = w[(size_t)c * ng + s];
        ka = ka + ((M.wl1 * Kabs_s[o1] + M.wl2 * Kabs_s[o2]) * ws) * 1.0;
        ke = ke + ((M.wl1 * Kext_s[o1] + M.wl2 * Kext_s[o2]) * ws) * 1.0;
        ks = ks + ((M.wl1 * Ksca_s[o1] + M.wl2 * Ksca_s[o2]) * ws) * 1.0;
    }
    Kabs[c] = ka;
    Kext[c] = ke;
    Ksca[c] = ks;
    Alb[c] = ks / ke;
}

__global__ void k_emis(SedGeom G, int nlam, int ng, SedMono M, const double *__restrict__ Kabs_s,
                       const double *__restrict__ w, const double *__restrict__ BB, const double *__restrict__ Tc,
                       const double *__restrict__ dens, const double *__restrict__ V, int nexits, int traceemis,
                       double *emis, double *Eem, double *EmisDis, double *vismass) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int nt = G.nt;
    if (c >= G.nr * nt) return;
    const int i = c / nt, jj = c % nt;
    const int l1 = M.l1, l2 = M.l2;
    const double wl1 = M.wl1, wl2 = M.wl2;
    // B(T) kappa summed over species, with and without the per-cell Kabs
    auto bk = [&](bool per_kabs) {
        double e = 0.0;
        const double T = Tc[c];
        int iT = (int)(T / 1.0);
        const double wT1 = (double)(float)(iT + 1) - T / 1.0;
        const double wT2 = T / 1.0 - (double)(float)iT;
        if (iT < 1) iT = 1;
        if (iT > TMAXD - 1) iT = TMAXD - 1;
        const double *B0 = BB + (size_t)iT * nlam, *B1 = BB + (size_t)(iT + 1) * nlam;
        for (int s = 0; s < ng; s++) {
            const double *K = Kabs_s + (size_t)s * nlam;
            const double t = ((wl1 * (wT1 * B0[l1] + wT2 * B1[l1]) * K[l1] + wl2 * (wT1 * B0[l2] + wT2 * B1[l2]) * K[l2]) *
                              w[(size_t)c * ng + s]) * 1.0;
            e = per_kabs ? (e + t / M.Kabs[c]) : (e + t);
        }
        return e;
    };
    // source function (TraceMono.f:229-305)
    double e = 0.0;
    if (traceemis) e = bk(true) * (1.0 - M.Alb[c]);
    emis[c] = e;
    if (i == 0) {
        Eem[c] = 0.0;
        EmisDis[c] = 0.0;
        vismass[c] = 0.0;
        return;
    }
    double ee = (Tc[c] != 0.0) ? bk(false) : 0.0;
    ee = (ee * dens[c]) * V[c];
    double taumin = 0.0;
    const int nex = nexits;
    for (int k = 1; k <= abs(nex); k++) {
        double Rad = 0.5;
        Rad = sqrt(G.R[i] * G.R[i] * Rad + G.R[i + 1] * G.R[i + 1] * (1.0 - Rad));
        double Th = 0.5;
        Th = G.Theta[jj] * Th + G.Theta[jj + 1] * (1.0 - Th);
        double z = Rad * Th;
        double phi = 1e-4;
        double x = Rad * sqrt(1.0 - Th * Th) * sin(phi);
        double y = Rad * sqrt(1.0 - Th * Th) * cos(phi);
        phi = (MCX_PI * (double)(k - 1)) / (double)(nex - 1);
        const double vx = 0.0, vy = cos(phi), vz = sin(phi);
        int qi = i, qj = jj, edgeNr = 0;
        bool onEdge = false;
        double tau = 0.0;
        for (int guard = 0; guard < 1000000; guard++) {
            double v;
            int in, jn;
            if (!exit_face(G.R2, G.T2, nt, x, y, z, vx, vy, vz, qi, qj, onEdge, edgeNr, v, in, jn)) break;
            onEdge = true;
            x = x + vx * v;
            y = y + vy * v;
            z = z + vz * v;
            const int cq = qi * nt + qj;
            tau = tau + ((v * dens[cq]) * MCX_AU) * M.Kabs[cq];
            if (in >= G.nr || in <= 0) break;
            qi = in;
            qj = jn;
        }
        if (tau < taumin || k == 1) taumin = tau;
    }
    const double tauexit = (nex == 0) ? 0.0 : taumin;
    const double vm = (nex >= 0) ? exp(-tauexit) : (1.0 - exp(-tauexit));
    vismass[c] = vm;
    Eem[c] = ee;
    EmisDis[c] = ee * vm;
}

// per species mix: Ksca-weighted Mueller matrix and its phi integrals
// (TraceMono.f:706-748); one thread per (class, element, angle)
__global__ void k_mix(int nclass, int nlam, int ng, SedMono M, const double *__restrict__ wcls,
                      const double *__restrict__ F_s, const double *__restrict__ Ksca_s,
                      const double *__restrict__ IF11_s, const double *__restrict__ IF12_s, int nel, double *F,
                      double *MIF11, double *MIF12) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nclass * 6 * 180) return;
    const int k = idx / (6 * 180), e = (idx / 180) % 6, a = idx % 180;
    const double *w = wcls + (size_t)k * ng;
    double ks = 0.0;
    for (int s = 0; s < ng; s++) {
        const size_t o1 = (size_t)s * nlam + M.l1, o2 = (size_t)s * nlam + M.l2;
        ks = ks + ((M.wl1 * Ksca_s[o1] + M.wl2 * Ksca_s[o2]) * w[s]) * 1.0;
    }
    double *Fk = F + (size_t)k * 6 * 180;
    if (ks == 0.0) {
        Fk[e * 180 + a] = (e == 0) ? 1.0 : 0.0;
        if (e == 0 && a == 0) {
            MIF11[k] = 0.0;
            MIF12[k] = 0.0;
        }
        return;
    }
    if (e < nel) {
        double acc = 0.0;
        for (int s = 0; s < ng; s++) {
            const size_t o1 = (size_t)s * nlam + M.l1, o2 = (size_t)s * nlam + M.l2;
            const double ml = (F_s[(o1 * 6 + e) * 180 + a] * Ksca_s[o1]) * M.wl1 +
                              (F_s[(o2 * 6 + e) * 180 + a] * Ksca_s[o2]) * M.wl2;
            acc = acc + (ml * w[s]) * 1.0;
        }
        Fk[e * 180 + a] = acc / ks;
    } else {
        Fk[e * 180 + a] = 0.0;
    }
    if (e == 0 && a == 0) {
        double i11 = 0.0, i12 = 0.0;
        for (int s = 0; s < ng; s++) {
            const size_t o1 = (size_t)s * nlam + M.l1, o2 = (size_t)s * nlam + M.l2;
            const double ml11 = (IF11_s[o1] * Ksca_s[o1]) * M.wl1 + (IF11_s[o2] * Ksca_s[o2]) * M.wl2;
            const double ml12 = (IF12_s[o1] * Ksca_s[o1]) * M.wl1 + (IF12_s[o2] * Ksca_s[o2]) * M.wl2;
            i11 = i11 + (ml11 * w[s]) * 1.0;
            i12 = i12 + (ml12 * w[s]) * 1.0;
        }
        MIF11[k] = i11 / ks;
        MIF12[k] = i12 / ks;
    }
}

// the emission cumulative over cells in the reference's order (i outer, j
// inner, TraceMono.f:1206-1212), and the totals Etot, Etot2.  The sums stay
// sequential (one thread adds, in cell order, so the cumulative is the
// reference's to the last bit); the block stages tiles of the inputs and
// the results through LDS so the adding thread never waits on HBM.
#define SCAN_TILE 1024
__global__ __launch_bounds__(256) void k_scan(int ncell, int nt, const double *__restrict__ Eem,
                                              const double *__restrict__ EmisDis, double *__restrict__ cum,
                                              double *__restrict__ tot) {
    __shared__ double se[SCAN_TILE], sd[SCAN_TILE], sc[SCAN_TILE];
    double Et = 0.0, E1 = 0.0, E2 = 0.0;
    for (int base = nt; base < ncell; base += SCAN_TILE) {
        const int n = min(SCAN_TILE, ncell - base);
        for (int k = threadIdx.x; k < n; k += blockDim.x) {
            se[k] = Eem[base + k];
            sd[k] = EmisDis[base + k];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int k = 0; k < n; k++) {
                E1 = E1 + se[k];
                E2 = E2 + sd[k];
                Et = Et + sd[k];
                sc[k] = Et;
            }
        }
        __syncthreads();
        for (int k = threadIdx.x; k < n; k += blockDim.x) cum[base + k] = sc[k];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        tot[0] = E1;
        tot[1] = E2;
    }
}

// ---------------------------------------------------------------------------
// scattering Monte Carlo: one wavefront per packet
struct McPhot {
    double x, y, z, vx, vy, vz, E, Q, U, V, Sx, Sy, Sz;
    int i, jj, irg, edgeNr;
    bool onEdge, scatt, pol;
};

struct McArgs {
    SedGeom G;
    SedMono M;
    SedObsDev O;
    const double *dens, *V, *cum, *tot, *vismass;
    double *sf, *sQ, *sU, *sV;
    const double *cos2phi, *sin2phi;
    double R0, Estar;
    int nphot_star, nphot_therm, ncell;
    unsigned long long seed;
};

// per observer direction (LDS): scattering-angle bin, rotated Stokes vector
// and the rotation to the image frame (TraceMono.f:1309-1335, 1646-1669)
struct ObsDirs {
    int ia[NPHIS];
    double E[NPHIS], Q[NPHIS], U[NPHIS], V[NPHIS], s2[NPHIS], c2[NPHIS];
};

__device__ void rotate_stokes(McPhot &p, double x, double y, double z) {
    double P = (p.U / p.E) * (p.U / p.E) + (p.Q / p.E) * (p.Q / p.E);
    if (P == 0.0) return;
    double cost = p.Sx * x + p.Sy * y + p.Sz * z;
    double sint = (x - (p.Sx * (p.vy * p.vy + p.vz * p.vz) - p.vx * (p.vy * p.Sy + p.vz * p.Sz)) * cost) /
                  (p.vy * p.Sz - p.vz * p.Sy);
    double cos2t = cost * cost - sint * sint;
    double sin2t = (2.0 * sint) * cost;
    double Q = p.Q * cos2t + p.U * sin2t;
    double U = -p.Q * sin2t + p.U * cos2t;
    double P2 = (U / p.E) * (U / p.E) + (Q / p.E) * (Q / p.E);
    double sPP2 = sqrt(P / P2);
    p.U = sPP2 * U;
    p.Q = sPP2 * Q;
}

// every lane fills its share of the 180 directions
__device__ void setup_obsdirs(const McArgs &A, const McPhot &p, bool rotate_pol, ObsDirs &od, int lane) {
    for (int i = lane; i < NPHIS; i += 64) {
        const double xs = A.O.xsf[i], ys = A.O.ysf[i], zs = A.O.zsf[i];
        od.ia[i] = (int)((180.0 * acos(p.vx * xs + p.vy * ys + p.vz * zs)) / MCX_PI + 1);
        double xn = p.vy * zs - p.vz * ys;
        double yn = p.vz * xs - p.vx * zs;
        double zn = p.vx * ys - p.vy * xs;
        double r = sqrt(xn * xn + yn * yn + zn * zn);
        xn /= r;
        yn /= r;
        zn /= r;
        McPhot q = p;
        if (rotate_pol && p.pol) rotate_stokes(q, xn, yn, zn);
        q.Sx = xn;
        q.Sy = yn;
        q.Sz = zn;
        q.vx = xs;
        q.vy = ys;
        q.vz = zs;
        // make_rotate_stokes (TraceMono.f:1561-1580)
        const double x = A.O.xin[i], y = A.O.yin[i], z = A.O.zin[i];
        double cost = q.Sx * x + q.Sy * y + q.Sz * z;
        double sint = (x - (q.Sx * (q.vy * q.vy + q.vz * q.vz) - q.vx * (q.vy * q.Sy + q.vz * q.Sz)) * cost) /
                      (q.vy * q.Sz - q.vz * q.Sy);
        double c2 = cost * cost - sint * sint;
        double s2 = (2.0 * sint) * cost;
        double een = sqrt(s2 * s2 + c2 * c2);
        od.s2[i] = s2 / een;
        od.c2[i] = c2 / een;
        od.E[i] = q.E;
        od.Q[i] = q.Q;
        od.U[i] = q.U;
        od.V[i] = q.V;
    }
    __syncthreads();
}

// angular part of makescattfield (TraceMono.f:1440-1509) and MakeStarScatter
// (1700-1750; star: F11/F12 only); lanes split the 180 directions
__device__ void add_angles(const McArgs &A, const McPhot &p, double v, int irg, int side, const ObsDirs &od,
                           bool star, int lane) {
    const int c = p.i * A.G.nt + p.jj;
    const double *F = A.M.F + (size_t)A.M.cls[c] * 6 * 180;
    const double x = p.x + p.vx * v / 2.0, y = p.y + p.vy * v / 2.0;
    double phi0 = (180.0 * acos(x / sqrt(x * x + y * y))) / MCX_PI;
    if (y < 0.0) phi0 = 360.0 - phi0;
    const double phi = 180.0 / (double)(float)NPHIS;
    const int j0 = (int)(((double)(float)NPHIS * (phi + phi0)) / 360.0);
    const double vAU = (v * MCX_AU) / 2.0;
    const bool usepol = A.O.usepol;
    for (int i = 1 + lane; i <= NPHIS; i += 64) {
        // the reference increments j from j0 and wraps it into 1..NPHIS
        int j = j0 + (i - 1);
        while (j < 1) j += NPHIS;
        while (j > NPHIS) j -= NPHIS;
        const int ia = od.ia[j - 1];
        const double F11 = F[ia - 1], F12 = F[180 + ia - 1];
        const double E1 = od.E[j - 1], Q1 = od.Q[j - 1], U1 = od.U[j - 1], V1 = od.V[j - 1];
        double sI, sQ = 0.0, sU = 0.0, sV = 0.0;
        if (star) {
            sI = F11 * E1;
            if (usepol) {
                const double Qt = F12 * E1;
                sQ = Qt * od.c2[j - 1];
                sU = -Qt * od.s2[j - 1];
            }
        } else {
            sI = F11 * E1 + F12 * Q1;
            if (usepol) {
                const double F22 = F[360 + ia - 1], F33 = F[540 + ia - 1], F34 = F[720 + ia - 1],
                             F44 = F[900 + ia - 1];
                const double Qt = F12 * E1 + F22 * Q1;
                const double Ut = F34 * V1 + F33 * U1;
                sQ = Qt * od.c2[j - 1] + Ut * od.s2[j - 1];
                sU = -Qt * od.s2[j - 1] + Ut * od.c2[j - 1];
                sV = -F34 * U1 + F44 * V1;
            }
        }
        const int a = (i <= NPHIS / 2) ? i : NPHIS + 1 - i;
        const double sgn = (i <= NPHIS / 2) ? 1.0 : -1.0;
        const size_t k = sf_idx(A.G.NRG, A.G.rgoff, p.i, p.jj, irg, a, side);
        atomicAdd(&A.sf[k], sI * vAU);
        if (usepol) {
            atomicAdd(&A.sQ[k], sQ * vAU);
            atomicAdd(&A.sU[k], sgn * sU * vAU);
            if (!star) atomicAdd(&A.sV[k], sgn * sV * vAU);
        }
    }
}

__device__ void star_position(const McArgs &A, PhiloxStream &g, McPhot &p) {
    random_direction(g, p.x, p.y, p.z);
    p.x = A.R0 * p.x;
    p.y = A.R0 * p.y;
    p.z = A.R0 * p.z;
    random_direction(g, p.vx, p.vy, p.vz);
    const double inp = p.x * p.vx + p.y * p.vy + p.z * p.vz;
    if (inp < 0.0) {
        p.vx = -p.vx;
        p.vy = -p.vy;
        p.vz = -p.vz;
    }
    p.edgeNr = 1;
    p.onEdge = true;
}

// scatangle with the cell's mixed Mueller matrix (Radiation.f:381-495), the
// theta cumulative summed bin by bin as the reference does
__device__ void scatter_mix(const McArgs &A, PhiloxStream &g, McPhot &p, int c) {
    const int k = A.M.cls[c];
    const double *Fm = A.M.F + (size_t)k * 6 * 180;
    const double MIF11 = A.M.MIF11[k], MIF12 = A.M.MIF12[k];
    const double *F11 = Fm, *F12 = Fm + 180, *F22 = Fm + 360, *F33 = Fm + 540, *F34 = Fm + 720, *F44 = Fm + 900;
    const double IF11 = MIF11 * 180.0, IF12 = MIF12 * 180.0;
    double phi;
    if (p.pol) {
        double Itot = 0.0;
        for (int i = 1; i <= 180; i++) Itot = Itot + IF11 * p.E + IF12 * (p.Q * A.cos2phi[i] + p.U * A.sin2phi[i]);
        Itot = g.next() * Itot;
        double I1 = 0.0, I0 = 0.0;
        phi = MCX_PI;
        for (int i = 1; i <= 180; i++) {
            I0 = I0 + IF11 * p.E + IF12 * (p.Q * A.cos2phi[i] + p.U * A.sin2phi[i]);
            if (I0 > Itot) {
                const double wp1 = (I0 - Itot) / (I0 - I1);
                phi = (MCX_PI * ((double)i - wp1)) / 180.0;
                break;
            }
            I1 = I0;
        }
    } else {
        phi = g.next() * MCX_PI;
    }
    double x = p.Sx, y = p.Sy, z = p.Sz;
    rotate_about(x, y, z, p.vx, p.vy, p.vz, phi);
    if (p.pol) rotate_stokes(p, x, y, z);
    p.Sx = x;
    p.Sy = y;
    p.Sz = z;
    double Fr = 180.0 * (p.E * MIF11 + p.Q * MIF12);
    Fr = g.next() * Fr;
    double Fi = 0.0, FiOld = 0.0, theta = MCX_PI;
    int i;
    bool found = false;
    for (i = 1; i <= 180; i++) {
        Fi = Fi + A.O.pisin[i - 1] * (F11[i - 1] * p.E + F12[i - 1] * p.Q);
        if (Fi > Fr) {
            theta = (double)i - (Fi - Fr) / (Fi - FiOld);
            theta = (theta * MCX_PI) / 180.0;
            found = true;
            break;
        }
        FiOld = Fi;
    }
    if (!found) {
        theta = MCX_PI;
        i = 180;
    }
    if (g.next() < 0.5) theta = -theta;
    double E = 0.0, Q = 0.0, U = 0.0, V = 0.0;
    E = E + F11[i - 1] * p.E;
    E = E + F12[i - 1] * p.Q;
    Q = Q + F12[i - 1] * p.E;
    Q = Q + F22[i - 1] * p.Q;
    U = U + F33[i - 1] * p.U;
    const double Fs = (theta >= 0.0) ? fabs(F34[i - 1]) : -fabs(F34[i - 1]);
    U = U + Fs * p.V;
    V = V - Fs * p.U;
    V = V + F44[i - 1] * p.V;
    if (E != 0.0) {
        p.Q = (p.E * Q) / E;
        p.U = (p.E * U) / E;
        p.V = (p.E * V) / E;
    }
    rotate_about(p.vx, p.vy, p.vz, p.Sx, p.Sy, p.Sz, theta);
    double r = sqrt(p.vx * p.vx + p.vy * p.vy + p.vz * p.vz);
    p.vx /= r;
    p.vy /= r;
    p.vz /= r;
    r = sqrt(p.Sx * p.Sx + p.Sy * p.Sy + p.Sz * p.Sz);
    p.Sx /= r;
    p.Sy /= r;
    p.Sz /= r;
    if (i != 1 && i != 180) p.pol = true;
}

// theta cell of a point on the star (strictly inside; MakeStarScatter and
// emitposition keep the previous cell otherwise)
__device__ inline bool star_cell(const SedGeom &G, double ct, int &jj) {
    for (int j = 1; j <= G.NT - 1; j++)
        if (ct < G.Theta[j - 1] && ct > G.Theta[j]) {
            jj = j - 1;
            return true;
        }
    return false;
}

// MakeStarScatter (TraceMono.f:1585-1773) for packet ip, and the TraceMono
// packet loop (583-910: emitposition 1124, trace2dmono 1291) for the thermal
// packets; block = one wavefront = one packet
__global__ void __launch_bounds__(64) k_mc(McArgs A) {
    __shared__ ObsDirs od;
    const int lane = threadIdx.x;
    const int ip = blockIdx.x;
    const SedGeom &G = A.G;
    const int nt = G.nt;
    const bool angular = (A.O.scat_how != 1);
    PhiloxStream g;
    g.init(A.seed, (uint64_t)ip);
    McPhot p;
    memset(&p, 0, sizeof(p));
    if (ip < A.nphot_star) {
        p.E = A.Estar / (double)(float)A.nphot_star;
        p.scatt = false;
        star_position(A, g, p);
        const double ct = fabs(p.z) / A.R0;
        p.i = 0;
        (void)star_cell(G, ct, p.jj);
        p.irg = irg_of(G.nrg, G.thet, p.jj, acos(ct));
        if (angular) setup_obsdirs(A, p, false, od, lane);
        for (int guard = 0; guard < 10000000; guard++) {
            double v;
            int in, jn, rn;
            if (!exit_rg(G, p.x, p.y, p.z, p.vx, p.vy, p.vz, p.i, p.jj, p.irg, p.onEdge, p.edgeNr, v, in, jn, rn))
                break;
            const int c = p.i * nt + p.jj;
            const double tau = (A.dens[c] * A.M.Kext[c]) * MCX_AU;
            double w = ((tau * v) > 1e-5) ? (1.0 - exp(-tau * v)) / tau : v;
            const int irg = p.irg < 1 ? 1 : p.irg;
            const int side = (p.z > 0.0) ? 1 : 2;
            if (!angular) {
                if (lane == 0) {
                    atomicAdd(&A.sf[sf_idx(G.NRG, G.rgoff, p.i, p.jj, 1, 0, 1)], ((p.E * w) * MCX_AU) / 2.0);
                    atomicAdd(&A.sf[sf_idx(G.NRG, G.rgoff, p.i, p.jj, 1, 0, 2)], ((p.E * w) * MCX_AU) / 2.0);
                }
            } else {
                if (lane == 0) atomicAdd(&A.sf[sf_idx(G.NRG, G.rgoff, p.i, p.jj, irg, 0, side)], (p.E * w) * MCX_AU);
                add_angles(A, p, v, irg, side, od, true, lane);
            }
            w = exp(-tau * v);
            p.E = p.E * w;
            if (angular) {
                __syncthreads();
                for (int k = lane; k < NPHIS; k += 64) od.E[k] = od.E[k] * w;
                __syncthreads();
            }
            p.x = p.x + p.vx * v;
            p.y = p.y + p.vy * v;
            p.z = p.z + p.vz * v;
            p.onEdge = true;
            if (in >= G.nr || in < 0) break;
            p.i = in;
            p.jj = jn;
            p.irg = rn;
        }
        return;
    }
    // ---- a thermal packet: emitposition (star and thermal emission only)
    const double Etot2 = A.tot[1];
    for (;;) {
        p.pol = false;
        const double Etot = Etot2 + A.Estar + 0.0 + 0.0;
        double Er = Etot * g.next();
        if (Er < A.Estar) {
            p.E = Etot;
            p.scatt = false;
            star_position(A, g, p);
            const double ct = fabs(p.z) / A.R0;
            int jj;
            if (star_cell(G, ct, jj)) {
                p.i = 0;
                p.jj = jj;
                const double th = acos(fabs(p.z) / sqrt(p.x * p.x + p.y * p.y + p.z * p.z));
                p.irg = irg_of(G.nrg, G.thet, p.jj, th);
                break;
            }
        } else {
            p.scatt = true;
            Er = Er - A.Estar - 0.0 - 0.0;
            // first cell whose running emission sum exceeds Er (TraceMono.f:1206-1212)
            int lo = nt, hi = A.ncell;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (Er < A.cum[mid]) hi = mid;
                else lo = mid + 1;
            }
            if (lo < A.ncell) {
                const int c = lo, i = c / nt, jj = c % nt;
                double Rad = g.next();
                Rad = sqrt(G.R[i] * G.R[i] * Rad + G.R[i + 1] * G.R[i + 1] * (1.0 - Rad));
                double Th = g.next();
                Th = G.Theta[jj] * Th + G.Theta[jj + 1] * (1.0 - Th);
                p.z = Rad * Th;
                if (g.next() < 0.5) p.z = -p.z;
                const double phi = (g.next() * MCX_PI) * 2.0;
                p.x = Rad * sqrt(1.0 - Th * Th) * sin(phi);
                p.y = Rad * sqrt(1.0 - Th * Th) * cos(phi);
                random_direction(g, p.vx, p.vy, p.vz);
                p.i = i;
                p.jj = jj;
                const double th = acos(fabs(p.z) / sqrt(p.x * p.x + p.y * p.y + p.z * p.z));
                p.irg = irg_of(G.nrg, G.thet, jj, th);
                p.onEdge = false;
                p.E = Etot / A.vismass[c];
                break;
            }
        }
    }
    p.E = p.E / (double)(float)A.nphot_therm;
    if (A.O.scat_how == 2) {
        p.pol = false;
        p.Q = 0.0;
        p.U = 0.0;
        p.V = 0.0;
        double x, y, z;
        if (p.vz != 1.0) {
            x = -p.vy;
            y = p.vx;
            z = 0.0;
        } else {
            x = 1.0;
            y = 0.0;
            z = 0.0;
        }
        const double r = sqrt(x * x + y * y + z * z);
        p.Sx = x / r;
        p.Sy = y / r;
        p.Sz = z / r;
    }
    const int maxvisit = (G.nr + G.NT + 360) * 2;
    for (int guard = 0; guard < 100000000; guard++) {
        const double tau0 = -log(g.next());
        // trace2dmono with makescattfield
        bool escape = false;
        {
            double tau = 0.0;
            if (angular && p.scatt) setup_obsdirs(A, p, true, od, lane);
            int nvisit = 0;
            for (;;) {
                nvisit++;
                if (nvisit > maxvisit) {
                    escape = true;
                    break;
                }
                double v;
                int in, jn, rn;
                if (!exit_rg(G, p.x, p.y, p.z, p.vx, p.vy, p.vz, p.i, p.jj, p.irg, p.onEdge, p.edgeNr, v, in, jn,
                             rn)) {
                    escape = true;
                    break;
                }
                const int c = p.i * nt + p.jj;
                const double tau1 = ((v * A.dens[c]) * A.M.Kext[c]) * MCX_AU;
                const bool inside = (tau + tau1) > tau0;
                if (inside) v = (v * (tau0 - tau)) / tau1;
                if (p.scatt) {
                    if (!angular) {
                        if (lane == 0) {
                            atomicAdd(&A.sf[sf_idx(G.NRG, G.rgoff, p.i, p.jj, 1, 0, 1)], ((p.E * v) * MCX_AU) / 2.0);
                            atomicAdd(&A.sf[sf_idx(G.NRG, G.rgoff, p.i, p.jj, 1, 0, 2)], ((p.E * v) * MCX_AU) / 2.0);
                        }
                    } else {
                        const int irg = p.irg < 1 ? 1 : p.irg;
                        const int side = (p.z > 0.0) ? 1 : 2;
                        if (lane == 0)
                            atomicAdd(&A.sf[sf_idx(G.NRG, G.rgoff, p.i, p.jj, irg, 0, side)], (p.E * v) * MCX_AU);
                        add_angles(A, p, v, irg, side, od, false, lane);
                    }
                }
                p.x = p.x + p.vx * v;
                p.y = p.y + p.vy * v;
                p.z = p.z + p.vz * v;
                if (inside) {
                    p.onEdge = false;
                    break;
                }
                p.onEdge = true;
                tau = tau + tau1;
                if (in >= G.nr || in < 0) {
                    escape = true;
                    break;
                }
                p.i = in;
                p.jj = jn;
                p.irg = rn;
            }
        }
        if (escape) break;
        const int c = p.i * nt + p.jj;
        const double fstop = pow(A.M.Alb[c], A.O.pfstop);
        const double fact = A.M.Alb[c] / fstop;
        p.E = p.E * fact;
        if (A.O.scat_how == 2) {
            p.Q *= fact;
            p.U *= fact;
            p.V *= fact;
        }
        if (g.next() > fstop) break;
        p.scatt = true;
        if (A.O.scat_how == 1) random_direction(g, p.vx, p.vy, p.vz);
        else scatter_mix(A, g, p, c);
        __syncthreads(); // od is rebuilt by the next trace
    }
}

// ---------------------------------------------------------------------------
// formal solution along the image rays (TraceMono.f:332-493), one thread per
// ray; stars contribute Fstar/(pi R*^2) behind the ray (486-488)
struct RayArgs {
    int nt, NRG, MAXRG, scat_how, mc, tracestar;
    const int *rgoff;
    const long long *off;
    const int *cnt, *seg_c, *hitstar;
    const unsigned char *seg_irg, *seg_k, *seg_j1, *seg_j2;
    const double *seg_v, *fracirg, *dens, *V, *Kext, *Alb, *emis, *sf, *sQ, *sU, *sV;
    double Istar;
    double *image, *scatim, *imageQ, *imageU, *imageV; // U, V may be null (SED)
    int nray;
};

__global__ void k_rays(RayArgs A) {
    const int ray = blockIdx.x * blockDim.x + threadIdx.x;
    if (ray >= A.nray) return;
    const long long s0 = A.off[ray];
    const int n = A.cnt[ray];
    double img = 0.0, sim = 0.0, imQ = 0.0, imU = 0.0, imV = 0.0, tau = 0.0, fact = 1.0;
    const bool uv = A.imageU != nullptr && A.scat_how == 2;
    bool stopped = false;
    const double tau_max = 1e8;
    for (int s = 0; s < n; s++) {
        const long long q = s0 + s;
        const int c = A.seg_c[q], ip = c / A.nt, jp = c % A.nt, irg = A.seg_irg[q], kp = A.seg_k[q];
        double x_scat = 0.0, x_scatQ = 0.0, x_scatU = 0.0, x_scatV = 0.0;
        if (A.mc) {
            const int jj1 = A.seg_j1[q], jj2 = A.seg_j2[q];
            const int djj = (jj1 <= jj2) ? 1 : -1;
            int njj = 0, jj = jj1;
            const double ww = A.Alb[c] / (A.V[c] * A.fracirg[(size_t)jp * (A.MAXRG + 1) + irg - 1]);
            for (;;) {
                const size_t k = sf_idx(A.NRG, A.rgoff, ip, jp, irg, jj, kp);
                x_scat = x_scat + ww * A.sf[k];
                if (A.scat_how == 2) x_scatQ = x_scatQ + ww * A.sQ[k];
                if (uv) {
                    x_scatU = x_scatU + ww * A.sU[k];
                    x_scatV = x_scatV + ww * A.sV[k];
                }
                njj++;
                if (jj == jj2) break;
                jj = jj + djj;
            }
            x_scat = x_scat / (double)(float)njj;
            x_scatQ = x_scatQ / (double)(float)njj;
            x_scatU = x_scatU / (double)(float)njj;
            x_scatV = x_scatV / (double)(float)njj;
        }
        const double tau_e = ((A.seg_v[q] * A.dens[c]) * A.Kext[c]) * MCX_AU;
        if (ip != 0) {
            if (tau_e < 1e-6) {
                img = img + ((2.0 * x_scat + A.emis[c]) * tau_e) * fact;
                imQ = imQ + ((2.0 * x_scatQ) * tau_e) * fact;
                imU = imU + ((2.0 * x_scatU) * tau_e) * fact;
                imV = imV + ((2.0 * x_scatV) * tau_e) * fact;
                sim = sim + ((2.0 * x_scat) * tau_e) * fact;
                fact = fact * (1.0 - tau_e);
            } else {
                const double ex = exp(-tau_e);
                const double frac = 1.0 - ex;
                img = img + ((2.0 * x_scat + A.emis[c]) * frac) * fact;
                imQ = imQ + ((2.0 * x_scatQ) * frac) * fact;
                imU = imU + ((2.0 * x_scatU) * frac) * fact;
                imV = imV + ((2.0 * x_scatV) * frac) * fact;
                sim = sim + ((2.0 * x_scat) * frac) * fact;
                fact = fact * ex;
            }
            tau = tau + tau_e;
        }
        if (tau > tau_max) {
            stopped = true;
            break;
        }
    }
    if (!stopped && A.hitstar[ray] == 1 && A.tracestar) img = img + (A.Istar * fact) * 1.0;
    A.image[ray] = img;
    A.scatim[ray] = sim;
    A.imageQ[ray] = imQ;
    if (A.imageU) {
        A.imageU[ray] = imU;
        A.imageV[ray] = imV;
    }
}

// MakeImage (TraceMono.f:2759-3020): the ray-traced image on its rings
// (nimg radii, nphi azimuths over half a turn) spread over a square of
// npix^2 pixels of half-width Rmax (AU).  Polar cell (ring i, azimuth j,
// 1-based as the reference) has nint samples at random radius and azimuth,
// each deposited at phi and at -phi (U, V change sign there).
// Full scattering interpolates I-|P|, and the two linear-polarisation
// amplitudes, in log between the four neighbouring rays where the ring
// spacing exceeds a pixel, falling back to linear interpolation where that
// yields a non-finite value; otherwise log-linear I.  Draws: sample k of
// cell t = (i-1)*nphi + j-1 takes Philox block k of stream (seed, t), its
// two doubles being the radius and azimuth draws.
//
// One block (one wave) per polar cell, lane l taking samples l, l+64, ...
// Most rings lie inside a few central pixels, and an outer cell's samples
// spread over tens: a lane sums its consecutive deposits on one pixel, and
// adds each run to an LDS tile over the pixels the cell can reach on that
// side (the +phi and the -phi half), which goes to the image once at the
// end; cells whose box exceeds the tile add their runs to the image.
struct ImgArgs {
    int nimg, nphi, npix, nint, scat_how;
    const double *R;                // [nimg] ring radii (AU)
    const double *I, *Q, *U, *V;    // [nimg*nphi] ray images, ray = ring*nphi + azimuth
    // per ray, from k_image_rays: log10(I - |P|) (log10 I without full
    // scattering), log10 of the amplitudes al, ar, and U, V / (2 al ar)
    const double *liv, *lal, *lar, *cv, *sv;
    double Rmax, pa, au;            // half field (AU), position angle (deg), AU in cm
    unsigned long long seed;
    double *im, *imQ, *imU, *imV;   // [npix*npix], pixel (ix, iy) at (ix-1)*npix + iy-1
};

// number_invalid (Settling.f:652-674): not a finite value below DBL_MAX/2
__device__ __forceinline__ bool img_invalid(double a) {
    const double b = (a * 2.0) / 2.0, c = a - 1e100;
    return !(b == a && c < a);
}

struct ImgAcc {
    int px;           // pixel of the sums, -1 for none
    double f, q, u, v;
};

__device__ __forceinline__ void img_add(const ImgArgs &A, const ImgAcc &c) {
    atomicAdd(&A.im[c.px], c.f);
    if (A.scat_how == 2) {
        atomicAdd(&A.imQ[c.px], c.q);
        atomicAdd(&A.imU[c.px], c.u);
        atomicAdd(&A.imV[c.px], c.v);
    }
}

// The per-ray factors of MakeImage's interpolation (TraceMono.f:2884-2927),
// once per ray instead of once per sample: with full scattering
// p = sqrt(Q^2 + U^2 + V^2), log10(I - p), log10 sqrt((p + Q)/2),
// log10 sqrt((p - Q)/2) and U, V over 2 al ar; otherwise log10 I.
__global__ __launch_bounds__(256) void k_image_rays(int nray, int scat_how, const double *I, const double *Q,
                                                    const double *U, const double *V, double *liv, double *lal,
                                                    double *lar, double *cv, double *sv) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nray) return;
    if (scat_how != 2) {
        liv[q] = log10(I[q]);
        return;
    }
    const double Iq = I[q], Qq = Q[q], Uq = U[q], Vq = V[q];
    const double pq = sqrt(Qq * Qq + Uq * Uq + Vq * Vq);
    const double al = sqrt((pq + Qq) / 2.0), ar = sqrt((pq - Qq) / 2.0);
    liv[q] = log10(Iq - pq);
    lal[q] = log10(al);
    lar[q] = log10(ar);
    cv[q] = Uq / ((2.0 * al) * ar);
    sv[q] = Vq / ((2.0 * al) * ar);
}

#define IMG_TILE 256   // pixels of a side's LDS tile

// a side's pixel box: the pixels that a cell's samples at angles a0..a1
// (radians, after the position angle) and radii Ri..Ri1 can reach, one
// pixel of margin each way, clipped to the image
struct ImgBox {
    int x0, y0, nx, ny;
    bool ok;   // fits the tile
};

__device__ __forceinline__ ImgBox img_box(const ImgArgs &A, double Ri, double Ri1, double a0, double a1,
                                          double fpix) {
    double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
    auto put = [&](double R, double a) {
        const double x = -R * cos(a), y = R * sin(a);
        xmin = fmin(xmin, x);
        xmax = fmax(xmax, x);
        ymin = fmin(ymin, y);
        ymax = fmax(ymax, y);
    };
    put(Ri, a0);
    put(Ri, a1);
    put(Ri1, a0);
    put(Ri1, a1);
    const double q = 0.5 * MCX_PI;
    for (double m = floor(a0 / q) + 1.0; m * q < a1; m += 1.0) {   // the axes the sector crosses
        put(Ri, m * q);
        put(Ri1, m * q);
    }
    auto pix = [&](double v) {   // pixel index of a coordinate, kept within [-1, npix + 2]
        const double t = (fpix * (v + A.Rmax)) / (2.0 * A.Rmax) + 1.0;
        return (int)floor(fmin(fmax(t, -1.0), fpix + 2.0));
    };
    ImgBox B;
    const int xl = max(1, pix(xmin) - 1), xh = min(A.npix, pix(xmax) + 1);
    const int yl = max(1, pix(ymin) - 1), yh = min(A.npix, pix(ymax) + 1);
    B.x0 = xl;
    B.y0 = yl;
    B.nx = max(0, xh - xl + 1);
    B.ny = max(0, yh - yl + 1);
    B.ok = B.nx * B.ny <= IMG_TILE;
    return B;
}

// a run's sums into the side's LDS tile when the pixel lies in its box,
// else straight into the image
__device__ __forceinline__ void img_put(const ImgArgs &A, double (*tl)[IMG_TILE], const ImgBox &B,
                                        const ImgAcc &c) {
    if (c.px < 0) return;
    const int ix = c.px / A.npix + 1, iy = c.px % A.npix + 1;
    const int dx = ix - B.x0, dy = iy - B.y0;
    if (B.ok && dx >= 0 && dx < B.nx && dy >= 0 && dy < B.ny) {
        const int e = dx * B.ny + dy;
        atomicAdd(&tl[0][e], c.f);
        if (A.scat_how == 2) {
            atomicAdd(&tl[1][e], c.q);
            atomicAdd(&tl[2][e], c.u);
            atomicAdd(&tl[3][e], c.v);
        }
    } else {
        img_add(A, c);
    }
}

// one block (one wave) per polar cell; lane l takes samples l, l+64, ...
__global__ __launch_bounds__(64, 3) void k_image(ImgArgs A) {
    __shared__ double tile[2][4][IMG_TILE];
    const int np = A.nphi;
    const long long cell = blockIdx.x;
    const int lane = threadIdx.x;
    const int i = (int)(cell / np) + 1, j = (int)(cell % np) + 1;
    const double Ri = A.R[i - 1], Ri1 = A.R[i];
    const double w2 = ((((2.0 * MCX_PI) * (A.au * A.au)) * (Ri1 - Ri)) / (double)(float)np);
    const double den = (double)(float)(2 * A.nint);
    const double fpix = (double)(float)A.npix;
    const double par = (A.pa * MCX_PI) / 180.0;
    const uint32_t key0 = (uint32_t)A.seed, key1 = (uint32_t)(A.seed >> 32);
    const double th0 = (MCX_PI * (double)(j - 1)) / (double)np, th1 = (MCX_PI * (double)j) / (double)np;
    const ImgBox box[2] = {img_box(A, Ri, Ri1, th0 + par, th1 + par, fpix),
                           img_box(A, Ri, Ri1, -th1 + par, -th0 + par, fpix)};
    for (int e = lane; e < 2 * 4 * IMG_TILE; e += 64) (&tile[0][0][0])[e] = 0.0;
    __syncthreads();
    ImgAcc acc[2] = {{-1, 0.0, 0.0, 0.0, 0.0}, {-1, 0.0, 0.0, 0.0, 0.0}};
    for (int k = lane; k < A.nint; k += 64) {
        double ranR, ranPhi;
        philox_block(key0, key1, (uint64_t)k, (uint64_t)cell, ranR, ranPhi);
        const double R = Ri + (Ri1 - Ri) * ranR;
        const double phi = (MCX_PI * ((double)(float)(j - 1) + ranPhi)) / (double)(float)np;
        double wp1, wp2;
        int jp1, jp2;
        if (ranPhi < 0.5) {
            wp1 = 0.5 - ranPhi;
            jp1 = j - 1;
            if (jp1 == 0) jp1 = 1;
            wp2 = ranPhi + 0.5;
            jp2 = j;
        } else {
            wp1 = ranPhi - 0.5;
            jp1 = j + 1;
            if (jp1 == np + 1) jp1 = np;
            wp2 = 1.5 - ranPhi;
            jp2 = j;
        }
        const double wr1 = 1.0 - ranR, wr2 = ranR;
        const int ir1 = i, ir2 = i + 1;
        double flux, fQ = 0.0, fU = 0.0, fV = 0.0;
        const size_t r11 = (size_t)(ir1 - 1) * np + (jp1 - 1), r12 = (size_t)(ir1 - 1) * np + (jp2 - 1);
        const size_t r21 = (size_t)(ir2 - 1) * np + (jp1 - 1), r22 = (size_t)(ir2 - 1) * np + (jp2 - 1);
        auto mix = [&](const double *X) {
            return wr1 * (wp1 * X[r11] + wp2 * X[r12]) + wr2 * (wp1 * X[r21] + wp2 * X[r22]);
        };
        if (A.scat_how == 2) {
            if (fabs(A.R[ir2 - 1] - A.R[ir1 - 1]) > (A.Rmax / fpix)) {
                const double ea = pow(10.0, mix(A.lal)), eb = pow(10.0, mix(A.lar));
                flux = ea * ea;
                flux = flux + eb * eb;
                flux = flux + pow(10.0, mix(A.liv));
                flux = (flux * w2) * R;
                fQ = ea * ea;
                fQ = fQ - eb * eb;
                fQ = (fQ * w2) * R;
                fU = ea;
                fU = fU * eb;
                fU = (fU * 2.0) * mix(A.cv);
                fU = (fU * w2) * R;
                fV = ea;
                fV = fV * eb;
                fV = (fV * 2.0) * mix(A.sv);
                fV = (fV * w2) * R;
                if (img_invalid(flux) || img_invalid(fQ) || img_invalid(fU) || img_invalid(fV)) {
                    flux = mix(A.I) * w2 * R;
                    fQ = mix(A.Q) * w2 * R;
                    fU = mix(A.U) * w2 * R;
                    fV = mix(A.V) * w2 * R;
                }
            } else {
                flux = mix(A.I) * w2 * R;
                fQ = mix(A.Q) * w2 * R;
                fU = mix(A.U) * w2 * R;
                fV = mix(A.V) * w2 * R;
            }
        } else {
            flux = mix(A.liv);
            flux = (pow(10.0, flux) * w2) * R;
        }
        for (int side = 0; side < 2; side++) {
            const double ph = side ? -phi : phi;
            double sn, cs;
            sincos(ph + par, &sn, &cs);
            double x = -R * cs;
            x = (fpix * (x + A.Rmax)) / (2.0 * A.Rmax) + 1.0;
            double y = R * sn;
            y = (fpix * (y + A.Rmax)) / (2.0 * A.Rmax) + 1.0;
            const int ix = (int)x, iy = (int)y;
            const int px = (ix <= A.npix && iy <= A.npix && ix > 0 && iy > 0) ? (ix - 1) * A.npix + (iy - 1) : -1;
            ImgAcc &c = acc[side];
            if (px != c.px) {
                img_put(A, tile[side], box[side], c);
                c.px = px;
                c.f = c.q = c.u = c.v = 0.0;
            }
            if (px >= 0) {
                c.f += flux / den;
                c.q += fQ / den;
                c.u += side ? -(fU / den) : fU / den;
                c.v += side ? -(fV / den) : fV / den;
            }
        }
    }
    img_put(A, tile[0], box[0], acc[0]);
    img_put(A, tile[1], box[1], acc[1]);
    __syncthreads();
    for (int side = 0; side < 2; side++) {
        const ImgBox &B = box[side];
        if (!B.ok) continue;
        for (int e = lane; e < B.nx * B.ny; e += 64) {
            const size_t px = (size_t)(B.x0 + e / B.ny - 1) * A.npix + (B.y0 + e % B.ny - 1);
            const double f = tile[side][0][e];
            if (f != 0.0) atomicAdd(&A.im[px], f);
            if (A.scat_how == 2) {
                const double q = tile[side][1][e], u = tile[side][2][e], v = tile[side][3][e];
                if (q != 0.0) atomicAdd(&A.imQ[px], q);
                if (u != 0.0) atomicAdd(&A.imU[px], u);
                if (v != 0.0) atomicAdd(&A.imV[px], v);
            }
        }
    }
}

// ===========================================================================
// host side
static void host_rotate(double &x, double &y, double &z, double u, double v, double w, double theta) {
    const double cost = cos(theta), sint = sin(theta);
    const double u2 = u * u, v2 = v * v, w2 = w * w;
    const double X = x, Y = y, Z = z;
    const double inp = X * u + Y * v + Z * w;
    x = u * inp + (X * (v2 + w2) - u * (v * Y + w * Z)) * cost + (v * Z - w * Y) * sint;
    y = v * inp + (Y * (u2 + w2) - v * (u * X + w * Z)) * cost + (w * X - u * Z) * sint;
    z = w * inp + (Z * (u2 + v2) - w * (u * X + v * Y)) * cost + (u * Y - v * X) * sint;
}

// monochromatic weights of lam0 on the grid (TraceMono.f:622-647), 0-based
static void mono_weights(const std::vector<double> &lam, double lam0, int &l1, int &l2, double &wl1, double &wl2) {
    const int nlam = (int)lam.size();
    if (lam0 <= lam[0]) {
        l1 = 0;
        l2 = 1;
        wl1 = 1.0;
        wl2 = 0.0;
    }
    for (int i = 1; i <= nlam - 1; i++)
        if (lam0 >= lam[i - 1] && lam0 <= lam[i]) {
            l1 = i - 1;
            l2 = i;
            wl1 = (lam[i] - lam0) / (lam[i] - lam[i - 1]);
            wl2 = (lam0 - lam[i - 1]) / (lam[i] - lam[i - 1]);
        }
    if (lam0 >= lam[nlam - 1]) {
        l1 = nlam - 2;
        l2 = nlam - 1;
        wl1 = 0.0;
        wl2 = 1.0;
    }
}

// tau01R of the midplane row at lam0: a radial ray from R(1) to tau = 0.1
// (TraceMono.f:2300-2345)
static double tau01_midplane(const mcx_ctx *ctx, double lam0) {
    int l1 = 0, l2 = 1;
    double wl1 = 1.0, wl2 = 0.0;
    mono_weights(ctx->h_lam, lam0, l1, l2, wl1, wl2);
    const int nt = ctx->md.nt, ng = ctx->ng, nlam = ctx->nlam;
    const std::vector<double> &R = ctx->h_R, &Th = ctx->h_Theta;
    std::vector<double> R2(R.size()), T2(Th.size());
    for (size_t k = 0; k < R.size(); k++) R2[k] = R[k] * R[k];
    for (size_t k = 0; k < Th.size(); k++) T2[k] = Th[k] * Th[k];
    const int jj0 = nt - 1;
    const double r = R[1];
    const double ct = (Th[jj0] + Th[jj0 + 1]) / 2.0;
    double x = sqrt(1.0 - ct * ct) * r, y = 0.0, z = ct * r;
    const double vx = x / r, vy = 0.0, vz = z / r;
    int i = 1, jj = jj0, edgeNr = 1;
    bool onEdge = true;
    const double tau0 = 0.1;
    double tau = 0.0;
    for (int guard = 0; guard < 100000; guard++) {
        const int c = i * nt + jj;
        double Kext = 0.0;
        for (int s = 0; s < ng; s++) {
            const double *Ke = ctx->h_Kext.data() + (size_t)s * nlam;
            Kext = Kext + ((Ke[l1] * wl1 + Ke[l2] * wl2) * ctx->h_w[(size_t)c * ng + s]) * 1.0;
        }
        double v;
        int in, jn;
        if (!exit_face(R2.data(), T2.data(), nt, x, y, z, vx, vy, vz, i, jj, onEdge, edgeNr, v, in, jn)) break;
        const double tau1 = ((Kext * v) * ctx->h_dens[c]) * MCX_AU;
        if (tau + tau1 > tau0) {
            v = (v * (tau0 - tau)) / tau1;
            x += vx * v;
            y += vy * v;
            z += vz * v;
            break;
        }
        x += vx * v;
        y += vy * v;
        z += vz * v;
        onEdge = true;
        tau += tau1;
        if (in >= ctx->md.nr || in < 0) {
            z = 0.0;
            break;
        }
        i = in;
        jj = jn;
    }
    double R01 = sqrt(x * x + z * z);
    if (R01 > R[ctx->md.nr - 1]) R01 = R[1];
    return R01;
}

// image radii of TracePath (TraceMono.f:2098-2229), nRfix = 0
static int image_radii(const mcx_ctx *ctx, const mcx_obs *o, const double *theta_av, double R01,
                       std::vector<double> &Rimg) {
    const int nt = ctx->md.nt, NR = ctx->md.nr;
    const std::vector<double> &Rg = ctx->h_R;
    int nimg = (o->nj0 > 0) ? o->nj0 * (NR - 1) + nt * 2 * 2 + 1 + NSTAR_RAYS
                            : (NR - 2) / (-o->nj0) + 1 + nt * 2 * 2 + 1 + NSTAR_RAYS;
    if (ctx->tb.scat_how == 2) nimg += 180;
    std::vector<double> R(nimg, 0.0);
    int nj = 0;
    const double ang = o->angle;
    for (int j = 0; j < nt; j++) {
        R[nj++] = R01 * fabs(cos(MCX_PI / 2.0 - theta_av[j] + ang));
        R[nj++] = fabs(R01 * cos(theta_av[j] - MCX_PI / 2.0 + ang));
    }
    for (int j = 0; j < nt; j++) {
        R[nj++] = R01 * fabs(cos(MCX_PI / 2.0 - theta_av[j]));
        R[nj++] = fabs(R01 * cos(theta_av[j] - MCX_PI / 2.0));
    }
    if (ctx->tb.scat_how == 2)
        for (int j = 1; j <= 180; j++) {
            const double theta = (double)j * (MCX_PI / 2.0) / 181.0;
            R[nj++] = Rg[1] * fabs(cos(MCX_PI / 2.0 - theta));
        }
    std::sort(R.begin(), R.begin() + nj);
    if (o->nj0 > 0) {
        for (int i = 1; i <= NR - 1; i++)
            for (int j = 1; j <= o->nj0; j++)
                R[nj++] = Rg[i] + (Rg[i + 1] - Rg[i]) * ((double)j - 0.5) / (double)o->nj0;
    } else {
        for (int i = 1; i <= NR - 1; i += -o->nj0) R[nj++] = pow(10.0, log10(Rg[i] * Rg[i + 1]) / 2.0);
    }
    for (int i = 1; i <= NSTAR_RAYS - 1; i++) R[nj++] = ((0.999 * Rg[0]) * (double)(i - 1)) / (double)(NSTAR_RAYS - 2);
    R[nj++] = Rg[0] * (double)1.001f;
    R[nimg - 1] = Rg[NR] * (double)0.9999f;
    std::sort(R.begin(), R.begin() + nimg);
    // separate coincident radii (TraceMono.f:2217-2229)
    int k = 0;
restart:
    k++;
    for (int i = nimg; i >= 2; i--) {
        if (R[i - 1] == R[i - 2]) {
            if (i != nimg) R[i - 1] = (R[i] + R[i - 1]) / 2.0;
            else R[i - 2] = (R[i - 2] + R[i - 1]) / 2.0;
            if (k < nimg * 10) goto restart;
        }
    }
    Rimg.assign(R.begin(), R.begin() + nimg);
    return nimg;
}

// the rays of the last wavelength, kept for an image
struct RayImages {
    std::vector<double> R;          // ring radii (AU)
    int nimg = 0, nphi = 0;
    std::vector<double> I, Q, U, V; // [nimg*nphi]
};

static int sed_core(mcx_ctx *ctx, const mcx_obs *o, const double *T, const double *theta_av, int32_t nlam_obs,
                    const double *lam_obs, double *flux, double *scatflux, double *fluxQ, double path_lam,
                    RayImages *ro);

extern "C" int mcx_sed(mcx_ctx *ctx, const mcx_obs *o, const double *T, const double *theta_av, int32_t nlam_obs,
                       const double *lam_obs, double *flux, double *scatflux, double *fluxQ) {
    // Telescope.f:140: the SED's rays are laid out at 0.55 micron
    return sed_core(ctx, o, T, theta_av, nlam_obs, lam_obs, flux, scatflux, fluxQ, 0.55, nullptr);
}

static int sed_core(mcx_ctx *ctx, const mcx_obs *o, const double *T, const double *theta_av, int32_t nlam_obs,
                    const double *lam_obs, double *flux, double *scatflux, double *fluxQ, double path_lam,
                    RayImages *ro) {
    if (!ctx || !o || !theta_av || nlam_obs < 0 || (nlam_obs > 0 && (!lam_obs || !flux))) return MCX_EINVAL;
    if (!ctx->have_model) return fail(ctx, MCX_ESTATE, "mcx_set_model first");
    if (!ctx->d_Kext || ctx->h_thet.empty())
        return fail(ctx, MCX_EINVAL, "the SED needs Kext (mcx_set_tables) and thet (mcx_set_model)");
    if (o->nphi < 1 || o->nj0 == 0 || o->nphot < 0 || o->nphotstar < 0)
        return fail(ctx, MCX_EINVAL, "bad observation");
    if (!T && !ctx->have_temps) return fail(ctx, MCX_ESTATE, "no temperatures: pass T or run mcx_temperature");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const DevModel &md = ctx->md;
    const int nr = md.nr, nt = md.nt, NT = nt + 1, ncell = md.ncell, ng = ctx->ng, nlam = ctx->nlam;
    const int scat_how = ctx->tb.scat_how;
    const bool scattering = ctx->tb.scattering != 0;
    const int nmc = o->tracescat ? o->nphot : 0, nmcs = o->tracescat ? o->nphotstar : 0;
    const bool mc = scattering && (nmc + nmcs) != 0;
    std::vector<void *> tmp;
    struct Guard {
        mcx_ctx *c;
        std::vector<void *> &l;
        ~Guard() { pfree_list(c, l); }
    } guard{ctx, tmp};
    int r;
    // ---- geometry of the observation -----------------------------------
    std::vector<int> nrg(nt), rgoff(nt);
    int NRG = 0, MAXRG = 1;
    for (int jj = 0; jj < nt; jj++) {
        int n = 1;
        if (scattering && (nmc + nmcs) != 0 && scat_how == 2) {
            const int q = (int)((ctx->h_thet[jj + 1] - ctx->h_thet[jj]) * 270.0 / MCX_PI);
            n = (q > 1) ? q : 1;
        }
        nrg[jj] = n;
        rgoff[jj] = NRG;
        NRG += n;
        if (n > MAXRG) MAXRG = n;
    }
    if (MAXRG > 255) return fail(ctx, MCX_EINVAL, "more than 255 theta sub-cells in a row");
    std::vector<double> thetrg((size_t)nt * (MAXRG + 1), 0.0), fracirg((size_t)nt * (MAXRG + 1), 0.0);
    for (int jj = 0; jj < nt; jj++) {
        const double t0 = ctx->h_thet[jj], t1 = ctx->h_thet[jj + 1];
        for (int k = 1; k <= nrg[jj] + 1; k++) {
            const double c = cos(t0 + ((double)(float)(k - 1) * (t1 - t0)) / (double)(float)nrg[jj]);
            thetrg[(size_t)jj * (MAXRG + 1) + k - 1] = c * c;
        }
        for (int k = 1; k <= nrg[jj]; k++)
            fracirg[(size_t)jj * (MAXRG + 1) + k - 1] =
                (cos(t0 + ((double)(float)(k - 1) * (t1 - t0)) / (double)(float)nrg[jj]) -
                 cos(t0 + ((double)(float)k * (t1 - t0)) / (double)(float)nrg[jj])) /
                (ctx->h_Theta[jj] - ctx->h_Theta[jj + 1]);
    }
    std::vector<double> dirs(6 * NPHIS), pisin(180);
    for (int i = 1; i <= NPHIS; i++) {
        const double phi = MCX_PI * ((double)i - 0.5) / ((double)(float)NPHIS / 2.0);
        double x = 0.0, y = 0.0, z = 1.0;
        host_rotate(x, y, z, 0.0, 1.0, 0.0, o->angle);
        host_rotate(x, y, z, 0.0, 0.0, 1.0, phi);
        double rr = sqrt(x * x + y * y + z * z);
        dirs[0 * NPHIS + i - 1] = x / rr;
        dirs[1 * NPHIS + i - 1] = y / rr;
        dirs[2 * NPHIS + i - 1] = z / rr;
        x = 0.0;
        y = 1.0;
        z = 0.0;
        host_rotate(x, y, z, 0.0, 0.0, 1.0, phi);
        rr = sqrt(x * x + y * y + z * z);
        dirs[3 * NPHIS + i - 1] = x / rr;
        dirs[4 * NPHIS + i - 1] = y / rr;
        dirs[5 * NPHIS + i - 1] = z / rr;
    }
    for (int a = 1; a <= 180; a++) pisin[a - 1] = MCX_PI * sin((MCX_PI * ((double)a - 0.5)) / 180.0);
    const double R01 = tau01_midplane(ctx, path_lam);
    std::vector<double> Rimg;
    const int nimg = image_radii(ctx, o, theta_av, R01, Rimg);
    const int nphi = o->nphi, nray = nimg * nphi;
    int *d_nrg, *d_rgoff;
    double *d_thetrg, *d_fracirg, *d_dirs, *d_pisin, *d_Rimg, *d_thet;
    if ((r = pupload(ctx, tmp, &d_nrg, nrg.data(), nrg.size()))) return r;
    if ((r = pupload(ctx, tmp, &d_rgoff, rgoff.data(), rgoff.size()))) return r;
    if ((r = pupload(ctx, tmp, &d_thetrg, thetrg.data(), thetrg.size()))) return r;
    if ((r = pupload(ctx, tmp, &d_fracirg, fracirg.data(), fracirg.size()))) return r;
    if ((r = pupload(ctx, tmp, &d_dirs, dirs.data(), dirs.size()))) return r;
    if ((r = pupload(ctx, tmp, &d_pisin, pisin.data(), pisin.size()))) return r;
    if ((r = pupload(ctx, tmp, &d_Rimg, Rimg.data(), Rimg.size()))) return r;
    if ((r = pupload(ctx, tmp, &d_thet, ctx->h_thet.data(), ctx->h_thet.size()))) return r;
    SedGeom G;
    G.nr = nr;
    G.nt = nt;
    G.NT = NT;
    G.NRG = NRG;
    G.MAXRG = MAXRG;
    G.nrg = d_nrg;
    G.rgoff = d_rgoff;
    G.thetrg = d_thetrg;
    G.thet = d_thet;
    G.R2 = md.R2;
    G.T2 = md.T2;
    G.R = ctx->rw.R ? ctx->rw.R : nullptr;
    G.Theta = md.Theta;
    if (!G.R) return fail(ctx, MCX_EINVAL, "the SED needs thet in mcx_set_model");
    // ---- image rays ------------------------------------------------------
    int *d_cnt, *d_hit;
    long long *d_off;
    if ((r = palloc(ctx, tmp, &d_cnt, (size_t)nray))) return r;
    if ((r = palloc(ctx, tmp, &d_hit, (size_t)nray))) return r;
    if ((r = palloc(ctx, tmp, &d_off, (size_t)nray))) return r;
    const unsigned rb = (unsigned)((nray + 127) / 128);
    hipLaunchKernelGGL(k_paths, dim3(rb), dim3(128), 0, st, G, d_Rimg, nimg, nphi, o->angle, 0, d_cnt,
                       (const long long *)nullptr, (int *)nullptr, (unsigned char *)nullptr,
                       (unsigned char *)nullptr, (double *)nullptr, (unsigned char *)nullptr,
                       (unsigned char *)nullptr, d_hit);
    HIPCHK(ctx, hipGetLastError());
    std::vector<int> cnt(nray);
    HIPCHK(ctx, hipMemcpyAsync(cnt.data(), d_cnt, sizeof(int) * nray, hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    std::vector<long long> off(nray);
    long long nseg = 0;
    for (int k = 0; k < nray; k++) {
        off[k] = nseg;
        nseg += cnt[k];
    }
    HIPCHK(ctx, hipMemcpyAsync(d_off, off.data(), sizeof(long long) * nray, hipMemcpyHostToDevice, st));
    int *seg_c;
    unsigned char *seg_irg, *seg_k, *seg_j1, *seg_j2;
    double *seg_v;
    if ((r = palloc(ctx, tmp, &seg_c, (size_t)nseg))) return r;
    if ((r = palloc(ctx, tmp, &seg_irg, (size_t)nseg))) return r;
    if ((r = palloc(ctx, tmp, &seg_k, (size_t)nseg))) return r;
    if ((r = palloc(ctx, tmp, &seg_j1, (size_t)nseg))) return r;
    if ((r = palloc(ctx, tmp, &seg_j2, (size_t)nseg))) return r;
    if ((r = palloc(ctx, tmp, &seg_v, (size_t)nseg))) return r;
    hipLaunchKernelGGL(k_paths, dim3(rb), dim3(128), 0, st, G, d_Rimg, nimg, nphi, o->angle, 1, d_cnt,
                       (const long long *)d_off, seg_c, seg_irg, seg_k, seg_v, seg_j1, seg_j2, d_hit);
    HIPCHK(ctx, hipGetLastError());
    std::vector<int> hit(nray);
    HIPCHK(ctx, hipMemcpyAsync(hit.data(), d_hit, sizeof(int) * nray, hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    for (int k = 0; k < nray; k++)
        if (hit[k] < 0) return fail(ctx, MCX_EGEOM, "an image ray found nothing to hit");
    // ---- per-wavelength buffers -------------------------------------------
    double *d_T = nullptr;
    if (T) {
        if ((r = pupload(ctx, tmp, &d_T, T, (size_t)ncell))) return r;
    } else {
        d_T = ctx->d_T;
    }
    double *Kabs, *Kext, *Ksca, *Alb, *emis, *Eem, *EmisDis, *vism, *cum, *tot, *Fmix, *MIF11, *MIF12;
    if ((r = palloc(ctx, tmp, &Kabs, (size_t)ncell))) return r;
    if ((r = palloc(ctx, tmp, &Kext, (size_t)ncell))) return r;
    if ((r = palloc(ctx, tmp, &Ksca, (size_t)ncell))) return r;
    if ((r = palloc(ctx, tmp, &Alb, (size_t)ncell))) return r;
    if ((r = palloc(ctx, tmp, &emis, (size_t)ncell))) return r;
    if ((r = palloc(ctx, tmp, &Eem, (size_t)ncell))) return r;
    if ((r = palloc(ctx, tmp, &EmisDis, (size_t)ncell))) return r;
    if ((r = palloc(ctx, tmp, &vism, (size_t)ncell))) return r;
    if ((r = palloc(ctx, tmp, &cum, (size_t)ncell))) return r;
    if ((r = palloc(ctx, tmp, &tot, 2))) return r;
    const int nclass = ctx->nclass;
    if ((r = palloc(ctx, tmp, &Fmix, (size_t)nclass * 6 * 180))) return r;
    if ((r = palloc(ctx, tmp, &MIF11, (size_t)nclass))) return r;
    if ((r = palloc(ctx, tmp, &MIF12, (size_t)nclass))) return r;
    const size_t nsf = (size_t)nr * NRG * 91 * 2;
    double *sf, *sQ, *sU, *sV;
    if ((r = palloc(ctx, tmp, &sf, nsf))) return r;
    if ((r = palloc(ctx, tmp, &sQ, nsf))) return r;
    if ((r = palloc(ctx, tmp, &sU, nsf))) return r;
    if ((r = palloc(ctx, tmp, &sV, nsf))) return r;
    double *image, *scatim, *imageQ, *imageU = nullptr, *imageV = nullptr;
    if ((r = palloc(ctx, tmp, &image, (size_t)nray))) return r;
    if ((r = palloc(ctx, tmp, &scatim, (size_t)nray))) return r;
    if ((r = palloc(ctx, tmp, &imageQ, (size_t)nray))) return r;
    if (ro) {
        if ((r = palloc(ctx, tmp, &imageU, (size_t)nray))) return r;
        if ((r = palloc(ctx, tmp, &imageV, (size_t)nray))) return r;
    }
    std::vector<double> h_img(nray), h_sim(nray), h_imQ(nray);
    const unsigned cb = (unsigned)((ncell + 255) / 256);
    for (int il = 0; il < nlam_obs; il++) {
        SedMono M;
        mono_weights(ctx->h_lam, lam_obs[il], M.l1, M.l2, M.wl1, M.wl2);
        M.Kabs = Kabs;
        M.Kext = Kext;
        M.Ksca = Ksca;
        M.Alb = Alb;
        M.cls = ctx->rw.cls;
        M.F = Fmix;
        M.MIF11 = MIF11;
        M.MIF12 = MIF12;
        hipLaunchKernelGGL(k_mono, dim3(cb), dim3(256), 0, st, nr, nt, nlam, ng, M, ctx->d_Kabs, ctx->d_Kext,
                           ctx->d_Ksca, md.w, Kabs, Kext, Ksca, Alb);
        hipLaunchKernelGGL(k_emis, dim3(cb), dim3(256), 0, st, G, nlam, ng, M, ctx->d_Kabs, md.w, ctx->d_BB, d_T,
                           md.dens, ctx->d_V, o->nexits, o->traceemis, emis, Eem, EmisDis, vism);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipMemsetAsync(sf, 0, sizeof(double) * nsf, st));
        HIPCHK(ctx, hipMemsetAsync(sQ, 0, sizeof(double) * nsf, st));
        HIPCHK(ctx, hipMemsetAsync(sU, 0, sizeof(double) * nsf, st));
        HIPCHK(ctx, hipMemsetAsync(sV, 0, sizeof(double) * nsf, st));
        const double Estar = ctx->h_Fstar[M.l1] * M.wl1 + ctx->h_Fstar[M.l2] * M.wl2;
        if (mc) {
            const int nel = o->usepol ? 6 : 2;
            const unsigned mb = (unsigned)((nclass * 6 * 180 + 255) / 256);
            hipLaunchKernelGGL(k_mix, dim3(mb), dim3(256), 0, st, nclass, nlam, ng, M, ctx->d_wcls, ctx->tb.F,
                               ctx->d_Ksca, ctx->tb.IF11, ctx->tb.IF12, nel, Fmix, MIF11, MIF12);
            hipLaunchKernelGGL(k_scan, dim3(1), dim3(256), 0, st, ncell, nt, Eem, EmisDis, cum, tot);
            HIPCHK(ctx, hipGetLastError());
            McArgs A;
            A.G = G;
            A.M = M;
            A.O.xsf = d_dirs;
            A.O.ysf = d_dirs + NPHIS;
            A.O.zsf = d_dirs + 2 * NPHIS;
            A.O.xin = d_dirs + 3 * NPHIS;
            A.O.yin = d_dirs + 4 * NPHIS;
            A.O.zin = d_dirs + 5 * NPHIS;
            A.O.pisin = d_pisin;
            A.O.usepol = o->usepol;
            A.O.scat_how = scat_how;
            A.O.pfstop = o->pfstop;
            A.dens = md.dens;
            A.V = ctx->d_V;
            A.cum = cum;
            A.tot = tot;
            A.vismass = vism;
            A.sf = sf;
            A.sQ = sQ;
            A.sU = sU;
            A.sV = sV;
            A.cos2phi = ctx->tb.cos2phi;
            A.sin2phi = ctx->tb.sin2phi;
            A.R0 = ctx->h_R[0];
            A.Estar = Estar;
            A.nphot_star = nmcs;
            A.nphot_therm = nmc;
            A.ncell = ncell;
            A.seed = o->seed + (unsigned long long)il;
            hipLaunchKernelGGL(k_mc, dim3((unsigned)(nmcs + nmc)), dim3(64), 0, st, A);
            HIPCHK(ctx, hipGetLastError());
        }
        RayArgs B;
        B.nt = nt;
        B.NRG = NRG;
        B.MAXRG = MAXRG;
        B.scat_how = scat_how;
        B.mc = mc ? 1 : 0;
        B.tracestar = o->tracestar;
        B.rgoff = d_rgoff;
        B.off = d_off;
        B.cnt = d_cnt;
        B.seg_c = seg_c;
        B.hitstar = d_hit;
        B.seg_irg = seg_irg;
        B.seg_k = seg_k;
        B.seg_j1 = seg_j1;
        B.seg_j2 = seg_j2;
        B.seg_v = seg_v;
        B.fracirg = d_fracirg;
        B.dens = md.dens;
        B.V = ctx->d_V;
        B.Kext = Kext;
        B.Alb = Alb;
        B.emis = emis;
        B.sf = sf;
        B.sQ = sQ;
        B.sU = sU;
        B.sV = sV;
        B.imageU = imageU;
        B.imageV = imageV;
        B.Istar = (((ctx->h_Fstar[M.l1] * M.wl1 + ctx->h_Fstar[M.l2] * M.wl2) * 1.0) * 1.0) /
                  (MCX_PI * ((ctx->h_R[0] * MCX_AU) * (ctx->h_R[0] * MCX_AU)));
        B.image = image;
        B.scatim = scatim;
        B.imageQ = imageQ;
        B.nray = nray;
        hipLaunchKernelGGL(k_rays, dim3(rb), dim3(128), 0, st, B);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipMemcpyAsync(h_img.data(), image, sizeof(double) * nray, hipMemcpyDeviceToHost, st));
        HIPCHK(ctx, hipMemcpyAsync(h_sim.data(), scatim, sizeof(double) * nray, hipMemcpyDeviceToHost, st));
        HIPCHK(ctx, hipMemcpyAsync(h_imQ.data(), imageQ, sizeof(double) * nray, hipMemcpyDeviceToHost, st));
        HIPCHK(ctx, hipStreamSynchronize(st));
        // flux quadrature over the rings (TraceMono.f:499-520), mask = 1
        double fl = 0.0, sfl = 0.0, fq = 0.0;
        for (int i = 1; i <= nimg - 1; i++)
            for (int k = 1; k <= nphi; k++) {
                const double w1 = (((2.0 * MCX_PI) * fabs(Rimg[i - 1])) * (MCX_AU * MCX_AU)) / (double)nphi;
                const double w2 = (((2.0 * MCX_PI) * fabs(Rimg[i])) * (MCX_AU * MCX_AU)) / (double)nphi;
                const size_t a = (size_t)(i - 1) * nphi + k - 1, b = (size_t)i * nphi + k - 1;
                fl = fl + ((Rimg[i] - Rimg[i - 1]) * (w1 * h_img[a] + w2 * h_img[b])) / 2.0;
                sfl = sfl + ((Rimg[i] - Rimg[i - 1]) * (w1 * h_sim[a] + w2 * h_sim[b])) / 2.0;
                if (scat_how == 2) fq = fq + ((Rimg[i] - Rimg[i - 1]) * (w1 * h_imQ[a] + w2 * h_imQ[b])) / 2.0;
            }
        flux[il] = fl;
        if (scatflux) scatflux[il] = sfl;
        if (fluxQ) fluxQ[il] = fq;
        if (ro && il == nlam_obs - 1) {
            ro->R = Rimg;
            ro->nimg = nimg;
            ro->nphi = nphi;
            ro->I = h_img;
            ro->Q = h_imQ;
            ro->U.assign(nray, 0.0);
            ro->V.assign(nray, 0.0);
            HIPCHK(ctx, hipMemcpyAsync(ro->U.data(), imageU, sizeof(double) * nray, hipMemcpyDeviceToHost, st));
            HIPCHK(ctx, hipMemcpyAsync(ro->V.data(), imageV, sizeof(double) * nray, hipMemcpyDeviceToHost, st));
            HIPCHK(ctx, hipStreamSynchronize(st));
        }
    }
    return MCX_OK;
}

extern "C" int mcx_image(mcx_ctx *ctx, const mcx_obs *o, const double *T, const double *theta_av, double lam,
                         const mcx_image_opts *io, double *im, double *imQ, double *imU, double *imV,
                         int32_t *nimg, int32_t ray_cap, double *Rimg, double *rays) {
    if (!ctx || !o || !io || !im || !nimg) return MCX_EINVAL;
    if (io->npix < 1 || io->nint < 1 || !(io->Rmax > 0.0) || !(lam > 0.0))
        return fail(ctx, MCX_EINVAL, "image: need npix, nint >= 1, Rmax > 0 and lam > 0");
    if (ctx->tb.scat_how == 2 && (!imQ || !imU || !imV))
        return fail(ctx, MCX_EINVAL, "image: full scattering needs the Q, U and V images");
    RayImages ro;
    double fl = 0.0, sfl = 0.0, fq = 0.0;
    const auto t0 = std::chrono::steady_clock::now();
    int r = sed_core(ctx, o, T, theta_av, 1, &lam, &fl, &sfl, &fq, lam, &ro);
    if (r) return r;
    ctx->image_rays_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    ctx->image_spread_ms = 0.0;
    *nimg = ro.nimg;
    const int nray = ro.nimg * ro.nphi;
    if (Rimg || rays) {
        if (ray_cap < ro.nimg) return fail(ctx, MCX_EINVAL, "image: ray_cap below the number of rings (see *nimg)");
        if (Rimg) memcpy(Rimg, ro.R.data(), sizeof(double) * ro.nimg);
        if (rays) {
            const std::vector<double> *src[4] = {&ro.I, &ro.Q, &ro.U, &ro.V};
            for (int q = 0; q < 4; q++) memcpy(rays + (size_t)q * nray, src[q]->data(), sizeof(double) * nray);
        }
    }
    hipStream_t st = ctx->stream;
    std::vector<void *> tmp;
    struct Guard {
        mcx_ctx *c;
        std::vector<void *> &l;
        ~Guard() { pfree_list(c, l); }
    } guard{ctx, tmp};
    const size_t npx = (size_t)io->npix * io->npix;
    double *dR, *dI, *dQ, *dU, *dV, *dim;
    if ((r = pupload(ctx, tmp, &dR, ro.R.data(), ro.R.size()))) return r;
    if ((r = pupload(ctx, tmp, &dI, ro.I.data(), (size_t)nray))) return r;
    if ((r = pupload(ctx, tmp, &dQ, ro.Q.data(), (size_t)nray))) return r;
    if ((r = pupload(ctx, tmp, &dU, ro.U.data(), (size_t)nray))) return r;
    if ((r = pupload(ctx, tmp, &dV, ro.V.data(), (size_t)nray))) return r;
    if ((r = palloc(ctx, tmp, &dim, 4 * npx))) return r;
    double *dfac;
    if ((r = palloc(ctx, tmp, &dfac, 5 * (size_t)std::max(nray, 1)))) return r;
    if (nray > 0) {
        hipLaunchKernelGGL(k_image_rays, dim3((unsigned)((nray + 255) / 256)), dim3(256), 0, st, nray,
                           ctx->tb.scat_how, dI, dQ, dU, dV, dfac, dfac + nray, dfac + 2 * (size_t)nray,
                           dfac + 3 * (size_t)nray, dfac + 4 * (size_t)nray);
        HIPCHK(ctx, hipGetLastError());
    }
    // zeroed by copying a host array of zeros
    std::vector<double> zero(4 * npx, 0.0);
    HIPCHK(ctx, hipMemcpyAsync(dim, zero.data(), sizeof(double) * 4 * npx, hipMemcpyHostToDevice, st));
    ImgArgs A;
    A.nimg = ro.nimg;
    A.nphi = ro.nphi;
    A.npix = io->npix;
    A.nint = io->nint;
    A.scat_how = ctx->tb.scat_how;
    A.R = dR;
    A.I = dI;
    A.Q = dQ;
    A.U = dU;
    A.V = dV;
    A.liv = dfac;
    A.lal = dfac + nray;
    A.lar = dfac + 2 * (size_t)nray;
    A.cv = dfac + 3 * (size_t)nray;
    A.sv = dfac + 4 * (size_t)nray;
    A.Rmax = io->Rmax;
    A.pa = io->pa_deg;
    A.au = MCX_AU;
    A.seed = io->seed;
    A.im = dim;
    A.imQ = dim + npx;
    A.imU = dim + 2 * npx;
    A.imV = dim + 3 * npx;
    const long long ncellp = (long long)(ro.nimg - 1) * ro.nphi;
    if (ncellp > 0x7fffffffLL) return fail(ctx, MCX_EINVAL, "image: too many ring cells for one launch");
    if (ncellp > 0) {
        HIPCHK(ctx, hipEventRecord(ctx->ev[0], st));
        hipLaunchKernelGGL(k_image, dim3((unsigned)ncellp), dim3(64), 0, st, A);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipEventRecord(ctx->ev[1], st));
        HIPCHK(ctx, hipEventSynchronize(ctx->ev[1]));
        float ms = 0.f;
        HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
        ctx->image_spread_ms = ms;
    }
    double *outs[4] = {im, imQ, imU, imV};
    for (int q = 0; q < 4; q++)
        if (outs[q])
            HIPCHK(ctx, hipMemcpyAsync(outs[q], dim + (size_t)q * npx, sizeof(double) * npx, hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    return MCX_OK;
}

extern "C" int mcx_last_image_timing(const mcx_ctx *ctx, double *rays_ms, double *spread_ms) {
    if (!ctx) return MCX_EINVAL;
    if (rays_ms) *rays_ms = ctx->image_rays_ms;
    if (spread_ms) *spread_ms = ctx->image_spread_ms;
    return MCX_OK;
}
