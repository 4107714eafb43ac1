// structure.hip -- the disk's vertical structure on the MI355X, behind
// mcx_temp_average / mcx_disk_structure / mcx_compare_cells of
// include/mcx.h: TempAverage, DiskStructure with CheckMinimumDensity and
// CheckCells, and compareCells (DiskStructure.f:25-584, 2659-3012), with the
// Numerical Recipes Cash-Karp integrator DiskStructure calls
// (odeint/rkqs/rkck, Subroutines.f:1933-2085).
//
// Work split: every (radius, species) column is an independent hydrostatic
// integration from the midplane up, one thread each; inside a column the
// adaptive step control carries over from interval to interval, so the
// column stays sequential.  The per-cell composition update and the
// checks then run one thread per cell.  Both kernels are latency-bound FP64
// chains on a few hundred threads; the HBM traffic is a few times the
// grid's size.
#include "mcx_internal.h"

namespace {

// derivs (DiskStructure.f:469-493): mu = 2.3 m_p with 2.3 a default REAL,
// G in cgs, kb = 1.3806503d-16 (Modules.f:11)
#define MU_GAS ((double)2.3f * 1.67262158e-24)
#define G_GRAV 6.67300e-8
#define KB_CGS 1.3806503e-16

// Cash-Karp coefficients exactly as the reference declares them: default-REAL
// (single precision) constant expressions converted to real*8
#define F32(a) ((double)(float)(a))

struct VSeg {
    double z1, z2, T1, T2, F1, F2, dlnT, scale, Mstar, R3;
};

__device__ __forceinline__ double vs_derivs(const VSeg &s, double x) {
    const double beta = s.F1 + (s.F2 - s.F1) * (x - s.z1) / (s.z2 - s.z1);
    const double T = s.T1 + (s.T2 - s.T1) * (x - s.z1) / (s.z2 - s.z1);
    return -((((MU_GAS * G_GRAV) * s.Mstar) * (1.0 + beta)) * x) / ((((s.scale * s.scale) * KB_CGS) * T) * s.R3) -
           s.dlnT;
}

// one rkck trial step (Subroutines.f:2042-2085); the right-hand side does not
// depend on y, so only the abscissae x + A h are evaluated
__device__ __forceinline__ void vs_rkck(const VSeg &s, double y, double dydx, double x, double h, double &yout,
                                        double &yerr) {
    const double A3 = F32(.3f), A4 = F32(.6f), A5 = 1.0, A6 = F32(.875f);
    const double C1 = F32(37.f / 378.f), C3 = F32(250.f / 621.f), C4 = F32(125.f / 594.f), C6 = F32(512.f / 1771.f);
    const double DC1 = C1 - F32(2825.f / 27648.f), DC3 = C3 - F32(18575.f / 48384.f),
                 DC4 = C4 - F32(13525.f / 55296.f), DC5 = F32(-277.f / 14336.f), DC6 = C6 - F32(.25f);
    const double ak3 = vs_derivs(s, x + A3 * h);
    const double ak4 = vs_derivs(s, x + A4 * h);
    const double ak5 = vs_derivs(s, x + A5 * h);
    const double ak6 = vs_derivs(s, x + A6 * h);
    yout = y + h * (C1 * dydx + C3 * ak3 + C4 * ak4 + C6 * ak6);
    yerr = h * (DC1 * dydx + DC3 * ak3 + DC4 * ak4 + DC5 * ak5 + DC6 * ak6);
}

// odeint + rkqs for one interval (Subroutines.f:1933-2036), kmax = 0 and
// hmin = 0; on "too many steps" y keeps its start value as in the reference
__device__ double vs_odeint(const VSeg &s, double ystart, double eps, double h1, int &nsteps) {
    const double SAFETY = F32(0.9f), PGROW = F32(-.2f), PSHRNK = F32(-.25f), ERRCON = F32(1.89e-4f);
    const double x1 = s.z1, x2 = s.z2;
    double x = x1, h = copysign(h1, x2 - x1), y = ystart;
    for (int nstp = 1; nstp <= 10000; nstp++) {
        const double dydx = vs_derivs(s, x);
        const double yscal = fabs(y) + fabs(h * dydx) + 1.e-150;
        if ((x + h - x2) * (x + h - x1) > 0.0) h = x2 - x;
        double hnext, ytemp, yerr;
        for (;;) { // rkqs
            vs_rkck(s, y, dydx, x, h, ytemp, yerr);
            double errmax = fmax(0.0, fabs(yerr / yscal));
            errmax = errmax / eps;
            if (errmax > 1.0) {
                const double htemp = (SAFETY * h) * pow(errmax, PSHRNK);
                h = copysign(fmax(fabs(htemp), F32(0.1f) * fabs(h)), h);
                continue;
            }
            hnext = (errmax > ERRCON) ? (SAFETY * h) * pow(errmax, PGROW) : 5.0 * h;
            x = x + h;
            y = ytemp;
            break;
        }
        nsteps++;
        if ((x - x2) * (x2 - x1) >= 0.0) return y;
        h = hnext; // |hnext| < hmin = 0 never holds
    }
    return ystart;
}

struct ColArgs {
    int nr, nt, ng, mpstr;
    unsigned settle, gascoupled;     // bit s: species s
    const double *R_av, *theta_av;   // [nr], [nt]
    const double *V;                 // [ncell]
    const double *shscale;           // [nr]
    const double *gshscale;          // [ng*nr]
    const double *FradZ;             // [ncell] or null
    const double *Temp;              // [ncell]
    const double *dens0, *gasdens, *w0; // old state
    double Mstar;
    double *rho;                     // [(ng+1)*nr*nt]
    unsigned long long *nsteps;
};

// DiskStructure.f:137-389 for column (i, ii): ii = 0 is the gas, ii = s+1
// dust species s (scale-height type DISK)
__global__ void __launch_bounds__(64) k_columns(ColArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (a.nr - 1) * (a.ng + 1)) return;
    const int i = 1 + t / (a.ng + 1), ii = t % (a.ng + 1), nt = a.nt, ng = a.ng;
    const size_t c0 = (size_t)i * nt;
    double *rho = a.rho + ((size_t)ii * a.nr + i) * nt;
    double scale;
    if (ii != 0) {
        scale = a.shscale[i];
        if ((a.settle >> (ii - 1)) & 1u) scale = scale * a.gshscale[(size_t)(ii - 1) * a.nr + i];
        if (((a.gascoupled >> (ii - 1)) & 1u) && scale < 1.0) scale = 1.0;
    } else {
        scale = 1.0 * a.shscale[i];
    }
    double M0 = 0.0;
    for (int j = 0; j < nt; j++) {
        const size_t c = c0 + j;
        if (ii != 0) M0 = M0 + (a.dens0[c] * a.V[c]) * a.w0[c * ng + ii - 1];
        else M0 = M0 + a.gasdens[c] * a.V[c];
    }
    const double Rav = a.R_av[i];
    auto Tcol = [&](int j) { return a.mpstr ? a.Temp[c0 + nt - 1] : a.Temp[c0 + j]; };
    rho[nt - 1] = 0.0;
    int nsteps = 0;
    if (scale > 0.0) {
        VSeg s;
        s.scale = scale;
        s.Mstar = a.Mstar;
        s.R3 = Rav * (Rav * Rav);
        double zhi = Rav * cos(a.theta_av[nt - 1]), Thi = Tcol(nt - 1), lnThi = log(Thi);
        double Fhi = a.FradZ ? a.FradZ[c0 + nt - 1] : 0.0, y = 0.0;
        for (int j = nt - 2; j >= 0; j--) {
            const double zlo = Rav * cos(a.theta_av[j]), Tlo = Tcol(j), lnTlo = log(Tlo);
            const double Flo = a.FradZ ? a.FradZ[c0 + j] : 0.0;
            s.z1 = zhi; s.z2 = zlo; s.T1 = Thi; s.T2 = Tlo; s.F1 = Fhi; s.F2 = Flo;
            s.dlnT = (lnThi - lnTlo) / (zhi - zlo);
            const double h1 = fabs(zlo - zhi) * 1e-6;
            y = vs_odeint(s, y, 1e-6, h1, nsteps);
            rho[j] = y;
            zhi = zlo; Thi = Tlo; lnThi = lnTlo; Fhi = Flo;
        }
    } else {
        for (int j = nt - 2; j >= 0; j--) rho[j] = ((1.0 / tan(a.theta_av[j])) < -scale) ? 0.0 : -60.0;
    }
    double M1 = 0.0;
    for (int j = 0; j < nt; j++) M1 = M1 + exp(rho[j]) * a.V[c0 + j];
    for (int j = 0; j < nt; j++) {
        double r = (exp(rho[j]) * M0) / M1;
        rho[j] = (r < 1e-50) ? 1e-50 : r;
    }
    if (nsteps) atomicAdd(a.nsteps, (unsigned long long)nsteps);
}

struct CellArgs {
    int nr, nt, ng;
    const double *rho;
    double *dens, *dens0, *gasdens, *w, *w0;
};

// DiskStructure.f:399-447 (densities and abundances of a cell from its
// columns), CheckMinimumDensity (2966-3012), then CheckCells (2659-2702),
// which also runs over the void row
__global__ void k_cell_update(CellArgs a) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nr * a.nt) return;
    const int i = c / a.nt, j = c % a.nt, ng = a.ng;
    double *wc = a.w + (size_t)c * ng, *w0c = a.w0 + (size_t)c * ng;
    double dens = a.dens[c], dens0 = a.dens0[c];
    if (i >= 1) {
        auto R = [&](int ii) { return a.rho[((size_t)ii * a.nr + i) * a.nt + j]; };
        double gasf[MAXNG];
        for (int q = 0; q < ng; q++) {
            double g = (w0c[q] >= 1e-100) ? 1.0 - (dens * wc[q]) / (dens0 * w0c[q]) : 0.0;
            if (g < 0.0) g = 0.0;
            if (g > 1.0) g = 1.0;
            if (isnan(g)) g = 0.0;
            gasf[q] = g;
        }
        double gasdens = R(0);
        double d0 = 0.0;
        for (int q = 0; q < ng; q++) d0 = d0 + R(q + 1);
        dens0 = d0;
        if (dens0 > 1e-50)
            for (int q = 0; q < ng; q++) w0c[q] = R(q + 1) / dens0;
        double tot = 0.0;
        for (int q = 0; q < ng; q++) tot = tot + w0c[q];
        for (int q = 0; q < ng; q++) {
            w0c[q] = w0c[q] / tot;
            wc[q] = w0c[q] * (1.0 - gasf[q]);
        }
        tot = 0.0;
        for (int q = 0; q < ng; q++) tot = tot + wc[q];
        if (tot > 1e-50) {
            for (int q = 0; q < ng; q++) wc[q] = wc[q] / tot;
        } else {
            for (int q = 0; q < ng; q++) wc[q] = w0c[q];
        }
        dens = dens0 * tot;
        // CheckMinimumDensity
        if (dens0 <= 1e-50) dens0 = 1e-50;
        if (dens <= 1e-50) dens = 1e-50;
        if (gasdens <= 1e-50) gasdens = 1e-50;
        double mt[MAXNG];
        for (int q = 0; q < ng; q++) mt[q] = dens * wc[q];
        for (int q = 0; q < ng; q++) {
            if (mt[q] < 1e-50) mt[q] = 1e-50;
            if (mt[q] > dens0 * w0c[q]) mt[q] = dens0 * w0c[q];
        }
        double d = 0.0;
        for (int q = 0; q < ng; q++) d = d + mt[q];
        dens = d;
        if (dens <= 1e-50) {
            dens = 1e-50;
            for (int q = 0; q < ng; q++) wc[q] = w0c[q];
        } else {
            for (int q = 0; q < ng; q++) wc[q] = mt[q] / dens;
        }
        a.gasdens[c] = gasdens;
        a.dens0[c] = dens0;
    }
    // CheckCells
    double mass[MAXNG];
    bool rescale = false;
    for (int q = 0; q < ng; q++) {
        mass[q] = dens * wc[q];
        const double mmax = dens0 * w0c[q];
        if (mass[q] > mmax) {
            mass[q] = mmax;
            rescale = true;
        }
    }
    if (rescale) {
        double d = 0.0;
        for (int q = 0; q < ng; q++) d = d + mass[q];
        dens = d;
        if (dens > 1e-50) {
            for (int q = 0; q < ng; q++) wc[q] = mass[q] / dens;
        } else {
            dens = 1e-50;
            for (int q = 0; q < ng; q++) wc[q] = w0c[q];
        }
    }
    a.dens[c] = dens;
}

// TempAverage (DiskStructure.f:25-79)
__global__ void k_temp_average(int ncell, const double *__restrict__ T, double *Tav, double fw, int first) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    if (first) {
        Tav[c] = T[c];
        return;
    }
    double t = T[c];
    if (t < (double)2.7f) t = (double)2.7f;
    Tav[c] = (1.0 - fw) * Tav[c] + fw * t;
}

struct CmpArgs {
    int nr, nt;
    const double *a[6], *b[6]; // T, dT, dens, ddens, EJv, dEJv
    double *out;               // [3]
};

// compareCells (DiskStructure.f:553-584): one block; fixed-order tree sums
__global__ void __launch_bounds__(256) k_compare(CmpArgs p) {
    __shared__ double sh[3][256];
    double e[3] = {0.0, 0.0, 0.0};
    const int n = (p.nr - 1) * p.nt;
    for (int k = threadIdx.x; k < n; k += 256) {
        const int c = p.nt + k; // radii 1..nr-1
        if (p.a[1][c] > 1e-200 && p.b[1][c] > 1e-200) {
            for (int q = 0; q < 3; q++) {
                const double d = p.a[2 * q][c] - p.b[2 * q][c];
                const double s1 = p.a[2 * q + 1][c], s2 = p.b[2 * q + 1][c];
                e[q] = e[q] + (d * d) / (s1 * s1 + s2 * s2);
            }
        }
    }
    for (int q = 0; q < 3; q++) sh[q][threadIdx.x] = e[q];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < 3; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double N = (double)((float)(p.nr - 1) * (float)p.nt) - 1.0; // single-precision REAL product
        p.out[0] = sh[0][0] / N;
        p.out[1] = sqrt(sh[1][0] / N);
        p.out[2] = sh[2][0] / N;
    }
}

struct TmpList {
    std::vector<void *> v;
    ~TmpList() { free_list(v); }
};

} // namespace

void release_structure(mcx_ctx *ctx) {
    if (ctx->d_Tav) (void)hipFree(ctx->d_Tav);
    ctx->d_Tav = nullptr;
    ctx->tav_ncell = 0;
    ctx->have_tav = false;
}

extern "C" int mcx_temp_average(mcx_ctx *ctx, double fw, int32_t first, double *Tav) {
    if (!ctx) return MCX_EINVAL;
    if (!ctx->have_temps) return fail(ctx, MCX_ESTATE, "no temperatures yet (mcx_temperature or mcx_set_temperature)");
    if (!(fw >= 0.0 && fw <= 1.0)) return fail(ctx, MCX_EINVAL, "fw must lie in [0, 1]");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int ncell = ctx->md.ncell;
    if (ctx->tav_ncell != ncell) {
        release_structure(ctx);
        HIPCHK(ctx, hipMalloc(&ctx->d_Tav, sizeof(double) * ncell));
        ctx->tav_ncell = ncell;
    }
    if (!first && !ctx->have_tav) return fail(ctx, MCX_ESTATE, "the first mcx_temp_average call must pass first = 1");
    hipLaunchKernelGGL(k_temp_average, dim3((ncell + 255) / 256), dim3(256), 0, ctx->stream, ncell, ctx->d_T,
                       ctx->d_Tav, fw, (int)(first != 0));
    HIPCHK(ctx, hipGetLastError());
    if (Tav) HIPCHK(ctx, hipMemcpyAsync(Tav, ctx->d_Tav, sizeof(double) * ncell, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->have_tav = true;
    return MCX_OK;
}

extern "C" int mcx_disk_structure(mcx_ctx *ctx, const mcx_struct *st, const double *Temp, mcx_density *d,
                                  int64_t *nsteps) {
    if (!ctx || !st || !d) return MCX_EINVAL;
    if (!ctx->have_model) return fail(ctx, MCX_ESTATE, "mcx_set_model first");
    if (!st->R_av || !st->theta_av || !(st->Mstar > 0.0))
        return fail(ctx, MCX_EINVAL, "mcx_struct needs R_av, theta_av and Mstar > 0");
    if (!d->dens || !d->dens0 || !d->gasdens || !d->w || !d->w0)
        return fail(ctx, MCX_EINVAL, "mcx_density needs dens, dens0, gasdens, w and w0");
    const int nr = ctx->md.nr, nt = ctx->md.nt, ng = ctx->ng, ncell = nr * nt;
    if (!Temp && !(ctx->have_tav && ctx->tav_ncell == ncell))
        return fail(ctx, MCX_ESTATE, "no averaged temperatures for this grid (mcx_temp_average)");
    if (ng > MAXNG) return fail(ctx, MCX_EINVAL, "too many species");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    TmpList tmp;
    int r;
    std::vector<double> ones_r(nr, 1.0), ones_g((size_t)ng * nr, 1.0);
    double *dRav, *dtav, *dsh, *dgsh, *dF = nullptr, *dTemp = ctx->d_Tav, *dd, *dd0, *dgas, *dw, *dw0, *drho;
    unsigned long long *dn;
    if ((r = dupload(ctx, tmp.v, &dRav, st->R_av, (size_t)nr))) return r;
    if ((r = dupload(ctx, tmp.v, &dtav, st->theta_av, (size_t)nt))) return r;
    if ((r = dupload(ctx, tmp.v, &dsh, st->shscale ? st->shscale : ones_r.data(), (size_t)nr))) return r;
    if ((r = dupload(ctx, tmp.v, &dgsh, st->gshscale ? st->gshscale : ones_g.data(), (size_t)ng * nr))) return r;
    if (st->FradZ && (r = dupload(ctx, tmp.v, &dF, st->FradZ, (size_t)ncell))) return r;
    if (Temp && (r = dupload(ctx, tmp.v, &dTemp, Temp, (size_t)ncell))) return r;
    if ((r = dupload(ctx, tmp.v, &dd, (const double *)d->dens, (size_t)ncell))) return r;
    if ((r = dupload(ctx, tmp.v, &dd0, (const double *)d->dens0, (size_t)ncell))) return r;
    if ((r = dupload(ctx, tmp.v, &dgas, (const double *)d->gasdens, (size_t)ncell))) return r;
    if ((r = dupload(ctx, tmp.v, &dw, (const double *)d->w, (size_t)ncell * ng))) return r;
    if ((r = dupload(ctx, tmp.v, &dw0, (const double *)d->w0, (size_t)ncell * ng))) return r;
    if ((r = dalloc(ctx, tmp.v, &drho, (size_t)(ng + 1) * ncell))) return r;
    if ((r = dalloc(ctx, tmp.v, &dn, 1))) return r;
    HIPCHK(ctx, hipMemsetAsync(dn, 0, sizeof(unsigned long long), s));
    ColArgs a{};
    a.nr = nr; a.nt = nt; a.ng = ng; a.mpstr = st->mpstr != 0;
    for (int q = 0; q < ng; q++) {
        if (st->settle && st->settle[q]) a.settle |= 1u << q;
        if (st->gascoupled && st->gascoupled[q]) a.gascoupled |= 1u << q;
    }
    a.R_av = dRav; a.theta_av = dtav; a.V = ctx->d_V; a.shscale = dsh; a.gshscale = dgsh; a.FradZ = dF;
    a.Temp = dTemp; a.dens0 = dd0; a.gasdens = dgas; a.w0 = dw0; a.Mstar = st->Mstar; a.rho = drho; a.nsteps = dn;
    const int ncol = (nr - 1) * (ng + 1);
    if (ncol > 0) {
        hipLaunchKernelGGL(k_columns, dim3((ncol + 63) / 64), dim3(64), 0, s, a);
        HIPCHK(ctx, hipGetLastError());
    }
    CellArgs b{nr, nt, ng, drho, dd, dd0, dgas, dw, dw0};
    hipLaunchKernelGGL(k_cell_update, dim3((ncell + 255) / 256), dim3(256), 0, s, b);
    HIPCHK(ctx, hipGetLastError());
    unsigned long long hn = 0;
    HIPCHK(ctx, hipMemcpyAsync(d->dens, dd, sizeof(double) * ncell, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(d->dens0, dd0, sizeof(double) * ncell, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(d->gasdens, dgas, sizeof(double) * ncell, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(d->w, dw, sizeof(double) * ncell * ng, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(d->w0, dw0, sizeof(double) * ncell * ng, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(&hn, dn, sizeof(hn), hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    if (nsteps) *nsteps = (int64_t)hn;
    return MCX_OK;
}

extern "C" int mcx_compare_cells(mcx_ctx *ctx, const mcx_cellstate *prev, const mcx_cellstate *cur, double *errT,
                                 double *errRho, double *errE) {
    if (!ctx || !prev || !cur) return MCX_EINVAL;
    if (!ctx->have_model) return fail(ctx, MCX_ESTATE, "mcx_set_model first");
    const double *pa[6] = {prev->T, prev->dT, prev->dens, prev->ddens, prev->EJv, prev->dEJv};
    const double *pb[6] = {cur->T, cur->dT, cur->dens, cur->ddens, cur->EJv, cur->dEJv};
    for (int k = 0; k < 6; k++)
        if (!pa[k] || !pb[k]) return fail(ctx, MCX_EINVAL, "mcx_cellstate arrays must all be given");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int nr = ctx->md.nr, nt = ctx->md.nt, ncell = nr * nt;
    hipStream_t s = ctx->stream;
    TmpList tmp;
    CmpArgs p{};
    p.nr = nr; p.nt = nt;
    int r;
    for (int k = 0; k < 6; k++) {
        double *q;
        if ((r = dupload(ctx, tmp.v, &q, pa[k], (size_t)ncell))) return r;
        p.a[k] = q;
        if ((r = dupload(ctx, tmp.v, &q, pb[k], (size_t)ncell))) return r;
        p.b[k] = q;
    }
    if ((r = dalloc(ctx, tmp.v, &p.out, 3))) return r;
    hipLaunchKernelGGL(k_compare, dim3(1), dim3(256), 0, s, p);
    HIPCHK(ctx, hipGetLastError());
    double h[3];
    HIPCHK(ctx, hipMemcpyAsync(h, p.out, sizeof(h), hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    if (errT) *errT = h[0];
    if (errRho) *errRho = h[1];
    if (errE) *errE = h[2];
    return MCX_OK;
}
