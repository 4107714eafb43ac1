// diffuse.hip -- MCMax's diffusion approximation for optically thick cells
// with poor photon statistics (Diffuse, Diffusion.f:9-329, flux limiter
// off) on the MI355X, behind mcx_diffuse() of include/mcx.h.
//
// The cells to solve and their neighbours are numbered on the host in the
// reference's order; per iteration the device computes each numbered cell's
// Rosseland coefficient from the tabulated DiffCoeffCell (one thread per
// cell), assembles the 5-point log-r / theta stencil of the T^4 equation
// into a dense column-major matrix (one thread per row), solves it with
// rocSOLVER's LU with partial pivoting (what DGESV computes), and updates
// the solved cells.
#include "mcx_internal.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

__device__ __forceinline__ double sq(double x) { return x * x; }

struct DiffArgs {
    int N, nr, nt, NTh, ng, nphotmin;
    double dTmax, Rout;
    const int *celc, *celi, *celj, *number; // number[(i)*(NTh+1) + j], 1-based, 0 = none
    const double *r, *t;                    // [N+1] cell centres (1-based)
    double *T, *EJv;                        // [ncell]
    const double *dT;
    const int *Ni;
    int *thick;
    KDSrc KD;
    const int *cls;
    const double *dens, *w, *R, *Kp;
    double *Dc; // [N+1]
    int *ok;    // [N+1]
    double *A, *b;
};

// Diffusion.f:122-149: Dc = 1/(KDext rho) at int(T), the thick flag's
// optical-depth check to the outer edge, and which cells are kept fixed
__global__ void k_diff_coef(DiffArgs D) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (k > D.N) return;
    const int c = D.celc[k], i = D.celi[k], j = D.celj[k];
    int iT = (int)(D.T[c] / 1.0);
    if (iT <= 1) iT = 1;
    if (iT >= TMAXD - 1) iT = TMAXD - 1;
    const double Kext = kd_at(D.KD, D.cls[c], iT).x - 0.0;
    double dens = 0.0;
    for (int s = 0; s < D.ng; s++) dens = dens + D.w[(size_t)c * D.ng + s] * D.dens[c];
    const double Dc = 1.0 / (Kext * dens);
    D.Dc[k] = Dc;
    if (((D.Rout - D.R[i]) * MCX_AU / Dc) < 2.0) D.thick[c] = 0;
    int ok;
    if (D.Ni[c] < D.nphotmin || (D.dT[c] / D.T[c]) > D.dTmax) {
        ok = 0;
        if (i == 1 || i == D.nr - 1 || j == 1) ok = 1;
        if (!D.thick[c] && D.Ni[c] > 1) ok = 1;
    } else {
        ok = 1;
    }
    D.ok[k] = ok;
}

// Diffusion.f:203-253: row k of the system, column-major A(N x N), zeroed
__global__ void k_diff_rows(DiffArgs D) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (k > D.N) return;
    const int N = D.N, c = D.celc[k], i = D.celi[k], j = D.celj[k];
    double *A = D.A;
    auto at = [&](int col) -> double & { return A[(size_t)(k - 1) + (size_t)(col - 1) * N]; };
    if (D.ok[k]) {
        at(k) = 1.0;
        const double T = D.T[c];
        D.b[k - 1] = sq(T) * sq(T);
        return;
    }
    const int W = D.NTh + 1;
    int j0 = D.number[i * W + j], jm = D.number[(i - 1) * W + j], jp = D.number[(i + 1) * W + j];
    {
        // every cell solved here had its four neighbours numbered when it
        // was selected; a row that would reach an unnumbered one stays fixed
        const int jt = (j != D.NTh - 1) ? D.number[i * W + j + 1] : j0;
        if (jm < 1 || jp < 1 || D.number[i * W + j - 1] < 1 || jt < 1) {
            at(k) = 1.0;
            D.b[k - 1] = sq(D.T[c]) * sq(D.T[c]);
            return;
        }
    }
    const double *r = D.r, *t = D.t;
    double dx0 = log(r[jp] / r[jm]), dx1 = log(r[jp] / r[j0]), dx2 = log(r[j0] / r[jm]);
    double f1 = D.Dc[jp] * r[jp], f2 = D.Dc[jm] * r[jm];
    const double r3 = r[j0] * sq(r[j0]);
    at(jm) = at(jm) + f2 / ((dx2 * dx0) * r3);
    at(j0) = at(j0) - f2 / ((dx2 * dx0) * r3);
    at(jp) = at(jp) + f1 / ((dx1 * dx0) * r3);
    at(j0) = at(j0) - f1 / ((dx1 * dx0) * r3);
    double tjm;
    j0 = D.number[i * W + j];
    if (j != D.NTh - 1) {
        jm = D.number[i * W + j + 1];
        tjm = t[jm];
    } else {
        jm = D.number[i * W + j];
        tjm = MCX_PI - t[jm];
    }
    jp = D.number[i * W + j - 1];
    dx0 = t[jp] - tjm;
    dx1 = t[jp] - t[j0];
    dx2 = t[j0] - tjm;
    f1 = D.Dc[jp] * sin(t[jp]);
    f2 = D.Dc[jm] * sin(tjm);
    const double r2s = sq(r[j0]) * sin(t[j0]);
    at(jm) = at(jm) + f2 / ((dx2 * dx0) * r2s);
    at(j0) = at(j0) - f2 / ((dx2 * dx0) * r2s);
    at(jp) = at(jp) + f1 / ((dx1 * dx0) * r2s);
    at(j0) = at(j0) - f1 / ((dx1 * dx0) * r2s);
    D.b[k - 1] = 0.0;
}

// Diffusion.f:264-284: the solved cells' T = (T^4)^(1/4) and EJv
__global__ void k_diff_update(DiffArgs D) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (k > D.N) return;
    if (D.ok[k]) return;
    const int c = D.celc[k];
    const double x = D.b[k - 1];
    double T = D.T[c];
    if (x > 0.0) T = pow(x, 0.25);
    if (x < 0.0) T = 1.0;
    int iT = (int)(T / 1.0);
    if (iT <= 1) iT = 1;
    if (iT >= TMAXD - 1) iT = TMAXD - 1;
    double Etot = 0.0;
    for (int s = 0; s < D.ng; s++) Etot = Etot + (D.w[(size_t)c * D.ng + s] * 1.0) * D.Kp[(size_t)s * (TMAXD + 1) + iT];
    D.EJv[c] = Etot;
    const double Tmx = (double)((float)(TMAXD - 1) * 1.0f);
    if (T > Tmx) T = Tmx;
    if (T < 1.0) T = 1.0;
    if (isnan(T) || isinf(T)) T = 1.0;
    D.T[c] = T;
}

extern "C" int mcx_diffuse(mcx_ctx *ctx, int32_t nphotmin, double dTmax, double factRW, const double *R_av,
                           const double *theta_av, double *T, int32_t *thick_out, double *EJv, int32_t *ncells) {
    if (!ctx || !R_av || !theta_av) return MCX_EINVAL;
    if (!ctx->have_temps) return fail(ctx, MCX_ESTATE, "mcx_temperature first");
    if (!ctx->rw.KD.on) return fail(ctx, MCX_EINVAL, "Diffuse needs Kext and gsym (mcx_set_tables), thet (mcx_set_model)");
    if (!(factRW > 0.0)) return fail(ctx, MCX_EINVAL, "factRW must be positive");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const DevModel &md = ctx->md;
    const int nr = md.nr, nt = md.nt, NTh = nt + 1, ncell = md.ncell, ng = ctx->ng;
    std::vector<void *> tmp;
    struct Guard {
        std::vector<void *> &l;
        ~Guard() { free_list(l); }
    } guard{tmp};
    int r;
    // C%thick as InitRandomWalk sets it (RandomWalk.f:174-189)
    hipLaunchKernelGGL(k_thick_cells, dim3((ncell + 255) / 256), dim3(256), 0, st, nr, nt, ctx->rw.R, md.dens,
                       ctx->rw.cls, ctx->rw.KD, factRW, ctx->d_thick);
    HIPCHK(ctx, hipGetLastError());
    std::vector<double> hT(ncell), hdT(ncell);
    std::vector<int> hNi(ncell), hth(ncell);
    HIPCHK(ctx, hipMemcpyAsync(hT.data(), ctx->d_T, sizeof(double) * ncell, hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipMemcpyAsync(hdT.data(), ctx->d_dT, sizeof(double) * ncell, hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipMemcpyAsync(hNi.data(), ctx->d_Ni, sizeof(int) * ncell, hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipMemcpyAsync(hth.data(), ctx->d_thick, sizeof(int) * ncell, hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    // number the cells to solve and their neighbours (Diffusion.f:40-73)
    const int W = NTh + 1;
    std::vector<int> number((size_t)(nr + 1) * W, 0);
    auto NUM = [&](int i, int j) -> int & { return number[(size_t)i * W + j]; };
    int N = 0, ncor = 0;
    for (int i = 2; i <= nr - 2; i++)
        for (int j = 2; j <= NTh - 1; j++) {
            const int c = i * nt + j - 1;
            if (hNi[c] < nphotmin || (hdT[c] / hT[c]) > dTmax) {
                if (hth[c] || hNi[c] <= 1) {
                    ncor++;
                    if (!NUM(i, j)) NUM(i, j) = ++N;
                    if (!NUM(i - 1, j)) NUM(i - 1, j) = ++N;
                    if (!NUM(i + 1, j)) NUM(i + 1, j) = ++N;
                    if (!NUM(i, j - 1)) NUM(i, j - 1) = ++N;
                    if (!NUM(i, j + 1) && j != NTh - 1) NUM(i, j + 1) = ++N;
                }
            }
        }
    if (ncells) *ncells = ncor;
    if (N >= 1) {
        if ((size_t)N * N > ((size_t)1 << 32)) return fail(ctx, MCX_EINVAL, "too many diffusion cells for a dense solve");
        std::vector<int> celc(N + 1, 0), celi(N + 1, 0), celj(N + 1, 0);
        std::vector<double> rr(N + 1, 0.0), tt(N + 1, 0.0);
        for (int i = 1; i <= nr - 1; i++)
            for (int j = 1; j <= NTh - 1; j++)
                if (NUM(i, j)) {
                    const int k = NUM(i, j);
                    celc[k] = i * nt + j - 1;
                    celi[k] = i;
                    celj[k] = j;
                    rr[k] = R_av[i];
                    tt[k] = theta_av[j - 1];
                }
        DiffArgs D;
        D.N = N;
        D.nr = nr;
        D.nt = nt;
        D.NTh = NTh;
        D.ng = ng;
        D.nphotmin = nphotmin;
        D.dTmax = dTmax;
        D.Rout = ctx->h_R[nr];
        int *dc, *di, *dj, *dn, *dok;
        double *dr, *dt, *dD, *dA, *db, *dEJv;
        if ((r = dupload(ctx, tmp, &dc, celc.data(), celc.size()))) return r;
        if ((r = dupload(ctx, tmp, &di, celi.data(), celi.size()))) return r;
        if ((r = dupload(ctx, tmp, &dj, celj.data(), celj.size()))) return r;
        if ((r = dupload(ctx, tmp, &dn, number.data(), number.size()))) return r;
        if ((r = dupload(ctx, tmp, &dr, rr.data(), rr.size()))) return r;
        if ((r = dupload(ctx, tmp, &dt, tt.data(), tt.size()))) return r;
        if ((r = dalloc(ctx, tmp, &dD, (size_t)N + 1))) return r;
        if ((r = dalloc(ctx, tmp, &dok, (size_t)N + 1))) return r;
        if ((r = dalloc(ctx, tmp, &dA, (size_t)N * N))) return r;
        if ((r = dalloc(ctx, tmp, &db, (size_t)N))) return r;
        if (EJv) {
            if ((r = dupload(ctx, tmp, &dEJv, EJv, (size_t)ncell))) return r;
        } else {
            if ((r = dalloc(ctx, tmp, &dEJv, (size_t)ncell))) return r;
            HIPCHK(ctx, hipMemsetAsync(dEJv, 0, sizeof(double) * ncell, st));
        }
        rocblas_int *ipiv, *info;
        if ((r = dalloc(ctx, tmp, &ipiv, (size_t)N))) return r;
        if ((r = dalloc(ctx, tmp, &info, 1))) return r;
        D.celc = dc;
        D.celi = di;
        D.celj = dj;
        D.number = dn;
        D.r = dr;
        D.t = dt;
        D.T = ctx->d_T;
        D.EJv = dEJv;
        D.dT = ctx->d_dT;
        D.Ni = ctx->d_Ni;
        D.thick = ctx->d_thick;
        D.KD = ctx->rw.KD;
        D.cls = ctx->rw.cls;
        D.dens = md.dens;
        D.w = md.w;
        D.R = ctx->rw.R;
        D.Kp = ctx->tb.Kp;
        D.Dc = dD;
        D.ok = dok;
        D.A = dA;
        D.b = db;
        if (!ctx->blas) {
            if (rocblas_create_handle((rocblas_handle *)&ctx->blas) != rocblas_status_success)
                return fail(ctx, MCX_EDEVICE, "rocblas_create_handle failed");
        }
        rocblas_handle h = (rocblas_handle)ctx->blas;
        if (rocblas_set_stream(h, st) != rocblas_status_success)
            return fail(ctx, MCX_EDEVICE, "rocblas_set_stream failed");
        const unsigned nb = (unsigned)((N + 255) / 256);
        for (int iter = 1; iter <= 3; iter++) {
            hipLaunchKernelGGL(k_diff_coef, dim3(nb), dim3(256), 0, st, D);
            HIPCHK(ctx, hipMemsetAsync(dA, 0, sizeof(double) * (size_t)N * N, st));
            hipLaunchKernelGGL(k_diff_rows, dim3(nb), dim3(256), 0, st, D);
            HIPCHK(ctx, hipGetLastError());
            if (rocsolver_dgesv(h, N, 1, dA, N, ipiv, db, N, info) != rocblas_status_success)
                return fail(ctx, MCX_EDEVICE, "rocsolver_dgesv failed");
            rocblas_int hinfo = 0;
            HIPCHK(ctx, hipMemcpyAsync(&hinfo, info, sizeof(rocblas_int), hipMemcpyDeviceToHost, st));
            HIPCHK(ctx, hipStreamSynchronize(st));
            if (hinfo != 0) return fail(ctx, MCX_EGEOM, "diffusion matrix is singular");
            hipLaunchKernelGGL(k_diff_update, dim3(nb), dim3(256), 0, st, D);
            HIPCHK(ctx, hipGetLastError());
        }
        if (EJv) HIPCHK(ctx, hipMemcpyAsync(EJv, dEJv, sizeof(double) * ncell, hipMemcpyDeviceToHost, st));
    }
    if (T) HIPCHK(ctx, hipMemcpyAsync(T, ctx->d_T, sizeof(double) * ncell, hipMemcpyDeviceToHost, st));
    if (thick_out) HIPCHK(ctx, hipMemcpyAsync(thick_out, ctx->d_thick, sizeof(int) * ncell, hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    return MCX_OK;
}

void release_blas(mcx_ctx *ctx) {
    if (ctx && ctx->blas) {
        (void)rocblas_destroy_handle((rocblas_handle)ctx->blas);
        ctx->blas = nullptr;
    }
}
