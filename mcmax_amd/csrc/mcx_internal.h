// mcx_internal.h -- definitions shared by the translation units of libmcx.so
// (transport: mcx.hip; ray-traced SED: sed.hip; diffusion: diffuse.hip;
// vertical structure: structure.hip): device-resident tables and
// model, tallies, and the host context behind the C-ABI of include/mcx.h.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>
#include <stdio.h>
#include <stdlib.h>
#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "../../include/mcx.h"
#include "mcx_device.h"

#define TMAXD MCX_DTMAX
#define NANGD MCX_DNANG
#define MAXNG 16

// fixed-point resolutions
#define EJV_BITS 31   // one EJv event of the cell's bound maps to 2^31
#define EJV2_BITS 40  // one squared bound maps to 2^40 (3 x 32-bit limbs)
#define EABS_BITS 24  // one nominal packet energy maps to 2^24

enum { C_NEMIT = 0, C_NINTERACT, C_NABS, C_NSCAT, C_NESCAPE, C_NSTARHIT, C_NSTEPS, C_NRAAR, C_NMAXINT,
       C_NGEOM, C_NOVERFLOW, C_NRW, C_NVISOVF, C_NTIMEOUT, C_NTALLYATOM, C_NCOUNT };

// per-lane table of the cells the lane's current packet has entered (Ni's
// distinct-packet count and the packet's own absorbed energy per cell):
// 2^VIS_LOG2 slots of {tag = packet + 1, cell, own energy lo, hi}, linear
// probing over at most VIS_PROBE slots
#define VIS_LOG2 10
#define VIS_PROBE 64

// slots of DevTally::counters beyond the event counters: the longest packet
// history (cell steps), packets finished in the current launch, and wall-clock
// stamps of the launch's start and of 50/90/99/99.9/100% finished
#define G_MAXSTEPS 15
#define G_DONE 16
#define G_TLINE 17
#define G_CLOCK 23
#define G_TOREP 24 // set by the first packet abandoned at the time limit
#define G_NSLOTS 25

// Stages of a concurrent pass (at most KSTAGE_MAX) and the readiness counters
// of their gates: gate[G_FIN + k] packets that finished with their last
// deposit in epoch k, gate[G_PASS + k] packets of earlier stages that went
// past the deposits stage k waits for, gate[G_CHUNK + k] / gate[G_CHDONE + k]
// chunks of stage k's snapshot claimed / built, gate[G_OPEN] the highest
// stage whose snapshot is complete.
#define KSTAGE_MAX 40
#define G_FIN 0
#define G_PASS (KSTAGE_MAX + 2)
#define G_CHUNK (2 * (KSTAGE_MAX + 2))
#define G_CHDONE (3 * (KSTAGE_MAX + 2))
#define G_OPEN (4 * (KSTAGE_MAX + 2))
#define G_TOPEN (4 * (KSTAGE_MAX + 2) + 1) // wall clock when stage k opened
#define G_GATE_SLOTS (5 * (KSTAGE_MAX + 2) + 1)

#define NY_RW 1000 // survival-table length (RandomWalk.f:6-7)

struct DevTables {
    int nlam, ng, scat_how, scattering;
    const double *nu;        // [nlam]
    const double *cumStar;   // [nlam-1]
    double Lstar;
    const double *CUM;       // [ng][TMAX+1][nlam-1]
    const double *Kp;        // [ng][TMAX+1]
    const double *Ksca;      // [ng][nlam]
    const double *F;         // [ng][nlam][6][180]
    const double *SC;        // [ng][nlam][180][8] per angle bin: the cumulatives
                             // sum pi sin(theta) F11, F12 up to the bin, then
                             // F11, F12, F22, F33, F34, F44 of the bin (one line)
    const double *IF11, *IF12;
    const double *cos2phi, *sin2phi; // [361]
    const double *pc2, *ps2;         // [181] prefix sums of cos 2phi, sin 2phi (bins 1..k)
};

struct DevModel {
    int nr, nt, ncell;
    const double *R2, *T2, *Theta; // [nr+1], [nt+1], [nt+1]
    double R0;                     // R[0] in AU
    const double *dens, *mass;     // [ncell]
    const double *w;               // [ncell*ng]
    const double2 *KT;             // [ncell*nlam] {KabsTot, KscaTot}
    const double *KpCls;           // [nclass*(TMAX+1)] sum_s w_s Kp_s(T) of each species-mix
                                   // class (computeE as a table; cell c: row cls[c]);
                                   // null: summed at each read (KpRow)
    const int *cls;                // [ncell] species-mix class of the cell
    const double *CUMmix;          // [nclass][TMAX+1][nlam-1] sum_s w_s CUM_s of each
                                   // class (several species; null: sum per event)
    const double *wcls;            // [nclass*ng] weights of each class
    const double *Kp;              // [ng][TMAX+1] the species' Planck means
    int ng;
};

// The Planck-mean row sum_s w_s Kp_s(T) of a species mix (computeE,
// Radiation.f:611-627): the class table, or where the context keeps none the
// species' rows summed at each read, in the order the tables are built from
// (so both give the same bits).
__device__ __attribute__((noinline)) inline double kp_sum(const double *Kp, const double *w, int ng, int j) {
    double e = 0.0;
    for (int s = 0; s < ng; s++) e = e + (Kp[(size_t)s * (TMAXD + 1) + j] * w[s]) * 1.0;
    return e;
}

struct KpRow {
    const double *t, *Kp, *w;
    int ng;
    __device__ __forceinline__ double operator[](int j) const { return t ? t[j] : kp_sum(Kp, w, ng, j); }
};

__device__ __forceinline__ KpRow kp_row(const DevModel &md, int k) {
    return KpRow{md.KpCls ? md.KpCls + (size_t)k * (TMAXD + 1) : nullptr, md.Kp, md.wcls + (size_t)k * md.ng, md.ng};
}

// Where the diffusion coefficients {KDext, KDabs} of a species mix at integer
// T come from: the class table t, or (t null) DiffCoeffCell evaluated at each
// read, from the class's per-wavelength terms AK and, at the four
// temperatures of the thick test, from K4.  on: Kext and g are known.
struct KDSrc {
    int on;
    const double2 *t;                  // [nclass*(TMAX+1)]
    const double2 *AK;                 // [nclass*nlam] {Kext - g^2 Ksca, Kabs} of the mix
    const double2 *K4;                 // [nclass*4] at T = 100, 600, 1100, 1600
    const double *Kabs, *Ksca, *Kext;  // [ng][nlam]
    const double *gsym, *wcls;         // [ng][nlam], [nclass*ng]
    const double *BB, *nu;             // [TMAX+1][nlam], [nlam]
    int nlam, ng, scattering;
};

// The mix's terms of DiffCoeffCell at wavelength l (Dtable.f:120-160): the
// Rosseland denominator Kext - g^2 Ksca and Kabs.  The species sums are
// those the MC pass sees (KabsTot/KscaTot), except at the four temperatures
// InitRandomWalk evaluates first (kabstot false: Kext - Ksca from the species
// tables, kept by the one-species cache of Dtable.f:107-112, 175-179).
__device__ __forceinline__ double2 kd_terms(const KDSrc &S, const double *w, int l, bool kabstot) {
    const int nlam = S.nlam, ng = S.ng;
    double ka = 0.0, ks = 0.0, ke = 0.0, gg = 0.0;
    if (kabstot) {
        for (int s = 0; s < ng; s++) {
            ka = ka + (S.Kabs[(size_t)s * nlam + l] * w[s]) * 1.0;
            ks = ks + (S.Ksca[(size_t)s * nlam + l] * w[s]) * 1.0;
        }
        ke = ka + ks;
    } else {
        for (int s = 0; s < ng; s++) {
            ke = ke + (S.Kext[(size_t)s * nlam + l] * w[s]) * 1.0;
            ks = ks + (S.Ksca[(size_t)s * nlam + l] * w[s]) * 1.0;
        }
        ka = ke - ks;
    }
    for (int s = 0; s < ng; s++)
        gg = gg + ((S.gsym[(size_t)s * nlam + l] * S.Ksca[(size_t)s * nlam + l]) * w[s]) * 1.0;
    const double gl = S.scattering ? gg / ks : 0.0;
    return make_double2(ke - (gl * gl) * ks, ka - 0.0);
}

// DiffCoeffCell (Dtable.f:97-182) of class k at integer T:
// KDext = int dB/dT dnu / int dB/dT / (Kext - g^2 Ksca) dnu,
// KDabs = int dB/dT Kabs dnu / int dB/dT dnu, trapezoids over nu.
// With AK the per-wavelength terms are read, else summed over the species
// (the same bits).
__device__ __attribute__((noinline)) inline double2 dev_diffcoeff(const KDSrc &S, int k, int iT) {
    if (iT < 1) iT = 1;
    if (iT > TMAXD - 1) iT = TMAXD - 1;
    const int nlam = S.nlam;
    const bool kabstot = !(iT == 100 || iT == 600 || iT == 1100 || iT == 1600);
    const bool tab = kabstot && S.AK;
    const double *w = S.wcls + (size_t)k * S.ng;
    const double2 *ak = tab ? S.AK + (size_t)k * nlam : nullptr;
    const double *B0 = S.BB + (size_t)iT * nlam, *B1 = S.BB + (size_t)(iT + 1) * nlam;
    double i1 = 0.0, i2 = 0.0, i3 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
    for (int l = 0; l < nlam; l++) {
        const double2 q = tab ? ak[l] : kd_terms(S, w, l, kabstot);
        const double dB = B1[l] - B0[l];
        const double s1 = dB / q.x, s2 = dB, s3 = dB * q.y;
        if (l > 0) {
            const double h = fabs(S.nu[l - 1] - S.nu[l]) * 0.5;
            i1 = i1 + h * (p1 + s1);
            i2 = i2 + h * (p2 + s2);
            i3 = i3 + h * (p3 + s3);
        }
        p1 = s1;
        p2 = s2;
        p3 = s3;
    }
    return make_double2(i2 / i1, i3 / i2);
}

__device__ __forceinline__ double2 kd_at(const KDSrc &S, int k, int iT) {
    if (S.t) return S.t[(size_t)k * (TMAXD + 1) + iT];
    if (S.K4 && (iT == 100 || iT == 600 || iT == 1100 || iT == 1600)) return S.K4[(size_t)k * 4 + (iT - 100) / 500];
    return dev_diffcoeff(S, k, iT);
}

struct DevTally {
    unsigned long long *EJv;    // [ncell]
    unsigned long long *EJv2;   // [ncell*3] limbs of weight 2^0, 2^32, 2^64
    unsigned long long *Eabs;   // [KSTAGE_MAX+2][ncell] absorbed energy by deposit epoch (one-lane
                                // passes: epoch 0 only); the pass's Eabs is the sum
    const double *invU1, *invU2;
    unsigned long long *Tcell;  // [ncell] bits of the latest BW T (monotone; one-lane passes)
    int *Ni;                    // [ncell] distinct packets per cell
    uint4 *vtab;                // [lanes << VIS_LOG2] visited cells of each lane's packet
    unsigned long long *Sk;     // [KSTAGE_MAX+1][ncell] snapshot S_k of stage k (epochs 1..k)
    float *Tk;                  // [KSTAGE_MAX+1][ncell] T(S_k/m): start of stage k's T searches
    unsigned long long *gate;   // [G_GATE_SLOTS] the stages' readiness counters
    unsigned long long *specE;  // [2][(NANG+1)*nlam]
    unsigned *ispec;            // [2][(NANG+1)*nlam]
    unsigned long long *counters;
    double *tautot;
    double *torep;  // [16] the first abandoned packet's state (time limit)
    // the pass's results as doubles (k_finalize)
    double *oEJv, *oEJv2, *oEabs, *oTbw;
    int *thick;     // C%thick of the pass (null without random-walk tables)
};

// modified random walk (RandomWalk.f)
struct DevRW {
    int on;
    double factRW;
    const int *thick;     // [ncell]
    const int *cls;       // [ncell] species-mix class of the cell
    KDSrc KD;             // {KDext, KDabs} of each class (DiffCoeffCell)
    const double *R;      // [nr+1] radial edges (AU)
    const double *sinT;   // [nt+1] sin(theta) of the cone edges
    const double *y, *phi; // [NY_RW] (InitRandomWalk)
};

struct mcx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::string err;
    std::vector<void *> tab_alloc, mod_alloc, run_alloc;
    bool have_tables = false, have_model = false, have_pass = false;
    // tables
    DevTables tb{};
    int nlam = 0, ng = 0;
    double Lstar = 0.0;
    std::vector<double> h_Kp, h_nu;
    double *d_Kabs = nullptr, *d_Ksca = nullptr;
    double *d_Kext = nullptr, *d_gsym = nullptr, *d_BB = nullptr; // Kext, g may stay null
    std::vector<double> h_lam, h_Fstar, h_Kext;                      // host copies for mcx_sed
    const double *d_rwy = nullptr, *d_rwphi = nullptr;
    // random walk, per model
    DevRW rw{};
    int nclass = 0;
    int *d_thick = nullptr;
    double *d_wcls = nullptr; // [nclass][ng] weights of each species-mix class
    std::vector<double> h_R, h_Theta, h_thet, h_dens, h_w;
    bool have_temps = false;
    void *blas = nullptr; // rocBLAS handle of the dense solves (mcx_diffuse)
    double *d_Tav = nullptr; // TempAverage's running average (mcx_temp_average)
    int tav_ncell = 0;
    bool have_tav = false;
    // model
    DevModel md{};
    double *d_V = nullptr, *d_R2 = nullptr;
    double Tvoid = 1.0;
    // run
    DevTally ty{};
    size_t vtab_lanes = 0;      // lanes the visit table holds
    unsigned long long *d_Sk = nullptr, *d_gate = nullptr;
    long long *d_stages = nullptr;  // the stage table of the last pass run from this context
    float *d_Tk = nullptr;
    int *d_Ni = nullptr;
    double *d_EJv = nullptr, *d_EJv2 = nullptr, *d_Eabs = nullptr, *d_Tbw = nullptr;
    double *d_T = nullptr, *d_dT = nullptr, *d_dEJv = nullptr;
    unsigned long long *d_next = nullptr;
    void *d_bms = nullptr;      // launch descriptors of the models of a pass run from this context
    size_t bm_cap = 0;
    void *d_bmap = nullptr;     // the model of each workgroup (explicit shares of a batched launch)
    size_t bmap_cap = 0;
    void *graph = nullptr, *gexec = nullptr; // captured pass (mcx_batch_transport with a graph)
    std::string graph_key;
    double unitE = 0.0, E0 = 0.0;
    double kernel_ms = 0.0, pass_ms = 0.0;
    double image_rays_ms = 0.0, image_spread_ms = 0.0;  // last mcx_image: ray part (host clock), k_image (events)
    int sm_count = 0;
    int wclk_khz = 0; // rate of the device's constant wall clock (wall_clock64)
    size_t lds_max = 65536; // LDS bytes one workgroup may allocate
};

static inline int fail(mcx_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}

#define HIPCHK(ctx, call)                                                                                     \
    do {                                                                                                       \
        hipError_t e_ = (call);                                                                                \
        if (e_ != hipSuccess) return fail(ctx, MCX_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
static int dalloc(mcx_ctx *ctx, std::vector<void *> &list, T **p, size_t n) {
    void *q = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&q, n * sizeof(T));
    if (e != hipSuccess) return fail(ctx, MCX_ENOMEM, "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed");
    list.push_back(q);
    *p = (T *)q;
    return 0;
}

template <typename T>
static int dupload(mcx_ctx *ctx, std::vector<void *> &list, T **p, const T *h, size_t n) {
    int r = dalloc(ctx, list, p, n);
    if (r) return r;
    HIPCHK(ctx, hipMemcpyAsync(*p, h, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return 0;
}

static inline void free_list(std::vector<void *> &list) {
    for (void *q : list) (void)hipFree(q);
    list.clear();
}

// C%thick (RandomWalk.f:174-189): the shell is wider than factRW Rosseland
// lengths at any of T = 100, 600, 1100, 1600 K.
static __global__ void k_thick_cells(int nr, int nt, const double *__restrict__ R, const double *__restrict__ dens,
                                     const int *__restrict__ cls, const KDSrc KD, double factRW, int *thick) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nr * nt) return;
    const int i = c / nt;
    int th = 0;
    if (i > 0) {
        for (int iT = 100; iT <= 1600; iT += 500) {
            const double lr = 1.0 / ((MCX_AU * kd_at(KD, cls[c], iT).x) * dens[c]);
            if ((R[i + 1] - R[i]) > factRW * lr) th = 1;
        }
    }
    thick[c] = th;
}

// frees the context's rocBLAS handle (defined with the dense solves)
void release_blas(mcx_ctx *ctx);
// frees the averaged temperatures (defined with the vertical structure)
void release_structure(mcx_ctx *ctx);
