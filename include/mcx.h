/*
 * mcx.h -- C-ABI of the MI355X Monte Carlo photon-transport library
 * (libmcx.so).  It replaces one MCMax stage: the Monte Carlo pass
 *     subroutine MCRadiation(NphotTot,niter,dofastvisc)   MCRadiation.f:6
 * together with the Lucy temperature inversion that follows it in the
 * driver (Main.f:498-558), and the steps either side of it: the diffusion
 * approximation (Diffuse), the vertical-structure iteration (TempAverage,
 * DiskStructure, compareCells), the ray-traced SED (Observe SPECTRUM,
 * TraceMono), the ray-traced image (Observe IMAGE, MakeImage) and the dust
 * opacities the pass reads (ComputePart, the DHS particle from optical
 * constants).  The reference passes everything through module
 * globals (Modules.f:141-150); here the caller hands the same quantities over
 * once as plain host arrays (mcx_set_tables, mcx_set_model) and the library
 * keeps them resident in HBM.
 *
 * Conventions
 *  - every call returns 0 on success or a negative MCX_E* code; the text of
 *    the last error is available from mcx_last_error(); nothing aborts;
 *  - all arrays are host memory owned by the caller and read only during the
 *    call; the context owns its device buffers;
 *  - one host thread per context; calls on one context are not re-entrant.
 *
 * Index conventions (identical to the reference's arrays, 0-based):
 *   radial cells  i = 0..nr-1 (i = 0 is the void between R* and Rin),
 *   theta cells   j = 0..nt-1 (reference j+1), cell c = i*nt + j;
 *   R[0..nr] radial edges in AU, R[0] = Rstar/AU          (Init.f:2787-2852)
 *   Theta[0..nt] cos(theta) edges, Theta[0]=1, Theta[nt]=0 (Init.f:2876-2886)
 *   BB[T*nlam + l], T = 0..tmax                          (Init.f:2616-2622)
 *   per-species tables [s*nlam + l]; Kp[s*(tmax+1) + T]  (Init.f:3784-3790)
 *   Mueller F[((s*nlam + l)*6 + k)*180 + a], k = F11,F12,F22,F33,F34,F44
 *                                                        (Modules.f:53-57)
 *   SED bins [a*nlam + l], a = 0..30 (MCRadiation.f:21, spec(nlam,0:nangle))
 */
#ifndef MCX_H
#define MCX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCX_OK 0
#define MCX_EINVAL (-1)   /* bad argument or inconsistent sizes          */
#define MCX_EDEVICE (-2)  /* HIP runtime error                           */
#define MCX_ENOMEM (-3)   /* device allocation failed                    */
#define MCX_ESTATE (-4)   /* call out of order (e.g. no model set)       */
#define MCX_ETIME (-6)    /* the pass reached its time limit (mcx_run.max_ms)
                             with packets unfinished; tallies partial   */
#define MCX_EGEOM (-5)    /* packets lost to "nothing to hit"
                             (MCRadiation.f:1223-1232 stops there)       */

#define MCX_TMAX 10000    /* Modules.f:5 */
#define MCX_NANGLE 30     /* MCRadiation.f:20 */

typedef struct mcx_ctx mcx_ctx;

/* Wavelength, dust and star tables: the module globals lam, nu, BB, Grain(:)
 * and D%Fstar/D%Lstar read by MCRadiation (Modules.f:13, 59-83, 109-113). */
typedef struct {
    int32_t nlam, ng, tmax;          /* tmax must equal MCX_TMAX            */
    const double *lam, *nu;          /* [nlam] micron, Hz                   */
    const double *BB;                /* [(tmax+1)*nlam]                     */
    const double *Kabs, *Ksca;       /* [ng*nlam] cm^2/g                    */
    const double *F;                 /* [ng*nlam*6*180] (may be NULL when
                                        scat_how != 2)                      */
    const double *IF11, *IF12;       /* [ng*nlam]                           */
    const double *Kp;                /* [ng*(tmax+1)]                       */
    const double *Fstar;             /* [nlam]                              */
    double Lstar;                    /* trapezoid integral of Fstar         */
    const double *cos2phi, *sin2phi; /* [361] (Init.f:3777-3781)            */
    int32_t scat_how;                /* 0 NONE, 1 ISOTROPIC, 2 FULL         */
    /* modified random walk only (may be NULL when it is not used):       */
    const double *Kext;              /* [ng*nlam] as read (Init.f:4886)     */
    const double *gsym;              /* [ng*nlam] asymmetry g (Init.f:4037) */
} mcx_tables;

/* Grid and density: D%R, D%Theta and C(i,j)%dens/%V/%w (Modules.f:85-121). */
typedef struct {
    int32_t nr, nt;
    const double *R, *Theta;         /* [nr+1], [nt+1]                      */
    const double *dens, *V;          /* [nr*nt] g/cm^3, cm^3                */
    const double *w;                 /* [nr*nt*ng] species abundances       */
    double Rstar_cm, Tstar;          /* void-cell temperature, Main.f:546   */
    const double *thet;              /* [nt+1] theta of the cone edges, for
                                        MinDist (D%SinTheta, Init.f:2878);
                                        may be NULL without random walk     */
} mcx_grid;

/* One MC pass: MCRadiation(NphotTot, ...) with nruns = 1. */
typedef struct {
    int64_t nphot;                   /* NphotTot                            */
    uint64_t seed;                   /* Philox4x32-10 key; packet p uses the
                                        counter stream (block, p)           */
    int32_t maxinteract;             /* >0: roulette, MCRadiation.f:568-571 */
    int32_t nlanes;                  /* 1: a single lane, packets in order,
                                        each Bjorkman-Wood absorption takes
                                        its slice of the cell's running
                                        total, as the serial reference loop
                                        (MCRadiation.f:320, Radiation.f:173-
                                        258).  Otherwise the staged schedule
                                        below, with at most this many
                                        packets in flight (0: the device's
                                        resident lanes); its results do not
                                        depend on the concurrency            */
    int64_t stage0;                  /* staged schedule: packets run in
                                        stages [P_k, P_k+1), P_1 = stage0
                                        (0: 64), P_k+1 = stage_ratio P_k;
                                        packet p of a stage takes its slice
                                        in cell c at S_k(c) p / P_k plus its
                                        own earlier deposits in c, S_k(c) the
                                        Eabs of the packets before the stage
                                        (the serial loop's running total in
                                        expectation; see epoch_deposits).
                                        All stages run in one launch; a
                                        pass is bit-reproducible for a given
                                        seed                                */
    int32_t randomwalk;              /* keyword randomwalk (RNDW): modified
                                        random walk in thick cells          */
    double factRW;                   /* keyword factRW (Init.f:205: 10)     */
    int32_t coop_lanes;              /* inverse-CDF lookups: a wavefront
                                        serves them together when at most
                                        this many lanes need one; -1 picks
                                        the default, 0 never (results do not
                                        depend on it)                       */
    int32_t stage_ratio;             /* 0: 2; -1: the ratio grows as
                                        min(16, 1 + sqrt(P_k / stage0))     */
    int32_t precision;               /* MCX_FP64: the reference's real*8
                                        throughout.  MCX_FP32: directions,
                                        their rotations and the Stokes
                                        vector, and the optical-depth draws
                                        in single precision (hardware sin,
                                        cos, log, rsqrt); positions, the
                                        crossing quadratics, optical-depth
                                        sums, tallies and temperatures FP64 */
    double max_ms;                   /* > 0: time limit of the pass (the
                                        reference's maxruntime, MCRadiation.f:
                                        323-342, as a hard stop): past it
                                        every lane abandons its packet and
                                        the call returns MCX_ETIME          */
    int64_t epoch_deposits;          /* staged schedule, B: a packet of
                                        stage j files its d-th Eabs deposit
                                        under epoch j + m, m the smallest
                                        m >= 1 with d < B (2^m - 1), and S_k
                                        sums epochs 1..k; stage k starts once
                                        every earlier packet has finished or
                                        gone past the deposits S_k counts,
                                        so a long history holds back only
                                        the next stages' starts by B, 3B,
                                        7B, ... of its deposits.  0: B =
                                        16000; < 0: S_k holds every deposit
                                        of the packets before P_k (each
                                        stage waits for all of them)        */
} mcx_run;

#define MCX_FP64 0
#define MCX_FP32 1

/* Pass tallies and statistics (MCRadiation.f:656-715). */
typedef struct {
    double *EJv, *EJv2, *Eabs, *Tbw; /* [nr*nt] raw sums; Tbw = the cell's
                                        Bjorkman-Wood T at the end          */
    int32_t *Ni;                     /* [nr*nt] distinct packets per cell   */
    double *spec, *scatspec;         /* [(MCX_NANGLE+1)*nlam], energy sums
                                        already weighted as spec(l,a)       */
    int32_t *ispec, *iscatspec;      /* [(MCX_NANGLE+1)*nlam] counts        */
    int64_t nemit, ninteract, nabs, nscat, nescape, nstarhit, nsteps,
            nraar, nmaxinteract, ngeomfail;
    double tautot, Eescape;          /* sum of optical depth; escaped energy
                                        (equals Lstar: energy conservation) */
    int64_t nrw;                     /* modified-random-walk steps          */
    int64_t noverflow;               /* fixed-point adds that saturated (0
                                        in every valid pass)                */
    int32_t *thick;                  /* [nr*nt] C%thick of the last pass
                                        (InitRandomWalk, RandomWalk.f:152)  */
    int64_t maxsteps;                /* cell steps of the longest packet
                                        history                             */
    double t_done[5];                /* ms after the packet kernel started
                                        when 50, 90, 99, 99.9 and 100% of
                                        the packets had finished (last
                                        launch of the pass; -1: unknown)    */
    int64_t nvisit_overflow;         /* cell entries a packet's visit table
                                        could not hold (Ni may count such a
                                        packet twice; 0 in every test)      */
    int64_t ntimeout;                /* packets abandoned at the time limit */
    int64_t ntallyatomic;            /* EJv/EJv2 atomics issued (one per
                                        non-zero limb at each cell exit)    */
} mcx_tallies;

/* Temperature products of Main.f:498-549 and the void cell (Main.f:546). */
typedef struct {
    double *T, *dT, *dEJv;           /* [nr*nt]                             */
} mcx_temps;

int  mcx_create(mcx_ctx **ctx, int device);
void mcx_destroy(mcx_ctx *ctx);
const char *mcx_last_error(const mcx_ctx *ctx);

/* Replaces the reads of lam, nu, BB, Grain(:)%Kabs/Ksca/F/Kp, D%Fstar/Lstar
 * that MCRadiation and interact make (MCRadiation.f:210-229,
 * Radiation.f:41-355, 940-1001).  Builds the device-resident derived tables
 * (emission CDFs, scattering-angle CDFs). */
int mcx_set_tables(mcx_ctx *ctx, const mcx_tables *tab);

/* Replaces the reads of D%R, D%Theta, C(i,j)%dens/V/w (MCRadiation.f:901,
 * 1150).  Builds KabsTot/KscaTot (MCRadiation.f:210-229) and, per distinct
 * species mix, tables over integer T: the Planck means sum_s w_s Kp_s(T)
 * (computeE, Radiation.f:611-627; 80 kB per mix), the diffusion
 * coefficients (DiffCoeffCell, Dtable.f:97-182; 160 kB) and with several
 * species the reemission cumulative sum_s w_s CUM_s (16 MB per mix for 200
 * wavelengths).  Each table is built while it fits in 4 GiB, or in the bytes
 * the environment variable MCX_CLASS_TABLE_BYTES gives (read at each call;
 * many contexts on one GPU set it low).  Without a table the Planck means
 * and diffusion coefficients are evaluated at each read (the same values
 * bit for bit) and the reemission sums the species at every probe (the same
 * up to rounding).  Models of one batched launch need the same reemission
 * choice. */
int mcx_set_model(mcx_ctx *ctx, const mcx_grid *grid);

/* One transport pass: MCRadiation.f:269-668 on the device.  Tallies stay in
 * HBM until mcx_get_tallies.  Returns MCX_EGEOM if any packet hit the
 * reference's "nothing to hit" stop (the pass still completes). */
int mcx_transport(mcx_ctx *ctx, const mcx_run *run);

/* Many disk models in common launches (no counterpart in the reference,
 * which holds one model in its module globals, Modules.f:141-150, and runs
 * MCRadiation for it, MCRadiation.f:6).  Each of the nmodels contexts holds
 * one model (mcx_set_tables + mcx_set_model, grids of any size, same species
 * case); one pass of run->nphot packets runs for all of them in one launch
 * with blocks_per_model workgroups per model (0: the device's resident
 * workgroups shared out), on the first context's stream.  The long
 * histories that end each model's stages then run side by side instead of
 * one model after another.  Model m uses the Philox key run->seed + m *
 * seed_stride; its tallies are bit-identical to those of mcx_transport on
 * its own context with that key.  Read them with mcx_get_tallies /
 * mcx_temperature on the model's context.  status[m] (may be NULL)
 * receives model m's code (MCX_OK, MCX_EGEOM, MCX_ETIME); the call returns
 * the first failure.  One-lane runs (nlanes = 1) take one model.
 * use_graph != 0: the pass's launches are captured in a HIP graph on the
 * first call (for these contexts and this run configuration) and replayed
 * on later calls with the new keys, without host synchronisation between
 * stages (no time limit in this mode; mcx_last_timing reports the graph's
 * duration).  The structure of the models may change between calls through
 * mcx_set_model only if the grid sizes stay the same. */
int mcx_batch_transport(mcx_ctx *const *ctxs, int32_t nmodels, const mcx_run *run, int32_t blocks_per_model,
                        int64_t seed_stride, int32_t use_graph, int32_t *status);

/* mcx_batch_transport with each model's own share of the launch: model m
 * runs on model_blocks[m] >= 1 workgroups (a sweep gives the models with
 * more work more of the device).  Results do not depend on the shares.  A
 * graph (use_graph) is replayed while the total number of workgroups stays
 * the same; the shares themselves may change between calls. */
int mcx_batch_transport_blocks(mcx_ctx *const *ctxs, int32_t nmodels, const mcx_run *run,
                               const int32_t *model_blocks, int64_t seed_stride, int32_t use_graph,
                               int32_t *status);

/* Copy the last pass's tallies to the host (any pointer may be NULL). */
int mcx_get_tallies(mcx_ctx *ctx, mcx_tallies *out);

/* Lucy temperature inversion of the last pass on the device
 * (Main.f:498-549, determineT Radiation.f:629-702); copies T, dT, dEJv to
 * the host when the pointers are non-NULL. */
int mcx_temperature(mcx_ctx *ctx, mcx_temps *out);

/* Observation of the SED (Telescope.f:590-657 defaults in brackets): the
 * reference's Observe with type SPECTRUM, one inclination. */
typedef struct {
    double angle;                    /* inclination [rad]                   */
    int32_t nphi, nj0;               /* image azimuths [45], rays per radial
                                        cell [2]                            */
    int32_t nphot, nphotstar;        /* scattering packets per wavelength
                                        from the disk / the star [1000]     */
    int32_t nexits;                  /* escape directions per cell [5]      */
    int32_t usepol;                  /* [1]                                 */
    int32_t tracestar, traceemis, tracescat; /* [1, 1, 1]                   */
    double pfstop;                   /* survival exponent [0.25]            */
    uint64_t seed;                   /* Philox key of wavelength k: seed+k  */
} mcx_obs;

/* The ray-traced SED at nlam_obs wavelengths: TracePath + TraceFlux
 * (TraceMono.f:8-578, 2032-2293) on the device.  T: cell temperatures
 * [nr*nt] (NULL: those of the last mcx_temperature); theta_av: the theta
 * cell centres [nt] (Init.f:2791-2815) that place the image radii.  Writes
 * flux, scatflux and fluxQ [nlam_obs] (spectrum before the distance
 * scaling, Telescope.f:160).  Needs Kext (mcx_set_tables) and thet
 * (mcx_set_model). */
int mcx_sed(mcx_ctx *ctx, const mcx_obs *obs, const double *T, const double *theta_av, int32_t nlam_obs,
            const double *lam_obs, double *flux, double *scatflux, double *fluxQ);

/* Load a temperature state instead of computing one (the driver's
 * restart from a structure file): T, and optionally dT and Ni [nr*nt]
 * (NULL: zero), as mcx_temperature would leave them. */
int mcx_set_temperature(mcx_ctx *ctx, const double *T, const double *dT, const int32_t *Ni);

/* The diffusion approximation that follows the inversion in the driver
 * (Diffuse, Diffusion.f:9-329, Main.f:561-563): thick cells (C%thick as
 * InitRandomWalk sets it with factRW) seen by fewer than nphotmin packets or
 * with dT/T > dTmax get T from the steady T^4 diffusion equation, solved
 * densely (3 passes).  Works on the device T of the last mcx_temperature,
 * which it updates; copies T, the thick flags and EJv (reference C%EJv: the
 * solved cells' sum_s w_s Kp_s(T); other cells keep the values passed in,
 * which may be NULL) back when the pointers are non-NULL.  R_av [nr],
 * theta_av [nt]: cell centres.  ncells: number of cells that triggered it. */
int mcx_diffuse(mcx_ctx *ctx, int32_t nphotmin, double dTmax, double factRW, const double *R_av,
                const double *theta_av, double *T, int32_t *thick, double *EJv, int32_t *ncells);

/* TempAverage(fw) (DiskStructure.f:25-79): the running average of the
 * temperatures the vertical structure is solved for, kept on the device.
 * first != 0 stores the current temperatures (the reference's first call,
 * which allocates the average); otherwise Tav = (1-fw) Tav + fw max(T, 2.7).
 * T is that of the last mcx_temperature / mcx_diffuse / mcx_set_temperature.
 * The driver's weights (Main.f:683-690): 1 on the first iteration, f_weight
 * [0.5] after, 1/(niter-niter0) beyond niter0.  Tav [nr*nt] may be NULL. */
int mcx_temp_average(mcx_ctx *ctx, double fw, int32_t first, double *Tav);

/* Vertical-structure settings (DiskStructure.f:137-169; Init.f defaults in
 * brackets).  Every species has the scale-height type DISK. */
typedef struct {
    const double *R_av, *theta_av;   /* cell centres [nr] cm, [nt] rad      */
    double Mstar;                    /* D%Mstar in g                        */
    const double *shscale;           /* [nr] shscale(i) (Init.f:2754-2763)
                                        [NULL: 1]; < 0 makes a wedge        */
    const double *gshscale;          /* [ng*nr] Grain(ii)%shscale(i), used
                                        where settle (Init.f:3413-3425)
                                        [NULL: 1]                           */
    const int32_t *settle;           /* [ng] Grain%settle [NULL: none]      */
    const int32_t *gascoupled;       /* [ng] Grain%gascoupled [NULL: none]  */
    int32_t mpstr;                   /* keyword mpstr: midplane T for the
                                        whole column [0]                    */
    const double *FradZ;             /* [nr*nt] C%FradZ (radiation pressure,
                                        MCRadiation.f:835) [NULL: 0]        */
} mcx_struct;

/* The density state DiskStructure reads and rewrites: C%dens, C%dens0,
 * C%gasdens [nr*nt] and C%w, C%w0 [nr*nt*ng] (Modules.f:85-121).  At set-up
 * dens0 = gasdens = dens and w0 = w (Init.f:3899-3904). */
typedef struct {
    double *dens, *dens0, *gasdens;
    double *w, *w0;
} mcx_density;

/* DiskStructure (DiskStructure.f:86-453): every column (radius, gas or
 * species) in hydrostatic equilibrium at the temperatures Temp [nr*nt]
 * (NULL: the device average of mcx_temp_average), integrated from the
 * midplane up with the reference's adaptive Cash-Karp scheme and
 * renormalised to the column's mass; then the abundances, the minimum
 * densities (CheckMinimumDensity) and CheckCells.  d is read and
 * rewritten.  The next transport pass sees the new structure after
 * mcx_set_model with d->dens and d->w.  nsteps (may be NULL) receives the
 * number of Runge-Kutta steps. */
int mcx_disk_structure(mcx_ctx *ctx, const mcx_struct *st, const double *Temp, mcx_density *d, int64_t *nsteps);

/* One state of the cells for compareCells: C%T, C%dT, C%dens, C%ddens,
 * C%EJv (per volume), C%dEJv [nr*nt]. */
typedef struct {
    const double *T, *dT, *dens, *ddens, *EJv, *dEJv;
} mcx_cellstate;

/* compareCells (DiskStructure.f:553-584): the driver's convergence test
 * (Main.f:734, errT < epsiter [3]).  Cells with dT <= 1e-200 in either
 * state do not count; errRho divides by ddens, which the reference never
 * sets (zero: errRho is then not finite, and the driver does not print it). */
int mcx_compare_cells(mcx_ctx *ctx, const mcx_cellstate *prev, const mcx_cellstate *cur, double *errT,
                      double *errRho, double *errE);

/* Duration in milliseconds of the last mcx_transport's packet kernel and of
 * the whole pass, both measured with HIP events on the context's stream. */
int mcx_last_timing(const mcx_ctx *ctx, double *kernel_ms, double *pass_ms);

/* ---- images ------------------------------------------------------------ */

/* The pixel image of an IMAGE observation (Telescope.f:166-178 with
 * MakeImage, TraceMono.f:2759-3020, no PSF, noise or planets). */
typedef struct {
    int32_t npix;                    /* tel%npixel [500]                    */
    int32_t nint;                    /* tel%nint: samples per polar cell
                                        [1000]                              */
    double Rmax;                     /* half width of the field in AU
                                        (fov/2)                             */
    double pa_deg;                   /* D%PA, keyword posangle [0]          */
    uint64_t seed;                   /* Philox key of the samples (the
                                        reference draws from ran2)          */
} mcx_image_opts;

/* An IMAGE observation at wavelength lam (micron): the image rays of
 * TracePath laid out at lam (TraceMono.f:2098-2229), their I, Q, U, V
 * (TraceFlux, TraceMono.f:325-480, as for mcx_sed), then MakeImage's
 * spread onto npix x npix pixels: im, imQ, imU, imV [npix*npix], pixel
 * (ix, iy) at (ix-1)*npix + iy-1, before MakeImage's final scaling by
 * 1e23 zscale / distance^2.  imQ, imU, imV are needed with full scattering
 * (scat_how = 2) and may be NULL otherwise.  *nimg receives the number of
 * rings; when Rimg [ray_cap] or rays [4][ray_cap*nphi] are given, the
 * rings' radii (AU) and the rays' I, Q, U, V (ray = ring*nphi + azimuth)
 * are copied there (MCX_EINVAL if ray_cap < *nimg). */
int mcx_image(mcx_ctx *ctx, const mcx_obs *obs, const double *T, const double *theta_av, double lam,
              const mcx_image_opts *opts, double *im, double *imQ, double *imU, double *imV, int32_t *nimg,
              int32_t ray_cap, double *Rimg, double *rays);

/* Durations of the last mcx_image: the ray images (rays_ms, host clock
 * around the rays' part, uploads and kernels included) and the spread onto
 * pixels (spread_ms, HIP events around that kernel on the context's
 * stream).  No counterpart in the reference (it prints no timings). */
int mcx_last_image_timing(const mcx_ctx *ctx, double *rays_ms, double *spread_ms);

/* ---- dust opacities from optical constants ---------------------------- */

/* One dust species for ComputePart (ComputePart.f:1-552, the particle that
 * Init.f:3372 computes for a `computepartNN` keyword with standard = FILE):
 * grains of nm materials mixed in volume, sizes amin..amax sampled at ns
 * points with n(a) ~ a^-apow, each size a distribution of hollow spheres
 * (DHS) with vacuum fractions 0.01..fmax at nf Gauss-Legendre points
 * (fmax = 0: solid spheres).  Optical constants as the lnk files give them
 * (n, k >= 0), already on the wavelength grid (readrefindCP,
 * ComputePart.f:1102-1175) and blended by the caller where blend is set. */
typedef struct {
    int32_t nm, ns, nf, pad;
    double amin, amax;          /* micron                                  */
    double apow;                /* size-distribution power                 */
    double fmax;                /* largest DHS vacuum fraction, < 1        */
    double rscale, mscale;      /* p%rscale, p%mscale (ComputePart.f:144-155):
                                   1, 1 unless normalize_abun is set       */
    const float *e1, *e2;       /* [nm][nlam] n and k (REAL)               */
    const float *frac;          /* [nm] volume fractions, summing to 1     */
    const float *rho;           /* [nm] material densities, g cm^-3        */
} mcx_part_spec;

/* Results for one species: p%Kabs, p%Ksca, p%Kext [nlam] (cm^2 g^-1),
 * p%F [nlam][6][180] (F11, F12, F22, F33, F34, F44 at 0.5..179.5 deg) and
 * p%rho (the grains' mean density). */
typedef struct {
    double *Kabs, *Ksca, *Kext, *F, *rho;
} mcx_part_out;

/* Work of one mcx_compute_parts call. */
typedef struct {
    int64_t ndhs;               /* layered-sphere evaluations (DMiLay)     */
    int64_t nhom;               /* homogeneous-sphere evaluations (MeerhoffMie) */
    int64_t nerr;               /* terms whose layered sphere did not converge
                                   and fell back to a homogeneous sphere  */
    int64_t dhs_terms, hom_terms;  /* series terms summed                  */
    int64_t scratch_bytes;      /* device scratch of the largest batch     */
    double ms_total;            /* wall time of the call (HIP events)      */
} mcx_part_stats;

/* ComputePart for nspec species on one wavelength grid lam [nlam] (micron),
 * all on the device: the layered-sphere series (Toon & Ackerman, DMiLay
 * ComputePart.f:1256-1715) for every (species, wavelength, material, size,
 * vacuum fraction), homogeneous spheres (MeerhoffMie, ComputePart.f:1883)
 * where the reference uses them (x > 10^4, fmax = 0, or no convergence),
 * and the size/shape averages with the reference's single-precision sums.
 * stats may be NULL.  Inputs outside what the reference handles (nf > 1
 * with 0 < fmax <= 0.01; |m| so close to 1 or below that the homogeneous
 * series would start inside its own length) return MCX_EINVAL. */
int mcx_compute_parts(mcx_ctx *ctx, int32_t nlam, const double *lam, int32_t nspec, const mcx_part_spec *specs,
                      mcx_part_out *out, mcx_part_stats *stats);

/* ---- warm-start network ------------------------------------------------ */

/* No counterpart in the reference, which starts every structure iteration
 * from the fixed guess Tstar sqrt(R_av(1)/R_av(i))/10 (Init.f:3935) and reads
 * the first pass's temperatures in the first vertical-structure solve
 * (DiskStructure.f:19, 36-44, 71; MCRadiation.f:276-282 zeroes T per pass).
 * A network maps a cell's features to log T so that the first solve can
 * start near the converged field; the Monte Carlo pass judges it.
 *
 * Shape: MCX_MLP_NIN inputs -> MCX_MLP_DEPTH hidden layers of MCX_MLP_WIDTH
 * (SiLU) -> 1 output.  Parameters are one FP32 vector [MCX_MLP_NPARAMS]:
 * W1 [nin][width], b1 [width], then per hidden layer l = 2..depth
 * W_l [width in][width out], b_l [width], then w_out [width], b_out.
 * Layers 2..depth and the output run in BF16 (weights and activations) on
 * the matrix cores with FP32 accumulation; layer 1 runs in FP32.  Arrays are
 * host memory owned by the caller; the handle owns its device buffers. */
#define MCX_MLP_NIN 8
#define MCX_MLP_WIDTH 256
#define MCX_MLP_DEPTH 4
#define MCX_MLP_NPARAMS 199937

typedef struct mcx_mlp mcx_mlp;

/* Other shapes than the one above return MCX_EINVAL. */
int mcx_mlp_create(mcx_mlp **net, int device, int32_t nin, int32_t width, int32_t depth);
void mcx_mlp_destroy(mcx_mlp *net);
const char *mcx_mlp_last_error(const mcx_mlp *net);

/* Load / read the FP32 master parameters (loading resets the Adam state). */
int mcx_mlp_set_params(mcx_mlp *net, const float *params);
int mcx_mlp_get_params(mcx_mlp *net, float *params);

/* y [n] = network(x [n][nin]). */
int mcx_mlp_forward(mcx_mlp *net, int64_t n, const float *x, float *y);

/* The loss sum_r (y_r - t_r)^2 / n over the n rows of x [n][nin], t [n], and
 * its gradient grad [MCX_MLP_NPARAMS] (may be NULL) with respect to the
 * parameters: forward and backward on the device. */
int mcx_mlp_gradient(mcx_mlp *net, int64_t n, const float *x, const float *t, double *loss, float *grad);

/* The training set (copied to the device): x [n][nin], t [n]. */
int mcx_mlp_set_data(mcx_mlp *net, int64_t n, const float *x, const float *t);

typedef struct {
    int32_t batch;                   /* rows per step, drawn with replacement
                                        from the training set               */
    int32_t nsteps;
    float lr, beta1, beta2, eps;     /* Adam                                */
    uint64_t seed;                   /* key of the batch draws              */
} mcx_mlp_train_opts;

/* nsteps Adam steps on the mean squared error of minibatches of the
 * training set; loss [nsteps] (may be NULL) receives each step's loss.
 * Deterministic for a given seed and starting state. */
int mcx_mlp_train(mcx_mlp *net, const mcx_mlp_train_opts *opts, float *loss);

/* No counterpart in the reference: mcx_mlp_forward for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops).  The same computation, bit for bit, with x and y staged through a
 * fixed device array of the library (MCX_MLP_NOALLOC_MAXROWS rows, 2.25 MiB
 * per GPU, held until the process ends); no allocation of any kind.  Larger
 * n returns MCX_EINVAL.  Every call on one GPU shares that array: calls that
 * overlap on one GPU (from several threads or handles) race. */
#define MCX_MLP_NOALLOC_MAXROWS 65536
int mcx_mlp_forward_noalloc(mcx_mlp *net, int64_t n, const float *x, float *y);

/* Kernel time of the last mcx_mlp_forward (HIP events, summed over its
 * chunks) and its row count, and the mean time of a training step of the
 * last mcx_mlp_train. */
int mcx_mlp_last_timing(const mcx_mlp *net, double *forward_ms, int64_t *forward_rows, double *step_ms);

#ifdef __cplusplus
}
#endif
#endif
