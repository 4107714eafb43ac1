// mcx.hip -- MI355X (gfx950) Monte Carlo photon transport for MCMax-style
// axisymmetric disk models, behind the C-ABI of include/mcx.h.
//
// Design (one HIP stream per context, everything resident in HBM):
//  * one lane per photon packet; lanes of a wavefront pull packet numbers
//    from a device counter with one wave-aggregated atomic, so a lane whose
//    packet escapes starts the next packet at once (persistent kernel);
//  * each packet draws from its own Philox4x32-10 counter stream
//    (key = seed, counter = (block, packet)), so a packet's random numbers do
//    not depend on scheduling;
//  * the Lucy path-length tallies EJv, EJv2 and the absorbed energy Eabs are
//    integer fixed-point sums (per-cell units), accumulated with 64-bit
//    integer atomics: the sums do not depend on the order of additions;
//  * Bjorkman-Wood reemission (Radiation.f:173-258): a one-lane pass takes
//    each absorption's slice of the cell's cumulative absorbed energy from
//    the value the Eabs fetch-add returns, exactly as the serial loop; a
//    concurrent pass runs the packets in stages [P_k, P_k+1) and packet p
//    takes its slice at S_k(c) p/P_k + (its own earlier deposits in c),
//    S_k the cell's Eabs after the packets before the stage -- the serial
//    loop's cumulative in expectation, taken from the completed prefix --
//    so every history depends only on its own draws and the stage's
//    snapshot, and a pass is bit-reproducible whatever the scheduling;
//  * reemission wavelengths are drawn by binary search in per-species
//    cumulative tables int_0^lambda B_nu(T) kappa dnu (one row per integer
//    T), instead of rebuilding the O(nlam) difference spectrum per
//    absorption (Radiation.f:187-213, 966-989);
//  * Ni (distinct packets per cell, MCRadiation.f:943-946): each lane keeps
//    a hash table of the cells its current packet has entered (with the
//    packet's own absorbed energy there); the first entry adds 1 to Ni;
//  * the modified random walk (RandomWalk.f:47-147) reads its diffusion
//    coefficients from a table per distinct species mix and integer T, built
//    once per model on the device, instead of recomputing DiffCoeffCell
//    (Dtable.f:97-182) whenever a cell's temperature moves.
#include "mcx_internal.h"

// One model of a launch: its tables, grid, tallies, random-walk data, energy
// units and packet counter.  A launch runs one model (a plain pass) or many
// (mcx_batch_transport): blocks [m*bpm, (m+1)*bpm) serve model m.
struct BatchModel {
    DevTables tb;
    DevModel md;
    DevTally ty;
    DevRW rw;
    unsigned long long *next_pid;
    double invE;   // 2^EABS_BITS / E0
    double unitE;  // E0 / 2^EABS_BITS
    unsigned long long seed; // the model's Philox key
};

struct TransportArgs {
    const BatchModel *bm; // [models]
    int bpm;              // blocks per model
    const int *bmap;      // [blocks] the model of each workgroup (null: blockIdx / bpm)
    int nmodels;
    long long pid_end;
    long long pid_base;
    long long nphot;
    // the stages of a concurrent pass: stage k holds packets [P[k], P[k+1]),
    // P[nstage] = nphot; depB: the deposits a packet makes per epoch unit
    // (0: every deposit of a stage-j packet is in epoch j+1)
    int nstage;
    unsigned depB;
    const long long *P;  // [nstage + 1] in device memory
    int maxinteract;
    unsigned long long deadline; // wall_clock64() past which lanes abandon their packets (0: none)
    int serial;    // only lane 0 of the grid takes packets (in packet order)
    int coop;      // wave-cooperative lookups when at most this many lanes need one
    int rb_mask;   // draws per lane ring minus one (a power of two minus one)
};

enum { S_NEED = 0, S_EMIT, S_STEP, S_INTERACT, S_LOOKUP, S_AFTER, S_ESCAPE, S_RW, S_WAIT, S_DONE };

// x**0.25 of the T^4 interpolation as two correctly rounded square roots:
// within 2 ulp of pow(x, 0.25) and an order of magnitude shorter on the GPU.
__device__ __forceinline__ double qroot(double x) { return sqrt(sqrt(x)); }

// ---------------------------------------------------------------------------
// Temperature from absorbed energy per gram (Radiation.f:537-562, 629-702).
template <class K> __device__ double dev_determineT(const K &kpc, double E1, double Tcell) {
    int iTmin = 1, iTmax = TMAXD;
    int iT = (int)(Tcell / 1.0);
    if (iT < 1) iT = 1;
    if (iT > TMAXD - 1) iT = TMAXD - 1;
    double E = kpc[iT];
    double Emax = kpc[iTmax];
    double Emin = kpc[1];
    if (isnan(E1)) return 3.0;
    if (E1 > Emax) return (double)(TMAXD - 1);
    if (E1 < Emin) return 1.0;
    int iT0 = iT, niter = 0;
    while (abs(iTmax - iTmin) > 1 && niter <= TMAXD) {
        niter++;
        iT = (int)(pow(E1 / E, 0.25) * (double)iT);
        if (iT == iT0) {
            if (E1 < E) iT = iT0 - 1;
            if (E1 > E) iT = iT0 + 1;
        }
        if (iT <= iTmin) iT = iTmin + 1;
        if (iT >= iTmax) iT = iTmax - 1;
        E = kpc[iT];
        if (E >= E1) { iTmax = iT; Emax = E; }
        if (E <= E1) { iTmin = iT; Emin = E; }
        iT0 = iT;
    }
    if (iTmin == iTmax) return (double)iTmin;
    float a4 = fpow4(iTmin), b4 = fpow4(iTmax);
    float d4 = b4 - a4;
    return qroot((double)a4 + ((double)d4 * (E1 - Emin)) / (Emax - Emin));
}

// increaseT: walk up from int(Tstart); when the start is already above E1
// the reference's walk can only fail and fall back to determineT, so do that.
template <class K> __device__ double dev_increaseT(const K &kpc, double E1, double Tstart) {
    int j = (int)(Tstart / 1.0);
    if (j > TMAXD) j = TMAXD;
    if (j < 1) j = 1;
    double kp0 = kpc[j];
    if (kp0 <= E1) {
        for (int i = j; i <= TMAXD - 1; i++) {
            double kp1 = kpc[i + 1];
            if (kp0 <= E1 && kp1 >= E1) {
                float a4 = fpow4(i), b4 = fpow4(i + 1);
                float d4 = b4 - a4;
                return qroot((double)a4 + ((double)d4 * (E1 - kp0)) / (kp1 - kp0));
            }
            kp0 = kp1;
        }
    }
    return dev_determineT(kpc, E1, Tstart);
}

// increaseT with entries jw..jw+2 of the cell's table already in registers
// (w0, w1, w2): the same walk and result, without the loads it can skip.
template <class K>
__device__ __forceinline__ double win_at(const K &kpc, int jw, double w0, double w1, double w2, int k) {
    if (k == jw) return w0;
    if (k == jw + 1) return w1;
    if (k == jw + 2) return w2;
    return kpc[k];
}

template <class K>
__device__ double dev_increaseT_w(const K &kpc, int jw, double w0, double w1, double w2, double E1, double Tstart) {
    int j = (int)(Tstart / 1.0);
    if (j > TMAXD) j = TMAXD;
    if (j < 1) j = 1;
    double kp0 = win_at(kpc, jw, w0, w1, w2, j);
    if (kp0 <= E1) {
        for (int i = j; i <= TMAXD - 1; i++) {
            double kp1 = win_at(kpc, jw, w0, w1, w2, i + 1);
            if (kp0 <= E1 && kp1 >= E1) {
                float a4 = fpow4(i), b4 = fpow4(i + 1);
                float d4 = b4 - a4;
                return qroot((double)a4 + ((double)d4 * (E1 - kp0)) / (kp1 - kp0));
            }
            kp0 = kp1;
        }
    }
    return dev_determineT(kpc, E1, Tstart);
}

// T of energy per gram E1: the bracket kp(i) <= E1 <= kp(i+1) and its T^4
// interpolation, the value increaseT and determineT both reach whatever
// their start; the window kpc[jw..jw+2] in (w0, w1, w2) is tried first.
__device__ __forceinline__ double t_interp(int i, double kp0, double kp1, double E1) {
    float a4 = fpow4(i), b4 = fpow4(i + 1);
    float d4 = b4 - a4;
    return qroot((double)a4 + ((double)d4 * (E1 - kp0)) / (kp1 - kp0));
}

template <class K> __device__ double dev_T_win(const K &kpc, int jw, double w0, double w1, double w2, double E1) {
    if (w0 <= E1 && w1 >= E1) return t_interp(jw, w0, w1, E1);
    if (w1 <= E1 && w2 >= E1) return t_interp(jw + 1, w1, w2, E1);
    return dev_determineT(kpc, E1, (double)((E1 > w2) ? jw + 2 : jw));
}

// the window start for a temperature guess
__device__ __forceinline__ int win_start(double g) {
    int jw = (g >= 1.0) ? (int)(g / 1.0) : 1;
    if (jw > TMAXD - 2) jw = TMAXD - 2;
    return jw;
}

// First bin b in [0, n) with cum(b) >= Lr; n when none.
template <typename F>
__device__ __forceinline__ int lower_bound_bins(int n, double Lr, F cum) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (cum(mid) >= Lr) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// emit()'s wavelength rule (Radiation.f:966-989): cumulative trapezoid, the
// weight of the lower grid point rises linearly through the chosen bin.
template <typename F>
__device__ __forceinline__ void pick_lambda(int nbins, double Lr, F cum, int &l1, double &wl1) {
    int b = lower_bound_bins(nbins, Lr, cum);
    if (b >= nbins) {
        l1 = nbins - 1;
        wl1 = 0.0;
        return;
    }
    double Lt = cum(b);
    double Ltold = (b > 0) ? cum(b - 1) : 0.0;
    wl1 = (Lr - Ltold) / (Lt - Ltold);
    l1 = b;
}

// Packet state in registers.  The Stokes vector (Q, U, V) and the
// polarisation reference direction S, which only emission and scattering
// touch, live in LDS (field-major, one slot per thread; see P_Q .. P_SZ).
struct Packet {
    double x, y, z, vx, vy, vz, E, wl1;
    double tau0, tau, nEJv;
    int i, jj, l1, edgeNr, ntrace, nstep;
    unsigned ndep; // Eabs deposits so far (their epochs, concurrent passes)
    bool onEdge, scatt, pol;
};

// Fixed-point Eabs units of energy E (2^EABS_BITS per nominal packet energy).
__device__ __forceinline__ unsigned long long eabs_q(double E, double invE, unsigned long long *scnt) {
    double qd = rint(E * invE);
    if (!(qd < 0x1p62)) {
        atomicAdd(&scnt[C_NOVERFLOW], 1ull);
        qd = 0x1p62;
    }
    return (unsigned long long)qd;
}

// Fixed-point Eabs add; returns the cell's total before the add.
__device__ __forceinline__ unsigned long long add_eabs(const DevTally &ty, int c, double E, double invE,
                                                       unsigned long long &q, unsigned long long *scnt) {
    double qd = rint(E * invE);
    if (!(qd < 0x1p62)) {
        atomicAdd(&scnt[C_NOVERFLOW], 1ull);
        qd = 0x1p62;
    }
    q = (unsigned long long)qd;
    return atomicAdd(&ty.Eabs[c], q);
}

// ---- the gated stages of a concurrent pass -------------------------------
// A packet of stage j files its d-th Eabs deposit (0-based) under epoch j + m,
// m the smallest m >= 1 with d < B (2^m - 1) (B = 0: m = 1), at most cap.
// Stage k's snapshot S_k sums epochs 1..k: a packet's deposits reach the
// snapshots of later stages in steps of B, 2B, 4B, ... deposits, so a long
// history does not hold the next stages back, and every snapshot is still a
// fixed function of the packets' own histories.
__device__ __forceinline__ int dep_epoch(int j, unsigned d, unsigned B, int cap) {
    const int m = B ? 32 - __clz(d / B + 1u) : 1;
    const int e = j + m;
    return e < cap ? e : cap;
}

// loads and stores that other XCDs see within the launch (the snapshots)
__device__ __forceinline__ unsigned long long ld_coh(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float ld_coh(const float *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coh(unsigned long long *p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coh(float *p, float v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// the stage of packet pid in the stage table sP[0..nstage] (LDS)
__device__ __forceinline__ int stage_of(const long long *sP, int nstage, long long pid) {
    int lo = 0, hi = nstage - 1;  // sP[lo] <= pid < sP[hi + 1]
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pid >= sP[mid]) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// A gated pass's deposit of q Eabs units in cell c by a packet of stage pk
// that has made ndep deposits so far: filed under its epoch; when the packet
// passes B (2^m - 1) deposits it is ready for stage pk + m (its deposits up
// to there performed before the count, release order)
__device__ __forceinline__ void gated_deposit(const DevTally &ty, const TransportArgs &A, int ncell, int c,
                                              unsigned long long q, int pk, unsigned &ndep) {
    atomicAdd(&ty.Eabs[(size_t)dep_epoch(pk, ndep, A.depB, A.nstage) * ncell + c], q);
    ndep++;
    if (A.depB && ndep % A.depB == 0u) {
        const unsigned r = ndep / A.depB + 1u;
        if ((r & (r - 1u)) == 0u) {
            const int k = pk + (31 - __clz(r));
            if (k < A.nstage) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                atomicAdd(&ty.gate[G_PASS + k], 1ull);
            }
        }
    }
}

// hunt (Subroutines.f:2086-2130) in the decreasing survival table, from jlo = 1;
// returns the 1-based jlo.
__device__ __forceinline__ int hunt_desc(const double *__restrict__ xx, int n, double x) {
    int jlo = 1, jhi, inc = 1;
    const bool ascnd = (xx[n - 1] >= xx[0]);
    if ((x >= xx[jlo - 1]) == ascnd) {
        for (;;) {
            jhi = jlo + inc;
            if (jhi > n) {
                jhi = n + 1;
                break;
            }
            if ((x >= xx[jhi - 1]) == ascnd) {
                jlo = jhi;
                inc += inc;
            } else {
                break;
            }
        }
    } else {
        jhi = jlo;
        for (;;) {
            jlo = jhi - inc;
            if (jlo < 1) {
                jlo = 0;
                break;
            }
            if ((x < xx[jlo - 1]) == ascnd) {
                jhi = jlo;
                inc += inc;
            } else {
                break;
            }
        }
    }
    while (jhi - jlo != 1) {
        int jm = (jhi + jlo) / 2;
        if ((x >= xx[jm - 1]) == ascnd) jlo = jm;
        else jhi = jm;
    }
    if (x == xx[n - 1]) jlo = n - 1;
    if (x == xx[0]) jlo = 1;
    return jlo;
}

// Per-lane integer partial sums of EJv and the three EJv2 limbs for one cell.
// Integer addition is associative, so adding them to the cell when the lane
// moves on gives the same totals as adding every contribution at once.
struct TallyAcc {
    int c;
    unsigned long long e1, l0, l1, l2;
};

__device__ __forceinline__ void acc_flush(const DevTally &ty, TallyAcc &t, unsigned long long *scnt) {
    if (t.c >= 0) {
        if (t.e1) atomicAdd(&ty.EJv[t.c], t.e1);
        unsigned long long *b = ty.EJv2 + 3 * (size_t)t.c;
        if (t.l0) atomicAdd(b, t.l0);
        if (t.l1) atomicAdd(b + 1, t.l1);
        if (t.l2) atomicAdd(b + 2, t.l2);
        // the tally atomics issued (the pass's atomic rate)
        const unsigned n = (t.e1 != 0ull) + (t.l0 != 0ull) + (t.l1 != 0ull) + (t.l2 != 0ull);
        if (n) atomicAdd(&scnt[C_NTALLYATOM], (unsigned long long)n);
    }
    t.e1 = t.l0 = t.l1 = t.l2 = 0ull;
}

// non-negative integral double to u64; the common case below 2^32 takes the
// short native conversion
__device__ __forceinline__ unsigned long long d2u64(double q) {
    return (q < 0x1p32) ? (unsigned long long)(unsigned)q : (unsigned long long)q;
}

// one EJv contribution x of the current cell, in units 1/invU1
__device__ __forceinline__ void acc_ejv(TallyAcc &t, double x, double invU1, unsigned long long *scnt) {
    double q = rint(x * invU1);
    if (!(q < 0x1p62)) {
        if (!(q > 0.0)) return;
        atomicAdd(&scnt[C_NOVERFLOW], 1ull);
        q = 0x1p62;
    }
    if (q > 0.0) t.e1 += d2u64(q);
}

// one EJv2 contribution x2 in units 1/invU2, split into 32-bit-weighted limbs
__device__ __forceinline__ void acc_ejv2(TallyAcc &t, double x2, double invU2, unsigned long long *scnt) {
    double q = x2 * invU2;
    if (!(q > 0.0)) return;
    if (q < 0x1p31) { // one limb: the value the general split below gives
        t.l0 += (unsigned long long)(unsigned)rint(q);
        return;
    }
    double l2 = floor(q * 0x1p-64);
    if (l2 > 0x1p31) {
        atomicAdd(&scnt[C_NOVERFLOW], 1ull);
        l2 = 0x1p31;
    }
    q = q - l2 * 0x1p64;
    double l1 = floor(q * 0x1p-32);
    double l0 = rint(q - l1 * 0x1p32);
    if (l0 > 0.0) t.l0 += d2u64(l0);
    if (l1 > 0.0) t.l1 += d2u64(l1);
    if (l2 > 0.0) t.l2 += d2u64(l2);
}

// An inverse-CDF lookup: the first bin b in [0, n) whose cumulative value
// f(b) reaches the target t.  Kinds:
//  L_SAME  f = ((R0[n+b] - R0[b]) ws) 1          reemission, T0 and T1 in one
//                                                 integer bin (Radiation.f:187-213)
//  L_DIFF  f = ((((e1 R1[n+b] + (1-e1) R1[b]) - e0 R0[n+b]) - (1-e0) R0[b]) ws) 1
//                                                 reemission, four Planck rows
//  L_SCAT  f = e1 C11[b] + e0 C12[b] > t          scattering angle, E F11 + Q F12
//                                                 (Radiation.f:436-452)
// R0, R1 point at cumulative rows; for L_SCAT R1 is the chosen (species,
// lambda) block of SC (records of 8: C11, C12, then F11 ... F44) and
// e1 = E, e0 = Q.  With several species
// (ng > 1) the reemission spectrum is the cell's mixture, sum_s w_s times
// species s's difference spectrum (Radiation.f:199-214 with KabsTot): R0, R1
// are species 0's rows, species s's lie (TMAX+1) n further per species, and
// ws carries the bits of a pointer to the cell's abundances (no register of
// its own); ng is the kernel's species count.
enum { L_SAME = 0, L_DIFF = 1, L_SCAT = 2 };

struct Lookup {
    const double *R0, *R1;
    double e0, e1, ws, t;
    int n, kind;
};

__device__ __forceinline__ double lk_eval(const Lookup &q, int b, int ng) {
    if (q.kind == L_SCAT) return q.e1 * q.R1[8 * b] + q.e0 * q.R1[8 * b + 1];
    if (ng <= 1) {
        if (q.kind == L_SAME) return ((q.R0[q.n + b] - q.R0[b]) * q.ws) * 1.0;
        const double a2 = 1.0 - q.e1, a4 = 1.0 - q.e0;
        return ((((q.e1 * q.R1[q.n + b] + a2 * q.R1[b]) - q.e0 * q.R0[q.n + b]) - a4 * q.R0[b]) * q.ws) * 1.0;
    }
    const double a2 = 1.0 - q.e1, a4 = 1.0 - q.e0;
    const size_t stride = (size_t)(TMAXD + 1) * q.n;
    const double *W = (const double *)__double_as_longlong(q.ws);
    double acc = 0.0;
#pragma unroll 1
    for (int s = 0; s < ng; s++) {
        const double *R0 = q.R0 + s * stride, *R1 = q.R1 + s * stride;
        const double d = (q.kind == L_SAME) ? (R0[q.n + b] - R0[b])
                                            : (((q.e1 * R1[q.n + b] + a2 * R1[b]) - q.e0 * R0[q.n + b]) - a4 * R0[b]);
        acc = acc + (d * W[s]) * 1.0;
    }
    return acc;
}

__device__ __forceinline__ bool lk_hit(const Lookup &q, double v) { return (q.kind == L_SCAT) ? (v > q.t) : (v >= q.t); }

// one lane alone: binary search (the CDFs are non-decreasing)
__device__ __forceinline__ void lk_single(const Lookup &q, int ng, int &b, double &fb, double &fbm1) {
    int lo = 0, hi = q.n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (lk_hit(q, lk_eval(q, mid, ng))) hi = mid;
        else lo = mid + 1;
    }
    b = lo;
    fb = (lo < q.n) ? lk_eval(q, lo, ng) : 0.0;
    fbm1 = (lo > 0 && lo < q.n) ? lk_eval(q, lo - 1, ng) : 0.0;
}

// lane src's value, made wave-uniform (scalar registers): every lane asks for
// the same lane's value
__device__ __forceinline__ int ushfl(int v, int src) { return __builtin_amdgcn_readfirstlane(__shfl(v, src)); }
__device__ __forceinline__ double ushfl(double v, int src) {
    const long long b = __double_as_longlong(v);
    const int lo = ushfl((int)(unsigned)b, src), hi = ushfl((int)(unsigned)(b >> 32), src);
    return __longlong_as_double(((long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ const double *shfl_ptr(const double *p, int src) {
    unsigned long long v = (unsigned long long)p;
    unsigned lo = (unsigned)ushfl((int)(unsigned)v, src), hi = (unsigned)ushfl((int)(unsigned)(v >> 32), src);
    return (const double *)(((unsigned long long)hi << 32) | lo);
}

// the whole wavefront serves lane src's lookup: every lane evaluates 4 bins
// per round (all loads in flight together), a ballot finds the first hit.
// Must be called by all 64 lanes.
__device__ void lk_wave(const Lookup &mine, int src, int lane, int ng, int &b, double &fb, double &fbm1) {
    Lookup q;
    q.R0 = shfl_ptr(mine.R0, src);
    q.R1 = shfl_ptr(mine.R1, src);
    q.e0 = ushfl(mine.e0, src);
    q.e1 = ushfl(mine.e1, src);
    q.ws = ushfl(mine.ws, src);
    q.t = ushfl(mine.t, src);
    q.n = ushfl(mine.n, src);
    q.kind = ushfl(mine.kind, src);
    b = q.n;
    fb = 0.0;
    fbm1 = 0.0;
    double prev_last = 0.0;
    for (int base = 0; base < q.n; base += 256) {
        double v[4];
        bool h[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int bb = base + 64 * k + lane;
            v[k] = (bb < q.n) ? lk_eval(q, bb, ng) : 0.0;
            h[k] = (bb < q.n) && lk_hit(q, v[k]);
        }
        bool done = false;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (done) break;
            const unsigned long long m = __ballot(h[k]);
            if (m) {
                const int bl = __ffsll((long long)m) - 1;
                b = base + 64 * k + bl;
                fb = __shfl(v[k], bl);
                double below = __shfl(v[k], (bl + 63) & 63);
                double prevk = (k > 0) ? __shfl(v[(k + 3) & 3], 63) : prev_last;
                fbm1 = (b == 0) ? 0.0 : ((bl > 0) ? below : prevk);
                done = true;
            }
        }
        if (done) break;
        prev_last = __shfl(v[3], 63);
    }
}

// A packet's Philox stream read through a ring of draws in LDS.  Draw d of
// packet p is word d&1 of block d>>1 of the stream (p's counter), so the
// draws are those of PhiloxStream; the blocks are computed ahead by the
// whole wavefront at one point of the loop (rng_service), and here only
// when the ring runs dry.
__device__ __noinline__ double2 philox_pair(uint32_t k0, uint32_t k1, unsigned long long blk, unsigned long long pid) {
    double2 r;
    philox_block(k0, k1, blk, pid, r.x, r.y);
    return r;
}

struct RingRng {
    double *ring;
    unsigned mask;
    uint32_t k0, k1;
    unsigned long long pid, head, fill; // next draw; draws in the ring up to fill
    __device__ __forceinline__ double next() {
        if (head >= fill) {
            const unsigned long long b = head >> 1;
            const double2 r = philox_pair(k0, k1, b, pid);
            ring[(unsigned)(2 * b) & mask] = r.x;
            ring[(unsigned)(2 * b + 1) & mask] = r.y;
            fill = 2 * b + 2;
        }
        const double v = ring[(unsigned)head & mask];
        head++;
        return v;
    }
};

#define NPL 23          // per-thread LDS doubles, all used (odd stride spreads the banks)
#define COOP_LANES 6 // default: a wave serves lookups together when at most this many lanes need one

// MODE 0: one dust species; 1: several, with per-class reemission tables;
// 2: several, their tables summed at every reemission probe
// the direction draw and the optical-depth draw of the kernel's precision
#define RDIR(gen, X, Y, Z)                                  \
    do {                                                    \
        if constexpr (F32) random_direction_f(gen, X, Y, Z); \
        else random_direction(gen, X, Y, Z);                \
    } while (0)
#define TAUDRAW(u) (F32 ? tau_draw_f(u) : -log(u))

// The workgroup's share of model mi's pass: its lanes take the model's
// packets until none is left, then the workgroup's tallies and counters are
// flushed.  The first workgroup to start on the model starts its timeline.
// F32: the direction arithmetic in single precision (precision mode MCX_FP32)
template <int MODE, bool F32>
__device__ __forceinline__ void transport_model(const TransportArgs &A, const int mi) {
    extern __shared__ double lds[];
    const BatchModel &Bm = A.bm[mi];
    const DevTables &tb = Bm.tb;
    const DevModel &md = Bm.md;
    const DevTally &ty = Bm.ty;
    const int nr = md.nr, nt = md.nt, nlam = tb.nlam, ng = tb.ng;
    constexpr bool ONE = MODE == 0, MIX = MODE == 1;
    // species summed per reemission probe (class tables: one row, as for one
    // species)
    const int lng = (MODE == 2) ? ng : 1;
    const int nbins = nlam - 1;
    double *sR2 = lds;
    double *sT2 = sR2 + (nr + 1);
    double *sTh = sT2 + (nt + 1);
    // prefix sums over phi bins 1..k of cos(2 phi), sin(2 phi), k = 0..180
    double *sPC2 = sTh + (nt + 1);
    double *sPS2 = sPC2 + 181;
    double *sCS = sPS2 + 181; // stellar emission cumulative, nbins entries
    // per-thread state kept out of registers (Stokes vector, S direction,
    // interaction cache): NPL doubles per thread, thread-major, so that every
    // field is a constant offset from the thread's base
    double *pl = sCS + nbins + (size_t)threadIdx.x * NPL;
    // rings of Philox draws, RB per thread
    const unsigned RB = (unsigned)A.rb_mask + 1u;
    double *sRing = sCS + nbins + (size_t)blockDim.x * NPL;
    // random walk: radial edges (AU), sin of the cone edges
    double *sRau = sRing + (size_t)blockDim.x * RB;
    double *sSinT = sRau + (nr + 1);
#define P_Q pl[0]
#define P_U pl[1]
#define P_V pl[2]
#define P_SX pl[3]
#define P_SY pl[4]
#define P_SZ pl[5]
#define P_TRW pl[6]  // cell T the next random-walk check reads
#define P_KDX pl[7]  // {KDext, KDabs} of the cell at kd_iT
#define P_KDY pl[8]
#define P_MASS pl[9] // mass of cell ic
#define P_IC pl[10]  // ic: the cell these entries belong to (-1: none)
#define P_TC pl[11]  // 2*class + C%thick of cell ic
#define P_KDIT pl[12] // kd_iT
#define P_FS pl[13]   // p / P_k of the packet (0 in the first stage)
#define P_BASE pl[14] // S_k(c) p / P_k of the visited cell, in Eabs units
#define P_OWN pl[15]  // bits of the packet's own Eabs units in the visited cell
#define P_VSLOT pl[16] // its slot in the lane's visit table (-1: none; +2^VIS_LOG2: own changed)
#define P_TG pl[17]   // a guess of the visited cell's T as the packet sees it
#define P_JW pl[18]   // the visited cell's table window kpc[jw..jw+2] (-1: none) ...
#define P_W0 pl[19]   // ... and its three values
#define P_W1 pl[20]
#define P_W2 pl[21]
#define P_PK pl[22]   // the stage of the lane's packet
    for (int k = threadIdx.x; k <= nr; k += blockDim.x) sR2[k] = md.R2[k];
    for (int k = threadIdx.x; k <= nt; k += blockDim.x) {
        sT2[k] = md.T2[k];
        sTh[k] = md.Theta[k];
    }
    for (int k = threadIdx.x; k <= 180; k += blockDim.x) {
        sPC2[k] = tb.pc2[k];
        sPS2[k] = tb.ps2[k];
    }
    for (int k = threadIdx.x; k < nbins; k += blockDim.x) sCS[k] = tb.cumStar[k];
    if (Bm.rw.on) {
        for (int k = threadIdx.x; k <= nr; k += blockDim.x) sRau[k] = Bm.rw.R[k];
        for (int k = threadIdx.x; k <= nt; k += blockDim.x) sSinT[k] = Bm.rw.sinT[k];
    }
    __syncthreads();

    const int lane = __lane_id();
    // event counters live in LDS (one set per block) to keep them out of VGPRs
    __shared__ unsigned long long scnt[C_NCOUNT];
    __shared__ double stau;
    for (int k = threadIdx.x; k < C_NCOUNT; k += blockDim.x) scnt[k] = 0ull;
    if (threadIdx.x == 0) stau = 0.0;
    if (threadIdx.x == 0) atomicCAS(&ty.counters[G_TLINE], 0ull, wall_clock64());
    __syncthreads();
#define CNT(k) atomicAdd(&scnt[k], 1ull)
    double tautot = 0.0;

    Packet p;
    memset(&p, 0, sizeof(p));
    RingRng g;
    g.ring = sRing + (size_t)threadIdx.x * RB;
    g.mask = RB - 1u;
    g.k0 = (uint32_t)Bm.seed;
    g.k1 = (uint32_t)(Bm.seed >> 32);
    g.pid = 0;
    g.head = g.fill = 0;
    // requests of the wave's lanes whose ring runs low (rng service)
    __shared__ unsigned long long rq_pid[256], rq_blk[256];
    __shared__ int rq_own[256];
    long long pid = -1;
    int st = S_NEED;
    // the highest stage whose snapshot each wave has seen complete
    __shared__ int s_gopen[16], s_nap[16], s_polls[16];
    if ((threadIdx.x & 63) == 0) {
        s_gopen[threadIdx.x >> 6] = 0;
        s_nap[threadIdx.x >> 6] = 0;
        s_polls[threadIdx.x >> 6] = 0;
    }
    // the stage table
    __shared__ long long sP[KSTAGE_MAX + 1];
    for (int q = threadIdx.x; q <= A.nstage; q += blockDim.x) sP[q] = A.P[q];
    __syncthreads();
    P_PK = 0.0;
    if (A.serial && (blockIdx.x * blockDim.x + threadIdx.x) != 0) st = S_DONE;
    const long long raar_limit = (long long)nr * (long long)(nt + 1) * 2;
    const double AU = MCX_AU;

    // cell cache: rho and the fixed-point units of cell cc, the opacity pair
    // of (kc, kl); tallies of cell cc accumulate in registers
    int cc = -1, kc = -1, kl = -1;
    double c_dens = 0.0, c_u1 = 0.0, c_u2 = 0.0;
    double c_Tl = 0.0; // BW T after this lane's last absorption in cell cc (0: none yet)
    // interaction state of cell ic (loaded at the first interaction in a
    // cell, alongside that interaction's own round trips): 2*class + C%thick,
    // the cell's mass, the cell T the next random-walk check reads (known to
    // the lane: its own update, or a load issued at the interaction) and
    // {KDext, KDabs} at kd_iT
    P_IC = -1.0;
    P_TC = 0.0;
    P_KDIT = -1.0;
    P_TRW = 0.0;
    P_MASS = 0.0;
    P_FS = 0.0;
    P_BASE = 0.0;
    P_OWN = 0.0;
    P_VSLOT = -1.0;
    P_TG = 0.0;
    P_JW = -1.0;
    // this lane's visit table
    uint4 *const vt = ty.vtab + ((size_t)(blockIdx.x * blockDim.x + threadIdx.x) << VIS_LOG2);
    double2 k1 = make_double2(0.0, 0.0), k2 = make_double2(0.0, 0.0);
    TallyAcc acc;
    acc.c = -1;
    acc.e1 = acc.l0 = acc.l1 = acc.l2 = 0ull;

    for (;;) {
        // ---- wave-aggregated packet acquisition -------------------------
        {
            bool need = (st == S_NEED);
            // packets that ended since the last acquisition: the pass's
            // timeline (when 50/90/99/99.9/100% had finished) and the
            // longest history
            const bool fin = need && pid >= 0;
            unsigned long long fmask = __ballot(fin);
            if (fin) {
                atomicMax(&ty.counters[G_MAXSTEPS], (unsigned long long)p.nstep);
                atomicAdd(&scnt[C_NSTEPS], (unsigned long long)p.nstep);
            }
            if (!A.serial) {
                // a finished packet is ready for every stage from the epoch
                // its next deposit would have had (release: its deposits first)
                const int kf = dep_epoch((int)P_PK, p.ndep, A.depB, A.nstage);
                const bool cred = fin && kf < A.nstage;
                if (__ballot(cred)) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                if (cred) atomicAdd(&ty.gate[G_FIN + kf], 1ull);
            }
            if (fmask && lane == __ffsll((long long)fmask) - 1) {
                const unsigned long long nf = (unsigned long long)__popcll(fmask);
                const unsigned long long d0 = atomicAdd(&ty.counters[G_DONE], nf);
                const unsigned long long n = (unsigned long long)(A.pid_end - A.pid_base);
                const unsigned long long num[5] = {1, 9, 99, 999, 1}, den[5] = {2, 10, 100, 1000, 1};
                for (int k = 0; k < 5; k++) {
                    const unsigned long long mk = (n * num[k] + den[k] - 1) / den[k];
                    if (d0 < mk && d0 + nf >= mk) ty.counters[G_TLINE + 1 + k] = wall_clock64();
                }
            }
            unsigned long long mask = __ballot(need);
            if (mask) {
                int leader = __ffsll((long long)mask) - 1;
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(Bm.next_pid, (unsigned long long)__popcll(mask));
                unsigned lo = __shfl((unsigned)base, leader);
                unsigned hi = __shfl((unsigned)(base >> 32), leader);
                base = ((unsigned long long)hi << 32) | lo;
                if (need) {
                    unsigned long long below = mask & ((1ull << lane) - 1ull);
                    long long mine = (long long)(base + (unsigned long long)__popcll(below));
                    if (mine >= A.pid_end) {
                        st = S_DONE;
                    } else {
                        pid = mine;
                        g.pid = (unsigned long long)pid;
                        g.head = g.fill = 0;
                        p.nstep = 0;
                        // a new packet: its own visit table entries, its own
                        // view of the cells (P_k = 0 in the first stage)
                        cc = -1;
                        P_IC = -1.0;
                        P_VSLOT = -1.0;
                        p.ndep = 0u;
                        const int pk = A.serial ? 0 : stage_of(sP, A.nstage, mine);
                        P_PK = (double)pk;
                        P_FS = (pk > 0) ? (double)pid / (double)sP[pk] : 0.0;
                        // a packet of a stage whose snapshot is not complete waits
                        st = (pk <= s_gopen[threadIdx.x >> 6]) ? S_EMIT : S_WAIT;
                    }
                }
            }
        }
        // ---- gated stages: lanes whose packet waits for its stage's
        //      snapshot; their wave checks the next stage's gate (every packet
        //      before the stage finished, or past the deposits the stage
        //      counts) and, once open, builds a chunk of 64 cells of its
        //      snapshot; the wave that completes the last chunk opens it
        if (!A.serial) {
            const unsigned long long wmask = __ballot(st == S_WAIT);
            if (wmask) {
                const int ld = __ffsll((long long)wmask) - 1;
                unsigned op = 0;
                if (lane == ld) op = (unsigned)ld_coh(&ty.gate[G_OPEN]);
                const int open = (int)__shfl(op, ld);
                int g_open = s_gopen[threadIdx.x >> 6];
                if (open > g_open) {
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                    g_open = open;
                    if (lane == 0) s_gopen[threadIdx.x >> 6] = open;
                }
                if (st == S_WAIT && (int)P_PK <= g_open) st = S_EMIT;
                const int gk = g_open + 1;
                // the gate's counters are summed on every fourth poll of a wave
                int polls = 0;
                if (lane == 0) polls = ++s_polls[threadIdx.x >> 6];
                polls = __shfl(polls, 0);
                if (__ballot(st == S_WAIT) && gk < A.nstage && (polls & 3) == 1) {
                    unsigned long long v = 0ull;
                    if (lane >= 1 && lane <= gk) v = ld_coh(&ty.gate[G_FIN + lane]);
                    if (lane == 0) v = ld_coh(&ty.gate[G_PASS + gk]);
                    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
                    const int nchunk = (md.ncell + 63) >> 6;
                    if (v == (unsigned long long)sP[gk]) {
                        unsigned ch = 0;
                        if (lane == 0) ch = (unsigned)atomicAdd(&ty.gate[G_CHUNK + gk], 1ull);
                        ch = __shfl(ch, 0);
                        if (ch < (unsigned)nchunk) {
                            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                            const int c = (int)ch * 64 + lane;
                            if (c < md.ncell) {
                                const size_t nc = (size_t)md.ncell;
                                const unsigned long long S =
                                    ((gk > 1) ? ld_coh(&ty.Sk[(size_t)(gk - 1) * nc + c]) : 0ull) +
                                    ld_coh(&ty.Eabs[(size_t)gk * nc + c]);
                                st_coh(&ty.Sk[(size_t)gk * nc + c], S);
                                const float Tg =
                                    S ? (float)dev_determineT(kp_row(md, md.cls[c]), ((double)S * Bm.unitE) / md.mass[c],
                                                              1.0)
                                      : 0.f;
                                st_coh(&ty.Tk[(size_t)gk * nc + c], Tg);
                            }
                            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                            if (lane == 0) {
                                const unsigned long long d = atomicAdd(&ty.gate[G_CHDONE + gk], 1ull);
                                if (d + 1ull == (unsigned long long)nchunk) {
                                    ty.gate[G_TOPEN + gk] = wall_clock64();
                                    __hip_atomic_store(&ty.gate[G_OPEN], (unsigned long long)gk, __ATOMIC_RELEASE,
                                                       __HIP_MEMORY_SCOPE_AGENT);
                                }
                            }
                        }
                    }
                }
                // a wave with nothing to do but wait polls less and less
                // often (up to ~8k cycles between polls)
                if (__all(st == S_WAIT || st == S_DONE)) {
                    const int nap = s_nap[threadIdx.x >> 6];
                    for (int z = 0; z < nap; z++) __builtin_amdgcn_s_sleep(127);
                    if (lane == 0) s_nap[threadIdx.x >> 6] = nap < 2 ? nap + 1 : 2;
                } else if (lane == 0) {
                    s_nap[threadIdx.x >> 6] = 0;
                }
            }
        }
        if (__all(st == S_DONE)) break;
        // the pass's time limit: past it every lane gives up its packet
        // (counted) and takes no other, so the launch ends
        if (A.deadline && wall_clock64() > A.deadline) {
            if (st != S_DONE) {
                CNT(C_NTIMEOUT);
                // the first abandoned packet's state, for the error message
                if (atomicCAS(&ty.counters[G_TOREP], 0ull, 1ull) == 0ull) {
                    double *d = ty.torep;
                    d[0] = (double)pid; d[1] = (double)st; d[2] = (double)p.nstep; d[3] = (double)p.i;
                    d[4] = (double)p.jj; d[5] = p.x; d[6] = p.y; d[7] = p.z; d[8] = p.vx; d[9] = p.vy;
                    d[10] = p.vz; d[11] = (double)p.onEdge; d[12] = (double)p.edgeNr; d[13] = p.tau0;
                    d[14] = p.tau; d[15] = (double)p.ntrace;
                }
                st = S_DONE;
            }
            acc_flush(ty, acc, scnt);
            break;
        }

        // ---- rng service: the wavefront computes, 64 blocks per pass, the
        //      next RB/4 Philox blocks of every lane whose ring is at most
        //      half full
        {
            const bool low = (st != S_DONE) && (g.fill - g.head <= (unsigned long long)(RB >> 1));
            const unsigned long long m = __ballot(low);
            if (m) {
                const int wbase = threadIdx.x & ~63;
                if (low) {
                    const int rank = __popcll(m & ((1ull << lane) - 1ull));
                    rq_pid[wbase + rank] = g.pid;
                    rq_blk[wbase + rank] = g.fill >> 1;
                    rq_own[wbase + rank] = threadIdx.x;
                }
                __builtin_amdgcn_wave_barrier();
                const int per = (int)(RB >> 2); // blocks per request
                const int items = __popcll(m) * per;
                for (int w = lane; w < items; w += 64) {
                    const int r = w / per, k = w - r * per;
                    const unsigned long long b = rq_blk[wbase + r] + (unsigned long long)k;
                    double d0, d1;
                    philox_block(g.k0, g.k1, b, rq_pid[wbase + r], d0, d1);
                    double *ring = sRing + (size_t)rq_own[wbase + r] * RB;
                    ring[(unsigned)(2 * b) & (RB - 1u)] = d0;
                    ring[(unsigned)(2 * b + 1) & (RB - 1u)] = d1;
                }
                __builtin_amdgcn_wave_barrier();
                if (low) g.fill += (unsigned long long)(RB >> 1);
            }
        }

        // ---- stellar emission (MCRadiation.f:347-383, 500-513, 554) ------
        if (st == S_EMIT) {
            p.nEJv = 0.0;
            (void)g.next(); // where_emit
            RDIR(g, p.x, p.y, p.z);
            p.x = md.R0 * p.x;
            p.y = md.R0 * p.y;
            p.z = md.R0 * p.z;
            CNT(C_NEMIT);
            RDIR(g, p.vx, p.vy, p.vz); // inside emit()
            double Lr = g.next() * tb.Lstar;
            pick_lambda(nbins, Lr, [&](int b) { return sCS[b]; }, p.l1, p.wl1);
            p.E = tb.Lstar / ((double)(float)A.nphot * ((1.0 - 0.0) - 0.0));
            RDIR(g, p.vx, p.vy, p.vz);
            if (p.x * p.vx + p.y * p.vy + p.z * p.vz < 0.0) {
                p.vx = -p.vx;
                p.vy = -p.vy;
                p.vz = -p.vz;
            }
            p.edgeNr = 1;
            p.onEdge = true;
            // the theta cell with Theta(j+1) < ct < Theta(j); none when ct sits
            // on an edge, as in the reference's scan (MCRadiation.f:376-381)
            double ct = fabs(p.z) / md.R0;
            {
                int lo = 0, hi = nt + 1; // first edge below ct (edges decrease)
                while (lo < hi) {
                    int mid = (lo + hi) >> 1;
                    if (sTh[mid] < ct) hi = mid;
                    else lo = mid + 1;
                }
                if (lo >= 1 && lo <= nt && sTh[lo - 1] > ct) {
                    p.i = 0;
                    p.jj = lo - 1;
                }
            }
            p.pol = false;
            p.scatt = false;
            P_Q = 0.0;
            P_U = 0.0;
            P_V = 0.0;
            {
                double sx = -p.vy, sy = p.vx, sz = 0.0;
                double r = sqrt(sx * sx + sy * sy + sz * sz);
                P_SX = sx / r;
                P_SY = sy / r;
                P_SZ = sz / r;
            }
            p.E = p.E * 1.0;
            p.tau0 = TAUDRAW(g.next());
            p.tau = 0.0;
            p.ntrace = 0;
            st = S_STEP;
        }

        // ---- one cell step of trace2d (MCRadiation.f:941-1138) ----------
        if (st == S_STEP) {
            const int c = p.i * nt + p.jj;
            // issue the loads of a new cell first; atomics only after them, so
            // that the loads do not wait behind the atomics' completion
            const bool newc = (c != cc), newk = (c != kc || p.l1 != kl);
            double nd = 0.0, n1 = 0.0, n2 = 0.0;
            double2 nk1 = k1, nk2 = k2;
            uint4 ve = make_uint4(0u, 0u, 0u, 0u);
            unsigned vh = 0u;
            unsigned long long nS = 0ull;
            float nTg = 0.f;
            if (newc) {
                nd = md.dens[c];
                n1 = ty.invU1[c];
                n2 = ty.invU2[c];
                // c's slot of the visit table, and the stage's snapshot
                vh = ((unsigned)c * 0x9E3779B1u) >> (32 - VIS_LOG2);
                ve = vt[vh];
                // plain loads: this wave passed an agent-scope acquire after
                // it saw the stage open (the builders' release), so stage
                // pk's snapshot is visible and may be cached in L2
                const int pk = (int)P_PK;
                if (!A.serial && pk > 0) {
                    nS = ty.Sk[(size_t)pk * md.ncell + c];
                    nTg = ty.Tk[(size_t)pk * md.ncell + c];
                }
            }
            if (newk) {
                nk1 = md.KT[(size_t)c * nlam + p.l1];
                nk2 = md.KT[(size_t)c * nlam + p.l1 + 1];
            }
            double v;
            int inext, jnext;
            bool ok = exit_face(sR2, sT2, nt, p.x, p.y, p.z, p.vx, p.vy, p.vz, p.i, p.jj, p.onEdge, p.edgeNr, v,
                                inext, jnext);
            if (newc) {
                // the previous visit's own energy back to its slot
                {
                    const int vs = (int)P_VSLOT;
                    if (vs >= (1 << VIS_LOG2)) {
                        const unsigned long long o = (unsigned long long)__double_as_longlong(P_OWN);
                        vt[vs - (1 << VIS_LOG2)] = make_uint4((unsigned)(pid + 1), (unsigned)cc, (unsigned)o,
                                                              (unsigned)(o >> 32));
                    }
                }
                acc_flush(ty, acc, scnt);
                cc = c;
                acc.c = c;
                c_Tl = 0.0;
                c_dens = nd;
                c_u1 = n1;
                c_u2 = n2;
                // this packet's entry for c: found, or inserted (its first
                // entry into c: Ni += 1, MCRadiation.f:943-946)
                const unsigned tag = (unsigned)(pid + 1);
                int slot = -1;
                unsigned long long own = 0ull;
                for (int k = 0; k < VIS_PROBE; k++) {
                    if (ve.x != tag) {
                        vt[vh] = make_uint4(tag, (unsigned)c, 0u, 0u);
                        atomicAdd(&ty.Ni[c], 1);
                        slot = (int)vh;
                        break;
                    }
                    if (ve.y == (unsigned)c) {
                        own = ((unsigned long long)ve.w << 32) | ve.z;
                        slot = (int)vh;
                        break;
                    }
                    vh = (vh + 1u) & ((1u << VIS_LOG2) - 1u);
                    ve = vt[vh];
                }
                if (slot < 0) {
                    // table full along the probe: counted as a first entry
                    CNT(C_NVISOVF);
                    atomicAdd(&ty.Ni[c], 1);
                }
                P_VSLOT = (double)slot;
                P_JW = -1.0;
                P_OWN = __longlong_as_double((long long)own);
                P_BASE = (double)nS * P_FS;
                P_TG = (nTg > 0.f) ? (double)(nTg * __powf((float)P_FS, 0.2f)) : 0.0;
            }
            if (newk) {
                k1 = nk1;
                k2 = nk2;
                kc = c;
                kl = p.l1;
            }
            const double wl2 = 1.0 - p.wl1;
            double Kabs = k1.x * p.wl1 + k2.x * wl2;
            double Kext = tb.scattering ? ((k1.x + k1.y) * p.wl1 + (k2.x + k2.y) * wl2) : Kabs;
            if (!ok) {
                CNT(C_NGEOM);
                st = S_NEED;
            } else {
                p.nstep++; // added to C_NSTEPS when the packet ends
                double tau1 = ((Kext * v) * c_dens) * AU;
                if ((p.tau + tau1) > p.tau0) {
                    v = (v * (p.tau0 - p.tau)) / tau1;
                    tautot += 2.0 * (p.tau0 - p.tau);
                    p.x = p.x + p.vx * v;
                    p.y = p.y + p.vy * v;
                    p.z = p.z + p.vz * v;
                    p.onEdge = false;
                    double ejv = ((p.E * v) * AU) * Kabs;
                    acc_ejv(acc, ejv, c_u1, scnt);
                    p.nEJv = p.nEJv + ejv;
                    st = S_INTERACT;
                } else {
                    p.x = p.x + p.vx * v;
                    p.y = p.y + p.vy * v;
                    p.z = p.z + p.vz * v;
                    double ejv = ((p.E * v) * AU) * Kabs;
                    p.nEJv = p.nEJv + ejv;
                    acc_ejv2(acc, p.nEJv * p.nEJv, c_u2, scnt);
                    p.nEJv = 0.0;
                    acc_ejv(acc, ejv, c_u1, scnt);
                    p.onEdge = true;
                    p.tau = p.tau + tau1;
                    tautot += tau1;
                    if (inext >= nr) {
                        st = S_ESCAPE;
                    } else if (inext < 0) {
                        CNT(C_NSTARHIT);
                        st = S_EMIT;
                    } else {
                        p.i = inext;
                        p.jj = jnext;
                        p.ntrace++;
                        if (p.ntrace > raar_limit) {
                            CNT(C_NRAAR);
                            st = S_EMIT;
                        }
                    }
                }
            }
        }

        // a pending inverse-CDF lookup and its result: set and consumed within
        // one trip of the loop, so they are not live across the cell step
        Lookup lk;
        lk.R0 = lk.R1 = nullptr;
        lk.e0 = lk.e1 = lk.ws = lk.t = 0.0;
        lk.n = 0;
        lk.kind = L_SAME;
        int lk_b = 0;
        double lk_fb = 0.0, lk_fbm1 = 0.0;

        // ---- interaction up to its CDF lookup (Radiation.f:41-355) ------
        if (st == S_INTERACT) {
            CNT(C_NINTERACT);
            const int c = p.i * nt + p.jj;
            if ((double)c != P_IC) {
                P_IC = (double)c;
                P_MASS = md.mass[c];
                P_TC = (double)((Bm.rw.on ? Bm.rw.thick[c] : 0) | (md.cls[c] << 1));
                P_KDIT = -1.0;
                if (!A.serial) P_TRW = -1.0; // not known yet for this cell
            }
            const int l1 = p.l1, l2 = p.l1 + 1;
            const double wl2 = 1.0 - p.wl1;
            double Ksca = k1.y * p.wl1 + k2.y * wl2;
            double Kext = Ksca + (k1.x * p.wl1 + k2.x * wl2);
            Kext = Kext + 0.0;
            double KabsL = Kext - Ksca;
            double Albedo = Ksca / Kext;
            bool absorb = true;
            if (tb.scattering) absorb = (g.next() > Albedo);
            if (absorb) {
                CNT(C_NABS);
                p.scatt = false;
                p.pol = false;
                P_Q = 0.0;
                P_U = 0.0;
                P_V = 0.0;
                double Ex = (p.E * (KabsL - 0.0)) / KabsL;
                const KpRow kpc = kp_row(md, (int)P_TC >> 1);
                const double *Kp0 = tb.Kp; // species 0
                const double mass = P_MASS;
                int jw;
                double w0, w1, w2, q0 = 0.0, q1 = 0.0, q2 = 0.0, T0 = 0.0, T1;
                if (A.serial) {
                    // The cell's T after this lane's previous absorption in
                    // it is the likely start: fetch that window of the tables
                    // together with the cell's energy (one round trip
                    // instead of several).
                    jw = (c_Tl >= 1.0) ? (int)(c_Tl / 1.0) : -1;
                    if (jw > TMAXD - 2) jw = TMAXD - 2;
                    w0 = w1 = w2 = 0.0;
                    if (jw >= 1) {
                        w0 = kpc[jw];
                        w1 = kpc[jw + 1];
                        w2 = kpc[jw + 2];
                        if (ONE) {
                            q0 = Kp0[jw];
                            q1 = Kp0[jw + 1];
                            q2 = Kp0[jw + 2];
                        }
                    }
                    const double Thint = __longlong_as_double((long long)__hip_atomic_load(
                        &ty.Tcell[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                    unsigned long long q;
                    unsigned long long old = add_eabs(ty, c, Ex, Bm.invE, q, scnt);
                    if (jw < 1) {
                        jw = (int)(Thint / 1.0);
                        if (jw < 1) jw = 1;
                        if (jw > TMAXD - 2) jw = TMAXD - 2;
                        w0 = kpc[jw];
                        w1 = kpc[jw + 1];
                        w2 = kpc[jw + 2];
                        if (ONE) {
                            q0 = Kp0[jw];
                            q1 = Kp0[jw + 1];
                            q2 = Kp0[jw + 2];
                        }
                    }
                    if (old != 0ull) T0 = dev_increaseT_w(kpc, jw, w0, w1, w2, ((double)old * Bm.unitE) / mass, Thint);
                    T1 = dev_increaseT_w(kpc, jw, w0, w1, w2, ((double)(old + q) * Bm.unitE) / mass, T0);
                    atomicMax(&ty.Tcell[c], (unsigned long long)__double_as_longlong(T1));
                    P_TRW = (T1 > Thint) ? T1 : Thint; // the cell's T after this update
                } else {
                    // staged: the slice (S_k p/P_k + own, + q) of this
                    // packet's view of the cell; the add itself returns nothing
                    // anyone waits for
                    const unsigned long long q = eabs_q(Ex, Bm.invE, scnt);
                    gated_deposit(ty, A, md.ncell, c, q, (int)P_PK, p.ndep);
                    const unsigned long long o0 = (unsigned long long)__double_as_longlong(P_OWN), o1 = o0 + q;
                    P_OWN = __longlong_as_double((long long)o1);
                    if (P_VSLOT >= 0.0 && P_VSLOT < (double)(1 << VIS_LOG2)) P_VSLOT += (double)(1 << VIS_LOG2);
                    const double base = P_BASE;
                    const double Eb = (base + (double)o0) * Bm.unitE, E1 = (base + (double)o1) * Bm.unitE;
                    // the window of the previous absorption in this visit
                    // still brackets both energies in a long walk inside the
                    // cell: no load
                    if (!ONE && P_JW >= 1.0 && P_W0 <= Eb / mass && E1 / mass <= P_W2) {
                        jw = (int)P_JW;
                        w0 = P_W0;
                        w1 = P_W1;
                        w2 = P_W2;
                    } else {
                        jw = win_start((c_Tl >= 1.0) ? c_Tl : P_TG);
                        w0 = kpc[jw];
                        w1 = kpc[jw + 1];
                        w2 = kpc[jw + 2];
                        if (ONE) {
                            q0 = Kp0[jw];
                            q1 = Kp0[jw + 1];
                            q2 = Kp0[jw + 2];
                        } else {
                            P_JW = (double)jw;
                            P_W0 = w0;
                            P_W1 = w1;
                            P_W2 = w2;
                        }
                    }
                    if (Eb != 0.0) T0 = dev_T_win(kpc, jw, w0, w1, w2, Eb / mass);
                    T1 = dev_T_win(kpc, jw, w0, w1, w2, E1 / mass);
                    P_TRW = T1;
                }
                c_Tl = T1;
                int iT0 = (int)(T0 / 1.0), iT1 = (int)(T1 / 1.0);
                double e0 = T0 - (double)(float)iT0;
                double e1 = T1 - (double)(float)iT1;
                if (iT0 >= TMAXD - 1) iT0 = TMAXD - 2;
                if (iT1 >= TMAXD - 1) iT1 = TMAXD - 1;
                if (iT0 < 1) iT0 = 1;
                if (iT1 < 2) iT1 = 2;
                const bool same = (iT0 == iT1);
                const double *wc = md.w + (size_t)c * ng;
                // coefficients of the four Planck rows of the difference spectrum
                const double a1 = e1, a2 = 1.0 - e1, a3 = e0, a4 = 1.0 - e0;
                // Planck-mean difference of species s, times its weight
                auto kps = [&](int s) {
                    const double *Kp = tb.Kp + (size_t)s * (TMAXD + 1);
                    double k00, k01, k10, k11;
                    if (ONE) {
                        k00 = win_at(Kp, jw, q0, q1, q2, iT0);
                        k01 = win_at(Kp, jw, q0, q1, q2, iT0 + 1);
                        k10 = win_at(Kp, jw, q0, q1, q2, iT1);
                        k11 = win_at(Kp, jw, q0, q1, q2, iT1 + 1);
                    } else {
                        k00 = Kp[iT0];
                        k01 = Kp[iT0 + 1];
                        k10 = Kp[iT1];
                        k11 = Kp[iT1 + 1];
                    }
                    double d = same ? (k01 - k00) : (((a1 * k11 + a2 * k10) - a3 * k01) - a4 * k00);
                    return (d * wc[s]) * 1.0;
                };
                double kp = 0.0;
                if (MIX) {
                    // the class tables: kp from the cell's mixture Kp
                    const double k00 = win_at(kpc, jw, w0, w1, w2, iT0), k01 = win_at(kpc, jw, w0, w1, w2, iT0 + 1);
                    const double k10 = win_at(kpc, jw, w0, w1, w2, iT1), k11 = win_at(kpc, jw, w0, w1, w2, iT1 + 1);
                    kp = same ? (k01 - k00) : (((a1 * k11 + a2 * k10) - a3 * k01) - a4 * k00);
                } else {
#pragma unroll 4
                    for (int s = 0; s < ng; s++) kp = kp + kps(s);
                }
                // emit(): new direction, then the wavelength
                CNT(C_NEMIT);
                RDIR(g, p.vx, p.vy, p.vz);
                {
                    double sx = -p.vy, sy = p.vx, sz = 0.0;
                    double r = sqrt(sx * sx + sy * sy + sz * sz);
                    P_SX = sx / r;
                    P_SY = sy / r;
                    P_SZ = sz / r;
                }
                double Lr = g.next() * kp;
                // the wavelength from the cell's mixture spectrum: its class
                // table, else the species' rows summed per probe (one
                // species: its own rows)
                const double *Cb = MIX ? md.CUMmix + (size_t)((int)P_TC >> 1) * (TMAXD + 1) * nbins : tb.CUM;
                lk.R0 = Cb + (size_t)iT0 * nbins;
                lk.R1 = Cb + (size_t)iT1 * nbins;
                lk.e0 = e0;
                lk.e1 = e1;
                lk.ws = MIX ? 1.0 : ONE ? wc[0] : __longlong_as_double((long long)wc);
                lk.t = Lr;
                lk.n = nbins;
                lk.kind = same ? L_SAME : L_DIFF;
                st = S_LOOKUP;
            } else {
                CNT(C_NSCAT);
                p.scatt = true;
                if ((int)P_TC & 1) {
                    if (A.serial) {
                        P_TRW = __longlong_as_double((long long)__hip_atomic_load(&ty.Tcell[c], __ATOMIC_RELAXED,
                                                                                   __HIP_MEMORY_SCOPE_AGENT));
                    } else if (P_TRW < 0.0) {
                        // the cell's T as this packet sees it (random-walk test)
                        const double Eb =
                            (P_BASE + (double)(unsigned long long)__double_as_longlong(P_OWN)) * Bm.unitE;
                        double Tv = 0.0;
                        if (Eb != 0.0) {
                            const KpRow kpc = kp_row(md, (int)P_TC >> 1);
                            const int jw = win_start((c_Tl >= 1.0) ? c_Tl : P_TG);
                            Tv = dev_T_win(kpc, jw, kpc[jw], kpc[jw + 1], kpc[jw + 2], Eb / P_MASS);
                        }
                        P_TRW = Tv;
                    }
                }
                if (tb.scat_how == 1) {
                    RDIR(g, p.vx, p.vy, p.vz);
                    st = S_AFTER;
                } else {
                    // pick the (species, lambda) Mueller matrix (Radiation.f:273-290)
                    double KscaR = Ksca * g.next();
                    const double *wc = md.w + (size_t)c * ng;
                    size_t moff = (size_t)(ng - 1) * nlam + l1;
                    double MIF11, MIF12;
                    if (ONE) {
                        // every table value the choice can need, in one round trip
                        const double wv = wc[0], ks2 = tb.Ksca[l2], ks1 = tb.Ksca[l1];
                        const double a11 = tb.IF11[l1], b11 = tb.IF11[l2], a12 = tb.IF12[l1], b12 = tb.IF12[l2];
                        double w1 = wv * 1.0;
                        double w2 = (w1 * wl2) * ks2;
                        KscaR = KscaR - w2;
                        if (KscaR < 0.0) {
                            moff = l2;
                        } else {
                            w1 = (w1 * p.wl1) * ks1;
                            KscaR = KscaR - w1;
                            moff = l1;
                        }
                        MIF11 = (moff == (size_t)l2) ? b11 : a11;
                        MIF12 = (moff == (size_t)l2) ? b12 : a12;
                    } else {
                        // every species' values are loaded whatever the draw
                        // (no early exit), so the loads are in flight together;
                        // the first running sum below zero picks, as in the
                        // reference's loop
                        // the integrals IF11, IF12 of every candidate are
                        // loaded with it, so the pick costs one round trip
                        bool found = false;
                        MIF11 = tb.IF11[moff];
                        MIF12 = tb.IF12[moff];
#pragma unroll 4
                        for (int s = 0; s < ng; s++) {
                            const size_t o1 = (size_t)s * nlam + l1, o2 = o1 + 1;
                            const double ws = wc[s], k2 = tb.Ksca[o2], k1 = tb.Ksca[o1];
                            const double a2 = tb.IF11[o2], b2 = tb.IF12[o2], a1 = tb.IF11[o1], b1 = tb.IF12[o1];
                            double w1 = ws * 1.0;
                            const double w2 = (w1 * wl2) * k2;
                            w1 = (w1 * p.wl1) * k1;
                            if (!found) {
                                KscaR = KscaR - w2;
                                if (KscaR < 0.0) {
                                    moff = o2;
                                    MIF11 = a2;
                                    MIF12 = b2;
                                    found = true;
                                } else {
                                    KscaR = KscaR - w1;
                                    if (KscaR < 0.0) {
                                        moff = o1;
                                        MIF11 = a1;
                                        MIF12 = b1;
                                        found = true;
                                    }
                                }
                            }
                        }
                    }
                    // scatangle (Radiation.f:381-495)
                    double IF11 = MIF11 * 180.0, IF12 = MIF12 * 180.0;
                    double phi;
                    if (p.pol) {
                        // the phi bins' cumulative intensity, bins 1..k:
                        // I(k) = k IF11 E + IF12 (Q sum cos 2phi + U sum sin 2phi);
                        // the first bin whose I exceeds the draw (I is
                        // non-decreasing: |IF12 (Q cos + U sin)| <= IF11 E)
                        const double aE = IF11 * p.E, Qp = P_Q, Up = P_U;
                        auto Ik = [&](int k) { return (double)k * aE + IF12 * (Qp * sPC2[k] + Up * sPS2[k]); };
                        const double Itot = g.next() * Ik(180);
                        int lo = 1, hi = 181;
                        while (lo < hi) {
                            const int mid = (lo + hi) >> 1;
                            if (Ik(mid) > Itot) hi = mid;
                            else lo = mid + 1;
                        }
                        phi = MCX_PI;
                        if (lo <= 180) {
                            const double I0 = Ik(lo), I1 = Ik(lo - 1);
                            const double wp1 = (I0 - Itot) / (I0 - I1);
                            phi = (MCX_PI * ((double)lo - wp1)) / 180.0;
                        }
                    } else {
                        phi = g.next() * MCX_PI;
                    }
                    double sx = P_SX, sy = P_SY, sz = P_SZ;
                    if constexpr (F32) rotate_about_f(sx, sy, sz, p.vx, p.vy, p.vz, phi);
                    else rotate_about(sx, sy, sz, p.vx, p.vy, p.vz, phi);
                    if (p.pol) {
                        // RotateStokes (Radiation.f:506-534)
                        double P = (P_U / p.E) * (P_U / p.E) + (P_Q / p.E) * (P_Q / p.E);
                        if (P != 0.0 && F32) {
                            const float fE = (float)p.E, fvx = (float)p.vx, fvy = (float)p.vy, fvz = (float)p.vz;
                            const float fSX = (float)P_SX, fSY = (float)P_SY, fSZ = (float)P_SZ;
                            const float fsx = (float)sx, fsy = (float)sy, fsz = (float)sz;
                            const float fQ = (float)P_Q, fU = (float)P_U;
                            const float cost = fSX * fsx + fSY * fsy + fSZ * fsz;
                            // the denominator (v x S)_x can round to zero in single
                            // precision: then the Stokes vector keeps its frame
                            const float den = fvy * fSZ - fvz * fSY;
                            const float sint = (fsx - (fSX * (fvy * fvy + fvz * fvz) - fvx * (fvy * fSY + fvz * fSZ)) *
                                                          cost) / (den != 0.0f ? den : 1.0f);
                            const float cos2t = cost * cost - sint * sint;
                            const float sin2t = (2.0f * sint) * cost;
                            const float Qn = fQ * cos2t + fU * sin2t;
                            const float Un = -fQ * sin2t + fU * cos2t;
                            const float P2 = (Un / fE) * (Un / fE) + (Qn / fE) * (Qn / fE);
                            const float sPP2 = sqrtf((float)P / P2);
                            if (den != 0.0f && isfinite(sPP2 * Un) && isfinite(sPP2 * Qn)) {
                                P_U = (double)(sPP2 * Un);
                                P_Q = (double)(sPP2 * Qn);
                            }
                        } else if (P != 0.0) {
                            double cost = P_SX * sx + P_SY * sy + P_SZ * sz;
                            double sint = (sx - (P_SX * (p.vy * p.vy + p.vz * p.vz) -
                                                 p.vx * (p.vy * P_SY + p.vz * P_SZ)) * cost) /
                                          (p.vy * P_SZ - p.vz * P_SY);
                            double cos2t = cost * cost - sint * sint;
                            double sin2t = (2.0 * sint) * cost;
                            double Qn = P_Q * cos2t + P_U * sin2t;
                            double Un = -P_Q * sin2t + P_U * cos2t;
                            double P2 = (Un / p.E) * (Un / p.E) + (Qn / p.E) * (Qn / p.E);
                            double sPP2 = sqrt(P / P2);
                            P_U = sPP2 * Un;
                            P_Q = sPP2 * Qn;
                        }
                    }
                    P_SX = sx;
                    P_SY = sy;
                    P_SZ = sz;
                    double Fr = 180.0 * (p.E * MIF11 + P_Q * MIF12);
                    Fr = g.next() * Fr;
                    // first angle bin whose cumulative E F11 + Q F12 exceeds Fr
                    lk.R1 = tb.SC + moff * 180 * 8;
                    lk.R0 = lk.R1;
                    lk.e1 = p.E;
                    lk.e0 = P_Q;
                    lk.ws = 1.0;
                    lk.t = Fr;
                    lk.n = 180;
                    lk.kind = L_SCAT;
                    st = S_LOOKUP;
                }
            }
        }

        // ---- inverse-CDF lookups: a wave whose few lanes need one serves
        //      them together, one lookup at a time; otherwise each lane
        //      searches alone.  Both give the first bin that reaches the
        //      target.  All lanes of the wave reach this point.
        {
            const unsigned long long sm = __ballot(st == S_LOOKUP);
            if (sm) {
                if (__popcll(sm) <= A.coop) {
                    unsigned long long rem = sm;
                    while (rem) {
                        const int src = __ffsll((long long)rem) - 1;
                        rem &= rem - 1;
                        int b;
                        double fb, fbm1;
                        lk_wave(lk, src, lane, lng, b, fb, fbm1);
                        if (lane == src) {
                            lk_b = b;
                            lk_fb = fb;
                            lk_fbm1 = fbm1;
                        }
                    }
                } else if (st == S_LOOKUP) {
                    lk_single(lk, lng, lk_b, lk_fb, lk_fbm1);
                }
            }
        }

        // ---- the interaction after its lookup ----------------------------
        if (st == S_LOOKUP) {
            if (lk.kind != L_SCAT) {
                // emit()'s wavelength rule (Radiation.f:966-989)
                if (lk_b >= nbins) {
                    p.l1 = nbins - 1;
                    p.wl1 = 0.0;
                } else {
                    p.wl1 = (lk.t - lk_fbm1) / (lk_fb - lk_fbm1);
                    p.l1 = lk_b;
                }
            } else {
                const double Fr = lk.t;
                double theta;
                int ia;
                if (lk_b >= 180) {
                    theta = MCX_PI;
                    ia = 180;
                } else {
                    ia = lk_b + 1;
                    theta = (double)ia - (lk_fb - Fr) / (lk_fb - lk_fbm1);
                    theta = (theta * MCX_PI) / 180.0;
                }
                if (g.next() < 0.5) theta = -theta;
                // the bin's Mueller elements share the line of its cumulatives
                const double *Mb = lk.R1 + 8 * (ia - 1) + 2;
                const double f11 = Mb[0], f12 = Mb[1], f22 = Mb[2], f33 = Mb[3], f34 = Mb[4], f44 = Mb[5];
                double En = 0.0, Qn = 0.0, Un = 0.0, Vn = 0.0;
                En = En + f11 * p.E;
                En = En + f12 * P_Q;
                Qn = Qn + f12 * p.E;
                Qn = Qn + f22 * P_Q;
                Un = Un + f33 * P_U;
                double Fs = (theta >= 0.0) ? fabs(f34) : -fabs(f34);
                Un = Un + Fs * P_V;
                Vn = Vn - Fs * P_U;
                Vn = Vn + f44 * P_V;
                if (En != 0.0) {
                    P_Q = (p.E * Qn) / En;
                    P_U = (p.E * Un) / En;
                    P_V = (p.E * Vn) / En;
                }
                if constexpr (F32) {
                    rotate_about_f(p.vx, p.vy, p.vz, P_SX, P_SY, P_SZ, theta);
                    normalize_f(p.vx, p.vy, p.vz);
                    double sx = P_SX, sy = P_SY, sz = P_SZ;
                    normalize_f(sx, sy, sz);
                    P_SX = sx;
                    P_SY = sy;
                    P_SZ = sz;
                } else {
                    rotate_about(p.vx, p.vy, p.vz, P_SX, P_SY, P_SZ, theta);
                    double r = sqrt(p.vx * p.vx + p.vy * p.vy + p.vz * p.vz);
                    p.vx /= r;
                    p.vy /= r;
                    p.vz /= r;
                    r = sqrt(P_SX * P_SX + P_SY * P_SY + P_SZ * P_SZ);
                    P_SX /= r;
                    P_SY /= r;
                    P_SZ /= r;
                }
                if (ia != 1 && ia != 180) p.pol = true;
            }
            st = S_AFTER;
        }

        // ---- after every interaction (MCRadiation.f:567-573) -------------
        if (st == S_AFTER) {
            st = S_STEP;
            if (A.maxinteract > 0) {
                if (g.next() < (1.0 / (double)(float)A.maxinteract)) {
                    CNT(C_NMAXINT);
                    st = S_NEED;
                } else {
                    p.E = p.E * (double)((float)A.maxinteract / (float)(A.maxinteract - 1));
                }
            }
            if (st == S_STEP) {
                p.tau0 = TAUDRAW(g.next());
                p.tau = 0.0;
                p.ntrace = 0;
                // trace2d starts with the random walk in thick cells (MCRadiation.f:916)
                if ((int)P_TC & 1) st = S_RW;
            }
        }

        // ---- modified random walk (MCRadiation.f:916-929, RandomWalk.f:47-147)
        if (st == S_RW) {
            const DevRW &rw = Bm.rw;
            const int c = p.i * nt + p.jj;
            // MinDist (RandomWalk.f:15-43)
            double rho = p.x * p.x + p.y * p.y;
            const double r = sqrt(rho + p.z * p.z);
            rho = sqrt(rho);
            const double az = fabs(p.z);
            const double vR1 = fabs(r - sRau[p.i]), vR2 = fabs(sRau[p.i + 1] - r);
            const double vT1 = fabs(sTh[p.jj] * rho - sSinT[p.jj] * az);
            const double vT2 = fabs(sTh[p.jj + 1] * rho - sSinT[p.jj + 1] * az);
            double dmin = vR1;
            if (vR2 < dmin) dmin = vR2;
            if (vT1 < dmin) dmin = vT1;
            if (vT2 < dmin) dmin = vT2;
            const double Thint = P_TRW;
            int iT = (int)(Thint / 1.0);
            if (iT == 0) iT = 1;
            if (iT > TMAXD - 1) iT = TMAXD - 1;
            if ((double)iT != P_KDIT) {
                const double2 kq = kd_at(rw.KD, (int)P_TC >> 1, iT);
                P_KDX = kq.x;
                P_KDY = kq.y;
                P_KDIT = (double)iT;
            }
            const double2 kd = make_double2(P_KDX, P_KDY);
            const double lr = 1.0 / ((AU * kd.x) * c_dens);
            if (dmin <= rw.factRW * lr) {
                st = S_STEP;
            } else {
                CNT(C_NRW);
                // the window of the cell's table at the walk's start T, fetched
                // together with the Eabs update
                const KpRow kpc = kp_row(md, (int)P_TC >> 1);
                int jw = (int)(Thint / 1.0);
                if (jw < 1) jw = 1;
                if (jw > TMAXD - 2) jw = TMAXD - 2;
                const double w0 = kpc[jw], w1 = kpc[jw + 1], w2 = kpc[jw + 2];
                const int iy = hunt_desc(rw.phi, NY_RW, g.next());
                // y(iy) of InitRandomWalk (RandomWalk.f:160-164): 1, then a
                // single-precision ratio times 0.75
                const double yv = (iy == 1) ? 1.0 : (double)((float)(NY_RW - iy + 1) / (float)(NY_RW - 1)) * 0.75;
                const double v = -((3.0 * log(yv)) * (dmin * dmin)) / (lr * (MCX_PI * MCX_PI));
                RDIR(g, p.vx, p.vy, p.vz);
                p.x = p.x + dmin * p.vx;
                p.y = p.y + dmin * p.vy;
                p.z = p.z + dmin * p.vz;
                const double ejv = ((p.E * v) * AU) * kd.y;
                acc_ejv(acc, ejv, c_u1, scnt);
                p.nEJv = p.nEJv + ejv; // the module-level nEJv of RandomWalk.f:115
                if (A.serial) {
                    unsigned long long q;
                    const unsigned long long old = add_eabs(ty, c, ejv * c_dens, Bm.invE, q, scnt);
                    const double T1 =
                        dev_increaseT_w(kpc, jw, w0, w1, w2, ((double)(old + q) * Bm.unitE) / P_MASS, Thint);
                    atomicMax(&ty.Tcell[c], (unsigned long long)__double_as_longlong(T1));
                    P_TRW = (T1 > Thint) ? T1 : Thint;
                } else {
                    const unsigned long long q = eabs_q(ejv * c_dens, Bm.invE, scnt);
                    gated_deposit(ty, A, md.ncell, c, q, (int)P_PK, p.ndep);
                    const unsigned long long o1 = (unsigned long long)__double_as_longlong(P_OWN) + q;
                    P_OWN = __longlong_as_double((long long)o1);
                    if (P_VSLOT >= 0.0 && P_VSLOT < (double)(1 << VIS_LOG2)) P_VSLOT += (double)(1 << VIS_LOG2);
                    P_TRW = dev_T_win(kpc, jw, w0, w1, w2, ((P_BASE + (double)o1) * Bm.unitE) / P_MASS);
                }
                tautot += v / dmin;
                // leave in a direction of the forward hemisphere (MCRadiation.f:919-925)
                double vx, vy, vz;
                for (;;) {
                    RDIR(g, vx, vy, vz);
                    const double inp = p.vx * vx + p.vy * vy + p.vz * vz;
                    if (!(inp < 0.0)) break;
                }
                p.vx = vx;
                p.vy = vy;
                p.vz = vz;
            }
        }

        // ---- escape: MC spectrum (MCRadiation.f:576-646) ----------------
        if (st == S_ESCAPE) {
            CNT(C_NESCAPE);
            int ia = (p.vz > 0.0) ? (int)(p.vz * 30.0 + 1.0) : (int)(-p.vz * 30.0 + 1.0);
            if (ia > NANGD) ia = NANGD;
            int il = (p.wl1 > 1.0 - p.wl1) ? p.l1 : p.l1 + 1;
            size_t bin = (size_t)ia * nlam + il;
            unsigned long long q = (unsigned long long)rint(p.E * Bm.invE);
            atomicAdd(&ty.specE[bin], q);
            atomicAdd(&ty.ispec[bin], 1u);
            if (p.scatt) {
                size_t off = (size_t)(NANGD + 1) * nlam;
                atomicAdd(&ty.specE[off + bin], q);
                atomicAdd(&ty.ispec[off + bin], 1u);
            }
            st = S_NEED;
        }
    }
    acc_flush(ty, acc, scnt);

    // ---- counters: one global atomic per block -------------------------
    if (tautot != 0.0) atomicAdd(&stau, tautot);
    __syncthreads();
    for (int k = threadIdx.x; k < C_NCOUNT; k += blockDim.x)
        if (scnt[k]) atomicAdd(&ty.counters[k], scnt[k]);
    if (threadIdx.x == 0 && stau != 0.0) atomicAdd(ty.tautot, stau);
    __syncthreads();
#undef CNT
#undef P_Q
#undef P_U
#undef P_V
#undef P_SX
#undef P_SY
#undef P_SZ
#undef P_TRW
#undef P_KDX
#undef P_KDY
#undef P_MASS
#undef P_IC
#undef P_TC
#undef P_KDIT
#undef P_FS
#undef P_BASE
#undef P_PK
#undef P_OWN
#undef P_VSLOT
#undef P_TG
#undef P_JW
#undef P_W0
#undef P_W1
#undef P_W2
}

// One launch for the pass of every model: workgroup b starts on its model
// (the block map, or b / bpm).  With help on, a workgroup whose model has no
// packet left to take joins the model with the most packets not yet taken
// (its own tables in LDS, its lanes' visit tables cleared), until every
// packet of every model is taken.  A packet's result does not depend on the
// lane or workgroup that runs it.
// (HELP: the instance for batched launches; the loop costs registers, so a
// single model's pass runs the instance without it.)  A single model's pass
// runs one workgroup of four waves per CU: the packet state needs more than
// the 256 VGPRs of two waves per SIMD, and at one wave per SIMD the rest
// lives in AGPRs instead of scratch (measured: the production pass 10% and
// a lone lane's step 5% faster than at two waves per SIMD with ~35 VGPRs
// spilled).  A batched launch is bound by its lanes' throughput instead
// (the 64-model sweep: 137 s at two workgroups per CU, 163 s at one), so
// the HELP instance runs two.
template <int MODE, bool F32, bool HELP>
__global__ void __launch_bounds__(256, HELP ? 2 : 1) k_transport(TransportArgs A) {
    int mi = A.bmap ? A.bmap[blockIdx.x] : (int)(blockIdx.x / A.bpm);
    if constexpr (!HELP) {
        transport_model<MODE, F32>(A, mi);
        return;
    }
    __shared__ int s_next;
    for (;;) {
        transport_model<MODE, F32>(A, mi);
        if (threadIdx.x == 0) {
            long long best = 0;
            int bm = -1;
            // past the time limit every packet is given up: no more joining
            const bool late = A.deadline && wall_clock64() > A.deadline;
            for (int m = 0; m < A.nmodels && !late; m++) {
                const unsigned long long c =
                    __hip_atomic_load(A.bm[m].next_pid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const long long left = A.pid_end - (long long)c;
                if (left > best) {
                    best = left;
                    bm = m;
                }
            }
            s_next = bm;
        }
        __syncthreads();
        mi = s_next;
        __syncthreads();
        if (mi < 0) return;
        // the visit tables are tagged by packet number: clear this
        // workgroup's lanes' tables before packets of another model
        uint4 *vb = A.bm[mi].ty.vtab + ((size_t)blockIdx.x * blockDim.x << VIS_LOG2);
        const size_t nv = (size_t)blockDim.x << VIS_LOG2;
        for (size_t k = threadIdx.x; k < nv; k += blockDim.x) vb[k] = make_uint4(0u, 0u, 0u, 0u);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// per-cell tables (MCRadiation.f:210-229; Radiation.f:611-627)
__global__ void k_cell_opacity(int ncell, int nlam, int ng, const double *__restrict__ Kabs,
                               const double *__restrict__ Ksca, const double *__restrict__ w, double2 *KT) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)ncell * nlam) return;
    int c = (int)(idx / nlam), l = (int)(idx % nlam);
    double ka = 0.0, ks = 0.0;
    for (int s = 0; s < ng; s++) {
        ka = ka + (Kabs[(size_t)s * nlam + l] * w[(size_t)c * ng + s]) * 1.0;
        ks = ks + (Ksca[(size_t)s * nlam + l] * w[(size_t)c * ng + s]) * 1.0;
    }
    KT[idx] = make_double2(ka, ks);
}

// computeE (Radiation.f:611-627) as a table per species-mix class (the cells
// of a class have bitwise equal weights, so their rows are equal)
__global__ void k_class_kp(int nclass, int ng, const double *__restrict__ Kp, const double *__restrict__ wcls,
                           double *KpCls) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)nclass * (TMAXD + 1)) return;
    int k = (int)(idx / (TMAXD + 1)), T = (int)(idx % (TMAXD + 1));
    KpCls[idx] = KpRow{nullptr, Kp, wcls + (size_t)k * ng, ng}[T];
}

// the reemission cumulatives of each species-mix class: sum_s w_s CUM_s
// (the cell's KabsTot spectrum, Radiation.f:199-214, per integer T)
__global__ void k_cum_mix(int nclass, int ng, int nbins, const double *__restrict__ CUM, const double *__restrict__ wcls,
                          double *CUMmix) {
    const size_t n1 = (size_t)(TMAXD + 1) * nbins;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)nclass * n1) return;
    const int k = (int)(idx / n1);
    const size_t tb = idx - (size_t)k * n1;
    double e = 0.0;
    for (int s = 0; s < ng; s++) e = e + (wcls[(size_t)k * ng + s] * CUM[(size_t)s * n1 + tb]) * 1.0;
    CUMmix[idx] = e;
}

// fixed-point units: a path segment inside a spherical shell is at most the
// tangent chord 2 sqrt(R2^2 - R1^2); with the cell's largest KabsTot this
// bounds one EJv contribution.
__global__ void k_cell_units(int nr, int nt, int nlam, const double *__restrict__ R2, const double2 *__restrict__ KT,
                             double E0, double *invU1, double *invU2) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nr * nt) return;
    int i = c / nt;
    double kmax = 0.0;
    for (int l = 0; l < nlam; l++) kmax = fmax(kmax, KT[(size_t)c * nlam + l].x);
    double chord = 2.0 * sqrt(fmax(R2[i + 1] - R2[i], 0.0));
    double B = E0 * MCX_AU * chord * kmax * (1.0 + 1.0 / 64.0);
    if (!(B > 0.0)) B = 1e-300;
    invU1[c] = ldexp(1.0, EJV_BITS) / B;
    invU2[c] = ldexp(1.0, EJV2_BITS) / (B * B);
}

// DiffCoeffCell (Dtable.f:97-182) for every species-mix class and integer T
__global__ void k_diffcoeff(int nclass, const KDSrc S, double2 *KD) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)nclass * (TMAXD + 1)) return;
    KD[idx] = dev_diffcoeff(S, (int)(idx / (TMAXD + 1)), (int)(idx % (TMAXD + 1)));
}

// without the table: each class's per-wavelength terms of DiffCoeffCell and
// its coefficients at the thick test's four temperatures
__global__ void k_diffcoeff_terms(int nclass, const KDSrc S, double2 *AK, double2 *K4) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nak = (size_t)nclass * S.nlam;
    if (idx < nak) {
        const int k = (int)(idx / S.nlam), l = (int)(idx % S.nlam);
        AK[idx] = kd_terms(S, S.wcls + (size_t)k * S.ng, l, true);
    } else if (idx < nak + (size_t)nclass * 4) {
        const size_t q = idx - nak;
        K4[q] = dev_diffcoeff(S, (int)(q / 4), 100 + 500 * (int)(q % 4));
    }
}

__global__ void k_fill_u4(size_t n, uint4 *p) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        p[k] = make_uint4(0u, 0u, 0u, 0u);
}

__global__ void k_set_u64(unsigned long long *p, unsigned long long v) { *p = v; }

__global__ void k_fill_u64(size_t n, unsigned long long *p, unsigned long long v) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) p[k] = v;
}

__global__ void k_fill_u32(size_t n, unsigned *p, unsigned v) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) p[k] = v;
}

__global__ void k_clock(unsigned long long *p) { *p = wall_clock64(); }

// raw tallies from the fixed-point sums; Tbw = the cell's Bjorkman-Wood T
// after the pass, T(Eabs/m) (the serial loop's last increaseT)
__global__ void k_finalize(int ncell, int nb, const unsigned long long *EJvq, const unsigned long long *EJv2q,
                           const unsigned long long *Eabsq, const double *invU1, const double *invU2, double unitE,
                           const DevModel md, double *EJv, double *EJv2, double *Eabs, double *Tbw) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    EJv[c] = (double)EJvq[c] / invU1[c];
    const unsigned long long *L = EJv2q + 3 * (size_t)c;
    double s = ((double)L[2] * 0x1p64 + (double)L[1] * 0x1p32) + (double)L[0];
    EJv2[c] = s / invU2[c];
    // the pass's absorbed energy: every epoch's deposits
    unsigned long long e = 0ull;
    for (int b = 0; b < nb; b++) e += Eabsq[(size_t)b * ncell + c];
    Eabs[c] = (double)e * unitE;
    Tbw[c] = e ? dev_determineT(kp_row(md, md.cls[c]), ((double)e * unitE) / md.mass[c], 1.0) : 0.0;
}

// Main.f:498-549 per cell; void cell temperature Main.f:546-548.
__global__ void k_temperature(int nr, int nt, int ng, const double *__restrict__ EJv, const double *__restrict__ EJv2,
                              const int *__restrict__ Ni, const double *__restrict__ Tbw, const double *__restrict__ V,
                              const DevModel md, const double *__restrict__ w, const double *__restrict__ Kp,
                              double Tvoid, double *T, double *dT, double *dEJv) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nr * nt) return;
    int i = c / nt;
    if (i == 0) {
        T[c] = Tvoid;
        dT[c] = 1.0;
        dEJv[c] = 0.0;
        return;
    }
    double Vc = V[c];
    double E = EJv[c] / Vc;
    double E2 = EJv2[c] / (Vc * Vc);
    double Tc = dev_determineT(kp_row(md, md.cls[c]), E, Tbw[c]);
    if (Tc < 1.0) Tc = 1.0;
    double dTc, dE;
    if (Ni[c] > 1) {
        float fN = (float)Ni[c];
        dE = (E + sqrt(fabs(E2 - (E * E) / (double)fN))) / (double)sqrtf(fN);
        dTc = ((0.25 * Tc) * dE) / E;
    } else {
        dTc = (double)(float)TMAXD;
        dE = 0.0;
        for (int s = 0; s < ng; s++) dE = dE + (w[(size_t)c * ng + s] * 1.0) * Kp[(size_t)s * (TMAXD + 1) + TMAXD];
    }
    if (isnan(Tc)) Tc = 1.0;
    if (isnan(dTc)) dTc = 1.0;
    T[c] = Tc;
    dT[c] = dTc;
    dEJv[c] = dE;
}

// ===========================================================================
// host side
static double powi_rl(double x, unsigned k) {
    double r = (k & 1u) ? x : 1.0;
    while (k >>= 1) {
        x = x * x;
        if (k & 1u) r = r * x;
    }
    return r;
}

// InitRandomWalk's tables (RandomWalk.f:160-169): y(i) and the survival
// probability phi(y) = 2 sum_n (-1)^(n+1) y^(n^2), 1000 terms.
static void random_walk_tables(std::vector<double> &y, std::vector<double> &phi) {
    const int NY = NY_RW, nmax = 1000;
    y.assign(NY, 0.0);
    phi.assign(NY, 0.0);
    y[0] = 1.0;
    phi[0] = 1.0;
    for (int i = 2; i <= NY; i++) {
        y[i - 1] = (double)((float)(NY - i + 1) / (float)(NY - 1)) * 0.75;
        double sum = 0.0;
        for (int n = 1; n <= nmax; n++) {
            double sg = ((n + 1) % 2 == 0) ? 1.0 : -1.0;
            sum = sum + sg * powi_rl(y[i - 1], (unsigned)(n * n));
        }
        phi[i - 1] = sum * 2.0;
    }
}

extern "C" int mcx_create(mcx_ctx **out, int device) {
    if (!out) return MCX_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MCX_EDEVICE;
    if (device < 0 || device >= ndev) return MCX_EINVAL;
    mcx_ctx *ctx = new mcx_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return MCX_EDEVICE;
    }
    for (int k = 0; k < 4; k++) (void)hipEventCreate(&ctx->ev[k]);
    hipDeviceProp_t prop;
    (void)hipGetDeviceProperties(&prop, device);
    ctx->sm_count = prop.multiProcessorCount;
    ctx->lds_max = prop.sharedMemPerBlock;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess) ctx->wclk_khz = khz;
    *out = ctx;
    return MCX_OK;
}

extern "C" void mcx_destroy(mcx_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    free_list(ctx->run_alloc);
    free_list(ctx->mod_alloc);
    free_list(ctx->tab_alloc);
    if (ctx->d_bms) (void)hipFree(ctx->d_bms);
    if (ctx->d_bmap) (void)hipFree(ctx->d_bmap);
    if (ctx->gexec) (void)hipGraphExecDestroy((hipGraphExec_t)ctx->gexec);
    if (ctx->graph) (void)hipGraphDestroy((hipGraph_t)ctx->graph);
    release_blas(ctx);
    release_structure(ctx);
    for (int k = 0; k < 4; k++)
        if (ctx->ev[k]) (void)hipEventDestroy(ctx->ev[k]);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" const char *mcx_last_error(const mcx_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" int mcx_set_tables(mcx_ctx *ctx, const mcx_tables *t) {
    if (!ctx || !t) return MCX_EINVAL;
    if (t->tmax != TMAXD) return fail(ctx, MCX_EINVAL, "tmax must be 10000");
    if (t->nlam < 2 || t->ng < 1 || t->ng > MAXNG) return fail(ctx, MCX_EINVAL, "nlam >= 2 and 1 <= ng <= 16 required");
    if (!t->nu || !t->BB || !t->Kabs || !t->Ksca || !t->Kp || !t->Fstar || !t->IF11 || !t->IF12 || !t->cos2phi ||
        !t->sin2phi)
        return fail(ctx, MCX_EINVAL, "missing table");
    if (t->scat_how == 2 && !t->F) return fail(ctx, MCX_EINVAL, "scat_how=2 needs the Mueller matrices");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    free_list(ctx->run_alloc);
    free_list(ctx->mod_alloc);
    free_list(ctx->tab_alloc);
    ctx->vtab_lanes = 0;
    ctx->ty.vtab = nullptr;
    ctx->have_model = ctx->have_pass = false;
    const int nlam = t->nlam, ng = t->ng, nb = nlam - 1;
    ctx->nlam = nlam;
    ctx->ng = ng;
    ctx->Lstar = t->Lstar;
    ctx->h_nu.assign(t->nu, t->nu + nlam);
    ctx->h_Kp.assign(t->Kp, t->Kp + (size_t)ng * (TMAXD + 1));
    // cumulative trapezoids (the running sum emit() forms, Radiation.f:970-971)
    std::vector<double> cumStar(nb), CUM((size_t)ng * (TMAXD + 1) * nb);
    {
        double Lt = 0.0;
        for (int b = 0; b < nb; b++) {
            Lt = Lt + (fabs(t->nu[b] - t->nu[b + 1]) * 0.5) * (t->Fstar[b] + t->Fstar[b + 1]);
            cumStar[b] = Lt;
        }
    }
    for (int s = 0; s < ng; s++) {
        const double *K = t->Kabs + (size_t)s * nlam;
        for (int T = 0; T <= TMAXD; T++) {
            const double *B = t->BB + (size_t)T * nlam;
            double *row = CUM.data() + ((size_t)s * (TMAXD + 1) + T) * nb;
            double Lt = 0.0;
            for (int b = 0; b < nb; b++) {
                Lt = Lt + (fabs(t->nu[b] - t->nu[b + 1]) * 0.5) * (B[b] * K[b] + B[b + 1] * K[b + 1]);
                row[b] = Lt;
            }
        }
    }
    // scattering-angle cumulatives sum_{m<=a} pi sin(theta_m) F1x(m), each
    // bin's record beside that bin's Mueller elements (one 64-byte line)
    std::vector<double> SC;
    const bool full = (t->scat_how == 2);
    if (full) {
        SC.resize((size_t)ng * nlam * 180 * 8);
        for (size_t sl = 0; sl < (size_t)ng * nlam; sl++) {
            const double *M = t->F + sl * 6 * 180;
            double a11 = 0.0, a12 = 0.0;
            for (int a = 1; a <= 180; a++) {
                double thet = (MCX_PI * ((double)a - 0.5)) / 180.0;
                double ps = MCX_PI * sin(thet);
                a11 = a11 + ps * M[a - 1];
                a12 = a12 + ps * M[180 + a - 1];
                double *r = SC.data() + (sl * 180 + a - 1) * 8;
                r[0] = a11;
                r[1] = a12;
                for (int k = 0; k < 6; k++) r[2 + k] = M[k * 180 + a - 1];
            }
        }
    }
    int scattering = 0;
    if (t->scat_how != 0)
        for (size_t k = 0; k < (size_t)ng * nlam; k++)
            if (t->Ksca[k] != 0.0) scattering = 1;
    DevTables &tb = ctx->tb;
    tb.nlam = nlam;
    tb.ng = ng;
    tb.scat_how = t->scat_how;
    tb.scattering = scattering;
    tb.Lstar = t->Lstar;
    double *p;
    int r;
#define UP(dst, src, n)                                                   \
    do {                                                                  \
        r = dupload(ctx, ctx->tab_alloc, &p, (const double *)(src), (n)); \
        if (r) return r;                                                  \
        dst = p;                                                          \
    } while (0)
    UP(tb.nu, t->nu, nlam);
    UP(tb.cumStar, cumStar.data(), nb);
    UP(tb.CUM, CUM.data(), CUM.size());
    UP(tb.Kp, t->Kp, (size_t)ng * (TMAXD + 1));
    UP(tb.Ksca, t->Ksca, (size_t)ng * nlam);
    UP(tb.IF11, t->IF11, (size_t)ng * nlam);
    UP(tb.IF12, t->IF12, (size_t)ng * nlam);
    UP(tb.cos2phi, t->cos2phi, 361);
    UP(tb.sin2phi, t->sin2phi, 361);
    {
        // prefix sums of cos 2phi(i), sin 2phi(i) over the scattering bins i = 1..180
        std::vector<double> pc(181, 0.0), ps(181, 0.0);
        for (int i = 1; i <= 180; i++) {
            pc[i] = pc[i - 1] + t->cos2phi[i];
            ps[i] = ps[i - 1] + t->sin2phi[i];
        }
        UP(tb.pc2, pc.data(), 181);
        UP(tb.ps2, ps.data(), 181);
    }
    if (full) {
        UP(tb.F, t->F, (size_t)ng * nlam * 6 * 180);
        UP(tb.SC, SC.data(), SC.size());
    } else {
        tb.F = tb.SC = nullptr;
    }
    UP(ctx->d_Kabs, t->Kabs, (size_t)ng * nlam);
    ctx->d_Ksca = (double *)tb.Ksca;
    ctx->d_Kext = ctx->d_gsym = ctx->d_BB = nullptr;
    UP(ctx->d_BB, t->BB, (size_t)(TMAXD + 1) * nlam);
    if (t->Kext) UP(ctx->d_Kext, t->Kext, (size_t)ng * nlam);
    if (t->gsym) UP(ctx->d_gsym, t->gsym, (size_t)ng * nlam);
    ctx->h_Kext.clear();
    if (t->Kext) ctx->h_Kext.assign(t->Kext, t->Kext + (size_t)ng * nlam);
    ctx->h_lam.assign(t->lam ? t->lam : t->nu, (t->lam ? t->lam : t->nu) + nlam);
    ctx->h_Fstar.assign(t->Fstar, t->Fstar + nlam);
    {
        std::vector<double> y, phi;
        random_walk_tables(y, phi);
        UP(ctx->d_rwy, y.data(), NY_RW);
        UP(ctx->d_rwphi, phi.data(), NY_RW);
    }
#undef UP
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->have_tables = true;
    return MCX_OK;
}

extern "C" int mcx_set_model(mcx_ctx *ctx, const mcx_grid *gr) {
    if (!ctx || !gr) return MCX_EINVAL;
    if (!ctx->have_tables) return fail(ctx, MCX_ESTATE, "mcx_set_tables first");
    if (gr->nr < 2 || gr->nt < 1 || !gr->R || !gr->Theta || !gr->dens || !gr->V || !gr->w)
        return fail(ctx, MCX_EINVAL, "bad grid");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    free_list(ctx->run_alloc);
    free_list(ctx->mod_alloc);
    ctx->vtab_lanes = 0;
    ctx->ty.vtab = nullptr;
    ctx->have_model = ctx->have_pass = false;
    const int nr = gr->nr, nt = gr->nt, ncell = nr * nt, ng = ctx->ng, nlam = ctx->nlam;
    std::vector<double> R2(nr + 1), T2(nt + 1), mass(ncell);
    for (int i = 0; i <= nr; i++) R2[i] = gr->R[i] * gr->R[i];
    for (int j = 0; j <= nt; j++) T2[j] = gr->Theta[j] * gr->Theta[j];
    for (int c = 0; c < ncell; c++) mass[c] = gr->V[c] * gr->dens[c];
    DevModel &md = ctx->md;
    md.nr = nr;
    md.nt = nt;
    md.ncell = ncell;
    md.R0 = gr->R[0];
    double *p;
    int r;
#define UP(dst, src, n)                                                   \
    do {                                                                  \
        r = dupload(ctx, ctx->mod_alloc, &p, (const double *)(src), (n)); \
        if (r) return r;                                                  \
        dst = p;                                                          \
    } while (0)
    UP(md.R2, R2.data(), nr + 1);
    UP(md.T2, T2.data(), nt + 1);
    UP(md.Theta, gr->Theta, nt + 1);
    UP(md.dens, gr->dens, ncell);
    UP(md.mass, mass.data(), ncell);
    UP(md.w, gr->w, (size_t)ncell * ng);
    UP(ctx->d_V, gr->V, ncell);
#undef UP
    ctx->d_R2 = (double *)md.R2;
    double2 *KT;
    if ((r = dalloc(ctx, ctx->mod_alloc, &KT, (size_t)ncell * nlam))) return r;

    {
        size_t n = (size_t)ncell * nlam;
        hipLaunchKernelGGL(k_cell_opacity, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, ncell, nlam, ng,
                           ctx->d_Kabs, ctx->d_Ksca, md.w, KT);
        HIPCHK(ctx, hipGetLastError());
    }
    md.KT = KT;
    // species-mix classes (cells with bitwise equal weights), radial edges;
    // the random walk's diffusion-coefficient table when Kext and g are known
    ctx->rw = DevRW{};
    ctx->nclass = 0;
    ctx->d_thick = nullptr;
    ctx->d_wcls = nullptr;
    {
        std::map<std::vector<uint64_t>, int> seen;
        std::vector<int> cls(ncell);
        std::vector<double> wcls;
        for (int c = 0; c < ncell; c++) {
            std::vector<uint64_t> key(ng);
            memcpy(key.data(), gr->w + (size_t)c * ng, sizeof(double) * ng);
            auto it = seen.find(key);
            if (it == seen.end()) {
                it = seen.emplace(key, (int)seen.size()).first;
                wcls.insert(wcls.end(), gr->w + (size_t)c * ng, gr->w + (size_t)(c + 1) * ng);
            }
            cls[c] = it->second;
        }
        const int nclass = (int)seen.size();
        const double *dR;
        int *dcls;
        if ((r = dupload(ctx, ctx->mod_alloc, &ctx->d_wcls, wcls.data(), wcls.size()))) return r;
        if ((r = dupload(ctx, ctx->mod_alloc, (double **)&dR, gr->R, (size_t)nr + 1))) return r;
        if ((r = dupload(ctx, ctx->mod_alloc, &dcls, cls.data(), (size_t)ncell))) return r;
        ctx->nclass = nclass;
        ctx->rw.cls = dcls;
        md.cls = dcls;
        md.wcls = ctx->d_wcls;
        md.Kp = ctx->tb.Kp;
        md.ng = ng;
        // Tables per class and integer T: the Planck means (8 B per entry),
        // the diffusion coefficients (16 B) and, with several species, the
        // reemission cumulatives (8 (nlam-1) B).  Each is built while it
        // fits in 4 GiB, or in the bytes of the environment's
        // MCX_CLASS_TABLE_BYTES (many contexts on one GPU); otherwise the
        // Planck means and diffusion coefficients are evaluated at each read
        // (the same bits) and the reemission sums the species per probe.
        size_t cap = (size_t)4 << 30;
        if (const char *e = getenv("MCX_CLASS_TABLE_BYTES")) cap = (size_t)strtoull(e, nullptr, 10);
        const size_t nT = (size_t)nclass * (TMAXD + 1);
        md.KpCls = nullptr;
        if (nT * sizeof(double) <= cap) {
            double *KpCls;
            if ((r = dalloc(ctx, ctx->mod_alloc, &KpCls, nT))) return r;
            hipLaunchKernelGGL(k_class_kp, dim3((unsigned)((nT + 255) / 256)), dim3(256), 0, ctx->stream, nclass, ng,
                               ctx->tb.Kp, ctx->d_wcls, KpCls);
            HIPCHK(ctx, hipGetLastError());
            md.KpCls = KpCls;
        }
        md.CUMmix = nullptr;
        const int nb = nlam - 1;
        const size_t nmix = nT * nb;
        if (ng > 1 && nmix * sizeof(double) <= cap) {
            double *mix;
            if ((r = dalloc(ctx, ctx->mod_alloc, &mix, nmix))) return r;
            hipLaunchKernelGGL(k_cum_mix, dim3((unsigned)((nmix + 255) / 256)), dim3(256), 0, ctx->stream, nclass, ng,
                               nb, ctx->tb.CUM, ctx->d_wcls, mix);
            HIPCHK(ctx, hipGetLastError());
            md.CUMmix = mix;
        }
        ctx->rw.R = dR;
        if (ctx->d_Kext && ctx->d_gsym && gr->thet) {
            std::vector<double> sinT(nt + 1);
            for (int j = 0; j <= nt; j++) sinT[j] = sin(gr->thet[j]);
            const double *dsin;
            if ((r = dupload(ctx, ctx->mod_alloc, (double **)&dsin, sinT.data(), (size_t)nt + 1))) return r;
            if ((r = dalloc(ctx, ctx->mod_alloc, &ctx->d_thick, (size_t)ncell))) return r;
            KDSrc S{};
            S.on = 1;
            S.Kabs = ctx->d_Kabs;
            S.Ksca = ctx->d_Ksca;
            S.Kext = ctx->d_Kext;
            S.gsym = ctx->d_gsym;
            S.wcls = ctx->d_wcls;
            S.BB = ctx->d_BB;
            S.nu = ctx->tb.nu;
            S.nlam = nlam;
            S.ng = ng;
            S.scattering = ctx->tb.scattering;
            if (nT * sizeof(double2) <= cap) {
                double2 *KD;
                if ((r = dalloc(ctx, ctx->mod_alloc, &KD, nT))) return r;
                hipLaunchKernelGGL(k_diffcoeff, dim3((unsigned)((nT + 255) / 256)), dim3(256), 0, ctx->stream, nclass,
                                   S, KD);
                HIPCHK(ctx, hipGetLastError());
                S.t = KD;
            } else {
                // the per-wavelength terms (16 nlam B per class) make an
                // evaluation ~ nlam divisions instead of 3 ng nlam loads
                double2 *AK, *K4;
                const size_t nak = (size_t)nclass * nlam, n = nak + (size_t)nclass * 4;
                if ((r = dalloc(ctx, ctx->mod_alloc, &AK, nak))) return r;
                if ((r = dalloc(ctx, ctx->mod_alloc, &K4, (size_t)nclass * 4))) return r;
                hipLaunchKernelGGL(k_diffcoeff_terms, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                                   nclass, S, AK, K4);
                HIPCHK(ctx, hipGetLastError());
                S.AK = AK;
                S.K4 = K4;
            }
            ctx->rw.KD = S;
            ctx->rw.sinT = dsin;
            ctx->rw.y = ctx->d_rwy;
            ctx->rw.phi = ctx->d_rwphi;
            ctx->rw.thick = ctx->d_thick;
        }
    }
    // host copies for the observation geometry (mcx_sed)
    ctx->h_R.assign(gr->R, gr->R + nr + 1);
    ctx->h_Theta.assign(gr->Theta, gr->Theta + nt + 1);
    ctx->h_thet.clear();
    if (gr->thet) ctx->h_thet.assign(gr->thet, gr->thet + nt + 1);
    ctx->h_dens.assign(gr->dens, gr->dens + ncell);
    ctx->h_w.assign(gr->w, gr->w + (size_t)ncell * ng);
    ctx->have_temps = false;
    {
        // Main.f:546
        const double AU = MCX_AU;
        double q = gr->Rstar_cm / (gr->R[1] * AU);
        double T0 = gr->Tstar * 0.5 * (1.0 - sqrt(1.0 - q * q));
        ctx->Tvoid = (T0 < 1.0) ? 1.0 : T0;
    }
    // per-pass tally buffers
    DevTally &ty = ctx->ty;
    double *invU1, *invU2;
    if ((r = dalloc(ctx, ctx->mod_alloc, &ty.EJv, ncell))) return r;
    if ((r = dalloc(ctx, ctx->mod_alloc, &ty.EJv2, (size_t)3 * ncell))) return r;
    if ((r = dalloc(ctx, ctx->mod_alloc, &ty.Eabs, (size_t)(KSTAGE_MAX + 2) * ncell))) return r;
    if ((r = dalloc(ctx, ctx->mod_alloc, &invU1, ncell))) return r;
    if ((r = dalloc(ctx, ctx->mod_alloc, &invU2, ncell))) return r;
    ty.invU1 = invU1;
    ty.invU2 = invU2;
    if ((r = dalloc(ctx, ctx->mod_alloc, &ty.Tcell, ncell))) return r;
    if ((r = dalloc(ctx, ctx->mod_alloc, &ty.specE, (size_t)2 * (NANGD + 1) * nlam))) return r;
    if ((r = dalloc(ctx, ctx->mod_alloc, &ty.ispec, (size_t)2 * (NANGD + 1) * nlam))) return r;
    if ((r = dalloc(ctx, ctx->mod_alloc, &ty.counters, G_NSLOTS))) return r;
    if ((r = dalloc(ctx, ctx->mod_alloc, &ty.tautot, 1))) return r;
    if ((r = dalloc(ctx, ctx->mod_alloc, &ty.torep, 16))) return r;
    if ((r = dalloc(ctx, ctx->mod_alloc, &ctx->d_Ni, ncell))) return r;
    if ((r = dalloc(ctx, ctx->mod_alloc, &ctx->d_Sk, (size_t)(KSTAGE_MAX + 1) * ncell))) return r;
    if ((r = dalloc(ctx, ctx->mod_alloc, &ctx->d_Tk, (size_t)(KSTAGE_MAX + 1) * ncell))) return r;
    if ((r = dalloc(ctx, ctx->mod_alloc, &ctx->d_gate, G_GATE_SLOTS))) return r;
    if ((r = dalloc(ctx, ctx->mod_alloc, &ctx->d_stages, KSTAGE_MAX + 1))) return r;
    if ((r = dalloc(ctx, ctx->mod_alloc, &ctx->d_EJv, ncell))) return r;
    if ((r = dalloc(ctx, ctx->mod_alloc, &ctx->d_EJv2, ncell))) return r;
    if ((r = dalloc(ctx, ctx->mod_alloc, &ctx->d_Eabs, ncell))) return r;
    if ((r = dalloc(ctx, ctx->mod_alloc, &ctx->d_Tbw, ncell))) return r;
    if ((r = dalloc(ctx, ctx->mod_alloc, &ctx->d_T, ncell))) return r;
    if ((r = dalloc(ctx, ctx->mod_alloc, &ctx->d_dT, ncell))) return r;
    if ((r = dalloc(ctx, ctx->mod_alloc, &ctx->d_dEJv, ncell))) return r;
    if ((r = dalloc(ctx, ctx->mod_alloc, &ctx->d_next, 1))) return r;

    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->have_model = true;
    return MCX_OK;
}

// the end of the stage that starts at packet P: ratio x P (ratio >= 2), or
// with ratio -1 the growing ratio min(16, 1 + sqrt(P / stage0)); at most N
static long long stage_next(long long P, long long N, long long stage0, int ratio) {
    double r = (ratio >= 2) ? (double)ratio : 1.0 + sqrt((double)P / (double)stage0);
    if (r > 16.0) r = 16.0;
    const double nx = ceil((double)P * r);
    return (nx >= (double)N) ? N : (long long)nx;
}

// The stage table of a concurrent pass of N packets: P[k] the first packet
// of stage k, P[nstage] = N; one stage for a one-lane pass.  Sets nstage and
// depB in A; false when the schedule has more than KSTAGE_MAX stages.
struct StageTab {
    long long P[KSTAGE_MAX + 1];
};

static bool stage_table(const mcx_run *run, bool serial, TransportArgs &A, StageTab &T) {
    const long long N = run->nphot;
    const long long s0 = run->stage0 > 0 ? run->stage0 : 64;
    const int ratio = run->stage_ratio == 0 ? 2 : run->stage_ratio;
    for (int k = 0; k <= KSTAGE_MAX; k++) T.P[k] = N;
    T.P[0] = 0;
    int ns = 0;
    if (serial) {
        ns = 1;
    } else {
        long long a = 0, b = std::min(N, s0);
        while (a < N) {
            if (ns >= KSTAGE_MAX) return false;
            T.P[ns++] = a;
            a = b;
            b = stage_next(b, N, s0, ratio);
        }
    }
    T.P[ns] = N;
    A.nstage = ns;
    // deposits per epoch step (0: the default; negative: every deposit of a
    // stage's packets counts for the next stage)
    A.depB = (run->epoch_deposits < 0) ? 0u : (run->epoch_deposits == 0) ? 16000u : (unsigned)run->epoch_deposits;
    return true;
}

// the stage table into device memory (a kernel, so that a graph captures it)
__global__ void k_set_stages(StageTab t, long long *out) {
    const int k = threadIdx.x;
    if (k <= KSTAGE_MAX) out[k] = t.P[k];
}

// ---- the pass's per-model kernels, reading model m's buffers from its
// launch descriptor (so that a captured pass stays valid when a model's
// buffers are reallocated between calls) ----------------------------------
__global__ void k_bm_prep(const BatchModel *__restrict__ bm, int m, double factRW) {
    const BatchModel &B = bm[m];
    const DevModel &md = B.md;
    const DevTally &ty = B.ty;
    const int ncell = md.ncell, nlam = B.tb.nlam;
    const size_t nspec = (size_t)2 * (NANGD + 1) * nlam;
    const size_t n = std::max((size_t)3 * ncell, nspec);
    const double E0 = B.unitE * ldexp(1.0, EABS_BITS);
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
        if (k < (size_t)ncell) {
            const int c = (int)k, i = c / md.nt;
            ty.EJv[c] = 0ull;
            ty.Eabs[c] = 0ull;
            ty.Tcell[c] = 0ull;
            ty.Ni[c] = 0;
            if (ty.thick) {
                int th = 0;
                if (B.rw.on && i > 0)
                    for (int iT = 100; iT <= 1600; iT += 500) {
                        const double lr = 1.0 / ((MCX_AU * kd_at(B.rw.KD, md.cls[c], iT).x) * md.dens[c]);
                        if ((B.rw.R[i + 1] - B.rw.R[i]) > factRW * lr) th = 1;
                    }
                ty.thick[c] = th;
            }
            // fixed-point units (see k_cell_units)
            double kmax = 0.0;
            for (int l = 0; l < nlam; l++) kmax = fmax(kmax, md.KT[(size_t)c * nlam + l].x);
            const double chord = 2.0 * sqrt(fmax(md.R2[i + 1] - md.R2[i], 0.0));
            double Bc = E0 * MCX_AU * chord * kmax * (1.0 + 1.0 / 64.0);
            if (!(Bc > 0.0)) Bc = 1e-300;
            ((double *)ty.invU1)[c] = ldexp(1.0, EJV_BITS) / Bc;
            ((double *)ty.invU2)[c] = ldexp(1.0, EJV2_BITS) / (Bc * Bc);
        }
        if (k < (size_t)3 * ncell) ty.EJv2[k] = 0ull;
        if (k < nspec) {
            ty.specE[k] = 0ull;
            ty.ispec[k] = 0u;
        }
        if (k < G_NSLOTS) ty.counters[k] = 0ull;
        if (k < G_GATE_SLOTS) ty.gate[k] = 0ull;
        if (k == 0) {
            *ty.tautot = 0.0;
            *B.next_pid = 0ull;
        }
    }
    // the other epochs' absorbed energy (epoch 0 above)
    for (size_t k = (size_t)ncell + (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)(KSTAGE_MAX + 2) * ncell;
         k += (size_t)gridDim.x * blockDim.x)
        ty.Eabs[k] = 0ull;
}

__global__ void k_bm_finalize(const BatchModel *__restrict__ bm, int m, int nb) {
    const BatchModel &B = bm[m];
    const DevModel &md = B.md;
    const DevTally &ty = B.ty;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= md.ncell) return;
    ty.oEJv[c] = (double)ty.EJv[c] / ty.invU1[c];
    const unsigned long long *L = ty.EJv2 + 3 * (size_t)c;
    const double s = ((double)L[2] * 0x1p64 + (double)L[1] * 0x1p32) + (double)L[0];
    ty.oEJv2[c] = s / ty.invU2[c];
    unsigned long long e = 0ull;
    for (int b = 0; b < nb; b++) e += ty.Eabs[(size_t)b * md.ncell + c];
    ty.oEabs[c] = (double)e * B.unitE;
    ty.oTbw[c] = e ? dev_determineT(kp_row(md, md.cls[c]), ((double)e * B.unitE) / md.mass[c], 1.0) : 0.0;
}

// Per-model set-up of a pass, host part: checks, the packet energy and the
// model's launch descriptor.
static int host_prep(mcx_ctx *ctx, const mcx_run *run, uint64_t seed, BatchModel &hb) {
    if (!ctx->have_model) return fail(ctx, MCX_ESTATE, "mcx_set_model first");
    DevRW rw = ctx->rw;
    rw.on = 0;
    if (run->randomwalk) {
        if (!rw.KD.on)
            return fail(ctx, MCX_EINVAL,
                        "random walk needs Kext and gsym (mcx_set_tables), thet (mcx_set_model)");
        if (!(run->factRW > 0.0)) return fail(ctx, MCX_EINVAL, "factRW must be positive");
        rw.on = 1;
        rw.factRW = run->factRW;
    }
    const double E0 = ctx->Lstar / ((double)(float)run->nphot * ((1.0 - 0.0) - 0.0));
    ctx->E0 = E0;
    ctx->unitE = ldexp(E0, -EABS_BITS);
    DevTally &ty = ctx->ty;
    ty.Ni = ctx->d_Ni;
    ty.Sk = ctx->d_Sk;
    ty.Tk = ctx->d_Tk;
    ty.gate = ctx->d_gate;
    ty.oEJv = ctx->d_EJv;
    ty.oEJv2 = ctx->d_EJv2;
    ty.oEabs = ctx->d_Eabs;
    ty.oTbw = ctx->d_Tbw;
    ty.thick = ctx->d_thick;
    hb.tb = ctx->tb;
    hb.md = ctx->md;
    hb.ty = ty;
    hb.rw = rw;
    hb.next_pid = ctx->d_next;
    hb.invE = ldexp(1.0, EABS_BITS) / E0;
    hb.unitE = ctx->unitE;
    hb.seed = seed;
    return MCX_OK;
}

// Per-model set-up, device part (kernels only, so that it can be captured in
// a graph): the thick flags, cleared tallies, the fixed-point units.
static int enqueue_prep(mcx_ctx *ctx, const BatchModel &hb, double factRW, hipStream_t s) {
    const DevModel &md = ctx->md;
    const DevTally &ty = ctx->ty;
    const int ncell = md.ncell, nlam = ctx->nlam;
    const dim3 g1((ncell + 255) / 256), b1(256);
    if (hb.rw.on) {
        hipLaunchKernelGGL(k_thick_cells, g1, b1, 0, s, md.nr, md.nt, hb.rw.R, md.dens, hb.rw.cls, hb.rw.KD, factRW,
                           ctx->d_thick);
    } else if (ctx->d_thick) {
        hipLaunchKernelGGL(k_fill_u32, dim3(64), b1, 0, s, (size_t)ncell, (unsigned *)ctx->d_thick, 0u);
    }
    hipLaunchKernelGGL(k_fill_u64, dim3(64), b1, 0, s, (size_t)ncell, ty.EJv, 0ull);
    hipLaunchKernelGGL(k_fill_u64, dim3(64), b1, 0, s, (size_t)3 * ncell, ty.EJv2, 0ull);
    hipLaunchKernelGGL(k_fill_u64, dim3(256), b1, 0, s, (size_t)(KSTAGE_MAX + 2) * ncell, ty.Eabs, 0ull);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), b1, 0, s, (size_t)G_GATE_SLOTS, ctx->d_gate, 0ull);
    hipLaunchKernelGGL(k_fill_u64, dim3(64), b1, 0, s, (size_t)ncell, ty.Tcell, 0ull);
    hipLaunchKernelGGL(k_fill_u64, dim3(64), b1, 0, s, (size_t)2 * (NANGD + 1) * nlam, ty.specE, 0ull);
    hipLaunchKernelGGL(k_fill_u32, dim3(64), b1, 0, s, (size_t)2 * (NANGD + 1) * nlam, ty.ispec, 0u);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), b1, 0, s, (size_t)G_NSLOTS, ty.counters, 0ull);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), b1, 0, s, (size_t)1, (unsigned long long *)ty.tautot, 0ull);
    hipLaunchKernelGGL(k_fill_u32, dim3(64), b1, 0, s, (size_t)ncell, (unsigned *)ctx->d_Ni, 0u);
    hipLaunchKernelGGL(k_cell_units, g1, b1, 0, s, md.nr, md.nt, nlam, md.R2, md.KT, ctx->E0, (double *)ty.invU1,
                       (double *)ty.invU2);
    HIPCHK(ctx, hipGetLastError());
    return MCX_OK;
}

// The code of a finished pass for one model (and its error text).
static int pass_status(mcx_ctx *ctx) {
    unsigned long long cnt[G_NSLOTS];
    HIPCHK(ctx, hipMemcpy(cnt, ctx->ty.counters, sizeof(cnt), hipMemcpyDeviceToHost));
    if (cnt[C_NTIMEOUT] > 0) {
        double d[16];
        HIPCHK(ctx, hipMemcpy(d, ctx->ty.torep, sizeof(d), hipMemcpyDeviceToHost));
        char buf[512];
        snprintf(buf, sizeof(buf),
                 "; first abandoned: packet %.0f state %.0f after %.0f cell steps in cell (%.0f, %.0f) at (%.9g, %.9g, "
                 "%.9g) AU, direction (%.9g, %.9g, %.9g), onEdge %.0f edge %.0f, tau %.6g of %.6g, %.0f crossings",
                 d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8], d[9], d[10], d[11], d[12], d[14], d[13], d[15]);
        return fail(ctx, MCX_ETIME,
                    "time limit: " + std::to_string(cnt[C_NTIMEOUT]) + " packets unfinished" + std::string(buf));
    }
    if (cnt[C_NGEOM] > 0) return fail(ctx, MCX_EGEOM, std::to_string(cnt[C_NGEOM]) + " packets found nothing to hit");
    return MCX_OK;
}

// Whether the workgroups of a batched launch move to other models once their
// own has no packet left (on, unless the environment's MCX_BATCH_HELP is 0).
static int batch_help(int nm) {
    static const bool off = [] {
        const char *e = getenv("MCX_BATCH_HELP");
        return e && e[0] == '0';
    }();
    return (nm > 1 && !off) ? 1 : 0;
}

// Launch shape of a pass.
struct LaunchCfg {
    void (*kern)(TransportArgs);
    int threads, rb, bpm;
    size_t lds, lanes;
    int blocks;       // workgroups of a batched launch with a block map
    const int *bmap;  // [blocks] model of each workgroup, or null (bpm each)
};

// The pass's device work after the descriptors are in place: set-up, the
// stages (snapshot, counters, transport launch) and the tallies' finish, all
// enqueued on s without host synchronisation (a graph captures it whole).
static void enqueue_pass(mcx_ctx *const *ctxs, int nm, const mcx_run *run, const LaunchCfg &L,
                         const std::vector<BatchModel> &hb, hipStream_t s) {
    mcx_ctx *ctx = ctxs[0];
    const BatchModel *dbm = (const BatchModel *)ctx->d_bms;
    for (int m = 0; m < nm; m++) {
        const int nc = ctxs[m]->md.ncell;
        hipLaunchKernelGGL(k_bm_prep, dim3((unsigned)std::min(1024, (3 * nc + 255) / 256 + 1)), dim3(256), 0, s, dbm, m,
                           run->factRW);
    }
    hipLaunchKernelGGL(k_fill_u4, dim3(1024), dim3(256), 0, s, L.lanes << VIS_LOG2, ctx->ty.vtab);
    TransportArgs A;
    A.bm = dbm;
    A.bpm = L.bpm;
    A.bmap = L.bmap;
    A.nmodels = nm;
    A.nphot = run->nphot;
    A.maxinteract = run->maxinteract;
    A.deadline = 0ull;
    A.serial = 0;
    A.coop = (run->coop_lanes < 0) ? COOP_LANES : run->coop_lanes;
    A.rb_mask = L.rb - 1;
    // every stage in one launch (the stages open inside it)
    StageTab T;
    (void)stage_table(run, false, A, T);
    hipLaunchKernelGGL(k_set_stages, dim3(1), dim3(64), 0, s, T, ctx->d_stages);
    A.P = ctx->d_stages;
    A.pid_base = 0;
    A.pid_end = run->nphot;
    const int blocks = (nm == 1) ? (int)std::max(1ll, std::min((long long)L.bpm, ((long long)run->nphot + L.threads - 1) / L.threads))
                       : L.bmap ? L.blocks
                                : L.bpm * nm;
    if (nm == 1) A.bpm = blocks;
    hipLaunchKernelGGL(L.kern, dim3(blocks), dim3(L.threads), L.lds, s, A);
    for (int m = 0; m < nm; m++)
        hipLaunchKernelGGL(k_bm_finalize, dim3((ctxs[m]->md.ncell + 255) / 256), dim3(256), 0, s, dbm, m, A.nstage + 1);
    (void)hb;
}

// A pass replayed from a graph: the whole sequence is captured once for a
// set of models and a run configuration and replayed for later calls (the
// descriptors, which hold the keys and energies, are rewritten before each
// replay).  No time limit and no per-stage timing in this mode.
static int graph_pass(mcx_ctx *const *ctxs, int nm, const mcx_run *run, const LaunchCfg &L,
                      const std::vector<BatchModel> &hb, int32_t *status) {
    mcx_ctx *ctx = ctxs[0];
    hipStream_t s = ctx->stream;
    char kb[256];
    snprintf(kb, sizeof(kb), "%d %lld %d %d %.17g %d %d %lld %d %d %d %d %zu %zu %d %p", nm, (long long)run->nphot,
             L.bpm, run->randomwalk, run->factRW, run->precision, run->maxinteract, (long long)run->stage0,
             run->stage_ratio, run->coop_lanes, L.threads, L.rb, L.lds, L.lanes, L.bmap ? L.blocks : 0,
             (const void *)L.bmap);
    std::string key(kb);
    // the descriptors carry every model buffer's address: the graph depends
    // on the models' sizes and on the places of the descriptors, the visit
    // tables and the stage table (which mcx_set_model reallocates)
    key += " " + std::to_string((unsigned long long)(uintptr_t)ctx->d_bms) + " " +
           std::to_string((unsigned long long)(uintptr_t)ctx->ty.vtab) + " " +
           std::to_string((unsigned long long)(uintptr_t)ctx->d_stages);
    for (int m = 0; m < nm; m++)
        key += " " + std::to_string(ctxs[m]->md.nr) + "x" + std::to_string(ctxs[m]->md.nt) + "x" +
               std::to_string(ctxs[m]->nlam) + (ctxs[m]->d_thick ? "t" : "");
    if (!ctx->gexec || ctx->graph_key != key) {
        if (ctx->gexec) (void)hipGraphExecDestroy((hipGraphExec_t)ctx->gexec);
        if (ctx->graph) (void)hipGraphDestroy((hipGraph_t)ctx->graph);
        ctx->gexec = ctx->graph = nullptr;
        ctx->graph_key.clear();
        HIPCHK(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        enqueue_pass(ctxs, nm, run, L, hb, s);
        hipGraph_t g = nullptr;
        const hipError_t e1 = hipStreamEndCapture(s, &g);
        if (e1 != hipSuccess || !g) return fail(ctx, MCX_EDEVICE, std::string("graph capture: ") + hipGetErrorString(e1));
        hipGraphExec_t ex = nullptr;
        const hipError_t e2 = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        if (e2 != hipSuccess) {
            (void)hipGraphDestroy(g);
            return fail(ctx, MCX_EDEVICE, std::string("graph instantiate: ") + hipGetErrorString(e2));
        }
        ctx->graph = g;
        ctx->gexec = ex;
        ctx->graph_key = key;
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev[0], s));
    HIPCHK(ctx, hipGraphLaunch((hipGraphExec_t)ctx->gexec, s));
    HIPCHK(ctx, hipEventRecord(ctx->ev[3], s));
    HIPCHK(ctx, hipEventSynchronize(ctx->ev[3]));
    float pms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&pms, ctx->ev[0], ctx->ev[3]));
    int rc = MCX_OK;
    for (int m = 0; m < nm; m++) {
        mcx_ctx *c = ctxs[m];
        c->kernel_ms = pms;
        c->pass_ms = pms;
        c->have_pass = true;
        const int r = pass_status(c);
        if (status) status[m] = r;
        if (r && !rc) rc = r;
    }
    if (rc && nm > 1) ctx->err = "a model's pass failed (see its status word and its context's error)";
    return rc;
}

// One pass of nm models in common launches (one per stage), on the first
// model's stream: bpm blocks per model (0: the device's resident blocks
// shared out), the staged schedule or one lane.
static int run_pass(mcx_ctx *const *ctxs, int nm, const mcx_run *run, int bpm_req, const int32_t *mblocks,
                    int64_t seed_stride, int use_graph, int32_t *status) {
    mcx_ctx *ctx = ctxs[0];
    if (run->nphot <= 0 || run->nphot >= (1ll << 31)) return fail(ctx, MCX_EINVAL, "nphot must be in 1 .. 2^31-1");
    if (run->precision != MCX_FP64 && run->precision != MCX_FP32) return fail(ctx, MCX_EINVAL, "unknown precision mode");
    if (run->stage0 < 0 || run->stage_ratio < -1 || run->stage_ratio == 1)
        return fail(ctx, MCX_EINVAL, "stage0 >= 0 and stage_ratio -1, 0 or >= 2 required");
    const bool serial = (run->nlanes == 1);
    if (serial && nm > 1) return fail(ctx, MCX_EINVAL, "a one-lane pass runs one model");
    {
        TransportArgs T;
        StageTab P;
        if (!stage_table(run, serial, T, P))
            return fail(ctx, MCX_EINVAL, "the stage schedule has more than " + std::to_string(KSTAGE_MAX) +
                                             " stages (raise stage0 or stage_ratio)");
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    // one kernel for all models: the same species case
    const int mode = (ctx->ng == 1) ? 0 : ctx->md.CUMmix ? 1 : 2;
    int nrmax = 0, ntmax = 0, nlmax = 0;
    bool anyrw = false;
    for (int m = 0; m < nm; m++) {
        mcx_ctx *c = ctxs[m];
        if (!c || !c->have_model) return fail(ctx, MCX_ESTATE, "model " + std::to_string(m) + ": mcx_set_model first");
        if (c->device != ctx->device) return fail(ctx, MCX_EINVAL, "all models on one device");
        const int mo = (c->ng == 1) ? 0 : c->md.CUMmix ? 1 : 2;
        if (mo != mode) return fail(ctx, MCX_EINVAL, "models of one launch need the same species case");
        nrmax = std::max(nrmax, c->md.nr);
        ntmax = std::max(ntmax, c->md.nt);
        nlmax = std::max(nlmax, c->nlam);
    }
    anyrw = run->randomwalk != 0;
    void (*kern)(TransportArgs);
    if (batch_help(serial ? 1 : nm)) {
        if (run->precision == MCX_FP32)
            kern = (mode == 0)   ? k_transport<0, true, true>
                   : (mode == 1) ? k_transport<1, true, true>
                                 : k_transport<2, true, true>;
        else
            kern = (mode == 0)   ? k_transport<0, false, true>
                   : (mode == 1) ? k_transport<1, false, true>
                                 : k_transport<2, false, true>;
    } else if (run->precision == MCX_FP32) {
        kern = (mode == 0) ? k_transport<0, true, false> : (mode == 1) ? k_transport<1, true, false> : k_transport<2, true, false>;
    } else {
        kern = (mode == 0) ? k_transport<0, false, false> : (mode == 1) ? k_transport<1, false, false> : k_transport<2, false, false>;
    }
    // launch shape: a full device of resident wavefronts, or one lane
    int threads = serial ? 64 : 256; // one wavefront: lane 0 traces, the others serve its lookups
    size_t lds = sizeof(double) * ((size_t)(nrmax + 1) + 2 * (size_t)(ntmax + 1) + 2 * 181 + (size_t)(nlmax - 1) +
                                   NPL * (size_t)threads);
    if (anyrw) lds += sizeof(double) * ((size_t)(nrmax + 1) + (size_t)(ntmax + 1));
    // rings of Philox draws: 16 per thread when they fit beside the rest in
    // the CU's LDS (one workgroup per CU; a batched launch's two workgroups
    // per CU share it), else fewer
    int rb = 16;
    const size_t lds_cap = batch_help(serial ? 1 : nm) ? std::min(ctx->lds_max, (size_t)80 * 1024) : ctx->lds_max;
    while (rb > 4 && lds + sizeof(double) * (size_t)rb * threads > lds_cap) rb >>= 1;
    lds += sizeof(double) * (size_t)rb * threads;
    if (lds > ctx->lds_max)
        return fail(ctx, MCX_EINVAL, "grid too large for the kernel's LDS tables (" + std::to_string(lds) + " bytes)");
    int resident = 1;
    if (!serial) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, lds) != hipSuccess || per_cu < 1)
            per_cu = 1;
        resident = per_cu * ctx->sm_count;
        if (run->nlanes > 1) resident = std::max(1, std::min(resident, (run->nlanes + threads - 1) / threads));
    }
    const int bpm = serial ? 1 : (bpm_req > 0 ? bpm_req : std::max(1, resident / nm));
    // an explicit share of the workgroups per model: the block map
    long long nblk = (long long)bpm * nm;
    if (mblocks) {
        if (serial) return fail(ctx, MCX_EINVAL, "a one-lane pass runs one model");
        nblk = 0;
        for (int m = 0; m < nm; m++) {
            if (mblocks[m] < 1) return fail(ctx, MCX_EINVAL, "model " + std::to_string(m) + ": blocks must be >= 1");
            nblk += mblocks[m];
        }
        if (nblk > (1ll << 20)) return fail(ctx, MCX_EINVAL, "more than 2^20 workgroups in one launch");
    }
    // the lanes' visit tables (cleared per pass: tags are packet numbers)
    const size_t lanes = (size_t)nblk * threads;
    if (lanes > ctx->vtab_lanes) {
        HIPCHK(ctx, hipStreamSynchronize(s));
        free_list(ctx->run_alloc);
        ctx->ty.vtab = nullptr;
        ctx->vtab_lanes = 0;
        int r = dalloc(ctx, ctx->run_alloc, &ctx->ty.vtab, lanes << VIS_LOG2);
        if (r) return r;
        ctx->vtab_lanes = lanes;
    }
    // the launch descriptors
    if ((size_t)nm > ctx->bm_cap) {
        HIPCHK(ctx, hipStreamSynchronize(s));
        if (ctx->d_bms) (void)hipFree(ctx->d_bms);
    if (ctx->gexec) (void)hipGraphExecDestroy((hipGraphExec_t)ctx->gexec);
    if (ctx->graph) (void)hipGraphDestroy((hipGraph_t)ctx->graph);
        ctx->d_bms = nullptr;
        ctx->bm_cap = 0;
        if (hipMalloc(&ctx->d_bms, sizeof(BatchModel) * (size_t)nm) != hipSuccess)
            return fail(ctx, MCX_ENOMEM, "launch descriptors");
        ctx->bm_cap = (size_t)nm;
    }
    std::vector<BatchModel> hb(nm);
    for (int m = 0; m < nm; m++) {
        HIPCHK(ctx, hipStreamSynchronize(ctxs[m]->stream));
        int r = host_prep(ctxs[m], run, run->seed + (uint64_t)(seed_stride * m), hb[m]);
        if (r) return (m == 0) ? r : fail(ctx, r, "model " + std::to_string(m) + ": " + ctxs[m]->err);
        hb[m].ty.vtab = ctx->ty.vtab;
    }
    // the descriptors hold each model's key: a graph replays with the keys
    // of this call
    HIPCHK(ctx, hipMemcpy(ctx->d_bms, hb.data(), sizeof(BatchModel) * (size_t)nm, hipMemcpyHostToDevice));
    const int *dmap = nullptr;
    if (mblocks) {
        std::vector<int> hmap((size_t)nblk);
        for (int m = 0, b = 0; m < nm; m++)
            for (int k = 0; k < mblocks[m]; k++) hmap[b++] = m;
        if ((size_t)nblk > ctx->bmap_cap) {
            HIPCHK(ctx, hipStreamSynchronize(s));
            if (ctx->d_bmap) (void)hipFree(ctx->d_bmap);
            ctx->d_bmap = nullptr;
            ctx->bmap_cap = 0;
            if (hipMalloc(&ctx->d_bmap, sizeof(int) * (size_t)nblk) != hipSuccess)
                return fail(ctx, MCX_ENOMEM, "block map");
            ctx->bmap_cap = (size_t)nblk;
        }
        HIPCHK(ctx, hipMemcpy(ctx->d_bmap, hmap.data(), sizeof(int) * (size_t)nblk, hipMemcpyHostToDevice));
        dmap = (const int *)ctx->d_bmap;
    }
    if (use_graph) {
        if (serial) return fail(ctx, MCX_EINVAL, "graph replay runs concurrent passes only");
        LaunchCfg L{kern, threads, rb, bpm, lds, lanes, (int)nblk, dmap};
        return graph_pass(ctxs, nm, run, L, hb, status);
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev[0], s));
    for (int m = 0; m < nm; m++) {
        int r = enqueue_prep(ctxs[m], hb[m], run->factRW, s);
        if (r) return r;
    }
    hipLaunchKernelGGL(k_fill_u4, dim3(1024), dim3(256), 0, s, lanes << VIS_LOG2, ctx->ty.vtab);
    TransportArgs A;
    A.bm = (const BatchModel *)ctx->d_bms;
    A.bpm = bpm;
    A.bmap = dmap;
    A.nmodels = nm;
    A.nphot = run->nphot;
    A.maxinteract = run->maxinteract;
    A.deadline = 0ull;
    if (run->max_ms > 0.0 && ctx->wclk_khz > 0) {
        unsigned long long now = 0;
        hipLaunchKernelGGL(k_clock, dim3(1), dim3(1), 0, s, ctx->ty.counters + G_CLOCK);
        HIPCHK(ctx, hipMemcpyAsync(&now, ctx->ty.counters + G_CLOCK, sizeof(now), hipMemcpyDeviceToHost, s));
        HIPCHK(ctx, hipStreamSynchronize(s));
        A.deadline = now + (unsigned long long)(run->max_ms * (double)ctx->wclk_khz);
    }
    A.serial = serial ? 1 : 0;
    A.coop = (run->coop_lanes < 0) ? COOP_LANES : run->coop_lanes;
    A.rb_mask = rb - 1;
    // one launch for the whole pass: the stages open inside it (one stage
    // when one lane runs it)
    StageTab T;
    (void)stage_table(run, serial, A, T);
    hipLaunchKernelGGL(k_set_stages, dim3(1), dim3(64), 0, s, T, ctx->d_stages);
    A.P = ctx->d_stages;
    for (int m = 0; m < nm; m++) {
        mcx_ctx *c = ctxs[m];
        hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, s, c->d_next, 0ull);
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, (size_t)7, c->ty.counters + G_DONE, 0ull);
    }
    A.pid_base = 0;
    A.pid_end = run->nphot;
    const int blocks = serial ? 1
                       : (nm == 1) ? (int)std::max(1ll, std::min((long long)bpm, ((long long)run->nphot + threads - 1) / threads))
                       : dmap      ? (int)nblk
                                   : bpm * nm;
    if (nm == 1) A.bpm = blocks;
    HIPCHK(ctx, hipEventRecord(ctx->ev[1], s));
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(threads), lds, s, A);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipEventRecord(ctx->ev[2], s));
    HIPCHK(ctx, hipEventSynchronize(ctx->ev[2]));
    float kms_total = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&kms_total, ctx->ev[1], ctx->ev[2]));
    // MCX_STAGE_LOG set: when each stage of each model opened (ms after the
    // launch's start)
    static const bool stage_log = getenv("MCX_STAGE_LOG") != nullptr;
    if (stage_log && !serial && ctx->wclk_khz > 0) {
        for (int m = 0; m < nm; m++) {
            std::vector<unsigned long long> g(G_GATE_SLOTS), cn(G_NSLOTS);
            HIPCHK(ctx, hipMemcpy(g.data(), ctxs[m]->d_gate, sizeof(unsigned long long) * G_GATE_SLOTS,
                                  hipMemcpyDeviceToHost));
            HIPCHK(ctx, hipMemcpy(cn.data(), ctxs[m]->ty.counters, sizeof(unsigned long long) * G_NSLOTS,
                                  hipMemcpyDeviceToHost));
            fprintf(stderr, "[mcx] model %d: %d stages, B %u; stage opened at (ms):", m, A.nstage, A.depB);
            for (int k = 1; k < A.nstage; k++)
                fprintf(stderr, " %.1f", (double)(g[G_TOPEN + k] - cn[G_TLINE]) / (double)ctx->wclk_khz);
            fprintf(stderr, "\n");
        }
    }
    for (int m = 0; m < nm; m++) {
        mcx_ctx *c = ctxs[m];
        const DevModel &md = c->md;
        const int ncell = md.ncell;
        hipLaunchKernelGGL(k_finalize, dim3((ncell + 255) / 256), dim3(256), 0, s, ncell, A.nstage + 1, c->ty.EJv,
                           c->ty.EJv2,
                           c->ty.Eabs, c->ty.invU1, c->ty.invU2, c->unitE, md, c->d_EJv,
                           c->d_EJv2, c->d_Eabs, c->d_Tbw);
        HIPCHK(ctx, hipGetLastError());
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev[3], s));
    HIPCHK(ctx, hipEventSynchronize(ctx->ev[3]));
    float pms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&pms, ctx->ev[0], ctx->ev[3]));
    int rc = MCX_OK;
    for (int m = 0; m < nm; m++) {
        mcx_ctx *c = ctxs[m];
        c->kernel_ms = kms_total;
        c->pass_ms = pms;
        c->have_pass = true;
        const int r = pass_status(c);
        if (status) status[m] = r;
        if (r && !rc) rc = r;
    }
    if (rc && nm > 1) ctx->err = "a model's pass failed (see its status word and its context's error)";
    return rc;
}

extern "C" int mcx_transport(mcx_ctx *ctx, const mcx_run *run) {
    if (!ctx || !run) return MCX_EINVAL;
    return run_pass(&ctx, 1, run, 0, nullptr, 0, 0, nullptr);
}

extern "C" int mcx_batch_transport(mcx_ctx *const *ctxs, int32_t nmodels, const mcx_run *run,
                                   int32_t blocks_per_model, int64_t seed_stride, int32_t use_graph,
                                   int32_t *status) {
    if (!ctxs || !run || nmodels < 1 || !ctxs[0]) return MCX_EINVAL;
    return run_pass(ctxs, nmodels, run, blocks_per_model, nullptr, seed_stride, use_graph, status);
}

extern "C" int mcx_batch_transport_blocks(mcx_ctx *const *ctxs, int32_t nmodels, const mcx_run *run,
                                          const int32_t *model_blocks, int64_t seed_stride, int32_t use_graph,
                                          int32_t *status) {
    if (!ctxs || !run || nmodels < 1 || !ctxs[0] || !model_blocks) return MCX_EINVAL;
    return run_pass(ctxs, nmodels, run, 0, model_blocks, seed_stride, use_graph, status);
}

extern "C" int mcx_get_tallies(mcx_ctx *ctx, mcx_tallies *o) {
    if (!ctx || !o) return MCX_EINVAL;
    if (!ctx->have_pass) return fail(ctx, MCX_ESTATE, "no transport pass yet");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    const int ncell = ctx->md.ncell, nlam = ctx->nlam;
    if (o->EJv) HIPCHK(ctx, hipMemcpy(o->EJv, ctx->d_EJv, sizeof(double) * ncell, hipMemcpyDeviceToHost));
    if (o->EJv2) HIPCHK(ctx, hipMemcpy(o->EJv2, ctx->d_EJv2, sizeof(double) * ncell, hipMemcpyDeviceToHost));
    if (o->Eabs) HIPCHK(ctx, hipMemcpy(o->Eabs, ctx->d_Eabs, sizeof(double) * ncell, hipMemcpyDeviceToHost));
    if (o->Tbw) HIPCHK(ctx, hipMemcpy(o->Tbw, ctx->d_Tbw, sizeof(double) * ncell, hipMemcpyDeviceToHost));
    if (o->Ni) HIPCHK(ctx, hipMemcpy(o->Ni, ctx->d_Ni, sizeof(int) * ncell, hipMemcpyDeviceToHost));
    if (o->thick) {
        if (ctx->d_thick) HIPCHK(ctx, hipMemcpy(o->thick, ctx->d_thick, sizeof(int) * ncell, hipMemcpyDeviceToHost));
        else memset(o->thick, 0, sizeof(int32_t) * ncell);
    }
    const size_t nb = (size_t)(NANGD + 1) * nlam;
    std::vector<unsigned long long> sE(2 * nb);
    std::vector<unsigned> sc(2 * nb);
    HIPCHK(ctx, hipMemcpy(sE.data(), ctx->ty.specE, sizeof(unsigned long long) * 2 * nb, hipMemcpyDeviceToHost));
    HIPCHK(ctx, hipMemcpy(sc.data(), ctx->ty.ispec, sizeof(unsigned) * 2 * nb, hipMemcpyDeviceToHost));
    // spec(l,a) = sum over escapes of wangle(a) E wlam(l)  (MCRadiation.f:54-78, 631-645)
    std::vector<double> wlam(nlam), wangle(NANGD + 1, 0.0), cosangle(NANGD + 1);
    const std::vector<double> &nu = ctx->h_nu;
    for (int i = 1; i <= nlam; i++) {
        if (i == 1) wlam[0] = -1e23 / (nu[1] / 2.0);
        else if (i == nlam) wlam[nlam - 1] = -1e23 / ((nu[nlam - 1] - nu[nlam - 2]) / 2.0);
        else wlam[i - 1] = -1e23 / ((nu[i] - nu[i - 2]) / 2.0);
    }
    for (int i = 0; i <= NANGD; i++) cosangle[i] = (double)i / (double)NANGD;
    for (int i = 1; i <= NANGD; i++) wangle[i] = 1.0 / (cosangle[i] - cosangle[i - 1]);
    double Eesc = 0.0;
    for (int a = 0; a <= NANGD; a++)
        for (int l = 0; l < nlam; l++) {
            size_t k = (size_t)a * nlam + l;
            double E = (double)sE[k] * ctx->unitE, Es = (double)sE[nb + k] * ctx->unitE;
            Eesc += E;
            if (o->spec) o->spec[k] = (wangle[a] * E) * wlam[l];
            if (o->scatspec) o->scatspec[k] = (wangle[a] * Es) * wlam[l];
            if (o->ispec) o->ispec[k] = (int32_t)sc[k];
            if (o->iscatspec) o->iscatspec[k] = (int32_t)sc[nb + k];
        }
    unsigned long long cnt[G_NSLOTS];
    double tt = 0.0;
    HIPCHK(ctx, hipMemcpy(cnt, ctx->ty.counters, sizeof(cnt), hipMemcpyDeviceToHost));
    HIPCHK(ctx, hipMemcpy(&tt, ctx->ty.tautot, sizeof(double), hipMemcpyDeviceToHost));
    o->nemit = (int64_t)cnt[C_NEMIT];
    o->ninteract = (int64_t)cnt[C_NINTERACT];
    o->nabs = (int64_t)cnt[C_NABS];
    o->nscat = (int64_t)cnt[C_NSCAT];
    o->nescape = (int64_t)cnt[C_NESCAPE];
    o->nstarhit = (int64_t)cnt[C_NSTARHIT];
    o->nsteps = (int64_t)cnt[C_NSTEPS];
    o->nraar = (int64_t)cnt[C_NRAAR];
    o->nmaxinteract = (int64_t)cnt[C_NMAXINT];
    o->ngeomfail = (int64_t)cnt[C_NGEOM];
    o->nrw = (int64_t)cnt[C_NRW];
    o->noverflow = (int64_t)cnt[C_NOVERFLOW];
    o->tautot = tt;
    o->Eescape = Eesc;
    o->maxsteps = (int64_t)cnt[G_MAXSTEPS];
    o->nvisit_overflow = (int64_t)cnt[C_NVISOVF];
    o->ntimeout = (int64_t)cnt[C_NTIMEOUT];
    o->ntallyatomic = (int64_t)cnt[C_NTALLYATOM];
    for (int k = 0; k < 5; k++) {
        const long long t0 = (long long)cnt[G_TLINE], t1 = (long long)cnt[G_TLINE + 1 + k];
        o->t_done[k] = (t1 != 0 && ctx->wclk_khz > 0) ? (double)(t1 - t0) / (double)ctx->wclk_khz : -1.0;
    }
    return MCX_OK;
}

extern "C" int mcx_temperature(mcx_ctx *ctx, mcx_temps *o) {
    if (!ctx) return MCX_EINVAL;
    if (!ctx->have_pass) return fail(ctx, MCX_ESTATE, "no transport pass yet");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const DevModel &md = ctx->md;
    const int ncell = md.ncell;
    hipLaunchKernelGGL(k_temperature, dim3((ncell + 255) / 256), dim3(256), 0, ctx->stream, md.nr, md.nt, ctx->ng,
                       ctx->d_EJv, ctx->d_EJv2, ctx->d_Ni, ctx->d_Tbw, ctx->d_V, md, md.w, ctx->tb.Kp,
                       ctx->Tvoid, ctx->d_T, ctx->d_dT, ctx->d_dEJv);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->have_temps = true;
    if (o) {
        if (o->T) HIPCHK(ctx, hipMemcpy(o->T, ctx->d_T, sizeof(double) * ncell, hipMemcpyDeviceToHost));
        if (o->dT) HIPCHK(ctx, hipMemcpy(o->dT, ctx->d_dT, sizeof(double) * ncell, hipMemcpyDeviceToHost));
        if (o->dEJv) HIPCHK(ctx, hipMemcpy(o->dEJv, ctx->d_dEJv, sizeof(double) * ncell, hipMemcpyDeviceToHost));
    }
    return MCX_OK;
}

extern "C" int mcx_set_temperature(mcx_ctx *ctx, const double *T, const double *dT, const int32_t *Ni) {
    if (!ctx || !T) return MCX_EINVAL;
    if (!ctx->have_model) return fail(ctx, MCX_ESTATE, "mcx_set_model first");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int ncell = ctx->md.ncell;
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_T, T, sizeof(double) * ncell, hipMemcpyHostToDevice, s));
    if (dT) HIPCHK(ctx, hipMemcpyAsync(ctx->d_dT, dT, sizeof(double) * ncell, hipMemcpyHostToDevice, s));
    else HIPCHK(ctx, hipMemsetAsync(ctx->d_dT, 0, sizeof(double) * ncell, s));
    if (Ni) HIPCHK(ctx, hipMemcpyAsync(ctx->d_Ni, Ni, sizeof(int) * ncell, hipMemcpyHostToDevice, s));
    else HIPCHK(ctx, hipMemsetAsync(ctx->d_Ni, 0, sizeof(int) * ncell, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    ctx->have_temps = true;
    return MCX_OK;
}

extern "C" int mcx_last_timing(const mcx_ctx *ctx, double *kernel_ms, double *pass_ms) {
    if (!ctx) return MCX_EINVAL;
    if (kernel_ms) *kernel_ms = ctx->kernel_ms;
    if (pass_ms) *pass_ms = ctx->pass_ms;
    return MCX_OK;
}
